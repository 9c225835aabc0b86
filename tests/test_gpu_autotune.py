"""The autotuner (MI355X only): tile-configuration tuning of the FP64 kernel and
the measured split depth.  Results after tuning must stay within the parity bar."""
import numpy as np
import pytest

from oracle import inputs
from psmma_git_amd import api

pytestmark = pytest.mark.gpu


def dev_cm(x):
    import torch

    return torch.from_numpy(np.ascontiguousarray(x.T)).cuda().t()


@pytest.mark.parametrize("kind", [api.KIND_TOEPLITZ, api.KIND_CAUCHY])
def test_autotune_f64_then_parity(gpu, ora, kind):
    import torch

    n = 1100  # its own size bucket (2048 x 2048 x 2048 after rounding up)
    cfg, ms = api.autotune_f64(kind, n, n, n)
    print(f"kind {kind}: tuned cfg {cfg}, {ms:.3f} ms")
    assert 0 <= cfg <= 6 and ms > 0
    rng = np.random.default_rng(31 + kind)
    a = rng.uniform(-1, 1, (n, n))
    c = api.colmajor_empty(n, n)
    if kind == api.KIND_TOEPLITZ:
        g = rng.standard_normal(2 * n - 1)
        api.toeplitz_dev(1.0, dev_cm(a), torch.from_numpy(g).cuda(), n, n, c)
        want = ora.structured_gemm_np(1, 1.0, a, dict(g=g, korder=n), 0, 0, n)
    else:
        b = inputs.cauchy_generators(n, rng)
        d, f, u, v = (torch.from_numpy(np.ascontiguousarray(b[:, i])).cuda() for i in range(4))
        api.cauchy_dev(1.0, dev_cm(a), d, f, u, v, n, c)
        want = ora.structured_gemm_np(0, 1.0, a, dict(d=b[:, 0], f=b[:, 1], u=b[:, 2],
                                                       v=b[:, 3]), 0, 0, n)
    torch.cuda.synchronize()
    assert ora.rel_fro(c.cpu().numpy(), want) < 1e-13


def test_split_depth_is_measured(gpu, ora):
    import torch

    n = 1536
    rng = np.random.default_rng(41)
    a = rng.uniform(-1, 1, (n, n))
    g = rng.standard_normal(2 * n - 1)
    da, dg = dev_cm(a), torch.from_numpy(g).cuda()
    c = torch.full((n, n), 5.0, dtype=torch.float64, device="cuda").t()
    p = api.make_problem(api.KIND_TOEPLITZ, 1.0, da, [dg], n, c, lg=2 * n - 1, korder=n)
    t12, e12, ok12 = api.split_slices_dev(p, api.PREC_SPLIT_I8, 1e-12)
    t6, e6, ok6 = api.split_slices_dev(p, api.PREC_SPLIT_I8, 1e-6)
    print(f"moduli for 1e-12: {t12} (sample err {e12:.2e}); for 1e-6: {t6} ({e6:.2e})")
    assert ok12 and ok6 and e12 <= 1e-12 and e6 <= 1e-6 and t6 < t12
    torch.cuda.synchronize()
    assert bool((c == 5.0).all())  # the problem's output is untouched
    # the whole product at the chosen depth meets the target; one modulus fewer does not
    api.toeplitz_dev(1.0, da, dg, n, n, c, prec=api.PREC_SPLIT_I8, slices=t12)
    torch.cuda.synchronize()
    want = ora.structured_gemm_np(1, 1.0, a, dict(g=g, korder=n), 0, 0, n)
    assert ora.rel_fro(c.cpu().numpy(), want) <= 1e-12
    api.toeplitz_dev(1.0, da, dg, n, n, c, prec=api.PREC_SPLIT_I8, slices=t12 - 1)
    torch.cuda.synchronize()
    assert ora.rel_fro(c.cpu().numpy(), want) > 1e-12 / 16
    tb, eb, okb = api.split_slices_dev(p, api.PREC_SPLIT_BF16, 1e-12)
    print(f"bf16 slices for 1e-12: {tb} ({eb:.2e})")
    assert okb and 3 <= tb <= 7


def test_split_plan_uses_measured_depth(gpu, ora):
    import torch

    rng = np.random.default_rng(42)
    probs, keep, alive = [], [], []  # alive: the device tensors the problems point to
    for i, kind in enumerate([0, 1, 2, 0, 1, 2]):
        n = int(rng.integers(200, 700))
        a = rng.uniform(-1, 1, (n, n))
        da, dc = dev_cm(a), api.colmajor_empty(n, n)
        if kind == 0:
            b = inputs.cauchy_generators(n, rng)
            gens = [torch.from_numpy(np.ascontiguousarray(b[:, j])).cuda() for j in range(4)]
            alive.append((da, gens))
            probs.append(api.make_problem(0, 1.0, da, gens, n, dc))
            want = ora.structured_gemm_np(0, 1.0, a, dict(d=b[:, 0], f=b[:, 1], u=b[:, 2],
                                                           v=b[:, 3]), 0, 0, n)
        else:
            g = rng.standard_normal(2 * n - 1)
            dg = torch.from_numpy(g).cuda()
            alive.append((da, dg))
            probs.append(api.make_problem(kind, 1.0, da, [dg], n, dc, lg=2 * n - 1, korder=n))
            want = ora.structured_gemm_np(kind, 1.0, a, dict(g=g, korder=n), 0, 0, n)
        keep.append((dc, want))
    plan = api.BatchedPlan(probs, prec=api.PREC_SPLIT_I8, slices=0)
    t, nodes = plan.info()
    print("measured moduli", t)
    assert 2 <= t <= 16
    plan.launch()
    torch.cuda.synchronize()
    errs = [ora.rel_fro(dc.cpu().numpy(), w) for dc, w in keep]
    print("errors", max(errs))
    assert max(errs) <= 1e-12
    plan.close()
