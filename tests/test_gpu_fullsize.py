"""Parity at the benchmark configurations' full sizes (MI355X only).

Each product runs whole on the GPU; 64 sampled columns (8 per problem for the
512-problem batch) are checked against the CPU oracle: the explicitly built
columns of the structured matrix times the host copy of the dense operand, in
FP64.  Bars: 1e-13 for the plain FP64 forms, 1e-12 for the split and
compressed forms.
"""
import numpy as np
import pytest

from oracle import inputs
from oracle import oracle as ora
from psmma_git_amd import api

pytestmark = pytest.mark.gpu

COLS = 64


def dev_cm(x):
    import torch

    return torch.from_numpy(np.ascontiguousarray(x.T)).cuda().t()


def sampled(c, cols):
    import torch

    return c[:, torch.from_numpy(cols).cuda()].cpu().numpy()


@pytest.mark.parametrize("kind", ["toeplitz", "hankel"])
def test_window_kinds_n8192(gpu, kind):
    import torch

    n = 8192
    print(f"case {kind} N={n}", flush=True)
    g_np, a_np = inputs.toeplitz_hankel_inputs(n, seed=0, kind=kind)
    a, g = dev_cm(a_np), torch.from_numpy(g_np).cuda()
    c = api.colmajor_empty(n, n)
    cols = np.linspace(0, n - 1, COLS).astype(np.int64)
    rows = np.arange(n)[:, None]
    h = g_np[(n - 1) + cols[None, :] - rows] if kind == "toeplitz" else g_np[rows + cols[None, :]]
    want = a_np @ h
    # the configuration's 3-slice bf16 form (FP32 accumulation) stays near 1e-7
    for prec, slices, bar in ((api.PREC_FP64, 0, 1e-13), (api.PREC_SPLIT_I8, 13, 1e-12),
                              (api.PREC_SPLIT_BF16, 3, 5e-7)):
        if kind == "toeplitz":
            api.toeplitz_dev(1.0, a, g, n, n, c, prec=prec, slices=slices)
        else:
            api.hankel_dev(1.0, a, g, n, c, prec=prec, slices=slices)
        torch.cuda.synchronize()
        err = ora.rel_fro(sampled(c, cols), want)
        print(f"  prec {prec} slices {slices}: rel err {err:.3e}", flush=True)
        assert err < bar


def test_eig_merge_n16384(gpu):
    import torch

    n = 16384
    print(f"case eigenvector merge N={n}", flush=True)
    d_np, z_np, _ = inputs.merge_poles_and_z(n, 0)
    td, tz = torch.from_numpy(d_np).cuda(), torch.from_numpy(z_np).cuda()
    difl, difr, u, v = api.secular_dev(td, tz, 1.0, loewner=True)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(0)
    a = torch.randn((n, n), dtype=torch.float64, device="cuda", generator=gen).t()
    cols = np.linspace(0, n - 1, COLS).astype(np.int64)
    host = {k: t.cpu().numpy() for k, t in (("difl", difl), ("difr", difr), ("u", u), ("v", v))}
    cm = np.stack([ora.cauchyeig_matrix(d_np, host["u"], host["v"], host["difl"], host["difr"],
                                        0, n, int(j), 1)[:, 0] for j in cols], axis=1)
    want = a.cpu().numpy() @ cm
    x = api.colmajor_empty(n, n)
    runs = [("fp64", lambda: api.cauchyeig_dev(1.0, a, td, u, v, difl, difr, n, x), 1e-13),
            ("split_i8_t14", lambda: api.cauchyeig_dev(1.0, a, td, u, v, difl, difr, n, x,
                                                       prec=api.PREC_SPLIT_I8, slices=14), 1e-12),
            ("blr_rrqr", lambda: api.cauchyeig_blr_dev(1.0, a, td, u, v, difl, difr, x, leaf=256,
                                                       tol=1e-12), 1e-12),
            ("blr_rrlu", lambda: api.cauchyeig_blr_dev(1.0, a, td, u, v, difl, difr, x, leaf=256,
                                                       tol=1e-12, compressor=api.BLR_RRLU), 1e-12)]
    for name, fn, bar in runs:
        x.fill_(float("nan"))
        fn()
        torch.cuda.synchronize()
        err = ora.rel_fro(sampled(x, cols), want)
        print(f"  {name}: rel err {err:.3e}", flush=True)
        assert err < bar


def test_batch_512(gpu):
    import torch

    print("case 512-problem batch", flush=True)
    probs_np = inputs.batch_inputs(512, seed=0)
    probs, keep = [], []
    rng = np.random.default_rng(5)
    for pr in probs_np:
        n = pr["n"]
        a = dev_cm(pr["dense"])
        c = api.colmajor_empty(n, n)
        if pr["kind"] == 0:
            gens = [torch.from_numpy(np.ascontiguousarray(pr["gens"][:, i])).cuda()
                    for i in range(4)]
            probs.append(api.make_problem(api.KIND_CAUCHY, 1.0, a, gens, n, c))
        else:
            gens = [torch.from_numpy(pr["gens"]).cuda()]
            probs.append(api.make_problem(pr["kind"], 1.0, a, gens, n, c, lg=2 * n - 1, korder=n))
        cols = np.sort(rng.choice(n, size=min(8, n), replace=False))
        if pr["kind"] == 0:
            b = pr["gens"]
            h = np.stack([ora.cauchy_matrix(b[:, 0], b[:, 1], b[:, 2], b[:, 3], 0, n, int(j), 1)[:, 0]
                          for j in cols], axis=1)
        elif pr["kind"] == 1:
            h = pr["gens"][(n - 1) + cols[None, :] - np.arange(n)[:, None]]
        else:
            h = pr["gens"][np.arange(n)[:, None] + cols[None, :]]
        keep.append((a, gens, c, cols, pr["dense"] @ h))
    for prec, bar in ((api.PREC_FP64, 1e-13), (api.PREC_SPLIT_I8, 1e-12)):
        plan = api.BatchedPlan(probs, prec=prec, slices=0)
        info = plan.info()
        for _, _, c, _, _ in keep:
            c.fill_(float("nan"))
        plan.launch()
        torch.cuda.synchronize()
        errs = [ora.rel_fro(sampled(c, cols), want) for _, _, c, cols, want in keep]
        print(f"  prec {prec} plan info {info}: max rel err {max(errs):.3e}", flush=True)
        assert max(errs) < bar
        plan.close()
