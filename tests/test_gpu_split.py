"""Split-precision (int8 modular) products against the FP64 oracle (MI355X only).

Bar: relative Frobenius error <= 1e-12 for t >= 13 moduli at the sizes below;
the error shrinks as t grows (each modulus adds ~8 bits to the exact integer
range, i.e. ~4 bits per operand).
"""
import numpy as np
import pytest

from psmma_git_amd import api
from test_gpu_parity import ORA_KIND, case_seed, dev_colmajor, host, make_gens, run_dev

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["cauchy", "toeplitz", "hankel", "cauchyeig"])
# (1100, 1037, 700): 25 tiles of 256 x 256 per modulus, so t >= 13 takes the
# direct-to-LDS kernel with ragged tiles, a padded residue pitch and K % 64 != 0
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (100, 130, 77), (513, 257, 1000), (1024, 1024, 1024),
                                   (1100, 1037, 700)])
def test_split_i8_parity(gpu, ora, kind, m, n, k):
    import torch

    print(f"case split kind={kind} m={m} n={n} k={k}", flush=True)
    rng = np.random.default_rng(case_seed("s", kind, m, n, k))
    nglob = max(k, n) + 5
    gens, dev = make_gens(kind, nglob, rng)
    a = rng.uniform(-1, 1, (m, k))
    c = rng.standard_normal((m, n))
    want = ora.structured_gemm_np(ORA_KIND[kind], 0.75, a, gens, 2, 1, n, beta=0.5, c=c)
    errs = {}
    for t in (10, 13, 14):
        da = dev_colmajor(a, ld=m + 1)
        dc = dev_colmajor(c, ld=m + 2)
        run_dev(kind, 0.75, da, dev, gens, n, dc, 0.5, 2, 1, prec=api.PREC_SPLIT_I8, slices=t)
        torch.cuda.synchronize()
        errs[t] = ora.rel_fro(host(dc), want)
    print(f"errs={errs}", flush=True)
    assert errs[13] <= 1e-12 and errs[14] <= 1e-12
    assert errs[14] <= max(errs[10], 1e-15)  # more moduli never lose accuracy beyond rounding


@pytest.mark.parametrize("kind", ["cauchy", "toeplitz"])
def test_split_i8_long_contraction(gpu, ora, kind):
    """K = 16600 pads to 16640 > 16384: the direct-to-LDS kernel's residue
    reduction takes its exact (integer remainder) path, and Toeplitz takes the
    streamed-B form (the generator window holds at most 16384 steps)."""
    import torch

    m, n, k = 1100, 1037, 16600
    print(f"case split long kind={kind} m={m} n={n} k={k}", flush=True)
    rng = np.random.default_rng(case_seed("long", kind))
    gens, dev = make_gens(kind, k + 5, rng)
    a = rng.uniform(-1, 1, (m, k))
    want = ora.structured_gemm_np(ORA_KIND[kind], 1.0, a, gens, 0, 0, n)
    da = dev_colmajor(a)
    dc = dev_colmajor(np.zeros((m, n)))
    # t = 15: 46 -> 50 bits per operand at this K (t = 14 leaves Cauchy at ~8e-13)
    run_dev(kind, 1.0, da, dev, gens, n, dc, 0.0, 0, 0, prec=api.PREC_SPLIT_I8, slices=15)
    torch.cuda.synchronize()
    err = ora.rel_fro(host(dc), want)
    print(f"err={err:.3e}", flush=True)
    assert err <= 1e-12


def test_split_i8_zero_rows_and_columns(gpu, ora):
    import torch

    rng = np.random.default_rng(3)
    n = 200
    g = rng.standard_normal(2 * n - 1)
    g[: n // 2] = 0.0  # makes the last columns of H... partly zero
    a = rng.uniform(-1, 1, (n, n))
    a[5, :] = 0.0
    a[17, :] = 0.0
    da = dev_colmajor(a)
    dc = dev_colmajor(np.zeros((n, n)))
    api.toeplitz_dev(1.0, da, torch.from_numpy(g).cuda(), n, n, dc, prec=api.PREC_SPLIT_I8, slices=14)
    torch.cuda.synchronize()
    want = ora.structured_gemm_np(1, 1.0, a, dict(g=g, korder=n), 0, 0, n)
    x = host(dc)
    assert np.all(x[5] == 0.0) and np.all(x[17] == 0.0)
    assert ora.rel_fro(x, want) <= 1e-12


def test_split_i8_wide_dynamic_range_cauchy(gpu, ora):
    # Cauchy entries near the poles reach ~1/(min gap): per-column scaling keeps accuracy
    import torch

    from oracle import inputs

    n = 2048
    b, a = inputs.cauchy_dense_inputs(n, seed=1)
    d, f, u, v = (torch.from_numpy(np.ascontiguousarray(b[:, i])).cuda() for i in range(4))
    da = dev_colmajor(a)
    dc = dev_colmajor(np.zeros((n, n)))
    api.cauchy_dev(1.0, da, d, f, u, v, n, dc, prec=api.PREC_SPLIT_I8, slices=14)
    torch.cuda.synchronize()
    gens = dict(d=b[:, 0], f=b[:, 1], u=b[:, 2], v=b[:, 3])
    err = ora.rel_fro(host(dc), ora.structured_gemm_np(0, 1.0, a, gens, 0, 0, n))
    print(f"cauchy_dense_inputs split err {err:.3e}")
    assert err <= 1e-12


def test_split_rejects_bad_slices(gpu):
    import torch

    a = dev_colmajor(np.ones((8, 8)))
    c = dev_colmajor(np.zeros((8, 8)))
    g = torch.ones(15, dtype=torch.float64, device="cuda")
    with pytest.raises(api.PsmmaError):
        api.toeplitz_dev(1.0, a, g, 8, 8, c, prec=api.PREC_SPLIT_I8, slices=1)


@pytest.mark.parametrize("kind", ["cauchy", "toeplitz", "hankel", "cauchyeig"])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (100, 130, 77), (600, 300, 700)])
def test_split_bf16_parity(gpu, ora, kind, m, n, k):
    # s bf16 slices per operand keep ~8s bits: 3 slices ~1e-7, 6 slices <= 1e-12
    import torch

    print(f"case bf16 kind={kind} m={m} n={n} k={k}", flush=True)
    rng = np.random.default_rng(case_seed("b", kind, m, n, k))
    nglob = max(k, n) + 5
    gens, dev = make_gens(kind, nglob, rng)
    a = rng.uniform(-1, 1, (m, k))
    c = rng.standard_normal((m, n))
    want = ora.structured_gemm_np(ORA_KIND[kind], -1.5, a, gens, 1, 3, n, beta=2.0, c=c)
    errs = {}
    for s in (3, 6):
        da = dev_colmajor(a)
        dc = dev_colmajor(c)
        run_dev(kind, -1.5, da, dev, gens, n, dc, 2.0, 1, 3, prec=api.PREC_SPLIT_BF16, slices=s)
        torch.cuda.synchronize()
        errs[s] = ora.rel_fro(host(dc), want)
    print(f"errs={errs}", flush=True)
    assert errs[3] <= 1e-6
    assert errs[6] <= 1e-12


@pytest.mark.parametrize("kind", ["cauchy", "toeplitz", "hankel", "cauchyeig"])
def test_split_bf16_fp32acc_ragged(gpu, ora, kind):
    """1-3 slices run every kept slice pair into one FP32 accumulator on the
    direct-to-LDS bf16 kernel: 1100 x 1037 x 1300 gives 5 x 5 ragged 256 x 256
    tiles and K % 32 != 0.  Error falls by about 2^-8 per slice down to the FP32
    accumulation floor; the sliced operands are exact, so any kept pair that was
    dropped or doubled would show up as an error far above these bounds."""
    import torch

    m, n, k = 1100, 1037, 1300
    print(f"case bf16 fp32acc kind={kind} m={m} n={n} k={k}", flush=True)
    rng = np.random.default_rng(case_seed("bf", kind, m, n, k))
    gens, dev = make_gens(kind, max(k, n) + 5, rng)
    a = rng.uniform(-1, 1, (m, k))
    c = rng.standard_normal((m, n))
    want = ora.structured_gemm_np(ORA_KIND[kind], 0.5, a, gens, 3, 2, n, beta=-1.0, c=c)
    errs = {}
    for s in (1, 2, 3):
        da = dev_colmajor(a, ld=m + 3)
        dc = dev_colmajor(c, ld=m + 1)
        run_dev(kind, 0.5, da, dev, gens, n, dc, -1.0, 3, 2, prec=api.PREC_SPLIT_BF16, slices=s)
        torch.cuda.synchronize()
        errs[s] = ora.rel_fro(host(dc), want)
    print(f"errs={errs}", flush=True)
    assert errs[1] <= 2e-2 and errs[2] <= 1e-4 and errs[3] <= 1e-6
    assert errs[3] < errs[2] < errs[1]
