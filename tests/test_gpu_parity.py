"""Parity of the HIP kernels with the oracle, through the C ABI (MI355X only).

Bar: plain FP64 results within 1e-13 relative Frobenius error of the FP64
oracle (explicit structured matrix x FP64 GEMM); the known-answer forms of the
reference drivers hold on the GPU path as well.
"""
import numpy as np
import pytest

from oracle import inputs
from psmma_git_amd import api

pytestmark = pytest.mark.gpu

TOL = 1e-13


def case_seed(*key):
    """A seed fixed by the test case (str hashes vary between processes)."""
    import zlib

    return zlib.crc32(repr(key).encode())
KINDS = ["cauchy", "toeplitz", "hankel", "cauchyeig"]


def dev_colmajor(x, ld=None):
    """Upload a numpy matrix as a column-major device tensor with leading dim ld."""
    import torch

    m, n = x.shape
    ld = ld or max(1, m)
    buf = torch.zeros((n, ld), dtype=torch.float64 if x.dtype != np.complex128 else torch.complex128,
                      device="cuda")
    view = buf[:, :m].t()
    view.copy_(torch.from_numpy(np.ascontiguousarray(x)))
    return view


def host(t):
    return t.cpu().numpy()


def make_gens(kind, nglob, rng):
    import torch

    if kind == "cauchy":
        b = inputs.cauchy_generators(nglob, rng)
        gens = dict(d=b[:, 0].copy(), f=b[:, 1].copy(), u=b[:, 2].copy(), v=b[:, 3].copy())
    elif kind in ("toeplitz", "hankel"):
        gens = dict(g=rng.standard_normal(2 * nglob - 1), korder=nglob)
    else:
        e = inputs.eig_merge_inputs(nglob, seed=int(rng.integers(1 << 30)), with_q=False)
        gens = dict(f=e["f"], u=e["u"], v=e["v"], difl=e["difl"], difr=e["difr"])
    dev = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in gens.items()
           if isinstance(v, np.ndarray)}
    return gens, dev


def run_dev(kind, alpha, da, dev, gens, n, dc, beta, r0, c0, prec=api.PREC_FP64, slices=7):
    if kind == "cauchy":
        api.cauchy_dev(alpha, da, dev["d"], dev["f"], dev["u"], dev["v"], n, dc, beta, r0, c0,
                       prec, slices)
    elif kind == "toeplitz":
        api.toeplitz_dev(alpha, da, dev["g"], gens["korder"], n, dc, beta, r0, c0, prec, slices)
    elif kind == "hankel":
        api.hankel_dev(alpha, da, dev["g"], n, dc, beta, r0, c0, prec, slices)
    else:
        api.cauchyeig_dev(alpha, da, dev["f"], dev["u"], dev["v"], dev["difl"], dev["difr"], n, dc,
                          beta, r0, c0, prec, slices)


ORA_KIND = dict(cauchy=0, toeplitz=1, hankel=2, cauchyeig=3)


@pytest.mark.parametrize("kind", KINDS)
# (2050, 2100, 2013): 289 tiles of 128 x 128 select the large tiles (64 x 256 with BK = 32 for
# Toeplitz / Hankel, the 8-wave 128 x 128 with BK = 32 for the Cauchy kinds) with ragged M, N and K
# (K mod 32 = 29, K mod 4 = 1); (4100, 4200, 1001): 1089 tiles select the 256 x 128 tiles of the
# Cauchy kinds (16 waves for plain Cauchy, 8 for eigen-Cauchy)
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (63, 63, 63), (64, 64, 64), (257, 129, 300),
                                   (130, 520, 77), (1024, 1024, 1024), (2050, 2100, 2013),
                                   (4100, 4200, 1001)])
def test_right_fp64_parity(gpu, ora, kind, m, n, k):
    import torch

    print(f"case kind={kind} m={m} n={n} k={k}", flush=True)
    rng = np.random.default_rng(case_seed(kind, m, n, k))
    nglob = max(k, n) + 5
    gens, dev = make_gens(kind, nglob, rng)
    r0, c0 = (3, 2) if nglob > max(k, n) + 3 else (0, 0)
    a = rng.uniform(-1, 1, (m, k))
    c = rng.standard_normal((m, n))
    alpha, beta = 1.25, -0.5
    da = dev_colmajor(a, ld=m + 3)          # odd leading dimensions
    dc = dev_colmajor(c, ld=m + 1)
    run_dev(kind, alpha, da, dev, gens, n, dc, beta, r0, c0)
    torch.cuda.synchronize()
    want = ora.structured_gemm_np(ORA_KIND[kind], alpha, a, gens, r0, c0, n, beta=beta, c=c)
    err = ora.rel_fro(host(dc), want)
    print(f"rel_fro={err:.3e}", flush=True)
    assert err < TOL


@pytest.mark.parametrize("kind", KINDS)
def test_beta_zero_ignores_nan_output(gpu, ora, kind):
    import torch

    rng = np.random.default_rng(2)
    n = 96
    gens, dev = make_gens(kind, n, rng)
    a = rng.uniform(-1, 1, (n, n))
    da = dev_colmajor(a)
    dc = dev_colmajor(np.full((n, n), np.nan))
    run_dev(kind, 1.0, da, dev, gens, n, dc, 0.0, 0, 0)
    torch.cuda.synchronize()
    want = ora.structured_gemm_np(ORA_KIND[kind], 1.0, a, gens, 0, 0, n)
    assert ora.rel_fro(host(dc), want) < TOL


def test_toeplitz_kat_n2_plus_skew(gpu, ora):
    # Test/psmma_testtoeplitz.f90:39-46 generator: A H = N^2 A + (A r)1' - (A 1) r'
    import torch

    n = 1024
    g = np.array([n - i for i in range(1, 2 * n)], dtype=np.float64)
    g[n - 1] = float(n * n)
    rng = np.random.default_rng(9)
    a = rng.uniform(-1, 1, (n, n))
    r = np.arange(1, n + 1, dtype=np.float64)
    want = n * n * a + np.outer(a @ r, np.ones(n)) - np.outer(a @ np.ones(n), r)
    da = dev_colmajor(a)
    dc = dev_colmajor(np.zeros((n, n)))
    api.toeplitz_dev(1.0, da, torch.from_numpy(g).cuda(), n, n, dc)
    torch.cuda.synchronize()
    assert ora.rel_fro(host(dc), want) < TOL


def test_hankel_equals_reversed_toeplitz_large(gpu):
    # size-independent property at a config size: A*Hankel(h) = (A*Toeplitz(rev h)) J
    import torch

    n = 4096
    rng = np.random.default_rng(10)
    h = torch.from_numpy(rng.standard_normal(2 * n - 1)).cuda()
    a = dev_colmajor(rng.uniform(-1, 1, (n, n)))
    x1 = dev_colmajor(np.zeros((n, n)))
    x2 = dev_colmajor(np.zeros((n, n)))
    api.hankel_dev(1.0, a, h, n, x1)
    api.toeplitz_dev(1.0, a, torch.flip(h, [0]).contiguous(), n, n, x2)
    torch.cuda.synchronize()
    d = (x1 - torch.flip(x2, [1])).norm() / x1.norm()
    assert d.item() < 1e-15


# ---------------------------------------------------------------------------
# DFT (left-structured, complex128)

@pytest.mark.parametrize("m,n,k,nd,r0,c0", [(64, 8, 64, 64, 0, 0), (256, 37, 256, 256, 0, 0),
                                           (100, 20, 90, 1000, 17, 301), (1, 1, 1, 5, 2, 3),
                                           (512, 64, 512, 512, 0, 0)])
def test_dft_parity(gpu, ora, m, n, k, nd, r0, c0):
    import torch

    print(f"case dft m={m} n={n} k={k} nd={nd}", flush=True)
    rng = np.random.default_rng(m + n + k)
    b = rng.standard_normal((k, n)) + 1j * rng.standard_normal((k, n))
    c = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
    db = dev_colmajor(b, ld=k + 1)
    dc = dev_colmajor(c, ld=m + 3)
    api.dft_dev(0.75, db, nd, m, dc, beta=0.5, r0=r0, c0=c0)
    torch.cuda.synchronize()
    want = 0.75 * (ora.dft_matrix(nd, r0, m, c0, k) @ b) + 0.5 * c
    assert ora.rel_fro(host(dc), want) < TOL


def test_dft_config_shape_vs_fft(gpu, ora):
    # 4096 x 512 complex: F/sqrt(N) * B against the FFT of the same columns
    import torch

    n, ncols = 4096, 512
    b = inputs.dft_inputs(n, ncols, seed=0)
    db = dev_colmajor(b)
    dc = dev_colmajor(np.zeros((n, ncols), complex))
    api.dft_dev(1.0 / np.sqrt(n), db, n, n, dc)
    torch.cuda.synchronize()
    ref = torch.fft.fft(torch.from_numpy(b).cuda(), dim=0) / np.sqrt(n)
    x = dc
    err = ((x - ref).norm() / ref.norm()).item()
    assert err < TOL


def test_dft_times_ones_kat_gpu(gpu):
    import torch

    n = 4800 // 2
    db = dev_colmajor(np.ones((n, 4), complex))
    dc = dev_colmajor(np.zeros((n, 4), complex))
    api.dft_dev(1.0, db, n, n, dc)
    torch.cuda.synchronize()
    x = host(dc)
    want = np.zeros((n, 4), complex)
    want[0] = n
    assert np.max(np.abs(x - want)) < 1e-12 * n


# ---------------------------------------------------------------------------
# Fortran drop-ins with host arrays.

def test_fortran_toeplitz_entry(gpu, ora):
    n, k = 300, 300
    rng = np.random.default_rng(12)
    lda, ldq = n + 5, n + 2
    a = np.asfortranarray(rng.uniform(-1, 1, (lda, k)))
    g = rng.standard_normal(2 * k - 1)
    q = np.asfortranarray(rng.standard_normal((ldq, k)))
    q0 = q.copy()
    desca = api.descinit(n, k, 64, 64, lda)
    descq = api.descinit(n, k, 64, 64, ldq)
    api.psmma_toeplitz(n, k, 2.0, a, 1, 1, desca, g, 2 * k - 1, 0.5, q, 1, 1, descq)
    assert api.last_status() == 0
    want = ora.structured_gemm_np(1, 2.0, a[:n], dict(g=g, korder=k), 0, 0, k, 0.5, q0[:n])
    assert ora.rel_fro(q[:n], want) < TOL
    assert np.array_equal(q[n:], q0[n:])  # rows beyond N untouched


def test_fortran_redist_writes_at_iq_jq(gpu, ora):
    # REDIST=.true.: A read from (1,1), result redistributed to Q(IQ,JQ)
    n, k = 128, 128
    rng = np.random.default_rng(13)
    a = np.asfortranarray(rng.uniform(-1, 1, (n, k)))
    g = rng.standard_normal(2 * k - 1)
    ldq = n + 10
    q = np.zeros((ldq, k + 7), order="F")
    desca = api.descinit(n, k, 32, 32, n)
    descq = api.descinit(n + 10, k + 7, 32, 32, ldq)
    api.psmma_hankel(n, k, 1.0, a, 1, 1, desca, g, 2 * k - 1, 0.0, q, 4, 6, descq, None, True)
    want = ora.structured_gemm_np(2, 1.0, a, dict(g=g), 0, 0, k)
    assert ora.rel_fro(q[3:3 + n, 5:5 + k], want) < TOL
    assert np.all(q[:3] == 0) and np.all(q[:, :5] == 0)


def test_fortran_cauchy_entry_symmetric_toeplitz_and_split_lda(gpu, ora):
    n = 200
    rng = np.random.default_rng(14)
    b = inputs.cauchy_generators(n + 4, rng)  # LDB > N
    a = np.asfortranarray(rng.uniform(-1, 1, (n, n)))
    q = np.zeros((n, n), order="F")
    d = api.descinit(n, n, 16, 16, n)
    api.psmma_cauchy(n, n, 1.0, a, 1, 1, d, b, n + 4, 0.0, q, 1, 1, d)
    gens = dict(d=b[:, 0], f=b[:, 1], u=b[:, 2], v=b[:, 3])
    assert ora.rel_fro(q, ora.structured_gemm_np(0, 1.0, a, gens, 0, 0, n)) < TOL
    # symmetric Toeplitz stored in LDB = K entries (first column bottom-to-top)
    gs = rng.standard_normal(n)
    full = np.concatenate([gs, gs[:-1][::-1]])
    q2 = np.zeros((n, n), order="F")
    api.psmma_toeplitz(n, n, 1.0, a, 1, 1, d, gs, n, 0.0, q2, 1, 1, d)
    want = ora.structured_gemm_np(1, 1.0, a, dict(g=full, korder=n), 0, 0, n)
    assert ora.rel_fro(q2, want) < TOL


def test_fortran_cauchyeig_entry(gpu, ora):
    n = 160
    e = inputs.eig_merge_inputs(n, seed=3, with_q=True)
    q = np.zeros((n + 4, n + 2), order="F")
    d = api.descinit(n, n, 32, 32, n)
    dq = api.descinit(n + 4, n + 2, 32, 32, n + 4)
    lam = e["f"] + e["difl"]
    api.psmma_cauchyeig(n, n, 1, 1.0, e["q"], 1, 1, d, e["f"], lam, e["u"], e["v"], e["difl"],
                        e["difr"], 0.0, q, 2, 3, dq)
    x = q[1:1 + n, 2:2 + n]
    want = ora.structured_gemm_np(3, 1.0, e["q"], e, 0, 0, n)
    assert ora.rel_fro(x, want) < TOL
    assert np.max(np.abs(x.T @ x - np.eye(n))) < 1e-10  # Q*C stays orthogonal


@pytest.mark.parametrize("entry", [api.pzmdftl1, api.pzmdftl])
def test_fortran_dft_entry(gpu, ora, entry):
    n = 192
    rng = np.random.default_rng(15)
    b = np.asfortranarray(rng.standard_normal((n + 1, n)) + 1j * rng.standard_normal((n + 1, n)))
    b0 = b.copy()
    c = np.asfortranarray(np.ones((n, n), complex))
    entry(n, n, n, n, 1.0, b, n + 1, 2.0, c, n)
    assert api.last_status() == 0
    want = ora.dft_matrix(n, 0, n, 0, n) @ b0[:n] + 2.0
    assert ora.rel_fro(c, want) < TOL
    assert np.array_equal(b, b0)  # one process row: no ring shift, B intact


# ---------------------------------------------------------------------------
# The compute-level routines (PS*_COMPUTE) called directly through the C ABI.

@pytest.mark.parametrize("kind", ["cauchy", "toeplitz", "hankel"])
def test_ps_compute_entries(gpu, ora, kind):
    # SRC/pscauchy_compute.f90:2-3 and siblings: C = alpha*A*H + beta*C with local
    # LDA / LDC and the generator array B(LDB, *) as the drivers pass it.
    n = 230
    rng = np.random.default_rng(case_seed("ps_compute", kind))
    lda, ldc = n + 3, n + 6
    a = np.asfortranarray(rng.uniform(-1, 1, (lda, n)))
    c = np.asfortranarray(rng.standard_normal((ldc, n)))
    c0 = c.copy()
    desca = api.descinit(n, n, 32, 32, lda)
    if kind == "cauchy":
        b = inputs.cauchy_generators(n, rng)
        gens = dict(d=b[:, 0], f=b[:, 1], u=b[:, 2], v=b[:, 3])
        api.pscauchy_compute(n, n, 1.5, a, lda, desca, b, n, -0.5, c, ldc)
    else:
        b = rng.standard_normal(2 * n - 1)
        gens = dict(g=b, korder=n)
        fn = api.pstoeplitz_compute if kind == "toeplitz" else api.pshankel_compute
        fn(n, n, 1.5, a, lda, desca, b, 2 * n - 1, -0.5, c, ldc)
    assert api.last_status() == 0
    want = ora.structured_gemm_np(ORA_KIND[kind], 1.5, a[:n], gens, 0, 0, n, -0.5, c0[:n])
    assert ora.rel_fro(c[:n], want) < TOL
    assert np.array_equal(c[n:], c0[n:])  # rows beyond N untouched


def test_pscauchyeig_compute_entry(gpu, ora):
    n = 176
    e = inputs.eig_merge_inputs(n, seed=case_seed("pscauchyeig_compute") % 1000, with_q=True)
    lda, ldc = n + 2, n + 5
    a = np.zeros((lda, n), order="F")
    a[:n] = e["q"]
    c = np.zeros((ldc, n), order="F")
    desca = api.descinit(n, n, 16, 16, lda)
    lam = e["f"] + e["difl"]
    api.pscauchyeig_compute(n, n, 16, 16, 1.0, a, lda, desca, e["f"], lam, e["u"], e["v"],
                            e["difl"], e["difr"], 0.0, c, ldc)
    assert api.last_status() == 0
    want = ora.structured_gemm_np(3, 1.0, e["q"], e, 0, 0, n)
    assert ora.rel_fro(c[:n], want) < TOL


def test_dev_wrappers_check_shapes_before_the_call(gpu):
    import torch

    a = torch.zeros((64, 64), dtype=torch.float64, device="cuda").t()
    g = torch.zeros(127, dtype=torch.float64, device="cuda")
    small = torch.zeros((60, 64), dtype=torch.float64, device="cuda").t()  # 64 x 60
    with pytest.raises(ValueError, match="smaller than"):
        api.toeplitz_dev(1.0, a, g, 64, 64, small)
    d = torch.zeros(63, dtype=torch.float64, device="cuda")  # one short of r0 + k
    with pytest.raises(ValueError, match="needs 64"):
        api.cauchy_dev(1.0, a, d, g, g, g, 64, a.clone())
