"""Pins the CPU oracle (no GPU).

The reference ships no golden vectors (its drivers print an error norm and
apply no threshold), so the oracle is pinned by the known-answer forms those
drivers' inputs admit (Test/psmma_testtoeplitz.f90:39-46,
Test/psmma_testcauchy.f90:45-53, Test/pzfox_fft2.f90:99), by brute force on tiny
inputs, and by invariants a wrong index, sign or transposition would break.
"""
import numpy as np
import pytest

from oracle import inputs


# ---------------------------------------------------------------------------
# Index maps (SRC/Auxil_2p5D.f90:2-66, :752-792)

@pytest.mark.parametrize("n,nb,npcol", [(1, 1, 1), (10, 3, 2), (64, 16, 4), (1000, 64, 3),
                                        (5 * 16 + 3, 16, 2), (7, 8, 3), (129, 4, 5)])
def test_constindex_is_blockcyclic_ownership(ora, n, nb, npcol):
    # brute force INDXL2G semantics: global index g (1-based) belongs to
    # process column ((g-1) // nb) % npcol; local order = increasing g.
    for mycol in range(npcol):
        want = [g for g in range(1, n + 1) if ((g - 1) // nb) % npcol == mycol]
        got = ora.constindex(n, nb, mycol, npcol).tolist()
        assert got == want
        assert ora.numroc(n, nb, mycol, npcol) == len(want)


def test_is_intersect_reference_sense(ora):
    # TRUE when the sets share no element (it then selects the low-rank branch)
    assert ora.is_intersect([1, 2, 3], [4, 5]) is True
    assert ora.is_intersect([1, 2, 3], [3, 9]) is False
    assert ora.is_intersect([7], [1, 2, 7, 9]) is False
    # only the first min(m,n) elements of the shorter set are probed: a common
    # element beyond that prefix is not seen (reference behaviour)
    assert ora.is_intersect([1, 2, 3, 4], [9, 4, 8]) is False
    assert ora.is_intersect([1, 2], [5, 6, 2]) is False


# ---------------------------------------------------------------------------
# Known answers from the reference drivers' inputs.

def toeplitz_kat_generator(n):
    # Test/psmma_testtoeplitz.f90:39-46: D(I) = N - I (I != N), D(N) = N*N
    g = np.array([n - i for i in range(1, 2 * n)], dtype=np.float64)
    g[n - 1] = float(n * n)
    return g


@pytest.mark.parametrize("n", [1, 2, 17, 64, 200])
def test_toeplitz_kat_closed_form(ora, n):
    # H = N^2 I + S, S(i,j) = i - j  =>  A H = N^2 A + (A r) 1' - (A 1) r'
    rng = np.random.default_rng(n)
    a = rng.uniform(-1, 1, (n + 3, n))
    g = toeplitz_kat_generator(n)
    r = np.arange(1, n + 1, dtype=np.float64)
    want = n * n * a + np.outer(a @ r, np.ones(n)) - np.outer(a @ np.ones(n), r)
    got_np = ora.structured_gemm_np(ora.TOEPLITZ, 1.0, a, dict(g=g, korder=n), 0, 0, n)
    got_c = ora.structured_gemm_c(ora.TOEPLITZ, 1.0, a, g, n, 0, 0, n)
    assert ora.rel_fro(got_np, want) < 1e-14
    assert ora.rel_fro(got_c, want) < 1e-14


@pytest.mark.parametrize("n", [5, 96, 768])
def test_cauchy_interlaced_kat(ora, n):
    # Test/psmma_testcauchy.f90:45-53: D_i - F_j = (2(i-j)+1) g exactly in real arithmetic
    gap = (9.0 - 0.1) / (2 * n)
    i = np.arange(1, n + 1)
    f = 0.1 + (2 * i - 1) * gap
    d = f + gap
    rng = np.random.default_rng(1)
    u, v = rng.random(n), rng.random(n)
    b = np.asfortranarray(np.stack([d, f, u, v], axis=1))
    exact = np.outer(u, v) / (((2 * (i[:, None] - i[None, :]) + 1)) * gap)
    eye = np.eye(n)
    t_c = ora.structured_gemm_c(ora.CAUCHY, 1.0, eye, b, 0, 0, 0, n)
    t_np = ora.cauchy_matrix(d, f, u, v, 0, n, 0, n)
    # d - f suffers one cancellation of O(1) operands; relative error ~ 9 ulp / gap
    tol = 40 * np.finfo(float).eps * 9.0 / gap
    assert np.max(np.abs(t_c - exact) / np.abs(exact)) < tol
    assert np.max(np.abs(t_np - exact) / np.abs(exact)) < tol


@pytest.mark.parametrize("n", [1, 8, 48, 4800 // 16])
def test_dft_times_ones_kat(ora, n):
    # Test/pzfox_fft2.f90:99: B = 1  =>  F B = N e1 in every column
    f = ora.dft_matrix(n, 0, n, 0, n)
    x = f @ np.ones((n, 3))
    want = np.zeros((n, 3), complex)
    want[0, :] = n
    assert np.max(np.abs(x - want)) < 1e-12 * n


def test_dft_matches_fft(ora):
    n = 240
    rng = np.random.default_rng(3)
    b = rng.standard_normal((n, 7)) + 1j * rng.standard_normal((n, 7))
    f = ora.dft_matrix(n, 0, n, 0, n)
    assert ora.rel_fro(f @ b, np.fft.fft(b, axis=0)) < 1e-14
    # sub-block offsets index the global matrix
    assert np.allclose(ora.dft_matrix(n, 17, 30, 5, 40), f[17:47, 5:45], rtol=0, atol=1e-15)


# ---------------------------------------------------------------------------
# Invariants.

def test_hankel_is_column_reversed_toeplitz(ora):
    k, m = 37, 11
    rng = np.random.default_rng(5)
    h = rng.standard_normal(2 * k - 1)
    a = rng.standard_normal((m, k))
    hk = ora.structured_gemm_c(ora.HANKEL, 1.0, a, h, 0, 0, 0, k)
    tp = ora.structured_gemm_c(ora.TOEPLITZ, 1.0, a, h[::-1].copy(), k, 0, 0, k)
    assert np.array_equal(hk[:, ::-1], tp)


def test_toeplitz_product_is_a_correlation(ora):
    # (A H)(i,j) = sum_k A(i,k) g[(K-1+j) - k] = (A(i,:) * g)[K-1+j]: each row of
    # the product is a slice of the linear convolution of a row of A with g
    k, m = 50, 9
    rng = np.random.default_rng(6)
    g = rng.standard_normal(2 * k - 1)
    a = rng.standard_normal((m, k))
    got = ora.structured_gemm_c(ora.TOEPLITZ, 1.0, a, g, k, 0, 0, k)
    L = 4 * k
    conv = np.fft.irfft(np.fft.rfft(a, L, axis=1) * np.fft.rfft(g, L)[None, :], L, axis=1)
    want = conv[:, np.arange(k) + (k - 1)]
    assert ora.rel_fro(got, want) < 1e-13


def test_c_and_numpy_restatements_agree_with_offsets(ora):
    rng = np.random.default_rng(7)
    n = 64
    b = inputs.cauchy_generators(n, rng)
    g = rng.standard_normal(2 * n - 1)
    a = rng.standard_normal((13, 20))
    c = rng.standard_normal((13, 17))
    for kind, gen, gens in [
        (ora.CAUCHY, b, dict(d=b[:, 0], f=b[:, 1], u=b[:, 2], v=b[:, 3])),
        (ora.TOEPLITZ, g, dict(g=g, korder=n)),
        (ora.HANKEL, g, dict(g=g)),
    ]:
        x_np = ora.structured_gemm_np(kind, 0.7, a, gens, 9, 30, 17, beta=-0.3, c=c)
        x_c = ora.structured_gemm_c(kind, 0.7, a, gen, n, 9, 30, 17, beta=-0.3, c=c)
        assert ora.rel_fro(x_c, x_np) < 1e-14, kind


def test_alpha_beta_linearity(ora):
    rng = np.random.default_rng(8)
    n = 40
    g = rng.standard_normal(2 * n - 1)
    a = rng.standard_normal((n, n))
    c = rng.standard_normal((n, n))
    x1 = ora.structured_gemm_c(ora.TOEPLITZ, 1.0, a, g, n, 0, 0, n)
    x = ora.structured_gemm_c(ora.TOEPLITZ, 2.5, a, g, n, 0, 0, n, beta=-1.5, c=c)
    assert ora.rel_fro(x, 2.5 * x1 - 1.5 * c) < 1e-14
    # beta == 0 must not read C (NaNs in C do not propagate)
    cn = np.full((n, n), np.nan)
    x0 = ora.structured_gemm_c(ora.TOEPLITZ, 1.0, a, g, n, 0, 0, n, beta=0.0, c=cn)
    assert np.isfinite(x0).all()


# ---------------------------------------------------------------------------
# The ring loop on emulated grids (dense branch): every grid gives the product.

@pytest.mark.parametrize("grid,nb", [((1, 1), 8), ((2, 2), 8), ((2, 3), 5), ((3, 2), 16),
                                     ((1, 4), 7)])
def test_ring_emulation_dense_equals_product(ora, grid, nb):
    rng = np.random.default_rng(11)
    n = 61
    b = inputs.cauchy_generators(n, rng)
    g = rng.standard_normal(2 * n - 1)
    a = rng.standard_normal((n, n))
    c = rng.standard_normal((n, n))
    for kind, gen, gens in [
        (ora.CAUCHY, b, dict(d=b[:, 0], f=b[:, 1], u=b[:, 2], v=b[:, 3])),
        (ora.TOEPLITZ, g, dict(g=g, korder=n)),
        (ora.HANKEL, g, dict(g=g)),
    ]:
        want = ora.structured_gemm_np(kind, 1.3, a, gens, 0, 0, n, beta=0.5, c=c)
        got, stats = ora.ps_compute(kind, grid[0], grid[1], nb, 1.3, a, gen, beta=0.5, c=c,
                                    mode=ora.MODE_DENSE, nthreads=2)
        assert ora.rel_fro(got, want) < 1e-14, (kind, grid)
        assert stats[0] == grid[0] * grid[1] * grid[1]  # every rank takes every step


def test_ring_emulation_cauchyeig(ora):
    n = 48
    e = inputs.eig_merge_inputs(n, seed=2, with_q=True)
    gen = np.asfortranarray(np.stack([e["f"], e["f"] + e["difl"], e["u"], e["v"], e["difl"],
                                      e["difr"]], axis=1))
    want = ora.structured_gemm_np(ora.CAUCHYEIG, 1.0, e["q"], e, 0, 0, n)
    got, _ = ora.ps_compute(ora.CAUCHYEIG, 2, 2, 8, 1.0, e["q"], gen, mode=ora.MODE_DENSE)
    assert ora.rel_fro(got, want) < 1e-14


# ---------------------------------------------------------------------------
# Eigenvector-Cauchy generators of the merge step: secular roots and orthogonality.

def test_secular_roots_match_dense_eigensolver():
    n = 120
    e = inputs.eig_merge_inputs(n, seed=4, with_q=False)
    d, z = e["f"], e["u"]
    lam = d + e["difl"]
    want = np.linalg.eigvalsh(np.diag(d) + np.outer(z, z))
    assert np.max(np.abs(np.sort(lam) - want)) < 1e-13
    # the two differences describe the same root
    assert np.max(np.abs((d[1:] + e["difr"][:-1]) - lam[:-1])) < 1e-14
    assert (e["difl"] > 0).all() and (e["difr"][:-1] < 0).all()


def test_cauchyeig_columns_are_orthonormal_eigenvectors(ora):
    n = 150
    e = inputs.eig_merge_inputs(n, seed=5, with_q=False)
    c = ora.cauchyeig_matrix(e["f"], e["u"], e["v"], e["difl"], e["difr"], 0, n, 0, n)
    assert np.max(np.abs(np.linalg.norm(c, axis=0) - 1.0)) < 1e-14
    # eigenvector check: (D + z z') c_j = lam_j c_j
    m = np.diag(e["f"]) + np.outer(e["u"], e["u"])
    lam = e["f"] + e["difl"]
    assert np.max(np.abs(m @ c - c * lam[None, :])) < 1e-12
    assert np.max(np.abs(c.T @ c - np.eye(n))) < 1e-10


def test_loewner_z_makes_roots_exact_and_vectors_orthonormal(ora):
    """Loewner's theorem: D + zhat zhat^T has exactly the computed roots as eigenvalues,
    and the eigen-Cauchy columns built from zhat are orthonormal to rounding level."""
    n = 160
    e = inputs.eig_merge_inputs(n, seed=6, with_q=False)
    d, z, difl, difr = e["f"], e["u"], e["difl"], e["difr"]
    zh = inputs.loewner_z(d, z, difl, difr)
    assert np.all(np.sign(zh) == np.sign(z))
    # the computed roots are accurate, so zhat is close to z
    assert np.max(np.abs(zh - z)) < 1e-12
    lam = d + difl
    want = np.linalg.eigvalsh(np.diag(d) + np.outer(zh, zh))
    assert np.max(np.abs(np.sort(lam) - want)) < 1e-14
    v = inputs.column_normalisers(d, zh, difl, difr)
    c = ora.cauchyeig_matrix(d, zh, v, difl, difr, 0, n, 0, n)
    assert np.max(np.abs(c.T @ c - np.eye(n))) < 1e-13
    # a planted sign error in one factor would break the product's magnitude
    bad = inputs.loewner_z(d, z, difl, -difr)
    assert not np.allclose(bad, zh, rtol=1e-6)


def test_column_normalisers_match_explicit_norms(ora):
    n = 90
    e = inputs.eig_merge_inputs(n, seed=7, with_q=False)
    c = ora.cauchyeig_matrix(e["f"], e["u"], np.ones(n), e["difl"], e["difr"], 0, n, 0, n)
    assert np.max(np.abs(e["v"] * np.linalg.norm(c, axis=0) - 1.0)) < 1e-14
