"""The oracle's truncated pivoted QR (MDGEQP3 / Compress_RRQR restatement),
pinned against LAPACK's DGEQP3 (scipy.linalg.qr with pivoting), against
reconstruction and orthogonality, and through the reference's low-rank ring
branch (faithful mode of the Toeplitz ring)."""
import numpy as np
import pytest
import scipy.linalg as sl

from oracle import inputs


@pytest.mark.parametrize("m,n,seed", [(50, 40, 0), (40, 50, 1), (64, 64, 2), (7, 3, 3),
                                      (1, 5, 4), (5, 1, 5), (200, 120, 6)])
def test_mdgeqp3_matches_lapack_dgeqp3(ora, m, n, seed):
    a = np.random.default_rng(seed).standard_normal((m, n))
    f, jp, tau, rk = ora.mdgeqp3(a)
    q, r, p = sl.qr(a, pivoting=True)
    k = min(m, n)
    assert rk == k
    assert np.array_equal(jp - 1, p)
    scale = abs(r[0, 0])
    assert np.max(np.abs(np.triu(f[:k]) - r[:k])) < 1e-13 * scale * max(m, n)
    # the reflectors rebuild Q: A P = Q R
    qq = np.eye(m)
    for i in range(k - 1, -1, -1):
        v = np.zeros(m)
        v[i] = 1.0
        v[i + 1:] = f[i + 1:, i]
        qq -= tau[i] * np.outer(v, v @ qq)
    assert np.max(np.abs(qq[:, :k] @ np.triu(f[:k]) - a[:, jp - 1])) < 1e-13 * scale * m


def test_mdgeqp3_fixed_columns_lead(ora):
    rng = np.random.default_rng(7)
    a = rng.standard_normal((30, 12))
    jp_in = np.zeros(12, np.int32)
    jp_in[[3, 8]] = 1
    f, jp, tau, _ = ora.mdgeqp3(a, jpvt=jp_in)
    assert list(jp[:2]) == [4, 9]
    # the free part is pivoted QR of the projected remainder: same as DGEQP3 on it
    qf, _ = np.linalg.qr(a[:, [3, 8]], mode="complete")
    rest = np.delete(np.arange(12), [3, 8])
    proj = (qf.T @ a[:, rest])[2:]
    _, r2, p2 = sl.qr(proj, pivoting=True)
    assert np.array_equal(jp[2:] - 1, rest[p2])
    assert np.allclose(np.abs(np.diag(f)[2:]), np.abs(np.diag(r2)), rtol=1e-12)


def _geometric(m, n, decay, rng):
    k = min(m, n)
    u, _ = np.linalg.qr(rng.standard_normal((m, k)))
    v, _ = np.linalg.qr(rng.standard_normal((n, k)))
    s = decay ** np.arange(k)
    return (u * s) @ v.T


@pytest.mark.parametrize("m,n,decay", [(96, 80, 0.5), (80, 96, 0.3), (128, 128, 0.7)])
def test_compress_rrqr_rank_and_reconstruction(ora, m, n, decay):
    a = _geometric(m, n, decay, np.random.default_rng(8))
    rk, q, rpt = ora.compress_rrqr(a)
    _, r, _ = sl.qr(a, pivoting=True)
    d = np.abs(np.diag(r))
    want = int(np.argmax(d <= ora.RRQR_TOL * d[0])) if (d <= ora.RRQR_TOL * d[0]).any() else len(d)
    assert rk == want and 0 < rk < min(m, n)
    assert np.max(np.abs(q.T @ q - np.eye(rk))) < 1e-14
    # the error of a truncated RRQR is bounded by the dropped part of R
    err = np.linalg.norm(a - q @ rpt, 2)
    assert err <= 4.0 * np.sqrt(n) * d[rk] + 1e-15 * d[0]
    # R P^T keeps the upper-trapezoidal pattern of R in pivot order
    _, jp, _, _ = ora.mdgeqp3(a)
    assert np.allclose(np.tril(rpt[:, jp - 1], -1), 0.0)


def test_compress_rrqr_tolerance_is_the_default_real_literal(ora):
    assert ora.RRQR_TOL == float(np.float32(1e-13)) and ora.RRQR_TOL != 1e-13


def test_compress_rrqr_full_rank_leaves_h_unwritten(ora):
    a = np.random.default_rng(9).standard_normal((40, 30))
    rk, q, rpt = ora.compress_rrqr(a)
    assert rk == 30 and q is None and rpt is None


def test_compress_rrqr_zero_matrix(ora):
    rk, q, rpt = ora.compress_rrqr(np.zeros((10, 6)))
    assert rk == 0 and q.shape == (10, 0) and rpt.shape == (0, 6)


def test_compress_rrqr_ranks_of_eig_cauchy_blocks(ora):
    """Off-diagonal blocks of the eigenvector matrix: ranks as LAPACK's R gives them."""
    n, leaf = 512, 128
    e = inputs.eig_merge_inputs(n, seed=3, with_q=False)
    zh = inputs.loewner_z(e["f"], e["u"], e["difl"], e["difr"])
    v = inputs.column_normalisers(e["f"], zh, e["difl"], e["difr"])
    c = ora.cauchyeig_matrix(e["f"], zh, v, e["difl"], e["difr"], 0, n, 0, n)
    tol = 1e-12
    for bi in range(n // leaf):
        for bj in range(n // leaf):
            if bi == bj:
                continue
            blk = c[bi * leaf:(bi + 1) * leaf, bj * leaf:(bj + 1) * leaf]
            rk, q, rpt = ora.compress_rrqr(blk, tol)
            _, r, _ = sl.qr(blk, pivoting=True)
            d = np.abs(np.diag(r))
            want = int(np.argmax(d <= tol * d[0]))
            assert rk == want, (bi, bj)
            assert np.linalg.norm(blk - q @ rpt) <= 1e-10 * np.linalg.norm(blk)


def test_faithful_toeplitz_ring_uses_rrqr_branch(ora):
    """A smooth (numerically low-rank off the diagonal) Toeplitz operator on a 1 x 4
    grid: the off-diagonal steps take the RRQR branch and the product stays within
    the compression tolerance of the dense one."""
    n = 256
    k = np.arange(-(n - 1), n)
    g = np.exp(-(k / 40.0) ** 2)
    a = np.random.default_rng(10).standard_normal((n, n))
    dense, s0 = ora.ps_compute(ora.TOEPLITZ, 1, 4, 16, 1.0, a, g, mode=ora.MODE_DENSE)
    faith, s1 = ora.ps_compute(ora.TOEPLITZ, 1, 4, 16, 1.0, a, g, mode=ora.MODE_FAITHFUL)
    assert s1[1] > 0  # low-rank steps happened
    assert ora.rel_fro(faith, dense) < 1e-11
    # a random (full-rank) generator: every off-diagonal tile is full rank and the
    # explicit tile is regenerated, so the result is the dense one
    g2 = np.random.default_rng(11).standard_normal(2 * n - 1)
    d2, _ = ora.ps_compute(ora.TOEPLITZ, 1, 4, 16, 1.0, a, g2, mode=ora.MODE_DENSE)
    f2, st2 = ora.ps_compute(ora.TOEPLITZ, 1, 4, 16, 1.0, a, g2, mode=ora.MODE_FAITHFUL)
    assert st2[2] > 0 and st2[1] == 0
    assert ora.rel_fro(f2, d2) < 1e-14


def test_faithful_hankel_ring_is_dense(ora):
    n = 64
    g = np.random.default_rng(12).standard_normal(2 * n - 1)
    a = np.random.default_rng(13).standard_normal((n, n))
    _, st = ora.ps_compute(ora.HANKEL, 2, 2, 8, 1.0, a, g, mode=ora.MODE_FAITHFUL)
    assert st[1] == 0 and st[2] == 0 and st[0] > 0
