"""The oracle's generator-only RRLU of Cauchy-like blocks (rrluCauchy /
compress_cauchy restatement) and the faithful Cauchy ring that uses it.

Pinned by: exact rank deficiency (rows with u_i = 0) found and the row ID
reconstructing the block to rounding level; numerically low-rank blocks
(separated poles) reconstructed within the tolerance; the swap pass (every
10th step) exercised; the reference's tolerance literal; full rank leaving H
unwritten; and, through the ring, the reference's stale-tile product on a 2 x 2
grid recomputed independently in numpy.

The eigen-Cauchy RRLU (rrluCauchyEigrow / Compress_cauchyEig2 restatement) is
pinned on blocks of the eigenvector matrix of a rank-one update (orthogonal to
rounding, an independent check of the tile formula): ranks against the singular
values, reconstruction at each tolerance, exact deficiency from zero rows, the
swap pass, the explicit block at full rank, and the faithful ring against the
dense product."""
import numpy as np
import pytest

from oracle import inputs


def _cauchy(d, f, u, v):
    return (u[:, None] * v[None, :]) / (d[:, None] - f[None, :])


def test_rrlu_finds_exact_rank_deficiency(ora):
    rng = np.random.default_rng(0)
    m, n = 40, 30
    b = inputs.cauchy_generators(m, rng)  # interlaced poles: numerically full rank
    d, u = b[:, 0], b[:, 2].copy()
    f, v = b[:n, 1], b[:n, 3]
    assert ora.compress_cauchy(d, f, u, v)[0] == n
    zero = rng.choice(m, 12, replace=False)
    u[zero] = 0.0
    rk, x, y, _ = ora.compress_cauchy(d, f, u, v)
    c = _cauchy(d, f, u, v)
    assert rk == 28  # min(m - 12, n) rows carry the block
    assert ora.rel_fro(x @ y, c) < 1e-13
    # the zero rows are interpolated by zero
    assert np.all(x[zero] == 0.0)


@pytest.mark.parametrize("tol", [1e-6, 1e-10, 1e-14])
def test_rrlu_separated_poles_within_tolerance(ora, tol):
    rng = np.random.default_rng(1)
    m, n = 80, 70
    d = np.sort(rng.random(m))
    f = 1.05 + np.sort(rng.random(n))
    u = rng.standard_normal(m)
    v = rng.standard_normal(n)
    c = _cauchy(d, f, u, v)
    rk, x, y, _ = ora.compress_cauchy(d, f, u, v, tol)
    assert 0 < rk < m
    sv = np.linalg.svd(c, compute_uv=False)
    assert rk >= int(np.sum(sv > 10 * tol * sv[0]))  # no fewer rows than the spectrum needs
    assert np.linalg.norm(c - x @ y, 2) <= 1e3 * tol * sv[0]
    if tol == 1e-14:
        assert rk >= 10  # the swap pass of step 10 has run


def test_rrlu_full_rank_leaves_h_unwritten(ora):
    rng = np.random.default_rng(2)
    n = 32
    b = inputs.cauchy_generators(n, rng)
    sentinel = np.full(2 * n * n, 7.25)
    rk, x, y, h = ora.compress_cauchy(b[:, 0], b[:, 1], b[:, 2], b[:, 3], h_init=sentinel)
    assert rk == n and x is None
    assert np.all(h == 7.25)


def test_rrlu_tolerance_is_the_default_real_literal(ora):
    assert ora.RRLU_TOL == float(np.float32(1e-17)) and ora.RRLU_TOL != 1e-17


def _local(ora, n, nb, P, Q, r, c):
    rows = ora.constindex(n, nb, r, P) - 1
    cols = ora.constindex(n, nb, c, Q) - 1
    return rows, cols


def test_faithful_cauchy_ring_multiplies_the_stale_tile(ora):
    """Cauchy generators on a 2 x 2 grid: an off-diagonal step whose tile the RRLU
    finds full rank multiplies the tile its previous step left in the work buffer,
    as the reference does (compressed steps use the row ID, ~exact at TOL 1e-17);
    the dense mode gives the true product."""
    rng = np.random.default_rng(3)
    n, nb, P, Q = 64, 8, 2, 2
    b = inputs.cauchy_generators(n, rng)
    a = rng.standard_normal((n, n))
    h = _cauchy(b[:, 0], b[:, 1], b[:, 2], b[:, 3])
    got, st = ora.ps_compute(ora.CAUCHY, P, Q, nb, 1.0, a, b, mode=ora.MODE_FAITHFUL)
    dense, _ = ora.ps_compute(ora.CAUCHY, P, Q, nb, 1.0, a, b, mode=ora.MODE_DENSE)
    assert ora.rel_fro(dense, a @ h) < 1e-14
    want = np.zeros((n, n))
    stale = 0
    for r in range(P):
        for c in range(Q):
            rows, cols = _local(ora, n, nb, P, Q, r, c)
            prev = None
            acc = np.zeros((len(rows), len(cols)))
            for k1 in range(Q):
                ridx = ora.constindex(n, nb, (c + k1) % Q, Q) - 1
                if k1 == 0:  # own block column: intersecting index sets, dense tile
                    tile = h[np.ix_(ridx, cols)]
                else:
                    rk = ora.compress_cauchy(b[ridx, 0], b[cols, 1], b[ridx, 2], b[cols, 3])[0]
                    if rk >= len(ridx):  # full rank: the buffer still holds the last tile
                        tile = prev
                        stale += 1
                    else:
                        tile = h[np.ix_(ridx, cols)]
                acc = acc + a[np.ix_(rows, ridx)] @ tile
                prev = tile
            want[np.ix_(rows, cols)] = acc
    assert stale == st[3] and stale > 0
    assert ora.rel_fro(got, want) < 1e-13
    assert ora.rel_fro(got, dense) > 1e-3  # the quirk changes the result


def test_faithful_cauchy_ring_low_rank_branch(ora):
    """Rows with u_i = 0 make the off-diagonal tiles rank deficient: the ring takes
    the compressed branch and reproduces the true product."""
    rng = np.random.default_rng(4)
    n, nb = 64, 8
    b = inputs.cauchy_generators(n, rng)
    b[rng.choice(n, 40, replace=False), 2] = 0.0
    a = rng.standard_normal((n, n))
    h = _cauchy(b[:, 0], b[:, 1], b[:, 2], b[:, 3])
    got, st = ora.ps_compute(ora.CAUCHY, 2, 2, nb, 1.0, a, b, mode=ora.MODE_FAITHFUL)
    assert st[1] > 0
    assert ora.rel_fro(got, a @ h) < 1e-12


# ---------------------------------------------------------------------------
# Eigen-Cauchy blocks (rrluCauchyEigrow / Compress_cauchyEig2).

def _eig(n, seed):
    e = inputs.eig_merge_inputs(n, seed=seed, with_q=False)
    f, dl, dr = e["f"], e["difl"], e["difr"]
    zh = inputs.loewner_z(f, e["u"], dl, dr)
    v = inputs.column_normalisers(f, zh, dl, dr)
    return f, dl, dr, zh, v


def test_eig_rrlu_blocks_of_an_orthogonal_matrix(ora):
    n = 200
    f, dl, dr, zh, v = _eig(n, 5)
    c = ora.cauchyeig_matrix(f, zh, v, dl, dr, 0, n, 0, n)
    assert np.max(np.abs(c.T @ c - np.eye(n))) < 1e-13
    # interlaced rows / columns: numerically full rank, the explicit block comes back
    rows, cols = np.arange(0, 120, 2), np.arange(1, 120, 2)
    rk, q, y, _ = ora.compress_cauchyeig2(f, dl, dr, zh, v, rows, cols)
    assert rk == 60 and q is None
    assert np.max(np.abs(y - c[np.ix_(rows, cols)]) / np.abs(c[np.ix_(rows, cols)])) < 1e-15


@pytest.mark.parametrize("tol", [1e-6, 1e-10, 1e-14, float(np.float32(1e-17))])
def test_eig_rrlu_separated_blocks_within_tolerance(ora, tol):
    n = 200
    f, dl, dr, zh, v = _eig(n, 5)
    rows, cols = np.arange(0, 80), np.arange(120, 200)
    blk = ora.cauchyeig_matrix(f, zh, v, dl, dr, 0, n, 0, n)[np.ix_(rows, cols)]
    rk, q, y, _ = ora.compress_cauchyeig2(f, dl, dr, zh, v, rows, cols, tol)
    sv = np.linalg.svd(blk, compute_uv=False)
    assert 0 < rk < 80
    assert int(np.sum(sv > 10 * tol * sv[0])) <= rk <= int(np.sum(sv > 0.01 * tol * sv[0])) + 2
    assert np.linalg.norm(blk - q @ y, 2) <= 1e3 * tol * sv[0] + 1e-15 * sv[0]
    # a row ID: Q holds an identity on the chosen rows, Y is those rows of the block
    unit = np.all((q == 0.0) | (q == 1.0), axis=1) & (np.sum(q == 1.0, axis=1) == 1)
    chosen = np.flatnonzero(unit)
    assert len(chosen) == rk
    order = np.argmax(q[chosen], axis=1)
    assert np.array_equal(y[order], blk[chosen])


def test_eig_rrlu_exact_rank_deficiency_and_swap_pass(ora):
    n = 200
    f, dl, dr, zh, v = _eig(n, 5)
    rows, cols = np.arange(0, 120, 2), np.arange(1, 120, 2)
    u = zh.copy()
    zero = np.random.default_rng(0).choice(rows, 12, replace=False)
    u[zero] = 0.0
    rk, q, y, nswap = ora.compress_cauchyeig2(f, dl, dr, u, v, rows, cols)
    blk = ora.cauchyeig_matrix(f, u, v, dl, dr, 0, n, 0, n)[np.ix_(rows, cols)]
    assert rk == 48 and nswap > 0  # the swap pass of step 5 moved a row
    assert ora.rel_fro(q @ y, blk) < 1e-14
    assert np.all(q[np.searchsorted(rows, zero)] == 0.0)


def test_eig_rrlu_block_cyclic_tiles(ora):
    """The index sets of an off-diagonal ring step (block-cyclic, disjoint)."""
    n = 200
    f, dl, dr, zh, v = _eig(n, 5)
    c = ora.cauchyeig_matrix(f, zh, v, dl, dr, 0, n, 0, n)
    swaps = 0
    for nb in (4, 8, 16):
        rows = ora.constindex(n, nb, 0, 4) - 1
        cols = ora.constindex(n, nb, 1, 4) - 1
        blk = c[np.ix_(rows, cols)]
        rk, q, y, nsw = ora.compress_cauchyeig2(f, dl, dr, zh, v, rows, cols)
        swaps += nsw
        sv = np.linalg.svd(blk, compute_uv=False)
        assert rk >= int(np.sum(sv > 1e-15 * sv[0]))
        got = y if q is None else q @ y
        assert np.linalg.norm(blk - got, 2) <= 1e-14 * sv[0]
    assert swaps > 0


@pytest.mark.parametrize("P,Q,nb", [(1, 4, 16), (2, 2, 8), (1, 8, 4), (2, 4, 16)])
def test_faithful_eig_ring_matches_dense(ora, P, Q, nb):
    """The eigen-Cauchy ring's low-rank branch (compressed steps, or the explicit
    tile rebuilt at full rank: no stale tile for this kind) gives the dense product."""
    n = 200
    f, dl, dr, zh, v = _eig(n, 5)
    gen = np.asfortranarray(np.stack([f, f + dl, zh, v, dl, dr], axis=1))
    a = np.random.default_rng(9).standard_normal((n, n))
    c = ora.cauchyeig_matrix(f, zh, v, dl, dr, 0, n, 0, n)
    dense, _ = ora.ps_compute(ora.CAUCHYEIG, P, Q, nb, 1.0, a, gen, mode=ora.MODE_DENSE)
    faith, st = ora.ps_compute(ora.CAUCHYEIG, P, Q, nb, 1.0, a, gen, mode=ora.MODE_FAITHFUL)
    assert st[1] + st[2] > 0 and st[3] == 0
    assert ora.rel_fro(dense, a @ c) < 1e-14
    assert ora.rel_fro(faith, dense) < 1e-14


# The swap pass (every prd-th step) repeats until no interpolation coefficient
# |Z_i W_j / (D_i - D_j)| of the chosen rows exceeds rho.  When the rank is a
# multiple of prd, the last step ran a pass and nothing changed the generators
# after it, so the returned interpolation matrix is bounded by rho; at other
# ranks it is not.  Over a seeded family this pins rho from both sides (no
# coefficient above it; some between the two kinds' constants) and the period
# (ranks that are a multiple of a shorter period are not bounded).
RHO_CAUCHY = 1.12                   # SRC/Cauchylowrank.f90:70-71 (rho = 1.12D0, prd = 10)
RHO_EIG = float(np.float32(1.2))    # SRC/CauchyHssEig.f90:98-99 (rho = 1.2E+0, prd = 5)


def test_rrlu_interpolation_bounded_by_rho_after_the_swap_pass(ora):
    rng = np.random.default_rng(0)
    ranks, maxima = [], []
    for _ in range(150):
        m, n = 60, 50
        d = np.sort(rng.random(m))
        f = 1.0 + rng.uniform(0.01, 0.3) + np.sort(rng.random(n))
        u, v = rng.standard_normal(m), rng.standard_normal(n)
        for tol in (1e-4, 1e-6, 1e-8, 1e-10, 1e-12, 1e-14):
            rk, x, _, _ = ora.compress_cauchy(d, f, u, v, tol)
            if x is not None:
                ranks.append(rk)
                maxima.append(np.abs(x).max())
    ranks, maxima = np.array(ranks), np.array(maxima)
    after_pass = maxima[ranks % 10 == 0]
    assert len(after_pass) >= 40
    assert after_pass.max() <= RHO_CAUCHY * (1 + 1e-12)
    assert after_pass.max() > 1.05                        # the bound is approached, not slack
    assert np.sum(maxima[ranks % 10 == 5] > RHO_CAUCHY) >= 5  # no pass at the 5th step
    assert maxima.max() > 1.3                              # unbounded between passes


def test_eig_rrlu_interpolation_bounded_by_rho_after_the_swap_pass(ora):
    ranks, maxima = [], []
    for seed in range(12):
        n = 240
        f, dl, dr, zh, v = _eig(n, seed)
        for r0, r1, c0, c1 in [(0, 80, 120, 240), (0, 100, 140, 240), (160, 240, 0, 100),
                               (40, 120, 150, 240), (0, 60, 60, 180)]:
            rows, cols = np.arange(r0, r1), np.arange(c0, c1)
            for tol in (1e-4, 1e-6, 1e-8, 1e-10, 1e-12, 1e-14, ora.RRLU_TOL):
                rk, q, _, _ = ora.compress_cauchyeig2(f, dl, dr, zh, v, rows, cols, tol)
                if q is not None:
                    ranks.append(rk)
                    maxima.append(np.abs(q).max())
    ranks, maxima = np.array(ranks), np.array(maxima)
    after_pass = maxima[ranks % 5 == 0]
    assert len(after_pass) >= 40
    assert after_pass.max() <= RHO_EIG * (1 + 1e-12)
    # some coefficients sit between the plain kind's 1.12 and 1.2: the constants differ
    assert np.sum(after_pass > RHO_CAUCHY) >= 3
    assert np.sum(maxima[ranks % 5 != 0] > RHO_EIG) >= 10  # unbounded between passes
