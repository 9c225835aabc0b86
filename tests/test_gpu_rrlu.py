"""Batched generator-only RRLU of eigen-Cauchy blocks on the GPU
(psmma_rrlu_cauchyeig_batched_dev) against the oracle's Compress_cauchyEig2 /
rrluCauchyEigrow restatement.

Bar: bitwise.  Every quantity the factorisation computes is a subtraction, a
quotient, a product or a maximum of values rebuilt from the same generators in
the same order on both sides (no fused multiply-add, IEEE division), so the
rank, the number of swap-pass exchanges, Q and Y must be identical, block for
block.  The oracle restatement itself is pinned in tests/test_oracle_rrlu.py
(reconstruction within the tolerance, singular-value counts, row-ID structure,
exact deficiency, the swap pass, block-cyclic ring tiles).
"""
import numpy as np
import pytest

from oracle import inputs
from psmma_git_amd import api

pytestmark = pytest.mark.gpu


def _eig(n, seed):
    e = inputs.eig_merge_inputs(n, seed=seed, with_q=False)
    f, dl, dr = e["f"], e["difl"], e["difr"]
    zh = inputs.loewner_z(f, e["u"], dl, dr)
    v = inputs.column_normalisers(f, zh, dl, dr)
    return f, dl, dr, zh, v


def _run(gens, blocks, tol, kmax):
    """gens: host (f, difl, difr, u, v); blocks: list of (rows, cols), 0-based, all of
    one shape.  Returns host (rank, nswap, q, y)."""
    import torch

    f, dl, dr, u, v = (torch.from_numpy(np.ascontiguousarray(x, np.float64)).cuda()
                       for x in gens)
    pl = torch.from_numpy(np.stack([r + 1 for r, _ in blocks]).astype(np.int32)).cuda()
    pu = torch.from_numpy(np.stack([c + 1 for _, c in blocks]).astype(np.int32)).cuda()
    rank, nswap, q, y = api.rrlu_cauchyeig_batched_dev(pl, pu, f, u, v, dl, dr, tol, kmax)
    torch.cuda.synchronize()
    return rank.cpu().numpy(), nswap.cpu().numpy(), q.cpu().numpy(), y.cpu().numpy()


def _check(ora, gens, blocks, tol, kmax=256):
    f, dl, dr, u, v = gens
    rank, nswap, q, y = _run(gens, blocks, tol, kmax)
    ranks = []
    for b, (rows, cols) in enumerate(blocks):
        rk, qo, yo, nsw = ora.compress_cauchyeig2(f, dl, dr, u, v, rows, cols, tol)
        assert rank[b] == rk, (b, rank[b], rk)
        assert nswap[b] == nsw, (b, nswap[b], nsw)
        if 0 < rk < min(len(rows), len(cols)) and rk <= kmax:
            qg = q[b].T[:, :rk]
            yg = y[b].T[:rk, :]
            assert np.array_equal(qg, qo), b
            assert np.array_equal(yg, yo), b
        ranks.append(rk)
    return np.array(ranks), nswap


@pytest.mark.parametrize("tol", [api.RRLU_TOL, 1e-12])
def test_rrlu_merge_leaf_blocks(gpu, ora, tol):
    """The off-diagonal leaf blocks of the eigenvector-merge matrix."""
    n, leaf = 1024, 256
    gens = _eig(n, 3)
    nb = n // leaf
    blocks = [(np.arange(i * leaf, (i + 1) * leaf), np.arange(j * leaf, (j + 1) * leaf))
              for i in range(nb) for j in range(nb) if i != j]
    print(f"rrlu merge blocks tol={tol}", flush=True)
    ranks, _ = _check(ora, gens, blocks, tol)
    assert (ranks > 0).all() and (ranks < leaf).all()


@pytest.mark.parametrize("tol", [api.RRLU_TOL, 1e-12])
def test_rrlu_blocks_from_gpu_secular_solver(gpu, ora, tol):
    """The merge's 240 leaf blocks at N = 4096 with the generators of the GPU secular
    solver (as the compressed product uses them), run twice: both runs bitwise equal
    to each other and to the oracle."""
    import torch

    n, leaf = 4096, 256
    d, z, _ = inputs.merge_poles_and_z(n, 2)
    td, tz = torch.from_numpy(d).cuda(), torch.from_numpy(z).cuda()
    difl, difr, u, v = api.secular_dev(td, tz, 1.0, loewner=True)
    torch.cuda.synchronize()
    gens = tuple(x.cpu().numpy() for x in (td, difl, difr, u, v))
    nb = n // leaf
    blocks = [(np.arange(i * leaf, (i + 1) * leaf), np.arange(j * leaf, (j + 1) * leaf))
              for i in range(nb) for j in range(nb) if i != j]
    first = _run(gens, blocks, tol, 32)
    second = _run(gens, blocks, tol, 32)
    for a, b in zip(first, second):
        assert np.array_equal(a, b)
    _check(ora, gens, blocks, tol, kmax=32)


@pytest.mark.parametrize("tol", [1e-6, 1e-10, 1e-14, api.RRLU_TOL])
def test_rrlu_separated_blocks(gpu, ora, tol):
    n = 200
    gens = _eig(n, 5)
    blocks = [(np.arange(0, 80), np.arange(120, 200)), (np.arange(120, 200), np.arange(0, 80)),
              (np.arange(40, 120), np.arange(120, 200))]
    ranks, _ = _check(ora, gens, blocks, tol)
    assert (ranks > 0).all() and (ranks < 80).all()


def test_rrlu_exact_deficiency_and_swap_pass(gpu, ora):
    n = 200
    f, dl, dr, zh, v = _eig(n, 5)
    rows, cols = np.arange(0, 120, 2), np.arange(1, 120, 2)
    u = zh.copy()
    u[np.random.default_rng(0).choice(rows, 12, replace=False)] = 0.0
    ranks, nswap = _check(ora, (f, dl, dr, u, v), [(rows, cols)], api.RRLU_TOL)
    assert ranks[0] == 48 and nswap[0] > 0


def test_rrlu_block_cyclic_ring_tiles(gpu, ora):
    n = 200
    gens = _eig(n, 5)
    blocks = [(ora.constindex(n, nb, 0, 4) - 1, ora.constindex(n, nb, 1, 4) - 1)
              for nb in (5, 10, 25)]  # 50 rows and 50 columns each
    _, nswap = _check(ora, gens, blocks, api.RRLU_TOL)
    assert nswap.sum() > 0


def test_rrlu_full_rank_and_shapes(gpu, ora):
    n = 400
    gens = _eig(n, 7)
    # interlaced rows / columns: full rank, outputs left unwritten
    ranks, _ = _check(ora, gens, [(np.arange(0, 120, 2), np.arange(1, 120, 2))], api.RRLU_TOL)
    assert ranks[0] == 60
    _check(ora, gens, [(np.arange(0, 37), np.arange(200, 290))], 1e-12)
    _check(ora, gens, [(np.arange(144, 400), np.arange(0, 100))], 1e-12)
    # rank cap: outputs only for rank <= kmax, rank reported regardless
    rows, cols = np.arange(0, 80), np.arange(120, 200)
    rank, _, _, _ = _run(gens, [(rows, cols)], api.RRLU_TOL, 4)
    rk, _, _, _ = ora.compress_cauchyeig2(*gens, rows, cols)
    assert rank[0] == rk > 4


def test_rrlu_tiny_and_degenerate_blocks(gpu, ora):
    """1 x 1, 1 x n, m x 1 and 2 x 2 blocks, and an all-zero block: with every pivot zero the
    reference's strict test |u_kk| < tol * max |u| never fires, so the zero block is reported
    at full rank (its explicit block is exact); the quotients it divides by zero run on both
    sides alike."""
    gens = _eig(64, 9)
    for rows, cols in [([3], [40]), ([3], [30, 31, 32, 33, 34]), ([0, 1, 2, 3, 4], [40]),
                       ([0, 1], [50, 51])]:
        _check(ora, gens, [(np.array(rows), np.array(cols))], api.RRLU_TOL)
    f, dl, dr, zh, v = gens
    u0 = zh.copy()
    u0[:8] = 0.0
    ranks, _ = _check(ora, (f, dl, dr, u0, v), [(np.arange(8), np.arange(40, 48))], api.RRLU_TOL)
    assert ranks[0] == 8


def test_rrlu_rejects_bad_arguments(gpu):
    import torch

    L = api.lib()
    assert L.psmma_rrlu_cauchyeig_batched_dev(1, 257, 4, None, None, None, None, None, None,
                                              None, 8, 1e-17, 4, None, None, None, 257, 0,
                                              None, 4, 0, None) == -2
    assert L.psmma_rrlu_cauchyeig_batched_dev(0, 4, 4, None, None, None, None, None, None,
                                              None, 8, 1e-17, 0, None, None, None, 4, 16,
                                              None, 4, 16, None) == -13
    # a global index outside [1, korder]: rank -2, nothing read through it
    gens = _eig(16, 1)
    rank, _, _, _ = _run(gens, [(np.arange(0, 4), np.array([10, 11, 12, 16]))], 1e-12, 4)
    assert rank[0] == -2
