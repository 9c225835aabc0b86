"""Batched generator-only RRLU of plain Cauchy blocks on the GPU
(psmma_rrlu_cauchy_batched_dev) against the oracle's compress_cauchy('R') /
rrluCauchy restatement.

Bar: bitwise, as for the eigen-Cauchy kernel (tests/test_gpu_rrlu.py): the rank,
X and Y must be identical block for block.  The oracle restatement is pinned in
tests/test_oracle_rrlu.py (exact deficiency, separated poles within the
tolerance, full rank leaving H unwritten, the faithful Cauchy ring).
"""
import numpy as np
import pytest

from oracle import inputs
from psmma_git_amd import api

pytestmark = pytest.mark.gpu


def _run(gens, blocks, tol, kmax):
    """gens: host (d, f, u, v) global vectors; blocks: (rows, cols) 0-based, one shape."""
    import torch

    d, f, u, v = (torch.from_numpy(np.ascontiguousarray(x, np.float64)).cuda() for x in gens)
    pl = torch.from_numpy(np.stack([r + 1 for r, _ in blocks]).astype(np.int32)).cuda()
    pu = torch.from_numpy(np.stack([c + 1 for _, c in blocks]).astype(np.int32)).cuda()
    rank, x, y = api.rrlu_cauchy_batched_dev(pl, pu, d, f, u, v, tol, kmax)
    torch.cuda.synchronize()
    return rank.cpu().numpy(), x.cpu().numpy(), y.cpu().numpy()


def _check(ora, gens, blocks, tol, kmax=256):
    d, f, u, v = gens
    rank, x, y = _run(gens, blocks, tol, kmax)
    ranks = []
    for b, (rows, cols) in enumerate(blocks):
        rk, xo, yo, _ = ora.compress_cauchy(d[rows], f[cols], u[rows], v[cols], tol)
        assert rank[b] == rk, (b, rank[b], rk)
        if 0 < rk < len(rows) and rk <= kmax:
            assert np.array_equal(x[b].T[:, :rk], xo), b
            assert np.array_equal(y[b].T[:rk, :], yo), b
        ranks.append(rk)
    return np.array(ranks)


@pytest.mark.parametrize("tol", [api.RRLU_TOL, 1e-12])
def test_rrlu_cauchy_leaf_blocks(gpu, ora, tol):
    """Off-diagonal 256-blocks of the interlaced Cauchy generators of the Cauchy
    benchmark (d sorted, f between consecutive d, u and v Gaussian)."""
    n, leaf = 1024, 256
    b = inputs.cauchy_generators(n, np.random.default_rng(4))
    gens = (b[:, 0], b[:, 1], b[:, 2], b[:, 3])
    nb = n // leaf
    blocks = [(np.arange(i * leaf, (i + 1) * leaf), np.arange(j * leaf, (j + 1) * leaf))
              for i in range(nb) for j in range(nb) if i != j]
    print(f"rrlu cauchy leaf blocks tol={tol}", flush=True)
    ranks = _check(ora, gens, blocks, tol)
    assert (ranks > 0).all() and (ranks < leaf).all()


@pytest.mark.parametrize("tol", [1e-6, 1e-10, 1e-14])
def test_rrlu_cauchy_separated_poles(gpu, ora, tol):
    rng = np.random.default_rng(1)
    m, n = 80, 70
    d = np.concatenate([np.sort(rng.random(m)), np.zeros(n)])
    f = np.concatenate([np.zeros(m), 1.05 + np.sort(rng.random(n))])
    u = np.concatenate([rng.standard_normal(m), np.zeros(n)])
    v = np.concatenate([np.zeros(m), rng.standard_normal(n)])
    ranks = _check(ora, (d, f, u, v), [(np.arange(m), np.arange(m, m + n))], tol)
    assert 0 < ranks[0] < m
    if tol == 1e-14:
        assert ranks[0] >= 10  # the swap pass of step 10 has run


def test_rrlu_cauchy_exact_deficiency_and_full_rank(gpu, ora):
    rng = np.random.default_rng(0)
    m, n = 40, 30
    b = inputs.cauchy_generators(m, rng)
    d, f, u, v = b[:, 0], b[:, 1], b[:, 2].copy(), b[:, 3]
    full = _check(ora, (d, f, u, v), [(np.arange(m), np.arange(n))], api.RRLU_TOL)
    assert full[0] == n  # rank n < m: written
    u[rng.choice(m, 12, replace=False)] = 0.0
    assert _check(ora, (d, f, u, v), [(np.arange(m), np.arange(n))], api.RRLU_TOL)[0] == 28
    # square and interlaced: rank m, outputs left unwritten
    sq = inputs.cauchy_generators(32, np.random.default_rng(2))
    gens = (sq[:, 0], sq[:, 1], sq[:, 2], sq[:, 3])
    assert _check(ora, gens, [(np.arange(32), np.arange(32))], api.RRLU_TOL)[0] == 32


def test_rrlu_cauchy_block_cyclic_ring_tiles(gpu, ora):
    """The index sets of the Cauchy benchmark's off-diagonal ring steps (block-cyclic)."""
    n = 200
    b = inputs.cauchy_generators(n, np.random.default_rng(6))
    gens = (b[:, 0], b[:, 1], b[:, 2], b[:, 3])
    blocks = [(ora.constindex(n, nb, 0, 4) - 1, ora.constindex(n, nb, 2, 4) - 1)
              for nb in (5, 10, 25)]
    _check(ora, gens, blocks, api.RRLU_TOL)
    _check(ora, gens, blocks, 1e-10)


def test_rrlu_cauchy_tiny_and_degenerate_blocks(gpu, ora):
    b = inputs.cauchy_generators(64, np.random.default_rng(3))
    d, f, u, v = b[:, 0], b[:, 1], b[:, 2].copy(), b[:, 3]
    for rows, cols in [([3], [40]), ([3], [30, 31, 32, 33, 34]), ([0, 1, 2, 3, 4], [40]),
                       ([0, 1], [50, 51])]:
        _check(ora, (d, f, u, v), [(np.array(rows), np.array(cols))], api.RRLU_TOL)
    u[:8] = 0.0  # an all-zero block: full rank, as the reference reports it
    assert _check(ora, (d, f, u, v), [(np.arange(8), np.arange(40, 48))], api.RRLU_TOL)[0] == 8


def test_rrlu_cauchy_rejects_bad_arguments(gpu):
    L = api.lib()
    assert L.psmma_rrlu_cauchy_batched_dev(1, 300, 4, None, None, None, None, None, None, 8, 8,
                                           1e-17, 4, None, None, 300, 0, None, 4, 0,
                                           None) == -2
    assert L.psmma_rrlu_cauchy_batched_dev(0, 4, 4, None, None, None, None, None, None, 8, 8,
                                           -1.0, 4, None, None, 4, 16, None, 4, 16,
                                           None) == -12
    b = inputs.cauchy_generators(16, np.random.default_rng(7))
    rank, _, _ = _run((b[:, 0], b[:, 1], b[:, 2], b[:, 3]),
                      [(np.array([0, 1, 2, 16]), np.arange(4, 8))], 1e-12, 4)
    assert rank[0] == -2
