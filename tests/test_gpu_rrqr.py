"""Batched truncated RRQR on the GPU (psmma_rrqr_batched_dev) against the oracle's
Compress_RRQR / MDGEQP3 restatement.

Bar.  The rank equals the oracle's whenever the tolerance separates the diagonal
of R clearly (|R(r-1,r-1)| >= 3 tol |R(0,0)| >= 9 |R(r,r)|; the test inputs are
built so), and is within 2 of it otherwise.  Pivots are equal over the
well-determined prefix (|R(k,k)| > 1e-8 |R(0,0)|: beyond it the partial norms
that choose a pivot carry relative rounding errors of their own size, and
near-ties are decided by rounding differently on each side).  Over that prefix
R P^T rows agree to 1e-12 |R(0,0)| and each column l of Q to
1e-12 |R(0,0)| / |R(l,l)| (a column of Q is as well determined as the residual it
normalises).  Q is orthonormal to 1e-13, and both low-rank products approximate
the block within the truncation bound.  Full rank and rank >= kmax leave the
outputs unwritten, as documented.
"""
import numpy as np
import pytest

from oracle import inputs
from psmma_git_amd import api

pytestmark = pytest.mark.gpu


def _geometric(m, n, decay, rng):
    k = min(m, n)
    u, _ = np.linalg.qr(rng.standard_normal((m, k)))
    v, _ = np.linalg.qr(rng.standard_normal((n, k)))
    return (u * decay ** np.arange(k)) @ v.T


def _run(blocks, tol, kmax=api.RRQR_KMAX):
    import torch

    nb = len(blocks)
    m, n = blocks[0].shape
    t = torch.from_numpy(np.stack([b.T for b in blocks])).cuda().contiguous()
    rank, jpvt, q, rpt = api.rrqr_batched_dev(t, tol, kmax)
    torch.cuda.synchronize()
    return (rank.cpu().numpy(), jpvt.cpu().numpy(), q.cpu().numpy().transpose(0, 2, 1),
            rpt.cpu().numpy().transpose(0, 2, 1))


def _check_against_oracle(ora, blocks, tol, kmax=api.RRQR_KMAX):
    rank, jpvt, q, rpt = _run(blocks, tol, kmax)
    for b, blk in enumerate(blocks):
        m, n = blk.shape
        rk, qo, ro = ora.compress_rrqr(blk, tol)
        f, jo, _, _ = ora.mdgeqp3(blk, tol)
        dg = np.abs(np.diag(f)[:min(m, n)])
        dg = dg / dg[0] if dg[0] > 0 else dg
        clear = rk == min(m, n) or rk == 0 or (dg[rk - 1] >= 3 * tol and dg[rk] <= tol / 3)
        if rk >= kmax and not (rk == min(m, n) <= kmax):
            assert rank[b] == -1, b  # the cap is reached before rank or full rank shows
            continue
        if clear:
            assert rank[b] == rk, (b, rank[b], rk)
        else:
            assert abs(rank[b] - rk) <= 2, (b, rank[b], rk)
        r = rank[b]
        if not 0 < r < min(m, n):
            continue
        qg, rg = q[b][:, :r], rpt[b][:r]
        assert np.max(np.abs(qg.T @ qg - np.eye(r))) < 1e-13, b
        scale = abs(f[0, 0])
        nb = np.linalg.norm(blk, 2)
        bound = 4.0 * np.sqrt(n) * tol * scale + 1e-13 * nb
        assert np.linalg.norm(blk - qg @ rg, 2) <= bound, b
        if 0 < rk < min(m, n):
            assert np.linalg.norm(blk - qo @ ro, 2) <= bound, b
        pre = int(np.sum(dg[:min(r, rk)] > 1e-8))  # well-determined prefix
        assert np.array_equal(jpvt[b][:pre], jo[:pre]), b
        if pre and 0 < rk < min(m, n):
            assert np.max(np.abs(rg[:pre] - ro[:pre])) < 1e-12 * scale, b
            rdiag = np.abs(ro[np.arange(pre), jo[:pre] - 1])
            dq = np.max(np.abs(qg[:, :pre] - qo[:, :pre]), axis=0)
            assert np.all(dq * rdiag <= 1e-12 * scale), (b, dq * rdiag / scale)
    return rank


def _exact_rank(m, n, r, rng, noise=1e-17):
    a = _geometric(m, r, 0.7, rng)[:, :r] @ rng.standard_normal((r, n))
    return a + noise * rng.standard_normal((m, n))


@pytest.mark.parametrize("m,n", [(256, 256), (200, 150), (37, 90), (64, 64)])
def test_rrqr_geometric_blocks(gpu, ora, m, n):
    print(f"rrqr geometric {m}x{n}", flush=True)
    rng = np.random.default_rng(m * 1000 + n)
    blocks = [_geometric(m, n, d, rng) for d in (0.2, 0.35, 0.5, 0.6, 0.4, 0.25)]
    rank = _check_against_oracle(ora, blocks, 1e-12)
    assert (rank != 0).all() and (rank > 0).any()


def test_rrqr_reference_tolerance_and_full_rank(gpu, ora):
    rng = np.random.default_rng(5)
    blocks = [rng.standard_normal((48, 40)), _exact_rank(48, 40, 13, rng), np.zeros((48, 40)),
              _geometric(48, 40, 0.3, rng)]
    rank = _check_against_oracle(ora, blocks, ora.RRQR_TOL)
    assert list(rank[:3]) == [-1, 13, 0]  # min(m,n) = 40 > kmax: the cap shows first


@pytest.mark.parametrize("m,n,r", [(256, 256, 20), (256, 100, 7), (90, 256, 31), (128, 64, 1)])
def test_rrqr_exact_rank_blocks(gpu, ora, m, n, r):
    rng = np.random.default_rng(r)
    blocks = [_exact_rank(m, n, r, rng) for _ in range(5)]
    rank = _check_against_oracle(ora, blocks, 1e-12)
    assert (rank == r).all()


def test_rrqr_rank_cap(gpu, ora):
    rng = np.random.default_rng(6)
    blocks = [_geometric(100, 100, 0.75, rng), _geometric(100, 100, 0.005, rng)]
    rank, _, _, _ = _run(blocks, 1e-12, kmax=8)
    assert rank[0] == -1 and 0 < rank[1] < 8
    _check_against_oracle(ora, blocks, 1e-12, kmax=8)


def test_rrqr_eig_cauchy_off_diagonal_blocks(gpu, ora):
    """The compression the eigenvector-merge variant relies on: off-diagonal blocks of
    the eigen-Cauchy matrix at tolerance 1e-12."""
    n, leaf = 1024, 256
    e = inputs.eig_merge_inputs(n, seed=3, with_q=False)
    zh = inputs.loewner_z(e["f"], e["u"], e["difl"], e["difr"])
    v = inputs.column_normalisers(e["f"], zh, e["difl"], e["difr"])
    c = ora.cauchyeig_matrix(e["f"], zh, v, e["difl"], e["difr"], 0, n, 0, n)
    blocks = [np.ascontiguousarray(c[i * leaf:(i + 1) * leaf, j * leaf:(j + 1) * leaf])
              for i in range(n // leaf) for j in range(n // leaf) if i != j]
    rank = _check_against_oracle(ora, blocks, 1e-12)
    assert (rank > 0).all() and (rank < 32).all()


def test_rrqr_rejects_bad_arguments(gpu):
    L = api.lib()
    assert L.psmma_rrqr_batched_dev(1, 257, 4, None, 257, 0, 1e-12, 8, None, None, None, 1, 0,
                                    None, 1, 0, None) == -2
    assert L.psmma_rrqr_batched_dev(1, 4, 4, None, 4, 16, 1e-12, 33, None, None, None, 4, 0,
                                    None, 1, 0, None) == -4
    assert L.psmma_rrqr_batched_dev(0, 4, 4, None, 4, 16, 1e-12, 33, None, None, None, 4, 0,
                                    None, 1, 0, None) == -8
