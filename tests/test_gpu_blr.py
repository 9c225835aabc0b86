"""Compressed eigenvector-merge product (psmma_cauchyeig_blr_dev): off-diagonal
leaf blocks of the eigen-Cauchy matrix compressed by the batched RRQR, against
the FP64 oracle product (explicit C, numpy GEMM).

Bar: relative Frobenius error <= 20 * tol (each block is approximated within
~tol of its largest R entry, the blocks' errors add in quadrature); the summed
RRQR ranks within 2% of the oracle's Compress_RRQR ranks over the same blocks
(near-tolerance pivots may differ by rounding, see the RRQR tests); columns of
Q * C orthonormal to 1e-11 when Q is orthogonal.
"""
import numpy as np
import pytest

from oracle import inputs
from psmma_git_amd import api

pytestmark = pytest.mark.gpu


def _setup(n, seed):
    import torch

    d, z, rng = inputs.merge_poles_and_z(n, seed)
    td, tz = torch.from_numpy(d).cuda(), torch.from_numpy(z).cuda()
    difl, difr, u, v = api.secular_dev(td, tz, 1.0, loewner=True)
    torch.cuda.synchronize()
    gens = [td, u, v, difl, difr]
    return gens, [t.cpu().numpy() for t in gens], rng


def _dev(x):
    import torch

    return torch.from_numpy(np.asfortranarray(x)).cuda().t().contiguous().t()


@pytest.mark.parametrize("n,leaf,m", [(2048, 256, 2048), (2048, 128, 700), (1024, 256, 1024),
                                      (1024, 256, 333)])
def test_blr_product_matches_oracle(gpu, ora, n, leaf, m):
    import torch

    print(f"blr n={n} leaf={leaf} m={m}", flush=True)
    tol = 1e-12
    gens, gh, rng = _setup(n, 31)
    if m == n:
        a, _ = np.linalg.qr(rng.standard_normal((n, n)))
    else:
        a = rng.standard_normal((m, n))
    x0 = rng.standard_normal((m, n))
    da, dx = _dev(a), _dev(x0)
    st = api.cauchyeig_blr_dev(1.3, da, *gens, dx, leaf=leaf, tol=tol, beta=0.5)
    torch.cuda.synchronize()
    c = ora.cauchyeig_matrix(*gh, 0, n, 0, n)
    want = 1.3 * (a @ c) + 0.5 * x0
    got = dx.cpu().numpy()
    err = ora.rel_fro(got, want)
    print("rel err", err, st, flush=True)
    assert err <= 20 * tol
    nb = n // leaf
    assert st["blocks"] == nb * (nb - 1)
    assert st["lowrank"] + st["dense"] + st["zero"] == st["blocks"]
    assert st["lowrank"] > 0.9 * st["blocks"]
    ranks = 0
    for i in range(nb):
        for j in range(nb):
            if i != j:
                blk = c[i * leaf:(i + 1) * leaf, j * leaf:(j + 1) * leaf]
                r, _, _ = ora.compress_rrqr(blk, tol)
                ranks += r if r < min(32, leaf) else 0
    assert abs(st["rank_sum"] - ranks) <= 0.02 * ranks


@pytest.mark.parametrize("n,leaf,m", [(2048, 256, 2048), (2048, 128, 700)])
def test_blr_rrlu_compressor_matches_oracle(gpu, ora, n, leaf, m):
    """The same product with the generator-only RRLU (Compress_cauchyEig2) as the
    compressor: error within the same bar, and, the GPU RRLU being bitwise the
    oracle's, the same low-rank blocks and rank sum as the oracle's RRLU."""
    import torch

    print(f"blr rrlu n={n} leaf={leaf} m={m}", flush=True)
    tol = 1e-12
    gens, gh, rng = _setup(n, 31)
    a = rng.standard_normal((m, n))
    x0 = rng.standard_normal((m, n))
    da, dx = _dev(a), _dev(x0)
    st = api.cauchyeig_blr_dev(1.3, da, *gens, dx, leaf=leaf, tol=tol, beta=0.5,
                               compressor=api.BLR_RRLU)
    torch.cuda.synchronize()
    c = ora.cauchyeig_matrix(*gh, 0, n, 0, n)
    err = ora.rel_fro(dx.cpu().numpy(), 1.3 * (a @ c) + 0.5 * x0)
    print("rel err", err, st, flush=True)
    assert err <= 20 * tol
    f, u, v, dl, dr = gh
    nb = n // leaf
    low, ranks = 0, 0
    for i in range(nb):
        for j in range(nb):
            if i != j:
                r, _, _, _ = ora.compress_cauchyeig2(f, dl, dr, u, v,
                                                     np.arange(i * leaf, (i + 1) * leaf),
                                                     np.arange(j * leaf, (j + 1) * leaf), tol)
                if 0 < r <= 32 and r < leaf:
                    low += 1
                    ranks += r
    assert st["lowrank"] == low and st["rank_sum"] == ranks


@pytest.mark.parametrize("compressor", [1, 0])
def test_cauchy_blr_product_matches_oracle(gpu, ora, compressor):
    """The compressed form of the plain Cauchy product (the Cauchy benchmark's
    interlaced generators): error within 20 tol; with the RRLU the same low-rank blocks
    and rank sum as the oracle's compress_cauchy, with the RRQR the rank sum within 2%
    of the oracle's Compress_RRQR."""
    import torch

    n, leaf, m, tol = 2048, 256, 700, 1e-12
    rng = np.random.default_rng(8)
    b = inputs.cauchy_generators(n, rng)
    d, f, u, v = (b[:, i].copy() for i in range(4))
    a = rng.standard_normal((m, n))
    x0 = rng.standard_normal((m, n))
    da, dx = _dev(a), _dev(x0)
    gd = [torch.from_numpy(np.ascontiguousarray(x)).cuda() for x in (d, f, u, v)]
    st = api.cauchy_blr_dev(1.3, da, *gd, dx, leaf=leaf, tol=tol, beta=0.5, compressor=compressor)
    torch.cuda.synchronize()
    c = ora.cauchy_matrix(d, f, u, v, 0, n, 0, n)
    err = ora.rel_fro(dx.cpu().numpy(), 1.3 * (a @ c) + 0.5 * x0)
    print("cauchy blr", compressor, "rel err", err, st, flush=True)
    assert err <= 20 * tol
    nb = n // leaf
    low, ranks = 0, 0
    for i in range(nb):
        for j in range(nb):
            if i == j:
                continue
            r, cc = slice(i * leaf, (i + 1) * leaf), slice(j * leaf, (j + 1) * leaf)
            if compressor == 1:
                rk, _, _, _ = ora.compress_cauchy(d[r], f[cc], u[r], v[cc], tol)
            else:
                rk, _, _ = ora.compress_rrqr(np.ascontiguousarray(c[r, cc]), tol)
            if 0 < rk <= 32 and rk < leaf:
                low += 1
                ranks += rk
    if compressor == 1:
        assert st["lowrank"] == low and st["rank_sum"] == ranks
    else:
        assert abs(st["rank_sum"] - ranks) <= 0.02 * ranks


def test_blr_rejects_unknown_compressor(gpu):
    n = 256
    gens, _, _ = _setup(n, 35)
    a = api.colmajor_empty(n, n)
    with pytest.raises(api.PsmmaError):
        api.cauchyeig_blr_dev(1.0, a, *gens, a, leaf=128, compressor=2)


def test_blr_merge_is_orthogonal(gpu, ora):
    import torch

    n, leaf = 2048, 256
    gens, gh, rng = _setup(n, 32)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    dx = api.colmajor_empty(n, n)
    api.cauchyeig_blr_dev(1.0, _dev(q), *gens, dx, leaf=leaf, tol=1e-12)
    torch.cuda.synchronize()
    x = dx.cpu().numpy()
    assert np.max(np.abs(x.T @ x - np.eye(n))) < 1e-11


def test_blr_structured_rrqr_equals_explicit(gpu, ora):
    """The product's RRQR evaluates each block from the generators once (a quotient
    equal to IEEE division for these operands) and reads that copy in every pass; run
    on the same blocks materialised by the oracle's CauchylikeEig formula, the
    explicit-input RRQR must make the same rank decisions, block for block."""
    import torch

    n, leaf = 1024, 256
    gens, gh, rng = _setup(n, 34)
    c = ora.cauchyeig_matrix(*gh, 0, n, 0, n)
    nb = n // leaf
    blocks = [c[i * leaf:(i + 1) * leaf, j * leaf:(j + 1) * leaf]
              for i in range(nb) for j in range(nb) if i != j]
    t = torch.from_numpy(np.stack([b.T for b in blocks])).cuda().contiguous()
    rank, _, _, _ = api.rrqr_batched_dev(t, 1e-12, 32)
    rank = rank.cpu().numpy()
    dx = api.colmajor_empty(n, n)
    st = api.cauchyeig_blr_dev(1.0, _dev(rng.standard_normal((n, n))), *gens, dx, leaf=leaf,
                               tol=1e-12)
    torch.cuda.synchronize()
    low = rank[(rank > 0) & (rank < leaf)]
    assert st["lowrank"] == len(low)
    assert st["rank_sum"] == int(low.sum())
    assert st["zero"] == int(np.sum(rank == 0))
    assert st["dense"] == int(np.sum((rank < 0) | (rank >= leaf)))


def test_blr_rejects_ragged_leaf(gpu):
    import torch

    n = 300
    gens, _, _ = _setup(n, 33)
    a = api.colmajor_empty(n, n)
    with pytest.raises(api.PsmmaError):
        api.cauchyeig_blr_dev(1.0, a, *gens, a, leaf=256)


def _toeplitz_kat_generator(n):
    # the reference driver's matrix N^2 I + (i - j) (Test/psmma_testtoeplitz.f90:39-46) as
    # H(r, c) = g[n - 1 + c - r]: g[n - 1 + d] = N^2 [d == 0] - d
    d = np.arange(-(n - 1), n, dtype=np.float64)
    return np.where(d == 0, float(n) ** 2, 0.0) - d


@pytest.mark.parametrize("kind", ["toeplitz", "hankel"])
@pytest.mark.parametrize("gen", ["kat", "smooth", "random"])
def test_window_blr_product_matches_oracle(gpu, ora, kind, gen):
    """Compressed Toeplitz / Hankel products against the explicit product: the driver's
    N^2 I + (i - j) matrix (every off-diagonal block of rank 2), a smooth generator
    (numerically low-rank blocks), and a random one (full-rank blocks: the structured
    FP64 kernel takes every one of them, as the reference regenerates the tile)."""
    import torch

    n, leaf, m = 1024, 256, 700
    print(f"window blr kind={kind} gen={gen}", flush=True)
    rng = np.random.default_rng(41)
    if gen == "kat":
        g = _toeplitz_kat_generator(n) if kind == "toeplitz" else np.arange(2 * n - 1, dtype=float)
    elif gen == "smooth":
        s = np.arange(2 * n - 1, dtype=float) - (n - 1)
        g = 1.0 / (1.0 + (s / 64.0) ** 2) if kind == "toeplitz" else 1.0 / (1.0 + s + n)
    else:
        g = rng.standard_normal(2 * n - 1)
    a = rng.standard_normal((m, n))
    x0 = rng.standard_normal((m, n))
    r = np.arange(n)[:, None]
    c = np.arange(n)[None, :]
    h = g[(n - 1) + c - r] if kind == "toeplitz" else g[r + c]
    want = -0.7 * (a @ h) + 1.5 * x0
    da, dx, dg = _dev(a), _dev(x0), torch.from_numpy(g).cuda()
    if kind == "toeplitz":
        st = api.toeplitz_blr_dev(-0.7, da, dg, n, dx, leaf=leaf, beta=1.5)
    else:
        st = api.hankel_blr_dev(-0.7, da, dg, dx, leaf=leaf, beta=1.5)
    torch.cuda.synchronize()
    err = ora.rel_fro(dx.cpu().numpy(), want)
    print("rel err", err, st, flush=True)
    nb = n // leaf
    assert st["blocks"] == nb * (nb - 1)
    if gen == "kat":
        assert st["lowrank"] == st["blocks"] and st["rank_sum"] == 2 * st["blocks"]
        assert err <= 1e-13
    elif gen == "smooth":
        assert st["lowrank"] >= st["blocks"] // 2
        assert err <= 1e-12
    else:
        assert st["dense"] == st["blocks"]
        assert err <= 1e-13
