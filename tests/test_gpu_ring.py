"""The reference's P x Q grid on one GPU (MI355X only).

* Ring steps: every emulated rank's steps of PS*_COMPUTE (index lists from
  psmma_constindex, step class from psmma_is_intersect, product from
  psmma_ring_step_dev) replayed on the GPU; the assembled C must match the
  oracle's ring (dense mode) and the explicit product to 1e-13.
* Redistribution: psmma_redistribute_dev between block-cyclic layouts equals
  the layouts built in numpy from the global matrix, bitwise, both ways.
"""
import numpy as np
import pytest

from oracle import inputs
from psmma_git_amd import api

pytestmark = pytest.mark.gpu

TOL = 1e-13


def dev_cm(x):
    import torch

    return torch.from_numpy(np.ascontiguousarray(x.T)).cuda().t()


def make_kind(kind, n, rng):
    """(oracle kind, gens dict for the numpy oracle, C-restatement generator array,
    device generator list, lg, korder)."""
    import torch

    if kind == "cauchy":
        b = inputs.cauchy_generators(n, rng)
        gens = dict(d=b[:, 0], f=b[:, 1], u=b[:, 2], v=b[:, 3])
        dev = [torch.from_numpy(np.ascontiguousarray(b[:, i])).cuda() for i in range(4)]
        return 0, gens, b, dev, n, 0
    if kind in ("toeplitz", "hankel"):
        g = rng.standard_normal(2 * n - 1)
        dev = [torch.from_numpy(g).cuda()]
        return (1, dict(g=g, korder=n), g, dev, 2 * n - 1, n) if kind == "toeplitz" else \
               (2, dict(g=g), g, dev, 2 * n - 1, 0)
    e = inputs.eig_merge_inputs(n, seed=int(rng.integers(1 << 20)), with_q=False)
    gens = dict(f=e["f"], u=e["u"], v=e["v"], difl=e["difl"], difr=e["difr"])
    gen = np.asfortranarray(np.stack([e["f"], e["f"] + e["difl"], e["u"], e["v"], e["difl"],
                                      e["difr"]], axis=1))
    dev = [torch.from_numpy(np.ascontiguousarray(gens[k])).cuda()
           for k in ("f", "u", "v", "difl", "difr")]
    return 3, gens, gen, dev, n, 0


def replay_ring(okind, dev, lg, korder, P, Q, nb, alpha, a, beta, c):
    """Every rank (myrow, mycol) runs its ring steps over the global A, C (MB = NB)."""
    import torch

    n, k = a.shape
    out = c.copy()
    steps = dict(dense=0, offdiag=0)
    for myrow in range(P):
        rows = api.constindex(n, nb, myrow, P)
        for mycol in range(Q):
            cidx = api.constindex(k, nb, mycol, Q)
            if rows.size == 0 or cidx.size == 0:
                continue
            dc = dev_cm(c[np.ix_(rows - 1, cidx - 1)])
            tbeta = beta
            for k1 in range(Q):
                ridx = api.constindex(k, nb, (mycol + k1) % Q, Q)
                if ridx.size == 0:
                    continue
                steps["offdiag" if api.is_intersect(cidx, ridx) else "dense"] += 1
                da = dev_cm(a[np.ix_(rows - 1, ridx - 1)])
                p = api.make_problem(okind, alpha, da, dev, cidx.size, dc, beta=tbeta, lg=lg,
                                     korder=korder)
                api.ring_step_dev(p, ridx, cidx)
                tbeta = 1.0
            torch.cuda.synchronize()
            out[np.ix_(rows - 1, cidx - 1)] = dc.cpu().numpy()
    return out, steps


@pytest.mark.parametrize("kind", ["cauchy", "toeplitz", "hankel", "cauchyeig"])
@pytest.mark.parametrize("grid,nb,n", [((2, 2), 16, 300), ((2, 3), 5, 301), ((1, 4), 7, 250),
                                       ((3, 2), 32, 257), ((1, 2), 64, 4100)])
def test_ring_replay_matches_oracle_ring(gpu, ora, kind, grid, nb, n):
    P, Q = grid
    if n > 1000 and kind != "toeplitz":
        pytest.skip("one large case (128 x 128 tiles of the index-list kernel) per run")
    print(f"case ring {kind} grid={grid} nb={nb} n={n}", flush=True)
    rng = np.random.default_rng(n + nb + P * 10 + Q)
    okind, gens, gen, dev, lg, korder = make_kind(kind, n, rng)
    a = rng.uniform(-1, 1, (n, n))
    c = rng.standard_normal((n, n))
    got, steps = replay_ring(okind, dev, lg, korder, P, Q, nb, 1.3, a, 0.5, c)
    want = ora.structured_gemm_np(okind, 1.3, a, gens, 0, 0, n, 0.5, c)
    assert ora.rel_fro(got, want) < TOL
    if n <= 1000:
        ring, stats = ora.ps_compute(okind, P, Q, nb, 1.3, a, gen, beta=0.5, c=c,
                                     mode=ora.MODE_DENSE)
        assert ora.rel_fro(got, ring) < TOL
        assert stats[0] == steps["dense"] + steps["offdiag"]
    assert steps["dense"] == P * Q  # one diagonal step per rank; the rest are off-diagonal
    assert steps["offdiag"] == P * Q * (Q - 1)


def test_ring_step_rejects_out_of_range_indices(gpu):
    import torch

    n = 64
    g = torch.zeros(2 * n - 1, dtype=torch.float64, device="cuda")
    a = torch.zeros((n, n), dtype=torch.float64, device="cuda").t()
    c = torch.full((n, n), 2.0, dtype=torch.float64, device="cuda").t()
    p = api.make_problem(api.KIND_HANKEL, 1.0, a, [g], n, c, lg=2 * n - 1)
    ok = np.arange(1, n + 1, dtype=np.int32)
    bad = ok.copy()
    bad[5] = n + 1
    with pytest.raises(api.PsmmaError, match="-2"):
        api.ring_step_dev(p, bad, ok)
    with pytest.raises(api.PsmmaError, match="-3"):
        api.ring_step_dev(p, ok, np.where(ok == 3, 0, ok))
    torch.cuda.synchronize()
    assert bool((c == 2.0).all())


def local_layout(x, P, Q, mb, nb):
    """All ranks' local arrays of x, rank-major, each column-major (numpy)."""
    m, n = x.shape
    parts = []
    for pr in range(P):
        r = api.constindex(m, mb, pr, P) - 1
        for pc in range(Q):
            cc = api.constindex(n, nb, pc, Q) - 1
            parts.append(x[np.ix_(r, cc)].ravel(order="F"))
    return np.concatenate(parts) if parts else np.zeros(0)


@pytest.mark.parametrize("m,n,P,Q,src,dst", [
    (100, 100, 2, 2, (8, 8), (50, 50)),      # BCDD (NB = 8) -> BDD (ceil(N/P) = 50)
    (100, 100, 3, 3, (8, 8), (34, 34)),
    (61, 47, 2, 3, (8, 5), (4, 16)),
    (1000, 700, 4, 2, (64, 64), (250, 350)),
    (5, 3, 4, 4, (1, 1), (2, 2)),
])
def test_redistribution_is_the_layout_permutation(gpu, m, n, P, Q, src, dst):
    import torch

    print(f"case redistribute {m}x{n} grid {P}x{Q} {src}->{dst}", flush=True)
    rng = np.random.default_rng(m * n + P)
    x = rng.standard_normal((m, n))
    xs = local_layout(x, P, Q, *src)
    xd = local_layout(x, P, Q, *dst)
    da = torch.from_numpy(xs).cuda()
    db = torch.full((xd.size,), np.nan, dtype=torch.float64, device="cuda")
    api.redistribute_dev(m, n, P, Q, src[0], src[1], da, dst[0], dst[1], db)
    back = torch.full_like(da, np.nan)
    api.redistribute_dev(m, n, P, Q, dst[0], dst[1], db, src[0], src[1], back)
    torch.cuda.synchronize()
    assert np.array_equal(db.cpu().numpy(), xd)
    assert np.array_equal(back.cpu().numpy(), xs)
