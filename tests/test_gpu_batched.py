"""Batched uneven products (grouped launch and hipGraph plans) against the oracle."""
import numpy as np
import pytest

from oracle import inputs
from psmma_git_amd import api
from test_gpu_parity import ORA_KIND, dev_colmajor, host

pytestmark = pytest.mark.gpu

KINDS = ["cauchy", "toeplitz", "hankel", "cauchyeig"]


def make_batch(rng, nprob, nmax, beta_zero=True, nmin=1):
    import torch

    probs, keep = [], []
    for i in range(nprob):
        kind = KINDS[i % 4]
        n = int(rng.integers(nmin, nmax + 1))
        m = int(rng.integers(nmin, nmax + 1))
        nglob = n + 3
        if kind == "cauchy":
            b = inputs.cauchy_generators(nglob, rng)
            gens = dict(d=b[:, 0].copy(), f=b[:, 1].copy(), u=b[:, 2].copy(), v=b[:, 3].copy())
            order = ["d", "f", "u", "v"]
        elif kind in ("toeplitz", "hankel"):
            gens = dict(g=rng.standard_normal(2 * nglob - 1), korder=nglob)
            order = ["g"]
        else:
            e = inputs.eig_merge_inputs(nglob, seed=int(rng.integers(1 << 30)), with_q=False)
            gens = dict(f=e["f"], u=e["u"], v=e["v"], difl=e["difl"], difr=e["difr"])
            order = ["f", "u", "v", "difl", "difr"]
        dgens = [torch.from_numpy(np.ascontiguousarray(gens[k])).cuda() for k in order]
        a = rng.uniform(-1, 1, (m, n))
        c = rng.standard_normal((m, n))
        beta = 0.0 if beta_zero else 0.5
        da = dev_colmajor(a, ld=m + 1)
        dc = dev_colmajor(c, ld=m + 2)
        r0, c0 = (2, 1) if kind != "toeplitz" else (1, 2)
        lg = len(gens["g"]) if "g" in gens else (nglob if kind == "cauchyeig" else 0)
        p = api.make_problem(ORA_KIND[kind], 1.5, da, dgens, n, dc, beta=beta, r0=r0, c0=c0,
                             lg=lg, korder=nglob if kind == "toeplitz" else 0)
        want = ora_want(kind, a, gens, r0, c0, n, beta, c)
        probs.append(p)
        keep.append((dc, want, dgens, da))
    return probs, keep


def ora_want(kind, a, gens, r0, c0, n, beta, c):
    from oracle import oracle as ora

    return ora.structured_gemm_np(ORA_KIND[kind], 1.5, a, gens, r0, c0, n, beta=beta, c=c)


def errors(keep):
    from oracle import oracle as ora

    return [ora.rel_fro(host(dc), want) for dc, want, _, _ in keep]


def test_batched_fp64_one_shot(gpu):
    import torch

    rng = np.random.default_rng(21)
    probs, keep = make_batch(rng, 24, 300, beta_zero=False)
    api.batched_dev(probs)
    torch.cuda.synchronize()
    errs = errors(keep)
    print("max err", max(errs))
    assert max(errs) < 1e-13


def test_batched_fp64_plan_replays(gpu):
    # graph replays with eager work and synchronisation between them; every
    # replay must equal the eager result
    import torch

    rng = np.random.default_rng(22)
    probs, keep = make_batch(rng, 40, 400)
    plan = api.BatchedPlan(probs)
    cfg, nodes = plan.info()
    print("plan cfg", cfg, "nodes", nodes)
    assert nodes == 1  # one persistent launch over every problem of every kind
    for rep in range(3):
        for dc, _, _, _ in keep:
            dc.fill_(float("nan"))  # eager work on the same stream between replays
        torch.cuda.synchronize()
        plan.launch()
        torch.cuda.synchronize()
        assert max(errors(keep)) < 1e-13, rep
    plan.close()


def test_batched_split_plan(gpu):
    import torch

    rng = np.random.default_rng(23)
    probs, keep = make_batch(rng, 16, 300)
    plan = api.BatchedPlan(probs, prec=api.PREC_SPLIT_I8, slices=0)
    for rep in range(2):
        plan.launch()
        torch.cuda.synchronize()
        errs = errors(keep)
        print("split errs max", max(errs))
        assert max(errs) <= 1e-12
    plan.close()


def test_batched_rejects_invalid_problem(gpu):
    import torch

    a = dev_colmajor(np.ones((4, 4)))
    c = dev_colmajor(np.zeros((4, 4)))
    g = torch.ones(3, dtype=torch.float64, device="cuda")  # too short for K = 4
    p = api.make_problem(api.KIND_TOEPLITZ, 1.0, a, [g], 4, c, lg=3, korder=4)
    with pytest.raises(api.PsmmaError):
        api.batched_dev([p])


def test_batched_split_one_shot_with_zero_sized(gpu):
    import torch

    rng = np.random.default_rng(24)
    probs, keep = make_batch(rng, 12, 260, beta_zero=False)
    # an empty problem in the batch is skipped
    a0 = dev_colmajor(np.ones((3, 0)).reshape(3, 0) if False else np.ones((3, 1)))
    c0 = dev_colmajor(np.zeros((3, 1)))
    g0 = torch.ones(4, dtype=torch.float64, device="cuda")
    empty = api.make_problem(api.KIND_HANKEL, 1.0, a0, [g0], 0, c0, lg=4)
    api.batched_dev(probs + [empty], prec=api.PREC_SPLIT_I8, slices=0)
    torch.cuda.synchronize()
    errs = errors(keep)
    print("split one-shot errs max", max(errs))
    assert max(errs) <= 1e-12


def test_batched_split_large_tiles(gpu):
    """Problems with min(m, n) >= 2048 and k >= 4096 take the 256 x 256 tile class of the
    grouped split GEMM (ragged edges included); small ones keep 128 x 128 in the same
    plan."""
    import torch

    rng = np.random.default_rng(25)
    big, kb = make_batch(rng, 4, 2100, nmin=2048)
    small, ks = make_batch(rng, 4, 200)
    plan = api.BatchedPlan(big + small, prec=api.PREC_SPLIT_I8, slices=0)
    plan.launch()
    torch.cuda.synchronize()
    errs = errors(kb + ks)
    print("split large-tile errs max", max(errs))
    assert max(errs) <= 1e-12
    plan.close()


def make_rect_batch(rng, shapes, beta):
    """Problems of shape (kind, m, n, k): rectangular H, block offsets, beta."""
    import torch

    probs, keep = [], []
    for kind, m, n, k in shapes:
        nglob = max(n, k) + 5
        r0, c0 = 3, 2
        if kind == "cauchy":
            b = inputs.cauchy_generators(nglob, rng)
            gens = dict(d=b[:, 0].copy(), f=b[:, 1].copy(), u=b[:, 2].copy(), v=b[:, 3].copy())
            order, lg, korder = ["d", "f", "u", "v"], 0, 0
        elif kind in ("toeplitz", "hankel"):
            gens = dict(g=rng.standard_normal(2 * nglob - 1), korder=nglob)
            order, lg = ["g"], 2 * nglob - 1
            korder = nglob if kind == "toeplitz" else 0
        else:
            e = inputs.eig_merge_inputs(nglob, seed=int(rng.integers(1 << 30)), with_q=False)
            gens = dict(f=e["f"], u=e["u"], v=e["v"], difl=e["difl"], difr=e["difr"])
            order, lg, korder = ["f", "u", "v", "difl", "difr"], nglob, 0
        dgens = [torch.from_numpy(np.ascontiguousarray(gens[q])).cuda() for q in order]
        a = rng.uniform(-1, 1, (m, k))
        c = rng.standard_normal((m, n))
        da = dev_colmajor(a, ld=m + 3)
        dc = dev_colmajor(c, ld=m + 1)
        p = api.make_problem(ORA_KIND[kind], -0.75, da, dgens, n, dc, beta=beta, r0=r0, c0=c0,
                             lg=lg, korder=korder)
        want = ora_want_k(kind, a, gens, r0, c0, n, beta, c)
        probs.append(p)
        keep.append((dc, want, dgens, da, c))
    return probs, keep


def ora_want_k(kind, a, gens, r0, c0, n, beta, c):
    from oracle import oracle as ora

    return ora.structured_gemm_np(ORA_KIND[kind], -0.75, a, gens, r0, c0, n, beta=beta, c=c)


@pytest.mark.parametrize("beta", [0.0, 0.5])
def test_batched_split_fused_matches_residue_pipeline(gpu, monkeypatch, beta):
    """The fused split form (all moduli of a 128 x 128 tile in one workgroup, the
    recombination as a 64-bit fraction in registers) against the oracle and against the
    residue-plane pipeline on the same batch: ragged tiles, K != N, block offsets, the
    generator-window Toeplitz / Hankel form (kp <= 2048) and its dense fallback
    (kp = 2304), every kind, beta = 0 and beta != 0."""
    import torch
    from oracle import oracle as ora

    rng = np.random.default_rng(31 if beta == 0.0 else 32)
    shapes = [("cauchy", 300, 257, 700), ("toeplitz", 129, 390, 1030), ("hankel", 517, 130, 64),
              ("cauchyeig", 250, 260, 300), ("toeplitz", 700, 650, 2100),
              ("hankel", 333, 200, 2049), ("cauchy", 1, 1, 1), ("toeplitz", 60, 70, 50),
              ("hankel", 1030, 1100, 2000), ("cauchy", 1000, 900, 1500)]
    probs, keep = make_rect_batch(rng, shapes, beta)
    outs = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("PSMMA_GSPLIT_FUSED", fused)
        plan = api.BatchedPlan(probs, prec=api.PREC_SPLIT_I8, slices=14)
        for rep in range(2):  # replays: C is reset in between (beta != 0 reads it)
            for dc, _, _, _, c in keep:
                dc.copy_(dev_colmajor(c, ld=dc.stride(1)))
            torch.cuda.synchronize()
            plan.launch()
            torch.cuda.synchronize()
        plan.close()
        outs[fused] = [host(dc) for dc, _, _, _, _ in keep]
    for i, (s, (_, want, _, _, _)) in enumerate(zip(shapes, keep)):
        ef = ora.rel_fro(outs["1"][i], want)
        eu = ora.rel_fro(outs["0"][i], want)
        d = ora.rel_fro(outs["1"][i], outs["0"][i])
        print(f"{s}: fused {ef:.2e} residue-plane {eu:.2e} between {d:.2e}", flush=True)
        assert ef <= 1e-12 and eu <= 1e-12
        # the fused form adds the fraction's rounding to the operand rounding the
        # residue-plane form shares: sum_l r_l d_l with |r_l| <= 128, |d_l| <= 1/2 unit of
        # 2^-65 M (W_l rounded), a few 1e-14 of the product for balanced columns and up
        # to about 1e-13 where a Cauchy column's norm sits in a few entries
        assert ef <= eu + 2e-13 and d <= 2e-13


def test_batched_split_long_contractions_take_residue_planes(gpu):
    """K > 16384 leaves the fused form (its residue reduction needs kp <= 16384) and the
    generator-window form (the window holds at most 16384 steps): the whole batch runs
    the residue-plane pipeline, with the Toeplitz / Hankel operands as dense residue planes
    and their column exponents from the generators."""
    import torch
    from oracle import oracle as ora

    rng = np.random.default_rng(33)
    shapes = [("toeplitz", 70, 50, 16385), ("hankel", 40, 60, 16390), ("cauchy", 30, 20, 16400),
              ("cauchyeig", 50, 40, 300)]
    probs, keep = make_rect_batch(rng, shapes, 0.0)
    plan = api.BatchedPlan(probs, prec=api.PREC_SPLIT_I8, slices=15)
    plan.launch()
    torch.cuda.synchronize()
    plan.close()
    for s, (dc, want, _, _, _) in zip(shapes, keep):
        err = ora.rel_fro(host(dc), want)
        print(f"{s}: {err:.2e}", flush=True)
        assert err <= 1e-12


def test_batched_split_fused_random_shapes(gpu):
    """Forty problems of random kind and shape (m, n, k in [1, 700], block offsets,
    beta = 0 or 1.25) through the fused split plan at the measured depth, each against
    the oracle."""
    import torch
    from oracle import oracle as ora

    rng = np.random.default_rng(34)
    kinds = ["cauchy", "toeplitz", "hankel", "cauchyeig"]
    shapes = [(kinds[int(rng.integers(4))], int(rng.integers(1, 701)), int(rng.integers(1, 701)),
               int(rng.integers(1, 701))) for _ in range(40)]
    for beta in (0.0, 1.25):
        probs, keep = make_rect_batch(rng, shapes, beta)
        plan = api.BatchedPlan(probs, prec=api.PREC_SPLIT_I8, slices=0)
        t, _ = plan.info()
        plan.launch()
        torch.cuda.synchronize()
        plan.close()
        errs = [ora.rel_fro(host(dc), want) for dc, want, _, _, _ in keep]
        print(f"beta {beta}: depth {t}, max err {max(errs):.2e}", flush=True)
        assert max(errs) <= 1e-12
