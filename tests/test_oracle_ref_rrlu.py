"""The oracle's C restatements of the generator-only RRLU routines against the
reference itself: rrluCauchy (SRC/Cauchylowrank.f90:7-134) and rrluCauchyEigrow
(SRC/CauchyHssEig.f90:7-212), compiled from the reference's module sources with
amdflang (oracle/Makefile, target ref; FP contraction off in both builds).

Every output must be bitwise equal: the rank, both permutations, the permuted
generators, the Schur-complement generators U / V and the interpolation
generators W (and for the eigen-Cauchy kind the row permutation PQ and the swap
count).  The golden fixture (tests/golden/rrlu_reference.npz, written by
tests/golden/make_rrlu_golden.py from the compiled reference) holds these
outputs for 17 blocks at four tolerances, and three explicit tiles of the
reference's Cauchylike / CauchylikeEig, so the comparison runs wherever the
oracle builds; where the reference driver is built as well, more blocks are
compared live.  The eigen-Cauchy blocks lie in the first block row (rows 1..M):
elsewhere CauchyPivtEig's early-return test reads the block's row-pole array at
a global index past its end (SRC/CauchyHssEig.f90:675-678), which the
restatement replaces with the global pole array."""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import reference as ref

GOLDEN = Path(__file__).resolve().parent / "golden" / "rrlu_reference.npz"
FLOAT_KEYS_C = ("d", "f", "a", "v", "u")
FLOAT_KEYS_E = ("f", "a", "v", "u")


@pytest.fixture(scope="module")
def golden():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _same(got, want, keys, ints):
    rk = int(want["rk"])
    assert int(got["rk"]) == rk
    for k in ints:
        assert np.array_equal(np.asarray(got[k]), np.asarray(want[k])), k
    for k in keys:
        assert np.array_equal(np.asarray(got[k]), np.asarray(want[k])), k
    # W holds the interpolation generators of the rk chosen rows
    assert np.array_equal(np.asarray(got["w"])[:rk], np.asarray(want["w"])[:rk])


def _cauchy_specs(golden):
    return json.loads(str(golden["cauchy_specs"]))


def _eig_specs(golden):
    return json.loads(str(golden["eig_specs"]))


def test_golden_covers_every_tolerance_and_the_swap_pass(golden):
    c, e = _cauchy_specs(golden), _eig_specs(golden)
    assert len(c) == 9 and len(e) == 8
    ranks = [int(golden[f"c{i}_rk"]) for i in range(len(c))]
    assert ranks[:4] == sorted(ranks[:4]) and ranks[0] < ranks[3]  # tighter tol, more rows
    assert max(ranks) >= 20  # the plain swap pass (every 10th step) ran at least twice
    assert max(int(golden[f"e{i}_nswap"]) for i in range(len(e))) >= 1  # an eigen swap moved rows


@pytest.mark.parametrize("i", range(9))
def test_rrlu_cauchy_equals_the_reference(ora, golden, i):
    kind, m, n, seed, tol = _cauchy_specs(golden)[i]
    want = {k[len(f"c{i}_"):]: golden[k] for k in golden if k.startswith(f"c{i}_")}
    got = ora.rrlu_cauchy(*ref.cauchy_case(kind, m, n, seed), tol)
    _same(got, want, FLOAT_KEYS_C, ("pl", "pu"))


@pytest.mark.parametrize("i", range(8))
def test_rrlu_cauchyeig_equals_the_reference(ora, golden, i):
    ng, m, n, seed, tol = _eig_specs(golden)[i]
    want = {k[len(f"e{i}_"):]: golden[k] for k in golden if k.startswith(f"e{i}_")}
    f, dl, dr, zh, v, rows, cols = ref.cauchyeig_case(ng, m, n, seed)
    got = ora.rrlu_cauchyeig_row(f, dl, dr, zh[rows], v[cols], rows + 1, cols + 1, tol)
    _same(got, want, FLOAT_KEYS_E, ("pl", "pu", "pq"))
    assert got["nswap"] == int(want["nswap"])


@pytest.mark.skipif(not ref.available(), reason="reference driver not built (make -C oracle ref)")
@pytest.mark.parametrize("seed", [11, 12, 13])
def test_live_reference_more_blocks(ora, seed):
    for kind, m, n in (("separated", 64, 48), ("interlaced", 128, 96), ("deficient", 72, 72)):
        for tol in (1e-8, 1e-13, ora.RRLU_TOL):
            args = ref.cauchy_case(kind, m, n, seed)
            _same(ora.rrlu_cauchy(*args, tol), ref.rrlu_cauchy(*args, tol), FLOAT_KEYS_C,
                  ("pl", "pu"))
    f, dl, dr, zh, v, rows, cols = ref.cauchyeig_case(768, 192, 192, seed)
    for tol in (1e-8, 1e-13, ora.RRLU_TOL):
        got = ora.rrlu_cauchyeig_row(f, dl, dr, zh[rows], v[cols], rows + 1, cols + 1, tol)
        want = ref.rrlu_cauchyeig_row(f, dl, dr, f[cols], zh[rows], v[cols], rows + 1,
                                      cols + 1, tol)
        _same(got, want, FLOAT_KEYS_E, ("pl", "pu", "pq"))
        assert got["nswap"] == want["nswap"]


# Explicit tiles: Cauchylike (SRC/Cauchylowrank.f90:509-521) and CauchylikeEig
# (SRC/CauchyHssEig.f90:1693-1774), the formulas every structured tile of the
# Cauchy kinds comes from (both branches of the eigen-Cauchy pole difference).

def _tile_specs(golden):
    return json.loads(str(golden["tile_specs"]))


def test_cauchy_tiles_equal_the_reference(ora, golden):
    for i, (ng, seed, r0, m, c0, n) in enumerate(_tile_specs(golden)["cauchy"]):
        d, f, u, v = ref.cauchy_tile_generators(ng, seed)
        got = ora.cauchy_matrix(d, f, u, v, r0, m, c0, n)
        assert np.array_equal(got, golden[f"tc{i}"])


def test_cauchyeig_tiles_equal_the_reference(ora, golden):
    for i, (ng, seed, r0, m, c0, n) in enumerate(_tile_specs(golden)["eig"]):
        f, dl, dr, zh, v, _, _ = ref.cauchyeig_case(ng, 1, 1, seed)
        rows, cols = np.arange(r0, r0 + m)[:, None], np.arange(c0, c0 + n)[None, :]
        assert np.any(rows <= cols) and np.any(rows > cols)  # both branches
        got = ora.cauchyeig_matrix(f, zh, v, dl, dr, r0, m, c0, n)
        assert np.array_equal(got, golden[f"te{i}"])


@pytest.mark.skipif(not ref.available(), reason="reference driver not built (make -C oracle ref)")
def test_live_reference_tiles(ora):
    f, dl, dr, zh, v, _, _ = ref.cauchyeig_case(600, 1, 1, 21)
    for r0, m, c0, n in ((0, 600, 0, 600), (250, 90, 10, 300)):
        want = ref.cauchyeig_tile(f, dl, dr, zh, v, np.arange(r0, r0 + m), np.arange(c0, c0 + n))
        assert np.array_equal(ora.cauchyeig_matrix(f, zh, v, dl, dr, r0, m, c0, n), want)
    d, ff, u, vv = ref.cauchy_tile_generators(500, 22)
    want = ref.cauchy_tile(d, ff, u, vv)
    assert np.array_equal(ora.cauchy_matrix(d, ff, u, vv, 0, 500, 0, 500), want)
