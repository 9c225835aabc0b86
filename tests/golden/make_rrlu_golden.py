"""Writes rrlu_reference.npz: every output of the reference's own rrluCauchy
(SRC/Cauchylowrank.f90:7-134) and rrluCauchyEigrow (SRC/CauchyHssEig.f90:7-212),
and the explicit tiles of its Cauchylike (SRC/Cauchylowrank.f90:509-521) and
CauchylikeEig (SRC/CauchyHssEig.f90:1693-1774), compiled from the reference
sources by `make -C oracle ref`, on the blocks listed below (inputs regenerated
from their seeds by oracle/reference.py).  tests/test_oracle_ref_rrlu.py compares
the oracle's restatements with it, tests/test_gpu_ref_tiles.py the tiles the GPU
kernels generate.

Run from the repository root: python tests/golden/make_rrlu_golden.py"""
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle import oracle as ora  # noqa: E402
from oracle import reference as ref  # noqa: E402

TOLS = [1e-6, 1e-10, 1e-14, ora.RRLU_TOL]
CAUCHY = ([("separated", 80, 70, 1, t) for t in TOLS] +
          [("interlaced", 256, 256, 0, ora.RRLU_TOL), ("interlaced", 100, 60, 2, ora.RRLU_TOL),
           ("deficient", 96, 64, 6, ora.RRLU_TOL), ("separated", 40, 120, 3, 1e-12),
           ("separated", 150, 30, 4, 1e-12)])
EIG = ([(200, 60, 80, 5, t) for t in TOLS] +
       [(512, 128, 256, 3, ora.RRLU_TOL), (1024, 256, 256, 7, 1e-12),
        (1024, 256, 256, 7, ora.RRLU_TOL), (2048, 256, 256, 1, 1e-12)])
# tiles: (generator size, seed, r0, rows, c0, cols); the eigen-Cauchy blocks cross
# the diagonal (both branches of the pole difference), one of them at the last column
CAUCHY_TILES = [(256, 0, 32, 64, 100, 80)]
EIG_TILES = [(1024, 7, 100, 64, 60, 96), (1024, 7, 960, 64, 980, 44)]


def main():
    if not ref.available():
        sys.exit("build the reference driver first: make -C oracle ref")
    out = {"cauchy_specs": np.array(json.dumps(CAUCHY)), "eig_specs": np.array(json.dumps(EIG))}
    for i, (kind, m, n, seed, tol) in enumerate(CAUCHY):
        r = ref.rrlu_cauchy(*ref.cauchy_case(kind, m, n, seed), tol)
        for key, val in r.items():
            out[f"c{i}_{key}"] = np.asarray(val)
    for i, (ng, m, n, seed, tol) in enumerate(EIG):
        f, dl, dr, zh, v, rows, cols = ref.cauchyeig_case(ng, m, n, seed)
        r = ref.rrlu_cauchyeig_row(f, dl, dr, f[cols], zh[rows], v[cols], rows + 1, cols + 1, tol)
        for key, val in r.items():
            out[f"e{i}_{key}"] = np.asarray(val)
    out["tile_specs"] = np.array(json.dumps({"cauchy": CAUCHY_TILES, "eig": EIG_TILES}))
    for i, (ng, seed, r0, m, c0, n) in enumerate(CAUCHY_TILES):
        d, f, u, v = ref.cauchy_tile_generators(ng, seed)
        out[f"tc{i}"] = ref.cauchy_tile(d[r0:r0 + m], f[c0:c0 + n], u[r0:r0 + m], v[c0:c0 + n])
    for i, (ng, seed, r0, m, c0, n) in enumerate(EIG_TILES):
        f, dl, dr, zh, v, _, _ = ref.cauchyeig_case(ng, 1, 1, seed)
        out[f"te{i}"] = ref.cauchyeig_tile(f, dl, dr, zh, v, np.arange(r0, r0 + m),
                                           np.arange(c0, c0 + n))
    np.savez_compressed(Path(__file__).with_name("rrlu_reference.npz"), **out)


if __name__ == "__main__":
    main()
