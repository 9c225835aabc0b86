"""The tiles the GPU kernels generate in LDS against the reference's own tile
generators: Cauchylike (SRC/Cauchylowrank.f90:509-521) and CauchylikeEig
(SRC/CauchyHssEig.f90:1693-1774, both branches of the pole difference), as
compiled from the reference sources into the golden fixture
(tests/golden/rrlu_reference.npz, tests/golden/make_rrlu_golden.py).

The product I * H(r0:r0+m, c0:c0+n) with an identity A returns the generated
block exactly (every product is h * 1 or h * 0 and every sum adds exact zeros),
so the FP64 structured kernel and the persistent batch kernel must reproduce the
reference's tiles bitwise."""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import reference as ref
from psmma_git_amd import api

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden" / "rrlu_reference.npz"


@pytest.fixture(scope="module")
def golden():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _dev(x):
    import torch

    return torch.from_numpy(np.ascontiguousarray(x, np.float64)).cuda()


def _eye(m):
    import torch

    return torch.eye(m, dtype=torch.float64, device="cuda").t()  # column-major view


def _host(c):
    import torch

    torch.cuda.synchronize()
    return c.cpu().numpy()


def test_cauchy_kernel_tiles_equal_the_reference(gpu, golden):
    specs = json.loads(str(golden["tile_specs"]))
    for i, (ng, seed, r0, m, c0, n) in enumerate(specs["cauchy"]):
        print(f"case cauchy tile {i}: {m} x {n} at ({r0}, {c0})", flush=True)
        d, f, u, v = (_dev(x) for x in ref.cauchy_tile_generators(ng, seed))
        eye = _eye(m)
        c = api.colmajor_empty(m, n)
        api.cauchy_dev(1.0, eye, d, f, u, v, n, c, r0=r0, c0=c0)
        assert np.array_equal(_host(c), golden[f"tc{i}"])
        cb = api.colmajor_empty(m, n)
        api.batched_dev([api.make_problem(api.KIND_CAUCHY, 1.0, eye, [d, f, u, v], n, cb,
                                          r0=r0, c0=c0)])
        assert np.array_equal(_host(cb), golden[f"tc{i}"])


def test_cauchyeig_kernel_tiles_equal_the_reference(gpu, golden):
    specs = json.loads(str(golden["tile_specs"]))
    for i, (ng, seed, r0, m, c0, n) in enumerate(specs["eig"]):
        print(f"case eigen-Cauchy tile {i}: {m} x {n} at ({r0}, {c0})", flush=True)
        f, dl, dr, zh, v, _, _ = ref.cauchyeig_case(ng, 1, 1, seed)
        gens = [_dev(x) for x in (f, zh, v, dl, dr)]
        eye = _eye(m)
        c = api.colmajor_empty(m, n)
        api.cauchyeig_dev(1.0, eye, *gens, n, c, r0=r0, c0=c0)
        assert np.array_equal(_host(c), golden[f"te{i}"])
        cb = api.colmajor_empty(m, n)
        api.batched_dev([api.make_problem(api.KIND_CAUCHYEIG, 1.0, eye, gens, n, cb,
                                          r0=r0, c0=c0, lg=ng)])
        assert np.array_equal(_host(cb), golden[f"te{i}"])
