"""Eigen-Cauchy generators on the GPU (psmma_secular_dev) against the oracle's
secular roots, Loewner z and column normalisers, and the merge product
Q * C(generators) that consumes them.

Bars: roots, zhat and normalisers within 1e-12 relative of the oracle (both
converge to within a few units in the last place of the shifted root); the
product within 1e-13 relative Frobenius error of the FP64 oracle product; the
columns of Q * C orthonormal to 1e-12 when Loewner's z is used.
"""
import numpy as np
import pytest

from oracle import inputs
from psmma_git_amd import api

pytestmark = pytest.mark.gpu


def _gpu_gens(d, z, loewner):
    import torch

    td = torch.from_numpy(d).cuda()
    tz = torch.from_numpy(z).cuda()
    difl, difr, u, v = api.secular_dev(td, tz, 1.0, loewner=loewner)
    torch.cuda.synchronize()
    return td, difl, difr, u, v


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.abs(b), initial=0.0))


@pytest.mark.parametrize("n,seed", [(1, 0), (2, 1), (7, 2), (300, 3), (1000, 4)])
def test_secular_roots_match_oracle(gpu, n, seed):
    print(f"secular roots n={n} seed={seed}", flush=True)
    d, z, _ = inputs.merge_poles_and_z(n, seed)
    wl, wr = inputs.secular_roots(d, z, 1.0)
    _, difl, difr, u, v = _gpu_gens(d, z, loewner=True)
    gl, gr = difl.cpu().numpy(), difr.cpu().numpy()
    assert _rel(gl, wl) < 1e-12
    assert _rel(gr[:-1], wr[:-1]) < 1e-12
    assert (gl > 0).all() and (gr[: n - 1] < 0).all()
    zh = inputs.loewner_z(d, z, wl, wr)
    assert _rel(u.cpu().numpy(), zh) < 1e-11
    vw = inputs.column_normalisers(d, zh, wl, wr)
    assert _rel(v.cpu().numpy(), vw) < 1e-11


def test_secular_without_loewner_normalises_z(gpu):
    n = 400
    d, z, _ = inputs.merge_poles_and_z(n, 11)
    _, difl, difr, u, v = _gpu_gens(d, z, loewner=False)
    assert np.array_equal(u.cpu().numpy(), z)
    vw = inputs.column_normalisers(d, z, difl.cpu().numpy(), difr.cpu().numpy())
    assert _rel(v.cpu().numpy(), vw) < 1e-13


def test_secular_rejects_bad_arguments(gpu):
    import torch

    d = torch.zeros(4, dtype=torch.float64, device="cuda")
    st = api.lib().psmma_secular_dev(4, d.data_ptr(), d.data_ptr(), 0.0, d.data_ptr(),
                                     d.data_ptr(), None, None, None)
    assert st == -4
    st = api.lib().psmma_secular_dev(-1, None, None, 1.0, None, None, None, None, None)
    assert st == -1


@pytest.mark.parametrize("prec,slices,tol", [(api.PREC_FP64, 0, 1e-13),
                                             (api.PREC_SPLIT_I8, 14, 1e-12)])
def test_merge_product_orthonormal(gpu, ora, prec, slices, tol):
    """X = Q * C with Q orthogonal and C the eigenvector matrix from GPU generators:
    X matches the oracle product and has orthonormal columns."""
    import torch

    n = 768
    print(f"merge product n={n} prec={prec}", flush=True)
    d, z, rng = inputs.merge_poles_and_z(n, 21)
    td, difl, difr, u, v = _gpu_gens(d, z, loewner=True)
    g = rng.standard_normal((n, n))
    q, _ = np.linalg.qr(g)
    tq = torch.from_numpy(np.asfortranarray(q)).cuda().t().contiguous().t()
    x = torch.empty((n, n), dtype=torch.float64, device="cuda").t()
    api.cauchyeig_dev(1.0, tq, td, u, v, difl, difr, n, x, prec=prec, slices=slices)
    torch.cuda.synchronize()
    gens = [t.cpu().numpy() for t in (td, u, v, difl, difr)]
    c = ora.cauchyeig_matrix(*gens, 0, n, 0, n)
    want = q @ c
    got = x.cpu().numpy()
    assert ora.rel_fro(got, want) < tol
    assert np.max(np.abs(got.T @ got - np.eye(n))) < 1e-12
