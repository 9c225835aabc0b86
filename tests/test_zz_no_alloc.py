"""psmma_toeplitz_noalloc_dev: the allocation-free variant of psmma_toeplitz_dev.

Called through the C ABI; its results must equal those of psmma_toeplitz_dev on
the same inputs (same kernels, same moduli: bitwise), on two calls in a row, the
second smaller than the first, and stay within the int8 form's bar against the
FP64 oracle.  Device free memory does not change across the calls.
"""
import numpy as np
import pytest

from psmma_git_amd import api

pytestmark = pytest.mark.gpu

SPLIT_TOL = 1e-12


def _case(n_rows, n_cols, k, korder, seed):
    import torch

    rng = np.random.default_rng(seed)
    a = rng.uniform(-1, 1, (n_rows, k))
    g = rng.standard_normal(2 * korder - 1)
    c0 = rng.standard_normal((n_rows, n_cols))
    da = torch.from_numpy(np.ascontiguousarray(a.T)).cuda().t()
    dg = torch.from_numpy(g).cuda()
    return a, g, c0, da, dg


def _dev_copy(x):
    import torch

    return torch.from_numpy(np.ascontiguousarray(x.T)).cuda().t()


@pytest.mark.parametrize("prec,slices", [(api.PREC_SPLIT_I8, 0), (api.PREC_SPLIT_I8, 14),
                                         (api.PREC_FP64, 0)])
def test_noalloc_matches_toeplitz_dev_on_two_calls(gpu, ora, prec, slices):
    import torch

    # first call: M = 1536, N = 1280, K = 1536 at offsets; the second one smaller
    shapes = [(1536, 1280, 1536, 2048, 300, 200, 1.25, 0.5),
              (700, 900, 1000, 1100, 37, 61, -0.75, 0.0)]
    cases = []
    for i, (m, n, k, korder, r0, c0, alpha, beta) in enumerate(shapes):
        a, g, cinit, da, dg = _case(m, n, k, korder, 10 + i)
        cases.append((a, g, cinit, da, dg, _dev_copy(cinit), _dev_copy(cinit)))
    # warm-up on a small product: loads the library's code objects (and the static scratch)
    wa, wg, wc, wda, wdg = _case(64, 64, 64, 64, 1)
    wdc = _dev_copy(wc)
    api.toeplitz_noalloc_dev(1.0, wda, wdg, 64, 64, wdc, 0.0, 0, 0, prec, slices)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for (m, n, k, korder, r0, c0, alpha, beta), case in zip(shapes, cases):
        a, g, cinit, da, dg, c_na, _ = case
        api.toeplitz_noalloc_dev(alpha, da, dg, korder, n, c_na, beta, r0, c0, prec, slices)
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0  # the two calls allocated nothing
    for (m, n, k, korder, r0, c0, alpha, beta), case in zip(shapes, cases):
        a, g, cinit, da, dg, c_na, c_ref = case
        api.toeplitz_dev(alpha, da, dg, korder, n, c_ref, beta, r0, c0, prec, slices)
        torch.cuda.synchronize()
        got, ref = c_na.cpu().numpy(), c_ref.cpu().numpy()
        assert np.array_equal(got, ref), (m, n, k)
        want = ora.structured_gemm_np(1, alpha, a, dict(g=g, korder=korder), r0, c0, n, beta,
                                      cinit)
        tol = 1e-13 if prec == api.PREC_FP64 else SPLIT_TOL
        assert ora.rel_fro(got, want) < tol, (m, n, k)


def test_noalloc_rejects_products_beyond_the_scratch(gpu):
    import torch

    m = n = k = 4096  # about 470 MB of int8 layout at t = 14 against 128 MiB of scratch
    da = torch.zeros((k, m), dtype=torch.float64, device="cuda").t()
    dg = torch.zeros(2 * k - 1, dtype=torch.float64, device="cuda")
    dc = torch.full((n, m), 3.0, dtype=torch.float64, device="cuda").t()
    with pytest.raises(api.PsmmaError, match="bytes of scratch"):
        api.toeplitz_noalloc_dev(1.0, da, dg, k, n, dc, 0.0, 0, 0, api.PREC_SPLIT_I8, 14)
    torch.cuda.synchronize()
    assert bool((dc == 3.0).all())  # rejected before any kernel ran
    with pytest.raises(api.PsmmaError, match="bf16"):
        api.toeplitz_noalloc_dev(1.0, da, dg, k, n, dc, 0.0, 0, 0, api.PREC_SPLIT_BF16, 6)
