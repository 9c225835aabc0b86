"""Argument checks of the device API that run before any HIP call (no GPU needed).

* the Python wrappers reject host tensors, wrong dtypes and malformed views with
  ValueError before the library is called;
* the C ABI rejects int8 split calls whose contraction exceeds the exact int32
  accumulation limit (single products and batches), and accepts slices = 0 as
  "the form's default".
"""
import ctypes

import numpy as np
import pytest
import torch

from psmma_git_amd import api

KMAX = 131071  # PSMMA_SPLIT_I8_KMAX


def _cm(m, n, dtype=torch.float64):
    return torch.zeros((n, m), dtype=dtype).t()


def test_wrappers_reject_host_tensors():
    a, c, g = _cm(8, 8), _cm(8, 8), torch.zeros(15, dtype=torch.float64)
    with pytest.raises(ValueError, match="device"):
        api.toeplitz_dev(1.0, a, g, 8, 8, c)
    with pytest.raises(ValueError, match="device"):
        api.hankel_dev(1.0, a, g, 8, c)
    with pytest.raises(ValueError, match="device"):
        api.cauchy_dev(1.0, a, g, g, g, g, 8, c)
    with pytest.raises(ValueError, match="device"):
        api.cauchyeig_dev(1.0, a, g, g, g, g, g, 8, c)


def test_wrappers_reject_wrong_dtypes():
    a32 = _cm(8, 8, torch.float32)
    c, g = _cm(8, 8), torch.zeros(15, dtype=torch.float64)
    with pytest.raises(ValueError, match="dtype"):
        api.toeplitz_dev(1.0, a32, g, 8, 8, c)
    b = _cm(8, 4)  # the DFT takes complex128
    with pytest.raises(ValueError, match="dtype"):
        api.dft_dev(1.0, b, 8, 8, _cm(8, 4, torch.complex128))
    with pytest.raises(ValueError, match="dtype"):
        api.rrlu_cauchy_batched_dev(torch.zeros((2, 4), dtype=torch.int64),
                                    torch.zeros((2, 4), dtype=torch.int32), g, g, g, g)


def _colmajor_host(t):
    """_colmajor without the device check (layout rules only)."""
    orig = api._on_device
    try:
        api._on_device = lambda t, name, dtype=None: t
        return api._colmajor(t, "a")
    finally:
        api._on_device = orig


def test_layout_rules():
    with pytest.raises(ValueError, match="column-major"):
        _colmajor_host(torch.zeros((8, 4), dtype=torch.float64))
    p, m, n, ld = _colmajor_host(_cm(8, 4))
    assert (m, n, ld) == (8, 4, 8)
    # a column view of a wider matrix keeps its leading dimension
    big = _cm(16, 4)
    p, m, n, ld = _colmajor_host(big[:8, :])
    assert (m, n, ld) == (8, 4, 16)
    # a single column: leading dimension is its row count
    p, m, n, ld = _colmajor_host(_cm(8, 1))
    assert ld == 8


def test_output_and_generator_sizes_checked():
    orig = api._on_device
    try:
        api._on_device = lambda t, name, dtype=None: t
        with pytest.raises(ValueError, match="smaller than"):
            api._out(_cm(7, 8), 8, 8)
        with pytest.raises(ValueError, match="smaller than"):
            api._out(_cm(8, 7), 8, 8)
        api._out(_cm(8, 8), 8, 8)
        with pytest.raises(ValueError, match="needs 9"):
            api._vec(torch.zeros(8, dtype=torch.float64), "d", 9)
        with pytest.raises(ValueError, match="1-D"):
            api._vec(torch.zeros((2, 4), dtype=torch.float64), "d", 1)
    finally:
        api._on_device = orig


def _fake(addr=256):
    return ctypes.c_void_p(addr)  # never dereferenced: the checks return first


@pytest.mark.parametrize("k", [KMAX + 1, 200000])
def test_split_i8_rejects_long_contractions(k):
    L = api.lib()
    st = L.psmma_toeplitz_dev(16, 16, k, 1.0, _fake(), 16, _fake(), 2 * k, k, 0, 0, 0.0, _fake(),
                              16, api.PREC_SPLIT_I8, 0, None)
    assert st == -3
    assert "exact-accumulation limit" in api.last_error()
    st = L.psmma_cauchy_dev(16, 16, k, 1.0, _fake(), 16, _fake(), _fake(), _fake(), _fake(), 0, 0,
                            0.0, _fake(), 16, api.PREC_SPLIT_I8, 13, None)
    assert st == -3


def test_batched_split_rejects_long_contractions():
    p = api.Problem(api.KIND_HANKEL, 16, 16, KMAX + 1, 1.0, 0.0, 256, 16, 256, 0, 0, 0, 0,
                    2 * (KMAX + 1) + 32, 0, 0, 0, 256, 16)
    arr = (api.Problem * 1)(p)
    st = api.lib().psmma_batched_dev(1, ctypes.cast(arr, ctypes.c_void_p), api.PREC_SPLIT_I8, 0,
                                     None)
    assert st == -1000


def test_default_slices_accepted():
    api.set_precision(api.PREC_SPLIT_I8, 0)
    api.set_precision(api.PREC_SPLIT_BF16, 0)
    api.set_precision(api.PREC_FP64, 0)
    with pytest.raises(api.PsmmaError):
        api.set_precision(api.PREC_SPLIT_I8, -1)
    with pytest.raises(api.PsmmaError):
        api.set_precision(api.PREC_SPLIT_I8, 17)
    assert np.isfinite(KMAX)
