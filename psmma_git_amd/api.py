"""Python mirror of the PSMMA entry points over libpsmma_hip.so.

Two levels, matching include/psmma.h:

* ``psmma_cauchy``, ``psmma_toeplitz``, ``psmma_hankel``, ``psmma_cauchyeig``,
  ``pzmdftl1`` / ``pzmdftl`` and the ``ps*_compute`` routines take the
  reference's Fortran argument lists (SRC/psmma_*.f90, SRC/pzmdftL1.f90) with
  host numpy arrays in Fortran order, and update the output array in place,
  exactly as a Fortran caller's arrays are updated.  Errors follow the
  reference's print-and-return convention; ``last_status()`` exposes the INFO
  code.
* ``*_dev`` functions take torch device tensors (column-major views are passed
  as their storage pointer plus leading dimension) and a stream, and raise
  ``PsmmaError`` on a nonzero status.

The library must have been built (``python -m psmma_git_amd.build``); nothing
here falls back to a CPU implementation.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libpsmma_hip.so"

PREC_FP64, PREC_SPLIT_I8, PREC_SPLIT_BF16, PREC_SPLIT_FP32 = 0, 1, 2, 3
KIND_CAUCHY, KIND_TOEPLITZ, KIND_HANKEL, KIND_CAUCHYEIG = 0, 1, 2, 3


class PsmmaError(RuntimeError):
    pass


_lib = None

_I = ctypes.c_int
_D = ctypes.c_double
_P = ctypes.c_void_p
_PI = ctypes.POINTER(ctypes.c_int)
_PD = ctypes.POINTER(ctypes.c_double)

# (name, argtypes, restype)
_SIGS = {
    "psmma_cauchy_": ([_PI, _PI, _PD, _P, _PI, _PI, _P, _P, _PI, _PD, _P, _PI, _PI, _P, _P, _PI], None),
    "psmma_toeplitz_": ([_PI, _PI, _PD, _P, _PI, _PI, _P, _P, _PI, _PD, _P, _PI, _PI, _P, _P, _PI], None),
    "psmma_hankel_": ([_PI, _PI, _PD, _P, _PI, _PI, _P, _P, _PI, _PD, _P, _PI, _PI, _P, _P, _PI], None),
    "psmma_cauchyeig_": ([_PI, _PI, _PI, _PD, _P, _PI, _PI, _P, _P, _P, _P, _P, _P, _P, _PD, _P,
                          _PI, _PI, _P, _P, _PI], None),
    "pzmdftl1_": ([_PI, _PI, _PI, _PI, _PD, _P, _PI, _PD, _P, _PI, _PI, _PI, _P, _P], None),
    "pzmdftl_": ([_PI, _PI, _PI, _PI, _PD, _P, _PI, _PD, _P, _PI, _PI, _PI, _P, _P], None),
    "pscauchy_compute_": ([_PI, _PI, _PD, _P, _PI, _P, _P, _PI, _PD, _P, _PI, _PI, _PI, _P], None),
    "pstoeplitz_compute_": ([_PI, _PI, _PD, _P, _PI, _P, _P, _PI, _PD, _P, _PI, _PI, _PI, _P], None),
    "pshankel_compute_": ([_PI, _PI, _PD, _P, _PI, _P, _P, _PI, _PD, _P, _PI, _PI, _PI, _P], None),
    "pscauchyeig_compute_": ([_PI, _PI, _PI, _PI, _PD, _P, _PI, _P, _P, _P, _P, _P, _P, _P, _PD,
                              _P, _PI, _PI, _PI, _P], None),
    "psmma_set_precision": ([_I, _I], _I),
    "psmma_last_status": ([], _I),
    "psmma_last_error": ([], ctypes.c_char_p),
    "psmma_version": ([], ctypes.c_char_p),
    "psmma_toeplitz_dev": ([_I, _I, _I, _D, _P, _I, _P, _I, _I, _I, _I, _D, _P, _I, _I, _I, _P], _I),
    "psmma_toeplitz_noalloc_dev": ([_I, _I, _I, _D, _P, _I, _P, _I, _I, _I, _I, _D, _P, _I, _I,
                                    _I, _P], _I),
    "psmma_hankel_dev": ([_I, _I, _I, _D, _P, _I, _P, _I, _I, _I, _D, _P, _I, _I, _I, _P], _I),
    "psmma_cauchy_dev": ([_I, _I, _I, _D, _P, _I, _P, _P, _P, _P, _I, _I, _D, _P, _I, _I, _I, _P], _I),
    "psmma_cauchyeig_dev": ([_I, _I, _I, _D, _P, _I, _P, _P, _P, _P, _P, _I, _I, _I, _D, _P, _I,
                             _I, _I, _P], _I),
    "pzmdft_dev": ([_I, _I, _I, _I, _D, _P, _I, _I, _I, _D, _P, _I, _I, _I, _P], _I),
    "psmma_secular_dev": ([_I, _P, _P, _D, _P, _P, _P, _P, _P], _I),
    "psmma_rrqr_batched_dev": ([_I, _I, _I, _P, _I, ctypes.c_longlong, _D, _I, _P, _P, _P, _I,
                                ctypes.c_longlong, _P, _I, ctypes.c_longlong, _P], _I),
    "psmma_rrlu_cauchyeig_batched_dev": ([_I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _I, _D, _I, _P,
                                          _P, _P, _I, ctypes.c_longlong, _P, _I,
                                          ctypes.c_longlong, _P], _I),
    "psmma_rrlu_cauchy_batched_dev": ([_I, _I, _I, _P, _P, _P, _P, _P, _P, _I, _I, _D, _I, _P, _P,
                                       _I, ctypes.c_longlong, _P, _I, ctypes.c_longlong, _P],
                                      _I),
    "psmma_cauchy_blr_dev": ([_I, _I, _D, _P, _I, _P, _P, _P, _P, _I, _D, _I, _D, _P, _I, _P, _P],
                             _I),
    "psmma_toeplitz_blr_dev": ([_I, _I, _D, _P, _I, _P, _I, _I, _I, _D, _D, _P, _I, _P, _P], _I),
    "psmma_hankel_blr_dev": ([_I, _I, _D, _P, _I, _P, _I, _I, _D, _D, _P, _I, _P, _P], _I),
    "psmma_cauchyeig_blr_ex_dev": ([_I, _I, _D, _P, _I, _P, _P, _P, _P, _P, _I, _D, _I, _D, _P,
                                    _I, _P, _P], _I),
    "psmma_cauchyeig_blr_dev": ([_I, _I, _D, _P, _I, _P, _P, _P, _P, _P, _I, _D, _D, _P, _I, _P,
                                 _P], _I),
    "psmma_batched_dev": ([_I, _P, _I, _I, _P], _I),
    "psmma_batched_plan_create": ([_I, _P, _I, _I, ctypes.POINTER(ctypes.c_void_p)], _I),
    "psmma_batched_plan_launch": ([_P, _P], _I),
    "psmma_batched_plan_info": ([_P, _PI, _PI], _I),
    "psmma_batched_plan_destroy": ([_P], _I),
    "psmma_autotune_f64": ([_I, _I, _I, _I, _P, _P], _I),
    "psmma_split_slices_dev": ([_P, _I, _D, _P, _P, _P], _I),
    "psmma_hss_build_dev": ([_I, _I, _P, _P, _P, _P, _P, _I, _D, _I, ctypes.c_ulonglong, _P, _P,
                             _P], _I),
    "psmma_hss_product_dev": ([_P, _I, _D, _P, _I, _D, _P, _I, _P], _I),
    "psmma_hss_layout": ([_P, _P, _P, _P, _P, _P, _P, _P, _P, _P], _I),
    "psmma_hss_destroy": ([_P], _I),
    "psmma_constindex": ([_I, _I, _I, _I, _P], _I),
    "psmma_is_intersect": ([_I, _P, _I, _P], _I),
    "psmma_ring_step_dev": ([_P, _P, _P, _P], _I),
    "psmma_local_offsets": ([_I, _I, _I, _I, _I, _I, _P], ctypes.c_longlong),
    "psmma_redistribute_dev": ([_I, _I, _I, _I, _I, _I, _P, _I, _I, _P, _P], _I),
}


def lib() -> ctypes.CDLL:
    """Load libpsmma_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise PsmmaError(f"{LIB_PATH} is missing: build it with `python -m psmma_git_amd.build`")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def declared_symbols() -> list[str]:
    return list(_SIGS)


# ---------------------------------------------------------------------------
# Fortran-list entry points on host arrays.

def _i(v):
    return ctypes.byref(ctypes.c_int(int(v)))


def _d(v):
    return ctypes.byref(ctypes.c_double(float(v)))


def _arr(x, dtype=np.float64):
    if x is None:
        return None
    if not isinstance(x, np.ndarray) or x.dtype != dtype:
        raise TypeError(f"expected a numpy {dtype} array")
    if x.ndim > 1 and not x.flags.f_contiguous:
        raise TypeError("arrays must be Fortran-ordered (column-major)")
    return x.ctypes.data_as(ctypes.c_void_p)


def _desc(desc):
    d = np.ascontiguousarray(desc, dtype=np.int32)
    if d.size < 9:
        raise ValueError("a ScaLAPACK descriptor has 9 entries")
    return d, d.ctypes.data_as(ctypes.c_void_p)


def descinit(m: int, n: int, mb: int, nb: int, lld: int, ictxt: int = 0) -> np.ndarray:
    """ScaLAPACK DESCINIT for a 1 x 1 grid (DTYPE=1, RSRC=CSRC=0)."""
    return np.array([1, ictxt, m, n, mb, nb, 0, 0, max(1, lld)], dtype=np.int32)


def last_status() -> int:
    return int(lib().psmma_last_status())


def last_error() -> str:
    return lib().psmma_last_error().decode()


def set_precision(prec: int, slices: int = 0) -> None:
    st = lib().psmma_set_precision(int(prec), int(slices))
    if st != 0:
        raise PsmmaError(f"psmma_set_precision({prec}, {slices}) -> {st}")


def _right_entry(name, N, K, ALPHA, A, IA, JA, DESCA, B, LDB, BETA, Q, IQ, JQ, DESCQ, WORK, REDIST):
    da, pa = _desc(DESCA)
    dq, pq = _desc(DESCQ)
    getattr(lib(), name)(_i(N), _i(K), _d(ALPHA), _arr(A), _i(IA), _i(JA), pa, _arr(B), _i(LDB),
                         _d(BETA), _arr(Q), _i(IQ), _i(JQ), pq, _arr(WORK), _i(1 if REDIST else 0))


def psmma_cauchy(N, K, ALPHA, A, IA, JA, DESCA, B, LDB, BETA, Q, IQ, JQ, DESCQ, WORK=None,
                 REDIST=False):
    """PSMMA_CAUCHY (SRC/psmma_cauchy.f90:2-3): Q = ALPHA*A*H + BETA*Q, B(LDB,4)=[D F U V]."""
    _right_entry("psmma_cauchy_", N, K, ALPHA, A, IA, JA, DESCA, B, LDB, BETA, Q, IQ, JQ, DESCQ,
                 WORK, REDIST)


def psmma_toeplitz(N, K, ALPHA, A, IA, JA, DESCA, B, LDB, BETA, Q, IQ, JQ, DESCQ, WORK=None,
                   REDIST=False):
    """PSMMA_TOEPLITZ (SRC/psmma_toeplitz.f90:2-3): H(i,j) = B(K+j-i)."""
    _right_entry("psmma_toeplitz_", N, K, ALPHA, A, IA, JA, DESCA, B, LDB, BETA, Q, IQ, JQ, DESCQ,
                 WORK, REDIST)


def psmma_hankel(N, K, ALPHA, A, IA, JA, DESCA, B, LDB, BETA, Q, IQ, JQ, DESCQ, WORK=None,
                 REDIST=False):
    """PSMMA_HANKEL (SRC/psmma_hankel.f90:2-3): H(i,j) = B(i+j-1)."""
    _right_entry("psmma_hankel_", N, K, ALPHA, A, IA, JA, DESCA, B, LDB, BETA, Q, IQ, JQ, DESCQ,
                 WORK, REDIST)


def psmma_cauchyeig(N, K, ID, ALPHA, A, IA, JA, DESCA, F, D, U, V, DIFL, DIFR, BETA, Q, IQ, JQ,
                    DESCQ, WORK=None, REDIST=False):
    """PSMMA_CAUCHYEIG (SRC/psmma_cauchyeig.f90:2-3): H(i,j) = U_i V_j / (F_i - D_j)."""
    da, pa = _desc(DESCA)
    dq, pq = _desc(DESCQ)
    lib().psmma_cauchyeig_(_i(N), _i(K), _i(ID), _d(ALPHA), _arr(A), _i(IA), _i(JA), pa, _arr(F),
                           _arr(D), _arr(U), _arr(V), _arr(DIFL), _arr(DIFR), _d(BETA), _arr(Q),
                           _i(IQ), _i(JQ), pq, _arr(WORK), _i(1 if REDIST else 0))


def _dft_entry(name, N, K, NB, KB, ALPHA, B, LDB, BETA, C, LDC, IMROW=0, IMCOL=0, WORK=None,
               RWORK=None):
    getattr(lib(), name)(_i(N), _i(K), _i(NB), _i(KB), _d(ALPHA), _arr(B, np.complex128), _i(LDB),
                         _d(BETA), _arr(C, np.complex128), _i(LDC), _i(IMROW), _i(IMCOL),
                         None, None)


def pzmdftl1(N, K, NB, KB, ALPHA, B, LDB, BETA, C, LDC, IMROW=0, IMCOL=0, WORK=None, RWORK=None):
    """PZMDFTL1 (SRC/pzmdftL1.f90:1-2): C = ALPHA*F*B + BETA*C, F the unnormalised DFT."""
    _dft_entry("pzmdftl1_", N, K, NB, KB, ALPHA, B, LDB, BETA, C, LDC, IMROW, IMCOL)


def pzmdftl(N, K, NB, KB, ALPHA, B, LDB, BETA, C, LDC, IMROW=0, IMCOL=0, WORK=None, RWORK=None):
    """PZMDFTL (SRC/pzmdftL.f90:1-2), SVD variant: same product on one process."""
    _dft_entry("pzmdftl_", N, K, NB, KB, ALPHA, B, LDB, BETA, C, LDC, IMROW, IMCOL)


def _compute_entry(name, N, K, ALPHA, A, LDA, DESCA, B, LDB, BETA, C, LDC, IMROW=0, IMCOL=0):
    da, pa = _desc(DESCA)
    getattr(lib(), name)(_i(N), _i(K), _d(ALPHA), _arr(A), _i(LDA), pa, _arr(B), _i(LDB),
                         _d(BETA), _arr(C), _i(LDC), _i(IMROW), _i(IMCOL), None)


def pscauchy_compute(*args, **kw):
    """PSCAUCHY_COMPUTE (SRC/pscauchy_compute.f90:2-3)."""
    _compute_entry("pscauchy_compute_", *args, **kw)


def pstoeplitz_compute(*args, **kw):
    """PSTOEPLITZ_COMPUTE (SRC/pstoeplitz_compute.f90:2-3)."""
    _compute_entry("pstoeplitz_compute_", *args, **kw)


def pshankel_compute(*args, **kw):
    """PSHANKEL_COMPUTE (SRC/pshankel_compute.f90:2-3)."""
    _compute_entry("pshankel_compute_", *args, **kw)


def pscauchyeig_compute(N, K, NB, KB, ALPHA, A, LDA, DESCA, F, D, U, V, DIFL, DIFR, BETA, C, LDC,
                        IMROW=0, IMCOL=0):
    """PSCAUCHYEIG_COMPUTE (SRC/pscauchyeig_compute.f90:2-3)."""
    da, pa = _desc(DESCA)
    lib().pscauchyeig_compute_(_i(N), _i(K), _i(NB), _i(KB), _d(ALPHA), _arr(A), _i(LDA), pa,
                               _arr(F), _arr(D), _arr(U), _arr(V), _arr(DIFL), _arr(DIFR),
                               _d(BETA), _arr(C), _i(LDC), _i(IMROW), _i(IMCOL), None)


# ---------------------------------------------------------------------------
# Device API on torch tensors.

def _check(st: int, what: str):
    if st != 0:
        raise PsmmaError(f"{what} returned {st}: {last_error()}")


def _on_device(t, name, dtype=None):
    """Checks that t is a torch tensor on a HIP device with the given dtype (float64 default)."""
    import torch
    dtype = dtype or torch.float64
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name}: expected a torch tensor, got {type(t).__name__}")
    if t.dtype != dtype:
        raise ValueError(f"{name}: dtype {t.dtype}, expected {dtype}")
    if t.device.type != "cuda":
        raise ValueError(f"{name}: expected a device (cuda/HIP) tensor, got one on {t.device}")
    return t


def _colmajor(t, name="matrix", dtype=None):
    """(pointer, rows, cols, ld) of a column-major 2-D device tensor view."""
    _on_device(t, name, dtype)
    if t.dim() != 2 or (t.stride(0) != 1 and t.shape[0] > 1):
        raise ValueError(f"{name}: expected a column-major 2-D tensor (stride(0) == 1)")
    m, n = t.shape
    if n > 1 and t.stride(1) < m:
        raise ValueError(f"{name}: column stride {t.stride(1)} is below the row count {m}")
    return t.data_ptr(), m, n, max(1, m, t.stride(1) if n > 1 else 1)


def _out(c, m, n, name="c", dtype=None):
    """Output view: column-major with at least m rows and n columns."""
    pc, mc, nc, ldc = _colmajor(c, name, dtype)
    if mc < m or nc < n:
        raise ValueError(f"{name}: {mc} x {nc} is smaller than the {m} x {n} product")
    return pc, ldc


def _vec(t, name, need=0, dtype=None):
    """Pointer of a contiguous 1-D device vector with at least `need` entries."""
    _on_device(t, name, dtype)
    if t.dim() != 1 or not t.is_contiguous():
        raise ValueError(f"{name}: expected a contiguous 1-D tensor")
    if t.numel() < need:
        raise ValueError(f"{name}: {t.numel()} entries, the block needs {need}")
    return t.data_ptr()


def _stream_ptr(stream):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def colmajor_empty(m: int, n: int, dtype=None, device="cuda"):
    """Allocate an m x n column-major tensor (a transposed row-major buffer)."""
    import torch
    return torch.empty((n, m), dtype=dtype or torch.float64, device=device).t()


# slices = 0: the numeric form's default (int8: the fewest moduli that keep 45 bits per
# operand at the call's K; bf16: 6 slices).


def toeplitz_dev(alpha, a, g, korder, n, c, beta=0.0, r0=0, c0=0, prec=PREC_FP64, slices=0,
                 stream=None):
    """C(m x n) = alpha * A(m x k) * T(r0:r0+k, c0:c0+n) + beta * C, T(r,c) = g[korder-1+c-r]."""
    pa, m, k, lda = _colmajor(a, "a")
    pc, ldc = _out(c, m, n)
    _check(lib().psmma_toeplitz_dev(m, n, k, alpha, pa, lda, _vec(g, "g"), g.numel(), korder, r0,
                                    c0, beta, pc, ldc, prec, slices, _stream_ptr(stream)),
           "psmma_toeplitz_dev")


NOALLOC_SCRATCH_BYTES = 128 << 20


def toeplitz_noalloc_dev(alpha, a, g, korder, n, c, beta=0.0, r0=0, c0=0, prec=PREC_SPLIT_I8,
                         slices=0, stream=None):
    """psmma_toeplitz_noalloc_dev: toeplitz_dev without any allocation (FP64 or the int8
    form in the library's static per-GPU scratch array)."""
    pa, m, k, lda = _colmajor(a, "a")
    pc, ldc = _out(c, m, n)
    _check(lib().psmma_toeplitz_noalloc_dev(m, n, k, alpha, pa, lda, _vec(g, "g"), g.numel(),
                                            korder, r0, c0, beta, pc, ldc, prec, slices,
                                            _stream_ptr(stream)), "psmma_toeplitz_noalloc_dev")


def hankel_dev(alpha, a, g, n, c, beta=0.0, r0=0, c0=0, prec=PREC_FP64, slices=0, stream=None):
    """C(m x n) = alpha * A(m x k) * H(r0:r0+k, c0:c0+n) + beta * C, H(r,c) = g[r+c]."""
    pa, m, k, lda = _colmajor(a, "a")
    pc, ldc = _out(c, m, n)
    _check(lib().psmma_hankel_dev(m, n, k, alpha, pa, lda, _vec(g, "g"), g.numel(), r0, c0, beta,
                                  pc, ldc, prec, slices, _stream_ptr(stream)), "psmma_hankel_dev")


def cauchy_dev(alpha, a, d, f, u, v, n, c, beta=0.0, r0=0, c0=0, prec=PREC_FP64, slices=0,
               stream=None):
    """C(m x n) = alpha * A(m x k) * H(r0:r0+k, c0:c0+n) + beta * C, H(r,c) = u_r v_c/(d_r-f_c)."""
    pa, m, k, lda = _colmajor(a, "a")
    pc, ldc = _out(c, m, n)
    rows, cols = (r0 + k, c0 + n) if m and n and k else (0, 0)
    _check(lib().psmma_cauchy_dev(m, n, k, alpha, pa, lda, _vec(d, "d", rows), _vec(f, "f", cols),
                                  _vec(u, "u", rows), _vec(v, "v", cols), r0, c0, beta, pc, ldc,
                                  prec, slices, _stream_ptr(stream)), "psmma_cauchy_dev")


def cauchyeig_dev(alpha, a, fpole, u, v, difl, difr, n, c, beta=0.0, r0=0, c0=0, prec=PREC_FP64,
                  slices=0, stream=None):
    """C(m x n) = alpha * A(m x k) * H(r0:r0+k, c0:c0+n) + beta * C, H the eigen-Cauchy matrix
    of order fpole.numel()."""
    pa, m, k, lda = _colmajor(a, "a")
    pc, ldc = _out(c, m, n)
    nglob = fpole.numel() if hasattr(fpole, "numel") else 0
    ptrs = [_vec(t, nm, nglob) for t, nm in
            ((fpole, "fpole"), (u, "u"), (v, "v"), (difl, "difl"), (difr, "difr"))]
    _check(lib().psmma_cauchyeig_dev(m, n, k, alpha, pa, lda, *ptrs, nglob, r0, c0, beta, pc, ldc,
                                     prec, slices, _stream_ptr(stream)), "psmma_cauchyeig_dev")


def dft_dev(alpha, b, nd, m, c, beta=0.0, r0=0, c0=0, prec=PREC_FP64, slices=0, stream=None):
    """C(m x n) = alpha * F(r0:r0+m, c0:c0+k) * B + beta*C; b, c complex128 column-major.
    prec: PREC_FP64 or PREC_SPLIT_FP32 (slices 1..6, 0 = 5)."""
    import torch
    pb, k, n, ldb = _colmajor(b, "b", torch.complex128)
    pc, ldc = _out(c, m, n, "c", torch.complex128)
    _check(lib().pzmdft_dev(m, n, k, nd, alpha, pb, ldb, r0, c0, beta, pc, ldc, prec, slices,
                            _stream_ptr(stream)), "pzmdft_dev")


def secular_dev(d, z, rho=1.0, loewner=True, stream=None):
    """Generators of the eigen-Cauchy operand of D + rho z z^T (d strictly increasing,
    float64 device tensors).  Returns (difl, difr, u, v): u is Loewner's z when
    loewner is true, else z itself; v the column normalisers for that u."""
    import torch
    n = d.numel()
    _vec(d, "d")
    _vec(z, "z", n)
    difl = torch.empty_like(d)
    difr = torch.empty_like(d)
    zhat = torch.empty_like(d) if loewner else None
    v = torch.empty_like(d)
    _check(lib().psmma_secular_dev(n, d.data_ptr(), z.data_ptr(), float(rho), difl.data_ptr(),
                                   difr.data_ptr(), zhat.data_ptr() if loewner else None,
                                   v.data_ptr(), _stream_ptr(stream)), "psmma_secular_dev")
    return difl, difr, (zhat if loewner else z), v


RRQR_KMAX = 32


BLR_RRQR, BLR_RRLU = 0, 1


def cauchyeig_blr_dev(alpha, a, fpole, u, v, difl, difr, c, leaf=256, tol=1e-12, beta=0.0,
                      stream=None, compressor=None):
    """X = alpha * A * C + beta * X with the off-diagonal leaf blocks of the eigen-Cauchy
    matrix C compressed by the batched RRQR (default: psmma_cauchyeig_blr_dev) or by the
    compressor given (BLR_RRQR / BLR_RRLU: psmma_cauchyeig_blr_ex_dev).  a, c:
    column-major (m x n) device views.  Returns the stats dict (blocks, lowrank, dense,
    zero, rank_sum)."""
    pa, m, n, lda = _colmajor(a, "a")
    pc, ldc = _out(c, m, n)
    gp = [_vec(t, nm, n) for t, nm in
          ((fpole, "fpole"), (u, "u"), (v, "v"), (difl, "difl"), (difr, "difr"))]
    st = (ctypes.c_longlong * 5)()
    if compressor is None:
        _check(lib().psmma_cauchyeig_blr_dev(m, n, alpha, pa, lda, *gp, leaf, float(tol), beta, pc,
                                             ldc, st, _stream_ptr(stream)),
               "psmma_cauchyeig_blr_dev")
    else:
        _check(lib().psmma_cauchyeig_blr_ex_dev(m, n, alpha, pa, lda, *gp, leaf, float(tol),
                                                int(compressor), beta, pc, ldc, st,
                                                _stream_ptr(stream)),
               "psmma_cauchyeig_blr_ex_dev")
    return dict(zip(("blocks", "lowrank", "dense", "zero", "rank_sum"), list(st)))


RRLU_TOL = float(np.float32(1.0e-17))  # the reference's default-real TOL


def _index_lists(pl, pu, nblk):
    """pl (nblk, m) and pu (nblk, n): contiguous int32 device tensors of 1-based indices."""
    import torch
    for t, nm in ((pl, "pl"), (pu, "pu")):
        _on_device(t, nm, torch.int32)
        if t.dim() != 2 or t.shape[0] != nblk or not t.is_contiguous():
            raise ValueError(f"{nm}: expected a contiguous ({nblk}, width) int32 tensor")


def cauchy_blr_dev(alpha, a, d, f, u, v, c, leaf=256, tol=1e-12, beta=0.0, stream=None,
                   compressor=1):
    """X = alpha * A * C + beta * X, C(i,j) = u_i v_j / (d_i - f_j), with the off-diagonal leaf
    blocks compressed by the RRLU (BLR_RRLU, default) or the RRQR (BLR_RRQR).  a, c:
    column-major (m x n) device views; d, f, u, v: float64 device vectors of length n.
    Returns the stats dict (blocks, lowrank, dense, zero, rank_sum)."""
    pa, m, n, lda = _colmajor(a, "a")
    pc, ldc = _out(c, m, n)
    gp = [_vec(t, nm, n) for t, nm in ((d, "d"), (f, "f"), (u, "u"), (v, "v"))]
    st = (ctypes.c_longlong * 5)()
    _check(lib().psmma_cauchy_blr_dev(m, n, alpha, pa, lda, *gp, leaf, float(tol),
                                      int(compressor), beta, pc, ldc, st, _stream_ptr(stream)),
           "psmma_cauchy_blr_dev")
    return dict(zip(("blocks", "lowrank", "dense", "zero", "rank_sum"), list(st)))


def toeplitz_blr_dev(alpha, a, g, korder, c, leaf=256, tol=1e-13, beta=0.0, stream=None):
    """X = alpha * A * H + beta * X, H(r, c) = g[korder - 1 + c - r] (n x n, n = a's
    columns), off-diagonal leaf blocks compressed by the batched RRQR from the generator
    (ConstToeplitzlowrank -> Compress_RRQR, SRC/Auxil_2p5D.f90:968-1189).  a, c:
    column-major (m x n) device views; g: float64 device vector.  Returns the stats dict."""
    pa, m, n, lda = _colmajor(a, "a")
    pc, ldc = _out(c, m, n)
    pg = _vec(g, "g", korder - 1 + n)
    st = (ctypes.c_longlong * 5)()
    _check(lib().psmma_toeplitz_blr_dev(m, n, alpha, pa, lda, pg, int(g.numel()), int(korder),
                                        leaf, float(tol), beta, pc, ldc, st,
                                        _stream_ptr(stream)), "psmma_toeplitz_blr_dev")
    return dict(zip(("blocks", "lowrank", "dense", "zero", "rank_sum"), list(st)))


def hankel_blr_dev(alpha, a, g, c, leaf=256, tol=1e-13, beta=0.0, stream=None):
    """X = alpha * A * H + beta * X, H(r, c) = g[r + c] (n x n), off-diagonal leaf blocks
    compressed by the batched RRQR from the generator.  Returns the stats dict."""
    pa, m, n, lda = _colmajor(a, "a")
    pc, ldc = _out(c, m, n)
    pg = _vec(g, "g", 2 * n - 1)
    st = (ctypes.c_longlong * 5)()
    _check(lib().psmma_hankel_blr_dev(m, n, alpha, pa, lda, pg, int(g.numel()), leaf,
                                      float(tol), beta, pc, ldc, st, _stream_ptr(stream)),
           "psmma_hankel_blr_dev")
    return dict(zip(("blocks", "lowrank", "dense", "zero", "rank_sum"), list(st)))


def rrlu_cauchyeig_batched_dev(pl, pu, fpole, u, v, difl, difr, tol=RRLU_TOL, kmax=32,
                               stream=None, out=None):
    """Batched generator-only RRLU of eigen-Cauchy blocks (Compress_cauchyEig2).  pl
    (nblk, m) / pu (nblk, n): int32 device tensors of 1-based global rows / columns;
    generators: float64 device vectors of the matrix order.  Returns (rank, nswap, q,
    y): rank, nswap (nblk,) int32; q (nblk, kq, m) and y (nblk, n, kq) hold, as
    column-major views q[b].T and y[b].T, Q (m x rank) and Y (rank x n) in their
    leading columns / rows when 0 < rank < min(m, n) and rank <= kmax,
    kq = min(kmax, m, n).  out: a previous call's (rank, nswap, q, y) to reuse."""
    import torch
    nblk, m = pl.shape
    n = pu.shape[1]
    _index_lists(pl, pu, nblk)
    korder = fpole.numel() if hasattr(fpole, "numel") else 0
    for t, nm in ((fpole, "fpole"), (u, "u"), (v, "v"), (difl, "difl"), (difr, "difr")):
        _vec(t, nm, korder)
    kq = max(1, min(kmax, m, n))
    dev = pl.device
    if out is not None:
        rank, nswap, q, y = out
        if q.shape != (nblk, kq, m) or y.shape != (nblk, n, kq):
            raise ValueError("out buffers do not match the blocks")
    else:
        rank = torch.empty(nblk, dtype=torch.int32, device=dev)
        nswap = torch.empty(nblk, dtype=torch.int32, device=dev)
        q = torch.zeros((nblk, kq, m), dtype=torch.float64, device=dev)
        y = torch.zeros((nblk, n, kq), dtype=torch.float64, device=dev)
    _check(lib().psmma_rrlu_cauchyeig_batched_dev(
        nblk, m, n, pl.data_ptr(), pu.data_ptr(), fpole.data_ptr(), u.data_ptr(), v.data_ptr(),
        difl.data_ptr(), difr.data_ptr(), fpole.numel(), float(tol), kmax, rank.data_ptr(),
        nswap.data_ptr(), q.data_ptr(), m, kq * m, y.data_ptr(), kq, n * kq,
        _stream_ptr(stream)), "psmma_rrlu_cauchyeig_batched_dev")
    return rank, nswap, q, y


def rrlu_cauchy_batched_dev(pl, pu, d, f, u, v, tol=RRLU_TOL, kmax=32, stream=None, out=None):
    """Batched generator-only RRLU of plain Cauchy blocks (compress_cauchy('R')).  pl (nblk,
    m) / pu (nblk, n): int32 device tensors of 1-based global rows / columns; d, u (rows)
    and f, v (columns): float64 device vectors.  Returns (rank, x, y) with x (nblk, kq, m)
    and y (nblk, n, kq) holding, as column-major views x[b].T and y[b].T, X (m x rank)
    and Y (rank x n) when 0 < rank < m and rank <= kmax, kq = min(kmax, m, n).  out: a
    previous call's (rank, x, y) to reuse."""
    import torch
    nblk, m = pl.shape
    n = pu.shape[1]
    _index_lists(pl, pu, nblk)
    _vec(d, "d"), _vec(u, "u", d.numel()), _vec(f, "f"), _vec(v, "v", f.numel())
    kq = max(1, min(kmax, m, n))
    dev = pl.device
    if out is not None:
        rank, x, y = out
        if x.shape != (nblk, kq, m) or y.shape != (nblk, n, kq):
            raise ValueError("out buffers do not match the blocks")
    else:
        rank = torch.empty(nblk, dtype=torch.int32, device=dev)
        x = torch.zeros((nblk, kq, m), dtype=torch.float64, device=dev)
        y = torch.zeros((nblk, n, kq), dtype=torch.float64, device=dev)
    _check(lib().psmma_rrlu_cauchy_batched_dev(
        nblk, m, n, pl.data_ptr(), pu.data_ptr(), d.data_ptr(), f.data_ptr(), u.data_ptr(),
        v.data_ptr(), d.numel(), f.numel(), float(tol), kmax, rank.data_ptr(), x.data_ptr(), m,
        kq * m, y.data_ptr(), kq, n * kq, _stream_ptr(stream)), "psmma_rrlu_cauchy_batched_dev")
    return rank, x, y


def rrqr_batched_dev(blocks, tol, kmax=RRQR_KMAX, stream=None):
    """Batched truncated RRQR of the blocks of a (nblk, n, m) float64 device tensor (block b
    is the column-major m x n matrix blocks[b].T).  Returns (rank, jpvt, q, rpt): rank
    (nblk,) int32; jpvt (nblk, n) 1-based; q (nblk, kq, m) and rpt (nblk, n, kq) hold, as
    column-major views q[b].T and rpt[b].T, Q (m x rank) and R P^T (rank x n) in their
    leading columns / rows, kq = min(kmax, m, n)."""
    import torch
    _on_device(blocks, "blocks")
    if blocks.dim() != 3:
        raise ValueError("blocks: expected a (nblk, n, m) tensor")
    nblk, n, m = blocks.shape
    if not blocks.is_contiguous():
        raise ValueError("blocks must be contiguous (nblk, n, m)")
    kq = max(1, min(kmax, m, n))
    dev = blocks.device
    rank = torch.empty(nblk, dtype=torch.int32, device=dev)
    jpvt = torch.empty((nblk, n), dtype=torch.int32, device=dev)
    q = torch.zeros((nblk, kq, m), dtype=torch.float64, device=dev)
    rpt = torch.zeros((nblk, n, kq), dtype=torch.float64, device=dev)
    _check(lib().psmma_rrqr_batched_dev(nblk, m, n, blocks.data_ptr(), max(1, m), max(1, m) * n,
                                        float(tol), kmax, rank.data_ptr(), jpvt.data_ptr(),
                                        q.data_ptr(), max(1, m), kq * m, rpt.data_ptr(), kq,
                                        kq * n, _stream_ptr(stream)), "psmma_rrqr_batched_dev")
    return rank, jpvt, q, rpt


# ---------------------------------------------------------------------------
# Batched uneven products (include/psmma.h section 3).

class Problem(ctypes.Structure):
    """psmma_problem: C = alpha * A * H(r0:r0+k, c0:c0+n) + beta * C on device pointers."""
    _fields_ = [
        ("kind", ctypes.c_int), ("m", ctypes.c_int), ("n", ctypes.c_int), ("k", ctypes.c_int),
        ("alpha", ctypes.c_double), ("beta", ctypes.c_double),
        ("a", ctypes.c_void_p), ("lda", ctypes.c_int),
        ("g0", ctypes.c_void_p), ("g1", ctypes.c_void_p), ("g2", ctypes.c_void_p),
        ("g3", ctypes.c_void_p), ("g4", ctypes.c_void_p),
        ("lg", ctypes.c_int), ("korder", ctypes.c_int), ("r0", ctypes.c_int), ("c0", ctypes.c_int),
        ("c", ctypes.c_void_p), ("ldc", ctypes.c_int),
    ]


def make_problem(kind, alpha, a, gens, n, c, beta=0.0, r0=0, c0=0, lg=0, korder=0):
    """Problem from torch tensors: a, c column-major; gens: list of up to five 1-D tensors."""
    pa, m, k, lda = _colmajor(a)
    pc, _, _, ldc = _colmajor(c)
    g = [t.data_ptr() for t in gens] + [0] * (5 - len(gens))
    return Problem(kind, m, n, k, float(alpha), float(beta), pa, lda, *g, int(lg), int(korder),
                   int(r0), int(c0), pc, ldc)


def batched_dev(problems, prec=PREC_FP64, slices=0, stream=None):
    arr = (Problem * len(problems))(*problems)
    _check(lib().psmma_batched_dev(len(problems), ctypes.cast(arr, ctypes.c_void_p), prec, slices,
                                   _stream_ptr(stream)), "psmma_batched_dev")


class BatchedPlan:
    """hipGraph-captured batch; launch() replays it on a stream."""

    def __init__(self, problems, prec=PREC_FP64, slices=0):
        arr = (Problem * len(problems))(*problems)
        h = ctypes.c_void_p()
        _check(lib().psmma_batched_plan_create(len(problems), ctypes.cast(arr, ctypes.c_void_p),
                                               prec, slices, ctypes.byref(h)),
               "psmma_batched_plan_create")
        self._h = h

    def launch(self, stream=None):
        _check(lib().psmma_batched_plan_launch(self._h, _stream_ptr(stream)),
               "psmma_batched_plan_launch")

    def info(self):
        cfg, nodes = ctypes.c_int(), ctypes.c_int()
        lib().psmma_batched_plan_info(self._h, ctypes.byref(cfg), ctypes.byref(nodes))
        return cfg.value, nodes.value

    def close(self):
        if self._h:
            lib().psmma_batched_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# The reference's P x Q grid, one rank's share at a time (include/psmma.h 3b).

def constindex(n: int, nb: int, mycol: int, npcol: int) -> np.ndarray:
    """ConstIndex (SRC/Auxil_2p5D.f90:2-66): 1-based global indices of process column mycol."""
    cnt = lib().psmma_constindex(n, nb, mycol, npcol, None)
    if cnt < 0:
        raise ValueError(f"constindex({n}, {nb}, {mycol}, {npcol}): invalid arguments")
    out = np.zeros(max(cnt, 1), dtype=np.int32)
    lib().psmma_constindex(n, nb, mycol, npcol, out.ctypes.data_as(ctypes.c_void_p))
    return out[:cnt]


def is_intersect(a, b) -> bool:
    """is_intersect (SRC/Auxil_2p5D.f90:752-792), the reference's sense: True iff disjoint."""
    a = np.ascontiguousarray(a, dtype=np.int32)
    b = np.ascontiguousarray(b, dtype=np.int32)
    return bool(lib().psmma_is_intersect(a.size, a.ctypes.data_as(ctypes.c_void_p), b.size,
                                         b.ctypes.data_as(ctypes.c_void_p)))


def ring_step_dev(problem: "Problem", rindex, cindex, stream=None):
    """One ring step C = alpha * A * H(rindex, cindex) + beta * C (problem.r0 / c0 ignored);
    rindex, cindex: 1-based global indices (host arrays)."""
    r = np.ascontiguousarray(rindex, dtype=np.int32)
    c = np.ascontiguousarray(cindex, dtype=np.int32)
    if r.size != problem.k or c.size != problem.n:
        raise ValueError("rindex / cindex lengths must equal the step's k / n")
    _check(lib().psmma_ring_step_dev(ctypes.byref(problem), r.ctypes.data_as(ctypes.c_void_p),
                                     c.ctypes.data_as(ctypes.c_void_p), _stream_ptr(stream)),
           "psmma_ring_step_dev")


def local_offsets(m, n, nprow, npcol, mb, nb) -> np.ndarray:
    """Offsets (elements) of the P*Q local arrays of a block-cyclic layout, rank-major."""
    out = np.zeros(nprow * npcol + 1, dtype=np.int64)
    tot = lib().psmma_local_offsets(m, n, nprow, npcol, mb, nb, out.ctypes.data_as(ctypes.c_void_p))
    if tot < 0:
        raise ValueError("local_offsets: invalid arguments")
    return out


def redistribute_dev(m, n, nprow, npcol, src_mb, src_nb, a, dst_mb, dst_nb, b, stream=None):
    """PDGEMR2D between two block-cyclic layouts of one grid; a, b: 1-D float64 device tensors
    holding all ranks' local arrays (local_offsets order)."""
    na = int(local_offsets(m, n, nprow, npcol, src_mb, src_nb)[-1])
    nb_ = int(local_offsets(m, n, nprow, npcol, dst_mb, dst_nb)[-1])
    _check(lib().psmma_redistribute_dev(m, n, nprow, npcol, src_mb, src_nb, _vec(a, "a", na),
                                        dst_mb, dst_nb, _vec(b, "b", nb_), _stream_ptr(stream)),
           "psmma_redistribute_dev")


# ---------------------------------------------------------------------------
# Autotuner (include/psmma.h 3c).

def autotune_f64(kind: int, m: int, n: int, k: int):
    """Times the FP64 tile configurations for this (kind, size bucket) and caches the fastest.
    Returns (cfg, ms)."""
    cfg, ms = ctypes.c_int(-1), ctypes.c_float(0.0)
    _check(lib().psmma_autotune_f64(kind, m, n, k, ctypes.byref(cfg), ctypes.byref(ms)),
           "psmma_autotune_f64")
    return cfg.value, ms.value


def split_slices_dev(problem: "Problem", prec=PREC_SPLIT_I8, target=1e-12, stream=None):
    """Fewest moduli / slices whose error on a sampled column block is <= target.
    Returns (slices, measured error, reached)."""
    t, err = ctypes.c_int(0), ctypes.c_double(0.0)
    st = lib().psmma_split_slices_dev(ctypes.byref(problem), prec, float(target), ctypes.byref(t),
                                      ctypes.byref(err), _stream_ptr(stream))
    if st < 0 or st > 1:
        raise PsmmaError(f"psmma_split_slices_dev returned {st}")
    return t.value, err.value, st == 0


# ---------------------------------------------------------------------------
# Nested HSS representation (include/psmma.h 6b).

class Hss:
    """HSS representation of the n x n eigen-Cauchy (kind KIND_CAUCHYEIG, gens = fpole, u, v,
    difl, difr) or plain Cauchy (KIND_CAUCHY, gens = d, f, u, v) matrix on the device."""

    STAT_KEYS = ("levels", "rank_max", "rank_sum", "saturated", "storage", "samples")

    def __init__(self, kind, gens, leaf=256, tol=1e-12, samples=0, seed=0, stream=None):
        n = gens[0].numel()
        ptrs = [_vec(t, f"gens[{i}]", n) for i, t in enumerate(gens)] + [None] * (5 - len(gens))
        h = ctypes.c_void_p()
        st = (ctypes.c_longlong * 6)()
        self._gens = gens  # the generators must outlive the handle
        _check(lib().psmma_hss_build_dev(kind, n, *ptrs, leaf, float(tol), samples, seed,
                                         ctypes.byref(h), st, _stream_ptr(stream)),
               "psmma_hss_build_dev")
        self._h, self.n = h, n
        self.stats = dict(zip(self.STAT_KEYS, list(st)))

    def product(self, alpha, a, c, beta=0.0, stream=None):
        """X = alpha * A * C + beta * X; a, c column-major (m x n) device views."""
        pa, m, k, lda = _colmajor(a, "a")
        if k != self.n:
            raise ValueError(f"a has {k} columns, the HSS matrix is {self.n} x {self.n}")
        pc, ldc = _out(c, m, self.n)
        _check(lib().psmma_hss_product_dev(self._h, m, float(alpha), pa, lda, float(beta), pc, ldc,
                                           _stream_ptr(stream)), "psmma_hss_product_dev")

    def layout(self):
        """Post-order node arrays and a host copy of the generator array H: node i has
        U_i = H[off[i,0]:] (udim[i] = rows x cols, column-major), V_i = H[off[i,1]:],
        B_i = H[off[i,2]:]; C(I_a, I_b) ~ Ut_a B_a Vt_b^T for siblings a, b."""
        import torch

        nn = ctypes.c_int()
        _check(lib().psmma_hss_layout(self._h, ctypes.byref(nn), None, None, None, None, None,
                                      None, None, None), "psmma_hss_layout")
        k = nn.value
        parent, lo, size = (np.zeros(k, np.int32) for _ in range(3))
        ud, vd, bd = (np.zeros(2 * k, np.int32) for _ in range(3))
        off = np.zeros(3 * k, np.int64)
        hp = ctypes.c_void_p()
        p_ = lambda x: x.ctypes.data_as(ctypes.c_void_p)
        _check(lib().psmma_hss_layout(self._h, ctypes.byref(nn), p_(parent), p_(lo), p_(size),
                                      p_(ud), p_(vd), p_(bd), p_(off), ctypes.byref(hp)),
               "psmma_hss_layout")
        off = off.reshape(k, 3)
        s = self.stats["samples"]
        total = int(off[:, 2].max()) + s * s if k > 1 else 0
        h = np.zeros(max(total, 1), np.float64)
        if total:
            torch.cuda.synchronize()
            # the runtime the library itself is linked against (dlsym through its handle)
            memcpy = lib().hipMemcpy
            memcpy.argtypes = [_P, _P, ctypes.c_size_t, _I]
            memcpy.restype = _I
            _check(memcpy(p_(h), hp, 8 * total, 2), "hipMemcpy")  # device to host
        return dict(parent=parent, lo=lo, size=size, udim=ud.reshape(k, 2),
                    vdim=vd.reshape(k, 2), bdim=bd.reshape(k, 2), off=off, H=h)

    def close(self):
        if getattr(self, "_h", None):
            lib().psmma_hss_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
