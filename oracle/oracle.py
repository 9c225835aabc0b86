"""CPU oracle for the PSMMA structured x dense products.

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and
by the CPU-baseline leg of bench.py, always as the checker, never as the thing
measured or shipped.

Two restatements of the reference algorithm live here:

* ``psmma_oracle.c`` (loaded through ctypes): the tile generators, index maps
  and the block-cyclic ring loop of PS*_COMPUTE on an emulated P x Q grid
  (reference files cited function by function in that source);
* numpy functions below that materialise a structured matrix explicitly from
  its generators and multiply it with FP64 BLAS.  They restate the same
  formulas (cited per function) and serve as the FP64 oracle at sizes where
  the C ring emulation would be slow.

Seeded input generators for the five benchmark configurations are in
``inputs.py`` next to this file.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libpsmma_oracle.so"

CAUCHY, TOEPLITZ, HANKEL, CAUCHYEIG = 0, 1, 2, 3
MODE_DENSE, MODE_FAITHFUL = 0, 1


def build() -> Path:
    """Compile the C oracle with its Makefile (gcc only)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = ctypes.CDLL(str(LIB_PATH))
        _declare(_lib)
    return _lib


class OraProblem(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int), ("nprow", ctypes.c_int), ("npcol", ctypes.c_int),
        ("nb", ctypes.c_int), ("n", ctypes.c_int), ("k", ctypes.c_int),
        ("alpha", ctypes.c_double), ("beta", ctypes.c_double),
        ("a", ctypes.c_void_p), ("lda", ctypes.c_int),
        ("gen", ctypes.c_void_p), ("ldgen", ctypes.c_int),
        ("c", ctypes.c_void_p), ("ldc", ctypes.c_int), ("mode", ctypes.c_int),
    ]


def _declare(L):
    I, D, P = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
    L.ora_numroc.argtypes = [I, I, I, I]
    L.ora_numroc.restype = I
    L.ora_constindex.argtypes = [I, I, I, I, P]
    L.ora_constindex.restype = I
    L.ora_is_intersect.argtypes = [I, P, I, P]
    L.ora_is_intersect.restype = I
    L.ora_dgemm.argtypes = [I, I, I, D, P, I, P, I, D, P, I]
    L.ora_dgemm.restype = None
    L.ora_set_blas_dgemm.argtypes = [P]
    L.ora_set_blas_dgemm.restype = None
    L.ora_ps_compute.argtypes = [ctypes.POINTER(OraProblem), I, P]
    L.ora_ps_compute.restype = I
    L.ora_structured_gemm.argtypes = [I, I, I, I, D, P, I, P, I, I, I, I, D, P, I]
    L.ora_structured_gemm.restype = I
    L.ora_mdgeqp3.argtypes = [I, I, P, I, P, P, D, I, P]
    L.ora_mdgeqp3.restype = I
    L.ora_compress_rrqr.argtypes = [I, I, P, I, P, D]
    L.ora_compress_rrqr.restype = I
    L.ora_compress_cauchy.argtypes = [P, P, P, P, D, I, I, P]
    L.ora_compress_cauchy.restype = I
    L.ora_compress_cauchyeig2.argtypes = [P, P, P, P, P, P, P, D, I, I, P, P]
    L.ora_compress_cauchyeig2.restype = I
    L.ora_rrlu_cauchy.argtypes = [P, P, P, P, D, P, P, P, P, I, I]
    L.ora_rrlu_cauchy.restype = I
    L.ora_rrlu_cauchyeig_row.argtypes = [P, P, P, P, P, P, D, P, P, P, P, P, I, I, P]
    L.ora_rrlu_cauchyeig_row.restype = I


def _f64(x):
    return np.asfortranarray(x, dtype=np.float64)


def _ptr(x):
    return x.ctypes.data_as(ctypes.c_void_p)


# ---------------------------------------------------------------------------
# Index maps (C restatement).

def constindex(n: int, nb: int, mycol: int, npcol: int) -> np.ndarray:
    """SRC/Auxil_2p5D.f90:2-66 (ConstIndex), 1-based global indices."""
    out = np.zeros(n + nb + 1, dtype=np.int32)
    ln = lib().ora_constindex(n, nb, mycol, npcol, _ptr(out))
    return out[:ln].copy()


def is_intersect(a, b) -> bool:
    """SRC/Auxil_2p5D.f90:752-792: TRUE when the sets are DISJOINT (reference sense)."""
    a = np.ascontiguousarray(a, dtype=np.int32)
    b = np.ascontiguousarray(b, dtype=np.int32)
    return bool(lib().ora_is_intersect(len(a), _ptr(a), len(b), _ptr(b)))


def numroc(n: int, nb: int, iproc: int, nprocs: int) -> int:
    return int(lib().ora_numroc(n, nb, iproc, nprocs))


# ---------------------------------------------------------------------------
# Explicit structured matrices (numpy restatement), 0-based block offsets.

def toeplitz_matrix(g, korder: int, r0: int, k: int, c0: int, n: int) -> np.ndarray:
    """SRC/Auxil_2p5D.f90:224-231: H(i,j) = G(K + j - i), 1-based."""
    r = np.arange(r0, r0 + k)[:, None]
    c = np.arange(c0, c0 + n)[None, :]
    return np.asarray(g, dtype=np.float64)[korder - 1 + c - r]


def hankel_matrix(g, r0: int, k: int, c0: int, n: int) -> np.ndarray:
    """SRC/Auxil_2p5D.f90:309-316: H(i,j) = G(i + j - 1), 1-based."""
    r = np.arange(r0, r0 + k)[:, None]
    c = np.arange(c0, c0 + n)[None, :]
    return np.asarray(g, dtype=np.float64)[r + c]


def cauchy_matrix(d, f, u, v, r0: int, k: int, c0: int, n: int) -> np.ndarray:
    """SRC/Cauchylowrank.f90:509-521: H(i,j) = U_i V_j / (D_i - F_j)."""
    d = np.asarray(d, np.float64)[r0:r0 + k, None]
    u = np.asarray(u, np.float64)[r0:r0 + k, None]
    f = np.asarray(f, np.float64)[None, c0:c0 + n]
    v = np.asarray(v, np.float64)[None, c0:c0 + n]
    return (u * v) / (d - f)


def cauchyeig_matrix(fpole, u, v, difl, difr, r0: int, k: int, c0: int, n: int) -> np.ndarray:
    """SRC/CauchyHssEig.f90:1758-1770 (CauchylikeEig) with PL=i, PU=j."""
    F = np.asarray(fpole, np.float64)
    r = np.arange(r0, r0 + k)[:, None]
    c = np.arange(c0, c0 + n)[None, :]
    cp1 = np.minimum(c + 1, len(F) - 1)
    den = np.where(r <= c, (F[r] - F[c]) - np.asarray(difl)[c],
                   (F[r] - F[cp1]) - np.asarray(difr)[c])
    return (np.asarray(u, np.float64)[r] * np.asarray(v, np.float64)[c]) / den


def dft_matrix(nd: int, r0: int, m: int, c0: int, k: int) -> np.ndarray:
    """F(r,c) = exp(-2 pi i (r c mod nd) / nd), unnormalised (SRC/pzmdftL1.f90:329-337).

    The exponent is reduced modulo nd in exact integer arithmetic before the
    angle is formed, so every entry is correctly rounded."""
    r = np.arange(r0, r0 + m, dtype=np.int64)[:, None]
    c = np.arange(c0, c0 + k, dtype=np.int64)[None, :]
    e = (r * c) % nd
    ang = 2.0 * np.pi * e.astype(np.float64) / nd
    return np.cos(ang) - 1j * np.sin(ang)


_TWO_PI_LD = np.longdouble("6.283185307179586476925286766559005768")


def dft_twiddles(nd: int) -> np.ndarray:
    """tw[e] = exp(-2 pi i e / nd) for e in [0, nd), each part rounded once from long double
    (the FP64 table the DFT kernels read)."""
    ang = _TWO_PI_LD * np.arange(nd).astype(np.longdouble) / np.longdouble(nd)
    return np.cos(ang).astype(np.float64) - 1j * np.sin(ang).astype(np.float64)


def slice8(x, slices: int) -> list:
    """Fixed-point slices of x in [-1, 1]: x_1 = rint(x 2^8) 2^-8, x_p = rint(r 2^(8p)) 2^(-8p)
    of the remainder r; |x - sum x_p| <= 2^(-8 slices - 1)."""
    out, r = [], np.array(x, np.float64, copy=True)
    for p in range(slices):
        sc = 2.0 ** (8 * (p + 1))
        q = np.rint(r * sc) / sc
        out.append(q)
        r = r - q
    return out


def dft_split_fp32(nd: int, r0: int, m: int, c0: int, b, slices: int) -> np.ndarray:
    """The split-FP32 DFT product F(r0:r0+m, c0:c0+k) * B as the split scheme defines it:
    columns of B scaled by 2^-e_j (max |Re|, |Im| < 2^e_j), F and the scaled B cut into
    `slices` 8-bit fixed-point slices, slice pairs p + q <= slices + 1 multiplied exactly and
    summed, then scaled back.  Test infrastructure for the split-FP32 form of the product of
    SRC/pzmdftL1.f90:488-498."""
    b = np.asarray(b, np.complex128)
    k, n = b.shape
    tw = dft_twiddles(nd)
    e = ((np.arange(r0, r0 + m, dtype=np.int64)[:, None] *
          np.arange(c0, c0 + k, dtype=np.int64)[None, :]) % nd)
    f = tw[e]
    amax = np.maximum(np.abs(b.real), np.abs(b.imag)).max(axis=0) if k else np.zeros(n)
    ex = np.frexp(amax)[1]
    scale = np.where(amax > 0, np.ldexp(1.0, -ex), 0.0)
    bs = b * scale[None, :]
    fr, fi = slice8(f.real, slices), slice8(f.imag, slices)
    br, bi = slice8(bs.real, slices), slice8(bs.imag, slices)
    x = np.zeros((m, n), np.complex128)
    for p in range(slices):
        for q in range(slices - p):
            x += (fr[p] + 1j * fi[p]) @ (br[q] + 1j * bi[q])
    return x * np.where(amax > 0, np.ldexp(1.0, ex), 0.0)[None, :]


def structured_gemm_np(kind: int, alpha, a, gens: dict, r0: int, c0: int, n: int, beta=0.0,
                       c=None) -> np.ndarray:
    """C = alpha * A * H(r0:r0+k, c0:c0+n) + beta * C in FP64 with an explicit H."""
    a = np.asarray(a, np.float64)
    k = a.shape[1]
    if kind == TOEPLITZ:
        h = toeplitz_matrix(gens["g"], gens["korder"], r0, k, c0, n)
    elif kind == HANKEL:
        h = hankel_matrix(gens["g"], r0, k, c0, n)
    elif kind == CAUCHY:
        h = cauchy_matrix(gens["d"], gens["f"], gens["u"], gens["v"], r0, k, c0, n)
    else:
        h = cauchyeig_matrix(gens["f"], gens["u"], gens["v"], gens["difl"], gens["difr"],
                             r0, k, c0, n)
    out = alpha * (a @ h)
    if beta != 0.0:
        out = out + beta * np.asarray(c, np.float64)
    return out


# ---------------------------------------------------------------------------
# C restatement entry points.

def structured_gemm_c(kind: int, alpha, a, gen, korder: int, r0: int, c0: int, n: int,
                      beta=0.0, c=None) -> np.ndarray:
    """Dense step of PS*_COMPUTE on a 1 x 1 grid (C restatement)."""
    a = _f64(a)
    m, k = a.shape
    gen = _f64(gen)
    ldgen = gen.shape[0]
    out = _f64(c) if c is not None else np.zeros((m, n), order="F")
    out = out.copy(order="F")
    st = lib().ora_structured_gemm(kind, m, n, k, float(alpha), _ptr(a), max(1, m), _ptr(gen),
                                   ldgen, korder, r0, c0, float(beta), _ptr(out), max(1, m))
    if st != 0:
        raise RuntimeError(f"ora_structured_gemm failed ({st})")
    return out


def ps_compute(kind: int, nprow: int, npcol: int, nb: int, alpha, a, gen, beta=0.0, c=None,
               mode: int = MODE_DENSE, nthreads: int = 1):
    """PS*_COMPUTE over an emulated nprow x npcol grid (C restatement).

    Returns (C, stats) with stats = [dense, lowrank, fullrank, stale] step counts."""
    a = _f64(a)
    n, k = a.shape
    gen = _f64(gen)
    out = _f64(c).copy(order="F") if c is not None else np.zeros((n, k), order="F")
    prob = OraProblem(kind, nprow, npcol, nb, n, k, float(alpha), float(beta), _ptr(a).value,
                      max(1, n), _ptr(gen).value, gen.shape[0], _ptr(out).value, max(1, n), mode)
    stats = np.zeros(4, dtype=np.int64)
    st = lib().ora_ps_compute(ctypes.byref(prob), int(nthreads), _ptr(stats))
    if st != 0:
        raise RuntimeError(f"ora_ps_compute failed ({st})")
    return out, stats


_BLAS = {"name": "the oracle's C DGEMM", "lib": None}

MKL_PATHS = ("/opt/conda/lib/libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so.1")


def use_host_blas() -> str:
    """Route the ring restatement's DGEMMs to the host BLAS, one sequential DGEMM per emulated
    rank thread as the reference runs one MPI rank per core (Test/run.sh:1): MKL's libmkl_rt
    (the reference's BLAS, SLMakes/SLmake.inc.intel:24-35) with its sequential threading layer
    when installed.  Returns the name of the BLAS in use."""
    if _BLAS["lib"] is not None:
        return _BLAS["name"]
    for path in MKL_PATHS:
        if os.path.exists(path):
            try:
                mkl = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
                mkl.MKL_Set_Interface_Layer(0)   # LP64: 32-bit INTEGER, as the reference links
                mkl.MKL_Set_Threading_Layer(1)   # sequential
                ver = ctypes.create_string_buffer(256)
                mkl.MKL_Get_Version_String(ver, 256)
                addr = ctypes.cast(mkl.dgemm_, ctypes.c_void_p).value
            except (OSError, AttributeError):
                continue
            lib().ora_set_blas_dgemm(addr)
            _BLAS.update(name=ver.value.decode().strip() + " (sequential, LP64)", lib=mkl)
            return _BLAS["name"]
    return _BLAS["name"]


def dgemm_c(alpha, a, b, beta=0.0, c=None) -> np.ndarray:
    a = _f64(a)
    b = _f64(b)
    m, k = a.shape
    n = b.shape[1]
    out = _f64(c).copy(order="F") if c is not None else np.zeros((m, n), order="F")
    lib().ora_dgemm(m, n, k, float(alpha), _ptr(a), max(1, m), _ptr(b), max(1, k), float(beta),
                    _ptr(out), max(1, m))
    return out


# ---------------------------------------------------------------------------
# Truncated pivoted QR (C restatement of MDGEQP3 / Compress_RRQR).

RRQR_TOL = float(np.float32(1e-13))  # Compress_RRQR's TOL: the default-real literal 1.0E-13


def mdgeqp3(a, tol: float = 0.0, nb: int = 32, jpvt=None):
    """SRC/mdgeqp3.f:153-370.  Returns (factored A, jpvt (1-based), tau, rk_bound)."""
    a = np.array(a, dtype=np.float64, order="F", copy=True)
    m, n = a.shape
    jp = np.zeros(n, np.int32) if jpvt is None else np.array(jpvt, np.int32)
    tau = np.zeros(max(1, min(m, n)))
    rk = ctypes.c_int(0)
    st = lib().ora_mdgeqp3(m, n, _ptr(a), max(1, m), _ptr(jp), _ptr(tau), tol, nb,
                           ctypes.byref(rk))
    if st != 0:
        raise RuntimeError(f"ora_mdgeqp3 returned {st}")
    return a, jp, tau[:min(m, n)], rk.value


def compress_rrqr(a, tol: float = RRQR_TOL):
    """SRC/Auxil_2p5D.f90:1060-1189.  Returns (rk, Q (m x rk), R P^T (rk x n)); Q and
    R P^T are None at full rank (the reference leaves H unwritten there)."""
    a = np.array(a, dtype=np.float64, order="F", copy=True)
    m, n = a.shape
    h = np.zeros(max(1, 2 * m * n))
    rk = lib().ora_compress_rrqr(m, n, _ptr(a), max(1, m), _ptr(h), tol)
    if rk < 0:
        raise RuntimeError("ora_compress_rrqr failed")
    if rk == min(m, n):
        return rk, None, None
    q = h[:m * rk].reshape((m, rk), order="F").copy()
    r = h[m * rk:m * rk + rk * n].reshape((rk, n), order="F").copy()
    return rk, q, r


RRLU_TOL = float(np.float32(1e-17))  # ConstCauchylowrank's TOL: the default-real literal 1.0E-17


def compress_cauchy(d, f, u, v, tol: float = RRLU_TOL, h_init=None):
    """SRC/Cauchylowrank.f90:387-491 ('R'): row ID of U_i V_j / (D_i - F_j).
    Returns (rk, X (m x rk), Y (rk x n), h): X, Y are None when rk >= m (H unwritten);
    h is the raw buffer (initialised from h_init, to observe what is left unwritten)."""
    d, f, u, v = (np.ascontiguousarray(x, dtype=np.float64) for x in (d, f, u, v))
    m, n = len(d), len(f)
    h = np.zeros(max(1, 2 * m * n)) if h_init is None else np.array(h_init, np.float64)
    rk = lib().ora_compress_cauchy(_ptr(d), _ptr(f), _ptr(u), _ptr(v), tol, m, n, _ptr(h))
    if rk < 0:
        raise RuntimeError("ora_compress_cauchy failed")
    if rk >= m:
        return rk, None, None, h
    x = h[:m * rk].reshape((m, rk), order="F").copy()
    y = h[m * rk:m * rk + rk * n].reshape((rk, n), order="F").copy()
    return rk, x, y, h


def compress_cauchyeig2(fpole, difl, difr, u, v, rows, cols, tol: float = RRLU_TOL):
    """SRC/CauchyHssEig.f90:2640-2829: row ID of the eigen-Cauchy block with global
    0-based rows / cols and the global generators u, v (gathered here, as
    ConstCauchyEiglowrank does).  Returns (rk, Q (m x rk), Y (rk x n), nswap); at
    full rank Q is None and Y is the explicit block."""
    ff, dl, dr, u, v = (np.ascontiguousarray(x, dtype=np.float64)
                        for x in (fpole, difl, difr, u, v))
    rows = np.asarray(rows, np.int64)
    cols = np.asarray(cols, np.int64)
    m, n = len(rows), len(cols)
    pl = np.ascontiguousarray(rows + 1, np.int32)
    pu = np.ascontiguousarray(cols + 1, np.int32)
    ug = np.ascontiguousarray(u[rows])
    vg = np.ascontiguousarray(v[cols])
    h = np.zeros(max(1, 2 * m * n))
    nsw = np.zeros(1, np.int32)
    rk = lib().ora_compress_cauchyeig2(_ptr(ff), _ptr(dl), _ptr(dr), _ptr(ug), _ptr(vg),
                                       _ptr(pl), _ptr(pu), tol, m, n, _ptr(h), _ptr(nsw))
    if rk < 0:
        raise RuntimeError("ora_compress_cauchyeig2 failed")
    if rk == min(m, n):
        return rk, None, h[:m * n].reshape((m, n), order="F").copy(), int(nsw[0])
    q = h[:m * rk].reshape((m, rk), order="F").copy()
    y = h[m * rk:m * rk + rk * n].reshape((rk, n), order="F").copy()
    return rk, q, y, int(nsw[0])


def rrlu_cauchy(d, f, a, v, tol: float = RRLU_TOL):
    """SRC/Cauchylowrank.f90:7-134 (rrluCauchy) on one block: every output.
    Returns dict(rk, pl, pu (1-based), d, f, a, v (permuted / updated in place),
    u (m), w (min(m, n)))."""
    d, f, a, v = (np.array(x, dtype=np.float64) for x in (d, f, a, v))
    m, n = len(d), len(f)
    u = np.zeros(m)
    w = np.zeros(max(1, min(m, n)))
    pl = np.zeros(m, np.int32)
    pu = np.zeros(n, np.int32)
    rk = lib().ora_rrlu_cauchy(_ptr(d), _ptr(f), _ptr(a), _ptr(v), tol, _ptr(u), _ptr(w),
                               _ptr(pl), _ptr(pu), m, n)
    return dict(rk=rk, pl=pl, pu=pu, d=d, f=f, a=a, v=v, u=u, w=w[:min(m, n)])


def rrlu_cauchyeig_row(fpole, difl, difr, a, v, pl, pu, tol: float = RRLU_TOL):
    """SRC/CauchyHssEig.f90:7-212 (rrluCauchyEigrow) on one block with global 1-based
    row / column indices pl, pu and the block's generators a (rows), v (columns).
    Returns dict(rk, pl, pu, pq, nswap, f (the block's row poles), a, v, u, w)."""
    ff, dl, dr = (np.ascontiguousarray(x, dtype=np.float64) for x in (fpole, difl, difr))
    a, v = np.array(a, np.float64), np.array(v, np.float64)
    pl, pu = np.array(pl, np.int32), np.array(pu, np.int32)
    m, n = len(pl), len(pu)
    f = np.ascontiguousarray(ff[pl - 1])
    u = np.zeros(m)
    w = np.zeros(max(1, min(m, n)))
    pq = np.zeros(m, np.int32)
    nsw = np.zeros(1, np.int32)
    rk = lib().ora_rrlu_cauchyeig_row(_ptr(ff), _ptr(dl), _ptr(dr), _ptr(f), _ptr(a), _ptr(v),
                                      tol, _ptr(u), _ptr(w), _ptr(pl), _ptr(pu), _ptr(pq), m, n,
                                      _ptr(nsw))
    return dict(rk=rk, pl=pl, pu=pu, pq=pq, nswap=int(nsw[0]), f=f, a=a, v=v, u=u,
                w=w[:min(m, n)])


def rel_fro(x, ref) -> float:
    """Relative Frobenius error ||x - ref||_F / ||ref||_F."""
    x = np.asarray(x)
    ref = np.asarray(ref)
    den = np.linalg.norm(ref)
    return float(np.linalg.norm(x - ref) / (den if den > 0 else 1.0))
