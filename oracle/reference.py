"""Runs the reference's own generator-only RRLU routines (rrluCauchy,
rrluCauchyEigrow) and explicit tile generators (Cauchylike, CauchylikeEig)
through the driver oracle/ref_rrlu.f90, built from the reference's module
sources by `make -C oracle ref` into oracle/_ref.

Test infrastructure only: the CPU tests compare the oracle's restatements of
the same routines with these outputs, and tests/golden/make_rrlu_golden.py
writes them into the committed fixture the GPU tests use.  Absent where the
reference sources are not (the driver is then not built and `available()` is
False)."""
from __future__ import annotations

import subprocess
import tempfile
from pathlib import Path

import numpy as np

DRIVER = Path(__file__).resolve().parent / "_ref" / "ref_rrlu"


def available() -> bool:
    return DRIVER.exists()


def _run(payload: bytes, timeout: float) -> bytes:
    with tempfile.TemporaryDirectory() as td:
        fin, fout = Path(td) / "in.bin", Path(td) / "out.bin"
        fin.write_bytes(payload)
        subprocess.run([str(DRIVER), str(fin), str(fout)], check=True, timeout=timeout,
                       capture_output=True)
        return fout.read_bytes()


class _Reader:
    def __init__(self, buf: bytes):
        self.buf, self.off = buf, 0

    def take(self, dtype, count):
        a = np.frombuffer(self.buf, dtype=dtype, count=count, offset=self.off).copy()
        self.off += a.nbytes
        return a


def rrlu_cauchy(d, f, a, v, tol: float, timeout: float = 60.0) -> dict:
    """rrluCauchy (SRC/Cauchylowrank.f90:7-134): the same dict as oracle.rrlu_cauchy."""
    d, f, a, v = (np.ascontiguousarray(x, np.float64) for x in (d, f, a, v))
    m, n = len(d), len(f)
    head = np.array([0, m, n], np.int32).tobytes() + np.array([tol], np.float64).tobytes()
    r = _Reader(_run(head + d.tobytes() + f.tobytes() + a.tobytes() + v.tobytes(), timeout))
    rk = int(r.take(np.int32, 1)[0])
    pl, pu = r.take(np.int32, m), r.take(np.int32, n)
    d2, f2, a2, v2 = r.take(np.float64, m), r.take(np.float64, n), r.take(np.float64, m), \
        r.take(np.float64, n)
    u, w = r.take(np.float64, m), r.take(np.float64, min(m, n))
    return dict(rk=rk, pl=pl, pu=pu, d=d2, f=f2, a=a2, v=v2, u=u, w=w)


def rrlu_cauchyeig_row(fpole, difl, difr, dcol, a, v, pl, pu, tol: float,
                       timeout: float = 60.0) -> dict:
    """rrluCauchyEigrow (SRC/CauchyHssEig.f90:7-212) on the block with global 1-based
    rows pl and columns pu (dcol: the column poles the routine carries along with the
    columns): the same dict as oracle.rrlu_cauchyeig_row."""
    ff, dl, dr, dc, a, v = (np.ascontiguousarray(x, np.float64)
                            for x in (fpole, difl, difr, dcol, a, v))
    pl, pu = np.ascontiguousarray(pl, np.int32), np.ascontiguousarray(pu, np.int32)
    m, n, ng = len(pl), len(pu), len(ff)
    head = (np.array([1, m, n], np.int32).tobytes() + np.array([tol], np.float64).tobytes() +
            np.array([ng], np.int32).tobytes())
    body = b"".join(x.tobytes() for x in (ff, dl, dr, dc, a, v, pl, pu))
    r = _Reader(_run(head + body, timeout))
    rk = int(r.take(np.int32, 1)[0])
    pl2, pu2 = r.take(np.int32, m), r.take(np.int32, n)
    pq, nswap = r.take(np.int32, m), int(r.take(np.int32, 1)[0])
    f2, a2, v2 = r.take(np.float64, m), r.take(np.float64, m), r.take(np.float64, n)
    u, w = r.take(np.float64, m), r.take(np.float64, min(m, n))
    return dict(rk=rk, pl=pl2, pu=pu2, pq=pq, nswap=nswap, f=f2, a=a2, v=v2, u=u, w=w)


# ---------------------------------------------------------------------------
# Blocks on which the restatements are pinned to the reference (the golden
# fixtures under tests/golden and the live comparison use the same inputs).
# Plain Cauchy: separated poles, and off-diagonal blocks of the benchmark's
# interlaced generators.  Eigen-Cauchy: blocks of the first block row of the
# merge matrix (rows 1..M): there CauchyPivtEig's early-return test, which reads
# the block's row-pole array at a global row index (SRC/CauchyHssEig.f90:675-678),
# stays inside that array.

def cauchy_case(kind: str, m: int, n: int, seed: int):
    """(d, f, a, v) of a plain Cauchy block."""
    from oracle import inputs
    rng = np.random.default_rng(seed)
    if kind == "separated":
        d = np.sort(rng.random(m))
        f = 1.05 + np.sort(rng.random(n))
        return d, f, rng.standard_normal(m), rng.standard_normal(n)
    b = inputs.cauchy_generators(m + n, rng)  # interlaced poles; rows 0..m-1, columns m..m+n-1
    d, f, a, v = b[:m, 0], b[m:, 1], b[:m, 2].copy(), b[m:, 3]
    if kind == "deficient":  # a quarter of the rows zero: exact rank deficiency
        a[rng.choice(m, m // 4, replace=False)] = 0.0
    return d, f, a, v


def cauchy_tile_generators(nglob: int, seed: int):
    """(d, f, u, v): the benchmark's interlaced plain Cauchy generators of size nglob."""
    from oracle import inputs
    b = inputs.cauchy_generators(nglob, np.random.default_rng(seed))
    return (np.ascontiguousarray(b[:, i]) for i in range(4))


def cauchyeig_case(nglob: int, m: int, n: int, seed: int):
    """(fpole, difl, difr, u, v, rows, cols) of the first-block-row block rows 0..m-1,
    columns m..m+n-1 (0-based) of the N = nglob merge matrix."""
    from oracle import inputs
    e = inputs.eig_merge_inputs(nglob, seed=seed, with_q=False)
    f, dl, dr = e["f"], e["difl"], e["difr"]
    zh = inputs.loewner_z(f, e["u"], dl, dr)
    v = inputs.column_normalisers(f, zh, dl, dr)
    return f, dl, dr, zh, v, np.arange(m), np.arange(m, m + n)


def cauchy_tile(d, f, u, v, timeout: float = 60.0) -> np.ndarray:
    """Cauchylike (SRC/Cauchylowrank.f90:509-521): the m x n block U_i V_j / (D_i - F_j)."""
    d, f, u, v = (np.ascontiguousarray(x, np.float64) for x in (d, f, u, v))
    m, n = len(d), len(f)
    head = np.array([2, m, n], np.int32).tobytes() + np.zeros(1).tobytes()
    out = _run(head + b"".join(x.tobytes() for x in (d, f, u, v)), timeout)
    return np.frombuffer(out, np.float64, m * n).reshape((m, n), order="F").copy()


def cauchyeig_tile(fpole, difl, difr, u, v, rows, cols, timeout: float = 60.0) -> np.ndarray:
    """CauchylikeEig (SRC/CauchyHssEig.f90:1693-1774): the block of the eigen-Cauchy matrix
    at global 0-based rows / cols; u, v are the global row / column generators."""
    ff, dl, dr = (np.ascontiguousarray(x, np.float64) for x in (fpole, difl, difr))
    rows, cols = np.asarray(rows), np.asarray(cols)
    m, n = len(rows), len(cols)
    ul = np.ascontiguousarray(np.asarray(u, np.float64)[rows])
    vl = np.ascontiguousarray(np.asarray(v, np.float64)[cols])
    pl = np.ascontiguousarray(rows + 1, np.int32)
    pu = np.ascontiguousarray(cols + 1, np.int32)
    head = (np.array([3, m, n], np.int32).tobytes() + np.zeros(1).tobytes() +
            np.array([len(ff)], np.int32).tobytes())
    out = _run(head + b"".join(x.tobytes() for x in (ff, dl, dr, ul, vl, pl, pu)), timeout)
    return np.frombuffer(out, np.float64, m * n).reshape((m, n), order="F").copy()
