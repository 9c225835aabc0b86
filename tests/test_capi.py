"""The C-ABI library loads and exports what include/psmma.h declares (no GPU).

Argument-check paths of the Fortran entry points return before any HIP call,
so the reference's INFO codes can be exercised on the CPU too.
"""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from psmma_git_amd import api

HEADER = Path(__file__).resolve().parents[1] / "include" / "psmma.h"


def header_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b([a-z_][a-z0-9_]*)\s*\(", text)) - {"if", "defined"})


def test_library_exports_every_declared_symbol():
    L = api.lib()
    names = header_functions()
    assert len(names) >= 19
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_mirror_binds_every_declared_symbol():
    assert set(header_functions()) <= set(api.declared_symbols())


FORTRAN_MODULE = Path(__file__).resolve().parents[1] / "psmma_git_amd" / "fortran" / "psmma_hip.f90"
FLANG = Path("/opt/rocm/llvm/bin/amdflang")


@pytest.mark.skipif(not FLANG.exists(), reason="amdflang not installed")
def test_fortran_interface_module_compiles_and_binds_declared_symbols(tmp_path):
    """The ISO_C_BINDING module compiles, and every C name it binds is declared in
    include/psmma.h and exported by the library."""
    import subprocess

    r = subprocess.run([str(FLANG), "-c", "-o", str(tmp_path / "m.o"), "-J", str(tmp_path),
                        str(FORTRAN_MODULE)], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    bound = re.findall(r'bind\(C, name="([a-z0-9_]+)"\)', FORTRAN_MODULE.read_text())
    assert len(bound) >= 14
    assert set(bound) <= set(header_functions())
    L = api.lib()
    assert all(hasattr(L, n) for n in bound)


def test_version_string():
    assert "gfx950" in api.lib().psmma_version().decode()


def _right_call(fn, n, k, lda, ldq, nb=4, ldb=None, redist=False):
    a = np.zeros((lda, max(k, 1)), order="F")
    q = np.full((ldq, max(k, 1)), 7.0, order="F")
    b = np.zeros((ldb if ldb is not None else 2 * max(k, 1), 4), order="F")
    desca = api.descinit(n, k, nb, nb, lda)
    descq = api.descinit(n, k, nb, nb, ldq)
    fn(n, k, 1.0, a, 1, 1, desca, b, b.shape[0], 0.0, q, 1, 1, descq, None, redist)
    return q


@pytest.mark.parametrize("fn,name", [(api.psmma_toeplitz, "PSTOEPLITZ_COMPUTE"),
                                     (api.psmma_hankel, "PSHANKEL_COMPUTE"),
                                     (api.psmma_cauchy, "PSCAUCHY_COMPUTE")])
def test_reference_info_codes(fn, name, capfd):
    # SRC/pscauchy_compute.f90:234-257,307-312: N<0 -> 2, K<0 -> 3, NB<1 -> 5,
    # LDA<MP -> 9, LDC<MP -> 14; the message names the routine, outputs untouched
    q = _right_call(fn, -1, 4, 4, 4)
    assert api.last_status() == 2
    assert np.all(q == 7.0)
    _right_call(fn, 4, -2, 4, 4)
    assert api.last_status() == 3
    _right_call(fn, 4, 4, 4, 4, nb=0)
    assert api.last_status() == 5
    q = _right_call(fn, 8, 4, 4, 8)
    assert api.last_status() == 9
    assert np.all(q == 7.0)
    _right_call(fn, 8, 4, 8, 4)
    assert api.last_status() == 14
    err = capfd.readouterr().err
    assert f"On entry to {name}" in err and "parameter number 14" in err


def test_quick_returns_do_not_touch_outputs():
    # N == 0 (MP == 0) and (alpha == 0, beta == 1) return at once
    a = np.zeros((4, 4), order="F")
    q = np.full((4, 4), 3.0, order="F")
    g = np.zeros(7)
    d = api.descinit(4, 4, 2, 2, 4)
    api.psmma_toeplitz(0, 4, 1.0, a, 1, 1, d, g, 7, 0.0, q, 1, 1, d)
    assert api.last_status() == 0 and np.all(q == 3.0)
    api.psmma_toeplitz(4, 4, 0.0, a, 1, 1, d, g, 7, 1.0, q, 1, 1, d)
    assert api.last_status() == 0 and np.all(q == 3.0)


def test_short_generators_rejected():
    a = np.zeros((4, 4), order="F")
    q = np.zeros((4, 4), order="F")
    d = api.descinit(4, 4, 2, 2, 4)
    api.psmma_hankel(4, 4, 1.0, a, 1, 1, d, np.zeros(5), 5, 0.0, q, 1, 1, d)
    assert api.last_status() == 8
    api.psmma_toeplitz(4, 4, 1.0, a, 1, 1, d, np.zeros(3), 3, 0.0, q, 1, 1, d)
    assert api.last_status() == 8


def test_dft_info_codes(capfd):
    b = np.zeros((4, 4), complex, order="F")
    c = np.zeros((4, 4), complex, order="F")
    api.pzmdftl1(-1, 4, 4, 4, 1.0, b, 4, 0.0, c, 4)
    assert api.last_status() == 1
    api.pzmdftl1(4, 4, 0, 4, 1.0, b, 4, 0.0, c, 4)
    assert api.last_status() == 3
    api.pzmdftl1(4, 4, 4, 4, 1.0, b, 4, 0.0, c, 2)
    assert api.last_status() == 10
    # the reference contracts over K but sizes columns by N: K != N is refused
    api.pzmdftl1(4, 3, 4, 4, 1.0, b, 4, 0.0, c, 4)
    assert api.last_status() == 2
    assert "PDMDFTL" in capfd.readouterr().err


def test_set_precision_validates():
    assert api.lib().psmma_set_precision(9, 3) == -1
    assert api.lib().psmma_set_precision(api.PREC_SPLIT_I8, -1) == -2
    assert api.lib().psmma_set_precision(api.PREC_SPLIT_I8, 17) == -2
    assert api.lib().psmma_set_precision(api.PREC_SPLIT_I8, 0) == 0  # 0: chosen from K
    assert api.lib().psmma_set_precision(api.PREC_FP64, 0) == 0


GRID_CHECK = Path(__file__).resolve().parent / "fortran" / "_build" / "grid_check"


@pytest.mark.skipif(not FLANG.exists(), reason="amdflang not installed")
def test_fortran_caller_on_a_larger_grid_is_refused():
    """A Fortran caller whose BLACS reports a 2 x 2 grid: PSMMA_TOEPLITZ (context from
    its descriptor) and PZMDFTL1 / PZMDFTL (context from COMMON /CONTEXT/) refuse it
    before any device work and leave their outputs untouched (tests/fortran/grid_check.f90,
    built with the library)."""
    import subprocess

    if not GRID_CHECK.exists():
        subprocess.run(["make", "-s", "-C", str(GRID_CHECK.parents[1])], check=True)
    r = subprocess.run([str(GRID_CHECK)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    vals = dict(line.split() for line in r.stdout.strip().splitlines())
    assert vals == {"toeplitz_status": "-2", "pzmdftl1_status": "-2", "pzmdftl_status": "-2",
                    "outputs_untouched": "1"}
    assert r.stderr.count("process grid 2 x 2") == 3


def test_hss_build_argument_codes():
    """Argument checks of psmma_hss_build_dev return before any device call (the
    pointers below are never dereferenced)."""
    L = api.lib()
    p = ctypes.c_void_p(64)
    h = ctypes.c_void_p()
    st = (ctypes.c_longlong * 6)()

    def call(kind=api.KIND_CAUCHYEIG, n=1024, gens=(p,) * 5, leaf=256, tol=1e-12, samples=0):
        return L.psmma_hss_build_dev(kind, n, *gens, leaf, tol, samples, 0, ctypes.byref(h), st,
                                     None)

    assert call(kind=api.KIND_TOEPLITZ) == -1
    assert call(n=0) == -2
    assert call(gens=(p, p, p, p, None)) == -3  # eigen-Cauchy needs DIFR
    assert call(leaf=300) == -8                  # leaf above 256
    assert call(leaf=1) == -8                    # a leaf of one index
    assert call(tol=-1.0) == -9
    assert call(samples=129) == -10
    assert L.psmma_hss_product_dev(None, 4, 1.0, p, 4, 0.0, p, 4, None) == -1
    assert L.psmma_hss_destroy(None) == 0
