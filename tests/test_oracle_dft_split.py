"""The split-FP32 DFT scheme as the oracle restates it (CPU only).

What pins it: the slices reconstruct each value within 2^(-8S-1) and are exact
in FP32; every slab class sum the kernel forms in FP32 stays below 2^24 grid
units (so it is exact); the scheme converges to the FP64 product as S grows,
within the closed-form bound; the twiddle table equals exp(-2 pi i e / nd).
"""
import numpy as np
import pytest

from oracle import oracle


def test_twiddle_table_is_the_dft_root_powers():
    nd = 1000
    tw = oracle.dft_twiddles(nd)
    e = np.arange(nd)
    assert np.max(np.abs(tw - np.exp(-2j * np.pi * e / nd))) < 2e-15  # float64 angle: ~1 ulp of 2 pi
    assert tw[0] == 1.0
    assert abs(tw[nd // 2] + 1.0) < 1e-19 and abs(tw[nd // 4] + 1j) < 1e-19
    assert np.allclose(tw[1] ** 7, tw[7], rtol=0, atol=1e-15)


@pytest.mark.parametrize("s", [1, 2, 3, 5, 6])
def test_slices_reconstruct_and_fit_fp32(s):
    rng = np.random.default_rng(s)
    x = np.concatenate([rng.uniform(-1, 1, 5000), [-1.0, 1.0, 0.0, 2.0 ** -60, 0.5 - 2.0 ** -54]])
    sl = oracle.slice8(x, s)
    assert np.max(np.abs(x - sum(sl))) <= 2.0 ** (-8 * s - 1)
    for p, xp in enumerate(sl):
        assert np.array_equal(xp.astype(np.float32).astype(np.float64), xp)  # exact in FP32
        units = xp * 2.0 ** (8 * (p + 1))
        assert np.array_equal(units, np.rint(units))
        assert np.max(np.abs(units)) <= (2 ** 8 if p == 0 else 2 ** 7)


def test_slab_class_sums_stay_exact_in_fp32():
    # the largest |real or imaginary| class sum over one K-slab of 8, in grid units of
    # 2^(-8 d), must stay below 2^24 for every class the kernel keeps (S = 6)
    s, slab = 6, 8
    rng = np.random.default_rng(7)
    f = oracle.dft_twiddles(97)[rng.integers(0, 97, (16, slab))]
    f[0, :] = 1.0 + 0.0j  # worst case: unit entries
    b = rng.uniform(-1, 1, (slab, 16)) + 1j * rng.uniform(-1, 1, (slab, 16))
    b[:, 0] = (1 - 2.0 ** -30) * (1 + 1j)
    fr, fi = oracle.slice8(f.real, s), oracle.slice8(f.imag, s)
    br, bi = oracle.slice8(b.real, s), oracle.slice8(b.imag, s)
    worst = 0.0
    for d in range(s):
        re = sum(np.abs(fr[p]) @ np.abs(br[d - p]) + np.abs(fi[p]) @ np.abs(bi[d - p])
                 for p in range(d + 1))
        im = sum(np.abs(fr[p]) @ np.abs(bi[d - p]) + np.abs(fi[p]) @ np.abs(br[d - p])
                 for p in range(d + 1))
        worst = max(worst, np.max(re) * 2.0 ** (8 * (d + 2)), np.max(im) * 2.0 ** (8 * (d + 2)))
    assert worst < 2.0 ** 24


@pytest.mark.parametrize("m,k,n,nd,r0,c0", [(40, 64, 9, 64, 0, 0), (33, 50, 7, 1000, 17, 301)])
def test_scheme_converges_to_the_fp64_product(m, k, n, nd, r0, c0):
    rng = np.random.default_rng(m + k)
    b = rng.standard_normal((k, n)) + 1j * rng.standard_normal((k, n))
    b[:, 0] *= 1e-30  # column scaling keeps small columns accurate
    want = oracle.dft_matrix(nd, r0, m, c0, k) @ b
    errs = []
    for s in range(1, 7):
        x = oracle.dft_split_fp32(nd, r0, m, c0, b, s)
        errs.append(oracle.rel_fro(x, want))
        # normwise bound: each operand within 2^(-8s-1), dropped pairs below 2^(-8(s+1))
        assert errs[-1] < 8 * k * 2.0 ** (-8 * s) / np.sqrt(k)
        assert oracle.rel_fro(x[:, :1], want[:, :1]) < 8 * 2.0 ** (-8 * s) * np.sqrt(k)
    assert errs[-1] < 1e-14
    assert all(errs[i + 1] < errs[i] for i in range(4))
