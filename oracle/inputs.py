"""Seeded synthetic inputs for the five benchmark configurations.

TEST INFRASTRUCTURE ONLY (tests/, bench.py).  numpy ``default_rng(seed)``
(PCG64); draws happen in the order listed per function and dense matrices are
filled in column-major order.  The same arrays go to the GPU path and to the
oracle.
"""
from __future__ import annotations

import numpy as np


def cauchy_generators(n: int, rng: np.random.Generator):
    """d sorted U[0,1); w_j = d_j + (d_{j+1}-d_j)/2, w_n = d_n + 0.5/n; u, v ~ N(0,1).

    Returns B(n,4) = [D F U V] as PSMMA_CAUCHY expects (SRC/psmma_cauchy.f90:28-30)."""
    d = np.sort(rng.random(n))
    w = np.empty(n)
    w[:-1] = d[:-1] + 0.5 * (d[1:] - d[:-1])
    w[-1] = d[-1] + 0.5 / n
    u = rng.standard_normal(n)
    v = rng.standard_normal(n)
    return np.asfortranarray(np.stack([d, w, u, v], axis=1))


def dense_uniform(m: int, n: int, rng: np.random.Generator) -> np.ndarray:
    return np.asfortranarray(rng.uniform(-1.0, 1.0, size=m * n).reshape((m, n), order="F"))


def cauchy_dense_inputs(n: int = 2048, seed: int = 0):
    """Plain Cauchy-like x dense.  Returns (gens B(n,4), dense (n x n))."""
    rng = np.random.default_rng(seed)
    b = cauchy_generators(n, rng)
    return b, dense_uniform(n, n, rng)


def toeplitz_hankel_inputs(n: int = 8192, seed: int = 0, kind: str = "toeplitz"):
    """Toeplitz (or Hankel) x dense: generator of length 2n-1 ~ N(0,1), dense U[-1,1]."""
    rng = np.random.default_rng(seed)
    g = rng.standard_normal(2 * n - 1)
    return g, dense_uniform(n, n, rng)


def dft_inputs(n: int = 4096, ncols: int = 512, seed: int = 0) -> np.ndarray:
    """Complex dense operand of the DFT product: re and im i.i.d. N(0,1)."""
    rng = np.random.default_rng(seed)
    re = rng.standard_normal(n * ncols).reshape((n, ncols), order="F")
    im = rng.standard_normal(n * ncols).reshape((n, ncols), order="F")
    return np.asfortranarray(re + 1j * im)


def secular_roots(d: np.ndarray, z: np.ndarray, rho: float = 1.0, chunk: int = 1024):
    """Roots of 1 + rho * sum z_i^2 / (d_i - lam) = 0 for sorted d, rho > 0.

    Root j lies in (d_j, d_{j+1}) (the last one in (d_n, d_n + rho z'z)).  Each
    root is located relative to its nearer pole: with origin s = d_j (or
    d_{j+1}) the secular function is evaluated on the shifted poles d_i - s, so
    DIFL_j = lam_j - d_j and DIFR_j = lam_j - d_{j+1} come out with small
    relative error.  Safeguarded bisection to full precision, vectorised over
    a chunk of roots.  Returns (difl, difr)."""
    d = np.asarray(d, np.float64)
    z2 = np.asarray(z, np.float64) ** 2
    n = len(d)
    nrm2 = float(z2.sum())
    dn = np.append(d, d[-1] + rho * nrm2)
    difl = np.empty(n)
    difr = np.empty(n)
    for j0 in range(0, n, chunk):
        js = np.arange(j0, min(n, j0 + chunk))
        lo_p = dn[js]
        hi_p = dn[js + 1]
        mid = 0.5 * (lo_p + hi_p)
        # secular function at the midpoint decides which pole is the origin
        fm = 1.0 + rho * np.sum(z2[None, :] / (d[None, :] - mid[:, None]), axis=1)
        use_left = fm >= 0.0  # root left of the midpoint -> nearer d_j
        origin = np.where(use_left, lo_p, hi_p)
        delta = d[None, :] - origin[:, None]  # shifted poles
        # tau in (a, b) relative to origin
        a = np.where(use_left, 0.0, lo_p - hi_p)
        b = np.where(use_left, mid - lo_p, 0.0)
        for _ in range(200):
            t = 0.5 * (a + b)
            f = 1.0 + rho * np.sum(z2[None, :] / (delta - t[:, None]), axis=1)
            right = f < 0.0  # f increases in tau between poles: root is right of t
            a = np.where(right, t, a)
            b = np.where(right, b, t)
            if np.all((b - a) <= 2.0 * np.finfo(float).eps * np.maximum(np.abs(a), np.abs(b))):
                break
        tau = 0.5 * (a + b)
        difl[js] = np.where(use_left, tau, tau + (hi_p - lo_p))
        difr[js] = np.where(use_left, tau - (hi_p - lo_p), tau)
    return difl, difr


def _lam_minus_d(d, difl, difr, rows, cols):
    """lambda_c - d_r for index arrays rows (column vector) and cols (row vector), from the
    shifted roots: (d_c - d_r) + DIFL_c for r <= c, (d_{c+1} - d_r) + DIFR_c otherwise."""
    n = len(d)
    cp1 = np.minimum(cols + 1, n - 1)
    return np.where(rows <= cols, (d[cols] - d[rows]) + difl[cols],
                    (d[cp1] - d[rows]) + difr[cols])


def loewner_z(d, z, difl, difr, rho: float = 1.0, chunk: int = 512):
    """Loewner's z for computed roots (the z that makes lambda exact eigenvalues of
    D + rho z z^T): zhat_i^2 = prod_j (lam_j - d_i) / prod_{j != i} (d_j - d_i) / rho, with
    the sign of z_i; factors paired as ratios (lam_j - d_i) / (d_j - d_i)."""
    d = np.asarray(d, np.float64)
    n = len(d)
    out = np.empty(n)
    cols = np.arange(n)[None, :]
    for i0 in range(0, n, chunk):
        rows = np.arange(i0, min(n, i0 + chunk))[:, None]
        num = _lam_minus_d(d, difl, difr, rows, cols)
        den = d[cols] - d[rows]
        ratio = np.where(rows == cols, num, num / np.where(rows == cols, 1.0, den))
        out[i0:i0 + rows.shape[0]] = np.prod(ratio, axis=1)
    return np.copysign(np.sqrt(np.abs(out) / rho), z)


def column_normalisers(d, u, difl, difr, chunk: int = 1024):
    """v_c = 1 / || u ./ (d - lam_c) || with the differences the tile generator forms."""
    d = np.asarray(d, np.float64)
    n = len(d)
    v = np.empty(n)
    rows = np.arange(n)[:, None]
    for j0 in range(0, n, chunk):
        cols = np.arange(j0, min(n, j0 + chunk))[None, :]
        den = -_lam_minus_d(d, difl, difr, rows, cols)
        v[j0:j0 + cols.shape[1]] = 1.0 / np.linalg.norm(u[:, None] / den, axis=0)
    return v


def merge_poles_and_z(n: int = 16384, seed: int = 0):
    """The first draws of eig_merge_inputs (same seed, same order): sorted poles d and unit z."""
    rng = np.random.default_rng(seed)
    d = np.sort(rng.random(n))
    z = rng.standard_normal(n)
    z /= np.linalg.norm(z)
    return d, z, rng


def eig_merge_inputs(n: int = 16384, seed: int = 0, with_q: bool = True):
    """Eigenvector merge: d sorted U[0,1), z ~ N(0,1) normalised, rho = 1.

    Returns dict with f (= d, the poles), u (= z), v (column normalisers),
    difl, difr and optionally the orthogonal q (QR of an n x n Gaussian)."""
    d, z, rng = merge_poles_and_z(n, seed)
    difl, difr = secular_roots(d, z, 1.0)
    # column normaliser: v_j = 1 / || z / (d - lam_j) ||, with d_i - lam_j formed
    # exactly as the tile generator does
    v = column_normalisers(d, z, difl, difr)
    out = dict(f=d, u=z, v=v, difl=difl, difr=difr)
    if with_q:
        g = rng.standard_normal(n * n).reshape((n, n), order="F")
        q, _ = np.linalg.qr(g)
        out["q"] = np.asfortranarray(q)
    return out


def batch_inputs(nprob: int = 512, seed: int = 0, nmin: int = 64, nmax: int = 2048):
    """Batch of uneven structured products: sizes log-uniform in [nmin, nmax], k = n;
    kinds in a seeded shuffle, half Cauchy, a quarter each Toeplitz and Hankel.
    Each problem: dict(kind, n, gens, dense)."""
    rng = np.random.default_rng(seed)
    sizes = np.rint(np.exp(rng.uniform(np.log(nmin), np.log(nmax), size=nprob))).astype(int)
    kinds = np.array([0] * (nprob // 2) + [1] * (nprob // 4) + [2] * (nprob - nprob // 2 - nprob // 4))
    rng.shuffle(kinds)
    probs = []
    for n, kd in zip(sizes, kinds):
        n = int(n)
        if kd == 0:
            gens = cauchy_generators(n, rng)
        else:
            gens = rng.standard_normal(2 * n - 1)
        probs.append(dict(kind=int(kd), n=n, gens=gens, dense=dense_uniform(n, n, rng)))
    return probs
