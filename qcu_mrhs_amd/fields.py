"""Seeded synthetic inputs: gauge links and multi-RHS sources.

Every generator is a pure function of (shape, seed) built on numpy's
Philox4x64 counter-based generator, so the GPU path and the CPU oracle
receive bit-identical inputs.

Layouts (the reference's, see include/point/qcu_point.cuh:67-87 and
include/kernel/copy_vector/copy_color_spinor.cuh:27-38):

* even/odd gauge   ``[mu 4][parity 2][t][z][y][x/2][Nc][Nc]`` complex
* lexicographic gauge ``[mu 4][t][z][y][x][Nc][Nc]`` complex
  (reader dims ``{Nd, Lt, Lz, Ly, Lx, Nc, 2*Nc}``, src/qcu_interface.cu:387-388)
* user fermion (one RHS) ``[parity 2][t][z][y][x/2][spin 4][Nc]`` complex,
  even half first (src/qcu_interface.cu:288-299)
* MRHS fermion ``[parity 2][t][z][y][x/2][spin 4][Nc][rhs R]`` complex
"""
from __future__ import annotations

import numpy as np

__all__ = [
    "rng",
    "haar_su_n",
    "exp_i_eps_h",
    "gauge_field",
    "gaussian_sources",
    "to_mrhs",
    "from_mrhs",
    "half_volume",
    "gauge_eo_to_lex",
    "gauge_lex_to_eo",
]


def rng(seed: int) -> np.random.Generator:
    """Philox4x64-10 generator keyed by ``seed``."""
    return np.random.Generator(np.random.Philox(key=int(seed)))


def half_volume(L) -> int:
    Lx, Ly, Lz, Lt = (int(v) for v in L)
    if Lx % 2:
        raise ValueError("Lx must be even for the even/odd layout")
    return Lx * Ly * Lz * Lt // 2


def _complex_gaussian(g: np.random.Generator, shape, scale=1.0) -> np.ndarray:
    re = g.standard_normal(shape)
    im = g.standard_normal(shape)
    out = np.empty(shape, dtype=np.complex128)
    out.real = re * scale
    out.imag = im * scale
    return out


def haar_su_n(count: int, Nc: int, g: np.random.Generator, chunk: int = 1 << 18) -> np.ndarray:
    """``count`` Haar-random SU(Nc) matrices.

    Mezzadri (2007), "How to generate random matrices from the classical
    compact groups": QR of a complex Gaussian matrix with the phases of R's
    diagonal moved into Q gives Haar U(N); dividing by det^(1/N) gives
    Haar SU(N) (the N choices of root differ by centre elements, under
    which Haar measure is invariant).
    """
    out = np.empty((count, Nc, Nc), dtype=np.complex128)
    for lo in range(0, count, chunk):
        hi = min(count, lo + chunk)
        z = _complex_gaussian(g, (hi - lo, Nc, Nc), 1.0 / np.sqrt(2.0))
        q, r = np.linalg.qr(z)
        d = np.diagonal(r, axis1=1, axis2=2)
        ph = d / np.abs(d)
        q = q * ph[:, None, :]
        det = np.linalg.det(q)
        q = q / (det ** (1.0 / Nc))[:, None, None]
        out[lo:hi] = q
    return out


def exp_i_eps_h(count: int, Nc: int, eps: float, g: np.random.Generator,
                chunk: int = 1 << 18) -> np.ndarray:
    """``count`` links U = exp(i*eps*H), H traceless Hermitian.

    H = (A + A^+)/2 - tr(.)/Nc with A a complex Gaussian matrix of unit
    variance per real component; the exponential is taken through the
    eigendecomposition H = W diag(l) W^+, U = W diag(exp(i eps l)) W^+,
    so det U = exp(i eps tr H) = 1.
    """
    out = np.empty((count, Nc, Nc), dtype=np.complex128)
    for lo in range(0, count, chunk):
        hi = min(count, lo + chunk)
        a = _complex_gaussian(g, (hi - lo, Nc, Nc))
        h = 0.5 * (a + np.conj(np.swapaxes(a, 1, 2)))
        tr = np.trace(h, axis1=1, axis2=2).real / Nc
        h[:, np.arange(Nc), np.arange(Nc)] -= tr[:, None]
        lam, w = np.linalg.eigh(h)
        ph = np.exp(1j * eps * lam)
        out[lo:hi] = (w * ph[:, None, :]) @ np.conj(np.swapaxes(w, 1, 2))
    return out


def gauge_field(L, Nc: int, seed: int, kind: str = "haar", eps: float = 0.0) -> np.ndarray:
    """Even/odd gauge field ``[4][2][Vh][Nc][Nc]`` (complex128).

    kind="haar": i.i.d. Haar SU(Nc) links.  kind="smooth": exp(i eps H).
    kind="unit": every link is the identity (free field).
    Links are i.i.d., generated in storage order.
    """
    Vh = half_volume(L)
    n = 4 * 2 * Vh
    if kind == "unit":
        u = np.zeros((n, Nc, Nc), dtype=np.complex128)
        u[:, np.arange(Nc), np.arange(Nc)] = 1.0
    elif kind == "haar":
        u = haar_su_n(n, Nc, rng(seed))
    elif kind == "smooth":
        u = exp_i_eps_h(n, Nc, eps, rng(seed))
    else:
        raise ValueError(f"unknown gauge kind {kind!r}")
    return u.reshape(4, 2, Vh, Nc, Nc)


def gaussian_sources(L, Nc: int, R: int, seed: int) -> np.ndarray:
    """R user-layout sources ``[R][2*Vh*4*Nc]`` complex128, i.i.d. N(0,1) per
    real component."""
    Vh = half_volume(L)
    return _complex_gaussian(rng(seed), (R, 2 * Vh * 4 * Nc))


def to_mrhs(vecs: np.ndarray, Nc: int) -> np.ndarray:
    """Stack R user vectors ``[R][nsites*4*Nc]`` into RHS-innermost
    ``[nsites][4][Nc][R]`` (flattened)."""
    R = vecs.shape[0]
    return np.ascontiguousarray(vecs.reshape(R, -1).T).reshape(-1)


def from_mrhs(mrhs: np.ndarray, R: int) -> np.ndarray:
    """Inverse of :func:`to_mrhs`."""
    return np.ascontiguousarray(mrhs.reshape(-1, R).T)


def _eo_site_maps(L):
    Lx, Ly, Lz, Lt = (int(v) for v in L)
    t, z, y, x = np.meshgrid(np.arange(Lt), np.arange(Lz), np.arange(Ly), np.arange(Lx),
                             indexing="ij")
    p = (x + y + z + t) & 1
    cb = ((t * Lz + z) * Ly + y) * (Lx // 2) + x // 2
    return p.reshape(-1), cb.reshape(-1)


def gauge_lex_to_eo(lex: np.ndarray, L, Nc: int) -> np.ndarray:
    """``[4][t][z][y][x][Nc][Nc]`` -> ``[4][2][Vh][Nc][Nc]``."""
    Vh = half_volume(L)
    p, cb = _eo_site_maps(L)
    lex = lex.reshape(4, 2 * Vh, Nc, Nc)
    eo = np.empty((4, 2, Vh, Nc, Nc), dtype=lex.dtype)
    eo[:, p, cb] = lex
    return eo


def gauge_eo_to_lex(eo: np.ndarray, L, Nc: int) -> np.ndarray:
    """Inverse of :func:`gauge_lex_to_eo`."""
    Vh = half_volume(L)
    p, cb = _eo_site_maps(L)
    eo = eo.reshape(4, 2, Vh, Nc, Nc)
    return np.ascontiguousarray(eo[:, p, cb])
