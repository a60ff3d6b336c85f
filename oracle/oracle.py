"""ctypes front end of the CPU FP64 oracle (oracle/qcu_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the CPU-baseline leg of ``bench.py``, always as the checker or the
timed CPU baseline, never as a product code path.

What pins each oracle function (asserted in tests/test_oracle_*.py):

* ``dslash``     free-field plane-wave closed form; exact adjoint under the
                 dagger flag; gamma5-hermiticity; gauge covariance under random
                 SU(N) transforms; parity structure; RHS independence; equality
                 with an independent dense operator built from the explicit
                 4x4 gamma matrices (oracle/dense.py) on small lattices.
* ``mat``        dense solve of the full M = 1 - kappa*D (numpy) reproduces b.
* ``schur``      equals the Schur complement taken from the dense full M.
* ``bicgstab``   true full-system residual ||b - M x|| / ||b|| <= tol
                 (computed independently through the dense M), agreement
                 with numpy.linalg.solve on small lattices.
* BLAS, dot, norm  the reference unit tests' semantics (tests/test_blas/*.cu,
                 tests/test_norm/test_norm.cu:47-68,
                 tests/test_inner_prod/test_inner_prod.cu:22-28) on integer data
                 of the same shape as those tests.
* gather/scatter the reference gather test (tests/test_gather/test_gather.cu:
                 32-85: R=1 gather is the identity) and the round trip.
* ``clover``     unit links give A = 1; a constant abelian field gives the
                 closed form 1 - kappa c_sw sigma_{mu nu} sin(B) tau; equality with
                 the whole-lattice numpy construction from the explicit 4x4
                 gamma matrices (oracle/dense.py); Hermitian blocks; gauge
                 covariance; the sign and normalisation of the term by the
                 O(a) cancellation it exists for (the would-be zero modes of a
                 constant-field background move to O(a^2) at c_sw = 1).
* ``clover_apply``, ``clover_invert``  the dense block-diagonal matrix and
                 numpy.linalg.inv.
* ``gauge_eo``   parity unpinned: the reference's converter lives in the
                 absent qcu_io submodule; only the round trip and agreement
                 with the site convention of include/point/qcu_point.cuh are
                 checked.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_c128p = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build() -> None:
    subprocess.run(["make", "-B", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, i, l, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_long, ctypes.c_double
        L.oracle_dslash.argtypes = [_c128p, _c128p, _c128p, _i32p, i, i, i, i]
        L.oracle_mat.argtypes = [_c128p, _c128p, _c128p, _i32p, i, i, d, i]
        L.oracle_schur.argtypes = [_c128p, _c128p, _c128p, _c128p, _i32p, i, i, d, i]
        L.oracle_bicgstab.argtypes = [_c128p, _c128p, _c128p, _i32p, i, i, d, i, d,
                                      ctypes.POINTER(ctypes.c_int), _f64p]
        L.oracle_bicgstab.restype = i
        L.oracle_cxsay.argtypes = [_c128p, _c128p, _c128p, _c128p, l, i]
        L.oracle_cxpay.argtypes = [_c128p, _c128p, _c128p, _c128p, l, i]
        L.oracle_caxpby.argtypes = [_c128p, _c128p, _c128p, _c128p, _c128p, l, i]
        L.oracle_caxpbypcz.argtypes = [_c128p] * 7 + [l, i]
        L.oracle_dotc.argtypes = [_c128p, _c128p, _c128p, l, i]
        L.oracle_norm.argtypes = [_f64p, _c128p, l, i]
        L.oracle_gather.argtypes = [_c128p, ctypes.POINTER(vp), l, i, i]
        L.oracle_scatter.argtypes = [ctypes.POINTER(vp), _c128p, l, i, i]
        L.oracle_gauge_eo.argtypes = [_c128p, _c128p, _i32p, i, i]
        L.oracle_clover.argtypes = [_c128p, _c128p, _i32p, i, d, d]
        L.oracle_clover_apply.argtypes = [_c128p, _c128p, _c128p, l, i, i]
        L.oracle_clover_invert.argtypes = [_c128p, _c128p, l, i]
        L.oracle_num_threads.restype = i
        L.oracle_set_num_threads.argtypes = [i]
        _lib = L
    return _lib


def _L(L):
    return np.ascontiguousarray(np.asarray(L, dtype=np.int32))


def _c(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def half_volume(L) -> int:
    return int(np.prod(np.asarray(L, dtype=np.int64))) // 2


def dslash(inp, gauge, L, Nc, R, parity, dagger=0):
    """One-parity MRHS hopping term; ``inp``/result are half vectors
    ``[Vh][4][Nc][R]``; ``parity`` is the output parity."""
    n = half_volume(L) * 4 * Nc * R
    out = np.empty(n, dtype=np.complex128)
    lib().oracle_dslash(out, _c(inp).reshape(-1), _c(gauge).reshape(-1), _L(L), Nc, R,
                        int(parity), int(dagger))
    return out


def mat(inp, gauge, L, Nc, R, kappa, dagger=0):
    """Full-lattice M = 1 - kappa*D on ``[2][Vh][4][Nc][R]``."""
    out = np.empty(2 * half_volume(L) * 4 * Nc * R, dtype=np.complex128)
    lib().oracle_mat(out, _c(inp).reshape(-1), _c(gauge).reshape(-1), _L(L), Nc, R,
                     float(kappa), int(dagger))
    return out


def schur(inp, gauge, L, Nc, R, kappa, dagger=0):
    """Odd Schur operator  in - kappa^2 D_oe D_eo in."""
    n = half_volume(L) * 4 * Nc * R
    out = np.empty(n, dtype=np.complex128)
    tmp = np.empty(n, dtype=np.complex128)
    lib().oracle_schur(out, _c(inp).reshape(-1), tmp, _c(gauge).reshape(-1), _L(L), Nc, R,
                       float(kappa), int(dagger))
    return out


def bicgstab(b, gauge, L, Nc, R, kappa, max_iter, tol):
    """Even/odd BiCGStab.  Returns (x, converged, iterations, rel_res[R])."""
    x = np.empty(2 * half_volume(L) * 4 * Nc * R, dtype=np.complex128)
    it = ctypes.c_int(0)
    res = np.zeros(R, dtype=np.float64)
    ok = lib().oracle_bicgstab(x, _c(b).reshape(-1), _c(gauge).reshape(-1), _L(L), Nc, R,
                               float(kappa), int(max_iter), float(tol), ctypes.byref(it), res)
    return x, bool(ok), int(it.value), res


def _scal(a, R):
    return _c(np.broadcast_to(np.asarray(a, dtype=np.complex128), (R,)))


def cxsay(x, a, y, R):
    res = np.empty_like(_c(x))
    lib().oracle_cxsay(res, _c(x), _scal(a, R), _c(y), x.size // R, R)
    return res


def cxpay(x, a, y, R):
    res = np.empty_like(_c(x))
    lib().oracle_cxpay(res, _c(x), _scal(a, R), _c(y), x.size // R, R)
    return res


def caxpby(a, x, b, y, R):
    res = np.empty_like(_c(x))
    lib().oracle_caxpby(res, _scal(a, R), _c(x), _scal(b, R), _c(y), x.size // R, R)
    return res


def caxpbypcz(a, x, b, y, c, z, R):
    res = np.empty_like(_c(x))
    lib().oracle_caxpbypcz(res, _scal(a, R), _c(x), _scal(b, R), _c(y), _scal(c, R), _c(z),
                           x.size // R, R)
    return res


def dotc(a, b, R):
    res = np.empty(R, dtype=np.complex128)
    lib().oracle_dotc(res, _c(a), _c(b), a.size // R, R)
    return res


def norm(a, R):
    res = np.empty(R, dtype=np.float64)
    lib().oracle_norm(res, _c(a), a.size // R, R)
    return res


def gather(vecs, Nc):
    """``vecs`` [R][nsites*4*Nc] -> MRHS [nsites][4][Nc][R]."""
    vecs = [_c(v) for v in vecs]
    R = len(vecs)
    nsites = vecs[0].size // (4 * Nc)
    ptrs = (ctypes.c_void_p * R)(*[v.ctypes.data for v in vecs])
    dst = np.empty(nsites * 4 * Nc * R, dtype=np.complex128)
    lib().oracle_gather(dst, ptrs, nsites, Nc, R)
    return dst


def scatter(mrhs, Nc, R):
    mrhs = _c(mrhs)
    nsites = mrhs.size // (4 * Nc * R)
    outs = [np.empty(nsites * 4 * Nc, dtype=np.complex128) for _ in range(R)]
    ptrs = (ctypes.c_void_p * R)(*[v.ctypes.data for v in outs])
    lib().oracle_scatter(ptrs, mrhs, nsites, Nc, R)
    return np.stack(outs)


def gauge_eo(lex, L, Nc):
    eo = np.empty(4 * 2 * half_volume(L) * Nc * Nc, dtype=np.complex128)
    lib().oracle_gauge_eo(eo, _c(lex).reshape(-1), _L(L), Nc, 0)
    return eo


def gauge_eo_reverse(eo, L, Nc):
    lex = np.empty(4 * 2 * half_volume(L) * Nc * Nc, dtype=np.complex128)
    lib().oracle_gauge_eo(_c(eo).reshape(-1), lex, _L(L), Nc, 1)
    return lex


def clover(gauge, L, Nc, csw, kappa):
    """Clover term A = 1 - kappa c_sw sum_{mu<nu} sigma_{mu nu} F_{mu nu} as
    its two chirality blocks per site: ``[parity 2][Vh][chirality 2][2Nc][2Nc]``,
    block row/column (spin in the pair {0,1} or {2,3}) * Nc + colour."""
    Vh = half_volume(L)
    n = 2 * Nc
    out = np.empty(2 * Vh * 2 * n * n, dtype=np.complex128)
    lib().oracle_clover(out, _c(gauge).reshape(-1), _L(L), Nc, float(csw), float(kappa))
    return out.reshape(2, Vh, 2, n, n)


def clover_apply(blocks, inp, Nc, R):
    """A in on one parity: ``blocks`` [Vh][2][2Nc][2Nc], ``inp`` a half
    vector [Vh][4][Nc][R]."""
    blocks = _c(blocks)
    Vh = blocks.size // (2 * (2 * Nc) ** 2)
    out = np.empty(Vh * 4 * Nc * R, dtype=np.complex128)
    lib().oracle_clover_apply(out, blocks.reshape(-1), _c(inp).reshape(-1), Vh, Nc, R)
    return out


def clover_invert(blocks):
    """Inverse of every trailing n x n block (Gauss-Jordan, partial pivoting)."""
    blocks = _c(blocks)
    n = blocks.shape[-1]
    out = np.empty_like(blocks)
    lib().oracle_clover_invert(out.reshape(-1), blocks.reshape(-1), blocks.size // (n * n), n)
    return out


def clover_mat(inp, gauge, clov, L, Nc, R, kappa, dagger=0):
    """Full-lattice M = A - kappa D on ``[2][Vh][4][Nc][R]`` (A Hermitian, so
    the dagger flag only reaches D)."""
    inp = _c(inp).reshape(-1)
    out = mat(inp, gauge, L, Nc, R, kappa, dagger)
    h = out.size // 2
    for p in (0, 1):
        out[p * h:(p + 1) * h] += clover_apply(clov[p], inp[p * h:(p + 1) * h], Nc, R) - inp[p * h:(p + 1) * h]
    return out


def num_threads() -> int:
    return int(lib().oracle_num_threads())


def set_num_threads(n: int) -> None:
    lib().oracle_set_num_threads(int(n))
