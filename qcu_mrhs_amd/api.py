"""Python mirror of the reference's C interface, over libqcu.so.

The functions keep the reference's names and argument meaning
(include/qcu.h:14-29, src/interface.cu:20-141): ``initGridSize``,
``pushBackFermions``, ``loadQcuGauge``, ``getDslash``, ``start_dslash``,
``mat_Qcu``, ``qcuInvert``, ``finalizeQcu``, ``gauge_eo_precondition``,
``gauge_reverse_eo_precondition``, ``read_gauge_from_file``.  Fermion and
gauge arguments are :class:`DeviceArray` objects (device memory allocated
through the library).  Where the reference aborts on misuse, these raise
:class:`QcuError` (the library records the error code).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import PRECISION, QcuGrid, QcuGroupInfo, QcuParam, QcuRhsGroup, QcuSolveInfo

__all__ = [
    "QcuError", "DeviceArray", "PRECISION", "initGridSize", "pushBackFermions", "loadQcuGauge",
    "getDslash", "start_dslash", "mat_Qcu", "qcuInvert", "qcuInvertEx", "finalizeQcu",
    "gauge_eo_precondition", "gauge_reverse_eo_precondition", "read_gauge_from_file",
    "set_solver_options", "dslash_mrhs", "to_storage", "from_storage", "synchronize", "invert_groups",
]


class QcuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"qcu error {code}: {msg}")
        self.code = code


def lib():
    return _lib.load()


def check(code=None):
    """Raise QcuError if the library recorded an error (and clear it)."""
    L = lib()
    err = L.qcuLastError() if code is None else code
    if err:
        msg = L.qcuLastErrorString().decode()
        L.qcuClearError()
        raise QcuError(err, msg)


def synchronize():
    check(lib().qcuDeviceSynchronize())


# ---------------------------------------------------------------- storage
def _bf16_bits(x: np.ndarray) -> np.ndarray:
    """float32 -> bfloat16 bit patterns, round to nearest even."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32)
    rounding = ((u >> 16) & 1) + 0x7FFF
    return ((u + rounding) >> 16).astype(np.uint16)


def _bf16_to_f32(bits: np.ndarray) -> np.ndarray:
    return (bits.astype(np.uint32) << 16).view(np.float32)


def to_storage(a: np.ndarray, prec: int) -> np.ndarray:
    """complex numpy array -> raw storage in precision ``prec`` (interleaved re/im)."""
    a = np.ascontiguousarray(a)
    if prec == PRECISION["double"]:
        return a.astype(np.complex128)
    if prec == PRECISION["single"]:
        return a.astype(np.complex64)
    f = a.astype(np.complex64).view(np.float32)
    if prec == PRECISION["half"]:
        return f.astype(np.float16)
    if prec == PRECISION["bf16"]:
        return _bf16_bits(f)
    raise ValueError(prec)


def from_storage(raw: np.ndarray, prec: int) -> np.ndarray:
    """raw storage -> complex128."""
    if prec in (PRECISION["double"], PRECISION["single"]):
        return raw.astype(np.complex128)
    f = raw.astype(np.float32) if prec == PRECISION["half"] else _bf16_to_f32(raw)
    return f.view(np.complex64).astype(np.complex128)


def storage_bytes(count: int, prec: int) -> int:
    return count * {0: 4, 1: 8, 2: 16, 4: 4}[prec]


class DeviceArray:
    """A device buffer from the library's allocator."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = lib().qcuDeviceMalloc(max(self.nbytes, 1))
        if not self.ptr:
            check()
            raise QcuError(5, "allocation failed")

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        d.upload(a)
        return d

    @classmethod
    def from_complex(cls, a: np.ndarray, prec: int = 2) -> "DeviceArray":
        return cls.from_numpy(to_storage(a, prec))

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(lib().qcuMemcpyHtoD(self.ptr, a.ctypes.data, a.nbytes))

    def download(self, dtype, count=None) -> np.ndarray:
        dtype = np.dtype(dtype)
        count = self.nbytes // dtype.itemsize if count is None else count
        out = np.empty(count, dtype=dtype)
        check(lib().qcuMemcpyDtoH(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def complex(self, count: int, prec: int = 2) -> np.ndarray:
        dt = {0: np.float16, 1: np.complex64, 2: np.complex128, 4: np.uint16}[prec]
        n = count * (2 if prec in (0, 4) else 1)
        return from_storage(self.download(dt, n), prec)

    def zero(self) -> None:
        check(lib().qcuMemsetD(self.ptr, 0, self.nbytes))

    def offset(self, nbytes: int) -> int:
        return self.ptr + int(nbytes)

    def free(self) -> None:
        if self.ptr:
            lib().qcuDeviceFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _p(x):
    return x.ptr if isinstance(x, DeviceArray) else x


# ------------------------------------------------------ reference entry points
def initGridSize(lattice, n_color, m_rhs, inputFloatPrecision=2, dslashFloatPrecision=2,
                 grid=(1, 1, 1, 1)):
    g = QcuGrid((ctypes.c_int * 4)(*grid))
    p = QcuParam((ctypes.c_int * 4)(*lattice))
    lib().initGridSize(ctypes.byref(g), ctypes.byref(p), int(n_color), int(m_rhs),
                       int(inputFloatPrecision), int(dslashFloatPrecision))
    check()


def pushBackFermions(fermionOut, fermionIn):
    lib().pushBackFermions(_p(fermionOut), _p(fermionIn))
    check()


def loadQcuGauge(gauge, floatPrecision=2):
    lib().loadQcuGauge(_p(gauge), int(floatPrecision))
    check()


def getDslash(dslashType, mass):
    lib().getDslash(int(dslashType), float(mass))
    check()


def start_dslash(parity, daggerFlag=0):
    lib().start_dslash(int(parity), int(daggerFlag))
    check()


def mat_Qcu(daggerFlag=0):
    lib().mat_Qcu(int(daggerFlag))
    check()


def qcuInvert(max_iteration, p_max_prec):
    lib().qcuInvert(int(max_iteration), float(p_max_prec))
    check()


def qcuInvertEx(max_iteration, tol, raise_on_nonconvergence=False):
    info = QcuSolveInfo()
    code = lib().qcuInvertEx(int(max_iteration), float(tol), ctypes.byref(info))
    res = {k: getattr(info, k) for k, _ in QcuSolveInfo._fields_}
    res["code"] = code
    if code and (code != 6 or raise_on_nonconvergence):
        check(code)
    if code == 6:
        lib().qcuClearError()
    return res


def invert_groups(groups, max_iteration, tol, raise_on_nonconvergence=False):
    """Batched solve of uneven RHS groups (qcuInvertGroups).  ``groups`` is a
    list of (xs, bs) pairs of equal-length lists of device buffers (full-lattice
    fermions); returns one dict per group (converged, iterations, outer_steps,
    sites_per_wave, true_rel_residual, seconds, batch_seconds)."""
    n = len(groups)
    arr = (QcuRhsGroup * max(n, 1))()
    keep = []
    for k, (xs, bs) in enumerate(groups):
        if len(xs) != len(bs):
            raise ValueError("each group needs as many solutions as sources")
        outs = (ctypes.c_void_p * max(len(xs), 1))(*[_p(x) for x in xs])
        ins = (ctypes.c_void_p * max(len(bs), 1))(*[_p(b) for b in bs])
        keep += [outs, ins]
        arr[k].m_rhs = len(xs)
        arr[k].fermion_out = ctypes.cast(outs, ctypes.POINTER(ctypes.c_void_p))
        arr[k].fermion_in = ctypes.cast(ins, ctypes.POINTER(ctypes.c_void_p))
    info = (QcuGroupInfo * max(n, 1))()
    code = lib().qcuInvertGroups(n, arr, int(max_iteration), float(tol), info)
    if code and (code != 6 or raise_on_nonconvergence):
        check(code)
    if code == 6:
        lib().qcuClearError()
    return [{k: getattr(info[g], k) for k, _ in QcuGroupInfo._fields_} for g in range(n)]


def set_solver_options(inner_precision=2, delta=0.0, use_preconditioner=False):
    check(lib().qcuSetSolverOptions(int(inner_precision), float(delta), int(bool(use_preconditioner))))


def set_clover_coefficient(csw):
    """c_sw of the clover operator (getDslash(DSLASH_CLOVER, mass))."""
    check(lib().qcuSetCloverCoefficient(float(csw)))


DSLASH_WILSON, DSLASH_CLOVER = 0, 1


def finalizeQcu():
    lib().finalizeQcu()


def gauge_eo_precondition(prec_gauge, non_prec_gauge, precision=2):
    lib().gauge_eo_precondition(_p(prec_gauge), _p(non_prec_gauge), int(precision))
    check()


def gauge_reverse_eo_precondition(non_prec_gauge, prec_gauge, precision=2):
    lib().gauge_reverse_eo_precondition(_p(non_prec_gauge), _p(prec_gauge), int(precision))
    check()


def read_gauge_from_file(gauge, path):
    lib().read_gauge_from_file(_p(gauge), str(path).encode())
    check()


# --------------------------------------------------------- stateless ops
DSLASH_VARIANT = {"auto": 0, "lane_rhs": 1, "wave": 2, "mfma": 3, "lds": 4}


def dslash_tune(variant="auto", by=0, bz=0, spw=0):
    """Force the dslash launch shape for every later call (variant 'auto' resets)."""
    v = DSLASH_VARIANT[variant] if isinstance(variant, str) else int(variant)
    check(lib().qcuDslashTune(v, int(by), int(bz), int(spw)))


def dslash_autotune(lattice, n_color, m_rhs, precision, repeats=10):
    """Autotune the dslash launch shape for (lattice, Nc, R, precision); the
    chosen shape enters the tuning table.  Returns (best, best_us, records)
    with best = {variant, by, bz, spw} and records = every timed candidate
    (the first is the built-in default)."""
    L4 = (ctypes.c_int * 4)(*lattice)
    best = (ctypes.c_int * 4)()
    us = ctypes.c_double(0.0)
    cap = 512
    rec = (ctypes.c_double * (5 * cap))()
    n = ctypes.c_int(0)
    check(lib().qcuDslashAutotune(L4, int(n_color), int(m_rhs), int(precision), int(repeats), best,
                                  ctypes.byref(us), rec, cap, ctypes.byref(n)))
    names = ("variant", "by", "bz", "spw")
    records = [dict(zip(names + ("us",), (int(rec[5 * k]), int(rec[5 * k + 1]), int(rec[5 * k + 2]),
                                           int(rec[5 * k + 3]), float(rec[5 * k + 4]))))
               for k in range(min(n.value, cap))]
    return dict(zip(names, list(best))), us.value, records


def dslash_tune_save(path):
    check(lib().qcuDslashTuneSave(str(path).encode()))


def dslash_tune_load(path):
    n = ctypes.c_int(0)
    check(lib().qcuDslashTuneLoad(str(path).encode(), ctypes.byref(n)))
    return n.value


def dslash_mrhs(out, inp, gauge, lattice, n_color, m_rhs, parity, dagger=0, precision=2,
                xpay=None, coef=1.0):
    L4 = (ctypes.c_int * 4)(*lattice)
    check(lib().qcuDslashMrhs(_p(out), _p(inp), _p(gauge), L4, int(n_color), int(m_rhs),
                              int(parity), int(dagger), int(precision),
                              _p(xpay) if xpay is not None else None, float(coef)))


def clover_build(clover, clover_inv, gauge, lattice, n_color, csw, kappa):
    """FP64 clover blocks [2][Vh][2][2Nc][2Nc] (and their inverse, if
    ``clover_inv`` is not None) from an FP64 even/odd gauge field."""
    L4 = (ctypes.c_int * 4)(*lattice)
    check(lib().qcuCloverBuild(_p(clover), _p(clover_inv) if clover_inv is not None else None, _p(gauge), L4,
                               int(n_color), float(csw), float(kappa)))


def clover_apply(out, inp, blocks, lattice, n_color, m_rhs, precision=2, clover_precision=2, xpay=None, coef=0.0):
    """out = A in + coef * xpay on one parity (``blocks`` = that parity's blocks)."""
    L4 = (ctypes.c_int * 4)(*lattice)
    check(lib().qcuCloverApply(_p(out), _p(inp), _p(blocks), L4, int(n_color), int(m_rhs), int(precision),
                               int(clover_precision), _p(xpay) if xpay is not None else None, float(coef)))
