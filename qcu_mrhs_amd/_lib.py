"""ctypes binding of libqcu.so (include/qcu.h, include/qcu_ops.h).

The shared library is the product: everything here only declares prototypes.
Loading fails loudly if the library has not been built -- there is no CPU
fallback.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqcu.so")

PRECISION = {"half": 0, "single": 1, "double": 2, "bf16": 4}
kQcuPrecisionHalf, kQcuPrecisionSingle, kQcuPrecisionDouble, kQcuPrecisionBFloat16 = 0, 1, 2, 4
kQcuDslashWilson = 0


class QcuParam(ctypes.Structure):
    _fields_ = [("lattice_size", ctypes.c_int * 4)]


class QcuGrid(ctypes.Structure):
    _fields_ = [("grid_size", ctypes.c_int * 4)]


class QcuSolveInfo(ctypes.Structure):
    _fields_ = [
        ("converged", ctypes.c_int),
        ("iterations", ctypes.c_int),
        ("reliable_updates", ctypes.c_int),
        ("restarts", ctypes.c_int),
        ("true_rel_residual", ctypes.c_double),
        ("seconds", ctypes.c_double),
        ("preconditioned", ctypes.c_int),
        ("fallback", ctypes.c_int),
    ]


class QcuRhsGroup(ctypes.Structure):
    _fields_ = [
        ("m_rhs", ctypes.c_int),
        ("fermion_out", ctypes.POINTER(ctypes.c_void_p)),
        ("fermion_in", ctypes.POINTER(ctypes.c_void_p)),
    ]


class QcuGroupInfo(ctypes.Structure):
    _fields_ = [
        ("converged", ctypes.c_int),
        ("iterations", ctypes.c_int),
        ("outer_steps", ctypes.c_int),
        ("sites_per_wave", ctypes.c_int),
        ("true_rel_residual", ctypes.c_double),
        ("seconds", ctypes.c_double),
        ("batch_seconds", ctypes.c_double),
    ]


_lib = None


def load() -> ctypes.CDLL:
    """Load libqcu.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    L = ctypes.CDLL(LIB_PATH)
    vp, i, d, l, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_long, ctypes.c_size_t
    ip4 = ctypes.POINTER(ctypes.c_int)
    sig = {
        "initGridSize": ([ctypes.POINTER(QcuGrid), ctypes.POINTER(QcuParam), i, i, i, i], None),
        "pushBackFermions": ([vp, vp], None),
        "loadQcuGauge": ([vp, i], None),
        "getDslash": ([i, d], None),
        "start_dslash": ([i, i], None),
        "mat_Qcu": ([i], None),
        "qcuInvert": ([i, d], None),
        "finalizeQcu": ([], None),
        "gauge_eo_precondition": ([vp, vp, i], None),
        "gauge_reverse_eo_precondition": ([vp, vp, i], None),
        "read_gauge_from_file": ([vp, ctypes.c_char_p], None),
        "qcuLastError": ([], i),
        "qcuLastErrorString": ([], ctypes.c_char_p),
        "qcuClearError": ([], None),
        "qcuInvertEx": ([i, d, ctypes.POINTER(QcuSolveInfo)], i),
        "qcuSetSolverOptions": ([i, d, i], i),
        "qcuSetCloverCoefficient": ([d], i),
        "qcuCloverBuild": ([vp, vp, vp, ip4, i, d, d], i),
        "qcuCloverApply": ([vp, vp, vp, ip4, i, i, i, i, vp, d], i),
        "qcuDeviceMalloc": ([sz], vp),
        "qcuDeviceFree": ([vp], i),
        "qcuMemcpyHtoD": ([vp, vp, sz], i),
        "qcuMemcpyDtoH": ([vp, vp, sz], i),
        "qcuMemsetD": ([vp, i, sz], i),
        "qcuDeviceSynchronize": ([], i),
        "qcuDslashMrhs": ([vp, vp, vp, ip4, i, i, i, i, i, vp, d], i),
        "qcuBlasXsay": ([vp, vp, vp, vp, l, i, i], i),
        "qcuBlasXpay": ([vp, vp, vp, vp, l, i, i], i),
        "qcuBlasAxpby": ([vp, vp, vp, vp, vp, l, i, i], i),
        "qcuBlasAxpbypcz": ([vp, vp, vp, vp, vp, vp, vp, l, i, i], i),
        "qcuBlasDotc": ([vp, vp, vp, l, i, i], i),
        "qcuBlasDotcNoAlloc": ([vp, vp, vp, l, i, i], i),
        "qcuBlasNorm": ([vp, vp, l, i, i], i),
        "qcuColorSpinorGather": ([vp, i, ctypes.POINTER(vp), i, l, i, i], i),
        "qcuColorSpinorScatter": ([ctypes.POINTER(vp), i, vp, i, l, i, i], i),
        "qcuCopyComplexVector": ([vp, i, vp, i, l], i),
        "qcuDslashTune": ([i, i, i, i], i),
        "qcuDslashAutotune": ([ip4, i, i, i, i, ip4, ctypes.POINTER(ctypes.c_double),
                               ctypes.POINTER(ctypes.c_double), i, ip4], i),
        "qcuDslashSetTuned": ([ip4, i, i, i, i, i, i, i], i),
        "qcuDslashTuneSave": ([ctypes.c_char_p], i),
        "qcuDslashTuneLoad": ([ctypes.c_char_p, ip4], i),
        "qcuInvertGroups": ([i, ctypes.POINTER(QcuRhsGroup), i, d, ctypes.POINTER(QcuGroupInfo)], i),
        "qcuGroupsDslashVariant": ([], i),
        "qcuGroupsDslashVariant64": ([], i),
        "qcuGroupsDslashTune": ([i, i], i),
        "qcuLgnnSetWeights": ([i, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)], i),
        "qcuLgnnGetWeights": ([ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)], i),
        "qcuLgnnApply": ([vp, vp, i], i),
        "qcuLgnnLossGrad": ([vp, vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                             ctypes.POINTER(ctypes.c_double)], i),
        "qcuLgnnCollect": ([i, d, i, i, ctypes.c_uint, ip4], i),
        "qcuLgnnTrain": ([i, d, ctypes.c_uint, ctypes.POINTER(ctypes.c_double)], i),
        "qcuLgnnClearData": ([], i),
        "qcuLgnnSetObjective": ([i], i),
        "qcuLgnnSetProjected": ([i], i),
        "qcuLgnnPowerLossGrad": ([vp, i, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                  ctypes.POINTER(ctypes.c_double)], i),
        "qcuTimerStart": ([], i),
        "qcuTimerStopMs": ([], d),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L
