"""CPU-side checks of the C-ABI library: it loads, exports every entry point
declared in include/*.h, and reports misuse without touching the GPU."""
import glob
import os
import re

import pytest

from qcu_mrhs_amd import _lib, api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        txt = re.sub(r"//[^\n]*", "", txt)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", txt, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    names = _declared()
    reference = {"initGridSize", "pushBackFermions", "loadQcuGauge", "getDslash", "start_dslash",
                 "mat_Qcu", "qcuInvert", "finalizeQcu", "gauge_eo_precondition",
                 "gauge_reverse_eo_precondition", "read_gauge_from_file"}
    assert reference <= names
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


@pytest.mark.parametrize("call", [lambda: api.start_dslash(0, 0), lambda: api.mat_Qcu(0),
                                  lambda: api.qcuInvert(10, 1e-10), lambda: api.getDslash(0, 0.1),
                                  lambda: api.pushBackFermions(0, 0)])
def test_calls_before_init_report_not_initialized(call):
    api.finalizeQcu()
    with pytest.raises(api.QcuError) as e:
        call()
    assert e.value.code == 1
    assert _lib.load().qcuLastError() == 0      # check() cleared it


def _c_params():
    """{function: [parameter C type strings]} from include/*.h (comments stripped)."""
    out = {}
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        txt = re.sub(r"//[^\n]*", "", txt)
        txt = re.sub(r"^\s*#.*$", "", txt, flags=re.M)
        for stmt in re.split(r"[;{}]", txt):
            m = re.fullmatch(r"\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(([^()]*)\)\s*", stmt,
                             flags=re.S)
            if not m or m.group(1) in ("if", "while", "for", "return", "sizeof"):
                continue
            params = [p.strip() for p in m.group(2).split(",") if p.strip() and p.strip() != "void"]
            out[m.group(1)] = params
    return out


def _ctype_of(param):
    """ctypes type expected for one C parameter declaration."""
    import ctypes
    p = re.sub(r"\s+", " ", param)
    p = re.sub(r"\b(const)\b", "", p).strip()
    arr = re.search(r"\[(\d*)\]\s*$", p)
    p = re.sub(r"\[\d*\]\s*$", "", p).strip()
    stars = p.count("*") + (1 if arr else 0)
    base = re.sub(r"\*", " ", p).split()
    base = base[:-1] if len(base) > 1 else base  # drop the parameter name
    base = " ".join(base)
    if stars == 0:
        return {"int": ctypes.c_int, "double": ctypes.c_double, "long": ctypes.c_long, "size_t": ctypes.c_size_t,
                "unsigned": ctypes.c_uint}[base]
    if base == "char" and stars == 1:
        return ctypes.c_char_p
    if stars == 1 and base in ("int", "double"):
        return ctypes.POINTER({"int": ctypes.c_int, "double": ctypes.c_double}[base])
    return "pointer"


def test_ctypes_prototypes_match_the_headers():
    """Every prototype the Python binding declares (qcu_mrhs_amd/_lib.py) has
    the parameter count and types of its C declaration in include/*.h: a
    mismatched prototype passes arguments of the wrong width or in the wrong
    slots, which corrupts host memory instead of raising."""
    import ctypes
    L = _lib.load()
    decl = _c_params()
    checked = 0
    for name, params in sorted(decl.items()):
        f = getattr(L, name, None)
        if f is None or f.argtypes is None:
            continue
        assert len(f.argtypes) == len(params), (name, params, f.argtypes)
        for p, t in zip(params, f.argtypes):
            want = _ctype_of(p)
            if want == "pointer":
                is_ptr = t is ctypes.c_void_p or t is ctypes.c_char_p or (
                    isinstance(t, type) and issubclass(t, ctypes._Pointer))
                assert is_ptr, (name, p, t)
            else:
                assert t == want, (name, p, t, want)
        checked += 1
    assert checked == len(decl) >= 50, (checked, len(decl))


def test_launch_shape_setters_validate_without_the_gpu():
    """qcuDslashTune and qcuGroupsDslashTune only record a process-wide choice:
    out-of-range variants are rejected with kQcuErrorInvalidValue (2) before
    any GPU work, valid ones accepted, and 0 restores the automatic choice."""
    L = _lib.load()
    for bad in ((4, 0), (0, 7), (1, 3), (-1, 0)):
        assert L.qcuGroupsDslashTune(*bad) == 2, bad
        assert L.qcuLastError() == 2
        L.qcuClearError()
    for ok in ((5, 5), (3, 2), (2, 3), (0, 0)):
        assert L.qcuGroupsDslashTune(*ok) == 0, ok
    for bad in ((9, 0, 0, 0), (3, -1, 0, 0), (2, 0, 0, -1), (4, 0, 0, -2)):
        assert L.qcuDslashTune(*bad) == 2, bad
        L.qcuClearError()
    assert L.qcuDslashTune(4, 4, 8, 12) == 0
    assert L.qcuDslashTune(0, 0, 0, 0) == 0
    assert L.qcuLastError() == 0
