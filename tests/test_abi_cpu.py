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
