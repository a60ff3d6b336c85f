"""The allocation-free dot product qcuBlasDotcNoAlloc (include/qcu_ops.h)
against qcuBlasDotc on the same device inputs, through the C ABI: two calls
in a row (the second on a smaller input, so the shared device scratch holds
the first call's partials when it starts), FP64 / FP32 / FP16 storage, RHS
counts on both sides of the 32 where the block count starts to shrink, and
the error for an RHS count beyond the scratch."""
import numpy as np
import pytest

from oracle import oracle
from qcu_mrhs_amd import api

pytestmark = pytest.mark.gpu


def _dots(n, R, prec, seed):
    g = np.random.default_rng(seed)
    a = g.standard_normal(n * R) + 1j * g.standard_normal(n * R)
    b = g.standard_normal(n * R) + 1j * g.standard_normal(n * R)
    da, db = api.DeviceArray.from_complex(a, prec), api.DeviceArray.from_complex(b, prec)
    d_ref, d_new = api.DeviceArray(16 * R), api.DeviceArray(16 * R)
    L = api.lib()
    api.check(L.qcuBlasDotc(d_ref.ptr, da.ptr, db.ptr, n, R, prec))
    api.check(L.qcuBlasDotcNoAlloc(d_new.ptr, da.ptr, db.ptr, n, R, prec))
    api.synchronize()
    # FP64 oracle on the inputs as stored
    a_s = api.from_storage(api.to_storage(a, prec), prec)
    b_s = api.from_storage(api.to_storage(b, prec), prec)
    return d_ref.complex(R), d_new.complex(R), oracle.dotc(a_s, b_s, R)


@pytest.mark.parametrize("prec", [2, 1, 0])
@pytest.mark.parametrize("sizes", [((4 * 3 * 20000 + 5, 16), (4 * 3 * 300 + 1, 5)),
                                   ((4 * 3 * 9000 + 3, 48), (4 * 3 * 100, 33))])
def test_no_alloc_dotc_matches_dotc_on_two_calls(prec, sizes):
    for k, (n, R) in enumerate(sizes):
        print("case no-alloc dotc", prec, n, R, flush=True)
        ref, new, ora = _dots(n, R, prec, 100 * k + R + prec)
        if R <= 32:
            np.testing.assert_array_equal(new, ref)      # same blocks, same summation order
        else:
            np.testing.assert_allclose(new, ref, rtol=1e-13, atol=0)
        np.testing.assert_allclose(new, ora, rtol=1e-12)


def test_no_alloc_dotc_rejects_too_many_rhs():
    d = api.DeviceArray(16 * 300)
    code = api.lib().qcuBlasDotcNoAlloc(d.ptr, d.ptr, d.ptr, 1, 257, 2)
    assert code == 2
    assert b"exceeds the maximum" in api.lib().qcuLastErrorString()
    api.lib().qcuClearError()
