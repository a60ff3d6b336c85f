"""GPU parity: dslash, BLAS, reductions and layouts through the C ABI
(libqcu.so) against the CPU oracle on the same seeded inputs."""
import sys

import numpy as np
import pytest

from oracle import oracle
from qcu_mrhs_amd import api, fields

pytestmark = pytest.mark.gpu

D, S, H, B = 2, 1, 0, 4


def _case(*a):
    print("case", *a, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("L,Nc,R", [((8, 8, 8, 8), 3, 4), ((4, 4, 4, 8), 2, 3), ((4, 4, 4, 4), 4, 5),
                                    ((4, 4, 2, 4), 8, 2), ((6, 4, 4, 4), 3, 1), ((4, 4, 4, 4), 3, 16),
                                    ((4, 2, 4, 6), 3, 12)])
def test_dslash_fp64_matches_oracle(L, Nc, R):
    _case("dslash fp64", L, Nc, R)
    gauge = fields.gauge_field(L, Nc, 100 + Nc, kind="haar")
    d_g = api.DeviceArray.from_complex(gauge.reshape(-1))
    Vh = fields.half_volume(L)
    src = fields.gaussian_sources(L, Nc, R, 200 + R).reshape(R, 2, -1)
    for parity in (0, 1):
        inp = fields.to_mrhs(src[:, 1 - parity], Nc)
        d_in = api.DeviceArray.from_complex(inp)
        d_out = api.DeviceArray(inp.nbytes)
        for dagger in (0, 1):
            api.dslash_mrhs(d_out, d_in, d_g, L, Nc, R, parity, dagger)
            got = d_out.complex(inp.size)
            ref = oracle.dslash(inp, gauge, L, Nc, R, parity, dagger)
            assert _rel(got, ref) < 1e-13
            assert np.abs(got - ref).max() < 1e-13 * np.abs(ref).max()


def test_dslash_fp64_production_shape():
    """16^3x32, SU(3), R=12, exp(i 0.3 H) links: relative error <= 1e-13."""
    L, Nc, R = (16, 16, 16, 32), 3, 12
    _case("dslash fp64 production", L, Nc, R)
    gauge = fields.gauge_field(L, Nc, 2024, kind="smooth", eps=0.3)
    src = fields.gaussian_sources(L, Nc, R, 2025).reshape(R, 2, -1)
    inp = fields.to_mrhs(src[:, 0], Nc)
    d_g = api.DeviceArray.from_complex(gauge.reshape(-1))
    d_in = api.DeviceArray.from_complex(inp)
    d_out = api.DeviceArray(inp.nbytes)
    api.dslash_mrhs(d_out, d_in, d_g, L, Nc, R, 1, 0)
    got = d_out.complex(inp.size)
    ref = oracle.dslash(inp, gauge, L, Nc, R, 1, 0)
    assert _rel(got, ref) < 1e-13


@pytest.mark.parametrize("prec,tol", [(S, 2e-7), (H, 2e-3), (B, 1.5e-2)])
def test_dslash_low_precision_error_bound(prec, tol):
    """Inner-solver dslash in FP32 / FP16 / BF16 storage (FP32 arithmetic):
    normwise error against the FP64 oracle below the storage rounding bound."""
    L, Nc, R = (8, 8, 8, 8), 3, 8
    _case("dslash lowprec", prec)
    gauge = fields.gauge_field(L, Nc, 5, kind="smooth", eps=0.5)
    src = fields.gaussian_sources(L, Nc, R, 6).reshape(R, 2, -1)
    inp = fields.to_mrhs(src[:, 0], Nc)
    d_g = api.DeviceArray.from_complex(gauge.reshape(-1), prec)
    d_in = api.DeviceArray.from_complex(inp, prec)
    d_out = api.DeviceArray(api.storage_bytes(inp.size, prec))
    api.dslash_mrhs(d_out, d_in, d_g, L, Nc, R, 1, 0, precision=prec)
    got = d_out.complex(inp.size, prec)
    ref = oracle.dslash(inp, gauge, L, Nc, R, 1, 0)
    err = _rel(got, ref)
    print(f"precision {prec}: dslash normwise relative error vs FP64 oracle = {err:.3e}")
    assert err < tol


def test_dslash_xpay_epilogue():
    L, Nc, R, kappa = (4, 4, 4, 8), 3, 3, 0.13
    gauge = fields.gauge_field(L, Nc, 9, kind="haar")
    src = fields.gaussian_sources(L, Nc, R, 10)
    full = fields.to_mrhs(src, Nc)
    n = full.size // 2
    d_g = api.DeviceArray.from_complex(gauge.reshape(-1))
    d_full = api.DeviceArray.from_complex(full)
    d_out = api.DeviceArray(n * 16)
    # out_1 = in_1 - kappa D_1 in_0   == odd half of M in
    api.dslash_mrhs(d_out, d_full, d_g, L, Nc, R, 1, 0, xpay=d_full.offset(n * 16), coef=-kappa)
    ref = oracle.mat(full, gauge, L, Nc, R, kappa)[n:]
    assert _rel(d_out.complex(n), ref) < 1e-13


@pytest.mark.parametrize("prec", [D, S])
def test_blas_reference_semantics(prec):
    """Shapes of the reference BLAS tests (R=3, n=1024^2+1)."""
    _case("blas", prec)
    g = np.random.default_rng(1)
    R, n = 3, 1024 * 1024 + 1
    x, y, z = (g.integers(0, 16, n * R) + 1j * (g.integers(0, 16, n * R) + 1) for _ in range(3))
    a, b, c = (g.standard_normal(R) + 1j * g.standard_normal(R) for _ in range(3))
    dx, dy, dz = (api.DeviceArray.from_complex(v, prec) for v in (x, y, z))
    da, db, dc = (api.DeviceArray.from_complex(v) for v in (a, b, c))
    dres = api.DeviceArray(api.storage_bytes(n * R, prec))
    L = api.lib()
    tol = 1e-15 if prec == D else 1e-6
    cases = [
        (lambda: L.qcuBlasXsay(dres.ptr, dx.ptr, da.ptr, dy.ptr, n, R, prec), oracle.cxsay(x, a, y, R)),
        (lambda: L.qcuBlasXpay(dres.ptr, dx.ptr, da.ptr, dy.ptr, n, R, prec), oracle.cxpay(x, a, y, R)),
        (lambda: L.qcuBlasAxpby(dres.ptr, da.ptr, dx.ptr, db.ptr, dy.ptr, n, R, prec),
         oracle.caxpby(a, x, b, y, R)),
        (lambda: L.qcuBlasAxpbypcz(dres.ptr, da.ptr, dx.ptr, db.ptr, dy.ptr, dc.ptr, dz.ptr, n, R, prec),
         oracle.caxpbypcz(a, x, b, y, c, z, R)),
    ]
    for call, ref in cases:
        api.check(call())
        assert _rel(dres.complex(n * R, prec), ref) < tol
    dd = api.DeviceArray(16 * R)
    api.check(L.qcuBlasDotc(dd.ptr, dx.ptr, dy.ptr, n, R, prec))
    dot1 = dd.complex(R)
    ref_dot = oracle.dotc(x, y, R) if prec == D else oracle.dotc(x.astype(np.complex64), y.astype(np.complex64), R)
    np.testing.assert_allclose(dot1, ref_dot, rtol=1e-14)
    api.check(L.qcuBlasDotc(dd.ptr, dx.ptr, dy.ptr, n, R, prec))
    np.testing.assert_array_equal(dd.complex(R), dot1)          # bit-reproducible
    dn = api.DeviceArray(8 * R)
    api.check(L.qcuBlasNorm(dn.ptr, dx.ptr, n, R, prec))
    np.testing.assert_allclose(dn.download(np.float64), oracle.norm(x, R), rtol=1e-14)


@pytest.mark.parametrize("R", [1, 5, 12, 16, 17, 64])
def test_reductions_odd_rhs_counts(R):
    g = np.random.default_rng(R)
    n = 4 * 3 * 1000 + 7
    a = g.standard_normal(n * R) + 1j * g.standard_normal(n * R)
    b = g.standard_normal(n * R) + 1j * g.standard_normal(n * R)
    da, db = api.DeviceArray.from_complex(a), api.DeviceArray.from_complex(b)
    dd = api.DeviceArray(16 * R)
    api.check(api.lib().qcuBlasDotc(dd.ptr, da.ptr, db.ptr, n, R, D))
    np.testing.assert_allclose(dd.complex(R), oracle.dotc(a, b, R), rtol=1e-12)


def test_gather_scatter_and_copy():
    L, Nc, R = (4, 4, 4, 4), 3, 5
    Vh = fields.half_volume(L)
    src = fields.gaussian_sources(L, Nc, R, 3)
    dsrc = [api.DeviceArray.from_complex(v) for v in src]
    ptrs = (api.ctypes.c_void_p * R)(*[d.ptr for d in dsrc]) if hasattr(api, "ctypes") else None
    import ctypes
    ptrs = (ctypes.c_void_p * R)(*[d.ptr for d in dsrc])
    nsites = 2 * Vh
    dm = api.DeviceArray(src.size * 16)
    api.check(api.lib().qcuColorSpinorGather(dm.ptr, D, ptrs, D, nsites, Nc, R))
    m = dm.complex(src.size)
    np.testing.assert_array_equal(m, oracle.gather(list(src), Nc))
    douts = [api.DeviceArray(v.nbytes) for v in src]
    optrs = (ctypes.c_void_p * R)(*[d.ptr for d in douts])
    api.check(api.lib().qcuColorSpinorScatter(optrs, D, dm.ptr, D, nsites, Nc, R))
    for r in range(R):
        np.testing.assert_array_equal(douts[r].complex(src[r].size), src[r])
    # R = 1 gather is the identity (tests/test_gather/test_gather.cu:32-85)
    one = (ctypes.c_void_p * 1)(dsrc[0].ptr)
    api.check(api.lib().qcuColorSpinorGather(dm.ptr, D, one, D, nsites, Nc, 1))
    np.testing.assert_array_equal(dm.complex(src[0].size), src[0])
    # FP64 -> FP32 -> FP64 copy
    d32 = api.DeviceArray(src[0].size * 8)
    api.check(api.lib().qcuCopyComplexVector(d32.ptr, S, dsrc[0].ptr, D, src[0].size))
    np.testing.assert_array_equal(d32.complex(src[0].size, S), src[0].astype(np.complex64))
