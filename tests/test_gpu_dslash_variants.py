"""GPU parity of every dslash kernel variant (lane-per-RHS, wave-per-site
VALU, wave-per-site MFMA, LDS-staged x-row MFMA) and launch shape against the
FP64 oracle."""
import numpy as np
import pytest

from oracle import oracle
from qcu_mrhs_amd import api, fields

pytestmark = pytest.mark.gpu

D, S, H, B = 2, 1, 0, 4
TOL = {D: 1e-13, S: 2e-7, H: 2e-3, B: 1.5e-2}

SHAPES = [((8, 8, 8, 8), 3, 4), ((4, 4, 4, 4), 3, 16), ((4, 2, 4, 6), 3, 12), ((4, 4, 4, 8), 2, 3),
          ((4, 4, 4, 4), 4, 5), ((4, 4, 2, 4), 8, 2), ((6, 4, 4, 4), 3, 17), ((4, 4, 4, 4), 3, 32)]


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(autouse=True)
def _reset_tune():
    yield
    api.dslash_tune("auto")


@pytest.mark.parametrize("variant", ["lane_rhs", "wave", "mfma"])
@pytest.mark.parametrize("L,Nc,R", SHAPES)
def test_variant_matches_oracle(variant, L, Nc, R):
    print("case", variant, L, Nc, R, flush=True)
    precs = [H, B] if variant == "mfma" else [D, S, H, B]
    gauge = fields.gauge_field(L, Nc, 300 + Nc + R, kind="haar")
    src = fields.gaussian_sources(L, Nc, R, 400 + R).reshape(R, 2, -1)
    api.dslash_tune(variant)
    for prec in precs:
        d_g = api.DeviceArray.from_complex(gauge.reshape(-1), prec)
        for parity, dagger in ((0, 0), (1, 1), (1, 0)):
            inp = fields.to_mrhs(src[:, 1 - parity], Nc)
            xp = fields.to_mrhs(src[:, parity], Nc)
            d_in = api.DeviceArray.from_complex(inp, prec)
            d_xp = api.DeviceArray.from_complex(xp, prec)
            d_out = api.DeviceArray(api.storage_bytes(inp.size, prec))
            api.dslash_mrhs(d_out, d_in, d_g, L, Nc, R, parity, dagger, precision=prec)
            ref = oracle.dslash(inp, gauge, L, Nc, R, parity, dagger)
            got = d_out.complex(inp.size, prec)
            assert _rel(got, ref) < TOL[prec], (prec, parity, dagger, _rel(got, ref))
            # fused epilogue out = x - 0.13 D in
            api.dslash_mrhs(d_out, d_in, d_g, L, Nc, R, parity, dagger, precision=prec, xpay=d_xp, coef=-0.13)
            got = d_out.complex(inp.size, prec)
            assert _rel(got, xp - 0.13 * ref) < TOL[prec], (prec, "xpay")


@pytest.mark.parametrize("variant,by,bz,spw",
                         [(v, *t) for v in ("wave", "mfma") for t in ((1, 1, 1), (2, 4, 3), (4, 2, 5), (8, 8, 2))]
                         + [("mfma", 2, 4, -1), ("mfma", 8, 8, -5)])
def test_traversal_tiles_and_sites_per_wave(variant, by, bz, spw):
    """Every (y, z) tiling and sites-per-wave count covers each site exactly
    once (spw < 0: the persistent grid of the MFMA kernel)."""
    L, Nc, R = (6, 8, 8, 6), 3, 8
    prec = H if variant == "mfma" else D
    gauge = fields.gauge_field(L, Nc, 7, kind="haar")
    inp = fields.to_mrhs(fields.gaussian_sources(L, Nc, R, 8).reshape(R, 2, -1)[:, 0], Nc)
    api.dslash_tune(variant, by, bz, spw)
    d_g = api.DeviceArray.from_complex(gauge.reshape(-1), prec)
    d_in = api.DeviceArray.from_complex(inp, prec)
    d_out = api.DeviceArray(api.storage_bytes(inp.size, prec))
    d_out.zero()
    api.dslash_mrhs(d_out, d_in, d_g, L, Nc, R, 1, 0, precision=prec)
    ref = oracle.dslash(inp, gauge, L, Nc, R, 1, 0)
    got = d_out.complex(inp.size, prec)
    per_site = np.abs(got - ref).reshape(-1, 4 * Nc * R).max(axis=1) / np.abs(ref).max()
    assert per_site.max() < TOL[prec] * 10, np.argmax(per_site)


@pytest.mark.parametrize("prec", [H, B])
@pytest.mark.parametrize("spw", [-4, -8])
def test_mfma_persistent_grid_production_shape(prec, spw):
    """Persistent MFMA grid (spw < 0: a fixed number of workgroups per CU
    striding through their XCD's share of the sites) at 16^3 x 32 with two
    RHS chunks (R = 20), both parities and the xpay form, against the oracle."""
    L, Nc, R = (16, 16, 16, 32), 3, 20
    print("case persistent", prec, spw, flush=True)
    gauge = fields.gauge_field(L, Nc, 61, kind="smooth", eps=0.3)
    src = fields.gaussian_sources(L, Nc, R, 62).reshape(R, 2, -1)
    api.dslash_tune("mfma", 6, 8, spw)
    d_g = api.DeviceArray.from_complex(gauge.reshape(-1), prec)
    for parity, dagger in ((1, 0), (0, 1)):
        inp = fields.to_mrhs(src[:, 1 - parity], Nc)
        xp = fields.to_mrhs(src[:, parity], Nc)
        d_in = api.DeviceArray.from_complex(inp, prec)
        d_xp = api.DeviceArray.from_complex(xp, prec)
        d_out = api.DeviceArray(api.storage_bytes(inp.size, prec))
        ref = oracle.dslash(inp, gauge, L, Nc, R, parity, dagger)
        api.dslash_mrhs(d_out, d_in, d_g, L, Nc, R, parity, dagger, precision=prec)
        assert _rel(d_out.complex(inp.size, prec), ref) < TOL[prec]
        api.dslash_mrhs(d_out, d_in, d_g, L, Nc, R, parity, dagger, precision=prec, xpay=d_xp, coef=-0.13)
        assert _rel(d_out.complex(inp.size, prec), xp - 0.13 * ref) < TOL[prec]


@pytest.mark.parametrize("prec", [S, D])
@pytest.mark.parametrize("L,Nc,R", [((4, 4, 4, 4), 8, 18), ((4, 4, 4, 8), 4, 5), ((6, 4, 4, 4), 3, 17),
                                    ((4, 4, 4, 8), 2, 3), ((16, 16, 16, 16), 8, 16)])
def test_fp32_fp64_matrix_core_dslash_matches_oracle(L, Nc, R, prec):
    """FP32 storage on the FP32 matrix cores (v_mfma_f32_16x16x4_f32) and FP64
    storage on the FP64 ones (v_mfma_f64_16x16x4_f64): both parities, dagger,
    the xpay form, one and several sites per wavefront, RHS counts with partial
    16-RHS chunks, against the FP64 oracle at the storage precision's bound."""
    print("case mfma32/64", L, Nc, R, prec, flush=True)
    gauge = fields.gauge_field(L, Nc, 500 + Nc + R, kind="haar")
    src = fields.gaussian_sources(L, Nc, R, 600 + R).reshape(R, 2, -1)
    d_g = api.DeviceArray.from_complex(gauge.reshape(-1), prec)
    for spw in (1, 4):
        api.dslash_tune("mfma", 0, 0, spw)
        for parity, dagger in ((1, 0), (0, 1)):
            inp = fields.to_mrhs(src[:, 1 - parity], Nc)
            xp = fields.to_mrhs(src[:, parity], Nc)
            d_in = api.DeviceArray.from_complex(inp, prec)
            d_xp = api.DeviceArray.from_complex(xp, prec)
            d_out = api.DeviceArray(api.storage_bytes(inp.size, prec))
            ref = oracle.dslash(inp, gauge, L, Nc, R, parity, dagger)
            api.dslash_mrhs(d_out, d_in, d_g, L, Nc, R, parity, dagger, precision=prec)
            assert _rel(d_out.complex(inp.size, prec), ref) < TOL[prec], (spw, parity, dagger)
            api.dslash_mrhs(d_out, d_in, d_g, L, Nc, R, parity, dagger, precision=prec, xpay=d_xp, coef=-0.13)
            assert _rel(d_out.complex(inp.size, prec), xp - 0.13 * ref) < TOL[prec], (spw, "xpay")


@pytest.mark.parametrize("L,Nc,R,ts,by,bz", [((8, 4, 4, 6), 3, 17, 4, 0, 0), ((4, 6, 4, 8), 2, 3, 5, 0, 0),
                                             ((4, 4, 6, 4), 4, 5, 1, 0, 0), ((12, 4, 4, 8), 3, 16, 3, 0, 0),
                                             ((36, 4, 4, 4), 3, 16, 2, 0, 0), ((20, 4, 4, 6), 4, 32, 6, 0, 0),
                                             ((6, 4, 4, 4), 3, 12, 48, 0, 0), ((2, 4, 2, 4), 3, 4, 3, 0, 0),
                                             ((10, 2, 4, 2), 2, 20, 1, 0, 0), ((8, 8, 4, 4), 3, 16, 2, 4, 2),
                                             ((4, 6, 6, 4), 3, 5, 3, 6, 3), ((6, 12, 2, 6), 4, 16, 48, 12, 2),
                                             ((4, 6, 4, 4), 2, 3, 2, 3, 4)])
def test_lds_staged_dslash_matches_oracle(L, Nc, R, ts, by, bz):
    """LDS-staged MFMA kernel (a workgroup per 4-site segment of an x-row,
    sweeping `ts` t steps with a ring of three haloed full rows and four
    spin-projected side rows): ragged last t runs, Ts = 1 and Ts beyond Lt,
    segments running past the row (Lx/2 = 18, 10, 5) and rows shorter than a
    segment (Lx/2 = 1, 2, 3), Nc 2/3/4, partial 16-RHS chunks, extents of 2,
    explicit (y, z) blocks of the workgroup order (the whole axis, odd
    blocks), both parities, dagger and the xpay form, FP16 and BF16, against
    the oracle."""
    print("case lds", L, Nc, R, ts, by, bz, flush=True)
    gauge = fields.gauge_field(L, Nc, 900 + Nc + R, kind="haar")
    src = fields.gaussian_sources(L, Nc, R, 910 + R).reshape(R, 2, -1)
    api.dslash_tune("lds", by, bz, ts)
    for prec in (H, B):
        d_g = api.DeviceArray.from_complex(gauge.reshape(-1), prec)
        for parity, dagger in ((0, 0), (1, 1), (1, 0), (0, 1)):
            inp = fields.to_mrhs(src[:, 1 - parity], Nc)
            xp = fields.to_mrhs(src[:, parity], Nc)
            d_in = api.DeviceArray.from_complex(inp, prec)
            d_xp = api.DeviceArray.from_complex(xp, prec)
            d_out = api.DeviceArray(api.storage_bytes(inp.size, prec))
            d_out.zero()
            ref = oracle.dslash(inp, gauge, L, Nc, R, parity, dagger)
            api.dslash_mrhs(d_out, d_in, d_g, L, Nc, R, parity, dagger, precision=prec)
            got = d_out.complex(inp.size, prec)
            per_site = np.abs(got - ref).reshape(-1, 4 * Nc * R).max(axis=1) / np.abs(ref).max()
            assert _rel(got, ref) < TOL[prec] and per_site.max() < TOL[prec] * 10, \
                (prec, parity, dagger, _rel(got, ref), np.argmax(per_site))
            api.dslash_mrhs(d_out, d_in, d_g, L, Nc, R, parity, dagger, precision=prec, xpay=d_xp, coef=-0.13)
            assert _rel(d_out.complex(inp.size, prec), xp - 0.13 * ref) < TOL[prec], (prec, parity, dagger, "xpay")


def test_lds_staged_dslash_production_shape():
    """LDS-staged kernel at 24^3 x 48, SU(3), 16 RHS (three segments per
    x-row) with Ts = 12 and 5, FP16 and BF16, both parities, dagger and xpay,
    against the oracle."""
    L, Nc, R = (24, 24, 24, 48), 3, 16
    gauge = fields.gauge_field(L, Nc, 71, kind="smooth", eps=0.5)
    src = fields.gaussian_sources(L, Nc, R, 72).reshape(R, 2, -1)
    cases = []
    for parity, dagger in ((1, 0), (0, 1)):
        inp = fields.to_mrhs(src[:, 1 - parity], Nc)
        xp = fields.to_mrhs(src[:, parity], Nc)
        cases.append((parity, dagger, inp, xp, oracle.dslash(inp, gauge, L, Nc, R, parity, dagger)))
    for prec in (H, B):
        d_g = api.DeviceArray.from_complex(gauge.reshape(-1), prec)
        for ts in (12, 5):
            print("case lds production", prec, ts, flush=True)
            api.dslash_tune("lds", 0, 0, ts)
            for parity, dagger, inp, xp, ref in cases:
                d_in = api.DeviceArray.from_complex(inp, prec)
                d_xp = api.DeviceArray.from_complex(xp, prec)
                d_out = api.DeviceArray(api.storage_bytes(inp.size, prec))
                api.dslash_mrhs(d_out, d_in, d_g, L, Nc, R, parity, dagger, precision=prec)
                assert _rel(d_out.complex(inp.size, prec), ref) < TOL[prec], (prec, ts, parity, dagger)
                api.dslash_mrhs(d_out, d_in, d_g, L, Nc, R, parity, dagger, precision=prec, xpay=d_xp, coef=-0.13)
                assert _rel(d_out.complex(inp.size, prec), xp - 0.13 * ref) < TOL[prec], (prec, ts, "xpay")


def test_lds_variant_falls_back_where_unsupported():
    """Shapes beyond the LDS kernel's limits (Nc = 8, FP64 and FP32
    storage) run the wave-per-site kernels instead of failing."""
    print("case lds fallback", flush=True)
    api.dslash_tune("lds", 0, 0, 4)
    for L, Nc, R, prec in (((4, 4, 4, 4), 8, 4, H), ((4, 4, 4, 4), 3, 4, S), ((4, 4, 4, 4), 3, 4, D)):
        gauge = fields.gauge_field(L, Nc, 5, kind="haar")
        inp = fields.to_mrhs(fields.gaussian_sources(L, Nc, R, 6).reshape(R, 2, -1)[:, 0], Nc)
        d_g = api.DeviceArray.from_complex(gauge.reshape(-1), prec)
        d_in = api.DeviceArray.from_complex(inp, prec)
        d_out = api.DeviceArray(api.storage_bytes(inp.size, prec))
        api.dslash_mrhs(d_out, d_in, d_g, L, Nc, R, 1, 0, precision=prec)
        assert _rel(d_out.complex(inp.size, prec), oracle.dslash(inp, gauge, L, Nc, R, 1, 0)) < TOL[prec]
