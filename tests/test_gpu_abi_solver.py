"""GPU parity of the reference C entry points (include/qcu.h) and of the
BiCGStab solve, against the CPU oracle on the same seeded inputs."""
import numpy as np
import pytest

from oracle import oracle
from qcu_mrhs_amd import api, fields

pytestmark = pytest.mark.gpu


def _case(*a):
    print("case", *a, flush=True)


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _setup(L, Nc, R, gauge, kappa, inner=2, user_prec=2):
    api.initGridSize(L, Nc, R, user_prec, inner)
    d_g = api.DeviceArray.from_complex(gauge.reshape(-1))
    api.loadQcuGauge(d_g, 2)
    api.getDslash(0, 1.0 / (2 * kappa) - 4.0)
    return d_g


def test_start_dslash_and_mat_qcu_reference_protocol():
    L, Nc, R, kappa = (8, 4, 4, 8), 3, 4, 0.125
    _case("start_dslash/mat_Qcu", L, Nc, R)
    gauge = fields.gauge_field(L, Nc, 42, kind="haar")
    _setup(L, Nc, R, gauge, kappa)
    src = fields.gaussian_sources(L, Nc, R, 43)
    half = src.shape[1] // 2
    for parity in (0, 1):
        for dagger in (0, 1):
            ins = [api.DeviceArray.from_complex(src[r].reshape(2, half)[1 - parity]) for r in range(R)]
            outs = [api.DeviceArray(half * 16) for _ in range(R)]
            for r in range(R):
                api.pushBackFermions(outs[r], ins[r])
            api.start_dslash(parity, dagger)
            got = np.stack([o.complex(half) for o in outs])
            ref = oracle.dslash(fields.to_mrhs(src.reshape(R, 2, half)[:, 1 - parity], Nc), gauge, L,
                                Nc, R, parity, dagger)
            assert _rel(fields.to_mrhs(got, Nc), ref) < 1e-13
    for dagger in (0, 1):
        ins = [api.DeviceArray.from_complex(src[r]) for r in range(R)]
        outs = [api.DeviceArray(src[r].nbytes) for r in range(R)]
        for r in range(R):
            api.pushBackFermions(outs[r], ins[r])
        api.mat_Qcu(dagger)
        got = fields.to_mrhs(np.stack([o.complex(src.shape[1]) for o in outs]), Nc)
        ref = oracle.mat(fields.to_mrhs(src, Nc), gauge, L, Nc, R, kappa, dagger)
        assert _rel(got, ref) < 1e-13
    api.finalizeQcu()


def _solve(L, Nc, R, gauge, b, kappa, tol, inner=2, max_iter=3000, delta=0.0):
    _setup(L, Nc, R, gauge, kappa, inner=inner)
    api.set_solver_options(inner, delta, False)
    bs = [api.DeviceArray.from_complex(b[r]) for r in range(R)]
    xs = [api.DeviceArray(b[r].nbytes) for r in range(R)]
    for r in range(R):
        api.pushBackFermions(xs[r], bs[r])
    info = api.qcuInvertEx(max_iter, tol)
    x = np.stack([xs[r].complex(b.shape[1]) for r in range(R)])
    api.finalizeQcu()
    return x, info


def test_invert_small_plain_fp64_matches_oracle():
    """8^3x8, SU(3) Haar, R=4, kappa=0.12, tol 1e-10 (the small plain case)."""
    L, Nc, R, kappa, tol = (8, 8, 8, 8), 3, 4, 0.12, 1e-10
    _case("invert small plain", L, R)
    gauge = fields.gauge_field(L, Nc, 1001, kind="haar")
    b = fields.gaussian_sources(L, Nc, R, 1002)
    x, info = _solve(L, Nc, R, gauge, b, kappa, tol)
    xo, ok, it, res = oracle.bicgstab(fields.to_mrhs(b, Nc), gauge, L, Nc, R, kappa, 3000, tol)
    assert ok and info["converged"]
    assert abs(info["iterations"] - it) <= 1, (info, it)
    assert _rel(fields.to_mrhs(x, Nc), xo) < 1e-8
    # independent FP64 true residual through the oracle operator
    r = fields.to_mrhs(b, Nc) - oracle.mat(fields.to_mrhs(x, Nc), gauge, L, Nc, R, kappa)
    true_rel = oracle.norm(r, R) / oracle.norm(fields.to_mrhs(b, Nc), R)
    assert true_rel.max() < 1.5e-10 and info["true_rel_residual"] < 1.5e-10, (true_rel, info)
    print(f"plain FP64: {info['iterations']} iterations (oracle {it}), true residual "
          f"{true_rel.max():.3e}, {info['seconds'] * 1e3:.1f} ms")


@pytest.mark.parametrize("inner", [1, 0, 4])
def test_invert_small_mixed_precision(inner):
    """Reliable-update mixed precision (FP32 / FP16 / BF16 inner) reaches a
    FP64 true residual <= 1e-10; error vs the FP64 oracle stated."""
    L, Nc, R, kappa, tol = (8, 8, 8, 8), 3, 4, 0.12, 1e-10
    _case("invert small mixed", inner)
    gauge = fields.gauge_field(L, Nc, 1001, kind="haar")
    b = fields.gaussian_sources(L, Nc, R, 1002)
    x, info = _solve(L, Nc, R, gauge, b, kappa, tol, inner=inner)
    xo, ok, it, _ = oracle.bicgstab(fields.to_mrhs(b, Nc), gauge, L, Nc, R, kappa, 3000, tol)
    r = fields.to_mrhs(b, Nc) - oracle.mat(fields.to_mrhs(x, Nc), gauge, L, Nc, R, kappa)
    true_rel = oracle.norm(r, R) / oracle.norm(fields.to_mrhs(b, Nc), R)
    err = [np.linalg.norm(x[k] - fields.from_mrhs(xo, R)[k]) / np.linalg.norm(x[k]) for k in range(R)]
    print(f"inner precision {inner}: {info['iterations']} iterations, {info['reliable_updates']} reliable "
          f"updates, true residual {true_rel.max():.3e}, per-RHS error vs FP64 oracle {np.max(err):.2e}")
    assert info["converged"] and true_rel.max() < 1.5e-10
    assert max(err) < 1e-8


@pytest.mark.parametrize("inner", [0, 4])
def test_invert_mixed_with_lds_staged_dslash(inner):
    """The FP16 / BF16 mixed solve with the LDS-staged dslash selected for
    every hop (variant 4; the Schur operator then forms omega's inner products
    in a separate pass): FP64 true residual <= 1e-10 through the oracle
    operator and the solution within 1e-8 of the oracle's."""
    L, Nc, R, kappa, tol = (8, 8, 8, 8), 3, 4, 0.12, 1e-10
    _case("invert mixed lds", inner)
    gauge = fields.gauge_field(L, Nc, 1001, kind="haar")
    b = fields.gaussian_sources(L, Nc, R, 1002)
    api.dslash_tune("lds", 0, 0, 3)
    try:
        x, info = _solve(L, Nc, R, gauge, b, kappa, tol, inner=inner)
    finally:
        api.dslash_tune("auto")
    xo, ok, it, _ = oracle.bicgstab(fields.to_mrhs(b, Nc), gauge, L, Nc, R, kappa, 3000, tol)
    r = fields.to_mrhs(b, Nc) - oracle.mat(fields.to_mrhs(x, Nc), gauge, L, Nc, R, kappa)
    true_rel = oracle.norm(r, R) / oracle.norm(fields.to_mrhs(b, Nc), R)
    err = [np.linalg.norm(x[k] - fields.from_mrhs(xo, R)[k]) / np.linalg.norm(x[k]) for k in range(R)]
    print(f"inner precision {inner} (LDS dslash): {info['iterations']} iterations, {info['reliable_updates']} "
          f"reliable updates, true residual {true_rel.max():.3e}, per-RHS error vs FP64 oracle {np.max(err):.2e}")
    assert info["converged"] and true_rel.max() < 1.5e-10
    assert max(err) < 1e-8


@pytest.mark.parametrize("Nc,R,inner", [(2, 17, 0), (4, 17, 4), (3, 33, 0)])
def test_invert_mixed_precision_rhs_chunks(Nc, R, inner):
    """Matrix-core inner operator over more than one 16-RHS chunk (R = 17, 33:
    the last chunk partly empty) and colour counts 2 / 4: FP64 true residual
    <= 1e-10 through the oracle operator, solution within 1e-8 of the oracle."""
    L, kappa, tol = (8, 8, 8, 8), 0.12, 1e-10
    _case("invert mixed chunks", Nc, R, inner)
    gauge = fields.gauge_field(L, Nc, 1100 + Nc, kind="haar")
    b = fields.gaussian_sources(L, Nc, R, 1200 + R)
    x, info = _solve(L, Nc, R, gauge, b, kappa, tol, inner=inner)
    bm, xm = fields.to_mrhs(b, Nc), fields.to_mrhs(x, Nc)
    r = bm - oracle.mat(xm, gauge, L, Nc, R, kappa)
    true_rel = oracle.norm(r, R) / oracle.norm(bm, R)
    xo, ok, it, _ = oracle.bicgstab(bm, gauge, L, Nc, R, kappa, 3000, tol)
    print(f"Nc {Nc} R {R} inner {inner}: {info['iterations']} iterations (oracle FP64 {it}), "
          f"{info['reliable_updates']} reliable updates, true residual {true_rel.max():.3e}, "
          f"error vs oracle {_rel(xm, xo):.2e}")
    assert info["converged"] and ok
    assert true_rel.max() < 1.5e-10 and info["true_rel_residual"] < 1.5e-10
    assert _rel(xm, xo) < 1e-8


def test_invert_production_shape_mixed_half():
    """16^3x32, R=12, exp(i 0.3 H), kappa=0.135 with FP16 inner vectors and
    the matrix-core operator: FP64 true residual <= 1e-10 through the oracle."""
    L, Nc, R, kappa, tol = (16, 16, 16, 32), 3, 12, 0.135, 1e-10
    _case("invert production mixed half", L, R)
    gauge = fields.gauge_field(L, Nc, 2024, kind="smooth", eps=0.3)
    b = fields.gaussian_sources(L, Nc, R, 2026)
    x, info = _solve(L, Nc, R, gauge, b, kappa, tol, inner=0, max_iter=20000)
    bm = fields.to_mrhs(b, Nc)
    r = bm - oracle.mat(fields.to_mrhs(x, Nc), gauge, L, Nc, R, kappa)
    true_rel = oracle.norm(r, R) / oracle.norm(bm, R)
    print(f"production mixed FP16: {info['iterations']} iterations, {info['reliable_updates']} reliable "
          f"updates, true residual {true_rel.max():.3e} (library {info['true_rel_residual']:.3e}), "
          f"{info['seconds']:.3f} s")
    assert info["converged"] and true_rel.max() < 1.5e-10 and info["true_rel_residual"] < 1.5e-10


def test_invert_production_shape_plain_fp64():
    """16^3x32, R=12, exp(i 0.3 H), kappa=0.135: converges to a FP64 true
    residual <= 1e-10 (checked independently through the oracle operator)."""
    L, Nc, R, kappa, tol = (16, 16, 16, 32), 3, 12, 0.135, 1e-10
    _case("invert production plain", L, R)
    gauge = fields.gauge_field(L, Nc, 2024, kind="smooth", eps=0.3)
    b = fields.gaussian_sources(L, Nc, R, 2026)
    x, info = _solve(L, Nc, R, gauge, b, kappa, tol, max_iter=20000)
    bm = fields.to_mrhs(b, Nc)
    r = bm - oracle.mat(fields.to_mrhs(x, Nc), gauge, L, Nc, R, kappa)
    true_rel = oracle.norm(r, R) / oracle.norm(bm, R)
    print(f"production plain FP64: {info['iterations']} iterations, true residual {true_rel.max():.3e}, "
          f"{info['seconds']:.3f} s")
    assert info["converged"] and true_rel.max() < 1.5e-10


def test_gauge_eo_precondition_round_trip():
    L, Nc, R = (4, 6, 4, 2), 3, 1
    api.initGridSize(L, Nc, R, 2, 2)
    g = np.random.default_rng(0)
    V = int(np.prod(L))
    lex = g.standard_normal((4, V, Nc, Nc)) + 1j * g.standard_normal((4, V, Nc, Nc))
    d_lex = api.DeviceArray.from_complex(lex.reshape(-1))
    d_eo = api.DeviceArray(lex.nbytes)
    api.gauge_eo_precondition(d_eo, d_lex, 2)
    np.testing.assert_array_equal(d_eo.complex(lex.size), oracle.gauge_eo(lex, L, Nc))
    d_back = api.DeviceArray(lex.nbytes)
    api.gauge_reverse_eo_precondition(d_back, d_eo, 2)
    np.testing.assert_array_equal(d_back.complex(lex.size), lex.reshape(-1))
    api.finalizeQcu()


def test_read_gauge_from_file_matches_oracle_eo(tmp_path):
    """read_gauge_from_file (src/qcu_interface.cu:364-413 of the reference reads
    through qcu_io): a raw little-endian complex128 file in the lexicographic
    layout [mu][t][z][y][x][Nc][Nc], written here from a Haar field, lands on
    the device in the even/odd layout bit for bit as the oracle's gauge_eo
    maps it; the loaded field then drives a dslash that agrees with the
    oracle on the original even/odd field; a truncated file is an error."""
    L, Nc, R = (4, 6, 4, 8), 3, 2
    _case("read_gauge_from_file", L)
    api.initGridSize(L, Nc, R, 2, 2)
    eo = fields.gauge_field(L, Nc, 77, kind="haar")
    lex = fields.gauge_eo_to_lex(eo, L, Nc)
    np.testing.assert_array_equal(oracle.gauge_eo(lex, L, Nc), eo.reshape(-1))
    path = tmp_path / "gauge.bin"
    np.ascontiguousarray(lex, dtype="<c16").tofile(path)
    d_eo = api.DeviceArray(lex.nbytes)
    api.read_gauge_from_file(d_eo, path)
    np.testing.assert_array_equal(d_eo.complex(lex.size), oracle.gauge_eo(lex, L, Nc))
    src = fields.gaussian_sources(L, Nc, R, 78).reshape(R, 2, -1)
    inp = fields.to_mrhs(src[:, 0], Nc)
    d_in = api.DeviceArray.from_complex(inp)
    d_out = api.DeviceArray(inp.nbytes)
    api.dslash_mrhs(d_out, d_in, d_eo, L, Nc, R, 1, 0)
    assert _rel(d_out.complex(inp.size), oracle.dslash(inp, eo, L, Nc, R, 1)) < 1e-13
    short = tmp_path / "short.bin"
    np.ascontiguousarray(lex.reshape(-1)[:-1], dtype="<c16").tofile(short)
    with pytest.raises(api.QcuError):
        api.read_gauge_from_file(d_eo, short)
    api.finalizeQcu()


def test_queue_misuse_is_reported():
    L, Nc, R = (4, 4, 4, 4), 3, 2
    api.initGridSize(L, Nc, R, 2, 2)
    api.getDslash(0, 0.0)
    d = api.DeviceArray(16)
    api.pushBackFermions(d, d)
    with pytest.raises(api.QcuError):
        api.start_dslash(0, 0)            # queue holds 1 of 2 pairs
    api.pushBackFermions(d, d)
    with pytest.raises(api.QcuError):
        api.pushBackFermions(d, d)        # queue is full
    api.finalizeQcu()
