"""GPU parity of the Wilson-clover operator M = A - kappa D against the FP64
CPU oracle on the same seeded inputs: the clover blocks and their inverse
(qcuCloverBuild), the block application in every storage precision
(qcuCloverApply), mat_Qcu with getDslash(clover), and qcuInvertEx on the
even/odd clover Schur system (plain FP64 and mixed FP32 / FP16 / BF16 inner)
with each RHS's true residual recomputed through the oracle operator and the
solution compared with a direct sparse solve of A - kappa D.

The reference declares the clover type and rejects it
(include/qcu_helper_macro.h:7-11, src/qcu_interface.cu:97-105); the oracle's
clover term is pinned in tests/test_oracle_clover.py."""
import numpy as np
import pytest
import scipy.sparse.linalg as sla

from oracle import dense, oracle
from qcu_mrhs_amd import api, fields

pytestmark = pytest.mark.gpu

D, S, H, B = 2, 1, 0, 4


def _case(*a):
    print("case", *a, flush=True)


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("Nc,kind", [(3, "smooth"), (2, "haar"), (4, "haar"), (8, "smooth")])
def test_build_and_inverse_match_oracle(Nc, kind):
    L, csw, kappa = (4, 4, 4, 8), 1.1, 0.125
    _case("clover build", Nc, kind)
    gauge = fields.gauge_field(L, Nc, 600 + Nc, kind=kind, eps=0.5)
    ref = oracle.clover(gauge, L, Nc, csw, kappa)
    ref_inv = oracle.clover_invert(ref)
    d_g = api.DeviceArray.from_complex(gauge.reshape(-1))
    d_c = api.DeviceArray(ref.size * 16)
    d_i = api.DeviceArray(ref.size * 16)
    api.clover_build(d_c, d_i, d_g, L, Nc, csw, kappa)
    got = d_c.complex(ref.size).reshape(ref.shape)
    got_inv = d_i.complex(ref.size).reshape(ref.shape)
    print(f"Nc {Nc}: term error {np.abs(got - ref).max():.2e}, inverse error {np.abs(got_inv - ref_inv).max():.2e}")
    assert np.abs(got - ref).max() < 1e-14
    assert np.abs(got_inv - ref_inv).max() < 1e-12 * np.abs(ref_inv).max()


@pytest.mark.parametrize("prec,tol", [(D, 1e-14), (S, 1e-6), (H, 2e-3), (B, 1.5e-2)])
@pytest.mark.parametrize("Nc,R", [(3, 16), (3, 5), (8, 3), (2, 1)])
def test_apply_matches_oracle(prec, tol, Nc, R):
    """out = A in + coef y on one parity, in place (out = in) as the solver
    applies A_ee^{-1}, and with out = y as the Schur operator applies A_oo;
    FP64 spinors read FP64 blocks, the other storage precisions FP32 blocks."""
    L, csw, kappa = (4, 4, 2, 4), 1.0, 0.13
    _case("clover apply", prec, Nc, R)
    gauge = fields.gauge_field(L, Nc, 700 + Nc, kind="smooth", eps=0.6)
    c = oracle.clover(gauge, L, Nc, csw, kappa)
    cprec = D if prec == D else S
    src = fields.gaussian_sources(L, Nc, R, 710 + R)
    y = fields.gaussian_sources(L, Nc, R, 720 + R)
    for parity in (0, 1):
        blocks = np.ascontiguousarray(c[parity]).reshape(-1)
        d_blk = api.DeviceArray.from_complex(blocks, cprec)
        inp = fields.to_mrhs(src.reshape(R, 2, -1)[:, parity], Nc)
        yv = fields.to_mrhs(y.reshape(R, 2, -1)[:, parity], Nc)
        # storage rounding of the inputs is part of what is compared
        inp_s = api.from_storage(api.to_storage(inp, prec), prec)
        y_s = api.from_storage(api.to_storage(yv, prec), prec)
        ref = oracle.clover_apply(c[parity], inp_s, Nc, R) - 0.7 * y_s
        d_in = api.DeviceArray.from_complex(inp, prec)
        d_y = api.DeviceArray.from_complex(yv, prec)
        api.clover_apply(d_y, d_in, d_blk, L, Nc, R, prec, cprec, xpay=d_y, coef=-0.7)
        err = _rel(d_y.complex(inp.size, prec), ref)
        api.clover_apply(d_in, d_in, d_blk, L, Nc, R, prec, cprec)
        err2 = _rel(d_in.complex(inp.size, prec), oracle.clover_apply(c[parity], inp_s, Nc, R))
        print(f"prec {prec} Nc {Nc} R {R} parity {parity}: xpay error {err:.2e}, in-place error {err2:.2e}")
        assert err < tol and err2 < tol


@pytest.mark.parametrize("prec,tol", [(D, 1e-13), (S, 2e-6)])
def test_mat_qcu_clover_matches_oracle(prec, tol):
    """mat_Qcu with getDslash(clover): out = (A - kappa D) in on full-lattice
    vectors, both dagger flags (A is Hermitian, so the dagger reaches D only)."""
    L, Nc, R, mass, csw = (8, 4, 4, 8), 3, 4, -0.2, 1.3
    kappa = 1.0 / (2 * (4 + mass))
    _case("clover mat_Qcu", prec)
    gauge = fields.gauge_field(L, Nc, 800, kind="smooth", eps=0.5)
    c = oracle.clover(gauge, L, Nc, csw, kappa)
    b = fields.gaussian_sources(L, Nc, R, 801)
    api.initGridSize(L, Nc, R, D, prec)
    d_g = api.DeviceArray.from_complex(gauge.reshape(-1))
    api.loadQcuGauge(d_g, D)
    api.getDslash(api.DSLASH_CLOVER, mass)
    api.set_clover_coefficient(csw)
    try:
        for dagger in (0, 1):
            ins = [api.DeviceArray.from_complex(b[r]) for r in range(R)]
            outs = [api.DeviceArray(b[r].nbytes) for r in range(R)]
            for r in range(R):
                api.pushBackFermions(outs[r], ins[r])
            api.mat_Qcu(dagger)
            got = fields.to_mrhs(np.stack([o.complex(b.shape[1]) for o in outs]), Nc)
            ref = oracle.clover_mat(fields.to_mrhs(b, Nc), gauge, c, L, Nc, R, kappa, dagger)
            err = _rel(got, ref)
            print(f"prec {prec} dagger {dagger}: relative error {err:.2e}")
            assert err < tol
    finally:
        api.finalizeQcu()


def _clover_solve(L, Nc, R, gauge, mass, csw, b, inner, tol, max_iter=5000):
    api.initGridSize(L, Nc, R, D, inner)
    d_g = api.DeviceArray.from_complex(gauge.reshape(-1))
    api.loadQcuGauge(d_g, D)
    api.getDslash(api.DSLASH_CLOVER, mass)
    api.set_clover_coefficient(csw)
    api.set_solver_options(inner, 0.0, False)
    n = b.shape[1]
    d_b = [api.DeviceArray.from_complex(b[r]) for r in range(R)]
    xs = [api.DeviceArray(n * 16) for _ in range(R)]
    for r in range(R):
        api.pushBackFermions(xs[r], d_b[r])
    info = api.qcuInvertEx(max_iter, tol)
    x = np.stack([v.complex(n) for v in xs])
    api.finalizeQcu()
    return x, info


@pytest.mark.parametrize("inner", [D, S, H, B])
def test_clover_solve_matches_direct_solve(inner):
    """8 x 4 x 4 x 8, SU(3) links exp(i 0.5 H), kappa 0.13, c_sw 1.2, 3 RHS:
    every RHS at a FP64 true residual <= 1e-10 through the oracle's clover
    operator and within 1e-8 of scipy's sparse direct solve of A - kappa D."""
    L, Nc, R, csw, kappa, tol = (8, 4, 4, 8), 3, 3, 1.2, 0.13, 1e-10
    mass = 1.0 / (2 * kappa) - 4.0
    _case("clover solve", inner)
    gauge = fields.gauge_field(L, Nc, 900, kind="smooth", eps=0.5)
    b = fields.gaussian_sources(L, Nc, R, 901)
    x, info = _clover_solve(L, Nc, R, gauge, mass, csw, b, inner, tol)
    c = oracle.clover(gauge, L, Nc, csw, kappa)
    bm = fields.to_mrhs(b, Nc)
    r = bm - oracle.clover_mat(fields.to_mrhs(x, Nc), gauge, c, L, Nc, R, kappa)
    tr = oracle.norm(r, R) / oracle.norm(bm, R)
    M = (dense.clover_matrix(gauge, L, Nc, csw, kappa) - kappa * dense.hopping_matrix(gauge, L, Nc)).tocsc()
    lu = sla.splu(M)
    err = np.array([_rel(x[k], lu.solve(b[k])) for k in range(R)])
    print(f"inner {inner}: {info['iterations']} iterations, {info['reliable_updates']} reliable updates, "
          f"true residual per RHS {' '.join(f'{v:.2e}' for v in tr)} (library {info['true_rel_residual']:.2e}), "
          f"error vs direct solve {' '.join(f'{v:.2e}' for v in err)}")
    assert info["converged"]
    assert tr.max() <= tol * 1.05 and info["true_rel_residual"] <= tol * 1.05
    assert err.max() < 1e-8


@pytest.mark.parametrize("inner", [D, H])
def test_clover_solve_16c3x32(inner):
    """16^3 x 32, SU(3) links exp(i 0.3 H), kappa 0.135, c_sw 1.0, 12 RHS:
    plain FP64 and FP16-inner solves to a FP64 true residual of 1e-10 per RHS
    through the oracle's clover operator."""
    L, Nc, R, csw, kappa, tol = (16, 16, 16, 32), 3, 12, 1.0, 0.135, 1e-10
    mass = 1.0 / (2 * kappa) - 4.0
    _case("clover solve 16^3x32", inner)
    gauge = fields.gauge_field(L, Nc, 2024, kind="smooth", eps=0.3)
    b = fields.gaussian_sources(L, Nc, R, 2026)
    x, info = _clover_solve(L, Nc, R, gauge, mass, csw, b, inner, tol)
    c = oracle.clover(gauge, L, Nc, csw, kappa)
    bm = fields.to_mrhs(b, Nc)
    r = bm - oracle.clover_mat(fields.to_mrhs(x, Nc), gauge, c, L, Nc, R, kappa)
    tr = oracle.norm(r, R) / oracle.norm(bm, R)
    print(f"inner {inner}: {info['iterations']} iterations, {info['reliable_updates']} reliable updates, "
          f"{info['seconds'] * 1e3:.1f} ms, true residual max {tr.max():.2e} "
          f"(library {info['true_rel_residual']:.2e})")
    assert info["converged"]
    assert tr.max() <= tol * 1.05 and info["true_rel_residual"] <= tol * 1.05


def test_clover_rejected_where_only_wilson_is_modelled():
    """The batched groups and the learned preconditioner model the Wilson
    operator: with the clover type they return kQcuErrorUnsupported."""
    L, Nc = (4, 4, 4, 4), 3
    _case("clover rejects")
    gauge = fields.gauge_field(L, Nc, 5, kind="smooth", eps=0.3)
    api.initGridSize(L, Nc, 2, D, S)
    d_g = api.DeviceArray.from_complex(gauge.reshape(-1))
    api.loadQcuGauge(d_g, D)
    api.getDslash(api.DSLASH_CLOVER, -0.1)
    b = fields.gaussian_sources(L, Nc, 2, 6)
    groups = [([api.DeviceArray(b[r].nbytes) for r in range(2)], [api.DeviceArray.from_complex(b[r]) for r in range(2)])]
    with pytest.raises(api.QcuError) as e:
        api.invert_groups(groups, 100, 1e-8)
    assert e.value.code == 4
    api.finalizeQcu()
