"""GPU parity at the production workloads, against the FP64 CPU oracle on the
same seeded inputs:

* 16^3 x 32, SU(3), 12 RHS, links exp(i 0.3 H), kappa 0.135: FP64 and FP32
  dslash, plain FP64 and FP32-inner solves;
* 24^3 x 48, SU(3), 16 RHS, links exp(i 0.5 H), kappa 0.1355: plain FP64 and
  FP16 / BF16-inner reliable-update solves, each RHS's true residual
  recomputed through the oracle operator and its error against the oracle's
  FP64 BiCGStab solution stated;
* 16^4, SU(4) and SU(8) Haar links, kappa 0.12, groups of 3, 8, 17 and 32
  sources in one batched call: every group's true residual through the oracle
  operator.

Stopping rule of the reference (src/solver/bicgstab_solve.cu:322-340): every
RHS at ||r|| / ||b|| <= tol; here the residual is the FP64 true residual of
the full system."""
import numpy as np
import pytest

from oracle import oracle
from qcu_mrhs_amd import api, fields

pytestmark = pytest.mark.gpu

D, S, H, B = 2, 1, 0, 4
TOL = 1e-10
# both solutions meet ||b - M x|| <= 1e-10 ||b||; their distance is bounded by
# cond(M) times that, and cond(M) stays below 1e3 at these kappa and links
ERR_VS_ORACLE = 1e-7


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


def _solve(d_b, n, inner, max_iter=20000, delta=0.0):
    """One qcuInvertEx on the loaded gauge field; returns (x [R][n], info)."""
    R = len(d_b)
    api.set_solver_options(inner, delta, False)
    xs = [api.DeviceArray(n * 16) for _ in range(R)]
    for r in range(R):
        api.pushBackFermions(xs[r], d_b[r])
    info = api.qcuInvertEx(max_iter, TOL)
    return np.stack([x.complex(n) for x in xs]), info


def _true_residual(x, bm, gauge, L, Nc, R, kappa):
    """Per-RHS ||b - M x|| / ||b|| with the oracle's FP64 operator."""
    r = bm - oracle.mat(fields.to_mrhs(x, Nc), gauge, L, Nc, R, kappa)
    return oracle.norm(r, R) / oracle.norm(bm, R)


def _per_rhs_error(x, xo_mrhs, R):
    xo = fields.from_mrhs(xo_mrhs, R)
    return np.array([np.linalg.norm(x[k] - xo[k]) / np.linalg.norm(xo[k]) for k in range(R)])


# ---- 16^3 x 32, 12 RHS ----------------------------------------------------

C2 = dict(L=(16, 16, 16, 32), Nc=3, R=12, kappa=0.135, eps=0.3, seeds=(2024, 2026))


@pytest.fixture(scope="module")
def c2_fields():
    g = fields.gauge_field(C2["L"], C2["Nc"], C2["seeds"][0], kind="smooth", eps=C2["eps"])
    b = fields.gaussian_sources(C2["L"], C2["Nc"], C2["R"], C2["seeds"][1])
    return g, b


def test_c2_dslash_fp64_and_fp32(c2_fields):
    """One-parity dslash, both parities and dagger flags: FP64 within 1e-13
    and FP32 storage within 2e-7 of the FP64 oracle (normwise relative)."""
    gauge, b = c2_fields
    L, Nc, R = C2["L"], C2["Nc"], C2["R"]
    src = b.reshape(R, 2, -1)
    for prec, tol in ((D, 1e-13), (S, 2e-7)):
        d_g = api.DeviceArray.from_complex(gauge.reshape(-1), prec)
        for parity, dagger in ((1, 0), (0, 1)):
            _case("c2 dslash", prec, parity, dagger)
            inp = fields.to_mrhs(src[:, 1 - parity], Nc)
            d_in = api.DeviceArray.from_complex(inp, prec)
            d_out = api.DeviceArray(api.storage_bytes(inp.size, prec))
            api.dslash_mrhs(d_out, d_in, d_g, L, Nc, R, parity, dagger, precision=prec)
            err = _rel(d_out.complex(inp.size, prec), oracle.dslash(inp, gauge, L, Nc, R, parity, dagger))
            print(f"precision {prec} parity {parity} dagger {dagger}: relative error {err:.2e}")
            assert err < tol


@pytest.mark.parametrize("inner", [D, S])
def test_c2_solve_fp64_and_fp32_inner(c2_fields, inner):
    """Plain FP64 and FP32-inner mixed solves to a FP64 true residual of 1e-10
    per RHS (oracle operator); plain FP64 iterates as the oracle's BiCGStab."""
    gauge, b = c2_fields
    L, Nc, R, kappa = C2["L"], C2["Nc"], C2["R"], C2["kappa"]
    _case("c2 solve", inner)
    _setup(L, Nc, R, gauge, kappa, inner=inner)
    d_b = [api.DeviceArray.from_complex(b[r]) for r in range(R)]
    x, info = _solve(d_b, b.shape[1], inner)
    api.finalizeQcu()
    bm = fields.to_mrhs(b, Nc)
    tr = _true_residual(x, bm, gauge, L, Nc, R, kappa)
    print(f"inner {inner}: {info['iterations']} iterations, {info['reliable_updates']} reliable updates, "
          f"true residual per RHS max {tr.max():.2e} (library {info['true_rel_residual']:.2e}), "
          f"{info['seconds'] * 1e3:.1f} ms")
    assert info["converged"]
    assert tr.max() <= TOL * 1.05 and info["true_rel_residual"] <= TOL * 1.05


# ---- 24^3 x 48, 16 RHS ------------------------------------------------------

C3 = dict(L=(24, 24, 24, 48), Nc=3, R=16, kappa=0.1355, eps=0.5, seeds=(3001, 3002))


@pytest.fixture(scope="module")
def c3_system():
    L, Nc, R, kappa = C3["L"], C3["Nc"], C3["R"], C3["kappa"]
    gauge = fields.gauge_field(L, Nc, C3["seeds"][0], kind="smooth", eps=C3["eps"])
    b = fields.gaussian_sources(L, Nc, R, C3["seeds"][1])
    bm = fields.to_mrhs(b, Nc)
    xo, ok, it, _ = oracle.bicgstab(bm, gauge, L, Nc, R, kappa, 2000, TOL)
    assert ok
    print(f"oracle FP64 BiCGStab: {it} iterations", flush=True)
    return gauge, b, bm, xo, it


@pytest.mark.parametrize("inner", [D, H, B])
def test_c3_solve_per_rhs_against_oracle(c3_system, inner):
    """Plain FP64 and FP16 / BF16-inner (matrix-core dslash, FP32
    accumulation, FP64 reliable updates) solves of the 24^3 x 48 system: each
    RHS's true residual through the oracle operator <= 1e-10 and its error
    against the oracle's FP64 solution <= 1e-7; plain FP64 takes the oracle's
    iteration count within one."""
    gauge, b, bm, xo, it_oracle = c3_system
    L, Nc, R, kappa = C3["L"], C3["Nc"], C3["R"], C3["kappa"]
    _case("c3 solve", inner)
    _setup(L, Nc, R, gauge, kappa, inner=inner)
    d_b = [api.DeviceArray.from_complex(b[r]) for r in range(R)]
    x, info = _solve(d_b, b.shape[1], inner)
    api.finalizeQcu()
    del d_b
    tr = _true_residual(x, bm, gauge, L, Nc, R, kappa)
    err = _per_rhs_error(x, xo, R)
    print(f"inner {inner}: {info['iterations']} iterations (oracle FP64 {it_oracle}), "
          f"{info['reliable_updates']} reliable updates, {info['seconds'] * 1e3:.1f} ms")
    print("  true residual per RHS:", " ".join(f"{v:.2e}" for v in tr))
    print("  error vs FP64 oracle per RHS:", " ".join(f"{v:.2e}" for v in err))
    assert info["converged"]
    assert tr.max() <= TOL * 1.05 and info["true_rel_residual"] <= TOL * 1.05
    assert err.max() <= ERR_VS_ORACLE
    if inner == D:
        assert abs(info["iterations"] - it_oracle) <= 1


# ---- 16^4 batched groups ----------------------------------------------------

GROUPS = (3, 8, 17, 32)


@pytest.mark.parametrize("Nc", [4, 8])
def test_c4_groups_true_residual_per_group(Nc):
    """SU(4) / SU(8) on 16^4, Haar links, kappa 0.12, groups of 3, 8, 17 and
    32 sources in one qcuInvertGroups call (FP32 inner, FP64 outer): every
    RHS of every group at a FP64 true residual <= 1e-10 through the oracle
    operator; the 3-source group's solution against the oracle's BiCGStab."""
    L, kappa = (16, 16, 16, 16), 0.12
    _case("c4 groups", Nc)
    gauge = fields.gauge_field(L, Nc, 5000 + Nc, kind="haar")
    _setup(L, Nc, 4, gauge, kappa, inner=S)
    api.set_solver_options(S, 0.0, False)
    srcs, groups = [], []
    for k, R in enumerate(GROUPS):
        b = fields.gaussian_sources(L, Nc, R, 5100 + 10 * Nc + k)
        srcs.append(b)
        groups.append(([api.DeviceArray(b[r].nbytes) for r in range(R)],
                       [api.DeviceArray.from_complex(b[r]) for r in range(R)]))
    infos = api.invert_groups(groups, 3000, TOL)
    print(f"SU({Nc}) group dslash kernel:", {2: "VALU wave per site", 3: "FP32 matrix cores", 5: "FP32 matrix cores, packed RHS"}.get(
        api.lib().qcuGroupsDslashVariant(), "?"), "; FP64 hops:", {2: "VALU wave per site", 3: "FP64 matrix cores", 5: "FP64 matrix cores, packed RHS"}.get(
        api.lib().qcuGroupsDslashVariant64(), "?"), flush=True)
    xs = [np.stack([x.complex(srcs[k].shape[1]) for x in groups[k][0]]) for k in range(len(GROUPS))]
    api.finalizeQcu()
    del groups
    for k, R in enumerate(GROUPS):
        bm = fields.to_mrhs(srcs[k], Nc)
        tr = _true_residual(xs[k], bm, gauge, L, Nc, R, kappa)
        print(f"SU({Nc}) group of {R}: {infos[k]['iterations']} inner iterations, {infos[k]['outer_steps']} outer "
              f"steps, true residual max {tr.max():.2e} (library {infos[k]['true_rel_residual']:.2e}), "
              f"{infos[k]['seconds'] * 1e3:.1f} ms")
        assert infos[k]["converged"]
        assert tr.max() <= TOL * 1.05 and infos[k]["true_rel_residual"] <= TOL * 1.05
        if R == 3:
            xo, ok, _, _ = oracle.bicgstab(bm, gauge, L, Nc, R, kappa, 3000, TOL)
            err = _per_rhs_error(xs[k], xo, R)
            print("  error vs FP64 oracle per RHS:", " ".join(f"{v:.2e}" for v in err))
            assert ok and err.max() <= ERR_VS_ORACLE
