"""GPU parity of the batched uneven-group solve (qcuInvertGroups): every group
of one batch against the FP64 CPU oracle on the same seeded inputs."""
import numpy as np
import pytest

from oracle import oracle
from qcu_mrhs_amd import api, fields

pytestmark = pytest.mark.gpu

SIZES = (3, 8, 17, 32)


def _case(*a):
    print("case", *a, flush=True)


def _run(L, Nc, kappa, sizes, seed, inner=1, tol=1e-10, user_prec=2):
    gauge = fields.gauge_field(L, Nc, seed, kind="haar")
    api.initGridSize(L, Nc, 4, user_prec, inner)
    d_g = api.DeviceArray.from_complex(gauge.reshape(-1))
    api.loadQcuGauge(d_g, 2)
    api.getDslash(0, 1.0 / (2 * kappa) - 4.0)
    api.set_solver_options(inner, 0.1, False)
    srcs, groups = [], []
    for k, R in enumerate(sizes):
        b = fields.gaussian_sources(L, Nc, R, seed + 100 + k)
        bs = [api.DeviceArray.from_complex(b[r], user_prec) for r in range(R)]
        xs = [api.DeviceArray(api.storage_bytes(b.shape[1], user_prec)) for _ in range(R)]
        srcs.append(b)
        groups.append((xs, bs))
    infos = api.invert_groups(groups, 3000, tol)
    print("group dslash kernel:", {2: "VALU wave per site", 3: "FP32 matrix cores", 5: "FP32 matrix cores, packed RHS"}.get(
        api.lib().qcuGroupsDslashVariant(), "?"), "; FP64 hops:", {2: "VALU wave per site", 3: "FP64 matrix cores", 5: "FP64 matrix cores, packed RHS"}.get(
        api.lib().qcuGroupsDslashVariant64(), "?"), flush=True)
    xs = [np.stack([x.complex(srcs[k].shape[1], user_prec) for x in groups[k][0]]) for k in range(len(sizes))]
    api.finalizeQcu()
    return gauge, srcs, xs, infos


@pytest.mark.parametrize("Nc", [4, 8, 3])
def test_groups_match_oracle(Nc):
    """8^4 lattice, Haar SU(N) links, kappa 0.12, groups of 3, 8, 17 and 32
    sources solved in one batch (FP32 inner, FP64 outer) to 1e-10: each group's
    FP64 true residual is recomputed through the oracle operator and its
    solution compared with the oracle's BiCGStab on the same sources."""
    L, kappa, tol = (8, 8, 8, 8), 0.12, 1e-10
    _case("groups", L, Nc, SIZES)
    gauge, srcs, xs, infos = _run(L, Nc, kappa, SIZES, 4100 + Nc)
    for k, R in enumerate(SIZES):
        bm = fields.to_mrhs(srcs[k], Nc)
        xm = fields.to_mrhs(xs[k], Nc)
        r = bm - oracle.mat(xm, gauge, L, Nc, R, kappa)
        true_rel = (oracle.norm(r, R) / oracle.norm(bm, R)).max()
        xo, ok, it, _ = oracle.bicgstab(bm, gauge, L, Nc, R, kappa, 3000, tol)
        err = np.linalg.norm(xm - xo) / np.linalg.norm(xo)
        print(f"Nc {Nc} group R={R}: {infos[k]['iterations']} inner iterations, {infos[k]['outer_steps']} outer "
              f"steps, sites/wave {infos[k]['sites_per_wave']}, true residual {true_rel:.2e} "
              f"(library {infos[k]['true_rel_residual']:.2e}), error vs oracle {err:.2e}, "
              f"{infos[k]['seconds'] * 1e3:.1f} ms")
        assert infos[k]["converged"] and ok
        assert true_rel < 1.5e-10 and infos[k]["true_rel_residual"] < 1.5e-10
        assert err < 1e-8


@pytest.mark.parametrize("Nc,inner,fp64", [(8, 5, 5), (4, 5, 3), (3, 3, 5), (8, 2, 3), (4, 3, 2)])
def test_groups_each_kernel_matches_oracle(Nc, inner, fp64):
    """The batched groups solve with its dslash kernels chosen explicitly
    (qcuGroupsDslashTune): inner hops and FP64 hops on the VALU wave kernel (2),
    the matrix cores per group (3) or the matrix cores over the groups' RHS
    packed into shared 16-RHS chunks (5; groups of 3, 8, 17, 32 share 4 chunks),
    every group to a FP64 true residual <= 1e-10 through the oracle operator."""
    L, kappa, tol = (8, 8, 8, 8), 0.12, 1e-10
    _case("groups forced", Nc, inner, fp64)
    api.check(api.lib().qcuGroupsDslashTune(inner, fp64))
    try:
        gauge, srcs, xs, infos = _run(L, Nc, kappa, SIZES, 4300 + Nc)
    finally:
        api.check(api.lib().qcuGroupsDslashTune(0, 0))
    for k, R in enumerate(SIZES):
        bm = fields.to_mrhs(srcs[k], Nc)
        r = bm - oracle.mat(fields.to_mrhs(xs[k], Nc), gauge, L, Nc, R, kappa)
        true_rel = (oracle.norm(r, R) / oracle.norm(bm, R)).max()
        print(f"Nc {Nc} inner {inner} fp64 {fp64} group R={R}: {infos[k]['iterations']} inner iterations, "
              f"{infos[k]['outer_steps']} outer steps, true residual {true_rel:.2e}")
        assert infos[k]["converged"] and true_rel < 1.5e-10


def test_groups_single_group_equals_plain_solver():
    """A batch of one group gives the same solution as qcuInvertEx on the same
    sources (both to 1e-10 FP64 residual)."""
    L, Nc, kappa = (8, 8, 8, 8), 3, 0.12
    _case("groups single")
    gauge, srcs, xs, infos = _run(L, Nc, kappa, (5,), 4200)
    assert infos[0]["converged"]
    xo, ok, _, _ = oracle.bicgstab(fields.to_mrhs(srcs[0], Nc), gauge, L, Nc, 5, kappa, 3000, 1e-10)
    assert np.linalg.norm(fields.to_mrhs(xs[0], Nc) - xo) / np.linalg.norm(xo) < 1e-8


def test_groups_reject_bad_arguments():
    L, Nc = (8, 8, 8, 8), 3
    api.initGridSize(L, Nc, 4, 2, 1)
    gauge = fields.gauge_field(L, Nc, 1, kind="haar")
    d_g = api.DeviceArray.from_complex(gauge.reshape(-1))
    api.loadQcuGauge(d_g, 2)
    api.getDslash(0, 0.1)
    with pytest.raises(api.QcuError):
        api.invert_groups([], 10, 1e-10)
    with pytest.raises(api.QcuError):
        api.invert_groups([([], [])], 10, 1e-10)
    api.finalizeQcu()
