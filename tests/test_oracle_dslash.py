"""Pins of the oracle's dslash (CPU only).

Each check is chosen so that a plausible restatement mistake fails it:
a wrong gamma sign or table entry (plane wave, dense operator), a wrong
neighbour index or wrap (plane wave with non-zero momenta in every
direction), U vs U^+ or U(x) vs U(x-mu) (gauge covariance, dense operator),
a wrong dagger sign (adjoint, gamma5-hermiticity), RHS mixing (RHS
independence).
"""
import numpy as np
import pytest

from oracle import dense, oracle
from qcu_mrhs_amd import fields


def _half(full, parity, Vh, per_site):
    return full.reshape(2, Vh * per_site)[parity]


@pytest.mark.parametrize("L,Nc", [((4, 4, 4, 8), 3), ((6, 4, 2, 4), 2), ((4, 2, 4, 4), 4)])
def test_free_field_plane_wave(L, Nc):
    """U = 1: D e^{ipx} chi = e^{ipx} sum_mu [2cos p_mu - 2i sin p_mu g_mu] chi."""
    g = np.random.default_rng(1)
    Vh = fields.half_volume(L)
    gauge = fields.gauge_field(L, Nc, 0, kind="unit")
    R = 2
    for dagger in (0, 1):
        for trial in range(3):
            mom = [g.integers(0, l) for l in L]
            chis = [g.standard_normal(4 * Nc) + 1j * g.standard_normal(4 * Nc) for _ in range(R)]
            psi = np.stack([dense.plane_wave(L, Nc, mom, chi) for chi in chis])  # [R][V*4Nc]
            sym = dense.free_field_symbol(L, Nc, mom, dagger=bool(dagger))
            expect = np.stack([dense.plane_wave(L, Nc, mom, sym @ chi) for chi in chis])
            for p in (0, 1):
                inp = fields.to_mrhs(psi.reshape(R, 2, -1)[:, 1 - p], Nc)
                out = oracle.dslash(inp, gauge, L, Nc, R, p, dagger)
                exp_p = fields.to_mrhs(expect.reshape(R, 2, -1)[:, p], Nc)
                np.testing.assert_allclose(out, exp_p, rtol=0, atol=1e-12 * np.abs(exp_p).max())


@pytest.mark.parametrize("dagger", [0, 1])
@pytest.mark.parametrize("L,Nc,kind", [((4, 4, 4, 4), 3, "haar"), ((4, 6, 4, 2), 2, "smooth"),
                                        ((2, 4, 2, 4), 4, "haar")])
def test_matches_dense_operator(L, Nc, kind, dagger):
    """Oracle (projection tables) == explicit 4x4 gamma-matrix operator."""
    gauge = fields.gauge_field(L, Nc, 11, kind=kind, eps=0.4)
    D = dense.hopping_matrix(gauge, L, Nc, dagger=bool(dagger))
    R = 3
    Vh = fields.half_volume(L)
    psi = fields.gaussian_sources(L, Nc, R, 5)
    ref = np.stack([D @ psi[r] for r in range(R)])
    for p in (0, 1):
        out = oracle.dslash(fields.to_mrhs(psi.reshape(R, 2, -1)[:, 1 - p], Nc), gauge, L, Nc, R,
                            p, dagger)
        np.testing.assert_allclose(out, fields.to_mrhs(ref.reshape(R, 2, -1)[:, p], Nc),
                                   rtol=0, atol=1e-12)
    # parity structure: D has no block within one parity
    n = Vh * 4 * Nc
    assert abs(D[:n, :n]).sum() == 0 and abs(D[n:, n:]).sum() == 0


def test_adjoint_and_gamma5_hermiticity():
    L, Nc, R = (8, 4, 4, 6), 3, 3
    gauge = fields.gauge_field(L, Nc, 3, kind="haar")
    Vh = fields.half_volume(L)
    g = np.random.default_rng(7)
    n = Vh * 4 * Nc * R
    phi = g.standard_normal(n) + 1j * g.standard_normal(n)
    psi = g.standard_normal(n) + 1j * g.standard_normal(n)
    g5 = np.repeat(np.array([1, 1, -1, -1], dtype=np.float64), Nc * R)
    g5 = np.tile(g5, Vh)
    for p in (0, 1):
        Dpsi = oracle.dslash(psi, gauge, L, Nc, R, p, 0)             # parity 1-p -> p
        Ddphi = oracle.dslash(phi, gauge, L, Nc, R, 1 - p, 1)        # parity p -> 1-p
        lhs = oracle.dotc(phi, Dpsi, R)
        rhs = oracle.dotc(Ddphi, psi, R)
        np.testing.assert_allclose(lhs, rhs, rtol=1e-12)
        # gamma5 D gamma5 = D^+
        np.testing.assert_allclose(g5 * oracle.dslash(g5 * psi, gauge, L, Nc, R, p, 0),
                                   oracle.dslash(psi, gauge, L, Nc, R, p, 1), rtol=0, atol=1e-12)


@pytest.mark.parametrize("Nc", [2, 3])
def test_gauge_covariance(Nc):
    L, R = (4, 4, 6, 4), 2
    V = int(np.prod(L))
    Vh = V // 2
    gauge = fields.gauge_field(L, Nc, 21, kind="haar")
    gx = dense.random_gauge_transform(V, Nc, np.random.default_rng(3))
    gauge_t = dense.transform_gauge(gauge, L, Nc, gx)
    psi = fields.gaussian_sources(L, Nc, R, 9)
    psi_t = np.stack([dense.transform_fermion(psi[r], gx, Nc) for r in range(R)])
    for p in (0, 1):
        out = oracle.dslash(fields.to_mrhs(psi.reshape(R, 2, -1)[:, 1 - p], Nc), gauge, L, Nc, R, p)
        out_t = oracle.dslash(fields.to_mrhs(psi_t.reshape(R, 2, -1)[:, 1 - p], Nc), gauge_t, L,
                              Nc, R, p)
        # g(x) applied to the untransformed result
        out_u = fields.from_mrhs(out, R)
        full = np.zeros((R, 2, Vh * 4 * Nc), dtype=np.complex128)
        full[:, p] = out_u
        expect = np.stack([dense.transform_fermion(full[r].reshape(-1), gx, Nc) for r in range(R)])
        np.testing.assert_allclose(out_t, fields.to_mrhs(expect.reshape(R, 2, -1)[:, p], Nc),
                                   rtol=0, atol=1e-12)


def test_rhs_independence_and_permutation():
    L, Nc, R = (4, 4, 4, 4), 3, 5
    gauge = fields.gauge_field(L, Nc, 4, kind="haar")
    psi = fields.gaussian_sources(L, Nc, R, 2).reshape(R, 2, -1)[:, 0]
    out = fields.from_mrhs(oracle.dslash(fields.to_mrhs(psi, Nc), gauge, L, Nc, R, 1), R)
    for r in range(R):
        single = oracle.dslash(fields.to_mrhs(psi[r:r + 1], Nc), gauge, L, Nc, 1, 1)
        np.testing.assert_array_equal(out[r], single)
    perm = np.random.default_rng(0).permutation(R)
    outp = fields.from_mrhs(oracle.dslash(fields.to_mrhs(psi[perm], Nc), gauge, L, Nc, R, 1), R)
    np.testing.assert_array_equal(outp, out[perm])


def test_single_site_source_reaches_only_neighbours():
    L, Nc, R = (6, 4, 4, 4), 3, 1
    Vh = fields.half_volume(L)
    gauge = fields.gauge_field(L, Nc, 8, kind="haar")
    inp = np.zeros(Vh * 4 * Nc, dtype=np.complex128)
    site = 17
    g = np.random.default_rng(5)
    inp[site * 4 * Nc:(site + 1) * 4 * Nc] = g.standard_normal(4 * Nc) + 1j * g.standard_normal(4 * Nc)
    out = oracle.dslash(inp, gauge, L, Nc, R, 0).reshape(Vh, -1)
    hit = np.nonzero(np.abs(out).sum(axis=1))[0]
    c = dense.coords(L)[Vh + site]
    nbs = set()
    for mu in range(4):
        for d in (1, -1):
            cc = c.copy()
            cc[mu] = (cc[mu] + d) % L[mu]
            nbs.add(int(dense.site_index(cc, L)))
    assert set(int(h) for h in hit) == nbs
