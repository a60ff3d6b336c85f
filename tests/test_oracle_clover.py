"""Pins of the oracle's clover term (CPU only).

The reference declares the clover dslash type but implements none
(include/qcu_helper_macro.h:7-11, src/qcu_interface.cu:97-105), so the oracle
is the textbook Sheikholeslami-Wohlert term and these checks pin it to the
mathematics: the free field, the closed form of a constant abelian field (leaf
orientation and the 1/8), the whole-lattice numpy construction from the
explicit 4x4 gamma matrices (site indexing, link placement, chirality blocks),
Hermiticity and gauge covariance, and the O(a) cancellation the term exists
for, which fixes its sign and normalisation relative to the Wilson hop.
"""
import numpy as np
import pytest
import scipy.linalg as sl
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from oracle import dense, oracle
from qcu_mrhs_amd import fields


def _full_blocks(c):
    """[2][Vh][2][n][n] chirality blocks -> [V][4Nc][4Nc] site matrices."""
    return np.stack([sl.block_diag(c[p, i, 0], c[p, i, 1]) for p in range(2) for i in range(c.shape[1])])


@pytest.mark.parametrize("Nc", [2, 3, 4])
def test_unit_links_give_identity(Nc):
    L = (4, 2, 2, 4)
    c = oracle.clover(fields.gauge_field(L, Nc, 0, kind="unit"), L, Nc, 1.3, 0.13)
    np.testing.assert_array_equal(c, np.broadcast_to(np.eye(2 * Nc), c.shape))


@pytest.mark.parametrize("plane", [(0, 1), (0, 3), (1, 2), (2, 3)])
@pytest.mark.parametrize("Nc", [2, 3])
def test_constant_abelian_field_closed_form(plane, Nc):
    """Every (mu, nu) plaquette is exp(i B tau): F_{mu nu} = sin(B) tau exactly,
    the other planes vanish, so A = 1 - kappa c_sw sigma_{mu nu} (x) sin(B) tau
    at every site."""
    mu, nu = plane
    L = (4, 6, 4, 2) if plane != (2, 3) else (2, 2, 4, 6)
    gauge, B = dense.constant_abelian_field(L, Nc, mu, nu, 1)
    csw, kappa = 1.7, 0.12
    tau = np.zeros((Nc, Nc))
    tau[0, 0], tau[1, 1] = 1.0, -1.0
    expect = np.eye(4 * Nc) - kappa * csw * np.kron(dense.sigma(mu, nu), np.sin(B) * tau)
    got = _full_blocks(oracle.clover(gauge, L, Nc, csw, kappa))
    np.testing.assert_allclose(got, np.broadcast_to(expect, got.shape), rtol=0, atol=1e-14)


@pytest.mark.parametrize("Nc,kind", [(3, "smooth"), (2, "haar"), (4, "haar")])
def test_matches_whole_lattice_dense_construction(Nc, kind):
    L = (4, 4, 2, 4)
    gauge = fields.gauge_field(L, Nc, 40 + Nc, kind=kind, eps=0.5)
    got = _full_blocks(oracle.clover(gauge, L, Nc, 1.1, 0.135))
    expect = dense.clover_site_matrices(gauge, L, Nc, 1.1, 0.135)
    np.testing.assert_allclose(got, expect, rtol=0, atol=1e-14)


def test_hermitian_and_gauge_covariant():
    """A(x) Hermitian; under U_mu(x) -> g(x) U_mu(x) g(x+mu)^+ the blocks go to
    (1 (x) g(x)) A(x) (1 (x) g(x))^+."""
    L, Nc = (4, 4, 4, 2), 3
    gauge = fields.gauge_field(L, Nc, 11, kind="haar")
    c = oracle.clover(gauge, L, Nc, 1.0, 0.13)
    np.testing.assert_allclose(c, np.conj(np.swapaxes(c, -1, -2)), rtol=0, atol=1e-14)
    V = int(np.prod(L))
    gx = dense.random_gauge_transform(V, Nc, np.random.default_rng(3))
    c2 = oracle.clover(dense.transform_gauge(gauge, L, Nc, gx), L, Nc, 1.0, 0.13)
    G = np.stack([np.kron(np.eye(2), g) for g in gx]).reshape(2, V // 2, 1, 2 * Nc, 2 * Nc)
    expect = G @ c @ np.conj(np.swapaxes(G, -1, -2))
    np.testing.assert_allclose(c2, expect, rtol=0, atol=1e-13)


def test_sign_and_normalisation_by_o_a_cancellation():
    """Constant field, one flux quantum in the (x, y) plane of an 8 x 8 x 2 x 2
    lattice, SU(2)-embedded, at kappa = 1/8 (m = 0): the continuum Dirac operator
    has exact zero modes; the Wilson term moves them by O(a B) (B/2 here) and
    the clover term at c_sw = 1 removes that shift, leaving O(a^2 B^2).  A sign
    error doubles the shift; a wrong normalisation leaves it O(a B)."""
    L, Nc, kappa = (8, 8, 2, 2), 2, 0.125
    gauge, B = dense.constant_abelian_field(L, Nc, 0, 1, 1)
    D = dense.hopping_matrix(gauge, L, Nc)
    lowest = {}
    for csw in (-1.0, 0.0, 0.5, 1.0, 1.5):
        A = sp.block_diag(list(_full_blocks(oracle.clover(gauge, L, Nc, csw, kappa))), format="csr")
        ev = sla.eigs((A - kappa * D).tocsc(), k=4, sigma=0, return_eigenvectors=False)
        lowest[csw] = np.abs(ev).min() / (2 * kappa)    # eigenvalue of D_W + m
    assert abs(lowest[0.0] - B / 2) < 0.02 * B
    assert lowest[1.0] < 0.1 * B * B
    assert abs(lowest[-1.0] - B) < 0.02 * B
    assert abs(lowest[0.5] - B / 4) < 0.02 * B and abs(lowest[1.5] - B / 4) < 0.02 * B


def test_apply_invert_and_full_operator():
    """clover_apply against the dense block-diagonal matrix (R = 3), the block
    inverse against numpy.linalg.inv, clover_mat against A - kappa D (dense)
    and gamma5-hermiticity of the clover operator."""
    L, Nc, R, kappa, csw = (4, 4, 2, 4), 3, 3, 0.13, 1.2
    gauge = fields.gauge_field(L, Nc, 21, kind="smooth", eps=0.6)
    c = oracle.clover(gauge, L, Nc, csw, kappa)
    inv = oracle.clover_invert(c)
    np.testing.assert_allclose(inv, np.linalg.inv(c), rtol=0, atol=1e-13)
    np.testing.assert_allclose(c @ inv, np.broadcast_to(np.eye(2 * Nc), c.shape), rtol=0, atol=1e-13)
    A = dense.clover_matrix(gauge, L, Nc, csw, kappa)
    M = A - kappa * dense.hopping_matrix(gauge, L, Nc)
    b = fields.gaussian_sources(L, Nc, R, 22)
    bm = fields.to_mrhs(b, Nc)
    h = bm.size // 2
    for p in (0, 1):
        got = oracle.clover_apply(c[p], bm[p * h:(p + 1) * h], Nc, R)
        expect = fields.to_mrhs(np.stack([(A @ b[r]).reshape(2, -1)[p] for r in range(R)]), Nc)
        np.testing.assert_allclose(got, expect, rtol=0, atol=1e-13)
    got = oracle.clover_mat(bm, gauge, c, L, Nc, R, kappa)
    expect = fields.to_mrhs(np.stack([M @ b[r] for r in range(R)]), Nc)
    np.testing.assert_allclose(got, expect, rtol=0, atol=1e-12)
    V = int(np.prod(L))
    g5 = sp.kron(sp.eye(V), sp.kron(sp.csr_matrix(dense.GAMMA5), sp.eye(Nc)))
    Md = M.toarray()
    np.testing.assert_allclose((g5 @ M @ g5).toarray(), np.conj(Md.T), rtol=0, atol=1e-13)
