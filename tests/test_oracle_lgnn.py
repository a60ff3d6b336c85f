"""Pins of the learned-preconditioner oracle (CPU only).

* With weights (1 -+ gamma_mu) on the hops the two layers reproduce
  1 + kappa^2 D_oe D_eo computed by the (independently pinned) dslash oracle,
  which fixes the transport, hop numbering and spin-matrix conventions.
* Gauge covariance of the network under random SU(N) transforms.
* The collapsed two-hop form equals the layer-by-layer evaluation.
* Analytic gradients agree with central finite differences.
"""
import numpy as np
import pytest

from oracle import dense, lgnn, oracle
from qcu_mrhs_amd import fields

L, Nc, R = (4, 4, 2, 4), 3, 2


def _setup(seed=1, kind="haar"):
    gauge = fields.gauge_field(L, Nc, seed, kind=kind, eps=0.4)
    tr = lgnn.Transport(gauge, L, Nc)
    Vh = fields.half_volume(L)
    g = np.random.default_rng(seed + 10)
    psi = g.standard_normal((Vh, 4, Nc, R)) + 1j * g.standard_normal((Vh, 4, Nc, R))
    return gauge, tr, psi


def _neumann_net(kappa):
    W1 = np.zeros((2, 1, 9, 4, 4), dtype=np.complex128)
    W2 = np.zeros((1, 2, 9, 4, 4), dtype=np.complex128)
    I4 = np.eye(4)
    W1[0, 0, 0] = I4
    W2[0, 0, 0] = I4
    for mu in range(4):
        W1[1, 0, 1 + 2 * mu] = kappa * (I4 - dense.GAMMA[mu])
        W1[1, 0, 2 + 2 * mu] = kappa * (I4 + dense.GAMMA[mu])
        W2[0, 1, 1 + 2 * mu] = kappa * (I4 - dense.GAMMA[mu])
        W2[0, 1, 2 + 2 * mu] = kappa * (I4 + dense.GAMMA[mu])
    return lgnn.Net(W1, W2)


def test_neumann_weights_reproduce_schur_neumann_term():
    kappa = 0.13
    gauge, tr, psi = _setup()
    net = _neumann_net(kappa)
    out, _ = net.forward(tr, psi, L, Nc, R)
    x = psi.reshape(-1)
    # psi + kappa^2 D_oe D_eo psi  ==  2 psi - schur(psi)
    expect = 2 * x - oracle.schur(x, gauge, L, Nc, R, kappa)
    np.testing.assert_allclose(out.reshape(-1), expect, rtol=0, atol=1e-12)


def test_gauge_covariance():
    gauge, tr, psi = _setup(3)
    net = lgnn.Net.random(8, 5)
    V = int(np.prod(L))
    gx = dense.random_gauge_transform(V, Nc, np.random.default_rng(7))
    gauge_t = dense.transform_gauge(gauge, L, Nc, gx)
    tr_t = lgnn.Transport(gauge_t, L, Nc)
    go = gx[V // 2:]
    psi_t = np.einsum("nab,nsbr->nsar", go, psi)
    out, _ = net.forward(tr, psi, L, Nc, R)
    out_t, _ = net.forward(tr_t, psi_t, L, Nc, R)
    np.testing.assert_allclose(out_t, np.einsum("nab,nsbr->nsar", go, out), rtol=0, atol=1e-11)


def test_collapsed_form_matches_layers():
    gauge, tr, psi = _setup(4)
    net = lgnn.Net.random(8, 9)
    out, _ = net.forward(tr, psi, L, Nc, R)
    M0, M = net.collapsed()
    V = int(np.prod(L))
    full = np.zeros((V, 4, Nc, R), dtype=np.complex128)
    full[V // 2:] = psi
    u = [tr.apply(full, h) for h in range(1, 9)]
    acc = np.einsum("st,vtcr->vscr", M0, full)
    for h in range(8):
        q = sum(np.einsum("st,vtcr->vscr", M[h, hp], u[hp]) for hp in range(8))
        q[V // 2:] = 0.0      # the hidden field feeding a hop lives on the even sites
        acc += tr.apply(q, h + 1)
    np.testing.assert_allclose(out, acc[V // 2:], rtol=0, atol=1e-11)


def test_gradients_match_finite_differences():
    gauge, tr, psi = _setup(6)
    g = np.random.default_rng(2)
    tgt = g.standard_normal(psi.shape) + 1j * g.standard_normal(psi.shape)
    net = lgnn.Net.random(3, 11, scale=0.3)
    loss, gW1, gW2 = net.loss_and_grad(tr, psi, tgt, L, Nc, R)
    eps = 1e-6
    for which, W, G in (("W1", net.W1, gW1), ("W2", net.W2, gW2)):
        for _ in range(6):
            idx = tuple(g.integers(0, n) for n in W.shape)
            for part, dz in (("re", 1.0), ("im", 1j)):
                Wp, Wm = W.copy(), W.copy()
                Wp[idx] += eps * dz
                Wm[idx] -= eps * dz
                netp = lgnn.Net(Wp, net.W2) if which == "W1" else lgnn.Net(net.W1, Wp)
                netm = lgnn.Net(Wm, net.W2) if which == "W1" else lgnn.Net(net.W1, Wm)
                lp = netp.loss_and_grad(tr, psi, tgt, L, Nc, R)[0]
                lm = netm.loss_and_grad(tr, psi, tgt, L, Nc, R)[0]
                fd = (lp - lm) / (2 * eps)
                an = G[idx].real if part == "re" else G[idx].imag
                assert abs(fd - an) <= 1e-5 * max(1.0, abs(an)), (which, idx, part, fd, an)


def test_adjoint_network_is_the_adjoint():
    """<y, N x> = <N^+ y, x> for random odd fields (pins adjoint_apply)."""
    gauge, tr, x = _setup(8)
    g = np.random.default_rng(3)
    y = g.standard_normal(x.shape) + 1j * g.standard_normal(x.shape)
    net = lgnn.Net.random(3, 12, scale=0.3)
    nx, _ = net.forward(tr, x, L, Nc, R)
    nty = net.adjoint_apply(tr, y, L, Nc, R)
    lhs, rhs = np.vdot(y, nx), np.vdot(nty, x)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


@pytest.mark.parametrize("K", [1, 3])
def test_power_loss_gradients_match_finite_differences(K):
    """Residual-power loss ||(1 - A N)^K v / ||v||||^2 with A the dslash
    oracle's odd Schur operator: the back-propagated gradient equals central
    finite differences of the loss; with the Neumann weights and K = 1 the
    loss is ||kappa^4 (D_oe D_eo)^2 w||^2 (1 - A (1 + K) = K^2, K = kappa^2 D_oe D_eo)."""
    kappa = 0.12
    gauge, tr, v = _setup(9, kind="smooth")

    def schur(x, dag):
        return oracle.schur(x.reshape(-1), gauge, L, Nc, R, kappa, dag)

    if K == 1:
        net = _neumann_net(kappa)
        loss, _, _ = net.power_loss_and_grad(tr, schur, v, 1, L, Nc, R)
        w = (v / np.sqrt((np.abs(v) ** 2).sum(axis=(0, 1, 2)))).reshape(-1)
        kw = w - schur(w, 0)
        k2w = kw - schur(kw, 0)
        assert abs(loss - np.vdot(k2w, k2w).real) <= 1e-10 * loss
    net = lgnn.Net.random(2, 13, scale=0.2)
    loss, gW1, gW2 = net.power_loss_and_grad(tr, schur, v, K, L, Nc, R)
    g = np.random.default_rng(4)
    eps = 1e-6
    for which, W, G in (("W1", net.W1, gW1), ("W2", net.W2, gW2)):
        for _ in range(4):
            idx = tuple(g.integers(0, n) for n in W.shape)
            for dz in (1.0, 1j):
                Wp, Wm = W.copy(), W.copy()
                Wp[idx] += eps * dz
                Wm[idx] -= eps * dz
                netp = lgnn.Net(Wp, net.W2) if which == "W1" else lgnn.Net(net.W1, Wp)
                netm = lgnn.Net(Wm, net.W2) if which == "W1" else lgnn.Net(net.W1, Wm)
                fd = (netp.power_loss_and_grad(tr, schur, v, K, L, Nc, R)[0] -
                      netm.power_loss_and_grad(tr, schur, v, K, L, Nc, R)[0]) / (2 * eps)
                an = G[idx].real if dz == 1.0 else G[idx].imag
                assert abs(fd - an) <= 1e-5 * max(1.0, abs(an)), (which, idx, dz, fd, an)


def _recon_table(h):
    """Spin reconstruction of the projected matrix-core gather for hop
    h = 1 + 2 mu (+1 backward): spin m[r] = i^kr[r] x spin r, r = 0, 1 -- the
    dslash's reconstruction without dagger (the (1 + gamma) form, plus a
    factor -1 forward)."""
    mu, back = (h - 1) >> 1, (h - 1) & 1
    m = (3, 2) if mu < 2 else (2, 3)
    k_plus = {0: (3, 3), 1: (2, 0), 2: (3, 1), 3: (0, 0)}[mu]
    return m, tuple((k + (0 if back else 2)) & 3 for k in k_plus)


def test_projected_family_has_two_spins_per_hop_field():
    """The package's P_h = (1 -+ gamma_mu) / 2 is the orthogonal projector
    onto the image of hop h's spin factor in the first Neumann term (the
    oracle's gamma matrices, pinned by the Neumann test above); its image is
    exactly the spinors whose spins 2, 3 follow from spins 0, 1 by the
    reconstruction the projected gather applies; second-layer weights moved
    into the family keep the Neumann term, and every collapsed block M[h][h']
    then maps into that image."""
    from qcu_mrhs_amd import lgnn as plgnn
    np.testing.assert_allclose(plgnn.GAMMA, dense.GAMMA, rtol=0, atol=0)
    g = np.random.default_rng(41)
    for h in range(1, 9):
        mu, back = (h - 1) >> 1, (h - 1) & 1
        P = plgnn.projector(h)
        S = np.eye(4) + (1 if back else -1) * dense.GAMMA[mu]
        np.testing.assert_allclose(P @ P, P, atol=1e-15)
        np.testing.assert_allclose(P, P.conj().T, atol=1e-15)
        np.testing.assert_allclose(P @ S, S, atol=1e-15)
        assert abs(np.trace(P) - 2) < 1e-15
        m, kr = _recon_table(h)
        v = P @ (g.standard_normal((4, 6)) + 1j * g.standard_normal((4, 6)))
        for r in (0, 1):
            np.testing.assert_allclose(v[m[r]], 1j ** kr[r] * v[r], atol=1e-14)
        # and a spinor obeying the reconstruction is in the image
        u = np.zeros(4, dtype=np.complex128)
        u[0], u[1] = 0.3 - 1.1j, 0.7 + 0.2j
        u[m[0]], u[m[1]] = 1j ** kr[0] * u[0], 1j ** kr[1] * u[1]
        np.testing.assert_allclose(P @ u, u, atol=1e-15)
    net = lgnn.Net.random(8, 11, scale=0.3)
    w2p = plgnn.project_w2(net.W2[0])
    np.testing.assert_allclose(plgnn.project_w2(w2p), w2p, atol=1e-15)
    np.testing.assert_array_equal(w2p[:, 0], net.W2[0][:, 0])
    _, M = lgnn.Net(net.W1, w2p[None]).collapsed()
    for k in range(8):
        P = plgnn.projector(k + 1)
        for kp in range(8):
            np.testing.assert_allclose(P @ M[k, kp], M[k, kp], atol=1e-13)
    neu = _neumann_net(0.13)
    np.testing.assert_allclose(plgnn.project_w2(neu.W2[0]), neu.W2[0], atol=1e-15)
