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
