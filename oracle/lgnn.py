"""CPU reference of the learned gauge-covariant preconditioner.

TEST INFRASTRUCTURE ONLY (see oracle/oracle.py): the checker for the GPU
forward and backward kernels of qcu_mrhs_amd/csrc/lgnn.hip.

Model (parallel-transport convolution layers of Lehner & Wettig 2023,
"Gauge-equivariant neural networks as preconditioners in lattice QCD"):

    layer(psi)_o(x) = sum_i sum_{h=0..8} W[o][i][h] (T_h psi_i)(x)

with T_0 = 1, T_{mu+} f(x) = U_mu(x) f(x+mu), T_{mu-} f(x) = U_mu(x-mu)^+ f(x-mu)
(hop index h = 1 + 2 mu + {0 forward, 1 backward}) and W a complex 4x4 spin
matrix acting on the spin index only, so the layer is gauge covariant.
The preconditioner is two linear layers 1 -> C -> 1 (C = 8 channels) applied
to a field that lives on the odd sites and restricted back to them.

This module evaluates that definition literally on full-lattice fields
(both parities, every hop), independently of the collapsed two-hop form the
GPU uses.  Gradients are obtained by explicit back-propagation through the
two layers; a finite-difference check pins them (tests/test_oracle_lgnn.py).

Field layout: full lattice [parity][Vh][spin 4][Nc][R] complex128, parity-major
(as qcu_mrhs_amd.fields.to_mrhs on per-RHS full vectors).
"""
from __future__ import annotations

import numpy as np

from . import dense

NHOP = 9


def hop_tables(L):
    """For every full-vector site: its 8 neighbour sites (hop 1..8)."""
    L = tuple(int(v) for v in L)
    c = dense.coords(L)
    Lv = np.array(L)
    nb = np.empty((c.shape[0], 8), dtype=np.int64)
    for mu in range(4):
        step = np.zeros(4, dtype=np.int64)
        step[mu] = 1
        nb[:, 2 * mu] = dense.site_index((c + step) % Lv, L)
        nb[:, 2 * mu + 1] = dense.site_index((c - step) % Lv, L)
    return nb


def _links(gauge_eo, L, Nc):
    """Per full-vector site and hop h=1..8 the transport matrix:
    forward U_mu(x), backward U_mu(x-mu)^+ ."""
    L = tuple(int(v) for v in L)
    V = int(np.prod(L))
    g = np.asarray(gauge_eo).reshape(4, V, Nc, Nc)
    nb = hop_tables(L)
    T = np.empty((V, 8, Nc, Nc), dtype=np.complex128)
    for mu in range(4):
        T[:, 2 * mu] = g[mu]
        T[:, 2 * mu + 1] = np.conj(np.swapaxes(g[mu][nb[:, 2 * mu + 1]], 1, 2))
    return T, nb


class Transport:
    """T_h on full-lattice fields f[V][4][Nc][R] (h = 0..8) and its adjoint."""

    def __init__(self, gauge_eo, L, Nc):
        self.T, self.nb = _links(gauge_eo, L, Nc)

    def apply(self, f, h):
        if h == 0:
            return f.copy()
        k = h - 1
        return np.einsum("vab,vsbr->vsar", self.T[:, k], f[self.nb[:, k]])

    def adjoint(self, f, h):
        """T_h^+ :  (T_{mu+})^+ = T_{mu-}, (T_{mu-})^+ = T_{mu+}."""
        if h == 0:
            return f.copy()
        k = h - 1
        return self.apply(f, 1 + (k ^ 1))


def ptc(tr, W, fields):
    """One layer: W[o][i][h] (4x4), fields[i][V][4][Nc][R] -> out[o]."""
    C_out, C_in = W.shape[0], W.shape[1]
    hopped = [[tr.apply(fields[i], h) for h in range(NHOP)] for i in range(C_in)]
    out = []
    for o in range(C_out):
        acc = np.zeros_like(fields[0])
        for i in range(C_in):
            for h in range(NHOP):
                acc += np.einsum("st,vtcr->vscr", W[o, i, h], hopped[i][h])
        out.append(acc)
    return out, hopped


def odd_mask(L, shape):
    V = int(np.prod(L))
    m = np.zeros(shape, dtype=bool)
    m[V // 2:] = True
    return m


class Net:
    """Two linear PTC layers 1 -> C -> 1 on odd-site fields."""

    def __init__(self, W1, W2):
        self.W1 = np.asarray(W1, dtype=np.complex128)   # [C][1][9][4][4]
        self.W2 = np.asarray(W2, dtype=np.complex128)   # [1][C][9][4][4]

    @staticmethod
    def random(C, seed, scale=0.1):
        g = np.random.default_rng(seed)
        W1 = (g.standard_normal((C, 1, NHOP, 4, 4)) + 1j * g.standard_normal((C, 1, NHOP, 4, 4))) * scale
        W2 = (g.standard_normal((1, C, NHOP, 4, 4)) + 1j * g.standard_normal((1, C, NHOP, 4, 4))) * scale
        return Net(W1, W2)

    def forward(self, tr, psi_odd, L, Nc, R):
        """psi_odd: [Vh][4][Nc][R] -> N psi on the odd sites, plus the cache."""
        V = int(np.prod(L))
        full = np.zeros((V, 4, Nc, R), dtype=np.complex128)
        full[V // 2:] = psi_odd.reshape(V // 2, 4, Nc, R)
        hid, hop1 = ptc(tr, self.W1, [full])
        out, hop2 = ptc(tr, self.W2, hid)
        return out[0][V // 2:].copy(), (full, hid, hop1, hop2)

    def loss_and_grad(self, tr, psi_odd, target_odd, L, Nc, R):
        """L = sum_r ||N psi_r - e_r||^2 / ||e_r||^2 ; gradients are
        dL/dRe W + i dL/dIm W (same shapes as W1, W2)."""
        V = int(np.prod(L))
        phi, _ = self.forward(tr, psi_odd, L, Nc, R)
        tgt = target_odd.reshape(V // 2, 4, Nc, R)
        nrm = (np.abs(tgt) ** 2).sum(axis=(0, 1, 2))          # per RHS
        diff = phi - tgt
        loss = float(((np.abs(diff) ** 2).sum(axis=(0, 1, 2)) / nrm).sum())
        gW1, gW2 = self.grad_given_dphi(tr, psi_odd, diff / nrm, L, Nc, R)
        return loss, gW1, gW2

    def adjoint_apply(self, tr, y_odd, L, Nc, R):
        """N^+ y on the odd sites: the adjoints of the two layers in reverse
        order (T_h^+ and W^+), on full-lattice fields, restricted to odd."""
        V = int(np.prod(L))
        full = np.zeros((V, 4, Nc, R), dtype=np.complex128)
        full[V // 2:] = y_odd.reshape(V // 2, 4, Nc, R)
        C = self.W1.shape[0]
        dhid = []
        for i in range(C):
            acc = np.zeros_like(full)
            for h in range(NHOP):
                acc += tr.adjoint(np.einsum("ts,vtcr->vscr", np.conj(self.W2[0, i, h]), full), h)
            dhid.append(acc)
        out = np.zeros_like(full)
        for o in range(C):
            for h in range(NHOP):
                out += tr.adjoint(np.einsum("ts,vtcr->vscr", np.conj(self.W1[o, 0, h]), dhid[o]), h)
        return out[V // 2:].copy()

    def power_loss_and_grad(self, tr, schur, v_odd, K, L, Nc, R):
        """Residual-power loss L = sum_r ||(1 - A N)^K w_r||^2 with
        w_r = v_r / ||v_r|| and A = schur(x, dagger) the odd Schur operator;
        gradients by back-propagation through the K factors:
        g_K = w_K, dL/d(N w_j)^* = -A^+ g_{j+1}, g_j = g_{j+1} - N^+ A^+ g_{j+1}."""
        V = int(np.prod(L))
        shape = (V // 2, 4, Nc, R)
        v = v_odd.reshape(shape)
        nrm = np.sqrt((np.abs(v) ** 2).sum(axis=(0, 1, 2)))
        w = [v / nrm]
        for _ in range(K):
            y, _ = self.forward(tr, w[-1], L, Nc, R)
            w.append(w[-1] - schur(y, 0).reshape(shape))
        loss = float((np.abs(w[K]) ** 2).sum())
        gW1 = np.zeros_like(self.W1)
        gW2 = np.zeros_like(self.W2)
        g = w[K]
        for j in range(K - 1, -1, -1):
            u = schur(g, 1).reshape(shape)
            a1, a2 = self.grad_given_dphi(tr, w[j], -u, L, Nc, R)
            gW1 += a1
            gW2 += a2
            if j > 0:
                g = g - self.adjoint_apply(tr, u, L, Nc, R)
        return loss, gW1, gW2

    def grad_given_dphi(self, tr, psi_odd, dphi_odd, L, Nc, R):
        """dL/dRe W + i dL/dIm W for a loss whose derivative with respect to
        the conjugate of the network output N psi is dphi (odd sites)."""
        V = int(np.prod(L))
        _, (full, hid, hop1, hop2) = self.forward(tr, psi_odd, L, Nc, R)
        dphi = np.zeros((V, 4, Nc, R), dtype=np.complex128)
        dphi[V // 2:] = dphi_odd.reshape(V // 2, 4, Nc, R)
        C = self.W1.shape[0]
        gW2 = np.zeros_like(self.W2)
        dhid = []
        for i in range(C):
            acc = np.zeros_like(dphi)
            for h in range(NHOP):
                # phi = sum W2[0][i][h] (T_h hid_i)
                gW2[0, i, h] = 2 * np.einsum("vscr,vtcr->st", dphi, np.conj(hop2[i][h]))
                acc += tr.adjoint(np.einsum("ts,vtcr->vscr", np.conj(self.W2[0, i, h]), dphi), h)
            dhid.append(acc)
        gW1 = np.zeros_like(self.W1)
        for o in range(C):
            for h in range(NHOP):
                gW1[o, 0, h] = 2 * np.einsum("vscr,vtcr->st", dhid[o], np.conj(hop1[0][h]))
        return gW1, gW2

    def collapsed(self):
        """(M0, M[h][h']) of the equivalent two-hop stencil on odd sites:
        N psi = M0 psi + sum_{h,h'>=1} T_h M[h][h'] T_h' psi."""
        C = self.W1.shape[0]
        M0 = sum(self.W2[0, o, 0] @ self.W1[o, 0, 0] for o in range(C))
        M = np.zeros((8, 8, 4, 4), dtype=np.complex128)
        for h in range(1, NHOP):
            for hp in range(1, NHOP):
                M[h - 1, hp - 1] = sum(self.W2[0, o, h] @ self.W1[o, 0, hp] for o in range(C))
        return M0, M
