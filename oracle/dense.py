"""Independent dense/sparse formulation of the Wilson hopping operator.

TEST INFRASTRUCTURE ONLY (see oracle/oracle.py).  This module does not use
the projection tables of qcu_oracle.c: it writes the operator with the full
4x4 spin matrices,

    D[x, x+mu] = (1 - gamma_mu) (x) U_mu(x)
    D[x, x-mu] = (1 + gamma_mu) (x) U_mu(x-mu)^+

using the chiral gamma basis whose (1 + gamma_mu) matrices are spelled out in
the reference's comments (include/kernel/qcu_gamma.cuh:10-18, 70-76, 126-132,
185-191).  Agreement between this module and the oracle therefore pins the
oracle's projection/reconstruction tables, neighbour indexing and link
placement to the textbook operator.

Vector ordering: full lattice ``[parity][Vh][spin 4][Nc]`` (R = 1).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

I = 1j
GAMMA = np.array([
    [[0, 0, 0, I], [0, 0, I, 0], [0, -I, 0, 0], [-I, 0, 0, 0]],      # gamma_1 (x)
    [[0, 0, 0, -1], [0, 0, 1, 0], [0, 1, 0, 0], [-1, 0, 0, 0]],      # gamma_2 (y)
    [[0, 0, I, 0], [0, 0, 0, -I], [-I, 0, 0, 0], [0, I, 0, 0]],      # gamma_3 (z)
    [[0, 0, 1, 0], [0, 0, 0, 1], [1, 0, 0, 0], [0, 1, 0, 0]],        # gamma_4 (t)
], dtype=np.complex128)
GAMMA5 = GAMMA[0] @ GAMMA[1] @ GAMMA[2] @ GAMMA[3]


def coords(L):
    """Arrays (x, y, z, t) of every lattice site in full-vector order
    ``[parity][t][z][y][x/2]``."""
    Lx, Ly, Lz, Lt = L
    Vh = Lx * Ly * Lz * Lt // 2
    cb = np.arange(Vh)
    xh = cb % (Lx // 2)
    y = (cb // (Lx // 2)) % Ly
    z = (cb // (Lx // 2 * Ly)) % Lz
    t = cb // (Lx // 2 * Ly * Lz)
    out = []
    for p in (0, 1):
        x = 2 * xh + ((y + z + t + p) & 1)
        out.append(np.stack([x, y, z, t], axis=1))
    return np.concatenate(out, axis=0)


def site_index(c, L):
    """Full-vector site index of coordinates c[..., 4]."""
    Lx, Ly, Lz, Lt = L
    Vh = Lx * Ly * Lz * Lt // 2
    x, y, z, t = c[..., 0], c[..., 1], c[..., 2], c[..., 3]
    p = (x + y + z + t) & 1
    return p * Vh + ((t * Lz + z) * Ly + y) * (Lx // 2) + x // 2


def link(gauge_eo, L, mu, c):
    """U_mu at coordinates c[n, 4] from an even/odd gauge array."""
    Lx, Ly, Lz, Lt = L
    Vh = Lx * Ly * Lz * Lt // 2
    s = site_index(c, L)
    g = gauge_eo.reshape(4, 2 * Vh, gauge_eo.shape[-2], gauge_eo.shape[-1])
    return g[mu, s]


def hopping_matrix(gauge_eo, L, Nc, dagger=False):
    """Sparse full-lattice hopping matrix D (dagger: gamma -> -gamma)."""
    L = tuple(int(v) for v in L)
    V = int(np.prod(L))
    c = coords(L)
    rows, cols, vals = [], [], []
    sgn = -1.0 if dagger else 1.0
    eye4 = np.eye(4)
    for mu in range(4):
        step = np.zeros(4, dtype=np.int64)
        step[mu] = 1
        Lv = np.array(L)
        cf = (c + step) % Lv
        cb = (c - step) % Lv
        s0 = site_index(c, L)
        sf = site_index(cf, L)
        sb = site_index(cb, L)
        Uf = link(gauge_eo, L, mu, c)                                  # U_mu(x)
        Ub = np.conj(np.swapaxes(link(gauge_eo, L, mu, cb), 1, 2))     # U_mu(x-mu)^+
        Pf = eye4 - sgn * GAMMA[mu]
        Pb = eye4 + sgn * GAMMA[mu]
        for (snb, U, P) in ((sf, Uf, Pf), (sb, Ub, Pb)):
            blk = np.einsum("ab,nij->naibj", P, U)                      # [n][4][Nc][4][Nc]
            blk = blk.reshape(V, 4 * Nc, 4 * Nc)
            r = (s0[:, None, None] * 4 * Nc + np.arange(4 * Nc)[None, :, None])
            q = (snb[:, None, None] * 4 * Nc + np.arange(4 * Nc)[None, None, :])
            r, q = np.broadcast_arrays(r, q)
            rows.append(r.reshape(-1))
            cols.append(q.reshape(-1))
            vals.append(blk.reshape(-1))
    n = V * 4 * Nc
    D = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(n, n)).tocsr()
    D.sum_duplicates()
    return D


def plane_wave(L, Nc, momentum, chi):
    """Full-lattice vector psi(x) = exp(i p.x) chi, p_mu = 2 pi n_mu / L_mu."""
    c = coords(L)
    p = 2 * np.pi * np.asarray(momentum, dtype=np.float64) / np.asarray(L, dtype=np.float64)
    ph = np.exp(1j * (c @ p))
    return (ph[:, None] * np.asarray(chi, dtype=np.complex128).reshape(1, 4 * Nc)).reshape(-1)


def free_field_symbol(L, Nc, momentum, dagger=False):
    """The 4Nc x 4Nc matrix  sum_mu [2 cos p_mu - 2 i sin p_mu gamma_mu] (x) 1,
    i.e. D acting on a plane wave with U = 1 (dagger flips gamma's sign)."""
    p = 2 * np.pi * np.asarray(momentum, dtype=np.float64) / np.asarray(L, dtype=np.float64)
    sgn = -1.0 if dagger else 1.0
    m = np.zeros((4, 4), dtype=np.complex128)
    for mu in range(4):
        m += 2 * np.cos(p[mu]) * np.eye(4) - sgn * 2j * np.sin(p[mu]) * GAMMA[mu]
    return np.kron(m, np.eye(Nc))


def random_gauge_transform(V, Nc, g):
    """V Haar SU(Nc) matrices (QR with phase fix, det normalised)."""
    z = (g.standard_normal((V, Nc, Nc)) + 1j * g.standard_normal((V, Nc, Nc))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r, axis1=1, axis2=2)
    q = q * (d / np.abs(d))[:, None, :]
    q = q / (np.linalg.det(q) ** (1.0 / Nc))[:, None, None]
    return q


def transform_gauge(gauge_eo, L, Nc, gx):
    """U_mu(x) -> g(x) U_mu(x) g(x+mu)^+ ; gx indexed by full-vector site."""
    L = tuple(int(v) for v in L)
    V = int(np.prod(L))
    c = coords(L)
    out = np.empty((4, V, Nc, Nc), dtype=np.complex128)
    for mu in range(4):
        step = np.zeros(4, dtype=np.int64)
        step[mu] = 1
        sf = site_index((c + step) % np.array(L), L)
        s0 = site_index(c, L)
        U = link(gauge_eo, L, mu, c)
        out[mu, s0] = gx[s0] @ U @ np.conj(np.swapaxes(gx[sf], 1, 2))
    return out.reshape(4, 2, V // 2, Nc, Nc)


def transform_fermion(psi_full, gx, Nc):
    """psi(x) -> g(x) psi(x) for a full-lattice R=1 vector."""
    V = gx.shape[0]
    v = psi_full.reshape(V, 4, Nc)
    return np.einsum("nab,nsb->nsa", gx, v).reshape(-1)


def sigma(mu, nu):
    """sigma_{mu nu} = (i/2) [gamma_mu, gamma_nu] (4x4)."""
    return 0.5j * (GAMMA[mu] @ GAMMA[nu] - GAMMA[nu] @ GAMMA[mu])


def field_strength(gauge_eo, L, Nc):
    """F_{mu nu}(x) for mu < nu, as [6][V][Nc][Nc] in full-vector site order:
    the clover average of the four plaquette leaves, made Hermitian and
    traceless, F = ((Q - Q^+) - tr(Q - Q^+)/Nc) / (8 i).  Written with whole-
    lattice coordinate arrays (no per-site loop), independently of the C
    oracle's path products."""
    L = tuple(int(v) for v in L)
    c = coords(L)
    Lv = np.array(L)

    def U(mu, shift):
        return link(gauge_eo, L, mu, (c + np.asarray(shift)) % Lv)

    def dag(m):
        return np.conj(np.swapaxes(m, -1, -2))

    out = []
    for mu in range(4):
        for nu in range(mu + 1, 4):
            e_mu = np.eye(4, dtype=np.int64)[mu]
            e_nu = np.eye(4, dtype=np.int64)[nu]
            z = np.zeros(4, dtype=np.int64)
            leaves = (
                U(mu, z) @ U(nu, e_mu) @ dag(U(mu, e_nu)) @ dag(U(nu, z)),
                U(nu, z) @ dag(U(mu, e_nu - e_mu)) @ dag(U(nu, -e_mu)) @ U(mu, -e_mu),
                dag(U(mu, -e_mu)) @ dag(U(nu, -e_mu - e_nu)) @ U(mu, -e_mu - e_nu) @ U(nu, -e_nu),
                dag(U(nu, -e_nu)) @ U(mu, -e_nu) @ U(nu, e_mu - e_nu) @ dag(U(mu, z)),
            )
            Q = sum(leaves)
            A = Q - dag(Q)
            tr = np.trace(A, axis1=1, axis2=2)
            A = A - (tr / Nc)[:, None, None] * np.eye(Nc)[None]
            out.append(A / 8j)
    return np.stack(out)


def clover_site_matrices(gauge_eo, L, Nc, csw, kappa):
    """A(x) = 1 - kappa c_sw sum_{mu<nu} sigma_{mu nu} (x) F_{mu nu}(x), the full
    4Nc x 4Nc matrix per site ([V][4Nc][4Nc], rows spin-major), built from the
    explicit 4x4 gamma matrices."""
    F = field_strength(gauge_eo, L, Nc)
    V = F.shape[1]
    A = np.broadcast_to(np.eye(4 * Nc, dtype=np.complex128), (V, 4 * Nc, 4 * Nc)).copy()
    q = 0
    for mu in range(4):
        for nu in range(mu + 1, 4):
            s = sigma(mu, nu)
            A -= kappa * csw * np.einsum("st,nab->nsatb", s, F[q]).reshape(V, 4 * Nc, 4 * Nc)
            q += 1
    return A


def clover_matrix(gauge_eo, L, Nc, csw, kappa):
    """Sparse block-diagonal full-lattice clover matrix A."""
    A = clover_site_matrices(gauge_eo, L, Nc, csw, kappa)
    return sp.block_diag(list(A), format="csr")


def constant_abelian_field(L, Nc, mu, nu, n_flux):
    """Links of a constant abelian field in the (mu, nu) plane: U_nu(x) =
    exp(i B x_mu tau), U_mu = 1 except U_mu(x_mu = L_mu - 1) = exp(-i B L_mu x_nu
    tau), the other links 1, with tau = diag(1, -1, 0, ...) and B = 2 pi n_flux /
    (L_mu L_nu) (n_flux quanta through the periodic plane).  Every (mu, nu)
    plaquette U_mu(x) U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+ equals exp(i B tau) and
    every other plaquette is 1.  Returns (even/odd gauge [4][2][Vh][Nc][Nc], B)."""
    L = tuple(int(v) for v in L)
    V = int(np.prod(L))
    B = 2 * np.pi * n_flux / (L[mu] * L[nu])
    c = coords(L)
    th = np.zeros((4, V))
    th[nu] = B * c[:, mu]
    th[mu] = np.where(c[:, mu] == L[mu] - 1, -B * L[mu] * c[:, nu], 0.0)
    tau = np.zeros(Nc)
    tau[0], tau[1] = 1.0, -1.0
    U = np.zeros((4, V, Nc, Nc), dtype=np.complex128)
    U[:, :, np.arange(Nc), np.arange(Nc)] = np.exp(1j * th[:, :, None] * tau[None, None, :])
    # coords() lists sites in full-vector order, which is the even/odd storage order
    return U.reshape(4, 2, V // 2, Nc, Nc), B
