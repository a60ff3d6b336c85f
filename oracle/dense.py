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
