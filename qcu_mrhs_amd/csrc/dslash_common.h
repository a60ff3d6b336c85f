// dslash_common.h -- spin tables, site geometry and lattice traversal shared
// by the Wilson hopping kernels.
#pragma once

#include "dslash.h"

namespace qcu {

// Projection / reconstruction tables as powers of i for the (1 + gamma) form;
// the forward hop and the dagger each add a factor -1 = i^2.
// mat2 (reference include/kernel/qcu_gamma.cuh:47-51,101-106,159-163,208-212),
// projection scale (:61-67,115-123,172-180,222-228), reconstruction scale
// (:52-59,107-114,164-171,213-220), sign convention (:278-324).
__host__ __device__ constexpr int kMat2(int mu, int r) { return mu < 2 ? 3 - r : 2 + r; }
__host__ __device__ constexpr int kProjK(int mu, int r) {
    return mu == 0 ? 1 : mu == 1 ? (r == 0 ? 2 : 0) : mu == 2 ? (r == 0 ? 1 : 3) : 0;
}
__host__ __device__ constexpr int kReconK(int mu, int r) {
    return mu == 0 ? 3 : mu == 1 ? (r == 0 ? 2 : 0) : mu == 2 ? (r == 0 ? 3 : 1) : 0;
}

// Constants of one hop, selected at run time per lane.
struct HopInfo {
    int kp0, kp1;   // projection factors i^kp for rows 0 and 1
    int kr0, kr1;   // reconstruction factors
    bool m0_is_2;   // row 0 pairs with spin 2 (z, t) or spin 3 (x, y)
};

__device__ __forceinline__ HopInfo hop_info(int mu, bool fwd, bool dagger) {
    const int sgn = (fwd != dagger) ? 2 : 0;
    HopInfo h;
    h.kp0 = (kProjK(0, 0) * (mu == 0) + kProjK(1, 0) * (mu == 1) + kProjK(2, 0) * (mu == 2) +
             kProjK(3, 0) * (mu == 3) + sgn) & 3;
    h.kp1 = (kProjK(0, 1) * (mu == 0) + kProjK(1, 1) * (mu == 1) + kProjK(2, 1) * (mu == 2) +
             kProjK(3, 1) * (mu == 3) + sgn) & 3;
    h.kr0 = (kReconK(0, 0) * (mu == 0) + kReconK(1, 0) * (mu == 1) + kReconK(2, 0) * (mu == 2) +
             kReconK(3, 0) * (mu == 3) + sgn) & 3;
    h.kr1 = (kReconK(0, 1) * (mu == 0) + kReconK(1, 1) * (mu == 1) + kReconK(2, 1) * (mu == 2) +
             kReconK(3, 1) * (mu == 3) + sgn) & 3;
    h.m0_is_2 = mu >= 2;
    return h;
}

// Checkerboard coordinates of site `site` at parity `parity`.
struct Site {
    int x, y, z, t, xh, row;
};

__device__ __forceinline__ Site site_coords(int site, const Geom &g, int parity) {
    Site s;
    s.xh = site % g.Lxh;
    int rest = site / g.Lxh;
    s.y = rest % g.L[1];
    rest /= g.L[1];
    s.z = rest % g.L[2];
    s.t = rest / g.L[2];
    s.x = 2 * s.xh + ((s.y + s.z + s.t + parity) & 1);
    s.row = site - s.xh;
    return s;
}

// checkerboard index of the neighbour of `site` one step along mu (dir +1 forward)
__device__ __forceinline__ int neighbour(int site, const Site &s, const Geom &g, int mu, bool fwd) {
    switch (mu) {
        case 0: {
            const int x = fwd ? (s.x + 1 == g.L[0] ? 0 : s.x + 1) : (s.x == 0 ? g.L[0] - 1 : s.x - 1);
            return s.row + (x >> 1);
        }
        case 1: {
            const int st = g.Lxh;
            return fwd ? site + (s.y + 1 == g.L[1] ? -(g.L[1] - 1) * st : st)
                       : site + (s.y == 0 ? (g.L[1] - 1) * st : -st);
        }
        case 2: {
            const int st = g.Lxh * g.L[1];
            return fwd ? site + (s.z + 1 == g.L[2] ? -(g.L[2] - 1) * st : st)
                       : site + (s.z == 0 ? (g.L[2] - 1) * st : -st);
        }
        default: {
            const int st = g.Lxh * g.L[1] * g.L[2];
            return fwd ? site + (s.t + 1 == g.L[3] ? -(g.L[3] - 1) * st : st)
                       : site + (s.t == 0 ? (g.L[3] - 1) * st : -st);
        }
    }
}

// Branch-free neighbour for a per-lane (run-time) direction: the lane
// constants are computed once, the per-site work is selects only.
struct NbConst {
    int mu;
    int L;       // extent along mu
    int stride;  // checkerboard index stride along mu (mu > 0)
    bool fwd;
};

__device__ __forceinline__ NbConst nb_const(const Geom &g, int mu, bool fwd) {
    NbConst c;
    c.mu = mu;
    c.fwd = fwd;
    c.L = mu == 0 ? g.L[0] : mu == 1 ? g.L[1] : mu == 2 ? g.L[2] : g.L[3];
    c.stride = mu == 1 ? g.Lxh : mu == 2 ? g.Lxh * g.L[1] : g.Lxh * g.L[1] * g.L[2];
    return c;
}

__device__ __forceinline__ int neighbour_bl(int site, const Site &s, const Geom &g, const NbConst &c) {
    const int xn = c.fwd ? (s.x + 1 == g.L[0] ? 0 : s.x + 1) : (s.x == 0 ? g.L[0] - 1 : s.x - 1);
    const int nbx = s.row + (xn >> 1);
    const int coord = c.mu == 1 ? s.y : c.mu == 2 ? s.z : s.t;
    const int off = c.fwd ? (coord + 1 == c.L ? -(c.L - 1) * c.stride : c.stride)
                          : (coord == 0 ? (c.L - 1) * c.stride : -c.stride);
    return c.mu == 0 ? nbx : site + off;
}

// Lattice traversal: traversal index q -> checkerboard site.  The (y, z)
// plane is cut into By x Bz tiles; a "tile column" (all x, By y, Bz z) is
// swept along t, so the t-neighbours of a site are revisited one tile slice
// (Lx/2 * By * Bz sites) later and stay in the XCD's L2.  By = Ly, Bz = Lz is
// the plain lexicographic order.  Requires Ly % By == 0 and Lz % Bz == 0.
__device__ __forceinline__ int traverse(int q, const Geom &g, int By, int Bz) {
    const int slice = g.Lxh * By * Bz;
    const int colsz = slice * g.L[3];
    const int col = q / colsz;
    int r = q - col * colsz;
    const int t = r / slice;
    r -= t * slice;
    const int xh = r % g.Lxh;
    r /= g.Lxh;
    const int yi = r % By;
    const int zi = r / By;
    const int nyb = g.L[1] / By;
    const int yb = col % nyb, zb = col / nyb;
    const int y = yb * By + yi, z = zb * Bz + zi;
    return ((t * g.L[2] + z) * g.L[1] + y) * g.Lxh + xh;
}

}  // namespace qcu
