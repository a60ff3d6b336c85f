// lgnn.hip -- GPU kernels and training of the learned preconditioner (lgnn.h).
//
// Hop numbering: k = 2 mu + d (d = 0 forward, 1 backward) = the column of the
// neighbour table; the network's hop index is h = k + 1 (h = 0 is the
// identity).  Fields are RHS-innermost half vectors [Vh][spin 4][Nc][R];
// the eight hop fields of one half vector are stored back to back.
#include <cmath>
#include <cstring>
#include <type_traits>

#include "dslash_common.h"
#include "geom.h"
#include "lgnn.h"

namespace qcu {
namespace {

using C32 = Cx<float>;

// out_k(y) = V_k(y) in(nb_k(y)) for every site y of parity `pout` and hop k;
// V = U_mu(y) forward, U_mu(y - mu)^+ backward (link on the neighbour parity).
template <typename CS, int NC>
__global__ __launch_bounds__(256) void k_transport8(C32 *__restrict__ out, const CS *__restrict__ in,
                                                   const C32 *__restrict__ gauge, const int *__restrict__ nbr,
                                                   int Vh, int R, int pout) {
    const int site = blockIdx.x * blockDim.y + threadIdx.y;
    if (site >= Vh) return;
    const int r = threadIdx.x;
    const long sl = 4L * NC * R;
    const long fsize = long(Vh) * sl;
    const int *nrow = nbr + (long(pout) * Vh + site) * 8;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int mu = k >> 1;
        const bool fwd = (k & 1) == 0;
        const int nb = nrow[k];
        const C32 *U = fwd ? gauge + ((long(mu) * 2 + pout) * Vh + site) * NC * NC
                           : gauge + ((long(mu) * 2 + 1 - pout) * Vh + nb) * NC * NC;
        const CS *src = in + nb * sl + r;
        C32 *dst = out + k * fsize + site * sl + r;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            C32 v[NC];
#pragma unroll
            for (int b = 0; b < NC; ++b) v[b] = load_cx<float>(src + (s * NC + b) * R);
#pragma unroll
            for (int a = 0; a < NC; ++a) {
                C32 acc = {0.f, 0.f};
#pragma unroll
                for (int b = 0; b < NC; ++b) {
                    if (fwd) cfma(acc, U[a * NC + b], v[b]);
                    else cfma_conj(acc, U[b * NC + a], v[b]);
                }
                dst[(s * NC + a) * R] = acc;
            }
        }
    }
}

// q[(k,s)] = sum_{(k',s')} M[(k,s)][(k',s')] u[(k',s')] per (even site, colour, rhs)
__global__ __launch_bounds__(256) void k_mix(C32 *__restrict__ q, const C32 *__restrict__ u,
                                            const C32 *__restrict__ mcol, long ncol, int Nc, int R, long fsize) {
    __shared__ C32 M[32 * 32];
    for (int i = threadIdx.x; i < 1024; i += blockDim.x) M[i] = mcol[i];
    __syncthreads();
    const long stride = long(gridDim.x) * blockDim.x;
    for (long j = long(blockIdx.x) * blockDim.x + threadIdx.x; j < ncol; j += stride) {
        // j enumerates (site, colour, rhs) with rhs fastest
        const long r = j % R;
        const long rest = j / R;
        const long c = rest % Nc;
        const long site = rest / Nc;
        C32 x[32];
#pragma unroll
        for (int k = 0; k < 8; ++k)
#pragma unroll
            for (int s = 0; s < 4; ++s) x[k * 4 + s] = u[k * fsize + ((site * 4 + s) * Nc + c) * R + r];
#pragma unroll 4
        for (int i = 0; i < 32; ++i) {
            C32 acc = {0.f, 0.f};
#pragma unroll
            for (int jj = 0; jj < 32; ++jj) cfma(acc, M[i * 32 + jj], x[jj]);
            q[(i >> 2) * fsize + ((site * 4 + (i & 3)) * Nc + c) * R + r] = acc;
        }
    }
}

// out(x) = M0 in(x) + sum_k V_k(x) q_k(nb_k(x)) on the odd sites
template <typename CS, typename CO, int NC>
__global__ __launch_bounds__(256) void k_gather8(CO *__restrict__ out, const CS *__restrict__ in,
                                                const C32 *__restrict__ q, const C32 *__restrict__ gauge,
                                                const int *__restrict__ nbr, const C32 *__restrict__ m0, int Vh,
                                                int R) {
    const int site = blockIdx.x * blockDim.y + threadIdx.y;
    if (site >= Vh) return;
    const int r = threadIdx.x;
    const long sl = 4L * NC * R;
    const long fsize = long(Vh) * sl;
    const int pout = 1;
    C32 acc[4][NC];
    {
        C32 v[4][NC];
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int c = 0; c < NC; ++c) v[s][c] = load_cx<float>(in + site * sl + (s * NC + c) * R + r);
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                C32 a = {0.f, 0.f};
#pragma unroll
                for (int t = 0; t < 4; ++t) cfma(a, m0[s * 4 + t], v[t][c]);
                acc[s][c] = a;
            }
    }
    const int *nrow = nbr + (long(pout) * Vh + site) * 8;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int mu = k >> 1;
        const bool fwd = (k & 1) == 0;
        const int nb = nrow[k];
        const C32 *U = fwd ? gauge + ((long(mu) * 2 + pout) * Vh + site) * NC * NC
                           : gauge + ((long(mu) * 2 + 1 - pout) * Vh + nb) * NC * NC;
        const C32 *src = q + k * fsize + nb * sl + r;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            C32 v[NC];
#pragma unroll
            for (int b = 0; b < NC; ++b) v[b] = src[(s * NC + b) * R];
#pragma unroll
            for (int a = 0; a < NC; ++a) {
#pragma unroll
                for (int b = 0; b < NC; ++b) {
                    if (fwd) cfma(acc[s][a], U[a * NC + b], v[b]);
                    else cfma_conj(acc[s][a], U[b * NC + a], v[b]);
                }
            }
        }
    }
    CO *o = out + site * sl + r;
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int c = 0; c < NC; ++c) store_cx(o + (s * NC + c) * R, acc[s][c]);
}

// Partial sums of G[(k,s)][(k',s')] = sum_j a_{k^1}[s](j) conj(b_k'[s'](j)) over
// columns j = (even site, colour, rhs); a = transported dL/dphi*, b = u.
// Block partials in FP64; thread t owns entries 4t .. 4t+3.  Each block walks
// a contiguous range of sites; per site the (colour, rhs) columns are one
// contiguous run of Nc*R entries in every (field, spin) row.
constexpr int kGramTile = 32;
__global__ __launch_bounds__(256) void k_gram(C64 *__restrict__ part, const C32 *__restrict__ a,
                                             const C32 *__restrict__ b, int Vh, int ncr, long fsize) {
    __shared__ C32 sa[32][kGramTile];
    __shared__ C32 sb[32][kGramTile];
    const int t = threadIdx.x;
    const int i = t >> 3;            // row (k, s)
    const int j0 = (t & 7) * 4;      // first of 4 columns (k', s')
    C64 acc[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    const int per_block = (Vh + int(gridDim.x) - 1) / int(gridDim.x);
    const int lo = per_block * int(blockIdx.x);
    const int hi = lo + per_block < Vh ? lo + per_block : Vh;
    for (int site = lo; site < hi; ++site) {
        const long sbase = long(site) * 4 * ncr;
        for (int cr0 = 0; cr0 < ncr; cr0 += kGramTile) {
            // stage kGramTile columns of both 32-vectors
            for (int e = t; e < 32 * kGramTile; e += blockDim.x) {
                const int row = e / kGramTile, col = e % kGramTile;
                const int cr = cr0 + col;
                C32 va = {0.f, 0.f}, vb = {0.f, 0.f};
                if (cr < ncr) {
                    const int k = row >> 2, sp = row & 3;
                    const long off = sbase + long(sp) * ncr + cr;
                    va = a[(k ^ 1) * fsize + off];
                    vb = b[k * fsize + off];
                }
                sa[row][col] = va;
                sb[row][col] = vb;
            }
            __syncthreads();
#pragma unroll 4
            for (int col = 0; col < kGramTile; ++col) {
                const C32 x = sa[i][col];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const C32 y = sb[j0 + q][col];
                    acc[q].re += double(x.re) * y.re + double(x.im) * y.im;
                    acc[q].im += double(x.im) * y.re - double(x.re) * y.im;
                }
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) part[long(blockIdx.x) * 1040 + i * 32 + j0 + q] = acc[q];
}

// Partial sums of G0[s][t] = sum_j d[s](j) conj(p[t](j)) over odd-site columns.
__global__ __launch_bounds__(256) void k_gram0(C64 *__restrict__ part, const C32 *__restrict__ d,
                                              const C32 *__restrict__ p, long ncol, int Nc, int R) {
    __shared__ C64 red[16][256];
    C64 acc[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[e] = {0.0, 0.0};
    const long stride = long(gridDim.x) * blockDim.x;
    for (long j = long(blockIdx.x) * blockDim.x + threadIdx.x; j < ncol; j += stride) {
        const long r = j % R;
        const long rest = j / R;
        const long c = rest % Nc;
        const long site = rest / Nc;
        C32 dv[4], pv[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const long off = ((site * 4 + s) * Nc + c) * R + r;
            dv[s] = d[off];
            pv[s] = p[off];
        }
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int t2 = 0; t2 < 4; ++t2) {
                acc[s * 4 + t2].re += double(dv[s].re) * pv[t2].re + double(dv[s].im) * pv[t2].im;
                acc[s * 4 + t2].im += double(dv[s].im) * pv[t2].re - double(dv[s].re) * pv[t2].im;
            }
    }
#pragma unroll
    for (int e = 0; e < 16; ++e) red[e][threadIdx.x] = acc[e];
    __syncthreads();
    if (threadIdx.x < 16) {
        C64 s = {0.0, 0.0};
        for (int k = 0; k < int(blockDim.x); ++k) {
            s.re += red[threadIdx.x][k].re;
            s.im += red[threadIdx.x][k].im;
        }
        part[long(blockIdx.x) * 1040 + 1024 + threadIdx.x] = s;
    }
}

// Sum block partials (fixed order) and apply the chain rule to W1, W2:
//   dL/dM* = G, dL/dW2[o][h]* = sum_h' G[h][h'] W1[o][h']^+, dL/dW1[o][h']* = sum_h W2[o][h]^+ G[h][h']
// (h = 0 pairs with M0).  The stored gradient is 2 dL/dW* = dL/dRe W + i dL/dIm W,
// accumulated over batches.
// G[e] = sum over the block partials of entry e (fixed order: strided lanes, then a tree)
__global__ __launch_bounds__(256) void k_sum_partials(C64 *__restrict__ G, const C64 *__restrict__ part, int nblk) {
    __shared__ double sr[256], si[256];
    const int e = blockIdx.x;
    double ar = 0.0, ai = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
        ar += part[long(b) * 1040 + e].re;
        ai += part[long(b) * 1040 + e].im;
    }
    sr[threadIdx.x] = ar;
    si[threadIdx.x] = ai;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) {
            sr[threadIdx.x] += sr[threadIdx.x + w];
            si[threadIdx.x] += si[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) G[e] = {sr[0], si[0]};
}

__global__ void k_chain(C64 *g1, C64 *g2, const C64 *w1, const C64 *w2, const C64 *Gsum, int C) {
    __shared__ C64 G[1040];
    for (int e = threadIdx.x; e < 1040; e += blockDim.x) G[e] = Gsum[e];
    __syncthreads();
    // entry (o, h, s, t) of W: flattened ((o * 9 + h) * 4 + s) * 4 + t
    const int total = C * 9 * 16;
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        const int t = e & 3, s = (e >> 2) & 3, h = (e >> 4) % 9, o = e / 144;
        C64 a2 = {0.0, 0.0}, a1 = {0.0, 0.0};
        if (h == 0) {
            // W2[o][0]: sum_u G0[s][u] conj(W1[o][0][t][u]);  W1[o][0]: sum_u conj(W2[o][0][u][s]) G0[u][t]
            for (int u = 0; u < 4; ++u) {
                const C64 gsu = G[1024 + s * 4 + u];
                const C64 w1tu = w1[(o * 9) * 16 + t * 4 + u];
                a2.re += gsu.re * w1tu.re + gsu.im * w1tu.im;
                a2.im += gsu.im * w1tu.re - gsu.re * w1tu.im;
                const C64 w2us = w2[(o * 9) * 16 + u * 4 + s];
                const C64 gut = G[1024 + u * 4 + t];
                a1.re += w2us.re * gut.re + w2us.im * gut.im;
                a1.im += w2us.re * gut.im - w2us.im * gut.re;
            }
        } else {
            const int k = h - 1;
            for (int kp = 0; kp < 8; ++kp)
                for (int u = 0; u < 4; ++u) {
                    // W2[o][h][s][t] += G[(k,s)][(kp,u)] conj(W1[o][kp+1][t][u])
                    const C64 gg = G[(k * 4 + s) * 32 + kp * 4 + u];
                    const C64 w = w1[(o * 9 + kp + 1) * 16 + t * 4 + u];
                    a2.re += gg.re * w.re + gg.im * w.im;
                    a2.im += gg.im * w.re - gg.re * w.im;
                    // W1[o][h][s][t] += conj(W2[o][kp+1][u][s]) G[(kp,u)][(k,t)]
                    const C64 w2v = w2[(o * 9 + kp + 1) * 16 + u * 4 + s];
                    const C64 g2v = G[(kp * 4 + u) * 32 + k * 4 + t];
                    a1.re += w2v.re * g2v.re + w2v.im * g2v.im;
                    a1.im += w2v.re * g2v.im - w2v.im * g2v.re;
                }
        }
        g2[e].re += 2.0 * a2.re;
        g2[e].im += 2.0 * a2.im;
        g1[e].re += 2.0 * a1.re;
        g1[e].im += 2.0 * a1.im;
    }
}

// collapsed stencil: M0 = sum_o W2[o][0] W1[o][0], M[(k,s)][(k',t)] = sum_o (W2[o][k+1] W1[o][k'+1])[s][t]
__global__ void k_collapse(Cx<float> *mcol, const C64 *w1, const C64 *w2, int C) {
    for (int e = threadIdx.x; e < 1040; e += blockDim.x) {
        int hA, hB, s, t;
        if (e < 1024) {
            const int i = e / 32, j = e % 32;
            hA = (i >> 2) + 1;
            s = i & 3;
            hB = (j >> 2) + 1;
            t = j & 3;
        } else {
            hA = hB = 0;
            s = (e - 1024) >> 2;
            t = (e - 1024) & 3;
        }
        C64 acc = {0.0, 0.0};
        for (int o = 0; o < C; ++o)
            for (int u = 0; u < 4; ++u) acc = acc + w2[(o * 9 + hA) * 16 + s * 4 + u] * w1[(o * 9 + hB) * 16 + u * 4 + t];
        mcol[e] = {float(acc.re), float(acc.im)};
    }
}

// Stencil of N^+ from that of N: with N = M0 + sum_{k,k'} T_k M[k][k'] T_k' and T_k^+ = T_{k^1},
// N^+ = M0^+ + sum_{a,b} T_a (M[b^1][a^1])^+ T_b, so block (a, b) of the adjoint stencil is the
// conjugate transpose of block (b^1, a^1).
__global__ void k_adjoint_stencil(Cx<float> *adj, const Cx<float> *m) {
    for (int e = threadIdx.x; e < 1040; e += blockDim.x) {
        Cx<float> v;
        if (e < 1024) {
            const int i = e / 32, j = e % 32;  // row (a, s), column (b, t)
            const int a = i >> 2, sp = i & 3, b = j >> 2, t = j & 3;
            v = m[((b ^ 1) * 4 + t) * 32 + (a ^ 1) * 4 + sp];
        } else {
            const int sp = (e - 1024) >> 2, t = (e - 1024) & 3;
            v = m[1024 + t * 4 + sp];
        }
        adj[e] = {v.re, -v.im};
    }
}

// W2[o][h] = P_h W2[o][h] for hops h = 1 + k, k = 2 mu + (0 forward | 1 backward), with P_h the
// orthogonal projector (1 -+ gamma_mu) / 2 onto the spinors whose spin m(r) is i^kr x spin r (r = 0, 1;
// m and kr from the dslash tables of the hop without dagger): P[r][r] = P[m][m] = 1/2, P[m][r] = i^kr / 2,
// P[r][m] = i^-kr / 2.  One thread per (o, h, column).
__global__ void k_project_w2(C64 *w2, int C) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C * 8 * 4) return;
    const int col = i & 3, k = (i >> 2) & 7, o = i >> 5;
    const int mu = k >> 1;
    const bool fwd = (k & 1) == 0;
    const int sgn = fwd ? 2 : 0;
    C64 *w = w2 + (size_t(o) * 9 + 1 + k) * 16;  // [4][4] row-major: w[s * 4 + col]
    C64 v[4];
    for (int s = 0; s < 4; ++s) v[s] = w[s * 4 + col];
    for (int r = 0; r < 2; ++r) {
        const int m = kMat2(mu, r), kr = (kReconK(mu, r) + sgn) & 3;
        // i^kr as (re, im)
        const double pr = kr == 0 ? 1.0 : kr == 2 ? -1.0 : 0.0, pi = kr == 1 ? 1.0 : kr == 3 ? -1.0 : 0.0;
        const C64 a = v[r], b = v[m];
        // (P v)_r = (a + i^-kr b) / 2, (P v)_m = (i^kr a + b) / 2
        const C64 up = {0.5 * (a.re + pr * b.re + pi * b.im), 0.5 * (a.im + pr * b.im - pi * b.re)};
        const C64 lo = {0.5 * (pr * a.re - pi * a.im + b.re), 0.5 * (pr * a.im + pi * a.re + b.im)};
        w[r * 4 + col] = up;
        w[m * 4 + col] = lo;
    }
}

__global__ void k_adam(C64 *w, C64 *m, C64 *v, const C64 *g, int n, double lr, double b1, double b2, double eps,
                       double c1, double c2, double inv_batches) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double gr = g[i].re * inv_batches, gi = g[i].im * inv_batches;
    m[i].re = b1 * m[i].re + (1 - b1) * gr;
    m[i].im = b1 * m[i].im + (1 - b1) * gi;
    v[i].re = b2 * v[i].re + (1 - b2) * gr * gr;
    v[i].im = b2 * v[i].im + (1 - b2) * gi * gi;
    w[i].re -= lr * (m[i].re / c1) / (sqrt(v[i].re / c2) + eps);
    w[i].im -= lr * (m[i].im / c1) / (sqrt(v[i].im / c2) + eps);
}

__global__ __launch_bounds__(256) void k_diff_scale(C32 *__restrict__ d, const C32 *__restrict__ phi,
                                                   const C32 *__restrict__ tgt, const double *__restrict__ inv_n,
                                                   long N, int R) {
    const long stride = long(gridDim.x) * blockDim.x;
    for (long e = long(blockIdx.x) * blockDim.x + threadIdx.x; e < N; e += stride) {
        const float sc = float(inv_n[e % R]);
        const C32 a = phi[e], b = tgt[e];
        d[e] = {(a.re - b.re) * sc, (a.im - b.im) * sc};
    }
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void k_gaussian(C64 *__restrict__ out, long count, uint64_t key) {
    const long stride = long(gridDim.x) * blockDim.x;
    for (long i = long(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) {
        const uint64_t z = splitmix64(key ^ (uint64_t(i) * 0xD1B54A32D192ED03ull));
        const double u1 = (double((z >> 11) & ((1ull << 26) - 1)) + 1.0) * (1.0 / 67108864.0);  // (0, 1]
        const double u2 = double(z >> 38) * (1.0 / 67108864.0);                                  // [0, 1)
        const double rad = sqrt(-2.0 * log(u1));
        double sn, cs;
        sincospi(2.0 * u2, &sn, &cs);
        out[i] = {rad * cs, rad * sn};
    }
}

template <typename F>
hipError_t dispatch_nc(int Nc, F &&f) {
    switch (Nc) {
        case 2: return f(std::integral_constant<int, 2>{});
        case 3: return f(std::integral_constant<int, 3>{});
        case 4: return f(std::integral_constant<int, 4>{});
        case 8: return f(std::integral_constant<int, 8>{});
        default: return hipErrorInvalidValue;
    }
}

void *dmal(size_t bytes, hipError_t &err) {
    void *p = nullptr;
    if (err == hipSuccess) err = hipMalloc(&p, bytes);
    return p;
}

}  // namespace

hipError_t launch_gaussian(Cx<double> *out, long count, uint64_t key, hipStream_t s) {
    const long nb = std::min<long>((count + 255) / 256, 8192);
    hipLaunchKernelGGL(k_gaussian, dim3(int(nb > 0 ? nb : 1)), dim3(256), 0, s, out, count, key);
    return hipGetLastError();
}

LgnnPreconditioner::LgnnPreconditioner(const Geom &g, int Nc, int R, int channels)
    : g_(g), Nc_(Nc), R_(R), C_(channels), n_(long(g.Vh) * 4 * Nc) {}

LgnnPreconditioner::~LgnnPreconditioner() {
    for (void *p : {(void *)w1_, (void *)w2_, (void *)g1_, (void *)g2_, (void *)m1_, (void *)m2_, (void *)v1_,
                    (void *)v2_, (void *)mcol_, (void *)madj_, u_, q_, inf_, phi_, dlt_, (void *)gm_, (void *)dnorm_,
                    pg_, pu_, pn_, ptmp_, (void *)pscale_,
                    (void *)dloss_, gauge32_, q16_, glo_[0], glo_[1]})
        if (p) (void)hipFree(p);
    for (void *p : pw_)
        if (p) (void)hipFree(p);
    for (void *p : phops_)
        if (p) (void)hipFree(p);
    if (hbuf_) (void)hipHostFree(hbuf_);
    reduce_work_free(red_);
}

bool LgnnPreconditioner::supports(int prec) const {
    return prec == kDouble || prec == kSingle || prec == kHalf || prec == kBF16;
}

hipError_t LgnnPreconditioner::ensure_work32(hipStream_t) {
    if (u_) return hipSuccess;
    hipError_t err = hipSuccess;
    const size_t half32 = size_t(n_) * R_ * sizeof(C32);
    u_ = dmal(8 * half32, err);
    q_ = dmal(8 * half32, err);
    inf_ = dmal(half32, err);
    phi_ = dmal(half32, err);
    dlt_ = dmal(half32, err);
    gram_blocks_ = 512;
    gm_ = static_cast<C64 *>(dmal(sizeof(C64) * 1040 * size_t(gram_blocks_ + 1), err));
    if (err != hipSuccess) set_error(5, "learned preconditioner: allocation failed (%s)", hipGetErrorString(err));
    return err;
}

hipError_t LgnnPreconditioner::ensure_fields(hipStream_t s) {
    if (w1_) return hipSuccess;
    hipError_t err = hipSuccess;
    const size_t wbytes = size_t(C_) * 9 * 16 * sizeof(C64);
    mcol_ = static_cast<Cx<float> *>(dmal(sizeof(Cx<float>) * 1040, err));
    madj_ = static_cast<Cx<float> *>(dmal(sizeof(Cx<float>) * 1040, err));
    w1_ = static_cast<C64 *>(dmal(wbytes, err));
    w2_ = static_cast<C64 *>(dmal(wbytes, err));
    g1_ = static_cast<C64 *>(dmal(wbytes, err));
    g2_ = static_cast<C64 *>(dmal(wbytes, err));
    m1_ = static_cast<C64 *>(dmal(wbytes, err));
    m2_ = static_cast<C64 *>(dmal(wbytes, err));
    v1_ = static_cast<C64 *>(dmal(wbytes, err));
    v2_ = static_cast<C64 *>(dmal(wbytes, err));
    dnorm_ = static_cast<double *>(dmal(sizeof(double) * 2 * R_, err));
    dloss_ = static_cast<double *>(dmal(sizeof(double) * 2 * R_, err));
    if (err != hipSuccess) {
        set_error(5, "learned preconditioner: allocation failed (%s)", hipGetErrorString(err));
        return err;
    }
    QCU_HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&hbuf_), sizeof(double) * 4 * R_));
    QCU_HIP_TRY(reduce_work_alloc(red_, R_));
    for (C64 *p : {g1_, g2_, m1_, m2_, v1_, v2_}) QCU_HIP_TRY(hipMemsetAsync(p, 0, wbytes, s));
    return hipSuccess;
}

hipError_t LgnnPreconditioner::set_gauge(const void *gauge64, hipStream_t s, long long gen) {
    if (gen >= 0 && gauge64 == gauge_ && gen == gauge_gen_ && gauge32_) return hipSuccess;  // copies current
    gauge_ = gauge64;
    gauge_gen_ = -1;
    glo_valid_[0] = glo_valid_[1] = false;
    const long count = 8L * g_.Vh * Nc_ * Nc_;
    if (!gauge32_) QCU_HIP_TRY(hipMalloc(&gauge32_, sizeof(C32) * count));
    QCU_HIP_TRY(launch_convert(gauge32_, kSingle, gauge64, kDouble, count, nullptr, 1, s));
    gauge_gen_ = gen;
    return hipSuccess;
}

hipError_t LgnnPreconditioner::set_weights(const double *w1, const double *w2, hipStream_t s) {
    QCU_HIP_TRY(ensure_fields(s));
    const size_t wbytes = size_t(C_) * 9 * 16 * sizeof(C64);
    QCU_HIP_TRY(hipMemcpyAsync(w1_, w1, wbytes, hipMemcpyHostToDevice, s));
    QCU_HIP_TRY(hipMemcpyAsync(w2_, w2, wbytes, hipMemcpyHostToDevice, s));
    for (C64 *p : {g1_, g2_, m1_, m2_, v1_, v2_}) QCU_HIP_TRY(hipMemsetAsync(p, 0, wbytes, s));
    adam_t_ = 0;
    have_weights_ = true;
    if (projected_) QCU_HIP_TRY(project_w2(s));
    QCU_HIP_TRY(collapse(s));
    return hipStreamSynchronize(s);
}

hipError_t LgnnPreconditioner::project_w2(hipStream_t s) {
    hipLaunchKernelGGL(k_project_w2, dim3(div_up(C_ * 32, 256)), dim3(256), 0, s, w2_, C_);
    return hipGetLastError();
}

hipError_t LgnnPreconditioner::set_projected(bool on, hipStream_t s) {
    projected_ = on;
    if (!on || !have_weights_) return hipSuccess;
    QCU_HIP_TRY(project_w2(s));
    QCU_HIP_TRY(collapse(s));
    return hipStreamSynchronize(s);
}

hipError_t LgnnPreconditioner::get_weights(double *w1, double *w2, hipStream_t s) {
    if (!have_weights_) return hipErrorInvalidValue;
    const size_t wbytes = size_t(C_) * 9 * 16 * sizeof(C64);
    QCU_HIP_TRY(hipMemcpyAsync(w1, w1_, wbytes, hipMemcpyDeviceToHost, s));
    QCU_HIP_TRY(hipMemcpyAsync(w2, w2_, wbytes, hipMemcpyDeviceToHost, s));
    return hipStreamSynchronize(s);
}

hipError_t LgnnPreconditioner::collapse(hipStream_t s) {
    hipLaunchKernelGGL(k_collapse, dim3(1), dim3(256), 0, s, mcol_, w1_, w2_, C_);
    QCU_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_adjoint_stencil, dim3(1), dim3(256), 0, s, madj_, mcol_);
    return hipGetLastError();
}

hipError_t LgnnPreconditioner::forward(const void *in, int prec, void *out, int out_prec, hipStream_t s,
                                       const Cx<float> *stencil, void *hops) {
    const Cx<float> *mc = stencil ? stencil : mcol_;
    QCU_HIP_TRY(ensure_work32(s));
    if (!in || !out || !u_ || !q_) return hipErrorInvalidValue;
    GeomTables tab;
    QCU_HIP_TRY(geom_tables(g_, g_.L[1], g_.L[2], tab, s));
    const int spb = R_ >= 256 ? 1 : 256 / R_;
    const dim3 block(R_, spb), grid(div_up(g_.Vh, spb));
    const long fsize = n_ * R_;
    C32 *u = static_cast<C32 *>(hops ? hops : u_), *q = static_cast<C32 *>(q_);
    const C32 *g32 = static_cast<const C32 *>(gauge32_);
    hipError_t e = dispatch_nc(Nc_, [&](auto ncv) -> hipError_t {
        constexpr int NC = decltype(ncv)::value;
        switch (prec) {
            case kDouble: hipLaunchKernelGGL((k_transport8<Cx<double>, NC>), grid, block, 0, s, u, static_cast<const Cx<double> *>(in), g32, tab.nbr, g_.Vh, R_, 0); break;
            case kSingle: hipLaunchKernelGGL((k_transport8<Cx<float>, NC>), grid, block, 0, s, u, static_cast<const Cx<float> *>(in), g32, tab.nbr, g_.Vh, R_, 0); break;
            case kHalf: hipLaunchKernelGGL((k_transport8<Half2Store, NC>), grid, block, 0, s, u, static_cast<const Half2Store *>(in), g32, tab.nbr, g_.Vh, R_, 0); break;
            case kBF16: hipLaunchKernelGGL((k_transport8<Bf162Store, NC>), grid, block, 0, s, u, static_cast<const Bf162Store *>(in), g32, tab.nbr, g_.Vh, R_, 0); break;
            default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    });
    if (e != hipSuccess) return e;
    const long ncol = long(g_.Vh) * Nc_ * R_;
    hipLaunchKernelGGL(k_mix, dim3(int(std::min<long>((ncol + 255) / 256, 8192))), dim3(256), 0, s, q, u, mc, ncol,
                       Nc_, R_, fsize);
    QCU_HIP_TRY(hipGetLastError());
    const C32 *m0 = mc + 1024;
    return dispatch_nc(Nc_, [&](auto ncv) -> hipError_t {
        constexpr int NC = decltype(ncv)::value;
#define QCU_G8(CS, CO) hipLaunchKernelGGL((k_gather8<CS, CO, NC>), grid, block, 0, s, static_cast<CO *>(out), static_cast<const CS *>(in), q, g32, tab.nbr, m0, g_.Vh, R_)
        if (out_prec == kSingle) {
            switch (prec) {
                case kDouble: QCU_G8(Cx<double>, Cx<float>); break;
                case kSingle: QCU_G8(Cx<float>, Cx<float>); break;
                case kHalf: QCU_G8(Half2Store, Cx<float>); break;
                case kBF16: QCU_G8(Bf162Store, Cx<float>); break;
                default: return hipErrorInvalidValue;
            }
        } else {
            if (prec != out_prec) return hipErrorInvalidValue;
            switch (prec) {
                case kDouble: QCU_G8(Cx<double>, Cx<double>); break;
                case kHalf: QCU_G8(Half2Store, Half2Store); break;
                case kBF16: QCU_G8(Bf162Store, Bf162Store); break;
                default: return hipErrorInvalidValue;
            }
        }
#undef QCU_G8
        return hipGetLastError();
    });
}

hipError_t LgnnPreconditioner::apply(void *out, const void *in, int prec, hipStream_t s) {
    if (!ready()) {
        set_error(1, "learned preconditioner has no weights or gauge field");
        return hipErrorInvalidValue;
    }
    QCU_HIP_TRY(ensure_fields(s));
    if ((prec == kHalf || prec == kBF16) && Nc_ <= 4) return apply_mfma(out, in, prec, s);
    return forward(in, prec, out, prec, s);
}

hipError_t LgnnPreconditioner::apply_mfma(void *out, const void *in, int prec, hipStream_t s) {
    const int slot = prec == kHalf ? 0 : 1;
    const long gcount = 8L * g_.Vh * Nc_ * Nc_;
    const size_t vec = size_t(n_) * R_ * 4;  // one 16-bit complex per word
    if (double(vec) * 8.0 >= 2147483647.0 * 4.0 || double(gcount) * 4.0 >= 2147483647.0) return hipErrorInvalidValue;
    if (!glo_[slot]) QCU_HIP_TRY(hipMalloc(&glo_[slot], size_t(gcount) * 4));
    if (!glo_valid_[slot]) {
        QCU_HIP_TRY(launch_convert(glo_[slot], prec, gauge_, kDouble, gcount, nullptr, 1, s));
        glo_valid_[slot] = true;
    }
    if (!q16_) QCU_HIP_TRY(hipMalloc(&q16_, 8 * vec + kLgnnFragBytes));
    GeomTables tab;
    QCU_HIP_TRY(geom_tables(g_, g_.L[1], g_.L[2], tab, s));
    return launch_lgnn_mfma(out, in, glo_[slot], q16_, mcol_, tab.walk, g_, Nc_, R_, prec, projected_, s);
}

hipError_t LgnnPreconditioner::zero_grad(hipStream_t s) {
    QCU_HIP_TRY(ensure_fields(s));
    const size_t wbytes = size_t(C_) * 9 * 16 * sizeof(C64);
    QCU_HIP_TRY(hipMemsetAsync(g1_, 0, wbytes, s));
    return hipMemsetAsync(g2_, 0, wbytes, s);
}

hipError_t LgnnPreconditioner::loss_grad(const void *in, const void *target, double *loss_host, hipStream_t s) {
    if (!ready()) return hipErrorInvalidValue;
    QCU_HIP_TRY(ensure_fields(s));
    // the FP32 work fields (phi_, dlt_, u_, q_, gm_) must exist before their pointers are passed on
    QCU_HIP_TRY(ensure_work32(s));
    const long N = n_ * R_;
    // forward in FP32 (keeps u for the gradient)
    QCU_HIP_TRY(forward(in, kSingle, phi_, kSingle, s));
    // per-RHS ||e||^2, then d = (phi - e) / ||e||^2 and loss = sum ||phi - e||^2 / ||e||^2
    QCU_HIP_TRY(launch_norm(dnorm_, target, n_, R_, kSingle, false, red_, 0, s));
    QCU_HIP_TRY(hipMemcpyAsync(hbuf_, dnorm_, sizeof(double) * R_, hipMemcpyDeviceToHost, s));
    QCU_HIP_TRY(hipStreamSynchronize(s));
    for (int r = 0; r < R_; ++r) hbuf_[R_ + r] = hbuf_[r] > 0 ? 1.0 / hbuf_[r] : 0.0;
    QCU_HIP_TRY(hipMemcpyAsync(dnorm_ + R_, hbuf_ + R_, sizeof(double) * R_, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_diff_scale, dim3(2048), dim3(256), 0, s, static_cast<C32 *>(dlt_),
                       static_cast<const C32 *>(phi_), static_cast<const C32 *>(target), dnorm_ + R_, N, R_);
    QCU_HIP_TRY(hipGetLastError());
    if (loss_host) {
        // ||d||^2 * ||e||^2 = ||phi - e||^2 / ||e||^2
        QCU_HIP_TRY(launch_norm(dloss_, dlt_, n_, R_, kSingle, false, red_, 1, s));
        QCU_HIP_TRY(hipMemcpyAsync(hbuf_ + 2 * R_, dloss_, sizeof(double) * R_, hipMemcpyDeviceToHost, s));
        QCU_HIP_TRY(hipStreamSynchronize(s));
        double L = 0.0;
        for (int r = 0; r < R_; ++r) L += hbuf_[2 * R_ + r] * hbuf_[r];
        *loss_host = L;
    }
    return accumulate_grad(in, dlt_, s);
}

hipError_t LgnnPreconditioner::accumulate_grad(const void *w, const void *dphi, hipStream_t s, const void *hops) {
    // a_k = T_k dphi on the even sites (the adjoint of hop k is hop k ^ 1), into q
    GeomTables tab;
    QCU_HIP_TRY(geom_tables(g_, g_.L[1], g_.L[2], tab, s));
    const int spb = R_ >= 256 ? 1 : 256 / R_;
    const dim3 block(R_, spb), grid(div_up(g_.Vh, spb));
    QCU_HIP_TRY(dispatch_nc(Nc_, [&](auto ncv) -> hipError_t {
        constexpr int NC = decltype(ncv)::value;
        hipLaunchKernelGGL((k_transport8<Cx<float>, NC>), grid, block, 0, s, static_cast<C32 *>(q_),
                           static_cast<const C32 *>(dphi), static_cast<const C32 *>(gauge32_), tab.nbr, g_.Vh, R_, 0);
        return hipGetLastError();
    }));
    const long ncol = long(g_.Vh) * Nc_ * R_;
    const long fsize = n_ * R_;
    hipLaunchKernelGGL(k_gram, dim3(gram_blocks_), dim3(256), 0, s, gm_, static_cast<const C32 *>(q_),
                       static_cast<const C32 *>(hops ? hops : u_), g_.Vh, Nc_ * R_, fsize);
    QCU_HIP_TRY(hipGetLastError());
    // the input itself is needed in FP32 for the M0 gradient
    const C32 *inf = static_cast<const C32 *>(w);
    hipLaunchKernelGGL(k_gram0, dim3(gram_blocks_), dim3(256), 0, s, gm_, static_cast<const C32 *>(dphi), inf, ncol,
                       Nc_, R_);
    QCU_HIP_TRY(hipGetLastError());
    // partial sums land after the per-block partials
    C64 *gsum = gm_ + 1040 * size_t(gram_blocks_);
    hipLaunchKernelGGL(k_sum_partials, dim3(1040), dim3(256), 0, s, gsum, gm_, gram_blocks_);
    QCU_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_chain, dim3(1), dim3(256), 0, s, g1_, g2_, w1_, w2_, gsum, C_);
    return hipGetLastError();
}

hipError_t LgnnPreconditioner::schur32(void *out, const void *in, double kappa, int dagger, hipStream_t s) {
    DslashArgs a;
    a.out = ptmp_;
    a.in = in;
    a.gauge = gauge32_;
    a.xpay = nullptr;
    a.coef = 1.0;
    a.g = g_;
    a.Nc = Nc_;
    a.R = R_;
    a.parity = 0;
    a.dagger = dagger;
    a.prec = kSingle;
    QCU_HIP_TRY(launch_dslash(a, s));
    a.out = out;
    a.in = ptmp_;
    a.xpay = in;
    a.coef = -kappa * kappa;
    a.parity = 1;
    return launch_dslash(a, s);
}

hipError_t LgnnPreconditioner::ensure_power(int K, hipStream_t s) {
    hipError_t err = hipSuccess;
    const size_t half32 = size_t(n_) * R_ * sizeof(C32);
    while (int(pw_.size()) < K + 1) {
        pw_.push_back(dmal(half32, err));
        if (err != hipSuccess) return err;
    }
    while (int(phops_.size()) < K) {  // the 8 hop fields of each factor's input, kept for the gradient
        phops_.push_back(dmal(8 * half32, err));
        if (err != hipSuccess) return err;
    }
    if (pg_) return hipSuccess;
    pg_ = dmal(half32, err);
    pu_ = dmal(half32, err);
    pn_ = dmal(half32, err);
    ptmp_ = dmal(half32, err);
    // [0, R): per-RHS scale, [R, 2R): -1, [2R, 4R): complex ones for the axpy kernels
    pscale_ = static_cast<double *>(dmal(sizeof(double) * 4 * R_, err));
    if (err != hipSuccess) return err;
    for (int r = 0; r < R_; ++r) {
        hbuf_[r] = -1.0;
        hbuf_[R_ + 2 * r] = 1.0;
        hbuf_[R_ + 2 * r + 1] = 0.0;
    }
    QCU_HIP_TRY(hipMemcpyAsync(pscale_ + R_, hbuf_, sizeof(double) * 3 * R_, hipMemcpyHostToDevice, s));
    return hipStreamSynchronize(s);
}

hipError_t LgnnPreconditioner::power_loss_grad(const void *v, int K, double kappa, double *loss_host,
                                               hipStream_t s) {
    if (!ready() || K < 1 || K > 8) return hipErrorInvalidValue;
    QCU_HIP_TRY(ensure_fields(s));
    QCU_HIP_TRY(ensure_work32(s));
    QCU_HIP_TRY(ensure_power(K, s));
    const long N = n_ * R_;
    const size_t half32 = size_t(N) * sizeof(C32);
    const C64 *ones = reinterpret_cast<const C64 *>(pscale_ + 2 * R_);
    // w_0 = v / ||v|| per RHS
    QCU_HIP_TRY(launch_norm(dnorm_, v, n_, R_, kSingle, true, red_, 0, s));
    QCU_HIP_TRY(hipMemcpyAsync(hbuf_, dnorm_, sizeof(double) * R_, hipMemcpyDeviceToHost, s));
    QCU_HIP_TRY(hipStreamSynchronize(s));
    for (int r = 0; r < R_; ++r) hbuf_[r] = hbuf_[r] > 0 ? 1.0 / hbuf_[r] : 0.0;
    QCU_HIP_TRY(hipMemcpyAsync(pscale_, hbuf_, sizeof(double) * R_, hipMemcpyHostToDevice, s));
    QCU_HIP_TRY(launch_convert(pw_[0], kSingle, v, kSingle, N, pscale_, R_, s));
    // w_{j+1} = w_j - A N w_j
    for (int j = 0; j < K; ++j) {
        QCU_HIP_TRY(forward(pw_[j], kSingle, phi_, kSingle, s, nullptr, phops_[j]));
        QCU_HIP_TRY(schur32(pu_, phi_, kappa, 0, s));
        AxpyArgs ax{kXsay, pw_[j + 1], pw_[j], pu_, nullptr, ones, nullptr, nullptr, n_, R_, kSingle};
        QCU_HIP_TRY(launch_axpy(ax, s));
    }
    if (loss_host) {
        QCU_HIP_TRY(launch_norm(dloss_, pw_[K], n_, R_, kSingle, false, red_, 1, s));
        QCU_HIP_TRY(hipMemcpyAsync(hbuf_ + 2 * R_, dloss_, sizeof(double) * R_, hipMemcpyDeviceToHost, s));
        QCU_HIP_TRY(hipStreamSynchronize(s));
        double L = 0.0;
        for (int r = 0; r < R_; ++r) L += hbuf_[2 * R_ + r];
        *loss_host = L;
    }
    // back-propagation: g_K = w_K; dL/d(N w_j)^* = -A^+ g_{j+1}; g_j = g_{j+1} - N^+ A^+ g_{j+1}
    QCU_HIP_TRY(hipMemcpyAsync(pg_, pw_[K], half32, hipMemcpyDeviceToDevice, s));
    for (int j = K - 1; j >= 0; --j) {
        QCU_HIP_TRY(schur32(pu_, pg_, kappa, 1, s));
        if (j > 0) QCU_HIP_TRY(forward(pu_, kSingle, pn_, kSingle, s, madj_));
        QCU_HIP_TRY(launch_convert(dlt_, kSingle, pu_, kSingle, N, pscale_ + R_, R_, s));
        QCU_HIP_TRY(accumulate_grad(pw_[j], dlt_, s, phops_[j]));
        if (j > 0) {
            AxpyArgs ax{kXsay, pg_, pg_, pn_, nullptr, ones, nullptr, nullptr, n_, R_, kSingle};
            QCU_HIP_TRY(launch_axpy(ax, s));
        }
    }
    return hipSuccess;
}

hipError_t LgnnPreconditioner::read_grad(double *gw1, double *gw2, hipStream_t s) {
    const size_t wbytes = size_t(C_) * 9 * 16 * sizeof(C64);
    QCU_HIP_TRY(hipMemcpyAsync(gw1, g1_, wbytes, hipMemcpyDeviceToHost, s));
    QCU_HIP_TRY(hipMemcpyAsync(gw2, g2_, wbytes, hipMemcpyDeviceToHost, s));
    return hipStreamSynchronize(s);
}

hipError_t LgnnPreconditioner::adam_step(double lr, int batches, hipStream_t s) {
    const double b1 = 0.9, b2 = 0.999, eps = 1e-12;
    ++adam_t_;
    const double c1 = 1.0 - std::pow(b1, adam_t_), c2 = 1.0 - std::pow(b2, adam_t_);
    const int n = C_ * 9 * 16;
    const double inv = batches > 0 ? 1.0 / batches : 1.0;
    hipLaunchKernelGGL(k_adam, dim3(div_up(n, 256)), dim3(256), 0, s, w1_, m1_, v1_, g1_, n, lr, b1, b2, eps, c1, c2, inv);
    hipLaunchKernelGGL(k_adam, dim3(div_up(n, 256)), dim3(256), 0, s, w2_, m2_, v2_, g2_, n, lr, b1, b2, eps, c1, c2, inv);
    QCU_HIP_TRY(hipGetLastError());
    if (projected_) QCU_HIP_TRY(project_w2(s));  // back into the projected family after the step
    return collapse(s);
}

}  // namespace qcu
