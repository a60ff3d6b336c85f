// lgnn_mfma.hip -- FP16 / BF16 application of the learned preconditioner on
// matrix cores (declared in lgnn.h).
//
// Network (collapsed, odd-parity input psi):
//     out(x) = M0 psi(x) + sum_k (T_k q_k)(x),   q_k(y) = sum_k' M[k][k'] (T_k' psi)(y)
// with hops k = 2 mu + (0 forward | 1 backward), T_k f(y) = V_k(y) f(nb_k(y)),
// V = U_mu(y) forward, U_mu(y - mu)^+ backward, M a 32 x 32 complex matrix over
// (hop, spin) and M0 a 4 x 4 spin matrix.
//
// k_lgnn_mix: one wavefront per (even site, 16-RHS chunk).  Lane (m, g)
// (m = RHS, g = lane group) transports the 4 spins of hop k' = 4 st + g on
// VALU (v_dot2 on 16-bit pairs, FP32 accumulation) and packs them as the MFMA B operand
// (k index = (k', spin, re|im)); the real 2x2 embedding of M is the A operand
// (rows = (k, s, re|im)), kept in registers for the whole wavefront.  8 MFMAs
// (4 row blocks x 2 k steps) per colour give all 32 q entries of that colour.
// k_lgnn_gather: one wavefront per (odd site, 16-RHS chunk), the dslash's MFMA
// structure without spin projection: links as the A operand, the 4 spins of
// q_k(nb_k) as B, one MFMA chain per spin; M0 psi(x) is added in the epilogue.
// Both kernels read a walk-table row (site + 8 neighbours) per site.
//
// Projected form (PROJ, for a network whose second-layer weights of hop k lie
// in the image of the Wilson projector P_k = (1 -+ gamma_mu) / 2 of that hop,
// so that M[k][.] = P_k M[k][.] and every q_k satisfies the reconstruction
// relation of the dslash: spin m(r) = i^kr(k) x spin r, r = 0, 1): the mix
// computes and stores only spins 0 and 1 of each q_k (two MFMA row blocks
// instead of four, fields of 2 spins), and the gather rebuilds spins 2 and 3
// from them with the hop's phases before its MFMAs -- half the field traffic.
#include <cstdint>

#include "dslash_common.h"
#include "lgnn.h"

namespace qcu {
namespace {

typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <typename S> struct W;
template <> struct W<__half> {
    __device__ __forceinline__ static float re(uint32_t w) { return float(__builtin_bit_cast(f16x2, w)[0]); }
    __device__ __forceinline__ static float im(uint32_t w) { return float(__builtin_bit_cast(f16x2, w)[1]); }
    __device__ __forceinline__ static uint32_t pack(float r, float i) {
        const f16x2 v = {_Float16(r), _Float16(i)};
        return __builtin_bit_cast(uint32_t, v);
    }
    __device__ __forceinline__ static f32x4 mma(u32x4 a, u32x4 b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b),
                                                      c, 0, 0, 0);
    }
    // c + a.lo * b.lo + a.hi * b.hi (FP32 accumulation)
    __device__ __forceinline__ static float dot(uint32_t a, uint32_t b, float c) {
        return __builtin_amdgcn_fdot2(__builtin_bit_cast(f16x2, a), __builtin_bit_cast(f16x2, b), c, false);
    }
};
template <> struct W<__hip_bfloat16> {
    __device__ __forceinline__ static float re(uint32_t w) { return __builtin_bit_cast(float, w << 16); }
    __device__ __forceinline__ static float im(uint32_t w) { return __builtin_bit_cast(float, w & 0xffff0000u); }
    __device__ __forceinline__ static uint32_t pack(float r, float i) {
        const bf16x2 v = {__bf16(r), __bf16(i)};
        return __builtin_bit_cast(uint32_t, v);
    }
    __device__ __forceinline__ static f32x4 mma(u32x4 a, u32x4 b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                       __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
    }
    __device__ __forceinline__ static float dot(uint32_t a, uint32_t b, float c) {
        return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, a), __builtin_bit_cast(bf16x2, b), c,
                                               false);
    }
};

constexpr int kRawBufferFlags = 0x00020000;  // gfx9 raw buffer resource, dword 3

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void *p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, int(bytes), kRawBufferFlags);
}
__device__ __forceinline__ uint32_t bload(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    return uint32_t(__builtin_amdgcn_raw_buffer_load_b32(r, int(voff), int(soff), 0));
}
__device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, uint32_t v, unsigned voff, unsigned soff) {
    __builtin_amdgcn_raw_buffer_store_b32(int(v), r, int(voff), int(soff), 0);
}

typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x3 __attribute__((ext_vector_type(3)));

// The hidden fields are written once and read once, by another kernel, long after the caches have
// turned over (1-2 GB per application): their stores and loads carry the non-temporal policy, which
// leaves the caches to the input spinors and links (per application 1285 -> 1210 us in the general
// family, 800 -> 790 us in the projected one, at 24^3 x 48, 16 RHS).
constexpr int kNonTemporal = 2;  // gfx950 cache-policy bit nt

// N = Nc (2..4) consecutive words in one load / store; words past N read as zero; AUX = cache policy
template <int N, int AUX>
__device__ __forceinline__ u32x4 bload_words(__amdgpu_buffer_rsrc_t r, unsigned voff) {
    if constexpr (N == 2) {
        const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, int(voff), 0, AUX);
        return u32x4{v[0], v[1], 0u, 0u};
    } else if constexpr (N == 3) {
        const u32x3 v = __builtin_amdgcn_raw_buffer_load_b96(r, int(voff), 0, AUX);
        return u32x4{v[0], v[1], v[2], 0u};
    } else {
        return __builtin_amdgcn_raw_buffer_load_b128(r, int(voff), 0, AUX);
    }
}
template <int N, int AUX>
__device__ __forceinline__ void bstore_words(__amdgpu_buffer_rsrc_t r, u32x4 w, unsigned voff) {
    if constexpr (N == 2)
        __builtin_amdgcn_raw_buffer_store_b64(u32x2{w[0], w[1]}, r, int(voff), 0, AUX);
    else if constexpr (N == 3)
        __builtin_amdgcn_raw_buffer_store_b96(u32x3{w[0], w[1], w[2]}, r, int(voff), 0, AUX);
    else
        __builtin_amdgcn_raw_buffer_store_b128(w, r, int(voff), 0, AUX);
}

// MFMA A-operand fragments of the real 2x2 embedding of M, in lane order:
// frag[(rb * 2 + st) * 64 + lane] = 4 words (8 values) of rows (k, s, re|im) =
// rb*16 + lane%16 and columns (k', t, re|im) = st*32 + (lane/16)*8 + 0..7.
// PROJ: the rows (k, s < 2, re|im) only, 2 row blocks.
template <typename S, bool PROJ>
__global__ void k_lgnn_frags(u32x4 *__restrict__ frag, const Cx<float> *__restrict__ mcol) {
    using P = W<S>;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (PROJ ? 4 : 8) * 64) return;
    const int lane = i & 63, rbst = i >> 6, rb = rbst >> 1, st = rbst & 1;
    const int m = lane & 15, g = lane >> 4;
    u32x4 a;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        float v[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int row = rb * 16 + m, col = st * 32 + g * 8 + 2 * w + h;
            const int rk = row >> 1;  // (k, s) row of M: k 4 + s, or (k, s < 2) = k 2 + s projected
            const int ks = PROJ ? (rk >> 1) * 4 + (rk & 1) : rk, rho = row & 1, kt = col >> 1, sig = col & 1;
            const Cx<float> c = mcol[ks * 32 + kt];
            v[h] = rho == sig ? c.re : (rho == 0 ? -c.im : c.im);
        }
        a[w] = P::pack(v[0], v[1]);
    }
    frag[i] = a;
}

// q fields: 8 back-to-back fields [8][Vh][spin 4][R][NC] of 32-bit words (colour
// innermost, unlike the solver's vectors: one lane's colours are one 8/12/16-byte
// access in both kernels, and 16 lanes cover a contiguous run), [8][Vh][spin 2][R][NC]
// in the projected form; `fbytes` = bytes of one odd vector, 4 spins (< 2^31, checked
// on the host).
constexpr int kMixSites = 2;  // sites per wavefront of k_lgnn_mix

template <typename S, int NC, bool PROJ>
__global__ __launch_bounds__(256) void k_lgnn_mix(uint32_t *__restrict__ q, const uint32_t *__restrict__ in,
                                                 const uint32_t *__restrict__ gauge, const u32x4 *__restrict__ frag,
                                                 const int *__restrict__ walk0, int Vh, int R, unsigned fbytes,
                                                 unsigned gauge_bytes) {
    using P = W<S>;
    // A operand: real embedding of M (8 fragments x 64 lanes x 16 bytes), staged once per
    // workgroup in LDS and read per row block, so that only one row block's accumulators
    // and A fragments are live at a time
    constexpr int NRB = PROJ ? 2 : 4;  // row blocks of 16 rows (k, s, re|im)
    __shared__ u32x4 sA[NRB * 2 * 64];
    sA[threadIdx.x] = frag[threadIdx.x];
    if constexpr (!PROJ) sA[threadIdx.x + 256] = frag[threadIdx.x + 256];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int m = lane & 15, g = lane >> 4;
    const int pos0 = (xcd_remap(blockIdx.x, gridDim.x) * 4 + wv) * kMixSites;
    if (pos0 >= Vh) return;  // uniform
    const int r = 16 * int(blockIdx.y) + m;
    const unsigned row_bytes = unsigned(R) * 4u, site_bytes = 4u * NC * row_bytes;
    const unsigned rcb = unsigned(min(r, R - 1)) * 4u;
    const auto r_in = rsrc(in, fbytes);
    const auto r_g = rsrc(gauge, gauge_bytes);
    const size_t fwords = fbytes / 4u;
    const unsigned spin_bytes = unsigned(R) * NC * 4u;
    const bool fwd0 = (g & 1) == 0;  // hop parity is the same in both k steps
    // word permutations: swap the two 16-bit halves; sign of the high / low half
    constexpr uint32_t kSwap = 0x01000302u, kSignHi = 0x80000000u, kSignLo = 0x00008000u;

    for (int j = 0; j < kMixSites; ++j) {
        const int pos = pos0 + j;
        if (pos >= Vh) break;  // uniform
        const int *row = walk0 + long(pos) * 9;
        const int site = row[0];
        // B operand of both k steps: this lane's hop k' = 4 st + g transported, (spin, re|im) packed
        u32x4 bop[2][NC];
#pragma unroll
        for (int st = 0; st < 2; ++st) {
            const int h = 4 * st + g, mu = h >> 1;
            const int lo = (g & 1) ? row[2 + 4 * st] : row[1 + 4 * st];
            const int hi = (g & 1) ? row[4 + 4 * st] : row[3 + 4 * st];
            const int nb = (g & 2) ? hi : lo;
            const unsigned va = unsigned(nb) * site_bytes + rcb;
            uint32_t psi[4][NC], lk[NC][NC];
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int b = 0; b < NC; ++b) psi[t][b] = bload(r_in, va + unsigned(t * NC + b) * row_bytes, 0u);
            // forward: U_mu(y) at parity 0; backward: U_mu(nb) at parity 1 (used as its adjoint)
            const unsigned gl = fwd0 ? (unsigned(mu) * 2u) * unsigned(Vh) + unsigned(site)
                                     : (unsigned(mu) * 2u + 1u) * unsigned(Vh) + unsigned(nb);
#pragma unroll
            for (int a = 0; a < NC; ++a)
#pragma unroll
                for (int b = 0; b < NC; ++b) lk[a][b] = bload(r_g, (gl * NC * NC + unsigned(a * NC + b)) * 4u, 0u);
            // u = V psi with V[c][b] = U[c][b] forward, conj(U[b][c]) backward, as two dot2 per term:
            //   forward:  re = (Ur, -Ui).(pr, pi), im = (Ui, Ur).(pr, pi)
            //   backward: re = (Ur,  Ui).(pr, pi), im = (-Ui, Ur).(pr, pi)
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                uint32_t la[NC], lb[NC];
#pragma unroll
                for (int b = 0; b < NC; ++b) {
                    const uint32_t uw = fwd0 ? lk[c][b] : lk[b][c];
                    const uint32_t sw = __builtin_amdgcn_perm(uw, uw, kSwap);
                    la[b] = fwd0 ? (uw ^ kSignHi) : uw;
                    lb[b] = fwd0 ? sw : (sw ^ kSignLo);
                }
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    float ur = 0.f, ui = 0.f;
#pragma unroll
                    for (int b = 0; b < NC; ++b) {
                        ur = P::dot(la[b], psi[t][b], ur);
                        ui = P::dot(lb[b], psi[t][b], ui);
                    }
                    bop[st][c][t] = P::pack(ur, ui);
                }
            }
        }
        // lane (m, g) holds rows 4g + e of block rb: (k, s) = (2 rb + g/2, 2 (g & 1) + e/2), re|im = e & 1;
        // fields 2 rb and 2 rb + 1 share one buffer resource, the lane's field is a 32-bit offset.
        // Projected: (k, s) = (4 rb + g, e/2), fields 4 rb .. 4 rb + 3 (of 2 spins) behind one resource.
        const unsigned qf = PROJ ? fbytes / 2u : fbytes;  // bytes of one q field
        const unsigned lbase = PROJ ? unsigned(g) * qf + unsigned(site) * (site_bytes / 2u) + unsigned(r) * NC * 4u
                                    : unsigned(g >> 1) * fbytes + unsigned(site) * site_bytes +
                                          unsigned(2 * (g & 1)) * spin_bytes + unsigned(r) * NC * 4u;
#pragma unroll
        for (int rb = 0; rb < NRB; ++rb) {
            const u32x4 a0 = sA[(rb * 2) * 64 + lane], a1 = sA[(rb * 2 + 1) * 64 + lane];
            f32x4 d[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                d[c] = P::mma(a0, bop[0][c], f32x4{0.f, 0.f, 0.f, 0.f});
                d[c] = P::mma(a1, bop[1][c], d[c]);
            }
            if (r < R) {
                const auto r_q = PROJ ? rsrc(q + size_t(4 * rb) * (qf / 4u), 4u * qf)
                                      : rsrc(q + size_t(2 * rb) * fwords, 2u * fbytes);
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    u32x4 w = {0u, 0u, 0u, 0u};
#pragma unroll
                    for (int c = 0; c < NC; ++c) w[c] = P::pack(d[c][2 * i], d[c][2 * i + 1]);
                    bstore_words<NC, kNonTemporal>(r_q, w, lbase + unsigned(i) * spin_bytes);
                }
            }
        }
    }
}

// w times i^k on a packed (re, im) word of two 16-bit values: odd k swaps the halves, then the sign
// bits of the negated halves flip (i: (-im, re), -1: (-re, -im), -i: (im, -re))
__device__ __forceinline__ uint32_t times_i_pow(uint32_t w, int k) {
    const uint32_t sw = __builtin_amdgcn_perm(w, w, (k & 1) ? 0x01000302u : 0x03020100u);
    const uint32_t mask = k == 0 ? 0u : k == 1 ? 0x00008000u : k == 2 ? 0x80008000u : 0x80000000u;
    return sw ^ mask;
}

template <typename S, int NC, bool WIDE, bool PROJ>
__global__ __launch_bounds__(256) void k_lgnn_gather(uint32_t *__restrict__ out, const uint32_t *__restrict__ in,
                                                    const uint32_t *__restrict__ q, const uint32_t *__restrict__ gauge,
                                                    const Cx<float> *__restrict__ m0, const int *__restrict__ walk1,
                                                    int Vh, int R, unsigned fbytes, unsigned gauge_bytes) {
    using P = W<S>;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int m = lane & 15, g = lane >> 4;
    const int pos = xcd_remap(blockIdx.x, gridDim.x) * 4 + wv;
    if (pos >= Vh) return;  // uniform
    const int r = 16 * int(blockIdx.y) + m;
    const unsigned row_bytes = unsigned(R) * 4u, site_bytes = 4u * NC * row_bytes;
    const int cb = m >> 1;                 // output colour of A row m
    const bool bre = (m & 1) == 0;
    const int cbc = cb < NC ? cb : NC - 1;  // clamped (row never stored)
    const auto r_g = rsrc(gauge, gauge_bytes);
    const auto r_in = rsrc(in, fbytes);
    const auto r_out = rsrc(out, fbytes);

    const int *row = walk1 + long(pos) * 9;
    const int site = row[0];
    f32x4 d[4];
#pragma unroll
    for (int sp = 0; sp < 4; ++sp) d[sp] = f32x4{0.f, 0.f, 0.f, 0.f};
    u32x4 qw[2][4];
    uint32_t lw[2][4];
#pragma unroll
    for (int st = 0; st < 2; ++st) {
        const int h = 4 * st + g, mu = h >> 1;
        const bool fwd = (h & 1) == 0;
        const int lo = (g & 1) ? row[2 + 4 * st] : row[1 + 4 * st];
        const int hi = (g & 1) ? row[4 + 4 * st] : row[3 + 4 * st];
        const int nb = (g & 2) ? hi : lo;
        const unsigned qf = PROJ ? fbytes / 2u : fbytes;  // bytes of one q field
        const unsigned va = unsigned(nb) * (PROJ ? site_bytes / 2u : site_bytes) + unsigned(min(r, R - 1)) * NC * 4u;
        // field h of q at the neighbour: this lane's RHS, colours c' < NC of each spin (padding zero).
        // WIDE: one buffer resource spans the step's 4 fields, uniform over the wavefront, and the
        // lane group's field is a byte offset; otherwise a resource per field, which differs between
        // lane groups and makes the compiler loop over the distinct resources for every load.
        __amdgpu_buffer_rsrc_t r_q;
        unsigned vq = va;
        if constexpr (WIDE) {
            r_q = rsrc(q + size_t(4 * st) * (qf / 4u), 4u * qf);
            vq = va + unsigned(g) * qf;
        } else {
            r_q = rsrc(q + size_t(h) * (qf / 4u), qf);
        }
#pragma unroll
        for (int sp = 0; sp < (PROJ ? 2 : 4); ++sp)
            qw[st][sp] = bload_words<NC, kNonTemporal>(r_q, vq + unsigned(sp) * unsigned(R) * NC * 4u);
        if constexpr (PROJ) {
            // q_h lies in the image of the hop's projector: spin m(r) = i^kr x spin r, with m(0) = 3,
            // m(1) = 2 for the x and y hops (step 0) and m(0) = 2, m(1) = 3 for z and t (step 1)
            const HopInfo hi = hop_info(mu, fwd, false);
            u32x4 lo0, lo1;
#pragma unroll
            for (int cc = 0; cc < 4; ++cc) {
                lo0[cc] = cc < NC ? times_i_pow(qw[st][0][cc], hi.kr0) : 0u;
                lo1[cc] = cc < NC ? times_i_pow(qw[st][1][cc], hi.kr1) : 0u;
            }
            qw[st][st == 0 ? 3 : 2] = lo0;
            qw[st][st == 0 ? 2 : 3] = lo1;
        }
        // A word from link V[cb][c']: forward U(x)[cb][c'] (parity 1 at x), backward conj(U(nb)[c'][cb]) (parity 0)
        const unsigned gl = fwd ? (unsigned(mu) * 2u + 1u) * unsigned(Vh) + unsigned(site)
                                : (unsigned(mu) * 2u) * unsigned(Vh) + unsigned(nb);
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) {
            uint32_t w = 0u;
            if (cc < NC) {
                const unsigned idx = fwd ? unsigned(cbc * NC + cc) : unsigned(cc * NC + cbc);
                const uint32_t uw = bload(r_g, (gl * NC * NC + idx) * 4u, 0u);
                const float vr = P::re(uw), vi = fwd ? P::im(uw) : -P::im(uw);
                // real row (Vr, -Vi) for re output, (Vi, Vr) for im output
                w = bre ? P::pack(vr, -vi) : P::pack(vi, vr);
            }
            lw[st][cc] = w;
        }
    }
#pragma unroll
    for (int st = 0; st < 2; ++st) {
        const u32x4 a = {lw[st][0], lw[st][1], lw[st][2], lw[st][3]};
#pragma unroll
        for (int sp = 0; sp < 4; ++sp) {
            const u32x4 b = qw[st][sp];
            d[sp] = P::mma(a, b, d[sp]);
        }
    }
    // epilogue: lane (m, g) holds rows 4g + e = (colour 2g + e/2, re|im) of RHS r; add M0 psi(x)
    if (r < R) {
        const unsigned xb = unsigned(site) * site_bytes + unsigned(r) * 4u;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int c = 2 * g + i;
            if (c < NC) {
                float pr[4], pi[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const uint32_t w = bload(r_in, xb + unsigned(t * NC + c) * row_bytes, 0u);
                    pr[t] = P::re(w);
                    pi[t] = P::im(w);
                }
#pragma unroll
                for (int sp = 0; sp < 4; ++sp) {
                    float vr = d[sp][2 * i], vi = d[sp][2 * i + 1];
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        const Cx<float> mm = m0[sp * 4 + t];
                        vr = fmaf(mm.re, pr[t], fmaf(-mm.im, pi[t], vr));
                        vi = fmaf(mm.re, pi[t], fmaf(mm.im, pr[t], vi));
                    }
                    bstore(r_out, P::pack(vr, vi), xb + unsigned(sp * NC + c) * row_bytes, 0u);
                }
            }
        }
    }
}

template <typename S, int NC, bool PROJ>
hipError_t launch_nc(void *out, const void *in, const void *gauge, void *q, const Cx<float> *mcol, const int *walk,
                     const Geom &g, int R, hipStream_t s) {
    const unsigned fbytes = unsigned(size_t(g.Vh) * 4 * NC * R * 4);
    const unsigned gbytes = unsigned(size_t(g.Vh) * 8 * NC * NC * 4);
    const dim3 grid(div_up(g.Vh, 4), div_up(R, 16)), block(256);
    // the A fragments live after the 8 q fields (of 4 spins: the same place in the projected form)
    auto *frag = reinterpret_cast<u32x4 *>(static_cast<char *>(q) + 8 * size_t(fbytes));
    hipLaunchKernelGGL((k_lgnn_frags<S, PROJ>), dim3(PROJ ? 1 : 2), dim3(256), 0, s, frag, mcol);
    QCU_HIP_TRY(hipGetLastError());
    const dim3 grid_mix(div_up(g.Vh, 4 * kMixSites), div_up(R, 16));
    hipLaunchKernelGGL((k_lgnn_mix<S, NC, PROJ>), grid_mix, block, 0, s, static_cast<uint32_t *>(q),
                       static_cast<const uint32_t *>(in), static_cast<const uint32_t *>(gauge), frag, walk, g.Vh, R,
                       fbytes, gbytes);
    QCU_HIP_TRY(hipGetLastError());
    // four fields behind one buffer resource while their bytes stay below 2^31
    const double qf = PROJ ? 0.5 * double(fbytes) : double(fbytes);
    auto gather = 4.0 * qf < 2147483647.0 ? k_lgnn_gather<S, NC, true, PROJ> : k_lgnn_gather<S, NC, false, PROJ>;
    hipLaunchKernelGGL(gather, grid, block, 0, s, static_cast<uint32_t *>(out), static_cast<const uint32_t *>(in),
                       static_cast<const uint32_t *>(q), static_cast<const uint32_t *>(gauge), mcol + 1024,
                       walk + size_t(g.Vh) * 9, g.Vh, R, fbytes, gbytes);
    return hipGetLastError();
}

template <typename S, bool PROJ>
hipError_t launch_s(void *out, const void *in, const void *gauge, void *q, const Cx<float> *mcol, const int *walk,
                    const Geom &g, int Nc, int R, hipStream_t s) {
    switch (Nc) {
        case 2: return launch_nc<S, 2, PROJ>(out, in, gauge, q, mcol, walk, g, R, s);
        case 3: return launch_nc<S, 3, PROJ>(out, in, gauge, q, mcol, walk, g, R, s);
        case 4: return launch_nc<S, 4, PROJ>(out, in, gauge, q, mcol, walk, g, R, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_lgnn_mfma(void *out, const void *in, const void *gauge, void *q, const Cx<float> *mcol,
                            const int *walk, const Geom &g, int Nc, int R, int prec, bool projected, hipStream_t s) {
    if (Nc < 2 || Nc > 4 || R < 1 || !out || !in || !gauge || !q || !mcol || !walk) return hipErrorInvalidValue;
    // two q fields share one buffer resource: 2 x (Vh 4 Nc R words) must stay below 2^31 bytes
    if (double(g.Vh) * 32.0 * Nc * R >= 2147483647.0 || double(g.Vh) * 32.0 * Nc * Nc >= 2147483647.0)
        return hipErrorInvalidValue;
    if (prec == kHalf)
        return projected ? launch_s<__half, true>(out, in, gauge, q, mcol, walk, g, Nc, R, s)
                         : launch_s<__half, false>(out, in, gauge, q, mcol, walk, g, Nc, R, s);
    if (prec == kBF16)
        return projected ? launch_s<__hip_bfloat16, true>(out, in, gauge, q, mcol, walk, g, Nc, R, s)
                         : launch_s<__hip_bfloat16, false>(out, in, gauge, q, mcol, walk, g, Nc, R, s);
    return hipErrorInvalidValue;
}

}  // namespace qcu
