// dslash_mfma.hip -- Wilson hopping term on MFMA for FP16 / BF16 storage
// (v_mfma_f32_16x16x32_{f16,bf16}, FP32 accumulation).
//
// One wavefront computes one output site.  For each output spin s the site's
// result is a real GEMM
//     D_s[rhs][(c, re|im)] = sum_k A_s[rhs][k] * B[k][(c, re|im)]
// with k running over (hop h, input colour c', re|im): A_s holds the spin-
// projected half spinor of the neighbour (h_r = psi_r + s psi_m(r) for the
// upper spins, the reconstruction-scaled copy for the lower ones), B the real
// 2x2 embedding of the link V_h (U_mu(x) forward, U_mu(x-mu)^+ backward):
//     B[(c',re)][(c,re)] = Re V[c][c'],  B[(c',im)][(c,re)] = -Im V[c][c'],
//     B[(c',re)][(c,im)] = Im V[c][c'],  B[(c',im)][(c,im)] =  Re V[c][c'].
// Each hop takes 8 k entries (2 Nc <= 8, zero padded) or 16 (Nc = 8), so one
// K = 32 instruction covers 4 hops (2 for Nc = 8): 8 MFMAs per site and 16
// right-hand sides.  The instruction is issued transposed, D^T = B^T A^T
// (links as the matrix-core A operand, spinors as B), so each lane's
// accumulator rows are whole (re, im) pairs of its own RHS and go straight to
// memory.  Lane l: operand row / column l % 16, k group l / 16; accumulator
// rows 4*(l/16)..+3, column l % 16 (verified on gfx950).
//
// Every complex number is one 32-bit word (re in the low half), so the
// spin projection and reconstruction are word operations: multiplying by
// i^k is a 16-bit rotate plus a sign-bit XOR, the sum a packed add -- the
// words go into the MFMA operands without any conversion.  Rows beyond R and
// columns beyond Nc read clamped, valid addresses and are never stored, so
// the kernel has no divergent branches around its loads.
#include <cstdint>
#include <type_traits>

#include "dslash_common.h"
#include "geom.h"

namespace qcu {
namespace {

typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <typename S> struct Pk;
template <typename S> struct PkPhase;
template <> struct Pk<__half> {
    __device__ __forceinline__ static uint32_t add(uint32_t a, uint32_t b) {
        return __builtin_bit_cast(uint32_t, __builtin_bit_cast(f16x2, a) + __builtin_bit_cast(f16x2, b));
    }
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
};
// Multiplication by i^k with the parity of k known at compile time (odd: the
// (re, im) halves swap) and its sign pattern per lane.  FP16: a packed sign
// vector, so z + i^k w is one v_pk_fma_f16 (half swap through op_sel) and
// i^k z one v_pk_mul_f16.
template <> struct PkPhase<__half> {
    f16x2 sv;
    // from a sign vector word already in the form make() builds
    __device__ __forceinline__ static PkPhase raw(uint32_t word) {
        PkPhase p;
        p.sv = __builtin_bit_cast(f16x2, word);
        return p;
    }
    __device__ __forceinline__ static PkPhase make(uint32_t mask) {
        PkPhase p;
        // (+-1, +-1) in FP16: 0x3c00 per half, the mask's sign bits on top
        p.sv = __builtin_bit_cast(f16x2, 0x3c003c00u | (mask & 0x80008000u));
        return p;
    }
    template <bool ODD>
    __device__ __forceinline__ uint32_t fma(uint32_t w, uint32_t z) const {  // w + i^k z
        const f16x2 zz = __builtin_bit_cast(f16x2, z);
        const f16x2 zr = ODD ? __builtin_shufflevector(zz, zz, 1, 0) : zz;
        return __builtin_bit_cast(uint32_t, __builtin_bit_cast(f16x2, w) + sv * zr);
    }
    template <bool ODD>
    __device__ __forceinline__ uint32_t mul(uint32_t z) const {  // i^k z
        const f16x2 zz = __builtin_bit_cast(f16x2, z);
        const f16x2 zr = ODD ? __builtin_shufflevector(zz, zz, 1, 0) : zz;
        return __builtin_bit_cast(uint32_t, sv * zr);
    }
};
// BF16 (no packed BF16 arithmetic): the word operation with a compile-time byte permutation
// (absent for even k) and the per-lane sign bits
template <> struct PkPhase<__hip_bfloat16> {
    uint32_t mask;
    __device__ __forceinline__ static PkPhase raw(uint32_t m) { return make(m); }
    __device__ __forceinline__ static PkPhase make(uint32_t m) {
        PkPhase p;
        p.mask = m;
        return p;
    }
    template <bool ODD>
    __device__ __forceinline__ uint32_t mul(uint32_t z) const {
        if constexpr (ODD) z = __builtin_amdgcn_perm(z, z, 0x01000302u);
        return z ^ mask;
    }
    template <bool ODD>
    __device__ __forceinline__ uint32_t fma(uint32_t w, uint32_t z) const;
};

template <> struct Pk<__hip_bfloat16> {
    __device__ __forceinline__ static float re(uint32_t w) { return __builtin_bit_cast(float, w << 16); }
    __device__ __forceinline__ static float im(uint32_t w) { return __builtin_bit_cast(float, w & 0xffff0000u); }
    __device__ __forceinline__ static uint32_t pack(float r, float i) {
        const bf16x2 v = {__bf16(r), __bf16(i)};
        return __builtin_bit_cast(uint32_t, v);
    }
    __device__ __forceinline__ static uint32_t add(uint32_t a, uint32_t b) {
        return pack(re(a) + re(b), im(a) + im(b));
    }
    __device__ __forceinline__ static f32x4 mma(u32x4 a, u32x4 b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                       __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
    }
};

template <bool ODD>
__device__ __forceinline__ uint32_t PkPhase<__hip_bfloat16>::fma(uint32_t w, uint32_t z) const {
    return Pk<__hip_bfloat16>::add(w, mul<ODD>(z));
}

constexpr int kRawBufferFlags = 0x00020000;  // gfx9 raw buffer resource, dword 3

template <typename T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const T *p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(p), 0, int(bytes), kRawBufferFlags);
}
__device__ __forceinline__ uint32_t bload(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    return uint32_t(__builtin_amdgcn_raw_buffer_load_b32(r, int(voff), int(soff), 0));
}

// v_perm_b32 byte selectors: identity and 16-bit rotate of a (re, im) word
constexpr uint32_t kPermId = 0x03020100u, kPermRot = 0x01000302u;

struct WordOp {  // z -> z * i^k on a packed word: byte permutation, then sign bits
    uint32_t sel, mask;
};
__device__ __forceinline__ WordOp word_op(int k) {
    // i^1: (-im, re)  i^2: (-re, -im)  i^3: (im, -re)
    const uint32_t odd = uint32_t(k & 1), hi = uint32_t(k >> 1) & 1u;
    WordOp w;
    w.sel = kPermId ^ (odd * (kPermId ^ kPermRot));
    w.mask = ((odd ^ hi) << 15) | (hi << 31);
    return w;
}
// directions in the order x, z, y, t: those with odd phases (x, z) first
__host__ __device__ constexpr int mu_of(int i) { return ((i & 1) << 1) | (i >> 1); }

// Sign masks of the projection (p0, p1) and reconstruction (r0, r1) phases i^k of
// every hop position hh = 0..7 (direction mu_of(hh / 2), forward for even hh),
// without and with the dagger: [dagger][hh][p0, p1, r0, r1].  A lane reads the
// row of its hop with one 16-byte load instead of deriving the four masks
// (about 45 vector instructions per wavefront, paid once per site at one site
// per wavefront).  m: the masks (BF16 word operations); h: the FP16 sign
// vectors (+-1, +-1) that PkPhase<__half> multiplies by.
struct PhaseTab {
    uint32_t m[2][8][4];
    uint32_t h[2][8][4];
};
__host__ __device__ constexpr uint32_t phase_mask(int k) {
    return ((uint32_t((k & 1) ^ ((k >> 1) & 1))) << 15) | (uint32_t((k >> 1) & 1) << 31);
}
__host__ __device__ constexpr PhaseTab make_phase_tab() {
    PhaseTab t{};
    for (int dag = 0; dag < 2; ++dag)
        for (int hh = 0; hh < 8; ++hh) {
            const int mu = mu_of(hh >> 1);
            const bool fwd = (hh & 1) == 0;
            const int sgn = (fwd != (dag != 0)) ? 2 : 0;
            t.m[dag][hh][0] = phase_mask((kProjK(mu, 0) + sgn) & 3);
            t.m[dag][hh][1] = phase_mask((kProjK(mu, 1) + sgn) & 3);
            t.m[dag][hh][2] = phase_mask((kReconK(mu, 0) + sgn) & 3);
            t.m[dag][hh][3] = phase_mask((kReconK(mu, 1) + sgn) & 3);
            for (int j = 0; j < 4; ++j) t.h[dag][hh][j] = 0x3c003c00u | (t.m[dag][hh][j] & 0x80008000u);
        }
    return t;
}
__constant__ PhaseTab kPhaseTab = make_phase_tab();

__device__ __forceinline__ uint32_t apply(const WordOp &o, uint32_t w) {
    return __builtin_amdgcn_perm(w, w, o.sel) ^ o.mask;
}

__device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, uint32_t v, unsigned voff, unsigned soff) {
    __builtin_amdgcn_raw_buffer_store_b32(int(v), r, int(voff), int(soff), 0);
}

// One wavefront = a run of traversal positions (`spw` consecutive ones, or a
// stride through its XCD's share in the persistent form) x one 16-RHS chunk
// (blockIdx.y).  Per site the wavefront needs one scalar row of the walk
// table (site + 8 neighbours), prefetched one site ahead, so the only exposed
// memory latency per site is that of its link and spinor loads, which are
// all issued back to back before any of them is waited for.  (Computing the
// row from the traversal index instead costs about 500 scalar instructions
// per site, which would make the scalar unit the bottleneck.)
//
// RED (xpay instance only): the epilogue also accumulates omega's inner
// products <t, s> and <t, t> over the values it stores (rounded to the storage
// precision, as a separate reduction would read them), for the unpreconditioned
// t = s - kappa^2 D_oe D_eo s whose xpay vector is s itself, so the partner
// words are the xpay words already loaded.  (An inner product with a vector the
// kernel does not otherwise read -- alpha's <r0, v> -- costs more in this
// memory-latency-bound kernel than the separate pass it would replace: 360 us
// against 268 + 105 us at 24^3 x 48, 16 RHS.)  Lanes sum over their entries,
// the four lane groups of an RHS combine by shuffles and the workgroup's
// wavefronts through LDS; each workgroup writes one row of per-RHS partials
// [blockIdx.x][R][4], which two small launches reduce in a fixed order.
template <typename S, int NC, bool DAG, bool XP, bool RED = false>
__global__ __launch_bounds__(256) void k_dslash_mfma(uint32_t *__restrict__ out, const uint32_t *__restrict__ in,
                                                    const uint32_t *__restrict__ gauge,
                                                    const uint32_t *__restrict__ xpay, float coef, int Vh, int R,
                                                    int parity, const int *__restrict__ walk, int spw,
                                                    unsigned vec_bytes, unsigned gauge_bytes,
                                                    double *__restrict__ part = nullptr) {
    static_assert(!RED || (XP && !DAG), "the fused reduction belongs to the xpay hop of the Schur operator");
    using P = Pk<S>;
    constexpr int KH = (2 * NC <= 8) ? 8 : 16;  // k entries per hop
    constexpr int HPS = 32 / KH;                 // hops per MFMA
    constexpr int STEPS = 8 / HPS;               // K steps per site
    constexpr int GPH = 4 / HPS;                 // lane groups per hop
    constexpr int CPL = NC < 4 ? NC : 4;         // input colours per lane and hop

    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int m = lane & 15;
    const int grp = lane >> 4;
    const int csub = (grp % GPH) * 4;           // first input colour of this lane's k entries
    const bool gb0 = (grp & 1) != 0, gb1 = (grp & 2) != 0;
    const int cb = m >> 1;                      // output colour of B / D column m
    const bool bre = (m & 1) == 0;
    const int cbc = cb < NC ? cb : NC - 1;      // clamped (column never stored)
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int r0 = 16 * int(blockIdx.y);
    const int r = r0 + m;
    const unsigned row_bytes = unsigned(R) * 4u;            // one (spin, colour) row of R words
    const unsigned site_bytes = 4u * NC * row_bytes;
    const unsigned rcb = unsigned(min(r, R - 1)) * 4u;      // rows beyond R are never stored
    const unsigned ll = unsigned(NC) * NC;
    const auto r_in = rsrc(in, vec_bytes);
    const auto r_g = rsrc(gauge, gauge_bytes);
    const auto r_out = rsrc(out, vec_bytes);
    const auto r_x = rsrc(xpay, xpay ? vec_bytes : 0u);

    // per-lane, per-step hop constants.  The hops are taken in the direction order x, z, y, t
    // (position hh = st * HPS + grp / GPH, direction mu_of(hh / 2)): the x and z hops have odd
    // projection and reconstruction phases i^k, the y and t hops even ones, so every step's
    // phases share their parity and the half swap of an odd phase is fixed at compile time.
    using PH = PkPhase<S>;
    int mu_s[STEPS];
    bool fwd_s[STEPS];
    PH p0_s[STEPS], p1_s[STEPS], r0_s[STEPS], r1_s[STEPS];
    WordOp bop_s[STEPS];
#pragma unroll
    for (int st = 0; st < STEPS; ++st) {
        const int hh = st * HPS + grp / GPH;
        const int mu = mu_of(hh >> 1);
        const bool fwd = (hh & 1) == 0;
        const u32x4 pm = *reinterpret_cast<const u32x4 *>(std::is_same<S, __half>::value
                                                              ? kPhaseTab.h[DAG ? 1 : 0][hh]
                                                              : kPhaseTab.m[DAG ? 1 : 0][hh]);
        mu_s[st] = mu;
        fwd_s[st] = fwd;
        p0_s[st] = PH::raw(pm[0]);
        p1_s[st] = PH::raw(pm[1]);
        r0_s[st] = PH::raw(pm[2]);
        r1_s[st] = PH::raw(pm[3]);
        // B word from the stored link word U: forward V = U[cb][c'], backward
        // V = conj(U[c'][cb]); real column (Vr, -Vi), imaginary column (Vi, Vr)
        bop_s[st].sel = bre ? kPermId : kPermRot;
        bop_s[st].mask = fwd ? (bre ? 0x80000000u : 0u) : (bre ? 0u : 0x00008000u);
    }

    const int *wtab = walk + long(parity) * Vh * 9;
    // Traversal positions of this wavefront: q_first, q_first + q_step, ... below q_end.
    // spw > 0: `spw` consecutive positions.  spw < 0 (persistent grid of -spw workgroups per
    // CU, a multiple of the 8 XCDs): the logical workgroups of one XCD (consecutive after
    // xcd_remap) share that XCD's contiguous eighth of the traversal and stride through it
    // together, so the sites in flight on an XCD stay one compact window while each
    // wavefront's set-up above is paid once for many sites.
    int q_first, q_step, q_end;
    if (spw > 0) {
        q_first = (b * 4 + wv) * spw;
        q_step = 1;
        q_end = min(q_first + spw, Vh);
    } else {
        const int wpx = int(gridDim.x) * 4 / kNumXcd;  // wavefronts per XCD
        const int w = b * 4 + wv;
        const int x = w / wpx;
        const int chunk = (Vh + kNumXcd - 1) / kNumXcd;
        q_first = x * chunk + (w - x * wpx);
        q_step = wpx;
        q_end = min((x + 1) * chunk, Vh);
    }
    if constexpr (!RED)
        if (q_first >= q_end) return;  // uniform
    // fused inner products: this lane's partial sums (Re, Im of <s, t>, |t|^2)
    float racc[3] = {0.f, 0.f, 0.f};
    int cur[9];
    if (q_first < q_end) {
#pragma unroll
        for (int k = 0; k < 9; ++k) cur[k] = wtab[long(q_first) * 9 + k];
    }

    for (int q = q_first; q < q_end; q += q_step) {
        const bool more = q + q_step < q_end;
        const int site = cur[0];

        unsigned voff_a[STEPS], voff_b0[STEPS], voff_b1[STEPS], gaddr[STEPS], gstep[STEPS];
#pragma unroll
        for (int st = 0; st < STEPS; ++st) {
            // this lane's hop: two-level select of the scalar neighbour row (no branches)
            // (walk row: cur[1 + 2 mu + (0 forward | 1 backward)])
            int nb;
            if constexpr (HPS == 4) {
                // lane groups 0, 1: direction mu_of(2 st) (x | y), groups 2, 3: mu_of(2 st + 1) (z | t)
                const int lo = gb0 ? cur[2 + 2 * st] : cur[1 + 2 * st];
                const int hi = gb0 ? cur[6 + 2 * st] : cur[5 + 2 * st];
                nb = gb1 ? hi : lo;
            } else {
                nb = gb1 ? cur[2 + 2 * mu_of(st)] : cur[1 + 2 * mu_of(st)];
            }
            const unsigned va = unsigned(nb) * site_bytes + rcb + unsigned(csub) * row_bytes;
            // spin m(0) is 3 for x, y hops and 2 for z, t hops: lower-spin rows of the projection
            const bool m02 = mu_s[st] >= 2;
            voff_a[st] = va;
            voff_b0[st] = va + (m02 ? 0u : unsigned(NC) * row_bytes);
            voff_b1[st] = va + (m02 ? unsigned(NC) * row_bytes : 0u);
            const unsigned gl = fwd_s[st] ? (unsigned(mu_s[st]) * 2 + parity) * unsigned(Vh) + unsigned(site)
                                          : (unsigned(mu_s[st]) * 2 + 1 - parity) * unsigned(Vh) + unsigned(nb);
            gstep[st] = fwd_s[st] ? 4u : 4u * NC;  // step of c' in bytes
            gaddr[st] = (gl * ll + unsigned(cbc) * (fwd_s[st] ? NC : 1)) * 4u + unsigned(csub) * gstep[st];
        }
        // prefetch the next walk row only now, after the last use of this one
        int nxt[9];
        if (more) {
#pragma unroll
            for (int k = 0; k < 9; ++k) nxt[k] = wtab[long(q + q_step) * 9 + k];
        }
        // every load of this site first: spinor rows, then links
        uint32_t w0[STEPS][4], w1[STEPS][4], wb0[STEPS][4], wb1[STEPS][4], lw[STEPS][4];
#pragma unroll
        for (int st = 0; st < STEPS; ++st)
#pragma unroll
            for (int cc = 0; cc < 4; ++cc) {
                if (cc < CPL) {
                    const unsigned c1 = unsigned(cc) * row_bytes;
                    w0[st][cc] = bload(r_in, voff_a[st] + c1, 0u);
                    w1[st][cc] = bload(r_in, voff_a[st] + c1, unsigned(NC) * row_bytes);
                    wb0[st][cc] = bload(r_in, voff_b0[st] + c1, 2u * NC * row_bytes);
                    wb1[st][cc] = bload(r_in, voff_b1[st] + c1, 2u * NC * row_bytes);
                } else {
                    w0[st][cc] = w1[st][cc] = wb0[st][cc] = wb1[st][cc] = 0u;
                }
            }
#pragma unroll
        for (int st = 0; st < STEPS; ++st)
#pragma unroll
            for (int cc = 0; cc < 4; ++cc) lw[st][cc] = cc < CPL ? bload(r_g, gaddr[st] + unsigned(cc) * gstep[st], 0u) : 0u;
        // the xpay words of this lane's output rows are needed only in the epilogue: issue them
        // now so their latency overlaps the MFMA work instead of serialising every store
        const unsigned xbase = unsigned(site) * site_bytes + rcb + unsigned(2 * grp) * row_bytes;
        uint32_t xw[2][4];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int sp = 0; sp < 4; ++sp)
                xw[i][sp] = (XP && 2 * grp + i < NC) ? bload(r_x, xbase, unsigned(sp * NC + i) * row_bytes) : 0u;

        // the arithmetic runs at raised wavefront priority: a wavefront whose operands have
        // arrived finishes its site ahead of wavefronts still issuing loads, so its next
        // site's loads go out sooner (FP16 256 -> 245 us, BF16 259 -> 241 us at 24^3x48,
        // 16 RHS; raising it for the load phase instead was slower)
        __builtin_amdgcn_s_setprio(1);
        f32x4 d[4];
#pragma unroll
        for (int sp = 0; sp < 4; ++sp) d[sp] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int st = 0; st < STEPS; ++st) {
            constexpr bool kOddFirst = true;  // steps over x and z hops come first
            const bool odd = kOddFirst && (st * HPS < 4);
            const bool m02 = mu_s[st] >= 2;  // per lane when a step holds two directions (HPS = 4)
            u32x4 a[4], bf;
#pragma unroll
            for (int cc = 0; cc < 4; ++cc) {
                if (cc >= CPL) {  // zero padding of the k entries (2 Nc < 8)
                    a[0][cc] = a[1][cc] = a[2][cc] = a[3][cc] = 0u;
                    bf[cc] = 0u;
                    continue;
                }
                uint32_t h0, h1, g0, g1;
                if (odd) {
                    h0 = p0_s[st].template fma<true>(w0[st][cc], wb0[st][cc]);
                    h1 = p1_s[st].template fma<true>(w1[st][cc], wb1[st][cc]);
                    g0 = r0_s[st].template mul<true>(h0);
                    g1 = r1_s[st].template mul<true>(h1);
                } else {
                    h0 = p0_s[st].template fma<false>(w0[st][cc], wb0[st][cc]);
                    h1 = p1_s[st].template fma<false>(w1[st][cc], wb1[st][cc]);
                    g0 = r0_s[st].template mul<false>(h0);
                    g1 = r1_s[st].template mul<false>(h1);
                }
                a[0][cc] = h0;
                a[1][cc] = h1;
                a[2][cc] = m02 ? g0 : g1;
                a[3][cc] = m02 ? g1 : g0;
                bf[cc] = apply(bop_s[st], lw[st][cc]);
            }
            // links as the A operand, spinors as B: D^T = B^T A^T, so the
            // accumulator rows are (colour, re|im) and its column is this lane's RHS
#pragma unroll
            for (int sp = 0; sp < 4; ++sp) d[sp] = P::mma(bf, a[sp], d[sp]);
        }
        __builtin_amdgcn_s_setprio(0);
        // lane (m, grp) holds rows 4*grp + e = (colour 2*grp + e/2, re|im) of RHS r
        const bool unit = coef == 1.f;  // the bare operator's first Schur hop: no scaling
        // lane (m, grp) holds rows 4*grp + e = (colour 2*grp + e/2, re|im) of RHS r
        auto store = [&](auto unit_c) {
            constexpr bool kUnit = decltype(unit_c)::value;
            if (r < R) {
                const unsigned obase = unsigned(site) * site_bytes + unsigned(r) * 4u + unsigned(2 * grp) * row_bytes;
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    if (2 * grp + i < NC) {
#pragma unroll
                        for (int sp = 0; sp < 4; ++sp) {
                            const unsigned so = unsigned(sp * NC + i) * row_bytes;
                            float vr = d[sp][2 * i], vi = d[sp][2 * i + 1];
                            if constexpr (!kUnit) {
                                vr *= coef;
                                vi *= coef;
                            }
                            if constexpr (XP) {
                                vr += P::re(xw[i][sp]);
                                vi += P::im(xw[i][sp]);
                            }
                            const uint32_t ow = P::pack(vr, vi);
                            bstore(r_out, ow, obase, so);
                            if constexpr (RED) {
                                const float tr = P::re(ow), ti = P::im(ow);
                                const float ur = P::re(xw[i][sp]), ui = P::im(xw[i][sp]);
                                racc[0] = fmaf(ur, tr, fmaf(ui, ti, racc[0]));
                                racc[1] = fmaf(ur, ti, fmaf(-ui, tr, racc[1]));
                                racc[2] = fmaf(tr, tr, fmaf(ti, ti, racc[2]));
                            }
                        }
                    }
                }
            }
        };
        if (unit) store(std::true_type{});
        else store(std::false_type{});
        if (more) {
#pragma unroll
            for (int k = 0; k < 9; ++k) cur[k] = nxt[k];
        }
    }
    if constexpr (RED) {
        // the four lane groups of RHS m, then the workgroup's four wavefronts (fixed order)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            racc[k] += __shfl_xor(racc[k], 16);
            racc[k] += __shfl_xor(racc[k], 32);
        }
        __shared__ double red[4][16][3];
        if (lane < 16)
#pragma unroll
            for (int k = 0; k < 3; ++k) red[wv][lane][k] = double(racc[k]);
        __syncthreads();
        if (wv == 0 && lane < 16 && r0 + lane < R) {
            double *row = part + (long(blockIdx.x) * R + r0 + lane) * 4;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double t = ((red[0][lane][k] + red[1][lane][k]) + red[2][lane][k]) + red[3][lane][k];
                // omega wants <t, s> = conj(<s, t>)
                row[k] = k == 1 ? -t : t;
            }
            row[3] = 0.0;
        }
    }
}

// ---------------------------------------------------------------------------
// LDS-staged form (variant kDslashLds, Nc <= 4).  One workgroup = a segment of
// kLdsSites consecutive output sites of one x-row (fixed y, z; one wavefront
// per site) x a run of Ts consecutive t x one 16-RHS chunk.  It sweeps t and
// keeps in LDS
//   * a ring of three input rows (y, z, t-1), (y, z, t), (y, z, t+1), each the
//     segment plus one halo site: the x hops of layer t read the middle one,
//     the t hops the outer ones.  The x neighbours of sites xh = a..a+S-1 are
//     xh - 1 + p and xh + p, p = (y + z + t + parity) & 1 the row's
//     checkerboard phase, so a row is staged from site a - 1 + p on;
//   * the four input rows (y-1, z, t), (y+1, z, t), (y, z-1, t), (y, z+1, t)
//     over the segment, each serving one hop.
// The rows travel from global memory straight into LDS (buffer loads with the
// LDS flag: each lane's word lands at M0 + 4 lane), so staging holds no
// registers and a layer's loads are issued as one batch: a staged row is laid
// out in spin planes [spin][site][colour][rhs] of whole 64-word chunks, and
// every row offset that depends only on the workgroup's position is scalar.
// Per t step every wavefront reads its operand words for layer t from LDS, the
// workgroup synchronises, the next layer's rows are issued into the slots that
// just went out of use while the matrix-core work of layer t runs, and a second
// synchronisation after their arrival closes the step.  The arithmetic is that
// of k_dslash_mfma (about 5.25 input rows per output row are read against its
// 8 neighbour spinors per site); small workgroups let several share a CU, so
// one workgroup's memory wait overlaps another's arithmetic.
//
// Measured at 24^3 x 48, SU(3), 16 RHS, FP16 (Ts = 12): 20 % fewer L2 requests
// per site than k_dslash_mfma (39.7 against 49.4) but a lower L2 hit rate (43
// against 69 %) with workgroups ordered row by row; ordered in (y, z) blocks,
// 232-235 us against 236-254 us row by row and 238-250 us for k_dslash_mfma in
// the same processes -- within the spread between GPUs.  The per-layer step
// chains an LDS read phase, a link round trip and a workgroup barrier on the
// staged loads, which 5 workgroups per CU hide about as well as the free-running
// wavefronts of k_dslash_mfma hide their L1 misses; and the solver's Schur
// operator would lose k_dslash_mfma's fused omega reduction.  So this form is
// selected explicitly (qcuDslashTune / qcuDslashSetTuned / a tuning file), not
// by the autotuner.  A workgroup over the segments of two adjacent rows (y hops
// between them read from LDS: 4 input rows per output row instead of 5) was
// slower, 262 us against 245 us for k_dslash_mfma: only two such 8-wavefront
// workgroups fit a CU (83 VGPRs, 48 KB of LDS) against five of this one, and
// capping the registers for a third spills (303 us).
constexpr int kLdsSites = 4;

using lds_void = __attribute__((address_space(3))) void;

// one 64-word chunk: lane l's word of the buffer at voffset + soffset lands at dst[l]
__device__ __forceinline__ void lds_chunk(__amdgpu_buffer_rsrc_t r, uint32_t *dst, unsigned voff, unsigned soff) {
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (lds_void *)dst, 4, int(voff), int(soff), 0, 0);
}

typedef unsigned u32x8 __attribute__((ext_vector_type(8)));

// centre-row plane: chunks 0..CC-1 of this wavefront's spin plane (voffsets v[k], row offset sb)
template <int CC>
__device__ __forceinline__ void stage_center_plane(__amdgpu_buffer_rsrc_t r, uint32_t *dst, u32x8 v, unsigned sb) {
#pragma unroll
    for (int k = 0; k < CC; ++k) lds_chunk(r, dst + 64 * k, v[k], sb);
}

// side rows: 4 rows x 4 planes x NC chunks; wavefront wv takes jobs q = wv + 4 i (i < 4 NC), job q =
// (row, plane) q / NC, chunk q % NC, whose chunk (and so voffset sv[i % NC]) repeats with period NC in i.
// rb[row] = byte offset of the row's first site; rows are ROW_W words apart, planes PS.
template <int NC, int PS, int ROW_W>
__device__ __forceinline__ void stage_side_rows(__amdgpu_buffer_rsrc_t r, uint32_t *side, u32x4 sv, int wv,
                                                unsigned rb0, unsigned rb1, unsigned rb2, unsigned rb3,
                                                unsigned row_bytes) {
#pragma unroll
    for (int i = 0; i < 4 * NC; ++i) {
        const int q = wv + 4 * i;
        const int rp = q / NC, ch = q - rp * NC;
        const int row = rp >> 2, sp = rp & 3;
        const unsigned rb = row == 0 ? rb0 : (row == 1 ? rb1 : (row == 2 ? rb2 : rb3));
        lds_chunk(r, side + row * ROW_W + sp * PS + 64 * ch, sv[i % NC], rb + unsigned(sp * NC) * row_bytes);
    }
}

template <typename S, int NC, bool DAG, bool XP>
__global__ __launch_bounds__(64 * kLdsSites) void k_dslash_lds(uint32_t *__restrict__ out,
                                                              const uint32_t *__restrict__ in,
                                                              const uint32_t *__restrict__ gauge,
                                                              const uint32_t *__restrict__ xpay, float coef, Geom g,
                                                              int R, int parity, int Ts, int by, int bz,
                                                              unsigned vec_bytes, unsigned gauge_bytes) {
    static_assert(NC <= 4, "one k step per four hops");
    using P = Pk<S>;
    using PH = PkPhase<S>;
    constexpr int STEPS = 2, CPL = NC, SEG = kLdsSites;
    static_assert(SEG == 4, "one wavefront per spin plane of a staged row");
    constexpr int EW = NC * 16;                // entries (colour, rhs) of one site in one spin plane
    constexpr int CEN_E = (SEG + 1) * EW;      // entries of a centre-row plane (SIDE_E = SEG EW = 64 NC)
    constexpr int CC = (CEN_E + 63) / 64;      // 64-word chunks of a centre-row plane
    constexpr int PS = CC * 64;                // words of one spin plane (centre and side rows alike)
    constexpr int ROW_W = 4 * PS;              // words of one staged row
    __shared__ __align__(16) uint32_t side[4][ROW_W];  // (y-, y+, z-, z+) rows of the current layer
    __shared__ __align__(16) uint32_t cen[3][ROW_W];   // ring of centre rows
    const int Lxh = g.Lxh, Ly = g.L[1], Lz = g.L[2], Lt = g.L[3];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);  // site a + wv of the segment
    const int m = lane & 15, grp = lane >> 4;
    const int cb = m >> 1;
    const bool bre = (m & 1) == 0;
    const int cbc = cb < NC ? cb : NC - 1;
    const bool gb1 = grp >= 2, fwd = (grp & 1) == 0;  // this lane's hops: forward for even groups
    const int r0 = 16 * int(blockIdx.y);
    const int r = r0 + m;
    const unsigned row_bytes = unsigned(R) * 4u;
    const unsigned site_bytes = 4u * NC * row_bytes;
    const unsigned ll4 = 4u * NC * NC;
    const auto r_in = rsrc(in, vec_bytes);
    const auto r_g = rsrc(gauge, gauge_bytes);
    const auto r_out = rsrc(out, vec_bytes);
    const auto r_x = rsrc(xpay, xpay ? vec_bytes : 0u);

    // workgroup -> (x segment, y, z, t run): consecutive workgroups on one XCD take the segments of the
    // rows of a by x bz block of the (y, z) plane, then the next block, so the workgroups an XCD runs at
    // once stage each other's rows as side rows while those are in its L2
    const int nxs = (Lxh + SEG - 1) / SEG;
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int xs = b % nxs;
    int rest = b / nxs;
    const int iy = rest % by;
    rest /= by;
    const int iz = rest % bz;
    rest /= bz;
    const int nty = Ly / by, ntz = Lz / bz;
    const int y = (rest % nty) * by + iy;
    rest /= nty;
    const int z = (rest % ntz) * bz + iz;
    const int seg = rest / ntz;
    const int a0 = xs * SEG;
    const int t0 = seg * Ts, t1 = min(t0 + Ts, Lt);
    const int yp = y + 1 == Ly ? 0 : y + 1, ym = y == 0 ? Ly - 1 : y - 1;
    const int zp = z + 1 == Lz ? 0 : z + 1, zm = z == 0 ? Lz - 1 : z - 1;
    // (the helpers capture by value: a local whose address reached a closure would stay in scratch
    // memory next to the LDS-load intrinsic)
    auto row_base = [=](int yy, int zz, int tt) __attribute__((always_inline)) {  // first site index of a row
        return ((tt * Lz + zz) * Ly + yy) * Lxh;
    };
    auto wrap_t = [=](int tl) __attribute__((always_inline)) {  // tl in [-1, 2 Lt)
        return tl < 0 ? tl + Lt : (tl >= Lt ? tl - Lt : tl);
    };
    auto phase = [=](int tt) __attribute__((always_inline)) { return (y + z + tt + parity) & 1; };

    // hop constants of this lane (the same for every site)
    PH p0_s[STEPS], p1_s[STEPS], r0_s[STEPS], r1_s[STEPS];
#pragma unroll
    for (int st = 0; st < STEPS; ++st) {
        const int hh = st * 4 + grp;
        const u32x4 pm = *reinterpret_cast<const u32x4 *>(std::is_same<S, __half>::value
                                                              ? kPhaseTab.h[DAG ? 1 : 0][hh]
                                                              : kPhaseTab.m[DAG ? 1 : 0][hh]);
        p0_s[st] = PH::raw(pm[0]);
        p1_s[st] = PH::raw(pm[1]);
        r0_s[st] = PH::raw(pm[2]);
        r1_s[st] = PH::raw(pm[3]);
    }
    WordOp bop;  // B word from the stored link word (see k_dslash_mfma)
    bop.sel = bre ? kPermId : kPermRot;
    bop.mask = fwd ? (bre ? 0x80000000u : 0u) : (bre ? 0u : 0x00008000u);
    // link word offsets of this lane: forward V = U[cb][c'], backward V = conj(U[c'][cb])
    const unsigned lk_off = unsigned(cbc) * (fwd ? NC : 1) * 4u, lk_step = fwd ? 4u : 4u * NC;
    // partner spin planes: 3, 2 for the x and y hops (lane groups 0, 1), 2, 3 for z and t (groups 2, 3)
    const int ob0 = (gb1 ? 2 : 3) * PS, ob1 = (gb1 ? 3 : 2) * PS;

    // ---- staging.  Entry e = (site j, colour c, rhs mm) of a plane, e = (j NC + c) 16 + mm; lane l of a
    // chunk k loads entry 64 k + l.  RHS beyond R and entries beyond a plane read clamped, valid
    // addresses into words that are never used.
    const unsigned st_rm = unsigned(min(r0 + m, R - 1)) * 4u;
    auto entry_off = [=](int e) __attribute__((always_inline)) { return unsigned((e >> 4) % NC) * row_bytes + st_rm; };
    // centre rows: wavefront wv loads spin plane wv, chunks 0..CC-1; voffset for a row of phase 0 / 1
    // (vector registers: an index that becomes constant only after unrolling stays out of memory)
    static_assert(CC <= 8 && NC <= 4, "staging offsets fit their vectors");
    u32x8 cv0, cv1;
#pragma unroll
    for (int k = 0; k < CC; ++k) {
        const int e = min(64 * k + lane, CEN_E - 1);
        const int j = e / EW;
        cv0[k] = unsigned(((a0 - 1 + j) % Lxh + Lxh) % Lxh) * site_bytes + entry_off(e);
        cv1[k] = unsigned((a0 + j) % Lxh) * site_bytes + entry_off(e);
    }
    // side rows: 4 rows x 4 planes x NC chunks; wavefront wv takes jobs q = wv + 4 i (i < 4 NC), job q =
    // (row, plane) q / NC, chunk q % NC, whose chunk repeats with period NC in i
    u32x4 sv;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        const int e = 64 * ((wv + 4 * i) % NC) + lane;
        sv[i] = unsigned(min(a0 + e / EW, Lxh - 1)) * site_bytes + entry_off(e);
    }
    auto stage_center = [=](int tl, int slot) __attribute__((always_inline)) {
        const int tt = wrap_t(tl);
        const unsigned sb = unsigned(row_base(y, z, tt)) * site_bytes + unsigned(wv * NC) * row_bytes;
        stage_center_plane<CC>(r_in, &cen[slot][wv * PS], phase(tt) ? cv1 : cv0, sb);
    };
    auto stage_side = [=](int tl) __attribute__((always_inline)) {
        const int tt = wrap_t(tl);
        stage_side_rows<NC, PS, ROW_W>(r_in, &side[0][0], sv, wv, unsigned(row_base(ym, z, tt)) * site_bytes,
                                       unsigned(row_base(yp, z, tt)) * site_bytes,
                                       unsigned(row_base(y, zm, tt)) * site_bytes,
                                       unsigned(row_base(y, zp, tt)) * site_bytes, row_bytes);
    };
    // Workgroup barrier ordering LDS accesses only (no wait for outstanding global stores).  The
    // staged form first waits for this wavefront's loads into LDS (vmcnt(0) in the gfx9 encoding:
    // vmcnt 0, expcnt 7, lgkmcnt 15), as an instruction the compiler's own wait tracking sees.
    auto lds_barrier = []() __attribute__((always_inline)) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    };
    auto staged_barrier = [=]() __attribute__((always_inline)) {
        __builtin_amdgcn_s_waitcnt(0x0F70);
        lds_barrier();
    };
    auto read_barrier = [=]() __attribute__((always_inline)) { lds_barrier(); };

    // prologue: layers t0 - 1, t0, t0 + 1 into ring slots 0, 1, 2; side rows of layer t0 (one batch)
    stage_center(t0 - 1, 0);
    stage_center(t0, 1);
    stage_center(t0 + 1, 2);
    stage_side(t0);
    staged_barrier();

    const int xh = a0 + wv;
    const bool live = xh < Lxh;  // wavefront-uniform: a segment may run past the row
    // per-lane parts of the xpay / output offsets (the site part is scalar)
    const unsigned lane_out = unsigned(min(r, R - 1)) * 4u + unsigned(2 * grp) * row_bytes;
    for (int t = t0; t < t1; ++t) {
        const int sl_m = (t - t0) % 3, sl_0 = (t - t0 + 1) % 3, sl_p = (t - t0 + 2) % 3;
        const int tf = t + 1 == Lt ? 0 : t + 1, tb = t == 0 ? Lt - 1 : t - 1;
        // ---- operand words of layer t.  Step 0: lane groups 0, 1 the x hops (middle row, window site
        // wv + 1 forward / wv), 2, 3 the z hops (side rows); step 1: groups 0, 1 the y hops (side rows),
        // 2, 3 the t hops (outer rows, window site wv + 1 - p(t +- 1))
        uint32_t w0[STEPS][4], w1[STEPS][4], wb0[STEPS][4], wb1[STEPS][4];
        {
            const uint32_t *a_s0 = gb1 ? &side[fwd ? 3 : 2][wv * EW] : &cen[sl_0][(wv + (fwd ? 1 : 0)) * EW];
            const uint32_t *a_s1 = gb1 ? &cen[fwd ? sl_p : sl_m][(wv + 1 - phase(fwd ? tf : tb)) * EW]
                                       : &side[fwd ? 1 : 0][wv * EW];
#pragma unroll
            for (int st = 0; st < STEPS; ++st) {
                const uint32_t *src = (st == 0 ? a_s0 : a_s1) + m;
#pragma unroll
                for (int cc = 0; cc < 4; ++cc) {
                    if (cc < CPL) {
                        w0[st][cc] = src[cc * 16];
                        w1[st][cc] = src[PS + cc * 16];
                        wb0[st][cc] = src[ob0 + cc * 16];
                        wb1[st][cc] = src[ob1 + cc * 16];
                    } else {
                        w0[st][cc] = w1[st][cc] = wb0[st][cc] = wb1[st][cc] = 0u;
                    }
                }
            }
        }
        read_barrier();
        // ---- this site's links and xpay words, then the next layer's rows into the freed slots
        const int site = row_base(y, z, t) + xh;
        const unsigned sbase = unsigned(site) * site_bytes;
        uint32_t lw[STEPS][4];
        uint32_t xw[2][4];
        if (live) {
            const int x = 2 * xh + phase(t);
            const int xhb = x == 0 ? Lxh - 1 : (x - 1) >> 1;
            const unsigned vh = unsigned(g.Vh);
            const unsigned f0 = unsigned(parity) * vh + unsigned(site), b0 = unsigned(1 - parity) * vh;
            // forward U_mu(x) at (mu, parity, site); backward U_mu(x - mu) at (mu, 1 - parity, neighbour)
            const unsigned gx = fwd ? f0 : b0 + unsigned(row_base(y, z, t) + xhb);
            const unsigned gz = 4u * vh + (fwd ? f0 : b0 + unsigned(row_base(y, zm, t) + xh));
            const unsigned gy = 2u * vh + (fwd ? f0 : b0 + unsigned(row_base(ym, z, t) + xh));
            const unsigned gt = 6u * vh + (fwd ? f0 : b0 + unsigned(row_base(y, z, tb) + xh));
            const unsigned gl[STEPS] = {gb1 ? gz : gx, gb1 ? gt : gy};
#pragma unroll
            for (int st = 0; st < STEPS; ++st) {
                const unsigned gaddr = gl[st] * ll4 + lk_off;
#pragma unroll
                for (int cc = 0; cc < 4; ++cc)
                    lw[st][cc] = cc < CPL ? bload(r_g, gaddr + unsigned(cc) * lk_step, 0u) : 0u;
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int sp = 0; sp < 4; ++sp)
                    xw[i][sp] = (XP && 2 * grp + i < NC) ? bload(r_x, lane_out, sbase + unsigned(sp * NC + i) * row_bytes) : 0u;
        }
        // issued on the last step of a run too (rows of the wrapped layer, never read): a conditional
        // issue would make the wait for this step's links also wait for these loads
        stage_center(t + 2, sl_m);
        stage_side(t + 1);
        uint32_t ow[2][4];
        if (live) {
            __builtin_amdgcn_s_setprio(1);
            f32x4 d[4];
#pragma unroll
            for (int sp = 0; sp < 4; ++sp) d[sp] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int st = 0; st < STEPS; ++st) {
                const bool odd = st == 0;  // step 0 holds the x and z hops (odd phases)
                u32x4 av[4], bf;
#pragma unroll
                for (int cc = 0; cc < 4; ++cc) {
                    if (cc >= CPL) {
                        av[0][cc] = av[1][cc] = av[2][cc] = av[3][cc] = 0u;
                        bf[cc] = 0u;
                        continue;
                    }
                    uint32_t h0, h1, g0, g1;
                    if (odd) {
                        h0 = p0_s[st].template fma<true>(w0[st][cc], wb0[st][cc]);
                        h1 = p1_s[st].template fma<true>(w1[st][cc], wb1[st][cc]);
                        g0 = r0_s[st].template mul<true>(h0);
                        g1 = r1_s[st].template mul<true>(h1);
                    } else {
                        h0 = p0_s[st].template fma<false>(w0[st][cc], wb0[st][cc]);
                        h1 = p1_s[st].template fma<false>(w1[st][cc], wb1[st][cc]);
                        g0 = r0_s[st].template mul<false>(h0);
                        g1 = r1_s[st].template mul<false>(h1);
                    }
                    av[0][cc] = h0;
                    av[1][cc] = h1;
                    av[2][cc] = gb1 ? g0 : g1;  // spin m(0) is 2 for the z and t hops
                    av[3][cc] = gb1 ? g1 : g0;
                    bf[cc] = apply(bop, lw[st][cc]);
                }
#pragma unroll
                for (int sp = 0; sp < 4; ++sp) d[sp] = P::mma(bf, av[sp], d[sp]);
            }
            __builtin_amdgcn_s_setprio(0);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int sp = 0; sp < 4; ++sp) {
                    float vr = d[sp][2 * i] * coef, vi = d[sp][2 * i + 1] * coef;
                    if constexpr (XP) {
                        vr += P::re(xw[i][sp]);
                        vi += P::im(xw[i][sp]);
                    }
                    ow[i][sp] = P::pack(vr, vi);
                }
        }
        // the next layer is in LDS; then this layer's results go out while the next one is read
        staged_barrier();
        if (live && r < R) {
#pragma unroll
            for (int i = 0; i < 2; ++i)
                if (2 * grp + i < NC) {
#pragma unroll
                    for (int sp = 0; sp < 4; ++sp)
                        bstore(r_out, ow[i][sp], lane_out, sbase + unsigned(sp * NC + i) * row_bytes);
                }
        }
    }
}

template <typename S, int NC>
hipError_t launch_lds(const DslashArgs &a, const DslashTune &t, hipStream_t stream) {
    const int Ts = t.spw > 0 ? t.spw : 12;
    const int nseg = (a.g.L[3] + Ts - 1) / Ts;
    const int nxs = (a.g.Lxh + kLdsSites - 1) / kLdsSites;
    dim3 grid(nxs * a.g.L[1] * a.g.L[2] * nseg, div_up(a.R, 16)), block(64 * kLdsSites);
    auto out = static_cast<uint32_t *>(a.out);
    auto in = static_cast<const uint32_t *>(a.in);
    auto gauge = static_cast<const uint32_t *>(a.gauge);
    auto xp = static_cast<const uint32_t *>(a.xpay);
    const unsigned vec_bytes = unsigned(size_t(a.g.Vh) * 4 * a.Nc * a.R * 4);
    const unsigned gauge_bytes = unsigned(size_t(a.g.Vh) * 8 * a.Nc * a.Nc * 4);
    auto k = a.dagger ? (xp ? k_dslash_lds<S, NC, true, true> : k_dslash_lds<S, NC, true, false>)
                      : (xp ? k_dslash_lds<S, NC, false, true> : k_dslash_lds<S, NC, false, false>);
    // (y, z) block of the workgroup order: by | Ly, bz | Lz (else rows one by one)
    const int by = t.by > 0 && a.g.L[1] % t.by == 0 ? t.by : 1;
    const int bz = t.bz > 0 && a.g.L[2] % t.bz == 0 ? t.bz : 1;
    hipLaunchKernelGGL(k, grid, block, 0, stream, out, in, gauge, xp, float(a.coef), a.g, a.R, a.parity, Ts, by, bz,
                       vec_bytes, gauge_bytes);
    return hipGetLastError();
}

template <typename S, int NC>
hipError_t launch(const DslashArgs &a, const DslashTune &t, hipStream_t stream) {
    GeomTables tab;
    QCU_HIP_TRY(geom_tables(a.g, t.by, t.bz, tab, stream));
    // spw < 0: persistent grid of -spw workgroups per CU
    const int nblk = t.spw > 0 ? div_up(a.g.Vh, 4 * t.spw) : -t.spw * num_cus();
    dim3 grid(nblk, div_up(a.R, 16)), block(256);
    auto out = static_cast<uint32_t *>(a.out);
    auto in = static_cast<const uint32_t *>(a.in);
    auto gauge = static_cast<const uint32_t *>(a.gauge);
    auto xp = static_cast<const uint32_t *>(a.xpay);
    const unsigned vec_bytes = unsigned(size_t(a.g.Vh) * 4 * a.Nc * a.R * 4);
    const unsigned gauge_bytes = unsigned(size_t(a.g.Vh) * 8 * a.Nc * a.Nc * 4);
    // xpay is a separate instance: its prefetched words cost registers the plain form keeps
    auto k = a.dagger ? (xp ? k_dslash_mfma<S, NC, true, true> : k_dslash_mfma<S, NC, true, false>)
                      : (xp ? k_dslash_mfma<S, NC, false, true> : k_dslash_mfma<S, NC, false, false>);
    hipLaunchKernelGGL(k, grid, block, 0, stream, out, in, gauge, xp, float(a.coef), a.g.Vh, a.R, a.parity,
                       tab.walk, t.spw, vec_bytes, gauge_bytes, nullptr);
    return hipGetLastError();
}

template <typename S, int NC>
hipError_t launch_dot(const DslashArgs &a, const DslashTune &t, double *part, hipStream_t stream, int &nblk_out) {
    GeomTables tab;
    QCU_HIP_TRY(geom_tables(a.g, t.by, t.bz, tab, stream));
    const int nblk = t.spw > 0 ? div_up(a.g.Vh, 4 * t.spw) : -t.spw * num_cus();
    nblk_out = nblk;
    dim3 grid(nblk, div_up(a.R, 16)), block(256);
    auto out = static_cast<uint32_t *>(a.out);
    auto in = static_cast<const uint32_t *>(a.in);
    auto gauge = static_cast<const uint32_t *>(a.gauge);
    auto xp = static_cast<const uint32_t *>(a.xpay);
    const unsigned vec_bytes = unsigned(size_t(a.g.Vh) * 4 * a.Nc * a.R * 4);
    const unsigned gauge_bytes = unsigned(size_t(a.g.Vh) * 8 * a.Nc * a.Nc * 4);
    hipLaunchKernelGGL((k_dslash_mfma<S, NC, false, true, true>), grid, block, 0, stream, out, in, gauge, xp,
                       float(a.coef), a.g.Vh, a.R, a.parity, tab.walk, t.spw, vec_bytes, gauge_bytes, part);
    return hipGetLastError();
}

template <typename S>
hipError_t dispatch_dot(const DslashArgs &a, const DslashTune &t, double *part, hipStream_t stream, int &nblk) {
    switch (a.Nc) {
        case 2: return launch_dot<S, 2>(a, t, part, stream, nblk);
        case 3: return launch_dot<S, 3>(a, t, part, stream, nblk);
        case 4: return launch_dot<S, 4>(a, t, part, stream, nblk);
        case 8: return launch_dot<S, 8>(a, t, part, stream, nblk);
        default: return hipErrorInvalidValue;
    }
}

template <typename S>
hipError_t dispatch_nc(const DslashArgs &a, const DslashTune &t, hipStream_t stream) {
    switch (a.Nc) {
        case 2: return launch<S, 2>(a, t, stream);
        case 3: return launch<S, 3>(a, t, stream);
        case 4: return launch<S, 4>(a, t, stream);
        case 8: return launch<S, 8>(a, t, stream);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

bool dslash_lds_supported(const Geom &g, int Nc, int prec) {
    return (prec == kHalf || prec == kBF16) && Nc >= 2 && Nc <= 4 && g.Lxh >= 1 && g.L[1] >= 2 && g.L[2] >= 2 &&
           g.L[3] >= 2;
}

hipError_t launch_dslash_lds(const DslashArgs &a, const DslashTune &t, hipStream_t stream) {
    if (!dslash_lds_supported(a.g, a.Nc, a.prec) || t.spw < 0) return hipErrorInvalidValue;
    if (double(a.g.Vh) * 16.0 * a.Nc * a.R >= 2147483647.0 || double(a.g.Vh) * 32.0 * a.Nc * a.Nc >= 2147483647.0)
        return hipErrorInvalidValue;
    switch (a.Nc) {
        case 2: return a.prec == kHalf ? launch_lds<__half, 2>(a, t, stream) : launch_lds<__hip_bfloat16, 2>(a, t, stream);
        case 3: return a.prec == kHalf ? launch_lds<__half, 3>(a, t, stream) : launch_lds<__hip_bfloat16, 3>(a, t, stream);
        case 4: return a.prec == kHalf ? launch_lds<__half, 4>(a, t, stream) : launch_lds<__hip_bfloat16, 4>(a, t, stream);
        default: return hipErrorInvalidValue;
    }
}

bool dslash_mfma_supported(int Nc, int prec) {
    return (prec == kHalf || prec == kBF16) && (Nc == 2 || Nc == 3 || Nc == 4 || Nc == 8);
}

hipError_t launch_dslash_mfma_omega(const DslashArgs &a, const DslashTune &t, double *partials, hipStream_t stream,
                                    int &nblk) {
    if (!dslash_mfma_supported(a.Nc, a.prec) || a.dagger || !a.xpay) return hipErrorInvalidValue;
    if (t.by <= 0 || t.bz <= 0 || t.spw == 0 || t.spw < -64 || a.g.L[1] % t.by || a.g.L[2] % t.bz)
        return hipErrorInvalidValue;
    if (t.spw < 0 && (num_cus() * 4) % kNumXcd != 0) return hipErrorInvalidValue;
    if (double(a.g.Vh) * 16.0 * a.Nc * a.R >= 2147483647.0 || double(a.g.Vh) * 32.0 * a.Nc * a.Nc >= 2147483647.0)
        return hipErrorInvalidValue;
    return a.prec == kHalf ? dispatch_dot<__half>(a, t, partials, stream, nblk)
                           : dispatch_dot<__hip_bfloat16>(a, t, partials, stream, nblk);
}

hipError_t launch_dslash_mfma(const DslashArgs &a, const DslashTune &t, hipStream_t stream) {
    if (!dslash_mfma_supported(a.Nc, a.prec)) return hipErrorInvalidValue;
    if (t.by <= 0 || t.bz <= 0 || t.spw == 0 || t.spw < -64 || a.g.L[1] % t.by || a.g.L[2] % t.bz)
        return hipErrorInvalidValue;
    if (t.spw < 0 && (num_cus() * 4) % kNumXcd != 0) return hipErrorInvalidValue;
    // buffer resources address at most 2^31 - 1 bytes
    if (double(a.g.Vh) * 16.0 * a.Nc * a.R >= 2147483647.0 || double(a.g.Vh) * 32.0 * a.Nc * a.Nc >= 2147483647.0)
        return hipErrorInvalidValue;
    return a.prec == kHalf ? dispatch_nc<__half>(a, t, stream) : dispatch_nc<__hip_bfloat16>(a, t, stream);
}

}  // namespace qcu
