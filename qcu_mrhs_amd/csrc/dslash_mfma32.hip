// dslash_mfma32.hip -- Wilson hopping term on the FP32 and FP64 matrix cores
// (v_mfma_f32_16x16x4_f32 for FP32 storage, v_mfma_f64_16x16x4_f64 for FP64
// storage: the same 16 x 16 x 4 tile and operand layout), Nc = 2, 3, 4, 8.
//
// One wavefront computes one output site for a chunk of 16 right-hand sides
// (lane l: RHS m = l % 16, lane group g = l / 16).  Per hop and projected row
// (the two spin rows h_0, h_1 of the spin-projected neighbour spinor) the
// product V h is a real matrix product
//     P[(c, re|im)][rhs] = sum_{c'} A[(c, re|im)][(c', part)] B[(c', part)][rhs]
// with the link V in its real 2x2 embedding as the A operand and the projected
// spinor as B.  The instruction's K = 4 runs over four input colours of one
// component: one instruction takes the real parts of h (A = (Re V, Im V) in the
// (re, im) rows), a second the imaginary parts (A = (-Im V, Re V)), so each
// lane loads every complex number it needs once (input colour c' = 4 cg + g,
// output row m).  Operand layout of the instruction: lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16] and receives D[4 (l / 16) + v][l % 16]
// (FP32; the FP64 instruction returns D[l / 16 + 4 v][l % 16], and its A rows
// are permuted to match), v = 0..3, i.e. output colours 2g, 2g + 1 (re, im) of RHS m -- whole complex
// numbers per lane, so the reconstruction of the lower spins (a phase i^k per
// hop, known at compile time) is register arithmetic: every hop's product
// starts from zero and is added to the upper spin and, with its phase, to the
// lower spin it reconstructs.
//
// Entries beyond Nc (rows or input colours) and RHS beyond R read clamped,
// valid addresses and are zeroed or never stored.
#include <algorithm>
#include <type_traits>

#include "dslash_common.h"
#include "geom.h"

namespace qcu {
namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
template <typename T> struct Acc;
template <> struct Acc<float> { using type = f32x4; };
template <> struct Acc<double> { using type = f64x4; };

__device__ __forceinline__ f32x4 mma4(float a, float b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f64x4 mma4(double a, double b, f64x4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// z * i^k for compile-time k
template <int K, typename T>
__device__ __forceinline__ Cx<T> rot(Cx<T> z) {
    if constexpr ((K & 3) == 0) return z;
    else if constexpr ((K & 3) == 1) return {-z.im, z.re};
    else if constexpr ((K & 3) == 2) return {-z.re, -z.im};
    else return {z.im, -z.re};
}

constexpr int proj_k(int mu, int row, bool fwd, bool dag) {
    return (kProjK(mu, row) + ((fwd != dag) ? 2 : 0)) & 3;
}
constexpr int recon_k(int mu, int row, bool fwd, bool dag) {
    return (kReconK(mu, row) + ((fwd != dag) ? 2 : 0)) & 3;
}

template <int NC>
struct Mfma32Lane {
    int m, g;         // RHS slot and lane group
    int c, part;      // output colour / component of this lane's A row
    bool arow;        // the A row exists (c < NC)
};

// One hop (mu, FWD) of one site: accumulates the four spins of acc (lane layout of D).
template <typename T, int NC, bool DAG, int MU, bool FWD>
__device__ __forceinline__ void mfma32_hop(typename Acc<T>::type (&acc)[4], const Cx<T> *__restrict__ nbsp,
                                           const Cx<T> *__restrict__ link, const Mfma32Lane<NC> &ln, long R,
                                           long r) {
    using V = typename Acc<T>::type;
    constexpr int NCG = (NC + 3) / 4;
    constexpr int M0 = kMat2(MU, 0), M1 = kMat2(MU, 1);
    constexpr int KP0 = proj_k(MU, 0, FWD, DAG), KP1 = proj_k(MU, 1, FWD, DAG);
    constexpr int KR0 = recon_k(MU, 0, FWD, DAG), KR1 = recon_k(MU, 1, FWD, DAG);
    V p0 = {T(0), T(0), T(0), T(0)}, p1 = {T(0), T(0), T(0), T(0)};
#pragma unroll
    for (int cg = 0; cg < NCG; ++cg) {
        const int cp = 4 * cg + ln.g;
        const bool kin = cp < NC;
        const int cpc = kin ? cp : NC - 1;
        // projected rows h_0 = psi_0 + i^KP0 psi_M0, h_1 = psi_1 + i^KP1 psi_M1 for (c', rhs)
        const Cx<T> a0 = nbsp[(0 * NC + cpc) * R + r];
        const Cx<T> a1 = nbsp[(1 * NC + cpc) * R + r];
        const Cx<T> b0 = nbsp[(M0 * NC + cpc) * R + r];
        const Cx<T> b1 = nbsp[(M1 * NC + cpc) * R + r];
        Cx<T> h0 = a0 + rot<KP0>(b0), h1 = a1 + rot<KP1>(b1);
        // link V[c][c']: forward U[c][c'], backward conj(U[c'][c])
        Cx<T> v = FWD ? link[ln.c * NC + cpc] : link[cpc * NC + ln.c];
        if (!FWD) v.im = -v.im;
        T are = ln.part ? v.im : v.re;   // multiplies Re h
        T aim = ln.part ? v.re : -v.im;  // multiplies Im h
        if (!(kin && ln.arow)) are = aim = T(0);
        if (!kin) h0 = h1 = {T(0), T(0)};
        p0 = mma4(are, h0.re, p0);
        p0 = mma4(aim, h0.im, p0);
        p1 = mma4(are, h1.re, p1);
        p1 = mma4(aim, h1.im, p1);
    }
    // upper spins take the products, the lower ones their reconstruction phases
    acc[0] += p0;
    acc[1] += p1;
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const Cx<T> q0 = rot<KR0>(Cx<T>{p0[2 * e], p0[2 * e + 1]});
        const Cx<T> q1 = rot<KR1>(Cx<T>{p1[2 * e], p1[2 * e + 1]});
        acc[M0][2 * e] += q0.re;
        acc[M0][2 * e + 1] += q0.im;
        acc[M1][2 * e] += q1.re;
        acc[M1][2 * e + 1] += q1.im;
    }
}

template <typename T, int NC>
__device__ __forceinline__ Mfma32Lane<NC> mfma32_lane() {
    const int lane = threadIdx.x & 63;
    Mfma32Lane<NC> ln;
    ln.m = lane & 15;
    ln.g = lane >> 4;
    if constexpr (std::is_same<T, double>::value) {
        // FP64 instruction: lane l receives D[l / 16 + 4 v][l % 16] (measured on gfx950), so A row m
        // holds (colour 2 (m & 3) + (m >> 3), part (m >> 2) & 1) to land colours 2g, 2g + 1 (re, im) in
        // v = 0..3 of lane group g, as the FP32 layout does with A row m = (colour m >> 1, part m & 1)
        ln.c = 2 * (ln.m & 3) + (ln.m >> 3);
        ln.part = (ln.m >> 2) & 1;
    } else {
        ln.c = ln.m >> 1;
        ln.part = ln.m & 1;
    }
    ln.arow = ln.c < NC;
    if (!ln.arow) ln.c = NC - 1;  // clamped address, A operand zeroed
    return ln;
}

// Sites q of the traversal handled by this wavefront: spw consecutive positions starting at
// (b * 4 + wavefront) * spw; every chunk of 16 RHS of each site.
template <typename T, int NC, bool DAG>
__device__ __forceinline__ void mfma32_block(int b, Cx<T> *__restrict__ out, const Cx<T> *__restrict__ in,
                                             const Cx<T> *__restrict__ gauge, const Cx<T> *__restrict__ xpay,
                                             T coef, int Vh, int R, int parity, const int *__restrict__ walk,
                                             int spw) {
    using V = typename Acc<T>::type;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const Mfma32Lane<NC> ln = mfma32_lane<T, NC>();
    const long site_len = 4L * NC * R;
    const long link_len = long(NC) * NC;
    const int *wtab = walk + long(parity) * Vh * 9;
    const int q0 = (b * 4 + wv) * spw;
    for (int j = 0; j < spw; ++j) {
        const int q = q0 + j;
        if (q >= Vh) break;  // uniform
        int nbr[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) nbr[k] = __builtin_amdgcn_readfirstlane(wtab[long(q) * 9 + k]);
        const int site = nbr[0];
        for (int r0 = 0; r0 < R; r0 += 16) {
            const long r = min(r0 + ln.m, R - 1);
            V acc[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) acc[s] = V{T(0), T(0), T(0), T(0)};
#define QCU_HOP(MU)                                                                                              \
    mfma32_hop<T, NC, DAG, MU, true>(acc, in + nbr[1 + 2 * MU] * site_len,                                       \
                                     gauge + ((MU * 2 + parity) * long(Vh) + site) * link_len, ln, R, r);        \
    mfma32_hop<T, NC, DAG, MU, false>(acc, in + nbr[2 + 2 * MU] * site_len,                                      \
                                      gauge + ((MU * 2 + 1 - parity) * long(Vh) + nbr[2 + 2 * MU]) * link_len, ln, R, r);
            QCU_HOP(0)
            QCU_HOP(1)
            QCU_HOP(2)
            QCU_HOP(3)
#undef QCU_HOP
            // lane (m, g) holds colours 2g, 2g + 1 of RHS r0 + m for every spin
            if (r0 + ln.m < R) {
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    const int c = 2 * ln.g + e;
                    if (c < NC) {
#pragma unroll
                        for (int s = 0; s < 4; ++s) {
                            const long o = site * site_len + (long(s) * NC + c) * R + r;
                            Cx<T> v = {coef * acc[s][2 * e], coef * acc[s][2 * e + 1]};
                            if (xpay) v = v + xpay[o];
                            out[o] = v;
                        }
                    }
                }
            }
        }
    }
}

template <typename T, int NC, bool DAG>
__global__ __launch_bounds__(256) void k_dslash_mfma32(Cx<T> *__restrict__ out, const Cx<T> *__restrict__ in,
                                                      const Cx<T> *__restrict__ gauge, const Cx<T> *__restrict__ xpay,
                                                      T coef, int Vh, int R, int parity, const int *__restrict__ walk,
                                                      int spw) {
    mfma32_block<T, NC, DAG>(xcd_remap(blockIdx.x, gridDim.x), out, in, gauge, xpay, coef, Vh, R, parity, walk, spw);
}

// every group of a batch in one launch: blockIdx.y = group
template <typename T, int NC, bool DAG>
__global__ __launch_bounds__(256) void k_dslash_mfma32_groups(Cx<T> *__restrict__ out, const Cx<T> *__restrict__ in,
                                                             const Cx<T> *__restrict__ gauge,
                                                             const Cx<T> *__restrict__ xpay, T coef, int Vh,
                                                             int parity, const int *__restrict__ walk, SegTab tab) {
    const int grp = blockIdx.y;
    if (tab.done && tab.done[grp]) return;
    const int spw = tab.spw[grp];
    const int nblk = (Vh + 4 * spw - 1) / (4 * spw);
    if (int(blockIdx.x) >= nblk) return;
    const long off = tab.n * tab.base[grp];
    mfma32_block<T, NC, DAG>(xcd_remap(blockIdx.x, nblk), out + off, in + off, gauge, xpay ? xpay + off : nullptr,
                             coef, Vh, tab.R[grp], parity, walk, spw);
}

// Every group of a batch in one launch with the groups' RHS packed into shared 16-RHS chunks
// (blockIdx.y = chunk): lane m of chunk k takes batch RHS q = 16 k + m, i.e. RHS q - base[g] of the
// group g holding q, at its own offset and stride in the group-concatenated [n][R_g] vectors.  Groups
// of 3, 8, 17 and 32 sources then fill 4 chunks per site instead of 6 (1 + 1 + 2 + 2), and the links
// and matrix-core work per site shrink with them.  Lanes of a finished group compute and store into
// its work vector (never read again); a chunk whose groups have all finished is skipped.
template <typename T, int NC, bool DAG>
__global__ __launch_bounds__(256) void k_dslash_mfma32_packed(Cx<T> *__restrict__ out, const Cx<T> *__restrict__ in,
                                                             const Cx<T> *__restrict__ gauge,
                                                             const Cx<T> *__restrict__ xpay, T coef, int Vh,
                                                             int parity, const int *__restrict__ walk, SegTab tab,
                                                             int spw) {
    using V = typename Acc<T>::type;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const Mfma32Lane<NC> ln = mfma32_lane<T, NC>();
    // this lane's batch RHS and its group
    const int qr = 16 * int(blockIdx.y) + ln.m;
    const bool valid = qr < tab.Rtot;
    const int q = valid ? qr : tab.Rtot - 1;
    int grp = 0;
#pragma unroll
    for (int k = 1; k < kMaxGroups; ++k)
        if (k < tab.G && tab.base[k] <= q) grp = k;
    const bool active = !(tab.done && tab.done[grp]);
    if (!__any(valid && active)) return;  // wavefront-uniform
    const long Rg = tab.R[grp];
    const long lane_off = tab.n * tab.base[grp] + (q - tab.base[grp]);  // complex entries
    const long site_len = 4L * NC * Rg;  // of this lane's group
    const long link_len = long(NC) * NC;
    const int *wtab = walk + long(parity) * Vh * 9;
    const int q0 = (xcd_remap(blockIdx.x, gridDim.x) * 4 + wv) * spw;
    const Cx<T> *inl = in + lane_off;
    for (int j = 0; j < spw; ++j) {
        const int pos = q0 + j;
        if (pos >= Vh) break;  // uniform
        int nbr[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) nbr[k] = __builtin_amdgcn_readfirstlane(wtab[long(pos) * 9 + k]);
        const int site = nbr[0];
        V acc[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) acc[s] = V{T(0), T(0), T(0), T(0)};
#define QCU_HOP(MU)                                                                                              \
    mfma32_hop<T, NC, DAG, MU, true>(acc, inl + nbr[1 + 2 * MU] * site_len,                                      \
                                     gauge + ((MU * 2 + parity) * long(Vh) + site) * link_len, ln, Rg, 0);      \
    mfma32_hop<T, NC, DAG, MU, false>(acc, inl + nbr[2 + 2 * MU] * site_len,                                     \
                                      gauge + ((MU * 2 + 1 - parity) * long(Vh) + nbr[2 + 2 * MU]) * link_len, ln, Rg, 0);
        QCU_HOP(0)
        QCU_HOP(1)
        QCU_HOP(2)
        QCU_HOP(3)
#undef QCU_HOP
        if (valid) {
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const int c = 2 * ln.g + e;
                if (c < NC) {
#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        const long o = lane_off + site * site_len + (long(s) * NC + c) * Rg;
                        Cx<T> v = {coef * acc[s][2 * e], coef * acc[s][2 * e + 1]};
                        if (xpay) v = v + xpay[o];
                        out[o] = v;
                    }
                }
            }
        }
    }
}

template <typename T, int NC>
hipError_t launch_nc(const DslashArgs &a, const DslashTune &t, hipStream_t stream) {
    GeomTables tab;
    QCU_HIP_TRY(geom_tables(a.g, t.by, t.bz, tab, stream));
    dim3 grid(div_up(a.g.Vh, 4 * t.spw)), block(256);
    auto out = static_cast<Cx<T> *>(a.out);
    auto in = static_cast<const Cx<T> *>(a.in);
    auto gauge = static_cast<const Cx<T> *>(a.gauge);
    auto xp = static_cast<const Cx<T> *>(a.xpay);
    auto k = a.dagger ? k_dslash_mfma32<T, NC, true> : k_dslash_mfma32<T, NC, false>;
    hipLaunchKernelGGL(k, grid, block, 0, stream, out, in, gauge, xp, T(a.coef), a.g.Vh, a.R, a.parity, tab.walk,
                       t.spw);
    return hipGetLastError();
}

template <typename T, int NC>
hipError_t launch_groups_nc(const DslashArgs &a, const SegTab &seg, int by, int bz, hipStream_t stream) {
    GeomTables tab;
    QCU_HIP_TRY(geom_tables(a.g, by, bz, tab, stream));
    int nmax = 1;
    for (int k = 0; k < seg.G; ++k) nmax = std::max(nmax, div_up(a.g.Vh, 4 * seg.spw[k]));
    dim3 grid(nmax, seg.G), block(256);
    auto out = static_cast<Cx<T> *>(a.out);
    auto in = static_cast<const Cx<T> *>(a.in);
    auto gauge = static_cast<const Cx<T> *>(a.gauge);
    auto xp = static_cast<const Cx<T> *>(a.xpay);
    auto k = a.dagger ? k_dslash_mfma32_groups<T, NC, true> : k_dslash_mfma32_groups<T, NC, false>;
    hipLaunchKernelGGL(k, grid, block, 0, stream, out, in, gauge, xp, T(a.coef), a.g.Vh, a.parity, tab.walk, seg);
    return hipGetLastError();
}

template <typename T, int NC>
hipError_t launch_packed_nc(const DslashArgs &a, const SegTab &seg, int by, int bz, int spw, hipStream_t stream) {
    GeomTables tab;
    QCU_HIP_TRY(geom_tables(a.g, by, bz, tab, stream));
    dim3 grid(div_up(a.g.Vh, 4 * spw), div_up(seg.Rtot, 16)), block(256);
    auto out = static_cast<Cx<T> *>(a.out);
    auto in = static_cast<const Cx<T> *>(a.in);
    auto gauge = static_cast<const Cx<T> *>(a.gauge);
    auto xp = static_cast<const Cx<T> *>(a.xpay);
    auto k = a.dagger ? k_dslash_mfma32_packed<T, NC, true> : k_dslash_mfma32_packed<T, NC, false>;
    hipLaunchKernelGGL(k, grid, block, 0, stream, out, in, gauge, xp, T(a.coef), a.g.Vh, a.parity, tab.walk, seg,
                       spw);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_packed_prec(const DslashArgs &a, const SegTab &tab, int by, int bz, int spw, hipStream_t stream) {
    switch (a.Nc) {
        case 2: return launch_packed_nc<T, 2>(a, tab, by, bz, spw, stream);
        case 3: return launch_packed_nc<T, 3>(a, tab, by, bz, spw, stream);
        case 4: return launch_packed_nc<T, 4>(a, tab, by, bz, spw, stream);
        case 8: return launch_packed_nc<T, 8>(a, tab, by, bz, spw, stream);
        default: return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t launch_prec(const DslashArgs &a, const DslashTune &t, hipStream_t stream) {
    switch (a.Nc) {
        case 2: return launch_nc<T, 2>(a, t, stream);
        case 3: return launch_nc<T, 3>(a, t, stream);
        case 4: return launch_nc<T, 4>(a, t, stream);
        case 8: return launch_nc<T, 8>(a, t, stream);
        default: return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t launch_groups_prec(const DslashArgs &a, const SegTab &tab, int by, int bz, hipStream_t stream) {
    switch (a.Nc) {
        case 2: return launch_groups_nc<T, 2>(a, tab, by, bz, stream);
        case 3: return launch_groups_nc<T, 3>(a, tab, by, bz, stream);
        case 4: return launch_groups_nc<T, 4>(a, tab, by, bz, stream);
        case 8: return launch_groups_nc<T, 8>(a, tab, by, bz, stream);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

bool dslash_mfma32_supported(int Nc, int prec) {
    return prec == kSingle && (Nc == 2 || Nc == 3 || Nc == 4 || Nc == 8);
}

bool dslash_mfma64_supported(int Nc, int prec) {
    return prec == kDouble && (Nc == 2 || Nc == 3 || Nc == 4 || Nc == 8);
}

hipError_t launch_dslash_mfma32(const DslashArgs &a, const DslashTune &t, hipStream_t stream) {
    if (!dslash_mfma32_supported(a.Nc, a.prec) && !dslash_mfma64_supported(a.Nc, a.prec)) return hipErrorInvalidValue;
    if (t.by <= 0 || t.bz <= 0 || t.spw <= 0 || a.g.L[1] % t.by || a.g.L[2] % t.bz) return hipErrorInvalidValue;
    return a.prec == kDouble ? launch_prec<double>(a, t, stream) : launch_prec<float>(a, t, stream);
}

hipError_t launch_dslash_mfma32_groups(const DslashArgs &a, const SegTab &tab, int by, int bz, hipStream_t stream) {
    if (!dslash_mfma32_supported(a.Nc, a.prec) && !dslash_mfma64_supported(a.Nc, a.prec)) return hipErrorInvalidValue;
    if (tab.G < 1 || tab.G > kMaxGroups || by <= 0 || bz <= 0 || a.g.L[1] % by || a.g.L[2] % bz)
        return hipErrorInvalidValue;
    for (int k = 0; k < tab.G; ++k)
        if (tab.spw[k] <= 0 || tab.R[k] < 1 || !dslash_supported(a.Nc, tab.R[k])) return hipErrorInvalidValue;
    return a.prec == kDouble ? launch_groups_prec<double>(a, tab, by, bz, stream)
                             : launch_groups_prec<float>(a, tab, by, bz, stream);
}

hipError_t launch_dslash_mfma32_packed(const DslashArgs &a, const SegTab &tab, int by, int bz, int spw,
                                       hipStream_t stream) {
    if (!dslash_mfma32_supported(a.Nc, a.prec) && !dslash_mfma64_supported(a.Nc, a.prec)) return hipErrorInvalidValue;
    if (tab.G < 1 || tab.G > kMaxGroups || tab.Rtot < 1 || spw <= 0 || by <= 0 || bz <= 0 || a.g.L[1] % by ||
        a.g.L[2] % bz)
        return hipErrorInvalidValue;
    for (int k = 0; k < tab.G; ++k)
        if (tab.R[k] < 1 || !dslash_supported(a.Nc, tab.R[k]) || (k > 0 && tab.base[k] != tab.base[k - 1] + tab.R[k - 1]))
            return hipErrorInvalidValue;
    return a.prec == kDouble ? launch_packed_prec<double>(a, tab, by, bz, spw, stream)
                             : launch_packed_prec<float>(a, tab, by, bz, spw, stream);
}

}  // namespace qcu
