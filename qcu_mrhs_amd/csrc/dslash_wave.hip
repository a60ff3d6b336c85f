// dslash_wave.hip -- wave-per-site Wilson hopping kernel (VALU).
//
// One wavefront computes one output site (looping over `spw` sites).  Its 64
// lanes are 4 groups x 16 right-hand sides: group mu handles the two hops
// along dimension mu (forward U_mu(x) psi(x+mu), backward U_mu(x-mu)^+ psi(x-mu))
// for its RHS, so every lane loads two neighbour spinors and every load
// instruction reads four contiguous 16-RHS rows.  The four partial sums are
// then reduced with a transposing butterfly (lane xor 32, then xor 16) that
// leaves output spin s in lane group s, so the store is one contiguous 16-RHS
// row per (spin, colour).  Only a few sites per XCD are in flight, and the
// traversal (tiles in y,z swept along t, contiguous per XCD) keeps every
// neighbour reuse within the XCD's L2.
#include <algorithm>

#include "dslash_common.h"

namespace qcu {
namespace {

template <typename T>
__device__ __forceinline__ Cx<T> shfl_xor_cx(Cx<T> v, int mask) {
    return {__shfl_xor(v.re, mask), __shfl_xor(v.im, mask)};
}

// Per-lane constants of the hops of one direction in arithmetic form:
// z * i^k = (c re - s im, s re + c im) with (c, s) = (1, 0), (0, 1), (-1, 0),
// (0, -1) for k = 0..3, and the projection's partner rows as element offsets.
// Multiplying by (c, s) costs four FMAs where a run-time select of the 16-byte
// FP64 components costs about ten VALU operations.  The constants are those of
// the forward hop; every factor of the backward hop is the negative (its sign
// exponent differs by 2), applied as an FMA operand modifier.
template <typename T>
struct HopMul {
    T cp0, sp0, cp1, sp1;  // projection factors of rows 0 and 1
    T cr0, sr0, cr1, sr1;  // reconstruction factors
    int ob0, ob1;          // element offsets of the partner rows of rows 0 and 1
};

template <typename T>
__device__ __forceinline__ void ipow_cs(int k, T &c, T &s) {
    c = (k & 1) ? T(0) : ((k & 2) ? T(-1) : T(1));
    s = (k & 1) ? ((k & 2) ? T(-1) : T(1)) : T(0);
}

template <typename T, int NC>
__device__ __forceinline__ HopMul<T> hop_mul(const HopInfo &h, int R) {
    HopMul<T> m;
    ipow_cs(h.kp0, m.cp0, m.sp0);
    ipow_cs(h.kp1, m.cp1, m.sp1);
    ipow_cs(h.kr0, m.cr0, m.sr0);
    ipow_cs(h.kr1, m.cr1, m.sr1);
    m.ob0 = (h.m0_is_2 ? 2 : 3) * NC * R;
    m.ob1 = (h.m0_is_2 ? 3 : 2) * NC * R;
    return m;
}

// a + (c, s) * b
template <typename T>
__device__ __forceinline__ Cx<T> add_rot(Cx<T> a, Cx<T> b, T c, T s) {
    return {fma(c, b.re, fma(-s, b.im, a.re)), fma(s, b.re, fma(c, b.im, a.im))};
}

// One hop of this lane's direction.  acc[0], acc[1] collect the upper spins,
// acc[2], acc[3] the reconstructed rows in projection order (partner of row 0,
// partner of row 1); the caller maps them to spins 2 and 3 once per site.
template <typename T, typename CS, int NC, bool FWD>
__device__ __forceinline__ void hop_rt(Cx<T> (&acc)[4][NC], const CS *__restrict__ psi, int R,
                                       const CS *__restrict__ U, const HopMul<T> &h) {
    Cx<T> h0[NC], h1[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const Cx<T> a0 = load_cx<T>(psi + (0 * NC + c) * R);
        const Cx<T> a1 = load_cx<T>(psi + (1 * NC + c) * R);
        const Cx<T> b0 = load_cx<T>(psi + h.ob0 + c * R);
        const Cx<T> b1 = load_cx<T>(psi + h.ob1 + c * R);
        constexpr T sg = FWD ? T(1) : T(-1);
        h0[c] = add_rot(a0, b0, sg * h.cp0, sg * h.sp0);
        h1[c] = add_rot(a1, b1, sg * h.cp1, sg * h.sp1);
    }
#pragma unroll
    for (int a = 0; a < NC; ++a) {
        Cx<T> s0 = {T(0), T(0)}, s1 = {T(0), T(0)};
#pragma unroll
        for (int b = 0; b < NC; ++b) {
            if constexpr (FWD) {
                const Cx<T> u = load_cx<T>(U + a * NC + b);
                cfma(s0, u, h0[b]);
                cfma(s1, u, h1[b]);
            } else {
                const Cx<T> u = load_cx<T>(U + b * NC + a);  // (U^+)[a][b] = conj(U[b][a])
                cfma_conj(s0, u, h0[b]);
                cfma_conj(s1, u, h1[b]);
            }
        }
        acc[0][a] = acc[0][a] + s0;
        acc[1][a] = acc[1][a] + s1;
        constexpr T sg = FWD ? T(1) : T(-1);
        acc[2][a] = add_rot(acc[2][a], s0, sg * h.cr0, sg * h.sr0);
        acc[3][a] = add_rot(acc[3][a], s1, sg * h.cr1, sg * h.sr1);
    }
}

// Workgroup body: 4 wavefronts x spw sites, starting at traversal position
// 4 * spw * b.
// Epilogue of the plain operator: out = coef D in (+ xpay).
struct StoreEpi {
    static constexpr bool kResidual = false;
};
// Fused residual epilogue (FP64 only): v = xpay + coef D in is not stored;
// instead d = b - v is rounded to the inner storage as lo = scale_r d, and
// |d|^2 is summed per RHS into the workgroup's LDS row of this wavefront
// (slot[wave][R], one writer per entry, so the order of the sums is fixed).
template <typename LO>
struct ResidualEpi {
    static constexpr bool kResidual = true;
    const Cx<double> *b;
    LO *lo;
    const double *scale;
    double *slot;
};

template <typename T, typename CS, int NC, bool DAG, typename EPI = StoreEpi>
__device__ __forceinline__ void wave_block(int b, CS *__restrict__ out, const CS *__restrict__ in,
                                           const CS *__restrict__ gauge, const CS *__restrict__ xpay,
                                           T coef, const Geom &g, int R, int parity, int by, int bz,
                                           int spw, const EPI &ep = EPI{}) {
    const int lane = threadIdx.x & 63;
    // wavefront-uniform: the site's coordinates and neighbours go to the scalar unit
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int mu = lane >> 4;
    const int rl = lane & 15;
    const long site_len = 4L * NC * R;
    const long link_len = long(NC) * NC;
    const long Vh = g.Vh;
    const HopInfo hif = hop_info(mu, true, DAG);
    const HopMul<T> hm = hop_mul<T, NC>(hif, R);  // forward; backward = negated factors
    const bool m02 = hif.m0_is_2;  // the same for both hops of a direction
    const NbConst ncf = nb_const(g, mu, true), ncb = nb_const(g, mu, false);
    const bool lo = lane < 32, odd = (mu & 1) != 0;

    for (int j = 0; j < spw; ++j) {
        const int q = (b * spw + j) * 4 + wv;
        if (q >= g.Vh) break;  // uniform within the wavefront
        const int site = traverse(q, g, by, bz);
        const Site s = site_coords(site, g, parity);
        const int nf = neighbour_bl(site, s, g, ncf);
        const int nb = neighbour_bl(site, s, g, ncb);
        const CS *Uf = gauge + ((mu * 2 + parity) * Vh + site) * link_len;
        const CS *Ub = gauge + ((mu * 2 + 1 - parity) * Vh + nb) * link_len;
        for (int r0 = 0; r0 < R; r0 += 16) {
            const int r = r0 + rl;
            const bool act = r < R;
            Cx<T> acc[4][NC];
#pragma unroll
            for (int sp = 0; sp < 4; ++sp)
#pragma unroll
                for (int c = 0; c < NC; ++c) acc[sp][c] = {T(0), T(0)};
            if (act) {
                hop_rt<T, CS, NC, true>(acc, in + nf * site_len + r, R, Uf, hm);
                hop_rt<T, CS, NC, false>(acc, in + nb * site_len + r, R, Ub, hm);
            }
            // reconstructed rows -> spins 2, 3: row 0 pairs with spin 2 for z, t hops, with spin 3 for x, y
            if (!m02) {
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    const Cx<T> t = acc[2][c];
                    acc[2][c] = acc[3][c];
                    acc[3][c] = t;
                }
            }
            // lanes < 32 keep spins {0,1}, lanes >= 32 keep spins {2,3}
            Cx<T> kept[2][NC];
#pragma unroll
            for (int s2 = 0; s2 < 2; ++s2)
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    const Cx<T> send = lo ? acc[2 + s2][c] : acc[s2][c];
                    const Cx<T> recv = shfl_xor_cx(send, 32);
                    kept[s2][c] = (lo ? acc[s2][c] : acc[2 + s2][c]) + recv;
                }
            // lane group mu keeps spin mu
            Cx<T> fin[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                const Cx<T> send = odd ? kept[0][c] : kept[1][c];
                const Cx<T> recv = shfl_xor_cx(send, 16);
                fin[c] = (odd ? kept[1][c] : kept[0][c]) + recv;
            }
            if constexpr (EPI::kResidual) {
                double a2 = 0.0;
                if (act) {
                    const long base = (site * 4L + mu) * NC * R + r;
                    const double sc = ep.scale[r];
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
                        Cx<double> v = load_cx<double>(xpay + base + c * R);
                        v.re = fma(coef, fin[c].re, v.re);
                        v.im = fma(coef, fin[c].im, v.im);
                        const Cx<double> d = ep.b[base + c * R] - v;
                        a2 = fma(d.re, d.re, fma(d.im, d.im, a2));
                        store_cx(ep.lo + base + c * R, sc * d);
                    }
                }
                // the 4 spins of one RHS sit in lane groups mu = 0..3
                a2 += __shfl_xor(a2, 16);
                a2 += __shfl_xor(a2, 32);
                if (mu == 0 && act) ep.slot[wv * R + r] += a2;
            } else if (act) {
                const long base = (site * 4L + mu) * NC * R + r;
                if (xpay) {
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
                        Cx<T> v = load_cx<T>(xpay + base + c * R);
                        v.re = fma(coef, fin[c].re, v.re);
                        v.im = fma(coef, fin[c].im, v.im);
                        store_cx(out + base + c * R, v);
                    }
                } else {
#pragma unroll
                    for (int c = 0; c < NC; ++c) store_cx(out + base + c * R, coef * fin[c]);
                }
            }
        }
    }
}

template <typename T, typename CS, int NC, bool DAG>
__global__ __launch_bounds__(256) void k_dslash_wave(CS *__restrict__ out, const CS *__restrict__ in,
                                                    const CS *__restrict__ gauge,
                                                    const CS *__restrict__ xpay, T coef, Geom g, int R,
                                                    int parity, int by, int bz, int spw) {
    wave_block<T, CS, NC, DAG>(xcd_remap(blockIdx.x, gridDim.x), out, in, gauge, xpay, coef, g, R, parity,
                               by, bz, spw);
}

// every group of a batch in one launch: blockIdx.y = group
// Fused residual form of the second Schur hop (FP64): lo = scale_r (b - (xpay + coef D in)),
// partials[block][r] = sum over the block's sites of |b - v|^2 for RHS r.  Dynamic LDS: 4 R doubles.
template <int NC, typename LO>
__global__ __launch_bounds__(256) void k_dslash_wave_residual(LO *__restrict__ lo, const Cx<double> *__restrict__ in,
                                                             const Cx<double> *__restrict__ gauge,
                                                             const Cx<double> *__restrict__ xpay,
                                                             const Cx<double> *__restrict__ b,
                                                             const double *__restrict__ scale,
                                                             double *__restrict__ partials, double coef, Geom g,
                                                             int R, int parity, int by, int bz, int spw) {
    extern __shared__ double slot[];
    for (int i = threadIdx.x; i < 4 * R; i += blockDim.x) slot[i] = 0.0;
    __syncthreads();
    wave_block<double, Cx<double>, NC, false, ResidualEpi<LO>>(xcd_remap(blockIdx.x, gridDim.x), nullptr, in, gauge,
                                                              xpay, coef, g, R, parity, by, bz, spw,
                                                              ResidualEpi<LO>{b, lo, scale, slot});
    __syncthreads();
    for (int i = threadIdx.x; i < R; i += blockDim.x)
        partials[long(blockIdx.x) * R + i] = (slot[i] + slot[R + i]) + (slot[2 * R + i] + slot[3 * R + i]);
}

// out[x][r] = sum of in[i][r] over rows i in [x per, (x + 1) per) (sqrt of it if `root`); the
// workgroup's threads are (row phase, RHS) so each row is one coalesced load; fixed summation order
__global__ __launch_bounds__(256) void k_rows_sum(double *__restrict__ out, const double *__restrict__ in, int nrows,
                                                 int R, int per, bool root) {
    __shared__ double red[256];
    const int cols = R < 256 ? R : 256, rows = 256 / cols;
    const int c = threadIdx.x % cols, rw = threadIdx.x / cols;
    const int r = blockIdx.y * cols + c;
    const int i0 = blockIdx.x * per, i1 = min(nrows, i0 + per);
    double a = 0.0;
    if (rw < rows && r < R)
        for (int i = i0 + rw; i < i1; i += rows) a += in[long(i) * R + r];
    red[threadIdx.x] = a;
    __syncthreads();
    if (rw == 0 && r < R) {
        double t = 0.0;
        for (int k = 0; k < rows; ++k) t += red[k * cols + c];
        out[long(blockIdx.x) * R + r] = root ? sqrt(t) : t;
    }
}

constexpr int kResidualSlices = 64;  // first-level slices of the per-workgroup partials

template <int NC, typename LO>
hipError_t launch_residual_nc(const DslashArgs &a, const DslashTune &t, const void *b, void *lo,
                              const double *scale, double *partials, double *nrm, hipStream_t stream) {
    const int nblk = div_up(a.g.Vh, 4 * t.spw);
    hipLaunchKernelGGL((k_dslash_wave_residual<NC, LO>), dim3(nblk), dim3(256), size_t(4) * a.R * sizeof(double),
                       stream, static_cast<LO *>(lo), static_cast<const Cx<double> *>(a.in),
                       static_cast<const Cx<double> *>(a.gauge), static_cast<const Cx<double> *>(a.xpay),
                       static_cast<const Cx<double> *>(b), scale, partials, a.coef, a.g, a.R, a.parity, t.by, t.bz,
                       t.spw);
    QCU_HIP_TRY(hipGetLastError());
    // partials [nblk][R] -> [kResidualSlices][R] (stored after them) -> nrm
    double *p2 = partials + long(nblk) * a.R;
    const int cols = a.R < 256 ? a.R : 256;
    hipLaunchKernelGGL(k_rows_sum, dim3(kResidualSlices, div_up(a.R, cols)), dim3(256), 0, stream, p2,
                       static_cast<const double *>(partials), nblk, a.R, div_up(nblk, kResidualSlices), false);
    QCU_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_rows_sum, dim3(1, div_up(a.R, cols)), dim3(256), 0, stream, nrm,
                       static_cast<const double *>(p2), kResidualSlices, a.R, kResidualSlices, true);
    return hipGetLastError();
}

template <typename LO>
hipError_t launch_residual_lo_t(const DslashArgs &a, const DslashTune &t, const void *b, void *lo,
                                const double *scale, double *partials, double *nrm, hipStream_t stream) {
    switch (a.Nc) {
        case 2: return launch_residual_nc<2, LO>(a, t, b, lo, scale, partials, nrm, stream);
        case 3: return launch_residual_nc<3, LO>(a, t, b, lo, scale, partials, nrm, stream);
        case 4: return launch_residual_nc<4, LO>(a, t, b, lo, scale, partials, nrm, stream);
        case 8: return launch_residual_nc<8, LO>(a, t, b, lo, scale, partials, nrm, stream);
        default: return hipErrorInvalidValue;
    }
}

template <typename T, typename CS, int NC, bool DAG>
__global__ __launch_bounds__(256) void k_dslash_wave_groups(CS *__restrict__ out, const CS *__restrict__ in,
                                                           const CS *__restrict__ gauge,
                                                           const CS *__restrict__ xpay, T coef, Geom g,
                                                           int parity, int by, int bz, SegTab tab) {
    const int grp = blockIdx.y;
    if (tab.done && tab.done[grp]) return;
    const int spw = tab.spw[grp];
    const int nblk = (g.Vh + 4 * spw - 1) / (4 * spw);
    if (int(blockIdx.x) >= nblk) return;
    const long off = tab.n * tab.base[grp];
    wave_block<T, CS, NC, DAG>(xcd_remap(blockIdx.x, nblk), out + off, in + off, gauge,
                               xpay ? xpay + off : nullptr, coef, g, tab.R[grp], parity, by, bz, spw);
}

template <typename T, typename S, int NC>
hipError_t launch_groups(const DslashArgs &a, const SegTab &tab, int by, int bz, hipStream_t stream) {
    using CS = typename CxStore<S>::type;
    int nmax = 1;
    for (int k = 0; k < tab.G; ++k) nmax = std::max(nmax, div_up(a.g.Vh, 4 * tab.spw[k]));
    dim3 grid(nmax, tab.G), block(256);
    auto out = static_cast<CS *>(a.out);
    auto in = static_cast<const CS *>(a.in);
    auto gauge = static_cast<const CS *>(a.gauge);
    auto xp = static_cast<const CS *>(a.xpay);
    if (a.dagger)
        hipLaunchKernelGGL((k_dslash_wave_groups<T, CS, NC, true>), grid, block, 0, stream, out, in, gauge, xp,
                           T(a.coef), a.g, a.parity, by, bz, tab);
    else
        hipLaunchKernelGGL((k_dslash_wave_groups<T, CS, NC, false>), grid, block, 0, stream, out, in, gauge, xp,
                           T(a.coef), a.g, a.parity, by, bz, tab);
    return hipGetLastError();
}

template <typename T, typename S, int NC>
hipError_t launch(const DslashArgs &a, const DslashTune &t, hipStream_t stream) {
    using CS = typename CxStore<S>::type;
    const int sites_per_block = 4 * t.spw;
    dim3 grid(div_up(a.g.Vh, sites_per_block)), block(256);
    auto out = static_cast<CS *>(a.out);
    auto in = static_cast<const CS *>(a.in);
    auto gauge = static_cast<const CS *>(a.gauge);
    auto xp = static_cast<const CS *>(a.xpay);
    if (a.dagger)
        hipLaunchKernelGGL((k_dslash_wave<T, CS, NC, true>), grid, block, 0, stream, out, in, gauge, xp,
                           T(a.coef), a.g, a.R, a.parity, t.by, t.bz, t.spw);
    else
        hipLaunchKernelGGL((k_dslash_wave<T, CS, NC, false>), grid, block, 0, stream, out, in, gauge, xp,
                           T(a.coef), a.g, a.R, a.parity, t.by, t.bz, t.spw);
    return hipGetLastError();
}

template <typename T, typename S>
hipError_t dispatch_nc_groups(const DslashArgs &a, const SegTab &tab, int by, int bz, hipStream_t stream) {
    switch (a.Nc) {
        case 2: return launch_groups<T, S, 2>(a, tab, by, bz, stream);
        case 3: return launch_groups<T, S, 3>(a, tab, by, bz, stream);
        case 4: return launch_groups<T, S, 4>(a, tab, by, bz, stream);
        case 8: return launch_groups<T, S, 8>(a, tab, by, bz, stream);
        default: return hipErrorInvalidValue;
    }
}

template <typename T, typename S>
hipError_t dispatch_nc(const DslashArgs &a, const DslashTune &t, hipStream_t stream) {
    switch (a.Nc) {
        case 2: return launch<T, S, 2>(a, t, stream);
        case 3: return launch<T, S, 3>(a, t, stream);
        case 4: return launch<T, S, 4>(a, t, stream);
        case 8: return launch<T, S, 8>(a, t, stream);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_dslash_wave(const DslashArgs &a, const DslashTune &t, hipStream_t stream) {
    if (t.by <= 0 || t.bz <= 0 || t.spw <= 0 || a.g.L[1] % t.by || a.g.L[2] % t.bz) return hipErrorInvalidValue;
    switch (a.prec) {
        case kDouble: return dispatch_nc<double, double>(a, t, stream);
        case kSingle: return dispatch_nc<float, float>(a, t, stream);
        case kHalf: return dispatch_nc<float, __half>(a, t, stream);
        case kBF16: return dispatch_nc<float, __hip_bfloat16>(a, t, stream);
        default: return hipErrorInvalidValue;
    }
}

long dslash_residual_partials(const Geom &g, int R) { return (long(div_up(g.Vh, 4)) + kResidualSlices) * R; }

hipError_t launch_rows_sum(double *out, const double *in, int nrows, int cols, int slices, hipStream_t stream) {
    if (nrows < 1 || cols < 1 || slices < 1) return hipErrorInvalidValue;
    const int c = cols < 256 ? cols : 256;
    hipLaunchKernelGGL(k_rows_sum, dim3(slices, div_up(cols, c)), dim3(256), 0, stream, out, in, nrows, cols,
                       div_up(nrows, slices), false);
    return hipGetLastError();
}

hipError_t launch_dslash_residual_lo(const DslashArgs &a, const void *b, void *lo, int lo_prec, const double *scale,
                                     double *partials, long partials_cap, double *nrm, hipStream_t stream) {
    if (a.prec != kDouble || a.dagger || !a.in || !a.gauge || !a.xpay || !b || !lo || !scale || !partials || !nrm ||
        a.R < 1 || 32L * a.R > 65536L || !dslash_supported(a.Nc, a.R))
        return hipErrorInvalidValue;
    if (a.g.Vh <= 0) return hipSuccess;
    // the wave kernel's launch shape for this operator (forced when the tuned variant is another)
    DslashTune t = dslash_resolve(a);
    if (t.variant != kDslashWaveSite) {
        DslashArgs w = a;
        w.tune = DslashTune{};
        w.tune.variant = kDslashWaveSite;
        t = dslash_resolve(w);
    }
    if (t.by <= 0 || t.bz <= 0 || t.spw <= 0 || a.g.L[1] % t.by || a.g.L[2] % t.bz) return hipErrorInvalidValue;
    if ((long(div_up(a.g.Vh, 4 * t.spw)) + kResidualSlices) * a.R > partials_cap) return hipErrorInvalidValue;
    switch (lo_prec) {
        case kSingle: return launch_residual_lo_t<Cx<float>>(a, t, b, lo, scale, partials, nrm, stream);
        case kHalf: return launch_residual_lo_t<Half2Store>(a, t, b, lo, scale, partials, nrm, stream);
        case kBF16: return launch_residual_lo_t<Bf162Store>(a, t, b, lo, scale, partials, nrm, stream);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_dslash_wave_groups(const DslashArgs &a, const SegTab &tab, int by, int bz, hipStream_t stream) {
    if (tab.G < 1 || tab.G > kMaxGroups || by <= 0 || bz <= 0 || a.g.L[1] % by || a.g.L[2] % bz)
        return hipErrorInvalidValue;
    for (int k = 0; k < tab.G; ++k)
        if (tab.spw[k] <= 0 || tab.R[k] < 1 || !dslash_supported(a.Nc, tab.R[k])) return hipErrorInvalidValue;
    switch (a.prec) {
        case kDouble: return dispatch_nc_groups<double, double>(a, tab, by, bz, stream);
        case kSingle: return dispatch_nc_groups<float, float>(a, tab, by, bz, stream);
        case kHalf: return dispatch_nc_groups<float, __half>(a, tab, by, bz, stream);
        case kBF16: return dispatch_nc_groups<float, __hip_bfloat16>(a, tab, by, bz, stream);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace qcu
