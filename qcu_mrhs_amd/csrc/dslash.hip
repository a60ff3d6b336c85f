// dslash.hip -- one-parity multi-RHS SU(N) Wilson hopping term on gfx950.
//
// Operator (include/kernel/sun_mrhs_wilson_dslash_simt.cuh:97-181 of the
// reference; output parity p, input on 1-p):
//   out(x) = sum_mu (1 - g_mu) U_mu(x) psi(x+mu) + (1 + g_mu) U_mu(x-mu)^+ psi(x-mu)
// dagger flips the sign of every gamma (ibid. :147,176).  Spin structure via
// two projected rows per hop: h_r = psi_r + s_r psi_m(r), out_r += U h_r,
// out_m(r) += c_r U h_r (include/kernel/qcu_gamma.cuh:47-324).
//
// This file holds the dispatcher and the lane-per-(site, rhs) VALU kernel
// (variant kDslashLaneRhs): R lanes of one site read R consecutive complex
// numbers for every (spin, colour); the link is a broadcast load shared by
// those lanes.  The wave-per-site kernels live in dslash_wave.hip (VALU) and
// dslash_mfma.hip (matrix cores).  Every variant has an optional epilogue
// out = x + coef * D psi (Schur operator, full-lattice M, source preparation).
#include <algorithm>
#include <array>
#include <map>
#include <mutex>
#include <tuple>

#include "blas.h"
#include "dslash_common.h"

namespace qcu {
namespace {

template <typename T, typename CS, typename CG, int NC, int MU, bool FWD, bool DAG>
__device__ __forceinline__ void hop(Cx<T> (&acc)[4][NC], const CS *__restrict__ psi, int R,
                                    const CG *__restrict__ U) {
    constexpr int sgn = (FWD != DAG) ? 2 : 0;
    constexpr int m0 = kMat2(MU, 0), m1 = kMat2(MU, 1);
    constexpr int kp0 = kProjK(MU, 0) + sgn, kp1 = kProjK(MU, 1) + sgn;
    constexpr int kr0 = kReconK(MU, 0) + sgn, kr1 = kReconK(MU, 1) + sgn;
    Cx<T> h0[NC], h1[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const Cx<T> a0 = load_cx<T>(psi + (0 * NC + c) * R);
        const Cx<T> a1 = load_cx<T>(psi + (1 * NC + c) * R);
        const Cx<T> b0 = load_cx<T>(psi + (m0 * NC + c) * R);
        const Cx<T> b1 = load_cx<T>(psi + (m1 * NC + c) * R);
        h0[c] = a0 + ipow<kp0>(b0);
        h1[c] = a1 + ipow<kp1>(b1);
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
                const Cx<T> u = load_cx<T>(U + b * NC + a);   // (U^+)[a][b] = conj(U[b][a])
                cfma_conj(s0, u, h0[b]);
                cfma_conj(s1, u, h1[b]);
            }
        }
        acc[0][a] = acc[0][a] + s0;
        acc[m0][a] = acc[m0][a] + ipow<kr0>(s0);
        acc[1][a] = acc[1][a] + s1;
        acc[m1][a] = acc[m1][a] + ipow<kr1>(s1);
    }
}

template <typename T, typename CS, typename CG, int NC, bool DAG>
__global__ __launch_bounds__(256) void k_wilson_dslash_valu(CS *__restrict__ out,
                                                           const CS *__restrict__ in,
                                                           const CG *__restrict__ gauge,
                                                           const CS *__restrict__ xpay, T coef,
                                                           Geom g, int R, int parity) {
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int site = b * blockDim.y + threadIdx.y;
    if (site >= g.Vh) return;
    const int r = threadIdx.x;

    const int Lx = g.L[0], Ly = g.L[1], Lz = g.L[2], Lt = g.L[3], Lxh = g.Lxh;
    const int xh = site % Lxh;
    int rest = site / Lxh;
    const int y = rest % Ly;
    rest /= Ly;
    const int z = rest % Lz;
    const int t = rest / Lz;
    const int x = 2 * xh + ((y + z + t + parity) & 1);
    const int row = site - xh;
    const int sy = Lxh, sz = Lxh * Ly, st = Lxh * Ly * Lz;

    int nbf[4], nbb[4];
    nbf[0] = row + ((x + 1 == Lx ? 0 : x + 1) >> 1);
    nbb[0] = row + ((x == 0 ? Lx - 1 : x - 1) >> 1);
    nbf[1] = site + (y + 1 == Ly ? -(Ly - 1) * sy : sy);
    nbb[1] = site + (y == 0 ? (Ly - 1) * sy : -sy);
    nbf[2] = site + (z + 1 == Lz ? -(Lz - 1) * sz : sz);
    nbb[2] = site + (z == 0 ? (Lz - 1) * sz : -sz);
    nbf[3] = site + (t + 1 == Lt ? -(Lt - 1) * st : st);
    nbb[3] = site + (t == 0 ? (Lt - 1) * st : -st);

    const long site_len = 4L * NC * R;
    const long link_len = long(NC) * NC;
    const CS *inr = in + r;
    const long Vh = g.Vh;

    Cx<T> acc[4][NC];
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[s][c] = {T(0), T(0)};

#define QCU_HOP(MU)                                                                          \
    hop<T, CS, CG, NC, MU, true, DAG>(acc, inr + nbf[MU] * site_len, R,                       \
                                      gauge + ((MU * 2 + parity) * Vh + site) * link_len);    \
    hop<T, CS, CG, NC, MU, false, DAG>(acc, inr + nbb[MU] * site_len, R,                      \
                                       gauge + ((MU * 2 + 1 - parity) * Vh + nbb[MU]) * link_len);
    QCU_HOP(0)
    QCU_HOP(1)
    QCU_HOP(2)
    QCU_HOP(3)
#undef QCU_HOP

    CS *o = out + site * site_len + r;
    if (xpay) {
        const CS *xp = xpay + site * site_len + r;
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                Cx<T> v = load_cx<T>(xp + (s * NC + c) * R);
                v.re = fma(coef, acc[s][c].re, v.re);
                v.im = fma(coef, acc[s][c].im, v.im);
                store_cx(o + (s * NC + c) * R, v);
            }
    } else {
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int c = 0; c < NC; ++c) store_cx(o + (s * NC + c) * R, coef * acc[s][c]);
    }
}

template <typename T, typename S, int NC>
hipError_t launch_valu(const DslashArgs &a, hipStream_t stream) {
    using CS = typename CxStore<S>::type;
    const int spb = a.R >= 256 ? 1 : 256 / a.R;
    dim3 block(a.R, spb);
    dim3 grid(div_up(a.g.Vh, spb));
    auto out = static_cast<CS *>(a.out);
    auto in = static_cast<const CS *>(a.in);
    auto gauge = static_cast<const CS *>(a.gauge);
    auto xp = static_cast<const CS *>(a.xpay);
    if (a.dagger)
        hipLaunchKernelGGL((k_wilson_dslash_valu<T, CS, CS, NC, true>), grid, block, 0, stream,
                           out, in, gauge, xp, T(a.coef), a.g, a.R, a.parity);
    else
        hipLaunchKernelGGL((k_wilson_dslash_valu<T, CS, CS, NC, false>), grid, block, 0, stream,
                           out, in, gauge, xp, T(a.coef), a.g, a.R, a.parity);
    return hipGetLastError();
}

template <typename T, typename S>
hipError_t dispatch_nc(const DslashArgs &a, hipStream_t stream) {
    switch (a.Nc) {
        case 2: return launch_valu<T, S, 2>(a, stream);
        case 3: return launch_valu<T, S, 3>(a, stream);
        case 4: return launch_valu<T, S, 4>(a, stream);
        case 8: return launch_valu<T, S, 8>(a, stream);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

bool dslash_supported(int Nc, int R) {
    return (Nc == 2 || Nc == 3 || Nc == 4 || Nc == 8) && R >= 1 && R <= 256;
}

namespace {
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int, int, int, int>, DslashTune> g_tuned;
DslashTune g_override;  // variant != auto: forced for every launch

int largest_divisor_at_most(int L, int cap) {
    for (int d = cap < L ? cap : L; d > 1; --d)
        if (L % d == 0) return d;
    return 1;
}
}  // namespace

void dslash_set_tuned(const int L[4], int Nc, int R, int prec, const DslashTune &t) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[std::make_tuple(L[0], L[1], L[2], L[3], Nc, R, prec)] = t;
}

void dslash_set_override(const DslashTune &t) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_override = t;
}

DslashTune dslash_resolve(const DslashArgs &a) {
    DslashTune t = a.tune;
    if (t.variant == kDslashAuto) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        if (g_override.variant != kDslashAuto) {
            t = g_override;
        } else {
            auto it = g_tuned.find(std::make_tuple(a.g.L[0], a.g.L[1], a.g.L[2], a.g.L[3], a.Nc, a.R, a.prec));
            if (it != g_tuned.end()) t = it->second;
        }
    }
    return dslash_fill_defaults(a, t);
}

bool dslash_get_tuned(const int L[4], int Nc, int R, int prec, DslashTune &t) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(std::make_tuple(L[0], L[1], L[2], L[3], Nc, R, prec));
    if (it == g_tuned.end()) return false;
    t = it->second;
    return true;
}

std::vector<std::pair<std::array<int, 7>, DslashTune>> dslash_tuned_entries() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    std::vector<std::pair<std::array<int, 7>, DslashTune>> v;
    for (const auto &e : g_tuned) {
        const auto &k = e.first;
        v.push_back({{std::get<0>(k), std::get<1>(k), std::get<2>(k), std::get<3>(k), std::get<4>(k),
                      std::get<5>(k), std::get<6>(k)},
                     e.second});
    }
    return v;
}

DslashTune dslash_fill_defaults(const DslashArgs &a, DslashTune t) {
    if (t.variant == kDslashAuto) t.variant = dslash_mfma_supported(a.Nc, a.prec) ? kDslashMfma : kDslashWaveSite;
    if (t.variant == kDslashMfma && !dslash_matrix_supported(a.Nc, a.prec)) t.variant = kDslashWaveSite;
    if (t.variant == kDslashLds && !dslash_lds_supported(a.g, a.Nc, a.prec))
        t.variant = dslash_matrix_supported(a.Nc, a.prec) ? kDslashMfma : kDslashWaveSite;
    if (t.variant == kDslashLds) {
        // t steps per workgroup and the (y, z) block of the workgroup order (measured at 24^3x48,
        // 16 RHS: blocks of 4 x 8 to 6 x 12 rows 232-235 us against 237-254 us row by row)
        if (t.spw <= 0) t.spw = 12;
        if (t.by <= 0 || a.g.L[1] % t.by) t.by = largest_divisor_at_most(a.g.L[1], 4);
        if (t.bz <= 0 || a.g.L[2] % t.bz) t.bz = largest_divisor_at_most(a.g.L[2], 8);
        return t;
    }
    if (t.variant != kDslashLaneRhs) {
        // spw < 0 (persistent grid, -spw workgroups per CU) exists for the FP16/BF16 MFMA kernel only
        if (t.spw == 0 || (t.spw < 0 && !(t.variant == kDslashMfma && dslash_mfma_supported(a.Nc, a.prec))))
            t.spw = t.variant == kDslashMfma ? 1 : 4;  // measured best at 24^3x48, R = 16
        // tile the (y, z) plane so one tile slice (the t-reuse distance) stays
        // around 256 KiB: bytes per site = 4 Nc R complex numbers (measured at
        // 24^3x48, 16 RHS: best among tiles of 0.06 - 5 MiB for FP16 and FP64)
        const double site_bytes = 4.0 * a.Nc * a.R * 2.0 * double(prec_bytes(a.prec));
        double area = (1 << 18) / (site_bytes * a.g.Lxh);
        if (area < 1) area = 1;
        if (t.by <= 0 || a.g.L[1] % t.by) {
            int cap = 1;
            while ((cap + 1) * (cap + 1) <= area) ++cap;
            t.by = largest_divisor_at_most(a.g.L[1], cap);
        }
        if (t.bz <= 0 || a.g.L[2] % t.bz) {
            int cap = int(area / t.by);
            t.bz = largest_divisor_at_most(a.g.L[2], cap < 1 ? 1 : cap);
        }
    }
    return t;
}

hipError_t launch_dslash(const DslashArgs &a, hipStream_t stream) {
    if (!dslash_supported(a.Nc, a.R)) return hipErrorInvalidValue;
    if (a.g.Vh <= 0) return hipSuccess;
    const DslashTune t = dslash_resolve(a);
    if (t.variant == kDslashWaveSite) return launch_dslash_wave(a, t, stream);
    if (t.variant == kDslashLds) return launch_dslash_lds(a, t, stream);
    if (t.variant == kDslashMfma)
        return (a.prec == kSingle || a.prec == kDouble) ? launch_dslash_mfma32(a, t, stream)
                                                         : launch_dslash_mfma(a, t, stream);
    switch (a.prec) {
        case kDouble: return dispatch_nc<double, double>(a, stream);
        case kSingle: return dispatch_nc<float, float>(a, stream);
        case kHalf: return dispatch_nc<float, __half>(a, stream);
        case kBF16: return dispatch_nc<float, __hip_bfloat16>(a, stream);
        default: return hipErrorInvalidValue;
    }
}

constexpr int kDotSlices = 512;  // first-level slices of the per-workgroup partials

long dslash_omega_partials(const Geom &g, int R) {
    // one row of R x 4 per workgroup (at most one per 4 sites, or 64 per CU for the persistent grid)
    const long nblk = std::max<long>(div_up(g.Vh, 4), 64L * num_cus());
    return (nblk + kDotSlices) * R * 4;
}

hipError_t launch_dslash_omega(const DslashArgs &a, double *partials, long cap, Cx<double> *st, hipStream_t stream) {
    if (!dslash_supported(a.Nc, a.R) || a.g.Vh <= 0 || (a.prec != kHalf && a.prec != kBF16) || a.dagger ||
        !a.xpay || !partials || !st || a.R > kReduceThreads)
        return hipErrorNotSupported;
    const DslashTune t = dslash_resolve(a);
    if (t.variant != kDslashMfma) return hipErrorNotSupported;
    if (cap < dslash_omega_partials(a.g, a.R)) return hipErrorInvalidValue;
    int nblk = 0;
    QCU_HIP_TRY(launch_dslash_mfma_omega(a, t, partials, stream, nblk));
    // [nblk][R][4] -> [kDotSlices][R][4] (after them) -> the finisher sums the slices in order
    double *p2 = partials + long(nblk) * a.R * 4;
    QCU_HIP_TRY(launch_rows_sum(p2, partials, nblk, 4 * a.R, kDotSlices, stream));
    return launch_finish_omega(st, a.R, p2, kDotSlices, stream);
}

}  // namespace qcu
