// blas.hip -- multi-RHS BLAS, single-pass deterministic reductions and the
// fused BiCGStab steps.
//
// Thread mapping: each workgroup keeps (256/R)*R active lanes, so a lane's
// RHS index is fixed for its whole grid-stride loop and consecutive lanes
// touch consecutive complex numbers of the [n][R] layout (coalesced).
//
// Reductions read their vectors once: each block reduces its lanes per RHS
// through LDS in a fixed order and writes R partials, and a one-block kernel
// sums the partials in block order and runs the op's finisher (e.g.
// alpha = rho / <r0, v>) on device, so no host round trip is needed.  Results
// are bit-reproducible for a given (n, R).  Accumulation is FP64 for every
// storage precision.
#include <algorithm>
#include <mutex>
#include <type_traits>

#include "blas.h"

// Device scratch of launch_dotc_noalloc: block partials [blocks][R][4] FP64.
// Each device holds its own copy for the life of the process; host code
// reaches it only through the address hipGetSymbolAddress returns for the
// current device (the symbol itself is a host-side stand-in).
static __device__ __align__(256) double scratch[qcu::kNoAllocScratchBytes / sizeof(double)];

namespace qcu {
namespace {

template <int P> struct PrecT;
template <> struct PrecT<kDouble> { using S = double; using T = double; };
template <> struct PrecT<kSingle> { using S = float; using T = float; };
template <> struct PrecT<kHalf> { using S = __half; using T = float; };
template <> struct PrecT<kBF16> { using S = __hip_bfloat16; using T = float; };

#define QCU_PREC_DISPATCH(prec, ...)                                \
    switch (prec) {                                                 \
        case kDouble: { constexpr int P = kDouble; __VA_ARGS__; }   \
        case kSingle: { constexpr int P = kSingle; __VA_ARGS__; }   \
        case kHalf: { constexpr int P = kHalf; __VA_ARGS__; }       \
        case kBF16: { constexpr int P = kBF16; __VA_ARGS__; }       \
        default: return hipErrorInvalidValue;                       \
    }

__device__ __forceinline__ bool rhs_lane(int R, long &g, long &stride, int &rhs) {
    const int per_block = (int(blockDim.x) / R) * R;
    if (int(threadIdx.x) >= per_block) return false;
    g = long(blockIdx.x) * per_block + threadIdx.x;
    stride = long(gridDim.x) * per_block;
    rhs = int(threadIdx.x) % R;
    return true;
}

template <typename T>
__device__ __forceinline__ Cx<T> tc(C64 v) { return {T(v.re), T(v.im)}; }

// ------------------------------------------------------------------ axpy
template <int KIND, typename CS, typename T>
__global__ __launch_bounds__(256) void k_axpy(CS *__restrict__ res, const CS *__restrict__ x,
                                             const CS *__restrict__ y, const CS *__restrict__ z,
                                             const C64 *__restrict__ a, const C64 *__restrict__ b,
                                             const C64 *__restrict__ c, long N, int R) {
    long g, stride;
    int rhs;
    if (!rhs_lane(R, g, stride, rhs)) return;
    const Cx<T> ar = tc<T>(a[rhs]);
    Cx<T> br = {T(0), T(0)}, crr = {T(0), T(0)};
    if constexpr (KIND == kAxpby || KIND == kAxpbypcz) br = tc<T>(b[rhs]);
    if constexpr (KIND == kAxpbypcz) crr = tc<T>(c[rhs]);
    for (long e = g; e < N; e += stride) {
        const Cx<T> xv = load_cx<T>(x + e);
        const Cx<T> yv = load_cx<T>(y + e);
        Cx<T> o;
        if constexpr (KIND == kXsay) o = xv - ar * yv;
        else if constexpr (KIND == kXpay) o = xv + ar * yv;
        else if constexpr (KIND == kAxpby) o = ar * xv + br * yv;
        else o = ar * xv + br * yv + crr * load_cx<T>(z + e);
        store_cx(res + e, o);
    }
}

// ------------------------------------------------------------- reduction
template <class Op, class = void> struct HasSplit { static constexpr bool value = false; };
template <class Op> struct HasSplit<Op, std::void_t<decltype(Op::kSplit)>> { static constexpr bool value = Op::kSplit; };
// entries per trip of the split form (default 4; an op with wide entries may ask for fewer)
template <class Op, class = void> struct SplitWidth { static constexpr int value = 4; };
template <class Op> struct SplitWidth<Op, std::void_t<decltype(Op::kSplitWidth)>> {
    static constexpr int value = Op::kSplitWidth;
};

// Two kernels per reduction: every block reduces its lanes per RHS through
// LDS in a fixed order and writes R partials; a one-block kernel then sums the
// partials in block order and runs the op's finisher.  The kernel boundary
// orders the partial writes before their reads: no device-scope fences or
// atomics inside the hot kernel (on a multi-XCD GPU a device-scope release
// writes back the XCD's L2, which made a per-block fence the dominant cost).
template <class Op>
__device__ __forceinline__ void reduce_partial(const Op &op, long N, int R, double *__restrict__ partials) {
    constexpr int NA = Op::NA;
    __shared__ double buf[NA][kReduceThreads];
    double acc[NA];
#pragma unroll
    for (int k = 0; k < NA; ++k) acc[k] = 0.0;
    long g, stride;
    int rhs;
    if (rhs_lane(R, g, stride, rhs)) {
        auto loc = op.local(rhs);
        long e = g;
        // several entries per trip keep more loads in flight (same summation order as one per trip)
        if constexpr (HasSplit<Op>::value) {
            constexpr int W = SplitWidth<Op>::value;
            for (; e + (W - 1) * stride < N; e += W * stride) {
                typename Op::V v[W];
#pragma unroll
                for (int u = 0; u < W; ++u) v[u] = op.load(e + u * stride);
#pragma unroll
                for (int u = 0; u < W; ++u) op.apply(e + u * stride, v[u], loc, acc);
            }
        } else {
            for (; e + 3 * stride < N; e += 4 * stride) {
#pragma unroll
                for (int u = 0; u < 4; ++u) op.elem(e + u * stride, loc, acc);
            }
        }
        for (; e < N; e += stride) op.elem(e, loc, acc);
    }
#pragma unroll
    for (int k = 0; k < NA; ++k) buf[k][threadIdx.x] = acc[k];
    __syncthreads();
    const int nch = kReduceThreads / R;
    if (int(threadIdx.x) < R) {
        double s[NA];
#pragma unroll
        for (int k = 0; k < NA; ++k) s[k] = 0.0;
        for (int j = 0; j < nch; ++j)
#pragma unroll
            for (int k = 0; k < NA; ++k) s[k] += buf[k][j * R + threadIdx.x];
#pragma unroll
        for (int k = 0; k < NA; ++k) partials[(long(blockIdx.x) * R + threadIdx.x) * 4 + k] = s[k];
    }
}

// One block: sum `nblk` block partials per RHS in a fixed order, then finish.
template <class Op>
__device__ __forceinline__ void reduce_finish(const Op &op, int nblk, int R, const double *__restrict__ partials) {
    constexpr int NA = Op::NA;
    __shared__ double buf[NA][kReduceThreads];
    const int t = threadIdx.x;
    const int per_block = (kReduceThreads / R) * R;
    const int nch = per_block / R;
    double acc[NA];
#pragma unroll
    for (int k = 0; k < NA; ++k) acc[k] = 0.0;
    if (t < per_block) {
        const int r = t % R, ch = t / R;
        // four independent partial loads per trip; fixed order bb, bb + nch, bb + 2 nch, bb + 3 nch
        int bb = ch;
        for (; bb + 3 * nch < nblk; bb += 4 * nch) {
            double v[4][NA];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int k = 0; k < NA; ++k) v[u][k] = partials[(long(bb + u * nch) * R + r) * 4 + k];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int k = 0; k < NA; ++k) acc[k] += v[u][k];
        }
        for (; bb < nblk; bb += nch)
#pragma unroll
            for (int k = 0; k < NA; ++k) acc[k] += partials[(long(bb) * R + r) * 4 + k];
    }
#pragma unroll
    for (int k = 0; k < NA; ++k) buf[k][t] = acc[k];
    __syncthreads();
    if (t < R) {
        double tot[NA];
#pragma unroll
        for (int k = 0; k < NA; ++k) tot[k] = 0.0;
        for (int j = 0; j < nch; ++j)
#pragma unroll
            for (int k = 0; k < NA; ++k) tot[k] += buf[k][j * R + t];
        op.finish(t, tot);
    }
}

template <class Op>
__global__ __launch_bounds__(kReduceThreads) void k_reduce(Op op, long N, int R, double *__restrict__ partials) {
    reduce_partial(op, N, R, partials);
}

template <class Op>
__global__ __launch_bounds__(kReduceThreads) void k_reduce_finish(Op op, int nblk, int R,
                                                                 const double *__restrict__ partials) {
    reduce_finish(op, nblk, R, partials);
}

template <typename CS>
struct DotOp {
    static constexpr int NA = 2;
    const CS *a, *b;
    C64 *res;
    struct L {};
    __device__ L local(int) const { return {}; }
    __device__ void elem(long e, L, double (&acc)[NA]) const {
        const Cx<double> u = load_cx<double>(a + e), v = load_cx<double>(b + e);
        acc[0] = fma(u.re, v.re, fma(u.im, v.im, acc[0]));
        acc[1] = fma(u.re, v.im, fma(-u.im, v.re, acc[1]));
    }
    __device__ void finish(int r, const double (&t)[NA]) const { res[r] = {t[0], t[1]}; }
    __device__ DotOp shifted(long off, int base) const { return {a + off, b + off, res + base}; }
};

template <typename CS>
struct NormOp {
    static constexpr int NA = 1;
    const CS *a;
    double *res;
    bool take_sqrt;
    struct L {};
    __device__ L local(int) const { return {}; }
    __device__ void elem(long e, L, double (&acc)[NA]) const {
        const Cx<double> u = load_cx<double>(a + e);
        acc[0] = fma(u.re, u.re, fma(u.im, u.im, acc[0]));
    }
    __device__ void finish(int r, const double (&t)[NA]) const {
        res[r] = take_sqrt ? sqrt(t[0]) : t[0];
    }
    __device__ NormOp shifted(long off, int base) const { return {a + off, res + base, take_sqrt}; }
};

template <typename CS>
struct AlphaOp {  // <r0, v>, alpha = rho / <r0, v>
    static constexpr int NA = 2;
    const CS *r0, *v;
    C64 *st;
    int R;
    struct L {};
    __device__ L local(int) const { return {}; }
    __device__ void elem(long e, L, double (&acc)[NA]) const {
        const Cx<double> u = load_cx<double>(r0 + e), w = load_cx<double>(v + e);
        acc[0] = fma(u.re, w.re, fma(u.im, w.im, acc[0]));
        acc[1] = fma(u.re, w.im, fma(-u.im, w.re, acc[1]));
    }
    __device__ void finish(int r, const double (&t)[NA]) const {
        const C64 r0v = {t[0], t[1]};
        st[kR0v * R + r] = r0v;
        // a zero denominator (converged or all-zero RHS) freezes that RHS
        st[kAlpha * R + r] = (r0v.re == 0.0 && r0v.im == 0.0) ? C64{0.0, 0.0}
                                                                : cdiv(st[kRho * R + r], r0v);
    }
    __device__ AlphaOp shifted(long off, int base) const { return {r0 + off, v + off, st + base, R}; }
};

template <typename CS>
struct OmegaOp {  // <t,s>, <t,t>, omega = <t,s>/<t,t>
    static constexpr int NA = 3;
    const CS *t, *s;
    C64 *st;
    int R;
    struct L {};
    __device__ L local(int) const { return {}; }
    __device__ void elem(long e, L, double (&acc)[NA]) const {
        const Cx<double> u = load_cx<double>(t + e), w = load_cx<double>(s + e);
        acc[0] = fma(u.re, w.re, fma(u.im, w.im, acc[0]));
        acc[1] = fma(u.re, w.im, fma(-u.im, w.re, acc[1]));
        acc[2] = fma(u.re, u.re, fma(u.im, u.im, acc[2]));
    }
    __device__ void finish(int r, const double (&tt)[NA]) const {
        const C64 ts = {tt[0], tt[1]}, t2 = {tt[2], 0.0};
        st[kTs * R + r] = ts;
        st[kTt * R + r] = t2;
        st[kOmega * R + r] = t2.re == 0.0 ? C64{0.0, 0.0} : cdiv(ts, t2);
    }
    __device__ OmegaOp shifted(long off, int base) const { return {t + off, s + off, st + base, R}; }
};

template <typename CS>
struct XsayNormOp {  // res = x - y ; ||res|| per RHS (FP64 accumulation of the stored values)
    static constexpr int NA = 1;
    CS *res;
    const CS *x, *y;
    double *nrm;
    bool take_sqrt;
    static constexpr bool kSplit = true;
    struct L {};
    struct V { CS x, y; };
    __device__ L local(int) const { return {}; }
    __device__ V load(long e) const { return {x[e], y[e]}; }
    __device__ void apply(long e, const V &v, L, double (&acc)[NA]) const {
        using T = decltype(load_cx<double>(&v.x).re);
        const Cx<T> d = load_cx<T>(&v.x) - load_cx<T>(&v.y);
        CS rs;
        store_cx(&rs, d);
        res[e] = rs;
        const Cx<double> u = load_cx<double>(&rs);
        acc[0] = fma(u.re, u.re, fma(u.im, u.im, acc[0]));
    }
    __device__ void elem(long e, L l, double (&acc)[NA]) const { apply(e, load(e), l, acc); }
    __device__ void finish(int r, const double (&t)[NA]) const { nrm[r] = take_sqrt ? sqrt(t[0]) : t[0]; }
    __device__ XsayNormOp shifted(long off, int base) const {
        return {res + off, x + off, y + off, nrm + base, take_sqrt};
    }
};

template <typename CL>
struct ResidualLoOp {  // d = b - y in FP64 ; ||d|| per RHS ; lo = scale_r d rounded to the inner storage
    static constexpr int NA = 1;
    static constexpr bool kSplit = true;
    CL *lo;
    const Cx<double> *b, *y;
    const double *scale;
    double *nrm;
    struct L { double sc; };
    struct V { Cx<double> b, y; };
    __device__ L local(int rhs) const { return {scale[rhs]}; }
    __device__ V load(long e) const { return {b[e], y[e]}; }
    __device__ void apply(long e, const V &v, L l, double (&acc)[NA]) const {
        const Cx<double> d = v.b - v.y;
        acc[0] = fma(d.re, d.re, fma(d.im, d.im, acc[0]));
        store_cx(lo + e, l.sc * d);
    }
    __device__ void elem(long e, L l, double (&acc)[NA]) const { apply(e, load(e), l, acc); }
    __device__ void finish(int r, const double (&t)[NA]) const { nrm[r] = sqrt(t[0]); }
    __device__ ResidualLoOp shifted(long off, int base) const {
        return {lo + off, b + off, y + off, scale + base, nrm + base};
    }
};

template <typename CS, typename T>
struct XROp {  // x += a p + w shat ; r = s - w t ; ||r||^2 ; <r0, r> ; beta
    static constexpr int NA = 3;
    CS *x, *r;
    const CS *p, *shat, *s, *t, *r0;
    C64 *st;
    int R;
    struct L { Cx<T> al, om; };
    __device__ L local(int rhs) const { return {tc<T>(st[kAlpha * R + rhs]), tc<T>(st[kOmega * R + rhs])}; }
    // split form: every load of a trip is issued before the first store (the stores could
    // alias the later loads as far as the compiler knows, which would serialise them); FP64
    // entries take two per trip (four would hold 20 16-byte loads per lane, which costs more
    // occupancy than it gains)
    static constexpr bool kSplit = true;
    static constexpr int kSplitWidth = sizeof(CS) <= 8 ? 4 : 2;
    struct V { CS x, p, s, sh, t, r0; };
    __device__ V load(long e) const {
        V v;
        v.s = s[e];
        v.sh = shat == s ? v.s : shat[e];
        v.x = x[e];
        v.p = p[e];
        v.t = t[e];
        v.r0 = r0[e];
        return v;
    }
    __device__ void apply(long e, const V &v, L l, double (&acc)[NA]) const {
        const Cx<T> sv = load_cx<T>(&v.s);
        const Cx<T> xv = load_cx<T>(&v.x) + l.al * load_cx<T>(&v.p) + l.om * load_cx<T>(&v.sh);
        store_cx(x + e, xv);
        const Cx<T> rv = sv - l.om * load_cx<T>(&v.t);
        // reductions use the stored (rounded) residual: round in registers, not through memory
        CS rs;
        store_cx(&rs, rv);
        r[e] = rs;
        const Cx<double> rd = load_cx<double>(&rs);
        const Cx<double> u = load_cx<double>(&v.r0);
        acc[0] = fma(rd.re, rd.re, fma(rd.im, rd.im, acc[0]));
        acc[1] = fma(u.re, rd.re, fma(u.im, rd.im, acc[1]));
        acc[2] = fma(u.re, rd.im, fma(-u.im, rd.re, acc[2]));
    }
    __device__ void elem(long e, L l, double (&acc)[NA]) const { apply(e, load(e), l, acc); }
    __device__ void finish(int rr, const double (&tt)[NA]) const {
        st[kRnorm2 * R + rr] = {tt[0], 0.0};
        const C64 rho_new = {tt[1], tt[2]};
        st[kRhoNew * R + rr] = rho_new;
        // beta = ((alpha * rho_new) / omega) / rho (src/solver/bicgstab_solve.cu:350-374)
        const C64 om = st[kOmega * R + rr], rho = st[kRho * R + rr];
        C64 beta = {0.0, 0.0};
        if (!(om.re == 0.0 && om.im == 0.0) && !(rho.re == 0.0 && rho.im == 0.0)) {
            beta = st[kAlpha * R + rr] * rho_new;
            beta = cdiv(beta, om);
            beta = cdiv(beta, rho);
        }
        st[kBeta * R + rr] = beta;
        st[kTmpA * R + rr] = st[kRho * R + rr];  // rho of this iteration (reliable updates)
        st[kRho * R + rr] = rho_new;
    }
    __device__ XROp shifted(long off, int base) const {
        return {x + off, r + off, p + off, shat + off, s + off, t + off, r0 + off, st + base, R};
    }
};

template <typename CS, typename T>
__global__ __launch_bounds__(256) void k_step_s(CS *__restrict__ sv, const CS *__restrict__ r,
                                               const CS *__restrict__ v, const C64 *__restrict__ st,
                                               long N, int R) {
    long g, stride;
    int rhs;
    if (!rhs_lane(R, g, stride, rhs)) return;
    const Cx<T> al = tc<T>(st[kAlpha * R + rhs]);
    for (long e = g; e < N; e += stride) store_cx(sv + e, load_cx<T>(r + e) - al * load_cx<T>(v + e));
}

template <typename CS, typename T>
__global__ __launch_bounds__(256) void k_step_p(CS *__restrict__ p, const CS *__restrict__ r,
                                               const CS *__restrict__ v, const C64 *__restrict__ st,
                                               long N, int R) {
    long g, stride;
    int rhs;
    if (!rhs_lane(R, g, stride, rhs)) return;
    const Cx<T> om = tc<T>(st[kOmega * R + rhs]);
    const Cx<T> be = tc<T>(st[kBeta * R + rhs] * st[kPScale * R + rhs]);
    for (long e = g; e < N; e += stride) {
        const Cx<T> q = load_cx<T>(p + e) - om * load_cx<T>(v + e);
        store_cx(p + e, load_cx<T>(r + e) + be * q);
    }
}

template <typename CD, typename CS>
__global__ __launch_bounds__(256) void k_convert(CD *__restrict__ dst, const CS *__restrict__ src,
                                                long N, const double *__restrict__ scale, int R) {
    long g, stride;
    int rhs;
    if (!rhs_lane(R, g, stride, rhs)) return;
    const double sc = scale ? scale[rhs] : 1.0;
    for (long e = g; e < N; e += stride) {
        Cx<double> v = load_cx<double>(src + e);
        if (scale) v = sc * v;
        store_cx(dst + e, v);
    }
}

__global__ void k_fill(C64 *dst, C64 v, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) dst[i] = v;
}

// ------------------------------------------------------------ batched groups
// blockIdx.y = group; the group's vectors start n * base entries into each
// batch vector and its scalars `base` entries into each slot.
__device__ __forceinline__ bool seg_group(const SegTab &tab, int &grp, int &R, long &off, int &base) {
    grp = blockIdx.y;
    if (tab.done && tab.done[grp]) return false;
    R = tab.R[grp];
    base = tab.base[grp];
    off = tab.n * base;
    return true;
}

// Reduction over every group (grid.y = group), partial pass.
template <class Op>
__global__ __launch_bounds__(kReduceThreads) void k_reduce_seg(Op op, SegTab tab, double *__restrict__ partials,
                                                              int max_blocks) {
    int grp, R, base;
    long off;
    if (!seg_group(tab, grp, R, off, base)) return;
    reduce_partial(op.shifted(off, base), tab.n * R, R, partials + size_t(grp) * max_blocks * tab.Rmax * 4);
}

// Finishing pass, one block per group; with GROUP_DONE it also counts the
// iteration and raises the group's done flag once every RHS has
// ||r||^2 <= st[kBnorm2] (the squared stopping threshold).
template <class Op, bool GROUP_DONE>
__global__ __launch_bounds__(kReduceThreads) void k_reduce_seg_finish(Op op, SegTab tab,
                                                                     const double *__restrict__ partials,
                                                                     int max_blocks, int nblk, const C64 *st,
                                                                     int *done, int *iters) {
    int grp, R, base;
    long off;
    if (!seg_group(tab, grp, R, off, base)) return;
    reduce_finish(op.shifted(off, base), nblk, R, partials + size_t(grp) * max_blocks * tab.Rmax * 4);
    if constexpr (GROUP_DONE) {
        __syncthreads();
        if (threadIdx.x == 0) {
            bool all = true;
            for (int r = 0; r < R; ++r)
                all = all && st[kRnorm2 * tab.Rtot + base + r].re <= st[kBnorm2 * tab.Rtot + base + r].re;
            iters[grp] += 1;
            if (all) done[grp] = 1;
        }
    }
}

template <typename CS, typename T>
__global__ __launch_bounds__(256) void k_step_s_seg(CS *sv, const CS *r, const CS *v, const C64 *st, SegTab tab) {
    int grp, R, base;
    long off, g, stride;
    int rhs;
    if (!seg_group(tab, grp, R, off, base) || !rhs_lane(R, g, stride, rhs)) return;
    const Cx<T> al = tc<T>(st[kAlpha * tab.Rtot + base + rhs]);
    const long N = tab.n * R;
    sv += off;
    r += off;
    v += off;
    for (long e = g; e < N; e += stride) store_cx(sv + e, load_cx<T>(r + e) - al * load_cx<T>(v + e));
}

template <typename CS, typename T>
__global__ __launch_bounds__(256) void k_step_p_seg(CS *p, const CS *r, const CS *v, const C64 *st, SegTab tab) {
    int grp, R, base;
    long off, g, stride;
    int rhs;
    if (!seg_group(tab, grp, R, off, base) || !rhs_lane(R, g, stride, rhs)) return;
    const Cx<T> om = tc<T>(st[kOmega * tab.Rtot + base + rhs]);
    const Cx<T> be = tc<T>(st[kBeta * tab.Rtot + base + rhs]);
    const long N = tab.n * R;
    p += off;
    r += off;
    v += off;
    for (long e = g; e < N; e += stride) {
        const Cx<T> q = load_cx<T>(p + e) - om * load_cx<T>(v + e);
        store_cx(p + e, load_cx<T>(r + e) + be * q);
    }
}

// dst = scale[rhs] * src (scale may be null)
template <typename CD, typename CS>
__global__ __launch_bounds__(256) void k_convert_seg(CD *dst, const CS *src, const double *scale, SegTab tab) {
    int grp, R, base;
    long off, g, stride;
    int rhs;
    if (!seg_group(tab, grp, R, off, base) || !rhs_lane(R, g, stride, rhs)) return;
    const double sc = scale ? scale[base + rhs] : 1.0;
    const long N = tab.n * R;
    dst += off;
    src += off;
    for (long e = g; e < N; e += stride) {
        Cx<double> v = load_cx<double>(src + e);
        if (scale) v = sc * v;
        store_cx(dst + e, v);
    }
}

// res = x - a[rhs] y
template <typename CS, typename T>
__global__ __launch_bounds__(256) void k_xsay_seg(CS *res, const CS *x, const CS *y, const C64 *a, SegTab tab) {
    int grp, R, base;
    long off, g, stride;
    int rhs;
    if (!seg_group(tab, grp, R, off, base) || !rhs_lane(R, g, stride, rhs)) return;
    const Cx<T> ar = tc<T>(a[base + rhs]);
    const long N = tab.n * R;
    res += off;
    x += off;
    y += off;
    for (long e = g; e < N; e += stride) store_cx(res + e, load_cx<T>(x + e) - ar * load_cx<T>(y + e));
}

// xhi += scale[rhs] * xlo (FP64 accumulation)
template <typename CS>
__global__ __launch_bounds__(256) void k_accum_seg(Cx<double> *xhi, const CS *xlo, const double *scale, SegTab tab) {
    int grp, R, base;
    long off, g, stride;
    int rhs;
    if (!seg_group(tab, grp, R, off, base) || !rhs_lane(R, g, stride, rhs)) return;
    const double sc = scale[base + rhs];
    const long N = tab.n * R;
    xhi += off;
    xlo += off;
    // four entries per pass, every load issued before the first use
    constexpr int U = 4;
    long e = g;
    for (; e + (U - 1) * stride < N; e += U * stride) {
        CS v[U];
        Cx<double> h[U];
#pragma unroll
        for (int k = 0; k < U; ++k) {
            v[k] = xlo[e + k * stride];
            h[k] = xhi[e + k * stride];
        }
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const Cx<double> w = load_cx<double>(&v[k]);
            h[k].re = fma(sc, w.re, h[k].re);
            h[k].im = fma(sc, w.im, h[k].im);
            xhi[e + k * stride] = h[k];
        }
    }
    for (; e < N; e += stride) {
        const Cx<double> v = load_cx<double>(xlo + e);
        Cx<double> h = xhi[e];
        h.re = fma(sc, v.re, h.re);
        h.im = fma(sc, v.im, h.im);
        xhi[e] = h;
    }
}

inline int stream_blocks(long N, int R) {
    const int per_block = (256 / R) * R;
    long want = (N + per_block - 1) / per_block;
    const long cap = 256L * 16;  // 16 workgroups per CU on 256 CUs
    return int(want < cap ? (want > 0 ? want : 1) : cap);
}

template <class Op>
hipError_t run_reduce(const Op &op, long n, int R, ReduceWork &w, int slot, hipStream_t s) {
    if (R < 1 || R > kReduceThreads || slot < 0 || slot >= kReduceSlots) return hipErrorInvalidValue;
    const int nb = reduce_blocks(n, R, w.max_blocks);
    double *part = w.partials + size_t(slot) * w.max_blocks * R * 4;
    hipLaunchKernelGGL(k_reduce<Op>, dim3(nb), dim3(kReduceThreads), 0, s, op, n * R, R, part);
    QCU_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_finish<Op>, dim3(1), dim3(kReduceThreads), 0, s, op, nb, R, part);
    return hipGetLastError();
}

}  // namespace

int reduce_blocks(long n, int R, int max_blocks) {
    const int per_block = (kReduceThreads / R) * R;
    const long N = n * R;
    long nb = (N + per_block * 8L - 1) / (per_block * 8L);  // >= 8 entries per lane
    if (nb > max_blocks) nb = max_blocks;
    if (nb < 1) nb = 1;
    return int(nb);
}

hipError_t reduce_work_alloc(ReduceWork &w, int R) {
    // 1024 blocks (4 per CU); 2048 measured slower for the 16-bit reductions and no faster
    // for the FP64 ones
    w.max_blocks = 1024;
    // one partials region per slot: reductions on different slots may be in flight together
    QCU_HIP_TRY(hipMalloc(&w.partials, sizeof(double) * 4 * size_t(w.max_blocks) * R * kReduceSlots));
    return hipSuccess;
}

void reduce_work_free(ReduceWork &w) {
    if (w.partials) (void)hipFree(w.partials);
    w = ReduceWork{};
}

hipError_t launch_axpy(const AxpyArgs &a, hipStream_t s) {
    if (a.R < 1 || a.R > 256) return hipErrorInvalidValue;
    const long N = a.n * a.R;
    if (N == 0) return hipSuccess;
    const int nb = stream_blocks(N, a.R);
    QCU_PREC_DISPATCH(a.prec, {
        using CS = typename CxStore<typename PrecT<P>::S>::type;
        using T = typename PrecT<P>::T;
        auto res = static_cast<CS *>(a.res);
        auto x = static_cast<const CS *>(a.x);
        auto y = static_cast<const CS *>(a.y);
        auto z = static_cast<const CS *>(a.z);
        switch (a.kind) {
            case kXsay: hipLaunchKernelGGL((k_axpy<kXsay, CS, T>), dim3(nb), dim3(256), 0, s, res, x, y, z, a.a, a.b, a.c, N, a.R); break;
            case kXpay: hipLaunchKernelGGL((k_axpy<kXpay, CS, T>), dim3(nb), dim3(256), 0, s, res, x, y, z, a.a, a.b, a.c, N, a.R); break;
            case kAxpby: hipLaunchKernelGGL((k_axpy<kAxpby, CS, T>), dim3(nb), dim3(256), 0, s, res, x, y, z, a.a, a.b, a.c, N, a.R); break;
            case kAxpbypcz: hipLaunchKernelGGL((k_axpy<kAxpbypcz, CS, T>), dim3(nb), dim3(256), 0, s, res, x, y, z, a.a, a.b, a.c, N, a.R); break;
            default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    })
}

hipError_t launch_dotc(C64 *res, const void *a, const void *b, long n, int R, int prec,
                       ReduceWork &w, int slot, hipStream_t s) {
    QCU_PREC_DISPATCH(prec, {
        using CS = typename CxStore<typename PrecT<P>::S>::type;
        DotOp<CS> op{static_cast<const CS *>(a), static_cast<const CS *>(b), res};
        return run_reduce(op, n, R, w, slot, s);
    })
}

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per device.
hipError_t scratch_address(double **out) {
    int dev = 0;
    QCU_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        QCU_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_dev[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_dev[dev];
    return hipSuccess;
}
}  // namespace

hipError_t launch_dotc_noalloc(C64 *res, const void *a, const void *b, long n, int R, int prec, hipStream_t s) {
    if (R < 1 || R > kReduceThreads) return hipErrorInvalidValue;
    double *part = nullptr;
    QCU_HIP_TRY(scratch_address(&part));
    // every block writes all of its R partials before the finish kernel reads
    // them, so the scratch needs no clearing between calls
    ReduceWork w;
    w.partials = part;
    w.max_blocks = std::min(1024, int(kNoAllocScratchBytes / (sizeof(double) * 4 * R)));
    return launch_dotc(res, a, b, n, R, prec, w, 0, s);
}

hipError_t launch_norm(double *res, const void *a, long n, int R, int prec, bool take_sqrt,
                       ReduceWork &w, int slot, hipStream_t s) {
    QCU_PREC_DISPATCH(prec, {
        using CS = typename CxStore<typename PrecT<P>::S>::type;
        NormOp<CS> op{static_cast<const CS *>(a), res, take_sqrt};
        return run_reduce(op, n, R, w, slot, s);
    })
}

hipError_t launch_convert(void *dst, int dst_prec, const void *src, int src_prec, long count,
                          const double *scale, int R, hipStream_t s) {
    if (count == 0) return hipSuccess;
    if (R < 1 || R > 256) return hipErrorInvalidValue;
    const int nb = stream_blocks(count, R);
    QCU_PREC_DISPATCH(dst_prec, {
        using CD = typename CxStore<typename PrecT<P>::S>::type;
        switch (src_prec) {
            case kDouble: hipLaunchKernelGGL((k_convert<CD, Cx<double>>), dim3(nb), dim3(256), 0, s, static_cast<CD *>(dst), static_cast<const Cx<double> *>(src), count, scale, R); break;
            case kSingle: hipLaunchKernelGGL((k_convert<CD, Cx<float>>), dim3(nb), dim3(256), 0, s, static_cast<CD *>(dst), static_cast<const Cx<float> *>(src), count, scale, R); break;
            case kHalf: hipLaunchKernelGGL((k_convert<CD, Half2Store>), dim3(nb), dim3(256), 0, s, static_cast<CD *>(dst), static_cast<const Half2Store *>(src), count, scale, R); break;
            case kBF16: hipLaunchKernelGGL((k_convert<CD, Bf162Store>), dim3(nb), dim3(256), 0, s, static_cast<CD *>(dst), static_cast<const Bf162Store *>(src), count, scale, R); break;
            default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    })
}

hipError_t launch_residual_lo(double *nrm, void *lo, int lo_prec, const void *b, const void *y, const double *scale,
                              long n, int R, ReduceWork &w, int slot, hipStream_t s) {
    const auto *bb = static_cast<const Cx<double> *>(b);
    const auto *yy = static_cast<const Cx<double> *>(y);
    switch (lo_prec) {
        case kSingle: return run_reduce(ResidualLoOp<Cx<float>>{static_cast<Cx<float> *>(lo), bb, yy, scale, nrm}, n, R, w, slot, s);
        case kHalf: return run_reduce(ResidualLoOp<Half2Store>{static_cast<Half2Store *>(lo), bb, yy, scale, nrm}, n, R, w, slot, s);
        case kBF16: return run_reduce(ResidualLoOp<Bf162Store>{static_cast<Bf162Store *>(lo), bb, yy, scale, nrm}, n, R, w, slot, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_xsay_norm(double *nrm, void *res, const void *x, const void *y, long n, int R, int prec,
                            bool take_sqrt, ReduceWork &w, int slot, hipStream_t s) {
    QCU_PREC_DISPATCH(prec, {
        using CS = typename CxStore<typename PrecT<P>::S>::type;
        XsayNormOp<CS> op{static_cast<CS *>(res), static_cast<const CS *>(x), static_cast<const CS *>(y), nrm,
                          take_sqrt};
        return run_reduce(op, n, R, w, slot, s);
    })
}

hipError_t launch_step_alpha(C64 *st, const void *r0, const void *v, long n, int R, int prec,
                             ReduceWork &w, hipStream_t s) {
    QCU_PREC_DISPATCH(prec, {
        using CS = typename CxStore<typename PrecT<P>::S>::type;
        AlphaOp<CS> op{static_cast<const CS *>(r0), static_cast<const CS *>(v), st, R};
        return run_reduce(op, n, R, w, 0, s);
    })
}

hipError_t launch_finish_alpha(C64 *st, int R, const double *partials, int nblk, hipStream_t s) {
    if (R < 1 || R > kReduceThreads || nblk < 1) return hipErrorInvalidValue;
    AlphaOp<Cx<double>> op{nullptr, nullptr, st, R};
    hipLaunchKernelGGL(k_reduce_finish<AlphaOp<Cx<double>>>, dim3(1), dim3(kReduceThreads), 0, s, op, nblk, R,
                       partials);
    return hipGetLastError();
}

hipError_t launch_finish_omega(C64 *st, int R, const double *partials, int nblk, hipStream_t s) {
    if (R < 1 || R > kReduceThreads || nblk < 1) return hipErrorInvalidValue;
    OmegaOp<Cx<double>> op{nullptr, nullptr, st, R};
    hipLaunchKernelGGL(k_reduce_finish<OmegaOp<Cx<double>>>, dim3(1), dim3(kReduceThreads), 0, s, op, nblk, R,
                       partials);
    return hipGetLastError();
}

hipError_t launch_step_s(void *sv, const void *r, const void *v, const C64 *st, long n, int R,
                         int prec, hipStream_t s) {
    const long N = n * R;
    const int nb = stream_blocks(N, R);
    QCU_PREC_DISPATCH(prec, {
        using CS = typename CxStore<typename PrecT<P>::S>::type;
        using T = typename PrecT<P>::T;
        hipLaunchKernelGGL((k_step_s<CS, T>), dim3(nb), dim3(256), 0, s, static_cast<CS *>(sv),
                           static_cast<const CS *>(r), static_cast<const CS *>(v), st, N, R);
        return hipGetLastError();
    })
}

hipError_t launch_step_omega(C64 *st, const void *t, const void *sv, long n, int R, int prec,
                             ReduceWork &w, hipStream_t s) {
    QCU_PREC_DISPATCH(prec, {
        using CS = typename CxStore<typename PrecT<P>::S>::type;
        OmegaOp<CS> op{static_cast<const CS *>(t), static_cast<const CS *>(sv), st, R};
        return run_reduce(op, n, R, w, 1, s);
    })
}

hipError_t launch_step_xr(void *x, void *r, const void *p, const void *shat, const void *sv,
                          const void *t, const void *r0, C64 *st, long n, int R, int prec,
                          ReduceWork &w, hipStream_t s) {
    QCU_PREC_DISPATCH(prec, {
        using CS = typename CxStore<typename PrecT<P>::S>::type;
        using T = typename PrecT<P>::T;
        XROp<CS, T> op{static_cast<CS *>(x), static_cast<CS *>(r), static_cast<const CS *>(p),
                       static_cast<const CS *>(shat), static_cast<const CS *>(sv),
                       static_cast<const CS *>(t),
                       static_cast<const CS *>(r0), st, R};
        return run_reduce(op, n, R, w, 2, s);
    })
}

hipError_t launch_step_p(void *p, const void *r, const void *v, const C64 *st, long n, int R,
                         int prec, hipStream_t s) {
    const long N = n * R;
    const int nb = stream_blocks(N, R);
    QCU_PREC_DISPATCH(prec, {
        using CS = typename CxStore<typename PrecT<P>::S>::type;
        using T = typename PrecT<P>::T;
        hipLaunchKernelGGL((k_step_p<CS, T>), dim3(nb), dim3(256), 0, s, static_cast<CS *>(p),
                           static_cast<const CS *>(r), static_cast<const CS *>(v), st, N, R);
        return hipGetLastError();
    })
}

// ------------------------------------------------------------ batched groups
hipError_t seg_work_alloc(SegWork &w, int G, int Rmax) {
    seg_work_free(w);
    w.max_blocks = 2048;
    w.G = G;
    w.Rmax = Rmax;
    QCU_HIP_TRY(hipMalloc(&w.partials, sizeof(double) * 4 * size_t(w.max_blocks) * Rmax * G));
    return hipSuccess;
}

void seg_work_free(SegWork &w) {
    if (w.partials) (void)hipFree(w.partials);
    w = SegWork{};
}

namespace {

bool seg_ok(const SegTab &tab, const SegWork *w) {
    if (tab.G < 1 || tab.G > kMaxGroups || tab.n <= 0) return false;
    for (int k = 0; k < tab.G; ++k)
        if (tab.R[k] < 1 || tab.R[k] > tab.Rmax || tab.R[k] > kReduceThreads) return false;
    return !w || (w->partials && tab.G <= w->G && tab.Rmax <= w->Rmax);
}

template <class Op, bool GD = false>
hipError_t run_seg_reduce(const Op &op, const SegTab &tab, SegWork &w, hipStream_t s, const C64 *st = nullptr,
                          int *done = nullptr, int *iters = nullptr) {
    if (!seg_ok(tab, &w)) return hipErrorInvalidValue;
    const int nb = reduce_blocks(tab.n, tab.Rmax, w.max_blocks);
    hipLaunchKernelGGL(k_reduce_seg<Op>, dim3(nb, tab.G), dim3(kReduceThreads), 0, s, op, tab, w.partials,
                       w.max_blocks);
    QCU_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL((k_reduce_seg_finish<Op, GD>), dim3(1, tab.G), dim3(kReduceThreads), 0, s, op, tab,
                       w.partials, w.max_blocks, nb, st, done, iters);
    return hipGetLastError();
}

dim3 seg_grid(const SegTab &tab) { return dim3(stream_blocks(tab.n * tab.Rmax, tab.Rmax), tab.G); }

}  // namespace

hipError_t seg_dotc(C64 *res, const void *a, const void *b, int prec, const SegTab &tab, SegWork &w, int slot,
                    hipStream_t s) {
    QCU_PREC_DISPATCH(prec, {
        using CS = typename CxStore<typename PrecT<P>::S>::type;
        DotOp<CS> op{static_cast<const CS *>(a), static_cast<const CS *>(b), res};
        (void)slot;
        return run_seg_reduce(op, tab, w, s);
    })
}

hipError_t seg_norm(double *res, const void *a, int prec, bool take_sqrt, const SegTab &tab, SegWork &w, int slot,
                    hipStream_t s) {
    QCU_PREC_DISPATCH(prec, {
        using CS = typename CxStore<typename PrecT<P>::S>::type;
        NormOp<CS> op{static_cast<const CS *>(a), res, take_sqrt};
        (void)slot;
        return run_seg_reduce(op, tab, w, s);
    })
}

hipError_t seg_step_alpha(C64 *st, const void *r0, const void *v, int prec, const SegTab &tab, SegWork &w,
                          hipStream_t s) {
    QCU_PREC_DISPATCH(prec, {
        using CS = typename CxStore<typename PrecT<P>::S>::type;
        AlphaOp<CS> op{static_cast<const CS *>(r0), static_cast<const CS *>(v), st, tab.Rtot};
        return run_seg_reduce(op, tab, w, s);
    })
}

hipError_t seg_step_omega(C64 *st, const void *t, const void *sv, int prec, const SegTab &tab, SegWork &w,
                          hipStream_t s) {
    QCU_PREC_DISPATCH(prec, {
        using CS = typename CxStore<typename PrecT<P>::S>::type;
        OmegaOp<CS> op{static_cast<const CS *>(t), static_cast<const CS *>(sv), st, tab.Rtot};
        return run_seg_reduce(op, tab, w, s);
    })
}

hipError_t seg_step_xr(void *x, void *r, const void *p, const void *sv, const void *t, const void *r0, C64 *st,
                       int *done, int *iters, int prec, const SegTab &tab, SegWork &w, hipStream_t s) {
    QCU_PREC_DISPATCH(prec, {
        using CS = typename CxStore<typename PrecT<P>::S>::type;
        using T = typename PrecT<P>::T;
        XROp<CS, T> op{static_cast<CS *>(x), static_cast<CS *>(r), static_cast<const CS *>(p),
                       static_cast<const CS *>(sv), static_cast<const CS *>(sv), static_cast<const CS *>(t),
                       static_cast<const CS *>(r0), st, tab.Rtot};
        return run_seg_reduce<decltype(op), true>(op, tab, w, s, st, done, iters);
    })
}

hipError_t seg_step_s(void *sv, const void *r, const void *v, const C64 *st, int prec, const SegTab &tab,
                      hipStream_t s) {
    if (!seg_ok(tab, nullptr)) return hipErrorInvalidValue;
    QCU_PREC_DISPATCH(prec, {
        using CS = typename CxStore<typename PrecT<P>::S>::type;
        using T = typename PrecT<P>::T;
        hipLaunchKernelGGL((k_step_s_seg<CS, T>), seg_grid(tab), dim3(256), 0, s, static_cast<CS *>(sv),
                           static_cast<const CS *>(r), static_cast<const CS *>(v), st, tab);
        return hipGetLastError();
    })
}

hipError_t seg_step_p(void *p, const void *r, const void *v, const C64 *st, int prec, const SegTab &tab,
                      hipStream_t s) {
    if (!seg_ok(tab, nullptr)) return hipErrorInvalidValue;
    QCU_PREC_DISPATCH(prec, {
        using CS = typename CxStore<typename PrecT<P>::S>::type;
        using T = typename PrecT<P>::T;
        hipLaunchKernelGGL((k_step_p_seg<CS, T>), seg_grid(tab), dim3(256), 0, s, static_cast<CS *>(p),
                           static_cast<const CS *>(r), static_cast<const CS *>(v), st, tab);
        return hipGetLastError();
    })
}

hipError_t seg_xsay(void *res, const void *x, const void *y, const C64 *a, int prec, const SegTab &tab,
                    hipStream_t s) {
    if (!seg_ok(tab, nullptr)) return hipErrorInvalidValue;
    QCU_PREC_DISPATCH(prec, {
        using CS = typename CxStore<typename PrecT<P>::S>::type;
        using T = typename PrecT<P>::T;
        hipLaunchKernelGGL((k_xsay_seg<CS, T>), seg_grid(tab), dim3(256), 0, s, static_cast<CS *>(res),
                           static_cast<const CS *>(x), static_cast<const CS *>(y), a, tab);
        return hipGetLastError();
    })
}

hipError_t seg_convert(void *dst, int dst_prec, const void *src, int src_prec, const double *scale,
                       const SegTab &tab, hipStream_t s) {
    if (!seg_ok(tab, nullptr)) return hipErrorInvalidValue;
    QCU_PREC_DISPATCH(dst_prec, {
        using CD = typename CxStore<typename PrecT<P>::S>::type;
        switch (src_prec) {
            case kDouble: hipLaunchKernelGGL((k_convert_seg<CD, Cx<double>>), seg_grid(tab), dim3(256), 0, s, static_cast<CD *>(dst), static_cast<const Cx<double> *>(src), scale, tab); break;
            case kSingle: hipLaunchKernelGGL((k_convert_seg<CD, Cx<float>>), seg_grid(tab), dim3(256), 0, s, static_cast<CD *>(dst), static_cast<const Cx<float> *>(src), scale, tab); break;
            case kHalf: hipLaunchKernelGGL((k_convert_seg<CD, Half2Store>), seg_grid(tab), dim3(256), 0, s, static_cast<CD *>(dst), static_cast<const Half2Store *>(src), scale, tab); break;
            case kBF16: hipLaunchKernelGGL((k_convert_seg<CD, Bf162Store>), seg_grid(tab), dim3(256), 0, s, static_cast<CD *>(dst), static_cast<const Bf162Store *>(src), scale, tab); break;
            default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    })
}

hipError_t seg_accum(void *xhi, const void *xlo, int lo_prec, const double *scale, const SegTab &tab,
                     hipStream_t s) {
    if (!seg_ok(tab, nullptr) || !scale) return hipErrorInvalidValue;
    auto *h = static_cast<Cx<double> *>(xhi);
    switch (lo_prec) {
        case kDouble: hipLaunchKernelGGL((k_accum_seg<Cx<double>>), seg_grid(tab), dim3(256), 0, s, h, static_cast<const Cx<double> *>(xlo), scale, tab); break;
        case kSingle: hipLaunchKernelGGL((k_accum_seg<Cx<float>>), seg_grid(tab), dim3(256), 0, s, h, static_cast<const Cx<float> *>(xlo), scale, tab); break;
        case kHalf: hipLaunchKernelGGL((k_accum_seg<Half2Store>), seg_grid(tab), dim3(256), 0, s, h, static_cast<const Half2Store *>(xlo), scale, tab); break;
        case kBF16: hipLaunchKernelGGL((k_accum_seg<Bf162Store>), seg_grid(tab), dim3(256), 0, s, h, static_cast<const Bf162Store *>(xlo), scale, tab); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_scalar_fill(C64 *dst, C64 value, int count, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill, dim3(div_up(count, 256)), dim3(256), 0, s, dst, value, count);
    return hipGetLastError();
}

}  // namespace qcu
