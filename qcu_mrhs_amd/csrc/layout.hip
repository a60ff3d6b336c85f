// layout.hip -- data-format kernels: gather/scatter between the caller's
// per-RHS vectors and the internal RHS-innermost buffers, and the gauge
// even/odd reordering.
//
// Gather (src/data_format/qcu_data_format_shift.cu:60-76,185-206;
// include/kernel/copy_vector/copy_color_spinor.cuh:16-39,70-84 of the
// reference): R vectors [site][spin][colour] -> [site][spin][colour][rhs],
// with precision conversion.  One lane per destination complex number, so
// the RHS-innermost side is written with unit stride; the R source pointers
// travel in the kernel arguments (no pointer table in device memory).
#include <algorithm>

#include "layout.h"

namespace qcu {
namespace {

template <typename CD, typename CS>
__global__ __launch_bounds__(256) void k_gather(CD *__restrict__ dst, PtrTable src, long n, int R) {
    const long N = n * R;
    const long stride = long(gridDim.x) * blockDim.x;
    for (long e = long(blockIdx.x) * blockDim.x + threadIdx.x; e < N; e += stride) {
        const long j = e / R;
        const int r = int(e - j * R);
        const CS *s = static_cast<const CS *>(src.p[r]);
        store_cx(dst + e, load_cx<double>(s + j));
    }
}

// Scatter through LDS: a block reads a tile of tj consecutive entries x all R
// RHS contiguously from the RHS-innermost source, then writes tj consecutive
// entries of each destination vector (one contiguous run per RHS).  Writing
// straight from the source order would give every destination vector 64-byte
// pieces per wavefront, with the halves of one 128-byte line written by
// different blocks (measured 3.4 TB/s against 4.9 TB/s for the gather).
// The LDS row pitch R + 1 keeps the transposed reads free of bank conflicts.
template <typename CD, typename CS>
__global__ __launch_bounds__(256) void k_scatter(PtrTable dst, const CS *__restrict__ src, long n, int R, int tj) {
    extern __shared__ Cx<double> tile[];  // [tj][R + 1]
    const int pitch = R + 1;
    for (long j0 = long(blockIdx.x) * tj; j0 < n; j0 += long(gridDim.x) * tj) {
        const int cnt = int(n - j0 < tj ? n - j0 : tj);
        const long base = j0 * R;
        for (int i = threadIdx.x; i < cnt * R; i += blockDim.x) {
            const int jj = i / R, r = i - jj * R;
            tile[jj * pitch + r] = load_cx<double>(src + base + i);
        }
        __syncthreads();
        for (int i = threadIdx.x; i < cnt * R; i += blockDim.x) {
            const int r = i / cnt, jj = i - r * cnt;
            CD *d = static_cast<CD *>(const_cast<void *>(dst.p[r]));
            store_cx(d + j0 + jj, tile[jj * pitch + r]);
        }
        __syncthreads();
    }
}

// lexicographic [mu][t][z][y][x][Nc][Nc] <-> even/odd [mu][parity][Vh][Nc][Nc]
template <typename C>
__global__ __launch_bounds__(256) void k_gauge_eo(C *__restrict__ eo, C *__restrict__ lex, Geom g,
                                                 int ll, int reverse) {
    const long V = 2L * g.Vh;
    const long total = 4L * V * ll;
    const long stride = long(gridDim.x) * blockDim.x;
    for (long e = long(blockIdx.x) * blockDim.x + threadIdx.x; e < total; e += stride) {
        const long k = e % ll;
        const long link = e / ll;
        const int mu = int(link / V);
        long s = link % V;
        const int x = int(s % g.L[0]);
        s /= g.L[0];
        const int y = int(s % g.L[1]);
        s /= g.L[1];
        const int z = int(s % g.L[2]);
        const int t = int(s / g.L[2]);
        const int p = (x + y + z + t) & 1;
        const long cb = ((long(t) * g.L[2] + z) * g.L[1] + y) * g.Lxh + x / 2;
        const long ie = ((long(mu) * 2 + p) * g.Vh + cb) * ll + k;
        if (reverse) lex[e] = eo[ie];
        else eo[ie] = lex[e];
    }
}

inline int grid_for(long N) {
    long nb = (N + 255) / 256;
    if (nb > 4096) nb = 4096;
    return int(nb > 0 ? nb : 1);
}

template <typename CD>
hipError_t gather_src(void *dst, const PtrTable &src, int src_prec, long n, int R, hipStream_t s) {
    const int nb = grid_for(n * R);
    switch (src_prec) {
        case kDouble: hipLaunchKernelGGL((k_gather<CD, Cx<double>>), dim3(nb), dim3(256), 0, s, static_cast<CD *>(dst), src, n, R); break;
        case kSingle: hipLaunchKernelGGL((k_gather<CD, Cx<float>>), dim3(nb), dim3(256), 0, s, static_cast<CD *>(dst), src, n, R); break;
        case kHalf: hipLaunchKernelGGL((k_gather<CD, Half2Store>), dim3(nb), dim3(256), 0, s, static_cast<CD *>(dst), src, n, R); break;
        case kBF16: hipLaunchKernelGGL((k_gather<CD, Bf162Store>), dim3(nb), dim3(256), 0, s, static_cast<CD *>(dst), src, n, R); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename CS>
hipError_t scatter_dst(const PtrTable &dst, int dst_prec, const void *src, long n, int R, hipStream_t s) {
    // tile of about 2176 complex numbers (34 KiB of LDS with the padded pitch)
    const int tj = std::max(1, 2176 / (R + 1));
    const size_t lds = size_t(tj) * (R + 1) * sizeof(Cx<double>);
    const long tiles = (n + tj - 1) / tj;
    const int nb = int(tiles < 8192 ? (tiles > 0 ? tiles : 1) : 8192);
    const auto *sv = static_cast<const CS *>(src);
    switch (dst_prec) {
        case kDouble: hipLaunchKernelGGL((k_scatter<Cx<double>, CS>), dim3(nb), dim3(256), lds, s, dst, sv, n, R, tj); break;
        case kSingle: hipLaunchKernelGGL((k_scatter<Cx<float>, CS>), dim3(nb), dim3(256), lds, s, dst, sv, n, R, tj); break;
        case kHalf: hipLaunchKernelGGL((k_scatter<Half2Store, CS>), dim3(nb), dim3(256), lds, s, dst, sv, n, R, tj); break;
        case kBF16: hipLaunchKernelGGL((k_scatter<Bf162Store, CS>), dim3(nb), dim3(256), lds, s, dst, sv, n, R, tj); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_gather(void *dst, int dst_prec, const PtrTable &src, int src_prec, long n, int R,
                         hipStream_t s) {
    if (R < 1 || R > kMaxPtrs) return hipErrorInvalidValue;
    switch (dst_prec) {
        case kDouble: return gather_src<Cx<double>>(dst, src, src_prec, n, R, s);
        case kSingle: return gather_src<Cx<float>>(dst, src, src_prec, n, R, s);
        case kHalf: return gather_src<Half2Store>(dst, src, src_prec, n, R, s);
        case kBF16: return gather_src<Bf162Store>(dst, src, src_prec, n, R, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_scatter(const PtrTable &dst, int dst_prec, const void *src, int src_prec, long n,
                          int R, hipStream_t s) {
    if (R < 1 || R > kMaxPtrs) return hipErrorInvalidValue;
    switch (src_prec) {
        case kDouble: return scatter_dst<Cx<double>>(dst, dst_prec, src, n, R, s);
        case kSingle: return scatter_dst<Cx<float>>(dst, dst_prec, src, n, R, s);
        case kHalf: return scatter_dst<Half2Store>(dst, dst_prec, src, n, R, s);
        case kBF16: return scatter_dst<Bf162Store>(dst, dst_prec, src, n, R, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_gauge_eo(void *eo, void *lex, const Geom &g, int Nc, int prec, int reverse,
                           hipStream_t s) {
    const int ll = Nc * Nc;
    const int nb = grid_for(4L * 2 * g.Vh * ll);
    switch (prec) {
        case kDouble: hipLaunchKernelGGL((k_gauge_eo<Cx<double>>), dim3(nb), dim3(256), 0, s, static_cast<Cx<double> *>(eo), static_cast<Cx<double> *>(lex), g, ll, reverse); break;
        case kSingle: hipLaunchKernelGGL((k_gauge_eo<Cx<float>>), dim3(nb), dim3(256), 0, s, static_cast<Cx<float> *>(eo), static_cast<Cx<float> *>(lex), g, ll, reverse); break;
        case kHalf: hipLaunchKernelGGL((k_gauge_eo<Half2Store>), dim3(nb), dim3(256), 0, s, static_cast<Half2Store *>(eo), static_cast<Half2Store *>(lex), g, ll, reverse); break;
        case kBF16: hipLaunchKernelGGL((k_gauge_eo<Bf162Store>), dim3(nb), dim3(256), 0, s, static_cast<Bf162Store *>(eo), static_cast<Bf162Store *>(lex), g, ll, reverse); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace qcu
