// qcu_common.h -- shared definitions of the gfx950 multi-RHS Wilson library.
//
// Data layouts (the reference's; include/point/qcu_point.cuh:67-87,
// include/kernel/copy_vector/copy_color_spinor.cuh:27-38):
//   MRHS half vector  [Vh][spin 4][color Nc][rhs R]      (RHS innermost)
//   gauge             [mu 4][parity 2][Vh][row Nc][col Nc]
// Checkerboard index i = ((t*Lz + z)*Ly + y)*(Lx/2) + x/2, parity (x+y+z+t)&1.
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

namespace qcu {

enum Prec : int { kHalf = 0, kSingle = 1, kDouble = 2, kUndefined = 3, kBF16 = 4 };

inline size_t prec_bytes(int p) {  // bytes of one real component
    switch (p) {
        case kHalf: return 2;
        case kBF16: return 2;
        case kSingle: return 4;
        case kDouble: return 8;
        default: return 0;
    }
}

// Error reporting shared by every translation unit (defined in qcu_api.hip).
void set_error(int code, const char *fmt, ...);
int last_error();

#define QCU_HIP_TRY(expr)                                                              \
    do {                                                                               \
        hipError_t qcu_e_ = (expr);                                                    \
        if (qcu_e_ != hipSuccess) {                                                    \
            ::qcu::set_error(5, "%s:%d %s -> %s", __FILE__, __LINE__, #expr,           \
                             hipGetErrorString(qcu_e_));                               \
            return qcu_e_;                                                             \
        }                                                                              \
    } while (0)

// Lattice geometry of one GPU's (periodic) lattice.
struct Geom {
    int L[4];    // X, Y, Z, T
    int Lxh;     // Lx / 2
    int Vh;      // half volume
};

// A batch of right-hand-side groups stored back to back ("segments"): in every
// batch half vector, group g occupies entries [n*base[g], n*(base[g]+R[g])) in
// the [n][R[g]] layout (n = Vh*4*Nc entries per RHS); its per-RHS solver
// scalars are st[slot*Rtot + base[g] + r].  Batched kernels put the group in
// blockIdx.y and skip a group whose done flag is set.
constexpr int kMaxGroups = 8;
struct SegTab {
    int G = 0;
    int Rtot = 0;                  // RHS over all groups (scalar slot stride)
    int Rmax = 0;
    long n = 0;                    // complex entries per RHS of a half vector
    int R[kMaxGroups] = {};
    int base[kMaxGroups] = {};     // first batch-wide RHS index of the group
    int spw[kMaxGroups] = {};      // dslash sites per wavefront for the group (autotuned)
    const int *done = nullptr;     // device flags [G]; null = every group active
};

inline Geom make_geom(const int L[4]) {
    Geom g;
    for (int i = 0; i < 4; ++i) g.L[i] = L[i];
    g.Lxh = L[0] / 2;
    g.Vh = g.Lxh * L[1] * L[2] * L[3];
    return g;
}

// ---------------------------------------------------------------------------
// complex arithmetic on plain structs (no CUDA/HIP complex headers)
// ---------------------------------------------------------------------------
template <typename T>
struct alignas(2 * sizeof(T)) Cx {
    T re, im;
};

template <typename T>
__host__ __device__ __forceinline__ Cx<T> cx(T r, T i) { return Cx<T>{r, i}; }
template <typename T>
__host__ __device__ __forceinline__ Cx<T> operator+(Cx<T> a, Cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T>
__host__ __device__ __forceinline__ Cx<T> operator-(Cx<T> a, Cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T>
__host__ __device__ __forceinline__ Cx<T> operator*(Cx<T> a, Cx<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__host__ __device__ __forceinline__ Cx<T> operator*(T s, Cx<T> a) { return {s * a.re, s * a.im}; }
template <typename T>
__host__ __device__ __forceinline__ Cx<T> conj(Cx<T> a) { return {a.re, -a.im}; }
// conj(a) * b
template <typename T>
__host__ __device__ __forceinline__ Cx<T> cdot(Cx<T> a, Cx<T> b) {
    return {a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}
// acc += a * b
template <typename T>
__host__ __device__ __forceinline__ void cfma(Cx<T> &acc, Cx<T> a, Cx<T> b) {
    acc.re = fma(a.re, b.re, acc.re);
    acc.re = fma(-a.im, b.im, acc.re);
    acc.im = fma(a.re, b.im, acc.im);
    acc.im = fma(a.im, b.re, acc.im);
}
// acc += conj(a) * b
template <typename T>
__host__ __device__ __forceinline__ void cfma_conj(Cx<T> &acc, Cx<T> a, Cx<T> b) {
    acc.re = fma(a.re, b.re, acc.re);
    acc.re = fma(a.im, b.im, acc.re);
    acc.im = fma(a.re, b.im, acc.im);
    acc.im = fma(-a.im, b.re, acc.im);
}
// complex division a / b  (a * conj(b) / |b|^2)
template <typename T>
__host__ __device__ __forceinline__ Cx<T> cdiv(Cx<T> a, Cx<T> b) {
    const T d = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}
// multiply by i^K
template <int K, typename T>
__host__ __device__ __forceinline__ Cx<T> ipow(Cx<T> z) {
    if constexpr ((K & 3) == 0) return z;
    else if constexpr ((K & 3) == 1) return {-z.im, z.re};
    else if constexpr ((K & 3) == 2) return {-z.re, -z.im};
    else return {z.im, -z.re};
}

// ---------------------------------------------------------------------------
// storage-type conversion (compute in float/double; store in any precision)
// ---------------------------------------------------------------------------
// Pair of 16-bit values stored as one 32-bit word.
struct alignas(4) Half2Store { __half re, im; };
struct alignas(4) Bf162Store { __hip_bfloat16 re, im; };

template <typename S> struct CxStore;        // complex storage type for real type S
template <> struct CxStore<double> { using type = Cx<double>; };
template <> struct CxStore<float> { using type = Cx<float>; };
template <> struct CxStore<__half> { using type = Half2Store; };
template <> struct CxStore<__hip_bfloat16> { using type = Bf162Store; };

template <typename T>
__device__ __forceinline__ Cx<T> load_cx(const Cx<double> *p) { Cx<double> v = *p; return {T(v.re), T(v.im)}; }
template <typename T>
__device__ __forceinline__ Cx<T> load_cx(const Cx<float> *p) { Cx<float> v = *p; return {T(v.re), T(v.im)}; }
typedef _Float16 qcu_f16x2 __attribute__((ext_vector_type(2)));
typedef __bf16 qcu_bf16x2 __attribute__((ext_vector_type(2)));
// 16-bit pairs travel as one 32-bit word (re in the low half)
template <typename T>
__device__ __forceinline__ Cx<T> load_cx(const Half2Store *p) {
    const qcu_f16x2 v = __builtin_bit_cast(qcu_f16x2, *reinterpret_cast<const uint32_t *>(p));
    return {T(float(v[0])), T(float(v[1]))};
}
template <typename T>
__device__ __forceinline__ Cx<T> load_cx(const Bf162Store *p) {
    const uint32_t w = *reinterpret_cast<const uint32_t *>(p);
    return {T(__builtin_bit_cast(float, w << 16)), T(__builtin_bit_cast(float, w & 0xffff0000u))};
}
template <typename T>
__device__ __forceinline__ void store_cx(Cx<double> *p, Cx<T> v) { *p = Cx<double>{double(v.re), double(v.im)}; }
template <typename T>
__device__ __forceinline__ void store_cx(Cx<float> *p, Cx<T> v) { *p = Cx<float>{float(v.re), float(v.im)}; }
template <typename T>
__device__ __forceinline__ void store_cx(Half2Store *p, Cx<T> v) {
    const qcu_f16x2 h = {_Float16(float(v.re)), _Float16(float(v.im))};
    *reinterpret_cast<uint32_t *>(p) = __builtin_bit_cast(uint32_t, h);
}
template <typename T>
__device__ __forceinline__ void store_cx(Bf162Store *p, Cx<T> v) {
    const qcu_bf16x2 h = {__bf16(float(v.re)), __bf16(float(v.im))};
    *reinterpret_cast<uint32_t *>(p) = __builtin_bit_cast(uint32_t, h);
}

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------
constexpr int kNumXcd = 8;  // MI355X: 8 accelerator complex dies, 32 CUs each

// Workgroups are dispatched round-robin over the XCDs (block i -> XCD i % 8).
// Remap so each XCD walks one contiguous range of logical blocks: neighbouring
// sites then share that XCD's L2.  Bijective for any block count.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
    const int xcd = bid % kNumXcd;
    const int q = nblocks / kNumXcd, rem = nblocks % kNumXcd;
    const int start = xcd * q + (xcd < rem ? xcd : rem);
    return start + bid / kNumXcd;
}

inline int div_up(long a, long b) { return int((a + b - 1) / b); }

// compute units of the current device (queried once; 256 on MI355X)
inline int num_cus() {
    static const int n = [] {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0)
            v = 256;
        return v;
    }();
    return n;
}

}  // namespace qcu
