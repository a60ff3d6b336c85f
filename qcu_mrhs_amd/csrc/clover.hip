// clover.hip -- build, block inverse and application of the clover term.
//
// Build and inverse run once per (gauge, kappa, c_sw): one thread per site
// (build) or per block (inverse), FP64, set-up cost only.  The application
// runs twice per Schur operator in the solver and is HBM-bound: a workgroup
// stages the clover blocks of its sites into LDS with contiguous loads
// (consecutive sites' blocks are adjacent in memory), then every lane owns
// one (site, chirality, RHS) triple: it reads the 2Nc spinor entries of its
// chirality (RHS-innermost, so the lanes of a wavefront read consecutive
// words), multiplies by the block read from LDS (the lanes of one block read
// the same LDS word: a broadcast) and writes 2Nc entries.  A lane reads all
// of its input before its first store and nothing else reads those entries,
// so out may alias in (A^{-1} applied in place) or y.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "clover.h"

namespace qcu {
namespace {

// sigma_{mu nu} = (i/2)[gamma_mu, gamma_nu] restricted to the two chirality
// blocks: [plane (01,02,03,12,13,23)][chirality][2][2].
struct SigmaTab {
    Cx<double> s[6][2][2][2];
};

// The chiral gamma basis of the hopping kernels (the (1 + gamma_mu) forms
// spelled out in the reference's include/kernel/qcu_gamma.cuh:10-18, 70-76,
// 126-132, 185-191), rows spin 0..3.
SigmaTab make_sigma() {
    using C = Cx<double>;
    const C o{0, 0}, p1{1, 0}, m1{-1, 0}, pi{0, 1}, mi{0, -1};
    const C gam[4][4][4] = {
        {{o, o, o, pi}, {o, o, pi, o}, {o, mi, o, o}, {mi, o, o, o}},
        {{o, o, o, m1}, {o, o, p1, o}, {o, p1, o, o}, {m1, o, o, o}},
        {{o, o, pi, o}, {o, o, o, mi}, {mi, o, o, o}, {o, pi, o, o}},
        {{o, o, p1, o}, {o, o, o, p1}, {p1, o, o, o}, {o, p1, o, o}},
    };
    SigmaTab t{};
    int q = 0;
    for (int mu = 0; mu < 4; ++mu)
        for (int nu = mu + 1; nu < 4; ++nu, ++q)
            for (int ch = 0; ch < 2; ++ch)
                for (int a = 0; a < 2; ++a)
                    for (int b = 0; b < 2; ++b) {
                        const int sa = 2 * ch + a, sb = 2 * ch + b;
                        C c{0, 0};
                        for (int k = 0; k < 4; ++k)
                            c = c + gam[mu][sa][k] * gam[nu][k][sb] - gam[nu][sa][k] * gam[mu][k][sb];
                        t.s[q][ch][a][b] = C{-0.5 * c.im, 0.5 * c.re};  // (i/2) c
                    }
    return t;
}

template <int NC>
struct Mat {
    Cx<double> m[NC][NC];
};

// U_mu at coordinates c (wrapped), FP64 even/odd gauge [4][2][Vh][NC][NC]
template <int NC>
__device__ __forceinline__ const Cx<double> *link_ptr(const Cx<double> *gauge, const Geom &g, int mu,
                                                    const int c[4]) {
    int w[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) w[k] = (c[k] % g.L[k] + g.L[k]) % g.L[k];
    const int p = (w[0] + w[1] + w[2] + w[3]) & 1;
    const long idx = ((long(w[3]) * g.L[2] + w[2]) * g.L[1] + w[1]) * g.Lxh + w[0] / 2;
    return gauge + ((long(mu) * 2 + p) * g.Vh + idx) * NC * NC;
}

// w = w * U^{dag ? +: 1}
template <int NC>
__device__ __forceinline__ void mul_right(Mat<NC> &w, const Cx<double> *u, bool dag) {
    Mat<NC> t;
    for (int i = 0; i < NC; ++i)
        for (int j = 0; j < NC; ++j) {
            Cx<double> s{0.0, 0.0};
            for (int k = 0; k < NC; ++k) {
                if (dag) cfma(s, w.m[i][k], conj(u[j * NC + k]));
                else cfma(s, w.m[i][k], u[k * NC + j]);
            }
            t.m[i][j] = s;
        }
    w = t;
}

// The four leaves of the (mu, nu) clover at x, each a path of four links:
// {direction (0 = mu, 1 = nu), offset along mu, offset along nu, daggered}.
__constant__ int kLeaf[4][4][4] = {
    {{0, 0, 0, 0}, {1, 1, 0, 0}, {0, 0, 1, 1}, {1, 0, 0, 1}},        // U_mu(x) U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
    {{1, 0, 0, 0}, {0, -1, 1, 1}, {1, -1, 0, 1}, {0, -1, 0, 0}},     // U_nu(x) U_mu(x-mu+nu)^+ U_nu(x-mu)^+ U_mu(x-mu)
    {{0, -1, 0, 1}, {1, -1, -1, 1}, {0, -1, -1, 0}, {1, 0, -1, 0}},  // U_mu(x-mu)^+ U_nu(x-mu-nu)^+ U_mu(x-mu-nu) U_nu(x-nu)
    {{1, 0, -1, 1}, {0, 0, -1, 0}, {1, 1, -1, 0}, {0, 0, 0, 1}},     // U_nu(x-nu)^+ U_mu(x-nu) U_nu(x+mu-nu) U_mu(x)^+
};

template <int NC>
__global__ __launch_bounds__(64) void k_clover_build(Cx<double> *__restrict__ clov, const Cx<double> *__restrict__ gauge,
                                                    Geom g, double coef, SigmaTab sig) {
    constexpr int n = 2 * NC;
    const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= 2L * g.Vh) return;
    const int parity = int(e / g.Vh), site = int(e % g.Vh);
    int x[4];
    {
        const int xh = site % g.Lxh;
        int rest = site / g.Lxh;
        x[1] = rest % g.L[1];
        rest /= g.L[1];
        x[2] = rest % g.L[2];
        x[3] = rest / g.L[2];
        x[0] = 2 * xh + ((x[1] + x[2] + x[3] + parity) & 1);
    }
    Cx<double> *blk = clov + e * 2 * n * n;
    for (int k = 0; k < 2 * n * n; ++k) {
        const int c = k % n, r = (k / n) % n;
        blk[k] = Cx<double>{r == c ? 1.0 : 0.0, 0.0};
    }
    int q = 0;
#pragma unroll 1
    for (int mu = 0; mu < 4; ++mu)
#pragma unroll 1
        for (int nu = mu + 1; nu < 4; ++nu, ++q) {
            Mat<NC> Q;
            for (int i = 0; i < NC; ++i)
                for (int j = 0; j < NC; ++j) Q.m[i][j] = Cx<double>{0.0, 0.0};
#pragma unroll 1
            for (int l = 0; l < 4; ++l) {
                Mat<NC> w;
                for (int i = 0; i < NC; ++i)
                    for (int j = 0; j < NC; ++j) w.m[i][j] = Cx<double>{i == j ? 1.0 : 0.0, 0.0};
#pragma unroll 1
                for (int h = 0; h < 4; ++h) {
                    const int *st = kLeaf[l][h];
                    int c[4] = {x[0], x[1], x[2], x[3]};
                    c[mu] += st[1];
                    c[nu] += st[2];
                    mul_right<NC>(w, link_ptr<NC>(gauge, g, st[0] ? nu : mu, c), st[3] != 0);
                }
                for (int i = 0; i < NC; ++i)
                    for (int j = 0; j < NC; ++j) Q.m[i][j] = Q.m[i][j] + w.m[i][j];
            }
            // F = ((Q - Q^+) - tr(Q - Q^+)/NC) / (8i); tr(Q - Q^+) = 2i Im tr Q
            double imtr = 0.0;
            for (int a = 0; a < NC; ++a) imtr += Q.m[a][a].im;
            for (int a = 0; a < NC; ++a)
                for (int b = a; b < NC; ++b) {
                    Cx<double> d = Q.m[a][b] - conj(Q.m[b][a]);  // anti-Hermitian part times 2
                    if (a == b) d.im -= 2.0 * imtr / NC;
                    // F_ab = d / (8i) = (d.im, -d.re) / 8
                    const Cx<double> f{d.im * 0.125, -d.re * 0.125};
                    for (int ch = 0; ch < 2; ++ch)
                        for (int s = 0; s < 2; ++s)
                            for (int t = 0; t < 2; ++t) {
                                const Cx<double> sv = sig.s[q][ch][s][t];
                                if (sv.re == 0.0 && sv.im == 0.0) continue;
                                Cx<double> *B = blk + ch * n * n;
                                const Cx<double> v = (-coef) * (sv * f);
                                B[(s * NC + a) * n + t * NC + b] = B[(s * NC + a) * n + t * NC + b] + v;
                                if (a != b) {  // F_ba = conj(F_ab)
                                    const Cx<double> v2 = (-coef) * (sv * conj(f));
                                    B[(s * NC + b) * n + t * NC + a] = B[(s * NC + b) * n + t * NC + a] + v2;
                                }
                            }
                }
        }
}

// In-place Gauss-Jordan inversion with partial pivoting (row interchanges
// recorded, undone as column interchanges at the end).
template <int N>
__global__ __launch_bounds__(64) void k_clover_invert(Cx<double> *inv, const Cx<double> *a, long nblocks) {
    const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= nblocks) return;
    Cx<double> *m = inv + e * N * N;
    if (inv != a)
        for (int k = 0; k < N * N; ++k) m[k] = a[e * N * N + k];
    int perm[N];
#pragma unroll 1
    for (int k = 0; k < N; ++k) {
        int p = k;
        double best = -1.0;
        for (int i = k; i < N; ++i) {
            const Cx<double> v = m[i * N + k];
            const double mag = v.re * v.re + v.im * v.im;
            if (mag > best) {
                best = mag;
                p = i;
            }
        }
        perm[k] = p;
        if (p != k)
            for (int j = 0; j < N; ++j) {
                const Cx<double> t = m[k * N + j];
                m[k * N + j] = m[p * N + j];
                m[p * N + j] = t;
            }
        const Cx<double> piv = cdiv(Cx<double>{1.0, 0.0}, m[k * N + k]);
        m[k * N + k] = Cx<double>{1.0, 0.0};
        for (int j = 0; j < N; ++j) m[k * N + j] = m[k * N + j] * piv;
        for (int i = 0; i < N; ++i) {
            if (i == k) continue;
            const Cx<double> f = m[i * N + k];
            m[i * N + k] = Cx<double>{0.0, 0.0};
            for (int j = 0; j < N; ++j) m[i * N + j] = m[i * N + j] - f * m[k * N + j];
        }
    }
#pragma unroll 1
    for (int k = N - 1; k >= 0; --k) {
        const int p = perm[k];
        if (p == k) continue;
        for (int i = 0; i < N; ++i) {
            const Cx<double> t = m[i * N + k];
            m[i * N + k] = m[i * N + p];
            m[i * N + p] = t;
        }
    }
}

extern __shared__ __align__(16) unsigned char qcu_clover_lds[];

// out = A in + coef * y; S sites per workgroup, lanes (site, chirality, rhs)
template <int NC, typename CS, typename TC>
__global__ __launch_bounds__(1024) void k_clover_apply(CS *out, const CS *in, const Cx<TC> *__restrict__ blocks,
                                                      const CS *y, TC coef, int Vh, int R, int S) {
    constexpr int n = 2 * NC;
    auto *lds = reinterpret_cast<Cx<TC> *>(qcu_clover_lds);
    const int site0 = blockIdx.x * S;
    const int ns = min(S, Vh - site0);
    const long base = long(site0) * 2 * n * n;
    const int total = ns * 2 * n * n;
    for (int k = threadIdx.x; k < total; k += blockDim.x) lds[k] = blocks[base + k];
    __syncthreads();
    const int tid = threadIdx.x;
    const int sl = tid / (2 * R);
    if (sl >= ns) return;
    const int ch = (tid / R) & 1, r = tid % R;
    const long site = site0 + sl;
    const long e0 = ((site * 4 + 2 * ch) * NC) * R + r;  // entry (spin 2ch, colour 0, rhs r)
    Cx<TC> v[n];
#pragma unroll
    for (int k = 0; k < n; ++k) v[k] = load_cx<TC>(in + e0 + long(k) * R);
    const Cx<TC> *L = lds + (sl * 2 + ch) * n * n;
#pragma unroll
    for (int row = 0; row < n; ++row) {
        Cx<TC> acc{TC(0), TC(0)};
#pragma unroll
        for (int col = 0; col < n; ++col) cfma(acc, L[row * n + col], v[col]);
        const long e = e0 + long(row) * R;
        if (y) {
            const Cx<TC> yv = load_cx<TC>(y + e);
            acc.re = fma(coef, yv.re, acc.re);
            acc.im = fma(coef, yv.im, acc.im);
        }
        store_cx(out + e, acc);
    }
}

template <int NC, typename CS, typename TC>
hipError_t apply_t(void *out, const void *in, const void *blocks, const void *y, double coef, int Vh, int R,
                   hipStream_t s) {
    constexpr int n = 2 * NC;
    const int per_site = 2 * n * n * int(sizeof(Cx<TC>));
    int S = 2 * R >= 256 ? 1 : 256 / (2 * R);
    S = std::min(S, std::max(1, 65536 / per_site));
    const int threads = (S * 2 * R + 63) / 64 * 64;
    if (threads > 1024) return hipErrorInvalidConfiguration;
    const int blocks_n = div_up(Vh, S);
    hipLaunchKernelGGL((k_clover_apply<NC, CS, TC>), dim3(blocks_n), dim3(threads), size_t(S) * per_site, s,
                       static_cast<CS *>(out), static_cast<const CS *>(in), static_cast<const Cx<TC> *>(blocks),
                       static_cast<const CS *>(y), TC(coef), Vh, R, S);
    return hipGetLastError();
}

template <int NC, typename TC>
hipError_t apply_prec(void *out, const void *in, const void *blocks, const void *y, double coef, int Vh, int R,
                      int prec, hipStream_t s) {
    switch (prec) {
        case kDouble: return apply_t<NC, Cx<double>, TC>(out, in, blocks, y, coef, Vh, R, s);
        case kSingle: return apply_t<NC, Cx<float>, TC>(out, in, blocks, y, coef, Vh, R, s);
        case kHalf: return apply_t<NC, Half2Store, TC>(out, in, blocks, y, coef, Vh, R, s);
        case kBF16: return apply_t<NC, Bf162Store, TC>(out, in, blocks, y, coef, Vh, R, s);
        default: return hipErrorInvalidValue;
    }
}

template <int NC>
hipError_t apply_nc(void *out, const void *in, const void *blocks, const void *y, double coef, int Vh, int R,
                    int prec, int clover_prec, hipStream_t s) {
    if (clover_prec == kDouble) return apply_prec<NC, double>(out, in, blocks, y, coef, Vh, R, prec, s);
    if (clover_prec == kSingle) return apply_prec<NC, float>(out, in, blocks, y, coef, Vh, R, prec, s);
    return hipErrorInvalidValue;
}

}  // namespace

bool clover_supported(int Nc) { return Nc == 2 || Nc == 3 || Nc == 4 || Nc == 8; }

hipError_t launch_clover_build(void *clover, const void *gauge, const Geom &g, int Nc, double csw, double kappa,
                               hipStream_t s) {
    static const SigmaTab sig = make_sigma();
    const double coef = kappa * csw;
    const int nb = div_up(2L * g.Vh, 64);
    auto *c = static_cast<Cx<double> *>(clover);
    const auto *u = static_cast<const Cx<double> *>(gauge);
    switch (Nc) {
        case 2: hipLaunchKernelGGL(k_clover_build<2>, dim3(nb), dim3(64), 0, s, c, u, g, coef, sig); break;
        case 3: hipLaunchKernelGGL(k_clover_build<3>, dim3(nb), dim3(64), 0, s, c, u, g, coef, sig); break;
        case 4: hipLaunchKernelGGL(k_clover_build<4>, dim3(nb), dim3(64), 0, s, c, u, g, coef, sig); break;
        case 8: hipLaunchKernelGGL(k_clover_build<8>, dim3(nb), dim3(64), 0, s, c, u, g, coef, sig); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_clover_invert(void *inv, const void *clover, long nblocks, int n, hipStream_t s) {
    const int nb = div_up(nblocks, 64);
    auto *o = static_cast<Cx<double> *>(inv);
    const auto *a = static_cast<const Cx<double> *>(clover);
    switch (n) {
        case 4: hipLaunchKernelGGL(k_clover_invert<4>, dim3(nb), dim3(64), 0, s, o, a, nblocks); break;
        case 6: hipLaunchKernelGGL(k_clover_invert<6>, dim3(nb), dim3(64), 0, s, o, a, nblocks); break;
        case 8: hipLaunchKernelGGL(k_clover_invert<8>, dim3(nb), dim3(64), 0, s, o, a, nblocks); break;
        case 16: hipLaunchKernelGGL(k_clover_invert<16>, dim3(nb), dim3(64), 0, s, o, a, nblocks); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_clover_apply(void *out, const void *in, const void *blocks, const void *y, double coef,
                               int Vh, int Nc, int R, int prec, int clover_prec, hipStream_t s) {
    if (R <= 0 || R > 512 || Vh <= 0) return hipErrorInvalidValue;
    switch (Nc) {
        case 2: return apply_nc<2>(out, in, blocks, y, coef, Vh, R, prec, clover_prec, s);
        case 3: return apply_nc<3>(out, in, blocks, y, coef, Vh, R, prec, clover_prec, s);
        case 4: return apply_nc<4>(out, in, blocks, y, coef, Vh, R, prec, clover_prec, s);
        case 8: return apply_nc<8>(out, in, blocks, y, coef, Vh, R, prec, clover_prec, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace qcu
