/*
 * qcu_oracle.c -- CPU FP64 restatement of the reference's multi-RHS SU(N)
 * Wilson dslash, full and even/odd Schur operators, BiCGStab recurrence,
 * per-RHS strided BLAS and the gather/scatter layouts.
 *
 * THIS FILE IS TEST INFRASTRUCTURE.  It is the correctness oracle: only the
 * tests, the smoke check and the CPU-baseline leg of bench.py load it, and
 * only as the checker / the timed CPU baseline.  The product library
 * (qcu_mrhs_amd/libqcu.so) never links or calls it.
 *
 * Conventions (all restated from the reference, cited per function):
 *   - lattice L[4] = {Lx, Ly, Lz, Lt}; Vh = Lx*Ly*Lz*Lt/2
 *   - checkerboard index  i = ((t*Lz + z)*Ly + y)*(Lx/2) + x/2,
 *     site parity p = (x+y+z+t) & 1, full x = 2*xh + ((y+z+t+p) & 1)
 *     (include/point/qcu_point.cuh:16-61,89-91)
 *   - MRHS half vector  [Vh][spin 4][color Nc][rhs R]  complex
 *     (include/kernel/copy_vector/copy_color_spinor.cuh:27-38)
 *   - gauge  [mu 4][parity 2][Vh][row Nc][col Nc]  complex
 *     (include/point/qcu_point.cuh:67-75)
 *   - "dslash" output parity p:  out(x) = sum_mu (1-g_mu) U_mu(x) psi(x+mu)
 *                                           + (1+g_mu) U_mu(x-mu)^+ psi(x-mu)
 *     (include/kernel/sun_mrhs_wilson_dslash_simt.cuh:97-181,200-202)
 *
 * Parity status of each function (what pins it) is listed in the header of
 * oracle/oracle.py and asserted by tests/test_oracle_*.py.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double _Complex cplx;

#define EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------ */
/* spin projection tables                                              */
/* ------------------------------------------------------------------ */
/* For direction mu and projected row r in {0,1}:
 *   mat2[mu][r]   = spin index combined with r
 *                   (qcu_gamma.cuh:47-51 Gamma1, :101-106 Gamma2,
 *                    :159-163 Gamma3, :208-212 Gamma4)
 *   proj1p[mu][r] = projection scale of (1+gamma)   (qcu_gamma.cuh:61-67,
 *                   :115-123, :172-180, :222-228)
 *   rec1p[mu][r]  = reconstruction scale of (1+gamma) (qcu_gamma.cuh:52-59,
 *                   :107-114, :164-171, :213-220)
 * The executed scale is multiplied by +1 for the backward hop and -1 for
 * the forward hop (qcu_gamma.cuh:278-324), and negated again when the
 * dagger flag is set (sun_mrhs_wilson_dslash_simt.cuh:147,176).          */
static const int mat2[4][2] = {{3, 2}, {3, 2}, {2, 3}, {2, 3}};
static cplx proj1p(int mu, int r) {
    switch (mu) {
        case 0: return I;                       /* ( i,  i) */
        case 1: return r == 0 ? -1.0 : 1.0;     /* (-1,  1) */
        case 2: return r == 0 ? I : -I;         /* ( i, -i) */
        default: return 1.0;                    /* ( 1,  1) */
    }
}
static cplx rec1p(int mu, int r) {
    switch (mu) {
        case 0: return -I;                      /* (-i, -i) */
        case 1: return r == 0 ? -1.0 : 1.0;     /* (-1,  1) */
        case 2: return r == 0 ? -I : I;         /* (-i,  i) */
        default: return 1.0;                    /* ( 1,  1) */
    }
}

/* ------------------------------------------------------------------ */
/* site indexing                                                      */
/* ------------------------------------------------------------------ */
static void cb_coords(long i, int p, const int L[4], int c[4]) {
    const int Lxh = L[0] / 2;
    int xh = (int)(i % Lxh);
    long rest = i / Lxh;
    c[1] = (int)(rest % L[1]); rest /= L[1];
    c[2] = (int)(rest % L[2]); rest /= L[2];
    c[3] = (int)rest;
    c[0] = 2 * xh + ((c[1] + c[2] + c[3] + p) & 1);
}
static long cb_index(const int c[4], const int L[4]) {
    return (((long)c[3] * L[2] + c[2]) * L[1] + c[1]) * (L[0] / 2) + c[0] / 2;
}
/* index of the neighbour x + dir*mu (dir = +1 / -1), periodic in every
 * direction (qcu_point.cuh:16-61) */
static long cb_neighbour(const int c[4], const int L[4], int mu, int dir) {
    int d[4] = {c[0], c[1], c[2], c[3]};
    d[mu] = (d[mu] + dir + L[mu]) % L[mu];
    return cb_index(d, L);
}

EXPORT long oracle_half_volume(const int L[4]) {
    return (long)L[0] * L[1] * L[2] * L[3] / 2;
}

/* ------------------------------------------------------------------ */
/* dslash                                                             */
/* ------------------------------------------------------------------ */
/* One-parity multi-RHS hopping term; `parity` is the parity of the OUTPUT
 * site, the input lives on 1-parity (sun_mrhs_wilson_dslash_simt.cuh:200-228).
 * Per direction the neighbour spinor is projected to two half-spinor rows
 * h_r = psi_r + s*psi_m(r), multiplied by U_mu(x) (forward) or
 * U_mu(x-mu)^+ (backward, link taken at the neighbour's parity), and
 * reconstructed: out_r += U h_r, out_m(r) += c * U h_r
 * (sun_mrhs_wilson_dslash_simt.cuh:97-181; complex_matmul.cuh:503-600). */
EXPORT void oracle_dslash(cplx *out, const cplx *in, const cplx *gauge,
                          const int L[4], int Nc, int R, int parity, int dagger) {
    const long Vh = oracle_half_volume(L);
    const long site_len = 4L * Nc * R;
    const long link_len = (long)Nc * Nc;
#pragma omp parallel
    {
        cplx *h = (cplx *)malloc(sizeof(cplx) * Nc * R);
        cplx *uh = (cplx *)malloc(sizeof(cplx) * Nc * R);
#pragma omp for schedule(static)
        for (long i = 0; i < Vh; ++i) {
            int c[4];
            cb_coords(i, parity, L, c);
            cplx *o = out + i * site_len;
            for (long e = 0; e < site_len; ++e) o[e] = 0;
            for (int mu = 0; mu < 4; ++mu) {
                for (int fb = 0; fb < 2; ++fb) {       /* 0 = forward, 1 = backward */
                    const int dir = fb == 0 ? +1 : -1;
                    const double sign = (fb == 1 ? 1.0 : -1.0) * (dagger ? -1.0 : 1.0);
                    const long nb = cb_neighbour(c, L, mu, dir);
                    const cplx *psi = in + nb * site_len;
                    const cplx *U = fb == 0
                        ? gauge + ((long)(mu * 2 + parity) * Vh + i) * link_len
                        : gauge + ((long)(mu * 2 + (1 - parity)) * Vh + nb) * link_len;
                    for (int r = 0; r < 2; ++r) {
                        const int m = mat2[mu][r];
                        const cplx ps = sign * proj1p(mu, r);
                        const cplx rs = sign * rec1p(mu, r);
                        for (int a = 0; a < Nc; ++a)
                            for (int k = 0; k < R; ++k)
                                h[a * R + k] = psi[((long)r * Nc + a) * R + k]
                                             + ps * psi[((long)m * Nc + a) * R + k];
                        for (int a = 0; a < Nc; ++a) {
                            for (int k = 0; k < R; ++k) uh[a * R + k] = 0;
                            for (int b = 0; b < Nc; ++b) {
                                /* forward: U[a][b]; backward: (U^+)[a][b] = conj(U[b][a]) */
                                const cplx u = fb == 0 ? U[a * Nc + b] : conj(U[b * Nc + a]);
                                for (int k = 0; k < R; ++k) uh[a * R + k] += u * h[b * R + k];
                            }
                        }
                        for (int a = 0; a < Nc; ++a)
                            for (int k = 0; k < R; ++k) {
                                o[((long)r * Nc + a) * R + k] += uh[a * R + k];
                                o[((long)m * Nc + a) * R + k] += rs * uh[a * R + k];
                            }
                    }
                }
            }
        }
        free(h);
        free(uh);
    }
}

/* ------------------------------------------------------------------ */
/* per-RHS strided BLAS (src/qcu_blas/qcu_complex_*.cu, kernel/element_wise) */
/* n = number of complex entries per RHS, R = stride = number of RHS.   */
/* ------------------------------------------------------------------ */
/* res = x - a_r y   (include/kernel/element_wise/c_xsay.cuh:14-42) */
EXPORT void oracle_cxsay(cplx *res, const cplx *x, const cplx *a, const cplx *y, long n, int R) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i)
        for (int r = 0; r < R; ++r) res[i * R + r] = x[i * R + r] - a[r] * y[i * R + r];
}
/* res = x + a_r y   (include/kernel/element_wise/c_xpay.cuh:14-42) */
EXPORT void oracle_cxpay(cplx *res, const cplx *x, const cplx *a, const cplx *y, long n, int R) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i)
        for (int r = 0; r < R; ++r) res[i * R + r] = x[i * R + r] + a[r] * y[i * R + r];
}
/* res = a_r x + b_r y   (include/kernel/element_wise/c_axpby.cuh:14-41) */
EXPORT void oracle_caxpby(cplx *res, const cplx *a, const cplx *x, const cplx *b, const cplx *y,
                          long n, int R) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i)
        for (int r = 0; r < R; ++r) res[i * R + r] = a[r] * x[i * R + r] + b[r] * y[i * R + r];
}
/* res = a_r x + b_r y + c_r z   (include/kernel/element_wise/c_axpbypcz.cuh:14-44) */
EXPORT void oracle_caxpbypcz(cplx *res, const cplx *a, const cplx *x, const cplx *b, const cplx *y,
                             const cplx *c, const cplx *z, long n, int R) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i)
        for (int r = 0; r < R; ++r)
            res[i * R + r] = a[r] * x[i * R + r] + b[r] * y[i * R + r] + c[r] * z[i * R + r];
}

/* Partial sums over fixed blocks of ORACLE_BLOCK entries, combined in block
 * order, so the result does not depend on the OpenMP thread count. */
#define ORACLE_BLOCK 4096L

/* <a,b>_r = sum_i conj(a[i*R+r]) * b[i*R+r]
 * (include/kernel/reduction/reduction.cuh:147-158; cublasZdotc semantics at
 *  src/qcu_blas/qcu_complex_dotc.cu:62-66) */
EXPORT void oracle_dotc(cplx *res, const cplx *a, const cplx *b, long n, int R) {
    const long nb = (n + ORACLE_BLOCK - 1) / ORACLE_BLOCK;
    cplx *part = (cplx *)calloc((size_t)(nb > 0 ? nb : 1) * R, sizeof(cplx));
#pragma omp parallel for schedule(static)
    for (long blk = 0; blk < nb; ++blk) {
        const long lo = blk * ORACLE_BLOCK, hi = lo + ORACLE_BLOCK < n ? lo + ORACLE_BLOCK : n;
        for (long i = lo; i < hi; ++i)
            for (int r = 0; r < R; ++r) part[blk * R + r] += conj(a[i * R + r]) * b[i * R + r];
    }
    for (int r = 0; r < R; ++r) res[r] = 0;
    for (long blk = 0; blk < nb; ++blk)
        for (int r = 0; r < R; ++r) res[r] += part[blk * R + r];
    free(part);
}
/* ||a||_r = sqrt(sum_i |a[i*R+r]|^2)
 * (reduction.cuh:85-126 followed by SqrtOp, src/qcu_blas/qcu_complex_norm.cu:44-54) */
EXPORT void oracle_norm(double *res, const cplx *a, long n, int R) {
    const long nb = (n + ORACLE_BLOCK - 1) / ORACLE_BLOCK;
    double *part = (double *)calloc((size_t)(nb > 0 ? nb : 1) * R, sizeof(double));
#pragma omp parallel for schedule(static)
    for (long blk = 0; blk < nb; ++blk) {
        const long lo = blk * ORACLE_BLOCK, hi = lo + ORACLE_BLOCK < n ? lo + ORACLE_BLOCK : n;
        for (long i = lo; i < hi; ++i)
            for (int r = 0; r < R; ++r) {
                const cplx v = a[i * R + r];
                part[blk * R + r] += creal(v) * creal(v) + cimag(v) * cimag(v);
            }
    }
    for (int r = 0; r < R; ++r) res[r] = 0;
    for (long blk = 0; blk < nb; ++blk)
        for (int r = 0; r < R; ++r) res[r] += part[blk * R + r];
    for (int r = 0; r < R; ++r) res[r] = sqrt(res[r]);
    free(part);
}

/* ------------------------------------------------------------------ */
/* layouts                                                            */
/* ------------------------------------------------------------------ */
/* Gather R per-RHS vectors [site][spin][color] into [site][spin][color][rhs]
 * over `nsites` sites (copy_color_spinor.cuh:16-39,70-84). */
EXPORT void oracle_gather(cplx *dst, const cplx *const *src, long nsites, int Nc, int R) {
    const long sl = 4L * Nc;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < nsites; ++i)
        for (int r = 0; r < R; ++r)
            for (long e = 0; e < sl; ++e) dst[(i * sl + e) * R + r] = src[r][i * sl + e];
}
/* Inverse of oracle_gather (copy_color_spinor.cuh:41-68,86-101). */
EXPORT void oracle_scatter(cplx *const *dst, const cplx *src, long nsites, int Nc, int R) {
    const long sl = 4L * Nc;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < nsites; ++i)
        for (int r = 0; r < R; ++r)
            for (long e = 0; e < sl; ++e) dst[r][i * sl + e] = src[(i * sl + e) * R + r];
}

/* Lexicographic gauge [mu][t][z][y][x][Nc][Nc] -> even/odd
 * [mu][parity][t][z][y][x/2][Nc][Nc].  The reference delegates this to the
 * absent qcu_io submodule (src/interface.cu:76-105); the mapping restated
 * here follows the even/odd site convention of qcu_point.cuh:89-91 and the
 * reader dims {Nd,Lt,Lz,Ly,Lx,Nc,2Nc} (src/qcu_interface.cu:387-388).
 * reverse != 0 performs the inverse mapping. */
EXPORT void oracle_gauge_eo(cplx *eo, cplx *lex, const int L[4], int Nc, int reverse) {
    const long V = (long)L[0] * L[1] * L[2] * L[3], Vh = V / 2, ll = (long)Nc * Nc;
    for (int mu = 0; mu < 4; ++mu)
        for (long s = 0; s < V; ++s) {
            int c[4];
            long rest = s;
            c[0] = (int)(rest % L[0]); rest /= L[0];
            c[1] = (int)(rest % L[1]); rest /= L[1];
            c[2] = (int)(rest % L[2]); rest /= L[2];
            c[3] = (int)rest;
            const int p = (c[0] + c[1] + c[2] + c[3]) & 1;
            cplx *a = eo + ((long)(mu * 2 + p) * Vh + cb_index(c, L)) * ll;
            cplx *b = lex + ((long)mu * V + s) * ll;
            if (reverse) memcpy(b, a, sizeof(cplx) * ll);
            else memcpy(a, b, sizeof(cplx) * ll);
        }
}

/* ------------------------------------------------------------------ */
/* operators built from the dslash                                    */
/* ------------------------------------------------------------------ */
/* Full-lattice M = 1 - kappa*D on a full MRHS vector [parity][Vh][4][Nc][R]
 * (src/qcu_interface.cu:174-255: out = in - kappa * dslash(in), both parities). */
EXPORT void oracle_mat(cplx *out, const cplx *in, const cplx *gauge, const int L[4], int Nc, int R,
                       double kappa, int dagger) {
    const long n = oracle_half_volume(L) * 4L * Nc * R;
    for (int p = 0; p < 2; ++p) {
        oracle_dslash(out + p * n, in + (1 - p) * n, gauge, L, Nc, R, p, dagger);
#pragma omp parallel for schedule(static)
        for (long e = 0; e < n; ++e) out[p * n + e] = in[p * n + e] - kappa * out[p * n + e];
    }
}

/* Odd Schur operator  out = in - kappa^2 D_oe D_eo in
 * (src/solver/bicgstab_solve.cu:26-70). tmp: half-vector scratch. */
EXPORT void oracle_schur(cplx *out, const cplx *in, cplx *tmp, const cplx *gauge, const int L[4],
                         int Nc, int R, double kappa, int dagger) {
    const long n = oracle_half_volume(L) * 4L * Nc * R;
    oracle_dslash(tmp, in, gauge, L, Nc, R, 0, dagger);
    oracle_dslash(out, tmp, gauge, L, Nc, R, 1, dagger);
    const double k2 = kappa * kappa;
#pragma omp parallel for schedule(static)
    for (long e = 0; e < n; ++e) out[e] = in[e] - k2 * out[e];
}

/* ------------------------------------------------------------------ */
/* BiCGStab                                                            */
/* ------------------------------------------------------------------ */
/* Even/odd preconditioned multi-RHS BiCGStab, restating
 *   reCalculate_b_even  (src/solver/bicgstab.cu:141-189):  b' = b_o + kappa D_oe b_e
 *   solve_odd_policy1   (src/solver/bicgstab_solve.cu:102-389): recurrence and
 *                       all-RHS stopping rule ||r_r|| / ||b'_r|| <= tol (:77-88)
 *   solve_even          (:688-740):  x_e = b_e + kappa D_eo x_o
 * with x0 = 0 (the intended initial guess, bicgstab_solve.cu:243-247).
 * x_full, b_full: [parity][Vh][4][Nc][R].  Returns 1 if converged.
 * *iters = number of completed iterations (the reference prints the loop
 * index, i.e. this value minus one).  res_out[r] = final ||r_r||/||b'_r||. */
EXPORT int oracle_bicgstab(cplx *x_full, const cplx *b_full, const cplx *gauge, const int L[4],
                           int Nc, int R, double kappa, int max_iter, double tol, int *iters,
                           double *res_out) {
    const long Vh = oracle_half_volume(L);
    const long n = Vh * 4L * Nc;          /* entries per RHS */
    const long nn = n * R;                /* entries per half vector */
    const cplx *b_e = b_full, *b_o = b_full + nn;
    cplx *x_e = x_full, *x_o = x_full + nn;
    cplx *bp = malloc(sizeof(cplx) * nn), *r0 = malloc(sizeof(cplx) * nn);
    cplx *r = malloc(sizeof(cplx) * nn), *p = malloc(sizeof(cplx) * nn);
    cplx *v = malloc(sizeof(cplx) * nn), *s = malloc(sizeof(cplx) * nn);
    cplx *t = malloc(sizeof(cplx) * nn), *tmp = malloc(sizeof(cplx) * nn);
    cplx *kap = malloc(sizeof(cplx) * R), *one = malloc(sizeof(cplx) * R);
    cplx *rho = malloc(sizeof(cplx) * R), *r0v = malloc(sizeof(cplx) * R);
    cplx *alpha = malloc(sizeof(cplx) * R), *omega = malloc(sizeof(cplx) * R);
    cplx *ts = malloc(sizeof(cplx) * R), *tt = malloc(sizeof(cplx) * R);
    cplx *rho_new = malloc(sizeof(cplx) * R), *beta = malloc(sizeof(cplx) * R);
    double *nb = malloc(sizeof(double) * R), *nr = malloc(sizeof(double) * R);
    for (int k = 0; k < R; ++k) { kap[k] = kappa; one[k] = 1.0; }

    /* b' = b_o + kappa * D_oe b_e */
    oracle_dslash(tmp, b_e, gauge, L, Nc, R, 1, 0);
    oracle_cxpay(bp, b_o, kap, tmp, n, R);
    oracle_norm(nb, bp, n, R);
    for (int k = 0; k < R; ++k) nr[k] = nb[k];

    memset(x_o, 0, sizeof(cplx) * nn);
    memcpy(r, bp, sizeof(cplx) * nn);
    memcpy(r0, bp, sizeof(cplx) * nn);
    memcpy(p, bp, sizeof(cplx) * nn);

    int converged = 0, it;
    for (it = 0; it < max_iter; ++it) {
        oracle_dotc(rho, r0, r, n, R);                               /* rho = <r0, r>   */
        oracle_schur(v, p, tmp, gauge, L, Nc, R, kappa, 0);          /* v = A p         */
        oracle_dotc(r0v, r0, v, n, R);
        for (int k = 0; k < R; ++k) alpha[k] = rho[k] / r0v[k];      /* alpha           */
        oracle_cxsay(s, r, alpha, v, n, R);                          /* s = r - alpha v */
        oracle_schur(t, s, tmp, gauge, L, Nc, R, kappa, 0);          /* t = A s         */
        oracle_dotc(ts, t, s, n, R);
        oracle_dotc(tt, t, t, n, R);
        for (int k = 0; k < R; ++k) omega[k] = ts[k] / tt[k];        /* omega           */
        oracle_caxpbypcz(x_o, one, x_o, alpha, p, omega, s, n, R);   /* x += a p + w s  */
        oracle_cxsay(r, s, omega, t, n, R);                          /* r = s - w t     */
        oracle_norm(nr, r, n, R);
        int ok = 1;
        for (int k = 0; k < R; ++k) if (!(nr[k] / nb[k] <= tol)) ok = 0;
        if (ok) { converged = 1; ++it; break; }
        oracle_dotc(rho_new, r0, r, n, R);
        for (int k = 0; k < R; ++k) {                                /* beta, 3 steps   */
            beta[k] = alpha[k] * rho_new[k];
            beta[k] = beta[k] / omega[k];
            beta[k] = beta[k] / rho[k];
        }
        oracle_cxsay(p, p, omega, v, n, R);                          /* p = p - w v     */
        oracle_cxpay(p, r, beta, p, n, R);                           /* p = r + beta p  */
    }
    if (iters) *iters = it;
    if (res_out) for (int k = 0; k < R; ++k) res_out[k] = nr[k] / nb[k];

    /* x_e = b_e + kappa D_eo x_o */
    oracle_dslash(tmp, x_o, gauge, L, Nc, R, 0, 0);
    oracle_cxpay(x_e, b_e, kap, tmp, n, R);

    free(bp); free(r0); free(r); free(p); free(v); free(s); free(t); free(tmp);
    free(kap); free(one); free(rho); free(r0v); free(alpha); free(omega); free(ts); free(tt);
    free(rho_new); free(beta); free(nb); free(nr);
    return converged;
}

EXPORT int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
EXPORT void oracle_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------ */
/* clover term                                                         */
/* ------------------------------------------------------------------ */
/* The reference declares kDslashClover (include/qcu_helper_macro.h:7-11) but
 * getDslash rejects it (src/qcu_interface.cu:97-105): there is no reference
 * algorithm to restate, so this is the textbook Sheikholeslami-Wohlert term
 * (Sheikholeslami & Wohlert, Nucl. Phys. B259 (1985) 572; normalisation of
 * Luescher, Sint, Sommer & Weisz, Nucl. Phys. B478 (1996) 365: the term
 * c_sw (i/4) sigma_{mu nu} Fhat_{mu nu} added to D_W, Fhat = (Q - Q^+)/8
 * anti-Hermitian) in the kappa normalisation M = 2 kappa (D_W + m) of the
 * Wilson operator M = 1 - kappa D, whose forward hop is (1 - gamma_mu) U_mu(x):
 *
 *     M_clover = A - kappa D,
 *     A(x) = 1 - kappa c_sw sum_{mu<nu} sigma_{mu nu} F_{mu nu}(x),
 *     sigma_{mu nu} = (i/2) [gamma_mu, gamma_nu]   (the kernels' chiral basis),
 *     F_{mu nu}(x) = ((Q - Q^+) - tr(Q - Q^+)/Nc) / (8 i)   (Hermitian, = Fhat / i),
 *
 * Q_{mu nu}(x) the sum of the four plaquette leaves at x in the (mu, nu) plane,
 * each starting at x and oriented mu -> nu.  The sign: D_W = Dslash - D^2/2
 * with D_mu D_mu = Dslash^2 + (i/2) sigma Fhat leaves -(i/4) sigma Fhat at O(a),
 * which c_sw = 1 cancels; 2 kappa c_sw (i/4) sum_{mu,nu} sigma Fhat
 * = -kappa c_sw sum_{mu<nu} sigma F.  In the chiral basis sigma is block
 * diagonal in the spin pairs {0, 1} and {2, 3}, so A(x) is two Hermitian
 * 2Nc x 2Nc blocks, stored as clover[parity][Vh][chirality 2][2Nc][2Nc] with
 * row/column index (spin in the pair) * Nc + colour.                        */
static const cplx *link_at(const cplx *gauge, const int L[4], int Nc, int mu, const int c0[4]) {
    int c[4];
    for (int k = 0; k < 4; ++k) c[k] = ((c0[k] % L[k]) + L[k]) % L[k];
    const int p = (c[0] + c[1] + c[2] + c[3]) & 1;
    return gauge + ((long)(mu * 2 + p) * oracle_half_volume(L) + cb_index(c, L)) * (long)Nc * Nc;
}

/* out = a b, a^+ b, a b^+ (Nc x Nc) */
static void mm(cplx *out, const cplx *a, int ta, const cplx *b, int tb, int Nc) {
    for (int i = 0; i < Nc; ++i)
        for (int j = 0; j < Nc; ++j) {
            cplx s = 0;
            for (int k = 0; k < Nc; ++k) {
                const cplx x = ta ? conj(a[k * Nc + i]) : a[i * Nc + k];
                const cplx y = tb ? conj(b[j * Nc + k]) : b[k * Nc + j];
                s += x * y;
            }
            out[i * Nc + j] = s;
        }
}

/* product of four links with daggers: w = l0^d0 l1^d1 l2^d2 l3^d3 */
static void path4(cplx *w, const cplx *l[4], const int d[4], int Nc, cplx *t1, cplx *t2) {
    mm(t1, l[0], d[0], l[1], d[1], Nc);
    /* t2 = t1 l2^d2 */
    for (int i = 0; i < Nc; ++i)
        for (int j = 0; j < Nc; ++j) {
            cplx s = 0;
            for (int k = 0; k < Nc; ++k) s += t1[i * Nc + k] * (d[2] ? conj(l[2][j * Nc + k]) : l[2][k * Nc + j]);
            t2[i * Nc + j] = s;
        }
    for (int i = 0; i < Nc; ++i)
        for (int j = 0; j < Nc; ++j) {
            cplx s = 0;
            for (int k = 0; k < Nc; ++k) s += t2[i * Nc + k] * (d[3] ? conj(l[3][j * Nc + k]) : l[3][k * Nc + j]);
            w[i * Nc + j] = s;
        }
}

static const double _Complex GAMMA_CH[4][4][4] = {
    {{0, 0, 0, I}, {0, 0, I, 0}, {0, -I, 0, 0}, {-I, 0, 0, 0}},
    {{0, 0, 0, -1}, {0, 0, 1, 0}, {0, 1, 0, 0}, {-1, 0, 0, 0}},
    {{0, 0, I, 0}, {0, 0, 0, -I}, {-I, 0, 0, 0}, {0, I, 0, 0}},
    {{0, 0, 1, 0}, {0, 0, 0, 1}, {1, 0, 0, 0}, {0, 1, 0, 0}},
};

EXPORT void oracle_clover(cplx *clov, const cplx *gauge, const int L[4], int Nc, double csw, double kappa) {
    const long Vh = oracle_half_volume(L);
    const int n = 2 * Nc;
    /* sigma_{mu nu} = (i/2)[gamma_mu, gamma_nu] */
    cplx sig[6][4][4];
    int pl = 0;
    for (int mu = 0; mu < 4; ++mu)
        for (int nu = mu + 1; nu < 4; ++nu, ++pl)
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b) {
                    cplx s = 0;
                    for (int k = 0; k < 4; ++k)
                        s += GAMMA_CH[mu][a][k] * GAMMA_CH[nu][k][b] - GAMMA_CH[nu][a][k] * GAMMA_CH[mu][k][b];
                    sig[pl][a][b] = 0.5 * I * s;
                }
#pragma omp parallel
    {
        cplx *Q = malloc(sizeof(cplx) * Nc * Nc), *W = malloc(sizeof(cplx) * Nc * Nc);
        cplx *t1 = malloc(sizeof(cplx) * Nc * Nc), *t2 = malloc(sizeof(cplx) * Nc * Nc);
        cplx *F = malloc(sizeof(cplx) * 6 * Nc * Nc);
#pragma omp for collapse(2) schedule(static)
        for (int p = 0; p < 2; ++p)
            for (long i = 0; i < Vh; ++i) {
                int x[4];
                cb_coords(i, p, L, x);
                int pl2 = 0;
                for (int mu = 0; mu < 4; ++mu)
                    for (int nu = mu + 1; nu < 4; ++nu, ++pl2) {
                        int xm[4], xn[4], xmn[4], xpm[4], xpn[4], xpmn[4], xnpm[4];
                        for (int k = 0; k < 4; ++k) {
                            xm[k] = x[k] - (k == mu);
                            xn[k] = x[k] - (k == nu);
                            xmn[k] = x[k] - (k == mu) - (k == nu);
                            xpm[k] = x[k] + (k == mu);
                            xpn[k] = x[k] + (k == nu);
                            xpmn[k] = x[k] - (k == mu) + (k == nu);  /* x - mu + nu */
                            xnpm[k] = x[k] + (k == mu) - (k == nu);  /* x + mu - nu */
                        }
                        for (int e = 0; e < Nc * Nc; ++e) Q[e] = 0;
                        /* U_mu(x) U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+ */
                        const cplx *l1[4] = {link_at(gauge, L, Nc, mu, x), link_at(gauge, L, Nc, nu, xpm),
                                             link_at(gauge, L, Nc, mu, xpn), link_at(gauge, L, Nc, nu, x)};
                        const int d1[4] = {0, 0, 1, 1};
                        path4(W, l1, d1, Nc, t1, t2);
                        for (int e = 0; e < Nc * Nc; ++e) Q[e] += W[e];
                        /* U_nu(x) U_mu(x+nu-mu)^+ U_nu(x-mu)^+ U_mu(x-mu) */
                        const cplx *l2[4] = {link_at(gauge, L, Nc, nu, x), link_at(gauge, L, Nc, mu, xpmn),
                                             link_at(gauge, L, Nc, nu, xm), link_at(gauge, L, Nc, mu, xm)};
                        const int d2[4] = {0, 1, 1, 0};
                        path4(W, l2, d2, Nc, t1, t2);
                        for (int e = 0; e < Nc * Nc; ++e) Q[e] += W[e];
                        /* U_mu(x-mu)^+ U_nu(x-mu-nu)^+ U_mu(x-mu-nu) U_nu(x-nu) */
                        const cplx *l3[4] = {link_at(gauge, L, Nc, mu, xm), link_at(gauge, L, Nc, nu, xmn),
                                             link_at(gauge, L, Nc, mu, xmn), link_at(gauge, L, Nc, nu, xn)};
                        const int d3[4] = {1, 1, 0, 0};
                        path4(W, l3, d3, Nc, t1, t2);
                        for (int e = 0; e < Nc * Nc; ++e) Q[e] += W[e];
                        /* U_nu(x-nu)^+ U_mu(x-nu) U_nu(x-nu+mu) U_mu(x)^+ */
                        const cplx *l4[4] = {link_at(gauge, L, Nc, nu, xn), link_at(gauge, L, Nc, mu, xn),
                                             link_at(gauge, L, Nc, nu, xnpm), link_at(gauge, L, Nc, mu, x)};
                        const int d4[4] = {1, 0, 0, 1};
                        path4(W, l4, d4, Nc, t1, t2);
                        for (int e = 0; e < Nc * Nc; ++e) Q[e] += W[e];
                        /* F = ((Q - Q^+) - tr(Q - Q^+)/Nc) / (8 i) */
                        cplx tr = 0;
                        for (int a = 0; a < Nc; ++a) tr += Q[a * Nc + a] - conj(Q[a * Nc + a]);
                        cplx *Fp = F + pl2 * Nc * Nc;
                        for (int a = 0; a < Nc; ++a)
                            for (int b = 0; b < Nc; ++b) {
                                cplx v = Q[a * Nc + b] - conj(Q[b * Nc + a]);
                                if (a == b) v -= tr / Nc;
                                Fp[a * Nc + b] = v / (8.0 * I);
                            }
                    }
                for (int ch = 0; ch < 2; ++ch) {
                    cplx *blk = clov + (((long)p * Vh + i) * 2 + ch) * n * n;
                    for (int s = 0; s < 2; ++s)
                        for (int a = 0; a < Nc; ++a)
                            for (int t = 0; t < 2; ++t)
                                for (int b = 0; b < Nc; ++b) {
                                    cplx v = (s == t && a == b) ? 1.0 : 0.0;
                                    for (int q = 0; q < 6; ++q)
                                        v -= kappa * csw * sig[q][2 * ch + s][2 * ch + t] * F[q * Nc * Nc + a * Nc + b];
                                    blk[(s * Nc + a) * n + t * Nc + b] = v;
                                }
                }
            }
        free(Q); free(W); free(t1); free(t2); free(F);
    }
}

/* out = A in on one parity's multi-RHS half vector, A given as the two chirality
 * blocks per site (clover[Vh][2][2Nc][2Nc] of that parity). */
EXPORT void oracle_clover_apply(cplx *out, const cplx *blocks, const cplx *in, long Vh, int Nc, int R) {
    const int n = 2 * Nc;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < Vh; ++i)
        for (int ch = 0; ch < 2; ++ch) {
            const cplx *blk = blocks + (i * 2 + ch) * n * n;
            for (int row = 0; row < n; ++row)
                for (int r = 0; r < R; ++r) {
                    cplx s = 0;
                    for (int col = 0; col < n; ++col) {
                        const int t = col / Nc, b = col % Nc;
                        s += blk[row * n + col] * in[((i * 4 + 2 * ch + t) * Nc + b) * R + r];
                    }
                    const int sp = row / Nc, a = row % Nc;
                    out[((i * 4 + 2 * ch + sp) * Nc + a) * R + r] = s;
                }
        }
}

/* inverse of each n x n block (Gauss-Jordan with partial pivoting) */
EXPORT void oracle_clover_invert(cplx *inv, const cplx *blocks, long nblocks, int n) {
#pragma omp parallel
    {
        cplx *a = malloc(sizeof(cplx) * n * n * 2);
#pragma omp for schedule(static)
        for (long k = 0; k < nblocks; ++k) {
            const cplx *m = blocks + k * n * n;
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < 2 * n; ++j) a[i * 2 * n + j] = j < n ? m[i * n + j] : (j - n == i ? 1.0 : 0.0);
            for (int c = 0; c < n; ++c) {
                int piv = c;
                for (int i = c + 1; i < n; ++i)
                    if (cabs(a[i * 2 * n + c]) > cabs(a[piv * 2 * n + c])) piv = i;
                if (piv != c)
                    for (int j = 0; j < 2 * n; ++j) {
                        const cplx t = a[c * 2 * n + j];
                        a[c * 2 * n + j] = a[piv * 2 * n + j];
                        a[piv * 2 * n + j] = t;
                    }
                const cplx d = a[c * 2 * n + c];
                for (int j = 0; j < 2 * n; ++j) a[c * 2 * n + j] /= d;
                for (int i = 0; i < n; ++i)
                    if (i != c) {
                        const cplx f = a[i * 2 * n + c];
                        for (int j = 0; j < 2 * n; ++j) a[i * 2 * n + j] -= f * a[c * 2 * n + j];
                    }
            }
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j) inv[k * n * n + i * n + j] = a[i * 2 * n + n + j];
        }
        free(a);
    }
}
