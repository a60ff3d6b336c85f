/*
 * qcu_ops.h -- stateless device-pointer operations of libqcu.so on the
 * internal multi-RHS layout.  They expose, as plain C, the internal operator
 * and BLAS interfaces that the reference keeps in C++:
 *
 *   qcuDslashMrhs       <- qcu::Dslash::apply(DslashParam) (include/qcd/qcu_dslash.h:10-74)
 *   qcuCloverBuild/Apply   the clover term the reference declares but does not implement
 *   qcuBlas*            <- qcu::qcu_blas functors (include/qcu_blas/qcu_blas_complex_*.h,
 *                          src/qcu_blas/qcu_complex_{xsay,xpay,axpby,axpbypcz,dotc,norm}.cu)
 *   qcuColorSpinorGather/Scatter <- colorSpinorGather/Scatter
 *                          (include/data_format/qcu_data_format_shift.cuh:8-17)
 *   qcuCopyComplexVector <- copyComplexVector_interface (same header)
 *
 * Layouts: MRHS half vector [Vh][4][Nc][R] complex, gauge [4][2][Vh][Nc][Nc]
 * complex, both in the storage precision `precision` (kQcuPrecision*).
 * Per-RHS scalar arrays are device arrays of R complex FP64 values (the
 * reference keeps them in the vector precision).  All calls are asynchronous
 * on the library stream unless stated; qcuDeviceSynchronize waits.
 * Return: 0 on success, else a kQcuError* code (see qcu.h).
 */
#pragma once

#include "qcu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* out (parity `parity`) = D in (parity 1 - parity); optionally
 * out = xpay + coef * D in when xpay != NULL (else out = coef * D in). */
int qcuDslashMrhs(void *out, const void *in, const void *gauge, const int lattice[4], int n_color,
                  int m_rhs, int parity, int dagger, int precision, const void *xpay, double coef);

/* Clover term of the Wilson-clover operator (no reference counterpart: the
 * reference declares the type, include/qcu_helper_macro.h:7-11, and rejects it,
 * src/qcu_interface.cu:97-105).  qcuCloverBuild writes
 * A = 1 - kappa c_sw sum_{mu<nu} sigma_{mu nu} F_{mu nu} from an FP64 even/odd
 * gauge field as FP64 blocks clover[parity 2][Vh][chirality 2][2Nc][2Nc]
 * (row/column = spin within the pair {0,1} or {2,3} times Nc + colour) and,
 * when clover_inv != NULL, the inverse of every block in the same layout. */
int qcuCloverBuild(void *clover, void *clover_inv, const void *gauge, const int lattice[4], int n_color,
                   double csw, double kappa);
/* out = A in + coef * xpay (xpay may be NULL) on one parity's MRHS half
 * vectors [Vh][4][Nc][R] in `precision`; `blocks` = that parity's
 * [Vh][2][2Nc][2Nc] (a term or its inverse) in clover_precision
 * (kQcuPrecisionDouble or kQcuPrecisionSingle).  out may equal in or xpay. */
int qcuCloverApply(void *out, const void *in, const void *blocks, const int lattice[4], int n_color, int m_rhs,
                   int precision, int clover_precision, const void *xpay, double coef);

/* res = x - a_r y  (src/qcu_blas/qcu_complex_xsay.cu) */
int qcuBlasXsay(void *res, const void *x, const void *a, const void *y, long n, int m_rhs,
                int precision);
/* res = x + a_r y  (src/qcu_blas/qcu_complex_xpay.cu) */
int qcuBlasXpay(void *res, const void *x, const void *a, const void *y, long n, int m_rhs,
                int precision);
/* res = a_r x + b_r y  (src/qcu_blas/qcu_complex_axpby.cu) */
int qcuBlasAxpby(void *res, const void *a, const void *x, const void *b, const void *y, long n,
                 int m_rhs, int precision);
/* res = a_r x + b_r y + c_r z  (src/qcu_blas/qcu_complex_axpbypcz.cu) */
int qcuBlasAxpbypcz(void *res, const void *a, const void *x, const void *b, const void *y,
                    const void *c, const void *z, long n, int m_rhs, int precision);
/* res[r] = sum_i conj(a[i*R+r]) b[i*R+r]  -> R complex FP64 (device)
 * (src/qcu_blas/qcu_complex_dotc.cu) */
int qcuBlasDotc(void *res, const void *a, const void *b, long n, int m_rhs, int precision);
/* qcuBlasDotc without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * block partials live in a fixed 1 MiB device scratch that the library holds
 * on every device it runs on until the process ends; m_rhs must be <= 256
 * (larger values are rejected with kQcuErrorInvalidValue).  For m_rhs <= 32
 * the result is bit-identical to qcuBlasDotc's; above that it uses fewer
 * blocks (same FP64 accumulation, another summation order).  Calls share the
 * scratch of their device: calls that overlap in time on one GPU race. */
int qcuBlasDotcNoAlloc(void *res, const void *a, const void *b, long n, int m_rhs, int precision);
/* res[r] = ||a_r||_2 -> R FP64 (device)  (src/qcu_blas/qcu_complex_norm.cu) */
int qcuBlasNorm(void *res, const void *a, long n, int m_rhs, int precision);

/* Gather R per-RHS vectors (host array of R device pointers) of nsites*4*Nc
 * complex numbers into [nsites][4][Nc][R], converting precision; scatter is
 * the inverse. */
int qcuColorSpinorGather(void *dst, int dst_precision, void *const *src_ptrs, int src_precision,
                         long nsites, int n_color, int m_rhs);
int qcuColorSpinorScatter(void *const *dst_ptrs, int dst_precision, const void *src,
                          int src_precision, long nsites, int n_color, int m_rhs);
/* dst = src converted between precisions, `count` complex numbers */
int qcuCopyComplexVector(void *dst, int dst_precision, const void *src, int src_precision,
                         long count);

/* Launch shape of the dslash for every later call in this process (the
 * stateless qcuDslashMrhs and the solver alike): variant 0 = automatic
 * (tuning table, else MFMA for FP16/BF16 and wave-per-site VALU otherwise),
 * 1 = lane per (site, rhs), 2 = wavefront per site (VALU), 3 = wavefront per
 * site on MFMA (FP16/BF16 only), 4 = MFMA with the input rows staged in LDS
 * (FP16/BF16, Nc <= 4; spw = t steps per workgroup, 0 = 12; other shapes fall
 * back to variant 3 or 2); by, bz = traversal tile in y and z and spw =
 * sites per wavefront (0 = choose; for variant 3 a negative spw = a persistent
 * grid of -spw workgroups per CU whose wavefronts stride through their XCD's
 * share of the sites, -64 <= spw).  No reference counterpart. */
int qcuDslashTune(int variant, int by, int bz, int spw);

/* Autotune the dslash for (lattice, n_color, m_rhs, precision): times the
 * built-in default shape and kernel variants 1-3 over sites-per-wavefront
 * and traversal-tile candidates, and records the fastest for later automatic
 * launches (the default stays unless a candidate is more than 2 % faster).  best[4] receives
 * {variant, by, bz, spw}, *best_us its kernel time; records (may be NULL,
 * capacity max_records, 5 doubles each: variant, by, bz, spw, us) and
 * *n_records list every candidate.  No reference counterpart. */
int qcuDslashAutotune(const int lattice[4], int n_color, int m_rhs, int precision, int repeats, int best[4],
                      double *best_us, double *records, int max_records, int *n_records);
/* Enter a launch shape into the tuning table (e.g. from a saved tuning cache). */
int qcuDslashSetTuned(const int lattice[4], int n_color, int m_rhs, int precision, int variant, int by, int bz,
                      int spw);
/* Write the tuning table to a text file (one entry per line: Lx Ly Lz Lt Nc
 * R precision variant by bz spw) / add the entries of such a file to the
 * table (*n_loaded receives their number).  No reference counterpart; the
 * reference's launch shapes are compile-time constants
 * (include/qcu_wmma_constant.h:7-22). */
int qcuDslashTuneSave(const char *path);
int qcuDslashTuneLoad(const char *path, int *n_loaded);

/* Device timing of the stateless operations: HIP events recorded on the
 * stream those operations run on.  qcuTimerStopMs waits for the stop event
 * and returns the elapsed milliseconds (negative on error). */
int qcuTimerStart(void);
double qcuTimerStopMs(void);

#ifdef __cplusplus
}
#endif
