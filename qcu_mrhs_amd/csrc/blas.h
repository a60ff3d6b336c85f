// blas.h -- multi-RHS BLAS and single-pass reductions on RHS-innermost vectors.
//
// A vector of `n` entries per RHS is stored [n][R]; per-RHS scalars are
// arrays of R FP64 complex numbers in device memory.  Every kernel handles
// all R right-hand sides in one launch (the reference launches one kernel per
// RHS, src/qcu_blas/qcu_complex_xsay.cu:21-26).
#pragma once

#include "qcu_common.h"

namespace qcu {

using C64 = Cx<double>;

// res = x - a_r y | x + a_r y | a_r x + b_r y | a_r x + b_r y + c_r z
enum AxpyKind : int { kXsay = 0, kXpay = 1, kAxpby = 2, kAxpbypcz = 3 };

struct AxpyArgs {
    int kind;
    void *res;
    const void *x, *y, *z;
    const C64 *a, *b, *c;  // device scalar arrays [R]
    long n;                // entries per RHS
    int R;
    int prec;              // storage precision of res/x/y/z
};
hipError_t launch_axpy(const AxpyArgs &a, hipStream_t s);

// Reduction workspace: per-block partial sums, one region per reduction slot.
struct ReduceWork {
    double *partials = nullptr;        // [kReduceSlots][max_blocks][R][4]
    int max_blocks = 0;
};
constexpr int kReduceSlots = 16;
constexpr int kReduceThreads = 256;

hipError_t reduce_work_alloc(ReduceWork &w, int R);
void reduce_work_free(ReduceWork &w);
// number of blocks a reduction over n*R entries uses (fixed => deterministic)
int reduce_blocks(long n, int R, int max_blocks);

// <a,b>_r -> res[r] (complex, FP64 accumulation)
hipError_t launch_dotc(C64 *res, const void *a, const void *b, long n, int R, int prec,
                       ReduceWork &w, int slot, hipStream_t s);
// res = x - y and ||res||_r -> nrm[r] in one pass (sqrt when take_sqrt)
hipError_t launch_xsay_norm(double *nrm, void *res, const void *x, const void *y, long n, int R, int prec,
                            bool take_sqrt, ReduceWork &w, int slot, hipStream_t s);
// FP64 residual d = b - y, ||d||_r -> nrm[r] (sqrt), and lo = scale[r] d in the
// inner storage `lo_prec` (FP32/FP16/BF16), in one pass; d itself is not stored
hipError_t launch_residual_lo(double *nrm, void *lo, int lo_prec, const void *b, const void *y, const double *scale,
                              long n, int R, ReduceWork &w, int slot, hipStream_t s);
// <a,b>_r -> res[r] like launch_dotc, with the block partials in the library's
// fixed device scratch instead of a ReduceWork (no device allocation; calls
// on one device share the scratch, so overlapping calls race).  R <= 256;
// for R <= kNoAllocFullRhs the block count, and so the result, is the same
// as launch_dotc's with a ReduceWork of 1024 blocks.
constexpr long kNoAllocScratchBytes = 1L << 20;
constexpr int kNoAllocFullRhs = int(kNoAllocScratchBytes / (8 * 4 * 1024));
hipError_t launch_dotc_noalloc(C64 *res, const void *a, const void *b, long n, int R, int prec, hipStream_t s);
// ||a||_r -> res[r] (sqrt applied when take_sqrt, else squared norm)
hipError_t launch_norm(double *res, const void *a, long n, int R, int prec, bool take_sqrt,
                       ReduceWork &w, int slot, hipStream_t s);

// Copy with precision conversion of `count` complex numbers, optionally
// scaled per RHS: dst[i] = scale[i % R] * src[i] (scale may be null).
hipError_t launch_convert(void *dst, int dst_prec, const void *src, int src_prec, long count,
                          const double *scale, int R, hipStream_t s);

// ---------------------------------------------------------------------------
// BiCGStab fused steps (all RHS, one launch each).  Scalar state layout:
// st[k*R + r] for the slots below, FP64 complex.
// ---------------------------------------------------------------------------
enum ScalarSlot : int {
    kRho = 0,      // <r0, r>
    kRhoNew,       // <r0, r_new>
    kAlpha,
    kOmega,
    kBeta,
    kR0v,          // <r0, v>
    kTs,           // <t, s>
    kTt,           // <t, t>   (real part)
    kRnorm2,       // ||r||^2  (real part)
    kBnorm2,       // ||b'||^2 (real part)
    kPScale,       // factor folded into the next p update (reliable-update rescale)
    kTmpA,
    kTmpB,
    kNumSlots
};

// alpha = rho / <r0, v>
hipError_t launch_step_alpha(C64 *st, const void *r0, const void *v, long n, int R, int prec,
                             ReduceWork &w, hipStream_t s);
// The same finishers from per-RHS partial sums computed elsewhere (the dslash
// epilogue): partials [nblk][R][4] in the layout of the reductions above,
// (Re, Im) of <r0, v> for alpha; (Re, Im) of <t, s> and <t, t> for omega.
hipError_t launch_finish_alpha(C64 *st, int R, const double *partials, int nblk, hipStream_t s);
hipError_t launch_finish_omega(C64 *st, int R, const double *partials, int nblk, hipStream_t s);
// s = r - alpha v
hipError_t launch_step_s(void *sv, const void *r, const void *v, const C64 *st, long n, int R,
                         int prec, hipStream_t s);
// omega = <t,s> / <t,t>
hipError_t launch_step_omega(C64 *st, const void *t, const void *sv, long n, int R, int prec,
                             ReduceWork &w, hipStream_t s);
// x += alpha p + omega shat ; r = s - omega t ; ||r||^2, <r0,r> ;
// beta = ((alpha * rho_new) / omega) / rho ; rho <- rho_new.
// Unpreconditioned: p = p, shat = s.  Right-preconditioned: p = N p, shat = N s.
hipError_t launch_step_xr(void *x, void *r, const void *p, const void *shat, const void *sv,
                          const void *t, const void *r0, C64 *st, long n, int R, int prec,
                          ReduceWork &w, hipStream_t s);
// p = r + beta * pscale * (p - omega v)
hipError_t launch_step_p(void *p, const void *r, const void *v, const C64 *st, long n, int R,
                         int prec, hipStream_t s);
// set every slot to the values needed at the start of a solve
hipError_t launch_scalar_fill(C64 *dst, C64 value, int count, hipStream_t s);

// ---------------------------------------------------------------------------
// Batched groups (SegTab layout, qcu_common.h): one launch covers every group
// (grid.y = group); groups whose done flag is set are skipped.  Scalars use
// slot stride tab.Rtot.  Reductions keep partials per group (the slot
// argument is accepted for symmetry with the single-group calls; a batch's
// reductions run in stream order).
// ---------------------------------------------------------------------------
struct SegWork {
    double *partials = nullptr;        // [G][max_blocks][Rmax][4]
    int max_blocks = 0, G = 0, Rmax = 0;
};
hipError_t seg_work_alloc(SegWork &w, int G, int Rmax);
void seg_work_free(SegWork &w);
hipError_t seg_dotc(C64 *res, const void *a, const void *b, int prec, const SegTab &tab, SegWork &w, int slot,
                    hipStream_t s);
hipError_t seg_norm(double *res, const void *a, int prec, bool take_sqrt, const SegTab &tab, SegWork &w, int slot,
                    hipStream_t s);
hipError_t seg_step_alpha(C64 *st, const void *r0, const void *v, int prec, const SegTab &tab, SegWork &w,
                          hipStream_t s);
hipError_t seg_step_s(void *sv, const void *r, const void *v, const C64 *st, int prec, const SegTab &tab,
                      hipStream_t s);
hipError_t seg_step_omega(C64 *st, const void *t, const void *sv, int prec, const SegTab &tab, SegWork &w,
                          hipStream_t s);
// x += alpha p + omega s ; r = s - omega t ; ||r||^2 ; <r0, r> ; beta ; then per group:
// iters[g] += 1 and done[g] = 1 once every RHS has ||r||^2 <= st[kBnorm2]
hipError_t seg_step_xr(void *x, void *r, const void *p, const void *sv, const void *t, const void *r0, C64 *st,
                       int *done, int *iters, int prec, const SegTab &tab, SegWork &w, hipStream_t s);
// p = r + beta (p - omega v)
hipError_t seg_step_p(void *p, const void *r, const void *v, const C64 *st, int prec, const SegTab &tab,
                      hipStream_t s);
// res = x - a[rhs] y
hipError_t seg_xsay(void *res, const void *x, const void *y, const C64 *a, int prec, const SegTab &tab,
                    hipStream_t s);
// dst = scale[rhs] src (scale may be null)
hipError_t seg_convert(void *dst, int dst_prec, const void *src, int src_prec, const double *scale,
                       const SegTab &tab, hipStream_t s);
// xhi (FP64) += scale[rhs] xlo
hipError_t seg_accum(void *xhi, const void *xlo, int lo_prec, const double *scale, const SegTab &tab,
                     hipStream_t s);

}  // namespace qcu
