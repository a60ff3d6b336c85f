/*
 * qcu.h -- C ABI of the MI355X (gfx950) multi-RHS SU(N) Wilson dslash and
 * BiCGStab library (libqcu.so).
 *
 * The first eleven entry points keep the names, argument meaning, layouts
 * and call protocol of the reference library's C ABI (include/qcu.h:7-29 and
 * src/interface.cu:12-141 of YellowBicycleee/qcu_mrhs), so a caller that
 * binds the reference can bind this library unchanged.  Every pointer named
 * "fermion" or "gauge" is DEVICE memory owned by the caller.
 *
 * Call protocol (src/interface.cu, src/qcu_interface.cu):
 *   initGridSize -> loadQcuGauge -> getDslash -> R x pushBackFermions
 *   -> exactly one of start_dslash / mat_Qcu / qcuInvert (each consumes the
 *   queue) -> ... -> finalizeQcu
 *
 * Errors: the reference aborts through errorQcu().  Here every entry point
 * records the failure (qcuLastError / qcuLastErrorString), prints it to
 * stderr and returns; set QCU_ABORT_ON_ERROR=1 in the environment to abort
 * like the reference.
 */
#pragma once

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* include/qcu.h:7-13 -- index order X, Y, Z, T (src/interface.cu:23-30) */
typedef struct QcuParam_s {
    int lattice_size[4];
} QcuParam;

typedef struct QcuGrid_t {
    int grid_size[4];
} QcuGrid;

/* Precision codes.  The reference's QcuPrecision enum lives in its absent
 * qcu_io submodule; these values follow the declaration order used by every
 * precision switch of the reference (src/data_format/qcu_data_format_shift.cu:
 * 82-93).  kQcuPrecisionBFloat16 has no reference counterpart. */
enum {
    kQcuPrecisionHalf = 0,
    kQcuPrecisionSingle = 1,
    kQcuPrecisionDouble = 2,
    kQcuPrecisionUndefined = 3,
    kQcuPrecisionBFloat16 = 4
};

/* include/qcu_helper_macro.h:7-11 */
enum { kQcuDslashWilson = 0, kQcuDslashClover = 1 };

/* Error codes returned by qcuLastError(). */
enum {
    kQcuSuccess = 0,
    kQcuErrorNotInitialized = 1,
    kQcuErrorInvalidValue = 2,
    kQcuErrorQueue = 3,
    kQcuErrorUnsupported = 4,
    kQcuErrorDevice = 5,
    kQcuErrorNotConverged = 6,
    kQcuErrorBreakdown = 7
};

/* ---- reference entry points (include/qcu.h:14-29) -------------------- */

/* src/interface.cu:20-38.  Allocates the internal multi-RHS buffers.
 * inputFloatPrecision: precision of the user fermion vectors (the
 * reference's "out_float_precision"); dslashFloatPrecision: precision the
 * solver iterates in (kQcuPrecisionDouble = plain FP64 BiCGStab; Single /
 * Half / BFloat16 = reliable-update mixed precision with that inner
 * precision).  grid_size must be all ones (one GPU). */
void initGridSize(QcuGrid *grid, QcuParam *param, int n_color, int m_rhs,
                  int inputFloatPrecision, int dslashFloatPrecision);
/* src/interface.cu:40-43: queue one (out, in) pair of device pointers. */
void pushBackFermions(void *fermionOut, void *fermionIn);
/* src/interface.cu:45-48 / src/qcu_interface.cu:257-268: copy an even/odd
 * gauge field [mu][parity][t][z][y][x/2][Nc][Nc] into internal storage. */
void loadQcuGauge(void *gauge, int floatPrecision);
/* src/interface.cu:50-53 / src/qcu_interface.cu:66-107: kappa = 1/(2(4+mass)).
 * dslashType kQcuDslashWilson: M = 1 - kappa D.  kQcuDslashClover (declared
 * by the reference, include/qcu_helper_macro.h:7-11, but rejected by its
 * getDslash, src/qcu_interface.cu:97-105): the Wilson-clover operator
 * M = A - kappa D, A = 1 - kappa c_sw sum_{mu<nu} sigma_{mu nu} F_{mu nu}
 * with the clover-leaf field strength; c_sw from qcuSetCloverCoefficient
 * (default 1).  The term and its inverse are built on the device at the
 * first mat_Qcu / qcuInvert after the gauge field, kappa or c_sw changed. */
void getDslash(int dslashType, double mass);
/* src/interface.cu:55-58 / src/qcu_interface.cu:109-172: out = D in on one
 * parity; queued pointers are half-lattice vectors, `parity` is the output
 * parity. */
void start_dslash(int parity, int daggerFlag);
/* src/interface.cu:59-62 / src/qcu_interface.cu:174-255: out = (1 - kappa D) in
 * on full-lattice vectors ((A - kappa D) in with the clover type).  start_dslash
 * applies the hopping term D alone for either type. */
void mat_Qcu(int daggerFlag);
/* src/interface.cu:70-73 / src/qcu_interface.cu:280-362: solve M x = b for
 * every queued (x, b) pair; even/odd preconditioned BiCGStab. */
void qcuInvert(int max_iteration, double p_max_prec);
/* src/interface.cu:64-68 */
void finalizeQcu(void);
/* src/interface.cu:76-105: lexicographic [mu][t][z][y][x][Nc][Nc] ->
 * even/odd gauge (device pointers). */
void gauge_eo_precondition(void *prec_gauge, void *non_prec_gauge, int precision);
/* src/interface.cu:107-134: inverse of gauge_eo_precondition. */
void gauge_reverse_eo_precondition(void *non_prec_gauge, void *prec_gauge, int precision);
/* src/interface.cu:137-141: the reference reads through its absent qcu_io
 * submodule; this library reads a raw little-endian complex128 file
 * "<prefix>" holding the lexicographic layout and writes the even/odd field
 * to the device pointer `gauge`. */
void read_gauge_from_file(void *gauge, const char *file_path_prefix);

/* ---- additions (no reference counterpart) ----------------------------- */

int qcuLastError(void);
const char *qcuLastErrorString(void);
void qcuClearError(void);

typedef struct QcuSolveInfo_s {
    int converged;             /* 1 if every RHS met the tolerance            */
    int iterations;            /* BiCGStab iterations (inner iterations)      */
    int reliable_updates;      /* FP64 true-residual recomputations           */
    int restarts;              /* breakdown restarts                          */
    double true_rel_residual;  /* max_r ||b_r - M x_r|| / ||b_r||, FP64, full */
    double seconds;            /* wall time of the solve (gather..scatter)    */
    int preconditioned;        /* 1 if the learned preconditioner produced x  */
    int fallback;              /* 1 if the preconditioned solve failed and
                                  plain BiCGStab (from x0 = 0) produced x     */
} QcuSolveInfo;

/* Clover coefficient c_sw of the kQcuDslashClover operator (default 1.0);
 * takes effect at the next mat_Qcu / qcuInvert.  Requires initGridSize. */
int qcuSetCloverCoefficient(double csw);

/* qcuInvert with status: returns kQcuSuccess or an error code, fills info. */
int qcuInvertEx(int max_iteration, double tol, QcuSolveInfo *info);

/* Solver options.  inner_precision: precision of the inner (sloppy) solve,
 * overriding initGridSize's dslashFloatPrecision; delta: reliable-update
 * threshold (update when ||r|| < delta * max ||r|| since the last update;
 * 0 = default for the inner precision: 1e-3 FP32, 5e-3 FP16, 1e-2 BF16);
 * use_preconditioner: 1 = apply the loaded learned right preconditioner; if
 * that solve breaks down, diverges or does not converge, plain BiCGStab
 * solves the same system again and QcuSolveInfo.fallback is set. */
int qcuSetSolverOptions(int inner_precision, double delta, int use_preconditioner);

/* ---- learned preconditioner (no reference counterpart) ---------------
 * A linear two-layer gauge-covariant network (parallel-transport
 * convolutions, 1 -> C -> 1 channels, nearest-neighbour hops, complex 4x4
 * spin-matrix weights) used as a fixed right preconditioner of the odd
 * Schur system when qcuSetSolverOptions(..., use_preconditioner = 1).
 * Weights: W1[C][9][4][4], W2[C][9][4][4] complex, passed as interleaved
 * (re, im) doubles; hop 0 = identity, hop 1 + 2 mu (+1) = forward (backward)
 * transport along mu.  All calls use the gauge field loaded by loadQcuGauge. */
int qcuLgnnSetWeights(int channels, const double *w1, const double *w2);
int qcuLgnnGetWeights(double *w1, double *w2);
/* out = N in on odd-parity multi-RHS half vectors [Vh][4][Nc][R] (device). */
int qcuLgnnApply(void *out, const void *in, int precision);
/* One forward/backward pass on a batch of R pairs (FP32 complex odd half
 * vectors): loss = sum_r ||N in_r - target_r||^2 / ||target_r||^2 and the
 * gradients dL/dRe W + i dL/dIm W (host arrays shaped like the weights). */
int qcuLgnnLossGrad(const void *in, const void *target, double *loss, double *gw1, double *gw2);
/* Solve the queued systems with plain FP64 BiCGStab and keep n_snapshots
 * (error, residual) pairs per RHS -- e_k = x - x_k, r_k = A e_k of the odd
 * Schur system at evenly spaced iterations -- plus n_random pairs with e a
 * complex Gaussian vector (drawn from `seed`) and r = A e, as training data
 * for the current gauge field.  *n_batches receives the total number of
 * stored batches (R pairs each).  The queue is consumed. */
int qcuLgnnCollect(int max_iteration, double tol, int n_snapshots, int n_random, unsigned seed, int *n_batches);
/* Adam over the stored batches (one step per batch, order shuffled with
 * `seed`, learning rate lr_e = learning_rate (1 + cos(pi e / epochs)) / 2 in
 * epoch e); loss_per_epoch[e] = mean loss per pair. */
int qcuLgnnTrain(int epochs, double learning_rate, unsigned seed, double *loss_per_epoch);
int qcuLgnnClearData(void);
/* Training objective of qcuLgnnTrain: power = 0 (default) is the pair loss
 * above; power = K in 1..8 is the residual-power loss
 *     sum_r ||(1 - A N)^K w_r||^2,  w_r = v_r / ||v_r||,
 * on a fixed complex Gaussian vector v per stored batch (A the FP32 odd Schur
 * operator of the batch's configuration): it weights most the modes that A N
 * leaves far from 1, which set the iteration count of the preconditioned
 * solve.  Back-propagated through the K factors with A^+ and N^+. */
int qcuLgnnSetObjective(int power);
/* One forward/backward pass of the residual-power loss on the loaded gauge
 * field for R FP32 odd vectors v: loss and gradients as for qcuLgnnLossGrad. */
int qcuLgnnPowerLossGrad(const void *v, int power, double *loss, double *gw1, double *gw2);
/* Projected network (on = 1; 0, the default, leaves the weights general):
 * every second-layer weight of hop h = 1..8 is kept in the image of that
 * hop's Wilson projector, W2[o][h] = P_h W2[o][h], P_h = (1 - gamma_mu) / 2
 * forward and (1 + gamma_mu) / 2 backward -- the spin structure of the
 * hopping term, which the first Neumann term already has.  The weights are
 * projected now, whenever they are set and after every training step, and
 * the FP16 / BF16 application stores two spins of each hidden hop field
 * instead of four (half its traffic).  The mode persists across
 * qcuLgnnSetWeights calls until switched off. */
int qcuLgnnSetProjected(int on);

/* ---- batched uneven right-hand-side groups (no reference counterpart) --
 * Solves M x = b for up to 8 groups of sources of different sizes on the
 * loaded gauge field in one batch: every step of the iteration is one launch
 * over all groups and the inner iteration is replayed from a captured HIP
 * graph; groups finish independently (device-side convergence flags).
 * Precision: inner BiCGStab in the inner precision of qcuSetSolverOptions /
 * initGridSize (FP32 storage recommended) inside an FP64 defect-correction
 * loop, so every group reaches the FP64 tolerance `tol`.  Pointers are device
 * full-lattice fermions in the user precision of initGridSize; the m_rhs of
 * initGridSize does not limit the group sizes (1..128 each). */
typedef struct QcuRhsGroup_s {
    int m_rhs;
    void **fermion_out;  /* m_rhs device pointers: solutions x */
    void **fermion_in;   /* m_rhs device pointers: sources b   */
} QcuRhsGroup;

typedef struct QcuGroupInfo_s {
    int converged;             /* every RHS of the group met tol (FP64 residual) */
    int iterations;            /* inner iterations summed over outer steps        */
    int outer_steps;           /* FP64 defect-correction steps                    */
    int sites_per_wave;        /* dslash launch shape chosen by the autotuner     */
    double true_rel_residual;  /* max_r ||b_r - M x_r|| / ||b_r||, FP64           */
    double seconds;            /* solve start -> this group's convergence          */
    double batch_seconds;      /* wall time of the whole call                     */
} QcuGroupInfo;

int qcuInvertGroups(int n_groups, const QcuRhsGroup *groups, int max_iteration, double tol, QcuGroupInfo *info);
/* Dslash kernel the last qcuInvertGroups batch ran, as chosen by its timing
 * of the candidates: 2 = VALU wave per site, 3 = FP32 matrix cores
 * (v_mfma_f32_16x16x4_f32, FP32 storage) per group, 5 = FP32 matrix cores
 * over the groups' right-hand sides packed into shared 16-RHS chunks; 0
 * before any batch, -1 without initGridSize. */
int qcuGroupsDslashVariant(void);
/* The same for the FP64 hops of the batch's defect correction: 2 = VALU wave
 * per site, 3 / 5 = FP64 matrix cores (v_mfma_f64_16x16x4_f64) per group /
 * over packed chunks; timed once per batch shape.  No reference counterpart. */
int qcuGroupsDslashVariant64(void);
/* Choose the batched groups' dslash kernels for every later qcuInvertGroups
 * call in this process (after each batch shape's own timing): variant for the
 * inner hops, variant64 for the FP64 hops; 0 = the timed choice, 2 = VALU wave
 * per site, 3 = matrix cores per group, 5 = matrix cores over packed RHS
 * chunks (3 and 5 need FP32 / FP64 storage; otherwise the timed choice
 * stays).  No reference counterpart. */
int qcuGroupsDslashTune(int variant, int variant64);

/* Device-memory plumbing for callers that have no other GPU runtime. */
void *qcuDeviceMalloc(size_t bytes);
int qcuDeviceFree(void *ptr);
int qcuMemcpyHtoD(void *dst, const void *src, size_t bytes);
int qcuMemcpyDtoH(void *dst, const void *src, size_t bytes);
int qcuMemsetD(void *dst, int value, size_t bytes);
int qcuDeviceSynchronize(void);

#ifdef __cplusplus
}
#endif
