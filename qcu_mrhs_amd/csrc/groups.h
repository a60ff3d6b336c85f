// groups.h -- batched solve of uneven right-hand-side groups.
//
// Several groups of right-hand sides (e.g. 3, 8, 17 and 32 sources) on the
// same gauge field are solved together: every vector of the batch holds the
// groups back to back (SegTab layout), so each step of the iteration is ONE
// launch over all groups (grid.y = group) and the inner iteration is captured
// once into a HIP graph and replayed.  Convergence is decided on the device:
// the fused x/r update raises a group's done flag when all of its RHS meet the
// threshold, and every later launch skips that group, so the graph needs no
// host decision between iterations.
//
// Precision scheme: an FP64 outer loop (defect correction) around an inner
// BiCGStab in the storage precision (FP32 by default): r = b' - A x in FP64,
// inner solve of A d = r / ||r|| to a relative 1e-5 (per RHS), x += ||r|| d in
// FP64, repeated until the FP64 residual of every RHS of a group meets the
// tolerance.  The dslash sites-per-wavefront of each group is chosen by
// autotune() (timed per group on the device).
#pragma once

#include <vector>

#include "blas.h"

namespace qcu {

struct GroupResult {
    int converged = 0;
    int iterations = 0;          // inner BiCGStab iterations, summed over outer steps
    int outer_steps = 0;         // FP64 residual recomputations that started an inner solve
    int spw = 0;                 // dslash sites per wavefront used for the group
    double true_rel_residual = -1.0;  // max_r ||b_r - M x_r|| / ||b_r||, FP64, full lattice
    double seconds = 0.0;        // from the start of the solve to the group's convergence
};

// groups dslash variant: matrix cores over the groups' RHS packed into shared 16-RHS chunks
constexpr int kGroupsPacked = 5;

// Process-wide override of the groups' dslash kernels (inner hops, FP64 hops): 0 = the timed choice,
// else 2 (VALU), 3 (matrix cores per group) or 5 (packed); applied after the timing of each batch shape.
void groups_force_variants(int inner, int fp64);

class GroupSolver {
public:
    GroupSolver(const Geom &g, int Nc, const std::vector<int> &Rs);
    ~GroupSolver();
    GroupSolver(const GroupSolver &) = delete;
    GroupSolver &operator=(const GroupSolver &) = delete;

    const std::vector<int> &sizes() const { return Rs_; }
    int total_rhs() const { return tab_.Rtot; }
    long half_n() const { return n_; }

    // Batch full vectors (FP64): [parity][group-concatenated half vector].
    void *b_full() const { return b_; }
    void *x_full() const { return x_; }

    // Allocates the batch buffers and picks each group's dslash shape (sites
    // per wavefront) by timing it on the inner precision; cached per precision.
    hipError_t autotune(const void *glo, int prec, hipStream_t s);
    // dslash kernel the autotuner chose for the batch (2 = VALU wave-per-site, 3 = FP32 matrix cores per
    // group, 5 = FP32 matrix cores over the groups' RHS packed into shared chunks)
    int dslash_variant() const { return variant_; }
    // and for the FP64 hops of the defect correction (2 = VALU, 3 / 5 = FP64 matrix cores; timed in solve)
    int dslash_variant64() const { return variant64_; }

    // Solve M x = b for the batch held in b_full(); result in x_full().
    hipError_t solve(const void *g64, const void *glo, int inner_prec, double kappa, int max_iter, double tol,
                     std::vector<GroupResult> &res, hipStream_t s);

private:
    hipError_t ensure(int prec);
    hipError_t dslash(void *out, const void *in, const void *gauge, int prec, int parity, const void *xpay,
                      double coef, const SegTab &tab, hipStream_t s);
    hipError_t iteration(const void *glo, int P, double kappa, hipStream_t s);
    hipError_t tune64(const void *g64, hipStream_t s);
    SegTab tab_with(const int *done) const;

    Geom g_;
    int Nc_;
    std::vector<int> Rs_;
    long n_;
    SegTab tab_;
    int by_ = 0, bz_ = 0;
    // group dslash kernel (autotuned): 2 = wave-per-site VALU, 3 = matrix cores per group, 5 = matrix cores
    // with the groups' RHS packed into shared 16-RHS chunks (sites per wavefront spw_packed_)
    int variant_ = 2;
    int spw_packed_ = 1;
    int variant64_ = 2;  // the same for the FP64 hops
    int spw_packed64_ = 1;
    bool tuned64_ = false;
    int lo_prec_ = -1;
    bool tuned_ = false;
    int tuned_prec_ = -1;
    // FP64 batch vectors
    void *b_ = nullptr, *x_ = nullptr, *bp_ = nullptr, *r64_ = nullptr, *t64_ = nullptr, *u64_ = nullptr;
    // inner-precision vectors
    void *r_ = nullptr, *r0_ = nullptr, *p_ = nullptr, *v_ = nullptr, *s_ = nullptr, *t_ = nullptr,
         *tmp_ = nullptr, *xl_ = nullptr;
    C64 *st_ = nullptr, *ones_ = nullptr;
    double *dn_ = nullptr, *scale_ = nullptr;
    int *done_ = nullptr, *odone_ = nullptr, *iters_ = nullptr;
    double *h_ = nullptr;  // pinned staging
    SegWork work_;
};

}  // namespace qcu
