// solver.h -- even/odd preconditioned multi-RHS BiCGStab on gfx950.
#pragma once

#include <functional>
#include <vector>

#include "blas.h"
#include "clover.h"
#include "dslash.h"

namespace qcu {

// Right preconditioner N for the odd Schur system: the solver iterates on
// A N y = b' and returns x = N y.
class Preconditioner {
public:
    virtual ~Preconditioner() = default;
    // out = N in on odd-parity MRHS half vectors stored in precision `prec`
    virtual hipError_t apply(void *out, const void *in, int prec, hipStream_t s) = 0;
    virtual bool supports(int prec) const = 0;
};

struct SolverOptions {
    int inner_prec = kDouble;  // kDouble: plain FP64 BiCGStab; else reliable-update mixed precision
    double delta = 0.0;        // reliable update when ||r|| < delta * max ||r|| since last update
                               // (0: 1e-3 FP32, 5e-3 FP16, 1e-2 BF16 inner precision)
    int use_precond = 0;       // apply the learned right preconditioner (inner solve only)
    double diverge = 1e6;      // stop (not converged) once a relative residual estimate exceeds this
    int verify_residual = 0;   // recompute b - M x on the full lattice after a mixed solve (else the
                               // final reliable update's FP64 residual, which equals it, is reported)
    int verbose = 0;
};

struct SolveResult {
    int converged = 0;
    int iterations = 0;
    int reliable_updates = 0;
    int restarts = 0;
    int diverged = 0;                 // stopped because a residual estimate exceeded SolverOptions::diverge
    double true_rel_residual = -1.0;  // max_r ||b - M x|| / ||b|| over the full lattice
    std::vector<double> rel_residual;  // per RHS, odd Schur system, last estimate
};

// Device fields the solver reads: the FP64 gauge field and (for a mixed solve)
// a copy in the inner precision.
struct GaugeSet {
    const void *g64 = nullptr;
    const void *glo = nullptr;
    int lo_prec = kDouble;
};

class BiCGStab {
public:
    BiCGStab(const Geom &g, int Nc, int R);
    ~BiCGStab();
    BiCGStab(const BiCGStab &) = delete;
    BiCGStab &operator=(const BiCGStab &) = delete;

    // b_full, x_full: FP64 MRHS full vectors [2][Vh][4][Nc][R] on the device.
    hipError_t solve(const void *b_full, void *x_full, const GaugeSet &gauge, double kappa,
                     int max_iter, double tol, const SolverOptions &opt, Preconditioner *prec,
                     SolveResult &res, hipStream_t s);

    // Wilson-clover operator M = A - kappa D (null: Wilson, A = 1).  The fields
    // must stay alive while set; the odd Schur system becomes
    // A_oo - kappa^2 D_oe A_ee^{-1} D_eo, b' = b_o + kappa D_oe A_ee^{-1} b_e and
    // x_e = A_ee^{-1} (b_e + kappa D_eo x_o).
    void set_clover(const CloverSet *c) { clover_ = c; }

    // FP64 full-lattice true residual max_r ||b - M x|| / ||b||.
    hipError_t true_residual(const void *b_full, const void *x_full, const void *g64, double kappa,
                             std::vector<double> &rel, hipStream_t s);

    // out = in - kappa^2 D_oe D_eo in in precision `prec` (tmp: half vector);
    // with a clover term out = A_oo in - kappa^2 D_oe A_ee^{-1} D_eo in
    hipError_t apply_schur(void *out, const void *in, void *tmp, const void *gauge, int prec,
                           double kappa, int dagger, hipStream_t s);

    // Called after every iteration of a plain (FP64) solve with the iteration
    // number and the current FP64 odd iterate x_o (device, [Vh][4][Nc][R]).
    std::function<hipError_t(int, const void *, hipStream_t)> iterate_hook;

    const Geom &geom() const { return g_; }
    int nc() const { return Nc_; }
    int nrhs() const { return R_; }

private:
    hipError_t ensure(int prec);
    // t = A s whose second hop also forms omega's <t, s>, <t, t> and finishes omega into st_
    // (FP16 / BF16, when the tuned dslash is the matrix-core kernel); fused = false: only t
    // was computed and the caller runs the separate reduction
    hipError_t schur_omega(void *out, const void *in, void *tmp, const void *gauge, int prec, double kappa,
                           bool &fused, hipStream_t s);
    // out = A_pp in + coef y (clover term or its inverse of parity p, storage `prec`)
    hipError_t clover_apply(void *out, const void *in, const void *blocks, const void *y, double coef, int prec,
                            hipStream_t s);
    hipError_t dslash(void *out, const void *in, const void *gauge, int prec, int parity,
                      const void *xpay, double coef, int dagger, hipStream_t s);
    // Reliable-update residual from tmp64_ = D_eo x_o: d = b' - (x_o - kappa^2 D_oe tmp64_)
    // (FP64), dn_ = ||d|| per RHS, r_ = dscale_[R + r] d in inner precision `prec`.
    hipError_t residual_lo(const void *x_o, const void *g64, double kappa, int prec, hipStream_t s);

    Geom g_;
    int Nc_, R_;
    long n_;        // complex entries per RHS in a half vector (Vh * 4 * Nc)
    int alloc_prec_ = -1;
    // inner (sloppy) vectors
    void *r0_ = nullptr, *r_ = nullptr, *p_ = nullptr, *v_ = nullptr, *s_ = nullptr, *t_ = nullptr,
         *tmp_ = nullptr, *xlo_ = nullptr, *ph_ = nullptr, *sh_ = nullptr;
    // FP64 vectors
    void *bp_ = nullptr, *tmp64_ = nullptr, *v64_ = nullptr;
    void *ax64_ = nullptr;     // A_oo x_o of the reliable update (clover only)
    const CloverSet *clover_ = nullptr;
    C64 *st_ = nullptr;        // scalar slots [kNumSlots][R]
    C64 *ones_ = nullptr;      // [R] of 1
    C64 *kap_ = nullptr;       // [R] of kappa
    double *dn_ = nullptr;     // [4][R] norms
    double *dscale_ = nullptr; // [2][R] per-RHS scales (accumulation, new inner residual)
    double *h_ = nullptr;      // pinned host staging [8][R]
    double *part_ = nullptr;   // per-workgroup partial sums of the fused residual (mixed solves)
    long part_cap_ = 0;
    double *dotp_ = nullptr;   // per-workgroup partials of the FP16/BF16 Schur hop's fused omega
    long dotp_cap_ = 0;
    ReduceWork red_;
};

}  // namespace qcu
