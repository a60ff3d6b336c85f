// dslash.h -- internal launch interface of the Wilson hopping kernels.
#pragma once

#include <array>
#include <utility>
#include <vector>

#include "qcu_common.h"

namespace qcu {

// Kernel variants of the one-parity hopping term.
enum DslashVariant : int {
    kDslashAuto = 0,      // choose per precision (and from the tuning table)
    kDslashLaneRhs = 1,   // one lane per (site, rhs), all 8 hops per lane
    kDslashWaveSite = 2,  // one wavefront per site: lanes = 16 rhs x 4 dimensions (VALU)
    kDslashMfma = 3,      // one wavefront per site on MFMA (FP16 / BF16 storage)
    kDslashLds = 4        // MFMA with the input rows staged in LDS: one workgroup per 4-site segment of an
                          // x-row sweeping `spw` consecutive t (FP16 / BF16, Nc <= 4); chosen explicitly
};

// Launch shape of the wave-per-site kernels.
struct DslashTune {
    int variant = kDslashAuto;
    int by = 0, bz = 0;   // traversal tile in y and z (0 = choose)
    int spw = 0;          // sites per wavefront (0 = choose)
};

struct DslashArgs {
    void *out;          // MRHS half vector, output parity
    const void *in;     // MRHS half vector, parity 1 - parity
    const void *gauge;  // [4][2][Vh][Nc][Nc], same storage precision as the spinors
    const void *xpay;   // optional: out = xpay + coef * D in (same layout as out)
    double coef;        // scale of D in (1 for the bare operator)
    Geom g;
    int Nc;
    int R;
    int parity;         // output parity
    int dagger;
    int prec;           // storage precision of out/in/gauge/xpay (Prec)
    DslashTune tune{};
};

bool dslash_supported(int Nc, int R);
hipError_t launch_dslash(const DslashArgs &a, hipStream_t stream);

// resolved launch shape (after defaults / tuning table)
DslashTune dslash_resolve(const DslashArgs &a);
// `t` with its unset fields (variant auto, tile or sites per wavefront <= 0)
// replaced by the built-in defaults for a's lattice, Nc, R and precision
DslashTune dslash_fill_defaults(const DslashArgs &a, DslashTune t);
// tuning-table lookup and listing
bool dslash_get_tuned(const int L[4], int Nc, int R, int prec, DslashTune &t);
std::vector<std::pair<std::array<int, 7>, DslashTune>> dslash_tuned_entries();
// per-(lattice, Nc, R, prec) tuning table consulted when tune.variant == auto
void dslash_set_tuned(const int L[4], int Nc, int R, int prec, const DslashTune &t);
// process-wide override (variant != auto forces that shape for every launch)
void dslash_set_override(const DslashTune &t);

// variant launchers (dslash_wave.hip, dslash_mfma.hip)
hipError_t launch_dslash_wave(const DslashArgs &a, const DslashTune &t, hipStream_t stream);
hipError_t launch_dslash_mfma(const DslashArgs &a, const DslashTune &t, hipStream_t stream);
bool dslash_mfma_supported(int Nc, int prec);
hipError_t launch_dslash_lds(const DslashArgs &a, const DslashTune &t, hipStream_t stream);
bool dslash_lds_supported(const Geom &g, int Nc, int prec);
// FP32 / FP64 storage on the FP32 / FP64 matrix cores (dslash_mfma32.hip), also for batched groups
hipError_t launch_dslash_mfma32(const DslashArgs &a, const DslashTune &t, hipStream_t stream);
hipError_t launch_dslash_mfma32_groups(const DslashArgs &a, const SegTab &tab, int by, int bz, hipStream_t stream);
// the groups' RHS packed into shared 16-RHS chunks (groups back to back: base[k] = base[k-1] + R[k-1])
hipError_t launch_dslash_mfma32_packed(const DslashArgs &a, const SegTab &tab, int by, int bz, int spw,
                                       hipStream_t stream);
bool dslash_mfma32_supported(int Nc, int prec);
bool dslash_mfma64_supported(int Nc, int prec);
// a matrix-core kernel exists for (Nc, prec)
inline bool dslash_matrix_supported(int Nc, int prec) {
    return dslash_mfma_supported(Nc, prec) || dslash_mfma32_supported(Nc, prec) || dslash_mfma64_supported(Nc, prec);
}
// the group kernels exist on matrix cores for (Nc, prec): FP32 or FP64 storage
inline bool dslash_groups_matrix_supported(int Nc, int prec) {
    return dslash_mfma32_supported(Nc, prec) || dslash_mfma64_supported(Nc, prec);
}

// Second hop of the FP64 Schur operator fused with the reliable-update residual
// (wave-per-site kernel, dagger 0): with v = a.xpay + a.coef D a.in on parity
// a.parity, writes lo = scale[r] (b - v) in storage lo_prec (single, half, bf16)
// and nrm[r] = ||b - v|| (deterministic: per-workgroup partials summed in a
// fixed order by two small launches).  v itself is not stored.  `partials` holds at least
// dslash_residual_partials(g, R) doubles.
long dslash_residual_partials(const Geom &g, int R);
hipError_t launch_dslash_residual_lo(const DslashArgs &a, const void *b, void *lo, int lo_prec, const double *scale,
                                     double *partials, long partials_cap, double *nrm, hipStream_t stream);

// out[x][c] = sum of in[i][c] over the rows i of slice x (nrows rows of `cols` doubles cut into `slices`
// consecutive slices), fixed summation order
hipError_t launch_rows_sum(double *out, const double *in, int nrows, int cols, int slices, hipStream_t stream);

// Second hop of the unpreconditioned FP16 / BF16 Schur operator, t = s + coef D in with
// a.xpay = s, with omega's inner products <t, s>, <t, t> in its epilogue (matrix-core
// kernel, dagger 0): omega = <t, s> / <t, t> lands in st like launch_step_omega.
// `partials` holds dslash_omega_partials(g, R) doubles.  Returns hipErrorNotSupported
// (and launches nothing) when the tuned launch for these arguments is not the
// matrix-core kernel; the caller then runs the separate reduction.
long dslash_omega_partials(const Geom &g, int R);
hipError_t launch_dslash_omega(const DslashArgs &a, double *partials, long cap, Cx<double> *st, hipStream_t stream);
hipError_t launch_dslash_mfma_omega(const DslashArgs &a, const DslashTune &t, double *partials, hipStream_t stream,
                                    int &nblk);

// One launch of the wave-per-site kernel over every group of a batch (grid.y =
// group): out/in/xpay are batch half vectors (SegTab layout), a.R is ignored,
// the group's sites per wavefront come from tab.spw, the traversal tile from
// (by, bz).  Groups whose done flag is set are skipped.
hipError_t launch_dslash_wave_groups(const DslashArgs &a, const SegTab &tab, int by, int bz, hipStream_t stream);

}  // namespace qcu
