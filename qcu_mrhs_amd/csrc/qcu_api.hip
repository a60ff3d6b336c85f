// qcu_api.hip -- the C ABI (include/qcu.h, include/qcu_ops.h) over a
// process-wide context, mirroring the reference's facade qcu::Qcu
// (include/qcu_interface.h:12-113, src/qcu_interface.cu:22-413) and its global
// singleton (src/interface.cu:12-18).
//
// Differences from the reference, by design:
//  * one GPU: the process grid must be 1x1x1x1 and no MPI is needed
//    (the reference asserts MPI_Comm_size == Gx*Gy*Gz*Gt, src/qcu_config/qcu_config.cu:76-86);
//  * misuse sets an error code (qcuLastError) instead of aborting, unless
//    QCU_ABORT_ON_ERROR=1;
//  * the R fermion pointers travel by value in kernel arguments (the
//    reference copies a device pointer table per call, src/qcu_interface.cu:123-126);
//  * the solve starts from x0 = 0 (the reference's memset is a no-op,
//    src/solver/bicgstab_solve.cu:243) and every call is stream ordered on one
//    library stream with a single synchronisation at the end.
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "clover.h"
#include "groups.h"
#include "layout.h"
#include "lgnn.h"
#include "qcu.h"
#include "qcu_ops.h"
#include "solver.h"
#include "tune.h"

namespace qcu {

namespace {
int g_err = 0;
char g_msg[1024] = "no error";
}  // namespace

void set_error(int code, const char *fmt, ...) {
    g_err = code;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
    va_end(ap);
    fprintf(stderr, "[qcu] error %d: %s\n", code, g_msg);
    const char *ab = getenv("QCU_ABORT_ON_ERROR");
    if (ab && ab[0] == '1') abort();
}

int last_error() { return g_err; }

struct Context {
    Geom g{};
    int Nc = 3, R = 1;
    int user_prec = kDouble, iter_prec = kDouble;
    double mass = 0.0, kappa = 0.125;
    bool have_dslash = false;
    // Wilson-clover operator (getDslash type kQcuDslashClover)
    bool clover = false;
    double csw = 1.0;
    bool clover_dirty = true;        // gauge, kappa or c_sw changed since the last build
    void *clov64 = nullptr, *cinv64 = nullptr, *clov32 = nullptr, *cinv32 = nullptr;
    void *cnorm = nullptr;           // one double: norm of the inverse (singularity check)
    CloverSet clover_set;
    void *gauge64 = nullptr;
    long long gauge_gen = 0;  // loads into gauge64 so far (copies derived from it compare this)
    void *gauge_lo[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    void *bufA = nullptr, *bufB = nullptr;  // full-lattice MRHS vectors, FP64-sized
    std::vector<void *> q_in, q_out;
    std::unique_ptr<BiCGStab> solver;
    std::unique_ptr<GroupSolver> groups;  // batched uneven RHS groups (qcuInvertGroups)
    SolverOptions opts;
    std::unique_ptr<LgnnPreconditioner> precond;
    struct LgnnBatch {
        int cfg;
        void *e, *r;  // FP32 odd half vectors
    };
    std::vector<void *> lgnn_gauges;  // FP64 copies, one per collected configuration
    std::vector<LgnnBatch> lgnn_batches;
    int lgnn_power = 0;  // training objective: 0 = (error, residual) pairs, K >= 1 = residual power K
    bool lgnn_projected = false;  // second-layer weights kept in the image of each hop's Wilson projector
    hipStream_t stream = nullptr;

    void clear_lgnn_data() {
        for (auto &b : lgnn_batches) {
            (void)hipFree(b.e);
            (void)hipFree(b.r);
        }
        for (void *p : lgnn_gauges) (void)hipFree(p);
        lgnn_batches.clear();
        lgnn_gauges.clear();
    }

    long half_n() const { return long(g.Vh) * 4 * Nc; }
    long gauge_count() const { return 4L * 2 * g.Vh * Nc * Nc; }

    ~Context() {
        clear_lgnn_data();
        groups.reset();
        solver.reset();
        precond.reset();
        for (void *p : {gauge64, bufA, bufB, clov64, cinv64, clov32, cinv32, cnorm}) if (p) (void)hipFree(p);
        for (void *p : gauge_lo) if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // gauge field in storage precision `prec` (converted lazily from FP64)
    const void *gauge(int prec) {
        if (prec == kDouble) return gauge64;
        if (!gauge_lo[prec]) {
            if (hipMalloc(&gauge_lo[prec], size_t(gauge_count()) * 2 * prec_bytes(prec)) != hipSuccess) {
                set_error(kQcuErrorDevice, "cannot allocate gauge copy");
                gauge_lo[prec] = nullptr;
                return nullptr;
            }
            if (launch_convert(gauge_lo[prec], prec, gauge64, kDouble, gauge_count(), nullptr, 1,
                               stream) != hipSuccess) {
                set_error(kQcuErrorDevice, "gauge conversion failed");
                return nullptr;
            }
        }
        return gauge_lo[prec];
    }
};

Context *g_ctx = nullptr;

// The clover term of the current (gauge, kappa, c_sw), built on first use:
// FP64 blocks and inverse for both parities, FP32 copies for the inner solves.
bool hip_ok(hipError_t e, const char *what);
ReduceWork g_ops_red;
bool ops_red_ready();

bool ensure_clover(Context *c) {
    if (!c->clover) {
        c->solver->set_clover(nullptr);
        return true;
    }
    if (c->clover_dirty) {
        const int n = 2 * c->Nc;
        const long nblocks = 2L * c->g.Vh * 2;
        const size_t cnt = size_t(nblocks) * n * n;
        if (!c->clov64) {
            if (!hip_ok(hipMalloc(&c->clov64, cnt * sizeof(Cx<double>)), "clover alloc") ||
                !hip_ok(hipMalloc(&c->cinv64, cnt * sizeof(Cx<double>)), "clover alloc") ||
                !hip_ok(hipMalloc(&c->clov32, cnt * sizeof(Cx<float>)), "clover alloc") ||
                !hip_ok(hipMalloc(&c->cinv32, cnt * sizeof(Cx<float>)), "clover alloc") ||
                !hip_ok(hipMalloc(&c->cnorm, sizeof(double)), "clover alloc"))
                return false;
        }
        bool ok = hip_ok(launch_clover_build(c->clov64, c->gauge64, c->g, c->Nc, c->csw, c->kappa, c->stream),
                         "clover build");
        ok = ok && hip_ok(launch_clover_invert(c->cinv64, c->clov64, nblocks, n, c->stream), "clover inverse");
        ok = ok && hip_ok(launch_convert(c->clov32, kSingle, c->clov64, kDouble, long(cnt), nullptr, 1, c->stream),
                          "clover copy");
        ok = ok && hip_ok(launch_convert(c->cinv32, kSingle, c->cinv64, kDouble, long(cnt), nullptr, 1, c->stream),
                          "clover copy");
        // a singular block (kappa c_sw |F| of order 1) leaves non-finite entries in the inverse
        double h = 0.0;
        ok = ok && ops_red_ready() &&
             hip_ok(launch_norm(static_cast<double *>(c->cnorm), c->cinv64, long(cnt), 1, kDouble, false, g_ops_red, 0,
                                c->stream), "clover check") &&
             hip_ok(hipMemcpyAsync(&h, c->cnorm, sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy") &&
             hip_ok(hipStreamSynchronize(c->stream), "clover");
        if (!ok) return false;
        if (!std::isfinite(h)) {
            set_error(kQcuErrorInvalidValue, "clover term is singular at kappa %.6g, c_sw %.6g", c->kappa, c->csw);
            return false;
        }
        c->clover_set.a64 = c->clov64;
        c->clover_set.inv64 = c->cinv64;
        c->clover_set.a32 = c->clov32;
        c->clover_set.inv32 = c->cinv32;
        c->clover_set.parity_stride = long(c->g.Vh) * 2 * n * n;
        c->clover_dirty = false;
    }
    c->solver->set_clover(&c->clover_set);
    return true;
}

// entry points whose kernels model the Wilson operator only
bool wilson_only(Context *c, const char *what) {
    if (!c->clover) return true;
    set_error(kQcuErrorUnsupported, "%s supports the Wilson operator only (getDslash type %d)", what,
              kQcuDslashWilson);
    return false;
}

Context *ctx_or_error() {
    if (!g_ctx) set_error(kQcuErrorNotInitialized, "Qcu is not initialized (call initGridSize first)");
    return g_ctx;
}

bool valid_storage(int p) { return p == kHalf || p == kSingle || p == kDouble || p == kBF16; }

bool queue_full(Context *c) {
    if (int(c->q_in.size()) != c->R || int(c->q_out.size()) != c->R) {
        set_error(kQcuErrorQueue, "fermion queue holds %zu pairs, expected %d", c->q_in.size(), c->R);
        return false;
    }
    return true;
}

PtrTable table(const std::vector<void *> &v, long offset_elems, int prec) {
    PtrTable t{};
    const size_t eb = 2 * prec_bytes(prec);
    for (size_t i = 0; i < v.size(); ++i)
        t.p[i] = static_cast<const char *>(v[i]) + offset_elems * eb;
    return t;
}

bool hip_ok(hipError_t e, const char *what) {
    if (e == hipSuccess) return true;
    if (last_error() == 0 || last_error() == kQcuErrorNotConverged)
        set_error(kQcuErrorDevice, "%s: %s", what, hipGetErrorString(e));
    return false;
}

// stateless ops: one workspace for reductions up to 256 RHS
bool ops_red_ready() {
    if (g_ops_red.partials) return true;
    return hip_ok(reduce_work_alloc(g_ops_red, 256), "reduction workspace");
}

}  // namespace qcu

using namespace qcu;

extern "C" {

int qcuLastError(void) { return g_err; }
const char *qcuLastErrorString(void) { return g_msg; }
void qcuClearError(void) {
    g_err = 0;
    strcpy(g_msg, "no error");
}

void initGridSize(QcuGrid *grid, QcuParam *param, int n_color, int m_rhs, int inputFloatPrecision,
                  int dslashFloatPrecision) {
    if (!grid || !param) return set_error(kQcuErrorInvalidValue, "null grid/param");
    for (int i = 0; i < 4; ++i)
        if (grid->grid_size[i] != 1)
            return set_error(kQcuErrorUnsupported, "process grid must be 1x1x1x1 (one GPU)");
    int L[4];
    for (int i = 0; i < 4; ++i) {
        L[i] = param->lattice_size[i];
        if (L[i] <= 0) return set_error(kQcuErrorInvalidValue, "lattice size must be positive");
    }
    if (L[0] % 2) return set_error(kQcuErrorInvalidValue, "Lx must be even");
    if (!dslash_supported(n_color, m_rhs) || m_rhs > kMaxPtrs)
        return set_error(kQcuErrorUnsupported, "n_color %d / m_rhs %d not supported (Nc in {2,3,4,8}, 1<=R<=%d)",
                         n_color, m_rhs, kMaxPtrs);
    if (!(inputFloatPrecision == kSingle || inputFloatPrecision == kDouble))
        return set_error(kQcuErrorUnsupported, "fermion precision %d not supported (Single or Double)",
                         inputFloatPrecision);
    if (!valid_storage(dslashFloatPrecision))
        return set_error(kQcuErrorUnsupported, "dslash precision %d not supported", dslashFloatPrecision);
    delete g_ctx;
    g_ctx = nullptr;
    auto *c = new Context();
    c->g = make_geom(L);
    c->Nc = n_color;
    c->R = m_rhs;
    c->user_prec = inputFloatPrecision;
    c->iter_prec = dslashFloatPrecision;
    c->opts.inner_prec = dslashFloatPrecision;
    const size_t full = size_t(2) * c->half_n() * m_rhs * sizeof(Cx<double>);
    if (!hip_ok(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "stream") ||
        !hip_ok(hipMalloc(&c->bufA, full), "alloc") || !hip_ok(hipMalloc(&c->bufB, full), "alloc") ||
        !hip_ok(hipMalloc(&c->gauge64, size_t(c->gauge_count()) * sizeof(Cx<double>)), "alloc")) {
        delete c;
        return;
    }
    c->solver = std::make_unique<BiCGStab>(c->g, n_color, m_rhs);
    g_ctx = c;
}

void pushBackFermions(void *fermionOut, void *fermionIn) {
    Context *c = ctx_or_error();
    if (!c) return;
    if (int(c->q_in.size()) >= c->R) return set_error(kQcuErrorQueue, "Fermion queue is full");
    c->q_in.push_back(fermionIn);
    c->q_out.push_back(fermionOut);
}

void loadQcuGauge(void *gauge, int floatPrecision) {
    Context *c = ctx_or_error();
    if (!c) return;
    if (!valid_storage(floatPrecision)) return set_error(kQcuErrorUnsupported, "gauge precision %d", floatPrecision);
    for (auto &p : c->gauge_lo)
        if (p) {
            (void)hipFree(p);
            p = nullptr;
        }
    c->clover_dirty = true;
    ++c->gauge_gen;  // before the copy: a failed load leaves no derived copy marked current
    if (!hip_ok(launch_convert(c->gauge64, kDouble, gauge, floatPrecision, c->gauge_count(), nullptr, 1,
                               c->stream), "loadQcuGauge"))
        return;
    (void)hip_ok(hipStreamSynchronize(c->stream), "loadQcuGauge");
}

void getDslash(int dslashType, double mass) {
    Context *c = ctx_or_error();
    if (!c) return;
    if (dslashType != kQcuDslashWilson && dslashType != kQcuDslashClover)
        return set_error(kQcuErrorUnsupported, "Unsupported dslash type %d", dslashType);
    if (dslashType == kQcuDslashClover && !clover_supported(c->Nc))
        return set_error(kQcuErrorUnsupported, "clover term: n_color %d not supported (2, 3, 4, 8)", c->Nc);
    c->clover = dslashType == kQcuDslashClover;
    c->clover_dirty = true;
    c->mass = mass;
    c->kappa = 1.0 / (2.0 * (4.0 + mass));
    c->have_dslash = true;
}

void start_dslash(int parity, int daggerFlag) {
    Context *c = ctx_or_error();
    if (!c) return;
    if (!c->have_dslash) return set_error(kQcuErrorNotInitialized, "Dslash is not initialized (call getDslash)");
    if (!queue_full(c)) return;
    if (parity != 0 && parity != 1) return set_error(kQcuErrorInvalidValue, "parity must be 0 or 1");
    const int P = c->iter_prec;
    const long n = c->half_n();
    bool ok = hip_ok(launch_gather(c->bufA, P, table(c->q_in, 0, c->user_prec), c->user_prec, n, c->R,
                                   c->stream), "gather");
    DslashArgs a{c->bufB, c->bufA, c->gauge(P), nullptr, 1.0, c->g, c->Nc, c->R, parity, daggerFlag ? 1 : 0, P};
    ok = ok && a.gauge && hip_ok(launch_dslash(a, c->stream), "dslash");
    ok = ok && hip_ok(launch_scatter(table(c->q_out, 0, c->user_prec), c->user_prec, c->bufB, P, n, c->R,
                                     c->stream), "scatter");
    ok = ok && hip_ok(hipStreamSynchronize(c->stream), "start_dslash");
    c->q_in.clear();
    c->q_out.clear();
}

void mat_Qcu(int daggerFlag) {
    Context *c = ctx_or_error();
    if (!c) return;
    if (!c->have_dslash) return set_error(kQcuErrorNotInitialized, "Dslash is not initialized (call getDslash)");
    if (!queue_full(c)) return;
    const int P = c->iter_prec;
    const long n = c->half_n();
    const size_t eb = 2 * prec_bytes(P);
    char *A = static_cast<char *>(c->bufA), *B = static_cast<char *>(c->bufB);
    bool ok = hip_ok(launch_gather(A, P, table(c->q_in, 0, c->user_prec), c->user_prec, 2 * n, c->R,
                                   c->stream), "gather");
    ok = ok && ensure_clover(c);
    for (int p = 0; p < 2 && ok; ++p) {
        char *in_p = A + p * n * c->R * eb, *out_p = B + p * n * c->R * eb;
        if (c->clover) {
            // out_p = A_pp in_p - kappa D_p in_{1-p}  (A Hermitian: the dagger reaches D only)
            DslashArgs a{out_p, A + (1 - p) * n * c->R * eb, c->gauge(P), nullptr, 1.0, c->g, c->Nc, c->R, p,
                         daggerFlag ? 1 : 0, P};
            ok = a.gauge && hip_ok(launch_dslash(a, c->stream), "dslash");
            ok = ok && hip_ok(launch_clover_apply(out_p, in_p, c->clover_set.term(P, p), out_p, -c->kappa, c->g.Vh,
                                                  c->Nc, c->R, P, CloverSet::storage(P), c->stream),
                              "clover");
            continue;
        }
        // out_p = in_p - kappa D_p in_{1-p}   (src/qcu_interface.cu:174-255)
        DslashArgs a{out_p, A + (1 - p) * n * c->R * eb, c->gauge(P), in_p,
                     -c->kappa, c->g, c->Nc, c->R, p, daggerFlag ? 1 : 0, P};
        ok = a.gauge && hip_ok(launch_dslash(a, c->stream), "dslash");
    }
    ok = ok && hip_ok(launch_scatter(table(c->q_out, 0, c->user_prec), c->user_prec, B, P, 2 * n, c->R,
                                     c->stream), "scatter");
    ok = ok && hip_ok(hipStreamSynchronize(c->stream), "mat_Qcu");
    c->q_in.clear();
    c->q_out.clear();
}

int qcuInvertEx(int max_iteration, double tol, QcuSolveInfo *info) {
    Context *c = ctx_or_error();
    if (!c) return kQcuErrorNotInitialized;
    if (!c->have_dslash) {
        set_error(kQcuErrorNotInitialized, "Dslash is not initialized (call getDslash)");
        return g_err;
    }
    if (!queue_full(c)) return g_err;
    const auto t0 = std::chrono::steady_clock::now();
    const long n = c->half_n();
    SolveResult res;
    GaugeSet gs;
    gs.g64 = c->gauge64;
    if (c->opts.inner_prec != kDouble) {
        gs.glo = c->gauge(c->opts.inner_prec);
        gs.lo_prec = c->opts.inner_prec;
    }
    if (!ensure_clover(c) || (c->opts.use_precond && !wilson_only(c, "the learned preconditioner"))) {
        c->q_in.clear();
        c->q_out.clear();
        return g_err;
    }
    Preconditioner *pc = nullptr;
    if (c->opts.use_precond) {
        if (!c->precond || !c->precond->ready()) {
            set_error(kQcuErrorNotInitialized, "no learned preconditioner loaded");
            c->q_in.clear();
            c->q_out.clear();
            return g_err;
        }
        pc = c->precond.get();
        if (!hip_ok(c->precond->set_gauge(c->gauge64, c->stream, c->gauge_gen), "preconditioner gauge")) {
            c->q_in.clear();
            c->q_out.clear();
            return g_err;
        }
    }
    bool ok = hip_ok(launch_gather(c->bufA, kDouble, table(c->q_in, 0, c->user_prec), c->user_prec, 2 * n,
                                   c->R, c->stream), "gather");
    int code = kQcuSuccess;
    int fallback = 0;
    if (ok) {
        hipError_t e = c->solver->solve(c->bufA, c->bufB, gs, c->kappa, max_iteration, tol, c->opts, pc, res,
                                        c->stream);
        if (pc && (e == hipErrorUnknown || (e == hipSuccess && !res.converged))) {
            // the learned preconditioner failed (breakdown, divergence or no convergence): plain
            // BiCGStab from x0 = 0 on the same system decides the result
            if (g_err == kQcuErrorBreakdown) qcuClearError();
            const int used = res.iterations;
            e = c->solver->solve(c->bufA, c->bufB, gs, c->kappa, max_iteration, tol, c->opts, nullptr, res,
                                 c->stream);
            res.iterations += used;
            res.restarts += 1;
            fallback = 1;
        }
        if (e != hipSuccess) {
            ok = false;
            code = g_err ? g_err : kQcuErrorDevice;
        }
    }
    ok = ok && hip_ok(launch_scatter(table(c->q_out, 0, c->user_prec), c->user_prec, c->bufB, kDouble, 2 * n,
                                     c->R, c->stream), "scatter");
    ok = ok && hip_ok(hipStreamSynchronize(c->stream), "qcuInvert");
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    c->q_in.clear();
    c->q_out.clear();
    if (!ok && code == kQcuSuccess) code = g_err ? g_err : kQcuErrorDevice;
    if (ok && !res.converged) {
        set_error(kQcuErrorNotConverged, "BiCGStab did not converge in %d iterations (max rel residual %.3e)",
                  res.iterations, res.true_rel_residual);
        code = kQcuErrorNotConverged;
    }
    if (info) {
        info->converged = res.converged;
        info->iterations = res.iterations;
        info->reliable_updates = res.reliable_updates;
        info->restarts = res.restarts;
        info->true_rel_residual = res.true_rel_residual;
        info->seconds = secs;
        info->preconditioned = pc && !fallback ? 1 : 0;
        info->fallback = fallback;
    }
    return code;
}

void qcuInvert(int max_iteration, double p_max_prec) { (void)qcuInvertEx(max_iteration, p_max_prec, nullptr); }

int qcuSetCloverCoefficient(double csw) {
    Context *c = ctx_or_error();
    if (!c) return kQcuErrorNotInitialized;
    if (!std::isfinite(csw)) {
        set_error(kQcuErrorInvalidValue, "c_sw must be finite");
        return g_err;
    }
    c->csw = csw;
    c->clover_dirty = true;
    return kQcuSuccess;
}

int qcuSetSolverOptions(int inner_precision, double delta, int use_preconditioner) {
    Context *c = ctx_or_error();
    if (!c) return kQcuErrorNotInitialized;
    if (!valid_storage(inner_precision)) {
        set_error(kQcuErrorUnsupported, "inner precision %d", inner_precision);
        return g_err;
    }
    if (!(delta >= 0.0 && delta < 1.0)) {
        set_error(kQcuErrorInvalidValue, "delta must be in [0, 1) (0 = default for the inner precision)");
        return g_err;
    }
    c->opts.inner_prec = inner_precision;
    c->opts.delta = delta;
    c->opts.use_precond = use_preconditioner ? 1 : 0;
    const char *v = getenv("QCU_VERBOSE");
    c->opts.verbose = (v && v[0] == '1') ? 1 : 0;
    return kQcuSuccess;
}

void finalizeQcu(void) {
    delete g_ctx;
    g_ctx = nullptr;
    reduce_work_free(g_ops_red);
}

void gauge_eo_precondition(void *prec_gauge, void *non_prec_gauge, int precision) {
    Context *c = ctx_or_error();
    if (!c) return;
    if (prec_gauge == non_prec_gauge) return set_error(kQcuErrorInvalidValue, "in-place conversion not supported");
    if (!valid_storage(precision)) return set_error(kQcuErrorUnsupported, "UNDEFINED precision");
    if (hip_ok(launch_gauge_eo(prec_gauge, non_prec_gauge, c->g, c->Nc, precision, 0, c->stream), "eo"))
        (void)hip_ok(hipStreamSynchronize(c->stream), "eo");
}

void gauge_reverse_eo_precondition(void *non_prec_gauge, void *prec_gauge, int precision) {
    Context *c = ctx_or_error();
    if (!c) return;
    if (prec_gauge == non_prec_gauge) return set_error(kQcuErrorInvalidValue, "in-place conversion not supported");
    if (!valid_storage(precision)) return set_error(kQcuErrorUnsupported, "UNDEFINED precision");
    if (hip_ok(launch_gauge_eo(prec_gauge, non_prec_gauge, c->g, c->Nc, precision, 1, c->stream), "eo"))
        (void)hip_ok(hipStreamSynchronize(c->stream), "eo");
}

void read_gauge_from_file(void *gauge, const char *file_path_prefix) {
    Context *c = ctx_or_error();
    if (!c) return;
    if (!file_path_prefix) return set_error(kQcuErrorInvalidValue, "null path");
    FILE *f = fopen(file_path_prefix, "rb");
    if (!f) return set_error(kQcuErrorInvalidValue, "cannot open %s", file_path_prefix);
    const size_t count = size_t(c->gauge_count());
    std::vector<Cx<double>> host(count);
    const size_t got = fread(host.data(), sizeof(Cx<double>), count, f);
    fclose(f);
    if (got != count)
        return set_error(kQcuErrorInvalidValue, "%s holds %zu complex values, expected %zu", file_path_prefix, got,
                         count);
    void *lex = nullptr;
    if (!hip_ok(hipMalloc(&lex, count * sizeof(Cx<double>)), "alloc")) return;
    bool ok = hip_ok(hipMemcpy(lex, host.data(), count * sizeof(Cx<double>), hipMemcpyHostToDevice), "copy");
    ok = ok && hip_ok(launch_gauge_eo(gauge, lex, c->g, c->Nc, kDouble, 0, c->stream), "eo");
    ok = ok && hip_ok(hipStreamSynchronize(c->stream), "eo");
    (void)hipFree(lex);
}

// ---- device memory plumbing ---------------------------------------------
void *qcuDeviceMalloc(size_t bytes) {
    void *p = nullptr;
    if (!hip_ok(hipMalloc(&p, bytes), "qcuDeviceMalloc")) return nullptr;
    return p;
}
int qcuDeviceFree(void *ptr) { return hip_ok(hipFree(ptr), "qcuDeviceFree") ? 0 : kQcuErrorDevice; }
int qcuMemcpyHtoD(void *dst, const void *src, size_t bytes) {
    return hip_ok(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "HtoD") ? 0 : kQcuErrorDevice;
}
int qcuMemcpyDtoH(void *dst, const void *src, size_t bytes) {
    return hip_ok(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "DtoH") ? 0 : kQcuErrorDevice;
}
int qcuMemsetD(void *dst, int value, size_t bytes) {
    return hip_ok(hipMemset(dst, value, bytes), "memset") ? 0 : kQcuErrorDevice;
}
int qcuDeviceSynchronize(void) { return hip_ok(hipDeviceSynchronize(), "sync") ? 0 : kQcuErrorDevice; }

// ---- stateless operations (qcu_ops.h) -----------------------------------
int qcuDslashMrhs(void *out, const void *in, const void *gauge, const int lattice[4], int n_color, int m_rhs,
                  int parity, int dagger, int precision, const void *xpay, double coef) {
    if (!lattice || lattice[0] % 2 || !dslash_supported(n_color, m_rhs) || !valid_storage(precision) ||
        (parity != 0 && parity != 1)) {
        set_error(kQcuErrorInvalidValue, "qcuDslashMrhs: invalid arguments");
        return g_err;
    }
    DslashArgs a{out, in, gauge, xpay, coef, make_geom(lattice), n_color, m_rhs, parity, dagger ? 1 : 0, precision};
    return hip_ok(launch_dslash(a, nullptr), "qcuDslashMrhs") ? 0 : g_err;
}

int qcuCloverBuild(void *clover, void *clover_inv, const void *gauge, const int lattice[4], int n_color,
                   double csw, double kappa) {
    if (!clover || !gauge || !lattice || lattice[0] % 2 || !clover_supported(n_color) || !std::isfinite(csw) ||
        !std::isfinite(kappa)) {
        set_error(kQcuErrorInvalidValue, "qcuCloverBuild: invalid arguments");
        return g_err;
    }
    for (int i = 0; i < 4; ++i)
        if (lattice[i] <= 0) {
            set_error(kQcuErrorInvalidValue, "qcuCloverBuild: invalid lattice");
            return g_err;
        }
    const Geom g = make_geom(lattice);
    bool ok = hip_ok(launch_clover_build(clover, gauge, g, n_color, csw, kappa, nullptr), "qcuCloverBuild");
    if (ok && clover_inv)
        ok = hip_ok(launch_clover_invert(clover_inv, clover, 2L * g.Vh * 2, 2 * n_color, nullptr), "qcuCloverBuild");
    return ok && hip_ok(hipStreamSynchronize(nullptr), "qcuCloverBuild") ? 0 : g_err;
}

int qcuCloverApply(void *out, const void *in, const void *blocks, const int lattice[4], int n_color, int m_rhs,
                   int precision, int clover_precision, const void *xpay, double coef) {
    if (!out || !in || !blocks || !lattice || lattice[0] % 2 || !clover_supported(n_color) || m_rhs < 1 ||
        m_rhs > 512 || !valid_storage(precision) || (clover_precision != kDouble && clover_precision != kSingle)) {
        set_error(kQcuErrorInvalidValue, "qcuCloverApply: invalid arguments");
        return g_err;
    }
    const Geom g = make_geom(lattice);
    return hip_ok(launch_clover_apply(out, in, blocks, xpay, coef, g.Vh, n_color, m_rhs, precision,
                                      clover_precision, nullptr), "qcuCloverApply") ? 0 : g_err;
}

static int blas_axpy(int kind, void *res, const void *x, const void *y, const void *z, const void *a,
                     const void *b, const void *cc, long n, int R, int prec) {
    if (!valid_storage(prec) || R < 1 || R > 256 || n < 0) {
        set_error(kQcuErrorInvalidValue, "blas: invalid arguments");
        return g_err;
    }
    AxpyArgs ax{kind, res, x, y, z, static_cast<const C64 *>(a), static_cast<const C64 *>(b),
                static_cast<const C64 *>(cc), n, R, prec};
    return hip_ok(launch_axpy(ax, nullptr), "blas") ? 0 : g_err;
}

int qcuBlasXsay(void *res, const void *x, const void *a, const void *y, long n, int m_rhs, int precision) {
    return blas_axpy(kXsay, res, x, y, nullptr, a, nullptr, nullptr, n, m_rhs, precision);
}
int qcuBlasXpay(void *res, const void *x, const void *a, const void *y, long n, int m_rhs, int precision) {
    return blas_axpy(kXpay, res, x, y, nullptr, a, nullptr, nullptr, n, m_rhs, precision);
}
int qcuBlasAxpby(void *res, const void *a, const void *x, const void *b, const void *y, long n, int m_rhs,
                 int precision) {
    return blas_axpy(kAxpby, res, x, y, nullptr, a, b, nullptr, n, m_rhs, precision);
}
int qcuBlasAxpbypcz(void *res, const void *a, const void *x, const void *b, const void *y, const void *c,
                    const void *z, long n, int m_rhs, int precision) {
    return blas_axpy(kAxpbypcz, res, x, y, z, a, b, c, n, m_rhs, precision);
}
int qcuBlasDotc(void *res, const void *a, const void *b, long n, int m_rhs, int precision) {
    if (!valid_storage(precision) || m_rhs < 1 || m_rhs > 256 || !ops_red_ready()) {
        set_error(kQcuErrorInvalidValue, "qcuBlasDotc: invalid arguments");
        return g_err;
    }
    return hip_ok(launch_dotc(static_cast<C64 *>(res), a, b, n, m_rhs, precision, g_ops_red, 14, nullptr), "dotc")
               ? 0 : g_err;
}
int qcuBlasDotcNoAlloc(void *res, const void *a, const void *b, long n, int m_rhs, int precision) {
    if (!valid_storage(precision) || !res || !a || !b || n < 0 || m_rhs < 1) {
        set_error(kQcuErrorInvalidValue, "qcuBlasDotcNoAlloc: invalid arguments");
        return g_err;
    }
    if (m_rhs > kReduceThreads) {
        set_error(kQcuErrorInvalidValue,
                  "qcuBlasDotcNoAlloc: m_rhs %d exceeds the maximum of %d (fixed %ld-byte device scratch)", m_rhs,
                  kReduceThreads, kNoAllocScratchBytes);
        return g_err;
    }
    return hip_ok(launch_dotc_noalloc(static_cast<C64 *>(res), a, b, n, m_rhs, precision, nullptr),
                  "dotc (no allocation)") ? 0 : g_err;
}
int qcuBlasNorm(void *res, const void *a, long n, int m_rhs, int precision) {
    if (!valid_storage(precision) || m_rhs < 1 || m_rhs > 256 || !ops_red_ready()) {
        set_error(kQcuErrorInvalidValue, "qcuBlasNorm: invalid arguments");
        return g_err;
    }
    return hip_ok(launch_norm(static_cast<double *>(res), a, n, m_rhs, precision, true, g_ops_red, 15, nullptr),
                  "norm") ? 0 : g_err;
}

int qcuColorSpinorGather(void *dst, int dst_precision, void *const *src_ptrs, int src_precision, long nsites,
                         int n_color, int m_rhs) {
    if (!src_ptrs || m_rhs < 1 || m_rhs > kMaxPtrs || !valid_storage(dst_precision) ||
        !valid_storage(src_precision)) {
        set_error(kQcuErrorInvalidValue, "gather: invalid arguments");
        return g_err;
    }
    PtrTable t{};
    for (int i = 0; i < m_rhs; ++i) t.p[i] = src_ptrs[i];
    return hip_ok(launch_gather(dst, dst_precision, t, src_precision, nsites * 4 * n_color, m_rhs, nullptr),
                  "gather") ? 0 : g_err;
}
int qcuColorSpinorScatter(void *const *dst_ptrs, int dst_precision, const void *src, int src_precision,
                          long nsites, int n_color, int m_rhs) {
    if (!dst_ptrs || m_rhs < 1 || m_rhs > kMaxPtrs || !valid_storage(dst_precision) ||
        !valid_storage(src_precision)) {
        set_error(kQcuErrorInvalidValue, "scatter: invalid arguments");
        return g_err;
    }
    PtrTable t{};
    for (int i = 0; i < m_rhs; ++i) t.p[i] = dst_ptrs[i];
    return hip_ok(launch_scatter(t, dst_precision, src, src_precision, nsites * 4 * n_color, m_rhs, nullptr),
                  "scatter") ? 0 : g_err;
}
int qcuDslashTune(int variant, int by, int bz, int spw) {
    if (variant < kDslashAuto || variant > kDslashLds || by < 0 || bz < 0 || spw < -64 ||
        (spw < 0 && variant != kDslashMfma)) {
        set_error(kQcuErrorInvalidValue, "qcuDslashTune: invalid arguments");
        return g_err;
    }
    DslashTune t;
    t.variant = variant;
    t.by = by;
    t.bz = bz;
    t.spw = spw;
    dslash_set_override(t);
    return 0;
}

int qcuDslashAutotune(const int lattice[4], int n_color, int m_rhs, int precision, int repeats, int best[4],
                      double *best_us, double *records, int max_records, int *n_records) {
    if (!lattice || lattice[0] % 2 || !dslash_supported(n_color, m_rhs) || !valid_storage(precision) || repeats < 1) {
        set_error(kQcuErrorInvalidValue, "qcuDslashAutotune: invalid arguments");
        return g_err;
    }
    DslashTune t;
    double us = 0.0;
    std::vector<TuneRecord> log;
    if (!hip_ok(dslash_autotune(lattice, n_color, m_rhs, precision, repeats, t, us, &log), "autotune")) return g_err;
    if (best) {
        best[0] = t.variant;
        best[1] = t.by;
        best[2] = t.bz;
        best[3] = t.spw;
    }
    if (best_us) *best_us = us;
    int n = 0;
    for (const auto &r : log) {
        if (records && n < max_records) {
            double *o = records + 5 * n;
            o[0] = r.tune.variant;
            o[1] = r.tune.by;
            o[2] = r.tune.bz;
            o[3] = r.tune.spw;
            o[4] = r.us;
        }
        ++n;
    }
    if (n_records) *n_records = n;
    return 0;
}

int qcuDslashTuneSave(const char *path) {
    if (!path) {
        set_error(kQcuErrorInvalidValue, "qcuDslashTuneSave: null path");
        return g_err;
    }
    if (dslash_tune_save(path) != hipSuccess) {
        set_error(kQcuErrorInvalidValue, "qcuDslashTuneSave: cannot write %s", path);
        return g_err;
    }
    return 0;
}

int qcuDslashTuneLoad(const char *path, int *n_loaded) {
    int n = 0;
    if (!path) {
        set_error(kQcuErrorInvalidValue, "qcuDslashTuneLoad: null path");
        return g_err;
    }
    const hipError_t e = dslash_tune_load(path, n);
    if (n_loaded) *n_loaded = n;
    if (e != hipSuccess) {
        set_error(kQcuErrorInvalidValue, "qcuDslashTuneLoad: cannot read %s or malformed entry after %d entries", path,
                  n);
        return g_err;
    }
    return 0;
}

int qcuDslashSetTuned(const int lattice[4], int n_color, int m_rhs, int precision, int variant, int by, int bz,
                      int spw) {
    if (!lattice || variant < kDslashLaneRhs || variant > kDslashLds || by < 0 || bz < 0 || spw < -64 ||
        (spw < 0 && variant != kDslashMfma)) {
        set_error(kQcuErrorInvalidValue, "qcuDslashSetTuned: invalid arguments");
        return g_err;
    }
    DslashTune t;
    t.variant = variant;
    t.by = by;
    t.bz = bz;
    t.spw = spw;
    dslash_set_tuned(lattice, n_color, m_rhs, precision, t);
    return 0;
}

// ---- batched uneven RHS groups ------------------------------------------
int qcuInvertGroups(int n_groups, const QcuRhsGroup *groups, int max_iteration, double tol, QcuGroupInfo *info) {
    Context *c = ctx_or_error();
    if (!c) return g_err;
    if (!c->have_dslash) {
        set_error(kQcuErrorNotInitialized, "Dslash is not initialized (call getDslash)");
        return g_err;
    }
    if (!wilson_only(c, "qcuInvertGroups")) return g_err;
    if (n_groups < 1 || n_groups > kMaxGroups || !groups || max_iteration < 1 || !(tol > 0.0)) {
        set_error(kQcuErrorInvalidValue, "qcuInvertGroups: 1..%d groups, max_iteration >= 1, tol > 0", kMaxGroups);
        return g_err;
    }
    std::vector<int> Rs;
    for (int k = 0; k < n_groups; ++k) {
        const QcuRhsGroup &gr = groups[k];
        if (gr.m_rhs < 1 || gr.m_rhs > kMaxPtrs || !gr.fermion_in || !gr.fermion_out || !dslash_supported(c->Nc, gr.m_rhs)) {
            set_error(kQcuErrorInvalidValue, "qcuInvertGroups: group %d has an invalid size or pointer table", k);
            return g_err;
        }
        for (int r = 0; r < gr.m_rhs; ++r)
            if (!gr.fermion_in[r] || !gr.fermion_out[r]) {
                set_error(kQcuErrorInvalidValue, "qcuInvertGroups: null fermion pointer in group %d", k);
                return g_err;
            }
        Rs.push_back(gr.m_rhs);
    }
    if (!c->groups || c->groups->sizes() != Rs) {
        c->groups.reset();
        c->groups = std::make_unique<GroupSolver>(c->g, c->Nc, Rs);
    }
    GroupSolver &gs = *c->groups;
    const int P = c->opts.inner_prec;
    const void *glo = c->gauge(P);
    if (!glo) return g_err;
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<GroupResult> res;
    // batch buffers and the per-group launch shapes (timed once per batch shape and precision)
    hipError_t e = gs.autotune(glo, P, c->stream);
    const long n = c->half_n();
    const long nR = n * gs.total_rhs();
    auto *b = static_cast<Cx<double> *>(gs.b_full());
    int base = 0;
    for (int k = 0; k < n_groups && e == hipSuccess; ++k) {
        std::vector<void *> in(groups[k].fermion_in, groups[k].fermion_in + Rs[k]);
        e = launch_gather(b + n * base, kDouble, table(in, 0, c->user_prec), c->user_prec, n, Rs[k], c->stream);
        if (e == hipSuccess)
            e = launch_gather(b + nR + n * base, kDouble, table(in, n, c->user_prec), c->user_prec, n, Rs[k], c->stream);
        base += Rs[k];
    }
    if (e == hipSuccess) e = gs.solve(c->gauge64, glo, P, c->kappa, max_iteration, tol, res, c->stream);
    auto *x = static_cast<Cx<double> *>(gs.x_full());
    base = 0;
    for (int k = 0; k < n_groups && e == hipSuccess; ++k) {
        std::vector<void *> out(groups[k].fermion_out, groups[k].fermion_out + Rs[k]);
        e = launch_scatter(table(out, 0, c->user_prec), c->user_prec, x + n * base, kDouble, n, Rs[k], c->stream);
        if (e == hipSuccess)
            e = launch_scatter(table(out, n, c->user_prec), c->user_prec, x + nR + n * base, kDouble, n, Rs[k], c->stream);
        base += Rs[k];
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (!hip_ok(e, "qcuInvertGroups")) return g_err;
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    bool all = true;
    for (int k = 0; k < n_groups; ++k) {
        all = all && res[k].converged;
        if (info) {
            info[k].converged = res[k].converged;
            info[k].iterations = res[k].iterations;
            info[k].outer_steps = res[k].outer_steps;
            info[k].sites_per_wave = res[k].spw;
            info[k].true_rel_residual = res[k].true_rel_residual;
            info[k].seconds = res[k].seconds;
            info[k].batch_seconds = secs;
        }
    }
    if (!all) {
        set_error(kQcuErrorNotConverged, "qcuInvertGroups: not every group converged in %d iterations", max_iteration);
        return kQcuErrorNotConverged;
    }
    return 0;
}

int qcuGroupsDslashVariant(void) {
    Context *c = ctx_or_error();
    if (!c) return -1;
    return c->groups ? c->groups->dslash_variant() : 0;
}

int qcuGroupsDslashTune(int variant, int variant64) {
    auto ok = [](int v) { return v == 0 || v == kDslashWaveSite || v == kDslashMfma || v == kGroupsPacked; };
    if (!ok(variant) || !ok(variant64)) {
        set_error(kQcuErrorInvalidValue, "qcuGroupsDslashTune: variants are 0, 2, 3 or 5");
        return g_err;
    }
    groups_force_variants(variant, variant64);
    return 0;
}

int qcuGroupsDslashVariant64(void) {
    Context *c = ctx_or_error();
    if (!c) return -1;
    return c->groups ? c->groups->dslash_variant64() : 0;
}

// ---- learned preconditioner --------------------------------------------
int qcuLgnnSetWeights(int channels, const double *w1, const double *w2) {
    Context *c = ctx_or_error();
    if (!c) return g_err;
    if (channels < 1 || channels > 64 || !w1 || !w2) {
        set_error(kQcuErrorInvalidValue, "qcuLgnnSetWeights: invalid arguments");
        return g_err;
    }
    if (!c->precond || c->precond->channels() != channels)
        c->precond = std::make_unique<LgnnPreconditioner>(c->g, c->Nc, c->R, channels);
    if (!hip_ok(c->precond->set_projected(c->lgnn_projected, c->stream), "lgnn projection")) return g_err;
    if (!hip_ok(c->precond->set_gauge(c->gauge64, c->stream, c->gauge_gen), "lgnn gauge")) return g_err;
    return hip_ok(c->precond->set_weights(w1, w2, c->stream), "lgnn weights") ? 0 : g_err;
}

int qcuLgnnGetWeights(double *w1, double *w2) {
    Context *c = ctx_or_error();
    if (!c) return g_err;
    if (!c->precond) {
        set_error(kQcuErrorNotInitialized, "no learned preconditioner");
        return g_err;
    }
    return hip_ok(c->precond->get_weights(w1, w2, c->stream), "lgnn weights") ? 0 : g_err;
}

int qcuLgnnApply(void *out, const void *in, int precision) {
    Context *c = ctx_or_error();
    if (!c) return g_err;
    if (!c->precond || !valid_storage(precision)) {
        set_error(kQcuErrorNotInitialized, "qcuLgnnApply: no preconditioner or invalid precision");
        return g_err;
    }
    bool ok = hip_ok(c->precond->set_gauge(c->gauge64, c->stream, c->gauge_gen), "lgnn gauge");
    ok = ok && hip_ok(c->precond->apply(out, in, precision, c->stream), "lgnn apply");
    ok = ok && hip_ok(hipStreamSynchronize(c->stream), "lgnn apply");
    return ok ? 0 : g_err;
}

int qcuLgnnLossGrad(const void *in, const void *target, double *loss, double *gw1, double *gw2) {
    Context *c = ctx_or_error();
    if (!c) return g_err;
    if (!c->precond) {
        set_error(kQcuErrorNotInitialized, "no learned preconditioner");
        return g_err;
    }
    bool ok = hip_ok(c->precond->set_gauge(c->gauge64, c->stream, c->gauge_gen), "lgnn gauge");
    ok = ok && hip_ok(c->precond->zero_grad(c->stream), "lgnn grad");
    ok = ok && hip_ok(c->precond->loss_grad(in, target, loss, c->stream), "lgnn grad");
    if (ok && gw1 && gw2) ok = hip_ok(c->precond->read_grad(gw1, gw2, c->stream), "lgnn grad");
    return ok ? 0 : g_err;
}

int qcuLgnnCollect(int max_iteration, double tol, int n_snapshots, int n_random, unsigned seed, int *n_batches) {
    Context *c = ctx_or_error();
    if (!c) return g_err;
    if (!c->have_dslash) {
        set_error(kQcuErrorNotInitialized, "Dslash is not initialized (call getDslash)");
        return g_err;
    }
    if (!wilson_only(c, "qcuLgnnCollect")) return g_err;
    if (n_snapshots < 0 || n_random < 0 || n_snapshots + n_random < 1) {
        set_error(kQcuErrorInvalidValue, "qcuLgnnCollect: need at least one snapshot or random pair");
        return g_err;
    }
    if (!queue_full(c)) return g_err;
    const long n = c->half_n();
    const long nR = n * c->R;
    const size_t b64 = size_t(nR) * sizeof(Cx<double>), b32 = size_t(nR) * sizeof(Cx<float>);
    SolverOptions opt;
    opt.inner_prec = kDouble;
    GaugeSet gs;
    gs.g64 = c->gauge64;
    SolveResult res;
    bool ok = hip_ok(launch_gather(c->bufA, kDouble, table(c->q_in, 0, c->user_prec), c->user_prec, 2 * n, c->R,
                                   c->stream), "gather");
    c->q_in.clear();
    c->q_out.clear();
    // pass 1: iteration count
    ok = ok && hip_ok(c->solver->solve(c->bufA, c->bufB, gs, c->kappa, max_iteration, tol, opt, nullptr, res,
                                       c->stream), "collect solve");
    if (!ok) return g_err;
    const int iters = res.iterations;
    std::vector<int> pick;
    for (int k = 0; k < n_snapshots; ++k) {
        const int it = 1 + int((long(k) * std::max(iters - 1, 1)) / n_snapshots);
        if (pick.empty() || pick.back() != it) pick.push_back(it);
    }
    // pass 2: the same (deterministic) iterates, FP64 snapshots at the picked iterations
    std::vector<void *> snaps(pick.size(), nullptr);
    for (auto &p : snaps)
        if (!hip_ok(hipMalloc(&p, b64), "snapshot")) ok = false;
    size_t next = 0;
    c->solver->iterate_hook = [&](int it, const void *x, hipStream_t s) -> hipError_t {
        if (next < pick.size() && it == pick[next]) return hipMemcpyAsync(snaps[next++], x, b64, hipMemcpyDeviceToDevice, s);
        return hipSuccess;
    };
    ok = ok && hip_ok(c->solver->solve(c->bufA, c->bufB, gs, c->kappa, max_iteration, tol, opt, nullptr, res,
                                       c->stream), "collect solve");
    c->solver->iterate_hook = nullptr;
    // gauge copy for this configuration
    void *gcopy = nullptr;
    ok = ok && hip_ok(hipMalloc(&gcopy, size_t(c->gauge_count()) * sizeof(Cx<double>)), "gauge copy");
    ok = ok && hip_ok(hipMemcpyAsync(gcopy, c->gauge64, size_t(c->gauge_count()) * sizeof(Cx<double>),
                                     hipMemcpyDeviceToDevice, c->stream), "gauge copy");
    const int cfg = int(c->lgnn_gauges.size());
    if (ok) c->lgnn_gauges.push_back(gcopy);
    const Cx<double> *xstar = static_cast<const Cx<double> *>(c->bufB) + nR;  // odd half of x
    void *d64 = nullptr, *r64 = nullptr, *tmp = nullptr;
    ok = ok && hip_ok(hipMalloc(&d64, b64), "alloc") && hip_ok(hipMalloc(&r64, b64), "alloc") &&
         hip_ok(hipMalloc(&tmp, b64), "alloc");
    C64 *minus1 = nullptr;
    ok = ok && hip_ok(hipMalloc(&minus1, sizeof(C64) * c->R), "alloc");
    ok = ok && hip_ok(launch_scalar_fill(minus1, C64{-1.0, 0.0}, c->R, c->stream), "fill");
    for (size_t k = 0; ok && k < next + size_t(n_random); ++k) {
        if (k < next) {
            // e = x* - x_k ;  r = A e  (odd Schur operator, FP64)
            AxpyArgs ax{kXpay, d64, xstar, snaps[k], nullptr, minus1, nullptr, nullptr, n, c->R, kDouble};
            ok = hip_ok(launch_axpy(ax, c->stream), "pair");
        } else {
            // e = complex Gaussian vector, r = A e: pins the network on the high modes too
            const uint64_t key = (uint64_t(seed) << 32) ^ (uint64_t(cfg) << 20) ^ uint64_t(k);
            ok = hip_ok(launch_gaussian(static_cast<Cx<double> *>(d64), nR, key, c->stream), "pair");
        }
        ok = ok && hip_ok(c->solver->apply_schur(r64, d64, tmp, c->gauge64, kDouble, c->kappa, 0, c->stream), "pair");
        Context::LgnnBatch bt{cfg, nullptr, nullptr};
        ok = ok && hip_ok(hipMalloc(&bt.e, b32), "pair") && hip_ok(hipMalloc(&bt.r, b32), "pair");
        ok = ok && hip_ok(launch_convert(bt.e, kSingle, d64, kDouble, nR, nullptr, c->R, c->stream), "pair");
        ok = ok && hip_ok(launch_convert(bt.r, kSingle, r64, kDouble, nR, nullptr, c->R, c->stream), "pair");
        if (ok) c->lgnn_batches.push_back(bt);
    }
    ok = ok && hip_ok(hipStreamSynchronize(c->stream), "collect");
    for (void *p : snaps)
        if (p) (void)hipFree(p);
    for (void *p : {d64, r64, tmp, static_cast<void *>(minus1)})
        if (p) (void)hipFree(p);
    if (n_batches) *n_batches = int(c->lgnn_batches.size());
    return ok ? 0 : g_err;
}

int qcuLgnnTrain(int epochs, double learning_rate, unsigned seed, double *loss_per_epoch) {
    Context *c = ctx_or_error();
    if (!c) return g_err;
    if (!c->precond || c->lgnn_batches.empty() || epochs < 1) {
        set_error(kQcuErrorNotInitialized, "qcuLgnnTrain: needs weights (qcuLgnnSetWeights) and collected data");
        return g_err;
    }
    const long nR = c->half_n() * c->R;
    std::vector<size_t> order(c->lgnn_batches.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    uint64_t st = 0x9E3779B97F4A7C15ull ^ seed;
    bool ok = true;
    for (int ep = 0; ok && ep < epochs; ++ep) {
        for (size_t i = order.size(); i > 1; --i) {  // Fisher-Yates, splitmix64
            st += 0x9E3779B97F4A7C15ull;
            uint64_t z = st;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            std::swap(order[i - 1], order[size_t(z % i)]);
        }
        const double lr = 0.5 * learning_rate * (1.0 + std::cos(M_PI * double(ep) / double(epochs)));
        double total = 0.0;
        int cur = -1;
        for (size_t i = 0; ok && i < order.size(); ++i) {
            const auto &bt = c->lgnn_batches[order[i]];
            if (bt.cfg != cur) {
                ok = hip_ok(c->precond->set_gauge(c->lgnn_gauges[bt.cfg], c->stream), "lgnn gauge");
                cur = bt.cfg;
            }
            double L = 0.0;
            ok = ok && hip_ok(c->precond->zero_grad(c->stream), "lgnn train");
            if (c->lgnn_power > 0) {
                // a fixed complex Gaussian vector per batch (the batch index is its key)
                const uint64_t key = (uint64_t(seed) << 32) ^ (uint64_t(order[i]) * 0x9E3779B97F4A7C15ull);
                ok = ok && hip_ok(launch_gaussian(static_cast<Cx<double> *>(c->bufA), nR, key, c->stream), "lgnn train");
                ok = ok && hip_ok(launch_convert(c->bufB, kSingle, c->bufA, kDouble, nR, nullptr, c->R, c->stream),
                                  "lgnn train");
                ok = ok && hip_ok(c->precond->power_loss_grad(c->bufB, c->lgnn_power, c->kappa, &L, c->stream),
                                  "lgnn train");
            } else {
                ok = ok && hip_ok(c->precond->loss_grad(bt.r, bt.e, &L, c->stream), "lgnn train");
            }
            ok = ok && hip_ok(c->precond->adam_step(lr, 1, c->stream), "lgnn train");
            total += L;
        }
        if (loss_per_epoch) loss_per_epoch[ep] = total / (double(order.size()) * c->R);
    }
    ok = ok && hip_ok(c->precond->set_gauge(c->gauge64, c->stream, c->gauge_gen), "lgnn gauge");
    ok = ok && hip_ok(hipStreamSynchronize(c->stream), "lgnn train");
    return ok ? 0 : g_err;
}

int qcuLgnnSetObjective(int power) {
    Context *c = ctx_or_error();
    if (!c) return g_err;
    if (power < 0 || power > 8) {
        set_error(kQcuErrorInvalidValue, "qcuLgnnSetObjective: power must be 0 (pair loss) or 1..8");
        return g_err;
    }
    c->lgnn_power = power;
    return 0;
}

int qcuLgnnSetProjected(int on) {
    Context *c = ctx_or_error();
    if (!c) return g_err;
    if (on != 0 && on != 1) {
        set_error(kQcuErrorInvalidValue, "qcuLgnnSetProjected: on must be 0 or 1");
        return g_err;
    }
    c->lgnn_projected = on == 1;
    if (c->precond && !hip_ok(c->precond->set_projected(c->lgnn_projected, c->stream), "lgnn projection"))
        return g_err;
    return 0;
}

int qcuLgnnPowerLossGrad(const void *v, int power, double *loss, double *gw1, double *gw2) {
    Context *c = ctx_or_error();
    if (!c) return g_err;
    if (!c->precond || !c->precond->ready() || !v || power < 1 || power > 8 || !c->have_dslash) {
        set_error(kQcuErrorInvalidValue, "qcuLgnnPowerLossGrad: needs weights, getDslash and 1 <= power <= 8");
        return g_err;
    }
    bool ok = hip_ok(c->precond->set_gauge(c->gauge64, c->stream, c->gauge_gen), "lgnn gauge");
    ok = ok && hip_ok(c->precond->zero_grad(c->stream), "lgnn power");
    ok = ok && hip_ok(c->precond->power_loss_grad(v, power, c->kappa, loss, c->stream), "lgnn power");
    if (ok && gw1 && gw2) ok = hip_ok(c->precond->read_grad(gw1, gw2, c->stream), "lgnn power");
    ok = ok && hip_ok(hipStreamSynchronize(c->stream), "lgnn power");
    return ok ? 0 : g_err;
}

int qcuLgnnClearData(void) {
    Context *c = ctx_or_error();
    if (!c) return g_err;
    c->clear_lgnn_data();
    return 0;
}

static hipEvent_t g_ev0 = nullptr, g_ev1 = nullptr;
int qcuTimerStart(void) {
    if (!g_ev0 && (!hip_ok(hipEventCreate(&g_ev0), "event") || !hip_ok(hipEventCreate(&g_ev1), "event")))
        return g_err;
    return hip_ok(hipEventRecord(g_ev0, nullptr), "event") ? 0 : g_err;
}
double qcuTimerStopMs(void) {
    if (!g_ev0) return -1.0;
    float ms = -1.0f;
    if (!hip_ok(hipEventRecord(g_ev1, nullptr), "event") || !hip_ok(hipEventSynchronize(g_ev1), "event") ||
        !hip_ok(hipEventElapsedTime(&ms, g_ev0, g_ev1), "event"))
        return -1.0;
    return double(ms);
}

int qcuCopyComplexVector(void *dst, int dst_precision, const void *src, int src_precision, long count) {
    if (!valid_storage(dst_precision) || !valid_storage(src_precision) || count < 0) {
        set_error(kQcuErrorInvalidValue, "copy: invalid arguments");
        return g_err;
    }
    return hip_ok(launch_convert(dst, dst_precision, src, src_precision, count, nullptr, 1, nullptr), "copy")
               ? 0 : g_err;
}

}  // extern "C"
