// lgnn.h -- learned gauge-covariant right preconditioner for the odd Schur
// system: two linear parallel-transport convolution layers 1 -> C -> 1
// (C = 8 channels, nearest-neighbour hops, complex spin-matrix weights).
//
// Because the layers are linear and their weights act on spin only, the
// network restricted to odd sites equals the two-hop stencil
//     N psi = M0 psi + sum_{h,h'=1..8} T_h M[h][h'] T_h' psi,
//     M0 = sum_o W2[o][0] W1[o][0],   M[h][h'] = sum_o W2[o][h] W1[o][h'],
// with T_h the parallel transport along hop h (forward U_mu(x) psi(x+mu),
// backward U_mu(x-mu)^+ psi(x-mu)).  The GPU evaluates it in three steps:
// transport of the odd input to the eight hop fields on the even sites, the
// (hop, spin) mixing by the 32 x 32 complex matrix M, and the transported
// gather back onto the odd sites.  Gradients with respect to M0 and M are
// outer-product reductions; the chain rule to W1, W2 and the Adam update run
// on the device as well.
#pragma once

#include <vector>

#include "blas.h"
#include "solver.h"

namespace qcu {

constexpr int kLgnnHops = 9;

// out[i] = complex Gaussian (unit variance per real component), a pure
// function of (key, i): SplitMix64 counter hashing + Box-Muller.
hipError_t launch_gaussian(Cx<double> *out, long count, uint64_t key, hipStream_t s);

// FP16 / BF16 network application on matrix cores (lgnn_mfma.hip), Nc <= 4:
// q_k(y) = sum_k' M[k][k'] (T_k' in)(y) on the even sites (transport on VALU,
// the 32 x 32 (hop, spin) mixing on MFMA), then out(x) = M0 in(x) + sum_k
// (T_k q_k)(x) on the odd sites (transport on MFMA).  q holds 8 fields of one
// 32-bit word per complex; mcol = the collapsed stencil (M [32][32], M0 [4][4]).
// projected: M[k][.] = P_k M[k][.] with P_k = (1 -+ gamma_mu) / 2 the Wilson projector of
// hop k (second-layer weights in its image), so q_k has two independent spins and the
// fields hold only spins 0 and 1 (half the field traffic); the gather rebuilds the others.
constexpr size_t kLgnnFragBytes = 8 * 64 * 16;  // MFMA fragments of M stored after the q fields
hipError_t launch_lgnn_mfma(void *out, const void *in, const void *gauge, void *q, const Cx<float> *mcol,
                            const int *walk, const Geom &g, int Nc, int R, int prec, bool projected,
                            hipStream_t s);

class LgnnPreconditioner : public Preconditioner {
public:
    LgnnPreconditioner(const Geom &g, int Nc, int R, int channels);
    ~LgnnPreconditioner() override;
    LgnnPreconditioner(const LgnnPreconditioner &) = delete;
    LgnnPreconditioner &operator=(const LgnnPreconditioner &) = delete;

    bool ready() const { return have_weights_ && gauge_ != nullptr; }
    bool supports(int prec) const override;
    int channels() const { return C_; }

    // Gauge field the transports use (FP64 even/odd layout, device); a
    // single-precision copy is kept.  Must be set before apply().  gen >= 0 names
    // the content (the caller's load count): the same pointer and generation as
    // the last call keep the converted copies; gen < 0 always converts.
    hipError_t set_gauge(const void *gauge64, hipStream_t s, long long gen = -1);
    // Weights: W1 [C][9][4][4], W2 [C][9][4][4] complex FP64 (host arrays).
    hipError_t set_weights(const double *w1, const double *w2, hipStream_t s);
    hipError_t get_weights(double *w1, double *w2, hipStream_t s);
    // Projected network: every second-layer weight of hop h = 1..8 is kept in the image of
    // that hop's Wilson projector, W2[o][h] = P_h W2[o][h] with P_h = (1 -+ gamma_mu) / 2
    // (1 - gamma_mu forward, 1 + gamma_mu backward, the spin structure of the Wilson hopping
    // term and of the first Neumann term): the weights are projected when set, when this mode
    // is switched on and after every training step, and the FP16 / BF16 application stores
    // half the hop fields.
    hipError_t set_projected(bool on, hipStream_t s);
    bool projected() const { return projected_; }

    // out = N in on odd-parity MRHS half vectors stored in precision `prec`
    hipError_t apply(void *out, const void *in, int prec, hipStream_t s) override;

    // One forward/backward pass on a batch of R pairs (FP32 odd MRHS vectors):
    // loss = sum_r ||N r_r - e_r||^2 / ||e_r||^2; gradients dL/dRe W + i dL/dIm W
    // accumulated into the device gradient buffers (zeroed by zero_grad()).
    hipError_t zero_grad(hipStream_t s);
    hipError_t loss_grad(const void *in, const void *target, double *loss_host, hipStream_t s);
    hipError_t read_grad(double *gw1, double *gw2, hipStream_t s);
    // Adam step with the accumulated gradient (divided by `batches`).
    hipError_t adam_step(double lr, int batches, hipStream_t s);

    // Residual-power objective on a batch of R FP32 odd vectors v:
    //     loss = sum_r ||(1 - A N)^K w_r||^2,  w_r = v_r / ||v_r||,
    // A = 1 - kappa^2 D_oe D_eo the odd Schur operator of the current gauge field in
    // FP32.  The K-th power weights the modes that the preconditioned operator A N
    // leaves far from 1 -- those that set the Krylov iteration count -- ever more
    // strongly as K grows.  Gradients as in loss_grad (accumulated).
    hipError_t power_loss_grad(const void *v, int K, double kappa, double *loss_host, hipStream_t s);

private:
    hipError_t ensure_fields(hipStream_t s);
    hipError_t ensure_work32(hipStream_t s);
    hipError_t apply_mfma(void *out, const void *in, int prec, hipStream_t s);
    hipError_t collapse(hipStream_t s);
    hipError_t project_w2(hipStream_t s);  // W2[o][h] = P_h W2[o][h], h = 1..8
    // stencil: collapsed stencil to apply (default the network's); hops: where the 8 hop
    // fields of the input go (default the work field u_)
    hipError_t forward(const void *in, int prec, void *out, int out_prec, hipStream_t s,
                       const Cx<float> *stencil = nullptr, void *hops = nullptr);
    // gradient of a loss whose derivative with respect to the network output N w is dphi
    // (FP32 odd field w; hops = the 8 hop fields forward(w) produced, default u_)
    hipError_t accumulate_grad(const void *w, const void *dphi, hipStream_t s, const void *hops = nullptr);
    // FP32 odd Schur operator with the FP32 gauge copy: out = A in (dagger: A^+)
    hipError_t schur32(void *out, const void *in, double kappa, int dagger, hipStream_t s);
    hipError_t ensure_power(int K, hipStream_t s);

    Geom g_;
    int Nc_, R_, C_;
    long n_;                        // complex entries per RHS of a half vector
    bool have_weights_ = false;
    bool projected_ = false;
    const void *gauge_ = nullptr;   // FP64 gauge (borrowed)
    long long gauge_gen_ = -1;      // generation of gauge_'s content the copies hold (-1: unknown)
    void *gauge32_ = nullptr;       // FP32 copy
    // parameters (FP64 complex) and Adam moments
    C64 *w1_ = nullptr, *w2_ = nullptr;          // [C][9][16]
    C64 *g1_ = nullptr, *g2_ = nullptr;          // gradients
    C64 *m1_ = nullptr, *m2_ = nullptr, *v1_ = nullptr, *v2_ = nullptr;  // Adam (v stores |g|^2 in re)
    int adam_t_ = 0;
    // collapsed stencil (FP32 complex): M0 [16], M [64][16] as a 32 x 32 matrix
    Cx<float> *mcol_ = nullptr;                  // [32*32 + 16]
    Cx<float> *madj_ = nullptr;                  // the stencil of N^+ (same layout)
    // residual-power objective work fields (FP32 odd half vectors): w_0..w_K, g, A^+ g, N^+ A^+ g,
    // the dslash intermediate on the even sites, and per-RHS scale factors
    std::vector<void *> pw_, phops_;
    void *pg_ = nullptr, *pu_ = nullptr, *pn_ = nullptr, *ptmp_ = nullptr;
    double *pscale_ = nullptr;
    // work fields (FP32): u[8] hop fields on even sites, q[8], odd scratch
    void *u_ = nullptr, *q_ = nullptr, *inf_ = nullptr, *phi_ = nullptr, *dlt_ = nullptr;
    C64 *gm_ = nullptr;                          // dL/dM* partials [nblk][32*32 + 16]
    int gram_blocks_ = 0;
    double *dnorm_ = nullptr, *dloss_ = nullptr, *hbuf_ = nullptr;
    ReduceWork red_;
    // FP16 / BF16 inference path: gauge copies [half, bf16] and the 8 mixed hop fields
    void *glo_[2] = {nullptr, nullptr};
    bool glo_valid_[2] = {false, false};
    void *q16_ = nullptr;
};

}  // namespace qcu
