// clover.h -- Sheikholeslami-Wohlert clover term of the Wilson-clover operator
// M = A - kappa D on gfx950.
//
// The reference declares the clover type (include/qcu_helper_macro.h:7-11) and
// gives up on it (src/qcu_interface.cu:97-105); this is the textbook term in
// the kappa normalisation of M = 1 - kappa D (forward hop (1 - gamma_mu) U):
//
//   A(x) = 1 - kappa c_sw sum_{mu<nu} sigma_{mu nu} F_{mu nu}(x),
//   sigma_{mu nu} = (i/2)[gamma_mu, gamma_nu],
//   F_{mu nu} = ((Q - Q^+) - tr(Q - Q^+)/Nc) / (8i),  Q = sum of the 4 leaves.
//
// In the chiral basis of the dslash kernels sigma is block diagonal in the
// spin pairs {0,1} and {2,3}, so A(x) is two Hermitian 2Nc x 2Nc blocks:
//   clover field  [parity 2][Vh][chirality 2][2Nc][2Nc]  complex,
//   block row/column index = (spin within the pair) * Nc + colour.
// Built once per (gauge, kappa, c_sw) in FP64 with its block inverse; the
// mixed-precision solves read FP32 copies.
#pragma once

#include "qcu_common.h"

namespace qcu {

// clover[2][Vh][2][n][n] (FP64) from the FP64 even/odd gauge field
hipError_t launch_clover_build(void *clover, const void *gauge, const Geom &g, int Nc, double csw, double kappa,
                               hipStream_t s);
// inverse of each of `nblocks` n x n FP64 blocks (Gauss-Jordan, partial pivoting);
// inv may equal clover
hipError_t launch_clover_invert(void *inv, const void *clover, long nblocks, int n, hipStream_t s);

// out = A in + coef * y on one parity's MRHS half vectors [Vh][4][Nc][R] stored in
// `prec`; `blocks` = that parity's [Vh][2][n][n] in `clover_prec` (kDouble or
// kSingle).  y may be null; out may alias in or y.
hipError_t launch_clover_apply(void *out, const void *in, const void *blocks, const void *y, double coef,
                               int Vh, int Nc, int R, int prec, int clover_prec, hipStream_t s);

bool clover_supported(int Nc);

// Device clover fields of one operator: FP64 term and inverse for both
// parities, and their FP32 copies for the mixed-precision inner solves.
struct CloverSet {
    const void *a64 = nullptr, *inv64 = nullptr;
    const void *a32 = nullptr, *inv32 = nullptr;
    long parity_stride = 0;  // complex entries per parity ([Vh][2][n][n])

    const void *term(int prec, int parity) const { return at(prec == kDouble ? a64 : a32, prec, parity); }
    const void *inverse(int prec, int parity) const { return at(prec == kDouble ? inv64 : inv32, prec, parity); }
    static int storage(int prec) { return prec == kDouble ? kDouble : kSingle; }

private:
    const void *at(const void *base, int prec, int parity) const {
        const size_t eb = prec == kDouble ? sizeof(Cx<double>) : sizeof(Cx<float>);
        return static_cast<const char *>(base) + size_t(parity) * size_t(parity_stride) * eb;
    }
};

}  // namespace qcu
