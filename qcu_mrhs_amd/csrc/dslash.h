// dslash.h -- internal launch interface of the Wilson hopping kernels.
#pragma once

#include "qcu_common.h"

namespace qcu {

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
};

bool dslash_supported(int Nc, int R);
hipError_t launch_dslash(const DslashArgs &a, hipStream_t stream);

}  // namespace qcu
