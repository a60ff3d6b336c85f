// layout.h -- gather/scatter and gauge reordering launchers.
#pragma once

#include "qcu_common.h"

namespace qcu {

constexpr int kMaxPtrs = 128;  // RHS pointers passed by value in kernel arguments

struct PtrTable {
    const void *p[kMaxPtrs];
};

// R user vectors of n complex numbers -> RHS-innermost [n][R]
hipError_t launch_gather(void *dst, int dst_prec, const PtrTable &src, int src_prec, long n, int R,
                         hipStream_t s);
// RHS-innermost [n][R] -> R user vectors
hipError_t launch_scatter(const PtrTable &dst, int dst_prec, const void *src, int src_prec, long n,
                          int R, hipStream_t s);
// lexicographic <-> even/odd gauge field (reverse = 1: eo -> lex)
hipError_t launch_gauge_eo(void *eo, void *lex, const Geom &g, int Nc, int prec, int reverse,
                           hipStream_t s);

}  // namespace qcu
