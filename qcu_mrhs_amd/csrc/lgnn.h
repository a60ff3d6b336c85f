// lgnn.h -- learned gauge-covariant right preconditioner (linear two-layer
// parallel-transport network) for the odd Schur system.
#pragma once

#include "solver.h"

namespace qcu {

class LgnnPreconditioner : public Preconditioner {
public:
    LgnnPreconditioner() = default;
    ~LgnnPreconditioner() override = default;
    bool ready() const { return false; }
    bool supports(int) const override { return false; }
    hipError_t apply(void *, const void *, int, hipStream_t) override { return hipErrorNotSupported; }
};

}  // namespace qcu
