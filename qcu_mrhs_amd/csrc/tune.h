// tune.h -- dslash launch-shape autotuner.
#pragma once

#include <vector>

#include "dslash.h"

namespace qcu {

struct TuneRecord {
    DslashTune tune;
    double us;
};

// Time the candidate shapes for (L, Nc, R, prec) and store the fastest in the
// tuning table.  `log` (optional) receives every candidate's time.
hipError_t dslash_autotune(const int L[4], int Nc, int R, int prec, int reps, DslashTune &best,
                           double &best_us, std::vector<TuneRecord> *log);
// Tuning table <-> text file, one entry per line:
// Lx Ly Lz Lt Nc R precision variant by bz spw
hipError_t dslash_tune_save(const char *path);
hipError_t dslash_tune_load(const char *path, int &n_loaded);

}  // namespace qcu
