// geom.h -- device index tables of a lattice: traversal order and the eight
// neighbour indices of every checkerboard site.  Built once per lattice (and
// per traversal tile) on the device and cached for the life of the process.
// They cost 4 + 32 bytes per site per dslash (about 2% of the compulsory
// spinor + link traffic of a 16-RHS FP16 dslash) and remove every integer
// division and wrap test from the hopping kernels.
#pragma once

#include "qcu_common.h"

namespace qcu {

struct GeomTables {
    const int *order = nullptr;  // [Vh]: traversal index -> checkerboard site
    const int *nbr = nullptr;    // [2][Vh][8]: output parity, site, hop 2*mu + (0 fwd | 1 bwd)
    const int *walk = nullptr;   // [2][Vh][9]: output parity, traversal index -> (site, its 8 neighbours)
};

// Tables for lattice g and traversal tile (by, bz); built on `stream` on first use.
hipError_t geom_tables(const Geom &g, int by, int bz, GeomTables &out, hipStream_t stream);
// Free every cached table (finalizeQcu).
void geom_tables_release();

}  // namespace qcu
