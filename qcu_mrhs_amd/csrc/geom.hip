// geom.hip -- construction and cache of the lattice index tables (geom.h).
#include <map>
#include <mutex>
#include <tuple>

#include "dslash_common.h"
#include "geom.h"

namespace qcu {
namespace {

__global__ void k_build_order(int *order, Geom g, int by, int bz) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < g.Vh) order[q] = traverse(q, g, by, bz);
}

__global__ void k_build_nbr(int *nbr, Geom g) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= 2L * g.Vh) return;
    const int parity = int(i / g.Vh);
    const int site = int(i - long(parity) * g.Vh);
    const Site s = site_coords(site, g, parity);
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) {
        nbr[i * 8 + 2 * mu] = neighbour(site, s, g, mu, true);
        nbr[i * 8 + 2 * mu + 1] = neighbour(site, s, g, mu, false);
    }
}

// one row per traversal position: the site and its 8 neighbours, so a
// wavefront gets everything it needs for a site with one scalar load
__global__ void k_build_walk(int *walk, const int *order, const int *nbr, int Vh) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= 2L * Vh) return;
    const int parity = int(i / Vh);
    const int q = int(i - long(parity) * Vh);
    const int site = order[q];
    int *row = walk + i * 9;
    row[0] = site;
#pragma unroll
    for (int k = 0; k < 8; ++k) row[1 + k] = nbr[(long(parity) * Vh + site) * 8 + k];
}

using Key4 = std::tuple<int, int, int, int>;
using Key6 = std::tuple<int, int, int, int, int, int>;
std::mutex g_mu;
std::map<Key4, int *> g_nbr;
std::map<Key6, int *> g_order;
std::map<Key6, int *> g_walk;

}  // namespace

hipError_t geom_tables(const Geom &g, int by, int bz, GeomTables &out, hipStream_t stream) {
    std::lock_guard<std::mutex> lk(g_mu);
    bool built = false;
    const Key4 k4{g.L[0], g.L[1], g.L[2], g.L[3]};
    auto it = g_nbr.find(k4);
    if (it == g_nbr.end()) {
        built = true;
        int *p = nullptr;
        QCU_HIP_TRY(hipMalloc(&p, sizeof(int) * 16 * size_t(g.Vh)));
        hipLaunchKernelGGL(k_build_nbr, dim3(div_up(2L * g.Vh, 256)), dim3(256), 0, stream, p, g);
        QCU_HIP_TRY(hipGetLastError());
        it = g_nbr.emplace(k4, p).first;
    }
    const Key6 k6{g.L[0], g.L[1], g.L[2], g.L[3], by, bz};
    auto jt = g_order.find(k6);
    if (jt == g_order.end()) {
        built = true;
        int *p = nullptr;
        QCU_HIP_TRY(hipMalloc(&p, sizeof(int) * size_t(g.Vh)));
        hipLaunchKernelGGL(k_build_order, dim3(div_up(g.Vh, 256)), dim3(256), 0, stream, p, g, by, bz);
        QCU_HIP_TRY(hipGetLastError());
        jt = g_order.emplace(k6, p).first;
    }
    auto wt = g_walk.find(k6);
    if (wt == g_walk.end()) {
        built = true;
        int *p = nullptr;
        QCU_HIP_TRY(hipMalloc(&p, sizeof(int) * 18 * size_t(g.Vh)));
        hipLaunchKernelGGL(k_build_walk, dim3(div_up(2L * g.Vh, 256)), dim3(256), 0, stream, p, jt->second,
                           it->second, g.Vh);
        QCU_HIP_TRY(hipGetLastError());
        wt = g_walk.emplace(k6, p).first;
    }
    out.nbr = it->second;
    out.order = jt->second;
    out.walk = wt->second;
    // the tables may be read from another stream later: finish building them now
    if (built) QCU_HIP_TRY(hipStreamSynchronize(stream));
    return hipSuccess;
}

void geom_tables_release() {
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto &kv : g_nbr) (void)hipFree(kv.second);
    for (auto &kv : g_order) (void)hipFree(kv.second);
    for (auto &kv : g_walk) (void)hipFree(kv.second);
    g_nbr.clear();
    g_order.clear();
    g_walk.clear();
}

}  // namespace qcu
