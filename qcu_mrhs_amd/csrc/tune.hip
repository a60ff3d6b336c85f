// tune.hip -- dslash launch-shape autotuner.
//
// For one (lattice, Nc, R, storage precision) it times every applicable
// kernel variant over a grid of sites-per-wavefront values and (y, z)
// traversal tiles on scratch buffers filled with bounded values, using HIP
// events around repeated launches, and records the fastest shape in the
// dslash tuning table (consulted whenever a launch asks for the automatic
// shape).  Candidates are limited to shapes every kernel accepts, so the
// tuner never launches an invalid configuration.
#include <algorithm>
#include <vector>

#include "dslash.h"
#include "tune.h"

namespace qcu {
namespace {

__global__ void k_fill_words(uint32_t *p, size_t n, uint32_t seed, int prec) {
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint32_t h = uint32_t(i) * 2654435761u ^ seed;
        h ^= h >> 13;
        h *= 0x5bd1e995u;
        h ^= h >> 15;
        const float v = float(h & 0xffffu) / 65536.0f - 0.5f;  // in [-0.5, 0.5)
        uint32_t w;
        if (prec == kHalf) {
            const _Float16 a = _Float16(v), b = _Float16(-v);
            w = uint32_t(__builtin_bit_cast(uint16_t, a)) | (uint32_t(__builtin_bit_cast(uint16_t, b)) << 16);
        } else if (prec == kBF16) {
            const __bf16 a = __bf16(v), b = __bf16(-v);
            w = uint32_t(__builtin_bit_cast(uint16_t, a)) | (uint32_t(__builtin_bit_cast(uint16_t, b)) << 16);
        } else {
            w = __builtin_bit_cast(uint32_t, v);  // FP32 words; FP64 buffers are filled as pairs
            if (prec == kDouble) w = (i & 1) ? 0x3fe00000u ^ (h & 0x000fffffu) : 0u;
        }
        p[i] = w;
    }
}

int divisors_upto(int L, int cap, std::vector<int> &out) {
    out.clear();
    for (int d = 1; d <= std::min(L, cap); ++d)
        if (L % d == 0) out.push_back(d);
    return int(out.size());
}

}  // namespace

hipError_t dslash_autotune(const int L[4], int Nc, int R, int prec, int reps, DslashTune &best,
                           double &best_us, std::vector<TuneRecord> *log) {
    if (!dslash_supported(Nc, R)) return hipErrorInvalidValue;
    const Geom g = make_geom(L);
    const size_t cb = 2 * prec_bytes(prec);
    const size_t vec = size_t(g.Vh) * 4 * Nc * R * cb;
    const size_t gauge = size_t(g.Vh) * 8 * Nc * Nc * cb;
    void *in = nullptr, *out = nullptr, *gg = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t err = hipMalloc(&in, vec);
    if (err == hipSuccess) err = hipMalloc(&out, vec);
    if (err == hipSuccess) err = hipMalloc(&gg, gauge);
    if (err == hipSuccess) err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
    if (err == hipSuccess) {
        hipLaunchKernelGGL(k_fill_words, dim3(1024), dim3(256), 0, nullptr, static_cast<uint32_t *>(in), vec / 4, 17u, prec);
        hipLaunchKernelGGL(k_fill_words, dim3(1024), dim3(256), 0, nullptr, static_cast<uint32_t *>(gg), gauge / 4, 29u, prec);
        err = hipDeviceSynchronize();
    }
    // candidate shapes
    std::vector<DslashTune> cands;
    std::vector<int> variants = {kDslashLaneRhs, kDslashWaveSite};
    if (dslash_mfma_supported(Nc, prec)) variants.push_back(kDslashMfma);
    std::vector<int> bys, bzs;
    divisors_upto(L[1], 12, bys);
    divisors_upto(L[2], 12, bzs);
    for (int v : variants) {
        if (v == kDslashLaneRhs) {
            DslashTune t;
            t.variant = v;
            cands.push_back(t);
            continue;
        }
        for (int spw : {2, 4, 8, 16}) {
            for (int by : bys)
                for (int bz : bzs) {
                    // keep the search small: square-ish tiles plus the plain order
                    if (!(by == bz || (by == L[1] && bz == L[2]) || by * 2 == bz || bz * 2 == by)) continue;
                    DslashTune t;
                    t.variant = v;
                    t.by = by;
                    t.bz = bz;
                    t.spw = spw;
                    cands.push_back(t);
                }
            DslashTune t;
            t.variant = v;
            t.by = L[1];
            t.bz = L[2];
            t.spw = spw;
            cands.push_back(t);
        }
    }
    best_us = 1e30;
    for (size_t i = 0; err == hipSuccess && i < cands.size(); ++i) {
        DslashArgs a{out, in, gg, nullptr, 1.0, g, Nc, R, 1, 0, prec};
        a.tune = cands[i];
        err = launch_dslash(a, nullptr);  // warm up (also builds index tables)
        if (err != hipSuccess) break;
        err = hipEventRecord(e0, nullptr);
        for (int k = 0; k < reps && err == hipSuccess; ++k) err = launch_dslash(a, nullptr);
        if (err == hipSuccess) err = hipEventRecord(e1, nullptr);
        if (err == hipSuccess) err = hipEventSynchronize(e1);
        float ms = 0.f;
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
        const double us = 1e3 * double(ms) / reps;
        if (log) log->push_back(TuneRecord{cands[i], us});
        if (err == hipSuccess && us < best_us) {
            best_us = us;
            best = cands[i];
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    for (void *p : {in, out, gg})
        if (p) (void)hipFree(p);
    if (err == hipSuccess) dslash_set_tuned(L, Nc, R, prec, best);
    return err;
}

}  // namespace qcu
