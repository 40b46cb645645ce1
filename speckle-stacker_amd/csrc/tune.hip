// tune.hip -- autotuner of the batched triple-product kernel.
//
// For one (frame extents, bispectrum depth) it times k_accum_tile on synthetic
// spectra: every supported vy tile at one batch size, then batch sizes with the
// fastest tile.  The choice is the smallest batch whose frames/s is within 3 %
// of the best measured (smaller batches need less HBM and finish sooner), and it
// is cached per (ncols, nrows, bdepth) for the process; jobs created with
// vy_tile or frames_per_batch 0 take the cached choice when there is one.
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "smip_internal.h"

using namespace smip_amd;

namespace {

std::mutex g_mu;
std::map<std::tuple<uint64_t, uint64_t, uint64_t>, smip_tune_result> g_cache;

// deterministic pseudo-random spectrum values in [-1, 1), zero outside the rows
// [pad, pad + nrows) of each stored column, as the forward transform leaves them
__global__ void k_fill_spectra(float2* H, int64_t n, int32_t rs, int32_t pad, int32_t nrows)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int32_t r = (int32_t)(i % rs);
    if (r < pad || r >= pad + nrows) {
        H[i] = make_float2(0.f, 0.f);
        return;
    }
    uint64_t z = (uint64_t)i * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const float a = (float)(z & 0xFFFFFF) * (1.f / 8388608.f) - 1.f;
    const float b = (float)((z >> 24) & 0xFFFFFF) * (1.f / 8388608.f) - 1.f;
    H[i] = make_float2(a, b);
}

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    Timer()
    {
        SMIP_HIP(hipEventCreate(&a));
        SMIP_HIP(hipEventCreate(&b));
    }
    ~Timer()
    {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
};

} // namespace

namespace smip_amd {

bool tune_lookup(uint64_t ncols, uint64_t nrows, uint64_t bdepth, int* vy_tile, int* batch)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_cache.find({ ncols, nrows, bdepth });
    if (it == g_cache.end())
        return false;
    *vy_tile = it->second.vy_tile;
    *batch = it->second.frames_per_batch;
    return true;
}

} // namespace smip_amd

extern "C" SMIP_API int smip_autotune(smip_ctx* ctx, uint64_t ncols, uint64_t nrows, uint64_t bdepth,
    uint64_t max_bytes, smip_tune_result* out)
{
    return guarded([&] {
        if (!ctx || !out)
            fail(SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        smip_job_params p {};
        p.ncols = ncols;
        p.nrows = nrows;
        p.bdepth = bdepth;
        p.reco_radius = 1.0;
        p.frame_dtype = SMIP_DTYPE_F32;
        p.precision = SMIP_PRECISION_MIXED;
        uint64_t fp0 = 0;
        if (smip_job_footprint(&p, &fp0) != SMIP_OK) // validates the geometry
            fail(SMIP_ERR_INVALID_ARGUMENT, smip_last_error());
        uint64_t free_b = 0, total_b = 0;
        if (smip_ctx_mem_info(ctx, &free_b, &total_b) != SMIP_OK)
            fail(SMIP_ERR_HIP, smip_last_error());
        if (max_bytes == 0)
            max_bytes = free_b / 4;

        const Desc d = make_desc(ncols, nrows, bdepth, bdepth);
        const LoopBounds L = loop_bounds(d, (int64_t)ncols, (int64_t)nrows);
        static const int kTiles[] = { 9, 13, 17, 25, 33 };
        static const int kBatches[] = { 32, 64, 128, 256, 512, 1024 };
        // spectra sized for the largest tile's padding
        int pad = 0, tail = 0;
        for (int t : kTiles) {
            pad = std::max(pad, tile_plan_pad_needed(d, t));
            tail = std::max(tail, tile_plan_tail_needed(d, (int)nrows, t));
        }
        const HalfPlane hp = make_half_plane((int32_t)ncols, (int32_t)nrows, pad, tail);
        const uint64_t per_frame = (uint64_t)hp.fs * sizeof(float2);
        const uint64_t fixed = d.base_size * sizeof(double2);
        if (fixed + 32 * per_frame > max_bytes)
            fail(SMIP_ERR_NO_MEMORY, "autotune: the bispectrum and 32 spectra exceed max_bytes");
        int max_batch = 32;
        for (int b : kBatches)
            if (fixed + (uint64_t)b * per_frame <= max_bytes)
                max_batch = b;

        ensure_device(ctx);
        hipStream_t s = ctx->stream;
        float2* H = static_cast<float2*>(dev_malloc((size_t)max_batch * per_frame));
        double2* B = nullptr;
        try {
            B = static_cast<double2*>(dev_malloc(sizeof(double2) * d.base_size));
        } catch (...) {
            dev_free(H);
            throw;
        }
        std::vector<TilePlan*> plans;
        auto cleanup = [&] {
            for (auto* t : plans)
                tile_plan_destroy(t);
            dev_free(H);
            dev_free(B);
        };
        try {
            const int64_t n = (int64_t)max_batch * hp.fs;
            hipLaunchKernelGGL(k_fill_spectra, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, H, n, hp.rs,
                hp.pad, hp.nrows);
            SMIP_HIP(hipGetLastError());
            SMIP_HIP(hipMemsetAsync(B, 0, sizeof(double2) * d.base_size, s));
            Timer tm;
            // ms per launch of nf frames with tile plan tp (one warm-up, best of two)
            auto time_launch = [&](TilePlan* tp, int nf) {
                launch_accum_tile(tp, hp, H, nf, B, s);
                float best = 1e30f;
                for (int rep = 0; rep < 2; ++rep) {
                    SMIP_HIP(hipEventRecord(tm.a, s));
                    launch_accum_tile(tp, hp, H, nf, B, s);
                    SMIP_HIP(hipEventRecord(tm.b, s));
                    SMIP_HIP(hipEventSynchronize(tm.b));
                    float ms = 0.f;
                    SMIP_HIP(hipEventElapsedTime(&ms, tm.a, tm.b));
                    best = std::min(best, ms);
                }
                return (double)best;
            };
            const int probe = std::min(max_batch, 256);
            int best_tile = kTiles[0];
            double best_ms = 1e30;
            TilePlan* best_plan = nullptr;
            for (int t : kTiles) {
                TilePlan* tp = tile_plan_create(d, L, hp, t, s);
                plans.push_back(tp);
                const double ms = time_launch(tp, probe);
                if (ms < best_ms) {
                    best_ms = ms;
                    best_tile = t;
                    best_plan = tp;
                }
            }
            std::vector<std::pair<int, double>> rate; // (batch, frames/s)
            double best_rate = 0.0;
            for (int b : kBatches) {
                if (b > max_batch)
                    break;
                const double ms = time_launch(best_plan, b);
                const double r = b / (ms * 1e-3);
                rate.emplace_back(b, r);
                best_rate = std::max(best_rate, r);
            }
            int pick = rate.back().first;
            double pick_rate = rate.back().second;
            for (auto& br : rate)
                if (br.second >= 0.97 * best_rate) {
                    pick = br.first;
                    pick_rate = br.second;
                    break;
                }
            smip_tune_result r {};
            r.vy_tile = best_tile;
            r.frames_per_batch = pick;
            r.frames_per_s = pick_rate;
            r.tile_tflops = pick_rate * 14.0 * (double)count_triples(L) * 1e-12;
            r.candidates = (int32_t)(std::size(kTiles) + rate.size());
            {
                std::lock_guard<std::mutex> lk(g_mu);
                g_cache[{ ncols, nrows, bdepth }] = r;
            }
            *out = r;
        } catch (...) {
            cleanup();
            throw;
        }
        cleanup();
    });
}

extern "C" SMIP_API int smip_autotune_forget(void)
{
    return guarded([&] {
        std::lock_guard<std::mutex> lk(g_mu);
        g_cache.clear();
    });
}
