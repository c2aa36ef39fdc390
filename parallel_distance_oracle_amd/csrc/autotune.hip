// autotune.hip -- picks build knobs by timing trial builds.
//
// None of the knobs changes the labels: BATCH_SIZE only delays prunes that
// cannot fire earlier (label sets are identical for every batch size), the
// grid sizes only change how work is spread, and both bit-parallel variants
// compute the same records.  So the tuner optimises wall time alone.
#include <chrono>
#include <vector>

#include "internal.hpp"

namespace {

double time_build(pado_gpu_ctx *ctx, const pado_gpu_graph *g, const pado_gpu_build_cfg &c, pado_gpu_status *st)
{
    pado_gpu_index *idx = nullptr;
    (void)hipStreamSynchronize(ctx->stream);
    const auto t0 = std::chrono::steady_clock::now();
    *st = pado_gpu_build(ctx, g, &c, &idx);
    (void)hipStreamSynchronize(ctx->stream);
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    pado_gpu_index_free(idx);
    return ms;
}

}  // namespace

extern "C" pado_gpu_status pado_gpu_autotune(pado_gpu_ctx *ctx, const pado_gpu_graph *graph,
                                             const pado_gpu_build_cfg *start, uint32_t max_trials,
                                             pado_gpu_build_cfg *best, double *best_ms, uint32_t *trials_out)
{
    pado::clear_error();
    if (!ctx || !graph || !best) {
        pado::set_error("null argument");
        return PADO_GPU_ERR_ARG;
    }
    pado_gpu_build_cfg cur{};
    if (start) cur = *start;
    if (!cur.batch_size) cur.batch_size = pado::kBatch;
    if (!cur.push_blocks) cur.push_blocks = pado::kPushGrid;
    if (!cur.add_blocks) cur.add_blocks = pado::kAddGrid;
    if (!cur.iters_per_sync) cur.iters_per_sync = 8;
    if (!max_trials) max_trials = 20;
    uint32_t trials = 0;
    pado_gpu_status st = PADO_GPU_OK;
    double cur_ms = time_build(ctx, graph, cur, &st);
    ++trials;
    if (st != PADO_GPU_OK) return st;
    // Coordinate search: one knob at a time, keep the best value.
    const std::vector<uint32_t> batches = {256, 512, 1024};
    const std::vector<uint32_t> pushes = {512, 1024, 2048, 4096};
    const std::vector<uint32_t> adds = {512, 1024, 2048};
    const std::vector<uint32_t> iters = {2, 4, 8, 16};
    const std::vector<uint32_t> flags = {0u, PADO_GPU_BUILD_STREAM, PADO_GPU_BUILD_BP_PER_SLOT,
                                         PADO_GPU_BUILD_STREAM | PADO_GPU_BUILD_BP_PER_SLOT};
    auto sweep = [&](uint32_t pado_gpu_build_cfg::*field, const std::vector<uint32_t> &vals) -> pado_gpu_status {
        for (uint32_t v : vals) {
            if (trials >= max_trials) return PADO_GPU_OK;
            if (cur.*field == v) continue;
            pado_gpu_build_cfg c = cur;
            c.*field = v;
            pado_gpu_status s = PADO_GPU_OK;
            const double ms = time_build(ctx, graph, c, &s);
            ++trials;
            if (s != PADO_GPU_OK) return s;
            if (ms < cur_ms) {
                cur_ms = ms;
                cur = c;
            }
        }
        return PADO_GPU_OK;
    };
    PADO_TRY(sweep(&pado_gpu_build_cfg::push_blocks, pushes));
    PADO_TRY(sweep(&pado_gpu_build_cfg::add_blocks, adds));
    PADO_TRY(sweep(&pado_gpu_build_cfg::batch_size, batches));
    PADO_TRY(sweep(&pado_gpu_build_cfg::iters_per_sync, iters));
    PADO_TRY(sweep(&pado_gpu_build_cfg::flags, flags));
    *best = cur;
    if (best_ms) *best_ms = cur_ms;
    if (trials_out) *trials_out = trials;
    return PADO_GPU_OK;
}
