// autotune.hip -- picks build knobs by timing trial builds.
//
// None of the knobs changes the labels: BATCH_SIZE only delays prunes that
// cannot fire earlier (label sets are identical for every batch size), the
// grid sizes, work-descriptor size and block shape only change how work is
// spread, and both bit-parallel variants compute the same records.  So the
// tuner optimises wall time alone.  Each configuration is timed as the
// median of three builds (run-to-run spread is a few percent), after one
// untimed build that takes the first-use costs (code-object loading, pool
// growth) out of the comparison.
#include <algorithm>
#include <chrono>
#include <functional>
#include <vector>

#include "internal.hpp"

namespace {

constexpr int kReps = 3;

using TrialFn = std::function<pado_gpu_status(const pado_gpu_build_cfg &, double *ms)>;

pado_gpu_status median_ms(const TrialFn &once, const pado_gpu_build_cfg &c, double *out)
{
    double t[kReps];
    for (int r = 0; r < kReps; ++r) PADO_TRY(once(c, &t[r]));
    std::sort(t, t + kReps);
    *out = t[kReps / 2];
    return PADO_GPU_OK;
}

struct Knob {
    uint32_t pado_gpu_build_cfg::*field;
    std::vector<uint32_t> vals;
};

// Coordinate search from `cur`: one knob at a time, keep the best value.
// `trials` counts configurations timed (each kReps builds).
pado_gpu_status coordinate_search(const TrialFn &once, const std::vector<Knob> &knobs, uint32_t max_trials,
                                  pado_gpu_build_cfg &cur, double &cur_ms, uint32_t &trials)
{
    double warm = 0.0;
    PADO_TRY(once(cur, &warm));
    PADO_TRY(median_ms(once, cur, &cur_ms));
    trials = 1;
    for (const Knob &k : knobs) {
        for (uint32_t v : k.vals) {
            if (trials >= max_trials) return PADO_GPU_OK;
            if (cur.*(k.field) == v) continue;
            pado_gpu_build_cfg c = cur;
            c.*(k.field) = v;
            double ms = 0.0;
            PADO_TRY(median_ms(once, c, &ms));
            ++trials;
            if (ms < cur_ms) {
                cur_ms = ms;
                cur = c;
            }
        }
    }
    return PADO_GPU_OK;
}

template <class F>
pado_gpu_status timed(pado_gpu_ctx *ctx, F &&build, double *ms)
{
    PADO_HIP(hipStreamSynchronize(ctx->stream));
    const auto t0 = std::chrono::steady_clock::now();
    pado_gpu_status s = build();
    (void)hipStreamSynchronize(ctx->stream);
    *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return s;
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
    const uint32_t n = graph->n;
    pado_gpu_build_cfg cur{};
    if (start) cur = *start;
    if (!cur.batch_size) cur.batch_size = pado::kBatch;
    if (!cur.push_blocks) cur.push_blocks = pado::default_push_grid(n);
    if (!cur.add_blocks) cur.add_blocks = pado::default_add_grid(n);
    if (!cur.iters_per_sync) cur.iters_per_sync = 8;
    if (!cur.push_chunk) cur.push_chunk = pado::kPushChunk;
    if (!cur.push_waves) cur.push_waves = 4;
    if (!max_trials) max_trials = 24;
    const TrialFn once = [&](const pado_gpu_build_cfg &c, double *ms) -> pado_gpu_status {
        pado_gpu_index *idx = nullptr;
        pado_gpu_status s = timed(ctx, [&] { return pado_gpu_build(ctx, graph, &c, &idx); }, ms);
        pado_gpu_index_free(idx);
        return s;
    };
    const uint32_t keep = cur.flags & PADO_GPU_BUILD_WIDE;
    std::vector<Knob> knobs = {
        {&pado_gpu_build_cfg::push_chunk, {4, 8, 16}},
        {&pado_gpu_build_cfg::push_waves, {2, 4, 8}},
        {&pado_gpu_build_cfg::push_blocks, {256, 512, 768, 1024, 2048}},
        {&pado_gpu_build_cfg::add_blocks, {256, 384, 512, 768, 1024}},
        {&pado_gpu_build_cfg::iters_per_sync, {2, 4, 8, 16}},
        {&pado_gpu_build_cfg::batch_size, {256, 512, 1024}},
        {&pado_gpu_build_cfg::flags, {keep, keep | PADO_GPU_BUILD_STREAM, keep | PADO_GPU_BUILD_BP_PER_SLOT}},
    };
    // The one-workgroup batch loop only pays off on small graphs (measured:
    // 7x faster at 1K vertices, 1.7x slower at 1.6K-vertex grids, 20-40x
    // slower at 20K-60K-vertex grids), so it is a candidate only there.
    if (n <= 4096) knobs.back().vals.push_back(keep | PADO_GPU_BUILD_DEVICE_LOOP);
    double cur_ms = 0.0;
    uint32_t trials = 0;
    PADO_TRY(coordinate_search(once, knobs, max_trials, cur, cur_ms, trials));
    *best = cur;
    if (best_ms) *best_ms = cur_ms;
    if (trials_out) *trials_out = trials;
    return PADO_GPU_OK;
}

extern "C" pado_gpu_status pado_gpu_autotune_multi(pado_gpu_ctx *ctx, uint32_t num_graphs, const uint32_t *n,
                                                   const uint64_t *off, const uint32_t *adj, uint32_t sample,
                                                   const pado_gpu_build_cfg *start, uint32_t max_trials,
                                                   pado_gpu_build_cfg *best, double *best_ms, uint32_t *trials_out)
{
    pado::clear_error();
    if (!ctx || !n || !off || !adj || !best || num_graphs == 0) {
        pado::set_error("null argument or no graphs");
        return PADO_GPU_ERR_ARG;
    }
    // Every stride-th graph, so the sample keeps the job's mix of kinds and
    // sizes (graphs of one kind often alternate).
    if (!sample || sample > num_graphs) sample = num_graphs;
    const uint32_t stride = num_graphs / sample;
    std::vector<uint32_t> sn;
    std::vector<uint64_t> soff;
    std::vector<uint32_t> sadj;
    uint64_t op = 0, ap = 0;
    for (uint32_t g = 0; g < num_graphs; ++g) {
        const uint64_t m = off[op + n[g]];
        if (g % stride == 0 && sn.size() < sample) {
            sn.push_back(n[g]);
            soff.insert(soff.end(), off + op, off + op + n[g] + 1);
            sadj.insert(sadj.end(), adj + ap, adj + ap + m);
        }
        op += (uint64_t)n[g] + 1;
        ap += m;
    }
    if (sadj.empty()) sadj.push_back(0);
    pado_gpu_build_cfg cur{};
    if (start) cur = *start;
    if (!cur.iters_per_sync) cur.iters_per_sync = 8;
    if (!cur.push_chunk) cur.push_chunk = pado::kPushChunk;
    if (!cur.push_waves) cur.push_waves = 4;
    if (!max_trials) max_trials = 12;
    const TrialFn once = [&](const pado_gpu_build_cfg &c, double *ms) -> pado_gpu_status {
        pado_gpu_multi *mi = nullptr;
        pado_gpu_status s = timed(ctx,
                                  [&] {
                                      return pado_gpu_build_multi(ctx, (uint32_t)sn.size(), sn.data(), soff.data(),
                                                                  sadj.data(), &c, &mi);
                                  },
                                  ms);
        pado_gpu_multi_free(mi);
        return s;
    };
    // Grids stay per graph (0: sized from each graph's vertex count); the
    // knobs that matter for many small builds are the iteration blocks and
    // the push shape.
    const std::vector<Knob> knobs = {
        {&pado_gpu_build_cfg::iters_per_sync, {4, 8, 16, 32}},
        {&pado_gpu_build_cfg::push_chunk, {4, 8, 16}},
        {&pado_gpu_build_cfg::push_waves, {2, 4, 8}},
        {&pado_gpu_build_cfg::batch_size, {256, 512, 1024}},
    };
    double cur_ms = 0.0;
    uint32_t trials = 0;
    PADO_TRY(coordinate_search(once, knobs, max_trials, cur, cur_ms, trials));
    *best = cur;
    if (best_ms) *best_ms = cur_ms;
    if (trials_out) *trials_out = trials;
    return PADO_GPU_OK;
}
