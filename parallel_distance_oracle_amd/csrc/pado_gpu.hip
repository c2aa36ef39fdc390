// pado_gpu.hip -- C ABI of the PADO GPU library: contexts, graph upload, index
// build / export / import / order, queries and index files.
//
// Each entry point's reference counterpart is named in include/pado_gpu.h.
#include <chrono>
#include <cstring>
#include <fstream>
#include <mutex>

#include "internal.hpp"

// Staging memory of pado_gpu_query_noalloc: one call's pair ids (2 x u32) and
// answers (u8), PADO_GPU_NOALLOC_MAX_PAIRS pairs at most (9 MiB).  Every GPU the library runs on
// holds its own copy until the process ends.  In host code this name is only
// a stand-in for the device symbol: kernels and copies use the address
// hipGetSymbolAddress returns for the current device (g_scratch below).
static __device__ __align__(256) uint8_t scratch[PADO_GPU_NOALLOC_MAX_PAIRS * 9ull];

namespace pado {

namespace {
// Device address of `scratch` per device, resolved once per device by
// pado_gpu_ctx_create with that device current.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch[kMaxDevices] = {};
}  // namespace

static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

pado_gpu_status launch_query(const pado_gpu_index *idx, bool rank_space, const uint32_t *qa, const uint32_t *qb,
                             uint64_t k, uint8_t *out, bool ref_compat, hipStream_t st);
pado_gpu_status launch_wquery(const pado_gpu_index *idx, bool rank_space, const uint32_t *qa, const uint32_t *qb,
                              uint64_t k, uint32_t *out, bool ref_compat, hipStream_t st);
pado_gpu_status launch_query_wide(const pado_gpu_index *idx, bool rank_space, const uint32_t *qa, const uint32_t *qb,
                                  uint64_t k, uint16_t *out, bool ref_compat, hipStream_t st);
pado_gpu_status launch_query_pair(const pado_gpu_index *idx, bool rank_space, uint32_t a, uint32_t b,
                                  bool ref_compat, uint32_t *out_dev, hipStream_t st);

namespace {

// order_labels (PARA:2198-2281): Index[rank2id[v]] = L[v].  Here new position
// of original id x is x itself, its source is rank-space vertex rank[x].
__global__ __launch_bounds__(256) void k_order_counts(uint32_t n, const uint32_t *rank, const uint64_t *loff,
                                                      uint32_t *cnt)
{
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < n; x += stride) {
        uint32_t v = rank[x];
        cnt[x] = (uint32_t)(loff[v + 1] - loff[v]);
    }
}

__global__ __launch_bounds__(256) void k_order_copy(uint32_t n, const uint32_t *rank, const uint8_t *bpd,
                                                    const ulonglong2 *bps, const uint64_t *loff, const uint32_t *lab,
                                                    const uint32_t *wdist, uint8_t *qbpd, ulonglong2 *qbps,
                                                    const uint64_t *qoff, uint32_t *qlab, uint32_t *qwdist)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t nw = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t x = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; x < n; x += nw) {
        const uint32_t v = rank[x];
        if (bpd) {
            qbpd[(uint64_t)x * kBpdStride + lane] = bpd[(uint64_t)v * kBpdStride + lane];
            if (lane < kBP) qbps[(uint64_t)x * kBP + lane] = bps[(uint64_t)v * kBP + lane];
        }
        const uint64_t s = loff[v], dsto = qoff[x];
        const uint32_t L = (uint32_t)(loff[v + 1] - s);
        for (uint32_t j = lane; j < L; j += 64) {
            qlab[dsto + j] = lab[s + j];
            if (wdist) qwdist[dsto + j] = wdist[s + j];
        }
    }
}

__global__ void k_pack_labels(uint64_t total, const uint32_t *hub, const uint8_t *dist, uint32_t *lab)
{
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride)
        lab[i] = (hub[i] << 8) | dist[i];
}

struct Timer {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double ms() const
    {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

pado_gpu_status check_ctx(pado_gpu_ctx *ctx)
{
    if (!ctx) {
        set_error("null context");
        return PADO_GPU_ERR_ARG;
    }
    PADO_HIP(hipSetDevice(ctx->device));
    return PADO_GPU_OK;
}

// Builds the original-id query copy from the rank-space index.
pado_gpu_status order_index(pado_gpu_index *idx, const uint32_t *h_rank)
{
    hipStream_t st = idx->ctx->stream;
    const uint32_t n = idx->n;
    // rank must be a permutation of [0, n).
    std::vector<uint8_t> seen(n, 0);
    for (uint32_t x = 0; x < n; ++x) {
        if (h_rank[x] >= n || seen[h_rank[x]]) {
            set_error("rank is not a permutation of [0, n)");
            return PADO_GPU_ERR_ARG;
        }
        seen[h_rank[x]] = 1;
    }
    DevBuf<uint32_t> rank, cnt;
    PADO_TRY(rank.alloc(n));
    PADO_TRY(cnt.alloc(n));
    PADO_HIP(hipMemcpyAsync(rank.p, h_rank, sizeof(uint32_t) * n, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_order_counts, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, n, rank.p, idx->loff.p, cnt.p);
    PADO_TRY(idx->qoff.alloc((uint64_t)n + 1));
    PADO_TRY(scan_u32_to_u64(cnt.p, idx->qoff.p, n, st));
    const bool w = idx->weighted != 0;
    if (!w) {
        PADO_TRY(idx->qbpd.alloc((uint64_t)n * kBpdStride));
        PADO_TRY(idx->qbps.alloc((uint64_t)n * kBP * 2));
    } else {
        PADO_TRY(idx->qwdist.alloc(idx->num_labels));
    }
    PADO_TRY(idx->qlab.alloc(idx->num_labels));
    hipLaunchKernelGGL(k_order_copy, dim3(grid_for((uint64_t)n * 64, 256, 8192)), dim3(256), 0, st, n, rank.p,
                       w ? nullptr : idx->bpd.p, w ? nullptr : (const ulonglong2 *)idx->bps.p, idx->loff.p,
                       w ? idx->whub.p : idx->lab.p, w ? idx->wdist.p : nullptr, w ? nullptr : idx->qbpd.p,
                       w ? nullptr : (ulonglong2 *)idx->qbps.p, idx->qoff.p, idx->qlab.p,
                       w ? idx->qwdist.p : nullptr);
    PADO_HIP(hipStreamSynchronize(st));
    idx->ordered = 1;
    return PADO_GPU_OK;
}

}  // namespace
}  // namespace pado

using namespace pado;

namespace pado {

pado_gpu_status build_bp_phase(pado_gpu_graph *g, pado_gpu_index *idx, const pado_gpu_build_cfg &c,
                               DevBuf<uint8_t> &used)
{
    std::vector<uint8_t> h_used;
    Timer t0;
    // The wide domain always uses the multi-source sweep (it stops at level 254).
    pado_gpu_status s = (c.flags & PADO_GPU_BUILD_BP_PER_SLOT) && !idx->wide ? build_bp(g, idx, used, h_used)
                                                                            : build_bp_multi(g, idx, used, h_used);
    idx->ms_bp = t0.ms();
    if (s == PADO_GPU_OK && idx->wide && idx->bp_max_level >= kInf8) {
        // The reference never pushes from a bit-parallel root or selected
        // neighbour (PARA:1117-1119): every path through one is covered by
        // its slot.  With the slot BFS stopped at level 254 that no longer
        // holds, so here those vertices take part in the batch phase like
        // any other (they become hubs and push); the slots still prune
        // wherever both ends were reached.
        if (hipMemsetAsync(used.p, 0, g->n, g->ctx->stream) != hipSuccess) s = PADO_GPU_ERR_HIP;
        idx->bp_pass_through = 1;
    }
    return s;
}

}  // namespace pado

extern "C" {

const char *pado_gpu_last_error(void) { return g_last_error.c_str(); }

int pado_gpu_device_count(void)
{
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

pado_gpu_status pado_gpu_ctx_create(int device, pado_gpu_ctx **out)
{
    clear_error();
    if (!out) {
        set_error("null output pointer");
        return PADO_GPU_ERR_ARG;
    }
    *out = nullptr;
    int count = 0;
    PADO_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) {
        set_error("device " + std::to_string(device) + " out of range (" + std::to_string(count) + " devices)");
        return PADO_GPU_ERR_ARG;
    }
    PADO_HIP(hipSetDevice(device));
    if (device < kMaxDevices) {
        // The allocation-free query's staging array: its address on this
        // device, looked up once per device (the device is current here).
        std::lock_guard<std::mutex> g(g_scratch_mu);
        if (!g_scratch[device]) {
            void *ptr = nullptr;
            const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !ptr) {
                set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
                return PADO_GPU_ERR_HIP;
            }
            g_scratch[device] = ptr;
        }
    }
    auto *c = new pado_gpu_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->side_ev, hipEventDisableTiming) != hipSuccess) {
        if (c->stream) (void)hipStreamDestroy(c->stream);
        if (c->side) (void)hipStreamDestroy(c->side);
        delete c;
        set_error("stream or event creation failed");
        return PADO_GPU_ERR_HIP;
    }
    *out = c;
    return PADO_GPU_OK;
}

void pado_gpu_ctx_destroy(pado_gpu_ctx *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    if (ctx->side_ev) (void)hipEventDestroy(ctx->side_ev);
    delete ctx;
}

void *pado_gpu_ctx_stream(pado_gpu_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

pado_gpu_status pado_gpu_ctx_sync(pado_gpu_ctx *ctx)
{
    clear_error();
    PADO_TRY(check_ctx(ctx));
    PADO_HIP(hipStreamSynchronize(ctx->stream));
    return PADO_GPU_OK;
}

pado_gpu_status pado_gpu_graph_upload(pado_gpu_ctx *ctx, uint32_t n, uint64_t m_dir, const uint64_t *off,
                                      const uint32_t *adj, const uint8_t *weights, pado_gpu_graph **out)
{
    clear_error();
    PADO_TRY(check_ctx(ctx));
    if (!out || !off || (m_dir && !adj)) {
        set_error("null argument");
        return PADO_GPU_ERR_ARG;
    }
    *out = nullptr;
    if (n == 0 || n >= kMaxVertices) {
        set_error("vertex count must be in [1, 2^24)");
        return PADO_GPU_ERR_ARG;
    }
    if (m_dir >= (1ull << 32)) {
        set_error("directed edge count must be < 2^32");
        return PADO_GPU_ERR_ARG;
    }
    if (off[0] != 0 || off[n] != m_dir) {
        set_error("offsets must start at 0 and end at m_dir");
        return PADO_GPU_ERR_ARG;
    }
    for (uint32_t v = 0; v < n; ++v)
        if (off[v + 1] < off[v]) {
            set_error("offsets must be non-decreasing");
            return PADO_GPU_ERR_ARG;
        }
    for (uint64_t e = 0; e < m_dir; ++e)
        if (adj[e] >= n) {
            set_error("adjacency id out of range");
            return PADO_GPU_ERR_ARG;
        }
    // id_transfer sorts every list descending (GR:192); the push early exit
    // and the bit-parallel neighbour scan rely on it.
    for (uint32_t v = 0; v < n; ++v)
        for (uint64_t e = off[v] + 1; e < off[v + 1]; ++e)
            if (adj[e] > adj[e - 1]) {
                set_error("adjacency lists must be sorted descending (rank space, after id_transfer)");
                return PADO_GPU_ERR_ARG;
            }
    auto *g = new pado_gpu_graph();
    g->ctx = ctx;
    g->n = n;
    g->m = m_dir;
    g->h_off.assign(off, off + n + 1);
    g->h_adj.assign(adj, adj + m_dir);
    if (weights) {
        // Pruning is exact only for positive lengths: a zero-length edge makes
        // two vertices equidistant to every hub.
        for (uint64_t e = 0; e < m_dir; ++e)
            if (weights[e] == 0) {
                delete g;
                set_error("edge weights must be in [1, 255]");
                return PADO_GPU_ERR_ARG;
            }
        g->h_w.assign(weights, weights + m_dir);
        g->weighted = true;
    }
    pado_gpu_status s = g->off.alloc((uint64_t)n + 1);
    if (s == PADO_GPU_OK) s = g->adj.alloc(m_dir);
    if (s == PADO_GPU_OK && weights) s = g->w.alloc(m_dir);
    if (s != PADO_GPU_OK) {
        delete g;
        return s;
    }
    hipStream_t st = ctx->stream;
    if (hipMemcpyAsync(g->off.p, off, sizeof(uint64_t) * (n + 1), hipMemcpyHostToDevice, st) != hipSuccess ||
        (m_dir && hipMemcpyAsync(g->adj.p, adj, sizeof(uint32_t) * m_dir, hipMemcpyHostToDevice, st) != hipSuccess) ||
        (weights && hipMemcpyAsync(g->w.p, weights, m_dir, hipMemcpyHostToDevice, st) != hipSuccess) ||
        hipStreamSynchronize(st) != hipSuccess) {
        delete g;
        set_error("graph upload failed");
        return PADO_GPU_ERR_HIP;
    }
    *out = g;
    return PADO_GPU_OK;
}

void pado_gpu_graph_free(pado_gpu_graph *g) { delete g; }

pado_gpu_status pado_gpu_build(pado_gpu_ctx *ctx, const pado_gpu_graph *graph, const pado_gpu_build_cfg *cfg,
                               pado_gpu_index **out)
{
    clear_error();
    PADO_TRY(check_ctx(ctx));
    if (!graph || !out) {
        set_error("null argument");
        return PADO_GPU_ERR_ARG;
    }
    *out = nullptr;
    pado_gpu_build_cfg c{};
    c.batch_size = kBatch;
    if (cfg) c = *cfg;
    if (c.batch_size == 0) c.batch_size = kBatch;
    auto *idx = new pado_gpu_index();
    idx->ctx = ctx;
    idx->n = graph->n;
    idx->wide = (c.flags & PADO_GPU_BUILD_WIDE) && !graph->weighted ? 1u : 0u;
    if (idx->wide && graph->n > kWideMaxVertices) {
        delete idx;
        set_error("graph too large for the wide hop domain (at most 2^20 vertices)");
        return PADO_GPU_ERR_DOMAIN;
    }
    pado_gpu_graph *g = const_cast<pado_gpu_graph *>(graph);
    pado_gpu_status s = PADO_GPU_OK;
    if (g->weighted) {
        // WeightedParaVertexCentricPLL<512> (WPARA:1231-1566): no
        // bit-parallel labels; batch_size above 512 is clamped to 512.
        Timer t1;
        s = build_weighted(g, idx, c.batch_size);
        idx->ms_batch = t1.ms();
    } else {
        DevBuf<uint8_t> used;
        s = build_bp_phase(g, idx, c, used);
        if (s == PADO_GPU_OK) {
            Timer t1;
            s = build_batches(g, idx, used, c);
            idx->ms_batch = t1.ms();
        }
    }
    if (s != PADO_GPU_OK) {
        delete idx;
        return s;
    }
    *out = idx;
    return PADO_GPU_OK;
}

pado_gpu_status pado_gpu_index_get_info(const pado_gpu_index *idx, pado_gpu_index_info *info)
{
    clear_error();
    if (!idx || !info) {
        set_error("null argument");
        return PADO_GPU_ERR_ARG;
    }
    std::memset(info, 0, sizeof(*info));
    info->num_v = idx->n;
    info->weighted = idx->weighted;
    info->num_labels = idx->num_labels;
    info->ordered = idx->ordered;
    info->iterations = idx->iterations;
    info->ms_bp = idx->ms_bp;
    info->ms_batch = idx->ms_batch;
    info->bp_max_level = idx->bp_max_level;
    for (uint32_t i = 0; i < kBP; ++i) info->bp_roots[i] = idx->bp_roots[i];
    info->pushes = idx->counters.pushes;
    info->ind_pass = idx->counters.ind_pass;
    info->bp_pruned = idx->counters.bp_pruned;
    info->cands = idx->counters.cands;
    info->dq_pruned = idx->counters.dq_pruned;
    info->dq_rows = idx->counters.dq_rows;
    info->actives = idx->counters.actives;
    info->edges = idx->counters.edges;
    info->tails = idx->counters.tails;
    info->dq_labels = idx->counters.dq_labels;
    info->labels_appended = idx->counters.labels;
    info->ms_push = idx->ktimes.ms_push;
    info->ms_add = idx->ktimes.ms_add;
    info->ms_bp_levels = idx->ktimes.ms_bp;
    info->n_push = idx->ktimes.n_push;
    info->n_add = idx->ktimes.n_add;
    info->n_bp_levels = idx->ktimes.n_bp;
    info->dist_bits = idx->wide ? Hop<true>::kBits : Hop<false>::kBits;
    info->pull_iters = (uint32_t)idx->counters.pulls;
    return PADO_GPU_OK;
}

}  // extern "C"

namespace {

// Canonical rank-space export; exactly one of d8 / d16 (or neither) receives
// the label distances.
pado_gpu_status export_labels(pado_gpu_ctx *ctx, const pado_gpu_index *idx, uint8_t *bp_dist, uint64_t *bp_sets,
                              uint64_t *label_off, uint32_t *hubs, uint8_t *d8, uint16_t *d16)
{
    PADO_TRY(check_ctx(ctx));
    if (!idx) {
        set_error("null index");
        return PADO_GPU_ERR_ARG;
    }
    if (idx->weighted) {
        set_error("weighted index: use pado_gpu_windex_export");
        return PADO_GPU_ERR_STATE;
    }
    if (idx->wide && d8) {
        set_error("wide-domain index: its distances need pado_gpu_index_export_wide");
        return PADO_GPU_ERR_STATE;
    }
    const uint32_t B = idx->wide ? Hop<true>::kBits : Hop<false>::kBits;
    const uint32_t M = (1u << B) - 1u;
    if (!idx->loff.p) {
        set_error("index has no rank-space labels (loaded from an ordered file)");
        return PADO_GPU_ERR_STATE;
    }
    const uint32_t n = idx->n;
    hipStream_t st = ctx->stream;
    if (bp_dist) {
        std::vector<uint8_t> tmp((uint64_t)n * kBpdStride);
        PADO_HIP(hipMemcpyAsync(tmp.data(), idx->bpd.p, tmp.size(), hipMemcpyDeviceToHost, st));
        PADO_HIP(hipStreamSynchronize(st));
        for (uint32_t v = 0; v < n; ++v) std::memcpy(bp_dist + (uint64_t)v * kBP, &tmp[(uint64_t)v * kBpdStride], kBP);
    }
    if (bp_sets) PADO_HIP(hipMemcpyAsync(bp_sets, idx->bps.p, sizeof(uint64_t) * n * kBP * 2, hipMemcpyDeviceToHost, st));
    if (label_off) PADO_HIP(hipMemcpyAsync(label_off, idx->loff.p, sizeof(uint64_t) * (n + 1), hipMemcpyDeviceToHost, st));
    if (hubs || d8 || d16) {
        std::vector<uint32_t> lab(idx->num_labels);
        if (idx->num_labels)
            PADO_HIP(hipMemcpyAsync(lab.data(), idx->lab.p, sizeof(uint32_t) * idx->num_labels, hipMemcpyDeviceToHost, st));
        PADO_HIP(hipStreamSynchronize(st));
        for (uint64_t i = 0; i < idx->num_labels; ++i) {
            if (hubs) hubs[i] = lab[i] >> B;
            if (d8) d8[i] = (uint8_t)(lab[i] & M);
            if (d16) d16[i] = (uint16_t)(lab[i] & M);
        }
    }
    PADO_HIP(hipStreamSynchronize(st));
    return PADO_GPU_OK;
}

}  // namespace

extern "C" {

pado_gpu_status pado_gpu_index_export(pado_gpu_ctx *ctx, const pado_gpu_index *idx, uint8_t *bp_dist,
                                      uint64_t *bp_sets, uint64_t *label_off, uint32_t *hubs, uint8_t *dists)
{
    clear_error();
    // A wide index is refused only if its distances are asked for.
    return export_labels(ctx, idx, bp_dist, bp_sets, label_off, hubs, dists, nullptr);
}

pado_gpu_status pado_gpu_index_export_wide(pado_gpu_ctx *ctx, const pado_gpu_index *idx, uint8_t *bp_dist,
                                           uint64_t *bp_sets, uint64_t *label_off, uint32_t *hubs, uint16_t *dists)
{
    clear_error();
    return export_labels(ctx, idx, bp_dist, bp_sets, label_off, hubs, nullptr, dists);
}

pado_gpu_status pado_gpu_index_import(pado_gpu_ctx *ctx, uint32_t n, const uint8_t *bp_dist, const uint64_t *bp_sets,
                                      const uint64_t *label_off, const uint32_t *hubs, const uint8_t *dists,
                                      pado_gpu_index **out)
{
    clear_error();
    PADO_TRY(check_ctx(ctx));
    if (!out || !bp_dist || !bp_sets || !label_off || n == 0 || n >= kMaxVertices) {
        set_error("bad argument");
        return PADO_GPU_ERR_ARG;
    }
    *out = nullptr;
    const uint64_t total = label_off[n];
    if (total && (!hubs || !dists)) {
        set_error("null label arrays");
        return PADO_GPU_ERR_ARG;
    }
    for (uint32_t v = 0; v < n; ++v) {
        if (label_off[v + 1] < label_off[v]) {
            set_error("label offsets must be non-decreasing");
            return PADO_GPU_ERR_ARG;
        }
        for (uint64_t i = label_off[v]; i < label_off[v + 1]; ++i) {
            if (hubs[i] >= n || (i > label_off[v] && hubs[i] <= hubs[i - 1])) {
                set_error("hubs must be < n and strictly ascending per vertex");
                return PADO_GPU_ERR_ARG;
            }
        }
    }
    hipStream_t st = ctx->stream;
    auto *idx = new pado_gpu_index();
    idx->ctx = ctx;
    idx->n = n;
    idx->num_labels = total;
    std::vector<uint8_t> tmp((uint64_t)n * kBpdStride, kInf8);
    for (uint32_t v = 0; v < n; ++v) std::memcpy(&tmp[(uint64_t)v * kBpdStride], bp_dist + (uint64_t)v * kBP, kBP);
    pado_gpu_status s = PADO_GPU_OK;
    DevBuf<uint32_t> dh;
    DevBuf<uint8_t> dd;
    if ((s = idx->bpd.alloc(tmp.size())) != PADO_GPU_OK || (s = idx->bps.alloc((uint64_t)n * kBP * 2)) != PADO_GPU_OK ||
        (s = idx->loff.alloc((uint64_t)n + 1)) != PADO_GPU_OK || (s = idx->lab.alloc(total)) != PADO_GPU_OK ||
        (s = dh.alloc(total)) != PADO_GPU_OK || (s = dd.alloc(total)) != PADO_GPU_OK) {
        delete idx;
        return s;
    }
    if (hipMemcpyAsync(idx->bpd.p, tmp.data(), tmp.size(), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(idx->bps.p, bp_sets, sizeof(uint64_t) * n * kBP * 2, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(idx->loff.p, label_off, sizeof(uint64_t) * (n + 1), hipMemcpyHostToDevice, st) != hipSuccess ||
        (total && hipMemcpyAsync(dh.p, hubs, sizeof(uint32_t) * total, hipMemcpyHostToDevice, st) != hipSuccess) ||
        (total && hipMemcpyAsync(dd.p, dists, total, hipMemcpyHostToDevice, st) != hipSuccess)) {
        delete idx;
        set_error("index import copy failed");
        return PADO_GPU_ERR_HIP;
    }
    if (total)
        hipLaunchKernelGGL(k_pack_labels, dim3(grid_for(total, 256, 8192)), dim3(256), 0, st, total, dh.p, dd.p,
                           idx->lab.p);
    if (hipStreamSynchronize(st) != hipSuccess) {
        delete idx;
        set_error("index import failed");
        return PADO_GPU_ERR_HIP;
    }
    *out = idx;
    return PADO_GPU_OK;
}

pado_gpu_status pado_gpu_index_order(pado_gpu_ctx *ctx, pado_gpu_index *idx, const uint32_t *rank, uint32_t n)
{
    clear_error();
    PADO_TRY(check_ctx(ctx));
    if (!idx || !rank || n != idx->n) {
        set_error("rank must have num_v entries");
        return PADO_GPU_ERR_ARG;
    }
    return order_index(idx, rank);
}

}  // extern "C"

namespace {

// Host-array queries move 8 bytes of ids in and 1 (or 4) bytes of answers out
// per pair.  In chunks of kHostChunk pairs alternating between the query
// stream and the context's side stream, chunk i + 1's copy in runs while
// chunk i's kernel does, and chunk i's answers are copied out after chunk
// i + 1 is queued.  `launch(da, db, n, dout, stream)` queues one chunk's
// kernel; es = answer bytes.
constexpr uint64_t kHostChunk = 4ull << 20;

template <typename Launch>
pado_gpu_status host_query_chunked(pado_gpu_ctx *ctx, hipStream_t st, const uint32_t *a, const uint32_t *b,
                                   uint64_t k, void *out, uint32_t es, uint32_t *da, uint32_t *db, uint8_t *dout,
                                   Launch launch)
{
    hipStream_t ss[2] = {st, ctx->side};
    // The side stream starts after whatever was queued on st before the call.
    PADO_HIP(hipEventRecord(ctx->side_ev, st));
    PADO_HIP(hipStreamWaitEvent(ctx->side, ctx->side_ev, 0));
    uint8_t *hout = static_cast<uint8_t *>(out);
    uint64_t prev0 = 0, prevn = 0;
    hipStream_t prevs = st;
    for (uint64_t c0 = 0, i = 0; c0 < k; c0 += kHostChunk, ++i) {
        const uint64_t n = k - c0 < kHostChunk ? k - c0 : kHostChunk;
        hipStream_t s = ss[i & 1];
        PADO_HIP(hipMemcpyAsync(da + c0, a + c0, sizeof(uint32_t) * n, hipMemcpyHostToDevice, s));
        PADO_HIP(hipMemcpyAsync(db + c0, b + c0, sizeof(uint32_t) * n, hipMemcpyHostToDevice, s));
        PADO_TRY(launch(da + c0, db + c0, n, dout + c0 * es, s));
        if (prevn)
            PADO_HIP(hipMemcpyAsync(hout + prev0 * es, dout + prev0 * es, prevn * es, hipMemcpyDeviceToHost, prevs));
        prev0 = c0;
        prevn = n;
        prevs = s;
    }
    PADO_HIP(hipMemcpyAsync(hout + prev0 * es, dout + prev0 * es, prevn * es, hipMemcpyDeviceToHost, prevs));
    PADO_HIP(hipStreamSynchronize(ctx->side));
    PADO_HIP(hipStreamSynchronize(st));
    return PADO_GPU_OK;
}

}  // namespace

extern "C" {

pado_gpu_status pado_gpu_query(pado_gpu_ctx *ctx, pado_gpu_index *idx, const uint32_t *a, const uint32_t *b, uint64_t k,
                               uint8_t *out, uint32_t flags, void *stream)
{
    clear_error();
    PADO_TRY(check_ctx(ctx));
    if (!idx || (k && (!a || !b || !out))) {
        set_error("null argument");
        return PADO_GPU_ERR_ARG;
    }
    if (idx->weighted) {
        set_error("weighted index: use pado_gpu_wquery");
        return PADO_GPU_ERR_STATE;
    }
    if (idx->wide) {
        set_error("wide-domain index: its distances need pado_gpu_query_wide");
        return PADO_GPU_ERR_STATE;
    }
    const bool rank_space = (flags & PADO_GPU_QUERY_RANK_SPACE) != 0;
    if (!rank_space && !idx->ordered) {
        set_error("query before order_labels: call pado_gpu_index_order or pass PADO_GPU_QUERY_RANK_SPACE");
        return PADO_GPU_ERR_STATE;
    }
    if (rank_space && !idx->loff.p) {
        set_error("index has no rank-space labels");
        return PADO_GPU_ERR_STATE;
    }
    if (k == 0) return PADO_GPU_OK;
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    const bool ref = (flags & PADO_GPU_QUERY_REF_COMPAT) != 0;
    if (flags & PADO_GPU_QUERY_DEVICE_PTRS) return launch_query(idx, rank_space, a, b, k, out, ref, st);
    if (idx->qa.n < k) {
        PADO_TRY(idx->qa.alloc(k));
        PADO_TRY(idx->qb.alloc(k));
        PADO_TRY(idx->qout.alloc(k));
    }
    return host_query_chunked(ctx, st, a, b, k, out, 1u, idx->qa.p, idx->qb.p, idx->qout.p,
                              [&](const uint32_t *da, const uint32_t *db, uint64_t n, uint8_t *dout, hipStream_t s) {
                                  return launch_query(idx, rank_space, da, db, n, dout, ref, s);
                              });
}

pado_gpu_status pado_gpu_query_pair(pado_gpu_ctx *ctx, pado_gpu_index *idx, uint32_t a, uint32_t b, uint32_t flags,
                                    uint32_t *out)
{
    clear_error();
    PADO_TRY(check_ctx(ctx));
    if (!idx || !out) {
        set_error("null argument");
        return PADO_GPU_ERR_ARG;
    }
    const bool rank_space = (flags & PADO_GPU_QUERY_RANK_SPACE) != 0;
    if (!rank_space && !idx->ordered) {
        set_error("query before order_labels: call pado_gpu_index_order or pass PADO_GPU_QUERY_RANK_SPACE");
        return PADO_GPU_ERR_STATE;
    }
    if (rank_space && !idx->loff.p) {
        set_error("index has no rank-space labels");
        return PADO_GPU_ERR_STATE;
    }
    std::lock_guard<std::mutex> g(idx->pair_mu);
    if (!idx->pair_out) {
        std::lock_guard<std::recursive_mutex> cg(capture_mutex());
        uint32_t *h = nullptr;
        void *d = nullptr;
        if (hipHostMalloc((void **)&h, sizeof(uint32_t), hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
            set_error("pinned answer slot allocation failed");
            return PADO_GPU_ERR_NOMEM;
        }
        if (hipHostGetDevicePointer(&d, h, 0) != hipSuccess) {
            (void)hipHostFree(h);
            set_error("answer slot is not mapped for the device");
            return PADO_GPU_ERR_HIP;
        }
        idx->pair_out = h;
        idx->pair_out_dev = static_cast<uint32_t *>(d);
    }
    PADO_TRY(launch_query_pair(idx, rank_space, a, b, (flags & PADO_GPU_QUERY_REF_COMPAT) != 0, idx->pair_out_dev,
                               ctx->stream));
    PADO_HIP(hipStreamSynchronize(ctx->stream));
    *out = __atomic_load_n(idx->pair_out, __ATOMIC_ACQUIRE);
    return PADO_GPU_OK;
}

pado_gpu_status pado_gpu_query_wide(pado_gpu_ctx *ctx, pado_gpu_index *idx, const uint32_t *a, const uint32_t *b,
                                    uint64_t k, uint16_t *out, uint32_t flags, void *stream)
{
    clear_error();
    PADO_TRY(check_ctx(ctx));
    if (!idx || (k && (!a || !b || !out))) {
        set_error("null argument");
        return PADO_GPU_ERR_ARG;
    }
    if (idx->weighted) {
        set_error("weighted index: use pado_gpu_wquery");
        return PADO_GPU_ERR_STATE;
    }
    const bool ref = (flags & PADO_GPU_QUERY_REF_COMPAT) != 0;
    if (idx->wide && ref) {
        set_error("the reference has no wide hop domain: REF_COMPAT does not apply");
        return PADO_GPU_ERR_ARG;
    }
    const bool rank_space = (flags & PADO_GPU_QUERY_RANK_SPACE) != 0;
    if (!rank_space && !idx->ordered) {
        set_error("query before order_labels: call pado_gpu_index_order or pass PADO_GPU_QUERY_RANK_SPACE");
        return PADO_GPU_ERR_STATE;
    }
    if (rank_space && !idx->loff.p) {
        set_error("index has no rank-space labels");
        return PADO_GPU_ERR_STATE;
    }
    if (k == 0) return PADO_GPU_OK;
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    if (flags & PADO_GPU_QUERY_DEVICE_PTRS) return launch_query_wide(idx, rank_space, a, b, k, out, ref, st);
    if (idx->qa.n < k) {
        PADO_TRY(idx->qa.alloc(k));
        PADO_TRY(idx->qb.alloc(k));
    }
    if (idx->qout32.n * 2 < k) PADO_TRY(idx->qout32.alloc((k + 1) / 2));
    return host_query_chunked(ctx, st, a, b, k, out, 2u, idx->qa.p, idx->qb.p,
                              reinterpret_cast<uint8_t *>(idx->qout32.p),
                              [&](const uint32_t *da, const uint32_t *db, uint64_t n, uint8_t *dout, hipStream_t s) {
                                  return launch_query_wide(idx, rank_space, da, db, n,
                                                           reinterpret_cast<uint16_t *>(dout), ref, s);
                              });
}

pado_gpu_status pado_gpu_query_noalloc(pado_gpu_ctx *ctx, pado_gpu_index *idx, const uint32_t *a, const uint32_t *b,
                                       uint64_t k, uint8_t *out, uint32_t flags)
{
    clear_error();
    PADO_TRY(check_ctx(ctx));
    if (!idx || (k && (!a || !b || !out))) {
        set_error("null argument");
        return PADO_GPU_ERR_ARG;
    }
    if (flags & PADO_GPU_QUERY_DEVICE_PTRS) {
        set_error("pado_gpu_query_noalloc takes host arrays (pado_gpu_query with device pointers allocates nothing)");
        return PADO_GPU_ERR_ARG;
    }
    if (k > PADO_GPU_NOALLOC_MAX_PAIRS) {
        set_error("pado_gpu_query_noalloc: " + std::to_string(k) + " pairs exceed the staging maximum of " +
                  std::to_string(PADO_GPU_NOALLOC_MAX_PAIRS) + " per call");
        return PADO_GPU_ERR_ARG;
    }
    if (idx->weighted || idx->wide) {
        set_error("pado_gpu_query_noalloc answers u8 distances: not for weighted or wide-domain indexes");
        return PADO_GPU_ERR_STATE;
    }
    const bool rank_space = (flags & PADO_GPU_QUERY_RANK_SPACE) != 0;
    if (!rank_space && !idx->ordered) {
        set_error("query before order_labels: call pado_gpu_index_order or pass PADO_GPU_QUERY_RANK_SPACE");
        return PADO_GPU_ERR_STATE;
    }
    if (rank_space && !idx->loff.p) {
        set_error("index has no rank-space labels");
        return PADO_GPU_ERR_STATE;
    }
    if (k == 0) return PADO_GPU_OK;
    void *base = ctx->device < kMaxDevices ? g_scratch[ctx->device] : nullptr;
    if (!base) {
        set_error("no staging array for this device (contexts on device ids >= 64 have none)");
        return PADO_GPU_ERR_STATE;
    }
    // Layout: a ids [0, M), b ids [M, 2M), answers at byte 8M (M = maximum
    // pairs).  Nothing needs clearing: each call writes what it reads.
    uint32_t *da = static_cast<uint32_t *>(base);
    uint32_t *db = da + PADO_GPU_NOALLOC_MAX_PAIRS;
    uint8_t *dout = static_cast<uint8_t *>(base) + 8ull * PADO_GPU_NOALLOC_MAX_PAIRS;
    hipStream_t st = ctx->stream;
    PADO_HIP(hipMemcpyAsync(da, a, sizeof(uint32_t) * k, hipMemcpyHostToDevice, st));
    PADO_HIP(hipMemcpyAsync(db, b, sizeof(uint32_t) * k, hipMemcpyHostToDevice, st));
    PADO_TRY(launch_query(idx, rank_space, da, db, k, dout, (flags & PADO_GPU_QUERY_REF_COMPAT) != 0, st));
    PADO_HIP(hipMemcpyAsync(out, dout, k, hipMemcpyDeviceToHost, st));
    PADO_HIP(hipStreamSynchronize(st));
    return PADO_GPU_OK;
}

pado_gpu_status pado_gpu_windex_export(pado_gpu_ctx *ctx, const pado_gpu_index *idx, uint64_t *label_off,
                                       uint32_t *hubs, uint32_t *dists)
{
    clear_error();
    PADO_TRY(check_ctx(ctx));
    if (!idx) {
        set_error("null index");
        return PADO_GPU_ERR_ARG;
    }
    if (!idx->weighted) {
        set_error("unweighted index: use pado_gpu_index_export");
        return PADO_GPU_ERR_STATE;
    }
    hipStream_t st = ctx->stream;
    const uint64_t L = idx->num_labels;
    if (label_off)
        PADO_HIP(hipMemcpyAsync(label_off, idx->loff.p, sizeof(uint64_t) * (idx->n + 1), hipMemcpyDeviceToHost, st));
    if (hubs && L) PADO_HIP(hipMemcpyAsync(hubs, idx->whub.p, sizeof(uint32_t) * L, hipMemcpyDeviceToHost, st));
    if (dists && L) PADO_HIP(hipMemcpyAsync(dists, idx->wdist.p, sizeof(uint32_t) * L, hipMemcpyDeviceToHost, st));
    PADO_HIP(hipStreamSynchronize(st));
    return PADO_GPU_OK;
}

pado_gpu_status pado_gpu_wquery(pado_gpu_ctx *ctx, pado_gpu_index *idx, const uint32_t *a, const uint32_t *b,
                                uint64_t k, uint32_t *out, uint32_t flags, void *stream)
{
    clear_error();
    PADO_TRY(check_ctx(ctx));
    if (!idx || (k && (!a || !b || !out))) {
        set_error("null argument");
        return PADO_GPU_ERR_ARG;
    }
    if (!idx->weighted) {
        set_error("unweighted index: use pado_gpu_query");
        return PADO_GPU_ERR_STATE;
    }
    const bool rank_space = (flags & PADO_GPU_QUERY_RANK_SPACE) != 0;
    if (!rank_space && !idx->ordered) {
        set_error("query before order_labels: call pado_gpu_index_order or pass PADO_GPU_QUERY_RANK_SPACE");
        return PADO_GPU_ERR_STATE;
    }
    if (k == 0) return PADO_GPU_OK;
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    const bool ref = (flags & PADO_GPU_QUERY_REF_COMPAT) != 0;
    if (flags & PADO_GPU_QUERY_DEVICE_PTRS) return launch_wquery(idx, rank_space, a, b, k, out, ref, st);
    if (idx->qa.n < k) {
        PADO_TRY(idx->qa.alloc(k));
        PADO_TRY(idx->qb.alloc(k));
    }
    if (idx->qout32.n < k) PADO_TRY(idx->qout32.alloc(k));
    return host_query_chunked(ctx, st, a, b, k, out, 4u, idx->qa.p, idx->qb.p,
                              reinterpret_cast<uint8_t *>(idx->qout32.p),
                              [&](const uint32_t *da, const uint32_t *db, uint64_t n, uint8_t *dout, hipStream_t s) {
                                  return launch_wquery(idx, rank_space, da, db, n, reinterpret_cast<uint32_t *>(dout),
                                                       ref, s);
                              });
}

void pado_gpu_index_free(pado_gpu_index *idx) { delete idx; }

// Index file in the reference's format (UNP:1121-1180 writer, 1183-1224
// reader): u32 num_v, u32 BITPARALLEL_SIZE, then per ORIGINAL id: 50 x (u8 d,
// u64 S^-1, u64 S^0) unpadded, u32 size, size x (u32 rank-space hub, u8 dist).
pado_gpu_status pado_gpu_index_store(pado_gpu_ctx *ctx, const pado_gpu_index *idx, const char *filename,
                                     const uint32_t *rank, uint32_t n)
{
    clear_error();
    PADO_TRY(check_ctx(ctx));
    if (!idx || !filename || !rank || n != idx->n || idx->weighted) {
        set_error("bad argument");
        return PADO_GPU_ERR_ARG;
    }
    if (idx->wide) {
        set_error("the reference's index file holds u8 distances: a wide-domain index cannot be stored in it");
        return PADO_GPU_ERR_DOMAIN;
    }
    std::vector<uint8_t> bd((uint64_t)n * kBP);
    std::vector<uint64_t> bs((uint64_t)n * kBP * 2), off((uint64_t)n + 1);
    std::vector<uint32_t> hub(idx->num_labels);
    std::vector<uint8_t> dist(idx->num_labels);
    PADO_TRY(pado_gpu_index_export(ctx, idx, bd.data(), bs.data(), off.data(), hub.data(), dist.data()));
    std::ofstream f(filename, std::ios::binary);
    if (!f.is_open()) {
        set_error(std::string("cannot open file ") + filename);
        return PADO_GPU_ERR_IO;
    }
    const uint32_t bp = kBP;
    f.write((const char *)&n, 4);
    f.write((const char *)&bp, 4);
    for (uint32_t x = 0; x < n; ++x) {
        const uint32_t v = rank[x];
        if (v >= n) {
            set_error("rank out of range");
            return PADO_GPU_ERR_ARG;
        }
        for (uint32_t i = 0; i < kBP; ++i) {
            f.write((const char *)&bd[(uint64_t)v * kBP + i], 1);
            f.write((const char *)&bs[((uint64_t)v * kBP + i) * 2], 8);
            f.write((const char *)&bs[((uint64_t)v * kBP + i) * 2 + 1], 8);
        }
        const uint32_t sz = (uint32_t)(off[v + 1] - off[v]);
        f.write((const char *)&sz, 4);
        for (uint64_t i = off[v]; i < off[v + 1]; ++i) {
            f.write((const char *)&hub[i], 4);
            f.write((const char *)&dist[i], 1);
        }
    }
    if (!f.good()) {
        set_error(std::string("write failed: ") + filename);
        return PADO_GPU_ERR_IO;
    }
    return PADO_GPU_OK;
}

pado_gpu_status pado_gpu_index_load(pado_gpu_ctx *ctx, const char *filename, pado_gpu_index **out)
{
    clear_error();
    PADO_TRY(check_ctx(ctx));
    if (!filename || !out) {
        set_error("bad argument");
        return PADO_GPU_ERR_ARG;
    }
    *out = nullptr;
    std::ifstream f(filename, std::ios::binary);
    if (!f.is_open()) {
        set_error(std::string("cannot open file ") + filename);
        return PADO_GPU_ERR_IO;
    }
    uint32_t n = 0, bp = 0;
    f.read((char *)&n, 4);
    f.read((char *)&bp, 4);
    if (!f.good() || bp != kBP || n == 0 || n >= kMaxVertices) {
        set_error("bad index file header");
        return PADO_GPU_ERR_IO;
    }
    std::vector<uint8_t> bd((uint64_t)n * kBP);
    std::vector<uint64_t> bs((uint64_t)n * kBP * 2), off((uint64_t)n + 1, 0);
    std::vector<uint32_t> hub;
    std::vector<uint8_t> dist;
    for (uint32_t x = 0; x < n; ++x) {
        for (uint32_t i = 0; i < kBP; ++i) {
            f.read((char *)&bd[(uint64_t)x * kBP + i], 1);
            f.read((char *)&bs[((uint64_t)x * kBP + i) * 2], 8);
            f.read((char *)&bs[((uint64_t)x * kBP + i) * 2 + 1], 8);
        }
        uint32_t sz = 0;
        f.read((char *)&sz, 4);
        if (!f.good()) {
            set_error("truncated index file");
            return PADO_GPU_ERR_IO;
        }
        for (uint32_t i = 0; i < sz; ++i) {
            uint32_t h = 0;
            uint8_t d = 0;
            f.read((char *)&h, 4);
            f.read((char *)&d, 1);
            hub.push_back(h);
            dist.push_back(d);
        }
        off[x + 1] = hub.size();
    }
    if (!f.good()) {
        set_error("truncated index file");
        return PADO_GPU_ERR_IO;
    }
    // The file is in original-id order: import it as-is and mark it ordered.
    pado_gpu_index *idx = nullptr;
    PADO_TRY(pado_gpu_index_import(ctx, n, bd.data(), bs.data(), off.data(), hub.data(), dist.data(), &idx));
    std::swap(idx->qbpd.p, idx->bpd.p);
    std::swap(idx->qbpd.n, idx->bpd.n);
    std::swap(idx->qbps.p, idx->bps.p);
    std::swap(idx->qbps.n, idx->bps.n);
    std::swap(idx->qoff.p, idx->loff.p);
    std::swap(idx->qoff.n, idx->loff.n);
    std::swap(idx->qlab.p, idx->lab.p);
    std::swap(idx->qlab.n, idx->lab.n);
    idx->ordered = 1;
    *out = idx;
    return PADO_GPU_OK;
}

// Weighted index file (WPARA:1982-2018 writer, 2021-2055 reader): u32 num_v,
// then per ORIGINAL id: u32 size, size x (u32 rank-space hub, i16 dist),
// hubs ascending.
pado_gpu_status pado_gpu_windex_store(pado_gpu_ctx *ctx, const pado_gpu_index *idx, const char *filename,
                                      const uint32_t *rank, uint32_t n)
{
    clear_error();
    PADO_TRY(check_ctx(ctx));
    if (!idx || !filename || !rank || n != idx->n || !idx->weighted || !idx->loff.p) {
        set_error("bad argument (a weighted rank-space index and num_v rank entries are required)");
        return PADO_GPU_ERR_ARG;
    }
    std::vector<uint64_t> off((uint64_t)n + 1);
    std::vector<uint32_t> hub(idx->num_labels ? idx->num_labels : 1), dist(idx->num_labels ? idx->num_labels : 1);
    PADO_TRY(pado_gpu_windex_export(ctx, idx, off.data(), hub.data(), dist.data()));
    for (uint64_t i = 0; i < idx->num_labels; ++i)
        if (dist[i] >= 32767u) {
            set_error("a label distance does not fit the file's int16 field");
            return PADO_GPU_ERR_DOMAIN;
        }
    std::ofstream f(filename, std::ios::binary);
    if (!f.is_open()) {
        set_error(std::string("cannot open file ") + filename);
        return PADO_GPU_ERR_IO;
    }
    f.write((const char *)&n, 4);
    for (uint32_t x = 0; x < n; ++x) {
        const uint32_t v = rank[x];
        if (v >= n) {
            set_error("rank out of range");
            return PADO_GPU_ERR_ARG;
        }
        const uint32_t sz = (uint32_t)(off[v + 1] - off[v]);
        f.write((const char *)&sz, 4);
        for (uint64_t i = off[v]; i < off[v + 1]; ++i) {
            const int16_t d = (int16_t)dist[i];
            f.write((const char *)&hub[i], 4);
            f.write((const char *)&d, 2);
        }
    }
    if (!f.good()) {
        set_error(std::string("write failed: ") + filename);
        return PADO_GPU_ERR_IO;
    }
    return PADO_GPU_OK;
}

pado_gpu_status pado_gpu_windex_load(pado_gpu_ctx *ctx, const char *filename, pado_gpu_index **out)
{
    clear_error();
    PADO_TRY(check_ctx(ctx));
    if (!filename || !out) {
        set_error("bad argument");
        return PADO_GPU_ERR_ARG;
    }
    *out = nullptr;
    std::ifstream f(filename, std::ios::binary);
    if (!f.is_open()) {
        set_error(std::string("cannot open file ") + filename);
        return PADO_GPU_ERR_IO;
    }
    uint32_t n = 0;
    f.read((char *)&n, 4);
    if (!f.good() || n == 0 || n >= kMaxVertices) {
        set_error("bad index file header");
        return PADO_GPU_ERR_IO;
    }
    std::vector<uint64_t> off((uint64_t)n + 1, 0);
    std::vector<uint32_t> hub, dist;
    for (uint32_t x = 0; x < n; ++x) {
        uint32_t sz = 0;
        f.read((char *)&sz, 4);
        if (!f.good()) {
            set_error("truncated index file");
            return PADO_GPU_ERR_IO;
        }
        for (uint32_t i = 0; i < sz; ++i) {
            uint32_t h = 0;
            int16_t d = 0;
            f.read((char *)&h, 4);
            f.read((char *)&d, 2);
            if (h >= n || d < 0) {
                set_error("bad label in index file");
                return PADO_GPU_ERR_IO;
            }
            hub.push_back(h);
            dist.push_back((uint32_t)d);
        }
        off[x + 1] = hub.size();
    }
    if (!f.good()) {
        set_error("truncated index file");
        return PADO_GPU_ERR_IO;
    }
    auto *idx = new pado_gpu_index();
    idx->ctx = ctx;
    idx->n = n;
    idx->weighted = 1;
    idx->num_labels = hub.size();
    const uint64_t L = hub.size();
    hipStream_t st = ctx->stream;
    pado_gpu_status s = idx->qoff.alloc((uint64_t)n + 1);
    if (s == PADO_GPU_OK) s = idx->qlab.alloc(L);
    if (s == PADO_GPU_OK) s = idx->qwdist.alloc(L);
    if (s != PADO_GPU_OK) {
        delete idx;
        return s;
    }
    if (hipMemcpyAsync(idx->qoff.p, off.data(), 8ull * (n + 1), hipMemcpyHostToDevice, st) != hipSuccess ||
        (L && hipMemcpyAsync(idx->qlab.p, hub.data(), 4ull * L, hipMemcpyHostToDevice, st) != hipSuccess) ||
        (L && hipMemcpyAsync(idx->qwdist.p, dist.data(), 4ull * L, hipMemcpyHostToDevice, st) != hipSuccess) ||
        hipStreamSynchronize(st) != hipSuccess) {
        delete idx;
        set_error("index upload failed");
        return PADO_GPU_ERR_HIP;
    }
    idx->ordered = 1;
    *out = idx;
    return PADO_GPU_OK;
}

}  // extern "C"
