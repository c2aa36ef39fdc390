// multi.hip -- labeling many independent graphs as one job.
//
// The reference labels one graph per process (pado.cpp:137-351).  Two facts
// make a single interleaved job exact for many graphs at once:
//   * pruning never crosses components (pushes follow edges), and
//   * the label sets do not depend on BATCH_SIZE (a batch only delays prunes
//     that cannot fire before the labels they need exist),
// so a union graph whose ids interleave the graphs while keeping each graph's
// own rank order yields, restricted to each graph, exactly that graph's
// labels.  Bit-parallel slots are per graph: slot i of every graph is seeded
// in the same multi-source sweep (each BFS stays in its component).  The
// union orders vertices by (local rank, graph), so batch 0 holds the top
// root of every graph and every batch carries roots of many graphs.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <mutex>
#include <numeric>
#include <thread>

#include "internal.hpp"

using namespace pado;

namespace {

// Concurrent builds of one job: one stream each; the device exposes four
// hardware queues per process by default, so more streams would share them.
constexpr uint32_t kMultiWorkers = 4;

struct GraphSlot {
    uint32_t n = 0;
    uint64_t eoff = 0;           // offset of this graph's adjacency in the input
    uint64_t voff = 0;           // offset of this graph's offset array in the input
    int union_id = -1;           // index into unions, or -1 (built alone)
    std::vector<uint32_t> gid;   // local rank -> union id
    pado_gpu_index *alone = nullptr;
    pado_gpu_status status = PADO_GPU_OK;
};

struct Union {
    pado_gpu_index *idx = nullptr;
    std::vector<uint32_t> owner;   // union id -> graph
    std::vector<uint32_t> local;   // union id -> local rank
    std::vector<uint64_t> loff_host;  // label offsets of the union (per-graph counts)
    // Host copy of the rank-space union export, filled on first export.
    bool exported = false;
    std::vector<uint8_t> bpd;
    std::vector<uint64_t> bps, loff;
    std::vector<uint32_t> hub;
    std::vector<uint8_t> dist;
};

// Upper bound on the graph's largest distance: per component, 2 x the BFS
// eccentricity of some vertex bounds every distance.  The vertex is the
// middle of a double-sweep path (first vertex -> farthest a -> farthest b,
// then the vertex halfway from b back to a), whose eccentricity is close to
// the radius on path- and grid-like components, so the bound is near the
// diameter instead of up to twice it.  Returns UINT32_MAX if any bound >= 255.
uint32_t distance_bound(uint32_t n, const uint64_t *off, const uint32_t *adj)
{
    std::vector<uint32_t> comp(n, UINT32_MAX), dist(n, UINT32_MAX), par(n), q(n);
    // BFS from s over its component; returns (eccentricity, farthest vertex).
    auto bfs = [&](uint32_t s) {
        uint32_t h = 0, t = 0, far = s;
        dist[s] = 0;
        par[s] = s;
        q[t++] = s;
        while (h < t) {
            const uint32_t v = q[h++];
            if (dist[v] > dist[far]) far = v;
            for (uint64_t e = off[v]; e < off[v + 1]; ++e) {
                const uint32_t u = adj[e];
                if (dist[u] == UINT32_MAX) {
                    dist[u] = dist[v] + 1;
                    par[u] = v;
                    q[t++] = u;
                }
            }
        }
        const uint32_t ecc = dist[far];
        for (uint32_t i = 0; i < t; ++i) dist[q[i]] = UINT32_MAX;  // ready for the next sweep
        return std::make_pair(ecc, far);
    };
    uint32_t bound = 0;
    for (uint32_t s = 0; s < n; ++s) {
        if (comp[s] != UINT32_MAX) continue;
        // Mark the component (BFS from s), then the double sweep.
        uint32_t h = 0, t = 0;
        comp[s] = s;
        q[t++] = s;
        while (h < t) {
            const uint32_t v = q[h++];
            for (uint64_t e = off[v]; e < off[v + 1]; ++e)
                if (comp[adj[e]] == UINT32_MAX) {
                    comp[adj[e]] = s;
                    q[t++] = adj[e];
                }
        }
        const auto e0 = bfs(s);
        const auto ea = bfs(e0.second);
        // Walk from b (farthest from a) back toward a by half the a-b distance.
        uint32_t mid = ea.second;
        {
            // par[] still holds the BFS tree rooted at a.
            for (uint32_t k = 0; k < ea.first / 2; ++k) mid = par[mid];
        }
        const auto em = bfs(mid);
        const uint32_t b = 2 * std::min(std::min(e0.first, ea.first), em.first);
        bound = std::max(bound, b);
        if (bound >= 255) return UINT32_MAX;
    }
    return bound;
}

// BFS eccentricity of vertex s (within its component).
uint32_t eccentricity(uint32_t n, const uint64_t *off, const uint32_t *adj, uint32_t s)
{
    std::vector<uint32_t> dist(n, UINT32_MAX), q(n);
    uint32_t h = 0, t = 0, ecc = 0;
    dist[s] = 0;
    q[t++] = s;
    while (h < t) {
        const uint32_t v = q[h++];
        ecc = std::max(ecc, dist[v]);
        for (uint64_t e = off[v]; e < off[v + 1]; ++e)
            if (dist[adj[e]] == UINT32_MAX) {
                dist[adj[e]] = dist[v] + 1;
                q[t++] = adj[e];
            }
    }
    return ecc;
}

double ms_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct pado_gpu_multi {
    pado_gpu_ctx *ctx = nullptr;
    std::vector<GraphSlot> graphs;
    std::vector<Union> unions;
    double ms_total = 0;
    uint64_t union_vertices = 0;
    ~pado_gpu_multi()
    {
        for (auto &u : unions) pado_gpu_index_free(u.idx);
        for (auto &g : graphs) pado_gpu_index_free(g.alone);
    }
};

namespace {

// One union of graphs inside the u8 domain, labeled as one interleaved graph
// on the worker's context: ids ordered by (local rank, graph), each graph with
// its own 50 bit-parallel slots.
pado_gpu_status build_union(pado_gpu_ctx *wctx, pado_gpu_multi *mi, const std::vector<uint32_t> &grp,
                            const uint64_t *off, const uint32_t *adj, const pado_gpu_build_cfg *cfg, Union &U)
{
    uint32_t maxn = 0;
    uint64_t total = 0, medges = 0;
    for (uint32_t g : grp) {
        maxn = std::max(maxn, mi->graphs[g].n);
        total += mi->graphs[g].n;
        medges += off[mi->graphs[g].voff + mi->graphs[g].n];
    }
    U.owner.resize(total);
    U.local.resize(total);
    uint32_t next = 0;
    for (uint32_t g : grp) mi->graphs[g].gid.assign(mi->graphs[g].n, 0);
    for (uint32_t r = 0; r < maxn; ++r)
        for (uint32_t g : grp)
            if (r < mi->graphs[g].n) {
                mi->graphs[g].gid[r] = next;
                U.owner[next] = g;
                U.local[next] = r;
                ++next;
            }
    // Union CSR: relabelled adjacency keeps each list descending because the
    // relabelling preserves the order inside a graph.
    std::vector<uint64_t> uoff(total + 1, 0);
    std::vector<uint32_t> uadj(medges);
    for (uint32_t x = 0; x < total; ++x) {
        const GraphSlot &G = mi->graphs[U.owner[x]];
        const uint64_t *o = off + G.voff;
        const uint32_t v = U.local[x];
        uoff[x + 1] = uoff[x] + (o[v + 1] - o[v]);
    }
    for (uint32_t x = 0; x < total; ++x) {
        const GraphSlot &G = mi->graphs[U.owner[x]];
        const uint64_t *o = off + G.voff;
        const uint32_t *a = adj + G.eoff;
        const uint32_t v = U.local[x];
        uint64_t w = uoff[x];
        for (uint64_t e = o[v]; e < o[v + 1]; ++e) uadj[w++] = G.gid[a[e]];
    }
    // Per-graph bit-parallel seeds, renamed into union ids.
    BpSeeds seeds;
    std::vector<uint8_t> used(total, 0), lused;
    for (uint32_t g : grp) {
        const GraphSlot &G = mi->graphs[g];
        select_bp_roots(G.n, off + G.voff, adj + G.eoff, lused, nullptr, seeds, G.gid.data());
        for (uint32_t v = 0; v < G.n; ++v)
            if (lused[v]) used[G.gid[v]] = 1;
    }
    pado_gpu_graph *ug = nullptr;
    PADO_TRY(pado_gpu_graph_upload(wctx, (uint32_t)total, medges, uoff.data(), uadj.data(), nullptr, &ug));
    auto *idx = new pado_gpu_index();
    idx->ctx = wctx;
    idx->n = (uint32_t)total;
    for (auto &r : idx->bp_roots) r = kEmpty;
    DevBuf<uint8_t> d_used;
    const auto tb = std::chrono::steady_clock::now();
    pado_gpu_status s = build_bp_seeded(ug, idx, d_used, used, seeds);
    idx->ms_bp = ms_since(tb);
    if (s == PADO_GPU_OK) {
        const auto tc = std::chrono::steady_clock::now();
        pado_gpu_build_cfg c{};
        if (cfg) c = *cfg;
        s = build_batches(ug, idx, d_used, c);
        idx->ms_batch = ms_since(tc);
    }
    pado_gpu_graph_free(ug);
    if (s != PADO_GPU_OK) {
        delete idx;
        return s;
    }
    U.loff_host.resize(total + 1);
    if (copy_sync(U.loff_host.data(), idx->loff.p, 8 * (total + 1), hipMemcpyDeviceToHost, wctx->stream) != hipSuccess) {
        delete idx;
        set_error("label offset copy failed");
        return PADO_GPU_ERR_HIP;
    }
    U.idx = idx;
    return PADO_GPU_OK;
}

// A graph outside the proven range, on its own.  Vertex 0 is the first
// bit-parallel root and its BFS runs to its eccentricity, so an eccentricity
// >= 255 means the reference's u8 domain is left: such a graph is built in
// the wide hop domain (12-bit distances) at once.  The others try the u8
// domain first and fall back to the wide one if a batch iteration reaches
// 255.  Only a graph the wide domain cannot hold (more than 2^20 vertices or
// distances past 4094) reports ERR_DOMAIN.
pado_gpu_status build_alone(pado_gpu_ctx *wctx, GraphSlot &G, const uint64_t *off, const uint32_t *adj,
                            const pado_gpu_build_cfg *cfg)
{
    pado_gpu_build_cfg base{};
    if (cfg) base = *cfg;
    pado_gpu_build_cfg wide_cfg = base;
    wide_cfg.flags |= PADO_GPU_BUILD_WIDE;
    const bool deep = eccentricity(G.n, off + G.voff, adj + G.eoff, 0) >= kInf8;
    pado_gpu_graph *gg = nullptr;
    PADO_TRY(pado_gpu_graph_upload(wctx, G.n, off[G.voff + G.n], off + G.voff, adj + G.eoff, nullptr, &gg));
    G.status = pado_gpu_build(wctx, gg, deep ? &wide_cfg : &base, &G.alone);
    if (G.status == PADO_GPU_ERR_DOMAIN && !deep) G.status = pado_gpu_build(wctx, gg, &wide_cfg, &G.alone);
    pado_gpu_graph_free(gg);
    if (G.status != PADO_GPU_OK && G.status != PADO_GPU_ERR_DOMAIN) return G.status;
    return PADO_GPU_OK;
}

}  // namespace

extern "C" {

pado_gpu_status pado_gpu_build_multi(pado_gpu_ctx *ctx, uint32_t num_graphs, const uint32_t *n, const uint64_t *off,
                                     const uint32_t *adj, const pado_gpu_build_cfg *cfg, pado_gpu_multi **out)
{
    clear_error();
    if (!ctx || !out || (num_graphs && (!n || !off))) {
        set_error("null argument");
        return PADO_GPU_ERR_ARG;
    }
    *out = nullptr;
    const auto t0 = std::chrono::steady_clock::now();
    auto *mi = new pado_gpu_multi();
    mi->ctx = ctx;
    mi->graphs.resize(num_graphs);
    uint64_t voff = 0, eoff = 0;
    for (uint32_t g = 0; g < num_graphs; ++g) {
        GraphSlot &G = mi->graphs[g];
        G.n = n[g];
        G.voff = voff;
        G.eoff = eoff;
        if (G.n == 0) {
            delete mi;
            set_error("graph " + std::to_string(g) + " has no vertices");
            return PADO_GPU_ERR_ARG;
        }
        if (off[voff] != 0) {
            delete mi;
            set_error("per-graph offsets must start at 0");
            return PADO_GPU_ERR_ARG;
        }
        voff += (uint64_t)G.n + 1;
        eoff += off[G.voff + G.n];
    }
    if (eoff && !adj) {
        delete mi;
        set_error("null adjacency");
        return PADO_GPU_ERR_ARG;
    }
    for (uint32_t g = 0; g < num_graphs; ++g) {
        const GraphSlot &G = mi->graphs[g];
        const uint64_t *o = off + G.voff;
        for (uint64_t e = 0; e < o[G.n]; ++e)
            if (adj[G.eoff + e] >= G.n) {
                set_error("adjacency id out of range in graph " + std::to_string(g));
                delete mi;
                return PADO_GPU_ERR_ARG;
            }
    }

    // Workers: each task runs on its own context (stream), so the unions and
    // the graphs built alone -- independent work, each a latency-bound chain
    // of small launches -- overlap on the chip.
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const uint32_t workers = std::max(1u, std::min<uint32_t>(kMultiWorkers, hw));

    // Graphs inside the u8 domain (2 x one vertex's eccentricity per component
    // below 255) go into `workers` unions balanced by vertex count (each below
    // 2^24 vertices and 2^32 directed edges); the others are built alone.
    std::vector<uint8_t> in_domain(num_graphs, 0);
    {
        std::atomic<uint32_t> next{0};
        std::vector<std::thread> th;
        for (uint32_t w = 0; w < std::min<uint32_t>(hw, 16u); ++w)
            th.emplace_back([&] {
                for (uint32_t g; (g = next.fetch_add(1)) < num_graphs;) {
                    const GraphSlot &G = mi->graphs[g];
                    in_domain[g] = distance_bound(G.n, off + G.voff, adj + G.eoff) != UINT32_MAX;
                }
            });
        for (auto &t : th) t.join();
    }
    std::vector<uint32_t> order;
    for (uint32_t g = 0; g < num_graphs; ++g)
        if (in_domain[g]) order.push_back(g);
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t x, uint32_t y) { return mi->graphs[x].n > mi->graphs[y].n; });
    const size_t k0 = std::min<size_t>(workers, order.size());
    std::vector<std::vector<uint32_t>> groups(k0);
    std::vector<uint64_t> gv(k0, 0), ge(k0, 0);
    for (uint32_t g : order) {
        const uint64_t m = off[mi->graphs[g].voff + mi->graphs[g].n];
        // The lightest union that can still take the graph; a new one if none can.
        int best = -1;
        for (size_t k = 0; k < groups.size(); ++k)
            if (gv[k] + mi->graphs[g].n < kMaxVertices && ge[k] + m < (1ull << 32) &&
                (best < 0 || gv[k] < gv[best]))
                best = (int)k;
        if (best < 0) {
            groups.emplace_back();
            gv.push_back(0);
            ge.push_back(0);
            best = (int)groups.size() - 1;
        }
        groups[best].push_back(g);
        gv[best] += mi->graphs[g].n;
        ge[best] += m;
    }
    // Each union lists its graphs in index order (the interleaving order).
    for (auto &grp : groups) std::sort(grp.begin(), grp.end());
    mi->unions.resize(groups.size());

    // Tasks: unions first (largest first), then the graphs built alone
    // (largest first).
    struct Task {
        int union_id;   // >= 0: a union, else -1
        uint32_t graph; // a graph built alone
        uint64_t size;
    };
    std::vector<Task> tasks;
    for (size_t k = 0; k < groups.size(); ++k) tasks.push_back({(int)k, 0u, gv[k]});
    std::vector<Task> alone;
    for (uint32_t g = 0; g < num_graphs; ++g)
        if (!in_domain[g]) alone.push_back({-1, g, mi->graphs[g].n});
    std::stable_sort(tasks.begin(), tasks.end(), [](const Task &a, const Task &b) { return a.size > b.size; });
    std::stable_sort(alone.begin(), alone.end(), [](const Task &a, const Task &b) { return a.size > b.size; });
    tasks.insert(tasks.end(), alone.begin(), alone.end());

    std::atomic<size_t> next_task{0};
    std::atomic<int> failed{0};
    std::mutex err_mu;
    pado_gpu_status first_status = PADO_GPU_OK;
    std::string first_error;
    auto worker = [&]() {
        pado_gpu_ctx *wctx = nullptr;
        pado_gpu_status s = pado_gpu_ctx_create(ctx->device, &wctx);
        for (size_t t; s == PADO_GPU_OK && !failed.load() && (t = next_task.fetch_add(1)) < tasks.size();) {
            const Task &T = tasks[t];
            if (T.union_id >= 0) {
                s = build_union(wctx, mi, groups[T.union_id], off, adj, cfg, mi->unions[T.union_id]);
                if (s == PADO_GPU_OK) {
                    mi->unions[T.union_id].idx->ctx = ctx;  // the job's context outlives the worker's
                    for (uint32_t g : groups[T.union_id]) mi->graphs[g].union_id = T.union_id;
                }
            } else {
                GraphSlot &G = mi->graphs[T.graph];
                s = build_alone(wctx, G, off, adj, cfg);
                if (G.alone) G.alone->ctx = ctx;
            }
        }
        if (s != PADO_GPU_OK) {
            std::lock_guard<std::mutex> g(err_mu);
            if (first_status == PADO_GPU_OK) {
                first_status = s;
                first_error = pado_gpu_last_error();
            }
            failed.store(1);
        }
        if (wctx) {
            (void)hipStreamSynchronize(wctx->stream);
            pado_gpu_ctx_destroy(wctx);
        }
    };
    std::vector<std::thread> pool;
    const uint32_t nthreads = std::max<uint32_t>(1, std::min<uint32_t>(workers, (uint32_t)tasks.size()));
    for (uint32_t w = 0; w < nthreads; ++w) pool.emplace_back(worker);
    for (auto &t : pool) t.join();
    (void)hipSetDevice(ctx->device);
    if (first_status != PADO_GPU_OK) {
        delete mi;
        set_error(first_error);
        return first_status;
    }
    for (const Union &U : mi->unions) mi->union_vertices += U.owner.size();
    mi->ms_total = ms_since(t0);
    *out = mi;
    return PADO_GPU_OK;
}

pado_gpu_status pado_gpu_multi_graph_info(const pado_gpu_multi *mi, uint32_t g, pado_gpu_status *status,
                                          uint64_t *num_labels, int *in_union, uint32_t *dist_bits)
{
    if (!mi || g >= mi->graphs.size()) {
        set_error("bad graph index");
        return PADO_GPU_ERR_ARG;
    }
    const GraphSlot &G = mi->graphs[g];
    if (status) *status = G.status;
    if (in_union) *in_union = G.union_id >= 0;
    if (dist_bits) *dist_bits = (G.alone && G.alone->wide) ? Hop<true>::kBits : Hop<false>::kBits;
    if (num_labels) {
        if (G.union_id < 0) {
            *num_labels = G.alone ? G.alone->num_labels : 0;
        } else {
            // Label count of this graph inside the union: sum over its vertices.
            const Union &U = mi->unions[G.union_id];
            uint64_t c = 0;
            for (uint32_t v = 0; v < G.n; ++v) c += U.loff_host[G.gid[v] + 1] - U.loff_host[G.gid[v]];
            *num_labels = c;
        }
    }
    return PADO_GPU_OK;
}

}  // extern "C"

namespace {

// Export of graph g with u8 (d8) or u16 (d16) label distances.
pado_gpu_status multi_export(pado_gpu_ctx *ctx, pado_gpu_multi *mi, uint32_t g, uint8_t *bp_dist, uint64_t *bp_sets,
                             uint64_t *label_off, uint32_t *hubs, uint8_t *d8, uint16_t *d16)
{
    if (!ctx || !mi || g >= mi->graphs.size()) {
        set_error("bad argument");
        return PADO_GPU_ERR_ARG;
    }
    GraphSlot &G = mi->graphs[g];
    if (G.union_id < 0) {
        if (G.status != PADO_GPU_OK || !G.alone) {
            set_error("graph " + std::to_string(g) + " has no index (outside every hop domain)");
            return PADO_GPU_ERR_DOMAIN;
        }
        if (d16) return pado_gpu_index_export_wide(ctx, G.alone, bp_dist, bp_sets, label_off, hubs, d16);
        return pado_gpu_index_export(ctx, G.alone, bp_dist, bp_sets, label_off, hubs, d8);
    }
    Union &U = mi->unions[G.union_id];
    if (!U.exported) {
        const uint32_t N = U.idx->n;
        U.bpd.resize((uint64_t)N * kBP);
        U.bps.resize((uint64_t)N * kBP * 2);
        U.loff.resize((uint64_t)N + 1);
        U.hub.resize(U.idx->num_labels);
        U.dist.resize(U.idx->num_labels);
        PADO_TRY(pado_gpu_index_export(ctx, U.idx, U.bpd.data(), U.bps.data(), U.loff.data(), U.hub.data(),
                                       U.dist.data()));
        U.exported = true;
    }
    uint64_t o = 0;
    if (label_off) label_off[0] = 0;
    for (uint32_t v = 0; v < G.n; ++v) {
        const uint32_t x = G.gid[v];
        if (bp_dist) std::memcpy(bp_dist + (uint64_t)v * kBP, &U.bpd[(uint64_t)x * kBP], kBP);
        if (bp_sets) std::memcpy(bp_sets + (uint64_t)v * kBP * 2, &U.bps[(uint64_t)x * kBP * 2], kBP * 16);
        for (uint64_t i = U.loff[x]; i < U.loff[x + 1]; ++i, ++o) {
            if (hubs) hubs[o] = U.local[U.hub[i]];
            if (d8) d8[o] = U.dist[i];
            if (d16) d16[o] = U.dist[i];
        }
        if (label_off) label_off[v + 1] = o;
    }
    return PADO_GPU_OK;
}

}  // namespace

extern "C" {

pado_gpu_status pado_gpu_multi_export(pado_gpu_ctx *ctx, pado_gpu_multi *mi, uint32_t g, uint8_t *bp_dist,
                                      uint64_t *bp_sets, uint64_t *label_off, uint32_t *hubs, uint8_t *dists)
{
    clear_error();
    return multi_export(ctx, mi, g, bp_dist, bp_sets, label_off, hubs, dists, nullptr);
}

pado_gpu_status pado_gpu_multi_export_wide(pado_gpu_ctx *ctx, pado_gpu_multi *mi, uint32_t g, uint8_t *bp_dist,
                                           uint64_t *bp_sets, uint64_t *label_off, uint32_t *hubs, uint16_t *dists)
{
    clear_error();
    return multi_export(ctx, mi, g, bp_dist, bp_sets, label_off, hubs, nullptr, dists);
}

pado_gpu_status pado_gpu_multi_query(pado_gpu_ctx *ctx, pado_gpu_multi *mi, uint32_t g, const uint32_t *a,
                                     const uint32_t *b, uint64_t k, uint16_t *out)
{
    clear_error();
    if (!ctx || !mi || g >= mi->graphs.size() || (k && (!a || !b || !out))) {
        set_error("bad argument");
        return PADO_GPU_ERR_ARG;
    }
    if (k == 0) return PADO_GPU_OK;
    GraphSlot &G = mi->graphs[g];
    if (G.union_id < 0) {
        if (G.status != PADO_GPU_OK || !G.alone) {
            set_error("graph " + std::to_string(g) + " has no index (outside every hop domain)");
            return PADO_GPU_ERR_DOMAIN;
        }
        return pado_gpu_query_wide(ctx, G.alone, a, b, k, out, PADO_GPU_QUERY_RANK_SPACE, nullptr);
    }
    // A union graph: local ids -> union ids; pairs with an id outside the
    // graph are answered here (a == b ? 0 : 255, as the query kernel does).
    std::vector<uint32_t> ua(k), ub(k);
    for (uint64_t i = 0; i < k; ++i) {
        const bool in = a[i] < G.n && b[i] < G.n;
        ua[i] = in ? G.gid[a[i]] : 0u;
        ub[i] = in ? G.gid[b[i]] : 0u;
    }
    PADO_TRY(pado_gpu_query_wide(ctx, mi->unions[G.union_id].idx, ua.data(), ub.data(), k, out,
                                 PADO_GPU_QUERY_RANK_SPACE, nullptr));
    for (uint64_t i = 0; i < k; ++i)
        if (a[i] >= G.n || b[i] >= G.n) out[i] = a[i] == b[i] ? 0 : 255;
    return PADO_GPU_OK;
}

pado_gpu_status pado_gpu_multi_info(const pado_gpu_multi *mi, uint32_t *num_unions, uint64_t *union_vertices,
                                    double *ms_total)
{
    if (!mi) return PADO_GPU_ERR_ARG;
    if (num_unions) *num_unions = (uint32_t)mi->unions.size();
    if (union_vertices) *union_vertices = mi->union_vertices;
    if (ms_total) *ms_total = mi->ms_total;
    return PADO_GPU_OK;
}

void pado_gpu_multi_free(pado_gpu_multi *mi) { delete mi; }

}  // extern "C"
