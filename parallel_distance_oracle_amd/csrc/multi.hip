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
#include <chrono>
#include <cstring>
#include <numeric>

#include "internal.hpp"

using namespace pado;

namespace {

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

    // Partition: graphs inside the u8 domain go into unions of < 2^24
    // vertices and < 2^32 directed edges; the others are built alone.
    std::vector<std::vector<uint32_t>> groups(1);
    uint64_t gv = 0, ge = 0;
    for (uint32_t g = 0; g < num_graphs; ++g) {
        const GraphSlot &G = mi->graphs[g];
        const uint64_t *o = off + G.voff;
        const uint64_t m = o[G.n];
        if (distance_bound(G.n, o, adj + G.eoff) == UINT32_MAX) continue;
        if (gv + G.n >= kMaxVertices || ge + m >= (1ull << 32)) {
            groups.emplace_back();
            gv = ge = 0;
        }
        groups.back().push_back(g);
        gv += G.n;
        ge += m;
    }

    pado_gpu_status s = PADO_GPU_OK;
    for (auto &grp : groups) {
        if (grp.empty()) continue;
        Union U;
        // Interleave: order (local rank, graph).
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
        // Union CSR: relabelled adjacency keeps each list descending because
        // the relabelling preserves the order inside a graph.
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
            for (uint64_t e = o[v]; e < o[v + 1]; ++e) {
                if (a[e] >= G.n) {
                    set_error("adjacency id out of range in graph " + std::to_string(U.owner[x]));
                    delete mi;
                    return PADO_GPU_ERR_ARG;
                }
                uadj[w++] = G.gid[a[e]];
            }
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
        s = pado_gpu_graph_upload(ctx, (uint32_t)total, medges, uoff.data(), uadj.data(), nullptr, &ug);
        if (s != PADO_GPU_OK) break;
        auto *idx = new pado_gpu_index();
        idx->ctx = ctx;
        idx->n = (uint32_t)total;
        for (auto &r : idx->bp_roots) r = kEmpty;
        DevBuf<uint8_t> d_used;
        const auto tb = std::chrono::steady_clock::now();
        s = build_bp_seeded(ug, idx, d_used, used, seeds);
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
            break;
        }
        U.idx = idx;
        U.loff_host.resize(total + 1);
        if (hipMemcpy(U.loff_host.data(), idx->loff.p, 8 * (total + 1), hipMemcpyDeviceToHost) != hipSuccess) {
            delete idx;
            set_error("label offset copy failed");
            s = PADO_GPU_ERR_HIP;
            break;
        }
        const int uid = (int)mi->unions.size();
        for (uint32_t g : grp) mi->graphs[g].union_id = uid;
        mi->union_vertices += total;
        mi->unions.push_back(std::move(U));
    }
    if (s != PADO_GPU_OK) {
        delete mi;
        return s;
    }
    // Graphs outside the proven range: on their own.  Vertex 0 is the first
    // bit-parallel root and its BFS runs to its eccentricity, so an
    // eccentricity >= 255 means the reference's u8 domain is left: such a
    // graph is built in the wide hop domain (12-bit distances) at once.  The
    // others try the u8 domain first and fall back to the wide one if a batch
    // iteration reaches 255.  Only a graph the wide domain cannot hold (more
    // than 2^20 vertices or distances past 4094) reports ERR_DOMAIN.
    pado_gpu_build_cfg base{};
    if (cfg) base = *cfg;
    pado_gpu_build_cfg wide_cfg = base;
    wide_cfg.flags |= PADO_GPU_BUILD_WIDE;
    for (uint32_t g = 0; g < num_graphs; ++g) {
        GraphSlot &G = mi->graphs[g];
        if (G.union_id >= 0) continue;
        const bool deep = eccentricity(G.n, off + G.voff, adj + G.eoff, 0) >= kInf8;
        pado_gpu_graph *gg = nullptr;
        s = pado_gpu_graph_upload(ctx, G.n, off[G.voff + G.n], off + G.voff, adj + G.eoff, nullptr, &gg);
        if (s != PADO_GPU_OK) {
            delete mi;
            return s;
        }
        G.status = pado_gpu_build(ctx, gg, deep ? &wide_cfg : &base, &G.alone);
        if (G.status == PADO_GPU_ERR_DOMAIN && !deep) G.status = pado_gpu_build(ctx, gg, &wide_cfg, &G.alone);
        pado_gpu_graph_free(gg);
        if (G.status != PADO_GPU_OK && G.status != PADO_GPU_ERR_DOMAIN) {
            s = G.status;
            delete mi;
            return s;
        }
    }
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
