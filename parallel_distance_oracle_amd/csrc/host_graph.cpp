// host_graph.cpp -- host-side graph ingest for the PADO GPU library.
//
//  * Deterministic workload generators (R-MAT, 2D grids with deleted edges,
//    query pairs) on a pinned splitmix64 stream, so every run and every
//    machine sees the same graphs.
//  * The reference's graph semantics, restated for the drop-in:
//      edge list -> undirected CSR      includes/graph.h:120-153
//      degree order with rand() jitter  includes/graph.h:156-174
//      renumber + descending adjacency  includes/graph.h:176-206
//
// Everything here is plain host C++; device upload happens in pado_gpu.hip.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <thread>
#include <vector>

#include "pado_gpu.h"

namespace {

inline uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline double unit53(uint64_t &s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

// Static partition of [0, count) over the host's cores.
template <class F>
void parallel_for(uint64_t count, F f)
{
    unsigned t = std::thread::hardware_concurrency();
    if (t == 0) t = 1;
    if (t > 32) t = 32;
    if (count < (1u << 16)) t = 1;
    std::vector<std::thread> th;
    const uint64_t per = (count + t - 1) / t;
    for (unsigned k = 0; k < t; ++k) {
        const uint64_t lo = k * per, hi = std::min<uint64_t>(count, lo + per);
        if (lo >= hi) break;
        th.emplace_back([=, &f] {
            for (uint64_t i = lo; i < hi; ++i) f(i);
        });
    }
    for (auto &x : th) x.join();
}

}  // namespace

extern "C" {

// R-MAT edge sampler: edge_factor << scale samples; quadrant probabilities
// (a, b, c, 1-a-b-c) at every level; self-loops dropped; undirected duplicates
// removed.  Output pairs are (u < v), sorted.  Returns the edge count, or the
// required capacity if `cap` is too small (nothing written then).
uint64_t pado_rmat_edges(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                         uint64_t seed, uint32_t *out, uint64_t cap)
{
    const uint64_t samples = (uint64_t)edge_factor << scale;
    // Sample i uses draws i*scale .. i*scale+scale-1 of one splitmix64 stream;
    // the stream's state after j draws is seed + j*gamma, so samples are
    // generated independently (in parallel) with the same result.
    std::vector<uint64_t> keys(samples);
    const double ab = a + b, abc = a + b + c;
    parallel_for(samples, [&](uint64_t i) {
        uint64_t s = seed + (i * scale) * 0x9E3779B97F4A7C15ull;
        uint32_t u = 0, v = 0;
        for (uint32_t l = 0; l < scale; ++l) {
            double r = unit53(s);
            uint32_t bu = 0, bv = 0;
            if (r < a) {
            } else if (r < ab) {
                bv = 1;
            } else if (r < abc) {
                bu = 1;
            } else {
                bu = 1;
                bv = 1;
            }
            u = (u << 1) | bu;
            v = (v << 1) | bv;
        }
        if (u > v) std::swap(u, v);
        keys[i] = (u == v) ? ~0ull : (((uint64_t)u << 32) | v);  // self-loops sort last, dropped
    });
    std::sort(keys.begin(), keys.end());
    while (!keys.empty() && keys.back() == ~0ull) keys.pop_back();
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    if ((uint64_t)keys.size() > cap || out == nullptr) return keys.size();
    for (size_t i = 0; i < keys.size(); ++i) {
        out[2 * i] = (uint32_t)(keys[i] >> 32);
        out[2 * i + 1] = (uint32_t)keys[i];
    }
    return keys.size();
}

// side x side grid (vertex id = row*side + col), each of its 2*side*(side-1)
// edges kept with probability (1 - delete_frac).  Same return convention.
uint64_t pado_grid_edges(uint32_t side, double delete_frac, uint64_t seed, uint32_t *out, uint64_t cap)
{
    uint64_t s = seed;
    std::vector<uint64_t> keys;
    for (uint32_t r = 0; r < side; ++r)
        for (uint32_t c = 0; c < side; ++c) {
            uint32_t v = r * side + c;
            if (c + 1 < side && unit53(s) >= delete_frac) keys.push_back(((uint64_t)v << 32) | (v + 1));
            if (r + 1 < side && unit53(s) >= delete_frac) keys.push_back(((uint64_t)v << 32) | (v + side));
        }
    if ((uint64_t)keys.size() > cap || out == nullptr) return keys.size();
    for (size_t i = 0; i < keys.size(); ++i) {
        out[2 * i] = (uint32_t)(keys[i] >> 32);
        out[2 * i + 1] = (uint32_t)keys[i];
    }
    return keys.size();
}

// Integer weights uniform in [lo, hi] for m edges, from the seed.
void pado_edge_weights(uint64_t m, uint32_t lo, uint32_t hi, uint64_t seed, uint32_t *w)
{
    uint64_t s = seed;
    const uint64_t span = (uint64_t)hi - lo + 1;
    for (uint64_t i = 0; i < m; ++i) w[i] = lo + (uint32_t)(splitmix64(s) % span);
}

// k pairs uniform over [0, n).
void pado_uniform_pairs(uint32_t n, uint64_t k, uint64_t seed, uint32_t *a, uint32_t *b)
{
    uint64_t s = seed;
    for (uint64_t i = 0; i < k; ++i) {
        a[i] = (uint32_t)(splitmix64(s) % n);
        b[i] = (uint32_t)(splitmix64(s) % n);
    }
}

// k draws from Zipf(s_exp) over ranks 1..n, mapped through rank2id so that
// rank r (0 = highest degree) -> original vertex id.  Inverse-CDF sampling on
// a precomputed table.
void pado_zipf_vertices(uint32_t n, double s_exp, uint64_t k, uint64_t seed, const uint32_t *rank2id,
                        uint32_t *out)
{
    std::vector<double> cdf(n);
    double acc = 0.0;
    for (uint32_t r = 0; r < n; ++r) {
        acc += std::pow((double)(r + 1), -s_exp);
        cdf[r] = acc;
    }
    uint64_t s = seed;
    for (uint64_t i = 0; i < k; ++i) {
        double u = unit53(s) * acc;
        uint32_t r = (uint32_t)(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
        if (r >= n) r = n - 1;
        out[i] = rank2id ? rank2id[r] : r;
    }
}

// ---------------------------------------------------------------------------
// Reference graph semantics.
// ---------------------------------------------------------------------------

// num_v = max id + 1 over the edge list (includes/graph.h:129-131).
uint32_t pado_graph_num_v(const uint32_t *edges, uint64_t m)
{
    uint32_t n = 0;
    for (uint64_t i = 0; i < 2 * m; ++i) n = std::max(n, edges[i] + 1);
    return n;
}

// Undirected CSR: both directions per edge, in edge order, no de-duplication
// and no self-loop removal (includes/graph.h:125-153).  off has n+1 entries.
void pado_graph_csr(const uint32_t *edges, uint64_t m, uint32_t n, uint64_t *off, uint32_t *adj)
{
    std::fill(off, off + n + 1, 0);
    for (uint64_t i = 0; i < m; ++i) {
        off[edges[2 * i] + 1]++;
        off[edges[2 * i + 1] + 1]++;
    }
    for (uint32_t v = 0; v < n; ++v) off[v + 1] += off[v];
    std::vector<uint64_t> pos(off, off + n);
    for (uint64_t i = 0; i < m; ++i) {
        uint32_t x = edges[2 * i], y = edges[2 * i + 1];
        adj[pos[x]++] = y;
        adj[pos[y]++] = x;
    }
}

// Degree order (includes/graph.h:156-174): key = float(deg) + float(rand())/
// RAND_MAX, sorted descending as pair<float, id>.  `use_global_rand` != 0
// consumes the process-wide rand() exactly as the reference does; otherwise a
// private glibc TYPE_3 generator with seed 1 produces the same sequence the
// reference sees in a fresh process, without touching global state.
int pado_make_rank(const uint64_t *off, uint32_t n, int use_global_rand, uint32_t *rank)
{
    struct random_data rd;
    std::memset(&rd, 0, sizeof(rd));
    char state[128];
    std::memset(state, 0, sizeof(state));
    if (!use_global_rand && initstate_r(1u, state, sizeof(state), &rd) != 0) return -1;
    std::vector<std::pair<float, uint32_t>> key(n);
    for (uint32_t v = 0; v < n; ++v) {
        int32_t r;
        if (use_global_rand) {
            r = rand();
        } else {
            random_r(&rd, &r);
        }
        float deg = (float)(uint32_t)(off[v + 1] - off[v]);
        key[v] = std::make_pair(deg + (float)r / (float)RAND_MAX, v);
    }
    std::sort(key.rbegin(), key.rend());
    for (uint32_t i = 0; i < n; ++i) rank[key[i].second] = i;
    return 0;
}

// Renumber into rank space and sort each adjacency list descending
// (includes/graph.h:176-206).  rank2id may be null.
void pado_id_transfer(uint32_t n, const uint64_t *off, const uint32_t *adj, const uint32_t *rank,
                      uint64_t *off_new, uint32_t *adj_new, uint32_t *rank2id)
{
    std::vector<uint32_t> r2i(n);
    for (uint32_t v = 0; v < n; ++v) r2i[rank[v]] = v;
    off_new[0] = 0;
    for (uint32_t nv = 0; nv < n; ++nv) {
        uint32_t v = r2i[nv];
        off_new[nv + 1] = off_new[nv] + (off[v + 1] - off[v]);
    }
    for (uint32_t nv = 0; nv < n; ++nv) {
        uint32_t v = r2i[nv];
        uint64_t o = off_new[nv];
        for (uint64_t e = off[v]; e < off[v + 1]; ++e) adj_new[o++] = rank[adj[e]];
        std::sort(adj_new + off_new[nv], adj_new + off_new[nv + 1], std::greater<uint32_t>());
    }
    if (rank2id) std::copy(r2i.begin(), r2i.end(), rank2id);
}

// WeightedGraph::construct (includes/graph.h:313-351): edge i appended to both
// endpoints' lists in input order, weight cast to the u8 edge weight type.
void pado_wgraph_csr(const uint32_t *edges, const uint32_t *weights, uint64_t m, uint32_t n, uint64_t *off,
                     uint32_t *adj, uint8_t *w)
{
    pado_graph_csr(edges, m, n, off, adj);
    std::vector<uint64_t> pos(off, off + n);
    for (uint64_t i = 0; i < m; ++i) {
        const uint8_t wi = (uint8_t)weights[i];
        w[pos[edges[2 * i]]++] = wi;
        w[pos[edges[2 * i + 1]]++] = wi;
    }
}

// WeightedGraph::id_transfer (includes/graph.h:374-408): each list of
// (rank-space id, weight) pairs sorted descending, so parallel edges to one
// neighbour keep the heavier weight first.
void pado_wid_transfer(uint32_t n, const uint64_t *off, const uint32_t *adj, const uint8_t *w,
                       const uint32_t *rank, uint64_t *off_new, uint32_t *adj_new, uint8_t *w_new,
                       uint32_t *rank2id)
{
    std::vector<uint32_t> r2i(n);
    for (uint32_t v = 0; v < n; ++v) r2i[rank[v]] = v;
    off_new[0] = 0;
    for (uint32_t nv = 0; nv < n; ++nv) {
        uint32_t v = r2i[nv];
        off_new[nv + 1] = off_new[nv] + (off[v + 1] - off[v]);
    }
    std::vector<uint64_t> key;
    for (uint32_t nv = 0; nv < n; ++nv) {
        uint32_t v = r2i[nv];
        key.clear();
        for (uint64_t e = off[v]; e < off[v + 1]; ++e) key.push_back(((uint64_t)rank[adj[e]] << 8) | w[e]);
        std::sort(key.begin(), key.end(), std::greater<uint64_t>());
        uint64_t o = off_new[nv];
        for (uint64_t k : key) {
            adj_new[o] = (uint32_t)(k >> 8);
            w_new[o++] = (uint8_t)(k & 255u);
        }
    }
    if (rank2id) std::copy(r2i.begin(), r2i.end(), rank2id);
}

}  // extern "C"
