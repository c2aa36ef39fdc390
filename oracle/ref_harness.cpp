// ref_harness.cpp -- drives the reference implementation, compiled from the
// reference's own headers, to produce label dumps, query answers and timings.
//
// TEST INFRASTRUCTURE ONLY (see oracle/build_ref.sh).  Built into oracle/_ref/
// and used by tests/ (golden fixtures) and by bench.py's cpu_baseline leg.
//
// The reference API it exercises (paths relative to the reference repo):
//   Graph(vector<pair<idi,idi>>&)      includes/graph.h:120-153
//   Graph::make_rank / id_transfer     includes/graph.h:156-206
//   ParaVertexCentricPLL<1024>(G)      includes/pado_unw_unv_para.201912311607.parallel_bp_labeling.h:309-313
//   order_labels / query_distance      same file :2198-2416
//   WeightedGraph / WeightedParaVertexCentricPLL<512>  includes/pado_weighted_unv_para.h
//
// Binary formats (all little endian):
//   edges : u64 m, then m pairs of u32 (original ids)
//   wedges: u64 m, then m triples (u32 a, u32 b, u32 w)
//   rank  : u32 n, then n u32 (rank of each original id)
//   labels: u32 n, u64 total, u64 off[n+1], u32 hub[total], u8 dist[total]
//           (canonical: rank-space vertex order, hubs ascending)
//   bp    : u32 n, u8 bp_dist[n*50], u64 bp_sets[n*50*2]
//   pairs : u64 k, u32 a[k], u32 b[k] (original ids); answers: u8[k] (u16 weighted)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstdint>
#include <string>
#include <vector>
#include <map>
#include <unordered_map>
#include <algorithm>
#include <fstream>
#include <sstream>
#include <iostream>
#include <bitset>
#include <atomic>
#include <chrono>
#include <cmath>
#include <climits>
#include <clocale>
#include <omp.h>
#include <xmmintrin.h>

// The reference headers use pair/sort/ofstream unqualified inside namespace
// PADO (written for icpc); make those names visible there.
namespace PADO { using std::pair; using std::sort; using std::make_pair; using std::ofstream; }

// The labels L and Index are private members; expose them to this harness.
#define private public
#include "globals.h"
#include "graph.h"
#include "pado_unw_unv_para.201912311607.parallel_bp_labeling.h"
#ifdef PADO_REF_WEIGHTED
#include "pado_weighted_unv_para.h"
#endif
#undef private

using namespace PADO;

static double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static std::vector<std::pair<idi, idi>> read_edges(const char *path)
{
    FILE *f = fopen(path, "rb");
    if (!f) { fprintf(stderr, "cannot open %s\n", path); exit(2); }
    uint64_t m = 0;
    if (fread(&m, 8, 1, f) != 1) exit(2);
    std::vector<uint32_t> raw(2 * m);
    if (m && fread(raw.data(), 4, 2 * m, f) != 2 * m) exit(2);
    fclose(f);
    std::vector<std::pair<idi, idi>> e(m);
    for (uint64_t i = 0; i < m; ++i) e[i] = std::make_pair(raw[2 * i], raw[2 * i + 1]);
    return e;
}

static std::vector<idi> read_rank(const char *path)
{
    FILE *f = fopen(path, "rb");
    if (!f) { fprintf(stderr, "cannot open %s\n", path); exit(2); }
    uint32_t n = 0;
    if (fread(&n, 4, 1, f) != 1) exit(2);
    std::vector<idi> r(n);
    if (n && fread(r.data(), 4, n, f) != n) exit(2);
    fclose(f);
    return r;
}

static void write_rank(const char *path, const std::vector<idi> &r)
{
    FILE *f = fopen(path, "wb");
    uint32_t n = r.size();
    fwrite(&n, 4, 1, f);
    fwrite(r.data(), 4, n, f);
    fclose(f);
}

template <class PLL>
static void dump_labels(PLL &p, idi n, const char *path)
{
    std::vector<uint64_t> off(n + 1, 0);
    std::vector<uint32_t> hub;
    std::vector<uint8_t> dist;
    for (idi v = 0; v < n; ++v) {
        auto &Lv = p.L[v];
        std::vector<std::pair<idi, uint8_t>> ol;
        for (size_t bi = 0; bi < Lv.batches.size(); ++bi) {
            idi id_offset = Lv.batches[bi].batch_id * 1024;
            idi ds = Lv.batches[bi].start_index;
            idi de = ds + Lv.batches[bi].size;
            for (idi di = ds; di < de; ++di) {
                idi vs = Lv.distances[di].start_index;
                idi ve = vs + Lv.distances[di].size;
                for (idi vi = vs; vi < ve; ++vi)
                    ol.push_back(std::make_pair(Lv.vertices[vi] + id_offset, Lv.distances[di].dist));
            }
        }
        std::sort(ol.begin(), ol.end());
        for (auto &x : ol) { hub.push_back(x.first); dist.push_back(x.second); }
        off[v + 1] = hub.size();
    }
    FILE *f = fopen(path, "wb");
    uint64_t tot = hub.size();
    fwrite(&n, 4, 1, f);
    fwrite(&tot, 8, 1, f);
    fwrite(off.data(), 8, n + 1, f);
    fwrite(hub.data(), 4, tot, f);
    fwrite(dist.data(), 1, tot, f);
    fclose(f);
}

template <class PLL>
static void dump_bp(PLL &p, idi n, const char *path)
{
    FILE *f = fopen(path, "wb");
    fwrite(&n, 4, 1, f);
    for (idi v = 0; v < n; ++v) fwrite(p.L[v].bp_dist, 1, 50, f);
    for (idi v = 0; v < n; ++v) fwrite(p.L[v].bp_sets, 8, 100, f);
    fclose(f);
}

static void usage()
{
    fprintf(stderr,
            "usage: pado_ref unw <edges.bin> <rank.bin|-> <out_prefix> [pairs.bin] [repeat]\n"
            "       pado_ref rank <edges.bin> <rank_out.bin>\n"
            "       pado_ref csr <edges.bin> <rank.bin> <csr_out.bin>\n"
            "       pado_ref text <edges.txt> <csr_out.bin>\n"
#ifdef PADO_REF_WEIGHTED
            "       pado_ref wtext <wedges.txt> <csr_out.bin>\n"
            "       pado_ref w <wedges.bin> <rank.bin|-> <out_prefix> [pairs.bin]\n"
#endif
            );
    exit(2);
}

int main(int argc, char **argv)
{
    if (argc < 4) usage();
    std::string mode = argv[1];
    if (mode == "rank") {
        // make_rank exactly as the reference calls it (rand() with the glibc
        // default seed; this is the first use of rand() in the process).
        auto e = read_edges(argv[2]);
        Graph G(e);
        write_rank(argv[3], G.make_rank());
        return 0;
    }
    if (mode == "text" || mode == "wtext") {
        // The reference's own text readers, Graph(const char*) (graph.h:95-118)
        // and WeightedGraph(const char*) (graph.h:277-304), on a text file:
        // u32 num_v, u64 num_e, u64 off[num_v + 1], u32 adj[num_e] (+ u8 w[num_e]).
#ifdef PADO_REF_WEIGHTED
        if (mode == "wtext") {
            WeightedGraph G(argv[2]);
            FILE *f = fopen(argv[3], "wb");
            uint32_t n = G.num_v;
            uint64_t m = G.num_e;
            fwrite(&n, 4, 1, f);
            fwrite(&m, 8, 1, f);
            for (idi v = 0; v < n; ++v) { uint64_t o = G.vertices[v]; fwrite(&o, 8, 1, f); }
            fwrite(&m, 8, 1, f);
            fwrite(G.out_edges, 4, m, f);
            fwrite(G.out_weights, 1, m, f);
            fclose(f);
            return 0;
        }
#endif
        Graph G(argv[2]);
        FILE *f = fopen(argv[3], "wb");
        uint32_t n = G.num_v;
        uint64_t m = G.num_e;
        fwrite(&n, 4, 1, f);
        fwrite(&m, 8, 1, f);
        for (idi v = 0; v < n; ++v) { uint64_t o = G.vertices[v]; fwrite(&o, 8, 1, f); }
        fwrite(&m, 8, 1, f);
        fwrite(G.out_edges, 4, m, f);
        fclose(f);
        return 0;
    }
    if (mode == "csr") {
        auto e = read_edges(argv[2]);
        Graph G(e);
        std::vector<idi> rank = read_rank(argv[3]);
        G.id_transfer(rank);
        FILE *f = fopen(argv[4], "wb");
        uint32_t n = G.num_v;
        uint64_t m = G.num_e;
        fwrite(&n, 4, 1, f);
        fwrite(&m, 8, 1, f);
        for (idi v = 0; v < n; ++v) { uint64_t o = G.vertices[v]; fwrite(&o, 8, 1, f); }
        fwrite(&m, 8, 1, f);
        fwrite(G.out_edges, 4, m, f);
        fclose(f);
        return 0;
    }
    if (mode == "unw") {
        if (argc < 5) usage();
        auto e = read_edges(argv[2]);
        Graph G(e);
        std::vector<idi> rank = (std::string(argv[3]) == "-") ? G.make_rank() : read_rank(argv[3]);
        std::vector<idi> rank2id = G.id_transfer(rank);
        idi n = G.num_v;
        int repeat = argc > 6 ? atoi(argv[6]) : 1;
        std::string pre = argv[4];
        double best = 1e30;
        ParaVertexCentricPLL<1024> *p = nullptr;
        for (int it = 0; it < repeat; ++it) {
            delete p;
            double t0 = now_s();
            p = new ParaVertexCentricPLL<1024>(G);
            double t1 = now_s();
            best = std::min(best, t1 - t0);
            printf("ref_build_wall_s: %.6f\n", t1 - t0);
        }
        printf("ref_build_best_s: %.6f threads: %d\n", best, omp_get_max_threads());
        if (pre != "-") {
            dump_labels(*p, n, (pre + ".labels.bin").c_str());
            dump_bp(*p, n, (pre + ".bp.bin").c_str());
        }
        if (argc > 5 && std::string(argv[5]) != "-") {
            p->order_labels(rank2id, rank);
            FILE *f = fopen(argv[5], "rb");
            uint64_t k = 0;
            if (fread(&k, 8, 1, f) != 1) exit(2);
            std::vector<uint32_t> a(k), b(k);
            if (k && fread(a.data(), 4, k, f) != k) exit(2);
            if (k && fread(b.data(), 4, k, f) != k) exit(2);
            fclose(f);
            std::vector<uint8_t> out(k);
            double q0 = now_s();
            for (uint64_t i = 0; i < k; ++i) out[i] = p->query_distance(a[i], b[i]);
            double q1 = now_s();
#pragma omp parallel for schedule(static)
            for (uint64_t i = 0; i < k; ++i) out[i] = p->query_distance(a[i], b[i]);
            double q2 = now_s();
            printf("ref_query_serial_s: %.6f ref_query_omp_s: %.6f pairs: %lu\n", q1 - q0, q2 - q1,
                   (unsigned long)k);
            if (pre != "-") {
                FILE *g = fopen((pre + ".answers.bin").c_str(), "wb");
                fwrite(out.data(), 1, k, g);
                fclose(g);
            }
        }
        delete p;
        return 0;
    }
#ifdef PADO_REF_WEIGHTED
    if (mode == "w") {
        if (argc < 5) usage();
        FILE *f = fopen(argv[2], "rb");
        uint64_t m = 0;
        if (fread(&m, 8, 1, f) != 1) exit(2);
        std::vector<uint32_t> raw(3 * m);
        if (m && fread(raw.data(), 4, 3 * m, f) != 3 * m) exit(2);
        fclose(f);
        std::vector<std::pair<idi, idi>> e(m);
        std::vector<weighti> w(m);
        for (uint64_t i = 0; i < m; ++i) {
            e[i] = std::make_pair(raw[3 * i], raw[3 * i + 1]);
            w[i] = static_cast<weighti>(raw[3 * i + 2]);
        }
        WeightedGraph G(e, w);
        std::vector<idi> rank = (std::string(argv[3]) == "-") ? G.make_rank() : read_rank(argv[3]);
        std::vector<idi> rank2id = G.id_transfer(rank);
        double t0 = now_s();
        auto *p = new WeightedParaVertexCentricPLL<512>(G);
        double t1 = now_s();
        printf("ref_build_wall_s: %.6f threads: %d\n", t1 - t0, omp_get_max_threads());
        if (std::string(argv[4]) != "-") {
            // Canonical weighted labels: rank-space vertex order, (hub, dist) sorted.
            idi n = G.num_v;
            std::vector<uint64_t> off(n + 1, 0);
            std::vector<uint32_t> hub;
            std::vector<int16_t> dist;
            for (idi v = 0; v < n; ++v) {
                std::vector<std::pair<idi, int16_t>> ol;
                for (size_t i = 0; i < p->L[v].vertices.size(); ++i)
                    ol.push_back(std::make_pair(p->L[v].vertices[i], p->L[v].distances[i]));
                std::sort(ol.begin(), ol.end());
                for (auto &x : ol) {
                    hub.push_back(x.first);
                    dist.push_back(x.second);
                }
                off[v + 1] = hub.size();
            }
            FILE *g = fopen((std::string(argv[4]) + ".wlabels.bin").c_str(), "wb");
            uint64_t tot = hub.size();
            fwrite(&n, 4, 1, g);
            fwrite(&tot, 8, 1, g);
            fwrite(off.data(), 8, n + 1, g);
            fwrite(hub.data(), 4, tot, g);
            fwrite(dist.data(), 2, tot, g);
            fclose(g);
        }
        if (argc > 5 && std::string(argv[5]) != "-") {
            p->order_labels(rank2id, rank);
            FILE *g = fopen(argv[5], "rb");
            uint64_t k = 0;
            if (fread(&k, 8, 1, g) != 1) exit(2);
            std::vector<uint32_t> a(k), b(k);
            if (k && fread(a.data(), 4, k, g) != k) exit(2);
            if (k && fread(b.data(), 4, k, g) != k) exit(2);
            fclose(g);
            std::vector<int16_t> out(k);
            for (uint64_t i = 0; i < k; ++i) out[i] = p->query_distance(a[i], b[i]);
            FILE *h = fopen((std::string(argv[4]) + ".answers.bin").c_str(), "wb");
            fwrite(out.data(), 2, k, h);
            fclose(h);
        }
        delete p;
        return 0;
    }
#endif
    usage();
    return 2;
}
