// ref_unp.cpp -- writes an index file with the reference's own writer.
//
// TEST INFRASTRUCTURE ONLY (built by oracle/build_ref.sh into oracle/_ref/).
// The compiled OpenMP class writes no index (its store_index_to_file has
// every write commented out), so the file format is produced here by the
// sequential VertexCentricPLL<1024>, whose writer works and whose label sets
// equal the OpenMP build's (paths relative to the reference repo):
//   Graph(vector<pair<idi,idi>>&), make_rank, id_transfer   includes/graph.h:120-206
//   VertexCentricPLL<1024>(G)                                includes/pado_unw_unv_unp.h:42-257
//   store_index_to_file(filename, rank)                      includes/pado_unw_unv_unp.h:1120-1180
//
//   ref_unp <edges.bin> <rank.bin> <out.idx>
// edges: u64 m, then m pairs of u32 (original ids); rank: u32 n, then n u32.
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <string>
#include <vector>
#include <map>
#include <unordered_map>
#include <algorithm>
#include <fstream>
#include <iostream>
#include <bitset>
#include <cmath>
#include <climits>
#include <xmmintrin.h>

// The reference headers use pair/sort/ofstream unqualified inside namespace
// PADO (written for icpc); make those names visible there.
namespace PADO { using std::pair; using std::sort; using std::make_pair; using std::ofstream; }

#include "globals.h"
#include "graph.h"
#include "pado_unw_unv_unp.h"

using namespace PADO;

int main(int argc, char **argv)
{
    if (argc != 4) {
        fprintf(stderr, "usage: ref_unp <edges.bin> <rank.bin> <out.idx>\n");
        return 2;
    }
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    uint64_t m = 0;
    if (fread(&m, 8, 1, f) != 1) return 2;
    std::vector<uint32_t> raw(2 * m);
    if (m && fread(raw.data(), 4, 2 * m, f) != 2 * m) return 2;
    fclose(f);
    std::vector<std::pair<idi, idi>> e(m);
    for (uint64_t i = 0; i < m; ++i) e[i] = std::make_pair(raw[2 * i], raw[2 * i + 1]);
    f = fopen(argv[2], "rb");
    if (!f) return 2;
    uint32_t n = 0;
    if (fread(&n, 4, 1, f) != 1) return 2;
    std::vector<idi> rank(n);
    if (n && fread(rank.data(), 4, n, f) != n) return 2;
    fclose(f);
    Graph G(e);
    G.id_transfer(rank);
    VertexCentricPLL<1024> pll(G);
    pll.store_index_to_file(argv[3], rank);
    return 0;
}
