// Mirrors the reference's weighted call pattern (pado.cpp:270-291) against the
// GPU drop-in headers, then queries the index in memory and after a reload.
//   wdropin_example <weighted_edges.txt> <index_out> <pairs.txt>
// prints "Q a b d" (reference-compatible int16 answer) and "X a b d" (exact
// u32) per pair, then "L d" per pair from the reloaded index file.
#include <stdio.h>
#include <stdlib.h>

#include <fstream>
#include <vector>

#include "pado/graph.h"
#include "pado/pado_weighted_unv_para.h"

using namespace PADO;

int main(int argc, char *argv[])
{
    if (argc < 4) {
        fprintf(stderr, "usage: %s <weighted_edges.txt> <index_out> <pairs.txt>\n", argv[0]);
        return EXIT_FAILURE;
    }
    WeightedGraph G(argv[1]);
    std::vector<idi> rank = G.make_rank();
    std::vector<idi> rank2id = G.id_transfer(rank);
    WeightedParaVertexCentricPLL<512> *VCPLL = new WeightedParaVertexCentricPLL<512>(G);
    VCPLL->store_index_to_file(argv[2], rank);
    VCPLL->order_labels(rank2id, rank);
    std::ifstream pin(argv[3]);
    idi a, b;
    std::vector<idi> qa, qb;
    while (pin >> a >> b) {
        qa.push_back(a);
        qb.push_back(b);
    }
    for (size_t i = 0; i < qa.size(); ++i)
        printf("Q %u %u %d\n", qa[i], qb[i], (int)VCPLL->query_distance(qa[i], qb[i]));
    std::vector<uint32_t> exact;
    VCPLL->query_distance(qa, qb, exact);
    for (size_t i = 0; i < qa.size(); ++i) printf("X %u %u %u\n", qa[i], qb[i], exact[i]);
    delete VCPLL;
    WeightedVertexCentricPLL<512> *L = new WeightedVertexCentricPLL<512>();
    L->load_index_from_file(argv[2]);
    for (size_t i = 0; i < qa.size(); ++i) printf("L %d\n", (int)L->query_distance(qa[i], qb[i]));
    delete L;
    return EXIT_SUCCESS;
}
