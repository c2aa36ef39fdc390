// Mirrors the reference's call pattern (pado.cpp:51-126 and :185-237, and
// query_distance.cpp:55-61) against the GPU drop-in headers.
//   dropin_example <edges.txt> <index_out> <pairs.txt>
// prints one distance per pair (exact), then the reference-compatible one.
#include <stdio.h>
#include <stdlib.h>

#include <fstream>
#include <vector>

#include "pado/graph.h"
#include "pado/pado_unw_unv_para.h"

using namespace PADO;

int main(int argc, char *argv[])
{
    if (argc < 4) {
        fprintf(stderr, "usage: %s <edges.txt> <index_out> <pairs.txt>\n", argv[0]);
        return EXIT_FAILURE;
    }
    Graph G(argv[1]);
    std::vector<idi> rank = G.make_rank();
    std::vector<idi> rank2id = G.id_transfer(rank);
    ParaVertexCentricPLL<1024> *VCPLL = new ParaVertexCentricPLL<1024>(G);
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
        printf("Q %u %u %u %u\n", qa[i], qb[i], (unsigned)VCPLL->query_distance(qa[i], qb[i]),
               (unsigned)VCPLL->query_distance_reference(qa[i], qb[i]));
    delete VCPLL;
    // query_distance.cpp's pattern: load the index file and query it.
    VertexCentricPLL<1024> *L = new VertexCentricPLL<1024>();
    L->load_index_from_file(argv[2]);
    for (size_t i = 0; i < qa.size(); ++i) printf("L %u\n", (unsigned)L->query_distance(qa[i], qb[i]));
    delete L;
    return EXIT_SUCCESS;
}
