// Reads a text edge list through the drop-in's Graph(const char*) or
// WeightedGraph(const char*) and writes the CSR it built:
// u32 num_v, u64 num_e, u64 off[num_v + 1], u32 adj[num_e] (+ u8 w[num_e]).
//   graph_text <text|wtext> <in.txt> <out.bin>
#include <cstdio>
#include <cstring>

#include "pado/graph.h"

template <class G>
static FILE *dump_csr(const G &g, const char *path)
{
    FILE *f = fopen(path, "wb");
    uint32_t n = g.num_v;
    uint64_t m = g.num_e;
    fwrite(&n, 4, 1, f);
    fwrite(&m, 8, 1, f);
    for (uint32_t v = 0; v < n; ++v) {
        uint64_t o = g.vertices[v];
        fwrite(&o, 8, 1, f);
    }
    fwrite(&m, 8, 1, f);
    fwrite(g.out_edges, 4, m, f);
    return f;
}

int main(int argc, char **argv)
{
    if (argc != 4) return 2;
    if (strcmp(argv[1], "wtext") == 0) {
        PADO::WeightedGraph g(argv[2]);
        FILE *f = dump_csr(g, argv[3]);
        fwrite(g.out_weights, 1, g.num_e, f);
        fclose(f);
    } else {
        PADO::Graph g(argv[2]);
        fclose(dump_csr(g, argv[3]));
    }
    return 0;
}
