// line_fields.cpp -- TEST INFRASTRUCTURE ONLY.
//
// The line loop of the reference's text readers, Graph(const char*) and
// WeightedGraph(const char*) (includes/graph.h:95-118, 277-304), restated
// with the C++ standard library doing the parsing: getline, skip lines whose
// first character is '#' or '%', then `std::istringstream(line) >> a >> b
// [>> c]` into uint32_t variables declared outside the loop.  The reference
// leaves those variables uninitialised; here they start at 0, so only files
// whose first edge line sets every field are comparable with it.
//
//   line_fields <ncols 2|3> <in.txt> <out.bin>
// writes u64 m, then m x ncols u32 values.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

int main(int argc, char **argv)
{
    if (argc != 4) {
        fprintf(stderr, "usage: line_fields <2|3> <in.txt> <out.bin>\n");
        return 2;
    }
    const int ncols = atoi(argv[1]);
    std::ifstream in(argv[2]);
    if (!in.is_open() || (ncols != 2 && ncols != 3)) return 1;
    std::string line;
    uint32_t a = 0, b = 0, c = 0;
    std::vector<uint32_t> rows;
    while (std::getline(in, line)) {
        if (line[0] == '#' || line[0] == '%') continue;
        std::istringstream lin(line);
        if (ncols == 2)
            lin >> a >> b;
        else
            lin >> a >> b >> c;
        rows.push_back(a);
        rows.push_back(b);
        if (ncols == 3) rows.push_back(c);
    }
    FILE *f = fopen(argv[3], "wb");
    if (!f) return 1;
    const uint64_t m = rows.size() / ncols;
    fwrite(&m, 8, 1, f);
    if (!rows.empty()) fwrite(rows.data(), 4, rows.size(), f);
    return fclose(f) == 0 ? 0 : 1;
}
