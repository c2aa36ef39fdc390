// edge_io.cpp -- edge-list files for the PADO GPU library's host ingest.
//
//  * Text, as Graph(const char*) / WeightedGraph(const char*) read it
//    (includes/graph.h:95-118, 277-304): lines starting with '#' or '%' are
//    skipped; every other line is one edge, `lin >> head >> tail [>> weight]`
//    on an istringstream of the line, with head / tail / weight declared
//    outside the loop.  The stream's rules therefore decide malformed lines:
//    a field that meets the end of the line before any character keeps the
//    previous line's value (the extraction never starts), a field that meets
//    a non-number stores 0 and ends the line, a number past 2^32 - 1 stores
//    2^32 - 1 and ends the line, and a field after one that ended the line
//    (failed, or ran into the end) keeps its previous value.  Values before
//    the first line are 0.  The file is memory-mapped and parsed in parallel
//    chunks cut at line breaks; a chunk's lines that still depend on values
//    from before it are patched in one ordered pass.
//  * Binary, as the reference's converter writes it
//    (distributed_version/tools/graph-convert/edgelist2binary.cpp:28-91):
//    u32 num_v, u64 num_e, then num_e (u32 head, u32 tail) pairs.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "pado_gpu.h"

namespace {

inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

// One field of `lin >> x` inside a line [p, e).  Returns the field's fate:
// 0 = the extraction did not start (end of line first: x unchanged, stream
// failed), 1 = x assigned and the stream still good, 2 = x assigned and the
// stream done (end of line reached, non-number, or overflow).
inline int extract(const char *&p, const char *e, uint32_t &x)
{
    while (p < e && is_space(*p)) ++p;
    if (p == e) return 0;
    bool neg = false;
    if (*p == '+' || *p == '-') {
        neg = *p == '-';
        ++p;
    }
    if (p == e || *p < '0' || *p > '9') {
        x = 0;
        return 2;
    }
    uint64_t v = 0;
    bool over = false;
    while (p < e && *p >= '0' && *p <= '9') {
        v = v * 10 + (uint64_t)(*p - '0');
        if (v > 0xFFFFFFFFull) over = true, v = 0xFFFFFFFFull;
        ++p;
    }
    if (over) {
        x = 0xFFFFFFFFu;
        return 2;
    }
    // An unsigned extraction of "-n" yields the value's negation modulo 2^32.
    x = neg ? (uint32_t)(0u - (uint32_t)v) : (uint32_t)v;
    return p == e ? 2 : 1;
}

struct Chunk {
    std::vector<uint32_t> rec;  // ncols values per edge
    // Fields whose value comes from before the chunk: for column c, the
    // edges [0, first_set[c]) take the incoming value.
    uint64_t first_set[3] = {~0ull, ~0ull, ~0ull};
    uint32_t last[3] = {0, 0, 0};  // values after the chunk's last line (valid where set)
};

void parse_range(const char *b, const char *e, uint32_t ncols, Chunk &ck)
{
    uint32_t cur[3] = {0, 0, 0};
    bool known[3] = {false, false, false};
    uint64_t edges = 0;
    const char *p = b;
    while (p < e) {
        const char *nl = (const char *)memchr(p, '\n', (size_t)(e - p));
        const char *le = nl ? nl : e;
        if (p < le && (*p == '#' || *p == '%')) {
            p = nl ? nl + 1 : e;
            continue;
        }
        // An empty last "line" after the file's final newline is no line.
        if (!nl && p == le) break;
        const char *q = p;
        bool good = true;
        for (uint32_t c = 0; c < ncols && good; ++c) {
            uint32_t x = cur[c];
            const int r = extract(q, le, x);
            if (r == 0) {
                good = false;  // unchanged from here on
                break;
            }
            cur[c] = x;
            if (!known[c]) {
                known[c] = true;
                ck.first_set[c] = edges;
            }
            if (r == 2) good = false;
        }
        for (uint32_t c = 0; c < ncols; ++c) ck.rec.push_back(cur[c]);
        ++edges;
        p = nl ? nl + 1 : e;
    }
    for (uint32_t c = 0; c < ncols; ++c) {
        ck.last[c] = cur[c];
        if (!known[c]) ck.first_set[c] = ~0ull;
    }
}

}  // namespace

extern "C" {

void pado_free(void *p) { free(p); }

int pado_read_edge_text(const char *path, uint32_t ncols, uint32_t **out, uint64_t *m)
{
    if (!path || !out || !m || ncols < 2 || ncols > 3) return -3;
    *out = nullptr;
    *m = 0;
    const int fd = open(path, O_RDONLY);
    if (fd < 0) return -1;
    struct stat sb;
    if (fstat(fd, &sb) != 0) {
        close(fd);
        return -1;
    }
    const size_t size = (size_t)sb.st_size;
    const char *data = nullptr;
    if (size) {
        void *mp = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
        if (mp == MAP_FAILED) {
            close(fd);
            return -1;
        }
        data = (const char *)mp;
    }
    // Chunks of about 4 MiB, cut after a newline.
    unsigned t = std::thread::hardware_concurrency();
    if (t == 0) t = 1;
    if (t > 32) t = 32;
    std::vector<size_t> cuts{0};
    const size_t step = std::max<size_t>(size / (8 * t) + 1, 1u << 22);
    while (cuts.back() < size) {
        size_t c = std::min(size, cuts.back() + step);
        if (c < size) {
            const char *nl = (const char *)memchr(data + c, '\n', size - c);
            c = nl ? (size_t)(nl - data) + 1 : size;
        }
        cuts.push_back(c);
    }
    const size_t nck = cuts.size() - 1;
    std::vector<Chunk> ck(nck);
    std::vector<std::thread> th;
    std::atomic<size_t> next{0};
    for (unsigned k = 0; k < std::min<size_t>(t, nck); ++k)
        th.emplace_back([&] {
            for (size_t i; (i = next.fetch_add(1)) < nck;) parse_range(data + cuts[i], data + cuts[i + 1], ncols, ck[i]);
        });
    for (auto &x : th) x.join();
    if (data) munmap((void *)data, size);
    close(fd);
    // Ordered pass: fields that depend on earlier chunks.
    uint32_t carry[3] = {0, 0, 0};
    uint64_t total = 0;
    for (size_t i = 0; i < nck; ++i) {
        const uint64_t ne = ck[i].rec.size() / ncols;
        for (uint32_t c = 0; c < ncols; ++c) {
            const uint64_t upto = std::min<uint64_t>(ck[i].first_set[c], ne);
            for (uint64_t j = 0; j < upto; ++j) ck[i].rec[j * ncols + c] = carry[c];
            if (ck[i].first_set[c] != ~0ull) carry[c] = ck[i].last[c];
        }
        total += ne;
    }
    uint32_t *buf = (uint32_t *)malloc(sizeof(uint32_t) * (total ? total * ncols : 1));
    if (!buf) return -4;
    uint64_t o = 0;
    for (size_t i = 0; i < nck; ++i) {
        if (!ck[i].rec.empty()) memcpy(buf + o, ck[i].rec.data(), sizeof(uint32_t) * ck[i].rec.size());
        o += ck[i].rec.size();
    }
    *out = buf;
    *m = total;
    return 0;
}

int pado_read_edge_binary(const char *path, uint32_t **out, uint64_t *m, uint32_t *num_v)
{
    if (!path || !out || !m) return -3;
    *out = nullptr;
    *m = 0;
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    uint32_t nv = 0;
    uint64_t ne = 0;
    if (fread(&nv, 4, 1, f) != 1 || fread(&ne, 8, 1, f) != 1) {
        fclose(f);
        return -2;
    }
    // The header must match the file's size before anything is allocated.
    if (fseek(f, 0, SEEK_END) != 0) {
        fclose(f);
        return -2;
    }
    const long size = ftell(f);
    if (size < 12 || (uint64_t)(size - 12) != ne * 8) {
        fclose(f);
        return -2;
    }
    fseek(f, 12, SEEK_SET);
    uint32_t *buf = (uint32_t *)malloc(sizeof(uint32_t) * (ne ? 2 * ne : 1));
    if (!buf) {
        fclose(f);
        return -4;
    }
    if (ne && fread(buf, 8, ne, f) != ne) {
        free(buf);
        fclose(f);
        return -2;
    }
    fclose(f);
    *out = buf;
    *m = ne;
    if (num_v) *num_v = nv;
    return 0;
}

int pado_write_edge_binary(const char *path, const uint32_t *edges, uint64_t m)
{
    if (!path || (m && !edges)) return -3;
    FILE *f = fopen(path, "wb");
    if (!f) return -1;
    uint32_t nv = 0;
    for (uint64_t i = 0; i < 2 * m; ++i) nv = std::max(nv, edges[i] + 1);
    const bool ok = fwrite(&nv, 4, 1, f) == 1 && fwrite(&m, 8, 1, f) == 1 && (!m || fwrite(edges, 8, m, f) == m);
    return fclose(f) == 0 && ok ? 0 : -1;
}

}  // extern "C"
