/*
 * pado_gpu.h -- C ABI of the MI355X-native PADO library (libpado_gpu.so).
 *
 * The library builds PADO's unweighted pruned landmark labeling (50
 * bit-parallel roots + 1024-root batched pruned BFS) and its weighted variant
 * (512-root batches) on one AMD Instinct MI355X, answers exact distance
 * queries from them (u8 hops / u32 weighted), and serves continuous query
 * streams over a resident index.  Plain C types only;
 * no exceptions or exit() cross this boundary.  Device memory is owned by the
 * opaque handles; host buffers are owned by the caller.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference repository johnpzh/parallel_distance_oracle):
 *   GR   = includes/graph.h
 *   PARA = includes/pado_unw_unv_para.201912311607.parallel_bp_labeling.h
 *   UNP  = includes/pado_unw_unv_unp.h
 *   WPARA= includes/pado_weighted_unv_para.h
 *
 * Id spaces: a graph handed to pado_gpu_graph_upload is in RANK space (the
 * output of Graph::id_transfer, GR:176-206: vertex 0 = highest degree, each
 * adjacency list sorted descending).  Label hubs are rank-space ids, as in the
 * reference's IndexOrdered (PARA:115-121).  Queries take ORIGINAL ids once
 * pado_gpu_index_order has been given the rank array (PARA:2198-2281).
 */
#ifndef PADO_GPU_H_
#define PADO_GPU_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    PADO_GPU_OK = 0,
    PADO_GPU_ERR_ARG = 1,     /* bad argument (null pointer, size, id range)  */
    PADO_GPU_ERR_HIP = 2,     /* HIP runtime error; see pado_gpu_last_error() */
    PADO_GPU_ERR_NOMEM = 3,   /* device or host allocation failed             */
    PADO_GPU_ERR_DOMAIN = 4,  /* graph outside the uint8 distance domain       */
    PADO_GPU_ERR_STATE = 5,   /* call out of order (e.g. query before order)  */
    PADO_GPU_ERR_IO = 6,      /* file could not be opened / read / written     */
    PADO_GPU_ERR_BUSY = 7,    /* serving admission refused (no slot / memory)  */
    PADO_GPU_ERR_TIMEOUT = 8  /* serving wait ended before the request was done */
} pado_gpu_status;

typedef struct pado_gpu_ctx pado_gpu_ctx;
typedef struct pado_gpu_graph pado_gpu_graph;
typedef struct pado_gpu_index pado_gpu_index;

/* Compile-time constants of the reference (PARA:42-47). */
#define PADO_GPU_BP_SIZE 50u
#define PADO_GPU_BATCH_SIZE 1024u
#define PADO_GPU_INF8 255u

/* Text of the last error raised on the calling thread ("" if none). */
const char *pado_gpu_last_error(void);

/* ---------------------------------------------------------------------------
 * Host-side graph ingest (no device needed).
 * ------------------------------------------------------------------------ */

/* Graph(vector<pair>) num_v (GR:129-131): max id + 1. */
uint32_t pado_graph_num_v(const uint32_t *edges, uint64_t m);
/* Graph::construct (GR:125-153): undirected CSR, off[n+1], adj[2m]. */
void pado_graph_csr(const uint32_t *edges, uint64_t m, uint32_t n, uint64_t *off, uint32_t *adj);
/* Graph::make_rank (GR:156-174).  use_global_rand = 1 consumes rand() as the
 * reference does; 0 uses a private generator with the same default seed. */
int pado_make_rank(const uint64_t *off, uint32_t n, int use_global_rand, uint32_t *rank);
/* Graph::id_transfer (GR:176-206); rank2id may be NULL. */
void pado_id_transfer(uint32_t n, const uint64_t *off, const uint32_t *adj, const uint32_t *rank,
                      uint64_t *off_new, uint32_t *adj_new, uint32_t *rank2id);
/* WeightedGraph::construct (GR:313-351): CSR plus u8 weights w[2m] (weights
 * cast to weighti as the reference's file reader does, GR:277-304).
 * WeightedGraph::make_rank (GR:353-372) is the same degree order as
 * Graph::make_rank, so pado_make_rank applies. */
void pado_wgraph_csr(const uint32_t *edges, const uint32_t *weights, uint64_t m, uint32_t n, uint64_t *off,
                     uint32_t *adj, uint8_t *w);
/* WeightedGraph::id_transfer (GR:374-408): lists of (id, weight) sorted
 * descending; rank2id may be NULL. */
void pado_wid_transfer(uint32_t n, const uint64_t *off, const uint32_t *adj, const uint8_t *w,
                       const uint32_t *rank, uint64_t *off_new, uint32_t *adj_new, uint8_t *w_new,
                       uint32_t *rank2id);

/* Edge-list files.  pado_read_edge_text reads a text edge list the way
 * Graph(const char*) (ncols = 2, GR:95-118) or WeightedGraph(const char*)
 * (ncols = 3, GR:277-304) does -- '#' / '%' lines skipped, one edge per other
 * line, malformed lines decided by the reference's istringstream extraction
 * (a missing field keeps the previous line's value) -- memory-mapped and
 * parsed on all host cores.  *out receives m x ncols u32 values (malloc'd;
 * release with pado_free; weights as read, before the u8 cast).
 * pado_read_edge_binary reads the reference's binary edge list
 * (distributed_version/tools/graph-convert/edgelist2binary.cpp:28-91: u32
 * num_v, u64 num_e, num_e u32 pairs); pado_write_edge_binary writes one.
 * Return 0, or -1 (cannot open), -2 (malformed binary), -3 (bad argument),
 * -4 (out of memory). */
int pado_read_edge_text(const char *path, uint32_t ncols, uint32_t **out, uint64_t *m);
int pado_read_edge_binary(const char *path, uint32_t **out, uint64_t *m, uint32_t *num_v);
int pado_write_edge_binary(const char *path, const uint32_t *edges, uint64_t m);
void pado_free(void *p);

/* Vertex ranges of the final canonical label sort (order_labels' per-vertex
 * sort, PARA:2240, done on the device in calls of bounded size): cuts[0] = 0
 * < ... < cuts[r] = n, each range [cuts[i], cuts[i+1]) holding at most
 * max_items labels of the CSR off[n+1].  Returns r, or 0 if one vertex alone
 * exceeds max_items or max_cuts (>= r + 1 entries) is too small. */
uint32_t pado_sort_plan(const uint64_t *off, uint32_t n, uint64_t max_items, uint32_t *cuts, uint32_t max_cuts);

/* Deterministic workload generators (no reference counterpart; they produce
 * the synthetic inputs the reference's scripts would download).  Return the
 * edge count, or the capacity needed if `cap` is too small. */
uint64_t pado_rmat_edges(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                         uint64_t seed, uint32_t *out, uint64_t cap);
uint64_t pado_grid_edges(uint32_t side, double delete_frac, uint64_t seed, uint32_t *out, uint64_t cap);
void pado_edge_weights(uint64_t m, uint32_t lo, uint32_t hi, uint64_t seed, uint32_t *w);
void pado_uniform_pairs(uint32_t n, uint64_t k, uint64_t seed, uint32_t *a, uint32_t *b);
void pado_zipf_vertices(uint32_t n, double s_exp, uint64_t k, uint64_t seed, const uint32_t *rank2id,
                        uint32_t *out);

/* ---------------------------------------------------------------------------
 * Device: context, graph, index build.
 * ------------------------------------------------------------------------ */

/* Number of visible HIP devices (0 if none). */
int pado_gpu_device_count(void);

/* One context per device: a HIP stream at the default priority.  Replaces the
 * reference's implicit OpenMP runtime (construct, PARA:1740-1895). */
pado_gpu_status pado_gpu_ctx_create(int device, pado_gpu_ctx **out);
void pado_gpu_ctx_destroy(pado_gpu_ctx *ctx);
/* The context's hipStream_t, for callers that enqueue their own work. */
void *pado_gpu_ctx_stream(pado_gpu_ctx *ctx);
/* Wait for all work queued on the context's stream. */
pado_gpu_status pado_gpu_ctx_sync(pado_gpu_ctx *ctx);
/* Device-to-device copy of `bytes` bytes with a plain copy kernel on the
 * context's stream (16-byte aligned buffers; no reference counterpart: it is
 * the known-size kernel the HBM byte counters are calibrated with). */
pado_gpu_status pado_gpu_copy_bytes(pado_gpu_ctx *ctx, void *dst, const void *src, uint64_t bytes);

/* Upload a rank-space CSR (the Graph after id_transfer: GR:31-92, 176-206).
 * off: n+1 entries, adj: m_dir entries (each undirected edge twice).
 * weights: NULL for unweighted graphs, else m_dir u8 weights (WeightedGraph,
 * GR:223-424).  The library keeps its own copies. */
pado_gpu_status pado_gpu_graph_upload(pado_gpu_ctx *ctx, uint32_t n, uint64_t m_dir, const uint64_t *off,
                                      const uint32_t *adj, const uint8_t *weights, pado_gpu_graph **out);
void pado_gpu_graph_free(pado_gpu_graph *g);

/* Graph(edge_list) / WeightedGraph(edge_list, weights) + make_rank +
 * id_transfer (GR:120-206, 313-408) on the device: m undirected edges
 * (edges[2i], edges[2i+1]) with optional weights[i] (cast to u8; 0 refused),
 * both directions, duplicates and self-loops kept.  rank_in (nullable, num_v
 * entries) fixes the order; otherwise make_rank runs on the host from the
 * degrees (use_global_rand as for pado_make_rank).  The result is the same
 * rank-space CSR the host path produces.  num_v = max id + 1 (as
 * pado_graph_num_v); rank_out / rank2id_out (nullable) receive num_v entries.
 * Fewer than 2^30 edges. */
pado_gpu_status pado_gpu_graph_from_edges(pado_gpu_ctx *ctx, const uint32_t *edges, uint64_t m,
                                          const uint32_t *weights, const uint32_t *rank_in, int use_global_rand,
                                          pado_gpu_graph **out, uint32_t *num_v, uint32_t *rank_out,
                                          uint32_t *rank2id_out);
/* Copies of a device graph: off[n+1], adj[m_dir], w[m_dir] (each nullable). */
pado_gpu_status pado_gpu_graph_export(pado_gpu_ctx *ctx, const pado_gpu_graph *g, uint64_t *off, uint32_t *adj,
                                      uint8_t *w);
pado_gpu_status pado_gpu_graph_get_info(const pado_gpu_graph *g, uint32_t *num_v, uint64_t *m_dir, int *weighted);

typedef struct {
    uint32_t batch_size;  /* roots per batch: 1..1024, 0 = 1024 (PARA:42); any value
                             gives the same labels, so it is a tuning knob */
    uint32_t flags;       /* PADO_GPU_BUILD_* */
    uint32_t push_blocks; /* grid of the push kernel (blocks of push_waves waves),
                             0 = 256 up to 2^17 vertices, else 768 */
    uint32_t add_blocks;  /* grid of the add kernel (256-thread blocks),
                             0 = 256 up to 2^17 vertices, else 512 */
    uint32_t iters_per_sync; /* batch iterations enqueued per host check of the
                                frontier (iterations past a batch's end are
                                empty launches), 0 = 8 */
    uint32_t sort_max_items; /* labels per call of the final canonical sort
                                (0 = 2^31 - 1025, hipCUB's int item count);
                                larger indexes are sorted in vertex ranges */
    uint32_t push_chunk;  /* edges per push work descriptor: 4, 8, 16 or 32, 0 = 8 */
    uint32_t push_waves;  /* waves per push block: 2, 4 or 8, 0 = 4 */
    uint32_t pull_ratio;  /* direction of the push step: an iteration whose frontier
                             has more push edges than pull_ratio / 1000 of the
                             adjacency entries of the vertices >= roots_start runs
                             in pull form (each tail gathers the roots offered to
                             it; same labels and counters).  0 = 0xFFFFFFFF =
                             always push: on R-MAT 20 every threshold measured
                             slower (400: +0.6%, 100: +11%) */
} pado_gpu_build_cfg;

/* Bit-parallel labeling one BFS per slot (as the reference runs it) instead of
 * the default single multi-source sweep over all 50 slots; same output. */
#define PADO_GPU_BUILD_BP_PER_SLOT 1u
/* The batch step runs by default as captured hipGraphs (batch prologue plus a
 * block of iterations, then iteration blocks until the frontier is empty).
 * With this flag it runs as per-kernel launches instead, each push / add
 * launch bracketed by HIP events (pado_gpu_index_info's ms_push / ms_add).
 * Same output either way. */
#define PADO_GPU_BUILD_STREAM 2u
/* Wide hop domain (no reference counterpart: the reference's u8 level and
 * iteration counters wrap at 255, GLB:29-40, PARA:711/1470).  Labels carry
 * 12-bit hop distances (up to 4094; hubs below 2^20, so at most 2^20
 * vertices), bit-parallel distances stay u8 with the bit-parallel BFS stopped
 * at level 254 (farther vertices keep 255 in that slot, which pruning and
 * queries then skip).  When that cut happens, the bit-parallel roots and
 * their selected neighbours are labeled as ordinary hubs and the batch phase
 * prunes by labels alone (the reference never pushes from those vertices,
 * PARA:1117-1119, and prunes by the slots, both exact only while every slot
 * reaches every vertex); queries still use the slots wherever both ends were
 * reached.  For graphs whose BFS depth reaches 255 -- large
 * grids, road-like graphs.  Every distance stays exact.  On a graph inside
 * the u8 domain (every batch iteration below 253) the label sets equal the
 * default build's.  Query with pado_gpu_query_wide; export with
 * pado_gpu_index_export_wide.  Ignored for weighted graphs. */
#define PADO_GPU_BUILD_WIDE 4u
/* Run the whole batch phase inside one workgroup (no reference counterpart):
 * the same kernels' work back to back with workgroup barriers in place of
 * kernel launches and host checks.  Same labels.  It removes the launch and
 * host-check latency of every iteration but runs on one compute unit, so it
 * pays off only on small graphs (measured on MI355X: 7x faster on a
 * 1K-vertex R-MAT, 1.7x slower on a 1.6K-vertex grid, 20-40x slower on
 * 20K-60K-vertex grids); the autotuner tries it below 4K vertices.  Its label
 * pool is sized up front; if it runs dry the build falls back to the launch
 * schedule.  Ignored with PADO_GPU_BUILD_STREAM. */
#define PADO_GPU_BUILD_DEVICE_LOOP 8u

/* ParaVertexCentricPLL<BATCH_SIZE>(const Graph&) -> construct (PARA:309-313,
 * 1740-1895): bit-parallel labeling (PARA:544-840) then batch_process over all
 * root batches (PARA:1413-1736).  cfg may be NULL.  Returns
 * PADO_GPU_ERR_DOMAIN if a BFS level or batch iteration reaches 255 (the
 * reference's u8 counters wrap there), or with PADO_GPU_BUILD_WIDE if an
 * iteration reaches 4095 or the graph has more than 2^20 vertices. */
pado_gpu_status pado_gpu_build(pado_gpu_ctx *ctx, const pado_gpu_graph *graph, const pado_gpu_build_cfg *cfg,
                               pado_gpu_index **out);

typedef struct {
    uint32_t num_v;
    uint32_t weighted;
    uint64_t num_labels;       /* normal labels (Label_size, PARA:2149) */
    uint32_t ordered;          /* order_labels has run (query by original id) */
    uint32_t bp_max_level;
    uint64_t iterations;       /* batch iterations summed over batches */
    double ms_bp;              /* bit-parallel labeling wall time */
    double ms_batch;           /* batched labeling wall time (incl. final sort) */
    uint32_t bp_roots[50];     /* slot roots (rank space), 0xFFFFFFFF if empty */
    uint64_t pushes;           /* (tail, root) pairs offered by push_labels */
    uint64_t ind_pass;         /* pairs that set a fresh indicator bit */
    uint64_t bp_pruned;        /* pairs pruned by the bit-parallel check (bp_hit_count, PARA:1179) */
    uint64_t cands;            /* candidates reaching distance_query */
    uint64_t dq_pruned;        /* candidates pruned by distance_query */
    uint64_t dq_rows;          /* root-distance rows scanned */
    /* Units of the labeling kernels' algorithmic byte counts: */
    uint64_t actives;          /* active vertices, summed over iterations */
    uint64_t edges;            /* pushed edges (tails >= roots_start) */
    uint64_t tails;            /* distinct tails with a fresh indicator bit, per iteration, summed */
    uint64_t dq_labels;        /* labels with dist < iter read by distance queries */
    uint64_t labels_appended;  /* labels written (== num_labels) */
    /* HIP-event times of the labeling kernels on the build stream: */
    double ms_push, ms_add, ms_bp_levels;
    uint64_t n_push, n_add, n_bp_levels;  /* launches */
    uint32_t dist_bits;        /* label distance bits: 8 (the reference's u8
                                  domain) or 12 (PADO_GPU_BUILD_WIDE) */
    uint32_t pull_iters;       /* batch iterations run in pull form */
} pado_gpu_index_info;

pado_gpu_status pado_gpu_index_get_info(const pado_gpu_index *idx, pado_gpu_index_info *info);

/* Canonical export in RANK space: bp_dist[n*50], bp_sets[n*50*2] (S^-1,
 * S^0\S^-1 per slot), label_off[n+1], hubs[num_labels] ascending per vertex,
 * dists[num_labels].  Any output may be NULL.  This is the (hub, dist) order
 * order_labels produces (PARA:2221-2251), indexed by rank instead of id. */
pado_gpu_status pado_gpu_index_export(pado_gpu_ctx *ctx, const pado_gpu_index *idx, uint8_t *bp_dist,
                                      uint64_t *bp_sets, uint64_t *label_off, uint32_t *hubs, uint8_t *dists);
/* The same export for either hop domain, with u16 label distances (a wide
 * index's labels do not fit export's u8 dists). */
pado_gpu_status pado_gpu_index_export_wide(pado_gpu_ctx *ctx, const pado_gpu_index *idx, uint8_t *bp_dist,
                                           uint64_t *bp_sets, uint64_t *label_off, uint32_t *hubs, uint16_t *dists);
/* Inverse of export (rank space); hubs must ascend per vertex. */
pado_gpu_status pado_gpu_index_import(pado_gpu_ctx *ctx, uint32_t n, const uint8_t *bp_dist, const uint64_t *bp_sets,
                                      const uint64_t *label_off, const uint32_t *hubs, const uint8_t *dists,
                                      pado_gpu_index **out);

/* order_labels(rank2id, rank) (PARA:2198-2281): builds the query copy indexed
 * by ORIGINAL id; rank[x] = rank-space id of original vertex x. */
pado_gpu_status pado_gpu_index_order(pado_gpu_ctx *ctx, pado_gpu_index *idx, const uint32_t *rank, uint32_t n);

#define PADO_GPU_QUERY_DEVICE_PTRS 1u  /* a, b, out are device pointers     */
#define PADO_GPU_QUERY_REF_COMPAT  2u  /* reproduce the reference's unsigned
                                          bit-parallel comparison            */
#define PADO_GPU_QUERY_RANK_SPACE  4u  /* ids are rank-space ids            */

/* query_distance(a, b) (PARA:2283-2416) for k pairs: out[i] = distance in hops,
 * 255 if unreachable or >= 253; ids >= num_v give (a == b ? 0 : 255).  Exact
 * (signed bit-parallel comparison) unless PADO_GPU_QUERY_REF_COMPAT.  With
 * DEVICE_PTRS the call is asynchronous on `stream` (NULL = context stream);
 * otherwise it stages through device buffers and returns when done. */
pado_gpu_status pado_gpu_query(pado_gpu_ctx *ctx, pado_gpu_index *idx, const uint32_t *a, const uint32_t *b, uint64_t k,
                               uint8_t *out, uint32_t flags, void *stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): pado_gpu_query on HOST arrays a, b, out with the same
 * answers and flags (REF_COMPAT, RANK_SPACE; DEVICE_PTRS is refused), staged
 * through a fixed device array the library holds on every GPU it runs on
 * (9 bytes per pair, 9 MiB) instead of allocated buffers.  At most
 * PADO_GPU_NOALLOC_MAX_PAIRS pairs per call (PADO_GPU_ERR_ARG beyond).  It
 * calls no hipMalloc / hipFree.  Runs on the context's stream and returns when
 * the answers are in `out`.  All calls on one GPU share the array: calls that
 * overlap on one GPU (from several threads or contexts) race and must be
 * serialised by the caller. */
/* query_distance(a, b) (PARA:2283-2416, WPARA:2125-2243) for one pair, as the
 * reference's drivers call it in a loop (pado.cpp:114-119): the ids go to
 * the kernel as arguments and the answer comes back through mapped host
 * memory, so a pair costs one launch and one stream synchronisation (no
 * copies).  *out: u8 hops (255 = unreachable or >= 253) for an index in the
 * reference's domain, u16 exact hops (0xFFFF = unreachable) for a wide one,
 * u32 distances for a weighted one (flags as for pado_gpu_query /
 * pado_gpu_wquery: PADO_GPU_QUERY_REF_COMPAT, PADO_GPU_QUERY_RANK_SPACE).
 * Calls on one index are serialised. */
pado_gpu_status pado_gpu_query_pair(pado_gpu_ctx *ctx, pado_gpu_index *idx, uint32_t a, uint32_t b, uint32_t flags,
                                    uint32_t *out);

/* query_distance for an index of either hop domain with u16 answers: a
 * wide-domain index answers exact hop counts, 0xFFFF if unreachable (its
 * bit-parallel part uses only slots reached on both sides; REF_COMPAT is
 * refused, the reference has no such domain); an index in the reference's
 * domain answers what pado_gpu_query answers, widened.  Other flags as for
 * pado_gpu_query. */
#define PADO_GPU_WIDE_INF 0xFFFFu
pado_gpu_status pado_gpu_query_wide(pado_gpu_ctx *ctx, pado_gpu_index *idx, const uint32_t *a, const uint32_t *b,
                                    uint64_t k, uint16_t *out, uint32_t flags, void *stream);

#define PADO_GPU_NOALLOC_MAX_PAIRS (1u << 20)
pado_gpu_status pado_gpu_query_noalloc(pado_gpu_ctx *ctx, pado_gpu_index *idx, const uint32_t *a, const uint32_t *b,
                                       uint64_t k, uint8_t *out, uint32_t flags);

void pado_gpu_index_free(pado_gpu_index *idx);

/* store_index_to_file / load_index_from_file in the reference's format
 * (UNP:1120-1224; PARA:2082-2195 prints only): u32 num_v, u32 50, then per
 * original id 50 x (u8, u64, u64), u32 size, size x (u32 hub, u8 dist). */
pado_gpu_status pado_gpu_index_store(pado_gpu_ctx *ctx, const pado_gpu_index *idx, const char *filename,
                                     const uint32_t *rank, uint32_t n);
/* The loaded index is ordered (query by original id) and has no rank-space
 * copy. */
pado_gpu_status pado_gpu_index_load(pado_gpu_ctx *ctx, const char *filename, pado_gpu_index **out);

/* ---------------------------------------------------------------------------
 * Weighted graphs (WeightedParaVertexCentricPLL<512>, WPARA).  A graph
 * uploaded with weights is built by pado_gpu_build into a weighted index:
 * labels (hub, u32 distance), no bit-parallel part.  The label sets are the
 * canonical pruned landmark labels -- (r, d(v,r)) is a label of v iff r ranks
 * at or above v and no vertex ranked above r lies on a shortest v-r path --
 * which is what the reference's weighted build produces.  cfg->batch_size
 * caps the roots per batch at 512 (0 and larger values mean 512); the first
 * batch takes a quarter of that.
 * ------------------------------------------------------------------------ */
#define PADO_GPU_WINF 0xFFFFFFFFu

/* Canonical rank-space export: label_off[n+1], hubs ascending, u32 dists. */
pado_gpu_status pado_gpu_windex_export(pado_gpu_ctx *ctx, const pado_gpu_index *idx, uint64_t *label_off,
                                       uint32_t *hubs, uint32_t *dists);
/* query_distance (WPARA:2125-2243) for k pairs: exact u32 distances,
 * PADO_GPU_WINF if unreachable; with PADO_GPU_QUERY_REF_COMPAT the reference's
 * int16 domain (>= 32765 -> 32767).  Flags as for pado_gpu_query. */
pado_gpu_status pado_gpu_wquery(pado_gpu_ctx *ctx, pado_gpu_index *idx, const uint32_t *a, const uint32_t *b,
                                uint64_t k, uint32_t *out, uint32_t flags, void *stream);

/* store_index_to_file / load_index_from_file of the weighted class
 * (WPARA:1982-2018, 2021-2055): u32 num_v, then per original id u32 size and
 * size x (u32 rank-space hub, i16 dist), hubs ascending.  store needs the
 * rank array (num_v entries) and returns PADO_GPU_ERR_DOMAIN if a distance
 * does not fit int16; load returns an ordered weighted index (queries by
 * original id through pado_gpu_wquery). */
pado_gpu_status pado_gpu_windex_store(pado_gpu_ctx *ctx, const pado_gpu_index *idx, const char *filename,
                                      const uint32_t *rank, uint32_t n);
pado_gpu_status pado_gpu_windex_load(pado_gpu_ctx *ctx, const char *filename, pado_gpu_index **out);

/* ---------------------------------------------------------------------------
 * Autotuner (no reference counterpart).  Every knob of pado_gpu_build_cfg
 * leaves the labels unchanged, so the tuner may pick any of them purely on
 * speed: it times full trial builds of `graph` in a coordinate search over
 * the push work-descriptor size {4, 8, 16}, push block shape {2, 4, 8}
 * waves, push grid {256..2048} and add grid {256..1024} blocks, iterations
 * per host check {2..16},
 * batch size {256, 512, 1024} and the flags (hipGraph or stream batch step,
 * bit-parallel variant, and below 4K vertices the one-workgroup batch loop;
 * PADO_GPU_BUILD_WIDE of `start` is kept), starting
 * from `start` (NULL = defaults).  One untimed build first, then every
 * configuration is timed as the median of three builds.  max_trials caps the
 * configurations timed (0 = 24); trials_out (nullable) receives their number,
 * best_ms the median of the returned configuration.
 * ------------------------------------------------------------------------ */
pado_gpu_status pado_gpu_autotune(pado_gpu_ctx *ctx, const pado_gpu_graph *graph, const pado_gpu_build_cfg *start,
                                  uint32_t max_trials, pado_gpu_build_cfg *best, double *best_ms,
                                  uint32_t *trials_out);

/* ---------------------------------------------------------------------------
 * Many independent graphs in one job (no reference counterpart: the
 * reference labels one graph per process).  Each graph g is a rank-space CSR
 * of n[g] vertices; `off` concatenates the graphs' offset arrays (n[g] + 1
 * entries each, starting at 0), `adj` their adjacency arrays (local ids,
 * descending per list).  Graphs whose every distance is provably below 255
 * (2 x the eccentricity of one vertex per component) are labeled together as
 * one interleaved union -- label sets depend on neither BATCH_SIZE nor other
 * components, so each graph gets exactly the labels of a build on its own --
 * with every graph keeping its own 50 bit-parallel slots.  The others are
 * built one by one: in the u8 domain if they stay inside it, else in the
 * wide hop domain (PADO_GPU_BUILD_WIDE: a graph whose first bit-parallel
 * root's eccentricity reaches 255 goes there directly); only a graph outside
 * the wide domain too reports PADO_GPU_ERR_DOMAIN.
 * ------------------------------------------------------------------------ */
typedef struct pado_gpu_multi pado_gpu_multi;

pado_gpu_status pado_gpu_build_multi(pado_gpu_ctx *ctx, uint32_t num_graphs, const uint32_t *n, const uint64_t *off,
                                     const uint32_t *adj, const pado_gpu_build_cfg *cfg, pado_gpu_multi **out);
/* The autotuner for a many-graph job: times pado_gpu_build_multi on `sample`
 * of the graphs (every num_graphs / sample-th one, so the mix of kinds and
 * sizes is kept; 0 = all) over iterations per host check {4..32}, push
 * work-descriptor size, push block shape and batch size (grids stay sized per
 * graph), three builds per configuration as pado_gpu_autotune does
 * (max_trials 0 = 12).  Same array layout as pado_gpu_build_multi. */
pado_gpu_status pado_gpu_autotune_multi(pado_gpu_ctx *ctx, uint32_t num_graphs, const uint32_t *n,
                                        const uint64_t *off, const uint32_t *adj, uint32_t sample,
                                        const pado_gpu_build_cfg *start, uint32_t max_trials,
                                        pado_gpu_build_cfg *best, double *best_ms, uint32_t *trials_out);
/* Per-graph result: status (OK or ERR_DOMAIN), label count, whether it was
 * labeled inside the union (1) or on its own (0), and its label distance
 * bits (8: the reference's u8 domain, 12: wide).  Outputs are nullable. */
pado_gpu_status pado_gpu_multi_graph_info(const pado_gpu_multi *mi, uint32_t g, pado_gpu_status *status,
                                          uint64_t *num_labels, int *in_union, uint32_t *dist_bits);
/* Canonical rank-space export of graph g (same layout as pado_gpu_index_export;
 * PADO_GPU_ERR_STATE for a wide-domain graph's u8 distances). */
pado_gpu_status pado_gpu_multi_export(pado_gpu_ctx *ctx, pado_gpu_multi *mi, uint32_t g, uint8_t *bp_dist,
                                      uint64_t *bp_sets, uint64_t *label_off, uint32_t *hubs, uint8_t *dists);
/* The same with u16 distances, for a graph of either domain. */
pado_gpu_status pado_gpu_multi_export_wide(pado_gpu_ctx *ctx, pado_gpu_multi *mi, uint32_t g, uint8_t *bp_dist,
                                           uint64_t *bp_sets, uint64_t *label_off, uint32_t *hubs, uint16_t *dists);
/* query_distance (PARA:2283-2416) on graph g for k pairs of its rank-space
 * ids (host arrays): u16 answers -- for a graph in the u8 domain the
 * reference's answers (255 = unreachable or >= 253), for a wide-domain graph
 * exact hops (0xFFFF = unreachable). */
pado_gpu_status pado_gpu_multi_query(pado_gpu_ctx *ctx, pado_gpu_multi *mi, uint32_t g, const uint32_t *a,
                                     const uint32_t *b, uint64_t k, uint16_t *out);
/* Whole-job figures: number of unions, union vertices, build ms. */
pado_gpu_status pado_gpu_multi_info(const pado_gpu_multi *mi, uint32_t *num_unions, uint64_t *union_vertices,
                                    double *ms_total);
void pado_gpu_multi_free(pado_gpu_multi *mi);

/* ---------------------------------------------------------------------------
 * Serving: a request loop over a resident index (no reference counterpart;
 * it turns query_distance, PARA:2283-2416, into a continuously batched
 * service).  Requests carry any number of (a, b) pairs.  submit() admits a
 * request only if a request slot and staging memory for its pairs are free
 * (PADO_GPU_ERR_BUSY otherwise); a dispatcher thread coalesces admitted
 * requests into query launches on several streams, so requests that arrive
 * while others run are admitted as slots and memory free up and their work
 * overlaps on the device.  Each launch runs the grid-stride label-merge
 * kernel (one wave per pair; the hardware's workgroup dispatch balances long
 * and short label lists), or on request a persistent one (resident
 * workgroups whose waves take pairs from a work counter).  Each answer is computed per pair, so it is
 * the same alone and in any mix.  Unweighted indexes answer u8 hops,
 * weighted ones u32 distances (as pado_gpu_query / pado_gpu_wquery).
 * ------------------------------------------------------------------------ */
typedef struct pado_gpu_server pado_gpu_server;

/* Serve with the grid-stride query kernel: the default; the flag is kept for
 * callers that name it. */
#define PADO_GPU_SERVE_GRID_STRIDE 0x10u
/* Serve with the persistent work-counter kernel instead (one counter per
 * launch, taken 16 pairs at a time; its contended atomics make it slower on
 * a multi-XCD device, kept for comparison). */
#define PADO_GPU_SERVE_PERSISTENT 0x20u
/* Serve with one resident kernel instead of per-launch kernels: the kernel
 * stays on the device while requests flow, reading each request's pairs from
 * the mapped staging ring and writing its answers straight back into mapped
 * host memory.  submit() cuts a request into descriptors of up to 64 pairs
 * spread over eight queues and publishes them in mapped memory; one wave of
 * the kernel mirrors the published counts into device memory, the others take
 * descriptors by ticket, and the wave that finishes a request's last
 * descriptor raises the request's completion flag, which a completer thread
 * polls.  No dispatcher, no copies, no launch per batch.  The kernel leaves
 * after 20 ms without work and at the latest after 100 ms (draining what it
 * took in), and is relaunched while requests are outstanding, so device-wide
 * synchronisation elsewhere in the process waits at most that long behind
 * it.  Requests larger than 2^32 - 1 pairs are not served in this mode. */
#define PADO_GPU_SERVE_RESIDENT 0x40u

typedef struct {
    uint64_t staging_pairs;    /* pair capacity of the admission ring; 0 = derived from free memory */
    uint64_t max_launch_pairs; /* coalescing cap per launch (0 = 4M) */
    uint32_t max_requests;     /* request slots (0 = 4096) */
    uint32_t streams;          /* launches in flight (0 = 3) */
    uint32_t flags;            /* PADO_GPU_QUERY_REF_COMPAT | PADO_GPU_SERVE_* */
    uint32_t min_launch_pairs; /* while launches are in flight, pending pairs that start a
                                  launch at once (0 = 256Ki) */
    uint32_t max_wait_us;      /* ... or the oldest pending request's wait that does (0 = 200) */
    uint32_t reserved[3];
} pado_gpu_serve_cfg;

typedef struct {
    uint64_t requests_done;
    uint64_t pairs_done;
    uint64_t launches;
    uint64_t rejected;         /* submit() calls refused with PADO_GPU_ERR_BUSY */
    uint64_t staging_pairs;    /* admission ring capacity in pairs */
    double busy_ms;            /* sum of launch (copy + kernel) times */
} pado_gpu_serve_stats;

/* The index must be ordered (pado_gpu_index_order); it stays resident and
 * must outlive the server. */
pado_gpu_status pado_gpu_serve_start(pado_gpu_ctx *ctx, pado_gpu_index *idx, const pado_gpu_serve_cfg *cfg,
                                     pado_gpu_server **out);
/* Copies the k pairs into the admission ring and queues them; *req_id
 * identifies the request.  k may be 0. */
pado_gpu_status pado_gpu_serve_submit(pado_gpu_server *srv, const uint32_t *a, const uint32_t *b, uint64_t k,
                                      uint64_t *req_id);
/* *done = 1 once the request's answers are ready. */
pado_gpu_status pado_gpu_serve_poll(pado_gpu_server *srv, uint64_t req_id, int *done);
/* Blocks until done, copies the k answers to out (k u8 values for an
 * unweighted index, k u16 for a wide-domain one, k u32 for a weighted one),
 * releases the slot.  latency_ms (nullable) = completion time - submit time.
 * A request is waited for once: another wait on the same id, concurrent or
 * later, returns PADO_GPU_ERR_ARG.  If its launch failed, the request is
 * released and PADO_GPU_ERR_HIP returned; after a failure the server admits
 * nothing more (submit returns PADO_GPU_ERR_STATE). */
pado_gpu_status pado_gpu_serve_wait(pado_gpu_server *srv, uint64_t req_id, void *out, double *latency_ms);
/* The same, waiting at most timeout_us microseconds (0 = a poll): on
 * PADO_GPU_ERR_TIMEOUT the request stays live, to be waited for again or
 * cancelled. */
pado_gpu_status pado_gpu_serve_wait_for(pado_gpu_server *srv, uint64_t req_id, void *out, double *latency_ms,
                                        uint64_t timeout_us);
/* Withdraws a request: queued, it is never launched; launched, its answers
 * are dropped when the launch ends; done, they are discarded.  Its slot and
 * ring space are released (a launched one's when the launch ends), and its
 * id becomes unknown.  PADO_GPU_ERR_STATE while a wait on it is in progress. */
pado_gpu_status pado_gpu_serve_cancel(pado_gpu_server *srv, uint64_t req_id);
pado_gpu_status pado_gpu_serve_get_stats(pado_gpu_server *srv, pado_gpu_serve_stats *st);
/* Open-loop load generator: request i (pairs [req_off[i], req_off[i+1]) of a
 * and b) is submitted at arrival_s[i] seconds after the start, whatever
 * happened to earlier requests (refused admissions are retried), and waited
 * for in order.  latency_ms[i] = completion - arrival (admission waits
 * included); answers (nullable) receive every request's answers at its
 * offset (u8 or u32 per index type); wall_s = last completion - start.
 * arrival_s must be non-decreasing. */
pado_gpu_status pado_gpu_serve_replay(pado_gpu_server *srv, const uint32_t *a, const uint32_t *b,
                                      const uint64_t *req_off, uint32_t num_req, const double *arrival_s,
                                      double *latency_ms, void *answers, double *wall_s);
/* Drains in-flight launches and joins the dispatcher. */
void pado_gpu_serve_stop(pado_gpu_server *srv);

#ifdef __cplusplus
}
#endif

#endif /* PADO_GPU_H_ */
