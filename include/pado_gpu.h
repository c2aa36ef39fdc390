/*
 * pado_gpu.h -- C ABI of the MI355X-native PADO library (libpado_gpu.so).
 *
 * The library builds PADO's unweighted pruned landmark labeling (50
 * bit-parallel roots + 1024-root batched pruned BFS) on one AMD Instinct
 * MI355X and answers exact hop-distance queries from it.  Plain C types only;
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
    PADO_GPU_ERR_BUSY = 7     /* serving admission refused (no slot / memory)  */
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

#ifdef __cplusplus
}
#endif

#endif /* PADO_GPU_H_ */
