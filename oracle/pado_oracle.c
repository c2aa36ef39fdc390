/*
 * pado_oracle.c -- CPU restatement of PADO's unweighted vertex-centric pruned
 * landmark labeling and its distance query.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product in
 * parallel_distance_oracle_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  It is never linked into, called
 * by, or used as a fallback for the product library.
 *
 * Every function follows a named reference function (paths relative to the
 * reference repository):
 *   GR    = includes/graph.h
 *   PARA  = includes/pado_unw_unv_para.201912311607.parallel_bp_labeling.h
 *   UNP   = includes/pado_unw_unv_unp.h
 *
 * The restatement is sequential.  The reference's OpenMP schedule only
 * changes the order of hubs inside one (batch, distance) group: every pruning
 * decision is a set operation (test-and-set indicators, a BP check that
 * depends only on the pair and the iteration, and a distance query over
 * labels of earlier iterations), so the canonical per-vertex label sets are a
 * pure function of the rank-space graph, BATCH_SIZE and BITPARALLEL_SIZE.
 *
 * Pinning (what checks each function independently of this file):
 *   po_graph_*, po_id_transfer : the C++ reference harness (oracle/_ref) dumps
 *       the CSR it builds; CSR-vs-edge-multiset invariants in tests.
 *   po_make_rank              : reference harness rank array; sortedness.
 *   po_bp_labeling            : closed-form BFS checks (bp_dist == BFS distance
 *       from the slot root; S^-1 / S^0 membership from BFS distances to the
 *       selected neighbours) and the reference harness dumps.
 *   po_labeling               : reference harness canonical label dumps
 *       (golden fixtures under tests/golden) and the 2-hop cover property
 *       (signed query == BFS on every pair of small graphs).
 *   po_query                  : BFS (signed mode) and the reference harness's
 *       query_distance answers (ref_compat mode).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#define PO_BP 50          /* BITPARALLEL_SIZE, PARA:45 */

/* Planted-mistake switch, set only by tests/test_planted.py to show that the
 * pinning checks fail on plausible errors.  0 = faithful restatement.
 *   1: bit-parallel neighbours scanned from the START of the adjacency
 *   2: stored S^0 not masked with ~S^-1 (dropped term)
 *   3: bit-parallel prune uses `<` instead of `<=`
 *   4: indicator set AFTER the bit-parallel check instead of before
 *   5: labels recorded with distance iter - 1 (off by one)
 *   6: query intersects labels but skips the bit-parallel part          */
static int po_mutation = 0;
void po_set_mutation(int m) { po_mutation = m; }
#define PO_INF8 255u      /* SMALLI_MAX / WEIGHTI_MAX, GLB:37-39 */

/* ------------------------------------------------------------------------ */
/* Graph construction: GR:120-153 (Graph(edge_list) / construct).            */
/* num_v = max id + 1 (GR:130); both directions appended, no de-duplication, */
/* no self-loop removal (GR:138-141); adjacency in input order.              */
/* ------------------------------------------------------------------------ */
uint32_t po_graph_num_v(const uint32_t *edges, uint64_t m)
{
    uint32_t n = 0;
    for (uint64_t i = 0; i < 2 * m; ++i)
        if (edges[i] + 1 > n) n = edges[i] + 1;
    return n;
}

/* off: n+1 entries, adj: 2m entries. */
void po_graph_csr(const uint32_t *edges, uint64_t m, uint32_t n,
                  uint64_t *off, uint32_t *adj)
{
    memset(off, 0, sizeof(uint64_t) * (n + 1));
    for (uint64_t i = 0; i < m; ++i) {
        off[edges[2 * i] + 1]++;
        off[edges[2 * i + 1] + 1]++;
    }
    for (uint32_t v = 0; v < n; ++v) off[v + 1] += off[v];
    uint64_t *pos = (uint64_t *)malloc(sizeof(uint64_t) * (n + 1));
    memcpy(pos, off, sizeof(uint64_t) * (n + 1));
    /* GR:138-141 pushes head->tail then tail->head per edge, in edge order. */
    for (uint64_t i = 0; i < m; ++i) {
        uint32_t a = edges[2 * i], b = edges[2 * i + 1];
        adj[pos[a]++] = b;
        adj[pos[b]++] = a;
    }
    free(pos);
}

/* ------------------------------------------------------------------------ */
/* make_rank: GR:156-174.  key = float(deg) + float(rand())/RAND_MAX, sorted */
/* descending as std::pair<float, idi> (ties: larger id first).  rand() is   */
/* glibc's with its default seed 1 (srand is never called); random_r with a  */
/* TYPE_3 state seeded with 1 yields the same sequence without touching the  */
/* process-wide generator.                                                   */
/* ------------------------------------------------------------------------ */
typedef struct { float key; uint32_t id; } po_key;

static int po_key_cmp_desc(const void *pa, const void *pb)
{
    const po_key *a = (const po_key *)pa, *b = (const po_key *)pb;
    if (a->key != b->key) return a->key < b->key ? 1 : -1;
    if (a->id != b->id) return a->id < b->id ? 1 : -1;
    return 0;
}

#include <stdlib.h>
struct random_data;
int po_make_rank(const uint64_t *off, uint32_t n, uint32_t *rank)
{
    struct random_data *rd = (struct random_data *)calloc(1, 64);
    char *state = (char *)calloc(1, 128);
    /* initstate_r/random_r are glibc extensions declared with _DEFAULT_SOURCE */
    extern int initstate_r(unsigned int, char *, size_t, struct random_data *);
    extern int random_r(struct random_data *, int32_t *);
    if (initstate_r(1u, state, 128, rd) != 0) { free(rd); free(state); return -1; }
    po_key *k = (po_key *)malloc(sizeof(po_key) * (n ? n : 1));
    for (uint32_t v = 0; v < n; ++v) {
        int32_t r;
        random_r(rd, &r);
        uint32_t deg = (uint32_t)(off[v + 1] - off[v]);
        /* float(rand()) / RAND_MAX: int RAND_MAX converts to float first. */
        float frac = (float)r / (float)RAND_MAX;
        k[v].key = (float)deg + frac;
        k[v].id = v;
    }
    qsort(k, n, sizeof(po_key), po_key_cmp_desc);
    for (uint32_t r = 0; r < n; ++r) rank[k[r].id] = r;
    free(k); free(rd); free(state);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* id_transfer: GR:176-206.  Renumber by rank, sort every adjacency list     */
/* DESCENDING ("lower rank first", GR:192), return rank2id.                  */
/* ------------------------------------------------------------------------ */
static int po_u32_desc(const void *a, const void *b)
{
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return x < y ? 1 : (x > y ? -1 : 0);
}

void po_id_transfer(uint32_t n, const uint64_t *off, const uint32_t *adj,
                    const uint32_t *rank, uint64_t *off_new, uint32_t *adj_new,
                    uint32_t *rank2id)
{
    for (uint32_t v = 0; v < n; ++v) rank2id[rank[v]] = v;
    off_new[0] = 0;
    for (uint32_t nv = 0; nv < n; ++nv) {
        uint32_t v = rank2id[nv];
        off_new[nv + 1] = off_new[nv] + (off[v + 1] - off[v]);
    }
    for (uint32_t nv = 0; nv < n; ++nv) {
        uint32_t v = rank2id[nv];
        uint64_t o = off_new[nv];
        for (uint64_t e = off[v]; e < off[v + 1]; ++e) adj_new[o++] = rank[adj[e]];
        qsort(adj_new + off_new[nv], off_new[nv + 1] - off_new[nv], sizeof(uint32_t), po_u32_desc);
    }
}

/* ------------------------------------------------------------------------ */
/* BFS checker (plain textbook BFS, hop counts, UINT32_MAX = unreachable).   */
/* ------------------------------------------------------------------------ */
void po_bfs(uint32_t n, const uint64_t *off, const uint32_t *adj, uint32_t src,
            uint32_t *dist)
{
    uint32_t *q = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (uint32_t v = 0; v < n; ++v) dist[v] = UINT32_MAX;
    uint32_t h = 0, t = 0;
    dist[src] = 0; q[t++] = src;
    while (h < t) {
        uint32_t v = q[h++];
        for (uint64_t e = off[v]; e < off[v + 1]; ++e) {
            uint32_t w = adj[e];
            if (dist[w] == UINT32_MAX) { dist[w] = dist[v] + 1; q[t++] = w; }
        }
    }
    free(q);
}

/* BFS answers for k (a, b) pairs: one BFS per distinct source.  out[i] = hop */
/* distance, UINT32_MAX if unreachable.                                      */
static const uint32_t *po_sort_key;
static int po_idx_by_key(const void *x, const void *y)
{
    uint32_t a = po_sort_key[*(const uint64_t *)x], b = po_sort_key[*(const uint64_t *)y];
    return a < b ? -1 : (a > b ? 1 : 0);
}

void po_bfs_pairs(uint32_t n, const uint64_t *off, const uint32_t *adj, const uint32_t *a,
                  const uint32_t *b, uint64_t k, uint32_t *out)
{
    uint64_t *ord = (uint64_t *)malloc(sizeof(uint64_t) * (k ? k : 1));
    for (uint64_t i = 0; i < k; ++i) ord[i] = i;
    po_sort_key = a;
    qsort(ord, k, sizeof(uint64_t), po_idx_by_key);
    uint32_t *dist = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t cur = UINT32_MAX;
    for (uint64_t j = 0; j < k; ++j) {
        uint64_t i = ord[j];
        if (a[i] >= n || b[i] >= n) { out[i] = (a[i] == b[i]) ? 0 : UINT32_MAX; continue; }
        if (a[i] != cur) { cur = a[i]; po_bfs(n, off, adj, cur, dist); }
        out[i] = dist[b[i]];
    }
    free(ord); free(dist);
}

/* ------------------------------------------------------------------------ */
/* Bit-parallel labeling: PARA:544-840 (semantics identical in its serial    */
/* branch PARA:553-659 and parallel branch PARA:660-836).                    */
/*   bp_d  : n*50 u8   (L[v].bp_dist[i])                                     */
/*   bp_s  : n*50*2 u64 ([v][i][0] = S^-1, [v][i][1] = S^0 \ S^-1)          */
/*   used  : n u8      (used_bp_roots)                                       */
/*   roots : 50 u32    (slot root, or UINT32_MAX if the slot is empty)       */
/* Returns the largest BFS level reached (the reference's level counter is   */
/* a u8 that wraps at 256, PARA:711; callers reject eccentricity >= 254).    */
/* ------------------------------------------------------------------------ */
uint32_t po_bp_labeling(uint32_t n, const uint64_t *off, const uint32_t *adj,
                        uint8_t *bp_d, uint64_t *bp_s, uint8_t *used,
                        uint32_t *roots)
{
    uint32_t *tmp_d = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint64_t *s1 = (uint64_t *)malloc(sizeof(uint64_t) * (n ? n : 1)); /* S^-1 */
    uint64_t *s0 = (uint64_t *)malloc(sizeof(uint64_t) * (n ? n : 1)); /* S^0  */
    uint32_t *que = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint64_t m = off[n];
    uint32_t (*sib)[2] = malloc(sizeof(uint32_t) * 2 * (m ? m : 1));
    uint32_t (*chd)[2] = malloc(sizeof(uint32_t) * 2 * (m ? m : 1));
    uint32_t max_level = 0;
    memset(bp_s, 0, sizeof(uint64_t) * 2 * PO_BP * (uint64_t)n);
    uint32_t r = 0;
    for (uint32_t i = 0; i < PO_BP; ++i) {
        while (r < n && used[r]) ++r;                         /* PARA:671-673 */
        if (r == n) {                                          /* PARA:674-679 */
            for (uint32_t v = 0; v < n; ++v) bp_d[(uint64_t)v * PO_BP + i] = PO_INF8;
            roots[i] = UINT32_MAX;
            continue;
        }
        roots[i] = r;
        used[r] = 1;
        for (uint32_t v = 0; v < n; ++v) { tmp_d[v] = PO_INF8; s1[v] = 0; s0[v] = 0; }
        uint32_t qt0 = 0, qt1, qh = 0;
        que[qh++] = r; tmp_d[r] = 0; qt1 = qh;
        /* Up to 64 unused neighbours, scanned from the END of the descending  */
        /* adjacency (highest rank first), PARA:690-709.                       */
        int ns = 0;
        uint64_t deg = off[r + 1] - off[r];
        for (uint64_t di = 0; di < deg; ++di) {
            uint32_t v = adj[off[r] + (po_mutation == 1 ? di : deg - 1 - di)];
            if (!used[v]) {
                used[v] = 1;
                que[qh++] = v;
                tmp_d[v] = 1;
                s1[v] = 1ULL << ns;
                if (++ns == 64) break;
            }
        }
        for (uint32_t d = 0; qt0 < qh; ++d) {                  /* PARA:711 */
            uint64_t nsib = 0, nchd = 0;
            if (d > max_level) max_level = d;
            for (uint32_t qi = qt0; qi < qt1; ++qi) {
                uint32_t v = que[qi];
                for (uint64_t e = off[v]; e < off[v + 1]; ++e) {
                    uint32_t tv = adj[e];
                    if (d > tmp_d[tv]) {
                    } else if (d == tmp_d[tv]) {
                        if (v < tv) { sib[nsib][0] = v; sib[nsib][1] = tv; ++nsib; }
                    } else {
                        if (tmp_d[tv] == PO_INF8) { tmp_d[tv] = d + 1; que[qh++] = tv; }
                        chd[nchd][0] = v; chd[nchd][1] = tv; ++nchd;
                    }
                }
            }
            for (uint64_t k = 0; k < nsib; ++k) {              /* PARA:783-792 */
                uint32_t v = sib[k][0], w = sib[k][1];
                s0[v] |= s1[w];
                s0[w] |= s1[v];
            }
            for (uint64_t k = 0; k < nchd; ++k) {              /* PARA:804-813 */
                uint32_t v = chd[k][0], c = chd[k][1];
                s1[c] |= s1[v];
                s0[c] |= s0[v];
            }
            qt0 = qt1; qt1 = qh;
        }
        for (uint32_t v = 0; v < n; ++v) {                     /* PARA:829-835 */
            bp_d[(uint64_t)v * PO_BP + i] = (uint8_t)tmp_d[v];
            bp_s[((uint64_t)v * PO_BP + i) * 2 + 0] = s1[v];
            bp_s[((uint64_t)v * PO_BP + i) * 2 + 1] = po_mutation == 2 ? s0[v] : (s0[v] & ~s1[v]);
        }
    }
    free(tmp_d); free(s1); free(s0); free(que); free(sib); free(chd);
    return max_level;
}

/* ------------------------------------------------------------------------ */
/* Batched labeling: PARA:848-1736 (initialize, push_labels, distance_query, */
/* insert_label_only, update_label_indices, reset_at_end, batch_process) and */
/* construct PARA:1740-1836.                                                 */
/* ------------------------------------------------------------------------ */
typedef struct {
    uint32_t n, batch;
    uint32_t *len, *cap;        /* per-vertex label list, insertion order */
    uint32_t **hub;             /* rank-space hub ids */
    uint8_t **dist;
    uint32_t *last_lo, *last_hi;/* last distance group (PARA:1109-1110) */
    /* statistics, per build */
    uint64_t n_iters, n_pushes, n_ind_pass, n_bp_pruned, n_cands, n_dq_pruned;
    uint64_t bytes_alg;         /* algorithmic HBM bytes of the batch phase */
} po_index;

static void po_lab_push(po_index *X, uint32_t v, uint32_t h, uint8_t d)
{
    if (X->len[v] == X->cap[v]) {
        uint32_t c = X->cap[v] ? X->cap[v] * 2 : 4;
        X->hub[v] = (uint32_t *)realloc(X->hub[v], sizeof(uint32_t) * c);
        X->dist[v] = (uint8_t *)realloc(X->dist[v], c);
        X->cap[v] = c;
    }
    X->hub[v][X->len[v]] = h;
    X->dist[v][X->len[v]] = d;
    X->len[v]++;
}

/* BP check inside push_labels, PARA:1164-1186, with its unsigned arithmetic:
 * `inti td = ...; if (td - 2 <= iter)` wraps for td < 2 (slot skipped). */
static int po_bp_prunes(const uint8_t *bp_d, const uint64_t *bp_s,
                        uint32_t a, uint32_t b, uint32_t iter)
{
    const uint8_t *da = bp_d + (uint64_t)a * PO_BP, *db = bp_d + (uint64_t)b * PO_BP;
    const uint64_t *sa = bp_s + (uint64_t)a * PO_BP * 2, *sb = bp_s + (uint64_t)b * PO_BP * 2;
    for (uint32_t i = 0; i < PO_BP; ++i) {
        uint32_t td = (uint32_t)da[i] + (uint32_t)db[i];
        if (td - 2u <= iter) {
            if (sa[2 * i] & sb[2 * i]) td -= 2u;
            else if ((sa[2 * i] & sb[2 * i + 1]) | (sa[2 * i + 1] & sb[2 * i])) td -= 1u;
            if (po_mutation == 3 ? (td < iter) : (td <= iter)) return 1;
        }
    }
    return 0;
}

po_index *po_labeling(uint32_t n, const uint64_t *off, const uint32_t *adj,
                      const uint8_t *bp_d, const uint64_t *bp_s, const uint8_t *used,
                      uint32_t batch)
{
    po_index *X = (po_index *)calloc(1, sizeof(po_index));
    X->n = n; X->batch = batch;
    X->len = calloc(n ? n : 1, sizeof(uint32_t));
    X->cap = calloc(n ? n : 1, sizeof(uint32_t));
    X->hub = calloc(n ? n : 1, sizeof(uint32_t *));
    X->dist = calloc(n ? n : 1, sizeof(uint8_t *));
    X->last_lo = calloc(n ? n : 1, sizeof(uint32_t));
    X->last_hi = calloc(n ? n : 1, sizeof(uint32_t));
    const uint32_t W = (batch + 63) / 64;      /* words per indicator row */
    uint64_t *ind = calloc((uint64_t)(n ? n : 1) * W, sizeof(uint64_t));
    uint8_t *once = calloc(n ? n : 1, 1);
    uint32_t *once_q = malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t n_once = 0;
    uint8_t *dm = malloc((uint64_t)batch * (n ? n : 1));   /* dist_matrix[B][n] */
    memset(dm, PO_INF8, (uint64_t)batch * n);
    uint32_t *active = malloc(sizeof(uint32_t) * (n ? n : 1));
    uint8_t *is_active = calloc(n ? n : 1, 1);
    uint32_t *cq = malloc(sizeof(uint32_t) * (n ? n : 1));  /* candidate_queue */
    uint8_t *got = calloc(n ? n : 1, 1);
    uint32_t *ncand = calloc(n ? n : 1, sizeof(uint32_t));
    uint32_t **cand = calloc(n ? n : 1, sizeof(uint32_t *));
    uint32_t *ccap = calloc(n ? n : 1, sizeof(uint32_t));
    uint8_t *tail_seen = calloc(n ? n : 1, 1);
    uint32_t *tail_q = malloc(sizeof(uint32_t) * (n ? n : 1));

    uint32_t nb = (n + batch - 1) / batch;
    for (uint32_t b = 0; b < nb; ++b) {
        uint32_t rs = b * batch;
        uint32_t rsz = (n - rs < batch) ? (n - rs) : batch;
        uint32_t n_act = 0;
        /* initialize, PARA:848-1080 */
        for (uint32_t r = rs; r < rs + rsz; ++r)
            if (!used[r]) active[n_act++] = r;
        for (uint32_t k = 0; k < n_once; ++k) {
            uint32_t v = once_q[k];
            memset(ind + (uint64_t)v * W, 0, sizeof(uint64_t) * W);
            once[v] = 0;
        }
        n_once = 0;
        for (uint32_t r = rs; r < rs + rsz; ++r) {
            if (used[r]) continue;
            uint32_t ro = r - rs;
            ind[(uint64_t)r * W + ro / 64] |= 1ULL << (ro % 64);
            X->last_lo[r] = X->len[r];
            po_lab_push(X, r, r, 0);                       /* PARA:932-949 */
            X->last_hi[r] = X->len[r];
            X->bytes_alg += 17ull * PO_BP;                 /* root BP record */
        }
        for (uint32_t r = rs; r < rs + rsz; ++r) {          /* PARA:986-1013 */
            if (used[r]) continue;
            uint8_t *row = dm + (uint64_t)(r - rs) * n;
            for (uint32_t l = 0; l < X->len[r]; ++l) row[X->hub[r][l]] = X->dist[r][l];
        }
        uint32_t iter = 0;
        while (n_act) {
            ++iter;
            X->n_iters++;
            uint32_t n_cq = 0, n_tail = 0;
            /* push, PARA:1486-1559 -> push_labels PARA:1084-1211 */
            for (uint32_t ai = 0; ai < n_act; ++ai) {
                uint32_t vh = active[ai];
                is_active[vh] = 0;
                X->bytes_alg += 8;
                if (used[vh]) continue;                    /* PARA:1117-1119 */
                for (uint64_t e = off[vh]; e < off[vh + 1]; ++e) {
                    uint32_t vt = adj[e];
                    if (vt < rs) break;                    /* PARA:1121-1123 */
                    X->bytes_alg += 4;
                    for (uint32_t li = X->last_lo[vh]; li < X->last_hi[vh]; ++li) {
                        uint32_t rr = X->hub[vh][li];      /* real id of the label root */
                        if (vt <= rr) continue;            /* PARA:1134-1139 */
                        uint32_t ro = rr - rs;
                        X->n_pushes++;
                        uint64_t *w = ind + (uint64_t)vt * W + ro / 64;
                        uint64_t bit = 1ULL << (ro % 64);
                        if (*w & bit) continue;            /* PARA:1147-1153 */
                        if (po_mutation == 4 && po_bp_prunes(bp_d, bp_s, rr, vt, iter)) { X->n_bp_pruned++; continue; }
                        *w |= bit;
                        X->n_ind_pass++;
                        if (!tail_seen[vt]) { tail_seen[vt] = 1; tail_q[n_tail++] = vt; }
                        if (!once[vt]) { once[vt] = 1; once_q[n_once++] = vt; }
                        if (po_mutation != 4 && po_bp_prunes(bp_d, bp_s, rr, vt, iter)) { X->n_bp_pruned++; continue; }
                        /* is_candidate + candidates_que: a (tail, root) pair
                         * passes the indicator at most once per batch, so the
                         * queue never holds duplicates. PARA:1194-1198 */
                        if (ncand[vt] == ccap[vt]) {
                            ccap[vt] = ccap[vt] ? ccap[vt] * 2 : 4;
                            cand[vt] = realloc(cand[vt], sizeof(uint32_t) * ccap[vt]);
                        }
                        cand[vt][ncand[vt]++] = ro;
                        X->n_cands++;
                        if (!got[vt]) { got[vt] = 1; cq[n_cq++] = vt; }   /* PARA:1201-1206 */
                    }
                }
            }
            for (uint32_t k = 0; k < n_tail; ++k) tail_seen[tail_q[k]] = 0;
            X->bytes_alg += (uint64_t)n_tail * 17ull * PO_BP;
            n_act = 0;
            if (n_cq == 0) break;                          /* PARA:1562-1564 */
            /* add, PARA:1568-1665 */
            for (uint32_t ci = 0; ci < n_cq; ++ci) {
                uint32_t v = cq[ci];
                got[v] = 0;
                uint32_t lo = X->len[v];
                uint32_t nprev = 0;
                for (uint32_t l = 0; l < lo; ++l) nprev += X->dist[v][l] < iter;
                X->bytes_alg += 5ull * nprev;
                for (uint32_t k = 0; k < ncand[v]; ++k) {
                    uint32_t ro = cand[v][k];
                    uint32_t rr = ro + rs;
                    const uint8_t *row = dm + (uint64_t)ro * n;
                    int prune = 0;
                    /* distance_query, PARA:1218-1270: labels inserted in this
                     * iteration (index >= lo) have dist == iter and are not
                     * yet visible through the distance groups. */
                    for (uint32_t l = 0; l < lo; ++l) {
                        uint32_t d = X->dist[v][l];
                        if (d >= iter) continue;
                        uint32_t h = X->hub[v][l];
                        if (h >= rr) continue;
                        if (d + (uint32_t)row[h] <= iter) { prune = 1; break; }
                    }
                    if (prune) { X->n_dq_pruned++; continue; }
                    if (!is_active[v]) { is_active[v] = 1; active[n_act++] = v; }
                    po_lab_push(X, v, rr, (uint8_t)(po_mutation == 5 ? iter - 1 : iter));  /* insert_label_only */
                    if (v >= rs && v - rs < rsz)           /* PARA:1287-1290 */
                        dm[(uint64_t)(v - rs) * n + rr] = (uint8_t)iter;
                }
                ncand[v] = 0;
                if (X->len[v] != lo) {                     /* update_label_indices */
                    X->last_lo[v] = lo;
                    X->last_hi[v] = X->len[v];
                    X->bytes_alg += 5ull * (X->len[v] - lo);
                }
            }
        }
        /* reset_at_end, PARA:1328-1410 */
        for (uint32_t r = rs; r < rs + rsz; ++r) {
            uint8_t *row = dm + (uint64_t)(r - rs) * n;
            for (uint32_t l = 0; l < X->len[r]; ++l) row[X->hub[r][l]] = PO_INF8;
        }
    }
    free(ind); free(once); free(once_q); free(dm); free(active); free(is_active);
    free(cq); free(got); free(ncand);
    for (uint32_t v = 0; v < n; ++v) free(cand[v]);
    free(cand); free(ccap); free(tail_seen); free(tail_q);
    return X;
}

uint64_t po_index_num_labels(const po_index *X)
{
    uint64_t s = 0;
    for (uint32_t v = 0; v < X->n; ++v) s += X->len[v];
    return s;
}

void po_index_stats(const po_index *X, uint64_t *out7)
{
    out7[0] = X->n_iters; out7[1] = X->n_pushes; out7[2] = X->n_ind_pass;
    out7[3] = X->n_bp_pruned; out7[4] = X->n_cands; out7[5] = X->n_dq_pruned;
    out7[6] = X->bytes_alg;
}

/* Canonical export: per rank-space vertex, (hub, dist) sorted by hub, which  */
/* is std::pair order since hubs are unique per vertex (order_labels,         */
/* PARA:2221-2240).  off: n+1 u64; hub/dist: num_labels entries.              */
typedef struct { uint32_t h; uint8_t d; } po_hd;
static int po_hd_cmp(const void *a, const void *b)
{
    const po_hd *x = (const po_hd *)a, *y = (const po_hd *)b;
    if (x->h != y->h) return x->h < y->h ? -1 : 1;
    return (int)x->d - (int)y->d;
}

void po_index_export(const po_index *X, uint64_t *off, uint32_t *hub, uint8_t *dist)
{
    off[0] = 0;
    po_hd *tmp = NULL; uint32_t tcap = 0;
    for (uint32_t v = 0; v < X->n; ++v) {
        uint32_t L = X->len[v];
        if (L > tcap) { tcap = L; tmp = realloc(tmp, sizeof(po_hd) * tcap); }
        for (uint32_t l = 0; l < L; ++l) { tmp[l].h = X->hub[v][l]; tmp[l].d = X->dist[v][l]; }
        qsort(tmp, L, sizeof(po_hd), po_hd_cmp);
        for (uint32_t l = 0; l < L; ++l) { hub[off[v] + l] = tmp[l].h; dist[off[v] + l] = tmp[l].d; }
        off[v + 1] = off[v] + L;
    }
    free(tmp);
}

void po_index_free(po_index *X)
{
    if (!X) return;
    for (uint32_t v = 0; v < X->n; ++v) { free(X->hub[v]); free(X->dist[v]); }
    free(X->hub); free(X->dist); free(X->len); free(X->cap);
    free(X->last_lo); free(X->last_hi); free(X);
}

/* ------------------------------------------------------------------------ */
/* query_distance: PARA:2283-2416 over a canonical index in rank space.      */
/* ref_compat = 1 reproduces the reference's unsigned comparison             */
/* `int td; inti d; if (td - 2 <= d)` (PARA:2306, 2315-2316), which skips a  */
/* slot whenever td < 2; ref_compat = 0 compares signed, which is exact.     */
/* ------------------------------------------------------------------------ */
uint32_t po_query(uint32_t n, const uint8_t *bp_d, const uint64_t *bp_s,
                  const uint64_t *off, const uint32_t *hub, const uint8_t *dist,
                  uint32_t a, uint32_t b, int ref_compat)
{
    if (a >= n || b >= n) return a == b ? 0u : PO_INF8;   /* PARA:2289-2291 */
    uint32_t d = PO_INF8;
    const uint8_t *da = bp_d + (uint64_t)a * PO_BP, *db = bp_d + (uint64_t)b * PO_BP;
    const uint64_t *sa = bp_s + (uint64_t)a * PO_BP * 2, *sb = bp_s + (uint64_t)b * PO_BP * 2;
    for (uint32_t i = 0; i < (po_mutation == 6 ? 0u : PO_BP); ++i) {
        int td = (int)da[i] + (int)db[i];
        int gate = ref_compat ? ((uint32_t)(td - 2) <= d) : (td - 2 <= (int)d);
        if (gate) {
            if (sa[2 * i] & sb[2 * i]) td -= 2;
            else if ((sa[2 * i] & sb[2 * i + 1]) | (sa[2 * i + 1] & sb[2 * i])) td -= 1;
            if ((uint32_t)td < d) d = (uint32_t)td;
        }
    }
    uint64_t i1 = off[a], e1 = off[a + 1], i2 = off[b], e2 = off[b + 1];
    while (i1 < e1 && i2 < e2) {                           /* PARA:2394-2410 */
        uint32_t v1 = hub[i1], v2 = hub[i2];
        if (v1 == v2) {
            uint32_t td = (uint32_t)dist[i1] + dist[i2];
            if (td < d) d = td;
            ++i1; ++i2;
        } else if (v1 < v2) ++i1; else ++i2;
    }
    if (d >= PO_INF8 - 2) d = PO_INF8;                     /* PARA:2412-2414 */
    return d;
}

void po_query_batch(uint32_t n, const uint8_t *bp_d, const uint64_t *bp_s,
                    const uint64_t *off, const uint32_t *hub, const uint8_t *dist,
                    const uint32_t *a, const uint32_t *b, uint64_t k,
                    uint8_t *out, int ref_compat)
{
    for (uint64_t i = 0; i < k; ++i)
        out[i] = (uint8_t)po_query(n, bp_d, bp_s, off, hub, dist, a[i], b[i], ref_compat);
}
