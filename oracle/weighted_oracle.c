/*
 * weighted_oracle.c -- CPU restatement of PADO's weighted vertex-centric
 * labeling (WeightedParaVertexCentricPLL<BATCH_SIZE>) and its query.
 *
 * TEST INFRASTRUCTURE ONLY (see pado_oracle.c's header for the rules).
 *
 * Reference (relative to the reference repository):
 *   WPARA = includes/pado_weighted_unv_para.h
 *   GR    = includes/graph.h (WeightedGraph, :223-424)
 *
 * This restates the reference's schedule with ONE thread, queue by queue: the
 * order of the active queue, of each vertex's candidate queue, of its
 * current-batch label queue and of the has-new-labels queue.  The reference's
 * outcome can depend on those orders (its distance query returns the FIRST
 * witness sum, which tightens the filter table, and filter_out_labels removes
 * labels in queue order), so parity is defined against the reference run with
 * OMP_NUM_THREADS=1.  It keeps the reference's index arithmetic as written,
 * including distance_query reading short_index[cand_root_id] with the root
 * OFFSET (WPARA:630; filter_out_labels uses the real id, WPARA:1132).
 *
 * Pinned by: the reference harness (oracle/_ref, weighted mode, one thread)
 * label dumps and answers; Dijkstra (exact query answers).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define WO_INF 32767  /* WEIGHTILARGE_MAX, GLB:34 */

typedef int16_t wl;   /* weightiLarge */

/* id_transfer for WeightedGraph (GR:374-408): adjacency of each vertex sorted
 * descending as (id, weight) pairs. */
typedef struct { uint32_t id; uint8_t w; } wo_pair;
static int wo_pair_desc(const void *a, const void *b)
{
    const wo_pair *x = (const wo_pair *)a, *y = (const wo_pair *)b;
    if (x->id != y->id) return x->id < y->id ? 1 : -1;
    return (int)y->w - (int)x->w;
}

void wo_id_transfer(uint32_t n, const uint64_t *off, const uint32_t *adj, const uint8_t *w, const uint32_t *rank,
                    uint64_t *off_new, uint32_t *adj_new, uint8_t *w_new, uint32_t *rank2id)
{
    for (uint32_t v = 0; v < n; ++v) rank2id[rank[v]] = v;
    off_new[0] = 0;
    for (uint32_t nv = 0; nv < n; ++nv) {
        uint32_t v = rank2id[nv];
        off_new[nv + 1] = off_new[nv] + (off[v + 1] - off[v]);
    }
    wo_pair *tmp = NULL;
    uint64_t cap = 0;
    for (uint32_t nv = 0; nv < n; ++nv) {
        uint32_t v = rank2id[nv];
        uint64_t d = off[v + 1] - off[v];
        if (d > cap) { cap = d; tmp = realloc(tmp, sizeof(wo_pair) * cap); }
        for (uint64_t e = 0; e < d; ++e) { tmp[e].id = rank[adj[off[v] + e]]; tmp[e].w = w[off[v] + e]; }
        qsort(tmp, d, sizeof(wo_pair), wo_pair_desc);
        for (uint64_t e = 0; e < d; ++e) { adj_new[off_new[nv] + e] = tmp[e].id; w_new[off_new[nv] + e] = tmp[e].w; }
    }
    free(tmp);
}

typedef struct {
    uint32_t n, B;
    uint32_t *len, *cap;
    uint32_t **hub;  /* real ids, insertion order */
    wl **dist;
} wo_index;

static void wo_push(wo_index *X, uint32_t v, uint32_t h, wl d)
{
    if (X->len[v] == X->cap[v]) {
        uint32_t c = X->cap[v] ? X->cap[v] * 2 : 4;
        X->hub[v] = realloc(X->hub[v], sizeof(uint32_t) * c);
        X->dist[v] = realloc(X->dist[v], sizeof(wl) * c);
        X->cap[v] = c;
    }
    X->hub[v][X->len[v]] = h;
    X->dist[v][X->len[v]] = d;
    X->len[v]++;
}

wo_index *wo_labeling(uint32_t n, const uint64_t *off, const uint32_t *adj, const uint8_t *wgt, uint32_t B)
{
    wo_index *X = calloc(1, sizeof(wo_index));
    X->n = n; X->B = B;
    X->len = calloc(n ? n : 1, 4); X->cap = calloc(n ? n : 1, 4);
    X->hub = calloc(n ? n : 1, sizeof(uint32_t *)); X->dist = calloc(n ? n : 1, sizeof(wl *));
    const uint64_t nB = (uint64_t)n * B;
    wl *rlb = malloc(sizeof(wl) * nB);      /* roots_labels_buffer[B][n] */
    wl *DT = malloc(sizeof(wl) * nB);       /* dists_table[n][B] */
    wl *CD = malloc(sizeof(wl) * nB);       /* candidates_dists[n][B] */
    wl *VD = malloc(sizeof(wl) * nB);       /* vertices_dists[n][B] */
    for (uint64_t i = 0; i < nB; ++i) { rlb[i] = WO_INF; DT[i] = WO_INF; CD[i] = WO_INF; VD[i] = WO_INF; }
    uint32_t *cq = malloc(4 * nB), *ncq = calloc(n ? n : 1, 4);     /* candidates_que */
    uint32_t *vq = malloc(4 * nB), *nvq = calloc(n ? n : 1, 4);     /* vertices_que */
    uint32_t *rq = malloc(4 * nB), *nrq = calloc(n ? n : 1, 4);     /* reached_roots_que */
    uint32_t *lq = malloc(4 * nB), *nlq = calloc(n ? n : 1, 4);     /* last_new_roots */
    uint32_t *active = malloc(4 * (n ? n : 1)), *hcq = malloc(4 * (n ? n : 1));
    uint32_t *hnl = malloc(4 * (n ? n : 1)), *ohc = malloc(4 * (n ? n : 1));
    uint8_t *is_active = calloc(n ? n : 1, 1), *has_cand = calloc(n ? n : 1, 1);
    uint8_t *has_new = calloc(n ? n : 1, 1), *once = calloc(n ? n : 1, 1);

    const uint32_t nb = n / B, rem = n % B;
    for (uint32_t b = 0; b < nb + (rem ? 1 : 0); ++b) {
        const uint32_t rs = b * B, rsz = (b < nb) ? B : rem;
        uint32_t n_act = 0, n_hcq = 0, n_hnl = 0, n_ohc = 0;
        /* initialize_tables, WPARA:305-383 */
        for (uint32_t ro = 0; ro < rsz; ++ro) {
            uint32_t r = rs + ro;
            for (uint32_t l = 0; l < X->len[r]; ++l) rlb[(uint64_t)ro * n + X->hub[r][l]] = X->dist[r][l];
            rlb[(uint64_t)ro * n + r] = 0;
        }
        for (uint32_t ro = 0; ro < rsz; ++ro) {
            uint32_t r = rs + ro;
            vq[(uint64_t)r * B + nvq[r]++] = ro;
            VD[(uint64_t)r * B + ro] = 0;
            lq[(uint64_t)r * B + nlq[r]++] = ro;
        }
        for (uint32_t ro = 0; ro < rsz; ++ro) active[n_act++] = rs + ro;
        for (uint32_t ro = 0; ro < rsz; ++ro) { hnl[n_hnl++] = rs + ro; has_new[rs + ro] = 1; }

        while (n_act) {
            /* send_messages, WPARA:435-522, over the active queue in order */
            n_hcq = 0;
            for (uint32_t ai = 0; ai < n_act; ++ai) {
                uint32_t vh = active[ai];
                is_active[vh] = 0;
                for (uint64_t e = off[vh]; e < off[vh + 1]; ++e) {
                    uint32_t vt = adj[e];
                    if (vt <= rs) break;
                    wl w = (wl)wgt[e];
                    int got = 0;
                    for (uint32_t ri = 0; ri < nlq[vh]; ++ri) {
                        uint32_t ro = lq[(uint64_t)vh * B + ri];
                        if (vt <= ro + rs) continue;
                        wl tmp = (wl)(VD[(uint64_t)vh * B + ro] + w);
                        wl *dt = &DT[(uint64_t)vt * B + ro], *cd = &CD[(uint64_t)vt * B + ro];
                        if (tmp < *dt && tmp < *cd) {
                            if (*cd == WO_INF) cq[(uint64_t)vt * B + ncq[vt]++] = ro;
                            if (tmp < *cd) *cd = tmp;
                            if (*dt == WO_INF) rq[(uint64_t)vt * B + nrq[vt]++] = ro;
                            if (tmp < *dt) *dt = tmp;
                            got = 1;
                        }
                    }
                    if (got) {
                        if (!has_cand[vt]) { has_cand[vt] = 1; hcq[n_hcq++] = vt; }
                        if (!once[vt]) { once[vt] = 1; ohc[n_ohc++] = vt; }
                    }
                }
                nlq[vh] = 0;
            }
            n_act = 0;
            /* add, WPARA:1371-1517 */
            for (uint32_t ci = 0; ci < n_hcq; ++ci) {
                uint32_t v = hcq[ci];
                int need = 0;
                for (uint32_t k = 0; k < ncq[v]; ++k) {
                    uint32_t co = cq[(uint64_t)v * B + k];
                    wl tmp = CD[(uint64_t)v * B + co];
                    /* distance_query, WPARA:527-914 */
                    int q = WO_INF;
                    uint32_t creal = co + rs;
                    for (uint32_t l = 0; l < X->len[v] && q == WO_INF; ++l) {
                        uint32_t r = X->hub[v][l];
                        wl rd = rlb[(uint64_t)co * n + r];
                        if (creal <= r || rd == WO_INF) continue;
                        wl s = (wl)(X->dist[v][l] + rd);  /* stored as weightiLarge, WPARA:600 */
                        if (s <= tmp) q = s;
                    }
                    if (q == WO_INF) {
                        /* the reference indexes short_index with the root offset here */
                        const uint32_t c_si = co;
                        for (uint32_t k2 = 0; k2 < nvq[v]; ++k2) {
                            uint32_t ro = vq[(uint64_t)v * B + k2];
                            if (creal <= ro + rs) continue;
                            wl cdv = VD[(uint64_t)c_si * B + ro];
                            if (cdv != WO_INF) {
                                wl s = (wl)(VD[(uint64_t)v * B + ro] + cdv);
                                if (s <= tmp) { q = s; break; }
                            }
                        }
                    }
                    if (q == WO_INF) {
                        if (VD[(uint64_t)v * B + co] == WO_INF) vq[(uint64_t)v * B + nvq[v]++] = co;
                        VD[(uint64_t)v * B + co] = tmp;
                        lq[(uint64_t)v * B + nlq[v]++] = co;
                        need = 1;
                    } else if (q < tmp) {
                        DT[(uint64_t)v * B + co] = (wl)q;
                    }
                }
                if (need) {
                    if (!is_active[v]) { is_active[v] = 1; active[n_act++] = v; }
                    if (!has_new[v]) { has_new[v] = 1; hnl[n_hnl++] = v; }
                }
            }
            for (uint32_t ci = 0; ci < n_hcq; ++ci) {
                uint32_t v = hcq[ci];
                has_cand[v] = 0;
                for (uint32_t k = 0; k < ncq[v]; ++k) CD[(uint64_t)v * B + cq[(uint64_t)v * B + k]] = WO_INF;
                ncq[v] = 0;
            }
        }
        /* filter_out_labels, WPARA:1100-1228 */
        for (uint32_t i = 0; i < n_hnl; ++i) {
            uint32_t v = hnl[i];
            for (uint32_t ic = 0; ic < nvq[v]; ++ic) {
                uint32_t co = vq[(uint64_t)v * B + ic];
                wl dvc = VD[(uint64_t)v * B + co];
                uint32_t c = co + rs;
                for (uint32_t ir = 0; ir < nvq[v]; ++ir) {
                    uint32_t ro = vq[(uint64_t)v * B + ir];
                    if (co <= ro) continue;
                    wl dvr = VD[(uint64_t)v * B + ro];
                    if (dvr > dvc) continue;
                    wl dcr = VD[(uint64_t)c * B + ro];
                    if (dcr > dvc) continue;
                    if ((int)dvr + (int)dcr <= dvc) VD[(uint64_t)v * B + co] = WO_INF;
                }
            }
        }
        /* reset_tables, WPARA:918-956 */
        for (uint32_t ro = 0; ro < rsz; ++ro) {
            uint32_t r = rs + ro;
            for (uint32_t l = 0; l < X->len[r]; ++l) rlb[(uint64_t)ro * n + X->hub[r][l]] = WO_INF;
            rlb[(uint64_t)ro * n + r] = WO_INF;
        }
        for (uint32_t i = 0; i < n_ohc; ++i) {
            uint32_t v = ohc[i];
            once[v] = 0;
            for (uint32_t k = 0; k < nrq[v]; ++k) DT[(uint64_t)v * B + rq[(uint64_t)v * B + k]] = WO_INF;
            nrq[v] = 0;
        }
        /* update_index, WPARA:961-990 */
        for (uint32_t i = 0; i < n_hnl; ++i) {
            uint32_t v = hnl[i];
            has_new[v] = 0;
            for (uint32_t k = 0; k < nvq[v]; ++k) {
                uint32_t ro = vq[(uint64_t)v * B + k];
                wl d = VD[(uint64_t)v * B + ro];
                if (d == WO_INF) continue;
                VD[(uint64_t)v * B + ro] = WO_INF;
                wo_push(X, v, ro + rs, d);
            }
            nvq[v] = 0;
        }
    }
    free(rlb); free(DT); free(CD); free(VD); free(cq); free(ncq); free(vq); free(nvq);
    free(rq); free(nrq); free(lq); free(nlq); free(active); free(hcq); free(hnl); free(ohc);
    free(is_active); free(has_cand); free(has_new); free(once);
    return X;
}

uint64_t wo_num_labels(const wo_index *X)
{
    uint64_t s = 0;
    for (uint32_t v = 0; v < X->n; ++v) s += X->len[v];
    return s;
}

/* order_labels (WPARA:2058-2089): per rank-space vertex, (hub, dist) sorted. */
typedef struct { uint32_t h; int16_t d; } wo_hd;
static int wo_hd_cmp(const void *a, const void *b)
{
    const wo_hd *x = a, *y = b;
    if (x->h != y->h) return x->h < y->h ? -1 : 1;
    return (int)x->d - (int)y->d;
}

void wo_export(const wo_index *X, uint64_t *off, uint32_t *hub, int16_t *dist)
{
    off[0] = 0;
    wo_hd *t = NULL;
    uint32_t cap = 0;
    for (uint32_t v = 0; v < X->n; ++v) {
        uint32_t L = X->len[v];
        if (L > cap) { cap = L; t = realloc(t, sizeof(wo_hd) * cap); }
        for (uint32_t l = 0; l < L; ++l) { t[l].h = X->hub[v][l]; t[l].d = X->dist[v][l]; }
        qsort(t, L, sizeof(wo_hd), wo_hd_cmp);
        for (uint32_t l = 0; l < L; ++l) { hub[off[v] + l] = t[l].h; dist[off[v] + l] = t[l].d; }
        off[v + 1] = off[v] + L;
    }
    free(t);
}

void wo_free(wo_index *X)
{
    if (!X) return;
    for (uint32_t v = 0; v < X->n; ++v) { free(X->hub[v]); free(X->dist[v]); }
    free(X->hub); free(X->dist); free(X->len); free(X->cap); free(X);
}

/* query_distance (WPARA:2125-2243) on rank-space ids: min over common hubs,
 * d >= 32765 -> 32767, ids out of range -> (a == b ? 0 : 32767). */
void wo_query_batch(uint32_t n, const uint64_t *off, const uint32_t *hub, const int16_t *dist, const uint32_t *qa,
                    const uint32_t *qb, uint64_t k, int16_t *out)
{
    for (uint64_t i = 0; i < k; ++i) {
        uint32_t a = qa[i], b = qb[i];
        if (a >= n || b >= n) { out[i] = (a == b) ? 0 : WO_INF; continue; }
        uint32_t d = WO_INF;
        uint64_t i1 = off[a], e1 = off[a + 1], i2 = off[b], e2 = off[b + 1];
        while (i1 < e1 && i2 < e2) {
            if (hub[i1] == hub[i2]) {
                uint32_t td = (uint32_t)(dist[i1] + dist[i2]);
                if (td < d) d = td;
                ++i1; ++i2;
            } else if (hub[i1] < hub[i2]) ++i1; else ++i2;
        }
        if (d >= WO_INF - 2) d = WO_INF;
        out[i] = (int16_t)d;
    }
}

/* Dijkstra checker (textbook, binary heap), u32 distances, UINT32_MAX =
 * unreachable. */
void wo_dijkstra(uint32_t n, const uint64_t *off, const uint32_t *adj, const uint8_t *w, uint32_t src,
                 uint32_t *dist)
{
    typedef struct { uint32_t d, v; } it;
    it *h = malloc(sizeof(it) * (off[n] + 2));
    uint64_t hn = 0;
    for (uint32_t v = 0; v < n; ++v) dist[v] = UINT32_MAX;
    dist[src] = 0;
    h[hn++] = (it){0, src};
    while (hn) {
        it top = h[0];
        h[0] = h[--hn];
        for (uint64_t i = 0;;) {
            uint64_t l = 2 * i + 1, r = l + 1, s = i;
            if (l < hn && h[l].d < h[s].d) s = l;
            if (r < hn && h[r].d < h[s].d) s = r;
            if (s == i) break;
            it t = h[i]; h[i] = h[s]; h[s] = t; i = s;
        }
        if (top.d != dist[top.v]) continue;
        for (uint64_t e = off[top.v]; e < off[top.v + 1]; ++e) {
            uint32_t u = adj[e], nd = top.d + w[e];
            if (nd < dist[u]) {
                dist[u] = nd;
                uint64_t i = hn++;
                h[i] = (it){nd, u};
                while (i && h[(i - 1) / 2].d > h[i].d) {
                    it t = h[i]; h[i] = h[(i - 1) / 2]; h[(i - 1) / 2] = t; i = (i - 1) / 2;
                }
            }
        }
    }
    free(h);
}

/* The canonical label set of ONE vertex v from one Dijkstra, for checks at
 * sizes where all-pairs Dijkstra is out of reach.  (h, d(v, h)) is a label of
 * v iff h ranks at or above v (h <= v in rank space), h is reachable, and no
 * vertex ranked above h lies on any shortest v-h path -- the set
 * WeightedParaVertexCentricPLL builds (WPARA:1100-1228 filters exactly the
 * pairs with such a vertex).  With worst[x] = the smallest id over all
 * shortest v-x paths (endpoints included), h qualifies iff worst[h] == h.
 * Edge weights are positive, so every shortest-path predecessor of x settles
 * before x and one pass in settle order computes worst[].  Writes dist[n] and
 * up to cap labels (hub ascending); returns the label count (which may
 * exceed cap). */
uint64_t wo_canonical_of(uint32_t n, const uint64_t *off, const uint32_t *adj, const uint8_t *w, uint32_t v,
                         uint32_t *dist, uint32_t *hub, uint32_t *hd, uint64_t cap)
{
    typedef struct { uint32_t d, v; } it;
    it *h = malloc(sizeof(it) * (off[n] + 2));
    uint32_t *order = malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t *worst = malloc(sizeof(uint32_t) * (n ? n : 1));
    uint64_t hn = 0, no = 0;
    for (uint32_t x = 0; x < n; ++x) dist[x] = UINT32_MAX;
    dist[v] = 0;
    h[hn++] = (it){0, v};
    while (hn) {
        it top = h[0];
        h[0] = h[--hn];
        for (uint64_t i = 0;;) {
            uint64_t l = 2 * i + 1, r = l + 1, s = i;
            if (l < hn && h[l].d < h[s].d) s = l;
            if (r < hn && h[r].d < h[s].d) s = r;
            if (s == i) break;
            it t = h[i]; h[i] = h[s]; h[s] = t; i = s;
        }
        if (top.d != dist[top.v]) continue;
        order[no++] = top.v;
        for (uint64_t e = off[top.v]; e < off[top.v + 1]; ++e) {
            uint32_t u = adj[e], nd = top.d + w[e];
            if (nd < dist[u]) {
                dist[u] = nd;
                uint64_t i = hn++;
                h[i] = (it){nd, u};
                while (i && h[(i - 1) / 2].d > h[i].d) {
                    it t = h[i]; h[i] = h[(i - 1) / 2]; h[(i - 1) / 2] = t; i = (i - 1) / 2;
                }
            }
        }
    }
    for (uint64_t k = 0; k < no; ++k) {
        const uint32_t x = order[k];
        uint32_t m = x;
        if (x != v)
            for (uint64_t e = off[x]; e < off[x + 1]; ++e) {
                const uint32_t p = adj[e];
                if (dist[p] != UINT32_MAX && (uint64_t)dist[p] + w[e] == dist[x] && worst[p] < m) m = worst[p];
            }
        worst[x] = m;
    }
    uint64_t cnt = 0;
    for (uint32_t x = 0; x <= v && x < n; ++x) {
        if (dist[x] == UINT32_MAX || worst[x] != x) continue;
        if (cnt < cap) {
            hub[cnt] = x;
            hd[cnt] = dist[x];
        }
        ++cnt;
    }
    free(h);
    free(order);
    free(worst);
    return cnt;
}
