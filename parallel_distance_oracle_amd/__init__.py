"""MI355X-native PADO: pruned landmark labeling + exact distance queries on one GPU.

Python mirror of the reference's C++ API (johnpzh/parallel_distance_oracle):

    G = Graph.from_edges(edges)          # Graph(vector<pair>)      includes/graph.h:120-153
    rank = G.make_rank()                 # Graph::make_rank         includes/graph.h:156-174
    rank2id = G.id_transfer(rank)        # Graph::id_transfer       includes/graph.h:176-206
    pll = ParaVertexCentricPLL(G)        # construct                pado_unw_unv_para*.h:309-313
    pll.order_labels(rank2id, rank)      # order_labels             pado_unw_unv_para*.h:2198-2281
    d = pll.query_distance(a, b)         # query_distance           pado_unw_unv_para*.h:2283-2416

Everything below the ctypes layer runs in libpado_gpu.so (HIP kernels for
gfx950 behind a C ABI, include/pado_gpu.h).  There is no CPU fallback: if the
library is missing or no GPU is visible, the device calls raise.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# $PADO_GPU_LIB selects another build of the library (e.g. a variant being
# timed); by default the in-tree build is loaded.
LIB_PATH = os.environ.get("PADO_GPU_LIB") or os.path.join(HERE, "libpado_gpu.so")

BP_SIZE = 50
BATCH_SIZE = 1024
INF8 = 255

QUERY_DEVICE_PTRS = 1
QUERY_REF_COMPAT = 2
QUERY_RANK_SPACE = 4
BUILD_BP_PER_SLOT = 1
BUILD_STREAM = 2  # per-kernel launches with HIP-event timing (default: hipGraph batch step)
BUILD_WIDE = 4  # wide hop domain: 12-bit label distances, u16 answers (graphs deeper than 254)
BUILD_DEVICE_LOOP = 8  # whole batch phase in one workgroup (deep graphs: many small iterations)
WIDE_INF = 0xFFFF  # wide-domain query: unreachable
SERVE_GRID_STRIDE = 0x10
SERVE_PERSISTENT = 0x20
SERVE_RESIDENT = 0x40
WINF = 0xFFFFFFFF  # weighted query: no common hub (unreachable)

_STATUS = {0: "OK", 1: "ERR_ARG", 2: "ERR_HIP", 3: "ERR_NOMEM", 4: "ERR_DOMAIN", 5: "ERR_STATE",
           6: "ERR_IO", 7: "ERR_BUSY", 8: "ERR_TIMEOUT"}


class PadoError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{_STATUS.get(status, status)}: {msg}")
        self.status = status


class BuildCfg(ctypes.Structure):
    _fields_ = [("batch_size", ctypes.c_uint32), ("flags", ctypes.c_uint32), ("push_blocks", ctypes.c_uint32),
                ("add_blocks", ctypes.c_uint32), ("iters_per_sync", ctypes.c_uint32),
                ("sort_max_items", ctypes.c_uint32), ("push_chunk", ctypes.c_uint32),
                ("push_waves", ctypes.c_uint32), ("pull_ratio", ctypes.c_uint32)]

    def as_dict(self):
        return {"batch_size": self.batch_size, "flags": self.flags, "push_blocks": self.push_blocks,
                "add_blocks": self.add_blocks, "iters_per_sync": self.iters_per_sync,
                "push_chunk": self.push_chunk, "push_waves": self.push_waves, "pull_ratio": self.pull_ratio}


class ServeCfg(ctypes.Structure):
    _fields_ = [("staging_pairs", ctypes.c_uint64), ("max_launch_pairs", ctypes.c_uint64),
                ("max_requests", ctypes.c_uint32), ("streams", ctypes.c_uint32), ("flags", ctypes.c_uint32),
                ("min_launch_pairs", ctypes.c_uint32), ("max_wait_us", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32 * 3)]


class ServeStats(ctypes.Structure):
    _fields_ = [("requests_done", ctypes.c_uint64), ("pairs_done", ctypes.c_uint64), ("launches", ctypes.c_uint64),
                ("rejected", ctypes.c_uint64), ("staging_pairs", ctypes.c_uint64), ("busy_ms", ctypes.c_double)]


class IndexInfo(ctypes.Structure):
    _fields_ = [
        ("num_v", ctypes.c_uint32), ("weighted", ctypes.c_uint32), ("num_labels", ctypes.c_uint64),
        ("ordered", ctypes.c_uint32), ("bp_max_level", ctypes.c_uint32), ("iterations", ctypes.c_uint64),
        ("ms_bp", ctypes.c_double), ("ms_batch", ctypes.c_double), ("bp_roots", ctypes.c_uint32 * 50),
        ("pushes", ctypes.c_uint64), ("ind_pass", ctypes.c_uint64), ("bp_pruned", ctypes.c_uint64),
        ("cands", ctypes.c_uint64), ("dq_pruned", ctypes.c_uint64), ("dq_rows", ctypes.c_uint64),
        ("actives", ctypes.c_uint64), ("edges", ctypes.c_uint64), ("tails", ctypes.c_uint64),
        ("dq_labels", ctypes.c_uint64), ("labels_appended", ctypes.c_uint64),
        ("ms_push", ctypes.c_double), ("ms_add", ctypes.c_double), ("ms_bp_levels", ctypes.c_double),
        ("n_push", ctypes.c_uint64), ("n_add", ctypes.c_uint64), ("n_bp_levels", ctypes.c_uint64),
        ("dist_bits", ctypes.c_uint32), ("pull_iters", ctypes.c_uint32),
    ]


_lib = None
_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64


def lib():
    """Load libpado_gpu.so (built by build.build()); raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} not built: run parallel_distance_oracle_amd.build.build()")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "pado_gpu_last_error": (ctypes.c_char_p, []),
        "pado_graph_num_v": (_u32, [_vp, _u64]),
        "pado_graph_csr": (None, [_vp, _u64, _u32, _vp, _vp]),
        "pado_make_rank": (ctypes.c_int, [_vp, _u32, ctypes.c_int, _vp]),
        "pado_id_transfer": (None, [_u32, _vp, _vp, _vp, _vp, _vp, _vp]),
        "pado_wgraph_csr": (None, [_vp, _vp, _u64, _u32, _vp, _vp, _vp]),
        "pado_wid_transfer": (None, [_u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
        "pado_sort_plan": (_u32, [_vp, _u32, _u64, _vp, _u32]),
        "pado_read_edge_text": (ctypes.c_int, [ctypes.c_char_p, _u32, _vp, ctypes.POINTER(_u64)]),
        "pado_read_edge_binary": (ctypes.c_int, [ctypes.c_char_p, _vp, ctypes.POINTER(_u64),
                                                 ctypes.POINTER(_u32)]),
        "pado_write_edge_binary": (ctypes.c_int, [ctypes.c_char_p, _vp, _u64]),
        "pado_free": (None, [_vp]),
        "pado_gpu_windex_export": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
        "pado_gpu_wquery": (ctypes.c_int, [_vp, _vp, _vp, _vp, _u64, _vp, _u32, _vp]),
        "pado_rmat_edges": (_u64, [_u32, _u32, ctypes.c_double, ctypes.c_double, ctypes.c_double, _u64, _vp, _u64]),
        "pado_grid_edges": (_u64, [_u32, ctypes.c_double, _u64, _vp, _u64]),
        "pado_edge_weights": (None, [_u64, _u32, _u32, _u64, _vp]),
        "pado_uniform_pairs": (None, [_u32, _u64, _u64, _vp, _vp]),
        "pado_zipf_vertices": (None, [_u32, ctypes.c_double, _u64, _u64, _vp, _vp]),
        "pado_gpu_device_count": (ctypes.c_int, []),
        "pado_gpu_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
        "pado_gpu_ctx_destroy": (None, [_vp]),
        "pado_gpu_ctx_stream": (_vp, [_vp]),
        "pado_gpu_ctx_sync": (ctypes.c_int, [_vp]),
        "pado_gpu_copy_bytes": (ctypes.c_int, [_vp, _vp, _vp, _u64]),
        "pado_gpu_graph_upload": (ctypes.c_int, [_vp, _u32, _u64, _vp, _vp, _vp, ctypes.POINTER(_vp)]),
        "pado_gpu_graph_free": (None, [_vp]),
        "pado_gpu_graph_from_edges": (ctypes.c_int, [_vp, _vp, _u64, _vp, _vp, ctypes.c_int, ctypes.POINTER(_vp),
                                                     ctypes.POINTER(_u32), _vp, _vp]),
        "pado_gpu_graph_export": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
        "pado_gpu_graph_get_info": (ctypes.c_int, [_vp, ctypes.POINTER(_u32), ctypes.POINTER(_u64),
                                                   ctypes.POINTER(ctypes.c_int)]),
        "pado_gpu_build": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(BuildCfg), ctypes.POINTER(_vp)]),
        "pado_gpu_index_get_info": (ctypes.c_int, [_vp, ctypes.POINTER(IndexInfo)]),
        "pado_gpu_index_export": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
        "pado_gpu_index_import": (ctypes.c_int, [_vp, _u32, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(_vp)]),
        "pado_gpu_index_order": (ctypes.c_int, [_vp, _vp, _vp, _u32]),
        "pado_gpu_query": (ctypes.c_int, [_vp, _vp, _vp, _vp, _u64, _vp, _u32, _vp]),
        "pado_gpu_query_noalloc": (ctypes.c_int, [_vp, _vp, _vp, _vp, _u64, _vp, _u32]),
        "pado_gpu_query_wide": (ctypes.c_int, [_vp, _vp, _vp, _vp, _u64, _vp, _u32, _vp]),
        "pado_gpu_query_pair": (ctypes.c_int, [_vp, _vp, _u32, _u32, _u32, ctypes.POINTER(_u32)]),
        "pado_gpu_index_export_wide": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
        "pado_gpu_index_free": (None, [_vp]),
        "pado_gpu_index_store": (ctypes.c_int, [_vp, _vp, ctypes.c_char_p, _vp, _u32]),
        "pado_gpu_index_load": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.POINTER(_vp)]),
        "pado_gpu_autotune": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(BuildCfg), _u32, ctypes.POINTER(BuildCfg),
                                             ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_u32)]),
        "pado_gpu_build_multi": (ctypes.c_int, [_vp, _u32, _vp, _vp, _vp, ctypes.POINTER(BuildCfg),
                                                ctypes.POINTER(_vp)]),
        "pado_gpu_autotune_multi": (ctypes.c_int, [_vp, _u32, _vp, _vp, _vp, _u32, ctypes.POINTER(BuildCfg), _u32,
                                                   ctypes.POINTER(BuildCfg), ctypes.POINTER(ctypes.c_double),
                                                   ctypes.POINTER(_u32)]),
        "pado_gpu_multi_graph_info": (ctypes.c_int, [_vp, _u32, ctypes.POINTER(ctypes.c_int),
                                                     ctypes.POINTER(_u64), ctypes.POINTER(ctypes.c_int),
                                                     ctypes.POINTER(_u32)]),
        "pado_gpu_multi_export": (ctypes.c_int, [_vp, _vp, _u32, _vp, _vp, _vp, _vp, _vp]),
        "pado_gpu_multi_export_wide": (ctypes.c_int, [_vp, _vp, _u32, _vp, _vp, _vp, _vp, _vp]),
        "pado_gpu_multi_query": (ctypes.c_int, [_vp, _vp, _u32, _vp, _vp, _u64, _vp]),
        "pado_gpu_multi_info": (ctypes.c_int, [_vp, ctypes.POINTER(_u32), ctypes.POINTER(_u64),
                                               ctypes.POINTER(ctypes.c_double)]),
        "pado_gpu_multi_free": (None, [_vp]),
        "pado_gpu_serve_start": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(ServeCfg), ctypes.POINTER(_vp)]),
        "pado_gpu_serve_submit": (ctypes.c_int, [_vp, _vp, _vp, _u64, ctypes.POINTER(_u64)]),
        "pado_gpu_serve_poll": (ctypes.c_int, [_vp, _u64, ctypes.POINTER(ctypes.c_int)]),
        "pado_gpu_serve_wait": (ctypes.c_int, [_vp, _u64, _vp, ctypes.POINTER(ctypes.c_double)]),
        "pado_gpu_serve_wait_for": (ctypes.c_int, [_vp, _u64, _vp, ctypes.POINTER(ctypes.c_double), _u64]),
        "pado_gpu_serve_cancel": (ctypes.c_int, [_vp, _u64]),
        "pado_gpu_serve_get_stats": (ctypes.c_int, [_vp, ctypes.POINTER(ServeStats)]),
        "pado_gpu_serve_replay": (ctypes.c_int, [_vp, _vp, _vp, _vp, _u32, _vp, _vp, _vp,
                                                 ctypes.POINTER(ctypes.c_double)]),
        "pado_gpu_serve_stop": (None, [_vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def exported_symbols():
    """Names the C ABI declares (include/pado_gpu.h) -- used by tests."""
    import re
    hdr = os.path.join(os.path.dirname(HERE), "include", "pado_gpu.h")
    txt = open(hdr).read()
    return sorted(set(re.findall(r"\b(pado_[a-z0-9_]+)\s*\(", txt)))


def _check(status):
    if status != 0:
        raise PadoError(status, lib().pado_gpu_last_error().decode())


def _p(a):
    return a.ctypes.data_as(_vp) if a is not None and a.size else None


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# ---------------------------------------------------------------------------
# Workload generators (deterministic; seed-pinned splitmix64)
# ---------------------------------------------------------------------------
def rmat_edges(scale, edge_factor, a=0.57, b=0.19, c=0.19, seed=1):
    """Undirected R-MAT edges (u < v), self-loops and duplicates removed."""
    L = lib()
    m = L.pado_rmat_edges(scale, edge_factor, a, b, c, seed, None, 0)
    out = np.zeros(2 * max(m, 1), np.uint32)
    L.pado_rmat_edges(scale, edge_factor, a, b, c, seed, _p(out), m)
    return out[: 2 * m].reshape(-1, 2)


def grid_edges(side, delete_frac=0.05, seed=1):
    L = lib()
    m = L.pado_grid_edges(side, delete_frac, seed, None, 0)
    out = np.zeros(2 * max(m, 1), np.uint32)
    L.pado_grid_edges(side, delete_frac, seed, _p(out), m)
    return out[: 2 * m].reshape(-1, 2)


def edge_weights(m, lo=1, hi=100, seed=1):
    w = np.zeros(max(m, 1), np.uint32)
    lib().pado_edge_weights(m, lo, hi, seed, _p(w))
    return w[:m]


def uniform_pairs(n, k, seed=1):
    a = np.zeros(max(k, 1), np.uint32)
    b = np.zeros(max(k, 1), np.uint32)
    lib().pado_uniform_pairs(n, k, seed, _p(a), _p(b))
    return a[:k], b[:k]


def read_edge_text(path, ncols=2):
    """Edge-list text as the reference's readers take it (pado_read_edge_text):
    an (m, ncols) u32 array."""
    buf = ctypes.POINTER(ctypes.c_uint32)()
    m = _u64()
    st = lib().pado_read_edge_text(str(path).encode(), ncols, ctypes.byref(buf), ctypes.byref(m))
    if st != 0:
        raise OSError(f"cannot read edge list {path} ({st})")
    try:
        out = np.ctypeslib.as_array(buf, shape=(m.value * ncols,)).copy() if m.value else np.zeros(0, np.uint32)
    finally:
        lib().pado_free(buf)
    return out.reshape(-1, ncols)


def read_edge_binary(path):
    """The reference's binary edge list (u32 num_v, u64 num_e, u32 pairs)."""
    buf = ctypes.POINTER(ctypes.c_uint32)()
    m = _u64()
    nv = _u32()
    st = lib().pado_read_edge_binary(str(path).encode(), ctypes.byref(buf), ctypes.byref(m), ctypes.byref(nv))
    if st != 0:
        raise OSError(f"cannot read binary edge list {path} ({st})")
    try:
        out = np.ctypeslib.as_array(buf, shape=(m.value * 2,)).copy() if m.value else np.zeros(0, np.uint32)
    finally:
        lib().pado_free(buf)
    return out.reshape(-1, 2)


def write_edge_binary(path, edges):
    e = _c(edges, np.uint32).reshape(-1, 2)
    if lib().pado_write_edge_binary(str(path).encode(), _p(e), e.shape[0]) != 0:
        raise OSError(f"cannot write {path}")


def mixed_graphs(count, seed=4):
    """The many-graph workload: `count` rank-space graphs with vertex counts
    log-uniform in [2^10, 2^16]; even-numbered ones R-MAT (edge factor 8,
    (0.57, 0.19, 0.19, 0.05)), odd-numbered ones 2D grids with 5% of the edges
    deleted (road-like, large diameter)."""
    rng = np.random.default_rng(seed)
    graphs = []
    for i in range(count):
        nv = 2.0 ** rng.uniform(10, 16)
        if i % 2 == 0:
            e = rmat_edges(max(1, int(round(np.log2(nv)))), 8, 0.57, 0.19, 0.19, seed + i)
        else:
            e = grid_edges(max(2, int(round(np.sqrt(nv)))), 0.05, seed + i)
        G = Graph.from_edges(e)
        G.id_transfer(G.make_rank())
        graphs.append(G)
    return graphs


def sort_plan(off, max_items):
    """Vertex ranges of the final label sort (pado_sort_plan): the cut list,
    or None if one vertex alone exceeds max_items."""
    off = _c(off, np.uint64)
    n = off.size - 1
    cuts = np.zeros(n + 2, np.uint32)
    r = lib().pado_sort_plan(_p(off), n, max_items, _p(cuts), n + 2)
    return None if r == 0 else cuts[: r + 1]


def zipf_vertices(n, k, s=1.1, seed=1, rank2id=None):
    out = np.zeros(max(k, 1), np.uint32)
    r2i = _c(rank2id, np.uint32) if rank2id is not None else None
    lib().pado_zipf_vertices(n, s, k, seed, _p(r2i), _p(out))
    return out[:k]


# ---------------------------------------------------------------------------
# Graph: the reference's Graph class (includes/graph.h:31-92)
# ---------------------------------------------------------------------------
class Graph:
    """Undirected CSR with the reference's semantics.

    num_v = max id + 1, both directions per input edge, no de-duplication
    (includes/graph.h:125-153).  make_rank / id_transfer renumber it into rank
    space with descending adjacency lists (includes/graph.h:156-206)."""

    def __init__(self, num_v, off, adj):
        self.num_v = int(num_v)
        self.off = _c(off, np.uint64)
        self.adj = _c(adj, np.uint32)

    @property
    def num_e(self):
        return int(self.adj.size)

    @property
    def out_degrees(self):
        return np.diff(self.off).astype(np.uint32)

    @classmethod
    def from_edges(cls, edges):
        e = _c(edges, np.uint32).reshape(-1)
        m = e.size // 2
        L = lib()
        n = L.pado_graph_num_v(_p(e), m) if m else 0
        off = np.zeros(n + 1, np.uint64)
        adj = np.zeros(max(2 * m, 1), np.uint32)
        if n:
            L.pado_graph_csr(_p(e), m, n, _p(off), _p(adj))
        return cls(n, off, adj[: 2 * m])

    @classmethod
    def from_file(cls, path):
        """Graph(const char*) (includes/graph.h:95-118): '#'/'%' lines skipped,
        the reference's rules for malformed lines; parsed on all host cores."""
        return cls.from_edges(read_edge_text(path, 2))

    @classmethod
    def from_binary(cls, path):
        """The reference's binary edge list (edgelist2binary.cpp:28-91)."""
        return cls.from_edges(read_edge_binary(path))

    def make_rank(self, use_global_rand=False):
        rank = np.zeros(max(self.num_v, 1), np.uint32)
        _check(0 if lib().pado_make_rank(_p(self.off), self.num_v, 1 if use_global_rand else 0, _p(rank)) == 0
               else 1)
        return rank[: self.num_v]

    def id_transfer(self, rank):
        """Renumber in place into rank space; returns rank2id."""
        rank = _c(rank, np.uint32)
        n = self.num_v
        off2 = np.zeros(n + 1, np.uint64)
        adj2 = np.zeros(max(self.adj.size, 1), np.uint32)
        r2i = np.zeros(max(n, 1), np.uint32)
        lib().pado_id_transfer(n, _p(self.off), _p(self.adj) if self.adj.size else _p(adj2), _p(rank),
                               _p(off2), _p(adj2), _p(r2i))
        self.off, self.adj = off2, adj2[: self.adj.size]
        return r2i[:n]


class WeightedGraph(Graph):
    """The reference's WeightedGraph (includes/graph.h:223-408): Graph's CSR
    plus one u8 weight per directed edge (weights cast to weighti, as its file
    reader does).  Its make_rank is Graph's degree order."""

    def __init__(self, num_v, off, adj, w):
        super().__init__(num_v, off, adj)
        self.w = _c(w, np.uint8)

    @classmethod
    def from_edges(cls, edges, weights):
        e = _c(edges, np.uint32).reshape(-1)
        m = e.size // 2
        wt = _c(weights, np.uint32).reshape(-1)
        if wt.size != m:
            raise ValueError(f"{wt.size} weights for {m} edges")
        L = lib()
        n = L.pado_graph_num_v(_p(e), m) if m else 0
        off = np.zeros(n + 1, np.uint64)
        adj = np.zeros(max(2 * m, 1), np.uint32)
        w = np.zeros(max(2 * m, 1), np.uint8)
        if n:
            L.pado_wgraph_csr(_p(e), _p(wt), m, n, _p(off), _p(adj), _p(w))
        return cls(n, off, adj[: 2 * m], w[: 2 * m])

    @classmethod
    def from_file(cls, path):
        """WeightedGraph(const char*) (includes/graph.h:277-304): 'a b w' rows,
        '#'/'%' lines skipped, weights cast to u8 as the reference does."""
        a = read_edge_text(path, 3)
        return cls.from_edges(a[:, :2].copy(), (a[:, 2] & 0xFF).astype(np.uint32))

    def id_transfer(self, rank):
        """Renumber in place into rank space (includes/graph.h:374-408);
        returns rank2id."""
        rank = _c(rank, np.uint32)
        n = self.num_v
        off2 = np.zeros(n + 1, np.uint64)
        adj2 = np.zeros(max(self.adj.size, 1), np.uint32)
        w2 = np.zeros(max(self.adj.size, 1), np.uint8)
        r2i = np.zeros(max(n, 1), np.uint32)
        src_adj = self.adj if self.adj.size else adj2
        src_w = self.w if self.w.size else w2
        lib().pado_wid_transfer(n, _p(self.off), _p(src_adj), _p(src_w), _p(rank), _p(off2), _p(adj2), _p(w2),
                                _p(r2i))
        self.off, self.adj, self.w = off2, adj2[: self.adj.size], w2[: self.adj.size]
        return r2i[:n]


# ---------------------------------------------------------------------------
# Device objects
# ---------------------------------------------------------------------------
def device_count():
    return lib().pado_gpu_device_count()


class Context:
    def __init__(self, device=0):
        h = _vp()
        _check(lib().pado_gpu_ctx_create(device, ctypes.byref(h)))
        self.h = h
        self.device = device

    @property
    def stream(self):
        return lib().pado_gpu_ctx_stream(self.h)

    def sync(self):
        _check(lib().pado_gpu_ctx_sync(self.h))

    def close(self):
        if self.h:
            lib().pado_gpu_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def copy_bytes(ctx, dst_ptr, src_ptr, nbytes):
    """Plain device copy kernel on the context's stream (asynchronous)."""
    _check(lib().pado_gpu_copy_bytes(ctx.h, _vp(dst_ptr), _vp(src_ptr), nbytes))


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class DeviceGraph:
    def __init__(self, ctx, G, weights=None):
        self.ctx = ctx
        h = _vp()
        if weights is None and isinstance(G, WeightedGraph):
            weights = G.w
        w = _c(weights, np.uint8) if weights is not None else None
        if w is not None and w.size == 0:
            w = np.ones(1, np.uint8)  # a non-null pointer marks the graph weighted
        _check(lib().pado_gpu_graph_upload(ctx.h, G.num_v, G.num_e, _p(G.off), _p(G.adj) if G.adj.size else None,
                                           _p(w), ctypes.byref(h)))
        self.h = h
        self.num_v = G.num_v

    @classmethod
    def from_edges(cls, ctx, edges, weights=None, rank=None):
        """Device ingest: the reference's Graph / WeightedGraph construct,
        make_rank and id_transfer (includes/graph.h:120-206, 313-408) run
        on the GPU.  Returns (device graph, rank, rank2id)."""
        e = _c(edges, np.uint32).reshape(-1)
        m = e.size // 2
        n = lib().pado_graph_num_v(_p(e), m) if m else 0
        w = _c(weights, np.uint32).reshape(-1) if weights is not None else None
        if w is not None and w.size != m:
            raise ValueError(f"{w.size} weights for {m} edges")
        r_in = _c(rank, np.uint32) if rank is not None else None
        rank_out = np.zeros(max(n, 1), np.uint32)
        r2i = np.zeros(max(n, 1), np.uint32)
        h = _vp()
        nv = _u32()
        _check(lib().pado_gpu_graph_from_edges(ctx.h, _p(e), m, _p(w), _p(r_in), 0, ctypes.byref(h),
                                               ctypes.byref(nv), _p(rank_out), _p(r2i)))
        self = cls.__new__(cls)
        self.ctx = ctx
        self.h = h
        self.num_v = nv.value
        return self, rank_out[:n], r2i[:n]

    def info(self):
        n, m, w = _u32(), _u64(), ctypes.c_int()
        _check(lib().pado_gpu_graph_get_info(self.h, ctypes.byref(n), ctypes.byref(m), ctypes.byref(w)))
        return n.value, m.value, bool(w.value)

    def export(self):
        """(off[n+1], adj[m_dir], w[m_dir] or None) of the device CSR."""
        n, m, weighted = self.info()
        off = np.zeros(n + 1, np.uint64)
        adj = np.zeros(max(m, 1), np.uint32)
        w = np.zeros(max(m, 1), np.uint8) if weighted else None
        _check(lib().pado_gpu_graph_export(self.ctx.h, self.h, _p(off), _p(adj), _p(w)))
        return off, adj[:m], (w[:m] if weighted else None)

    def close(self):
        if self.h:
            lib().pado_gpu_graph_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Index:
    """A device-resident PADO index (rank space; plus an original-id copy
    after order())."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.h = handle
        self.info = self.get_info()
        self.num_v = self.info.num_v

    @classmethod
    def build(cls, ctx, dgraph, batch_size=BATCH_SIZE, bp_per_slot=False, cfg=None, wide=False):
        if cfg is None:
            cfg = BuildCfg()
            cfg.batch_size = batch_size
            cfg.flags = (BUILD_BP_PER_SLOT if bp_per_slot else 0) | (BUILD_WIDE if wide else 0)
        h = _vp()
        _check(lib().pado_gpu_build(ctx.h, dgraph.h, ctypes.byref(cfg), ctypes.byref(h)))
        return cls(ctx, h)

    @classmethod
    def import_(cls, ctx, n, bp_d, bp_s, loff, hub, dist):
        h = _vp()
        bp_d = _c(bp_d, np.uint8)
        bp_s = _c(bp_s, np.uint64)
        loff = _c(loff, np.uint64)
        hub = _c(hub, np.uint32)
        dist = _c(dist, np.uint8)
        _check(lib().pado_gpu_index_import(ctx.h, n, _p(bp_d), _p(bp_s), _p(loff), _p(hub), _p(dist),
                                           ctypes.byref(h)))
        return cls(ctx, h)

    @classmethod
    def load(cls, ctx, filename):
        h = _vp()
        _check(lib().pado_gpu_index_load(ctx.h, filename.encode(), ctypes.byref(h)))
        return cls(ctx, h)

    def get_info(self):
        info = IndexInfo()
        _check(lib().pado_gpu_index_get_info(self.h, ctypes.byref(info)))
        return info

    def stats(self):
        i = self.get_info()
        return {k: getattr(i, k) for k, _ in IndexInfo._fields_ if k != "bp_roots"}

    def export(self):
        """Canonical rank-space arrays: (bp_d[n,50], bp_s[n,50,2], loff[n+1], hub, dist); dist is u8, or
        u16 for a wide-domain index."""
        n = self.num_v
        tot = self.info.num_labels
        bp_d = np.zeros(n * BP_SIZE, np.uint8)
        bp_s = np.zeros(n * BP_SIZE * 2, np.uint64)
        loff = np.zeros(n + 1, np.uint64)
        hub = np.zeros(max(tot, 1), np.uint32)
        if self.wide:
            dist = np.zeros(max(tot, 1), np.uint16)
            fn = lib().pado_gpu_index_export_wide
        else:
            dist = np.zeros(max(tot, 1), np.uint8)
            fn = lib().pado_gpu_index_export
        _check(fn(self.ctx.h, self.h, _p(bp_d), _p(bp_s), _p(loff), _p(hub), _p(dist)))
        return bp_d.reshape(n, BP_SIZE), bp_s.reshape(n, BP_SIZE, 2), loff, hub[:tot], dist[:tot]

    @property
    def weighted(self):
        return bool(self.info.weighted)

    @property
    def wide(self):
        """Labels in the wide hop domain (12-bit distances, u16 answers)."""
        return self.info.dist_bits == 12

    def wexport(self):
        """Weighted index, canonical rank space: (loff[n+1], hub u32, dist u32)."""
        n = self.num_v
        tot = self.info.num_labels
        loff = np.zeros(n + 1, np.uint64)
        hub = np.zeros(max(tot, 1), np.uint32)
        dist = np.zeros(max(tot, 1), np.uint32)
        _check(lib().pado_gpu_windex_export(self.ctx.h, self.h, _p(loff), _p(hub), _p(dist)))
        return loff, hub[:tot], dist[:tot]

    def order(self, rank):
        rank = _c(rank, np.uint32)
        _check(lib().pado_gpu_index_order(self.ctx.h, self.h, _p(rank), rank.size))
        self.info = self.get_info()

    def query(self, a, b, ref_compat=False, rank_space=False):
        """Distances for pairs (a[i], b[i]): u8 for an unweighted index, u16
        (WIDE_INF = unreachable) for a wide-domain one, u32 (WINF =
        unreachable) for a weighted one."""
        a = _c(a, np.uint32).reshape(-1)
        b = _c(b, np.uint32).reshape(-1)
        if a.size != b.size:
            raise ValueError("a and b differ in length")
        flags = (QUERY_REF_COMPAT if ref_compat else 0) | (QUERY_RANK_SPACE if rank_space else 0)
        if self.weighted:
            out = np.zeros(max(a.size, 1), np.uint32)
            _check(lib().pado_gpu_wquery(self.ctx.h, self.h, _p(a), _p(b), a.size, _p(out), flags, None))
        elif self.wide:
            out = np.zeros(max(a.size, 1), np.uint16)
            _check(lib().pado_gpu_query_wide(self.ctx.h, self.h, _p(a), _p(b), a.size, _p(out), flags, None))
        else:
            out = np.zeros(max(a.size, 1), np.uint8)
            _check(lib().pado_gpu_query(self.ctx.h, self.h, _p(a), _p(b), a.size, _p(out), flags, None))
        return out[: a.size]

    def query_pair(self, a, b, ref_compat=False, rank_space=False):
        """One pair through pado_gpu_query_pair (one launch, no copies)."""
        out = _u32()
        flags = (QUERY_REF_COMPAT if ref_compat else 0) | (QUERY_RANK_SPACE if rank_space else 0)
        _check(lib().pado_gpu_query_pair(self.ctx.h, self.h, int(a), int(b), flags, ctypes.byref(out)))
        return out.value

    def query_device(self, a_ptr, b_ptr, k, out_ptr, ref_compat=False, rank_space=False, stream=None):
        """Asynchronous query on device buffers (device pointers of the library's
        HIP runtime, e.g. hip.DeviceArray.ptr);
        out holds u8 (unweighted), u16 (wide domain) or u32 (weighted) distances."""
        flags = QUERY_DEVICE_PTRS | (QUERY_REF_COMPAT if ref_compat else 0) | (QUERY_RANK_SPACE if rank_space else 0)
        fn = (lib().pado_gpu_wquery if self.weighted else
              lib().pado_gpu_query_wide if self.wide else lib().pado_gpu_query)
        _check(fn(self.ctx.h, self.h, _vp(a_ptr), _vp(b_ptr), k, _vp(out_ptr), flags,
                  _vp(stream) if stream else None))

    def store(self, filename, rank):
        rank = _c(rank, np.uint32)
        _check(lib().pado_gpu_index_store(self.ctx.h, self.h, filename.encode(), _p(rank), rank.size))

    def close(self):
        if self.h:
            lib().pado_gpu_index_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(ctx, dgraph, max_trials=0, start=None):
    """Times trial builds (median of three each); returns (best BuildCfg,
    its median ms, configurations timed)."""
    best = BuildCfg()
    ms = ctypes.c_double()
    tr = _u32()
    _check(lib().pado_gpu_autotune(ctx.h, dgraph.h, ctypes.byref(start) if start is not None else None,
                                   max_trials, ctypes.byref(best), ctypes.byref(ms), ctypes.byref(tr)))
    return best, ms.value, tr.value


def _multi_arrays(graphs):
    ns = np.array([G.num_v for G in graphs], np.uint32)
    off = np.concatenate([G.off for G in graphs]).astype(np.uint64)
    adj = np.concatenate([G.adj for G in graphs]).astype(np.uint32) if graphs else np.zeros(0, np.uint32)
    return ns, off, adj


def autotune_multi(ctx, graphs, sample=64, max_trials=0, start=None):
    """The many-graph job's tuner on `sample` of the graphs; returns (best
    BuildCfg, its median ms on the sample, configurations timed)."""
    ns, off, adj = _multi_arrays(graphs)
    best = BuildCfg()
    ms = ctypes.c_double()
    tr = _u32()
    _check(lib().pado_gpu_autotune_multi(ctx.h, len(graphs), _p(ns), _p(off), _p(adj) if adj.size else _p(ns),
                                         sample, ctypes.byref(start) if start is not None else None, max_trials,
                                         ctypes.byref(best), ctypes.byref(ms), ctypes.byref(tr)))
    return best, ms.value, tr.value


class MultiIndex:
    """Many independent rank-space graphs labeled as one job."""

    def __init__(self, ctx, graphs, batch_size=BATCH_SIZE, cfg=None):
        self.ctx = ctx
        ns, off, adj = _multi_arrays(graphs)
        if cfg is None:
            cfg = BuildCfg()
            cfg.batch_size = batch_size
        h = _vp()
        _check(lib().pado_gpu_build_multi(ctx.h, len(graphs), _p(ns), _p(off), _p(adj) if adj.size else None,
                                          ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self.sizes = ns

    def graph_info(self, g):
        st = ctypes.c_int()
        nl = _u64()
        iu = ctypes.c_int()
        db = _u32()
        _check(lib().pado_gpu_multi_graph_info(self.h, g, ctypes.byref(st), ctypes.byref(nl), ctypes.byref(iu),
                                               ctypes.byref(db)))
        return {"status": st.value, "num_labels": nl.value, "in_union": bool(iu.value), "wide": db.value == 12}

    def export(self, g):
        """Graph g's canonical rank-space arrays; dist is u8, or u16 for a
        graph labeled in the wide hop domain."""
        n = int(self.sizes[g])
        gi = self.graph_info(g)
        tot = gi["num_labels"]
        bp_d = np.zeros(n * BP_SIZE, np.uint8)
        bp_s = np.zeros(n * BP_SIZE * 2, np.uint64)
        loff = np.zeros(n + 1, np.uint64)
        hub = np.zeros(max(tot, 1), np.uint32)
        if gi["wide"]:
            dist = np.zeros(max(tot, 1), np.uint16)
            fn = lib().pado_gpu_multi_export_wide
        else:
            dist = np.zeros(max(tot, 1), np.uint8)
            fn = lib().pado_gpu_multi_export
        _check(fn(self.ctx.h, self.h, g, _p(bp_d), _p(bp_s), _p(loff), _p(hub), _p(dist)))
        return bp_d.reshape(n, BP_SIZE), bp_s.reshape(n, BP_SIZE, 2), loff, hub[:tot], dist[:tot]

    def query(self, g, a, b):
        """Distances on graph g between its rank-space vertices a[i], b[i] (u16:
        the reference's u8 answers for a graph in its domain, exact hops with
        WIDE_INF = unreachable for a wide-domain graph)."""
        a = _c(a, np.uint32).reshape(-1)
        b = _c(b, np.uint32).reshape(-1)
        out = np.zeros(max(a.size, 1), np.uint16)
        _check(lib().pado_gpu_multi_query(self.ctx.h, self.h, g, _p(a), _p(b), a.size, _p(out)))
        return out[: a.size]

    def info(self):
        nu = _u32()
        uv = _u64()
        ms = ctypes.c_double()
        _check(lib().pado_gpu_multi_info(self.h, ctypes.byref(nu), ctypes.byref(uv), ctypes.byref(ms)))
        return {"unions": nu.value, "union_vertices": uv.value, "ms_total": ms.value}

    def close(self):
        if self.h:
            lib().pado_gpu_multi_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Busy(PadoError):
    pass


class Timeout(PadoError):
    pass


class Server:
    """Continuous-batching query service over an ordered, resident Index."""

    def __init__(self, ctx, index, staging_pairs=0, max_launch_pairs=0, max_requests=0, streams=0,
                 ref_compat=False, persistent=False, min_launch_pairs=0, max_wait_us=0, resident=False):
        cfg = ServeCfg()
        cfg.min_launch_pairs = min_launch_pairs
        cfg.max_wait_us = max_wait_us
        cfg.staging_pairs = staging_pairs
        cfg.max_launch_pairs = max_launch_pairs
        cfg.max_requests = max_requests
        cfg.streams = streams
        cfg.flags = ((QUERY_REF_COMPAT if ref_compat else 0) | (SERVE_PERSISTENT if persistent else 0)
                     | (SERVE_RESIDENT if resident else 0))
        self.dtype = np.uint32 if index.weighted else (np.uint16 if index.wide else np.uint8)
        h = _vp()
        _check(lib().pado_gpu_serve_start(ctx.h, index.h, ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self.index = index  # keep the index alive while serving

    def submit(self, a, b):
        """Returns a request id; raises Busy when admission refuses."""
        a = _c(a, np.uint32).reshape(-1)
        b = _c(b, np.uint32).reshape(-1)
        rid = _u64()
        st = lib().pado_gpu_serve_submit(self.h, _p(a), _p(b), a.size, ctypes.byref(rid))
        if st == 7:
            raise Busy(st, lib().pado_gpu_last_error().decode())
        _check(st)
        return rid.value

    def poll(self, rid):
        d = ctypes.c_int()
        _check(lib().pado_gpu_serve_poll(self.h, rid, ctypes.byref(d)))
        return bool(d.value)

    def wait(self, rid, k, timeout_us=None):
        """Blocks (at most timeout_us, then raises Timeout); returns
        (answers[k] -- u8 hops, u16 for a wide-domain index, u32 for a
        weighted one --, latency_ms from submit to completion)."""
        out = np.zeros(max(k, 1), self.dtype)
        lat = ctypes.c_double()
        if timeout_us is None:
            st = lib().pado_gpu_serve_wait(self.h, rid, _p(out), ctypes.byref(lat))
        else:
            st = lib().pado_gpu_serve_wait_for(self.h, rid, _p(out), ctypes.byref(lat), int(timeout_us))
        if st == 8:
            raise Timeout(st, lib().pado_gpu_last_error().decode())
        _check(st)
        return out[:k], lat.value

    def cancel(self, rid):
        """Withdraws a request (its slot and ring space are released)."""
        _check(lib().pado_gpu_serve_cancel(self.h, rid))

    def replay(self, a, b, req_off, arrival_s):
        """Native open-loop load generator: request i = pairs req_off[i]:req_off[i+1]
        arriving at arrival_s[i] s.  Returns (latency_ms per request from
        arrival to completion, answers for all pairs, wall seconds)."""
        a = _c(a, np.uint32).reshape(-1)
        b = _c(b, np.uint32).reshape(-1)
        off = _c(req_off, np.uint64).reshape(-1)
        arr = _c(arrival_s, np.float64).reshape(-1)
        nreq = off.size - 1
        # Output buffers are written once before the run, so that first-touch
        # page faults (hundreds of MB of answers) are not taken while requests
        # are being timed.
        lat = np.empty(max(nreq, 1), np.float64)
        lat.fill(0.0)
        ans = np.empty(max(int(off[-1]), 1), self.dtype)
        ans.fill(0)
        wall = ctypes.c_double()
        _check(lib().pado_gpu_serve_replay(self.h, _p(a), _p(b), _p(off), nreq, _p(arr), _p(lat), _p(ans),
                                           ctypes.byref(wall)))
        return lat[:nreq], ans[:int(off[-1])], wall.value

    def stats(self):
        s = ServeStats()
        _check(lib().pado_gpu_serve_get_stats(self.h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in ServeStats._fields_}

    def stop(self):
        if self.h:
            lib().pado_gpu_serve_stop(self.h)
            self.h = None

    def __del__(self):
        try:
            self.stop()
        except Exception:
            pass


class ParaVertexCentricPLL:
    """ParaVertexCentricPLL<BATCH_SIZE> (pado_unw_unv_para*.h:42-304) on the GPU.

    Constructing from a rank-space Graph builds the index (PARA:309-313);
    order_labels / query_distance / store_index_to_file / load_index_from_file
    keep the reference's meaning."""

    def __init__(self, G=None, batch_size=BATCH_SIZE, ctx=None):
        self.ctx = ctx or default_context()
        self.index = None
        self.BATCH_SIZE = batch_size
        if G is not None:
            dg = DeviceGraph(self.ctx, G)
            try:
                self.index = Index.build(self.ctx, dg, batch_size)
            finally:
                dg.close()

    def order_labels(self, rank2id, rank):
        self.index.order(rank)

    def query_distance(self, a, b):
        """Scalar (reference signature) or array arguments."""
        scalar = np.isscalar(a) and np.isscalar(b)
        out = self.index.query(np.atleast_1d(a), np.atleast_1d(b))
        return int(out[0]) if scalar else out

    def store_index_to_file(self, filename, rank):
        self.index.store(filename, rank)

    def load_index_from_file(self, filename):
        self.index = Index.load(self.ctx, filename)


VertexCentricPLL = ParaVertexCentricPLL
