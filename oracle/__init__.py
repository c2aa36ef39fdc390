"""CPU oracle for the PADO GPU library -- TEST INFRASTRUCTURE ONLY.

Wraps two checkers:
  * liboracle.so: the C restatement in pado_oracle.c (sequential, exact
    reference semantics; every function cites the reference file:line it
    follows);
  * _ref/pado_ref: the reference implementation itself, compiled from the
    reference's headers by build_ref.sh (present only where it was built).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package.  The product (parallel_distance_oracle_amd) never does.
"""
import ctypes
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "pado_ref")
BP = 50

_lib = None

u8p = np.ctypeslib.ndpointer(np.uint8, flags="C")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C")


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.po_graph_num_v.restype = ctypes.c_uint32
        L.po_graph_num_v.argtypes = [u32p, ctypes.c_uint64]
        L.po_graph_csr.argtypes = [u32p, ctypes.c_uint64, ctypes.c_uint32, u64p, u32p]
        L.po_make_rank.argtypes = [u64p, ctypes.c_uint32, u32p]
        L.po_id_transfer.argtypes = [ctypes.c_uint32, u64p, u32p, u32p, u64p, u32p, u32p]
        L.po_bfs.argtypes = [ctypes.c_uint32, u64p, u32p, ctypes.c_uint32, u32p]
        L.po_bfs_pairs.argtypes = [ctypes.c_uint32, u64p, u32p, u32p, u32p, ctypes.c_uint64, u32p]
        L.po_bp_labeling.restype = ctypes.c_uint32
        L.po_bp_labeling.argtypes = [ctypes.c_uint32, u64p, u32p, u8p, u64p, u8p, u32p]
        L.po_labeling.restype = ctypes.c_void_p
        L.po_labeling.argtypes = [ctypes.c_uint32, u64p, u32p, u8p, u64p, u8p, ctypes.c_uint32]
        L.po_index_num_labels.restype = ctypes.c_uint64
        L.po_index_num_labels.argtypes = [ctypes.c_void_p]
        L.po_index_stats.argtypes = [ctypes.c_void_p, u64p]
        L.po_index_export.argtypes = [ctypes.c_void_p, u64p, u32p, u8p]
        L.po_index_free.argtypes = [ctypes.c_void_p]
        L.po_query.restype = ctypes.c_uint32
        L.po_query.argtypes = [ctypes.c_uint32, u8p, u64p, u64p, u32p, u8p,
                               ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int]
        L.po_query_batch.argtypes = [ctypes.c_uint32, u8p, u64p, u64p, u32p, u8p,
                                     u32p, u32p, ctypes.c_uint64, u8p, ctypes.c_int]
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# --- GR:120-206 ----------------------------------------------------------
def graph_csr(edges):
    """Graph(edge_list) (GR:120-153): returns (n, off[n+1] u64, adj[2m] u32)."""
    e = _c(edges, np.uint32).reshape(-1)
    m = e.size // 2
    n = lib().po_graph_num_v(e, m) if m else 0
    off = np.zeros(n + 1, np.uint64)
    adj = np.zeros(max(2 * m, 1), np.uint32)
    if n:
        lib().po_graph_csr(e, m, n, off, adj)
    return n, off, adj[: 2 * m]


def make_rank(n, off):
    """Graph::make_rank (GR:156-174) with glibc rand() default seed."""
    rank = np.zeros(max(n, 1), np.uint32)
    lib().po_make_rank(_c(off, np.uint64), n, rank)
    return rank[:n]


def id_transfer(n, off, adj, rank):
    """Graph::id_transfer (GR:176-206): (off', adj', rank2id)."""
    off2 = np.zeros(n + 1, np.uint64)
    adj2 = np.zeros(max(adj.size, 1), np.uint32)
    r2i = np.zeros(max(n, 1), np.uint32)
    lib().po_id_transfer(n, _c(off, np.uint64), _c(adj, np.uint32) if adj.size else np.zeros(1, np.uint32),
                         _c(rank, np.uint32), off2, adj2, r2i)
    return off2, adj2[: adj.size], r2i[:n]


def bfs(n, off, adj, src):
    d = np.zeros(max(n, 1), np.uint32)
    lib().po_bfs(n, _c(off, np.uint64), _c(adj, np.uint32) if adj.size else np.zeros(1, np.uint32), src, d)
    return d[:n]


def bfs_pairs(n, off, adj, a, b):
    """Exact hop distances for pairs (uint32, 0xFFFFFFFF = unreachable)."""
    a = _c(a, np.uint32).reshape(-1)
    b = _c(b, np.uint32).reshape(-1)
    out = np.zeros(max(a.size, 1), np.uint32)
    lib().po_bfs_pairs(n, _c(off, np.uint64), _c(adj, np.uint32) if adj.size else np.zeros(1, np.uint32),
                       a if a.size else np.zeros(1, np.uint32), b if b.size else np.zeros(1, np.uint32),
                       a.size, out)
    return out[: a.size]


def bfs_to_u8(d):
    """Map exact hop distances to the reference's u8 answer domain:
    unreachable or >= 253 -> 255 (PARA:2412-2414)."""
    d = np.asarray(d, np.uint64)
    return np.where(d >= 253, 255, d).astype(np.uint8)


# --- PARA:544-1836 -------------------------------------------------------
def bp_labeling(n, off, adj):
    """bit_parallel_labeling (PARA:544-840) -> (bp_d[n,50], bp_s[n,50,2], used[n], roots[50], max_level)."""
    bp_d = np.zeros(max(n, 1) * BP, np.uint8)
    bp_s = np.zeros(max(n, 1) * BP * 2, np.uint64)
    used = np.zeros(max(n, 1), np.uint8)
    roots = np.zeros(BP, np.uint32)
    lvl = lib().po_bp_labeling(n, _c(off, np.uint64), _c(adj, np.uint32) if adj.size else np.zeros(1, np.uint32),
                               bp_d, bp_s, used, roots)
    return (bp_d[: n * BP].reshape(n, BP), bp_s[: n * BP * 2].reshape(n, BP, 2), used[:n], roots, lvl)


def labeling(n, off, adj, bp_d, bp_s, used, batch=1024):
    """construct (PARA:1740-1836) after BP: canonical labels (off, hub, dist) + stats."""
    L = lib()
    h = L.po_labeling(n, _c(off, np.uint64), _c(adj, np.uint32) if adj.size else np.zeros(1, np.uint32),
                      _c(bp_d, np.uint8).reshape(-1), _c(bp_s, np.uint64).reshape(-1), _c(used, np.uint8), batch)
    try:
        tot = L.po_index_num_labels(h)
        loff = np.zeros(n + 1, np.uint64)
        hub = np.zeros(max(tot, 1), np.uint32)
        dist = np.zeros(max(tot, 1), np.uint8)
        L.po_index_export(h, loff, hub, dist)
        st = np.zeros(7, np.uint64)
        L.po_index_stats(h, st)
    finally:
        L.po_index_free(h)
    keys = ["iters", "pushes", "ind_pass", "bp_pruned", "cands", "dq_pruned", "bytes_alg"]
    return loff, hub[:tot], dist[:tot], dict(zip(keys, (int(x) for x in st)))


class Index:
    """A complete oracle index over a rank-space graph."""

    def __init__(self, n, off, adj, batch=1024):
        self.n = n
        self.bp_d, self.bp_s, self.used, self.roots, self.max_level = bp_labeling(n, off, adj)
        self.loff, self.hub, self.dist, self.stats = labeling(n, off, adj, self.bp_d, self.bp_s, self.used, batch)

    def query(self, a, b, ref_compat=False):
        """query_distance (PARA:2283-2416) on rank-space ids."""
        a = _c(a, np.uint32).reshape(-1)
        b = _c(b, np.uint32).reshape(-1)
        out = np.zeros(max(a.size, 1), np.uint8)
        hub = self.hub if self.hub.size else np.zeros(1, np.uint32)
        dist = self.dist if self.dist.size else np.zeros(1, np.uint8)
        lib().po_query_batch(self.n, self.bp_d.reshape(-1), self.bp_s.reshape(-1), self.loff, hub, dist,
                             a, b, a.size, out, 1 if ref_compat else 0)
        return out[: a.size]


# --- reference harness (oracle/_ref/pado_ref) --------------------------------
def have_ref():
    return os.path.exists(REF_BIN)


def _write_edges(path, edges):
    e = _c(edges, np.uint32).reshape(-1)
    with open(path, "wb") as f:
        f.write(np.uint64(e.size // 2).tobytes())
        f.write(e.tobytes())


def _write_rank(path, rank):
    r = _c(rank, np.uint32)
    with open(path, "wb") as f:
        f.write(np.uint32(r.size).tobytes())
        f.write(r.tobytes())


def read_labels(path):
    with open(path, "rb") as f:
        n = int(np.frombuffer(f.read(4), np.uint32)[0])
        tot = int(np.frombuffer(f.read(8), np.uint64)[0])
        off = np.frombuffer(f.read(8 * (n + 1)), np.uint64).copy()
        hub = np.frombuffer(f.read(4 * tot), np.uint32).copy()
        dist = np.frombuffer(f.read(tot), np.uint8).copy()
    return off, hub, dist


def read_bp(path):
    with open(path, "rb") as f:
        n = int(np.frombuffer(f.read(4), np.uint32)[0])
        d = np.frombuffer(f.read(n * BP), np.uint8).reshape(n, BP).copy()
        s = np.frombuffer(f.read(n * BP * 16), np.uint64).reshape(n, BP, 2).copy()
    return d, s


def run_ref(edges, rank=None, pairs=None, threads=None, repeat=1, dump=True, env_extra=None):
    """Run the reference build on `edges` (original ids) with `rank` (or its own
    make_rank when None).  Returns dict with labels/bp (rank space), answers
    (original-id pairs) and the timing lines it printed: ref_build_wall_s
    (constructor wall time, last build), ref_build_best_s, build_walls (every
    build), total_labeling_s (the reference's own Total_labeling_time per
    build, PARA:1770-1837) and, with pairs, ref_query_serial_s /
    ref_query_omp_s."""
    if not have_ref():
        raise FileNotFoundError(REF_BIN)
    env = dict(os.environ)
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    if env_extra:
        env.update(env_extra)
    with tempfile.TemporaryDirectory() as td:
        ep = os.path.join(td, "e.bin")
        _write_edges(ep, edges)
        rp = "-"
        if rank is not None:
            rp = os.path.join(td, "r.bin")
            _write_rank(rp, rank)
        pp = "-"
        if pairs is not None:
            pp = os.path.join(td, "p.bin")
            a = _c(pairs[0], np.uint32)
            b = _c(pairs[1], np.uint32)
            with open(pp, "wb") as f:
                f.write(np.uint64(a.size).tobytes())
                f.write(a.tobytes())
                f.write(b.tobytes())
        pre = os.path.join(td, "out") if dump else "-"
        cp = subprocess.run([REF_BIN, "unw", ep, rp, pre, pp, str(repeat)], env=env,
                            capture_output=True, text=True, check=True)
        res = {"stdout": cp.stdout, "build_walls": [], "total_labeling_s": []}
        for line in cp.stdout.splitlines():
            toks = line.split()
            if toks[:1] == ["ref_build_wall_s:"]:
                res["build_walls"].append(float(toks[1]))
            if "Total_labeling_time:" in toks:
                res["total_labeling_s"].append(float(toks[toks.index("Total_labeling_time:") + 1]))
            if "bp_hit_count:" in toks:
                # printed by construct (PARA:1889); one line per build
                res["bp_hit_count"] = int(toks[toks.index("bp_hit_count:") + 1].replace(",", ""))
            for i, t in enumerate(toks[:-1]):
                if t.endswith(":") and t.startswith("ref_"):
                    try:
                        res[t[:-1]] = float(toks[i + 1])
                    except ValueError:
                        pass
        if dump:
            res["labels"] = read_labels(pre + ".labels.bin")
            res["bp"] = read_bp(pre + ".bp.bin")
            if pairs is not None:
                res["answers"] = np.fromfile(pre + ".answers.bin", np.uint8)
    return res


def ref_unp_file(edges, rank):
    """The index file the reference's own writer produces
    (VertexCentricPLL<1024>::store_index_to_file, UNP:1120-1180) for `edges`
    in the vertex order `rank`, as bytes."""
    unp = os.path.join(HERE, "_ref", "pado_ref_unp")
    if not os.path.exists(unp):
        raise FileNotFoundError(unp)
    with tempfile.TemporaryDirectory() as td:
        ep, rp, ip = (os.path.join(td, x) for x in ("e.bin", "r.bin", "i.idx"))
        _write_edges(ep, edges)
        _write_rank(rp, rank)
        subprocess.run([unp, ep, rp, ip], check=True, capture_output=True)
        with open(ip, "rb") as f:
            return f.read()


def ref_rank(edges):
    """The reference's own make_rank output for `edges` (fresh process)."""
    with tempfile.TemporaryDirectory() as td:
        ep = os.path.join(td, "e.bin")
        _write_edges(ep, edges)
        rp = os.path.join(td, "r.bin")
        subprocess.run([REF_BIN, "rank", ep, rp], check=True, capture_output=True)
        with open(rp, "rb") as f:
            n = int(np.frombuffer(f.read(4), np.uint32)[0])
            return np.frombuffer(f.read(4 * n), np.uint32).copy()


def read_text_edges(path):
    """Edge list text as the reference's Graph(const char*) reads it (GR:95-118):
    lines starting with '#' or '%' skipped; first two integers per line."""
    out = []
    with open(path) as f:
        for line in f:
            if not line or line[0] in "#%":
                continue
            t = line.split()
            if len(t) >= 2:
                out.append((int(t[0]), int(t[1])))
    return np.array(out, dtype=np.uint32).reshape(-1, 2)


def set_mutation(m):
    """Planted-mistake switch (tests only); 0 restores the faithful oracle."""
    lib().po_set_mutation(int(m))


# --- weighted (WPARA) ----------------------------------------------------------
i16p = np.ctypeslib.ndpointer(np.int16, flags="C")


def _wlib():
    L = lib()
    if not getattr(L, "_w_ready", False):
        L.wo_id_transfer.argtypes = [ctypes.c_uint32, u64p, u32p, u8p, u32p, u64p, u32p, u8p, u32p]
        L.wo_labeling.restype = ctypes.c_void_p
        L.wo_labeling.argtypes = [ctypes.c_uint32, u64p, u32p, u8p, ctypes.c_uint32]
        L.wo_num_labels.restype = ctypes.c_uint64
        L.wo_num_labels.argtypes = [ctypes.c_void_p]
        L.wo_export.argtypes = [ctypes.c_void_p, u64p, u32p, i16p]
        L.wo_free.argtypes = [ctypes.c_void_p]
        L.wo_query_batch.argtypes = [ctypes.c_uint32, u64p, u32p, i16p, u32p, u32p, ctypes.c_uint64, i16p]
        L.wo_dijkstra.argtypes = [ctypes.c_uint32, u64p, u32p, u8p, ctypes.c_uint32, u32p]
        L.wo_canonical_of.restype = ctypes.c_uint64
        L.wo_canonical_of.argtypes = [ctypes.c_uint32, u64p, u32p, u8p, ctypes.c_uint32, u32p, u32p, u32p,
                                      ctypes.c_uint64]
        L._w_ready = True
    return L


def weighted_graph_csr(edges, weights):
    """WeightedGraph(edge_list, weight_list) (GR:306-351): CSR + u8 weights."""
    e = _c(edges, np.uint32).reshape(-1, 2)
    w = _c(weights, np.uint32).astype(np.uint8)
    n = int(e.max()) + 1 if e.size else 0
    deg = np.bincount(e.reshape(-1), minlength=n)
    off = np.zeros(n + 1, np.uint64)
    off[1:] = np.cumsum(deg)
    adj = np.zeros(2 * e.shape[0], np.uint32)
    ww = np.zeros(2 * e.shape[0], np.uint8)
    pos = off[:-1].copy()
    for i in range(e.shape[0]):  # edge order as the reference appends
        a, b = int(e[i, 0]), int(e[i, 1])
        adj[pos[a]] = b; ww[pos[a]] = w[i]; pos[a] += 1
        adj[pos[b]] = a; ww[pos[b]] = w[i]; pos[b] += 1
    return n, off, adj, ww


def weighted_id_transfer(n, off, adj, w, rank):
    off2 = np.zeros(n + 1, np.uint64)
    adj2 = np.zeros(max(adj.size, 1), np.uint32)
    w2 = np.zeros(max(adj.size, 1), np.uint8)
    r2i = np.zeros(max(n, 1), np.uint32)
    _wlib().wo_id_transfer(n, _c(off, np.uint64), _c(adj, np.uint32), _c(w, np.uint8), _c(rank, np.uint32),
                           off2, adj2, w2, r2i)
    return off2, adj2[: adj.size], w2[: adj.size], r2i[:n]


class WeightedIndex:
    """WeightedParaVertexCentricPLL<B> restated (single-thread schedule)."""

    def __init__(self, n, off, adj, w, batch=512):
        L = _wlib()
        h = L.wo_labeling(n, _c(off, np.uint64), _c(adj, np.uint32), _c(w, np.uint8), batch)
        try:
            tot = L.wo_num_labels(h)
            self.loff = np.zeros(n + 1, np.uint64)
            self.hub = np.zeros(max(tot, 1), np.uint32)
            self.dist = np.zeros(max(tot, 1), np.int16)
            L.wo_export(h, self.loff, self.hub, self.dist)
            self.hub, self.dist = self.hub[:tot], self.dist[:tot]
        finally:
            L.wo_free(h)
        self.n = n

    def query(self, a, b):
        a = _c(a, np.uint32).reshape(-1)
        b = _c(b, np.uint32).reshape(-1)
        out = np.zeros(max(a.size, 1), np.int16)
        hub = self.hub if self.hub.size else np.zeros(1, np.uint32)
        dist = self.dist if self.dist.size else np.zeros(1, np.int16)
        _wlib().wo_query_batch(self.n, self.loff, hub, dist, a, b, a.size, out)
        return out[: a.size]


def dijkstra(n, off, adj, w, src):
    d = np.zeros(max(n, 1), np.uint32)
    _wlib().wo_dijkstra(n, _c(off, np.uint64), _c(adj, np.uint32), _c(w, np.uint8), src, d)
    return d[:n]


def canonical_of(n, off, adj, w, v):
    """Dijkstra distances from rank-space vertex v and v's canonical label
    set (hubs ascending, distances) from the same run: (dist[n], hub, hd)."""
    L = _wlib()
    d = np.zeros(max(n, 1), np.uint32)
    off, adj, w = _c(off, np.uint64), _c(adj, np.uint32), _c(w, np.uint8)
    cap = 4096
    while True:
        hub = np.zeros(cap, np.uint32)
        hd = np.zeros(cap, np.uint32)
        cnt = int(L.wo_canonical_of(n, off, adj, w, v, d, hub, hd, cap))
        if cnt <= cap:
            return d[:n], hub[:cnt], hd[:cnt]
        cap = cnt


def read_wlabels(path):
    with open(path, "rb") as f:
        n = int(np.frombuffer(f.read(4), np.uint32)[0])
        tot = int(np.frombuffer(f.read(8), np.uint64)[0])
        off = np.frombuffer(f.read(8 * (n + 1)), np.uint64).copy()
        hub = np.frombuffer(f.read(4 * tot), np.uint32).copy()
        dist = np.frombuffer(f.read(2 * tot), np.int16).copy()
    return off, hub, dist


def run_ref_weighted(edges, weights, rank=None, pairs=None, threads=1):
    """Reference WeightedParaVertexCentricPLL<512> build (rank-space labels) and
    query answers for original-id pairs."""
    env = dict(os.environ)
    env["OMP_NUM_THREADS"] = str(threads)
    e = _c(edges, np.uint32).reshape(-1, 2)
    w = _c(weights, np.uint32)
    with tempfile.TemporaryDirectory() as td:
        ep = os.path.join(td, "w.bin")
        with open(ep, "wb") as f:
            f.write(np.uint64(e.shape[0]).tobytes())
            f.write(np.column_stack([e, w]).astype(np.uint32).tobytes())
        rp = "-"
        if rank is not None:
            rp = os.path.join(td, "r.bin")
            _write_rank(rp, rank)
        pp = "-"
        if pairs is not None:
            pp = os.path.join(td, "p.bin")
            with open(pp, "wb") as f:
                f.write(np.uint64(len(pairs[0])).tobytes())
                f.write(_c(pairs[0], np.uint32).tobytes())
                f.write(_c(pairs[1], np.uint32).tobytes())
        pre = os.path.join(td, "out")
        cp = subprocess.run([REF_BIN, "w", ep, rp, pre, pp], env=env, capture_output=True, text=True, check=True)
        res = {"stdout": cp.stdout, "labels": read_wlabels(pre + ".wlabels.bin")}
        if pairs is not None:
            res["answers"] = np.fromfile(pre + ".answers.bin", np.int16)
    return res


def ref_text_csr(path, weighted=False):
    """The reference's own text reader (Graph(const char*), GR:95-118, or
    WeightedGraph(const char*), GR:277-304) on `path`: (num_v, off, adj[, w])
    of the CSR it builds."""
    if not have_ref():
        raise FileNotFoundError(REF_BIN)
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "csr.bin")
        subprocess.run([REF_BIN, "wtext" if weighted else "text", str(path), out], check=True,
                       capture_output=True)
        with open(out, "rb") as f:
            n = int(np.frombuffer(f.read(4), np.uint32)[0])
            m = int(np.frombuffer(f.read(8), np.uint64)[0])
            off = np.frombuffer(f.read(8 * (n + 1)), np.uint64).copy()
            adj = np.frombuffer(f.read(4 * m), np.uint32).copy()
            if weighted:
                w = np.frombuffer(f.read(m), np.uint8).copy()
                return n, off, adj, w
    return n, off, adj


LINE_FIELDS = os.path.join(HERE, "line_fields")


def line_fields(path, ncols=2):
    """The reference's text-reader line loop (GR:95-118, 277-304) run by the
    C++ standard library's own istringstream (line_fields.cpp): an (m, ncols)
    u32 array of the values each line leaves in the loop's variables."""
    if not os.path.exists(LINE_FIELDS):
        subprocess.run(["make", "-s", "-C", HERE, "line_fields"], check=True)
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "rows.bin")
        subprocess.run([LINE_FIELDS, str(ncols), str(path), out], check=True, capture_output=True)
        with open(out, "rb") as f:
            m = int(np.frombuffer(f.read(8), np.uint64)[0])
            rows = np.frombuffer(f.read(4 * m * ncols), np.uint32).copy()
    return rows.reshape(-1, ncols)
