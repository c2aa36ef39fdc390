"""The benchmark workloads at their full size, on the GPU, through the C ABI.

R-MAT 20, edge factor 16 (the unweighted GPU workload): the build's
bit-parallel records, canonical label sets and bp_hit_count equal the
reference OpenMP build's (SHA-256 digests in tests/golden, written by
tests/golden/make_golden.py from the reference's own ParaVertexCentricPLL);
10M uniform pairs answered in one launch -- the first 100K in reference mode
equal the reference's query_distance answers, the exact answers of 64 sampled
sources against every target equal BFS, and all 10M are symmetric.

Weighted R-MAT 22, edge factor 16, weights U[1, 100] (the weighted serving
workload): distances from sampled sources equal Dijkstra for every target,
the labels of sampled vertices are exactly their canonical pruned-landmark
sets (computed from one Dijkstra each), and served answers equal the plain
batch query's.
"""
import hashlib
import sys

import numpy as np
import pytest

import oracle
import parallel_distance_oracle_amd as pado
from helpers import bp_digest, label_digest, load_rmat, rank_space, rmat_edges

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def rmat20(gpu_ctx):
    meta, arr = load_rmat(20, 16)
    say("R-MAT 20 ef16 ingest")
    G, rank, r2i = rank_space(rmat_edges(20, 16))
    assert G.num_v == meta["num_v"]
    say(f"R-MAT 20 build n={G.num_v}")
    dg = pado.DeviceGraph(gpu_ctx, G)
    idx = pado.Index.build(gpu_ctx, dg)
    dg.close()
    yield meta, arr, G, rank, idx
    idx.close()


@pytest.mark.timeout(600)
def test_rmat20_labels_equal_reference(rmat20):
    meta, _, _, _, idx = rmat20
    bp_d, bp_s, loff, hub, dist = idx.export()
    assert hub.size == meta["num_labels"]
    assert bp_digest(bp_d, bp_s) == meta["bp_sha256"]
    assert label_digest(loff, hub, dist) == meta["labels_sha256"]
    assert idx.stats()["bp_pruned"] == meta["bp_hit_count"]


@pytest.mark.timeout(600)
def test_rmat20_ten_million_pairs_one_launch(gpu_ctx, rmat20):
    from parallel_distance_oracle_amd.hip import DeviceArray

    meta, arr, G, rank, idx = rmat20
    idx.order(rank)
    k = 10_000_000
    qa, qb = pado.uniform_pairs(G.num_v, k, seed=1)
    say("10M pairs, one launch")
    da, db = DeviceArray.from_numpy(qa), DeviceArray.from_numpy(qb)
    out = DeviceArray(k, np.uint8)
    idx.query_device(da.ptr, db.ptr, k, out.ptr)
    gpu_ctx.sync()
    got = out.numpy()
    # Symmetry over all 10M: d(a, b) = d(b, a), in a second launch.
    idx.query_device(db.ptr, da.ptr, k, out.ptr)
    gpu_ctx.sync()
    assert np.array_equal(out.numpy(), got)
    da.free()
    db.free()
    out.free()
    # The reference's answers on the first 100K pairs (reference mode), and the
    # exact answers differ from them only where the reference is wrong.
    n_ref = meta["n_pairs"]
    ref = idx.query(qa[:n_ref], qb[:n_ref], ref_compat=True)
    assert hashlib.sha256(ref.tobytes()).hexdigest() == meta["answers_sha256"]
    assert np.array_equal(ref, arr["answers"])
    # BFS from 64 sources: the sources of the first 64 pairs, every target.
    n = G.num_v
    tgt = np.arange(n, dtype=np.uint32)
    say("BFS from 64 sources")
    for i in range(64):
        s = int(qa[i])
        d = oracle.bfs_to_u8(oracle.bfs(n, G.off, G.adj, int(rank[s])))
        assert got[i] == d[rank[qb[i]]], i
        all_t = idx.query(np.full(n, s, np.uint32), tgt)
        assert np.array_equal(all_t, d[rank]), s
    # Where the exact answer and the reference's differ, BFS sides with exact.
    diff = np.nonzero(ref != got[:n_ref])[0][:64]
    for i in diff:
        d = oracle.bfs_to_u8(oracle.bfs(n, G.off, G.adj, int(rank[qa[i]])))
        assert got[i] == d[rank[qb[i]]] and ref[i] > got[i]


@pytest.fixture(scope="module")
def wrmat22(gpu_ctx):
    say("weighted R-MAT 22 ef16 ingest")
    e = pado.rmat_edges(22, 16, 0.57, 0.19, 0.19, 1)
    w = pado.edge_weights(e.shape[0], 1, 100, seed=1)
    G = pado.WeightedGraph.from_edges(e, w)
    del e, w
    rank = G.make_rank()
    r2i = G.id_transfer(rank)
    say(f"weighted R-MAT 22 build n={G.num_v} m={G.num_e}")
    dg = pado.DeviceGraph(gpu_ctx, G)
    X = pado.Index.build(gpu_ctx, dg)
    dg.close()
    X.order(rank)
    yield G, rank, r2i, X
    X.close()


@pytest.mark.timeout(900)
def test_weighted_rmat22_dijkstra_and_canonical_labels(gpu_ctx, wrmat22):
    G, rank, r2i, X = wrmat22
    n = G.num_v
    assert X.info.num_labels > 0
    loff, hub, dist = X.wexport()
    rng = np.random.default_rng(22)
    # Sources: two hubs (ranks 0 and 1), two random vertices, and the
    # lowest-ranked vertex that has an edge (rank order is by degree).
    deg = np.diff(G.off.astype(np.int64))
    low = int(np.nonzero(deg > 0)[0][-1])
    srcs = [0, 1, int(rng.integers(0, n)), int(rng.integers(0, n)), low]
    tgt = np.arange(n, dtype=np.uint32)
    for v in srcs:
        say(f"Dijkstra + canonical labels of rank-space vertex {v}")
        d, ch, cd = oracle.canonical_of(n, G.off, G.adj, G.w, v)
        got = X.query(np.full(n, v, np.uint32), tgt, rank_space=True)
        exp = np.where(d == 0xFFFFFFFF, pado.WINF, d).astype(np.uint32)
        assert np.array_equal(got, exp), v
        s, e = int(loff[v]), int(loff[v + 1])
        assert np.array_equal(hub[s:e], ch), v
        assert np.array_equal(dist[s:e], cd), v
    # Original-id queries agree with rank-space ones on random pairs.
    a, b = pado.uniform_pairs(n, 200000, seed=9)
    assert np.array_equal(X.query(a, b), X.query(rank[a], rank[b], rank_space=True))


@pytest.mark.timeout(600)
def test_weighted_rmat22_served_equals_plain(gpu_ctx, wrmat22):
    G, rank, r2i, X = wrmat22
    n = G.num_v
    rng = np.random.default_rng(5)
    reqs = []
    for i in range(48):
        k = int(np.floor(np.exp(rng.uniform(0, np.log(65536)))))
        half = k // 2
        a = np.concatenate([rng.integers(0, n, half), pado.zipf_vertices(n, k - half, 1.1, 100 + i, r2i)])
        b = np.concatenate([rng.integers(0, n, half), pado.zipf_vertices(n, k - half, 1.1, 200 + i, r2i)])
        reqs.append((a.astype(np.uint32), b.astype(np.uint32)))
    say("weighted R-MAT 22 serving")
    srv = pado.Server(gpu_ctx, X, staging_pairs=1 << 21, streams=3)
    try:
        ids = [srv.submit(a, b) for a, b in reqs]
        outs = [srv.wait(r, a.size)[0] for r, (a, _) in zip(ids, reqs)]
    finally:
        srv.stop()
    for (a, b), o in zip(reqs, outs):
        assert o.dtype == np.uint32 and np.array_equal(o, X.query(a, b))


@pytest.mark.timeout(900)
def test_multi_graph_job_1024_graphs(gpu_ctx):
    """1,024 graphs with vertex counts log-uniform in [2^10, 2^16], half R-MAT
    (edge factor 8), half grids with 5% of edges deleted, labeled as one job:
    every graph is labeled (the deep grids in the wide hop domain); every 4th
    graph inside the u8 domain equals the oracle's build of it alone bit for
    bit; every graph's answers from two sources to all targets equal BFS."""
    say("1,024-graph workload ingest")
    graphs = pado.mixed_graphs(1024, seed=4)
    say(f"1,024-graph job: {sum(G.num_v for G in graphs)} vertices")
    mi = pado.MultiIndex(gpu_ctx, graphs)
    infos = [mi.graph_info(g) for g in range(len(graphs))]
    assert all(x["status"] == 0 for x in infos)
    wide = [g for g, x in enumerate(infos) if x["wide"]]
    narrow = [g for g, x in enumerate(infos) if not x["wide"]]
    say(f"{len(narrow)} graphs in the u8 domain, {len(wide)} wide")
    assert len(wide) > 50 and all(g % 2 == 1 for g in wide)  # the deep ones are grids
    for g in narrow[::4]:
        G = graphs[g]
        bp_d, bp_s, loff, hub, dist = mi.export(g)
        X = oracle.Index(G.num_v, G.off, G.adj)
        assert np.array_equal(bp_d, X.bp_d) and np.array_equal(bp_s, X.bp_s), g
        assert np.array_equal(loff, X.loff) and np.array_equal(hub, X.hub) and np.array_equal(dist, X.dist), g
    say("BFS checks of every graph")
    far = 0
    for g, G in enumerate(graphs):
        n = G.num_v
        tgt = np.arange(n, dtype=np.uint32)
        for s in (0, n - 1):
            d = oracle.bfs(n, G.off, G.adj, s).astype(np.int64)
            got = mi.query(g, np.full(n, s, np.uint32), tgt)
            if infos[g]["wide"]:
                exp = np.where(d == 0xFFFFFFFF, pado.WIDE_INF, d)
                far += int(np.count_nonzero((d > 254) & (d != 0xFFFFFFFF)))
            else:
                exp = np.where(d >= 253, 255, d)
            assert np.array_equal(got.astype(np.int64), exp), (g, s)
    assert far > 0
    # Labels of wide graphs: hubs rank above their vertex; exports are u16.
    for g in wide[:8]:
        _, _, loff, hub, dist = mi.export(g)
        assert dist.dtype == np.uint16
        assert np.all(hub <= np.repeat(np.arange(graphs[g].num_v), np.diff(loff.astype(np.int64))))
    mi.close()
