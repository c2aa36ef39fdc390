"""GPU distance queries through the C ABI: exact against BFS, bit-exact to the
reference's query_distance in ref-compat mode, original-id ordering, device
pointers, file round trip."""
import os
import sys

import numpy as np
import pytest

import oracle
import parallel_distance_oracle_amd as pado
from helpers import GOLDEN, RMAT_SMALL, TINY, load_rmat, load_tiny, rank_space, rmat_edges

pytestmark = pytest.mark.gpu


def _pll(ctx, edges, rank=None):
    G, rank, r2i = rank_space(edges, rank)
    print(f"[gpu] build+order n={G.num_v}", flush=True)
    sys.stdout.flush()
    p = pado.ParaVertexCentricPLL(G, ctx=ctx)
    p.order_labels(r2i, rank)
    return p, G, rank, r2i


@pytest.mark.parametrize("name", TINY)
def test_tiny_all_pairs(gpu_ctx, name):
    g = load_tiny(name)
    p, G, rank, _ = _pll(gpu_ctx, g["edges"], g["rank"])
    qa, qb = g["qa"], g["qb"]
    assert np.array_equal(p.index.query(qa, qb, ref_compat=True), g["answers"])
    exact = oracle.bfs_to_u8(oracle.bfs_pairs(G.num_v, G.off, G.adj, rank[qa], rank[qb]))
    assert np.array_equal(p.query_distance(qa, qb), exact)


@pytest.mark.parametrize("scale,ef", RMAT_SMALL)
def test_rmat_uniform_pairs(gpu_ctx, scale, ef):
    meta, arr = load_rmat(scale, ef)
    p, G, rank, _ = _pll(gpu_ctx, rmat_edges(scale, ef))
    qa, qb = pado.uniform_pairs(G.num_v, meta["n_pairs"], seed=1)
    ref = p.index.query(qa, qb, ref_compat=True)
    assert np.array_equal(ref, arr["answers"])
    got = p.query_distance(qa, qb)
    exact = oracle.bfs_to_u8(oracle.bfs_pairs(G.num_v, G.off, G.adj, rank[qa], rank[qb]))
    assert np.array_equal(got, exact)
    # The two modes differ exactly on the reference's defect set.
    assert np.all((got == ref) | (ref != exact))
    # Rank-space queries agree with original-id queries.
    assert np.array_equal(p.index.query(rank[qa], rank[qb], rank_space=True), got)


def test_zipf_pairs_hit_the_defect_set(gpu_ctx):
    """Hub-heavy pairs (Zipf over degree rank) exercise the bit-parallel roots."""
    p, G, rank, r2i = _pll(gpu_ctx, rmat_edges(14, 8))
    a = pado.zipf_vertices(G.num_v, 50000, 1.1, seed=2, rank2id=r2i)
    b = pado.zipf_vertices(G.num_v, 50000, 1.1, seed=3, rank2id=r2i)
    exact = oracle.bfs_to_u8(oracle.bfs_pairs(G.num_v, G.off, G.adj, rank[a], rank[b]))
    assert np.array_equal(p.query_distance(a, b), exact)
    X = oracle.Index(G.num_v, G.off, G.adj)
    assert np.array_equal(p.index.query(a, b, ref_compat=True), X.query(rank[a], rank[b], ref_compat=True))


def test_out_of_range_and_self(gpu_ctx):
    p, G, rank, _ = _pll(gpu_ctx, rmat_edges(10, 8))
    n = G.num_v
    a = np.array([0, n, n + 5, n, 3, 3], np.uint32)
    b = np.array([0, n, n, 2, n + 1, 3], np.uint32)
    out = p.query_distance(a, b)
    assert list(out) == [0, 0, 255, 255, 255, 0]
    assert p.query_distance(5, 5) == 0


def test_query_before_order_is_an_error(gpu_ctx):
    G, _, _ = rank_space(rmat_edges(10, 8))
    p = pado.ParaVertexCentricPLL(G, ctx=gpu_ctx)
    with pytest.raises(pado.PadoError):
        p.query_distance(np.array([1], np.uint32), np.array([2], np.uint32))


def test_device_pointer_queries(gpu_ctx):
    """Pairs already in HBM (buffers of the library's own HIP runtime), answers
    written to HBM, on the context's stream."""
    from parallel_distance_oracle_amd.hip import DeviceArray

    p, G, rank, _ = _pll(gpu_ctx, rmat_edges(12, 8))
    qa, qb = pado.uniform_pairs(G.num_v, 30000, seed=9)
    da, db = DeviceArray.from_numpy(qa), DeviceArray.from_numpy(qb)
    out = DeviceArray(qa.size, np.uint8)
    p.index.query_device(da.ptr, db.ptr, qa.size, out.ptr, stream=gpu_ctx.stream)
    gpu_ctx.sync()
    assert np.array_equal(out.numpy(), p.query_distance(qa, qb))


def test_store_load_roundtrip_and_format(gpu_ctx, tmp_path):
    g = load_tiny("edgelist_in")
    p, G, rank, _ = _pll(gpu_ctx, g["edges"], g["rank"])
    path = str(tmp_path / "idx.bin")
    p.store_index_to_file(path, rank)
    # Byte layout of the reference's writer (UNP:1121-1180).
    n = G.num_v
    exp = bytearray(np.uint32(n).tobytes() + np.uint32(50).tobytes())
    for x in range(n):
        v = rank[x]
        for i in range(50):
            exp += np.uint8(g["bp_d"][v, i]).tobytes() + np.uint64(g["bp_s"][v, i, 0]).tobytes() + \
                np.uint64(g["bp_s"][v, i, 1]).tobytes()
        lo, hi = int(g["loff"][v]), int(g["loff"][v + 1])
        exp += np.uint32(hi - lo).tobytes()
        for j in range(lo, hi):
            exp += np.uint32(g["hub"][j]).tobytes() + np.uint8(g["dist"][j]).tobytes()
    assert open(path, "rb").read() == bytes(exp)
    q = pado.ParaVertexCentricPLL(ctx=gpu_ctx)
    q.load_index_from_file(path)
    assert np.array_equal(q.query_distance(g["qa"], g["qb"]), p.query_distance(g["qa"], g["qb"]))
    os.remove(path)


def test_export_import_roundtrip(gpu_ctx):
    G, rank, r2i = rank_space(rmat_edges(12, 8))
    idx = pado.Index.build(gpu_ctx, pado.DeviceGraph(gpu_ctx, G))
    bp_d, bp_s, loff, hub, dist = idx.export()
    idx2 = pado.Index.import_(gpu_ctx, G.num_v, bp_d, bp_s, loff, hub, dist)
    e2 = idx2.export()
    for x, y in zip((bp_d, bp_s, loff, hub, dist), e2):
        assert np.array_equal(x, y)
    qa, qb = pado.uniform_pairs(G.num_v, 5000, seed=4)
    assert np.array_equal(idx.query(qa, qb, rank_space=True), idx2.query(qa, qb, rank_space=True))


@pytest.mark.parametrize("weighted", [False, True])
def test_host_array_queries_in_chunks(gpu_ctx, weighted):
    """Host-array queries longer than one 4M-pair chunk (copies and kernels of
    consecutive chunks on two streams, a ragged last chunk) give the answers of
    one device-pointer launch over the same pairs, exact and
    reference-compatible; out-of-range ids and self pairs sit in several chunks."""
    from parallel_distance_oracle_amd.hip import DeviceArray

    e = rmat_edges(12, 8)
    if weighted:
        G = pado.WeightedGraph.from_edges(e, pado.edge_weights(e.shape[0], 1, 100, seed=1))
    else:
        G = pado.Graph.from_edges(e)
    rank = G.make_rank()
    G.id_transfer(rank)
    print(f"[gpu] chunked host queries weighted={weighted} n={G.num_v}", flush=True)
    idx = pado.Index.build(gpu_ctx, pado.DeviceGraph(gpu_ctx, G))
    idx.order(rank)
    k = 9_000_001  # chunks of 4M, 4M and 1M + 1 pairs
    qa, qb = pado.uniform_pairs(G.num_v, k, seed=13)
    for pos in (5, 4_194_300, 4_194_310, 8_388_700, k - 1):
        qa[pos] = G.num_v + 3
        qb[pos + 1 if pos + 1 < k else pos - 1] = qa[pos + 1 if pos + 1 < k else pos - 1]
    dt = np.uint32 if weighted else np.uint8
    da, db = DeviceArray.from_numpy(qa), DeviceArray.from_numpy(qb)
    out = DeviceArray(k, dt)
    for ref in (False, True):
        host = idx.query(qa, qb, ref_compat=ref)
        idx.query_device(da.ptr, db.ptr, k, out.ptr, ref_compat=ref, stream=gpu_ctx.stream)
        gpu_ctx.sync()
        dev = out.numpy()
        assert host.dtype == dt and host.size == k
        assert np.array_equal(host, dev)
    da.free()
    db.free()
    out.free()
    idx.close()


@pytest.mark.parametrize("tag", ["edgelist_in", "input_dat", "rmat10_ef8"])
def test_reference_written_index_file(gpu_ctx, tmp_path, tag):
    """An index file written by the reference's own writer
    (VertexCentricPLL<1024>::store_index_to_file, UNP:1120-1180; committed
    bytes) loads into the library and answers exactly (BFS); the library's
    own build of the same graph, stored, is byte-identical to it."""
    g = np.load(os.path.join(GOLDEN, f"unp_{tag}.npz"))
    ref_bytes = g["index"].tobytes()
    path = str(tmp_path / "ref.idx")
    with open(path, "wb") as f:
        f.write(ref_bytes)
    print(f"[gpu] reference-written index {tag}", flush=True)
    X = pado.Index.load(gpu_ctx, path)
    G, rank, _ = rank_space(g["edges"], g["rank"])
    n = G.num_v
    a, b = np.meshgrid(np.arange(n, dtype=np.uint32), np.arange(n, dtype=np.uint32), indexing="ij")
    a, b = a.reshape(-1), b.reshape(-1)
    if a.size > 200000:
        sel = np.random.default_rng(1).choice(a.size, 200000, replace=False)
        a, b = a[sel], b[sel]
    exact = oracle.bfs_to_u8(oracle.bfs_pairs(n, G.off, G.adj, rank[a], rank[b]))
    assert np.array_equal(X.query(a, b), exact)
    X.close()
    p = pado.ParaVertexCentricPLL(G, ctx=gpu_ctx)
    mine = str(tmp_path / "mine.idx")
    p.store_index_to_file(mine, rank)
    with open(mine, "rb") as f:
        assert f.read() == ref_bytes
