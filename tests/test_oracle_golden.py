"""The CPU oracle against the reference's own outputs (golden fixtures written
by tests/golden/make_golden.py from the reference build)."""
import os

import numpy as np
import pytest

import oracle
from helpers import (GOLDEN, RMAT_SMALL, TINY, bp_digest, digest, have_rmat, label_digest, load_rmat, load_tiny,
                     oracle_rank_space, rmat_edges)


@pytest.mark.parametrize("name", TINY)
def test_tiny_bp_and_labels(name):
    g = load_tiny(name)
    n, off, adj, _ = oracle_rank_space(g["edges"], g["rank"])
    X = oracle.Index(n, off, adj)
    assert np.array_equal(X.bp_d, g["bp_d"])
    assert np.array_equal(X.bp_s, g["bp_s"])
    assert np.array_equal(X.loff, g["loff"])
    assert np.array_equal(X.hub, g["hub"]) and np.array_equal(X.dist, g["dist"])


@pytest.mark.parametrize("name", TINY)
def test_tiny_query_ref_compat_equals_reference(name):
    """All pairs: the oracle's ref_compat query reproduces query_distance."""
    g = load_tiny(name)
    rank = g["rank"]
    n, off, adj, _ = oracle_rank_space(g["edges"], rank)
    X = oracle.Index(n, off, adj)
    got = X.query(rank[g["qa"]], rank[g["qb"]], ref_compat=True)
    assert np.array_equal(got, g["answers"])


@pytest.mark.parametrize("name", TINY)
def test_tiny_reference_defect_is_the_unsigned_bp_gate(name):
    """Where the reference disagrees with BFS, the signed oracle agrees with
    BFS and the disagreement disappears: the defect set is exactly the pairs
    the unsigned comparison (PARA:2316) mis-answers."""
    g = load_tiny(name)
    rank = g["rank"]
    n, off, adj, _ = oracle_rank_space(g["edges"], rank)
    X = oracle.Index(n, off, adj)
    a, b = rank[g["qa"]], rank[g["qb"]]
    exact = oracle.bfs_to_u8(oracle.bfs_pairs(n, off, adj, a, b))
    signed = X.query(a, b)
    assert np.array_equal(signed, exact)
    defect = g["answers"] != exact
    # Every defect pair involves a bit-parallel root (distance 0 or 1 to a slot root).
    roots = set(int(r) for r in X.roots if r != 0xFFFFFFFF)
    for i in np.nonzero(defect)[0][:200]:
        da = X.bp_d[a[i]]
        db = X.bp_d[b[i]]
        assert np.any((da.astype(int) + db.astype(int)) < 2), (int(a[i]), int(b[i]))
    assert len(roots) > 0


@pytest.mark.parametrize("scale,ef", RMAT_SMALL)
def test_rmat_labels_and_bp(scale, ef):
    meta, arr = load_rmat(scale, ef)
    edges = rmat_edges(scale, ef)
    assert digest(edges.astype(np.uint32)) == meta["edges_sha256"]
    n, off0, adj0 = oracle.graph_csr(edges)
    rank = oracle.make_rank(n, off0)
    assert digest(rank) == meta["rank_sha256"]
    n, off, adj, _ = oracle_rank_space(edges, rank)
    X = oracle.Index(n, off, adj)
    assert X.hub.size == meta["num_labels"]
    assert X.stats["bp_pruned"] == meta["bp_hit_count"]  # reference's bp_hit_count (PARA:1179, 1889)
    assert np.array_equal(X.loff, arr["loff"])
    assert np.array_equal(X.hub, arr["hub"]) and np.array_equal(X.dist, arr["dist"])
    assert bp_digest(X.bp_d, X.bp_s) == meta["bp_sha256"]


@pytest.mark.parametrize("scale,ef", RMAT_SMALL)
def test_rmat_queries(scale, ef):
    import parallel_distance_oracle_amd as pado
    meta, arr = load_rmat(scale, ef)
    edges = rmat_edges(scale, ef)
    n, off0, adj0 = oracle.graph_csr(edges)
    rank = oracle.make_rank(n, off0)
    n, off, adj, _ = oracle_rank_space(edges, rank)
    X = oracle.Index(n, off, adj)
    qa, qb = pado.uniform_pairs(n, meta["n_pairs"], seed=1)
    ra, rb = rank[qa], rank[qb]
    assert np.array_equal(X.query(ra, rb, ref_compat=True), arr["answers"])
    exact = oracle.bfs_to_u8(oracle.bfs_pairs(n, off, adj, ra, rb))
    assert np.array_equal(X.query(ra, rb), exact)


@pytest.mark.slow
@pytest.mark.skipif(not have_rmat(16, 16), reason="R-MAT 16 digest fixture not generated")
def test_rmat16_digests():
    meta, arr = load_rmat(16, 16)
    edges = rmat_edges(16, 16)
    n, off0, _ = oracle.graph_csr(edges)
    rank = oracle.make_rank(n, off0)
    n, off, adj, _ = oracle_rank_space(edges, rank)
    X = oracle.Index(n, off, adj)
    assert label_digest(X.loff, X.hub, X.dist) == meta["labels_sha256"]
    assert bp_digest(X.bp_d, X.bp_s) == meta["bp_sha256"]


def _parse_unp(buf):
    """The reference's index file (UNP:1120-1180): u32 num_v, u32 50, then per
    original id 50 x (u8 dist, u64 S^-1, u64 S^0), u32 size, size x (u32 hub,
    u8 dist)."""
    n, bp = np.frombuffer(buf[:8], np.uint32)
    pos = 8
    rec = np.dtype([("d", "u1"), ("s0", "<u8"), ("s1", "<u8")])
    bp_d = np.zeros((n, bp), np.uint8)
    bp_s = np.zeros((n, bp, 2), np.uint64)
    off = [0]
    hubs, dists = [], []
    lab = np.dtype([("h", "<u4"), ("d", "u1")])
    for x in range(n):
        r = np.frombuffer(buf[pos:pos + 17 * bp], rec)
        pos += 17 * bp
        bp_d[x], bp_s[x, :, 0], bp_s[x, :, 1] = r["d"], r["s0"], r["s1"]
        sz = int(np.frombuffer(buf[pos:pos + 4], np.uint32)[0])
        pos += 4
        L = np.frombuffer(buf[pos:pos + 5 * sz], lab)
        pos += 5 * sz
        hubs.append(L["h"])
        dists.append(L["d"])
        off.append(off[-1] + sz)
    assert pos == len(buf)
    return bp_d, bp_s, np.array(off, np.uint64), np.concatenate(hubs), np.concatenate(dists)


@pytest.mark.parametrize("tag", ["edgelist_in", "input_dat", "rmat10_ef8"])
def test_reference_written_index_file_equals_oracle(tag):
    """Index files written by the reference's own writer (committed bytes)
    hold, per original id, the oracle's bit-parallel records and canonical
    labels of rank[id] (hubs in rank space)."""
    g = np.load(os.path.join(GOLDEN, f"unp_{tag}.npz"))
    bp_d, bp_s, off, hub, dist = _parse_unp(g["index"].tobytes())
    n, off0, adj0 = oracle.graph_csr(g["edges"])
    rank = g["rank"]
    off2, adj2, _ = oracle.id_transfer(n, off0, adj0, rank)
    X = oracle.Index(n, off2, adj2)
    assert np.array_equal(bp_d, X.bp_d[rank]) and np.array_equal(bp_s, X.bp_s[rank])
    for x in range(n):
        v = int(rank[x])
        s, e = int(off[x]), int(off[x + 1])
        assert np.array_equal(hub[s:e], X.hub[int(X.loff[v]):int(X.loff[v + 1])]), x
        assert np.array_equal(dist[s:e], X.dist[int(X.loff[v]):int(X.loff[v + 1])]), x
