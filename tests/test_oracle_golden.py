"""The CPU oracle against the reference's own outputs (golden fixtures written
by tests/golden/make_golden.py from the reference build)."""
import numpy as np
import pytest

import oracle
from helpers import (RMAT_SMALL, TINY, bp_digest, digest, have_rmat, label_digest, load_rmat, load_tiny,
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
