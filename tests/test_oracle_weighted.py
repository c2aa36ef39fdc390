"""The weighted oracle (WeightedParaVertexCentricPLL<512> restated) against the
reference's own outputs, Dijkstra, and the canonical form of pruned landmark
labels: (r, d(v,r)) is a label of v iff r ranks at or above v and no vertex
ranked above r lies on any shortest v-r path."""
import os

import numpy as np
import pytest

import oracle
from helpers import GOLDEN

WEIGHTED = ("weighted_edgelist_in", "7_w_edgelist_in", "1_w_edgelist_in", "triangle_in", "rectangle_in",
            "rmat10_ef8", "rmat12_ef8")


def _load(tag):
    return dict(np.load(os.path.join(GOLDEN, f"weighted_{tag}.npz")))


def _rank_space(g):
    n, off, adj, w = oracle.weighted_graph_csr(g["edges"], g["weights"])
    off2, adj2, w2, r2i = oracle.weighted_id_transfer(n, off, adj, w, g["rank"])
    return n, off2, adj2, w2


@pytest.mark.parametrize("tag", WEIGHTED)
def test_weighted_labels_and_answers_equal_reference(tag):
    g = _load(tag)
    n, off, adj, w = _rank_space(g)
    X = oracle.WeightedIndex(n, off, adj, w)
    assert np.array_equal(X.loff, g["loff"])
    assert np.array_equal(X.hub, g["hub"]) and np.array_equal(X.dist, g["dist"])
    rank = g["rank"]
    assert np.array_equal(X.query(rank[g["qa"]], rank[g["qb"]]), g["answers"])


@pytest.mark.parametrize("tag", WEIGHTED)
def test_weighted_answers_equal_dijkstra(tag):
    g = _load(tag)
    n, off, adj, w = _rank_space(g)
    X = oracle.WeightedIndex(n, off, adj, w)
    for s in range(0, n, max(1, n // 25)):
        d = oracle.dijkstra(n, off, adj, w, s).astype(np.int64)
        exp = np.where(d == 0xFFFFFFFF, 32767, d)
        exp = np.where(exp >= 32765, 32767, exp)
        got = X.query(np.full(n, s, np.uint32), np.arange(n, dtype=np.uint32)).astype(np.int64)
        assert np.array_equal(got, exp), s


def canonical_labels(n, off, adj, w):
    """Canonical pruned landmark labels from all-pairs Dijkstra (small n)."""
    inf = 1 << 40
    D = np.zeros((n, n), np.int64)
    for s in range(n):
        d = oracle.dijkstra(n, off, adj, w, s).astype(np.int64)
        d[d == 0xFFFFFFFF] = inf
        D[s] = d
    hubs = [[] for _ in range(n)]
    for r in range(n):
        via = (D[:, :r] + D[:r, r][None, :]).min(axis=1) if r else np.full(n, 2 * inf)
        ok = (np.arange(n) >= r) & (D[:, r] < inf) & (via > D[:, r])
        for v in np.nonzero(ok)[0]:
            hubs[v].append((r, int(D[v, r])))
    return hubs


@pytest.mark.parametrize("tag", ["weighted_edgelist_in", "triangle_in", "rmat10_ef8"])
def test_weighted_labels_are_canonical(tag):
    g = _load(tag)
    n, off, adj, w = _rank_space(g)
    X = oracle.WeightedIndex(n, off, adj, w)
    C = canonical_labels(n, off, adj, w)
    for v in range(n):
        got = list(zip(X.hub[X.loff[v]:X.loff[v + 1]].tolist(), X.dist[X.loff[v]:X.loff[v + 1]].tolist()))
        assert got == C[v], v


@pytest.mark.parametrize("tag", ["weighted_edgelist_in", "7_w_edgelist_in", "triangle_in", "rectangle_in",
                                 "rmat10_ef8", "rmat12_ef8"])
def test_canonical_of_one_vertex(tag):
    """The one-Dijkstra canonical label set of a vertex (the checker of the
    full-size weighted GPU test) equals the reference's labels of that vertex
    on every vertex of the reference's fixtures, and its distances equal
    Dijkstra's."""
    g = _load(tag)
    n, off, adj, w = _rank_space(g)
    step = max(1, n // 60)
    for v in range(0, n, step):
        d, hub, hd = oracle.canonical_of(n, off, adj, w, v)
        assert np.array_equal(d, oracle.dijkstra(n, off, adj, w, v))
        s, e = int(g["loff"][v]), int(g["loff"][v + 1])
        assert np.array_equal(hub, g["hub"][s:e]), v
        assert np.array_equal(hd.astype(np.int64), g["dist"][s:e].astype(np.int64)), v
