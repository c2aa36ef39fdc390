"""Pins of the oracle that do not use the reference at all: closed forms of the
bit-parallel labels from plain BFS, the 2-hop cover property, and label
invariants of pruned landmark labeling."""
import numpy as np
import pytest

import oracle
from helpers import RMAT_SMALL, TINY, load_tiny, oracle_rank_space, rmat_edges

INF = 0xFFFFFFFF


def _graphs():
    for name in TINY:
        g = load_tiny(name)
        yield name, g["edges"], g["rank"]
    for scale, ef in RMAT_SMALL[:2]:
        e = rmat_edges(scale, ef)
        n, off0, _ = oracle.graph_csr(e)
        yield f"rmat{scale}", e, oracle.make_rank(n, off0)


def _select_bp_roots(n, off, adj):
    """Root/neighbour selection written from PARA:670-709: slot root = first
    unused vertex in rank order; then up to 64 unused neighbours taken from
    the END of its descending adjacency list."""
    used = np.zeros(n, bool)
    slots = []
    r = 0
    for _ in range(50):
        while r < n and used[r]:
            r += 1
        if r == n:
            slots.append(None)
            continue
        used[r] = True
        sel = []
        for v in adj[int(off[r]):int(off[r + 1])][::-1]:
            if not used[v]:
                used[v] = True
                sel.append(int(v))
                if len(sel) == 64:
                    break
        slots.append((r, sel))
    return slots, used


@pytest.mark.parametrize("case", list(_graphs()), ids=lambda c: c[0])
def test_bp_closed_form(case):
    """bp_dist = BFS distance from the slot root; bit j in S^-1 iff
    d(u_j, v) = d(r, v) - 1; bit j in S^0 \\ S^-1 iff d(u_j, v) = d(r, v)."""
    _, edges, rank = case
    n, off, adj, _ = oracle_rank_space(edges, rank)
    bp_d, bp_s, used, roots, _ = oracle.bp_labeling(n, off, adj)
    slots, used_py = _select_bp_roots(n, off, adj)
    assert np.array_equal(used.astype(bool), used_py)
    for i, slot in enumerate(slots):
        if slot is None:
            assert np.all(bp_d[:, i] == 255) and np.all(bp_s[:, i] == 0)
            continue
        r, sel = slot
        assert roots[i] == r
        dr = oracle.bfs(n, off, adj, r).astype(np.int64)
        exp_d = np.where(dr == INF, 255, dr)
        assert np.array_equal(bp_d[:, i], exp_d.astype(np.uint8))
        s1 = np.zeros(n, np.uint64)
        s0 = np.zeros(n, np.uint64)
        for j, u in enumerate(sel):
            du = oracle.bfs(n, off, adj, u).astype(np.int64)
            reach = dr != INF
            s1[reach & (du == dr - 1)] |= np.uint64(1 << j)
            s0[reach & (du == dr)] |= np.uint64(1 << j)
        assert np.array_equal(bp_s[:, i, 0], s1)
        assert np.array_equal(bp_s[:, i, 1], s0)


@pytest.mark.parametrize("case", list(_graphs()), ids=lambda c: c[0])
def test_two_hop_cover_equals_bfs(case):
    """Signed query over the oracle's labels == BFS on every pair."""
    _, edges, rank = case
    n, off, adj, _ = oracle_rank_space(edges, rank)
    X = oracle.Index(n, off, adj)
    m = min(n, 300)
    a, b = np.meshgrid(np.arange(m, dtype=np.uint32), np.arange(n, dtype=np.uint32), indexing="ij")
    a, b = a.reshape(-1), b.reshape(-1)
    exact = oracle.bfs_to_u8(oracle.bfs_pairs(n, off, adj, a, b))
    assert np.array_equal(X.query(a, b), exact)


@pytest.mark.parametrize("scale,ef", RMAT_SMALL)
def test_label_invariants(scale, ef):
    """Hubs rank above their vertex; non-BP vertices hold (v, 0); hubs unique
    and ascending; every label distance is the exact BFS distance."""
    e = rmat_edges(scale, ef)
    n, off0, _ = oracle.graph_csr(e)
    n, off, adj, _ = oracle_rank_space(e, oracle.make_rank(n, off0))
    X = oracle.Index(n, off, adj)
    v_of = np.repeat(np.arange(n, dtype=np.uint32), np.diff(X.loff).astype(np.int64))
    assert np.all(X.hub <= v_of)
    self_lab = X.hub == v_of
    assert np.all(X.dist[self_lab] == 0)
    has_self = np.zeros(n, bool)
    has_self[v_of[self_lab]] = True
    assert np.array_equal(has_self, X.used == 0)
    for v in range(n):
        h = X.hub[int(X.loff[v]):int(X.loff[v + 1])]
        assert np.all(np.diff(h.astype(np.int64)) > 0)
    exact = oracle.bfs_pairs(n, off, adj, X.hub, v_of)
    assert np.array_equal(X.dist.astype(np.uint32), exact)


def test_labels_independent_of_batch_size():
    """BATCH_SIZE = 1 is sequential pruned landmark labeling (one pruned BFS
    per root, in rank order).  A prune that a later batch iteration could
    see is never needed earlier (a hub h of the same batch can only prune at
    iteration k through d(v,h) + d(r,h) <= k, and both labels exist by
    iteration k-1), so every batch size yields the same label sets -- and the
    same exact distances."""
    e = rmat_edges(12, 8)
    n, off0, _ = oracle.graph_csr(e)
    n, off, adj, _ = oracle_rank_space(e, oracle.make_rank(n, off0))
    bp_d, bp_s, used, _, _ = oracle.bp_labeling(n, off, adj)
    a, b = np.meshgrid(np.arange(0, n, 7, dtype=np.uint32), np.arange(n, dtype=np.uint32), indexing="ij")
    a, b = a.reshape(-1), b.reshape(-1)
    exact = oracle.bfs_to_u8(oracle.bfs_pairs(n, off, adj, a, b))
    ref = oracle.labeling(n, off, adj, bp_d, bp_s, used, 1024)
    for bs in (1, 3, 64, 1000):
        loff, hub, dist, _ = oracle.labeling(n, off, adj, bp_d, bp_s, used, bs)
        assert np.array_equal(loff, ref[0]) and np.array_equal(hub, ref[1]) and np.array_equal(dist, ref[2])
        out = np.zeros(a.size, np.uint8)
        oracle.lib().po_query_batch(n, bp_d.reshape(-1), bp_s.reshape(-1), loff, hub, dist, a, b, a.size, out, 0)
        assert np.array_equal(out, exact)
