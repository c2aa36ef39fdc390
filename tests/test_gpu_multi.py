"""Many graphs labeled in one job: every graph's bit-parallel records and label
sets equal the oracle's build of that graph alone; a graph outside the
reference's u8 distance domain is labeled in the wide hop domain, with
answers equal to BFS."""
import sys

import numpy as np
import pytest

import oracle
import parallel_distance_oracle_amd as pado
from helpers import rank_space

pytestmark = pytest.mark.gpu


def _graphs(count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        if i % 2 == 0:
            scale = int(rng.integers(5, 11))
            e = pado.rmat_edges(scale, 8, 0.57, 0.19, 0.19, seed + i)
        else:
            side = int(rng.integers(4, 40))
            e = pado.grid_edges(side, 0.05, seed + i)
        if e.shape[0] == 0:
            e = np.array([[0, 1]], np.uint32)
        G, _, _ = rank_space(e)
        out.append(G)
    return out


def test_multi_graph_job_matches_single_builds(gpu_ctx):
    graphs = _graphs(40, seed=3)
    # One graph far outside the u8 domain (a 300 x 300 grid: diameter ~598).
    big, _, _ = rank_space(pado.grid_edges(300, 0.0, 1))
    graphs.append(big)
    print(f"[gpu] multi build of {len(graphs)} graphs", flush=True)
    sys.stdout.flush()
    mi = pado.MultiIndex(gpu_ctx, graphs)
    info = mi.info()
    assert info["unions"] >= 1
    for gi, G in enumerate(graphs[:-1]):
        gi_info = mi.graph_info(gi)
        assert gi_info["status"] == 0, gi
        bp_d, bp_s, loff, hub, dist = mi.export(gi)
        X = oracle.Index(G.num_v, G.off, G.adj)
        assert np.array_equal(bp_d, X.bp_d) and np.array_equal(bp_s, X.bp_s), gi
        assert np.array_equal(loff, X.loff) and np.array_equal(hub, X.hub) and np.array_equal(dist, X.dist), gi
    # The deep grid is labeled on its own in the wide hop domain: u16 label
    # distances, answers equal to BFS (beyond 255 too).
    g = len(graphs) - 1
    last = mi.graph_info(g)
    assert last["status"] == 0 and not last["in_union"] and last["wide"]
    _, _, loff, hub, dist = mi.export(g)
    assert dist.dtype == np.uint16 and hub.size == last["num_labels"]
    n = big.num_v
    for s in (0, n // 2, n - 1):
        d = oracle.bfs(n, big.off, big.adj, s).astype(np.int64)
        got = mi.query(g, np.full(n, s, np.uint32), np.arange(n, dtype=np.uint32)).astype(np.int64)
        assert np.array_equal(got, np.where(d == 0xFFFFFFFF, pado.WIDE_INF, d)), s
        assert d.max() > 255
    # Union graphs answer through the union's index, as their own build would.
    G = graphs[0]
    got = mi.query(0, np.zeros(G.num_v, np.uint32), np.arange(G.num_v, dtype=np.uint32))
    assert np.array_equal(got, oracle.bfs_to_u8(oracle.bfs(G.num_v, G.off, G.adj, 0)))
    assert mi.query(0, [G.num_v, G.num_v, 0], [G.num_v, G.num_v + 3, G.num_v]).tolist() == [0, 255, 255]
    mi.close()
