"""Device ingest (Graph / WeightedGraph construct + make_rank + id_transfer on
the GPU, includes/graph.h:120-206, 313-408) produces the same rank-space CSR,
rank and rank2id as the host path, which is pinned to the reference's rank
and the oracle's CSR by the CPU tests."""
import numpy as np
import pytest

import parallel_distance_oracle_amd as pado
from helpers import load_rmat, load_tiny, rmat_edges

pytestmark = pytest.mark.gpu


def _host(edges, weights=None):
    if weights is None:
        G = pado.Graph.from_edges(edges)
        rank = G.make_rank()
        r2i = G.id_transfer(rank)
        return G.off, G.adj, None, rank, r2i
    G = pado.WeightedGraph.from_edges(edges, weights)
    rank = G.make_rank()
    r2i = G.id_transfer(rank)
    return G.off, G.adj, G.w, rank, r2i


@pytest.mark.parametrize("case", ["edgelist_in", "dup_and_loops", "rmat12", "rmat16"])
def test_device_ingest_equals_host(gpu_ctx, case):
    if case == "edgelist_in":
        e = load_tiny(case)["edges"]
    elif case == "dup_and_loops":
        rng = np.random.default_rng(9)
        e = rng.integers(0, 500, size=(3000, 2)).astype(np.uint32)
        e = np.vstack([e, e[:100], [[3, 3]], [[499, 2000]]]).astype(np.uint32)
    else:
        e = rmat_edges(int(case[4:]), 8)
    print(f"[gpu] ingest {case} m={e.shape[0]}", flush=True)
    off, adj, _, rank, r2i = _host(e)
    dg, rank_d, r2i_d = pado.DeviceGraph.from_edges(gpu_ctx, e)
    off_d, adj_d, w_d = dg.export()
    assert np.array_equal(rank_d, rank) and np.array_equal(r2i_d, r2i)
    assert np.array_equal(off_d, off) and np.array_equal(adj_d, adj) and w_d is None
    dg.close()


@pytest.mark.parametrize("scale", [10, 14])
def test_device_ingest_weighted_equals_host(gpu_ctx, scale):
    e = rmat_edges(scale, 8)
    w = pado.edge_weights(e.shape[0], 1, 100, seed=1)
    # parallel edges with different weights exercise the (id, weight) order
    e = np.vstack([e, e[:50]]).astype(np.uint32)
    w = np.concatenate([w, (w[:50] % 7) + 1]).astype(np.uint32)
    off, adj, ww, rank, r2i = _host(e, w)
    dg, rank_d, r2i_d = pado.DeviceGraph.from_edges(gpu_ctx, e, weights=w)
    off_d, adj_d, w_d = dg.export()
    assert np.array_equal(rank_d, rank) and np.array_equal(r2i_d, r2i)
    assert np.array_equal(off_d, off) and np.array_equal(adj_d, adj) and np.array_equal(w_d, ww)
    dg.close()


def test_device_ingest_builds_reference_labels(gpu_ctx):
    """Device-ingested R-MAT 12 builds the reference's labels."""
    meta, arr = load_rmat(12, 8)
    dg, rank, _ = pado.DeviceGraph.from_edges(gpu_ctx, rmat_edges(12, 8))
    idx = pado.Index.build(gpu_ctx, dg)
    _, _, loff, hub, dist = idx.export()
    assert np.array_equal(loff, arr["loff"]) and np.array_equal(hub, arr["hub"]) and np.array_equal(dist, arr["dist"])
    idx.close()
    dg.close()


def test_device_ingest_given_rank_and_errors(gpu_ctx):
    e = rmat_edges(10, 8)
    off, adj, _, rank, _ = _host(e)
    dg, rank_d, _ = pado.DeviceGraph.from_edges(gpu_ctx, e, rank=rank)
    off_d, adj_d, _ = dg.export()
    assert np.array_equal(rank_d, rank) and np.array_equal(off_d, off) and np.array_equal(adj_d, adj)
    dg.close()
    bad = rank.copy()
    bad[0] = bad[1]
    with pytest.raises(pado.PadoError):
        pado.DeviceGraph.from_edges(gpu_ctx, e, rank=bad)
    with pytest.raises(pado.PadoError):
        pado.DeviceGraph.from_edges(gpu_ctx, e[:3], weights=np.array([1, 0, 2]))
