"""Weighted index on the GPU (WeightedParaVertexCentricPLL<512>, includes/
pado_weighted_unv_para.h) through the C ABI: labels bit-exact against the
reference's label dumps (tests/golden/weighted_*.npz), answers against the
reference's query_distance (REF_COMPAT) and against Dijkstra (exact)."""
import os

import numpy as np
import pytest

import oracle
import parallel_distance_oracle_amd as pado
from helpers import GOLDEN

pytestmark = pytest.mark.gpu

TAGS = ("weighted_edgelist_in", "7_w_edgelist_in", "1_w_edgelist_in", "triangle_in", "rectangle_in",
        "rmat10_ef8", "rmat12_ef8")


def _fixture(tag):
    return dict(np.load(os.path.join(GOLDEN, f"weighted_{tag}.npz")))


def _build(ctx, edges, weights, batch=512):
    G = pado.WeightedGraph.from_edges(edges, weights)
    rank = G.make_rank()
    G.id_transfer(rank)
    dg = pado.DeviceGraph(ctx, G)
    X = pado.Index.build(ctx, dg, batch_size=batch)
    return G, rank, X


@pytest.mark.parametrize("tag", TAGS)
def test_weighted_labels_bit_exact(gpu_ctx, tag):
    g = _fixture(tag)
    print("weighted build", tag, flush=True)
    G, rank, X = _build(gpu_ctx, g["edges"], g["weights"])
    assert np.array_equal(rank, g["rank"])
    assert X.weighted
    loff, hub, dist = X.wexport()
    assert np.array_equal(loff, g["loff"])
    assert np.array_equal(hub, g["hub"])
    assert np.array_equal(dist.astype(np.int64), g["dist"].astype(np.int64))
    X.order(rank)
    got = X.query(g["qa"], g["qb"], ref_compat=True)
    assert np.array_equal(got.astype(np.int64), g["answers"].astype(np.int64))


@pytest.mark.parametrize("batch", [1, 7, 64, 512])
def test_weighted_batch_size_invariant(gpu_ctx, batch):
    g = _fixture("rmat10_ef8")
    print("weighted batch", batch, flush=True)
    _, _, X = _build(gpu_ctx, g["edges"], g["weights"], batch=batch)
    loff, hub, dist = X.wexport()
    assert np.array_equal(loff, g["loff"]) and np.array_equal(hub, g["hub"])
    assert np.array_equal(dist.astype(np.int64), g["dist"].astype(np.int64))


@pytest.mark.parametrize("tag", ["rmat10_ef8", "rmat12_ef8", "weighted_edgelist_in"])
def test_weighted_exact_answers_equal_dijkstra(gpu_ctx, tag):
    g = _fixture(tag)
    G, rank, X = _build(gpu_ctx, g["edges"], g["weights"])
    n = G.num_v
    srcs = list(range(0, n, max(1, n // 16)))
    for s in srcs:
        d = oracle.dijkstra(n, G.off, G.adj, G.w, s)
        got = X.query(np.full(n, s, np.uint32), np.arange(n, dtype=np.uint32), rank_space=True)
        assert np.array_equal(got, d), s


@pytest.mark.parametrize("scale", [14, 16])
def test_weighted_rmat_matches_oracle(gpu_ctx, scale):
    """R-MAT 14 / 16 (edge factor 8, weights 1..100): many 512-root batches and
    label-pool growth; the oracle (single-thread WPARA schedule, pinned to the
    reference on the smaller fixtures) is the checker here."""
    print("weighted rmat", scale, flush=True)
    e = pado.rmat_edges(scale, 8, 0.57, 0.19, 0.19, 1)
    w = pado.edge_weights(e.shape[0], 1, 100, seed=1)
    G, rank, X = _build(gpu_ctx, e, w)
    W = oracle.WeightedIndex(G.num_v, G.off, G.adj, G.w)
    loff, hub, dist = X.wexport()
    assert np.array_equal(loff, W.loff) and np.array_equal(hub, W.hub)
    assert np.array_equal(dist.astype(np.int64), W.dist.astype(np.int64))
    qa, qb = pado.uniform_pairs(G.num_v, 50000, seed=3)
    X.order(rank)
    got = X.query(qa, qb, ref_compat=True)
    assert np.array_equal(got.astype(np.int64), W.query(rank[qa], rank[qb]).astype(np.int64))


def test_weighted_errors(gpu_ctx):
    e = np.array([[0, 1], [1, 2]], np.uint32)
    G = pado.WeightedGraph.from_edges(e, np.array([1, 0]))
    G.id_transfer(G.make_rank())
    with pytest.raises(pado.PadoError):
        pado.DeviceGraph(gpu_ctx, G)
    G = pado.WeightedGraph.from_edges(e, np.array([3, 4]))
    rank = G.make_rank()
    G.id_transfer(rank)
    X = pado.Index.build(gpu_ctx, pado.DeviceGraph(gpu_ctx, G))
    with pytest.raises(pado.PadoError):
        X.export()  # the unweighted export refuses a weighted index
    with pytest.raises(pado.PadoError):
        X.query([0], [1])  # original ids before order()
    a, c = int(rank[0]), int(rank[2])  # original 0 -(3)- 1 -(4)- 2
    assert X.query([a, a, 5], [c, a, 1], rank_space=True).tolist() == [7, 0, pado.WINF]
    assert X.query([5], [1], rank_space=True, ref_compat=True).tolist() == [32767]
