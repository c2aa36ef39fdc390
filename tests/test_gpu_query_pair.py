"""One-pair queries (pado_gpu_query_pair, the drop-in's query_distance(a, b))
equal the batched query for every index kind and flag, including ids out of
range; the per-pair cost is printed."""
import time

import numpy as np
import pytest

import parallel_distance_oracle_amd as pado
from helpers import rank_space, rmat_edges

pytestmark = pytest.mark.gpu


def _pairs(n, k, seed):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, n, k).astype(np.uint32)
    b = rng.integers(0, n, k).astype(np.uint32)
    a[:3] = [0, n, n + 5]  # ids out of range
    b[:3] = [n + 1, n, 1]
    return a, b


def _check(idx, n, kw_list, seed):
    a, b = _pairs(n, 300, seed)
    for kw in kw_list:
        exp = idx.query(a, b, **kw)
        got = np.array([idx.query_pair(x, y, **kw) for x, y in zip(a, b)], dtype=exp.dtype)
        assert np.array_equal(got, exp), kw
    t = time.perf_counter()
    for x, y in zip(a[:200], b[:200]):
        idx.query_pair(x, y)
    return (time.perf_counter() - t) / 200 * 1e6


def test_query_pair_unweighted(gpu_ctx):
    G, rank, _ = rank_space(rmat_edges(12, 8))
    idx = pado.Index.build(gpu_ctx, pado.DeviceGraph(gpu_ctx, G))
    idx.order(rank)
    us = _check(idx, G.num_v, [{}, {"ref_compat": True}, {"rank_space": True}], 3)
    print(f"[gpu] query_pair: {us:.1f} us per pair (Python loop)", flush=True)


def test_query_pair_wide_and_weighted(gpu_ctx):
    G, rank, _ = rank_space(pado.grid_edges(150, 0.0, 1))
    Y = pado.Index.build(gpu_ctx, pado.DeviceGraph(gpu_ctx, G), wide=True)
    Y.order(rank)
    _check(Y, G.num_v, [{}], 4)
    e = rmat_edges(11, 8)
    w = pado.edge_weights(e.shape[0], 1, 100, seed=2)
    W = pado.WeightedGraph.from_edges(e, w)
    wr = W.make_rank()
    W.id_transfer(wr)
    X = pado.Index.build(gpu_ctx, pado.DeviceGraph(gpu_ctx, W))
    X.order(wr)
    _check(X, W.num_v, [{}, {"ref_compat": True}], 5)
