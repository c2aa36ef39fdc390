"""The build knobs the autotuner sweeps leave the labels unchanged, and the
configurations the tuners return build label-identical indexes (single graph
and many-graph job)."""
import numpy as np
import pytest

import parallel_distance_oracle_amd as pado
from helpers import rank_space

pytestmark = pytest.mark.gpu

SET_STATS = ("iterations", "pushes", "ind_pass", "bp_pruned", "cands", "dq_pruned", "num_labels")


def _same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


def _build(ctx, dg, **kw):
    cfg = pado.BuildCfg()
    for k, v in kw.items():
        setattr(cfg, k, v)
    idx = pado.Index.build(ctx, dg, cfg=cfg)
    out = idx.export(), {k: idx.stats()[k] for k in SET_STATS}
    idx.close()
    return out


@pytest.mark.parametrize("graph", ["rmat12", "grid_wide"])
def test_push_shapes_give_identical_labels(gpu_ctx, graph):
    """Every (descriptor size, block waves) instantiation of the push kernel,
    with small and large grids: same labels, same set counters."""
    wide = graph == "grid_wide"
    e = pado.grid_edges(150, 0.02, 5) if wide else pado.rmat_edges(12, 8, 0.57, 0.19, 0.19, 7)
    G, _, _ = rank_space(e)
    dg = pado.DeviceGraph(gpu_ctx, G)
    flags = pado.BUILD_WIDE if wide else 0
    ref, ref_st = _build(gpu_ctx, dg, flags=flags)
    for chunk, waves, grid in ((4, 4, 768), (4, 2, 64), (8, 2, 64), (8, 8, 1024), (16, 2, 4096), (32, 4, 256),
                               (32, 8, 512), (16, 8, 128)):
        print(f"[gpu] {graph} chunk={chunk} waves={waves} push_blocks={grid}", flush=True)
        got, st = _build(gpu_ctx, dg, flags=flags, push_chunk=chunk, push_waves=waves, push_blocks=grid,
                         add_blocks=grid)
        assert _same(got, ref), (chunk, waves, grid)
        assert st == ref_st, (chunk, waves, grid)
    dg.close()


def test_bad_push_shape_refused(gpu_ctx):
    G, _, _ = rank_space(pado.rmat_edges(8, 8, 0.57, 0.19, 0.19, 1))
    dg = pado.DeviceGraph(gpu_ctx, G)
    for kw in ({"push_chunk": 12}, {"push_chunk": 64}, {"push_waves": 3}, {"push_waves": 16}):
        cfg = pado.BuildCfg()
        for k, v in kw.items():
            setattr(cfg, k, v)
        with pytest.raises(pado.PadoError):
            pado.Index.build(gpu_ctx, dg, cfg=cfg)
    dg.close()


def test_autotune_result_builds_identical_labels(gpu_ctx):
    G, _, _ = rank_space(pado.rmat_edges(11, 8, 0.57, 0.19, 0.19, 2))
    dg = pado.DeviceGraph(gpu_ctx, G)
    print("[gpu] autotune R-MAT 11 (8 configurations)", flush=True)
    best, ms, trials = pado.autotune(gpu_ctx, dg, max_trials=8)
    assert 1 <= trials <= 8 and ms > 0
    assert best.push_chunk in (4, 8, 16, 32) and best.push_waves in (2, 4, 8)
    assert best.batch_size in (256, 512, 1024)
    ref, ref_st = _build(gpu_ctx, dg)
    idx = pado.Index.build(gpu_ctx, dg, cfg=best)
    assert _same(idx.export(), ref)
    idx.close()
    dg.close()


def test_autotune_multi_result_builds_identical_labels(gpu_ctx):
    graphs = pado.mixed_graphs(12, seed=9)
    print("[gpu] autotune_multi on 6 of 12 graphs (4 configurations)", flush=True)
    best, ms, trials = pado.autotune_multi(gpu_ctx, graphs, sample=6, max_trials=4)
    assert 1 <= trials <= 4 and ms > 0
    a = pado.MultiIndex(gpu_ctx, graphs)
    b = pado.MultiIndex(gpu_ctx, graphs, cfg=best)
    for g in range(len(graphs)):
        ia, ib = a.graph_info(g), b.graph_info(g)
        assert ia == ib, g
        assert _same(a.export(g), b.export(g)), g
    a.close()
    b.close()
