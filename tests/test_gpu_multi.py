"""Many graphs labeled in one job: every graph's bit-parallel records and label
sets equal the oracle's build of that graph alone; graphs outside the
reference's u8 distance domain report PADO_GPU_ERR_DOMAIN."""
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
    last = mi.graph_info(len(graphs) - 1)
    assert last["status"] == 4 and not last["in_union"]
    with pytest.raises(pado.PadoError):
        mi.export(len(graphs) - 1)
    mi.close()
