"""The batch phase run by one workgroup per graph (PADO_GPU_BUILD_DEVICE_LOOP)
gives the labels, bit-parallel records and set counters of the launch
schedule, in the reference's u8 domain and in the wide hop domain."""
import time

import numpy as np
import pytest

import parallel_distance_oracle_amd as pado
from helpers import rank_space

pytestmark = pytest.mark.gpu

SET_STATS = ("iterations", "pushes", "ind_pass", "bp_pruned", "cands", "dq_pruned", "num_labels")


def _build(ctx, dg, flags):
    cfg = pado.BuildCfg()
    cfg.flags = flags
    ctx.sync()
    t = time.perf_counter()
    idx = pado.Index.build(ctx, dg, cfg=cfg)
    ctx.sync()
    dt = time.perf_counter() - t
    st = idx.stats()
    out = idx.export(), {k: st[k] for k in SET_STATS}, dt
    idx.close()
    return out


@pytest.mark.parametrize("case", ["rmat10", "rmat13", "grid40", "grid_deep150", "grid_deep240_holes"])
def test_device_loop_equals_launch_schedule(gpu_ctx, case):
    if case.startswith("rmat"):
        e = pado.rmat_edges(int(case[4:]), 8, 0.57, 0.19, 0.19, 5)
        wide = False
    elif case == "grid40":
        e = pado.grid_edges(40, 0.05, 2)
        wide = False
    elif case == "grid_deep150":
        e = pado.grid_edges(150, 0.0, 3)
        wide = True
    else:
        e = pado.grid_edges(240, 0.05, 4)
        wide = True
    G, _, _ = rank_space(e)
    dg = pado.DeviceGraph(gpu_ctx, G)
    base = pado.BUILD_WIDE if wide else 0
    print(f"[gpu] device loop {case} n={G.num_v}", flush=True)
    ref, ref_st, t_ref = _build(gpu_ctx, dg, base)
    got, st, t_loop = _build(gpu_ctx, dg, base | pado.BUILD_DEVICE_LOOP)
    print(f"[gpu]   launch schedule {t_ref:.3f} s, device loop {t_loop:.3f} s, iterations {st['iterations']}",
          flush=True)
    for x, y in zip(ref, got):
        assert np.array_equal(x, y)
    assert st == ref_st
    dg.close()
