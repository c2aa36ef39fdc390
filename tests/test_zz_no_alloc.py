"""pado_gpu_query_noalloc (the allocation-free variant of pado_gpu_query on host
arrays, staged through the library's fixed device array) called through the
C ABI: its answers equal pado_gpu_query's bit for bit, on two calls in a row
(the second smaller, so stale staging contents would show), in exact and
reference-compatible mode, and sizes past the maximum are refused."""
import ctypes
import sys

import numpy as np
import pytest

import parallel_distance_oracle_amd as pado
from helpers import rank_space, rmat_edges

pytestmark = pytest.mark.gpu

MAX_PAIRS = 1 << 20  # PADO_GPU_NOALLOC_MAX_PAIRS


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _query(L, ctx, idx, a, b, flags):
    out = np.full(a.size, 7, np.uint8)
    st = L.pado_gpu_query(ctx.h, idx.h, _p(a), _p(b), a.size, _p(out), flags, None)
    assert st == 0, L.pado_gpu_last_error()
    return out


def _noalloc(L, ctx, idx, a, b, flags):
    out = np.full(a.size, 7, np.uint8)
    st = L.pado_gpu_query_noalloc(ctx.h, idx.h, _p(a), _p(b), a.size, _p(out), flags)
    assert st == 0, L.pado_gpu_last_error()
    return out


def test_noalloc_query_equals_query(gpu_ctx):
    L = pado.lib()
    G, rank, r2i = rank_space(rmat_edges(14, 8))
    print(f"[gpu] noalloc index n={G.num_v}", flush=True)
    sys.stdout.flush()
    dg = pado.DeviceGraph(gpu_ctx, G)
    idx = pado.Index.build(gpu_ctx, dg)
    idx.order(rank)
    for flags in (0, pado.QUERY_REF_COMPAT):
        # Full size, then a smaller call over different pairs.
        a1, b1 = pado.uniform_pairs(G.num_v, MAX_PAIRS, seed=5)
        half = MAX_PAIRS // 2
        z = pado.zipf_vertices(G.num_v, 2 * half, 1.1, 9, r2i)
        a1[:half], b1[:half] = z[:half], z[half:]
        a2, b2 = pado.uniform_pairs(G.num_v, 12345, seed=6)
        a2[:10] = np.arange(10, dtype=np.uint32) + G.num_v  # out-of-range ids
        got1 = _noalloc(L, gpu_ctx, idx, a1, b1, flags)
        got2 = _noalloc(L, gpu_ctx, idx, a2, b2, flags)
        assert np.array_equal(got1, _query(L, gpu_ctx, idx, a1, b1, flags))
        assert np.array_equal(got2, _query(L, gpu_ctx, idx, a2, b2, flags))
    # Rank-space ids too.
    a3, b3 = pado.uniform_pairs(G.num_v, 4096, seed=7)
    fl = pado.QUERY_RANK_SPACE
    assert np.array_equal(_noalloc(L, gpu_ctx, idx, a3, b3, fl), _query(L, gpu_ctx, idx, a3, b3, fl))
    # Zero pairs is a no-op; one past the maximum is refused with ERR_ARG.
    e = np.zeros(0, np.uint32)
    assert L.pado_gpu_query_noalloc(gpu_ctx.h, idx.h, None, None, 0, None, 0) == 0
    big = np.zeros(MAX_PAIRS + 1, np.uint32)
    out = np.zeros(MAX_PAIRS + 1, np.uint8)
    assert L.pado_gpu_query_noalloc(gpu_ctx.h, idx.h, _p(big), _p(big), big.size, _p(out), 0) == 1
    assert b"maximum" in L.pado_gpu_last_error()
    assert e.size == 0
    idx.close()
    dg.close()
