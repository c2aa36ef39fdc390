"""Resident serving (one kernel that stays on the device while requests flow):
answers equal the plain batch query alone and in a mix, for u8, wide (u16)
and weighted (u32) indexes, exact and reference-compatible; admission, ring
wrap, cancel and timed waits behave as in the launch-per-batch mode; the
kernel leaves when idle and is relaunched for new work."""
import sys
import threading
import time

import numpy as np
import pytest

import parallel_distance_oracle_amd as pado
from helpers import rank_space, rmat_edges

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def served(gpu_ctx):
    G, rank, r2i = rank_space(rmat_edges(14, 8))
    print(f"[gpu] resident serve index n={G.num_v}", flush=True)
    sys.stdout.flush()
    idx = pado.Index.build(gpu_ctx, pado.DeviceGraph(gpu_ctx, G))
    idx.order(rank)
    return G, idx, r2i


def _requests(n, r2i, count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        k = int(np.floor(np.exp(rng.uniform(0, np.log(65536))))) if i % 7 else 0
        half = k // 2
        a = np.concatenate([rng.integers(0, n, half), pado.zipf_vertices(n, k - half, 1.1, seed + i, r2i)])
        b = np.concatenate([rng.integers(0, n, half), pado.zipf_vertices(n, k - half, 1.1, seed + 7 * i, r2i)])
        out.append((a.astype(np.uint32), b.astype(np.uint32)))
    return out


def _submit_retry(srv, a, b):
    for _ in range(20000):
        try:
            return srv.submit(a, b)
        except pado.Busy:
            time.sleep(0.0005)
    raise AssertionError("admission never freed up")


@pytest.mark.parametrize("ref", [False, True])
def test_resident_alone_equals_mix_equals_batch(gpu_ctx, served, ref):
    G, idx, r2i = served
    reqs = _requests(G.num_v, r2i, 50, seed=11 + ref)
    print(f"[gpu] resident alone/mix ref_compat={ref}", flush=True)
    srv = pado.Server(gpu_ctx, idx, staging_pairs=1 << 20, resident=True, ref_compat=ref)
    try:
        alone = []
        for a, b in reqs:
            rid = srv.submit(a, b)
            alone.append(srv.wait(rid, a.size)[0])
        mix = [None] * len(reqs)

        def client(indices):
            for i in indices:
                a, b = reqs[i]
                rid = _submit_retry(srv, a, b)
                mix[i] = srv.wait(rid, a.size)[0]

        th = [threading.Thread(target=client, args=(range(j, len(reqs), 3),)) for j in range(3)]
        for t in th:
            t.start()
        for t in th:
            t.join(timeout=120)
        st = srv.stats()
    finally:
        srv.stop()
    for (a, b), x, y in zip(reqs, alone, mix):
        assert y is not None
        assert np.array_equal(x, y)
        assert np.array_equal(x, idx.query(a, b, ref_compat=ref))
    assert st["requests_done"] == 2 * len(reqs)


def test_resident_admission_and_ring_wrap(gpu_ctx, served):
    G, idx, r2i = served
    srv = pado.Server(gpu_ctx, idx, staging_pairs=4096, max_requests=4, resident=True)
    try:
        a = np.arange(3000, dtype=np.uint32) % G.num_v
        r1 = srv.submit(a, a)
        with pytest.raises(pado.Busy):
            srv.submit(a, a)
        out, lat = srv.wait(r1, a.size)
        assert np.all(out == 0) and lat >= 0
        big = np.arange(4096, dtype=np.uint32) % G.num_v
        r3 = srv.submit(big[:3001], big[:3001])
        assert np.all(srv.wait(r3, 3001)[0] == 0)
        r4 = srv.submit(big, big)
        assert np.all(srv.wait(r4, 4096)[0] == 0)
        ids = [srv.submit(a[:10], a[:10]) for _ in range(4)]
        with pytest.raises(pado.Busy):
            srv.submit(a[:10], a[:10])
        for r in ids:
            srv.wait(r, 10)
    finally:
        srv.stop()
    # Requests that end exactly at the ring's end, back to back.
    rng = np.random.default_rng(41)
    cap, k, count = 4096, 256, 300
    A = rng.integers(0, G.num_v, count * k).astype(np.uint32)
    B = pado.zipf_vertices(G.num_v, count * k, 1.1, 41, r2i).astype(np.uint32)
    off = (np.arange(count + 1) * k).astype(np.uint64)
    srv = pado.Server(gpu_ctx, idx, staging_pairs=cap, resident=True)
    try:
        lat, ans, _ = srv.replay(A, B, off, np.zeros(count))
        st = srv.stats()
    finally:
        srv.stop()
    assert np.array_equal(ans, idx.query(A, B))
    assert st["requests_done"] == count and np.all(lat >= 0)


def test_resident_idle_exit_and_relaunch(gpu_ctx, served):
    """After 20 ms without work the kernel leaves; the next request relaunches
    it and is answered."""
    G, idx, r2i = served
    reqs = _requests(G.num_v, r2i, 6, seed=19)
    srv = pado.Server(gpu_ctx, idx, staging_pairs=1 << 18, resident=True)
    try:
        for a, b in reqs:
            rid = srv.submit(a, b)
            assert np.array_equal(srv.wait(rid, a.size)[0], idx.query(a, b))
            time.sleep(0.06)
        launches = srv.stats()["launches"]
    finally:
        srv.stop()
    assert launches >= len([1 for a, _ in reqs if a.size]) - 1


def test_resident_cancel_and_timed_wait(gpu_ctx, served):
    G, idx, _ = served
    srv = pado.Server(gpu_ctx, idx, staging_pairs=4096, max_requests=8, resident=True)
    try:
        a = (np.arange(4096, dtype=np.uint32) * 7) % G.num_v
        b = (np.arange(4096, dtype=np.uint32) * 13) % G.num_v
        exp = idx.query(a, b)
        for _ in range(5):
            r = _submit_retry(srv, a[:3000], b[:3000])
            srv.cancel(r)
            with pytest.raises(pado.PadoError):
                srv.wait(r, 3000)
        r = _submit_retry(srv, a, b)
        assert np.array_equal(srv.wait(r, 4096)[0], exp)
        r = _submit_retry(srv, a, b)
        got = None
        for _ in range(200000):
            try:
                got = srv.wait(r, 4096, timeout_us=0)[0]
                break
            except pado.Timeout:
                pass
        assert got is not None and np.array_equal(got, exp)
    finally:
        srv.stop()


def test_resident_weighted_and_wide(gpu_ctx):
    e = rmat_edges(12, 8)
    w = pado.edge_weights(e.shape[0], 1, 100, seed=1)
    WG = pado.WeightedGraph.from_edges(e, w)
    rank = WG.make_rank()
    r2i = WG.id_transfer(rank)
    print(f"[gpu] resident weighted n={WG.num_v}", flush=True)
    X = pado.Index.build(gpu_ctx, pado.DeviceGraph(gpu_ctx, WG))
    X.order(rank)
    reqs = _requests(WG.num_v, r2i, 16, seed=5)
    srv = pado.Server(gpu_ctx, X, staging_pairs=1 << 20, resident=True)
    try:
        ids = [srv.submit(a, b) for a, b in reqs]
        outs = [srv.wait(r, a.size)[0] for r, (a, _) in zip(ids, reqs)]
    finally:
        srv.stop()
    for (a, b), o in zip(reqs, outs):
        assert o.dtype == np.uint32 and np.array_equal(o, X.query(a, b))
    # A wide-domain index (a 200 x 200 grid): u16 exact hops.
    G, grank, gr2i = rank_space(pado.grid_edges(200, 0.0, 1))
    print(f"[gpu] resident wide n={G.num_v}", flush=True)
    Y = pado.Index.build(gpu_ctx, pado.DeviceGraph(gpu_ctx, G), wide=True)
    Y.order(grank)
    reqs = _requests(G.num_v, gr2i, 10, seed=8)
    srv = pado.Server(gpu_ctx, Y, staging_pairs=1 << 20, resident=True)
    try:
        ids = [srv.submit(a, b) for a, b in reqs]
        outs = [srv.wait(r, a.size)[0] for r, (a, _) in zip(ids, reqs)]
    finally:
        srv.stop()
    for (a, b), o in zip(reqs, outs):
        assert o.dtype == np.uint16 and np.array_equal(o, Y.query(a, b))


def test_resident_replay(gpu_ctx, served):
    G, idx, r2i = served
    reqs = _requests(G.num_v, r2i, 300, seed=31)
    A = np.concatenate([a for a, _ in reqs]).astype(np.uint32)
    B = np.concatenate([b for _, b in reqs]).astype(np.uint32)
    off = np.concatenate([[0], np.cumsum([a.size for a, _ in reqs])]).astype(np.uint64)
    arrivals = np.cumsum(np.random.default_rng(3).exponential(2e-4, size=len(reqs)))
    srv = pado.Server(gpu_ctx, idx, staging_pairs=1 << 21, resident=True)
    try:
        lat, ans, wall = srv.replay(A, B, off, arrivals)
        st = srv.stats()
    finally:
        srv.stop()
    assert np.array_equal(ans, idx.query(A, B))
    assert np.all(lat >= 0) and st["requests_done"] == len(reqs)
