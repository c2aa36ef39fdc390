"""The wide hop domain (PADO_GPU_BUILD_WIDE: 12-bit label distances, u16
answers) for graphs deeper than the reference's u8 domain.

The reference has no such domain, so it is pinned two ways:
  * on graphs inside the u8 domain (every batch iteration below 253) the wide
    build's bit-parallel records and label sets equal the default build's --
    which the reference's own fixtures pin -- and its answers equal BFS;
  * on deep graphs (grids of diameter up to ~600, a long path) every answer
    from sampled sources to all targets equals BFS, every label of sampled
    hubs carries the BFS distance, the hubs rank above their vertices, and
    the bit-parallel distances equal BFS up to level 254 (255 beyond).
"""
import sys

import numpy as np
import pytest

import oracle
import parallel_distance_oracle_amd as pado
from helpers import load_rmat, load_tiny, rank_space, rmat_edges

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu]", *a, flush=True)
    sys.stdout.flush()


def _graph(kind, p, seed=1):
    if kind == "rmat":
        e = rmat_edges(p, 8) if seed == 1 else pado.rmat_edges(p, 8, 0.57, 0.19, 0.19, seed)
    elif kind == "grid":
        e = pado.grid_edges(p, 0.05, seed)
    elif kind == "path":
        v = np.arange(p - 1, dtype=np.uint32)
        e = np.stack([v, v + 1], axis=1)
    else:
        e = load_tiny(p)["edges"]
    G, rank, r2i = rank_space(e)
    return G, rank, r2i


def _exact(G, src):
    d = oracle.bfs(G.num_v, G.off, G.adj, src).astype(np.int64)
    return np.where(d == 0xFFFFFFFF, pado.WIDE_INF, d).astype(np.uint16)


@pytest.mark.parametrize("kind,p", [("tiny", "edgelist_in"), ("rmat", 12), ("rmat", 14), ("grid", 40),
                                    ("grid", 100)])
def test_wide_equals_default_inside_u8_domain(gpu_ctx, kind, p):
    G, rank, _ = _graph(kind, p)
    say(f"wide vs default {kind} {p} n={G.num_v}")
    dg = pado.DeviceGraph(gpu_ctx, G)
    X = pado.Index.build(gpu_ctx, dg)
    W = pado.Index.build(gpu_ctx, dg, wide=True)
    assert W.wide and not X.wide
    ex, ew = X.export(), W.export()
    assert ew[4].dtype == np.uint16
    for x, w in zip(ex, ew):
        assert np.array_equal(x, w)
    if kind == "rmat":
        meta, arr = load_rmat(p, 8)
        assert np.array_equal(ew[3], arr["hub"]) and np.array_equal(ew[4], arr["dist"])
    n = G.num_v
    qa, qb = pado.uniform_pairs(n, 20000, seed=2)
    gx = X.query(qa, qb, rank_space=True)
    gw = W.query(qa, qb, rank_space=True)
    assert gw.dtype == np.uint16
    ok = gx != 255  # the default domain's 255 means unreachable (or >= 253)
    assert np.array_equal(gw[ok], gx[ok])
    assert np.all((gw[~ok] == pado.WIDE_INF) | ((gw[~ok] >= 253) & (gw[~ok] < pado.WIDE_INF)))
    for s in (0, n // 2, n - 1):
        assert np.array_equal(W.query(np.full(n, s, np.uint32), np.arange(n, dtype=np.uint32), rank_space=True),
                              _exact(G, s))
    X.close()
    W.close()
    dg.close()


@pytest.mark.parametrize("kind,p", [("grid", 180), ("grid", 300), ("path", 1500)])
def test_wide_deep_graphs_exact(gpu_ctx, kind, p):
    G, rank, r2i = _graph(kind, p, seed=7)
    n = G.num_v
    say(f"wide deep {kind} {p} n={n}")
    dg = pado.DeviceGraph(gpu_ctx, G)
    with pytest.raises(pado.PadoError):
        pado.Index.build(gpu_ctx, dg)  # the reference's u8 domain cannot hold it
    W = pado.Index.build(gpu_ctx, dg, wide=True)
    bp_d, bp_s, loff, hub, dist = W.export()
    lab_of = np.repeat(np.arange(n), np.diff(loff.astype(np.int64)))
    assert np.all(hub <= lab_of)  # hubs rank at or above their vertex
    # Bit-parallel distances: BFS from the slot root, up to level 254.
    info = W.get_info()
    for slot in (0, 1, 49):
        r = info.bp_roots[slot]
        if r == 0xFFFFFFFF:
            continue
        d = oracle.bfs(n, G.off, G.adj, int(r)).astype(np.int64)
        exp = np.where(d <= 254, d, 255)
        assert np.array_equal(bp_d[:, slot].astype(np.int64), exp), slot
    # Labels of sampled hubs carry the BFS distance.
    rng = np.random.default_rng(3)
    for h in rng.choice(np.unique(hub), size=min(12, np.unique(hub).size), replace=False):
        sel = hub == h
        d = oracle.bfs(n, G.off, G.adj, int(h)).astype(np.int64)
        assert np.array_equal(dist[sel].astype(np.int64), d[lab_of[sel]]), h
    # Answers: sampled sources to every target equal BFS, rank space and
    # original ids.
    W.order(rank)
    tgt = np.arange(n, dtype=np.uint32)
    far = 0
    for s in [0, 1, n // 3, n - 1] + list(rng.integers(0, n, 4)):
        s = int(s)
        exp = _exact(G, s)
        far += int(np.count_nonzero((exp > 255) & (exp < pado.WIDE_INF)))
        assert np.array_equal(W.query(np.full(n, s, np.uint32), tgt, rank_space=True), exp), s
        s_orig = int(r2i[s])
        assert np.array_equal(W.query(np.full(n, s_orig, np.uint32), r2i[tgt]), exp), s
    assert far > 0  # answers beyond the u8 domain were checked
    # Out-of-range ids.
    assert W.query([n, n, 0], [n, n + 1, n]).tolist() == [0, pado.WIDE_INF, pado.WIDE_INF]
    W.close()
    dg.close()


@pytest.mark.parametrize("kind,p,seed", [("path", 700, 1), ("grid", 24, 5)])
def test_wide_all_pairs_exact(gpu_ctx, kind, p, seed):
    """Every pair of a deep path (700 vertices, diameter 699: the bit-parallel
    BFS is cut at level 254, so bit-parallel vertices become hubs) and of a
    small grid (no cut) equals BFS."""
    G, rank, _ = _graph(kind, p, seed=seed)
    n = G.num_v
    say(f"wide all pairs {kind} {p} n={n}")
    W = pado.Index.build(gpu_ctx, pado.DeviceGraph(gpu_ctx, G), wide=True)
    a = np.repeat(np.arange(n, dtype=np.uint32), n)
    b = np.tile(np.arange(n, dtype=np.uint32), n)
    got = W.query(a, b, rank_space=True).reshape(n, n)
    for s in range(n):
        assert np.array_equal(got[s], _exact(G, s)), s
    W.close()


def test_wide_domain_limits_and_refusals(gpu_ctx):
    # More than 2^20 vertices does not fit the wide labels' 20 hub bits.
    n = (1 << 20) + 5
    v = np.arange(n - 1, dtype=np.uint32)
    G, _, _ = rank_space(np.stack([v, v + 1], axis=1))
    dg = pado.DeviceGraph(gpu_ctx, G)
    with pytest.raises(pado.PadoError) as e:
        pado.Index.build(gpu_ctx, dg, wide=True)
    assert e.value.status == 4
    dg.close()
    # A wide index refuses the u8 query / export / file paths.
    G, rank, _ = _graph("grid", 30)
    dg = pado.DeviceGraph(gpu_ctx, G)
    W = pado.Index.build(gpu_ctx, dg, wide=True)
    W.order(rank)
    L = pado.lib()
    out = np.zeros(4, np.uint8)
    a = np.zeros(4, np.uint32)
    assert L.pado_gpu_query(gpu_ctx.h, W.h, pado._p(a), pado._p(a), 4, pado._p(out), 0, None) == 5
    assert L.pado_gpu_query_noalloc(gpu_ctx.h, W.h, pado._p(a), pado._p(a), 4, pado._p(out), 0) == 5
    with pytest.raises(pado.PadoError):
        W.store("/tmp/pado_wide_refused.idx", rank)
    W.close()
    dg.close()


def test_wide_index_served(gpu_ctx):
    """A wide index is served with u16 answers equal to the batch query's."""
    G, rank, r2i = _graph("grid", 200, seed=11)
    n = G.num_v
    say(f"wide serve n={n}")
    W = pado.Index.build(gpu_ctx, pado.DeviceGraph(gpu_ctx, G), wide=True)
    W.order(rank)
    rng = np.random.default_rng(4)
    reqs = [(rng.integers(0, n, k).astype(np.uint32), rng.integers(0, n, k).astype(np.uint32))
            for k in (1, 17, 300, 5000, 40000, 2)]
    srv = pado.Server(gpu_ctx, W, staging_pairs=1 << 18, streams=2)
    try:
        ids = [srv.submit(a, b) for a, b in reqs]
        outs = [srv.wait(r, a.size)[0] for r, (a, _) in zip(ids, reqs)]
    finally:
        srv.stop()
    for (a, b), o in zip(reqs, outs):
        assert o.dtype == np.uint16 and np.array_equal(o, W.query(a, b))
    W.close()
