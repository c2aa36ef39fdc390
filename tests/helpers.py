"""Shared test helpers: fixture loading, rank-space graphs, digests."""
import hashlib
import json
import os

import numpy as np

import oracle
import parallel_distance_oracle_amd as pado

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TINY = ("edgelist_in", "input_dat", "deleted_graph_in")
RMAT_SMALL = ((10, 8), (12, 8), (14, 8))


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def label_digest(loff, hub, dist):
    return digest(np.asarray(loff, np.uint64), np.asarray(hub, np.uint32), np.asarray(dist, np.uint8))


def bp_digest(bp_d, bp_s):
    return digest(np.asarray(bp_d, np.uint8), np.asarray(bp_s, np.uint64))


def load_tiny(name):
    return dict(np.load(os.path.join(GOLDEN, f"tiny_{name}.npz")))


def load_rmat(scale, ef):
    tag = f"rmat{scale}_ef{ef}"
    with open(os.path.join(GOLDEN, tag + ".json")) as f:
        meta = json.load(f)
    arr = dict(np.load(os.path.join(GOLDEN, tag + ".npz")))
    return meta, arr


def have_rmat(scale, ef):
    return os.path.exists(os.path.join(GOLDEN, f"rmat{scale}_ef{ef}.json"))


def rank_space(edges, rank=None):
    """Product-side pipeline: Graph(edges) -> make_rank -> id_transfer.
    Returns (G in rank space, rank, rank2id)."""
    G = pado.Graph.from_edges(edges)
    if rank is None:
        rank = G.make_rank()
    rank2id = G.id_transfer(rank)
    return G, np.asarray(rank, np.uint32), rank2id


def oracle_rank_space(edges, rank):
    n, off, adj = oracle.graph_csr(edges)
    off2, adj2, r2i = oracle.id_transfer(n, off, adj, rank)
    return n, off2, adj2, r2i


def rmat_edges(scale, ef):
    return pado.rmat_edges(scale, ef, 0.57, 0.19, 0.19, 1)
