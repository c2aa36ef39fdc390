"""Writes the golden fixtures under tests/golden/ from the REFERENCE build.

    python tests/golden/make_golden.py [--big]

Inputs: the reference's own tiny edge lists (tools/*.in, copied here as
arrays) and R-MAT graphs from this package's seeded generator.  Outputs come
from oracle/_ref/pado_ref (the reference's ParaVertexCentricPLL<1024>,
compiled from its headers by oracle/build_ref.sh): its make_rank order, its
bit-parallel arrays, its canonical label sets and its query_distance answers.
Large graphs are recorded as SHA-256 digests (see digest()).

Run in the build container (the reference tree is not on the GPU box); the
fixtures are committed.
"""
import argparse
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
import parallel_distance_oracle_amd as pado  # noqa: E402

REF_TOOLS = "/root/reference/tools"


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def label_digest(loff, hub, dist):
    return digest(np.asarray(loff, np.uint64), np.asarray(hub, np.uint32), np.asarray(dist, np.uint8))


def bp_digest(bp_d, bp_s):
    return digest(np.asarray(bp_d, np.uint8), np.asarray(bp_s, np.uint64))


def rmat_config(scale, ef, seed=1):
    return pado.rmat_edges(scale, ef, 0.57, 0.19, 0.19, seed)


def tiny(name):
    edges = oracle.read_text_edges(os.path.join(REF_TOOLS, name))
    rank = oracle.ref_rank(edges)
    n = int(edges.max()) + 1
    a, b = np.meshgrid(np.arange(n, dtype=np.uint32), np.arange(n, dtype=np.uint32), indexing="ij")
    a, b = a.reshape(-1), b.reshape(-1)
    R = oracle.run_ref(edges, rank, pairs=(a, b), threads=4)
    loff, hub, dist = R["labels"]
    bp_d, bp_s = R["bp"]
    out = os.path.join(HERE, "tiny_" + name.replace(".", "_") + ".npz")
    np.savez_compressed(out, edges=edges, rank=rank, bp_d=bp_d, bp_s=bp_s, loff=loff, hub=hub, dist=dist,
                        qa=a, qb=b, answers=R["answers"])
    print("wrote", out, "n", n, "labels", hub.size)


def rmat(scale, ef, n_pairs, full_labels, threads=8):
    edges = rmat_config(scale, ef)
    rank = oracle.ref_rank(edges)
    n = int(edges.max()) + 1
    qa, qb = pado.uniform_pairs(n, n_pairs, seed=1)
    R = oracle.run_ref(edges, rank, pairs=(qa, qb), threads=threads)
    loff, hub, dist = R["labels"]
    bp_d, bp_s = R["bp"]
    meta = {
        "scale": scale, "edge_factor": ef, "seed": 1, "num_v": n, "num_edges": int(edges.shape[0]),
        "edges_sha256": digest(edges.astype(np.uint32)),
        "rank_sha256": digest(rank.astype(np.uint32)),
        "num_labels": int(hub.size),
        "labels_sha256": label_digest(loff, hub, dist),
        "bp_sha256": bp_digest(bp_d, bp_s),
        "n_pairs": int(n_pairs), "pairs_seed": 1,
        "answers_sha256": digest(R["answers"]),
        "bp_hit_count": R.get("bp_hit_count"),
        "ref_build_s": R.get("ref_build_best_s"),
    }
    tag = f"rmat{scale}_ef{ef}"
    with open(os.path.join(HERE, tag + ".json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    arrays = {"answers": R["answers"]}
    if full_labels:
        arrays.update(loff=loff, hub=hub, dist=dist)
    np.savez_compressed(os.path.join(HERE, tag + ".npz"), **arrays)
    print("wrote", tag, meta["num_labels"], "labels")


def read_weighted(path):
    rows = []
    with open(path) as f:
        for line in f:
            if not line or line[0] in "#%":
                continue
            t = line.split()
            if len(t) >= 3:
                rows.append((int(t[0]), int(t[1]), int(t[2])))
    a = np.array(rows, np.uint32)
    return a[:, :2], a[:, 2]


def weighted(tag, edges, weights, n_pairs=None):
    """Reference WeightedParaVertexCentricPLL<512>, one OpenMP thread."""
    n = int(edges.max()) + 1
    rank = oracle.ref_rank(edges)
    if n_pairs is None:
        a, b = np.meshgrid(np.arange(n, dtype=np.uint32), np.arange(n, dtype=np.uint32), indexing="ij")
        qa, qb = a.reshape(-1), b.reshape(-1)
    else:
        qa, qb = pado.uniform_pairs(n, n_pairs, seed=1)
    R = oracle.run_ref_weighted(edges, weights, rank, pairs=(qa, qb), threads=1)
    loff, hub, dist = R["labels"]
    out = os.path.join(HERE, f"weighted_{tag}.npz")
    np.savez_compressed(out, edges=edges, weights=weights, rank=rank, loff=loff, hub=hub, dist=dist, qa=qa, qb=qb,
                        answers=R["answers"])
    print("wrote", out, "n", n, "labels", hub.size)


def unp_index(tag, edges, rank):
    """An index file written by the reference's own writer (the sequential
    VertexCentricPLL<1024>, UNP:1120-1180), stored as its bytes."""
    data = np.frombuffer(oracle.ref_unp_file(edges, rank), np.uint8)
    out = os.path.join(HERE, f"unp_{tag}.npz")
    np.savez_compressed(out, edges=np.asarray(edges, np.uint32), rank=np.asarray(rank, np.uint32), index=data)
    print("wrote", out, data.size, "bytes")


def unp_fixtures():
    for name in ("edgelist_in", "input_dat"):
        g = np.load(os.path.join(HERE, f"tiny_{name}.npz"))
        unp_index(name, g["edges"], g["rank"])
    e = rmat_config(10, 8)
    unp_index("rmat10_ef8", e, oracle.ref_rank(e))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--big", action="store_true", help="also R-MAT 16/18/20 digests (minutes)")
    ap.add_argument("--unp-only", action="store_true", help="only the reference-written index files")
    args = ap.parse_args()
    if not oracle.have_ref():
        sys.exit("oracle/_ref/pado_ref is not built (run oracle/build_ref.sh)")
    if args.unp_only:
        unp_fixtures()
        return
    unp_fixtures()
    for name in ("edgelist.in", "input.dat", "deleted_graph.in"):
        tiny(name)
    rmat(10, 8, 20000, True)
    rmat(12, 8, 20000, True)
    rmat(14, 8, 100000, True)
    for name in ("weighted_edgelist.in", "7_w_edgelist.in", "1_w_edgelist.in", "triangle.in", "rectangle.in"):
        e, w = read_weighted(os.path.join(REF_TOOLS, name))
        weighted(name.replace(".", "_"), e, w)
    for scale in (10, 12):
        e = rmat_config(scale, 8)
        w = pado.edge_weights(e.shape[0], 1, 100, seed=1)
        weighted(f"rmat{scale}_ef8", e, w, n_pairs=20000)
    if args.big:
        rmat(16, 16, 100000, False)
        rmat(18, 16, 100000, False)
        rmat(20, 16, 100000, False)


if __name__ == "__main__":
    main()
