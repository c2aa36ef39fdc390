"""The C++ drop-in (include/pado/*.h) driven the way the reference's own
drivers call PADO (pado.cpp, query_distance.cpp), on the GPU."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
import parallel_distance_oracle_amd as pado
from helpers import load_tiny

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_cpp_dropin_end_to_end(tmp_path):
    g = load_tiny("edgelist_in")
    edges = g["edges"]
    ep = tmp_path / "edges.txt"
    ep.write_text("".join(f"{a} {b}\n" for a, b in edges))
    n = int(edges.max()) + 1
    pp = tmp_path / "pairs.txt"
    pp.write_text("".join(f"{a} {b}\n" for a, b in zip(g["qa"], g["qb"])))
    exe = tmp_path / "dropin"
    libdir = os.path.dirname(pado.LIB_PATH)
    cmd = ["g++", "-std=c++14", "-O1", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "dropin_example.cpp"), "-o", str(exe), "-L", libdir, "-lpado_gpu",
           "-Wl,-rpath," + libdir]
    subprocess.run(cmd, check=True)
    print("[gpu] running C++ drop-in", flush=True)
    r = subprocess.run([str(exe), str(ep), str(tmp_path / "idx.bin"), str(pp)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    q = [line.split() for line in r.stdout.splitlines() if line.startswith("Q ")]
    lq = [int(line.split()[1]) for line in r.stdout.splitlines() if line.startswith("L ")]
    assert len(q) == n * n and len(lq) == n * n
    exact = np.array([int(x[3]) for x in q], np.uint8)
    refc = np.array([int(x[4]) for x in q], np.uint8)
    assert np.array_equal(refc, g["answers"])
    rank = g["rank"]
    nn, off0, adj0 = oracle.graph_csr(edges)
    off2, adj2, _ = oracle.id_transfer(nn, off0, adj0, rank)
    bfs = oracle.bfs_to_u8(oracle.bfs_pairs(nn, off2, adj2, rank[g["qa"]], rank[g["qb"]]))
    assert np.array_equal(exact, bfs)
    assert np.array_equal(np.array(lq, np.uint8), exact)
    assert "Total_labeling_time:" in r.stdout and "Label_size:" in r.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
@pytest.mark.parametrize("tag", ["weighted_edgelist_in", "7_w_edgelist_in"])
def test_cpp_weighted_dropin_end_to_end(tmp_path, tag):
    """WeightedGraph -> WeightedParaVertexCentricPLL<512> -> store / order /
    query_distance, then load_index_from_file, as pado.cpp and
    query_distance.cpp drive the reference; answers equal the reference's."""
    from helpers import GOLDEN

    g = dict(np.load(os.path.join(GOLDEN, f"weighted_{tag}.npz")))
    ep = tmp_path / "wedges.txt"
    ep.write_text("".join(f"{a} {b} {w}\n" for (a, b), w in zip(g["edges"], g["weights"])))
    pp = tmp_path / "pairs.txt"
    pp.write_text("".join(f"{a} {b}\n" for a, b in zip(g["qa"], g["qb"])))
    exe = tmp_path / "wdropin"
    libdir = os.path.dirname(pado.LIB_PATH)
    cmd = ["g++", "-std=c++14", "-O1", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "wdropin_example.cpp"), "-o", str(exe), "-L", libdir, "-lpado_gpu",
           "-Wl,-rpath," + libdir]
    subprocess.run(cmd, check=True)
    print(f"[gpu] running weighted C++ drop-in {tag}", flush=True)
    r = subprocess.run([str(exe), str(ep), str(tmp_path / "widx.bin"), str(pp)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    q = np.array([int(x.split()[3]) for x in lines if x.startswith("Q ")], np.int64)
    xq = np.array([int(x.split()[3]) for x in lines if x.startswith("X ")], np.int64)
    lq = np.array([int(x.split()[1]) for x in lines if x.startswith("L ")], np.int64)
    ans = g["answers"].astype(np.int64)
    assert np.array_equal(q, ans) and np.array_equal(lq, ans)
    reach = ans != 32767
    assert np.array_equal(xq[reach], ans[reach])
    assert np.all(xq[~reach] == pado.WINF)  # tiny graphs: 32767 only for unreachable pairs
