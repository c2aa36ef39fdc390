"""Edge-list files on the host: the parallel text reader against the C++
standard library's own line loop (oracle/line_fields.cpp) and against the
reference's Graph(const char*) / WeightedGraph(const char*) CSR
(includes/graph.h:95-118, 277-304, built under oracle/_ref), and the binary
edge list of the reference's converter (edgelist2binary.cpp:28-91)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
import parallel_distance_oracle_amd as pado

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
need_ref = pytest.mark.skipif(not oracle.have_ref(), reason="reference harness not built")

# Lines the reference's loop reads in unusual ways.  The first edge line sets
# every field: the reference's variables start uninitialised.
CRAFTED = (
    "3 4\n"
    "# comment\n"
    "% matrix-market comment\n"
    "\n"                      # nothing extracted: previous (3, 4) again
    "5\n"                     # head 5, tail kept
    "6 \n"                    # trailing blank: tail kept
    "7 x\n"                   # tail not a number: 0
    "x 9\n"                   # head not a number: 0, tail kept
    "  10\t11  \n"
    "12 13\r\n"               # CRLF
    "+7 8\n"
    " # not a comment\n"      # head '#': 0, tail kept
    "8 9 10 extra\n"
    "9\t\t1\n"
    "5x 6\n"                  # head 5, then 'x' fails tail: 0
    "1 2"                     # no final newline
)

OVERFLOW = (
    "1 2\n"
    "4294967295 3\n"
    "4294967296 4\n"          # past 2^32 - 1: max, tail kept
    "99999999999999999999999 5\n"
    "5 -1\n"                  # unsigned extraction of -1: 2^32 - 1
    "-3 4\n"
    "- 4\n"                   # sign alone: 0
    "-4294967296 6\n"
    "7 8\n"
)

WEIGHTED = (
    "1 2 3\n"
    "# c\n"
    "2 3 300\n"               # weighti is u8: 44
    "3 4\n"                   # weight kept (300)
    "4 5 x\n"                 # weight 0
    "5 6 -1\n"                # 2^32 - 1 -> 255
    "6 7 8 9\n"
    "\n"
    "7 1 256\r\n"
    "2 6 17"
)


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_bytes(text.encode())
    return p


@pytest.mark.parametrize("text", [CRAFTED, OVERFLOW], ids=["crafted", "overflow"])
def test_text_fields_match_istream(tmp_path, text):
    p = _write(tmp_path, "e.txt", text)
    got = pado.read_edge_text(p, 2)
    exp = oracle.line_fields(p, 2)
    assert got.shape == exp.shape and np.array_equal(got, exp)


def test_weighted_fields_match_istream(tmp_path):
    p = _write(tmp_path, "w.txt", WEIGHTED)
    got = pado.read_edge_text(p, 3)
    exp = oracle.line_fields(p, 3)
    assert np.array_equal(got, exp)


def test_empty_and_comment_only(tmp_path):
    for text in ("", "# only\n% comments\n"):
        p = _write(tmp_path, "e.txt", text)
        assert pado.read_edge_text(p).shape == (0, 2)
        assert np.array_equal(pado.read_edge_text(p), oracle.line_fields(p))
        G = pado.Graph.from_file(p)
        assert G.num_v == 0 and G.num_e == 0


def test_missing_file_and_bad_arity(tmp_path):
    with pytest.raises(OSError):
        pado.read_edge_text(tmp_path / "absent.txt")
    p = _write(tmp_path, "e.txt", "1 2\n")
    with pytest.raises(OSError):
        pado.read_edge_text(p, 4)


def _fixed(a, b):
    return f"{a:>11d} {b:>11d}\n"


def test_chunk_boundaries(tmp_path):
    """A file of several parse chunks (cut after the first newline at or past
    each 4 MiB step) whose chunks start with lines that keep earlier values,
    and one chunk made only of such lines."""
    rng = np.random.default_rng(11)
    W = 24
    nlines = (13 << 20) // W
    a = rng.integers(0, 1 << 20, nlines)
    b = rng.integers(0, 1 << 20, nlines)
    kinds = rng.choice(5, nlines, p=[0.8, 0.05, 0.05, 0.05, 0.05])
    kinds[0] = 0
    blank = " " * (W - 1) + "\n"

    def render(i):
        k = kinds[i]
        if k == 0:
            return _fixed(a[i], b[i])
        if k == 1:
            return blank
        if k == 2:
            return f"{a[i]:>11d}" + " " * 12 + "\n"
        if k == 3:
            return f"{a[i]:>11d} {'x':>11s}\n"
        return f"{'#':<23s}\n"

    step = 1 << 22
    size = nlines * W
    cuts = [0]
    while cuts[-1] < size:
        c = min(size, cuts[-1] + step)
        if c < size:
            c = c + (W - 1 - c) % W + 1  # one past the first newline at or after c
        cuts.append(c)
    starts = [c // W for c in cuts[1:-1]]
    assert len(starts) >= 3
    kinds[starts[0]] = 1
    kinds[starts[1]] = 2
    kinds[starts[1] + 1] = 1
    kinds[starts[2]:] = 1  # the last chunk: nothing set at all
    kinds[starts[2] - 1] = 0
    p = tmp_path / "big.txt"
    with open(p, "w") as f:
        f.write("".join(render(i) for i in range(nlines)))
    assert os.path.getsize(p) == size
    got = pado.read_edge_text(p, 2)
    exp = oracle.line_fields(p, 2)
    assert got.shape == exp.shape and np.array_equal(got, exp)


@need_ref
@pytest.mark.parametrize("weighted", [False, True])
def test_graph_from_file_matches_reference(tmp_path, weighted):
    """Graph.from_file / WeightedGraph.from_file build the CSR the reference's
    file constructors build from the same file."""
    p = _write(tmp_path, "e.txt", WEIGHTED if weighted else CRAFTED)
    exp = oracle.ref_text_csr(p, weighted)
    G = (pado.WeightedGraph if weighted else pado.Graph).from_file(p)
    assert G.num_v == exp[0]
    assert np.array_equal(G.off, exp[1]) and np.array_equal(G.adj, exp[2])
    if weighted:
        assert np.array_equal(G.w, exp[3])


@need_ref
@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_cpp_dropin_file_constructors(tmp_path):
    """include/pado/graph.h's Graph(const char*) / WeightedGraph(const char*)
    against the reference's, on the same files; a missing file exits 1 with
    the reference's message."""
    exe = tmp_path / "graph_text"
    libdir = os.path.dirname(pado.LIB_PATH)
    subprocess.run(["g++", "-std=c++14", "-O1", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "graph_text.cpp"), "-o", str(exe), "-L", libdir,
                    "-lpado_gpu", "-Wl,-rpath," + libdir], check=True)
    for mode, text in (("text", CRAFTED), ("wtext", WEIGHTED)):
        p = _write(tmp_path, mode + ".txt", text)
        out = tmp_path / (mode + ".bin")
        subprocess.run([str(exe), mode, str(p), str(out)], check=True, timeout=60)
        exp = oracle.ref_text_csr(p, mode == "wtext")
        with open(out, "rb") as f:
            n = int(np.frombuffer(f.read(4), np.uint32)[0])
            m = int(np.frombuffer(f.read(8), np.uint64)[0])
            off = np.frombuffer(f.read(8 * (n + 1)), np.uint64)
            adj = np.frombuffer(f.read(4 * m), np.uint32)
            w = np.frombuffer(f.read(m), np.uint8) if mode == "wtext" else None
        assert n == exp[0] and np.array_equal(off, exp[1]) and np.array_equal(adj, exp[2])
        if w is not None:
            assert np.array_equal(w, exp[3])
    r = subprocess.run([str(exe), "text", str(tmp_path / "absent.txt"), str(tmp_path / "x.bin")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "cannot open file" in r.stderr


def test_binary_round_trip(tmp_path):
    e = pado.rmat_edges(12, 8, 0.57, 0.19, 0.19, 3)
    p = tmp_path / "e.bin"
    pado.write_edge_binary(p, e)
    raw = p.read_bytes()
    assert np.frombuffer(raw[:4], np.uint32)[0] == e.max() + 1
    assert np.frombuffer(raw[4:12], np.uint64)[0] == e.shape[0]
    assert np.array_equal(pado.read_edge_binary(p), e.reshape(-1, 2))
    G1, G2 = pado.Graph.from_binary(p), pado.Graph.from_edges(e)
    assert np.array_equal(G1.off, G2.off) and np.array_equal(G1.adj, G2.adj)
    # A header that disagrees with the file's size is refused.
    q = tmp_path / "short.bin"
    q.write_bytes(raw[:-8])
    with pytest.raises(OSError):
        pado.read_edge_binary(q)
    q.write_bytes(raw[:6])
    with pytest.raises(OSError):
        pado.read_edge_binary(q)
    empty = tmp_path / "empty.bin"
    pado.write_edge_binary(empty, np.zeros((0, 2), np.uint32))
    assert pado.read_edge_binary(empty).shape == (0, 2)
