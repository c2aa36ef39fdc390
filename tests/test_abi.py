"""The C ABI library loads without a GPU and exports every declared symbol;
the C++ drop-in headers compile against it."""
import ctypes
import os
import shutil
import subprocess

import pytest

import parallel_distance_oracle_amd as pado

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    L = pado.lib()
    names = pado.exported_symbols()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_error_path_without_device_is_a_status_not_a_crash():
    L = pado.lib()
    h = ctypes.c_void_p()
    st = L.pado_gpu_ctx_create(10_000, ctypes.byref(h))
    assert st != 0
    assert L.pado_gpu_last_error() != b""


def test_null_arguments_are_rejected():
    L = pado.lib()
    assert L.pado_gpu_graph_upload(None, 0, 0, None, None, None, None) == 1
    assert L.pado_gpu_query(None, None, None, None, 0, None, 0, None) == 1


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
@pytest.mark.parametrize("example", ["dropin_example.cpp", "wdropin_example.cpp"])
def test_cpp_drop_in_compiles(tmp_path, example):
    src = os.path.join(ROOT, "tests", "cpp", example)
    if not os.path.exists(src):
        pytest.skip("C++ drop-in example not present")
    out = tmp_path / "dropin"
    cmd = ["g++", "-std=c++14", "-O1", "-I", os.path.join(ROOT, "include"), src, "-o", str(out),
           "-L", os.path.dirname(pado.LIB_PATH), "-lpado_gpu", "-Wl,-rpath," + os.path.dirname(pado.LIB_PATH)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no host C compiler")
def test_ctypes_structs_match_the_header(tmp_path):
    """Every struct the Python mirror passes across the ABI has the header's
    size and field offsets (a C compiler reads include/pado_gpu.h)."""
    structs = {"pado_gpu_build_cfg": pado.BuildCfg, "pado_gpu_index_info": pado.IndexInfo,
               "pado_gpu_serve_cfg": pado.ServeCfg, "pado_gpu_serve_stats": pado.ServeStats}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "pado_gpu.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["return 0; }"]
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    r = subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)
