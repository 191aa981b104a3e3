"""The ctypes mirrors in eggpu_amd/native.py must lay out every struct of
include/egx.h exactly as a C compiler does: same size, and each field at the
same offset under the same name.  A small C program built with gcc against the
header prints sizeof/offsetof; no GPU and no library load are needed."""
import ctypes
import os
import shutil
import subprocess

import pytest

from eggpu_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PAIRS = {
    "egx_graph_info": native.GraphInfo,
    "egx_bin_opts": native.BinOpts,
    "egx_solve_opts": native.SolveOpts,
    "egx_solve_stats": native.SolveStats,
    "egx_ingest_report": native.IngestReport,
    "egx_server_opts": native.ServerOpts,
    "egx_request_info": native.RequestInfo,
    "egx_server_stats": native.ServerStats,
    "egx_cache_stats": native.CacheStats,
    "egx_eig_opts": native.EigOpts,
}


def _c_layout(tmp_path):
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "egx.h"', "int main(void) {"]
    for cname, cls in PAIRS.items():
        lines.append(f'    printf("%s size %zu\\n", "{cname}", sizeof({cname}));')
        for field, _ in cls._fields_:
            lines.append(f'    printf("%s %s %zu\\n", "{cname}", "{field}", offsetof({cname}, {field}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    sizes, offsets = {}, {}
    for line in out.splitlines():
        a, b, c = line.split()
        if b == "size":
            sizes[a] = int(c)
        else:
            offsets[(a, b)] = int(c)
    return sizes, offsets


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_ctypes_mirrors_match_the_header(tmp_path):
    sizes, offsets = _c_layout(tmp_path)
    for cname, cls in PAIRS.items():
        assert ctypes.sizeof(cls) == sizes[cname], cname
        for field, _ in cls._fields_:
            assert getattr(cls, field).offset == offsets[(cname, field)], (cname, field)
