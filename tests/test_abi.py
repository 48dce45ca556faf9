"""The C ABI: libdistger_hip.so loads without a GPU and exports every entry
point include/distger_hip.h declares; the Python binding declares each of
them; a missing library is an error, never a fallback."""

import ctypes
import os
import re

import pytest

import distger
from distger import ffi

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "distger_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b(dg_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for n in ("dg_open", "dg_close", "dg_graph_from_edges", "dg_walk_round", "dg_walk_run",
              "dg_corpus_write_text", "dg_sgns_init", "dg_sgns_train", "dg_emb_write",
              "dg_sgns_tuples_forward", "dg_sgns_tuples_backward", "dg_last_error"):
        assert n in names
    assert len(names) >= 40


def test_library_exports_every_declared_symbol():
    path = ffi.lib_path()
    assert os.path.exists(path), "build with __graft_entry__.build() first"
    lib = ctypes.CDLL(path)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_the_header():
    L = distger.load_library()
    for n in declared_functions():
        fn = getattr(L, n)
        assert fn.argtypes is not None or n.endswith("_default"), f"{n} has no argtypes"


def test_missing_library_is_an_error(monkeypatch, tmp_path):
    monkeypatch.setattr(ffi, "_lib", None)
    monkeypatch.setattr(ffi, "lib_path", lambda: str(tmp_path / "libdistger_hip.so"))
    with pytest.raises(distger.DgError):
        ffi.load_library()


def test_struct_layouts_match_the_header(tmp_path):
    """Every struct the binding passes by pointer has the C compiler's size
    and field offsets (a field added on one side only would shift the rest)."""
    import subprocess

    pairs = [("dg_walk_params", ffi.WalkParams), ("dg_round_stats", ffi.RoundStats),
             ("dg_sgns_params", ffi.SgnsParams), ("dg_sgns_stats", ffi.SgnsStats)]
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "distger_hip.h"',
             "int main(void) {"]
    for cname, cls in pairs:
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append('  printf("dg_edge size %zu\\n", sizeof(dg_edge));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(REPO, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        a, b, v = line.split()
        got[(a, b)] = int(v)
    for cname, cls in pairs:
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)
    assert got[("dg_edge", "size")] == ffi.EDGE_DTYPE.itemsize


def test_walk_length_limit_default():
    """The header names the length limit a zero length_limit stands for, the
    walker takes it from there, and the reference defaults leave the field 0."""
    text = open(HEADER).read()
    m = re.search(r"#define DG_WALK_LENGTH_LIMIT_DEFAULT (\d+)", text)
    assert m and int(m.group(1)) == 1024
    assert "0 = DG_WALK_LENGTH_LIMIT_DEFAULT" in text
    src = open(os.path.join(REPO, "gdistger-re_amd", "csrc", "walker.hip")).read()
    assert "kDefaultLengthLimit = DG_WALK_LENGTH_LIMIT_DEFAULT;" in src
    p = ffi.WalkParams()
    distger.load_library().dg_walk_params_default(ctypes.byref(p), 1000)
    assert p.length_limit == 0 and p.max_length == 0 and p.min_length == 20
