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
