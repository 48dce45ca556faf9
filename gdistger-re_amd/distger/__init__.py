"""Python bindings (ctypes) for libdistger_hip.so, the MI355X DistGER
embedding hot path.  The library is the product; this package only forwards
calls through its C ABI (include/distger_hip.h) and raises DgError on any
non-zero status.  There is no fallback: if the shared library is missing the
import of `distger.ffi` functions raises."""

from .ffi import (  # noqa: F401
    DgError,
    EDGE_DTYPE,
    Context,
    WalkParams,
    RoundStats,
    lib_path,
    load_library,
    gen_chung_lu,
    gen_rmat,
    gen_zipf,
    gen_normal,
    lp_split,
    auc,
    write_edges_bin,
    read_edges_bin,
    convert_text_edges,
    lp_write_files,
    lr_value,
)
