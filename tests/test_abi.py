"""CPU checks of the C ABI: the library builds, loads, and exports every
entry point include/thv.h declares; the Python mirror binds them all.  No
compute call is made here (the container has no GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "thv.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(thv_[a-z_0-9]+)\s*\(", text)))


@pytest.fixture(scope="module")
def libpath():
    import thievory_amd.build as B
    if not os.path.exists(B.LIB):
        B.build(verbose=False)
    return B.LIB


def test_header_declares_the_boundary():
    syms = declared_symbols()
    for s in ("thv_graph_open", "thv_bfs", "thv_cc", "thv_sssp", "thv_pr", "thv_batch",
              "thv_serve_start", "thv_serve_submit", "thv_serve_poll", "thv_serve_stop",
              "thv_last_error", "thv_graph_close", "thv_store_config"):
        assert s in syms


def test_library_exports_every_declared_symbol(libpath):
    lib = ctypes.CDLL(libpath)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", libpath], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (thv_\w+)", out))
    assert set(declared_symbols()) <= exported


def test_python_mirror_binds_every_symbol(libpath):
    import thievory_amd as T
    L = T.lib()
    for s in declared_symbols():
        assert getattr(L, s).argtypes is not None or s in ("thv_last_error", "thv_version"), s


def test_header_compiles_as_c(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "thv.h"\nint main(void){thv_stats s; (void)s; return 0;}\n')
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.dirname(HEADER),
                    "-c", str(src), "-o", str(tmp_path / "t.o")], check=True)
