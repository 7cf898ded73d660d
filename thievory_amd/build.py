"""Build recipe for the native parts of thievory_amd.

    python -m thievory_amd.build          # library + CLI
    python -m thievory_amd.build --force  # rebuild everything

Produces, in-tree:
    thievory_amd/_lib/libthv_hip.so   the C-ABI library (include/thv.h)
    thievory_amd/_lib/thv_main        the drop-in CLI (links only the ABI)
Objects go to thievory_amd/_build/.  Every HIP source is compiled for gfx950
with packed-FP32 instruction selection disabled.
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIB_DIR, "libthv_hip.so")
CLI = os.path.join(LIB_DIR, "thv_main")
INCLUDE = os.path.join(ROOT, "include")

ARCH = os.environ.get("THV_OFFLOAD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-result", "-I", INCLUDE]

LIB_SOURCES = ["graph.cpp", "io.cpp", "abi.cpp", "frontier.hip", "expand.hip",
               "algos.hip", "rmat.hip", "batch.hip", "serve.cpp", "autotune.cpp",
               "bottomup.hip", "cc.hip", "noalloc.hip", "wide.hip", "probe.hip"]
CLI_SOURCES = ["cli.cpp"]
# stand-alone host tools (no GPU): format converter and values.bin comparer
TOOLS = {"thv_convert": "convert.cpp", "thv_compare": "compare.cpp"}


def _obj(src):
    return os.path.join(BUILD, os.path.splitext(src)[0] + ".o")


def _stale(target, deps):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps)


def _headers():
    return [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".h"))] + \
        [os.path.join(INCLUDE, "thv.h")]


def _compile(src, force):
    path = os.path.join(CSRC, src)
    obj = _obj(src)
    if not force and not _stale(obj, [path] + _headers()):
        return obj, None
    if src in CLI_SOURCES:  # plain C++ against the C ABI only
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", INCLUDE, "-c", path, "-o", obj]
    else:
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = [HIPCC] + COMMON + lang + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return obj, f"{' '.join(cmd)}\n{r.stdout}\n{r.stderr}"
    return obj, None


def build(force=False, jobs=None, verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    srcs = LIB_SOURCES + [s for s in CLI_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = jobs or min(8, os.cpu_count() or 4)
    errors = []
    with cf.ThreadPoolExecutor(jobs) as ex:
        for obj, err in ex.map(lambda s: _compile(s, force), srcs):
            if err:
                errors.append(err)
    if errors:
        raise RuntimeError("compilation failed:\n" + "\n".join(errors))
    lib_objs = [_obj(s) for s in LIB_SOURCES]
    if force or _stale(LIB, lib_objs):
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", LIB] + lib_objs + ["-lpthread"]
        subprocess.run(cmd, check=True)
        if verbose:
            print("built", LIB)
    for exe, src in TOOLS.items():
        out, path = os.path.join(LIB_DIR, exe), os.path.join(CSRC, src)
        if force or _stale(out, [path]):
            subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-o", out, path], check=True)
            if verbose:
                print("built", out)
    cli_objs = [_obj(s) for s in CLI_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    if cli_objs and (force or _stale(CLI, cli_objs + [LIB])):
        cmd = ["g++", "-o", CLI] + cli_objs + ["-L", LIB_DIR, "-lthv_hip",
                                               "-Wl,-rpath,$ORIGIN", "-lpthread"]
        subprocess.run(cmd, check=True)
        if verbose:
            print("built", CLI)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("-j", type=int, default=None)
    a = ap.parse_args()
    try:
        build(force=a.force, jobs=a.j)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
