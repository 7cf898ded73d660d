"""The drop-in CLI (thv_main): same flags and results/values.bin contract as
the reference's `main` (src/main.cu, src/graph.cuh:817-852)."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "thievory_amd", "_lib", "thv_main")


@pytest.fixture(scope="module")
def cli():
    if not os.path.exists(CLI):
        import thievory_amd.build as B
        B.build(verbose=False)
    return CLI


def _run(cli, cwd, *args):
    r = subprocess.run([cli] + [str(a) for a in args], cwd=cwd, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def _values(cwd, dtype, n):
    v = np.fromfile(os.path.join(cwd, "results", "values.bin"), dtype=dtype)
    assert len(v) == n
    return v


def test_cli_without_input_exits_quietly(cli, tmp_path):
    r = subprocess.run([cli], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout == ""


@pytest.mark.gpu
def test_cli_tiny4(cli, tmp_path, golden):
    (tmp_path / "results").mkdir()
    f = os.path.join(golden, "tiny4.csr")
    out = _run(cli, tmp_path, "--input", f, "--algo", "bfs", "--source", 1, "--runs", 2)
    assert "Average execution time:" in out and "Num Vertices: 4" in out
    assert _values(tmp_path, np.uint32, 4).tolist() == [3, 1, 2, 5]
    _run(cli, tmp_path, "--input", f, "--algo", "cc")
    assert _values(tmp_path, np.uint32, 4).tolist() == [0, 0, 0, 3]
    _run(cli, tmp_path, "--input", f, "--algo", "sssp", "--source", 1)
    assert _values(tmp_path, np.uint32, 4).tolist() == [5, 1, 5, 5]
    _run(cli, tmp_path, "--input", os.path.join(golden, "tiny4.wcsr"), "--algo", "sssp",
         "--edgeSize", 8, "--source", 1)
    assert _values(tmp_path, np.uint32, 4).tolist() == [5, 1, 5, 5]


@pytest.mark.gpu
def test_cli_rand64_split_store_and_pagerank(cli, tmp_path, golden):
    (tmp_path / "results").mkdir()
    og = O.read_graph(os.path.join(golden, "rand64.csr"), "csr")
    f = os.path.join(golden, "rand64.csr")
    for src in (0, 9, 63):
        _run(cli, tmp_path, "--input", f, "--algo", "bfs", "--source", src, "--budget", 0.0,
             "--partition", 16)
        assert np.array_equal(_values(tmp_path, np.uint32, og.n), O.bfs(og, src))
        _run(cli, tmp_path, "--input", f, "--algo", "sssp", "--source", src, "--budget", 0.3,
             "--partition", 16)
        assert np.array_equal(_values(tmp_path, np.uint32, og.n), O.sssp(og, src))
    ref, _ = O.pr_exact(og)
    _run(cli, tmp_path, "--input", f, "--algo", "pr", "--type", "push")
    assert O.normalized_l1(_values(tmp_path, np.float64, og.n), ref) < 1e-5
    _run(cli, tmp_path, "--input", os.path.join(golden, "rand64.csc"), "--algo", "pr",
         "--type", "pull")
    assert O.normalized_l1(_values(tmp_path, np.float64, og.n), ref) < 1e-5


@pytest.mark.gpu
def test_cli_no_results_dir(cli, tmp_path, golden):
    out = _run(cli, tmp_path, "--input", os.path.join(golden, "tiny4.csr"))
    assert "Average execution time:" in out
    assert not (tmp_path / "results").exists()


@pytest.mark.gpu
def test_cli_bad_file(cli, tmp_path):
    bad = tmp_path / "bad.csr"
    bad.write_bytes(np.array([4, 100], np.uint64).tobytes())
    r = subprocess.run([cli, "--input", str(bad)], cwd=tmp_path, capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1 and "EFORMAT" not in r.stdout
    assert "file size" in r.stderr


@pytest.mark.gpu
def test_cli_symmetric_flag(cli, tmp_path):
    """--symmetric 1 on an undirected file (every edge both ways): bottom-up
    BFS levels and union-find CC give the oracle's values."""
    (tmp_path / "results").mkdir()
    og = O.rmat(12, 8, seed=21)
    path = str(tmp_path / "sym.csr")
    O.write_graph(path, "csr", og)
    for src in (0, 1, 777):
        _run(cli, tmp_path, "--input", path, "--algo", "bfs", "--source", src,
             "--symmetric", 1, "--budget", 0.25, "--partition", 256)
        assert np.array_equal(_values(tmp_path, np.uint32, og.n), O.bfs(og, src))
    _run(cli, tmp_path, "--input", path, "--algo", "cc", "--symmetric", 1, "--budget", 0.25)
    assert np.array_equal(_values(tmp_path, np.uint32, og.n), O.cc(og))
