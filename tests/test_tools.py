"""Format tooling (CPU): thv_convert reproduces the reference converter's
bytes; thv_compare applies the reference comparer's rules."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "thievory_amd", "_lib")
REF_CONVERT = os.path.join(ROOT, "oracle", "_ref", "convert")


@pytest.fixture(scope="module")
def tools():
    conv, cmpb = os.path.join(LIB, "thv_convert"), os.path.join(LIB, "thv_compare")
    if not (os.path.exists(conv) and os.path.exists(cmpb)):
        import thievory_amd.build as B
        B.build(verbose=False)
    return conv, cmpb


def _convert(conv, cwd, name, *extra):
    subprocess.run([conv, "--input", name, "--skip", "1", "--type", "all", *map(str, extra)],
                   cwd=cwd, check=True, capture_output=True)


@pytest.mark.parametrize("name", ["tiny4", "rand64"])
def test_convert_matches_committed_fixtures(tools, tmp_path, golden, name):
    shutil.copy(os.path.join(golden, f"{name}.txt"), tmp_path / f"{name}.txt")
    _convert(tools[0], tmp_path, f"{name}.txt")
    for ext in ("csr", "csc", "wcsr"):
        assert (tmp_path / f"{name}.{ext}").read_bytes() == \
            open(os.path.join(golden, f"{name}.{ext}"), "rb").read(), ext


@pytest.mark.skipif(not os.path.exists(REF_CONVERT), reason="reference converter not built")
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_convert_matches_reference_on_odd_inputs(tools, tmp_path, seed):
    rng = np.random.default_rng(seed)
    lines = ["# header"]
    for _ in range(400):
        r = rng.random()
        if r < 0.05:
            lines.append("")                       # blank line
        elif r < 0.08:
            lines.append("% comment in the middle")
        elif r < 0.1:
            lines.append(f"{rng.integers(0, 90)} {rng.integers(0, 90)} 7")  # extra column
        else:
            lines.append(f"{rng.integers(0, 90)} {rng.integers(0, 90)}")
    text = "\n".join(lines) + "\n"
    (tmp_path / "a.txt").write_text(text)
    (tmp_path / "b.txt").write_text(text)
    _convert(tools[0], tmp_path, "a.txt")
    for t in ("csr", "csc", "wcsr"):
        subprocess.run([REF_CONVERT, "--input", "b.txt", "--skip", "1", "--type", t],
                       cwd=tmp_path, check=True, capture_output=True)
    assert (tmp_path / "a.csr").read_bytes() == (tmp_path / "b.csr").read_bytes()
    assert (tmp_path / "a.csc").read_bytes() == (tmp_path / "b.csc").read_bytes()
    # the reference names its interleaved output "<base>..wcsr"
    assert (tmp_path / "a.wcsr").read_bytes() == (tmp_path / "b..wcsr").read_bytes()


def test_convert_weighted_and_undirected(tools, tmp_path):
    (tmp_path / "w.txt").write_text("# h\n0 1 5\n1 2 9\n2 0 3\n")
    _convert(tools[0], tmp_path, "w.txt", "--weighted", 1, "--undirected", 1)
    g = O.read_graph(str(tmp_path / "w.csr"), "csr")
    assert g.n == 3 and g.m == 6
    assert g.off.tolist() == [0, 2, 4, 6]
    assert g.adj.tolist() == [1, 2, 0, 2, 0, 1]
    assert g.w.tolist() == [5, 3, 5, 9, 3, 9]
    gw = O.read_graph(str(tmp_path / "w.wcsr"), "wcsr")
    assert gw.adj.tolist() == g.adj.tolist() and gw.w.tolist() == g.w.tolist()
    c = O.read_graph(str(tmp_path / "w.csc"), "csc")
    t = O.transpose(g)
    assert c.off.tolist() == t.off.tolist() and c.adj.tolist() == t.adj.tolist()


def _cmp(cmpb, a, b, t, n):
    return subprocess.run([cmpb, "--input1", str(a), "--input2", str(b), "--type", str(t),
                           "--vertices", str(n)], capture_output=True, text=True)


def test_compare_rules(tools, tmp_path):
    cmpb = tools[1]
    a = np.array([1, 2, 3, 4], np.uint32)
    a.tofile(tmp_path / "a")
    a.tofile(tmp_path / "b")
    assert _cmp(cmpb, tmp_path / "a", tmp_path / "b", 0, 4).returncode == 0
    b = a.copy()
    b[2] = 9
    b.tofile(tmp_path / "b")
    r = _cmp(cmpb, tmp_path / "a", tmp_path / "b", 0, 4)
    assert r.returncode == 1 and "Mismatches: 1 of 4" in r.stdout and "at: 2" in r.stdout
    x = np.array([1.0, 2.0], np.float64)
    x.tofile(tmp_path / "x")
    (x + np.array([5e-4, 0.0])).tofile(tmp_path / "y")
    assert _cmp(cmpb, tmp_path / "x", tmp_path / "y", 2, 2).returncode == 0  # |d| <= 1e-3
    (x + np.array([2e-3, 0.0])).tofile(tmp_path / "y")
    assert _cmp(cmpb, tmp_path / "x", tmp_path / "y", 2, 2).returncode == 1
    u = np.array([1, 2], np.uint64)
    u.tofile(tmp_path / "u")
    assert _cmp(cmpb, tmp_path / "u", tmp_path / "u", 1, 2).returncode == 0
    assert _cmp(cmpb, tmp_path / "u", tmp_path / "u", 1, 3).returncode == 2  # short file


def test_profsum_tables(tmp_path, capsys):
    """The rocprofv3 CSV summariser keeps template instances apart and
    divides counters by the calibration factor."""
    from thievory_amd import profsum
    tr = tmp_path / "t_kernel_trace.csv"
    tr.write_text("Kernel_Name,Start_Timestamp,End_Timestamp\n"
                  "k_expand<0>,0,1000000\nk_expand<0>,0,3000000\nk_expand<5>,10,20\n")
    profsum.main(["trace", str(tr), "--match", "k_expand"])
    out = capsys.readouterr().out
    assert "| `k_expand<0>` | 2 | 4.000 | 2.0000 |" in out
    assert "`k_expand<5>` | 1 |" in out
    pm = tmp_path / "p_counter_collection.csv"
    pm.write_text("Dispatch_Id,Kernel_Name,Counter_Name,Counter_Value\n"
                  "1,k,FETCH_SIZE,100\n2,k,FETCH_SIZE,300\n")
    profsum.main(["pmc", str(pm), "--factor", "0.5"])
    out = capsys.readouterr().out
    assert "| `k` | FETCH_SIZE | 2 | 400 | 200.0 | 400.0 |" in out
    st = tmp_path / "s_kernel_stats.csv"
    st.write_text("Name,Calls,TotalDurationNs,AverageNs\nk,4,8000000,2000000\n")
    profsum.main(["stats", str(st)])
    assert "| `k` | 4 | 8.000 | 2.0000 | 100.0% |" in capsys.readouterr().out
