"""The square_root / parallel_cc / approx_cut / karger_stein command lines: same argv and CSV
shape as the reference executables (src/executables/*.cpp).  Argument and file
errors are checked on the CPU (they fail before any GPU call); the fixture runs
need a GPU."""
import subprocess
from pathlib import Path

import numpy as np
import pytest
from graphfile import write_graph

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "comm-avoiding-cuts-cc_amd" / "bin"
GOLD = Path(__file__).resolve().parent / "golden" / "reference_inputs"
F32_95 = float(np.float32(0.95))


def run(*args, check=False):
    r = subprocess.run([str(a) for a in args], capture_output=True, text=True, timeout=300)
    if check and r.returncode != 0:
        raise AssertionError(f"{args}: rc {r.returncode}\n{r.stdout}\n{r.stderr}")
    return r


@pytest.mark.parametrize("tool,nargs", [("square_root", 0), ("parallel_cc", 0), ("approx_cut", 0),
                                       ("karger_stein", 0)])
def test_usage(tool, nargs):
    r = run(BIN / tool)
    assert r.returncode == 1 and r.stdout.startswith("Usage:")


def test_missing_and_malformed_files(tmp_path):
    r = run(BIN / "square_root", "0.95", tmp_path / "nope.graph", "0")
    assert r.returncode == 1 and "cannot open" in r.stderr
    bad = tmp_path / "bad.graph"
    bad.write_text("# c\n3 2\n0 1 5\n1 x 2\n")
    for tool in ("square_root", "approx_cut"):
        r = run(BIN / tool, "0.95", bad, "0")
        assert r.returncode == 1 and "malformed" in r.stderr
    short = tmp_path / "short.graph"
    short.write_text("# c\n3 3\n0 1 5\n1 2 2\n")
    r = run(BIN / "parallel_cc", short, "0")
    assert r.returncode == 1 and "fewer edge lines" in r.stderr
    r = run(BIN / "square_root", "zz", GOLD / "cycle_4.graph", "0")
    assert r.returncode == 1
    r = run(BIN / "karger_stein", tmp_path / "nope.graph", "0")
    assert r.returncode == 1 and "cannot open" in r.stderr
    r = run(BIN / "karger_stein", GOLD / "cycle_4.graph", "x")
    assert r.returncode == 1


@pytest.mark.gpu
@pytest.mark.parametrize("name,cut", [("cycle_4.graph", 200), ("cycle_6.graph", 10000), ("cycle_10.graph", 10),
                                      ("cycle_20.graph", 100), ("complete_4.graph", 85)])
def test_square_root_fixtures(name, cut):
    out = run(BIN / "square_root", "0.95", GOLD / name, "0", check=True).stdout.strip().split(",")
    assert len(out) == 10 and out[0].endswith(name) and out[1] == "0" and out[2] == "1"
    assert out[8] == "low" and int(out[9]) == cut
    float(out[5]) and float(out[6]) == 0.0


@pytest.mark.gpu
def test_square_root_er1000(tmp_path, orc):
    from graphgen import dense, erdos_renyi

    rng = np.random.default_rng(0)
    e = erdos_renyi(rng, 1000, 0.05, wmax=1)
    f = tmp_path / "er1000.in"
    write_graph(f, 1000, e)
    out = run(BIN / "square_root", "0.95", f, "0", check=True).stdout.strip().split(",")
    assert int(out[3]) == 1000 and int(out[4]) == len(e)
    assert int(out[7]) == orc.sqrt_trials(1000, len(e), F32_95)
    assert int(out[9]) == orc.stoer_wagner(dense(1000, e))


@pytest.mark.gpu
def test_parallel_cc_fixtures():
    for name in ("cc_8.graph", "cc_13.graph"):
        lines = run(BIN / "parallel_cc", GOLD / name, "7", "3", check=True).stdout.strip().splitlines()
        assert len(lines) == 3
        for k, line in enumerate(lines):
            f = line.split(",")
            assert len(f) == 9 and f[1] == str(7 + k) and f[7] == "cc" and f[8] == "5"


@pytest.mark.gpu
def test_approx_cut_cli_matches_oracle(orc):
    from graphfile import read_graph

    n, e = read_graph(GOLD / "cycle_10.graph")
    out = run(BIN / "approx_cut", "0.95", GOLD / "cycle_10.graph", "3", check=True).stdout.strip().split(",")
    want = orc.approx_cut(n, e, F32_95, 3)
    assert len(out) == 9 and (int(out[8]), int(out[7])) == want


@pytest.mark.gpu
def test_click_mode():
    out = run(BIN / "square_root", "0.95", "CLICK", "60", "1", check=True).stdout.strip().split(",")
    assert out[0] == "CLICK" and int(out[3]) == 60 and int(out[4]) == 60 * 59 // 2 and int(out[9]) >= 0


@pytest.mark.gpu
@pytest.mark.parametrize("name,cut", [("cycle_4.graph", 200), ("cycle_6.graph", 10000), ("cycle_10.graph", 10),
                                      ("cycle_20.graph", 100), ("complete_4.graph", 85)])
def test_karger_stein_fixtures(name, cut):
    out = run(BIN / "karger_stein", GOLD / name, "5", check=True).stdout.strip().split(",")
    assert len(out) == 9 and out[0].endswith(name) and out[1] == "5" and out[2] == "1"
    assert out[6] == "?" and out[7] == "co-karger-stein" and int(out[8]) == cut
    float(out[5])


@pytest.mark.gpu
def test_karger_stein_sparse_ws_matches_oracle(tmp_path, orc):
    # Watts-Strogatz, n = 3000, 8 ring neighbours a side: the sparse recursion
    import cac

    n = 3000
    e = cac.gen_ws(n, 8, 4)
    E = np.stack([e["u"], e["v"], e["w"]], 1).astype(np.uint64)
    f = tmp_path / "ws3000.in"
    write_graph(f, n, E)
    out = run(BIN / "karger_stein", f, "6", check=True).stdout.strip().split(",")
    T = orc.ks_trials(n, 0.90)
    assert int(out[3]) == n and int(out[4]) == len(E)
    assert int(out[8]) == min(orc.ks_sparse_trial(n, E, 6, t) for t in range(T))


def test_edge_list_cache(tmp_path):
    # CAC_EDGE_CACHE: the first run writes <file>.cacbin holding the parsed
    # edges; a later run reads it while the text keeps its size and mtime,
    # shown by garbling the text in place (same size, mtime restored): with the
    # cache the garbage is not seen, without it the parse fails.  The file is
    # read before any GPU call, so this runs on the CPU (where the tool then
    # stops for want of a GPU).
    import os
    import shutil

    src = tmp_path / "g.graph"
    shutil.copy(GOLD / "cc_13.graph", src)
    env = dict(os.environ, CAC_EDGE_CACHE="1")

    def tool(use_cache):
        return subprocess.run([str(BIN / "parallel_cc"), str(src), "0"], capture_output=True, text=True, timeout=300,
                              env=env if use_cache else {k: v for k, v in os.environ.items() if k != "CAC_EDGE_CACHE"})

    tool(True)
    cache = tmp_path / "g.graph.cacbin"
    raw = cache.read_bytes()
    assert raw[:8] == b"CACEDGE1"
    n, m = int.from_bytes(raw[8:12], "little"), int.from_bytes(raw[16:24], "little")
    lines = src.read_text().splitlines()
    want_n, want_m = (int(x) for x in lines[1].split())
    want = np.array([[int(x) for x in ln.split()] for ln in lines[2:2 + want_m]], dtype=np.uint64)
    rec = np.frombuffer(raw[64:], dtype=np.dtype([("u", "<u4"), ("v", "<u4"), ("w", "<u8")]))
    assert (n, m) == (want_n, want_m) and len(rec) == m
    assert np.array_equal(rec["u"], want[:, 0]) and np.array_equal(rec["v"], want[:, 1])
    assert np.array_equal(rec["w"], want[:, 2])
    st = src.stat()
    text = src.read_bytes()
    body = text.index(b"\n", text.index(b"\n") + 1) + 1
    garbled = text[:body] + text[body:].replace(b"1", b"x", 1)
    assert len(garbled) == len(text)
    src.write_bytes(garbled)
    os.utime(src, ns=(st.st_atime_ns, st.st_mtime_ns))
    assert "malformed" not in tool(True).stderr
    assert "malformed" in tool(False).stderr
