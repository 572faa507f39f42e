"""C ABI surface, input contracts and CLI behaviour that need no GPU."""
import os
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import gpu_betweennesscentrality_amd as bcx
from gpu_betweennesscentrality_amd import graphgen as gg

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "gpu_betweennesscentrality_amd" / "bin"


def declared_functions():
    text = (ROOT / "include" / "bcx.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bcx_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) > 30
    L = bcx.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", str(ROOT / "gpu_betweennesscentrality_amd" / "libbcx.so")],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (bcx_\w+)", out))
    assert set(names) <= exported


def test_python_mirror_declares_every_symbol():
    assert set(declared_functions()) == set(bcx._SIGS)


def write(tmp_path, text, name="g.txt"):
    p = tmp_path / name
    p.write_text(text)
    return p


@pytest.mark.parametrize("text,match", [
    ("3 2\n1 2 5\n2 3 5\n", "no reverse"),
    ("2 2\n1 2 5\n2 1 6\n", "different weights"),
    ("2 2\n1 2 0\n2 1 0\n", "weight <= 0"),
    ("2 3\n1 2 5\n2 1 5\n", "truncated"),
    ("2 2\n1 3 5\n3 1 5\n", "out of range"),
    ("2 2\n1 1 5\n1 1 5\n", "self loop"),
    ("x y\n", "bad header"),
    ("2 4\n1 2 5\n2 1 5\n1 2 7\n2 1 7\n", "listed twice"),
])
def test_graph_reader_rejects_bad_input(tmp_path, text, match):
    with pytest.raises(bcx.BcxError, match=match):
        bcx.Graph.read(write(tmp_path, text))


def test_graph_reader_duplicates_first_wins(tmp_path):
    g = bcx.Graph.read(write(tmp_path, "2 4\n1 2 5\n2 1 5\n1 2 5\n2 1 5\n"))
    assert g.n == 2 and g.arcs == 2


def test_missing_file(tmp_path):
    with pytest.raises(bcx.BcxError, match="cannot open"):
        bcx.Graph.read(tmp_path / "nope.txt")


def test_written_graph_roundtrip(tmp_path):
    g = gg.make("grid1k")
    p = tmp_path / "g.txt"
    gg.write_graph(g, p)
    G = bcx.Graph.read(p)
    assert G.n == g.n and G.arcs == g.arcs and G.connected()
    head = p.read_text().splitlines()[:2]
    assert head[0] == f"{g.n} {g.arcs}"


def test_order_file_errors(tmp_path):
    g = gg.make("grid1k")
    G = bcx.Graph.from_arcs(g.n, g.u, g.v, g.w)
    bad = write(tmp_path, f"{g.n}\n0 0\n1 0\n", "o.txt")
    with pytest.raises(bcx.BcxError, match="repeated"):
        bcx.Index.build_order(G, bad)
    short = write(tmp_path, f"{g.n}\n0 0\n", "o2.txt")
    with pytest.raises(bcx.BcxError):
        bcx.Index.build_order(G, short)
    with pytest.raises(bcx.BcxError, match="permutation"):
        bcx.Index.build_order(G, np.zeros(g.n, np.int32))


def test_no_device_context_fails_loudly():
    from conftest import HAS_GPU

    if HAS_GPU:
        pytest.skip("device present")
    g = gg.make("grid1k")
    G = bcx.Graph.from_arcs(g.n, g.u, g.v, g.w)
    ix = bcx.Index.build_mhl(G)
    with pytest.raises(bcx.BcxError, match="no HIP device"):
        bcx.Context(G, ix, 0)


def test_graphgen_is_deterministic():
    a = gg.road_graph(2000, 2500, seed=7)
    b = gg.road_graph(2000, 2500, seed=7)
    assert np.array_equal(a.u, b.u) and np.array_equal(a.w, b.w)
    assert a.m == 2500 and a.arcs == 5000
    g = gg.make("grid1k")
    assert g.n == 1024 and g.m == 1834  # 1984 - 198 deleted + 48 diagonals
    assert g.w.min() >= 1 and g.w.max() <= 1000


@pytest.mark.parametrize("value", ["5:2", "x", "0:99999", "3:4:5", "-1:4"])
def test_cli_refuses_a_bad_source_range(tmp_path, value):
    """BCX_SOURCES is checked right after the graph is read, before any index
    or device work."""
    g = gg.make("grid1k")
    gp = tmp_path / "g.txt"
    gg.write_graph(g, gp)
    env = dict(os.environ, BCX_SOURCES=value)
    for cmd in ([BIN / "MHLQBC", gp, "0"], [BIN / "QBC", gp, tmp_path / "order.txt", "0"]):
        r = subprocess.run(cmd, capture_output=True, text=True, env=env)
        assert r.returncode == 1 and "BCX_SOURCES must be begin:end" in r.stderr, r.stderr


def test_cli_usage_and_query_file_contract(tmp_path):
    r = subprocess.run([BIN / "MHLQBC"], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr
    g = gg.make("grid1k")
    gp = tmp_path / "g.txt"
    gg.write_graph(g, gp)
    q = write(tmp_path, "5\n1\n2\n", "q.txt")
    r = subprocess.run([BIN / "RQ", gp, q, "0"], capture_output=True, text=True)
    assert r.returncode == 1
    assert "randNumber.size()<2  querySize=5" in r.stdout
    r = subprocess.run([BIN / "RQ", gp, tmp_path / "none.txt", "0"], capture_output=True, text=True)
    assert r.returncode == 1 and "Cannot open:" in r.stdout
    r = subprocess.run([BIN / "QBC", tmp_path / "none.txt", "o", "0"], capture_output=True, text=True)
    assert r.returncode == 1 and "cannot open" in r.stderr
