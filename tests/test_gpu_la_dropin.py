"""Drop-in on the reference's own shipped inputs (README.md:53-98): the
LosAngeles graph, its order file and its query batch (data/roadData,
data/orderFile, data/queryBatch, stored as arrays in la_graph.npz), run
through the three executables and the C ABI, against the oracle's BC of
every vertex (la_bc.npz, all 49,374 sources)."""
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

import gpu_betweennesscentrality_amd as bcx
from gpu_betweennesscentrality_amd import graphgen as gg

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent
GOLDEN = ROOT / "tests" / "golden"
BIN = ROOT / "gpu_betweennesscentrality_amd" / "bin"


def close_bc(a, b, tol=1e-9):
    return bool(np.all(np.abs(a - b) <= tol * np.maximum(np.abs(b), 1.0)))


@pytest.fixture(scope="module")
def la(tmp_path_factory):
    d = tmp_path_factory.mktemp("la")
    a = np.load(GOLDEN / "la_graph.npz")
    n = int(a["n"])
    g = gg.ArcGraph(n, a["u"], a["v"], a["w"])
    gp = d / "LosAngeles.txt"
    gg.write_graph(g, gp)  # arcs in the shipped file's order
    op = d / "LosAngelesOrder.txt"
    op.write_text(f"{n}\n" + "".join(f"{v} {r}\n" for v, r in enumerate(a["order"].tolist())))
    qp = d / "queries.txt"
    gg.write_queries(a["queries"], qp)
    ref = np.load(GOLDEN / "la_bc.npz")["bc"]
    return dict(n=n, g=g, graph=gp, order=op, queries=qp, q=a["queries"], ref=ref, dir=d)


def test_mhlqbc_full_bc(la):
    env = dict(os.environ, BCX_BC_DUMP=str(la["dir"] / "mhl.bin"))
    out = subprocess.run([BIN / "MHLQBC", la["graph"], "0", la["dir"] / "mhl.txt"], env=env,
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    assert "seconds (---totalTime---)" in out.stdout
    got = np.fromfile(la["dir"] / "mhl.bin", np.float64)
    assert close_bc(got, la["ref"])
    lines = (la["dir"] / "mhl.txt").read_text().splitlines()
    assert lines[0] == "vertexID  BCValue" and len(lines) == la["n"] + 1
    k = int(np.argmax(la["ref"]))
    assert lines[k + 1] == f"{k + 1} {la['ref'][k]:.2f}"


def test_qbc_order_file_and_check_file(la):
    check = la["dir"] / "check.txt"
    check.write_text(f"{la['n']} {la['g'].arcs}\n" +
                     "".join(f"{v + 1} {x:.6f}\n" for v, x in enumerate(la["ref"].tolist())))
    env = dict(os.environ, BCX_BC_DUMP=str(la["dir"] / "qbc.bin"))
    out = subprocess.run([BIN / "QBC", la["graph"], la["order"], "0", la["dir"] / "qbc.txt", check], env=env,
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "BC value is correct" in out.stdout
    assert close_bc(np.fromfile(la["dir"] / "qbc.bin", np.float64), la["ref"])


def test_rq_query_batch_table(la):
    out = subprocess.run([BIN / "RQ", la["graph"], la["queries"], "0"], capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    q = la["q"]
    assert f"The number of vertices to query: {len(q)}" in out.stdout
    lines = out.stdout.splitlines()
    k = lines.index("-" * 20)
    rows = [ln.split() for ln in lines[k + 1:k + 1 + len(q)]]
    assert [int(r[0]) for r in rows] == (q + 1).tolist()
    got = np.array([float(r[1]) for r in rows])
    want = la["ref"][q]
    assert np.all(np.abs(got - want) <= 1e-9 * np.maximum(np.abs(want), 1.0) + 5e-9)


def test_full_bc_every_form_through_the_abi(la):
    G = bcx.Graph.read(la["graph"])
    ctx = bcx.Context(G, bcx.Index.build_mhl(G), 0)
    for mode in (bcx.MODE_SWEEP, bcx.MODE_BATCHED):
        bc, ovf = ctx.full_bc(mode)
        assert not ovf and close_bc(bc, la["ref"]), f"mode {mode}"
    ctx.close()
