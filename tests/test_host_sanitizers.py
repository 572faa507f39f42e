"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the graph
reader, both index builders, the H2H build, the index image and both
host-array seams (with damaged input), and the CPU oracle, driven by
tests/host_sanitize_driver.cpp on the seeded grid, a tie-heavy grid, the
saturating 35x35 unit grid and the reference's LosAngeles graph with its
shipped order file.  Built with g++/gcc from the sources (no device code)."""
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

from gpu_betweennesscentrality_amd import graphgen as gg

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "gpu_betweennesscentrality_amd" / "csrc"
FLAGS = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
         "-fopenmp"]


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    d = tmp_path_factory.mktemp("san")
    obj = d / "oracle.o"
    subprocess.run(["gcc", "-std=c11", *FLAGS, "-c", str(ROOT / "oracle" / "oracle.c"), "-o", str(obj)],
                   check=True)
    exe = d / "host_sanitize_driver"
    subprocess.run(["g++", "-std=c++17", *FLAGS, f"-I{ROOT / 'include'}", f"-I{CSRC}",
                    str(ROOT / "tests" / "host_sanitize_driver.cpp"), str(CSRC / "graph_io.cpp"),
                    str(CSRC / "index_build.cpp"), str(obj), "-o", str(exe), "-lm"], check=True)
    return exe


def _write_order(path, order):
    with open(path, "w") as f:
        f.write(f"{len(order)}\n")
        for v, r in enumerate(np.asarray(order).tolist()):
            f.write(f"{v} {r}\n")


CASES = {
    "grid1k": lambda: (gg.make("grid1k"), None),
    "tie_grid": lambda: (gg.grid_graph(k=20, seed=3, wmax=2), None),
    "unit_grid35": lambda: (gg.unit_grid(35), None),
}


def _run(driver, tmp_path, graph_path, order_path):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="4")
    r = subprocess.run([str(driver), str(graph_path), str(order_path) if order_path else "-", str(tmp_path)],
                       capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0 and r.stdout.startswith("ok"), (r.stdout[-2000:] + r.stderr[-4000:])


@pytest.mark.parametrize("name", sorted(CASES))
def test_generated_graphs(driver, tmp_path, name):
    g, _ = CASES[name]()
    gp = tmp_path / "g.txt"
    gg.write_graph(g, gp)
    _run(driver, tmp_path, gp, None)


def test_losangeles_with_shipped_order(driver, tmp_path, golden_dir):
    la = np.load(golden_dir / "la_graph.npz")
    gp, op = tmp_path / "la.txt", tmp_path / "la_order.txt"
    g = gg.ArcGraph(int(la["n"]), la["u"], la["v"], la["w"])
    gg.write_graph(g, gp)
    _write_order(op, la["order"])
    _run(driver, tmp_path, gp, op)
