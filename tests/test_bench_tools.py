"""bench.py's host-side tools (no GPU): the one-GPU contract, the HBM-traffic
summary of rocprofv3 counter files, the committed counter summaries it reads
back, and the sweep's algorithmic-bytes bound."""
import re
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "measures one GPU" in r.stderr


def _csv(path, counter, rows):
    import csv

    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Kernel_Name", "Counter_Name", "Counter_Value"])
        for k, v in rows:
            w.writerow([k, counter, v])


def test_summarize_pmc_scales_by_the_copy(tmp_path):
    """FETCH_SIZE and WRITE_SIZE (KiB) are scaled by what each reads for the
    1 GiB copy; the streaming read cross-checks the fetch scale."""
    kern = "void (anonymous namespace)::sweep_tile_kernel<1, 1, 8>(...)"
    _csv(tmp_path / "f.csv", "FETCH_SIZE", [(kern, 1000), (kern, 3000), (bench.COPY_KERNEL, 524288),
                                            (bench.READ_KERNEL, 2097152)])
    _csv(tmp_path / "w.csv", "WRITE_SIZE", [(kern, 500), (bench.COPY_KERNEL, 1048576)])
    out = bench.summarize_pmc(str(tmp_path / "f.csv"), str(tmp_path / "w.csv"))
    # The largest row of the kernel (one dispatch), times 2 for FETCH_SIZE.
    assert "| FETCH_SIZE | 3000 | 524288 | 2.000 | 6.1440e+06 |" in out
    assert "| WRITE_SIZE | 500 | 1048576 | 1.000 | 5.1200e+05 |" in out
    assert "scale 2.000" in out
    assert re.search(r"traffic_bytes_per_launch: 6\.656000e\+06", out)


@pytest.mark.parametrize("path", sorted((ROOT / "profiles").glob("sweep_tile_*_pmc.md")), ids=lambda p: p.name)
def test_committed_traffic_summaries_parse(path):
    """Each committed counter summary names its workload and launch size, and
    bench.py reads back the traffic its FETCH_SIZE and WRITE_SIZE rows add up
    to."""
    text = path.read_text()
    workload = re.search(r"^workload: (\S+)$", text, re.M).group(1)
    sources = int(re.search(r"^sources: (\d+)$", text, re.M).group(1))
    assert path.name == f"sweep_tile_{workload}_pmc.md"
    traffic = bench.committed_traffic(sources, workload)
    rows = dict(re.findall(r"^\| (FETCH_SIZE|WRITE_SIZE) \|.*\| ([0-9.e+]+) \|$", text, re.M))
    assert traffic and traffic == pytest.approx(float(rows["FETCH_SIZE"]) + float(rows["WRITE_SIZE"]), rel=1e-4)
    assert bench.committed_traffic(sources + 64, workload) is None  # another launch size: not this summary


def test_algorithmic_bytes_bound():
    # 8 per bag entry, 4 per node, 12 per node with descendants, 8 per such
    # node that is not one of the source's at most `height` ancestors.
    assert bench.sweep_algorithmic_bytes(n=10, bags=20, inner=4, height=3, sources=2) == 2 * (160 + 40 + 48 + 8)
    assert bench.sweep_algorithmic_bytes(n=10, bags=20, inner=2, height=3, sources=1) == 160 + 40 + 24
