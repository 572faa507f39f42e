"""Parity at the 1,070,000-vertex road-graph configuration (batched MHL-QBC
with arterials): sampled sources against the oracle while the label-scan
workspace is held, a job whose launch is cut short by device memory (the
workspace released, whole tiles per compute unit, several launches) against
the same sources in launches that fit, and the forward pass's rows bit for
bit."""
import numpy as np
import pytest

import gpu_betweennesscentrality_amd as bcx
import oracle
from _bigcfg import close_bc, max_rel, oracle_rows, rows_equal, sample_sources, workload

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cfg():
    g, csr, G, ix, ctx = workload("road1070k")
    yield g, csr, ix, ctx
    ctx.close()


def test_sampled_sources_vs_oracle_with_workspace_held(cfg):
    g, csr, ix, ctx = cfg
    tree = ix.tree_order()
    ctx.bc_sources(tree[:4096].astype(np.int32), bcx.MODE_BATCHED)  # allocates the workspace
    src = sample_sources(ix, g.n, 0, 64, seed=51)
    want, _ = oracle.brandes(csr, src, threads=oracle.default_threads())
    got, ovf = ctx.bc_sources(src, bcx.MODE_SWEEP)
    assert not ovf
    assert close_bc(got, want), f"max rel err {max_rel(got, want):.3e}"


def test_memory_capped_job_equals_fitting_launches(cfg):
    g, csr, ix, ctx = cfg
    tree = ix.tree_order()
    _, cus = ctx.sweep_slots()
    default = ctx.tuning()["sweep_batch"]
    batch = 4 * cus * 64  # four tiles per compute unit: more than device memory holds at this size
    ctx.set_tuning(sweep_batch=batch)
    ctx.bc_sources(tree[:4096].astype(np.int32), bcx.MODE_BATCHED)  # holds the workspace
    src = tree[:batch].astype(np.int32)
    ctx.reset_timing()
    capped, o1 = ctx.bc_sources(src, bcx.MODE_SWEEP)
    launches = ctx.timing()["scan_launches"]
    print(f"[cfg] capped job: {launches} launches for {batch // 64} tiles", flush=True)
    assert launches >= 2, "the job must not fit one launch at this size"
    parts = np.zeros(g.n)
    step = 64 * cus
    for k in range(0, batch, step):
        part, o2 = ctx.bc_sources(src[k:k + step], bcx.MODE_SWEEP)
        assert not o2
        parts += part
    ctx.set_tuning(sweep_batch=default)
    assert not o1
    assert close_bc(capped, parts), f"max rel err {max_rel(capped, parts):.3e}"


def test_rows_bit_exact(cfg):
    g, csr, ix, ctx = cfg
    src = sample_sources(ix, g.n, 64, 0, seed=52)
    want = oracle_rows(csr, src)
    err = rows_equal(ctx.query_rows(src, bcx.MODE_SWEEP), want)
    assert err is None, err
