"""Parity at the 100,000-vertex road-graph configuration (plain QBC, all
sources; MHLQBC.cu:495-537 over its label index): the bag sweep and the
label scan against the oracle, their distance/count rows bit for bit, and a
full BC of every source in both forms."""
import numpy as np
import pytest

import gpu_betweennesscentrality_amd as bcx
import oracle
from _bigcfg import close_bc, max_rel, oracle_rows, rows_equal, sample_sources, workload

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cfg():
    g, csr, G, ix, ctx = workload("road100k")
    yield g, csr, ix, ctx
    ctx.close()


def test_sampled_sources_both_forms_vs_oracle(cfg):
    g, csr, ix, ctx = cfg
    src = sample_sources(ix, g.n, 256, 256, seed=41)
    want, _ = oracle.brandes(csr, src, threads=oracle.default_threads())
    for mode in (bcx.MODE_SWEEP, bcx.MODE_BATCHED):
        got, ovf = ctx.bc_sources(src, mode)
        assert not ovf
        assert close_bc(got, want), f"mode {mode}: max rel err {max_rel(got, want):.3e}"


def test_rows_bit_exact_every_form(cfg):
    g, csr, ix, ctx = cfg
    src = sample_sources(ix, g.n, 64, 16, seed=42)
    want = oracle_rows(csr, src)
    assert rows_equal(ctx.query_rows(src, bcx.MODE_BATCHED), want) is None
    for form in (bcx.SWEEP_FORM_ONE_PER_CU, bcx.SWEEP_FORM_TWO_PER_CU):
        ctx.set_tuning(sweep_form=form)
        err = rows_equal(ctx.query_rows(src, bcx.MODE_SWEEP), want)
        assert err is None, f"form {form}: {err}"
    ctx.set_tuning(sweep_form=bcx.SWEEP_FORM_AUTO)


def test_full_bc_sweep_equals_label_scan(cfg):
    """All 100,000 sources: the bag sweep (one launch of 1,563 tiles on 512
    state slots) against the reference's label-scan method."""
    g, csr, ix, ctx = cfg
    sweep, o1 = ctx.full_bc(bcx.MODE_SWEEP)
    scan, o2 = ctx.full_bc(bcx.MODE_BATCHED)
    assert not o1 and not o2
    assert close_bc(sweep, scan), f"max rel err {max_rel(sweep, scan):.3e}"
    assert np.all(sweep >= 0)
