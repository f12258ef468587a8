"""Robustness of the device context: admission that fails part way, shared
geometry under uploads, edited host tables, a case whose dt aborts, and the
ghost-depth contract.  Each check compares with the case run alone (bitwise)
or with the documented error."""
import ctypes as C

import numpy as np
import pytest

from beta1_amd import HostCase, MhdContext, OTDesc, WindDesc
from beta1_amd.capi import MhdError, MhdOptions, lib

pytestmark = pytest.mark.gpu

FIELDS = ("U_H", "B_xi", "B_eta", "B_zeta")


def alone(desc, steps, refresh_first=False, **kw):
    ctx = MhdContext(**kw)
    cid = ctx.add(desc)
    ctx.finalize()
    if refresh_first:
        ctx.refresh()
    ctx.step(steps)
    nb = ctx.info(cid)[0]
    out = {(f, b): ctx.field(cid, f, b) for f in FIELDS for b in range(nb)}
    ctx.close()
    return out


def same(ctx, cid, ref):
    for (f, b), a in ref.items():
        assert np.array_equal(ctx.field(cid, f, b), a), (cid, f, b)


def test_failed_admission_leaves_the_captured_step_valid():
    """Eight cases fill the per-case tables; a ninth that exceeds the HBM
    budget makes them grow before the budget check fails.  The captured
    graph named the old tables: the next graph step must re-capture and still
    equal eager stepping."""
    descs = [OTDesc(n=8, seed=s) for s in range(8)]
    probe = MhdContext()
    for d in descs:
        probe.add(d)
    probe.finalize()
    need = probe.memory_bytes()
    probe.close()
    ctx = MhdContext(hbm_budget_gb=1.02 * need / 1e9)
    ids = [ctx.add(d) for d in descs]
    ctx.finalize()
    ctx.step(4, graph=True)
    assert ctx.try_add(WindDesc(nang=16, nrad=32)) is None  # over budget
    assert ctx.try_add(OTDesc(n=8, seed=99)) is None        # its state alone is over budget too
    ctx.step(4, graph=True)
    for d, cid in zip(descs, ids):
        same(ctx, cid, alone(d, 8))


def test_geometry_upload_is_copy_on_write():
    """Two resident cases of one grid share a geometry copy; an upload of a
    geometry field into one of them gives that case its own copy and leaves
    the other stepping exactly as alone.  The face metric the B_cell weights
    are formed from cannot be replaced."""
    a, b = WindDesc(nang=8, nrad=16, v_kms=350.0), WindDesc(nang=8, nrad=16, n_cc=12.0)
    ctx = MhdContext()
    ia, ib = ctx.add(a), ctx.add(b)
    ctx.finalize()
    shared = ctx.memory_bytes()
    beta = ctx.field(ia, "beta_xi", 0)
    ctx.set_field(ia, "beta_xi", 0, beta * 1.25)
    assert ctx.memory_bytes() > shared                     # a private copy now
    assert np.array_equal(ctx.field(ia, "beta_xi", 0), beta * 1.25)
    assert np.array_equal(ctx.field(ib, "beta_xi", 0), beta)
    ctx.step(3)
    same(ctx, ib, alone(b, 3))
    with pytest.raises(MhdError):
        ctx.set_field(ib, "JDxi", 0, ctx.field(ib, "JDxi", 0))
    # the edited case's steps follow its new Hodge weights (J edges change)
    ref = alone(a, 3)
    assert not np.array_equal(ctx.field(ia, "B_eta", 0), ref[("B_eta", 0)])


def test_geometry_is_shared_by_content_not_by_name():
    """Cases share a geometry copy exactly when their grids are equal:
    shells whose outer radius differs in the 10th digit do not."""
    def mem(descs):
        ctx = MhdContext()
        for d in descs:
            ctx.add(d)
        ctx.finalize()
        m = ctx.memory_bytes()
        ctx.close()
        return m
    one = mem([WindDesc(nang=8, nrad=16)])
    two_same = mem([WindDesc(nang=8, nrad=16), WindDesc(nang=8, nrad=16, v_kms=600.0)])
    two_diff = mem([WindDesc(nang=8, nrad=16), WindDesc(nang=8, nrad=16, r_out=8.0 + 1e-9)])
    assert two_same < 1.7 * one < two_diff


def test_edited_weight_tables_are_not_dropped_silently():
    """A host case whose weight table was replaced is refused (the device
    forms the tables from the grid's least-squares data); re-uploading the
    builder's own table is accepted."""
    desc = WindDesc(nang=8, nrad=16)
    hc = HostCase(desc)
    w = hc.field("Jcell_from_Jedge_w", 2)
    hc.set_field("Jcell_from_Jedge_w", 2, w * 1.01)
    ctx = MhdContext()
    ctx.finalize()
    with pytest.raises(MhdError):
        ctx.adopt(hc)
    hc.set_field("Jcell_from_Jedge_w", 2, w)
    cid = ctx.adopt(hc)
    assert cid is not None
    ctx.step(2)
    same(ctx, cid, alone(desc, 2))


def poisoned(ctx, cid):
    """One cell of block 0 with a velocity of 1e9: the case's dt falls below
    the abort floor (all values stay finite in the first step)."""
    U = ctx.field(cid, "U_H", 0)
    g = ctx.ghost
    U[g + 3, g + 3, g + 5, 1] = 1e9 * U[g + 3, g + 3, g + 5, 0]
    U[g + 3, g + 3, g + 5, 4] = 1e19
    ctx.set_field(cid, "U_H", 0, U)
    ctx.refresh()


def test_dt_abort_is_reported_per_case_and_cleared_on_retire():
    good, bad = WindDesc(nang=8, nrad=16, v_kms=500.0), WindDesc(nang=8, nrad=16)
    ctx = MhdContext()
    ig, ibad = ctx.add(good), ctx.add(bad)
    ctx.finalize()
    poisoned(ctx, ibad)
    with pytest.raises(MhdError) as e:
        ctx.step(1)
    assert e.value.code == -3
    assert ctx.case_error(ibad) and not ctx.case_error(ig)
    ctx.retire(ibad)
    ctx.step(2)                       # no error once the failed case left
    # (the refresh after the upload synced the good case's ghosts as well)
    same(ctx, ig, alone(good, 3, refresh_first=True))
    assert not ctx.case_error(ig)


def test_server_retires_a_failing_request_and_serves_the_rest():
    from beta1_amd.serve import Request, ServingEngine

    descs = [WindDesc(nang=8, nrad=16, v_kms=420.0), WindDesc(nang=8, nrad=16, v_kms=2.0e9),
             WindDesc(nang=8, nrad=16, n_cc=9.0)]
    reqs = [Request(i, d, steps=6, arrival=0.0) for i, d in enumerate(descs)]
    eng = ServingEngine(precision="fp64", hbm_budget_gb=0.0, chunk=3, workers=2, keep_results=True)
    out = eng.run(reqs)
    eng.close()
    assert out["requests"] == 3 and out["failed"] == 1
    for r in out["records"]:
        if r.rid == 1:
            assert r.failed and not r.result
        else:
            assert not r.failed
            for key, a in alone(r.desc, r.steps).items():
                assert np.array_equal(r.result[key], a), (r.rid, key)


def test_ghost_depth_below_two_is_rejected():
    opt = MhdOptions(0, 0, 0, 1, 0.0, 0, 1, 1)
    ctx = C.c_void_p()
    assert lib().mhd_create(C.byref(opt), C.byref(ctx)) == -1


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 1e-5)])
def test_served_results_match_the_oracle(precision, tol):
    """Requests served by continuous batching (Lunar and two-fluid Mercury,
    one of them with the conducting interior r_in = 0.85, mixed sizes, a
    budget that holds about two at a time) against the FP64 oracle run for
    the same number of steps."""
    from beta1_amd import MercuryDesc
    from beta1_amd.fields import interior
    from beta1_amd.serve import Request, ServingEngine
    from oracle_bind import Oracle

    descs = [WindDesc(nang=8, nrad=16, v_kms=380.0), MercuryDesc(nang=8, nrad=16, n_cc=30.0),
             WindDesc(nang=12, nrad=16, n_cc=9.0), MercuryDesc(nang=8, nrad=24, v_kms=450.0, r_in=0.85)]
    probe = MhdContext(precision=precision)
    probe.add(descs[3])
    probe.finalize()
    big = probe.memory_bytes()
    probe.close()
    reqs = [Request(i, d, steps=4 + 2 * i, arrival=0.01 * i, kind="mercury" if isinstance(d, MercuryDesc) else
                    "lunar") for i, d in enumerate(descs)]
    eng = ServingEngine(precision=precision, hbm_budget_gb=2.2 * big / 1e9, chunk=3, workers=2, keep_results=True)
    out = eng.run(reqs)
    eng.close()
    assert out["requests"] == len(reqs) and out["failed"] == 0
    for r in out["records"]:
        hc = HostCase(r.desc, ghost=2)
        orc = Oracle(hc, threads=8)
        orc.step(r.steps)
        for f in FIELDS:
            for b in range(hc.nblocks):
                sl = interior(f, hc.n, hc.g)
                a, ref = r.result[(f, b)][sl], orc.view(f, b)[sl]
                e = np.abs(a - ref).max() / max(np.abs(ref).max(), 1e-300)
                assert e <= tol, (r.rid, f, b, e)
