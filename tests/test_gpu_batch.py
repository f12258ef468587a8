"""Batching and serving on the GPU: every case of a mixed batch, and every
case admitted into a running context, evolves exactly as it does alone."""
import numpy as np
import pytest

from beta1_amd import MercuryDesc, MhdContext, OTDesc, WindDesc

pytestmark = pytest.mark.gpu

FIELDS = ("U_H", "B_xi", "B_eta", "B_zeta")


def alone(desc, steps, **kw):
    ctx = MhdContext(**kw)
    cid = ctx.add(desc)
    ctx.finalize()
    ctx.step(steps)
    nb = ctx.info(cid)[0]
    out = {(f, b): ctx.field(cid, f, b) for f in FIELDS for b in range(nb)}
    ctx.close()
    return out


def same(ctx, cid, ref):
    for (f, b), a in ref.items():
        assert np.array_equal(ctx.field(cid, f, b), a), (cid, f, b)


DESCS = [WindDesc(nang=8, nrad=16, v_kms=350.0), WindDesc(nang=12, nrad=16, n_cc=12.0),
         WindDesc(nang=8, nrad=24, B_nT=(1.0, -4.0, 3.0)), OTDesc(n=12)]


def test_mixed_batch_each_case_as_alone():
    ctx = MhdContext(autotune=True)
    ids = [ctx.add(d) for d in DESCS]
    ctx.finalize()
    ctx.step(5, graph=True)
    for d, cid in zip(DESCS, ids):
        same(ctx, cid, alone(d, 5))


def test_admission_and_retirement_mid_run():
    a, b, c = DESCS[0], DESCS[1], DESCS[3]
    ctx = MhdContext()
    ia = ctx.add(a)
    ctx.finalize()
    ctx.step(3)                       # odd: the running case sits in the second U buffer
    ib = ctx.try_add(b)
    ctx.step(4, graph=True)
    ic = ctx.try_add(c)
    ctx.step(2)
    assert ctx.case_steps(ia) == 9 and ctx.case_steps(ib) == 6 and ctx.case_steps(ic) == 2
    same(ctx, ia, alone(a, 9))
    same(ctx, ib, alone(b, 6))
    ctx.retire(ia)
    ctx.step(2)
    same(ctx, ib, alone(b, 8))
    same(ctx, ic, alone(c, 4))


def test_admission_respects_hbm_budget():
    probe = MhdContext()
    probe.add(DESCS[0])
    probe.finalize()
    one = probe.memory_bytes()
    probe.close()
    # one case is state (37 slots) + geometry (38 slots) of its shape: a
    # second case of the same shape adds its state only (~0.49 of `one`), a
    # case of the 1.5x larger shape adds ~1.4 of `one`
    ctx = MhdContext(hbm_budget_gb=1.6 * one / 1e9)
    ctx.finalize()
    first = ctx.try_add(DESCS[0])
    assert first is not None
    assert ctx.try_add(DESCS[2]) is None      # a second grid shape does not fit
    second = ctx.try_add(DESCS[0])            # same shape shares geometry: fits
    assert second is not None
    ctx.retire(first)
    assert ctx.try_add(DESCS[2]) is None
    ctx.retire(second)
    assert ctx.try_add(DESCS[2]) is not None


def test_serving_results_equal_alone():
    """Continuous batching with a budget that holds about two cases: queued
    requests wait, are admitted as others retire, and each result equals the
    case run alone for the same number of steps."""
    from beta1_amd.serve import Request, ServingEngine

    descs = [WindDesc(nang=8, nrad=16, v_kms=420.0), MercuryDesc(nang=8, nrad=16),
             WindDesc(nang=12, nrad=16, n_cc=9.0), MercuryDesc(nang=8, nrad=24, n_cc=30.0),
             WindDesc(nang=8, nrad=16, B_nT=(2.0, -3.0, 1.0))]
    probe = MhdContext()
    probe.add(descs[2])
    probe.finalize()
    big = probe.memory_bytes()
    probe.close()
    reqs = [Request(i, d, steps=5 + 2 * i, arrival=0.02 * i) for i, d in enumerate(descs)]
    eng = ServingEngine(precision="fp64", hbm_budget_gb=1.8 * big / 1e9, chunk=3, workers=2, keep_results=True)
    out = eng.run(reqs)
    eng.close()
    assert out["requests"] == len(reqs)
    for r in out["records"]:
        assert r.admitted >= r.arrival and r.done >= r.admitted
        ref = alone(r.desc, r.steps)
        for key, a in ref.items():
            assert np.array_equal(r.result[key], a), (r.rid, key)


def test_xcd_order_and_tuning_do_not_change_results():
    d = DESCS[1]
    ref = alone(d, 4, xcd_order=False)
    ctx = MhdContext(autotune=True, xcd_order=True)
    cid = ctx.add(d)
    ctx.finalize()
    ctx.step(4)
    same(ctx, cid, ref)
    assert {t[0] for t in ctx.tuning().values()} <= {1, 2, 4}


SHARED = [WindDesc(nang=8, nrad=16, v_kms=350.0), WindDesc(nang=8, nrad=16, n_cc=12.0),
          WindDesc(nang=8, nrad=16, B_nT=(1.0, -4.0, 3.0)), OTDesc(n=12), WindDesc(nang=12, nrad=16)]
KERNELS = ("k_face_emf", "k_edge_emf", "k_fluid", "k_curl_update", "k_derived", "k_jcell")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_geometry_shared_tiles_leave_each_case_as_alone(precision):
    """Cases of one geometry set run each geometry tile back to back (three
    shells share one set here); every case still evolves as it does alone,
    also after a member of the set retires."""
    ctx = MhdContext(precision=precision)
    ids = [ctx.add(d) for d in SHARED]
    ctx.finalize()
    for n, k in enumerate(KERNELS):
        ctx.set_tuning(k, rows=(1, 2, 4)[n % 3], i_inner=n % 2 == 1, geo_shared=True)
    assert all(ctx.tuning()[k][2] == "geo-shared" for k in KERNELS)
    ctx.step(3, graph=True)
    ctx.step(2)
    for d, cid in zip(SHARED, ids):
        same(ctx, cid, alone(d, 5, precision=precision))
    ctx.retire(ids[1])
    ctx.step(3, graph=True)
    for d, cid in zip(SHARED, ids):
        if cid != ids[1]:
            same(ctx, cid, alone(d, 8, precision=precision))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("rows", [1, 4])
def test_member_looping_derived_leaves_each_case_as_alone(precision, rows):
    """k_derived with member-looping tiles: one workgroup loads a geometry
    tile once and updates the three shells sharing it in turn (each case's dt
    reduced on its own); every case evolves bit for bit as alone, also after a
    member retires.  Only k_derived has the variant, j-inner."""
    from beta1_amd.capi import MhdError

    ctx = MhdContext(precision=precision)
    ids = [ctx.add(d) for d in SHARED]
    ctx.finalize()
    ctx.set_tuning("k_derived", rows=rows, geo_shared=2)
    assert ctx.tuning()["k_derived"] == (rows, "j-inner", "geo-loop")
    for bad in (dict(kernel="k_jcell", rows=rows, geo_shared=2), dict(kernel="k_derived", rows=rows, i_inner=True,
                                                                       geo_shared=2)):
        with pytest.raises(MhdError):
            ctx.set_tuning(**bad)
    ctx.step(3, graph=True)
    ctx.step(2)
    for d, cid in zip(SHARED, ids):
        same(ctx, cid, alone(d, 5, precision=precision))
        assert ctx.last_dt(cid) == alone_dt(d, 5, precision)
    ctx.retire(ids[0])
    ctx.step(3, graph=True)
    for d, cid in zip(SHARED, ids):
        if cid != ids[0]:
            same(ctx, cid, alone(d, 8, precision=precision))


def alone_dt(desc, steps, precision):
    ctx = MhdContext(precision=precision)
    cid = ctx.add(desc)
    ctx.finalize()
    ctx.step(steps)
    dt = ctx.last_dt(cid)
    ctx.close()
    return dt


def test_autotune_after_admission_keeps_results():
    ctx = MhdContext(autotune=True)
    ids = [ctx.add(d) for d in SHARED[:3]]
    ctx.finalize()
    ctx.step(2)
    ctx.autotune()
    ctx.step(2, graph=True)
    for d, cid in zip(SHARED[:3], ids):
        same(ctx, cid, alone(d, 4))
