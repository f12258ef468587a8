"""The allocation-free diagnostics entry point (mhd_diagnostics_noalloc)
against mhd_diagnostics: the same 17 values bit for bit, on two calls in a
row (the second over fewer and smaller cases than the first, so stale keys of
the first call would show), and a clear error beyond its static scratch."""
import ctypes as C

import numpy as np
import pytest

from beta1_amd import MercuryDesc, MhdContext, OTDesc, WindDesc
from beta1_amd.capi import lib

pytestmark = pytest.mark.gpu


def test_noalloc_diagnostics_equal_allocating_entry_point():
    big = MhdContext()
    ids = [big.add(d) for d in (WindDesc(nang=16, nrad=32), MercuryDesc(nang=12, nrad=24), OTDesc(n=16))]
    big.finalize()
    big.step(3)
    first = big.diagnostics_noalloc(ids)
    for q, cid in enumerate(ids):
        assert np.array_equal(first[q], big.diagnostics_raw(cid)), cid
    # second call: one small case of another context on the same device
    small = MhdContext()
    sid = small.add(WindDesc(nang=8, nrad=16, v_kms=600.0))
    small.finalize()
    small.step(2)
    second = small.diagnostics_noalloc([sid])
    assert second.shape == (1, 17)
    assert np.array_equal(second[0], small.diagnostics_raw(sid))
    assert not np.array_equal(second[0], first[0])
    # and the first context again, one case, order reversed in a pair
    again = big.diagnostics_noalloc([ids[2], ids[0]])
    assert np.array_equal(again[0], big.diagnostics_raw(ids[2]))
    assert np.array_equal(again[1], big.diagnostics_raw(ids[0]))


def test_noalloc_diagnostics_rejects_what_its_scratch_cannot_hold():
    ctx = MhdContext()
    cid = ctx.add(OTDesc(n=8))
    ctx.finalize()
    out = np.zeros(17 * 65)
    for ids in ([cid] * 65, [], [cid, cid], [cid + 7]):
        arr = (C.c_int * max(len(ids), 1))(*ids)
        rc = lib().mhd_diagnostics_noalloc(ctx.ctx, arr, len(ids), out.ctypes.data_as(C.POINTER(C.c_double)))
        assert rc == -1, ids
        assert lib().mhd_last_error(ctx.ctx).decode().startswith("mhd_diagnostics_noalloc")
    assert np.array_equal(ctx.diagnostics_noalloc([cid])[0], ctx.diagnostics_raw(cid))
