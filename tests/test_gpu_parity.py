"""GPU parity: the HIP step (through the C-ABI) against the CPU FP64 oracle.

Bar: FP64 path bit-identical to the oracle (the test asserts relative
L-inf <= 1e-12 per conserved variable and reports the bitwise status);
mixed precision within 1e-5 relative L-inf of the FP64 path; div B at
round-off.
"""
import numpy as np
import pytest

from beta1_amd import HostCase, MercuryDesc, MhdContext, OTDesc, WindDesc
from beta1_amd.fields import interior
from oracle_bind import Oracle

pytestmark = pytest.mark.gpu

CONSERVED = ["U_H", "B_xi", "B_eta", "B_zeta"]


def rel_linf(a, b):
    den = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() / den


def compare(ctx, cid, orc, nblocks, n, g, tol=1e-12):
    worst = {}
    for name in CONSERVED:
        for b in range(nblocks):
            sl = interior(name, n, g)
            gpu = ctx.field(cid, name, b)[sl]
            ref = orc.view(name, b)[sl]
            assert np.all(np.isfinite(gpu)), f"{name} block {b} not finite"
            if name == "U_H":
                for m in range(5):
                    e = rel_linf(gpu[..., m], ref[..., m])
                    worst[f"{name}[{m}]"] = max(worst.get(f"{name}[{m}]", 0.0), e)
            else:
                e = rel_linf(gpu, ref)
                worst[name] = max(worst.get(name, 0.0), e)
    for k, e in worst.items():
        assert e <= tol, f"{k}: rel Linf {e:.3e} > {tol}"
    return worst


def run_pair(desc, steps, recon=0, graph=False, precision="fp64"):
    g = 2
    hc = HostCase(desc, ghost=g)
    orc = Oracle(hc, recon=recon)
    ctx = MhdContext(precision=precision, recon=recon, ghost=g)
    cid = ctx.add(desc)
    ctx.finalize()
    dts_o = orc.step(steps)
    ctx.step(steps, graph=graph)
    return hc, orc, ctx, cid, dts_o


@pytest.mark.parametrize("n,steps", [(16, 10), (32, 6)])
def test_orszag_tang_fp64_matches_oracle(n, steps):
    desc = OTDesc(n=n)
    hc, orc, ctx, cid, dts = run_pair(desc, steps)
    assert ctx.last_dt(cid) == dts[-1]
    compare(ctx, cid, orc, 1, hc.n, hc.g)
    assert ctx.divb_max(cid) < 1e-11


def test_orszag_tang_minmod_matches_oracle():
    desc = OTDesc(n=16)
    hc, orc, ctx, cid, dts = run_pair(desc, 8, recon=1)
    compare(ctx, cid, orc, 1, hc.n, hc.g)


@pytest.mark.parametrize("nang,nrad,steps", [(8, 16, 12), (12, 24, 6)])
def test_lunar_shell_fp64_matches_oracle(nang, nrad, steps):
    desc = WindDesc(nang=nang, nrad=nrad)
    hc, orc, ctx, cid, dts = run_pair(desc, steps)
    assert ctx.last_dt(cid) == dts[-1]
    compare(ctx, cid, orc, 6, hc.n, hc.g)
    assert ctx.divb_max(cid) < 1e-9


def test_mercury_shell_fp64_matches_oracle():
    desc = MercuryDesc(nang=8, nrad=16)
    hc, orc, ctx, cid, dts = run_pair(desc, 8)
    assert ctx.last_dt(cid) == dts[-1]
    compare(ctx, cid, orc, 6, hc.n, hc.g)
    assert ctx.divb_max(cid) < 1e-9


def test_graph_step_equals_eager():
    desc = WindDesc(nang=8, nrad=16)
    a = MhdContext()
    ca = a.add(desc)
    a.finalize()
    b = MhdContext()
    cb = b.add(desc)
    b.finalize()
    a.step(6 + 25 + 21)
    b.step(2, graph=True)
    b.step(4, graph=True)  # graph replays with eager syncs in between
    b.step(25, graph=True)  # two ten-step graphs, two two-step graphs, one eager step
    b.step(21, graph=True)  # from the odd buffer: one eager step first
    for name in CONSERVED:
        for blk in range(6):
            assert np.array_equal(a.field(ca, name, blk), b.field(cb, name, blk))


def test_second_stream_overlap_leaves_results_unchanged():
    """The fluid update beside the EMF chain and J_cell beside the derived
    fields (two streams, eager and captured) equal the one-stream order."""
    descs = [WindDesc(nang=8, nrad=16), MercuryDesc(nang=8, nrad=16)]
    runs = []
    for overlap in (False, True):
        ctx = MhdContext(overlap=overlap)
        ids = [ctx.add(d) for d in descs]
        ctx.finalize()
        ctx.step(3)
        ctx.step(4, graph=True)
        ctx.step(1)
        runs.append({(c, f, b): ctx.field(c, f, b) for c in ids for f in CONSERVED + ["J_cell", "B_cell"]
                     for b in range(6)})
        runs[-1]["dt"] = [ctx.last_dt(c) for c in ids]
        ctx.close()
    for k in runs[0]:
        assert np.array_equal(runs[0][k], runs[1][k]), k


def test_mixed_precision_close_to_fp64():
    desc = WindDesc(nang=8, nrad=16)
    f = MhdContext("fp64")
    cf = f.add(desc)
    f.finalize()
    m = MhdContext("mixed")
    cm = m.add(desc)
    m.finalize()
    f.step(10)
    m.step(10)
    hc = HostCase(desc)
    for name in CONSERVED:
        for blk in range(6):
            sl = interior(name, hc.n, hc.g)
            a, b = m.field(cm, name, blk)[sl], f.field(cf, name, blk)[sl]
            assert rel_linf(a, b) <= 1e-5, name
    assert m.divb_max(cm) < 1e-9


def test_compute_dt_matches_oracle_and_leaves_step_untouched():
    desc = WindDesc(nang=8, nrad=16)
    hc = HostCase(desc)
    orc = Oracle(hc)
    ctx = MhdContext()
    cid = ctx.add(desc)
    ctx.finalize()
    assert ctx.compute_dt(cid) == orc.compute_dt()
    ctx.step(3)
    orc.step(3)
    assert ctx.compute_dt(cid) == orc.compute_dt()
    ctx.step(1)
    assert ctx.last_dt(cid) == orc.step(1)[0]


def test_device_bcell_weights_equal_host_table():
    """The device keeps Mi and mu of the B_cell least squares instead of the
    18-weight table; the table it reports (formed from them and the face
    area vectors, as the B_cell kernel forms each weight) equals the host
    builder's table bit for bit."""
    desc = WindDesc(nang=8, nrad=16)
    ctx = MhdContext("fp64", recon=0, ghost=2, device=0)
    cid = ctx.add(desc)
    ctx.finalize()
    hc = HostCase(desc, ghost=2)
    for b in range(6):
        dev = ctx.field(cid, "Bcell_from_Bface_w", b)
        host = hc.field("Bcell_from_Bface_w", b)
        assert np.abs(host).max() > 0.0
        assert np.array_equal(dev, host), b
