"""The drop-in boundary on the GPU.

* A case assembled on the device from a driver's own arrays (mhd_case_begin,
  mhd_add_block, mhd_set_topology, mhd_case_end) steps bit for bit like the
  generated case (mhd_add_case_lunar), before and after mhd_finalize.
* The per-stage entry points in StepOnce's order give mhd_step's result bit
  for bit, and at every stage boundary the device arrays equal the oracle's
  after the same stage (FP64: relative L-inf <= 1e-12; they are bitwise).
"""
import numpy as np
import pytest

from beta1_amd import HostCase, MercuryDesc, MhdContext, OTDesc, WindDesc, case_parts
from beta1_amd.capi import MhdError
from beta1_amd.fields import interior
from oracle_bind import Oracle

pytestmark = pytest.mark.gpu

FIELDS = ("U_H", "B_xi", "B_eta", "B_zeta")
TOL = 1e-12


def run(ctx, cid, steps):
    ctx.step(steps)
    nb = ctx.info(cid)[0]
    return {(f, b): ctx.field(cid, f, b) for f in FIELDS for b in range(nb)}


@pytest.mark.parametrize("live", [False, True])
def test_device_case_from_driver_arrays_steps_like_generated(live):
    desc = WindDesc(nang=8, nrad=16, v_kms=450.0)
    ref = MhdContext()
    rid = ref.add(desc)
    ref.finalize()
    want = run(ref, rid, 6)
    k, blocks = case_parts(HostCase(desc))
    ctx = MhdContext()
    if live:
        ctx.finalize()
    cid = ctx.add_parts(k, blocks)
    if not live:
        ctx.finalize()
    got = run(ctx, cid, 6)
    for key, a in want.items():
        assert np.array_equal(got[key], a), key
    assert ctx.last_dt(cid) == ref.last_dt(rid)


def test_device_drop_in_refuses_bad_arrays_and_unfinished_cases():
    k, blocks = case_parts(HostCase(WindDesc(nang=8, nrad=16)))
    blocks[1]["geometry"]["alpha_eta"] = blocks[1]["geometry"]["alpha_eta"] * 1.01
    ctx = MhdContext()
    with pytest.raises(MhdError, match="alpha_eta"):
        ctx.add_parts(k, blocks)
    ctx.finalize()
    ctx.step(1)  # nothing resident: the failed draft is not a case


STAGES = ["dt", "fluid", "face_emf", "sync:Eface", "edge_emf", "sync:Eedge", "curl", "update", "sync:Ucell",
          "sync:Bface", "derived"]


@pytest.mark.parametrize("desc", [WindDesc(nang=8, nrad=16), MercuryDesc(nang=8, nrad=16), OTDesc(n=12)])
def test_staged_step_equals_mhd_step(desc):
    a, b = MhdContext(), MhdContext()
    ia, ib = a.add(desc), b.add(desc)
    a.finalize()
    b.finalize()
    a.step(4)
    for _ in range(3):
        for s in STAGES:
            b.stage(s)
    b.step(1)  # and back to whole steps
    assert a.case_steps(ia) == b.case_steps(ib) == 4
    nb = a.info(ia)[0]
    for f in FIELDS + ("B_cell", "J_cell"):
        for blk in range(nb):
            assert np.array_equal(a.field(ia, f, blk), b.field(ib, f, blk)), (f, blk)
    assert a.last_dt(ia) == b.last_dt(ib)


def test_open_step_blocks_whole_steps():
    ctx = MhdContext()
    ctx.add(WindDesc(nang=8, nrad=16))
    ctx.finalize()
    ctx.stage("dt")
    ctx.stage("fluid")
    for call in (lambda: ctx.step(1), lambda: ctx.refresh(), lambda: ctx.stage("fluid"), lambda: ctx.stage("derived")):
        with pytest.raises(MhdError):
            call()
    for s in STAGES[2:]:
        ctx.stage(s)
    ctx.step(2)
    with pytest.raises(MhdError):
        ctx.stage("update")  # no open step


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def compare(ctx, cid, orc, names, n, g, sl_of=None, comps=None):
    worst = 0.0
    for name in names:
        for blk in range(orc.nblocks):
            sl = sl_of(name) if sl_of else interior(name, n, g)
            a, r = ctx.field(cid, name, blk)[sl], orc.view(name, blk)[sl]
            if comps is not None:
                a, r = a[..., comps[name]], r[..., comps[name]]
            assert np.all(np.isfinite(a)), (name, blk)
            worst = max(worst, rel(a, r))
    return worst


def tangential(name):
    ax = {"Eface_xi": 0, "Eface_eta": 1, "Eface_zeta": 2}[name]
    return [c for c in range(3) if c != ax]


@pytest.mark.parametrize("desc", [WindDesc(nang=8, nrad=16), MercuryDesc(nang=8, nrad=16)])
def test_every_stage_matches_the_oracle(desc):
    """After each stage the device's arrays equal the oracle's after the same
    stage: face EMF candidates (the two tangential components the device
    keeps), edge EMFs before and after the far-field handler, the updated
    state, its synced ghosts (the single-hop ghosts both fill), B_cell and
    J_cell, and each step's dt."""
    hc = HostCase(desc, ghost=2)
    orc = Oracle(hc)
    ctx = MhdContext()
    cid = ctx.add(desc)
    ctx.finalize()
    n, g = hc.n, hc.g
    ef = ["Eface_xi", "Eface_eta", "Eface_zeta"]
    ee = ["E_xi", "E_eta", "E_zeta"]
    comps = {nm: tangential(nm) for nm in ef}
    two = isinstance(desc, MercuryDesc) and desc.two_fluid
    state = ["U_H", "B_xi", "B_eta", "B_zeta"] + (["U_Na"] if two else [])
    for step in range(3):
        dt_o = orc.compute_dt()
        for s in STAGES:
            ctx.stage(s)
            orc.stage(s)
            if s == "dt":
                assert ctx.last_dt(cid) == dt_o, step
            elif s == "face_emf":
                assert compare(ctx, cid, orc, ef, n, g, comps=comps) <= TOL, (step, s)
            elif s in ("edge_emf", "sync:Eedge"):
                assert compare(ctx, cid, orc, ee, n, g) <= TOL, (step, s)
            elif s == "update":
                assert compare(ctx, cid, orc, state, n, g) <= TOL, (step, s)
            elif s == "derived":
                assert compare(ctx, cid, orc, ["B_cell", "J_cell"], n, g) <= TOL, (step, s)
            elif s == "sync:Bface":
                # the synced ghosts: one layer outside each block face
                for name in state:
                    assert face_ghosts_rel(ctx, cid, orc, name, n, g) <= TOL, (step, name)


def face_ghosts_rel(ctx, cid, orc, name, n, g):
    """Relative L-inf over the DOFs one layer outside exactly one face of
    each block (interface and physical ghosts alike)."""
    from beta1_amd.fields import LOCATION, delta
    worst = 0.0
    for blk in range(orc.nblocks):
        a, r = ctx.field(cid, name, blk), orc.view(name, blk)
        cnt = [n[q] + 1 - delta(LOCATION[name], q) for q in range(3)]
        idx = np.meshgrid(*[np.arange(cnt[q] + 2 * g) - g for q in range(3)], indexing="ij")
        out = sum(((idx[q] == -1) | (idx[q] == cnt[q])).astype(int) for q in range(3))
        inside = np.all([(idx[q] >= -1) & (idx[q] <= cnt[q]) for q in range(3)], axis=0)
        m = (out == 1) & inside
        worst = max(worst, rel(a[m], r[m]))
    return worst


def test_blocks_of_unequal_extent_match_the_oracle():
    """Two blocks of 4 and 6 cells along x (one slot size for the case) built
    from a driver's arrays, a perturbed state, 6 steps: bitwise the oracle."""
    from test_dropin import two_box_parts

    k, blocks = two_box_parts()
    rng = np.random.default_rng(9)
    made = []
    for _ in range(2):
        hc = HostCase.from_parts(k, blocks, ghost=2)
        r = np.random.default_rng(9)
        for b in range(hc.nblocks):
            U = hc.field("U_H", b)
            n = hc.ns[b]
            U[2:2 + n[0], 2:2 + n[1], 2:2 + n[2]] *= 1.0 + 1e-3 * r.uniform(-1, 1, (n[0], n[1], n[2], 1))
            hc.set_field("U_H", b, U)
        for sy in (3, 4, 5):
            hc.apply_sync(sy)
        made.append(hc)
    del rng
    orc = Oracle(made[0])
    ctx = MhdContext()
    cid = ctx.adopt(made[1])
    ctx.finalize()
    dts = orc.step(6)
    ctx.step(6)
    assert ctx.last_dt(cid) == dts[-1]
    for f in FIELDS:
        for b in range(2):
            n = orc.ns[b]
            sl = interior(f, n, 2)
            a, r = ctx.field(cid, f, b)[sl], orc.view(f, b)[sl]
            assert np.array_equal(a, r), (f, b)


def kwall_box_parts(nk, hi_bc):
    """One block of 4 x 4 x nk cells over [0,1]^2 x [0,2], periodic in x and
    y, an absorbing wall at z = 0 and `hi_bc` (1 far field, 2 wall) at z = 2;
    the Lunar constants of the default wind."""
    from test_dropin import G
    k = HostCase(WindDesc(nang=4, nrad=6), ghost=G).constants()
    ni = nj = 4
    xs, ys, zs = np.linspace(0, 1, ni + 1), np.linspace(0, 1, nj + 1), np.linspace(0, 2, nk + 1) ** 1.2
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    f = np.zeros((6, 14), dtype=np.int64)
    ident = [0, 1, 2, 1, 1, 1]
    f[0] = [1, 0, 0, 1, 1] + ident + [ni, 0, 0]
    f[1] = [1, 0, 0, 0, 1] + ident + [-ni, 0, 0]
    f[2] = [1, 0, 0, 3, 1] + ident + [0, nj, 0]
    f[3] = [1, 0, 0, 2, 1] + ident + [0, -nj, 0]
    f[4] = [0, 2, -1, -1, 0] + [0] * 9
    f[5] = [0, hi_bc, -1, -1, 0] + [0] * 9
    return k, [{"n": (ni, nj, nk), "nodes": [X, Y, Z], "faces": f}]


@pytest.mark.parametrize("nk,hi_bc", [(8, 2), (8, 1), (65, 2)])
def test_walls_on_k_faces_match_the_oracle(nk, hi_bc):
    """The step's kernels apply the wall and far-field work of the k faces
    themselves (the fluid kernel from the neighbour lane's new state); a box
    with walls or far field on both k faces steps bit for bit like the
    oracle, and like the per-stage entries (which leave that work to the
    syncs).  nk = 65 puts the high wall's last cell at the start of a
    64-lane tile, where the fluid kernel leaves the wall to the op list."""
    k, blocks = kwall_box_parts(nk, hi_bc)
    made = []
    for _ in range(3):
        hc = HostCase.from_parts(k, blocks, ghost=2)
        r = np.random.default_rng(3)
        U = hc.field("U_H", 0)
        U[2:6, 2:6, 2:2 + nk] *= 1.0 + 2e-3 * r.uniform(-1, 1, (4, 4, nk, 1))
        hc.set_field("U_H", 0, U)
        for sy in (3, 4, 5):
            hc.apply_sync(sy)
        made.append(hc)
    orc = Oracle(made[0])
    ctx, staged = MhdContext(), MhdContext()
    cid, sid = ctx.adopt(made[1]), staged.adopt(made[2])
    ctx.finalize()
    staged.finalize()
    dts = orc.step(5)
    ctx.step(3)
    ctx.step(2, graph=True)
    for _ in range(5):
        for s in STAGES:
            staged.stage(s)
    assert ctx.last_dt(cid) == dts[-1] == staged.last_dt(sid)
    for f in FIELDS:
        sl = interior(f, (4, 4, nk), 2)
        a, r, b = ctx.field(cid, f, 0), orc.view(f, 0), staged.field(sid, f, 0)
        assert np.array_equal(a[sl], r[sl]), f
        assert np.array_equal(a, b), f  # ghosts too


@pytest.mark.parametrize("n,sides", [((6, 5, 7), ((2, 1), (1, 1), "p")), ((5, 6, 66), ("p", (2, 2), (1, 1)))])
def test_walls_on_i_and_j_faces_match_the_oracle(n, sides):
    """Walls and far field on the i and j faces (the shells have them on k
    only), meeting far-field faces along edges: the whole step, whose
    one-pass EMF kernel takes the wall's zero face candidates inline and the
    far-field ghosts from the Eface array, steps bit for bit like the oracle
    and like the per-stage entries."""
    from test_dropin import walled_box_parts
    k, blocks = walled_box_parts(n, sides)
    made = []
    for _ in range(3):
        hc = HostCase.from_parts(k, blocks, ghost=2)
        r = np.random.default_rng(5)
        U = hc.field("U_H", 0)
        g = 2
        U[g:g + n[0], g:g + n[1], g:g + n[2]] *= 1.0 + 2e-3 * r.uniform(-1, 1, (*n, 1))
        hc.set_field("U_H", 0, U)
        for sy in (3, 4, 5):
            hc.apply_sync(sy)
        made.append(hc)
    orc = Oracle(made[0])
    ctx, staged = MhdContext(), MhdContext()
    cid, sid = ctx.adopt(made[1]), staged.adopt(made[2])
    ctx.finalize()
    staged.finalize()
    dts = orc.step(5)
    ctx.step(3)
    ctx.step(2, graph=True)
    for _ in range(5):
        for s in STAGES:
            staged.stage(s)
    assert ctx.last_dt(cid) == dts[-1] == staged.last_dt(sid)
    for f in FIELDS:
        sl = interior(f, n, 2)
        a, r, b = ctx.field(cid, f, 0), orc.view(f, 0), staged.field(sid, f, 0)
        assert np.array_equal(a[sl], r[sl]), f
        assert np.array_equal(a, b), f  # ghosts too
