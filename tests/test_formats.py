"""On-disk formats (CPU only): the reference's CASE grid directory and the
MPCNSRST restart file.

The CASE directory round-trips (generated shell -> files -> case): nodes,
topology, geometry, op lists and initial state are identical.  Independently
of the transform derivation, every interface read back from the files maps
each face node onto a node with the same coordinates in the neighbour block.
The restart file is parsed here from its published layout
(src/5_io/IOModule_Bin.cpp:168-330) and compared with the AoS fields.
"""
import os
import struct

import numpy as np
import pytest

from beta1_amd import CaseDirDesc, HostCase, OTDesc, WindDesc
from beta1_amd.capi import MhdError
from beta1_amd.fields import interior

G = 2
FIELDS = ["U_H", "B_xi", "B_eta", "B_zeta", "Badd_xi", "Badd_eta", "Badd_zeta", "Jac", "JDxi", "JDet", "JDze",
          "beta_xi", "alpha_zeta", "Bcell_from_Bface_w", "Jcell_from_Jedge_w"]


def shell_dir(tmp_path, desc):
    hc = HostCase(desc, ghost=G)
    d = str(tmp_path / "CASE")
    hc.write_dir(d)
    return hc, d


def test_case_dir_files_have_the_reference_layout(tmp_path):
    hc, d = shell_dir(tmp_path, WindDesc(nang=4, nrad=6))
    raw = open(os.path.join(d, "geometry", "grid0000.grd"), "rb").read()
    nb = struct.unpack_from("<i", raw, 0)[0]
    assert nb == 6
    off = 4
    for _ in range(nb):
        n = struct.unpack_from("<i", raw, off)[0]
        name = raw[off + 4: off + 4 + n]
        assert name.rstrip(b"\0") == b"Fluid"
        off += 4 + n
    dims = struct.unpack_from("<18i", raw, off)
    assert dims == (5, 5, 7) * 6
    off += 72
    # x of block 0, i fastest, equals the node coordinates
    x = np.frombuffer(raw, dtype="<f8", count=5 * 5 * 7, offset=off).reshape(7, 5, 5)
    nodes = hc.nodes(0)[0]
    gn = G + 1
    assert np.array_equal(x, nodes[gn:gn + 5, gn:gn + 5, gn:gn + 7].transpose(2, 1, 0))
    assert len(raw) == off + 6 * 3 * 5 * 5 * 7 * 8
    bc = open(os.path.join(d, "geometry", "boundary_condition", "boundary0000.txt")).read().split()
    assert bc[0] == "2"  # two physical faces per block: the body and the far field
    names = open(os.path.join(d, "geometry", "boundary_condition", "boundary_name.fvbnd")).read().split("\n")
    assert names[0] == "FVBND 1 3" and "Solid_Surface" in names and "Farfield" in names
    inner = open(os.path.join(d, "geometry", "boundary_condition", "inner0000.txt")).read().split("\n")
    assert int(inner[0]) == 24 and inner[1].split() == ["4"] * 6


@pytest.mark.parametrize("desc", [WindDesc(nang=6, nrad=8), WindDesc(nang=5, nrad=7, r_out=4.0)])
def test_case_dir_round_trip_is_identical(tmp_path, desc):
    hc, d = shell_dir(tmp_path, desc)
    back = HostCase(CaseDirDesc(d, desc), ghost=G)
    assert back.nblocks == hc.nblocks and back.n == hc.n
    for b in range(hc.nblocks):
        assert np.array_equal(back.faces(b), hc.faces(b)), b
        for a, c in zip(back.nodes(b), hc.nodes(b)):
            assert np.array_equal(a, c)
        for f in FIELDS:
            assert np.array_equal(back.field(f, b), hc.field(f, b)), (b, f)
    for sy in range(14):
        assert back.op_count(sy) == hc.op_count(sy)


def test_periodic_box_round_trip_keeps_self_links(tmp_path):
    hc = HostCase(OTDesc(n=6), ghost=G)
    d = str(tmp_path / "BOX")
    hc.write_dir(d)
    inner = open(os.path.join(d, "geometry", "boundary_condition", "inner0000.txt")).read().split()
    assert inner[0] == "6" and inner[1] == "6"
    back = HostCase(CaseDirDesc(d, WindDesc()), ghost=G)
    assert np.array_equal(back.faces(0), hc.faces(0))
    assert all(back.faces(0)[f][4] == 1 for f in range(6))  # periodic flag


def test_interfaces_read_from_files_map_coincident_nodes(tmp_path):
    """Geometric pin of the transform derivation: for every interface, node
    (i, j, k) of the face and node T(i, j, k) of the neighbour coincide."""
    desc = WindDesc(nang=5, nrad=4)
    _, d = shell_dir(tmp_path, desc)
    back = HostCase(CaseDirDesc(d, desc), ghost=G)
    gn = G + 1
    n = back.n
    checked = 0
    for b in range(back.nblocks):
        X = np.stack(back.nodes(b), -1)
        for f in range(6):
            row = back.faces(b)[f]
            if row[0] != 1:
                continue
            nb, perm, sign, off = row[2], row[5:8], row[8:11], row[11:14]
            Y = np.stack(back.nodes(nb), -1)
            ax = f // 2
            rng = [range(n[a] + 1) for a in range(3)]
            rng[ax] = [0 if f % 2 == 0 else n[ax]]
            for i in rng[0]:
                for j in rng[1]:
                    for k in rng[2]:
                        loc = (i, j, k)
                        t = [0, 0, 0]
                        for a in range(3):
                            t[perm[a]] = sign[a] * loc[a] + off[a]
                        p = X[i + gn, j + gn, k + gn]
                        q = Y[t[0] + gn, t[1] + gn, t[2] + gn]
                        assert np.allclose(p, q, rtol=0, atol=1e-12), (b, f, loc)
                        checked += 1
    assert checked == 6 * 4 * 6 * 5  # 4 interfaces per panel, 6 x 5 face nodes


def test_case_dir_rejects_missing_files(tmp_path):
    with pytest.raises(MhdError):
        HostCase(CaseDirDesc(str(tmp_path / "nothing"), WindDesc()), ghost=G)


def read_restart_py(path):
    """Restart parser written from the reference's file layout."""
    raw = open(path, "rb").read()
    assert raw[:8] == b"MPCNSRST"
    version, step = struct.unpack_from("<ii", raw, 8)
    time = struct.unpack_from("<d", raw, 16)[0]
    nb, nf = struct.unpack_from("<ii", raw, 24)
    off = 32
    out = {}
    for _ in range(nf):
        n = struct.unpack_from("<i", raw, off)[0]
        name = raw[off + 4: off + 4 + n].decode()
        off += 4 + n
        loc, nc, ng = struct.unpack_from("<iii", raw, off)
        off += 12
        blocks = []
        for _ in range(nb):
            lo = struct.unpack_from("<3i", raw, off)
            hi = struct.unpack_from("<3i", raw, off + 12)
            active = struct.unpack_from("<i", raw, off + 24)[0]
            off += 28
            shp = tuple(h - l for l, h in zip(lo, hi)) + (nc,)
            cnt = int(np.prod(shp))
            blocks.append((lo, hi, np.frombuffer(raw, "<f8", cnt, off).reshape(shp) if active else None))
            off += 8 * cnt if active else 0
        out[name] = (loc, nc, ng, blocks)
    assert off == len(raw)
    return version, step, time, out


def test_restart_layout_and_round_trip(tmp_path):
    hc = HostCase(WindDesc(nang=4, nrad=6), ghost=G)
    for sy in (3, 4, 5, 6, 7):  # U wall/far field, U halo, B wall, B halo: a synced state
        hc.apply_sync(sy)
    path = str(tmp_path / "restart.bin")
    hc.write_restart(path, step=1234, time=5.5)
    version, step, time, f = read_restart_py(path)
    assert (version, step, time) == (1, 1234, 5.5)
    assert list(f) == ["U_H", "B_xi", "B_eta", "B_zeta"]
    assert [f[k][0] for k in f] == [0, 1, 2, 3] and [f[k][1] for k in f] == [5, 1, 1, 1]
    for name, (loc, nc, ng, blocks) in f.items():
        assert ng == G
        for b, (lo, hi, v) in enumerate(blocks):
            assert lo == (-G, -G, -G)
            assert np.array_equal(v, hc.field(name, b))
    # perturb the state, read the file back: interiors and ghosts restored
    ref = {(n, b): hc.field(n, b) for n in f for b in range(6)}
    for b in range(6):
        U = hc.field("U_H", b)
        hc.set_field("U_H", b, U * 1.5)
    assert hc.read_restart(path) == (1234, 5.5)
    for (n, b), a in ref.items():
        assert np.array_equal(hc.field(n, b), a), (n, b)


def test_restart_reads_across_ghost_depths(tmp_path):
    desc = WindDesc(nang=4, nrad=6)
    a = HostCase(desc, ghost=2)
    rng = np.random.default_rng(5)
    for b in range(6):
        U = a.field("U_H", b)
        sl = interior("U_H", a.n, 2)
        U[sl] *= 1.0 + 0.01 * rng.random(U[sl].shape)
        a.set_field("U_H", b, U)
    for sy in (3, 4, 5, 6, 7):
        a.apply_sync(sy)
    path = str(tmp_path / "r.bin")
    a.write_restart(path)
    c = HostCase(desc, ghost=3)
    c.read_restart(path)
    for b in range(6):
        assert np.array_equal(c.field("U_H", b)[interior("U_H", c.n, 3)], a.field("U_H", b)[interior("U_H", a.n, 2)])


def test_two_fluid_restart_holds_na(tmp_path):
    from beta1_amd import MercuryDesc
    hc = HostCase(MercuryDesc(nang=4, nrad=6), ghost=G)
    for sy in (3, 4, 5, 6, 7):
        hc.apply_sync(sy)
    path = str(tmp_path / "m.bin")
    hc.write_restart(path, step=7, time=0.5)
    _, _, _, f = read_restart_py(path)
    assert list(f) == ["U_H", "U_Na", "B_xi", "B_eta", "B_zeta"]  # the Mercury whitelist
    for b, (lo, hi, v) in enumerate(f["U_Na"][3]):
        assert np.array_equal(v, hc.field("U_Na", b))


def restart_with_inactive_block(src, dst, field, block):
    """Rewrite a restart file with `block` of `field` marked inactive (no data)."""
    raw = bytearray(open(src, "rb").read())
    nb, nf = struct.unpack_from("<ii", raw, 24)
    off, out = 32, bytearray(raw[:32])
    for _ in range(nf):
        n = struct.unpack_from("<i", raw, off)[0]
        name = raw[off + 4: off + 4 + n].decode()
        head = raw[off: off + 4 + n + 12]
        nc = struct.unpack_from("<i", raw, off + 4 + n + 4)[0]
        off += 4 + n + 12
        out += head
        for b in range(nb):
            lo = struct.unpack_from("<3i", raw, off)
            hi = struct.unpack_from("<3i", raw, off + 12)
            active = struct.unpack_from("<i", raw, off + 24)[0]
            cnt = int(np.prod([h - l for l, h in zip(lo, hi)])) * nc if active else 0
            data = raw[off + 28: off + 28 + 8 * cnt]
            if name == field and b == block:
                out += raw[off: off + 24] + struct.pack("<i", 0)
            else:
                out += raw[off: off + 28] + data
            off += 28 + 8 * cnt
    open(dst, "wb").write(bytes(out))


def test_restart_rejects_bad_files_and_writes_nothing(tmp_path):
    """IOModule::ReadRestartBinFile fails on a file block without data for a
    block the case holds (src/5_io/IOModule_Bin.cpp:340-350) and on a cut
    file; here every field is checked before any is written, so a rejected
    file leaves the case as it was."""
    hc = HostCase(WindDesc(nang=4, nrad=6), ghost=G)
    for sy in (3, 4, 5, 6, 7):
        hc.apply_sync(sy)
    good = str(tmp_path / "good.bin")
    hc.write_restart(good, step=3, time=0.25)
    before = {(n, b): hc.field(n, b) for n in ("U_H", "B_xi", "B_eta", "B_zeta") for b in range(6)}
    # a file from another state, then spoiled
    other = HostCase(WindDesc(nang=4, nrad=6, v_kms=650.0), ghost=G)
    for sy in (3, 4, 5, 6, 7):
        other.apply_sync(sy)
    src = str(tmp_path / "other.bin")
    other.write_restart(src)
    inactive = str(tmp_path / "inactive.bin")
    restart_with_inactive_block(src, inactive, "B_zeta", 4)
    cut = str(tmp_path / "cut.bin")
    open(cut, "wb").write(open(src, "rb").read()[:-1000])
    for bad in (inactive, cut):
        with pytest.raises(MhdError):
            hc.read_restart(bad)
        for (n, b), a in before.items():
            assert np.array_equal(hc.field(n, b), a), (bad, n, b)
    assert hc.read_restart(good) == (3, 0.25)
    # the spoiled file's other fields are fine: reading the unspoiled one works
    hc.read_restart(src)
    assert not np.array_equal(hc.field("U_H", 0), before[("U_H", 0)])
