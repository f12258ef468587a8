"""Faces shared by several interface patches (CPU only).

The reference's inner-boundary file may split a block face among several
neighbours, each patch over part of the face (src/1_grid/1_MPCNS_Grid.cpp:
392-505).  The check here is a decomposition invariance that the mathematics
fixes: the same periodic, stretched box (a different length and a
non-uniform spacing per axis) as one block and as three blocks -- A on the
low i side, B and C splitting the high i side along j, so A's two i faces are
each shared by B and C, one through plain and one through periodic patches --
must step to the same values bit for bit, every interior cell and face, with
the same dt.  A wrong part chosen for a ghost, a wrong offset or a wrong
split line moves a value and fails it.  The case directory round trips with
its patches, the same case handed over as a driver's arrays (split faces
through mhd_host_set_face_parts) is that case bit for bit, and faces that
mix patch kinds, leave a gap or link back elsewhere are refused.
"""
import os
import shutil
import struct

import numpy as np
import pytest

from beta1_amd import CaseDirDesc, HostCase, OTDesc, WindDesc, case_parts
from beta1_amd.capi import MhdError
from oracle_bind import Oracle

G = 2
N = 8
CI, CJ = 3, 5  # A: i < CI; B: i >= CI, j < CJ; C: i >= CI, j >= CJ
STRETCH = ((1.0, 0.3), (0.7, -0.25), (1.3, 0.2))  # (length, amplitude) per axis
FACES = ("B_xi", "B_eta", "B_zeta")


def read_grd(path):
    raw = open(path, "rb").read()
    nb = struct.unpack_from("<i", raw, 0)[0]
    off = 4
    for _ in range(nb):
        off += 4 + struct.unpack_from("<i", raw, off)[0]
    dims = [struct.unpack_from("<3i", raw, off + 12 * b) for b in range(nb)]
    off += 12 * nb
    blocks = []
    for d in dims:
        cnt = d[0] * d[1] * d[2]
        xyz = []
        for _ in range(3):
            a = np.frombuffer(raw[off:off + 8 * cnt], dtype="<f8").reshape(d[2], d[1], d[0])  # k, j, i
            xyz.append(a.transpose(2, 1, 0).copy())  # [i][j][k]
            off += 8 * cnt
        blocks.append(xyz)
    return blocks


def write_grd(path, blocks):
    out = bytearray(struct.pack("<i", len(blocks)))
    for b in range(len(blocks)):
        name = f"blk{b}".encode()
        out += struct.pack("<i", len(name)) + name
    for xyz in blocks:
        out += struct.pack("<3i", *xyz[0].shape)
    for xyz in blocks:
        for a in xyz:
            out += np.ascontiguousarray(a.transpose(2, 1, 0)).astype("<f8").tobytes()
    open(path, "wb").write(bytes(out))


def stretched(t, amp):
    return t + amp * np.sin(2 * np.pi * t) / (2 * np.pi)


def one_block_dir(tmp_path, curved=False):
    """The periodic N^3 box (self-linked faces) with stretched coordinates;
    `curved` adds to each coordinate a term periodic in the other two, so
    the grid lines bend and cross at varying angles (a non-orthogonal,
    curvilinear grid whose periodic faces stay translates of each other)."""
    d = str(tmp_path / ("CURVED" if curved else "ONE"))
    HostCase(OTDesc(n=N), ghost=G).write_dir(d)
    path = os.path.join(d, "geometry", "grid0000.grd")
    (x, y, z), = read_grd(path)
    t = [x, y, z]
    xyz = [STRETCH[a][0] * stretched(t[a], STRETCH[a][1]) for a in range(3)]
    if curved:
        w = 2 * np.pi
        xyz[0] = xyz[0] + 0.03 * np.sin(w * t[1]) * np.cos(w * t[2])
        xyz[1] = xyz[1] + 0.025 * np.sin(w * t[2] + 0.3) * np.cos(w * t[0])
        xyz[2] = xyz[2] + 0.03 * np.sin(w * t[0]) * np.sin(w * t[1])
    write_grd(path, [xyz])
    return d


def code(v, neg):
    return -(v + 1) if neg else v + 1


def patch_line(b, lo, hi, nb, tlo, thi, periodic):
    """One inner patch in the reference's signed 1-based coding, for an
    identity orientation: the first in-plane axis is marked on both sides."""
    ax = [a for a in range(3) if lo[a] == hi[a]][0]
    mark = 1 if ax == 0 else 0
    src = -(b + 1) if periodic else b + 1
    tar = -(nb + 1) if periodic else nb + 1
    s = " ".join(f"{code(lo[a], a == mark)} {code(hi[a], a == mark)}" for a in range(3))
    t = " ".join(f"{code(tlo[a], a == mark)} {code(thi[a], a == mark)}" for a in range(3))
    return f"{s} {src} blk{b}   {t} {tar} blk{nb}"


# Block b holds the nodes global = ORG[b] + SGN[b] * local: A (cells i < CI),
# B (i >= CI, j < CJ) and C (i >= CI, j >= CJ), C either aligned with the box
# or turned half a turn about k (its i and j axes against the global ones,
# still right-handed).
DIMS = ((CI, N, N), (N - CI, CJ, N), (N - CI, N - CJ, N))
LAYOUT = {"aligned": (((0, 0, 0), (CI, 0, 0), (CI, CJ, 0)), ((1, 1, 1),) * 3),
          "turned": (((0, 0, 0), (CI, 0, 0), (N, N, 0)), ((1, 1, 1), (1, 1, 1), (-1, -1, 1)))}
ORG, SGN = LAYOUT["aligned"]


def use(layout):
    global ORG, SGN
    ORG, SGN = LAYOUT[layout]


def to_global(b, p):
    return tuple(o + s * q for o, s, q in zip(ORG[b], SGN[b], p))


def to_local(b, g):
    return tuple((x - o) * s for o, s, x in zip(ORG[b], SGN[b], g))


def patches():
    """Every interface patch, from the blocks' global face rectangles: a face
    on the plane x_ax = v meets the faces of the partner plane (v itself
    inside the box; 0 <-> N across the periodic boundary) wherever their
    rectangles overlap.  Both sides list the same global corners."""
    out = []  # (block, lo, hi, neighbour, tlo, thi, periodic)
    faces = []
    for b, n in enumerate(DIMS):
        for ax in range(3):
            for side in (0, 1):
                lo = [0, 0, 0]
                hi = list(n)
                lo[ax] = hi[ax] = side * n[ax]
                g0, g1 = to_global(b, lo), to_global(b, hi)
                faces.append((b, ax, g0[ax], [min(g0[a], g1[a]) for a in range(3)],
                              [max(g0[a], g1[a]) for a in range(3)]))
    for b, ax, v, lo, hi in faces:
        partner = N - v if v in (0, N) else v
        periodic = int(v in (0, N))
        for b2, ax2, v2, lo2, hi2 in faces:
            if ax2 != ax or v2 != partner or (b2 == b and v2 == v):
                continue
            c0 = [max(lo[a], lo2[a]) for a in range(3)]
            c1 = [min(hi[a], hi2[a]) for a in range(3)]
            if any(c1[a] <= c0[a] for a in range(3) if a != ax):
                continue
            t0, t1 = list(c0), list(c1)
            c0[ax] = c1[ax] = v
            t0[ax] = t1[ax] = partner
            out.append((b, to_local(b, c0), to_local(b, c1), b2, to_local(b2, t0), to_local(b2, t1), periodic))
    return out


def three_block_dir(tmp_path, one, extra=None, drop=None):
    """The box of `one` as the blocks A, B and C: A's i faces are each shared
    by B and C (one through plain patches, one through periodic ones)."""
    d = str(tmp_path / "THREE")
    shutil.copytree(one, d)
    (x, y, z), = read_grd(os.path.join(one, "geometry", "grid0000.grd"))
    blocks = []
    for b, n in enumerate(DIMS):
        ii = np.array([to_global(b, (q, 0, 0))[0] for q in range(n[0] + 1)])
        jj = np.array([to_global(b, (0, q, 0))[1] for q in range(n[1] + 1)])
        kk = np.array([to_global(b, (0, 0, q))[2] for q in range(n[2] + 1)])
        blocks.append([a[np.ix_(ii, jj, kk)] for a in (x, y, z)])
    write_grd(os.path.join(d, "geometry", "grid0000.grd"), blocks)
    P = patches()
    if drop is not None:
        P = [p for p in P if not drop(p)]
    lines = [patch_line(*p) for p in sorted(P, key=lambda p: p[0])]  # the file lists each block's patches in turn
    cnt = [sum(1 for p in P if p[0] == b) for b in range(3)]
    with open(os.path.join(d, "geometry", "boundary_condition", "inner0000.txt"), "w") as f:
        f.write(f"{len(P)}\n{' '.join(map(str, cnt))}\n")
        f.write("imin imax jmin jmax kmin kmax block name   (this side, then the target side)\n")
        f.write("\n".join(lines) + "\n")
    bnd = ["0"] * 3 if extra is None else extra
    with open(os.path.join(d, "geometry", "boundary_condition", "boundary0000.txt"), "w") as f:
        f.write("\n".join(bnd) + "\n")
    return d


def cell_of(b, n):
    """Global indices of block b's cells (lo node of each cell's global span)."""
    ax = [np.array([min(to_global(b, tuple(q if a == c else 0 for c in range(3)))[a],
                        to_global(b, tuple(q + 1 if a == c else 0 for c in range(3)))[a]) for q in range(n[a])])
          for a in range(3)]
    return np.meshgrid(*ax, indexing="ij")


def face_of(b, n, f):
    """Global indices of block b's faces of family f (the node plane along f)."""
    ax = []
    for a in range(3):
        if a == f:
            ax.append(np.array([to_global(b, tuple(q if a == c else 0 for c in range(3)))[a] for q in range(n[a] + 1)]))
        else:
            ax.append(np.array([min(to_global(b, tuple(q if a == c else 0 for c in range(3)))[a],
                                    to_global(b, tuple(q + 1 if a == c else 0 for c in range(3)))[a])
                                for q in range(n[a])]))
    return np.meshgrid(*ax, indexing="ij")


def fill(orc, blocks):
    """A smooth, non-uniform state as functions of the global cell and face
    indices (periodic in each), written into the blocks of either case; a
    face flux changes sign where the block's face normal runs against the
    global axis."""
    t = 2 * np.pi / N

    def cell(i, j, k):
        rho = 1.0 + 0.2 * np.sin(t * i) * np.cos(t * j) + 0.1 * np.cos(t * k)
        u = 0.3 * np.cos(t * j + 0.4), -0.2 * np.sin(t * i), 0.1 * np.sin(t * (i + k))
        p = 0.6 + 0.1 * np.cos(t * (i - j))
        return rho, u, p

    def face(a, i, j, k):
        idx = (i % N, j % N, k % N)
        return 1e-2 * (0.5 + 0.3 * np.sin(t * idx[(a + 1) % 3]) + 0.2 * np.cos(t * idx[(a + 2) % 3] + a))

    for b in blocks:
        n = DIMS[b] if len(blocks) > 1 else (N, N, N)
        rho, u, p = cell(*cell_of(b, n))
        U = orc.view("U_H", len(blocks) > 1 and b or 0)
        sl = tuple(slice(G, G + n[a]) for a in range(3))
        U[sl + (0,)] = rho
        for a in range(3):
            U[sl + (1 + a,)] = rho * u[a]
        U[sl + (4,)] = p / (5.0 / 3.0 - 1) + 0.5 * rho * (u[0] ** 2 + u[1] ** 2 + u[2] ** 2)
        for a, name in enumerate(FACES):
            m = [n[q] + (1 if q == a else 0) for q in range(3)]
            B = orc.view(name, len(blocks) > 1 and b or 0)
            B[tuple(slice(G, G + m[q]) for q in range(3)) + (0,)] = SGN[b][a] * face(a, *face_of(b, n, a))
    orc.sync("Ucell")
    orc.sync("Bface")
    olib_update(orc)


def olib_update(orc):
    from oracle_bind import olib
    olib().or_update_derived(orc.c)


@pytest.mark.parametrize("curved", [False, True])
@pytest.mark.parametrize("layout", ["aligned", "turned"])
def test_split_faces_step_like_the_single_block(tmp_path, layout, curved):
    """Aligned, every block computes its cells in the single block's order,
    so the values are equal bit for bit; with C turned, C sums its face and
    edge terms in the reversed order along i and j, so they agree to
    round-off (relative 1e-12 of each field's largest value).  On the curved
    grid the turned block's metric, B_cell / J_cell weights and UCT edge
    sums run with its i and j axes reversed on skewed cells, which a sign
    or index slip in any of them breaks."""
    use(layout)
    one = one_block_dir(tmp_path, curved)
    three = three_block_dir(tmp_path, one)
    h1 = HostCase(CaseDirDesc(one, WindDesc()), ghost=G)
    h3 = HostCase(CaseDirDesc(three, WindDesc()), ghost=G)
    assert h3.faces(0)[0][0] == 2 and h3.faces(0)[1][0] == 2  # A's i faces: shared by B and C
    assert [len(h3.face_parts(0, f)) for f in range(6)] == [2, 2, 0, 0, 0, 0]
    assert h3.emf_shell_ok()
    o1, o3 = Oracle(h1), Oracle(h3)
    fill(o1, [0])
    fill(o3, [0, 1, 2])
    u0 = o1.view("U_H", 0).copy()
    dt1, dt3 = o1.step(3), o3.step(3)
    assert dt1 == dt3
    assert np.abs(o1.view("U_H", 0) - u0).max() > 1e-4  # the state did evolve
    U1 = o1.view("U_H", 0)
    Bs1 = [o1.view(f, 0) for f in FACES]

    def same(got, ref, what):
        if layout == "aligned":
            assert np.array_equal(got, ref), what
        else:
            scale = np.abs(ref).max(axis=tuple(range(ref.ndim - 1))) if ref.ndim == 4 else np.abs(ref).max()
            assert (np.abs(got - ref) <= 1e-12 * scale).all(), what

    for b, n in enumerate(DIMS):
        I, J, K = cell_of(b, n)
        same(o3.view("U_H", b)[G:G + n[0], G:G + n[1], G:G + n[2]], U1[G + I, G + J, G + K], ("U_H", b))
        for a, name in enumerate(FACES):
            m = [n[q] + (1 if q == a else 0) for q in range(3)]
            I, J, K = face_of(b, n, a)
            got = o3.view(name, b)[G:G + m[0], G:G + m[1], G:G + m[2], 0]
            same(got, SGN[b][a] * Bs1[a][G + I, G + J, G + K, 0], (name, b))


def test_split_faces_round_trip_through_the_case_directory(tmp_path):
    use("turned")
    three = three_block_dir(tmp_path, one_block_dir(tmp_path))
    h3 = HostCase(CaseDirDesc(three, WindDesc()), ghost=G)
    back = str(tmp_path / "BACK")
    h3.write_dir(back)
    hb = HostCase(CaseDirDesc(back, WindDesc()), ghost=G)
    for b in range(3):
        assert np.array_equal(h3.faces(b), hb.faces(b))
        for f in range(6):
            assert np.array_equal(h3.face_parts(b, f), hb.face_parts(b, f))
        for name in ("Jac", "JDxi", "Bcell_from_Bface_w"):
            assert np.array_equal(h3.field(name, b), hb.field(name, b))


def test_faces_mixing_patch_kinds_or_leaving_gaps_are_refused(tmp_path):
    use("aligned")
    one = one_block_dir(tmp_path)
    # drop the patch linking A's +i face to C: the face is half covered
    a_to_c = lambda p: p[0] == 0 and p[3] == 2 and p[6] == 0  # noqa: E731
    with pytest.raises(MhdError, match="cover the face"):
        HostCase(CaseDirDesc(three_block_dir(tmp_path / "gap", one, drop=a_to_c), WindDesc()), ghost=G)
    # a far-field patch in its place (type 2: the second name of the
    # fvbnd list): interfaces and a condition on one face
    ff = ["1", f"{CI} {CI} {CJ} {N} 0 {N} 2", "0", "0"]
    with pytest.raises(MhdError, match="mixes"):
        HostCase(CaseDirDesc(three_block_dir(tmp_path / "mix", one, extra=ff, drop=a_to_c), WindDesc()), ghost=G)


def test_split_faces_through_the_drop_in_calls(tmp_path):
    """The turned three-block case as a driver hands it over: faces of kind 2
    with their patches.  Nodes, faces, patches, geometry, op lists and the
    oracle's steps equal those of the case its CASE directory gives."""
    use("turned")
    h3 = HostCase(CaseDirDesc(three_block_dir(tmp_path, one_block_dir(tmp_path)), WindDesc()), ghost=G)
    k, blocks = case_parts(h3, geometry=False)
    assert sorted(blocks[0]["parts"]) == [0, 1]
    back = HostCase.from_parts(k, blocks, ghost=G)
    for b in range(3):
        assert np.array_equal(h3.faces(b), back.faces(b))
        for f in range(6):
            assert np.array_equal(h3.face_parts(b, f), back.face_parts(b, f))
        for p, q in zip(h3.nodes(b), back.nodes(b)):
            assert np.array_equal(p, q)
        for name in ("Jac", "JDxi", "JDet", "JDze", "Bcell_from_Bface_w", "Jcell_from_Jedge_w", "U_H"):
            assert np.array_equal(h3.field(name, b), back.field(name, b)), (name, b)
    for sy in range(14):
        assert h3.op_count(sy) == back.op_count(sy), sy
    o3, ob = Oracle(h3), Oracle(back)
    fill(o3, [0, 1, 2])
    fill(ob, [0, 1, 2])
    assert o3.step(2) == ob.step(2)
    for b in range(3):
        for name in ("U_H",) + FACES:
            assert np.array_equal(o3.view(name, b), ob.view(name, b)), (name, b)


def test_inconsistent_split_faces_are_refused(tmp_path):
    use("aligned")
    h3 = HostCase(CaseDirDesc(three_block_dir(tmp_path, one_block_dir(tmp_path)), WindDesc()), ghost=G)

    def refused(change, reason):
        k, blocks = case_parts(h3, geometry=False)
        change(blocks)
        with pytest.raises(MhdError, match=reason):
            HostCase.from_parts(k, blocks, ghost=G)

    def drop_part(blocks):
        blocks[0]["parts"][1] = blocks[0]["parts"][1][:1]
    refused(drop_part, "two or more patches")

    def shrink(blocks):
        p = blocks[0]["parts"][1].copy()
        p[0, 3] -= 1  # the first patch stops one cell short along k
        blocks[0]["parts"][1] = p
    refused(shrink, "cover the face")

    def shift(blocks):
        p = blocks[0]["parts"][1].copy()
        p[0, 15] += 1  # the first patch lands one cell further along k
        blocks[0]["parts"][1] = p
    refused(shift, "link back")

    def no_parts(blocks):
        del blocks[0]["parts"]
    refused(no_parts, "patches missing")

    def kind_one(blocks):
        f = blocks[0]["faces"].copy()
        f[1][0] = 1  # A's +i face a single interface to B: B's part of it links back, C's does not
        blocks[0]["faces"] = f
        del blocks[0]["parts"][1]
    refused(kind_one, "link back|another position|outside")
