"""The drop-in set-up path (CPU only): a case assembled from a driver's own
arrays -- interior nodes, faces in the IndexTransform convention, optional
geometry fields -- through mhd_host_case_begin / add_block / set_topology /
case_end.  Given the arrays of a generated case it reproduces that case bit
for bit (nodes incl. the ghosts it fills, geometry, weight tables, op lists,
initial state); inconsistent input is refused with a reason."""
import numpy as np
import pytest

from beta1_amd import HostCase, OTDesc, WindDesc, case_parts
from beta1_amd.capi import MhdError
from oracle_bind import Oracle

G = 2
FIELDS = ["U_H", "B_xi", "B_eta", "B_zeta", "Badd_xi", "Badd_eta", "Badd_zeta", "Jac", "JDxi", "JDet", "JDze",
          "beta_xi", "beta_eta", "beta_zeta", "alpha_xi", "alpha_eta", "alpha_zeta", "Bcell_from_Bface_w",
          "Jcell_from_Jedge_w"]


def assert_same_case(a, b, fields=FIELDS):
    assert a.nblocks == b.nblocks and a.n == b.n and a.g == b.g
    for blk in range(a.nblocks):
        assert np.array_equal(a.faces(blk), b.faces(blk)), blk
        for p, q in zip(a.nodes(blk), b.nodes(blk)):
            assert np.array_equal(p, q), blk
        for f in fields:
            assert np.array_equal(a.field(f, blk), b.field(f, blk)), (blk, f)
    for sy in range(14):
        assert a.op_count(sy) == b.op_count(sy), sy


@pytest.mark.parametrize("geometry", [True, False])
def test_parts_of_a_shell_reproduce_it_bitwise(geometry):
    hc = HostCase(WindDesc(nang=6, nrad=8), ghost=G)
    k, blocks = case_parts(hc, geometry=geometry)
    assert_same_case(hc, HostCase.from_parts(k, blocks, ghost=G))


def test_parts_of_a_periodic_box_reproduce_its_grid():
    """Self-links with a negative block id in the reference are periodic
    links here; the ghost nodes stay mirrored.  The state is the Lunar
    initial state of the constants; the box's own state uploads back."""
    hc = HostCase(OTDesc(n=8), ghost=G)
    k, blocks = case_parts(hc)
    back = HostCase.from_parts(k, blocks, ghost=G)
    assert_same_case(hc, back, fields=[f for f in FIELDS if f not in ("U_H", "B_xi", "B_eta", "B_zeta")])
    for f in ("U_H", "B_xi", "B_eta", "B_zeta"):
        back.set_field(f, 0, hc.field(f, 0))
        assert np.array_equal(back.field(f, 0), hc.field(f, 0))


def test_oracle_of_parts_steps_like_oracle_of_generated_case():
    hc = HostCase(WindDesc(nang=6, nrad=8), ghost=G)
    back = HostCase.from_parts(*case_parts(hc), ghost=G)
    a, b = Oracle(hc), Oracle(back)
    assert a.step(3) == b.step(3)
    for blk in range(6):
        for f in ("U_H", "B_xi", "B_eta", "B_zeta", "J_cell"):
            assert np.array_equal(a.view(f, blk), b.view(f, blk)), (blk, f)


def spoiled(change):
    hc = HostCase(WindDesc(nang=6, nrad=8), ghost=G)
    k, blocks = case_parts(hc)
    change(k, blocks)
    return k, blocks


def _link_not_back(k, blocks):
    f = blocks[0]["faces"]
    iface = [q for q in range(6) if f[q][0] == 1][0]
    f[iface][3] = (f[iface][3] + 1) % 6


def _offset(k, blocks):
    f = blocks[1]["faces"]
    iface = [q for q in range(6) if f[q][0] == 1][0]
    f[iface][11] += 1


def _metric(k, blocks):
    blocks[2]["geometry"]["JDxi"] = blocks[2]["geometry"]["JDxi"] * (1 + 1e-6)


def _table(k, blocks):
    blocks[3]["geometry"]["Jcell_from_Jedge_w"] = blocks[3]["geometry"]["Jcell_from_Jedge_w"] * (1 + 1e-6)


def _bc(k, blocks):
    blocks[0]["faces"][4][1] = 7


def _perm(k, blocks):
    f = blocks[0]["faces"]
    iface = [q for q in range(6) if f[q][0] == 1][0]
    f[iface][5:8] = [0, 0, 1]


def _node(k, blocks):
    blocks[4]["nodes"][0][0, 0, 0] += 1e-3  # a corner node no longer matches its neighbours


def _gamma(k, blocks):
    k[0] = 1.0


@pytest.mark.parametrize("change,reason", [
    (_link_not_back, "does not link back"), (_offset, "does not link back|another position|outside the neighbour"),
    (_metric, "JDxi"), (_table, "Jcell_from_Jedge_w"), (_bc, "bc 1"), (_perm, "permutation"),
    (_node, "another position"), (_gamma, "mhd_host_case_begin")])
def test_inconsistent_parts_are_refused_with_a_reason(change, reason):
    import re
    k, blocks = spoiled(change)
    with pytest.raises(MhdError) as e:
        HostCase.from_parts(k, blocks, ghost=G)
    assert re.search(reason, str(e.value)), str(e.value)


def test_missing_block_or_wrong_node_count_is_refused():
    hc = HostCase(WindDesc(nang=6, nrad=8), ghost=G)
    k, blocks = case_parts(hc)
    with pytest.raises(MhdError, match="nodes or topology missing"):
        _five_of_six(k, blocks)
    bad = dict(blocks[0])
    bad["nodes"] = [a[:-1] for a in blocks[0]["nodes"]]
    with pytest.raises(ValueError):
        HostCase.from_parts(k, [bad] + blocks[1:], ghost=G)


def _five_of_six(k, blocks):
    """Begin six blocks, give five, end: the sixth is missing."""
    import ctypes as C

    from beta1_amd.capi import lib
    from beta1_amd.solver import _dptr, _parts_args

    h = C.c_void_p()
    kk = np.ascontiguousarray(k, dtype=np.float64)
    assert lib().mhd_host_case_begin(6, G, _dptr(kk), C.byref(h)) == 0
    try:
        for ib, (ni, nj, nk, xyz, keep, ptrs, faces, _parts) in enumerate(_parts_args(blocks[:5], G)):
            assert lib().mhd_host_add_block(h, ib, ni, nj, nk, _dptr(xyz[0]), _dptr(xyz[1]), _dptr(xyz[2]),
                                            xyz[0].size, ptrs) == 0
            assert lib().mhd_host_set_topology(h, ib, faces) == 0
        rc = lib().mhd_host_case_end(h)
        if rc:
            raise MhdError(lib().mhd_host_last_error(h).decode(), rc)
    finally:
        lib().mhd_host_case_destroy(h)


def two_box_parts(perturb=1e-3, seed=4):
    """A box [0,1] x [0,0.6] x [0,0.8] in two blocks of unequal extent
    (4 and 6 cells along x, the interface at x = 0.4), periodic in y and z,
    far field at both x ends; the Lunar constants of the default wind."""
    k = HostCase(WindDesc(nang=4, nrad=6), ghost=G).constants()
    nj, nk = 6, 8
    blocks = []
    for ni, x0, x1 in ((4, 0.0, 0.4), (6, 0.4, 1.0)):
        xs = np.linspace(x0, x1, ni + 1)
        ys, zs = np.linspace(0.0, 0.6, nj + 1), np.linspace(0.0, 0.8, nk + 1)
        X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
        blocks.append({"n": (ni, nj, nk), "nodes": [X, Y, Z], "faces": np.zeros((6, 14), dtype=np.int64)})
    ident = [0, 1, 2, 1, 1, 1]
    for b, blk in enumerate(blocks):
        ni = blk["n"][0]
        f = blk["faces"]
        for face, shift in ((2, (0, nj, 0)), (3, (0, -nj, 0)), (4, (0, 0, nk)), (5, (0, 0, -nk))):
            f[face] = [1, 0, b, face ^ 1, 1] + ident + list(shift)
        if b == 0:
            f[0] = [0, 1, -1, -1, 0] + [0] * 9                      # far field at x = 0
            f[1] = [1, 0, 1, 0, 0] + ident + [-ni, 0, 0]            # -> block 1, x = 0.4
        else:
            f[0] = [1, 0, 0, 1, 0] + ident + [blocks[0]["n"][0], 0, 0]
            f[1] = [0, 1, -1, -1, 0] + [0] * 9                      # far field at x = 1
    return k, blocks


def walled_box_parts(n, sides, stretch=1.2):
    """One block of n = (ni, nj, nk) cells over [0,1] x [0,1] x [0,2] (z
    stretched), each axis either periodic ("p") or a pair of physical sides
    (1 far field, 2 absorbing wall) for its low and high face; the Lunar
    constants of the default wind."""
    k = HostCase(WindDesc(nang=4, nrad=6), ghost=G).constants()
    ni, nj, nk = n
    xs, ys, zs = np.linspace(0, 1, ni + 1), np.linspace(0, 1, nj + 1), np.linspace(0, 2, nk + 1) ** stretch
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    f = np.zeros((6, 14), dtype=np.int64)
    ident = [0, 1, 2, 1, 1, 1]
    for ax, kind in enumerate(sides):
        lo, hi = 2 * ax, 2 * ax + 1
        if kind == "p":
            shift = [0, 0, 0]
            shift[ax] = n[ax]
            f[lo] = [1, 0, 0, hi, 1] + ident + shift
            f[hi] = [1, 0, 0, lo, 1] + ident + [-v for v in shift]
        else:
            f[lo] = [0, kind[0], -1, -1, 0] + [0] * 9
            f[hi] = [0, kind[1], -1, -1, 0] + [0] * 9
    return k, [{"n": tuple(n), "nodes": [X, Y, Z], "faces": f}]


def test_blocks_of_unequal_extent():
    """Blocks of different sizes share one slot size in a case; the
    geometry of each is its own box's (closed form) and the halo carries the
    neighbour's cells across the interface."""
    k, blocks = two_box_parts()
    hc = HostCase.from_parts(k, blocks, ghost=G)
    assert hc.nblocks == 2
    for b, (ni, h) in enumerate(((4, 0.1), (6, 0.1))):
        jac = hc.field("Jac", b)[G:G + ni, G:G + 6, G:G + 8, 0]
        assert np.allclose(jac, h * 0.1 * 0.1, rtol=1e-12)
    # an oracle step on it keeps a uniform flow uniform (free stream)
    orc = Oracle(hc)
    U0 = [orc.view("U_H", b).copy() for b in range(2)]
    orc.step(2)
    for b, ni in enumerate((4, 6)):
        a, r = orc.view("U_H", b)[G:G + ni, G:-G, G:-G], U0[b][G:G + ni, G:-G, G:-G]
        assert np.abs(a - r).max() <= 1e-13 * np.abs(r).max()


@pytest.mark.parametrize("make", [
    lambda: HostCase(WindDesc(nang=6, nrad=8), ghost=G),
    lambda: HostCase(OTDesc(n=8), ghost=G),
    lambda: HostCase(WindDesc(nang=6, nrad=8), ghost=3),
    lambda: HostCase.from_parts(*two_box_parts(), ghost=G),
])
def test_op_lists_are_order_independent(make):
    """Every op list of a case writes each element once and reads none it
    writes, so the device may apply its records in any order and load an
    element's sources before its stores."""
    hc = make()
    for sy in range(14):
        if hc.op_count(sy):
            assert hc.ops_independent(sy), sy


@pytest.mark.parametrize("make", [
    lambda: HostCase(WindDesc(nang=6, nrad=8), ghost=G),
    lambda: HostCase(OTDesc(n=8), ghost=G),
    lambda: HostCase(WindDesc(nang=6, nrad=8), ghost=3),
    lambda: HostCase.from_parts(*two_box_parts(), ghost=G),
    lambda: HostCase.from_parts(*walled_box_parts((5, 6, 7), ((2, 1), (1, 1), "p")), ghost=G),
    lambda: HostCase.from_parts(*walled_box_parts((4, 5, 6), ("p", (2, 2), (1, 1))), ghost=3),
])
def test_eface_sync_is_what_the_one_pass_emf_assumes(make):
    """The whole step forms the face EMF candidates inside the edge EMF
    kernel and only the shell's through memory: every Eface halo copy must
    read a face on the shell (the first layer inside an interface side) and
    write a ghost face the kernel loads, the shell enumeration must list each
    shell point once, and the wall handler must zero exactly the faces the
    kernel takes as zero."""
    assert make().emf_shell_ok()


def test_mercury_op_lists_are_order_independent():
    from beta1_amd import MercuryDesc
    hc = HostCase(MercuryDesc(nang=6, nrad=8), ghost=G)
    assert hc.op_count(10) and hc.op_count(12)
    for sy in range(14):
        if hc.op_count(sy):
            assert hc.ops_independent(sy), sy


def test_walls_meeting_along_an_edge_are_refused():
    """Walls on two axes of one block share an edge, where the reference's
    face-by-face wall order is not what the library's order-free wall phases
    give: such a case is refused with the reason, walls on one axis are not."""
    from beta1_amd.capi import MhdError
    with pytest.raises(MhdError, match="meet along an edge"):
        HostCase.from_parts(*walled_box_parts((5, 6, 7), ((2, 1), (1, 2), "p")), ghost=G)
    HostCase.from_parts(*walled_box_parts((5, 6, 7), ((2, 2), (1, 1), "p")), ghost=G)
