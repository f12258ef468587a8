"""Tecplot binary output (CPU only): the reference's Mixed-mode .plt writer.

The file is parsed here from the published #!TDV112 layout the reference
writes (src/5_io/IOModule_Tecplot.cpp:262-579): header, one zone record per
block, the 357.0 end-of-header marker, then per zone a data-section header,
the min/max of every variable and the float32 data, variable-major, i
fastest.  Cell variables cover IMax*JMax*(KMax-1) values with zeros outside
the cells.  Values are pinned by invariants rather than by the writer's
formulas: node coordinates are the grid's, u*rho gives back the momentum and
p/(gamma-1) + ke the energy, a uniform induced field is reconstructed from
its face fluxes, and the Mercury neutral density times the photoionization
frequency gives back the stored production rate.
"""
import struct

import numpy as np
import pytest

from beta1_amd import HostCase, MercuryDesc, WindDesc
from beta1_amd.fields import interior

G = 2
LUNAR_VARS = ["x", "y", "z", "u_H", "v_H", "w_H", "p_H", "T_H", "Bx", "By", "Bz", "bx_ind", "by_ind", "bz_ind",
              "Jx", "Jy", "Jz"]
MERCURY_VARS = ["x", "y", "z", "u_H", "v_H", "w_H", "p_H", "T_H", "u_Na", "v_Na", "w_Na", "p_Na", "T_Na",
                "Bx", "By", "Bz", "bx_ind", "by_ind", "bz_ind", "n_Na", "Jx", "Jy", "Jz"]


def read_plt(path):
    raw = open(path, "rb").read()
    off = 0

    def take(fmt):
        nonlocal off
        v = struct.unpack_from("<" + fmt, raw, off)
        off += struct.calcsize("<" + fmt)
        return v if len(v) > 1 else v[0]

    def string():
        s = []
        while True:
            c = take("i")
            if c == 0:
                return "".join(map(chr, s))
            s.append(c)

    assert raw[:8] == b"#!TDV112"
    off = 8
    assert take("ii") == (1, 0)
    title = string()
    nvar = take("i")
    names = [string() for _ in range(nvar)]
    zones = []
    while True:
        marker = take("f")
        if marker == 357.0:
            break
        assert marker == 299.0
        z = {"name": string()}
        z["parent"], z["strand"] = take("ii")
        z["time"] = take("d")
        z["color"], z["ordered"], z["has_loc"] = take("iii")
        z["loc"] = list(take(f"{nvar}i")) if nvar > 1 else [take("i")]
        z["neigh"] = take("ii")
        z["dims"] = take("iii")
        z["aux"] = take("i")
        zones.append(z)
    for z in zones:
        assert take("f") == 299.0
        assert list(take(f"{nvar}i")) == [1] * nvar  # float data
        assert take("iii") == (0, 0, -1)
        z["minmax"] = [take("dd") for _ in range(nvar)]
        I, J, K = z["dims"]
        z["data"] = []
        for v in range(nvar):
            shp = (K, J, I) if z["loc"][v] == 0 else (K - 1, J, I)
            n = int(np.prod(shp))
            z["data"].append(np.frombuffer(raw, "<f4", n, off).reshape(shp).transpose(2, 1, 0))
            off += 4 * n
    assert off == len(raw)
    return title, names, zones


def perturb(hc, seed):
    rng = np.random.default_rng(seed)
    for b in range(hc.nblocks):
        for name in ["U_H", "U_Na"] if hc.desc.__class__ is MercuryDesc else ["U_H"]:
            U = hc.field(name, b)
            rho = U[..., 0] * (1.0 + 0.2 * rng.random(U.shape[:3]))
            vel = U[..., 1:4] / U[..., :1] + 0.3 * rng.standard_normal(U.shape[:3] + (3,))
            p = 0.1 + rng.random(U.shape[:3])
            U[..., 0] = rho
            U[..., 1:4] = rho[..., None] * vel
            U[..., 4] = p / (5.0 / 3.0 - 1.0) + 0.5 * rho * (vel * vel).sum(-1)
            U[G + 1, G + 1, G + 1, 4] = 0.0  # internal energy below zero: the pressure floor
            hc.set_field(name, b, U)
        for name in ["B_cell", "J_cell"]:
            hc.set_field(name, b, rng.standard_normal(hc.field(name, b).shape))


def check_zone_values(hc, z, names, gamma, coeffs):
    gn = G + 1
    ni, nj, nk = hc.n
    b = int(z["name"][4:].split("_")[0])
    d = dict(zip(names, z["data"]))
    mm = dict(zip(names, z["minmax"]))
    X = hc.nodes(b)
    for c, a in zip("xyz", X):
        node = a[gn:gn + ni + 1, gn:gn + nj + 1, gn:gn + nk + 1]
        assert np.array_equal(d[c], node.astype(np.float32))
        assert mm[c] == (node.min(), node.max())
    sl = interior("U_H", hc.n, G)
    for sp, coeff in coeffs.items():
        U = hc.field("U_" + sp, b)[sl]
        u, v, w, p, T = (d[f"{q}_{sp}"][:ni, :nj, :nk].astype(np.float64) for q in "uvwpT")
        rho = U[..., 0]
        for q, m in zip((u, v, w), (1, 2, 3)):
            assert np.allclose(q * rho, U[..., m], rtol=3e-7, atol=1e-30)
        ke = 0.5 * rho * (u * u + v * v + w * w)
        assert p[1, 1, 1] == np.float32(1e-12)
        ok = np.ones(p.shape, bool)
        ok[1, 1, 1] = False
        assert np.allclose((p / (gamma - 1) + ke)[ok], U[..., 4][ok], rtol=1e-5)
        assert np.allclose(T * rho * coeff, p, rtol=1e-6, atol=1e-30)
    for name, comps in (("B_cell", ("Bx", "By", "Bz")), ("J_cell", ("Jx", "Jy", "Jz"))):
        F = hc.field(name, b)[sl]
        for m, c in enumerate(comps):
            assert np.array_equal(d[c][:ni, :nj, :nk], F[..., m].astype(np.float32))
    for name in names[3:]:
        a = d[name]
        assert a.shape == (ni + 1, nj + 1, nk)
        # padding outside the cells is zero and takes part in min/max
        assert not a[ni].any() and not a[:, nj].any()
        lo, hi = mm[name]
        assert np.float32(lo) == a.min() and np.float32(hi) == a.max()
        assert lo <= 0.0 <= hi


@pytest.mark.parametrize("desc, names", [(WindDesc(nang=4, nrad=6), LUNAR_VARS),
                                         (MercuryDesc(nang=4, nrad=6), MERCURY_VARS)])
def test_tecplot_layout_and_values(tmp_path, desc, names):
    hc = HostCase(desc, ghost=G)
    perturb(hc, 3)
    path = str(tmp_path / "flow.plt")
    hc.write_tecplot(path, time=2.5)
    title, got_names, zones = read_plt(path)
    assert title == "flow_field" and got_names == names
    assert [z["name"] for z in zones] == [f"Zone{b}_Fluid" for b in range(6)]
    ni, nj, nk = hc.n
    nvar = len(names)
    for z in zones:
        assert (z["parent"], z["strand"], z["time"], z["color"], z["ordered"], z["has_loc"]) == (-1, -2, 2.5, -1,
                                                                                               0, 1)
        assert z["loc"] == [0, 0, 0] + [1] * (nvar - 3)
        assert z["neigh"] == (0, 0) and z["aux"] == 0
        assert z["dims"] == (ni + 1, nj + 1, nk + 1)
    k = hc.constants()
    coeffs = {"H": k[7]}
    if desc.__class__ is MercuryDesc:
        coeffs["Na"] = hc.constants_two_fluid()[3]
    for z in zones:
        check_zone_values(hc, z, names, k[0], coeffs)


def test_induced_cell_field_reconstructs_a_uniform_field(tmp_path):
    """Face fluxes of a uniform b0 (b0 . face area vector) give bx_ind..bz_ind = b0."""
    hc = HostCase(WindDesc(nang=4, nrad=6), ghost=G)
    b0 = np.array([0.3, -1.1, 0.7])
    for b in range(hc.nblocks):
        for f, A in (("B_xi", "JDxi"), ("B_eta", "JDet"), ("B_zeta", "JDze")):
            hc.set_field(f, b, hc.field(A, b) @ b0[:, None])
    path = str(tmp_path / "u.plt")
    hc.write_tecplot(path)
    _, names, zones = read_plt(path)
    ni, nj, nk = hc.n
    for z in zones:
        d = dict(zip(names, z["data"]))
        for m, c in enumerate(("bx_ind", "by_ind", "bz_ind")):
            assert np.allclose(d[c][:ni, :nj, :nk], b0[m], rtol=0, atol=2e-6), (z["name"], c)


def test_mercury_neutral_density_gives_the_production_rate(tmp_path):
    hc = HostCase(MercuryDesc(nang=4, nrad=6), ghost=G)
    path = str(tmp_path / "m.plt")
    hc.write_tecplot(path)
    _, names, zones = read_plt(path)
    ni, nj, nk = hc.n
    sl = interior("U_H", hc.n, G)
    for z in zones:
        b = int(z["name"][4:].split("_")[0])
        n = z["data"][names.index("n_Na")][:ni, :nj, :nk].astype(np.float64)
        photo = hc.field("Photo_rate", b)[sl][..., 0]
        drag = hc.field("drag_rate", b)[sl][..., 0]  # the photoionization frequency of the cell
        assert n.max() > 0
        assert np.allclose(n * drag, photo, rtol=1e-6, atol=0)


def test_tecplot_rejects_unwritable_path(tmp_path):
    from beta1_amd.capi import MhdError
    hc = HostCase(WindDesc(nang=4, nrad=6), ghost=G)
    with pytest.raises(MhdError):
        hc.write_tecplot(str(tmp_path / "no" / "such" / "dir.plt"))
