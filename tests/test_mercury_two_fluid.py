"""Two-fluid (H+ / Na+) Mercury step on the CPU: the host set-up against the
reference's set-up numbers and exosphere model, and the oracle against
invariants of the Mercury source terms (Z5_Mercury/src/physics/
MercurySolver_FluidRHS.cpp:206-527, MercurySolver_MagExplicit_Ambipo.cpp).

Pins (each would catch a dropped term, a wrong sign or a wrong species):
* constants recomputed from Z5_Mercury/9999setup and calc_physical_constant;
* photoionization rate against an independent numpy evaluation of the
  Exner et al. (2020) exosphere at the reference's dual-grid cell centres;
* the induction force only exchanges momentum between the species (their sum
  vanishes because u_+ is the density-weighted mean velocity);
* where the flow is uniform the Na+ mass source equals a1_Na * Photo and the
  drag removes momentum at the rate a4 * nu;
* the ambipolar edge EMF is a discrete gradient where n_e is uniform, so
  its curl (the change it makes to B) vanishes to round-off.
"""
import numpy as np
import pytest

from beta1_amd import HostCase, MercuryDesc
from beta1_amd.fields import interior
from oracle_bind import Oracle

G = 2
IN = (slice(2, -2), slice(2, -2), slice(3, -3))  # cells clear of walls and block edges


def case(**kw):
    return HostCase(MercuryDesc(nang=6, nrad=10, **kw), ghost=G)


def zero(hc, name):
    for b in range(hc.nblocks):
        hc.set_field(name, b, np.zeros_like(hc.field(name, b)))


def test_two_fluid_constants_follow_the_reference_setup():
    hc = case()
    t = hc.constants_two_fluid()
    k = hc.constants()
    U, T, n, M, B, L = 4.031e5, 1.7407e5, 4e7, 0.001, 1e-9, 2.44e6
    NA, R, mu0, qe, M_Na = 6.02e23, 8.314, 1.256637061e-6, 1.602e-19, 0.023
    rho_ref = M * n / NA
    kB = R / NA
    gm = k[0]
    want = {2: M_Na, 3: R * T / (M_Na * U * U), 9: qe * L * B * n / (rho_ref * U),
            10: L / U * (M_Na / NA / rho_ref) * 1e6, 11: L / U, 12: 3 * L * kB * n / (rho_ref * U ** 3),
            13: 1e6 * L * kB / (rho_ref * U ** 3 * (gm - 1)), 14: T, 15: 1000.0,
            16: rho_ref * U / (qe * L * B * n)}
    for i, v in want.items():
        assert t[i] == pytest.approx(v, rel=1e-14), i
    assert t[0] == 1.0 and t[1] == 1.0
    # Na+ seed state: rho_small * rho0 at rest, p0 * rho_small / 23
    rho0, p0 = k[8], (k[12] - 0.5 * (k[9] ** 2 + k[10] ** 2 + k[11] ** 2) / k[8]) * (gm - 1)
    assert t[4] == pytest.approx(1e-6 * rho0, rel=1e-14)
    assert t[5] == t[6] == t[7] == 0.0
    assert t[8] == pytest.approx(p0 * 1e-6 / 23 / (gm - 1), rel=1e-9)
    assert case(ambipolar=False).constants_two_fluid()[1] == 0.0


def exner_photo(x, y, z, L=2.44e6):
    """Exner et al. (2020, Table 2) sodium exosphere times the
    photoionization frequency (5e-5 /s, 1e-5 /s in the geometric shadow)."""
    r = np.sqrt(x * x + y * y + z * z)
    out = np.zeros_like(r)
    ok = r > 1.0
    rh = np.stack([x, y, z]) / np.where(ok, r, 1.0)
    dr = (r - 1.0) * L
    deg = np.pi / 180

    def bell(lat, lon, width):
        p = np.array([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)])
        ang = np.arccos(np.clip(np.einsum("i...,i->...", rh, p), -1, 1))
        return np.exp(-(ang / width) ** 2)

    n = 7.84e6 * np.exp(-dr / 431e3)
    n = n + 8.86e9 * np.exp(-dr / 100e3) * bell(0, 0, 15 * deg)
    n = n + 4.06e10 * np.exp(-dr / 232e3) * (bell(50 * deg, 0, 20 * deg) + bell(-50 * deg, 0, 20 * deg))
    sp = 5.67e6 * np.exp(-dr / 748e3)
    n = n + sp * (bell(80 * deg, 0, 15 * deg) + bell(-80 * deg, 0, 15 * deg))
    n = n + sp * (bell(15 * deg, np.pi, 10 * deg) + bell(-15 * deg, np.pi, 10 * deg))
    n = n / (r * r) * 1e-6
    shadow = (x < 0) & (np.sqrt(y * y + z * z) < 1)
    out[ok] = (np.where(shadow, 1e-5, 5e-5) * n)[ok]
    return out


def test_photo_rate_and_drag_follow_the_exosphere_model():
    hc = case()
    gn = G + 1
    ni, nj, nk = hc.n
    for b in range(hc.nblocks):
        X = hc.nodes(b)
        # cell centres: mean of the 8 corner nodes, cells -G .. n+G-1
        c = [0.125 * sum(a[gn - G + di:gn + ni + G + di, gn - G + dj:gn + nj + G + dj, gn - G + dk:gn + nk + G + dk]
                         for di in (0, 1) for dj in (0, 1) for dk in (0, 1)) for a in X]
        want = exner_photo(*c)
        got = hc.field("Photo_rate", b)[..., 0]
        assert np.allclose(got, want, rtol=1e-11, atol=0), b
        shadow = (c[0] < 0) & (np.sqrt(c[1] ** 2 + c[2] ** 2) < 1)
        assert np.array_equal(hc.field("drag_rate", b)[..., 0], np.where(shadow, 1e-5, 5e-5))
    assert max(hc.field("Photo_rate", b).max() for b in range(6)) > 0


def test_induction_force_only_exchanges_momentum():
    hc = case(ambipolar=False)
    zero(hc, "drag_rate")
    zero(hc, "Photo_rate")
    o = Oracle(hc)
    o.step(1)
    for b in range(hc.nblocks):
        RH, RN = o.view("RHS_H", b)[IN], o.view("RHS_Na", b)[IN]
        scale = np.abs(RH[..., 1:4]).max()
        assert scale > 1e-9
        assert np.abs(RH[..., 1:4] + RN[..., 1:4]).max() < 1e-7 * scale


def test_photoionization_mass_and_drag_rates():
    hc = case(ambipolar=False)
    t = hc.constants_two_fluid()
    zero(hc, "drag_rate")
    o = Oracle(hc)
    o.step(1)
    sl = interior("U_H", hc.n, G)
    for b in range(hc.nblocks):
        ph = hc.field("Photo_rate", b)[sl][IN][..., 0]
        assert np.array_equal(o.view("RHS_Na", b)[IN][..., 0], t[10] * ph)
    hc = case(ambipolar=False)
    zero(hc, "Photo_rate")
    o = Oracle(hc)
    o.step(1)
    for b in range(hc.nblocks):
        U = hc.field("U_H", b)[sl][IN]
        nu = hc.field("drag_rate", b)[sl][IN][..., 0]
        tot = o.view("RHS_H", b)[IN][..., 1:4] + o.view("RHS_Na", b)[IN][..., 1:4]
        want = -t[11] * nu[..., None] * U[..., 1:4]
        assert np.abs(tot - want).max() < 1e-9 * np.abs(want).max()


def test_ambipolar_emf_is_curl_free_where_density_is_uniform():
    """With uniform densities and a varying pressure, switching the ambipolar
    terms on changes the momentum but not B (away from the far field, whose
    edges the boundary condition overwrites, and from the wall, whose ghost
    cells hold other densities)."""
    out = {}
    for amb in (False, True):
        hc = case(ambipolar=amb)
        rng = np.random.default_rng(3)
        for b in range(hc.nblocks):
            U = hc.field("U_H", b)
            U[..., 4] *= 1.0 + 0.05 * rng.random(U.shape[:3])
            hc.set_field("U_H", b, U)
        for sy in (3, 4, 5):
            hc.apply_sync(sy)
        o = Oracle(hc)
        o.step(1)
        out[amb] = [(o.view("B_xi", b).copy(), o.view("U_H", b).copy()) for b in range(hc.nblocks)]
    for b in range(6):
        (b0, u0), (b1, u1) = out[False][b], out[True][b]
        assert np.abs(u1 - u0).max() > 1e-8
        # interior faces clear of the wall (whose ghost densities differ) and
        # of the far field
        assert np.abs(b1 - b0)[G:-G, G:-G, G + 2:-G - 3].max() < 1e-13 * np.abs(b0).max()


def test_two_fluid_oracle_threads_equal_serial():
    hc = case()
    a, b = Oracle(hc), Oracle(hc, threads=4)
    a.step(2)
    b.step(2)
    for blk in range(6):
        for name in ("U_H", "U_Na", "B_xi", "B_eta", "B_zeta"):
            assert np.array_equal(a.view(name, blk), b.view(name, blk))
