// Case builders: Orszag-Tang periodic box and the solar-wind shell cases.
// Initial states follow the reference's Lunar_Initial
// (Z4_Lunar/include/2_Initial.h:16-99) and LunarBackgroundState
// (Z4_Lunar/include/0_BackgroundState.h:16-67); the Orszag-Tang vortex is the
// standard 3-D extension used as the Cartesian periodic test.
#include "setup.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace mpcns
{

// Execute an op list on host state arrays (case-relative offsets).
void apply_ops_host(std::vector<BlockArrays> &blk, const std::vector<MhdOp> &ops, const MhdCaseConst &cst)
{
    const int64_t bs = (int64_t)st_nslot(blk[0].d.model) * blk[0].d.slot;
    auto at = [&](int64_t rel) -> double & { return blk[rel / bs].st[rel % bs]; };
    for (const MhdOp &run : ops)
        for (int e = 0; e < std::max(run.run, 1); ++e)
    {
        MhdOp o = run;
        o.dst += (int64_t)e * run.dstep;
        o.src += (int64_t)e * run.sstep;
        if (o.kind == OP_SOLID_FIRST || o.kind == OP_SOLID_GHOST)
        {
            // BC_Solid_Surface_ (Z4_Lunar/src/boundary/Boundary_Condition_Fluid.cpp:19-184),
            // the same arithmetic as the device op
            const double gm = cst.gamma;
            const int64_t L = o.stride;
            const double rho_r = std::max(at(o.src), 2.5e-3);
            const double r2 = std::max(rho_r, 2.5e-3);
            const double mx = at(o.src + L), my = at(o.src + 2 * L), mz = at(o.src + 3 * L), E = at(o.src + 4 * L);
            double eint = E - 0.5 * (mx * mx + my * my + mz * mz) / r2;
            if (eint < 0.0)
                eint = 0.0;
            const double p_r = std::max((gm - 1.0) * eint, 1e-12);
            at(o.dst) = rho_r;
            at(o.dst + L) = rho_r * 0.0;
            at(o.dst + 2 * L) = rho_r * 0.0;
            at(o.dst + 3 * L) = rho_r * 0.0;
            at(o.dst + 4 * L) = p_r / (gm - 1.0) + 0.5 * rho_r * (0.0 * 0.0 + 0.0 * 0.0 + 0.0 * 0.0);
            continue;
        }
        for (int c = 0; c < o.ncomp; ++c)
        {
            const int64_t d = o.dst + c * o.stride;
            switch (o.kind)
            {
            case OP_COPY:
                at(d) = at(o.src + c * o.stride);
                break;
            case OP_NEG:
                at(d) = -at(o.src + c * o.stride);
                break;
            case OP_ZERO:
                at(d) = 0.0;
                break;
            case OP_CONST:
                at(d) = o.val;
                break;
            case OP_QINF:
                at(d) = cst.qinf[c];
                break;
            case OP_QINF_NA:
                at(d) = cst.qinfs[c];
                break;
            case OP_NA_FLOOR:
                at(d) = c == 0 ? 1.0e-12 : (c == 4 ? 1.0e-12 / (cst.gamma - 1.0) : 0.0);
                break;
            default:
                throw std::runtime_error("apply_ops_host: op kind not supported on host");
            }
        }
    }
}

static void alloc_state(std::vector<BlockArrays> &blk)
{
    for (auto &A : blk)
        A.st.assign((size_t)st_nslot(A.d.model) * A.d.slot, 0.0);
}

// Copy-type boundary (DefaultPhysicalCopy, reference
// src/6_boundary/Boundary.cpp:303-331) plus halo for the three face slots
// starting at `slot0` -- used for the one-off Badd sync at start-up.
static std::vector<MhdOp> face_copy_sync(CaseSetup &cs, int slot0)
{
    std::vector<MhdOp> ops;
    const Grid &grd = cs.grid;
    const int g = grd.g;
    const int64_t bs = (int64_t)st_nslot(cs.blk[0].d.model) * cs.blk[0].d.slot;
    for (int b = 0; b < (int)grd.blocks.size(); ++b)
    {
        const Block &B = grd.blocks[b];
        const MhdBlockDev &d = cs.blk[b].d;
        for (int f = 0; f < 6; ++f)
        {
            if (B.face[f].kind != 0 || B.face[f].bc == BC_NONE)
                continue;
            const int ax = f / 2, sgn = (f % 2) ? 1 : -1;
            for (int a = 0; a < 3; ++a)
            {
                const int loc = LOC_FACE_XI + a;
                int lo[3], hi[3];
                for (int q = 0; q < 3; ++q)
                {
                    lo[q] = 0;
                    hi[q] = loc_count(loc, q, B.n(q));
                }
                const int ref = sgn < 0 ? 0 : hi[ax] - 1;
                lo[ax] = ref;
                hi[ax] = ref + 1;
                for (int i = lo[0]; i < hi[0]; ++i)
                    for (int j = lo[1]; j < hi[1]; ++j)
                        for (int k = lo[2]; k < hi[2]; ++k)
                            for (int l = 1; l <= g; ++l)
                            {
                                int q[3] = {i, j, k};
                                q[ax] += sgn * l;
                                MhdOp o{};
                                o.dst = b * bs + (int64_t)(slot0 + a) * d.slot + blk_off(d, q[0], q[1], q[2]);
                                o.src = b * bs + (int64_t)(slot0 + a) * d.slot + blk_off(d, i, j, k);
                                o.stride = d.slot;
                                o.kind = OP_COPY;
                                o.ncomp = 1;
                                ops.push_back(o);
                            }
            }
        }
    }
    // halo: reuse the B halo ops, relocated from ST_B to slot0
    for (MhdOp o : cs.ops[SY_B_HALO])
    {
        const int64_t sl = cs.blk[0].d.slot;
        auto shift = [&](int64_t rel) {
            const int64_t b = rel / bs, r = rel % bs;
            const int64_t s = r / sl, off = r % sl;
            return b * bs + (s - ST_B + slot0) * sl + off;
        };
        o.dst = shift(o.dst);
        o.src = shift(o.src);
        ops.push_back(o);
    }
    return ops;
}

static void finalize_grid(CaseSetup &cs, int model = 0)
{
    fill_ghost_nodes(cs.grid);
    build_geometry(cs.grid, cs.blk);
    cs.shape_key = geometry_key(cs.blk);
    for (auto &A : cs.blk)
    {
        A.d.case_id = 0;
        A.d.model = model;
    }
    alloc_state(cs.blk);
}

CaseSetup make_case_orszag_tang(int n, int g, double gamma, double cfl, uint64_t seed, double noise)
{
    CaseSetup cs;
    cs.grid = make_periodic_box(n, n, n, 1.0, 1.0, 1.0, g);
    finalize_grid(cs);

    MhdCaseConst &c = cs.cst;
    c.gamma = gamma;
    c.cfl = cfl;
    c.inver_MA2 = 1.0;
    c.momentum_hall_coeff = 1.0;
    c.M_H = 1.0;
    c.M_ref = 1.0;
    c.ne_floor = 0.01;
    c.state_coeff = 1.0;
    c.dt_floor = 1e-7;
    const double pi = std::acos(-1.0);
    const double rho0 = 25.0 / (36.0 * pi), p0 = 5.0 / (12.0 * pi), B0 = 1.0 / std::sqrt(4.0 * pi);
    c.qinf[0] = rho0;
    c.qinf[1] = c.qinf[2] = c.qinf[3] = 0.0;
    c.qinf[4] = p0 / (gamma - 1.0);
    c.B_imf[0] = c.B_imf[1] = c.B_imf[2] = 0.0;
    c.u_inf[0] = c.u_inf[1] = c.u_inf[2] = 0.0;

    build_ops(cs);

    BlockArrays &A = cs.blk[0];
    const MhdBlockDev &d = A.d;
    const double h = 1.0 / n;
    auto S = [&](int s, int i, int j, int k) -> double & { return A.st[(size_t)s * d.slot + blk_off(d, i, j, k)]; };
    uint64_t rs = seed;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < n; ++k)
            {
                const double x = (i + 0.5) * h, y = (j + 0.5) * h;
                const double fr = 1.0 + noise * (2.0 * uniform01(rs) - 1.0);
                const double fp = 1.0 + noise * (2.0 * uniform01(rs) - 1.0);
                const double rho = rho0 * fr, p = p0 * fp;
                const double u = -std::sin(2 * pi * y), v = std::sin(2 * pi * x), w = 0.0;
                S(ST_U0 + 0, i, j, k) = rho;
                S(ST_U0 + 1, i, j, k) = rho * u;
                S(ST_U0 + 2, i, j, k) = rho * v;
                S(ST_U0 + 3, i, j, k) = rho * w;
                S(ST_U0 + 4, i, j, k) = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w);
            }
    // face fluxes from the zeta-edge circulation a = A_z dz  (curl of edge 1-form)
    auto az = [&](int i, int j) {
        const double x = i * h, y = j * h;
        return B0 * (std::cos(4 * pi * x) / (4 * pi) + std::cos(2 * pi * y) / (2 * pi)) * h;
    };
    for (int i = 0; i <= n; ++i)
        for (int j = 0; j <= n; ++j)
            for (int k = 0; k <= n; ++k)
            {
                if (j < n && k < n)
                    S(ST_B + 0, i, j, k) = az(i, j + 1) - az(i, j);
                if (i < n && k < n)
                    S(ST_B + 1, i, j, k) = az(i, j) - az(i + 1, j);
                if (i < n && j < n)
                    S(ST_B + 2, i, j, k) = 0.0;
            }
    apply_ops_host(cs.blk, cs.ops[SY_U_HALO], c);
    apply_ops_host(cs.blk, cs.ops[SY_B_HALO], c);
    return cs;
}

// Nondimensionalisation of the Lunar set-up: reference values of
// Z4_Lunar/9999setup/1_Initial_condition.txt:4-5 and constants of
// Z4_Lunar/9999setup/0_General_Setting.txt:26-27.
struct LunarRef
{
    double U = 4.61e5, T = 1.2e5, n = 7e6, M = 0.001, B = 1e-9, L = 1.7374e6;
    double NA = 6.02e23, R = 8.314, mu0 = 1.256637061e-6, qe = 1.602e-19;
};

CaseSetup make_case_lunar(int nang, int nrad, double r_in, double r_out, int g, double gamma, double cfl,
                          const WindParams &w)
{
    return make_case_lunar_on(make_cubed_sphere_shell(nang, nrad, r_in, r_out, g, BC_SOLID, BC_FARFIELD), gamma, cfl,
                              w);
}

CaseSetup make_case_lunar_on(Grid grid, double gamma, double cfl, const WindParams &w)
{
    return make_case_with(std::move(grid), lunar_constants(gamma, cfl, w));
}

MhdCaseConst lunar_constants(double gamma, double cfl, const WindParams &w)
{
    const LunarRef R;
    MhdCaseConst c{};
    const double rho_ref = R.M * R.n / R.NA;
    const double M_H = 0.001 * w.m_amu;
    c.gamma = gamma;
    c.cfl = cfl;
    // calc_physical_constant (Z4_Lunar/src/physics/LunarSolver_Tools.cpp:28-44)
    c.inver_MA2 = R.B * R.B / (R.U * R.U * R.mu0 * rho_ref);
    c.momentum_hall_coeff = (R.B * R.B) / (R.mu0 * rho_ref * R.U * R.U);
    c.M_H = M_H;
    c.M_ref = R.M;
    c.ne_floor = 0.01;
    c.state_coeff = (R.R * R.T) / (M_H * R.U * R.U);
    c.dt_floor = 1e-7;
    // BuildLunarBackgroundState (Z4_Lunar/include/0_BackgroundState.h:16-67)
    const double k_B = R.R / R.NA;
    const double n_m3 = w.n_cc * 1e6;
    const double T_K = w.T_eV * 11604.518;
    const double rho0 = (n_m3 / R.n) * (M_H / R.M);
    const double V = w.v_kms * 1e3;
    const double u0 = w.vdir[0] * V / R.U, v0 = w.vdir[1] * V / R.U, w0 = w.vdir[2] * V / R.U;
    const double p0 = n_m3 * k_B * T_K / (rho_ref * R.U * R.U);
    c.qinf[0] = rho0;
    c.qinf[1] = rho0 * u0;
    c.qinf[2] = rho0 * v0;
    c.qinf[3] = rho0 * w0;
    c.qinf[4] = 0.5 * rho0 * (u0 * u0 + v0 * v0 + w0 * w0) + p0 / (gamma - 1.0);
    c.u_inf[0] = u0;
    c.u_inf[1] = v0;
    c.u_inf[2] = w0;
    for (int a = 0; a < 3; ++a)
        c.B_imf[a] = w.B_nT[a] * 1e-9 / R.B;
    return c;
}

CaseSetup make_case_with(Grid grid, const MhdCaseConst &cst)
{
    CaseSetup cs;
    cs.grid = std::move(grid);
    finalize_grid(cs);
    cs.cst = cst;
    const MhdCaseConst &c = cs.cst;
    build_ops(cs);

    // U = q_inf everywhere, B_face = 0, Badd = B_imf . S_face
    // (Z4_Lunar/include/2_Initial.h:50-90; Z4_Lunar/include/operators/Dipole.h:334-375)
    for (auto &A : cs.blk)
    {
        const MhdBlockDev &d = A.d;
        for (int m = 0; m < 5; ++m)
            for (int64_t o = 0; o < d.slot; ++o)
                A.st[(size_t)(ST_U0 + m) * d.slot + o] = c.qinf[m];
        for (int a = 0; a < 3; ++a)
            for (int64_t o = 0; o < d.slot; ++o)
            {
                const double *jd = &A.geo[0];
                const double sx = jd[(size_t)(G_JD + 3 * a + 0) * d.slot + o];
                const double sy = jd[(size_t)(G_JD + 3 * a + 1) * d.slot + o];
                const double sz = jd[(size_t)(G_JD + 3 * a + 2) * d.slot + o];
                A.st[(size_t)(ST_BADD + a) * d.slot + o] = c.B_imf[0] * sx + c.B_imf[1] * sy + c.B_imf[2] * sz;
            }
    }
    apply_ops_host(cs.blk, face_copy_sync(cs, ST_BADD), c);
    return cs;
}

// Cell centre as the reference's dual grid sums it (Block::calc_Dual_Grids,
// src/1_grid/1_MPCNS_Grid_GCL.cpp:6-23): dual(i+1, j+1, k+1) of cell (i, j, k).
void dual_centre(const Block &B, int i, int j, int k, double out[3])
{
    const int I = i + 1, J = j + 1, K = k + 1;
    const int64_t o[8] = {B.node_off(I, J, K),         B.node_off(I - 1, J, K),     B.node_off(I, J - 1, K),
                          B.node_off(I - 1, J - 1, K), B.node_off(I, J, K - 1),     B.node_off(I - 1, J, K - 1),
                          B.node_off(I, J - 1, K - 1), B.node_off(I - 1, J - 1, K - 1)};
    const std::vector<double> *X[3] = {&B.x, &B.y, &B.z};
    for (int a = 0; a < 3; ++a)
    {
        const std::vector<double> &v = *X[a];
        out[a] = 0.125 * (v[o[0]] + v[o[1]] + v[o[2]] + v[o[3]] + v[o[4]] + v[o[5]] + v[o[6]] + v[o[7]]);
    }
}

// Neutral Na density [cm^-3] at a cell centre: the exosphere of Exner et al.
// (2020, Table 2) -- a radially symmetric part plus Gaussian bells for
// thermal desorption, photon-stimulated desorption and sputtering; the Na+
// photoionization production rate [cm^-3 s^-1] is that density times the
// photoionization frequency, lower in the geometric shadow
// (Mercury_Initial::Initial_Na_And_PhotoRate_, Z5_Mercury/include/2_Initial.h:162-392).
double mercury_na_density(const double x[3], double L_ref)
{
    const double r_nd = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    if (!(r_nd > 1.0) || !std::isfinite(r_nd))
        return 0.0;
    const double pi = std::acos(-1.0), deg = pi / 180.0;
    const double rx = x[0] * (1.0 / r_nd), ry = x[1] * (1.0 / r_nd), rz = x[2] * (1.0 / r_nd);
    const double dr = L_ref * r_nd - L_ref;
    auto gauss2 = [](double dth, double H) { return H <= 0.0 ? 0.0 : std::exp(-((dth / H) * (dth / H))); };
    auto angle_to = [&](double lat, double lon) {
        const double cl = std::cos(lat);
        double d = rx * (cl * std::cos(lon)) + ry * (cl * std::sin(lon)) + rz * std::sin(lat);
        d = std::max(-1.0, std::min(1.0, d));
        return std::acos(d);
    };
    const double day = 0.0, night = pi;
    const double n_miv = 7.84e6 * std::exp(-dr / 431.0e3);
    const double n_td = 8.86e9 * std::exp(-dr / 100.0e3) * gauss2(angle_to(0.0, day), 15.0 * deg);
    const double n_psd = 4.06e10 * std::exp(-dr / 232.0e3) *
                         (gauss2(angle_to(50.0 * deg, day), 20.0 * deg) + gauss2(angle_to(-50.0 * deg, day), 20.0 * deg));
    const double sp = 5.67e6 * std::exp(-dr / 748.0e3);
    const double n_sp_day = sp * (gauss2(angle_to(80.0 * deg, day), 15.0 * deg) + gauss2(angle_to(-80.0 * deg, day), 15.0 * deg));
    const double n_sp_night =
        sp * (gauss2(angle_to(15.0 * deg, night), 10.0 * deg) + gauss2(angle_to(-15.0 * deg, night), 10.0 * deg));
    double n_m3 = (1.0 / (r_nd * r_nd)) * (n_miv + n_td + n_psd + n_sp_day + n_sp_night);
    if (!std::isfinite(n_m3) || n_m3 < 0.0)
        n_m3 = 0.0;
    return n_m3 * 1.0e-6;
}

double mercury_photo_rate(const double x[3], double L_ref)
{
    const double r_nd = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    if (!(r_nd > 1.0) || !std::isfinite(r_nd))
        return 0.0;
    const double n_cm3 = mercury_na_density(x, L_ref);
    const bool shadow = x[0] < 0.0 && std::sqrt(x[1] * x[1] + x[2] * x[2]) < 1.0;
    double q = (shadow ? 1.0e-5 : 5.0e-5) * n_cm3;
    if (!std::isfinite(q) || q < 0.0)
        q = 0.0;
    return q;
}

// Mercury-configured shell: reference values of Z5_Mercury/9999setup
// (REF U=4.031e5 m/s, T=1.7407e5 K, n=4e7 m^-3, M=1e-3 kg/mol, B=1 nT,
// L=2.44e6 m), an offset dipole added to Badd through the curl of its edge
// circulation (Z5_Mercury/include/operators/Dipole.h:89-333, vector potential
// A = m x dr / (|dr|^2 + a^2)^(3/2)), plus the IMF flux.  Two-fluid (the
// reference's physics): H+ solar wind and a Na+ seed state, the Na+
// photoionization source and the ion-neutral drag; single fluid: the shared
// Lunar step with the dipole background.  The body is the absorbing wall of
// the Lunar set-up in both (the reference's Mercury grid continues into a
// coupled resistive interior, which this shell does not have).
CaseSetup make_case_mercury(int nang, int nrad, double r_in, double r_out, int g, double gamma, double cfl,
                            const WindParams &w, const DipoleParams &dp)
{
    CaseSetup cs;
    cs.grid = make_cubed_sphere_shell(nang, nrad, r_in, r_out, g, BC_SOLID, BC_FARFIELD);
    finalize_grid(cs, dp.two_fluid ? 1 : 0);

    struct
    {
        double U = 4.031e5, T = 1.7407e5, n = 4e7, M = 0.001, B = 1e-9, L = 2.44e6;
        double NA = 6.02e23, R = 8.314, mu0 = 1.256637061e-6;
    } R;
    cs.L_ref = R.L;
    MhdCaseConst &c = cs.cst;
    const double rho_ref = R.M * R.n / R.NA;
    const double M_H = 0.001 * w.m_amu;
    c.gamma = gamma;
    c.cfl = cfl;
    c.inver_MA2 = R.B * R.B / (R.U * R.U * R.mu0 * rho_ref);
    c.momentum_hall_coeff = (R.B * R.B) / (R.mu0 * rho_ref * R.U * R.U);
    c.M_H = M_H;
    c.M_ref = R.M;
    c.ne_floor = 0.01;
    c.state_coeff = (R.R * R.T) / (M_H * R.U * R.U);
    c.dt_floor = 1e-7;
    const double k_B = R.R / R.NA;
    const double n_m3 = w.n_cc * 1e6;
    const double T_K = w.T_eV * 11604.518;
    const double rho0 = (n_m3 / R.n) * (M_H / R.M);
    const double V = w.v_kms * 1e3;
    const double u0 = w.vdir[0] * V / R.U, v0 = w.vdir[1] * V / R.U, w0 = w.vdir[2] * V / R.U;
    const double p0 = n_m3 * k_B * T_K / (rho_ref * R.U * R.U);
    c.qinf[0] = rho0;
    c.qinf[1] = rho0 * u0;
    c.qinf[2] = rho0 * v0;
    c.qinf[3] = rho0 * w0;
    c.qinf[4] = 0.5 * rho0 * (u0 * u0 + v0 * v0 + w0 * w0) + p0 / (gamma - 1.0);
    c.u_inf[0] = u0;
    c.u_inf[1] = v0;
    c.u_inf[2] = w0;
    for (int a = 0; a < 3; ++a)
        c.B_imf[a] = w.B_nT[a] * 1e-9 / R.B;
    if (dp.two_fluid)
    {
        // calc_physical_constant (Z5_Mercury/src/physics/MercurySolver_Tools.cpp:3-46),
        // source coefficients (MercurySolver_FluidRHS.cpp:210-225), seed
        // state (Z5_Mercury/include/0_BackgroundState.h:65-81, rho_small 1e-6)
        const double qe = 1.602e-19, M_Na = 0.023, m_Na = M_Na / R.NA, U3 = R.U * R.U * R.U;
        c.model = 1.0;
        c.ambipolar = dp.ambipolar ? 1.0 : 0.0;
        c.M_Na = M_Na;
        c.state_coeff_Na = (R.R * R.T) / (M_Na * R.U * R.U);
        c.induce = (qe * R.L * R.B * R.n) / (rho_ref * R.U);
        c.a1_Na = (R.L / R.U) * (m_Na / rho_ref) * 1e6;
        c.a4 = R.L / R.U;
        c.a5 = (3.0 * R.L * k_B * R.n) / (rho_ref * U3);
        c.a6 = (1.0e6 * R.L * k_B) / (rho_ref * U3 * (gamma - 1.0));
        c.T_ref = R.T;
        c.Tn0 = 1000.0;
        c.pe_coef = (rho_ref * R.U) / (qe * R.L * R.B * R.n);
        const double rho_small = 1e-6;
        const double ps = p0 * rho_small / 23.0;
        c.qinfs[0] = rho_small * rho0;
        c.qinfs[1] = c.qinfs[2] = c.qinfs[3] = 0.0;
        c.qinfs[4] = ps / (gamma - 1.0) + 0.5 * c.qinfs[0] * (0.0 * 0.0 + 0.0 * 0.0 + 0.0 * 0.0);
        c.hall = dp.hall ? 1.0 : 0.0;
        c.hall_coef = R.B * R.M / (R.U * qe * R.L * R.mu0 * rho_ref * R.NA);
        c.hall_rmin = dp.hall_rmin;
        c.hall_rmax = dp.hall_rmax;
        c.local_arti = (dp.loc_eta_max > 0.0 && dp.loc_r_decay > 0.0) ? 1.0 : 0.0;
        c.loc_eta_max = dp.loc_eta_max;
        for (int a = 0; a < 3; ++a)
            c.loc_center[a] = dp.loc_center[a];
        c.loc_r_decay = dp.loc_r_decay;
        c.loc_r_cutoff = dp.loc_r_cutoff;
        c.resist = dp.resist ? 1.0 : 0.0;
        c.resist_nsub = dp.resist_nsub;
        c.inver_Rem = dp.resist_eta_max / (R.mu0 * R.U * R.L);
    }

    build_ops(cs);

    // Na+ seed state, photoionization rate and drag frequency of every cell
    if (dp.two_fluid)
        for (size_t b = 0; b < cs.blk.size(); ++b)
        {
            BlockArrays &A = cs.blk[b];
            const Block &B = cs.grid.blocks[b];
            const MhdBlockDev &d = A.d;
            for (int q = 0; q < 5; ++q)
                for (int64_t o = 0; o < d.slot; ++o)
                    A.st[(size_t)(ST_UNA0 + q) * d.slot + o] = c.qinfs[q];
            for (int i = -g; i < B.ni + g; ++i)
                for (int j = -g; j < B.nj + g; ++j)
                    for (int k = -g; k < B.nk + g; ++k)
                    {
                        double xc[3];
                        dual_centre(B, i, j, k, xc);
                        const int64_t o = blk_off(d, i, j, k);
                        A.st[(size_t)ST_PHOTO * d.slot + o] = mercury_photo_rate(xc, R.L);
                        // drag frequency: photoionization frequency of the
                        // geometric shadow test (MercurySolver_FluidRHS.cpp:360-367)
                        const bool shadow = xc[0] < 0.0 && std::sqrt(xc[1] * xc[1] + xc[2] * xc[2]) < 1.0;
                        A.st[(size_t)ST_DRAG * d.slot + o] = shadow ? 1.0e-5 : 5.0e-5;
                    }
            // per edge: the edge vector (dr of src/3_field/4_MPCNS_Geometry.cpp:1251-1320),
            // the radial Hall taper at the midpoint (HallRadialTaperEdge_,
            // MercurySolver_MagExplicit_Hall.cpp:7-37) and the local
            // artificial resistivity (MercurySolver_MagExplicit_Resis.cpp:86-121)
            for (int e = 0; e < 3; ++e)
                for (int i = -g; i <= B.ni + g; ++i)
                    for (int j = -g; j <= B.nj + g; ++j)
                        for (int k = -g; k <= B.nk + g; ++k)
                        {
                            const int64_t o = blk_off(d, i, j, k);
                            const int64_t n0 = B.node_off(i, j, k), n1 = B.node_off(i + (e == 0), j + (e == 1), k + (e == 2));
                            const double p0[3] = {B.x[n0], B.y[n0], B.z[n0]}, p1[3] = {B.x[n1], B.y[n1], B.z[n1]};
                            double m[3];
                            for (int q = 0; q < 3; ++q)
                                m[q] = 0.5 * (p0[q] + p1[q]);
                            const double r = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
                            const double width = c.hall_rmax - c.hall_rmin;
                            const double taper =
                                width <= 0.0 ? (r >= c.hall_rmin ? 1.0 : 0.0)
                                             : 0.5 * (1.0 + std::tanh(2.0 * (r - 0.5 * (c.hall_rmin + c.hall_rmax)) / width));
                            A.st[(size_t)(ST_TAPER + e) * d.slot + o] = taper;
                            double eta = 0.0;
                            if (c.local_arti != 0.0)
                            {
                                const double dx = m[0] - c.loc_center[0], dy = m[1] - c.loc_center[1],
                                             dz = m[2] - c.loc_center[2];
                                const double rl = std::sqrt(dx * dx + dy * dy + dz * dz);
                                eta = (c.loc_r_cutoff > 0.0 && rl >= c.loc_r_cutoff)
                                          ? 0.0
                                          : c.loc_eta_max * std::exp(-rl / c.loc_r_decay);
                            }
                            A.st[(size_t)(ST_ETALOC + e) * d.slot + o] = eta;
                            // inver_Rem yita0(r) (MercurySolver_MagExplicit_Resis.cpp:196-206, 268-275)
                            double yita0 = 0.0;
                            if (!(r <= 0.84 || r >= 1.04))
                                yita0 = 0.5 * (std::tanh((r - 0.84) / 0.02) - std::tanh((r - 1.04) / 0.02));
                            A.st[(size_t)(ST_ETARES + e) * d.slot + o] =
                                c.resist != 0.0 && yita0 != 0.0 ? c.inver_Rem * yita0 : 0.0;
                        }
        }

    // nondimensional dipole: |B| = |m| / r^3 at the equator of the body
    const double m[3] = {0.0, 0.0, -dp.B_eq_nT * 1e-9 / R.B};
    const double r0[3] = {0.0, 0.0, dp.offset_z};
    const double a2 = dp.softening * dp.softening;
    auto A_of = [&](const double *x, double *A) {
        const double d[3] = {x[0] - r0[0], x[1] - r0[1], x[2] - r0[2]};
        double f = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + a2;
        f = std::sqrt(f) * f;
        A[0] = (m[1] * d[2] - m[2] * d[1]) / f;
        A[1] = (m[2] * d[0] - m[0] * d[2]) / f;
        A[2] = (m[0] * d[1] - m[1] * d[0]) / f;
    };
    for (size_t b = 0; b < cs.blk.size(); ++b)
    {
        BlockArrays &A = cs.blk[b];
        const Block &B = cs.grid.blocks[b];
        const MhdBlockDev &d = A.d;
        for (int q = 0; q < 5; ++q)
            for (int64_t o = 0; o < d.slot; ++o)
                A.st[(size_t)(ST_U0 + q) * d.slot + o] = c.qinf[q];
        // IMF flux everywhere (the outermost ghost faces keep only this part)
        for (int a = 0; a < 3; ++a)
            for (int64_t o = 0; o < d.slot; ++o)
            {
                const double sx = A.geo[(size_t)(G_JD + 3 * a + 0) * d.slot + o];
                const double sy = A.geo[(size_t)(G_JD + 3 * a + 1) * d.slot + o];
                const double sz = A.geo[(size_t)(G_JD + 3 * a + 2) * d.slot + o];
                A.st[(size_t)(ST_BADD + a) * d.slot + o] = c.B_imf[0] * sx + c.B_imf[1] * sy + c.B_imf[2] * sz;
            }
        // edge circulation of the dipole potential (midpoint rule), boxes [-g, n+g]
        std::vector<double> Ed[3];
        for (int e = 0; e < 3; ++e)
            Ed[e].assign((size_t)d.slot, 0.0);
        for (int i = -g; i <= B.ni + g; ++i)
            for (int j = -g; j <= B.nj + g; ++j)
                for (int k = -g; k <= B.nk + g; ++k)
                    for (int e = 0; e < 3; ++e)
                    {
                        const int64_t o0 = B.node_off(i, j, k);
                        const int64_t o1 = B.node_off(i + (e == 0), j + (e == 1), k + (e == 2));
                        const double x1[3] = {B.x[o0], B.y[o0], B.z[o0]}, x2[3] = {B.x[o1], B.y[o1], B.z[o1]};
                        const double xm[3] = {0.5 * (x1[0] + x2[0]), 0.5 * (x1[1] + x2[1]), 0.5 * (x1[2] + x2[2])};
                        double Av[3];
                        A_of(xm, Av);
                        Ed[e][blk_off(d, i, j, k)] =
                            Av[0] * (x2[0] - x1[0]) + Av[1] * (x2[1] - x1[1]) + Av[2] * (x2[2] - x1[2]);
                    }
        const int64_t di = (int64_t)d.sj * d.sk, dj = d.sk, dk = 1;
        for (int i = -g; i < B.ni + g; ++i)
            for (int j = -g; j < B.nj + g; ++j)
                for (int k = -g; k < B.nk + g; ++k)
                {
                    const int64_t o = blk_off(d, i, j, k);
                    // curl of the edge 1-form (CurlEdgeToFace with multiper +1)
                    const double fx = (Ed[1][o] - Ed[1][o + dk]) + (Ed[2][o + dj] - Ed[2][o]);
                    const double fy = (Ed[0][o + dk] - Ed[0][o]) + (Ed[2][o] - Ed[2][o + di]);
                    const double fz = (Ed[0][o] - Ed[0][o + dj]) + (Ed[1][o + di] - Ed[1][o]);
                    const double f[3] = {fx, fy, fz};
                    for (int a = 0; a < 3; ++a)
                    {
                        const double sx = A.geo[(size_t)(G_JD + 3 * a + 0) * d.slot + o];
                        const double sy = A.geo[(size_t)(G_JD + 3 * a + 1) * d.slot + o];
                        const double sz = A.geo[(size_t)(G_JD + 3 * a + 2) * d.slot + o];
                        A.st[(size_t)(ST_BADD + a) * d.slot + o] =
                            f[a] + (c.B_imf[0] * sx + c.B_imf[1] * sy + c.B_imf[2] * sz);
                    }
                }
    }
    apply_ops_host(cs.blk, face_copy_sync(cs, ST_BADD), c);
    return cs;
}

} // namespace mpcns
