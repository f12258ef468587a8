// Tecplot binary (.plt, "#!TDV112") output of one case in the "Mixed" mode
// both solvers select (Z4_Lunar/src/LunarSolver.cpp:43-64,
// Z5_Mercury/src/MercurySolver.cpp:36-65): node coordinates x, y, z, then
// the whitelisted cell fields at their own location, float32, variable-major,
// one ordered zone per block.  The layout follows
// IOModule::WriteTecplotBinFile (src/5_io/IOModule_Tecplot.cpp:262-579).
//
// Fields the step does not keep resident are formed here from the state, as
// the reference's output-time derived fields are:
//   PV_H / PV_Na  u, v, w, p, T from U (calc_PV, Z4_Lunar/src/physics/LunarSolver_Tools.cpp:58-110,
//                 Z5_Mercury/src/physics/MercurySolver_Tools.cpp:138-187)
//   Bind_cell     the induced cell field W_B . (face B without Badd)
//                 (calc_Bcell, LunarSolver_Tools.cpp:162-238)
//   Na            neutral Na density [cm^-3] (Z5_Mercury/include/2_Initial.h:264-361)
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "host_case.h"

namespace mpcns
{
namespace
{
struct Out
{
    std::vector<char> buf;
    void raw(const void *p, size_t n)
    {
        const char *c = static_cast<const char *>(p);
        buf.insert(buf.end(), c, c + n);
    }
    void i32(int32_t v) { raw(&v, 4); }
    void f32(float v) { raw(&v, 4); }
    void f64(double v) { raw(&v, 8); }
    // a Tecplot string: one int32 per character, zero-terminated
    void str(const std::string &s)
    {
        for (char c : s)
            i32((int32_t)c);
        i32(0);
    }
};

// One AoS FieldBlock downloaded through FieldIO: (i, j, k, comp), comp fastest.
struct Aos
{
    std::vector<double> v;
    int nc = 0, g = 0, c1 = 0, c2 = 0;
    void load(FieldIO &io, int b, int field)
    {
        int region, slot, loc;
        field_slot(field, 0, region, slot, loc, nc);
        const MhdBlockDev &d = io.block(b);
        const int rc = io.get(b, field, v);
        if (rc)
            throw std::runtime_error("tecplot: field read failed");
        g = d.g;
        c1 = loc_count(loc, 1, d.nj) + 2 * g;
        c2 = loc_count(loc, 2, d.nk) + 2 * g;
    }
    double operator()(int i, int j, int k, int m) const
    {
        return v[(((size_t)(i + g) * c1 + (j + g)) * c2 + (k + g)) * nc + m];
    }
};

enum VarKind
{
    V_COORD,
    V_PV_H,
    V_PV_NA,
    V_BCELL,
    V_BIND,
    V_NA,
    V_JCELL
};
struct Var
{
    std::string name;
    VarKind kind;
    int comp;
};

std::vector<Var> tecplot_vars(int model)
{
    std::vector<Var> v = {{"x", V_COORD, 0}, {"y", V_COORD, 1}, {"z", V_COORD, 2}};
    auto add = [&](VarKind k, std::initializer_list<const char *> names) {
        int c = 0;
        for (const char *n : names)
            v.push_back({n, k, c++});
    };
    add(V_PV_H, {"u_H", "v_H", "w_H", "p_H", "T_H"});
    if (model == 1)
        add(V_PV_NA, {"u_Na", "v_Na", "w_Na", "p_Na", "T_Na"});
    add(V_BCELL, {"Bx", "By", "Bz"});
    add(V_BIND, {"bx_ind", "by_ind", "bz_ind"});
    if (model == 1)
        add(V_NA, {"n_Na"});
    add(V_JCELL, {"Jx", "Jy", "Jz"});
    return v;
}

// u, v, w, p, T of one cell with the floors of calc_PV.
void primitive(const Aos &U, int i, int j, int k, double gamma, double coeff, double out[5])
{
    const double rho = std::max(U(i, j, k, 0), 1e-12);
    const double u = U(i, j, k, 1) / rho, v = U(i, j, k, 2) / rho, w = U(i, j, k, 3) / rho;
    const double ke = 0.5 * rho * (u * u + v * v + w * w);
    double eint = U(i, j, k, 4) - ke;
    if (eint < 0.0)
        eint = 0.0;
    const double p = std::max((gamma - 1.0) * eint, 1e-12);
    out[0] = u;
    out[1] = v;
    out[2] = w;
    out[3] = p;
    out[4] = p / (rho * coeff);
}
} // namespace

int tecplot_write(FieldIO &io, const Grid &grd, const MhdCaseConst &cst, double L_ref, const char *path,
                  double time)
{
    const int model = io.block(0).model;
    const std::vector<Var> vars = tecplot_vars(model);
    const int nvar = (int)vars.size();
    const int nb = io.nblocks();
    if ((int)grd.blocks.size() != nb)
        throw std::runtime_error("tecplot: grid and state block counts differ");

    std::unique_ptr<FILE, int (*)(FILE *)> fp(std::fopen(path, "wb"), &std::fclose);
    if (!fp)
        throw std::runtime_error(std::string("tecplot: cannot open ") + path);
    auto flush = [&](Out &o) {
        if (!o.buf.empty() && std::fwrite(o.buf.data(), 1, o.buf.size(), fp.get()) != o.buf.size())
            throw std::runtime_error(std::string("tecplot: write failed on ") + path);
        o.buf.clear();
    };

    Out o;
    o.raw("#!TDV112", 8);
    o.i32(1); // byte order
    o.i32(0); // full file
    o.str("flow_field");
    o.i32(nvar);
    for (const Var &v : vars)
        o.str(v.name);
    for (int b = 0; b < nb; ++b)
    {
        const Block &B = grd.blocks[b];
        o.f32(299.0f);
        o.str("Zone" + std::to_string(b) + "_Fluid");
        o.i32(-1); // parent zone
        o.i32(-2); // strand
        o.f64(time);
        o.i32(-1); // zone colour
        o.i32(0);  // ordered
        o.i32(1);  // variable locations follow
        for (const Var &v : vars)
            o.i32(v.kind == V_COORD ? 0 : 1);
        o.i32(0); // raw face neighbours
        o.i32(0); // miscellaneous neighbours
        o.i32(B.ni + 1);
        o.i32(B.nj + 1);
        o.i32(B.nk + 1);
        o.i32(0); // auxiliary data
    }
    o.f32(357.0f);
    flush(o);

    for (int b = 0; b < nb; ++b)
    {
        const Block &B = grd.blocks[b];
        const int NI = B.ni + 1, NJ = B.nj + 1, NK = B.nk + 1;
        const size_t nnode = (size_t)NI * NJ * NK;
        // cell variables are written over IMax * JMax * (KMax - 1), zero outside the cells
        const size_t ncell = (size_t)NI * NJ * (NK - 1);

        Aos UH, UN, BC, JC, BF[3], W;
        UH.load(io, b, MHD_F_U_H);
        if (model == 1)
            UN.load(io, b, MHD_F_U_NA);
        BC.load(io, b, MHD_F_B_CELL);
        JC.load(io, b, MHD_F_J_CELL);
        for (int a = 0; a < 3; ++a)
            BF[a].load(io, b, MHD_F_B_XI + a);
        W.load(io, b, MHD_F_BCELL_W);

        // every cell's derived values, evaluated once
        const size_t nc = (size_t)B.ni * B.nj * B.nk;
        std::vector<double> pvh(5 * nc), pvn(model == 1 ? 5 * nc : 0), bind(3 * nc), na(model == 1 ? nc : 0);
        auto cid = [&](int i, int j, int k) { return ((size_t)k * B.nj + j) * B.ni + i; };
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 0; k < B.nk; ++k)
            for (int j = 0; j < B.nj; ++j)
                for (int i = 0; i < B.ni; ++i)
                {
                    const size_t q = cid(i, j, k);
                    primitive(UH, i, j, k, cst.gamma, cst.state_coeff, &pvh[5 * q]);
                    if (model == 1)
                    {
                        primitive(UN, i, j, k, cst.gamma, cst.state_coeff_Na, &pvn[5 * q]);
                        double xc[3];
                        dual_centre(B, i, j, k, xc);
                        na[q] = mercury_na_density(xc, L_ref);
                    }
                    const double phi[6] = {-BF[0](i, j, k, 0), BF[0](i + 1, j, k, 0), -BF[1](i, j, k, 0),
                                           BF[1](i, j + 1, k, 0), -BF[2](i, j, k, 0), BF[2](i, j, k + 1, 0)};
                    for (int c = 0; c < 3; ++c)
                    {
                        double s = 0.0;
                        for (int n = 0; n < 6; ++n)
                            s += W(i, j, k, 6 * c + n) * phi[n];
                        bind[3 * q + c] = s;
                    }
                }

        auto value = [&](const Var &v, int i, int j, int k) -> double {
            if (v.kind == V_COORD)
            {
                const std::vector<double> &X = v.comp == 0 ? B.x : (v.comp == 1 ? B.y : B.z);
                return X[B.node_off(i, j, k)];
            }
            if (i >= B.ni || j >= B.nj || k >= B.nk)
                return 0.0;
            const size_t q = cid(i, j, k);
            switch (v.kind)
            {
            case V_PV_H:
                return pvh[5 * q + v.comp];
            case V_PV_NA:
                return pvn[5 * q + v.comp];
            case V_BCELL:
                return BC(i, j, k, v.comp);
            case V_BIND:
                return bind[3 * q + v.comp];
            case V_NA:
                return na[q];
            default:
                return JC(i, j, k, v.comp);
            }
        };

        o.f32(299.0f);
        for (int v = 0; v < nvar; ++v)
            o.i32(1); // float data
        o.i32(0);     // passive variables
        o.i32(0);     // shared variables
        o.i32(-1);    // shared connectivity
        std::vector<float> data;
        data.reserve(3 * nnode + (size_t)(nvar - 3) * ncell);
        for (const Var &v : vars)
        {
            const int nk = v.kind == V_COORD ? NK : NK - 1;
            double lo = 0.0, hi = 0.0;
            bool first = true;
            for (int k = 0; k < nk; ++k)
                for (int j = 0; j < NJ; ++j)
                    for (int i = 0; i < NI; ++i)
                    {
                        const double x = value(v, i, j, k);
                        if (first)
                            lo = hi = x, first = false;
                        else
                        {
                            if (x < lo)
                                lo = x;
                            if (x > hi)
                                hi = x;
                        }
                        data.push_back((float)x);
                    }
            o.f64(lo);
            o.f64(hi);
        }
        o.raw(data.data(), data.size() * sizeof(float));
        flush(o);
    }
    return MHD_OK;
}
} // namespace mpcns
