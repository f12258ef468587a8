// A case assembled from the caller's own grid, topology and geometry: the
// path of a reference driver that already holds Grid, TOPO::Topology and
// Field objects (Z4_Lunar/main.cpp:46-123) and hands them over block by
// block instead of letting the library generate the grid.
//
// The caller gives each block's interior node coordinates and six faces in
// the reference's IndexTransform convention; the library fills the ghost
// nodes from them (mirror at physical faces, copy across interfaces, as
// Grid_Preprocess does), builds the geometry and weight tables, and checks
// any geometry array the caller also gives (its own Field's Jac, JD*, beta*,
// alpha*, Bcell_from_Bface_w, Jcell_from_Jedge_w) against what it built.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/beta1_mhd.h"
#include "host_case.h"
#include "setup.h"

using namespace mpcns;

namespace
{
constexpr int kGeoFields = 12; // MHD_F_JAC .. MHD_F_JCELL_W

struct DraftBlock
{
    bool have = false, topo = false;
    int ni = 0, nj = 0, nk = 0;
    std::vector<double> x, y, z;             // interior nodes, [i][j][k], k fastest
    FaceLink face[6];
    bool split[6] = {false, false, false, false, false, false}; // kind 2: patches to come (mhd_host_set_face_parts)
    std::vector<std::vector<double>> given;  // kGeoFields AoS arrays (empty: not given)
};
} // namespace

struct mhd_case_draft
{
    int ghost = 2;
    MhdCaseConst cst{};
    std::vector<DraftBlock> blk;
};

namespace
{
int fail(mhd_host_case *h, int code, const std::string &msg)
{
    if (h)
        h->err = msg;
    return code;
}

bool perm_ok(const int *p)
{
    bool seen[3] = {false, false, false};
    for (int a = 0; a < 3; ++a)
    {
        if (p[a] < 0 || p[a] > 2 || seen[p[a]])
            return false;
        seen[p[a]] = true;
    }
    return true;
}

// Patches of a split face: each a non-empty cell range [lo, hi) of the
// face's two tangential axes (n1 cells along the first, n2 along the
// second), together covering every cell face once, all periodic or all not.
// Empty string when they do.
std::string face_parts_problem(const std::vector<FacePart> &parts, int n1, int n2)
{
    std::vector<int> cover((size_t)n1 * n2, 0);
    for (const FacePart &P : parts)
    {
        if (P.lo[0] < 0 || P.hi[0] > n1 || P.lo[1] < 0 || P.hi[1] > n2 || P.lo[0] >= P.hi[0] || P.lo[1] >= P.hi[1])
            return "patch range outside the face";
        if (P.periodic != parts[0].periodic)
            return "periodic and plain patches on one face";
        for (int u = P.lo[0]; u < P.hi[0]; ++u)
            for (int v = P.lo[1]; v < P.hi[1]; ++v)
                cover[(size_t)u * n2 + v] += 1;
    }
    for (int c : cover)
        if (c != 1)
            return "patches do not cover the face exactly once";
    return "";
}

// The links of an interface face with the cell range each covers: the
// parts of a split face, else the whole face as one.
std::vector<FacePart> face_links(const DraftBlock &A, int f)
{
    const FaceLink &L = A.face[f];
    if (!L.parts.empty())
        return L.parts;
    const int n[3] = {A.ni, A.nj, A.nk}, ax = f / 2;
    FacePart P;
    P.lo[0] = P.lo[1] = 0;
    P.hi[0] = n[ax == 0 ? 1 : 0];
    P.hi[1] = n[ax == 2 ? 1 : 2];
    P.nb_block = L.nb_block, P.nb_face = L.nb_face, P.periodic = L.periodic;
    P.trans = L.trans;
    return {P};
}

// Node (i, j, k) of face f's node plane, the transform image in the
// neighbour, both as coordinates; false when the image is outside it.
bool face_node_image(const std::vector<DraftBlock> &B, int b, int nb, const IndexTransform &T, const int loc[3],
                     double p[3], double q[3])
{
    const DraftBlock &A = B[b];
    const DraftBlock &N = B[nb];
    int t[3];
    for (int a = 0; a < 3; ++a)
        t[T.perm[a]] = T.sign[a] * loc[a] + T.offset[a];
    const int nn[3] = {N.ni, N.nj, N.nk};
    for (int a = 0; a < 3; ++a)
        if (t[a] < 0 || t[a] > nn[a])
            return false;
    auto at = [](const DraftBlock &D, int i, int j, int k, double out[3]) {
        const size_t o = ((size_t)i * (D.nj + 1) + j) * (D.nk + 1) + k;
        out[0] = D.x[o];
        out[1] = D.y[o];
        out[2] = D.z[o];
    };
    at(A, loc[0], loc[1], loc[2], p);
    at(N, t[0], t[1], t[2], q);
    return true;
}

// Field numbers of the optional geometry arrays and their (location, comps).
void geo_field(int q, int &field, int &loc, int &nc)
{
    static const int kF[kGeoFields] = {MHD_F_JAC,     MHD_F_JDXI,     MHD_F_JDET,      MHD_F_JDZE,
                                       MHD_F_BETA_XI, MHD_F_BETA_ETA, MHD_F_BETA_ZETA, MHD_F_ALPHA_XI,
                                       MHD_F_ALPHA_ETA, MHD_F_ALPHA_ZETA, MHD_F_BCELL_W, MHD_F_JCELL_W};
    field = kF[q];
    int region, slot;
    field_slot(field, 0, region, slot, loc, nc);
}

MhdBlockDev dims_of(int ni, int nj, int nk, int g)
{
    MhdBlockDev d{};
    d.ni = ni;
    d.nj = nj;
    d.nk = nk;
    d.g = g;
    d.sj = nj + 1 + 2 * g;
    d.sk = nk + 1 + 2 * g;
    return d;
}

// Interior DOF range of a location: [0, n + 1 - delta) per axis.
bool compare_field(const BlockArrays &A, int q, const std::vector<double> &given, double rtol, std::string &why)
{
    int field, loc, nc;
    geo_field(q, field, loc, nc);
    int region, slot;
    field_slot(field, 0, region, slot, loc, nc);
    std::vector<double> mine(aos_count(A.d, loc, nc));
    soa_aos(A.d, loc, nc, const_cast<double *>(A.geo.data()) + (size_t)slot * A.d.slot, mine.data(), true);
    const int N[3] = {A.d.ni, A.d.nj, A.d.nk}, g = A.d.g;
    int ext[3], cnt[3];
    for (int a = 0; a < 3; ++a)
    {
        cnt[a] = loc_count(loc, a, N[a]);
        ext[a] = cnt[a] + 2 * g;
    }
    double scale = 0.0;
    for (int i = 0; i < cnt[0]; ++i)
        for (int j = 0; j < cnt[1]; ++j)
            for (int k = 0; k < cnt[2]; ++k)
                for (int m = 0; m < nc; ++m)
                {
                    const size_t o = ((((size_t)(i + g)) * ext[1] + (j + g)) * ext[2] + (k + g)) * nc + m;
                    scale = std::max(scale, std::fabs(mine[o]));
                }
    for (int i = 0; i < cnt[0]; ++i)
        for (int j = 0; j < cnt[1]; ++j)
            for (int k = 0; k < cnt[2]; ++k)
                for (int m = 0; m < nc; ++m)
                {
                    const size_t o = ((((size_t)(i + g)) * ext[1] + (j + g)) * ext[2] + (k + g)) * nc + m;
                    if (!(std::fabs(given[o] - mine[o]) <= rtol * scale))
                    {
                        static const char *kName[10] = {"Jac",      "JDxi",      "JDet",       "JDze",
                                                        "beta_xi",  "beta_eta",  "beta_zeta",  "alpha_xi",
                                                        "alpha_eta", "alpha_zeta"};
                        why = std::string(kName[q]) + " component " + std::to_string(m) + " at (" +
                              std::to_string(i) + ", " + std::to_string(j) + ", " + std::to_string(k) +
                              "): given " + std::to_string(given[o]) + ", the nodes give " + std::to_string(mine[o]);
                        return false;
                    }
                }
    return true;
}
} // namespace

extern "C" int mhd_host_case_begin(int nblocks, int ghost, const double *k, mhd_host_case **out)
{
    if (!k || !out || nblocks < 1 || nblocks > 100000 || ghost < 1)
        return MHD_ERR_ARG;
    MhdCaseConst c{};
    c.gamma = k[0];
    c.cfl = k[1];
    c.inver_MA2 = k[2];
    c.momentum_hall_coeff = k[3];
    c.M_H = k[4];
    c.M_ref = k[5];
    c.ne_floor = k[6];
    c.state_coeff = k[7];
    for (int m = 0; m < 5; ++m)
        c.qinf[m] = k[8 + m];
    for (int a = 0; a < 3; ++a)
    {
        c.B_imf[a] = k[13 + a];
        c.u_inf[a] = k[16 + a];
    }
    c.dt_floor = k[19];
    if (!(c.gamma > 1.0) || !(c.cfl > 0.0) || !(c.M_H > 0.0) || !(c.M_ref > 0.0) || !(c.qinf[0] > 0.0) ||
        !(c.dt_floor >= 0.0) || !(c.ne_floor >= 0.0))
        return MHD_ERR_ARG;
    auto h = new mhd_host_case();
    h->draft = std::make_shared<mhd_case_draft>();
    h->draft->ghost = ghost;
    h->draft->cst = c;
    h->draft->blk.resize(nblocks);
    *out = h;
    return MHD_OK;
}

extern "C" int mhd_host_add_block(mhd_host_case *h, int ib, int ni, int nj, int nk, const double *x, const double *y,
                                  const double *z, size_t nnode, const double *const *geometry)
{
    if (!h || !h->draft)
        return fail(h, MHD_ERR_STATE, "mhd_host_add_block: not a case begun with mhd_host_case_begin");
    mhd_case_draft &D = *h->draft;
    if (ib < 0 || ib >= (int)D.blk.size() || ni < 1 || nj < 1 || nk < 1 || !x || !y || !z)
        return fail(h, MHD_ERR_ARG, "mhd_host_add_block: bad block index, extent or node pointer");
    if (nnode != (size_t)(ni + 1) * (nj + 1) * (nk + 1))
        return fail(h, MHD_ERR_ARG, "mhd_host_add_block: expected the (ni+1)(nj+1)(nk+1) interior nodes");
    DraftBlock &B = D.blk[ib];
    B.ni = ni;
    B.nj = nj;
    B.nk = nk;
    B.x.assign(x, x + nnode);
    B.y.assign(y, y + nnode);
    B.z.assign(z, z + nnode);
    B.given.assign(kGeoFields, std::vector<double>());
    const MhdBlockDev d = dims_of(ni, nj, nk, D.ghost);
    for (int q = 0; geometry && q < kGeoFields; ++q)
        if (geometry[q])
        {
            int field, loc, nc;
            geo_field(q, field, loc, nc);
            const size_t n = aos_count(d, loc, nc);
            B.given[q].assign(geometry[q], geometry[q] + n);
        }
    B.have = true;
    return MHD_OK;
}

extern "C" int mhd_host_set_topology(mhd_host_case *h, int ib, const int *faces)
{
    if (!h || !h->draft)
        return fail(h, MHD_ERR_STATE, "mhd_host_set_topology: not a case begun with mhd_host_case_begin");
    mhd_case_draft &D = *h->draft;
    if (ib < 0 || ib >= (int)D.blk.size() || !faces)
        return fail(h, MHD_ERR_ARG, "mhd_host_set_topology: bad block index");
    DraftBlock &B = D.blk[ib];
    for (int f = 0; f < 6; ++f)
    {
        const int *o = faces + 14 * f;
        FaceLink L;
        L.kind = o[0];
        L.bc = o[1];
        L.nb_block = o[2];
        L.nb_face = o[3];
        L.periodic = o[4];
        for (int a = 0; a < 3; ++a)
        {
            L.trans.perm[a] = o[5 + a];
            L.trans.sign[a] = o[8 + a];
            L.trans.offset[a] = o[11 + a];
        }
        const std::string where = "block " + std::to_string(ib) + " face " + std::to_string(f) + ": ";
        if (L.kind == 0)
        {
            if (L.bc != BC_FARFIELD && L.bc != BC_SOLID)
                return fail(h, MHD_ERR_ARG, where + "physical faces take bc 1 (Farfield) or 2 (Solid_Surface)");
        }
        else if (L.kind == 1)
        {
            if (L.nb_block < 0 || L.nb_block >= (int)D.blk.size() || L.nb_face < 0 || L.nb_face > 5 ||
                !perm_ok(L.trans.perm))
                return fail(h, MHD_ERR_ARG, where + "bad neighbour or permutation");
            for (int a = 0; a < 3; ++a)
                if (L.trans.sign[a] != 1 && L.trans.sign[a] != -1)
                    return fail(h, MHD_ERR_ARG, where + "transform signs must be +1 or -1");
            L.bc = BC_NONE;
        }
        else if (L.kind == 2) // several interface patches, given by mhd_host_set_face_parts
        {
            L = FaceLink{};
            L.kind = 1;
        }
        else
            return fail(h, MHD_ERR_ARG, where + "kind must be 0 (physical), 1 (interface) or 2 (several interface "
                                                "patches)");
        B.face[f] = L;
        B.split[f] = o[0] == 2;
    }
    B.topo = true;
    return MHD_OK;
}

extern "C" int mhd_host_set_face_parts(mhd_host_case *h, int ib, int face, int nparts, const int *parts)
{
    if (!h || !h->draft)
        return fail(h, MHD_ERR_STATE, "mhd_host_set_face_parts: not a case begun with mhd_host_case_begin");
    mhd_case_draft &D = *h->draft;
    if (ib < 0 || ib >= (int)D.blk.size() || face < 0 || face > 5)
        return fail(h, MHD_ERR_ARG, "mhd_host_set_face_parts: bad block or face index");
    DraftBlock &B = D.blk[ib];
    const std::string where = "block " + std::to_string(ib) + " face " + std::to_string(face) + ": ";
    if (!B.topo || !B.split[face])
        return fail(h, MHD_ERR_STATE, where + "declare the face kind 2 with mhd_host_set_topology first");
    if (nparts < 2 || !parts)
        return fail(h, MHD_ERR_ARG, where + "a split face takes two or more patches");
    FaceLink L;
    L.kind = 1;
    for (int q = 0; q < nparts; ++q)
    {
        const int *o = parts + 16 * q;
        FacePart P;
        P.lo[0] = o[0], P.lo[1] = o[1], P.hi[0] = o[2], P.hi[1] = o[3];
        P.nb_block = o[4], P.nb_face = o[5], P.periodic = o[6];
        for (int a = 0; a < 3; ++a)
        {
            P.trans.perm[a] = o[7 + a];
            P.trans.sign[a] = o[10 + a];
            P.trans.offset[a] = o[13 + a];
        }
        if (P.nb_block < 0 || P.nb_block >= (int)D.blk.size() || P.nb_face < 0 || P.nb_face > 5 ||
            !perm_ok(P.trans.perm))
            return fail(h, MHD_ERR_ARG, where + "patch " + std::to_string(q) + ": bad neighbour or permutation");
        for (int a = 0; a < 3; ++a)
            if (P.trans.sign[a] != 1 && P.trans.sign[a] != -1)
                return fail(h, MHD_ERR_ARG, where + "transform signs must be +1 or -1");
        L.parts.push_back(P);
    }
    const int ax = face / 2, n[3] = {B.ni, B.nj, B.nk};
    const std::string bad = face_parts_problem(L.parts, n[ax == 0 ? 1 : 0], n[ax == 2 ? 1 : 2]);
    if (!bad.empty())
        return fail(h, MHD_ERR_ARG, where + bad);
    L.nb_block = L.parts[0].nb_block;
    L.nb_face = L.parts[0].nb_face;
    L.periodic = L.parts[0].periodic;
    L.trans = L.parts[0].trans;
    B.face[face] = L;
    B.split[face] = false;
    return MHD_OK;
}

extern "C" int mhd_host_case_end(mhd_host_case *h)
{
    if (!h || !h->draft)
        return fail(h, MHD_ERR_STATE, "mhd_host_case_end: not a case begun with mhd_host_case_begin");
    const mhd_case_draft &D = *h->draft;
    const int nb = (int)D.blk.size();
    for (int b = 0; b < nb; ++b)
    {
        if (!D.blk[b].have || !D.blk[b].topo)
            return fail(h, MHD_ERR_STATE, "block " + std::to_string(b) + ": nodes or topology missing");
        for (int f = 0; f < 6; ++f)
            if (D.blk[b].split[f])
                return fail(h, MHD_ERR_STATE, "block " + std::to_string(b) + " face " + std::to_string(f) +
                                                  ": declared split, patches missing (mhd_host_set_face_parts)");
    }
    // interfaces: each link (a whole face, or each patch of a split face)
    // is met by a neighbour link that maps every one of its cell faces back
    // (same periodic flag, inverse transform), and (unless periodic) maps
    // every node of its range onto a coincident node of the neighbour
    for (int b = 0; b < nb; ++b)
        for (int f = 0; f < 6; ++f)
        {
            const DraftBlock &A = D.blk[b];
            if (A.face[f].kind != 1)
                continue;
            const std::string where = "block " + std::to_string(b) + " face " + std::to_string(f) + ": ";
            const int n[3] = {A.ni, A.nj, A.nk}, ax = f / 2, t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
            double scale = 0.0;
            for (size_t q = 0; q < A.x.size(); ++q)
                scale = std::max(scale, std::max(std::fabs(A.x[q]), std::max(std::fabs(A.y[q]), std::fabs(A.z[q]))));
            for (const FacePart &P : face_links(A, f))
            {
                if (D.blk[P.nb_block].face[P.nb_face].kind != 1)
                    return fail(h, MHD_ERR_ARG, where + "the neighbour face does not link back");
                const std::vector<FacePart> back = face_links(D.blk[P.nb_block], P.nb_face);
                const int bax = P.nb_face / 2, b1 = bax == 0 ? 1 : 0, b2 = bax == 2 ? 1 : 2;
                // cell-face centres in doubled node coordinates
                int dc[3];
                dc[ax] = (f & 1) ? 2 * n[ax] : 0;
                for (int u = P.lo[0]; u < P.hi[0]; ++u)
                    for (int v = P.lo[1]; v < P.hi[1]; ++v)
                    {
                        dc[t1] = 2 * u + 1;
                        dc[t2] = 2 * v + 1;
                        int tc[3];
                        for (int a = 0; a < 3; ++a)
                            tc[P.trans.perm[a]] = P.trans.sign[a] * dc[a] + 2 * P.trans.offset[a];
                        const FacePart *R = nullptr;
                        for (const FacePart &Q : back)
                            if (2 * Q.lo[0] < tc[b1] && tc[b1] < 2 * Q.hi[0] && 2 * Q.lo[1] < tc[b2] &&
                                tc[b2] < 2 * Q.hi[1])
                                R = &Q;
                        bool ok = R && R->nb_block == b && R->nb_face == f && R->periodic == P.periodic;
                        if (ok)
                        {
                            int rc[3];
                            for (int a = 0; a < 3; ++a)
                                rc[R->trans.perm[a]] = R->trans.sign[a] * tc[a] + 2 * R->trans.offset[a];
                            ok = rc[0] == dc[0] && rc[1] == dc[1] && rc[2] == dc[2];
                        }
                        if (!ok)
                            return fail(h, MHD_ERR_ARG, where + "the neighbour face does not link back");
                    }
                if (P.periodic)
                    continue;
                int loc[3];
                loc[ax] = (f & 1) ? n[ax] : 0;
                for (loc[t1] = P.lo[0]; loc[t1] <= P.hi[0]; ++loc[t1])
                    for (loc[t2] = P.lo[1]; loc[t2] <= P.hi[1]; ++loc[t2])
                    {
                        double p[3], q[3];
                        if (!face_node_image(D.blk, b, P.nb_block, P.trans, loc, p, q))
                            return fail(h, MHD_ERR_ARG, where + "the transform maps a face node outside the neighbour");
                        for (int a = 0; a < 3; ++a)
                            if (!(std::fabs(p[a] - q[a]) <= 1e-9 * std::max(scale, 1.0)))
                                return fail(h, MHD_ERR_ARG, where + "the transform maps a face node onto a node "
                                                                    "at another position");
                    }
            }
        }
    try
    {
        Grid grd;
        grd.g = D.ghost;
        grd.blocks.resize(nb);
        for (int b = 0; b < nb; ++b)
        {
            const DraftBlock &S = D.blk[b];
            Block &B = grd.blocks[b];
            B.ni = S.ni;
            B.nj = S.nj;
            B.nk = S.nk;
            B.gn = D.ghost + 1;
            B.x.assign(B.node_count(), 0.0);
            B.y.assign(B.node_count(), 0.0);
            B.z.assign(B.node_count(), 0.0);
            for (int i = 0; i <= S.ni; ++i)
                for (int j = 0; j <= S.nj; ++j)
                    for (int k = 0; k <= S.nk; ++k)
                    {
                        const size_t src = ((size_t)i * (S.nj + 1) + j) * (S.nk + 1) + k;
                        const int64_t o = B.node_off(i, j, k);
                        B.x[o] = S.x[src];
                        B.y[o] = S.y[src];
                        B.z[o] = S.z[src];
                    }
            for (int f = 0; f < 6; ++f)
                B.face[f] = S.face[f];
        }
        CaseSetup cs = make_case_with(std::move(grd), D.cst);
        // the caller's own geometry, where given, must be what its nodes give
        for (int b = 0; b < nb; ++b)
        {
            const DraftBlock &S = D.blk[b];
            std::string why;
            for (int q = 0; q < 10; ++q)
                if (!S.given[q].empty() && !compare_field(cs.blk[b], q, S.given[q], 1e-9, why))
                    return fail(h, MHD_ERR_ARG, "block " + std::to_string(b) + ": " + why);
            const double *wb = S.given[10].empty() ? nullptr : S.given[10].data();
            const double *wj = S.given[11].empty() ? nullptr : S.given[11].data();
            if ((wb || wj) && !weight_tables_match(cs.blk[b], wb, wj, 1e-9, why))
                return fail(h, MHD_ERR_ARG, "block " + std::to_string(b) + ": " + why);
        }
        h->cs = std::move(cs);
        h->draft.reset();
        return MHD_OK;
    }
    catch (const std::exception &e)
    {
        return fail(h, MHD_ERR_ARG, e.what());
    }
}

// the reason the calling thread's last host-case constructor failed
static thread_local std::string g_host_ctor_err;
void mpcns::host_ctor_error(const std::string &msg) { g_host_ctor_err = msg; }

extern "C" const char *mhd_host_last_error(mhd_host_case *h)
{
    return h ? h->err.c_str() : (g_host_ctor_err.empty() ? "null host case" : g_host_ctor_err.c_str());
}
