// Host-only C-ABI: build a case on the CPU (grid, topology, geometry,
// weights, op lists, initial state) and read it back in the reference AoS
// FieldBlock layout.  Used where no device is present (set-up checks, CPU
// baselines) and by the device context, which consumes the same builders.
#include <cstddef>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_set>
#include <vector>

#include "formats.h"
#include "host_case.h"

using namespace mpcns;

namespace mpcns
{
WindParams wind_params(const mhd_wind_desc &d)
{
    WindParams w;
    w.n_cc = d.n_cc;
    w.v_kms = d.v_kms;
    for (int a = 0; a < 3; ++a)
    {
        w.vdir[a] = d.vdir[a];
        w.B_nT[a] = d.B_nT[a];
    }
    w.T_eV = d.T_eV;
    w.m_amu = d.m_amu > 0 ? d.m_amu : 1.0;
    return w;
}

DipoleParams dipole_params(const mhd_dipole_desc &d)
{
    DipoleParams p;
    p.B_eq_nT = d.B_eq_nT;
    p.offset_z = d.offset_z;
    p.softening = d.softening;
    p.two_fluid = d.single_fluid == 0;
    p.ambipolar = d.no_ambipolar == 0;
    p.hall = d.no_hall == 0;
    p.hall_rmin = d.hall_r_min;
    p.hall_rmax = d.hall_r_max;
    p.loc_eta_max = d.local_eta_max;
    for (int a = 0; a < 3; ++a)
        p.loc_center[a] = d.local_center[a];
    p.loc_r_decay = d.local_r_decay;
    p.loc_r_cutoff = d.local_r_cutoff;
    p.resist = d.no_resistance == 0 && d.resist_eta_max > 0.0;
    p.resist_nsub = d.resist_subcycles > 1 ? d.resist_subcycles : 1;
    p.resist_eta_max = d.resist_eta_max;
    return p;
}

bool wind_desc_ok(const mhd_wind_desc *d)
{
    return d && d->nang >= 4 && d->nrad >= 4 && d->r_in > 0 && d->r_out > d->r_in;
}

// Field number -> (region, slot, location, components) in the device layout.
bool field_slot(int f, int ucur, int &region, int &slot, int &loc, int &nc)
{
    region = 0;
    switch (f)
    {
    case MHD_F_U_H:
        slot = ST_U0 + 5 * ucur, loc = LOC_CELL, nc = 5;
        return true;
    case MHD_F_B_XI:
    case MHD_F_B_ETA:
    case MHD_F_B_ZETA:
        slot = ST_B + (f - MHD_F_B_XI), loc = LOC_FACE_XI + (f - MHD_F_B_XI), nc = 1;
        return true;
    case MHD_F_BADD_XI:
    case MHD_F_BADD_ETA:
    case MHD_F_BADD_ZETA:
        slot = ST_BADD + (f - MHD_F_BADD_XI), loc = LOC_FACE_XI + (f - MHD_F_BADD_XI), nc = 1;
        return true;
    case MHD_F_B_CELL:
        slot = ST_BCELL, loc = LOC_CELL, nc = 3;
        return true;
    case MHD_F_J_CELL:
        slot = ST_JCELL, loc = LOC_CELL, nc = 3;
        return true;
    case MHD_F_E_XI:
    case MHD_F_E_ETA:
    case MHD_F_E_ZETA:
        slot = ST_E + (f - MHD_F_E_XI), loc = LOC_EDGE_XI + (f - MHD_F_E_XI), nc = 1;
        return true;
    case MHD_F_EFACE_XI:
    case MHD_F_EFACE_ETA:
    case MHD_F_EFACE_ZETA:
        slot = ST_EFACE + 3 * (f - MHD_F_EFACE_XI), loc = LOC_FACE_XI + (f - MHD_F_EFACE_XI), nc = 3;
        return true;
    // two-fluid cases only (the caller checks the block's slot count)
    case MHD_F_U_NA:
        slot = ST_UNA0 + 5 * ucur, loc = LOC_CELL, nc = 5;
        return true;
    case MHD_F_PHOTO_RATE:
        slot = ST_PHOTO, loc = LOC_CELL, nc = 1;
        return true;
    case MHD_F_DRAG_RATE:
        slot = ST_DRAG, loc = LOC_CELL, nc = 1;
        return true;
    default:
        break;
    }
    region = 1;
    switch (f)
    {
    case MHD_F_JAC:
        slot = G_JAC, loc = LOC_CELL, nc = 1;
        return true;
    case MHD_F_JDXI:
    case MHD_F_JDET:
    case MHD_F_JDZE:
        slot = G_JD + 3 * (f - MHD_F_JDXI), loc = LOC_FACE_XI + (f - MHD_F_JDXI), nc = 3;
        return true;
    case MHD_F_BETA_XI:
    case MHD_F_BETA_ETA:
    case MHD_F_BETA_ZETA:
        slot = G_BETA + (f - MHD_F_BETA_XI), loc = LOC_FACE_XI + (f - MHD_F_BETA_XI), nc = 1;
        return true;
    case MHD_F_ALPHA_XI:
    case MHD_F_ALPHA_ETA:
    case MHD_F_ALPHA_ZETA:
        slot = G_ALPHA + (f - MHD_F_ALPHA_XI), loc = LOC_EDGE_XI + (f - MHD_F_ALPHA_XI), nc = 1;
        return true;
    case MHD_F_BCELL_W: // region 3: the full table, held by host cases (device cases form it from Mi, mu, JD)
        region = 3, slot = 0, loc = LOC_CELL, nc = 18;
        return true;
    case MHD_F_JCELL_W: // region 2: the full table, held by host cases only
        region = 2, slot = 0, loc = LOC_CELL, nc = 36;
        return true;
    default:
        region = -1;
        return false;
    }
}

size_t aos_count(const MhdBlockDev &d, int loc, int nc)
{
    size_t n = (size_t)nc;
    const int N[3] = {d.ni, d.nj, d.nk};
    for (int a = 0; a < 3; ++a)
        n *= (size_t)(loc_count(loc, a, N[a]) + 2 * d.g);
    return n;
}

// SoA slots (nc consecutive slots starting at `soa`) <-> AoS FieldBlock.
void soa_aos(const MhdBlockDev &d, int loc, int nc, double *soa, double *aos, bool to_aos)
{
    const int N[3] = {d.ni, d.nj, d.nk};
    int cnt[3];
    for (int a = 0; a < 3; ++a)
        cnt[a] = loc_count(loc, a, N[a]);
    const int g = d.g;
    size_t q = 0;
    for (int i = -g; i < cnt[0] + g; ++i)
        for (int j = -g; j < cnt[1] + g; ++j)
            for (int k = -g; k < cnt[2] + g; ++k)
            {
                const int64_t o = blk_off(d, i, j, k);
                for (int m = 0; m < nc; ++m, ++q)
                {
                    if (to_aos)
                        aos[q] = soa[(size_t)m * d.slot + o];
                    else
                        soa[(size_t)m * d.slot + o] = aos[q];
                }
            }
}
} // namespace mpcns

extern "C" int mhd_host_case_ot(const mhd_ot_desc *d, int ghost, mhd_host_case **out)
{
    if (!d || !out || d->n < 4 || ghost < 1)
        return MHD_ERR_ARG;
    try
    {
        auto h = std::make_unique<mhd_host_case>();
        h->cs = make_case_orszag_tang(d->n, ghost, d->gamma, d->cfl, d->seed, d->noise);
        *out = h.release();
        return MHD_OK;
    }
    catch (const std::exception &e)
    {
        host_ctor_error(e.what());
        return MHD_ERR_ARG;
    }
}

extern "C" int mhd_host_case_lunar(const mhd_wind_desc *d, int ghost, mhd_host_case **out)
{
    if (!wind_desc_ok(d) || !out || ghost < 1)
        return MHD_ERR_ARG;
    try
    {
        auto h = std::make_unique<mhd_host_case>();
        h->cs = make_case_lunar(d->nang, d->nrad, d->r_in, d->r_out, ghost, d->gamma, d->cfl, wind_params(*d));
        *out = h.release();
        return MHD_OK;
    }
    catch (const std::exception &e)
    {
        host_ctor_error(e.what());
        return MHD_ERR_ARG;
    }
}

extern "C" int mhd_host_case_mercury(const mhd_wind_desc *d, const mhd_dipole_desc *dp, int ghost,
                                     mhd_host_case **out)
{
    if (!wind_desc_ok(d) || !dp || !out || ghost < 1 || dp->softening < 0)
        return MHD_ERR_ARG;
    try
    {
        auto h = std::make_unique<mhd_host_case>();
        h->cs = make_case_mercury(d->nang, d->nrad, d->r_in, d->r_out, ghost, d->gamma, d->cfl, wind_params(*d),
                                  dipole_params(*dp));
        *out = h.release();
        return MHD_OK;
    }
    catch (const std::exception &e)
    {
        host_ctor_error(e.what());
        return MHD_ERR_ARG;
    }
}

extern "C" int mhd_host_case_destroy(mhd_host_case *h)
{
    delete h;
    return MHD_OK;
}

extern "C" int mhd_host_case_info(mhd_host_case *h, int *nblocks, int *ni, int *nj, int *nk, int *ghost)
{
    if (!h)
        return MHD_ERR_ARG;
    if (h->cs.blk.empty())
        return MHD_ERR_STATE; // a case begun with mhd_host_case_begin and not yet ended
    const MhdBlockDev &d = h->cs.blk[0].d;
    if (nblocks)
        *nblocks = (int)h->cs.blk.size();
    if (ni)
        *ni = d.ni;
    if (nj)
        *nj = d.nj;
    if (nk)
        *nk = d.nk;
    if (ghost)
        *ghost = d.g;
    return MHD_OK;
}

extern "C" int mhd_host_block_info(mhd_host_case *h, int block, int *ni, int *nj, int *nk)
{
    if (!h || !ni || !nj || !nk || block < 0 || block >= (int)h->cs.blk.size())
        return MHD_ERR_ARG;
    const MhdBlockDev &d = h->cs.blk[block].d;
    *ni = d.ni;
    *nj = d.nj;
    *nk = d.nk;
    return MHD_OK;
}

extern "C" int mhd_host_field_count(mhd_host_case *h, int field, size_t *n)
{
    int region, slot, loc, nc;
    if (!h || !n || h->cs.blk.empty() || !field_slot(field, 0, region, slot, loc, nc))
        return MHD_ERR_ARG;
    if (region == 0 && slot + nc > st_nslot(h->cs.blk[0].d.model))
        return MHD_ERR_ARG;
    *n = aos_count(h->cs.blk[0].d, loc, nc);
    return MHD_OK;
}

extern "C" int mhd_host_download(mhd_host_case *h, int block, int field, double *aos, size_t n)
{
    int region, slot, loc, nc;
    if (!h || !aos || block < 0 || block >= (int)h->cs.blk.size() || !field_slot(field, 0, region, slot, loc, nc))
        return MHD_ERR_ARG;
    BlockArrays &A = h->cs.blk[block];
    if (n != aos_count(A.d, loc, nc) || (region == 0 && slot + nc > st_nslot(A.d.model)))
        return MHD_ERR_ARG;
    std::vector<double> &v = region == 0 ? A.st : (region == 1 ? A.geo : (region == 2 ? A.wj : A.wb));
    soa_aos(A.d, loc, nc, v.data() + (size_t)slot * A.d.slot, aos, true);
    return MHD_OK;
}

extern "C" int mhd_host_upload(mhd_host_case *h, int block, int field, const double *aos, size_t n)
{
    int region, slot, loc, nc;
    if (!h || !aos || block < 0 || block >= (int)h->cs.blk.size() || !field_slot(field, 0, region, slot, loc, nc))
        return MHD_ERR_ARG;
    BlockArrays &A = h->cs.blk[block];
    if (n != aos_count(A.d, loc, nc) || (region == 0 && slot + nc > st_nslot(A.d.model)))
        return MHD_ERR_ARG;
    std::vector<double> &v = region == 0 ? A.st : (region == 1 ? A.geo : (region == 2 ? A.wj : A.wb));
    soa_aos(A.d, loc, nc, v.data() + (size_t)slot * A.d.slot, const_cast<double *>(aos), false);
    if (region != 0)
        h->cs.shape_key.clear(); // edited geometry: no longer the same as any resident copy
    if (region >= 2)
        h->cs.tables_edited = true;
    return MHD_OK;
}

extern "C" int mhd_host_faces(mhd_host_case *h, int block, int *out)
{
    if (!h || !out || block < 0 || block >= (int)h->cs.grid.blocks.size())
        return MHD_ERR_ARG;
    const Block &B = h->cs.grid.blocks[block];
    for (int f = 0; f < 6; ++f)
    {
        const FaceLink &L = B.face[f];
        int *o = out + 14 * f;
        o[0] = L.parts.empty() ? L.kind : 2; // 2: several interface patches, the first described
        o[1] = L.bc;
        o[2] = L.nb_block;
        o[3] = L.nb_face;
        o[4] = L.periodic;
        for (int a = 0; a < 3; ++a)
        {
            o[5 + a] = L.trans.perm[a];
            o[8 + a] = L.trans.sign[a];
            o[11 + a] = L.trans.offset[a];
        }
    }
    return MHD_OK;
}

extern "C" int mhd_host_face_parts(mhd_host_case *h, int block, int face, int *out, int cap, int *nparts)
{
    if (!h || !nparts || block < 0 || block >= (int)h->cs.grid.blocks.size() || face < 0 || face > 5)
        return MHD_ERR_ARG;
    const FaceLink &L = h->cs.grid.blocks[block].face[face];
    *nparts = (int)L.parts.size();
    if (*nparts > cap || (*nparts > 0 && !out))
        return MHD_ERR_ARG;
    for (int q = 0; q < *nparts; ++q)
    {
        const FacePart &P = L.parts[q];
        int *o = out + 16 * q;
        o[0] = P.lo[0], o[1] = P.lo[1], o[2] = P.hi[0], o[3] = P.hi[1];
        o[4] = P.nb_block, o[5] = P.nb_face, o[6] = P.periodic;
        for (int a = 0; a < 3; ++a)
        {
            o[7 + a] = P.trans.perm[a];
            o[10 + a] = P.trans.sign[a];
            o[13 + a] = P.trans.offset[a];
        }
    }
    return MHD_OK;
}

extern "C" int mhd_host_nodes(mhd_host_case *h, int block, int comp, double *out, size_t n)
{
    if (!h || !out || block < 0 || block >= (int)h->cs.grid.blocks.size() || comp < 0 || comp > 2)
        return MHD_ERR_ARG;
    const Block &B = h->cs.grid.blocks[block];
    const std::vector<double> &v = comp == 0 ? B.x : (comp == 1 ? B.y : B.z);
    if (n != v.size())
        return MHD_ERR_ARG;
    std::memcpy(out, v.data(), sizeof(double) * n);
    return MHD_OK;
}

// Constants of the two-fluid step, in MhdCaseConst order from `model`:
// model, ambipolar, M_Na, state_coeff_Na, qinfs[5], induce, a1_Na, a4, a5, a6,
// T_ref, Tn0, pe_coef, hall, hall_coef, hall_rmin, hall_rmax, local_arti,
// loc_eta_max, loc_center[3], loc_r_decay, loc_r_cutoff, resist, resist_nsub,
// inver_Rem (31 doubles).
extern "C" int mhd_host_constants_two_fluid(mhd_host_case *h, double *o)
{
    if (!h || !o)
        return MHD_ERR_ARG;
    const MhdCaseConst &k = h->cs.cst;
    static_assert(offsetof(MhdCaseConst, inver_Rem) - offsetof(MhdCaseConst, model) == 30 * sizeof(double),
                  "two-fluid constants are 31 consecutive doubles");
    std::memcpy(o, &k.model, 31 * sizeof(double));
    return MHD_OK;
}

extern "C" int mhd_host_constants(mhd_host_case *h, double *o)
{
    if (!h || !o)
        return MHD_ERR_ARG;
    const MhdCaseConst &k = h->cs.cst;
    const double v[22] = {k.gamma,    k.cfl,      k.inver_MA2, k.momentum_hall_coeff, k.M_H,     k.M_ref,
                          k.ne_floor, k.state_coeff, k.qinf[0], k.qinf[1],            k.qinf[2], k.qinf[3],
                          k.qinf[4],  k.B_imf[0], k.B_imf[1],  k.B_imf[2],            k.u_inf[0], k.u_inf[1],
                          k.u_inf[2], k.dt_floor, 0.0,         0.0};
    std::memcpy(o, v, sizeof(v));
    return MHD_OK;
}

// Apply one of the case's halo/boundary op lists to its host arrays
// (copy / negate / zero / constant ops only).
extern "C" int mhd_host_apply_sync(mhd_host_case *h, int sync)
{
    if (!h || sync < 0 || sync >= SY_COUNT)
        return MHD_ERR_ARG;
    try
    {
        apply_ops_host(h->cs.blk, h->cs.ops[sync], h->cs.cst);
    }
    catch (const std::exception &)
    {
        return MHD_ERR_ARG;
    }
    return MHD_OK;
}

extern "C" int mhd_host_op_count(mhd_host_case *h, int sync, size_t *n)
{
    if (!h || !n || sync < 0 || sync >= SY_COUNT)
        return MHD_ERR_ARG;
    *n = h->cs.ops[sync].size();
    return MHD_OK;
}

// The condition under which the device applies a list's records (and the
// elements of a run) in any order, loading every source before its stores:
// no element is written twice and no written element is read by the list.
extern "C" int mhd_host_ops_independent(mhd_host_case *h, int sync, int *independent)
{
    if (!h || !independent || sync < 0 || sync >= SY_COUNT)
        return MHD_ERR_ARG;
    const std::vector<MhdOp> &ops = h->cs.ops[sync];
    std::unordered_set<int64_t> dst;
    bool ok = true;
    for (const MhdOp &o : ops)
    {
        const int n = o.run > 1 ? o.run : 1;
        const int nc = o.kind == OP_CONST ? 1 : o.kind == OP_NA_FLOOR || o.kind == OP_SOLID_FIRST ||
                                                         o.kind == OP_SOLID_GHOST ? 5 : o.ncomp;
        for (int e = 0; e < n && ok; ++e)
            for (int c = 0; c < nc; ++c)
                if (!dst.insert(o.dst + (int64_t)e * o.dstep + (int64_t)c * o.stride).second)
                    ok = false;
    }
    for (const MhdOp &o : ops)
    {
        if (!ok)
            break;
        if (o.kind != OP_COPY && o.kind != OP_NEG && o.kind != OP_SOLID_FIRST && o.kind != OP_SOLID_GHOST)
            continue;
        const int n = o.run > 1 ? o.run : 1;
        const int nc = o.kind == OP_COPY || o.kind == OP_NEG ? o.ncomp : 5;
        for (int e = 0; e < n && ok; ++e)
            for (int c = 0; c < nc; ++c)
                if (dst.count(o.src + (int64_t)e * o.sstep + (int64_t)c * o.stride))
                    ok = false;
    }
    *independent = ok ? 1 : 0;
    return MHD_OK;
}

namespace
{
// face_source evaluated in one piece (the kernels compose it from per-axis
// parts): 0 in range and on no wall slab, 1 on a wall slab or out of range
// along one axis only, beyond a wall, 2 otherwise
int face_source_whole(const MhdBlockDev &b, int a, int i, int j, int k)
{
    const int p[3] = {i, j, k}, n[3] = {b.ni, b.nj, b.nk};
    int nout = 0, oax = 0;
    for (int t = 0; t < 3; ++t)
        if (p[t] < 0 || p[t] >= n[t] + (t == a ? 1 : 0))
        {
            ++nout;
            oax = t;
        }
    if (nout == 0)
    {
        for (int t = 0; t < 3; ++t)
            if ((b.bc[2 * t] == BC_SOLID && p[t] == 0) ||
                (b.bc[2 * t + 1] == BC_SOLID && p[t] == n[t] + (t == a ? 1 : 0) - 1))
                return 1;
        return 0;
    }
    return nout == 1 && b.bc[2 * oax + (p[oax] < 0 ? 0 : 1)] == BC_SOLID ? 1 : 2;
}
} // namespace

// What the one-pass EMF of a whole step assumes of the Eface sync, checked on
// the op lists: face_source's per-axis composition is its whole evaluation;
// the shell enumeration lists each shell point once and only those; every halo copy reads a face on the shell and writes a ghost face
// that face_source sends to the Eface array; every face the wall handler
// zeroes is one face_source makes zero, and every in-range face it makes zero
// is zeroed by the handler.
extern "C" int mhd_host_emf_shell_check(mhd_host_case *h, int *ok)
{
    if (!h || !ok)
        return MHD_ERR_ARG;
    const std::vector<BlockArrays> &blk = h->cs.blk;
    bool good = !blk.empty();
    for (const BlockArrays &B : blk)
    {
        const MhdBlockDev &d = B.d;
        if (d.ni < 2 || d.nj < 2 || d.nk < 2)
            continue;
        std::unordered_set<int64_t> seen;
        const int n = shell_count(d);
        for (int t = 0; t < n && good; ++t)
        {
            int i, j, k;
            shell_point(d, t, i, j, k);
            good = i >= 0 && i <= d.ni && j >= 0 && j <= d.nj && k >= 0 && k <= d.nk && in_shell(d, i, j, k) &&
                   seen.insert(blk_off(d, i, j, k)).second;
        }
        int brute = 0;
        for (int i = 0; i <= d.ni; ++i)
            for (int j = 0; j <= d.nj; ++j)
                for (int k = 0; k <= d.nk; ++k)
                    brute += in_shell(d, i, j, k) ? 1 : 0;
        good = good && brute == n;
        // the per-axis composition equals the whole evaluation on the stencil's reach
        for (int a = 0; a < 3 && good; ++a)
            for (int i = -2; i <= d.ni + 2 && good; ++i)
                for (int j = -2; j <= d.nj + 2 && good; ++j)
                    for (int k = -2; k <= d.nk + 2 && good; ++k)
                        good = face_source(d, a, i, j, k) == face_source_whole(d, a, i, j, k);
    }
    const int64_t sl = blk[0].d.slot, bs = (int64_t)st_nslot(blk[0].d.model) * sl;
    // element offset -> block, Eface family and component, point
    auto decode = [&](int64_t off, int &b, int &fam, int &comp, int p[3]) {
        b = (int)(off / bs);
        const int64_t r = off % bs, s = r / sl, q = r % sl;
        const MhdBlockDev &d = blk[b].d;
        fam = (int)(s - ST_EFACE) / 3;
        comp = (int)(s - ST_EFACE) % 3;
        p[0] = (int)(q / ((int64_t)d.sj * d.sk)) - d.g;
        p[1] = (int)((q / d.sk) % d.sj) - d.g;
        p[2] = (int)(q % d.sk) - d.g;
        return s >= ST_EFACE && s < ST_EFACE + 9;
    };
    for (const MhdOp &o : h->cs.ops[SY_EFACE_HALO])
    {
        const int n = o.run > 1 ? o.run : 1;
        for (int e = 0; e < n && good; ++e)
        {
            int b, fam, comp, p[3];
            good = decode(o.src + (int64_t)e * o.sstep, b, fam, comp, p) && in_shell(blk[b].d, p[0], p[1], p[2]) &&
                   face_source(blk[b].d, fam, p[0], p[1], p[2]) != 2;
            good = good && decode(o.dst + (int64_t)e * o.dstep, b, fam, comp, p) &&
                   face_source(blk[b].d, fam, p[0], p[1], p[2]) == 2;
        }
    }
    std::unordered_set<int64_t> zeroed;
    for (const MhdOp &o : h->cs.ops[SY_EFACE_BC])
    {
        const int n = o.run > 1 ? o.run : 1;
        for (int e = 0; e < n && good; ++e)
            for (int c = 0; c < o.ncomp && good; ++c)
            {
                const int64_t a = o.dst + (int64_t)e * o.dstep + (int64_t)c * o.stride;
                int b, fam, comp, p[3];
                good = o.kind == OP_ZERO && decode(a, b, fam, comp, p) &&
                       (comp == fam || face_source(blk[b].d, fam, p[0], p[1], p[2]) == 1);
                zeroed.insert(a);
            }
    }
    for (size_t b = 0; b < blk.size() && good; ++b)
    {
        const MhdBlockDev &d = blk[b].d;
        for (int fam = 0; fam < 3 && good; ++fam)
            for (int i = 0; i <= d.ni && good; ++i)
                for (int j = 0; j <= d.nj && good; ++j)
                    for (int k = 0; k <= d.nk && good; ++k)
                    {
                        if (i > d.ni - (fam == 0 ? 0 : 1) || j > d.nj - (fam == 1 ? 0 : 1) ||
                            k > d.nk - (fam == 2 ? 0 : 1) || face_source(d, fam, i, j, k) != 1)
                            continue;
                        for (int comp = 0; comp < 3 && good; ++comp)
                            if (comp != fam)
                                good = zeroed.count((int64_t)b * bs + (int64_t)(ST_EFACE + 3 * fam + comp) * sl +
                                                    blk_off(d, i, j, k)) > 0;
                    }
    }
    *ok = good ? 1 : 0;
    return MHD_OK;
}

// ------------------------------------------------- CASE directory, restart
namespace
{
// The restart whitelists of the Lunar solver (Z4_Lunar/src/LunarSolver.cpp:38-39)
// and of the two-fluid Mercury solver (Z5_Mercury/src/MercurySolver.cpp:31-34).
struct RestartSpec
{
    const char *name;
    int field;
};
const RestartSpec kRestartOne[4] = {
    {"U_H", MHD_F_U_H}, {"B_xi", MHD_F_B_XI}, {"B_eta", MHD_F_B_ETA}, {"B_zeta", MHD_F_B_ZETA}};
const RestartSpec kRestartTwo[5] = {{"U_H", MHD_F_U_H},
                                    {"U_Na", MHD_F_U_NA},
                                    {"B_xi", MHD_F_B_XI},
                                    {"B_eta", MHD_F_B_ETA},
                                    {"B_zeta", MHD_F_B_ZETA}};
std::vector<RestartSpec> restart_fields(int model)
{
    return model == 1 ? std::vector<RestartSpec>(kRestartTwo, kRestartTwo + 5)
                      : std::vector<RestartSpec>(kRestartOne, kRestartOne + 4);
}

void box_of(const MhdBlockDev &d, int loc, int *lo, int *hi)
{
    const int N[3] = {d.ni, d.nj, d.nk};
    for (int a = 0; a < 3; ++a)
    {
        lo[a] = -d.g;
        hi[a] = loc_count(loc, a, N[a]) + d.g;
    }
}
} // namespace

namespace mpcns
{
int restart_write(FieldIO &io, const char *path, long long step, double time)
{
    std::vector<RestartField> fields;
    for (const RestartSpec &S : restart_fields(io.block(0).model))
    {
        int region, slot, loc, nc;
        field_slot(S.field, 0, region, slot, loc, nc);
        RestartField F;
        F.name = S.name;
        F.loc_code = restart_loc_code(loc);
        F.ncomp = nc;
        F.nghost = io.block(0).g;
        for (int b = 0; b < io.nblocks(); ++b)
        {
            RestartBlock B;
            box_of(io.block(b), loc, B.lo, B.hi);
            const int rc = io.get(b, S.field, B.v);
            if (rc)
                return rc;
            F.blocks.push_back(std::move(B));
        }
        fields.push_back(std::move(F));
    }
    write_restart_file(path, step, time, fields);
    return MHD_OK;
}

int restart_read(FieldIO &io, const char *path, long long *step, double *time)
{
    long long st = 0;
    double t = 0.0;
    const std::vector<RestartField> fields = read_restart_file(path, st, t);
    // every whitelisted field is checked before any is written
    // (IOModule::ReadRestartBinFile, src/5_io/IOModule_Bin.cpp:234-370: the
    // location, components and extents must agree, and a block the file holds
    // no data for cannot restore a block the case holds); the ghost depth may
    // differ (the overlap is copied)
    const std::vector<RestartSpec> wl = restart_fields(io.block(0).model);
    std::vector<std::pair<const RestartField *, const RestartSpec *>> use;
    for (const RestartField &F : fields)
    {
        const RestartSpec *spec = nullptr;
        for (const RestartSpec &S : wl)
            if (F.name == S.name)
                spec = &S;
        if (!spec)
            continue; // not part of this step's state: skipped, as the reference does
        int region, slot, loc, nc;
        field_slot(spec->field, 0, region, slot, loc, nc);
        if (F.loc_code != restart_loc_code(loc) || F.ncomp != nc)
            throw std::runtime_error("restart: location or component mismatch for " + F.name);
        if ((int)F.blocks.size() != io.nblocks())
            throw std::runtime_error("restart: block count mismatch");
        for (int b = 0; b < io.nblocks(); ++b)
        {
            const RestartBlock &B = F.blocks[b];
            int lo[3], hi[3];
            box_of(io.block(b), loc, lo, hi);
            for (int a = 0; a < 3; ++a)
                if (B.hi[a] - B.lo[a] - 2 * F.nghost != hi[a] - lo[a] - 2 * io.block(b).g)
                    throw std::runtime_error("restart: block extent mismatch for " + F.name);
            if (B.v.empty())
                throw std::runtime_error("restart: block " + std::to_string(b) + " of " + F.name +
                                         " is inactive in the file but held by the case");
        }
        use.emplace_back(&F, spec);
    }
    for (const auto &fs : use)
    {
        const RestartField &F = *fs.first;
        const RestartSpec *spec = fs.second;
        int region, slot, loc, nc;
        field_slot(spec->field, 0, region, slot, loc, nc);
        for (int b = 0; b < io.nblocks(); ++b)
        {
            const RestartBlock &B = F.blocks[b];
            int lo[3], hi[3];
            box_of(io.block(b), loc, lo, hi);
            std::vector<double> aos;
            int rc = io.get(b, spec->field, aos);
            if (rc)
                return rc;
            copy_overlap(B, nc, lo, hi, aos.data());
            if ((rc = io.put(b, spec->field, aos)))
                return rc;
        }
    }
    if (step)
        *step = st;
    if (time)
        *time = t;
    return MHD_OK;
}
} // namespace mpcns

namespace
{
struct HostIO : FieldIO
{
    mhd_host_case *h;
    explicit HostIO(mhd_host_case *h_) : h(h_) {}
    int nblocks() override { return (int)h->cs.blk.size(); }
    const MhdBlockDev &block(int b) override { return h->cs.blk[b].d; }
    // state slots, geometry slots (region 1) or the full J_cell / B_cell tables (regions 2, 3)
    double *base(int b, int field, int &loc, int &nc)
    {
        int region, slot;
        field_slot(field, 0, region, slot, loc, nc);
        BlockArrays &A = h->cs.blk[b];
        return (region == 0 ? A.st.data() : (region == 1 ? A.geo.data() : (region == 2 ? A.wj.data() : A.wb.data()))) +
               (size_t)slot * A.d.slot;
    }
    int get(int b, int field, std::vector<double> &aos) override
    {
        int loc, nc;
        double *p = base(b, field, loc, nc);
        const MhdBlockDev &d = h->cs.blk[b].d;
        aos.assign(aos_count(d, loc, nc), 0.0);
        soa_aos(d, loc, nc, p, aos.data(), true);
        return MHD_OK;
    }
    int put(int b, int field, const std::vector<double> &aos) override
    {
        int loc, nc;
        double *p = base(b, field, loc, nc);
        soa_aos(h->cs.blk[b].d, loc, nc, p, const_cast<double *>(aos.data()), false);
        return MHD_OK;
    }
};
} // namespace

extern "C" int mhd_host_case_write_dir(mhd_host_case *h, const char *dir)
{
    if (!h || !dir || h->cs.blk.empty() || h->cs.blk[0].st.empty())
        return MHD_ERR_ARG;
    try
    {
        write_case_dir(h->cs.grid, dir);
        return MHD_OK;
    }
    catch (const std::exception &e)
    {
        h->err = e.what();
        return MHD_ERR_ARG;
    }
}

extern "C" int mhd_host_case_from_dir(const char *dir, const mhd_wind_desc *d, int ghost, mhd_host_case **out)
{
    if (!dir || !d || !out || ghost < 1)
        return MHD_ERR_ARG;
    try
    {
        auto h = std::make_unique<mhd_host_case>();
        h->cs = make_case_lunar_on(read_case_dir(dir, ghost), d->gamma, d->cfl, wind_params(*d));
        *out = h.release();
        return MHD_OK;
    }
    catch (const std::exception &e)
    {
        host_ctor_error(e.what());
        return MHD_ERR_ARG;
    }
}

extern "C" int mhd_host_write_restart(mhd_host_case *h, const char *path, long long step, double time)
{
    if (!h || !path || h->cs.blk.empty() || h->cs.blk[0].st.empty())
        return MHD_ERR_ARG;
    try
    {
        HostIO io(h);
        return restart_write(io, path, step, time);
    }
    catch (const std::exception &e)
    {
        h->err = e.what();
        return MHD_ERR_ARG;
    }
}

extern "C" int mhd_host_write_tecplot(mhd_host_case *h, const char *path, double time)
{
    if (!h || !path || h->cs.blk.empty() || h->cs.blk[0].st.empty())
        return MHD_ERR_ARG;
    try
    {
        HostIO io(h);
        return tecplot_write(io, h->cs.grid, h->cs.cst, h->cs.L_ref, path, time);
    }
    catch (const std::exception &e)
    {
        h->err = e.what();
        return MHD_ERR_ARG;
    }
}

extern "C" int mhd_host_read_restart(mhd_host_case *h, const char *path, long long *step, double *time)
{
    if (!h || !path || h->cs.blk.empty() || h->cs.blk[0].st.empty())
        return MHD_ERR_ARG;
    try
    {
        HostIO io(h);
        const int rc = restart_read(io, path, step, time);
        if (rc)
            return rc;
        // ghosts from the new interior: U and B syncs (physical, then halo)
        for (int sy : {SY_U_BC0, SY_U_BC1, SY_U_HALO, SY_B_BC, SY_B_HALO})
            apply_ops_host(h->cs.blk, h->cs.ops[sy], h->cs.cst);
        return MHD_OK;
    }
    catch (const std::exception &e)
    {
        h->err = e.what();
        return MHD_ERR_ARG;
    }
}
