// C-ABI of the explicit MHD step (include/beta1_mhd.h): context, case
// placement in device arenas, the step sequence and data exchange.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../../include/beta1_mhd.h"
#include "../host/setup.h"
#include "../layout.h"
#include "kernels.h"

using namespace mpcns;

namespace
{
struct CaseRec
{
    CaseSetup cs;         // grid, constants, op lists (arrays freed after upload)
    int geo_set = -1;     // geometry set index
    int64_t st_base = 0;  // arena offset of block 0 state
    int first_block = 0;  // index into the global block table
    int nblocks = 0;
    int active = 1;
};

struct GeoSet
{
    std::string key;
    int64_t base = 0;
    int64_t block_stride = 0;
};

enum KernelId
{
    KP_DT = 0,
    KP_DTFIN,
    KP_UC,
    KP_FACE,
    KP_EDGE,
    KP_FLUID,
    KP_CURL,
    KP_DERIVED,
    KP_OPS,
    KP_COUNT
};
const char *kKernelName[KP_COUNT] = {"k_dt", "k_dt_finalize", "k_ucontra", "k_face_emf", "k_edge_emf",
                                     "k_fluid", "k_curl_update", "k_derived", "k_ops"};
} // namespace

struct mhd_ctx
{
    mhd_options opt{};
    std::string err;
    hipStream_t q = nullptr;
    std::vector<std::unique_ptr<CaseRec>> cases;
    std::vector<GeoSet> geos;
    bool finalized = false;
    int ucur = 0;

    double *d_st = nullptr;
    int64_t st_n = 0;
    double *d_geo = nullptr;
    int64_t geo_n = 0;
    MhdBlockDev *d_blk = nullptr;
    std::vector<MhdBlockDev> h_blk;
    MhdCaseConst *d_cst = nullptr;
    double *d_dt = nullptr;
    unsigned long long *d_dtbits = nullptr;
    unsigned long long *d_dtscratch = nullptr; // mhd_compute_dt (leaves the step's dt untouched)
    unsigned long long *d_divb = nullptr;
    int *d_err = nullptr;
    int *d_active = nullptr;
    MhdOp *d_ops = nullptr;
    int op_start[SY_COUNT] = {0};
    int op_count[SY_COUNT] = {0};
    mhdk::LaunchShape shape{};

    hipGraphExec_t gexec = nullptr;
    int graph_ucur = -1;

    bool prof = false;
    hipEvent_t ev[2 * KP_COUNT * 4] = {};
    double prof_ms[KP_COUNT] = {0};
    int prof_n[KP_COUNT] = {0};
};

#define HIPCHK(ctx, x)                                                                                                 \
    do                                                                                                                 \
    {                                                                                                                  \
        hipError_t e_ = (x);                                                                                           \
        if (e_ != hipSuccess)                                                                                          \
        {                                                                                                              \
            (ctx)->err = std::string(#x) + ": " + hipGetErrorString(e_);                                               \
            return MHD_ERR_HIP;                                                                                        \
        }                                                                                                              \
    } while (0)

static int fail(mhd_ctx *c, int code, const std::string &msg)
{
    if (c)
        c->err = msg;
    return code;
}

extern "C" int mhd_create(const mhd_options *opt, mhd_ctx **out)
{
    if (!opt || !out)
        return MHD_ERR_ARG;
    if (opt->ghost < 1 || (opt->recon == 1 && opt->ghost < 2))
        return MHD_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= opt->device)
        return MHD_ERR_HIP;
    if (hipSetDevice(opt->device) != hipSuccess)
        return MHD_ERR_HIP;
    mhd_ctx *c = new mhd_ctx();
    c->opt = *opt;
    if (hipStreamCreateWithFlags(&c->q, hipStreamNonBlocking) != hipSuccess)
    {
        delete c;
        return MHD_ERR_HIP;
    }
    *out = c;
    return MHD_OK;
}

static void free_device(mhd_ctx *c)
{
    if (c->gexec)
        (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
    void *ptrs[] = {c->d_st,   c->d_geo, c->d_blk, c->d_cst,    c->d_dt, c->d_dtbits,
                    c->d_divb, c->d_err, c->d_active, c->d_ops, c->d_dtscratch};
    for (void *p : ptrs)
        if (p)
            (void)hipFree(p);
    c->d_st = c->d_geo = nullptr;
    c->d_blk = nullptr;
    c->d_cst = nullptr;
    c->d_dt = nullptr;
    c->d_dtbits = c->d_divb = c->d_dtscratch = nullptr;
    c->d_err = c->d_active = nullptr;
    c->d_ops = nullptr;
}

extern "C" int mhd_destroy(mhd_ctx *c)
{
    if (!c)
        return MHD_OK;
    (void)hipSetDevice(c->opt.device);
    (void)hipStreamSynchronize(c->q);
    free_device(c);
    for (auto &e : c->ev)
        if (e)
            (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->q);
    delete c;
    return MHD_OK;
}

extern "C" const char *mhd_last_error(mhd_ctx *c) { return c ? c->err.c_str() : "null context"; }

static int add_case(mhd_ctx *c, CaseSetup &&cs, int *case_id)
{
    if (c->finalized)
        return fail(c, MHD_ERR_STATE, "cases must be added before mhd_finalize");
    auto r = std::make_unique<CaseRec>();
    r->cs = std::move(cs);
    r->nblocks = (int)r->cs.blk.size();
    if (case_id)
        *case_id = (int)c->cases.size();
    c->cases.push_back(std::move(r));
    return MHD_OK;
}

extern "C" int mhd_add_case_orszag_tang(mhd_ctx *c, const mhd_ot_desc *d, int *case_id)
{
    if (!c || !d || d->n < 4)
        return MHD_ERR_ARG;
    try
    {
        return add_case(c, make_case_orszag_tang(d->n, c->opt.ghost, d->gamma, d->cfl, d->seed, d->noise), case_id);
    }
    catch (const std::exception &e)
    {
        return fail(c, MHD_ERR_ARG, e.what());
    }
}

extern "C" int mhd_add_case_lunar(mhd_ctx *c, const mhd_wind_desc *d, int *case_id)
{
    if (!c || !d || d->nang < 4 || d->nrad < 4)
        return MHD_ERR_ARG;
    try
    {
        WindParams w;
        w.n_cc = d->n_cc;
        w.v_kms = d->v_kms;
        for (int a = 0; a < 3; ++a)
        {
            w.vdir[a] = d->vdir[a];
            w.B_nT[a] = d->B_nT[a];
        }
        w.T_eV = d->T_eV;
        w.m_amu = d->m_amu > 0 ? d->m_amu : 1.0;
        return add_case(c, make_case_lunar(d->nang, d->nrad, d->r_in, d->r_out, c->opt.ghost, d->gamma, d->cfl, w),
                        case_id);
    }
    catch (const std::exception &e)
    {
        return fail(c, MHD_ERR_ARG, e.what());
    }
}

// ------------------------------------------------------------- step driver
static int run_ops(mhd_ctx *c, int sy, int ushift)
{
    mhdk::launch_ops(c->d_ops + c->op_start[sy], c->op_count[sy], c->d_st, c->d_cst, ushift, c->q);
    return MHD_OK;
}

struct ProfScope
{
    mhd_ctx *c;
    int id;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(mhd_ctx *c_, int id_) : c(c_), id(id_)
    {
        if (c->prof)
        {
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
            (void)hipEventRecord(a, c->q);
        }
    }
    ~ProfScope()
    {
        if (c->prof)
        {
            (void)hipEventRecord(b, c->q);
            (void)hipEventSynchronize(b);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, a, b);
            c->prof_ms[id] += ms;
            c->prof_n[id] += 1;
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    }
};

static void derived(mhd_ctx *c)
{
    {
        ProfScope p(c, KP_DERIVED);
        mhdk::launch_derived(c->shape, c->d_blk, c->d_st, c->d_geo, c->d_cst, c->d_dtbits, c->ucur, c->q);
    }
    ProfScope p(c, KP_OPS);
    run_ops(c, SY_JCELL_BC, 0);
}

// One StepOnce (Z4_Lunar/src/LunarSolver_Driver.cpp:21-92) for every case.
static void enqueue_step(mhd_ctx *c)
{
    const int mixed = c->opt.precision == MHD_PREC_MIXED;
    const int ncase = (int)c->cases.size();
    // dt and the contravariant velocities of this step were computed by the
    // previous step's derived-field kernel (or by mhd_finalize)
    {
        ProfScope p(c, KP_DTFIN);
        mhdk::launch_dt_finalize(c->d_dtbits, c->d_dt, c->d_err, c->d_cst, ncase, c->d_active, c->q);
    }
    {
        ProfScope p(c, KP_OPS);
        run_ops(c, SY_UC_CORNER, 0);
    }
    {
        ProfScope p(c, KP_FACE);
        mhdk::launch_face_emf(c->shape, c->d_blk, c->d_st, mixed, c->q);
    }
    {
        ProfScope p(c, KP_OPS);
        run_ops(c, SY_EFACE_BC, 0);
        run_ops(c, SY_EFACE_HALO, 0);
    }
    {
        ProfScope p(c, KP_EDGE);
        mhdk::launch_edge_emf(c->shape, c->d_blk, c->d_st, c->q);
    }
    {
        ProfScope p(c, KP_OPS);
        run_ops(c, SY_EEDGE_BC, 0);
    }
    {
        ProfScope p(c, KP_FLUID);
        mhdk::launch_fluid(c->shape, c->d_blk, c->d_st, c->d_geo, c->d_cst, c->d_dt, c->ucur, mixed, c->opt.recon, c->q);
    }
    {
        ProfScope p(c, KP_CURL);
        mhdk::launch_curl_update(c->shape, c->d_blk, c->d_st, c->d_dt, c->q);
    }
    c->ucur ^= 1;
    {
        ProfScope p(c, KP_OPS);
        run_ops(c, SY_U_BC0, c->ucur);
        run_ops(c, SY_U_BC1, c->ucur);
        run_ops(c, SY_U_HALO, c->ucur);
        run_ops(c, SY_B_BC, 0);
        run_ops(c, SY_B_HALO, 0);
    }
    derived(c);
}

extern "C" int mhd_finalize(mhd_ctx *c)
{
    if (!c)
        return MHD_ERR_ARG;
    if (c->finalized)
        return fail(c, MHD_ERR_STATE, "already finalized");
    if (c->cases.empty())
        return fail(c, MHD_ERR_STATE, "no cases");
    HIPCHK(c, hipSetDevice(c->opt.device));
    // geometry sets and arena placement
    int64_t geo_n = 0, st_n = 0;
    std::map<std::string, int> geo_index;
    c->h_blk.clear();
    for (int ci = 0; ci < (int)c->cases.size(); ++ci)
    {
        CaseRec &r = *c->cases[ci];
        const int64_t slot = r.cs.blk[0].d.slot;
        auto it = geo_index.find(r.cs.shape_key);
        if (it == geo_index.end())
        {
            GeoSet gs;
            gs.key = r.cs.shape_key;
            gs.base = geo_n;
            gs.block_stride = (int64_t)G_NSLOT * slot;
            geo_n += gs.block_stride * r.nblocks;
            geo_index[gs.key] = (int)c->geos.size();
            c->geos.push_back(gs);
        }
        r.geo_set = geo_index[r.cs.shape_key];
        r.st_base = st_n;
        st_n += (int64_t)ST_NSLOT * slot * r.nblocks;
        r.first_block = (int)c->h_blk.size();
        for (int b = 0; b < r.nblocks; ++b)
        {
            MhdBlockDev d = r.cs.blk[b].d;
            d.case_id = ci;
            d.st = r.st_base + (int64_t)b * ST_NSLOT * slot;
            d.geo = c->geos[r.geo_set].base + (int64_t)b * c->geos[r.geo_set].block_stride;
            d.geo32 = -1;
            c->h_blk.push_back(d);
        }
    }
    double budget = c->opt.hbm_budget_gb * 1e9;
    if (budget > 0 && (double)(geo_n + st_n) * 8.0 > budget)
        return fail(c, MHD_ERR_MEM, "device memory budget exceeded");
    c->geo_n = geo_n;
    c->st_n = st_n;
    HIPCHK(c, hipMalloc(&c->d_geo, sizeof(double) * geo_n));
    HIPCHK(c, hipMalloc(&c->d_st, sizeof(double) * st_n));
    // upload geometry (once per set) and state
    std::vector<bool> geo_done(c->geos.size(), false);
    for (auto &rp : c->cases)
    {
        CaseRec &r = *rp;
        const int64_t slot = r.cs.blk[0].d.slot;
        for (int b = 0; b < r.nblocks; ++b)
        {
            const MhdBlockDev &d = c->h_blk[r.first_block + b];
            if (!geo_done[r.geo_set])
                HIPCHK(c, hipMemcpy(c->d_geo + d.geo, r.cs.blk[b].geo.data(), sizeof(double) * G_NSLOT * slot,
                                    hipMemcpyHostToDevice));
            HIPCHK(c, hipMemcpy(c->d_st + d.st, r.cs.blk[b].st.data(), sizeof(double) * ST_NSLOT * slot,
                                hipMemcpyHostToDevice));
            std::vector<double>().swap(r.cs.blk[b].geo);
            std::vector<double>().swap(r.cs.blk[b].st);
        }
        geo_done[r.geo_set] = true;
    }
    // block table, constants, dt
    const int nblk = (int)c->h_blk.size(), ncase = (int)c->cases.size();
    HIPCHK(c, hipMalloc(&c->d_blk, sizeof(MhdBlockDev) * nblk));
    HIPCHK(c, hipMemcpy(c->d_blk, c->h_blk.data(), sizeof(MhdBlockDev) * nblk, hipMemcpyHostToDevice));
    std::vector<MhdCaseConst> cst(ncase);
    std::vector<int> act(ncase);
    for (int ci = 0; ci < ncase; ++ci)
    {
        cst[ci] = c->cases[ci]->cs.cst;
        act[ci] = c->cases[ci]->active;
    }
    HIPCHK(c, hipMalloc(&c->d_cst, sizeof(MhdCaseConst) * ncase));
    HIPCHK(c, hipMemcpy(c->d_cst, cst.data(), sizeof(MhdCaseConst) * ncase, hipMemcpyHostToDevice));
    HIPCHK(c, hipMalloc(&c->d_active, sizeof(int) * ncase));
    HIPCHK(c, hipMemcpy(c->d_active, act.data(), sizeof(int) * ncase, hipMemcpyHostToDevice));
    HIPCHK(c, hipMalloc(&c->d_dt, sizeof(double) * ncase));
    HIPCHK(c, hipMalloc(&c->d_dtbits, sizeof(unsigned long long) * ncase));
    HIPCHK(c, hipMalloc(&c->d_divb, sizeof(unsigned long long) * ncase));
    HIPCHK(c, hipMalloc(&c->d_dtscratch, sizeof(unsigned long long) * ncase));
    HIPCHK(c, hipMalloc(&c->d_err, sizeof(int)));
    HIPCHK(c, hipMemset(c->d_err, 0, sizeof(int)));
    HIPCHK(c, hipMemset(c->d_dt, 0, sizeof(double) * ncase));
    {
        std::vector<unsigned long long> init(ncase);
        const double big = 1e100;
        unsigned long long bits;
        std::memcpy(&bits, &big, 8);
        std::fill(init.begin(), init.end(), bits);
        HIPCHK(c, hipMemcpy(c->d_dtbits, init.data(), 8 * ncase, hipMemcpyHostToDevice));
    }
    // op lists: relocate to global offsets, concatenate per sync group
    std::vector<MhdOp> all;
    for (int sy = 0; sy < SY_COUNT; ++sy)
    {
        c->op_start[sy] = (int)all.size();
        for (int ci = 0; ci < ncase; ++ci)
        {
            const CaseRec &r = *c->cases[ci];
            for (MhdOp o : r.cs.ops[sy])
            {
                o.dst += r.st_base;
                if (o.kind == OP_COPY || o.kind == OP_NEG || o.kind == OP_SOLID_FIRST || o.kind == OP_SOLID_GHOST)
                    o.src += r.st_base;
                o.cas = ci;
                all.push_back(o);
            }
        }
        c->op_count[sy] = (int)all.size() - c->op_start[sy];
    }
    HIPCHK(c, hipMalloc(&c->d_ops, sizeof(MhdOp) * std::max<size_t>(1, all.size())));
    if (!all.empty())
        HIPCHK(c, hipMemcpy(c->d_ops, all.data(), sizeof(MhdOp) * all.size(), hipMemcpyHostToDevice));
    for (auto &rp : c->cases)
        rp->cs.ops.assign(SY_COUNT, {});
    c->shape.nblocks = nblk;
    c->shape.max_ni = c->shape.max_nj = c->shape.max_nk = 0;
    for (const auto &d : c->h_blk)
    {
        c->shape.max_ni = std::max(c->shape.max_ni, d.ni);
        c->shape.max_nj = std::max(c->shape.max_nj, d.nj);
        c->shape.max_nk = std::max(c->shape.max_nk, d.nk);
    }
    c->ucur = 0;
    c->finalized = true;
    derived(c);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->q));
    return MHD_OK;
}

static int check_err(mhd_ctx *c)
{
    int e = 0;
    HIPCHK(c, hipMemcpyAsync(&e, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->q));
    HIPCHK(c, hipStreamSynchronize(c->q));
    if (e)
        return fail(c, MHD_ERR_DT, "dt too small or NaN");
    return MHD_OK;
}

extern "C" int mhd_step(mhd_ctx *c, int nsteps)
{
    if (!c || nsteps < 0)
        return MHD_ERR_ARG;
    if (!c->finalized)
        return fail(c, MHD_ERR_STATE, "mhd_finalize first");
    HIPCHK(c, hipSetDevice(c->opt.device));
    for (int s = 0; s < nsteps; ++s)
        enqueue_step(c);
    HIPCHK(c, hipGetLastError());
    return check_err(c);
}

extern "C" int mhd_step_graph(mhd_ctx *c, int nsteps)
{
    if (!c || nsteps < 0)
        return MHD_ERR_ARG;
    if (!c->finalized)
        return fail(c, MHD_ERR_STATE, "mhd_finalize first");
    HIPCHK(c, hipSetDevice(c->opt.device));
    if (c->prof)
        return mhd_step(c, nsteps);
    int s = 0;
    if (c->ucur != 0 && nsteps > 0)
    {
        enqueue_step(c);
        ++s;
    }
    if (nsteps - s >= 2)
    {
        if (!c->gexec)
        {
            hipGraph_t g;
            HIPCHK(c, hipStreamBeginCapture(c->q, hipStreamCaptureModeThreadLocal));
            enqueue_step(c);
            enqueue_step(c);
            HIPCHK(c, hipStreamEndCapture(c->q, &g));
            HIPCHK(c, hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0));
            (void)hipGraphDestroy(g);
        }
        for (; s + 2 <= nsteps; s += 2)
            HIPCHK(c, hipGraphLaunch(c->gexec, c->q));
    }
    for (; s < nsteps; ++s)
        enqueue_step(c);
    HIPCHK(c, hipGetLastError());
    return check_err(c);
}

extern "C" int mhd_synchronize(mhd_ctx *c)
{
    if (!c)
        return MHD_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->q));
    return MHD_OK;
}

extern "C" int mhd_compute_dt(mhd_ctx *c, int case_id, double *dt)
{
    if (!c || !dt || case_id < 0 || case_id >= (int)c->cases.size())
        return MHD_ERR_ARG;
    if (!c->finalized)
        return fail(c, MHD_ERR_STATE, "mhd_finalize first");
    const double big = 1e100;
    unsigned long long bits;
    std::memcpy(&bits, &big, 8);
    mhdk::launch_fill_u64(c->d_dtscratch, (int)c->cases.size(), bits, c->q);
    mhdk::launch_dt(c->shape, c->d_blk, c->d_st, c->d_geo, c->d_cst, c->d_dtscratch, c->ucur, c->q);
    HIPCHK(c, hipMemcpyAsync(&bits, c->d_dtscratch + case_id, 8, hipMemcpyDeviceToHost, c->q));
    HIPCHK(c, hipStreamSynchronize(c->q));
    std::memcpy(dt, &bits, 8);
    return MHD_OK;
}

extern "C" int mhd_last_dt(mhd_ctx *c, int case_id, double *dt)
{
    if (!c || !dt || case_id < 0 || case_id >= (int)c->cases.size() || !c->finalized)
        return MHD_ERR_ARG;
    HIPCHK(c, hipMemcpyAsync(dt, c->d_dt + case_id, 8, hipMemcpyDeviceToHost, c->q));
    HIPCHK(c, hipStreamSynchronize(c->q));
    return MHD_OK;
}

extern "C" int mhd_divb_max(mhd_ctx *c, int case_id, double *out)
{
    if (!c || !out || case_id < 0 || case_id >= (int)c->cases.size() || !c->finalized)
        return MHD_ERR_ARG;
    mhdk::launch_fill_u64(c->d_divb, (int)c->cases.size(), 0ull, c->q);
    mhdk::launch_divb(c->shape, c->d_blk, c->d_st, c->d_geo, c->d_divb, c->q);
    unsigned long long bits = 0;
    HIPCHK(c, hipMemcpyAsync(&bits, c->d_divb + case_id, 8, hipMemcpyDeviceToHost, c->q));
    HIPCHK(c, hipStreamSynchronize(c->q));
    std::memcpy(out, &bits, 8);
    return MHD_OK;
}

// ----------------------------------------------------------- data exchange
static int exchange(mhd_ctx *c, int case_id, int block, int field, double *aos, size_t n, bool down)
{
    if (!c || !aos || case_id < 0 || case_id >= (int)c->cases.size())
        return MHD_ERR_ARG;
    if (!c->finalized)
        return fail(c, MHD_ERR_STATE, "mhd_finalize first");
    CaseRec &r = *c->cases[case_id];
    if (block < 0 || block >= r.nblocks)
        return MHD_ERR_ARG;
    int region, slot, loc, nc;
    if (!field_slot(field, c->ucur, region, slot, loc, nc))
        return fail(c, MHD_ERR_ARG, "field not stored on the device");
    const MhdBlockDev &d = c->h_blk[r.first_block + block];
    if (n != aos_count(d, loc, nc))
        return fail(c, MHD_ERR_ARG, "buffer size does not match the FieldBlock extent");
    HIPCHK(c, hipStreamSynchronize(c->q));
    double *dev = (region == 0 ? c->d_st + d.st : c->d_geo + d.geo) + (int64_t)slot * d.slot;
    std::vector<double> soa((size_t)nc * d.slot);
    HIPCHK(c, hipMemcpy(soa.data(), dev, sizeof(double) * soa.size(), hipMemcpyDeviceToHost));
    soa_aos(d, loc, nc, soa.data(), aos, down);
    if (!down)
        HIPCHK(c, hipMemcpy(dev, soa.data(), sizeof(double) * soa.size(), hipMemcpyHostToDevice));
    return MHD_OK;
}

extern "C" int mhd_download(mhd_ctx *c, int case_id, int block, int field, double *aos, size_t n)
{
    return exchange(c, case_id, block, field, aos, n, true);
}

extern "C" int mhd_upload(mhd_ctx *c, int case_id, int block, int field, const double *aos, size_t n)
{
    return exchange(c, case_id, block, field, const_cast<double *>(aos), n, false);
}

extern "C" int mhd_field_count(mhd_ctx *c, int case_id, int field, size_t *n)
{
    if (!c || !n || case_id < 0 || case_id >= (int)c->cases.size())
        return MHD_ERR_ARG;
    int region, slot, loc, nc;
    if (!field_slot(field, 0, region, slot, loc, nc))
        return MHD_ERR_ARG;
    *n = aos_count(c->cases[case_id]->cs.blk[0].d, loc, nc);
    return MHD_OK;
}

extern "C" int mhd_case_info(mhd_ctx *c, int case_id, int *nblocks, int *ni, int *nj, int *nk, int *ghost)
{
    if (!c || case_id < 0 || case_id >= (int)c->cases.size())
        return MHD_ERR_ARG;
    const CaseRec &r = *c->cases[case_id];
    const MhdBlockDev &d = r.cs.blk[0].d;
    if (nblocks)
        *nblocks = r.nblocks;
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

extern "C" int mhd_get_faces(mhd_ctx *c, int case_id, int block, int *out)
{
    if (!c || !out || case_id < 0 || case_id >= (int)c->cases.size())
        return MHD_ERR_ARG;
    const CaseRec &r = *c->cases[case_id];
    if (block < 0 || block >= r.nblocks)
        return MHD_ERR_ARG;
    const Block &B = r.cs.grid.blocks[block];
    for (int f = 0; f < 6; ++f)
    {
        const FaceLink &L = B.face[f];
        int *o = out + 14 * f;
        o[0] = L.kind;
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

extern "C" int mhd_get_nodes(mhd_ctx *c, int case_id, int block, int comp, double *out, size_t n)
{
    if (!c || !out || case_id < 0 || case_id >= (int)c->cases.size() || comp < 0 || comp > 2)
        return MHD_ERR_ARG;
    const CaseRec &r = *c->cases[case_id];
    if (block < 0 || block >= r.nblocks)
        return MHD_ERR_ARG;
    const Block &B = r.cs.grid.blocks[block];
    const std::vector<double> &v = comp == 0 ? B.x : (comp == 1 ? B.y : B.z);
    if (n != v.size())
        return fail(c, MHD_ERR_ARG, "node buffer size mismatch");
    std::memcpy(out, v.data(), sizeof(double) * n);
    return MHD_OK;
}

extern "C" int mhd_get_constants(mhd_ctx *c, int case_id, double *o)
{
    if (!c || !o || case_id < 0 || case_id >= (int)c->cases.size())
        return MHD_ERR_ARG;
    const MhdCaseConst &k = c->cases[case_id]->cs.cst;
    const double v[22] = {k.gamma,   k.cfl,     k.inver_MA2, k.momentum_hall_coeff, k.M_H,     k.M_ref,   k.ne_floor, k.state_coeff,
                          k.qinf[0], k.qinf[1], k.qinf[2],   k.qinf[3],             k.qinf[4], k.B_imf[0], k.B_imf[1], k.B_imf[2],
                          k.u_inf[0], k.u_inf[1], k.u_inf[2], k.dt_floor,           0.0,       0.0};
    std::memcpy(o, v, sizeof(v));
    return MHD_OK;
}

extern "C" int mhd_profile(mhd_ctx *c, int enable)
{
    if (!c)
        return MHD_ERR_ARG;
    c->prof = enable != 0;
    for (int k = 0; k < KP_COUNT; ++k)
    {
        c->prof_ms[k] = 0;
        c->prof_n[k] = 0;
    }
    return MHD_OK;
}

extern "C" int mhd_profile_read(mhd_ctx *c, int kernel, double *ms, int *launches, const char **name)
{
    if (!c || kernel < 0 || kernel >= KP_COUNT)
        return MHD_ERR_ARG;
    if (ms)
        *ms = c->prof_ms[kernel];
    if (launches)
        *launches = c->prof_n[kernel];
    if (name)
        *name = kKernelName[kernel];
    return MHD_OK;
}

extern "C" int mhd_memory(mhd_ctx *c, double *bytes)
{
    if (!c || !bytes)
        return MHD_ERR_ARG;
    *bytes = 8.0 * (double)(c->st_n + c->geo_n);
    return MHD_OK;
}
