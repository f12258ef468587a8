// C-ABI of the explicit MHD step (include/beta1_mhd.h): context, case
// placement on the device, live admission / retirement, launch tiling and
// tuning, the step sequence and data exchange.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "../../../include/beta1_mhd.h"
#include "../host/host_case.h"
#include "../host/setup.h"
#include "../layout.h"
#include "kernels.h"

using namespace mpcns;

namespace
{
struct CaseRec
{
    CaseSetup cs; // grid, topology, constants; arrays and op lists freed once placed
    int id = -1;
    int geo_set = -1;
    double *d_st = nullptr; // the case's state allocation
    int64_t slot = 0;
    int nblocks = 0;
    int state = 0;        // 0 pending, 1 active, 2 retired, 3 draft (mhd_case_begin .. mhd_case_end)
    int first_block = -1; // index into the block table while active
    // segments in the op sub-lists (kOpCls per sync, see OpGroup)
    size_t op_off[SY_COUNT * 4] = {0}, op_n[SY_COUNT * 4] = {0};
    long long steps = 0;
    size_t st_bytes = 0;
    mhd_host_case *draft = nullptr; // the blocks given so far (state 3)
    CaseRec() = default;
    CaseRec(const CaseRec &) = delete;
    CaseRec &operator=(const CaseRec &) = delete;
    ~CaseRec()
    {
        if (draft)
            mhd_host_case_destroy(draft);
    }
};

struct GeoSet
{
    std::string key;
    char *d = nullptr; // per block: metric slots in double, weight slots of weight_esize bytes
    int64_t block_bytes = 0;
    int refs = 0;
    size_t bytes = 0;
};

struct TileArr
{
    int *d_start = nullptr;
    size_t cap_start = 0;
    int n = 0;
};


// Op list of one sync group on the device: the cases' ops back to back,
// each case's segment appended when it is placed and cut out when it retires.
// The records of one sync, in sub-lists by run length: runs shorter than
// kLongRun on a thread each, longer ones on a wavefront each (its lanes take
// the run's elements in turn), so a list mixing single points and long rows
// has no thread walking a row alone; and apart, the records the step's
// kernels apply themselves (MhdOp::fold), which only refresh and the
// per-group syncs run.
constexpr int kOpCls = 4, kLongRun = 16;
static int op_cls(const MhdOp &o) { return (o.fold ? 2 : 0) + (o.run >= kLongRun ? 1 : 0); }
struct OpGroup
{
    MhdOp *d = nullptr;
    size_t cap = 0, n = 0;
};

enum Range
{
    R_CELL = 0, // cells [0, N)
    R_FACE,     // [0, N]   (faces / edges incl. both boundary layers)
    R_EXT,      // [-1, N]  (cells incl. one ghost ring)
    R_COUNT
};
const int kRangeLo[R_COUNT] = {0, 0, -1}, kRangeHi[R_COUNT] = {0, 1, 1};
const int kJt[3] = {1, 2, 4};
// launch variants: workgroup rows kJt[v % 3], tile order i-inner if
// v % 6 >= 3, geometry-shared tile map (cases of one geometry set
// interleaved per tile) if v >= 6; v = 12 + rows index: geometry-shared map
// with member-looping tiles (a tile covers every case of its geometry set,
// which the kernel loads once), for the kernels in loops_members
const int kVariants = 12, kLoopVariants = 15;

// Launch tiling over a set of blocks.  An entry is one block or, in a
// geometry-shared set, the same block of every case sharing a geometry set
// (lead_k cases, block-table indices members[lead_moff ..]).  Empty members:
// entry e is block e of the whole table.
struct TileSet
{
    std::vector<int> lead_k, lead_moff, members;
    int *d_lead_k = nullptr, *d_lead_moff = nullptr, *d_members = nullptr;
    size_t cap_k = 0, cap_moff = 0, cap_mem = 0;
    int nent = 0;
    TileArr tiles[R_COUNT][3];  // [range][rows]
    TileArr chunks[2][R_COUNT][3]; // [kChunkCi index][range][rows]
    TileArr cols[3];           // column tiles of the interior cells (kColCi[q] planes each)
    TileArr shell;             // 256 shell points of the face range per tile
    TileArr pad[R_COUNT][3];   // padded-row plane chunks (kChunkCi[0] planes) [range][rows]
    TileArr kl[R_COUNT][3];    // geometry-shared sets: member-looping tiles [range][rows]
};
// planes along i per column-tile workgroup
// planes per column tile of the first-order fluid kernels, by the launch
// variant's row index: a small grid (the 64^3 box: 128 tiles of 8 planes
// for 256 CUs) needs the shorter columns, large ones the longer
constexpr int kColCi[3] = {8, 4, 2};
// planes along i per plane-chunk tile (kernels that march along i and carry
// the values plane i - 1 shares with plane i in registers): two sizes
constexpr int kChunkCi[2] = {4, 8};
// block sets: every block, the single-fluid blocks, the two-fluid blocks
enum BlockSet
{
    BS_ALL = 0,
    BS_ONE,
    BS_TWO,
    BS_COUNT
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
    KP_JCELL,
    KP_FLUID2,
    KP_DERIVED2,
    KP_AMBI,
    KP_ERES,
    KP_JCELL_RES,
    KP_CURL_RES,
    KP_EMF,   // face + edge EMF in one pass (whole steps)
    KP_SHELL, // face EMF of the shell points before the Eface sync (whole steps)
    KP_COUNT
};
const char *kKernelName[KP_COUNT] = {"k_dt",      "k_dt_finalize", "k_ucontra", "k_face_emf", "k_edge_emf",
                                     "k_fluid",   "k_curl_update", "k_derived", "k_ops",      "k_jcell",
                                     "k_fluid2",  "k_derived2",    "k_emf_extra", "k_eres",     "k_jcell_res",
                                     "k_curl_res", "k_emf",       "k_face_shell"};
const int kKernelRange[KP_COUNT] = {R_CELL, R_CELL, R_EXT,  R_FACE, R_FACE, R_CELL, R_FACE, R_EXT, R_CELL,
                                    R_CELL, R_CELL, R_EXT,  R_FACE, R_FACE, R_CELL, R_FACE, R_FACE, R_FACE};
// kernels with a member-looping variant (k_derived_k)
bool loops_members(int kernel) { return kernel == KP_DERIVED; }
// kernels over padded-row plane chunks (their lanes take neighbours along k
// from the lanes beside them), kChunkCi[0] planes a tile
bool uses_padded_tiles(int kernel)
{
    return kernel == KP_EMF || kernel == KP_AMBI || kernel == KP_JCELL || kernel == KP_JCELL_RES;
}
} // namespace

struct mhd_ctx
{
    mhd_options opt{};
    std::string err;
    hipStream_t q = nullptr;
    // second stream for the independent branches of a step (opt.overlap):
    // fork[0]/join[0] around the fluid update, fork[1]/join[1] around J_cell
    hipStream_t q2 = nullptr;
    hipEvent_t fork[2] = {nullptr, nullptr}, join[2] = {nullptr, nullptr};
    std::vector<std::unique_ptr<CaseRec>> cases;
    std::vector<GeoSet> geos;
    bool finalized = false;
    int ucur = 0;
    size_t bytes_used = 0;

    // block table and launch tiling (rebuilt whenever the active set changes)
    std::vector<MhdBlockDev> h_blk;
    MhdBlockDev *d_blk = nullptr;
    size_t blk_cap = 0;
    TileSet ts[BS_COUNT][2]; // [block set][geometry-shared]
    int jt_idx[KP_COUNT];
    int n_two = 0; // two-fluid blocks in the table
    // whole steps form the face and edge EMF in one pass (k_face_shell, the
    // Eface sync, k_emf); needs every block extent >= 2 (shell enumeration)
    bool emf_fused = false;

    // case-indexed device arrays
    int case_cap = 0;
    MhdCaseConst *d_cst = nullptr;
    double *d_dt = nullptr, *d_zero = nullptr;
    unsigned long long *d_dtbits = nullptr, *d_dtscratch = nullptr, *d_divb = nullptr;
    double *d_dtres = nullptr; // per-case dt of the current resistive subcycle
    unsigned long long *d_diag = nullptr; // 17 order keys per case (min/max diagnostics)
    size_t cap_diag = 0;
    // this device's static scratch of mhd_diagnostics_noalloc (17 keys for
    // each of up to mhdk::DIAG_SEL_MAX cases; not owned, never freed)
    unsigned long long *d_scratch = nullptr;
    int max_nsub = 0;          // resistive subcycles of the resident cases (0: none)
    int *d_err = nullptr, *d_active = nullptr;

    // op lists of the active cases, one device array per sync group
    OpGroup og[SY_COUNT * kOpCls]; // sub-list cls of sync sy at sy * kOpCls + cls

    // captured steps: kGraphShort and kGraphLong consecutive steps (even: the U double buffer alternates)
    hipGraphExec_t gexec = nullptr, gexec_long = nullptr;
    // per-stage entry points: mhd_fluid_rhs has written the next U buffer and
    // mhd_apply_update has not yet made it current
    bool fluid_pending = false;
    // enqueue_step is recording a whole step: its kernels fold the k-face
    // boundary work (MhdTileMap::fold) and its syncs skip those records
    bool in_step = false;

    bool prof = false;
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

// The metric (Jac, JD) is double in both precisions; mixed-precision contexts
// store the Hodge weights and the B_cell / J_cell weight tables in float.
static int mixed_of(const mhd_ctx *c) { return c->opt.precision == MHD_PREC_MIXED; }
static int weight_esize(const mhd_ctx *c) { return mixed_of(c) ? (int)sizeof(float) : (int)sizeof(double); }

static int fail(mhd_ctx *c, int code, const std::string &msg)
{
    if (c)
        c->err = msg;
    return code;
}

static unsigned long long dbits(double v)
{
    unsigned long long b;
    std::memcpy(&b, &v, 8);
    return b;
}

template <class T> static void free_t(T *&p)
{
    if (p)
        (void)hipFree((void *)p);
    p = nullptr;
}

// ------------------------------------------------------------ life cycle
extern "C" int mhd_create(const mhd_options *opt, mhd_ctx **out)
{
    if (!opt || !out)
        return MHD_ERR_ARG;
    // two ghost layers at least: the stencil kernels issue their loads ahead
    // of the per-family range checks, and the slot padding of two layers is
    // what keeps every such load inside the block's box
    if (opt->ghost < 2 || opt->precision < 0 || opt->precision > 1 || opt->recon < 0 || opt->recon > 1)
        return MHD_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= opt->device)
        return MHD_ERR_HIP;
    if (hipSetDevice(opt->device) != hipSuccess)
        return MHD_ERR_HIP;
    mhd_ctx *c = new mhd_ctx();
    c->opt = *opt;
    for (int k = 0; k < KP_COUNT; ++k)
        c->jt_idx[k] = 2; // 64 x 4, j-tile before i, until tuned
    bool ok = hipStreamCreateWithFlags(&c->q, hipStreamNonBlocking) == hipSuccess &&
              hipStreamCreateWithFlags(&c->q2, hipStreamNonBlocking) == hipSuccess;
    for (int e = 0; e < 2 && ok; ++e)
        ok = hipEventCreateWithFlags(&c->fork[e], hipEventDisableTiming) == hipSuccess &&
             hipEventCreateWithFlags(&c->join[e], hipEventDisableTiming) == hipSuccess;
    // the device address of the static diagnostics scratch (device current)
    ok = ok && mhdk::diag_scratch(opt->device, &c->d_scratch) == hipSuccess && c->d_scratch;
    if (!ok)
    {
        mhd_destroy(c);
        return MHD_ERR_HIP;
    }
    *out = c;
    return MHD_OK;
}

extern "C" int mhd_destroy(mhd_ctx *c)
{
    if (!c)
        return MHD_OK;
    (void)hipSetDevice(c->opt.device);
    if (c->q)
        (void)hipStreamSynchronize(c->q);
    if (c->q2)
        (void)hipStreamSynchronize(c->q2);
    if (c->gexec)
        (void)hipGraphExecDestroy(c->gexec);
    if (c->gexec_long)
        (void)hipGraphExecDestroy(c->gexec_long);
    for (auto &r : c->cases)
        free_t(r->d_st);
    for (auto &g : c->geos)
        free_t(g.d);
    for (auto &set : c->ts)
        for (TileSet &T : set)
        {
            for (auto &rr : T.tiles)
                for (auto &t : rr)
                    free_t(t.d_start);
            for (auto &cc : T.chunks)
                for (auto &rr : cc)
                    for (auto &t : rr)
                        free_t(t.d_start);
            for (TileArr &A : T.cols)
                free_t(A.d_start);
            free_t(T.shell.d_start);
            for (auto &rr : T.pad)
                for (auto &t : rr)
                    free_t(t.d_start);
            free_t(T.d_lead_k);
            free_t(T.d_lead_moff);
            free_t(T.d_members);
        }
    for (auto &g : c->og)
        free_t(g.d);
    free_t(c->d_blk);
    free_t(c->d_cst);
    free_t(c->d_dt);
    free_t(c->d_zero);
    free_t(c->d_dtbits);
    free_t(c->d_dtscratch);
    free_t(c->d_divb);
    free_t(c->d_dtres);
    free_t(c->d_diag);
    free_t(c->d_err);
    free_t(c->d_active);
    for (int e = 0; e < 2; ++e)
    {
        if (c->fork[e])
            (void)hipEventDestroy(c->fork[e]);
        if (c->join[e])
            (void)hipEventDestroy(c->join[e]);
    }
    if (c->q2)
        (void)hipStreamDestroy(c->q2);
    if (c->q)
        (void)hipStreamDestroy(c->q);
    delete c;
    return MHD_OK;
}

extern "C" const char *mhd_last_error(mhd_ctx *c) { return c ? c->err.c_str() : "null context"; }

// ------------------------------------------------------- device tables
template <class T> static int grow(mhd_ctx *c, T *&p, size_t &cap, size_t need)
{
    if (need <= cap && p)
        return MHD_OK;
    const size_t ncap = std::max<size_t>(std::max<size_t>(need, 2 * cap), 1);
    free_t(p);
    HIPCHK(c, hipMalloc((void **)&p, sizeof(T) * ncap));
    cap = ncap;
    return MHD_OK;
}

// Forget the captured step (its kernel arguments name the block table, tile
// maps, op groups and per-case arrays as they are now).
static void drop_graph(mhd_ctx *c)
{
    for (hipGraphExec_t *g : {&c->gexec, &c->gexec_long})
        if (*g)
        {
            (void)hipGraphExecDestroy(*g);
            *g = nullptr;
        }
}

template <class T> static int regrow(mhd_ctx *c, T *&p, int old, int cap, bool keep)
{
    T *np = nullptr;
    HIPCHK(c, hipMalloc((void **)&np, sizeof(T) * cap));
    HIPCHK(c, hipMemset(np, 0, sizeof(T) * cap));
    if (keep && p && old > 0)
        HIPCHK(c, hipMemcpy(np, p, sizeof(T) * old, hipMemcpyDeviceToDevice));
    free_t(p);
    p = np;
    return MHD_OK;
}

static int ensure_case_capacity(mhd_ctx *c)
{
    const int need = (int)c->cases.size();
    if (need <= c->case_cap && c->case_cap > 0)
        return MHD_OK;
    const int old = c->case_cap;
    const int cap = std::max(8, std::max(need, 2 * old));
    HIPCHK(c, hipStreamSynchronize(c->q));
    // the captured step holds the per-case arrays below as kernel arguments:
    // it must not be replayed once they move
    drop_graph(c);
    int rc;
    if ((rc = regrow(c, c->d_cst, old, cap, true)) || (rc = regrow(c, c->d_dt, old, cap, true)) ||
        (rc = regrow(c, c->d_zero, old, cap, false)) || (rc = regrow(c, c->d_dtbits, old, cap, true)) ||
        (rc = regrow(c, c->d_dtscratch, old, cap, false)) || (rc = regrow(c, c->d_divb, old, cap, false)) ||
        (rc = regrow(c, c->d_dtres, old, cap, false)) ||
        (rc = regrow(c, c->d_err, old, cap, true)) || (rc = regrow(c, c->d_active, old, cap, true)))
        return rc;
    std::vector<unsigned long long> big(cap - old, dbits(1e100));
    HIPCHK(c, hipMemcpy(c->d_dtbits + old, big.data(), 8 * (size_t)(cap - old), hipMemcpyHostToDevice));
    c->case_cap = cap;
    return MHD_OK;
}

// Entries of every block set (plain: one block each; geometry-shared: per
// geometry set in order of first use, per block, the set's cases in case
// order) and their tile starts per launch range and workgroup height.
static int build_tiles(mhd_ctx *c)
{
    const int nblk = (int)c->h_blk.size();
    c->n_two = 0;
    c->emf_fused = nblk > 0;
    for (const MhdBlockDev &d : c->h_blk)
    {
        c->n_two += d.model == 1;
        c->emf_fused = c->emf_fused && d.ni >= 2 && d.nj >= 2 && d.nk >= 2;
    }
    for (int s = 0; s < BS_COUNT; ++s)
    {
        auto in_set = [&](int model) { return s == BS_ALL || (s == BS_ONE) == (model == 0); };
        for (int sh = 0; sh < 2; ++sh)
        {
            TileSet &T = c->ts[s][sh];
            T.lead_k.clear();
            T.lead_moff.clear();
            T.members.clear();
            if (sh == 0)
            {
                if (s == BS_ALL)
                    T.nent = nblk; // identity: no member list
                else
                    for (int b = 0; b < nblk; ++b)
                        if (in_set(c->h_blk[b].model))
                        {
                            T.lead_k.push_back(1);
                            T.lead_moff.push_back((int)T.members.size());
                            T.members.push_back(b);
                        }
            }
            else
            {
                std::vector<std::vector<int>> by_geo(c->geos.size());
                std::vector<int> geo_order;
                for (auto &rp : c->cases)
                    if (rp->state == 1 && in_set(rp->cs.blk[0].d.model))
                    {
                        if (by_geo[rp->geo_set].empty())
                            geo_order.push_back(rp->geo_set);
                        by_geo[rp->geo_set].push_back(rp->id);
                    }
                for (int g : geo_order)
                {
                    const std::vector<int> &cs = by_geo[g];
                    for (int b = 0; b < c->cases[cs[0]]->nblocks; ++b)
                    {
                        T.lead_k.push_back((int)cs.size());
                        T.lead_moff.push_back((int)T.members.size());
                        for (int ci : cs)
                            T.members.push_back(c->cases[ci]->first_block + b);
                    }
                }
            }
            if (!(s == BS_ALL && sh == 0))
                T.nent = (int)T.lead_k.size();
            int rc;
            if ((rc = grow(c, T.d_lead_k, T.cap_k, std::max<size_t>(T.lead_k.size(), 1))) ||
                (rc = grow(c, T.d_lead_moff, T.cap_moff, std::max<size_t>(T.lead_moff.size(), 1))) ||
                (rc = grow(c, T.d_members, T.cap_mem, std::max<size_t>(T.members.size(), 1))))
                return rc;
            if (!T.lead_k.empty())
            {
                HIPCHK(c, hipMemcpy(T.d_lead_k, T.lead_k.data(), sizeof(int) * T.lead_k.size(), hipMemcpyHostToDevice));
                HIPCHK(c, hipMemcpy(T.d_lead_moff, T.lead_moff.data(), sizeof(int) * T.lead_moff.size(),
                                    hipMemcpyHostToDevice));
                HIPCHK(c, hipMemcpy(T.d_members, T.members.data(), sizeof(int) * T.members.size(),
                                    hipMemcpyHostToDevice));
            }
            const bool ident = T.members.empty();
            for (int r = 0; r < R_COUNT; ++r)
                for (int ji = 0; ji < 9; ++ji)
                {
                    const int ch = ji >= 3 ? kChunkCi[ji / 3 - 1] : 1;
                    TileArr &A = ji >= 3 ? T.chunks[ji / 3 - 1][r][ji % 3] : T.tiles[r][ji];
                    const int lo = kRangeLo[r], hi = kRangeHi[r], jt = kJt[ji % 3];
                    std::vector<int> start(std::max(T.nent, 1), 0);
                    long long total = 0;
                    for (int e = 0; e < T.nent; ++e)
                    {
                        const MhdBlockDev &d = c->h_blk[ident ? e : T.members[T.lead_moff[e]]];
                        const int ei = d.ni - lo + hi, ej = d.nj - lo + hi, ek = d.nk - lo + hi;
                        start[e] = (int)total;
                        total += (long long)(ident ? 1 : T.lead_k[e]) * ((ei + ch - 1) / ch) *
                                 ((ej * ek + 64 * jt - 1) / (64 * jt));
                    }
                    if (total > (1LL << 31) - 64)
                        return fail(c, MHD_ERR_MEM, "too many tiles for one launch");
                    A.n = (int)total;
                    if ((rc = grow(c, A.d_start, A.cap_start, start.size())))
                        return rc;
                    HIPCHK(c, hipMemcpy(A.d_start, start.data(), sizeof(int) * start.size(), hipMemcpyHostToDevice));
                }
            if (sh == 1)
                for (int r = 0; r < R_COUNT; ++r)
                    for (int ji = 0; ji < 3; ++ji)
                    {
                        TileArr &A = T.kl[r][ji];
                        const int lo = kRangeLo[r], hi = kRangeHi[r], jt = kJt[ji];
                        std::vector<int> start(std::max(T.nent, 1), 0);
                        long long total = 0;
                        for (int e = 0; e < T.nent; ++e)
                        {
                            const MhdBlockDev &d = c->h_blk[T.members[T.lead_moff[e]]];
                            const int ei = d.ni - lo + hi, ej = d.nj - lo + hi, ek = d.nk - lo + hi;
                            start[e] = (int)total;
                            total += (long long)ei * ((ej * ek + 64 * jt - 1) / (64 * jt));
                        }
                        if (total > (1LL << 31) - 64)
                            return fail(c, MHD_ERR_MEM, "too many tiles for one launch");
                        A.n = (int)total;
                        if ((rc = grow(c, A.d_start, A.cap_start, start.size())))
                            return rc;
                        HIPCHK(c, hipMemcpy(A.d_start, start.data(), sizeof(int) * start.size(), hipMemcpyHostToDevice));
                    }
            // padded-row tiles of each range (mhdk::pad_tiles_per_plane tiles of
            // 62 owned points a wavefront per plane, kChunkCi[0] planes per tile)
            for (int r = 0; r < R_COUNT; ++r)
            for (int ji = 0; ji < 3; ++ji)
            {
                TileArr &A = T.pad[r][ji];
                const int lo = kRangeLo[r], hi = kRangeHi[r];
                std::vector<int> start(std::max(T.nent, 1), 0);
                long long total = 0;
                for (int e = 0; e < T.nent; ++e)
                {
                    const MhdBlockDev &d = c->h_blk[ident ? e : T.members[T.lead_moff[e]]];
                    start[e] = (int)total;
                    total += (long long)(ident ? 1 : T.lead_k[e]) * ((d.ni - lo + hi + kChunkCi[0] - 1) / kChunkCi[0]) *
                             mhdk::pad_tiles_per_plane(d.nj - lo + hi, d.nk - lo + hi, kJt[ji]);
                }
                if (total > (1LL << 31) - 64)
                    return fail(c, MHD_ERR_MEM, "too many tiles for one launch");
                A.n = (int)total;
                if ((rc = grow(c, A.d_start, A.cap_start, start.size())))
                    return rc;
                HIPCHK(c, hipMemcpy(A.d_start, start.data(), sizeof(int) * start.size(), hipMemcpyHostToDevice));
            }
            {
                TileArr &A = T.shell;
                std::vector<int> start(std::max(T.nent, 1), 0);
                long long total = 0;
                for (int e = 0; e < T.nent; ++e)
                {
                    const MhdBlockDev &d = c->h_blk[ident ? e : T.members[T.lead_moff[e]]];
                    start[e] = (int)total;
                    if (d.ni >= 2 && d.nj >= 2 && d.nk >= 2)
                        total += (long long)(ident ? 1 : T.lead_k[e]) * ((shell_count(d) + 255) / 256);
                }
                if (total > (1LL << 31) - 64)
                    return fail(c, MHD_ERR_MEM, "too many tiles for one launch");
                A.n = (int)total;
                if ((rc = grow(c, A.d_start, A.cap_start, start.size())))
                    return rc;
                HIPCHK(c, hipMemcpy(A.d_start, start.data(), sizeof(int) * start.size(), hipMemcpyHostToDevice));
            }
            for (int q = 0; q < 3; ++q)
            {
                TileArr &A = T.cols[q];
                const int tj = mhdk::col_tile_rows();
                std::vector<int> start(std::max(T.nent, 1), 0);
                long long total = 0;
                for (int e = 0; e < T.nent; ++e)
                {
                    const MhdBlockDev &d = c->h_blk[ident ? e : T.members[T.lead_moff[e]]];
                    start[e] = (int)total;
                    total += (long long)(ident ? 1 : T.lead_k[e]) * ((d.ni + kColCi[q] - 1) / kColCi[q]) *
                             ((d.nj + tj - 1) / tj) * ((d.nk + 63) / 64);
                }
                if (total > (1LL << 31) - 64)
                    return fail(c, MHD_ERR_MEM, "too many tiles for one launch");
                A.n = (int)total;
                if ((rc = grow(c, A.d_start, A.cap_start, start.size())))
                    return rc;
                HIPCHK(c, hipMemcpy(A.d_start, start.data(), sizeof(int) * start.size(), hipMemcpyHostToDevice));
            }
        }
    }
    return MHD_OK;
}

// The block set a kernel kind runs over: the model-specific kernels over
// their model's blocks (the single-fluid ones over every block when no
// two-fluid case is resident), the shared kernels over every block.
static int set_of(const mhd_ctx *c, int kernel)
{
    if (kernel == KP_FLUID2 || kernel == KP_DERIVED2 || kernel == KP_AMBI || kernel == KP_ERES ||
        kernel == KP_JCELL_RES || kernel == KP_CURL_RES)
        return BS_TWO;
    if ((kernel == KP_FLUID || kernel == KP_DERIVED) && c->n_two > 0)
        return BS_ONE;
    return BS_ALL;
}

// Plane-chunk size of a kernel (index into kChunkCi), -1 for one plane per
// tile; measured per kernel on the 6 x 64 x 64 x 128 shell: the face EMF
// 152 -> 102 us at 4 planes; the edge EMF is faster per point (its carried
// values halve the occupancy).  The padded-row kernels take kChunkCi[0].
static int chunk_of(int kernel)
{
    if (kernel == KP_FACE || kernel == KP_EMF)
        return 0;
    return -1;
}

static mhdk::Launch launch_of(const mhd_ctx *c, int kernel, int jt_index = -1)
{
    const int r = kKernelRange[kernel];
    const int v = jt_index >= 0 ? jt_index : c->jt_idx[kernel];
    const int ji = v % 3, sh = v >= 6 ? 1 : 0;
    const bool kl = v >= kVariants;
    const TileSet &S = c->ts[set_of(c, kernel)][sh];
    const int ch = chunk_of(kernel);
    const bool padded = uses_padded_tiles(kernel) && !kl;
    const TileArr &T = padded ? S.pad[r][ji] : kl ? S.kl[r][ji] : ch >= 0 ? S.chunks[ch][r][ji] : S.tiles[r][ji];
    const bool ident = S.members.empty();
    mhdk::Launch L;
    L.tm.blks = c->d_blk;
    L.tm.block_start = T.d_start;
    L.tm.lead_k = ident ? nullptr : S.d_lead_k;
    L.tm.lead_moff = ident ? nullptr : S.d_lead_moff;
    L.tm.members = ident ? nullptr : S.d_members;
    L.tm.nblk = S.nent;
    L.tm.ntiles = T.n;
    L.tm.lo = kRangeLo[r];
    L.tm.hi = kRangeHi[r];
    L.tm.xcd_chunk = c->opt.xcd_order ? (T.n + 7) / 8 : 0;
    L.tm.i_inner = !kl && (v % 6) >= 3;
    L.tm.ci = padded ? kChunkCi[0] : ch >= 0 ? kChunkCi[ch] : 0;
    L.tm.kloop = kl;
    L.tm.fold = c->in_step ? 1 : 0;
    L.jt = kJt[ji];
    return L;
}

// Plane-chunk launch (kChunkCi[ch] planes per tile) of a kernel's range and
// block set, per-case entries, with its tuned workgroup rows.
static mhdk::Launch launch_of_chunk(const mhd_ctx *c, int kernel, int ch)
{
    const int ji = c->jt_idx[kernel] % 3;
    mhdk::Launch L = launch_of(c, kernel, ji);
    const TileArr &T = c->ts[set_of(c, kernel)][0].chunks[ch][kKernelRange[kernel]][ji];
    L.tm.block_start = T.d_start;
    L.tm.ntiles = T.n;
    L.tm.xcd_chunk = c->opt.xcd_order ? (T.n + 7) / 8 : 0;
    L.tm.i_inner = 0;
    L.tm.ci = kChunkCi[ch];
    L.tm.kloop = 0;
    return L;
}

// Column-tile launch of a kernel's block set (per-case or geometry-shared
// entries as the kernel's tuned variant says; the variant's row index picks
// the column length kColCi, the column tile's rows are fixed).
static mhdk::Launch launch_of_col(const mhd_ctx *c, int kernel, int jt_index = -1)
{
    const int v = jt_index >= 0 ? jt_index : c->jt_idx[kernel];
    mhdk::Launch L = launch_of(c, kernel, v);
    const TileSet &S = c->ts[set_of(c, kernel)][v >= 6 ? 1 : 0];
    const TileArr &T = S.cols[v % 3];
    L.tm.block_start = T.d_start;
    L.tm.ntiles = T.n;
    L.tm.xcd_chunk = c->opt.xcd_order ? (T.n + 7) / 8 : 0;
    L.tm.i_inner = 0;
    L.tm.ci = kColCi[v % 3];
    L.tm.kloop = 0;
    L.jt = mhdk::col_tile_rows();
    return L;
}

// Shell-point launch of the face EMF (every block, per-case entries).
static mhdk::Launch launch_of_shell(const mhd_ctx *c)
{
    mhdk::Launch L = launch_of(c, KP_SHELL, 0);
    const TileSet &S = c->ts[BS_ALL][0];
    L.tm.block_start = S.shell.d_start;
    L.tm.ntiles = S.shell.n;
    L.tm.xcd_chunk = c->opt.xcd_order ? (S.shell.n + 7) / 8 : 0;
    L.tm.i_inner = 0;
    L.tm.ci = 0;
    L.tm.kloop = 0;
    return L;
}

// The fluid updates: column tiles for the first-order scheme, cell tiles for minmod.
static void launch_fluid_any(mhd_ctx *c, int variant, const double *dts, hipStream_t q)
{
    const int mixed = mixed_of(c);
    if (c->opt.recon == 0)
        mhdk::launch_fluid_col(launch_of_col(c, KP_FLUID, variant), c->d_cst, dts, c->ucur, mixed, q);
    else
        mhdk::launch_fluid(launch_of(c, KP_FLUID, variant), c->d_cst, dts, c->ucur, mixed, c->opt.recon, q);
}
static void launch_fluid2_any(mhd_ctx *c, int variant, const double *dts, hipStream_t q)
{
    const int mixed = mixed_of(c);
    if (c->opt.recon == 0)
        mhdk::launch_fluid2_col(launch_of_col(c, KP_FLUID2, variant), c->d_cst, dts, c->ucur, mixed, q);
    else
        mhdk::launch_fluid2(launch_of(c, KP_FLUID2, variant), c->d_cst, dts, c->ucur, mixed, c->opt.recon, q);
}

// Rebuild the block table, tiles, op lists and per-case constants for the
// current set of active cases.
static int rebuild(mhd_ctx *c)
{
    int rc = ensure_case_capacity(c);
    if (rc)
        return rc;
    HIPCHK(c, hipStreamSynchronize(c->q));
    drop_graph(c);
    c->h_blk.clear();
    c->max_nsub = 0;
    std::vector<MhdCaseConst> cst(c->case_cap);
    std::vector<int> act(c->case_cap, 0);
    for (int ci = 0; ci < (int)c->cases.size(); ++ci)
    {
        CaseRec &r = *c->cases[ci];
        cst[ci] = r.cs.cst;
        if (r.state != 1)
        {
            r.first_block = -1;
            continue;
        }
        act[ci] = 1;
        if (r.cs.cst.model == 1.0 && r.cs.cst.resist != 0.0)
            c->max_nsub = std::max(c->max_nsub, (int)r.cs.cst.resist_nsub);
        r.first_block = (int)c->h_blk.size();
        const GeoSet &G = c->geos[r.geo_set];
        for (int b = 0; b < r.nblocks; ++b)
        {
            MhdBlockDev d = r.cs.blk[b].d;
            d.case_id = ci;
            d.st = (int64_t)(uintptr_t)(r.d_st + (int64_t)b * st_nslot(d.model) * r.slot);
            d.geo = (int64_t)(uintptr_t)(G.d + (int64_t)b * G.block_bytes);
            c->h_blk.push_back(d);
        }
    }
    if ((rc = grow(c, c->d_blk, c->blk_cap, std::max<size_t>(c->h_blk.size(), 1))))
        return rc;
    if (!c->h_blk.empty())
        HIPCHK(c, hipMemcpy(c->d_blk, c->h_blk.data(), sizeof(MhdBlockDev) * c->h_blk.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_cst, cst.data(), sizeof(MhdCaseConst) * c->case_cap, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_active, act.data(), sizeof(int) * c->case_cap, hipMemcpyHostToDevice));
    return build_tiles(c);
}

// ------------------------------------------------------------- placement
static int og_append(mhd_ctx *c, int q, const std::vector<MhdOp> &ops, size_t &off)
{
    OpGroup &G = c->og[q];
    off = G.n;
    if (ops.empty())
        return MHD_OK;
    if (G.n + ops.size() > G.cap)
    {
        const size_t ncap = std::max(G.n + ops.size(), 2 * G.cap);
        MhdOp *np = nullptr;
        HIPCHK(c, hipMalloc((void **)&np, sizeof(MhdOp) * ncap));
        if (G.n)
            HIPCHK(c, hipMemcpy(np, G.d, sizeof(MhdOp) * G.n, hipMemcpyDeviceToDevice));
        free_t(G.d);
        G.d = np;
        G.cap = ncap;
    }
    HIPCHK(c, hipMemcpy(G.d + G.n, ops.data(), sizeof(MhdOp) * ops.size(), hipMemcpyHostToDevice));
    G.n += ops.size();
    return MHD_OK;
}

// Remove [off, off + cnt) from a group: later ops move down in chunks of at
// most cnt, so no chunk overlaps its destination.
static int og_cut(mhd_ctx *c, int q, size_t off, size_t cnt)
{
    OpGroup &G = c->og[q];
    if (cnt == 0)
        return MHD_OK;
    for (size_t s = off + cnt; s < G.n; s += cnt)
    {
        const size_t m = std::min(cnt, G.n - s);
        HIPCHK(c, hipMemcpy(G.d + s - cnt, G.d + s, sizeof(MhdOp) * m, hipMemcpyDeviceToDevice));
    }
    G.n -= cnt;
    for (auto &r : c->cases)
        if (r->state == 1 && r->op_off[q] > off)
            r->op_off[q] -= cnt;
    return MHD_OK;
}

// Undo a partial placement: the case's op segments, its state and its
// reference to the geometry set (freed with the last reference).
static void unplace(mhd_ctx *c, CaseRec &r)
{
    for (int q = 0; q < SY_COUNT * kOpCls; ++q)
        if (r.op_n[q])
        {
            (void)og_cut(c, q, r.op_off[q], r.op_n[q]);
            r.op_n[q] = 0;
        }
    if (r.d_st)
    {
        free_t(r.d_st);
        c->bytes_used -= r.st_bytes;
        r.st_bytes = 0;
    }
    if (r.geo_set >= 0)
    {
        GeoSet &G = c->geos[r.geo_set];
        if (--G.refs == 0)
        {
            free_t(G.d);
            c->bytes_used -= G.bytes;
        }
        r.geo_set = -1;
    }
}

static int place_case_(mhd_ctx *c, CaseRec &r);
static int place_case(mhd_ctx *c, CaseRec &r)
{
    const int rc = place_case_(c, r);
    if (rc && rc != MHD_ERR_MEM)
        unplace(c, r); // a HIP failure part way: nothing of the case stays resident
    return rc;
}

static int place_case_(mhd_ctx *c, CaseRec &r)
{
    const int64_t slot = r.cs.blk[0].d.slot;
    r.slot = slot;
    if (r.cs.shape_key.empty())
        r.cs.shape_key = geometry_key(r.cs.blk);
    int gi = -1;
    for (int g = 0; g < (int)c->geos.size(); ++g)
        if (c->geos[g].refs > 0 && c->geos[g].key == r.cs.shape_key)
            gi = g;
    const int nslot = st_nslot(r.cs.blk[0].d.model);
    const size_t st_bytes = sizeof(double) * (size_t)nslot * slot * r.nblocks;
    const size_t geo_bytes = gi < 0 ? (size_t)geo_block_bytes(slot, weight_esize(c)) * r.nblocks : 0;
    const double budget = c->opt.hbm_budget_gb * 1e9;
    if (budget > 0 && (double)(c->bytes_used + st_bytes + geo_bytes) > budget)
        return fail(c, MHD_ERR_MEM, "HBM budget exceeded");
    if (gi >= 0)
    {
        c->geos[gi].refs += 1;
        r.geo_set = gi;
    }
    else
    {
        GeoSet G;
        G.key = r.cs.shape_key;
        G.block_bytes = geo_block_bytes(slot, weight_esize(c));
        G.bytes = geo_bytes;
        HIPCHK(c, hipMalloc((void **)&G.d, geo_bytes));
        c->bytes_used += geo_bytes;
        G.refs = 1;
        gi = 0;
        while (gi < (int)c->geos.size() && c->geos[gi].refs > 0)
            ++gi;
        if (gi == (int)c->geos.size())
            c->geos.push_back(G);
        else
            c->geos[gi] = G;
        r.geo_set = gi; // registered before the copies: a failed copy is rolled back with it
        std::vector<float> f32;
        const size_t metric = (size_t)G_NMETRIC * slot;
        for (int b = 0; b < r.nblocks; ++b)
        {
            const std::vector<double> &geo = r.cs.blk[b].geo;
            char *dst = c->geos[gi].d + (int64_t)b * c->geos[gi].block_bytes;
            if (mixed_of(c))
            {
                HIPCHK(c, hipMemcpy(dst, geo.data(), sizeof(double) * metric, hipMemcpyHostToDevice));
                f32.assign(geo.begin() + metric, geo.end());
                HIPCHK(c, hipMemcpy(dst + sizeof(double) * metric, f32.data(), sizeof(float) * f32.size(),
                                    hipMemcpyHostToDevice));
            }
            else
                HIPCHK(c, hipMemcpy(dst, geo.data(), sizeof(double) * geo.size(), hipMemcpyHostToDevice));
        }
    }
    HIPCHK(c, hipMalloc((void **)&r.d_st, st_bytes));
    r.st_bytes = st_bytes;
    c->bytes_used += st_bytes;
    for (int b = 0; b < r.nblocks; ++b)
    {
        double *dst = r.d_st + (int64_t)b * nslot * slot;
        HIPCHK(c, hipMemcpy(dst, r.cs.blk[b].st.data(), sizeof(double) * nslot * slot, hipMemcpyHostToDevice));
        if (c->ucur == 1) // admitted mid-run: the current U buffer is the second one
        {
            HIPCHK(c, hipMemcpy(dst + ST_U1 * slot, dst + ST_U0 * slot, sizeof(double) * 5 * slot,
                                hipMemcpyDeviceToDevice));
            if (nslot == ST_NSLOT2)
                HIPCHK(c, hipMemcpy(dst + ST_UNA1 * slot, dst + ST_UNA0 * slot, sizeof(double) * 5 * slot,
                                    hipMemcpyDeviceToDevice));
        }
        std::vector<double>().swap(r.cs.blk[b].geo);
        std::vector<double>().swap(r.cs.blk[b].wj);
        std::vector<double>().swap(r.cs.blk[b].wb);
        std::vector<double>().swap(r.cs.blk[b].st);
    }
    // op lists: case-relative offsets -> absolute addresses, appended to the
    // device op groups
    for (int sy = 0; sy < SY_COUNT; ++sy)
    {
        std::vector<MhdOp> ops = sy < (int)r.cs.ops.size() ? std::move(r.cs.ops[sy]) : std::vector<MhdOp>();
        for (MhdOp &o : ops)
        {
            o.dst = (int64_t)(uintptr_t)(r.d_st + o.dst);
            if (o.kind == OP_COPY || o.kind == OP_NEG || o.kind == OP_SOLID_FIRST || o.kind == OP_SOLID_GHOST)
                o.src = (int64_t)(uintptr_t)(r.d_st + o.src);
            else
                o.src = o.dst;
            o.cas = r.id;
        }
        std::vector<MhdOp> by_cls[kOpCls];
        for (const MhdOp &o : ops)
            by_cls[op_cls(o)].push_back(o);
        for (int cl = 0; cl < kOpCls; ++cl)
        {
            const int q = sy * kOpCls + cl;
            int rc = og_append(c, q, by_cls[cl], r.op_off[q]);
            if (rc)
                return rc;
            r.op_n[q] = by_cls[cl].size();
        }
        // the host copy is consumed: a failure after this point cannot be retried
        r.cs.ops[sy].clear();
    }
    r.cs.ops.clear();
    r.state = 1;
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

// `folded`: also the records the step's kernels apply themselves.
static void add_sync(const mhd_ctx *c, mhdk::OpLaunch &L, int sy, int ushift, bool folded)
{
    for (int cl = 0; cl < kOpCls; ++cl)
    {
        const OpGroup &G = c->og[sy * kOpCls + cl];
        if (G.n && (folded || cl < 2))
            L.seg[L.nseg++] = {G.d, (int)G.n, (cl & 1) ? 6 : 0, ushift};
    }
}
static bool has_ops(const mhd_ctx *c, int sy, bool folded)
{
    for (int cl = 0; cl < (folded ? kOpCls : 2); ++cl)
        if (c->og[sy * kOpCls + cl].n)
            return true;
    return false;
}
static void run_ops(mhd_ctx *c, int sy, int ushift, bool folded = true)
{
    mhdk::OpLaunch L{};
    add_sync(c, L, sy, ushift, folded);
    mhdk::launch_ops(L, c->d_cst, c->q);
}
// Two op lists in one launch: only for lists with disjoint destinations of
// which neither reads what the other writes (U-slot and B-slot lists, the
// B_cell and J_cell ghost lists).
static void run_ops2(mhd_ctx *c, int sya, int ushift_a, int syb, int ushift_b, bool folded = true)
{
    mhdk::OpLaunch L{};
    add_sync(c, L, sya, ushift_a, folded);
    add_sync(c, L, syb, ushift_b, folded);
    mhdk::launch_ops(L, c->d_cst, c->q);
}

// The U and B syncs of a step: two chains of dependent lists (U wall first
// layer, U wall ghosts, U halo; B wall, B halo), the k-face work the kernels
// folded left out (U's only with the first-order fluid kernel); the lists
// left in each chain go out in order, the two chains side by side, so a
// shell whose only walls are k faces needs one launch.
static void run_ub_chains(mhd_ctx *c, bool fold_u)
{
    const int uc[3] = {SY_U_BC0, SY_U_BC1, SY_U_HALO}, bc[2] = {SY_B_BC, SY_B_HALO};
    int u[3], nu = 0, bb[2], nbb = 0;
    for (int q : uc)
        if (has_ops(c, q, !fold_u))
            u[nu++] = q;
    for (int q : bc)
        if (has_ops(c, q, false))
            bb[nbb++] = q;
    for (int t = 0; t < std::max(nu, nbb); ++t)
    {
        mhdk::OpLaunch L{};
        if (t < nu)
            add_sync(c, L, u[t], c->ucur, !fold_u);
        if (t < nbb)
            add_sync(c, L, bb[t], 0, false);
        mhdk::launch_ops(L, c->d_cst, c->q);
    }
}

// Derived fields (contravariant velocities, B_cell and the next dt, then
// J_cell); the J_cell wall condition (the SY_JCELL_BC op list of the host
// set-up) is applied inside the J_cell kernel.
static bool runs(const mhd_ctx *c, int kernel) { return c->ts[set_of(c, kernel)][0].nent > 0; }

// Branch b of a step on the second stream: it starts after everything
// enqueued on the context stream so far; join() makes the context stream wait
// for it.  Without overlap (or while profiling, which times kernels on the
// context stream) both run on the context stream in order.
static bool overlapped(const mhd_ctx *c) { return c->opt.overlap && !c->prof; }
static hipStream_t fork_branch(mhd_ctx *c, int b)
{
    if (!overlapped(c))
        return c->q;
    (void)hipEventRecord(c->fork[b], c->q);
    (void)hipStreamWaitEvent(c->q2, c->fork[b], 0);
    return c->q2;
}
static void join_branch(mhd_ctx *c, int b)
{
    if (!overlapped(c))
        return;
    (void)hipEventRecord(c->join[b], c->q2);
    (void)hipStreamWaitEvent(c->q, c->join[b], 0);
}

static void derived(mhd_ctx *c)
{
    {
        // J_cell (with its J edges) reads only B (synced) and the geometry: beside B_cell / U_plus / dt
        const hipStream_t qb = fork_branch(c, 1);
        ProfScope p(c, KP_JCELL);
        mhdk::launch_jcell(launch_of(c, KP_JCELL), mixed_of(c), qb);
    }
    if (runs(c, KP_DERIVED))
    {
        ProfScope p(c, KP_DERIVED);
        mhdk::launch_derived(launch_of(c, KP_DERIVED), c->d_cst, c->d_dtbits, c->ucur, mixed_of(c), c->q);
    }
    if (runs(c, KP_DERIVED2))
    {
        ProfScope p(c, KP_DERIVED2);
        mhdk::launch_derived2(launch_of(c, KP_DERIVED2), c->d_cst, c->d_dtbits, c->ucur, mixed_of(c), c->q);
    }
    join_branch(c, 1);
    // B_cell and J_cell ghosts of the two-fluid cases with the Hall EMF
    // (the reference's B_cell and J_cell syncs; empty for the other cases)
    ProfScope p(c, KP_OPS);
    run_ops2(c, SY_BCELL_BC, 0, SY_JCELL_GH, 0);
    run_ops2(c, SY_BCELL_HALO, 0, SY_JCELL_HALO, 0);
}

// The stages of one StepOnce (Z4_Lunar/src/LunarSolver_Driver.cpp:21-92)
// over every active case, in the reference's order.  enqueue_step runs them
// back to back; the per-stage entry points of the C-ABI run them one at a time.

// Compute_Timestep: the min the previous derived pass accumulated becomes
// each case's dt (LunarSolver_Timestep.cpp:8-140, abort check :128-139).
static void stage_dt(mhd_ctx *c)
{
    ProfScope p(c, KP_DTFIN);
    mhdk::launch_dt_finalize(c->d_dtbits, c->d_dt, c->d_err, c->d_cst, c->case_cap, c->d_active, c->q);
}

// Scheme_U_ + AddSourceToRHS_Fluid + the U part of ApplyUpdate_Euler_
// (LunarSolver_FluidRHS.cpp:3-201,528-568, LunarSolver_Advance.cpp:61-122):
// reads U[cur], B_cell, J_cell, p_e and dt, writes U + dt RHS into U[1-cur].
static void stage_fluid(mhd_ctx *c, hipStream_t qf)
{
    if (runs(c, KP_FLUID))
    {
        ProfScope p(c, KP_FLUID);
        launch_fluid_any(c, -1, c->d_dt, qf);
    }
    if (runs(c, KP_FLUID2))
    {
        ProfScope p(c, KP_FLUID2);
        launch_fluid2_any(c, -1, c->d_dt, qf);
    }
}

// AssembleOneDirectionEMF_ for the three face families
// (LunarSolver_MagExplicit_Ideal.cpp:348-496), after the cube-vertex cells
// took the contravariant velocity of their source cells.
static void stage_uc_corner(mhd_ctx *c)
{
    ProfScope p(c, KP_OPS);
    run_ops(c, SY_UC_CORNER, 0);
}
static void stage_face_emf(mhd_ctx *c, bool corner = true)
{
    if (corner)
        stage_uc_corner(c);
    ProfScope p(c, KP_FACE);
    mhdk::launch_face_emf(launch_of(c, KP_FACE), mixed_of(c), c->q);
}

// Sync("Eface"): the wall handler, then the ghost candidates the edge EMF reads.
static void stage_sync_eface(mhd_ctx *c)
{
    ProfScope p(c, KP_OPS);
    run_ops(c, SY_EFACE_BC, 0);
    run_ops(c, SY_EFACE_HALO, 0);
}

// AssembleEdgeEMF_FromFaceE_Ideal_ (LunarSolver_MagExplicit_Ideal.cpp:53-287),
// then the two-fluid EMF terms added before the Eedge sync
// (Z5_Mercury/src/physics/MercurySolver_MagneticRHS.cpp:30-33).
static void stage_edge_emf(mhd_ctx *c)
{
    {
        ProfScope p(c, KP_EDGE);
        mhdk::launch_edge_emf(launch_of(c, KP_EDGE), c->q);
    }
    if (runs(c, KP_AMBI))
    {
        ProfScope p(c, KP_AMBI);
        mhdk::launch_emf_extra(launch_of(c, KP_AMBI), c->d_cst, c->ucur, mixed_of(c), c->q);
    }
}

// The face and edge EMF of a whole step in one pass: the shell's face
// candidates (zero on a wall's inner slab), the halo copy of the Eface sync,
// then k_emf, which forms every in-range candidate itself and takes the wall
// handler's zeros inline; then the two-fluid EMF terms.  E equals the three
// stages' bit for bit.
static void stage_emf(mhd_ctx *c)
{
    {
        ProfScope p(c, KP_SHELL);
        mhdk::launch_face_shell(launch_of_shell(c), c->q);
    }
    {
        ProfScope p(c, KP_OPS);
        run_ops(c, SY_EFACE_HALO, 0);
    }
    {
        ProfScope p(c, KP_EMF);
        mhdk::launch_emf(launch_of(c, KP_EMF), c->q);
    }
    if (runs(c, KP_AMBI))
    {
        ProfScope p(c, KP_AMBI);
        mhdk::launch_emf_extra(launch_of(c, KP_AMBI), c->d_cst, c->ucur, mixed_of(c), c->q);
    }
}

// Sync("Eedge"): the far-field edge handler.
static void stage_sync_eedge(mhd_ctx *c)
{
    ProfScope p(c, KP_OPS);
    run_ops(c, SY_EEDGE_BC, 0);
}

// CurlEdgeToFace(-1) + the B part of ApplyUpdate_Euler_ (CTOperators.cpp:6-56).
static void stage_curl(mhd_ctx *c)
{
    ProfScope p(c, KP_CURL);
    mhdk::launch_curl_update(launch_of(c, KP_CURL), c->d_dt, c->q);
}

// Interior resistivity of the two-fluid cases: explicit subcycles on B
// (ResistiveDiffusionSubcycles_, Z5_Mercury/src/physics/MercurySolver_MagneticRHS.cpp:79-141).
static void stage_resistive(mhd_ctx *c)
{
    const int mixed = mixed_of(c);
    for (int isub = 0; isub < c->max_nsub; ++isub)
    {
        mhdk::launch_dt_res(c->d_dtres, c->d_dt, c->d_cst, c->case_cap, isub, c->q);
        {
            ProfScope p(c, KP_OPS);
            run_ops(c, SY_B_BC, 0, !c->in_step); // in a step the curl kernel has copied the k-face wall ghosts
            run_ops(c, SY_B_HALO, 0);
        }
        {
            ProfScope p(c, KP_JCELL_RES);
            mhdk::launch_jcell(launch_of(c, KP_JCELL_RES), mixed, c->q);
        }
        {
            ProfScope p(c, KP_OPS);
            run_ops(c, SY_JCELL_GH, 0);
            run_ops(c, SY_JCELL_HALO, 0);
        }
        {
            ProfScope p(c, KP_ERES);
            mhdk::launch_eres(launch_of(c, KP_ERES), c->d_cst, c->q);
        }
        {
            ProfScope p(c, KP_OPS);
            run_ops(c, SY_EEDGE_BC, 0); // Eres far field: the Eedge handler (Boundary_Build.cpp:200-205)
        }
        {
            ProfScope p(c, KP_CURL_RES);
            mhdk::launch_curl_update(launch_of(c, KP_CURL_RES), c->d_dtres, c->q);
        }
        ProfScope p(c, KP_OPS);
        run_ops(c, SY_B_BC, 0, !c->in_step);
        run_ops(c, SY_B_HALO, 0);
    }
}

// One StepOnce for every active case.  dt and the contravariant velocities
// were computed by the previous step's derived-field kernel (or at
// placement).
static void enqueue_step(mhd_ctx *c)
{
    c->in_step = true;
    stage_dt(c);
    // The corner velocities the face EMF reads go first: launched after the
    // fluid kernel they wait for its workgroups to drain (77 us, on the EMF
    // chain's critical path, instead of 6).
    stage_uc_corner(c);
    // The fluid update reads U[cur], B_cell, J_cell, p_e and dt and writes
    // U[1-cur]; the EMF / curl / resistive chain reads U[cur] and writes
    // Eface, E, B faces, J_cell (subcycles) -- J_cell only after the fluid
    // update has read it, so the subcycles wait for the fluid branch.
    stage_fluid(c, fork_branch(c, 0));
    if (c->emf_fused)
        stage_emf(c);
    else
    {
        stage_face_emf(c, false);
        stage_sync_eface(c);
        stage_edge_emf(c);
    }
    stage_sync_eedge(c);
    stage_curl(c);
    join_branch(c, 0);
    stage_resistive(c);
    c->ucur ^= 1;
    {
        // U and B syncs: each field physical -> halo (the U wall ghosts after
        // the wall's first layer); the U-slot and B-slot lists share launches
        ProfScope p(c, KP_OPS);
        run_ub_chains(c, c->opt.recon == 0);
    }
    derived(c);
    c->in_step = false;
}

// Autotuner: per kernel, time each workgroup height (64x1, 64x2, 64x4) on the
// placed cases and keep the fastest.  Extra launches must not change the
// state: face/edge EMF and the fluid kernel write scratch the next step
// overwrites, the curl update runs with dt = 0, the derived kernel recomputes
// the same values.
static int autotune(mhd_ctx *c)
{
    if (c->h_blk.empty())
        return MHD_OK;
    const int mixed = c->opt.precision == MHD_PREC_MIXED;
    const int kernels[] = {KP_FACE,  KP_EDGE,     KP_FLUID, KP_CURL, KP_DERIVED,
                           KP_JCELL, KP_FLUID2, KP_DERIVED2, KP_AMBI, KP_EMF};
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    for (int k : kernels)
    {
        if (!runs(c, k) || (k == KP_EMF && !c->emf_fused))
            continue;
        double best = 1e30;
        int best_j = c->jt_idx[k];
        const int nv = loops_members(k) ? kLoopVariants : kVariants;
        for (int ji = 0; ji < nv; ++ji)
        {
            auto launch = [&]() {
                const mhdk::Launch L = launch_of(c, k, ji);
                switch (k)
                {
                case KP_FACE:
                    mhdk::launch_face_emf(L, mixed, c->q);
                    break;
                case KP_EDGE:
                    mhdk::launch_edge_emf(L, c->q);
                    break;
                case KP_EMF:
                    mhdk::launch_emf(L, c->q); // writes E only, step scratch
                    break;
                case KP_FLUID:
                    launch_fluid_any(c, ji, c->d_zero, c->q);
                    break;
                case KP_CURL:
                    mhdk::launch_curl_update(L, c->d_zero, c->q);
                    break;
                case KP_JCELL:
                    mhdk::launch_jcell(L, mixed, c->q);
                    break;
                case KP_FLUID2:
                    launch_fluid2_any(c, ji, c->d_zero, c->q);
                    break;
                case KP_DERIVED2:
                    mhdk::launch_derived2(L, c->d_cst, c->d_dtbits, c->ucur, mixed, c->q);
                    break;
                case KP_AMBI:
                    // E is step scratch: the next step's edge EMF rewrites it
                    mhdk::launch_emf_extra(L, c->d_cst, c->ucur, mixed, c->q);
                    break;
                default:
                    mhdk::launch_derived(L, c->d_cst, c->d_dtbits, c->ucur, mixed, c->q);
                    break;
                }
            };
            launch(); // warm
            HIPCHK(c, hipEventRecord(e0, c->q));
            for (int rep = 0; rep < 3; ++rep)
                launch();
            HIPCHK(c, hipEventRecord(e1, c->q));
            HIPCHK(c, hipEventSynchronize(e1));
            float ms = 0;
            HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
            if (ms < best)
            {
                best = ms;
                best_j = ji;
            }
        }
        c->jt_idx[k] = best_j;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return MHD_OK;
}

// Live admission of the pending case `id` into the running context: place,
// rebuild the tables, compute the new case's derived fields, dt and
// contravariant velocities (the derived kernel is idempotent for the cases
// already running).  *placed tells whether anything of it reached the device.
static int admit(mhd_ctx *c, int id, bool *placed)
{
    *placed = false;
    if (c->fluid_pending)
        return fail(c, MHD_ERR_STATE, "a step is in progress (mhd_fluid_rhs without mhd_apply_update)");
    HIPCHK(c, hipSetDevice(c->opt.device));
    int rc = ensure_case_capacity(c);
    if (rc)
        return rc;
    if ((rc = place_case(c, *c->cases[id])))
        return rc; // rolled back by place_case
    *placed = true;
    if ((rc = rebuild(c)))
        return rc;
    derived(c);
    HIPCHK(c, hipGetLastError());
    return MHD_OK;
}

static int add_case(mhd_ctx *c, CaseSetup &&cs, int *case_id)
{
    auto r = std::make_unique<CaseRec>();
    r->cs = std::move(cs);
    r->nblocks = (int)r->cs.blk.size();
    const int id = (int)c->cases.size();
    r->id = id;
    c->cases.push_back(std::move(r));
    if (case_id)
        *case_id = id;
    if (!c->finalized)
        return MHD_OK;
    bool placed = false;
    const int rc = admit(c, id, &placed);
    if (rc && !placed)
        c->cases.pop_back(); // nothing was placed: the id is free again
    else if (rc)
        c->cases[id]->state = 2;
    return rc;
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
    if (!c || !wind_desc_ok(d))
        return MHD_ERR_ARG;
    try
    {
        return add_case(c,
                        make_case_lunar(d->nang, d->nrad, d->r_in, d->r_out, c->opt.ghost, d->gamma, d->cfl,
                                        wind_params(*d)),
                        case_id);
    }
    catch (const std::exception &e)
    {
        return fail(c, MHD_ERR_ARG, e.what());
    }
}

extern "C" int mhd_add_case_mercury(mhd_ctx *c, const mhd_wind_desc *d, const mhd_dipole_desc *dp, int *case_id)
{
    if (!c || !wind_desc_ok(d) || !dp || dp->softening < 0)
        return MHD_ERR_ARG;
    try
    {
        return add_case(c,
                        make_case_mercury(d->nang, d->nrad, d->r_in, d->r_out, c->opt.ghost, d->gamma, d->cfl,
                                          wind_params(*d), dipole_params(*dp)),
                        case_id);
    }
    catch (const std::exception &e)
    {
        return fail(c, MHD_ERR_ARG, e.what());
    }
}

// Bytes a case would add to the context (its geometry only when no resident
// case shares the grid shape).
static size_t case_bytes(const mhd_ctx *c, const CaseSetup &cs)
{
    const size_t slot = (size_t)cs.blk[0].d.slot, nb = cs.blk.size();
    size_t need = sizeof(double) * st_nslot(cs.blk[0].d.model) * slot * nb;
    bool shared = false;
    for (const GeoSet &G : c->geos)
        shared = shared || (G.refs > 0 && G.key == cs.shape_key);
    if (!shared)
        need += (size_t)geo_block_bytes((int64_t)slot, weight_esize(c)) * nb;
    return need;
}

// A host case whose full weight tables were replaced: the device forms the
// weights from the least-squares data, so the tables must be the ones that
// data gives (else the edit would be dropped silently).
static int check_edited_tables(mhd_ctx *c, const CaseSetup &cs)
{
    if (!cs.tables_edited)
        return MHD_OK;
    for (size_t b = 0; b < cs.blk.size(); ++b)
    {
        const BlockArrays &A = cs.blk[b];
        std::vector<double> wb(aos_count(A.d, LOC_CELL, 18)), wj(aos_count(A.d, LOC_CELL, 36));
        soa_aos(A.d, LOC_CELL, 18, const_cast<double *>(A.wb.data()), wb.data(), true);
        soa_aos(A.d, LOC_CELL, 36, const_cast<double *>(A.wj.data()), wj.data(), true);
        std::string why;
        if (!weight_tables_match(A, wb.data(), wj.data(), 1e-12, why))
            return fail(c, MHD_ERR_ARG,
                        "block " + std::to_string(b) + ": " + why +
                            "; the device forms the weight tables from the grid's least-squares data, so edited "
                            "tables cannot be adopted");
    }
    return MHD_OK;
}

extern "C" int mhd_adopt_host_case(mhd_ctx *c, mhd_host_case *h, int *case_id)
{
    if (!c || !h || h->cs.blk.empty())
        return MHD_ERR_ARG;
    if (h->cs.blk[0].d.g != c->opt.ghost)
        return fail(c, MHD_ERR_ARG, "host case built with another ghost width");
    if (h->cs.shape_key.empty())
        h->cs.shape_key = geometry_key(h->cs.blk);
    if (int rc = check_edited_tables(c, h->cs))
        return rc;
    const double budget = c->opt.hbm_budget_gb * 1e9;
    if (c->finalized && budget > 0 && (double)(c->bytes_used + case_bytes(c, h->cs)) > budget)
        return fail(c, MHD_ERR_MEM, "HBM budget exceeded"); // the host case stays intact
    return add_case(c, std::move(h->cs), case_id);
}

extern "C" int mhd_finalize(mhd_ctx *c)
{
    if (!c)
        return MHD_ERR_ARG;
    if (c->finalized)
        return fail(c, MHD_ERR_STATE, "already finalized");
    HIPCHK(c, hipSetDevice(c->opt.device));
    int rc = ensure_case_capacity(c);
    if (rc)
        return rc;
    for (auto &r : c->cases)
        if (r->state == 0 && (rc = place_case(c, *r)))
            return rc;
    if ((rc = rebuild(c)))
        return rc;
    c->ucur = 0;
    c->finalized = true;
    derived(c);
    if (c->opt.autotune && (rc = autotune(c)))
        return rc;
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->q));
    return MHD_OK;
}

extern "C" int mhd_retire(mhd_ctx *c, int case_id)
{
    if (!c || case_id < 0 || case_id >= (int)c->cases.size())
        return MHD_ERR_ARG;
    if (c->fluid_pending)
        return fail(c, MHD_ERR_STATE, "a staged step is open: mhd_apply_update first");
    CaseRec &r = *c->cases[case_id];
    if (r.state != 1)
        return fail(c, MHD_ERR_STATE, "case is not active");
    HIPCHK(c, hipStreamSynchronize(c->q));
    free_t(r.d_st);
    c->bytes_used -= r.st_bytes;
    GeoSet &G = c->geos[r.geo_set];
    if (--G.refs == 0)
    {
        free_t(G.d);
        c->bytes_used -= G.bytes;
    }
    r.state = 2;
    for (int q = 0; q < SY_COUNT * kOpCls; ++q)
    {
        int rc = og_cut(c, q, r.op_off[q], r.op_n[q]);
        if (rc)
            return rc;
        r.op_n[q] = 0;
    }
    const int zero = 0;
    HIPCHK(c, hipMemcpy(c->d_err + case_id, &zero, sizeof(int), hipMemcpyHostToDevice));
    return rebuild(c);
}

extern "C" int mhd_case_error(mhd_ctx *c, int case_id, int *dt_error)
{
    if (!c || !dt_error || case_id < 0 || case_id >= (int)c->cases.size() || !c->finalized)
        return MHD_ERR_ARG;
    int e = 0;
    if (c->cases[case_id]->state == 1)
    {
        HIPCHK(c, hipMemcpyAsync(&e, c->d_err + case_id, sizeof(int), hipMemcpyDeviceToHost, c->q));
        HIPCHK(c, hipStreamSynchronize(c->q));
    }
    *dt_error = e;
    return MHD_OK;
}

// ------------------------------------------------------------- stepping
static int check_err(mhd_ctx *c)
{
    std::vector<int> e(c->case_cap, 0);
    HIPCHK(c, hipMemcpyAsync(e.data(), c->d_err, sizeof(int) * c->case_cap, hipMemcpyDeviceToHost, c->q));
    HIPCHK(c, hipStreamSynchronize(c->q));
    for (int ci = 0; ci < (int)c->cases.size(); ++ci)
        if (e[ci] && c->cases[ci]->state == 1)
            return fail(c, MHD_ERR_DT, "dt too small or NaN in case " + std::to_string(ci));
    return MHD_OK;
}

static void count_steps(mhd_ctx *c, int n)
{
    for (auto &r : c->cases)
        if (r->state == 1)
            r->steps += n;
}

extern "C" int mhd_step(mhd_ctx *c, int nsteps)
{
    if (!c || nsteps < 0)
        return MHD_ERR_ARG;
    if (!c->finalized)
        return fail(c, MHD_ERR_STATE, "mhd_finalize first");
    if (c->fluid_pending)
        return fail(c, MHD_ERR_STATE, "a staged step is open: mhd_apply_update first");
    if (c->h_blk.empty())
        return MHD_OK;
    HIPCHK(c, hipSetDevice(c->opt.device));
    for (int s = 0; s < nsteps; ++s)
        enqueue_step(c);
    HIPCHK(c, hipGetLastError());
    count_steps(c, nsteps);
    return check_err(c);
}

constexpr int kGraphShort = 2, kGraphLong = 10;
extern "C" int mhd_step_graph(mhd_ctx *c, int nsteps)
{
    if (!c || nsteps < 0)
        return MHD_ERR_ARG;
    if (!c->finalized)
        return fail(c, MHD_ERR_STATE, "mhd_finalize first");
    if (c->fluid_pending)
        return fail(c, MHD_ERR_STATE, "a staged step is open: mhd_apply_update first");
    if (c->h_blk.empty())
        return MHD_OK;
    if (c->prof)
        return mhd_step(c, nsteps);
    HIPCHK(c, hipSetDevice(c->opt.device));
    int s = 0;
    if (c->ucur != 0 && nsteps > 0)
    {
        enqueue_step(c);
        ++s;
    }
    // long runs replay kGraphLong steps per launch (each graph launch costs
    // about the gap between two steps), the rest two at a time
    auto capture = [&](hipGraphExec_t &ge, int n) -> int {
        if (ge)
            return MHD_OK;
        hipGraph_t g;
        HIPCHK(c, hipStreamBeginCapture(c->q, hipStreamCaptureModeThreadLocal));
        for (int q = 0; q < n; ++q)
            enqueue_step(c);
        HIPCHK(c, hipStreamEndCapture(c->q, &g));
        HIPCHK(c, hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
        (void)hipGraphDestroy(g);
        return MHD_OK;
    };
    if (nsteps - s >= 2 * kGraphLong)
    {
        if (int rc = capture(c->gexec_long, kGraphLong))
            return rc;
        for (; s + kGraphLong <= nsteps; s += kGraphLong)
            HIPCHK(c, hipGraphLaunch(c->gexec_long, c->q));
    }
    if (nsteps - s >= kGraphShort)
    {
        if (int rc = capture(c->gexec, kGraphShort))
            return rc;
        for (; s + kGraphShort <= nsteps; s += kGraphShort)
            HIPCHK(c, hipGraphLaunch(c->gexec, c->q));
    }
    for (; s < nsteps; ++s)
        enqueue_step(c);
    HIPCHK(c, hipGetLastError());
    count_steps(c, nsteps);
    return check_err(c);
}

extern "C" int mhd_synchronize(mhd_ctx *c)
{
    if (!c)
        return MHD_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->q));
    return MHD_OK;
}

// ------------------------------------------------- per-stage entry points
// The stages of mhd_step one call at a time, on the context stream, over
// every active case (the fluid update then runs on the context stream too).
static int stage_enter(mhd_ctx *c)
{
    if (!c)
        return MHD_ERR_ARG;
    if (!c->finalized)
        return fail(c, MHD_ERR_STATE, "mhd_finalize first");
    HIPCHK(c, hipSetDevice(c->opt.device));
    return MHD_OK;
}
static int stage_leave(mhd_ctx *c)
{
    HIPCHK(c, hipGetLastError());
    return MHD_OK;
}

extern "C" int mhd_step_dt(mhd_ctx *c)
{
    if (int rc = stage_enter(c))
        return rc;
    if (c->fluid_pending)
        return fail(c, MHD_ERR_STATE, "mhd_step_dt inside an open step");
    if (c->h_blk.empty())
        return MHD_OK;
    stage_dt(c);
    if (int rc = stage_leave(c))
        return rc;
    return check_err(c);
}

extern "C" int mhd_fluid_rhs(mhd_ctx *c)
{
    if (int rc = stage_enter(c))
        return rc;
    if (c->fluid_pending)
        return fail(c, MHD_ERR_STATE, "mhd_fluid_rhs twice in one step (mhd_apply_update ends a step)");
    stage_fluid(c, c->q);
    c->fluid_pending = true;
    return stage_leave(c);
}

extern "C" int mhd_face_emf(mhd_ctx *c)
{
    if (int rc = stage_enter(c))
        return rc;
    stage_face_emf(c);
    return stage_leave(c);
}

extern "C" int mhd_edge_emf(mhd_ctx *c)
{
    if (int rc = stage_enter(c))
        return rc;
    stage_edge_emf(c);
    return stage_leave(c);
}

extern "C" int mhd_curl_update(mhd_ctx *c)
{
    if (int rc = stage_enter(c))
        return rc;
    stage_curl(c);
    return stage_leave(c);
}

extern "C" int mhd_apply_update(mhd_ctx *c)
{
    if (int rc = stage_enter(c))
        return rc;
    if (!c->fluid_pending)
        return fail(c, MHD_ERR_STATE, "mhd_apply_update without mhd_fluid_rhs in this step");
    stage_resistive(c);
    c->ucur ^= 1;
    c->fluid_pending = false;
    count_steps(c, 1);
    return stage_leave(c);
}

extern "C" int mhd_sync(mhd_ctx *c, const char *group)
{
    if (int rc = stage_enter(c))
        return rc;
    if (!group)
        return MHD_ERR_ARG;
    const std::string g(group);
    if (g == "Eface")
        stage_sync_eface(c);
    else if (g == "Eedge")
        stage_sync_eedge(c);
    else if (g == "Ucell")
    {
        run_ops(c, SY_U_BC0, c->ucur);
        run_ops(c, SY_U_BC1, c->ucur);
        run_ops(c, SY_U_HALO, c->ucur);
    }
    else if (g == "Bface")
    {
        run_ops(c, SY_B_BC, 0);
        run_ops(c, SY_B_HALO, 0);
    }
    else if (g == "B_cell")
    {
        run_ops(c, SY_BCELL_BC, 0);
        run_ops(c, SY_BCELL_HALO, 0);
    }
    else if (g == "J_cell")
    {
        run_ops(c, SY_JCELL_GH, 0);
        run_ops(c, SY_JCELL_HALO, 0);
    }
    else if (g != "Jedge" && g != "Uplus" && g != "Ehall" && g != "Badd")
        return fail(c, MHD_ERR_ARG, "unknown sync group " + g);
    return stage_leave(c);
}

extern "C" int mhd_derived(mhd_ctx *c)
{
    if (int rc = stage_enter(c))
        return rc;
    if (c->fluid_pending)
        return fail(c, MHD_ERR_STATE, "mhd_derived inside an open step: mhd_apply_update first");
    derived(c);
    return stage_leave(c);
}

// --------------------------------------- a case from the caller's arrays
static CaseRec *draft_of(mhd_ctx *c, int case_id)
{
    if (!c || case_id < 0 || case_id >= (int)c->cases.size())
        return nullptr;
    CaseRec *r = c->cases[case_id].get();
    return r->state == 3 && r->draft ? r : nullptr;
}

extern "C" int mhd_case_begin(mhd_ctx *c, int nblocks, const double *consts22, int *case_id)
{
    if (!c || !case_id)
        return MHD_ERR_ARG;
    mhd_host_case *h = nullptr;
    const int rc = mhd_host_case_begin(nblocks, c->opt.ghost, consts22, &h);
    if (rc)
        return fail(c, rc, "mhd_case_begin: bad block count or constants");
    auto r = std::make_unique<CaseRec>();
    r->id = (int)c->cases.size();
    r->state = 3;
    r->draft = h;
    r->nblocks = nblocks;
    *case_id = r->id;
    c->cases.push_back(std::move(r));
    return MHD_OK;
}

extern "C" int mhd_add_block(mhd_ctx *c, int case_id, int ib, int ni, int nj, int nk, const double *x, const double *y,
                             const double *z, size_t nnode, const double *const *geometry)
{
    CaseRec *r = draft_of(c, case_id);
    if (!r)
        return c ? fail(c, MHD_ERR_STATE, "mhd_add_block: not a case begun with mhd_case_begin") : MHD_ERR_ARG;
    const int rc = mhd_host_add_block(r->draft, ib, ni, nj, nk, x, y, z, nnode, geometry);
    return rc ? fail(c, rc, r->draft->err) : MHD_OK;
}

extern "C" int mhd_set_topology(mhd_ctx *c, int case_id, int ib, const int *faces14x6)
{
    CaseRec *r = draft_of(c, case_id);
    if (!r)
        return c ? fail(c, MHD_ERR_STATE, "mhd_set_topology: not a case begun with mhd_case_begin") : MHD_ERR_ARG;
    const int rc = mhd_host_set_topology(r->draft, ib, faces14x6);
    return rc ? fail(c, rc, r->draft->err) : MHD_OK;
}

extern "C" int mhd_set_face_parts(mhd_ctx *c, int case_id, int ib, int face, int nparts, const int *parts16)
{
    CaseRec *r = draft_of(c, case_id);
    if (!r)
        return c ? fail(c, MHD_ERR_STATE, "mhd_set_face_parts: not a case begun with mhd_case_begin") : MHD_ERR_ARG;
    const int rc = mhd_host_set_face_parts(r->draft, ib, face, nparts, parts16);
    return rc ? fail(c, rc, r->draft->err) : MHD_OK;
}

extern "C" int mhd_case_end(mhd_ctx *c, int case_id)
{
    CaseRec *r = draft_of(c, case_id);
    if (!r)
        return c ? fail(c, MHD_ERR_STATE, "mhd_case_end: not a case begun with mhd_case_begin") : MHD_ERR_ARG;
    int rc = mhd_host_case_end(r->draft);
    if (rc)
        return fail(c, rc, r->draft->err); // the draft stays: blocks or topology may be given again
    r->cs = std::move(r->draft->cs);
    mhd_host_case_destroy(r->draft);
    r->draft = nullptr;
    r->nblocks = (int)r->cs.blk.size();
    r->state = 0;
    if (!c->finalized)
        return MHD_OK;
    bool placed = false;
    rc = admit(c, case_id, &placed);
    if (rc)
        r->state = 2; // not admitted (e.g. MHD_ERR_MEM); the id is not reused
    return rc;
}

static CaseRec *active_case(mhd_ctx *c, int case_id)
{
    if (!c || case_id < 0 || case_id >= (int)c->cases.size() || !c->finalized)
        return nullptr;
    CaseRec *r = c->cases[case_id].get();
    return r->state == 1 ? r : nullptr;
}

extern "C" int mhd_compute_dt(mhd_ctx *c, int case_id, double *dt)
{
    if (!dt || !active_case(c, case_id))
        return MHD_ERR_ARG;
    mhdk::launch_fill_u64(c->d_dtscratch, c->case_cap, dbits(1e100), c->q);
    mhdk::launch_dt(launch_of(c, KP_DT), c->d_cst, c->d_dtscratch, c->ucur, mixed_of(c), c->q);
    unsigned long long bits = 0;
    HIPCHK(c, hipMemcpyAsync(&bits, c->d_dtscratch + case_id, 8, hipMemcpyDeviceToHost, c->q));
    HIPCHK(c, hipStreamSynchronize(c->q));
    std::memcpy(dt, &bits, 8);
    return MHD_OK;
}

extern "C" int mhd_last_dt(mhd_ctx *c, int case_id, double *dt)
{
    if (!dt || !active_case(c, case_id))
        return MHD_ERR_ARG;
    HIPCHK(c, hipMemcpyAsync(dt, c->d_dt + case_id, 8, hipMemcpyDeviceToHost, c->q));
    HIPCHK(c, hipStreamSynchronize(c->q));
    return MHD_OK;
}

extern "C" int mhd_case_steps(mhd_ctx *c, int case_id, long long *steps)
{
    if (!c || !steps || case_id < 0 || case_id >= (int)c->cases.size())
        return MHD_ERR_ARG;
    *steps = c->cases[case_id]->steps;
    return MHD_OK;
}

extern "C" int mhd_divb_max(mhd_ctx *c, int case_id, double *out)
{
    if (!out || !active_case(c, case_id))
        return MHD_ERR_ARG;
    mhdk::launch_fill_u64(c->d_divb, c->case_cap, 0ull, c->q);
    mhdk::launch_divb(launch_of_chunk(c, KP_DT, 1), c->d_divb, mixed_of(c), c->q);
    unsigned long long bits = 0;
    HIPCHK(c, hipMemcpyAsync(&bits, c->d_divb + case_id, 8, hipMemcpyDeviceToHost, c->q));
    HIPCHK(c, hipStreamSynchronize(c->q));
    std::memcpy(out, &bits, 8);
    return MHD_OK;
}

static unsigned long long order_key_host(double x)
{
    unsigned long long b;
    std::memcpy(&b, &x, 8);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

extern "C" int mhd_diagnostics(mhd_ctx *c, int case_id, double *out)
{
    if (!out || !active_case(c, case_id))
        return MHD_ERR_ARG;
    int rc = grow(c, c->d_diag, c->cap_diag, (size_t)17 * c->case_cap);
    if (rc)
        return rc;
    std::vector<unsigned long long> init((size_t)17 * c->case_cap);
    for (size_t q = 0; q < init.size(); ++q)
        init[q] = order_key_host((q % 17) < 8 ? HUGE_VAL : -HUGE_VAL);
    HIPCHK(c, hipMemcpyAsync(c->d_diag, init.data(), 8 * init.size(), hipMemcpyHostToDevice, c->q));
    mhdk::launch_diag(launch_of(c, KP_DT), c->d_cst, c->d_diag, c->ucur, c->q);
    unsigned long long keys[17];
    HIPCHK(c, hipMemcpyAsync(keys, c->d_diag + 17 * (size_t)case_id, sizeof(keys), hipMemcpyDeviceToHost, c->q));
    HIPCHK(c, hipStreamSynchronize(c->q));
    for (int q = 0; q < 17; ++q)
    {
        const unsigned long long b = (keys[q] >> 63) ? (keys[q] & 0x7fffffffffffffffull) : ~keys[q];
        std::memcpy(&out[q], &b, 8);
    }
    return MHD_OK;
}

// The reduction of mhd_diagnostics for a list of cases into the device's
// static scratch: no device (or host heap) allocation, one fill kernel, one
// reduction kernel, one copy of 17 * n keys to the caller's stack.
extern "C" int mhd_diagnostics_noalloc(mhd_ctx *c, const int *case_ids, int n, double *out)
{
    if (!c || !case_ids || !out)
        return MHD_ERR_ARG;
    if (n < 1 || n > mhdk::DIAG_SEL_MAX)
        return fail(c, MHD_ERR_ARG,
                    "mhd_diagnostics_noalloc: " + std::to_string(n) + " cases requested; its static scratch holds 1 to " +
                        std::to_string(mhdk::DIAG_SEL_MAX));
    mhdk::DiagSel sel;
    sel.n = n;
    for (int q = 0; q < n; ++q)
    {
        if (!active_case(c, case_ids[q]))
            return fail(c, MHD_ERR_ARG, "mhd_diagnostics_noalloc: case " + std::to_string(case_ids[q]) + " is not active");
        for (int p = 0; p < q; ++p)
            if (case_ids[p] == case_ids[q])
                return fail(c, MHD_ERR_ARG, "mhd_diagnostics_noalloc: case " + std::to_string(case_ids[q]) +
                                                " requested twice");
        sel.ids[q] = case_ids[q];
    }
    HIPCHK(c, hipSetDevice(c->opt.device));
    mhdk::launch_diag_sel(launch_of(c, KP_DT), c->d_cst, c->d_scratch, c->ucur, sel, c->q);
    unsigned long long keys[17 * mhdk::DIAG_SEL_MAX];
    HIPCHK(c, hipMemcpyAsync(keys, c->d_scratch, sizeof(unsigned long long) * 17 * (size_t)n, hipMemcpyDeviceToHost,
                             c->q));
    HIPCHK(c, hipStreamSynchronize(c->q));
    for (int q = 0; q < 17 * n; ++q)
    {
        const unsigned long long b = (keys[q] >> 63) ? (keys[q] & 0x7fffffffffffffffull) : ~keys[q];
        std::memcpy(&out[q], &b, 8);
    }
    return MHD_OK;
}

// ----------------------------------------------------------- data exchange
// Geometry slots [slot, slot + nc) of a block as doubles (weight slots of a
// mixed-precision context are stored as float).
static int geo_download(mhd_ctx *c, const MhdBlockDev &d, int slot, int nc, std::vector<double> &out)
{
    out.assign((size_t)nc * d.slot, 0.0);
    if (mixed_of(c) && slot >= G_NMETRIC)
    {
        std::vector<float> f32(out.size());
        HIPCHK(c, hipMemcpy(f32.data(), reinterpret_cast<const void *>((uintptr_t)(d.geo + geo_slot_byte(slot, d.slot, 4))),
                            sizeof(float) * f32.size(), hipMemcpyDeviceToHost));
        out.assign(f32.begin(), f32.end());
        return MHD_OK;
    }
    HIPCHK(c, hipMemcpy(out.data(), reinterpret_cast<const void *>((uintptr_t)(d.geo + geo_slot_byte(slot, d.slot, 8))),
                        sizeof(double) * out.size(), hipMemcpyDeviceToHost));
    return MHD_OK;
}

// Bcell_from_Bface_w of a device block, formed from Mi, mu and the face area
// vectors with the B_cell kernel's (and the host builder's) arithmetic;
// interior cells only, as the host table.  The device keeps no table to upload.
static int bcell_weights_down(mhd_ctx *c, const MhdBlockDev &d, bool down, double *aos)
{
    if (!down)
        return fail(c, MHD_ERR_ARG, "Bcell_from_Bface_w is formed on the device from its least-squares data: "
                                    "it cannot be uploaded");
    std::vector<double> mi, mu, jd;
    int rc;
    if ((rc = geo_download(c, d, G_WBM, 6, mi)) || (rc = geo_download(c, d, G_WBU, 6, mu)) ||
        (rc = geo_download(c, d, G_JD, 9, jd)))
        return rc;
    const int64_t sl = d.slot, step[3] = {(int64_t)d.sj * d.sk, (int64_t)d.sk, 1};
    std::vector<double> soa((size_t)18 * sl, 0.0);
    for (int i = 0; i < d.ni; ++i)
        for (int j = 0; j < d.nj; ++j)
            for (int k = 0; k < d.nk; ++k)
            {
                const int64_t o = blk_off(d, i, j, k);
                const double m00 = mi[o], m01 = mi[sl + o], m02 = mi[2 * sl + o], m11 = mi[3 * sl + o],
                             m12 = mi[4 * sl + o], m22 = mi[5 * sl + o];
                for (int n = 0; n < 6; ++n)
                {
                    const int ax = n >> 1;
                    const int64_t oo = o + ((n & 1) ? step[ax] : 0);
                    double sv[3];
                    for (int a = 0; a < 3; ++a)
                    {
                        const double v = jd[(size_t)(3 * ax + a) * sl + oo];
                        sv[a] = (n & 1) ? v : -v;
                    }
                    const double m = mu[(size_t)n * sl + o];
                    const double rx = m * sv[0], ry = m * sv[1], rz = m * sv[2];
                    soa[(size_t)n * sl + o] = m00 * rx + m01 * ry + m02 * rz;
                    soa[(size_t)(6 + n) * sl + o] = m01 * rx + m11 * ry + m12 * rz;
                    soa[(size_t)(12 + n) * sl + o] = m02 * rx + m12 * ry + m22 * rz;
                }
            }
    soa_aos(d, LOC_CELL, 18, soa.data(), aos, true);
    return MHD_OK;
}

// Copy-on-write of a case's geometry before an upload changes it: a set
// other resident cases share is copied for this case alone (a key no later
// case can match), so their steps are not touched.
static int own_geometry(mhd_ctx *c, CaseRec &r)
{
    GeoSet &S = c->geos[r.geo_set];
    if (S.refs <= 1)
    {
        S.key = "own:" + std::to_string(r.id);
        return MHD_OK;
    }
    const double budget = c->opt.hbm_budget_gb * 1e9;
    if (budget > 0 && (double)(c->bytes_used + S.bytes) > budget)
        return fail(c, MHD_ERR_MEM, "HBM budget exceeded (copy of a shared geometry before an upload)");
    GeoSet G;
    G.key = "own:" + std::to_string(r.id);
    G.block_bytes = S.block_bytes;
    G.bytes = S.bytes;
    G.refs = 1;
    HIPCHK(c, hipMalloc((void **)&G.d, G.bytes));
    HIPCHK(c, hipMemcpy(G.d, S.d, G.bytes, hipMemcpyDeviceToDevice));
    c->bytes_used += G.bytes;
    S.refs -= 1;
    int gi = 0;
    while (gi < (int)c->geos.size() && c->geos[gi].refs > 0)
        ++gi;
    if (gi == (int)c->geos.size())
        c->geos.push_back(G);
    else
        c->geos[gi] = G;
    r.geo_set = gi;
    return rebuild(c); // block table: this case's new geometry addresses
}

static int exchange(mhd_ctx *c, int case_id, int block, int field, double *aos, size_t n, bool down)
{
    if (!aos || !c)
        return MHD_ERR_ARG;
    CaseRec *r = active_case(c, case_id);
    if (!r)
        return fail(c, MHD_ERR_STATE, "case not active (or context not finalized)");
    if (block < 0 || block >= r->nblocks)
        return MHD_ERR_ARG;
    int region, slot, loc, nc;
    if (!field_slot(field, c->ucur, region, slot, loc, nc) || region == 2)
        return fail(c, MHD_ERR_ARG, "field not stored on the device");
    {
        const MhdBlockDev &d0 = c->h_blk[r->first_block + block];
        if (region == 0 && slot + nc > st_nslot(d0.model))
            return fail(c, MHD_ERR_ARG, "field not held by this case (two-fluid cases only)");
        if (n != aos_count(d0, loc, nc))
            return fail(c, MHD_ERR_ARG, "buffer size does not match the FieldBlock extent");
    }
    HIPCHK(c, hipStreamSynchronize(c->q));
    if (region == 3)
        return bcell_weights_down(c, c->h_blk[r->first_block + block], down, aos);
    if (region == 1 && !down)
    {
        if (slot >= G_JD && slot < G_JD + 9)
            return fail(c, MHD_ERR_ARG, "the B_cell weights of a device case are formed from its face area vectors: "
                                        "JDxi/JDet/JDze cannot be replaced (build the case with the new metric)");
        int rc = own_geometry(c, *r);
        if (rc)
            return rc;
    }
    const MhdBlockDev &d = c->h_blk[r->first_block + block]; // (re-read: own_geometry may move it)
    std::vector<double> soa((size_t)nc * d.slot);
    if (region == 1 && mixed_of(c) && slot >= G_NMETRIC)
    {
        // weight tables of a mixed-precision context are stored as float
        float *dev = reinterpret_cast<float *>((uintptr_t)(d.geo + geo_slot_byte(slot, d.slot, sizeof(float))));
        std::vector<float> f32(soa.size());
        HIPCHK(c, hipMemcpy(f32.data(), dev, sizeof(float) * f32.size(), hipMemcpyDeviceToHost));
        soa.assign(f32.begin(), f32.end());
        soa_aos(d, loc, nc, soa.data(), aos, down);
        if (!down)
        {
            f32.assign(soa.begin(), soa.end());
            HIPCHK(c, hipMemcpy(dev, f32.data(), sizeof(float) * f32.size(), hipMemcpyHostToDevice));
        }
        return MHD_OK;
    }
    double *dev = region == 0 ? st_ptr(d) + (int64_t)slot * d.slot
                              : reinterpret_cast<double *>((uintptr_t)(d.geo + geo_slot_byte(slot, d.slot, 8)));
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
    if (c->cases[case_id]->state == 3)
        return fail(c, MHD_ERR_STATE, "case not ended (mhd_case_end)");
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
    if (c->cases[case_id]->state == 3)
        return fail(c, MHD_ERR_STATE, "case not ended (mhd_case_end)");
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

extern "C" int mhd_block_info(mhd_ctx *c, int case_id, int block, int *ni, int *nj, int *nk)
{
    if (!c || !ni || !nj || !nk || case_id < 0 || case_id >= (int)c->cases.size())
        return MHD_ERR_ARG;
    const CaseRec &r = *c->cases[case_id];
    if (r.state == 3 || block < 0 || block >= (int)r.cs.blk.size())
        return fail(c, MHD_ERR_ARG, "no such block (or case not ended)");
    const MhdBlockDev &d = r.cs.blk[block].d;
    *ni = d.ni;
    *nj = d.nj;
    *nk = d.nk;
    return MHD_OK;
}

extern "C" int mhd_get_faces(mhd_ctx *c, int case_id, int block, int *out)
{
    if (!c || !out || case_id < 0 || case_id >= (int)c->cases.size())
        return MHD_ERR_ARG;
    if (c->cases[case_id]->state == 3)
        return fail(c, MHD_ERR_STATE, "case not ended (mhd_case_end)");
    const CaseRec &r = *c->cases[case_id];
    if (block < 0 || block >= r.nblocks)
        return MHD_ERR_ARG;
    const Block &B = r.cs.grid.blocks[block];
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

extern "C" int mhd_get_nodes(mhd_ctx *c, int case_id, int block, int comp, double *out, size_t n)
{
    if (!c || !out || case_id < 0 || case_id >= (int)c->cases.size() || comp < 0 || comp > 2)
        return MHD_ERR_ARG;
    if (c->cases[case_id]->state == 3)
        return fail(c, MHD_ERR_STATE, "case not ended (mhd_case_end)");
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
    if (c->cases[case_id]->state == 3)
        return fail(c, MHD_ERR_STATE, "case not ended (mhd_case_end)");
    const MhdCaseConst &k = c->cases[case_id]->cs.cst;
    const double v[22] = {k.gamma,    k.cfl,         k.inver_MA2, k.momentum_hall_coeff, k.M_H,      k.M_ref,
                          k.ne_floor, k.state_coeff, k.qinf[0],   k.qinf[1],             k.qinf[2],  k.qinf[3],
                          k.qinf[4],  k.B_imf[0],    k.B_imf[1],  k.B_imf[2],            k.u_inf[0], k.u_inf[1],
                          k.u_inf[2], k.dt_floor,    0.0,         0.0};
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

extern "C" int mhd_tuning(mhd_ctx *c, int kernel, int *rows, int *i_inner, int *geo_shared)
{
    if (!c || !rows || !i_inner || !geo_shared || kernel < 0 || kernel >= KP_COUNT)
        return MHD_ERR_ARG;
    const int v = c->jt_idx[kernel];
    *rows = kJt[v % 3];
    *i_inner = v < kVariants && (v % 6) >= 3;
    *geo_shared = v >= kVariants ? 2 : v >= 6;
    return MHD_OK;
}

extern "C" int mhd_set_tuning(mhd_ctx *c, int kernel, int rows, int i_inner, int geo_shared)
{
    if (!c || kernel < 0 || kernel >= KP_COUNT)
        return MHD_ERR_ARG;
    int ji = -1;
    for (int q = 0; q < 3; ++q)
        if (kJt[q] == rows)
            ji = q;
    if (ji < 0 || (i_inner != 0 && i_inner != 1) || geo_shared < 0 || geo_shared > 2)
        return fail(c, MHD_ERR_ARG, "launch shape: rows 1, 2 or 4; i_inner 0 or 1; geo_shared 0, 1 or 2");
    if (geo_shared == 2 && (i_inner || !loops_members(kernel)))
        return fail(c, MHD_ERR_ARG, std::string("launch shape: member-looping tiles (geo_shared 2) are j-inner and only for ") +
                                        kKernelName[KP_DERIVED]);
    c->jt_idx[kernel] = geo_shared == 2 ? kVariants + ji : ji + 3 * i_inner + 6 * geo_shared;
    (void)hipStreamSynchronize(c->q);
    drop_graph(c);
    return MHD_OK;
}

extern "C" int mhd_autotune(mhd_ctx *c)
{
    if (!c)
        return MHD_ERR_ARG;
    if (!c->finalized)
        return fail(c, MHD_ERR_STATE, "context not finalized");
    if (c->fluid_pending)
        return fail(c, MHD_ERR_STATE, "a staged step is open: mhd_apply_update first");
    HIPCHK(c, hipSetDevice(c->opt.device));
    int rc = autotune(c);
    if (rc)
        return rc;
    drop_graph(c);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->q));
    return check_err(c);
}

extern "C" int mhd_memory(mhd_ctx *c, double *bytes)
{
    if (!c || !bytes)
        return MHD_ERR_ARG;
    *bytes = (double)c->bytes_used;
    return MHD_OK;
}

// ---------------------------------------------------- refresh and restart
// Ghosts from the current interior (U and B syncs: physical, then halo) and
// the derived fields, for every resident case: what the reference runs after
// a restart read (Sync of U_H and B faces, UpdateDerivedFields_,
// Z4_Lunar/src/physics/LunarSolver_Tools.cpp:143-161).  Idempotent for cases
// whose ghosts are already current.
extern "C" int mhd_refresh(mhd_ctx *c)
{
    if (!c)
        return MHD_ERR_ARG;
    if (!c->finalized)
        return fail(c, MHD_ERR_STATE, "context not finalized");
    if (c->fluid_pending)
        return fail(c, MHD_ERR_STATE, "a staged step is open: mhd_apply_update first");
    HIPCHK(c, hipSetDevice(c->opt.device));
    run_ops(c, SY_U_BC0, c->ucur);
    run_ops(c, SY_U_BC1, c->ucur);
    run_ops(c, SY_U_HALO, c->ucur);
    run_ops(c, SY_B_BC, 0);
    run_ops(c, SY_B_HALO, 0);
    // the next step's dt is a min accumulated by the derived kernel: restart
    // it, and the dt error flags with it (the state is a new one)
    mhdk::launch_fill_u64(c->d_dtbits, c->case_cap, dbits(1e100), c->q);
    HIPCHK(c, hipMemsetAsync(c->d_err, 0, sizeof(int) * c->case_cap, c->q));
    derived(c);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->q));
    return MHD_OK;
}

namespace
{
struct DeviceIO : FieldIO
{
    mhd_ctx *c;
    int cid;
    CaseRec *r;
    DeviceIO(mhd_ctx *c_, int cid_, CaseRec *r_) : c(c_), cid(cid_), r(r_) {}
    int nblocks() override { return r->nblocks; }
    const MhdBlockDev &block(int b) override { return c->h_blk[r->first_block + b]; }
    int get(int b, int field, std::vector<double> &aos) override
    {
        int region, slot, loc, nc;
        field_slot(field, 0, region, slot, loc, nc);
        aos.assign(aos_count(block(b), loc, nc), 0.0);
        return exchange(c, cid, b, field, aos.data(), aos.size(), true);
    }
    int put(int b, int field, const std::vector<double> &aos) override
    {
        return exchange(c, cid, b, field, const_cast<double *>(aos.data()), aos.size(), false);
    }
};
} // namespace

extern "C" int mhd_write_restart(mhd_ctx *c, int case_id, const char *path, long long step, double time)
{
    if (!c || !path)
        return MHD_ERR_ARG;
    CaseRec *r = active_case(c, case_id);
    if (!r)
        return fail(c, MHD_ERR_STATE, "case not active (or context not finalized)");
    try
    {
        DeviceIO io(c, case_id, r);
        return restart_write(io, path, step, time);
    }
    catch (const std::exception &e)
    {
        return fail(c, MHD_ERR_ARG, e.what());
    }
}

extern "C" int mhd_write_tecplot(mhd_ctx *c, int case_id, const char *path, double time)
{
    if (!c || !path)
        return MHD_ERR_ARG;
    CaseRec *r = active_case(c, case_id);
    if (!r)
        return fail(c, MHD_ERR_STATE, "case not active (or context not finalized)");
    try
    {
        DeviceIO io(c, case_id, r);
        return tecplot_write(io, r->cs.grid, r->cs.cst, r->cs.L_ref, path, time);
    }
    catch (const std::exception &e)
    {
        return fail(c, MHD_ERR_ARG, e.what());
    }
}

extern "C" int mhd_read_restart(mhd_ctx *c, int case_id, const char *path, long long *step, double *time)
{
    if (!c || !path)
        return MHD_ERR_ARG;
    CaseRec *r = active_case(c, case_id);
    if (!r)
        return fail(c, MHD_ERR_STATE, "case not active (or context not finalized)");
    try
    {
        DeviceIO io(c, case_id, r);
        const int rc = restart_read(io, path, step, time);
        if (rc)
        {
            // a copy failed after the file was checked: bring the ghosts and
            // derived fields in line with what was written
            (void)mhd_refresh(c);
            return rc;
        }
    }
    catch (const std::exception &e)
    {
        return fail(c, MHD_ERR_ARG, e.what()); // the file was rejected before any field was written
    }
    return mhd_refresh(c);
}

extern "C" int mhd_add_case_from_dir(mhd_ctx *c, const char *dir, const mhd_wind_desc *d, int *case_id)
{
    if (!c || !dir || !d)
        return MHD_ERR_ARG;
    mhd_host_case *h = nullptr;
    const int rc = mhd_host_case_from_dir(dir, d, c->opt.ghost, &h);
    if (rc)
        return fail(c, rc, std::string("cannot build a case from ") + dir);
    const int ra = mhd_adopt_host_case(c, h, case_id);
    mhd_host_case_destroy(h);
    return ra;
}
