// HIP kernels of the explicit MHD step for gfx950 (CDNA4).
//
// Memory-bound stencils over the padded per-block box of layout.h.  A
// launch is a 1-D grid over tiles of 64 x rows consecutive points (k
// fastest) of one (block, i) plane of the launch range, so a wavefront's
// slot loads and stores are contiguous runs along k; see MhdTileMap.
//
// FP64 arithmetic is written in the same operation order as the reference
// (and as the CPU oracle) and the library is compiled with
// -ffp-contract=off, so the FP64 path reproduces the reference bit for bit
// wherever division and square root are correctly rounded.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>

#include "../host/setup.h"
#include "../layout.h"
#include "kernels.h"

using namespace mpcns;

namespace
{
// Device addresses travel as integers (block table, op lists).  Converting
// them through a global-address-space pointer lets the compiler emit global
// (not flat) loads and stores, whose completion it can count per access
// instead of draining every outstanding access at each wait.
typedef __attribute__((address_space(1))) double gdouble;
__device__ __forceinline__ double *gptr(int64_t a) { return (double *)(gdouble *)(uintptr_t)a; }
__device__ __forceinline__ double *dst_ptr(const MhdBlockDev &b) { return gptr(b.st); }
// The metric (Jac, JD) is stored in double; the Hodge weights and weight
// tables after it as WT: double for the FP64 step, float for the mixed one.
// dweights<WT>(b) + (s - G_NMETRIC) * slot is weight slot s.
__device__ __forceinline__ const double *dmetric(const MhdBlockDev &b) { return gptr(b.geo); }
template <typename WT> __device__ __forceinline__ const WT *dweights(const MhdBlockDev &b)
{
    return (const WT *)(__attribute__((address_space(1))) WT *)(uintptr_t)(b.geo + geo_slot_byte(G_NMETRIC, b.slot, 8));
}

__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }
// Division and square root of the flux arithmetic: correctly rounded in
// FP64 (the reference's operations); in the FP32 flux of the mixed step the
// hardware reciprocal and square root (1 ulp), well inside FP32 rounding.
template <typename R> __device__ __forceinline__ R fdiv(R a, R b) { return a / b; }
template <> __device__ __forceinline__ float fdiv<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
template <typename R> __device__ __forceinline__ R fsqrt(R a) { return sqrt(a); }
template <> __device__ __forceinline__ float fsqrt<float>(float a) { return __builtin_amdgcn_sqrtf(a); }

// Map a workgroup/thread to (block, i, j, k) of its tile.  `wg` is false for
// padding workgroups of the XCD-ordered grid (no block at all); the return
// value is false for padding threads.
// Workgroup -> (block, tile index within the block's entry); false for
// padding workgroups of the XCD-ordered grid.
__device__ __forceinline__ bool tile_block(const MhdTileMap &tm, MhdBlockDev &b, int &lt)
{
    int t = (int)blockIdx.x;
    if (tm.xcd_chunk > 0)
        t = (t & 7) * tm.xcd_chunk + (t >> 3);
    if (t >= tm.ntiles)
        return false;
    int bi = 0, hi = tm.nblk - 1;
    while (bi < hi)
    {
        const int mid = (bi + hi + 1) >> 1;
        if (tm.block_start[mid] <= t)
            bi = mid;
        else
            hi = mid - 1;
    }
    lt = t - tm.block_start[bi];
    if (tm.members)
    {
        const int kk = tm.lead_k[bi];
        const int m = lt % kk;
        lt /= kk;
        b = tm.blks[tm.members[tm.lead_moff[bi] + m]];
    }
    else
        b = tm.blks[bi];
    return true;
}

// Column tiles: a workgroup of 64 x COL_TJ threads owns the cells
// (i0 .. i0 + ci - 1) x (j0 .. j0 + COL_TJ - 1) x (k0 .. k0 + 63) of a block
// and marches along i; one thread per (j, k) column.
constexpr int COL_TJ = 4, COL_TK = 64; // 4 x 64 cells: 187 us vs 196 at 8 x 64 (6 x 64^2 x 128 shell)
__device__ __forceinline__ bool tile_col(const MhdTileMap &tm, MhdBlockDev &b, int &i0, int &j0, int &k0)
{
    int lt;
    if (!tile_block(tm, b, lt))
        return false;
    const int ntk = (b.nk + COL_TK - 1) / COL_TK, ntj = (b.nj + COL_TJ - 1) / COL_TJ;
    k0 = (lt % ntk) * COL_TK;
    lt /= ntk;
    j0 = (lt % ntj) * COL_TJ;
    i0 = (lt / ntj) * tm.ci;
    return true;
}

// Workgroup -> (entry, tile index within the entry); false for padding
// workgroups of the XCD-ordered grid.
__device__ __forceinline__ bool tile_entry(const MhdTileMap &tm, int &bi, int &lt)
{
    int t = (int)blockIdx.x;
    if (tm.xcd_chunk > 0)
        t = (t & 7) * tm.xcd_chunk + (t >> 3);
    if (t >= tm.ntiles)
        return false;
    bi = 0;
    int hi = tm.nblk - 1;
    while (bi < hi)
    {
        const int mid = (bi + hi + 1) >> 1;
        if (tm.block_start[mid] <= t)
            bi = mid;
        else
            hi = mid - 1;
    }
    lt = t - tm.block_start[bi];
    return true;
}

// The point (i, j, k) of tile lt of block b's launch range: a tile is 64 *
// rows consecutive points of one (j, k) plane of the range, k fastest, so
// every wavefront is full however the extents divide by 64.  False for
// padding threads.
__device__ __forceinline__ bool tile_point(const MhdTileMap &tm, const MhdBlockDev &b, int lt, int &i, int &j, int &k)
{
    const int per_tile = 64 * (int)blockDim.y;
    const int ext_i = b.ni - tm.lo + tm.hi, ext_j = b.nj - tm.lo + tm.hi, ext_k = b.nk - tm.lo + tm.hi;
    const int plane = ext_j * ext_k;
    const int ntp = (plane + per_tile - 1) / per_tile;
    // i-inner order: a tile's i-neighbours are the next tiles in launch order
    // (still in L2 when they are read again)
    const int pt = tm.i_inner ? lt / ext_i : lt % ntp;
    const int iy = tm.i_inner ? lt % ext_i : lt / ntp;
    const int f = pt * per_tile + (int)threadIdx.y * 64 + (int)threadIdx.x;
    i = tm.lo + iy;
    j = tm.lo + f / ext_k;
    k = tm.lo + f % ext_k;
    return f < plane;
}

__device__ __forceinline__ bool tile_cell(const MhdTileMap &tm, MhdBlockDev &b, int &i, int &j, int &k, bool &wg)
{
    int bi, lt;
    wg = tile_entry(tm, bi, lt);
    if (!wg)
        return false;
    if (tm.members)
    {
        const int kk = tm.lead_k[bi];
        const int m = lt % kk;
        lt /= kk;
        b = tm.blks[tm.members[tm.lead_moff[bi] + m]];
    }
    else
        b = tm.blks[bi];
    return tile_point(tm, b, lt, i, j, k);
}

// Member-looping tiles of a geometry-shared map (tm.kloop): a tile is a point
// tile of the entry's first case; the K cases of the entry are the blocks
// tm.blks[tm.members[m0 + m]], m < K, which share its geometry and extents.
__device__ __forceinline__ bool tile_cell_k(const MhdTileMap &tm, MhdBlockDev &b, int &i, int &j, int &k, bool &wg,
                                            int &m0, int &K)
{
    int bi, lt;
    wg = tile_entry(tm, bi, lt);
    if (!wg)
        return false;
    m0 = tm.lead_moff[bi];
    K = tm.lead_k[bi];
    b = tm.blks[tm.members[m0]];
    return tile_point(tm, b, lt, i, j, k);
}

// Plane-chunk tiles: as tile_cell, but a tile spans the planes
// [i0, i1) (tm.ci planes, fewer in the block's last chunk) at one (j, k),
// so a thread marches along i and keeps what plane i - 1 shares with plane
// i in registers.  With tm.ci <= 1 a tile is one plane (i1 = i0 + 1).
// With nsub > 1 the launch holds nsub workgroups per tile (sub = 0 ..
// nsub - 1, consecutive after the XCD mapping, so the nsub workgroups of a
// tile share an XCD's L2).
__device__ __forceinline__ bool tile_planes(const MhdTileMap &tm, MhdBlockDev &b, int &i0, int &i1, int &j, int &k,
                                            bool &wg, int nsub = 1, int *sub = nullptr)
{
    const int ci = tm.ci > 1 ? tm.ci : 1;
    int t = (int)blockIdx.x;
    if (tm.xcd_chunk > 0)
        t = (t & 7) * tm.xcd_chunk + (t >> 3);
    wg = t < nsub * tm.ntiles;
    if (!wg)
        return false;
    if (nsub > 1)
    {
        *sub = t % nsub;
        t /= nsub;
    }
    int bi = 0, hi = tm.nblk - 1;
    while (bi < hi)
    {
        const int mid = (bi + hi + 1) >> 1;
        if (tm.block_start[mid] <= t)
            bi = mid;
        else
            hi = mid - 1;
    }
    int lt = t - tm.block_start[bi];
    if (tm.members)
    {
        const int kk = tm.lead_k[bi];
        const int m = lt % kk;
        lt /= kk;
        b = tm.blks[tm.members[tm.lead_moff[bi] + m]];
    }
    else
        b = tm.blks[bi];
    const int per_tile = 64 * (int)blockDim.y;
    const int ext_i = b.ni - tm.lo + tm.hi, ext_j = b.nj - tm.lo + tm.hi, ext_k = b.nk - tm.lo + tm.hi;
    const int nch = (ext_i + ci - 1) / ci;
    const int plane = ext_j * ext_k;
    const int ntp = (plane + per_tile - 1) / per_tile;
    const int pt = tm.i_inner ? lt / nch : lt % ntp;
    const int iy = tm.i_inner ? lt % nch : lt / ntp;
    const int f = pt * per_tile + (int)threadIdx.y * 64 + (int)threadIdx.x;
    i0 = tm.lo + iy * ci;
    i1 = min(i0 + ci, tm.lo + ext_i);
    j = tm.lo + f / ext_k;
    k = tm.lo + f % ext_k;
    return f < plane;
}

// Padded-row plane-chunk tiles of a launch range: its (j, k) plane with
// each row padded by one point on either side (k = lo - 1 and the point past
// the range's last) and flattened, k fastest.  Wavefront w of a plane takes the 64
// consecutive padded points from 62 w - 1: lanes 1 .. 62 own 62 points, and
// every owned point's neighbours along k are the lanes beside it, so a
// kernel can take what they formed by a lane shift instead of loading it.
// Lanes past the plane are clamped onto its last point (own = false); with
// nsub > 1, nsub workgroups per tile as in tile_planes.
constexpr int PAD_OWN = 62;
__device__ __forceinline__ bool tile_padded(const MhdTileMap &tm, MhdBlockDev &b, int &i0, int &i1, int &j, int &k,
                                            bool &own, int nsub, int &sub)
{
    const int ci = tm.ci > 1 ? tm.ci : 1;
    int t = (int)blockIdx.x;
    if (tm.xcd_chunk > 0)
        t = (t & 7) * tm.xcd_chunk + (t >> 3);
    if (t >= nsub * tm.ntiles)
        return false;
    sub = t % nsub;
    t /= nsub;
    int bi = 0, hi = tm.nblk - 1;
    while (bi < hi)
    {
        const int mid = (bi + hi + 1) >> 1;
        if (tm.block_start[mid] <= t)
            bi = mid;
        else
            hi = mid - 1;
    }
    int lt = t - tm.block_start[bi];
    if (tm.members)
    {
        const int kk = tm.lead_k[bi];
        const int m = lt % kk;
        lt /= kk;
        b = tm.blks[tm.members[tm.lead_moff[bi] + m]];
    }
    else
        b = tm.blks[bi];
    const int jt = (int)blockDim.y, lo = tm.lo;
    const int ext_i = b.ni - lo + tm.hi, ext_j = b.nj - lo + tm.hi, ext_k = b.nk - lo + tm.hi;
    const int Kp = ext_k + 2, P = ext_j * Kp;
    const int nch = (ext_i + ci - 1) / ci, nw = (P + PAD_OWN - 1) / PAD_OWN, ntp = (nw + jt - 1) / jt;
    const int pt = tm.i_inner ? lt / nch : lt % ntp;
    const int iy = tm.i_inner ? lt % nch : lt / ntp;
    const int w = pt * jt + (int)threadIdx.y, lane = (int)threadIdx.x;
    int F = PAD_OWN * w + lane - 1;
    own = lane >= 1 && lane <= PAD_OWN && F < P;
    F = min(max(F, 0), P - 1);
    const int r = F / Kp;
    j = lo + r;
    k = lo - 1 + (F - r * Kp);
    own = own && k >= lo && k < lo + ext_k;
    i0 = lo + iy * ci;
    i1 = min(i0 + ci, lo + ext_i);
    return true;
}

// The value lane - 1 (lane_prev) or lane + 1 (lane_next) of the wavefront
// holds (DPP wave shifts by one lane; lane 0, resp. 63, keeps its own).
__device__ __forceinline__ double lane_prev(double v)
{
    const long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp((int)x, (int)x, 0x138, 0xf, 0xf, false); // wave_shr:1
    const int hi = __builtin_amdgcn_update_dpp((int)(x >> 32), (int)(x >> 32), 0x138, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double lane_next(double v)
{
    const long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp((int)x, (int)x, 0x130, 0xf, 0xf, false); // wave_shl:1
    const int hi = __builtin_amdgcn_update_dpp((int)(x >> 32), (int)(x >> 32), 0x130, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// ------------------------------------------------------------------ state
struct Prim
{
    double rho_raw, u, v, w, p;
};

// calc_PV (Z4_Lunar/src/physics/LunarSolver_Tools.cpp:58-110).  EXACT: the
// velocities are quotients as in the reference (FP64 step); otherwise
// products with one reciprocal of rho (within 2 ulp of FP64, for the mixed
// step's FP32 flux).
template <bool EXACT = true>
__device__ __forceinline__ Prim prim_from(double r0, double m1, double m2, double m3, double E, double gm)
{
    Prim q;
    q.rho_raw = r0;
    const double rho = dmax(r0, 1e-12);
    if (EXACT)
    {
        q.u = m1 / rho;
        q.v = m2 / rho;
        q.w = m3 / rho;
    }
    else
    {
        const double ir = 1.0 / rho;
        q.u = m1 * ir;
        q.v = m2 * ir;
        q.w = m3 * ir;
    }
    const double ke = 0.5 * rho * (q.u * q.u + q.v * q.v + q.w * q.w);
    double eint = E - ke;
    if (eint < 0.0)
        eint = 0.0;
    double p = (gm - 1.0) * eint;
    q.p = dmax(p, 1e-12);
    return q;
}

template <bool EXACT = true>
__device__ __forceinline__ Prim prim_of(const double *__restrict__ U, int64_t sl, int64_t o, double gm)
{
    return prim_from<EXACT>(U[o], U[sl + o], U[2 * sl + o], U[3 * sl + o], U[4 * sl + o], gm);
}

// -------------------------------------------------------- Rusanov flux
// Scheme_U_ lambdas (Z4_Lunar/src/physics/LunarSolver_FluidRHS.cpp:5-102)
template <typename R>
__device__ __forceinline__ void rusanov(R gm, const R *WL, const R *WR, R k1, R k2, R k3, R *out)
{
    const R rhoL = WL[0], uL = WL[1], vL = WL[2], wL = WL[3], pL = WL[4];
    const R rhoR = WR[0], uR = WR[1], vR = WR[2], wR = WR[3], pR = WR[4];
    const R one = (R)1.0, half = (R)0.5, zero = (R)0.0;
    const R eL = fdiv(pL, gm - one) + half * rhoL * (uL * uL + vL * vL + wL * wL);
    const R eR = fdiv(pR, gm - one) + half * rhoR * (uR * uR + vR * vR + wR * wR);
    const R kk = k1 * k1 + k2 * k2 + k3 * k3;
    const R uvL = k1 * uL + k2 * vL + k3 * wL;
    const R uvR = k1 * uR + k2 * vR + k3 * wR;
    const R tL = fdiv(gm * pL, rhoL < (R)1e-30 ? (R)1e-30 : rhoL);
    const R tR = fdiv(gm * pR, rhoR < (R)1e-30 ? (R)1e-30 : rhoR);
    const R cL = fsqrt(zero < tL ? tL : zero);
    const R cR = fsqrt(zero < tR ? tR : zero);
    const R sk = fsqrt(kk);
    const R radL = fabs(uvL) + cL * sk;
    const R radR = fabs(uvR) + cR * sk;
    const R rad = (radL < radR) ? radR : radL;
    const R hr = half * rad;
    out[0] = half * (rhoL * uvL + rhoR * uvR) - hr * (rhoR - rhoL);
    out[1] = half * ((rhoL * uvL * uL + k1 * pL) + (rhoR * uvR * uR + k1 * pR)) - hr * (rhoR * uR - rhoL * uL);
    out[2] = half * ((rhoL * uvL * vL + k2 * pL) + (rhoR * uvR * vR + k2 * pR)) - hr * (rhoR * vR - rhoL * vL);
    out[3] = half * ((rhoL * uvL * wL + k3 * pL) + (rhoR * uvR * wR + k3 * pR)) - hr * (rhoR * wR - rhoL * wL);
    out[4] = half * (uvL * (eL + pL) + uvR * (eR + pR)) - hr * (eR - eL);
}

// The per-cell part of the Rusanov flux -- primitive state, total energy and
// sound speed -- evaluated once per cell and shared by every face the cell
// touches; rusanov_pre combines two of them with the same arithmetic as
// rusanov / rusanov_mixed, so a face's flux is bit for bit the same.
struct FluxCell
{
    double rho, u, v, w, p, e, c;
};
template <typename R>
__device__ __forceinline__ FluxCell flux_cell(const double *__restrict__ U, int64_t sl, int64_t o, double gm)
{
    const Prim q = prim_of<sizeof(R) == 8>(U, sl, o, gm);
    FluxCell f;
    f.rho = q.rho_raw;
    f.u = q.u;
    f.v = q.v;
    f.w = q.w;
    f.p = q.p;
    if constexpr (sizeof(R) == 8)
    {
        f.e = fdiv(f.p, gm - 1.0) + 0.5 * f.rho * (f.u * f.u + f.v * f.v + f.w * f.w);
        const double t = fdiv(gm * f.p, f.rho < 1e-30 ? 1e-30 : f.rho);
        f.c = fsqrt(0.0 < t ? t : 0.0);
    }
    else
    {
        const double igm = 1.0 / (gm - 1.0);
        f.e = f.p * igm + 0.5 * f.rho * (f.u * f.u + f.v * f.v + f.w * f.w);
        const float t = fdiv((float)(gm * f.p), (float)(f.rho < 1e-30 ? 1e-30 : f.rho));
        f.c = (double)fsqrt(0.0f < t ? t : 0.0f);
    }
    return f;
}

template <typename R>
__device__ __forceinline__ void rusanov_pre(const FluxCell &L, const FluxCell &Rc, double k1, double k2, double k3,
                                            double *out)
{
    const double rhoL = L.rho, uL = L.u, vL = L.v, wL = L.w, pL = L.p, eL = L.e;
    const double rhoR = Rc.rho, uR = Rc.u, vR = Rc.v, wR = Rc.w, pR = Rc.p, eR = Rc.e;
    const double uvL = k1 * uL + k2 * vL + k3 * wL;
    const double uvR = k1 * uR + k2 * vR + k3 * wR;
    if constexpr (sizeof(R) == 8)
    {
        const double kk = k1 * k1 + k2 * k2 + k3 * k3;
        const double sk = fsqrt(kk);
        const double radL = fabs(uvL) + L.c * sk;
        const double radR = fabs(uvR) + Rc.c * sk;
        const double rad = (radL < radR) ? radR : radL;
        const double hr = 0.5 * rad;
        out[0] = 0.5 * (rhoL * uvL + rhoR * uvR) - hr * (rhoR - rhoL);
        out[1] = 0.5 * ((rhoL * uvL * uL + k1 * pL) + (rhoR * uvR * uR + k1 * pR)) - hr * (rhoR * uR - rhoL * uL);
        out[2] = 0.5 * ((rhoL * uvL * vL + k2 * pL) + (rhoR * uvR * vR + k2 * pR)) - hr * (rhoR * vR - rhoL * vL);
        out[3] = 0.5 * ((rhoL * uvL * wL + k3 * pL) + (rhoR * uvR * wR + k3 * pR)) - hr * (rhoR * wR - rhoL * wL);
        out[4] = 0.5 * (uvL * (eL + pL) + uvR * (eR + pR)) - hr * (eR - eL);
    }
    else
    {
        const float fk1 = (float)k1, fk2 = (float)k2, fk3 = (float)k3;
        const float sk = fsqrt(fk1 * fk1 + fk2 * fk2 + fk3 * fk3);
        const float radL = fabsf((float)uvL) + (float)L.c * sk;
        const float radR = fabsf((float)uvR) + (float)Rc.c * sk;
        const float hr = 0.5f * ((radL < radR) ? radR : radL);
        out[0] = 0.5 * (rhoL * uvL + rhoR * uvR) - (double)(hr * (float)(rhoR - rhoL));
        out[1] = 0.5 * ((rhoL * uvL * uL + k1 * pL) + (rhoR * uvR * uR + k1 * pR)) -
                 (double)(hr * (float)(rhoR * uR - rhoL * uL));
        out[2] = 0.5 * ((rhoL * uvL * vL + k2 * pL) + (rhoR * uvR * vR + k2 * pR)) -
                 (double)(hr * (float)(rhoR * vR - rhoL * vL));
        out[3] = 0.5 * ((rhoL * uvL * wL + k3 * pL) + (rhoR * uvR * wR + k3 * pR)) -
                 (double)(hr * (float)(rhoR * wR - rhoL * wL));
        out[4] = 0.5 * (uvL * (eL + pL) + uvR * (eR + pR)) - (double)(hr * (float)(eR - eL));
    }
}

// Absorbing-wall state of BC_Solid_Surface_ (Z4_Lunar/src/boundary/Boundary_Condition_Fluid.cpp:19-184):
// d = (rho_r, 0, 0, 0, p_r / (gamma - 1)) from the cell state s; the wall op
// lists and the fluid kernel's folded k-face walls share it.
__device__ __forceinline__ double solid_p(double gm, double rho, double mx, double my, double mz, double E)
{
    rho = dmax(rho, 2.5e-3);
    const double ke = 0.5 * (mx * mx + my * my + mz * mz) / rho;
    double eint = E - ke;
    if (eint < 0.0)
        eint = 0.0;
    const double p = (gm - 1.0) * eint;
    return dmax(p, 1e-12);
}
__device__ __forceinline__ void solid_state(double gm, const double *s, double *d)
{
    const double rho_r = dmax(s[0], 2.5e-3);
    const double p_r = solid_p(gm, rho_r, s[1], s[2], s[3], s[4]);
    d[0] = rho_r;
    d[1] = rho_r * 0.0;
    d[2] = rho_r * 0.0;
    d[3] = rho_r * 0.0;
    const double ke_g = 0.5 * rho_r * (0.0 * 0.0 + 0.0 * 0.0 + 0.0 * 0.0);
    d[4] = p_r / (gm - 1.0) + ke_g;
}

template <typename T> __device__ __forceinline__ T minmod(T a, T b)
{
    if (a * b <= (T)0)
        return (T)0;
    return fabs(a) < fabs(b) ? a : b;
}

// Rusanov flux of the mixed-precision step: the central part 1/2 (F(W_L) +
// F(W_R)) in FP64, the wave speeds and the dissipation 1/2 rad (U_R - U_L) in
// FP32 on the FP64 differences of the conserved states.  A uniform state
// keeps its flux to FP64 rounding and the FP32 error is relative to the
// dissipation, which is small where the flow is smooth.
__device__ __forceinline__ void rusanov_mixed(double gm, const double *WL, const double *WR, double k1, double k2,
                                              double k3, double *out)
{
    const double rhoL = WL[0], uL = WL[1], vL = WL[2], wL = WL[3], pL = WL[4];
    const double rhoR = WR[0], uR = WR[1], vR = WR[2], wR = WR[3], pR = WR[4];
    const double igm = 1.0 / (gm - 1.0);
    const double eL = pL * igm + 0.5 * rhoL * (uL * uL + vL * vL + wL * wL);
    const double eR = pR * igm + 0.5 * rhoR * (uR * uR + vR * vR + wR * wR);
    const double uvL = k1 * uL + k2 * vL + k3 * wL;
    const double uvR = k1 * uR + k2 * vR + k3 * wR;
    const float fk1 = (float)k1, fk2 = (float)k2, fk3 = (float)k3;
    const float sk = fsqrt(fk1 * fk1 + fk2 * fk2 + fk3 * fk3);
    const float tL = fdiv((float)(gm * pL), (float)(rhoL < 1e-30 ? 1e-30 : rhoL));
    const float tR = fdiv((float)(gm * pR), (float)(rhoR < 1e-30 ? 1e-30 : rhoR));
    const float cL = fsqrt(0.0f < tL ? tL : 0.0f);
    const float cR = fsqrt(0.0f < tR ? tR : 0.0f);
    const float radL = fabsf((float)uvL) + cL * sk;
    const float radR = fabsf((float)uvR) + cR * sk;
    const float hr = 0.5f * ((radL < radR) ? radR : radL);
    out[0] = 0.5 * (rhoL * uvL + rhoR * uvR) - (double)(hr * (float)(rhoR - rhoL));
    out[1] = 0.5 * ((rhoL * uvL * uL + k1 * pL) + (rhoR * uvR * uR + k1 * pR)) -
             (double)(hr * (float)(rhoR * uR - rhoL * uL));
    out[2] = 0.5 * ((rhoL * uvL * vL + k2 * pL) + (rhoR * uvR * vR + k2 * pR)) -
             (double)(hr * (float)(rhoR * vR - rhoL * vL));
    out[3] = 0.5 * ((rhoL * uvL * wL + k3 * pL) + (rhoR * uvR * wR + k3 * pR)) -
             (double)(hr * (float)(rhoR * wR - rhoL * wL));
    out[4] = 0.5 * (uvL * (eL + pL) + uvR * (eR + pR)) - (double)(hr * (float)(eR - eL));
}

} // namespace

// ======================================================================
// dt: Compute_Timestep (Z4_Lunar/src/physics/LunarSolver_Timestep.cpp:8-140).
// Per-cell CFL limit; the caller reduces with a wavefront min and one 64-bit
// atomicMin per workgroup on the bit pattern (positive doubles order like
// their bits).  Cells with V <= 0 or a non-positive / NaN denominator are
// skipped, as the reference does.
//
// The values a cell's CFL limit and contravariant velocities read: U (5),
// Jac and the area vectors JD of its 6 faces (f = 2 * axis + side).  Loaded
// up front, unconditionally, so no load sits behind a data-dependent branch.
struct CellGeo
{
    double u[5];
    double jac;
    double a[6][3];
};

__device__ __forceinline__ void load_cell_geo(const MhdBlockDev &b, const double *__restrict__ G, int64_t o,
                                              CellGeo &q)
{
    const int64_t sl = b.slot;
    const int64_t step[3] = {(int64_t)b.sj * b.sk, (int64_t)b.sk, 1};
    q.jac = G[(int64_t)G_JAC * sl + o];
#pragma unroll
    for (int f = 0; f < 6; ++f)
    {
        const int ax = f >> 1;
        const int64_t oo = o + ((f & 1) ? step[ax] : 0);
        const double *A = G + (int64_t)(G_JD + 3 * ax) * sl;
        q.a[f][0] = A[oo];
        q.a[f][1] = A[sl + oo];
        q.a[f][2] = A[2 * sl + oo];
    }
}

__device__ __forceinline__ void load_cell(const MhdBlockDev &b, const double *__restrict__ U,
                                          const double *__restrict__ G, int64_t o, CellGeo &q)
{
    const int64_t sl = b.slot;
#pragma unroll
    for (int m = 0; m < 5; ++m)
        q.u[m] = U[m * sl + o];
    load_cell_geo(b, G, o, q);
}

// B_cell = sum_n W[c][n] pt[n] (calc_Bcell, Z4_Lunar/src/physics/LunarSolver_Tools.cpp:
// 194-236) with Bcell_from_Bface_w formed in place from the stored
// least-squares data: W[c][n] = Mi[c] . (mu_n s_n), s_n = -+ the area vector
// of face n (q.a), in the host builder's order of operations.
// The least-squares data of a cell's B_cell weights: Mi (6: 00, 01, 02, 11,
// 12, 22) and mu of the six faces.
template <typename WT>
__device__ __forceinline__ void load_bcell_w(const MhdBlockDev &b, int64_t o, double *mi, double *mu)
{
    const int64_t sl = b.slot;
    const WT *Wm = dweights<WT>(b) + (int64_t)(G_WBM - G_NMETRIC) * sl;
    const WT *Wu = dweights<WT>(b) + (int64_t)(G_WBU - G_NMETRIC) * sl;
#pragma unroll
    for (int n = 0; n < 6; ++n)
    {
        mi[n] = Wm[(int64_t)n * sl + o];
        mu[n] = (double)Wu[(int64_t)n * sl + o];
    }
}

__device__ __forceinline__ void bcell_from_w(const double *mi, const double *muq, const CellGeo &q, const double *pt,
                                             double *bc)
{
    const double m00 = mi[0], m01 = mi[1], m02 = mi[2], m11 = mi[3], m12 = mi[4], m22 = mi[5];
    bc[0] = bc[1] = bc[2] = 0.0;
#pragma unroll
    for (int n = 0; n < 6; ++n)
    {
        const double mu = muq[n];
        const double s0 = (n & 1) ? q.a[n][0] : -q.a[n][0];
        const double s1 = (n & 1) ? q.a[n][1] : -q.a[n][1];
        const double s2 = (n & 1) ? q.a[n][2] : -q.a[n][2];
        const double rx = mu * s0, ry = mu * s1, rz = mu * s2;
        const double w0 = m00 * rx + m01 * ry + m02 * rz;
        const double w1 = m01 * rx + m11 * ry + m12 * rz;
        const double w2 = m02 * rx + m12 * ry + m22 * rz;
        bc[0] += w0 * pt[n];
        bc[1] += w1 * pt[n];
        bc[2] += w2 * pt[n];
    }
}

template <typename WT>
__device__ __forceinline__ void bcell_from_faces(const MhdBlockDev &b, int64_t o, const CellGeo &q, const double *pt,
                                                 double *bc)
{
    double mi[6], mu[6];
    load_bcell_w<WT>(b, o, mi, mu);
    bcell_from_w(mi, mu, q, pt, bc);
}

// |A_f| and the outward unit normal of each face (zero where |A_f| = 0):
// geometry only, so a kernel updating several cases of one grid forms them
// once per cell.
struct FaceNormals
{
    double An[6], n[6][3];
};
__device__ __forceinline__ void face_normals(const CellGeo &q, FaceNormals &fn)
{
#pragma unroll
    for (int f = 0; f < 6; ++f)
    {
        const double axv = q.a[f][0], ayv = q.a[f][1], azv = q.a[f][2];
        const double An = sqrt(axv * axv + ayv * ayv + azv * azv);
        fn.An[f] = An;
        double nx = 0.0, ny = 0.0, nz = 0.0;
        if (An > 0.0)
        {
            nx = axv / An;
            ny = ayv / An;
            nz = azv / An;
            if (!(f & 1))
            {
                nx = -nx;
                ny = -ny;
                nz = -nz;
            }
        }
        fn.n[f][0] = nx;
        fn.n[f][1] = ny;
        fn.n[f][2] = nz;
    }
}

__device__ __forceinline__ double cell_dt_n(const MhdCaseConst &c, const CellGeo &q, const FaceNormals &fn, double Bx,
                                            double By, double Bz)
{
    double local = 1e100;
    const double V = fabs(q.jac);
    if (V > 0.0)
    {
        const double gm = c.gamma;
        const double rho = dmax(q.u[0], 1e-2);
        const double ux = q.u[1] / rho, uy = q.u[2] / rho, uz = q.u[3] / rho;
        const double E = q.u[4];
        const double ke = 0.5 * rho * (ux * ux + uy * uy + uz * uz);
        const double eint = E - ke;
        const double p = (gm - 1.0) * eint;
        const double cs2 = gm * p / rho;
        const double B2 = Bx * Bx + By * By + Bz * Bz;
        const double vA2 = c.inver_MA2 * B2 / rho;
        double denom = 0.0;
#pragma unroll
        for (int f = 0; f < 6; ++f)
        {
            const double An = fn.An[f];
            double term = 0.0;
            if (An > 0.0)
            {
                const double nx = fn.n[f][0], ny = fn.n[f][1], nz = fn.n[f][2];
                const double un = ux * nx + uy * ny + uz * nz;
                const double Bn = Bx * nx + By * ny + Bz * nz;
                const double vAn2 = c.inver_MA2 * (Bn * Bn) / rho;
                const double tt = cs2 + vA2;
                double disc = tt * tt - 4.0 * cs2 * vAn2;
                if (disc < 0.0)
                    disc = 0.0;
                const double cf = sqrt(0.5 * (tt + sqrt(disc)));
                term = (fabs(un) + cf) * An;
            }
            denom += term;
        }
        if (denom > 0.0)
        {
            const double dtc = c.cfl * V / denom;
            local = dmin(local, dtc);
        }
    }
    return local;
}

// Compute_Timestep's per-cell limit (LunarSolver_Timestep.cpp:8-140)
__device__ __forceinline__ double cell_dt(const MhdCaseConst &c, const CellGeo &q, double Bx, double By, double Bz)
{
    FaceNormals fn;
    face_normals(q, fn);
    return cell_dt_n(c, q, fn, Bx, By, Bz);
}

__device__ __forceinline__ void reduce_dt(double local, unsigned long long *__restrict__ dtbits, int case_id)
{
#pragma unroll
    for (int s = 32; s > 0; s >>= 1)
        local = dmin(local, __shfl_xor(local, s, 64));
    __shared__ double red[4];
    if (threadIdx.x == 0)
        red[threadIdx.y] = local;
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0)
    {
        double m = red[0];
        for (int q = 1; q < (int)blockDim.y; ++q)
            m = dmin(m, red[q]);
        if (m < 1e100)
            atomicMin(&dtbits[case_id], (unsigned long long)__double_as_longlong(m));
    }
}

__global__ __launch_bounds__(256) void k_dt(const MhdTileMap tm, const MhdCaseConst *__restrict__ cst,
                                            unsigned long long *__restrict__ dtbits, int ucur)
{
    MhdBlockDev b;
    int i, j, k;
    bool wg;
    const bool ok = tile_cell(tm, b, i, j, k, wg);
    double local = 1e100;
    if (ok)
    {
        const int64_t sl = b.slot, o = blk_off(b, i, j, k);
        const double *U = dst_ptr(b) + (int64_t)(ST_U0 + 5 * ucur) * sl;
        const double *Bc = dst_ptr(b) + (int64_t)ST_BCELL * sl;
        CellGeo q;
        load_cell(b, U, dmetric(b), o, q);
        local = cell_dt(cst[b.case_id], q, Bc[o], Bc[sl + o], Bc[2 * sl + o]);
        if (b.model == 1) // the two-fluid step also scans U_Na (MercurySolver_Timestep.cpp:118-119)
        {
            const double *UN = dst_ptr(b) + (int64_t)(ST_UNA0 + 5 * ucur) * sl;
#pragma unroll
            for (int m = 0; m < 5; ++m)
                q.u[m] = UN[m * sl + o];
            local = dmin(local, cell_dt(cst[b.case_id], q, Bc[o], Bc[sl + o], Bc[2 * sl + o]));
        }
    }
    if (wg)
        reduce_dt(local, dtbits, b.case_id);
}

// finalize: bits -> dt, abort check per case, reset for the next reduction
__global__ void k_dt_finalize(unsigned long long *dtbits, double *dt, int *err, const MhdCaseConst *cst, int ncase,
                              const int *active)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncase)
        return;
    const double v = __longlong_as_double((long long)dtbits[c]);
    if (active[c])
    {
        dt[c] = v;
        if (!(v == v) || v < cst[c].dt_floor || isinf(v))
            err[c] = 1;
    }
    dtbits[c] = (unsigned long long)__double_as_longlong(1e100);
}

// ======================================================================
// contravariant cell velocity per axis, cells [-1, N] (step scratch):
// cell_u_contra (Z4_Lunar/src/physics/LunarSolver_MagExplicit_Ideal.cpp:414-428)
// with U_plus from calc_Uplus (Z4_Lunar/src/physics/LunarSolver_Tools.cpp:112-141).
__device__ __forceinline__ void cell_ucontra_of(const CellGeo &c, double up0, double up1, double up2, double *uc)
{
    const double inv_jac = 1.0 / (fabs(c.jac) + 1.0e-30);
#pragma unroll
    for (int ax = 0; ax < 3; ++ax)
    {
        const double kx = 0.5 * (c.a[2 * ax][0] + c.a[2 * ax + 1][0]);
        const double ky = 0.5 * (c.a[2 * ax][1] + c.a[2 * ax + 1][1]);
        const double kz = 0.5 * (c.a[2 * ax][2] + c.a[2 * ax + 1][2]);
        uc[ax] = (kx * up0 + ky * up1 + kz * up2) * inv_jac;
    }
}

__device__ __forceinline__ void cell_ucontra(double gm, const CellGeo &c, double *uc)
{
    const Prim q = prim_from(c.u[0], c.u[1], c.u[2], c.u[3], c.u[4], gm);
    const bool has = q.rho_raw > 1e-20;
    cell_ucontra_of(c, has ? q.u : 0.0, has ? q.v : 0.0, has ? q.w : 0.0, uc);
}

__global__ __launch_bounds__(256) void k_ucontra(const MhdTileMap tm, const MhdCaseConst *__restrict__ cst,
                                                 int ucur)
{
    MhdBlockDev b;
    int i, j, k;
    bool wg;
    if (!tile_cell(tm, b, i, j, k, wg))
        return;
    const int64_t sl = b.slot, o = blk_off(b, i, j, k);
    CellGeo q;
    load_cell(b, dst_ptr(b) + (int64_t)(ST_U0 + 5 * ucur) * sl, dmetric(b), o, q);
    double uc[3];
    cell_ucontra(cst[b.case_id].gamma, q, uc);
    double *UC = dst_ptr(b) + (int64_t)ST_UC * sl;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax)
        UC[ax * sl + o] = uc[ax];
}

// ======================================================================
// Face EMF candidates: AssembleOneDirectionEMF_
// (Z4_Lunar/src/physics/LunarSolver_MagExplicit_Ideal.cpp:348-496).  FP64 in
// both precisions: the EMFs feed the constrained-transport curl.
// One candidate of a face from its left and right cells: bt = face-B average
// with the background along the transverse axis, bi = the induced face-B
// average, uf / ub = contravariant velocities along the face normal / the
// transverse axis (G = u_f bt - u_b B_face,tot; the result is
// 1/2 (G_L + G_R) - 1/2 max(|u_f|) (bi_R - bi_L), :430-450).
__device__ __forceinline__ double emf_pair(double btL, double btR, double biL, double biR, double ufL, double ufR,
                                           double ubL, double ubR, double bff)
{
    const double GL = ufL * btL - ubL * bff;
    const double GR = ufR * btR - ubR * bff;
    const double radius = dmax(fabs(ufL), fabs(ufR));
    return 0.5 * (GL + GR) - 0.5 * radius * (biR - biL);
}

// Face EMF candidates over plane-chunk tiles of the points [0, N]^3.  At
// point o the xi, eta and zeta faces share the right cell o; their left
// cells are o - di, o - dj, o - dk.  A cell's averages along axis a are
//   bt_a = 1/2 ((B_a + Badd_a)(X) + (B_a + Badd_a)(X + s_a)),
//   bi_a = 1/2 (B_a(X) + B_a(X + s_a)),
// each formed once per cell and shared by the faces that need it.  The
// cell o - di of plane i is cell o of plane i - 1, and the xi fluxes at
// o, o - dj, o - dk of plane i were loaded at + di on plane i - 1: both are
// carried in registers.  Every load is in bounds at every point of the range
// (stencil -1 .. +1, slots span [-g, N + g], g >= 1); each family stores on
// its own range only.
__global__ __launch_bounds__(256) void k_face_emf(const MhdTileMap tm)
{
    MhdBlockDev b;
    int i0, i1, j, k;
    bool wg;
    if (!tile_planes(tm, b, i0, i1, j, k, wg))
        return;
    const int64_t sl = b.slot;
    const int64_t di = (int64_t)b.sj * b.sk, dj = b.sk, dk = 1;
    double *S = dst_ptr(b);
    const double *Bx = S + (int64_t)ST_B * sl, *By = Bx + sl, *Bz = Bx + 2 * sl;
    const double *Ax = S + (int64_t)ST_BADD * sl, *Ay = Ax + sl, *Az = Ax + 2 * sl;
    const double *U0 = S + (int64_t)ST_UC * sl, *U1 = U0 + sl, *U2 = U0 + 2 * sl;
    double *E0 = S + (int64_t)ST_EFACE * sl, *E1 = E0 + 3 * sl, *E2 = E0 + 6 * sl;
    const bool jin = j < b.nj, kin = k < b.nk, jon = j <= b.nj, kon = k <= b.nk;
    int64_t o = blk_off(b, i0, j, k);
    // left cell of the xi face (o - di): eta / zeta averages and velocities
    double lbt1, lbt2, lbi1, lbi2, lu0, lu1, lu2;
    {
        const int64_t x = o - di;
        const double y0 = By[x], y1 = By[x + dj], z0 = Bz[x], z1 = Bz[x + dk];
        lbt1 = 0.5 * ((y0 + Ay[x]) + (y1 + Ay[x + dj]));
        lbt2 = 0.5 * ((z0 + Az[x]) + (z1 + Az[x + dk]));
        lbi1 = 0.5 * (y0 + y1);
        lbi2 = 0.5 * (z0 + z1);
        lu0 = U0[x];
        lu1 = U1[x];
        lu2 = U2[x];
    }
    // xi fluxes (B, Badd) at o, o - dj, o - dk of the current plane
    double bx = Bx[o], ax = Ax[o], bxj = Bx[o - dj], axj = Ax[o - dj], bxk = Bx[o - dk], axk = Ax[o - dk];
#pragma unroll 1
    for (int i = i0; i < i1; ++i, o += di)
    {
        // right cell o
        const double nbx = Bx[o + di], nax = Ax[o + di];
        const double y0 = By[o], y1 = By[o + dj], z0 = Bz[o], z1 = Bz[o + dk];
        const double ay0 = Ay[o], az0 = Az[o];
        const double rbt0 = 0.5 * ((bx + ax) + (nbx + nax));
        const double rbt1 = 0.5 * ((y0 + ay0) + (y1 + Ay[o + dj]));
        const double rbt2 = 0.5 * ((z0 + az0) + (z1 + Az[o + dk]));
        const double rbi0 = 0.5 * (bx + nbx), rbi1 = 0.5 * (y0 + y1), rbi2 = 0.5 * (z0 + z1);
        const double ru0 = U0[o], ru1 = U1[o], ru2 = U2[o];
        // cell o - dj (left of the eta face): xi and zeta averages
        const double nbxj = Bx[o - dj + di], naxj = Ax[o - dj + di];
        const double jz0 = Bz[o - dj], jz1 = Bz[o - dj + dk];
        const double jbt0 = 0.5 * ((bxj + axj) + (nbxj + naxj));
        const double jbt2 = 0.5 * ((jz0 + Az[o - dj]) + (jz1 + Az[o - dj + dk]));
        const double jbi0 = 0.5 * (bxj + nbxj), jbi2 = 0.5 * (jz0 + jz1);
        const double ju0 = U0[o - dj], ju1 = U1[o - dj], ju2 = U2[o - dj];
        // cell o - dk (left of the zeta face): xi and eta averages
        const double nbxk = Bx[o - dk + di], naxk = Ax[o - dk + di];
        const double ky0 = By[o - dk], ky1 = By[o - dk + dj];
        const double kbt0 = 0.5 * ((bxk + axk) + (nbxk + naxk));
        const double kbt1 = 0.5 * ((ky0 + Ay[o - dk]) + (ky1 + Ay[o - dk + dj]));
        const double kbi0 = 0.5 * (bxk + nbxk), kbi1 = 0.5 * (ky0 + ky1);
        const double ku0 = U0[o - dk], ku1 = U1[o - dk], ku2 = U2[o - dk];
        // xi face: candidates along zeta (e1) and eta (e2)
        if (i <= b.ni && jin && kin)
        {
            const double bff = bx + ax;
            const double Ge = emf_pair(lbt1, rbt1, lbi1, rbi1, lu0, ru0, lu1, ru1, bff);
            const double Gz = emf_pair(lbt2, rbt2, lbi2, rbi2, lu0, ru0, lu2, ru2, bff);
            E0[sl + o] = Gz;
            E0[2 * sl + o] = -Ge;
        }
        // eta face: candidates along xi (e0) and zeta (e2)
        if (i < b.ni && jon && kin)
        {
            const double bff = y0 + ay0;
            const double Gx = emf_pair(jbt0, rbt0, jbi0, rbi0, ju1, ru1, ju0, ru0, bff);
            const double Gz = emf_pair(jbt2, rbt2, jbi2, rbi2, ju1, ru1, ju2, ru2, bff);
            E1[o] = -Gz;
            E1[2 * sl + o] = Gx;
        }
        // zeta face: candidates along xi (e0) and eta (e1)
        if (i < b.ni && jin && kon)
        {
            const double bff = z0 + az0;
            const double Gx = emf_pair(kbt0, rbt0, kbi0, rbi0, ku2, ru2, ku0, ru0, bff);
            const double Ge = emf_pair(kbt1, rbt1, kbi1, rbi1, ku2, ru2, ku1, ru1, bff);
            E2[o] = Ge;
            E2[sl + o] = -Gx;
        }
        lbt1 = rbt1;
        lbt2 = rbt2;
        lbi1 = rbi1;
        lbi2 = rbi2;
        lu0 = ru0;
        lu1 = ru1;
        lu2 = ru2;
        bx = nbx;
        ax = nax;
        bxj = nbxj;
        axj = naxj;
        bxk = nbxk;
        axk = naxk;
    }
}

// ======================================================================
// Edge EMF with UCT upwinding: AssembleEdgeEMF_FromFaceE_Ideal_
// (Z4_Lunar/src/physics/LunarSolver_MagExplicit_Ideal.cpp:53-287).
__device__ __forceinline__ double max_abs4(const double *UCa, int64_t o0, int64_t o1, int64_t o2, int64_t o3)
{
    double a = 0.0;
    a = dmax(a, fabs(UCa[o0]));
    a = dmax(a, fabs(UCa[o1]));
    a = dmax(a, fabs(UCa[o2]));
    a = dmax(a, fabs(UCa[o3]));
    return a;
}

__global__ __launch_bounds__(256) void k_edge_emf(const MhdTileMap tm)
{
    MhdBlockDev b;
    int i, j, k;
    bool wg;
    if (!tile_cell(tm, b, i, j, k, wg))
        return;
    const int64_t sl = b.slot, o = blk_off(b, i, j, k);
    const int64_t di = (int64_t)b.sj * b.sk, dj = b.sk, dk = 1;
    double *S = dst_ptr(b);
    const double *Fx = S + (int64_t)(ST_EFACE + 0) * sl, *Fy = S + (int64_t)(ST_EFACE + 3) * sl,
                 *Fz = S + (int64_t)(ST_EFACE + 6) * sl;
    const double *Bx = S + (int64_t)ST_B * sl, *By = Bx + sl, *Bz = Bx + 2 * sl;
    const double *UCx = S + (int64_t)ST_UC * sl, *UCy = UCx + sl, *UCz = UCx + 2 * sl;
    double *E = S + (int64_t)ST_E * sl;
    const double Cuct = 0.5;
    // The point range is [0, N]^3 and the stencil reaches -1 in each axis, so
    // the three families' loads are in bounds everywhere and issue together;
    // each family stores on its own range only.
    const double Ec0 = 0.25 * (Fy[o] + Fy[o - dk] + Fz[o] + Fz[o - dj]);
    const double Ec1 = 0.25 * (Fx[sl + o] + Fx[sl + o - dk] + Fz[sl + o] + Fz[sl + o - di]);
    const double Ec2 = 0.25 * (Fx[2 * sl + o] + Fx[2 * sl + o - dj] + Fy[2 * sl + o] + Fy[2 * sl + o - di]);
    {
        const double ae = max_abs4(UCy, o, o - dj, o - dk, o - dj - dk);
        const double az = max_abs4(UCz, o, o - dj, o - dk, o - dj - dk);
        const double dBz = Bz[o] - Bz[o - dj];
        const double dBe = By[o] - By[o - dk];
        const double v = Ec0 + Cuct * (0.5 * ae * dBz - 0.5 * az * dBe);
        if (i < b.ni && j <= b.nj && k <= b.nk)
            E[o] = v;
    }
    {
        const double az = max_abs4(UCz, o, o - di, o - dk, o - di - dk);
        const double ax = max_abs4(UCx, o, o - di, o - dk, o - di - dk);
        const double dBx = Bx[o] - Bx[o - dk];
        const double dBz = Bz[o] - Bz[o - di];
        const double v = Ec1 + Cuct * (0.5 * az * dBx - 0.5 * ax * dBz);
        if (i <= b.ni && j < b.nj && k <= b.nk)
            E[sl + o] = v;
    }
    {
        const double ax = max_abs4(UCx, o, o - di, o - dj, o - di - dj);
        const double ae = max_abs4(UCy, o, o - di, o - dj, o - di - dj);
        const double dBe = By[o] - By[o - di];
        const double dBx = Bx[o] - Bx[o - dj];
        const double v = Ec2 + Cuct * (0.5 * ax * dBe - 0.5 * ae * dBx);
        if (i <= b.ni && j <= b.nj && k < b.nk)
            E[2 * sl + o] = v;
    }
}

// ======================================================================
// Face and edge EMF in one pass (a whole step's form of k_face_emf, the
// Eface sync and k_edge_emf).
//
// The pass forms every in-range face candidate from the same loads with the
// same expressions as k_face_emf, takes zero where the wall handler zeroes
// and reads the other ghost faces from Eface, which the shell pass below and
// the halo copy fill first (face_source, shell_layers in layout.h); so E is
// bit for bit what the three stages give, and only the shell's candidates
// go through memory.

// Face EMF candidates of the shell points (k_face_emf's expressions at one
// point, zero on a wall's inner slab as the wall handler leaves them): 256
// shell points per workgroup.
__global__ __launch_bounds__(256) void k_face_shell(const MhdTileMap tm)
{
    MhdBlockDev b;
    int lt;
    if (!tile_block(tm, b, lt))
        return;
    const int t = lt * 256 + (int)threadIdx.x;
    if (t >= shell_count(b))
        return;
    int i, j, k;
    shell_point(b, t, i, j, k);
    const int64_t sl = b.slot;
    const int64_t di = (int64_t)b.sj * b.sk, dj = b.sk, dk = 1;
    double *S = dst_ptr(b);
    const double *Bx = S + (int64_t)ST_B * sl, *By = Bx + sl, *Bz = Bx + 2 * sl;
    const double *Ax = S + (int64_t)ST_BADD * sl, *Ay = Ax + sl, *Az = Ax + 2 * sl;
    const double *U0 = S + (int64_t)ST_UC * sl, *U1 = U0 + sl, *U2 = U0 + 2 * sl;
    double *E0 = S + (int64_t)ST_EFACE * sl, *E1 = E0 + 3 * sl, *E2 = E0 + 6 * sl;
    const int64_t o = blk_off(b, i, j, k);
    // right cell o: averages along all three axes
    const double bx = Bx[o], ax = Ax[o], nbx = Bx[o + di], nax = Ax[o + di];
    const double y0 = By[o], y1 = By[o + dj], z0 = Bz[o], z1 = Bz[o + dk];
    const double ay0 = Ay[o], az0 = Az[o];
    const double rbt0 = 0.5 * ((bx + ax) + (nbx + nax));
    const double rbt1 = 0.5 * ((y0 + ay0) + (y1 + Ay[o + dj]));
    const double rbt2 = 0.5 * ((z0 + az0) + (z1 + Az[o + dk]));
    const double rbi0 = 0.5 * (bx + nbx), rbi1 = 0.5 * (y0 + y1), rbi2 = 0.5 * (z0 + z1);
    const double ru0 = U0[o], ru1 = U1[o], ru2 = U2[o];
    if (i <= b.ni && j < b.nj && k < b.nk)
    {
        const int64_t x = o - di;
        const double ly0 = By[x], ly1 = By[x + dj], lz0 = Bz[x], lz1 = Bz[x + dk];
        const double lbt1 = 0.5 * ((ly0 + Ay[x]) + (ly1 + Ay[x + dj]));
        const double lbt2 = 0.5 * ((lz0 + Az[x]) + (lz1 + Az[x + dk]));
        const double lbi1 = 0.5 * (ly0 + ly1), lbi2 = 0.5 * (lz0 + lz1);
        const double bff = bx + ax;
        const double Ge = emf_pair(lbt1, rbt1, lbi1, rbi1, U0[x], ru0, U1[x], ru1, bff);
        const double Gz = emf_pair(lbt2, rbt2, lbi2, rbi2, U0[x], ru0, U2[x], ru2, bff);
        const bool z = face_source(b, 0, i, j, k) == 1;
        E0[sl + o] = z ? 0.0 : Gz;
        E0[2 * sl + o] = z ? 0.0 : -Ge;
    }
    if (i < b.ni && j <= b.nj && k < b.nk)
    {
        const int64_t x = o - dj;
        const double bxj = Bx[x], nbxj = Bx[x + di], jz0 = Bz[x], jz1 = Bz[x + dk];
        const double jbt0 = 0.5 * ((bxj + Ax[x]) + (nbxj + Ax[x + di]));
        const double jbt2 = 0.5 * ((jz0 + Az[x]) + (jz1 + Az[x + dk]));
        const double jbi0 = 0.5 * (bxj + nbxj), jbi2 = 0.5 * (jz0 + jz1);
        const double bff = y0 + ay0;
        const double Gx = emf_pair(jbt0, rbt0, jbi0, rbi0, U1[x], ru1, U0[x], ru0, bff);
        const double Gz = emf_pair(jbt2, rbt2, jbi2, rbi2, U1[x], ru1, U2[x], ru2, bff);
        const bool z = face_source(b, 1, i, j, k) == 1;
        E1[o] = z ? 0.0 : -Gz;
        E1[2 * sl + o] = z ? 0.0 : Gx;
    }
    if (i < b.ni && j < b.nj && k <= b.nk)
    {
        const int64_t x = o - dk;
        const double bxk = Bx[x], nbxk = Bx[x + di], ky0 = By[x], ky1 = By[x + dj];
        const double kbt0 = 0.5 * ((bxk + Ax[x]) + (nbxk + Ax[x + di]));
        const double kbt1 = 0.5 * ((ky0 + Ay[x]) + (ky1 + Ay[x + dj]));
        const double kbi0 = 0.5 * (bxk + nbxk), kbi1 = 0.5 * (ky0 + ky1);
        const double bff = z0 + az0;
        const double Gx = emf_pair(kbt0, rbt0, kbi0, rbi0, U2[x], ru2, U0[x], ru0, bff);
        const double Ge = emf_pair(kbt1, rbt1, kbi1, rbi1, U2[x], ru2, U1[x], ru1, bff);
        const bool z = face_source(b, 2, i, j, k) == 1;
        E2[o] = z ? 0.0 : Ge;
        E2[sl + o] = z ? 0.0 : -Gx;
    }
}

// max(0, |a|, |b|, |c|, |d|) in max_abs4's order
__device__ __forceinline__ double max_abs_of4(double a, double b, double c, double d)
{
    double m = 0.0;
    m = dmax(m, fabs(a));
    m = dmax(m, fabs(b));
    m = dmax(m, fabs(c));
    m = dmax(m, fabs(d));
    return m;
}

// The candidate of a face from face_source: formed (0), zero (1) or Eface (2).
// Element idx of a slot through a uniform base and a 32-bit byte offset
// (idx < 2^29), which the compiler issues as one SGPR base plus one VGPR
// offset instead of a 64-bit address per access.
template <typename T> __device__ __forceinline__ T ld32(const T *base, uint32_t idx)
{
    return *(const T *)((const char *)base + idx * (uint32_t)sizeof(T));
}
__device__ __forceinline__ void st32(double *base, uint32_t idx, double v)
{
    *(double *)((char *)base + (idx << 3)) = v;
}

__device__ __forceinline__ double face_pick(int src, double formed, const double *F, uint32_t o)
{
    return src == 0 ? formed : (src == 1 ? 0.0 : ld32(F, o));
}

// The fused pass over padded-row plane-chunk tiles of the points [0, N]^3
// (tile_padded), three workgroups per tile, one per edge family.  Each
// family takes its four face candidates and upwind speeds from the cells
// around it in one coordinate plane: the xi edge at o from the cells o,
// o - dj, o - dk, o - dj - dk of plane i; the eta edge from o, o - dk and
// their plane-(i - 1) neighbours, the zeta edge from o, o - dj and theirs,
// which a thread marching along i carries in registers (with the candidate
// of its own zeta or eta face).  What a point needs at o - dk (cell averages,
// velocities, the candidate of the face there) is what the lane before it
// formed as its own o, and the averages along zeta need B at o + dk, which
// the lane after it loaded: lane shifts instead of loads, the same values.
// A face candidate used by two families is a different component for each,
// so splitting by family repeats only cell averages and loads (served by
// the L2 the tile's three workgroups share).
// Every load is in bounds: the stencil reaches -1 .. +1 from points of
// [-1, N] x [0, N] x [-1, N + 1] and each slot spans [-g, N + g], g >= 2.
struct EmfArrays
{
    const double *Bx, *By, *Bz, *Ax, *Ay, *Az, *U0, *U1, *U2, *F0, *F1, *F2;
    double *E;
    int64_t sl;
    uint32_t di, dj;
};

// xi edges of the planes [i0, i1) at (j, k)
__device__ __forceinline__ void emf_xi(const EmfArrays &A, const MhdBlockDev &b, int i0, int i1, int j, int k, bool own)
{
    const uint32_t di = A.di, dj = A.dj;
    const int sy0 = face_jk_bits(b, 1, j, k), sz0 = face_jk_bits(b, 2, j, k), sz1 = face_jk_bits(b, 2, j - 1, k);
    const bool store = own && j <= b.nj;
    uint32_t o = (uint32_t)blk_off(b, i0, j, k);
#pragma unroll 1
    for (int i = i0; i < i1; ++i, o += di)
    {
        const int ib = face_axis_bits(b, 1, 0, i); // eta and zeta faces share the i range
        const double yo = ld32(A.By, o), ayo = ld32(A.Ay, o), yoj = ld32(A.By, o + dj), ayoj = ld32(A.Ay, o + dj);
        const double yj = ld32(A.By, o - dj), ayj = ld32(A.Ay, o - dj);
        const double zo = ld32(A.Bz, o), azo = ld32(A.Az, o), zj = ld32(A.Bz, o - dj), azj = ld32(A.Az, o - dj);
        const double u1 = ld32(A.U1, o), u2 = ld32(A.U2, o), ju1 = ld32(A.U1, o - dj), ju2 = ld32(A.U2, o - dj);
        // eta and zeta averages of the cells o and o - dj (zeta: B at + dk from the next lane)
        const double zok = lane_next(zo), azok = lane_next(azo), zjk1 = lane_next(zj), azjk1 = lane_next(azj);
        const double bt1 = 0.5 * ((yo + ayo) + (yoj + ayoj)), bi1 = 0.5 * (yo + yoj);
        const double bt2 = 0.5 * ((zo + azo) + (zok + azok)), bi2 = 0.5 * (zo + zok);
        const double jbt1 = 0.5 * ((yj + ayj) + (yo + ayo)), jbi1 = 0.5 * (yj + yo);
        const double jbt2 = 0.5 * ((zj + azj) + (zjk1 + azjk1)), jbi2 = 0.5 * (zj + zjk1);
        // the cells o - dk and o - dj - dk: the previous lane's o and o - dj
        const double kbt1 = lane_prev(bt1), kbi1 = lane_prev(bi1), ku1 = lane_prev(u1), ku2 = lane_prev(u2);
        const double jkbt1 = lane_prev(jbt1), jkbi1 = lane_prev(jbi1), jku1 = lane_prev(ju1), jku2 = lane_prev(ju2);
        const double yk = lane_prev(yo);
        // eta face at o (left cell o - dj), the previous lane's at o - dk;
        // zeta faces at o (left o - dk) and o - dj (left o - dj - dk)
        const double Fy0 = face_pick(face_source_of(ib, sy0),
                                     -emf_pair(jbt2, bt2, jbi2, bi2, ju1, u1, ju2, u2, yo + ayo), A.F1, o);
        const double nFy0 = lane_prev(Fy0);
        const double Fz0 = face_pick(face_source_of(ib, sz0),
                                     emf_pair(kbt1, bt1, kbi1, bi1, ku2, u2, ku1, u1, zo + azo), A.F2, o);
        const double nFz0 = face_pick(face_source_of(ib, sz1),
                                      emf_pair(jkbt1, jbt1, jkbi1, jbi1, jku2, ju2, jku1, ju1, zj + azj), A.F2, o - dj);
        const double Ec0 = 0.25 * (Fy0 + nFy0 + Fz0 + nFz0);
        const double ae = max_abs_of4(u1, ju1, ku1, jku1);
        const double az = max_abs_of4(u2, ju2, ku2, jku2);
        const double v = Ec0 + 0.5 * (0.5 * ae * (zo - zj) - 0.5 * az * (yo - yk));
        if (i < b.ni && store)
            st32(A.E, o, v);
    }
}

// eta edges of the planes [i0, i1) at (j, k): cells o, o - dk and, carried,
// o - di, o - di - dk
__device__ __forceinline__ void emf_eta(const EmfArrays &A, const MhdBlockDev &b, int i0, int i1, int j, int k, bool own)
{
    const int64_t sl = A.sl;
    const uint32_t di = A.di;
    const int sx0 = face_jk_bits(b, 0, j, k), sz0 = face_jk_bits(b, 2, j, k);
    const bool store = own && j < b.nj;
    uint32_t o = (uint32_t)blk_off(b, i0 - 1, j, k);
    // plane i0 - 1: zeta average and velocity of o, the zeta face's
    // candidate along eta, B_zeta at o; the xi flux of plane i0
    double cx = ld32(A.Bx, o + di), ca = ld32(A.Ax, o + di);
    double p_bt2, p_bi2, pu0, pu2, pFz1, pBz;
    {
        const double x0 = ld32(A.Bx, o), a0 = ld32(A.Ax, o);
        const double zo = ld32(A.Bz, o), azo = ld32(A.Az, o);
        pu0 = ld32(A.U0, o);
        pu2 = ld32(A.U2, o);
        const double zok = lane_next(zo), azok = lane_next(azo);
        const double bt0 = 0.5 * ((x0 + a0) + (cx + ca)), bi0 = 0.5 * (x0 + cx);
        const double kbt0 = lane_prev(bt0), kbi0 = lane_prev(bi0), ku0 = lane_prev(pu0), ku2 = lane_prev(pu2);
        p_bt2 = 0.5 * ((zo + azo) + (zok + azok));
        p_bi2 = 0.5 * (zo + zok);
        pFz1 = face_pick(face_source_of(face_axis_bits(b, 2, 0, i0 - 1), sz0),
                         -emf_pair(kbt0, bt0, kbi0, bi0, ku2, pu2, ku0, pu0, zo + azo), A.F2 + sl, o);
        pBz = zo;
        o += di;
    }
#pragma unroll 1
    for (int i = i0; i < i1; ++i, o += di)
    {
        const int ibx = face_axis_bits(b, 0, 0, i), ib = face_axis_bits(b, 2, 0, i);
        const double nx = ld32(A.Bx, o + di), na = ld32(A.Ax, o + di);
        const double zo = ld32(A.Bz, o), azo = ld32(A.Az, o);
        const double u0 = ld32(A.U0, o), u2 = ld32(A.U2, o);
        const double zok = lane_next(zo), azok = lane_next(azo);
        const double bt0 = 0.5 * ((cx + ca) + (nx + na)), bi0 = 0.5 * (cx + nx);
        const double bt2 = 0.5 * ((zo + azo) + (zok + azok)), bi2 = 0.5 * (zo + zok);
        // the cell o - dk (and o - di - dk): the previous lane's o
        const double kbt0 = lane_prev(bt0), kbi0 = lane_prev(bi0), ku0 = lane_prev(u0), ku2 = lane_prev(u2);
        const double pku0 = lane_prev(pu0), pku2 = lane_prev(pu2), cxk = lane_prev(cx);
        // xi face at o (left cell o - di), the previous lane's at o - dk; zeta face at o
        const double Fx1 = face_pick(face_source_of(ibx, sx0),
                                     emf_pair(p_bt2, bt2, p_bi2, bi2, pu0, u0, pu2, u2, cx + ca), A.F0 + sl, o);
        const double nFx1 = lane_prev(Fx1);
        const double Fz1 = face_pick(face_source_of(ib, sz0),
                                     -emf_pair(kbt0, bt0, kbi0, bi0, ku2, u2, ku0, u0, zo + azo), A.F2 + sl, o);
        const double Ec1 = 0.25 * (Fx1 + nFx1 + Fz1 + pFz1);
        const double az = max_abs_of4(u2, pu2, ku2, pku2);
        const double ax = max_abs_of4(u0, pu0, ku0, pku0);
        const double v = Ec1 + 0.5 * (0.5 * az * (cx - cxk) - 0.5 * ax * (zo - pBz));
        if (i <= b.ni && store)
            st32(A.E + sl, o, v);
        p_bt2 = bt2;
        p_bi2 = bi2;
        pu0 = u0;
        pu2 = u2;
        pFz1 = Fz1;
        pBz = zo;
        cx = nx;
        ca = na;
    }
}

// zeta edges of the planes [i0, i1) at (j, k): cells o, o - dj and, carried,
// o - di, o - di - dj
__device__ __forceinline__ void emf_zeta(const EmfArrays &A, const MhdBlockDev &b, int i0, int i1, int j, int k, bool own)
{
    const int64_t sl = A.sl;
    const uint32_t di = A.di, dj = A.dj;
    const int sx0 = face_jk_bits(b, 0, j, k), sx2 = face_jk_bits(b, 0, j - 1, k), sy0 = face_jk_bits(b, 1, j, k);
    const bool store = own && j <= b.nj && k < b.nk;
    uint32_t o = (uint32_t)blk_off(b, i0 - 1, j, k);
    // plane i0 - 1: eta averages and velocities of o and o - dj, the eta
    // face's candidate along zeta, B_eta at o; xi fluxes of plane i0
    double cx = ld32(A.Bx, o + di), ca = ld32(A.Ax, o + di), cxj = ld32(A.Bx, o - dj + di), caj = ld32(A.Ax, o - dj + di);
    double p_bt1, p_bi1, pj_bt1, pj_bi1, pu0, pu1, pju0, pju1, pFy2, pBy;
    {
        const double x0 = ld32(A.Bx, o), a0 = ld32(A.Ax, o), xj = ld32(A.Bx, o - dj), aj = ld32(A.Ax, o - dj);
        const double yo = ld32(A.By, o), ayo = ld32(A.Ay, o), yoj = ld32(A.By, o + dj), ayoj = ld32(A.Ay, o + dj);
        const double yj = ld32(A.By, o - dj), ayj = ld32(A.Ay, o - dj);
        pu0 = ld32(A.U0, o);
        pu1 = ld32(A.U1, o);
        pju0 = ld32(A.U0, o - dj);
        pju1 = ld32(A.U1, o - dj);
        const double bt0 = 0.5 * ((x0 + a0) + (cx + ca)), bi0 = 0.5 * (x0 + cx);
        const double jbt0 = 0.5 * ((xj + aj) + (cxj + caj)), jbi0 = 0.5 * (xj + cxj);
        p_bt1 = 0.5 * ((yo + ayo) + (yoj + ayoj));
        p_bi1 = 0.5 * (yo + yoj);
        pj_bt1 = 0.5 * ((yj + ayj) + (yo + ayo));
        pj_bi1 = 0.5 * (yj + yo);
        pFy2 = face_pick(face_source_of(face_axis_bits(b, 1, 0, i0 - 1), sy0),
                         emf_pair(jbt0, bt0, jbi0, bi0, pju1, pu1, pju0, pu0, yo + ayo), A.F1 + 2 * sl, o);
        pBy = yo;
        o += di;
    }
#pragma unroll 1
    for (int i = i0; i < i1; ++i, o += di)
    {
        const int ibx = face_axis_bits(b, 0, 0, i), ib = face_axis_bits(b, 1, 0, i);
        const double nx = ld32(A.Bx, o + di), na = ld32(A.Ax, o + di);
        const double nxj = ld32(A.Bx, o - dj + di), naj = ld32(A.Ax, o - dj + di);
        const double yo = ld32(A.By, o), ayo = ld32(A.Ay, o), yoj = ld32(A.By, o + dj), ayoj = ld32(A.Ay, o + dj);
        const double yj = ld32(A.By, o - dj), ayj = ld32(A.Ay, o - dj);
        const double u0 = ld32(A.U0, o), u1 = ld32(A.U1, o), ju0 = ld32(A.U0, o - dj), ju1 = ld32(A.U1, o - dj);
        const double bt0 = 0.5 * ((cx + ca) + (nx + na)), bi0 = 0.5 * (cx + nx);
        const double jbt0 = 0.5 * ((cxj + caj) + (nxj + naj)), jbi0 = 0.5 * (cxj + nxj);
        const double bt1 = 0.5 * ((yo + ayo) + (yoj + ayoj)), bi1 = 0.5 * (yo + yoj);
        const double jbt1 = 0.5 * ((yj + ayj) + (yo + ayo)), jbi1 = 0.5 * (yj + yo);
        // xi faces at o and o - dj (left o - dj - di); eta face at o
        const double Fx2 = face_pick(face_source_of(ibx, sx0),
                                     -emf_pair(p_bt1, bt1, p_bi1, bi1, pu0, u0, pu1, u1, cx + ca), A.F0 + 2 * sl, o);
        const double nFx2 = face_pick(face_source_of(ibx, sx2),
                                      -emf_pair(pj_bt1, jbt1, pj_bi1, jbi1, pju0, ju0, pju1, ju1, cxj + caj),
                                      A.F0 + 2 * sl, o - dj);
        const double Fy2 = face_pick(face_source_of(ib, sy0),
                                     emf_pair(jbt0, bt0, jbi0, bi0, ju1, u1, ju0, u0, yo + ayo), A.F1 + 2 * sl, o);
        const double Ec2 = 0.25 * (Fx2 + nFx2 + Fy2 + pFy2);
        const double ax = max_abs_of4(u0, pu0, ju0, pju0);
        const double ae = max_abs_of4(u1, pu1, ju1, pju1);
        const double v = Ec2 + 0.5 * (0.5 * ax * (yo - pBy) - 0.5 * ae * (cx - cxj));
        if (i <= b.ni && store)
            st32(A.E + 2 * sl, o, v);
        p_bt1 = bt1;
        p_bi1 = bi1;
        pj_bt1 = jbt1;
        pj_bi1 = jbi1;
        pu0 = u0;
        pu1 = u1;
        pju0 = ju0;
        pju1 = ju1;
        pFy2 = Fy2;
        pBy = yo;
        cx = nx;
        ca = na;
        cxj = nxj;
        caj = naj;
    }
}

__global__ __launch_bounds__(256) void k_emf(const MhdTileMap tm)
{
    MhdBlockDev b;
    int i0, i1, j, k, fam = 0;
    bool own;
    if (!tile_padded(tm, b, i0, i1, j, k, own, 3, fam)) // uniform over the workgroup
        return;
    // 32-bit element indices within a slot (a slot is < 2^29 elements): with
    // the slot bases uniform, the loads take an SGPR base and one VGPR offset
    EmfArrays A;
    A.sl = b.slot;
    A.di = (uint32_t)b.sj * (uint32_t)b.sk;
    A.dj = (uint32_t)b.sk;
    double *S = dst_ptr(b);
    A.Bx = S + (int64_t)ST_B * A.sl;
    A.By = A.Bx + A.sl;
    A.Bz = A.Bx + 2 * A.sl;
    A.Ax = S + (int64_t)ST_BADD * A.sl;
    A.Ay = A.Ax + A.sl;
    A.Az = A.Ax + 2 * A.sl;
    A.U0 = S + (int64_t)ST_UC * A.sl;
    A.U1 = A.U0 + A.sl;
    A.U2 = A.U0 + 2 * A.sl;
    A.F0 = S + (int64_t)ST_EFACE * A.sl;
    A.F1 = A.F0 + 3 * A.sl;
    A.F2 = A.F0 + 6 * A.sl;
    A.E = S + (int64_t)ST_E * A.sl;
    if (fam == 0)
        emf_xi(A, b, i0, i1, j, k, own);
    else if (fam == 1)
        emf_eta(A, b, i0, i1, j, k, own);
    else
        emf_zeta(A, b, i0, i1, j, k, own);
}

// ======================================================================
// B_face += dt * (-curl E): CurlEdgeToFace with multiper -1
// (Z4_Lunar/src/operators/CTOperators.cpp:6-56) and ApplyUpdate_Euler_
// (Z4_Lunar/src/physics/LunarSolver_Advance.cpp:61-122).
__global__ __launch_bounds__(256) void k_curl_update(const MhdTileMap tm, const double *__restrict__ dts)
{
    MhdBlockDev b;
    int i, j, k;
    bool wg;
    if (!tile_cell(tm, b, i, j, k, wg))
        return;
    const int64_t sl = b.slot, o = blk_off(b, i, j, k);
    const int64_t di = (int64_t)b.sj * b.sk, dj = b.sk, dk = 1;
    double *S = dst_ptr(b);
    const double *Ex = S + (int64_t)ST_E * sl, *Ey = Ex + sl, *Ez = Ex + 2 * sl;
    double *Bx = S + (int64_t)ST_B * sl, *By = Bx + sl, *Bz = Bx + 2 * sl;
    const double dt = dts[b.case_id];
    const double mult = -1.0;
    // stencil o .. o + 1 in each axis over the points [0, N]^3: in bounds
    // everywhere, so the loads of the three families issue together
    const double rx = mult * ((Ey[o] - Ey[o + dk]) + (Ez[o + dj] - Ez[o]));
    const double ry = mult * ((Ex[o + dk] - Ex[o]) + (Ez[o] - Ez[o + di]));
    const double rz = mult * ((Ex[o] - Ex[o + dj]) + (Ey[o + di] - Ey[o]));
    const double bx = Bx[o], by = By[o], bz = Bz[o];
    const double nbx = bx + dt * rx, nby = by + dt * ry, nbz = bz + dt * rz;
    const bool sx = i <= b.ni && j < b.nj, sy = i < b.ni && j <= b.nj, sz = i < b.ni && j < b.nj;
    if (sx && k < b.nk)
        Bx[o] = nbx;
    if (sy && k < b.nk)
        By[o] = nby;
    if (sz && k <= b.nk)
        Bz[o] = nbz;
    // solid walls on the k faces (folds_kface_b): the ghost faces copy the
    // inner layer -- the cell layer for B_xi, B_eta, the wall face for B_zeta
    // (Boundary_Build.cpp:29-33 with the Solid_Surface handler)
    if (tm.fold && folds_kface_b(b, 4) && k == 0)
        for (int l = 1; l <= b.g; ++l)
        {
            if (sx)
                Bx[o - l] = nbx;
            if (sy)
                By[o - l] = nby;
            if (sz)
                Bz[o - l] = nbz;
        }
    if (tm.fold && folds_kface_b(b, 5))
        for (int l = 1; l <= b.g; ++l)
        {
            if (k == b.nk - 1 && sx)
                Bx[o + l] = nbx;
            if (k == b.nk - 1 && sy)
                By[o + l] = nby;
            if (k == b.nk && sz)
                Bz[o + l] = nbz;
        }
}

// ======================================================================
// Fluid: Rusanov fluxes, divergence, J x B source and Euler update fused:
// Scheme_U_ + AddSourceToRHS_Fluid + ApplyUpdate_Euler_
// (Z4_Lunar/src/physics/LunarSolver_FluidRHS.cpp:3-201,528-568;
//  Z4_Lunar/src/physics/LunarSolver_Advance.cpp:61-122).
// U is read from buffer `ucur` and written to buffer 1-ucur.
template <typename R, int RECON>
__device__ __forceinline__ void face_flux(const double *__restrict__ U, int64_t sl, const double *__restrict__ A,
                                          int64_t oL, int64_t oR, int64_t step, double gm, double *F)
{
    double WL[5], WR[5];
    {
        const Prim a = prim_of<sizeof(R) == 8>(U, sl, oL, gm), c = prim_of<sizeof(R) == 8>(U, sl, oR, gm);
        WL[0] = a.rho_raw; WL[1] = a.u; WL[2] = a.v; WL[3] = a.w; WL[4] = a.p;
        WR[0] = c.rho_raw; WR[1] = c.u; WR[2] = c.v; WR[3] = c.w; WR[4] = c.p;
    }
    if (RECON == 1)
    {
        const Prim a = prim_of<sizeof(R) == 8>(U, sl, oL - step, gm),
                   c = prim_of<sizeof(R) == 8>(U, sl, oR + step, gm);
        const double WLL[5] = {a.rho_raw, a.u, a.v, a.w, a.p};
        const double WRR[5] = {c.rho_raw, c.u, c.v, c.w, c.p};
        double L2[5], R2[5];
#pragma unroll
        for (int m = 0; m < 5; ++m)
        {
            if constexpr (sizeof(R) == 8)
            {
                L2[m] = WL[m] + 0.5 * minmod(WL[m] - WLL[m], WR[m] - WL[m]);
                R2[m] = WR[m] - 0.5 * minmod(WR[m] - WL[m], WRR[m] - WR[m]);
            }
            else // FP32 slopes, added to the FP64 states
            {
                L2[m] = WL[m] + (double)(0.5f * minmod((float)(WL[m] - WLL[m]), (float)(WR[m] - WL[m])));
                R2[m] = WR[m] - (double)(0.5f * minmod((float)(WR[m] - WL[m]), (float)(WRR[m] - WR[m])));
            }
        }
        if (L2[0] > 0.0 && L2[4] > 0.0 && R2[0] > 0.0 && R2[4] > 0.0)
        {
#pragma unroll
            for (int m = 0; m < 5; ++m)
            {
                WL[m] = L2[m];
                WR[m] = R2[m];
            }
        }
    }
    const int64_t of = oR;
    if constexpr (sizeof(R) == 8)
    {
        rusanov<double>(gm, WL, WR, A[of], A[sl + of], A[2 * sl + of], F);
    }
    else
    {
        rusanov_mixed(gm, WL, WR, A[of], A[sl + of], A[2 * sl + of], F);
    }
}

template <typename R, int RECON>
__global__ __launch_bounds__(256) void k_fluid(const MhdTileMap tm, const MhdCaseConst *__restrict__ cst,
                                               const double *__restrict__ dts, int ucur)
{
    MhdBlockDev b;
    int i, j, k;
    bool wg;
    if (!tile_cell(tm, b, i, j, k, wg))
        return;
    const MhdCaseConst &c = cst[b.case_id];
    const double gm = c.gamma;
    const int64_t sl = b.slot, o = blk_off(b, i, j, k);
    const int64_t step[3] = {(int64_t)b.sj * b.sk, (int64_t)b.sk, 1};
    double *S = dst_ptr(b);
    const double *U = S + (int64_t)(ST_U0 + 5 * ucur) * sl;
    double *Un = S + (int64_t)(ST_U0 + 5 * (1 - ucur)) * sl;
    const double *G = dmetric(b);
    double rhs[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const double invJ = 1.0 / G[(int64_t)G_JAC * sl + o];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax)
    {
        const double *A = G + (int64_t)(G_JD + 3 * ax) * sl;
        double Flo[5], Fhi[5];
        face_flux<R, RECON>(U, sl, A, o - step[ax], o, step[ax], gm, Flo);
        face_flux<R, RECON>(U, sl, A, o, o + step[ax], step[ax], gm, Fhi);
#pragma unroll
        for (int m = 0; m < 5; ++m)
            rhs[m] -= (Fhi[m] - Flo[m]) * invJ;
    }
    // J x B (Hall_Num_Limiter weighting)
    const Prim q = prim_of<sizeof(R) == 8>(U, sl, o, gm);
    const double n_true = dmax(q.rho_raw, 0.0) / c.M_H * c.M_ref;
    const double ne_eff = sqrt(n_true * n_true + c.ne_floor * c.ne_floor);
    const double wgt = n_true / ne_eff;
    const double fc = c.momentum_hall_coeff * wgt;
    const double *Bc = S + (int64_t)ST_BCELL * sl, *Jc = S + (int64_t)ST_JCELL * sl;
    const double B0 = Bc[o], B1 = Bc[sl + o], B2 = Bc[2 * sl + o];
    const double J0 = Jc[o], J1 = Jc[sl + o], J2 = Jc[2 * sl + o];
    const double jx = J1 * B2 - J2 * B1;
    const double jy = J2 * B0 - J0 * B2;
    const double jz = J0 * B1 - J1 * B0;
    rhs[1] += fc * jx;
    rhs[2] += fc * jy;
    rhs[3] += fc * jz;
    rhs[4] += fc * (jx * q.u + jy * q.v + jz * q.w);
    const double dt = dts[b.case_id];
#pragma unroll
    for (int m = 0; m < 5; ++m)
        Un[m * sl + o] = U[m * sl + o] + dt * rhs[m];
}

// The first-order fluid update of k_fluid in column tiles.  Per plane i the
// workgroup holds flux_cell of its 64 x COL_TJ cells and their one-cell halo
// in j and k (LDS); a thread forms flux_cell of its own column's cell i + 1
// as it closes cell i, which is that cell's entry for the next plane, so only
// the halo is formed by the fill.  Each i-face flux is computed once (the
// upper face of cell i is carried as the lower face of cell i + 1), and the
// update adds the terms in k_fluid's order, ((0 - t_i) - t_j) - t_k, then the
// J x B source, as each term is formed: every face flux is rusanov_pre of the
// same two cells as in k_fluid, so the result is bit for bit k_fluid's, with
// about a fifth of its divisions and only the cell and its lower i-face
// carried from plane to plane.
template <typename R>
__device__ __forceinline__ void col_put(double (&P)[7][COL_TJ + 2][COL_TK + 2], int pj, int pk, const FluxCell &f)
{
    P[0][pj][pk] = f.rho;
    P[1][pj][pk] = f.u;
    P[2][pj][pk] = f.v;
    P[3][pj][pk] = f.w;
    P[4][pj][pk] = f.p;
    P[5][pj][pk] = f.e;
    P[6][pj][pk] = f.c;
}
__device__ __forceinline__ FluxCell col_get(const double (&P)[7][COL_TJ + 2][COL_TK + 2], int pj, int pk)
{
    FluxCell f;
    f.rho = P[0][pj][pk];
    f.u = P[1][pj][pk];
    f.v = P[2][pj][pk];
    f.w = P[3][pj][pk];
    f.p = P[4][pj][pk];
    f.e = P[5][pj][pk];
    f.c = P[6][pj][pk];
    return f;
}
// halo entry h of a column tile's LDS plane: the rows j0 - 1 and j0 + COL_TJ
// (every k of the tile and its halo), then the k halo of the tile's rows
__device__ __forceinline__ void col_halo(int h, int &pj, int &pk)
{
    constexpr int PJ = COL_TJ + 2, PK = COL_TK + 2;
    if (h < 2 * PK)
    {
        pj = h < PK ? 0 : PJ - 1;
        pk = h < PK ? h : h - PK;
    }
    else
    {
        pj = 1 + (h - 2 * PK) / 2;
        pk = (h & 1) ? PK - 1 : 0;
    }
}
constexpr int COL_NHALO = 2 * (COL_TK + 2) + 2 * COL_TJ;

template <typename R>
__global__ __launch_bounds__(64 * COL_TJ) __attribute__((amdgpu_waves_per_eu(3))) void k_fluid_col(const MhdTileMap tm, const MhdCaseConst *__restrict__ cst,
                                                           const double *__restrict__ dts, int ucur)
{
    // FP64: two LDS planes used alternately (the next plane is written while
    // the current one is read, one barrier per plane); the mixed step's
    // lighter threads keep one plane and 4 workgroups per CU
    constexpr int NB = sizeof(R) == 8 ? 2 : 1;
    __shared__ double P2[NB][7][COL_TJ + 2][COL_TK + 2];
    MhdBlockDev b;
    int i0, j0, k0;
    if (!tile_col(tm, b, i0, j0, k0)) // uniform over the workgroup
        return;
    const int tx = (int)threadIdx.x, ty = (int)threadIdx.y, tid = ty * 64 + tx;
    const int j = j0 + ty, k = k0 + tx;
    const bool act = j < b.nj && k < b.nk;
    // the thread's LDS entry is a cell (interior, or the ghost layer at N)
    const bool mine = j <= b.nj && k <= b.nk;
    const int i1 = min(i0 + tm.ci, b.ni);
    const MhdCaseConst &c = cst[b.case_id];
    const double gm = c.gamma;
    const int64_t sl = b.slot, di = (int64_t)b.sj * b.sk, dj = b.sk;
    double *S = dst_ptr(b);
    const double *U = S + (int64_t)(ST_U0 + 5 * ucur) * sl;
    double *Un = S + (int64_t)(ST_U0 + 5 * (1 - ucur)) * sl;
    const double *G = dmetric(b);
    const double *Ax = G + (int64_t)G_JD * sl, *Ay = Ax + 3 * sl, *Az = Ax + 6 * sl;
    const double *Bc = S + (int64_t)ST_BCELL * sl, *Jc = S + (int64_t)ST_JCELL * sl;
    const double dt = dts[b.case_id];
    // the J x B weight of k_fluid (Hall_Num_Limiter)
    const double fcoef = c.momentum_hall_coeff;
    int64_t o = blk_off(b, i0, mine ? j : j0, mine ? k : k0); // cell of plane `pl` in the loop
    // plane i0: the thread's own cell, its lower i-face flux, the halo
    FluxCell cur{};
    double Flo[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (mine)
        cur = flux_cell<R>(U, sl, o, gm);
    if (act)
        rusanov_pre<R>(flux_cell<R>(U, sl, o - di, gm), cur, Ax[o], Ax[sl + o], Ax[2 * sl + o], Flo);
    if (mine)
        col_put<R>(P2[0], ty + 1, tx + 1, cur);
    if (tid < COL_NHALO)
    {
        int pj, pk;
        col_halo(tid, pj, pk);
        const int jj = j0 - 1 + pj, kk = k0 - 1 + pk;
        if (jj <= b.nj && kk <= b.nk)
            col_put<R>(P2[0], pj, pk, flux_cell<R>(U, sl, blk_off(b, i0, jj, kk), gm));
    }
    __syncthreads();
    // the k-face walls and far field this kernel applies itself (folds_kface_u)
    const bool fu_lo = tm.fold && folds_kface_u(b, 4), fu_hi = tm.fold && folds_kface_u(b, 5);
    const bool solid_lo = fu_lo && b.bc[4] == BC_SOLID, solid_hi = fu_hi && b.bc[5] == BC_SOLID;
    for (int pl = i0; pl < i1; ++pl, o += di)
    {
        // the next plane's entry of this thread: cell pl + 1 (the ghost layer at N at the end)
        FluxCell nxt{};
        if (mine)
            nxt = flux_cell<R>(U, sl, o + di, gm);
        const int cb = (pl - i0) & (NB - 1);
        auto &P = P2[cb];
        if (NB == 2 && pl + 1 < i1) // the LDS plane pl + 1 (its buffer's last reads were before the previous barrier)
        {
            if (mine)
                col_put<R>(P2[cb ^ (NB - 1)], ty + 1, tx + 1, nxt);
            if (tid < COL_NHALO)
            {
                int pj, pk;
                col_halo(tid, pj, pk);
                const int jj = j0 - 1 + pj, kk = k0 - 1 + pk;
                if (jj <= b.nj && kk <= b.nk)
                    col_put<R>(P2[cb ^ (NB - 1)], pj, pk, flux_cell<R>(U, sl, blk_off(b, pl + 1, jj, kk), gm));
            }
        }
        double unc[5] = {0.0, 0.0, 0.0, 0.0, 0.0}; // the closed cell's new state
        if (act)
        {
            double r[5], F[5], Fl[5], Fh[5];
            const double invJ = 1.0 / G[(int64_t)G_JAC * sl + o];
            rusanov_pre<R>(cur, nxt, Ax[o + di], Ax[sl + o + di], Ax[2 * sl + o + di], F);
#pragma unroll
            for (int m = 0; m < 5; ++m)
            {
                r[m] = 0.0;
                r[m] -= (F[m] - Flo[m]) * invJ;
                Flo[m] = F[m];
            }
            rusanov_pre<R>(col_get(P, ty, tx + 1), cur, Ay[o], Ay[sl + o], Ay[2 * sl + o], Fl);
            rusanov_pre<R>(cur, col_get(P, ty + 2, tx + 1), Ay[o + dj], Ay[sl + o + dj], Ay[2 * sl + o + dj], Fh);
#pragma unroll
            for (int m = 0; m < 5; ++m)
                r[m] -= (Fh[m] - Fl[m]) * invJ;
            rusanov_pre<R>(col_get(P, ty + 1, tx), cur, Az[o], Az[sl + o], Az[2 * sl + o], Fl);
            rusanov_pre<R>(cur, col_get(P, ty + 1, tx + 2), Az[o + 1], Az[sl + o + 1], Az[2 * sl + o + 1], Fh);
#pragma unroll
            for (int m = 0; m < 5; ++m)
                r[m] -= (Fh[m] - Fl[m]) * invJ;
            const double n_true = dmax(cur.rho, 0.0) / c.M_H * c.M_ref;
            const double ne_eff = sqrt(n_true * n_true + c.ne_floor * c.ne_floor);
            const double fc = fcoef * (n_true / ne_eff);
            const double B0 = Bc[o], B1 = Bc[sl + o], B2 = Bc[2 * sl + o];
            const double J0 = Jc[o], J1 = Jc[sl + o], J2 = Jc[2 * sl + o];
            const double jx = J1 * B2 - J2 * B1;
            const double jy = J2 * B0 - J0 * B2;
            const double jz = J0 * B1 - J1 * B0;
            r[1] += fc * jx;
            r[2] += fc * jy;
            r[3] += fc * jz;
            r[4] += fc * (jx * cur.u + jy * cur.v + jz * cur.w);
#pragma unroll
            for (int m = 0; m < 5; ++m)
                unc[m] = U[m * sl + o] + dt * r[m];
        }
        // store the closed cell (the lanes of a wavefront together)
        {
            // a solid k wall's first layer takes the wall state of its
            // neighbour's new value (OP_SOLID_FIRST), its ghosts that of the
            // first layer (OP_SOLID_GHOST); far-field ghosts take q_inf (OP_QINF)
            double nb[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
            if (solid_lo)
            {
#pragma unroll
                for (int m = 0; m < 5; ++m)
                    nb[m] = __shfl_down(unc[m], 1, 64);
            }
            if (solid_hi)
            {
#pragma unroll
                for (int m = 0; m < 5; ++m)
                {
                    const double v = __shfl_up(unc[m], 1, 64);
                    if (k == b.nk - 1)
                        nb[m] = v;
                }
            }
            if (act)
            {
                const bool wall_lo = solid_lo && k == 0, wall_hi = solid_hi && k == b.nk - 1;
                if (wall_lo || wall_hi)
                {
                    double w[5], wg[5];
                    solid_state(gm, nb, w);
                    solid_state(gm, w, wg);
                    const int64_t out = wall_lo ? -1 : 1;
#pragma unroll
                    for (int m = 0; m < 5; ++m)
                    {
                        Un[m * sl + o] = w[m];
                        for (int l = 1; l <= b.g; ++l)
                            Un[m * sl + o + out * l] = wg[m];
                    }
                }
                else
                {
#pragma unroll
                    for (int m = 0; m < 5; ++m)
                        Un[m * sl + o] = unc[m];
                }
                const bool far_lo = fu_lo && !solid_lo && k == 0, far_hi = fu_hi && !solid_hi && k == b.nk - 1;
                if (far_lo || far_hi)
                {
                    const int64_t out = far_lo ? -1 : 1;
#pragma unroll
                    for (int m = 0; m < 5; ++m)
                        for (int l = 1; l <= b.g; ++l)
                            Un[m * sl + o + out * l] = c.qinf[m];
                }
            }
        }
        if (NB == 1 && pl + 1 < i1) // one plane: every read of plane pl, then plane pl + 1
        {
            __syncthreads();
            if (mine)
                col_put<R>(P, ty + 1, tx + 1, nxt);
            if (tid < COL_NHALO)
            {
                int pj, pk;
                col_halo(tid, pj, pk);
                const int jj = j0 - 1 + pj, kk = k0 - 1 + pk;
                if (jj <= b.nj && kk <= b.nk)
                    col_put<R>(P, pj, pk, flux_cell<R>(U, sl, blk_off(b, pl + 1, jj, kk), gm));
            }
        }
        if (pl + 1 < i1)
            __syncthreads();
        cur = nxt;
    }
}

// ======================================================================
// Derived fields, part 1: contravariant velocities of cells [-1, N] and, on
// interior cells, B_cell (calc_Bcell, Z4_Lunar/src/physics/LunarSolver_Tools.cpp:
// 163-236) fused with the next step's CFL limit (dt depends on U after the
// Ucell sync and on this B_cell; nothing changes them before the next step).
// Each result is stored as soon as it is known, which keeps few values live.
template <typename WT>
__global__ __launch_bounds__(256) void k_derived(const MhdTileMap tm, const MhdCaseConst *__restrict__ cst,
                                                 unsigned long long *__restrict__ dtbits, int ucur)
{
    MhdBlockDev b;
    int i, j, k;
    bool wg;
    const bool ok = tile_cell(tm, b, i, j, k, wg);
    double local = 1e100;
    if (ok)
    {
        const int64_t sl = b.slot, o = blk_off(b, i, j, k);
        const int64_t di = (int64_t)b.sj * b.sk, dj = b.sk, dk = 1;
        double *S = dst_ptr(b);
        CellGeo q;
        load_cell(b, S + (int64_t)(ST_U0 + 5 * ucur) * sl, dmetric(b), o, q);
        {
            double uc[3];
            cell_ucontra(cst[b.case_id].gamma, q, uc);
            double *UC = S + (int64_t)ST_UC * sl;
#pragma unroll
            for (int a = 0; a < 3; ++a)
                UC[a * sl + o] = uc[a];
        }
        // B_cell on the interior cells; its loads (face fluxes o .. o + 1, the
        // cell's weights) are in bounds on the ring too, so they issue with
        // the velocity's and only the stores and dt are limited to the interior
        const bool inner = i >= 0 && j >= 0 && k >= 0 && i < b.ni && j < b.nj && k < b.nk;
        {
            const double *Bx = S + (int64_t)ST_B * sl, *By = Bx + sl, *Bz = Bx + 2 * sl;
            const double *Ax = S + (int64_t)ST_BADD * sl, *Ay = Ax + sl, *Az = Ax + 2 * sl;
            const double pt[6] = {-Bx[o] - Ax[o],          Bx[o + di] + Ax[o + di], -By[o] - Ay[o],
                                  By[o + dj] + Ay[o + dj], -Bz[o] - Az[o],          Bz[o + dk] + Az[o + dk]};
            double bc[3];
            bcell_from_faces<WT>(b, o, q, pt, bc);
            if (inner)
            {
                double *Bc = S + (int64_t)ST_BCELL * sl;
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    Bc[a * sl + o] = bc[a];
                local = cell_dt(cst[b.case_id], q, bc[0], bc[1], bc[2]);
            }
        }
    }
    if (wg)
        reduce_dt(local, dtbits, b.case_id);
}

// k_derived over a geometry-shared map with member-looping tiles: the
// cell's geometry (Jac, the six face area vectors, the B_cell least-squares
// data) is loaded once and the K cases sharing it are updated in turn with
// k_derived's arithmetic, each case's dt reduced over the workgroup.
template <typename WT>
__global__ __launch_bounds__(256) void k_derived_k(const MhdTileMap tm, const MhdCaseConst *__restrict__ cst,
                                                   unsigned long long *__restrict__ dtbits, int ucur)
{
    MhdBlockDev b;
    int i, j, k, m0 = 0, K = 0;
    bool wg;
    const bool ok = tile_cell_k(tm, b, i, j, k, wg, m0, K);
    if (!wg) // a whole workgroup
        return;
    const int64_t sl = b.slot, o = ok ? blk_off(b, i, j, k) : 0;
    const int64_t di = (int64_t)b.sj * b.sk, dj = b.sk, dk = 1;
    CellGeo q;
    FaceNormals fn;
    double mi[6], mu[6];
    if (ok)
    {
        load_cell_geo(b, dmetric(b), o, q);
        load_bcell_w<WT>(b, o, mi, mu);
        face_normals(q, fn);
    }
    const bool inner = ok && i >= 0 && j >= 0 && k >= 0 && i < b.ni && j < b.nj && k < b.nk;
#pragma unroll 1
    for (int m = 0; m < K; ++m)
    {
        const MhdBlockDev bm = tm.blks[tm.members[m0 + m]];
        double local = 1e100;
        if (ok)
        {
            double *S = dst_ptr(bm);
            const double *U = S + (int64_t)(ST_U0 + 5 * ucur) * sl;
#pragma unroll
            for (int c = 0; c < 5; ++c)
                q.u[c] = U[c * sl + o];
            {
                double uc[3];
                cell_ucontra(cst[bm.case_id].gamma, q, uc);
                double *UC = S + (int64_t)ST_UC * sl;
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    UC[a * sl + o] = uc[a];
            }
            const double *Bx = S + (int64_t)ST_B * sl, *By = Bx + sl, *Bz = Bx + 2 * sl;
            const double *Ax = S + (int64_t)ST_BADD * sl, *Ay = Ax + sl, *Az = Ax + 2 * sl;
            const double pt[6] = {-Bx[o] - Ax[o],          Bx[o + di] + Ax[o + di], -By[o] - Ay[o],
                                  By[o + dj] + Ay[o + dj], -Bz[o] - Az[o],          Bz[o + dk] + Az[o + dk]};
            double bc[3];
            bcell_from_w(mi, mu, q, pt, bc);
            if (inner)
            {
                double *Bc = S + (int64_t)ST_BCELL * sl;
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    Bc[a * sl + o] = bc[a];
                local = cell_dt_n(cst[bm.case_id], q, fn, bc[0], bc[1], bc[2]);
            }
        }
        reduce_dt(local, dtbits, bm.case_id);
        __syncthreads(); // the reduction's shared slots are reused for the next case
    }
}

// Derived fields, part 2, interior cells: Calc_J_Edge then calc_Jcell
// (Z4_Lunar/src/physics/LunarSolver_Tools.cpp:777-891) in one pass.  A J
// edge is alpha CurlAdj of the beta-weighted face fluxes
// (CTOperators::CurlAdjFaceToEdge, Z4_Lunar/src/operators/CTOperators.cpp:
// 59-110); J_cell is sum_n W[c][n] J_n over the cell's 12 edges, with
// Jcell_from_Jedge_w formed in place: W[c][n] = Mi[c] . e_n for the active
// edges, e_n = node(end) - node(start), with the host builder's arithmetic,
// accumulated in the reference's order n = 0..11.
//
// The cube's nodes are nd[0..3] = o + {0, dj, dk, dj + dk} on plane i and
// nd[4..7] the same + di; edge n runs from node kEs[n] to kEe[n] (xi edges
// 0-3, eta 4-7, zeta 8-11).  Over padded-row plane chunks (tile_padded) of
// the cells [0, N)^3, a thread marching along i forms the J edges of plane
// i + 1 at (j, k) and (j + 1, k) -- the xi and zeta edges there, the eta
// edge at (j, k) -- and carries those of plane i; the edges and nodes at
// + dk are what the lane after it formed, the beta B products at - dk what
// the lane before it formed.  Each J edge is the value the per-edge form
// stores, by the same operations, so J_cell is bit for bit as from stored
// edges.  Loads reach planes i0 - 1 .. N, rows -1 .. N and k = -1 .. N (the
// padded lanes).
struct JEdges
{
    double jx, jxj, jy, jz, jzj; // xi at q, q + dj; eta at q; zeta at q, q + dj
    double py, pz, pyj;          // beta B products at q (eta, zeta faces) and q + dj (eta face)
};

template <typename WT> struct JArrays
{
    const double *Bx, *By, *Bz;
    const WT *bx, *by, *bz, *ax, *ay, *az;
};

// the J edges of plane q given the beta B products of plane q - di
template <typename WT>
__device__ __forceinline__ JEdges jedges_of(const JArrays<WT> &A, uint32_t q, uint32_t dj, double ppy, double ppz,
                                            double ppyj)
{
    JEdges r;
    const double px = ld32(A.bx, q) * ld32(A.Bx, q);
    r.py = ld32(A.by, q) * ld32(A.By, q);
    r.pz = ld32(A.bz, q) * ld32(A.Bz, q);
    const double pxj = ld32(A.bx, q + dj) * ld32(A.Bx, q + dj);
    r.pyj = ld32(A.by, q + dj) * ld32(A.By, q + dj);
    const double pzj = ld32(A.bz, q + dj) * ld32(A.Bz, q + dj);
    const double pxm = ld32(A.bx, q - dj) * ld32(A.Bx, q - dj), pzm = ld32(A.bz, q - dj) * ld32(A.Bz, q - dj);
    const double pxk = lane_prev(px), pyk = lane_prev(r.py), pyjk = lane_prev(r.pyj);
    // xi: (by By(-dk) - by By) + (bz Bz - bz Bz(-dj)); eta: (bx Bx - bx Bx(-dk)) +
    // (bz Bz(-di) - bz Bz); zeta: (bx Bx(-dj) - bx Bx) + (by By - by By(-di))
    r.jx = (1.0 * ((pyk - r.py) + (r.pz - pzm))) * ld32(A.ax, q);
    r.jxj = (1.0 * ((pyjk - r.pyj) + (pzj - r.pz))) * ld32(A.ax, q + dj);
    r.jy = (1.0 * ((px - pxk) + (ppz - r.pz))) * ld32(A.ay, q);
    r.jz = (1.0 * ((pxm - px) + (r.py - ppy))) * ld32(A.az, q);
    r.jzj = (1.0 * ((px - pxj) + (r.pyj - ppyj))) * ld32(A.az, q + dj);
    return r;
}

template <typename WT> __global__ __launch_bounds__(256) void k_jcell(const MhdTileMap tm)
{
    constexpr int kEs[12] = {0, 1, 2, 3, 0, 4, 2, 6, 0, 4, 1, 5};
    constexpr int kEe[12] = {4, 5, 6, 7, 1, 5, 3, 7, 2, 6, 3, 7};
    MhdBlockDev b;
    int i0, i1, j, k, sub;
    bool own;
    if (!tile_padded(tm, b, i0, i1, j, k, own, 1, sub)) // uniform over the workgroup
        return;
    const int64_t sl = b.slot;
    const uint32_t di = (uint32_t)b.sj * (uint32_t)b.sk, dj = (uint32_t)b.sk;
    double *S = dst_ptr(b);
    const WT *G = dweights<WT>(b);
    JArrays<WT> A;
    A.Bx = S + (int64_t)ST_B * sl;
    A.By = A.Bx + sl;
    A.Bz = A.Bx + 2 * sl;
    A.bx = G + (int64_t)(G_BETA - G_NMETRIC) * sl;
    A.by = A.bx + sl;
    A.bz = A.bx + 2 * sl;
    A.ax = G + (int64_t)(G_ALPHA - G_NMETRIC) * sl;
    A.ay = A.ax + sl;
    A.az = A.ax + 2 * sl;
    const double *X = dmetric(b) + (int64_t)G_NODE * sl, *Y = X + sl, *Z = X + 2 * sl;
    const WT *Mw = G + (int64_t)(G_WJM - G_NMETRIC) * sl;
    const WT *Aw = G + (int64_t)(G_WJA - G_NMETRIC) * sl;
    double *Jc = S + (int64_t)ST_JCELL * sl;
    // solid wall: J_cell = 0 on the two cell layers next to it
    // (Z4_Lunar/src/boundary/Boundary_Condition_EM.cpp:382-427)
    bool wall_jk = false;
    if (b.bc[2] == BC_SOLID && j < 2)
        wall_jk = true;
    if (b.bc[3] == BC_SOLID && j >= b.nj - 2)
        wall_jk = true;
    if (b.bc[4] == BC_SOLID && k < 2)
        wall_jk = true;
    if (b.bc[5] == BC_SOLID && k >= b.nk - 2)
        wall_jk = true;
    uint32_t o = (uint32_t)blk_off(b, i0, j, k);
    // plane i0: its J edges from the products of plane i0 - 1, its nodes
    JEdges cur;
    {
        const uint32_t m = o - di;
        const double ppy = ld32(A.by, m) * ld32(A.By, m), ppz = ld32(A.bz, m) * ld32(A.Bz, m);
        const double ppyj = ld32(A.by, m + dj) * ld32(A.By, m + dj);
        cur = jedges_of(A, o, dj, ppy, ppz, ppyj);
    }
    double nd[8][3];
    nd[0][0] = ld32(X, o);
    nd[0][1] = ld32(Y, o);
    nd[0][2] = ld32(Z, o);
    nd[1][0] = ld32(X, o + dj);
    nd[1][1] = ld32(Y, o + dj);
    nd[1][2] = ld32(Z, o + dj);
#pragma unroll 1
    for (int i = i0; i < i1; ++i, o += di)
    {
        const JEdges nx = jedges_of(A, o + di, dj, cur.py, cur.pz, cur.pyj);
        nd[4][0] = ld32(X, o + di);
        nd[4][1] = ld32(Y, o + di);
        nd[4][2] = ld32(Z, o + di);
        nd[5][0] = ld32(X, o + di + dj);
        nd[5][1] = ld32(Y, o + di + dj);
        nd[5][2] = ld32(Z, o + di + dj);
#pragma unroll
        for (int c = 0; c < 3; ++c)
        {
            nd[2][c] = lane_next(nd[0][c]);
            nd[3][c] = lane_next(nd[1][c]);
            nd[6][c] = lane_next(nd[4][c]);
            nd[7][c] = lane_next(nd[5][c]);
        }
        // Mi is symmetric: rows (00 01 02) (01 11 12) (02 12 22)
        const double m00 = ld32(Mw, o), m01 = ld32(Mw + sl, o), m02 = ld32(Mw + 2 * sl, o),
                     m11 = ld32(Mw + 3 * sl, o), m12 = ld32(Mw + 4 * sl, o), m22 = ld32(Mw + 5 * sl, o);
        const double mi[9] = {m00, m01, m02, m01, m11, m12, m02, m12, m22};
        const unsigned act = (unsigned)ld32(Aw, o);
        const double s[12] = {cur.jx, cur.jxj, lane_next(cur.jx), lane_next(cur.jxj),
                              cur.jy, nx.jy,   lane_next(cur.jy), lane_next(nx.jy),
                              cur.jz, nx.jz,   cur.jzj,           nx.jzj};
        double jc[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int n = 0; n < 12; ++n)
        {
            const double rx = nd[kEe[n]][0] - nd[kEs[n]][0], ry = nd[kEe[n]][1] - nd[kEs[n]][1],
                         rz = nd[kEe[n]][2] - nd[kEs[n]][2];
            const bool on = (act >> n) & 1u;
            const double w0 = on ? mi[0] * rx + mi[1] * ry + mi[2] * rz : 0.0;
            const double w1 = on ? mi[3] * rx + mi[4] * ry + mi[5] * rz : 0.0;
            const double w2 = on ? mi[6] * rx + mi[7] * ry + mi[8] * rz : 0.0;
            jc[0] += w0 * s[n];
            jc[1] += w1 * s[n];
            jc[2] += w2 * s[n];
        }
        const bool wall = wall_jk || (b.bc[0] == BC_SOLID && i < 2) || (b.bc[1] == BC_SOLID && i >= b.ni - 2);
        if (wall)
            jc[0] = jc[1] = jc[2] = 0.0;
        if (own)
        {
#pragma unroll
            for (int a = 0; a < 3; ++a)
                st32(Jc + a * sl, o, jc[a]);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c)
        {
            nd[0][c] = nd[4][c];
            nd[1][c] = nd[5][c];
        }
        cur = nx;
    }
}

// ======================================================================
// Two-fluid (Mercury) step: H+ and Na+ share the Rusanov fluxes, the CT
// induction and the derived magnetic fields of the single-fluid step; these
// kernels add the second species, the composition-weighted ion velocity,
// the Mercury sources and the ambipolar EMF.

// Hall_Num_Limiter (Z5_Mercury/src/physics/MercurySolver_Tools.cpp:90-136)
struct NumInfo
{
    double nH, nNa, ne, ne_eff, chiH, chiNa, wH, wNa;
};
__device__ __forceinline__ NumInfo num_limiter(const MhdCaseConst &c, double rhoH, double rhoNa)
{
    NumInfo n;
    n.nH = dmax(rhoH, 0.0) / c.M_H * c.M_ref;
    n.nNa = dmax(rhoNa, 0.0) / c.M_Na * c.M_ref;
    n.ne = n.nH + n.nNa;
    n.ne_eff = sqrt(n.ne * n.ne + c.ne_floor * c.ne_floor);
    if (n.ne > 1.0e-300)
    {
        n.chiH = n.nH / n.ne;
        n.chiNa = n.nNa / n.ne;
    }
    else
    {
        n.chiH = 1.0;
        n.chiNa = 0.0;
    }
    n.wH = n.nH / n.ne_eff;
    n.wNa = n.nNa / n.ne_eff;
    return n;
}

// Derived fields of the two-fluid step, cells [-1, N]: contravariant
// velocity of U_plus = chi_H u_H + chi_Na u_Na (calc_Uplus,
// Z5_Mercury/src/physics/MercurySolver_Tools.cpp:194-277), the electron
// pressure p_H + p_Na, and on interior cells B_cell and the next dt over both
// species (Z5_Mercury/src/physics/MercurySolver_Timestep.cpp:8-140).
template <typename WT>
__global__ __launch_bounds__(256) void k_derived2(const MhdTileMap tm, const MhdCaseConst *__restrict__ cst,
                                                  unsigned long long *__restrict__ dtbits, int ucur)
{
    MhdBlockDev b;
    int i, j, k;
    bool wg;
    const bool ok = tile_cell(tm, b, i, j, k, wg);
    double local = 1e100;
    if (ok)
    {
        const MhdCaseConst &c = cst[b.case_id];
        const int64_t sl = b.slot, o = blk_off(b, i, j, k);
        const int64_t di = (int64_t)b.sj * b.sk, dj = b.sk, dk = 1;
        double *S = dst_ptr(b);
        CellGeo q;
        load_cell(b, S + (int64_t)(ST_U0 + 5 * ucur) * sl, dmetric(b), o, q);
        const double *UN = S + (int64_t)(ST_UNA0 + 5 * ucur) * sl;
        double un[5];
#pragma unroll
        for (int m = 0; m < 5; ++m)
            un[m] = UN[m * sl + o];
        const Prim ph = prim_from(q.u[0], q.u[1], q.u[2], q.u[3], q.u[4], c.gamma);
        const Prim pn = prim_from(un[0], un[1], un[2], un[3], un[4], c.gamma);
        {
            const double rH = dmax(q.u[0], 0.0), rN = dmax(un[0], 0.0);
            const NumInfo n = num_limiter(c, rH, rN);
            const bool hH = rH > 1e-20, hN = rN > 1e-20;
            const double up0 = n.chiH * (hH ? ph.u : 0.0) + n.chiNa * (hN ? pn.u : 0.0);
            const double up1 = n.chiH * (hH ? ph.v : 0.0) + n.chiNa * (hN ? pn.v : 0.0);
            const double up2 = n.chiH * (hH ? ph.w : 0.0) + n.chiNa * (hN ? pn.w : 0.0);
            double uc[3];
            cell_ucontra_of(q, up0, up1, up2, uc);
            double *UC = S + (int64_t)ST_UC * sl;
#pragma unroll
            for (int a = 0; a < 3; ++a)
                UC[a * sl + o] = uc[a];
            S[(int64_t)ST_PE * sl + o] = ph.p + pn.p;
        }
        // B_cell on the interior cells; its loads (face fluxes o .. o + 1, the
        // cell's weights) are in bounds on the ring too, so they issue with
        // the velocity's and only the stores and dt are limited to the interior
        const bool inner = i >= 0 && j >= 0 && k >= 0 && i < b.ni && j < b.nj && k < b.nk;
        {
            const double *Bx = S + (int64_t)ST_B * sl, *By = Bx + sl, *Bz = Bx + 2 * sl;
            const double *Ax = S + (int64_t)ST_BADD * sl, *Ay = Ax + sl, *Az = Ax + 2 * sl;
            const double pt[6] = {-Bx[o] - Ax[o],          Bx[o + di] + Ax[o + di], -By[o] - Ay[o],
                                  By[o + dj] + Ay[o + dj], -Bz[o] - Az[o],          Bz[o + dk] + Az[o + dk]};
            double bc[3];
            bcell_from_faces<WT>(b, o, q, pt, bc);
            if (inner)
            {
                double *Bc = S + (int64_t)ST_BCELL * sl;
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    Bc[a * sl + o] = bc[a];
                FaceNormals fn;
                face_normals(q, fn);
                local = cell_dt_n(c, q, fn, bc[0], bc[1], bc[2]);
#pragma unroll
                for (int m = 0; m < 5; ++m)
                    q.u[m] = un[m];
                local = dmin(local, cell_dt_n(c, q, fn, bc[0], bc[1], bc[2]));
            }
        }
    }
    if (wg)
        reduce_dt(local, dtbits, b.case_id);
}

// The sources and the Euler update of the two-fluid cell o, given the flux
// divergence of a species in rhs (((0 - t_xi) - t_eta) - t_zeta) and both
// primitive states ph, pn (calc_PV of U[cur]): the Mercury terms in the order
// of MercurySolver_FluidRHS.cpp:206-527, then U[1-cur] = U[cur] + dt rhs
// (MercurySolver_Advance.cpp:73-112).  two_fluid_cell forms what both
// species read; two_fluid_species applies one species' terms.  Shared by the
// cell and column forms of the update.
struct TwoFluidCell
{
    NumInfo num;
    double upx, upy, upz, Bx, By, Bz, sjbx, sjby, sjbz, dpex, dpey, dpez, sss, vst;
};
__device__ __forceinline__ TwoFluidCell two_fluid_cell(const MhdCaseConst &c, const double *__restrict__ S,
                                                       const double *__restrict__ G, int64_t sl, int64_t o,
                                                       const int64_t *step, int ucur, const Prim &ph, const Prim &pn)
{
    TwoFluidCell t;
    const double *UH = S + (int64_t)(ST_U0 + 5 * ucur) * sl, *UN = S + (int64_t)(ST_UNA0 + 5 * ucur) * sl;
    t.num = num_limiter(c, UH[o], UN[o]);
    {
        // U_plus of this cell (calc_Uplus), from the same U the step started
        // from; its limiter of the clamped densities is t.num (num_limiter
        // clamps at 0 itself, and dmax(dmax(x, 0), 0) == dmax(x, 0) for every x)
        const double rH = dmax(UH[o], 0.0), rN = dmax(UN[o], 0.0);
        const NumInfo &n = t.num;
        const bool hH = rH > 1e-20, hN = rN > 1e-20;
        t.upx = n.chiH * (hH ? ph.u : 0.0) + n.chiNa * (hN ? pn.u : 0.0);
        t.upy = n.chiH * (hH ? ph.v : 0.0) + n.chiNa * (hN ? pn.v : 0.0);
        t.upz = n.chiH * (hH ? ph.w : 0.0) + n.chiNa * (hN ? pn.w : 0.0);
    }
    const double *Bc = S + (int64_t)ST_BCELL * sl, *Jc = S + (int64_t)ST_JCELL * sl;
    t.Bx = Bc[o];
    t.By = Bc[sl + o];
    t.Bz = Bc[2 * sl + o];
    const double Jx = Jc[o], Jy = Jc[sl + o], Jz = Jc[2 * sl + o];
    t.sjbx = Jy * t.Bz - Jz * t.By;
    t.sjby = Jz * t.Bx - Jx * t.Bz;
    t.sjbz = Jx * t.By - Jy * t.Bx;
    t.dpex = t.dpey = t.dpez = 0.0;
    if (c.ambipolar != 0.0)
    {
        double m9[9];
        const double V = fabs(G[(int64_t)G_JAC * sl + o]);
#pragma unroll
        for (int ax = 0; ax < 3; ++ax)
        {
            const double *A = G + (int64_t)(G_JD + 3 * ax) * sl;
#pragma unroll
            for (int q = 0; q < 3; ++q)
                m9[3 * ax + q] = V <= 0.0 ? 0.0 : 0.5 * (A[q * sl + o - step[ax]] + A[q * sl + o]) / V;
        }
        const double *PE = S + (int64_t)ST_PE * sl;
        const double gi = PE[o + step[0]] - PE[o - step[0]];
        const double gj = PE[o + step[1]] - PE[o - step[1]];
        const double gk = PE[o + step[2]] - PE[o - step[2]];
        t.dpex = 0.5 * (m9[0] * gi + m9[3] * gj + m9[6] * gk);
        t.dpey = 0.5 * (m9[1] * gi + m9[4] * gj + m9[7] * gk);
        t.dpez = 0.5 * (m9[2] * gi + m9[5] * gj + m9[8] * gk);
    }
    t.sss = S[(int64_t)ST_PHOTO * sl + o];
    t.vst = S[(int64_t)ST_DRAG * sl + o];
    return t;
}

__device__ __forceinline__ void two_fluid_species(const MhdCaseConst &c, double *__restrict__ S, int64_t sl,
                                                  int64_t o, int ucur, int sp, const Prim &p, const TwoFluidCell &t,
                                                  double *R_, double dt)
{
    const double *U = S + (int64_t)((sp ? ST_UNA0 : ST_U0) + 5 * ucur) * sl;
    const double n_s = sp ? t.num.nNa : t.num.nH;
    const double chi = sp ? t.num.wNa : t.num.wH;
    const double rho = dmax(U[o], 0.0);
    const double u = p.u, v = p.v, w = p.w;
    const double u2 = u * u + v * v + w * w;
    const double subx = (v - t.upy) * t.Bz - (w - t.upz) * t.By;
    const double suby = (w - t.upz) * t.Bx - (u - t.upx) * t.Bz;
    const double subz = (u - t.upx) * t.By - (v - t.upy) * t.Bx;
    const double subu = subx * u + suby * v + subz * w;
    const double sjbu = t.sjbx * u + t.sjby * v + t.sjbz * w;
    const double dpeu = t.dpex * u + t.dpey * v + t.dpez * w;
    const double T = p.p / (dmax(p.rho_raw, 1e-12) * (sp ? c.state_coeff_Na : c.state_coeff));
    const double Ts0 = T * c.T_ref;
    const double vst = t.vst, sss = t.sss;
    const double b1 = sp ? vst * (c.Tn0 - Ts0) * (c.M_Na / (c.M_Na + c.M_Na))
                         : vst * (c.Tn0 - Ts0) * (c.M_H / (c.M_H + c.M_Na));
    const double b2 = sp ? (c.M_Na / (c.M_Na + c.M_Na)) * vst : (c.M_Na / (c.M_H + c.M_Na)) * vst;
    R_[1] += c.induce * n_s * subx + c.momentum_hall_coeff * chi * t.sjbx;
    R_[2] += c.induce * n_s * suby + c.momentum_hall_coeff * chi * t.sjby;
    R_[3] += c.induce * n_s * subz + c.momentum_hall_coeff * chi * t.sjbz;
    R_[4] += c.induce * n_s * subu + c.momentum_hall_coeff * chi * sjbu;
    if (c.ambipolar != 0.0)
    {
        R_[1] -= chi * t.dpex;
        R_[2] -= chi * t.dpey;
        R_[3] -= chi * t.dpez;
        R_[4] -= chi * dpeu;
    }
    if (sp)
    {
        R_[0] += c.a1_Na * sss;
        R_[4] += c.a6 * sss * c.Tn0;
    }
    R_[1] -= c.a4 * rho * u * vst;
    R_[2] -= c.a4 * rho * v * vst;
    R_[3] -= c.a4 * rho * w * vst;
    R_[4] -= c.a4 * rho * u2 * vst;
    R_[4] += c.a5 * n_s * b1;
    R_[4] += c.a4 * rho * u2 * b2;
    double *Un = S + (int64_t)((sp ? ST_UNA0 : ST_U0) + 5 * (1 - ucur)) * sl;
#pragma unroll
    for (int m = 0; m < 5; ++m)
        Un[m * sl + o] = U[m * sl + o] + dt * R_[m];
}

__device__ __forceinline__ void two_fluid_update(const MhdCaseConst &c, double *__restrict__ S,
                                                 const double *__restrict__ G, int64_t sl, int64_t o,
                                                 const int64_t *step, int ucur, const Prim &ph, const Prim &pn,
                                                 double (&rhs)[2][5], double dt)
{
    const TwoFluidCell t = two_fluid_cell(c, S, G, sl, o, step, ucur, ph, pn);
#pragma unroll
    for (int sp = 0; sp < 2; ++sp)
        two_fluid_species(c, S, sl, o, ucur, sp, sp ? pn : ph, t, rhs[sp], dt);
}

// Fluid update of the two-fluid step: Rusanov fluxes of each species
// (Z5_Mercury/src/physics/MercurySolver_FluidRHS.cpp:3-202), the Mercury
// sources (:206-527: induction force n_s (u_s - u_+) x B, composition-
// weighted J x B, electron-pressure force, Na+ photoionization mass and
// energy, drag on the neutrals and thermal exchange) and the Euler update of
// both species (MercurySolver_Advance.cpp:73-112).  The pressure-gradient
// metric averages faces i-1 and i (j-1 and j, k-1 and k) as the reference
// does.
// (At 3 waves per SIMD -- 168 VGPRs, a few spilled -- it runs 0.52 ms on the
// 6 x 64^2 x 128 Mercury shell, 0.55 ms with the compiler's 182 VGPRs at 2.)
template <typename R, int RECON>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_fluid2(
    const MhdTileMap tm, const MhdCaseConst *__restrict__ cst, const double *__restrict__ dts, int ucur)
{
    MhdBlockDev b;
    int i, j, k;
    bool wg;
    if (!tile_cell(tm, b, i, j, k, wg))
        return;
    const MhdCaseConst &c = cst[b.case_id];
    const double gm = c.gamma;
    const int64_t sl = b.slot, o = blk_off(b, i, j, k);
    const int64_t step[3] = {(int64_t)b.sj * b.sk, (int64_t)b.sk, 1};
    double *S = dst_ptr(b);
    const double *G = dmetric(b);
    const double invJ = 1.0 / G[(int64_t)G_JAC * sl + o];
    double rhs[2][5];
    // Kept rolled: unrolling the two species doubles the live flux state
    // (256 VGPRs, occupancy 2) and measured 22% slower on the 6x64x64x128 shell.
#pragma unroll 1
    for (int sp = 0; sp < 2; ++sp)
    {
        const double *U = S + (int64_t)((sp ? ST_UNA0 : ST_U0) + 5 * ucur) * sl;
#pragma unroll
        for (int m = 0; m < 5; ++m)
            rhs[sp][m] = 0.0;
#pragma unroll
        for (int ax = 0; ax < 3; ++ax)
        {
            const double *A = G + (int64_t)(G_JD + 3 * ax) * sl;
            double Flo[5], Fhi[5];
            face_flux<R, RECON>(U, sl, A, o - step[ax], o, step[ax], gm, Flo);
            face_flux<R, RECON>(U, sl, A, o, o + step[ax], step[ax], gm, Fhi);
#pragma unroll
            for (int m = 0; m < 5; ++m)
                rhs[sp][m] -= (Fhi[m] - Flo[m]) * invJ;
        }
    }
    const double *UH = S + (int64_t)(ST_U0 + 5 * ucur) * sl, *UN = S + (int64_t)(ST_UNA0 + 5 * ucur) * sl;
    const Prim ph = prim_of<sizeof(R) == 8>(UH, sl, o, gm), pn = prim_of<sizeof(R) == 8>(UN, sl, o, gm);
    two_fluid_update(c, S, G, sl, o, step, ucur, ph, pn, rhs, dts[b.case_id]);
}

// The first-order two-fluid update in column tiles (k_fluid_col's form for
// both species): per plane i the workgroup evaluates flux_cell of H+ and Na+
// once for its 64 x COL_TJ cells and their one-cell halo in j and k (LDS); a
// thread keeps its column's cell i-1 of both species in registers, so each
// face flux is computed once (k_fluid2 evaluates every face twice, once from
// each side).  Each flux is rusanov_pre of the same two cells as face_flux
// in k_fluid2, the divergence adds the terms in its order and
// two_fluid_update applies the same sources, so the result is bit for bit
// k_fluid2's.
template <typename R>
__global__ __launch_bounds__(64 * COL_TJ) void k_fluid2_col(const MhdTileMap tm, const MhdCaseConst *__restrict__ cst,
                                                            const double *__restrict__ dts, int ucur)
{
    constexpr int PJ = COL_TJ + 2, PK = COL_TK + 2;
    __shared__ double P[2][7][PJ][PK];
    MhdBlockDev b;
    int i0, j0, k0;
    if (!tile_col(tm, b, i0, j0, k0)) // uniform over the workgroup
        return;
    const int tx = (int)threadIdx.x, ty = (int)threadIdx.y, tid = ty * 64 + tx;
    const int j = j0 + ty, k = k0 + tx;
    const bool act = j < b.nj && k < b.nk;
    const int i1 = min(i0 + tm.ci, b.ni);
    const MhdCaseConst &c = cst[b.case_id];
    const double gm = c.gamma;
    const int64_t sl = b.slot, di = (int64_t)b.sj * b.sk, dj = b.sk;
    const int64_t step[3] = {di, dj, 1};
    double *S = dst_ptr(b);
    const double *Us[2] = {S + (int64_t)(ST_U0 + 5 * ucur) * sl, S + (int64_t)(ST_UNA0 + 5 * ucur) * sl};
    const double *G = dmetric(b);
    const double *Ax = G + (int64_t)G_JD * sl, *Ay = Ax + 3 * sl, *Az = Ax + 6 * sl;
    const double dt = dts[b.case_id];
    int64_t o = blk_off(b, i0, act ? j : j0, act ? k : k0); // cell of plane `pl` in the loop
    FluxCell prev[2] = {};
    if (act)
    {
#pragma unroll
        for (int sp = 0; sp < 2; ++sp)
            prev[sp] = flux_cell<R>(Us[sp], sl, o - di, gm);
    }
    // the open cell (plane pl - 1): lower i-face flux, j and k terms, 1/J
    double Flo[2][5] = {}, tj[2][5] = {}, tk[2][5] = {};
    double invJ = 0.0;
    for (int pl = i0; pl <= i1; ++pl, o += di)
    {
        if (pl < i1)
        {
            for (int p = tid; p < PJ * PK; p += 64 * COL_TJ)
            {
                const int pj = p / PK, pk = p - pj * PK;
                const int jj = j0 - 1 + pj, kk = k0 - 1 + pk;
                if (jj <= b.nj && kk <= b.nk)
                {
#pragma unroll
                    for (int sp = 0; sp < 2; ++sp)
                    {
                        const FluxCell f = flux_cell<R>(Us[sp], sl, blk_off(b, pl, jj, kk), gm);
                        P[sp][0][pj][pk] = f.rho;
                        P[sp][1][pj][pk] = f.u;
                        P[sp][2][pj][pk] = f.v;
                        P[sp][3][pj][pk] = f.w;
                        P[sp][4][pj][pk] = f.p;
                        P[sp][5][pj][pk] = f.e;
                        P[sp][6][pj][pk] = f.c;
                    }
                }
            }
            __syncthreads();
        }
        if (act)
        {
            auto lds = [&](int sp, int pj, int pk) {
                FluxCell f;
                f.rho = P[sp][0][pj][pk];
                f.u = P[sp][1][pj][pk];
                f.v = P[sp][2][pj][pk];
                f.w = P[sp][3][pj][pk];
                f.p = P[sp][4][pj][pk];
                f.e = P[sp][5][pj][pk];
                f.c = P[sp][6][pj][pk];
                return f;
            };
            FluxCell cur[2];
            double F[2][5];
#pragma unroll
            for (int sp = 0; sp < 2; ++sp)
            {
                cur[sp] = pl < i1 ? lds(sp, ty + 1, tx + 1) : flux_cell<R>(Us[sp], sl, o, gm);
                rusanov_pre<R>(prev[sp], cur[sp], Ax[o], Ax[sl + o], Ax[2 * sl + o], F[sp]);
            }
            if (pl > i0) // close cell pl - 1
            {
                double rhs[2][5];
#pragma unroll
                for (int sp = 0; sp < 2; ++sp)
#pragma unroll
                    for (int m = 0; m < 5; ++m)
                    {
                        double r = 0.0;
                        r -= (F[sp][m] - Flo[sp][m]) * invJ;
                        r -= tj[sp][m];
                        r -= tk[sp][m];
                        rhs[sp][m] = r;
                    }
                const Prim ph{prev[0].rho, prev[0].u, prev[0].v, prev[0].w, prev[0].p};
                const Prim pn{prev[1].rho, prev[1].u, prev[1].v, prev[1].w, prev[1].p};
                two_fluid_update(c, S, G, sl, o - di, step, ucur, ph, pn, rhs, dt);
            }
            if (pl < i1) // open cell pl
            {
                invJ = 1.0 / G[(int64_t)G_JAC * sl + o];
#pragma unroll
                for (int sp = 0; sp < 2; ++sp)
                {
#pragma unroll
                    for (int m = 0; m < 5; ++m)
                        Flo[sp][m] = F[sp][m];
                    double Fl[5], Fh[5];
                    rusanov_pre<R>(lds(sp, ty, tx + 1), cur[sp], Ay[o], Ay[sl + o], Ay[2 * sl + o], Fl);
                    rusanov_pre<R>(cur[sp], lds(sp, ty + 2, tx + 1), Ay[o + dj], Ay[sl + o + dj], Ay[2 * sl + o + dj],
                                   Fh);
#pragma unroll
                    for (int m = 0; m < 5; ++m)
                        tj[sp][m] = (Fh[m] - Fl[m]) * invJ;
                    rusanov_pre<R>(lds(sp, ty + 1, tx), cur[sp], Az[o], Az[sl + o], Az[2 * sl + o], Fl);
                    rusanov_pre<R>(cur[sp], lds(sp, ty + 1, tx + 2), Az[o + 1], Az[sl + o + 1], Az[2 * sl + o + 1], Fh);
#pragma unroll
                    for (int m = 0; m < 5; ++m)
                        tk[sp][m] = (Fh[m] - Fl[m]) * invJ;
                }
            }
#pragma unroll
            for (int sp = 0; sp < 2; ++sp)
                prev[sp] = cur[sp];
        }
        if (pl < i1)
            __syncthreads();
    }
}

// Two-fluid edge EMF terms added to the ideal EMF before the Eedge sync, in
// the reference's order (Z5_Mercury/src/physics/MercurySolver_MagneticRHS.cpp:30-56):
//  * AddAmbipolarEdgeEMF_ (MercurySolver_MagExplicit_Ambipo.cpp:3-144):
//    E -= pe_coef (pe(node 1) - pe(node 0)) / ne_eff, pe averaged over the 8
//    cells around a node, ne_eff of the mean density of the 4 cells around
//    the edge;
//  * AddLocalArtificialResistivityToEdgeEMF_ (MercurySolver_MagExplicit_Resis.cpp:67-186):
//    E += eta(edge) J_edge, J_edge = alpha CurlAdj(beta B) as in J_cell;
//  * AddHallEdgeEMF_ (MercurySolver_MagExplicit_Hall.cpp:39-440): E +=
//    taper (<alpha J x B>_4 cells . dr + whistler LLF diffusion of the
//    transverse face-flux jumps), alpha = hall_coef / ne_eff.  The Ehall sync
//    of the reference only fills ghosts, which E never reads.
// The kernel runs over padded-row plane chunks of the points [0, N]^3
// (tile_padded) and marches along i.  The edges of the three families at o
// meet the cells o + s0 di + s1 dj + s2 dk, s in {-1, 0}^3 with s_a = 0 for
// family a: 7 distinct cells (code (s0 + 1) 4 + (s1 + 1) 2 + (s2 + 1)).  A
// point forms the densities and Hall terms of the cells o (code 7) and o - dj
// (5); those at o - dk (6, 4) are what the lane before it formed, and those of
// plane i - 1 (3, 1, and 2 from the lane before) it carries from its previous
// plane.  Node pressures (the mean of the 8 cells around a node), beta B
// products, node coordinates and face-area norms are shared the same way:
// each value is the one the per-point form computes, by the same operations,
// so E is bit for bit as before.  Every load stays in [-1, N + 1] (the padded
// lanes sit at k = -1 and N + 1, the warm-up plane at i0 - 1 >= -1).
struct XCell
{
    double rh, rn, e0, e1, e2, bt;
};

__device__ __forceinline__ XCell xcell_form(const MhdCaseConst &c, bool hall, const double *RH, const double *RN,
                                            const double *Jc, const double *Bc, int64_t sl, uint32_t p)
{
    XCell r;
    r.rh = ld32(RH, p);
    r.rn = ld32(RN, p);
    r.e0 = r.e1 = r.e2 = r.bt = 0.0;
    if (hall)
    {
        const NumInfo nm = num_limiter(c, r.rh, r.rn);
        const double alpha = c.hall_coef / nm.ne_eff;
        const double Jx = ld32(Jc, p), Jy = ld32(Jc + sl, p), Jz = ld32(Jc + 2 * sl, p);
        const double Bx = ld32(Bc, p), By = ld32(Bc + sl, p), Bz = ld32(Bc + 2 * sl, p);
        r.e0 = alpha * (Jy * Bz - Jz * By);
        r.e1 = alpha * (Jz * Bx - Jx * Bz);
        r.e2 = alpha * (Jx * By - Jy * Bx);
        r.bt = fabs(alpha) * sqrt(Bx * Bx + By * By + Bz * Bz);
    }
    return r;
}

__device__ __forceinline__ XCell xcell_prev(const XCell &a, bool hall)
{
    XCell r;
    r.rh = lane_prev(a.rh);
    r.rn = lane_prev(a.rn);
    r.e0 = r.e1 = r.e2 = r.bt = 0.0;
    if (hall)
    {
        r.e0 = lane_prev(a.e0);
        r.e1 = lane_prev(a.e1);
        r.e2 = lane_prev(a.e2);
        r.bt = lane_prev(a.bt);
    }
    return r;
}

// mean of the 8 cells around a node, summed in the order (i, j, k) offsets
// (-1,-1,-1), (-1,-1,0), (-1,0,-1), ... (0,0,0)
__device__ __forceinline__ double node_mean8(double v0, double v1, double v2, double v3, double v4, double v5,
                                             double v6, double v7)
{
    double sum = 0.0;
    sum += v0;
    sum += v1;
    sum += v2;
    sum += v3;
    sum += v4;
    sum += v5;
    sum += v6;
    sum += v7;
    return 0.125 * sum;
}

__device__ __forceinline__ double area_norm(const double *A, int64_t sl, uint32_t p)
{
    const double a0 = ld32(A, p), a1 = ld32(A + sl, p), a2 = ld32(A + 2 * sl, p);
    return sqrt(a0 * a0 + a1 * a1 + a2 * a2);
}

// ambipolar term of one edge: the 4 cells q0..q3 in the order (s_t1, s_t2) =
// (-1,-1), (-1,0), (0,-1), (0,0), node pressures at the edge's ends
__device__ __forceinline__ double ambi_term(const MhdCaseConst &c, const XCell &q0, const XCell &q1, const XCell &q2,
                                            const XCell &q3, double pe0, double pe1)
{
    double rhoH = 0.0, rhoNa = 0.0;
    rhoH += q0.rh;
    rhoNa += q0.rn;
    rhoH += q1.rh;
    rhoNa += q1.rn;
    rhoH += q2.rh;
    rhoNa += q2.rn;
    rhoH += q3.rh;
    rhoNa += q3.rn;
    const NumInfo nm = num_limiter(c, 0.25 * rhoH, 0.25 * rhoNa);
    return c.pe_coef * (pe1 - pe0) / nm.ne_eff;
}

// Hall term of one edge before the taper: <alpha J x B> . dr over the 4 cells
// plus the whistler diffusion of the transverse face-flux jumps.  mu1 goes
// with d21 and mu2 with d12 as (Ecen + 0.5 mu1 d21) - 0.5 mu2 d12; the
// whistler speeds come from the face-area norms (A0, A1) of the two face
// families along the edge.
__device__ __forceinline__ void hall_avg(const XCell &q0, const XCell &q1, const XCell &q2, const XCell &q3,
                                         double d0, double d1, double d2, double &Ecen, double &bmax, double &L)
{
    const double Ec0 = 0.25 * (q0.e0 + q1.e0 + q2.e0 + q3.e0);
    const double Ec1 = 0.25 * (q0.e1 + q1.e1 + q2.e1 + q3.e1);
    const double Ec2 = 0.25 * (q0.e2 + q1.e2 + q2.e2 + q3.e2);
    bmax = dmax(dmax(q0.bt, q1.bt), dmax(q2.bt, q3.bt));
    Ecen = Ec0 * d0 + Ec1 * d1 + Ec2 * d2;
    L = dmax(sqrt(d0 * d0 + d1 * d1 + d2 * d2), 1e-14);
}
__device__ __forceinline__ double whistler_mu(double bmax, double L, double A0, double A1)
{
    const double eps = 1e-14, Cwh = 0.1;
    const double Abar = 0.5 * (A0 + A1);
    const double inv_h = L / dmax(Abar, eps);
    return Cwh * bmax * inv_h * inv_h;
}

template <typename WT>
__global__ __launch_bounds__(256) void k_emf_extra(const MhdTileMap tm, const MhdCaseConst *__restrict__ cst,
                                                   int ucur)
{
    MhdBlockDev b;
    int i0, i1, j, k, sub;
    bool own;
    if (!tile_padded(tm, b, i0, i1, j, k, own, 1, sub)) // uniform over the workgroup
        return;
    const MhdCaseConst &c = cst[b.case_id];
    const bool amb = c.ambipolar != 0.0, loc = c.local_arti != 0.0, hall = c.hall != 0.0;
    if (!amb && !loc && !hall) // one case per block: uniform
        return;
    const bool cells = amb || hall, faces = loc || hall;
    // 32-bit element indices within a slot, slot bases uniform
    const int64_t sl = b.slot;
    const uint32_t di = (uint32_t)b.sj * (uint32_t)b.sk, dj = (uint32_t)b.sk;
    double *S = dst_ptr(b);
    const double *PE = S + (int64_t)ST_PE * sl;
    const double *RH = S + (int64_t)(ST_U0 + 5 * ucur) * sl, *RN = S + (int64_t)(ST_UNA0 + 5 * ucur) * sl;
    const double *Bx = S + (int64_t)ST_B * sl, *By = Bx + sl, *Bz = Bx + 2 * sl;
    const double *Bc = S + (int64_t)ST_BCELL * sl, *Jc = S + (int64_t)ST_JCELL * sl;
    const double *TP = S + (int64_t)ST_TAPER * sl, *EL = S + (int64_t)ST_ETALOC * sl;
    double *E = S + (int64_t)ST_E * sl;
    const double *G = dmetric(b);
    const double *X = G + (int64_t)G_NODE * sl, *JD = G + (int64_t)G_JD * sl;
    const WT *Wt = dweights<WT>(b);
    const WT *beta = Wt + (int64_t)(G_BETA - G_NMETRIC) * sl, *alph = Wt + (int64_t)(G_ALPHA - G_NMETRIC) * sl;

    // warm-up, plane i0 - 1: cells (i - 1, j, k) and (i - 1, j - 1, k); the
    // pressures of planes i - 1 and i around node (i, j, k); node (i, j, k);
    // what the families take from the faces at o - di
    uint32_t o = (uint32_t)blk_off(b, i0 - 1, j, k);
    XCell c3 = {}, c1 = {};
    if (cells)
    {
        c3 = xcell_form(c, hall, RH, RN, Jc, Bc, sl, o);
        c1 = xcell_form(c, hall, RH, RN, Jc, Bc, sl, o - dj);
    }
    double pe0 = 0.0, pm = 0.0, pj = 0.0, pjj = 0.0, qj = 0.0; // P[i][j-1], P[i][j], P[i-1][j+1], P[i-1][j]
    if (amb)
    {
        const double a0 = ld32(PE, o - dj), a1 = ld32(PE, o), b0 = ld32(PE, o + di - dj), b1 = ld32(PE, o + di);
        pe0 = node_mean8(lane_prev(a0), a0, lane_prev(a1), a1, lane_prev(b0), b0, lane_prev(b1), b1);
        pm = b0;
        pj = b1;
        pjj = ld32(PE, o + dj);
        qj = a1;
    }
    double xo0 = 0.0, xo1 = 0.0, xo2 = 0.0, pn1 = 0.0, pn2 = 0.0, pBy = 0.0, pBz = 0.0, ppy = 0.0, ppz = 0.0;
    if (hall)
    {
        xo0 = ld32(X, o + di);
        xo1 = ld32(X + sl, o + di);
        xo2 = ld32(X + 2 * sl, o + di);
        pn1 = area_norm(JD + 3 * sl, sl, o);
        pn2 = area_norm(JD + 6 * sl, sl, o);
    }
    if (faces)
    {
        pBy = ld32(By, o);
        pBz = ld32(Bz, o);
    }
    if (loc)
    {
        ppy = ld32(beta + sl, o) * pBy;
        ppz = ld32(beta + 2 * sl, o) * pBz;
    }
    o += di;
    const bool st0 = own && j <= b.nj, st1 = own && j < b.nj, st2 = own && j <= b.nj && k < b.nk;
#pragma unroll 1
    for (int i = i0; i < i1; ++i, o += di)
    {
        double e0 = ld32(E, o), e1 = ld32(E + sl, o), e2 = ld32(E + 2 * sl, o);
        XCell c7 = {}, c5 = {};
        if (cells)
        {
            c7 = xcell_form(c, hall, RH, RN, Jc, Bc, sl, o);
            c5 = xcell_form(c, hall, RH, RN, Jc, Bc, sl, o - dj);
        }
        const XCell c6 = xcell_prev(c7, hall), c4 = xcell_prev(c5, hall), c2 = xcell_prev(c3, hall);
        double pex = 0.0;
        if (amb)
        {
            // nodes (i + 1, j, k), (i, j + 1, k), (i, j, k + 1)
            const double a0 = ld32(PE, o + di - dj), a1 = ld32(PE, o + di), cj = ld32(PE, o + dj);
            pex = node_mean8(lane_prev(pm), pm, lane_prev(pj), pj, lane_prev(a0), a0, lane_prev(a1), a1);
            const double pey = node_mean8(lane_prev(qj), qj, lane_prev(pjj), pjj, lane_prev(pj), pj, lane_prev(cj), cj);
            const double pez = lane_next(pe0);
            e0 -= ambi_term(c, c4, c5, c6, c7, pe0, pex);
            e1 -= ambi_term(c, c2, c3, c6, c7, pe0, pey);
            e2 -= ambi_term(c, c1, c3, c5, c7, pe0, pez);
            qj = pj;
            pj = a1;
            pm = a0;
            pjj = cj;
        }
        double bxo = 0.0, byo = 0.0, bzo = 0.0, bxj = 0.0, bzj = 0.0;
        if (faces)
        {
            bxo = ld32(Bx, o);
            byo = ld32(By, o);
            bzo = ld32(Bz, o);
            bxj = ld32(Bx, o - dj);
            bzj = ld32(Bz, o - dj);
        }
        double py = 0.0, pz = 0.0;
        if (loc)
        {
            // J edges alpha CurlAdj(beta B) (CurlAdjFaceToEdge, Z4_Lunar/src/operators/CTOperators.cpp:59-110)
            const double px = ld32(beta, o) * bxo;
            py = ld32(beta + sl, o) * byo;
            pz = ld32(beta + 2 * sl, o) * bzo;
            const double pxj = ld32(beta, o - dj) * bxj, pzj = ld32(beta + 2 * sl, o - dj) * bzj;
            const double pxk = lane_prev(px), pyk = lane_prev(py);
            const double J0 = (1.0 * ((pyk - py) + (pz - pzj))) * ld32(alph, o);
            const double J1 = (1.0 * ((px - pxk) + (ppz - pz))) * ld32(alph + sl, o);
            const double J2 = (1.0 * ((pxj - px) + (py - ppy))) * ld32(alph + 2 * sl, o);
            e0 += ld32(EL, o) * J0;
            e1 += ld32(EL + sl, o) * J1;
            e2 += ld32(EL + 2 * sl, o) * J2;
        }
        double n1 = 0.0, n2 = 0.0, xn0 = 0.0, xn1 = 0.0, xn2 = 0.0;
        if (hall)
        {
            // edge vectors dr = node(o + st[a]) - node(o) (src/3_field/4_MPCNS_Geometry.cpp:1251-1320)
            xn0 = ld32(X, o + di);
            xn1 = ld32(X + sl, o + di);
            xn2 = ld32(X + 2 * sl, o + di);
            const double xj0 = ld32(X, o + dj), xj1 = ld32(X + sl, o + dj), xj2 = ld32(X + 2 * sl, o + dj);
            const double xk0 = lane_next(xo0), xk1 = lane_next(xo1), xk2 = lane_next(xo2);
            const double n0 = area_norm(JD, sl, o);
            n1 = area_norm(JD + 3 * sl, sl, o);
            n2 = area_norm(JD + 6 * sl, sl, o);
            const double n0j = area_norm(JD, sl, o - dj), n2j = area_norm(JD + 6 * sl, sl, o - dj);
            const double n0k = lane_prev(n0), n1k = lane_prev(n1);
            const double byk = lane_prev(byo), bxk = lane_prev(bxo);
            double Ecen, bmax, L;
            {
                hall_avg(c4, c5, c6, c7, xn0 - xo0, xn1 - xo1, xn2 - xo2, Ecen, bmax, L);
                const double mu2 = whistler_mu(bmax, L, n1, n1k), mu1 = whistler_mu(bmax, L, n2, n2j);
                const double h = Ecen + 0.5 * mu1 * (bzo - bzj) - 0.5 * mu2 * (byo - byk);
                e0 += ld32(TP, o) * h;
            }
            {
                hall_avg(c2, c3, c6, c7, xj0 - xo0, xj1 - xo1, xj2 - xo2, Ecen, bmax, L);
                const double mu2 = whistler_mu(bmax, L, n0, n0k), mu0 = whistler_mu(bmax, L, n2, pn2);
                const double h = Ecen + 0.5 * mu2 * (bxo - bxk) - 0.5 * mu0 * (bzo - pBz);
                e1 += ld32(TP + sl, o) * h;
            }
            {
                hall_avg(c1, c3, c5, c7, xk0 - xo0, xk1 - xo1, xk2 - xo2, Ecen, bmax, L);
                const double mu1 = whistler_mu(bmax, L, n0, n0j), mu0 = whistler_mu(bmax, L, n1, pn1);
                const double h = Ecen + 0.5 * mu0 * (byo - pBy) - 0.5 * mu1 * (bxo - bxj);
                e2 += ld32(TP + 2 * sl, o) * h;
            }
        }
        if (i < b.ni && st0)
            st32(E, o, e0);
        if (i <= b.ni && st1)
            st32(E + sl, o, e1);
        if (i <= b.ni && st2)
            st32(E + 2 * sl, o, e2);
        // carry plane i to plane i + 1
        c3 = c7;
        c1 = c5;
        pe0 = pex;
        xo0 = xn0;
        xo1 = xn1;
        xo2 = xn2;
        pn1 = n1;
        pn2 = n2;
        pBy = byo;
        pBz = bzo;
        ppy = py;
        ppz = pz;
    }
}

// Interior resistivity, one explicit subcycle's EMF: AddResistiveEdgeEMF_To_
// (Z5_Mercury/src/physics/MercurySolver_MagExplicit_Resis.cpp:188-330),
// E = inver_Rem yita0 (J_cell averaged over the 4 cells around the edge) . dr,
// 0 where yita0 vanishes; E holds Eres (the step's EMF has been consumed).
__global__ __launch_bounds__(256) void k_eres(const MhdTileMap tm, const MhdCaseConst *__restrict__ cst)
{
    MhdBlockDev b;
    int i, j, k;
    bool wg;
    if (!tile_cell(tm, b, i, j, k, wg))
        return;
    if (cst[b.case_id].resist == 0.0)
        return;
    const int64_t sl = b.slot, o = blk_off(b, i, j, k);
    const int64_t st[3] = {(int64_t)b.sj * b.sk, (int64_t)b.sk, 1};
    double *S = dst_ptr(b);
    const double *Jc = S + (int64_t)ST_JCELL * sl;
    double *E = S + (int64_t)ST_E * sl;
    const int idx[3] = {i, j, k}, n[3] = {b.ni, b.nj, b.nk};
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
        bool in = true;
#pragma unroll
        for (int q = 0; q < 3; ++q)
            in = in && (idx[q] < n[q] + (q == a ? 0 : 1));
        if (!in)
            continue;
        const double coef = S[(int64_t)(ST_ETARES + a) * sl + o];
        double e = 0.0;
        if (coef != 0.0)
        {
            const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
            // cells (i,j,k), minus t1, minus t2, minus both (get_edge_cells, :232-261)
            const int64_t q[4] = {o, o - st[t1], o - st[t2], o - st[t1] - st[t2]};
            double J[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int p = 0; p < 4; ++p)
#pragma unroll
                for (int m = 0; m < 3; ++m)
                    J[m] += Jc[m * sl + q[p]];
#pragma unroll
            for (int m = 0; m < 3; ++m)
                J[m] *= 0.25;
            const double *X = dmetric(b) + (int64_t)G_NODE * sl;
            const int64_t o2 = o + st[a];
            const double Jedge =
                J[0] * (X[o2] - X[o]) + J[1] * (X[sl + o2] - X[sl + o]) + J[2] * (X[2 * sl + o2] - X[2 * sl + o]);
            e = 0.0 + coef * Jedge;
        }
        E[a * sl + o] = e;
    }
}

// Per-case time step of resistive subcycle `isub`: dt / nsub while the case
// has subcycles left, else 0 (its B is left unchanged by the update).
__global__ void k_dt_res(double *dtres, const double *dt, const MhdCaseConst *cst, int ncase, int isub)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncase)
        return;
    const MhdCaseConst &k = cst[c];
    const bool on = k.model == 1.0 && k.resist != 0.0 && isub < (int)k.resist_nsub;
    dtres[c] = on ? dt[c] / (double)(int)k.resist_nsub : 0.0;
}

// ======================================================================
// Boundary / halo op lists (one phase per launch).

// Op dst/src hold absolute device addresses; components are `stride`
// elements apart; element e of a run acts on dst + e dstep (src + e sstep).
// `ushift` moves U ops to the second U buffer.
__device__ __forceinline__ void apply_op(const MhdOp &op, int e, const MhdCaseConst *__restrict__ cst, int ushift)
{
    const int64_t sh = ushift ? 5 * op.stride : 0;
    double *d = gptr(op.dst) + sh + (int64_t)e * op.dstep;
    const double *s = gptr(op.src) + sh + (int64_t)e * op.sstep;
    switch (op.kind)
    {
    case OP_COPY:
        for (int c = 0; c < op.ncomp; ++c)
            d[c * op.stride] = s[c * op.stride];
        break;
    case OP_NEG:
        for (int c = 0; c < op.ncomp; ++c)
            d[c * op.stride] = -s[c * op.stride];
        break;
    case OP_ZERO:
        for (int c = 0; c < op.ncomp; ++c)
            d[c * op.stride] = 0.0;
        break;
    case OP_CONST:
        d[0] = op.val;
        break;
    case OP_QINF:
        for (int c = 0; c < op.ncomp; ++c)
            d[c * op.stride] = cst[op.cas].qinf[c];
        break;
    case OP_QINF_NA:
        for (int c = 0; c < op.ncomp; ++c)
            d[c * op.stride] = cst[op.cas].qinfs[c];
        break;
    case OP_NA_FLOOR:
        // Na+ absorbing wall (Z5_Mercury/src/boundary/Boundary_Condition_Fluid.cpp:144-160)
        d[0] = 1.0e-12;
        d[op.stride] = 0.0;
        d[2 * op.stride] = 0.0;
        d[3 * op.stride] = 0.0;
        d[4 * op.stride] = 1.0e-12 / (cst[op.cas].gamma - 1.0);
        break;
    case OP_SOLID_FIRST:
    case OP_SOLID_GHOST:
    {
        const int64_t L = op.stride;
        const double sv[5] = {s[0], s[L], s[2 * L], s[3 * L], s[4 * L]};
        double dv[5];
        solid_state(cst[op.cas].gamma, sv, dv);
#pragma unroll
        for (int c = 0; c < 5; ++c)
            d[c * L] = dv[c];
        break;
    }
    default:
        break;
    }
}

// Records of a list on 2^lg lanes each (lg 0: a thread per record, lg 6: a
// wavefront per record); the lanes of a record take its elements in turn.
// Copies (and negated copies) of up to 5 components: a lane loads every
// component of two elements before it stores any.  Written through one
// pointer type the compiler must assume a store may feed a later load and
// would order each component's load after the previous store; the
// destinations of a list never overlap its sources, so loading first gives
// the same values without the serialised round trips.
constexpr int kCopyComp = 5;
__device__ __forceinline__ void copy_run(const MhdOp &op, int e0, int lanes, int n, int ushift)
{
    const int64_t sh = ushift ? 5 * op.stride : 0;
    const bool neg = op.kind == OP_NEG;
    for (int e = e0; e < n; e += 2 * lanes)
    {
        const bool two = e + lanes < n;
        const double *s0 = gptr(op.src) + sh + (int64_t)e * op.sstep, *s1 = s0 + (int64_t)lanes * op.sstep;
        double *d0 = gptr(op.dst) + sh + (int64_t)e * op.dstep, *d1 = d0 + (int64_t)lanes * op.dstep;
        double v0[kCopyComp], v1[kCopyComp];
#pragma unroll
        for (int c = 0; c < kCopyComp; ++c)
            if (c < op.ncomp)
            {
                v0[c] = s0[c * op.stride];
                v1[c] = two ? s1[c * op.stride] : 0.0;
            }
#pragma unroll
        for (int c = 0; c < kCopyComp; ++c)
            if (c < op.ncomp)
            {
                d0[c * op.stride] = neg ? -v0[c] : v0[c];
                if (two)
                    d1[c * op.stride] = neg ? -v1[c] : v1[c];
            }
    }
}

__device__ __forceinline__ void apply_rec(const MhdOp *__restrict__ ops, int t, int lg,
                                          const MhdCaseConst *__restrict__ cst, int ushift)
{
    const MhdOp &op = ops[t >> lg];
    const int lanes = 1 << lg, n = op.run > 1 ? op.run : 1;
    if ((op.kind == OP_COPY || op.kind == OP_NEG) && op.ncomp <= kCopyComp)
    {
        copy_run(op, t & (lanes - 1), lanes, n, ushift);
        return;
    }
    for (int e = t & (lanes - 1); e < n; e += lanes)
        apply_op(op, e, cst, ushift);
}

__global__ __launch_bounds__(256) void k_ops(const mhdk::OpLaunch L, const MhdCaseConst *__restrict__ cst)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
#pragma unroll
    for (int g = 0; g < mhdk::OPSEG_MAX; ++g)
    {
        if (g >= L.nseg)
            return;
        const int m = L.seg[g].n << L.seg[g].lg;
        if (t < m)
        {
            apply_rec(L.seg[g].ops, t, L.seg[g].lg, cst, L.seg[g].ushift);
            return;
        }
        t -= m;
    }
}

// div B: calc_divB (Z4_Lunar/src/physics/LunarSolver_Tools.cpp:685-775), max |.| per case
// Over plane-chunk tiles: a thread takes the max over its column of planes,
// so a workgroup's one atomicMax covers 8 planes (the per-cell form spent
// its time queueing 12 k atomics on one address per case).
__global__ __launch_bounds__(256) void k_divb(const MhdTileMap tm, unsigned long long *__restrict__ out)
{
    MhdBlockDev b;
    int i0, i1, j, k;
    bool wg;
    const bool ok = tile_planes(tm, b, i0, i1, j, k, wg);
    double v = 0.0;
    if (ok)
    {
        const int64_t sl = b.slot, di = (int64_t)b.sj * b.sk;
        const double *Bx = dst_ptr(b) + (int64_t)ST_B * sl, *By = Bx + sl, *Bz = Bx + 2 * sl;
        const double *Jac = dmetric(b) + (int64_t)G_JAC * sl;
        int64_t o = blk_off(b, i0, j, k);
        for (int i = i0; i < i1; ++i, o += di)
        {
            const double dv = (Bx[o + di] - Bx[o] + By[o + b.sk] - By[o] + Bz[o + 1] - Bz[o]) / Jac[o];
            double a = fabs(dv);
            if (!(a == a))
                a = 1e300;
            v = dmax(v, a);
        }
    }
#pragma unroll
    for (int s = 32; s > 0; s >>= 1)
        v = dmax(v, __shfl_xor(v, s, 64));
    __shared__ double red[4];
    if (threadIdx.x == 0)
        red[threadIdx.y] = v;
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0 && wg)
    {
        double m = red[0];
        for (int q = 1; q < (int)blockDim.y; ++q)
            m = dmax(m, red[q]);
        atomicMax(&out[b.case_id], (unsigned long long)__double_as_longlong(m));
    }
}

// Min/max diagnostics of PrintMinMaxDiagnostics_ (Z4_Lunar/src/tools/LunarSolver_PrintDiag.cpp:7-60,
// Z5_Mercury/src/tools/MercurySolver_PrintDiag.cpp:7-92) over interior cells:
// mins of rho_H, p_H, rho_Na, p_Na, Bx, By, Bz, |B|^2/2, then maxes of the
// same and of |div B| (17 values per case).  Doubles are reduced as
// order-preserving 64-bit keys with atomicMin / atomicMax.
__device__ __forceinline__ unsigned long long order_key(double x)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

// Keys go to out + 17 * slot: slot = case id when sel.n < 0, else the
// position of the block's case in sel.ids (workgroups of other cases return
// at once; a tile lies in one block, so the return is uniform).
__global__ __launch_bounds__(256) void k_diag(const MhdTileMap tm, const MhdCaseConst *__restrict__ cst,
                                              unsigned long long *__restrict__ out, int ucur, const mhdk::DiagSel sel)
{
    MhdBlockDev b;
    int i, j, k;
    bool wg;
    const bool ok = tile_cell(tm, b, i, j, k, wg);
    int slot = wg ? b.case_id : -1;
    if (sel.n >= 0 && wg)
    {
        slot = -1;
        for (int q = 0; q < sel.n; ++q)
            if (sel.ids[q] == b.case_id)
                slot = q;
    }
    if (slot < 0)
        return;
    double v[17];
#pragma unroll
    for (int q = 0; q < 8; ++q)
    {
        v[q] = __longlong_as_double(0x7ff0000000000000ll);      // +inf
        v[8 + q] = __longlong_as_double((long long)0xfff0000000000000ull); // -inf
    }
    v[16] = __longlong_as_double((long long)0xfff0000000000000ull);
    if (ok)
    {
        const MhdCaseConst &c = cst[b.case_id];
        const int64_t sl = b.slot, o = blk_off(b, i, j, k);
        const double *S = dst_ptr(b);
        const double *U = S + (int64_t)(ST_U0 + 5 * ucur) * sl;
        const Prim ph = prim_of(U, sl, o, c.gamma);
        double x[8] = {U[o], ph.p, v[2], v[3], 0.0, 0.0, 0.0, 0.0};
        if (b.model == 1)
        {
            const double *UN = S + (int64_t)(ST_UNA0 + 5 * ucur) * sl;
            x[2] = UN[o];
            x[3] = prim_of(UN, sl, o, c.gamma).p;
        }
        const double *Bc = S + (int64_t)ST_BCELL * sl;
        x[4] = Bc[o];
        x[5] = Bc[sl + o];
        x[6] = Bc[2 * sl + o];
        x[7] = 0.5 * (x[4] * x[4] + x[5] * x[5] + x[6] * x[6]);
#pragma unroll
        for (int q = 0; q < 8; ++q)
        {
            if (b.model != 1 && (q == 2 || q == 3))
                continue;
            v[q] = dmin(v[q], x[q]);
            v[8 + q] = dmax(v[8 + q], x[q]);
        }
        const double *Bx = S + (int64_t)ST_B * sl, *By = Bx + sl, *Bz = Bx + 2 * sl;
        const double dv = (Bx[o + (int64_t)b.sj * b.sk] - Bx[o] + By[o + b.sk] - By[o] + Bz[o + 1] - Bz[o]) /
                          dmetric(b)[(int64_t)G_JAC * sl + o];
        v[16] = dmax(v[16], fabs(dv));
    }
#pragma unroll
    for (int q = 0; q < 17; ++q)
    {
        double m = v[q];
#pragma unroll
        for (int s = 32; s > 0; s >>= 1)
        {
            const double w = __shfl_xor(m, s, 64);
            m = q < 8 ? dmin(m, w) : dmax(m, w);
        }
        v[q] = m;
    }
    if (threadIdx.x == 0 && wg)
    {
#pragma unroll
        for (int q = 0; q < 17; ++q)
        {
            if (q < 8)
                atomicMin(&out[17 * slot + q], order_key(v[q]));
            else
                atomicMax(&out[17 * slot + q], order_key(v[q]));
        }
    }
}

__global__ void k_fill_u64(unsigned long long *p, int n, unsigned long long v)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n)
        p[t] = v;
}

// Neutral keys of n cases' diagnostics: +inf for the 8 minima, -inf for the
// 9 maxima of each case.
__global__ void k_diag_reset(unsigned long long *p, int n17)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n17)
        p[t] = order_key((t % 17) < 8 ? __longlong_as_double(0x7ff0000000000000ll)
                                      : __longlong_as_double((long long)0xfff0000000000000ull));
}

// Scratch of mhd_diagnostics_noalloc: the order keys of up to
// DIAG_SEL_MAX cases, one copy per GPU for the life of the process.  Host
// code reaches it only through the address hipGetSymbolAddress gives for
// the current device (diag_scratch below).
static __device__ __align__(256) unsigned long long scratch[17 * mhdk::DIAG_SEL_MAX];

// ======================================================================
// launch helpers
namespace mhdk
{
static dim3 grid_of(const Launch &L)
{
    const int n = L.tm.xcd_chunk > 0 ? 8 * L.tm.xcd_chunk : L.tm.ntiles;
    return dim3((unsigned)std::max(n, 1));
}
static dim3 block_of(const Launch &L) { return dim3(64, (unsigned)L.jt, 1); }

void launch_dt(const Launch &L, const MhdCaseConst *cst, unsigned long long *dtbits, int ucur, int mixed,
               hipStream_t q)
{
    (void)mixed; // the metric is double in both precisions
    hipLaunchKernelGGL(k_dt, grid_of(L), block_of(L), 0, q, L.tm, cst, dtbits, ucur);
}
void launch_dt_finalize(unsigned long long *dtbits, double *dt, int *err, const MhdCaseConst *cst, int ncase,
                        const int *active, hipStream_t q)
{
    hipLaunchKernelGGL(k_dt_finalize, dim3((ncase + 63) / 64), dim3(64), 0, q, dtbits, dt, err, cst, ncase, active);
}
void launch_ucontra(const Launch &L, const MhdCaseConst *cst, int ucur, int mixed, hipStream_t q)
{
    (void)mixed;
    hipLaunchKernelGGL(k_ucontra, grid_of(L), block_of(L), 0, q, L.tm, cst, ucur);
}
void launch_face_emf(const Launch &L, int mixed, hipStream_t q)
{
    (void)mixed;
    hipLaunchKernelGGL(k_face_emf, grid_of(L), block_of(L), 0, q, L.tm);
}
void launch_edge_emf(const Launch &L, hipStream_t q)
{
    hipLaunchKernelGGL(k_edge_emf, grid_of(L), block_of(L), 0, q, L.tm);
}
void launch_face_shell(const Launch &L, hipStream_t q)
{
    hipLaunchKernelGGL(k_face_shell, grid_of(L), dim3(256), 0, q, L.tm);
}
void launch_emf(const Launch &L, hipStream_t q)
{
    // three workgroups per tile (one per edge family)
    MhdTileMap tm = L.tm;
    const int n = 3 * tm.ntiles;
    tm.xcd_chunk = tm.xcd_chunk > 0 ? (n + 7) / 8 : 0;
    const int nwg = tm.xcd_chunk > 0 ? 8 * tm.xcd_chunk : n;
    hipLaunchKernelGGL(k_emf, dim3((unsigned)std::max(nwg, 1)), block_of(L), 0, q, tm);
}
void launch_curl_update(const Launch &L, const double *dts, hipStream_t q)
{
    hipLaunchKernelGGL(k_curl_update, grid_of(L), block_of(L), 0, q, L.tm, dts);
}
void launch_fluid(const Launch &L, const MhdCaseConst *cst, const double *dts, int ucur, int mixed, int recon,
                  hipStream_t q)
{
    const dim3 g = grid_of(L), bl = block_of(L);
    if (mixed && recon)
        hipLaunchKernelGGL((k_fluid<float, 1>), g, bl, 0, q, L.tm, cst, dts, ucur);
    else if (mixed)
        hipLaunchKernelGGL((k_fluid<float, 0>), g, bl, 0, q, L.tm, cst, dts, ucur);
    else if (recon)
        hipLaunchKernelGGL((k_fluid<double, 1>), g, bl, 0, q, L.tm, cst, dts, ucur);
    else
        hipLaunchKernelGGL((k_fluid<double, 0>), g, bl, 0, q, L.tm, cst, dts, ucur);
}
void launch_fluid_col(const Launch &C, const MhdCaseConst *cst, const double *dts, int ucur, int mixed, hipStream_t q)
{
    const dim3 g = grid_of(C), bl(64, COL_TJ, 1);
    if (mixed)
        hipLaunchKernelGGL((k_fluid_col<float>), g, bl, 0, q, C.tm, cst, dts, ucur);
    else
        hipLaunchKernelGGL((k_fluid_col<double>), g, bl, 0, q, C.tm, cst, dts, ucur);
}
int col_tile_rows() { return COL_TJ; }
int pad_tiles_per_plane(int ext_j, int ext_k, int rows)
{
    const int P = ext_j * (ext_k + 2), nw = (P + PAD_OWN - 1) / PAD_OWN;
    return (nw + rows - 1) / rows;
}
void launch_jcell(const Launch &L, int mixed, hipStream_t q)
{
    if (mixed)
        hipLaunchKernelGGL(k_jcell<float>, grid_of(L), block_of(L), 0, q, L.tm);
    else
        hipLaunchKernelGGL(k_jcell<double>, grid_of(L), block_of(L), 0, q, L.tm);
}

void launch_derived(const Launch &L, const MhdCaseConst *cst, unsigned long long *dtbits, int ucur, int mixed,
                    hipStream_t q)
{
    if (L.tm.kloop)
    {
        if (mixed)
            hipLaunchKernelGGL(k_derived_k<float>, grid_of(L), block_of(L), 0, q, L.tm, cst, dtbits, ucur);
        else
            hipLaunchKernelGGL(k_derived_k<double>, grid_of(L), block_of(L), 0, q, L.tm, cst, dtbits, ucur);
        return;
    }
    if (mixed)
        hipLaunchKernelGGL(k_derived<float>, grid_of(L), block_of(L), 0, q, L.tm, cst, dtbits, ucur);
    else
        hipLaunchKernelGGL(k_derived<double>, grid_of(L), block_of(L), 0, q, L.tm, cst, dtbits, ucur);
}
void launch_derived2(const Launch &L, const MhdCaseConst *cst, unsigned long long *dtbits, int ucur, int mixed,
                     hipStream_t q)
{
    if (mixed)
        hipLaunchKernelGGL(k_derived2<float>, grid_of(L), block_of(L), 0, q, L.tm, cst, dtbits, ucur);
    else
        hipLaunchKernelGGL(k_derived2<double>, grid_of(L), block_of(L), 0, q, L.tm, cst, dtbits, ucur);
}
void launch_fluid2(const Launch &L, const MhdCaseConst *cst, const double *dts, int ucur, int mixed, int recon,
                   hipStream_t q)
{
    const dim3 g = grid_of(L), bl = block_of(L);
    if (mixed && recon)
        hipLaunchKernelGGL((k_fluid2<float, 1>), g, bl, 0, q, L.tm, cst, dts, ucur);
    else if (mixed)
        hipLaunchKernelGGL((k_fluid2<float, 0>), g, bl, 0, q, L.tm, cst, dts, ucur);
    else if (recon)
        hipLaunchKernelGGL((k_fluid2<double, 1>), g, bl, 0, q, L.tm, cst, dts, ucur);
    else
        hipLaunchKernelGGL((k_fluid2<double, 0>), g, bl, 0, q, L.tm, cst, dts, ucur);
}
void launch_fluid2_col(const Launch &C, const MhdCaseConst *cst, const double *dts, int ucur, int mixed,
                       hipStream_t q)
{
    const dim3 g = grid_of(C), bl(64, COL_TJ, 1);
    if (mixed)
        hipLaunchKernelGGL((k_fluid2_col<float>), g, bl, 0, q, C.tm, cst, dts, ucur);
    else
        hipLaunchKernelGGL((k_fluid2_col<double>), g, bl, 0, q, C.tm, cst, dts, ucur);
}
void launch_emf_extra(const Launch &L, const MhdCaseConst *cst, int ucur, int mixed, hipStream_t q)
{
    if (mixed)
        hipLaunchKernelGGL(k_emf_extra<float>, grid_of(L), block_of(L), 0, q, L.tm, cst, ucur);
    else
        hipLaunchKernelGGL(k_emf_extra<double>, grid_of(L), block_of(L), 0, q, L.tm, cst, ucur);
}
void launch_eres(const Launch &L, const MhdCaseConst *cst, hipStream_t q)
{
    hipLaunchKernelGGL(k_eres, grid_of(L), block_of(L), 0, q, L.tm, cst);
}
void launch_dt_res(double *dtres, const double *dt, const MhdCaseConst *cst, int ncase, int isub, hipStream_t q)
{
    hipLaunchKernelGGL(k_dt_res, dim3((ncase + 63) / 64), dim3(64), 0, q, dtres, dt, cst, ncase, isub);
}
void launch_ops(const OpLaunch &L, const MhdCaseConst *cst, hipStream_t q)
{
    long long t = 0;
    for (int g = 0; g < L.nseg; ++g)
        t += (long long)L.seg[g].n << L.seg[g].lg;
    if (t <= 0)
        return;
    hipLaunchKernelGGL(k_ops, dim3((unsigned)((t + 255) / 256)), dim3(256), 0, q, L, cst);
}
void launch_divb(const Launch &L, unsigned long long *out, int mixed, hipStream_t q)
{
    (void)mixed;
    hipLaunchKernelGGL(k_divb, grid_of(L), block_of(L), 0, q, L.tm, out);
}
void launch_diag(const Launch &L, const MhdCaseConst *cst, unsigned long long *out, int ucur, hipStream_t q)
{
    DiagSel all;
    all.n = -1;
    hipLaunchKernelGGL(k_diag, grid_of(L), block_of(L), 0, q, L.tm, cst, out, ucur, all);
}
void launch_diag_sel(const Launch &L, const MhdCaseConst *cst, unsigned long long *out, int ucur, const DiagSel &sel,
                     hipStream_t q)
{
    hipLaunchKernelGGL(k_diag_reset, dim3((17 * sel.n + 255) / 256), dim3(256), 0, q, out, 17 * sel.n);
    hipLaunchKernelGGL(k_diag, grid_of(L), block_of(L), 0, q, L.tm, cst, out, ucur, sel);
}
hipError_t diag_scratch(int device, unsigned long long **ptr)
{
    // looked up once per device, with that device current
    static std::mutex mu;
    static std::map<int, unsigned long long *> by_device;
    std::lock_guard<std::mutex> lock(mu);
    auto it = by_device.find(device);
    if (it == by_device.end())
    {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess)
            return e;
        if (cur != device)
            return hipErrorInvalidDevice;
        unsigned long long *p = nullptr;
        e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return e;
        it = by_device.emplace(device, p).first;
    }
    *ptr = it->second;
    return hipSuccess;
}
void launch_fill_u64(unsigned long long *p, int n, unsigned long long v, hipStream_t q)
{
    hipLaunchKernelGGL(k_fill_u64, dim3((n + 255) / 256), dim3(256), 0, q, p, n, v);
}
} // namespace mhdk
