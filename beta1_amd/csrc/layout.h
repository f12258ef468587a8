// Shared host/device definitions for the explicit MHD step: staggering,
// per-block device layout, slot numbering and the small POD records that
// travel between the host setup code and the HIP kernels.
//
// Staggering follows the reference's StaggerLocation
// (include/3_field/Field_Type.h:4-14) and FieldBlock::compute_extent
// (src/3_field/1_Field_Block.cpp:51-131): a block of Ni x Nj x Nk cells has
// N+1 node-like and N cell-like DOFs per axis, plus `g` ghost layers.
//
// Device layout (ours, not the reference's AoS): every field of a block is
// stored structure-of-arrays, one "slot" per component, and every slot of
// a block shares ONE padded index box (Ni+1+2g) x (Nj+1+2g) x (Nk+1+2g)
// with k fastest.  All locations therefore share one index function and
// neighbour offsets are the constants 1, SK and SJ*SK.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define MHD_HD __host__ __device__ inline
#else
#define MHD_HD inline
#endif

enum MhdLoc
{
    LOC_CELL = 0,
    LOC_NODE = 1,
    LOC_FACE_XI = 2,
    LOC_FACE_ET = 3,
    LOC_FACE_ZE = 4,
    LOC_EDGE_XI = 5,
    LOC_EDGE_ET = 6,
    LOC_EDGE_ZE = 7
};

// 1 = cell-like on that axis, 0 = node-like
// (src/4_halo/0_MPCNS_Halo_Build_BoxMakers.cpp:9-33).
MHD_HD int loc_delta(int loc, int ax)
{
    switch (loc)
    {
    case LOC_CELL:
        return 1;
    case LOC_NODE:
        return 0;
    case LOC_FACE_XI:
    case LOC_FACE_ET:
    case LOC_FACE_ZE:
        return (ax == loc - LOC_FACE_XI) ? 0 : 1;
    default:
        return (ax == loc - LOC_EDGE_XI) ? 1 : 0;
    }
}

// Number of interior DOFs of `loc` along axis `ax` for N cells.
MHD_HD int loc_count(int loc, int ax, int n) { return n + 1 - loc_delta(loc, ax); }

// ---------------------------------------------------------------------------
// State slots (per case, per block).  U is double-buffered: the fluid kernel
// reads U[cur] and writes U[1-cur].
enum MhdStateSlot
{
    ST_U0 = 0,     // 5 comps: rho, rho u, rho v, rho w, E
    ST_U1 = 5,     // 5 comps (second buffer)
    ST_B = 10,     // 3 slots: B_xi (FaceXi), B_eta (FaceEt), B_zeta (FaceZe)
    ST_BADD = 13,  // 3 slots: Badd_xi/eta/zeta
    ST_BCELL = 16, // 3 comps: total cell B
    ST_JCELL = 19, // 3 comps: cell J
    ST_EFACE = 22, // 9 slots: Eface_xi(3 comps), Eface_eta(3), Eface_zeta(3)
    ST_E = 31,     // 3 slots: E_xi (EdgeXi), E_eta, E_zeta
    ST_UC = 34,    // 3 slots: cell contravariant velocity per axis (step scratch)
    ST_NSLOT = 37, // (J edges are formed inside the J_cell pass, never stored)
    // two-fluid (Mercury, H+ and Na+) cases add:
    ST_UNA0 = 37,  // 5 comps: Na+ conservative state (U_Na)
    ST_UNA1 = 42,  // 5 comps (second buffer)
    ST_PHOTO = 47, // Na+ photoionization production rate [cm^-3 s^-1] (static)
    ST_DRAG = 48,  // ion-neutral drag frequency [1/s] (static)
    ST_PE = 49,    // electron pressure p_H + p_Na of cells [-1, N] (step scratch)
    // (edge vectors dr are differences of the geometry's node coordinates)
    ST_TAPER = 50, // 3 slots: radial Hall taper at the xi / eta / zeta edge midpoints (static)
    ST_ETALOC = 53, // 3 slots: local artificial resistivity of the edges (static)
    ST_ETARES = 56, // 3 slots: interior resistivity inver_Rem yita0(r) of the edges (static)
    ST_NSLOT2 = 59
};

// State slots per block for a physics model (0 single fluid, 1 two-fluid).
MHD_HD int st_nslot(int model) { return model == 1 ? (int)ST_NSLOT2 : (int)ST_NSLOT; }

// Geometry slots (shared by every case with the same grid).  The two
// least-squares weight tables of the reference are not stored: both are
// W[c][q] = sum_b Mi[c][b] mu_q s_q[b] over the active terms q, with Mi the
// (symmetric) truncated inverse of the normal matrix.  Bcell_from_Bface_w
// (18 per cell): s_q = the oriented area vector of face q, which the B_cell
// kernel loads anyway, mu_q its distance weight; Jcell_from_Jedge_w (36 per
// cell): s_q = the edge vector (difference of two node coordinates), mu_q = 1.
// The kernels form each weight with the host builder's arithmetic, so they
// are bit-identical to the tables (which host cases still hold in full for
// the C-ABI), from 12 and 7 stored values instead of 18 and 36.
enum MhdGeoSlot
{
    G_JAC = 0,     // cell volume
    G_JD = 1,      // 9 slots: JDxi(3), JDet(3), JDze(3) face area vectors
    G_NODE = 10,   // 3 slots: coordinates x, y, z of node (i, j, k)
    G_BETA = 13,   // 3 slots: face Hodge weight beta = |l*|/|S|
    G_ALPHA = 16,  // 3 slots: edge Hodge weight alpha = |e|/|S*|
    G_WBM = 19,    // 6 comps: Mi of the B_cell least squares (00, 01, 02, 11, 12, 22)
    G_WBU = 25,    // 6 comps: mu_q of the six faces (-xi, +xi, -eta, +eta, -zeta, +zeta), 0 if inactive
    G_WJM = 31,    // 6 comps: Mi of the J_cell least squares (00, 01, 02, 11, 12, 22)
    G_WJA = 37,    // 1 comp: active J edges of the cell, bit q for edge q (0..11)
    G_NSLOT = 38,
    G_NMETRIC = 13 // slots [0, G_NMETRIC): metric and nodes, always double
};

// Geometry storage of a block: the metric slots in double, then the Hodge
// weights and the weight tables (slots G_NMETRIC..G_NSLOT) in elements of
// `wsize` bytes (8, or 4 where the tables are stored as float).
MHD_HD int64_t geo_slot_byte(int s, int64_t slot, int wsize)
{
    return s < G_NMETRIC ? (int64_t)s * slot * 8
                         : (int64_t)G_NMETRIC * slot * 8 + (int64_t)(s - G_NMETRIC) * slot * wsize;
}
MHD_HD int64_t geo_block_bytes(int64_t slot, int wsize) { return geo_slot_byte(G_NSLOT, slot, wsize); }

// Physical boundary kinds on a block face.
enum MhdBc
{
    BC_NONE = 0, // interface (halo) or unused
    BC_FARFIELD = 1,
    BC_SOLID = 2
};

// Per-block record seen by every kernel.
struct MhdBlockDev
{
    int ni, nj, nk;  // cells
    int g;           // ghost depth
    int sj, sk;      // strides: off = ((i+g)*sj + (j+g))*sk + (k+g)
    int case_id;     // which case (dt, constants) this block belongs to
    int bc[6];       // BC kind per face: -i,+i,-j,+j,-k,+k
    int64_t slot;    // elements per slot
    int64_t st;      // device address of slot 0 of the block's state (0 before placement)
    int64_t geo;     // device address of slot 0 of the block's geometry
    int model;       // 0 single fluid (Lunar step), 1 two-fluid H+ / Na+ (Mercury step)
    int pad_;
};

// Shell pass of the one-pass face/edge EMF of a whole step.  The halo copy
// of the Eface sync fills a ghost face one layer outside an interface side
// from the neighbour face it coincides with, which lies on the neighbour's
// first layer inside its own interface side: index 0 along the axis of a low
// side, n - 1 of a high one.  Those layers of every interface side (bc
// BC_NONE) are the shell; the shell pass forms their face candidates before
// the sync, the one-pass kernel forms every in-range candidate itself.
// With every extent >= 2 the layers of an axis are distinct.
struct ShellLayers
{
    int n;    // 0, 1 or 2 layers along the axis
    int v[2]; // ascending
};
MHD_HD ShellLayers shell_layers(const MhdBlockDev &b, int ax)
{
    const int n = ax == 0 ? b.ni : (ax == 1 ? b.nj : b.nk);
    ShellLayers L;
    L.n = 0;
    L.v[0] = L.v[1] = 0;
    if (b.bc[2 * ax] == BC_NONE)
        L.v[L.n++] = 0;
    if (b.bc[2 * ax + 1] == BC_NONE)
        L.v[L.n++] = n - 1;
    return L;
}
// The m-th index of [0, n] that is not a layer of L.
MHD_HD int shell_skip(const ShellLayers &L, int m)
{
    for (int q = 0; q < L.n; ++q)
        if (m >= L.v[q])
            ++m;
    return m;
}
// Shell points of the face range [0, N]^3: the layer planes along i in full;
// in each other plane the layer rows along j in full and the layer points
// along k of the other rows.
MHD_HD int shell_count(const MhdBlockDev &b)
{
    const ShellLayers Li = shell_layers(b, 0), Lj = shell_layers(b, 1), Lk = shell_layers(b, 2);
    const int P = (b.nj + 1) * (b.nk + 1), Q = Lj.n * (b.nk + 1) + (b.nj + 1 - Lj.n) * Lk.n;
    return Li.n * P + (b.ni + 1 - Li.n) * Q;
}
MHD_HD void shell_point(const MhdBlockDev &b, int t, int &i, int &j, int &k)
{
    const ShellLayers Li = shell_layers(b, 0), Lj = shell_layers(b, 1), Lk = shell_layers(b, 2);
    const int nk1 = b.nk + 1, P = (b.nj + 1) * nk1, Q = Lj.n * nk1 + (b.nj + 1 - Lj.n) * Lk.n;
    if (t < Li.n * P)
    {
        const int p = t / P, r = t - p * P;
        i = Li.v[p];
        j = r / nk1;
        k = r - j * nk1;
        return;
    }
    t -= Li.n * P;
    const int m = t / Q;
    i = shell_skip(Li, m);
    int r = t - m * Q;
    if (r < Lj.n * nk1)
    {
        const int row = r / nk1;
        j = Lj.v[row];
        k = r - row * nk1;
        return;
    }
    r -= Lj.n * nk1;
    const int mr = r / Lk.n;
    j = shell_skip(Lj, mr);
    k = Lk.v[r - mr * Lk.n];
}
MHD_HD bool in_shell(const MhdBlockDev &b, int i, int j, int k)
{
    const int p[3] = {i, j, k};
    for (int ax = 0; ax < 3; ++ax)
    {
        const ShellLayers L = shell_layers(b, ax);
        for (int q = 0; q < L.n; ++q)
            if (p[ax] == L.v[q])
                return true;
    }
    return false;
}

// Where the one-pass EMF kernel takes the candidates of a face of family a
// (0 xi, 1 eta, 2 zeta) at point (i, j, k), as the Eface sync leaves them:
// 0 formed from the face's cells (an in-range face), 1 zero (a solid wall's
// inner slab and the ghost faces beyond it, which the wall handler zeroes),
// 2 the Eface array (another ghost face: the halo copy, or a value no kept
// edge depends on).
//
// face_source is composed of per-axis parts so that a kernel can form the
// (j, k) part of a point once and the i part once per plane: along axis t a
// face index p is out of range (bit 0), out beyond a solid wall (bit 1) or
// on a solid wall's inner slab (bit 2).
MHD_HD int face_axis_bits(const MhdBlockDev &b, int a, int t, int p)
{
    const int n = t == 0 ? b.ni : (t == 1 ? b.nj : b.nk);
    const int cnt = n + (t == a ? 1 : 0);
    const bool lo = p < 0, hi = p >= cnt;
    const bool wlo = b.bc[2 * t] == BC_SOLID, whi = b.bc[2 * t + 1] == BC_SOLID;
    const bool out = lo || hi;
    const bool wall_out = out && (lo ? wlo : whi);
    const bool slab = !out && ((wlo && p == 0) || (whi && p == cnt - 1));
    return (out ? 1 : 0) | (wall_out ? 2 : 0) | (slab ? 4 : 0);
}
// (j, k) part: axes out of range (bits 0-1), the one out of range is beyond
// a wall (bit 2), on a wall's slab (bit 3)
MHD_HD int face_jk_bits(const MhdBlockDev &b, int a, int j, int k)
{
    const int bj = face_axis_bits(b, a, 1, j), bk = face_axis_bits(b, a, 2, k);
    const int nout = (bj & 1) + (bk & 1);
    const bool wall_out = ((bj & 3) == 3) || ((bk & 3) == 3);
    const bool slab = ((bj | bk) & 4) != 0;
    return nout | (wall_out ? 4 : 0) | (slab ? 8 : 0);
}
MHD_HD int face_source_of(int ibits, int jkbits)
{
    const int nout = (ibits & 1) + (jkbits & 3);
    if (nout == 0)
        return ((ibits & 4) || (jkbits & 8)) ? 1 : 0;
    if (nout == 1)
        return ((ibits & 1) ? (ibits & 2) : (jkbits & 4)) ? 1 : 2;
    return 2;
}
MHD_HD int face_source(const MhdBlockDev &b, int a, int i, int j, int k)
{
    return face_source_of(face_axis_bits(b, a, 0, i), face_jk_bits(b, a, j, k));
}

// Boundary work on the k faces (faces 4 and 5: k is the lanes' axis of
// every kernel) that the step's kernels do themselves, so the step skips
// those ops (refresh and the per-group syncs still apply them; re-applying
// them is idempotent).  A solid wall's B_face ghost copies (k_curl_update,
// every block); the single-fluid U wall layer, its ghosts and far-field
// ghosts (k_fluid_col, first-order scheme), where the wall cell's neighbour
// is the next lane of the same 64-lane wavefront -- not for a high wall whose
// last cell opens a tile.  (Zeroing the wall's face EMFs inside k_face_emf
// cost that kernel more than the op launch it saved.)
MHD_HD bool folds_kface_b(const MhdBlockDev &b, int face)
{
    return (face == 4 || face == 5) && b.bc[face] == BC_SOLID;
}
MHD_HD bool folds_kface_u(const MhdBlockDev &b, int face)
{
    if (b.model != 0 || (face != 4 && face != 5) || (b.bc[face] != BC_SOLID && b.bc[face] != BC_FARFIELD))
        return false;
    return face == 4 || b.bc[face] != BC_SOLID || (b.nk - 1) % 64 != 0;
}

// Per-case constants (the reference's calc_physical_constant,
// Z4_Lunar/src/physics/LunarSolver_Tools.cpp:3-45).
struct MhdCaseConst
{
    double gamma;
    double cfl;
    double inver_MA2;           // B_ref^2/(U_ref^2 mu0 rho_ref)
    double momentum_hall_coeff; // B_ref^2/(mu0 rho_ref U_ref^2)
    double M_H;                 // ion molar mass [kg/mol]
    double M_ref;               // reference molar mass [kg/mol]
    double ne_floor;            // ne_hall_floor (nondimensional)
    double state_coeff;         // p = rho T coeff
    double qinf[5];             // far-field conservative state
    double B_imf[3];            // far-field total B
    double u_inf[3];            // far-field velocity (primitive)
    double dt_floor;            // abort threshold (reference: 1e-7)
    // two-fluid step (Z5_Mercury/src/physics/MercurySolver_Tools.cpp:3-46,
    // MercurySolver_FluidRHS.cpp:210-225, MercurySolver_MagExplicit_Ambipo.cpp:8-9)
    double model;          // 0 single fluid, 1 two-fluid
    double ambipolar;      // 1: electron-pressure force and ambipolar edge EMF
    double M_Na;           // Na+ molar mass [kg/mol]
    double state_coeff_Na; // p = rho T coeff of Na+
    double qinfs[5];       // Na+ far-field (seed) state
    double induce;         // momentum_induce_coeff
    double a1_Na, a4, a5, a6; // photoionization mass / drag / thermal coefficients
    double T_ref;          // reference temperature [K]
    double Tn0;            // neutral temperature [K]
    double pe_coef;        // ambipolar pressure coefficient
    // explicit Hall EMF (MercurySolver_MagExplicit_Hall.cpp) and local
    // artificial resistivity (MercurySolver_MagExplicit_Resis.cpp:67-187)
    double hall;           // 1: explicit Hall EMF on
    double hall_coef;      // B_ref M_ref / (U_ref q_e L_ref mu0 rho_ref N_A)
    double hall_rmin, hall_rmax; // radial taper of the Hall EMF [body radii]
    double local_arti;     // 1: local artificial resistivity on
    double loc_eta_max, loc_center[3], loc_r_decay, loc_r_cutoff;
    // interior resistivity by explicit subcycles (ResistiveDiffusionSubcycles_,
    // Z5_Mercury/src/physics/MercurySolver_MagneticRHS.cpp:79-141)
    double resist;         // 1: on
    double resist_nsub;    // subcycles per step
    double inver_Rem;      // eta_max / (mu0 U_ref L_ref)
};

MHD_HD double *st_ptr(const MhdBlockDev &b) { return reinterpret_cast<double *>((uintptr_t)b.st); }

// Launch tiling: a workgroup of 64 x jt threads covers 64 * jt consecutive
// points (k fastest) of one (block, i) plane of a launch range;
// block_start[e] is the first tile of entry e (ascending, nblk entries) and
// a tile finds its entry by binary search.  An entry is one block, or, in a
// geometry-shared map (members != 0), the same block of the K cases that
// share one geometry set: entry e then covers lead_k[e] blocks,
// blks[members[lead_moff[e] + m]] for m < lead_k[e], and its tiles run the K
// cases of one geometry tile back to back, so the K - 1 later reads of that
// geometry tile find it in L2.
// With xcd_chunk > 0 workgroup w runs tile (w % 8) * xcd_chunk + w / 8, so
// each of the 8 XCDs (workgroups are dealt round-robin) sweeps a contiguous
// range of tiles and neighbour planes stay in its own L2.
struct MhdTileMap
{
    const MhdBlockDev *blks;
    const int *block_start;
    const int *lead_k;    // geometry-shared map only (else null)
    const int *lead_moff; //
    const int *members;   //
    int nblk;
    int ntiles;
    int lo, hi; // range [lo, n + hi) on every axis
    int xcd_chunk;
    int i_inner; // 1: consecutive tiles step along i, 0: along the (j, k) plane
    int ci;      // column and plane-chunk tiles: planes per workgroup along i
    int kloop;   // geometry-shared map: 1 = a tile covers every case of its entry (the kernel
                 // loops over them with the geometry loaded once), 0 = one case per tile
    int fold;    // 1: the kernel applies the k-face boundary work of folds_kface_* itself (a
                 // whole step, which then skips those op records); 0: the syncs do it
};

MHD_HD int64_t blk_off(const MhdBlockDev &b, int i, int j, int k)
{
    return ((int64_t)(i + b.g) * b.sj + (j + b.g)) * b.sk + (k + b.g);
}
