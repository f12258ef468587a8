// Host-side setup for the explicit MHD step: multi-block grids, topology,
// geometry, mimetic weight tables, halo / boundary operation lists and
// initial states.  Everything here runs once per grid shape or per case;
// the per-step work is on the device.
//
// The C++ types mirror the reference's layers 1-6 (Grid/Block, TOPO
// IndexTransform, FieldBlock extents, Halo regions, BoundaryCore slabs) in
// reduced form: every block face is one patch (interface or physical).
#pragma once
#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "../layout.h"

namespace mpcns
{

struct Int3
{
    int i, j, k;
};

// Node-space index transform, same convention as the reference
// (include/2_topology/2_MPCNS_Topology.h:22-29):
//   nb[perm[a]] = sign[a] * local[a] + offset[a]
struct IndexTransform
{
    int perm[3];
    int sign[3];
    int offset[3];
};

// One face of a block: -i,+i,-j,+j,-k,+k  (face = 2*axis + side).
// One of several interface patches that share a block face (the reference's
// inner patches may cover part of a face, src/1_grid/1_MPCNS_Grid.cpp:392-505):
// the node range it covers along the face's two tangential axes (in axis
// order, inclusive) and its link.
struct FacePart
{
    int lo[2] = {0, 0}, hi[2] = {0, 0};
    int nb_block = -1;
    int nb_face = -1;
    int periodic = 0;
    IndexTransform trans{};
};

struct FaceLink
{
    int kind = 0;     // 0 = physical, 1 = interface
    int bc = BC_NONE; // MhdBc for physical faces
    int nb_block = -1;
    int nb_face = -1;
    int periodic = 0; // periodic interface: ghost nodes are mirrored (reference src/1_grid/1_MPCNS_Grid.cpp:1268)
    IndexTransform trans{};
    // several interface patches (kind 1, one periodic flag): a point crosses
    // the face through the part that covers it (nb_block .. trans above are
    // parts[0]'s); empty: one link for the whole face
    std::vector<FacePart> parts;
};

struct Block
{
    int ni = 0, nj = 0, nk = 0; // cells
    int gn = 0;                 // ghost node layers
    std::vector<double> x, y, z;
    FaceLink face[6];

    int n(int ax) const { return ax == 0 ? ni : (ax == 1 ? nj : nk); }
    int64_t node_off(int i, int j, int k) const
    {
        const int sj = nj + 1 + 2 * gn, sk = nk + 1 + 2 * gn;
        return ((int64_t)(i + gn) * sj + (j + gn)) * sk + (k + gn);
    }
    int64_t node_count() const
    {
        return (int64_t)(ni + 1 + 2 * gn) * (nj + 1 + 2 * gn) * (nk + 1 + 2 * gn);
    }
};

struct Grid
{
    int g = 2; // field ghost depth
    std::vector<Block> blocks;
};

// Per-block arrays in the device layout (see layout.h).
struct BlockArrays
{
    MhdBlockDev d{};
    std::vector<double> geo; // G_NSLOT * slot
    std::vector<double> st;  // st_nslot(d.model) * slot
    std::vector<double> wj;  // 36 * slot: Jcell_from_Jedge_w in full (host only)
    std::vector<double> wb;  // 18 * slot: Bcell_from_Bface_w in full (host only)
};

// One element-wise operation of a boundary / halo phase, or a run of them:
// element e < run acts on dst + e * dstep (reading src + e * sstep).
// Offsets are element offsets relative to the arena the kernel receives.
struct MhdOp
{
    int64_t dst;
    int64_t src;
    int64_t stride; // distance between components (slot size)
    double val;
    int32_t kind;  // MhdOpKind
    int32_t ncomp; // components handled by this op
    int32_t cas;   // case index (constants of OP_QINF / OP_SOLID_*)
    int32_t run;   // elements in the run (1: a single op)
    int32_t dstep; // destination and source step between the run's elements
    int32_t sstep;
    int32_t fold;  // 1: boundary work the step's kernels do themselves (folds_kface_*);
                   // the step skips it, refresh and the per-group syncs apply it too
};

enum MhdOpKind
{
    OP_COPY = 0,        // dst[c] = src[c]
    OP_NEG = 1,         // dst[c] = -src[c]
    OP_ZERO = 2,        // dst[c] = 0
    OP_CONST = 3,       // dst[0] = val
    OP_SOLID_FIRST = 4, // absorbing wall, first interior layer from second
    OP_SOLID_GHOST = 5, // absorbing wall, ghost from first interior layer
    OP_QINF = 6,        // dst[c] = qinf[c] of the case
    OP_QINF_NA = 7,     // dst[c] = qinfs[c] (Na+ far field) of the case
    OP_NA_FLOOR = 8     // Na+ absorbing wall ghost: rho = p = 1e-12, u = 0
};

// Op lists of one case; offsets are relative to the case's first block
// state slot 0, in "global state element" units once the case is placed.
enum MhdSync
{
    SY_EFACE_BC = 0,
    SY_EFACE_HALO,
    SY_EEDGE_BC,
    SY_U_BC0,
    SY_U_BC1,
    SY_U_HALO,
    SY_B_BC,
    SY_B_HALO,
    SY_JCELL_BC,
    SY_UC_CORNER, // corner-diagonal cells take the contravariant velocity of their source
    SY_BCELL_BC,   // two-fluid with Hall: B_cell physical ghosts
    SY_BCELL_HALO, //   B_cell halo
    SY_JCELL_GH,   //   J_cell physical ghosts (copies of the first layer)
    SY_JCELL_HALO, //   J_cell halo
    SY_COUNT
};

struct CaseSetup
{
    Grid grid;
    MhdCaseConst cst{};
    std::vector<BlockArrays> blk;          // geometry + initial state
    // SY_COUNT op lists; offsets are case-relative state offsets
    // (block * ST_NSLOT * slot + state_slot * slot + blk_off)
    std::vector<std::vector<MhdOp>> ops;
    std::string shape_key; // identical keys share geometry (geometry_key; empty: recompute)
    bool tables_edited = false; // the full weight tables were replaced on the host
    double L_ref = 0.0;    // body radius [m] of a Mercury case (neutral exosphere for output)
};

// ---------------------------------------------------------------------------
// Grid generators.
Grid make_periodic_box(int nx, int ny, int nz, double lx, double ly, double lz, int g);
Grid make_cubed_sphere_shell(int nang, int nrad, double r_in, double r_out, int g,
                             int bc_inner, int bc_outer);

// Link interface faces by matching node coordinates (used for multi-block
// grids built from separate panels).
void link_faces_by_coordinates(Grid &grd, double tol);

// Fill ghost nodes: mirror at physical / periodic faces, copy across
// interfaces.
void fill_ghost_nodes(Grid &grd);

// Geometry (reference src/3_field/4_MPCNS_Geometry.cpp) and mimetic weights
// (reference Z4_Lunar/src/LunarSolver.cpp:243-930) into the device layout.
void build_geometry(const Grid &grd, std::vector<BlockArrays> &blk);

// Locate the source of a ghost DOF through interfaces.
struct DofSource
{
    int kind; // 0 interior, 1 physical ghost of src block, 2 corner clamp, 3 own physical slab (not a halo dof)
    int block;
    int loc;
    int idx[3];
    int perm[3];
    int sign[3];
};
DofSource resolve_dof(const Grid &grd, int b, int loc, const int idx[3]);

// Build halo + boundary op lists for a case (offsets relative to the case).
void build_ops(CaseSetup &cs);

// Execute copy/zero/const op lists on host state arrays (start-up syncs).
void apply_ops_host(std::vector<BlockArrays> &blk, const std::vector<MhdOp> &ops, const MhdCaseConst &cst);
// Merge the ops of one list into runs: ops of one kind, component count and
// case whose destinations (and sources) advance by constant steps within one
// block slot become one record.  The ops of a list are independent (no op
// reads what another writes, no two write one element), so their order is
// free; runs are taken along k first, then along j and i for what is left.
void compact_ops(const std::vector<BlockArrays> &blk, std::vector<MhdOp> &ops);

// Field number (mhd_field) -> (region 0 state / 1 geometry / 2 the host-only
// J_cell weight table, first slot, location, components); false when the
// field is not stored.
bool field_slot(int field, int ucur, int &region, int &slot, int &loc, int &nc);
// Elements of a field block in the reference AoS layout incl. ghosts.
size_t aos_count(const MhdBlockDev &d, int loc, int nc);
// Convert between nc consecutive SoA slots and the AoS FieldBlock layout.
void soa_aos(const MhdBlockDev &d, int loc, int nc, double *soa, double *aos, bool to_aos);

// Case builders.
struct WindParams
{
    double n_cc;       // number density [cm^-3]
    double v_kms;      // speed [km/s]
    double vdir[3];    // unit flow direction
    double T_eV;       // temperature [eV]
    double B_nT[3];    // IMF [nT]
    double m_amu;      // ion mass [amu] (H+ = 1)
};

// Offset dipole of the Mercury set-up (nondimensional lengths).
struct DipoleParams
{
    double B_eq_nT;   // equatorial surface field of the dipole [nT]
    double offset_z;  // dipole centre along +z [body radii]
    double softening; // core softening length [body radii]
    bool two_fluid = true;  // H+ and Na+ (the reference's Mercury step)
    bool ambipolar = true;  // electron-pressure force and ambipolar EMF (two-fluid)
    bool hall = true;       // explicit Hall EMF (two-fluid)
    double hall_rmin = 1.1, hall_rmax = 1.2; // radial taper of the Hall EMF
    // local artificial resistivity (two-fluid; off when eta_max <= 0 or r_decay <= 0)
    double loc_eta_max = 0.1, loc_center[3] = {1.53, 0.0, 0.25}, loc_r_decay = 0.05, loc_r_cutoff = 0.05;
    // interior resistivity (two-fluid), explicit subcycles
    bool resist = true;
    int resist_nsub = 1;
    double resist_eta_max = 1.25e7; // eta_max_mercury [m^2/s]
};

CaseSetup make_case_mercury(int nang, int nrad, double r_in, double r_out, int g, double gamma, double cfl,
                            const WindParams &w, const DipoleParams &dp);
// Mercury exosphere: neutral Na density [cm^-3] and Na+ photoionization rate
// at a point (body radii), and the reference's dual-grid cell centre.
double mercury_na_density(const double x[3], double L_ref);
double mercury_photo_rate(const double x[3], double L_ref);
void dual_centre(const Block &B, int i, int j, int k, double out[3]);
CaseSetup make_case_orszag_tang(int n, int g, double gamma, double cfl, uint64_t seed, double noise);
CaseSetup make_case_lunar(int nang, int nrad, double r_in, double r_out, int g, double gamma,
                          double cfl, const WindParams &w);
// Lunar physics and initial state on a given grid (ghost nodes are filled
// here).
CaseSetup make_case_lunar_on(Grid grid, double gamma, double cfl, const WindParams &w);
// The single-fluid (Lunar) constants of a solar wind (calc_physical_constant
// and BuildLunarBackgroundState), and the Lunar set-up on a given grid with
// given constants: geometry, weights, op lists and the initial state
// (U = q_inf, B_face = 0, Badd = B_imf . S).
MhdCaseConst lunar_constants(double gamma, double cfl, const WindParams &w);
CaseSetup make_case_with(Grid grid, const MhdCaseConst &cst);

// Content key of a case's geometry: block count, extents, ghost depth and a
// 64-bit hash of every stored geometry word.  Cases share one device copy of
// their geometry only when their keys are equal.
std::string geometry_key(const std::vector<BlockArrays> &blk);

// The weight tables the kernels form from the stored least-squares data (the
// host builder's arithmetic, so bit-identical to its tables): 18
// Bcell_from_Bface_w weights (x of the six faces, then y, then z) and 36
// Jcell_from_Jedge_w weights (x of the twelve edges, then y, then z) of
// interior cell (i, j, k).
void formed_bcell_weights(const BlockArrays &A, int i, int j, int k, double W[18]);
void formed_jcell_weights(const BlockArrays &A, int i, int j, int k, double W[36]);
// Compare tables given in the AoS FieldBlock layout (18 / 36 components,
// either may be null) with the formed ones over the interior cells: true
// when every weight is within rtol of the cell's largest formed weight, else
// false with the first offending block cell in `why`.
bool weight_tables_match(const BlockArrays &A, const double *wb_aos, const double *wj_aos, double rtol,
                         std::string &why);

// Records why a host-case constructor failed (mhd_host_last_error(NULL)
// reports it on the calling thread).
void host_ctor_error(const std::string &msg);

// Deterministic generator (SplitMix64).
uint64_t splitmix64(uint64_t &s);
double uniform01(uint64_t &s);

} // namespace mpcns
