/* beta1_mhd.h -- C-ABI of the MI355X explicit MHD step.
 *
 * Drop-in boundary for the explicit step of the reference's solver drivers
 * (Z4_Lunar, Z5_Mercury).  Each entry point names the reference interface it
 * replaces.  Conventions: int status (0 = ok, negative = error, never
 * exit()); caller-owned host buffers; device memory owned by the context;
 * one HIP stream per context; a context is not thread-safe.
 *
 * Field numbering (MHD_F_*) follows the reference's registered field names
 * (Z4_Lunar/main.cpp:61-121, src/3_field/4_MPCNS_Geometry.cpp:6-73); host
 * buffers use the reference AoS FieldBlock layout [i][j][k][m] over the
 * FieldBlock extent with `ghost` layers (src/3_field/1_Field_Block.cpp:51-131).
 */
#ifndef BETA1_MHD_H
#define BETA1_MHD_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mhd_ctx mhd_ctx;

enum mhd_status
{
    MHD_OK = 0,
    MHD_ERR_ARG = -1,   /* bad argument / unknown id */
    MHD_ERR_HIP = -2,   /* HIP runtime error */
    MHD_ERR_DT = -3,    /* dt NaN or below the abort floor
                           (LunarSolver_Timestep.cpp:128-139) */
    MHD_ERR_STATE = -4, /* call out of order (e.g. step before finalize) */
    MHD_ERR_MEM = -5    /* device memory budget exceeded */
};

enum mhd_precision
{
    MHD_PREC_FP64 = 0, /* plain FP64 step, bit-compatible with the reference */
    MHD_PREC_MIXED = 1 /* FP32 flux / EMF arithmetic, FP64 state, metric, curl, dt */
};

enum mhd_field
{
    MHD_F_U_H = 0, MHD_F_PV_H, MHD_F_U_PLUS, MHD_F_B_XI, MHD_F_B_ETA, MHD_F_B_ZETA,
    MHD_F_BADD_XI, MHD_F_BADD_ETA, MHD_F_BADD_ZETA, MHD_F_B_CELL, MHD_F_BIND_CELL,
    MHD_F_J_XI, MHD_F_J_ETA, MHD_F_J_ZETA, MHD_F_J_CELL, MHD_F_E_XI, MHD_F_E_ETA, MHD_F_E_ZETA,
    MHD_F_EFACE_XI, MHD_F_EFACE_ETA, MHD_F_EFACE_ZETA, MHD_F_F_XI, MHD_F_F_ETA, MHD_F_F_ZETA,
    MHD_F_RHS_H, MHD_F_RHS_B_XI, MHD_F_RHS_B_ETA, MHD_F_RHS_B_ZETA, MHD_F_DIVB,
    MHD_F_JAC, MHD_F_JDXI, MHD_F_JDET, MHD_F_JDZE, MHD_F_BETA_XI, MHD_F_BETA_ETA, MHD_F_BETA_ZETA,
    MHD_F_ALPHA_XI, MHD_F_ALPHA_ETA, MHD_F_ALPHA_ZETA, MHD_F_BCELL_W, MHD_F_JCELL_W,
    /* two-fluid (Mercury) cases: Z5_Mercury/main.cpp:54-55 and the Na+
       source fields of Z5_Mercury/include/2_Initial.h:162-392; drag_rate is
       the ion-neutral drag frequency of each cell (a field of this build) */
    MHD_F_U_NA, MHD_F_PV_NA, MHD_F_RHS_NA, MHD_F_PHOTO_RATE, MHD_F_DRAG_RATE, MHD_F_COUNT
};

typedef struct
{
    int device;    /* HIP device ordinal */
    int precision; /* mhd_precision */
    int recon;     /* 0 first order (reference), 1 minmod MUSCL */
    int ghost;     /* ghost layers (>= 2; the first-order stencil reads one, minmod two) */
    double hbm_budget_gb; /* admission budget: bytes of state + geometry (0 = no limit) */
    int autotune;         /* 1: time workgroup heights per kernel at mhd_finalize */
    int xcd_order;        /* 1: deal tiles to the 8 XCDs in contiguous ranges */
    int overlap;          /* 1: run the fluid update beside the EMF / curl chain and J_cell beside
                             the other derived fields, on a second stream of the context (they
                             share no array); results are unchanged */
} mhd_options;

/* Orszag-Tang periodic box (n^3 cells on [0,1)^3). */
typedef struct
{
    int n;
    double gamma, cfl;
    unsigned long long seed;
    double noise; /* relative uniform noise on rho and p */
} mhd_ot_desc;

/* Solar-wind shell around an absorbing insulating body: 6-block cubed
   sphere, nang x nang x nrad cells per block, log-stretched radius. */
typedef struct
{
    int nang, nrad;
    double r_in, r_out, gamma, cfl;
    double n_cc, v_kms, vdir[3], T_eV, B_nT[3], m_amu;
} mhd_wind_desc;

/* Planetary dipole of the Mercury set-up: moment along -z with equatorial
   surface field B_eq_nT, centre offset_z body radii along +z, core softening
   length (body radii) in the vector potential. */
typedef struct
{
    double B_eq_nT, offset_z, softening;
    int single_fluid; /* 1: the shared single-fluid step with the dipole
                         background; 0: H+ and Na+ (MercurySolver's step) */
    int no_ambipolar; /* two-fluid only: 1 turns the electron-pressure force
                         and ambipolar EMF off (the reference's
                         is_Ambipolar_Efield 0) */
    int no_hall;      /* two-fluid only: 1 turns the explicit Hall EMF off */
    double hall_r_min, hall_r_max; /* radial taper of the Hall EMF [body radii] */
    /* local artificial resistivity eta_max exp(-r / r_decay) within r_cutoff
       of a centre (body radii); off when eta_max or r_decay is not > 0 */
    double local_eta_max, local_center[3], local_r_decay, local_r_cutoff;
    /* interior resistivity eta_max_mercury [m^2/s] in 0.84 < r < 1.04,
       applied by resist_subcycles (>= 1) explicit subcycles per step (the
       reference's ResistiveDiffusionSubcycles_; its PETSc solve is not
       built); no_resistance = 1 or eta 0 turns it off */
    int no_resistance, resist_subcycles;
    double resist_eta_max;
} mhd_dipole_desc;

typedef struct mhd_host_case mhd_host_case;

/* Context lifetime.  Replaces the solver constructor's set-up
   (LunarSolver::LunarSolver, Z4_Lunar/src/LunarSolver.cpp:12-235). */
int mhd_create(const mhd_options *opt, mhd_ctx **out);
int mhd_destroy(mhd_ctx *ctx);
const char *mhd_last_error(mhd_ctx *ctx);

/* Case set-up: grid, topology (Grid_Preprocess + TOPO::build_topology,
   include/1_grid/1_MPCNS_Grid.h:152, include/2_topology/2_MPCNS_Topology.h:154),
   geometry (Field::build_geometry, src/3_field/4_MPCNS_Geometry.cpp:4),
   weight tables (SetupHallFaceScratch_, Z4_Lunar/src/LunarSolver.cpp:243),
   halo patterns (Halo::build_registered_patterns, include/4_halo/1_MPCNS_Halo.h:9-43),
   boundary registry (LunarBoundary::Setup, Z4_Lunar/include/1_Boundary.h:29)
   and the initial state (Lunar_Initial::Initialization, Z4_Lunar/include/2_Initial.h:16). */
int mhd_add_case_orszag_tang(mhd_ctx *ctx, const mhd_ot_desc *d, int *case_id);
int mhd_add_case_lunar(mhd_ctx *ctx, const mhd_wind_desc *d, int *case_id);
/* Mercury-configured shell (Z5_Mercury/main.cpp set-up with the reference
   values of Z5_Mercury/9999setup): solar wind as for the Lunar case plus an
   offset dipole in the background field Badd (DipoleField,
   Z5_Mercury/include/operators/Dipole.h:89-333).  Single fluid: the step is
   the shared explicit step (MercurySolver::StepOnce,
   Z5_Mercury/src/MercurySolver_Driver.cpp:21-92, without the Na+ species,
   Hall, ambipolar and resistive terms). */
int mhd_add_case_mercury(mhd_ctx *ctx, const mhd_wind_desc *d, const mhd_dipole_desc *dp, int *case_id);
/* Move a case built by mhd_host_case_* into the context (same effect as
   mhd_add_case_*; the host set-up can run on another thread meanwhile).
   On MHD_ERR_MEM the host case is left intact for a later attempt; on
   success its arrays are consumed and only mhd_host_case_destroy remains
   valid on it.  No counterpart in the reference (its set-up runs once, in
   the solver constructor). */
int mhd_adopt_host_case(mhd_ctx *ctx, mhd_host_case *h, int *case_id);
/* Lunar physics and initial state on a grid read from a reference CASE
   directory: <dir>/geometry/grid0000.grd (Plot3D binary,
   Grid::Read_Grid_Binary, src/1_grid/1_MPCNS_Grid.cpp:123-221) and
   <dir>/geometry/boundary_condition/{boundary0000.txt, boundary_name.fvbnd,
   inner0000.txt} (Grid::Read_Phy_Boundary / Read_Phy_Name /
   Read_Inner_Boundary, :224-505), interface transforms derived as
   TOPO::build_topology does (src/2_topology/2_MPCNS_Topology.cpp:77-165).
   Farfield and Solid_Surface conditions, one per face (patches of one
   type may tile a face); a face may be split among several interface
   patches (mhd_get_faces kind 2, mhd_host_face_parts), not mix a condition
   with interfaces.  The grid fields of d (nang, nrad, r_in, r_out) are
   ignored. */
int mhd_add_case_from_dir(mhd_ctx *ctx, const char *dir, const mhd_wind_desc *d, int *case_id);
/* A case from the caller's own grid, topology and geometry: what a reference
   driver holds after Grid::Grid_Preprocess, TOPO::build_topology and the
   Field constructor (Z4_Lunar/main.cpp:46-123), handed over block by block
   instead of a generated grid.  Single-fluid (Lunar) physics.
   mhd_case_begin replaces the Field / LunarSolver constructor's set-up for
   nblocks blocks (src/3_field/2_MPCNS_Field.cpp, Z4_Lunar/src/LunarSolver.cpp:
   12-235) with the case constants consts22 in the order mhd_get_constants
   reports them (gamma, cfl, inver_MA2, momentum_hall_coeff, M_H, M_ref,
   ne_floor, state_coeff, qinf[5], B_imf[3], u_inf[3], dt_floor, 2 unused;
   calc_physical_constant, Z4_Lunar/src/physics/LunarSolver_Tools.cpp:3-45).
   mhd_add_block gives block ib (Block, include/1_grid/1_MPCNS_Grid.h:34-125):
   ni x nj x nk cells, the (ni+1)(nj+1)(nk+1) interior node coordinates
   x, y, z ([i][j][k], k fastest), and optionally the geometry the caller's
   Field holds (Field::build_geometry, src/3_field/4_MPCNS_Geometry.cpp:4;
   SetupHallFaceScratch_, Z4_Lunar/src/LunarSolver.cpp:243-930): geometry[q]
   for q = 0..11 is the AoS FieldBlock of MHD_F_JAC + q (Jac, JDxi, JDet,
   JDze, beta_xi/eta/zeta, alpha_xi/eta/zeta, Bcell_from_Bface_w,
   Jcell_from_Jedge_w) with the context's ghost depth, or NULL.
   mhd_set_topology gives its six faces (-i,+i,-j,+j,-k,+k) as 14 ints each,
   the layout of mhd_get_faces: {kind (0 physical, 1 interface, 2 several
   interface patches), bc (1 Farfield, 2 Solid_Surface), nb_block, nb_face,
   periodic, perm[3], sign[3], offset[3]} with the reference's IndexTransform
   nb[perm[a]] = sign[a]*local[a] + offset[a]
   (include/2_topology/2_MPCNS_Topology.h:24-151); the other fields of a
   kind 2 face are ignored.
   mhd_set_face_parts gives the nparts >= 2 interface patches of a face set
   to kind 2, 16 ints each in the layout of mhd_host_face_parts: the cell
   range [lo0, hi0) x [lo1, hi1) of the face's two tangential axes (j, k on
   an i face; i, k on a j face; i, j on a k face), then nb_block, nb_face,
   periodic, perm[3], sign[3], offset[3] (the inner-file patches of one face,
   Grid::Read_Inner_File, src/1_grid/2_MPCNS_Grid_Read.cpp).  The patches
   cover every cell face of the face once and are all periodic or all not.
   mhd_case_end checks that every interface (each patch of a split face)
   links back, each of its cell faces mapped back onto itself by the
   neighbour's patch, and maps each of its nodes onto a node at the same
   position (periodic links excepted), fills
   the ghost nodes (mirror at physical faces, copy across interfaces), builds
   the geometry, weights, halo / boundary ops and the Lunar initial state
   (U = qinf, B_face = 0, Badd = B_imf . S; Z4_Lunar/include/2_Initial.h:
   16-99), and compares every geometry array the caller gave with the one its
   nodes give (rel 1e-9 of the array's largest value; MHD_ERR_ARG naming the
   array and cell otherwise).  The library's own values are kept.  Upload the
   driver's state with mhd_upload and call mhd_refresh.  After mhd_finalize,
   mhd_case_end admits the case like mhd_add_case_*. */
int mhd_case_begin(mhd_ctx *ctx, int nblocks, const double *consts22, int *case_id);
int mhd_add_block(mhd_ctx *ctx, int case_id, int ib, int ni, int nj, int nk, const double *x, const double *y,
                  const double *z, size_t nnode, const double *const *geometry);
int mhd_set_topology(mhd_ctx *ctx, int case_id, int ib, const int *faces14x6);
int mhd_set_face_parts(mhd_ctx *ctx, int case_id, int ib, int face, int nparts, const int *parts16);
int mhd_case_end(mhd_ctx *ctx, int case_id);

/* Upload everything added so far and compute the derived fields
   (UpdateDerivedFields_, Z4_Lunar/src/physics/LunarSolver_Tools.cpp:143).
   After mhd_finalize, mhd_add_case_* admits a case into the running context
   (between steps): it fails with MHD_ERR_MEM when the HBM budget is full. */
int mhd_finalize(mhd_ctx *ctx);
/* Free a case's device state (its geometry when no other case shares it). */
int mhd_retire(mhd_ctx *ctx, int case_id);
/* Steps a case has been advanced by. */
int mhd_case_steps(mhd_ctx *ctx, int case_id, long long *steps);
/* 1 in *dt_error when a step of the case met the dt abort condition
   (NaN or below the floor, LunarSolver_Timestep.cpp:128-139: where the
   reference exits, mhd_step returns MHD_ERR_DT while such a case is
   resident; the other cases keep their own dt and are advanced normally).
   Cleared by mhd_retire and mhd_refresh.  0 for a retired case. */
int mhd_case_error(mhd_ctx *ctx, int case_id, int *dt_error);
/* Min/max diagnostics over the interior cells of a case
   (PrintMinMaxDiagnostics_, Z4_Lunar/src/tools/LunarSolver_PrintDiag.cpp:7-60,
   Z5_Mercury/src/tools/MercurySolver_PrintDiag.cpp:7-92): out[0..7] minima
   and out[8..15] maxima of rho_H, p_H, rho_Na, p_Na, B_cell x, y, z and
   |B_cell|^2/2, out[16] max |div B| (Na entries +inf / -inf for a single
   fluid). */
int mhd_diagnostics(mhd_ctx *ctx, int case_id, double *out17);
/* No counterpart in the reference.  mhd_diagnostics of several cases in one
   pass, for callers that cannot allocate device memory (a serving loop that
   polls its resident cases between steps, a harness that forbids hipMalloc):
   out + 17*q receives case_ids[q]'s 17 values, 1 <= n <= 64 distinct active
   cases (MHD_ERR_ARG otherwise).  It allocates nothing: the order keys live in
   one static device array of 17 * 64 keys (8.7 KB) per GPU, held until the
   process ends, and each call resets what it uses.  Calls that overlap on one
   GPU (other threads, other contexts on the same device) race on that array. */
int mhd_diagnostics_noalloc(mhd_ctx *ctx, const int *case_ids, int n, double *out);

/* Advance every case by nsteps explicit steps, each case with its own dt
   (LunarSolver::StepOnce, Z4_Lunar/src/LunarSolver_Driver.cpp:21-92).
   Returns MHD_ERR_DT if any case hit the dt abort condition. */
int mhd_step(mhd_ctx *ctx, int nsteps);
/* Same as mhd_step through a captured HIP graph of two steps. */
int mhd_step_graph(mhd_ctx *ctx, int nsteps);
/* The stages of mhd_step one at a time, over every active case, in the order
   of LunarSolver::StepOnce (Z4_Lunar/src/LunarSolver_Driver.cpp:21-92):
     mhd_step_dt        Compute_Timestep (LunarSolver_Timestep.cpp:8-140): each
                        case's dt from the min the last derived pass took;
                        MHD_ERR_DT as mhd_step
     mhd_fluid_rhs      Scheme_U_ + AddSourceToRHS_Fluid
                        (LunarSolver_FluidRHS.cpp:3-201,528-568) and the U part of
                        ApplyUpdate_Euler_: U + dt RHS into the context's second U
                        buffer (MHD_F_U_H still reads the current one)
     mhd_face_emf       AssembleOneDirectionEMF_, three directions
                        (LunarSolver_MagExplicit_Ideal.cpp:348-496)
     mhd_sync "Eface"   LunarBoundary::Sync (Z4_Lunar/src/boundary/Boundary_Sync_.cpp:4-65)
     mhd_edge_emf       AssembleEdgeEMF_FromFaceE_Ideal_ (:53-287) and the
                        two-fluid EMF terms (MercurySolver_MagneticRHS.cpp:30-33)
     mhd_sync "Eedge"
     mhd_curl_update    CurlEdgeToFace(-1) (Z4_Lunar/src/operators/CTOperators.cpp:6-56)
                        and the B part of ApplyUpdate_Euler_
     mhd_apply_update   ResistiveDiffusionSubcycles_ (two-fluid cases), then the
                        updated U becomes current; ends the step (step count)
     mhd_sync "Ucell", "Bface"
     mhd_derived        UpdateDerivedFields_ (LunarSolver_Tools.cpp:143-161) and
                        the min of the next dt
   The sequence gives mhd_step's result bit for bit.  mhd_sync takes the
   reference's group names; "Eface", "Eedge", "Ucell", "Bface", "B_cell" and
   "J_cell" run the device's op lists (the last two are empty unless a
   two-fluid case with the Hall EMF is resident), "Jedge", "Uplus", "Ehall"
   and "Badd" are accepted and do nothing (the device forms U_plus from U
   where it is used and never reads those ghosts).  Between mhd_fluid_rhs and
   mhd_apply_update a step is open: mhd_step, admission, retirement, refresh
   and tuning return MHD_ERR_STATE. */
int mhd_step_dt(mhd_ctx *ctx);
int mhd_fluid_rhs(mhd_ctx *ctx);
int mhd_face_emf(mhd_ctx *ctx);
int mhd_edge_emf(mhd_ctx *ctx);
int mhd_curl_update(mhd_ctx *ctx);
int mhd_apply_update(mhd_ctx *ctx);
int mhd_sync(mhd_ctx *ctx, const char *group);
int mhd_derived(mhd_ctx *ctx);
/* Compute_Timestep (Z4_Lunar/src/physics/LunarSolver_Timestep.cpp:8-140)
   on the current state, without stepping. */
int mhd_compute_dt(mhd_ctx *ctx, int case_id, double *dt);
/* dt used by the last step of a case. */
int mhd_last_dt(mhd_ctx *ctx, int case_id, double *dt);
/* max |div B| over interior cells (calc_divB, LunarSolver_Tools.cpp:685-775). */
int mhd_divb_max(mhd_ctx *ctx, int case_id, double *out);
int mhd_synchronize(mhd_ctx *ctx);
/* Re-sync U and B ghosts from the current interior and recompute the derived
   fields of every case (after mhd_upload of state, as the reference does
   after a restart read). */
int mhd_refresh(mhd_ctx *ctx);
/* MPCNSRST restart of a case's U_H, B_xi, B_eta, B_zeta
   (IOModule::WriteRestartBinFile / ReadRestartBinFile,
   src/5_io/IOModule_Bin.cpp:168-330; whitelist of
   Z4_Lunar/src/LunarSolver.cpp:38-39).  Reading accepts any ghost depth in
   the file, copies the overlap and calls mhd_refresh. */
int mhd_write_restart(mhd_ctx *ctx, int case_id, const char *path, long long step, double time);
int mhd_read_restart(mhd_ctx *ctx, int case_id, const char *path, long long *step, double *time);
/* Tecplot binary (#!TDV112) of a case in the reference's Mixed mode
   (IOModule::WriteTecplotBinFile, src/5_io/IOModule_Tecplot.cpp:320-579;
   field lists of Z4_Lunar/src/LunarSolver.cpp:43-64 and
   Z5_Mercury/src/MercurySolver.cpp:36-65): nodal x, y, z, then cell
   u_H..T_H, [u_Na..T_Na], Bx..Bz, bx_ind..bz_ind, [n_Na], Jx..Jz as float32,
   one zone "Zone<b>_Fluid" per block, `time` as the zone's solution time. */
int mhd_write_tecplot(mhd_ctx *ctx, int case_id, const char *path, double time);

/* Introspection and data exchange (FieldBlock::raw_data round trip,
   include/3_field/1_Field_Block.h:47). */
int mhd_case_info(mhd_ctx *ctx, int case_id, int *nblocks, int *ni, int *nj, int *nk, int *ghost);
/* cells of one block (blocks of a case may differ; mhd_case_info gives block 0) */
int mhd_block_info(mhd_ctx *ctx, int case_id, int block, int *ni, int *nj, int *nk);
int mhd_field_count(mhd_ctx *ctx, int case_id, int field, size_t *n);
int mhd_download(mhd_ctx *ctx, int case_id, int block, int field, double *aos, size_t n);
/* A device case keeps the least-squares data of Bcell_from_Bface_w, not the
   table: MHD_F_BCELL_W downloads the table formed from it (bit-identical to
   the host builder's) and cannot be uploaded (MHD_ERR_ARG); MHD_F_JCELL_W is
   held by host cases only. */
int mhd_upload(mhd_ctx *ctx, int case_id, int block, int field, const double *aos, size_t n);
/* Topology of a block: 6 faces (-i,+i,-j,+j,-k,+k) x 14 ints
   {kind (0 physical, 1 interface), bc, nb_block, nb_face, periodic,
    perm[3], sign[3], offset[3]} with the reference's IndexTransform
   convention nb[perm[a]] = sign[a]*local[a] + offset[a]
   (include/2_topology/2_MPCNS_Topology.h:22-29).  kind 2 marks a face
   several interface patches share (read from a CASE directory's inner file,
   each patch over part of the face); the other fields describe the first. */
int mhd_get_faces(mhd_ctx *ctx, int case_id, int block, int *faces14x6);
/* node coordinates incl. ghost+1 layers, comp 0..2 */
int mhd_get_nodes(mhd_ctx *ctx, int case_id, int block, int comp, double *out, size_t n);
/* case constants: gamma, cfl, inver_MA2, momentum_hall_coeff, M_H, M_ref,
   ne_floor, state_coeff, qinf[5], B_imf[3], u_inf[3], dt_floor (22 doubles) */
int mhd_get_constants(mhd_ctx *ctx, int case_id, double *out22);
/* Per-kernel timing of the last mhd_step call (ms, summed per kernel kind)
   measured with HIP events on the context's stream; names via index. */
int mhd_profile(mhd_ctx *ctx, int enable);
int mhd_profile_read(mhd_ctx *ctx, int kernel, double *ms, int *launches, const char **name);
/* device bytes of state + geometry held by the context */
int mhd_memory(mhd_ctx *ctx, double *bytes);
/* launch shape the tuner chose for a kernel index of mhd_profile_read:
   workgroup height (rows of 64 threads along j), tile order (i_inner 1:
   consecutive tiles step along i before the next row group of j) and
   geo_shared (1: the cases sharing one geometry set run each geometry tile
   back to back, so it is fetched from HBM once for all of them while it is
   in L2; 2: one workgroup loads the geometry tile into registers and updates
   every case sharing it in turn -- k_derived only, i_inner 0).  The
   first-order fluid kernels run column tiles of fixed height; for them
   rows 1, 2 and 4 select columns of 8, 4 and 2 planes and i_inner has no
   effect. */
int mhd_tuning(mhd_ctx *ctx, int kernel, int *rows, int *i_inner, int *geo_shared);
/* pin a kernel's launch shape (rows 1, 2 or 4; i_inner 0/1; geo_shared 0, 1
   or 2 as above) */
int mhd_set_tuning(mhd_ctx *ctx, int kernel, int rows, int i_inner, int geo_shared);
/* Re-run the launch-shape tuner on the cases resident now (the tuner of
   mhd_options.autotune runs once, at mhd_finalize).  No counterpart in the
   reference (it has no tuner); state is left as it was. */
int mhd_autotune(mhd_ctx *ctx);

/* ---- host-only case builder (no device needed) ------------------------
   The same grid / topology / geometry / weights / initial-state builders the
   device context uses, run on the CPU and read back in the AoS layout. */
int mhd_host_case_ot(const mhd_ot_desc *d, int ghost, mhd_host_case **out);
int mhd_host_case_lunar(const mhd_wind_desc *d, int ghost, mhd_host_case **out);
int mhd_host_case_mercury(const mhd_wind_desc *d, const mhd_dipole_desc *dp, int ghost, mhd_host_case **out);
/* CASE directory of the case's grid, readable by the reference
   (Grid::Read_* above) and by mhd_host_case_from_dir / mhd_add_case_from_dir. */
int mhd_host_case_write_dir(mhd_host_case *h, const char *dir);
int mhd_host_case_from_dir(const char *dir, const mhd_wind_desc *d, int ghost, mhd_host_case **out);
/* mhd_case_begin .. mhd_case_end on the host (no device): the host case can
   be adopted by a context later (mhd_adopt_host_case) or kept on the host;
   mhd_host_last_error gives the reason of a failed call; with a NULL case,
   the reason the calling thread's last host-case constructor
   (mhd_host_case_lunar / _mercury / _ot / _from_dir) failed. */
int mhd_host_case_begin(int nblocks, int ghost, const double *consts22, mhd_host_case **out);
int mhd_host_add_block(mhd_host_case *h, int ib, int ni, int nj, int nk, const double *x, const double *y,
                       const double *z, size_t nnode, const double *const *geometry);
int mhd_host_set_topology(mhd_host_case *h, int ib, const int *faces14x6);
int mhd_host_set_face_parts(mhd_host_case *h, int ib, int face, int nparts, const int *parts16);
int mhd_host_case_end(mhd_host_case *h);
const char *mhd_host_last_error(mhd_host_case *h);
/* MPCNSRST restart of the host arrays (read also re-syncs U and B ghosts). */
int mhd_host_write_restart(mhd_host_case *h, const char *path, long long step, double time);
int mhd_host_read_restart(mhd_host_case *h, const char *path, long long *step, double *time);
/* Tecplot binary of the host arrays, as mhd_write_tecplot. */
int mhd_host_write_tecplot(mhd_host_case *h, const char *path, double time);
int mhd_host_case_destroy(mhd_host_case *h);
int mhd_host_case_info(mhd_host_case *h, int *nblocks, int *ni, int *nj, int *nk, int *ghost);
int mhd_host_block_info(mhd_host_case *h, int block, int *ni, int *nj, int *nk);
int mhd_host_field_count(mhd_host_case *h, int field, size_t *n);
int mhd_host_download(mhd_host_case *h, int block, int field, double *aos, size_t n);
int mhd_host_upload(mhd_host_case *h, int block, int field, const double *aos, size_t n);
int mhd_host_faces(mhd_host_case *h, int block, int *faces14x6);
/* The patches of a face several interface patches share (mhd_host_faces
   kind 2; a CASE directory's inner file), 16 ints each: {lo[2], hi[2] (node
   range along the face's two tangential axes in axis order, inclusive),
   nb_block, nb_face, periodic, perm[3], sign[3], offset[3]}; *nparts = 0 for
   other faces.  No counterpart in the reference, which keeps its inner
   patches in TOPO::Topology (include/2_topology/2_MPCNS_Topology.h:154);
   the tests hand them to the oracle. */
int mhd_host_face_parts(mhd_host_case *h, int block, int face, int *out16xn, int cap, int *nparts);
int mhd_host_nodes(mhd_host_case *h, int block, int comp, double *out, size_t n);
int mhd_host_constants(mhd_host_case *h, double *out22);
/* two-fluid constants: model, ambipolar, M_Na, state_coeff_Na, qinfs[5],
   momentum_induce_coeff, a1_Na, a4, a5, a6, T_ref, Tn0, pe_coef, hall,
   hall_coef, hall_rmin, hall_rmax, local_arti, loc_eta_max, loc_center[3],
   loc_r_decay, loc_r_cutoff, resist, resist_nsub, inver_Rem (31 doubles;
   Z5_Mercury/src/physics/MercurySolver_Tools.cpp:3-46, 9999setup/3_MHD_Control.txt) */
int mhd_host_constants_two_fluid(mhd_host_case *h, double *out31);
/* apply one halo / boundary op list (index as in the library's sync table:
   0 Eface BC, 1 Eface halo, 2 Eedge BC, 3-5 Ucell BC0/BC1/halo, 6-7 Bface
   BC/halo, 8 J_cell BC, 9 corner velocities, 10-11 B_cell BC/halo and
   12-13 J_cell ghosts/halo of two-fluid cases with the Hall EMF) to the host
   arrays */
int mhd_host_apply_sync(mhd_host_case *h, int sync);
int mhd_host_op_count(mhd_host_case *h, int sync, size_t *n);
/* 1 in *independent when no element of op list `sync` is written twice and
   none it writes is read by the same list (the device applies a list's
   records in any order and loads an element's sources before its stores;
   no counterpart in the reference, whose Sync loops run in order) */
int mhd_host_ops_independent(mhd_host_case *h, int sync, int *independent);
/* 1 in *ok when the Eface sync of case h is what the one-pass face/edge EMF
   of a whole step assumes: its shell enumeration lists each shell point
   (the first layer inside every interface side) once, every Eface halo
   copy reads a shell face and writes a ghost face the kernel loads, and the
   wall handler zeroes exactly the faces the kernel takes as zero (no
   counterpart in the reference: a check of this library's fused form of
   Sync("Eface"), Z4_Lunar/src/LunarSolver_Driver.cpp:21-92) */
int mhd_host_emf_shell_check(mhd_host_case *h, int *ok);

#ifdef __cplusplus
}
#endif
#endif
