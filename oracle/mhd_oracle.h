/* CPU FP64 oracle of the explicit MHD step -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's LunarSolver::StepOnce
 * (Z4_Lunar/src/LunarSolver_Driver.cpp:21-92) on field blocks stored in the
 * reference's own AoS layout [i][j][k][m] (include/3_field/Field_Array.h:177)
 * with the extents of FieldBlock::compute_extent
 * (src/3_field/1_Field_Block.cpp:51-131).  It is used by the tests, by the
 * smoke check and by the bench's CPU-baseline leg, as the reference result to
 * compare against (test infrastructure); nothing in
 * the product links or calls it.
 */
#ifndef MHD_ORACLE_H
#define MHD_ORACLE_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct
{
    int kind; /* 0 physical, 1 interface */
    int bc;   /* 1 farfield, 2 solid (physical faces) */
    int nb_block, nb_face, periodic;
    int perm[3], sign[3], offset[3]; /* nb[perm[a]] = sign[a]*local[a] + offset[a] (node space) */
} or_face;

/* one of several interface patches sharing a block face (the reference's
   inner patches may cover part of a face, src/1_grid/1_MPCNS_Grid.cpp:392-505):
   its node range along the face's two tangential axes (axis order,
   inclusive) and its link (kind 1) */
typedef struct
{
    int lo[2], hi[2];
    or_face link;
} or_face_part;

typedef struct
{
    int nblocks;
    int g;     /* ghost depth of every field */
    int recon; /* 0 first order (reference), 1 minmod */
    double gamma, cfl, inver_MA2, momentum_hall_coeff, M_H, M_ref, ne_floor, state_coeff;
    double qinf[5], B_imf[3], u_inf[3];
    double dt_floor;
    /* two-fluid H+ / Na+ step of MercurySolver (model 1); constants of
       Z5_Mercury/src/physics/MercurySolver_Tools.cpp:3-46 and
       MercurySolver_FluidRHS.cpp:210-225 */
    int model;     /* 0 single fluid (Lunar), 1 two-fluid (Mercury) */
    int ambipolar; /* model 1: electron-pressure terms in the momentum source and the edge EMF */
    double M_Na, state_coeff_Na, qinfs[5];
    double induce, a1_Na, a4, a5, a6, T_ref, Tn0, pe_coef;
    /* explicit Hall EMF (Z5_Mercury/src/physics/MercurySolver_MagExplicit_Hall.cpp)
       and local artificial resistivity (MercurySolver_MagExplicit_Resis.cpp:67-187) */
    int hall, local_arti;
    double hall_coef, hall_rmin, hall_rmax;
    double loc_eta_max, loc_center[3], loc_r_decay, loc_r_cutoff;
    /* interior resistivity by explicit subcycles (ResistiveDiffusionSubcycles_,
       Z5_Mercury/src/physics/MercurySolver_MagneticRHS.cpp:79-141) */
    int resist, resist_nsub;
    double inver_Rem;
} or_desc;

enum
{
    OF_UH = 0, OF_PV, OF_UPLUS, OF_BXI, OF_BET, OF_BZE, OF_BADDXI, OF_BADDET, OF_BADDZE,
    OF_BCELL, OF_BIND, OF_JXI, OF_JET, OF_JZE, OF_JCELL, OF_EXI, OF_EET, OF_EZE,
    OF_EFXI, OF_EFET, OF_EFZE, OF_FXI, OF_FET, OF_FZE, OF_RHSH, OF_RHSBXI, OF_RHSBET, OF_RHSBZE,
    OF_DIVB, OF_JAC, OF_JDXI, OF_JDET, OF_JDZE, OF_BETAXI, OF_BETAET, OF_BETAZE,
    OF_ALXI, OF_ALET, OF_ALZE, OF_WB, OF_WJ,
    OF_UNA, OF_PVNA, OF_RHSNA, OF_PHOTO, OF_DRAG, OF_EHXI, OF_EHET, OF_EHZE,
    OF_ERXI, OF_ERET, OF_ERZE, OF_RBRXI, OF_RBRET, OF_RBRZE, OF_COUNT
};

typedef struct or_ctx or_ctx;

or_ctx *or_create(const or_desc *d);
void or_destroy(or_ctx *c);
/* faces: 6 entries (-i,+i,-j,+j,-k,+k) */
int or_set_block(or_ctx *c, int b, int ni, int nj, int nk, const or_face *faces);
/* a face several interface patches share (after or_set_block): a ghost
   crosses it through the patch covering its position in the face (clamped
   onto the face; on the line between two patches, the first listed) */
int or_set_face_parts(or_ctx *c, int b, int f, int n, const or_face_part *parts);
/* pointer to the AoS storage of a field (allocated by or_set_block) */
double *or_field(or_ctx *c, int b, int fid, long *count);
/* node coordinates, box [-(g+1), n+g+1]^3, k fastest */
double *or_nodes(or_ctx *c, int b, int comp, long *count);
/* UpdateDerivedFields_ (Z4_Lunar/src/physics/LunarSolver_Tools.cpp:143-161) */
int or_update_derived(or_ctx *c);
/* Compute_Timestep (Z4_Lunar/src/physics/LunarSolver_Timestep.cpp:8-140) */
int or_compute_dt(or_ctx *c, double *dt);
/* one StepOnce; returns 0 or a negative error (dt underflow / NaN) */
int or_step(or_ctx *c, double *dt_used);
/* max |div B| over interior cells (calc_divB, LunarSolver_Tools.cpp:685-775) */
double or_divb_max(or_ctx *c);
/* PrintMinMaxDiagnostics_: minima [0..7] and maxima [8..15] of rho_H, p_H,
   rho_Na, p_Na, Bx, By, Bz, |B|^2/2 over interior cells, [16] max |div B| */
void or_diagnostics(or_ctx *c, double *out17);
/* 1: the single-block fully periodic halo uses an index wrap instead of
   the generic interface resolution */
void or_use_periodic_wrap(or_ctx *c, int on);
/* OpenMP threads for the per-cell loops (default 1); results are identical
   for every thread count */
void or_set_threads(or_ctx *c, int n);
/* run only the named sync group on the current arrays (tests) */
int or_sync(or_ctx *c, const char *group);
/* the interior-resistivity EMF of the current J_cell into Eres (tests) */
int or_resistive_emf(or_ctx *c);
/* one stage of or_step ("dt", "fluid", "face_emf", "edge_emf", "curl",
   "update", "derived"; syncs through or_sync) on the current arrays (tests) */
int or_stage(or_ctx *c, const char *stage);

#ifdef __cplusplus
}
#endif
#endif
