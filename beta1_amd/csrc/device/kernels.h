// Launch wrappers of the step kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "../host/setup.h"
#include "../layout.h"

namespace mhdk
{
// One launch: the tile map of its range plus the workgroup height (rows of
// 64-lane wavefronts along j).
struct Launch
{
    MhdTileMap tm;
    int jt;
};

// `mixed`: geometry stored as float (mixed-precision contexts), else double.
void launch_dt(const Launch &L, const MhdCaseConst *cst, unsigned long long *dtbits, int ucur, int mixed,
               hipStream_t q);
void launch_dt_finalize(unsigned long long *dtbits, double *dt, int *err, const MhdCaseConst *cst, int ncase,
                        const int *active, hipStream_t q);
void launch_ucontra(const Launch &L, const MhdCaseConst *cst, int ucur, int mixed, hipStream_t q);
void launch_face_emf(const Launch &L, int mixed, hipStream_t q);
void launch_edge_emf(const Launch &L, hipStream_t q);
// Face EMF candidates of each block's shell points (256 per workgroup; the
// tile map's entries count ceil(shell_count / 256) tiles each), then the
// face and edge EMF in one pass over plane-chunk tiles of [0, N]^3 (after
// the halo copy of the shell's candidates); every extent >= 2.
void launch_face_shell(const Launch &L, hipStream_t q);
void launch_emf(const Launch &L, hipStream_t q);
// tiles per plane of the padded-row tiles (62 owned points per wavefront,
// rows wavefronts per tile) of a launch range ext_j x ext_k points a plane
int pad_tiles_per_plane(int ext_j, int ext_k, int rows);
void launch_curl_update(const Launch &L, const double *dts, hipStream_t q);
void launch_fluid(const Launch &L, const MhdCaseConst *cst, const double *dts, int ucur, int mixed, int recon,
                  hipStream_t q);
// First-order fluid update in column tiles (C.tm.ci planes per workgroup of
// 64 x col_tile_rows() threads); bit-identical to launch_fluid with recon 0.
void launch_fluid_col(const Launch &C, const MhdCaseConst *cst, const double *dts, int ucur, int mixed, hipStream_t q);
int col_tile_rows();
// J edges (face range), then J_cell from them (interior cells).
void launch_jcell(const Launch &L, int mixed, hipStream_t q);
void launch_derived(const Launch &L, const MhdCaseConst *cst, unsigned long long *dtbits, int ucur, int mixed,
                    hipStream_t q);
// two-fluid (Mercury) step
void launch_derived2(const Launch &L, const MhdCaseConst *cst, unsigned long long *dtbits, int ucur, int mixed,
                     hipStream_t q);
void launch_fluid2_col(const Launch &C, const MhdCaseConst *cst, const double *dts, int ucur, int mixed,
                       hipStream_t q);
void launch_fluid2(const Launch &L, const MhdCaseConst *cst, const double *dts, int ucur, int mixed, int recon,
                   hipStream_t q);
void launch_emf_extra(const Launch &L, const MhdCaseConst *cst, int ucur, int mixed, hipStream_t q);
void launch_eres(const Launch &L, const MhdCaseConst *cst, hipStream_t q);
void launch_dt_res(double *dtres, const double *dt, const MhdCaseConst *cst, int ncase, int isub, hipStream_t q);
// Up to OPSEG_MAX segments of op records in one launch: segment s holds n
// records on 2^lg lanes each (lg 0 or 6), U-slot ops shifted to the other U
// buffer when ushift is set.  The segments' destinations are disjoint and no
// segment reads what another writes.
struct OpSeg
{
    const mpcns::MhdOp *ops;
    int n, lg, ushift;
};
constexpr int OPSEG_MAX = 8;
struct OpLaunch
{
    OpSeg seg[OPSEG_MAX];
    int nseg;
};
void launch_ops(const OpLaunch &L, const MhdCaseConst *cst, hipStream_t q);
void launch_divb(const Launch &L, unsigned long long *out, int mixed, hipStream_t q);
void launch_fill_u64(unsigned long long *p, int n, unsigned long long v, hipStream_t q);
// 17 order keys per case (min/max diagnostics) at out + 17 * case id; the
// caller sets them to the neutral keys first
void launch_diag(const Launch &L, const MhdCaseConst *cst, unsigned long long *out, int ucur, hipStream_t q);
// Diagnostics of the cases sel.ids[0 .. sel.n) into out + 17 * q (neutral
// keys written first, by a kernel): the reduction of launch_diag.
constexpr int DIAG_SEL_MAX = 64;
struct DiagSel
{
    int n; // < 0: every case, at its case id
    int ids[DIAG_SEL_MAX];
};
void launch_diag_sel(const Launch &L, const MhdCaseConst *cst, unsigned long long *out, int ucur, const DiagSel &sel,
                     hipStream_t q);
// Device address of the static 17 * DIAG_SEL_MAX-key scratch array on
// `device`, which must be current; looked up once per device.
hipError_t diag_scratch(int device, unsigned long long **ptr);
} // namespace mhdk
