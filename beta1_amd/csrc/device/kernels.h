// Launch wrappers of the step kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "../host/setup.h"
#include "../layout.h"

namespace mhdk
{
// Largest block extents among the blocks of one launch (grid sizing).
struct LaunchShape
{
    int nblocks;
    int max_ni, max_nj, max_nk;
};

void launch_dt(const LaunchShape &s, const MhdBlockDev *blks, double *st, const double *geo, const MhdCaseConst *cst,
               unsigned long long *dtbits, int ucur, hipStream_t q);
void launch_dt_finalize(unsigned long long *dtbits, double *dt, int *err, const MhdCaseConst *cst, int ncase,
                        const int *active, hipStream_t q);
void launch_ucontra(const LaunchShape &s, const MhdBlockDev *blks, double *st, const double *geo,
                    const MhdCaseConst *cst, int ucur, hipStream_t q);
void launch_face_emf(const LaunchShape &s, const MhdBlockDev *blks, double *st, int mixed, hipStream_t q);
void launch_edge_emf(const LaunchShape &s, const MhdBlockDev *blks, double *st, hipStream_t q);
void launch_curl_update(const LaunchShape &s, const MhdBlockDev *blks, double *st, const double *dts, hipStream_t q);
void launch_fluid(const LaunchShape &s, const MhdBlockDev *blks, double *st, const double *geo,
                  const MhdCaseConst *cst, const double *dts, int ucur, int mixed, int recon, hipStream_t q);
void launch_derived(const LaunchShape &s, const MhdBlockDev *blks, double *st, const double *geo,
                    const MhdCaseConst *cst, unsigned long long *dtbits, int ucur, hipStream_t q);
void launch_ops(const mpcns::MhdOp *ops, int n, double *st, const MhdCaseConst *cst, int ushift, hipStream_t q);
void launch_divb(const LaunchShape &s, const MhdBlockDev *blks, const double *st, const double *geo,
                 unsigned long long *out, hipStream_t q);
void launch_fill_u64(unsigned long long *p, int n, unsigned long long v, hipStream_t q);
} // namespace mhdk
