// A case built on the CPU and held behind the opaque mhd_host_case handle of
// the C-ABI; the device context adopts its contents without rebuilding them.
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "../../../include/beta1_mhd.h"
#include "setup.h"

struct mhd_case_draft; // blocks and topology given so far (mhd_host_case_begin .. _end)

struct mhd_host_case
{
    mpcns::CaseSetup cs;
    std::string err;
    std::shared_ptr<mhd_case_draft> draft; // set between mhd_host_case_begin and mhd_host_case_end
};

namespace mpcns
{
// C descriptors -> builder parameters (m_amu <= 0 means protons).
WindParams wind_params(const mhd_wind_desc &d);
DipoleParams dipole_params(const mhd_dipole_desc &d);
bool wind_desc_ok(const mhd_wind_desc *d);

// Whitelisted restart fields of one case through an AoS get/put interface
// (host arrays or device memory); get/put return an mhd_status.
struct FieldIO
{
    virtual ~FieldIO() = default;
    virtual int nblocks() = 0;
    virtual const MhdBlockDev &block(int b) = 0;
    virtual int get(int b, int field, std::vector<double> &aos) = 0;
    virtual int put(int b, int field, const std::vector<double> &aos) = 0;
};
int restart_write(FieldIO &io, const char *path, long long step, double time);
int restart_read(FieldIO &io, const char *path, long long *step, double *time);
// Tecplot binary output in the reference's Mixed mode; throws on I/O errors.
int tecplot_write(FieldIO &io, const Grid &grd, const MhdCaseConst &cst, double L_ref, const char *path,
                  double time);
} // namespace mpcns
