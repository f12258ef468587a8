// A case built on the CPU and held behind the opaque mhd_host_case handle of
// the C-ABI; the device context adopts its contents without rebuilding them.
#pragma once
#include <string>

#include "../../../include/beta1_mhd.h"
#include "setup.h"

struct mhd_host_case
{
    mpcns::CaseSetup cs;
    std::string err;
};

namespace mpcns
{
// C descriptors -> builder parameters (m_amu <= 0 means protons).
WindParams wind_params(const mhd_wind_desc &d);
DipoleParams dipole_params(const mhd_dipole_desc &d);
bool wind_desc_ok(const mhd_wind_desc *d);
} // namespace mpcns
