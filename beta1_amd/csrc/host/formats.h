// The reference's CASE grid directory and MPCNSRST restart formats.
#pragma once
#include <string>
#include <vector>

#include "setup.h"

namespace mpcns
{
// CASE/geometry/grid0000.grd, boundary_condition/{boundary0000.txt,
// boundary_name.fvbnd, inner0000.txt}: whole-face patches, Farfield and
// Solid_Surface conditions.
void write_case_dir(const Grid &grd, const std::string &dir, const std::string &block_name = "Fluid");
// Reads the same files back into a grid (node coordinates, boundary
// conditions, interface links with their index transforms); ghost nodes are
// not filled.
Grid read_case_dir(const std::string &dir, int g);

struct RestartBlock
{
    int lo[3], hi[3];
    std::vector<double> v; // [i][j][k][m] over [lo, hi); empty if inactive
};
struct RestartField
{
    std::string name;
    int loc_code = 0; // StaggerLocation code of the reference's restart files
    int ncomp = 0, nghost = 0;
    std::vector<RestartBlock> blocks;
};
void write_restart_file(const std::string &path, long long step, double time, const std::vector<RestartField> &fields);
std::vector<RestartField> read_restart_file(const std::string &path, long long &step, double &time);
int restart_loc_code(int loc);
// Copy the part of `src` inside the box [lo, hi) into dst (AoS over that box).
void copy_overlap(const RestartBlock &src, int ncomp, const int *lo, const int *hi, double *dst);
} // namespace mpcns
