// The reference's on-disk formats for this step: the CASE grid directory
// (Plot3D-binary grid, physical-boundary ranges, boundary names, inner
// patches) and the MPCNSRST restart file.  Written and read here so a case
// can come from a reference CASE directory and a state can be exchanged with
// the reference solver.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "formats.h"

namespace mpcns
{
namespace
{
const char *kBcName[3] = {"No_boundary_Conditions", "Farfield", "Solid_Surface"};

void mkdirs(const std::string &path)
{
    std::string cur;
    std::stringstream ss(path);
    std::string part;
    if (!path.empty() && path[0] == '/')
        cur = "/";
    while (std::getline(ss, part, '/'))
    {
        if (part.empty())
            continue;
        cur += part + "/";
        ::mkdir(cur.c_str(), 0755);
    }
}

template <class T> void put(std::ofstream &o, T v) { o.write(reinterpret_cast<const char *>(&v), sizeof v); }
template <class T> T get(std::ifstream &in)
{
    T v{};
    in.read(reinterpret_cast<char *>(&v), sizeof v);
    if (!in)
        throw std::runtime_error("unexpected end of file");
    return v;
}

// Block face (2*axis+side) of a node range with one constant axis.
int face_of_range(const int *lo, const int *hi, const Block &B)
{
    for (int a = 0; a < 3; ++a)
        if (lo[a] == hi[a])
        {
            if (lo[a] == 0)
                return 2 * a;
            if (lo[a] == B.n(a))
                return 2 * a + 1;
        }
    return -1;
}

// The reference's inner-file coding (Inner_Boundary::Pre_process,
// include/1_grid/Grid_Boundary.h:244): |value| - 1 keeps the sign; the
// in-plane axis written negative pairs with the target's negative axis.
int minus1(int p) { return p > 0 ? p - 1 : p + 1; }

void transform_codes(const int *sub, const int *sup, int *code)
{
    // 1 marks the normal (constant) axis, 2 the negative in-plane axis, 0 the
    // other one (three-dimensional grids)
    if (sup[2] == sub[2])
    {
        code[2] = 1;
        if (sup[1] < 0 || sub[1] < 0)
            code[1] = 2, code[0] = 0;
        else
            code[1] = 0, code[0] = 2;
    }
    else if (sup[2] < 0 || sub[2] < 0)
    {
        code[2] = 2;
        if (sup[1] == sub[1])
            code[1] = 1, code[0] = 0;
        else
            code[1] = 0, code[0] = 1;
    }
    else
    {
        code[2] = 0;
        if (sup[1] < 0 || sub[1] < 0)
            code[1] = 2, code[0] = 1;
        else
            code[1] = 1, code[0] = 2;
    }
}
} // namespace

// ---------------------------------------------------------------- writing
void write_case_dir(const Grid &grd, const std::string &dir, const std::string &block_name)
{
    const std::string geo = dir + "/geometry", bcd = geo + "/boundary_condition";
    mkdirs(bcd);
    const int nb = (int)grd.blocks.size();
    {
        // Grid::Read_Grid_Binary (src/1_grid/1_MPCNS_Grid.cpp:123-221)
        std::ofstream o(geo + "/grid0000.grd", std::ios::binary);
        if (!o)
            throw std::runtime_error("cannot write " + geo + "/grid0000.grd");
        put<int32_t>(o, nb);
        for (int b = 0; b < nb; ++b)
        {
            put<int32_t>(o, (int32_t)block_name.size() + 1); // the reader takes a C string
            o.write(block_name.c_str(), (std::streamsize)block_name.size() + 1);
        }
        for (const Block &B : grd.blocks)
        {
            put<int32_t>(o, B.ni + 1);
            put<int32_t>(o, B.nj + 1);
            put<int32_t>(o, B.nk + 1);
        }
        for (const Block &B : grd.blocks)
            for (const std::vector<double> *X : {&B.x, &B.y, &B.z})
                for (int k = 0; k <= B.nk; ++k)
                    for (int j = 0; j <= B.nj; ++j)
                        for (int i = 0; i <= B.ni; ++i)
                            put<double>(o, (*X)[B.node_off(i, j, k)]);
    }
    {
        // Grid::Read_Phy_Boundary (src/1_grid/1_MPCNS_Grid.cpp:224-290):
        // 0-based node ranges and a type index into boundary_name.fvbnd
        std::ofstream o(bcd + "/boundary0000.txt");
        for (const Block &B : grd.blocks)
        {
            int n = 0;
            for (int f = 0; f < 6; ++f)
                n += B.face[f].kind == 0;
            o << n << "\n";
            for (int f = 0; f < 6; ++f)
            {
                if (B.face[f].kind != 0)
                    continue;
                const int ax = f / 2;
                int lo[3] = {0, 0, 0}, hi[3] = {B.ni, B.nj, B.nk};
                lo[ax] = hi[ax] = (f & 1) ? B.n(ax) : 0;
                const int type = B.face[f].bc == BC_NONE ? 0 : B.face[f].bc + 1; // names from index 2
                o << lo[0] << " " << hi[0] << " " << lo[1] << " " << hi[1] << " " << lo[2] << " " << hi[2] << " "
                  << type << "\n";
            }
        }
    }
    {
        // Grid::Read_Phy_Name (src/1_grid/1_MPCNS_Grid.cpp:292-335)
        std::ofstream o(bcd + "/boundary_name.fvbnd");
        o << "FVBND 1 3\n" << kBcName[0] << "\n" << kBcName[1] << "\n" << kBcName[2] << "\nBOUNDARIES\n";
    }
    {
        // Grid::Read_Inner_Boundary (src/1_grid/1_MPCNS_Grid.cpp:392-505):
        // 1-based signed node ranges; the in-plane axis written negative on
        // both sides is the pair the target's index frame maps it to
        std::ofstream o(bcd + "/inner0000.txt");
        std::vector<int> cnt(nb, 0);
        int total = 0;
        for (int b = 0; b < nb; ++b)
            for (int f = 0; f < 6; ++f)
                if (grd.blocks[b].face[f].kind == 1)
                {
                    const int np = std::max(1, (int)grd.blocks[b].face[f].parts.size());
                    cnt[b] += np, total += np;
                }
        o << total << "\n";
        for (int b = 0; b < nb; ++b)
            o << cnt[b] << (b + 1 < nb ? " " : "\n");
        o << "imin imax jmin jmax kmin kmax block name   (this side, then the target side)\n";
        for (int b = 0; b < nb; ++b)
        {
            const Block &B = grd.blocks[b];
            for (int f = 0; f < 6; ++f)
            {
                const FaceLink &F = B.face[f];
                if (F.kind != 1)
                    continue;
                const int ax = f / 2, t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
                std::vector<FacePart> parts = F.parts;
                if (parts.empty())
                {
                    FacePart W;
                    W.hi[0] = B.n(t1), W.hi[1] = B.n(t2);
                    W.nb_block = F.nb_block, W.nb_face = F.nb_face, W.periodic = F.periodic, W.trans = F.trans;
                    parts.push_back(W);
                }
                for (const FacePart &L : parts)
                {
                    const IndexTransform &T = L.trans;
                    int sub[3] = {0, 0, 0}, sup[3] = {B.ni, B.nj, B.nk};
                    sub[ax] = sup[ax] = (f & 1) ? B.n(ax) : 0;
                    sub[t1] = L.lo[0], sup[t1] = L.hi[0], sub[t2] = L.lo[1], sup[t2] = L.hi[1];
                    int tsub[3], tsup[3];
                    for (int a = 0; a < 3; ++a)
                    {
                        tsub[T.perm[a]] = T.sign[a] * sub[a] + T.offset[a];
                        tsup[T.perm[a]] = T.sign[a] * sup[a] + T.offset[a];
                    }
                    const int mark = ax == 0 ? 1 : 0; // first in-plane axis
                    auto code = [](int v, bool neg) { return neg ? -(v + 1) : v + 1; };
                    const int src = L.periodic ? -(b + 1) : b + 1, tar = L.periodic ? -(L.nb_block + 1) : L.nb_block + 1;
                    for (int a = 0; a < 3; ++a)
                        o << code(sub[a], a == mark) << " " << code(sup[a], a == mark) << " ";
                    o << src << " " << block_name << "   ";
                    for (int a = 0; a < 3; ++a)
                        o << code(tsub[a], a == T.perm[mark]) << " " << code(tsup[a], a == T.perm[mark]) << " ";
                    o << tar << " " << block_name << "\n";
                }
            }
        }
    }
}

// ---------------------------------------------------------------- reading
Grid read_case_dir(const std::string &dir, int g)
{
    const std::string geo = dir + "/geometry", bcd = geo + "/boundary_condition";
    Grid grd;
    grd.g = g;
    {
        std::ifstream in(geo + "/grid0000.grd", std::ios::binary);
        if (!in)
            throw std::runtime_error("cannot read " + geo + "/grid0000.grd");
        const int nb = get<int32_t>(in);
        if (nb <= 0 || nb > 100000)
            throw std::runtime_error("grid: bad block count");
        grd.blocks.resize(nb);
        for (int b = 0; b < nb; ++b)
        {
            const int len = get<int32_t>(in);
            if (len < 0 || len > 4096)
                throw std::runtime_error("grid: bad block name length");
            std::string name((size_t)len, '\0');
            in.read(&name[0], len);
        }
        std::vector<int> m(3 * nb);
        for (int &v : m)
            v = get<int32_t>(in);
        for (int b = 0; b < nb; ++b)
        {
            Block &B = grd.blocks[b];
            if (m[3 * b] < 2 || m[3 * b + 1] < 2 || m[3 * b + 2] < 2)
                throw std::runtime_error("grid: three-dimensional blocks only");
            B.ni = m[3 * b] - 1, B.nj = m[3 * b + 1] - 1, B.nk = m[3 * b + 2] - 1;
            B.gn = g + 1;
            const int64_t n = B.node_count();
            B.x.assign(n, 0.0), B.y.assign(n, 0.0), B.z.assign(n, 0.0);
            for (std::vector<double> *X : {&B.x, &B.y, &B.z})
                for (int k = 0; k <= B.nk; ++k)
                    for (int j = 0; j <= B.nj; ++j)
                        for (int i = 0; i <= B.ni; ++i)
                            (*X)[B.node_off(i, j, k)] = get<double>(in);
        }
    }
    std::vector<std::string> names;
    {
        std::ifstream in(bcd + "/boundary_name.fvbnd");
        if (!in)
            throw std::runtime_error("cannot read boundary_name.fvbnd");
        std::string line, w;
        std::getline(in, line); // FVBND 1 3
        std::getline(in, line);
        names.push_back(line); // index 0
        names.push_back("");   // index 1
        while (in >> w && w != "BOUNDARIES")
            names.push_back(w);
    }
    auto bc_of = [&](int type) {
        if (type < 0 || type >= (int)names.size())
            throw std::runtime_error("boundary: type index out of range");
        const std::string &n = names[type];
        if (n == "Farfield")
            return (int)BC_FARFIELD;
        if (n == "Solid_Surface")
            return (int)BC_SOLID;
        throw std::runtime_error("boundary: unsupported condition '" + n + "'");
    };
    const int nb = (int)grd.blocks.size();
    for (auto &B : grd.blocks)
        for (auto &f : B.face)
            f = FaceLink{};
    // the patches of each block face, physical (kind 0) or interface (1),
    // with their node ranges along the face's tangential axes
    struct Patch
    {
        int kind, bc;
        FacePart part;
    };
    std::vector<std::array<std::vector<Patch>, 6>> patches(nb);
    auto tangential = [](int f, const int lo[3], const int hi[3], FacePart &P) {
        const int ax = f / 2, t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
        P.lo[0] = lo[t1], P.hi[0] = hi[t1], P.lo[1] = lo[t2], P.hi[1] = hi[t2];
    };
    {
        std::ifstream in(bcd + "/boundary0000.txt");
        if (!in)
            throw std::runtime_error("cannot read boundary0000.txt");
        for (int b = 0; b < nb; ++b)
        {
            int n;
            in >> n;
            for (int l = 0; l < n; ++l)
            {
                int r[6], type;
                in >> r[0] >> r[1] >> r[2] >> r[3] >> r[4] >> r[5] >> type;
                if (!in)
                    throw std::runtime_error("boundary: truncated file");
                const int lo[3] = {std::min(r[0], r[1]), std::min(r[2], r[3]), std::min(r[4], r[5])};
                const int hi[3] = {std::max(r[0], r[1]), std::max(r[2], r[3]), std::max(r[4], r[5])};
                const int f = face_of_range(lo, hi, grd.blocks[b]);
                if (f < 0)
                    throw std::runtime_error("boundary: range is not a block face");
                Patch P{0, bc_of(type), {}};
                tangential(f, lo, hi, P.part);
                patches[b][f].push_back(P);
            }
        }
    }
    {
        std::ifstream in(bcd + "/inner0000.txt");
        if (!in)
            throw std::runtime_error("cannot read inner0000.txt");
        std::string line;
        std::getline(in, line);
        std::vector<int> cnt(nb);
        for (int &c : cnt)
            in >> c;
        std::getline(in, line);
        std::getline(in, line); // comment
        for (int b = 0; b < nb; ++b)
            for (int l = 0; l < cnt[b]; ++l)
            {
                int s[6], t[6], sb, tb;
                std::string sname, tname;
                in >> s[0] >> s[1] >> s[2] >> s[3] >> s[4] >> s[5] >> sb >> sname;
                in >> t[0] >> t[1] >> t[2] >> t[3] >> t[4] >> t[5] >> tb >> tname;
                if (!in)
                    throw std::runtime_error("inner: truncated file");
                int sub[3], sup[3], tsub[3], tsup[3];
                for (int a = 0; a < 3; ++a)
                {
                    sub[a] = minus1(s[2 * a]), sup[a] = minus1(s[2 * a + 1]);
                    tsub[a] = minus1(t[2 * a]), tsup[a] = minus1(t[2 * a + 1]);
                }
                const int me = std::abs(sb) - 1, nbk = std::abs(tb) - 1;
                if (me != b || nbk < 0 || nbk >= nb)
                    throw std::runtime_error("inner: block numbers out of range");
                // IndexTransform as TOPO::build_topology derives it
                // (src/2_topology/2_MPCNS_Topology.cpp:104-163)
                int code[3], tcode[3], inv[3];
                transform_codes(sub, sup, code);
                transform_codes(tsub, tsup, tcode);
                for (int d = 0; d < 3; ++d)
                    inv[tcode[d]] = d;
                IndexTransform T{};
                int alo[3], ahi[3], blo[3], bhi[3];
                for (int d = 0; d < 3; ++d)
                    T.perm[d] = inv[code[d]];
                int dir = 0, tdir = 0;
                for (int d = 0; d < 3; ++d)
                {
                    alo[d] = std::abs(sub[d]), ahi[d] = std::abs(sup[d]);
                    blo[d] = std::abs(tsub[d]), bhi[d] = std::abs(tsup[d]);
                }
                for (int d = 0; d < 3; ++d)
                {
                    if (alo[d] == ahi[d])
                        dir = alo[d] == 0 ? -(d + 1) : d + 1;
                    if (blo[d] == bhi[d])
                        tdir = blo[d] == 0 ? -(d + 1) : d + 1;
                }
                for (int d = 0; d < 3; ++d)
                {
                    const int p = T.perm[d];
                    if (alo[d] == ahi[d])
                        T.sign[d] = (dir * tdir > 0) ? -1 : 1;
                    else
                    {
                        if (std::abs(ahi[d] - alo[d]) != std::abs(bhi[p] - blo[p]))
                            throw std::runtime_error("inner: ranges of different length");
                        T.sign[d] = ((ahi[d] - alo[d]) * (bhi[p] - blo[p]) > 0) ? 1 : -1;
                    }
                    T.offset[d] = -T.sign[d] * alo[d] + blo[p];
                }
                const Block &B = grd.blocks[b];
                const int f = face_of_range(alo, ahi, B);
                const int tf = face_of_range(blo, bhi, grd.blocks[nbk]);
                if (f < 0 || tf < 0)
                    throw std::runtime_error("inner: range is not a block face");
                int lo[3], hi[3];
                for (int a = 0; a < 3; ++a)
                    lo[a] = std::min(alo[a], ahi[a]), hi[a] = std::max(alo[a], ahi[a]);
                Patch P{1, BC_NONE, {}};
                tangential(f, lo, hi, P.part);
                P.part.nb_block = nbk;
                P.part.nb_face = tf;
                P.part.periodic = sb < 0;
                P.part.trans = T;
                patches[b][f].push_back(P);
            }
    }
    // Each face: one condition, or several patches of one kind that tile it.
    // Interface patches may split a face among neighbours (FaceLink::parts);
    // physical patches of one type merge; a face mixing a physical condition
    // with interfaces, or two conditions, is refused (the boundary phases and
    // the kernels' folded boundary work hold one condition per face).
    for (int b = 0; b < nb; ++b)
        for (int f = 0; f < 6; ++f)
        {
            Block &B = grd.blocks[b];
            const std::vector<Patch> &ps = patches[b][f];
            const std::string where = "block " + std::to_string(b) + " face " + std::to_string(f);
            if (ps.empty())
                throw std::runtime_error(where + " has neither a boundary condition nor an inner patch");
            const int ax = f / 2, n1 = B.n(ax == 0 ? 1 : 0), n2 = B.n(ax == 2 ? 1 : 2);
            std::vector<int> cover((size_t)n1 * n2, 0);
            for (const Patch &P : ps)
            {
                if (P.part.lo[0] < 0 || P.part.hi[0] > n1 || P.part.lo[1] < 0 || P.part.hi[1] > n2 ||
                    P.part.lo[0] >= P.part.hi[0] || P.part.lo[1] >= P.part.hi[1])
                    throw std::runtime_error(where + ": patch range outside the face");
                for (int u = P.part.lo[0]; u < P.part.hi[0]; ++u)
                    for (int v = P.part.lo[1]; v < P.part.hi[1]; ++v)
                        cover[(size_t)u * n2 + v] += 1;
            }
            for (int c : cover)
                if (c != 1)
                    throw std::runtime_error(where + ": patches do not cover the face exactly once");
            FaceLink &L = B.face[f];
            if (ps[0].kind == 0)
            {
                for (const Patch &P : ps)
                    if (P.kind != 0 || P.bc != ps[0].bc)
                        throw std::runtime_error(where + " mixes boundary conditions or interfaces");
                L.kind = 0;
                L.bc = ps[0].bc;
                continue;
            }
            for (const Patch &P : ps)
                if (P.kind != 1 || P.part.periodic != ps[0].part.periodic)
                    throw std::runtime_error(where + " mixes interfaces with boundary conditions or periodic links "
                                                     "with plain ones");
            L.kind = 1;
            L.bc = BC_NONE;
            L.nb_block = ps[0].part.nb_block;
            L.nb_face = ps[0].part.nb_face;
            L.periodic = ps[0].part.periodic;
            L.trans = ps[0].part.trans;
            if (ps.size() > 1)
                for (const Patch &P : ps)
                    L.parts.push_back(P.part);
        }
    return grd;
}

// ---------------------------------------------------------------- restart
// IOModule::WriteRestartBinFile / ReadRestartBinFile
// (src/5_io/IOModule_Bin.cpp:168-330): magic, version 1, step, time, block
// and field counts; per field name, location code, components, ghosts, then
// per block the [lo, hi) extent, an active flag and the values [i][j][k][m].
void write_restart_file(const std::string &path, long long step, double time, const std::vector<RestartField> &fields)
{
    std::ofstream o(path, std::ios::binary);
    if (!o)
        throw std::runtime_error("cannot write " + path);
    o.write("MPCNSRST", 8);
    put<int32_t>(o, 1);
    put<int32_t>(o, (int32_t)step);
    put<double>(o, time);
    const int nb = fields.empty() ? 0 : (int)fields[0].blocks.size();
    put<int32_t>(o, nb);
    put<int32_t>(o, (int32_t)fields.size());
    for (const RestartField &F : fields)
    {
        put<int32_t>(o, (int32_t)F.name.size());
        o.write(F.name.data(), (std::streamsize)F.name.size());
        put<int32_t>(o, F.loc_code);
        put<int32_t>(o, F.ncomp);
        put<int32_t>(o, F.nghost);
        for (const RestartBlock &B : F.blocks)
        {
            for (int a = 0; a < 3; ++a)
                put<int32_t>(o, B.lo[a]);
            for (int a = 0; a < 3; ++a)
                put<int32_t>(o, B.hi[a]);
            put<int32_t>(o, 1);
            o.write(reinterpret_cast<const char *>(B.v.data()), (std::streamsize)(B.v.size() * sizeof(double)));
        }
    }
    if (!o)
        throw std::runtime_error("write failed: " + path);
}

std::vector<RestartField> read_restart_file(const std::string &path, long long &step, double &time)
{
    std::ifstream in(path, std::ios::binary);
    if (!in)
        throw std::runtime_error("cannot read " + path);
    char magic[8];
    in.read(magic, 8);
    if (!in || std::memcmp(magic, "MPCNSRST", 8) != 0)
        throw std::runtime_error("restart: magic mismatch");
    if (get<int32_t>(in) != 1)
        throw std::runtime_error("restart: unsupported version");
    step = get<int32_t>(in);
    time = get<double>(in);
    const int nb = get<int32_t>(in), nf = get<int32_t>(in);
    if (nb < 0 || nf < 0 || nb > 100000 || nf > 1000)
        throw std::runtime_error("restart: bad counts");
    std::vector<RestartField> out(nf);
    for (RestartField &F : out)
    {
        const int len = get<int32_t>(in);
        if (len < 0 || len > 4096)
            throw std::runtime_error("restart: bad name length");
        F.name.assign((size_t)len, '\0');
        in.read(&F.name[0], len);
        F.loc_code = get<int32_t>(in);
        F.ncomp = get<int32_t>(in);
        F.nghost = get<int32_t>(in);
        F.blocks.resize(nb);
        for (RestartBlock &B : F.blocks)
        {
            for (int a = 0; a < 3; ++a)
                B.lo[a] = get<int32_t>(in);
            for (int a = 0; a < 3; ++a)
                B.hi[a] = get<int32_t>(in);
            const int active = get<int32_t>(in);
            if (!active)
                continue;
            size_t n = (size_t)std::max(0, F.ncomp);
            for (int a = 0; a < 3; ++a)
            {
                if (B.hi[a] < B.lo[a])
                    throw std::runtime_error("restart: negative extent");
                n *= (size_t)(B.hi[a] - B.lo[a]);
            }
            B.v.resize(n);
            in.read(reinterpret_cast<char *>(B.v.data()), (std::streamsize)(n * sizeof(double)));
            if (!in)
                throw std::runtime_error("restart: truncated data");
        }
    }
    return out;
}

int restart_loc_code(int loc)
{
    // StaggerLocation codes of IOModule::LocToI32_ (src/5_io/IOModule_Bin.cpp:62-84)
    switch (loc)
    {
    case LOC_CELL: return 0;
    case LOC_FACE_XI: return 1;
    case LOC_FACE_ET: return 2;
    case LOC_FACE_ZE: return 3;
    case LOC_EDGE_XI: return 4;
    case LOC_EDGE_ET: return 5;
    case LOC_EDGE_ZE: return 6;
    case LOC_NODE: return 7;
    default: return -999;
    }
}

void copy_overlap(const RestartBlock &src, int ncomp, const int *lo, const int *hi, double *dst)
{
    const int ext[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    const int sext[3] = {src.hi[0] - src.lo[0], src.hi[1] - src.lo[1], src.hi[2] - src.lo[2]};
    if (src.v.empty())
        return;
    for (int i = std::max(lo[0], src.lo[0]); i < std::min(hi[0], src.hi[0]); ++i)
        for (int j = std::max(lo[1], src.lo[1]); j < std::min(hi[1], src.hi[1]); ++j)
            for (int k = std::max(lo[2], src.lo[2]); k < std::min(hi[2], src.hi[2]); ++k)
            {
                const size_t d = (((size_t)(i - lo[0]) * ext[1] + (j - lo[1])) * ext[2] + (k - lo[2])) * ncomp;
                const size_t s = (((size_t)(i - src.lo[0]) * sext[1] + (j - src.lo[1])) * sext[2] + (k - src.lo[2])) * ncomp;
                for (int m = 0; m < ncomp; ++m)
                    dst[d + m] = src.v[s + m];
            }
}
} // namespace mpcns
