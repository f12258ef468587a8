// Host-side setup: grids, topology, ghost nodes, geometry, mimetic weights,
// halo / boundary op lists and initial states.  See setup.h.
#include "setup.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>

namespace mpcns
{

// ------------------------------------------------------------------ random
uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double uniform01(uint64_t &s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

// ------------------------------------------------------------------ grids
static void alloc_nodes(Block &b, int ni, int nj, int nk, int gn)
{
    b.ni = ni;
    b.nj = nj;
    b.nk = nk;
    b.gn = gn;
    const int64_t n = b.node_count();
    b.x.assign(n, 0.0);
    b.y.assign(n, 0.0);
    b.z.assign(n, 0.0);
}

Grid make_periodic_box(int nx, int ny, int nz, double lx, double ly, double lz, int g)
{
    Grid grd;
    grd.g = g;
    grd.blocks.resize(1);
    Block &b = grd.blocks[0];
    alloc_nodes(b, nx, ny, nz, g + 1);
    for (int i = 0; i <= nx; ++i)
        for (int j = 0; j <= ny; ++j)
            for (int k = 0; k <= nz; ++k)
            {
                const int64_t o = b.node_off(i, j, k);
                b.x[o] = lx * i / nx;
                b.y[o] = ly * j / ny;
                b.z[o] = lz * k / nz;
            }
    // Self-periodic interfaces: pure translations by N nodes, as a periodic
    // inner patch with the block as its own neighbour
    // (reference src/1_grid/1_MPCNS_Grid.cpp:456).
    for (int ax = 0; ax < 3; ++ax)
    {
        const int n = b.n(ax);
        for (int side = 0; side < 2; ++side)
        {
            FaceLink &f = b.face[2 * ax + side];
            f.kind = 1;
            f.nb_block = 0;
            f.nb_face = 2 * ax + (1 - side);
            f.periodic = 1;
            for (int a = 0; a < 3; ++a)
            {
                f.trans.perm[a] = a;
                f.trans.sign[a] = 1;
                f.trans.offset[a] = 0;
            }
            f.trans.offset[ax] = side == 0 ? n : -n;
        }
    }
    return grd;
}

// Equiangular cubed-sphere panels.  Panel frames (centre c, u, v) are
// right-handed (u x v = c) so (i, j, k) = (u-angle, v-angle, radius) is a
// right-handed index frame with k pointing outward.
static const double kPanel[6][3][3] = {
    {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}},
    {{0, 1, 0}, {-1, 0, 0}, {0, 0, 1}},
    {{-1, 0, 0}, {0, -1, 0}, {0, 0, 1}},
    {{0, -1, 0}, {1, 0, 0}, {0, 0, 1}},
    {{0, 0, 1}, {0, 1, 0}, {-1, 0, 0}},
    {{0, 0, -1}, {0, 1, 0}, {1, 0, 0}},
};

Grid make_cubed_sphere_shell(int nang, int nrad, double r_in, double r_out, int g, int bc_inner, int bc_outer)
{
    Grid grd;
    grd.g = g;
    grd.blocks.resize(6);
    const double q = std::acos(-1.0) / 4.0;
    for (int p = 0; p < 6; ++p)
    {
        Block &b = grd.blocks[p];
        alloc_nodes(b, nang, nang, nrad, g + 1);
        const double *c = kPanel[p][0], *u = kPanel[p][1], *v = kPanel[p][2];
        for (int i = 0; i <= nang; ++i)
        {
            const double X = std::tan(-q + 2.0 * q * i / nang);
            for (int j = 0; j <= nang; ++j)
            {
                const double Y = std::tan(-q + 2.0 * q * j / nang);
                double d[3];
                for (int a = 0; a < 3; ++a)
                    d[a] = c[a] + X * u[a] + Y * v[a];
                const double inv = 1.0 / std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                for (int k = 0; k <= nrad; ++k)
                {
                    // log-stretched radius r_k = r_in (r_out/r_in)^(k/Nk)
                    const double r = r_in * std::pow(r_out / r_in, (double)k / nrad);
                    const int64_t o = b.node_off(i, j, k);
                    b.x[o] = r * d[0] * inv;
                    b.y[o] = r * d[1] * inv;
                    b.z[o] = r * d[2] * inv;
                }
            }
        }
        b.face[4].kind = 0;
        b.face[4].bc = bc_inner;
        b.face[5].kind = 0;
        b.face[5].bc = bc_outer;
    }
    link_faces_by_coordinates(grd, 1e-9 * r_in);
    return grd;
}

// ------------------------------------------------------------ face linking
static void face_tangents(int face, int &ax, int &t1, int &t2)
{
    ax = face / 2;
    t1 = ax == 0 ? 1 : 0;
    t2 = ax == 2 ? 1 : 2;
}

// The link a point crosses through interface face f: the face's own or, on
// a face several interface patches share, the part covering the point.  dc
// are doubled node coordinates (2 idx + the location's delta per axis),
// clamped onto the face; a point on the line between two parts takes the
// first part that covers it (both map it onto the same physical point).
struct LinkView
{
    int nb_block, nb_face, periodic;
    IndexTransform trans;
};
static LinkView link_at(const Block &B, int f, const int dc[3])
{
    const FaceLink &L = B.face[f];
    if (L.parts.empty())
        return {L.nb_block, L.nb_face, L.periodic, L.trans};
    int ax, t1, t2;
    face_tangents(f, ax, t1, t2);
    const int c1 = std::max(0, std::min(dc[t1], 2 * B.n(t1))), c2 = std::max(0, std::min(dc[t2], 2 * B.n(t2)));
    for (const FacePart &P : L.parts)
        if (2 * P.lo[0] <= c1 && c1 <= 2 * P.hi[0] && 2 * P.lo[1] <= c2 && c2 <= 2 * P.hi[1])
            return {P.nb_block, P.nb_face, P.periodic, P.trans};
    throw std::runtime_error("face parts do not cover the face");
}

static void face_point(const Block &b, int face, int a1, int a2, int idx[3])
{
    int ax, t1, t2;
    face_tangents(face, ax, t1, t2);
    idx[ax] = (face % 2) ? b.n(ax) : 0;
    idx[t1] = a1;
    idx[t2] = a2;
}

static bool find_node(const Block &b, int face, double px, double py, double pz, double tol, int out[3])
{
    int ax, t1, t2;
    face_tangents(face, ax, t1, t2);
    for (int a1 = 0; a1 <= b.n(t1); ++a1)
        for (int a2 = 0; a2 <= b.n(t2); ++a2)
        {
            int idx[3];
            face_point(b, face, a1, a2, idx);
            const int64_t o = b.node_off(idx[0], idx[1], idx[2]);
            if (std::fabs(b.x[o] - px) < tol && std::fabs(b.y[o] - py) < tol && std::fabs(b.z[o] - pz) < tol)
            {
                out[0] = idx[0];
                out[1] = idx[1];
                out[2] = idx[2];
                return true;
            }
        }
    return false;
}

void link_faces_by_coordinates(Grid &grd, double tol)
{
    const int nb = (int)grd.blocks.size();
    for (int b = 0; b < nb; ++b)
        for (int f = 0; f < 6; ++f)
        {
            Block &B = grd.blocks[b];
            if (B.face[f].kind == 0 && B.face[f].bc != BC_NONE)
                continue;
            if (B.face[f].kind == 1)
                continue;
            int ax, t1, t2;
            face_tangents(f, ax, t1, t2);
            int c00[3], c10[3], c01[3];
            face_point(B, f, 0, 0, c00);
            face_point(B, f, B.n(t1), 0, c10);
            face_point(B, f, 0, B.n(t2), c01);
            bool linked = false;
            for (int b2 = 0; b2 < nb && !linked; ++b2)
                for (int f2 = 0; f2 < 6 && !linked; ++f2)
                {
                    if (b2 == b && f2 == f)
                        continue;
                    const Block &C = grd.blocks[b2];
                    if (C.face[f2].kind == 0 && C.face[f2].bc != BC_NONE)
                        continue;
                    int n00[3], n10[3], n01[3];
                    const int64_t o00 = B.node_off(c00[0], c00[1], c00[2]);
                    const int64_t o10 = B.node_off(c10[0], c10[1], c10[2]);
                    const int64_t o01 = B.node_off(c01[0], c01[1], c01[2]);
                    if (!find_node(C, f2, B.x[o00], B.y[o00], B.z[o00], tol, n00))
                        continue;
                    if (!find_node(C, f2, B.x[o10], B.y[o10], B.z[o10], tol, n10))
                        continue;
                    if (!find_node(C, f2, B.x[o01], B.y[o01], B.z[o01], tol, n01))
                        continue;
                    IndexTransform T{};
                    // tangential axes from the corner matches
                    const int ts[2] = {t1, t2};
                    const int *nn[2] = {n10, n01};
                    for (int q = 0; q < 2; ++q)
                    {
                        const int t = ts[q];
                        int moved = -1, s = 0;
                        for (int a = 0; a < 3; ++a)
                            if (nn[q][a] != n00[a])
                            {
                                moved = a;
                                s = nn[q][a] > n00[a] ? 1 : -1;
                            }
                        if (moved < 0 || std::abs(nn[q][moved] - n00[moved]) != B.n(t))
                            throw std::runtime_error("link_faces: inconsistent tangent match");
                        T.perm[t] = moved;
                        T.sign[t] = s;
                    }
                    const int ax2 = f2 / 2;
                    T.perm[ax] = ax2;
                    // outward from b maps to inward into b2
                    // (same rule as reference src/2_topology/2_MPCNS_Topology.cpp:131)
                    T.sign[ax] = ((f % 2) ? 1 : -1) * ((f2 % 2) ? -1 : 1);
                    for (int a = 0; a < 3; ++a)
                        T.offset[a] = n00[T.perm[a]] - T.sign[a] * c00[a];
                    // verify every face node
                    for (int a1 = 0; a1 <= B.n(t1); ++a1)
                        for (int a2 = 0; a2 <= B.n(t2); ++a2)
                        {
                            int idx[3], nbi[3];
                            face_point(B, f, a1, a2, idx);
                            for (int a = 0; a < 3; ++a)
                                nbi[T.perm[a]] = T.sign[a] * idx[a] + T.offset[a];
                            const int64_t o = B.node_off(idx[0], idx[1], idx[2]);
                            const int64_t o2 = C.node_off(nbi[0], nbi[1], nbi[2]);
                            if (std::fabs(B.x[o] - C.x[o2]) > tol || std::fabs(B.y[o] - C.y[o2]) > tol ||
                                std::fabs(B.z[o] - C.z[o2]) > tol)
                                throw std::runtime_error("link_faces: node mismatch");
                        }
                    FaceLink &L = B.face[f];
                    L.kind = 1;
                    L.bc = BC_NONE;
                    L.nb_block = b2;
                    L.nb_face = f2;
                    L.periodic = 0;
                    L.trans = T;
                    linked = true;
                }
            if (!linked)
                throw std::runtime_error("link_faces: unmatched block face");
        }
}

// ---------------------------------------------------------- ghost nodes
// Ghost node coordinates.  Out-of-range axes facing an interface are crossed
// once (the neighbour's interior node, as the reference's interface mesh
// copy); what is still out of range afterwards -- physical and periodic faces,
// and the second axis at a cube vertex where only three blocks meet -- is
// mirrored, x(-g) = 2 x(0) - x(g) (reference src/1_grid/1_MPCNS_Grid.cpp:635).
static void node_get(const Grid &grd, int b, const int n[3], double p[3], int depth, bool hopped)
{
    if (depth > 32)
        throw std::runtime_error("node_get: resolution did not terminate");
    const Block &B = grd.blocks[b];
    int out_ax = -1, out_side = 0;
    for (int a = 0; a < 3 && out_ax < 0 && !hopped; ++a)
    {
        const int side = n[a] < 0 ? 0 : (n[a] > B.n(a) ? 1 : -1);
        if (side < 0)
            continue;
        const FaceLink &L = B.face[2 * a + side];
        if (L.kind == 1 && !L.periodic)
        {
            out_ax = a;
            out_side = side;
        }
    }
    if (out_ax >= 0)
    {
        const int dc[3] = {2 * n[0], 2 * n[1], 2 * n[2]};
        const LinkView L = link_at(B, 2 * out_ax + out_side, dc);
        int m[3];
        for (int a = 0; a < 3; ++a)
            m[L.trans.perm[a]] = L.trans.sign[a] * n[a] + L.trans.offset[a];
        node_get(grd, L.nb_block, m, p, depth + 1, true);
        return;
    }
    for (int a = 0; a < 3; ++a)
    {
        if (n[a] >= 0 && n[a] <= B.n(a))
            continue;
        const int edge = n[a] < 0 ? 0 : B.n(a);
        int m0[3] = {n[0], n[1], n[2]}, m1[3] = {n[0], n[1], n[2]};
        m0[a] = edge;
        m1[a] = 2 * edge - n[a];
        double p0[3], p1[3];
        node_get(grd, b, m0, p0, depth + 1, hopped);
        node_get(grd, b, m1, p1, depth + 1, hopped);
        for (int c = 0; c < 3; ++c)
            p[c] = 2 * p0[c] - p1[c];
        return;
    }
    const int64_t o = B.node_off(n[0], n[1], n[2]);
    p[0] = B.x[o];
    p[1] = B.y[o];
    p[2] = B.z[o];
}

void fill_ghost_nodes(Grid &grd)
{
    // Resolve into a copy so every ghost is computed from interior nodes.
    std::vector<Block> out = grd.blocks;
    for (int b = 0; b < (int)grd.blocks.size(); ++b)
    {
        const Block &B = grd.blocks[b];
        const int gn = B.gn;
#pragma omp parallel for schedule(dynamic)
        for (int i = -gn; i <= B.ni + gn; ++i)
            for (int j = -gn; j <= B.nj + gn; ++j)
                for (int k = -gn; k <= B.nk + gn; ++k)
                {
                    if (i >= 0 && i <= B.ni && j >= 0 && j <= B.nj && k >= 0 && k <= B.nk)
                        continue;
                    const int n[3] = {i, j, k};
                    double p[3];
                    node_get(grd, b, n, p, 0, false);
                    const int64_t o = B.node_off(i, j, k);
                    out[b].x[o] = p[0];
                    out[b].y[o] = p[1];
                    out[b].z[o] = p[2];
                }
    }
    grd.blocks.swap(out);
}

// --------------------------------------------------------------- geometry
namespace
{
using V3 = std::array<double, 3>;
inline V3 vsub(const V3 &a, const V3 &b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
inline V3 vadd(const V3 &a, const V3 &b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
inline V3 vcross(const V3 &a, const V3 &b)
{
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
inline double vdot(const V3 &a, const V3 &b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline double vnorm(const V3 &a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

struct NodeView
{
    const Block &B;
    V3 P(int i, int j, int k) const
    {
        const int64_t o = B.node_off(i, j, k);
        return {B.x[o], B.y[o], B.z[o]};
    }
    // cell centre = average of the 8 nodes
    V3 C(int i, int j, int k) const
    {
        V3 s{0, 0, 0};
        const int di[8] = {0, 1, 0, 1, 0, 1, 0, 1}, dj[8] = {0, 0, 1, 1, 0, 0, 1, 1}, dk[8] = {0, 0, 0, 0, 1, 1, 1, 1};
        double acc[3] = {0, 0, 0};
        for (int q = 0; q < 8; ++q)
        {
            const V3 p = P(i + di[q], j + dj[q], k + dk[q]);
            acc[0] += p[0];
            acc[1] += p[1];
            acc[2] += p[2];
        }
        s = {0.125 * acc[0], 0.125 * acc[1], 0.125 * acc[2]};
        return s;
    }
    // quad area vector from origin o and its three neighbours (two triangles)
    static V3 quad(const V3 &ori, const V3 &oxp, const V3 &oyp, const V3 &oxyp)
    {
        const V3 A = vadd(vcross(vsub(oxp, ori), vsub(oyp, ori)), vcross(vsub(oyp, oxyp), vsub(oxp, oxyp)));
        return {0.5 * A[0], 0.5 * A[1], 0.5 * A[2]};
    }
    // face area vectors JD (reference src/3_field/4_MPCNS_Geometry.cpp:98-201)
    V3 JDxi(int i, int j, int k) const { return quad(P(i, j, k), P(i, j + 1, k), P(i, j, k + 1), P(i, j + 1, k + 1)); }
    V3 JDet(int i, int j, int k) const { return quad(P(i, j, k), P(i, j, k + 1), P(i + 1, j, k), P(i + 1, j, k + 1)); }
    V3 JDze(int i, int j, int k) const { return quad(P(i, j, k), P(i + 1, j, k), P(i, j + 1, k), P(i + 1, j + 1, k)); }
    V3 face_centre(int ax, int i, int j, int k) const
    {
        V3 a, b, c, d;
        if (ax == 0)
        {
            a = P(i, j, k), b = P(i, j + 1, k), c = P(i, j, k + 1), d = P(i, j + 1, k + 1);
        }
        else if (ax == 1)
        {
            a = P(i, j, k), b = P(i + 1, j, k), c = P(i, j, k + 1), d = P(i + 1, j, k + 1);
        }
        else
        {
            a = P(i, j, k), b = P(i + 1, j, k), c = P(i, j + 1, k), d = P(i + 1, j + 1, k);
        }
        return {0.25 * (a[0] + b[0] + c[0] + d[0]), 0.25 * (a[1] + b[1] + c[1] + d[1]),
                0.25 * (a[2] + b[2] + c[2] + d[2])};
    }
    // Jac = 1/3 sum_f A_f . (x_f - x_c)  (reference :203-315)
    double Jac(int i, int j, int k) const
    {
        const V3 xc = C(i, j, k);
        double V = 0.0;
        auto nf = [](V3 a) { return V3{-a[0], -a[1], -a[2]}; };
        V += vdot(nf(JDxi(i, j, k)), vsub(face_centre(0, i, j, k), xc));
        V += vdot(JDxi(i + 1, j, k), vsub(face_centre(0, i + 1, j, k), xc));
        V += vdot(nf(JDet(i, j, k)), vsub(face_centre(1, i, j, k), xc));
        V += vdot(JDet(i, j + 1, k), vsub(face_centre(1, i, j + 1, k), xc));
        V += vdot(nf(JDze(i, j, k)), vsub(face_centre(2, i, j, k), xc));
        V += vdot(JDze(i, j, k + 1), vsub(face_centre(2, i, j, k + 1), xc));
        return V / 3.0;
    }
};

// Symmetric 3x3 Jacobi eigen-solver, 12 cyclic sweeps, eigenvalues sorted
// descending (reference Z4_Lunar/src/LunarSolver.cpp:303-371).
void sym_eig3(double A[3][3], double V[3][3], double d[3])
{
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            V[a][b] = a == b ? 1.0 : 0.0;
    auto rotate = [&](int p, int q) {
        if (std::fabs(A[p][q]) < 1e-300)
            return;
        const double app = A[p][p], aqq = A[q][q], apq = A[p][q];
        const double tau = (aqq - app) / (2.0 * apq);
        const double t = tau >= 0.0 ? 1.0 / (tau + std::sqrt(1.0 + tau * tau)) : -1.0 / (-tau + std::sqrt(1.0 + tau * tau));
        const double c = 1.0 / std::sqrt(1.0 + t * t);
        const double s = t * c;
        for (int k = 0; k < 3; ++k)
        {
            if (k == p || k == q)
                continue;
            const double akp = A[k][p], akq = A[k][q];
            A[k][p] = c * akp - s * akq;
            A[p][k] = A[k][p];
            A[k][q] = s * akp + c * akq;
            A[q][k] = A[k][q];
        }
        A[p][p] = c * c * app - 2.0 * s * c * apq + s * s * aqq;
        A[q][q] = s * s * app + 2.0 * s * c * apq + c * c * aqq;
        A[p][q] = A[q][p] = 0.0;
        for (int k = 0; k < 3; ++k)
        {
            const double vkp = V[k][p], vkq = V[k][q];
            V[k][p] = c * vkp - s * vkq;
            V[k][q] = s * vkp + c * vkq;
        }
    };
    for (int it = 0; it < 12; ++it)
    {
        rotate(0, 1);
        rotate(0, 2);
        rotate(1, 2);
    }
    for (int a = 0; a < 3; ++a)
        d[a] = A[a][a];
    for (int a = 0; a < 2; ++a)
        for (int b = a + 1; b < 3; ++b)
            if (d[b] > d[a])
            {
                std::swap(d[a], d[b]);
                for (int k = 0; k < 3; ++k)
                    std::swap(V[k][a], V[k][b]);
            }
}

// Pseudo-inverse least-squares weights: W_n = M^+ (mu_n s_n),
// M = sum mu_n s_n s_n^T, with relative activity threshold tau_act and
// eigenvalue truncation tau_eig (reference :489-622 and :794-912).  M^+ is
// formed on and above the diagonal and mirrored (it is symmetric).  Mi_out
// (6: 00, 01, 02, 11, 12, 22), mu_out (n: mu_q, 0 where q is inactive) and
// active_out, when given, receive M^+, the weights and the mask of active q
// (all zero where every weight is zero), from which
// W[c n + q] = Mi[c][0] (mu_q s_q[0]) + Mi[c][1] (mu_q s_q[1]) + Mi[c][2] (mu_q s_q[2]).
void pinv_weights(int n, const double (*s)[3], const double *mu_in, double tau_act, double *W /*3*n*/,
                  double *Mi_out = nullptr, unsigned *active_out = nullptr, double *mu_out = nullptr)
{
    for (int q = 0; q < 3 * n; ++q)
        W[q] = 0.0;
    if (Mi_out)
        for (int q = 0; q < 6; ++q)
            Mi_out[q] = 0.0;
    if (mu_out)
        for (int q = 0; q < n; ++q)
            mu_out[q] = 0.0;
    if (active_out)
        *active_out = 0;
    double mag[36], ref = 0.0;
    for (int q = 0; q < n; ++q)
    {
        mag[q] = std::sqrt(s[q][0] * s[q][0] + s[q][1] * s[q][1] + s[q][2] * s[q][2]);
        ref = std::max(ref, mag[q]);
    }
    if (ref <= 1e-300)
        return;
    bool active[36];
    double mu[36];
    double M[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int q = 0; q < n; ++q)
    {
        active[q] = mag[q] > tau_act * ref;
        if (!active[q])
        {
            mu[q] = 0.0;
            continue;
        }
        mu[q] = mu_in[q];
        M[0][0] += mu[q] * s[q][0] * s[q][0];
        M[0][1] += mu[q] * s[q][0] * s[q][1];
        M[0][2] += mu[q] * s[q][0] * s[q][2];
        M[1][1] += mu[q] * s[q][1] * s[q][1];
        M[1][2] += mu[q] * s[q][1] * s[q][2];
        M[2][2] += mu[q] * s[q][2] * s[q][2];
    }
    M[1][0] = M[0][1];
    M[2][0] = M[0][2];
    M[2][1] = M[1][2];
    double A[3][3], V[3][3], lam[3];
    std::memcpy(A, M, sizeof(A));
    sym_eig3(A, V, lam);
    const double lmax = std::max(lam[0], std::max(lam[1], lam[2]));
    if (lmax <= 1e-300)
        return;
    double il[3];
    for (int q = 0; q < 3; ++q)
        il[q] = lam[q] > 1e-10 * lmax ? 1.0 / lam[q] : 0.0;
    double Mi[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int a = 0; a < 3; ++a)
        for (int b = a; b < 3; ++b)
        {
            for (int q = 0; q < 3; ++q)
                Mi[a][b] += V[a][q] * il[q] * V[b][q];
            Mi[b][a] = Mi[a][b];
        }
    if (Mi_out)
    {
        Mi_out[0] = Mi[0][0];
        Mi_out[1] = Mi[0][1];
        Mi_out[2] = Mi[0][2];
        Mi_out[3] = Mi[1][1];
        Mi_out[4] = Mi[1][2];
        Mi_out[5] = Mi[2][2];
    }
    for (int q = 0; q < n; ++q)
    {
        if (!active[q])
            continue;
        if (active_out)
            *active_out |= 1u << q;
        if (mu_out)
            mu_out[q] = mu[q];
        const double rx = mu[q] * s[q][0], ry = mu[q] * s[q][1], rz = mu[q] * s[q][2];
        W[q] = Mi[0][0] * rx + Mi[0][1] * ry + Mi[0][2] * rz;
        W[n + q] = Mi[1][0] * rx + Mi[1][1] * ry + Mi[1][2] * rz;
        W[2 * n + q] = Mi[2][0] * rx + Mi[2][1] * ry + Mi[2][2] * rz;
    }
}

// alpha = |e|/|S*| with the reference's axis cap
// (src/3_field/4_MPCNS_Geometry.cpp:532-566).
double capped_alpha(double L, const V3 &p00, const V3 &p10, const V3 &p01, const V3 &p11, double Amag, double eps)
{
    const double alpha_raw = Amag < eps ? 0.0 : L / Amag;
    const double d00_10 = vnorm(vsub(p10, p00)), d00_01 = vnorm(vsub(p01, p00));
    const double d10_11 = vnorm(vsub(p11, p10)), d01_11 = vnorm(vsub(p11, p01));
    const double d00_11 = vnorm(vsub(p11, p00)), d10_01 = vnorm(vsub(p01, p10));
    const double href = std::max(std::max(std::max(d00_10, d00_01), std::max(d10_11, d01_11)),
                                 std::max(std::max(d00_11, d10_01), std::max(L, eps)));
    const double side_min = std::min(std::min(d00_10, d00_01), std::min(d10_11, d01_11));
    const double diag_min = std::min(d00_11, d10_01);
    const bool touches = L < 1e-12 || side_min < 1e-3 * href || diag_min < 1e-3 * href || Amag < 1e-3 * href * href;
    return touches ? 0.0 : alpha_raw;
}
} // namespace

void build_geometry(const Grid &grd, std::vector<BlockArrays> &blk)
{
    const int g = grd.g;
    const double eps = 1e-25;
    blk.resize(grd.blocks.size());
    // one slot size for every block of the case (the largest box, 512-byte
    // aligned): a case's blocks sit at multiples of it in one allocation and
    // its op lists address them so
    int64_t case_slot = 0;
    for (const Block &B : grd.blocks)
    {
        const int64_t raw = (int64_t)(B.ni + 1 + 2 * g) * (B.nj + 1 + 2 * g) * (B.nk + 1 + 2 * g);
        case_slot = std::max(case_slot, (raw + 63) / 64 * 64);
    }
    for (size_t b = 0; b < grd.blocks.size(); ++b)
    {
        const Block &B = grd.blocks[b];
        if (B.gn < g + 1)
            throw std::runtime_error("build_geometry: need g+1 ghost node layers");
        BlockArrays &A = blk[b];
        MhdBlockDev &d = A.d;
        d.ni = B.ni;
        d.nj = B.nj;
        d.nk = B.nk;
        d.g = g;
        d.sj = B.nj + 1 + 2 * g;
        d.sk = B.nk + 1 + 2 * g;
        d.slot = case_slot;
        for (int f = 0; f < 6; ++f)
            d.bc[f] = B.face[f].kind == 0 ? B.face[f].bc : BC_NONE;
        A.geo.assign((size_t)G_NSLOT * d.slot, 0.0);
        A.wj.assign((size_t)36 * d.slot, 0.0);
        A.wb.assign((size_t)18 * d.slot, 0.0);
        NodeView nv{B};
        auto G = [&](int s, int i, int j, int k) -> double & { return A.geo[(size_t)s * d.slot + blk_off(d, i, j, k)]; };
#pragma omp parallel for schedule(dynamic)
        for (int i = -g; i <= B.ni + g; ++i)
            for (int j = -g; j <= B.nj + g; ++j)
                for (int k = -g; k <= B.nk + g; ++k)
                {
                    const V3 jx = nv.JDxi(i, j, k), je = nv.JDet(i, j, k), jz = nv.JDze(i, j, k);
                    for (int c = 0; c < 3; ++c)
                    {
                        G(G_JD + c, i, j, k) = jx[c];
                        G(G_JD + 3 + c, i, j, k) = je[c];
                        G(G_JD + 6 + c, i, j, k) = jz[c];
                    }
                    G(G_JAC, i, j, k) = nv.Jac(i, j, k);
                    const V3 node = nv.P(i, j, k);
                    for (int c = 0; c < 3; ++c)
                        G(G_NODE + c, i, j, k) = node[c];
                    // face Hodge beta = |l*|/|S|  (reference :392-478)
                    const V3 cc = nv.C(i, j, k);
                    const V3 jd[3] = {jx, je, jz};
                    const V3 cl[3] = {nv.C(i - 1, j, k), nv.C(i, j - 1, k), nv.C(i, j, k - 1)};
                    for (int a = 0; a < 3; ++a)
                    {
                        const double Smag = vnorm(jd[a]);
                        const double Ls = vnorm(vsub(cc, cl[a]));
                        G(G_BETA + a, i, j, k) = Smag < eps ? 0.0 : Ls / Smag;
                    }
                    // edge Hodge alpha = |e|/|S*|  (reference :568-656)
                    {
                        const double L = vnorm(vsub(nv.P(i + 1, j, k), nv.P(i, j, k)));
                        const V3 p00 = nv.C(i, j - 1, k - 1), p10 = nv.C(i, j, k - 1), p01 = nv.C(i, j - 1, k), p11 = cc;
                        const V3 Av = NodeView::quad(p00, p10, p01, p11);
                        G(G_ALPHA + 0, i, j, k) = capped_alpha(L, p00, p10, p01, p11, vnorm(Av), eps);
                    }
                    {
                        const double L = vnorm(vsub(nv.P(i, j + 1, k), nv.P(i, j, k)));
                        const V3 p00 = nv.C(i - 1, j, k - 1), p10 = nv.C(i, j, k - 1), p01 = nv.C(i - 1, j, k), p11 = cc;
                        const V3 Av = NodeView::quad(p00, p10, p01, p11);
                        G(G_ALPHA + 1, i, j, k) = capped_alpha(L, p00, p10, p01, p11, vnorm(Av), eps);
                    }
                    {
                        const double L = vnorm(vsub(nv.P(i, j, k + 1), nv.P(i, j, k)));
                        const V3 p00 = nv.C(i - 1, j - 1, k), p10 = nv.C(i, j - 1, k), p01 = nv.C(i - 1, j, k), p11 = cc;
                        const V3 Av = NodeView::quad(p00, p10, p01, p11);
                        G(G_ALPHA + 2, i, j, k) = capped_alpha(L, p00, p10, p01, p11, vnorm(Av), eps);
                    }
                }
        // mimetic weights on interior cells only
#pragma omp parallel for schedule(dynamic)
        for (int i = 0; i < B.ni; ++i)
            for (int j = 0; j < B.nj; ++j)
                for (int k = 0; k < B.nk; ++k)
                {
                    const V3 xc = nv.C(i, j, k);
                    // B_cell from six oriented face fluxes: -xi,+xi,-eta,+eta,-zeta,+zeta
                    double S[6][3], mu[6];
                    const V3 f[6] = {nv.JDxi(i, j, k), nv.JDxi(i + 1, j, k), nv.JDet(i, j, k),
                                     nv.JDet(i, j + 1, k), nv.JDze(i, j, k), nv.JDze(i, j, k + 1)};
                    const V3 fc[6] = {nv.face_centre(0, i, j, k), nv.face_centre(0, i + 1, j, k),
                                      nv.face_centre(1, i, j, k), nv.face_centre(1, i, j + 1, k),
                                      nv.face_centre(2, i, j, k), nv.face_centre(2, i, j, k + 1)};
                    for (int q = 0; q < 6; ++q)
                    {
                        const double sg = (q % 2 == 0) ? -1.0 : 1.0;
                        for (int c = 0; c < 3; ++c)
                            S[q][c] = sg * f[q][c];
                        mu[q] = 1.0 / std::max(vnorm(vsub(fc[q], xc)), 1e-14);
                    }
                    double Wb[18], Mb[6], mub[6];
                    pinv_weights(6, S, mu, 1e-10, Wb, Mb, nullptr, mub);
                    const int64_t ob = blk_off(d, i, j, k);
                    for (int q = 0; q < 18; ++q)
                        A.wb[(size_t)q * d.slot + ob] = Wb[q];
                    for (int q = 0; q < 6; ++q)
                    {
                        G(G_WBM + q, i, j, k) = Mb[q];
                        G(G_WBU + q, i, j, k) = mub[q];
                    }
                    // J_cell from twelve edges: xi (0..3), eta (4..7), zeta (8..11)
                    double Le[12][3], one[12];
                    const int e[12][6] = {{0, 0, 0, 1, 0, 0}, {0, 1, 0, 1, 1, 0}, {0, 0, 1, 1, 0, 1}, {0, 1, 1, 1, 1, 1},
                                          {0, 0, 0, 0, 1, 0}, {1, 0, 0, 1, 1, 0}, {0, 0, 1, 0, 1, 1}, {1, 0, 1, 1, 1, 1},
                                          {0, 0, 0, 0, 0, 1}, {1, 0, 0, 1, 0, 1}, {0, 1, 0, 0, 1, 1}, {1, 1, 0, 1, 1, 1}};
                    for (int q = 0; q < 12; ++q)
                    {
                        const V3 d0 = vsub(nv.P(i + e[q][3], j + e[q][4], k + e[q][5]), nv.P(i + e[q][0], j + e[q][1], k + e[q][2]));
                        Le[q][0] = d0[0];
                        Le[q][1] = d0[1];
                        Le[q][2] = d0[2];
                        one[q] = 1.0;
                    }
                    double Wj[36], Mi[6];
                    unsigned act = 0;
                    pinv_weights(12, Le, one, 1e-10, Wj, Mi, &act);
                    const int64_t o = blk_off(d, i, j, k);
                    for (int q = 0; q < 36; ++q)
                        A.wj[(size_t)q * d.slot + o] = Wj[q];
                    for (int q = 0; q < 6; ++q)
                        G(G_WJM + q, i, j, k) = Mi[q];
                    G(G_WJA, i, j, k) = (double)act;
                }
    }
}

// ------------------------------------------------- geometry key and tables
namespace
{
// 64-bit hash of a run of words (four interleaved multiply-rotate lanes).
uint64_t hash_words(const uint64_t *w, size_t n, uint64_t seed)
{
    const uint64_t P1 = 0x9E3779B185EBCA87ull, P2 = 0xC2B2AE3D27D4EB4Full;
    uint64_t h[4] = {seed + P1 + P2, seed + P2, seed, seed - P1};
    auto round = [&](uint64_t acc, uint64_t x) {
        acc += x * P2;
        acc = (acc << 31) | (acc >> 33);
        return acc * P1;
    };
    size_t q = 0;
    for (; q + 4 <= n; q += 4)
        for (int l = 0; l < 4; ++l)
            h[l] = round(h[l], w[q + l]);
    uint64_t r = h[0] ^ (h[1] * 3) ^ (h[2] * 5) ^ (h[3] * 7) ^ n;
    for (; q < n; ++q)
        r = round(r, w[q]);
    r ^= r >> 33;
    r *= 0xff51afd7ed558ccdull;
    r ^= r >> 33;
    return r;
}
} // namespace

std::string geometry_key(const std::vector<BlockArrays> &blk)
{
    const size_t chunk = (size_t)1 << 20;
    std::vector<std::pair<size_t, size_t>> runs; // (block, first word)
    for (size_t b = 0; b < blk.size(); ++b)
        for (size_t o = 0; o < blk[b].geo.size(); o += chunk)
            runs.emplace_back(b, o);
    std::vector<uint64_t> part(runs.size());
#pragma omp parallel for schedule(dynamic)
    for (long r = 0; r < (long)runs.size(); ++r)
    {
        const std::vector<double> &g = blk[runs[r].first].geo;
        const size_t o = runs[r].second, n = std::min(chunk, g.size() - o);
        part[r] = hash_words(reinterpret_cast<const uint64_t *>(g.data() + o), n, (uint64_t)r);
    }
    const uint64_t h = hash_words(part.data(), part.size(), 0x243F6A8885A308D3ull);
    std::string key = "geo:" + std::to_string(blk.size());
    if (!blk.empty())
    {
        const MhdBlockDev &d = blk[0].d;
        key += ":" + std::to_string(d.ni) + "x" + std::to_string(d.nj) + "x" + std::to_string(d.nk) + ":g" +
               std::to_string(d.g);
    }
    char hex[24];
    std::snprintf(hex, sizeof hex, ":%016llx", (unsigned long long)h);
    return key + hex;
}

void formed_bcell_weights(const BlockArrays &A, int i, int j, int k, double W[18])
{
    const MhdBlockDev &d = A.d;
    const int64_t sl = d.slot, o = blk_off(d, i, j, k), step[3] = {(int64_t)d.sj * d.sk, (int64_t)d.sk, 1};
    const double *g = A.geo.data();
    const double *mi = g + (size_t)G_WBM * sl, *mu = g + (size_t)G_WBU * sl, *jd = g + (size_t)G_JD * sl;
    const double m00 = mi[o], m01 = mi[sl + o], m02 = mi[2 * sl + o], m11 = mi[3 * sl + o], m12 = mi[4 * sl + o],
                 m22 = mi[5 * sl + o];
    for (int n = 0; n < 6; ++n)
    {
        const int ax = n >> 1;
        const int64_t oo = o + ((n & 1) ? step[ax] : 0);
        double sv[3];
        for (int a = 0; a < 3; ++a)
        {
            const double v = jd[(size_t)(3 * ax + a) * sl + oo];
            sv[a] = (n & 1) ? v : -v;
        }
        const double m = mu[(size_t)n * sl + o];
        const double rx = m * sv[0], ry = m * sv[1], rz = m * sv[2];
        W[n] = m00 * rx + m01 * ry + m02 * rz;
        W[6 + n] = m01 * rx + m11 * ry + m12 * rz;
        W[12 + n] = m02 * rx + m12 * ry + m22 * rz;
    }
}

void formed_jcell_weights(const BlockArrays &A, int i, int j, int k, double W[36])
{
    // edge q runs from node kEs[q] to kEe[q] of the cell, node m = (i + (m >> 2),
    // j + (m & 1), k + ((m >> 1) & 1)): xi edges 0..3, eta 4..7, zeta 8..11
    static const int kEs[12] = {0, 1, 2, 3, 0, 4, 2, 6, 0, 4, 1, 5};
    static const int kEe[12] = {4, 5, 6, 7, 1, 5, 3, 7, 2, 6, 3, 7};
    const MhdBlockDev &d = A.d;
    const int64_t sl = d.slot, o = blk_off(d, i, j, k);
    const double *g = A.geo.data();
    const double *X = g + (size_t)G_NODE * sl, *mw = g + (size_t)G_WJM * sl;
    const int64_t q[4] = {0, (int64_t)d.sk, 1, (int64_t)d.sk + 1}, di = (int64_t)d.sj * d.sk;
    double nd[8][3];
    for (int m = 0; m < 8; ++m)
        for (int c = 0; c < 3; ++c)
            nd[m][c] = X[(size_t)c * sl + o + (m >= 4 ? di : 0) + q[m & 3]];
    const double m00 = mw[o], m01 = mw[sl + o], m02 = mw[2 * sl + o], m11 = mw[3 * sl + o], m12 = mw[4 * sl + o],
                 m22 = mw[5 * sl + o];
    const double mi[9] = {m00, m01, m02, m01, m11, m12, m02, m12, m22};
    const unsigned act = (unsigned)g[(size_t)G_WJA * sl + o];
    for (int n = 0; n < 12; ++n)
    {
        const double rx = nd[kEe[n]][0] - nd[kEs[n]][0], ry = nd[kEe[n]][1] - nd[kEs[n]][1],
                     rz = nd[kEe[n]][2] - nd[kEs[n]][2];
        const bool on = (act >> n) & 1u;
        W[n] = on ? mi[0] * rx + mi[1] * ry + mi[2] * rz : 0.0;
        W[12 + n] = on ? mi[3] * rx + mi[4] * ry + mi[5] * rz : 0.0;
        W[24 + n] = on ? mi[6] * rx + mi[7] * ry + mi[8] * rz : 0.0;
    }
}

bool weight_tables_match(const BlockArrays &A, const double *wb_aos, const double *wj_aos, double rtol,
                         std::string &why)
{
    const MhdBlockDev &d = A.d;
    // AoS FieldBlock of a cell field with ghost g: [i][j][k][m] over (n + 2g)^3
    const int ej = d.nj + 2 * d.g, ek = d.nk + 2 * d.g;
    for (int i = 0; i < d.ni; ++i)
        for (int j = 0; j < d.nj; ++j)
            for (int k = 0; k < d.nk; ++k)
            {
                const size_t cell = ((size_t)(i + d.g) * ej + (size_t)(j + d.g)) * ek + (size_t)(k + d.g);
                for (int t = 0; t < 2; ++t)
                {
                    const double *given = t == 0 ? wb_aos : wj_aos;
                    if (!given)
                        continue;
                    const int nw = t == 0 ? 18 : 36;
                    double W[36];
                    if (t == 0)
                        formed_bcell_weights(A, i, j, k, W);
                    else
                        formed_jcell_weights(A, i, j, k, W);
                    double scale = 0.0;
                    for (int q = 0; q < nw; ++q)
                        scale = std::max(scale, std::fabs(W[q]));
                    for (int q = 0; q < nw; ++q)
                    {
                        const double v = given[cell * nw + q];
                        if (!(std::fabs(v - W[q]) <= rtol * scale))
                        {
                            char buf[192];
                            std::snprintf(buf, sizeof buf, "%s weight %d of cell (%d, %d, %d): %.17g, the geometry gives %.17g",
                                          t == 0 ? "Bcell_from_Bface_w" : "Jcell_from_Jedge_w", q, i, j, k, v, W[q]);
                            why = buf;
                            return false;
                        }
                    }
                }
            }
    return true;
}

// ------------------------------------------------------------- halo rules
static int map_loc(int loc, const int perm[3])
{
    if (loc >= LOC_FACE_XI && loc <= LOC_FACE_ZE)
        return LOC_FACE_XI + perm[loc - LOC_FACE_XI];
    if (loc >= LOC_EDGE_XI && loc <= LOC_EDGE_ZE)
        return LOC_EDGE_XI + perm[loc - LOC_EDGE_XI];
    return loc;
}

DofSource resolve_dof(const Grid &grd, int b, int loc, const int idx_in[3])
{
    DofSource r{};
    r.block = b;
    r.loc = loc;
    for (int a = 0; a < 3; ++a)
    {
        r.idx[a] = idx_in[a];
        r.perm[a] = a;
        r.sign[a] = 1;
    }
    int hops = 0;
    for (;;)
    {
        const Block &B = grd.blocks[r.block];
        int nout = 0, hop_ax = -1, hop_side = 0;
        for (int a = 0; a < 3; ++a)
        {
            const int cnt = loc_count(r.loc, a, B.n(a));
            const int side = r.idx[a] < 0 ? 0 : (r.idx[a] >= cnt ? 1 : -1);
            if (side < 0)
                continue;
            ++nout;
            if (hop_ax < 0 && B.face[2 * a + side].kind == 1)
            {
                hop_ax = a;
                hop_side = side;
            }
        }
        if (hop_ax >= 0)
        {
            if (++hops > 8)
                throw std::runtime_error("resolve_dof: too many hops");
            int dc[3], ndc[3];
            for (int a = 0; a < 3; ++a)
                dc[a] = 2 * r.idx[a] + loc_delta(r.loc, a);
            const LinkView L = link_at(B, 2 * hop_ax + hop_side, dc);
            const IndexTransform &T = L.trans;
            for (int a = 0; a < 3; ++a)
                ndc[T.perm[a]] = T.sign[a] * dc[a] + 2 * T.offset[a];
            const int nloc = map_loc(r.loc, T.perm);
            for (int a = 0; a < 3; ++a)
                r.idx[a] = (ndc[a] - loc_delta(nloc, a)) / 2;
            r.loc = nloc;
            // compose frame maps (start axis -> current block axis)
            for (int a = 0; a < 3; ++a)
            {
                const int pa = r.perm[a];
                r.sign[a] *= T.sign[pa];
                r.perm[a] = T.perm[pa];
            }
            r.block = L.nb_block;
            continue;
        }
        if (nout == 0)
            r.kind = 0;
        else if (nout == 1)
            r.kind = hops == 0 ? 3 : 1;
        else
        {
            r.kind = 2;
            for (int a = 0; a < 3; ++a)
            {
                const int cnt = loc_count(r.loc, a, B.n(a));
                r.idx[a] = std::max(0, std::min(r.idx[a], cnt - 1));
            }
        }
        return r;
    }
}

// ------------------------------------------------------------- op lists
namespace
{
struct OpCtx
{
    const Grid &grd;
    const std::vector<BlockArrays> &blk;
    int64_t blk_stride() const { return (int64_t)st_nslot(blk[0].d.model) * blk[0].d.slot; }
    int64_t ref(int b, int slot, int i, int j, int k) const
    {
        return (int64_t)b * blk_stride() + (int64_t)slot * blk[b].d.slot + blk_off(blk[b].d, i, j, k);
    }
};

MhdOp mk(int64_t dst, int64_t src, int kind, int ncomp, int64_t stride, double val = 0.0)
{
    MhdOp o;
    o.dst = dst;
    o.src = src;
    o.stride = stride;
    o.val = val;
    o.kind = kind;
    o.ncomp = ncomp;
    o.cas = 0;
    o.run = 1;
    o.dstep = 0;
    o.sstep = 0;
    o.fold = 0;
    return o;
}

// Iterate over the ghost DOFs of `loc` in block b (depth g) and add halo
// copies.  `slot_of(loc)` gives the state slot of the field for that
// location; `vec3` marks 3-component edge-direction data (Eface) that is
// permuted with the frame; `cellvec` marks Cartesian cell data copied as is.
template <class SlotOf>
void halo_ops(const OpCtx &C, int b, int loc, int ncomp, bool eface, SlotOf slot_of, std::vector<MhdOp> &out)
{
    const Block &B = C.grd.blocks[b];
    const int g = C.grd.g;
    const int cnt[3] = {loc_count(loc, 0, B.ni), loc_count(loc, 1, B.nj), loc_count(loc, 2, B.nk)};
    const int64_t stride = C.blk[b].d.slot;
    for (int i = -g; i < cnt[0] + g; ++i)
        for (int j = -g; j < cnt[1] + g; ++j)
            for (int k = -g; k < cnt[2] + g; ++k)
            {
                const int idx[3] = {i, j, k};
                bool ghost = false;
                for (int a = 0; a < 3; ++a)
                    ghost |= idx[a] < 0 || idx[a] >= cnt[a];
                if (!ghost)
                    continue;
                const DofSource s = resolve_dof(C.grd, b, loc, idx);
                if (s.kind == 3)
                    continue; // written by the block's own boundary handler
                const int64_t dst = C.ref(b, slot_of(loc), i, j, k);
                if (loc == LOC_CELL || loc == LOC_NODE)
                {
                    out.push_back(mk(dst, C.ref(s.block, slot_of(s.loc), s.idx[0], s.idx[1], s.idx[2]), OP_COPY, ncomp, stride));
                }
                else if (!eface)
                {
                    const int ax = (loc >= LOC_EDGE_XI) ? loc - LOC_EDGE_XI : loc - LOC_FACE_XI;
                    const int sg = s.sign[ax];
                    out.push_back(mk(dst, C.ref(s.block, slot_of(s.loc), s.idx[0], s.idx[1], s.idx[2]), sg > 0 ? OP_COPY : OP_NEG, 1,
                                     stride));
                }
                else
                {
                    // Eface: component c is the candidate of the edge along axis c.
                    // The edge EMF reads component c != a of a face with normal a
                    // only one layer outside the block along the third axis t,
                    // every other index interior; no other ghost value of this
                    // step scratch is read, so only those are exchanged.
                    const int a = loc - LOC_FACE_XI;
                    for (int c = 0; c < 3; ++c)
                    {
                        if (c == a)
                            continue;
                        const int t = 3 - a - c;
                        if (idx[a] < 0 || idx[a] >= cnt[a] || idx[c] < 0 || idx[c] >= cnt[c] ||
                            (idx[t] != -1 && idx[t] != cnt[t]))
                            continue;
                        const int64_t d = dst + (int64_t)c * stride;
                        const int64_t sr = C.ref(s.block, slot_of(s.loc) + s.perm[c], s.idx[0], s.idx[1], s.idx[2]);
                        out.push_back(mk(d, sr, s.sign[c] > 0 ? OP_COPY : OP_NEG, 1, stride));
                    }
                }
            }
}

// Tangential-interior x normal box of a physical face for location `loc`
// (reference BoundaryCore::MakeInnerSlabBox_OneLayer, src/6_boundary/Boundary_Tools.cpp:93-196).
struct Slab
{
    int lo[3], hi[3];
    int ax, sgn; // normal axis, +1 high face / -1 low face
    int ref;     // normal index of the inner layer
};
Slab inner_slab(const Block &B, int loc, int face)
{
    Slab s;
    s.ax = face / 2;
    s.sgn = (face % 2) ? 1 : -1;
    for (int a = 0; a < 3; ++a)
    {
        s.lo[a] = 0;
        s.hi[a] = loc_count(loc, a, B.n(a));
    }
    const int hi_in = loc_count(loc, s.ax, B.n(s.ax));
    s.ref = s.sgn < 0 ? 0 : hi_in - 1;
    s.lo[s.ax] = s.ref;
    s.hi[s.ax] = s.ref + 1;
    return s;
}
} // namespace

void compact_ops(const std::vector<BlockArrays> &blk, std::vector<MhdOp> &ops)
{
    if (ops.size() < 2 || blk.empty())
        return;
    const int64_t bs = (int64_t)st_nslot(blk[0].d.model) * blk[0].d.slot, sl = blk[0].d.slot;
    struct Item
    {
        MhdOp op;
        int64_t grp;  // block * slots + slot of the destination
        int c[3];     // (i, j, k) box coordinates of the destination
    };
    std::vector<Item> left;
    left.reserve(ops.size());
    for (const MhdOp &o : ops)
    {
        Item it;
        it.op = o;
        const int64_t b = o.dst / bs, r = o.dst % bs, s = r / sl, p = r % sl;
        const MhdBlockDev &d = blk[b].d;
        it.grp = b * (bs / sl + 1) + s;
        it.c[0] = (int)(p / ((int64_t)d.sj * d.sk));
        it.c[1] = (int)((p / d.sk) % d.sj);
        it.c[2] = (int)(p % d.sk);
        left.push_back(it);
    }
    auto same = [](const MhdOp &a, const MhdOp &b) {
        return a.kind == b.kind && a.ncomp == b.ncomp && a.stride == b.stride && a.cas == b.cas && a.run == 1 &&
               b.run == 1 && a.fold == b.fold;
    };
    const bool uses_src[9] = {true, true, false, false, true, true, false, false, false};
    std::vector<MhdOp> out;
    out.reserve(ops.size() / 8 + 16);
    const int order[3][3] = {{0, 1, 2}, {0, 2, 1}, {1, 2, 0}}; // slow .. fast axes: runs along k, j, i
    for (int pass = 0; pass < 3 && !left.empty(); ++pass)
    {
        const int *ax = order[pass];
        std::sort(left.begin(), left.end(), [&](const Item &a, const Item &b) {
            if (a.op.kind != b.op.kind) return a.op.kind < b.op.kind;
            if (a.op.ncomp != b.op.ncomp) return a.op.ncomp < b.op.ncomp;
            if (a.op.cas != b.op.cas) return a.op.cas < b.op.cas;
            if (a.op.fold != b.op.fold) return a.op.fold < b.op.fold;
            if (a.grp != b.grp) return a.grp < b.grp;
            for (int q = 0; q < 3; ++q)
                if (a.c[ax[q]] != b.c[ax[q]]) return a.c[ax[q]] < b.c[ax[q]];
            return false;
        });
        std::vector<Item> rest;
        const size_t min_run = pass < 2 ? 8 : 2; // short runs are retried along the next axis
        size_t q = 0;
        while (q < left.size())
        {
            const MhdOp &o = left[q].op;
            size_t e = q + 1;
            const bool src = o.kind >= 0 && o.kind < 9 && uses_src[o.kind];
            int64_t ds = 0, ss = 0;
            if (o.kind != OP_CONST && e < left.size() && same(o, left[e].op) && left[e].grp == left[q].grp &&
                left[e].c[ax[0]] == left[q].c[ax[0]] && left[e].c[ax[1]] == left[q].c[ax[1]])
            {
                ds = left[e].op.dst - o.dst;
                ss = src ? left[e].op.src - o.src : ds;
                while (e < left.size() && same(o, left[e].op) && left[e].grp == left[q].grp &&
                       left[e].c[ax[0]] == left[q].c[ax[0]] && left[e].c[ax[1]] == left[q].c[ax[1]] &&
                       left[e].op.dst - left[e - 1].op.dst == ds &&
                       (!src || left[e].op.src - left[e - 1].op.src == ss) && e - q < (1u << 30))
                    ++e;
            }
            const size_t n = e - q;
            if (n >= min_run && ds != 0 && ds == (int32_t)ds && ss == (int32_t)ss)
            {
                MhdOp r = o;
                r.run = (int32_t)n;
                r.dstep = (int32_t)ds;
                r.sstep = (int32_t)(src ? ss : ds);
                out.push_back(r);
            }
            else if (pass < 2)
                rest.insert(rest.end(), left.begin() + q, left.begin() + e);
            else
                for (size_t t = q; t < e; ++t)
                    out.push_back(left[t].op);
            q = e;
        }
        left.swap(rest);
    }
    for (const Item &it : left)
        out.push_back(it.op);
    ops.swap(out);
}

void build_ops(CaseSetup &cs)
{
    const Grid &grd = cs.grid;
    const int g = grd.g;
    // The reference's absorbing wall writes its first cell layer from the
    // second and then its ghosts from the first, face after face
    // (Z4_Lunar/src/boundary/Boundary_Condition_Fluid.cpp:20-184); the op
    // lists apply every first layer, then every ghost, each in any order.
    // The two agree unless walls on two axes of one block share an edge (the
    // later wall rewrites a cell the earlier one's ghosts were taken from),
    // so such a block is refused.
    for (size_t b = 0; b < grd.blocks.size(); ++b)
    {
        int walls = 0, first = -1;
        for (int ax = 0; ax < 3; ++ax)
            if ((grd.blocks[b].face[2 * ax].kind == 0 && grd.blocks[b].face[2 * ax].bc == BC_SOLID) ||
                (grd.blocks[b].face[2 * ax + 1].kind == 0 && grd.blocks[b].face[2 * ax + 1].bc == BC_SOLID))
            {
                if (walls++ == 0)
                    first = ax;
                else
                    throw std::invalid_argument("block " + std::to_string(b) + ": solid walls on axes " +
                                                std::to_string(first) + " and " + std::to_string(ax) +
                                                " meet along an edge (the wall handler's face-by-face order "
                                                "there is not supported)");
            }
    }
    OpCtx C{grd, cs.blk};
    cs.ops.assign(SY_COUNT, {});
    const int nb = (int)grd.blocks.size();

    // ops of one face, marked when the step's kernels fold them
    auto push = [](std::vector<MhdOp> &L, MhdOp o, bool fold) {
        o.fold = fold ? 1 : 0;
        L.push_back(o);
    };
    for (int b = 0; b < nb; ++b)
    {
        const Block &B = grd.blocks[b];
        const int64_t stride = cs.blk[b].d.slot;
        const MhdBlockDev &D = cs.blk[b].d;
        for (int f = 0; f < 6; ++f)
        {
            if (B.face[f].kind != 0 || B.face[f].bc == BC_NONE)
                continue;
            const int bc = B.face[f].bc;
            const bool fold_b = folds_kface_b(D, f), fold_u = folds_kface_u(D, f);
            // ---- Eface: solid wall zeroes the inner slab and its ghosts
            //      (reference Z4_Lunar/src/boundary/Boundary_Condition_EM.cpp:787-803)
            if (bc == BC_SOLID)
                for (int a = 0; a < 3; ++a)
                {
                    const int loc = LOC_FACE_XI + a;
                    Slab s = inner_slab(B, loc, f);
                    for (int i = s.lo[0]; i < s.hi[0]; ++i)
                        for (int j = s.lo[1]; j < s.hi[1]; ++j)
                            for (int k = s.lo[2]; k < s.hi[2]; ++k)
                                for (int l = 0; l <= g; ++l)
                                {
                                    int q[3] = {i, j, k};
                                    q[s.ax] += s.sgn * l;
                                    cs.ops[SY_EFACE_BC].push_back(mk(C.ref(b, ST_EFACE + 3 * a, q[0], q[1], q[2]), 0, OP_ZERO, 3, stride));
                                }
                }
            // ---- E edges: far-field sets -(u x B)_inf . dr on the inner slab
            //      (reference Z4_Lunar/src/boundary/Boundary_Condition.cpp:573-623)
            if (bc == BC_FARFIELD)
            {
                const double *ui = cs.cst.u_inf, *Bi = cs.cst.B_imf;
                const double Einf[3] = {Bi[1] * ui[2] - Bi[2] * ui[1], Bi[2] * ui[0] - Bi[0] * ui[2],
                                        Bi[0] * ui[1] - Bi[1] * ui[0]};
                for (int a = 0; a < 3; ++a)
                {
                    const int loc = LOC_EDGE_XI + a;
                    Slab s = inner_slab(B, loc, f);
                    for (int i = s.lo[0]; i < s.hi[0]; ++i)
                        for (int j = s.lo[1]; j < s.hi[1]; ++j)
                            for (int k = s.lo[2]; k < s.hi[2]; ++k)
                            {
                                const int64_t o0 = B.node_off(i, j, k);
                                const int64_t o1 = B.node_off(i + (a == 0), j + (a == 1), k + (a == 2));
                                const double val = Einf[0] * (B.x[o1] - B.x[o0]) + Einf[1] * (B.y[o1] - B.y[o0]) +
                                                   Einf[2] * (B.z[o1] - B.z[o0]);
                                cs.ops[SY_EEDGE_BC].push_back(mk(C.ref(b, ST_E + a, i, j, k), 0, OP_CONST, 1, stride, val));
                            }
                }
            }
            // ---- U: far-field constant ghosts; absorbing wall
            //      (reference Z4_Lunar/src/boundary/Boundary_Condition_Fluid.cpp:3-184)
            {
                Slab s = inner_slab(B, LOC_CELL, f);
                for (int i = s.lo[0]; i < s.hi[0]; ++i)
                    for (int j = s.lo[1]; j < s.hi[1]; ++j)
                        for (int k = s.lo[2]; k < s.hi[2]; ++k)
                        {
                            if (bc == BC_FARFIELD)
                            {
                                for (int l = 1; l <= g; ++l)
                                {
                                    int q[3] = {i, j, k};
                                    q[s.ax] += s.sgn * l;
                                    push(cs.ops[SY_U_BC0], mk(C.ref(b, ST_U0, q[0], q[1], q[2]), 0, OP_QINF, 5, stride), fold_u);
                                }
                            }
                            else
                            {
                                int r[3] = {i, j, k};
                                r[s.ax] -= s.sgn; // second interior layer
                                push(cs.ops[SY_U_BC0],
                                     mk(C.ref(b, ST_U0, i, j, k), C.ref(b, ST_U0, r[0], r[1], r[2]), OP_SOLID_FIRST, 5, stride),
                                     fold_u);
                                for (int l = 1; l <= g; ++l)
                                {
                                    int q[3] = {i, j, k};
                                    q[s.ax] += s.sgn * l;
                                    push(cs.ops[SY_U_BC1],
                                         mk(C.ref(b, ST_U0, q[0], q[1], q[2]), C.ref(b, ST_U0, i, j, k), OP_SOLID_GHOST, 5, stride),
                                         fold_u);
                                }
                            }
                        }
            }
            // ---- B faces: solid copies the inner layer into the ghosts,
            //      far-field leaves ghosts untouched (reference Boundary_Build.cpp:29-33,120-122)
            if (bc == BC_SOLID)
                for (int a = 0; a < 3; ++a)
                {
                    const int loc = LOC_FACE_XI + a;
                    Slab s = inner_slab(B, loc, f);
                    for (int i = s.lo[0]; i < s.hi[0]; ++i)
                        for (int j = s.lo[1]; j < s.hi[1]; ++j)
                            for (int k = s.lo[2]; k < s.hi[2]; ++k)
                                for (int l = 1; l <= g; ++l)
                                {
                                    int q[3] = {i, j, k};
                                    q[s.ax] += s.sgn * l;
                                    push(cs.ops[SY_B_BC],
                                         mk(C.ref(b, ST_B + a, q[0], q[1], q[2]), C.ref(b, ST_B + a, i, j, k), OP_COPY, 1, stride),
                                         fold_b);
                                }
                }
            // ---- J_cell: solid wall zeroes two interior layers
            //      (reference Z4_Lunar/src/boundary/Boundary_Condition_EM.cpp:382-427)
            if (bc == BC_SOLID)
            {
                Slab s = inner_slab(B, LOC_CELL, f);
                for (int i = s.lo[0]; i < s.hi[0]; ++i)
                    for (int j = s.lo[1]; j < s.hi[1]; ++j)
                        for (int k = s.lo[2]; k < s.hi[2]; ++k)
                            for (int l = 0; l <= 1; ++l)
                            {
                                int q[3] = {i, j, k};
                                q[s.ax] -= s.sgn * l;
                                cs.ops[SY_JCELL_BC].push_back(mk(C.ref(b, ST_JCELL, q[0], q[1], q[2]), 0, OP_ZERO, 3, stride));
                            }
            }
        }
        // ---- corner-diagonal ghost cells (outside two non-periodic
        //      interfaces): contravariant velocity of the source cell
        for (int i = -1; i <= B.ni; ++i)
            for (int j = -1; j <= B.nj; ++j)
                for (int k = -1; k <= B.nk; ++k)
                {
                    const int idx[3] = {i, j, k};
                    int nif = 0;
                    for (int a = 0; a < 3; ++a)
                    {
                        const int side = idx[a] < 0 ? 0 : (idx[a] >= B.n(a) ? 1 : -1);
                        if (side >= 0 && B.face[2 * a + side].kind == 1 && !B.face[2 * a + side].periodic)
                            ++nif;
                    }
                    if (nif < 2)
                        continue;
                    const DofSource s = resolve_dof(grd, b, LOC_CELL, idx);
                    for (int a = 0; a < 3; ++a)
                        cs.ops[SY_UC_CORNER].push_back(mk(C.ref(b, ST_UC + a, i, j, k),
                                                          C.ref(s.block, ST_UC + s.perm[a], s.idx[0], s.idx[1], s.idx[2]),
                                                          s.sign[a] > 0 ? OP_COPY : OP_NEG, 1, stride));
                }
        // ---- Na+ of the two-fluid step: far-field seed state, absorbing
        //      wall ghosts at the floor state (reference
        //      Z5_Mercury/src/boundary/Boundary_Condition_Fluid.cpp:19-165)
        if (cs.blk[b].d.model == 1)
            for (int f = 0; f < 6; ++f)
            {
                if (B.face[f].kind != 0 || B.face[f].bc == BC_NONE)
                    continue;
                Slab s = inner_slab(B, LOC_CELL, f);
                for (int i = s.lo[0]; i < s.hi[0]; ++i)
                    for (int j = s.lo[1]; j < s.hi[1]; ++j)
                        for (int k = s.lo[2]; k < s.hi[2]; ++k)
                            for (int l = 1; l <= g; ++l)
                            {
                                int q[3] = {i, j, k};
                                q[s.ax] += s.sgn * l;
                                cs.ops[SY_U_BC0].push_back(mk(C.ref(b, ST_UNA0, q[0], q[1], q[2]), 0,
                                                              B.face[f].bc == BC_FARFIELD ? OP_QINF_NA : OP_NA_FLOOR, 5,
                                                              stride));
                            }
            }
        // ---- B_cell and J_cell ghosts, read by the Hall EMF and the interior
        //      resistivity of the two-fluid step: far-field B_cell = B_imf, wall B_cell and both J_cell
        //      copies of the first layer (reference Z5_Mercury/src/boundary/
        //      Boundary_Build.cpp:112-124, Boundary_Condition.cpp:625-657,
        //      Boundary_Condition_EM.cpp:118-163), then the halos
        const bool hall_ghosts = cs.blk[b].d.model == 1 && (cs.cst.hall != 0.0 || cs.cst.resist != 0.0);
        if (hall_ghosts)
        {
            for (int f = 0; f < 6; ++f)
            {
                if (B.face[f].kind != 0 || B.face[f].bc == BC_NONE)
                    continue;
                Slab s = inner_slab(B, LOC_CELL, f);
                for (int i = s.lo[0]; i < s.hi[0]; ++i)
                    for (int j = s.lo[1]; j < s.hi[1]; ++j)
                        for (int k = s.lo[2]; k < s.hi[2]; ++k)
                            for (int l = 1; l <= g; ++l)
                            {
                                int q[3] = {i, j, k};
                                q[s.ax] += s.sgn * l;
                                if (B.face[f].bc == BC_FARFIELD)
                                    for (int m = 0; m < 3; ++m)
                                        cs.ops[SY_BCELL_BC].push_back(mk(C.ref(b, ST_BCELL + m, q[0], q[1], q[2]), 0,
                                                                         OP_CONST, 1, stride, cs.cst.B_imf[m]));
                                else
                                    cs.ops[SY_BCELL_BC].push_back(mk(C.ref(b, ST_BCELL, q[0], q[1], q[2]),
                                                                     C.ref(b, ST_BCELL, i, j, k), OP_COPY, 3, stride));
                                cs.ops[SY_JCELL_GH].push_back(mk(C.ref(b, ST_JCELL, q[0], q[1], q[2]),
                                                                 C.ref(b, ST_JCELL, i, j, k), OP_COPY, 3, stride));
                            }
            }
            halo_ops(C, b, LOC_CELL, 3, false, [](int) { return (int)ST_BCELL; }, cs.ops[SY_BCELL_HALO]);
            halo_ops(C, b, LOC_CELL, 3, false, [](int) { return (int)ST_JCELL; }, cs.ops[SY_JCELL_HALO]);
        }
        // ---- halos
        halo_ops(C, b, LOC_CELL, 5, false, [](int) { return (int)ST_U0; }, cs.ops[SY_U_HALO]);
        if (cs.blk[b].d.model == 1)
            halo_ops(C, b, LOC_CELL, 5, false, [](int) { return (int)ST_UNA0; }, cs.ops[SY_U_HALO]);
        for (int a = 0; a < 3; ++a)
        {
            halo_ops(C, b, LOC_FACE_XI + a, 1, false, [](int l) { return (int)ST_B + (l - LOC_FACE_XI); }, cs.ops[SY_B_HALO]);
            halo_ops(C, b, LOC_FACE_XI + a, 3, true, [](int l) { return (int)ST_EFACE + 3 * (l - LOC_FACE_XI); },
                     cs.ops[SY_EFACE_HALO]);
        }
    }
    for (std::vector<MhdOp> &L : cs.ops)
        compact_ops(cs.blk, L);
}

} // namespace mpcns
