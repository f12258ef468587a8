/* CPU FP64 oracle of the explicit MHD step -- TEST INFRASTRUCTURE ONLY.
 * See mhd_oracle.h.  Every routine cites the reference file:line it
 * restates.  Loop order, accumulation order and floors follow the
 * reference so the FP64 GPU path can be compared bit for bit.
 * Build with -ffp-contract=off.
 */
#include "mhd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* std::max / std::min semantics (first argument wins ties) */
/* Per-cell loops over the i planes of a block run on c->nthreads OpenMP
 * threads; every iteration writes its own cells only, so the result does not
 * depend on the thread count. */
#define OMP_PRAGMA(x) _Pragma(#x)
#define OMP_FOR OMP_PRAGMA(omp parallel for schedule(static) num_threads(c->nthreads) if (c->nthreads > 1))

static inline double dmax(double a, double b) { return (a < b) ? b : a; }
static inline double dmin(double a, double b) { return (b < a) ? b : a; }

enum { L_CELL = 0, L_NODE, L_FXI, L_FET, L_FZE, L_EXI, L_EET, L_EZE };

static int delta_of(int loc, int ax)
{
    if (loc == L_CELL) return 1;
    if (loc == L_NODE) return 0;
    if (loc >= L_FXI && loc <= L_FZE) return ax == loc - L_FXI ? 0 : 1;
    return ax == loc - L_EXI ? 1 : 0;
}

typedef struct
{
    double *d;
    int loc, nc, g;
    int cnt[3]; /* interior counts */
    int ext[3]; /* cnt + 2g */
} fblk;

#define AT(f, i, j, k, m) \
    ((f)->d[((((size_t)((i) + (f)->g)) * (f)->ext[1] + (size_t)((j) + (f)->g)) * (f)->ext[2] + (size_t)((k) + (f)->g)) * (f)->nc + (m)])

typedef struct
{
    int n[3];
    or_face face[6];
    or_face_part *parts[6]; /* faces several interface patches share */
    int nparts[6];
    fblk f[OF_COUNT];
    double *node[3];
    int gn;
} oblock;

typedef struct
{
    int block, loc, idx[3], perm[3], sign[3], kind;
} osrc;

typedef struct
{
    int dst_idx[3];
    osrc s;
} hentry;

typedef struct
{
    hentry *e;
    long n;
    int built;
} hmap;

struct or_ctx
{
    or_desc d;
    oblock *b;
    double dt;
    int wrap;
    int nthreads; /* OpenMP threads of the per-cell loops (1: serial) */
    hmap hm[8];   /* per location */
};

static const struct
{
    int loc, nc, g0;
} kField[OF_COUNT] = {
    {L_CELL, 5, 0}, {L_CELL, 5, 0}, {L_CELL, 3, 0}, {L_FXI, 1, 0}, {L_FET, 1, 0}, {L_FZE, 1, 0},
    {L_FXI, 1, 0}, {L_FET, 1, 0}, {L_FZE, 1, 0}, {L_CELL, 3, 0}, {L_CELL, 3, 0}, {L_EXI, 1, 0},
    {L_EET, 1, 0}, {L_EZE, 1, 0}, {L_CELL, 3, 0}, {L_EXI, 1, 0}, {L_EET, 1, 0}, {L_EZE, 1, 0},
    {L_FXI, 3, 0}, {L_FET, 3, 0}, {L_FZE, 3, 0}, {L_FXI, 5, 1}, {L_FET, 5, 1}, {L_FZE, 5, 1},
    {L_CELL, 5, 1}, {L_FXI, 1, 1}, {L_FET, 1, 1}, {L_FZE, 1, 1}, {L_CELL, 1, 0}, {L_CELL, 1, 0},
    {L_FXI, 3, 0}, {L_FET, 3, 0}, {L_FZE, 3, 0}, {L_FXI, 1, 0}, {L_FET, 1, 0}, {L_FZE, 1, 0},
    {L_EXI, 1, 0}, {L_EET, 1, 0}, {L_EZE, 1, 0}, {L_CELL, 18, 0}, {L_CELL, 36, 0},
    {L_CELL, 5, 0}, {L_CELL, 5, 0}, {L_CELL, 5, 1}, {L_CELL, 1, 0}, {L_CELL, 1, 0},
    {L_EXI, 1, 0}, {L_EET, 1, 0}, {L_EZE, 1, 0},
    {L_EXI, 1, 0}, {L_EET, 1, 0}, {L_EZE, 1, 0}, {L_FXI, 1, 1}, {L_FET, 1, 1}, {L_FZE, 1, 1},
};
/* g0 = 1 marks fields the reference registers with zero ghosts
   (Z4_Lunar/main.cpp:107-121); their interior range is what matters. */

or_ctx *or_create(const or_desc *d)
{
    or_ctx *c = (or_ctx *)calloc(1, sizeof(or_ctx));
    c->d = *d;
    c->nthreads = 1;
    c->b = (oblock *)calloc((size_t)d->nblocks, sizeof(oblock));
    return c;
}

void or_destroy(or_ctx *c)
{
    if (!c) return;
    for (int b = 0; b < c->d.nblocks; ++b)
    {
        for (int f = 0; f < OF_COUNT; ++f) free(c->b[b].f[f].d);
        for (int q = 0; q < 3; ++q) free(c->b[b].node[q]);
        for (int f = 0; f < 6; ++f) free(c->b[b].parts[f]);
    }
    for (int l = 0; l < 8; ++l) free(c->hm[l].e);
    free(c->b);
    free(c);
}

int or_set_block(or_ctx *c, int b, int ni, int nj, int nk, const or_face *faces)
{
    if (b < 0 || b >= c->d.nblocks) return -1;
    oblock *B = &c->b[b];
    B->n[0] = ni; B->n[1] = nj; B->n[2] = nk;
    memcpy(B->face, faces, 6 * sizeof(or_face));
    for (int f = 0; f < OF_COUNT; ++f)
    {
        fblk *F = &B->f[f];
        F->loc = kField[f].loc;
        F->nc = kField[f].nc;
        F->g = kField[f].g0 ? 0 : c->d.g;
        size_t tot = (size_t)F->nc;
        for (int a = 0; a < 3; ++a)
        {
            F->cnt[a] = B->n[a] + 1 - delta_of(F->loc, a);
            F->ext[a] = F->cnt[a] + 2 * F->g;
            tot *= (size_t)F->ext[a];
        }
        free(F->d);
        F->d = (double *)calloc(tot, sizeof(double));
    }
    B->gn = c->d.g + 1;
    for (int q = 0; q < 3; ++q)
    {
        size_t tot = 1;
        for (int a = 0; a < 3; ++a) tot *= (size_t)(B->n[a] + 1 + 2 * B->gn);
        free(B->node[q]);
        B->node[q] = (double *)calloc(tot, sizeof(double));
    }
    return 0;
}

int or_set_face_parts(or_ctx *c, int b, int f, int n, const or_face_part *parts)
{
    if (b < 0 || b >= c->d.nblocks || f < 0 || f > 5 || n < 0) return -1;
    oblock *B = &c->b[b];
    free(B->parts[f]);
    B->parts[f] = NULL;
    B->nparts[f] = n;
    if (n > 0)
    {
        B->parts[f] = (or_face_part *)malloc((size_t)n * sizeof(or_face_part));
        memcpy(B->parts[f], parts, (size_t)n * sizeof(or_face_part));
    }
    return 0;
}

double *or_field(or_ctx *c, int b, int fid, long *count)
{
    fblk *F = &c->b[b].f[fid];
    if (count) *count = (long)F->nc * F->ext[0] * F->ext[1] * F->ext[2];
    return F->d;
}

double *or_nodes(or_ctx *c, int b, int comp, long *count)
{
    oblock *B = &c->b[b];
    if (count) *count = (long)(B->n[0] + 1 + 2 * B->gn) * (B->n[1] + 1 + 2 * B->gn) * (B->n[2] + 1 + 2 * B->gn);
    return B->node[comp];
}

void or_use_periodic_wrap(or_ctx *c, int on) { c->wrap = on; }
void or_set_threads(or_ctx *c, int n) { c->nthreads = n > 1 ? n : 1; }

static inline double NODE(const oblock *B, int q, int i, int j, int k)
{
    const int gn = B->gn;
    const size_t sj = (size_t)(B->n[1] + 1 + 2 * gn), sk = (size_t)(B->n[2] + 1 + 2 * gn);
    return B->node[q][((size_t)(i + gn) * sj + (size_t)(j + gn)) * sk + (size_t)(k + gn)];
}

/* ------------------------------------------------------------------ halo
 * Ghost DOFs are filled from the DOF they coincide with: each out-of-range
 * axis that faces an interface is crossed through its IndexTransform
 * (lowest axis first), which is the reference's copy
 * (src/4_halo/1_MPCNS_Halo_Inner.cpp:60-69) for translation interfaces and
 * extends it with the staggered-index shift and component orientation for
 * rotated ones.  What is left out of range on physical axes is a physical
 * ghost of the block reached (one axis) or a corner filled by the clamped
 * copy of ApplyPhysicalCornerDefault (src/6_boundary/Boundary.cpp:178-260). */
static int map_loc(int loc, const int *perm)
{
    if (loc >= L_FXI && loc <= L_FZE) return L_FXI + perm[loc - L_FXI];
    if (loc >= L_EXI && loc <= L_EZE) return L_EXI + perm[loc - L_EXI];
    return loc;
}

/* the link a ghost crosses through interface face f: the face's own, or
   the patch covering its doubled node coordinates dc (clamped onto the
   face) on a face several patches share -- the reference copies each inner
   patch's range (src/4_halo/1_MPCNS_Halo_Inner.cpp:60-69) */
static const or_face *face_link(const oblock *B, int f, const int dc[3])
{
    if (B->nparts[f] == 0) return &B->face[f];
    const int ax = f / 2, t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
    int c1 = dc[t1], c2 = dc[t2];
    if (c1 < 0) c1 = 0;
    if (c1 > 2 * B->n[t1]) c1 = 2 * B->n[t1];
    if (c2 < 0) c2 = 0;
    if (c2 > 2 * B->n[t2]) c2 = 2 * B->n[t2];
    for (int q = 0; q < B->nparts[f]; ++q)
    {
        const or_face_part *P = &B->parts[f][q];
        if (2 * P->lo[0] <= c1 && c1 <= 2 * P->hi[0] && 2 * P->lo[1] <= c2 && c2 <= 2 * P->hi[1])
            return &P->link;
    }
    return &B->face[f];
}

static osrc resolve(const or_ctx *c, int b, int loc, const int idx[3])
{
    osrc r;
    r.block = b; r.loc = loc;
    for (int a = 0; a < 3; ++a) { r.idx[a] = idx[a]; r.perm[a] = a; r.sign[a] = 1; }
    int hops = 0;
    for (;;)
    {
        const oblock *B = &c->b[r.block];
        int nout = 0, hax = -1, hside = 0;
        for (int a = 0; a < 3; ++a)
        {
            const int cnt = B->n[a] + 1 - delta_of(r.loc, a);
            int side = -1;
            if (r.idx[a] < 0) side = 0;
            else if (r.idx[a] >= cnt) side = 1;
            if (side < 0) continue;
            ++nout;
            if (hax < 0 && B->face[2 * a + side].kind == 1) { hax = a; hside = side; }
        }
        if (hax >= 0 && hops < 8)
        {
            int dc[3], nd[3];
            for (int a = 0; a < 3; ++a) dc[a] = 2 * r.idx[a] + delta_of(r.loc, a);
            const or_face *F = face_link(B, 2 * hax + hside, dc);
            for (int a = 0; a < 3; ++a) nd[F->perm[a]] = F->sign[a] * dc[a] + 2 * F->offset[a];
            const int nl = map_loc(r.loc, F->perm);
            for (int a = 0; a < 3; ++a) r.idx[a] = (nd[a] - delta_of(nl, a)) / 2;
            r.loc = nl;
            for (int a = 0; a < 3; ++a)
            {
                const int pa = r.perm[a];
                r.sign[a] *= F->sign[pa];
                r.perm[a] = F->perm[pa];
            }
            r.block = F->nb_block;
            ++hops;
            continue;
        }
        if (nout == 0) r.kind = 0;
        else if (nout == 1) r.kind = hops == 0 ? 3 : 1;
        else
        {
            r.kind = 2;
            for (int a = 0; a < 3; ++a)
            {
                const int cnt = B->n[a] + 1 - delta_of(r.loc, a);
                if (r.idx[a] < 0) r.idx[a] = 0;
                if (r.idx[a] > cnt - 1) r.idx[a] = cnt - 1;
            }
        }
        return r;
    }
}

static void build_hmap(or_ctx *c, int loc)
{
    hmap *H = &c->hm[loc];
    if (H->built) return;
    const int g = c->d.g;
    long cap = 1024, n = 0;
    hentry *e = (hentry *)malloc((size_t)cap * sizeof(hentry));
    for (int b = 0; b < c->d.nblocks; ++b)
    {
        const oblock *B = &c->b[b];
        int cnt[3];
        for (int a = 0; a < 3; ++a) cnt[a] = B->n[a] + 1 - delta_of(loc, a);
        for (int i = -g; i < cnt[0] + g; ++i)
            for (int j = -g; j < cnt[1] + g; ++j)
                for (int k = -g; k < cnt[2] + g; ++k)
                {
                    const int idx[3] = {i, j, k};
                    int ghost = 0;
                    for (int a = 0; a < 3; ++a) ghost |= (idx[a] < 0 || idx[a] >= cnt[a]);
                    if (!ghost) continue;
                    osrc s = resolve(c, b, loc, idx);
                    if (s.kind == 3) continue;
                    if (n == cap) { cap *= 2; e = (hentry *)realloc(e, (size_t)cap * sizeof(hentry)); }
                    e[n].dst_idx[0] = i; e[n].dst_idx[1] = j; e[n].dst_idx[2] = k;
                    e[n].s = s;
                    /* the destination block is implicit in the loop; store it */
                    e[n].s.kind = s.kind | (b << 4);
                    ++n;
                }
    }
    H->e = e; H->n = n; H->built = 1;
}

/* fids: one field per location slot of the group; for face/edge groups
   fid0 is the xi member and the others follow (xi, eta, zeta). */
static void halo_apply(or_ctx *c, int loc, int fid_base, int is_triple, int eface)
{
    const int g = c->d.g;
    if (c->wrap)
    {
        /* single fully periodic block: ghost index wraps modulo N */
        oblock *B = &c->b[0];
        const int fid = is_triple ? fid_base + (loc >= L_EXI ? loc - L_EXI : loc - L_FXI) : fid_base;
        fblk *F = &B->f[fid];
        for (int i = -g; i < F->cnt[0] + g; ++i)
            for (int j = -g; j < F->cnt[1] + g; ++j)
                for (int k = -g; k < F->cnt[2] + g; ++k)
                {
                    const int idx[3] = {i, j, k};
                    int s[3], ghost = 0;
                    for (int a = 0; a < 3; ++a)
                    {
                        const int N = B->n[a];
                        ghost |= (idx[a] < 0 || idx[a] >= F->cnt[a]);
                        s[a] = ((idx[a] % N) + N) % N;
                        if (idx[a] >= 0 && idx[a] < F->cnt[a]) s[a] = idx[a];
                    }
                    if (!ghost) continue;
                    for (int m = 0; m < F->nc; ++m) AT(F, i, j, k, m) = AT(F, s[0], s[1], s[2], m);
                }
        return;
    }
    build_hmap(c, loc);
    const hmap *H = &c->hm[loc];
    /* each entry writes its own ghost; sources are interior DOFs or ghosts
       written by the physical phase, none written in this pass */
    OMP_FOR
    for (long q = 0; q < H->n; ++q)
    {
        const hentry *E = &H->e[q];
        const int db = E->s.kind >> 4;
        const osrc *s = &E->s;
        int dfid = fid_base, sfid = fid_base;
        if (is_triple)
        {
            dfid = fid_base + (loc >= L_EXI ? loc - L_EXI : loc - L_FXI);
            sfid = fid_base + (s->loc >= L_EXI ? s->loc - L_EXI : s->loc - L_FXI);
        }
        fblk *D = &c->b[db].f[dfid];
        fblk *S = &c->b[s->block].f[sfid];
        const int *di = E->dst_idx;
        if (loc == L_CELL)
        {
            for (int m = 0; m < D->nc; ++m) AT(D, di[0], di[1], di[2], m) = AT(S, s->idx[0], s->idx[1], s->idx[2], m);
        }
        else if (!eface)
        {
            const int ax = loc >= L_EXI ? loc - L_EXI : loc - L_FXI;
            const double v = AT(S, s->idx[0], s->idx[1], s->idx[2], 0);
            AT(D, di[0], di[1], di[2], 0) = s->sign[ax] > 0 ? v : -v;
        }
        else
        {
            for (int m = 0; m < 3; ++m)
            {
                const double v = AT(S, s->idx[0], s->idx[1], s->idx[2], s->perm[m]);
                AT(D, di[0], di[1], di[2], m) = s->sign[m] > 0 ? v : -v;
            }
        }
    }
}

/* --------------------------------------------------- physical handlers
 * Inner slab (src/6_boundary/Boundary_Tools.cpp:93-196) and ghost slab
 * (:198-247) of one physical face for a field. */
typedef struct { int lo[3], hi[3], ax, sgn, ref; } slab;

static slab inner_slab(const fblk *F, int face)
{
    slab s;
    s.ax = face / 2;
    s.sgn = (face % 2) ? 1 : -1;
    for (int a = 0; a < 3; ++a) { s.lo[a] = 0; s.hi[a] = F->cnt[a]; }
    s.ref = s.sgn < 0 ? 0 : F->cnt[s.ax] - 1;
    s.lo[s.ax] = s.ref;
    s.hi[s.ax] = s.ref + 1;
    return s;
}

/* BoundaryCore::DefaultPhysicalCopy (src/6_boundary/Boundary.cpp:303-331) */
static void bc_copy(fblk *F, int face, int ngh)
{
    slab s = inner_slab(F, face);
    for (int i = s.lo[0]; i < s.hi[0]; ++i)
        for (int j = s.lo[1]; j < s.hi[1]; ++j)
            for (int k = s.lo[2]; k < s.hi[2]; ++k)
                for (int l = 1; l <= ngh; ++l)
                {
                    int q[3] = {i, j, k};
                    q[s.ax] += s.sgn * l;
                    for (int m = 0; m < F->nc; ++m) AT(F, q[0], q[1], q[2], m) = AT(F, i, j, k, m);
                }
}

/* LunarBoundary::BC_UH_Farfield_H_ (Z4_Lunar/src/boundary/Boundary_Condition_Fluid.cpp:3-17) */
static void bc_u_farfield(const or_ctx *c, fblk *F, int face, int ngh)
{
    slab s = inner_slab(F, face);
    for (int i = s.lo[0]; i < s.hi[0]; ++i)
        for (int j = s.lo[1]; j < s.hi[1]; ++j)
            for (int k = s.lo[2]; k < s.hi[2]; ++k)
                for (int l = 1; l <= ngh; ++l)
                {
                    int q[3] = {i, j, k};
                    q[s.ax] += s.sgn * l;
                    for (int m = 0; m < 5; ++m) AT(F, q[0], q[1], q[2], m) = c->d.qinf[m];
                }
}

/* MercuryBoundary::BC_UH_Farfield_Na_ (Z5_Mercury/src/boundary/Boundary_Condition_Fluid.cpp:19-34) */
static void bc_una_farfield(const or_ctx *c, fblk *F, int face, int ngh)
{
    slab s = inner_slab(F, face);
    for (int i = s.lo[0]; i < s.hi[0]; ++i)
        for (int j = s.lo[1]; j < s.hi[1]; ++j)
            for (int k = s.lo[2]; k < s.hi[2]; ++k)
                for (int l = 1; l <= ngh; ++l)
                {
                    int q[3] = {i, j, k};
                    q[s.ax] += s.sgn * l;
                    for (int m = 0; m < 5; ++m) AT(F, q[0], q[1], q[2], m) = c->d.qinfs[m];
                }
}

/* MercuryBoundary::BC_Solid_Surface_ for Na+ (Z5_Mercury/src/boundary/
 * Boundary_Condition_Fluid.cpp:36-165, the is_Na branch): absorbing wall, the
 * first layer is left alone and the ghosts hold the floor state. */
static void bc_una_solid(const or_ctx *c, fblk *U, int face, int ngh)
{
    const double rho_floor = 1.0e-12, p_floor = 1.0e-12;
    slab s = inner_slab(U, face);
    for (int i = s.lo[0]; i < s.hi[0]; ++i)
        for (int j = s.lo[1]; j < s.hi[1]; ++j)
            for (int k = s.lo[2]; k < s.hi[2]; ++k)
                for (int l = 1; l <= ngh; ++l)
                {
                    int q[3] = {i, j, k};
                    q[s.ax] += s.sgn * l;
                    AT(U, q[0], q[1], q[2], 0) = rho_floor;
                    AT(U, q[0], q[1], q[2], 1) = 0.0;
                    AT(U, q[0], q[1], q[2], 2) = 0.0;
                    AT(U, q[0], q[1], q[2], 3) = 0.0;
                    AT(U, q[0], q[1], q[2], 4) = p_floor / (c->d.gamma - 1.0);
                }
}

/* LunarBoundary::BC_Solid_Surface_ (Z4_Lunar/src/boundary/Boundary_Condition_Fluid.cpp:19-184):
 * absorbing wall -- first interior layer takes rho and p of the second
 * layer with zero velocity, ghosts copy that layer (again with v = 0). */
static double solid_p(double gamma, double rho, double mx, double my, double mz, double E)
{
    rho = dmax(rho, 2.5e-3);
    const double ke = 0.5 * (mx * mx + my * my + mz * mz) / rho;
    double eint = E - ke;
    if (eint < 0.0) eint = 0.0;
    const double p = (gamma - 1.0) * eint;
    return dmax(p, 1e-12);
}

static void bc_u_solid(const or_ctx *c, fblk *U, int face, int ngh)
{
    const double gm = c->d.gamma;
    slab s = inner_slab(U, face);
    for (int i = s.lo[0]; i < s.hi[0]; ++i)
        for (int j = s.lo[1]; j < s.hi[1]; ++j)
            for (int k = s.lo[2]; k < s.hi[2]; ++k)
            {
                int r[3] = {i, j, k};
                r[s.ax] -= s.sgn;
                {
                    const double rho_r = dmax(AT(U, r[0], r[1], r[2], 0), 2.5e-3);
                    const double p_r = solid_p(gm, rho_r, AT(U, r[0], r[1], r[2], 1), AT(U, r[0], r[1], r[2], 2),
                                               AT(U, r[0], r[1], r[2], 3), AT(U, r[0], r[1], r[2], 4));
                    AT(U, i, j, k, 0) = rho_r;
                    AT(U, i, j, k, 1) = rho_r * 0.0;
                    AT(U, i, j, k, 2) = rho_r * 0.0;
                    AT(U, i, j, k, 3) = rho_r * 0.0;
                    const double ke_g = 0.5 * rho_r * (0.0 * 0.0 + 0.0 * 0.0 + 0.0 * 0.0);
                    AT(U, i, j, k, 4) = p_r / (gm - 1.0) + ke_g;
                }
                for (int l = 1; l <= ngh; ++l)
                {
                    const double rho_r = dmax(AT(U, i, j, k, 0), 2.5e-3);
                    const double p_r = solid_p(gm, rho_r, AT(U, i, j, k, 1), AT(U, i, j, k, 2), AT(U, i, j, k, 3),
                                               AT(U, i, j, k, 4));
                    int q[3] = {i, j, k};
                    q[s.ax] += s.sgn * l;
                    AT(U, q[0], q[1], q[2], 0) = rho_r;
                    AT(U, q[0], q[1], q[2], 1) = rho_r * 0.0;
                    AT(U, q[0], q[1], q[2], 2) = rho_r * 0.0;
                    AT(U, q[0], q[1], q[2], 3) = rho_r * 0.0;
                    const double ke_g = 0.5 * rho_r * (0.0 * 0.0 + 0.0 * 0.0 + 0.0 * 0.0);
                    AT(U, q[0], q[1], q[2], 4) = p_r / (gm - 1.0) + ke_g;
                }
            }
}

/* BC_Solid_Surface_Eface_ (Z4_Lunar/src/boundary/Boundary_Condition_EM.cpp:787-803) */
static void bc_eface_solid(fblk *F, int face, int ngh)
{
    slab s = inner_slab(F, face);
    for (int i = s.lo[0]; i < s.hi[0]; ++i)
        for (int j = s.lo[1]; j < s.hi[1]; ++j)
            for (int k = s.lo[2]; k < s.hi[2]; ++k)
                for (int m = 0; m < 3; ++m) AT(F, i, j, k, m) = 0.0;
    bc_copy(F, face, ngh);
}

/* BC_Solid_Surface_Jcell (Z4_Lunar/src/boundary/Boundary_Condition_EM.cpp:382-427) */
static void bc_jcell_solid(fblk *F, int face, int ngh)
{
    slab s = inner_slab(F, face);
    for (int i = s.lo[0]; i < s.hi[0]; ++i)
        for (int j = s.lo[1]; j < s.hi[1]; ++j)
            for (int k = s.lo[2]; k < s.hi[2]; ++k)
            {
                for (int m = 0; m < F->nc; ++m) AT(F, i, j, k, m) = 0.0;
                int q[3] = {i, j, k};
                q[s.ax] -= s.sgn;
                for (int m = 0; m < F->nc; ++m) AT(F, q[0], q[1], q[2], m) = 0.0;
            }
    bc_copy(F, face, ngh);
}

/* BC_Farfield_Bface for B_cell / Bind_cell (Z4_Lunar/src/boundary/Boundary_Condition.cpp:625-657) */
static void bc_bcell_farfield(fblk *F, int face, int ngh, const double *val)
{
    slab s = inner_slab(F, face);
    for (int i = s.lo[0]; i < s.hi[0]; ++i)
        for (int j = s.lo[1]; j < s.hi[1]; ++j)
            for (int k = s.lo[2]; k < s.hi[2]; ++k)
                for (int l = 1; l <= ngh; ++l)
                {
                    int q[3] = {i, j, k};
                    q[s.ax] += s.sgn * l;
                    for (int m = 0; m < 3; ++m) AT(F, q[0], q[1], q[2], m) = val[m];
                }
}

/* BC_Farfield_Eedge_set_zerocurl (Z4_Lunar/src/boundary/Boundary_Condition.cpp:573-623) */
static void bc_eedge_farfield(const or_ctx *c, const oblock *B, fblk *F, int face, int ngh, int dir)
{
    const double *ui = c->d.u_inf, *Bi = c->d.B_imf;
    double E[3];
    E[0] = Bi[1] * ui[2] - Bi[2] * ui[1];
    E[1] = Bi[2] * ui[0] - Bi[0] * ui[2];
    E[2] = Bi[0] * ui[1] - Bi[1] * ui[0];
    slab s = inner_slab(F, face);
    for (int i = s.lo[0]; i < s.hi[0]; ++i)
        for (int j = s.lo[1]; j < s.hi[1]; ++j)
            for (int k = s.lo[2]; k < s.hi[2]; ++k)
            {
                const int ir = i + (dir == 0), jr = j + (dir == 1), kr = k + (dir == 2);
                AT(F, i, j, k, 0) = E[0] * (NODE(B, 0, ir, jr, kr) - NODE(B, 0, i, j, k)) +
                                    E[1] * (NODE(B, 1, ir, jr, kr) - NODE(B, 1, i, j, k)) +
                                    E[2] * (NODE(B, 2, ir, jr, kr) - NODE(B, 2, i, j, k));
            }
    bc_copy(F, face, ngh);
}

/* corner default for ghost DOFs outside on >= 2 physical axes */
static void corner_default(or_ctx *c, int fid)
{
    const int g = c->d.g;
    for (int b = 0; b < c->d.nblocks; ++b)
    {
        oblock *B = &c->b[b];
        fblk *F = &B->f[fid];
        for (int i = -g; i < F->cnt[0] + g; ++i)
            for (int j = -g; j < F->cnt[1] + g; ++j)
                for (int k = -g; k < F->cnt[2] + g; ++k)
                {
                    const int idx[3] = {i, j, k};
                    int nphys = 0, nif = 0, s[3];
                    for (int a = 0; a < 3; ++a)
                    {
                        s[a] = idx[a];
                        int side = -1;
                        if (idx[a] < 0) side = 0;
                        else if (idx[a] >= F->cnt[a]) side = 1;
                        if (side < 0) continue;
                        if (B->face[2 * a + side].kind == 1) ++nif;
                        else ++nphys;
                        s[a] = idx[a] < 0 ? 0 : F->cnt[a] - 1;
                    }
                    if (nif > 0 || nphys < 2) continue;
                    for (int m = 0; m < F->nc; ++m) AT(F, i, j, k, m) = AT(F, s[0], s[1], s[2], m);
                }
    }
}

/* Sync groups (Z4_Lunar/src/boundary/Boundary_Sync_.cpp:4-65 with the
 * handler table of Z4_Lunar/src/boundary/Boundary_Build.cpp:5-271). */
static void physical_phase(or_ctx *c, int fid)
{
    const int g = c->d.g;
    for (int b = 0; b < c->d.nblocks; ++b)
    {
        oblock *B = &c->b[b];
        for (int f = 0; f < 6; ++f)
        {
            if (B->face[f].kind != 0) continue;
            const int bc = B->face[f].bc;
            fblk *F = &B->f[fid];
            switch (fid)
            {
            case OF_UH:
                if (bc == 1) bc_u_farfield(c, F, f, g);
                else if (bc == 2) bc_u_solid(c, F, f, g);
                break;
            case OF_UNA:
                if (bc == 1) bc_una_farfield(c, F, f, g);
                else if (bc == 2) bc_una_solid(c, F, f, g);
                break;
            case OF_BXI: case OF_BET: case OF_BZE:
                if (bc == 2) bc_copy(F, f, g); /* far field: nop */
                break;
            case OF_BADDXI: case OF_BADDET: case OF_BADDZE:
            case OF_JXI: case OF_JET: case OF_JZE:
            case OF_EHXI: case OF_EHET: case OF_EHZE: /* Mercury default copy (Boundary_Build.cpp:28-36) */
                bc_copy(F, f, g);
                break;
            case OF_BCELL:
                if (bc == 1) bc_bcell_farfield(F, f, g, c->d.B_imf);
                else bc_copy(F, f, g);
                break;
            case OF_BIND:
            {
                const double z[3] = {0.0, 0.0, 0.0};
                if (bc == 1) bc_bcell_farfield(F, f, g, z);
                else bc_copy(F, f, g);
                break;
            }
            case OF_JCELL:
                if (bc == 2) bc_jcell_solid(F, f, g);
                else bc_copy(F, f, g);
                break;
            case OF_EXI: case OF_EET: case OF_EZE:
                if (bc == 1) bc_eedge_farfield(c, B, F, f, g, fid - OF_EXI); /* solid: nop */
                break;
            case OF_ERXI: case OF_ERET: case OF_ERZE: /* Z5_Mercury/src/boundary/Boundary_Build.cpp:200-205 */
                if (bc == 1) bc_eedge_farfield(c, B, F, f, g, fid - OF_ERXI);
                break;
            case OF_EFXI: case OF_EFET: case OF_EFZE:
                if (bc == 2) bc_eface_solid(F, f, g); /* far field: nop */
                break;
            default:
                break;
            }
        }
    }
}

static void sync_fields(or_ctx *c, const int *fids, int n, int do_physical)
{
    if (do_physical)
    {
        for (int q = 0; q < n; ++q) physical_phase(c, fids[q]);
        for (int q = 0; q < n; ++q) corner_default(c, fids[q]);
    }
    for (int q = 0; q < n; ++q)
    {
        const int fid = fids[q];
        const int loc = kField[fid].loc;
        int base = fid, triple = 0, eface = 0;
        if (fid >= OF_BXI && fid <= OF_BZE) { base = OF_BXI; triple = 1; }
        if (fid >= OF_BADDXI && fid <= OF_BADDZE) { base = OF_BADDXI; triple = 1; }
        if (fid >= OF_JXI && fid <= OF_JZE) { base = OF_JXI; triple = 1; }
        if (fid >= OF_EXI && fid <= OF_EZE) { base = OF_EXI; triple = 1; }
        if (fid >= OF_EFXI && fid <= OF_EFZE) { base = OF_EFXI; triple = 1; eface = 1; }
        if (fid >= OF_EHXI && fid <= OF_EHZE) { base = OF_EHXI; triple = 1; }
        if (fid >= OF_ERXI && fid <= OF_ERZE) { base = OF_ERXI; triple = 1; }
        halo_apply(c, loc, base, triple, eface);
    }
}

int or_sync(or_ctx *c, const char *group)
{
    static const int fU[] = {OF_UH, OF_UNA}, fB[] = {OF_BXI, OF_BET, OF_BZE}, fBadd[] = {OF_BADDXI, OF_BADDET, OF_BADDZE};
    static const int fEf[] = {OF_EFXI, OF_EFET, OF_EFZE}, fE[] = {OF_EXI, OF_EET, OF_EZE};
    static const int fBc[] = {OF_BCELL, OF_BIND}, fJ[] = {OF_JXI, OF_JET, OF_JZE}, fJc[] = {OF_JCELL};
    static const int fUp[] = {OF_UPLUS}, fEh[] = {OF_EHXI, OF_EHET, OF_EHZE}, fEr[] = {OF_ERXI, OF_ERET, OF_ERZE};
    /* the Mercury Ucell group holds U_H and U_Na (Z5_Mercury/src/boundary/Boundary_Build.cpp:571-578) */
    if (!strcmp(group, "Ucell")) sync_fields(c, fU, c->d.model == 1 ? 2 : 1, 1);
    else if (!strcmp(group, "Bface")) sync_fields(c, fB, 3, 1);
    else if (!strcmp(group, "Badd")) sync_fields(c, fBadd, 3, 1);
    else if (!strcmp(group, "Eface")) sync_fields(c, fEf, 3, 1);
    else if (!strcmp(group, "Eedge")) sync_fields(c, fE, 3, 1);
    else if (!strcmp(group, "B_cell")) sync_fields(c, fBc, 2, 1);
    else if (!strcmp(group, "Jedge")) sync_fields(c, fJ, 3, 1);
    else if (!strcmp(group, "J_cell")) sync_fields(c, fJc, 1, 1);
    else if (!strcmp(group, "Uplus")) sync_fields(c, fUp, 1, 0);
    else if (!strcmp(group, "Ehall")) sync_fields(c, fEh, 3, 1);
    else if (!strcmp(group, "Eres")) sync_fields(c, fEr, 3, 1);
    else return -1;
    return 0;
}

/* ------------------------------------------------------- derived fields */
/* calc_PV (Z4_Lunar/src/physics/LunarSolver_Tools.cpp:58-110; per species in
 * Z5_Mercury/src/physics/MercurySolver_Tools.cpp:138-192): full extent incl. ghosts */
static void calc_pv_one(or_ctx *c, int fu, int fpv, double coeff)
{
    const double gm = c->d.gamma;
    for (int b = 0; b < c->d.nblocks; ++b)
    {
        fblk *U = &c->b[b].f[fu], *PV = &c->b[b].f[fpv];
        const int g = U->g;
        OMP_FOR
        for (int i = -g; i < U->cnt[0] + g; ++i)
            for (int j = -g; j < U->cnt[1] + g; ++j)
                for (int k = -g; k < U->cnt[2] + g; ++k)
                {
                    const double rho = dmax(AT(U, i, j, k, 0), 1e-12);
                    const double u = AT(U, i, j, k, 1) / rho;
                    const double v = AT(U, i, j, k, 2) / rho;
                    const double w = AT(U, i, j, k, 3) / rho;
                    const double E = AT(U, i, j, k, 4);
                    const double ke = 0.5 * rho * (u * u + v * v + w * w);
                    double eint = E - ke;
                    if (eint < 0.0) eint = 0.0;
                    double p = (gm - 1.0) * eint;
                    p = dmax(p, 1e-12);
                    const double T = p / (rho * coeff);
                    AT(PV, i, j, k, 0) = u;
                    AT(PV, i, j, k, 1) = v;
                    AT(PV, i, j, k, 2) = w;
                    AT(PV, i, j, k, 3) = p;
                    AT(PV, i, j, k, 4) = T;
                }
    }
}

static void calc_pv(or_ctx *c)
{
    calc_pv_one(c, OF_UH, OF_PV, c->d.state_coeff);
    if (c->d.model == 1) calc_pv_one(c, OF_UNA, OF_PVNA, c->d.state_coeff_Na);
}

/* Hall_Num_Limiter (Z5_Mercury/src/physics/MercurySolver_Tools.cpp:90-136) */
typedef struct { double nH, nNa, ne, ne_eff, chiH, chiNa, wH, wNa; } numinfo;
static numinfo num_limiter(const or_ctx *c, double rhoH, double rhoNa)
{
    numinfo n;
    const double rhoH_pos = dmax(rhoH, 0.0), rhoNa_pos = dmax(rhoNa, 0.0);
    n.nH = rhoH_pos / c->d.M_H * c->d.M_ref;
    n.nNa = rhoNa_pos / c->d.M_Na * c->d.M_ref;
    n.ne = n.nH + n.nNa;
    n.ne_eff = sqrt(n.ne * n.ne + c->d.ne_floor * c->d.ne_floor);
    if (n.ne > 1.0e-300) { n.chiH = n.nH / n.ne; n.chiNa = n.nNa / n.ne; }
    else { n.chiH = 1.0; n.chiNa = 0.0; }
    n.wH = n.nH / n.ne_eff;
    n.wNa = n.nNa / n.ne_eff;
    return n;
}

/* calc_Uplus (Z4_Lunar/src/physics/LunarSolver_Tools.cpp:112-141); two-fluid:
 * composition-weighted ion velocity (Z5_Mercury/src/physics/MercurySolver_Tools.cpp:194-277) */
static void calc_uplus(or_ctx *c)
{
    for (int b = 0; b < c->d.nblocks && c->d.model == 1; ++b)
    {
        fblk *UH = &c->b[b].f[OF_UH], *UN = &c->b[b].f[OF_UNA], *PH = &c->b[b].f[OF_PV], *PN = &c->b[b].f[OF_PVNA];
        fblk *Up = &c->b[b].f[OF_UPLUS];
        const int g = UH->g;
        OMP_FOR
        for (int i = -g; i < UH->cnt[0] + g; ++i)
            for (int j = -g; j < UH->cnt[1] + g; ++j)
                for (int k = -g; k < UH->cnt[2] + g; ++k)
                {
                    const double rhoH0 = dmax(AT(UH, i, j, k, 0), 0.0), rhoNa0 = dmax(AT(UN, i, j, k, 0), 0.0);
                    double uh[3], un[3];
                    for (int m = 0; m < 3; ++m)
                    {
                        uh[m] = rhoH0 > 1e-20 ? AT(PH, i, j, k, m) : 0.0;
                        un[m] = rhoNa0 > 1e-20 ? AT(PN, i, j, k, m) : 0.0;
                    }
                    const numinfo n = num_limiter(c, rhoH0, rhoNa0);
                    for (int m = 0; m < 3; ++m) AT(Up, i, j, k, m) = n.chiH * uh[m] + n.chiNa * un[m];
                }
    }
    for (int b = 0; b < c->d.nblocks && c->d.model == 0; ++b)
    {
        fblk *U = &c->b[b].f[OF_UH], *PV = &c->b[b].f[OF_PV], *Up = &c->b[b].f[OF_UPLUS];
        const int g = U->g;
        OMP_FOR
        for (int i = -g; i < U->cnt[0] + g; ++i)
            for (int j = -g; j < U->cnt[1] + g; ++j)
                for (int k = -g; k < U->cnt[2] + g; ++k)
                {
                    const int has = AT(U, i, j, k, 0) > 1e-20;
                    for (int m = 0; m < 3; ++m) AT(Up, i, j, k, m) = has ? AT(PV, i, j, k, m) : 0.0;
                }
    }
    or_sync(c, "Uplus");
}

/* calc_Bcell (Z4_Lunar/src/physics/LunarSolver_Tools.cpp:163-236) */
static void calc_bcell(or_ctx *c)
{
    for (int b = 0; b < c->d.nblocks; ++b)
    {
        oblock *B = &c->b[b];
        fblk *Bc = &B->f[OF_BCELL], *Bi = &B->f[OF_BIND], *W = &B->f[OF_WB];
        fblk *Bx = &B->f[OF_BXI], *By = &B->f[OF_BET], *Bz = &B->f[OF_BZE];
        fblk *Ax = &B->f[OF_BADDXI], *Ay = &B->f[OF_BADDET], *Az = &B->f[OF_BADDZE];
        OMP_FOR
        for (int i = 0; i < B->n[0]; ++i)
            for (int j = 0; j < B->n[1]; ++j)
                for (int k = 0; k < B->n[2]; ++k)
                {
                    const double pt[6] = {
                        -AT(Bx, i, j, k, 0) - AT(Ax, i, j, k, 0), AT(Bx, i + 1, j, k, 0) + AT(Ax, i + 1, j, k, 0),
                        -AT(By, i, j, k, 0) - AT(Ay, i, j, k, 0), AT(By, i, j + 1, k, 0) + AT(Ay, i, j + 1, k, 0),
                        -AT(Bz, i, j, k, 0) - AT(Az, i, j, k, 0), AT(Bz, i, j, k + 1, 0) + AT(Az, i, j, k + 1, 0)};
                    const double pi[6] = {-AT(Bx, i, j, k, 0), AT(Bx, i + 1, j, k, 0), -AT(By, i, j, k, 0),
                                          AT(By, i, j + 1, k, 0), -AT(Bz, i, j, k, 0), AT(Bz, i, j, k + 1, 0)};
                    double xt = 0.0, yt = 0.0, zt = 0.0, xi = 0.0, yi = 0.0, zi = 0.0;
                    for (int n = 0; n < 6; ++n)
                    {
                        const double wx = AT(W, i, j, k, n), wy = AT(W, i, j, k, 6 + n), wz = AT(W, i, j, k, 12 + n);
                        xt += wx * pt[n];
                        yt += wy * pt[n];
                        zt += wz * pt[n];
                        xi += wx * pi[n];
                        yi += wy * pi[n];
                        zi += wz * pi[n];
                    }
                    AT(Bc, i, j, k, 0) = xt; AT(Bc, i, j, k, 1) = yt; AT(Bc, i, j, k, 2) = zt;
                    AT(Bi, i, j, k, 0) = xi; AT(Bi, i, j, k, 1) = yi; AT(Bi, i, j, k, 2) = zi;
                }
    }
    or_sync(c, "B_cell");
}

/* Calc_J_Edge (Z4_Lunar/src/physics/LunarSolver_Tools.cpp:777-835) with
 * CTOperators::CurlAdjFaceToEdge (Z4_Lunar/src/operators/CTOperators.cpp:59-110) */
static void calc_jedge(or_ctx *c)
{
    for (int b = 0; b < c->d.nblocks; ++b)
    {
        oblock *B = &c->b[b];
        fblk *Fx = &B->f[OF_BXI], *Fy = &B->f[OF_BET], *Fz = &B->f[OF_BZE];
        fblk *bx = &B->f[OF_BETAXI], *by = &B->f[OF_BETAET], *bz = &B->f[OF_BETAZE];
        fblk *Jx = &B->f[OF_JXI], *Jy = &B->f[OF_JET], *Jz = &B->f[OF_JZE];
        fblk *ax = &B->f[OF_ALXI], *ay = &B->f[OF_ALET], *az = &B->f[OF_ALZE];
        OMP_FOR
        for (int i = 0; i < Jx->cnt[0]; ++i)
            for (int j = 0; j < Jx->cnt[1]; ++j)
                for (int k = 0; k < Jx->cnt[2]; ++k)
                {
                    const double val = (AT(by, i, j, k - 1, 0) * AT(Fy, i, j, k - 1, 0) - AT(by, i, j, k, 0) * AT(Fy, i, j, k, 0)) +
                                       (AT(bz, i, j, k, 0) * AT(Fz, i, j, k, 0) - AT(bz, i, j - 1, k, 0) * AT(Fz, i, j - 1, k, 0));
                    AT(Jx, i, j, k, 0) = 1.0 * val;
                }
        OMP_FOR
        for (int i = 0; i < Jy->cnt[0]; ++i)
            for (int j = 0; j < Jy->cnt[1]; ++j)
                for (int k = 0; k < Jy->cnt[2]; ++k)
                {
                    const double val = (AT(bx, i, j, k, 0) * AT(Fx, i, j, k, 0) - AT(bx, i, j, k - 1, 0) * AT(Fx, i, j, k - 1, 0)) +
                                       (AT(bz, i - 1, j, k, 0) * AT(Fz, i - 1, j, k, 0) - AT(bz, i, j, k, 0) * AT(Fz, i, j, k, 0));
                    AT(Jy, i, j, k, 0) = 1.0 * val;
                }
        OMP_FOR
        for (int i = 0; i < Jz->cnt[0]; ++i)
            for (int j = 0; j < Jz->cnt[1]; ++j)
                for (int k = 0; k < Jz->cnt[2]; ++k)
                {
                    const double val = (AT(bx, i, j - 1, k, 0) * AT(Fx, i, j - 1, k, 0) - AT(bx, i, j, k, 0) * AT(Fx, i, j, k, 0)) +
                                       (AT(by, i, j, k, 0) * AT(Fy, i, j, k, 0) - AT(by, i - 1, j, k, 0) * AT(Fy, i - 1, j, k, 0));
                    AT(Jz, i, j, k, 0) = 1.0 * val;
                }
        OMP_FOR
        for (int i = 0; i < Jx->cnt[0]; ++i)
            for (int j = 0; j < Jx->cnt[1]; ++j)
                for (int k = 0; k < Jx->cnt[2]; ++k) AT(Jx, i, j, k, 0) *= AT(ax, i, j, k, 0);
        OMP_FOR
        for (int i = 0; i < Jy->cnt[0]; ++i)
            for (int j = 0; j < Jy->cnt[1]; ++j)
                for (int k = 0; k < Jy->cnt[2]; ++k) AT(Jy, i, j, k, 0) *= AT(ay, i, j, k, 0);
        OMP_FOR
        for (int i = 0; i < Jz->cnt[0]; ++i)
            for (int j = 0; j < Jz->cnt[1]; ++j)
                for (int k = 0; k < Jz->cnt[2]; ++k) AT(Jz, i, j, k, 0) *= AT(az, i, j, k, 0);
    }
    or_sync(c, "Jedge");
}

/* calc_Jcell (Z4_Lunar/src/physics/LunarSolver_Tools.cpp:837-891) */
static void calc_jcell(or_ctx *c)
{
    for (int b = 0; b < c->d.nblocks; ++b)
    {
        oblock *B = &c->b[b];
        fblk *Jc = &B->f[OF_JCELL], *W = &B->f[OF_WJ];
        fblk *Jx = &B->f[OF_JXI], *Jy = &B->f[OF_JET], *Jz = &B->f[OF_JZE];
        OMP_FOR
        for (int i = 0; i < B->n[0]; ++i)
            for (int j = 0; j < B->n[1]; ++j)
                for (int k = 0; k < B->n[2]; ++k)
                {
                    const double s[12] = {AT(Jx, i, j, k, 0), AT(Jx, i, j + 1, k, 0), AT(Jx, i, j, k + 1, 0), AT(Jx, i, j + 1, k + 1, 0),
                                          AT(Jy, i, j, k, 0), AT(Jy, i + 1, j, k, 0), AT(Jy, i, j, k + 1, 0), AT(Jy, i + 1, j, k + 1, 0),
                                          AT(Jz, i, j, k, 0), AT(Jz, i + 1, j, k, 0), AT(Jz, i, j + 1, k, 0), AT(Jz, i + 1, j + 1, k, 0)};
                    double x = 0.0, y = 0.0, z = 0.0;
                    for (int n = 0; n < 12; ++n)
                    {
                        x += AT(W, i, j, k, n) * s[n];
                        y += AT(W, i, j, k, 12 + n) * s[n];
                        z += AT(W, i, j, k, 24 + n) * s[n];
                    }
                    AT(Jc, i, j, k, 0) = x; AT(Jc, i, j, k, 1) = y; AT(Jc, i, j, k, 2) = z;
                }
    }
    or_sync(c, "J_cell");
}

int or_update_derived(or_ctx *c)
{
    calc_bcell(c);
    calc_jedge(c);
    calc_jcell(c);
    calc_pv(c);
    calc_uplus(c);
    return 0;
}

/* ------------------------------------------------------------ time step */
/* Compute_Timestep scans U_H, and for the two-fluid step also U_Na with the
 * same B_cell (Z5_Mercury/src/physics/MercurySolver_Timestep.cpp:8-140) */
int or_compute_dt(or_ctx *c, double *dt_out)
{
    const double gm = c->d.gamma, iMA2 = c->d.inver_MA2, CFL = c->d.cfl;
    double dt_local = 1e100;
    const int nsp = c->d.model == 1 ? 2 : 1;
    for (int sp = 0; sp < nsp; ++sp)
    for (int b = 0; b < c->d.nblocks; ++b)
    {
        oblock *B = &c->b[b];
        fblk *U = &B->f[sp ? OF_UNA : OF_UH], *Bc = &B->f[OF_BCELL], *Jac = &B->f[OF_JAC];
        fblk *Ax = &B->f[OF_JDXI], *Ay = &B->f[OF_JDET], *Az = &B->f[OF_JDZE];
        /* per-thread minima combined with the same min: the minimum of finite
           values does not depend on the order they are taken in */
        OMP_PRAGMA(omp parallel num_threads(c->nthreads) if (c->nthreads > 1))
        {
        double dt_mine = 1e100;
        OMP_PRAGMA(omp for schedule(static))
        for (int i = 0; i < B->n[0]; ++i)
            for (int j = 0; j < B->n[1]; ++j)
                for (int k = 0; k < B->n[2]; ++k)
                {
                    const double V = fabs(AT(Jac, i, j, k, 0));
                    if (V <= 0.0) continue;
                    const double rho = dmax(AT(U, i, j, k, 0), 1e-2);
                    const double ux = AT(U, i, j, k, 1) / rho, uy = AT(U, i, j, k, 2) / rho, uz = AT(U, i, j, k, 3) / rho;
                    const double E = AT(U, i, j, k, 4);
                    const double ke = 0.5 * rho * (ux * ux + uy * uy + uz * uz);
                    const double eint = E - ke;
                    const double p = (gm - 1.0) * eint;
                    const double cs2 = gm * p / rho;
                    const double Bx = AT(Bc, i, j, k, 0), By = AT(Bc, i, j, k, 1), Bz = AT(Bc, i, j, k, 2);
                    const double B2 = Bx * Bx + By * By + Bz * Bz;
                    double denom = 0.0;
                    for (int f = 0; f < 6; ++f)
                    {
                        fblk *A = f < 2 ? Ax : (f < 4 ? Ay : Az);
                        const int up = f & 1;
                        const int ii = i + (f == 1), jj = j + (f == 3), kk = k + (f == 5);
                        const double ax = AT(A, ii, jj, kk, 0), ay = AT(A, ii, jj, kk, 1), az = AT(A, ii, jj, kk, 2);
                        const double An = sqrt(ax * ax + ay * ay + az * az);
                        double term = 0.0;
                        if (An > 0.0)
                        {
                            double nx = ax / An, ny = ay / An, nz = az / An;
                            if (!up) { nx = -nx; ny = -ny; nz = -nz; }
                            const double un = ux * nx + uy * ny + uz * nz;
                            const double Bn = Bx * nx + By * ny + Bz * nz;
                            const double vA2 = iMA2 * B2 / rho;
                            const double vAn2 = iMA2 * (Bn * Bn) / rho;
                            const double tt = cs2 + vA2;
                            double disc = tt * tt - 4.0 * cs2 * vAn2;
                            if (disc < 0.0) disc = 0.0;
                            const double cf = sqrt(0.5 * (tt + sqrt(disc)));
                            term = (fabs(un) + cf) * An;
                        }
                        denom += term;
                    }
                    if (denom > 0.0)
                    {
                        const double dtc = CFL * V / denom;
                        dt_mine = dmin(dt_mine, dtc);
                    }
                }
        OMP_PRAGMA(omp critical(or_dt_min))
        dt_local = dmin(dt_local, dt_mine);
        }
    }
    *dt_out = dt_local;
    if (!isfinite(dt_local) || dt_local < c->d.dt_floor) return -2;
    return 0;
}

/* ------------------------------------------------------- fluid fluxes */
static inline double minmod(double a, double b)
{
    if (a * b <= 0.0) return 0.0;
    return fabs(a) < fabs(b) ? a : b;
}

/* Rusanov flux of Scheme_U_ (Z4_Lunar/src/physics/LunarSolver_FluidRHS.cpp:5-102) */
static void rusanov(double gm, const double *WL, const double *WR, double k1, double k2, double k3, double *out)
{
    const double rhoL = WL[0], uL = WL[1], vL = WL[2], wL = WL[3], pL = WL[4];
    const double rhoR = WR[0], uR = WR[1], vR = WR[2], wR = WR[3], pR = WR[4];
    double UL[5], UR[5], FL[5], FR[5];
    UL[0] = rhoL; UL[1] = rhoL * uL; UL[2] = rhoL * vL; UL[3] = rhoL * wL;
    UL[4] = pL / (gm - 1.0) + 0.5 * rhoL * (uL * uL + vL * vL + wL * wL);
    UR[0] = rhoR; UR[1] = rhoR * uR; UR[2] = rhoR * vR; UR[3] = rhoR * wR;
    UR[4] = pR / (gm - 1.0) + 0.5 * rhoR * (uR * uR + vR * vR + wR * wR);
    const double kk = k1 * k1 + k2 * k2 + k3 * k3;
    const double uvL = k1 * uL + k2 * vL + k3 * wL;
    const double cL = sqrt(dmax(0.0, gm * pL / dmax(rhoL, 1e-30)));
    const double radL = fabs(uvL) + cL * sqrt(kk);
    const double uvR = k1 * uR + k2 * vR + k3 * wR;
    const double cR = sqrt(dmax(0.0, gm * pR / dmax(rhoR, 1e-30)));
    const double radR = fabs(uvR) + cR * sqrt(kk);
    const double rad = dmax(radL, radR);
    {
        const double rhoe = pL / (gm - 1.0) + 0.5 * rhoL * (uL * uL + vL * vL + wL * wL);
        FL[0] = rhoL * uvL; FL[1] = rhoL * uvL * uL + k1 * pL; FL[2] = rhoL * uvL * vL + k2 * pL;
        FL[3] = rhoL * uvL * wL + k3 * pL; FL[4] = uvL * (rhoe + pL);
    }
    {
        const double rhoe = pR / (gm - 1.0) + 0.5 * rhoR * (uR * uR + vR * vR + wR * wR);
        FR[0] = rhoR * uvR; FR[1] = rhoR * uvR * uR + k1 * pR; FR[2] = rhoR * uvR * vR + k2 * pR;
        FR[3] = rhoR * uvR * wR + k3 * pR; FR[4] = uvR * (rhoe + pR);
    }
    for (int m = 0; m < 5; ++m) out[m] = 0.5 * (FL[m] + FR[m]) - 0.5 * rad * (UR[m] - UL[m]);
}

static void cell_w(const fblk *U, const fblk *PV, int i, int j, int k, double *W)
{
    W[0] = AT(U, i, j, k, 0);
    W[1] = AT(PV, i, j, k, 0); W[2] = AT(PV, i, j, k, 1); W[3] = AT(PV, i, j, k, 2); W[4] = AT(PV, i, j, k, 3);
}

/* Scheme_U_; the two-fluid step runs it for H+ then Na+
 * (Z5_Mercury/src/physics/MercurySolver_FluidRHS.cpp:115-202) */
static void scheme_u_one(or_ctx *c, int fu, int fpv, int frhs)
{
    const double gm = c->d.gamma;
    for (int b = 0; b < c->d.nblocks; ++b)
    {
        oblock *B = &c->b[b];
        fblk *U = &B->f[fu], *PV = &B->f[fpv], *RHS = &B->f[frhs], *Jac = &B->f[OF_JAC];
        fblk *F[3] = {&B->f[OF_FXI], &B->f[OF_FET], &B->f[OF_FZE]};
        fblk *M[3] = {&B->f[OF_JDXI], &B->f[OF_JDET], &B->f[OF_JDZE]};
        for (int d = 0; d < 3; ++d)
        {
            fblk *Fd = F[d];
            OMP_FOR
            for (int i = 0; i < Fd->cnt[0]; ++i)
                for (int j = 0; j < Fd->cnt[1]; ++j)
                    for (int k = 0; k < Fd->cnt[2]; ++k)
                    {
                        const int di = d == 0, dj = d == 1, dk = d == 2;
                        double WL[5], WR[5], flux[5];
                        cell_w(U, PV, i - di, j - dj, k - dk, WL);
                        cell_w(U, PV, i, j, k, WR);
                        if (c->d.recon == 1)
                        {
                            /* MUSCL-minmod on (rho, u, v, w, p); first order
                               where the limited state is not positive */
                            double WLL[5], WRR[5], L2[5], R2[5];
                            cell_w(U, PV, i - 2 * di, j - 2 * dj, k - 2 * dk, WLL);
                            cell_w(U, PV, i + di, j + dj, k + dk, WRR);
                            for (int m = 0; m < 5; ++m)
                            {
                                L2[m] = WL[m] + 0.5 * minmod(WL[m] - WLL[m], WR[m] - WL[m]);
                                R2[m] = WR[m] - 0.5 * minmod(WR[m] - WL[m], WRR[m] - WR[m]);
                            }
                            if (L2[0] > 0.0 && L2[4] > 0.0 && R2[0] > 0.0 && R2[4] > 0.0)
                                for (int m = 0; m < 5; ++m) { WL[m] = L2[m]; WR[m] = R2[m]; }
                        }
                        rusanov(gm, WL, WR, AT(M[d], i, j, k, 0), AT(M[d], i, j, k, 1), AT(M[d], i, j, k, 2), flux);
                        for (int m = 0; m < 5; ++m) AT(Fd, i, j, k, m) = flux[m];
                    }
        }
        OMP_FOR
        for (int i = 0; i < B->n[0]; ++i)
            for (int j = 0; j < B->n[1]; ++j)
                for (int k = 0; k < B->n[2]; ++k)
                {
                    const double invJ = 1.0 / AT(Jac, i, j, k, 0);
                    for (int m = 0; m < 5; ++m)
                    {
                        AT(RHS, i, j, k, m) -= (AT(F[0], i + 1, j, k, m) - AT(F[0], i, j, k, m)) * invJ;
                        AT(RHS, i, j, k, m) -= (AT(F[1], i, j + 1, k, m) - AT(F[1], i, j, k, m)) * invJ;
                        AT(RHS, i, j, k, m) -= (AT(F[2], i, j, k + 1, m) - AT(F[2], i, j, k, m)) * invJ;
                    }
                }
    }
}

static void scheme_u(or_ctx *c)
{
    scheme_u_one(c, OF_UH, OF_PV, OF_RHSH);
    if (c->d.model == 1) scheme_u_one(c, OF_UNA, OF_PVNA, OF_RHSNA);
}

/* AddSourceToRHS_Fluid (Z4_Lunar/src/physics/LunarSolver_FluidRHS.cpp:528-568)
 * with Hall_Num_Limiter (Z4_Lunar/src/physics/LunarSolver_Tools.cpp:47-56) */
static void add_source_mercury(or_ctx *c);
static void add_source(or_ctx *c)
{
    if (c->d.model == 1)
    {
        add_source_mercury(c);
        return;
    }
    for (int b = 0; b < c->d.nblocks; ++b)
    {
        oblock *B = &c->b[b];
        fblk *U = &B->f[OF_UH], *PV = &B->f[OF_PV], *Bc = &B->f[OF_BCELL], *J = &B->f[OF_JCELL], *RHS = &B->f[OF_RHSH];
        OMP_FOR
        for (int i = 0; i < B->n[0]; ++i)
            for (int j = 0; j < B->n[1]; ++j)
                for (int k = 0; k < B->n[2]; ++k)
                {
                    const double n_true = dmax(AT(U, i, j, k, 0), 0.0) / c->d.M_H * c->d.M_ref;
                    const double ne = n_true;
                    const double ne_eff = sqrt(ne * ne + c->d.ne_floor * c->d.ne_floor);
                    const double wgt = n_true / ne_eff;
                    const double fc = c->d.momentum_hall_coeff * wgt;
                    const double jx = AT(J, i, j, k, 1) * AT(Bc, i, j, k, 2) - AT(J, i, j, k, 2) * AT(Bc, i, j, k, 1);
                    const double jy = AT(J, i, j, k, 2) * AT(Bc, i, j, k, 0) - AT(J, i, j, k, 0) * AT(Bc, i, j, k, 2);
                    const double jz = AT(J, i, j, k, 0) * AT(Bc, i, j, k, 1) - AT(J, i, j, k, 1) * AT(Bc, i, j, k, 0);
                    AT(RHS, i, j, k, 1) += fc * jx;
                    AT(RHS, i, j, k, 2) += fc * jy;
                    AT(RHS, i, j, k, 3) += fc * jz;
                    AT(RHS, i, j, k, 4) += fc * (jx * AT(PV, i, j, k, 0) + jy * AT(PV, i, j, k, 1) + jz * AT(PV, i, j, k, 2));
                }
    }
}

/* AddSourceToRHS_Fluid of the two-fluid step (Z5_Mercury/src/physics/
 * MercurySolver_FluidRHS.cpp:206-527): per species the induction force
 * n_s (u_s - u_+) x B, the composition-weighted J x B, the ambipolar
 * electron-pressure force, photoionization (Na+ only), drag on the neutrals
 * and thermal exchange.  The metric of the pressure gradient averages faces
 * i-1 and i (j-1 and j, k-1 and k) as the reference does. */
static void add_source_mercury(or_ctx *c)
{
    const double sk1 = 5.0e-5, sk2 = 1.0e-5; /* drag frequency, illuminated / shadow [1/s] (:211-212) */
    (void)sk1; (void)sk2;
    const double induce = c->d.induce, hall = c->d.momentum_hall_coeff;
    const double a1_Na = c->d.a1_Na, a4 = c->d.a4, a5 = c->d.a5, a6 = c->d.a6, Tn0 = c->d.Tn0;
    const double M_H = c->d.M_H, M_Na = c->d.M_Na, T_ref = c->d.T_ref;
    for (int b = 0; b < c->d.nblocks; ++b)
    {
        oblock *B = &c->b[b];
        fblk *UH = &B->f[OF_UH], *UN = &B->f[OF_UNA], *PH = &B->f[OF_PV], *PN = &B->f[OF_PVNA];
        fblk *Up = &B->f[OF_UPLUS], *Bt = &B->f[OF_BCELL], *Jc = &B->f[OF_JCELL];
        fblk *Ph = &B->f[OF_PHOTO], *Dr = &B->f[OF_DRAG], *RH = &B->f[OF_RHSH], *RN = &B->f[OF_RHSNA];
        fblk *Jac = &B->f[OF_JAC], *Ax = &B->f[OF_JDXI], *Ay = &B->f[OF_JDET], *Az = &B->f[OF_JDZE];
        OMP_FOR
        for (int i = 0; i < B->n[0]; ++i)
            for (int j = 0; j < B->n[1]; ++j)
                for (int k = 0; k < B->n[2]; ++k)
                {
                    const double upx = AT(Up, i, j, k, 0), upy = AT(Up, i, j, k, 1), upz = AT(Up, i, j, k, 2);
                    const double Bx = AT(Bt, i, j, k, 0), By = AT(Bt, i, j, k, 1), Bz = AT(Bt, i, j, k, 2);
                    const double Jx = AT(Jc, i, j, k, 0), Jy = AT(Jc, i, j, k, 1), Jz = AT(Jc, i, j, k, 2);
                    const double sjbx = Jy * Bz - Jz * By;
                    const double sjby = Jz * Bx - Jx * Bz;
                    const double sjbz = Jx * By - Jy * Bx;
                    const numinfo num = num_limiter(c, AT(UH, i, j, k, 0), AT(UN, i, j, k, 0));
                    double dpex = 0.0, dpey = 0.0, dpez = 0.0;
                    if (c->d.ambipolar)
                    {
                        /* derv_at (:238-261) and dd (:228-235) */
                        double ax, ay, az, bx, by, bz, cx, cy, cz;
                        const double V = fabs(AT(Jac, i, j, k, 0));
                        if (V <= 0.0) ax = ay = az = bx = by = bz = cx = cy = cz = 0.0;
                        else
                        {
                            ax = 0.5 * (AT(Ax, i - 1, j, k, 0) + AT(Ax, i, j, k, 0)) / V;
                            ay = 0.5 * (AT(Ax, i - 1, j, k, 1) + AT(Ax, i, j, k, 1)) / V;
                            az = 0.5 * (AT(Ax, i - 1, j, k, 2) + AT(Ax, i, j, k, 2)) / V;
                            bx = 0.5 * (AT(Ay, i, j - 1, k, 0) + AT(Ay, i, j, k, 0)) / V;
                            by = 0.5 * (AT(Ay, i, j - 1, k, 1) + AT(Ay, i, j, k, 1)) / V;
                            bz = 0.5 * (AT(Ay, i, j - 1, k, 2) + AT(Ay, i, j, k, 2)) / V;
                            cx = 0.5 * (AT(Az, i, j, k - 1, 0) + AT(Az, i, j, k, 0)) / V;
                            cy = 0.5 * (AT(Az, i, j, k - 1, 1) + AT(Az, i, j, k, 1)) / V;
                            cz = 0.5 * (AT(Az, i, j, k - 1, 2) + AT(Az, i, j, k, 2)) / V;
                        }
                        const double pip = AT(PH, i + 1, j, k, 3) + AT(PN, i + 1, j, k, 3);
                        const double pim = AT(PH, i - 1, j, k, 3) + AT(PN, i - 1, j, k, 3);
                        const double pjp = AT(PH, i, j + 1, k, 3) + AT(PN, i, j + 1, k, 3);
                        const double pjm = AT(PH, i, j - 1, k, 3) + AT(PN, i, j - 1, k, 3);
                        const double pkp = AT(PH, i, j, k + 1, 3) + AT(PN, i, j, k + 1, 3);
                        const double pkm = AT(PH, i, j, k - 1, 3) + AT(PN, i, j, k - 1, 3);
                        dpex = 0.5 * (ax * (pip - pim) + bx * (pjp - pjm) + cx * (pkp - pkm));
                        dpey = 0.5 * (ay * (pip - pim) + by * (pjp - pjm) + cy * (pkp - pkm));
                        dpez = 0.5 * (az * (pip - pim) + bz * (pjp - pjm) + cz * (pkp - pkm));
                    }
                    const double sss = AT(Ph, i, j, k, 0);
                    const double vst = AT(Dr, i, j, k, 0);
                    for (int sp = 0; sp < 2; ++sp)
                    {
                        fblk *U = sp ? UN : UH, *PV = sp ? PN : PH, *R = sp ? RN : RH;
                        const double n_s = sp ? num.nNa : num.nH;
                        const double chi = sp ? num.wNa : num.wH;
                        const double rho = dmax(AT(U, i, j, k, 0), 0.0);
                        const double u = AT(PV, i, j, k, 0), v = AT(PV, i, j, k, 1), w = AT(PV, i, j, k, 2);
                        const double u2 = u * u + v * v + w * w;
                        const double subx = (v - upy) * Bz - (w - upz) * By;
                        const double suby = (w - upz) * Bx - (u - upx) * Bz;
                        const double subz = (u - upx) * By - (v - upy) * Bx;
                        const double subu = subx * u + suby * v + subz * w;
                        const double sjbu = sjbx * u + sjby * v + sjbz * w;
                        const double dpeu = dpex * u + dpey * v + dpez * w;
                        const double Ts0 = AT(PV, i, j, k, 4) * T_ref;
                        const double b1 = sp ? vst * (Tn0 - Ts0) * (M_Na / (M_Na + M_Na)) : vst * (Tn0 - Ts0) * (M_H / (M_H + M_Na));
                        const double b2 = sp ? (M_Na / (M_Na + M_Na)) * vst : (M_Na / (M_H + M_Na)) * vst;
                        AT(R, i, j, k, 1) += induce * n_s * subx + hall * chi * sjbx;
                        AT(R, i, j, k, 2) += induce * n_s * suby + hall * chi * sjby;
                        AT(R, i, j, k, 3) += induce * n_s * subz + hall * chi * sjbz;
                        AT(R, i, j, k, 4) += induce * n_s * subu + hall * chi * sjbu;
                        if (c->d.ambipolar)
                        {
                            AT(R, i, j, k, 1) -= chi * dpex;
                            AT(R, i, j, k, 2) -= chi * dpey;
                            AT(R, i, j, k, 3) -= chi * dpez;
                            AT(R, i, j, k, 4) -= chi * dpeu;
                        }
                        if (sp)
                        {
                            AT(R, i, j, k, 0) += a1_Na * sss;
                            AT(R, i, j, k, 4) += a6 * sss * Tn0;
                        }
                        AT(R, i, j, k, 1) -= a4 * rho * u * vst;
                        AT(R, i, j, k, 2) -= a4 * rho * v * vst;
                        AT(R, i, j, k, 3) -= a4 * rho * w * vst;
                        AT(R, i, j, k, 4) -= a4 * rho * u2 * vst;
                        AT(R, i, j, k, 4) += a5 * n_s * b1;
                        AT(R, i, j, k, 4) += a4 * rho * u2 * b2;
                    }
                }
    }
}

/* AddAmbipolarEdgeEMF_ (Z5_Mercury/src/physics/MercurySolver_MagExplicit_Ambipo.cpp:3-144):
 * E_edge -= pe_coef (pe(node 1) - pe(node 0)) / ne_eff(edge), pe = p_H + p_Na
 * averaged over the 8 cells around a node, ne from the 4 cells around the edge. */
static double pe_node(const fblk *PH, const fblk *PN, int i, int j, int k)
{
    double sum = 0.0;
    for (int di = -1; di <= 0; ++di)
        for (int dj = -1; dj <= 0; ++dj)
            for (int dk = -1; dk <= 0; ++dk)
                sum += AT(PH, i + di, j + dj, k + dk, 3) + AT(PN, i + di, j + dj, k + dk, 3);
    return 0.125 * sum;
}

static void add_ambipolar_emf(or_ctx *c)
{
    for (int b = 0; b < c->d.nblocks; ++b)
    {
        oblock *B = &c->b[b];
        fblk *UH = &B->f[OF_UH], *UN = &B->f[OF_UNA], *PH = &B->f[OF_PV], *PN = &B->f[OF_PVNA];
        for (int a = 0; a < 3; ++a)
        {
            fblk *E = &B->f[OF_EXI + a];
            const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2; /* the two transverse axes, ascending */
            OMP_FOR
            for (int i = 0; i < E->cnt[0]; ++i)
                for (int j = 0; j < E->cnt[1]; ++j)
                    for (int k = 0; k < E->cnt[2]; ++k)
                    {
                        const int d1[3] = {a == 0, a == 1, a == 2};
                        const double pe0 = pe_node(PH, PN, i, j, k);
                        const double pe1 = pe_node(PH, PN, i + d1[0], j + d1[1], k + d1[2]);
                        double rhoH = 0.0, rhoNa = 0.0;
                        for (int s1 = -1; s1 <= 0; ++s1)
                            for (int s2 = -1; s2 <= 0; ++s2)
                            {
                                int q[3] = {i, j, k};
                                q[t1] += s1;
                                q[t2] += s2;
                                rhoH += AT(UH, q[0], q[1], q[2], 0);
                                rhoNa += AT(UN, q[0], q[1], q[2], 0);
                            }
                        const numinfo num = num_limiter(c, 0.25 * rhoH, 0.25 * rhoNa);
                        AT(E, i, j, k, 0) -= c->d.pe_coef * (pe1 - pe0) / num.ne_eff;
                    }
        }
    }
}

static void zero_fblk(fblk *F);

/* edge midpoint of node (i,j,k) -> (i,j,k) + e_a (HallRadialTaperEdge_,
 * MercurySolver_MagExplicit_Hall.cpp:19-37; edge_midpoint, Resis.cpp:91-106) */
static void edge_mid(const oblock *B, int a, int i, int j, int k, double *m)
{
    const int ip = i + (a == 0), jp = j + (a == 1), kp = k + (a == 2);
    for (int q = 0; q < 3; ++q) m[q] = 0.5 * (NODE(B, q, i, j, k) + NODE(B, q, ip, jp, kp));
}

/* AddLocalArtificialResistivityToEdgeEMF_ (Z5_Mercury/src/physics/
 * MercurySolver_MagExplicit_Resis.cpp:67-186): E += eta(edge) J_edge with
 * eta = eta_max exp(-r / r_decay) inside r < r_cutoff of the centre. */
static void add_local_arti_emf(or_ctx *c)
{
    const or_desc *d = &c->d;
    for (int b = 0; b < d->nblocks; ++b)
    {
        oblock *B = &c->b[b];
        for (int a = 0; a < 3; ++a)
        {
            fblk *E = &B->f[OF_EXI + a], *J = &B->f[OF_JXI + a];
            for (int i = 0; i < E->cnt[0]; ++i)
                for (int j = 0; j < E->cnt[1]; ++j)
                    for (int k = 0; k < E->cnt[2]; ++k)
                    {
                        double m[3];
                        edge_mid(B, a, i, j, k, m);
                        const double dx = m[0] - d->loc_center[0], dy = m[1] - d->loc_center[1], dz = m[2] - d->loc_center[2];
                        const double r = sqrt(dx * dx + dy * dy + dz * dz);
                        double eta;
                        if (d->loc_r_cutoff > 0.0 && r >= d->loc_r_cutoff) eta = 0.0;
                        else eta = d->loc_eta_max * exp(-r / d->loc_r_decay);
                        AT(E, i, j, k, 0) += eta * AT(J, i, j, k, 0);
                    }
        }
    }
}

/* HallRadialTaper_ (MercurySolver_MagExplicit_Hall.cpp:7-17) */
static double hall_taper(const or_desc *d, const double *m)
{
    const double r = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
    const double width = d->hall_rmax - d->hall_rmin;
    if (width <= 0.0) return (r >= d->hall_rmin) ? 1.0 : 0.0;
    const double mid = 0.5 * (d->hall_rmin + d->hall_rmax);
    return 0.5 * (1.0 + tanh(2.0 * (r - mid) / width));
}

/* AddHallEdgeEMF_ / BuildHallFaceEMF_Rusanov_diff_ (MercurySolver_MagExplicit_Hall.cpp:39-440):
 * cell Hall field alpha (J x B) averaged over the 4 cells around an edge and
 * integrated along it, plus a whistler LLF diffusion of the transverse face
 * flux jumps, tapered in radius; then Sync("Ehall") and E += Ehall on the
 * edges of the block (MercurySolver_MagneticRHS.cpp:35-56). */
static void add_hall_emf(or_ctx *c)
{
    const or_desc *d = &c->d;
    const double eps = 1e-14, Cwh = 0.1;
    for (int b = 0; b < d->nblocks; ++b)
    {
        oblock *B = &c->b[b];
        fblk *Bc = &B->f[OF_BCELL], *Jc = &B->f[OF_JCELL], *UH = &B->f[OF_UH], *UN = &B->f[OF_UNA];
        fblk *Bf[3] = {&B->f[OF_BXI], &B->f[OF_BET], &B->f[OF_BZE]};
        fblk *JD[3] = {&B->f[OF_JDXI], &B->f[OF_JDET], &B->f[OF_JDZE]};
        for (int a = 0; a < 3; ++a) zero_fblk(&B->f[OF_EHXI + a]);
        /* Ehc and beta of every cell incl. ghosts */
        const int g = Bc->g;
        const int ext[3] = {Bc->ext[0], Bc->ext[1], Bc->ext[2]};
        const size_t ncell = (size_t)ext[0] * ext[1] * ext[2];
        double *Ehc = (double *)malloc(sizeof(double) * 3 * ncell), *beta = (double *)malloc(sizeof(double) * ncell);
#define CIX(i, j, k) ((((size_t)((i) + g)) * ext[1] + (size_t)((j) + g)) * ext[2] + (size_t)((k) + g))
        for (int i = -g; i < Bc->cnt[0] + g; ++i)
            for (int j = -g; j < Bc->cnt[1] + g; ++j)
                for (int k = -g; k < Bc->cnt[2] + g; ++k)
                {
                    const numinfo n = num_limiter(c, AT(UH, i, j, k, 0), AT(UN, i, j, k, 0));
                    const double alpha = d->hall_coef / n.ne_eff;
                    const double Jx = AT(Jc, i, j, k, 0), Jy = AT(Jc, i, j, k, 1), Jz = AT(Jc, i, j, k, 2);
                    const double Bx = AT(Bc, i, j, k, 0), By = AT(Bc, i, j, k, 1), Bz = AT(Bc, i, j, k, 2);
                    const size_t q = CIX(i, j, k);
                    Ehc[3 * q + 0] = alpha * (Jy * Bz - Jz * By);
                    Ehc[3 * q + 1] = alpha * (Jz * Bx - Jx * Bz);
                    Ehc[3 * q + 2] = alpha * (Jx * By - Jy * Bx);
                    beta[q] = fabs(alpha) * sqrt(Bx * Bx + By * By + Bz * Bz);
                }
        for (int a = 0; a < 3; ++a)
        {
            fblk *Eh = &B->f[OF_EHXI + a];
            const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2; /* transverse axes, ascending */
            for (int i = 0; i < Eh->cnt[0]; ++i)
                for (int j = 0; j < Eh->cnt[1]; ++j)
                    for (int k = 0; k < Eh->cnt[2]; ++k)
                    {
                        size_t cq[4];
                        int nq = 0;
                        for (int s1 = -1; s1 <= 0; ++s1)
                            for (int s2 = -1; s2 <= 0; ++s2)
                            {
                                int q[3] = {i, j, k};
                                q[t1] += s1;
                                q[t2] += s2;
                                cq[nq++] = CIX(q[0], q[1], q[2]);
                            }
                        double Ec[3];
                        for (int m = 0; m < 3; ++m)
                            Ec[m] = 0.25 * (Ehc[3 * cq[0] + m] + Ehc[3 * cq[1] + m] + Ehc[3 * cq[2] + m] + Ehc[3 * cq[3] + m]);
                        const int ip = i + (a == 0), jp = j + (a == 1), kp = k + (a == 2);
                        const double dr[3] = {NODE(B, 0, ip, jp, kp) - NODE(B, 0, i, j, k),
                                              NODE(B, 1, ip, jp, kp) - NODE(B, 1, i, j, k),
                                              NODE(B, 2, ip, jp, kp) - NODE(B, 2, i, j, k)};
                        const double Ecen = Ec[0] * dr[0] + Ec[1] * dr[1] + Ec[2] * dr[2];
                        const double beta_e = dmax(dmax(beta[cq[0]], beta[cq[1]]), dmax(beta[cq[2]], beta[cq[3]]));
                        const double L = dmax(sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]), eps);
                        /* face areas |S| of the two transverse families, averaged
                           over the two faces along the other transverse axis */
                        double mu[3] = {0.0, 0.0, 0.0};
                        for (int f = 0; f < 3; ++f)
                        {
                            if (f == a) continue;
                            const int o_ax = 3 - a - f; /* face family f is averaged along axis o_ax */
                            int q0[3] = {i, j, k}, q1[3] = {i, j, k};
                            q1[o_ax] -= 1;
                            const double *s0 = &AT(JD[f], q0[0], q0[1], q0[2], 0), *s1 = &AT(JD[f], q1[0], q1[1], q1[2], 0);
                            const double A0 = sqrt(s0[0] * s0[0] + s0[1] * s0[1] + s0[2] * s0[2]);
                            const double A1 = sqrt(s1[0] * s1[0] + s1[1] * s1[1] + s1[2] * s1[2]);
                            const double Abar = 0.5 * (A0 + A1);
                            const double inv_h = L / dmax(Abar, eps);
                            mu[o_ax] = Cwh * beta_e * inv_h * inv_h; /* mu along the averaging axis */
                        }
                        /* jumps of the face fluxes: family t2 across t1, family t1 across t2 */
                        int p1[3] = {i, j, k}, p2[3] = {i, j, k};
                        p1[t1] -= 1;
                        p2[t2] -= 1;
                        const double dB_t2_t1 = AT(Bf[t2], i, j, k, 0) - AT(Bf[t2], p1[0], p1[1], p1[2], 0);
                        const double dB_t1_t2 = AT(Bf[t1], i, j, k, 0) - AT(Bf[t1], p2[0], p2[1], p2[2], 0);
                        double m[3];
                        edge_mid(B, a, i, j, k, m);
                        double e;
                        if (a == 1) /* eta edge: + mu_zeta dBxi_zeta - mu_xi dBzeta_xi */
                            e = Ecen + 0.5 * mu[t2] * dB_t1_t2 - 0.5 * mu[t1] * dB_t2_t1;
                        else /* xi: + mu_eta dBzeta_eta - mu_zeta dBeta_zeta; zeta: + mu_xi dBeta_xi - mu_eta dBxi_eta */
                            e = Ecen + 0.5 * mu[t1] * dB_t2_t1 - 0.5 * mu[t2] * dB_t1_t2;
                        AT(Eh, i, j, k, 0) = hall_taper(d, m) * e;
                    }
        }
#undef CIX
        free(Ehc);
        free(beta);
    }
    or_sync(c, "Ehall");
    for (int b = 0; b < d->nblocks; ++b)
        for (int a = 0; a < 3; ++a)
        {
            fblk *E = &c->b[b].f[OF_EXI + a], *Eh = &c->b[b].f[OF_EHXI + a];
            for (int i = 0; i < E->cnt[0]; ++i)
                for (int j = 0; j < E->cnt[1]; ++j)
                    for (int k = 0; k < E->cnt[2]; ++k) AT(E, i, j, k, 0) += AT(Eh, i, j, k, 0);
        }
}

/* ------------------------------------------------------------ CT / EMF */
typedef struct
{
    fblk *B[3], *Badd[3], *JD[3], *Jac, *Up;
    const or_ctx *c;
    int b;
} emf_view;

static void emf_view_of(const or_ctx *c, int b, emf_view *v);

static inline double face_ind(const emf_view *v, int ax, int i, int j, int k) { return AT(v->B[ax], i, j, k, 0); }
static inline double face_tot(const emf_view *v, int ax, int i, int j, int k)
{
    return AT(v->B[ax], i, j, k, 0) + AT(v->Badd[ax], i, j, k, 0);
}
static inline double cell_beta_ind(const emf_view *v, int ax, int i, int j, int k)
{
    return 0.5 * (face_ind(v, ax, i, j, k) + face_ind(v, ax, i + (ax == 0), j + (ax == 1), k + (ax == 2)));
}
static inline double cell_beta_tot(const emf_view *v, int ax, int i, int j, int k)
{
    return 0.5 * (face_tot(v, ax, i, j, k) + face_tot(v, ax, i + (ax == 0), j + (ax == 1), k + (ax == 2)));
}
/* cell_u_contra (Z4_Lunar/src/physics/LunarSolver_MagExplicit_Ideal.cpp:414-428).
 * A ghost cell outside two (non-periodic) interfaces at once sits at a
 * corner where only three blocks meet; its node-built geometry is
 * degenerate, so it takes the contravariant velocity of the cell the halo
 * resolves it to, along the mapped axis with the mapped orientation. */
static int corner_cell(const emf_view *v, int i, int j, int k)
{
    const oblock *B = &v->c->b[v->b];
    const int idx[3] = {i, j, k};
    int n = 0;
    for (int a = 0; a < 3; ++a)
    {
        int side = -1;
        if (idx[a] < 0) side = 0;
        else if (idx[a] >= B->n[a]) side = 1;
        if (side < 0) continue;
        const or_face *F = &B->face[2 * a + side];
        if (F->kind == 1 && !F->periodic) ++n;
    }
    return n >= 2;
}

static double cell_u_contra(const emf_view *v, int ax, int i, int j, int k)
{
    if (corner_cell(v, i, j, k))
    {
        const int idx[3] = {i, j, k};
        const osrc s = resolve(v->c, v->b, L_CELL, idx);
        emf_view w;
        emf_view_of(v->c, s.block, &w);
        const double u = cell_u_contra(&w, s.perm[ax], s.idx[0], s.idx[1], s.idx[2]);
        return s.sign[ax] > 0 ? u : -u;
    }
    const int ip = i + (ax == 0), jp = j + (ax == 1), kp = k + (ax == 2);
    const double kx = 0.5 * (AT(v->JD[ax], i, j, k, 0) + AT(v->JD[ax], ip, jp, kp, 0));
    const double ky = 0.5 * (AT(v->JD[ax], i, j, k, 1) + AT(v->JD[ax], ip, jp, kp, 1));
    const double kz = 0.5 * (AT(v->JD[ax], i, j, k, 2) + AT(v->JD[ax], ip, jp, kp, 2));
    const double inv_jac = 1.0 / (fabs(AT(v->Jac, i, j, k, 0)) + 1.0e-30);
    return (kx * AT(v->Up, i, j, k, 0) + ky * AT(v->Up, i, j, k, 1) + kz * AT(v->Up, i, j, k, 2)) * inv_jac;
}

/* flux lambda of AssembleOneDirectionEMF_ (:430-450) */
static double emf_flux(const emf_view *v, int bax, int fax, const int *L, const int *R, double bff)
{
    const double btL = cell_beta_tot(v, bax, L[0], L[1], L[2]);
    const double btR = cell_beta_tot(v, bax, R[0], R[1], R[2]);
    const double biL = cell_beta_ind(v, bax, L[0], L[1], L[2]);
    const double biR = cell_beta_ind(v, bax, R[0], R[1], R[2]);
    const double ufL = cell_u_contra(v, fax, L[0], L[1], L[2]);
    const double ufR = cell_u_contra(v, fax, R[0], R[1], R[2]);
    const double ubL = cell_u_contra(v, bax, L[0], L[1], L[2]);
    const double ubR = cell_u_contra(v, bax, R[0], R[1], R[2]);
    const double GL = ufL * btL - ubL * bff;
    const double GR = ufR * btR - ubR * bff;
    const double radius = dmax(fabs(ufL), fabs(ufR));
    return 0.5 * (GL + GR) - 0.5 * radius * (biR - biL);
}

static void emf_view_of(const or_ctx *c, int b, emf_view *v)
{
    oblock *B = &c->b[b];
    v->c = c;
    v->b = b;
    for (int a = 0; a < 3; ++a)
    {
        v->B[a] = &B->f[OF_BXI + a];
        v->Badd[a] = &B->f[OF_BADDXI + a];
        v->JD[a] = &B->f[OF_JDXI + a];
    }
    v->Jac = &B->f[OF_JAC];
    v->Up = &B->f[OF_UPLUS];
}

/* AssembleOneDirectionEMF_ (Z4_Lunar/src/physics/LunarSolver_MagExplicit_Ideal.cpp:348-496) */
static void face_emf(or_ctx *c)
{
    for (int b = 0; b < c->d.nblocks; ++b)
    {
        oblock *B = &c->b[b];
        emf_view v;
        emf_view_of(c, b, &v);
        for (int fax = 0; fax < 3; ++fax)
        {
            fblk *Ef = &B->f[OF_EFXI + fax];
            OMP_FOR
            for (int i = 0; i < Ef->cnt[0]; ++i)
                for (int j = 0; j < Ef->cnt[1]; ++j)
                    for (int k = 0; k < Ef->cnt[2]; ++k)
                    {
                        const int R[3] = {i, j, k};
                        const int L[3] = {i - (fax == 0), j - (fax == 1), k - (fax == 2)};
                        const double bff = face_tot(&v, fax, i, j, k);
                        double e[3] = {0.0, 0.0, 0.0};
                        if (fax == 0)
                        {
                            const double Ge = emf_flux(&v, 1, 0, L, R, bff);
                            const double Gz = emf_flux(&v, 2, 0, L, R, bff);
                            e[1] = Gz;
                            e[2] = -Ge;
                        }
                        else if (fax == 1)
                        {
                            const double Gx = emf_flux(&v, 0, 1, L, R, bff);
                            const double Gz = emf_flux(&v, 2, 1, L, R, bff);
                            e[0] = -Gz;
                            e[2] = Gx;
                        }
                        else
                        {
                            const double Gx = emf_flux(&v, 0, 2, L, R, bff);
                            const double Ge = emf_flux(&v, 1, 2, L, R, bff);
                            e[0] = Ge;
                            e[1] = -Gx;
                        }
                        for (int m = 0; m < 3; ++m) AT(Ef, i, j, k, m) = e[m];
                    }
        }
    }
}

static double max_abs_u4(const emf_view *v, int ax, const int q[4][3])
{
    double a = 0.0;
    for (int n = 0; n < 4; ++n) a = dmax(a, fabs(cell_u_contra(v, ax, q[n][0], q[n][1], q[n][2])));
    return a;
}

/* AssembleEdgeEMF_FromFaceE_Ideal_ (Z4_Lunar/src/physics/LunarSolver_MagExplicit_Ideal.cpp:53-287) */
static void edge_emf(or_ctx *c)
{
    const double Cuct = 0.5;
    for (int b = 0; b < c->d.nblocks; ++b)
    {
        oblock *B = &c->b[b];
        emf_view v;
        emf_view_of(c, b, &v);
        fblk *Fx = &B->f[OF_EFXI], *Fy = &B->f[OF_EFET], *Fz = &B->f[OF_EFZE];
        fblk *Bx = v.B[0], *By = v.B[1], *Bz = v.B[2];
        fblk *Ex = &B->f[OF_EXI], *Ey = &B->f[OF_EET], *Ez = &B->f[OF_EZE];
        OMP_FOR
        for (int i = 0; i < Ex->cnt[0]; ++i)
            for (int j = 0; j < Ex->cnt[1]; ++j)
                for (int k = 0; k < Ex->cnt[2]; ++k)
                {
                    const double Ec = 0.25 * (AT(Fy, i, j, k, 0) + AT(Fy, i, j, k - 1, 0) + AT(Fz, i, j, k, 0) + AT(Fz, i, j - 1, k, 0));
                    const int q[4][3] = {{i, j, k}, {i, j - 1, k}, {i, j, k - 1}, {i, j - 1, k - 1}};
                    const double ae = max_abs_u4(&v, 1, q), az = max_abs_u4(&v, 2, q);
                    const double dBz = AT(Bz, i, j, k, 0) - AT(Bz, i, j - 1, k, 0);
                    const double dBe = AT(By, i, j, k, 0) - AT(By, i, j, k - 1, 0);
                    AT(Ex, i, j, k, 0) = Ec + Cuct * (0.5 * ae * dBz - 0.5 * az * dBe);
                }
        OMP_FOR
        for (int i = 0; i < Ey->cnt[0]; ++i)
            for (int j = 0; j < Ey->cnt[1]; ++j)
                for (int k = 0; k < Ey->cnt[2]; ++k)
                {
                    const double Ec = 0.25 * (AT(Fx, i, j, k, 1) + AT(Fx, i, j, k - 1, 1) + AT(Fz, i, j, k, 1) + AT(Fz, i - 1, j, k, 1));
                    const int q[4][3] = {{i, j, k}, {i - 1, j, k}, {i, j, k - 1}, {i - 1, j, k - 1}};
                    const double az = max_abs_u4(&v, 2, q), ax = max_abs_u4(&v, 0, q);
                    const double dBx = AT(Bx, i, j, k, 0) - AT(Bx, i, j, k - 1, 0);
                    const double dBz = AT(Bz, i, j, k, 0) - AT(Bz, i - 1, j, k, 0);
                    AT(Ey, i, j, k, 0) = Ec + Cuct * (0.5 * az * dBx - 0.5 * ax * dBz);
                }
        OMP_FOR
        for (int i = 0; i < Ez->cnt[0]; ++i)
            for (int j = 0; j < Ez->cnt[1]; ++j)
                for (int k = 0; k < Ez->cnt[2]; ++k)
                {
                    const double Ec = 0.25 * (AT(Fx, i, j, k, 2) + AT(Fx, i, j - 1, k, 2) + AT(Fy, i, j, k, 2) + AT(Fy, i - 1, j, k, 2));
                    const int q[4][3] = {{i, j, k}, {i - 1, j, k}, {i, j - 1, k}, {i - 1, j - 1, k}};
                    const double ax = max_abs_u4(&v, 0, q), ae = max_abs_u4(&v, 1, q);
                    const double dBe = AT(By, i, j, k, 0) - AT(By, i - 1, j, k, 0);
                    const double dBx = AT(Bx, i, j, k, 0) - AT(Bx, i, j - 1, k, 0);
                    AT(Ez, i, j, k, 0) = Ec + Cuct * (0.5 * ax * dBe - 0.5 * ae * dBx);
                }
    }
}

/* CTOperators::CurlEdgeToFace (Z4_Lunar/src/operators/CTOperators.cpp:6-56) */
static void curl_edges(or_ctx *c, double mult, int fe, int fr)
{
    for (int b = 0; b < c->d.nblocks; ++b)
    {
        oblock *B = &c->b[b];
        fblk *Ex = &B->f[fe], *Ey = &B->f[fe + 1], *Ez = &B->f[fe + 2];
        fblk *Rx = &B->f[fr], *Ry = &B->f[fr + 1], *Rz = &B->f[fr + 2];
        OMP_FOR
        for (int i = 0; i < Rx->cnt[0]; ++i)
            for (int j = 0; j < Rx->cnt[1]; ++j)
                for (int k = 0; k < Rx->cnt[2]; ++k)
                    AT(Rx, i, j, k, 0) = mult * ((AT(Ey, i, j, k, 0) - AT(Ey, i, j, k + 1, 0)) + (AT(Ez, i, j + 1, k, 0) - AT(Ez, i, j, k, 0)));
        OMP_FOR
        for (int i = 0; i < Ry->cnt[0]; ++i)
            for (int j = 0; j < Ry->cnt[1]; ++j)
                for (int k = 0; k < Ry->cnt[2]; ++k)
                    AT(Ry, i, j, k, 0) = mult * ((AT(Ex, i, j, k + 1, 0) - AT(Ex, i, j, k, 0)) + (AT(Ez, i, j, k, 0) - AT(Ez, i + 1, j, k, 0)));
        OMP_FOR
        for (int i = 0; i < Rz->cnt[0]; ++i)
            for (int j = 0; j < Rz->cnt[1]; ++j)
                for (int k = 0; k < Rz->cnt[2]; ++k)
                    AT(Rz, i, j, k, 0) = mult * ((AT(Ex, i, j, k, 0) - AT(Ex, i, j + 1, k, 0)) + (AT(Ey, i + 1, j, k, 0) - AT(Ey, i, j, k, 0)));
    }
}

static void curl_edge_to_face(or_ctx *c, double mult) { curl_edges(c, mult, OF_EXI, OF_RHSBXI); }

static void zero_fblk(fblk *F)
{
    memset(F->d, 0, sizeof(double) * (size_t)F->nc * F->ext[0] * F->ext[1] * F->ext[2]);
}

/* AssembleRHS_Induction_CT_ (Z4_Lunar/src/physics/LunarSolver_MagneticRHS.cpp:5-51) */
static void assemble_ct(or_ctx *c)
{
    for (int b = 0; b < c->d.nblocks; ++b)
        for (int a = 0; a < 3; ++a) zero_fblk(&c->b[b].f[OF_EXI + a]);
    face_emf(c);
    or_sync(c, "Eface");
    edge_emf(c);
    if (c->d.model == 1 && c->d.ambipolar) add_ambipolar_emf(c);
    if (c->d.model == 1 && c->d.local_arti) add_local_arti_emf(c);
    if (c->d.model == 1 && c->d.hall) add_hall_emf(c);
    or_sync(c, "Eedge");
    curl_edge_to_face(c, -1.0);
}

/* ApplyUpdate_Euler_ (Z4_Lunar/src/physics/LunarSolver_Advance.cpp:61-122) */
static void apply_update(or_ctx *c, double dt)
{
    for (int b = 0; b < c->d.nblocks; ++b)
    {
        oblock *B = &c->b[b];
        for (int sp = 0; sp < (c->d.model == 1 ? 2 : 1); ++sp)
        {
            fblk *U = &B->f[sp ? OF_UNA : OF_UH], *R = &B->f[sp ? OF_RHSNA : OF_RHSH];
            OMP_FOR
            for (int i = 0; i < B->n[0]; ++i)
                for (int j = 0; j < B->n[1]; ++j)
                    for (int k = 0; k < B->n[2]; ++k)
                        for (int m = 0; m < 5; ++m) AT(U, i, j, k, m) += dt * AT(R, i, j, k, m);
        }
        for (int a = 0; a < 3; ++a)
        {
            fblk *Bf = &B->f[OF_BXI + a], *RB = &B->f[OF_RHSBXI + a];
            OMP_FOR
            for (int i = 0; i < Bf->cnt[0]; ++i)
                for (int j = 0; j < Bf->cnt[1]; ++j)
                    for (int k = 0; k < Bf->cnt[2]; ++k) AT(Bf, i, j, k, 0) += dt * AT(RB, i, j, k, 0);
        }
    }
}

/* AddResistiveEdgeEMF_To_ (Z5_Mercury/src/physics/MercurySolver_MagExplicit_Resis.cpp:188-330):
 * E += inver_Rem yita0(r) (J_cell averaged over the 4 cells around the edge) . dr
 * where 0.84 < r < 1.04 at the edge midpoint. */
static void add_resistive_emf(or_ctx *c)
{
    for (int b = 0; b < c->d.nblocks; ++b)
    {
        oblock *B = &c->b[b];
        fblk *Jc = &B->f[OF_JCELL];
        for (int a = 0; a < 3; ++a)
        {
            fblk *E = &B->f[OF_ERXI + a];
            const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
            for (int i = 0; i < E->cnt[0]; ++i)
                for (int j = 0; j < E->cnt[1]; ++j)
                    for (int k = 0; k < E->cnt[2]; ++k)
                    {
                        double m[3];
                        edge_mid(B, a, i, j, k, m);
                        const double r = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
                        double yita0 = 0.0;
                        if (!(r <= 0.84 || r >= 1.04)) yita0 = 0.5 * (tanh((r - 0.84) / 0.02) - tanh((r - 1.04) / 0.02));
                        if (yita0 == 0.0) continue;
                        /* cells (i,j,k), minus t1, minus t2, minus both (get_edge_cells, :232-261) */
                        int q[4][3] = {{i, j, k}, {i, j, k}, {i, j, k}, {i, j, k}};
                        q[1][t1] -= 1;
                        q[2][t2] -= 1;
                        q[3][t1] -= 1;
                        q[3][t2] -= 1;
                        double J[3] = {0.0, 0.0, 0.0};
                        for (int n = 0; n < 4; ++n)
                            for (int m2 = 0; m2 < 3; ++m2) J[m2] += AT(Jc, q[n][0], q[n][1], q[n][2], m2);
                        for (int m2 = 0; m2 < 3; ++m2) J[m2] *= 0.25;
                        const int ip = i + (a == 0), jp = j + (a == 1), kp = k + (a == 2);
                        const double dr0 = NODE(B, 0, ip, jp, kp) - NODE(B, 0, i, j, k);
                        const double dr1 = NODE(B, 1, ip, jp, kp) - NODE(B, 1, i, j, k);
                        const double dr2 = NODE(B, 2, ip, jp, kp) - NODE(B, 2, i, j, k);
                        const double Jedge = J[0] * dr0 + J[1] * dr1 + J[2] * dr2;
                        AT(E, i, j, k, 0) += (c->d.inver_Rem * yita0) * Jedge;
                    }
        }
    }
}

int or_resistive_emf(or_ctx *c)
{
    for (int b = 0; b < c->d.nblocks; ++b)
        for (int a = 0; a < 3; ++a) zero_fblk(&c->b[b].f[OF_ERXI + a]);
    add_resistive_emf(c);
    return 0;
}

/* ResistiveDiffusionSubcycles_ (Z5_Mercury/src/physics/MercurySolver_MagneticRHS.cpp:79-141) */
static void resistive_subcycles(or_ctx *c, double dt)
{
    const int nsub = c->d.resist_nsub > 1 ? c->d.resist_nsub : 1;
    const double dt_res = dt / (double)nsub;
    for (int isub = 0; isub < nsub; ++isub)
    {
        or_sync(c, "Bface");
        calc_bcell(c);
        calc_jedge(c);
        calc_jcell(c);
        for (int b = 0; b < c->d.nblocks; ++b)
            for (int a = 0; a < 3; ++a)
            {
                zero_fblk(&c->b[b].f[OF_ERXI + a]);
                zero_fblk(&c->b[b].f[OF_RBRXI + a]);
            }
        add_resistive_emf(c);
        or_sync(c, "Eres");
        curl_edges(c, -1.0, OF_ERXI, OF_RBRXI);
        /* ApplyUpdate_Euler_BfaceOnly_ (MercurySolver_MagExplicit_Hall.cpp:441-479) */
        for (int b = 0; b < c->d.nblocks; ++b)
            for (int a = 0; a < 3; ++a)
            {
                fblk *Bf = &c->b[b].f[OF_BXI + a], *R = &c->b[b].f[OF_RBRXI + a];
                for (int i = 0; i < Bf->cnt[0]; ++i)
                    for (int j = 0; j < Bf->cnt[1]; ++j)
                        for (int k = 0; k < Bf->cnt[2]; ++k) AT(Bf, i, j, k, 0) += dt_res * AT(R, i, j, k, 0);
            }
        or_sync(c, "Bface");
    }
}

/* StepOnce (Z4_Lunar/src/LunarSolver_Driver.cpp:21-92) */
int or_step(or_ctx *c, double *dt_used)
{
    double dt;
    const int rc = or_compute_dt(c, &dt);
    if (dt_used) *dt_used = dt;
    if (rc) return rc;
    c->dt = dt;
    for (int b = 0; b < c->d.nblocks; ++b)
    {
        zero_fblk(&c->b[b].f[OF_RHSH]);
        zero_fblk(&c->b[b].f[OF_RHSNA]);
        for (int a = 0; a < 3; ++a) zero_fblk(&c->b[b].f[OF_RHSBXI + a]);
    }
    scheme_u(c);
    add_source(c);
    assemble_ct(c);
    apply_update(c, dt);
    if (c->d.model == 1 && c->d.resist) resistive_subcycles(c, dt);
    or_sync(c, "Ucell");
    or_sync(c, "Bface");
    or_update_derived(c);
    return 0;
}

/* One stage of or_step on the current arrays (tests), in or_step's order:
 * "dt" (Compute_Timestep into the step's dt), "fluid" (RHS zeroed, Scheme_U_,
 * AddSourceToRHS_Fluid), "face_emf" (E zeroed, the three face-EMF
 * directions), "edge_emf" (UCT edge EMF and the two-fluid EMF terms), "curl"
 * (CurlEdgeToFace(-1) into RHS_B), "update" (ApplyUpdate_Euler_ with the
 * step's dt and the resistive subcycles), "derived" (UpdateDerivedFields_);
 * syncs through or_sync.  Returns 0, -1 for an unknown stage, or the dt
 * error of "dt". */
int or_stage(or_ctx *c, const char *stage)
{
    if (!strcmp(stage, "dt"))
    {
        double dt;
        const int rc = or_compute_dt(c, &dt);
        c->dt = dt;
        return rc;
    }
    if (!strcmp(stage, "fluid"))
    {
        for (int b = 0; b < c->d.nblocks; ++b)
        {
            zero_fblk(&c->b[b].f[OF_RHSH]);
            zero_fblk(&c->b[b].f[OF_RHSNA]);
            for (int a = 0; a < 3; ++a) zero_fblk(&c->b[b].f[OF_RHSBXI + a]);
        }
        scheme_u(c);
        add_source(c);
        return 0;
    }
    if (!strcmp(stage, "face_emf"))
    {
        for (int b = 0; b < c->d.nblocks; ++b)
            for (int a = 0; a < 3; ++a) zero_fblk(&c->b[b].f[OF_EXI + a]);
        face_emf(c);
        return 0;
    }
    if (!strcmp(stage, "edge_emf"))
    {
        edge_emf(c);
        if (c->d.model == 1 && c->d.ambipolar) add_ambipolar_emf(c);
        if (c->d.model == 1 && c->d.local_arti) add_local_arti_emf(c);
        if (c->d.model == 1 && c->d.hall) add_hall_emf(c);
        return 0;
    }
    if (!strcmp(stage, "curl"))
    {
        curl_edge_to_face(c, -1.0);
        return 0;
    }
    if (!strcmp(stage, "update"))
    {
        apply_update(c, c->dt);
        if (c->d.model == 1 && c->d.resist) resistive_subcycles(c, c->dt);
        return 0;
    }
    if (!strcmp(stage, "derived")) return or_update_derived(c);
    return -1;
}

/* PrintMinMaxDiagnostics_ (Z4_Lunar/src/tools/LunarSolver_PrintDiag.cpp:7-60,
 * Z5_Mercury/src/tools/MercurySolver_PrintDiag.cpp:7-92) */
void or_diagnostics(or_ctx *c, double *out)
{
    for (int q = 0; q < 8; ++q) { out[q] = INFINITY; out[8 + q] = -INFINITY; }
    out[16] = -INFINITY;
    for (int b = 0; b < c->d.nblocks; ++b)
    {
        oblock *B = &c->b[b];
        fblk *UH = &B->f[OF_UH], *PH = &B->f[OF_PV], *UN = &B->f[OF_UNA], *PN = &B->f[OF_PVNA], *Bc = &B->f[OF_BCELL];
        fblk *Bx = &B->f[OF_BXI], *By = &B->f[OF_BET], *Bz = &B->f[OF_BZE], *Jac = &B->f[OF_JAC];
        for (int k = 0; k < B->n[2]; ++k)
            for (int j = 0; j < B->n[1]; ++j)
                for (int i = 0; i < B->n[0]; ++i)
                {
                    const double bx = AT(Bc, i, j, k, 0), by = AT(Bc, i, j, k, 1), bz = AT(Bc, i, j, k, 2);
                    const double x[8] = {AT(UH, i, j, k, 0), AT(PH, i, j, k, 3), AT(UN, i, j, k, 0), AT(PN, i, j, k, 3),
                                         bx, by, bz, 0.5 * (bx * bx + by * by + bz * bz)};
                    for (int q = 0; q < 8; ++q)
                    {
                        if (c->d.model != 1 && (q == 2 || q == 3)) continue;
                        out[q] = dmin(out[q], x[q]);
                        out[8 + q] = dmax(out[8 + q], x[q]);
                    }
                    const double dv = (AT(Bx, i + 1, j, k, 0) - AT(Bx, i, j, k, 0) + AT(By, i, j + 1, k, 0) - AT(By, i, j, k, 0) +
                                       AT(Bz, i, j, k + 1, 0) - AT(Bz, i, j, k, 0)) / AT(Jac, i, j, k, 0);
                    out[16] = dmax(out[16], fabs(dv));
                }
    }
}

double or_divb_max(or_ctx *c)
{
    double mx = 0.0;
    for (int b = 0; b < c->d.nblocks; ++b)
    {
        oblock *B = &c->b[b];
        fblk *Bx = &B->f[OF_BXI], *By = &B->f[OF_BET], *Bz = &B->f[OF_BZE], *Jac = &B->f[OF_JAC];
        for (int i = 0; i < B->n[0]; ++i)
            for (int j = 0; j < B->n[1]; ++j)
                for (int k = 0; k < B->n[2]; ++k)
                {
                    const double dv = (AT(Bx, i + 1, j, k, 0) - AT(Bx, i, j, k, 0) + AT(By, i, j + 1, k, 0) - AT(By, i, j, k, 0) +
                                       AT(Bz, i, j, k + 1, 0) - AT(Bz, i, j, k, 0)) / AT(Jac, i, j, k, 0);
                    mx = dmax(mx, fabs(dv));
                }
    }
    return mx;
}
