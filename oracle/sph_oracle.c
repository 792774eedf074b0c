/*
 * sph_oracle.c -- CPU FP64 restatement of sph07's per-timestep particle loop
 * (MSolution::Solution, MSolution.cpp:238-324) and of the initial-state
 * pipeline (MSimulationInit::Initialize, MSimulationInit.cpp:31-99).
 *
 * TEST INFRASTRUCTURE ONLY: it is the checker for the HIP product path and
 * the CPU baseline timed beside it.  Nothing in the product links or calls it.
 *
 * Written from the reference's behaviour, not its code: plain C arrays, the
 * reference's 1-based particle numbering, its linked-list cell order and its
 * floating-point evaluation order (every expression below keeps the operand
 * order of the cited line; compile with -ffp-contract=off so no FMA is formed).
 *
 * Pinning (what checks this file, beyond reading it against the reference;
 * each check fails for a planted mistake of the kind it names):
 *   - kernel W/dW: closed form, partition of unity, continuity at z=1,2,
 *     derivative check (tests/test_oracle_kernel.py);
 *   - grid, linked lists, neighbour sets incl. periodic images: brute-force
 *     O(N^2) search, lattice counts, list order (tests/test_oracle_neighbours.py);
 *   - strain: tr D of a linear field (tests/test_oracle_physics.py); the
 *     half-step shift from the closed-form kernel gradient of two approaching
 *     particles (tests/test_oracle_pins.py);
 *   - RhoUpdate: one-step uniform dilation, rho0 (1 - D eps dt);
 *   - EOS 29 / 28 with the pressure term: hydrostatic balance of a column
 *     (a = g rho0/rho), p(rho0) = 0, c^2 = dp/drho;
 *   - momentum: pair antisymmetry (sum m a = 0), Poiseuille start-up series
 *     (Morris 1997) for the whole loop; Monaghan term: zero for separating
 *     pairs, dissipative otherwise;
 *   - CSF: Laplace law of a relaxed droplet, |n|-weighted curvature -1/R;
 *   - rigid body: force and torque balance with every other particle,
 *     rigid-body kinematics v = V + Omega x r;
 *   - critical step: closed forms of options 1, 2 (incl. the step-1 1e-9)
 *     and 3 (sound, viscous, acceleration, capillary terms); integration and
 *     the dt rule: one-step closed forms of a lone particle.
 * The reference itself cannot be built here (it needs Boost uBLAS, which is
 * not installed), so no compiled-reference vectors exist; parity for the
 * uBLAS "v = 0" paths (CSF normals, rigid A/ALPHA) follows the evident intent
 * (zero fill) and is pinned by the invariants above.
 */
#include "sph_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI       3.141592654      /* global.h:12 (truncated literal) */
#define ORC_LLEND    (-1000000)       /* global.h:13 */
#define ORC_MAXMAT   16

#define SQR(x)   ((x) * (x))
#define CUBE(x)  ((x) * (x) * (x))
#define MINV(x, y) (((x) < (y)) ? (x) : (y))
#define MAXV(x, y) (((x) > (y)) ? (x) : (y))
#define AVG(x, y) (0.5 * ((x) + (y)))

typedef struct {
    double x[3], v[3];
    double m, h, rho, c, vabs;
    int mat, color, par, llp;
    int shift[3];             /* image shift in periods, for canonical ids */
} orc_ghost;

struct orc_scene {
    /* options */
    int ndim, disctype, tcrit_opt, maxnbr, st, rigid, boundary, baseaccel;
    double tssf, endtime, itss;
    int bcode[2][3];
    double bx[2][3];
    int btype[3];
    double base_a[3];
    double minadd[3][3], maxadd[3][3];
    int nummat;
    orc_material mat[ORC_MAXMAT + 1];   /* 1-based */
    /* rigid body (MRigidBody.h:18-30) */
    int rb_id;
    double rb_M, rb_I, rbR[3], rbV[3], rbW[3], rbA[3], rbAL[3];
    /* globals (MGlobalVars.h:18-31) */
    double dt, dtold, ptime, critts;
    int timestep;
    double cmm[2][3];
    double hmax;
    int done;
    /* grid (MNeighbourVars.h:19-28) */
    double gsize[3], gmin[3];
    int glim[3];
    int *grid;
    size_t gcap;
    /* particles, 1-based (index 0 unused) */
    int np;
    double *x, *v, *a;                  /* 3 per particle */
    double *rho, *rhoold, *rho0, *p, *c, *m, *h, *vabs, *critts_p, *trd, *mindist;
    double *normal, *nnormal, *N, *kappa;
    double *rod, *sig, *q;              /* disctype 0: 3x3 row-major rod and sigma, q (diagonal) */
    int *mat_id, *color, *dispbc, *active, *llp, *nnbr, *nbr;
    /* ghosts, 1-based */
    int ngp, gcapn;
    orc_ghost *g;
    char err[256];
};

static int fail(orc_scene *s, const char *msg)
{
    snprintf(s->err, sizeof s->err, "%s", msg);
    return 1;
}

static int rigid_p(const orc_scene *s, int mat)
{
    /* MSolution.h:59-60; with the rigid body off the reference compares
       against an uninitialised id, taken here as "no body". */
    return s->rigid && mat == s->rb_id;
}

/* ---- kernel: MKernelBSpline::GradW (static), MKernelBSpline.cpp:137-187 ---- */
static void orc_gradw_r(double dwdx[3], double *dwdr, const double *xi, const double *xj, int ndim,
                        double havg);
static void orc_gradw(double dwdx[3], const double *xi, const double *xj, int ndim, double havg)
{
    double dwdr;
    orc_gradw_r(dwdx, &dwdr, xi, xj, ndim, havg);
}

/* the full static GradW: dWdx and dWdr (dW/havg, 0 at r = 0, :174-185) */
static void orc_gradw_r(double dwdx[3], double *dwdr, const double *xi, const double *xj, int ndim,
                        double havg)
{
    double const1 = 0.0, const2, xij[3], z, rad, dW;
    int n;
    dwdx[0] = dwdx[1] = dwdx[2] = 0.0;
    switch (ndim) {
    case 3: const1 = 1.0 / (ORC_PI * CUBE(havg)); break;
    case 2: const1 = 10.0 / (7.0 * ORC_PI * SQR(havg)); break;
    case 1: const1 = 2.0 / (3.0 * havg); break;
    }
    const2 = 0.25 * const1;
    for (n = 0, rad = 0.0; n < ndim; n++) {
        xij[n] = xi[n] - xj[n];
        rad += SQR(xij[n]);
    }
    rad = sqrt(rad);
    z = rad / havg;
    if (z < 1.0)
        dW = const1 * ((9.0 / 4.0 * z - 3.0) * z);
    else if (z < 2.0)
        dW = -3.0 * const2 * SQR(2.0 - z);
    else
        dW = 0.0;
    for (n = 0; n < ndim; n++)
        dwdx[n] = (rad == 0.0) ? 0.0 : dW * xij[n] / (rad * havg);
    *dwdr = (rad == 0.0) ? 0.0 : dW / havg;
}

/* MKernelBSpline::Kernel (static), MKernelBSpline.cpp:98-135 */
static double orc_w(const double *xi, const double *xj, int ndim, double havg)
{
    double const1 = 0.0, const2, xij, z, rad;
    int n;
    switch (ndim) {
    case 3: const1 = 1.0 / (ORC_PI * CUBE(havg)); break;
    case 2: const1 = 10.0 / (7.0 * ORC_PI * SQR(havg)); break;
    case 1: const1 = 2.0 / (3.0 * havg); break;
    }
    const2 = 0.25 * const1;
    for (n = 0, rad = 0.0; n < ndim; n++) {
        xij = xi[n] - xj[n];
        rad += SQR(xij);
    }
    rad = sqrt(rad);
    z = rad / havg;
    if (z < 1.0) return const1 * (((0.75 * z - 1.5) * z) * z + 1.0);
    if (z < 2.0) return const2 * CUBE(2.0 - z);
    return 0.0;
}

/* Exported for the kernel closed-form tests. */
double orc_kernel_w(const double *xi, const double *xj, int ndim, double havg)
{
    return orc_w(xi, xj, ndim, havg);
}
void orc_kernel_gradw(double *dwdx, const double *xi, const double *xj, int ndim, double havg)
{
    orc_gradw(dwdx, xi, xj, ndim, havg);
}
double orc_kernel_dwdr(const double *xi, const double *xj, int ndim, double havg)
{
    double dwdx[3], dwdr;
    orc_gradw_r(dwdx, &dwdr, xi, xj, ndim, havg);
    return dwdr;
}

/* ---- neighbour record access: id>0 real, id<0 ghost (MSolution.cpp:187-188) ---- */
typedef struct {
    const double *x, *v;
    double m, h, rho, p, c, N, q;
    const double *nn, *sig;   /* sig: 3x3 row-major (disctype 0) */
    int mat, color;
} orc_view;

static const double orc_zero3[3] = {0.0, 0.0, 0.0};
static const double orc_zero9[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};

static orc_view view_of(const orc_scene *s, int id)
{
    orc_view r;
    if (id > 0) {
        r.x = &s->x[3 * id]; r.v = &s->v[3 * id];
        r.m = s->m[id]; r.h = s->h[id]; r.rho = s->rho[id]; r.p = s->p[id];
        r.c = s->c[id]; r.N = s->N[id]; r.nn = &s->nnormal[3 * id];
        r.mat = s->mat_id[id]; r.color = s->color[id];
        r.sig = s->sig ? &s->sig[9 * (size_t)id] : orc_zero9;
        r.q = s->q ? s->q[id] : 0.0;
    } else {
        const orc_ghost *g = &s->g[-id];
        r.x = g->x; r.v = g->v; r.m = g->m; r.h = g->h; r.rho = g->rho;
        /* ghosts keep constructor zeros for p, N and normals
           (MGhostParticleData.cpp:16-28, MCommonParticle.cpp:12,17-19) */
        r.p = 0.0; r.N = 0.0; r.nn = orc_zero3; r.c = g->c;
        r.mat = g->mat; r.color = g->color;
        r.sig = orc_zero9; r.q = 0.0;       /* sigma and q: not copied either */
    }
    return r;
}

/* ---- grid: MNeighbourVars::InitGrid / GetBoxCoords, MNeighbourVars.cpp:18-89 ---- */
static int init_grid(orc_scene *s)
{
    int i;
    size_t ncell;
    for (i = 0; i < 3; i++) {
        s->gmin[i] = 0.0;
        s->glim[i] = 1;
        s->gsize[i] = 1.0 / (2.0 * s->hmax);
    }
    for (i = 0; i < s->ndim; i++) {
        if (s->btype[i] == 1) {
            s->gmin[i] = s->cmm[0][i] - 0.01 * s->hmax;
            s->glim[i] = (int)(s->gsize[i] * (s->cmm[1][i] - s->gmin[i])) + 1;
        } else if (s->btype[i] == 4) {
            s->glim[i] = (int)(s->gsize[i] * (s->bx[1][i] - s->bx[0][i]));
            s->glim[i] = MAXV(s->glim[i], 0);
            s->gsize[i] = s->glim[i] / (s->bx[1][i] - s->bx[0][i]);
            s->gmin[i] = s->bx[0][i] - 1.0 / s->gsize[i];
            s->glim[i] += 2;
        } else {
            return fail(s, "grid: boundary type not supported");
        }
    }
    if (s->glim[0] <= 0 || s->glim[1] <= 0 || s->glim[2] <= 0)
        return fail(s, "grid: empty grid");
    ncell = (size_t)s->glim[0] * s->glim[1] * s->glim[2];
    if (ncell > ((size_t)1 << 31)) return fail(s, "grid: too many cells");
    if (ncell > s->gcap) {
        free(s->grid);
        s->grid = (int *)malloc(ncell * sizeof(int));
        if (!s->grid) return fail(s, "grid: out of memory");
        s->gcap = ncell;
    }
    for (size_t c = 0; c < ncell; c++) s->grid[c] = ORC_LLEND;
    return 0;
}

static int cell_index(const orc_scene *s, int i, int j, int k)
{
    return s->glim[2] * (s->glim[1] * i + j) + k;   /* MNeighbourVars.cpp:72 */
}

static void box_coords(const orc_scene *s, const double *x, int bc[3])
{
    for (int i = 0; i < 3; i++) bc[i] = (int)(s->gsize[i] * (x[i] - s->gmin[i]));
}

static int in_grid(const orc_scene *s, const int bc[3])
{
    for (int i = 0; i < 3; i++)
        if (bc[i] < 0 || bc[i] >= s->glim[i]) return 0;
    return 1;
}

/* MSolution::SetupLinkedList, MSolution.cpp:116-148 (one rank: the process
   slab is the whole grid, InitializeMpiProcess :1066-1087) */
static int setup_linked_list(orc_scene *s)
{
    int bc[3];
    if (init_grid(s)) return 1;
    for (int i = 1; i <= s->np; i++) {
        if (!s->active[i]) { s->llp[i] = ORC_LLEND; continue; }
        box_coords(s, &s->x[3 * i], bc);
        if (s->timestep == 1 && (bc[0] < 0 || bc[0] > s->glim[0] - 1)) {
            s->active[i] = 0;
            continue;
        }
        if (!in_grid(s, bc)) return fail(s, "neighbours: particle outside the grid");
        int c = cell_index(s, bc[0], bc[1], bc[2]);
        s->llp[i] = s->grid[c];
        s->grid[c] = i;
    }
    return 0;
}

/* MGhostParticleData::CreateGhostFromReal, MGhostParticleData.cpp:12-36 */
static int create_ghost(orc_scene *s, int id, const double add[3], const int shift[3], int cell)
{
    if (s->ngp + 1 >= s->gcapn) {
        int nc = s->gcapn ? 2 * s->gcapn : 1024;
        orc_ghost *ng = (orc_ghost *)realloc(s->g, (size_t)nc * sizeof(orc_ghost));
        if (!ng) return fail(s, "ghosts: out of memory");
        s->g = ng;
        s->gcapn = nc;
    }
    orc_ghost *g = &s->g[++s->ngp];
    for (int n = 0; n < 3; n++) {
        g->x[n] = s->x[3 * id + n] + add[n];
        g->v[n] = s->v[3 * id + n];
        g->shift[n] = shift[n];
    }
    g->m = s->m[id]; g->h = s->h[id]; g->rho = s->rho[id]; g->c = s->c[id];
    g->vabs = s->vabs[id]; g->mat = s->mat_id[id]; g->color = s->color[id];
    g->par = id;
    g->llp = s->grid[cell];
    s->grid[cell] = -s->ngp;
    return 0;
}

/* MSolution::GenerateGhostParticles, MSolution.cpp:878-918 */
static int generate_ghosts(orc_scene *s, const int lo[3], const int hi[3], const double add[3])
{
    int off[3] = {0, 0, 0}, shift[3] = {0, 0, 0};
    int min_lim, max_lim;
    if (s->bcode[0][0] != 2) { min_lim = 0; max_lim = s->glim[0] - 1; }
    else { min_lim = lo[0]; max_lim = hi[0]; }
    for (int n = 0; n < 3; n++)
        if (add[n] != 0.0) {
            int sg = (add[n] >= 0) ? 1 : -1;
            off[n] = sg * (s->glim[n] - 2);
            shift[n] = sg;
        }
    for (int k = lo[2]; k <= hi[2]; k++)
        for (int j = lo[1]; j <= hi[1]; j++)
            for (int i = min_lim; i <= max_lim; i++) {
                int id = s->grid[cell_index(s, i, j, k)];
                if (id < 0) continue;
                int tc = cell_index(s, i + off[0], j + off[1], k + off[2]);
                while (id > 0) {
                    if (create_ghost(s, id, add, shift, tc)) return 1;
                    id = s->llp[id];
                }
            }
    return 0;
}

/* MSolution::GhostSetup, MSolution.cpp:811-875 */
static int ghost_setup(orc_scene *s)
{
    int lo[3], hi[3];
    double add[3];
    s->ngp = 0;
    for (int m = 0; m < 3; m++) {
        for (int side = 0; side < 2; side++) {
            /* Xmin/Xmax both test code(0,0); Y/Z test their own side */
            int code = (m == 0) ? s->bcode[0][0] : s->bcode[side][m];
            if (code != 2) continue;
            for (int n = 0; n < 3; n++) {
                lo[n] = 0;
                hi[n] = s->glim[n] - 1;
                add[n] = side == 0 ? s->minadd[n][m] : s->maxadd[n][m];
            }
            lo[m] = hi[m] = (side == 0) ? 1 : s->glim[m] - 2;
            if (generate_ghosts(s, lo, hi, add)) return 1;
        }
    }
    return 0;
}

/* MSolution::LinkedListNeighbours, MSolution.cpp:150-215 */
static int linked_list_neighbours(orc_scene *s)
{
    int overflow = 0;
    for (int i = 1; i <= s->np; i++) s->nnbr[i] = 0;
    /* particles are independent: OpenMP threads over i give the serial result */
#pragma omp parallel for schedule(dynamic, 256) reduction(|: overflow)
    for (int i = 1; i <= s->np; i++) {
        int bc[3], lo[3], hi[3];
        if (!s->active[i] || overflow) continue;
        const double *xi = &s->x[3 * i];
        s->mindist[i] = 1.0e20;
        box_coords(s, xi, bc);
        for (int m = 0; m < 3; m++) {
            lo[m] = MAXV(bc[m] - 1, 0);
            hi[m] = MINV(bc[m] + 1, s->glim[m] - 1);
        }
        for (int m = s->ndim; m < 3; m++) lo[m] = hi[m] = 0;
        for (int r = lo[0]; r <= hi[0]; r++)
            for (int q = lo[1]; q <= hi[1]; q++)
                for (int t = lo[2]; t <= hi[2]; t++) {
                    int id = s->grid[cell_index(s, r, q, t)];
                    while (id != ORC_LLEND) {
                        orc_view pj = view_of(s, id);
                        int next = (id > 0) ? s->llp[id] : s->g[-id].llp;
                        if (id == i) { id = next; continue; }
                        double d2 = 0.0;
                        for (int n = 0; n < s->ndim; n++) d2 += SQR(xi[n] - pj.x[n]);
                        double hav = 2.0 * AVG(s->h[i], pj.h);
                        if (d2 < SQR(hav)) {
                            s->mindist[i] = MINV(s->mindist[i], sqrt(d2));
                            if (s->nnbr[i] < s->maxnbr) {
                                s->nbr[(size_t)i * s->maxnbr + s->nnbr[i]++] = id;
                            } else {
                                overflow = 1;
                                id = ORC_LLEND;
                                r = hi[0] + 1; q = hi[1] + 1; t = hi[2] + 1;
                                continue;
                            }
                        }
                        id = next;
                    }
                }
    }
    if (overflow) return fail(s, "ERROR: Neighbour list exceeded maximum.");
    return 0;
}

static int phase_neighbours(orc_scene *s, int with_ghosts)
{
    if (setup_linked_list(s)) return 1;
    if (with_ghosts && s->boundary) {
        if (ghost_setup(s)) return 1;
    } else {
        s->ngp = 0;
    }
    return linked_list_neighbours(s);
}

/* MSolution::Strain + ClassicStrain, MSolution.cpp:327-413 (trace only is
   used by the Morris discretisation; MParticle::CalcTraceROD :43-47) */
static int phase_strain(orc_scene *s)
{
    for (int i = 1; i <= s->np; i++) {
        if (!s->active[i]) continue;
        for (int n = 0; n < 3; n++) s->x[3 * i + n] = s->x[3 * i + n] - 0.5 * s->v[3 * i + n] * s->dt;
    }
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 1; i <= s->np; i++) {
        double dwdx[3];
        if (!s->active[i]) continue;
        const double *xi = &s->x[3 * i], *vi = &s->v[3 * i];
        /* grad v (MSolution.cpp:397-398); Morris needs its diagonal only */
        double g[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
        const int full = s->disctype == 0;
        for (int k = 0; k < s->nnbr[i]; k++) {
            int id = s->nbr[(size_t)i * s->maxnbr + k];
            orc_view pj = view_of(s, id);
            double volj = pj.m / pj.rho;
            double havg = AVG(s->h[i], pj.h);
            orc_gradw(dwdx, xi, pj.x, s->ndim, havg);
            if (full) {
                for (int r = 0; r < s->ndim; r++)
                    for (int c = 0; c < s->ndim; c++) g[r][c] += volj * (pj.v[r] - vi[r]) * dwdx[c];
            } else {
                for (int r = 0; r < s->ndim; r++) g[r][r] += volj * (pj.v[r] - vi[r]) * dwdx[r];
            }
        }
        /* rod = 0.5*(g + g^T), spin unused (:406-407); trace over ndim
           (MParticle::CalcTraceROD, MParticle.cpp:43-47) */
        if (full)
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) s->rod[9 * (size_t)i + 3 * r + c] = 0.5 * (g[r][c] + g[c][r]);
        double tr = 0.0;
        for (int r = 0; r < s->ndim; r++) tr += 0.5 * (g[r][r] + g[r][r]);
        s->trd[i] = tr;
    }
    for (int i = 1; i <= s->np; i++) {
        if (!s->active[i]) continue;
        for (int n = 0; n < 3; n++) s->x[3 * i + n] = s->x[3 * i + n] + 0.5 * s->v[3 * i + n] * s->dt;
    }
    return 0;
}

/* MSolution::CalculateSurfaceNormals, MSolution.cpp:1429-1493 (Morris 2000) */
static int phase_normals(orc_scene *s)
{
    double dwdx[3];
    if (!s->st) return 0;
    for (int i = 1; i <= s->np; i++) {
        if (!s->active[i]) continue;
        const double *xi = &s->x[3 * i];
        double *nrm = &s->normal[3 * i];
        nrm[0] = nrm[1] = nrm[2] = 0.0;
        for (int k = 0; k < s->nnbr[i]; k++) {
            orc_view pj = view_of(s, s->nbr[(size_t)i * s->maxnbr + k]);
            double volj = pj.m / pj.rho;
            double havg = AVG(s->h[i], pj.h);
            orc_gradw(dwdx, xi, pj.x, s->ndim, havg);
            double f = volj * (double)(pj.color - s->color[i]);
            for (int n = 0; n < 3; n++) nrm[n] += f * dwdx[n];
        }
        double in = 0.0;
        for (int k = 0; k < s->ndim; ++k) in += SQR(nrm[k]);
        in = sqrt(in);
        s->N[i] = in > 1.e-2 * s->h[i] ? 1 : 0;
        for (int n = 0; n < 3; n++)
            s->nnormal[3 * i + n] = (s->N[i] == 1) ? (1. / in) * nrm[n] : 0.0;
    }
    for (int i = 1; i <= s->np; i++) {
        if (!s->active[i]) continue;
        const double *xi = &s->x[3 * i], *nni = &s->nnormal[3 * i];
        double div = 0.0;
        for (int k = 0; k < s->nnbr[i]; k++) {
            orc_view pj = view_of(s, s->nbr[(size_t)i * s->maxnbr + k]);
            double volj = pj.m / pj.rho;
            double havg = AVG(s->h[i], pj.h);
            orc_gradw(dwdx, xi, pj.x, s->ndim, havg);
            double prod = 0;
            for (int n = 0; n < s->ndim; ++n) prod += (pj.nn[n] - nni[n]) * dwdx[n];
            div += MINV(s->N[i], pj.N) * volj * prod;
        }
        s->kappa[i] = div;
    }
    return 0;
}

/* MSolution::Constitutive + RhoUpdate, MSolution.cpp:436-477; model 9 EOS
   part of MConstitutive::Constitutive, MConstitutive.cpp:10-42,135-186.
   Deviatoric stress, q and energy are not read by the Morris momentum form. */
static int phase_constitutive(orc_scene *s)
{
    for (int i = 1; i <= s->np; i++) {
        if (!s->active[i]) continue;
        s->rhoold[i] = s->rho[i];
        s->rho[i] *= (1.0 - s->trd[i] * s->dt);
    }
    for (int i = 1; i <= s->np; i++) {
        if (!s->active[i]) continue;
        const orc_material *mt = &s->mat[s->mat_id[i]];
        if (mt->eos == 29) {
            s->c[i] = mt->eosinput[0];
            s->p[i] = SQR(s->c[i]) * (s->rho[i] - s->rho0[i]);
        } else if (mt->eos == 28) {
            double b = mt->eosinput[0], gamma = mt->eosinput[1];
            s->p[i] = b * (pow(s->rho[i] / s->rho0[i], gamma) - 1.0);
            s->c[i] = sqrt(b * gamma / s->rho0[i]);
        } else {
            return fail(s, "constitutive: EOS not supported");
        }
    }
    return 0;
}

/* Model 9 for the Classic discretisation, MConstitutive.cpp:10-42 per
   particle: F3DM9 (:44-68) deviatoric stress from rod, AViscosity (:123-133)
   q only when tr D < 0 (else q keeps its value), the EOS, StressUp
   (:210-222) sigma = -p I + s.  HieUpdate (:225-262) only moves the trial
   energy, which EOS 28/29 never read.  After RhoUpdate (MSolution.cpp:444). */
static int phase_constitutive_classic(orc_scene *s)
{
    for (int i = 1; i <= s->np; i++) {
        if (!s->active[i]) continue;
        s->rhoold[i] = s->rho[i];
        s->rho[i] *= (1.0 - s->trd[i] * s->dt);
    }
    for (int i = 1; i <= s->np; i++) {
        if (!s->active[i]) continue;
        const orc_material *mt = &s->mat[s->mat_id[i]];
        const double *rod = &s->rod[9 * (size_t)i];
        double *sg = &s->sig[9 * (size_t)i];
        double sdev[9];
        /* F3DM9: xmu = av_l * rho of the material (:49-52) */
        const double xmu = mt->av_l * mt->rho;
        if (fabs(xmu) < 1.e-10) {
            for (int k = 0; k < 9; k++) sdev[k] = 0.0;
        } else {
            const double davg = 1.0 / 3.0 * s->trd[i];
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) sdev[3 * r + c] = 2 * xmu * (rod[3 * r + c] - (r == c ? davg : 0.0));
        }
        /* AViscosity: c is the particle's sound speed before this EOS call */
        if (s->trd[i] < 0.0)
            s->q[i] = mt->av_l * s->rho[i] * s->c[i] * s->h[i] * fabs(s->trd[i]) +
                      mt->av_q * s->rho[i] * SQR(s->h[i] * s->trd[i]);
        if (mt->eos == 29) {
            s->c[i] = mt->eosinput[0];
            s->p[i] = SQR(s->c[i]) * (s->rho[i] - s->rho0[i]);
        } else if (mt->eos == 28) {
            double b = mt->eosinput[0], gamma = mt->eosinput[1];
            s->p[i] = b * (pow(s->rho[i] / s->rho0[i], gamma) - 1.0);
            s->c[i] = sqrt(b * gamma / s->rho0[i]);
        } else {
            return fail(s, "constitutive: EOS not supported");
        }
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) sg[3 * r + c] = -(r == c ? s->p[i] : 0.0) + sdev[3 * r + c];
    }
    return 0;
}

/* MSolution::CalcCriticalTimestep, MSolution.cpp:34-94 */
static int calc_critical_timestep(orc_scene *s)
{
    if (s->tcrit_opt < 1 || s->tcrit_opt > 3) return fail(s, "ERROR: Wrong critical time step option.");
    double critts = 1.e20;
#pragma omp parallel for schedule(static) reduction(min: critts)
    for (int i = 1; i <= s->np; i++) {
        double ct;
        const orc_material *mt = &s->mat[s->mat_id[i]];
        switch (s->tcrit_opt) {
        case 1:
            ct = s->h[i] / (s->c[i] + s->vabs[i]);
            if (s->timestep == 1) ct = 1.e-09;
            break;
        case 2:
            ct = s->mindist[i] / (s->c[i] + s->vabs[i]);
            if (s->timestep == 1) ct = 1.e-09;
            break;
        case 3: {
            double visc = mt->av_l, sig = mt->surface_tension, aabs = 0.0;
            ct = MINV(0.25 * s->h[i] / s->c[i], 0.125 * SQR(s->h[i]) / visc);
            for (int j = 0; j < s->ndim; j++) aabs += SQR(s->a[3 * i + j]);
            aabs = sqrt(aabs);
            ct = MINV(ct, 0.25 * sqrt(s->h[i] / aabs));
            if (s->st) ct = MINV(ct, 0.25 * sqrt(s->rho[i] * CUBE(s->h[i]) / 2 / ORC_PI / sig));
            break;
        }
        default:
            ct = 1.e20;
            break;
        }
        s->critts_p[i] = ct;
        if (s->active[i] && !rigid_p(s, s->mat_id[i]) && ct < critts) critts = ct;
    }
    s->critts = critts;
    return 0;
}

static int phase_timestep(orc_scene *s)
{
    if (calc_critical_timestep(s)) return 1;
    s->dtold = s->dt;                                         /* MSolution.cpp:275-277 */
    s->dt = MINV(s->tssf * s->critts, 1.1 * s->dtold);
    s->ptime += s->dt;
    return 0;
}

/* MSolution::RigidBodyAcceleration, MSolution.cpp:1496-1564;
   MRigidBody::CalculateCenterOfMass, MRigidBody.cpp:37-49 */
static int rigid_acceleration(orc_scene *s)
{
    double R[3] = {0.0, 0.0, 0.0};
    int count = 0;
    for (int i = 1; i <= s->np; i++)
        if (s->mat_id[i] == s->rb_id) {
            for (int n = 0; n < 3; n++) R[n] += s->x[3 * i + n];
            count++;
        }
    if (count == 0) return fail(s, "rigid body: no particles");
    for (int n = 0; n < 3; n++) s->rbR[n] = R[n] * (1. / count);
    double A[3] = {0.0, 0.0, 0.0}, AL[3] = {0.0, 0.0, 0.0};
    for (int k = 1; k <= s->np; k++) {
        if (!s->active[k] || !rigid_p(s, s->mat_id[k])) continue;
        const double *ak = &s->a[3 * k];
        double rl[3], mk = s->m[k];
        for (int n = 0; n < 3; n++) rl[n] = s->x[3 * k + n] - s->rbR[n];
        for (int n = 0; n < 3; n++) A[n] += mk * (0.0 + ak[n]);
        AL[0] += mk * (rl[1] * ak[2] - rl[2] * ak[1]);
        AL[1] += mk * (rl[2] * ak[0] - rl[0] * ak[2]);
        AL[2] += mk * (rl[0] * ak[1] - rl[1] * ak[0]);
    }
    for (int n = 0; n < 3; n++) {
        A[n] *= 1. / s->rb_M;
        AL[n] *= 1. / s->rb_I;
    }
    for (int n = 0; n < 3; n++) {
        s->rbA[n] = A[n];
        s->rbAL[n] = AL[n];
        if (s->baseaccel) s->rbA[n] += s->base_a[n];
    }
    return 0;
}

/* MSolution::MorrisAcceleration, MSolution.cpp:550-600.  visc_type 2 adds the
   Monaghan (1992) artificial viscosity term, an extension the reference lacks
   (its visc_type field is parsed and unused, MSimulationData.cpp:454-455). */
static int phase_momentum(orc_scene *s)
{
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 1; i <= s->np; i++) {
        double dwdx[3];
        double *ai = &s->a[3 * i];
        ai[0] = ai[1] = ai[2] = 0.0;
        if (!s->active[i]) continue;
        const double *xi = &s->x[3 * i], *vi = &s->v[3 * i];
        const orc_material *mi = &s->mat[s->mat_id[i]];
        double nu_i = mi->av_l, rho_i = s->rho[i], p_i = s->p[i];
        int rig_i = rigid_p(s, s->mat_id[i]);
        for (int k = 0; k < s->nnbr[i]; k++) {
            orc_view pj = view_of(s, s->nbr[(size_t)i * s->maxnbr + k]);
            const orc_material *mj = &s->mat[pj.mat];
            double nu_j = mj->av_l;
            double havg = AVG(s->h[i], pj.h);
            orc_gradw(dwdx, xi, pj.x, s->ndim, havg);
            double rab2 = 0.0, prod = 0.0;
            for (int r = 0; r < s->ndim; r++) {
                rab2 += SQR(xi[r] - pj.x[r]);
                prod += (xi[r] - pj.x[r]) * dwdx[r];
            }
            double sp = -pj.m * (pj.p / SQR(pj.rho) + p_i / SQR(rho_i));
            for (int n = 0; n < 3; n++) ai[n] = ai[n] + sp * dwdx[n];
            if (rig_i && rigid_p(s, pj.mat)) continue;
            double tv = pj.m * (nu_i * rho_i + nu_j * pj.rho) * prod / (rho_i * pj.rho) /
                        (rab2 + 0.01 * SQR(havg));
            for (int n = 0; n < 3; n++) ai[n] = ai[n] + tv * (vi[n] - pj.v[n]);
            if (mi->visc_type == 2) {
                /* Pi_ij = -alpha cbar mu_ij / rhobar for v_ij . r_ij < 0 */
                double vr = 0.0;
                for (int r = 0; r < s->ndim; r++) vr += (vi[r] - pj.v[r]) * (xi[r] - pj.x[r]);
                if (vr < 0.0) {
                    double alpha = AVG(mi->av_alpha, mj->av_alpha);
                    double mu = havg * vr / (rab2 + 0.01 * SQR(havg));
                    double pi_ij = -alpha * AVG(s->c[i], pj.c) * mu / AVG(rho_i, pj.rho);
                    double sv = -pj.m * pi_ij;
                    for (int n = 0; n < 3; n++) ai[n] = ai[n] + sv * dwdx[n];
                }
            }
        }
        if (s->st) {
            double f = -mi->surface_tension / rho_i * s->kappa[i];
            for (int n = 0; n < 3; n++) ai[n] = ai[n] + f * s->normal[3 * i + n];
        }
    }
    if (s->rigid) return rigid_acceleration(s);
    return 0;
}

/* MSolution::ClassicAcceleration, MSolution.cpp:506-548: a is cleared for
   computable particles only; images carry sigma = q = 0. */
static int phase_momentum_classic(orc_scene *s)
{
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 1; i <= s->np; i++) {
        double dwdx[3];
        if (!s->active[i]) continue;
        double *ai = &s->a[3 * i];
        ai[0] = ai[1] = ai[2] = 0.0;
        const double *xi = &s->x[3 * i], *sgi = &s->sig[9 * (size_t)i];
        const double rho_i = s->rho[i], q_i = s->q[i];
        for (int k = 0; k < s->nnbr[i]; k++) {
            orc_view pj = view_of(s, s->nbr[(size_t)i * s->maxnbr + k]);
            double havg = AVG(s->h[i], pj.h);
            orc_gradw(dwdx, xi, pj.x, s->ndim, havg);
            double ds[3] = {0.0, 0.0, 0.0};
            for (int n = 0; n < s->ndim; n++)
                for (int m = 0; m < s->ndim; m++)
                    ds[n] += ((pj.sig[3 * m + n] - (m == n ? pj.q : 0.0)) / SQR(pj.rho) +
                              (sgi[3 * m + n] - (m == n ? q_i : 0.0)) / SQR(rho_i)) * dwdx[m];
            for (int n = 0; n < 3; n++) ai[n] = ai[n] + pj.m * ds[n];
        }
    }
    return 0;
}

/* MSolution::ConstrainQuantity, MSolution.cpp:729-767; code 8 is the
   prescribed-velocity wall (extension): acceleration fully constrained. */
static void constrain(double *q, int code)
{
    switch (code) {
    case 1: q[0] = 0.0; break;
    case 2: q[1] = 0.0; break;
    case 3: q[2] = 0.0; break;
    case 4: q[0] = 0.0; q[1] = 0.0; break;
    case 5: q[1] = 0.0; q[2] = 0.0; break;
    case 6: q[2] = 0.0; q[0] = 0.0; break;
    case 7: case 8: q[0] = q[1] = q[2] = 0.0; break;
    }
}

/* MSolution::BoundCond, MSolution.cpp:700-726 */
static int phase_boundcond(orc_scene *s)
{
    if (s->baseaccel)
        for (int i = 1; i <= s->np; i++) {
            if (!s->active[i]) continue;
            if (!rigid_p(s, s->mat_id[i]))
                for (int n = 0; n < 3; n++) s->a[3 * i + n] += s->base_a[n];
        }
    for (int i = 1; i <= s->np; i++) {
        if (!s->active[i]) continue;
        constrain(&s->a[3 * i], s->dispbc[i]);
    }
    return 0;
}

/* MSolution::UpdateVelocity, MSolution.cpp:602-698 (XSPH off), and
   RigidBodyVelocity :1567-1591 */
static int phase_velocity(orc_scene *s)
{
    double dtn = AVG(s->dt, s->dtold);
    for (int i = 1; i <= s->np; i++) {
        if (!s->active[i] || rigid_p(s, s->mat_id[i])) continue;
        double *v = &s->v[3 * i];
        for (int n = 0; n < 3; n++) v[n] = v[n] + dtn * s->a[3 * i + n];
        double va = 0.0;
        for (int j = 0; j < s->ndim; j++) va += SQR(v[j]);
        s->vabs[i] = sqrt(va);
    }
    if (s->rigid) {
        for (int n = 0; n < 3; n++) {
            s->rbV[n] = s->rbV[n] + s->rbA[n] * dtn;
            s->rbW[n] = s->rbW[n] + s->rbAL[n] * dtn;
        }
        const double *V = s->rbV, *W = s->rbW;
        for (int k = 1; k <= s->np; k++) {
            if (!s->active[k] || !rigid_p(s, s->mat_id[k])) continue;
            double rl[3];
            for (int n = 0; n < 3; n++) rl[n] = s->x[3 * k + n] - s->rbR[n];
            s->v[3 * k + 0] = V[0] + W[1] * rl[2] - W[2] * rl[1];
            s->v[3 * k + 1] = V[1] + W[2] * rl[0] - W[0] * rl[2];
            s->v[3 * k + 2] = V[2] + W[0] * rl[1] - W[1] * rl[0];
        }
    }
    return 0;
}

/* MSolution::MoveParticles, MSolution.cpp:769-809; the step counter and stop
   test of Solution :317-319.  dispbc 8 (extension) keeps its position. */
static int phase_move(orc_scene *s)
{
    for (int m = 0; m < 3; m++) {
        s->cmm[0][m] = 1.0e20;
        s->cmm[1][m] = -1.0e20;
    }
    for (int i = 1; i <= s->np; i++)
        if (s->active[i] && s->dispbc[i] != 8)
            for (int n = 0; n < 3; n++) s->x[3 * i + n] += s->v[3 * i + n] * s->dt;
    for (int i = 1; i <= s->np; i++)
        for (int m = 0; m < s->ndim; m++) {
            double *xm = &s->x[3 * i + m];
            if (s->bcode[1][m] == 2 && *xm >= s->bx[1][m]) {
                double over = *xm - s->bx[1][m];
                *xm = s->bx[0][m] + over;
            }
            if (s->bcode[0][m] == 2 && *xm <= s->bx[0][m]) {
                double over = s->bx[0][m] - *xm;
                *xm = s->bx[0][m] - over;
            }
            s->cmm[0][m] = MINV(s->cmm[0][m], *xm);
            s->cmm[1][m] = MAXV(s->cmm[1][m], *xm);
        }
    s->timestep++;
    if (s->ptime - s->dt >= s->endtime) s->done = 1;
    return 0;
}

int orc_phase(orc_scene *s, int phase)
{
    switch (phase) {
    case ORC_PH_NEIGHBOURS: return phase_neighbours(s, 1);
    case ORC_PH_STRAIN: return phase_strain(s);
    case ORC_PH_NORMALS: return phase_normals(s);
    case ORC_PH_CONSTITUTIVE: return s->disctype == 0 ? phase_constitutive_classic(s) : phase_constitutive(s);
    case ORC_PH_TIMESTEP: return phase_timestep(s);
    case ORC_PH_MOMENTUM: return s->disctype == 0 ? phase_momentum_classic(s) : phase_momentum(s);
    case ORC_PH_BOUNDCOND: return phase_boundcond(s);
    case ORC_PH_VELOCITY: return phase_velocity(s);
    case ORC_PH_MOVE: return phase_move(s);
    }
    return fail(s, "unknown phase");
}

int orc_run(orc_scene *s, int nsteps, int *steps_done, int *done)
{
    int k = 0;
    for (; k < nsteps && !s->done; k++)
        for (int ph = 0; ph < ORC_NPHASES; ph++)
            if (orc_phase(s, ph)) {
                if (steps_done) *steps_done = k;
                if (done) *done = s->done;
                return 1;
            }
    if (steps_done) *steps_done = k;
    if (done) *done = s->done;
    return 0;
}

/* ---- initial state: MSimulationInit::Initialize, MSimulationInit.cpp:31-99 ---- */
int orc_deck_size(void) { return (int)sizeof(orc_deck); }
int orc_material_size(void) { return (int)sizeof(orc_material); }

orc_scene *orc_create(const orc_deck *d, char *err, int errlen)
{
    orc_scene *s = (orc_scene *)calloc(1, sizeof(orc_scene));
    const char *msg = NULL;
    if (!s) return NULL;
#define BAD(m) do { msg = (m); goto bad; } while (0)
    if (d->ndim < 1 || d->ndim > 3) BAD("ndim must be 1, 2 or 3");
    if (d->disctype != 0 && d->disctype != 2) BAD("disctype must be 0 (Classic) or 2 (Morris)");
    if (d->disctype == 0 && d->rigid_body) BAD("the rigid body needs disctype 2 (MSolution.cpp:492-500)");
    if (d->veloc_opt == 1) BAD("XSPH (veloc_opt 1) is out of scope");
    if (d->h_opt != 0) BAD("variable h (h_opt > 0) is out of scope");
    if (d->init_h_opt != 0) BAD("init_h_opt 1 is not supported");
    if (d->nummat < 1 || d->nummat > ORC_MAXMAT) BAD("bad material count");
    if (d->np < 1) BAD("no particles");
    s->ndim = d->ndim; s->disctype = d->disctype; s->tcrit_opt = d->tcrit_opt;
    s->maxnbr = d->maxnbr > 0 ? d->maxnbr : 40;
    s->st = d->surface_tension != 0; s->rigid = d->rigid_body != 0;
    s->boundary = d->boundary != 0; s->baseaccel = d->baseaccel != 0;
    s->tssf = d->tssf == 0.0 ? 0.8 : d->tssf;
    s->endtime = d->endtime; s->itss = d->itss;
    for (int side = 0; side < 2; side++)
        for (int m = 0; m < 3; m++) {
            s->bcode[side][m] = s->boundary ? d->boundary_code[side][m] : 0;
            s->bx[side][m] = d->boundary_x[side][m];
        }
    for (int m = 0; m < 3; m++) {
        /* BoundaryInput type rule, MSimulationData.cpp:1000-1004 */
        s->btype[m] = 1;
        if (s->bcode[0][m] > 0) s->btype[m] += 1;
        if (s->bcode[1][m] > 0) s->btype[m] += 2;
        if ((s->bcode[0][m] == 2) != (s->bcode[1][m] == 2)) BAD("periodic codes must come in pairs");
        if (s->bcode[0][m] == 1 || s->bcode[1][m] == 1) BAD("symmetry planes are not implemented");
        if ((s->bcode[0][m] != 0 && s->bcode[0][m] != 2) || (s->bcode[1][m] != 0 && s->bcode[1][m] != 2))
            BAD("ERROR: Boundary conditions code wrong.");
        s->base_a[m] = s->baseaccel ? d->base_a[m] : 0.0;
    }
    s->nummat = d->nummat;
    for (int k = 1; k <= d->nummat; k++) {
        s->mat[k] = d->mat[k - 1];
        if (s->mat[k].model != 9) BAD("only fluid material model 9 is on the path");
        if (s->mat[k].eos != 28 && s->mat[k].eos != 29) BAD("EOS must be 28 or 29");
    }
    s->rb_id = d->rb_id; s->rb_M = d->rb_total_mass; s->rb_I = d->rb_inertia;

    int np = s->np = d->np;
    size_t n1 = (size_t)np + 1;
    s->x = calloc(3 * n1, sizeof(double)); s->v = calloc(3 * n1, sizeof(double));
    s->a = calloc(3 * n1, sizeof(double)); s->normal = calloc(3 * n1, sizeof(double));
    s->nnormal = calloc(3 * n1, sizeof(double));
    double **scal[] = {&s->rho, &s->rhoold, &s->rho0, &s->p, &s->c, &s->m, &s->h, &s->vabs,
                       &s->critts_p, &s->trd, &s->mindist, &s->N, &s->kappa};
    for (size_t k = 0; k < sizeof scal / sizeof scal[0]; k++) *scal[k] = calloc(n1, sizeof(double));
    int **iscal[] = {&s->mat_id, &s->color, &s->dispbc, &s->active, &s->llp, &s->nnbr};
    for (size_t k = 0; k < sizeof iscal / sizeof iscal[0]; k++) *iscal[k] = calloc(n1, sizeof(int));
    s->nbr = calloc(n1 * (size_t)s->maxnbr, sizeof(int));
    if (s->disctype == 0) {
        s->rod = calloc(9 * n1, sizeof(double));
        s->sig = calloc(9 * n1, sizeof(double));
        s->q = calloc(n1, sizeof(double));
        if (!s->rod || !s->sig || !s->q) BAD("out of memory");
    }
    if (!s->x || !s->v || !s->a || !s->nnormal || !s->normal || !s->kappa || !s->nbr || !s->nnbr)
        BAD("out of memory");

    for (int i = 1; i <= np; i++) {
        for (int n = 0; n < 3; n++) {
            s->x[3 * i + n] = n < d->ndim ? d->x[3 * (i - 1) + n] : 0.0;
            s->v[3 * i + n] = (d->v && n < d->ndim) ? d->v[3 * (i - 1) + n] : 0.0;
        }
        s->mat_id[i] = d->mat_id[i - 1];
        if (s->mat_id[i] < 1 || s->mat_id[i] > s->nummat) BAD("Node has incorrect material type.");
        s->color[i] = d->color ? d->color[i - 1] : 1;
        s->dispbc[i] = d->dispbc ? d->dispbc[i - 1] : 0;
        s->active[i] = 1;
        s->llp[i] = ORC_LLEND;
    }
    /* InitBoundaryMultipliersAndAddtions, MSimulationInit.cpp:482-518 */
    for (int m = 0; m < 3; m++)
        if (s->bcode[0][m] == 2) {
            s->minadd[m][m] = s->bx[1][m] - s->bx[0][m];
            s->maxadd[m][m] = -s->bx[1][m] + s->bx[0][m];
        }
    /* InitMass / CalcMassFromTotal, MSimulationInit.cpp:107-177 */
    if (d->massopt == 1) {
        if (!d->node_mass) BAD("massopt 1 needs node masses");
        for (int i = 1; i <= np; i++) s->m[i] = d->node_mass[i - 1];
    } else {
        int cnt[ORC_MAXMAT + 1] = {0};
        for (int i = 1; i <= np; i++) cnt[s->mat_id[i]]++;
        for (int i = 1; i <= np; i++) s->m[i] = s->mat[s->mat_id[i]].mass / cnt[s->mat_id[i]];
    }
    /* InitVariables: vabs and bounding box, MSimulationInit.cpp:180-231 */
    for (int m = 0; m < 3; m++) { s->cmm[0][m] = 1.0e20; s->cmm[1][m] = -1.0e20; }
    for (int i = 1; i <= np; i++) {
        double va = 0.0;
        for (int j = 0; j < s->ndim; j++) va += SQR(s->v[3 * i + j]);
        s->vabs[i] = sqrt(va);
        for (int j = 0; j < s->ndim; j++) {
            s->cmm[0][j] = MINV(s->cmm[0][j], s->x[3 * i + j]);
            s->cmm[1][j] = MAXV(s->cmm[1][j], s->x[3 * i + j]);
        }
    }
    /* InitRho_P_C + EOScalc, MSimulationInit.cpp:267-311,384-440 */
    for (int i = 1; i <= np; i++) {
        orc_material *mt = &s->mat[s->mat_id[i]];
        s->rho0[i] = mt->rho;
        if (mt->eosinput[3] == 0.0) mt->eosinput[3] = 1.0;
        if (d->init_rhoe == 1 && d->node_rho) s->rho[i] = d->node_rho[i - 1];
        else s->rho[i] = s->rho0[i] / mt->eosinput[3];
        if (mt->eos == 28) {
            double B = mt->eosinput[0], gamma = mt->eosinput[1];
            s->p[i] = B * (pow(s->rho[i] / s->rho0[i], gamma) - 1.0);
            s->c[i] = sqrt(B * gamma / s->rho0[i]);
        } else {
            s->c[i] = mt->eosinput[0];
            s->p[i] = s->c[i] * (s->rho[i] - s->rho0[i]);
        }
        if (s->p[i] > mt->strinput0) s->p[i] = mt->strinput0;
    }
    /* InitH option 0, MSimulationInit.cpp:329-365 */
    s->hmax = 0.0;
    for (int i = 1; i <= np; i++) s->h[i] = s->mat[s->mat_id[i]].h;
    for (int k = 1; k <= s->nummat; k++)
        if (s->hmax < s->mat[k].h) s->hmax = s->mat[k].h;
    s->timestep = 1;
    /* first neighbours (no ghosts), critical step, first dt */
    if (phase_neighbours(s, 0)) { msg = s->err; goto bad; }
    if (calc_critical_timestep(s)) { msg = s->err; goto bad; }
    if (fabs(s->itss) > 1.e-10) s->dt = MINV(s->itss, s->critts * s->tssf);
    else s->dt = s->critts * s->tssf;
    for (int i = 1; i <= np; i++) s->rhoold[i] = s->rho[i];
    s->ptime = 0.0;
    return s;
bad:
    if (err && errlen > 0) snprintf(err, (size_t)errlen, "%s", msg ? msg : "error");
    orc_destroy(s);
    return NULL;
#undef BAD
}

void orc_destroy(orc_scene *s)
{
    if (!s) return;
    void *ptrs[] = {s->x, s->v, s->a, s->normal, s->nnormal, s->rho, s->rhoold, s->rho0, s->p,
                    s->c, s->m, s->h, s->vabs, s->critts_p, s->trd, s->mindist, s->N, s->kappa,
                    s->mat_id, s->color, s->dispbc, s->active, s->llp, s->nnbr, s->nbr, s->g,
                    s->grid, s->rod, s->sig, s->q};
    for (size_t k = 0; k < sizeof ptrs / sizeof ptrs[0]; k++) free(ptrs[k]);
    free(s);
}

const char *orc_last_error(orc_scene *s) { return s ? s->err : "null scene"; }
int orc_np(orc_scene *s) { return s->np; }
#ifdef _OPENMP
#include <omp.h>
int orc_threads(void) { return omp_get_max_threads(); }
#else
int orc_threads(void) { return 1; }
#endif
int orc_ngp(orc_scene *s) { return s->ngp; }

int orc_globals(orc_scene *s, double *out, int n)
{
    double g[32];
    int k = 0;
    g[k++] = s->dt; g[k++] = s->dtold; g[k++] = s->ptime; g[k++] = s->critts;
    g[k++] = s->timestep; g[k++] = s->hmax;
    for (int m = 0; m < 3; m++) g[k++] = s->cmm[0][m];
    for (int m = 0; m < 3; m++) g[k++] = s->cmm[1][m];
    g[k++] = s->ngp;
    for (int m = 0; m < 3; m++) g[k++] = s->rbR[m];
    for (int m = 0; m < 3; m++) g[k++] = s->rbV[m];
    for (int m = 0; m < 3; m++) g[k++] = s->rbW[m];
    for (int m = 0; m < 3; m++) g[k++] = s->rbA[m];
    for (int m = 0; m < 3; m++) g[k++] = s->rbAL[m];
    g[k++] = s->done;
    for (int i = 0; i < n && i < k; i++) out[i] = g[i];
    return k;
}

/* Overwrites the velocities (field ORC_F_V, np*3) between phases. */
int orc_set_field(orc_scene *s, int f, const double *in)
{
    if (f != ORC_F_V) return 1;
    memcpy(s->v + 3, in, (size_t)s->np * 3 * sizeof(double));
    for (int i = 1; i <= s->np; i++) {
        double va = 0.0;
        for (int j = 0; j < s->ndim; j++) va += SQR(s->v[3 * i + j]);
        s->vabs[i] = sqrt(va);
    }
    return 0;
}

int orc_get_field(orc_scene *s, int f, double *out)
{
    const double *src = NULL;
    int vec = 0;
    switch (f) {
    case ORC_F_X: src = s->x; vec = 1; break;
    case ORC_F_V: src = s->v; vec = 1; break;
    case ORC_F_A: src = s->a; vec = 1; break;
    case ORC_F_NORMAL: src = s->normal; vec = 1; break;
    case ORC_F_NNORMAL: src = s->nnormal; vec = 1; break;
    case ORC_F_RHO: src = s->rho; break;
    case ORC_F_P: src = s->p; break;
    case ORC_F_C: src = s->c; break;
    case ORC_F_MASS: src = s->m; break;
    case ORC_F_H: src = s->h; break;
    case ORC_F_TRD: src = s->trd; break;
    case ORC_F_MINDIST: src = s->mindist; break;
    case ORC_F_N: src = s->N; break;
    case ORC_F_KAPPA: src = s->kappa; break;
    case ORC_F_CRITTS: src = s->critts_p; break;
    case ORC_F_RHOOLD: src = s->rhoold; break;
    case ORC_F_VABS: src = s->vabs; break;
    case ORC_F_ACTIVE:
        for (int i = 1; i <= s->np; i++) out[i - 1] = s->active[i];
        return 0;
    case ORC_F_Q:
        for (int i = 1; i <= s->np; i++) out[i - 1] = s->q ? s->q[i] : 0.0;
        return 0;
    case ORC_F_SIGMA_DIAG: case ORC_F_SIGMA_OFF: case ORC_F_ROD_DIAG: case ORC_F_ROD_OFF: {
        /* diagonal (xx, yy, zz) or off-diagonal (xy, xz, yz); Morris: sigma = -p I */
        static const int dg[3] = {0, 4, 8}, od[3] = {1, 2, 5};
        const int *ix = (f == ORC_F_SIGMA_DIAG || f == ORC_F_ROD_DIAG) ? dg : od;
        const double *t = (f == ORC_F_SIGMA_DIAG || f == ORC_F_SIGMA_OFF) ? s->sig : s->rod;
        for (int i = 1; i <= s->np; i++)
            for (int k = 0; k < 3; k++) {
                double v;
                if (t) v = t[9 * (size_t)i + ix[k]];
                else if (f == ORC_F_SIGMA_DIAG) v = -s->p[i];
                else v = 0.0;
                out[3 * (size_t)(i - 1) + k] = v;
            }
        return 0;
    }
    default: return 1;
    }
    if (vec) memcpy(out, src + 3, (size_t)s->np * 3 * sizeof(double));
    else memcpy(out, src + 1, (size_t)s->np * sizeof(double));
    return 0;
}

int orc_get_neighbours(orc_scene *s, int32_t *nnbr, int64_t *canon, int maxnbr)
{
    for (int i = 1; i <= s->np; i++) {
        nnbr[i - 1] = s->nnbr[i];
        for (int k = 0; k < s->nnbr[i] && k < maxnbr; k++) {
            int id = s->nbr[(size_t)i * s->maxnbr + k];
            int64_t cid;
            if (id > 0) cid = id - 1;
            else {
                const orc_ghost *g = &s->g[-id];
                cid = -(1 + (int64_t)(g->par - 1) * 27 + (g->shift[0] + 1) * 9 +
                        (g->shift[1] + 1) * 3 + (g->shift[2] + 1));
            }
            canon[(size_t)(i - 1) * maxnbr + k] = cid;
        }
    }
    return 0;
}

/* The linked-list state of the last SetupLinkedList + GhostSetup
   (MNeighbourVars.cpp:57-80, MSolution.cpp:135-144, MGhostParticleData.cpp:16,30-31):
   grid geometry, sph_llgrid, llpointer of the real particles (by 0-based
   deck number) and of the images, and each image's source particle (1-based).
   Any output may be NULL.  Returns 1 if an array is too small. */
int orc_get_linked_list(orc_scene *s, double *gmin, double *gsize, int32_t *glim, int32_t *llgrid,
                        long long ncell, int32_t *real_next, int32_t *ghost_next, int32_t *ghost_par, int ngp)
{
    const long long nc = (long long)s->glim[0] * s->glim[1] * s->glim[2];
    for (int m = 0; m < 3; m++) {
        if (gmin) gmin[m] = s->gmin[m];
        if (gsize) gsize[m] = s->gsize[m];
        if (glim) glim[m] = s->glim[m];
    }
    if ((llgrid && ncell < nc) || ((ghost_next || ghost_par) && ngp < s->ngp)) return 1;
    if (llgrid)
        for (long long c = 0; c < nc; c++) llgrid[c] = s->grid[c];
    if (real_next)
        for (int i = 1; i <= s->np; i++) real_next[i - 1] = s->llp[i];
    for (int g = 1; g <= s->ngp; g++) {
        if (ghost_next) ghost_next[g - 1] = s->g[g].llp;
        if (ghost_par) ghost_par[g - 1] = s->g[g].par;
    }
    return 0;
}
