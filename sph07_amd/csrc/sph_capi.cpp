// sph_capi.cpp -- extern "C" boundary of the MI355X step loop and the host
// side of MSimulationInit::Initialize (MSimulationInit.cpp:31-99).
//
// A context owns one scene's device state (structure of arrays in HBM), its
// HIP stream and a pinned mirror of the device control block.  The phase
// entry points synchronise and return the reference's status codes; the
// fused sph_solution queues whole steps and synchronises once per chunk.
#include "../../include/sph07_amd.h"
#include "sph_kernels.hpp"

#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

using namespace sph;

namespace {

constexpr double kPI = 3.141592654;   // global.h:12

// B-spline normalisations for the averaged smoothing length havg, in the
// reference's operation order (MKernelBSpline.cpp:145-156); eps2 = 0.01 havg^2
// as in the Morris term (MSolution.cpp:592).  cut2 is left to the caller.
PairConst kernel_const(int ndim, double havg)
{
    PairConst pc{};
    pc.havg = havg;
    if (ndim == 3) pc.const1 = 1.0 / (kPI * (havg * havg * havg));
    else if (ndim == 2) pc.const1 = 10.0 / (7.0 * kPI * (havg * havg));
    else pc.const1 = 2.0 / (3.0 * havg);
    pc.const2 = 0.25 * pc.const1;
    pc.eps2 = 0.01 * (havg * havg);
    return pc;
}

// Kernel groups timed with HIP events on the context's stream.
enum KId { K_BIN, K_IMAGES, K_NBR, K_STRAIN, K_NORMALS, K_ADVANCE, K_MOMENTUM, K_INTEGRATE,
           K_FINISH, K_COUNT };
const char *const kKernelNames[K_COUNT] = {
    "bin (grid, count, scan, scatter, rank)", "periodic images", "k_neighbours", "k_strain",
    "k_normals+k_curvature", "k_advance", "k_momentum (+k_rigid_reduce)", "k_integrate", "k_finish"};

struct Timing {
    bool on = false;
    std::vector<hipEvent_t> pool;
    struct Pending { int id; hipEvent_t a, b; };
    std::vector<Pending> pending;
    double total_ms[K_COUNT] = {};
    long long count[K_COUNT] = {};
};

struct Context {
    hipStream_t st = nullptr;
    Params P{};
    // two buffer sets of the persistent particle fields: the per-step sort
    // moves them from AA[cur] to AA[cur ^ 1]; everything else is shared
    Arrays AA[2]{};
    int cur = 0;
    Ctl *hctl = nullptr;          // pinned mirror of the control block
    std::string err;
    int N = 0, ndim = 0, maxnbr = 0, arm = 0;
    std::vector<void *> part_allocs, cell_allocs;
    long long *canon = nullptr;
    Timing tm;
    // captured step sequences (hipGraphs), one per starting buffer set: the
    // kernels' arguments name the buffer sets, and a step moves the state to
    // the other set, so a graph is only valid from the set it was captured
    // from.  Both are dropped after any buffer or launch-shape change.
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    int gsteps = 0;
    bool use_graph = true;
    // launch shapes (timing only, never results) and whether they were tuned
    Tune tune;
    bool tuned = false;
};

// scenes at least this large are tuned on their first sph_solution call
constexpr int kAutotuneMinN = 65536;

inline Arrays &cur(Context *c) { return c->AA[c->cur]; }

void drop_graph(Context *c)
{
    for (hipGraphExec_t &g : c->gexec) {
        if (g) hipGraphExecDestroy(g);
        g = nullptr;
    }
    c->gsteps = 0;
}

hipEvent_t take_event(Context *c)
{
    if (!c->tm.pool.empty()) {
        hipEvent_t e = c->tm.pool.back();
        c->tm.pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    hipEventCreate(&e);
    return e;
}

// Called after a stream synchronisation: folds finished event pairs into the
// per-group totals.
void harvest_timing(Context *c)
{
    for (auto &p : c->tm.pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            c->tm.total_ms[p.id] += ms;
            c->tm.count[p.id]++;
        }
        c->tm.pool.push_back(p.a);
        c->tm.pool.push_back(p.b);
    }
    c->tm.pending.clear();
}

template <class F>
inline void timed(Context *c, int id, F &&f)
{
    if (!c->tm.on) {
        f();
        return;
    }
    hipEvent_t a = take_event(c), b = take_event(c);
    hipEventRecord(a, c->st);
    f();
    hipEventRecord(b, c->st);
    c->tm.pending.push_back({id, a, b});
}

#define HIPCHK(ctx, expr)                                                          \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess) {                                                    \
            (ctx)->err = std::string("HIP error: ") + hipGetErrorString(e_) +      \
                         " at " #expr;                                             \
            return SPH_ERR_HIP;                                                    \
        }                                                                          \
    } while (0)

template <class T>
int dalloc(Context *c, std::vector<void *> &list, T **p, size_t n)
{
    *p = nullptr;
    if (n == 0) n = 1;
    HIPCHK(c, hipMalloc((void **)p, n * sizeof(T)));
    list.push_back(*p);
    return 0;
}

void free_list(std::vector<void *> &list)
{
    for (void *p : list) hipFree(p);
    list.clear();
}

// per-particle arrays of capacity Ncap = N + gcap; persistent fields twice
int alloc_particles(Context *c, int gcap)
{
    Params &P = c->P;
    Arrays &A = c->AA[0], &B = c->AA[1];
    const size_t ncap = (size_t)c->N + gcap;
    if (ncap > (1ull << 30)) {                  // neighbour-list chunk stride is 32-bit
        c->err = "too many particles and images for one context";
        return 1;
    }
    if (c->arm == SPH_ARM_FP32C && ncap * 32 > 0xffffffffull) {   // 32-bit record offsets
        c->err = "the FP32c arm takes at most 2^27 particles and images per context";
        return 1;
    }
    std::vector<void *> &L = c->part_allocs;
    P.gcap = gcap;
    P.Ncap = (int)ncap;
    P.nstride = 1;
    P.nshift = 5;                              // log2(kNbrChunk * sizeof(int))
    while ((size_t)P.nstride < ncap) {
        P.nstride <<= 1;
        P.nshift++;
    }
    int rc = 0;
    for (int d = 0; d < 3; d++) {
        A.x[d] = A.xs[d] = A.v[d] = A.a[d] = A.nrm[d] = A.nn[d] = nullptr;
        B.x[d] = B.v[d] = B.a[d] = nullptr;
        if (d >= c->ndim) continue;
        for (Arrays *S : {&A, &B}) {
            rc |= dalloc(c, L, &S->x[d], ncap);
            rc |= dalloc(c, L, &S->v[d], ncap);
            rc |= dalloc(c, L, &S->a[d], c->N);
        }
        if (d == 0)
            for (Arrays *S : {&A, &B}) rc |= dalloc(c, L, &S->aa, c->N);
        rc |= dalloc(c, L, &A.xs[d], ncap);
        if (P.st) {
            rc |= dalloc(c, L, &A.nrm[d], c->N);
            rc |= dalloc(c, L, &A.nn[d], c->N);
        }
    }
    for (Arrays *S : {&A, &B}) {
        rc |= dalloc(c, L, &S->rho, ncap);
        rc |= dalloc(c, L, &S->m, ncap);
        rc |= dalloc(c, L, &S->info, ncap);
        rc |= dalloc(c, L, &S->orig, c->N);
        if (P.rigid) rc |= dalloc(c, L, &S->rrank, c->N);
        else S->rrank = nullptr;
    }
    if (P.disctype == 0) {
        rc |= dalloc(c, L, &A.q, c->N);
        rc |= dalloc(c, L, &B.q, c->N);
        for (int k = 0; k < 6; k++) {
            rc |= dalloc(c, L, &A.rod[k], c->N);
            rc |= dalloc(c, L, &A.sig[k], c->N);
        }
        rc |= dalloc(c, L, &A.rcC, ncap);
    } else {
        A.q = B.q = nullptr;
        for (int k = 0; k < 6; k++) A.rod[k] = A.sig[k] = nullptr;
        A.rcC = nullptr;
    }
    rc |= dalloc(c, L, &A.rsA, ncap);
    rc |= dalloc(c, L, &A.rsB, ncap);
    rc |= dalloc(c, L, &A.rmA, ncap);
    rc |= dalloc(c, L, &A.rmB, ncap);
    rc |= dalloc(c, L, &A.rmC, ncap);
    if (c->arm == SPH_ARM_FP32C) {
        rc |= dalloc(c, L, &A.fs, 2 * ncap);
        rc |= dalloc(c, L, &A.fm, 2 * ncap);
        rc |= dalloc(c, L, &A.fmC, ncap);
    } else {
        A.fs = A.fm = nullptr;
        A.fmC = nullptr;
    }
    rc |= dalloc(c, L, &A.rho_new, ncap);
    rc |= dalloc(c, L, &A.p, ncap);
    rc |= dalloc(c, L, &A.prho2, ncap);
    rc |= dalloc(c, L, &A.trd, c->N);
    rc |= dalloc(c, L, &A.mindist, c->N);
    rc |= dalloc(c, L, &A.critts_p, c->N);
    A.rhoold = nullptr;
    if (P.st) {
        rc |= dalloc(c, L, &A.Nf, c->N);
        rc |= dalloc(c, L, &A.kappa, c->N);
    } else {
        A.Nf = A.kappa = nullptr;
    }
    rc |= dalloc(c, L, &A.nnbr, c->N);
    rc |= dalloc(c, L, &A.nbr, (size_t)nbr_rows(c->maxnbr) * (size_t)P.nstride);
    rc |= dalloc(c, L, &A.cell_of, c->N);
    rc |= dalloc(c, L, &A.occ, (size_t)c->N + 1);
    rc |= dalloc(c, L, &A.ovf, (size_t)c->N + 1);
    rc |= dalloc(c, L, &A.slot, c->N);
    rc |= dalloc(c, L, &A.tmp, c->N);
    rc |= dalloc(c, L, &A.gsrc, gcap);
    rc |= dalloc(c, L, &A.gshift, gcap);
    // shared fields: the second set points at the same arrays
    for (int d = 0; d < 3; d++) {
        B.xs[d] = A.xs[d];
        B.nrm[d] = A.nrm[d];
        B.nn[d] = A.nn[d];
    }
    B.rsA = A.rsA; B.rsB = A.rsB; B.rmA = A.rmA; B.rmB = A.rmB; B.rmC = A.rmC;
    B.fs = A.fs; B.fm = A.fm; B.fmC = A.fmC;
    B.rho_new = A.rho_new; B.p = A.p; B.prho2 = A.prho2;
    B.trd = A.trd; B.mindist = A.mindist; B.critts_p = A.critts_p; B.rhoold = A.rhoold;
    B.Nf = A.Nf; B.kappa = A.kappa; B.nnbr = A.nnbr; B.nbr = A.nbr;
    B.cell_of = A.cell_of; B.slot = A.slot; B.tmp = A.tmp; B.occ = A.occ; B.ovf = A.ovf;
    B.gsrc = A.gsrc; B.gshift = A.gshift;
    for (int k = 0; k < 6; k++) {
        B.rod[k] = A.rod[k];
        B.sig[k] = A.sig[k];
    }
    B.rcC = A.rcC;
    return rc;
}

// cell arrays, control block and rigid list are shared by both sets
void share_cells(Context *c)
{
    Arrays &A = c->AA[0], &B = c->AA[1];
    B.cell_count = A.cell_count; B.cell_start = A.cell_start;
    B.gcount = A.gcount; B.gfirst = A.gfirst;
    B.slab_cnt = A.slab_cnt; B.slab_off = A.slab_off; B.scan_aux = A.scan_aux;
    B.rigid_idx = A.rigid_idx;
    B.ctl = A.ctl;
    B.bbox = A.bbox;
}

int alloc_cells(Context *c, int cap)
{
    Arrays &A = c->AA[0];
    std::vector<void *> &L = c->cell_allocs;
    free_list(L);
    c->P.ncell_cap = cap;
    int rc = 0;
    rc |= dalloc(c, L, &A.cell_count, (size_t)cap + 1);
    rc |= dalloc(c, L, &A.cell_start, (size_t)cap + 1);
    rc |= dalloc(c, L, &A.gcount, (size_t)cap);
    rc |= dalloc(c, L, &A.gfirst, (size_t)cap);
    rc |= dalloc(c, L, &A.slab_cnt, (size_t)cap);
    rc |= dalloc(c, L, &A.slab_off, (size_t)cap);
    rc |= dalloc(c, L, &A.scan_aux, (size_t)cap / 2048 + 4);
    // the binning leaves the counts zero after each scan; they start zero
    if (!rc) HIPCHK(c, hipMemset(A.cell_count, 0, ((size_t)cap + 1) * sizeof(int)));
    share_cells(c);
    return rc;
}

// Grows the image capacity: per-particle arrays re-allocated; the current
// set's persistent fields move to set 0 of the new allocation.
int grow_particles(Context *c, int new_gcap)
{
    Arrays old = cur(c);
    std::vector<void *> old_list;
    old_list.swap(c->part_allocs);
    if (alloc_particles(c, new_gcap)) return SPH_ERR_HIP;
    share_cells(c);
    c->cur = 0;
    Arrays &A = c->AA[0];
    const size_t n = (size_t)c->N;
    auto cp = [&](void *dst, const void *src, size_t bytes) -> int {
        if (!dst || !src) return 0;
        HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->st));
        return 0;
    };
    int rc = 0;
    for (int d = 0; d < c->ndim; d++) {
        rc |= cp(A.x[d], old.x[d], n * 8);
        rc |= cp(A.xs[d], old.xs[d], n * 8);
        rc |= cp(A.v[d], old.v[d], n * 8);
        rc |= cp(A.a[d], old.a[d], n * 8);
        if (d == 0) rc |= cp(A.aa, old.aa, n * 8);
        if (c->P.st) {
            rc |= cp(A.nrm[d], old.nrm[d], n * 8);
            rc |= cp(A.nn[d], old.nn[d], n * 8);
        }
    }
    rc |= cp(A.rho, old.rho, n * 8);
    rc |= cp(A.rho_new, old.rho_new, n * 8);
    rc |= cp(A.p, old.p, n * 8);
    rc |= cp(A.prho2, old.prho2, n * 8);
    rc |= cp(A.m, old.m, n * 8);
    rc |= cp(A.info, old.info, n * 4);
    rc |= cp(A.orig, old.orig, n * 4);
    if (c->P.rigid) rc |= cp(A.rrank, old.rrank, n * 4);
    if (c->P.disctype == 0) {
        rc |= cp(A.q, old.q, n * 8);
        for (int k = 0; k < 6; k++) {
            rc |= cp(A.rod[k], old.rod[k], n * 8);
            rc |= cp(A.sig[k], old.sig[k], n * 8);
        }
    }
    rc |= cp(A.trd, old.trd, n * 8);
    rc |= cp(A.mindist, old.mindist, n * 8);
    rc |= cp(A.critts_p, old.critts_p, n * 8);
    rc |= cp(A.nnbr, old.nnbr, n * 4);
    if (c->P.st) {
        rc |= cp(A.Nf, old.Nf, n * 8);
        rc |= cp(A.kappa, old.kappa, n * 8);
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    free_list(old_list);
    return rc;
}

int sync_ctl(Context *c)
{
    HIPCHK(c, hipMemcpyAsync(c->hctl, cur(c).ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    harvest_timing(c);
    return 0;
}

int push_ctl(Context *c)
{
    HIPCHK(c, hipMemcpyAsync(cur(c).ctl, c->hctl, sizeof(Ctl), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}

const char *err_text(int code)
{
    switch (code) {
    case kErrNbrOverflow: return "ERROR: Neighbour list exceeded maximum.";
    case kErrOutOfGrid: return "ERROR: particle outside the neighbour grid.";
    case 1: return "ERROR: a cell received periodic images from two passes (unsupported).";
    default: return "ERROR: device error.";
    }
}

// Checks the device error flag after a synchronisation.  Capacity errors are
// repaired (capacity grown, flag cleared) and reported as 'retry' (-1).
int check_device(Context *c)
{
    if (sync_ctl(c)) return SPH_ERR_HIP;
    Ctl &h = *c->hctl;
    if (h.err == 0) return 0;
    // a binning stopped by the error may have left counts that no scan cleared
    HIPCHK(c, hipMemsetAsync(cur(c).cell_count, 0, ((size_t)c->P.ncell_cap + 1) * sizeof(int), c->st));
    if (h.err == kErrGridCap) {
        long long want = (long long)h.ncell * 3 / 2 + 1024;   // h.ncell: cells or codes needed
        if (want > 0x7fffffffLL - 1) {
            c->err = "ERROR: neighbour grid too large.";
            return SPH_ERROR;
        }
        drop_graph(c);
        if (alloc_cells(c, (int)want)) return SPH_ERR_HIP;
        h.err = 0;
        if (push_ctl(c)) return SPH_ERR_HIP;
        return -1;
    }
    if (h.err == kErrGhostCap) {
        int want = std::max(2 * c->P.gcap, h.err_particle + 1024);
        drop_graph(c);
        if (grow_particles(c, want)) return SPH_ERR_HIP;
        h.err = 0;
        if (push_ctl(c)) return SPH_ERR_HIP;
        return -1;
    }
    c->err = err_text(h.err);
    if (h.err_particle >= 0) c->err += " (particle " + std::to_string(h.err_particle + 1) + ")";
    return h.err == kErrNbrOverflow ? SPH_ERR_NBR_OVERFLOW
           : h.err == kErrOutOfGrid ? SPH_ERR_OUT_OF_GRID
                                    : SPH_ERROR;
}

// Runs a phase sequence, repairing capacity errors by re-running it.
template <class F>
int run_phase(Context *c, F &&f)
{
    for (int attempt = 0; attempt < 8; attempt++) {
        f();
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            c->err = std::string("HIP launch error: ") + hipGetErrorString(e);
            return SPH_ERR_HIP;
        }
        int rc = check_device(c);
        if (rc != -1) return rc;
    }
    c->err = "ERROR: capacity repair did not converge.";
    return SPH_ERROR;
}

// MSolution::SetupLinkedList: binning, then the move of the persistent
// fields into cell order in the other buffer set, which becomes current.
void sort_step(Context *c, bool grid = true, bool move_a = true)
{
    launch_setup_linked_list(c->P, cur(c), c->st, grid);
    launch_permute(c->P, cur(c), c->AA[c->cur ^ 1], c->st, move_a);
    c->cur ^= 1;
}

// A chain of fused steps (first .. last) sets up each binning's grid in the
// previous step's k_finish, one launch fewer per step; the chain's first step
// sets up its own, and its last leaves the grid of its own binning in place
// for the downloads that follow.
void fused_step(Context *c, bool first = true, bool last = true)
{
    const Params &P = c->P;
    // the Morris momentum sweeps rewrite every a; the Classic one only the
    // computable particles' (the others keep theirs, so a moves)
    timed(c, K_BIN, [&] { sort_step(c, first, P.disctype == 0); });
    const Arrays &A = cur(c);
    hipStream_t st = c->st;
    if (P.periodic) timed(c, K_IMAGES, [&] { launch_ghost_setup(P, A, st); });
    timed(c, K_NBR, [&] { launch_neighbours(P, A, st, c->tune.cell_slots); });
    if (c->arm == SPH_ARM_FP32C)
        timed(c, K_STRAIN, [&] { launch_strain32(P, A, st, c->tune.strain_batch); });
    else
        timed(c, K_STRAIN, [&] { launch_strain(P, A, M_STRAIN | M_CONST | M_CRIT, st, c->tune.strain_batch); });
    if (P.st) timed(c, K_NORMALS, [&] { launch_normals(P, A, st); });
    timed(c, K_ADVANCE, [&] { launch_advance(P, A, st); });
    // the FP32c sweep integrates its particles itself where no rigid body
    // needs the whole sweep's forces first (k_integrate is not launched)
    const bool fuse_int = c->arm == SPH_ARM_FP32C && !P.rigid;
    timed(c, K_MOMENTUM, [&] {
        if (c->arm == SPH_ARM_FP32C) {
            launch_momentum32(P, A, st, c->tune.momentum_batch, fuse_int ? 1 : 0);
            launch_rigid_reduce(P, A, 1, st);
        } else {
            launch_momentum(P, A, 1, st, c->tune.momentum_batch);
        }
    });
    if (!fuse_int) timed(c, K_INTEGRATE, [&] { launch_integrate(P, A, M_BC | M_VEL | M_MOVE | M_RHO, st); });
    timed(c, K_FINISH, [&] { launch_finish(P, A, 1, 1, st, last ? 0 : 1); });
}

// ---- host part of MSimulationInit::Initialize -------------------------------
int build_context(Context *c, const sph_scene_desc *d)
{
    Params &P = c->P;
    if (d->ndim < 1 || d->ndim > 3) { c->err = "ndim must be 1, 2 or 3"; return SPH_ERROR; }
    if (d->disctype != 0 && d->disctype != 2) { c->err = "disctype must be 0 (Classic) or 2 (Morris)"; return SPH_ERR_UNSUPPORTED; }
    if (d->disctype == 0 && c->arm != SPH_ARM_FP64) { c->err = "the Classic discretisation (disctype 0) runs in the FP64 arm"; return SPH_ERR_UNSUPPORTED; }
    if (d->disctype == 0 && d->rigid_body) { c->err = "the rigid body needs disctype 2 (MSolution.cpp:492-500)"; return SPH_ERR_UNSUPPORTED; }
    if (d->veloc_opt == 1) { c->err = "XSPH (veloc_opt 1) is outside the GPU path"; return SPH_ERR_UNSUPPORTED; }
    if (d->h_opt != 0) { c->err = "variable h (h_opt > 0) is outside the GPU path"; return SPH_ERR_UNSUPPORTED; }
    if (d->init_h_opt != 0) { c->err = "init_h_opt 1 is not supported"; return SPH_ERR_UNSUPPORTED; }
    if (d->nummat < 1 || d->nummat > kMaxMat) { c->err = "material count out of range"; return SPH_ERROR; }
    if (d->np < 1) { c->err = "no particles"; return SPH_ERROR; }
    if (d->tcrit_opt < 1 || d->tcrit_opt > 3) { c->err = "ERROR: Wrong critical time step option."; return SPH_ERROR; }
    const int N = d->np, D = d->ndim;
    c->N = N;
    c->ndim = D;
    c->maxnbr = d->maxnbr > 0 ? d->maxnbr : 40;
    P.N = N;
    P.ndim = D;
    P.tcrit_opt = d->tcrit_opt;
    P.maxnbr = c->maxnbr;
    P.arm = c->arm;
    P.disctype = d->disctype;
    P.st = d->surface_tension != 0;
    P.rigid = d->rigid_body != 0;
    P.baseaccel = d->baseaccel != 0;
    P.nmat = d->nummat;
    P.rb_id = d->rb_id;
    P.rb_M = d->rb_total_mass;
    P.rb_I = d->rb_inertia;
    P.tssf = d->tssf == 0.0 ? 0.8 : d->tssf;
    P.endtime = d->endtime;
    P.want_mindist = d->tcrit_opt == 2;
    P.periodic = 0;
    for (int side = 0; side < 2; side++)
        for (int m = 0; m < 3; m++) {
            P.bcode[side][m] = d->boundary ? d->boundary_code[side][m] : 0;
            P.bx[side][m] = d->boundary_x[side][m];
        }
    for (int m = 0; m < 3; m++) {
        const int c0 = P.bcode[0][m], c1 = P.bcode[1][m];
        if ((c0 != 0 && c0 != 2) || (c1 != 0 && c1 != 2)) {
            c->err = c0 == 1 || c1 == 1 ? "ERROR: Symmetry boundary conditions not implemented yet."
                                        : "ERROR: Boundary conditions code wrong.";
            return SPH_ERROR;
        }
        if ((c0 == 2) != (c1 == 2)) { c->err = "Error in definition of periodic boundary conditions."; return SPH_ERROR; }
        if (c0 == 2 && m >= D) { c->err = "periodic axis beyond ndim"; return SPH_ERROR; }
        if (c0 == 2 && !(P.bx[1][m] > P.bx[0][m])) { c->err = "periodic planes must satisfy xmin < xmax"; return SPH_ERROR; }
        P.btype[m] = 1 + (c0 > 0 ? 1 : 0) + (c1 > 0 ? 2 : 0);
        if (c0 == 2) {
            P.periodic = 1;
            P.minadd[m][m] = P.bx[1][m] - P.bx[0][m];
            P.maxadd[m][m] = -P.bx[1][m] + P.bx[0][m];
        }
        P.base_a[m] = P.baseaccel ? d->base_a[m] : 0.0;
        if (m >= D && P.base_a[m] != 0.0) { c->err = "base acceleration beyond ndim"; return SPH_ERR_UNSUPPORTED; }
    }
    // materials (MMaterial cards; EOScalc sound speed, MSimulationInit.cpp:384-440)
    std::vector<double> mat_eos3(kMaxMat + 1, 1.0);
    P.hmax = 0.0;
    for (int k = 1; k <= d->nummat; k++) {
        const sph_material_desc &md = d->mat[k - 1];
        if (md.model != 9) { c->err = "only fluid material model 9 is on the GPU path"; return SPH_ERR_UNSUPPORTED; }
        if (md.eos != 28 && md.eos != 29) { c->err = "EOS must be 28 or 29"; return SPH_ERR_UNSUPPORTED; }
        Mat &mt = P.mat[k];
        mt.rho0 = md.rho;
        mt.nu = md.av_l;
        mt.sigma = md.surface_tension;
        mt.h = md.h;
        mt.eos = md.eos;
        mt.visc_type = md.visc_type;
        mt.alpha = md.av_alpha;
        mt.av_q = md.av_q;
        mt.eos_b = md.eosinput[0];
        mt.eos_gamma = md.eosinput[1];
        mt.c = md.eos == 29 ? md.eosinput[0] : sqrt(md.eosinput[0] * md.eosinput[1] / md.rho);
        mat_eos3[k] = md.eosinput[3] == 0.0 ? 1.0 : md.eosinput[3];
        if (P.hmax < md.h) P.hmax = md.h;   // InitH option 0 (:336-341)
    }
    if (P.hmax <= 0.0) { c->err = "smoothing length must be positive"; return SPH_ERROR; }
    // material-pair kernel constants (MKernelBSpline.cpp:145-156)
    for (int a = 1; a <= d->nummat; a++)
        for (int b = 1; b <= d->nummat; b++) {
            PairConst &pc = P.pc[a][b];
            pc = kernel_const(D, 0.5 * (P.mat[a].h + P.mat[b].h));
            const double hav = 2.0 * (0.5 * (P.mat[a].h + P.mat[b].h));
            pc.cut2 = hav * hav;
        }
    P.cut2max = 0.0;
    P.cut2min = 1e300;
    for (int a = 1; a <= d->nummat; a++)
        for (int b = 1; b <= d->nummat; b++) {
            P.cut2max = std::max(P.cut2max, P.pc[a][b].cut2);
            P.cut2min = std::min(P.cut2min, P.pc[a][b].cut2);
        }
    {
        // quantum of the FP32c record positions: cut = f 2^e, f in [0.5, 1),
        // q = 2^(e - 23), so cut / q = f 2^23
        int e = 0;
        frexp(sqrt(P.cut2max), &e);
        P.qexp = e - 23;
        P.qinv = ldexp(1.0, -P.qexp);
    }
    for (int m = 0; m < D; m++)
        if (P.btype[m] == 4) {
            const double gs = 1.0 / (2.0 * P.hmax);
            const int gl = (int)(gs * (P.bx[1][m] - P.bx[0][m]));
            if (gl + 2 < 3) { c->err = "periodic extent shorter than one cell"; return SPH_ERR_UNSUPPORTED; }
        }
    return SPH_OK;
}

int create_impl(Context *c, const sph_scene_desc *d)
{
    int rc = build_context(c, d);
    if (rc) return rc;
    Params &P = c->P;
    const int N = c->N, D = c->ndim;
    // ---- particle state in deck order ----
    std::vector<double> hx[3], hv[3], hm(N), hrho(N), hp(N);
    std::vector<uint32_t> hinfo(N);
    std::vector<int> rigid_idx;
    std::vector<int> per_mat(kMaxMat + 1, 0);
    for (int i = 0; i < N; i++) {
        const int mt = d->mat_id[i];
        if (mt < 1 || mt > d->nummat) {
            c->err = "Node " + std::to_string(i + 1) + " has incorrect material type.";
            return SPH_ERROR;
        }
        per_mat[mt]++;
    }
    for (int dd = 0; dd < D; dd++) {
        hx[dd].resize(N);
        hv[dd].resize(N);
    }
    double cmm[2][3];
    for (int m = 0; m < 3; m++) { cmm[0][m] = 1.0e20; cmm[1][m] = -1.0e20; }
    for (int i = 0; i < N; i++) {
        const int mt = d->mat_id[i];
        for (int dd = 0; dd < D; dd++) {
            hx[dd][i] = d->x[3 * (size_t)i + dd];
            hv[dd][i] = d->v ? d->v[3 * (size_t)i + dd] : 0.0;
            // MIN / MAX of global.h:21-22 (InitVariables, MSimulationInit.cpp:222-227)
            cmm[0][dd] = cmm[0][dd] < hx[dd][i] ? cmm[0][dd] : hx[dd][i];
            cmm[1][dd] = cmm[1][dd] > hx[dd][i] ? cmm[1][dd] : hx[dd][i];
            if (P.btype[dd] == 4 && (hx[dd][i] < P.bx[0][dd] || hx[dd][i] >= P.bx[1][dd])) {
                c->err = "particle outside the periodic interval at start";
                return SPH_ERR_UNSUPPORTED;
            }
        }
        // InitMass (MSimulationInit.cpp:144-177)
        if (d->massopt == 1) {
            if (!d->node_mass) { c->err = "massopt 1 needs node masses"; return SPH_ERROR; }
            hm[i] = d->node_mass[i];
        } else {
            hm[i] = d->mat[mt - 1].mass / per_mat[mt];
        }
        // InitRho_P_C / EOScalc (MSimulationInit.cpp:267-311, 384-440)
        const sph_material_desc &md = d->mat[mt - 1];
        const double rho0 = md.rho;
        const double eos3 = md.eosinput[3] == 0.0 ? 1.0 : md.eosinput[3];
        hrho[i] = (d->init_rhoe == 1 && d->node_rho) ? d->node_rho[i] : rho0 / eos3;
        double p;
        if (md.eos == 28) p = md.eosinput[0] * (pow(hrho[i] / rho0, md.eosinput[1]) - 1.0);
        else p = md.eosinput[0] * (hrho[i] - rho0);
        if (p > md.strinput0) p = md.strinput0;
        hp[i] = p;
        const int disp = d->dispbc ? d->dispbc[i] : 0;
        if (disp < 0 || disp > 8) { c->err = "dispbc out of range"; return SPH_ERROR; }
        const int color = d->color ? d->color[i] : 1;
        if (color < -32768 || color > 32767) { c->err = "colour out of range"; return SPH_ERROR; }
        uint32_t f = (uint32_t)mt | ((uint32_t)disp << kDispShift) | kActive |
                     ((uint32_t)(uint16_t)(int16_t)color << kColorShift);
        if (P.rigid && mt == P.rb_id) {
            f |= kRigid;
            rigid_idx.push_back(i);
        }
        hinfo[i] = f;
    }
    if (P.rigid && rigid_idx.empty()) { c->err = "rigid body has no particles"; return SPH_ERROR; }
    // compact FP32c momentum records: one mass per material and the linear EOS
    {
        bool uniform = c->arm == SPH_ARM_FP32C;
        std::vector<double> mref(kMaxMat + 1, -1.0);
        for (int i = 0; i < N && uniform; i++) {
            const int mt = d->mat_id[i];
            if (mref[mt] < 0.0) mref[mt] = hm[i];
            else if (hm[i] != mref[mt]) uniform = false;
        }
        for (int m = 1; m <= d->nummat && uniform; m++)
            if (d->mat[m - 1].eos != 29) uniform = false;
        // 2: one material and no rigid body, so the record needs no material
        // or rigid bits and carries 1/rho and p/rho^2 directly
        P.mcompact = uniform ? ((d->nummat == 1 && !P.rigid) ? 2 : 1) : 0;
        for (int m = 0; m <= kMaxMat; m++) P.mass32[m] = mref[m] > 0.0 ? (float)mref[m] : 0.0f;
    }
    // compact FP64 momentum records: one material with one particle mass, no
    // rigid body, the Morris discretisation (the mass is the particles' own)
    {
        bool one = c->arm == SPH_ARM_FP64 && d->nummat == 1 && !P.rigid && d->disctype == 2;
        for (int i = 1; i < N && one; i++)
            if (hm[i] != hm[0]) one = false;
        P.mcompact64 = one ? 1 : 0;
        P.mass64 = one ? hm[0] : 0.0;
    }
    P.n_rigid = (int)rigid_idx.size();

    // ---- capacities ----
    long long ncell0 = 1, ncode0 = 1;
    {
        const double gs = 1.0 / (2.0 * P.hmax);
        for (int m = 0; m < D; m++) {
            long long gl;
            if (P.btype[m] == 4) gl = (long long)(gs * (P.bx[1][m] - P.bx[0][m])) + 2;
            else gl = (long long)(gs * (cmm[1][m] - (cmm[0][m] - 0.01 * P.hmax))) + 1;
            gl = std::max(gl, 1LL);
            ncell0 *= gl;
            long long p2 = 1;
            while (p2 < gl) p2 <<= 1;
            ncode0 *= p2;          // cell-code space (power of two per axis)
        }
    }
    long long cap = std::max(ncell0, ncode0);
    cap += cap / 4 + 4096;
    if (cap > 0x7ffffff0LL) { c->err = "neighbour grid too large"; return SPH_ERROR; }
    int gcap = 0;
    if (P.periodic) {
        const double gs = 1.0 / (2.0 * P.hmax);
        long long g = 4096;
        for (int m = 0; m < D; m++)
            if (P.btype[m] == 4) {
                const long long gl = (long long)(gs * (P.bx[1][m] - P.bx[0][m]));
                g += 2 * ((long long)N * 3 / std::max(gl, 1LL)) ;
            }
        gcap = (int)std::min<long long>(g, 2LL * N + 4096);
    }
    HIPCHK(c, hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
    HIPCHK(c, hipHostMalloc((void **)&c->hctl, sizeof(Ctl), hipHostMallocDefault));
    memset(c->hctl, 0, sizeof(Ctl));
    // the control block and the bounding-box slots live as long as the context
    // (not part of the particle allocations): take them off the list
    if (dalloc(c, c->part_allocs, &c->AA[0].ctl, 1)) return SPH_ERR_HIP;
    c->part_allocs.pop_back();
    if (dalloc(c, c->part_allocs, &c->AA[0].bbox, (size_t)kBboxSlots * kBboxStride)) return SPH_ERR_HIP;
    {
        std::vector<unsigned long long> hb((size_t)kBboxSlots * kBboxStride, 0ull);
        for (int s = 0; s < kBboxSlots; s++)
            for (int m = 0; m < 3; m++) {
                hb[(size_t)s * kBboxStride + m] = dkey(1.0e20);
                hb[(size_t)s * kBboxStride + 3 + m] = dkey(-1.0e20);
                hb[(size_t)s * kBboxStride + kCritSlot] = dkey(1.0e20);
            }
        HIPCHK(c, hipMemcpy(c->AA[0].bbox, hb.data(), hb.size() * sizeof(unsigned long long), hipMemcpyHostToDevice));
    }
    c->part_allocs.pop_back();
    if (alloc_particles(c, gcap)) return SPH_ERR_HIP;
    if (alloc_cells(c, (int)cap)) return SPH_ERR_HIP;
    if (P.rigid) {
        if (dalloc(c, c->cell_allocs, &c->AA[0].rigid_idx, rigid_idx.size())) return SPH_ERR_HIP;
        // rigid index list must survive cell re-allocation: move to its own slot
        c->cell_allocs.pop_back();
        HIPCHK(c, hipMemcpy(c->AA[0].rigid_idx, rigid_idx.data(), rigid_idx.size() * sizeof(int),
                            hipMemcpyHostToDevice));
    }
    share_cells(c);
    c->cur = 0;
    Arrays &A = c->AA[0];
    {
        std::vector<int> orig(N), rrank(N, -1);
        for (int i = 0; i < N; i++) orig[i] = i;
        for (size_t k = 0; k < rigid_idx.size(); k++) rrank[rigid_idx[k]] = (int)k;
        HIPCHK(c, hipMemcpy(A.orig, orig.data(), N * 4, hipMemcpyHostToDevice));
        if (P.rigid) HIPCHK(c, hipMemcpy(A.rrank, rrank.data(), N * 4, hipMemcpyHostToDevice));
    }
    for (int dd = 0; dd < D; dd++) {
        HIPCHK(c, hipMemcpy(A.x[dd], hx[dd].data(), N * 8, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(A.xs[dd], hx[dd].data(), N * 8, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(A.v[dd], hv[dd].data(), N * 8, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemset(A.a[dd], 0, N * 8));
        if (dd == 0) HIPCHK(c, hipMemset(A.aa, 0, N * 8));
        if (P.st) {
            HIPCHK(c, hipMemset(A.nrm[dd], 0, N * 8));
            HIPCHK(c, hipMemset(A.nn[dd], 0, N * 8));
        }
    }
    HIPCHK(c, hipMemcpy(A.m, hm.data(), N * 8, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(A.rho, hrho.data(), N * 8, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(A.rho_new, hrho.data(), N * 8, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(A.p, hp.data(), N * 8, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemset(A.prho2, 0, N * 8));
    HIPCHK(c, hipMemcpy(A.info, hinfo.data(), N * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemset(A.trd, 0, N * 8));
    HIPCHK(c, hipMemset(A.nnbr, 0, N * 4));
    if (P.disctype == 0) {
        // q starts at zero (AViscosity at initialisation sees tr D = 0,
        // MSimulationInit.cpp:62-68); sigma and rod are written every step
        HIPCHK(c, hipMemset(A.q, 0, N * 8));
        for (int k = 0; k < 6; k++) {
            HIPCHK(c, hipMemset(A.rod[k], 0, N * 8));
            HIPCHK(c, hipMemset(A.sig[k], 0, N * 8));
        }
    }
    if (P.st) {
        HIPCHK(c, hipMemset(A.Nf, 0, N * 8));
        HIPCHK(c, hipMemset(A.kappa, 0, N * 8));
    }
    // ---- control block ----
    Ctl &h = *c->hctl;
    h.timestep = 1;
    h.ptime = 0.0;
    h.dt = h.dtold = 0.0;
    for (int s = 0; s < 2; s++)
        for (int m = 0; m < 3; m++) {
            h.cmm[s][m] = cmm[s][m];
        }
    h.critts_key = dkey(1.0e20);
    if (push_ctl(c)) return SPH_ERR_HIP;
    // ---- first neighbours (no images), critical step, first dt (:71-82) ----
    rc = run_phase(c, [&] {
        sort_step(c);
        launch_neighbours(c->P, cur(c), c->st);
        launch_accel_sq(c->P, cur(c), c->st);
        launch_strain(c->P, cur(c), M_CRIT, c->st);
        launch_crit_fold(c->P, cur(c), c->st);
    });
    if (rc) return rc;
    const double critts = dunkey(h.critts_key);
    h.critts = critts;
    h.critts_key = dkey(1.0e20);
    if (fabs(d->itss) > 1.e-10) h.dt = d->itss < critts * P.tssf ? d->itss : critts * P.tssf;
    else h.dt = critts * P.tssf;
    return push_ctl(c);
}

Context *ctx_of(void *p) { return (Context *)p; }

// deck number of every storage slot of the current order
int slot_orig(Context *c, std::vector<int> &orig)
{
    orig.resize(c->N);
    HIPCHK(c, hipMemcpy(orig.data(), cur(c).orig, c->N * 4, hipMemcpyDeviceToHost));
    return 0;
}

// ncomp: components kept on the device (the rest are written as zero);
// -1 for ndim
int download_vec(Context *c, double *const *src, double *out, int ncomp = -1)
{
    const int N = c->N;
    const int nc = ncomp < 0 ? c->ndim : ncomp;
    std::vector<double> tmp(N);
    std::vector<int> orig;
    if (slot_orig(c, orig)) return SPH_ERR_HIP;
    for (int d = 0; d < 3; d++) {
        if (d < nc && src[d]) {
            HIPCHK(c, hipMemcpy(tmp.data(), src[d], N * 8, hipMemcpyDeviceToHost));
            for (int i = 0; i < N; i++) out[3 * (size_t)orig[i] + d] = tmp[i];
        } else {
            for (int i = 0; i < N; i++) out[3 * (size_t)i + d] = 0.0;
        }
    }
    return 0;
}

}  // namespace

// =============================== C ABI =====================================
extern "C" {

int sph_desc_size(void) { return (int)sizeof(sph_scene_desc); }
int sph_material_desc_size(void) { return (int)sizeof(sph_material_desc); }

int sph_create(const sph_scene_desc *desc, int arm, void **out, char *err, int errlen)
{
    if (out) *out = nullptr;
    if (!desc || !out) return SPH_ERROR;
    Context *c = new Context();
    c->arm = arm;
    if (arm == SPH_ARM_FP64) {
        c->tune.strain_batch = kStrain64Batch;
        c->tune.momentum_batch = kMomentum64Batch;
    }
    const bool arm_ok = arm == SPH_ARM_FP64 || arm == SPH_ARM_FP32C;
    int rc = arm_ok ? create_impl(c, desc) : SPH_ERR_UNSUPPORTED;
    if (!arm_ok) c->err = "unknown arithmetic arm";
    if (rc) {
        if (err && errlen > 0) snprintf(err, (size_t)errlen, "%s", c->err.c_str());
        sph_destroy(c);
        return rc;
    }
    *out = c;
    return SPH_OK;
}

void sph_destroy(void *ctx)
{
    Context *c = ctx_of(ctx);
    if (!c) return;
    if (c->st) hipStreamSynchronize(c->st);
    drop_graph(c);
    free_list(c->part_allocs);
    free_list(c->cell_allocs);
    if (c->AA[0].ctl) hipFree(c->AA[0].ctl);
    if (c->AA[0].bbox) hipFree(c->AA[0].bbox);
    if (c->AA[0].rigid_idx) hipFree(c->AA[0].rigid_idx);
    if (c->canon) hipFree(c->canon);
    if (c->hctl) hipHostFree(c->hctl);
    harvest_timing(c);
    for (hipEvent_t e : c->tm.pool) hipEventDestroy(e);
    if (c->st) hipStreamDestroy(c->st);
    delete c;
}

const char *sph_last_error(void *ctx)
{
    Context *c = ctx_of(ctx);
    return c ? c->err.c_str() : "null context";
}

int sph_neighbours(void *ctx)
{
    Context *c = ctx_of(ctx);
    return run_phase(c, [&] { sort_step(c); });
}

int sph_ghost_setup(void *ctx)
{
    Context *c = ctx_of(ctx);
    if (!c->P.periodic) return SPH_OK;
    return run_phase(c, [&] { launch_ghost_setup(c->P, cur(c), c->st); });
}

int sph_linked_list_neighbours(void *ctx)
{
    Context *c = ctx_of(ctx);
    return run_phase(c, [&] { launch_neighbours(c->P, cur(c), c->st, c->tune.cell_slots); });
}

int sph_strain(void *ctx)
{
    Context *c = ctx_of(ctx);
    return run_phase(c, [&] {
        launch_shift(c->P, cur(c), c->st);
        launch_strain(c->P, cur(c), M_STRAIN, c->st);
    });
}

int sph_calculate_surface_normals(void *ctx)
{
    Context *c = ctx_of(ctx);
    if (!c->P.st) return SPH_OK;
    return run_phase(c, [&] { launch_normals(c->P, cur(c), c->st); });
}

int sph_constitutive(void *ctx)
{
    Context *c = ctx_of(ctx);
    return run_phase(c, [&] { launch_strain(c->P, cur(c), M_CONST | M_INPLACE, c->st); });
}

int sph_calc_critical_timestep(void *ctx, double *critts)
{
    Context *c = ctx_of(ctx);
    int rc = run_phase(c, [&] {
        launch_accel_sq(c->P, cur(c), c->st);
        launch_strain(c->P, cur(c), M_CRIT, c->st);
        launch_crit_fold(c->P, cur(c), c->st);
    });
    if (rc == 0 && critts) *critts = dunkey(c->hctl->critts_key);
    return rc;
}

int sph_time_advance(void *ctx)
{
    Context *c = ctx_of(ctx);
    return run_phase(c, [&] { launch_advance(c->P, cur(c), c->st); });
}

int sph_momentum(void *ctx)
{
    Context *c = ctx_of(ctx);
    return run_phase(c, [&] { launch_momentum(c->P, cur(c), 0, c->st); });
}

int sph_boundcond(void *ctx)
{
    Context *c = ctx_of(ctx);
    return run_phase(c, [&] { launch_integrate(c->P, cur(c), M_BC, c->st); });
}

int sph_update_velocity(void *ctx)
{
    Context *c = ctx_of(ctx);
    return run_phase(c, [&] {
        launch_rigid_kin(c->P, cur(c), c->st);
        launch_integrate(c->P, cur(c), M_VEL, c->st);
    });
}

int sph_move_particles(void *ctx)
{
    Context *c = ctx_of(ctx);
    return run_phase(c, [&] {
        launch_integrate(c->P, cur(c), M_MOVE, c->st);
        launch_finish(c->P, cur(c), 1, 0, c->st);
    });
}

int sph_end_step(void *ctx)
{
    Context *c = ctx_of(ctx);
    return run_phase(c, [&] { launch_finish(c->P, cur(c), 0, 1, c->st); });
}

// Captures kGraphSteps fused steps on the context's stream into the graph of
// the current buffer set (an even step count ends on the set it starts from).
constexpr int kGraphSteps = 16;
static_assert(kGraphSteps % 2 == 0, "a graph ends on the buffer set it starts from");

int build_graph(Context *c, int from)
{
    hipGraph_t g = nullptr;
    // capture records launches without running them: the steps start from
    // buffer set `from` whatever the current one is (an even count ends there)
    const int cur = c->cur;
    HIPCHK(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
    c->cur = from;
    for (int k = 0; k < kGraphSteps; k++) fused_step(c, k == 0, k == kGraphSteps - 1);
    c->cur = cur;
    HIPCHK(c, hipStreamEndCapture(c->st, &g));
    hipError_t e = hipGraphInstantiate(&c->gexec[from], g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) {
        c->gexec[from] = nullptr;
        c->use_graph = false;   // fall back to direct launches of the same kernels
        return 0;
    }
    c->gsteps = kGraphSteps;
    return 0;
}

// Autotuner: times the alternatives of each launch shape on the scene's
// current state and keeps the fastest.  For each cell storage order (Morton,
// row-major) it begins a step (sort and images, repairing capacities as the
// step loop does) and re-runs kernels that are idempotent on one step's
// inputs: the neighbour search (every slot count), the continuity sweep
// (whose rewrite of the positions is undone from a saved copy) and the
// momentum sweep.  The order with the smallest sum of the three fastest
// kernels wins.  Neither the order nor a shape changes a result (the lists
// are in the reference's traversal order whatever the storage order), and
// the following step re-sorts, so results are those of an untuned run.
constexpr float kTuneMargin = 0.01f;
int autotune(Context *c, int reps)
{
    Params &P = c->P;
    hipStream_t st = c->st;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    // one untimed launch (caches as the timed ones find them), then reps
    auto time_ms = [&](auto &&launch) -> float {
        float ms = 1e30f;
        launch();
        if (hipEventRecord(e0, st) != hipSuccess) return ms;
        for (int r = 0; r < reps; r++) launch();
        if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return ms;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) ms = 1e30f;
        return ms;
    };
    auto finish = [&](int rc) {
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        return rc;
    };
    // positions (the continuity sweep moves them a half step) and
    // accelerations (the momentum sweep rewrites them; the next step's
    // critical step reads them) are saved and put back
    const size_t n = (size_t)c->N;
    double *save = nullptr;
    if (hipMalloc((void **)&save, n * sizeof(double) * 2 * c->ndim) != hipSuccess) {
        c->err = "ERROR: autotuner allocation failed.";
        return finish(SPH_ERR_HIP);
    }
    Tune best = c->tune;
    float best_total = 1e30f;
    int rc = SPH_OK;
    bool stopped = false;
    for (int order : {0, 1}) {
        P.corder = order;
        for (int guard = 0;; guard++) {
            if (sync_ctl(c)) { rc = SPH_ERR_HIP; break; }
            if (c->hctl->done || c->hctl->err) { stopped = true; break; }
            sort_step(c);
            if (P.periodic) launch_ghost_setup(P, cur(c), st);
            const int r = check_device(c);
            if (r > 0) { rc = r; break; }
            if (r == 0) break;
            if (guard > 16) {
                c->err = "ERROR: capacity repair did not converge.";
                rc = SPH_ERROR;
                break;
            }
        }
        if (rc || stopped) break;
        const Arrays &A = cur(c);
        Tune t = c->tune;
        t.corder = order;
        // each shape's default first; another replaces it only when faster
        // by more than kTuneMargin (timing noise must not pick a slower one)
        float tn = 1e30f;
        for (int slots : {32, 16, 64, 128, 0}) {     // 0: the per-particle walk
            const float ms = time_ms([&] {
                hipMemsetAsync(&A.ctl->novf, 0, sizeof(int), st);
                launch_neighbours(P, A, st, slots);
            });
            if (ms < tn * (1.0f - kTuneMargin)) { tn = ms; t.cell_slots = slots; }
        }
        hipMemsetAsync(&A.ctl->novf, 0, sizeof(int), st);
        launch_neighbours(P, A, st, t.cell_slots);     // lists of this order for the sweeps
        for (int d = 0; d < c->ndim; d++) {
            hipMemcpyAsync(save + d * n, A.x[d], n * sizeof(double), hipMemcpyDeviceToDevice, st);
            hipMemcpyAsync(save + (c->ndim + d) * n, A.a[d], n * sizeof(double), hipMemcpyDeviceToDevice, st);
        }
        // records in flight per lane, the arm's default first
        const bool fp32c = c->arm == SPH_ARM_FP32C;
        const int gs_def = fp32c ? 8 : kStrain64Batch, gm_def = fp32c ? 8 : kMomentum64Batch;
        t.strain_batch = gs_def;
        t.momentum_batch = gm_def;
        float ts = 1e30f, tm = 1e30f;
        for (int g : {gs_def, 2, 4, 8}) {
            if (g == gs_def && ts < 1e30f) continue;
            if (fp32c && g == 2) continue;
            const float ms = time_ms([&] {
                if (fp32c) launch_strain32(P, A, st, g);
                else launch_strain(P, A, M_STRAIN | M_CONST | M_CRIT, st, g);
            });
            if (ms < ts * (1.0f - kTuneMargin)) { ts = ms; t.strain_batch = g; }
        }
        for (int g : {gm_def, 2, 4, 8}) {
            if (g == gm_def && tm < 1e30f) continue;
            const float ms = time_ms([&] {
                if (fp32c) launch_momentum32(P, A, st, g);
                else launch_momentum(P, A, -1, st, g);
            });
            if (ms < tm * (1.0f - kTuneMargin)) { tm = ms; t.momentum_batch = g; }
        }
        for (int d = 0; d < c->ndim; d++) {
            hipMemcpyAsync(A.x[d], save + d * n, n * sizeof(double), hipMemcpyDeviceToDevice, st);
            hipMemcpyAsync(A.a[d], save + (c->ndim + d) * n, n * sizeof(double), hipMemcpyDeviceToDevice, st);
        }
        // the timed continuity sweeps folded their critical steps into the
        // reduction slots: the step that follows starts from empty slots
        launch_crit_reset(A, st);
        if (hipStreamSynchronize(st) != hipSuccess) { rc = SPH_ERR_HIP; break; }
        if (tn + ts + tm < best_total) {
            best_total = tn + ts + tm;
            best = t;
        }
    }
    hipFree(save);
    finish(0);
    if (rc) return rc;
    c->tune = best;
    P.corder = best.corder;
    c->tuned = true;
    drop_graph(c);
    rc = check_device(c);
    return rc > 0 ? rc : SPH_OK;
}

int sph_solution(void *ctx, int max_steps, int *steps_done)
{
    Context *c = ctx_of(ctx);
    if (steps_done) *steps_done = 0;
    if (!c->tuned && c->N >= kAutotuneMinN && max_steps > 0) {
        const int rc = autotune(c, 5);
        if (rc) return rc;
    }
    // both buffer sets' graphs are captured on the first call (capture runs
    // nothing), so that later calls of any length launch ready graphs
    if (c->use_graph && !c->tm.on && max_steps > 0)
        for (const int from : {c->cur, c->cur ^ 1})
            if (!c->gexec[from] && build_graph(c, from)) return SPH_ERR_HIP;
    if (sync_ctl(c)) return SPH_ERR_HIP;
    const int t0 = c->hctl->timestep;
    int remaining = max_steps;
    int guard = 0;
    while (remaining > 0 && !c->hctl->done) {
        const int chunk = std::min(remaining, 64);
        int k = 0;
        if (c->use_graph && !c->tm.on && chunk >= kGraphSteps) {
            if (!c->gexec[c->cur] && build_graph(c, c->cur)) return SPH_ERR_HIP;
            if (c->gexec[c->cur])
                for (; k + c->gsteps <= chunk; k += c->gsteps)
                    HIPCHK(c, hipGraphLaunch(c->gexec[c->cur], c->st));
        }
        for (const int k1 = k; k < chunk; k++) fused_step(c, k == k1, k == chunk - 1);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            c->err = std::string("HIP launch error: ") + hipGetErrorString(e);
            return SPH_ERR_HIP;
        }
        int rc = check_device(c);
        if (rc > 0) {
            if (steps_done) *steps_done = c->hctl->timestep - t0;
            return rc;
        }
        if (rc == -1 && ++guard > 16) {
            c->err = "ERROR: capacity repair did not converge.";
            return SPH_ERROR;
        }
        remaining = max_steps - (c->hctl->timestep - t0);
    }
    if (steps_done) *steps_done = c->hctl->timestep - t0;
    return SPH_OK;
}

int sph_set_launch_shape(void *ctx, const int32_t *shape)
{
    Context *c = ctx_of(ctx);
    if (!c || !shape) return SPH_ERROR;
    const int slots = shape[0], gs = shape[1], gm = shape[2], order = shape[3];
    if ((slots != 0 && (slots < 4 || slots > 4096)) || (gs != 2 && gs != 4 && gs != 8) || (gm != 2 && gm != 4 && gm != 8) ||
        (order != 0 && order != 1)) {
        c->err = "sph_set_launch_shape: shape out of range";
        return SPH_ERROR;
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->tune.cell_slots = slots;
    c->tune.strain_batch = gs;
    c->tune.momentum_batch = gm;
    c->tune.corder = order;
    c->P.corder = order;
    c->tuned = true;
    drop_graph(c);
    return SPH_OK;
}

int sph_autotune(void *ctx, int reps, int32_t *shape)
{
    Context *c = ctx_of(ctx);
    int rc = SPH_OK;
    if (reps > 0) rc = autotune(c, reps);
    else if (reps == 0) c->tuned = true;
    if (shape) {
        shape[0] = c->tune.cell_slots;
        shape[1] = c->tune.strain_batch;
        shape[2] = c->tune.momentum_batch;
        shape[3] = c->tune.corder;
    }
    return rc;
}

int sph_synchronize(void *ctx)
{
    Context *c = ctx_of(ctx);
    HIPCHK(c, hipStreamSynchronize(c->st));
    return SPH_OK;
}

int sph_get_globals(void *ctx, sph_globals *out)
{
    Context *c = ctx_of(ctx);
    if (sync_ctl(c)) return SPH_ERR_HIP;
    const Ctl &h = *c->hctl;
    memset(out, 0, sizeof *out);
    out->dt = h.dt;
    out->dtold = h.dtold;
    out->ptime = h.ptime;
    out->critts = h.critts;
    for (int s = 0; s < 2; s++)
        for (int m = 0; m < 3; m++) out->coord_minmax[s][m] = h.cmm[s][m];
    out->hmax = c->P.hmax;
    for (int m = 0; m < 3; m++) {
        out->rb_R[m] = h.rbR[m];
        out->rb_V[m] = h.rbV[m];
        out->rb_OMEGA[m] = h.rbW[m];
        out->rb_A[m] = h.rbA[m];
        out->rb_ALPHA[m] = h.rbAL[m];
        out->grid_lim[m] = h.glim[m];
    }
    out->timestep = h.timestep;
    out->ngp = h.ngp;
    out->done = h.done;
    out->error = h.err;
    out->np = c->N;
    return SPH_OK;
}

int sph_download(void *ctx, int field, double *out)
{
    Context *c = ctx_of(ctx);
    HIPCHK(c, hipStreamSynchronize(c->st));
    const int N = c->N;
    const Arrays &A = cur(c);
    std::vector<uint32_t> info;
    std::vector<int> orig;
    std::vector<double> tmp;
    auto scalar = [&](const double *src) -> int {
        if (!src) {
            c->err = "field not kept for this scene";
            return SPH_ERROR;
        }
        if (slot_orig(c, orig)) return SPH_ERR_HIP;
        tmp.resize(N);
        HIPCHK(c, hipMemcpy(tmp.data(), src, N * 8, hipMemcpyDeviceToHost));
        for (int i = 0; i < N; i++) out[orig[i]] = tmp[i];
        return 0;
    };
    auto get_info = [&]() -> int {
        if (slot_orig(c, orig)) return SPH_ERR_HIP;
        std::vector<uint32_t> raw(N);
        HIPCHK(c, hipMemcpy(raw.data(), A.info, N * 4, hipMemcpyDeviceToHost));
        info.resize(N);
        for (int i = 0; i < N; i++) info[orig[i]] = raw[i];
        return 0;
    };
    switch (field) {
    case SPH_F_X: return download_vec(c, A.x, out);
    case SPH_F_V: return download_vec(c, A.v, out);
    case SPH_F_A: return download_vec(c, A.a, out);
    case SPH_F_NORMAL: return download_vec(c, A.nrm, out);
    case SPH_F_NNORMAL: return download_vec(c, A.nn, out);
    case SPH_F_Q:
        if (c->P.disctype != 0) {
            for (int i = 0; i < N; i++) out[i] = 0.0;
            return SPH_OK;
        }
        return scalar(A.q);
    case SPH_F_SIGMA_DIAG:
    case SPH_F_SIGMA_OFF:
    case SPH_F_ROD_DIAG:
    case SPH_F_ROD_OFF: {
        const bool sig = field == SPH_F_SIGMA_DIAG || field == SPH_F_SIGMA_OFF;
        const int k0 = (field == SPH_F_SIGMA_DIAG || field == SPH_F_ROD_DIAG) ? 0 : 3;
        if (c->P.disctype == 0) {
            double *const *t = sig ? A.sig : A.rod;
            double *src[3] = {t[k0], t[k0 + 1], t[k0 + 2]};
            return download_vec(c, src, out, 3);   // every component (a 2D zz too)
        }
        // disctype 2: sigma = -p I, no rate of deformation kept
        std::vector<double> p(N, 0.0);
        if (field == SPH_F_SIGMA_DIAG) {
            if (slot_orig(c, orig)) return SPH_ERR_HIP;
            tmp.resize(N);
            HIPCHK(c, hipMemcpy(tmp.data(), A.p, N * 8, hipMemcpyDeviceToHost));
            for (int i = 0; i < N; i++) p[orig[i]] = -tmp[i];
        }
        for (int i = 0; i < N; i++)
            for (int d = 0; d < 3; d++) out[3 * (size_t)i + d] = p[i];
        return SPH_OK;
    }
    case SPH_F_RHO: return scalar(A.rho);
    case SPH_F_P: return scalar(A.p);
    case SPH_F_MASS: return scalar(A.m);
    case SPH_F_TRD: return scalar(A.trd);
    case SPH_F_MINDIST: return scalar(A.mindist);
    case SPH_F_N: return scalar(A.Nf);
    case SPH_F_KAPPA: return scalar(A.kappa);
    case SPH_F_CRITTS: return scalar(A.critts_p);
    case SPH_F_C:
    case SPH_F_H:
    case SPH_F_ACTIVE:
        if (get_info()) return SPH_ERR_HIP;
        for (int i = 0; i < N; i++) {
            const Mat &mt = c->P.mat[info_mat(info[i])];
            out[i] = field == SPH_F_C ? mt.c : field == SPH_F_H ? mt.h : ((info[i] & kActive) ? 1.0 : 0.0);
        }
        return SPH_OK;
    case SPH_F_NNBR: {
        if (slot_orig(c, orig)) return SPH_ERR_HIP;
        std::vector<int> raw(N);
        HIPCHK(c, hipMemcpy(raw.data(), A.nnbr, N * 4, hipMemcpyDeviceToHost));
        for (int i = 0; i < N; i++) out[orig[i]] = raw[i];
        return SPH_OK;
    }
    case SPH_F_VABS: {
        std::vector<double> v(3 * (size_t)N);
        if (download_vec(c, A.v, v.data())) return SPH_ERR_HIP;
        for (int i = 0; i < N; i++) {
            double s = 0.0;
            for (int d = 0; d < c->ndim; d++) s += v[3 * (size_t)i + d] * v[3 * (size_t)i + d];
            out[i] = sqrt(s);
        }
        return SPH_OK;
    }
    default:
        c->err = "field not available";
        return SPH_ERROR;
    }
}

int sph_get_neighbours(void *ctx, int32_t *nnbr, int64_t *canon, int maxnbr)
{
    Context *c = ctx_of(ctx);
    const int N = c->N;
    if (!c->canon) HIPCHK(c, hipMalloc((void **)&c->canon, (size_t)N * c->maxnbr * 8));
    const int mn = std::min(maxnbr, c->maxnbr);
    launch_canon(c->P, cur(c), c->canon, mn, c->st);
    HIPCHK(c, hipStreamSynchronize(c->st));
    {
        std::vector<int> orig, raw(N);
        if (slot_orig(c, orig)) return SPH_ERR_HIP;
        HIPCHK(c, hipMemcpy(raw.data(), cur(c).nnbr, N * 4, hipMemcpyDeviceToHost));
        for (int i = 0; i < N; i++) nnbr[orig[i]] = raw[i];
    }
    std::vector<long long> tmp((size_t)N * mn);
    HIPCHK(c, hipMemcpy(tmp.data(), c->canon, tmp.size() * 8, hipMemcpyDeviceToHost));
    for (int i = 0; i < N; i++)
        for (int k = 0; k < nnbr[i] && k < mn; k++) canon[(size_t)i * maxnbr + k] = tmp[(size_t)i * mn + k];
    return SPH_OK;
}

// ---- MNeighbourVars surface ------------------------------------------------

int sph_grid_info(void *ctx, double *gridmin, double *gridsize, int32_t *gridlim)
{
    Context *c = ctx_of(ctx);
    if (!c) return SPH_ERROR;
    if (sync_ctl(c)) return SPH_ERR_HIP;
    for (int m = 0; m < 3; m++) {
        if (gridmin) gridmin[m] = c->hctl->gmin[m];
        if (gridsize) gridsize[m] = c->hctl->gsize[m];
        if (gridlim) gridlim[m] = c->hctl->glim[m];
    }
    return SPH_OK;
}

int sph_get_linked_list(void *ctx, int32_t *llgrid, long long ncell, int32_t *real_next, int32_t *ghost_next,
                        int32_t *ghost_par, int ngp)
{
    Context *c = ctx_of(ctx);
    if (!c) return SPH_ERROR;
    if (sync_ctl(c)) return SPH_ERR_HIP;
    const Ctl &h = *c->hctl;
    const long long n_cell = (long long)h.glim[0] * h.glim[1] * h.glim[2];
    const int n_img = c->P.periodic ? h.ngp : 0;
    if (llgrid && ncell < n_cell) {
        c->err = "sph_get_linked_list: llgrid holds fewer entries than the grid has cells";
        return SPH_ERROR;
    }
    if ((ghost_next || ghost_par) && ngp < n_img) {
        c->err = "sph_get_linked_list: image arrays hold fewer entries than there are images";
        return SPH_ERROR;
    }
    if ((ghost_next == nullptr) != (ghost_par == nullptr)) {
        c->err = "sph_get_linked_list: ghost_next and ghost_par go together";
        return SPH_ERROR;
    }
    const int N = c->N;
    const size_t bytes = (llgrid ? (size_t)n_cell : 0) + (real_next ? (size_t)N : 0) +
                         (ghost_next ? 2 * (size_t)n_img : 0);
    int *buf = nullptr;
    HIPCHK(c, hipMalloc((void **)&buf, (bytes ? bytes : 1) * sizeof(int)));
    int *d_grid = llgrid ? buf : nullptr;
    int *d_real = real_next ? buf + (llgrid ? n_cell : 0) : nullptr;
    int *d_gnext = ghost_next ? buf + (llgrid ? n_cell : 0) + (real_next ? N : 0) : nullptr;
    int *d_gpar = ghost_next ? d_gnext + n_img : nullptr;
    launch_linked_list(c->P, cur(c), d_grid, d_real, d_gnext, d_gpar, n_cell, c->st);
    hipError_t e = hipStreamSynchronize(c->st);
    if (e == hipSuccess && llgrid) e = hipMemcpy(llgrid, d_grid, (size_t)n_cell * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && real_next) e = hipMemcpy(real_next, d_real, (size_t)N * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && ghost_next && n_img) {
        e = hipMemcpy(ghost_next, d_gnext, (size_t)n_img * 4, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(ghost_par, d_gpar, (size_t)n_img * 4, hipMemcpyDeviceToHost);
    }
    hipFree(buf);
    HIPCHK(c, e);
    return SPH_OK;
}

int sph_get_grid_value(void *ctx, int i, int j, int k, int32_t *value)
{
    Context *c = ctx_of(ctx);
    if (!c || !value) return SPH_ERROR;
    if (sync_ctl(c)) return SPH_ERR_HIP;
    const Ctl &h = *c->hctl;
    if (i < 0 || j < 0 || k < 0 || i >= h.glim[0] || j >= h.glim[1] || k >= h.glim[2]) {
        c->err = "GetGridValue: cell outside the grid";
        return SPH_ERR_OUT_OF_GRID;
    }
    std::vector<int32_t> grid((size_t)h.glim[0] * h.glim[1] * h.glim[2]);
    const int rc = sph_get_linked_list(ctx, grid.data(), (long long)grid.size(), nullptr, nullptr, nullptr, 0);
    if (rc) return rc;
    *value = grid[(size_t)h.glim[2] * ((size_t)h.glim[1] * i + j) + k];
    return SPH_OK;
}

int sph_get_box_coords(void *ctx, const double *x, int32_t *boxcrd)
{
    Context *c = ctx_of(ctx);
    if (!c || !x || !boxcrd) return SPH_ERROR;
    if (sync_ctl(c)) return SPH_ERR_HIP;
    // RealVector temp = x - gridmin; boxcrd(i) = gridsize(i) * temp(i), truncated
    for (int m = 0; m < 3; m++) boxcrd[m] = (int32_t)(c->hctl->gsize[m] * (x[m] - c->hctl->gmin[m]));
    return SPH_OK;
}

int sph_get_loop_limits(void *ctx, int ndim, const int32_t *boxcrd, int32_t *looplim)
{
    Context *c = ctx_of(ctx);
    if (!c || !boxcrd || !looplim) return SPH_ERROR;
    if (sync_ctl(c)) return SPH_ERR_HIP;
    for (int m = 0; m < 3; m++) {
        looplim[m] = std::max(boxcrd[m] - 1, 0);
        looplim[3 + m] = std::min(boxcrd[m] + 1, c->hctl->glim[m] - 1);
    }
    for (int m = ndim; m < 3; m++) looplim[m] = looplim[3 + m] = 0;
    return SPH_OK;
}

// ---- MKernel surface -------------------------------------------------------

int sph_kernel_eval(int ndim, double havg, int n, const double *xi, const double *xj, double *w, double *dwdx,
                    double *dwdr)
{
    if (ndim < 1 || ndim > 3 || !(havg > 0.0) || n < 0 || (n > 0 && (!xi || !xj))) return SPH_ERROR;
    if (n == 0) return SPH_OK;
    const size_t v3 = 3 * (size_t)n * sizeof(double), v1 = (size_t)n * sizeof(double);
    double *d = nullptr;
    if (hipMalloc((void **)&d, 2 * v3 + v3 + 2 * v1) != hipSuccess) return SPH_ERR_HIP;
    double *dxi = d, *dxj = d + 3 * (size_t)n, *ddw = d + 6 * (size_t)n, *dw = d + 9 * (size_t)n,
           *ddr = d + 10 * (size_t)n;
    const PairConst pc = kernel_const(ndim, havg);
    // a private stream: the call neither waits for nor delays the scenes' streams
    hipStream_t st = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMemcpyAsync(dxi, xi, v3, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dxj, xj, v3, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        launch_kernel_eval(ndim, pc, n, dxi, dxj, w ? dw : nullptr, dwdx ? ddw : nullptr, dwdr ? ddr : nullptr, st);
        e = hipGetLastError();
    }
    if (e == hipSuccess && w) e = hipMemcpyAsync(w, dw, v1, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && dwdx) e = hipMemcpyAsync(dwdx, ddw, v3, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && dwdr) e = hipMemcpyAsync(dwdr, ddr, v1, hipMemcpyDeviceToHost, st);
    if (st) {
        const hipError_t es = hipStreamSynchronize(st);
        if (e == hipSuccess) e = es;
        hipStreamDestroy(st);
    }
    hipFree(d);
    return e == hipSuccess ? SPH_OK : SPH_ERR_HIP;
}

}  // extern "C"

// ---- sph_kernel_eval without allocation -------------------------------------
// The staging of sph_kernel_eval (coordinates in, W / dW/dx / dW/dr out: 11
// doubles per pair) lives in one static device array per GPU instead of a
// per-call hipMalloc.  Host code never uses the array's name as an address:
// its device address is looked up once per device with hipGetSymbolAddress
// (with that device current) and kept in scratch_dev[].
namespace {

constexpr int kNoAllocMaxPairs = SPH_KERNEL_EVAL_NOALLOC_MAX;
constexpr size_t kScratchDoubles = 11 * (size_t)kNoAllocMaxPairs;
constexpr int kMaxDevices = 64;

__device__ __align__(256) double scratch[kScratchDoubles];

std::mutex scratch_mu;
double *scratch_dev[kMaxDevices] = {};

// The current device's address of `scratch`, looked up on the first call on
// that device; nullptr on failure.
double *scratch_ptr()
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        scratch_dev[dev] = (double *)p;
    }
    return scratch_dev[dev];
}

}  // namespace

extern "C" int sph_kernel_eval_noalloc(int ndim, double havg, int n, const double *xi, const double *xj,
                                       double *w, double *dwdx, double *dwdr)
{
    if (ndim < 1 || ndim > 3 || !(havg > 0.0) || n < 0 || (n > 0 && (!xi || !xj))) return SPH_ERROR;
    if (n > kNoAllocMaxPairs) return SPH_ERR_CAPACITY;
    if (n == 0) return SPH_OK;
    double *d = scratch_ptr();
    if (!d) return SPH_ERR_HIP;
    const size_t v3 = 3 * (size_t)n * sizeof(double), v1 = (size_t)n * sizeof(double);
    // the same layout as sph_kernel_eval's staging, packed at the front
    double *dxi = d, *dxj = d + 3 * (size_t)n, *ddw = d + 6 * (size_t)n, *dw = d + 9 * (size_t)n,
           *ddr = d + 10 * (size_t)n;
    const PairConst pc = kernel_const(ndim, havg);
    // the calling thread's default stream: nothing is created
    hipStream_t st = hipStreamPerThread;
    hipError_t e = hipMemcpyAsync(dxi, xi, v3, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dxj, xj, v3, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        launch_kernel_eval(ndim, pc, n, dxi, dxj, w ? dw : nullptr, dwdx ? ddw : nullptr, dwdr ? ddr : nullptr, st);
        e = hipGetLastError();
    }
    if (e == hipSuccess && w) e = hipMemcpyAsync(w, dw, v1, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && dwdx) e = hipMemcpyAsync(dwdx, ddw, v3, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && dwdr) e = hipMemcpyAsync(dwdr, ddr, v1, hipMemcpyDeviceToHost, st);
    const hipError_t es = hipStreamSynchronize(st);
    if (e == hipSuccess) e = es;
    return e == hipSuccess ? SPH_OK : SPH_ERR_HIP;
}

// ---- kernel timing (HIP events on the context's stream) --------------------
extern "C" {

int sph_set_timing(void *ctx, int enable)
{
    Context *c = ctx_of(ctx);
    HIPCHK(c, hipStreamSynchronize(c->st));
    harvest_timing(c);
    c->tm.on = enable != 0;
    for (int k = 0; k < K_COUNT; k++) {
        c->tm.total_ms[k] = 0.0;
        c->tm.count[k] = 0;
    }
    return SPH_OK;
}

int sph_get_timing(void *ctx, int slot, char *name, int namelen, double *total_ms, long long *count)
{
    Context *c = ctx_of(ctx);
    if (slot < 0) return K_COUNT;
    if (slot >= K_COUNT) return SPH_ERROR;
    HIPCHK(c, hipStreamSynchronize(c->st));
    harvest_timing(c);
    if (name && namelen > 0) snprintf(name, (size_t)namelen, "%s", kKernelNames[slot]);
    if (total_ms) *total_ms = c->tm.total_ms[slot];
    if (count) *count = c->tm.count[slot];
    return SPH_OK;
}

}  // extern "C"

// ============================ batched scenes =================================
// Many independent scenes advanced by one launch sequence per step (the
// ensemble and serving workloads).  Each scene keeps its own device state
// and control block; the batch holds device tables of their parameter and
// array blocks and the block maps of the batched kernels (launch_batch_step).
namespace {

// candidates of the batch tuner: (cell storage order, slots per wave of the
// cell kernel; 0: the per-particle walk)
constexpr int kBatchCand[][2] = {{0, 64}, {1, 64}, {0, 32}, {1, 32}, {0, 128}, {1, 128}, {0, 0}, {1, 0}};
constexpr int kBatchNCand = (int)(sizeof(kBatchCand) / sizeof(kBatchCand[0]));

struct Batch {
    hipStream_t st = nullptr;
    int ndim = 2, arm = SPH_ARM_FP64;
    std::vector<Context *> slots;          // nullptr = free slot
    std::vector<int> live;                 // resident, not finished, no error (in the device tables)
    std::vector<int> pending;              // admitted, joining the live set at the next launch
    Params *dP = nullptr;
    Arrays *dA = nullptr, *dA2 = nullptr;   // per slot: current / other buffer set
    int cap_slots = 0;
    int *dscenes = nullptr;
    int2 *dpmap = nullptr, *dtmap = nullptr, *dwmap = nullptr;
    int2 *dvmap[kSweep32Variants] = {};
    int cap_vmap[kSweep32Variants] = {};
    int4 *dcmap = nullptr;
    int cap_scenes = 0, cap_pmap = 0, cap_tmap = 0, cap_cmap = 0, cap_wmap = 0;
    Ctl *dctl = nullptr, *hctl = nullptr; // gathered control blocks
    int cap_ctl = 0;
    BatchView view{};
    bool dirty = true;
    hipGraphExec_t gexec2[2] = {nullptr, nullptr};   // captured steps, per starting parity
    int phase = 0;                                   // 1: tables swapped since the rebuild
    std::string err;
    // launch shapes of the batch (timing only, never results): cell storage
    // order of every scene and slots per wave of the cell kernel, chosen
    // online (see sph_batch_step)
    int corder = 0, cell_slots = kBatchCellSlots;
    bool tuned = false;                  // shapes fixed (tuned, or tuning turned off)
    int trial = 0;                       // next candidate to time
    double trial_cost[kBatchNCand] = {0};   // ms per million particle-steps of each candidate
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool trial_pending = false;          // a timed trial is in flight (read at the next sync)
    long long trial_steps = 0;           // particle-steps of that trial
    // steps queued on the stream and not yet synchronised (sph_batch_launch):
    // scenes removed meanwhile are freed at the next synchronisation, because
    // the queued kernels still read their tables and control blocks
    bool inflight = false;
    std::vector<char> zombie;            // per slot: removed, freed at the next sync
};

// batches of at least this many live particles tune their shapes online
constexpr long long kBatchTuneMinN = 65536;

#define BCHK(b, expr)                                                               \
    do {                                                                            \
        hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess) {                                                     \
            (b)->err = std::string("HIP error: ") + hipGetErrorString(e_) + " at " #expr; \
            return SPH_ERR_HIP;                                                     \
        }                                                                           \
    } while (0)

void batch_drop_graph(Batch *b)
{
    for (auto &g : b->gexec2) {
        if (g) hipGraphExecDestroy(g);
        g = nullptr;
    }
}

template <class T>
int grow_dev(Batch *b, T **p, int *cap, int need)
{
    if (need <= *cap) return 0;
    int nc = std::max(need, 2 * *cap + 16);
    if (*p) BCHK(b, hipFree(*p));
    BCHK(b, hipMalloc((void **)p, (size_t)nc * sizeof(T)));
    *cap = nc;
    return 0;
}

// Rebuilds the scene list, the block maps and the per-slot tables.
int batch_rebuild(Batch *b)
{
    BCHK(b, hipStreamSynchronize(b->st));
    batch_drop_graph(b);
    const int nslot = (int)b->slots.size();
    if (grow_dev(b, &b->dP, &b->cap_slots, nslot)) return SPH_ERR_HIP;
    {
        int cap2 = b->cap_slots - 1;
        (void)cap2;
    }
    // Arrays table has the same capacity as the Params table
    static_assert(sizeof(Arrays) > 0, "");
    std::vector<int> scenes;
    std::vector<int2> pmap, tmap, wmap, vmap[kSweep32Variants];
    std::vector<int4> cmap;
    int any_p = 0, any_st = 0, any_r = 0, mask = 0;
    for (int s : b->live) {
        Context *c = b->slots[s];
        c->P.corder = b->corder;           // the next step sorts in this order
        scenes.push_back(s);
        const int vv = sweep32_variant(c->P);
        for (int lb = 0; lb < (c->N + kBlock - 1) / kBlock; lb++) {
            pmap.push_back(make_int2(s, lb));
            vmap[vv].push_back(make_int2(s, lb));
        }
        // cell_slots 0: the per-particle walk for every scene (a tuning candidate)
        if (cell_lists_fit(c->P) && b->cell_slots > 0) {
            const int nb = batch_cell_blocks(c->N, b->cell_slots);
            for (int lb = 0; lb < nb; lb++) cmap.push_back(make_int4(s, lb, nb, 0));
        } else {
            for (int lb = 0; lb < (c->N + kBlock - 1) / kBlock; lb++) wmap.push_back(make_int2(s, lb));
        }
        for (int t = 0; t < (c->P.ncell_cap + kScanTileItems - 1) / kScanTileItems; t++) tmap.push_back(make_int2(s, t));
        any_p |= c->P.periodic;
        any_st |= c->P.st;
        any_r |= c->P.rigid;
        for (int m = 0; m < 3; m++)
            for (int side = 0; side < 2; side++) {
                const int code = (m == 0) ? c->P.bcode[0][0] : c->P.bcode[side][m];
                if (c->P.periodic && code == 2) mask |= 1 << (2 * m + side);
            }
    }
    if (grow_dev(b, &b->dscenes, &b->cap_scenes, (int)scenes.size() + 1)) return SPH_ERR_HIP;
    if (grow_dev(b, &b->dpmap, &b->cap_pmap, (int)pmap.size() + 1)) return SPH_ERR_HIP;
    for (int vv = 0; vv < kSweep32Variants; vv++)
        if (grow_dev(b, &b->dvmap[vv], &b->cap_vmap[vv], (int)vmap[vv].size() + 1)) return SPH_ERR_HIP;
    if (grow_dev(b, &b->dtmap, &b->cap_tmap, (int)tmap.size() + 1)) return SPH_ERR_HIP;
    if (grow_dev(b, &b->dcmap, &b->cap_cmap, (int)cmap.size() + 1)) return SPH_ERR_HIP;
    if (grow_dev(b, &b->dwmap, &b->cap_wmap, (int)wmap.size() + 1)) return SPH_ERR_HIP;
    if (grow_dev(b, &b->dctl, &b->cap_ctl, nslot + 1)) return SPH_ERR_HIP;
    if (b->hctl) hipHostFree(b->hctl);
    BCHK(b, hipHostMalloc((void **)&b->hctl, (size_t)b->cap_ctl * sizeof(Ctl), hipHostMallocDefault));
    // parameter / array tables indexed by slot
    std::vector<Params> hp(nslot);
    std::vector<Arrays> ha(nslot), ha2(nslot);
    for (int s = 0; s < nslot; s++)
        if (b->slots[s]) {
            Context *c = b->slots[s];
            hp[s] = c->P;
            ha[s] = c->AA[c->cur];
            ha2[s] = c->AA[c->cur ^ 1];
        }
    if (b->dA) BCHK(b, hipFree(b->dA));
    if (b->dA2) BCHK(b, hipFree(b->dA2));
    BCHK(b, hipMalloc((void **)&b->dA, (size_t)std::max(nslot, 1) * sizeof(Arrays)));
    BCHK(b, hipMalloc((void **)&b->dA2, (size_t)std::max(nslot, 1) * sizeof(Arrays)));
    if (nslot) {
        BCHK(b, hipMemcpy(b->dP, hp.data(), nslot * sizeof(Params), hipMemcpyHostToDevice));
        BCHK(b, hipMemcpy(b->dA, ha.data(), nslot * sizeof(Arrays), hipMemcpyHostToDevice));
        BCHK(b, hipMemcpy(b->dA2, ha2.data(), nslot * sizeof(Arrays), hipMemcpyHostToDevice));
    }
    if (!scenes.empty()) {
        BCHK(b, hipMemcpy(b->dscenes, scenes.data(), scenes.size() * sizeof(int), hipMemcpyHostToDevice));
        BCHK(b, hipMemcpy(b->dpmap, pmap.data(), pmap.size() * sizeof(int2), hipMemcpyHostToDevice));
        for (int vv = 0; vv < kSweep32Variants; vv++)
            if (!vmap[vv].empty())
                BCHK(b, hipMemcpy(b->dvmap[vv], vmap[vv].data(), vmap[vv].size() * sizeof(int2), hipMemcpyHostToDevice));
        BCHK(b, hipMemcpy(b->dtmap, tmap.data(), tmap.size() * sizeof(int2), hipMemcpyHostToDevice));
        if (!cmap.empty())
            BCHK(b, hipMemcpy(b->dcmap, cmap.data(), cmap.size() * sizeof(int4), hipMemcpyHostToDevice));
        if (!wmap.empty())
            BCHK(b, hipMemcpy(b->dwmap, wmap.data(), wmap.size() * sizeof(int2), hipMemcpyHostToDevice));
    }
    BatchView &v = b->view;
    v.P = b->dP;
    v.A = b->dA;
    v.A2 = b->dA2;
    v.scenes = b->dscenes;
    v.pmap = b->dpmap;
    for (int vv = 0; vv < kSweep32Variants; vv++) {
        v.vmap[vv] = b->dvmap[vv];
        v.nvblocks[vv] = (int)vmap[vv].size();
    }
    v.tmap = b->dtmap;
    v.nscene = (int)scenes.size();
    v.npart_blocks = (int)pmap.size();
    v.ntile_blocks = (int)tmap.size();
    v.cmap = b->dcmap;
    v.wmap = b->dwmap;
    v.ncell_blocks = (int)cmap.size();
    v.nwalk_blocks = (int)wmap.size();
    v.zero_nb = 64;
    v.slab_nb = 8;
    v.ndim = b->ndim;
    v.arm = b->arm;
    v.any_periodic = any_p;
    v.any_st = any_st;
    v.any_rigid = any_r;
    v.pass_mask = mask;
    b->dirty = false;
    return 0;
}

// Gathers every live scene's control block, repairs capacity errors and
// drops finished or failed scenes from the live set.
int batch_check(Batch *b)
{
    const int ns = (int)b->live.size();
    if (ns == 0) return 0;
    launch_gather_ctl(b->dA, b->dscenes, ns, b->dctl, b->st);
    BCHK(b, hipMemcpyAsync(b->hctl, b->dctl, ns * sizeof(Ctl), hipMemcpyDeviceToHost, b->st));
    BCHK(b, hipStreamSynchronize(b->st));
    std::vector<int> keep;
    for (int k = 0; k < ns; k++) {
        const int s = b->live[k];
        Context *c = b->slots[s];
        *c->hctl = b->hctl[k];
        Ctl &h = *c->hctl;
        if (h.err == kErrGridCap || h.err == kErrGhostCap) {
            int rc = 0;
            if (h.err == kErrGridCap) {
                long long want = (long long)h.ncell * 3 / 2 + 1024;
                if (want > 0x7fffffffLL - 1) rc = SPH_ERROR;
                else rc = alloc_cells(c, (int)want);
            } else {
                rc = grow_particles(c, std::max(2 * c->P.gcap, h.err_particle + 1024));
            }
            if (rc == 0) {
                h.err = 0;
                rc = push_ctl(c);
            }
            b->dirty = true;
            if (rc) {
                c->err = "capacity repair failed";
                continue;
            }
            keep.push_back(s);
            continue;
        }
        if (h.err) {
            c->err = err_text(h.err);
            b->dirty = true;
            continue;
        }
        if (h.done) {
            b->dirty = true;
            continue;
        }
        keep.push_back(s);
    }
    b->live.swap(keep);
    return 0;
}

}  // namespace

extern "C" {

int sph_batch_create(int ndim, int arm, void **out)
{
    if (out) *out = nullptr;
    if (ndim < 1 || ndim > 3 || (arm != SPH_ARM_FP64 && arm != SPH_ARM_FP32C)) return SPH_ERROR;
    Batch *b = new Batch();
    b->ndim = ndim;
    b->arm = arm;
    if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return SPH_ERR_HIP;
    }
    *out = b;
    return SPH_OK;
}

const char *sph_batch_last_error(void *batch)
{
    return batch ? ((Batch *)batch)->err.c_str() : "null batch";
}

// Bytes of device memory a scene of this deck will hold (admission control).
long long sph_scene_footprint(const sph_scene_desc *d)
{
    const long long n = d->np;
    const long long D = d->ndim;
    const long long maxnbr = d->maxnbr > 0 ? d->maxnbr : 40;
    // per slot (reals and images): x, v in both buffer sets, xs, rho and m in
    // both sets, info in both sets, the two gather records, rho_new/p/prho2,
    // the neighbour list
    // (the list is counted twice: its chunk stride is a power of two >= slots)
    const long long per_slot = 32 * D + 8 * D + 32 + 8 + 144 + 24 + 8 * ((long long)maxnbr + 16);
    // per real particle: a and aa in both sets, orig, trd/mindist/critts, nnbr, sort
    // scratch, CSF fields
    const long long per_real = 16 * D + 16 + 8 + 24 + 4 + 12 + (d->surface_tension ? 16 * D + 16 : 0);
    const long long img = d->boundary ? n / 4 + 4096 : 0;
    return (n + img) * per_slot + n * per_real + n * 16 /* cell arrays, estimate */ + 65536;
}

int sph_batch_add(void *batch, const sph_scene_desc *desc, int step_limit, int *slot, char *err, int errlen)
{
    Batch *b = (Batch *)batch;
    if (slot) *slot = -1;
    if (desc->ndim != b->ndim) {
        if (err && errlen > 0) snprintf(err, (size_t)errlen, "scene ndim differs from the batch");
        return SPH_ERROR;
    }
    if (desc->disctype != 2) {
        if (err && errlen > 0) snprintf(err, (size_t)errlen, "batched scenes use disctype 2 (Morris)");
        return SPH_ERR_UNSUPPORTED;
    }
    // no drain: the scene is initialised on its own stream while queued
    // batch steps run, and joins the live set at the next launch (the tables
    // are rebuilt there)
    void *ctx = nullptr;
    int rc = sph_create(desc, b->arm, &ctx, err, errlen);
    if (rc) return rc;
    Context *c = ctx_of(ctx);
    c->hctl->step_limit = step_limit;
    c->P.corder = b->corder;
    if (push_ctl(c)) {
        if (err && errlen > 0) snprintf(err, (size_t)errlen, "%s", c->err.c_str());
        sph_destroy(c);
        return SPH_ERR_HIP;
    }
    int s = -1;
    for (int k = 0; k < (int)b->slots.size(); k++)
        if (!b->slots[k]) { s = k; break; }
    if (s < 0) {
        s = (int)b->slots.size();
        b->slots.push_back(nullptr);
        b->zombie.push_back(0);
    }
    b->slots[s] = c;
    b->pending.push_back(s);
    if (slot) *slot = s;
    return SPH_OK;
}

int sph_batch_remove(void *batch, int slot)
{
    Batch *b = (Batch *)batch;
    if (slot < 0 || slot >= (int)b->slots.size() || !b->slots[slot] || b->zombie[slot]) return SPH_ERROR;
    const auto pend = std::find(b->pending.begin(), b->pending.end(), slot);
    if (pend != b->pending.end()) {
        b->pending.erase(pend);    // never in the device tables
    } else if (b->inflight) {
        b->zombie[slot] = 1;       // queued steps still read it: freed at the next sync
        return SPH_OK;
    } else {
        b->live.erase(std::remove(b->live.begin(), b->live.end(), slot), b->live.end());
        b->dirty = true;
    }
    sph_destroy(b->slots[slot]);
    b->slots[slot] = nullptr;
    return SPH_OK;
}

// Advances every live scene by up to nsteps steps; scenes stop on their own
// step limit or stop test.  Returns the number of scenes still live in *live.
// Online tuning: while a batch of at least kBatchTuneMinN live particles is
// untuned, each step chunk that is one captured graph runs with the next
// candidate shape and is timed with events; its cost per particle-step is
// kept, and after the last candidate the cheapest is fixed.  Every candidate
// gives the same results, so the tuning is done on the run's own steps.
static long long batch_particles(const Batch *b)
{
    long long n = 0;
    for (int s : b->live) n += b->slots[s]->N;
    return n;
}

constexpr int kBatchGraphSteps = 8;    // even: a graph starts and ends on the same parity

// Queues nsteps steps of every live scene on the batch stream (a rebuild of
// the tables, after admissions or retirements, drains the stream first).
// With online tuning pending, the first captured graph runs the next
// candidate shape between two events, read at the next sync.
// Admitted scenes join the live set (and the tables at the next rebuild);
// only while no steps are queued, so the tables match the live set at sync.
static void merge_pending(Batch *b)
{
    if (b->pending.empty()) return;
    b->live.insert(b->live.end(), b->pending.begin(), b->pending.end());
    b->pending.clear();
    b->dirty = true;
}

static int batch_launch(Batch *b, int nsteps)
{
    merge_pending(b);
    if (b->live.empty() || nsteps <= 0) return 0;
    const long long npart = batch_particles(b);
    const bool tuning = !b->tuned && !b->trial_pending && npart >= kBatchTuneMinN && nsteps >= kBatchGraphSteps;
    if (tuning && (b->corder != kBatchCand[b->trial][0] || b->cell_slots != kBatchCand[b->trial][1])) {
        b->corder = kBatchCand[b->trial][0];
        b->cell_slots = kBatchCand[b->trial][1];
        b->dirty = true;
    }
    if (b->dirty) {
        if (batch_rebuild(b)) return SPH_ERR_HIP;
        b->phase = 0;
    }
    if (tuning && !b->ev[0]) {
        BCHK(b, hipEventCreate(&b->ev[0]));
        BCHK(b, hipEventCreate(&b->ev[1]));
    }
    int k = 0;
    BatchView views[2] = {b->view, b->view};
    std::swap(views[1].A, views[1].A2);
    if (nsteps >= kBatchGraphSteps) {
        hipGraphExec_t &gx = b->gexec2[b->phase];
        if (!gx) {
            hipGraph_t g = nullptr;
            BCHK(b, hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
            for (int q = 0; q < kBatchGraphSteps; q++) launch_batch_step(views[(b->phase + q) & 1], b->st);
            BCHK(b, hipStreamEndCapture(b->st, &g));
            if (hipGraphInstantiate(&gx, g, nullptr, nullptr, 0) != hipSuccess) gx = nullptr;
            hipGraphDestroy(g);
        }
        if (gx) {
            // (timed after the capture, so a trial's cost is its steps only)
            if (tuning) BCHK(b, hipEventRecord(b->ev[0], b->st));
            for (; k + kBatchGraphSteps <= nsteps; k += kBatchGraphSteps) {
                BCHK(b, hipGraphLaunch(gx, b->st));
                if (tuning && k == 0) {
                    BCHK(b, hipEventRecord(b->ev[1], b->st));
                    b->trial_pending = true;
                    b->trial_steps = npart * kBatchGraphSteps;
                }
            }
        }
    }
    // each remaining step sorts every live scene into its other buffer set
    for (; k < nsteps; k++) {
        launch_batch_step(views[b->phase], b->st);
        b->phase ^= 1;
        for (int s : b->live) b->slots[s]->cur ^= 1;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        b->err = std::string("HIP launch error: ") + hipGetErrorString(e);
        return SPH_ERR_HIP;
    }
    b->inflight = true;
    return 0;
}

// Waits for the queued steps, folds a pending tuning trial, updates the live
// set (finished and failed scenes leave it) and frees the scenes removed
// while steps were queued.
static int batch_sync(Batch *b)
{
    BCHK(b, hipStreamSynchronize(b->st));
    b->inflight = false;
    if (b->trial_pending) {
        float ms = 0.f;
        BCHK(b, hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
        b->trial_cost[b->trial] = (double)ms * 1e6 / (double)std::max(b->trial_steps, 1LL);
        b->trial_pending = false;
        if (++b->trial == kBatchNCand) {
            int best = 0;
            for (int k = 1; k < kBatchNCand; k++)
                if (b->trial_cost[k] < b->trial_cost[best]) best = k;
            b->corder = kBatchCand[best][0];
            b->cell_slots = kBatchCand[best][1];
            b->tuned = true;
            b->dirty = true;
        }
    }
    if (batch_check(b)) return SPH_ERR_HIP;
    for (int s = 0; s < (int)b->slots.size(); s++)
        if (b->zombie[s]) {
            b->live.erase(std::remove(b->live.begin(), b->live.end(), s), b->live.end());
            b->dirty = true;
            sph_destroy(b->slots[s]);
            b->slots[s] = nullptr;
            b->zombie[s] = 0;
        }
    return 0;
}

int sph_batch_step(void *batch, int nsteps, int *live)
{
    Batch *b = (Batch *)batch;
    int done_steps = 0;
    if (b->inflight && batch_sync(b)) return SPH_ERR_HIP;
    merge_pending(b);
    while (done_steps < nsteps && !b->live.empty()) {
        const bool tuning = !b->tuned && batch_particles(b) >= kBatchTuneMinN && nsteps - done_steps >= kBatchGraphSteps;
        const int chunk = tuning ? kBatchGraphSteps : std::min(nsteps - done_steps, 32);
        if (batch_launch(b, chunk)) return SPH_ERR_HIP;
        done_steps += chunk;
        if (batch_sync(b)) return SPH_ERR_HIP;
    }
    if (live) *live = (int)(b->live.size() + b->pending.size());
    return SPH_OK;
}

int sph_batch_launch(void *batch, int nsteps)
{
    Batch *b = (Batch *)batch;
    if (b->inflight && batch_sync(b)) return SPH_ERR_HIP;
    return batch_launch(b, nsteps) ? SPH_ERR_HIP : SPH_OK;
}

int sph_batch_sync(void *batch, int *live)
{
    Batch *b = (Batch *)batch;
    if (b->inflight && batch_sync(b)) return SPH_ERR_HIP;
    if (live) *live = (int)(b->live.size() + b->pending.size());
    return SPH_OK;
}

// Reports the batch's launch shapes; reps == 0 fixes the current ones and
// turns the online tuning off, reps < 0 only reports.
int sph_batch_autotune(void *batch, int reps, int32_t *shape)
{
    Batch *b = (Batch *)batch;
    if (!b) return SPH_ERROR;
    if (reps == 0) b->tuned = true;
    if (shape) {
        shape[0] = b->cell_slots;
        shape[1] = b->corder;
        shape[2] = b->tuned ? 1 : 0;
    }
    return SPH_OK;
}

// Status of one slot: its globals (timestep, ptime, done, error).
int sph_batch_status(void *batch, int slot, sph_globals *out)
{
    Batch *b = (Batch *)batch;
    if (slot < 0 || slot >= (int)b->slots.size() || !b->slots[slot]) return SPH_ERROR;
    BCHK(b, hipStreamSynchronize(b->st));
    return sph_get_globals(b->slots[slot], out);
}

// The scene context of a slot, for sph_download / sph_get_neighbours.
void *sph_batch_scene(void *batch, int slot)
{
    Batch *b = (Batch *)batch;
    if (slot < 0 || slot >= (int)b->slots.size()) return nullptr;
    hipStreamSynchronize(b->st);
    return b->slots[slot];
}

// Slots that are resident but no longer live (finished or failed), so the
// caller can retire them; returns how many were written.
int sph_batch_finished(void *batch, int *slots, int max)
{
    Batch *b = (Batch *)batch;
    int n = 0;
    for (int s = 0; s < (int)b->slots.size() && n < max; s++)
        if (b->slots[s] && !b->zombie[s] && std::find(b->live.begin(), b->live.end(), s) == b->live.end() &&
            std::find(b->pending.begin(), b->pending.end(), s) == b->pending.end())
            slots[n++] = s;
    return n;
}

int sph_batch_live(void *batch)
{
    const Batch *b = (const Batch *)batch;
    return (int)(b->live.size() + b->pending.size());
}

void sph_batch_destroy(void *batch)
{
    Batch *b = (Batch *)batch;
    if (!b) return;
    hipStreamSynchronize(b->st);
    batch_drop_graph(b);
    for (Context *c : b->slots)
        if (c) sph_destroy(c);
    hipFree(b->dP);
    hipFree(b->dA);
    hipFree(b->dA2);
    hipFree(b->dscenes);
    hipFree(b->dpmap);
    for (int2 *m : b->dvmap) hipFree(m);
    hipFree(b->dtmap);
    hipFree(b->dcmap);
    hipFree(b->dwmap);
    for (hipEvent_t e : b->ev)
        if (e) hipEventDestroy(e);
    hipFree(b->dctl);
    if (b->hctl) hipHostFree(b->hctl);
    hipStreamDestroy(b->st);
    delete b;
}

}  // extern "C"
