// sph_kernels.hpp -- launchers of the step-loop kernels (sph_kernels.hip).
#pragma once

#include "sph_state.hpp"

namespace sph {

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// A launch of xcd_grid(nb) workgroups read through xcd_block() hands every
// XCD one contiguous range of logical blocks: with cell-sorted storage the
// neighbour records a range gathers stay in that XCD's L2.
constexpr int kXcd = 8;
inline int xcd_grid(int nb) { return (nb + kXcd - 1) / kXcd * kXcd; }
#ifdef __HIPCC__
__device__ __forceinline__ int xcd_block(int nb)
{
    const int per = (nb + kXcd - 1) / kXcd;
    return (int)(blockIdx.x % kXcd) * per + (int)(blockIdx.x / kXcd);
}
#endif

#ifdef __HIPCC__
// A pointer into global memory (address space 1): loads and stores through it
// are global instructions (a 32-bit offset from a base in scalar registers),
// where a generic pointer read from a table of pointers (the batched kernels)
// would give flat ones.
template <class T>
using GPtr = __attribute__((address_space(1))) T *;
template <class T>
__device__ __forceinline__ GPtr<T> gptr(T *p)
{
    return (GPtr<T>)p;
}

// *p as a global load (T a vector type of 8, 16 or 32 bytes, naturally aligned)
template <class T>
__device__ __forceinline__ T gload(const T *p)
{
    typedef unsigned int W __attribute__((ext_vector_type(sizeof(T) / 4)));
    return __builtin_bit_cast(T, *reinterpret_cast<GPtr<const W>>(gptr(p)));
}

// the first of the two 16-byte halves of chunk row k0 / 8 of particle i's list
__device__ __forceinline__ const int4 *nbr_chunk(const int *nbr, int k0, int i, int ncap)
{
    return reinterpret_cast<const int4 *>(nbr + nbr_slot(k0, i, ncap));
}

// Calls f(j) for the nn entries of particle i in list order, reading the list
// one chunk (two 16-byte loads) at a time.
template <class F>
__device__ __forceinline__ void for_each_nbr(const int *nbr, int i, int nn, int ncap, F &&f)
{
    for (int k0 = 0; k0 < nn; k0 += kNbrChunk) {
        const int4 *p = nbr_chunk(nbr, k0, i, ncap);
        const int4 c0 = gload(p), c1 = gload(p + 1);
        const int e[kNbrChunk] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
        const int m = nn - k0;
#pragma unroll
        for (int u = 0; u < kNbrChunk; u++)
            if (u < m) f(e[u]);
    }
}

// As for_each_nbr, with the record loads of G consecutive entries issued
// before any is used: load(j) returns a neighbour's record, use(rec) consumes
// it; the entries past the list end up to the chunk boundary hold the
// particle's own index (padding written with the count), so they load a valid
// record that is not used.  The use order is the list order.  LOADS_FIRST
// pins the G loads ahead of every use with a scheduling barrier (the compiler
// otherwise interleaves them with the arithmetic to save registers, which
// leaves fewer gathers in flight).
template <int G, bool LOADS_FIRST = false, class Load, class Use>
__device__ __forceinline__ void for_each_nbr_batched(const int *nbr, int i, int nn, int ncap, Load &&load, Use &&use)
{
    static_assert(kNbrChunk % G == 0, "batch divides the chunk");
    for (int k0 = 0; k0 < nn; k0 += kNbrChunk) {
        const int4 *p = nbr_chunk(nbr, k0, i, ncap);
        const int4 c0 = gload(p), c1 = gload(p + 1);
        const int e[kNbrChunk] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
        const int m = nn - k0;
#pragma unroll
        for (int g0 = 0; g0 < kNbrChunk; g0 += G) {
            if (g0 >= m) break;
            decltype(load(0)) rec[G];
#pragma unroll
            for (int u = 0; u < G; u++) rec[u] = load(e[g0 + u]);   // entries past nn: padding (i)
            if (LOADS_FIRST) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int u = 0; u < G; u++)
                if (g0 + u < m) use(rec[u]);
        }
    }
}

// As for_each_nbr_batched, consuming the entries two at a time in list
// order: use2(rec_a, rec_b) for entries u and u + 1 whenever entry u is in
// the list.  Entry u + 1 may be padding (the particle's own index): the
// caller's pair term of a particle with itself must be exactly +0.
template <int G, bool LOADS_FIRST = false, class Load, class Use2>
__device__ __forceinline__ void for_each_nbr_pairs(const int *nbr, int i, int nn, int ncap, Load &&load, Use2 &&use2)
{
    static_assert(kNbrChunk % G == 0 && G % 2 == 0, "batch divides the chunk, entries go in twos");
    for (int k0 = 0; k0 < nn; k0 += kNbrChunk) {
        const int4 *p = nbr_chunk(nbr, k0, i, ncap);
        const int4 c0 = gload(p), c1 = gload(p + 1);
        const int e[kNbrChunk] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
        const int m = nn - k0;
#pragma unroll
        for (int g0 = 0; g0 < kNbrChunk; g0 += G) {
            if (g0 >= m) break;
            decltype(load(0)) rec[G];
#pragma unroll
            for (int u = 0; u < G; u++) rec[u] = load(e[g0 + u]);
            if (LOADS_FIRST) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int u = 0; u < G; u += 2)
                if (g0 + u < m) use2(rec[u], rec[u + 1]);
        }
    }
}


// As for_each_nbr_pairs, four entries at a time: use4(ra, rb, rc, rd) for
// entries u .. u + 3 whenever entry u is in the list (the entries past it up
// to the chunk boundary are padding, whose terms must be exactly +0).
template <int G, bool LOADS_FIRST = false, class Load, class Use4>
__device__ __forceinline__ void for_each_nbr_quads(const int *nbr, int i, int nn, int ncap, Load &&load, Use4 &&use4)
{
    static_assert(kNbrChunk % G == 0 && G % 4 == 0, "batch divides the chunk, entries go in fours");
    for (int k0 = 0; k0 < nn; k0 += kNbrChunk) {
        const int4 *p = nbr_chunk(nbr, k0, i, ncap);
        const int4 c0 = gload(p), c1 = gload(p + 1);
        const int e[kNbrChunk] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
        const int m = nn - k0;
#pragma unroll
        for (int g0 = 0; g0 < kNbrChunk; g0 += G) {
            if (g0 >= m) break;
            decltype(load(0)) rec[G];
#pragma unroll
            for (int u = 0; u < G; u++) rec[u] = load(e[g0 + u]);
            if (LOADS_FIRST) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int u = 0; u < G; u += 4)
                if (g0 + u < m) use4(rec[u], rec[u + 1], rec[u + 2], rec[u + 3]);
        }
    }
}

#endif

// masks of launch_strain
enum { M_STRAIN = 1, M_CONST = 2, M_CRIT = 4, M_INPLACE = 8 };
// masks of launch_integrate
enum { M_BC = 1, M_VEL = 2, M_MOVE = 4, M_RHO = 8 };

// grid = false: the grid of this binning was set up by the previous step's
// launch_finish (next_grid)
void launch_setup_linked_list(const Params &P, const Arrays &A, hipStream_t st, bool grid = true);
// move_a = false: the accelerations stay behind (a fused Morris step rewrites
// all of them before any read)
void launch_permute(const Params &P, const Arrays &S, const Arrays &T, hipStream_t st, bool move_a = true);
// aa (|a|^2 per particle) from the current accelerations
void launch_accel_sq(const Params &P, const Arrays &A, hipStream_t st);
void launch_ghost_setup(const Params &P, const Arrays &A, hipStream_t st);
// Launch shapes that change timing, never results (chosen per scene by the
// autotuner of sph_capi.cpp; defaults measured on the 3D dam break).
struct Tune {
    int cell_slots = 32;       // slots per wave of the cell-centric neighbour kernel
    int strain_batch = 8;      // neighbour records in flight per lane, continuity sweep
    int momentum_batch = 8;    // the same, momentum sweep
    int corder = 0;            // cell storage order: 0 Morton, 1 row-major
};

void launch_neighbours(const Params &P, const Arrays &A, hipStream_t st, int cell_slots = 32);
void launch_shift(const Params &P, const Arrays &A, hipStream_t st);
void launch_strain(const Params &P, const Arrays &A, int mask, hipStream_t st, int batch = 0);
void launch_normals(const Params &P, const Arrays &A, hipStream_t st);
void launch_advance(const Params &P, const Arrays &A, hipStream_t st);
void launch_crit_fold(const Params &P, const Arrays &A, hipStream_t st);
// rigid_kin < 0: the sweep alone, without the rigid-body reduction
void launch_momentum(const Params &P, const Arrays &A, int rigid_kin, hipStream_t st, int batch = 0);
// records in flight per lane of the FP64 sweeps when none is chosen
constexpr int kStrain64Batch = 4, kMomentum64Batch = 2;
// resets the critical-step reduction slots (results of sweeps run for timing only)
void launch_crit_reset(const Arrays &A, hipStream_t st);
void launch_rigid_kin(const Params &P, const Arrays &A, hipStream_t st);
void launch_rigid_reduce(const Params &P, const Arrays &A, int rigid_kin, hipStream_t st);
void launch_integrate(const Params &P, const Arrays &A, int mask, hipStream_t st);
// next_grid: also sets up the grid of the next step's binning from the new
// bounding box (a chained step then skips k_grid_setup)
void launch_finish(const Params &P, const Arrays &A, int bbox, int end, hipStream_t st, int next_grid = 0);
void launch_canon(const Params &P, const Arrays &A, long long *out, int maxnbr, hipStream_t st);
// the reference's linked-list arrays (sph_llgrid, llpointer of reals and
// images, image sources) from the last binning; any output may be null
void launch_linked_list(const Params &P, const Arrays &A, int *llgrid, int *real_next, int *ghost_next,
                        int *ghost_par, long long ncell, hipStream_t st);
// MKernelBSpline::Kernel / GradW statics over n pairs (device buffers)
void launch_kernel_eval(int ndim, const PairConst &pc, int n, const double *xi, const double *xj, double *w,
                        double *dwdx, double *dwdr, hipStream_t st);

// Device tables of one batched launch: per-scene parameter and array blocks,
// the scene list, and block maps (scene, local block) over particles and
// scan tiles.  All scenes share ndim and arithmetic arm.
// record variants of the FP32c sweeps: the momentum record (Params::mcompact)
// and whether the scene has one material
constexpr int kSweep32Variants = 5;
inline int sweep32_variant(const Params &P)
{
    const bool one = P.nmat == 1;
    return P.mcompact == 2 ? 0 : (P.mcompact == 1 ? (one ? 1 : 2) : (one ? 3 : 4));
}

struct BatchView {
    const Params *P;
    const Arrays *A;         // current buffer sets (binning reads these)
    const Arrays *A2;        // the other buffer sets (sorted into; the step continues on them)
    const int *scenes;       // scene indices into P/A, nscene entries
    const int2 *pmap;        // npart_blocks entries
    const int2 *tmap;        // ntile_blocks entries (scan tiles over cells)
    // neighbour search: the cell kernel's blocks (scene, block, scene's blocks)
    // and, for scenes whose lists exceed its 32-bit offsets, particle blocks
    // of the per-particle walk
    const int4 *cmap;
    const int2 *wmap;
    // particle blocks of the FP32c sweeps by record variant (sweep32_variant):
    // each variant has a kernel of its own, so that a kernel holds the
    // registers of its own variant, not the largest one's
    const int2 *vmap[kSweep32Variants];
    int nvblocks[kSweep32Variants];
    int nscene, npart_blocks, ntile_blocks, ncell_blocks, nwalk_blocks;
    int zero_nb, slab_nb;    // fixed blocks per scene for grid-stride kernels
    int ndim, arm;
    int any_periodic, any_st, any_rigid, pass_mask;
};
constexpr int kScanTileItems = kBlock * 8;

void launch_batch_step(const BatchView &B, hipStream_t st);
// slots per wave of the batched cell kernel
constexpr int kBatchCellSlots = 64;
// blocks of the cell kernel for one scene of n particles
int batch_cell_blocks(int n, int slots);
// whether a scene's lists fit the cell kernel's 32-bit offsets
bool cell_lists_fit(const Params &P);
void launch_gather_ctl(const Arrays *At, const int *sc, int ns, Ctl *out, hipStream_t st);
void launch_batch_strain32(const BatchView &B, hipStream_t st);
void launch_batch_momentum32(const BatchView &B, hipStream_t st);

// ---------------------------------------------------------------------------
// Per-particle integration: MSolution::BoundCond (:700-726), UpdateVelocity
// (:602-637) + the particle half of RigidBodyVelocity (:1578-1588),
// MoveParticles (:769-809) with the periodic wrap; lo/hi receive the moved
// position (the caller folds the bounding box).  dispbc 8 (extension) keeps
// its position.  Inline in both sources that call it (k_integrate and the
// FP32c momentum sweep's fused form) with contraction off, so that both
// compute the same bits.
// ---------------------------------------------------------------------------
enum { I_BC = 1, I_VEL = 2, I_MOVE = 4, I_RHO = 8 };

__device__ __forceinline__ void constrain(double *q, int code, int D)
{
    switch (code) {
    case 1: q[0] = 0.0; break;
    case 2: if (D > 1) q[1] = 0.0; break;
    case 3: if (D > 2) q[2] = 0.0; break;
    case 4: q[0] = 0.0; if (D > 1) q[1] = 0.0; break;
    case 5: if (D > 1) q[1] = 0.0; if (D > 2) q[2] = 0.0; break;
    case 6: if (D > 2) q[2] = 0.0; q[0] = 0.0; break;
    case 7: case 8: for (int d = 0; d < D; d++) q[d] = 0.0; break;
    }
}

// a: the particle's acceleration as the momentum sweep left it
template <int D>
__device__ __forceinline__ void integrate_particle(const Params &P, const Arrays &A, const Ctl *ctl, int i, uint32_t fi,
                                                   int mask, double (&a)[D], double (&lo)[D], double (&hi)[D])
{
#pragma clang fp contract(off)
    const bool act = fi & kActive, rig = fi & kRigid;
    const int disp = info_disp(fi);
    if ((mask & I_BC) && act) {
        if (P.baseaccel && !rig)
#pragma unroll
            for (int d = 0; d < D; d++) a[d] += P.base_a[d];
        constrain(a, disp, D);
#pragma unroll
        for (int d = 0; d < D; d++) A.a[d][i] = a[d];
    }
    if (mask & I_VEL) {
        // the next critical step's |a|^2 (CalcCriticalTimestep's sum, :41-44)
        double s = 0.0;
#pragma unroll
        for (int d = 0; d < D; d++) s += a[d] * a[d];
        A.aa[i] = s;
    }
    double x[D], v[D];
#pragma unroll
    for (int d = 0; d < D; d++) {
        x[d] = A.x[d][i];
        v[d] = A.v[d][i];
    }
    if ((mask & I_VEL) && act) {
        if (!rig) {
            const double dtn = 0.5 * (ctl->dt + ctl->dtold);
#pragma unroll
            for (int d = 0; d < D; d++) v[d] = v[d] + dtn * a[d];
        } else {
            double rl[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int d = 0; d < D; d++) rl[d] = x[d] - ctl->rbR[d];
            const double *V = ctl->rbV, *W = ctl->rbW;
            double vr[3];
            vr[0] = V[0] + W[1] * rl[2] - W[2] * rl[1];
            vr[1] = V[1] + W[2] * rl[0] - W[0] * rl[2];
            vr[2] = V[2] + W[0] * rl[1] - W[1] * rl[0];
#pragma unroll
            for (int d = 0; d < D; d++) v[d] = vr[d];
        }
#pragma unroll
        for (int d = 0; d < D; d++) A.v[d][i] = v[d];
    }
    if (mask & I_MOVE) {
        if (act && disp != 8) {
            const double dt = ctl->dt;
#pragma unroll
            for (int d = 0; d < D; d++) x[d] += v[d] * dt;
        }
#pragma unroll
        for (int d = 0; d < D; d++) {
            if (P.bcode[1][d] == 2 && x[d] >= P.bx[1][d]) {
                const double over = x[d] - P.bx[1][d];
                x[d] = P.bx[0][d] + over;
            }
            if (P.bcode[0][d] == 2 && x[d] <= P.bx[0][d]) {
                const double over = P.bx[0][d] - x[d];
                x[d] = P.bx[0][d] - over;
            }
            A.x[d][i] = x[d];
            lo[d] = x[d];
            hi[d] = x[d];
        }
    }
    if (mask & I_RHO) A.rho[i] = A.rho_new[i];
}

// FP32-compensated pair sweeps (sph_fp32c.hip); fused-step forms only
void launch_strain32(const Params &P, const Arrays &A, hipStream_t st, int batch = 8);
// fuse != 0: each particle is also integrated (k_integrate's full pass, no
// rigid body) and the moved bounding box folded; k_integrate is not launched
void launch_momentum32(const Params &P, const Arrays &A, hipStream_t st, int batch = 4, int fuse = 0);

}  // namespace sph
