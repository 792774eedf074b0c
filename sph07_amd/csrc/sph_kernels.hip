// sph_kernels.hip -- HIP kernels of the per-timestep particle loop, FP64 arm.
//
// Compiled with -ffp-contract=off: every FP64 expression keeps the operand
// order of the reference line it restates and no FMA is formed, so a
// particle's sums are bit-identical to the reference's sequential loops
// (the pair sweeps visit neighbours in the reference's list order).
//
// One thread per particle.  Kernels read the device control block and return
// at once when a previous kernel of the step flagged an error or the stop
// test fired, so a queued sequence of steps is a no-op after either.
#include "sph_cells.hpp"

#include <stdlib.h>

#include <math.h>

namespace sph {

#ifndef SPH_NBR_BOX_FILTER
#define SPH_NBR_BOX_FILTER 1
#endif
constexpr bool kNbrBoxFilter = SPH_NBR_BOX_FILTER;   // cell-box filter of the staged candidates
#ifndef SPH_NBR_TPOS_LDS
#define SPH_NBR_TPOS_LDS 1
#endif
constexpr bool kNbrTposLds = SPH_NBR_TPOS_LDS;       // targets' positions broadcast from LDS
#ifndef SPH_NBR_SMASK
#define SPH_NBR_SMASK 1
#endif
#ifndef SPH_NBR_FILT_BLK
#define SPH_NBR_FILT_BLK 4
#endif
constexpr int kFiltBlk = SPH_NBR_FILT_BLK;           // box-filter chunks with loads in flight together

#define CTL_STOP(ctl) ((ctl)->err != 0 || (ctl)->done != 0)

__device__ __forceinline__ void set_err(Ctl *ctl, int code, int who)
{
    if (atomicCAS(&ctl->err, 0, code) == 0) ctl->err_particle = who;
}

// ---------------------------------------------------------------------------
// wave / block primitives (64-lane wavefronts)
// ---------------------------------------------------------------------------
// Exclusive scan over the block; *total receives the block sum.
__device__ __forceinline__ int block_excl_scan(int v, int *total)
{
    __shared__ int wsum[kBlock / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int inc = wave_incl_scan(v);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    if (w == 0) {
        int s = lane < kBlock / 64 ? wsum[lane] : 0;
        s = wave_incl_scan(s);
        if (lane < kBlock / 64) wsum[lane] = s;
    }
    __syncthreads();
    int base = w ? wsum[w - 1] : 0;
    *total = wsum[kBlock / 64 - 1];
    __syncthreads();
    return base + inc - v;
}

__device__ __forceinline__ double wave_min(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// ---------------------------------------------------------------------------
// B-spline gradient, MKernelBSpline::GradW (MKernelBSpline.cpp:137-187)
// with the material-pair constants precomputed on the host.
// ---------------------------------------------------------------------------
template <int D>
__device__ __forceinline__ void gradw64(double dw[D], const double *xi, const double *xj,
                                        const PairConst &pc)
{
    double xij[D], rad = 0.0;
#pragma unroll
    for (int n = 0; n < D; n++) {
        xij[n] = xi[n] - xj[n];
        rad += xij[n] * xij[n];
    }
    rad = sqrt(rad);
    const double z = rad / pc.havg;
    double dW;
    if (z < 1.0)
        dW = pc.const1 * ((2.25 * z - 3.0) * z);
    else if (z < 2.0)
        dW = -3.0 * pc.const2 * ((2.0 - z) * (2.0 - z));
    else
        dW = 0.0;
    const double den = rad * pc.havg;
#pragma unroll
    for (int n = 0; n < D; n++) dw[n] = (rad == 0.0) ? 0.0 : dW * xij[n] / den;
}

// ---------------------------------------------------------------------------
// gather records (the same FP64 values as the per-field arrays)
// ---------------------------------------------------------------------------
template <int D>
__device__ __forceinline__ void put_srec(const Arrays &A, int j, const double *xs, double vol,
                                         const double *v, uint32_t info)
{
    double4 a, b;
    a.x = xs[0]; a.y = D > 1 ? xs[1] : 0.0; a.z = D > 2 ? xs[2] : 0.0; a.w = vol;
    b.x = v[0];  b.y = D > 1 ? v[1] : 0.0;  b.z = D > 2 ? v[2] : 0.0;  b.w = (double)info;
    A.rsA[j] = a;
    A.rsB[j] = b;
}

template <int D>
__device__ __forceinline__ void put_mrec(const Params &P, const Arrays &A, int j, const double *x, double m,
                                         const double *v, double rho, double prho2, uint32_t info)
{
    double4 a, b;
    a.x = x[0]; a.y = D > 1 ? x[1] : 0.0; a.z = D > 2 ? x[2] : 0.0; a.w = P.mcompact64 ? prho2 : m;
    b.x = v[0]; b.y = D > 1 ? v[1] : 0.0; b.z = D > 2 ? v[2] : 0.0; b.w = rho;
    A.rmA[j] = a;
    A.rmB[j] = b;
    if (!P.mcompact64) A.rmC[j] = make_double2(prho2, (double)info);
}

// symmetric tensors of the Classic discretisation: xx, yy, zz, xy, xz, yz
__device__ __forceinline__ int sym6(int r, int c)
{
    return r == c ? r : (r + c == 1 ? 3 : (r + c == 2 ? 4 : 5));
}

template <int D>
__device__ __forceinline__ void put_crec(const Arrays &A, int j, const double *T, uint32_t info)
{
    A.rmB[j] = make_double4(T[0], T[1], T[2], (double)info);
    A.rcC[j] = make_double4(T[3], T[4], T[5], 0.0);
}

template <int D>
__device__ __forceinline__ void vec_of(const double4 &r, double *out)
{
    out[0] = r.x;
    if (D > 1) out[1] = r.y;
    if (D > 2) out[2] = r.z;
}

// ---------------------------------------------------------------------------
// grid: MNeighbourVars::InitGrid (MNeighbourVars.cpp:18-66)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void grid_setup_body(const Params &P, Ctl *ctl)
{
    if (CTL_STOP(ctl)) return;
    long long ncell = 1;
    for (int i = 0; i < 3; i++) {
        ctl->gmin[i] = 0.0;
        ctl->glim[i] = 1;
        ctl->gsize[i] = 1.0 / (2.0 * P.hmax);
    }
    for (int i = 0; i < P.ndim; i++) {
        if (P.btype[i] == 4) {
            int gl = (int)(ctl->gsize[i] * (P.bx[1][i] - P.bx[0][i]));
            gl = gl > 0 ? gl : 0;
            ctl->gsize[i] = gl / (P.bx[1][i] - P.bx[0][i]);
            ctl->gmin[i] = P.bx[0][i] - 1.0 / ctl->gsize[i];
            ctl->glim[i] = gl + 2;
        } else {
            ctl->gmin[i] = ctl->cmm[0][i] - 0.01 * P.hmax;
            ctl->glim[i] = (int)(ctl->gsize[i] * (ctl->cmm[1][i] - ctl->gmin[i])) + 1;
        }
    }
    for (int i = 0; i < 3; i++) ncell *= (ctl->glim[i] > 0 ? ctl->glim[i] : 0);
    // Storage order of the cells, as the bits of a cell code that each axis
    // owns (each axis contributing ceil(log2(glim)) bits).  Morton: the bits
    // interleaved, the innermost traversal axis in the lowest bit of every
    // level, so that a contiguous range of sorted particles is a compact
    // region of space.  Row-major: z in the lowest bits, then y, then x.  The
    // traversal order of the neighbour search is unaffected (it walks cells
    // by coordinates), so results do not depend on the choice.
    int mb[3] = {0, 0, 0};
    for (int d = 0; d < 3; d++)
        while (mb[d] < 31 && (1LL << mb[d]) < (long long)ctl->glim[d]) mb[d]++;
    unsigned int msk[3] = {0u, 0u, 0u};
    int pos = 0;
    if (P.corder == 1) {
        for (int e = 2; e >= 0; e--)
            for (int l = 0; l < mb[e] && pos <= 30; l++) msk[e] |= 1u << pos++;
    } else {
        for (int l = 0; l < 31 && pos <= 30; l++)
            for (int e = 2; e >= 0; e--)
                if (l < mb[e] && pos <= 30) msk[e] |= 1u << pos++;
    }
    const long long ncode = pos <= 29 ? (1LL << pos) : 0x7fffffffLL;
    const long long need = ncode > ncell ? ncode : ncell;
    if (ncell <= 0 || pos > 29 || need + 2 > P.ncell_cap) {
        ctl->ncell = (int)(need > 0x7ffffff0LL ? 0x7ffffff0 : need);
        set_err(ctl, kErrGridCap, -1);
        return;
    }
    ctl->ncell = (int)ncell;
    ctl->ncode = (int)ncode;
    ctl->novf = 0;

    ctl->nbucket = (int)ncode + 1;
    for (int d = 0; d < 3; d++) ctl->cmask[d] = msk[d];
}

__device__ __forceinline__ int cell_code(const Ctl *ctl, int i, int j, int k)
{
    return code_dep(i, ctl->cmask[0]) | code_dep(j, ctl->cmask[1]) | code_dep(k, ctl->cmask[2]);
}

// bucket counts (cells + the inactive bucket) and image counts cleared
__device__ __forceinline__ void zero_cells_body(const Params &P, const Arrays &A, const int bid, const int nb)
{
    const Ctl *ctl = A.ctl;
    if (CTL_STOP(ctl)) return;
    const int n = ctl->ncell, nc = ctl->ncode;
    const int top = (nc > n ? nc : n) + 1;
    for (int c = bid * kBlock + threadIdx.x; c <= top; c += nb * kBlock) {
        A.cell_count[c] = 0;
        if (P.periodic && c < n) A.gcount[c] = 0;
    }
}

// MSolution::SetupLinkedList (MSolution.cpp:116-148): cell of each computable
// particle (MNeighbourVars::GetBoxCoords :83-89, truncation, no clamp) and an
// arbitrary slot inside the cell; the order inside a cell is fixed by k_permute.
template <int D>
__device__ __forceinline__ void bin_count_body(const Params &P, const Arrays &A, const int bid)
{
    Ctl *ctl = A.ctl;
    if (CTL_STOP(ctl)) return;
    const int i = bid * kBlock + threadIdx.x;
    const int lane = threadIdx.x & 63;
    // bucket of this particle (-1: none), every lane of the wave takes part
    int key = -1;
    if (i < P.N) {
        const uint32_t f = A.info[i];
        const int inactive_bucket = ctl->ncode;
        key = inactive_bucket;
        if (f & kActive) {
            int bc[3] = {0, 0, 0};
#pragma unroll
            for (int d = 0; d < D; d++) bc[d] = (int)(ctl->gsize[d] * (A.x[d][i] - ctl->gmin[d]));
            if (ctl->timestep == 1 && (bc[0] < 0 || bc[0] > ctl->glim[0] - 1)) {
                A.info[i] = f & ~kActive;          // outside the (single) process slab
            } else {
                bool out = false;
#pragma unroll
                for (int d = 0; d < 3; d++) out |= bc[d] < 0 || bc[d] >= ctl->glim[d];
                if (out) {
                    set_err(ctl, kErrOutOfGrid, A.orig[i]);
                    key = -1;
                } else {
                    key = cell_code(ctl, bc[0], bc[1], bc[2]);
                }
            }
        }
        A.cell_of[i] = key;
    }
    // The storage is sorted by last step's cells, so equal keys come in runs
    // of neighbouring lanes: one atomic per run, slots handed out in the run.
    const int prev = __shfl(key, lane > 0 ? lane - 1 : 0, 64);
    const bool head = lane == 0 || prev != key;
    const unsigned long long hm = __ballot(head);
    const unsigned long long upto = lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1ull);
    const int start = 63 - __clzll((long long)(hm & upto));
    const unsigned long long after = hm & ~upto;
    const int end = after ? __ffsll((long long)after) - 1 : 64;
    int base = 0;
    if (head && key >= 0) base = atomicAdd(&A.cell_count[key], end - start);
    base = __shfl(base, start, 64);
    if (i < P.N && key >= 0) A.slot[i] = base + (lane - start);
}

__device__ __forceinline__ void bin_scatter_body(const Params &P, const Arrays &A, const int bid)
{
    if (CTL_STOP(A.ctl)) return;
    const int i = bid * kBlock + threadIdx.x;
    if (i >= P.N) return;
    const int c = A.cell_of[i];
    if (c < 0) return;
    // the cell's members' deck numbers, the keys of their order
    A.tmp[A.cell_start[c] + A.slot[i]] = A.orig[i];
}


// Moves every persistent per-particle field to its sorted slot (src -> dst
// buffer set).  After an error or the stop test the move is the identity, so
// the two buffer sets still alternate once per queued step.
template <int D>
__device__ __forceinline__ void permute_body(const Params &P, const Arrays &S, const Arrays &T, const int bid,
                                             bool move_a = true)
{
    const int i = bid * kBlock + threadIdx.x;
    if (i >= P.N) return;
    const bool moving = !CTL_STOP(S.ctl);
    // Every source field is read before any store: the destination arrays are
    // other allocations, but the compiler cannot know it, and with loads and
    // stores interleaved each load waited for the store before it.
    const int c = moving ? S.cell_of[i] : -1;  // < 0 only with the out-of-grid error set
    const int oi = S.orig[i];
    const uint32_t fi = S.info[i];
    // a fused step rewrites every a in its momentum sweep before reading it
    // (the critical step reads aa): a moves only for the other paths, and
    // when the loop has stopped (nothing rewrites it then)
    const bool mva = move_a || !moving;
    double xv[D], vv[D], av[D];
#pragma unroll
    for (int d = 0; d < D; d++) {
        xv[d] = S.x[d][i];
        vv[d] = S.v[d][i];
        av[d] = mva ? S.a[d][i] : 0.0;
    }
    const double aa = S.aa[i], rho = S.rho[i], m = S.m[i];
    const double q = P.disctype == 0 ? S.q[i] : 0.0;
    const int rr = P.rigid ? S.rrank[i] : -1;
    const double dt = S.ctl->dt;
    // Order inside a cell: descending deck number, the order head insertion
    // leaves in the reference's linked lists (MSolution.cpp:143-144).  The
    // destination slot is the cell's first slot plus the number of its
    // members with a higher deck number (k_bin_scatter left their numbers in
    // tmp; with an error the storage stays as it is).
    int j = i;
    if (c >= 0) {
        const int s0 = S.cell_start[c], e0 = S.cell_start[c + 1];
        int r = 0;
        for (int t = s0; t < e0; t++) r += S.tmp[t] > oi;
        j = s0 + r;
    }
#pragma unroll
    for (int d = 0; d < D; d++) {
        T.x[d][j] = xv[d];
        T.v[d][j] = vv[d];
        if (mva) T.a[d][j] = av[d];
    }
    T.aa[j] = aa;
    T.rho[j] = rho;
    T.m[j] = m;
    T.info[j] = fi;
    T.orig[j] = oi;
    if (P.disctype == 0) T.q[j] = q;
    if (P.rigid) {
        T.rrank[j] = rr;
        if (rr >= 0) T.rigid_idx[rr] = j;
    }
    if (!moving) return;
    // cell code of every sorted slot (the inactive bucket is ncode)
    T.occ[j] = c;
    if (!(fi & kActive)) T.nnbr[j] = 0;
    {
        // MSolution::Strain's half step back (:334-338) and the continuity
        // record, written with the move (the neighbour search reads x only)
        const bool act = fi & kActive;
        double xs[D];
#pragma unroll
        for (int d = 0; d < D; d++) {
            xs[d] = act ? xv[d] - 0.5 * vv[d] * dt : xv[d];
            // the FP32c continuity sweep derives xs from x itself
            if (P.arm != 1) T.xs[d][j] = xs[d];
        }
        const double vol = m / rho;
        if (P.arm == 1) {
            put_srec32<D>(P, T, j, xs, vol, vv, fi);
        } else {
            put_srec<D>(T, j, xs, vol, vv, fi);
        }
    }
}

// ---------------------------------------------------------------------------
// device-wide exclusive scan of n = *n_ptr ints (wave scans + LDS)
// ---------------------------------------------------------------------------
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

__device__ __forceinline__ void scan_reduce_body(const int *in, const int *n_ptr, int *aux, const Ctl *ctl, const int bid)
{
    if (CTL_STOP(ctl)) return;
    const int n = *n_ptr;
    const int base = bid * kScanTile + threadIdx.x * kScanItems;
    int s = 0;
#pragma unroll
    for (int q = 0; q < kScanItems; q++) s += (base + q < n) ? in[base + q] : 0;
    int total;
    block_excl_scan(s, &total);
    if (threadIdx.x == 0) aux[bid] = total;
}

// single block: exclusive scan of the tile sums, total written to out[n]
__device__ __forceinline__ void scan_aux_body(int *aux, int nblocks, const int *n_ptr, int *out, const Ctl *ctl)
{
    if (CTL_STOP(ctl)) return;
    int carry = 0;
    for (int b0 = 0; b0 < nblocks; b0 += kBlock) {
        const int b = b0 + threadIdx.x;
        int v = b < nblocks ? aux[b] : 0;
        int total;
        int ex = block_excl_scan(v, &total);
        if (b < nblocks) aux[b] = carry + ex;
        carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) out[*n_ptr] = carry;
}

__device__ __forceinline__ void scan_down_body(const int *in, int *out, const int *n_ptr, const int *aux, const Ctl *ctl, const int bid)
{
    if (CTL_STOP(ctl)) return;
    const int n = *n_ptr;
    const int base = bid * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems], s = 0;
#pragma unroll
    for (int q = 0; q < kScanItems; q++) {
        v[q] = (base + q < n) ? in[base + q] : 0;
        s += v[q];
    }
    int total;
    int ex = block_excl_scan(s, &total) + aux[bid];
#pragma unroll
    for (int q = 0; q < kScanItems; q++) {
        if (base + q < n) out[base + q] = ex;
        ex += v[q];
    }
}

// The down pass of a scene's scan without the single-block pass over the tile
// sums: each block adds the totals of the tiles before it itself (a few
// hundred integers, cached) and the last block writes the total to out[n].
// It also clears the counts it has read, so that the next binning starts
// from zero without a clearing launch (the counts are zeroed at allocation
// and by the host after any device error, which can leave a binning unscanned).
__device__ __forceinline__ void scan_down_clear_body(int *in, int *out, const int *n_ptr, const int *aux, int nblocks,
                                                     const Ctl *ctl, const int bid)
{
    if (CTL_STOP(ctl)) return;
    const int n = *n_ptr;
    int pre = 0;
    for (int b = threadIdx.x; b < bid; b += kBlock) pre += aux[b];
    int prefix;
    block_excl_scan(pre, &prefix);
    const int base = bid * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems], s = 0;
#pragma unroll
    for (int q = 0; q < kScanItems; q++) {
        v[q] = (base + q < n) ? in[base + q] : 0;
        s += v[q];
    }
    int total;
    int ex = block_excl_scan(s, &total) + prefix;
#pragma unroll
    for (int q = 0; q < kScanItems; q++) {
        if (base + q < n) {
            out[base + q] = ex;
            in[base + q] = 0;
        }
        ex += v[q];
    }
    if (bid == nblocks - 1 && threadIdx.x == 0) out[n] = prefix + total;
}

// ---------------------------------------------------------------------------
// periodic images: MSolution::GhostSetup / GenerateGhostParticles
// (MSolution.cpp:811-918), MGhostParticleData::CreateGhostFromReal
// (MGhostParticleData.cpp:12-36).  One pass per periodic side, in the
// reference's pass order; the source slab is enumerated z-major, x fastest.
// ---------------------------------------------------------------------------
struct PassGeom {
    int lo[3], hi[3], off[3], n;
};

__device__ __forceinline__ PassGeom pass_geom(const Ctl *ctl, int m, int side, const double *add)
{
    PassGeom g;
    for (int d = 0; d < 3; d++) {
        g.lo[d] = 0;
        g.hi[d] = ctl->glim[d] - 1;
        g.off[d] = 0;
        if (add[d] != 0.0) g.off[d] = (add[d] >= 0 ? 1 : -1) * (ctl->glim[d] - 2);
    }
    g.lo[m] = g.hi[m] = side == 0 ? 1 : ctl->glim[m] - 2;
    g.n = (g.hi[0] - g.lo[0] + 1) * (g.hi[1] - g.lo[1] + 1) * (g.hi[2] - g.lo[2] + 1);
    return g;
}

__device__ __forceinline__ void slab_coords(const PassGeom &g, int s, int &i, int &j, int &k)
{
    const int nx = g.hi[0] - g.lo[0] + 1, ny = g.hi[1] - g.lo[1] + 1;
    i = g.lo[0] + s % nx;
    j = g.lo[1] + (s / nx) % ny;
    k = g.lo[2] + s / (nx * ny);
}

// A source cell contributes only when its list head is a real particle:
// it holds particles and no image yet (`if (id<0) continue`, :906-909).
__device__ __forceinline__ void reset_ngp_body(Ctl *ctl)
{
    if (CTL_STOP(ctl)) return;
    ctl->ngp = 0;   // MGhostParticleData::DeleteContents (MGhostParticleData.cpp:48-53)
}

__device__ __forceinline__ void ghost_count_body(const Params &P, const Arrays &A, int m, int side, const int bid, const int nb)
{
    const Ctl *ctl = A.ctl;
    if (CTL_STOP(ctl)) return;
    const double *add = side == 0 ? &P.minadd[0][m] : &P.maxadd[0][m];
    double addv[3] = {add[0], add[3], add[6]};
    PassGeom g = pass_geom(ctl, m, side, addv);
    for (int s = bid * kBlock + threadIdx.x; s < g.n; s += nb * kBlock) {
        int i, j, k;
        slab_coords(g, s, i, j, k);
        const int c = ctl->glim[2] * (ctl->glim[1] * i + j) + k;
        const int q = cell_code(ctl, i, j, k);
        const int cnt = A.cell_start[q + 1] - A.cell_start[q];
        A.slab_cnt[s] = (cnt > 0 && A.gcount[c] == 0) ? cnt : 0;
    }
}

__device__ __forceinline__ void ghost_scan_body(const Params &P, const Arrays &A, int m, int side)
{
    Ctl *ctl = A.ctl;
    if (CTL_STOP(ctl)) return;
    const double *add = side == 0 ? &P.minadd[0][m] : &P.maxadd[0][m];
    double addv[3] = {add[0], add[3], add[6]};
    PassGeom g = pass_geom(ctl, m, side, addv);
    int carry = ctl->ngp;
    for (int s0 = 0; s0 < g.n; s0 += kBlock) {
        const int s = s0 + threadIdx.x;
        int v = s < g.n ? A.slab_cnt[s] : 0;
        int total;
        int ex = block_excl_scan(v, &total);
        if (s < g.n) A.slab_off[s] = carry + ex;
        carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (carry > P.gcap) set_err(ctl, kErrGhostCap, carry);
        else ctl->ngp = carry;
    }
}

__device__ __forceinline__ void ghost_emit_body(const Params &P, const Arrays &A, int m, int side, const int bid, const int nb)
{
    Ctl *ctl = A.ctl;
    if (CTL_STOP(ctl)) return;
    const double *add = side == 0 ? &P.minadd[0][m] : &P.maxadd[0][m];
    double addv[3] = {add[0], add[3], add[6]};
    PassGeom g = pass_geom(ctl, m, side, addv);
    int shift[3];
    for (int d = 0; d < 3; d++) shift[d] = addv[d] != 0.0 ? (addv[d] >= 0 ? 1 : -1) : 0;
    const int shift_code = (shift[0] + 1) * 9 + (shift[1] + 1) * 3 + (shift[2] + 1);
    // one wave per source cell, one lane per image (the cell's images are
    // consecutive sources and consecutive image slots)
    const int lane = threadIdx.x & 63;
    for (int s = bid * (kBlock / 64) + (int)(threadIdx.x >> 6); s < g.n; s += nb * (kBlock / 64)) {
        const int cnt = A.slab_cnt[s];
        if (cnt == 0) continue;
        int i, j, k;
        slab_coords(g, s, i, j, k);
        const int c = ctl->glim[2] * (ctl->glim[1] * i + j) + k;
        const int tc = ctl->glim[2] * (ctl->glim[1] * (i + g.off[0]) + (j + g.off[1])) + (k + g.off[2]);
        if (A.gcount[tc] != 0) {
            // a cell fed by two passes would need a merge of image runs
            if (lane == 0) set_err(ctl, 1, -2);
            continue;
        }
        const int g0 = A.slab_off[s];
        const int s0 = A.cell_start[cell_code(ctl, i, j, k)];
        for (int r = lane; r < cnt; r += 64) {
            const int src = s0 + r;              // descending deck number
            const int gi = P.N + g0 + r;
            // the source's fields are read before any store (the compiler
            // cannot tell the image slots from the source slots)
            double gx[3] = {0.0, 0.0, 0.0}, gv[3] = {0.0, 0.0, 0.0};
            for (int d = 0; d < P.ndim; d++) {
                gx[d] = A.x[d][src] + addv[d];
                gv[d] = A.v[d][src];
            }
            const double rho = A.rho[src], mass = A.m[src];
            const uint32_t ginfo = (A.info[src] & (kMatMask | kRigid | (0xFFFFu << kColorShift))) | kGhost;
            for (int d = 0; d < P.ndim; d++) {
                A.x[d][gi] = gx[d];
                A.xs[d][gi] = gx[d];
                A.v[d][gi] = gv[d];
            }
            A.m[gi] = mass;
            A.rho[gi] = rho;
            A.rho_new[gi] = rho;
            A.p[gi] = 0.0;        // images keep p = 0 (not copied, :28)
            A.prho2[gi] = 0.0;
            A.info[gi] = ginfo;
            switch (P.ndim) {
            case 1: put_srec<1>(A, gi, gx, mass / rho, gv, ginfo); put_mrec<1>(P, A, gi, gx, mass, gv, rho, 0.0, ginfo); break;
            case 2: put_srec<2>(A, gi, gx, mass / rho, gv, ginfo); put_mrec<2>(P, A, gi, gx, mass, gv, rho, 0.0, ginfo); break;
            default: put_srec<3>(A, gi, gx, mass / rho, gv, ginfo); put_mrec<3>(P, A, gi, gx, mass, gv, rho, 0.0, ginfo);
            }
            if (P.disctype == 0) {
                // images carry sigma = q = 0 (not copied, MGhostParticleData.cpp:16-28)
                const double T0[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
                put_crec<3>(A, gi, T0, ginfo);
            }
            if (P.arm == 1) {
                switch (P.ndim) {
                case 1: put_srec32<1>(P, A, gi, gx, mass / rho, gv, ginfo); put_mrec32<1>(P, A, gi, gx, mass, gv, rho, 0.0, ginfo); break;
                case 2: put_srec32<2>(P, A, gi, gx, mass / rho, gv, ginfo); put_mrec32<2>(P, A, gi, gx, mass, gv, rho, 0.0, ginfo); break;
                default: put_srec32<3>(P, A, gi, gx, mass / rho, gv, ginfo); put_mrec32<3>(P, A, gi, gx, mass, gv, rho, 0.0, ginfo);
                }
            }
            A.gsrc[g0 + r] = src;
            A.gshift[g0 + r] = shift_code;
        }
        // after every lane has read the old count
        if (lane == 0) {
            A.gfirst[tc] = g0;
            A.gcount[tc] = cnt;
        }
    }
}

// ---------------------------------------------------------------------------
// MSolution::LinkedListNeighbours (MSolution.cpp:150-215): cells r(x) ->
// s(y) -> t(z) ascending, in each cell images (newest first) then real
// particles (descending number); strict d^2 < (2*AVG(h_i,h_j))^2.
// The half step back of MSolution::Strain (:334-338) and the continuity
// records are written by k_permute, with the move into cell order.
// ---------------------------------------------------------------------------
template <int D>
__device__ __forceinline__ void nbr_epilogue(const Params &P, const Arrays &A, int i, int nn, double mind)
{
    A.nnbr[i] = nn;
    // pad the last chunk with the particle's own index: the batched sweeps
    // load whole chunks without a per-entry bound
    for (int k = nn; k & (kNbrChunk - 1); k++) A.nbr[nbr_slot(k, i, P.nstride)] = i;
    if (P.want_mindist) A.mindist[i] = mind;
}

// One particle's walk (the batched path, the fallback of the cell kernel for
// oversized stencils, and a launch shape of its own).  Every lane steps
// through the 3^D stencil cells in the same order, x -> y -> z ascending
// (z fastest), so that a wave's lanes, whose particles lie in different
// cells, run one loop body: per cell its periodic images newest first, then
// its slots (descending deck number), the order of the reference's linked
// lists.  A lane's loop over a cell's candidates has its own trip count.  The
// next cell's range is loaded while the current cell's candidates are tested.
// A stencil cell whose box lies beyond the largest cut-off (with a relative
// margin far above FP64 rounding) holds no neighbour and is not walked.
// candidates of a stencil cell whose positions the walk loads together
#ifndef SPH_WALK_BATCH
#define SPH_WALK_BATCH 2
#endif
constexpr int kWalkBatch = SPH_WALK_BATCH;

struct WalkCell {
    int s0, len, g0, ng;
};

// the grid and cell tables the walk reads, loaded once per particle
struct WalkGrid {
    int glim[3];
    unsigned int cmask[3];
    int periodic;
    GPtr<const int> cstart, gcount, gfirst;
};

template <int D>
__device__ __forceinline__ WalkCell walk_cell(const WalkGrid &g, int k, const int (&bc)[3], unsigned int q,
                                              const double (&dl)[3], const double (&dh)[3], double skip2)
{
    constexpr int W1 = D == 3 ? 3 : 1, W0 = D == 3 ? 9 : (D == 2 ? 3 : 1);
    const int off[3] = {k / W0 - 1, D > 1 ? (k / W1) % 3 - 1 : 0, D > 2 ? k % 3 - 1 : 0};
    WalkCell w{0, 0, 0, 0};
    bool in = true;
    double box2 = 0.0;
#pragma unroll
    for (int d = 0; d < D; d++) {
        const int c = bc[d] + off[d];
        in = in && c >= 0 && c < g.glim[d];
        box2 += off[d] < 0 ? dl[d] : (off[d] > 0 ? dh[d] : 0.0);
    }
    if (!in || box2 > skip2) return w;
    const unsigned int qc = code_step(q, g.cmask[0], off[0]) | code_step(q, g.cmask[1], off[1]) |
                            code_step(q, g.cmask[2], off[2]);
    w.s0 = g.cstart[qc];
    w.len = g.cstart[qc + 1] - w.s0;
    if (g.periodic) {
        const int c = g.glim[2] * (g.glim[1] * (bc[0] + off[0]) + (bc[1] + off[1])) + (bc[2] + off[2]);
        w.ng = g.gcount[c];
        if (w.ng) w.g0 = g.gfirst[c];
    }
    return w;
}

// FIT: every list and position byte offset fits 32 bits (lists_fit32), so
// the loads and stores take a 32-bit offset from a base in scalar registers
template <int D, bool ONEMAT, bool MIND, bool FIT>
__device__ __forceinline__ void neighbours_one(const Params &P, const Arrays &A, const int i)
{
    Ctl *ctl = A.ctl;
    const uint32_t fi = A.info[i];
    if (!(fi & kActive)) { A.nnbr[i] = 0; return; }
    double xi[D];
#pragma unroll
    for (int d = 0; d < D; d++) xi[d] = A.x[d][i];
    int bc[3] = {0, 0, 0};
    // squared distance from xi to the low and high faces of its own cell
    double dl[3] = {0.0, 0.0, 0.0}, dh[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int d = 0; d < D; d++) {
        bc[d] = (int)(ctl->gsize[d] * (xi[d] - ctl->gmin[d]));
        const double f0 = xi[d] - (ctl->gmin[d] + bc[d] / ctl->gsize[d]);
        const double f1 = (ctl->gmin[d] + (bc[d] + 1) / ctl->gsize[d]) - xi[d];
        dl[d] = f0 > 0.0 ? f0 * f0 : 0.0;
        dh[d] = f1 > 0.0 ? f1 * f1 : 0.0;
    }
    WalkGrid g;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        g.glim[d] = ctl->glim[d];
        g.cmask[d] = ctl->cmask[d];
    }
    g.periodic = P.periodic;
    g.cstart = gptr<const int>(A.cell_start);
    g.gcount = gptr<const int>(A.gcount);
    g.gfirst = gptr<const int>(A.gfirst);
    const unsigned int q = (unsigned int)(code_dep(bc[0], g.cmask[0]) | code_dep(bc[1], g.cmask[1]) |
                                          code_dep(bc[2], g.cmask[2]));
    const int mat_i = info_mat(fi);
    const double cut2_0 = P.pc[mat_i][mat_i].cut2;
    const double skip2 = P.cut2max * (1.0 + 1.0e-9);
    const int nreal = P.N;
    // list entry k of slot i: entry i * 8 + (k / 8) * nstride * 8 + k % 8;
    // entries past the capacity (an overflow, reported below) all land in
    // the spare chunk row past it
    const unsigned int maxr = (unsigned int)nbr_round(P.maxnbr), cshift = (unsigned int)P.nshift - 2u;
    const unsigned int tbase = (unsigned int)i * kNbrChunk;
    const size_t rstride = (size_t)P.nstride * kNbrChunk;
    const GPtr<char> lst = gptr(reinterpret_cast<char *>(A.nbr));
    const GPtr<const char> xb[3] = {gptr(reinterpret_cast<const char *>(A.x[0])),
                                    gptr(reinterpret_cast<const char *>(A.x[D > 1 ? 1 : 0])),
                                    gptr(reinterpret_cast<const char *>(A.x[D > 2 ? 2 : 0]))};
    auto pos = [&](int j, int d) -> double {
        if (FIT) return *reinterpret_cast<GPtr<const double>>(xb[d] + ((unsigned int)j << 3));
        return reinterpret_cast<GPtr<const double>>(xb[d])[j];
    };
    double mind = 1.0e20;
    unsigned int nn = 0;
    // valid: the candidate is in the cell (the test runs either way, so that
    // the loaded positions are consumed in every pass and no wait for the
    // list stores lands at the top of the next pass)
    auto test = [&](int cand, const double (&xc)[D], bool valid) {
        double d2 = 0.0;
#pragma unroll
        for (int d = 0; d < D; d++) {
            const double dd = xi[d] - xc[d];
            d2 += dd * dd;
        }
        const double cut2 = ONEMAT ? cut2_0 : P.pc[mat_i][info_mat(A.info[cand])].cut2;
        if (d2 < cut2 && cand != i && valid) {
            if (MIND) {
                const double r1 = sqrt(d2);
                mind = mind < r1 ? mind : r1;
            }
            const unsigned int ps = nn < maxr ? nn : maxr;
            if (FIT)
                *reinterpret_cast<GPtr<int>>(lst + ((tbase + ((ps >> 3) << cshift) + (ps & 7u)) << 2)) = cand;
            else
                reinterpret_cast<GPtr<int>>(lst)[(size_t)(ps >> 3) * rstride + tbase + (ps & 7u)] = cand;
            nn++;
        }
    };
    constexpr int NC = D == 3 ? 27 : (D == 2 ? 9 : 3);
    WalkCell w = walk_cell<D>(g, 0, bc, q, dl, dh, skip2);
    for (int k = 0; k < NC; k++) {
        const WalkCell cur = w;
        if (k + 1 < NC) w = walk_cell<D>(g, k + 1, bc, q, dl, dh, skip2);
        // the cell's images, newest first (one at a time: only the cells
        // next to a periodic boundary have any)
        for (int u = 0; u < cur.ng; u++) {
            const int cand = nreal + cur.g0 + cur.ng - 1 - u;
            double xc[D];
#pragma unroll
            for (int d = 0; d < D; d++) xc[d] = pos(cand, d);
            test(cand, xc, true);
        }
        // its slots, kWalkBatch per pass: their positions are loaded before
        // any is tested (past the cell's last slot a pass reloads that one),
        // then they are tested and appended in order
        for (int u0 = 0; u0 < cur.len; u0 += kWalkBatch) {
            int cd[kWalkBatch];
            double xc[kWalkBatch][D];
#pragma unroll
            for (int b = 0; b < kWalkBatch; b++) {
                cd[b] = cur.s0 + (u0 + b < cur.len ? u0 + b : cur.len - 1);
#pragma unroll
                for (int d = 0; d < D; d++) xc[b][d] = pos(cd[b], d);
            }
#pragma unroll
            for (int b = 0; b < kWalkBatch; b++) test(cd[b], xc[b], u0 + b < cur.len);
        }
    }
    if (nn > (unsigned int)P.maxnbr) {
        set_err(ctl, kErrNbrOverflow, A.orig[i]);
        A.nnbr[i] = P.maxnbr;
        return;
    }
    nbr_epilogue<D>(P, A, i, (int)nn, mind);
}

// the walk specialised for one material (one cut-off), for the
// minimum-distance output and for 32-bit offsets (wave-uniform choices)
template <int D, bool FIT>
__device__ __forceinline__ void neighbours_walk_fit(const Params &P, const Arrays &A, const int i)
{
    if (P.nmat == 1) {
        if (P.want_mindist) neighbours_one<D, true, true, FIT>(P, A, i);
        else neighbours_one<D, true, false, FIT>(P, A, i);
    } else {
        if (P.want_mindist) neighbours_one<D, false, true, FIT>(P, A, i);
        else neighbours_one<D, false, false, FIT>(P, A, i);
    }
}
template <int D>
__device__ __forceinline__ void neighbours_walk(const Params &P, const Arrays &A, const int i)
{
    if (lists_fit32(P)) neighbours_walk_fit<D, true>(P, A, i);
    else neighbours_walk_fit<D, false>(P, A, i);
}

template <int D>
__device__ __forceinline__ void neighbours_body(const Params &P, const Arrays &A, const int bid)
{
    if (CTL_STOP(A.ctl)) return;
    const int i = bid * kBlock + threadIdx.x;
    if (i >= P.N) return;
    neighbours_walk<D>(P, A, i);
}

// ---------------------------------------------------------------------------
// Cell-centric neighbour search (single-scene path).  One wave per occupied
// cell: the stencil's candidates, in the traversal order of the walk above
// (cells x -> y -> z, images newest first, then the cell's slots), are staged
// as slot indices in LDS.  Up to kNbrBlk chunks of 64 candidates are then
// held in registers (lane = candidate, coalesced position loads) and every
// target of the cell is broadcast in turn against all of them: a ballot of
// the pairs inside the cut-off and a prefix count append them to the target's
// list in candidate order, i.e. in the reference's order.  The target's count
// stays wave-uniform across its chunks.
// ---------------------------------------------------------------------------
// waves per block of the cell kernel: the waves of a block hold its LDS until
// the slowest of them is done, and waves of one cell kernel vary with their
// cells' work, so blocks of one wave free each wave's LDS by itself
#ifndef SPH_NBR_CELL_WAVES
#define SPH_NBR_CELL_WAVES 1
#endif
constexpr int kCellWaves = SPH_NBR_CELL_WAVES, kCellBlock = 64 * kCellWaves;
// candidate chunks held in registers per pass: 5 for the common one-material
// search (69 VGPRs; 6 spill), 2 where per-candidate materials or distances
// take registers too
#ifndef SPH_NBR_BLK
#define SPH_NBR_BLK 5
#endif
template <bool ONEMAT, bool MIND> constexpr int nbr_blk() { return ONEMAT && !MIND ? SPH_NBR_BLK : 2; }

__device__ __forceinline__ double readlane_d(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}

// Target t of a target group against the nb (<= NB) resident chunks
// that start at staged index kb: candidates inside the cut-off are appended
// to the target's list in candidate order.  selfk: the target's own staged
// index relative to kb.  Entry pos of the list of slot i is entry
// i * 8 + ((pos / 8) << cshift) + pos % 8 of the list array, with
// 1 << cshift = Params::nstride * 8 (a shift, not a multiply); its byte offset
// is 32-bit (the launcher keeps larger lists on the per-particle walk), so the
// store takes the list base in scalar registers.
template <int D, bool ONEMAT, bool MIND, int NB>
__device__ __forceinline__ void nbr_target_block(const Params &P, const Arrays &A, int t, int tg, int lane, int kb,
                                                 int nb, int ncand, int selfk, const double (&xi)[D], uint32_t fi,
                                                 const int (&cv)[NB], const double (&xcv)[NB][D],
                                                 const int (&matcv)[NB], double cut2_0,
                                                 unsigned int cshift, unsigned int maxnbr, int &nn, double &mind,
                                                 const double (&xt)[D])
{
    unsigned int nnt = (unsigned int)__builtin_amdgcn_readlane(nn, t);
    char *lst = reinterpret_cast<char *>(A.nbr);
    const unsigned int tbase = (unsigned int)(tg + t) * kNbrChunk;
    const int matt = ONEMAT ? 0 : info_mat((uint32_t)__builtin_amdgcn_readlane((int)fi, t));
    double r1 = 1.0e20;
#pragma unroll
    for (int c = 0; c < NB; c++) {
        if (c < nb) {
            double d2 = 0.0;
#pragma unroll
            for (int d = 0; d < D; d++) {
                const double dd = xt[d] - xcv[c][d];
                d2 += dd * dd;
            }
            const double cut2 = ONEMAT ? cut2_0 : P.pc[matt][matcv[c]].cut2;
#if SPH_NBR_SMASK
            // lanes past the candidates and the target itself as a
            // wave-uniform mask; the lane's own bit of the ballot is its test
            const int nv = ncand - kb - c * 64, sl = selfk - c * 64;
            const unsigned long long vm = nv >= 64 ? ~0ull : ((1ull << nv) - 1ull);
            const unsigned long long pm =
                __builtin_amdgcn_ballot_w64(d2 < cut2) & vm & ~(sl >= 0 && sl < 64 ? 1ull << sl : 0ull);
            const bool acc = __builtin_amdgcn_inverse_ballot_w64(pm);
#else
            const int k = c * 64 + lane;
            const bool acc = d2 < cut2 && kb + k < ncand && k != selfk;
            const unsigned long long pm = __builtin_amdgcn_ballot_w64(acc);
#endif
            if (acc) {
                const unsigned int pos = nnt + __builtin_amdgcn_mbcnt_hi((unsigned)(pm >> 32),
                                                                         __builtin_amdgcn_mbcnt_lo((unsigned)pm, 0u));
                // entries past the list capacity (an overflow, reported by the
                // caller) all land in the spare chunk row past it
                const unsigned int ps = pos < maxnbr ? pos : maxnbr;
                *reinterpret_cast<int *>(lst + ((tbase + ((ps >> 3) << cshift) + (ps & 7u)) << 2)) = cv[c];
                if (MIND) {
                    const double r = sqrt(d2);
                    r1 = r < r1 ? r : r1;
                }
            }
            nnt += (unsigned int)__popcll(pm);
        }
    }
    nn = lane == t ? (int)nnt : nn;
    if (MIND) {
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            const double y = __shfl_xor(r1, o, 64);
            r1 = y < r1 ? y : r1;
        }
        if (lane == t) mind = r1 < mind ? r1 : mind;
    }
}

template <int D, bool ONEMAT, bool MIND>
__device__ __forceinline__ void nbr_cells_body(const Params &P, const Arrays &A, int wave, int nwave,
                                               int *stage, double *tpos)
{
    Ctl *ctl = A.ctl;
    if (CTL_STOP(ctl)) return;
    const int lane = threadIdx.x & 63;
    // this wave's cells: those whose first slot lies in its share of the slots
    const int j0 = (int)((long long)P.N * wave / nwave), j1 = (int)((long long)P.N * (wave + 1) / nwave);
    const int ncode = ctl->ncode;
    const unsigned int m0 = ctl->cmask[0], m1 = ctl->cmask[1], m2 = ctl->cmask[2];
    // the spare row's first entry (lists hold nbr_round(maxnbr) + kNbrChunk entries)
    const unsigned int maxnbr = (unsigned int)nbr_round(P.maxnbr);
    const unsigned int cshift = (unsigned int)P.nshift - 2u;       // log2(nstride * kNbrChunk)
    int jn = j0;
    while (jn < j1) {
        const int q = __builtin_amdgcn_readfirstlane(A.occ[jn]);
        if (q >= ncode) break;                              // inactive bucket: the end
        const int head = __builtin_amdgcn_readfirstlane(A.cell_start[q]);
        jn = __builtin_amdgcn_readfirstlane(A.cell_start[q + 1]);
        if (head < j0) continue;                            // cell owned by the previous wave
        const int bc[3] = {code_ext(q, m0), code_ext(q, m1), code_ext(q, m2)};
        const int c0 = head, c1 = jn;
        // the cell's own positions for the box filter, loaded ahead of the
        // staging so that their latency overlaps it
        double xo[D];
        {
            const int it = c0 + lane < c1 ? c0 + lane : c0;
#pragma unroll
            for (int d = 0; d < D; d++) xo[d] = A.x[d][it];
        }
        int own;
        int ncand = stage_cell<D>(P, A, ctl, q, bc, lane, stage, own);
        if (ncand > kCandCap) {
            // left to k_nbr_ovf (per-particle walk)
            if (lane == 0) A.ovf[atomicAdd(&ctl->novf, 1)] = q;
            continue;
        }
        if (kNbrBoxFilter && ncand > 64) {
            // Drop the candidates farther than the largest cut-off from the
            // box spanned by the cell's own particles (parts of the edge and
            // corner cells, and of the faces where the particles do not fill
            // the cell): no target of the cell can reach them.  The margin is
            // far above FP64 rounding; order is kept (in-place compaction).
            // A stencil of one chunk (the 2D scenes) is tested as it is: the
            // filter could not save a chunk there.
            double blo[3] = {0.0, 0.0, 0.0}, bhi[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int d = 0; d < D; d++) {
                blo[d] = xo[d];
                bhi[d] = xo[d];
            }
            for (int tg = c0 + 64; tg < c1; tg += 64) {
                const int it = tg + lane < c1 ? tg + lane : c0;
#pragma unroll
                for (int d = 0; d < D; d++) {
                    const double xv = A.x[d][it];
                    blo[d] = xv < blo[d] ? xv : blo[d];
                    bhi[d] = xv > bhi[d] ? xv : bhi[d];
                }
            }
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) {
#pragma unroll
                for (int d = 0; d < D; d++) {
                    const double lo = __shfl_xor(blo[d], o, 64), hi = __shfl_xor(bhi[d], o, 64);
                    blo[d] = lo < blo[d] ? lo : blo[d];
                    bhi[d] = hi > bhi[d] ? hi : bhi[d];
                }
            }
            const double reach2 = P.cut2max * (1.0 + 1.0e-9);
            int nkeep = 0, own_new = 0;
            // blocks of kFiltBlk chunks: every load of a block is issued
            // before any is used.  The compacted writes of a block land below
            // its first unread entry (a kept count never exceeds the entries
            // read), and its reads precede its writes in program order.
            for (int kb0 = 0; kb0 < ncand; kb0 += 64 * kFiltBlk) {
                int jb[kFiltBlk];
                double xb[kFiltBlk][D];
#pragma unroll
                for (int c = 0; c < kFiltBlk; c++) {
                    const int kk = kb0 + c * 64 + lane;
                    jb[c] = kk < ncand ? stage[kk] : c0;
#pragma unroll
                    for (int d = 0; d < D; d++) xb[c][d] = A.x[d][jb[c]];
                }
#pragma unroll
                for (int c = 0; c < kFiltBlk; c++) {
                    const int k0 = kb0 + c * 64;
                    if (k0 < ncand) {
                        const int nv = ncand - k0;          // lanes holding candidates
                        double e2 = 0.0;
#pragma unroll
                        for (int d = 0; d < D; d++) {
                            // distance outside the box on this axis (one term is positive at most)
                            const double e = fmax(blo[d] - xb[c][d], 0.0) + fmax(xb[c][d] - bhi[d], 0.0);
                            e2 += e * e;
                        }
                        const unsigned long long km = __builtin_amdgcn_ballot_w64(e2 <= reach2) &
                                                      (nv >= 64 ? ~0ull : ((1ull << nv) - 1ull));
                        const int pos = nkeep + (int)__builtin_amdgcn_mbcnt_hi(
                                                    (unsigned)(km >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)km, 0u));
                        if (__builtin_amdgcn_inverse_ballot_w64(km)) stage[pos] = jb[c];
                        const int ol = own - k0;            // the cell's own particles are always kept
                        if (ol >= 0 && ol < 64) own_new = nkeep + (int)__popcll(km & ((1ull << ol) - 1ull));
                        nkeep += (int)__popcll(km);
                    }
                }
            }
            ncand = nkeep;
            own = own_new;
            __builtin_amdgcn_wave_barrier();
        }
        for (int tg = c0; tg < c1; tg += 64) {
            const int nt = __builtin_amdgcn_readfirstlane(c1 - tg < 64 ? c1 - tg : 64);
            const bool tv = lane < nt;
            const int it = tv ? tg + lane : tg;         // lanes past the cell read a valid slot
            double xi[D];
#pragma unroll
            for (int d = 0; d < D; d++) xi[d] = A.x[d][it];
            if (kNbrTposLds) {
                // the targets' positions, read back as broadcasts by the target loop
#pragma unroll
                for (int d = 0; d < D; d++) tpos[d * 64 + lane] = xi[d];
                __builtin_amdgcn_wave_barrier();
            }
            const uint32_t fi = A.info[it];
            const double mi = A.m[it];
            int nn = 0;
            double mind = 1.0e20;
            const int mat0 = info_mat((uint32_t)__builtin_amdgcn_readfirstlane((int)fi));
            const double cut2_0 = P.pc[mat0][mat0].cut2;
            const int self0 = own + (tg - c0);          // staged index of target 0
            // 2D and 1D stencils mostly fit one chunk: two are held there
            constexpr int NB = D == 3 ? nbr_blk<ONEMAT, MIND>() : 2;
            for (int kb = 0; kb < ncand; kb += 64 * NB) {
                int cv[NB];
                double xcv[NB][D];
                int matcv[NB];
                const int nb = min(NB, (ncand - kb + 63) >> 6);
#pragma unroll
                for (int c = 0; c < NB; c++) {
                    cv[c] = tg;
#pragma unroll
                    for (int d = 0; d < D; d++) xcv[c][d] = 0.0;
                    matcv[c] = 0;
                    if (c < nb) {                           // chunks holding candidates only
                        const int k = kb + c * 64 + lane;
                        cv[c] = k < ncand ? stage[k] : tg;  // lanes past the list read a valid slot
#pragma unroll
                        for (int d = 0; d < D; d++) xcv[c][d] = A.x[d][cv[c]];
                        matcv[c] = ONEMAT ? 0 : info_mat(A.info[cv[c]]);
                    }
                }
                for (int t = 0; t < nt; t++) {
                    // target t's position: a broadcast read from LDS (LDS
                    // instructions, where readlane would take six VALU ones)
                    double xt[D];
#pragma unroll
                    for (int d = 0; d < D; d++) xt[d] = kNbrTposLds ? tpos[d * 64 + t] : readlane_d(xi[d], t);
                    nbr_target_block<D, ONEMAT, MIND, NB>(P, A, t, tg, lane, kb, nb, ncand, self0 + t - kb, xi, fi, cv,
                                                          xcv, matcv, cut2_0, cshift, maxnbr, nn, mind, xt);
                }
            }
            if (tv) {
                if (nn > P.maxnbr) {
                    set_err(ctl, kErrNbrOverflow, A.orig[it]);
                    A.nnbr[it] = P.maxnbr;
                } else {
                    nbr_epilogue<D>(P, A, it, nn, mind);
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// MSolution::Strain half-step back (MSolution.cpp:334-338), unfused path
template <int D>
__device__ __forceinline__ void shift_body(const Params &P, const Arrays &A, const int bid)
{
    const Ctl *ctl = A.ctl;
    if (CTL_STOP(ctl)) return;
    const int i = bid * kBlock + threadIdx.x;
    if (i >= P.N) return;
    const uint32_t fi = A.info[i];
    const bool act = fi & kActive;
    const double dt = ctl->dt;
    double xs[D], vi[D];
#pragma unroll
    for (int d = 0; d < D; d++) {
        vi[d] = A.v[d][i];
        xs[d] = act ? A.x[d][i] - 0.5 * vi[d] * dt : A.x[d][i];
        A.xs[d][i] = xs[d];
    }
    put_srec<D>(A, i, xs, A.m[i] / A.rho[i], vi, fi);
}

// ---------------------------------------------------------------------------
// EOS of model 9 (MConstitutive.cpp:135-186)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double eos_p(const Mat &mt, double rho)
{
    if (mt.eos == 29) return (mt.c * mt.c) * (rho - mt.rho0);
    return mt.eos_b * (pow(rho / mt.rho0, mt.eos_gamma) - 1.0);
}

// critical step of one particle, MSolution::CalcCriticalTimestep (:34-94)
template <int D>
__device__ __forceinline__ double crit_one(const Params &P, const Arrays &A, const Ctl *ctl, int i,
                                           const Mat &mt, double rho_i)
{
    double ct;
    const double h = mt.h;
    if (P.tcrit_opt == 3) {
        ct = 0.25 * h / mt.c;
        const double b = 0.125 * (h * h) / mt.nu;
        ct = ct < b ? ct : b;
        // |a|^2 summed over the axes in order, after the last step (aa)
        const double aabs = sqrt(A.aa[i]);
        const double c3 = 0.25 * sqrt(h / aabs);
        ct = ct < c3 ? ct : c3;
        if (P.st) {
            const double c4 = 0.25 * sqrt(rho_i * ((h * h) * h) / 2 / 3.141592654 / mt.sigma);
            ct = ct < c4 ? ct : c4;
        }
    } else {
        double va = 0.0;
#pragma unroll
        for (int d = 0; d < D; d++) va += A.v[d][i] * A.v[d][i];
        va = sqrt(va);
        const double len = P.tcrit_opt == 1 ? h : A.mindist[i];
        ct = len / (mt.c + va);
        if (ctl->timestep == 1) ct = 1.e-09;
    }
    return ct;
}

// the minimum over the block's waves into the reduction slots, one atomic per
// wave in slot (wave % kBboxSlots) (slots on separate lines): no block barrier,
// so a wave that is done leaves while the block's other waves still sweep
// (the minimum does not depend on the grouping)
__device__ __forceinline__ void block_min_to(const Arrays &A, int bid, double v)
{
    v = wave_min(v);
    const int w = bid * (kBlock / 64) + (int)(threadIdx.x >> 6);
    if ((threadIdx.x & 63) == 0)
        atomicMin(A.bbox + (size_t)(w % kBboxSlots) * kBboxStride + kCritSlot, dkey(v));
}

// ---------------------------------------------------------------------------
// Sweep 1: MSolution::ClassicStrain (:362-413) -> trace of D; optionally the
// RhoUpdate + EOS epilogue (:463-477, MConstitutive.cpp:178-186) and the
// per-particle critical step with a block min into the control block.
// ---------------------------------------------------------------------------
enum { S_STRAIN = 1, S_CONST = 2, S_CRIT = 4, S_INPLACE = 8 };

template <int D, int G = kStrain64Batch>
__device__ __forceinline__ void strain_body(const Params &P, const Arrays &A, int mask, const int bid)
{
    Ctl *ctl = A.ctl;
    if (CTL_STOP(ctl)) return;
    const int i = bid * kBlock + threadIdx.x;
    double crit = 1.0e20;
    if (i < P.N) {
        const uint32_t fi = A.info[i];
        const bool act = fi & kActive;
        const int mat_i = info_mat(fi);
        const Mat &mt = P.mat[mat_i];
        const double dt = ctl->dt;
        double tr = 0.0;
        if ((mask & S_STRAIN) && act) {
            double xi[D], vi[D], g[D];
#pragma unroll
            for (int d = 0; d < D; d++) {
                xi[d] = A.xs[d][i];
                vi[d] = A.v[d][i];
                g[d] = 0.0;
            }
            struct Rec { double4 a, b; };
            auto load = [&](int j) { return Rec{gload(A.rsA + j), gload(A.rsB + j)}; };
            for_each_nbr_batched<G>(A.nbr, i, A.nnbr[i], P.nstride, load, [&](const Rec &rec) {
                const double4 ra = rec.a, rb = rec.b;
                const double volj = ra.w;
                const PairConst &pc = P.pc[mat_i][P.nmat == 1 ? mat_i : info_mat((uint32_t)rb.w)];
                double xj[D], vj[D], dw[D];
                vec_of<D>(ra, xj);
                vec_of<D>(rb, vj);
                gradw64<D>(dw, xi, xj, pc);
#pragma unroll
                for (int d = 0; d < D; d++) g[d] += volj * (vj[d] - vi[d]) * dw[d];
            });
#pragma unroll
            for (int d = 0; d < D; d++) tr += 0.5 * (g[d] + g[d]);
            A.trd[i] = tr;
            // half step forward again (:353-357), not bit-idempotent by design
#pragma unroll
            for (int d = 0; d < D; d++) A.x[d][i] = xi[d] + 0.5 * vi[d] * dt;
        } else if (mask & S_CONST) {
            tr = A.trd[i];
        }
        double rho_i = A.rho[i];
        if (mask & S_CONST) {
            if (act) {
                rho_i = rho_i * (1.0 - tr * dt);
                const double p = eos_p(mt, rho_i);
                const double pr = p / (rho_i * rho_i);
                A.p[i] = p;
                A.prho2[i] = pr;
                double xr[D], vr[D];
#pragma unroll
                for (int d = 0; d < D; d++) {
                    xr[d] = A.x[d][i];
                    vr[d] = A.v[d][i];
                }
                put_mrec<D>(P, A, i, xr, A.m[i], vr, rho_i, pr, fi);
            }
            A.rho_new[i] = rho_i;
            if (mask & S_INPLACE) A.rho[i] = rho_i;
        }
        if (mask & S_CRIT) {
            const double ct = crit_one<D>(P, A, ctl, i, mt, rho_i);
            if (!(mask & S_CONST)) A.critts_p[i] = ct;
            if (act && !(fi & kRigid)) crit = ct;
        }
    }
    if (mask & S_CRIT) block_min_to(A, bid, crit);
}

// ---------------------------------------------------------------------------
// CSF, MSolution::CalculateSurfaceNormals (:1429-1493): colour-gradient normal
// and its normalised form, then the divergence of the normalised normal.
// ---------------------------------------------------------------------------
template <int D>
__device__ __forceinline__ void normals_body(const Params &P, const Arrays &A, const int bid)
{
    const Ctl *ctl = A.ctl;
    if (CTL_STOP(ctl)) return;
    const int i = bid * kBlock + threadIdx.x;
    if (i >= P.N) return;
    const uint32_t fi = A.info[i];
    if (!(fi & kActive)) return;
    const int mat_i = info_mat(fi), col_i = info_color(fi);
    double xi[D], nrm[D];
#pragma unroll
    for (int d = 0; d < D; d++) {
        xi[d] = A.x[d][i];
        nrm[d] = 0.0;
    }
    for_each_nbr(A.nbr, i, A.nnbr[i], P.nstride, [&](int j) {
        const uint32_t fj = A.info[j];
        const double volj = A.m[j] / A.rho[j];
        const PairConst &pc = P.pc[mat_i][info_mat(fj)];
        double xj[D], dw[D];
#pragma unroll
        for (int d = 0; d < D; d++) xj[d] = A.x[d][j];
        gradw64<D>(dw, xi, xj, pc);
        const double f = volj * (double)(info_color(fj) - col_i);
#pragma unroll
        for (int d = 0; d < D; d++) nrm[d] += f * dw[d];
    });
    double in = 0.0;
#pragma unroll
    for (int d = 0; d < D; d++) in += nrm[d] * nrm[d];
    in = sqrt(in);
    const double Nf = in > 1.e-2 * P.mat[mat_i].h ? 1.0 : 0.0;
    A.Nf[i] = Nf;
#pragma unroll
    for (int d = 0; d < D; d++) {
        A.nrm[d][i] = nrm[d];
        A.nn[d][i] = Nf == 1.0 ? (1. / in) * nrm[d] : 0.0;
    }
}

template <int D>
__device__ __forceinline__ void curvature_body(const Params &P, const Arrays &A, const int bid)
{
    const Ctl *ctl = A.ctl;
    if (CTL_STOP(ctl)) return;
    const int i = bid * kBlock + threadIdx.x;
    if (i >= P.N) return;
    const uint32_t fi = A.info[i];
    if (!(fi & kActive)) return;
    const int mat_i = info_mat(fi);
    double xi[D], nni[D];
#pragma unroll
    for (int d = 0; d < D; d++) {
        xi[d] = A.x[d][i];
        nni[d] = A.nn[d][i];
    }
    const double Ni = A.Nf[i];
    double div = 0.0;
    for_each_nbr(A.nbr, i, A.nnbr[i], P.nstride, [&](int j) {
        const uint32_t fj = A.info[j];
        const bool img = fj & kGhost;        // images: N = 0, normals zero
        const double volj = A.m[j] / A.rho[j];
        const PairConst &pc = P.pc[mat_i][info_mat(fj)];
        double xj[D], dw[D];
#pragma unroll
        for (int d = 0; d < D; d++) xj[d] = A.x[d][j];
        gradw64<D>(dw, xi, xj, pc);
        double prod = 0;
#pragma unroll
        for (int d = 0; d < D; d++) prod += ((img ? 0.0 : A.nn[d][j]) - nni[d]) * dw[d];
        const double Nj = img ? 0.0 : A.Nf[j];
        div += (Ni < Nj ? Ni : Nj) * volj * prod;
    });
    A.kappa[i] = div;
}

// dt rule of MSolution::Solution (:275-277); resets the min key
// one wave: folds the critical-step slots into Ctl::critts_key and resets them
__device__ __forceinline__ void crit_fold_body(const Arrays &A)
{
    Ctl *ctl = A.ctl;
    if (CTL_STOP(ctl)) return;
    unsigned long long *slot = A.bbox + (size_t)threadIdx.x * kBboxStride + kCritSlot;
    unsigned long long k = *slot;
    *slot = dkey(1.0e20);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const unsigned long long y = __shfl_xor(k, o, 64);
        k = y < k ? y : k;
    }
    if (threadIdx.x == 0 && k < ctl->critts_key) ctl->critts_key = k;
}

// one wave (lane 0 applies the rule after the fold)
__device__ __forceinline__ void advance_body(const Params &P, const Arrays &A)
{
    Ctl *ctl = A.ctl;
    if (CTL_STOP(ctl)) return;
    crit_fold_body(A);
    if (threadIdx.x != 0) return;
    const double critts = dunkey(ctl->critts_key);
    ctl->critts = critts;
    ctl->critts_key = dkey(1.0e20);
    ctl->dtold = ctl->dt;
    const double a = P.tssf * critts, b = 1.1 * ctl->dtold;
    ctl->dt = a < b ? a : b;
    ctl->ptime += ctl->dt;
}

// ---------------------------------------------------------------------------
// Sweep 2: MSolution::MorrisAcceleration (:550-600); visc_type 2 adds the
// Monaghan (1992) artificial viscosity, Pi_ij = -alpha*cbar*mu_ij/rhobar for
// approaching pairs (extension; the reference's visc_type field is unused).
// ---------------------------------------------------------------------------
template <int D, int G = kMomentum64Batch, bool CMP = false>
__device__ __forceinline__ void momentum_body(const Params &P, const Arrays &A, const int bid)
{
    const Ctl *ctl = A.ctl;
    if (CTL_STOP(ctl)) return;
    const int i = bid * kBlock + threadIdx.x;
    if (i >= P.N) return;
    const uint32_t fi = A.info[i];
    double ai[D];
#pragma unroll
    for (int d = 0; d < D; d++) ai[d] = 0.0;
    if (fi & kActive) {
        const int mat_i = info_mat(fi);
        const Mat &mi = P.mat[mat_i];
        const bool rig_i = fi & kRigid;
        const double nu_i = mi.nu, rho_i = A.rho_new[i], pr_i = A.prho2[i];
        double xi[D], vi[D];
#pragma unroll
        for (int d = 0; d < D; d++) {
            xi[d] = A.x[d][i];
            vi[d] = A.v[d][i];
        }
        struct Rec { double4 a, b; double2 c; };
        auto load = [&](int j) {
            return CMP ? Rec{gload(A.rmA + j), gload(A.rmB + j), make_double2(0.0, 0.0)}
                       : Rec{gload(A.rmA + j), gload(A.rmB + j), gload(A.rmC + j)};
        };
        for_each_nbr_batched<G>(A.nbr, i, A.nnbr[i], P.nstride, load, [&](const Rec &rec) {
            const double4 ra = rec.a, rb = rec.b;
            // the compact record: one material and mass, no rigid body; p/rho^2 in a.w
            const double2 rc = CMP ? make_double2(ra.w, 0.0) : rec.c;
            const uint32_t fj = CMP ? 0u : (uint32_t)rc.y;
            const int mat_j = (CMP || P.nmat == 1) ? mat_i : info_mat(fj);
            const PairConst &pc = P.pc[mat_i][mat_j];
            const double mj = CMP ? P.mass64 : ra.w, rho_j = rb.w;
            double xj[D], vj[D], dw[D];
            vec_of<D>(ra, xj);
            vec_of<D>(rb, vj);
            gradw64<D>(dw, xi, xj, pc);
            double rab2 = 0.0, prod = 0.0;
#pragma unroll
            for (int d = 0; d < D; d++) {
                rab2 += (xi[d] - xj[d]) * (xi[d] - xj[d]);
                prod += (xi[d] - xj[d]) * dw[d];
            }
            const double sp = -mj * (rc.x + pr_i);
#pragma unroll
            for (int d = 0; d < D; d++) ai[d] = ai[d] + sp * dw[d];
            if (!CMP && rig_i && (fj & kRigid)) return;
            const double nu_j = P.mat[mat_j].nu;
            const double tv = mj * (nu_i * rho_i + nu_j * rho_j) * prod / (rho_i * rho_j) / (rab2 + pc.eps2);
#pragma unroll
            for (int d = 0; d < D; d++) ai[d] = ai[d] + tv * (vi[d] - vj[d]);
            if (mi.visc_type == 2) {
                double vr = 0.0;
#pragma unroll
                for (int d = 0; d < D; d++) vr += (vi[d] - vj[d]) * (xi[d] - xj[d]);
                if (vr < 0.0) {
                    const Mat &mj_ = P.mat[mat_j];
                    const double alpha = 0.5 * (mi.alpha + mj_.alpha);
                    const double mu = pc.havg * vr / (rab2 + pc.eps2);
                    const double pij = -alpha * (0.5 * (mi.c + mj_.c)) * mu / (0.5 * (rho_i + rho_j));
                    const double sv = -mj * pij;
#pragma unroll
                    for (int d = 0; d < D; d++) ai[d] = ai[d] + sv * dw[d];
                }
            }
        });
        if (P.st) {
            const double f = -mi.sigma / rho_i * A.kappa[i];
#pragma unroll
            for (int d = 0; d < D; d++) ai[d] = ai[d] + f * A.nrm[d][i];
        }
    }
#pragma unroll
    for (int d = 0; d < D; d++) A.a[d][i] = ai[d];
}

// ---------------------------------------------------------------------------
// Classic discretisation (disctype 0).
// Sweep 1: ClassicStrain (MSolution.cpp:362-413) with the full velocity
// gradient, rod = (grad v + grad v^T) / 2; with S_CONST the model-9
// constitutive update of MConstitutive::Constitutive (MConstitutive.cpp:10-42)
// after RhoUpdate: F3DM9 (:44-68) s = 2 mu (rod - tr/3 I) with mu = av_l rho0,
// AViscosity (:123-133) q only when tr D < 0 (q persists otherwise), the EOS,
// StressUp (:210-222) sigma = -p I + s; the momentum record holds
// T = (sigma - q I) / rho^2.  Symmetric tensors are kept as xx, yy, zz, xy,
// xz, yz (index sym6).
// ---------------------------------------------------------------------------
template <int D, int G = kStrain64Batch>
__device__ __forceinline__ void strain_classic_body(const Params &P, const Arrays &A, int mask, const int bid)
{
    Ctl *ctl = A.ctl;
    if (CTL_STOP(ctl)) return;
    const int i = bid * kBlock + threadIdx.x;
    double crit = 1.0e20;
    if (i < P.N) {
        const uint32_t fi = A.info[i];
        const bool act = fi & kActive;
        const int mat_i = info_mat(fi);
        const Mat &mt = P.mat[mat_i];
        const double dt = ctl->dt;
        double tr = 0.0, rod[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if ((mask & S_STRAIN) && act) {
            double xi[D], vi[D], g[D][D];
#pragma unroll
            for (int d = 0; d < D; d++) {
                xi[d] = A.xs[d][i];
                vi[d] = A.v[d][i];
#pragma unroll
                for (int e = 0; e < D; e++) g[d][e] = 0.0;
            }
            struct Rec { double4 a, b; };
            auto load = [&](int j) { return Rec{gload(A.rsA + j), gload(A.rsB + j)}; };
            for_each_nbr_batched<G>(A.nbr, i, A.nnbr[i], P.nstride, load, [&](const Rec &rec) {
                const double4 ra = rec.a, rb = rec.b;
                const double volj = ra.w;
                const PairConst &pc = P.pc[mat_i][P.nmat == 1 ? mat_i : info_mat((uint32_t)rb.w)];
                double xj[D], vj[D], dw[D];
                vec_of<D>(ra, xj);
                vec_of<D>(rb, vj);
                gradw64<D>(dw, xi, xj, pc);
#pragma unroll
                for (int r = 0; r < D; r++)
#pragma unroll
                    for (int c = 0; c < D; c++) g[r][c] += volj * (vj[r] - vi[r]) * dw[c];
            });
#pragma unroll
            for (int r = 0; r < D; r++)
#pragma unroll
                for (int c = r; c < D; c++) rod[sym6(r, c)] = 0.5 * (g[r][c] + g[c][r]);
#pragma unroll
            for (int d = 0; d < D; d++) tr += 0.5 * (g[d][d] + g[d][d]);
            A.trd[i] = tr;
#pragma unroll
            for (int k = 0; k < 6; k++) A.rod[k][i] = rod[k];
#pragma unroll
            for (int d = 0; d < D; d++) A.x[d][i] = xi[d] + 0.5 * vi[d] * dt;
        } else if (mask & S_CONST) {
            tr = A.trd[i];
#pragma unroll
            for (int k = 0; k < 6; k++) rod[k] = A.rod[k][i];
        }
        double rho_i = A.rho[i];
        if (mask & S_CONST) {
            if (act) {
                rho_i = rho_i * (1.0 - tr * dt);
                // F3DM9
                double sdev[6];
                const double xmu = mt.nu * mt.rho0;
                if (fabs(xmu) < 1.e-10) {
#pragma unroll
                    for (int k = 0; k < 6; k++) sdev[k] = 0.0;
                } else {
                    const double davg = 1.0 / 3.0 * tr;
#pragma unroll
                    for (int k = 0; k < 6; k++) sdev[k] = 2 * xmu * (rod[k] - (k < 3 ? davg : 0.0));
                }
                // AViscosity: the sound speed before this step's EOS (constant for EOS 28/29)
                double q = A.q[i];
                if (tr < 0.0)
                    q = mt.nu * rho_i * mt.c * mt.h * fabs(tr) + mt.av_q * rho_i * ((mt.h * tr) * (mt.h * tr));
                A.q[i] = q;
                const double p = eos_p(mt, rho_i);
                A.p[i] = p;
                A.prho2[i] = p / (rho_i * rho_i);
                double sig[6], T[6];
                const double r2 = rho_i * rho_i;
#pragma unroll
                for (int k = 0; k < 6; k++) {
                    sig[k] = -(k < 3 ? p : 0.0) + sdev[k];
                    A.sig[k][i] = sig[k];
                    T[k] = (sig[k] - (k < 3 ? q : 0.0)) / r2;
                }
                double xr[D];
#pragma unroll
                for (int d = 0; d < D; d++) xr[d] = A.x[d][i];
                double4 ra;
                ra.x = xr[0]; ra.y = D > 1 ? xr[1] : 0.0; ra.z = D > 2 ? xr[2] : 0.0; ra.w = A.m[i];
                A.rmA[i] = ra;
                put_crec<D>(A, i, T, fi);
            }
            A.rho_new[i] = rho_i;
            if (mask & S_INPLACE) A.rho[i] = rho_i;
        }
        if (mask & S_CRIT) {
            const double ct = crit_one<D>(P, A, ctl, i, mt, rho_i);
            if (!(mask & S_CONST)) A.critts_p[i] = ct;
            if (act && !(fi & kRigid)) crit = ct;
        }
    }
    if (mask & S_CRIT) block_min_to(A, bid, crit);
}

// Sweep 2: MSolution::ClassicAcceleration (:506-548); a is cleared for
// computable particles only (:519-521).
template <int D, int G = kMomentum64Batch>
__device__ __forceinline__ void momentum_classic_body(const Params &P, const Arrays &A, const int bid)
{
    const Ctl *ctl = A.ctl;
    if (CTL_STOP(ctl)) return;
    const int i = bid * kBlock + threadIdx.x;
    if (i >= P.N) return;
    const uint32_t fi = A.info[i];
    if (!(fi & kActive)) return;
    const int mat_i = info_mat(fi);
    double ai[D], xi[D], Ti[6];
    {
        const double4 a = A.rmA[i], b = A.rmB[i], c = A.rcC[i];
        vec_of<D>(a, xi);
        Ti[0] = b.x; Ti[1] = b.y; Ti[2] = b.z; Ti[3] = c.x; Ti[4] = c.y; Ti[5] = c.z;
    }
#pragma unroll
    for (int d = 0; d < D; d++) ai[d] = 0.0;
    struct Rec { double4 a, b, c; };
    auto load = [&](int j) { return Rec{gload(A.rmA + j), gload(A.rmB + j), gload(A.rcC + j)}; };
    for_each_nbr_batched<G>(A.nbr, i, A.nnbr[i], P.nstride, load, [&](const Rec &rec) {
        const double4 ra = rec.a, rb = rec.b, rc = rec.c;
        const int mat_j = P.nmat == 1 ? mat_i : info_mat((uint32_t)rb.w);
        const PairConst &pc = P.pc[mat_i][mat_j];
        const double Tj[6] = {rb.x, rb.y, rb.z, rc.x, rc.y, rc.z};
        double xj[D], dw[D];
        vec_of<D>(ra, xj);
        gradw64<D>(dw, xi, xj, pc);
        double ds[D];
#pragma unroll
        for (int n = 0; n < D; n++) {
            ds[n] = 0.0;
#pragma unroll
            for (int m = 0; m < D; m++) ds[n] += (Tj[sym6(m, n)] + Ti[sym6(m, n)]) * dw[m];
        }
        const double mj = ra.w;
#pragma unroll
        for (int d = 0; d < D; d++) ai[d] = ai[d] + mj * ds[d];
    });
#pragma unroll
    for (int d = 0; d < D; d++) A.a[d][i] = ai[d];
}

// ---------------------------------------------------------------------------
// Rigid body: MSolution::RigidBodyAcceleration (:1496-1564) with
// MRigidBody::CalculateCenterOfMass (MRigidBody.cpp:37-49); one block,
// wave-shuffle FP64 reductions in a fixed order (deterministic).
// ---------------------------------------------------------------------------
constexpr int kRigidBlock = 1024;

__device__ __forceinline__ double block_sum_big(double v, double *sh)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < kRigidBlock / 64; k++) s += sh[k];
    return s;   // valid in thread 0
}

__device__ __forceinline__ void rigid_reduce_body(const Params &P, const Arrays &A, int kin)
{
    Ctl *ctl = A.ctl;
    if (CTL_STOP(ctl)) return;
    __shared__ double sh[kRigidBlock / 64];
    __shared__ double R[3];
    const int nr = P.n_rigid;
    double px[3] = {0.0, 0.0, 0.0};
    for (int q = threadIdx.x; q < nr; q += kRigidBlock) {
        const int k = A.rigid_idx[q];
        for (int d = 0; d < P.ndim; d++) px[d] += A.x[d][k];
    }
    for (int d = 0; d < 3; d++) {
        double s = block_sum_big(px[d], sh);
        if (threadIdx.x == 0) R[d] = s * (1. / nr);
    }
    __syncthreads();
    double pa[3] = {0.0, 0.0, 0.0}, pal[3] = {0.0, 0.0, 0.0};
    for (int q = threadIdx.x; q < nr; q += kRigidBlock) {
        const int k = A.rigid_idx[q];
        if (!(A.info[k] & kActive)) continue;
        double ak[3] = {0.0, 0.0, 0.0}, rl[3] = {0.0, 0.0, 0.0};
        for (int d = 0; d < P.ndim; d++) {
            ak[d] = A.a[d][k];
            rl[d] = A.x[d][k] - R[d];
        }
        const double mk = A.m[k];
        for (int d = 0; d < 3; d++) pa[d] += mk * (0.0 + ak[d]);
        pal[0] += mk * (rl[1] * ak[2] - rl[2] * ak[1]);
        pal[1] += mk * (rl[2] * ak[0] - rl[0] * ak[2]);
        pal[2] += mk * (rl[0] * ak[1] - rl[1] * ak[0]);
    }
    double sa[3], sal[3];
    for (int d = 0; d < 3; d++) {
        sa[d] = block_sum_big(pa[d], sh);
        sal[d] = block_sum_big(pal[d], sh);
    }
    if (threadIdx.x == 0) {
        const double dtn = 0.5 * (ctl->dt + ctl->dtold);
        for (int d = 0; d < 3; d++) {
            ctl->rbR[d] = R[d];
            double Ad = sa[d] * (1. / P.rb_M);
            ctl->rbAL[d] = sal[d] * (1. / P.rb_I);
            if (P.baseaccel) Ad += P.base_a[d];
            ctl->rbA[d] = Ad;
            if (kin) {
                ctl->rbV[d] = ctl->rbV[d] + ctl->rbA[d] * dtn;
                ctl->rbW[d] = ctl->rbW[d] + ctl->rbAL[d] * dtn;
            }
        }
    }
}

// first half of MSolution::RigidBodyVelocity (:1567-1577), unfused path
__device__ __forceinline__ void rigid_kin_body(Ctl *ctl)
{
    if (CTL_STOP(ctl)) return;
    const double dtn = 0.5 * (ctl->dt + ctl->dtold);
    for (int d = 0; d < 3; d++) {
        ctl->rbV[d] = ctl->rbV[d] + ctl->rbA[d] * dtn;
        ctl->rbW[d] = ctl->rbW[d] + ctl->rbAL[d] * dtn;
    }
}

// ---------------------------------------------------------------------------
// Per-particle integration: MSolution::BoundCond (:700-726), UpdateVelocity
// (:602-637) + the particle half of RigidBodyVelocity (:1578-1588),
// MoveParticles (:769-809): integrate_particle per particle, and the block's
// bounding box of the moved positions folded into one reduction slot.
// ---------------------------------------------------------------------------
template <int D>
__device__ __forceinline__ void integrate_body(const Params &P, const Arrays &A, int mask, const int bid)
{
    Ctl *ctl = A.ctl;
    if (CTL_STOP(ctl)) return;
    const int i = bid * kBlock + threadIdx.x;
    double lo[D], hi[D];
#pragma unroll
    for (int d = 0; d < D; d++) {
        lo[d] = 1.0e20;
        hi[d] = -1.0e20;
    }
    if (i < P.N) {
        double a[D];
#pragma unroll
        for (int d = 0; d < D; d++) a[d] = (mask & (I_BC | I_VEL)) ? A.a[d][i] : 0.0;
        integrate_particle<D>(P, A, ctl, i, A.info[i], mask, a, lo, hi);
    }
    if (mask & I_MOVE) {
        __shared__ double red[2][D][kBlock / 64];
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
        for (int d = 0; d < D; d++) {
            double l = wave_min(lo[d]), h = wave_max(hi[d]);
            if (lane == 0) {
                red[0][d][w] = l;
                red[1][d][w] = h;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long *slot = A.bbox + (size_t)(bid % kBboxSlots) * kBboxStride;
#pragma unroll
            for (int d = 0; d < D; d++) {
                double l = red[0][d][0], h = red[1][d][0];
                for (int k = 1; k < kBlock / 64; k++) {
                    l = fmin(l, red[0][d][k]);
                    h = fmax(h, red[1][d][k]);
                }
                atomicMin(&slot[d], dkey(l));
                atomicMax(&slot[3 + d], dkey(h));
            }
        }
    }
}

// bounding box of MoveParticles into coord_minmax; with end != 0 also the
// step counter and stop test of MSolution::Solution (:317-319)
// one wave: lane l reduces and resets bounding-box slot l
__device__ __forceinline__ void finish_body(const Params &P, const Arrays &A, int bbox, int end, int next_grid = 0)
{
    Ctl *ctl = A.ctl;
    if (CTL_STOP(ctl)) return;
    const int lane = threadIdx.x;
    if (bbox) {
        static_assert(kBboxSlots == 64, "one slot per lane");
        unsigned long long *slot = A.bbox + (size_t)lane * kBboxStride;
        unsigned long long kmin[3], kmax[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            kmin[d] = slot[d];
            kmax[d] = slot[3 + d];
            slot[d] = dkey(1.0e20);
            slot[3 + d] = dkey(-1.0e20);
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1)
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const unsigned long long a = __shfl_xor(kmin[d], o, 64), b = __shfl_xor(kmax[d], o, 64);
                kmin[d] = a < kmin[d] ? a : kmin[d];
                kmax[d] = b > kmax[d] ? b : kmax[d];
            }
        if (lane == 0)
            for (int d = 0; d < 3; d++) {
                ctl->cmm[0][d] = d < P.ndim ? dunkey(kmin[d]) : 1.0e20;
                ctl->cmm[1][d] = d < P.ndim ? dunkey(kmax[d]) : -1.0e20;
            }
    }
    if (end && lane == 0) {
        ctl->timestep++;
        if (ctl->ptime - ctl->dt >= P.endtime) ctl->done = 1;
        if (ctl->step_limit > 0 && ctl->timestep - 1 >= ctl->step_limit) ctl->done = 1;
    }
    // the next binning's grid, from the bounding box just written (lane 0
    // wrote it; grid_setup_body returns if the loop has stopped)
    if (next_grid && lane == 0) grid_setup_body(P, ctl);
}

// ---------------------------------------------------------------------------
// kernels: single-scene form (parameters by value) and batched form (one
// launch over many scenes; block b works for scene map[b].x on its local
// block map[b].y, so every block belongs to one scene and per-scene
// parameters are wave-uniform loads from the scene tables).
// ---------------------------------------------------------------------------
#define SCENE_OF(e) const Params &Pb = Pt[(e)]; const Arrays &Ab = At[(e)]

__global__ void k_grid_setup(Params P, Ctl *ctl) { grid_setup_body(P, ctl); }
__global__ void kb_grid_setup(const Params *__restrict__ Pt, const Arrays *__restrict__ At, const int *sc, int ns)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < ns) { SCENE_OF(sc[t]); grid_setup_body(Pb, Ab.ctl); }
}
__global__ void k_zero_cells(Params P, Arrays A) { zero_cells_body(P, A, blockIdx.x, gridDim.x); }
__global__ void kb_zero_cells(const Params *__restrict__ Pt, const Arrays *__restrict__ At, const int *sc, int nb)
{
    SCENE_OF(sc[blockIdx.x / nb]);
    zero_cells_body(Pb, Ab, blockIdx.x % nb, nb);
}
template <int D> __global__ __launch_bounds__(kBlock) void k_bin_count(Params P, Arrays A) { bin_count_body<D>(P, A, blockIdx.x); }
template <int D> __global__ __launch_bounds__(kBlock) void kb_bin_count(const Params *__restrict__ Pt, const Arrays *__restrict__ At, const int2 *map)
{
    const int2 e = map[blockIdx.x];
    SCENE_OF(e.x);
    bin_count_body<D>(Pb, Ab, e.y);
}
__global__ __launch_bounds__(kBlock) void k_bin_scatter(Params P, Arrays A) { bin_scatter_body(P, A, blockIdx.x); }
__global__ __launch_bounds__(kBlock) void kb_bin_scatter(const Params *__restrict__ Pt, const Arrays *__restrict__ At, const int2 *map)
{
    const int2 e = map[blockIdx.x];
    SCENE_OF(e.x);
    bin_scatter_body(Pb, Ab, e.y);
}
__global__ __launch_bounds__(kBlock) void k_scan_reduce(const int *in, const int *n_ptr, int *aux, const Ctl *ctl)
{
    scan_reduce_body(in, n_ptr, aux, ctl, blockIdx.x);
}
__global__ __launch_bounds__(kBlock) void kb_scan_reduce(const Arrays *__restrict__ At, const int2 *map)
{
    const int2 e = map[blockIdx.x];
    const Arrays &Ab = At[e.x];
    scan_reduce_body(Ab.cell_count, &Ab.ctl->nbucket, Ab.scan_aux, Ab.ctl, e.y);
}
__global__ __launch_bounds__(kBlock) void kb_scan_aux(const Params *__restrict__ Pt, const Arrays *__restrict__ At, const int *sc)
{
    SCENE_OF(sc[blockIdx.x]);
    const int nb = (Pb.ncell_cap + kScanTile - 1) / kScanTile;
    scan_aux_body(Ab.scan_aux, nb, &Ab.ctl->nbucket, Ab.cell_start, Ab.ctl);
}
__global__ __launch_bounds__(kBlock) void kb_scan_down(const Arrays *__restrict__ At, const int2 *map)
{
    const int2 e = map[blockIdx.x];
    const Arrays &Ab = At[e.x];
    scan_down_body(Ab.cell_count, Ab.cell_start, &Ab.ctl->nbucket, Ab.scan_aux, Ab.ctl, e.y);
}
__global__ __launch_bounds__(kBlock) void k_scan_down_clear(int *in, int *out, const int *n_ptr, const int *aux, int nblocks,
                                                            const Ctl *ctl)
{
    scan_down_clear_body(in, out, n_ptr, aux, nblocks, ctl, blockIdx.x);
}
__global__ void k_reset_ngp(Ctl *ctl) { reset_ngp_body(ctl); }
__global__ void kb_reset_ngp(const Arrays *__restrict__ At, const int *sc, int ns)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < ns) reset_ngp_body(At[sc[t]].ctl);
}
__global__ __launch_bounds__(kBlock) void k_ghost_count(Params P, Arrays A, int m, int side)
{
    ghost_count_body(P, A, m, side, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(kBlock) void k_ghost_scan(Params P, Arrays A, int m, int side) { ghost_scan_body(P, A, m, side); }
__global__ __launch_bounds__(kBlock) void k_ghost_emit(Params P, Arrays A, int m, int side)
{
    ghost_emit_body(P, A, m, side, blockIdx.x, gridDim.x);
}
// batched image passes: scenes without this periodic side skip the pass
__device__ __forceinline__ bool has_pass(const Params &P, int m, int side)
{
    if (!P.periodic) return false;
    const int code = (m == 0) ? P.bcode[0][0] : P.bcode[side][m];
    return code == 2;
}
__global__ __launch_bounds__(kBlock) void kb_ghost_count(const Params *__restrict__ Pt, const Arrays *__restrict__ At, const int *sc, int nb, int m, int side)
{
    SCENE_OF(sc[blockIdx.x / nb]);
    if (has_pass(Pb, m, side)) ghost_count_body(Pb, Ab, m, side, blockIdx.x % nb, nb);
}
__global__ __launch_bounds__(kBlock) void kb_ghost_scan(const Params *__restrict__ Pt, const Arrays *__restrict__ At, const int *sc, int m, int side)
{
    SCENE_OF(sc[blockIdx.x]);
    if (has_pass(Pb, m, side)) ghost_scan_body(Pb, Ab, m, side);
}
__global__ __launch_bounds__(kBlock) void kb_ghost_emit(const Params *__restrict__ Pt, const Arrays *__restrict__ At, const int *sc, int nb, int m, int side)
{
    SCENE_OF(sc[blockIdx.x / nb]);
    if (has_pass(Pb, m, side)) ghost_emit_body(Pb, Ab, m, side, blockIdx.x % nb, nb);
}
template <int D> __global__ __launch_bounds__(kBlock) void k_permute(Params P, Arrays S, Arrays T, int move_a)
{
    permute_body<D>(P, S, T, blockIdx.x, move_a != 0);
}
// aa from the current accelerations (before a critical step outside the fused loop)
template <int D> __global__ __launch_bounds__(kBlock) void k_accel_sq(Params P, Arrays A)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= P.N) return;
    double s = 0.0;
#pragma unroll
    for (int d = 0; d < D; d++) s += A.a[d][i] * A.a[d][i];
    A.aa[i] = s;
}
template <int D> __global__ __launch_bounds__(kBlock) void kb_permute(const Params *__restrict__ Pt, const Arrays *__restrict__ At, const Arrays *__restrict__ At2, const int2 *map)
{
    const int2 e = map[blockIdx.x];
    permute_body<D>(Pt[e.x], At[e.x], At2[e.x], e.y, false);   // batched scenes are Morris (a rewritten)
}
#define XCD_BID(P)                                       \
    const int nb_ = ((P).N + kBlock - 1) / kBlock;       \
    const int bid_ = xcd_block(nb_);                     \
    if (bid_ >= nb_) return;
template <int D> __global__ __launch_bounds__(kBlock) void k_neighbours(Params P, Arrays A) { XCD_BID(P); neighbours_body<D>(P, A, bid_); }
// The grid is a multiple of 8 blocks; each wave takes an equal share of the
// slots.  The one-material search without distances fits 72 VGPRs (7 waves
// per SIMD) without spills.
#ifndef SPH_NBR_WAVES
#define SPH_NBR_WAVES 7
#endif
// the variants with per-candidate materials or distances, and the batched
// kernel that holds all four, need more registers: one wave fewer
template <bool ONEMAT, bool MIND> constexpr int nbr_waves() { return ONEMAT && !MIND ? SPH_NBR_WAVES : SPH_NBR_WAVES - 1; }
template <int D, bool ONEMAT, bool MIND>
__global__ __launch_bounds__(kCellBlock) __attribute__((amdgpu_waves_per_eu(nbr_waves<ONEMAT, MIND>(), 8))) void k_nbr_cells(Params P, Arrays A)
{
    __shared__ int stage[kCellWaves][kCandCap];
    __shared__ double tpos[kCellWaves][kNbrTposLds ? 3 * 64 : 1];
    const int nb = (int)gridDim.x;
    const int b = xcd_block(nb);
    const int wv = (int)(threadIdx.x >> 6);
    nbr_cells_body<D, ONEMAT, MIND>(P, A, b * kCellWaves + wv, nb * kCellWaves, stage[wv], tpos[wv]);
}

// Batched cell kernel: block b serves scene m.x, as its block m.y of m.z
// (the scene's waves split its slots as in k_nbr_cells); the variant is
// chosen per scene (uniform per block).
template <int D>
__global__ __launch_bounds__(kCellBlock) __attribute__((amdgpu_waves_per_eu(SPH_NBR_WAVES - 1, 8))) void kb_nbr_cells(const Params *__restrict__ Pt, const Arrays *__restrict__ At, const int4 *map)
{
    __shared__ int stage[kCellWaves][kCandCap];
    __shared__ double tpos[kCellWaves][kNbrTposLds ? 3 * 64 : 1];
    const int4 m = map[blockIdx.x];
    SCENE_OF(m.x);
    const int wv = (int)(threadIdx.x >> 6);
    const int w = m.y * kCellWaves + wv, nw = m.z * kCellWaves;
    if (Pb.nmat == 1 && !Pb.want_mindist) nbr_cells_body<D, true, false>(Pb, Ab, w, nw, stage[wv], tpos[wv]);
    else if (Pb.nmat == 1) nbr_cells_body<D, true, true>(Pb, Ab, w, nw, stage[wv], tpos[wv]);
    else if (!Pb.want_mindist) nbr_cells_body<D, false, false>(Pb, Ab, w, nw, stage[wv], tpos[wv]);
    else nbr_cells_body<D, false, true>(Pb, Ab, w, nw, stage[wv], tpos[wv]);
}

// grid of the cell kernel: about `slots` slots per wave (a few cells), so
// that waves of uneven cost balance over several rounds
static int cell_blocks(int n, int slots)
{
    slots = slots < 4 ? 4 : slots;
    long long nb = ((long long)n + (long long)slots * kCellWaves - 1) / ((long long)slots * kCellWaves);
    nb = nb < 8 ? 8 : (nb > (1 << 20) ? (1 << 20) : nb);
    return xcd_grid((int)nb);
}
// cells whose stencil did not fit the staging space: one block per cell,
// the per-particle walk for each of its particles
constexpr int kOvfBlocks = 256;
template <int D> __global__ __launch_bounds__(kBlock) void k_nbr_ovf(Params P, Arrays A)
{
    const Ctl *ctl = A.ctl;
    if (CTL_STOP(ctl)) return;
    const int n = ctl->novf;
    for (int w = blockIdx.x; w < n; w += gridDim.x) {
        const int q = A.ovf[w];
        const int c0 = A.cell_start[q], c1 = A.cell_start[q + 1];
        for (int i = c0 + (int)threadIdx.x; i < c1; i += kBlock) neighbours_walk<D>(P, A, i);
    }
}
// overflow cells of the batched cell kernel: kOvfBlocksPerScene blocks per scene
constexpr int kOvfBlocksPerScene = 8;
template <int D> __global__ __launch_bounds__(kBlock) void kb_nbr_ovf(const Params *__restrict__ Pt, const Arrays *__restrict__ At, const int *sc)
{
    const int s = sc[blockIdx.x / kOvfBlocksPerScene], part = (int)blockIdx.x % kOvfBlocksPerScene;
    SCENE_OF(s);
    const Ctl *ctl = Ab.ctl;
    if (CTL_STOP(ctl)) return;
    const int n = ctl->novf;
    for (int w = part; w < n; w += kOvfBlocksPerScene) {
        const int q = Ab.ovf[w];
        const int c0 = Ab.cell_start[q], c1 = Ab.cell_start[q + 1];
        for (int i = c0 + (int)threadIdx.x; i < c1; i += kBlock) neighbours_walk<D>(Pb, Ab, i);
    }
}
template <int D> __global__ __launch_bounds__(kBlock) void kb_neighbours(const Params *__restrict__ Pt, const Arrays *__restrict__ At, const int2 *map)
{
    const int2 e = map[blockIdx.x];
    SCENE_OF(e.x);
    neighbours_body<D>(Pb, Ab, e.y);
}
template <int D> __global__ __launch_bounds__(kBlock) void k_shift(Params P, Arrays A) { shift_body<D>(P, A, blockIdx.x); }
template <int D, int G> __global__ __launch_bounds__(kBlock) void k_strain(Params P, Arrays A, int mask) { XCD_BID(P); strain_body<D, G>(P, A, mask, bid_); }
template <int D> __global__ __launch_bounds__(kBlock) void kb_strain(const Params *__restrict__ Pt, const Arrays *__restrict__ At, const int2 *map, int mask)
{
    const int2 e = map[blockIdx.x];
    SCENE_OF(e.x);
    strain_body<D>(Pb, Ab, mask, e.y);
}
template <int D> __global__ __launch_bounds__(kBlock) void k_normals(Params P, Arrays A) { XCD_BID(P); normals_body<D>(P, A, bid_); }
template <int D> __global__ __launch_bounds__(kBlock) void kb_normals(const Params *__restrict__ Pt, const Arrays *__restrict__ At, const int2 *map)
{
    const int2 e = map[blockIdx.x];
    SCENE_OF(e.x);
    if (Pb.st) normals_body<D>(Pb, Ab, e.y);
}
template <int D> __global__ __launch_bounds__(kBlock) void k_curvature(Params P, Arrays A) { XCD_BID(P); curvature_body<D>(P, A, bid_); }
template <int D> __global__ __launch_bounds__(kBlock) void kb_curvature(const Params *__restrict__ Pt, const Arrays *__restrict__ At, const int2 *map)
{
    const int2 e = map[blockIdx.x];
    SCENE_OF(e.x);
    if (Pb.st) curvature_body<D>(Pb, Ab, e.y);
}
__global__ __launch_bounds__(64) void k_advance(Params P, Arrays A) { advance_body(P, A); }
__global__ __launch_bounds__(64) void k_crit_fold(Arrays A) { crit_fold_body(A); }
// one wave per scene
__global__ __launch_bounds__(64) void kb_advance(const Params *__restrict__ Pt, const Arrays *__restrict__ At, const int *sc)
{
    SCENE_OF(sc[blockIdx.x]);
    advance_body(Pb, Ab);
}
template <int D, int G, bool CMP> __global__ __launch_bounds__(kBlock) void k_momentum(Params P, Arrays A) { XCD_BID(P); momentum_body<D, G, CMP>(P, A, bid_); }
template <int D, int G> __global__ __launch_bounds__(kBlock) void k_strain_classic(Params P, Arrays A, int mask) { XCD_BID(P); strain_classic_body<D, G>(P, A, mask, bid_); }
template <int D, int G> __global__ __launch_bounds__(kBlock) void k_momentum_classic(Params P, Arrays A) { XCD_BID(P); momentum_classic_body<D, G>(P, A, bid_); }
template <int D> __global__ __launch_bounds__(kBlock) void kb_momentum(const Params *__restrict__ Pt, const Arrays *__restrict__ At, const int2 *map)
{
    const int2 e = map[blockIdx.x];
    SCENE_OF(e.x);
    if (Pb.mcompact64) momentum_body<D, kMomentum64Batch, true>(Pb, Ab, e.y);
    else momentum_body<D, kMomentum64Batch, false>(Pb, Ab, e.y);
}
__global__ __launch_bounds__(kRigidBlock) void k_rigid_reduce(Params P, Arrays A, int kin) { rigid_reduce_body(P, A, kin); }
__global__ __launch_bounds__(kRigidBlock) void kb_rigid_reduce(const Params *__restrict__ Pt, const Arrays *__restrict__ At, const int *sc, int kin)
{
    SCENE_OF(sc[blockIdx.x]);
    if (Pb.rigid) rigid_reduce_body(Pb, Ab, kin);
}
__global__ void k_rigid_kin(Ctl *ctl) { rigid_kin_body(ctl); }
template <int D> __global__ __launch_bounds__(kBlock) void k_integrate(Params P, Arrays A, int mask) { integrate_body<D>(P, A, mask, blockIdx.x); }
template <int D> __global__ __launch_bounds__(kBlock) void kb_integrate(const Params *__restrict__ Pt, const Arrays *__restrict__ At, const int2 *map, int mask)
{
    const int2 e = map[blockIdx.x];
    SCENE_OF(e.x);
    integrate_body<D>(Pb, Ab, mask, e.y);
}
__global__ __launch_bounds__(64) void k_finish(Params P, Arrays A, int bbox, int end, int next_grid)
{
    finish_body(P, A, bbox, end, next_grid);
}
// one wave per scene
__global__ __launch_bounds__(64) void kb_finish(const Params *__restrict__ Pt, const Arrays *__restrict__ At, const int *sc, int bbox, int end)
{
    SCENE_OF(sc[blockIdx.x]);
    finish_body(Pb, Ab, bbox, end);
}

// canonical neighbour ids (tests and the sph_get_neighbours entry point)
__global__ void k_canon(Params P, Arrays A, long long *out, int maxnbr)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.N) return;
    const int nn = A.nnbr[i];
    for (int k = 0; k < nn && k < maxnbr; k++) {
        const int j = A.nbr[nbr_slot(k, i, P.nstride)];
        long long cid;
        if (j < P.N) cid = A.orig[j];
        else {
            const int g = j - P.N;
            cid = -(1 + (long long)A.orig[A.gsrc[g]] * 27 + A.gshift[g]);
        }
        out[(size_t)A.orig[i] * maxnbr + k] = cid;
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline int nblk(int n) { return (n + kBlock - 1) / kBlock; }

#define DISPATCH_D(D_, CALL)                                   \
    switch (D_) {                                              \
    case 1: { constexpr int DD = 1; CALL; } break;             \
    case 2: { constexpr int DD = 2; CALL; } break;             \
    default: { constexpr int DD = 3; CALL; } break;            \
    }

// exclusive scan of the counts into the starts (total at out[n]); the counts
// are left zero
void launch_scan(const Params &P, const Arrays &A, int *in, int *out, const int *n_ptr, int cap, hipStream_t st)
{
    const int nb = (cap + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kBlock), 0, st, in, n_ptr, A.scan_aux, A.ctl);
    hipLaunchKernelGGL(k_scan_down_clear, dim3(nb), dim3(kBlock), 0, st, in, out, n_ptr, A.scan_aux, nb, A.ctl);
}

void launch_setup_linked_list(const Params &P, const Arrays &A, hipStream_t st, bool grid)
{
    if (grid) hipLaunchKernelGGL(k_grid_setup, dim3(1), dim3(1), 0, st, P, A.ctl);
    // the counts start at zero (left so by the previous scan); the image
    // counts of periodic scenes are cleared here
    if (P.periodic)
        hipLaunchKernelGGL(k_zero_cells, dim3(nblk(P.ncell_cap + 1) < 1024 ? nblk(P.ncell_cap + 1) : 1024),
                           dim3(kBlock), 0, st, P, A);
    DISPATCH_D(P.ndim, hipLaunchKernelGGL(k_bin_count<DD>, dim3(nblk(P.N)), dim3(kBlock), 0, st, P, A));
    launch_scan(P, A, A.cell_count, A.cell_start, &A.ctl->nbucket, P.ncell_cap, st);
    hipLaunchKernelGGL(k_bin_scatter, dim3(nblk(P.N)), dim3(kBlock), 0, st, P, A);
}

void launch_permute(const Params &P, const Arrays &S, const Arrays &T, hipStream_t st, bool move_a)
{
    DISPATCH_D(P.ndim, hipLaunchKernelGGL(k_permute<DD>, dim3(nblk(P.N)), dim3(kBlock), 0, st, P, S, T, move_a ? 1 : 0));
}

void launch_accel_sq(const Params &P, const Arrays &A, hipStream_t st)
{
    DISPATCH_D(P.ndim, hipLaunchKernelGGL(k_accel_sq<DD>, dim3(nblk(P.N)), dim3(kBlock), 0, st, P, A));
}

void launch_ghost_setup(const Params &P, const Arrays &A, hipStream_t st)
{
    hipLaunchKernelGGL(k_reset_ngp, dim3(1), dim3(1), 0, st, A.ctl);
    if (!P.periodic) return;
    for (int m = 0; m < 3; m++)
        for (int side = 0; side < 2; side++) {
            const int code = (m == 0) ? P.bcode[0][0] : P.bcode[side][m];
            if (code != 2) continue;
            hipLaunchKernelGGL(k_ghost_count, dim3(256), dim3(kBlock), 0, st, P, A, m, side);
            hipLaunchKernelGGL(k_ghost_scan, dim3(1), dim3(kBlock), 0, st, P, A, m, side);
            hipLaunchKernelGGL(k_ghost_emit, dim3(256), dim3(kBlock), 0, st, P, A, m, side);
        }
}

bool cell_lists_fit(const Params &P)
{
    return lists_fit32(P);
}

int batch_cell_blocks(int n, int slots)
{
    const long long per = (long long)slots * kCellWaves;
    const long long nb = ((long long)n + per - 1) / per;
    return (int)(nb < 1 ? 1 : nb);
}

void launch_neighbours(const Params &P, const Arrays &A, hipStream_t st, int cell_slots)
{
    if (!cell_lists_fit(P) || cell_slots == 0) {
        // the cell kernel's list byte offsets are 32-bit; cell_slots 0 asks
        // for the per-particle walk (the same lists in the same order)
        DISPATCH_D(P.ndim, hipLaunchKernelGGL(k_neighbours<DD>, dim3(xcd_grid(nblk(P.N))), dim3(kBlock), 0, st, P, A));
        return;
    }
    const dim3 grid(cell_blocks(P.N, cell_slots)), blk(kCellBlock);
    if (P.nmat == 1 && !P.want_mindist) {
        DISPATCH_D(P.ndim, hipLaunchKernelGGL((k_nbr_cells<DD, true, false>), grid, blk, 0, st, P, A));
    } else if (P.nmat == 1) {
        DISPATCH_D(P.ndim, hipLaunchKernelGGL((k_nbr_cells<DD, true, true>), grid, blk, 0, st, P, A));
    } else if (!P.want_mindist) {
        DISPATCH_D(P.ndim, hipLaunchKernelGGL((k_nbr_cells<DD, false, false>), grid, blk, 0, st, P, A));
    } else {
        DISPATCH_D(P.ndim, hipLaunchKernelGGL((k_nbr_cells<DD, false, true>), grid, blk, 0, st, P, A));
    }
    DISPATCH_D(P.ndim, hipLaunchKernelGGL(k_nbr_ovf<DD>, dim3(kOvfBlocks), dim3(kBlock), 0, st, P, A));
}

void launch_shift(const Params &P, const Arrays &A, hipStream_t st)
{
    DISPATCH_D(P.ndim, hipLaunchKernelGGL(k_shift<DD>, dim3(nblk(P.N)), dim3(kBlock), 0, st, P, A));
}

// records in flight per lane: 2, 4 or 8 (0: the default); the use order is
// the list order whatever the batch, so the sums are the same bits
#define DISPATCH_G(G_, DEF, CALL)                              \
    switch ((G_) ? (G_) : (DEF)) {                             \
    case 2: { constexpr int GG = 2; CALL; } break;             \
    case 8: { constexpr int GG = 8; CALL; } break;             \
    default: { constexpr int GG = 4; CALL; } break;            \
    }

void launch_strain(const Params &P, const Arrays &A, int mask, hipStream_t st, int batch)
{
    if (P.disctype == 0) {
        DISPATCH_G(batch, kStrain64Batch,
                   DISPATCH_D(P.ndim, hipLaunchKernelGGL((k_strain_classic<DD, GG>), dim3(xcd_grid(nblk(P.N))),
                                                         dim3(kBlock), 0, st, P, A, mask)));
        return;
    }
    DISPATCH_G(batch, kStrain64Batch,
               DISPATCH_D(P.ndim, hipLaunchKernelGGL((k_strain<DD, GG>), dim3(xcd_grid(nblk(P.N))), dim3(kBlock),
                                                     0, st, P, A, mask)));
}

void launch_normals(const Params &P, const Arrays &A, hipStream_t st)
{
    DISPATCH_D(P.ndim, hipLaunchKernelGGL(k_normals<DD>, dim3(xcd_grid(nblk(P.N))), dim3(kBlock), 0, st, P, A));
    DISPATCH_D(P.ndim, hipLaunchKernelGGL(k_curvature<DD>, dim3(xcd_grid(nblk(P.N))), dim3(kBlock), 0, st, P, A));
}

void launch_advance(const Params &P, const Arrays &A, hipStream_t st)
{
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, st, P, A);
}

void launch_crit_fold(const Params &P, const Arrays &A, hipStream_t st)
{
    (void)P;
    hipLaunchKernelGGL(k_crit_fold, dim3(1), dim3(64), 0, st, A);
}

void launch_momentum(const Params &P, const Arrays &A, int rigid_kin, hipStream_t st, int batch)
{
    if (P.disctype == 0) {   // no rigid body on this path (rejected at creation)
        DISPATCH_G(batch, kMomentum64Batch,
                   DISPATCH_D(P.ndim, hipLaunchKernelGGL((k_momentum_classic<DD, GG>), dim3(xcd_grid(nblk(P.N))),
                                                         dim3(kBlock), 0, st, P, A)));
        return;
    }
    if (P.mcompact64) {
        DISPATCH_G(batch, kMomentum64Batch,
                   DISPATCH_D(P.ndim, hipLaunchKernelGGL((k_momentum<DD, GG, true>), dim3(xcd_grid(nblk(P.N))),
                                                         dim3(kBlock), 0, st, P, A)));
    } else {
        DISPATCH_G(batch, kMomentum64Batch,
                   DISPATCH_D(P.ndim, hipLaunchKernelGGL((k_momentum<DD, GG, false>), dim3(xcd_grid(nblk(P.N))),
                                                         dim3(kBlock), 0, st, P, A)));
    }
    if (P.rigid && rigid_kin >= 0)
        hipLaunchKernelGGL(k_rigid_reduce, dim3(1), dim3(kRigidBlock), 0, st, P, A, rigid_kin);
}

__global__ __launch_bounds__(64) void k_crit_reset(Arrays A)
{
    A.bbox[(size_t)threadIdx.x * kBboxStride + kCritSlot] = dkey(1.0e20);
}

void launch_crit_reset(const Arrays &A, hipStream_t st)
{
    static_assert(kBboxSlots == 64, "one slot per lane");
    hipLaunchKernelGGL(k_crit_reset, dim3(1), dim3(64), 0, st, A);
}

void launch_rigid_reduce(const Params &P, const Arrays &A, int rigid_kin, hipStream_t st)
{
    if (P.rigid)
        hipLaunchKernelGGL(k_rigid_reduce, dim3(1), dim3(kRigidBlock), 0, st, P, A, rigid_kin);
}

void launch_rigid_kin(const Params &P, const Arrays &A, hipStream_t st)
{
    if (P.rigid) hipLaunchKernelGGL(k_rigid_kin, dim3(1), dim3(1), 0, st, A.ctl);
}

void launch_integrate(const Params &P, const Arrays &A, int mask, hipStream_t st)
{
    DISPATCH_D(P.ndim, hipLaunchKernelGGL(k_integrate<DD>, dim3(nblk(P.N)), dim3(kBlock), 0, st, P, A, mask));
}

void launch_finish(const Params &P, const Arrays &A, int bbox, int end, hipStream_t st, int next_grid)
{
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(64), 0, st, P, A, bbox, end, next_grid);
}

void launch_canon(const Params &P, const Arrays &A, long long *out, int maxnbr, hipStream_t st)
{
    hipLaunchKernelGGL(k_canon, dim3(nblk(P.N)), dim3(kBlock), 0, st, P, A, out, maxnbr);
}

// ---------------------------------------------------------------------------
// The reference's linked-list state, rebuilt from the cell-sorted storage of
// the last binning: sph_llgrid (MNeighbourVars.cpp:57-66, 70-80), the
// llpointer of every real particle (MSolution.cpp:135-144) and of every
// periodic image (MGhostParticleData.cpp:30-31).  Ids are the reference's:
// real particles 1-based deck numbers, image g as -g, LINKEDLISTEND at the
// end of a list.  A cell's list is its images newest first, then its real
// particles by descending number; the images of one cell are consecutive in
// creation order (one source cell feeds them).
// ---------------------------------------------------------------------------
constexpr int kLinkedListEnd = -1000000;   // global.h:13

__global__ __launch_bounds__(kBlock) void k_linked_list(Params P, Arrays A, int *llgrid, int *real_next,
                                                        int *ghost_next, int *ghost_par)
{
    const Ctl *ctl = A.ctl;
    const long long gl1 = ctl->glim[1], gl2 = ctl->glim[2];
    const long long ncell = (long long)ctl->glim[0] * gl1 * gl2;
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long t0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    for (long long lin = t0; lin < ncell; lin += stride) {
        const int k = (int)(lin % gl2), j = (int)((lin / gl2) % gl1), i = (int)(lin / (gl2 * gl1));
        const int q = cell_code(ctl, i, j, k);
        const int s0 = A.cell_start[q], s1 = A.cell_start[q + 1];
        const int real_head = s0 < s1 ? A.orig[s0] + 1 : kLinkedListEnd;
        int head = real_head;
        const int cnt = P.periodic ? A.gcount[lin] : 0;
        if (cnt > 0) {
            const int g0 = A.gfirst[lin];          // image g0 + r has id -(g0 + r + 1)
            head = -(g0 + cnt);
            if (ghost_next)
                for (int r = 0; r < cnt; r++) {
                    ghost_next[g0 + r] = r > 0 ? -(g0 + r) : real_head;
                    ghost_par[g0 + r] = A.orig[A.gsrc[g0 + r]] + 1;
                }
        }
        if (llgrid) llgrid[lin] = head;
    }
    if (real_next)
        for (long long s = t0; s < P.N; s += stride) {
            const int q = A.occ[s];
            int nx = kLinkedListEnd;                // inactive particles end their list at once
            if (q < ctl->ncode && s + 1 < A.cell_start[q + 1]) nx = A.orig[s + 1] + 1;
            real_next[A.orig[s]] = nx;
        }
}

void launch_linked_list(const Params &P, const Arrays &A, int *llgrid, int *real_next, int *ghost_next,
                        int *ghost_par, long long ncell, hipStream_t st)
{
    const long long items = ncell > P.N ? ncell : P.N;
    long long nb = (items + kBlock - 1) / kBlock;
    nb = nb < 1 ? 1 : (nb > 65536 ? 65536 : nb);
    hipLaunchKernelGGL(k_linked_list, dim3((unsigned)nb), dim3(kBlock), 0, st, P, A, llgrid, real_next, ghost_next,
                       ghost_par);
}

// ---------------------------------------------------------------------------
// MKernelBSpline::Kernel and ::GradW, the static forms
// (MKernelBSpline.cpp:98-187), for n pairs: xi, xj are n*3 coordinates, w n
// values, dwdx n*3, dwdr n (any output may be null).  The normalisations come
// in pc, computed on the host with the reference's operation order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_kernel_eval(int ndim, PairConst pc, int n, const double *xi,
                                                        const double *xj, double *w, double *dwdx, double *dwdr)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double xij[3] = {0.0, 0.0, 0.0}, rad = 0.0;
    for (int d = 0; d < ndim; d++) {
        xij[d] = xi[3 * (size_t)p + d] - xj[3 * (size_t)p + d];
        rad += xij[d] * xij[d];
    }
    rad = sqrt(rad);
    const double z = rad / pc.havg;
    if (w) {
        double W;
        if (z < 1.0)
            W = pc.const1 * (((0.75 * z - 1.5) * z) * z + 1.0);
        else if (z < 2.0)
            W = pc.const2 * ((2.0 - z) * (2.0 - z) * (2.0 - z));
        else
            W = 0.0;
        w[p] = W;
    }
    if (dwdx || dwdr) {
        double dW;
        if (z < 1.0)
            dW = pc.const1 * ((2.25 * z - 3.0) * z);
        else if (z < 2.0)
            dW = -3.0 * pc.const2 * ((2.0 - z) * (2.0 - z));
        else
            dW = 0.0;
        if (dwdx)
            for (int d = 0; d < 3; d++)
                dwdx[3 * (size_t)p + d] = (d >= ndim || rad == 0.0) ? 0.0 : dW * xij[d] / (rad * pc.havg);
        if (dwdr) dwdr[p] = rad == 0.0 ? 0.0 : dW / pc.havg;
    }
}

void launch_kernel_eval(int ndim, const PairConst &pc, int n, const double *xi, const double *xj, double *w,
                        double *dwdx, double *dwdr, hipStream_t st)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_kernel_eval, dim3(nblk(n)), dim3(kBlock), 0, st, ndim, pc, n, xi, xj, w, dwdx, dwdr);
}

// One batched step over the scenes of a BatchView (all of one ndim and arm).
void launch_batch_step(const BatchView &B, hipStream_t st)
{
    const int ns = B.nscene, sb = (ns + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(kb_grid_setup, dim3(sb), dim3(kBlock), 0, st, B.P, B.A, B.scenes, ns);
    hipLaunchKernelGGL(kb_zero_cells, dim3(ns * B.zero_nb), dim3(kBlock), 0, st, B.P, B.A, B.scenes, B.zero_nb);
    DISPATCH_D(B.ndim, hipLaunchKernelGGL(kb_bin_count<DD>, dim3(B.npart_blocks), dim3(kBlock), 0, st, B.P, B.A, B.pmap));
    hipLaunchKernelGGL(kb_scan_reduce, dim3(B.ntile_blocks), dim3(kBlock), 0, st, B.A, B.tmap);
    hipLaunchKernelGGL(kb_scan_aux, dim3(ns), dim3(kBlock), 0, st, B.P, B.A, B.scenes);
    hipLaunchKernelGGL(kb_scan_down, dim3(B.ntile_blocks), dim3(kBlock), 0, st, B.A, B.tmap);
    hipLaunchKernelGGL(kb_bin_scatter, dim3(B.npart_blocks), dim3(kBlock), 0, st, B.P, B.A, B.pmap);
    DISPATCH_D(B.ndim, hipLaunchKernelGGL(kb_permute<DD>, dim3(B.npart_blocks), dim3(kBlock), 0, st, B.P, B.A, B.A2, B.pmap));
    const Arrays *A2 = B.A2;     // the sorted buffer set from here on
    if (B.any_periodic) {
        hipLaunchKernelGGL(kb_reset_ngp, dim3(sb), dim3(kBlock), 0, st, A2, B.scenes, ns);
        for (int m = 0; m < 3; m++)
            for (int side = 0; side < 2; side++) {
                if (!(B.pass_mask & (1 << (2 * m + side)))) continue;
                hipLaunchKernelGGL(kb_ghost_count, dim3(ns * B.slab_nb), dim3(kBlock), 0, st, B.P, A2, B.scenes, B.slab_nb, m, side);
                hipLaunchKernelGGL(kb_ghost_scan, dim3(ns), dim3(kBlock), 0, st, B.P, A2, B.scenes, m, side);
                hipLaunchKernelGGL(kb_ghost_emit, dim3(ns * B.slab_nb), dim3(kBlock), 0, st, B.P, A2, B.scenes, B.slab_nb, m, side);
            }
    }
    if (B.ncell_blocks > 0) {
        DISPATCH_D(B.ndim, hipLaunchKernelGGL(kb_nbr_cells<DD>, dim3(B.ncell_blocks), dim3(kCellBlock), 0, st, B.P, A2, B.cmap));
        DISPATCH_D(B.ndim, hipLaunchKernelGGL(kb_nbr_ovf<DD>, dim3(ns * kOvfBlocksPerScene), dim3(kBlock), 0, st, B.P, A2, B.scenes));
    }
    if (B.nwalk_blocks > 0)
        DISPATCH_D(B.ndim, hipLaunchKernelGGL(kb_neighbours<DD>, dim3(B.nwalk_blocks), dim3(kBlock), 0, st, B.P, A2, B.wmap));
    if (B.arm == 1) launch_batch_strain32(B, st);
    else DISPATCH_D(B.ndim, hipLaunchKernelGGL(kb_strain<DD>, dim3(B.npart_blocks), dim3(kBlock), 0, st, B.P, A2, B.pmap, M_STRAIN | M_CONST | M_CRIT));
    if (B.any_st) {
        DISPATCH_D(B.ndim, hipLaunchKernelGGL(kb_normals<DD>, dim3(B.npart_blocks), dim3(kBlock), 0, st, B.P, A2, B.pmap));
        DISPATCH_D(B.ndim, hipLaunchKernelGGL(kb_curvature<DD>, dim3(B.npart_blocks), dim3(kBlock), 0, st, B.P, A2, B.pmap));
    }
    if (ns > 0) hipLaunchKernelGGL(kb_advance, dim3(ns), dim3(64), 0, st, B.P, A2, B.scenes);
    if (B.arm == 1) launch_batch_momentum32(B, st);
    else DISPATCH_D(B.ndim, hipLaunchKernelGGL(kb_momentum<DD>, dim3(B.npart_blocks), dim3(kBlock), 0, st, B.P, A2, B.pmap));
    if (B.any_rigid) hipLaunchKernelGGL(kb_rigid_reduce, dim3(ns), dim3(kRigidBlock), 0, st, B.P, A2, B.scenes, 1);
    DISPATCH_D(B.ndim, hipLaunchKernelGGL(kb_integrate<DD>, dim3(B.npart_blocks), dim3(kBlock), 0, st, B.P, A2, B.pmap, M_BC | M_VEL | M_MOVE | M_RHO));
    if (ns > 0) hipLaunchKernelGGL(kb_finish, dim3(ns), dim3(64), 0, st, B.P, A2, B.scenes, 1, 1);
}

// gathers the control blocks of a batch into one array for a single copy
__global__ void kb_gather_ctl(const Arrays *__restrict__ At, const int *sc, int ns, Ctl *out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < ns) out[t] = *At[sc[t]].ctl;
}

void launch_gather_ctl(const Arrays *At, const int *sc, int ns, Ctl *out, hipStream_t st)
{
    hipLaunchKernelGGL(kb_gather_ctl, dim3((ns + kBlock - 1) / kBlock), dim3(kBlock), 0, st, At, sc, ns, out);
}

}  // namespace sph
