// sph_state.hpp -- device-resident state of one scene and the by-value
// kernel parameter blocks shared by the HIP kernels and the host driver.
//
// Layout in HBM (structure of arrays, one array per field, FP64):
//   [0, N)          real particles, sorted every step by neighbour cell (the
//                   reference's cell order, z fastest) and, inside a cell, by
//                   descending deck number -- so each cell is one contiguous
//                   range in the reference's traversal order; orig[] holds
//                   the deck number (0-based) of each slot
//   [N, N + Gcap)   periodic images ("ghosts"), rebuilt every step; image g
//                   (1-based creation number, MGhostParticleData.cpp:30-31)
//                   lives at N + g - 1.
// Neighbour lists are neighbour-major, nbr[k * Ncap + i], so that entry k of
// 64 consecutive particles is one coalesced 256-byte row.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sph {

constexpr int kMaxMat = 8;
constexpr int kBlock = 256;

// Neighbour lists are stored in chunks of kNbrChunk entries: entry k of
// particle i sits at ((k / kNbrChunk) * nstride + i) * kNbrChunk + k % kNbrChunk
// (nstride: Params::nstride, a power of two, so a chunk offset is a shift),
// so a particle's consecutive entries share one 32-byte sector (its writes
// merge in L2) and a wave reads a chunk of 64 lists in whole lines.
constexpr int kNbrChunk = 8;
// candidates one wave stages for a cell of the cell-centric neighbour search
constexpr int kCandCap = 1024;
constexpr int kBboxSlots = 64, kBboxStride = 16, kCritSlot = 8;
inline __host__ __device__ size_t nbr_slot(int k, int i, int ncap)
{
    return ((size_t)(k / kNbrChunk) * (size_t)ncap + (size_t)i) * kNbrChunk + (size_t)(k % kNbrChunk);
}
inline __host__ __device__ int nbr_round(int maxnbr) { return (maxnbr + kNbrChunk - 1) / kNbrChunk * kNbrChunk; }
// entries per slot of the list memory: the rounded capacity and one spare
// chunk, where the cell search stores entries past the capacity (an overflow)
inline __host__ __device__ int nbr_rows(int maxnbr) { return nbr_round(maxnbr) + kNbrChunk; }

// packed per-particle flags (info[])
constexpr uint32_t kMatMask = 0xFFu;         // bits 0-7 material (1-based)
constexpr int kDispShift = 8;                // bits 8-11 dispbc
constexpr uint32_t kActive = 1u << 12;       // MCommonParticle::ToCompute()
constexpr uint32_t kRigid = 1u << 13;        // MSolution::InRigidBody
constexpr uint32_t kGhost = 1u << 14;        // periodic image
constexpr int kColorShift = 16;              // bits 16-31 colour (int16)

__host__ __device__ inline int info_mat(uint32_t f) { return (int)(f & kMatMask); }
__host__ __device__ inline int info_disp(uint32_t f) { return (int)((f >> kDispShift) & 0xFu); }
__host__ __device__ inline int info_color(uint32_t f) { return (int)(int16_t)(f >> kColorShift); }

// error codes mirrored from include/sph07_amd.h
constexpr int kErrNbrOverflow = 2;
constexpr int kErrOutOfGrid = 3;
constexpr int kErrGridCap = 16;   // grid larger than the cell arrays (host grows, retries)
constexpr int kErrGhostCap = 17;  // more images than the image arrays (host grows, retries)

// Device control block: every scalar the step loop mutates.
struct Ctl {
    double dt, dtold, ptime, critts;
    double gmin[3], gsize[3];
    double cmm[2][3];
    double rbR[3], rbV[3], rbW[3], rbA[3], rbAL[3];
    unsigned long long critts_key;
    int glim[3];
    int ncell;
    int ncode;            // size of the cell-code space (storage order of the cells)
    int nbucket;          // ncode + 1: the last bucket holds inactive particles
    unsigned int cmask[3];// code bits owned by each axis (bit-interleaved cell code)
    int timestep, done, err, ngp;
    int novf;             // cells whose stencil exceeds the cell kernel's staging space
    int err_particle;
    int step_limit;       // > 0: the scene stops after this many steps (batched runs)
};

// Per-material constants (MMaterial.h:17-40 after CheckDefaults).
struct Mat {
    double rho0, nu, sigma, c, eos_b, eos_gamma, alpha, h;
    double av_q;          // quadratic artificial-viscosity coefficient (Classic discretisation)
    int eos, visc_type;
};

// Per material-pair kernel constants: havg = AVG(h_i,h_j) and the B-spline
// normalisations of MKernelBSpline.cpp:145-156, evaluated once with the
// reference's operation order (bit-identical to recomputing them per pair).
struct PairConst {
    double havg, const1, const2, eps2, cut2;   // eps2 = 0.01*SQR(havg), cut2 = SQR(2*havg)
};

struct Params {
    int N, Ncap;          // real particles, stride of per-particle arrays (N + Gcap)
    // neighbour lists: chunk stride in particles, the power of two >= Ncap,
    // and log2 of the chunk stride in bytes (nstride * kNbrChunk * 4)
    int nstride, nshift;
    int gcap;             // image capacity
    int ndim, tcrit_opt, maxnbr;
    int st, rigid, baseaccel, periodic;
    int nmat, rb_id, n_rigid;
    int ncell_cap;
    int want_mindist;
    int arm;              // 0 FP64, 1 FP32-compensated (selects the records the fused step writes)
    int disctype;         // 2 Morris (MorrisAcceleration), 0 Classic (full stress, ClassicAcceleration)
    // storage order of the cells: 0 Morton (bit-interleaved code), 1 row-major
    // (z fastest); results do not depend on it, only the memory locality
    int corder;
    // FP32c momentum record: 0 the 40-byte form; 1 32 bytes when every
    // material has one particle mass (mass32) and the linear EOS, so m and
    // p/rho^2 follow from the material and the stored rho - rho0; 2 32 bytes
    // when besides there is one material and no rigid body: 1/rho and p/rho^2
    // are stored (sph_rec32.hpp)
    int mcompact;
    float mass32[kMaxMat + 1];
    // FP32c record positions are fixed point in units of q = 2^qexp, the
    // power of two with (largest cut-off) / q in [2^22, 2^23); qinv = 1 / q
    int qexp;
    double qinv;
    // FP64 momentum record: 0 {x, m}, {v, rho}, {p/rho^2, info} (80 bytes);
    // 1 {x, p/rho^2}, {v, rho} (64 bytes) for one material with one particle
    // mass (mass64), no rigid body and the Morris discretisation
    int mcompact64;
    double mass64;
    int btype[3];
    int bcode[2][3];
    double bx[2][3];
    double base_a[3];
    double minadd[3][3], maxadd[3][3];
    double hmax, tssf, endtime;
    double cut2max, cut2min;                   // over material pairs
    double rb_M, rb_I;
    Mat mat[kMaxMat + 1];                      // 1-based
    PairConst pc[kMaxMat + 1][kMaxMat + 1];    // 1-based
};

// whether the neighbour lists (nstride * nbr_rows(maxnbr) entries of 4 bytes)
// fit 32-bit byte offsets; the per-particle arrays (fewer than nstride
// entries of at most 8 bytes) then fit too
inline __host__ __device__ bool lists_fit32(const Params &P)
{
    return (unsigned long long)P.nstride * (unsigned long long)nbr_rows(P.maxnbr) * 4ull <= 0xffffffffull;
}

struct Arrays {
    double *x[3], *xs[3], *v[3], *a[3];
    double *rho, *rho_new, *p, *prho2, *m;
    // |a|^2 after the last step's BoundCond (what the critical step of option 3
    // reads; written with the velocity update, moved with the particle)
    double *aa;
    double *trd, *mindist, *critts_p, *rhoold;
    double *nrm[3], *nn[3], *Nf, *kappa;
    // gather records of the pair sweeps (one or two 32-byte loads per
    // neighbour instead of one load per field):
    //   continuity: rsA = {xs[0..2], m/rho}, rsB = {v[0..2], info}
    //   momentum:   rmA = {x[0..2], m}, rmB = {v[0..2], rho}, rmC = {p/rho^2, info}
    double4 *rsA, *rsB, *rmA, *rmB;
    double2 *rmC;
    // FP32 records of the FP32-compensated arm (fixed-point positions X),
    // two float4 per particle side by side (32 bytes, one sector):
    //   continuity: fs[2j] = {X[0..2], vol}, fs[2j+1] = {v[0..2], packed}
    //   momentum:   fm[2j] = {X[0..2], m}, fm[2j+1] = {v[0..2], rho}, fmC[j] = {p/rho^2, packed}
    float4 *fs, *fm;
    float2 *fmC;
    // Classic discretisation (disctype 0): q, the diagonal artificial viscosity
    // that persists between steps (moved with the particle); the rate of
    // deformation and stress (symmetric: xx, yy, zz, xy, xz, yz); the momentum
    // record is rmA = {x, m}, rmB = {T_xx, T_yy, T_zz, info}, rcC = {T_xy,
    // T_xz, T_yz, 0} with T = (sigma - q I) / rho^2 (zero for images)
    double *q;
    double *rod[6], *sig[6];
    double4 *rcC;

    uint32_t *info;
    int *nnbr, *nbr;
    // deck number of each slot, rigid-body rank (or -1) of each slot
    int *orig, *rrank;
    // binning of real particles (slot[] doubles as the sorted destination)
    int *cell_of, *slot, *cell_count, *cell_start, *tmp;
    // periodic images: per target cell the first image index and the count
    // (the images of one cell are consecutive, created from one source cell)
    int *gcount, *gfirst, *gsrc, *gshift;
    int *slab_cnt, *slab_off;
    int *scan_aux;
    int *rigid_idx;
    // per-block reductions spread over kBboxSlots slots on separate cache lines
    // (ordered keys): [0..2] min and [3..5] max of the moved positions (reduced
    // and reset by the step's finish), [kCritSlot] the critical step (folded
    // into Ctl::critts_key); a block updates slot (block % kBboxSlots)
    unsigned long long *bbox;
    // cell code of each sorted slot (written while the particles are sorted)
    int *occ;
    // cells left to the per-particle neighbour walk (oversized stencils)
    int *ovf;
    Ctl *ctl;
};

// Ordered 64-bit key of a double: unsigned order == numeric order.
__host__ __device__ inline unsigned long long dkey(double d)
{
    unsigned long long b;
    __builtin_memcpy(&b, &d, 8);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__host__ __device__ inline double dunkey(unsigned long long k)
{
    unsigned long long b = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
    double d;
    __builtin_memcpy(&d, &b, 8);
    return d;
}

}  // namespace sph
