// sph_fp32c.hip -- FP32-compensated pair sweeps.
//
// Positions, velocities, densities, the time integration and every
// per-particle update stay FP64 (the same k_integrate / k_advance / rigid
// kernels as the FP64 arm).  Only the bulk pair sums run in FP32: each
// neighbour is one compact FP32 record (sph_rec32.hpp) holding its position
// as a fixed-point integer, so a pair separation is one integer subtraction
// and one conversion, without cancellation of large coordinates; the
// kernel gradient, pressure, viscous and Monaghan terms are FP32 with hardware
// reciprocal and reciprocal square root; each particle's sums are accumulated
// with the TwoSum error-free transformation (Knuth 1969) into (sum,
// compensation) FP32 pairs (two neighbours' terms per addition in the
// continuity sweep, four in the momentum sweep), and
// converted to FP64 once at the end.
//
// Neighbour sets and their order are those of the FP64 arm (shared lists).
#include "sph_kernels.hpp"
#include "sph_rec32.hpp"

#include <math.h>

namespace sph {

#define CTL_STOP32(ctl) ((ctl)->err != 0 || (ctl)->done != 0)

// neighbour records in flight per lane (more hides more gather latency, at
// the cost of registers and occupancy; measured on the 3D dam break)
#ifndef SPH_STRAIN32_BATCH
#define SPH_STRAIN32_BATCH 8
#endif
#ifndef SPH_MOMENTUM32_BATCH
#define SPH_MOMENTUM32_BATCH 4
#endif
constexpr int kStrain32Batch = SPH_STRAIN32_BATCH, kMomentum32Batch = SPH_MOMENTUM32_BATCH;


// TwoSum: s + e == a + b exactly (Knuth, TAOCP vol. 2, 4.2.2)
__device__ __forceinline__ void two_sum_acc(float &s, float &c, float x)
{
    const float t = s + x;
    const float bp = t - s;
    const float err = (s - (t - bp)) + (x - bp);
    s = t;
    c += err;
}

// Pair constants for separations in units of q (qinv = 1 / q, a power of
// two): ih = q / havg, so z = r' ih with r' = r / q; eps2 = 0.01 havg^2 / q^2;
// havg / q.  c1ih = const1 / havg is unchanged: grad W = f dx = f' dx' with
// f' = dW/dz c1ih / r' (gradw_factor of r'^2).
struct PairConst32 {
    float ih, c1ih, eps2, havg;
};

__device__ __forceinline__ PairConst32 pc32(const PairConst &pc, double qinv)
{
    PairConst32 q;
    q.havg = (float)(pc.havg * qinv);
    q.ih = (float)(1.0 / (pc.havg * qinv));
    q.c1ih = (float)(pc.const1 / pc.havg);
    q.eps2 = (float)(pc.eps2 * qinv * qinv);
    return q;
}

// dW/dr / r in FP32 from r^2 (so that grad W = f * dx; with r^2 and dx in
// units of q, pc32's constants give f' with f' dx' = grad W).  The B-spline
// derivative (MKernelBSpline.cpp:166-172) in branch-free form:
//   dW/dz = c1 (3 (1-z)+^2 - 0.75 (2-z)+^2),
// which is c1 (2.25 z^2 - 3 z) for z < 1, -3 c2 (2-z)^2 for 1 <= z < 2
// (c2 = c1 / 4) and 0 beyond; f = 0 at r = 0 as in the reference.  r^2 is
// in units of q^2 and separations are whole quanta, so any r > 0 has r^2 >= 1:
// r^2 = 0 takes rinv = 1, z = 0 and dW/dz = 3 - 0.75 * 4 = 0 exactly.
__device__ __forceinline__ float gradw_factor(float r2, const PairConst32 &q)
{
    const float rinv = __builtin_amdgcn_rsqf(fmaxf(r2, 1.0f));
    const float z = r2 * rinv * q.ih;
    const float a = fmaxf(2.0f - z, 0.0f), b = fmaxf(1.0f - z, 0.0f);
    const float dwz = 3.0f * (b * b) - 0.75f * (a * a);
    return dwz * (rinv * q.c1ih);
}

__device__ __forceinline__ double eos_p32(const Mat &mt, double rho)
{
    if (mt.eos == 29) return (mt.c * mt.c) * (rho - mt.rho0);
    return mt.eos_b * (pow(rho / mt.rho0, mt.eos_gamma) - 1.0);
}

// critical step of one particle, MSolution::CalcCriticalTimestep (:34-94),
// from values its caller has read: aa (|a|^2 after the last step, option 3),
// v, mindist (option 2), whether this is the first step
template <int D>
__device__ __forceinline__ double crit_of32(const Params &P, const Mat &mt, double rho_i, double aa,
                                            const double (&v)[D], double mind, bool step1)
{
    double ct;
    const double h = mt.h;
    if (P.tcrit_opt == 3) {
        ct = 0.25 * h / mt.c;
        const double b = 0.125 * (h * h) / mt.nu;
        ct = ct < b ? ct : b;
        // |a|^2 summed over the axes in order, after the last step (aa)
        const double aabs = sqrt(aa);
        const double c3 = 0.25 * sqrt(h / aabs);
        ct = ct < c3 ? ct : c3;
        if (P.st) {
            const double c4 = 0.25 * sqrt(rho_i * ((h * h) * h) / 2 / 3.141592654 / mt.sigma);
            ct = ct < c4 ? ct : c4;
        }
    } else {
        double va = 0.0;
#pragma unroll
        for (int d = 0; d < D; d++) va += v[d] * v[d];
        va = sqrt(va);
        const double len = P.tcrit_opt == 1 ? h : mind;
        ct = len / (mt.c + va);
        if (step1) ct = 1.e-09;
    }
    return ct;
}

// The particle's epilogue of the continuity sweep: the half step forward
// again, RhoUpdate, EOS, the momentum record and the critical step.  Every
// input is read before any store (the compiler cannot tell the arrays apart
// and would wait for each store before the next load).
template <int D>
__device__ __forceinline__ double strain32_finish(const Params &P, const Arrays &A, const Ctl *ctl, int i, bool act,
                                                  uint32_t fi, const Mat &mt, double dt, float gs, float gc)
{
    double rho_i = A.rho[i];
    double x0[D], vi[D];
#pragma unroll
    for (int d = 0; d < D; d++) {
        x0[d] = A.x[d][i];
        vi[d] = A.v[d][i];
    }
    const double mi = A.m[i];
    const double aa = P.tcrit_opt == 3 ? A.aa[i] : 0.0;
    const double mind = P.tcrit_opt == 2 ? A.mindist[i] : 0.0;
    const bool step1 = ctl->timestep == 1;
    if (act) {
        const double tr = (double)gs + (double)gc;
        A.trd[i] = tr;
        double xr[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int d = 0; d < D; d++) {
            // MSolution::Strain's half step back and forward again
            // (:334-338, :353-357), each rounded as written (no FMA): the
            // back step is the one k_permute used for the records
#pragma clang fp contract(off)
            const double xi = x0[d] - 0.5 * vi[d] * dt;
            xr[d] = xi + 0.5 * vi[d] * dt;
            A.x[d][i] = xr[d];
        }
        rho_i = rho_i * (1.0 - tr * dt);
        const double p = eos_p32(mt, rho_i);
        const double pr = p / (rho_i * rho_i);
        A.p[i] = p;
        A.prho2[i] = pr;
        put_mrec32<D>(P, A, i, xr, mi, vi, rho_i, pr, fi);
    }
    A.rho_new[i] = rho_i;
    return crit_of32<D>(P, mt, rho_i, aa, vi, mind, step1);
}

// a wave-uniform float kept in a scalar register
__device__ __forceinline__ float uniform_f(float v)
{
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

__device__ __forceinline__ double wave_min32(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}

// The wave's minimum critical step into reduction slot (w % kBboxSlots) with
// one atomic per wave: no block barrier, so a wave that is done leaves while
// the other waves of its block still sweep (the minimum does not depend on
// the grouping)
__device__ __forceinline__ void crit_min_wave(const Arrays &A, int w, double crit)
{
    crit = wave_min32(crit);
    if ((threadIdx.x & 63) == 0)
        atomicMin(A.bbox + (size_t)(w % kBboxSlots) * kBboxStride + kCritSlot, dkey(crit));
}

// continuity sweep + RhoUpdate/EOS + critical step (fused mode of k_strain);
// G neighbour records in flight per lane
template <int D, int G = kStrain32Batch, bool ONEMAT = false>
__device__ __forceinline__ void strain32_body(const Params &P, const Arrays &A, const int bid)
{
    Ctl *ctl = A.ctl;
    if (CTL_STOP32(ctl)) return;
    const int i = bid * kBlock + threadIdx.x;
    // The particle's own loads are issued by every lane, ahead of any branch
    // (lanes past the particles read the last one and store nothing): with
    // the setup outside the divergent region the sweep ran 7 % faster.
    const bool inr = i < P.N;
    const int ic = inr ? i : P.N - 1;
    double crit = 1.0e20;
    const uint32_t fi = A.info[ic];
    const bool act = inr && (fi & kActive);
    const int mat_i = info_mat(fi);
    const Mat &mt = P.mat[mat_i];
    const double dt = ctl->dt;
    float ri[D], vi32[D];
    float gs = 0.0f, gc = 0.0f;
    const float4 fa = rec32_at<0>(A.fs, (uint32_t)ic), fb = rec32_at<1>(A.fs, (uint32_t)ic);
    const float ri3[3] = {fa.x, fa.y, fa.z}, vi3[3] = {fb.x, fb.y, fb.z};
#pragma unroll
    for (int d = 0; d < D; d++) {
        ri[d] = ri3[d];
        vi32[d] = vi3[d];
    }
    const PairConst32 q0 = pc32(P.pc[mat_i][mat_i], P.qinv);
    auto pair = [&](const float4 ra, const float4 rb) {
        const uint32_t pk_j = __float_as_uint(rb.w);
        const PairConst32 q = ONEMAT ? q0 : pc32(P.pc[mat_i][rec_mat(pk_j)], P.qinv);
        const float rj[3] = {ra.x, ra.y, ra.z}, vj[3] = {rb.x, rb.y, rb.z};
        // separation and r^2 in units of q
        float dx[D], r2 = 0.0f;
#pragma unroll
        for (int d = 0; d < D; d++) {
            dx[d] = qsep(ri[d], rj[d]);
            r2 += dx[d] * dx[d];
        }
        // the pair's share of tr D, sum over d of (v_j - v_i)_d * vol_j * dW/dx_d,
        // goes into one compensated accumulator
        float dvx = 0.0f;
#pragma unroll
        for (int d = 0; d < D; d++) dvx += (vj[d] - vi32[d]) * dx[d];
        return (gradw_factor(r2, q) * ra.w) * dvx;
    };
    if (act) {
        struct Rec { float4 a, b; };
        auto load = [&](int j) { return Rec{rec32_at<0>(A.fs, (uint32_t)j), rec32_at<1>(A.fs, (uint32_t)j)}; };
        // two pairs' terms per compensated addition (a padding entry's term
        // with the particle itself is exactly +0)
        for_each_nbr_pairs<G, true>(A.nbr, ic, A.nnbr[ic], P.nstride, load, [&](const Rec &a, const Rec &b) {
            two_sum_acc(gs, gc, pair(a.a, a.b) + pair(b.a, b.b));
        });
    }
    if (inr) {
        const double ct = strain32_finish<D>(P, A, ctl, i, act, fi, mt, dt, gs, gc);
        if (act && !(fi & kRigid)) crit = ct;
    }
    crit_min_wave(A, bid * (kBlock / 64) + (int)(threadIdx.x >> 6), crit);
}

// The continuity sweep with two lanes per particle.  A 32-byte record takes
// two 16-byte loads, and the sweep is bound by the vector L1's tag lookups
// (one cache line per lookup, about one lookup per cycle per CU): with one
// lane per particle each load instruction touches up to 64 records.  Here
// lanes 2t and 2t + 1 both serve particle t and take its list entries two
// at a time: the first load gives lane 2t half 0 and lane 2t + 1 half 1 of
// entry a's record, the second gives lane 2t half 1 and lane 2t + 1 half 0
// of entry b's record, so every load touches one record per lane pair.  A
// swap of one half between the two lanes (DPP) completes record a in lane 2t
// and record b in lane 2t + 1; each lane computes its pair's term, and the
// sum of the two terms (the same in both lanes: addition commutes) goes into
// both lanes' compensated accumulators, exactly the sum and order of the
// one-lane form.  Block: kBlock particles on 2 * kBlock lanes.  Taken in 3D
// (the 3D dam break: 239 -> 229 us per sweep); the 2D scenes, with a third of
// the neighbours, run faster with one lane per particle (the ensemble's
// batched step lost 8 % with two).
#ifndef SPH_STRAIN32_PAIRS
#define SPH_STRAIN32_PAIRS 1
#endif
template <int D>
constexpr bool strain32_pairs()
{
    return SPH_STRAIN32_PAIRS != 0 && D == 3;
}
template <int D>
constexpr int strain32_threads()
{
    return strain32_pairs<D>() ? 2 * kBlock : kBlock;
}
// the momentum sweep's lane pairs, in 3D as well (the 3D dam break: 203 ->
// 197.5 us; the 2D ensemble's batched step took 1.45x as long with them)
template <int D>
constexpr bool momentum32_pairs()
{
    return strain32_pairs<D>();
}
template <int D>
constexpr int momentum32_threads()
{
    return momentum32_pairs<D>() ? 2 * kBlock : kBlock;
}

// v from the other lane of the lane pair (lanes 2t and 2t + 1 swap)
__device__ __forceinline__ float pair_swap(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, true));
}

// half `part` (0 or 1) of record j
__device__ __forceinline__ float4 rec32_half(const float4 *base, uint32_t j, uint32_t part)
{
    return gload(reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(base) + ((j << 5) + (part << 4))));
}

template <int D, int G = kStrain32Batch, bool ONEMAT = false>
__device__ __forceinline__ void strain32_pair_body(const Params &P, const Arrays &A, const int bid)
{
    static_assert(G == 2 || G == 4 || G == 8, "records in flight per lane");
    Ctl *ctl = A.ctl;
    if (CTL_STOP32(ctl)) return;
    const uint32_t par = threadIdx.x & 1u;
    const int i = bid * kBlock + (int)(threadIdx.x >> 1);
    const bool inr = i < P.N;
    const int ic = inr ? i : P.N - 1;
    double crit = 1.0e20;
    const uint32_t fi = A.info[ic];
    const bool act = inr && (fi & kActive);
    const int mat_i = info_mat(fi);
    const Mat &mt = P.mat[mat_i];
    const double dt = ctl->dt;
    float ri[D], vi32[D];
    float gs = 0.0f, gc = 0.0f;
    const float4 fa = rec32_at<0>(A.fs, (uint32_t)ic), fb = rec32_at<1>(A.fs, (uint32_t)ic);
    const float ri3[3] = {fa.x, fa.y, fa.z}, vi3[3] = {fb.x, fb.y, fb.z};
#pragma unroll
    for (int d = 0; d < D; d++) {
        ri[d] = ri3[d];
        vi32[d] = vi3[d];
    }
    const PairConst32 q0 = pc32(P.pc[mat_i][mat_i], P.qinv);
    // the pair term of the one-lane form
    auto pair = [&](const float4 ra, const float4 rb) {
        const uint32_t pk_j = __float_as_uint(rb.w);
        const PairConst32 q = ONEMAT ? q0 : pc32(P.pc[mat_i][rec_mat(pk_j)], P.qinv);
        const float rj[3] = {ra.x, ra.y, ra.z}, vj[3] = {rb.x, rb.y, rb.z};
        float dx[D], r2 = 0.0f;
#pragma unroll
        for (int d = 0; d < D; d++) {
            dx[d] = qsep(ri[d], rj[d]);
            r2 += dx[d] * dx[d];
        }
        float dvx = 0.0f;
#pragma unroll
        for (int d = 0; d < D; d++) dvx += (vj[d] - vi32[d]) * dx[d];
        return (gradw_factor(r2, q) * ra.w) * dvx;
    };
    if (act) {
        // both lanes of a pair have the same particle, so the same trip
        // counts.  A pass loads E = 2G entries' records (G per lane): for G = 8
        // two chunks of the list (the chunks past the list read as the
        // particle itself), else E entries of one chunk.
        constexpr int E = 2 * G, CH = E > kNbrChunk ? E / kNbrChunk : 1;
        const int nn = A.nnbr[ic];
        for (int k0 = 0; k0 < nn; k0 += kNbrChunk * CH) {
            int e[kNbrChunk * CH];
#pragma unroll
            for (int c = 0; c < CH; c++) {
                int4 c0 = make_int4(ic, ic, ic, ic), c1 = c0;
                if (c == 0 || k0 + c * kNbrChunk < nn) {
                    const int4 *lp = nbr_chunk(A.nbr, k0 + c * kNbrChunk, ic, P.nstride);
                    c0 = gload(lp);
                    c1 = gload(lp + 1);
                }
                const int ec[kNbrChunk] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
#pragma unroll
                for (int u = 0; u < kNbrChunk; u++) e[c * kNbrChunk + u] = ec[u];
            }
            const int m = nn - k0;
#pragma unroll
            for (int g0 = 0; g0 < kNbrChunk * CH; g0 += E) {
                if (g0 >= m) break;
                // entries past nn are padding (the particle itself: a term of +0)
                float4 ha[G], hb[G];
#pragma unroll
                for (int u = 0; u < G; u++) {
                    ha[u] = rec32_half(A.fs, (uint32_t)e[g0 + 2 * u], par);
                    hb[u] = rec32_half(A.fs, (uint32_t)e[g0 + 2 * u + 1], par ^ 1u);
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int u = 0; u < G; u++) {
                    if (g0 + 2 * u >= m) break;
                    // lane 2t: record a = (ha, swapped ha of lane 2t + 1);
                    // lane 2t + 1: record b = (hb, swapped hb of lane 2t)
                    const float4 own = par ? hb[u] : ha[u], snd = par ? ha[u] : hb[u];
                    const float4 oth = make_float4(pair_swap(snd.x), pair_swap(snd.y), pair_swap(snd.z), pair_swap(snd.w));
                    const float t = pair(own, oth);
                    two_sum_acc(gs, gc, t + pair_swap(t));
                }
            }
        }
    }
    if (inr && par == 0) {
        const double ct = strain32_finish<D>(P, A, ctl, i, act, fi, mt, dt, gs, gc);
        if (act && !(fi & kRigid)) crit = ct;
    }
    crit_min_wave(A, bid * (2 * kBlock / 64) + (int)(threadIdx.x >> 6), crit);
}

// momentum sweep: pressure gradient, Morris viscosity, Monaghan term.
// KIND: the momentum record of Params::mcompact (sph_rec32.hpp): 0 full
// (40 bytes), 1 {R, rho - rho0}, {v, packed}, 2 {R, 1/rho}, {v, p/rho^2}
// (one material, no rigid body; implies ONEMAT).
// PAIRS: two lanes per particle, as in the continuity sweep's lane pairs
// (the same sums in the same order: four terms per compensated addition,
// (t_a + t_b) + (t_c + t_d), with t_a, t_c from lane 2t and t_b, t_d from
// lane 2t + 1); block: kBlock particles on 2 * kBlock lanes.
// fuse: the storing lane also integrates its particle (integrate_particle,
// the full pass of k_integrate without a rigid body) and leaves the moved
// position in lo/hi for the caller's bounding box.
template <int D, int G = kMomentum32Batch, int KIND = 0, bool ONEMAT = false, bool PAIRS = false>
__device__ __forceinline__ void momentum32_body(const Params &P, const Arrays &A, const int bid, int fuse,
                                                double (&lo)[D], double (&hi)[D])
{
    static_assert(KIND != 2 || ONEMAT, "the one-material record has no material bits");
    const Ctl *ctl = A.ctl;
    if (CTL_STOP32(ctl)) return;
    const uint32_t par = PAIRS ? (threadIdx.x & 1u) : 0u;
    const int i = bid * kBlock + (int)(PAIRS ? threadIdx.x >> 1 : threadIdx.x);
    if (i >= P.N) return;
    const uint32_t fi = A.info[i];
    double ai[D];
#pragma unroll
    for (int d = 0; d < D; d++) ai[d] = 0.0;
    if (fi & kActive) {
        const int mat_i = info_mat(fi);
        const Mat &mi = P.mat[mat_i];
        const bool rig_i = fi & kRigid;
        const double rho_i64 = A.rho_new[i];
        const float rho_i = (float)rho_i64, pr_i = (float)A.prho2[i];
        // viscosities carry 1 / q: the Morris term's r . grad W / (r^2 + 0.01 h^2)
        // is 1 / q times its value in units of q
        const float q_i = (float)(mi.nu * P.qinv * rho_i64);
        const bool monaghan = mi.visc_type == 2;
        float ri[D], vi[D], as[D], ac[D];
        const float4 fa = rec32_at<0>(A.fm, (uint32_t)i), fb = rec32_at<1>(A.fm, (uint32_t)i);
        const float ri3[3] = {fa.x, fa.y, fa.z}, vi3[3] = {fb.x, fb.y, fb.z};
#pragma unroll
        for (int d = 0; d < D; d++) {
            ri[d] = ri3[d];
            vi[d] = vi3[d];
            as[d] = 0.0f;
            ac[d] = 0.0f;
        }
        const PairConst32 q0 = pc32(P.pc[mat_i][mat_i], P.qinv);
        struct Rec { float4 a, b; float2 c; };
        auto load = [&](int j) {
            Rec r;
            r.a = rec32_at<0>(A.fm, (uint32_t)j);
            r.b = rec32_at<1>(A.fm, (uint32_t)j);
            if (KIND == 0) r.c = gload(A.fmC + j);
            return r;
        };
        const float rho0_i = (float)mi.rho0, c2_i = (float)(mi.c * mi.c), m_i = P.mass32[mat_i];
        const float irho_i = (float)(1.0 / rho_i64);
        // one material: the partner's constants are the particle's own
        const float nu_i = (float)(mi.nu * P.qinv);
        const float mnu_i = m_i * nu_i;
        const float acbar_i = (float)(mi.alpha * mi.c);     // alpha_bar * c_bar for equal materials
        // one material and the material's mass (not the per-particle masses
        // of the 40-byte record): m_j alpha_bar c_bar havg of the Monaghan
        // term, a constant
        constexpr bool kConstMach = ONEMAT && KIND != 0;
        const float mach_i = m_i * acbar_i * q0.havg;
        struct Term { float t[D]; };
        auto term = [&](const Rec &rec) {
            const float4 ra = rec.a, rb = rec.b;
            const uint32_t fj = KIND == 1 ? __float_as_uint(rb.w) : (KIND == 0 ? __float_as_uint(rec.c.y) : 0u);
            const int mat_j = ONEMAT ? mat_i : rec_mat(fj);
            const PairConst32 q = ONEMAT ? q0 : pc32(P.pc[mat_i][mat_j], P.qinv);
            float mj, rho_j = 0.0f, prho2_j, ir_j;
            if (KIND == 2) {
                mj = m_i;
                ir_j = ra.w;
                prho2_j = rb.w;
            } else if (KIND == 1) {
                const float rho0_j = ONEMAT ? rho0_i : (float)P.mat[mat_j].rho0;
                const float c2_j = ONEMAT ? c2_i : (float)(P.mat[mat_j].c * P.mat[mat_j].c);
                mj = ONEMAT ? m_i : P.mass32[mat_j];
                rho_j = rho0_j + ra.w;
                ir_j = __builtin_amdgcn_rcpf(rho_j);
                prho2_j = (fj & kRecGhost) ? 0.0f : c2_j * ra.w * ir_j * ir_j;
            } else {
                mj = ra.w;
                rho_j = rb.w;
                ir_j = __builtin_amdgcn_rcpf(rho_j);
                prho2_j = rec.c.x;
            }
            const float rj[3] = {ra.x, ra.y, ra.z}, vj[3] = {rb.x, rb.y, rb.z};
            float dx[D], dv[D], r2 = 0.0f, vr = 0.0f;
#pragma unroll
            for (int d = 0; d < D; d++) {
                dx[d] = qsep(ri[d], rj[d]);
                dv[d] = vi[d] - vj[d];
                r2 += dx[d] * dx[d];
                vr += dv[d] * dx[d];
            }
            const float f = gradw_factor(r2, q);            // grad W = f * dx (dx, r2 in units of q)
            float sdw = -mj * (prho2_j + pr_i);             // coefficient of grad W
            // viscous terms, skipped between two rigid particles; evaluated
            // for every pair and selected (no divergent branch per pair)
            const bool visc = KIND == 2 || !(rig_i && (fj & kRecRigid));
            const float ir2e = __builtin_amdgcn_rcpf(r2 + q.eps2);
            // Morris: m_j (mu_i + mu_j) (r . grad W) / (rho_i rho_j (r^2 + 0.01 h^2));
            // with one material mu = nu rho, so (mu_i + mu_j) / (rho_i rho_j) = nu (1/rho_i + 1/rho_j)
            float tv;
            if (KIND == 2) {
                tv = mnu_i * (irho_i + ir_j) * (f * r2) * ir2e;
            } else {
                const float qj = (ONEMAT ? nu_i : (float)(P.mat[mat_j].nu * P.qinv)) * rho_j;
                tv = visc ? mj * (q_i + qj) * (f * r2) * (irho_i * ir_j * ir2e) : 0.0f;
            }
            if (monaghan) {
                // Monaghan: -m_j Pi_ij grad W, Pi_ij = -alpha cbar mu / rho_bar for approaching pairs
                float acbar = acbar_i;
                if (!ONEMAT) {
                    const Mat &mj_ = P.mat[mat_j];
                    acbar = (float)(0.5 * (mi.alpha + mj_.alpha)) * (float)(0.5 * (mi.c + mj_.c));
                }
                // 1 / rho_bar = 2 / (rho_i + rho_j) = 2 (1/rho_i)(1/rho_j) / (1/rho_i + 1/rho_j)
                const float irbar = KIND == 2 ? 2.0f * irho_i * ir_j * __builtin_amdgcn_rcpf(irho_i + ir_j)
                                              : 2.0f * __builtin_amdgcn_rcpf(rho_i + rho_j);
                // m_j alpha_bar c_bar mu / rho_bar, mu = havg (v . r) / (r^2 + 0.01 havg^2)
                const float pterm = kConstMach ? mach_i * (vr * ir2e) * irbar : mj * (acbar * (q.havg * vr * ir2e)) * irbar;
                sdw += (visc && vr < 0.0f) ? pterm : 0.0f;
            }
            const float sf = sdw * f;
            Term tm;
#pragma unroll
            for (int d = 0; d < D; d++) tm.t[d] = sf * dx[d] + tv * dv[d];
            return tm;
        };
        // four pairs' terms per compensated addition, summed as two sums of
        // two (a padding entry's term with the particle itself is exactly +0:
        // r = 0 gives grad W = 0); at least four records in flight per lane
        if constexpr (!PAIRS) {
            constexpr int GQ = G < 4 ? 4 : G;
            for_each_nbr_quads<GQ, true>(A.nbr, i, A.nnbr[i], P.nstride, load,
                                         [&](const Rec &a, const Rec &b, const Rec &c, const Rec &e) {
                const Term ta = term(a), tb = term(b), tc = term(c), te = term(e);
#pragma unroll
                for (int d = 0; d < D; d++) two_sum_acc(as[d], ac[d], (ta.t[d] + tb.t[d]) + (tc.t[d] + te.t[d]));
            });
        } else {
            // entries a, b of an entry pair: lane 2t takes half 0 of a's record
            // and half 1 of b's, lane 2t + 1 the other halves; one swap gives
            // lane 2t record a and lane 2t + 1 record b.  A pass loads E = 2GP
            // entries' records (GP = max(G, 4) per lane, a whole chunk at least:
            // a pass over half a chunk is not unrolled and its entries spill):
            // for G = 8 two chunks of the list (the chunks past the list read as
            // the particle itself).
            constexpr int GP = G < 4 ? 4 : G;
            constexpr int E = 2 * GP, CH = E > kNbrChunk ? E / kNbrChunk : 1;
            const int nn = A.nnbr[i];
            for (int k0 = 0; k0 < nn; k0 += kNbrChunk * CH) {
                int e[kNbrChunk * CH];
#pragma unroll
                for (int c = 0; c < CH; c++) {
                    int4 c0 = make_int4(i, i, i, i), c1 = c0;
                    if (c == 0 || k0 + c * kNbrChunk < nn) {
                        const int4 *lp = nbr_chunk(A.nbr, k0 + c * kNbrChunk, i, P.nstride);
                        c0 = gload(lp);
                        c1 = gload(lp + 1);
                    }
                    const int ec[kNbrChunk] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
#pragma unroll
                    for (int u = 0; u < kNbrChunk; u++) e[c * kNbrChunk + u] = ec[u];
                }
                const int m = nn - k0;
#pragma unroll
                for (int g0 = 0; g0 < kNbrChunk * CH; g0 += E) {
                    if (g0 >= m) break;
                    float4 ha[GP], hb[GP];
                    float2 hc[GP];
#pragma unroll
                    for (int u = 0; u < GP; u++) {
                        ha[u] = rec32_half(A.fm, (uint32_t)e[g0 + 2 * u], par);
                        hb[u] = rec32_half(A.fm, (uint32_t)e[g0 + 2 * u + 1], par ^ 1u);
                        if (KIND == 0) hc[u] = gload(A.fmC + e[g0 + 2 * u + (int)par]);
                    }
                    __builtin_amdgcn_sched_barrier(0);
                    auto own = [&](int w) {
                        const float4 o = par ? hb[w] : ha[w], sn = par ? ha[w] : hb[w];
                        Rec r;
                        r.a = o;
                        r.b = make_float4(pair_swap(sn.x), pair_swap(sn.y), pair_swap(sn.z), pair_swap(sn.w));
                        if (KIND == 0) r.c = hc[w];
                        return r;
                    };
                    // quads of entries (a quad never straddles a chunk)
#pragma unroll
                    for (int u = 0; u < GP; u += 2) {
                        if (g0 + 2 * u >= m) break;
                        const Term t1 = term(own(u)), t2 = term(own(u + 1));
                        float q4[D];
#pragma unroll
                        for (int d = 0; d < D; d++) q4[d] = (t1.t[d] + pair_swap(t1.t[d])) + (t2.t[d] + pair_swap(t2.t[d]));
                        if constexpr (D == 3) {
                            // the quad sums are the same in both lanes: lane 2t
                            // accumulates x, lane 2t + 1 z (both in slot 0), both y
                            two_sum_acc(as[0], ac[0], par ? q4[2] : q4[0]);
                            two_sum_acc(as[1], ac[1], q4[1]);
                        } else {
#pragma unroll
                            for (int d = 0; d < D; d++) two_sum_acc(as[d], ac[d], q4[d]);
                        }
                    }
                }
            }
            if constexpr (D == 3) {
                as[2] = pair_swap(as[0]);
                ac[2] = pair_swap(ac[0]);
            }
        }
#pragma unroll
        for (int d = 0; d < D; d++) ai[d] = (double)as[d] + (double)ac[d];
        if (P.st) {
            const double fst = -mi.sigma / rho_i64 * A.kappa[i];
#pragma unroll
            for (int d = 0; d < D; d++) ai[d] = ai[d] + fst * A.nrm[d][i];
        }
    }
    if (PAIRS && par) return;          // both lanes hold the sums; one stores them
    if (fuse) {
        // k_integrate rewrites a for the computable particles only
        if (!(fi & kActive))
#pragma unroll
            for (int d = 0; d < D; d++) A.a[d][i] = ai[d];
        integrate_particle<D>(P, A, ctl, i, fi, I_BC | I_VEL | I_MOVE | I_RHO, ai, lo, hi);
        return;
    }
#pragma unroll
    for (int d = 0; d < D; d++) A.a[d][i] = ai[d];
}

// the wave's bounding box of moved positions into reduction slot
// (w % kBboxSlots), one atomic per bound (every lane of the wave takes part)
template <int D>
__device__ __forceinline__ void bbox_wave(const Arrays &A, int w, double (&lo)[D], double (&hi)[D])
{
#pragma unroll
    for (int d = 0; d < D; d++) {
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            lo[d] = fmin(lo[d], __shfl_xor(lo[d], o, 64));
            hi[d] = fmax(hi[d], __shfl_xor(hi[d], o, 64));
        }
    }
    if ((threadIdx.x & 63) == 0) {
        unsigned long long *slot = A.bbox + (size_t)(w % kBboxSlots) * kBboxStride;
#pragma unroll
        for (int d = 0; d < D; d++) {
            atomicMin(&slot[d], dkey(lo[d]));
            atomicMax(&slot[3 + d], dkey(hi[d]));
        }
    }
}

template <int D, int G, bool ONEMAT> __global__ __launch_bounds__(strain32_threads<D>()) void k_strain32(Params P, Arrays A)
{
    const int nb = (P.N + kBlock - 1) / kBlock, b = xcd_block(nb);
    if (b >= nb) return;
    if constexpr (strain32_pairs<D>()) strain32_pair_body<D, G, ONEMAT>(P, A, b);
    else strain32_body<D, G, ONEMAT>(P, A, b);
}
template <int D, bool ONEMAT>
__global__ __launch_bounds__(strain32_threads<D>()) void kb_strain32(const Params *__restrict__ Pt, const Arrays *__restrict__ At,
                                                                     const int2 *map)
{
    const int2 e = map[blockIdx.x];
    if constexpr (strain32_pairs<D>()) strain32_pair_body<D, kStrain32Batch, ONEMAT>(Pt[e.x], At[e.x], e.y);
    else strain32_body<D, kStrain32Batch, ONEMAT>(Pt[e.x], At[e.x], e.y);
}
template <int D, int G, int KIND, bool ONEMAT>
__global__ __launch_bounds__(momentum32_threads<D>()) void k_momentum32(Params P, Arrays A, int fuse)
{
    const int nb = (P.N + kBlock - 1) / kBlock, b = xcd_block(nb);
    if (b >= nb) return;
    double lo[D], hi[D];
#pragma unroll
    for (int d = 0; d < D; d++) {
        lo[d] = 1.0e20;
        hi[d] = -1.0e20;
    }
    momentum32_body<D, G, KIND, ONEMAT, momentum32_pairs<D>()>(P, A, b, fuse, lo, hi);
    if (fuse) bbox_wave<D>(A, (int)(blockIdx.x * (momentum32_threads<D>() / 64) + (threadIdx.x >> 6)), lo, hi);
}
template <int D, int KIND, bool ONEMAT>
__global__ __launch_bounds__(momentum32_threads<D>()) void kb_momentum32(const Params *__restrict__ Pt, const Arrays *__restrict__ At,
                                                                      const int2 *map)
{
    const int2 e = map[blockIdx.x];
    double lo[D], hi[D];
    momentum32_body<D, kMomentum32Batch, KIND, ONEMAT, momentum32_pairs<D>()>(Pt[e.x], At[e.x], e.y, 0, lo, hi);
}

template <int D>
static void batch_strain32_d(const BatchView &B, hipStream_t st)
{
    for (int v = 0; v < kSweep32Variants; v++) {
        if (!B.nvblocks[v]) continue;
        const dim3 grid(B.nvblocks[v]), blk(strain32_threads<D>());
        if (v == 0 || v == 1 || v == 3)     // one material
            hipLaunchKernelGGL((kb_strain32<D, true>), grid, blk, 0, st, B.P, B.A2, B.vmap[v]);
        else
            hipLaunchKernelGGL((kb_strain32<D, false>), grid, blk, 0, st, B.P, B.A2, B.vmap[v]);
    }
}

template <int D>
static void batch_momentum32_d(const BatchView &B, hipStream_t st)
{
    for (int v = 0; v < kSweep32Variants; v++) {
        if (!B.nvblocks[v]) continue;
        const dim3 grid(B.nvblocks[v]), blk(momentum32_threads<D>());
        switch (v) {
        case 0: hipLaunchKernelGGL((kb_momentum32<D, 2, true>), grid, blk, 0, st, B.P, B.A2, B.vmap[v]); break;
        case 1: hipLaunchKernelGGL((kb_momentum32<D, 1, true>), grid, blk, 0, st, B.P, B.A2, B.vmap[v]); break;
        case 2: hipLaunchKernelGGL((kb_momentum32<D, 1, false>), grid, blk, 0, st, B.P, B.A2, B.vmap[v]); break;
        case 3: hipLaunchKernelGGL((kb_momentum32<D, 0, true>), grid, blk, 0, st, B.P, B.A2, B.vmap[v]); break;
        default: hipLaunchKernelGGL((kb_momentum32<D, 0, false>), grid, blk, 0, st, B.P, B.A2, B.vmap[v]); break;
        }
    }
}

// one launch per record variant present in the batch (sweep32_variant)
void launch_batch_strain32(const BatchView &B, hipStream_t st)
{
    switch (B.ndim) {
    case 1: batch_strain32_d<1>(B, st); break;
    case 2: batch_strain32_d<2>(B, st); break;
    default: batch_strain32_d<3>(B, st); break;
    }
}

void launch_batch_momentum32(const BatchView &B, hipStream_t st)
{
    switch (B.ndim) {
    case 1: batch_momentum32_d<1>(B, st); break;
    case 2: batch_momentum32_d<2>(B, st); break;
    default: batch_momentum32_d<3>(B, st); break;
    }
}

static inline int nblk32(int n) { return (n + kBlock - 1) / kBlock; }

template <int G, bool ONEMAT>
static void strain32_gm(const Params &P, const Arrays &A, hipStream_t st)
{
    const dim3 grid(xcd_grid(nblk32(P.N)));
    switch (P.ndim) {
    case 1: hipLaunchKernelGGL((k_strain32<1, G, ONEMAT>), grid, dim3(strain32_threads<1>()), 0, st, P, A); break;
    case 2: hipLaunchKernelGGL((k_strain32<2, G, ONEMAT>), grid, dim3(strain32_threads<2>()), 0, st, P, A); break;
    default: hipLaunchKernelGGL((k_strain32<3, G, ONEMAT>), grid, dim3(strain32_threads<3>()), 0, st, P, A); break;
    }
}

// one-material scenes take the form with every material constant hoisted
template <int G>
static void strain32_g(const Params &P, const Arrays &A, hipStream_t st)
{
    if (P.nmat == 1) strain32_gm<G, true>(P, A, st);
    else strain32_gm<G, false>(P, A, st);
}

template <int G, int KIND, bool ONEMAT>
static void momentum32_gc(const Params &P, const Arrays &A, hipStream_t st, int fuse)
{
    const dim3 grid(xcd_grid(nblk32(P.N)));
    switch (P.ndim) {
    case 1: hipLaunchKernelGGL((k_momentum32<1, G, KIND, ONEMAT>), grid, dim3(momentum32_threads<1>()), 0, st, P, A, fuse); break;
    case 2: hipLaunchKernelGGL((k_momentum32<2, G, KIND, ONEMAT>), grid, dim3(momentum32_threads<2>()), 0, st, P, A, fuse); break;
    default: hipLaunchKernelGGL((k_momentum32<3, G, KIND, ONEMAT>), grid, dim3(momentum32_threads<3>()), 0, st, P, A, fuse); break;
    }
}

template <int G>
static void momentum32_g(const Params &P, const Arrays &A, hipStream_t st, int fuse)
{
    switch (P.mcompact) {
    case 2: momentum32_gc<G, 2, true>(P, A, st, fuse); break;
    case 1:
        if (P.nmat == 1) momentum32_gc<G, 1, true>(P, A, st, fuse);
        else momentum32_gc<G, 1, false>(P, A, st, fuse);
        break;
    default:
        if (P.nmat == 1) momentum32_gc<G, 0, true>(P, A, st, fuse);
        else momentum32_gc<G, 0, false>(P, A, st, fuse);
    }
}

void launch_strain32(const Params &P, const Arrays &A, hipStream_t st, int batch)
{
    switch (batch) {
    case 2: strain32_g<2>(P, A, st); break;
    case 4: strain32_g<4>(P, A, st); break;
    default: strain32_g<8>(P, A, st); break;
    }
}

void launch_momentum32(const Params &P, const Arrays &A, hipStream_t st, int batch, int fuse)
{
    // two lanes per particle pass over a whole chunk at least: batch 2 is batch 4
    if (P.ndim == 3 && momentum32_pairs<3>() && batch == 2) batch = 4;
    switch (batch) {
    case 2: momentum32_g<2>(P, A, st, fuse); break;
    case 8: momentum32_g<8>(P, A, st, fuse); break;
    default: momentum32_g<4>(P, A, st, fuse); break;
    }
}

}  // namespace sph
