// sph_rec32.hpp -- compact FP32 gather records of the FP32-compensated arm.
//
// Positions are fixed-point integers: X = rint(x / q) modulo 2^32 in units of
// q = 2^Params::qexp, the power of two with cut / q in [2^22, 2^23) for the
// largest cut-off.  A pair separation is the 32-bit difference X_i - X_j
// (wrap-around included: any two points closer than 2^31 q, ~256 cut-offs,
// subtract correctly, whatever the size of the domain), exact in integers
// and exact again when converted to FP32, since neighbours lie within 2^24 q.
// The only rounding is the quantisation of each position, at most q / 2 =
// 2^-24 of the cut-off: the precision FP32 has for coordinates relative to a
// cell corner, without their cell bookkeeping.  The sweeps work in units of
// q (separations, r^2); the pair constants are scaled to match (PairConst32),
// the scalings being exact powers of two.
//
// Packed word: bits 0-3 material, bit 6 periodic image, bit 7 rigid.
#pragma once

#include "sph_kernels.hpp"

namespace sph {

constexpr uint32_t kRecMat = 0x0Fu, kRecGhost = 0x40u, kRecRigid = 0x80u;
static_assert(kMaxMat < 16, "material fits the packed word");

// X of one coordinate (qinv = 2^-qexp: the product is exact), its bits as a float
__device__ __forceinline__ float fixq(double x, double qinv)
{
    const long long k = (long long)rint(x * qinv);
    return __uint_as_float((uint32_t)(unsigned long long)k);
}

// separation component in units of q from two fixed-point coordinates
__device__ __forceinline__ float qsep(float xi, float xj)
{
    return (float)(int)(__float_as_uint(xi) - __float_as_uint(xj));
}

__device__ __forceinline__ uint32_t pack32(uint32_t info)
{
    return (info & kRecMat) | ((info & kRigid) ? kRecRigid : 0u) | ((info & kGhost) ? kRecGhost : 0u);
}

__device__ __forceinline__ int rec_mat(uint32_t pk) { return (int)(pk & kRecMat); }

// inverse of the cell-code deposit: the cell coordinate held in the bits m of v
__device__ __forceinline__ int code_ext(int v, unsigned int m)
{
    if (m == 0u) return 0;
    if (((m + (m & (0u - m))) & m) == 0u)                       // one run of bits (row-major order)
        return (int)(((unsigned int)v & m) >> __builtin_ctz(m));
    int r = 0, b = 0;
    while (m) {
        const unsigned int low = m & (0u - m);
        if ((unsigned)v & low) r |= 1 << b;
        b++;
        m ^= low;
    }
    return r;
}

// The two 16-byte halves of a record side by side, one 32-byte sector per
// particle (entries 2j and 2j + 1)
template <int PART>
__device__ __forceinline__ size_t rec32_index(int j)
{
    return 2 * (size_t)j + PART;
}

// continuity record: {X[0..2], vol}, {v[0..2], packed}
template <int D>
__device__ __forceinline__ void put_srec32(const Params &P, const Arrays &A, int j, const double *xs, double vol,
                                           const double *v, uint32_t info)
{
    float4 a = make_float4(0.f, 0.f, 0.f, (float)vol), b = make_float4(0.f, 0.f, 0.f, 0.f);
    a.x = fixq(xs[0], P.qinv);
    b.x = (float)v[0];
    if (D > 1) { a.y = fixq(xs[1], P.qinv); b.y = (float)v[1]; }
    if (D > 2) { a.z = fixq(xs[2], P.qinv); b.z = (float)v[2]; }
    b.w = __uint_as_float(pack32(info));
    A.fs[rec32_index<0>(j)] = a;
    A.fs[rec32_index<1>(j)] = b;
}

// momentum record, by Params::mcompact:
//   0: {X[0..2], m}, {v[0..2], rho}, fmC = {p/rho^2, packed} (40 bytes)
//   1: {X[0..2], rho - rho0}, {v[0..2], packed} (32 bytes; one mass per
//      material and the linear EOS: m and p/rho^2 follow from the material)
//   2: {X[0..2], 1/rho}, {v[0..2], p/rho^2} (32 bytes; one material, no rigid
//      body: nothing else differs between particles; images carry p = 0)
template <int D>
__device__ __forceinline__ void put_mrec32(const Params &P, const Arrays &A, int j, const double *x, double m,
                                           const double *v, double rho, double prho2, uint32_t info)
{
    float4 a = make_float4(0.f, 0.f, 0.f, (float)m), b = make_float4(0.f, 0.f, 0.f, (float)rho);
    a.x = fixq(x[0], P.qinv);
    b.x = (float)v[0];
    if (D > 1) { a.y = fixq(x[1], P.qinv); b.y = (float)v[1]; }
    if (D > 2) { a.z = fixq(x[2], P.qinv); b.z = (float)v[2]; }
    if (P.mcompact == 2) {
        a.w = (float)(1.0 / rho);
        b.w = (float)prho2;
        A.fm[rec32_index<0>(j)] = a;
        A.fm[rec32_index<1>(j)] = b;
        return;
    }
    if (P.mcompact == 1) {
        a.w = (float)(rho - P.mat[info & kMatMask].rho0);
        b.w = __uint_as_float(pack32(info));
        A.fm[rec32_index<0>(j)] = a;
        A.fm[rec32_index<1>(j)] = b;
        return;
    }
    A.fm[rec32_index<0>(j)] = a;
    A.fm[rec32_index<1>(j)] = b;
    A.fmC[j] = make_float2((float)prho2, __uint_as_float(pack32(info)));
}

// Half PART (0 or 1) of particle j's record: a 32-bit byte offset from the
// wave-uniform base of the half, so the load takes the base in scalar
// registers and one vector register of offset (the arm limits scenes to 2^27
// slots).
template <int PART>
__device__ __forceinline__ float4 rec32_at(const float4 *base, uint32_t j)
{
    return gload(reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(base) + ((j << 5) + PART * 16u)));
}

}  // namespace sph
