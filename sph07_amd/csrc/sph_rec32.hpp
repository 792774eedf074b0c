// sph_rec32.hpp -- compact FP32 gather records of the FP32-compensated arm.
//
// Positions are stored relative to a cell corner, so that a pair separation
// is rebuilt without cancellation of large coordinates (FP32 keeps them to
// ~1e-7 of the cut-off).  origin(c) = gmin + c * cs32 with cs32 the cell size
// rounded to FP32; neighbours lie less than one cell size apart on every axis.
//   * continuity records: r = x - origin(c) of the particle's bin cell c; the
//     packed word holds the low two bits of each cell coordinate, and the
//     separation is off * cs32 + (r_i - r_j) with off in {-1, 0, 1} from them
//     (off * cs32 is exact);
//   * momentum records: R = x - origin(c & ~1), the corner of the even cell at
//     or below c (0 <= R < 2 cs32); even-cell origins of neighbours differ by
//     0 or +-2 cs32 exactly, so the separation is R_i - R_j - 2 cs32 rint((R_i -
//     R_j) / (2 cs32)) (the subtraction of +-2 cs32 is exact, Sterbenz).  This
//     form needs no cell bits, which frees the packed word of the one-material
//     record for p/rho^2.
//
// Packed word: bits 0-3 material, bit 6 periodic image, bit 7 rigid, bytes
// 1-3 (continuity records) the low two bits of the cell coordinates, one byte
// per axis so that the three offsets come out of one byte-lane subtraction.
#pragma once

#include "sph_state.hpp"

namespace sph {

constexpr uint32_t kRecMat = 0x0Fu, kRecGhost = 0x40u, kRecRigid = 0x80u;
static_assert(kMaxMat < 16, "material fits the packed word");

__device__ __forceinline__ float cell_size32(const Ctl *ctl, int d) { return (float)(1.0 / ctl->gsize[d]); }

__device__ __forceinline__ double cell_origin(const Ctl *ctl, int d, int c)
{
    return ctl->gmin[d] + (double)c * (double)cell_size32(ctl, d);
}

// R of one coordinate: relative to the origin of the even cell at or below c
__device__ __forceinline__ float rel32(const Ctl *ctl, int d, int c, double x)
{
    return (float)(x - cell_origin(ctl, d, c & ~1));
}

__device__ __forceinline__ uint32_t pack32(uint32_t info)
{
    return (info & kRecMat) | ((info & kRigid) ? kRecRigid : 0u) | ((info & kGhost) ? kRecGhost : 0u);
}

// with the low two bits of the cell coordinates (continuity records)
__device__ __forceinline__ uint32_t pack32c(uint32_t info, const int *bc)
{
    return pack32(info) | ((uint32_t)(bc[0] & 3) << 8) | ((uint32_t)(bc[1] & 3) << 16) |
           ((uint32_t)(bc[2] & 3) << 24);
}

// Cell offsets (-1, 0, 1) of particle a relative to particle b on the three
// axes: per byte (4 + ca - cb + 1) mod 4 = offset + 1, no borrow across bytes.
__device__ __forceinline__ void cell_offs(uint32_t pa, uint32_t pb, float off[3])
{
    const uint32_t x = (((pa | 0x04040400u) - (pb & 0x03030300u)) + 0x01010100u) & 0x03030300u;
    off[0] = (float)((x >> 8) & 0xFFu) - 1.0f;
    off[1] = (float)((x >> 16) & 0xFFu) - 1.0f;
    off[2] = (float)(x >> 24) - 1.0f;
}

__device__ __forceinline__ int rec_mat(uint32_t pk) { return (int)(pk & kRecMat); }

// Separation component from two R values (see the header): tcs = 2 cs32,
// itcs = 1 / tcs
__device__ __forceinline__ float sep32(float ri, float rj, float tcs, float itcs)
{
    const float raw = ri - rj;
    return raw - tcs * __builtin_rintf(raw * itcs);
}

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

// continuity record: {r[0..2], vol}, {v[0..2], packed with the cell bits}
template <int D>
__device__ __forceinline__ void put_srec32(const Arrays &A, const Ctl *ctl, int j, const double *xs, double vol,
                                           const double *v, uint32_t info, const int *bc)
{
    float4 a = make_float4(0.f, 0.f, 0.f, (float)vol), b = make_float4(0.f, 0.f, 0.f, 0.f);
    a.x = (float)(xs[0] - cell_origin(ctl, 0, bc[0]));
    b.x = (float)v[0];
    if (D > 1) { a.y = (float)(xs[1] - cell_origin(ctl, 1, bc[1])); b.y = (float)v[1]; }
    if (D > 2) { a.z = (float)(xs[2] - cell_origin(ctl, 2, bc[2])); b.z = (float)v[2]; }
    b.w = __uint_as_float(pack32c(info, bc));
    A.fs[2 * (size_t)j] = a;
    A.fs[2 * (size_t)j + 1] = b;
}

// momentum record, by Params::mcompact:
//   0: {R[0..2], m}, {v[0..2], rho}, fmC = {p/rho^2, packed} (40 bytes)
//   1: {R[0..2], rho - rho0}, {v[0..2], packed} (32 bytes; one mass per
//      material and the linear EOS: m and p/rho^2 follow from the material)
//   2: {R[0..2], 1/rho}, {v[0..2], p/rho^2} (32 bytes; one material, no rigid
//      body: nothing else differs between particles; images carry p = 0)
template <int D>
__device__ __forceinline__ void put_mrec32(const Params &P, const Arrays &A, const Ctl *ctl, int j, const double *x,
                                           double m, const double *v, double rho, double prho2, uint32_t info,
                                           const int *bc)
{
    float4 a = make_float4(0.f, 0.f, 0.f, (float)m), b = make_float4(0.f, 0.f, 0.f, (float)rho);
    a.x = rel32(ctl, 0, bc[0], x[0]);
    b.x = (float)v[0];
    if (D > 1) { a.y = rel32(ctl, 1, bc[1], x[1]); b.y = (float)v[1]; }
    if (D > 2) { a.z = rel32(ctl, 2, bc[2], x[2]); b.z = (float)v[2]; }
    if (P.mcompact == 2) {
        a.w = (float)(1.0 / rho);
        b.w = (float)prho2;
        A.fm[2 * (size_t)j] = a;
        A.fm[2 * (size_t)j + 1] = b;
        return;
    }
    if (P.mcompact == 1) {
        a.w = (float)(rho - P.mat[info & kMatMask].rho0);
        b.w = __uint_as_float(pack32(info));
        A.fm[2 * (size_t)j] = a;
        A.fm[2 * (size_t)j + 1] = b;
        return;
    }
    A.fm[2 * (size_t)j] = a;
    A.fm[2 * (size_t)j + 1] = b;
    A.fmC[j] = make_float2((float)prho2, __uint_as_float(pack32(info)));
}

// Half PART (0 or 1) of particle j's 32-byte record: a 32-bit byte offset from
// the wave-uniform base, so the load takes the base in scalar registers and
// one vector register of offset (the arm limits scenes to 2^27 slots).
template <int PART>
__device__ __forceinline__ float4 rec32_at(const float4 *base, uint32_t j)
{
    return *reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(base) + ((j << 5) + PART * 16u));
}

}  // namespace sph
