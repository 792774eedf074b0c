// sph_rec32.hpp -- compact FP32 gather records of the FP32-compensated arm.
//
// Positions are stored relative to the lower corner of the particle's bin
// cell, origin(c) = gmin + c * cs32, where cs32 is the cell size rounded to
// FP32.  Two adjacent cells' origins then differ by exactly +-cs32, so a pair
// separation is rebuilt as off * cs32 + (r_i - r_j) with off in {-1, 0, 1}
// taken from the low two bits of each cell coordinate (kept in the packed
// word).  Relative positions are below one cell size, so FP32 keeps them to
// ~1e-7 of the cut-off.
//
// Packed word: bits 0-3 material, bit 7 rigid, bytes 1-3 the low two bits of
// the cell coordinates (one byte per axis, so that the three offsets come out
// of one byte-lane subtraction).
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

__device__ __forceinline__ uint32_t pack32(uint32_t info, const int *bc)
{
    return (info & kRecMat) | ((info & kRigid) ? kRecRigid : 0u) | ((info & kGhost) ? kRecGhost : 0u) |
           ((uint32_t)(bc[0] & 3) << 8) |
           ((uint32_t)(bc[1] & 3) << 16) | ((uint32_t)(bc[2] & 3) << 24);
}

__device__ __forceinline__ int rec_mat(uint32_t pk) { return (int)(pk & kRecMat); }

// Cell offsets (-1, 0, 1) of particle a relative to particle b on the three
// axes: per byte (4 + ca - cb + 1) mod 4 = offset + 1, no borrow across bytes.
__device__ __forceinline__ void cell_offs(uint32_t pa, uint32_t pb, float off[3])
{
    const uint32_t x = (((pa | 0x04040400u) - (pb & 0x03030300u)) + 0x01010100u) & 0x03030300u;
    off[0] = (float)((x >> 8) & 0xFFu) - 1.0f;
    off[1] = (float)((x >> 16) & 0xFFu) - 1.0f;
    off[2] = (float)(x >> 24) - 1.0f;
}

// inverse of the cell-code deposit: the cell coordinate held in the bits m of v
__device__ __forceinline__ int code_ext(int v, unsigned int m)
{
    int r = 0, b = 0;
    while (m) {
        const unsigned int low = m & (0u - m);
        if ((unsigned)v & low) r |= 1 << b;
        b++;
        m ^= low;
    }
    return r;
}

// continuity record: {r[0..2], vol}, {v[0..2], packed}
template <int D>
__device__ __forceinline__ void put_srec32(const Arrays &A, const Ctl *ctl, int j, const double *xs, double vol,
                                           const double *v, uint32_t info, const int *bc)
{
    float4 a = make_float4(0.f, 0.f, 0.f, (float)vol), b = make_float4(0.f, 0.f, 0.f, 0.f);
    a.x = (float)(xs[0] - cell_origin(ctl, 0, bc[0]));
    b.x = (float)v[0];
    if (D > 1) { a.y = (float)(xs[1] - cell_origin(ctl, 1, bc[1])); b.y = (float)v[1]; }
    if (D > 2) { a.z = (float)(xs[2] - cell_origin(ctl, 2, bc[2])); b.z = (float)v[2]; }
    b.w = __uint_as_float(pack32(info, bc));
    A.fs[2 * (size_t)j] = a;
    A.fs[2 * (size_t)j + 1] = b;
}

// momentum record: {r[0..2], m}, {v[0..2], rho}, {p/rho^2, packed}; with
// Params::mcompact {r[0..2], rho - rho0}, {v[0..2], packed} (32 bytes)
template <int D>
__device__ __forceinline__ void put_mrec32(const Params &P, const Arrays &A, const Ctl *ctl, int j, const double *x,
                                           double m, const double *v, double rho, double prho2, uint32_t info,
                                           const int *bc)
{
    float4 a = make_float4(0.f, 0.f, 0.f, (float)m), b = make_float4(0.f, 0.f, 0.f, (float)rho);
    a.x = (float)(x[0] - cell_origin(ctl, 0, bc[0]));
    b.x = (float)v[0];
    if (D > 1) { a.y = (float)(x[1] - cell_origin(ctl, 1, bc[1])); b.y = (float)v[1]; }
    if (D > 2) { a.z = (float)(x[2] - cell_origin(ctl, 2, bc[2])); b.z = (float)v[2]; }
    if (P.mcompact) {
        a.w = (float)(rho - P.mat[info & kMatMask].rho0);
        b.w = __uint_as_float(pack32(info, bc));
        A.fm[2 * (size_t)j] = a;
        A.fm[2 * (size_t)j + 1] = b;
        return;
    }
    A.fm[2 * (size_t)j] = a;
    A.fm[2 * (size_t)j + 1] = b;
    A.fmC[j] = make_float2((float)prho2, __uint_as_float(pack32(info, bc)));
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
