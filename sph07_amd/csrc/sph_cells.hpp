// sph_cells.hpp -- cell-centric building blocks of the neighbour search: wave
// scans, cell codes and the staging of a cell's stencil candidates in LDS.
#pragma once

#include "sph_kernels.hpp"
#include "sph_rec32.hpp"

namespace sph {

__device__ __forceinline__ int wave_incl_scan(int v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// deposits the low bits of v on the set bits of m (cell coordinate -> code bits)
__device__ __forceinline__ int code_dep(int v, unsigned int m)
{
    if (m == 0u) return 0;
    if (((m + (m & (0u - m))) & m) == 0u)                       // one run of bits (row-major order)
        return (int)(((unsigned int)v << __builtin_ctz(m)) & m);
    int r = 0;
    while (m) {
        const unsigned int low = m & (0u - m);
        if (v & 1) r |= (int)low;
        v >>= 1;
        m ^= low;
    }
    return r;
}

// Code bits of axis m of the cell with code q stepped by -1, 0, +1 along
// that axis (Morton increment and decrement on the axis's bits; the caller
// keeps the coordinate inside the grid).
__device__ __forceinline__ unsigned int code_step(unsigned int q, unsigned int m, int step)
{
    const unsigned int c = q & m;
    const unsigned int dn = (c - 1u) & m, up = ((c | ~m) + 1u) & m;
    return step < 0 ? dn : (step > 0 ? up : c);
}

// Stages the candidate slots of the stencil of the cell with code q (at
// coordinates bc) in stage[] in the traversal order of the reference's walk:
// cells x -> y -> z, in each cell its periodic images newest first, then its
// slots.  One lane per stencil cell loads the cell's range in a single round
// of loads; a wave scan gives each cell's offset.  Returns the candidate
// count; above kCandCap nothing is staged.  own: staged index of the cell's
// own first slot.
template <int D>
__device__ __forceinline__ int stage_cell(const Params &P, const Arrays &A, const Ctl *ctl, int q,
                                          const int (&bc)[3], int lane, int *stage, int &own)
{
    const unsigned int m0 = ctl->cmask[0], m1 = ctl->cmask[1], m2 = ctl->cmask[2];
    const int gl1 = ctl->glim[1], gl2 = ctl->glim[2];
    int sr = 0, ss = 0, sz = 0;
    bool inst;
    if (D == 3) {
        inst = lane < 27;
        sr = lane / 9 - 1; ss = (lane / 3) % 3 - 1; sz = lane % 3 - 1;
    } else if (D == 2) {
        inst = lane < 9;
        sr = lane / 3 - 1; ss = lane % 3 - 1;
    } else {
        inst = lane < 3;
        sr = lane - 1;
    }
    const int cr = bc[0] + sr, cs = bc[1] + ss, cz = bc[2] + sz;
    inst = inst && cr >= 0 && cr < ctl->glim[0] && cs >= 0 && cs < gl1 && cz >= 0 && cz < gl2;
    int ng = 0, g0 = 0, s0 = 0, len = 0;
    if (inst) {
        const unsigned int qu = (unsigned int)q;
        const int qc = (int)(code_step(qu, m0, sr) | code_step(qu, m1, ss) | code_step(qu, m2, sz));
        s0 = A.cell_start[qc];
        len = A.cell_start[qc + 1] - s0;
        if (P.periodic) {
            const int c = gl2 * (gl1 * cr + cs) + cz;
            ng = A.gcount[c];
            if (ng) g0 = A.gfirst[c];
        }
    }
    const int cnt = ng + len;
    const int incl = wave_incl_scan(cnt);
    const int off = incl - cnt;
    const int ncand = __builtin_amdgcn_readlane(incl, 63);
    own = __builtin_amdgcn_readlane(off + ng, D == 3 ? 13 : (D == 2 ? 4 : 1));
    if (ncand > kCandCap) return ncand;
    for (int u = 0; u < ng; u++) stage[off + u] = P.N + g0 + ng - 1 - u;
    for (int u = 0; u < len; u++) stage[off + ng + u] = s0 + u;
    // staged slots are read back by other lanes of the same wave (LDS
    // operations of one wave complete in order)
    __builtin_amdgcn_wave_barrier();
    return ncand;
}

}  // namespace sph
