// Wavefront-wide reductions (64 lanes) used by the setup, forward and adjoint
// kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace ocrt {

// Largest value / sum over one wavefront (every lane receives it).
__device__ inline double wave_max(double v)
{
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    return v;
}
__device__ inline double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// One halving exchange of a transposed wavefront reduction: lanes with
// (lane & off) == 0 keep the first H values, their partners the last H, and
// each adds what the partner sends for the half it keeps.
template <int H>
__device__ inline void halve_exchange(double *v, int off, bool upper)
{
#pragma unroll
    for (int i = 0; i < H; ++i) {
        const double give = upper ? v[i] : v[H + i];
        const double keep = upper ? v[H + i] : v[i];
        v[i] = keep + __shfl_xor(give, off);
    }
}

// Sums 8 C per-lane values over the wavefront in 7 C + 3 C exchanges
// (instead of 8 C x 6 for one reduction per value).  With the values ordered
// [8][C], every lane of the group of eight lanes l = 8 r .. 8 r + 7 ends with
// the totals of row r in v[0..C-1].
template <int C>
__device__ inline void wave_sum_8xC(double (&v)[8 * C], int lane)
{
    halve_exchange<4 * C>(v, 32, (lane & 32) != 0);
    halve_exchange<2 * C>(v, 16, (lane & 16) != 0);
    halve_exchange<C>(v, 8, (lane & 8) != 0);
#pragma unroll
    for (int off = 4; off > 0; off >>= 1)
#pragma unroll
        for (int i = 0; i < C; ++i) v[i] += __shfl_xor(v[i], off);
}

}  // namespace ocrt
