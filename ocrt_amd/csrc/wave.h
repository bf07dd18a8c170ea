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

// Cross-lane moves of a double as two 32-bit halves on the VALU (no LDS):
// gfx950's v_permlane32_swap / v_permlane16_swap and DPP row operations.
typedef unsigned lanes_u2 __attribute__((ext_vector_type(2)));

__device__ inline unsigned lo32(double v) { return (unsigned)__double_as_longlong(v); }
__device__ inline unsigned hi32(double v) { return (unsigned)((unsigned long long)__double_as_longlong(v) >> 32); }
__device__ inline double join32(unsigned hi, unsigned lo)
{
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// permlane swaps of the pair (x, y): with OFF = 32, x becomes
// [x lanes 0-31 | y lanes 0-31] and y becomes [x lanes 32-63 | y lanes 32-63];
// with OFF = 16 the same between rows 0/1 and 2/3 of 16 lanes.
template <int OFF> __device__ inline void lanes_swap(double &x, double &y)
{
    lanes_u2 l, h;
    if constexpr (OFF == 32) {
        l = __builtin_amdgcn_permlane32_swap(lo32(x), lo32(y), false, false);
        h = __builtin_amdgcn_permlane32_swap(hi32(x), hi32(y), false, false);
    } else {
        l = __builtin_amdgcn_permlane16_swap(lo32(x), lo32(y), false, false);
        h = __builtin_amdgcn_permlane16_swap(hi32(x), hi32(y), false, false);
    }
    x = join32(h.x, l.x);
    y = join32(h.y, l.y);
}

// DPP move within rows: 0x128 row_ror:8 (lane ^ 8), 0xB1 quad_perm 1,0,3,2
// (lane ^ 1), 0x4E quad_perm 2,3,0,1 (lane ^ 2), 0x141 row_half_mirror
// (lane ^ 7 within eight lanes).
template <int CTRL> __device__ inline double lanes_dpp(double v)
{
    const unsigned l = (unsigned)__builtin_amdgcn_update_dpp(0, (int)lo32(v), CTRL, 0xF, 0xF, false);
    const unsigned h = (unsigned)__builtin_amdgcn_update_dpp(0, (int)hi32(v), CTRL, 0xF, 0xF, false);
    return join32(h, l);
}

// The same moves for a float: one 32-bit lane value.
template <int OFF> __device__ inline void lanes_swap(float &x, float &y)
{
    lanes_u2 r;
    if constexpr (OFF == 32)
        r = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(y), false, false);
    else
        r = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(y), false, false);
    x = __uint_as_float(r.x);
    y = __uint_as_float(r.y);
}

template <int CTRL> __device__ inline float lanes_dpp(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// One halving exchange of a transposed wavefront reduction: lanes with
// (lane & off) == 0 keep the first H values, their partners the last H, and
// each adds what the partner sends for the half it keeps.  After a permlane
// swap of (v[i], v[H + i]) both halves are in place: every lane adds the two.
template <int H, int OFF, typename T>
__device__ inline void halve_exchange(T *v, bool upper)
{
#pragma unroll
    for (int i = 0; i < H; ++i) {
        if constexpr (OFF >= 16) {
            T x = v[i], y = v[H + i];
            lanes_swap<OFF>(x, y);
            v[i] = x + y;
        } else {
            const T give = upper ? v[i] : v[H + i];
            const T keep = upper ? v[H + i] : v[i];
            v[i] = keep + lanes_dpp<0x128>(give);   // OFF = 8
        }
    }
}

// Sums 8 C per-lane values over the wavefront in 7 C + 3 C exchanges
// (instead of 8 C x 6 for one reduction per value), all on the VALU.  With
// the values ordered [8][C], every lane of the group of eight lanes
// l = 8 r .. 8 r + 7 ends with the totals of row r in v[0..C-1].
template <int C, typename T>
__device__ inline void wave_sum_8xC(T (&v)[8 * C], int lane)
{
    halve_exchange<4 * C, 32>(v, (lane & 32) != 0);
    halve_exchange<2 * C, 16>(v, (lane & 16) != 0);
    halve_exchange<C, 8>(v, (lane & 8) != 0);
#pragma unroll
    for (int i = 0; i < C; ++i) {
        v[i] += lanes_dpp<0xB1>(v[i]);    // pairs
        v[i] += lanes_dpp<0x4E>(v[i]);    // quads
        v[i] += lanes_dpp<0x141>(v[i]);   // the other quad of the eight
    }
}

// The same reduction with LDS-routed exchanges (ds_bpermute through
// __shfl_xor).  The adjoint's per-level reductions keep this form: measured
// against the VALU form on one box, the adjoint ran at 68.6 against 71.1 ms
// (256 instances), while the Cox-Munk kernels gain from the VALU form.
template <int H>
__device__ inline void halve_exchange_lds(double *v, int off, bool upper)
{
#pragma unroll
    for (int i = 0; i < H; ++i) {
        const double give = upper ? v[i] : v[H + i];
        const double keep = upper ? v[H + i] : v[i];
        v[i] = keep + __shfl_xor(give, off);
    }
}

template <int C>
__device__ inline void wave_sum_8xC_lds(double (&v)[8 * C], int lane)
{
    halve_exchange_lds<4 * C>(v, 32, (lane & 32) != 0);
    halve_exchange_lds<2 * C>(v, 16, (lane & 16) != 0);
    halve_exchange_lds<C>(v, 8, (lane & 8) != 0);
#pragma unroll
    for (int off = 4; off > 0; off >>= 1)
#pragma unroll
        for (int i = 0; i < C; ++i) v[i] += __shfl_xor(v[i], off);
}

}  // namespace ocrt
