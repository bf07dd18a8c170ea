// Order loop of the SOS solve.
//   order_one      : single scattering and first-order surface fields, fp64
//                    (rt_primary_source.c:17-198, rt_layer_transport.c:107-208,
//                    rt_surface_boundary.c:210-527)
//   order_status   : per-item convergence test of the previous order and the
//                    lower-boundary reflection of the current field
//                    (rt_solver.c:3053-3129, rt_surface_boundary.c:137-208)
//   gemm_f64 / gemm_bf16 : scattering integral J = A . diag-mixed(F) on MFMA
//                    (rt_sos_operator.c:206-424 as a dense product)
//   sweep          : formal solution of every (Stokes, direction) column, one
//                    thread per column scanning the levels, fused with the
//                    order sum and its max-norms
//   extract / finalize : view sampling, Fourier reconstruction, reflectance,
//                    IPSS scaling, direct glint and transmittances
//                    (rt_solver.c:1010-1276, rt_fourier.c:9-33, surface.c:1813-1907)
#include <type_traits>

#include "ocrt_amd.h"
#include "sos_common.h"

namespace ocrt {

// Order sum of the mixed form: the two output rows (k = 0, k = nt) in fp64
// plus an fp32 Neumaier term; the interior rows in fp32.  The interior sum
// feeds only the convergence test's max-norm over all levels (rt_solver.c:
// 1855-1868): fp32 rounding over even 150 orders moves that norm by about
// 1e-5 of itself, so the test decides alike except for a residual within that
// fraction of the tolerance.  4 bytes read and 4 written per element and
// order where an fp64 sum takes 16.

// Formal solution of one column with the reference's linear-source closed form
// (rt_layer_transport.c:89-99, 144-203), in fp64.  S(k) gives the source at
// level k; emit(k, value, tr, where) receives the field at level k, with tr
// the Beer-Lambert factor of the layer just crossed and where = 0 for the
// starting row, 1 for an interior level, 2 for the far row (a compile-time
// constant, so the interior levels carry no row tests).  The source's slope
// takes the layer's 1 / dtau from rdt (wavefront-uniform scalar loads) in
// place of an fp64 division per column and level.
template <typename SrcFn, typename EmitFn>
__device__ inline void sweep_column_f64(int hemi, int nt, double mu, double bc, const double *h,
                                        const double *rdt, const double *T, int n, int o, SrcFn S, EmitFn emit)
{
    using first_t = std::integral_constant<int, 0>;
    using inner_t = std::integral_constant<int, 1>;
    using far_t = std::integral_constant<int, 2>;
    const int dL = hemi ? 1 : -1;
    const int k_first = hemi ? 0 : nt;
    double run = hemi ? 0.0 : bc;
    emit(k_first, run, 1.0, first_t{});
    double s_in = S(k_first);
    double hu = h[k_first];   // depth of the upstream level
    auto step = [&](double s_out, double hk, double tr, double rd) {
        if (hemi == 0) {
            const double a = (s_in - s_out) * rd;
            const double b = s_out - a * hk;
            run = tr * run + ((1.0 - tr) * (b + a * mu) + a * (hk - hu * tr)) * 0.5;
        } else {
            const double a = (s_out - s_in) * rd;
            const double b = s_out - a * hk;
            run = tr * run + ((1.0 - tr) * (b - a * mu) + a * (hk - hu * tr)) * 0.5;
        }
        s_in = s_out;
        hu = hk;
    };
    // level i steps along the scan: k = k_first + dL i, crossing layer
    // hemi ? k - 1 : k.  Interior levels i = 1 .. nt - 1 in blocks of 4 whose
    // sources, depths and transmittances are gathered before the recurrence
    // runs over them, then one at a time, then the far row.
    constexpr int kB = 4;
    auto layer_of = [&](int k) { return hemi ? k - 1 : k; };
    int i = 1;
    for (; i + kB <= nt; i += kB) {
        double sv[kB], hk[kB], tr[kB], rd[kB];
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            const int k = k_first + dL * (i + u);
            sv[u] = S(k);
            hk[u] = h[k];
            tr[u] = T[(long long)layer_of(k) * n + o];
            rd[u] = rdt[layer_of(k)];
        }
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            step(sv[u], hk[u], tr[u], rd[u]);
            emit(k_first + dL * (i + u), run, tr[u], inner_t{});
        }
    }
    for (; i < nt; ++i) {
        const int k = k_first + dL * i;
        const double tr = T[(long long)layer_of(k) * n + o];
        step(S(k), h[k], tr, rdt[layer_of(k)]);
        emit(k, run, tr, inner_t{});
    }
    const int kl = k_first + dL * nt;
    const double tr = T[(long long)layer_of(kl) * n + o];
    step(S(kl), h[kl], tr, rdt[layer_of(kl)]);
    emit(kl, run, tr, far_t{});
}

// Field store of the mixed form: bf16, plus the bf16 residual v - hi for the
// items whose field is kept as a pair.
__device__ inline void store_field_pair(__hip_bfloat16 *hi, __hip_bfloat16 *lo, long long ix, double v)
{
    const __hip_bfloat16 h = __float2bfloat16((float)v);
    hi[ix] = h;
    if (lo) lo[ix] = __float2bfloat16((float)(v - (double)__bfloat162float(h)));
}

// ---------------------------------------------------------------------------
// Order 1: one thread per state column; each wavefront takes one 64-column
// tile of one item from the plan's tile list.
template <typename FT, bool MIXED>
__global__ __launch_bounds__(256) void order_one(const DevInst *__restrict__ inst,
                                                 const DevItem *__restrict__ items, Pools P,
                                                 ItemState *__restrict__ st, const ColTile *__restrict__ tiles,
                                                 int n_tiles, const double *__restrict__ lev,
                                                 const double *__restrict__ Tpool)
{
    const int tix = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (tix >= n_tiles) return;
    const ColTile ct = tiles[tix];
    // the tile is wavefront-uniform: its item, and so every per-level array
    // read of the scan, stays in scalar registers
    const int item = __builtin_amdgcn_readfirstlane(ct.item);
    if (item < 0) return;
    const DevItem &it = items[item];
    const DevInst &I = inst[it.inst];
    const int n = it.n, nt = it.nt, Pp = it.P, m = it.m;
    const int col = ct.col0 + lane;
    if (ct.col0 == 0 && lane == 0) {
        ItemState s;
        s.active = I.max_orders >= 2 ? 1 : 0;
        s.orders = 1;
        s.converged = I.max_orders >= 2 ? 0 : 1;
        s.status = 0;
        s.residual = 0.0;
        s.max_cur = 0ull;
        s.max_tot = 0ull;
        s.reduced = 0;
        s.rnext = 0;
        st[item] = s;
    }
    if (col >= ct.col_end) return;
    FT *F = reinterpret_cast<FT *>(P.F) + it.off_F;
    // the fp64 form keeps the order sum at every level; the mixed form only at
    // its two output rows (fp64 with an fp32 compensation term)
    double *tot = MIXED ? nullptr : P.tot + it.off_tot;
    if (MIXED) {
        P.comp[it.off_cmp + col] = 0.f;
        P.comp[it.off_cmp + Pp + col] = 0.f;
    }
    __hip_bfloat16 *Flo = (MIXED && it.fsplit) ? reinterpret_cast<__hip_bfloat16 *>(P.Flo) + it.off_Flo : nullptr;
    if (col >= 6 * n) {
        for (int k = 0; k <= nt; ++k) {
            F[(long long)k * Pp + col] = to_storage<FT>(0.0);
            if (Flo) Flo[(long long)k * Pp + col] = __float2bfloat16(0.f);
            if (!MIXED) tot[(long long)k * Pp + col] = 0.0;
        }
        return;
    }
    // level and Beer-Lambert tables through the __restrict__ arguments (P.lev,
    // P.T): the per-level reads of the scan stay scalar loads
    const LevelView lv(lev, I);
    int c, hemi, o;
    col_decode(col, n, c, hemi, o);
    hemi = __builtin_amdgcn_readfirstlane(hemi);   // column tiles are direction-pure
    const double *mu_ring = P.ring + I.off_ring;
    const double mu = mu_ring[o];
    const double *T = Tpool + I.off_T;
    const ColumnSource src(I, it, P, lv, col);
    // one pass: the sweep is linear in its source, so single scattering and
    // the flat-surface seed (the specularly reflected direct beam scattered
    // once, rt_surface_boundary.c:210-428) share it; over the rough ocean the
    // direct beam reflected into the upward nodes is added level by level
    // (rt_surface_boundary.c:430-493)
    // The seed's path attenuation exp(-(tau - h[k]) / mu) from the surface up
    // to level k is carried as the running product of the layers' Beer-Lambert
    // factors, which the upward scan visits in order (k = nt .. 0).
    // An external bottom source (the coupled ocean's water-leaving field,
    // rt_surface_boundary.c:495-527) enters the same way: its value at the
    // node, attenuated to level k.  bsrc_only drops the solar source.
    // The diffuse field entering an in-water column at the top (the
    // transmitted skylight) enters the downward columns alike, attenuated
    // from the top to level k.
    const bool rough_seed = rough_seed_on(I) && hemi == 0;
    const bool ext = (I.surf != 0 && m <= I.bsrc_mmax && hemi == 0) || (m <= I.tsrc_mmax && hemi == 1);
    const bool solar = !I.bsrc_only;
    const double tau = lv.h[nt], down = exp(-tau / I.mus);
    const double FS = ((m == 0) ? OCRT_PI : 0.5 * OCRT_PI) * I.mus;
    const double rsun = rough_seed ? P.R[I.off_Rsun + ((long long)m * n + o) * 9 + 3 * c] : 0.0;
    const double seed0 = rsun * (FS * down);
    const bool top_up = I.top_kind != 0;
    const double ext0 = !ext ? 0.0
                        : hemi == 0 ? P.greek[I.off_bsrc + ((long long)c * (I.bsrc_mmax + 1) + m) * n + o]
                                    : P.greek[I.off_tsrc + ((long long)c * (I.tsrc_mmax + 1) + m) * n + o];
    double att = 1.0;
    sweep_column_f64(hemi, nt, mu, 0.0, lv.h, lv.rdt, T, n, o,
                     [&](int k) { return solar ? src.primary(lv, k) + src.flat_value(lv, I.mus, k) : 0.0; },
                     [&](int k, double v, double tr, auto where) {
                         constexpr int kWhere = decltype(where)::value;   // 0 starting row, 1 interior, 2 far row
                         if (rough_seed || ext) {
                             // the running product of the layers crossed since the start
                             if constexpr (kWhere != 0) att *= tr;
                             if (rough_seed) v += seed0 * att;
                             if (ext) v += ext0 * att;
                         }
                         const long long ix = (long long)k * Pp + col;
                         if constexpr (MIXED) {
                             if constexpr (kWhere != 1) P.trow[it.off_cmp + (k == 0 ? 0 : Pp) + col] = v;
                             store_field_pair(reinterpret_cast<__hip_bfloat16 *>(F), Flo, ix, v);
                         } else {
                             tot[ix] = v;
                             F[ix] = to_storage<FT>(v);
                         }
                         if constexpr (kWhere != 1) {
                             if (hemi == 1 && k == nt) P.bc[it.off_bc + 3 * n + c * n + o] = v;
                             if (top_up && hemi == 0 && k == 0) P.bc[it.off_bc + 9 * n + c * n + o] = v;
                         }
                     });
}

// ---------------------------------------------------------------------------
// Convergence of the order just finished and the lower boundary condition of
// the next one.  One workgroup per item.
template <typename FT>
__global__ void order_status(const DevInst *__restrict__ inst, const DevItem *__restrict__ items, Pools P,
                             ItemState *__restrict__ st, int it_next)
{
    const int item = blockIdx.x;
    const DevItem &it = items[item];
    const DevInst &I = inst[it.inst];
    __shared__ int act;
    if (threadIdx.x == 0) {
        ItemState s = st[item];
        if (it_next > 2 && s.active) {
            const double mc = __longlong_as_double((long long)s.max_cur);
            const double mt = __longlong_as_double((long long)s.max_tot);
            const double res = (mt > 0.0) ? mc / mt : 0.0;
            // the orders still to come, estimated from the last decay rate
            // (none before two residuals exist, none unless it decays)
            const double prev = s.residual, r = prev > 0.0 ? res / prev : 1.0;
            const double tail = r < 1.0 ? res * r / (1.0 - r) : INFINITY;
            s.orders = it_next - 1;
            s.residual = res;
            if (!isfinite(mc) || !isfinite(mt)) { s.status = -2; s.active = 0; }
            else if (res < I.tol) { s.converged = 1; s.active = 0; }
            else if (it_next - 1 >= I.max_orders) s.active = 0;
            // the next order's operator product runs reduced (gemm_bf16)
            s.rnext = (s.active && tail < (double)P.lo_tau) ? 1 : 0;
            s.reduced += s.rnext;
        }
        s.max_cur = 0ull;
        s.max_tot = 0ull;
        st[item] = s;
        act = s.active;
    }
    __syncthreads();
    if (!act) return;
    if (I.surf != 0) surface_bc_item(I, it, P);
    if (I.top_kind != 0) top_bc_item(I, it, P);
}

// ---------------------------------------------------------------------------
// Scattering integral in fp64 on v_mfma_f64_16x16x4f64.  Tile 64 (outgoing
// columns) x 64 (levels), K step 16, four waves of 32 x 32, one accumulator
// set per operator part.  Epilogue:
//   kGemmMix   J[lev][r]  = s0(lev) (A_0 F)[r][lev] + ydel(lev) (A_1 F)[r][lev]
//   kGemmParts J[lev][r]  = (A_0 F)[r][lev], J2[lev][r] = (A_1 F)[r][lev]
//   kGemmBoth  J as kGemmMix, and rp0 / rp1 as J / J2 of kGemmParts (an
//              order the adjoint rebuilds between checkpoints)
// with s0 = xdel for a particle first part, ydel for a molecular one.
template <int MODE>
__global__ __launch_bounds__(256) void gemm_f64(const DevInst *__restrict__ inst,
                                                const DevItem *__restrict__ items, Pools P,
                                                const GemmTile *__restrict__ tiles,
                                                const ItemState *__restrict__ st, int order)
{
    const GemmTile tl = tiles[blockIdx.x];
    if (tl.item < 0 || !item_runs(st[tl.item], order)) return;
    const DevItem &it = items[tl.item];
    const DevInst &I = inst[it.inst];
    const int K = it.K, Pp = it.P, L = it.L;
    const int nparts = K / Pp;
    const double *A = reinterpret_cast<const double *>(P.A) + it.off_A;
    const double *F = reinterpret_cast<const double *>(P.F) + it.off_F;
    double *J = reinterpret_cast<double *>(P.J) + it.off_F;
    const LevelView lv(P, I);
    __shared__ double As[64][17];
    __shared__ double Bs[16][65];
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int row0 = tl.tm * 64, lev0 = tl.tn * 64;
    double4_ acc[2][2][2];
    for (int q = 0; q < 2; ++q)
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) acc[q][a][b] = (double4_){0.0, 0.0, 0.0, 0.0};
    const int lr = t >> 2, lk = (t & 3) * 4;
    const int lev_load = lev0 + lr;
    // one K sweep per operator part; the part index is a compile-time
    // constant so the accumulators stay in registers
    auto run_part = [&](auto part) {
        constexpr int q = decltype(part)::value;
        for (int k0 = 0; k0 < Pp; k0 += 16) {
            const double *src = A + (long long)(row0 + lr) * K + q * Pp + k0 + lk;
            for (int e = 0; e < 4; ++e) As[lr][lk + e] = src[e];
            if (lev_load < L) {
                const double *fs = F + (long long)lev_load * Pp + k0 + lk;
                for (int e = 0; e < 4; ++e) Bs[lk + e][lr] = fs[e];
            } else {
                for (int e = 0; e < 4; ++e) Bs[lk + e][lr] = 0.0;
            }
            __syncthreads();
#pragma unroll
            for (int ks = 0; ks < 16; ks += 4) {
                double av[2], bv[2];
                for (int a = 0; a < 2; ++a) av[a] = As[wm + a * 16 + (lane & 15)][ks + (lane >> 4)];
                for (int b = 0; b < 2; ++b) bv[b] = Bs[ks + (lane >> 4)][wn + b * 16 + (lane & 15)];
                for (int a = 0; a < 2; ++a)
                    for (int b = 0; b < 2; ++b)
                        acc[q][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[q][a][b], 0, 0, 0);
            }
            __syncthreads();
        }
    };
    run_part(std::integral_constant<int, 0>{});
    if (nparts > 1) run_part(std::integral_constant<int, 1>{});
    // f64 16x16 result layout: lane holds rows (lane >> 4) + 4 i of column lane & 15
    const int kind0 = it.parts & 0xff;
    double *J2 = reinterpret_cast<double *>(P.J2) + it.off_F;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            const int lev = lev0 + wn + b * 16 + (lane & 15);
            const int r = row0 + wm + a * 16 + (lane >> 4);
            if (lev >= L) continue;
            const double x = lv.xd[lev], y = lv.yd[lev];
            const double s0 = (kind0 == 1) ? x : y;
            double *dst = J + (long long)lev * Pp;
            double *dst2 = J2 + (long long)lev * Pp;
            for (int i = 0; i < 4; ++i) {
                if (r + 4 * i >= Pp) continue;
                if (MODE == kGemmMix || MODE == kGemmBoth) {
                    double v = s0 * acc[0][a][b][i];
                    if (nparts > 1) v += y * acc[1][a][b][i];
                    dst[r + 4 * i] = v;
                    if (MODE == kGemmBoth) {
                        const long long ix = it.off_F + (long long)lev * Pp + r + 4 * i;
                        P.rp0[ix] = acc[0][a][b][i];
                        if (nparts > 1) P.rp1[ix] = acc[1][a][b][i];
                    }
                } else {
                    dst[r + 4 * i] = acc[0][a][b][i];
                    if (nparts > 1) dst2[r + 4 * i] = acc[1][a][b][i];
                }
            }
        }
}

// Scattering integral on bf16 MFMA (v_mfma_f32_16x16x32_bf16, gfx950) with
// fp32 accumulation.  Tile 64 (outgoing columns) x 128 (levels), K step 32:
// the operator tile is read once per 128 levels, which halves its share of
// the L2 traffic against a 64 x 64 tile.  Four waves of 32 x 64, each holding
// 2 x 4 accumulator blocks of 16 x 16 per operator part.  The operator of
// each part (particle, molecular) is stored as a bf16 pair A = A_hi + A_lo,
// [part][hi, lo][Mp][P]; both halves multiply the same bf16 field fragment
// (and, for the low Fourier modes, the hi half also its bf16 residual).  The
// level scattering fractions are applied to the fp32 accumulators in the
// epilogue:
//   J[lev][r] = xdel(lev) (A_p F)[r][lev] + ydel(lev) (A_r F)[r][lev].
typedef __bf16 bf16x8_ __attribute__((ext_vector_type(8)));

__device__ inline bf16x8_ lds_frag(const unsigned short *p)
{
    return __builtin_bit_cast(bf16x8_, *reinterpret_cast<const uint4 *>(p));
}

// LDS staging of one K slab of a tile: [part][hi, lo][row][k] of the operator
// and [hi, lo][level][k] of the field, 32 entries (four 16-byte chunks) per
// row, chunk c of row r stored at position c ^ ((r >> 1) & 3).  On gfx950
// that swizzle keeps both the staging stores (thread t: row t / 4, chunk
// t % 4) and the fragment reads (lane l: row l % 16, chunk l / 16) free of
// bank conflicts; a padded 40-entry row (80 bytes) conflicts on both, which
// doubles the LDS cycles of the tile loop.
// K slabs of the bf16 scattering integral in flight per tile: each stage of
// LDS holds one slab of 32, filled by direct global -> LDS loads.
constexpr int kGemmStages = 2;

template <int NP, bool SPLIT, int TM = 64, int TN = kGemmLevels> struct GemmLds {
    unsigned short A[kGemmStages][NP][2][TM][32];
    unsigned short B[kGemmStages][SPLIT ? 2 : 1][TN][32];
};

// One 16-byte chunk per lane of a wavefront straight from global memory into
// LDS (gfx950 global_load_lds_dwordx4): lane l's chunk lands at lds + 16 l.
// The destination base goes in M0 and must be wavefront-uniform.  The load
// counts in vmcnt like any vector load; nothing else in the kernels that use it
// sets M0.
__device__ inline void lds_dma16(const unsigned short *g, unsigned short *lds)
{
    const uint32_t la = (uint32_t)(uintptr_t)(__attribute__((address_space(3))) void *)lds;
    asm volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dwordx4 %1, off" ::"s"(la), "v"(g) : "memory");
}

// Wait until at most N vector-memory operations of this wavefront are
// outstanding and every LDS read has returned, then the workgroup barrier.
template <int N> __device__ inline void wait_vm_barrier()
{
    static_assert(N >= 0 && N < 64, "vmcnt range");
    asm volatile("s_waitcnt vmcnt(%0) lgkmcnt(0)\n\ts_barrier" ::"n"(N) : "memory");
}

// One tile of the bf16 scattering integral: operator rows row0 .. row0 + TM - 1
// and levels lev0 .. lev0 + TN - 1 of one item with NP operator parts (the
// item's K is NP P), four waves of TM/2 x TN/2.  The tile shapes (64 x 128,
// 64 x 64, 128 x 64, 128 x 128) give the same bits: every output takes the same K slabs
// in the same order.  NP is a template parameter so that single-part items
// hold half the accumulators (higher occupancy); SPLIT = false compiles out
// the bf16 pair field (its LDS plane) for tiles whose items keep a single
// bf16 field (SPLIT = true tiles hold only items that keep the pair).
// K slabs of 32 go from global memory straight into one of kGemmStages LDS
// stages (lds_dma16), kGemmStages - 1 slabs ahead of the MFMAs, so no staging
// registers are held and several slabs' loads are in flight.  Each row of a
// slab is 32 entries (four 16-byte chunks); chunk c of row r sits at position
// c ^ ((r >> 1) & 3), which keeps the fragment reads (lane l: row l % 16,
// chunk l / 16) free of bank conflicts on gfx950: lane l of a load writes
// position l % 4 of row l / 4, so it fetches the chunk that belongs there.
// Rows past the operator's Mp rows load its last row and levels past L the
// item's last level (their results are never used).  EPI = 0: the fp32
// results, scaled by the level fractions, go to out(lev, r, v) for the four
// rows r .. r + 3 of v; EPI = 1: the parts go unscaled to out(lev, r, v0, v1)
// (v1 zero for one part); EPI = 2: both, out(lev, r, v, v0, v1); levels < L
// only.  LO = false is the reduced product: the operator's hi halves times the
// bf16 field alone (no A_lo, no field residual), one MFMA per block.
template <int NP, bool SPLIT, int EPI = 0, int TM = 64, int TN = kGemmLevels, bool LO = true, typename Out>
__device__ inline void gemm_tile_bf16(const DevItem &it, const LevelView &lv, const unsigned short *A,
                                      const unsigned short *F, const unsigned short *Flo, float *Wp, int row0,
                                      int lev0, GemmLds<NP, SPLIT, TM, TN> &S, Out out)
{
    constexpr int AB = TM / 32, BB = TN / 32;   // 16 x 16 blocks per wave along rows and levels
    constexpr int AR = TM / 64, FR = TN / 64;   // 64-row groups of the operator and the field per slab
    constexpr bool split = LO && SPLIT;
    constexpr bool two = NP > 1;
    // loads per thread and slab
    constexpr int kLoads = NP * (LO ? 2 : 1) * AR + (split ? 2 : 1) * FR;
    constexpr int ST = kGemmStages;
    const int Pp = it.P, L = it.L, Mp = it.Mp;
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int wm = (wave >> 1) * (TM / 2), wn = (wave & 1) * (TN / 2);
    float4_ acc[NP][AB][BB];
#pragma unroll
    for (int q = 0; q < NP; ++q)
#pragma unroll
        for (int a = 0; a < AB; ++a)
#pragma unroll
            for (int b = 0; b < BB; ++b) acc[q][a][b] = (float4_){0.f, 0.f, 0.f, 0.f};
    // this lane's row within its wavefront's 16-row group and the chunk it
    // fetches for its position (the swizzle's inverse; rows 16 apart share it)
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    const int rl = lane >> 2;
    const int ck = ((lane & 3) ^ ((rl >> 1) & 3)) * 8;
    const long long slab = (long long)Mp * Pp;
    // one slab (k0) into stage g: the operator planes [part][hi, lo] and the
    // field planes [hi, lo], 64-row groups j, wavefront wv taking rows
    // 64 j + 16 wv .. + 15 of each
    auto issue = [&](int k0, int g) {
#pragma unroll
        for (int j = 0; j < AR; ++j) {
            const int r = 64 * j + 16 * wv + rl;
            const unsigned short *src = A + (long long)min(row0 + r, Mp - 1) * Pp + k0 + ck;
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                lds_dma16(src + (2 * q) * slab, &S.A[g][q][0][64 * j + 16 * wv][0]);
                if constexpr (LO) lds_dma16(src + (2 * q + 1) * slab, &S.A[g][q][1][64 * j + 16 * wv][0]);
            }
        }
#pragma unroll
        for (int j = 0; j < FR; ++j) {
            const int r = 64 * j + 16 * wv + rl;
            const long long off = (long long)min(lev0 + r, L - 1) * Pp + k0 + ck;
            lds_dma16(F + off, &S.B[g][0][64 * j + 16 * wv][0]);
            if constexpr (split) lds_dma16(Flo + off, &S.B[g][split ? 1 : 0][64 * j + 16 * wv][0]);
        }
    };
    // fragment of a 16 x 32 block: lane l holds row (l & 15), k 8 (l >> 4) .. + 7
    const int fr = lane & 15, fk = (((lane >> 4) ^ (fr >> 1)) & 3) * 8;   // k chunk lane / 16, swizzled
    auto mfma_part = [&](int g, auto part) {
        constexpr int q = decltype(part)::value;
#pragma unroll
        for (int a = 0; a < AB; ++a) {
            const bf16x8_ hi = lds_frag(&S.A[g][q][0][wm + a * 16 + fr][fk]);
            bf16x8_ lo;
            if constexpr (LO) lo = lds_frag(&S.A[g][q][1][wm + a * 16 + fr][fk]);
#pragma unroll
            for (int b = 0; b < BB; ++b) {
                const bf16x8_ fb = lds_frag(&S.B[g][0][wn + b * 16 + fr][fk]);
                if constexpr (split) {
                    const bf16x8_ fl = lds_frag(&S.B[g][split ? 1 : 0][wn + b * 16 + fr][fk]);
                    acc[q][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(hi, fl, acc[q][a][b], 0, 0, 0);
                }
                if constexpr (LO) acc[q][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(lo, fb, acc[q][a][b], 0, 0, 0);
                acc[q][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(hi, fb, acc[q][a][b], 0, 0, 0);
            }
        }
    };
    // slabs 0 .. ST - 2 first; step s then waits for slab s (the later slabs'
    // loads may stay in flight), passes the barrier (every wavefront is done
    // with the stage read in step s - 1), refills that stage with slab
    // s + ST - 1 and runs slab s's MFMAs
    const int ns = Pp / 32;
#pragma unroll
    for (int j = 0; j < ST - 1; ++j)
        if (j < ns) issue(32 * j, j);
    for (int s = 0; s < ns; ++s) {
        const int g = s % ST;
        if (s + ST - 2 < ns) wait_vm_barrier<(ST - 2) * kLoads>();
        else wait_vm_barrier<0>();
        if (s + ST - 1 < ns) issue(32 * (s + ST - 1), (s + ST - 1) % ST);
        mfma_part(g, std::integral_constant<int, 0>{});
        if constexpr (NP > 1) mfma_part(g, std::integral_constant<int, 1>{});
    }
    // 16 x 16 result layout: lane holds rows 4 (lane >> 4) + i of column lane & 15
    if constexpr (EPI == 1) {
#pragma unroll
        for (int b = 0; b < BB; ++b) {
            const int lev = lev0 + wn + b * 16 + (lane & 15);
            if (lev >= L) continue;
#pragma unroll
            for (int a = 0; a < AB; ++a) {
                const int r = row0 + wm + a * 16 + 4 * (lane >> 4);
                const float4 v0 = {acc[0][a][b][0], acc[0][a][b][1], acc[0][a][b][2], acc[0][a][b][3]};
                float4 v1 = {0.f, 0.f, 0.f, 0.f};
                if constexpr (two) v1 = {acc[NP - 1][a][b][0], acc[NP - 1][a][b][1], acc[NP - 1][a][b][2],
                                         acc[NP - 1][a][b][3]};
                out(lev, r, v0, v1);
            }
        }
    } else {
        const int kind0 = it.parts & 0xff;
#pragma unroll
        for (int b = 0; b < BB; ++b) {
            const int lev = lev0 + wn + b * 16 + (lane & 15);
            if (lev >= L) continue;
            const float x = (float)lv.xd[lev], y = (float)lv.yd[lev];
            const float s0 = (kind0 == 1) ? x : y;
#pragma unroll
            for (int a = 0; a < AB; ++a) {
                const int rb = row0 + wm + a * 16, r = rb + 4 * (lane >> 4);
                // projection rows of a rank-3 molecular item: W = G F, unscaled,
                // rows pbase .. pbase + 2 and a zero fourth (pbase is a multiple
                // of 4, so one lane group holds all four); those rows lie past
                // the item's 6 n columns, so their J entries below are unused
                if (Wp && r == it.pbase)
                    *reinterpret_cast<float4 *>(Wp + 4LL * lev) =
                        make_float4(acc[0][a][b][0], acc[0][a][b][1], acc[0][a][b][2], acc[0][a][b][3]);
                float4 v;
                v.x = s0 * acc[0][a][b][0]; v.y = s0 * acc[0][a][b][1];
                v.z = s0 * acc[0][a][b][2]; v.w = s0 * acc[0][a][b][3];
                if constexpr (two) {
                    v.x = fmaf(y, acc[NP - 1][a][b][0], v.x); v.y = fmaf(y, acc[NP - 1][a][b][1], v.y);
                    v.z = fmaf(y, acc[NP - 1][a][b][2], v.z); v.w = fmaf(y, acc[NP - 1][a][b][3], v.w);
                }
                if constexpr (EPI == 2) {
                    // the scaled sum and the unscaled parts
                    const float4 p0 = {acc[0][a][b][0], acc[0][a][b][1], acc[0][a][b][2], acc[0][a][b][3]};
                    float4 p1 = {0.f, 0.f, 0.f, 0.f};
                    if constexpr (two) p1 = {acc[NP - 1][a][b][0], acc[NP - 1][a][b][1], acc[NP - 1][a][b][2],
                                             acc[NP - 1][a][b][3]};
                    out(lev, r, v, p0, p1);
                } else {
                    out(lev, r, v);
                }
            }
        }
    }
}

// NP: operator parts of every item in the tile list; SPLIT: whether its items
// may keep their field as a bf16 pair (the plan keeps one list per kind).
// Single-part unsplit tiles of up to 64 x 128 fit 96 registers and 20 KB of
// LDS: five waves per SIMD; their 128 x 128 tiles two.
template <int NP, bool SPLIT, int TM, int TN>
__global__ __launch_bounds__(256, (NP == 1 && !SPLIT) ? (TM * TN > 8192 ? 2 : 5) : 1) void gemm_bf16(const DevInst *__restrict__ inst,
                                                 const DevItem *__restrict__ items, Pools P,
                                                 const GemmTile *__restrict__ tiles,
                                                 const ItemState *__restrict__ st, int order)
{
    const GemmTile tl = tiles[blockIdx.x];
    if (tl.item < 0) return;
    const ItemState is = st[tl.item];
    if (!item_runs(is, order)) return;
    // the reduced product once the item's estimated remaining share fell
    // below lo_tau (order_status); the test is workgroup-uniform
    const bool reduced = is.rnext != 0;
    const DevItem &it = items[tl.item];
    const DevInst &I = inst[it.inst];
    const int Pp = it.P;
    const unsigned short *A = reinterpret_cast<const unsigned short *>(P.A) + it.off_A;
    const unsigned short *F = reinterpret_cast<const unsigned short *>(P.F) + it.off_F;
    const unsigned short *Flo = reinterpret_cast<const unsigned short *>(P.Flo) + it.off_Flo;
    // J is stored in fp16 (11 significant bits), scaled by the instance's
    // power of two jscale: half the bytes of the write here and of the sweep's
    // read.  The items that keep their field as a pair (16 significant bits)
    // keep an fp32 source.
    _Float16 *J = reinterpret_cast<_Float16 *>(P.J) + it.off_F;
    float *Jf = P.Jf + it.off_Flo;
    const float js = I.jscale;
    const LevelView lv(P, I);
    __shared__ GemmLds<NP, SPLIT, TM, TN> S;
    float *Wp = (SPLIT && it.lowrank) ? P.W + it.off_W : nullptr;
    auto out = [&](int lev, int r, float4 v) {
        if (r >= Pp) return;
        if constexpr (SPLIT) {
            *reinterpret_cast<float4 *>(Jf + (long long)lev * Pp + r) = v;
        } else {
            typedef _Float16 half4_ __attribute__((ext_vector_type(4)));
            const half4_ h = {(_Float16)(v.x * js), (_Float16)(v.y * js), (_Float16)(v.z * js),
                              (_Float16)(v.w * js)};
            *reinterpret_cast<half4_ *>(J + (long long)lev * Pp + r) = h;
        }
    };
    if (reduced)
        gemm_tile_bf16<NP, SPLIT, 0, TM, TN, false>(it, lv, A, F, Flo, Wp, tl.tm * TM, tl.tn * TN, S,
                                                           out);
    else
        gemm_tile_bf16<NP, SPLIT, 0, TM, TN, true>(it, lv, A, F, Flo, Wp, tl.tm * TM, tl.tn * TN, S,
                                                          out);
}

// Operator products of the mixed adjoint plans on bf16 MFMA, three MFMAs per
// block (A_hi F_hi + A_lo F_hi + A_hi F_lo), fp32 results:
//   MODE kGemmParts: the forward field u_{o-1}, kept as its bf16 pair
//     (P.F, P.Flo), times the operator A, the parts unscaled:
//     J[lev][r] = (A_0 u)[r][lev], J2[lev][r] = (A_1 u)[r][lev]
//   MODE kGemmMix: the adjoint source Sbar, as its bf16 pair (P.F and P.Flo),
//     times the transposed operator (P.A):
//     J[lev][r] = s0(lev) (A_0^T Sbar)[r][lev] + ydel(lev) (A_1^T Sbar)[r][lev]
// with s0 = xdel for a particle first part, ydel for a molecular one.
template <int NP, int MODE>
__global__ __launch_bounds__(256, 1) void gemm_adj(const DevInst *__restrict__ inst, const DevItem *__restrict__ items,
                                                   Pools P, const GemmTile *__restrict__ tiles,
                                                   const ItemState *__restrict__ st, int order)
{
    const GemmTile tl = tiles[blockIdx.x];
    if (tl.item < 0 || !item_runs(st[tl.item], order)) return;
    const DevItem &it = items[tl.item];
    const DevInst &I = inst[it.inst];
    const int Pp = it.P;
    const unsigned short *A = reinterpret_cast<const unsigned short *>(P.A) + it.off_A;
    const unsigned short *F = reinterpret_cast<const unsigned short *>(P.F) + it.off_F;
    const unsigned short *Flo = reinterpret_cast<const unsigned short *>(P.Flo) + it.off_Flo;
    float *J = reinterpret_cast<float *>(P.J) + it.off_F;
    float *J2 = MODE == kGemmParts ? reinterpret_cast<float *>(P.J2) + it.off_F : nullptr;
    const LevelView lv(P, I);
    __shared__ GemmLds<NP, true> S;
    if constexpr (MODE == kGemmMix) {
        gemm_tile_bf16<NP, true, 0>(it, lv, A, F, Flo, nullptr, tl.tm * 64, tl.tn * kGemmLevels, S,
                                         [&](int lev, int r, float4 v) {
                                             if (r >= Pp) return;
                                             *reinterpret_cast<float4 *>(J + (long long)lev * Pp + r) = v;
                                         });
    } else {
        gemm_tile_bf16<NP, true, 1>(it, lv, A, F, Flo, nullptr, tl.tm * 64, tl.tn * kGemmLevels, S,
                                         [&](int lev, int r, float4 v0, float4 v1) {
                                             if (r >= Pp) return;
                                             *reinterpret_cast<float4 *>(J + (long long)lev * Pp + r) = v0;
                                             if (NP > 1) *reinterpret_cast<float4 *>(J2 + (long long)lev * Pp + r) = v1;
                                         });
    }
}

// Scattering integral of an order rebuilt by the mixed adjoint between two
// checkpoints (its items all keep the field pair): the source of the rebuilt
// sweep, J[lev][r] = s0(lev) (A_0 F)[r][lev] + ydel(lev) (A_1 F)[r][lev] in the
// fp32 source pool, and the unscaled parts the same order's backward step reads
// (P.J = A_0 F, P.J2 = A_1 F, fp32), from one product.  64 x 128 tiles.
template <int NP>
__global__ __launch_bounds__(256, 1) void gemm_rebuild(const DevInst *__restrict__ inst,
                                                       const DevItem *__restrict__ items, Pools P,
                                                       const GemmTile *__restrict__ tiles,
                                                       const ItemState *__restrict__ st, int order)
{
    const GemmTile tl = tiles[blockIdx.x];
    if (tl.item < 0 || !item_runs(st[tl.item], order)) return;
    const DevItem &it = items[tl.item];
    const DevInst &I = inst[it.inst];
    const int Pp = it.P;
    const unsigned short *A = reinterpret_cast<const unsigned short *>(P.A) + it.off_A;
    const unsigned short *F = reinterpret_cast<const unsigned short *>(P.F) + it.off_F;
    const unsigned short *Flo = reinterpret_cast<const unsigned short *>(P.Flo) + it.off_Flo;
    float *Jf = P.Jf + it.off_Flo;
    float *J0 = reinterpret_cast<float *>(P.J) + it.off_F;
    float *J1 = reinterpret_cast<float *>(P.J2) + it.off_F;
    const LevelView lv(P, I);
    __shared__ GemmLds<NP, true> S;
    gemm_tile_bf16<NP, true, 2>(it, lv, A, F, Flo, nullptr, tl.tm * 64, tl.tn * kGemmLevels, S,
                                   [&](int lev, int r, float4 v, float4 p0, float4 p1) {
                                       if (r >= Pp) return;
                                       const long long ix = (long long)lev * Pp + r;
                                       *reinterpret_cast<float4 *>(Jf + ix) = v;
                                       *reinterpret_cast<float4 *>(J0 + ix) = p0;
                                       if (NP > 1) *reinterpret_cast<float4 *>(J1 + ix) = p1;
                                   });
}

// Formal solution of the current order for every column, fused with the order
// sum total += cur and the two max-norms of the convergence test
// (rt_solver.c:1855-1868, 3093-3124).  One thread per column, one 64-column
// tile per wavefront, W wavefronts per workgroup; the levels are walked in
// blocks of U, every load of a block issued before its dependent recurrence
// steps, so the sequential scan overlaps its memory latency instead of
// exposing it once per level.  U and W do not change the arithmetic (every
// variant gives the same bits); the plan's autotuner picks them.
//
// The order sum is fp64 over the whole field (it feeds the convergence test).
// In the mixed form the two rows that become outputs, k = 0 (TOA) and k = nt
// (BOA), keep it in fp64 with an fp32 Neumaier compensation term and the
// interior levels in fp32; the scan takes the boundary row, the interior
// levels without per-level row tests, then the far row.
__device__ inline int clamp_level(int k, int nt) { return k < 0 ? 0 : (k > nt ? nt : k); }
__device__ inline int clamp_layer(int l, int nt) { return l < 0 ? 0 : (l > nt - 1 ? nt - 1 : l); }

// One level of the mixed-form scan in fp32, written with om = 1 - T (accurate
// from expm1):
//   run <- run - om run + ((S_out - S_in) + om (S_in + slope mu_signed)) / 2
// the linear-source recurrence with no 1 - T cancellation; slope is the source
// gradient in optical depth along the scan.  The multiply-adds are explicit,
// so every kernel that scans computes the same bits.
__device__ inline float mixed_step(float run, float s_in, float s_out, float om, float idt, float smu, bool down)
{
    const float a = (down ? (s_out - s_in) : (s_in - s_out)) * idt;
    return fmaf(0.5f, fmaf(om, fmaf(a, smu, s_in), s_out - s_in), fmaf(-om, run, run));
}

// One column of the mixed form: its field (bf16, plus the bf16 residual for
// the items that keep the pair), the order sum of its two output rows k = 0
// and k = nt (fp64 with an fp32 Neumaier term) and the column's two
// max-norms.  The sweep walks the interior levels itself, branch-free.
// The mixed form keeps no order sum at the interior levels: its convergence
// test divides the order's largest value over every level by the largest
// order sum over the output rows, a denominator no larger than the one over
// every level, so the test stops at the same order as the all-level test or
// later, never earlier.
struct MixedColumn {
    __hip_bfloat16 *F, *Flo;   // level stride P; Flo null unless the item keeps the pair
    double *row_top, *row_bot;    // fp64 order sum of the output rows k = 0 and k = nt
    double *bottom, *top;         // top: upward field at the top of an in-water column, else null
    float *cmp_top, *cmp_bot;
    long long Pp;
    int nt, down;
    bool acc;                     // false: recomputation, the field alone (no order sum)
    float mcur, mtot;

    __device__ MixedColumn(const Pools &P, const DevItem &it, int col, int c, int o, int down_, bool acc_ = true)
    {
        const int n = it.n;
        F = reinterpret_cast<__hip_bfloat16 *>(P.F) + it.off_F + col;
        Flo = it.fsplit ? reinterpret_cast<__hip_bfloat16 *>(P.Flo) + it.off_Flo + col : nullptr;
        row_top = P.trow + it.off_cmp + col;
        row_bot = row_top + it.P;
        bottom = P.bc + it.off_bc + 3 * n + c * n + o;
        top = nullptr;
        cmp_top = P.comp + it.off_cmp + col;
        cmp_bot = P.comp + it.off_cmp + it.P + col;
        Pp = it.P;
        nt = it.nt;
        down = down_;
        acc = acc_;
        mcur = mtot = 0.f;
    }

    // the field value v at the output row k (0 or nt)
    __device__ void emit_row(int k, double v)
    {
        const long long ix = (long long)k * Pp;
        store_field_pair(F, Flo, ix, v);
        mcur = fmaxf(mcur, fabsf((float)v));
        if (down && k == nt) *bottom = v;
        if (top && k == 0) *top = v;
        if (!acc) return;
        double *rs = k == 0 ? row_top : row_bot;
        const double t_old = *rs, tsum = t_old + v;
        *rs = tsum;
        mtot = fmaxf(mtot, fabsf((float)tsum));
        // Neumaier: the rounding error of this addition, kept in fp32
        float *cp = k == 0 ? cmp_top : cmp_bot;
        const double e = (fabs(t_old) >= fabs(v)) ? ((t_old - tsum) + v) : ((v - tsum) + t_old);
        *cp = (float)((double)*cp + e);
    }
};

template <typename FT, typename JT, bool MIXED, int U, int W>
__global__ __launch_bounds__(64 * W) void sweep(const DevInst *__restrict__ inst,
                                                const DevItem *__restrict__ items, Pools P,
                                                ItemState *__restrict__ st, const ColTile *__restrict__ tiles,
                                                int n_tiles, int order, const double *__restrict__ lev,
                                                const float *__restrict__ idtp, const float *__restrict__ wpool)
{
    const int tix = blockIdx.x * W + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (tix >= n_tiles) return;
    const ColTile ct = tiles[tix];
    // the tile is wavefront-uniform: keep its item in scalar registers
    const int item = __builtin_amdgcn_readfirstlane(ct.item);
    if (item < 0 || !item_runs(st[item], order)) return;
    const bool accumulate = order == 0;   // recomputation only rebuilds the field
    const DevItem &it = items[item];
    const DevInst &I = inst[it.inst];
    // every descriptor field the level loop needs is read once here: the loop's
    // stores could otherwise alias the descriptors and force reloads
    const int n = it.n, nt = it.nt, Pp = it.P;
    const long long off_F = it.off_F, off_bc = it.off_bc, off_cmp = it.off_cmp;
    const int col = ct.col0 + lane;
    double mcur = 0.0, mtot = 0.0;
    if (col < ct.col_end && col < 6 * n) {
        // level tables, 1 / dtau and the rank-3 projections through the __restrict__
        // arguments (P.lev, P.idt, P.W): the per-level uniform reads are scalar loads
        const LevelView lv(lev, I);
        int c, hemi, o;
        col_decode(col, n, c, hemi, o);
        hemi = __builtin_amdgcn_readfirstlane(hemi);   // column tiles are direction-pure
        const JT *__restrict__ J = reinterpret_cast<const JT *>(P.J) + off_F + col;
        const double mu = P.ring[I.off_ring + o];
        // boundary values: the surface's at the bottom of the upward columns,
        // the water-side reflection's at the top of an in-water column's
        // downward ones
        const double bc = hemi == 0 ? (I.surf != 0 ? P.bc[off_bc + c * n + o] : 0.0)
                                    : (I.top_kind != 0 ? P.bc[off_bc + 6 * n + c * n + o] : 0.0);
        double *top_up = (I.top_kind != 0 && hemi == 0) ? P.bc + off_bc + 9 * n + c * n + o : nullptr;
        // level order of the scan and the layer crossed when stepping to level k
        const int k_first = hemi ? 0 : nt;
        const int step = hemi ? 1 : -1;
        auto layer_of = [&](int k) { return hemi ? k - 1 : k; };
        if constexpr (MIXED) {
            const float *OM = P.Tf + I.off_T;
            const float *IDT = idtp + I.off_lev;
            MixedColumn mc(P, it, col, c, o, hemi, accumulate);
            mc.top = top_up;
            const float smu = hemi ? -(float)mu : (float)mu;
            // the source: fp32 for the items that keep their field as a pair,
            // else fp16 scaled by jscale
            const float ji = I.jinv;
            const float *__restrict__ Jf = P.Jf + it.off_Flo + col;
            // rank-3 molecular items add ydel(k) (U W[k]) to the dense source:
            // the column's three output factors, the item's projections
            const float *__restrict__ Wl = wpool + it.off_W;
            const double *__restrict__ YD = lv.yd;
            float u0 = 0.f, u1 = 0.f, u2 = 0.f;
            if (it.lowrank) {
                const float *Ul = P.lr + it.off_lr + col;
                u0 = Ul[0]; u1 = Ul[Pp]; u2 = Ul[2LL * Pp];
            }
            auto mol = [&](int k) -> float {
                const float4 w = *reinterpret_cast<const float4 *>(Wl + 4LL * k);
                return (float)YD[k] * fmaf(u2, w.z, fmaf(u1, w.y, u0 * w.x));
            };
            // kind 0: fp16 source; 1: fp32 source; 2: fp32 source plus the
            // rank-3 molecular term
            auto scan = [&](auto kind_tag) {
                constexpr int KIND = decltype(kind_tag)::value;
                constexpr bool SPLIT = KIND >= 1;
                auto jload = [&](int k) -> float {
                    if constexpr (KIND == 2) return Jf[(long long)k * Pp] + mol(k);
                    else if constexpr (SPLIT) return Jf[(long long)k * Pp];
                    else return (float)J[(long long)k * Pp] * ji;
                };
                float run = (float)bc;
                float s_in = jload(k_first);
                mc.emit_row(k_first, bc);
                float cur = 0.f;
                // interior levels, i = 1 .. nt - 1 steps along the scan: running
                // pointers at the level (source, field) and at the layer crossed
                // to reach it (1 - T, 1 / dtau)
                const long long dP = hemi ? (long long)Pp : -(long long)Pp;
                const int dL = hemi ? 1 : -1;
                const int k1 = k_first + dL, l1 = layer_of(k1);
                const long long ix1 = (long long)k1 * Pp;
                const _Float16 *__restrict__ pJ = J + ix1;
                const float *__restrict__ pJf = Jf + ix1;
                __hip_bfloat16 *__restrict__ pF = mc.F + ix1;
                __hip_bfloat16 *__restrict__ pFlo = SPLIT ? mc.Flo + ix1 : nullptr;
                const float *__restrict__ pOM = OM + (long long)l1 * n + o;
                const float *__restrict__ pIDT = IDT + l1;
                auto level = [&](long long e, float sj, float om, float idt) {
                    run = mixed_step(run, s_in, sj, om, idt, smu, hemi);
                    const __hip_bfloat16 h = __float2bfloat16(run);
                    pF[e] = h;
                    if constexpr (SPLIT) pFlo[e] = __float2bfloat16(run - __bfloat162float(h));
                    cur = fmaxf(cur, fabsf(run));
                    s_in = sj;
                };
                auto src = [&](long long e, int k) -> float {
                    if constexpr (KIND == 2) return pJf[e] + mol(k);
                    else if constexpr (SPLIT) return pJf[e];
                    else return (float)pJ[e] * ji;
                };
                int i = 1;
                // whole blocks of U levels: every load of the block is issued
                // before its dependent steps
                for (; i + U <= nt; i += U) {
                    float sj[U], om[U], idt[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        sj[u] = src(u * dP, k1 + dL * (i - 1 + u));
                        om[u] = pOM[u * dL * n];
                        idt[u] = pIDT[u * dL];
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u) level(u * dP, sj[u], om[u], idt[u]);
                    pJ += U * dP; pJf += U * dP; pF += U * dP;
                    if constexpr (SPLIT) pFlo += U * dP;
                    pOM += U * dL * n; pIDT += U * dL;
                }
                // the remaining interior levels one at a time
                for (; i < nt; ++i) {
                    level(0, src(0, k1 + dL * (i - 1)), pOM[0], pIDT[0]);
                    pJ += dP; pJf += dP; pF += dP;
                    if constexpr (SPLIT) pFlo += dP;
                    pOM += dL * n; pIDT += dL;
                }
                // the far output row
                const int kl = hemi ? nt : 0, ll = layer_of(kl);
                run = mixed_step(run, s_in, jload(kl), OM[(long long)ll * n + o], IDT[ll], smu, hemi);
                mc.mcur = fmaxf(mc.mcur, cur);
                mc.emit_row(kl, (double)run);
            };
            if (it.lowrank)
                scan(std::integral_constant<int, 2>{});
            else if (it.fsplit)
                scan(std::integral_constant<int, 1>{});
            else
                scan(std::integral_constant<int, 0>{});
            mcur = mc.mcur;
            mtot = mc.mtot;
        } else {
            FT *__restrict__ F = reinterpret_cast<FT *>(P.F) + off_F + col;
            double *__restrict__ tot = P.tot + it.off_tot + col;
            double *__restrict__ bottom = P.bc + off_bc + 3 * n + c * n + o;
            // t_old: the order sum before this order at level k
            auto emit = [&](int k, double v, double t_old) {
                const long long ix = (long long)k * Pp;
                F[ix] = to_storage<FT>(v);
                mcur = fmax(mcur, fabs(v));
                if (hemi == 1 && k == nt) *bottom = v;
                if (top_up && k == 0) *top_up = v;
                if (!accumulate) return;
                const double tsum = t_old + v;
                tot[ix] = tsum;
                mtot = fmax(mtot, fabs(tsum));
            };
            emit(k_first, bc, accumulate ? tot[(long long)k_first * Pp] : 0.0);
            const double *T = P.T + I.off_T;
            double run = bc;
            double s_in = (double)J[(long long)k_first * Pp];
            const double smu = hemi ? -mu : mu;
            for (int k0 = k_first + step; hemi ? (k0 <= nt) : (k0 >= 0); k0 += step * U) {
                double sj[U], tr[U], to[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    // clamped indices keep the block's loads unconditional, so
                    // they are all in flight before the first one is used
                    const int k = clamp_level(k0 + step * u, nt);
                    sj[u] = (double)J[(long long)k * Pp];
                    tr[u] = T[(long long)clamp_layer(layer_of(k0 + step * u), nt) * n + o];
                    to[u] = tot[(long long)k * Pp];
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int k = k0 + step * u;
                    if (k < 0 || k > nt) break;
                    const int ku = k - step;                // upstream level
                    const double s_out = sj[u];
                    const double rd = lv.rdt[hemi ? k - 1 : k];   // 1 / dtau of the layer crossed
                    const double a = hemi ? (s_out - s_in) * rd : (s_in - s_out) * rd;
                    const double b = s_out - a * lv.h[k];
                    run = tr[u] * run + ((1.0 - tr[u]) * (b + a * smu) + a * (lv.h[k] - lv.h[ku] * tr[u])) * 0.5;
                    emit(k, run, to[u]);
                    s_in = s_out;
                }
            }
        }
    } else if (!MIXED && col < ct.col_end) {
        // pad columns (6n <= col < P) of the fp64 field are written too:
        // adjoint plans sweep every order into a buffer of its own, and the
        // next scattering integral multiplies these columns by zero operator
        // entries, which must not meet leftover NaN or Inf bit patterns
        FT *__restrict__ F = reinterpret_cast<FT *>(P.F) + off_F + col;
        for (int k = 0; k <= nt; ++k) F[(long long)k * Pp] = to_storage<FT>(0.0);
    }
    if (!accumulate) return;
    const double bm = wave_max(mcur);
    const double bt = wave_max(mtot);
    if (lane == 0) {
        atomicMax(&st[item].max_cur, (unsigned long long)__double_as_longlong(bm));
        atomicMax(&st[item].max_tot, (unsigned long long)__double_as_longlong(bt));
    }
}

// ---------------------------------------------------------------------------
// Order sum at an output row (k = 0 or k = nt) of column col: the mixed form
// adds the row's compensation term.
__device__ inline double total_at(const Pools &P, const DevItem &it, bool mixed, int k, int col)
{
    if (mixed) {
        const long long r = it.off_cmp + (k == 0 ? 0 : it.P) + col;
        return P.trow[r] + (double)P.comp[r];
    }
    return P.tot[it.off_tot + (long long)k * it.P + col];
}

// Per-item outputs after the order loop (rt_solver.c:1010-1061).
__global__ void extract(const DevInst *__restrict__ inst, const DevItem *__restrict__ items, Pools P,
                        const ItemState *__restrict__ st, ItemOut *__restrict__ out, int mixed)
{
    const int item = blockIdx.x;
    const DevItem &it = items[item];
    const DevInst &I = inst[it.inst];
    const int n = it.n, nt = it.nt, Pp = it.P, m = it.m, v = I.vidx;
    const double *mu = P.ring + I.off_ring, *w = mu + n;
    const LevelView lv(P, I);
    // ordered sum over nodes for the hemispheric flux
    if (threadIdx.x == 0) {
        double F = 0.0;
        for (int o = 0; o < n; ++o) F += mu[o] * w[o] * total_at(P, it, mixed, nt, col_dn(n, 0, o));
        const ItemState s = st[item];
        ItemOut r;
        const double tau = lv.h[nt];
        for (int c = 0; c < 3; ++c) {
            double val = total_at(P, it, mixed, 0, col_up(n, c, v));
            if (rough_seed_on(I)) {
                const double down = exp(-tau / I.mus);
                const double FS = ((m == 0) ? OCRT_PI : 0.5 * OCRT_PI) * I.mus;
                const double rk = P.R[I.off_Rsun + ((long long)m * n + v) * 9 + 3 * c];
                val -= rk * (FS * (down * exp(-(tau - lv.h[0]) / mu[v])));
            }
            r.mode[c] = val;
        }
        r.boa_view = total_at(P, it, mixed, nt, col_dn(n, 0, v));
        r.boa_hemi = F;
        r.orders = s.orders;
        r.converged = s.converged;
        r.residual = s.residual;
        r.reduced = s.reduced;
        out[item] = r;
    }
    if (P.nodes) {
        // every node of the TOA upward field (rough seed removed) and of the
        // BOA downward field, [6][n]: up I Q U, down I Q U (the LUT's view
        // sampling reads the first four rows)
        double *dst = P.nodes + it.off_nodes;
        for (int q = threadIdx.x; q < 6 * n; q += blockDim.x) {
            const int c = q / n, o = q % n;
            double val;
            if (c < 3) {
                val = total_at(P, it, mixed, 0, col_up(n, c, o));
                if (rough_seed_on(I)) {
                    const double tau = lv.h[nt], down = exp(-tau / I.mus);
                    const double FS = ((m == 0) ? OCRT_PI : 0.5 * OCRT_PI) * I.mus;
                    val -= P.R[I.off_Rsun + ((long long)m * n + o) * 9 + 3 * c] *
                           (FS * (down * exp(-(tau - lv.h[0]) / mu[o])));
                }
            } else {
                val = total_at(P, it, mixed, nt, col_dn(n, c - 3, o));
            }
            dst[q] = val;
        }
    }
}

// Instance results (rt_solver.c:1062-1276): Fourier reconstruction at the view
// azimuth, rho = X / mu_s, IPSS kappa, direct glint and transmittances.


__global__ void finalize(const DevInst *__restrict__ inst, int n_inst, const ItemOut *__restrict__ io,
                         const InstOut *__restrict__ ipss, const Pools P, ocrt_result_t *__restrict__ res)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_inst) return;
    const DevInst &I = inst[i];
    const LevelView lv(P, I);
    const int mm = I.mmax;
    const double phi = I.raa * OCRT_PI / 180.0 + OCRT_PI;
    double Iv = io[I.item0].mode[0], Qv = io[I.item0].mode[1], Uv = 0.0, Iboa = io[I.item0].boa_view;
    int maxo = 0, conv = 1, status = 0;
    for (int m = 0; m <= mm; ++m) {
        const ItemOut &r = io[I.item0 + m];
        if (m > 0) {
            const double cm = cos((double)m * phi), sm = sin((double)m * phi);
            Iv += 2.0 * r.mode[0] * cm;
            Qv += 2.0 * r.mode[1] * cm;
            Uv += 2.0 * r.mode[2] * sm;
            Iboa += 2.0 * r.boa_view * cm;
        }
        if (r.orders > maxo) maxo = r.orders;
        if (!r.converged) conv = 0;
    }
    ocrt_result_t o;
    memset(&o, 0, sizeof(o));
    const double mus = I.mus, muv = I.muv;
    o.I_TOA = Iv; o.Q_TOA = Qv; o.U_TOA = Uv;
    o.rho_I = Iv / mus; o.rho_Q = Qv / mus; o.rho_U = Uv / mus;
    o.tau_R_total = I.tauR;
    double kappa = 1.0, tdir = -1.0;
    if (I.ipss) { kappa = ipss[i].kappa; tdir = ipss[i].tau_dir_ipss; }
    o.kappa = kappa;
    if (I.ipss) {
        o.rho_I *= kappa; o.rho_Q *= kappa; o.rho_U *= kappa;
        o.I_TOA *= kappa; o.Q_TOA *= kappa; o.U_TOA *= kappa;
    }
    const double tau = lv.h[I.nt];
    const double phv = (I.raa - 180.0) * OCRT_PI / 180.0;
    const int bfo = (I.surface_kind == 2);
    if (bfo && I.decouple == 0) {
        double rg[3];
        sunglint_rho(I.vza > 0.0 ? muv : 1.0, phv, mus, I.wind, I.sigma_type, I.n_water, tau, I.q_conv, rg);
        if (I.ipss && tdir >= 0.0) { const double c = exp(-(tdir - tau / mus)); rg[0] *= c; rg[1] *= c; rg[2] *= c; }
        o.rho_I += rg[0]; o.rho_Q += rg[1]; o.rho_U += -rg[2];
    }
    o.n_orders_used = maxo;
    o.converged = conv;
    o.T_dir_dn = (I.ipss && tdir >= 0.0) ? exp(-tdir) : exp(-tau / mus);
    o.T_diff_dn_hemi = (io[I.item0].boa_hemi * (2.0 * OCRT_PI)) / (OCRT_PI * mus);
    o.T_total_dn_hemi = o.T_dir_dn + o.T_diff_dn_hemi;
    o.T_diff_dn_dir = (2.0 * OCRT_PI * muv * Iboa) / (OCRT_PI * mus);
    o.T_total_up_dir = o.rho_I * mus;
    o.T_dir_up_view = exp(-tau / (muv > 0.0 ? muv : 1.0));
    o.T_sg_up_dir = 0.0;
    o.rho_glint_direct_I = o.rho_glint_direct_Q = o.rho_glint_direct_U = 0.0;
    if (bfo && I.wind > 0.0) {
        double rg[3];
        sunglint_rho(I.vza > 0.0 ? muv : 1.0, phv, mus, I.wind, I.sigma_type, I.n_water, tau, I.q_conv, rg);
        if (I.ipss && tdir >= 0.0) { const double c = exp(-(tdir - tau / mus)); rg[0] *= c; rg[1] *= c; rg[2] *= c; }
        o.T_sg_up_dir = rg[0] * mus;
        o.rho_glint_direct_I = rg[0]; o.rho_glint_direct_Q = rg[1]; o.rho_glint_direct_U = rg[2];
    }
    if (!(o.rho_I == o.rho_I) || !(o.rho_Q == o.rho_Q) || !(o.rho_U == o.rho_U) || !isfinite(o.rho_I))
        status = -2;
    if (I.input_bad) status = OCRT_ERR_ARGS;
    o.status = status;
    o.n_ring = I.n;
    res[i] = o;
}

// explicit instantiations
template __global__ void order_one<double, false>(const DevInst *, const DevItem *, Pools, ItemState *,
                                                  const ColTile *, int, const double *,
                                                  const double *);
template __global__ void order_one<__hip_bfloat16, true>(const DevInst *, const DevItem *, Pools, ItemState *,
                                                         const ColTile *, int, const double *,
                                                  const double *);
template __global__ void order_status<double>(const DevInst *, const DevItem *, Pools, ItemState *, int);
template __global__ void order_status<__hip_bfloat16>(const DevInst *, const DevItem *, Pools, ItemState *, int);
#define OCRT_SWEEP_INST(FT, JT, MX, U, W)                                                                   \
    template __global__ void sweep<FT, JT, MX, U, W>(const DevInst *, const DevItem *, Pools, ItemState *,    \
                                                     const ColTile *, int, int, const double *, const float *,   \
                                                     const float *);
OCRT_SWEEP_INST(double, double, false, 8, 4)
OCRT_SWEEP_INST(__hip_bfloat16, _Float16, true, 4, 2)
OCRT_SWEEP_INST(__hip_bfloat16, _Float16, true, 4, 4)
OCRT_SWEEP_INST(__hip_bfloat16, _Float16, true, 8, 2)
OCRT_SWEEP_INST(__hip_bfloat16, _Float16, true, 8, 4)
OCRT_SWEEP_INST(__hip_bfloat16, _Float16, true, 16, 2)
OCRT_SWEEP_INST(__hip_bfloat16, _Float16, true, 16, 4)
#undef OCRT_SWEEP_INST
#define OCRT_GEMM_INST(NP, SP, TM, TN)                                                                     \
    template __global__ void gemm_bf16<NP, SP, TM, TN>(const DevInst *, const DevItem *, Pools, const GemmTile *, \
                                                       const ItemState *, int);
OCRT_GEMM_INST(1, false, 64, 128)
OCRT_GEMM_INST(1, true, 64, 128)
OCRT_GEMM_INST(2, true, 64, 128)
OCRT_GEMM_INST(1, false, 64, 64)
OCRT_GEMM_INST(1, true, 64, 64)
OCRT_GEMM_INST(2, true, 64, 64)
OCRT_GEMM_INST(1, false, 128, 64)
OCRT_GEMM_INST(1, true, 128, 64)
OCRT_GEMM_INST(2, true, 128, 64)
OCRT_GEMM_INST(1, false, 128, 128)
OCRT_GEMM_INST(1, true, 128, 128)
OCRT_GEMM_INST(2, true, 128, 128)
#undef OCRT_GEMM_INST
template __global__ void gemm_rebuild<1>(const DevInst *, const DevItem *, Pools, const GemmTile *, const ItemState *,
                                          int);
template __global__ void gemm_rebuild<2>(const DevInst *, const DevItem *, Pools, const GemmTile *, const ItemState *,
                                          int);
template __global__ void gemm_adj<1, kGemmParts>(const DevInst *, const DevItem *, Pools, const GemmTile *,
                                                      const ItemState *, int);
template __global__ void gemm_adj<2, kGemmParts>(const DevInst *, const DevItem *, Pools, const GemmTile *,
                                                      const ItemState *, int);
template __global__ void gemm_adj<1, kGemmMix>(const DevInst *, const DevItem *, Pools, const GemmTile *,
                                                    const ItemState *, int);
template __global__ void gemm_adj<2, kGemmMix>(const DevInst *, const DevItem *, Pools, const GemmTile *,
                                                    const ItemState *, int);
template __global__ void gemm_f64<kGemmMix>(const DevInst *, const DevItem *, Pools, const GemmTile *,
                                            const ItemState *, int);
template __global__ void gemm_f64<kGemmParts>(const DevInst *, const DevItem *, Pools, const GemmTile *,
                                              const ItemState *, int);
template __global__ void gemm_f64<kGemmBoth>(const DevInst *, const DevItem *, Pools, const GemmTile *,
                                              const ItemState *, int);

}  // namespace ocrt
