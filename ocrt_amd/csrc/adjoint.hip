// Adjoint of the SOS solve.  For one Fourier mode the forward solve is
//   u_1 = sweep(S_1) (+ rough-surface seed),
//   u_o = sweep(M u_{o-1}, R B u_{o-1}),  o = 2..N,
//   X   = sum_o u_o at the TOA view node (minus the rough seed there),
// linear in the fields, so the adjoint runs the order recursion backward:
//   lambda_N = e,  lambda_{o-1} = e + M^T Sbar_o + B^T R^T bbar_o,
// where e carries the reconstruction weights of the TOA view entries and
// (Sbar_o, bbar_o) is the adjoint sweep of lambda_o.  The forward fields are
// kept for every order, or checkpointed every 2 or 4 orders and recomputed
// segment by segment (plan.hip).
// Two storage forms share these kernels: fp64 plans keep fp64 fields and an
// fp64 adjoint (FieldF64, double); mixed plans keep every forward field as its
// bf16 pair and the adjoint fields in fp32 (FieldPair, float).  The arithmetic
// of every kernel here is fp64 in both.
// Gradients accumulate per instance with respect to the level optical depths
// (explicit dependence gH, layer steps gD), the scattering fractions gX, gY
// and the wind speed gW.  The reference has no adjoint; the finite-difference
// checks of tests/ and the oracle's reverse mode pin this one.
#include "ocrt_amd.h"
#include "sos_common.h"

namespace ocrt {

template <typename T>
__global__ void fill(T *__restrict__ p, long long n, T v)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

// Reconstruction weight of mode m for Stokes component c over 1 / mu_s
// (rt_fourier.c:9-33, rho = X / mu_s).
__device__ inline double seed_weight(const DevInst &I, const Pools &P, int m, int c)
{
    const double phi = I.raa * OCRT_PI / 180.0 + OCRT_PI;
    double wt;
    if (c < 2) wt = (m == 0) ? 1.0 : 2.0 * cos(m * phi);
    else wt = (m == 0) ? 0.0 : 2.0 * sin(m * phi);
    return P.gout[3LL * I.gidx + c] * wt / I.mus;
}

// lambda += e at the TOA view entries of one item (LT: lambda's storage).
template <typename LT>
__device__ inline void add_seed(const DevInst &I, const DevItem &it, const Pools &P)
{
    LT *lam = reinterpret_cast<LT *>(P.lam) + it.off_F;
    for (int c = 0; c < 3; ++c) lam[col_up(it.n, c, I.vidx)] += (LT)seed_weight(I, P, it.m, c);
}

// Zero the instance accumulators (one thread per entry of one instance).
__global__ void adj_zero_grad(const DevInst *__restrict__ inst, Pools P)
{
    const DevInst &I = inst[blockIdx.y];
    const int L = I.nt + 1, cnt = 4 * L;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < cnt; e += gridDim.x * blockDim.x)
        P.grad[I.off_grad + e] = 0.0;
}

template <typename LT>
__global__ void adj_seed(const DevInst *__restrict__ inst, const DevItem *__restrict__ items, Pools P)
{
    const DevItem &it = items[blockIdx.x];
    if (threadIdx.x == 0) add_seed<LT>(inst[it.inst], it, P);
}

// Bottom downward field of the field in P.F into bc + 3n (order `order`
// reads the field of order - 1; FV: the field's storage).
template <typename FV>
__global__ void adj_bottom(const DevInst *__restrict__ inst, const DevItem *__restrict__ items, Pools P,
                           const ItemState *__restrict__ st, int order)
{
    const DevItem &it = items[blockIdx.x];
    if (!item_runs(st[blockIdx.x], order)) return;
    const int n = it.n;
    const FV F(P, it, 0);
    const long long bot = (long long)it.nt * it.P;
    for (int q = threadIdx.x; q < 3 * n; q += blockDim.x) {
        const int c = q / n, o = q % n;
        P.bc[it.off_bc + 3 * n + q] = F[bot + col_dn(n, c, o)];
    }
}

__global__ void adj_surface_bc(const DevInst *__restrict__ inst, const DevItem *__restrict__ items, Pools P,
                               const ItemState *__restrict__ st, int order)
{
    const DevItem &it = items[blockIdx.x];
    const DevInst &I = inst[it.inst];
    if (!item_runs(st[blockIdx.x], order) || I.surf == 0) return;
    surface_bc_item(I, it, P);
}

// Adjoint plans: the coefficients of one scan step over layer k at node o
// (step_coefs: Beer-Lambert factor, the linear-source weights of the two end
// levels and their derivatives by the layer depth), once per layer and node
// instead of once per column and order.
__global__ void adj_step_table(const DevInst *__restrict__ inst, Pools P)
{
    const DevInst &I = inst[blockIdx.y];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= I.nt) return;
    const double *h = P.lev + I.off_lev;
    const double *mu = P.ring + I.off_ring;
    const double dt = fabs(h[k] - h[k + 1]);
    double *c = P.sc + 6 * (I.off_T + (long long)k * I.n);
    float *cf = P.scf ? P.scf + 6 * (I.off_T + (long long)k * I.n) : nullptr;
    for (int o = 0; o < I.n; ++o) {
        step_coefs(dt, mu[o], c[6 * o], c[6 * o + 1], c[6 * o + 2], c[6 * o + 3], c[6 * o + 4], c[6 * o + 5]);
        if (cf)
            for (int e = 0; e < 6; ++e) cf[6 * o + e] = (float)c[6 * o + e];
    }
}

// Adjoint sweep of order `order` >= 2.  P.F holds u_o (FV: its storage), P.J /
// P.J2 the unscaled operator parts A_0 u_{o-1} / A_1 u_{o-1}, P.lam lambda_o
// (both LT: double in fp64 plans, float in mixed ones; the arithmetic is fp64).  Writes
// Sbar_o (fp64 plans: over P.J; mixed plans: as its bf16 pair to P.sb_hi /
// P.sb_lo, the form the A^T product reads), bbar_o over the boundary slots of P.bc, and accumulates
// gX, gY (Sbar . part) per level and gD per layer.  One thread per column,
// one 64-column tile (of one direction) per wavefront, walking each column
// against its forward direction; the per-level sums of eight steps are
// reduced over the wavefront together.
template <typename FV, typename LT>
__global__ __launch_bounds__(256) void sweep_adj(const DevInst *__restrict__ inst, const DevItem *__restrict__ items,
                                                 Pools P, const ItemState *__restrict__ st,
                                                 const ColTile *__restrict__ tiles, int n_tiles, int order,
                                                 const double *__restrict__ lev)
{
    const int tix = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (tix >= n_tiles) return;
    const ColTile ct = tiles[tix];
    const int item = __builtin_amdgcn_readfirstlane(ct.item);
    if (item < 0 || !item_runs(st[item], order)) return;
    const DevItem &it = items[item];
    const DevInst &I = inst[it.inst];
    const int n = it.n, nt = it.nt, Pp = it.P, L = nt + 1;
    const int nparts = it.K / Pp, kind0 = it.parts & 0xff;
    const long long off_F = it.off_F, off_bc = it.off_bc;
    if (ct.col0 >= 6 * n) return;   // a tile of pad columns
    const int col = ct.col0 + lane;
    const bool live = col < ct.col_end && col < 6 * n;
    // the dead lanes of a partial tile load a live column of the item (their
    // values reach no output), so every load is unconditional and a block's
    // loads are all in flight before its first step
    const int colr = live ? col : ct.col0;
    // tiles never mix directions: the whole wavefront walks one way
    const int hemi = __builtin_amdgcn_readfirstlane(ct.col0 >= 3 * n ? 1 : 0);
    int c = 0, hh = 0, o = 0;
    col_decode(colr, n, c, hh, o);
    // the level tables through the __restrict__ argument (P.lev): scalar loads
    const LevelView lv(lev, I);
    const FV U(P, it, colr);
    const bool two = nparts > 1;
    LT *__restrict__ J0 = reinterpret_cast<LT *>(P.J) + off_F + colr;
    // single-part items have no second part: read (and drop) the first
    const LT *__restrict__ J1 = two ? reinterpret_cast<const LT *>(P.J2) + off_F + colr : J0;
    const LT *__restrict__ lam = reinterpret_cast<const LT *>(P.lam) + off_F + colr;
    double *gX = P.grad + I.off_grad + L, *gY = gX + L, *gD = gY + L;
    constexpr bool MX = sizeof(LT) == 4;
    auto put_sbar = [&](int k, double v) {
        const long long ix = (long long)k * Pp;
        if constexpr (MX) {
            // the fp32 value's bf16 pair, rounded to nearest even on the bits
            // (finite values; one fp64 -> fp32 conversion, no double rounding)
            const float vf = (float)v;
            const unsigned ub = __float_as_uint(vf);
            const unsigned hb = (ub + 0x7fffu + ((ub >> 16) & 1u)) & 0xffff0000u;
            const unsigned ur = __float_as_uint(vf - __uint_as_float(hb));
            const unsigned lb = (ur + 0x7fffu + ((ur >> 16) & 1u)) >> 16;
            reinterpret_cast<unsigned short *>(P.sb_hi)[off_F + colr + ix] = (unsigned short)(hb >> 16);
            reinterpret_cast<unsigned short *>(P.sb_lo)[off_F + colr + ix] = (unsigned short)lb;
        } else {
            J0[ix] = (LT)v;
        }
    };
    auto source = [&](int k, double p0, double p1) {
        const double s0 = (kind0 == 1) ? lv.xd[k] : lv.yd[k];
        return s0 * p0 + lv.yd[k] * p1;
    };
    // d(Sbar . J) / d xdel and / d ydel at one level, summed over the wavefront
    auto level_sums = [&](int k, double sbar, double p0, double p1) {
        double vx, vy;
        if (kind0 == 1) { vx = sbar * p0; vy = sbar * p1; }
        else { vx = 0.0; vy = sbar * p0; }
        vx = wave_sum(live ? vx : 0.0);
        vy = wave_sum(live ? vy : 0.0);
        if (lane == 0) {
            if (vx != 0.0) atomicAdd(&gX[k], vx);
            if (vy != 0.0) atomicAdd(&gY[k], vy);
        }
    };
    const int k0 = hemi ? nt : 0;
    double rho = (double)lam[(long long)k0 * Pp];
    double pk0 = (double)J0[(long long)k0 * Pp];
    const double j1k0 = (double)J1[(long long)k0 * Pp];
    double pk1 = two ? j1k0 : 0.0;
    double sk = source(k0, pk0, pk1);
    double pending = 0.0;
    constexpr int kB = 8;
    for (int t0 = 0; t0 < nt; t0 += kB) {
        // gather the block's upstream levels first (clamped: no branches)
        double q0[kB], q1[kB], uv[kB], lv_[kB];
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            const int tc = t0 + u < nt ? t0 + u : nt - 1;
            const int kp = hemi ? nt - tc - 1 : tc + 1;
            const long long ix = (long long)kp * Pp;
            q0[u] = (double)J0[ix];
            const double j1 = (double)J1[ix];
            q1[u] = two ? j1 : 0.0;
            uv[u] = U[ix];
            lv_[u] = (double)lam[ix];
        }
        // mixed plans: the block's step coefficients from the fp32 table, with
        // its other loads (no load waits inside the steps)
        float cfv[MX ? 6 * kB : 1];
        if constexpr (MX) {
#pragma unroll
            for (int u = 0; u < kB; ++u) {
                const int tc = t0 + u < nt ? t0 + u : nt - 1;
                const long long layer = hemi ? nt - tc - 1 : tc;   // the layer between k and kp
                const float2 *c2 = reinterpret_cast<const float2 *>(P.scf + 6 * (I.off_T + layer * n + o));
                const float2 a = c2[0], b = c2[1], e = c2[2];
                cfv[6 * u + 0] = a.x; cfv[6 * u + 1] = a.y; cfv[6 * u + 2] = b.x;
                cfv[6 * u + 3] = b.y; cfv[6 * u + 4] = e.x; cfv[6 * u + 5] = e.y;
            }
        }
        // per step: d/d xdel, d/d ydel at level k and d/d(layer step)
        double red[3 * kB];
#pragma unroll
        for (int e = 0; e < 3 * kB; ++e) red[e] = 0.0;
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            const int t = t0 + u;
            if (t >= nt) break;
            const int k = hemi ? nt - t : t;
            const int kp = hemi ? k - 1 : k + 1;
            const double sp = source(kp, q0[u], q1[u]);
            // the step's coefficients over the layer between k and kp (table)
            double T, al, be, dT, dal, dbe;
            if constexpr (MX) {
                T = cfv[6 * u]; al = cfv[6 * u + 1]; be = cfv[6 * u + 2];
                dT = cfv[6 * u + 3]; dal = cfv[6 * u + 4]; dbe = cfv[6 * u + 5];
            } else {
                const double *cf = P.sc + 6 * (I.off_T + (long long)(k < kp ? k : kp) * n + o);
                T = cf[0]; al = cf[1]; be = cf[2]; dT = cf[3]; dal = cf[4]; dbe = cf[5];
            }
            const double sbar = pending + al * rho;
            pending = be * rho;
            if (live) {
                put_sbar(k, sbar);
                red[3 * u + 0] = (kind0 == 1) ? sbar * pk0 : 0.0;
                red[3 * u + 1] = (kind0 == 1) ? sbar * pk1 : sbar * pk0;
                red[3 * u + 2] = rho * (dT * uv[u] + dal * sk + dbe * sp);
            }
            rho = lv_[u] + T * rho;
            pk0 = q0[u]; pk1 = q1[u]; sk = sp;
        }
        wave_sum_8xC_lds<3>(red, lane);
        const int u = lane >> 3, t = t0 + u;
        if ((lane & 7) == 0 && t < nt) {
            const int k = hemi ? nt - t : t;
            if (red[0] != 0.0) atomicAdd(&gX[k], red[0]);
            if (red[1] != 0.0) atomicAdd(&gY[k], red[1]);
            if (red[2] != 0.0) atomicAdd(&gD[hemi ? k - 1 : k], red[2]);
        }
    }
    // last level of the walk: nt for upward columns, 0 for downward ones
    const int kl = hemi ? 0 : nt;
    if (live) put_sbar(kl, pending);
    level_sums(kl, pending, pk0, pk1);
    // the upward columns start from the boundary condition: its adjoint is rho
    if (live && hemi == 0 && I.surf != 0) P.bc[off_bc + c * n + o] = rho;
}

// After M^T Sbar_o has been written to P.lam: lambda_{o-1} += e (TOA view
// entries) + B^T R^T bbar_o (bottom downward entries), and the wind gradient
// bbar_o . dR/dwind . u_{o-1}(bottom), with P.F = u_{o-1}.  One workgroup
// (64 threads) per item.
template <typename FV, typename LT>
__global__ void surface_adj(const DevInst *__restrict__ inst, const DevItem *__restrict__ items, Pools P,
                            const ItemState *__restrict__ st, int order)
{
    const DevItem &it = items[blockIdx.x];
    if (!item_runs(st[blockIdx.x], order)) return;
    const DevInst &I = inst[it.inst];
    if (threadIdx.x == 0) add_seed<LT>(I, it, P);
    if (I.surf == 0) return;
    const int n = it.n, nt = it.nt, m = it.m, L = nt + 1;
    const double *mu = P.ring + I.off_ring, *w = mu + n;
    const double *bb = P.bc + it.off_bc;   // [3][n] adjoint of the boundary condition
    LT *lamb = reinterpret_cast<LT *>(P.lam) + it.off_F + (long long)nt * it.P;
    const FV Ub(P, it, 0);
    const long long bot = (long long)nt * it.P;
    const double W = (I.surf == 1) ? 2.0 * OCRT_PI : ((m == 0) ? 2.0 * OCRT_PI : OCRT_PI);
    double gw = 0.0;
    if (I.surf == 1) {
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            double acc[3];
            double Fr[9];
            fresnel_R(mu[i], I.n_water, I.q_conv, Fr);
            const double f = (w[i] == 0.0) ? 1.0
                             : (W * mu[i] * w[i]) * (1.0 / (2.0 * OCRT_PI * mu[i] * w[i]));
            for (int c2 = 0; c2 < 3; ++c2)
                acc[c2] = f * (bb[i] * Fr[c2] + bb[n + i] * Fr[3 + c2] + bb[2 * n + i] * Fr[6 + c2]);
            for (int c2 = 0; c2 < 3; ++c2) lamb[col_dn(n, c2, i)] += (LT)acc[c2];
        }
    } else {
        // rough surface, one wavefront: for every incoming node i the lanes
        // take the outgoing nodes op, whose kernel entries R^m[op][i] are
        // consecutive in op; eight nodes' three sums are reduced together
        const int lane = threadIdx.x & 63;
        for (int i0 = 0; i0 < n; i0 += 8) {
            double v[8 * 3];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int i = i0 + j < n ? i0 + j : n - 1;
                double a0 = 0.0, a1 = 0.0, a2 = 0.0;
                if (i0 + j < n && w[i] != 0.0) {
                    const double pw = W * mu[i] * w[i];
                    const double u0 = Ub[bot + col_dn(n, 0, i)], u1 = Ub[bot + col_dn(n, 1, i)],
                                 u2 = Ub[bot + col_dn(n, 2, i)];
                    // entry kl of R^m[op][i] at R[kl n + op]
                    const double *R = P.R + r_index(I, m, 0, i, 0);
                    const double *dR = P.dR + r_index(I, m, 0, i, 0);
                    for (int op = lane; op < n; op += 64) {
                        const double b0 = bb[op], b1 = bb[n + op], b2 = bb[2 * n + op];
                        a0 += b0 * R[op] + b1 * R[3 * n + op] + b2 * R[6 * n + op];
                        a1 += b0 * R[n + op] + b1 * R[4 * n + op] + b2 * R[7 * n + op];
                        a2 += b0 * R[2 * n + op] + b1 * R[5 * n + op] + b2 * R[8 * n + op];
                        gw += pw * ((b0 * dR[op] + b1 * dR[3 * n + op] + b2 * dR[6 * n + op]) * u0 +
                                    (b0 * dR[n + op] + b1 * dR[4 * n + op] + b2 * dR[7 * n + op]) * u1 +
                                    (b0 * dR[2 * n + op] + b1 * dR[5 * n + op] + b2 * dR[8 * n + op]) * u2);
                    }
                    a0 *= pw; a1 *= pw; a2 *= pw;
                }
                v[3 * j] = a0; v[3 * j + 1] = a1; v[3 * j + 2] = a2;
            }
            wave_sum_8xC_lds<3>(v, lane);
            const int i = i0 + (lane >> 3);
            if ((lane & 7) == 0 && i < n)
                for (int c2 = 0; c2 < 3; ++c2) lamb[col_dn(n, c2, i)] += (LT)v[c2];
        }
    }
    if (I.surf == 2) {
        gw = wave_sum(gw);
        if (threadIdx.x == 0 && gw != 0.0)
            atomicAdd(&P.grad[I.off_grad + 4 * L - 1], gw * slope_variance_dwind(I.wind, I.sigma_type));
    }
}

// Adjoint of order 1 with P.F = u_1 and P.lam = lambda_1: the adjoint sweep of
// the combined single-scattering and flat-seed source, the explicit
// dependence of that source on xdel, ydel and h, and the rough direct-beam
// seed (which enters u_1 and leaves the TOA output).
template <typename FV, typename LT>
__global__ __launch_bounds__(256) void order_one_adj(const DevInst *__restrict__ inst,
                                                     const DevItem *__restrict__ items, Pools P,
                                                     const ColTile *__restrict__ tiles, int n_tiles,
                                                     const double *__restrict__ lev)
{
    const int tix = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (tix >= n_tiles) return;
    const ColTile ct = tiles[tix];
    const int item = __builtin_amdgcn_readfirstlane(ct.item);
    if (item < 0) return;
    const DevItem &it = items[item];
    const DevInst &I = inst[it.inst];
    const int n = it.n, nt = it.nt, Pp = it.P, L = nt + 1, m = it.m;
    if (ct.col0 >= 6 * n) return;   // a tile of pad columns
    const int col = ct.col0 + lane;
    const bool live = col < ct.col_end && col < 6 * n;
    // dead lanes load a live column (unconditional loads); lmask zeroes their
    // adjoint, the only value of theirs that reaches the reductions unmasked
    const int colr = live ? col : ct.col0;
    const double lmask = live ? 1.0 : 0.0;
    // tiles never mix directions: the whole wavefront walks one way
    const int hemi = __builtin_amdgcn_readfirstlane(ct.col0 >= 3 * n ? 1 : 0);
    int c = 0, hh = 0, o = 0;
    col_decode(colr, n, c, hh, o);
    // the level tables through the __restrict__ argument (P.lev): scalar loads
    const LevelView lv(lev, I);
    const ColumnSource src(I, it, P, lv, colr);
    const double mus = I.mus, amu = P.ring[I.off_ring + o];
    const FV U(P, it, colr);
    const LT *__restrict__ lam = reinterpret_cast<const LT *>(P.lam) + it.off_F + colr;
    double *gH = P.grad + I.off_grad, *gX = gH + L, *gY = gX + L, *gD = gY + L, *gW = gD + nt;
    // rough direct-beam seed of the upward columns
    const bool rough = live && I.surf == 2 && hemi == 0;
    const double tau = lv.h[nt], down = exp(-tau / mus);
    const double FS = ((m == 0) ? OCRT_PI : 0.5 * OCRT_PI) * mus;
    const double rsun = rough ? P.R[I.off_Rsun + ((long long)m * n + o) * 9 + 3 * c] : 0.0;
    const double drsun = rough ? P.dR[I.off_Rsun + ((long long)m * n + o) * 9 + 3 * c] : 0.0;
    const double e_here = (live && hemi == 0 && o == I.vidx) ? seed_weight(I, P, m, c) : 0.0;
    auto seed_f = [&](int k) { return FS * (down * exp(-(tau - lv.h[k]) / amu)); };
    auto S = [&](int k) { return live ? src.primary(lv, k) + src.flat_value(lv, mus, k) : 0.0; };
    double g_nt = 0.0, g_w = 0.0;   // per-lane d/dh[nt] and d/dsigma^2 terms
    // explicit level terms of the source and the seed at level k for adjoint
    // sbar; lam_k is lambda_1 at level k, f the seed's attenuation seed_f(k)
    auto level_terms = [&](int k, double sbar, double lam_k, double f, double &vx, double &vy, double &vh) {
        vx = vy = vh = 0.0;
        if (!live) return;
        const double ch = lv.ch[k];
        vx = sbar * src.sgn * ch * src.pc;
        vy = sbar * src.sgn * ch * src.mc;
        vh = -sbar * src.primary(lv, k) / mus;
        if (src.flat && !src.flat_zero(k)) {
            const double b = src.flat_beam(lv, mus, k);
            vx += sbar * b * (src.RI * src.pA + src.RQ * src.pB);
            vy += sbar * b * (src.RI * src.mA + src.RQ * src.mB);
            const double fv = sbar * src.flat_value(lv, mus, k);
            vh += fv / mus;
            g_nt += -2.0 * fv / mus;
        }
        if (rough) {
            const double gd = lam_k - (k == 0 ? e_here : 0.0);
            const double dv = rsun * f;
            vh += gd * dv / amu;
            g_nt += -gd * dv * (1.0 / mus + 1.0 / amu);
            g_w += gd * drsun * f;
        }
    };
    const int k0 = hemi ? nt : 0;
    double rho = (double)lam[(long long)k0 * Pp] * lmask;
    double lam_k = rho;
    double sk = S(k0);
    double pending = 0.0;
    // the seed's attenuation at the current level, carried to the next step
    double fk = rough ? seed_f(k0) : 0.0;
    constexpr int kB = 8;
    for (int t0 = 0; t0 < nt; t0 += kB) {
        // gather the block's upstream levels first (clamped: no branches)
        double uv[kB], lv_[kB];
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            const int tc = t0 + u < nt ? t0 + u : nt - 1;
            const long long ix = (long long)(hemi ? nt - tc - 1 : tc + 1) * Pp;
            uv[u] = U[ix];
            lv_[u] = (double)lam[ix] * lmask;
        }
        // mixed plans: the block's step coefficients from the fp32 table, with
        // its other loads
        constexpr bool MX = sizeof(LT) == 4;
        float cfv[MX ? 6 * kB : 1];
        if constexpr (MX) {
#pragma unroll
            for (int u = 0; u < kB; ++u) {
                const int tc = t0 + u < nt ? t0 + u : nt - 1;
                const long long layer = hemi ? nt - tc - 1 : tc;   // the layer between k and kp
                const float2 *c2 = reinterpret_cast<const float2 *>(P.scf + 6 * (I.off_T + layer * n + o));
                const float2 a = c2[0], b = c2[1], e = c2[2];
                cfv[6 * u + 0] = a.x; cfv[6 * u + 1] = a.y; cfv[6 * u + 2] = b.x;
                cfv[6 * u + 3] = b.y; cfv[6 * u + 4] = e.x; cfv[6 * u + 5] = e.y;
            }
        }
        // per step: d/d xdel, d/d ydel, d/dh at level k and d/d(layer step)
        double red[4 * kB];
#pragma unroll
        for (int e = 0; e < 4 * kB; ++e) red[e] = 0.0;
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            const int t = t0 + u;
            if (t >= nt) break;
            const int k = hemi ? nt - t : t;
            const int kp = hemi ? k - 1 : k + 1;
            const double sp = S(kp);
            // the step's coefficients over the layer between k and kp (table)
            double T, al, be, dT, dal, dbe;
            if constexpr (MX) {
                T = cfv[6 * u]; al = cfv[6 * u + 1]; be = cfv[6 * u + 2];
                dT = cfv[6 * u + 3]; dal = cfv[6 * u + 4]; dbe = cfv[6 * u + 5];
            } else {
                const double *cf = P.sc + 6 * (I.off_T + (long long)(k < kp ? k : kp) * n + o);
                T = cf[0]; al = cf[1]; be = cf[2]; dT = cf[3]; dal = cf[4]; dbe = cf[5];
            }
            const double sbar = pending + al * rho;
            pending = be * rho;
            const double fkp = rough ? seed_f(kp) : 0.0;
            const double rk = live ? uv[u] - rsun * fkp : 0.0;
            level_terms(k, sbar, lam_k, fk, red[4 * u + 0], red[4 * u + 1], red[4 * u + 2]);
            fk = fkp;
            red[4 * u + 3] = rho * (dT * rk + dal * sk + dbe * sp);
            rho = lv_[u] + T * rho;
            lam_k = lv_[u];
            sk = sp;
        }
        wave_sum_8xC_lds<4>(red, lane);
        const int u = lane >> 3, t = t0 + u;
        if ((lane & 7) == 0 && t < nt) {
            const int k = hemi ? nt - t : t;
            if (red[0] != 0.0) atomicAdd(&gX[k], red[0]);
            if (red[1] != 0.0) atomicAdd(&gY[k], red[1]);
            if (red[2] != 0.0) atomicAdd(&gH[k], red[2]);
            if (red[3] != 0.0) atomicAdd(&gD[hemi ? k - 1 : k], red[3]);
        }
    }
    // last level of the walk: nt for upward columns, 0 for downward ones
    const int kl = hemi ? 0 : nt;
    double vx, vy, vh;
    level_terms(kl, pending, lam_k, fk, vx, vy, vh);
    vx = wave_sum(vx);
    vy = wave_sum(vy);
    vh = wave_sum(vh);
    const double snt = wave_sum(g_nt), sw = wave_sum(g_w);
    if (lane == 0) {
        if (vx != 0.0) atomicAdd(&gX[kl], vx);
        if (vy != 0.0) atomicAdd(&gY[kl], vy);
        if (vh != 0.0) atomicAdd(&gH[kl], vh);
        if (snt != 0.0) atomicAdd(&gH[nt], snt);
        if (sw != 0.0) atomicAdd(gW, sw * slope_variance_dwind(I.wind, I.sigma_type));
    }
}

// Direct glint in rho (black_fresnel_ocean with the glint coupled): its
// dependence on the column optical depth and the wind.  One thread per instance.
__global__ void glint_adj(const DevInst *__restrict__ inst, int n_inst, Pools P)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_inst) return;
    const DevInst &I = inst[i];
    if (I.surface_kind != OCRT_SURFACE_BLACK_FRESNEL_OCEAN || I.decouple) return;
    const LevelView lv(P, I);
    const int nt = I.nt, L = nt + 1;
    const double muv1 = I.vza > 0.0 ? I.muv : 1.0, phv = (I.raa - 180.0) * OCRT_PI / 180.0;
    double rg[3], dlog;
    sunglint_rho(muv1, phv, I.mus, I.wind, I.sigma_type, I.n_water, lv.h[nt], I.q_conv, rg, &dlog);
    const double *g = P.gout + 3LL * I.gidx;
    const double v = g[0] * rg[0] + g[1] * rg[1] - g[2] * rg[2];
    double *gH = P.grad + I.off_grad;
    gH[nt] += -v * (1.0 / I.mus + 1.0 / muv1);
    if (I.wind >= 1e-6) gH[4 * L - 1] += v * dlog * slope_variance_dwind(I.wind, I.sigma_type);
}

template __global__ void fill<double>(double *, long long, double);
template __global__ void fill<float>(float *, long long, float);
template __global__ void adj_seed<double>(const DevInst *, const DevItem *, Pools);
template __global__ void adj_seed<float>(const DevInst *, const DevItem *, Pools);
#define OCRT_ADJ_INST(FV, LT)                                                                                  \
    template __global__ void adj_bottom<FV>(const DevInst *, const DevItem *, Pools, const ItemState *, int);   \
    template __global__ void sweep_adj<FV, LT>(const DevInst *, const DevItem *, Pools, const ItemState *,      \
                                               const ColTile *, int, int, const double *);                     \
    template __global__ void surface_adj<FV, LT>(const DevInst *, const DevItem *, Pools, const ItemState *, int); \
    template __global__ void order_one_adj<FV, LT>(const DevInst *, const DevItem *, Pools, const ColTile *, int, \
                                                   const double *);
OCRT_ADJ_INST(FieldF64, double)
OCRT_ADJ_INST(FieldPair, float)
#undef OCRT_ADJ_INST

}  // namespace ocrt
