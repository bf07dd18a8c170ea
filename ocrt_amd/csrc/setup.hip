// Once-per-instance and once-per-mode kernels of the SOS solve, all fp64:
//   medium_levels / medium_fractions : uniform-tau level grid, US62 + aerosol
//       altitude mapping and the Beer-Lambert factors (rt_atm.c:177-400,
//       rt_molecular_profile.c, rt_layer_transport.c:36-87)
//   coxmunk_fourier : azimuthal Fourier kernels of the rough-ocean Mueller
//       BRDF for every mode at once (shared/surface.c:514-646, 839-938, 1263-1380)
//   angular_basis   : generalized-spherical-function recurrences
//       (rt_angular_basis.c:126-276)
//   phase_tables    : phase-matrix Fourier tables (rt_angular_basis.c:278-358)
//   pack_operator   : dense scattering operator with quadrature weights and
//       the Stokes sign pattern folded in (rt_sos_operator.c:50-126, 353-407)
#include <type_traits>

#include "common.h"
#include "device_math.h"
#include "wave.h"

namespace ocrt {

// ---------------------------------------------------------------------------
// Medium.  One thread per (instance, level).  The aerosol altitude of each
// level is found by bisection on [0, 100 km] independently per level (the
// reference walks levels sequentially with the previous altitude as the
// upper bracket; both converge to the same root to the last bits).
__global__ void medium_levels(const DevInst *__restrict__ inst, Pools P)
{
    const DevInst &I = inst[blockIdx.y];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int nt = I.nt;
    if (k > nt) return;
    const int L = nt + 1;
    double *h = P.lev + I.off_lev, *ch = h + L, *xd = ch + L, *yd = xd + L, *z = yd + L;
    if (I.water) {
        // in-water column (rt_water_rt.c:2156-2176): levels uniform in optical
        // depth, fixed particle and pure-water albedos at every level; the
        // solar source carries the transmitted beam's weight
        const double hk = (double)k * I.wtau / (double)nt;
        h[k] = hk;
        ch[k] = 0.5 * I.beam_scale * exp(-hk / I.mus);
        xd[k] = I.aer ? I.womega_p : 0.0;
        yd[k] = I.womega_w;
        z[k] = 1.0;
        return;
    }
    if (I.layered) {
        // running extinction sum from the top, in layer order, so every level
        // carries exactly the sequential prefix sum
        const double *R = P.layers + I.off_layers, *A = R + nt, *W = A + nt, *G = W + nt;
        double acc = 0.0;
        for (int l = 0; l < k; ++l) acc += R[l] + A[l] + G[l];
        h[k] = acc;
        ch[k] = 0.5 * exp(-acc / I.mus);
        z[k] = 0.0;
        if (k == 0) {
            if (I.tauR > 0.0) { xd[0] = 0.0; yd[0] = 1.0; }
            else { xd[0] = I.aer ? W[0] : 0.0; yd[0] = 0.0; }
        } else {
            const double dt = R[k - 1] + A[k - 1] + G[k - 1];
            xd[k] = I.aer ? W[k - 1] * A[k - 1] / dt : 0.0;
            yd[k] = R[k - 1] / dt;
        }
        return;
    }
    const double tau = I.tauR + (I.aer ? I.taua : 0.0);
    const double hk = (double)k * tau / (double)nt;
    h[k] = hk;
    ch[k] = 0.5 * exp(-hk / I.mus);
    double zk;
    if (!I.aer) {
        zk = (k == 0) ? 100.0 : us62_altitude_from_fraction(hk / tau);
        xd[k] = 0.0;
        yd[k] = 1.0;
    } else if (k == 0) {
        zk = 100.0;
    } else if (k == nt) {
        zk = 0.0;
    } else {
        const double H = I.aer_h > 0.0 ? I.aer_h : 2.0;
        double lo = 0.0, hi = 100.0;
        for (int it = 0; it < 80; ++it) {
            const double mid = 0.5 * (lo + hi);
            const double above = I.tauR * us62_grid_fraction_above(mid) +
                                 I.taua * aerosol_fraction_above(mid, 100.0, H);
            if (above > hk) lo = mid; else hi = mid;
        }
        zk = 0.5 * (lo + hi);
    }
    z[k] = zk;
}

// Scattering fractions of each level (aerosol case).
__global__ void medium_fractions(const DevInst *__restrict__ inst, Pools P)
{
    const DevInst &I = inst[blockIdx.y];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int nt = I.nt;
    if (k > nt) return;
    const int L = nt + 1;
    const double *h = P.lev + I.off_lev, *z = h + 4 * L;
    double *xd = P.lev + I.off_lev + 2 * L, *yd = xd + L;
    if (I.aer && !I.layered && !I.water) {
        if (k == 0) {
            if (I.tauR == 0.0 && I.taua > 0.0) { xd[0] = I.ssa; yd[0] = 0.0; }
            else { xd[0] = 0.0; yd[0] = 1.0; }
        } else {
            const double H = I.aer_h > 0.0 ? I.aer_h : 2.0;
            const double ray = I.tauR * us62_grid_fraction_above(z[k]);
            const double ae = I.taua * aerosol_fraction_above(z[k], 100.0, H);
            const double rp = (k == 1) ? 0.0 : I.tauR * us62_grid_fraction_above(z[k - 1]);
            const double ap = (k == 1) ? 0.0 : I.taua * aerosol_fraction_above(z[k - 1], 100.0, H);
            const double dr = ray - rp, da = ae - ap, dt = dr + da;
            if (dt > 0.0) { xd[k] = da * I.ssa / dt; yd[k] = dr / dt; }
            else { xd[k] = 0.0; yd[k] = (I.tauR > 0.0) ? 1.0 : 0.0; }
        }
    }
}

// ---------------------------------------------------------------------------
// Gas absorption (rt_atm.c:437-548, rt_absorption.c:344-455).  One workgroup
// per instance: every thread takes layers k (between levels k and k + 1,
// altitudes z[k] > z[k + 1]) and sums over the six gases sigma(z_mid, wl)
// times the gas column of the layer; then one thread adds the cumulative
// absorption to h and ch and rescales the scattering fractions by
// dtau_old / dtau_new, level by level in the reference's order.
__device__ inline double gas_sigma_row(const DevGas &G, const double *pool, int g, int layer, double wl)
{
    const int n = G.n_wl[g];
    if (!G.avail[g] || n < 2) return 0.0;
    layer = layer < 0 ? 0 : (layer >= G.n_layers[g] ? G.n_layers[g] - 1 : layer);
    const double *w = pool + G.off_wl[g], *s = pool + G.off_sig[g] + (long long)layer * n;
    if (wl < w[0] || wl > w[n - 1]) return __builtin_nan("");
    if (wl == w[0]) return s[0];
    if (wl == w[n - 1]) return s[n - 1];
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (w[mid] <= wl) lo = mid; else hi = mid;
    }
    const double u = (wl - w[lo]) / (w[hi] - w[lo]);
    return (1.0 - u) * s[lo] + u * s[hi];
}

__device__ inline double gas_sigma(const DevGas &G, const double *pool, int g, double z, double wl)
{
    if (!G.avail[g]) return 0.0;
    const int K = G.n_layers[g];
    if (K == 1) return gas_sigma_row(G, pool, g, 0, wl);
    const double *zg = pool + G.off_z;
    if (z <= zg[0]) return gas_sigma_row(G, pool, g, 0, wl);
    if (z >= zg[K - 1]) return gas_sigma_row(G, pool, g, K - 1, wl);
    int lo = 0, hi = K - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (zg[mid] <= z) lo = mid; else hi = mid;
    }
    const double dz = zg[hi] - zg[lo];
    if (dz <= 0.0) return gas_sigma_row(G, pool, g, lo, wl);
    const double u = (z - zg[lo]) / dz;
    return (1.0 - u) * gas_sigma_row(G, pool, g, lo, wl) + u * gas_sigma_row(G, pool, g, hi, wl);
}

// Column of gas g above altitude z (molecules cm^-2): trapezoid of the
// linearly interpolated number density over the AFGL levels.
__device__ inline double gas_column_above(const DevGas &G, const double *pool, int g, double z)
{
    const double *zz = pool + G.off_z, *nn = pool + G.off_n + (long long)g * G.n_lev;
    double cum = 0.0;
    for (int i = 0; i + 1 < G.n_lev; ++i) {
        const double z1 = zz[i], z2 = zz[i + 1];
        if (z2 <= z) continue;
        const double lo = z1 < z ? z : z1;
        if (z2 <= lo) continue;
        const double f1 = (lo - z1) / (z2 - z1), f2 = (z2 - z1) / (z2 - z1);
        const double a = (1.0 - f1) * nn[i] + f1 * nn[i + 1];
        const double b = (1.0 - f2) * nn[i] + f2 * nn[i + 1];
        cum += 0.5 * (a + b) * ((z2 - lo) * 1.0e5);
    }
    return cum;
}

__global__ __launch_bounds__(256) void medium_gas(const DevInst *__restrict__ inst, Pools P)
{
    const DevInst &I = inst[blockIdx.x];
    if (I.gas < 0 || I.layered) return;
    const DevGas &G = P.gas_tab[I.gas];
    const double *pool = P.gas_pool;
    const int nt = I.nt, L = nt + 1;
    double *h = P.lev + I.off_lev, *ch = h + L, *xd = ch + L, *yd = xd + L;
    const double *z = yd + L;
    double *dab = h + 5 * L;
    for (int k = threadIdx.x; k < nt; k += blockDim.x) {
        double lo = z[k + 1], hi = z[k];
        if (lo > hi) { const double t = lo; lo = hi; hi = t; }
        const double zm = 0.5 * (lo + hi);
        double acc = 0.0;
        for (int g = 0; g < 6; ++g) {
            double N;
            if (G.overridden[g]) {
                N = 0.0;
                if (G.col_def[g] > 0.0) {
                    const double Nd = hi > lo ? gas_column_above(G, pool, g, lo) - gas_column_above(G, pool, g, hi)
                                              : 0.0;
                    N = G.col_eff[g] * (Nd / G.col_def[g]);
                }
            } else {
                N = hi > lo ? gas_column_above(G, pool, g, lo) - gas_column_above(G, pool, g, hi) : 0.0;
            }
            acc += gas_sigma(G, pool, g, zm, I.wl) * N;
        }
        dab[k] = acc;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double cum = 0.0, h_prev = h[0];
    for (int k = 0; k <= nt; ++k) {
        const double h_old = h[k];
        h[k] = h_old + cum;
        ch[k] = 0.5 * exp(-h[k] / I.mus);
        if (k >= 1) {
            const double dt_old = h_old - h_prev, dt_new = dt_old + dab[k - 1];
            if (dt_new > 0.0 && dt_old > 0.0) {
                xd[k] = xd[k] * dt_old / dt_new;
                yd[k] = yd[k] * dt_old / dt_new;
            }
        }
        h_prev = h_old;
        if (k < nt) cum += dab[k];
    }
}

// Beer-Lambert factors T[layer][node] = exp(-dtau / mu_node) in fp64, 1 - T
// in fp32, and 1 / dtau in fp32 and in fp64 (row 6 of the level arrays).
__global__ void medium_transmit(const DevInst *__restrict__ inst, Pools P)
{
    const DevInst &I = inst[blockIdx.y];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int nt = I.nt;
    const double *h = P.lev + I.off_lev;
    if (k < nt) {
        const double *mu = P.ring + I.off_ring;
        double *T = P.T + I.off_T + (long long)k * I.n;
        float *Tf = P.Tf + I.off_T + (long long)k * I.n;
        const double dtau = h[k + 1] - h[k];
        P.idt[I.off_lev + k] = (float)(1.0 / dtau);
        P.lev[I.off_lev + 6LL * (nt + 1) + k] = 1.0 / dtau;
        for (int o = 0; o < I.n; ++o) {
            T[o] = exp(-dtau / fabs(mu[o]));
            // fp32 sweeps carry 1 - T: with dtau / mu ~ 1e-4 per layer, T itself
            // in fp32 would leave 1 - T with only ~1e-3 relative accuracy
            Tf[o] = (float)(-expm1(-dtau / fabs(mu[o])));
        }
    }
}

// cos(m phi_p), sin(m phi_p) on the midpoint azimuth rule, [m][2][kNPhi],
// and (wf non-null) their fp32 roundings in the same layout.
__global__ void fourier_weights(double *w, float *wf, int mmax)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = blockIdx.y;
    if (p >= kNPhi || m > mmax) return;
    const double phi = (p + 0.5) * (2.0 * OCRT_PI / (double)kNPhi);
    const double c = cos(m * phi), s = sin(m * phi);
    w[((long long)m * 2 + 0) * kNPhi + p] = c;
    w[((long long)m * 2 + 1) * kNPhi + p] = s;
    if (wf) {
        wf[((long long)m * 2 + 0) * kNPhi + p] = (float)c;
        wf[((long long)m * 2 + 1) * kNPhi + p] = (float)s;
    }
}

// cos(m phi_q) on the value kernel's half circle, [q][w_rows] (mode fastest):
// phi_q = (q + 1/2) 2 pi / 720, q < 360.
__global__ void value_weights(double *w, int w_rows)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= kVkHalf * w_rows) return;
    const int q = e / w_rows, m = e % w_rows;
    const double phi = 2.0 * OCRT_PI * ((double)q + 0.5) / (double)kVkPhi;
    w[e] = cos((double)m * phi);
}

// ---------------------------------------------------------------------------
// Rough-ocean Fourier kernels.  One workgroup per (instance, outgoing node,
// incoming node or the sun), over the pairs of the instances that own their
// kernels only.  The isotropic-slope BRDF is even in the
// relative azimuth for the II IQ QI QQ UU entries and odd for the others, so
// the midpoint rule over [0, 2 pi) folds onto its first half (phi_p and
// 2 pi - phi_p pair up) and only kNPhi / 2 azimuths are evaluated.  The
// shadowing factor is hoisted out of the azimuth loop.  Instances of a batch
// that share the ring, the sun and the surface (the bands of one geometry)
// share one kernel: only owners (surf_owner) compute it.
//
// DERIV also writes d R^m / d sigma^2 (pool dR, the layout of pool R) for the
// adjoint: every entry is proportional to the brdf, so its derivative is the
// entry times d ln(brdf) / d sigma^2.
//
// Thread t of the 256 evaluates the Mueller matrix at azimuths t and t + 256
// and keeps both in registers.  The Fourier sums are then taken in batches of
// 8 rows (8 modes, or 4 modes x {value, derivative}) x 9 entries: every thread
// reads its two samples of the cos / sin rows of the batch (consecutive
// azimuths over the lanes), forms its products, and transposed wavefront
// reductions (three entries at a time) leave each row's totals in one group of
// eight lanes; the four wavefronts' totals are added through LDS.
//
// RT = float (mixed-form plans, diffuse pairs only): the Mueller matrices, the
// Fourier products and the wavefront reductions in fp32 from the fp32 weight
// table, the four wavefronts' totals added in fp64, and only the fp32 copy
// Rf written (the mixed form reads no other copy of the diffuse kernels); the
// mixed adjoint plans keep no fp32 copy and take the fp64 arrays R and dR.
template <bool DERIV, typename RT = double>
__device__ inline void coxmunk_pair(const DevInst &I, const Pools &P, int pair)
{
    constexpr int kHalf = kNPhi / 2;
    static_assert(kHalf == 2 * kCoxMunkThreads, "two azimuths per thread");
    constexpr bool F32 = sizeof(RT) == 4;
    constexpr int kWaves = kCoxMunkThreads / 64;
    constexpr int kRowModes = DERIV ? 4 : 8;   // Fourier modes per batch
    __shared__ RT red[kWaves][8][9];
    const RT *cmw = F32 ? reinterpret_cast<const RT *>(P.cm_wf) : reinterpret_cast<const RT *>(P.cm_w);
    const int n = I.n;
    const double *mu = P.ring + I.off_ring;
    const double *w = mu + n;
    int o, i;
    double mo, mi;
    const bool solar = pair >= n * n;
    if (!solar) {
        o = pair / n; i = pair % n;
        if (w[i] == 0.0) {   // zero-weight incoming columns never enter the boundary sum
            const int nout = (I.mmax + 1) * 9;
            for (int t = threadIdx.x; t < nout; t += blockDim.x) {
                const int m = t / 9, kl = t % 9;
                if (!F32 || !P.Rf) P.R[r_index(I, m, o, i, kl)] = 0.0;
                if (P.Rf) P.Rf[r_index(I, m, o, i, kl)] = 0.f;
                if (DERIV) P.dR[r_index(I, m, o, i, kl)] = 0.0;
            }
            return;
        }
        mo = mu[o]; mi = mu[i];
    } else {
        o = pair - n * n; i = 0;
        mo = mu[o]; mi = I.mus;
    }
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const double dphi = 2.0 * OCRT_PI / (double)kNPhi;
    const double s2v = slope_variance(I.wind, I.sigma_type);
    const double shadow = sancer_shadow(mo, mi, s2v);
    const double dls = DERIV ? sancer_dlog(mo, mi, s2v, shadow) : 0.0;
    const CoxMunkPairT<RT> cm(mo, mi, s2v, shadow, I.n_water, I.q_conv);
    RT r[2][9];
    double dl[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        // azimuth (p + 1/2) dphi: its cosine and sine are the m = 1 row of the
        // Fourier weight table
        const int p = t + kCoxMunkThreads * q;
        RT tb2;
        cm.eval(cmw[2 * kNPhi + p], cmw[3 * kNPhi + p], r[q], tb2);
        dl[q] = (DERIV && tb2 >= 0.0) ? (double)tb2 / (s2v * s2v) - 1.0 / s2v + dls : 0.0;
    }
    const int M = I.mmax;
    // one reduction over the entries [E0, E0 + C) of a batch's eight rows
    auto reduce_entries = [&](int m0, auto e0, auto cnt) {
        constexpr int E0 = decltype(e0)::value, C = decltype(cnt)::value;
        RT v[8 * C];   // [row][entry - E0]
#pragma unroll
        for (int j = 0; j < kRowModes; ++j) {
            const int m = m0 + j <= M ? m0 + j : M;   // rows past the last mode are computed, not written
            const RT *wc = cmw + (long long)m * 2 * kNPhi, *ws = wc + kNPhi;
            const int t1 = t + kCoxMunkThreads;
            const RT c0 = wc[t], c1 = wc[t1], s0 = ws[t], s1 = ws[t1];
#pragma unroll
            for (int e = 0; e < C; ++e) {
                const int kl = E0 + e;
                const bool use_cos = (kl == 0 || kl == 1 || kl == 3 || kl == 4 || kl == 8);
                const RT a0 = r[0][kl] * (use_cos ? c0 : s0), a1 = r[1][kl] * (use_cos ? c1 : s1);
                if (DERIV) {
                    v[(2 * j) * C + e] = a0 + a1;
                    v[(2 * j + 1) * C + e] = a0 * (RT)dl[0] + a1 * (RT)dl[1];
                } else {
                    v[j * C + e] = a0 + a1;
                }
            }
        }
        wave_sum_8xC<C>(v, lane);
        if ((lane & 7) == 0)
#pragma unroll
            for (int e = 0; e < C; ++e) red[wave][lane >> 3][E0 + e] = v[e];
    };
    // the last one or two modes of the value form fit one reduction: its 24
    // slots hold (mode, entry) pairs q = 9 (m - m0) + kl instead of 8 rows of
    // three entries, so the tail costs one reduction instead of three
    const int tail = (!DERIV && (M + 1) % kRowModes != 0 && (M + 1) % kRowModes <= 2) ? (M + 1) % kRowModes : 0;
    const int m_tail = M + 1 - tail;
    for (int m0 = 0; m0 < m_tail; m0 += kRowModes) {
        reduce_entries(m0, std::integral_constant<int, 0>{}, std::integral_constant<int, 3>{});
        reduce_entries(m0, std::integral_constant<int, 3>{}, std::integral_constant<int, 3>{});
        reduce_entries(m0, std::integral_constant<int, 6>{}, std::integral_constant<int, 3>{});
        __syncthreads();
        if (t < 72) {
            const int row = t / 9, kl = t % 9;
            double tot = 0.0;
#pragma unroll
            for (int wv = 0; wv < kWaves; ++wv) tot += (double)red[wv][row][kl];
            const int m = m0 + (DERIV ? row >> 1 : row);
            if (m <= M) {
                const double nrm = (m == 0) ? 1.0 / (2.0 * OCRT_PI) : 1.0 / OCRT_PI;
                double val = 2.0 * tot * (dphi * nrm);
                if (m > 0 && (kl == 2 || kl == 5)) val = -val;
                const long long at = !solar ? r_index(I, m, o, i, kl)
                                            : I.off_Rsun + ((long long)m * n + o) * 9 + kl;
                if (DERIV && (row & 1)) {
                    P.dR[at] = val;
                } else {
                    if (!F32 || !P.Rf) P.R[at] = val;
                    if (P.Rf && !solar) P.Rf[at] = (float)val;
                }
            }
        }
        __syncthreads();
    }
    if (tail > 0) {
        RT v[24];
        const int t1 = t + kCoxMunkThreads;
#pragma unroll
        for (int q = 0; q < 24; ++q) {
            const int mq = m_tail + (q < 9 * tail ? q / 9 : 0), kl = q % 9;
            const RT *wc = cmw + (long long)mq * 2 * kNPhi, *ws = wc + kNPhi;
            const bool use_cos = (kl == 0 || kl == 1 || kl == 3 || kl == 4 || kl == 8);
            v[q] = (q < 9 * tail) ? r[0][kl] * (use_cos ? wc[t] : ws[t]) + r[1][kl] * (use_cos ? wc[t1] : ws[t1])
                                  : (RT)0;
        }
        wave_sum_8xC<3>(v, lane);
        if ((lane & 7) == 0)
#pragma unroll
            for (int e = 0; e < 3; ++e) red[wave][lane >> 3][e] = v[e];
        __syncthreads();
        if (t < 9 * tail) {
            const int row = t / 3, e = t % 3, m = m_tail + t / 9, kl = t % 9;
            double tot = 0.0;
#pragma unroll
            for (int wv = 0; wv < kWaves; ++wv) tot += (double)red[wv][row][e];
            const double nrm = (m == 0) ? 1.0 / (2.0 * OCRT_PI) : 1.0 / OCRT_PI;
            double val = 2.0 * tot * (dphi * nrm);
            if (m > 0 && (kl == 2 || kl == 5)) val = -val;
            const long long at = !solar ? r_index(I, m, o, i, kl) : I.off_Rsun + ((long long)m * n + o) * 9 + kl;
            if (!F32) P.R[at] = val;
            if (P.Rf && !solar) P.Rf[at] = (float)val;
        }
    }
}

// SET 0: every pair of an owner (n^2 diffuse, then n solar); 1: its solar
// pairs only; 2: its diffuse pairs only.  first[] counts the pairs of the set.
template <bool DERIV, typename RT, int SET>
// three workgroups per CU: the register budget that keeps the Mueller
// evaluations free of spills (measured faster than two or four)
__global__ __launch_bounds__(kCoxMunkThreads, 3) void coxmunk_fourier(const DevInst *__restrict__ inst, Pools P,
                                                                   const int *__restrict__ owners,
                                                                   const int *__restrict__ first, int n_owners,
                                                                   int pair_base)
{
    // pair b of the chunk belongs to owner k with first[k] <= b < first[k + 1]
    const int b = pair_base + blockIdx.x;
    if (b >= first[n_owners]) return;
    int lo = 0, hi = n_owners;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (first[mid] <= b) lo = mid; else hi = mid;
    }
    const DevInst &I = inst[owners[lo]];
    coxmunk_pair<DERIV, RT>(I, P, b - first[lo] + (SET == 1 ? I.n * I.n : 0));
}

template __global__ void coxmunk_fourier<false, double, 0>(const DevInst *, Pools, const int *, const int *, int, int);
template __global__ void coxmunk_fourier<true, double, 0>(const DevInst *, Pools, const int *, const int *, int, int);
template __global__ void coxmunk_fourier<false, double, 1>(const DevInst *, Pools, const int *, const int *, int, int);
template __global__ void coxmunk_fourier<false, float, 2>(const DevInst *, Pools, const int *, const int *, int, int);
template __global__ void coxmunk_fourier<true, double, 1>(const DevInst *, Pools, const int *, const int *, int, int);
template __global__ void coxmunk_fourier<true, float, 2>(const DevInst *, Pools, const int *, const int *, int, int);

// ---------------------------------------------------------------------------
// Signed direction cosine of direction index d = j + n, j in [-n, n]; j = 0 is
// the solar slot (-mu_sun).
__device__ inline double dir_mu(const double *ring_mu, int n, double mus, int d)
{
    const int j = d - n;
    if (j == 0) return -mus;
    return j > 0 ? ring_mu[j - 1] : -ring_mu[-j - 1];
}

// Value kernel of the aerosol I -> I phase table (rt_solver.c:204-240,
// selected at :922-925), for every Fourier mode of an instance:
//   P^m[j][k] = (1/720) sum_q P11(mu_j mu_k + s_j s_k cos phi_q) cos(m phi_q)
// over the 720 azimuth midpoints, P11 from the instance's angle table.  One
// workgroup per (owner instance, row j = 0..n, column direction k): the
// integrand is even in phi (phi_q and 2 pi - phi_q pair up), so the 360
// azimuths of the half circle are evaluated once into LDS and thread m folds
// them onto cos(m phi_q) for mode m, writing the entry of that mode's item.
// Runs after phase_tables and replaces its P table.
__global__ __launch_bounds__(256) void value_phase(const DevInst *__restrict__ inst, const DevItem *__restrict__ items,
                                                   Pools P, const int *__restrict__ owners,
                                                   const int *__restrict__ first, int n_owners, int pair_base,
                                                   const double *__restrict__ vw, int w_rows)
{
    __shared__ double val[kVkHalf];
    const int b = pair_base + blockIdx.x;
    if (b >= first[n_owners]) return;
    int lo = 0, hi = n_owners;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (first[mid] <= b) lo = mid; else hi = mid;
    }
    const DevInst &I = inst[owners[lo]];
    const int n = I.n, dirs = 2 * n + 1, pair = b - first[lo];
    const int j = pair / dirs, kd = pair % dirs;
    const double *ring = P.ring + I.off_ring;
    const double mj = dir_mu(ring, n, I.mus, j + n), mk = dir_mu(ring, n, I.mus, kd);
    const double sj = sqrt(fmax(0.0, 1.0 - mj * mj)), sk = sqrt(fmax(0.0, 1.0 - mk * mk));
    const double *th = P.greek + I.off_ang, *p11 = th + I.n_ang;
    for (int q = threadIdx.x; q < kVkHalf; q += blockDim.x) {
        const double phi = 2.0 * OCRT_PI * ((double)q + 0.5) / (double)kVkPhi;
        val[q] = p11_table_at(I.n_ang, th, p11, mj * mk + sj * sk * cos(phi));
    }
    __syncthreads();
    const long long cnt = (long long)(n + 1) * dirs;
    for (int m = threadIdx.x; m <= I.mmax; m += blockDim.x) {
        double acc = 0.0;
        for (int q = 0; q < kVkHalf; ++q) acc = fma(val[q], vw[(long long)q * w_rows + m], acc);
        const DevItem &it = items[I.item0 + m];
        P.tab[it.off_tab + 0 * cnt + (long long)j * dirs + kd] = 2.0 * acc / (double)kVkPhi;
    }
}


// Basis of one item: one thread per direction runs the three recurrences in l.
__global__ void angular_basis(const DevInst *__restrict__ inst, const DevItem *__restrict__ items, Pools P)
{
    const DevItem &it = items[blockIdx.x];
    if (!it.op_owner) return;   // shares the basis of an identical item
    const DevInst &I = inst[it.inst];
    const int n = it.n, dirs = 2 * n + 1, m = it.m, lmax = it.lmax;
    double *plm = P.basis + it.off_basis;
    double *rrl = plm + (long long)(lmax + 1) * dirs;
    double *rtl = rrl + (long long)(lmax + 1) * dirs;
    const double *mu = P.ring + I.off_ring;
    for (int d = threadIdx.x; d < dirs; d += blockDim.x) {
        const double x = dir_mu(mu, n, I.mus, d);
        for (int l = 0; l < m && l <= lmax; ++l) {
            plm[(long long)l * dirs + d] = 0.0;
            rrl[(long long)l * dirs + d] = 0.0;
            rtl[(long long)l * dirs + d] = 0.0;
        }
        double p0, p1, r0 = 0.0, r1, t0 = 0.0, t1;   // values at l-1 and l
        int l0;
        if (m == 0) {
            plm[0 * dirs + d] = 1.0; rrl[0 * dirs + d] = 0.0; rtl[0 * dirs + d] = 0.0;
            plm[1 * dirs + d] = x; rrl[1 * dirs + d] = 0.0; rtl[1 * dirs + d] = 0.0;
            p0 = x; p1 = 0.5 * (3.0 * x * x - 1.0);
            r1 = 3.0 * (1.0 - x * x) / (2.0 * sqrt(6.0)); t1 = 0.0;
            l0 = 2;
        } else if (m == 1) {
            const double s = sqrt(0.5 * (1.0 - x * x));
            p0 = s; p1 = x * s * sqrt(3.0);
            const double rt = sqrt(1.0 - x * x);
            r1 = -x * rt * 0.5; t1 = -rt * 0.5;
            plm[1 * dirs + d] = s; rrl[1 * dirs + d] = 0.0; rtl[1 * dirs + d] = 0.0;
            l0 = 2;
        } else {
            const double nm = diag_norm(m);
            const double sn = nm * sqrt((double)m / (double)(m + 1)) * sqrt((double)(m - 1) / (double)(m + 2));
            const double omx = 1.0 - x * x;
            p0 = 0.0; p1 = nm * pow(omx, 0.5 * (double)m);
            const double env = pow(omx, 0.5 * (double)m - 1.0);
            r1 = sn * (1.0 + x * x) * env; t1 = sn * 2.0 * x * env;
            l0 = m;
        }
        plm[(long long)l0 * dirs + d] = p1;
        rrl[(long long)l0 * dirs + d] = r1;
        rtl[(long long)l0 * dirs + d] = t1;
        for (int l = l0; l < lmax; ++l) {
            const double fw = (2.0 * l + 1.0) / sqrt((double)(l + m + 1) * (double)(l - m + 1));
            const double bw = sqrt((double)(l + m) * (double)(l - m)) / (2.0 * l + 1.0);
            const double sc = (double)(l + 1) * (double)(2 * l + 1) /
                              sqrt((double)(l + 3) * (double)(l - 1) * (double)(l + m + 1) * (double)(l - m + 1));
            const double dc = sqrt((double)(l + 2) * (double)(l - 2) * (double)(l + m) * (double)(l - m)) /
                              ((double)l * (double)(2 * l + 1));
            const double spc = 2.0 * (double)m / ((double)l * (double)(l + 1));
            const double p2 = fw * (x * p1 - bw * p0);
            const double r2 = sc * (x * r1 - spc * t1 - dc * r0);
            const double t2 = sc * (x * t1 - spc * r1 - dc * t0);
            plm[(long long)(l + 1) * dirs + d] = p2;
            rrl[(long long)(l + 1) * dirs + d] = r2;
            rtl[(long long)(l + 1) * dirs + d] = t2;
            p0 = p1; p1 = p2; r0 = r1; r1 = r2; t0 = t1; t1 = t2;
        }
    }
}

// Six phase tables [(n+1)][2n+1] of one item (rt_angular_basis.c:278-358):
//   P[j][k]   = sum_l plm_j plm_k beta_l       GR[j][k] = sum_l plm_j rrl_k gamma_l
//   GT[j][k]  = sum_l plm_j rtl_k gamma_l      ATT      = sum_l rtl_j rtl_k alpha_l + rrl_j rrl_k zeta_l
//   ARR       = sum_l rtl_j rtl_k zeta_l + rrl_j rrl_k alpha_l
//   ART       = sum_l rtl_j rrl_k alpha_l + rrl_j rtl_k zeta_l
// for rows j = 0..n (directions j + n) and every column direction k.  One
// wavefront per (item, group of kTabRows rows), the groups of an item on one
// XCD.  The rows' basis values and the Greek coefficients of each order l are
// wavefront-uniform (scalar loads); the lanes take the column directions in
// passes of 64 and each accumulates its column for all rows of the group,
// with the next order's column values loaded before the current order's
// multiply-adds.
__global__ __launch_bounds__(64) void phase_tables(const DevInst *__restrict__ inst,
                                                   const DevItem *__restrict__ items, Pools P, int n_items,
                                                   int groups)
{
    int item, g;
    xcd_item_block(groups, item, g);
    if (item >= n_items) return;
    const DevItem &it = items[item];
    if (!it.op_owner) return;   // shares the tables of an identical item
    const DevInst &I = inst[it.inst];
    const int n = it.n, dirs = 2 * n + 1, m = it.m, lmax = it.lmax;
    const int j0 = g * kTabRows;
    if (j0 > n) return;
    const int nr = min(kTabRows, n + 1 - j0);
    const long long cnt = (long long)(n + 1) * dirs;
    double *__restrict__ tab = P.tab + it.off_tab;
    const int lane = threadIdx.x;
    if (!I.aer) {
        for (int e = lane; e < nr * dirs; e += 64)
            for (int t = 0; t < 6; ++t) tab[t * cnt + (long long)j0 * dirs + e] = 0.0;
        return;
    }
    const double *__restrict__ plm = P.basis + it.off_basis;
    const double *__restrict__ rrl = plm + (long long)(lmax + 1) * dirs;
    const double *__restrict__ rtl = rrl + (long long)(lmax + 1) * dirs;
    const double *__restrict__ gk = P.greek + I.off_greek;
    const int q = I.La + 1;
    const int lend = lmax < I.La ? lmax : I.La;
    // the rows' values through the constant address space: wavefront-uniform
    // addresses there become scalar loads
    typedef const __attribute__((address_space(4))) double cdouble;
    cdouble *plm_s = (cdouble *)plm, *rrl_s = (cdouble *)rrl, *rtl_s = (cdouble *)rtl, *gk_s = (cdouble *)gk;
    for (int k0 = 0; k0 < dirs; k0 += 64) {
        const int k = k0 + lane;
        const int kc = k < dirs ? k : dirs - 1;   // clamped column: loads stay in bounds
        double acc[kTabRows][6];
#pragma unroll
        for (int r = 0; r < kTabRows; ++r)
#pragma unroll
            for (int t = 0; t < 6; ++t) acc[r][t] = 0.0;
        // one order's terms for every row of the group
        auto term = [&](int l, double pk, double rk, double tk) {
            const long long bj = (long long)l * dirs + n + j0;
            const double cb = gk_s[l], cg = gk_s[q + l], ca = gk_s[2 * q + l], cz = gk_s[3 * q + l];
            const double pkb = pk * cb, rkg = rk * cg, tkg = tk * cg;
            const double tka = tk * ca, tkz = tk * cz, rka = rk * ca, rkz = rk * cz;
#pragma unroll
            for (int r = 0; r < kTabRows; ++r) {
                // rows past the group's end read on into the next order's row (or
                // the pool's padding); their sums are never stored
                const double pj = plm_s[bj + r], rj = rrl_s[bj + r], tj = rtl_s[bj + r];
                acc[r][0] += pj * pkb;
                acc[r][1] += pj * rkg;
                acc[r][2] += pj * tkg;
                acc[r][3] += tj * tkz + rj * rka;
                acc[r][4] += tj * rka + rj * tkz;
                acc[r][5] += tj * tka + rj * rkz;
            }
        };
        // orders in pairs: both orders' column values are loaded before the
        // first order's multiply-adds, so the second load overlaps them
        int l = m;
        for (; l + 1 <= lend; l += 2) {
            const long long b0 = (long long)l * dirs + kc, b1 = b0 + dirs;
            const double p0 = plm[b0], r0 = rrl[b0], t0 = rtl[b0];
            const double p1 = plm[b1], r1 = rrl[b1], t1 = rtl[b1];
            term(l, p0, r0, t0);
            term(l + 1, p1, r1, t1);
        }
        if (l <= lend) {
            const long long b0 = (long long)l * dirs + kc;
            term(l, plm[b0], rrl[b0], rtl[b0]);
        }
        if (k < dirs) {
#pragma unroll
            for (int r = 0; r < kTabRows; ++r) {
                if (r >= nr) break;
                const long long idx = (long long)(j0 + r) * dirs + k;
#pragma unroll
                for (int t = 0; t < 6; ++t) tab[t * cnt + idx] = acc[r][t];
            }
        }
    }
}

// Coefficient selector of the operator: for output type (row block) and input
// type (column block) in the order [I_up I_dn Q_up Q_dn U_up U_dn], which of
// the 18 pair coefficients enters and with which sign
// (rt_sos_operator.c:383-406).
__constant__ signed char c_sel_q[6][6] = {
    {0, 1, 8, 9, 4, 5}, {1, 0, 9, 8, 5, 4}, {6, 7, 10, 11, 12, 13},
    {7, 6, 11, 10, 13, 12}, {2, 3, 14, 15, 16, 17}, {3, 2, 15, 14, 17, 16}};
__constant__ signed char c_sel_s[6][6] = {
    {1, 1, 1, 1, -1, -1}, {1, 1, 1, 1, 1, 1}, {1, 1, 1, 1, -1, 1},
    {1, 1, 1, 1, -1, 1}, {-1, 1, -1, -1, 1, 1}, {-1, 1, 1, 1, 1, 1}};

// Particle pair coefficient q for outgoing node o and incoming node i (1-based).
__device__ inline double pair_particle(const double *tab, long long cnt, int dirs, int n, int q, int o, int i)
{
    // table t, row, signed column
    int t, row, col;
    switch (q) {
    case 0: t = 0; row = i; col = o; break;
    case 1: t = 0; row = i; col = -o; break;
    case 2: t = 2; row = i; col = o; break;
    case 3: t = 2; row = i; col = -o; break;
    case 4: t = 2; row = o; col = i; break;
    case 5: t = 2; row = o; col = -i; break;
    case 6: t = 1; row = i; col = o; break;
    case 7: t = 1; row = i; col = -o; break;
    case 8: t = 1; row = o; col = i; break;
    case 9: t = 1; row = o; col = -i; break;
    case 10: t = 3; row = i; col = o; break;
    case 11: t = 3; row = i; col = -o; break;
    case 12: t = 4; row = i; col = o; break;
    case 13: t = 4; row = i; col = -o; break;
    case 14: t = 4; row = o; col = i; break;
    case 15: t = 4; row = o; col = -i; break;
    case 16: t = 5; row = i; col = o; break;
    default: t = 5; row = i; col = -o; break;
    }
    return tab[t * cnt + (long long)row * dirs + col + n];
}

// Molecular pair coefficient q (rt_sos_operator.c:102-119).
__device__ inline double pair_molecular(const DevInst &I, int m, const double *plm2, const double *rrl2,
                                        const double *rtl2, int n, int q, int o, int i)
{
    const double b0 = (m == 0) ? I.beta0 : 0.0;
    const double po = plm2[n + o], pom = plm2[n - o], pi = plm2[n + i];
    const double ro = rrl2[n + o], rom = rrl2[n - o], ri = rrl2[n + i], rim = rrl2[n - i];
    const double to = rtl2[n + o], tom = rtl2[n - o], ti = rtl2[n + i], tim = rtl2[n - i];
    switch (q) {
    case 0: return b0 + I.beta2 * pi * po;
    case 1: return b0 + I.beta2 * pi * pom;
    case 2: return I.gamma2 * pi * to;
    case 3: return I.gamma2 * pi * tom;
    case 4: return I.gamma2 * po * ti;
    case 5: return I.gamma2 * po * tim;
    case 6: return I.gamma2 * pi * ro;
    case 7: return I.gamma2 * pi * rom;
    case 8: return I.gamma2 * po * ri;
    case 9: return I.gamma2 * po * rim;
    case 10: return I.alpha2 * ri * ro;
    case 11: return I.alpha2 * ri * rom;
    case 12: return I.alpha2 * ti * ro;
    case 13: return I.alpha2 * ti * rom;
    case 14: return I.alpha2 * to * ri;
    case 15: return I.alpha2 * to * rim;
    case 16: return I.alpha2 * ti * to;
    default: return I.alpha2 * ti * tom;
    }
}

// The rank-3 molecular operator of Fourier mode m <= 2 (the reference's
// outer-product channels, rt_sos_operator.c:102-119): over the state columns
// c = (type t, node o), t in [I_up I_dn Q_up Q_dn U_up U_dn] with component
// t / 2 and hemisphere h = t % 2, and the l = 2 basis p, r, t at +o, with the
// parities of the generalised spherical functions (eps_p = eps_r = (-1)^m,
// eps_t = -(-1)^m) folding the mirrored directions in:
//   out factors  e1 = [I]  e2 = [I] eps_p^h p  e3 = [Q] eps_r^h r - [U] eps_t^h t
//   in factors   g_b = w_o x (the same functions)
//   A_mol = E C G^T,  C = [[beta0 (m = 0), 0, 0], [0, beta2, gamma2], [0, gamma2, alpha2]]
// (the selector signs of the dense form, c_sel_s, are reproduced by these
// parities; the dense operator of pack_operator equals E C G^T to rounding).
// U = E C are the output factors the sweep applies; G the input factors the
// operator's projection rows carry.
__device__ inline void molecular_factors(const DevInst &I, const DevItem &it, const Pools &P, int c, double E[3],
                                         double G[3])
{
    E[0] = E[1] = E[2] = G[0] = G[1] = G[2] = 0.0;
    const int n = it.n;
    if (c >= 6 * n) return;
    const int t = col_sel_type(c, n), comp = t / 2, h = t % 2, o = c - (c / n) * n + 1;
    const int dirs = 2 * n + 1;
    const double *plm = P.basis + it.off_basis;
    const double p2 = plm[2LL * dirs + n + o];
    const double r2 = plm[(long long)(it.lmax + 1) * dirs + 2LL * dirs + n + o];
    const double t2 = plm[2LL * (it.lmax + 1) * dirs + 2LL * dirs + n + o];
    const double ep = (it.m & 1) ? -1.0 : 1.0, et = -ep;
    const double w = P.ring[I.off_ring + n + (o - 1)];
    const double sp = h ? ep : 1.0, sr = h ? ep : 1.0, st = h ? et : 1.0;
    if (comp == 0) { E[0] = 1.0; E[1] = sp * p2; }
    else if (comp == 1) E[2] = sr * r2;
    else E[2] = -st * t2;
    if (w == 0.0) return;
    G[0] = w * E[0]; G[1] = w * E[1]; G[2] = w * E[2];
}

// Output factors U = E C of every rank-3 item that owns its operator, fp32.
__global__ void lowrank_factors(const DevInst *__restrict__ inst, const DevItem *__restrict__ items, Pools P)
{
    const DevItem &it = items[blockIdx.x];
    if (!it.lowrank || !it.op_owner) return;
    const DevInst &I = inst[it.inst];
    const double b0 = (it.m == 0) ? I.beta0 : 0.0;
    float *U = P.lr + it.off_lr;
    for (int c = threadIdx.x; c < it.P; c += blockDim.x) {
        double E[3], G[3];
        molecular_factors(I, it, P, c, E, G);
        U[c] = (float)(b0 * E[0]);
        U[it.P + c] = (float)(I.beta2 * E[1] + I.gamma2 * E[2]);
        U[2 * it.P + c] = (float)(I.gamma2 * E[1] + I.alpha2 * E[2]);
    }
}

// Dense operator A[Mp][K] of one item; element (r, c) with row r = outgoing
// state column and c = (part, incoming state column).  Part 0/1 of the
// column range is the particle (scaled by xdel at apply time) or molecular
// (scaled by ydel) block.
// TRANS packs the transposed operator (A_part^T of every part) in the same
// layout, for the adjoint's M^T products.
//
// Each wavefront takes whole rows (wave-uniform row index, so the row's
// selectors sit in scalar registers); its lanes take CW adjacent columns each
// (CW = 2 for bf16: one 32-bit store per half), with the column selectors
// decoded once per lane and pass.  The bf16 pair A = hi + lo is formed in
// registers: lo = bf16(v - hi).
__device__ inline double operator_entry(const DevInst &I, const DevItem &it, const Pools &P, int kind, int ro,
                                        int ci)
{
    const int n = it.n, dirs = 2 * n + 1;
    const int to = col_sel_type(ro, n), o = ro - (ro / n) * n + 1;
    const int ti = col_sel_type(ci, n), i = ci - (ci / n) * n + 1;
    const double w = P.ring[I.off_ring + n + (i - 1)];
    if (w == 0.0) return 0.0;
    const int q = c_sel_q[to][ti];
    const int s = c_sel_s[to][ti];
    double e;
    if (kind == 1) {
        e = pair_particle(P.tab + it.off_tab, (long long)(n + 1) * dirs, dirs, n, q, o, i);
    } else {
        const double *plm = P.basis + it.off_basis;
        const double *plm2 = plm + 2LL * dirs;
        const double *rrl2 = plm + (long long)(it.lmax + 1) * dirs + 2LL * dirs;
        const double *rtl2 = plm + 2LL * (it.lmax + 1) * dirs + 2LL * dirs;
        e = pair_molecular(I, it.m, plm2, rrl2, rtl2, n, q, o, i);
    }
    return w * (double)s * e;
}

template <typename AT, bool TRANS>
__global__ __launch_bounds__(256) void pack_operator(const DevInst *__restrict__ inst,
                                                     const DevItem *__restrict__ items, Pools P, int n_items,
                                                     int blocks_per_item)
{
    int item, bx;
    xcd_item_block(blocks_per_item, item, bx);   // an item's tables stay in one XCD's L2
    if (item >= n_items) return;
    const DevItem &it = items[item];
    if (!it.op_owner) return;   // shares the operator of an identical item
    const DevInst &I = inst[it.inst];
    constexpr bool BF = sizeof(AT) == 2;
    constexpr int CW = BF ? 2 : 1;
    const int six_n = 6 * it.n, Pp = it.P, Mp = it.Mp, K = it.K;
    const int nparts = K / Pp;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int waves = blocks_per_item * (blockDim.x >> 6);
    for (int r = __builtin_amdgcn_readfirstlane(bx * (int)(blockDim.x >> 6) + wave); r < Mp; r += waves) {
        for (int part = 0; part < nparts; ++part) {
            const int kind = (it.parts >> (8 * part)) & 0xff;   // 1 particle, 2 molecular
            for (int cc = CW * lane; cc < Pp; cc += CW * 64) {
                double v[CW];
#pragma unroll
                for (int e = 0; e < CW; ++e) {
                    const int c = cc + e;
                    if (it.lowrank && r >= it.pbase) {
                        // projection rows: the rank-3 molecular input factors G
                        double E[3], G[3];
                        molecular_factors(I, it, P, c, E, G);
                        v[e] = (r - it.pbase < 3) ? G[r - it.pbase] : 0.0;
                    } else {
                        v[e] = (r < six_n && c < six_n)
                                   ? operator_entry(I, it, P, kind, TRANS ? c : r, TRANS ? r : c)
                                   : 0.0;
                    }
                }
                if constexpr (BF) {
                    // fp32 rounding first, as every bf16 conversion of the solver
                    unsigned short hi[CW], lo[CW];
#pragma unroll
                    for (int e = 0; e < CW; ++e) {
                        const __hip_bfloat16 h = __float2bfloat16((float)v[e]);
                        const __hip_bfloat16 l = __float2bfloat16((float)(v[e] - (double)__bfloat162float(h)));
                        hi[e] = __builtin_bit_cast(unsigned short, h);
                        lo[e] = __builtin_bit_cast(unsigned short, l);
                    }
                    // [part][hi, lo][Mp][P]
                    unsigned *dst = reinterpret_cast<unsigned *>(reinterpret_cast<unsigned short *>(P.A) + it.off_A +
                                                                 ((long long)(part * 2) * Mp + r) * Pp + cc);
                    dst[0] = (unsigned)hi[0] | ((unsigned)hi[1] << 16);
                    dst[(long long)Mp * Pp / 2] = (unsigned)lo[0] | ((unsigned)lo[1] << 16);
                } else {
                    // [Mp][K], parts side by side along K
                    reinterpret_cast<double *>(P.A)[it.off_A + (long long)r * K + part * Pp + cc] = v[0];
                }
            }
        }
    }
}

template __global__ void pack_operator<double, false>(const DevInst *, const DevItem *, Pools, int, int);
template __global__ void pack_operator<__hip_bfloat16, false>(const DevInst *, const DevItem *, Pools, int, int);
template __global__ void pack_operator<double, true>(const DevInst *, const DevItem *, Pools, int, int);
template __global__ void pack_operator<__hip_bfloat16, true>(const DevInst *, const DevItem *, Pools, int, int);

}  // namespace ocrt
