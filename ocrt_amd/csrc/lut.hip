// LUT / full-grid epilogue of a solve (rt_solve_case_pol_lut,
// rt_solver.c:1031-1043 and 1280-1424): the per-mode TOA upward field at every
// ring node is sampled at each viewing zenith by Fritsch-Carlson PCHIP
// (rt_solver.c:409-495; linear sampling for the BOA downward intensity,
// rt_solver.c:386-408), reconstructed at each relative azimuth
// (rt_fourier.c:9-33), scaled by the IPSS factor of that direction and
// completed with the direct glint and the transmittance columns.
#include "ocrt_amd.h"
#include "sos_common.h"

namespace ocrt {

__device__ inline double pchip_int(double h0, double h1, double s0, double s1)
{
    if (s0 == 0.0 || s1 == 0.0 || (s0 > 0.0) != (s1 > 0.0)) return 0.0;
    const double w1 = 2.0 * h1 + h0, w2 = h1 + 2.0 * h0;
    return (w1 + w2) / (w1 / s0 + w2 / s1);
}

__device__ inline double pchip_edge(double h0, double h1, double s0, double s1)
{
    double d = ((2.0 * h0 + h1) * s0 - h0 * s1) / (h0 + h1);
    if (d == 0.0 || s0 == 0.0 || (d > 0.0) != (s0 > 0.0)) {
        if (s0 == 0.0 || (d > 0.0) != (s0 > 0.0)) d = 0.0;
    } else if (((s0 > 0.0) != (s1 > 0.0)) && fabs(d) > 3.0 * fabs(s0)) {
        d = 3.0 * s0;
    }
    return d;
}

// Nodes x[0..n-1] ascending; linear inside, linear extrapolation outside.
__device__ inline double sample_lin(const double *y, const double *x, int n, double mv)
{
    int jh = 1;
    while (jh < n && x[jh - 1] < mv) jh++;
    int jl = jh - 1;
    if (jl < 1) { jl = 1; jh = 2; }
    if (jh > n) { jh = n; jl = n - 1; }
    const double t = (mv - x[jl - 1]) / (x[jh - 1] - x[jl - 1]);
    return (1.0 - t) * y[jl - 1] + t * y[jh - 1];
}

__device__ inline double sample_pchip(const double *y, const double *x, int n, double mv)
{
    int jh = 1;
    while (jh < n && x[jh - 1] < mv) jh++;
    const int jl = jh - 1;
    if (jl < 1 || jh > n || n < 4) return sample_lin(y, x, n, mv);
    const double x1 = x[jl - 1], x2 = x[jh - 1], y1 = y[jl - 1], y2 = y[jh - 1];
    const double h = x2 - x1, s = (y2 - y1) / h;
    double d1, d2;
    if (jl == 1) {
        const double x3 = x[jh], s2 = (y[jh] - y2) / (x3 - x2);
        d1 = pchip_edge(h, x3 - x2, s, s2);
        d2 = pchip_int(h, x3 - x2, s, s2);
    } else if (jh == n) {
        const double x0 = x[jl - 2], s0 = (y1 - y[jl - 2]) / (x1 - x0);
        d1 = pchip_int(x1 - x0, h, s0, s);
        d2 = pchip_edge(h, x1 - x0, s, s0);
    } else {
        const double x0 = x[jl - 2], x3 = x[jh];
        const double s0 = (y1 - y[jl - 2]) / (x1 - x0), s2 = (y[jh] - y2) / (x3 - x2);
        d1 = pchip_int(x1 - x0, h, s0, s);
        d2 = pchip_int(h, x3 - x2, s, s2);
    }
    const double t = (mv - x1) / h, t2 = t * t, t3 = t2 * t;
    return (2.0 * t3 - 3.0 * t2 + 1.0) * y1 + (t3 - 2.0 * t2 + t) * h * d1 + (-2.0 * t3 + 3.0 * t2) * y2 +
           (t3 - t2) * h * d2;
}

// One workgroup per (viewing zenith, instance of the chunk).  Dynamic LDS:
// 4 x (m_max + 1) sampled modes.
//   grid   [instance][n_vza][n_raa][6]: rho_I, rho_Q, rho_U, T_diff_dn_dir,
//          T_sg_up_dir, T_total_up_dir
//   scal   [instance][2]: T_dir_dn, T_diff_dn_hemi
//   raw    [instance][n_vza][n_raa][3]: TOA (I, Q, U) before rho, IPSS and glint
//   vpm    [instance][n_vza][m_stride][3]: sampled Fourier modes at each zenith
// kappa_lut[job_of[i] + iv * n_raa + ir] is the IPSS factor of a direction
// (job_of[i] < 0: IPSS off).
__global__ __launch_bounds__(128) void lut_eval(const DevInst *__restrict__ inst, const DevItem *__restrict__ items,
                                                Pools P, const InstOut *__restrict__ iout,
                                                const int *__restrict__ job_of, const double *__restrict__ kappa_lut,
                                                int n_vza, const double *__restrict__ vza, int n_raa,
                                                const double *__restrict__ raa, double *__restrict__ grid,
                                                double *__restrict__ scal, double *__restrict__ raw,
                                                double *__restrict__ vpm, int m_stride)
{
    extern __shared__ double vals[];
    const int iv = blockIdx.x, i = blockIdx.y;
    const DevInst &I = inst[i];
    const int n = I.n, mm = I.mmax, M = mm + 1;
    const double *mu = P.ring + I.off_ring, *w = mu + n;
    const double mv = cos(vza[iv] * OCRT_PI / 180.0);
    for (int q = threadIdx.x; q < 4 * M; q += blockDim.x) {
        const int m = q >> 2, comp = q & 3;
        const double *y = P.nodes + items[I.item0 + m].off_nodes + (long long)comp * n;
        vals[comp * M + m] = comp < 3 ? sample_pchip(y, mu, n, mv) : sample_lin(y, mu, n, mv);
    }
    __syncthreads();
    // per-view Fourier samples [instance][vza][m_stride][3] (zero above m_max)
    double *vq = vpm + ((long long)I.gidx * n_vza + iv) * m_stride * 3;
    for (int q = threadIdx.x; q < 3 * m_stride; q += blockDim.x) {
        const int m = q / 3, comp = q % 3;
        vq[q] = m < M ? vals[comp * M + m] : 0.0;
    }
    const LevelView lv(P, I);
    const double mus = I.mus, tau = lv.h[I.nt];
    const bool ip = I.ipss && job_of[i] >= 0;
    const double tdir = iout[i].tau_dir_ipss;
    const double cg = (ip && tdir >= 0.0) ? exp(-(tdir - tau / mus)) : 1.0;
    double *g = grid + ((long long)I.gidx * n_vza + iv) * n_raa * 6;
    for (int ir = threadIdx.x; ir < n_raa; ir += blockDim.x) {
        const double phi = raa[ir] * OCRT_PI / 180.0 + OCRT_PI;
        double rI = vals[0], rQ = vals[M], rU = 0.0, bd = vals[3 * M];
        for (int m = 1; m <= mm; ++m) {
            const double cm = cos(m * phi), sm = sin(m * phi);
            rI += 2.0 * vals[m] * cm;
            rQ += 2.0 * vals[M + m] * cm;
            rU += 2.0 * vals[2 * M + m] * sm;
            bd += 2.0 * vals[3 * M + m] * cm;
        }
        double *w3 = raw + (((long long)I.gidx * n_vza + iv) * n_raa + ir) * 3;
        w3[0] = rI; w3[1] = rQ; w3[2] = rU;
        rI /= mus; rQ /= mus; rU /= mus;
        if (ip) {
            const double kg = kappa_lut[job_of[i] + (long long)iv * n_raa + ir];
            rI *= kg; rQ *= kg; rU *= kg;
        }
        double sgI = 0.0;
        if (I.surface_kind == OCRT_SURFACE_BLACK_FRESNEL_OCEAN && I.wind > 0.0) {
            double rg[3];
            sunglint_rho(vza[iv] > 0.0 ? mv : 1.0, (raa[ir] - 180.0) * OCRT_PI / 180.0, mus, I.wind, I.sigma_type,
                         I.n_water, tau, I.q_conv, rg);
            rg[0] *= cg; rg[1] *= cg; rg[2] *= cg;
            if (!I.decouple) { rI += rg[0]; rQ += rg[1]; rU += -rg[2]; }
            sgI = rg[0];
        }
        double *o = g + (long long)ir * 6;
        o[0] = rI; o[1] = rQ; o[2] = rU;
        o[3] = (2.0 * OCRT_PI * mv * bd) / (OCRT_PI * mus);
        o[4] = sgI * mus;
        o[5] = rI * mus;
    }
    if (iv == 0 && threadIdx.x == 0) {
        const double *bdn = P.nodes + items[I.item0].off_nodes + 3LL * n;
        double F = 0.0;
        for (int o = 0; o < n; ++o) F += mu[o] * w[o] * bdn[o];
        F *= 2.0 * OCRT_PI;
        scal[2LL * I.gidx] = (ip && tdir >= 0.0) ? exp(-tdir) : exp(-tau / mus);
        scal[2LL * I.gidx + 1] = F / (OCRT_PI * mus);
    }
}

}  // namespace ocrt
