// IPSS kappa = I1_ss / I1_pp (Zhai & Hu 2022 Eq. 7) in fp64, one workgroup
// per (instance, viewing direction) job (rt_ipss.c:24-644).
//
// The reference traces every ray by building a sorted list of shell
// crossings and assigning each interval to the layer of its midpoint.  On the
// GPU the same path lengths are formed without per-thread lists:
//   - the surface-anchored line of sight descends monotonically from the TOA
//     to the pixel, so it crosses shell k once, at t_-(k), and layer j is the
//     interval [t_-(j+1), t_-(j)];
//   - the length of a solar ray inside layer j is f(R_{j+1}) - f(R_j) with
//     f(R) = |{t >= 0 : |r0 + t u| <= R}|, a closed form per shell, and the
//     ray is shadowed when it meets the ground.
#include "common.h"
#include "device_math.h"

namespace ocrt {

constexpr int kIpssMaxLayers = 1536;
constexpr double kEarthRadius = 6371.0;

__device__ inline double shell_len_below(double b, double c0, double R)
{
    const double disc = b * b - c0 + R * R;
    if (disc < 0.0) return 0.0;
    const double s = sqrt(disc);
    const double t1 = -b - s, t2 = -b + s;
    return fmax(0.0, t2 - fmax(0.0, t1));
}

// Solar slant optical depth from r0 to the TOA; INFINITY when shadowed.
__device__ inline double solar_depth(const double *h_edge, const double *beta, int nl, double mus,
                                     double x, double y, double zc)
{
    const double sn = sqrt(fmax(0.0, 1.0 - mus * mus));
    double ux = -sn, uy = 0.0, uz = mus;
    const double un = sqrt(ux * ux + uy * uy + uz * uz);
    ux /= un; uy /= un; uz /= un;
    const double b = x * ux + y * uy + zc * uz;
    const double c0 = x * x + y * y + zc * zc;
    const double Re = kEarthRadius;
    const double dg = b * b - (c0 - Re * Re);
    if (dg >= 0.0) {
        const double q = sqrt(dg);
        if (-b - q > 1e-9 || -b + q > 1e-9) return INFINITY;
    }
    double tau = 0.0;
    double f_lo = shell_len_below(b, c0, Re + h_edge[0]);
    for (int j = 0; j < nl; ++j) {
        const double f_hi = shell_len_below(b, c0, Re + h_edge[j + 1]);
        tau += beta[j] * (f_hi - f_lo);
        f_lo = f_hi;
    }
    return tau;
}

__global__ __launch_bounds__(256) void ipss_kappa(const DevInst *__restrict__ inst, const IpssJob *__restrict__ jobs,
                                                  Pools P, double *__restrict__ kappa_out,
                                                  double *__restrict__ taudir_out)
{
    __shared__ double h_edge[kIpssMaxLayers + 1];
    __shared__ double beta[kIpssMaxLayers];
    __shared__ double src[kIpssMaxLayers];
    __shared__ double tabove[kIpssMaxLayers + 1];
    __shared__ double s_top[kIpssMaxLayers];
    __shared__ double red[8];
    __shared__ double geo[8];
    const IpssJob jb = jobs[blockIdx.x];
    const DevInst &I = inst[jb.inst];
    const int nl = I.nt;
    if (nl > kIpssMaxLayers) {
        if (threadIdx.x == 0) { kappa_out[blockIdx.x] = NAN; if (taudir_out) taudir_out[blockIdx.x] = -1.0; }
        return;
    }
    const int L = nl + 1;
    const double *h = P.lev + I.off_lev, *xd = h + 2 * L, *yd = h + 3 * L, *z = h + 4 * L;
    const double mus = I.mus, Re = kEarthRadius;
    // phase weights of this geometry (rt_ipss.c:178-196)
    if (threadIdx.x == 0) {
        const double sn = sqrt(fmax(0.0, 1.0 - mus * mus));
        const double tv = jb.vza * OCRT_PI / 180.0, fp = (180.0 - jb.raa) * OCRT_PI / 180.0;
        const double ct = sn * sin(tv) * cos(fp) - mus * cos(tv);
        const double dp = 0.0279;
        const double b2 = (1.0 - dp) / (2.0 + dp);
        geo[0] = 1.0 + b2 * 0.5 * (3.0 * ct * ct - 1.0);
        double pa = 1.0;
        if (I.n_ang > 1) {
            // value-kernel table: interpolated, not clipped (rt_ipss.c:134-146)
            const double *th = P.greek + I.off_ang;
            pa = p11_table_at(I.n_ang, th, th + I.n_ang, ct);
        } else if (I.aer) {
            const double *g = P.greek + I.off_greek;
            double p0 = 1.0, p1 = ct, s = g[0];
            if (I.La >= 1) s += g[1] * p1;
            for (int l = 2; l <= I.La; ++l) {
                const double p2 = ((2.0 * l - 1.0) * ct * p1 - (l - 1.0) * p0) / l;
                s += g[l] * p2;
                p0 = p1; p1 = p2;
            }
            pa = s > 0.0 ? s : 0.0;
        }
        geo[1] = pa;
        int have = 0;
        for (int k = 0; k < nl; ++k) if (xd[k + 1] > 0.0) { have = 1; break; }
        geo[2] = (have && I.aer) ? 1.0 : 0.0;
    }
    __syncthreads();
    const double PR = geo[0], PA = geo[1];
    const bool weighted = geo[2] != 0.0;
    for (int j = threadIdx.x; j <= nl; j += blockDim.x) h_edge[j] = z[nl - j];
    __syncthreads();
    for (int j = threadIdx.x; j < nl; j += blockDim.x) {
        const int k = nl - 1 - j;
        const double dz = h_edge[j + 1] - h_edge[j];
        const double dtau = h[k + 1] - h[k];
        beta[j] = (dtau > 0.0) ? dtau / dz : 0.0;
        if (weighted) {
            const double wr = yd[k + 1] > 0.0 ? yd[k + 1] : 0.0;
            const double wa = xd[k + 1] > 0.0 ? xd[k + 1] : 0.0;
            src[j] = wr * PR + wa * PA;
        } else {
            // single-scattering albedo of the layer: gas absorption removed
            // (rt_ipss.c:64-70)
            double ta = I.gas >= 0 ? h[5 * L + k] : 0.0;
            ta = ta < 0.0 ? 0.0 : (ta > dtau ? dtau : ta);
            src[j] = dtau > 0.0 ? (dtau - ta) / dtau : 0.0;
        }
    }
    __syncthreads();
    const double H = h_edge[nl];
    // line of sight anchored at the pixel (0, 0, Re) (rt_ipss.c:493-517)
    const double tv = jb.vza * OCRT_PI / 180.0, fp = (180.0 - jb.raa) * OCRT_PI / 180.0;
    const double Vx = sin(tv) * cos(fp), Vy = sin(tv) * sin(fp), Vz = cos(tv);
    const double Rt = Re + H, bq = Re * Vz;
    const double disc = bq * bq + Rt * Rt - Re * Re;
    const double dmax = disc > 0.0 ? -bq + sqrt(disc) : 0.0;
    const double rAx = dmax * Vx, rAy = dmax * Vy, rAz = Re + dmax * Vz;
    const double bl = -(rAx * Vx + rAy * Vy + rAz * Vz);
    const double cl = rAx * rAx + rAy * rAy + rAz * rAz;
    for (int j = threadIdx.x; j < nl; j += blockDim.x) {
        // entry of layer j from above = crossing of shell j+1
        double t = 0.0;
        if (j + 1 < nl) {
            const double R = Re + h_edge[j + 1];
            const double d = bl * bl - (cl - R * R);
            t = d >= 0.0 ? -bl - sqrt(d) : 0.0;
        }
        s_top[j] = t;
    }
    __syncthreads();
    // cumulative tau along the line of sight and the vertical tau above each edge
    if (threadIdx.x == 0) {
        double tau = 0.0;
        for (int j = nl - 1; j >= 0; --j) {
            const double sb = (j == 0) ? dmax : s_top[j - 1];
            tabove[j + 1] = tau;                 // los tau at the layer entry (temporary)
            tau += (sb - s_top[j]) * beta[j];
        }
    }
    __syncthreads();
    double I1 = 0.0;
    const double gx = 0.5773502691896257645091488, gw = 1.0;
    if (dmax > 0.0) {
        for (int q = threadIdx.x; q < 2 * nl; q += blockDim.x) {
            const int j = q >> 1;
            if (!(beta[j] > 0.0)) continue;
            const double sa = s_top[j], sbb = (j == 0) ? dmax : s_top[j - 1];
            const double half = 0.5 * (sbb - sa), mid = 0.5 * (sa + sbb);
            const double s = mid + half * ((q & 1) ? gx : -gx);
            const double wg = half * gw;
            const double bx = rAx - s * Vx, by = rAy - s * Vy, bz = rAz - s * Vz;
            const double hB = sqrt(bx * bx + by * by + bz * bz) - Re;
            if (hB < 0.0 || hB > H) continue;
            const double tview = tabove[j + 1] + (s - sa) * beta[j];
            const double tsun = solar_depth(h_edge, beta, nl, mus, bx, by, bz);
            if (!isfinite(tsun)) continue;
            I1 += wg * beta[j] * src[j] * exp(-tview - tsun);
        }
    }
    for (int off = 32; off > 0; off >>= 1) I1 += __shfl_xor(I1, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = I1;
    __syncthreads();
    if (threadIdx.x == 0) {
        double ss = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) ss += red[w];
        geo[3] = ss;
        // vertical optical depth above each ascending edge
        tabove[nl] = 0.0;
        for (int j = nl - 1; j >= 0; --j) tabove[j] = tabove[j + 1] + beta[j] * (h_edge[j + 1] - h_edge[j]);
    }
    __syncthreads();
    const double mv = cos(jb.vza * OCRT_PI / 180.0);
    double pp = 0.0;
    if (mv > 0.0) {
        const double C = 1.0 / mv + 1.0 / mus;
        for (int j = threadIdx.x; j < nl; j += blockDim.x) {
            if (!(beta[j] > 0.0)) continue;
            pp += src[j] / (mv * C) * (exp(-C * tabove[j + 1]) - exp(-C * tabove[j]));
        }
    }
    for (int off = 32; off > 0; off >>= 1) pp += __shfl_xor(pp, off);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = pp;
    __syncthreads();
    if (threadIdx.x == 0) {
        double ptot = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) ptot += red[w];
        const double ss = geo[3];
        kappa_out[blockIdx.x] = (!(ptot > 1.0e-300) || !isfinite(ss)) ? 1.0 : ss / ptot;
        if (taudir_out) {
            const double t = solar_depth(h_edge, beta, nl, mus, 0.0, 0.0, Re);
            taudir_out[blockIdx.x] = isfinite(t) ? t : -1.0;
        }
    }
}

}  // namespace ocrt
