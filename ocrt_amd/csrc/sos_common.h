// Device helpers shared by the forward order loop (sos.hip) and its adjoint
// (adjoint.hip).
#pragma once
#include "common.h"
#include "device_math.h"
#include "wave.h"

namespace ocrt {

typedef double double4_ __attribute__((ext_vector_type(4)));
typedef float float4_ __attribute__((ext_vector_type(4)));
typedef short short4_ __attribute__((ext_vector_type(4)));

struct LevelView {
    const double *h, *ch, *xd, *yd, *z, *rdt;   // rdt[l] = 1 / (h[l + 1] - h[l])
    __device__ LevelView(const Pools &P, const DevInst &I)
    {
        const int L = I.nt + 1;
        h = P.lev + I.off_lev; ch = h + L; xd = ch + L; yd = xd + L; z = yd + L; rdt = h + 6 * L;
    }
    // from the level pool passed to a kernel as a __restrict__ argument: the
    // kernel's stores cannot alias it, so its wavefront-uniform reads compile
    // to scalar loads
    __device__ LevelView(const double *lev, const DevInst &I)
    {
        const int L = I.nt + 1;
        h = lev + I.off_lev; ch = h + L; xd = ch + L; yd = xd + L; z = yd + L; rdt = h + 6 * L;
    }
};

// Read access to the forward field of an adjoint plan, as fp64, at column
// `col` of one item: the fp64 field (fp64 plans), or the mixed form's bf16
// field plus, for the items that keep the pair, its low half (Pools::F and
// Pools::Flo).  Items with a bf16 field alone read the field twice and weigh
// the second read by zero, so every load is unconditional.
struct FieldF64 {
    const double *p;
    __device__ FieldF64(const Pools &P, const DevItem &it, int col)
        : p(reinterpret_cast<const double *>(P.F) + it.off_F + col) {}
    __device__ double operator[](long long ix) const { return p[ix]; }
};
struct FieldPair {
    const __hip_bfloat16 *hi, *lo;
    float wlo;
    __device__ FieldPair(const Pools &P, const DevItem &it, int col)
        : hi(reinterpret_cast<const __hip_bfloat16 *>(P.F) + it.off_F + col),
          lo(it.fsplit ? reinterpret_cast<const __hip_bfloat16 *>(P.Flo) + it.off_Flo + col : hi),
          wlo(it.fsplit ? 1.f : 0.f) {}
    __device__ double operator[](long long ix) const
    {
        return (double)__bfloat162float(hi[ix]) + (double)(wlo * __bfloat162float(lo[ix]));
    }
};

// Whether an item takes part in a launch: the forward order loop follows the
// convergence flag (order == 0); recomputation and the adjoint pass an order
// and take every item whose forward solve reached it.
__device__ inline bool item_runs(const ItemState &s, int order)
{
    return order > 0 ? (order <= s.orders) : (s.active != 0);
}

// Order-1 source of one state column (rt_primary_source.c:61-70, 164-171) and,
// over the flat Fresnel surface, the source of the specularly reflected direct
// beam scattered once (rt_surface_boundary.c:210-428):
//   primary(k) = sgn ch[k] (xdel[k] pc + ydel[k] mc)
//   flat(k)    = sb e^{h[k]/mu_s} (RI (pA xdel[k] + mA ydel[k]) + RQ (pB xdel[k] + mB ydel[k]))
// with sb = e^{-2 h[nt]/mu_s} / 2 and flat(k) = 0 at the column's boundary level.
struct ColumnSource {
    double pc, mc, sgn;
    bool flat;
    double pA, mA, pB, mB, RI, RQ, sb;
    int hemi, nt;

    __device__ ColumnSource(const DevInst &I, const DevItem &it, const Pools &P, const LevelView &lv, int col)
    {
        const int n = it.n, dirs = 2 * n + 1, m = it.m, lmax = it.lmax;
        nt = it.nt;
        int c, o;
        col_decode(col, n, c, hemi, o);
        const int d = hemi ? n - (o + 1) : n + (o + 1);
        const double *plm = P.basis + it.off_basis;
        const double *rrl = plm + (long long)(lmax + 1) * dirs;
        const double *rtl = rrl + (long long)(lmax + 1) * dirs;
        const double *plm2 = plm + 2LL * dirs, *rrl2 = rrl + 2LL * dirs, *rtl2 = rtl + 2LL * dirs;
        const long long cnt = (long long)(n + 1) * dirs;
        const double *tab = P.tab + it.off_tab;
        const int mol = (m <= 2);
        const double b0 = (m == 0) ? I.beta0 : 0.0;
        const double xs = plm2[n];
        if (c == 0) { pc = tab[0 * cnt + d]; mc = b0 + I.beta2 * plm2[d] * xs; sgn = 1.0; }
        else if (c == 1) { pc = tab[1 * cnt + d]; mc = mol ? I.gamma2 * rrl2[d] * xs : 0.0; sgn = 1.0; }
        else { pc = tab[2 * cnt + d]; mc = mol ? I.gamma2 * rtl2[d] * xs : 0.0; sgn = -1.0; }
        flat = I.surf == 1;
        pA = mA = pB = mB = RI = RQ = sb = 0.0;
        if (!flat) return;
        const double mus = I.mus, nw = I.n_water, n2 = nw * nw;
        const double tn = sqrt(n2 - 1.0 + mus * mus);
        const double rp = (n2 * mus - tn) / (n2 * mus + tn), rs = (mus - tn) / (mus + tn);
        RI = (rp * rp + rs * rs) / 2.0;
        RQ = (I.q_conv == 1) ? (rp * rp - rs * rs) / 2.0 : (rs * rs - rp * rp) / 2.0;
        sb = 0.5 * exp(-2.0 * lv.h[nt] / mus);
        const int dk = hemi ? n + (o + 1) : n - (o + 1);   // mirrored direction
        const double bP = xs, bR = rrl2[n];
        double mgr = 0.0, mart = 0.0;
        if (I.aer) {
            const double *g = P.greek + I.off_greek;
            const int q = I.La + 1;
            const int lend = lmax < I.La ? lmax : I.La;
            for (int l = m; l <= lend; ++l) {
                const long long b = (long long)l * dirs;
                mgr += plm[b + dk] * rrl[b + n] * g[q + l];
                mart += rtl[b + dk] * rrl[b + n] * g[2 * q + l] + rrl[b + dk] * rtl[b + n] * g[3 * q + l];
            }
        }
        if (c == 0) {
            pA = tab[0 * cnt + dk]; mA = mol ? b0 + I.beta2 * plm2[dk] * bP : 0.0;
            pB = mgr; mB = mol ? I.gamma2 * bR * plm2[dk] : 0.0;
        } else if (c == 1) {
            pA = tab[1 * cnt + dk]; mA = mol ? I.gamma2 * rrl2[dk] * bP : 0.0;
            pB = tab[3 * cnt + dk]; mB = mol ? I.alpha2 * bR * rrl2[dk] : 0.0;
        } else {
            pA = tab[2 * cnt + dk]; mA = mol ? I.gamma2 * rtl2[dk] * bP : 0.0;
            pB = mart; mB = mol ? I.alpha2 * rtl2[dk] * bR : 0.0;
        }
    }
    __device__ double primary(const LevelView &lv, int k) const
    {
        return sgn * lv.ch[k] * (lv.xd[k] * pc + lv.yd[k] * mc);
    }
    __device__ bool flat_zero(int k) const { return (hemi == 0 && k == nt) || (hemi == 1 && k == 0); }
    __device__ double flat_beam(const LevelView &lv, double mus, int k) const { return sb * exp(lv.h[k] / mus); }
    __device__ double flat_value(const LevelView &lv, double mus, int k) const
    {
        if (!flat || flat_zero(k)) return 0.0;
        const double x = lv.xd[k], y = lv.yd[k];
        return flat_beam(lv, mus, k) * (RI * (pA * x + mA * y) + RQ * (pB * x + mB * y));
    }
};

// Downward boundary condition at the top of an in-water column for the next
// order, from the upward field there of the order just finished
// (off_bc + 9n), threads over the reflected nodes:
//   flat  : the node's own water-side Fresnel matrix (rt_solver.c:2594-2606);
//   rough : sum_k C mu_k w_k K^m[j][k] up(k), C = 2 pi (m = 0) or pi
//           (rt_solver.c:2797-2822), K the water-side Cox-Munk kernel.
__device__ inline void top_bc_item(const DevInst &I, const DevItem &it, const Pools &P)
{
    const int n = it.n, m = it.m;
    const double *up = P.bc + it.off_bc + 9 * n;   // [3][n] upward field at the top
    double *tb = P.bc + it.off_bc + 6 * n;
    const double *mu = P.ring + I.off_ring, *w = mu + n;
    if (I.top_kind == 1) {
        for (int o = threadIdx.x; o < n; o += blockDim.x) {
            double M[9];
            fresnel_R_water_side(mu[o], I.n_water, I.q_conv, M);
            const double a = up[o], b = up[n + o], cc = up[2 * n + o];
            tb[o] = M[0] * a + M[1] * b + M[2] * cc;
            tb[n + o] = M[3] * a + M[4] * b + M[5] * cc;
            tb[2 * n + o] = M[6] * a + M[7] * b + M[8] * cc;
        }
        return;
    }
    const double C = (m == 0) ? 2.0 * OCRT_PI : OCRT_PI;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        double aI = 0.0, aQ = 0.0, aU = 0.0;
        const double *K = P.rtop + I.off_rtop + ((long long)m * n + j) * n * 9;
        for (int k = 0; k < n; ++k) {
            if (w[k] == 0.0) continue;
            const double a = up[k], b = up[n + k], cc = up[2 * n + k];
            const double wk = C * mu[k] * w[k];
            const double *Kk = K + 9LL * k;
            aI += wk * (Kk[0] * a + Kk[1] * b + Kk[2] * cc);
            aQ += wk * (Kk[3] * a + Kk[4] * b + Kk[5] * cc);
            aU += wk * (Kk[6] * a + Kk[7] * b + Kk[8] * cc);
        }
        tb[j] = aI;
        tb[n + j] = aQ;
        tb[2 * n + j] = aU;
    }
}

// Lower boundary condition of order `it` for the upward columns from the
// bottom downward field of the previous order (bc + 3n), threads over the
// outgoing nodes (rt_surface_boundary.c:137-208):
//   flat  : the Fresnel matrix of the node itself,
//   rough : sum_i W mu_i w_i R^m[o][i] down(i), W = 2 pi (m = 0) or pi.
__device__ inline void surface_bc_item(const DevInst &I, const DevItem &it, const Pools &P)
{
    const int n = it.n, m = it.m;
    const double *bot = P.bc + it.off_bc + 3 * n;   // [3][n] downward field at BOA
    const double *mu = P.ring + I.off_ring, *w = mu + n;
    double *bc = P.bc + it.off_bc;
    const double W = (I.surf == 1) ? 2.0 * OCRT_PI : ((m == 0) ? 2.0 * OCRT_PI : OCRT_PI);
    if (I.surf == 1) {
        for (int o = threadIdx.x; o < n; o += blockDim.x) {
            double si, sq, su;
            double Fr[9];
            fresnel_R(mu[o], I.n_water, I.q_conv, Fr);
            const double a = bot[o], b = bot[n + o], cc = bot[2 * n + o];
            if (w[o] == 0.0) {
                si = Fr[0] * a + Fr[1] * b + Fr[2] * cc;
                sq = Fr[3] * a + Fr[4] * b + Fr[5] * cc;
                su = Fr[6] * a + Fr[7] * b + Fr[8] * cc;
            } else {
                const double dd = 1.0 / (2.0 * OCRT_PI * mu[o] * w[o]);
                const double pw = W * mu[o] * w[o];
                double R[9];
                for (int e = 0; e < 9; ++e) R[e] = dd * Fr[e];
                si = pw * (R[0] * a + R[1] * b + R[2] * cc);
                sq = pw * (R[3] * a + R[4] * b + R[5] * cc);
                su = pw * (R[6] * a + R[7] * b + R[8] * cc);
            }
            bc[o] = si;
            bc[n + o] = sq;
            bc[2 * n + o] = su;
        }
        return;
    }
    // rough surface, threads over the outgoing nodes: entry kl of R^m[o][i]
    // sits at R[kl n] from the node's own base, consecutive over the lanes.
    // Mixed-form plans read the fp32 copy of the kernels (half the bytes of
    // this per-order product), accumulating in fp64.
    if (P.Rf) {
        for (int o = threadIdx.x; o < n; o += blockDim.x) {
            double si = 0.0, sq = 0.0, su = 0.0;
            for (int i = 0; i < n; ++i) {
                if (w[i] == 0.0) continue;
                const double a = bot[i], b = bot[n + i], cc = bot[2 * n + i];
                const double pw = W * mu[i] * w[i];
                const float *R = P.Rf + r_index(I, m, o, i, 0);
                si += pw * ((double)R[0] * a + (double)R[n] * b + (double)R[2 * n] * cc);
                sq += pw * ((double)R[3 * n] * a + (double)R[4 * n] * b + (double)R[5 * n] * cc);
                su += pw * ((double)R[6 * n] * a + (double)R[7 * n] * b + (double)R[8 * n] * cc);
            }
            bc[o] = si;
            bc[n + o] = sq;
            bc[2 * n + o] = su;
        }
        return;
    }
    for (int o = threadIdx.x; o < n; o += blockDim.x) {
        double si = 0.0, sq = 0.0, su = 0.0;
        for (int i = 0; i < n; ++i) {
            if (w[i] == 0.0) continue;
            const double a = bot[i], b = bot[n + i], cc = bot[2 * n + i];
            const double pw = W * mu[i] * w[i];
            const double *R = P.R + r_index(I, m, o, i, 0);
            si += pw * (R[0] * a + R[n] * b + R[2 * n] * cc);
            sq += pw * (R[3 * n] * a + R[4 * n] * b + R[5 * n] * cc);
            su += pw * (R[6 * n] * a + R[7 * n] * b + R[8 * n] * cc);
        }
        bc[o] = si;
        bc[n + o] = sq;
        bc[2 * n + o] = su;
    }
}

// Direct sunglint reflectance (surface.c:1813-1907); dlog_out (optional)
// receives d ln(rho) / d sigma^2 on the rough branch (0 elsewhere).
__device__ inline void sunglint_rho(double mu_v, double phi_v, double mu_0, double ws, int st, double nw,
                                    double tau, int qc, double *rho, double *dlog_out = nullptr)
{
    const double trans = exp(-tau / mu_0) * exp(-tau / mu_v);
    if (dlog_out) *dlog_out = 0.0;
    if (ws < 1e-6) {
        const double sf = 0.5 * OCRT_PI / 180.0;
        const double st0 = sqrt(fmax(0.0, 1.0 - mu_0 * mu_0));
        const double smu = sf * fmax(st0, 1e-3);
        double de = phi_v;
        while (de > OCRT_PI) de -= 2.0 * OCRT_PI;
        while (de < -OCRT_PI) de += 2.0 * OCRT_PI;
        const double dmu = mu_v - mu_0;
        const double ex = -0.5 * ((dmu * dmu) / (smu * smu) + (de * de) / (sf * sf));
        if (ex < -50.0) { rho[0] = rho[1] = rho[2] = 0.0; return; }
        double MF[9];
        fresnel_R(mu_0, nw, qc, MF);
        const double den = fmax(2.0 * mu_0 * st0 * mu_0 * mu_v, 1e-12);
        const double r = 4.0 * OCRT_PI * trans / den * exp(ex);
        rho[0] = MF[0] * r; rho[1] = MF[3] * r; rho[2] = MF[6] * r;
        return;
    }
    double R[9], tb2;
    const double s2v = slope_variance(ws, st);
    const double shadow = sancer_shadow(mu_v, mu_0, s2v);
    coxmunk_R_core(mu_v, mu_0, cos(phi_v), sin(phi_v), s2v, shadow, nw, qc, R, &tb2);
    rho[0] = OCRT_PI * R[0] * trans; rho[1] = OCRT_PI * R[3] * trans; rho[2] = OCRT_PI * R[6] * trans;
    if (dlog_out && tb2 >= 0.0)
        *dlog_out = tb2 / (s2v * s2v) - 1.0 / s2v + sancer_dlog(mu_v, mu_0, s2v, shadow);
}

// One linear-source step of the formal solution written as
//   r_k = T r_prev + al S_k + be S_prev,  al = (1 - q) / 2,  be = (q - T) / 2,
//   q = (1 - T) |mu| / dtau  (rt_layer_transport.c:89-99 rearranged),
// and the derivatives of T, al, be in dtau (adjoint).
__device__ inline void step_coefs(double dt, double amu, double &T, double &al, double &be, double &dT,
                                  double &dal, double &dbe)
{
    const double x = dt / amu, t = exp(-x), om = -expm1(-x), q = om / x;
    const double dq = (t - q) / dt;
    T = t; al = 0.5 * (1.0 - q); be = 0.5 * (q - t);
    dT = -t / amu; dal = -0.5 * dq; dbe = 0.5 * (dq + t / amu);
}

}  // namespace ocrt
