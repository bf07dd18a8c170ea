/*
 * ocrt_oracle.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * Double-precision CPU restatement of the reference OCRT atmospheric SOS
 * path, used as the checker of the HIP library.  Only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py load it; the
 * product never calls it.  Every function names the reference code it
 * restates (paths relative to code/OCRT_C/src of the reference).
 *
 * Pinning (how each part is checked, see tests/test_oracle_*.py):
 *   - whole solve vs the reference solver compiled from its own sources
 *     (oracle/_ref) on configs 0-3, all surfaces, IPSS on/off: rel <= 1e-12;
 *   - whole LUT solve vs the reference-held fixtures
 *     validation/ipss_2026-09-05/cases/*.csv (Rayleigh, black, tau 0.25/1.0,
 *     nt 400, IPSS on/off) and docs/validation/cross_check_fr631_20260823/
 *     atm_rhoIQU_bfo.csv (rough ocean, wind 3);
 *   - quadrature: exact polynomial moments; basis: closed forms at l = m and
 *     l = 2, orthogonality; kernel tables: addition theorem vs brute-force
 *     azimuth quadrature of the HG phase function; sweep: Beer-Lambert closed
 *     form for constant and linear sources; Fresnel: R + T = 1; Fourier
 *     reconstruction: the reference's own KAT values (tests/test_raa_convention.c);
 *     IPSS: the reference fixture tests/fixtures/ipss_oracle_geometry.csv and
 *     the plane-parallel limit kappa -> 1.
 */
#define _POSIX_C_SOURCE 200809L
#include "ocrt_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define PI 3.14159265358979323846

/* Emulated rounding of the scattering integral's operands, packed as
 * (q << 16) | (operator significant bits << 8) | field significant bits;
 * Fourier modes m < q round the field to 16 bits instead (a two-term split). */
static int g_bf16 = 0;
void oracle_set_bf16_emulation(int bf16) { g_bf16 = (bf16 == 1) ? ((8 << 8) | 8) : bf16; }

/* round to p significant bits, nearest-even (p = 8: bfloat16, 11: binary16) */
static double round_bits(double x, int p)
{
    if (p <= 0 || x == 0.0 || !isfinite(x)) return x;
    int e;
    const double f = frexp(x, &e);
    return ldexp(nearbyint(ldexp(f, p)), e - p);
}


/* ------------------------------------------------------------------------ */
/* Quadrature: positive half of the 2n-point Gauss-Legendre rule, ascending,
 * weights summing to 1 (rt_quadrature.c:20-61: Tricomi start, Newton on the
 * three-term recurrence, tolerance 1e-15 relative). */
static void legendre_pn(int N, double t, double *pn, double *pnm1)
{
    double p0 = 1.0, p1 = t;
    for (int k = 1; k < N; ++k) {
        const double p2 = ((2.0 * k + 1.0) * t * p1 - (double)k * p0) / (double)(k + 1);
        p0 = p1;
        p1 = p2;
    }
    *pn = p1;
    *pnm1 = p0;
}

int oracle_gauss_legendre_pos(int n, double *mu, double *w)
{
    if (n < 1 || !mu || !w) return -1;
    const int N = 2 * n;
    for (int k = 1; k <= n; ++k) {
        double t = cos(PI * ((double)k - 0.25) / ((double)N + 0.5));
        double pn, pm, dp;
        int ok = 0;
        for (int it = 0; it < 100; ++it) {
            legendre_pn(N, t, &pn, &pm);
            dp = (double)N * (t * pn - pm) / (t * t - 1.0);
            const double dt = -pn / dp;
            t += dt;
            if (fabs(dt) < 1.0e-15 * (fabs(t) + 1.0)) { ok = 1; break; }
        }
        if (!ok) return -2;
        legendre_pn(N, t, &pn, &pm);
        dp = (double)N * (t * pn - pm) / (t * t - 1.0);
        mu[n - k] = t;
        w[n - k] = 2.0 / ((1.0 - t * t) * dp * dp);
    }
    return 0;
}

/* Ring = GL nodes plus the view cosine at its sorted position with zero
 * weight, deduplicated within 1e-12 (rt_angle_grid.c:21-73). */
int oracle_ring(int n_gauss, double mu_view, double *mu, double *w, int *view_index)
{
    if (oracle_gauss_legendre_pos(n_gauss, mu, w) != 0) return -1;
    int n = n_gauss;
    if (view_index) *view_index = -1;
    if (mu_view < 0.0) return n;
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (mu[mid] < mu_view) lo = mid + 1; else hi = mid;
    }
    const int pos = lo;
    if (pos > 0 && fabs(mu[pos - 1] - mu_view) < 1.0e-12) { if (view_index) *view_index = pos - 1; return n; }
    if (pos < n && fabs(mu[pos] - mu_view) < 1.0e-12) { if (view_index) *view_index = pos; return n; }
    for (int i = n; i > pos; --i) { mu[i] = mu[i - 1]; w[i] = w[i - 1]; }
    mu[pos] = mu_view;
    w[pos] = 0.0;
    if (view_index) *view_index = pos;
    return n + 1;
}

/* ------------------------------------------------------------------------ */
/* US Standard Atmosphere 1962 column profile (rt_molecular_profile.c) and
 * the gravity formula it uses (rt_rayleigh.c:46-52). */
static const double US62_Z[8] = {0.0, 11.0, 20.0, 32.0, 47.0, 51.0, 71.0, 84.852};
static const double US62_P[8] = {101325.0, 22632.1, 5474.89, 868.019,
                                 110.906, 66.9389, 3.95642, 0.3734};

static double gravity_wgs84(double lat_deg, double alt_m)
{
    const double lat = lat_deg * PI / 180.0;
    const double s = sin(lat) * sin(lat);
    const double s2 = sin(2.0 * lat) * sin(2.0 * lat);
    const double g0 = 9.780327 * (1.0 + 0.0053024 * s - 0.0000058 * s2);
    return g0 - 3.086e-6 * alt_m;
}

static double clamp01(double x) { return x <= 0.0 ? 0.0 : (x >= 1.0 ? 1.0 : x); }

static double us62_pressure(double z)
{
    if (z <= US62_Z[0]) return US62_P[0];
    if (z >= US62_Z[7]) return US62_P[7];
    int lo = 0, hi = 7;
    while (hi - lo > 1) {
        const int mid = lo + (hi - lo) / 2;
        if (z < US62_Z[mid]) hi = mid; else lo = mid;
    }
    const double t = (z - US62_Z[lo]) / (US62_Z[hi] - US62_Z[lo]);
    const double a = log(US62_P[lo]), b = log(US62_P[hi]);
    return exp(a + t * (b - a));
}

static double us62_fraction_above(double z)
{
    if (z <= 0.0) return 1.0;
    const double g0 = gravity_wgs84(45.0, 0.0);
    if (z >= US62_Z[7]) {
        const double gt = gravity_wgs84(45.0, US62_Z[7] * 1000.0);
        return clamp01((US62_P[7] / gt) / (US62_P[0] / g0));
    }
    const double gz = gravity_wgs84(45.0, z * 1000.0);
    return clamp01((us62_pressure(z) / gz) / (US62_P[0] / g0));
}

double oracle_us62_grid_fraction_above(double z)
{
    if (z <= 0.0) return 1.0;
    if (z >= US62_Z[7]) return 0.0;
    const double ft = us62_fraction_above(US62_Z[7]);
    return clamp01((us62_fraction_above(z) - ft) / (1.0 - ft));
}

static double us62_altitude_from_fraction(double f)
{
    if (f >= 1.0) return 0.0;
    if (f <= 0.0) return US62_Z[7];
    double lo = 0.0, hi = US62_Z[7];
    for (int it = 0; it < 64; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (oracle_us62_grid_fraction_above(mid) > f) lo = mid; else hi = mid;
    }
    return 0.5 * (lo + hi);
}

static double aer_fraction_above(double z, double ztop, double H)
{
    if (z <= 0.0) return 1.0;
    if (z >= ztop) return 0.0;
    const double et = exp(-ztop / H);
    return (exp(-z / H) - et) / (1.0 - et);
}

static double layers_sum(const ocrt_instance_t *in, int row)
{
    double s = 0.0;
    for (int l = 0; l < in->n_layers; ++l) s += in->layers[(size_t)row * in->n_layers + l];
    return s;
}

static int has_aerosol(const ocrt_instance_t *in)
{
    if (in->water_tau > 0.0) return in->L_max > 0 && in->greek != NULL && in->water_omega_p > 0.0;
    if (in->layers) return in->L_max > 0 && in->greek != NULL && layers_sum(in, 1) > 0.0;
    return in->L_max > 0 && in->tau_a > 0.0 && in->greek != NULL;
}

/* Explicit layering: h[k] is the running sum of the layer extinctions
 * (Rayleigh + aerosol + absorption) from the top; level k >= 1 carries the
 * scattering fractions of the layer above it, xdel = ssa_a tau_a / dtau and
 * ydel = tau_R / dtau (rt_atm.c:343-360, renormalised for absorption as in
 * rt_atm.c:521-538); level 0 keeps the generator's top value (pure Rayleigh
 * when any Rayleigh is present, rt_atm.c:310-316). */
static int medium_layers(const ocrt_instance_t *in, int aer, double mus, double *h, double *ch, double *xdel,
                         double *ydel, double *z_km)
{
    const int nt = in->n_layers;
    const double *R = in->layers, *A = R + nt, *W = A + nt, *G = W + nt;
    double acc = 0.0;
    h[0] = 0.0;
    for (int l = 0; l < nt; ++l) {
        if (!(R[l] >= 0.0 && A[l] >= 0.0 && G[l] >= 0.0 && W[l] >= 0.0 && W[l] <= 1.0)) return -1;
        const double dt = R[l] + A[l] + G[l];
        if (!(dt > 0.0)) return -1;
        acc += dt;
        h[l + 1] = acc;
        xdel[l + 1] = aer ? W[l] * A[l] / dt : 0.0;
        ydel[l + 1] = R[l] / dt;
    }
    if (layers_sum(in, 0) > 0.0) { xdel[0] = 0.0; ydel[0] = 1.0; }
    else { xdel[0] = aer ? W[0] : 0.0; ydel[0] = 0.0; }
    for (int k = 0; k <= nt; ++k) {
        ch[k] = 0.5 * exp(-h[k] / mus);
        z_km[k] = 0.0;
    }
    return 0;
}

/* Medium (rt_atm.c:177-400): uniform-tau levels h[k] = k tau / nt,
 * ch = exp(-h/mu_s)/2, Rayleigh/aerosol scattering fractions by altitude
 * bisection of the US62 + exponential-aerosol columns. */
int oracle_medium(const ocrt_instance_t *in, double *h, double *ch, double *xdel,
                  double *ydel, double *z_km)
{
    const int nt = in->n_layers;
    const double mus = cos(in->sza_deg * PI / 180.0);
    const int aer = has_aerosol(in);
    if (in->water_tau > 0.0) {
        /* in-water column (rt_water_rt.c:2156-2176): uniform levels, fixed
         * albedos; the solar source scaled by the transmitted beam's weight */
        for (int k = 0; k <= nt; ++k) {
            const double hk = (double)k * in->water_tau / (double)nt;
            h[k] = hk;
            ch[k] = 0.5 * in->beam_scale * exp(-hk / mus);
            xdel[k] = aer ? in->water_omega_p : 0.0;
            ydel[k] = in->water_omega_w;
            z_km[k] = 1.0;
        }
        return 0;
    }
    if (in->layers) return medium_layers(in, aer, mus, h, ch, xdel, ydel, z_km);
    const double tauR = in->tau_R > 0.0 ? in->tau_R : 0.0;
    const double taua = aer ? in->tau_a : 0.0;
    const double tau = tauR + taua;
    if (!(tau > 0.0)) return -1;
    for (int k = 0; k <= nt; ++k) {
        const double hk = (double)k * tau / (double)nt;
        h[k] = hk;
        ch[k] = 0.5 * exp(-hk / mus);
        xdel[k] = 0.0;
        ydel[k] = 1.0;
        z_km[k] = (k == 0) ? 100.0 : us62_altitude_from_fraction(hk / tau);
    }
    if (!aer) return 0;
    const double ztop = 100.0;
    const double H = in->aer_h_km > 0.0 ? in->aer_h_km : 2.0;
    if (tauR == 0.0 && taua > 0.0) { xdel[0] = in->ssa_a; ydel[0] = 0.0; }
    else { xdel[0] = 0.0; ydel[0] = 1.0; }
    double zprev = ztop, ray_prev = 0.0, aer_prev = 0.0;
    z_km[0] = ztop;
    for (int j = 1; j <= nt; ++j) {
        double zj;
        if (j == nt) {
            zj = 0.0;
        } else {
            double lo = 0.0, hi = zprev;
            for (int it = 0; it < 64; ++it) {
                const double mid = 0.5 * (lo + hi);
                const double above = tauR * oracle_us62_grid_fraction_above(mid) +
                                     taua * aer_fraction_above(mid, ztop, H);
                if (above > h[j]) lo = mid; else hi = mid;
            }
            zj = 0.5 * (lo + hi);
        }
        z_km[j] = zj;
        const double ray = tauR * oracle_us62_grid_fraction_above(zj);
        const double ae = taua * aer_fraction_above(zj, ztop, H);
        const double dr = ray - ray_prev, da = ae - aer_prev, dt = dr + da;
        if (dt > 0.0) { xdel[j] = da * in->ssa_a / dt; ydel[j] = dr / dt; }
        else { xdel[j] = 0.0; ydel[j] = (tauR > 0.0) ? 1.0 : 0.0; }
        zprev = zj; ray_prev = ray; aer_prev = ae;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Angular basis (rt_angular_basis.c:126-276): factorial-normalised
 * associated Legendre functions without Condon-Shortley phase and the real
 * spin-2 pair, by their three-term recurrences from the diagonal l = m. */
static double diag_norm(int m)
{
    double v = 1.0;
    for (int i = 1; i <= m; ++i) v *= sqrt((double)(i + m) / (double)i) * 0.5;
    return v;
}

int oracle_basis(int l_max, int m, int nd, const double *mu, double *plm, double *rrl, double *rtl)
{
    if (m < 0 || l_max < m || l_max < 2) return -1;
    memset(plm, 0, sizeof(double) * (size_t)(l_max + 1) * nd);
    memset(rrl, 0, sizeof(double) * (size_t)(l_max + 1) * nd);
    memset(rtl, 0, sizeof(double) * (size_t)(l_max + 1) * nd);
#define B(a, l, j) a[(size_t)(l) * nd + (j)]
    const double nm = diag_norm(m);
    const double sn = (m >= 2) ? nm * sqrt((double)m / (double)(m + 1)) *
                                     sqrt((double)(m - 1) / (double)(m + 2)) : 0.0;
    for (int j = 0; j < nd; ++j) {
        const double x = mu[j];
        if (m == 0) {
            B(plm, 0, j) = 1.0;
            B(plm, 1, j) = x;
            B(plm, 2, j) = 0.5 * (3.0 * x * x - 1.0);
            B(rrl, 2, j) = 3.0 * (1.0 - x * x) / (2.0 * sqrt(6.0));
        } else if (m == 1) {
            B(plm, 1, j) = sqrt(0.5 * (1.0 - x * x));
            B(plm, 2, j) = x * B(plm, 1, j) * sqrt(3.0);
            const double r = sqrt(1.0 - x * x);
            B(rrl, 2, j) = -x * r * 0.5;
            B(rtl, 2, j) = -r * 0.5;
        } else {
            const double omx = 1.0 - x * x;
            B(plm, m, j) = nm * pow(omx, 0.5 * (double)m);
            const double env = pow(omx, 0.5 * (double)m - 1.0);
            B(rrl, m, j) = sn * (1.0 + x * x) * env;
            B(rtl, m, j) = sn * 2.0 * x * env;
        }
    }
    const int l0 = (m < 2) ? 2 : m;
    for (int l = l0; l < l_max; ++l) {
        const double fw = (2.0 * l + 1.0) / sqrt((double)(l + m + 1) * (double)(l - m + 1));
        const double bw = sqrt((double)(l + m) * (double)(l - m)) / (2.0 * l + 1.0);
        const double sc = (double)(l + 1) * (double)(2 * l + 1) /
                          sqrt((double)(l + 3) * (double)(l - 1) * (double)(l + m + 1) *
                               (double)(l - m + 1));
        const double dc = sqrt((double)(l + 2) * (double)(l - 2) * (double)(l + m) *
                               (double)(l - m)) / ((double)l * (double)(2 * l + 1));
        const double spc = 2.0 * (double)m / ((double)l * (double)(l + 1));
        for (int j = 0; j < nd; ++j) {
            const double x = mu[j];
            B(plm, l + 1, j) = fw * (x * B(plm, l, j) - bw * B(plm, l - 1, j));
            B(rrl, l + 1, j) = sc * (x * B(rrl, l, j) - spc * B(rtl, l, j) - dc * B(rrl, l - 1, j));
            B(rtl, l + 1, j) = sc * (x * B(rtl, l, j) - spc * B(rrl, l, j) - dc * B(rtl, l - 1, j));
        }
    }
#undef B
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Fresnel and Cox-Munk (shared/surface.c:57-64, 385-468, 514-646). */
static double slope_variance(double ws, int sigma_type)
{
    if (sigma_type == 1) return 0.003 + 0.00512 * fmax(0.01, ws);
    const double s = 0.0731 * sqrt(fmax(0.0, ws));
    return fmax(s * s, 1e-10);
}

static void fresnel_R(double mu_i, double n_water, int qc, double *M)
{
    for (int i = 0; i < 9; ++i) M[i] = 0.0;
    const double sin2 = fmax(0.0, 1.0 - mu_i * mu_i);
    const double st2 = (1.0 / n_water) * (1.0 / n_water) * sin2;
    if (st2 >= 1.0) { M[0] = M[4] = M[8] = 1.0; return; }
    const double mt = sqrt(1.0 - st2);
    const double rs = (1.0 * mu_i - n_water * mt) / (1.0 * mu_i + n_water * mt);
    const double rp = (n_water * mu_i - 1.0 * mt) / (n_water * mu_i + 1.0 * mt);
    const double Rs = rs * rs, Rp = rp * rp;
    const double q = (qc == 1) ? 0.5 * (Rp - Rs) : 0.5 * (Rs - Rp);
    M[0] = 0.5 * (Rs + Rp); M[1] = q; M[3] = q; M[4] = 0.5 * (Rs + Rp); M[8] = rs * rp;
}

void oracle_flat_fresnel(double mu, double n_water, int q_convention, double *MF)
{
    fresnel_R(mu, n_water, q_convention, MF);
}

static void rot_L(double c, double s, double *L)
{
    const double c2 = c * c - s * s, s2 = 2.0 * c * s;
    for (int i = 0; i < 9; ++i) L[i] = 0.0;
    L[0] = 1.0; L[4] = c2; L[5] = -s2; L[7] = s2; L[8] = c2;
}

static void mul3(const double *A, const double *B, double *C)
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += A[i * 3 + k] * B[k * 3 + j];
            C[i * 3 + j] = s;
        }
}

void oracle_coxmunk_R(double mu_out, double mu_in_dn, double cphi, double sphi, double ws,
                      int sigma_type, double n_water, int qc, double *R)
{
    for (int i = 0; i < 9; ++i) R[i] = 0.0;
    const double s2v = slope_variance(ws, sigma_type);
    const double mia = -mu_in_dn;
    const double s1 = sqrt(fmax(0.0, 1.0 - mu_out * mu_out));
    const double s2 = sqrt(fmax(0.0, 1.0 - mia * mia));
    double cT = mu_out * mia + s1 * s2 * cphi;
    if (cT > 1.0) cT = 1.0; else if (cT < -1.0) cT = -1.0;
    const double co = sqrt(fmax(0.0, 0.5 * (1.0 - cT)));
    if (co <= 1e-6) return;
    const double cb = (mu_out + mu_in_dn) / (2.0 * co);
    if (cb <= 1e-6) return;
    const double sT = sqrt(fmax(0.0, 1.0 - cT * cT));
    const double den = fmax(1e-12, sT);
    const double s2s = fmax(1e-12, s2), s1s = fmax(1e-12, s1);
    double ci1, si1, ci2, si2;
    if (s2 <= 1.0e-8) {
        ci1 = -copysign(1.0, mia) * cphi; si1 = sphi;
    } else {
        ci1 = (mu_out - mia * cT) / (den * s2s);
        si1 = s1 * sphi / den;
        const double q = sqrt(ci1 * ci1 + si1 * si1);
        if (q < 1e-12) { ci1 = 1.0; si1 = 0.0; } else { ci1 /= q; si1 /= q; }
    }
    if (s1 <= 1.0e-8) {
        ci2 = -copysign(1.0, mu_out) * cphi; si2 = sphi;
    } else {
        ci2 = (mia - mu_out * cT) / (den * s1s);
        si2 = s2 * sphi / den;
        const double q = sqrt(ci2 * ci2 + si2 * si2);
        if (q < 1e-12) { ci2 = 1.0; si2 = 0.0; } else { ci2 /= q; si2 /= q; }
    }
    double MF[9];
    fresnel_R(co, n_water, qc, MF);
    const double cb2 = cb * cb;
    const double tb2 = (1.0 - cb2) / fmax(cb2, 1e-8);
    const double Pslope = (1.0 / (PI * s2v)) * exp(-tb2 / s2v);
    const double sl = sqrt(s2v);
    double li = 0.0, lv = 0.0;
    if (mu_in_dn < 1.0 - 1e-10) {
        const double si = sqrt(fmax(1e-20, 1.0 - mu_in_dn * mu_in_dn));
        const double nu = (mu_in_dn / si) / sl;
        if (nu <= 20.0) li = 0.5 * (exp(-nu * nu) / (nu * sqrt(PI)) - erfc(nu));
    }
    if (mu_out < 1.0 - 1e-10) {
        const double sv = sqrt(fmax(1e-20, 1.0 - mu_out * mu_out));
        const double nu = (mu_out / sv) / sl;
        if (nu <= 20.0) lv = 0.5 * (exp(-nu * nu) / (nu * sqrt(PI)) - erfc(nu));
    }
    const double S = 1.0 / (1.0 + li + lv);
    const double brdf = (Pslope * S) / (4.0 * mu_in_dn * mu_out * cb2 * cb2);
    double L1[9], L2[9], T1[9], T2[9];
    rot_L(ci1, si1, L1);
    rot_L(ci2, si2, L2);
    mul3(MF, L1, T1);
    mul3(L2, T1, T2);
    for (int i = 0; i < 9; ++i) R[i] = T2[i] * brdf;
}

/* ------------------------------------------------------------------------ */
/* Rough-interface Mueller kernels of the coupled ocean (shared/surface.c).
 * kind 0 R_aa: air-side reflection (:514-646), oracle_coxmunk_R above;
 * kind 1 R_ww: water-side internal reflection (:1136-1249);
 * kind 2 T_wa: water -> air transmission (:648-773);
 * kind 3 T_aw: air -> water transmission (:80-136, :1476-1566).
 * Arguments follow the reference's Fourier-kernel layout: mu_o is the
 * outgoing and mu_i the incident zenith cosine of the pair (for T_aw mu_o is
 * the in-water and mu_i the air cosine). */

/* Flat-interface reflection between indices n1 (incidence) and n2, meridian
 * basis; identity under total internal reflection (surface.c:385-430). */
static void fresnel_R_general(double mu_i, double n1, double n2, int qc, double *M)
{
    for (int i = 0; i < 9; ++i) M[i] = 0.0;
    const double sin2 = fmax(0.0, 1.0 - mu_i * mu_i);
    const double st2 = (n1 / n2) * (n1 / n2) * sin2;
    if (st2 >= 1.0) { M[0] = M[4] = M[8] = 1.0; return; }
    const double mt = sqrt(1.0 - st2);
    const double rs = (n1 * mu_i - n2 * mt) / (n1 * mu_i + n2 * mt);
    const double rp = (n2 * mu_i - n1 * mt) / (n2 * mu_i + n1 * mt);
    const double Rs = rs * rs, Rp = rp * rp;
    const double q = (qc == 1) ? 0.5 * (Rp - Rs) : 0.5 * (Rs - Rp);
    M[0] = 0.5 * (Rs + Rp); M[1] = q; M[3] = q; M[4] = 0.5 * (Rs + Rp); M[8] = rs * rp;
}

/* Sancer bistatic shadowing 1 / (1 + Lambda_a + Lambda_b) (surface.c:591-612). */
static double sancer(double mu_a, double mu_b, double s2v)
{
    const double sl = sqrt(s2v);
    double lam[2] = {0.0, 0.0};
    const double mus[2] = {mu_a, mu_b};
    for (int q = 0; q < 2; ++q)
        if (mus[q] < 1.0 - 1e-10) {
            const double sn = sqrt(fmax(1e-20, 1.0 - mus[q] * mus[q]));
            const double nu = (mus[q] / sn) / sl;
            if (nu <= 20.0) lam[q] = 0.5 * (exp(-nu * nu) / (nu * sqrt(PI)) - erfc(nu));
        }
    return 1.0 / (1.0 + lam[0] + lam[1]);
}

/* cos/sin of the meridian-to-scattering-plane angle at one end of the path;
 * the pole limit and the normalisation follow surface.c:731-755. */
static void meridian_angle(double s_self, double mu_self, double num, double s_other_sphi, double den,
                           double cphi, double sphi, double *c, double *s)
{
    if (s_self <= 1.0e-8) { *c = -copysign(1.0, mu_self) * cphi; *s = sphi; return; }
    *c = num / (den * fmax(1e-12, s_self));
    *s = s_other_sphi / den;
    const double q = sqrt(*c * *c + *s * *s);
    if (q < 1e-12) { *c = 1.0; *s = 0.0; } else { *c /= q; *s /= q; }
}

/* L(2 i_out) . M . L(2 i_in) */
static void rotate_pair(const double *M, double c_in, double s_in, double c_out, double s_out, double *out)
{
    double L1[9], L2[9], T1[9];
    rot_L(c_in, s_in, L1);
    rot_L(c_out, s_out, L2);
    mul3(M, L1, T1);
    mul3(L2, T1, out);
}

static void rww_trig(double mu_out, double mu_in, double cphi, double sphi, double ws, int st, double nw,
                     int qc, double *R)
{
    for (int i = 0; i < 9; ++i) R[i] = 0.0;
    const double s2v = slope_variance(ws, st);
    const double mia = -mu_in;
    const double s1 = sqrt(fmax(0.0, 1.0 - mu_out * mu_out));
    const double s2 = sqrt(fmax(0.0, 1.0 - mia * mia));
    double cT = mu_out * mia + s1 * s2 * cphi;
    if (cT > 1.0) cT = 1.0; else if (cT < -1.0) cT = -1.0;
    const double co = sqrt(fmax(0.0, 0.5 * (1.0 - cT)));
    if (co <= 1e-6) return;
    const double cb = (mu_out + mu_in) / (2.0 * co);
    if (cb <= 1e-6) return;
    const double den = fmax(1e-12, sqrt(fmax(0.0, 1.0 - cT * cT)));
    double c1, n1, c2, n2;
    meridian_angle(s2, mia, mu_out - mia * cT, s1 * sphi, den, cphi, sphi, &c1, &n1);
    meridian_angle(s1, mu_out, mia - mu_out * cT, s2 * sphi, den, cphi, sphi, &c2, &n2);
    double MF[9];
    fresnel_R_general(co, nw, 1.0, qc, MF);   /* water -> air facet, TIR included */
    const double cb2 = cb * cb;
    const double tb2 = (1.0 - cb2) / fmax(cb2, 1e-8);
    const double Pslope = (1.0 / (PI * s2v)) * exp(-tb2 / s2v);
    const double brdf = Pslope * sancer(mu_in, mu_out, s2v) / (4.0 * mu_in * mu_out * cb2 * cb2);
    double T[9];
    rotate_pair(MF, c1, n1, c2, n2, T);
    for (int i = 0; i < 9; ++i) R[i] = T[i] * brdf;
}

/* Water -> air Walter BTDF with Cox-Munk slopes and Sancer shadowing; the
 * refraction half vector N = n p1 - p2 gives |N|^2 = n^2 + 1 - 2 n cos(Psi). */
static void twa_trig(double mu_air, double mu_water, double cphi, double sphi, double ws, int st, double nw,
                     int qc, double *T)
{
    for (int i = 0; i < 9; ++i) T[i] = 0.0;
    const double m1 = mu_water, m2 = mu_air, n = nw;
    if (m1 <= 1e-9 || m2 <= 1e-9) return;
    const double s2v = slope_variance(ws, st);
    const double s1 = sqrt(fmax(0.0, 1.0 - m1 * m1)), s2 = sqrt(fmax(0.0, 1.0 - m2 * m2));
    double cP = m1 * m2 + s1 * s2 * cphi;
    if (cP > 1.0) cP = 1.0; else if (cP < -1.0) cP = -1.0;
    const double N2 = n * n + 1.0 - 2.0 * n * cP;
    if (N2 <= 1e-12) return;
    const double N = sqrt(N2);
    const double cb = (n * m1 - m2) / N, cw = (n - cP) / N;
    if (cb <= 1e-6 || cw <= 1e-6) return;
    /* facet transmission water(n) -> air(1) at incidence cw */
    const double st2 = n * n * fmax(0.0, 1.0 - cw * cw);
    if (st2 >= 1.0) return;
    const double ma = sqrt(1.0 - st2);
    const double rs = (n * cw - ma) / (n * cw + ma);
    const double rp = (cw - n * ma) / (cw + n * ma);
    const double ts = 1.0 + rs, tp = n * (1.0 + rp);
    const double cb2 = cb * cb;
    const double tb2 = (1.0 - cb2) / fmax(cb2, 1e-8);
    const double Pslope = (1.0 / (PI * s2v)) * exp(-tb2 / s2v);
    const double common = (ma * ma) / (n * m1 * m2 * N2) * (Pslope / (cb2 * cb2)) * sancer(m1, m2, s2v);
    const double fs = common * ts * ts, fp = common * tp * tp;
    double MT[9] = {0};
    const double q = (qc == 1) ? 0.5 * (fp - fs) : 0.5 * (fs - fp);
    MT[0] = MT[4] = 0.5 * (fs + fp);
    MT[1] = MT[3] = q;
    MT[8] = common * ts * tp;
    const double den = fmax(1e-12, sqrt(fmax(0.0, 1.0 - cP * cP)));
    double c1, n1, c2, n2;
    meridian_angle(s1, m1, m2 - m1 * cP, s2 * sphi, den, cphi, sphi, &c1, &n1);
    meridian_angle(s2, m2, m1 - m2 * cP, s1 * sphi, den, cphi, sphi, &c2, &n2);
    rotate_pair(MT, c1, n1, c2, n2, T);
}

/* Air -> water microfacet transmission without shadowing (Walter et al. 2007,
 * Beckmann density with the Cox-Munk variance).  Directions point away from
 * the interface: i = (s_i, 0, mu_i) in air, o = (s_o cos phi, s_o sin phi,
 * -mu_o) in water, h = -(i + n o) normalised to the upper hemisphere. */
static void taw_trig(double mu_water, double mu_air, double cphi, double sphi, double ws, int st, double nw,
                     int qc, double *T)
{
    for (int i = 0; i < 9; ++i) T[i] = 0.0;
    const double s2v = slope_variance(ws, st);
    const double mi = mu_air, mo = mu_water, n = nw;
    if (s2v <= 1e-10 || mi <= 1e-9 || mo <= 1e-9) return;
    const double si = sqrt(fmax(0.0, 1.0 - mi * mi)), so = sqrt(fmax(0.0, 1.0 - mo * mo));
    const double ox = so * cphi, oy = so * sphi;
    double h[3] = {-(si + n * ox), -(n * oy), -(mi - n * mo)};
    const double hn = sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
    if (hn < 1e-12) return;
    for (int k = 0; k < 3; ++k) h[k] /= hn;
    if (h[2] < 0.0) for (int k = 0; k < 3; ++k) h[k] = -h[k];
    const double ih = si * h[0] + mi * h[2];
    const double oh = ox * h[0] + oy * h[1] - mo * h[2];
    if (ih <= 0.0 || h[2] <= 1e-9) return;
    const double tan2 = (1.0 - h[2] * h[2]) / (h[2] * h[2]);
    const double D = exp(-tan2 / s2v) / (PI * s2v * h[2] * h[2] * h[2] * h[2]);
    const double dh = (ih + n * oh) * (ih + n * oh);
    /* facet Fresnel air(1) -> water(n) at incidence ih */
    const double sti = sqrt(fmax(0.0, 1.0 - ih * ih)), stt = sti / n;
    double Ts = 0.0, Tp = 0.0;
    if (stt < 1.0) {
        const double ctt = sqrt(fmax(0.0, 1.0 - stt * stt));
        const double rs = (ih - n * ctt) / (ih + n * ctt), rp = (n * ih - ctt) / (n * ih + ctt);
        Ts = 1.0 - rs * rs;
        Tp = 1.0 - rp * rp;
    }
    if (dh < 1e-12) return;
    const double common = (fabs(ih) * fabs(oh)) / (mi * mo) * (n * n * D) / dh;
    const double fs = common * Ts, fp = common * Tp;
    double MT[9] = {0};
    const double q = (qc == 1) ? 0.5 * (fp - fs) : 0.5 * (fs - fp);
    MT[0] = MT[4] = 0.5 * (fs + fp);
    MT[1] = MT[3] = q;
    MT[8] = common * sqrt(fmax(0.0, Ts * Tp));
    /* rotations along the photon path: both signed cosines negative and the
     * azimuth shifted by pi (surface.c:1512-1524) */
    const double cr = -cphi, sr = -sphi, a = -mi, b = -mo;
    double cP = a * b + si * so * cr;
    if (cP > 1.0) cP = 1.0; else if (cP < -1.0) cP = -1.0;
    const double den = fmax(1e-12, sqrt(fmax(0.0, 1.0 - cP * cP)));
    double c1, n1, c2, n2;
    meridian_angle(si, a, b - a * cP, so * sr, den, cr, sr, &c1, &n1);
    meridian_angle(so, b, a - b * cP, si * sr, den, cr, sr, &c2, &n2);
    rotate_pair(MT, c1, n1, c2, n2, T);
}

void oracle_interface_trig(int kind, double mu_o, double mu_i, double cphi, double sphi, double ws,
                           int st, double nw, int qc, double *M)
{
    switch (kind) {
    case 0: oracle_coxmunk_R(mu_o, mu_i, cphi, sphi, ws, st, nw, qc, M); break;
    case 1: rww_trig(mu_o, mu_i, cphi, sphi, ws, st, nw, qc, M); break;
    case 2: twa_trig(mu_o, mu_i, cphi, sphi, ws, st, nw, qc, M); break;
    default: taw_trig(mu_o, mu_i, cphi, sphi, ws, st, nw, qc, M); break;
    }
}

/* Azimuthal Fourier coefficients of one interface kernel for every mode at
 * once (surface.c:839-938, 1386-1447, 1583-1721, 1731-1811 per mode; the
 * all-mode form :1263-1380): midpoint phi rule, cos for II IQ QI QQ UU and
 * sin for the other entries, normalisation 1/(2 pi) at m = 0 and 1/pi above;
 * IU and QU negated for m > 0 except for R_ww.  T_aw of a flat surface
 * (slope variance <= 1e-10) has no rough kernel: -2. */
int oracle_interface_fourier(int kind, const double *mu_o, int n_o, const double *mu_i, int n_i, int m_max,
                             int n_phi, double ws, int st, double nw, int qc, double *out)
{
    static const int cos_entry[9] = {1, 1, 0, 1, 1, 0, 0, 0, 1};
    if (n_o < 1 || n_i < 1 || m_max < 0 || n_phi < 4) return -1;
    if (kind == 3 && slope_variance(ws, st) <= 1e-10) return -2;
    const double dphi = 2.0 * PI / (double)n_phi;
    const size_t slab = (size_t)n_o * n_i * 9;
    double *W = (double *)malloc(sizeof(double) * 2 * (size_t)(m_max + 1) * n_phi);
    double *acc = (double *)malloc(sizeof(double) * 9 * (size_t)(m_max + 1));
    if (!W || !acc) { free(W); free(acc); return -3; }
    for (int m = 0; m <= m_max; ++m)
        for (int p = 0; p < n_phi; ++p) {
            const double phi = (p + 0.5) * dphi;
            W[(2 * (size_t)m) * n_phi + p] = cos(m * phi);
            W[(2 * (size_t)m + 1) * n_phi + p] = sin(m * phi);
        }
    for (int o = 0; o < n_o; ++o)
        for (int i = 0; i < n_i; ++i) {
            memset(acc, 0, sizeof(double) * 9 * (size_t)(m_max + 1));
            for (int p = 0; p < n_phi; ++p) {
                const double phi = (p + 0.5) * dphi;
                double R[9];
                oracle_interface_trig(kind, mu_o[o], mu_i[i], cos(phi), sin(phi), ws, st, nw, qc, R);
                for (int m = 0; m <= m_max; ++m) {
                    const double cm = W[(2 * (size_t)m) * n_phi + p];
                    const double sm = W[(2 * (size_t)m + 1) * n_phi + p];
                    for (int kl = 0; kl < 9; ++kl) acc[m * 9 + kl] += R[kl] * (cos_entry[kl] ? cm : sm);
                }
            }
            for (int m = 0; m <= m_max; ++m) {
                const double nrm = (m == 0) ? 1.0 / (2.0 * PI) : 1.0 / PI;
                double *d = out + (size_t)m * slab + ((size_t)o * n_i + i) * 9;
                for (int kl = 0; kl < 9; ++kl) d[kl] = acc[m * 9 + kl] * (dphi * nrm);
                if (m > 0 && kind != 1) { d[2] = -d[2]; d[5] = -d[5]; }
            }
        }
    free(W);
    free(acc);
    return 0;
}

int oracle_coxmunk_fourier(const double *mu_o, int n_o, const double *mu_i, int n_i, int m_max,
                           int n_phi, double ws, int st, double nw, int qc, double *out)
{
    return oracle_interface_fourier(0, mu_o, n_o, mu_i, n_i, m_max, n_phi, ws, st, nw, qc, out);
}

/* Slope-integrated direct-beam irradiance terms (no shadowing), trapezoid rule
 * on 101 x 101 slopes over +-6 sigma:
 *   kind 3: air -> water transmittance T_aw(mu_sun) (surface.c:138-198);
 *   kind 1: water-side reflectance R_ww(mu_up), the first-encounter form
 *           (surface.c:229-238, 302-368). */
/* facet Fresnel term of the direct integrals at local incidence cosine c:
 * air -> water transmittance 1 - R (kind 3) or the water -> air reflectance,
 * 1 beyond the critical angle (kind 1) */
static double direct_facet(int kind, double c, double nw)
{
    if (kind == 3) {
        const double stt = sqrt(fmax(0.0, 1.0 - c * c)) / nw, ctt = sqrt(fmax(0.0, 1.0 - stt * stt));
        const double rs = (c - nw * ctt) / (c + nw * ctt), rp = (nw * c - ctt) / (nw * c + ctt);
        return 1.0 - 0.5 * (rs * rs + rp * rp);
    }
    const double sa = nw * sqrt(fmax(0.0, 1.0 - c * c));
    if (sa >= 1.0) return 1.0;
    const double ca = sqrt(fmax(0.0, 1.0 - sa * sa));
    const double rs = (nw * c - ca) / (nw * c + ca), rp = (c - nw * ca) / (c + nw * ca);
    return 0.5 * (rs * rs + rp * rp);
}

double oracle_interface_direct(int kind, double mu, double ws, int st, double nw)
{
    if (mu <= 1e-9) return kind == 3 ? 0.0 : 1.0;
    const double s2 = slope_variance(ws, st);
    const double sn = sqrt(fmax(0.0, 1.0 - mu * mu));
    if (s2 <= 1e-10) return direct_facet(kind, mu, nw);
    double acc = 0.0;
    const double sig = sqrt(s2), L = 6.0 * sig;
    const int N = 101;
    const double dz = 2.0 * L / (double)(N - 1);
    for (int i = 0; i < N; ++i) {
        const double zx = -L + i * dz;
        const double wx = (i == 0 || i == N - 1) ? 0.5 : 1.0;
        const double proj = mu + zx * sn;
        if (proj <= 0.0) continue;
        for (int j = 0; j < N; ++j) {
            const double zy = -L + j * dz;
            const double wy = (j == 0 || j == N - 1) ? 0.5 : 1.0;
            const double r2 = zx * zx + zy * zy;
            const double F = direct_facet(kind, proj / sqrt(1.0 + r2), nw);
            acc += wx * wy * ((1.0 / (PI * s2)) * exp(-r2 / s2)) * proj * F;
        }
    }
    const double v = acc * dz * dz / mu;
    return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
}

/* Direct sunglint (surface.c:1813-1907). */
static void sunglint_rho(double mu_v, double phi_v, double mu_0, double ws, int st, double nw,
                         double tau, int qc, double *rho)
{
    const double dphi = phi_v;
    const double trans = exp(-tau / mu_0) * exp(-tau / mu_v);
    if (ws < 1e-6) {
        const double sfov = 0.5 * PI / 180.0;
        const double st0 = sqrt(fmax(0.0, 1.0 - mu_0 * mu_0));
        const double smu = sfov * fmax(st0, 1e-3);
        double de = dphi;
        while (de > PI) de -= 2.0 * PI;
        while (de < -PI) de += 2.0 * PI;
        const double dmu = mu_v - mu_0;
        const double ex = -0.5 * ((dmu * dmu) / (smu * smu) + (de * de) / (sfov * sfov));
        if (ex < -50.0) { rho[0] = rho[1] = rho[2] = 0.0; return; }
        double MF[9];
        fresnel_R(mu_0, nw, qc, MF);
        const double den = fmax(2.0 * mu_0 * st0 * mu_0 * mu_v, 1e-12);
        const double r = 4.0 * PI * trans / den * exp(ex);
        rho[0] = MF[0] * r; rho[1] = MF[3] * r; rho[2] = MF[6] * r;
        return;
    }
    double R[9];
    oracle_coxmunk_R(mu_v, mu_0, cos(dphi), sin(dphi), ws, st, nw, qc, R);
    rho[0] = PI * R[0] * trans; rho[1] = PI * R[3] * trans; rho[2] = PI * R[6] * trans;
}

/* ------------------------------------------------------------------------ */
/* Fourier reconstruction (rt_fourier.c:9-33). */
double oracle_reconstruct_cos(const double *v, int m_max, double raa)
{
    const double phi = raa + PI;
    double s = v[0];
    for (int m = 1; m <= m_max; ++m) s += 2.0 * v[m] * cos((double)m * phi);
    return s;
}

double oracle_reconstruct_sin(const double *v, int m_max, double raa)
{
    if (m_max < 1) return 0.0;
    const double phi = raa + PI;
    double s = 0.0;
    for (int m = 1; m <= m_max; ++m) s += 2.0 * v[m] * sin((double)m * phi);
    return s;
}

/* Spin-2 completion of the diagonal families (rt_phase_moments.c:22-64). */
int oracle_complete_spherical(int L, const double *b11, const double *b22, const double *d33,
                              double *alpha, double *zeta)
{
    for (int l = 0; l <= L; ++l) alpha[l] = zeta[l] = 0.0;
    for (int l = 2; l <= L; ++l) {
        const double e = (double)l;
        const double lower = e * (e - 1.0) / ((e + 1.0) * (e + 2.0));
        const double cpl = 4.0 * (2.0 * e + 1.0) / (e * (e - 1.0) * (e + 1.0) * (e + 2.0));
        const double base = (e - 1.0) * (e - 1.0);
        double sb = 0.0, sd = 0.0, ob = 0.0, od = 0.0;
        for (int o = l - 1; o >= 0; --o) {
            const int gap = l - o;
            const double wgt = base - 1.5 * (double)(gap - 1) * (double)(l + o);
            if ((gap & 1) == 0) { sb += wgt * b11[o]; sd += wgt * d33[o]; }
            else { ob += wgt * b22[o]; od += wgt * d33[o]; }
        }
        zeta[l] = lower * d33[l] - cpl * (sd - ob);
        alpha[l] = lower * b11[l] - cpl * (sb - od);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* IPSS spherical-shell single scattering and kappa (rt_ipss.c:24-644). */
typedef struct {
    int nlay;
    double *h_edge, *beta, *omega, *w_ray, *w_aer, *w_src, *tau_above;
    double Re, H, mu_sun, depol;
    int L_aer;
    const double *betal;
    int n_ang;                       /* value-kernel P11 table, 0: none */
    const double *th_ang, *p11_ang;
    double *ts, *los_t, *los_tau;
    int *los_lay;
    int los_n;
    double glx[2], glw[2];
} ipss_t;

static void ipss_free(ipss_t *G)
{
    free(G->h_edge); free(G->beta); free(G->omega); free(G->w_ray); free(G->w_aer);
    free(G->w_src); free(G->tau_above); free(G->ts); free(G->los_t); free(G->los_tau);
    free(G->los_lay);
    memset(G, 0, sizeof(*G));
}

static void ipss_gl2(double *x, double *w)
{
    const int n = 2;
    for (int i = 0; i < (n + 1) / 2; ++i) {
        double z = cos(PI * (i + 0.75) / (n + 0.5)), pp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = 0.0;
            for (int j = 0; j < n; ++j) {
                const double p2 = p1;
                p1 = p0;
                p0 = ((2.0 * j + 1.0) * z * p1 - j * p2) / (j + 1.0);
            }
            pp = n * (z * p0 - p1) / (z * z - 1.0);
            const double dz = -p0 / pp;
            z += dz;
            if (fabs(dz) < 1e-15) break;
        }
        x[i] = -z; x[n - 1 - i] = z;
        w[i] = 2.0 / ((1.0 - z * z) * pp * pp); w[n - 1 - i] = w[i];
    }
}

static int ipss_init(ipss_t *G, int nt, const double *h, const double *xdel, const double *ydel,
                     const double *z, double mus, double depol, int aer, int L, const double *betal)
{
    memset(G, 0, sizeof(*G));
    G->nlay = nt; G->Re = 6371.0; G->mu_sun = mus; G->depol = depol; G->L_aer = -1;
    G->h_edge = malloc(sizeof(double) * (nt + 1));
    G->beta = malloc(sizeof(double) * nt);
    G->omega = malloc(sizeof(double) * nt);
    G->tau_above = malloc(sizeof(double) * (nt + 1));
    G->ts = malloc(sizeof(double) * (2 * nt + 8));
    G->los_t = malloc(sizeof(double) * (2 * nt + 8));
    G->los_tau = malloc(sizeof(double) * (2 * nt + 8));
    G->los_lay = malloc(sizeof(int) * (2 * nt + 8));
    if (!G->h_edge || !G->beta || !G->omega || !G->tau_above || !G->ts || !G->los_t ||
        !G->los_tau || !G->los_lay) { ipss_free(G); return -1; }
    for (int k = 0; k <= nt; ++k) G->h_edge[k] = z[nt - k];
    G->H = G->h_edge[nt];
    for (int j = 0; j < nt; ++j) {
        const int k = nt - 1 - j;
        const double dz = G->h_edge[j + 1] - G->h_edge[j];
        const double dtau = h[k + 1] - h[k];
        if (!(dz > 0.0)) { ipss_free(G); return -2; }
        G->beta[j] = dtau > 0.0 ? dtau / dz : 0.0;
        G->omega[j] = dtau > 0.0 ? 1.0 : 0.0;
    }
    int have_aer = 0;
    for (int k = 0; k < nt; ++k) if (xdel[k + 1] > 0.0) { have_aer = 1; break; }
    if (have_aer && aer) {
        G->w_ray = malloc(sizeof(double) * nt);
        G->w_aer = malloc(sizeof(double) * nt);
        G->w_src = malloc(sizeof(double) * nt);
        for (int j = 0; j < nt; ++j) {
            const int k = nt - 1 - j;
            G->w_ray[j] = ydel[k + 1] > 0.0 ? ydel[k + 1] : 0.0;
            G->w_aer[j] = xdel[k + 1] > 0.0 ? xdel[k + 1] : 0.0;
            G->w_src[j] = G->w_ray[j] + G->w_aer[j];
        }
        G->L_aer = L;
        G->betal = betal;
    }
    G->tau_above[nt] = 0.0;
    for (int j = nt - 1; j >= 0; --j)
        G->tau_above[j] = G->tau_above[j + 1] + G->beta[j] * (G->h_edge[j + 1] - G->h_edge[j]);
    ipss_gl2(G->glx, G->glw);
    return 0;
}

static double ipss_src(const ipss_t *G, int j) { return G->w_src ? G->w_src[j] : G->omega[j]; }

/* Linear interpolation in scattering angle on an ascending P11 table,
 * clamped to the end values (rt_solver.c:190-202, rt_ipss.c:134-146). */
static double p11_table_at(int n, const double *th, const double *p11, double cth)
{
    if (cth > 1.0) cth = 1.0; else if (cth < -1.0) cth = -1.0;
    const double theta = acos(cth) * 180.0 / PI;
    if (theta <= th[0]) return p11[0];
    if (theta >= th[n - 1]) return p11[n - 1];
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) { const int mid = (lo + hi) >> 1; if (th[mid] <= theta) lo = mid; else hi = mid; }
    const double t = (theta - th[lo]) / (th[hi] - th[lo]);
    return p11[lo] + t * (p11[hi] - p11[lo]);
}

static double ipss_phase_aer(const ipss_t *G, double ct)
{
    if (G->n_ang > 1) return p11_table_at(G->n_ang, G->th_ang, G->p11_ang, ct);
    if (G->L_aer >= 0 && G->betal) {
        double p0 = 1.0, p1 = ct, s = G->betal[0];
        if (G->L_aer >= 1) s += G->betal[1] * p1;
        for (int l = 2; l <= G->L_aer; ++l) {
            const double p2 = ((2.0 * l - 1.0) * ct * p1 - (l - 1.0) * p0) / l;
            s += G->betal[l] * p2;
            p0 = p1; p1 = p2;
        }
        return s > 0.0 ? s : 0.0;
    }
    return 1.0;
}

static void ipss_weights(ipss_t *G, double vza, double raa)
{
    if (!G->w_src) return;
    const double sn = sqrt(fmax(0.0, 1.0 - G->mu_sun * G->mu_sun));
    const double tv = vza * PI / 180.0, fp = (180.0 - raa) * PI / 180.0;
    const double ct = sn * sin(tv) * cos(fp) - G->mu_sun * cos(tv);
    const double d = (G->depol >= 0.0 && G->depol < 1.0) ? G->depol : 0.0;
    const double b2 = (1.0 - d) / (2.0 + d);
    const double PR = 1.0 + b2 * 0.5 * (3.0 * ct * ct - 1.0);
    const double PA = ipss_phase_aer(G, ct);
    for (int j = 0; j < G->nlay; ++j) G->w_src[j] = G->w_ray[j] * PR + G->w_aer[j] * PA;
}

/* Optical depth along the ray r0 + t u from ray-shell crossings sorted by
 * construction, each interval assigned to the layer of its midpoint, the ray
 * truncated at the first ground hit (rt_ipss.c:209-316). */
static double ipss_ray_tau(const ipss_t *G, const double r0[3], const double uin[3], int *hit_out)
{
    const double un = sqrt(uin[0] * uin[0] + uin[1] * uin[1] + uin[2] * uin[2]);
    const double u[3] = {uin[0] / un, uin[1] / un, uin[2] / un};
    const double b = r0[0] * u[0] + r0[1] * u[1] + r0[2] * u[2];
    const double c0 = r0[0] * r0[0] + r0[1] * r0[1] + r0[2] * r0[2];
    const int nl = G->nlay;
    double tstop = INFINITY;
    int hit = 0;
    const double dg = b * b - (c0 - G->Re * G->Re);
    if (dg >= 0.0) {
        const double q = sqrt(dg), c[2] = {-b - q, -b + q};
        for (int i = 0; i < 2; ++i) if (c[i] > 1e-9 && c[i] < tstop) { tstop = c[i]; hit = 1; }
    }
    double *ts = G->ts;
    const int cap = 2 * nl + 8;
    int n = 0;
    ts[n++] = 0.0;
    for (int k = nl; k >= 0; --k) {
        const double R = G->Re + G->h_edge[k], d = b * b - (c0 - R * R);
        if (d < 0.0) continue;
        const double t = -b - sqrt(d);
        if (t > 1e-9 && t <= tstop && n < cap) ts[n++] = t;
    }
    for (int k = 0; k <= nl; ++k) {
        const double R = G->Re + G->h_edge[k], d = b * b - (c0 - R * R);
        if (d < 0.0) continue;
        const double t = -b + sqrt(d);
        if (t > 1e-9 && t <= tstop && n < cap) ts[n++] = t;
    }
    if (hit && n < cap) ts[n++] = tstop;
    int m = 0;
    for (int i = 0; i < n; ++i) if (i == 0 || ts[i] > ts[m - 1] + 1e-12) ts[m++] = ts[i];
    double tau = 0.0;
    for (int i = 0; i + 1 < m; ++i) {
        const double tm = 0.5 * (ts[i] + ts[i + 1]);
        const double hm = sqrt(tm * tm + 2.0 * b * tm + c0) - G->Re;
        if (hm < 0.0 || hm > G->h_edge[nl]) continue;
        int lo = 0, hi = nl;
        while (hi - lo > 1) { const int mid = (lo + hi) / 2; if (G->h_edge[mid] <= hm) lo = mid; else hi = mid; }
        if (lo > nl - 1) lo = nl - 1;
        tau += (ts[i + 1] - ts[i]) * G->beta[lo];
    }
    if (hit_out) *hit_out = hit;
    return tau;
}

/* Solar slant optical depth from r0 to TOA; INFINITY in the Earth's shadow
 * (rt_ipss.c:320-333). */
static double ipss_tau_sun(const ipss_t *G, const double r0[3])
{
    const double mu = G->mu_sun, sn = sqrt(fmax(0.0, 1.0 - mu * mu));
    const double u[3] = {-sn, 0.0, mu};
    int hit = 0;
    const double t = ipss_ray_tau(G, r0, u, &hit);
    return hit ? INFINITY : t;
}

static int ipss_layer_of(const ipss_t *G, double h)
{
    if (h <= G->h_edge[0]) return 0;
    if (h >= G->h_edge[G->nlay]) return G->nlay - 1;
    int lo = 0, hi = G->nlay;
    while (hi - lo > 1) { const int mid = (lo + hi) / 2; if (G->h_edge[mid] <= h) lo = mid; else hi = mid; }
    return lo > G->nlay - 1 ? G->nlay - 1 : lo;
}

/* Surface-anchored spherical single scattering (rt_ipss.c:427-567). */
static double ipss_ss_surface(ipss_t *G, double vza, double raa)
{
    const double tv = vza * PI / 180.0, fp = (180.0 - raa) * PI / 180.0;
    const double V[3] = {sin(tv) * cos(fp), sin(tv) * sin(fp), cos(tv)};
    const double R = G->Re + G->H, bq = G->Re * V[2];
    const double disc = bq * bq + R * R - G->Re * G->Re;
    if (disc <= 0.0) return 0.0;
    const double s = -bq + sqrt(disc);
    if (!(s > 0.0)) return 0.0;
    const double rA[3] = {s * V[0], s * V[1], G->Re + s * V[2]};
    const double dmax = s;
    /* line-of-sight crossing table */
    const int nl = G->nlay, cap = 2 * nl + 8;
    const double b = -(rA[0] * V[0] + rA[1] * V[1] + rA[2] * V[2]);
    const double c0 = rA[0] * rA[0] + rA[1] * rA[1] + rA[2] * rA[2];
    int n = 0;
    G->los_t[n++] = 0.0;
    for (int k = nl; k >= 0; --k) {
        const double Rk = G->Re + G->h_edge[k], d = b * b - (c0 - Rk * Rk);
        if (d < 0.0) continue;
        const double t = -b - sqrt(d);
        if (t > 1e-12 && t <= dmax && n < cap) G->los_t[n++] = t;
    }
    for (int k = 0; k <= nl; ++k) {
        const double Rk = G->Re + G->h_edge[k], d = b * b - (c0 - Rk * Rk);
        if (d < 0.0) continue;
        const double t = -b + sqrt(d);
        if (t > 1e-12 && t <= dmax && n < cap) G->los_t[n++] = t;
    }
    if (n < cap) G->los_t[n++] = dmax;
    int m = 0;
    for (int i = 0; i < n; ++i) if (i == 0 || G->los_t[i] > G->los_t[m - 1] + 1e-12) G->los_t[m++] = G->los_t[i];
    double tau = 0.0;
    for (int i = 0; i + 1 < m; ++i) {
        const double tm = 0.5 * (G->los_t[i] + G->los_t[i + 1]);
        const double rm = sqrt(tm * tm + 2.0 * b * tm + c0);
        const int j = ipss_layer_of(G, rm - G->Re);
        G->los_lay[i] = j;
        G->los_tau[i] = tau;
        tau += (G->los_t[i + 1] - G->los_t[i]) * G->beta[j];
    }
    G->los_tau[m - 1] = tau;
    G->los_lay[m - 1] = G->los_lay[m > 1 ? m - 2 : 0];
    G->los_n = m;
    double I1 = 0.0;
    for (int i = 0; i + 1 < G->los_n; ++i) {
        const int j = G->los_lay[i];
        if (!(G->beta[j] > 0.0)) continue;
        const double sa = G->los_t[i], sb = G->los_t[i + 1];
        const double half = 0.5 * (sb - sa), mid = 0.5 * (sa + sb);
        for (int q = 0; q < 2; ++q) {
            const double sq = mid + half * G->glx[q];
            const double wg = half * G->glw[q];
            const double rB[3] = {rA[0] - sq * V[0], rA[1] - sq * V[1], rA[2] - sq * V[2]};
            const double hB = sqrt(rB[0] * rB[0] + rB[1] * rB[1] + rB[2] * rB[2]) - G->Re;
            if (hB < 0.0 || hB > G->H) continue;
            const double tview = G->los_tau[i] + (sq - sa) * G->beta[j];
            const double tsun = ipss_tau_sun(G, rB);
            if (!isfinite(tsun)) continue;
            I1 += wg * G->beta[j] * ipss_src(G, j) * exp(-tview - tsun);
        }
    }
    return I1;
}

static double ipss_ss_pp(const ipss_t *G, double vza)
{
    const double mv = cos(vza * PI / 180.0);
    if (!(mv > 0.0)) return 0.0;
    const double C = 1.0 / mv + 1.0 / G->mu_sun;
    double I1 = 0.0;
    for (int j = 0; j < G->nlay; ++j) {
        if (!(G->beta[j] > 0.0)) continue;
        I1 += ipss_src(G, j) / (mv * C) * (exp(-C * G->tau_above[j + 1]) - exp(-C * G->tau_above[j]));
    }
    return I1;
}

static double ipss_kappa(ipss_t *G, double vza, double raa)
{
    ipss_weights(G, vza, raa);
    const double ss = ipss_ss_surface(G, vza, raa);
    const double pp = ipss_ss_pp(G, vza);
    if (!(pp > 1.0e-300) || !isfinite(ss)) return 1.0;
    return ss / pp;
}

static double ipss_tau_dir(const ipss_t *G)
{
    const double rB[3] = {0.0, 0.0, G->Re};
    const double t = ipss_tau_sun(G, rB);
    return isfinite(t) ? t : -1.0;
}

/* ------------------------------------------------------------------------ */
/* The SOS solve. */
typedef struct {
    int n, nt, L, dirs, lmax, aer, La, m_max, vidx;
    double mus, muv, tau_tot;
    double beta0, beta2, gamma2, alpha2;
    const double *gb_, *gg_, *ga_, *gz_;   /* greek beta, gamma, alpha, zeta */
    int n_ang;                              /* value-kernel P11 table, 0: moment kernel */
    const double *th_ang, *p11_ang;
    double *rm_s, *gb_s, *rm, *gb;         /* signed ring, offset pointers [-n..n] */
    double *h, *ch, *xd, *yd, *z;
    double *T;                              /* [n][nt] Beer-Lambert factors */
} med_t;

typedef struct {
    double *plm, *rrl, *rtl;                /* [(lmax+1)][dirs] */
    double *P, *GR, *GT, *ARR, *ART, *ATT;  /* [(n+1)][dirs] rows j = 0..n */
} ang_t;

#define FI(k, j) ((size_t)(k) * dirs + (size_t)((j) + n))

static double coef(const double *a, int l, int La) { return (a && l <= La) ? a[l] : 0.0; }

/* Kernel tables (rt_angular_basis.c:278-358) for m. */
static void build_tables(const med_t *M, int m, ang_t *A)
{
    const int n = M->n, dirs = M->dirs, lm = M->lmax;
    const size_t cnt = (size_t)(n + 1) * dirs;
    memset(A->P, 0, cnt * sizeof(double)); memset(A->GR, 0, cnt * sizeof(double));
    memset(A->GT, 0, cnt * sizeof(double)); memset(A->ARR, 0, cnt * sizeof(double));
    memset(A->ART, 0, cnt * sizeof(double)); memset(A->ATT, 0, cnt * sizeof(double));
    if (!M->aer) return;
    for (int j = 0; j <= n; ++j)
        for (int k = -n; k <= n; ++k) {
            double p = 0, gr = 0, gt = 0, rr = 0, rt = 0, tt = 0;
            for (int l = m; l <= lm; ++l) {
                const double pj = A->plm[(size_t)l * dirs + j + n], pk = A->plm[(size_t)l * dirs + k + n];
                const double rj = A->rrl[(size_t)l * dirs + j + n], rk = A->rrl[(size_t)l * dirs + k + n];
                const double tj = A->rtl[(size_t)l * dirs + j + n], tk = A->rtl[(size_t)l * dirs + k + n];
                const double be = coef(M->gb_, l, M->La), ga = coef(M->gg_, l, M->La);
                const double al = coef(M->ga_, l, M->La), ze = coef(M->gz_, l, M->La);
                p += pj * pk * be;
                gr += pj * rk * ga;
                gt += pj * tk * ga;
                tt += tj * tk * al + rj * rk * ze;
                rr += tj * tk * ze + rj * rk * al;
                rt += tj * rk * al + rj * tk * ze;
            }
            const size_t ix = (size_t)j * dirs + k + n;
            A->P[ix] = p; A->GR[ix] = gr; A->GT[ix] = gt;
            A->ARR[ix] = rr; A->ART[ix] = rt; A->ATT[ix] = tt;
        }
}

/* Value kernel of the aerosol I -> I table (rt_solver.c:204-240): the
 * azimuth mean of the interpolated P11 table over 720 midpoints,
 * P[j][k] = (1/720) sum_q P11(mu_j mu_k + s_j s_k cos phi_q) cos(m phi_q),
 * rows j = 0..n (row 0 the solar direction -mu_s) and every column k. */
static void value_table(const med_t *M, int m, ang_t *A)
{
    const int n = M->n, dirs = M->dirs, nphi = 720;
    const double two_pi = 2.0 * PI;
    for (int j = 0; j <= n; ++j) {
        const double mj = M->rm[j], sj = sqrt(fmax(0.0, 1.0 - mj * mj));
        for (int k = -n; k <= n; ++k) {
            const double mk = M->rm[k], sk = sqrt(fmax(0.0, 1.0 - mk * mk));
            double acc = 0.0;
            for (int q = 0; q < nphi; ++q) {
                const double phi = two_pi * ((double)q + 0.5) / (double)nphi;
                const double cth = mj * mk + sj * sk * cos(phi);
                acc += p11_table_at(M->n_ang, M->th_ang, M->p11_ang, cth) * cos((double)m * phi);
            }
            A->P[(size_t)j * dirs + (size_t)(k + n)] = acc / (double)nphi;
        }
    }
}

#define TB(tab, j, k) (tab)[(size_t)(j) * dirs + (size_t)((k) + n)]
#define XPL(j) A->plm[(size_t)2 * dirs + (j) + n]
#define XRL(j) A->rrl[(size_t)2 * dirs + (j) + n]
#define XTL(j) A->rtl[(size_t)2 * dirs + (j) + n]

/* Single-scattering source (rt_primary_source.c:17-198), beam_q = 0. */
static void primary_source(const med_t *M, int m, const ang_t *A, double *SI, double *SQ, double *SU)
{
    const int n = M->n, dirs = M->dirs;
    const double b0 = (m == 0) ? M->beta0 : 0.0;
    const int mol = (m <= 2);
    for (int k = 0; k <= M->nt; ++k) {
        const double beam = M->ch[k], x = M->xd[k], y = M->yd[k];
        for (int j = -n; j <= n; ++j) {
            const double mi = b0 + M->beta2 * XPL(j) * XPL(0);
            const double mq = mol ? M->gamma2 * XRL(j) * XPL(0) : 0.0;
            const double mu_ = mol ? M->gamma2 * XTL(j) * XPL(0) : 0.0;
            SI[FI(k, j)] = beam * (x * TB(A->P, 0, j) + y * mi);
            SQ[FI(k, j)] = beam * (x * TB(A->GR, 0, j) + y * mq);
            SU[FI(k, j)] = -beam * (x * TB(A->GT, 0, j) + y * mu_);
        }
        SI[FI(k, 0)] = SQ[FI(k, 0)] = SU[FI(k, 0)] = NAN;
    }
}

/* Formal solution with a linear-in-tau source (rt_layer_transport.c:89-208). */
/* Formal solution with boundary values: bc at the bottom of the upward
 * columns, tb at the top of the downward ones (either may be NULL). */
static int sweep_bc(const med_t *M, const double *S, const double *bc, const double *tb, double *F)
{
    const int n = M->n, dirs = M->dirs, nt = M->nt;
    const double *tau = M->h;
    for (size_t i = 0; i < (size_t)(nt + 1) * dirs; ++i) F[i] = NAN;
    for (int j = -n; j <= n; ++j) {
        if (j == 0) continue;
        const double smu = M->rm[j];
        const double *T = M->T + (size_t)(abs(j) - 1) * nt;
        if (j > 0) {
            double run = bc ? bc[j - 1] : 0.0;
            F[FI(nt, j)] = run;
            for (int k = nt - 1; k >= 0; --k) {
                const double so = S[FI(k, j)], si = S[FI(k + 1, j)];
                if (!isfinite(so) || !isfinite(si)) return -2;
                const double dt = tau[k + 1] - tau[k], tr = T[k];
                const double a = (si - so) / dt, b = so - a * tau[k];
                run = tr * run + ((1.0 - tr) * (b + a * smu) + a * (tau[k] - tau[k + 1] * tr)) * 0.5;
                F[FI(k, j)] = run;
            }
        } else {
            double run = tb ? tb[-j - 1] : 0.0;
            F[FI(0, j)] = run;
            for (int k = 1; k <= nt; ++k) {
                const double so = S[FI(k, j)], si = S[FI(k - 1, j)];
                if (!isfinite(so) || !isfinite(si)) return -2;
                const double dt = tau[k] - tau[k - 1], tr = T[k - 1];
                const double a = (so - si) / dt, b = so - a * tau[k];
                run = tr * run + ((1.0 - tr) * (b + a * smu) + a * (tau[k] - tau[k - 1] * tr)) * 0.5;
                F[FI(k, j)] = run;
            }
        }
    }
    return 0;
}

static int sweep(const med_t *M, const double *S, const double *bc, double *F)
{
    return sweep_bc(M, S, bc, NULL, F);
}

/* Scattering integral (rt_sos_operator.c:206-424). */
static int apply_operator(const med_t *M, int m, const ang_t *A, const double *Ip, const double *Qp,
                          const double *Up, double *JI, double *JQ, double *JU)
{
    const int n = M->n, dirs = M->dirs, nt = M->nt;
    const int mol = (m <= 2);
    const double b0 = (m == 0) ? M->beta0 : 0.0;
    const double be2 = M->beta2, g2 = M->gamma2, a2 = M->alpha2;
    const int emu = g_bf16;
    for (size_t i = 0; i < (size_t)(nt + 1) * dirs; ++i) JI[i] = JQ[i] = JU[i] = NAN;
    for (int k = 0; k <= nt; ++k)
        for (int i = 1; i <= n; ++i)
            if (!isfinite(Ip[FI(k, i)]) || !isfinite(Ip[FI(k, -i)]) || !isfinite(Qp[FI(k, i)]) ||
                !isfinite(Qp[FI(k, -i)]) || !isfinite(Up[FI(k, i)]) || !isfinite(Up[FI(k, -i)]))
                return -2;
    for (int o = 1; o <= n; ++o) {
        const double po = XPL(o), ro = XRL(o), to = XTL(o);
        const double pom = XPL(-o), rom = XRL(-o), tom = XTL(-o);
        for (int k = 0; k <= nt; ++k) {
            const double x = M->xd[k], y = M->yd[k];
            double iu = 0, id = 0, qu = 0, qd = 0, uu = 0, ud = 0;
            for (int i = 1; i <= n; ++i) {
                const double w = M->gb[i];
                if (w == 0.0) continue;
                const double pi_ = XPL(i), ri = XRL(i), ti = XTL(i);
                const double rim = XRL(-i), tim = XTL(-i);
                double c[18], r[18];
                c[0] = TB(A->P, i, o); c[1] = TB(A->P, i, -o);
                c[2] = TB(A->GT, i, o); c[3] = TB(A->GT, i, -o);
                c[4] = TB(A->GT, o, i); c[5] = TB(A->GT, o, -i);
                c[6] = TB(A->GR, i, o); c[7] = TB(A->GR, i, -o);
                c[8] = TB(A->GR, o, i); c[9] = TB(A->GR, o, -i);
                c[10] = TB(A->ARR, i, o); c[11] = TB(A->ARR, i, -o);
                c[12] = TB(A->ART, i, o); c[13] = TB(A->ART, i, -o);
                c[14] = TB(A->ART, o, i); c[15] = TB(A->ART, o, -i);
                c[16] = TB(A->ATT, i, o); c[17] = TB(A->ATT, i, -o);
                if (mol) {
                    r[0] = b0 + be2 * pi_ * po; r[1] = b0 + be2 * pi_ * pom;
                    r[2] = g2 * pi_ * to; r[3] = g2 * pi_ * tom;
                    r[4] = g2 * po * ti; r[5] = g2 * po * tim;
                    r[6] = g2 * pi_ * ro; r[7] = g2 * pi_ * rom;
                    r[8] = g2 * po * ri; r[9] = g2 * po * rim;
                    r[10] = a2 * ri * ro; r[11] = a2 * ri * rom;
                    r[12] = a2 * ti * ro; r[13] = a2 * ti * rom;
                    r[14] = a2 * to * ri; r[15] = a2 * to * rim;
                    r[16] = a2 * ti * to; r[17] = a2 * ti * tom;
                } else {
                    for (int q = 0; q < 18; ++q) r[q] = 0.0;
                }
                double e[18];
                for (int q = 0; q < 18; ++q) {
                    double pc = c[q] * x + y * r[q];
                    e[q] = emu ? round_bits(pc, (emu >> 8) & 0xff) : pc;
                }
                double a0 = Ip[FI(k, i)], a1 = Ip[FI(k, -i)], a2_ = Qp[FI(k, i)], a3 = Qp[FI(k, -i)];
                double a4 = Up[FI(k, i)], a5 = Up[FI(k, -i)];
                if (emu) { const int pf = (m < ((emu >> 16) & 0xff)) ? 16 : (emu & 0xff);
                           a0 = round_bits(a0, pf); a1 = round_bits(a1, pf); a2_ = round_bits(a2_, pf);
                           a3 = round_bits(a3, pf); a4 = round_bits(a4, pf); a5 = round_bits(a5, pf); }
                iu += w * (a0 * e[0] + a2_ * e[8] - a4 * e[4]) + w * (a1 * e[1] + a3 * e[9] - a5 * e[5]);
                id += w * (a0 * e[1] + a2_ * e[9] + a4 * e[5]) + w * (a1 * e[0] + a3 * e[8] + a5 * e[4]);
                qu += w * (a0 * e[6] + a2_ * e[10] - a4 * e[12]) + w * (a1 * e[7] + a3 * e[11] + a5 * e[13]);
                qd += w * (a0 * e[7] + a2_ * e[11] - a4 * e[13]) + w * (a1 * e[6] + a3 * e[10] + a5 * e[12]);
                uu += -w * (a0 * e[2] + a2_ * e[14] - a4 * e[16]) - w * (-a1 * e[3] + a3 * e[15] - a5 * e[17]);
                ud += -w * (a0 * e[3] - a2_ * e[15] - a4 * e[17]) - w * (-a1 * e[2] - a3 * e[14] - a5 * e[16]);
            }
            JI[FI(k, o)] = iu; JI[FI(k, -o)] = id;
            JQ[FI(k, o)] = qu; JQ[FI(k, -o)] = qd;
            JU[FI(k, o)] = uu; JU[FI(k, -o)] = ud;
        }
    }
    return 0;
}

static double field_max_abs(const med_t *M, const double *F)
{
    const int n = M->n, dirs = M->dirs;
    double mx = 0.0;
    for (int k = 0; k <= M->nt; ++k)
        for (int j = -n; j <= n; ++j) {
            if (j == 0) continue;
            const double a = fabs(F[FI(k, j)]);
            if (a > mx) mx = a;
        }
    return mx;
}

static double moment_gr(const med_t *M, const ang_t *A, int m, int first, int second)
{
    if (!M->aer) return 0.0;
    const int n = M->n, dirs = M->dirs;
    double v = 0.0;
    for (int l = m; l <= M->lmax; ++l)
        v += A->plm[(size_t)l * dirs + first + n] * A->rrl[(size_t)l * dirs + second + n] * coef(M->gg_, l, M->La);
    return v;
}

static double moment_art(const med_t *M, const ang_t *A, int m, int first, int second)
{
    if (!M->aer) return 0.0;
    const int n = M->n, dirs = M->dirs;
    double v = 0.0;
    for (int l = m; l <= M->lmax; ++l)
        v += A->rtl[(size_t)l * dirs + first + n] * A->rrl[(size_t)l * dirs + second + n] * coef(M->ga_, l, M->La) +
             A->rrl[(size_t)l * dirs + first + n] * A->rtl[(size_t)l * dirs + second + n] * coef(M->gz_, l, M->La);
    return v;
}

typedef struct {
    int kind;          /* 0 none (black), 1 flat, 2 rough */
    double *R;         /* [n][n][9] for this m */
    double *Rsun;      /* [n][9] for this m (rough) */
    double W;          /* Fourier weight */
    int top;           /* in-water column: water-side reflection at the top, 0 none, 1 flat, 2 rough */
    const double *Rtop;/* [n][n][9] rough water-side kernel for this m */
} surf_t;

/* Surface-coupled order loop with first-order seeds (rt_solver.c:2923-3151,
 * rt_surface_boundary.c:65-598); the black variant (rt_solver.c:1987-2235)
 * is the same loop with no boundary operator and no seeds. */
static int sos_loop(const med_t *M, int m, const ang_t *A, const surf_t *S, const ocrt_instance_t *in,
                    double *tI, double *tQ, double *tU, int *orders, double *residual,
                    int *converged)
{
    const int n = M->n, dirs = M->dirs, nt = M->nt;
    const size_t cnt = (size_t)(nt + 1) * dirs;
    double *buf = (double *)calloc(cnt * 15 + 6 * (size_t)n, sizeof(double));
    if (!buf) return -3;
    double *sI = buf, *sQ = sI + cnt, *sU = sQ + cnt;
    double *pI = sU + cnt, *pQ = pI + cnt, *pU = pQ + cnt;
    double *o1I = pU + cnt, *o1Q = o1I + cnt, *o1U = o1Q + cnt;
    double *dI = o1U + cnt, *dQ = dI + cnt, *dU = dQ + cnt;
    double *cI = dU + cnt, *cQ = cI + cnt, *cU = cQ + cnt;
    double *bcI = cU + cnt, *bcQ = bcI + n, *bcU = bcQ + n;
    int rc = 0;
    const int bso = in->bottom_source_only != 0;
    primary_source(M, m, A, sI, sQ, sU);
    if (sweep(M, sI, NULL, pI) || sweep(M, sQ, NULL, pQ) || sweep(M, sU, NULL, pU)) { rc = -2; goto done; }
    if (bso) {   /* only the bottom source drives the field (rt_solver.c:949-954) */
        memset(pI, 0, cnt * sizeof(double)); memset(pQ, 0, cnt * sizeof(double)); memset(pU, 0, cnt * sizeof(double));
    }
    memcpy(tI, pI, cnt * sizeof(double));
    memcpy(tQ, pQ, cnt * sizeof(double));
    memcpy(tU, pU, cnt * sizeof(double));
    const int rough = (S->kind == 2) && !bso;
    if (S->kind == 1 && !bso) {
        /* reflected direct beam scattered once (rt_surface_boundary.c:210-428) */
        const double mus = M->mus, nw = in->n_water, n2 = nw * nw;
        const double tn = sqrt(n2 - 1.0 + mus * mus);
        const double rp = (n2 * mus - tn) / (n2 * mus + tn), rs = (mus - tn) / (mus + tn);
        const double RI = (rp * rp + rs * rs) / 2.0;
        const double RQ = (in->q_convention == 1) ? (rp * rp - rs * rs) / 2.0 : (rs * rs - rp * rp) / 2.0;
        const double tau = M->h[nt];
        const double sb = 0.5 * exp(-2.0 * tau / mus);
        const double bP = XPL(0), bR = XRL(0);
        const double b0 = (m == 0) ? M->beta0 : 0.0;
        const int mol = (m <= 2);
        double *fI = dI, *fQ = dQ, *fU = dU;   /* reuse as source buffers */
        memset(fI, 0, cnt * sizeof(double)); memset(fQ, 0, cnt * sizeof(double)); memset(fU, 0, cnt * sizeof(double));
        for (int L = 0; L < nt; ++L) {
            const double yt = M->yd[L], xt = M->xd[L], yb = M->yd[L + 1], xb = M->xd[L + 1];
            const double bt = sb * exp(M->h[L] / mus), bb = sb * exp(+M->h[L + 1] / mus);
            for (int d = 1; d <= n; ++d) {
                const double pu = XPL(d), pd = XPL(-d), ru = XRL(d), rd = XRL(-d), tu = XTL(d), td = XTL(-d);
                const double kPd = TB(A->P, 0, -d) * xt + (mol ? b0 + M->beta2 * pd * bP : 0.0) * yt;
                const double kRd = TB(A->GR, 0, -d) * xt + (mol ? M->gamma2 * rd * bP : 0.0) * yt;
                const double kTd = TB(A->GT, 0, -d) * xt + (mol ? M->gamma2 * td * bP : 0.0) * yt;
                const double kRTd = moment_gr(M, A, m, -d, 0) * xt + (mol ? M->gamma2 * bR * pd : 0.0) * yt;
                const double kRRd = TB(A->ARR, 0, -d) * xt + (mol ? M->alpha2 * bR * rd : 0.0) * yt;
                const double kTRd = moment_art(M, A, m, -d, 0) * xt + (mol ? M->alpha2 * td * bR : 0.0) * yt;
                const double kPu = TB(A->P, 0, d) * xb + (mol ? b0 + M->beta2 * pu * bP : 0.0) * yb;
                const double kRu = TB(A->GR, 0, d) * xb + (mol ? M->gamma2 * ru * bP : 0.0) * yb;
                const double kTu = TB(A->GT, 0, d) * xb + (mol ? M->gamma2 * tu * bP : 0.0) * yb;
                const double kRTu = moment_gr(M, A, m, d, 0) * xb + (mol ? M->gamma2 * bR * pu : 0.0) * yb;
                const double kRRu = TB(A->ARR, 0, d) * xb + (mol ? M->alpha2 * bR * ru : 0.0) * yb;
                const double kTRu = moment_art(M, A, m, d, 0) * xb + (mol ? M->alpha2 * tu * bR : 0.0) * yb;
                fI[FI(L, d)] = bt * (RI * kPd + RQ * kRTd);
                fQ[FI(L, d)] = bt * (RI * kRd + RQ * kRRd);
                fU[FI(L, d)] = bt * (RI * kTd + RQ * kTRd);
                fI[FI(L + 1, -d)] = bb * (RI * kPu + RQ * kRTu);
                fQ[FI(L + 1, -d)] = bb * (RI * kRu + RQ * kRRu);
                fU[FI(L + 1, -d)] = bb * (RI * kTu + RQ * kTRu);
            }
        }
        for (int k = 0; k <= nt; ++k) fI[FI(k, 0)] = fQ[FI(k, 0)] = fU[FI(k, 0)] = NAN;
        if (sweep(M, fI, NULL, cI) || sweep(M, fQ, NULL, cQ) || sweep(M, fU, NULL, cU)) { rc = -2; goto done; }
        for (int k = 0; k <= nt; ++k)
            for (int d = -n; d <= n; ++d) {
                if (d == 0) continue;
                tI[FI(k, d)] += cI[FI(k, d)]; tQ[FI(k, d)] += cQ[FI(k, d)]; tU[FI(k, d)] += cU[FI(k, d)];
                o1I[FI(k, d)] += cI[FI(k, d)]; o1Q[FI(k, d)] += cQ[FI(k, d)]; o1U[FI(k, d)] += cU[FI(k, d)];
            }
        memset(dI, 0, cnt * sizeof(double)); memset(dQ, 0, cnt * sizeof(double)); memset(dU, 0, cnt * sizeof(double));
    } else if (rough) {
        /* direct beam reflected by the rough surface (rt_surface_boundary.c:430-493) */
        const double tau = M->h[nt], down = exp(-tau / M->mus);
        const double FS = ((m == 0) ? PI : 0.5 * PI) * M->mus;
        for (int k = 0; k <= nt; ++k) {
            const double up = tau - M->h[k];
            for (int o = 0; o < n; ++o) {
                const double f = FS * (down * exp(-up / M->rm[o + 1]));
                const double vI = S->Rsun[o * 9 + 0] * f, vQ = S->Rsun[o * 9 + 3] * f, vU = S->Rsun[o * 9 + 6] * f;
                const size_t ix = FI(k, o + 1);
                tI[ix] += vI; tQ[ix] += vQ; tU[ix] += vU;
                o1I[ix] += vI; o1Q[ix] += vQ; o1U[ix] += vU;
                dI[ix] = vI; dQ[ix] = vQ; dU[ix] = vU;
            }
        }
    }
    if (S->kind != 0 && in->bottom_source && m <= in->bottom_m_max) {
        /* external upward source at the bottom, attenuated to every level
         * (rt_surface_boundary.c:495-527) */
        const int Mb = in->bottom_m_max + 1;
        const double *bI = in->bottom_source + (size_t)m * n;
        const double *bQ = bI + (size_t)Mb * n, *bU = bQ + (size_t)Mb * n;
        const double tau = M->h[nt];
        for (int k = 0; k <= nt; ++k) {
            const double up = tau - M->h[k];
            for (int o = 0; o < n; ++o) {
                const double a = exp(-up / M->rm[o + 1]);
                const double vI = bI[o] * a, vQ = bQ[o] * a, vU = bU[o] * a;
                const size_t ix = FI(k, o + 1);
                tI[ix] += vI; tQ[ix] += vQ; tU[ix] += vU;
                o1I[ix] += vI; o1Q[ix] += vQ; o1U[ix] += vU;
            }
        }
    }
    if (in->top_source && m <= in->top_m_max) {
        /* diffuse field entering an in-water column at the top, attenuated
         * down to every level */
        const int Mt = in->top_m_max + 1;
        const double *aI = in->top_source + (size_t)m * n;
        const double *aQ = aI + (size_t)Mt * n, *aU = aQ + (size_t)Mt * n;
        for (int k = 0; k <= nt; ++k)
            for (int o = 0; o < n; ++o) {
                const double a = exp(-M->h[k] / M->rm[o + 1]);
                const double vI = aI[o] * a, vQ = aQ[o] * a, vU = aU[o] * a;
                const size_t ix = FI(k, -(o + 1));
                tI[ix] += vI; tQ[ix] += vQ; tU[ix] += vU;
                o1I[ix] += vI; o1Q[ix] += vQ; o1U[ix] += vU;
            }
    }
    *orders = 1; *residual = 0.0; *converged = 1;
    if (in->max_orders == 1) goto finish;
    for (size_t i = 0; i < cnt; ++i) { pI[i] += o1I[i]; pQ[i] += o1Q[i]; pU[i] += o1U[i]; }
    *converged = 0;
    for (int it = 2; it <= in->max_orders; ++it) {
        rc = apply_operator(M, m, A, pI, pQ, pU, sI, sQ, sU);
        if (rc) goto done;
        const double *bI = NULL, *bQ = NULL, *bU = NULL;
        if (S->kind != 0) {
            for (int o = 1; o <= n; ++o) {
                double si = 0, sq = 0, su = 0;
                if (S->kind == 1 && M->gb[o] == 0.0) {
                    double F[9];
                    fresnel_R(M->rm[o], in->n_water, in->q_convention, F);
                    const double a = pI[FI(nt, -o)], b = pQ[FI(nt, -o)], c = pU[FI(nt, -o)];
                    si = F[0] * a + F[1] * b + F[2] * c;
                    sq = F[3] * a + F[4] * b + F[5] * c;
                    su = F[6] * a + F[7] * b + F[8] * c;
                } else {
                    for (int i = 1; i <= n; ++i) {
                        const double w = M->gb[i];
                        if (w == 0.0) continue;
                        const double a = pI[FI(nt, -i)], b = pQ[FI(nt, -i)], c = pU[FI(nt, -i)];
                        const double pw = S->W * M->rm[i] * w;
                        const double *R = S->R + ((size_t)(o - 1) * n + (i - 1)) * 9;
                        si += pw * (R[0] * a + R[1] * b + R[2] * c);
                        sq += pw * (R[3] * a + R[4] * b + R[5] * c);
                        su += pw * (R[6] * a + R[7] * b + R[8] * c);
                    }
                }
                bcI[o - 1] = si; bcQ[o - 1] = sq; bcU[o - 1] = su;
            }
            bI = bcI; bQ = bcQ; bU = bcU;
        }
        const double *tbI = NULL, *tbQ = NULL, *tbU = NULL;
        if (S->top) {
            /* water-side reflection of the previous order's upward field at the
             * top (rt_solver.c:2594-2606 flat, 2797-2822 rough) */
            double *tb = bcI + 3 * (size_t)n;
            for (int j = 1; j <= n; ++j) {
                double aI = 0, aQ = 0, aU = 0;
                if (S->top == 1) {
                    double F[9];
                    fresnel_R_general(M->rm[j], in->n_water, 1.0, in->q_convention, F);
                    const double a = pI[FI(0, j)], b = pQ[FI(0, j)], c = pU[FI(0, j)];
                    aI = F[0] * a + F[1] * b + F[2] * c;
                    aQ = F[3] * a + F[4] * b + F[5] * c;
                    aU = F[6] * a + F[7] * b + F[8] * c;
                } else {
                    const double C = (m == 0) ? 2.0 * PI : PI;
                    for (int k = 1; k <= n; ++k) {
                        const double wk = C * M->rm[k] * M->gb[k];
                        const double a = pI[FI(0, k)], b = pQ[FI(0, k)], c = pU[FI(0, k)];
                        const double *K = S->Rtop + ((size_t)(j - 1) * n + (k - 1)) * 9;
                        aI += wk * (K[0] * a + K[1] * b + K[2] * c);
                        aQ += wk * (K[3] * a + K[4] * b + K[5] * c);
                        aU += wk * (K[6] * a + K[7] * b + K[8] * c);
                    }
                }
                tb[j - 1] = aI; tb[n + j - 1] = aQ; tb[2 * n + j - 1] = aU;
            }
            tbI = tb; tbQ = tb + n; tbU = tb + 2 * n;
        }
        if (sweep_bc(M, sI, bI, tbI, cI) || sweep_bc(M, sQ, bQ, tbQ, cQ) || sweep_bc(M, sU, bU, tbU, cU)) {
            rc = -2;
            goto done;
        }
        for (int k = 0; k <= nt; ++k)
            for (int d = -n; d <= n; ++d) {
                if (d == 0) continue;
                tI[FI(k, d)] += cI[FI(k, d)]; tQ[FI(k, d)] += cQ[FI(k, d)]; tU[FI(k, d)] += cU[FI(k, d)];
            }
        *orders = it;
        double mi = field_max_abs(M, cI), t;
        t = field_max_abs(M, cQ); if (t > mi) mi = t;
        t = field_max_abs(M, cU); if (t > mi) mi = t;
        double mt = field_max_abs(M, tI);
        t = field_max_abs(M, tQ); if (t > mt) mt = t;
        t = field_max_abs(M, tU); if (t > mt) mt = t;
        *residual = (mt > 0.0) ? mi / mt : 0.0;
        if (*residual < in->tolerance) { *converged = 1; break; }
        memcpy(pI, cI, cnt * sizeof(double));
        memcpy(pQ, cQ, cnt * sizeof(double));
        memcpy(pU, cU, cnt * sizeof(double));
    }
finish:
    if (rough)
        for (int k = 0; k <= nt; ++k)
            for (int o = 1; o <= n; ++o) {
                tI[FI(k, o)] -= dI[FI(k, o)]; tQ[FI(k, o)] -= dQ[FI(k, o)]; tU[FI(k, o)] -= dU[FI(k, o)];
            }
done:
    free(buf);
    return rc;
}

static double pchip_int(double h0, double h1, double s0, double s1)
{
    if (s0 == 0.0 || s1 == 0.0 || (s0 > 0.0) != (s1 > 0.0)) return 0.0;
    const double w1 = 2.0 * h1 + h0, w2 = h1 + 2.0 * h0;
    return (w1 + w2) / (w1 / s0 + w2 / s1);
}

static double pchip_edge(double h0, double h1, double s0, double s1)
{
    double d = ((2.0 * h0 + h1) * s0 - h0 * s1) / (h0 + h1);
    if (d == 0.0 || s0 == 0.0 || (d > 0.0) != (s0 > 0.0)) {
        if (s0 == 0.0 || (d > 0.0) != (s0 > 0.0)) d = 0.0;
    } else if (((s0 > 0.0) != (s1 > 0.0)) && fabs(d) > 3.0 * fabs(s0)) {
        d = 3.0 * s0;
    }
    return d;
}

/* Linear view sampling over positive nodes (rt_solver.c:386-408). */
static double interp_lin(const double *y, const double *x, int n, double mv)
{
    int jh = 1;
    while (jh < n && x[jh - 1] < mv) jh++;
    int jl = jh - 1;
    if (jl < 1) { jl = 1; jh = 2; }
    if (jh > n) { jh = n; jl = n - 1; }
    const double t = (mv - x[jl - 1]) / (x[jh - 1] - x[jl - 1]);
    return (1.0 - t) * y[jl - 1] + t * y[jh - 1];
}

/* PCHIP view sampling (rt_solver.c:409-495); y[j-1], x[j-1] for j = 1..n. */
static double interp_pchip(const double *y, const double *x, int n, double mv)
{
    int jh = 1;
    while (jh < n && x[jh - 1] < mv) jh++;
    int jl = jh - 1;
    if (jl < 1 || jh > n || n < 4) return interp_lin(y, x, n, mv);
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
    return (2.0 * t3 - 3.0 * t2 + 1.0) * y1 + (t3 - 2.0 * t2 + t) * h * d1 +
           (-2.0 * t3 + 3.0 * t2) * y2 + (t3 - t2) * h * d2;
}

typedef struct {
    med_t M;
    double *modes;      /* [m_max+1][3] TOA at view node */
    double *boa_dn;     /* [m_max+1][n] I downward at BOA */
    double *toa_up;     /* [m_max+1][3][n] TOA upward at every node (LUT sampling) */
    double *boa_dn3;    /* [m_max+1][3][n] BOA downward I, Q, U at every node */
    int *orders;
    double *resid;
    int max_orders, all_conv;
} solve_t;

static void solve_free(solve_t *s)
{
    free(s->M.rm_s); free(s->M.gb_s); free(s->M.h); free(s->M.ch); free(s->M.xd); free(s->M.yd);
    free(s->M.z); free(s->M.T); free(s->modes); free(s->boa_dn); free(s->toa_up); free(s->boa_dn3);
    free(s->orders); free(s->resid);
    memset(s, 0, sizeof(*s));
}

/* Driver body shared by the single-view and LUT entry points
 * (rt_solver.c:618-1062). */
static int solve_core(const ocrt_instance_t *in, solve_t *s)
{
    memset(s, 0, sizeof(*s));
    if (!isfinite(in->wavelength_nm) || in->wavelength_nm < 330.0 - 1e-9 || in->wavelength_nm > 1100.0 + 1e-9) return -1;
    if (in->sza_deg < 0.0 || in->sza_deg >= 90.0) return -1;
    if (in->vza_deg < 0.0 || in->vza_deg >= 90.0) return -1;
    if (in->m_max < 0 || in->n_mu < 2 || in->n_layers < 2) return -1;
    if (in->max_orders < 1 || !(in->tolerance > 0.0)) return -1;
    const int aer = has_aerosol(in);
    if (aer && (in->tau_a < 0.0 || in->ssa_a < 0.0 || in->ssa_a > 1.0 || in->L_max < 2)) return -1;
    if (in->layers && in->ipss) return -1;
    med_t *M = &s->M;
    M->mus = cos(in->sza_deg * PI / 180.0);
    M->muv = cos(in->vza_deg * PI / 180.0);
    M->aer = aer;
    M->La = aer ? in->L_max : 0;
    M->m_max = in->m_max;
    if (aer) { const int q = in->L_max + 1; M->gb_ = in->greek; M->gg_ = in->greek + q; M->ga_ = in->greek + 2 * q; M->gz_ = in->greek + 3 * q; }
    if (in->n_ang_phase > 0 && (in->n_ang_phase < 2 || !in->theta_phase || !in->p11_phase)) return -1;
    if (in->bottom_source && in->bottom_m_max < 0) return -7;
    if (in->water_tau > 0.0) {
        if (in->layers || in->ipss || in->bottom_source || in->surface != OCRT_SURFACE_BLACK) return -1;
        if (in->water_omega_w < 0.0 || in->water_omega_p < 0.0 || in->water_omega_w + in->water_omega_p > 1.0) return -1;
        if (in->top_reflect < 0 || in->top_reflect > 2 || (in->top_reflect && !(in->n_water > 1.0))) return -1;
        if (in->top_reflect == 2 && !(in->wind_speed > 0.0)) return -1;
    } else if (in->top_reflect || in->top_source) {
        return -1;
    }
    if (aer && in->n_ang_phase >= 2) { M->n_ang = in->n_ang_phase; M->th_ang = in->theta_phase; M->p11_ang = in->p11_phase; }
    double mu[700], w[700];
    if (in->n_mu > 639) return -1;
    int vi;
    const int n = oracle_ring(in->n_mu, M->muv, mu, w, &vi);
    if (n < 2) return -1;
    M->n = n; M->nt = in->n_layers; M->L = M->nt + 1; M->dirs = 2 * n + 1; M->vidx = vi + 1;
    if (in->bottom_source && in->bottom_n_mu != n) return -7;   /* rt_solver.c:698-706 */
    if (in->top_source && (in->top_n_mu != n || in->top_m_max < 0)) return -7;
    M->lmax = in->m_max > 2 ? in->m_max : 2;
    if (aer && in->L_max > M->lmax) M->lmax = in->L_max;
    /* fixed in the reference's atmosphere driver (rt_solver.c:646); the water
     * body's own (0.039 for water, rt_water_rt.c:2122) in an in-water column */
    const double d = (in->water_tau > 0.0) ? in->depol : 0.0279;
    const double an = 2.0 * (1.0 - d) / (2.0 + d);
    M->beta0 = 1.0; M->beta2 = 0.5 * an; M->gamma2 = -an * sqrt(1.5); M->alpha2 = 3.0 * an;
    const int nt = M->nt, dirs = M->dirs;
    M->rm_s = malloc(sizeof(double) * dirs); M->gb_s = malloc(sizeof(double) * dirs);
    M->h = malloc(sizeof(double) * (nt + 1)); M->ch = malloc(sizeof(double) * (nt + 1));
    M->xd = malloc(sizeof(double) * (nt + 1)); M->yd = malloc(sizeof(double) * (nt + 1));
    M->z = malloc(sizeof(double) * (nt + 1)); M->T = malloc(sizeof(double) * (size_t)n * nt);
    const int mm = in->m_max;
    s->modes = calloc((size_t)3 * (mm + 1), sizeof(double));
    s->boa_dn = calloc((size_t)(mm + 1) * n, sizeof(double));
    s->toa_up = calloc((size_t)(mm + 1) * 3 * n, sizeof(double));
    s->boa_dn3 = calloc((size_t)(mm + 1) * 3 * n, sizeof(double));
    s->orders = calloc((size_t)mm + 1, sizeof(int));
    s->resid = calloc((size_t)mm + 1, sizeof(double));
    if (!M->rm_s || !M->gb_s || !M->h || !M->ch || !M->xd || !M->yd || !M->z || !M->T || !s->modes ||
        !s->boa_dn || !s->toa_up || !s->boa_dn3 || !s->orders || !s->resid) return -3;
    M->rm = M->rm_s + n; M->gb = M->gb_s + n;
    M->rm[0] = -M->mus; M->gb[0] = 0.0;
    for (int j = 1; j <= n; ++j) { M->rm[j] = mu[j - 1]; M->rm[-j] = -mu[j - 1]; M->gb[j] = M->gb[-j] = w[j - 1]; }
    if (oracle_medium(in, M->h, M->ch, M->xd, M->yd, M->z) != 0) return -1;
    M->tau_tot = M->h[nt];
    for (int o = 1; o <= n; ++o)
        for (int k = 0; k < nt; ++k) M->T[(size_t)(o - 1) * nt + k] = exp(-(M->h[k + 1] - M->h[k]) / fabs(M->rm[o]));

    const size_t cnt = (size_t)(nt + 1) * dirs;
    const size_t bcount = (size_t)(M->lmax + 1) * dirs, kcount = (size_t)(n + 1) * dirs;
    ang_t A;
    double *abuf = calloc(3 * bcount + 6 * kcount + 3 * cnt, sizeof(double));
    if (!abuf) return -3;
    A.plm = abuf; A.rrl = A.plm + bcount; A.rtl = A.rrl + bcount;
    A.P = A.rtl + bcount; A.GR = A.P + kcount; A.GT = A.GR + kcount;
    A.ARR = A.GT + kcount; A.ART = A.ARR + kcount; A.ATT = A.ART + kcount;
    double *tI = A.ATT + kcount, *tQ = tI + cnt, *tU = tQ + cnt;

    surf_t S;
    memset(&S, 0, sizeof(S));
    double *Rall = NULL, *Rsun = NULL;
    const int surf_aware = (in->surface == OCRT_SURFACE_BLACK_FRESNEL_OCEAN || in->surface == OCRT_SURFACE_FLAT);
    if (surf_aware && !(in->n_water > 0.0)) { free(abuf); return -1; }
    if (surf_aware) {
        const int flat = (in->surface == OCRT_SURFACE_FLAT) ||
                         (in->surface == OCRT_SURFACE_BLACK_FRESNEL_OCEAN && in->wind_speed <= 0.0);
        S.kind = flat ? 1 : 2;
        S.R = calloc((size_t)n * n * 9, sizeof(double));
        if (!S.R) { free(abuf); return -3; }
        if (!flat) {
            Rall = malloc(sizeof(double) * (size_t)(mm + 1) * n * n * 9);
            Rsun = malloc(sizeof(double) * (size_t)(mm + 1) * n * 9);
            if (!Rall || !Rsun) { free(abuf); free(S.R); free(Rall); free(Rsun); return -3; }
            double smu = M->mus;
            oracle_coxmunk_fourier(&M->rm[1], n, &M->rm[1], n, mm, 1024, in->wind_speed, in->sigma_type,
                                   in->n_water, in->q_convention, Rall);
            oracle_coxmunk_fourier(&M->rm[1], n, &smu, 1, mm, 1024, in->wind_speed, in->sigma_type,
                                   in->n_water, in->q_convention, Rsun);
        } else {
            for (int o = 0; o < n; ++o) {
                const double wo = M->gb[o + 1];
                if (wo <= 0.0) continue;
                double F[9];
                fresnel_R(M->rm[o + 1], in->n_water, in->q_convention, F);
                const double dd = 1.0 / (2.0 * PI * M->rm[o + 1] * wo);
                for (int e = 0; e < 9; ++e) S.R[((size_t)o * n + o) * 9 + e] = dd * F[e];
            }
            S.W = 2.0 * PI;
        }
    }
    double *Rtop = NULL;
    S.top = in->water_tau > 0.0 ? in->top_reflect : 0;
    if (S.top == 2) {
        /* water-side Cox-Munk kernels of every mode, 128 azimuths (rt_water_rt.c:5605) */
        Rtop = malloc(sizeof(double) * (size_t)(mm + 1) * n * n * 9);
        if (!Rtop) { free(abuf); free(S.R); free(Rall); free(Rsun); return -3; }
        oracle_interface_fourier(1, &M->rm[1], n, &M->rm[1], n, mm, 128, in->wind_speed, in->sigma_type,
                                 in->n_water, in->q_convention, Rtop);
    }
    int rc = 0;
    s->max_orders = 0; s->all_conv = 1;
    for (int m = 0; m <= mm; ++m) {
        if (Rtop) S.Rtop = Rtop + (size_t)m * n * n * 9;
        oracle_basis(M->lmax, m, dirs, M->rm_s, A.plm, A.rrl, A.rtl);
        build_tables(M, m, &A);
        if (M->n_ang > 1) value_table(M, m, &A);
        if (S.kind == 2) {
            memcpy(S.R, Rall + (size_t)m * n * n * 9, sizeof(double) * (size_t)n * n * 9);
            S.Rsun = Rsun + (size_t)m * n * 9;
            S.W = (m == 0) ? 2.0 * PI : PI;
        }
        int ords, conv;
        double res;
        rc = sos_loop(M, m, &A, &S, in, tI, tQ, tU, &ords, &res, &conv);
        if (rc) { rc = (rc == -3) ? -3 : -2; break; }
        s->orders[m] = ords; s->resid[m] = res;
        if (ords > s->max_orders) s->max_orders = ords;
        if (!conv) s->all_conv = 0;
        s->modes[3 * m + 0] = tI[FI(0, M->vidx)];
        s->modes[3 * m + 1] = tQ[FI(0, M->vidx)];
        s->modes[3 * m + 2] = tU[FI(0, M->vidx)];
        for (int jp = 1; jp <= n; ++jp) {
            s->boa_dn[(size_t)m * n + jp - 1] = tI[FI(nt, -jp)];
            s->toa_up[((size_t)m * 3 + 0) * n + jp - 1] = tI[FI(0, jp)];
            s->toa_up[((size_t)m * 3 + 1) * n + jp - 1] = tQ[FI(0, jp)];
            s->toa_up[((size_t)m * 3 + 2) * n + jp - 1] = tU[FI(0, jp)];
            s->boa_dn3[((size_t)m * 3 + 0) * n + jp - 1] = tI[FI(nt, -jp)];
            s->boa_dn3[((size_t)m * 3 + 1) * n + jp - 1] = tQ[FI(nt, -jp)];
            s->boa_dn3[((size_t)m * 3 + 2) * n + jp - 1] = tU[FI(nt, -jp)];
        }
    }
    free(abuf); free(S.R); free(Rall); free(Rsun); free(Rtop);
    return rc;
}

static int ipss_for(const ocrt_instance_t *in, const med_t *M, ipss_t *G)
{
    const int rc = ipss_init(G, M->nt, M->h, M->xd, M->yd, M->z, M->mus, 0.0279, M->aer, M->La, M->gb_);
    if (rc == 0 && G->w_src && M->n_ang > 1) {   /* the table has priority (rt_ipss.c:81-99, 134) */
        G->n_ang = M->n_ang; G->th_ang = M->th_ang; G->p11_ang = M->p11_ang;
    }
    return rc;
}

int oracle_solve(const ocrt_instance_t *in, ocrt_result_t *out, double *toa_modes, int *orders,
                 double *residuals)
{
    memset(out, 0, sizeof(*out));
    solve_t s;
    int rc = solve_core(in, &s);
    if (rc) { out->status = rc; solve_free(&s); return rc; }
    const med_t *M = &s.M;
    const int mm = in->m_max, n = M->n;
    const double mus = M->mus, muv = M->muv, raa = in->raa_deg * PI / 180.0;
    double *pm = malloc(sizeof(double) * 3 * (mm + 1));
    for (int m = 0; m <= mm; ++m) for (int c = 0; c < 3; ++c) pm[c * (mm + 1) + m] = s.modes[3 * m + c];
    double I = oracle_reconstruct_cos(pm, mm, raa);
    double Q = oracle_reconstruct_cos(pm + (mm + 1), mm, raa);
    double U = oracle_reconstruct_sin(pm + 2 * (mm + 1), mm, raa);
    out->tau_R_total = in->layers ? layers_sum(in, 0) : (in->tau_R > 0.0 ? in->tau_R : 0.0);
    out->I_TOA = I; out->Q_TOA = Q; out->U_TOA = U;
    out->rho_I = I / mus; out->rho_Q = Q / mus; out->rho_U = U / mus;
    double kappa = 1.0, tdir_ipss = -1.0;
    ipss_t G;
    int ip = 0;
    if (in->ipss) {
        if (ipss_for(in, M, &G) != 0) { free(pm); solve_free(&s); out->status = -1; return -1; }
        ip = 1;
        tdir_ipss = ipss_tau_dir(&G);
        kappa = ipss_kappa(&G, in->vza_deg, in->raa_deg);
        out->rho_I *= kappa; out->rho_Q *= kappa; out->rho_U *= kappa;
        out->I_TOA *= kappa; out->Q_TOA *= kappa; out->U_TOA *= kappa;
    }
    out->kappa = kappa;
    const double tau = M->h[M->nt];
    const double phv = (in->raa_deg - 180.0) * PI / 180.0;
    if (in->surface == OCRT_SURFACE_BLACK_FRESNEL_OCEAN && in->decouple_sunglint == 0) {
        double rg[3];
        sunglint_rho(in->vza_deg > 0 ? muv : 1.0, phv, mus, in->wind_speed, in->sigma_type, in->n_water,
                     tau, in->q_convention, rg);
        if (ip && tdir_ipss >= 0.0) { const double c = exp(-(tdir_ipss - tau / mus)); rg[0] *= c; rg[1] *= c; rg[2] *= c; }
        out->rho_I += rg[0]; out->rho_Q += rg[1]; out->rho_U += -rg[2];
    }
    out->n_orders_used = s.max_orders;
    out->converged = s.all_conv;
    out->T_dir_dn = (ip && tdir_ipss >= 0.0) ? exp(-tdir_ipss) : exp(-tau / mus);
    double F = 0.0;
    for (int jp = 1; jp <= n; ++jp) F += M->rm[jp] * M->gb[jp] * s.boa_dn[jp - 1];
    F *= 2.0 * PI;
    out->T_diff_dn_hemi = F / (PI * mus);
    out->T_total_dn_hemi = out->T_dir_dn + out->T_diff_dn_hemi;
    double *idv = malloc(sizeof(double) * (mm + 1));
    for (int m = 0; m <= mm; ++m) idv[m] = interp_lin(s.boa_dn + (size_t)m * n, &M->rm[1], n, muv);
    out->T_diff_dn_dir = (2.0 * PI * muv * oracle_reconstruct_cos(idv, mm, raa)) / (PI * mus);
    out->T_total_up_dir = out->rho_I * mus;
    out->T_dir_up_view = exp(-tau / (muv > 0.0 ? muv : 1.0));
    if (in->surface == OCRT_SURFACE_BLACK_FRESNEL_OCEAN && in->wind_speed > 0.0) {
        double rg[3];
        sunglint_rho(in->vza_deg > 0.0 ? muv : 1.0, phv, mus, in->wind_speed, in->sigma_type, in->n_water,
                     tau, in->q_convention, rg);
        if (ip && tdir_ipss >= 0.0) { const double c = exp(-(tdir_ipss - tau / mus)); rg[0] *= c; rg[1] *= c; rg[2] *= c; }
        out->T_sg_up_dir = rg[0] * mus;
        out->rho_glint_direct_I = rg[0]; out->rho_glint_direct_Q = rg[1]; out->rho_glint_direct_U = rg[2];
    }
    out->n_ring = n;
    out->status = 0;
    if (toa_modes) memcpy(toa_modes, s.modes, sizeof(double) * 3 * (mm + 1));
    if (orders) memcpy(orders, s.orders, sizeof(int) * (mm + 1));
    if (residuals) memcpy(residuals, s.resid, sizeof(double) * (mm + 1));
    free(idv); free(pm);
    if (ip) ipss_free(&G);
    solve_free(&s);
    return 0;
}

int oracle_solve_lut(const ocrt_instance_t *in, int n_vza, const double *vza, int n_raa, const double *raa,
                     double *rho_I, double *rho_Q, double *rho_U, double *T_diff_dn_dir,
                     double *T_sg_up_dir, double *T_total_up_dir, double *T_dir_dn, double *T_diff_dn_hemi)
{
    for (int iv = 0; iv < n_vza; ++iv) if (vza[iv] < 0.0 || vza[iv] >= 90.0) return -1;
    solve_t s;
    int rc = solve_core(in, &s);
    if (rc) { solve_free(&s); return rc; }
    const med_t *M = &s.M;
    const int mm = in->m_max, n = M->n;
    const double mus = M->mus, tau = M->h[M->nt];
    ipss_t G;
    int ip = 0;
    double tdir_ipss = -1.0;
    if (in->ipss) {
        if (ipss_for(in, M, &G) != 0) { solve_free(&s); return -1; }
        ip = 1;
        tdir_ipss = ipss_tau_dir(&G);
    }
    const double cg = (ip && tdir_ipss >= 0.0) ? exp(-(tdir_ipss - tau / mus)) : 1.0;
    const double tdd = (ip && tdir_ipss >= 0.0) ? exp(-tdir_ipss) : exp(-tau / mus);
    double F = 0.0;
    for (int jp = 1; jp <= n; ++jp) F += M->rm[jp] * M->gb[jp] * s.boa_dn[jp - 1];
    F *= 2.0 * PI;
    if (T_dir_dn) *T_dir_dn = tdd;
    if (T_diff_dn_hemi) *T_diff_dn_hemi = F / (PI * mus);
    double *pI = malloc(sizeof(double) * (mm + 1) * 4);
    double *pQ = pI + (mm + 1), *pU = pQ + (mm + 1), *pd = pU + (mm + 1);
    for (int iv = 0; iv < n_vza; ++iv) {
        const double mv = cos(vza[iv] * PI / 180.0);
        for (int m = 0; m <= mm; ++m) {
            pI[m] = interp_pchip(s.toa_up + ((size_t)m * 3 + 0) * n, &M->rm[1], n, mv);
            pQ[m] = interp_pchip(s.toa_up + ((size_t)m * 3 + 1) * n, &M->rm[1], n, mv);
            pU[m] = interp_pchip(s.toa_up + ((size_t)m * 3 + 2) * n, &M->rm[1], n, mv);
            pd[m] = interp_lin(s.boa_dn + (size_t)m * n, &M->rm[1], n, mv);
        }
        for (int ir = 0; ir < n_raa; ++ir) {
            const double ph = raa[ir] * PI / 180.0;
            double rI = oracle_reconstruct_cos(pI, mm, ph) / mus;
            double rQ = oracle_reconstruct_cos(pQ, mm, ph) / mus;
            double rU = oracle_reconstruct_sin(pU, mm, ph) / mus;
            if (ip) {
                const double kg = ipss_kappa(&G, vza[iv], raa[ir]);
                rI *= kg; rQ *= kg; rU *= kg;
            }
            double sgI = 0.0;
            if (in->surface == OCRT_SURFACE_BLACK_FRESNEL_OCEAN && in->wind_speed > 0.0) {
                double rg[3];
                sunglint_rho(vza[iv] > 0.0 ? mv : 1.0, (raa[ir] - 180.0) * PI / 180.0, mus, in->wind_speed,
                             in->sigma_type, in->n_water, tau, in->q_convention, rg);
                rg[0] *= cg; rg[1] *= cg; rg[2] *= cg;
                if (in->decouple_sunglint == 0) { rI += rg[0]; rQ += rg[1]; rU += -rg[2]; }
                sgI = rg[0];
            }
            const int ix = iv * n_raa + ir;
            rho_I[ix] = rI; rho_Q[ix] = rQ; rho_U[ix] = rU;
            if (T_diff_dn_dir) T_diff_dn_dir[ix] = (2.0 * PI * mv * oracle_reconstruct_cos(pd, mm, ph)) / (PI * mus);
            if (T_sg_up_dir) T_sg_up_dir[ix] = sgI * mus;
            if (T_total_up_dir) T_total_up_dir[ix] = rI * mus;
        }
    }
    free(pI);
    if (ip) ipss_free(&G);
    solve_free(&s);
    return 0;
}

/* Per-node Fourier fields of one solve: top_up [m_max+1][3][n] upward I, Q, U
 * at the top (k = 0; rough direct seed not removed), bot_dn [m_max+1][3][n]
 * downward at the bottom, orders [m_max+1]; n = ring size (returned). */
int oracle_solve_nodes(const ocrt_instance_t *in, double *top_up, double *bot_dn, int *orders)
{
    solve_t s;
    int rc = solve_core(in, &s);
    if (rc == 0) {
        const size_t cnt = (size_t)(in->m_max + 1) * 3 * s.M.n;
        if (top_up) memcpy(top_up, s.toa_up, cnt * sizeof(double));
        if (bot_dn) memcpy(bot_dn, s.boa_dn3, cnt * sizeof(double));
        if (orders) memcpy(orders, s.orders, (size_t)(in->m_max + 1) * sizeof(int));
        rc = s.M.n;
    }
    solve_free(&s);
    return rc;
}

int oracle_solve_batch(const ocrt_instance_t *in, int n, ocrt_result_t *out, int n_threads)
{
    int failed = 0;
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 8) reduction(+ : failed)
    for (int i = 0; i < n; ++i)
        if (oracle_solve(&in[i], &out[i], NULL, NULL, NULL) != 0) ++failed;
    return failed;
}

double oracle_ipss_kappa(const ocrt_instance_t *in, double vza_deg, double raa_deg)
{
    solve_t s;
    memset(&s, 0, sizeof(s));
    ocrt_instance_t tmp = *in;
    tmp.m_max = 0; tmp.max_orders = 1;
    if (solve_core(&tmp, &s) != 0) { solve_free(&s); return NAN; }
    ipss_t G;
    if (ipss_for(in, &s.M, &G) != 0) { solve_free(&s); return NAN; }
    const double k = ipss_kappa(&G, vza_deg, raa_deg);
    ipss_free(&G);
    solve_free(&s);
    return k;
}

double oracle_ipss_tau_dir_dn(const ocrt_instance_t *in)
{
    solve_t s;
    ocrt_instance_t tmp = *in;
    tmp.m_max = 0; tmp.max_orders = 1;
    if (solve_core(&tmp, &s) != 0) { solve_free(&s); return NAN; }
    ipss_t G;
    if (ipss_for(in, &s.M, &G) != 0) { solve_free(&s); return NAN; }
    const double t = ipss_tau_dir(&G);
    ipss_free(&G);
    solve_free(&s);
    return t;
}

/* Homogeneous 0-100 km shell atmosphere of total vertical depth tau_t in nlay
 * equal layers, used by the reference's IPSS geometry gate
 * (tests/test_ipss_gates.c:43-58, 75-130). */
static int ipss_homog(ipss_t *G, double tau_t, int nlay, double sza)
{
    memset(G, 0, sizeof(*G));
    G->nlay = nlay; G->Re = 6371.0; G->H = 100.0; G->mu_sun = cos(sza * PI / 180.0);
    G->L_aer = -1;
    G->h_edge = malloc(sizeof(double) * (nlay + 1));
    G->beta = malloc(sizeof(double) * nlay);
    G->omega = malloc(sizeof(double) * nlay);
    G->tau_above = malloc(sizeof(double) * (nlay + 1));
    G->ts = malloc(sizeof(double) * (2 * nlay + 8));
    G->los_t = malloc(sizeof(double) * (2 * nlay + 8));
    G->los_tau = malloc(sizeof(double) * (2 * nlay + 8));
    G->los_lay = malloc(sizeof(int) * (2 * nlay + 8));
    for (int k = 0; k <= nlay; ++k) G->h_edge[k] = 100.0 * (double)k / nlay;
    for (int j = 0; j < nlay; ++j) { G->beta[j] = tau_t / 100.0; G->omega[j] = 1.0; }
    G->tau_above[nlay] = 0.0;
    for (int j = nlay - 1; j >= 0; --j)
        G->tau_above[j] = G->tau_above[j + 1] + G->beta[j] * (G->h_edge[j + 1] - G->h_edge[j]);
    ipss_gl2(G->glx, G->glw);
    return 0;
}

double oracle_ipss_homog_tau_view(double tau_t, int nlay, double vza_deg)
{
    ipss_t G;
    ipss_homog(&G, tau_t, nlay, 0.0);
    const double tv = vza_deg * PI / 180.0;
    const double r0[3] = {0.0, 0.0, G.Re + G.H};
    const double u[3] = {-sin(tv), 0.0, -cos(tv)};
    int hit = 0;
    const double t = ipss_ray_tau(&G, r0, u, &hit);
    ipss_free(&G);
    return t;
}

double oracle_ipss_homog_tau_sun(double tau_t, int nlay, double sza_deg, const double *rB)
{
    ipss_t G;
    ipss_homog(&G, tau_t, nlay, sza_deg);
    const double t = ipss_tau_sun(&G, rB);
    ipss_free(&G);
    return isfinite(t) ? t : -1.0;
}

/* ======================================================================== */
/* Reverse-mode derivative of the solve (test infrastructure; the reference
 * has no adjoint: it is pinned by central finite differences of the forward
 * solve above).  Every order of every mode is stored, then the order
 * recursion is run backward:
 *   lambda_N = e,  lambda_{o-1} = e + M^T Sbar_o + B^T R^T bbar_o,
 * where e seeds the TOA view-node entries with the reconstruction weights,
 * (Sbar_o, bbar_o) is the adjoint sweep of lambda_o, M the scattering
 * integral and R the surface boundary operator.  Parameters: the per-level
 * medium h[k], xdel[k], ydel[k] and the wind speed. */

/* Per-step coefficients of the linear-source recurrence written as
 * r_k = T r_prev + al S_k + be S_prev, al = (1 - q) / 2, be = (q - T) / 2,
 * q = (1 - T) |mu| / dtau, and their derivatives in dtau. */
static void step_coefs(double dt, double amu, double *T, double *al, double *be, double *dT, double *dal,
                       double *dbe)
{
    const double x = dt / amu, t = exp(-x), om = -expm1(-x), q = om / x;
    const double dq = (t - q) / dt;
    *T = t; *al = 0.5 * (1.0 - q); *be = 0.5 * (q - t);
    *dT = -t / amu; *dal = -0.5 * dq; *dbe = 0.5 * (dq + t / amu);
}

/* Adjoint of F = sweep(S, bc): lam is the adjoint of F, Sbar (zeroed by the
 * caller) receives the adjoint of S, bcbar[n] that of the bottom boundary of
 * the upward columns (may be NULL), gh the h-gradient of the layer steps. */
static void sweep_adj(const med_t *M, const double *lam, const double *F, const double *S, double *Sbar,
                      double *bcbar, double *gh)
{
    const int n = M->n, dirs = M->dirs, nt = M->nt;
    const double *h = M->h;
    for (int j = -n; j <= n; ++j) {
        if (j == 0) continue;
        const double amu = fabs(M->rm[j]);
        if (j > 0) {
            double rho = lam[FI(0, j)];
            for (int k = 0; k < nt; ++k) {
                double T, al, be, dT, dal, dbe;
                step_coefs(h[k + 1] - h[k], amu, &T, &al, &be, &dT, &dal, &dbe);
                Sbar[FI(k, j)] += al * rho;
                Sbar[FI(k + 1, j)] += be * rho;
                const double dd = rho * (dT * F[FI(k + 1, j)] + dal * S[FI(k, j)] + dbe * S[FI(k + 1, j)]);
                gh[k + 1] += dd; gh[k] -= dd;
                rho = lam[FI(k + 1, j)] + T * rho;
            }
            if (bcbar) bcbar[j - 1] = rho;
        } else {
            double rho = lam[FI(nt, j)];
            for (int k = nt; k >= 1; --k) {
                double T, al, be, dT, dal, dbe;
                step_coefs(h[k] - h[k - 1], amu, &T, &al, &be, &dT, &dal, &dbe);
                Sbar[FI(k, j)] += al * rho;
                Sbar[FI(k - 1, j)] += be * rho;
                const double dd = rho * (dT * F[FI(k - 1, j)] + dal * S[FI(k, j)] + dbe * S[FI(k - 1, j)]);
                gh[k] += dd; gh[k - 1] -= dd;
                rho = lam[FI(k - 1, j)] + T * rho;
            }
        }
    }
}

/* Adjoint of J = M(x, y) F (apply_operator): adds M^T Jbar to Fbar and the
 * x / y gradients sum_q c_q E_q, sum_q r_q E_q, E_q = d(Jbar . J) / d e_q. */
static void apply_operator_adj(const med_t *M, int m, const ang_t *A, const double *JbI, const double *JbQ,
                               const double *JbU, const double *Ip, const double *Qp, const double *Up,
                               double *FbI, double *FbQ, double *FbU, double *gx, double *gy)
{
    const int n = M->n, dirs = M->dirs, nt = M->nt;
    const int mol = (m <= 2);
    const double b0 = (m == 0) ? M->beta0 : 0.0;
    const double be2 = M->beta2, g2 = M->gamma2, a2 = M->alpha2;
    for (int o = 1; o <= n; ++o) {
        const double po = XPL(o), ro = XRL(o), to = XTL(o);
        const double pom = XPL(-o), rom = XRL(-o), tom = XTL(-o);
        for (int k = 0; k <= nt; ++k) {
            const double x = M->xd[k], y = M->yd[k];
            const double biu = JbI[FI(k, o)], bid = JbI[FI(k, -o)], bqu = JbQ[FI(k, o)], bqd = JbQ[FI(k, -o)];
            const double buu = JbU[FI(k, o)], bud = JbU[FI(k, -o)];
            for (int i = 1; i <= n; ++i) {
                const double w = M->gb[i];
                if (w == 0.0) continue;
                const double pi_ = XPL(i), ri = XRL(i), ti = XTL(i), rim = XRL(-i), tim = XTL(-i);
                double c[18], r[18], e[18], E[18];
                c[0] = TB(A->P, i, o); c[1] = TB(A->P, i, -o);
                c[2] = TB(A->GT, i, o); c[3] = TB(A->GT, i, -o);
                c[4] = TB(A->GT, o, i); c[5] = TB(A->GT, o, -i);
                c[6] = TB(A->GR, i, o); c[7] = TB(A->GR, i, -o);
                c[8] = TB(A->GR, o, i); c[9] = TB(A->GR, o, -i);
                c[10] = TB(A->ARR, i, o); c[11] = TB(A->ARR, i, -o);
                c[12] = TB(A->ART, i, o); c[13] = TB(A->ART, i, -o);
                c[14] = TB(A->ART, o, i); c[15] = TB(A->ART, o, -i);
                c[16] = TB(A->ATT, i, o); c[17] = TB(A->ATT, i, -o);
                if (mol) {
                    r[0] = b0 + be2 * pi_ * po; r[1] = b0 + be2 * pi_ * pom;
                    r[2] = g2 * pi_ * to; r[3] = g2 * pi_ * tom;
                    r[4] = g2 * po * ti; r[5] = g2 * po * tim;
                    r[6] = g2 * pi_ * ro; r[7] = g2 * pi_ * rom;
                    r[8] = g2 * po * ri; r[9] = g2 * po * rim;
                    r[10] = a2 * ri * ro; r[11] = a2 * ri * rom;
                    r[12] = a2 * ti * ro; r[13] = a2 * ti * rom;
                    r[14] = a2 * to * ri; r[15] = a2 * to * rim;
                    r[16] = a2 * ti * to; r[17] = a2 * ti * tom;
                } else {
                    for (int q = 0; q < 18; ++q) r[q] = 0.0;
                }
                for (int q = 0; q < 18; ++q) e[q] = c[q] * x + y * r[q];
                const double a0 = Ip[FI(k, i)], a1 = Ip[FI(k, -i)], a2_ = Qp[FI(k, i)], a3 = Qp[FI(k, -i)];
                const double a4 = Up[FI(k, i)], a5 = Up[FI(k, -i)];
                FbI[FI(k, i)] += w * (biu * e[0] + bid * e[1] + bqu * e[6] + bqd * e[7] - buu * e[2] - bud * e[3]);
                FbI[FI(k, -i)] += w * (biu * e[1] + bid * e[0] + bqu * e[7] + bqd * e[6] + buu * e[3] + bud * e[2]);
                FbQ[FI(k, i)] += w * (biu * e[8] + bid * e[9] + bqu * e[10] + bqd * e[11] - buu * e[14] + bud * e[15]);
                FbQ[FI(k, -i)] += w * (biu * e[9] + bid * e[8] + bqu * e[11] + bqd * e[10] - buu * e[15] + bud * e[14]);
                FbU[FI(k, i)] += w * (-biu * e[4] + bid * e[5] - bqu * e[12] - bqd * e[13] + buu * e[16] + bud * e[17]);
                FbU[FI(k, -i)] += w * (-biu * e[5] + bid * e[4] + bqu * e[13] + bqd * e[12] + buu * e[17] + bud * e[16]);
                E[0] = biu * a0 + bid * a1;    E[1] = biu * a1 + bid * a0;
                E[2] = -buu * a0 + bud * a1;   E[3] = buu * a1 - bud * a0;
                E[4] = -biu * a4 + bid * a5;   E[5] = -biu * a5 + bid * a4;
                E[6] = bqu * a0 + bqd * a1;    E[7] = bqu * a1 + bqd * a0;
                E[8] = biu * a2_ + bid * a3;   E[9] = biu * a3 + bid * a2_;
                E[10] = bqu * a2_ + bqd * a3;  E[11] = bqu * a3 + bqd * a2_;
                E[12] = -bqu * a4 + bqd * a5;  E[13] = bqu * a5 - bqd * a4;
                E[14] = -buu * a2_ + bud * a3; E[15] = -buu * a3 + bud * a2_;
                E[16] = buu * a4 + bud * a5;   E[17] = buu * a5 + bud * a4;
                double sx = 0.0, sy = 0.0;
                for (int q = 0; q < 18; ++q) { sx += c[q] * E[q]; sy += r[q] * E[q]; }
                gx[k] += w * sx;
                gy[k] += w * sy;
            }
        }
    }
}

/* Direct sunglint of the result (rho_I, rho_Q, rho_U) contributions. */
static void glint3(const ocrt_instance_t *in, const med_t *M, double ws, double tau, double *g)
{
    const double muv = M->muv, phv = (in->raa_deg - 180.0) * PI / 180.0;
    double rg[3];
    sunglint_rho(in->vza_deg > 0 ? muv : 1.0, phv, M->mus, ws, in->sigma_type, in->n_water, tau,
                 in->q_convention, rg);
    g[0] = rg[0]; g[1] = rg[1]; g[2] = -rg[2];
}

int oracle_adjoint(const ocrt_instance_t *in, const double *gout, double *g_h, double *g_x, double *g_y,
                   double *g_wind)
{
    if (in->bottom_source || in->bottom_source_only || in->water_tau > 0.0) return -1;   /* forward-only options */
    if (in->ipss) return -1;
    solve_t s;
    int rc = solve_core(in, &s);
    if (rc) { solve_free(&s); return rc; }
    const med_t *M = &s.M;
    const int n = M->n, nt = M->nt, dirs = M->dirs, mm = in->m_max;
    const size_t cnt = (size_t)(nt + 1) * dirs;
    const double mus = M->mus, phi = in->raa_deg * PI / 180.0 + PI;
    for (int k = 0; k <= nt; ++k) g_h[k] = g_x[k] = g_y[k] = 0.0;
    *g_wind = 0.0;
    const int surf_aware = (in->surface == OCRT_SURFACE_BLACK_FRESNEL_OCEAN || in->surface == OCRT_SURFACE_FLAT);
    const int flat = in->surface == OCRT_SURFACE_FLAT ||
                     (in->surface == OCRT_SURFACE_BLACK_FRESNEL_OCEAN && in->wind_speed <= 0.0);
    const int kind = surf_aware ? (flat ? 1 : 2) : 0;
    /* surface kernels and their wind derivative (central difference in wind) */
    const size_t rsz = (size_t)(mm + 1) * n * n * 9, ssz = (size_t)(mm + 1) * n * 9;
    double *Rall = NULL, *Rsun = NULL, *dRall = NULL, *dRsun = NULL, *tmp = NULL;
    if (kind == 2) {
        Rall = malloc(sizeof(double) * rsz); dRall = malloc(sizeof(double) * rsz);
        Rsun = malloc(sizeof(double) * ssz); dRsun = malloc(sizeof(double) * ssz);
        tmp = malloc(sizeof(double) * (rsz + ssz));
        const double smu = mus, ws = in->wind_speed, dw = 1e-5 * ws;
        oracle_coxmunk_fourier(&M->rm[1], n, &M->rm[1], n, mm, 1024, ws, in->sigma_type, in->n_water,
                               in->q_convention, Rall);
        oracle_coxmunk_fourier(&M->rm[1], n, &smu, 1, mm, 1024, ws, in->sigma_type, in->n_water,
                               in->q_convention, Rsun);
        oracle_coxmunk_fourier(&M->rm[1], n, &M->rm[1], n, mm, 1024, ws + dw, in->sigma_type, in->n_water,
                               in->q_convention, dRall);
        oracle_coxmunk_fourier(&M->rm[1], n, &M->rm[1], n, mm, 1024, ws - dw, in->sigma_type, in->n_water,
                               in->q_convention, tmp);
        for (size_t e = 0; e < rsz; ++e) dRall[e] = (dRall[e] - tmp[e]) / (2.0 * dw);
        oracle_coxmunk_fourier(&M->rm[1], n, &smu, 1, mm, 1024, ws + dw, in->sigma_type, in->n_water,
                               in->q_convention, dRsun);
        oracle_coxmunk_fourier(&M->rm[1], n, &smu, 1, mm, 1024, ws - dw, in->sigma_type, in->n_water,
                               in->q_convention, tmp);
        for (size_t e = 0; e < ssz; ++e) dRsun[e] = (dRsun[e] - tmp[e]) / (2.0 * dw);
    }
    const size_t bcount = (size_t)(M->lmax + 1) * dirs, kcount = (size_t)(n + 1) * dirs;
    ang_t Ast;
    ang_t *const A = &Ast;
    double *abuf = calloc(3 * bcount + 6 * kcount, sizeof(double));
    A->plm = abuf; A->rrl = A->plm + bcount; A->rtl = A->rrl + bcount;
    A->P = A->rtl + bcount; A->GR = A->P + kcount; A->GT = A->GR + kcount;
    A->ARR = A->GT + kcount; A->ART = A->ARR + kcount; A->ATT = A->ART + kcount;
    int maxN = 1;
    for (int m = 0; m <= mm; ++m) if (s.orders[m] > maxN) maxN = s.orders[m];
    /* per order: field u (3 cnt), source J (3 cnt); work: lam, Sbar, Fbar, seeds */
    double *U = calloc((size_t)(maxN + 1) * 6 * cnt + 15 * cnt + 6 * (size_t)n, sizeof(double));
    if (!U || !abuf) { free(U); free(abuf); free(Rall); free(Rsun); free(dRall); free(dRsun); free(tmp);
                       solve_free(&s); return -3; }
    double *lam = U + (size_t)(maxN + 1) * 6 * cnt;
    double *sb3 = lam + 3 * cnt, *fb3 = sb3 + 3 * cnt, *f3 = fb3 + 3 * cnt, *d3 = f3 + 3 * cnt;
    double *bc3 = d3 + 3 * cnt;                      /* [3][n] */
    double *bcb3 = bc3 + 3 * n;                      /* [3][n] */
    double wt[3];
    for (int m = 0; m <= mm && rc == 0; ++m) {
        oracle_basis(M->lmax, m, dirs, M->rm_s, A->plm, A->rrl, A->rtl);
        build_tables(M, m, A);
        if (M->n_ang > 1) value_table(M, m, A);
        const double W = (kind == 1) ? 2.0 * PI : ((m == 0) ? 2.0 * PI : PI);
        const double *Rm = kind == 2 ? Rall + (size_t)m * n * n * 9 : NULL;
        const double *dRm = kind == 2 ? dRall + (size_t)m * n * n * 9 : NULL;
        const double *Rs = kind == 2 ? Rsun + (size_t)m * n * 9 : NULL;
        const double *dRs = kind == 2 ? dRsun + (size_t)m * n * 9 : NULL;
        wt[0] = wt[1] = (m == 0) ? 1.0 : 2.0 * cos(m * phi);
        wt[2] = (m == 0) ? 0.0 : 2.0 * sin(m * phi);
        /* ---- forward with storage: u_o at U + o*6cnt, J_o right after */
#define UF(o, c) (U + (size_t)(o) * 6 * cnt + (size_t)(c) * cnt)
#define JF(o, c) (U + (size_t)(o) * 6 * cnt + (size_t)(3 + (c)) * cnt)
        memset(U, 0, sizeof(double) * (size_t)(maxN + 1) * 6 * cnt);
        memset(f3, 0, sizeof(double) * 3 * cnt);
        memset(d3, 0, sizeof(double) * 3 * cnt);
        primary_source(M, m, A, JF(1, 0), JF(1, 1), JF(1, 2));
        double fl_RI = 0, fl_RQ = 0, fl_sb = 0;
        if (kind == 1) {
            /* flat seed source, as in sos_loop */
            const double nw = in->n_water, n2 = nw * nw;
            const double tn = sqrt(n2 - 1.0 + mus * mus);
            const double rp = (n2 * mus - tn) / (n2 * mus + tn), rs = (mus - tn) / (mus + tn);
            fl_RI = (rp * rp + rs * rs) / 2.0;
            fl_RQ = (in->q_convention == 1) ? (rp * rp - rs * rs) / 2.0 : (rs * rs - rp * rp) / 2.0;
            fl_sb = 0.5 * exp(-2.0 * M->h[nt] / mus);
            const double bP = XPL(0), bR = XRL(0), b0 = (m == 0) ? M->beta0 : 0.0;
            const int mol = (m <= 2);
            for (int L = 0; L < nt; ++L) {
                const double yt = M->yd[L], xt = M->xd[L], yb = M->yd[L + 1], xb = M->xd[L + 1];
                const double bt = fl_sb * exp(M->h[L] / mus), bb = fl_sb * exp(M->h[L + 1] / mus);
                for (int d = 1; d <= n; ++d) {
                    const double pu = XPL(d), pd = XPL(-d), ru = XRL(d), rd = XRL(-d), tu = XTL(d), td = XTL(-d);
                    const double kPd = TB(A->P, 0, -d) * xt + (mol ? b0 + M->beta2 * pd * bP : 0.0) * yt;
                    const double kRd = TB(A->GR, 0, -d) * xt + (mol ? M->gamma2 * rd * bP : 0.0) * yt;
                    const double kTd = TB(A->GT, 0, -d) * xt + (mol ? M->gamma2 * td * bP : 0.0) * yt;
                    const double kRTd = moment_gr(M, A, m, -d, 0) * xt + (mol ? M->gamma2 * bR * pd : 0.0) * yt;
                    const double kRRd = TB(A->ARR, 0, -d) * xt + (mol ? M->alpha2 * bR * rd : 0.0) * yt;
                    const double kTRd = moment_art(M, A, m, -d, 0) * xt + (mol ? M->alpha2 * td * bR : 0.0) * yt;
                    const double kPu = TB(A->P, 0, d) * xb + (mol ? b0 + M->beta2 * pu * bP : 0.0) * yb;
                    const double kRu = TB(A->GR, 0, d) * xb + (mol ? M->gamma2 * ru * bP : 0.0) * yb;
                    const double kTu = TB(A->GT, 0, d) * xb + (mol ? M->gamma2 * tu * bP : 0.0) * yb;
                    const double kRTu = moment_gr(M, A, m, d, 0) * xb + (mol ? M->gamma2 * bR * pu : 0.0) * yb;
                    const double kRRu = TB(A->ARR, 0, d) * xb + (mol ? M->alpha2 * bR * ru : 0.0) * yb;
                    const double kTRu = moment_art(M, A, m, d, 0) * xb + (mol ? M->alpha2 * tu * bR : 0.0) * yb;
                    f3[FI(L, d)] = bt * (fl_RI * kPd + fl_RQ * kRTd);
                    f3[cnt + FI(L, d)] = bt * (fl_RI * kRd + fl_RQ * kRRd);
                    f3[2 * cnt + FI(L, d)] = bt * (fl_RI * kTd + fl_RQ * kTRd);
                    f3[FI(L + 1, -d)] = bb * (fl_RI * kPu + fl_RQ * kRTu);
                    f3[cnt + FI(L + 1, -d)] = bb * (fl_RI * kRu + fl_RQ * kRRu);
                    f3[2 * cnt + FI(L + 1, -d)] = bb * (fl_RI * kTu + fl_RQ * kTRu);
                }
            }
        }
        /* order-1 source = primary + flat seed source (the sweep is linear) */
        for (int c = 0; c < 3; ++c)
            for (size_t i = 0; i < cnt; ++i) JF(1, c)[i] += f3[c * cnt + i];
        for (int k = 0; k <= nt; ++k) for (int c = 0; c < 3; ++c) JF(1, c)[FI(k, 0)] = 0.0;
        for (int c = 0; c < 3; ++c) sweep(M, JF(1, c), NULL, UF(1, c));
        if (kind == 2) {
            const double tau = M->h[nt], down = exp(-tau / mus);
            const double FS = ((m == 0) ? PI : 0.5 * PI) * mus;
            for (int k = 0; k <= nt; ++k)
                for (int o = 0; o < n; ++o) {
                    const double f = FS * (down * exp(-(tau - M->h[k]) / M->rm[o + 1]));
                    for (int c = 0; c < 3; ++c) {
                        d3[c * cnt + FI(k, o + 1)] = Rs[o * 9 + 3 * c] * f;
                        UF(1, c)[FI(k, o + 1)] += d3[c * cnt + FI(k, o + 1)];
                    }
                }
        }
        const int N = s.orders[m];
        for (int it = 2; it <= N; ++it) {
            apply_operator(M, m, A, UF(it - 1, 0), UF(it - 1, 1), UF(it - 1, 2), JF(it, 0), JF(it, 1), JF(it, 2));
            for (int k = 0; k <= nt; ++k) for (int c = 0; c < 3; ++c) JF(it, c)[FI(k, 0)] = 0.0;
            const double *bptr[3] = {NULL, NULL, NULL};
            if (kind != 0) {
                for (int o = 1; o <= n; ++o) {
                    double sv[3] = {0, 0, 0};
                    if (kind == 1) {
                        double F[9];
                        fresnel_R(M->rm[o], in->n_water, in->q_convention, F);
                        const double a = UF(it - 1, 0)[FI(nt, -o)], b = UF(it - 1, 1)[FI(nt, -o)],
                                     c = UF(it - 1, 2)[FI(nt, -o)];
                        /* flat: diagonal discrete delta, W mu w (1 / (2 pi mu w)) F = F */
                        if (M->gb[o] == 0.0) {
                            for (int r = 0; r < 3; ++r) sv[r] = F[3 * r] * a + F[3 * r + 1] * b + F[3 * r + 2] * c;
                        } else {
                            const double dd = 1.0 / (2.0 * PI * M->rm[o] * M->gb[o]), pw = W * M->rm[o] * M->gb[o];
                            for (int r = 0; r < 3; ++r)
                                sv[r] = pw * (dd * F[3 * r] * a + dd * F[3 * r + 1] * b + dd * F[3 * r + 2] * c);
                        }
                    } else {
                        for (int i = 1; i <= n; ++i) {
                            const double w = M->gb[i];
                            if (w == 0.0) continue;
                            const double pw = W * M->rm[i] * w;
                            const double *R = Rm + ((size_t)(o - 1) * n + (i - 1)) * 9;
                            for (int r = 0; r < 3; ++r)
                                sv[r] += pw * (R[3 * r] * UF(it - 1, 0)[FI(nt, -i)] + R[3 * r + 1] * UF(it - 1, 1)[FI(nt, -i)] +
                                               R[3 * r + 2] * UF(it - 1, 2)[FI(nt, -i)]);
                        }
                    }
                    for (int r = 0; r < 3; ++r) bc3[r * n + o - 1] = sv[r];
                }
                bptr[0] = bc3; bptr[1] = bc3 + n; bptr[2] = bc3 + 2 * n;
            }
            for (int c = 0; c < 3; ++c) sweep(M, JF(it, c), bptr[c], UF(it, c));
        }
        /* ---- backward */
        double ev[3];
        for (int c = 0; c < 3; ++c) ev[c] = gout[c] * wt[c] / mus;
        memset(lam, 0, sizeof(double) * 3 * cnt);
        for (int c = 0; c < 3; ++c) lam[c * cnt + FI(0, M->vidx)] = ev[c];
        for (int it = N; it >= 2; --it) {
            memset(sb3, 0, sizeof(double) * 3 * cnt);
            memset(bcb3, 0, sizeof(double) * 3 * n);
            for (int c = 0; c < 3; ++c)
                sweep_adj(M, lam + c * cnt, UF(it, c), JF(it, c), sb3 + c * cnt, kind ? bcb3 + c * n : NULL, g_h);
            for (int k = 0; k <= nt; ++k) for (int c = 0; c < 3; ++c) sb3[c * cnt + FI(k, 0)] = 0.0;
            memset(fb3, 0, sizeof(double) * 3 * cnt);
            apply_operator_adj(M, m, A, sb3, sb3 + cnt, sb3 + 2 * cnt, UF(it - 1, 0), UF(it - 1, 1), UF(it - 1, 2),
                               fb3, fb3 + cnt, fb3 + 2 * cnt, g_x, g_y);
            if (kind != 0) {
                for (int o = 1; o <= n; ++o) {
                    const double bb[3] = {bcb3[o - 1], bcb3[n + o - 1], bcb3[2 * n + o - 1]};
                    if (kind == 1) {
                        double F[9];
                        fresnel_R(M->rm[o], in->n_water, in->q_convention, F);
                        const double sc = (M->gb[o] == 0.0) ? 1.0
                                          : (W * M->rm[o] * M->gb[o]) / (2.0 * PI * M->rm[o] * M->gb[o]);
                        for (int c2 = 0; c2 < 3; ++c2)
                            fb3[c2 * cnt + FI(nt, -o)] += sc * (bb[0] * F[c2] + bb[1] * F[3 + c2] + bb[2] * F[6 + c2]);
                    } else {
                        for (int i = 1; i <= n; ++i) {
                            const double w = M->gb[i];
                            if (w == 0.0) continue;
                            const double pw = W * M->rm[i] * w;
                            const double *R = Rm + ((size_t)(o - 1) * n + (i - 1)) * 9;
                            const double *dR = dRm + ((size_t)(o - 1) * n + (i - 1)) * 9;
                            for (int c2 = 0; c2 < 3; ++c2) {
                                fb3[c2 * cnt + FI(nt, -i)] += pw * (bb[0] * R[c2] + bb[1] * R[3 + c2] + bb[2] * R[6 + c2]);
                                *g_wind += pw * (bb[0] * dR[c2] + bb[1] * dR[3 + c2] + bb[2] * dR[6 + c2]) *
                                           UF(it - 1, c2)[FI(nt, -i)];
                            }
                        }
                    }
                }
            }
            for (int c = 0; c < 3; ++c)
                for (size_t i = 0; i < cnt; ++i) lam[c * cnt + i] = fb3[c * cnt + i];
            for (int c = 0; c < 3; ++c) lam[c * cnt + FI(0, M->vidx)] += ev[c];
        }
        /* order 1: adjoint sweep through the combined primary + flat source */
        memset(sb3, 0, sizeof(double) * 3 * cnt);
        for (int c = 0; c < 3; ++c) {
            /* the sweep part of u_1 excludes the rough seed */
            double *Fp = fb3 + c * cnt;
            for (size_t i = 0; i < cnt; ++i) Fp[i] = UF(1, c)[i] - d3[c * cnt + i];
            sweep_adj(M, lam + c * cnt, Fp, JF(1, c), sb3 + c * cnt, NULL, g_h);
        }
        {
            const double b0 = (m == 0) ? M->beta0 : 0.0;
            const int mol = (m <= 2);
            for (int k = 0; k <= nt; ++k) {
                const double beam = M->ch[k], x = M->xd[k], y = M->yd[k];
                for (int j = -n; j <= n; ++j) {
                    if (j == 0) continue;
                    const double mi = b0 + M->beta2 * XPL(j) * XPL(0);
                    const double mq = mol ? M->gamma2 * XRL(j) * XPL(0) : 0.0;
                    const double mu_ = mol ? M->gamma2 * XTL(j) * XPL(0) : 0.0;
                    const double sI = sb3[FI(k, j)], sQ = sb3[cnt + FI(k, j)], sU = sb3[2 * cnt + FI(k, j)];
                    g_x[k] += beam * (sI * TB(A->P, 0, j) + sQ * TB(A->GR, 0, j) - sU * TB(A->GT, 0, j));
                    g_y[k] += beam * (sI * mi + sQ * mq - sU * mu_);
                    const double vI = beam * (x * TB(A->P, 0, j) + y * mi);
                    const double vQ = beam * (x * TB(A->GR, 0, j) + y * mq);
                    const double vU = -beam * (x * TB(A->GT, 0, j) + y * mu_);
                    g_h[k] += -(sI * vI + sQ * vQ + sU * vU) / mus;
                }
            }
            if (kind == 1) {
                const double bP = XPL(0), bR = XRL(0);
                for (int L = 0; L < nt; ++L) {
                    const double bt = fl_sb * exp(M->h[L] / mus), bb = fl_sb * exp(M->h[L + 1] / mus);
                    for (int d = 1; d <= n; ++d) {
                        const double pu = XPL(d), pd = XPL(-d), ru = XRL(d), rd = XRL(-d), tu = XTL(d), td = XTL(-d);
                        /* upward entries at level L use (xt, yt) = level L */
                        const double axI = TB(A->P, 0, -d) * fl_RI + moment_gr(M, A, m, -d, 0) * fl_RQ;
                        const double axQ = TB(A->GR, 0, -d) * fl_RI + TB(A->ARR, 0, -d) * fl_RQ;
                        const double axU = TB(A->GT, 0, -d) * fl_RI + moment_art(M, A, m, -d, 0) * fl_RQ;
                        const double ayI = (mol ? b0 + M->beta2 * pd * bP : 0.0) * fl_RI + (mol ? M->gamma2 * bR * pd : 0.0) * fl_RQ;
                        const double ayQ = (mol ? M->gamma2 * rd * bP : 0.0) * fl_RI + (mol ? M->alpha2 * bR * rd : 0.0) * fl_RQ;
                        const double ayU = (mol ? M->gamma2 * td * bP : 0.0) * fl_RI + (mol ? M->alpha2 * td * bR : 0.0) * fl_RQ;
                        const size_t iu = FI(L, d);
                        const double su[3] = {sb3[iu], sb3[cnt + iu], sb3[2 * cnt + iu]};
                        g_x[L] += bt * (su[0] * axI + su[1] * axQ + su[2] * axU);
                        g_y[L] += bt * (su[0] * ayI + su[1] * ayQ + su[2] * ayU);
                        const double fu = su[0] * f3[iu] + su[1] * f3[cnt + iu] + su[2] * f3[2 * cnt + iu];
                        g_h[L] += fu / mus; g_h[nt] += -2.0 * fu / mus;
                        /* downward entries at level L + 1 use level L + 1 */
                        const double bxI = TB(A->P, 0, d) * fl_RI + moment_gr(M, A, m, d, 0) * fl_RQ;
                        const double bxQ = TB(A->GR, 0, d) * fl_RI + TB(A->ARR, 0, d) * fl_RQ;
                        const double bxU = TB(A->GT, 0, d) * fl_RI + moment_art(M, A, m, d, 0) * fl_RQ;
                        const double byI = (mol ? b0 + M->beta2 * pu * bP : 0.0) * fl_RI + (mol ? M->gamma2 * bR * pu : 0.0) * fl_RQ;
                        const double byQ = (mol ? M->gamma2 * ru * bP : 0.0) * fl_RI + (mol ? M->alpha2 * bR * ru : 0.0) * fl_RQ;
                        const double byU = (mol ? M->gamma2 * tu * bP : 0.0) * fl_RI + (mol ? M->alpha2 * tu * bR : 0.0) * fl_RQ;
                        const size_t id = FI(L + 1, -d);
                        const double sd[3] = {sb3[id], sb3[cnt + id], sb3[2 * cnt + id]};
                        g_x[L + 1] += bb * (sd[0] * bxI + sd[1] * bxQ + sd[2] * bxU);
                        g_y[L + 1] += bb * (sd[0] * byI + sd[1] * byQ + sd[2] * byU);
                        const double fd = sd[0] * f3[id] + sd[1] * f3[cnt + id] + sd[2] * f3[2 * cnt + id];
                        g_h[L + 1] += fd / mus; g_h[nt] += -2.0 * fd / mus;
                    }
                }
            }
            if (kind == 2) {
                /* the seed enters u_1 (adjoint lam) and is removed from the output (-e) */
                const double tau = M->h[nt], down = exp(-tau / mus);
                const double FS = ((m == 0) ? PI : 0.5 * PI) * mus;
                for (int k = 0; k <= nt; ++k)
                    for (int o = 0; o < n; ++o) {
                        const double mo = M->rm[o + 1];
                        const double f = FS * (down * exp(-(tau - M->h[k]) / mo));
                        for (int c = 0; c < 3; ++c) {
                            double gd = lam[c * cnt + FI(k, o + 1)];
                            if (k == 0 && o + 1 == M->vidx) gd -= ev[c];
                            const double dv = d3[c * cnt + FI(k, o + 1)];
                            g_h[k] += gd * dv / mo;
                            g_h[nt] += -gd * dv * (1.0 / mus + 1.0 / mo);
                            *g_wind += gd * dRs[o * 9 + 3 * c] * f;
                        }
                    }
            }
        }
#undef UF
#undef JF
    }
    /* direct glint in rho (black_fresnel_ocean, glint not decoupled) */
    if (in->surface == OCRT_SURFACE_BLACK_FRESNEL_OCEAN && in->decouple_sunglint == 0) {
        double g[3];
        const double tau = M->h[nt], muv1 = in->vza_deg > 0 ? M->muv : 1.0;
        glint3(in, M, in->wind_speed, tau, g);
        g_h[nt] += -(gout[0] * g[0] + gout[1] * g[1] + gout[2] * g[2]) * (1.0 / mus + 1.0 / muv1);
        if (in->wind_speed > 0.0) {
            const double dw = 1e-5 * in->wind_speed;
            double gp[3], gm[3];
            glint3(in, M, in->wind_speed + dw, tau, gp);
            glint3(in, M, in->wind_speed - dw, tau, gm);
            for (int c = 0; c < 3; ++c) *g_wind += gout[c] * (gp[c] - gm[c]) / (2.0 * dw);
        }
    }
    free(U); free(abuf); free(Rall); free(Rsun); free(dRall); free(dRsun); free(tmp);
    solve_free(&s);
    return rc;
}
