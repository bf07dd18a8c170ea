// Device-side fp64 helpers: US Standard Atmosphere 1962 column profile,
// Fresnel and Cox-Munk surface Mueller matrices, storage conversions.
#pragma once
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

namespace ocrt {

// US62 anchors (rt_molecular_profile.c:10-16) and WGS84 gravity (rt_rayleigh.c:46-52).
__constant__ static const double c_us62_z[8] = {0.0, 11.0, 20.0, 32.0, 47.0, 51.0, 71.0, 84.852};
__constant__ static const double c_us62_p[8] = {101325.0, 22632.1, 5474.89, 868.019,
                                                110.906, 66.9389, 3.95642, 0.3734};

__device__ inline double gravity_wgs84(double lat_deg, double alt_m)
{
    const double lat = lat_deg * OCRT_PI / 180.0;
    const double s = sin(lat) * sin(lat);
    const double s2 = sin(2.0 * lat) * sin(2.0 * lat);
    const double g0 = 9.780327 * (1.0 + 0.0053024 * s - 0.0000058 * s2);
    return g0 - 3.086e-6 * alt_m;
}

__device__ inline double clamp01(double x) { return x <= 0.0 ? 0.0 : (x >= 1.0 ? 1.0 : x); }

__device__ inline double us62_pressure(double z)
{
    if (z <= c_us62_z[0]) return c_us62_p[0];
    if (z >= c_us62_z[7]) return c_us62_p[7];
    int lo = 0, hi = 7;
    while (hi - lo > 1) {
        const int mid = lo + (hi - lo) / 2;
        if (z < c_us62_z[mid]) hi = mid; else lo = mid;
    }
    const double t = (z - c_us62_z[lo]) / (c_us62_z[hi] - c_us62_z[lo]);
    const double a = log(c_us62_p[lo]), b = log(c_us62_p[hi]);
    return exp(a + t * (b - a));
}

__device__ inline double us62_fraction_above(double z)
{
    if (z <= 0.0) return 1.0;
    const double g0 = gravity_wgs84(45.0, 0.0);
    if (z >= c_us62_z[7]) {
        const double gt = gravity_wgs84(45.0, c_us62_z[7] * 1000.0);
        return clamp01((c_us62_p[7] / gt) / (c_us62_p[0] / g0));
    }
    const double gz = gravity_wgs84(45.0, z * 1000.0);
    return clamp01((us62_pressure(z) / gz) / (c_us62_p[0] / g0));
}

__device__ inline double us62_grid_fraction_above(double z)
{
    if (z <= 0.0) return 1.0;
    if (z >= c_us62_z[7]) return 0.0;
    const double ft = us62_fraction_above(c_us62_z[7]);
    return clamp01((us62_fraction_above(z) - ft) / (1.0 - ft));
}

__device__ inline double us62_altitude_from_fraction(double f)
{
    if (f >= 1.0) return 0.0;
    if (f <= 0.0) return c_us62_z[7];
    double lo = 0.0, hi = c_us62_z[7];
    for (int it = 0; it < 64; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (us62_grid_fraction_above(mid) > f) lo = mid; else hi = mid;
    }
    return 0.5 * (lo + hi);
}

__device__ inline double aerosol_fraction_above(double z, double ztop, double H)
{
    if (z <= 0.0) return 1.0;
    if (z >= ztop) return 0.0;
    const double et = exp(-ztop / H);
    return (exp(-z / H) - et) / (1.0 - et);
}

__host__ __device__ inline double diag_norm(int m)
{
    double v = 1.0;
    for (int i = 1; i <= m; ++i) v *= sqrt((double)(i + m) / (double)i) * 0.5;
    return v;
}

// Slope variance of the isotropic Gaussian facet distribution (surface.c:57-64).
__device__ inline double slope_variance(double ws, int sigma_type)
{
    if (sigma_type == 1) return 0.003 + 0.00512 * fmax(0.01, ws);
    const double s = 0.0731 * sqrt(fmax(0.0, ws));
    return fmax(s * s, 1e-10);
}

// Air-side Fresnel reflection Mueller matrix, meridian basis (surface.c:385-468).
__device__ inline void fresnel_R(double mu_i, double n_water, int qc, double *M)
{
    for (int i = 0; i < 9; ++i) M[i] = 0.0;
    const double sin2 = fmax(0.0, 1.0 - mu_i * mu_i);
    const double st2 = (1.0 / n_water) * (1.0 / n_water) * sin2;
    if (st2 >= 1.0) { M[0] = M[4] = M[8] = 1.0; return; }
    const double mt = sqrt(1.0 - st2);
    const double rs = (mu_i - n_water * mt) / (mu_i + n_water * mt);
    const double rp = (n_water * mu_i - mt) / (n_water * mu_i + mt);
    const double Rs = rs * rs, Rp = rp * rp;
    const double q = (qc == 1) ? 0.5 * (Rp - Rs) : 0.5 * (Rs - Rp);
    M[0] = 0.5 * (Rs + Rp); M[1] = q; M[3] = q; M[4] = 0.5 * (Rs + Rp); M[8] = rs * rp;
}

// Cox-Munk Mueller BRDF with Sancer shadowing and the meridian/facet-plane
// rotations, including the exact pole limits (surface.c:514-646).
// Sancer bistatic shadowing S = 1 / (1 + Lambda_i + Lambda_v) (surface.c:591-612);
// it depends on the two zenith cosines only.
// Fresnel reflection Mueller [3][3] of light inside the water meeting the
// surface at cosine mu_w (n1 = n_water, n2 = 1), meridian basis; total
// internal reflection gives the identity (shared/surface.c:385-429,
// rt_air_water.c:70-74).
__device__ inline void fresnel_R_water_side(double mu_w, double n_water, int qc, double *M)
{
    for (int e = 0; e < 9; ++e) M[e] = 0.0;
    const double sin2 = fmax(0.0, 1.0 - mu_w * mu_w);
    const double st2 = n_water * n_water * sin2;
    if (st2 >= 1.0) { M[0] = M[4] = M[8] = 1.0; return; }
    const double mt = sqrt(1.0 - st2);
    const double rs = (n_water * mu_w - mt) / (n_water * mu_w + mt);
    const double rp = (mu_w - n_water * mt) / (mu_w + n_water * mt);
    const double Rs = rs * rs, Rp = rp * rp;
    const double q = (qc == 1) ? 0.5 * (Rp - Rs) : 0.5 * (Rs - Rp);
    M[0] = 0.5 * (Rs + Rp); M[1] = q; M[3] = q; M[4] = 0.5 * (Rs + Rp); M[8] = rs * rp;
}

__device__ inline double sancer_shadow(double mu_out, double mu_in_dn, double s2v)
{
    const double sl = sqrt(s2v);
    double li = 0.0, lv = 0.0;
    if (mu_in_dn < 1.0 - 1e-10) {
        const double si = sqrt(fmax(1e-20, 1.0 - mu_in_dn * mu_in_dn));
        const double nu = (mu_in_dn / si) / sl;
        if (nu <= 20.0) li = 0.5 * (exp(-nu * nu) / (nu * sqrt(OCRT_PI)) - erfc(nu));
    }
    if (mu_out < 1.0 - 1e-10) {
        const double sv = sqrt(fmax(1e-20, 1.0 - mu_out * mu_out));
        const double nu = (mu_out / sv) / sl;
        if (nu <= 20.0) lv = 0.5 * (exp(-nu * nu) / (nu * sqrt(OCRT_PI)) - erfc(nu));
    }
    return 1.0 / (1.0 + li + lv);
}

// d ln S / d sigma^2 of the Sancer factor S = 1 / (1 + Lambda_i + Lambda_v):
// Lambda(nu) = (exp(-nu^2) / (nu sqrt(pi)) - erfc(nu)) / 2 with
// nu = cot(theta) / sigma, so dLambda / dsigma^2 = exp(-nu^2) / (4 sqrt(pi) nu sigma^2).
__device__ inline double sancer_dlog(double mu_out, double mu_in_dn, double s2v, double shadow)
{
    const double sl = sqrt(s2v);
    double d = 0.0;
    const double mus[2] = {mu_in_dn, mu_out};
    for (int q = 0; q < 2; ++q) {
        const double m = mus[q];
        if (m < 1.0 - 1e-10) {
            const double sn = sqrt(fmax(1e-20, 1.0 - m * m));
            const double nu = (m / sn) / sl;
            if (nu <= 20.0) d += exp(-nu * nu) / (4.0 * sqrt(OCRT_PI) * nu * s2v);
        }
    }
    return -shadow * d;
}

// Body of the Cox-Munk Mueller BRDF for a given slope variance and shadowing
// factor (shadow < 0: computed here).  tb2_out (optional) receives tan^2 of
// the facet tilt, from which d ln(brdf) / d sigma^2 = tb2 / sigma^4 - 1 / sigma^2
// + d ln S / d sigma^2 follows (-1 when the entry is identically zero).
__device__ inline void coxmunk_R_core(double mu_out, double mu_in_dn, double cphi, double sphi, double s2v,
                                      double shadow, double nw, int qc, double *R, double *tb2_out = nullptr)
{
    for (int i = 0; i < 9; ++i) R[i] = 0.0;
    if (tb2_out) *tb2_out = -1.0;
    const double mia = -mu_in_dn;
    const double s1 = sqrt(fmax(0.0, 1.0 - mu_out * mu_out));
    const double s2 = sqrt(fmax(0.0, 1.0 - mia * mia));
    double cT = mu_out * mia + s1 * s2 * cphi;
    cT = fmin(1.0, fmax(-1.0, cT));
    const double co = sqrt(fmax(0.0, 0.5 * (1.0 - cT)));
    if (co <= 1e-6) return;
    const double cb = (mu_out + mu_in_dn) / (2.0 * co);
    if (cb <= 1e-6) return;
    const double sT = sqrt(fmax(0.0, 1.0 - cT * cT));
    const double den = fmax(1e-12, sT);
    double ci1, si1, ci2, si2;
    if (s2 <= 1.0e-8) {
        ci1 = -copysign(1.0, mia) * cphi; si1 = sphi;
    } else {
        ci1 = (mu_out - mia * cT) / (den * fmax(1e-12, s2));
        si1 = s1 * sphi / den;
        const double q = sqrt(ci1 * ci1 + si1 * si1);
        if (q < 1e-12) { ci1 = 1.0; si1 = 0.0; } else { ci1 /= q; si1 /= q; }
    }
    if (s1 <= 1.0e-8) {
        ci2 = -copysign(1.0, mu_out) * cphi; si2 = sphi;
    } else {
        ci2 = (mia - mu_out * cT) / (den * fmax(1e-12, s1));
        si2 = s2 * sphi / den;
        const double q = sqrt(ci2 * ci2 + si2 * si2);
        if (q < 1e-12) { ci2 = 1.0; si2 = 0.0; } else { ci2 /= q; si2 /= q; }
    }
    double MF[9];
    fresnel_R(co, nw, qc, MF);
    const double cb2 = cb * cb;
    const double tb2 = (1.0 - cb2) / fmax(cb2, 1e-8);
    if (tb2_out) *tb2_out = tb2;
    const double Ps = (1.0 / (OCRT_PI * s2v)) * exp(-tb2 / s2v);
    const double S = shadow >= 0.0 ? shadow : sancer_shadow(mu_out, mu_in_dn, s2v);
    const double brdf = (Ps * S) / (4.0 * mu_in_dn * mu_out * cb2 * cb2);
    // R = L(i2) . MF . L(i1) . brdf with L(i) = diag(1, rot2(2i))
    const double c1 = ci1 * ci1 - si1 * si1, sn1 = 2.0 * ci1 * si1;
    const double c2 = ci2 * ci2 - si2 * si2, sn2 = 2.0 * ci2 * si2;
    const double L1[9] = {1, 0, 0, 0, c1, -sn1, 0, sn1, c1};
    const double L2[9] = {1, 0, 0, 0, c2, -sn2, 0, sn2, c2};
    double T1[9];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += MF[a * 3 + k] * L1[k * 3 + b];
            T1[a * 3 + b] = s;
        }
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += L2[a * 3 + k] * T1[k * 3 + b];
            R[a * 3 + b] = s * brdf;
        }
}

// coxmunk_R_core for one (outgoing, incoming) pair of zenith cosines at many
// azimuths.  Everything independent of the azimuth is formed once; the two
// meridian/facet-plane rotations enter through cos 2i and sin 2i taken as
// (x^2 - y^2) / (x^2 + y^2) and 2 x y / (x^2 + y^2) of the unnormalised
// direction components, which is the normalise-then-double-angle step of
// coxmunk_R_core without its square roots and divisions; the Fresnel
// amplitudes share one reciprocal, and R = L(i2) . MF . L(i1) . brdf is
// written out with the zeros of MF and L dropped.  Same branches, limits and
// thresholds as coxmunk_R_core; the results agree to rounding.
// T = double for every kernel of the fp64 form and the parity paths; the
// mixed form's diffuse surface kernels evaluate it in float (coxmunk_pair
// with float R), its fp32 copy being all the mixed form reads of them.
template <typename T> struct CoxMunkPairT {
    T mo, mia, s1, s2, mm, ss, msum, inv_s1, inv_s2, ps_pref, inv_s2v, fac, nw, inv_nw2;
    int qc;

    __device__ CoxMunkPairT(double mu_out, double mu_in_dn, double s2v, double shadow, double n_water, int q_conv)
    {
        // the pair's constants in fp64, then rounded to T
        const double mia_ = -mu_in_dn;
        const double s1_ = sqrt(fmax(0.0, 1.0 - mu_out * mu_out));
        const double s2_ = sqrt(fmax(0.0, 1.0 - mia_ * mia_));
        mo = (T)mu_out;
        mia = (T)mia_;
        s1 = (T)s1_;
        s2 = (T)s2_;
        mm = (T)(mu_out * mia_);
        ss = (T)(s1_ * s2_);
        msum = (T)(mu_out + mu_in_dn);
        inv_s1 = (T)(1.0 / fmax(1e-12, s1_));
        inv_s2 = (T)(1.0 / fmax(1e-12, s2_));
        ps_pref = (T)(1.0 / (OCRT_PI * s2v));
        inv_s2v = (T)(1.0 / s2v);
        fac = (T)(shadow / (4.0 * mu_in_dn * mu_out));
        nw = (T)n_water;
        inv_nw2 = (T)((1.0 / n_water) * (1.0 / n_water));
        qc = q_conv;
    }

    // cos 2i, sin 2i of one rotation from the unnormalised (x, y)
    __device__ static void double_angle(T x, T y, T &c, T &s)
    {
        const T qq = x * x + y * y;
        if (qq < (T)1e-24) { c = (T)1; s = (T)0; return; }
        const T iq = (T)1 / qq;
        c = (x * x - y * y) * iq;
        s = (T)2 * x * y * iq;
    }

    __device__ void eval(T cphi, T sphi, T *R, T &tb2) const
    {
        const T one = (T)1, zero = (T)0, half = (T)0.5;
#pragma unroll
        for (int i = 0; i < 9; ++i) R[i] = zero;
        tb2 = -one;
        T cT = mm + ss * cphi;
        cT = fmin(one, fmax(-one, cT));
        const T co = sqrt(fmax(zero, half * (one - cT)));
        if (co <= (T)1e-6) return;
        const T cb = msum / ((T)2 * co);
        if (cb <= (T)1e-6) return;
        const T rden = one / fmax((T)1e-12, sqrt(fmax(zero, one - cT * cT)));
        T c1, sn1, c2, sn2;
        if (s2 <= (T)1.0e-8) {
            const T ci = -copysign(one, mia) * cphi;
            c1 = ci * ci - sphi * sphi; sn1 = (T)2 * ci * sphi;
        } else {
            double_angle((mo - mia * cT) * rden * inv_s2, s1 * sphi * rden, c1, sn1);
        }
        if (s1 <= (T)1.0e-8) {
            const T ci = -copysign(one, mo) * cphi;
            c2 = ci * ci - sphi * sphi; sn2 = (T)2 * ci * sphi;
        } else {
            double_angle((mia - mo * cT) * rden * inv_s1, s2 * sphi * rden, c2, sn2);
        }
        // Fresnel reflection at the facet (fresnel_R at mu = co)
        T A = one, q = zero, m8 = one;
        const T st2 = inv_nw2 * fmax(zero, one - co * co);
        if (st2 < one) {
            const T mt = sqrt(one - st2);
            const T a = co - nw * mt, b = co + nw * mt, c = nw * co - mt, d = nw * co + mt;
            const T inv = one / (b * d);
            const T rs = a * d * inv, rp = c * b * inv;
            const T Rs = rs * rs, Rp = rp * rp;
            A = half * (Rs + Rp);
            q = (qc == 1) ? half * (Rp - Rs) : half * (Rs - Rp);
            m8 = rs * rp;
        }
        const T cb2 = cb * cb;
        tb2 = (one - cb2) / fmax(cb2, (T)1e-8);
        const T brdf = ps_pref * exp(-tb2 * inv_s2v) * fac / (cb2 * cb2);
        R[0] = A * brdf;
        R[1] = q * c1 * brdf;
        R[2] = -q * sn1 * brdf;
        R[3] = c2 * q * brdf;
        R[4] = (c2 * A * c1 - sn2 * m8 * sn1) * brdf;
        R[5] = (-c2 * A * sn1 - sn2 * m8 * c1) * brdf;
        R[6] = sn2 * q * brdf;
        R[7] = (sn2 * A * c1 + c2 * m8 * sn1) * brdf;
        R[8] = (-sn2 * A * sn1 + c2 * m8 * c1) * brdf;
    }
};
using CoxMunkPair = CoxMunkPairT<double>;

__device__ inline void coxmunk_R(double mu_out, double mu_in_dn, double cphi, double sphi, double ws,
                                 int sigma_type, double nw, int qc, double *R)
{
    coxmunk_R_core(mu_out, mu_in_dn, cphi, sphi, slope_variance(ws, sigma_type), -1.0, nw, qc, R);
}

// d sigma^2 / d wind of slope_variance.
__device__ inline double slope_variance_dwind(double ws, int sigma_type)
{
    if (sigma_type == 1) return ws > 0.01 ? 0.00512 : 0.0;
    const double s = 0.0731 * sqrt(fmax(0.0, ws));
    return (s * s > 1e-10) ? 0.0731 * 0.0731 : 0.0;
}

template <typename T> __device__ inline T to_storage(double v);
template <> __device__ inline double to_storage<double>(double v) { return v; }
template <> __device__ inline float to_storage<float>(double v) { return (float)v; }
template <> __device__ inline __hip_bfloat16 to_storage<__hip_bfloat16>(double v)
{
    return __float2bfloat16((float)v);
}

__device__ inline double as_double(double v) { return v; }
__device__ inline double as_double(float v) { return (double)v; }

}  // namespace ocrt
