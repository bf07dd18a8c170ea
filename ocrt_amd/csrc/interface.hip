// Rough air-water interface matrices of the coupled ocean: azimuthal Fourier
// kernels of the Cox-Munk Mueller matrices for
//   R_aa  air-side reflection               (code/OCRT_C/src/shared/surface.c:514-646, 839-1134)
//   R_ww  water-side internal reflection    (surface.c:1136-1249, 1386-1447)
//   T_wa  water -> air transmission         (surface.c:648-773, 1731-1811)
//   T_aw  air -> water transmission         (surface.c:80-136, 1476-1721)
// and the slope-integrated direct terms T_aw(mu_sun) and R_ww(mu_up)
// (surface.c:138-198, 229-368).
//
// One workgroup per (outgoing, incoming) node pair evaluates the 3x3 Mueller
// matrix at 256 azimuths at a time into LDS; thread m then projects the block
// onto cos(m phi) / sin(m phi) (weights from a table built once per call), so
// one pass over the azimuth rule serves every mode.
// Everything is fp64.  The entry points take host arrays, like the
// reference's functions, and own their device buffers.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "common.h"
#include "device_math.h"
#include "ocrt_amd.h"
#include "ocrt_rt.h"
#include "wave.h"

namespace ocrt {
namespace {

constexpr int kIfThreads = 256;       // azimuths per block and threads per pair
constexpr int kIfMaxModes = 256;      // m_max + 1 <= threads: one mode per thread
constexpr int kDirectThreads = 128;   // >= the 101 slope rows of the direct integrals
constexpr long long kIfMaxPairs = (1LL << 31) / kIfThreads - 1;   // work-items of one launch < 2^31

// Symmetric Fresnel-form matrix [[A q 0] [q A 0] [0 0 d]] rotated into the
// meridian planes: L(c2, s2) . M . L(c1, s1), with L = diag(1, [[c -s] [s c]])
// and (c, s) = (cos 2i, sin 2i) of each rotation angle.
__device__ inline void rotated_mueller(double A, double q, double d, double c1, double s1, double c2, double s2,
                                       double *R)
{
    R[0] = A;
    R[1] = q * c1;
    R[2] = -q * s1;
    R[3] = c2 * q;
    R[4] = c2 * A * c1 - s2 * d * s1;
    R[5] = -c2 * A * s1 - s2 * d * c1;
    R[6] = s2 * q;
    R[7] = s2 * A * c1 + c2 * d * s1;
    R[8] = -s2 * A * s1 + c2 * d * c1;
}

// cos 2i and sin 2i of the angle between the meridian plane of a direction
// (sine s_self, cosine mu_self) and the scattering plane.  Away from the pole
// the angle's (cos, sin) are (x, y) normalised (x = num / (den s_self),
// y = s_other sin(phi) / den; identity when |(x, y)| < 1e-12); at the pole
// (s_self <= 1e-8) they are (-sign(mu_self) cos phi, sin phi)
// (surface.c:731-755).
__device__ inline void plane_rotation(double s_self, double mu_self, double x, double y, double cphi, double sphi,
                                      double &c2, double &s2)
{
    if (s_self <= 1.0e-8) {
        const double ci = -copysign(1.0, mu_self) * cphi;
        c2 = ci * ci - sphi * sphi;
        s2 = 2.0 * ci * sphi;
        return;
    }
    CoxMunkPair::double_angle(x, y, c2, s2);
}

// Water-side internal reflection: the air-side geometry with the facet
// Fresnel from water (n) to air (1), identity beyond the critical angle.
struct RwwPair {
    double mo, mia, s1, s2, mm, ss, msum, inv_s1, inv_s2, ps_pref, inv_s2v, fac, nw;
    int qc;
    __device__ RwwPair(double mu_out, double mu_in, double s2v, double n_water, int q_conv)
    {
        mo = mu_out;
        mia = -mu_in;
        s1 = sqrt(fmax(0.0, 1.0 - mu_out * mu_out));
        s2 = sqrt(fmax(0.0, 1.0 - mia * mia));
        mm = mu_out * mia;
        ss = s1 * s2;
        msum = mu_out + mu_in;
        inv_s1 = 1.0 / fmax(1e-12, s1);
        inv_s2 = 1.0 / fmax(1e-12, s2);
        ps_pref = 1.0 / (OCRT_PI * s2v);
        inv_s2v = 1.0 / s2v;
        fac = sancer_shadow(mu_out, mu_in, s2v) / (4.0 * mu_in * mu_out);
        nw = n_water;
        qc = q_conv;
    }
    __device__ void eval(double cphi, double sphi, double *R) const
    {
#pragma unroll
        for (int k = 0; k < 9; ++k) R[k] = 0.0;
        const double cT = fmin(1.0, fmax(-1.0, mm + ss * cphi));
        const double co = sqrt(fmax(0.0, 0.5 * (1.0 - cT)));
        if (co <= 1e-6) return;
        const double cb = msum / (2.0 * co);
        if (cb <= 1e-6) return;
        const double rden = 1.0 / fmax(1e-12, sqrt(fmax(0.0, 1.0 - cT * cT)));
        double c1, n1, c2, n2;
        plane_rotation(s2, mia, (mo - mia * cT) * rden * inv_s2, s1 * sphi * rden, cphi, sphi, c1, n1);
        plane_rotation(s1, mo, (mia - mo * cT) * rden * inv_s1, s2 * sphi * rden, cphi, sphi, c2, n2);
        double A = 1.0, q = 0.0, d = 1.0;
        const double st2 = nw * nw * fmax(0.0, 1.0 - co * co);
        if (st2 < 1.0) {
            const double mt = sqrt(1.0 - st2);
            const double rs = (nw * co - mt) / (nw * co + mt), rp = (co - nw * mt) / (co + nw * mt);
            const double Rs = rs * rs, Rp = rp * rp;
            A = 0.5 * (Rs + Rp);
            q = (qc == 1) ? 0.5 * (Rp - Rs) : 0.5 * (Rs - Rp);
            d = rs * rp;
        }
        const double cb2 = cb * cb;
        const double tb2 = (1.0 - cb2) / fmax(cb2, 1e-8);
        const double brdf = ps_pref * exp(-tb2 * inv_s2v) * fac / (cb2 * cb2);
        rotated_mueller(A * brdf, q * brdf, d * brdf, c1, n1, c2, n2, R);
    }
};

// Water -> air Walter BTDF with Cox-Munk slopes and Sancer shadowing: the
// refraction half vector N = n p_water - p_air has |N|^2 = n^2 + 1 - 2 n cos(Psi),
// the facet tilt is (n mu_w - mu_a) / |N| and the in-water facet incidence
// (n - cos Psi) / |N|; zero beyond the critical angle.
struct TwaPair {
    double m1, m2, s1, s2, mm, ss, inv_s1, inv_s2, ps_pref, inv_s2v, pref, n;
    int qc;
    bool live;
    __device__ TwaPair(double mu_air, double mu_water, double s2v, double n_water, int q_conv)
    {
        m1 = mu_water;
        m2 = mu_air;
        live = m1 > 1e-9 && m2 > 1e-9;
        s1 = sqrt(fmax(0.0, 1.0 - m1 * m1));
        s2 = sqrt(fmax(0.0, 1.0 - m2 * m2));
        mm = m1 * m2;
        ss = s1 * s2;
        inv_s1 = 1.0 / fmax(1e-12, s1);
        inv_s2 = 1.0 / fmax(1e-12, s2);
        ps_pref = 1.0 / (OCRT_PI * s2v);
        inv_s2v = 1.0 / s2v;
        n = n_water;
        pref = live ? sancer_shadow(m1, m2, s2v) / (n * m1 * m2) : 0.0;
        qc = q_conv;
    }
    __device__ void eval(double cphi, double sphi, double *R) const
    {
#pragma unroll
        for (int k = 0; k < 9; ++k) R[k] = 0.0;
        if (!live) return;
        const double cP = fmin(1.0, fmax(-1.0, mm + ss * cphi));
        const double N2 = n * n + 1.0 - 2.0 * n * cP;
        if (N2 <= 1e-12) return;
        const double iN = 1.0 / sqrt(N2);
        const double cb = (n * m1 - m2) * iN, cw = (n - cP) * iN;
        if (cb <= 1e-6 || cw <= 1e-6) return;
        const double st2 = n * n * fmax(0.0, 1.0 - cw * cw);
        if (st2 >= 1.0) return;
        const double ma = sqrt(1.0 - st2);
        const double rs = (n * cw - ma) / (n * cw + ma), rp = (cw - n * ma) / (cw + n * ma);
        const double ts = 1.0 + rs, tp = n * (1.0 + rp);
        const double cb2 = cb * cb;
        const double tb2 = (1.0 - cb2) / fmax(cb2, 1e-8);
        const double common = (ma * ma) / N2 * pref * ps_pref * exp(-tb2 * inv_s2v) / (cb2 * cb2);
        const double fs = common * ts * ts, fp = common * tp * tp;
        const double rden = 1.0 / fmax(1e-12, sqrt(fmax(0.0, 1.0 - cP * cP)));
        double c1, n1, c2, n2;
        plane_rotation(s1, m1, (m2 - m1 * cP) * rden * inv_s1, s2 * sphi * rden, cphi, sphi, c1, n1);
        plane_rotation(s2, m2, (m1 - m2 * cP) * rden * inv_s2, s1 * sphi * rden, cphi, sphi, c2, n2);
        rotated_mueller(0.5 * (fs + fp), (qc == 1) ? 0.5 * (fp - fs) : 0.5 * (fs - fp), common * ts * tp, c1, n1,
                        c2, n2, R);
    }
};

// Air -> water microfacet transmission without shadowing (Walter et al. 2007,
// Beckmann density with the Cox-Munk variance).  Directions point away from
// the interface: i = (s_i, 0, mu_i) in air, o = (s_o cos phi, s_o sin phi,
// -mu_o) in water, h = -(i + n o) normalised into the upper hemisphere.  The
// Stokes rotations follow the photon path: both signed cosines negative and
// the azimuth shifted by pi (surface.c:1512-1524).
struct TawPair {
    double mi, mo, si, so, n, inv_s2v, dpref, inv_si, inv_so, ab;
    int qc;
    bool live;
    __device__ TawPair(double mu_water, double mu_air, double s2v, double n_water, int q_conv)
    {
        mi = mu_air;
        mo = mu_water;
        live = s2v > 1e-10 && mi > 1e-9 && mo > 1e-9;
        si = sqrt(fmax(0.0, 1.0 - mi * mi));
        so = sqrt(fmax(0.0, 1.0 - mo * mo));
        n = n_water;
        inv_s2v = 1.0 / s2v;
        dpref = 1.0 / (OCRT_PI * s2v);
        inv_si = 1.0 / fmax(1e-12, si);
        inv_so = 1.0 / fmax(1e-12, so);
        ab = mi * mo;   // (-mu_i)(-mu_o)
        qc = q_conv;
    }
    __device__ void eval(double cphi, double sphi, double *R) const
    {
#pragma unroll
        for (int k = 0; k < 9; ++k) R[k] = 0.0;
        if (!live) return;
        const double ox = so * cphi, oy = so * sphi;
        double hx = -(si + n * ox), hy = -(n * oy), hz = -(mi - n * mo);
        const double hn = sqrt(hx * hx + hy * hy + hz * hz);
        if (hn < 1e-12) return;
        const double ihn = (hz < 0.0 ? -1.0 : 1.0) / hn;
        hx *= ihn; hy *= ihn; hz *= ihn;
        const double ih = si * hx + mi * hz;
        const double oh = ox * hx + oy * hy - mo * hz;
        if (ih <= 0.0 || hz <= 1e-9) return;
        const double hz2 = hz * hz;
        const double D = exp(-((1.0 - hz2) / hz2) * inv_s2v) * dpref / (hz2 * hz2);
        const double dh = (ih + n * oh) * (ih + n * oh);
        if (dh < 1e-12) return;
        const double stt = sqrt(fmax(0.0, 1.0 - ih * ih)) / n;
        double Ts = 0.0, Tp = 0.0;
        if (stt < 1.0) {
            const double ctt = sqrt(fmax(0.0, 1.0 - stt * stt));
            const double rs = (ih - n * ctt) / (ih + n * ctt), rp = (n * ih - ctt) / (n * ih + ctt);
            Ts = 1.0 - rs * rs;
            Tp = 1.0 - rp * rp;
        }
        const double common = (fabs(ih) * fabs(oh)) / (mi * mo) * (n * n * D) / dh;
        const double fs = common * Ts, fp = common * Tp;
        const double cr = -cphi, sr = -sphi;
        const double cP = fmin(1.0, fmax(-1.0, ab + si * so * cr));
        const double rden = 1.0 / fmax(1e-12, sqrt(fmax(0.0, 1.0 - cP * cP)));
        double c1, n1, c2, n2;
        plane_rotation(si, -mi, (-mo + mi * cP) * rden * inv_si, so * sr * rden, cr, sr, c1, n1);
        plane_rotation(so, -mo, (-mi + mo * cP) * rden * inv_so, si * sr * rden, cr, sr, c2, n2);
        rotated_mueller(0.5 * (fs + fp), (qc == 1) ? 0.5 * (fp - fs) : 0.5 * (fs - fp),
                        common * sqrt(fmax(0.0, Ts * Tp)), c1, n1, c2, n2, R);
    }
};

// Air-side reflection, through the Cox-Munk evaluator of the atmosphere's
// boundary kernels.
struct RaaPair {
    CoxMunkPair cm;
    __device__ RaaPair(double mu_out, double mu_in, double s2v, double n_water, int q_conv)
        : cm(mu_out, mu_in, s2v, sancer_shadow(mu_out, mu_in, s2v), n_water, q_conv)
    {
    }
    __device__ void eval(double cphi, double sphi, double *R) const
    {
        double tb2;
        cm.eval(cphi, sphi, R, tb2);
    }
};

template <int KIND> struct PairOf;
template <> struct PairOf<OCRT_IFACE_R_AA> { using type = RaaPair; };
template <> struct PairOf<OCRT_IFACE_R_WW> { using type = RwwPair; };
template <> struct PairOf<OCRT_IFACE_T_WA> { using type = TwaPair; };
template <> struct PairOf<OCRT_IFACE_T_AW> { using type = TawPair; };

// W[p][m] = (cos(m phi_p), sin(m phi_p)), phi_p = (p + 1/2) 2 pi / n_phi:
// mode-fastest, so the threads of a wavefront (consecutive modes) read one
// contiguous run per azimuth.
__global__ void interface_weights(int n_phi, int n_rows, double2 *__restrict__ W)
{
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)n_rows * n_phi) return;
    const int p = (int)(e / n_rows), m = (int)(e % n_rows);
    const double phi = (p + 0.5) * (2.0 * OCRT_PI / (double)n_phi);
    W[e] = make_double2(cos(m * phi), sin(m * phi));
}

// One workgroup per node pair (blockIdx.x = o * n_i + i); output
// out[m][o][i][9] for m = 0..m_max.  Thread m projects onto mode m, summing
// the azimuths in their natural order, so every mode's bits are the same
// whatever m_max the launch covers.
template <int KIND>
__global__ __launch_bounds__(kIfThreads) void interface_fourier(const double *__restrict__ mu_o, int n_o,
                                                                const double *__restrict__ mu_i, int n_i,
                                                                int m_max, int n_phi, double s2v, double nw,
                                                                int qc, const double2 *__restrict__ W, int w_rows,
                                                                double *__restrict__ out)
{
    __shared__ double tab[9][kIfThreads];
    const int pair = blockIdx.x, t = threadIdx.x;
    const int o = pair / n_i, i = pair % n_i;
    const typename PairOf<KIND>::type pr(mu_o[o], mu_i[i], s2v, nw, qc);
    const bool owner = t <= m_max;
    double acc[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) acc[k] = 0.0;
    for (int p0 = 0; p0 < n_phi; p0 += kIfThreads) {
        const int p = p0 + t;
        double R[9];
        if (p < n_phi) {
            // cos and sin of the azimuth itself are the m = 1 entries of W
            const double2 cs = W[(long long)p * w_rows + 1];
            pr.eval(cs.x, cs.y, R);
        } else {
#pragma unroll
            for (int k = 0; k < 9; ++k) R[k] = 0.0;
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) tab[k][t] = R[k];
        __syncthreads();
        if (owner) {
            const int len = min(kIfThreads, n_phi - p0);
            const double2 *wt = W + (long long)p0 * w_rows + t;
            for (int q = 0; q < len; ++q) {
                const double2 w = wt[(long long)q * w_rows];
                const double c = w.x, s = w.y;
                acc[0] = fma(tab[0][q], c, acc[0]);
                acc[1] = fma(tab[1][q], c, acc[1]);
                acc[2] = fma(tab[2][q], s, acc[2]);
                acc[3] = fma(tab[3][q], c, acc[3]);
                acc[4] = fma(tab[4][q], c, acc[4]);
                acc[5] = fma(tab[5][q], s, acc[5]);
                acc[6] = fma(tab[6][q], s, acc[6]);
                acc[7] = fma(tab[7][q], s, acc[7]);
                acc[8] = fma(tab[8][q], c, acc[8]);
            }
        }
        __syncthreads();
    }
    if (!owner) return;
    const double dphi = 2.0 * OCRT_PI / (double)n_phi;
    const double nrm = (t == 0) ? 1.0 / (2.0 * OCRT_PI) : 1.0 / OCRT_PI;
    double *dst = out + (long long)t * n_o * n_i * 9 + (long long)pair * 9;
#pragma unroll
    for (int kl = 0; kl < 9; ++kl) {
        double v = acc[kl] * (dphi * nrm);
        // incoming-U column in the sine basis of the reference's Fourier fields
        if (KIND != OCRT_IFACE_R_WW && t > 0 && (kl == 2 || kl == 5)) v = -v;
        dst[kl] = v;
    }
}

// Facet term of the direct integrals at local incidence cosine c: air -> water
// transmittance 1 - R (T_aw) or the water -> air reflectance, 1 beyond the
// critical angle (R_ww).
template <int KIND> __device__ inline double direct_facet(double c, double nw)
{
    if (KIND == OCRT_IFACE_T_AW) {
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

// Slope-integrated direct term for one cosine per workgroup: trapezoid rule
// on 101 x 101 facet slopes over +-6 sigma, one slope row per thread.
template <int KIND>
__global__ __launch_bounds__(kDirectThreads) void interface_direct(const double *__restrict__ mu, double s2v,
                                                                   double nw, double *__restrict__ out)
{
    __shared__ double red[kDirectThreads / 64];
    const double m = mu[blockIdx.x];
    const int t = threadIdx.x;
    if (m <= 1e-9 || s2v <= 1e-10) {
        if (t == 0) out[blockIdx.x] = m <= 1e-9 ? (KIND == OCRT_IFACE_T_AW ? 0.0 : 1.0) : direct_facet<KIND>(m, nw);
        return;
    }
    constexpr int N = 101;
    const double sn = sqrt(fmax(0.0, 1.0 - m * m));
    const double L = 6.0 * sqrt(s2v), dz = 2.0 * L / (double)(N - 1);
    const double dens = 1.0 / (OCRT_PI * s2v);
    double acc = 0.0;
    if (t < N) {
        const double zx = -L + t * dz;
        const double wx = (t == 0 || t == N - 1) ? 0.5 : 1.0;
        const double proj = m + zx * sn;
        if (proj > 0.0)
            for (int j = 0; j < N; ++j) {
                const double zy = -L + j * dz;
                const double wy = (j == 0 || j == N - 1) ? 0.5 : 1.0;
                const double r2 = zx * zx + zy * zy;
                acc += wx * wy * (dens * exp(-r2 / s2v)) * proj * direct_facet<KIND>(proj / sqrt(1.0 + r2), nw);
            }
    }
    acc = wave_sum(acc);
    if ((t & 63) == 0) red[t >> 6] = acc;
    __syncthreads();
    if (t == 0) {
        double s = 0.0;
        for (int w = 0; w < kDirectThreads / 64; ++w) s += red[w];
        const double v = s * dz * dz / m;
        out[blockIdx.x] = v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
    }
}

double host_slope_variance(double ws, int sigma_type)
{
    if (sigma_type == 1) return 0.003 + 0.00512 * std::fmax(0.01, ws);
    const double s = 0.0731 * std::sqrt(std::fmax(0.0, ws));
    return std::fmax(s * s, 1e-10);
}

bool cosines_ok(const double *mu, int n)
{
    for (int k = 0; k < n; ++k)
        if (!(mu[k] > 0.0 && mu[k] <= 1.0)) return false;
    return true;
}

// Device buffers of one call, released on every exit path.
struct DevBufs {
    std::vector<void *> p;
    ~DevBufs()
    {
        for (void *q : p) (void)hipFree(q);
    }
    template <typename T> hipError_t get(T **ptr, size_t bytes)
    {
        void *q = nullptr;
        const hipError_t e = hipMalloc(&q, bytes ? bytes : 256);
        if (e == hipSuccess) p.push_back(q);
        *ptr = (T *)q;
        return e;
    }
};

#define IFCHK(x)                                                                          \
    do {                                                                                  \
        const hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                           \
            fprintf(stderr, "ocrt_amd: %s failed: %s\n", #x, hipGetErrorString(e_));      \
            return OCRT_ERR_DEVICE;                                                       \
        }                                                                                 \
    } while (0)

}  // namespace
}  // namespace ocrt

using namespace ocrt;

namespace ocrt {

// Rows of the azimuth weight table for modes 0..m_max: the m = 1 entries
// carry cos/sin of the azimuth itself, so at least two rows.
int interface_weight_rows(int m_max) { return m_max > 1 ? m_max + 1 : 2; }

// Fill the azimuth weight table W[n_phi][interface_weight_rows(m_max)]
// (caller-owned device memory) on st.
hipError_t interface_weights_launch(int n_phi, int m_max, double2 *d_w, hipStream_t st)
{
    const int w_rows = interface_weight_rows(m_max);
    const long long n_w = (long long)w_rows * n_phi;
    interface_weights<<<(unsigned)((n_w + 255) / 256), 256, 0, st>>>(n_phi, w_rows, d_w);
    return hipGetLastError();
}

// Launch the Fourier projection of one interface matrix over every node pair
// against a weight table made by interface_weights_launch for the same m_max
// and n_phi.  Asynchronous: no allocation, no synchronisation.
hipError_t interface_fourier_launch(int kind, const double *d_mo, int n_o, const double *d_mi, int n_i, int m_max,
                                    int n_phi, double s2v, double n_water, int q_convention, const double2 *d_w,
                                    double *d_out, hipStream_t st)
{
    const int w_rows = interface_weight_rows(m_max);
    const unsigned grid = (unsigned)((long long)n_o * n_i);
    switch (kind) {
    case OCRT_IFACE_R_AA:
        interface_fourier<OCRT_IFACE_R_AA><<<grid, kIfThreads, 0, st>>>(d_mo, n_o, d_mi, n_i, m_max, n_phi, s2v,
                                                                         n_water, q_convention, d_w, w_rows, d_out);
        break;
    case OCRT_IFACE_R_WW:
        interface_fourier<OCRT_IFACE_R_WW><<<grid, kIfThreads, 0, st>>>(d_mo, n_o, d_mi, n_i, m_max, n_phi, s2v,
                                                                         n_water, q_convention, d_w, w_rows, d_out);
        break;
    case OCRT_IFACE_T_WA:
        interface_fourier<OCRT_IFACE_T_WA><<<grid, kIfThreads, 0, st>>>(d_mo, n_o, d_mi, n_i, m_max, n_phi, s2v,
                                                                         n_water, q_convention, d_w, w_rows, d_out);
        break;
    default:
        interface_fourier<OCRT_IFACE_T_AW><<<grid, kIfThreads, 0, st>>>(d_mo, n_o, d_mi, n_i, m_max, n_phi, s2v,
                                                                         n_water, q_convention, d_w, w_rows, d_out);
        break;
    }
    return hipGetLastError();
}

int interface_fourier_device(int kind, const double *d_mo, int n_o, const double *d_mi, int n_i, int m_max,
                             int n_phi, double wind, int sigma_type, double n_water, int q_convention,
                             double *d_out, hipStream_t st)
{
    const double s2v = host_slope_variance(wind, sigma_type);
    DevBufs B;
    double2 *d_w;
    IFCHK(B.get(&d_w, sizeof(double2) * (size_t)interface_weight_rows(m_max) * n_phi));
    IFCHK(interface_weights_launch(n_phi, m_max, d_w, st));
    IFCHK(interface_fourier_launch(kind, d_mo, n_o, d_mi, n_i, m_max, n_phi, s2v, n_water, q_convention, d_w, d_out,
                                   st));
    // the weight table is released when this returns: finish with it first
    IFCHK(hipStreamSynchronize(st));
    return OCRT_OK;
}

}  // namespace ocrt

extern "C" int ocrt_interface_fourier(int kind, const double *mu_o, int n_o, const double *mu_i, int n_i, int m_max,
                                      int n_phi, double wind, int sigma_type, double n_water, int q_convention,
                                      double *out, void *hip_stream)
{
    if (kind < OCRT_IFACE_R_AA || kind > OCRT_IFACE_T_AW || !mu_o || !mu_i || !out) return OCRT_ERR_ARGS;
    if (n_o < 1 || n_i < 1 || (long long)n_o * n_i > kIfMaxPairs) return OCRT_ERR_ARGS;
    if (m_max < 0 || m_max + 1 > kIfMaxModes || n_phi < 4 || n_phi > OCRT_IFACE_MAX_PHI) return OCRT_ERR_ARGS;
    if (!std::isfinite(wind) || !(n_water > 0.0) || !std::isfinite(n_water)) return OCRT_ERR_ARGS;
    if (!cosines_ok(mu_o, n_o) || !cosines_ok(mu_i, n_i)) return OCRT_ERR_ARGS;
    if (kind == OCRT_IFACE_T_AW && host_slope_variance(wind, sigma_type) <= 1e-10) return OCRT_ERR_FLAT;
    hipStream_t st = (hipStream_t)hip_stream;
    const size_t n_out = (size_t)(m_max + 1) * n_o * n_i * 9;
    DevBufs B;
    double *d_mo, *d_mi, *d_out;
    IFCHK(B.get(&d_mo, sizeof(double) * n_o));
    IFCHK(B.get(&d_mi, sizeof(double) * n_i));
    IFCHK(B.get(&d_out, sizeof(double) * n_out));
    IFCHK(hipMemcpyAsync(d_mo, mu_o, sizeof(double) * n_o, hipMemcpyHostToDevice, st));
    IFCHK(hipMemcpyAsync(d_mi, mu_i, sizeof(double) * n_i, hipMemcpyHostToDevice, st));
    const int rc = interface_fourier_device(kind, d_mo, n_o, d_mi, n_i, m_max, n_phi, wind, sigma_type, n_water,
                                            q_convention, d_out, st);
    if (rc != OCRT_OK) return rc;
    IFCHK(hipMemcpyAsync(out, d_out, sizeof(double) * n_out, hipMemcpyDeviceToHost, st));
    IFCHK(hipStreamSynchronize(st));
    return OCRT_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free form of ocrt_interface_fourier.  Its one scratch buffer, the
// azimuth weight table, lives in a file-scope device array: 2^18 (cos, sin)
// pairs, 4 MiB on every GPU the library runs on, held until the process ends.
// Host code never passes the array's name to a kernel (on the host it is a
// stand-in, not a device address); it uses the address hipGetSymbolAddress
// returns on each device, looked up once per device at the first call there.
namespace ocrt {
namespace {
constexpr long long kIfScratchPairs = 1LL << 18;
constexpr int kScratchDevices = 64;
__device__ __align__(256) double2 scratch[kIfScratchPairs];
std::mutex g_scratch_mu;
double2 *g_scratch_addr[kScratchDevices];

// The scratch table's device address on the current device (nullptr and a
// message on failure).
double2 *scratch_table()
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kScratchDevices) {
        fprintf(stderr, "ocrt_amd: no usable current device for the scratch table\n");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            fprintf(stderr, "ocrt_amd: hipGetSymbolAddress(scratch) failed on device %d: %s\n", dev,
                    hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_addr[dev] = (double2 *)p;
    }
    return g_scratch_addr[dev];
}
}  // namespace
}  // namespace ocrt

extern "C" int ocrt_interface_fourier_noalloc(int kind, const double *d_mu_o, int n_o, const double *d_mu_i,
                                              int n_i, int m_max, int n_phi, double wind, int sigma_type,
                                              double n_water, int q_convention, double *d_out, void *hip_stream)
{
    if (kind < OCRT_IFACE_R_AA || kind > OCRT_IFACE_T_AW || !d_mu_o || !d_mu_i || !d_out) return OCRT_ERR_ARGS;
    if (n_o < 1 || n_i < 1 || (long long)n_o * n_i > kIfMaxPairs) return OCRT_ERR_ARGS;
    if (m_max < 0 || m_max + 1 > kIfMaxModes || n_phi < 4 || n_phi > OCRT_IFACE_MAX_PHI) return OCRT_ERR_ARGS;
    if (!std::isfinite(wind) || !(n_water > 0.0) || !std::isfinite(n_water)) return OCRT_ERR_ARGS;
    if ((long long)interface_weight_rows(m_max) * n_phi > kIfScratchPairs) {
        fprintf(stderr,
                "ocrt_interface_fourier_noalloc: (m_max + 1) x n_phi = %lld exceeds the %lld-entry scratch table "
                "(e.g. m_max 255 with n_phi 1024)\n",
                (long long)interface_weight_rows(m_max) * n_phi, kIfScratchPairs);
        return OCRT_ERR_ARGS;
    }
    const double s2v = host_slope_variance(wind, sigma_type);
    if (kind == OCRT_IFACE_T_AW && s2v <= 1e-10) return OCRT_ERR_FLAT;
    double2 *w = scratch_table();
    if (!w) return OCRT_ERR_DEVICE;
    hipStream_t st = (hipStream_t)hip_stream;
    // every entry of the table is rewritten before the projection reads it
    IFCHK(interface_weights_launch(n_phi, m_max, w, st));
    IFCHK(interface_fourier_launch(kind, d_mu_o, n_o, d_mu_i, n_i, m_max, n_phi, s2v, n_water, q_convention, w,
                                   d_out, st));
    return OCRT_OK;
}

extern "C" int ocrt_interface_direct(int kind, const double *mu, int n, double wind, int sigma_type, double n_water,
                                     double *out, void *hip_stream)
{
    if ((kind != OCRT_IFACE_T_AW && kind != OCRT_IFACE_R_WW) || !mu || !out || n < 1 || n >= (1 << 24))
        return OCRT_ERR_ARGS;
    if (!std::isfinite(wind) || !(n_water > 0.0) || !std::isfinite(n_water)) return OCRT_ERR_ARGS;
    for (int k = 0; k < n; ++k)
        if (!std::isfinite(mu[k]) || mu[k] > 1.0) return OCRT_ERR_ARGS;
    const double s2v = host_slope_variance(wind, sigma_type);
    hipStream_t st = (hipStream_t)hip_stream;
    DevBufs B;
    double *d_mu, *d_out;
    IFCHK(B.get(&d_mu, sizeof(double) * n));
    IFCHK(B.get(&d_out, sizeof(double) * n));
    IFCHK(hipMemcpyAsync(d_mu, mu, sizeof(double) * n, hipMemcpyHostToDevice, st));
    if (kind == OCRT_IFACE_T_AW)
        interface_direct<OCRT_IFACE_T_AW><<<n, kDirectThreads, 0, st>>>(d_mu, s2v, n_water, d_out);
    else
        interface_direct<OCRT_IFACE_R_WW><<<n, kDirectThreads, 0, st>>>(d_mu, s2v, n_water, d_out);
    IFCHK(hipGetLastError());
    IFCHK(hipMemcpyAsync(out, d_out, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    IFCHK(hipStreamSynchronize(st));
    return OCRT_OK;
}

// ---------------------------------------------------------------------------
// The reference's per-mode surface functions (shared/surface.h:40-61, 222-281),
// same names, arguments and return codes.  The first call for a node set and
// surface computes every mode 0..max(m, build cap) on the GPU in one launch
// and keeps them in a per-thread cache; later modes of the same key are
// copied from it, as the reference's all-mode cache does (surface.c:1263-1380).
namespace {

constexpr int kFkcModes = 64;   // default build cap, the reference's OCRT_FKC_MMAX
thread_local int g_build_cap = kFkcModes;

struct ModeCache {
    bool valid = false;
    int n_o = 0, n_i = 0, n_phi = 0, st = 0, qc = 0, built = -1;
    double ws = 0.0, nw = 0.0;
    std::vector<double> mo, mi, coef;
};
thread_local ModeCache g_cache[4];

int serve_mode(int kind, const double *mu_o, int n_o, const double *mu_i, int n_i, int m, int n_phi, double ws,
               int st, double nw, int qc, double *R_m)
{
    if (!mu_o || !mu_i || !R_m) return -1;
    if (n_o < 1 || n_i < 1 || m < 0 || n_phi < 4) return -1;
    if (kind == OCRT_IFACE_T_AW && host_slope_variance(ws, st) <= 1e-10) return -2;
    ModeCache &C = g_cache[kind];
    const size_t slab = (size_t)n_o * n_i * 9;
    const bool hit = C.valid && C.n_o == n_o && C.n_i == n_i && C.n_phi == n_phi && C.st == st && C.qc == qc &&
                     C.ws == ws && C.nw == nw && m <= C.built &&
                     std::memcmp(C.mo.data(), mu_o, sizeof(double) * n_o) == 0 &&
                     std::memcmp(C.mi.data(), mu_i, sizeof(double) * n_i) == 0;
    if (!hit) {
        const int top = std::min(kIfMaxModes - 1, std::max(m, g_build_cap));
        if (m > top) return -1;
        C.valid = false;
        C.coef.assign((size_t)(top + 1) * slab, 0.0);
        const int rc = ocrt_interface_fourier(kind, mu_o, n_o, mu_i, n_i, top, n_phi, ws, st, nw, qc,
                                              C.coef.data(), nullptr);
        if (rc != OCRT_OK) return rc == OCRT_ERR_ARGS ? -1 : -3;
        C.mo.assign(mu_o, mu_o + n_o);
        C.mi.assign(mu_i, mu_i + n_i);
        C.n_o = n_o; C.n_i = n_i; C.n_phi = n_phi; C.st = st; C.qc = qc; C.ws = ws; C.nw = nw;
        C.built = top;
        C.valid = true;
    }
    std::memcpy(R_m, C.coef.data() + (size_t)m * slab, sizeof(double) * slab);
    return 0;
}

}  // namespace

extern "C" void surface_fkc_set_build_mmax(int m_max)
{
    g_build_cap = (m_max >= 0 && m_max <= kFkcModes) ? m_max : kFkcModes;
}

extern "C" int surface_coxmunk_fourier_kernel(const double *mu_o, int n_o, const double *mu_i, int n_i, int m,
                                              int n_phi_quad, double ws, int sigma_type, double n_water,
                                              int q_convention, double *R_m)
{
    return serve_mode(OCRT_IFACE_R_AA, mu_o, n_o, mu_i, n_i, m, n_phi_quad, ws, sigma_type, n_water, q_convention,
                      R_m);
}

extern "C" int surface_R_ww_coxmunk_fourier_kernel(const double *mu_o, int n_o, const double *mu_i, int n_i, int m,
                                                   int n_phi_quad, double ws, int sigma_type, double n_water,
                                                   int q_convention, double *R_m)
{
    return serve_mode(OCRT_IFACE_R_WW, mu_o, n_o, mu_i, n_i, m, n_phi_quad, ws, sigma_type, n_water, q_convention,
                      R_m);
}

extern "C" int surface_T_wa_coxmunk_fourier_kernel(const double *mu_o, int n_o, const double *mu_i, int n_i, int m,
                                                   int n_phi_quad, double ws, int sigma_type, double n_water,
                                                   int q_convention, double *T_m)
{
    return serve_mode(OCRT_IFACE_T_WA, mu_o, n_o, mu_i, n_i, m, n_phi_quad, ws, sigma_type, n_water, q_convention,
                      T_m);
}

extern "C" int surface_T_aw_coxmunk_fourier_kernel(const double *mu_o, int n_o, const double *mu_i, int n_i, int m,
                                                   int n_phi_quad, double ws, int sigma_type, double n_water,
                                                   int q_convention, double *T_m)
{
    return serve_mode(OCRT_IFACE_T_AW, mu_o, n_o, mu_i, n_i, m, n_phi_quad, ws, sigma_type, n_water, q_convention,
                      T_m);
}

extern "C" double surface_T_aw_coxmunk_direct(double mu_sun_air, double n_water, double wind_speed, int sigma_type)
{
    double v = 0.0;
    if (ocrt_interface_direct(OCRT_IFACE_T_AW, &mu_sun_air, 1, wind_speed, sigma_type, n_water, &v, nullptr) != 0)
        return NAN;
    return v;
}

extern "C" double surface_R_ww_coxmunk_direct(double mu_up_water, double n_water, double wind_speed, int sigma_type)
{
    double v = 0.0;
    if (ocrt_interface_direct(OCRT_IFACE_R_WW, &mu_up_water, 1, wind_speed, sigma_type, n_water, &v, nullptr) != 0)
        return NAN;
    return v;
}

// ---------------------------------------------------------------------------
// Atmosphere -> water coupling of the coupled ocean: the atmosphere's BOA
// downward Fourier field mapped onto the in-water nodes
// (code/OCRT_C/src/rt_air_water_coupling.c:244-478).
//   rough surface with air weights: S_w^m(i) = C_m sum_j mu_j w_j T_aw^m[i][j] S_a^m(j),
//     C_0 = 2 pi, C_m = pi, T_aw from 128 azimuths (:268-315);
//   otherwise the flat map: Snell refraction to mu_air, the flat radiance
//     transmission Mueller and linear interpolation of S_a in mu_air (zero in
//     the total-reflection cone, :317-396); with wind > 0 the m = 0 intensity
//     is then replaced by the Cox-Munk BTDF transmission of the m = 0
//     skylight on a 192-node air rule with 96 azimuths (:398-476).
namespace ocrt {
namespace {

constexpr int kTawPhi = 128;      // azimuths of the rough transmission kernel
constexpr int kFineAir = 192;     // air nodes of the m = 0 BTDF integral
constexpr int kFineAzimuth = 96;  // its azimuths, dphi = 2 pi p / 96

// Linear interpolation on an ascending, duplicate-free grid with linear
// extrapolation from the end pairs (rt_air_water_coupling.c:102-114).
__device__ inline double interp_ascending(const double *ms, const double *v, const int *src, int w, double x)
{
    if (w < 2) return (w == 1) ? v[src[0]] : 0.0;
    int hi = 1;
    while (hi < w && ms[hi] < x) ++hi;
    int lo = hi - 1;
    if (hi >= w) { hi = w - 1; lo = w - 2; }
    const double t = (x - ms[lo]) / (ms[hi] - ms[lo]);
    return (1.0 - t) * v[src[lo]] + t * v[src[hi]];
}

// Rough path: one thread per (mode, water node).  field [3][m_max + 1][n_a].
__global__ void couple_rough(const double *__restrict__ T, const double *__restrict__ field,
                             const double *__restrict__ mu_a, const double *__restrict__ w_a, int n_w, int n_a,
                             int m_max, double *__restrict__ out)
{
    // instance blockIdx.y of a batch sharing nodes and surface
    field += (long long)blockIdx.y * 3 * (m_max + 1) * n_a;
    out += (long long)blockIdx.y * 3 * (m_max + 1) * n_w;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (m_max + 1) * n_w) return;
    const int m = e / n_w, iw = e % n_w;
    const double az = (m == 0) ? 2.0 * OCRT_PI : OCRT_PI;
    const long long M = m_max + 1;
    const double *Ia = field + (long long)m * n_a, *Qa = Ia + M * n_a, *Ua = Qa + M * n_a;
    double sI = 0.0, sQ = 0.0, sU = 0.0;
    for (int j = 0; j < n_a; ++j) {
        const double wj = az * mu_a[j] * w_a[j];
        if (wj == 0.0) continue;
        const double *t = T + (((long long)m * n_w + iw) * n_a + j) * 9;
        sI += wj * (t[0] * Ia[j] + t[1] * Qa[j] + t[2] * Ua[j]);
        sQ += wj * (t[3] * Ia[j] + t[4] * Qa[j] + t[5] * Ua[j]);
        sU += wj * (t[6] * Ia[j] + t[7] * Qa[j] + t[8] * Ua[j]);
    }
    out[(0 * M + m) * n_w + iw] = sI;
    out[(1 * M + m) * n_w + iw] = sQ;
    out[(2 * M + m) * n_w + iw] = sU;
}

// Flat path: one thread per (mode, water node).  ms/src: the air nodes
// sorted ascending with duplicates (within 1e-12) dropped, and the source
// index of each.
__global__ void couple_flat(const double *__restrict__ field, const double *__restrict__ ms,
                            const int *__restrict__ src, int w, const double *__restrict__ mu_w, int n_w, int n_a,
                            int m_max, double nw, int qc, double *__restrict__ out)
{
    field += (long long)blockIdx.y * 3 * (m_max + 1) * n_a;
    out += (long long)blockIdx.y * 3 * (m_max + 1) * n_w;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (m_max + 1) * n_w) return;
    const int m = e / n_w, iw = e % n_w;
    const long long M = m_max + 1;
    double r[3] = {0.0, 0.0, 0.0};
    const double mw = mu_w[iw];
    const double s2a = nw * nw * (1.0 - mw * mw);
    if (mw > 0.0 && mw <= 1.0 && s2a < 1.0) {
        const double ma = sqrt(1.0 - s2a);
        // flat radiance transmission air (1) -> water (n) at incidence ma
        // (surface.c:431-460): t_s = 1 + r_s, t_p = (1 / n)(1 + r_p), n^2 law
        const double st2 = (1.0 / nw) * (1.0 / nw) * fmax(0.0, 1.0 - ma * ma);
        const double mt = sqrt(1.0 - st2);
        const double rs = (ma - nw * mt) / (ma + nw * mt), rp = (nw * ma - mt) / (nw * ma + mt);
        const double ts = 1.0 + rs, tp = (1.0 / nw) * (1.0 + rp);
        const double rad = (nw * nw) * ((nw * mt) / ma);
        const double Ts = rad * ts * ts, Tp = rad * tp * tp;
        const double A = 0.5 * (Ts + Tp), q = (qc == 1) ? 0.5 * (Tp - Ts) : 0.5 * (Ts - Tp), d = rad * ts * tp;
        const double *Ia = field + (long long)m * n_a, *Qa = Ia + M * n_a, *Ua = Qa + M * n_a;
        const double I = interp_ascending(ms, Ia, src, w, ma), Q = interp_ascending(ms, Qa, src, w, ma),
                     U = interp_ascending(ms, Ua, src, w, ma);
        r[0] = A * I + q * Q;
        r[1] = q * I + A * Q;
        r[2] = d * U;
    }
    for (int c = 0; c < 3; ++c) out[(c * M + m) * n_w + iw] = r[c];
}

// The m = 0 intensity under a rough surface: one workgroup per water node,
// one fine air node per thread; each thread sums its 96 azimuths in order.
__global__ __launch_bounds__(256) void couple_btdf_m0(const double *__restrict__ field, const double *__restrict__ ms,
                                                      const int *__restrict__ src, int w,
                                                      const double *__restrict__ mu_w, const double *__restrict__ fa_mu,
                                                      const double *__restrict__ fa_w, double s2v, double nw,
                                                      double *__restrict__ out, long long in_stride,
                                                      long long out_stride)
{
    __shared__ double red[4];
    field += (long long)blockIdx.y * in_stride;
    out += (long long)blockIdx.y * out_stride;
    const int iw = blockIdx.x, t = threadIdx.x;
    const double mw = mu_w[iw];
    if (!(mw > 1e-9 && mw <= 1.0)) return;   // keeps the flat value
    double term = 0.0;
    if (t < kFineAir) {
        const double mua = fa_mu[t];
        const double L0 = interp_ascending(ms, field, src, w, mua);
        const TawPair pr(mw, mua, s2v, nw, 1);
        const double dph = 2.0 * OCRT_PI / (double)kFineAzimuth;
        double K0 = 0.0;
        for (int p = 0; p < kFineAzimuth; ++p) {
            const double phi = dph * (double)p;
            double R[9];
            pr.eval(cos(phi), sin(phi), R);
            K0 += fmax(R[0], 0.0);
        }
        K0 *= dph;
        term = fa_w[t] * mua * K0 * L0;
    }
    term = wave_sum(term);
    if ((t & 63) == 0) red[t >> 6] = term;
    __syncthreads();
    if (t == 0) out[iw] = (red[0] + red[1]) + (red[2] + red[3]);
}

}  // namespace
}  // namespace ocrt

extern "C" int ocrt_couple_atm_to_water_batch(int n_inst, int m_max, int n_a, const double *mu_a, const double *w_a,
                                              const double *fields, double n_water, int q_convention,
                                              const double *mu_w, int n_w, double wind, int sigma_type, double *out,
                                              void *hip_stream)
{
    if (n_inst < 1 || n_inst > 65535 || !mu_a || !fields || !mu_w || !out) return OCRT_ERR_ARGS;
    if (!(n_water > 1.0) || n_a < 2 || n_w < 1 || m_max < 0 || m_max + 1 > kIfMaxModes) return OCRT_ERR_ARGS;
    if ((long long)n_w * n_a > kIfMaxPairs || !std::isfinite(wind)) return OCRT_ERR_ARGS;
    hipStream_t st = (hipStream_t)hip_stream;
    const long long M = m_max + 1;
    const size_t nf = (size_t)M * n_a, no = (size_t)M * n_w;
    DevBufs B;
    double *d_field, *d_ma, *d_mw, *d_out;
    IFCHK(B.get(&d_field, sizeof(double) * 3 * nf * n_inst));
    IFCHK(B.get(&d_ma, sizeof(double) * n_a));
    IFCHK(B.get(&d_mw, sizeof(double) * n_w));
    IFCHK(B.get(&d_out, sizeof(double) * 3 * no * n_inst));
    IFCHK(hipMemcpyAsync(d_field, fields, sizeof(double) * 3 * nf * n_inst, hipMemcpyHostToDevice, st));
    IFCHK(hipMemcpyAsync(d_ma, mu_a, sizeof(double) * n_a, hipMemcpyHostToDevice, st));
    IFCHK(hipMemcpyAsync(d_mw, mu_w, sizeof(double) * n_w, hipMemcpyHostToDevice, st));
    const double s2v = host_slope_variance(wind, sigma_type);
    const dim3 grid((unsigned)((no + 255) / 256), (unsigned)n_inst);
    if (w_a && wind > 0.0 && s2v > 1e-10) {
        // one set of transmission kernels serves every instance of the batch
        double *d_T, *d_wa;
        IFCHK(B.get(&d_T, sizeof(double) * (size_t)M * n_w * n_a * 9));
        IFCHK(B.get(&d_wa, sizeof(double) * n_a));
        IFCHK(hipMemcpyAsync(d_wa, w_a, sizeof(double) * n_a, hipMemcpyHostToDevice, st));
        const int rc = interface_fourier_device(OCRT_IFACE_T_AW, d_mw, n_w, d_ma, n_a, m_max, kTawPhi, wind,
                                                sigma_type, n_water, q_convention, d_T, st);
        if (rc != OCRT_OK) return rc;
        couple_rough<<<grid, 256, 0, st>>>(d_T, d_field, d_ma, d_wa, n_w, n_a, m_max, d_out);
        IFCHK(hipGetLastError());
    } else {
        // the air nodes sorted ascending, duplicates within 1e-12 dropped
        // (rt_air_water_coupling.c:149-209)
        std::vector<int> ord(n_a);
        for (int a = 0; a < n_a; ++a) ord[a] = a;
        for (int a = 1; a < n_a; ++a) {
            const int k = ord[a];
            int b = a - 1;
            while (b >= 0 && mu_a[ord[b]] > mu_a[k]) { ord[b + 1] = ord[b]; --b; }
            ord[b + 1] = k;
        }
        std::vector<double> ms;
        std::vector<int> src;
        for (int a = 0; a < n_a; ++a) {
            const double m = mu_a[ord[a]];
            if (!ms.empty() && std::fabs(m - ms.back()) < 1e-12) continue;
            ms.push_back(m);
            src.push_back(ord[a]);
        }
        double *d_ms;
        int *d_src;
        IFCHK(B.get(&d_ms, sizeof(double) * ms.size()));
        IFCHK(B.get(&d_src, sizeof(int) * src.size()));
        IFCHK(hipMemcpyAsync(d_ms, ms.data(), sizeof(double) * ms.size(), hipMemcpyHostToDevice, st));
        IFCHK(hipMemcpyAsync(d_src, src.data(), sizeof(int) * src.size(), hipMemcpyHostToDevice, st));
        couple_flat<<<grid, 256, 0, st>>>(d_field, d_ms, d_src, (int)ms.size(), d_mw, n_w, n_a, m_max, n_water,
                                          q_convention, d_out);
        IFCHK(hipGetLastError());
        if (wind > 0.0) {
            std::vector<double> fa(2 * kFineAir);
            if (gauss_legendre_pos(kFineAir, fa.data(), fa.data() + kFineAir) != 0) return OCRT_ERR_ARGS;
            double *d_fa;
            IFCHK(B.get(&d_fa, sizeof(double) * fa.size()));
            IFCHK(hipMemcpyAsync(d_fa, fa.data(), sizeof(double) * fa.size(), hipMemcpyHostToDevice, st));
            couple_btdf_m0<<<dim3((unsigned)n_w, (unsigned)n_inst), 256, 0, st>>>(
                d_field, d_ms, d_src, (int)ms.size(), d_mw, d_fa, d_fa + kFineAir, s2v, n_water, d_out,
                (long long)(3 * nf), (long long)(3 * no));
            IFCHK(hipGetLastError());
        }
    }
    IFCHK(hipMemcpyAsync(out, d_out, sizeof(double) * 3 * no * n_inst, hipMemcpyDeviceToHost, st));
    IFCHK(hipStreamSynchronize(st));
    return OCRT_OK;
}

extern "C" int ocrt_couple_atm_to_water(int m_max, int n_a, const double *mu_a, const double *w_a, const double *I_a,
                                        const double *Q_a, const double *U_a, double n_water, int q_convention,
                                        const double *mu_w, int n_w, double wind, int sigma_type, double *I_w,
                                        double *Q_w, double *U_w, void *hip_stream)
{
    if (!mu_a || !I_a || !Q_a || !U_a || !mu_w || !I_w || !Q_w || !U_w || m_max < 0 || n_a < 1 || n_w < 1)
        return OCRT_ERR_ARGS;
    const size_t nf = (size_t)(m_max + 1) * n_a, no = (size_t)(m_max + 1) * n_w;
    std::vector<double> f(3 * nf), o(3 * no);
    std::copy(I_a, I_a + nf, f.begin());
    std::copy(Q_a, Q_a + nf, f.begin() + nf);
    std::copy(U_a, U_a + nf, f.begin() + 2 * nf);
    const int rc = ocrt_couple_atm_to_water_batch(1, m_max, n_a, mu_a, w_a, f.data(), n_water, q_convention, mu_w,
                                                  n_w, wind, sigma_type, o.data(), hip_stream);
    if (rc != OCRT_OK) return rc;
    std::copy(o.begin(), o.begin() + no, I_w);
    std::copy(o.begin() + no, o.begin() + 2 * no, Q_w);
    std::copy(o.begin() + 2 * no, o.end(), U_w);
    return OCRT_OK;
}

// The reference's coupling entry point and its output helpers, same names,
// structs and return codes (rt_air_water_coupling.h:26-68).
extern "C" int rt_aw_coupled_field_alloc(rt_aw_coupled_field_t *f, int m_max, int n_mu_water)
{
    if (!f || m_max < 0 || n_mu_water < 1) return -1;
    std::memset(f, 0, sizeof(*f));
    f->m_max = m_max;
    f->n_mu_water = n_mu_water;
    const size_t sz = (size_t)(m_max + 1) * (size_t)n_mu_water;
    f->mu_water_pos = (double *)std::calloc((size_t)n_mu_water, sizeof(double));
    f->I_inwater_per_m = (double *)std::calloc(sz, sizeof(double));
    f->Q_inwater_per_m = (double *)std::calloc(sz, sizeof(double));
    f->U_inwater_per_m = (double *)std::calloc(sz, sizeof(double));
    if (!f->mu_water_pos || !f->I_inwater_per_m || !f->Q_inwater_per_m || !f->U_inwater_per_m) {
        rt_aw_coupled_field_free(f);
        return -1;
    }
    return 0;
}

extern "C" void rt_aw_coupled_field_free(rt_aw_coupled_field_t *f)
{
    if (!f) return;
    std::free(f->mu_water_pos);
    std::free(f->I_inwater_per_m);
    std::free(f->Q_inwater_per_m);
    std::free(f->U_inwater_per_m);
    std::memset(f, 0, sizeof(*f));
}

extern "C" int rt_air_water_couple_atm_to_water(const rt_atm_boa_export_t *boa, double n_water, int q_convention,
                                                const double *mu_water_pos, int n_mu_water, const double *w_atm_pos,
                                                double wind_speed, int sigma_type, rt_aw_coupled_field_t *coupled)
{
    if (!boa || !mu_water_pos || !coupled) return -1;
    if (n_water <= 1.0 || boa->n_mu < 2) return -1;
    if (coupled->n_mu_water != n_mu_water || coupled->m_max != boa->m_max) return -1;
    if (!coupled->I_inwater_per_m || !coupled->Q_inwater_per_m || !coupled->U_inwater_per_m ||
        !coupled->mu_water_pos)
        return -1;
    const int rc = ocrt_couple_atm_to_water(boa->m_max, boa->n_mu, boa->mu_quad_pos, w_atm_pos, boa->I_per_m,
                                            boa->Q_per_m, boa->U_per_m, n_water, q_convention, mu_water_pos,
                                            n_mu_water, wind_speed, sigma_type, coupled->I_inwater_per_m,
                                            coupled->Q_inwater_per_m, coupled->U_inwater_per_m, nullptr);
    if (rc != OCRT_OK) return rc == OCRT_ERR_ARGS ? -1 : -3;
    std::memcpy(coupled->mu_water_pos, mu_water_pos, sizeof(double) * (size_t)n_mu_water);
    return 0;
}

// ---------------------------------------------------------------------------
// Water -> atmosphere coupling: the in-water upwelling Fourier field just
// below the surface mapped onto the air nodes (rt_air_water_coupling.c:598-780).
//   rough (wind > 0): per air node mu_a, the water -> air Cox-Munk Fourier
//     transmission (128 azimuths) over a composite water grid: 1024 trapezoid
//     points within 0.3 of the refracted cosine plus the 96-node Gauss rule
//     outside that window, contracted with the water field interpolated
//     linearly onto the grid, times 2 pi (m = 0) or pi;
//   flat: reverse Snell to mu_w, the flat water -> air radiance transmission
//     Mueller and the interpolated water field.
// The reference's multi-encounter re-escape term is opt-in there (an
// environment switch, off by default) and is not added here.
namespace ocrt {
namespace {

constexpr int kWinPoints = 1024;     // trapezoid window around the refracted cosine
constexpr int kGloPoints = 96;       // Gauss nodes on the rest of (0, 1)
constexpr double kWinHalf = 0.30;
constexpr int kGridPoints = kWinPoints + kGloPoints;

// Composite water grid of one air node: xi, wi [kGridPoints] (:640-660).
__global__ void water_grid(const double *__restrict__ mu_a, int n_a, double n2, const double *__restrict__ gl,
                           double *__restrict__ xi, double *__restrict__ wi)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_a * kGridPoints) return;
    const int ja = e / kGridPoints, i = e % kGridPoints;
    const double ma = mu_a[ja];
    const double mws = sqrt(fmax(0.0, 1.0 - (1.0 - ma * ma) / n2));
    double lo = mws - kWinHalf, hi = mws + kWinHalf;
    if (lo < 1e-6) lo = 1e-6;
    if (hi > 1.0 - 1e-9) hi = 1.0 - 1e-9;
    const double h = (hi - lo) / (double)(kWinPoints - 1);
    double x, w;
    if (i < kWinPoints) {
        x = lo + h * (double)i;
        w = ((i == 0 || i == kWinPoints - 1) ? 0.5 : 1.0) * h;
    } else {
        x = gl[i - kWinPoints];
        w = (x > lo && x < hi) ? 0.0 : gl[kGloPoints + i - kWinPoints];
    }
    xi[e] = x;
    wi[e] = w;
}

// Rough contraction of one air node: one workgroup per mode.  T [m][1][grid][9],
// water field [3][m_max + 1][n_w] on the sorted node map (ms, src, w).
__global__ __launch_bounds__(256) void couple_rough_up(const double *__restrict__ T, const double *__restrict__ xi,
                                                       const double *__restrict__ wi,
                                                       const double *__restrict__ field, const double *__restrict__ ms,
                                                       const int *__restrict__ src, int w, int n_w, int m_max, int ja,
                                                       int n_a, double *__restrict__ out)
{
    __shared__ double red[3][4];
    field += (long long)blockIdx.y * 3 * (m_max + 1) * n_w;
    out += (long long)blockIdx.y * 3 * (m_max + 1) * n_a;
    const int m = blockIdx.x, t = threadIdx.x;
    const long long M = m_max + 1;
    const double az = (m == 0) ? 2.0 * OCRT_PI : OCRT_PI;
    const double *Iw = field + (long long)m * n_w, *Qw = Iw + M * n_w, *Uw = Qw + M * n_w;
    double s[3] = {0.0, 0.0, 0.0};
    for (int i = t; i < kGridPoints; i += blockDim.x) {
        if (wi[i] == 0.0) continue;
        const double wj = az * xi[i] * wi[i];
        const double *tr = T + ((long long)m * kGridPoints + i) * 9;
        const double I = interp_ascending(ms, Iw, src, w, xi[i]), Q = interp_ascending(ms, Qw, src, w, xi[i]),
                     U = interp_ascending(ms, Uw, src, w, xi[i]);
        s[0] += wj * (tr[0] * I + tr[1] * Q + tr[2] * U);
        s[1] += wj * (tr[3] * I + tr[4] * Q + tr[5] * U);
        s[2] += wj * (tr[6] * I + tr[7] * Q + tr[8] * U);
    }
    for (int c = 0; c < 3; ++c) {
        const double v = wave_sum(s[c]);
        if ((t & 63) == 0) red[c][t >> 6] = v;
    }
    __syncthreads();
    if (t < 3) out[((long long)t * M + m) * n_a + ja] = (red[t][0] + red[t][1]) + (red[t][2] + red[t][3]);
}

// Flat map: one thread per (mode, air node).
__global__ void couple_flat_up(const double *__restrict__ field, const double *__restrict__ ms,
                               const int *__restrict__ src, int w, const double *__restrict__ mu_a, int n_a, int n_w,
                               int m_max, double nw, int qc, double *__restrict__ out)
{
    field += (long long)blockIdx.y * 3 * (m_max + 1) * n_w;
    out += (long long)blockIdx.y * 3 * (m_max + 1) * n_a;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (m_max + 1) * n_a) return;
    const int m = e / n_a, ja = e % n_a;
    const long long M = m_max + 1;
    const double ma = mu_a[ja];
    const double mw = sqrt(fmax(0.0, 1.0 - (1.0 - ma * ma) / (nw * nw)));
    // flat radiance transmission water (n) -> air (1) at incidence mw
    // (surface.c:431-460): t_s = 1 + r_s, t_p = n (1 + r_p), zero beyond the
    // critical angle
    double A = 0.0, q = 0.0, d = 0.0;
    const double st2 = nw * nw * fmax(0.0, 1.0 - mw * mw);
    if (st2 < 1.0) {
        const double mt = sqrt(1.0 - st2);
        const double rs = (nw * mw - mt) / (nw * mw + mt), rp = (mw - nw * mt) / (mw + nw * mt);
        const double ts = 1.0 + rs, tp = nw * (1.0 + rp);
        const double rad = (1.0 / (nw * nw)) * (mt / (nw * mw));
        const double Ts = rad * ts * ts, Tp = rad * tp * tp;
        A = 0.5 * (Ts + Tp);
        q = (qc == 1) ? 0.5 * (Tp - Ts) : 0.5 * (Ts - Tp);
        d = rad * ts * tp;
    }
    const double *Iw = field + (long long)m * n_w, *Qw = Iw + M * n_w, *Uw = Qw + M * n_w;
    const double I = interp_ascending(ms, Iw, src, w, mw), Q = interp_ascending(ms, Qw, src, w, mw),
                 U = interp_ascending(ms, Uw, src, w, mw);
    out[(0 * M + m) * n_a + ja] = A * I + q * Q;
    out[(1 * M + m) * n_a + ja] = q * I + A * Q;
    out[(2 * M + m) * n_a + ja] = d * U;
}

// Ascending, duplicate-free view of a node set (insertion sort, 1e-12 dedup;
// rt_air_water_coupling.c:149-209).
void sorted_nodes(const double *mu, int n, std::vector<double> &ms, std::vector<int> &src)
{
    std::vector<int> ord(n);
    for (int a = 0; a < n; ++a) ord[a] = a;
    for (int a = 1; a < n; ++a) {
        const int k = ord[a];
        int b = a - 1;
        while (b >= 0 && mu[ord[b]] > mu[k]) { ord[b + 1] = ord[b]; --b; }
        ord[b + 1] = k;
    }
    ms.clear();
    src.clear();
    for (int a = 0; a < n; ++a) {
        const double m = mu[ord[a]];
        if (!ms.empty() && std::fabs(m - ms.back()) < 1e-12) continue;
        ms.push_back(m);
        src.push_back(ord[a]);
    }
}

}  // namespace
}  // namespace ocrt

extern "C" int ocrt_couple_water_to_atm_batch(int n_inst, int m_max, int n_w, const double *mu_w,
                                              const double *fields, double n_water, int q_convention, double wind,
                                              int sigma_type, int n_a, const double *mu_a, double *out,
                                              void *hip_stream)
{
    if (n_inst < 1 || n_inst > 65535 || !fields || !mu_w || !mu_a || !out) return OCRT_ERR_ARGS;
    if (!(n_water > 1.0) || n_w < 2 || n_a < 1 || m_max < 0 || m_max + 1 > kIfMaxModes) return OCRT_ERR_ARGS;
    if (!std::isfinite(wind)) return OCRT_ERR_ARGS;
    hipStream_t st = (hipStream_t)hip_stream;
    const long long M = m_max + 1;
    const size_t nf = (size_t)M * n_w, no = (size_t)M * n_a;
    std::vector<double> ms;
    std::vector<int> src;
    sorted_nodes(mu_w, n_w, ms, src);
    DevBufs B;
    double *d_field, *d_ma, *d_out, *d_ms;
    int *d_src;
    IFCHK(B.get(&d_field, sizeof(double) * 3 * nf * n_inst));
    IFCHK(B.get(&d_ma, sizeof(double) * n_a));
    IFCHK(B.get(&d_out, sizeof(double) * 3 * no * n_inst));
    IFCHK(B.get(&d_ms, sizeof(double) * ms.size()));
    IFCHK(B.get(&d_src, sizeof(int) * src.size()));
    IFCHK(hipMemcpyAsync(d_field, fields, sizeof(double) * 3 * nf * n_inst, hipMemcpyHostToDevice, st));
    IFCHK(hipMemcpyAsync(d_ma, mu_a, sizeof(double) * n_a, hipMemcpyHostToDevice, st));
    IFCHK(hipMemcpyAsync(d_ms, ms.data(), sizeof(double) * ms.size(), hipMemcpyHostToDevice, st));
    IFCHK(hipMemcpyAsync(d_src, src.data(), sizeof(int) * src.size(), hipMemcpyHostToDevice, st));
    if (wind > 0.0) {
        std::vector<double> gl(2 * kGloPoints);
        if (gauss_legendre_pos(kGloPoints, gl.data(), gl.data() + kGloPoints) != 0) return OCRT_ERR_ARGS;
        double *d_gl, *d_xi, *d_wi, *d_T;
        IFCHK(B.get(&d_gl, sizeof(double) * gl.size()));
        IFCHK(B.get(&d_xi, sizeof(double) * (size_t)n_a * kGridPoints));
        IFCHK(B.get(&d_wi, sizeof(double) * (size_t)n_a * kGridPoints));
        IFCHK(B.get(&d_T, sizeof(double) * (size_t)M * kGridPoints * 9));
        IFCHK(hipMemcpyAsync(d_gl, gl.data(), sizeof(double) * gl.size(), hipMemcpyHostToDevice, st));
        IFCHK(hipMemsetAsync(d_out, 0, sizeof(double) * 3 * no * n_inst, st));
        const long long ng = (long long)n_a * kGridPoints;
        water_grid<<<(unsigned)((ng + 255) / 256), 256, 0, st>>>(d_ma, n_a, n_water * n_water, d_gl, d_xi, d_wi);
        IFCHK(hipGetLastError());
        // one azimuth weight table serves every air node, and one kernel row
        // per air node every instance of the batch; the launches below follow
        // each other on st, so d_T is reused without synchronising
        double2 *d_w;
        IFCHK(B.get(&d_w, sizeof(double2) * (size_t)interface_weight_rows(m_max) * kTawPhi));
        IFCHK(interface_weights_launch(kTawPhi, m_max, d_w, st));
        const double s2v = host_slope_variance(wind, sigma_type);
        for (int ja = 0; ja < n_a; ++ja) {
            if (!(mu_a[ja] > 0.0) || mu_a[ja] > 1.0) continue;   // outside (0, 1]: stays zero, as the reference
            const double *xi = d_xi + (size_t)ja * kGridPoints, *wi = d_wi + (size_t)ja * kGridPoints;
            IFCHK(interface_fourier_launch(OCRT_IFACE_T_WA, d_ma + ja, 1, xi, kGridPoints, m_max, kTawPhi, s2v,
                                           n_water, q_convention, d_w, d_T, st));
            couple_rough_up<<<dim3((unsigned)M, (unsigned)n_inst), 256, 0, st>>>(
                d_T, xi, wi, d_field, d_ms, d_src, (int)ms.size(), n_w, m_max, ja, n_a, d_out);
            IFCHK(hipGetLastError());
        }
    } else {
        couple_flat_up<<<dim3((unsigned)((no + 255) / 256), (unsigned)n_inst), 256, 0, st>>>(
            d_field, d_ms, d_src, (int)ms.size(), d_ma, n_a, n_w, m_max, n_water, q_convention, d_out);
        IFCHK(hipGetLastError());
    }
    IFCHK(hipMemcpyAsync(out, d_out, sizeof(double) * 3 * no * n_inst, hipMemcpyDeviceToHost, st));
    IFCHK(hipStreamSynchronize(st));
    return OCRT_OK;
}

extern "C" int ocrt_couple_water_to_atm(int m_max, int n_w, const double *mu_w, const double *I_w, const double *Q_w,
                                        const double *U_w, double n_water, int q_convention, double wind,
                                        int sigma_type, int n_a, const double *mu_a, double *I_a, double *Q_a,
                                        double *U_a, void *hip_stream)
{
    if (!I_w || !Q_w || !U_w || !mu_w || !mu_a || !I_a || !Q_a || !U_a || m_max < 0 || n_w < 1 || n_a < 1)
        return OCRT_ERR_ARGS;
    const size_t nf = (size_t)(m_max + 1) * n_w, no = (size_t)(m_max + 1) * n_a;
    std::vector<double> f(3 * nf), o(3 * no);
    std::copy(I_w, I_w + nf, f.begin());
    std::copy(Q_w, Q_w + nf, f.begin() + nf);
    std::copy(U_w, U_w + nf, f.begin() + 2 * nf);
    const int rc = ocrt_couple_water_to_atm_batch(1, m_max, n_w, mu_w, f.data(), n_water, q_convention, wind,
                                                  sigma_type, n_a, mu_a, o.data(), hip_stream);
    if (rc != OCRT_OK) return rc;
    std::copy(o.begin(), o.begin() + no, I_a);
    std::copy(o.begin() + no, o.begin() + 2 * no, Q_a);
    std::copy(o.begin() + 2 * no, o.end(), U_a);
    return OCRT_OK;
}

extern "C" int rt_air_water_couple_water_to_atm(const double *I_water_per_m, const double *Q_water_per_m,
                                                const double *U_water_per_m, const double *mu_water_pos,
                                                const double *w_water_pos, int n_mu_water, int m_max,
                                                double n_water, int q_convention, double wind_speed,
                                                int sigma_type, const double *mu_atm_pos, int n_mu_atm,
                                                const double *w_atm_pos, double *I_air_per_m,
                                                double *Q_air_per_m, double *U_air_per_m)
{
    (void)w_water_pos;   // weights enter only the reference's opt-in re-escape term
    (void)w_atm_pos;
    const int rc = ocrt_couple_water_to_atm(m_max, n_mu_water, mu_water_pos, I_water_per_m, Q_water_per_m,
                                            U_water_per_m, n_water, q_convention, wind_speed, sigma_type, n_mu_atm,
                                            mu_atm_pos, I_air_per_m, Q_air_per_m, U_air_per_m, nullptr);
    return rc == OCRT_OK ? 0 : (rc == OCRT_ERR_ARGS ? -1 : -3);
}
