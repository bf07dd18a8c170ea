// Gas absorption tables on the host: the reference's data formats
// (inputs/afgl_atm/afgl_<profile>.dat, inputs/xsec/xsec_<gas>.dat), its
// rt_absorption_t state (include/ocrt_rt.h), the ocrt_gas tables the plans
// use, and their device packing.  Host code only; the per-layer optical
// depths of a solve are computed on the GPU (setup.hip: medium_gas).
//
// Formats and rules follow code/OCRT_C/src/rt_absorption.c:
//   AFGL rows  "z P T n_total mr_h2o mr_o3 mr_o2 mr_co2 mr_no2 mr_ch4", '#'
//              comments; gas number density n_g = mr * 1e-6 * n_total; the
//              default column is the trapezoid of n_g over the levels
//              (rt_absorption.c:40-167)
//   xsec       "# n_layers = K" declares K layers: one row of wavelengths,
//              then K rows of sigma; without it, rows "wl sigma"; the grid
//              must be ascending and cover 330..1100 nm (:180-333)
//   sigma(z, wl) bilinear: linear in wl inside a layer row (NaN outside the
//              grid), linear in z between the bracketing layers, clamped to
//              the end layers (:344-415)
//   column above z: trapezoid of the linearly interpolated n_g (:425-455)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gas.h"
#include "ocrt_amd.h"
#include "ocrt_rt.h"

namespace {

const char *kGas[6] = {"h2o", "o3", "o2", "co2", "no2", "ch4"};
const char *kProfile[7] = {"usstd76", "tropical", "mlsumm", "mlwint", "sasumm", "sawint", "userdef"};

bool comment_or_blank(const char *line) { return line[0] == '#' || line[0] == '\n' || line[0] == '\0'; }

double *dup(const std::vector<double> &v)
{
    double *p = (double *)std::calloc(v.size() ? v.size() : 1, sizeof(double));
    if (p && !v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(double));
    return p;
}

int load_afgl(const char *dir, int profile, rt_afgl_atm_t *atm)
{
    std::memset(atm, 0, sizeof(*atm));
    if (profile < 0 || profile > 6) return -1;
    const std::string path = std::string(dir) + "/afgl_" + kProfile[profile] + ".dat";
    FILE *f = std::fopen(path.c_str(), "r");
    if (!f) {
        std::fprintf(stderr, "rt_afgl_load: cannot open %s\n", path.c_str());
        return -2;
    }
    std::vector<double> col[10];
    char line[1024];
    while (std::fgets(line, sizeof line, f)) {
        if (comment_or_blank(line)) continue;
        double v[10];
        if (std::sscanf(line, "%lf %lf %lf %lf %lf %lf %lf %lf %lf %lf", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5],
                        &v[6], &v[7], &v[8], &v[9]) != 10)
            continue;
        for (int c = 0; c < 10; ++c) col[c].push_back(v[c]);
    }
    std::fclose(f);
    const int n = (int)col[0].size();
    atm->n_levels = n;
    atm->z_km = dup(col[0]);
    atm->P_mbar = dup(col[1]);
    atm->T_K = dup(col[2]);
    atm->n_total_cm3 = dup(col[3]);
    for (int g = 0; g < 6; ++g) atm->mr[g] = dup(col[4 + g]);
    for (int g = 0; g < 6; ++g) {
        double s = 0.0;
        for (int i = 0; i + 1 < n; ++i) {
            const double a = atm->mr[g][i] * 1e-6 * atm->n_total_cm3[i];
            const double b = atm->mr[g][i + 1] * 1e-6 * atm->n_total_cm3[i + 1];
            s += 0.5 * (a + b) * ((atm->z_km[i + 1] - atm->z_km[i]) * 1.0e5);
        }
        atm->column_default[g] = s;
    }
    return 0;
}

bool grid_ok(const std::vector<double> &wl)
{
    if (wl.size() < 2) return false;
    for (size_t i = 0; i < wl.size(); ++i)
        if (!std::isfinite(wl[i]) || (i > 0 && !(wl[i] > wl[i - 1]))) return false;
    return wl.front() <= 330.0 + 1e-9 && wl.back() >= 1100.0 - 1e-9;
}

int load_xsec(const char *dir, int g, rt_xsec_t *x)
{
    std::memset(x, 0, sizeof(*x));
    const std::string path = std::string(dir) + "/xsec_" + kGas[g] + ".dat";
    FILE *f = std::fopen(path.c_str(), "r");
    if (!f) return 1;   // missing table: this gas contributes nothing
    std::vector<char> buf(1 << 16);
    int layers = 1;
    std::vector<std::vector<double>> rows;
    while (std::fgets(buf.data(), (int)buf.size(), f)) {
        const char *line = buf.data();
        if (line[0] == '#') {
            if (const char *p = std::strstr(line, "n_layers")) {
                if (std::sscanf(p, "n_layers = %d", &layers) != 1 || layers < 1) layers = 1;
            }
            if (const char *p = std::strstr(line, "T_ref")) std::sscanf(p, "T_ref = %lf", &x->T_ref_K);
            if (const char *p = std::strstr(line, "P_ref")) std::sscanf(p, "P_ref = %lf", &x->P_ref_atm);
            continue;
        }
        if (comment_or_blank(line)) continue;
        std::vector<double> r;
        const char *p = line;
        for (;;) {
            char *e;
            const double v = std::strtod(p, &e);
            if (e == p) break;
            r.push_back(v);
            p = e;
        }
        rows.push_back(std::move(r));
    }
    std::fclose(f);
    std::vector<double> wl, sig;
    if (layers == 1) {
        for (const auto &r : rows)
            if (r.size() >= 2) { wl.push_back(r[0]); sig.push_back(r[1]); }
    } else {
        if (rows.empty() || rows[0].size() < 2) return -2;
        wl = rows[0];
        sig.assign((size_t)layers * wl.size(), 0.0);
        for (int k = 0; k < layers && k + 1 < (int)rows.size(); ++k)
            for (size_t i = 0; i < wl.size() && i < rows[k + 1].size(); ++i) sig[k * wl.size() + i] = rows[k + 1][i];
    }
    if (!grid_ok(wl)) {
        std::fprintf(stderr, "rt_xsec_load: %s does not cover 330-1100 nm\n", path.c_str());
        return -3;
    }
    x->n_layers = layers;
    x->n_wl = (int)wl.size();
    x->wl_nm = dup(wl);
    x->sigma_cm2 = dup(sig);
    x->available = 1;
    return 0;
}

// ---- evaluation on the host (the same rules the device code follows) -----
double sigma_layer(const ocrt_gas::Table &t, int layer, double wl)
{
    if (!t.avail || t.n_wl < 2) return 0.0;
    layer = layer < 0 ? 0 : (layer >= t.n_layers ? t.n_layers - 1 : layer);
    const double *w = t.wl.data(), *s = t.sig.data() + (size_t)layer * t.n_wl;
    const int n = t.n_wl;
    if (wl < w[0] || wl > w[n - 1]) return NAN;
    if (wl == w[0]) return s[0];
    if (wl == w[n - 1]) return s[n - 1];
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        (w[mid] <= wl ? lo : hi) = mid;
    }
    const double u = (wl - w[lo]) / (w[hi] - w[lo]);
    return (1.0 - u) * s[lo] + u * s[hi];
}

double sigma_at(const ocrt_gas &g, int gi, double z, double wl)
{
    const ocrt_gas::Table &t = g.xs[gi];
    if (!t.avail) return 0.0;
    const int K = t.n_layers;
    if (K == 1) return sigma_layer(t, 0, wl);
    const double *zg = g.z.data();
    if (z <= zg[0]) return sigma_layer(t, 0, wl);
    if (z >= zg[K - 1]) return sigma_layer(t, K - 1, wl);
    int lo = 0, hi = K - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        (zg[mid] <= z ? lo : hi) = mid;
    }
    const double dz = zg[hi] - zg[lo];
    if (dz <= 0.0) return sigma_layer(t, lo, wl);
    const double u = (z - zg[lo]) / dz;
    return (1.0 - u) * sigma_layer(t, lo, wl) + u * sigma_layer(t, hi, wl);
}

double column_above(const ocrt_gas &g, int gi, double z)
{
    double cum = 0.0;
    const double *zz = g.z.data(), *nn = g.ngas[gi].data();
    for (int i = 0; i + 1 < g.n_lev; ++i) {
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

// Column optical depth of one gas over the AFGL layers (rt_absorption.c:520-545).
double tau_gas(const ocrt_gas &g, int gi, double wl)
{
    if (gi < 0 || gi > 5 || !g.xs[gi].avail || g.n_lev < 2) return 0.0;
    const bool ov = g.overridden[gi] && g.col_def[gi] > 0.0;
    const double scale = ov ? g.col_eff[gi] / g.col_def[gi] : 1.0;
    double tau = 0.0;
    for (int k = 0; k + 1 < g.n_lev; ++k) {
        const double lo = g.z[k], hi = g.z[k + 1];
        const double sig = sigma_at(g, gi, 0.5 * (lo + hi), wl);
        double N = hi > lo ? column_above(g, gi, lo) - column_above(g, gi, hi) : 0.0;
        if (ov) N *= scale;
        tau += sig * N;
    }
    return tau;
}

}  // namespace

// rt_absorption_t -> the solver's table (copies every array).
ocrt_gas *ocrt_gas_from_state(const rt_absorption_t *st)
{
    if (!st || !st->initialized || st->atm.n_levels < 2 || !st->atm.z_km) return nullptr;
    ocrt_gas *g = new ocrt_gas();
    const int n = st->atm.n_levels;
    g->n_lev = n;
    g->z.assign(st->atm.z_km, st->atm.z_km + n);
    for (int gi = 0; gi < 6; ++gi) {
        g->ngas[gi].resize(n);
        for (int i = 0; i < n; ++i) g->ngas[gi][i] = st->atm.mr[gi][i] * 1e-6 * st->atm.n_total_cm3[i];
        g->col_def[gi] = st->atm.column_default[gi];
        g->col_eff[gi] = st->column_eff[gi];
        g->overridden[gi] = st->column_overridden[gi];
        const rt_xsec_t &x = st->xsec[gi];
        ocrt_gas::Table &t = g->xs[gi];
        if (x.available && x.n_wl >= 2 && x.wl_nm && x.sigma_cm2) {
            if (x.n_layers > n) {   // layer i sits at AFGL level i
                delete g;
                return nullptr;
            }
            t.avail = 1;
            t.n_wl = x.n_wl;
            t.n_layers = x.n_layers;
            t.wl.assign(x.wl_nm, x.wl_nm + x.n_wl);
            t.sig.assign(x.sigma_cm2, x.sigma_cm2 + (size_t)x.n_layers * x.n_wl);
        }
    }
    return g;
}

namespace ocrt {

DevGas pack_gas(const ocrt_gas &g, std::vector<double> &pool)
{
    DevGas d{};
    d.n_lev = g.n_lev;
    d.off_z = (long long)pool.size();
    pool.insert(pool.end(), g.z.begin(), g.z.end());
    d.off_n = (long long)pool.size();
    for (int gi = 0; gi < 6; ++gi) pool.insert(pool.end(), g.ngas[gi].begin(), g.ngas[gi].end());
    for (int gi = 0; gi < 6; ++gi) {
        const ocrt_gas::Table &t = g.xs[gi];
        d.avail[gi] = t.avail;
        d.n_wl[gi] = t.n_wl;
        d.n_layers[gi] = t.n_layers;
        d.overridden[gi] = g.overridden[gi];
        d.col_def[gi] = g.col_def[gi];
        d.col_eff[gi] = g.col_eff[gi];
        d.off_wl[gi] = (long long)pool.size();
        pool.insert(pool.end(), t.wl.begin(), t.wl.end());
        d.off_sig[gi] = (long long)pool.size();
        pool.insert(pool.end(), t.sig.begin(), t.sig.end());
    }
    return d;
}

}  // namespace ocrt

extern "C" int rt_absorption_init(const char *afgl_dir, const char *xsec_dir, rt_afgl_profile_t profile,
                                  const double gas_column_override[RT_N_GAS], rt_absorption_t *st)
{
    if (!st || !afgl_dir || !xsec_dir) return -1;
    std::memset(st, 0, sizeof(*st));
    int rc = load_afgl(afgl_dir, (int)profile, &st->atm);
    if (rc) return rc;
    int missing = 0;
    for (int g = 0; g < RT_N_GAS; ++g) {
        const int r = load_xsec(xsec_dir, g, &st->xsec[g]);
        if (r > 0) ++missing;
        else if (r < 0) {
            rt_absorption_free(st);
            return r;
        }
    }
    if (missing == RT_N_GAS)
        std::fprintf(stderr, "rt_absorption_init: all %d cross-section files missing from '%s'; absorption is 0\n",
                     RT_N_GAS, xsec_dir);
    else if (missing)
        std::fprintf(stderr, "rt_absorption_init: %d/%d cross-section files missing\n", missing, RT_N_GAS);
    for (int g = 0; g < RT_N_GAS; ++g) {
        const bool ov = gas_column_override && gas_column_override[g] >= 0.0;
        st->column_eff[g] = ov ? gas_column_override[g] : st->atm.column_default[g];
        st->column_overridden[g] = ov ? 1 : 0;
    }
    st->initialized = 1;
    return 0;
}

extern "C" void rt_absorption_free(rt_absorption_t *st)
{
    if (!st) return;
    std::free(st->atm.z_km); std::free(st->atm.P_mbar); std::free(st->atm.T_K); std::free(st->atm.n_total_cm3);
    for (int g = 0; g < RT_N_GAS; ++g) {
        std::free(st->atm.mr[g]);
        std::free(st->xsec[g].wl_nm);
        std::free(st->xsec[g].sigma_cm2);
    }
    std::memset(st, 0, sizeof(*st));
}

extern "C" double rt_absorption_tau_per_gas(const rt_absorption_t *st, int gi, double wl)
{
    ocrt_gas *g = ocrt_gas_from_state(st);
    if (!g) return 0.0;
    const double t = tau_gas(*g, gi, wl);
    delete g;
    return t;
}

extern "C" double rt_absorption_tau_total(const rt_absorption_t *st, double wl)
{
    ocrt_gas *g = ocrt_gas_from_state(st);
    if (!g) return 0.0;
    double t = 0.0;
    for (int gi = 0; gi < RT_N_GAS; ++gi) t += tau_gas(*g, gi, wl);
    delete g;
    return t;
}

extern "C" int ocrt_gas_create(const char *afgl_dir, const char *xsec_dir, int profile, const double *override_,
                               ocrt_gas **out)
{
    if (!out) return OCRT_ERR_ARGS;
    *out = nullptr;
    rt_absorption_t st;
    if (rt_absorption_init(afgl_dir, xsec_dir, (rt_afgl_profile_t)profile, override_, &st) != 0)
        return OCRT_ERR_ARGS;
    *out = ocrt_gas_from_state(&st);
    rt_absorption_free(&st);
    return *out ? OCRT_OK : OCRT_ERR_ARGS;
}

extern "C" double ocrt_gas_tau(const ocrt_gas *g, int gi, double wl)
{
    if (!g) return 0.0;
    if (gi >= 0) return tau_gas(*g, gi, wl);
    double t = 0.0;
    for (int k = 0; k < 6; ++k) t += tau_gas(*g, k, wl);
    return t;
}

extern "C" void ocrt_gas_destroy(ocrt_gas *g) { delete g; }
