// ocrt -- command-line front end of the MI355X SOS solver.
//
// Option names, defaults, gates (OCRT_ADVANCED / OCRT_DEBUG), run modes and
// output formats follow OCRT's CLI (code/OCRT_C/src/main.c:1880-3906) and its
// printers (code/OCRT_C/src/rt_io.c:64-531) for the atmosphere surfaces
// black, flat and black_fresnel_ocean, with and without --pssa:
//   single geometry  --sza S --wavelength W --vza V --raa R --surface K ...
//                    -> one result line on stdout (rt_io.c:513-531)
//   LUT grid         the same without --vza/--raa -> 13-column CSV
//                    (main.c:3780-3892), file result<N>.csv unless
//                    --output-full-grid names one
//   batch            --batch IN.csv --output OUT.csv (rt_io.c:64-448); all rows
//                    are solved together by one GPU plan
// The solves go through the reference-compatible entry points of
// libocrt_amd.so (include/ocrt_rt.h); gas absorption reads the reference's
// data files under ./inputs like the reference does.  What this build does not
// serve (the coupled ocean, the Mie aerosol pipeline) ends with an error
// message and exit status 2.
#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <dirent.h>
#include <sys/stat.h>

#include "ocrt_rt.h"

namespace {

const char *kVersion = "ocrt_amd 0.3 (MI355X SOS solver; OCRT CLI interface v1.2)";

bool env_on(const char *name)
{
    const char *e = std::getenv(name);
    return e && *e && std::strcmp(e, "0") != 0;
}
bool advanced_on() { return env_on("OCRT_ADVANCED"); }
bool debug_on() { return env_on("OCRT_DEBUG"); }

[[noreturn]] void need_value(const char *flag)
{
    std::fprintf(stderr, "error: %s requires a value\n", flag);
    std::exit(2);
}
double next_double(int argc, char **argv, int &i, const char *flag)
{
    if (i + 1 >= argc) need_value(flag);
    return std::strtod(argv[++i], nullptr);
}
int next_int(int argc, char **argv, int &i, const char *flag)
{
    if (i + 1 >= argc) need_value(flag);
    return (int)std::strtol(argv[++i], nullptr, 10);
}
const char *next_str(int argc, char **argv, int &i, const char *flag)
{
    if (i + 1 >= argc) need_value(flag);
    return argv[++i];
}

void usage(FILE *fp, const char *prog)
{
    std::fprintf(fp,
        "%s\n\n"
        "Usage:\n"
        "  LUT grid (default)  %s --sza S --wavelength W --surface K [opts]\n"
        "      without --vza/--raa: rho and transmittances on a uniform (vza, raa) grid from one\n"
        "      solve (--lut-vza-step, --lut-raa-step, --lut-vza-max; CSV to --output-full-grid F,\n"
        "      else result<N>.csv)\n"
        "  single geometry     ... --vza V --raa R   (both required)\n"
        "  batch               %s --batch IN.csv --output OUT.csv --surface K [opts]\n\n"
        "Geometry / spectrum:\n"
        "  --sza DEG  --vza DEG  --raa DEG (180 = direct-glint branch)  --wavelength NM (330..1100)\n"
        "  --surface black | flat | black_fresnel_ocean      (ocean: coupled model, not served here)\n"
        "  --wind-speed MS   required for black_fresnel_ocean (0 = flat Fresnel)\n"
        "  --n-water X (1.34)   --decouple-sunglint   --pssa (IPSS spherical-shell correction)\n"
        "  --pressure HPA (1013.25; 0 = no Rayleigh)   --no-rayleigh   --rayleigh on|off\n"
        "Gas absorption (always on; AFGL profile + cross sections read from ./inputs/afgl_atm, ./inputs/xsec):\n"
        "  --atm-profile usstd76|tropical|mlsumm|mlwint|sasumm|sawint|userdef\n"
        "  --gas-column-h2o G (g/cm2, alias --pwv-g-cm)  --gas-column-o3 DU  --gas-column-no2 DU\n"
        "  --gas-column-o2 N  --gas-column-co2 N  --gas-column-ch4 N (molecules/cm2); all six 0 = no absorption\n"
        "Aerosol (--aod-555/--aod-865/--aod > 0 with --mie FILE): the Mie pipeline is not served;\n"
        "  aerosol solves go through rt_solve_case_pol_aerosol with Greek coefficients.\n"
        "Output: --output-full-grid F  --output-advanced (batch CSV detail columns)  --verbose\n"
        "Convergence knobs (values other than the defaults need OCRT_ADVANCED=1):\n"
        "  --n-mu 24 --n-layers 40 --m-max 2 --max-orders 20 --sos-max-orders 20\n"
        "  --sos-tolerance 1e-7 --conv-tol 1e-6 --l-max auto --sigma-model ocrt-floor\n"
        "  --sigma-type 1 --q-convention 1\n"
        "Diagnostics (OCRT_DEBUG=1): --integration-method constant, --sos-save-orders, --nt-tau,\n"
        "  --trunc-aer-fit, --trunc-aer-m\n"
        "  --help  --version  --examples\n",
        kVersion, prog, prog);
}

void examples(FILE *fp, const char *prog)
{
    std::fprintf(fp,
        "# single geometry, Rayleigh and US76 gases over a rough ocean (run where ./inputs holds the data)\n"
        "%s --surface black_fresnel_ocean --wind-speed 5 --sza 40 --vza 30 --raa 90 --wavelength 555\n"
        "# LUT grid with IPSS over a black surface, 300 DU of ozone\n"
        "%s --surface black --sza 70 --wavelength 443 --pssa --gas-column-o3 300 --output-full-grid lut.csv\n"
        "# batch CSV: case_id,sza_deg,vza_deg,raa_deg,wavelength_nm[,tau_R_ref[,rho_I_ref]]\n"
        "%s --surface flat --batch in.csv --output out.csv\n", prog, prog, prog);
}

int resolve_csv_name(const char *given, std::string &out)
{
    if (given) {
        const char *base = std::strrchr(given, '/');
        base = base ? base + 1 : given;
        const char *dot = std::strrchr(base, '.');
        if (!dot) { out = std::string(given) + ".csv"; return 0; }
        if (std::strcmp(dot, ".csv") == 0) { out = given; return 0; }
        std::fprintf(stderr, "error: output extension '%s' is not supported (csv only)\n", dot);
        return 2;
    }
    long maxn = 0;
    if (DIR *d = opendir(".")) {
        while (dirent *e = readdir(d)) {
            long n;
            char tail[8] = {0};
            if (std::sscanf(e->d_name, "result%ld%7s", &n, tail) == 2 && std::strcmp(tail, ".csv") == 0 && n > maxn)
                maxn = n;
        }
        closedir(d);
    }
    out = "result" + std::to_string(maxn + 1) + ".csv";
    return 0;
}

// rt_io.c:513-531 (atmosphere surfaces)
void print_single(const rt_case_t &cs, const rt_result_t &r)
{
    std::printf("%.10e %.10e %.10e  "
                "TOA_rho_I=%.10e TOA_rho_Q=%.10e TOA_rho_U=%.10e "
                "TOA_rho_glint_direct_I=%.10e TOA_rho_glint_direct_Q=%.10e TOA_rho_glint_direct_U=%.10e "
                "T_dir_dn=%.10e T_diff_dn_hemi=%.10e T_total_dn_hemi=%.10e T_diff_dn_dir=%.10e "
                "T_dir_up_view=%.10e T_diff_up_view=%.10e T_total_up_view=%.10e TOA_water_signal_I=%.10e "
                "T_up_rt_valid=%d "
                "diag_sunglint_up_flux_ratio=%.10e diag_TOA_up_flux_ratio=%.10e "
                "AOD_ref=%.10e AOD_ref_nm=%.3f AOD_band=%.10e AOD_ext_ratio=%.10e "
                "tau_R=%.10e sza=%.3f vza=%.3f raa=%.3f wl=%.3f orders=%d conv=%d\n",
                r.rho_I, r.rho_Q, r.rho_U, r.rho_I, r.rho_Q, r.rho_U,
                r.rho_glint_direct_I, r.rho_glint_direct_Q, r.rho_glint_direct_U,
                r.T_dir_dn, r.T_diff_dn_hemi, r.T_total_dn_hemi, r.T_diff_dn_dir,
                r.T_dir_up_view, r.T_diff_up_view, r.T_total_up_view, r.I_TOA_water_signal, r.T_up_rt_valid,
                r.T_sg_up_dir, r.T_total_up_dir,
                r.aerosol_aod_ref, r.aerosol_aod_ref_nm, r.aerosol_aod_band, r.aerosol_extinction_ratio,
                r.tau_R_total, cs.sza_deg, cs.vza_deg, cs.raa_deg, cs.wavelength_nm, r.n_orders_used, r.converged);
}

const char *kLutHeader = "sza_deg,wavelength_nm,vza_deg,raa_deg,rho_I,rho_Q,rho_U,"
                         "direct_transmittance,irradiance_transmittance,diffuse_dn_hemi,diffuse_dn_dir,"
                         "sunglint_up_flux_ratio_diag,toa_up_flux_ratio_diag\n";

struct Grid {
    std::vector<double> vza, raa, I, Q, U, tdd, tsg, ttu;
    rt_lut_grid_out_t lut{};
    Grid(double vmax, double vstep, double rstep)
    {
        const int nv = (int)(vmax / vstep) + 1, nr = (int)(360.0 / rstep);
        for (int i = 0; i < nv; ++i) vza.push_back(i * vstep);
        for (int i = 0; i < nr; ++i) raa.push_back(i * rstep);
        const size_t n = (size_t)nv * nr;
        for (auto *v : {&I, &Q, &U, &tdd, &tsg, &ttu}) v->assign(n, 0.0);
        lut.n_vza = nv; lut.vza_deg = vza.data();
        lut.n_raa = nr; lut.raa_deg = raa.data();
        lut.rho_I = I.data(); lut.rho_Q = Q.data(); lut.rho_U = U.data();
        lut.T_diff_dn_dir = tdd.data(); lut.T_sg_up_dir = tsg.data(); lut.T_total_up_dir = ttu.data();
        lut.view_m_max = -1;
    }
    // main.c:3859-3885 / rt_io.c:296-319
    void write(FILE *fp, const rt_case_t &cs) const
    {
        std::fputs(kLutHeader, fp);
        const double irr = lut.T_dir_dn + lut.T_diff_dn_hemi;
        for (int iv = 0; iv < lut.n_vza; ++iv)
            for (int ir = 0; ir < lut.n_raa; ++ir) {
                const int k = iv * lut.n_raa + ir;
                std::fprintf(fp, "%.6f,%.3f,%.6f,%.6f,%.10e,%.10e,%.10e,%.10e,%.10e,%.10e,%.10e,%.10e,%.10e\n",
                             cs.sza_deg, cs.wavelength_nm, vza[iv], raa[ir], I[k], Q[k], U[k], lut.T_dir_dn, irr,
                             lut.T_diff_dn_hemi, tdd[k], tsg[k], ttu[k]);
            }
    }
};

// ---- batch CSV (rt_io.c:64-448) -------------------------------------------
struct Row {
    int case_id = 0;
    double sza = 0, vza = 0, raa = 0, wl = 0, tau_ref = 0, rho_I_ref = 0, rho_Q_ref = 0, rho_U_ref = 0;
    bool have_tau = false, have_rho = false, have_QU = false;
    rt_result_t res{};
    int rc = 0;
    double ms = 0.0;
};

void rstrip(char *s)
{
    size_t n = std::strlen(s);
    while (n > 0 && (s[n - 1] == '\n' || s[n - 1] == '\r' || s[n - 1] == ' ' || s[n - 1] == '\t')) s[--n] = '\0';
}

int count_fields(const char *s)
{
    int n = 1;
    for (; *s; ++s) n += (*s == ',');
    return n;
}

int run_batch(const char *in_csv, const char *out_csv, const rt_options_t &opts, const rt_case_t &tmpl, bool mie)
{
    FILE *fi = std::fopen(in_csv, "r");
    if (!fi) { std::fprintf(stderr, "rt_io_run_batch: cannot open input '%s'\n", in_csv); return -1; }
    char header[2048], line[2048];
    if (!std::fgets(header, sizeof header, fi)) {
        std::fprintf(stderr, "rt_io_run_batch: empty input file\n");
        std::fclose(fi);
        return -1;
    }
    rstrip(header);
    const int nc = count_fields(header);
    const bool aer_layout = nc == 9 && (std::strstr(header, "aod") || std::strstr(header, "os_I") ||
                                        std::strstr(header, "osoaa") || std::strstr(header, "decoup_I"));
    if (nc < 5 || nc > 9 || nc == 8) {
        std::fprintf(stderr, "rt_io_run_batch: header must have 5-7 or 9 columns (got %d in '%s')\n", nc, header);
        std::fclose(fi);
        return -1;
    }
    if (aer_layout) {
        std::fprintf(stderr, "rt_io_run_batch: aerosol CSV requires --mie and aerosol config%s\n",
                     mie ? " (the Mie pipeline is not served by this build)" : "");
        std::fclose(fi);
        return -1;
    }
    const bool six_is_rho = nc == 6 && std::strstr(header, "rho_I");
    std::vector<Row> rows;
    while (std::fgets(line, sizeof line, fi)) {
        rstrip(line);
        if (line[0] == '\0' || line[0] == '#') continue;
        Row r;
        int got;
        if (nc == 9) {
            got = std::sscanf(line, "%d,%lf,%lf,%lf,%lf,%lf,%lf,%lf,%lf", &r.case_id, &r.sza, &r.vza, &r.raa, &r.wl,
                              &r.tau_ref, &r.rho_I_ref, &r.rho_Q_ref, &r.rho_U_ref);
            if (got == 9) r.have_tau = r.have_rho = r.have_QU = true;
        } else if (nc == 7) {
            got = std::sscanf(line, "%d,%lf,%lf,%lf,%lf,%lf,%lf", &r.case_id, &r.sza, &r.vza, &r.raa, &r.wl,
                              &r.tau_ref, &r.rho_I_ref);
            if (got == 7) r.have_tau = r.have_rho = true;
        } else if (nc == 6) {
            got = std::sscanf(line, "%d,%lf,%lf,%lf,%lf,%lf", &r.case_id, &r.sza, &r.vza, &r.raa, &r.wl,
                              six_is_rho ? &r.rho_I_ref : &r.tau_ref);
            if (got == 6) (six_is_rho ? r.have_rho : r.have_tau) = true;
        } else {
            got = std::sscanf(line, "%d,%lf,%lf,%lf,%lf", &r.case_id, &r.sza, &r.vza, &r.raa, &r.wl);
        }
        if (got < nc) {
            std::fprintf(stderr, "rt_io_run_batch: skip malformed row %d: '%s'\n", (int)rows.size() + 1, line);
            continue;
        }
        rows.push_back(r);
    }
    std::fclose(fi);
    if (rows.empty()) { std::fprintf(stderr, "rt_io_run_batch: no data rows\n"); return -1; }
    const int n = (int)rows.size();
    // each row takes the template's surface fields and its own geometry (rt_io.c:199-216)
    std::vector<rt_case_t> cases(n);
    for (int i = 0; i < n; ++i) {
        rt_case_t &c = cases[i];
        std::memset(&c, 0, sizeof c);
        c.sza_deg = rows[i].sza; c.vza_deg = rows[i].vza; c.raa_deg = rows[i].raa; c.wavelength_nm = rows[i].wl;
        c.surface = tmpl.surface; c.rayleigh_on = tmpl.rayleigh_on; c.aerosol_on = tmpl.aerosol_on;
        c.n_water = tmpl.n_water; c.wind_speed = tmpl.wind_speed; c.sigma_type = tmpl.sigma_type;
        c.q_convention = tmpl.q_convention; c.decouple_sunglint = tmpl.decouple_sunglint;
        c.tau_R_input = rows[i].have_tau ? rows[i].tau_ref : 0.0;
    }
    bool any_bad = false;
    if (opts.lut_enable && opts.vector_mode) {
        // one LUT solve per row, CSV per case
        for (int i = 0; i < n; ++i) {
            const auto t0 = std::chrono::steady_clock::now();
            Grid g(opts.lut_vza_max > 0 ? opts.lut_vza_max : 85.0, opts.lut_vza_step > 0 ? opts.lut_vza_step : 2.5,
                   opts.lut_raa_step > 0 ? opts.lut_raa_step : 5.0);
            rows[i].rc = rt_solve_case_pol_lut(&cases[i], &opts, nullptr, &rows[i].res, &g.lut);
            if (rows[i].rc == 0 && opts.lut_output_dir) {
                const std::string path = std::string(opts.lut_output_dir) + "/case_" +
                                         std::to_string(rows[i].case_id) + ".csv";
                if (FILE *fl = std::fopen(path.c_str(), "w")) {
                    g.write(fl, cases[i]);
                    std::fclose(fl);
                }
            }
            rows[i].ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
    } else {
        // every row in one plan
        std::vector<rt_result_t> res(n);
        std::vector<int> rcs(n, 0);
        const auto t0 = std::chrono::steady_clock::now();
        ocrt_rt_solve_batch(cases.data(), &opts, nullptr, n, res.data(), rcs.data());
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        for (int i = 0; i < n; ++i) {
            rows[i].res = res[i];
            rows[i].rc = rcs[i];
            rows[i].ms = ms / n;   // the batch wall time shared over its rows
        }
    }
    for (const Row &r : rows) any_bad |= (r.rc == -2 || r.rc == -3);
    FILE *fo = std::fopen(out_csv, "w");
    if (!fo) { std::fprintf(stderr, "rt_io_run_batch: cannot open output '%s'\n", out_csv); return -1; }
    std::fprintf(fo, "case_id,sza_deg,vza_deg,raa_deg,wavelength_nm,");
    const bool debug = opts.output_mode_debug != 0;
    const bool have_QU_ref = rows[0].have_QU;
    if (debug) {
        std::fprintf(fo, "rho_I_v2,rho_I_ref,delta_abs,delta_rel_pct,delta_rel_abs_pct,"
                         "tau_R_v2,tau_R_ref,n_orders_m0,n_orders_m1,n_orders_m2,walltime_ms");
        if (opts.vector_mode && have_QU_ref)
            std::fprintf(fo, ",rho_Q_v2,rho_Q_ref,delta_abs_Q,delta_rel_pct_Q,delta_rel_abs_pct_Q"
                             ",rho_U_v2,rho_U_ref,delta_abs_U,delta_rel_pct_U,delta_rel_abs_pct_U");
        else if (opts.vector_mode)
            std::fprintf(fo, ",rho_Q_v2,rho_U_v2");
    } else {
        std::fprintf(fo, "rho_I_v2");
        if (opts.vector_mode) std::fprintf(fo, ",rho_Q_v2,rho_U_v2");
    }
    std::fprintf(fo, "\n");
    for (const Row &r : rows) {
        std::fprintf(fo, "%d,%.6f,%.6f,%.6f,%.3f,", r.case_id, r.sza, r.vza, r.raa, r.wl);
        if (!debug) {
            std::fprintf(fo, "%.10e", r.res.rho_I);
            if (opts.vector_mode) std::fprintf(fo, ",%.10e,%.10e", r.res.rho_Q, r.res.rho_U);
            std::fprintf(fo, "\n");
            continue;
        }
        std::fprintf(fo, "%.10e,", r.res.rho_I);
        if (r.have_rho) {
            const double da = r.res.rho_I - r.rho_I_ref;
            const double drs = r.rho_I_ref != 0.0 ? 100.0 * da / r.rho_I_ref : 0.0;
            std::fprintf(fo, "%.10e,%.6e,%.6e,%.6e,", r.rho_I_ref, da, drs, std::fabs(drs));
        } else {
            std::fprintf(fo, ",,,,");
        }
        std::fprintf(fo, "%.10e,", r.res.tau_R_total);
        if (r.have_tau) std::fprintf(fo, "%.10e,", r.tau_ref);
        else std::fprintf(fo, ",");
        std::fprintf(fo, "%d,%d,%d,%.3f", r.res.sos_orders_per_m[0], r.res.sos_orders_per_m[1],
                     r.res.sos_orders_per_m[2], r.ms);
        if (opts.vector_mode && have_QU_ref) {
            const double iref = std::fabs(r.rho_I_ref);
            const double dq = r.res.rho_Q - r.rho_Q_ref, du = r.res.rho_U - r.rho_U_ref;
            const double pq = iref > 0.0 ? 100.0 * dq / iref : 0.0, pu = iref > 0.0 ? 100.0 * du / iref : 0.0;
            std::fprintf(fo, ",%.10e,%.10e,%.6e,%.6e,%.6e,%.10e,%.10e,%.6e,%.6e,%.6e", r.res.rho_Q, r.rho_Q_ref, dq,
                         pq, std::fabs(pq), r.res.rho_U, r.rho_U_ref, du, pu, std::fabs(pu));
        } else if (opts.vector_mode) {
            std::fprintf(fo, ",%.10e,%.10e", r.res.rho_Q, r.res.rho_U);
        }
        std::fprintf(fo, "\n");
    }
    std::fclose(fo);
    int compared = 0;
    double worst = 0.0;
    for (const Row &r : rows) {
        if (!r.have_rho || r.rho_I_ref == 0.0) continue;
        ++compared;
        worst = std::fmax(worst, std::fabs(100.0 * (r.res.rho_I - r.rho_I_ref) / r.rho_I_ref));
    }
    std::printf("rt_io_run_batch: n=%d, compared=%d, max|drel_I|=%.4f%%, output=%s\n", n, compared, worst, out_csv);
    return any_bad ? -2 : 0;
}

bool water_option(const char *a)
{
    static const char *names[] = {
        "--water-temperature", "--water-salinity", "--cdom-a440", "--cdom-slope", "--cdom-ref-lambda",
        "--water-model", "--ocrt-chl", "--ccrr-chl", "--ocrt-tsm", "--ccrr-tsm", "--ocrt-adom440",
        "--ccrr-adom440", "--ocrt-adom-slope", "--ccrr-adom-slope", "--ocrt-phyto-group", "--ocrt-tsm-species",
        "--ocrt-detritus-a440", "--ocrt-detritus-slope", "--ccrr-phase-moments", "--ccrr-particle-phase-lut",
        "--ccrr-particle-phase-case-id", "--ccrr-particle-phase-lmax", "--ccrr-particle-phase-nphi", "--iop-a",
        "--iop-b", "--iop-bb", "--fixed-bulk-iop", "--water-fixed-bulk-iop", "--iop-phase-lut",
        "--fixed-bulk-phase-lut", "--iop-mie-phase", "--water-mie-phase", "--ocrt-mie-moment-mode",
        "--iop-mie-moment-mode", "--water-mie-moment-mode", "--ocrt-mie-truncation", "--iop-mie-truncation",
        "--water-mie-truncation", "--ocrt-mie-ss-mode", "--iop-mie-ss-mode", "--water-mie-ss-mode",
        "--ocrt-mie-moment-nmu", "--iop-mie-moment-nmu", "--water-mie-moment-nmu", "--iop-phase-nphi",
        "--fixed-bulk-phase-nphi", "--n-mu-water", "--n-mu-water-sky", "--water-m-max", "--water-shared-grid",
        "--batch-full-grid", "--water-view-as-node", "--water-view-as-node-off", "--simple-compare",
        "--ccrr-compare", "--simple-output", "--ccrr-output"};
    for (const char *n : names)
        if (!std::strcmp(a, n)) return true;
    return !std::strncmp(a, "--debug-water-", 14) || !std::strncmp(a, "--scalar-ff-", 12) ||
           !std::strncmp(a, "--ff-", 5) || !std::strncmp(a, "--debug-fixed-bulk-", 19);
}

}  // namespace

int main(int argc, char **argv)
{
    rt_case_t cs;
    std::memset(&cs, 0, sizeof cs);
    cs.surface = RT_SURFACE_BLACK;
    cs.pressure_hpa = 1013.25;
    cs.n_water = 1.34;
    cs.sigma_type = 1;
    cs.q_convention = 1;
    cs.F_sun = RT_F_SOLAR_PI;
    rt_options_t opts = rt_options_default();

    const char *batch_in = nullptr, *batch_out = nullptr, *mie_path = nullptr, *lut_output = nullptr;
    bool have_sza = false, have_vza = false, have_raa = false, have_wl = false, have_surface = false,
         have_wind = false, have_m_max = false, have_n_layers = false, have_aer_h = false, have_aer_L = false,
         have_vza_step = false, have_raa_step = false, lut_mode = false, trunc_fit = false, trunc_m = false,
         nt_tau = false;
    double user_aod = 0.0, user_aod_ref_nm = 555.0;
    double lut_vza_step = 2.5, lut_raa_step = 2.5, lut_vza_max = 85.0;
    int aer_L_max = 32;
    double gas_col[6] = {-1, -1, -1, -1, -1, -1};   // h2o o3 o2 co2 no2 ch4 (mol/cm^2)
    int atm_profile = RT_AFGL_USSTD76;

    for (int i = 1; i < argc; ++i) {
        const char *a = argv[i];
        if (!std::strcmp(a, "--debug")) { setenv("OCRT_DEBUG", "1", 1); continue; }
        if (!std::strcmp(a, "--help") || !std::strcmp(a, "-h")) { usage(stdout, argv[0]); return 0; }
        else if (!std::strcmp(a, "--examples")) { examples(stdout, argv[0]); return 0; }
        else if (!std::strcmp(a, "--version")) { std::printf("ocrt %s\n", kVersion); return 0; }
        else if (!std::strcmp(a, "--sza")) { cs.sza_deg = next_double(argc, argv, i, a); have_sza = true; }
        else if (!std::strcmp(a, "--vza")) { cs.vza_deg = next_double(argc, argv, i, a); have_vza = true; }
        else if (!std::strcmp(a, "--raa")) { cs.raa_deg = next_double(argc, argv, i, a); have_raa = true; }
        else if (!std::strcmp(a, "--wavelength")) { cs.wavelength_nm = next_double(argc, argv, i, a); have_wl = true; }
        else if (!std::strcmp(a, "--surface")) {
            const char *s = next_str(argc, argv, i, a);
            if (!std::strcmp(s, "black")) cs.surface = RT_SURFACE_BLACK;
            else if (!std::strcmp(s, "lambert")) {
                std::fprintf(stderr, "error: --surface lambert not implemented (boundary condition not coded). Use "
                                     "--surface black for atm-only or --surface flat/black_fresnel_ocean for a "
                                     "Fresnel ocean boundary.\n");
                return 2;
            } else if (!std::strcmp(s, "black_fresnel_ocean") || !std::strcmp(s, "black-fresnel-ocean")) {
                cs.surface = RT_SURFACE_BLACK_FRESNEL_OCEAN;
            } else if (!std::strcmp(s, "coxmunk")) {
                std::fprintf(stderr, "warning: --surface coxmunk is deprecated; use --surface black_fresnel_ocean.\n");
                cs.surface = RT_SURFACE_BLACK_FRESNEL_OCEAN;
            } else if (!std::strcmp(s, "flat")) cs.surface = RT_SURFACE_FLAT;
            else if (!std::strcmp(s, "ocean")) {
                std::fprintf(stderr, "error: --surface ocean (coupled in-water RT) is not served by this solver; "
                                     "use black, flat or black_fresnel_ocean\n");
                return 2;
            } else { std::fprintf(stderr, "error: unknown surface '%s'\n", s); return 2; }
            have_surface = true;
        } else if (!std::strcmp(a, "--wind-speed")) { cs.wind_speed = next_double(argc, argv, i, a); have_wind = true; }
        else if (!std::strcmp(a, "--n-water")) cs.n_water = next_double(argc, argv, i, a);
        else if (water_option(a)) {
            std::fprintf(stderr, "error: %s belongs to the coupled ocean model (--surface ocean), which this solver "
                                 "does not serve\n", a);
            return 2;
        } else if (!std::strcmp(a, "--sigma-model")) {
            const char *v = next_str(argc, argv, i, a);
            if (!std::strcmp(v, "ocrt-floor") || !std::strcmp(v, "ocrt_floor")) cs.sigma_type = 1;
            else if (!std::strcmp(v, "nakajima-tanaka") || !std::strcmp(v, "nakajima_tanaka")) cs.sigma_type = 0;
            else { std::fprintf(stderr, "error: --sigma-model must be ocrt-floor|nakajima-tanaka\n"); return 2; }
            if (cs.sigma_type != 1 && !advanced_on()) {
                std::fprintf(stderr, "error: non-default --sigma-model requires OCRT_ADVANCED=1\n");
                return 2;
            }
        } else if (!std::strcmp(a, "--sigma-type")) {
            cs.sigma_type = next_int(argc, argv, i, a);
            if (cs.sigma_type != 1 && !advanced_on()) {
                std::fprintf(stderr, "error: --sigma-type %d differs from the default 1 (OCRT floor slope law; "
                                     "selector requires OCRT_ADVANCED=1)\n", cs.sigma_type);
                return 2;
            }
        } else if (!std::strcmp(a, "--q-convention")) {
            cs.q_convention = next_int(argc, argv, i, a);
            if (cs.q_convention != 1 && !advanced_on()) {
                std::fprintf(stderr, "error: --q-convention %d differs from the default 1 (Hansen/Mishchenko; "
                                     "convention selector requires OCRT_ADVANCED=1)\n", cs.q_convention);
                return 2;
            }
        } else if (!std::strcmp(a, "--pssa")) opts.pssa = 1;
        else if (!std::strcmp(a, "--pssa-mode")) {
            std::fprintf(stderr, "error: --pssa-mode was removed; IPSS (Zhai & Hu 2022) is the only spherical "
                                 "algorithm. Use plain --pssa.\n");
            return 2;
        } else if (!std::strcmp(a, "--decouple-sunglint")) cs.decouple_sunglint = 1;
        else if (!std::strcmp(a, "--mie")) mie_path = next_str(argc, argv, i, a);
        else if (!std::strcmp(a, "--aod-555")) { user_aod = next_double(argc, argv, i, a); user_aod_ref_nm = 555.0; }
        else if (!std::strcmp(a, "--aod-865")) { user_aod = next_double(argc, argv, i, a); user_aod_ref_nm = 865.0; }
        else if (!std::strcmp(a, "--aod-ref-wavelength")) user_aod_ref_nm = next_double(argc, argv, i, a);
        else if (!std::strcmp(a, "--aod")) user_aod = next_double(argc, argv, i, a);
        else if (!std::strcmp(a, "--aod-band")) {
            std::fprintf(stderr, "error: --aod-band was removed because band-specific AOD breaks spectral "
                                 "consistency; use --aod-555 or --aod-865\n");
            return 2;
        } else if (!std::strcmp(a, "--trunc-aer-fit") || !std::strcmp(a, "--trunc-aer-m")) {
            if (!debug_on()) {
                std::fprintf(stderr, "error: %s requires OCRT_DEBUG=1 (non-validated research truncation)\n", a);
                return 2;
            }
            (void)next_str(argc, argv, i, a);
            (!std::strcmp(a, "--trunc-aer-fit") ? trunc_fit : trunc_m) = true;
        } else if (!std::strcmp(a, "--aer-h-km")) {
            opts.aer_h_km = next_double(argc, argv, i, a);
            if (!(opts.aer_h_km > 0.0)) {
                std::fprintf(stderr, "error: --aer-h-km must be positive (OCRT default: 2.0 km exponential "
                                     "profile)\n");
                return 2;
            }
            have_aer_h = true;
        } else if (!std::strcmp(a, "--output-advanced")) opts.output_mode_debug = 1;
        else if (!std::strcmp(a, "--output-full-grid")) { lut_output = next_str(argc, argv, i, a); lut_mode = true; }
        else if (!std::strcmp(a, "--lut-vza-step") || !std::strcmp(a, "--vza-gap")) {
            lut_vza_step = next_double(argc, argv, i, a); have_vza_step = true;
        } else if (!std::strcmp(a, "--lut-raa-step") || !std::strcmp(a, "--raa-gap")) {
            lut_raa_step = next_double(argc, argv, i, a); have_raa_step = true;
        } else if (!std::strcmp(a, "--lut-vza-max")) lut_vza_max = next_double(argc, argv, i, a);
        else if (!std::strcmp(a, "--atm-profile") || !std::strcmp(a, "--gas-profile")) {
            const char *s = next_str(argc, argv, i, a);
            static const char *ok[] = {"usstd76", "tropical", "mlsumm", "mlwint", "sasumm", "sawint", "userdef"};
            int found = -1;
            for (int q = 0; q < 7; ++q)
                if (!std::strcmp(s, ok[q])) found = q;
            atm_profile = found;
            if (found < 0) {
                std::fprintf(stderr, "error: --atm-profile must be one of "
                                     "usstd76|tropical|mlsumm|mlwint|sasumm|sawint|userdef\n");
                return 2;
            }
        } else if (!std::strcmp(a, "--gas-column-h2o") || !std::strcmp(a, "--pwv-g-cm")) {
            gas_col[0] = next_double(argc, argv, i, a) * 6.02214076e23 / 18.01528;
        } else if (!std::strcmp(a, "--gas-column-o3")) gas_col[1] = next_double(argc, argv, i, a) * 2.6868e16;
        else if (!std::strcmp(a, "--gas-column-o2")) gas_col[2] = next_double(argc, argv, i, a);
        else if (!std::strcmp(a, "--gas-column-co2")) gas_col[3] = next_double(argc, argv, i, a);
        else if (!std::strcmp(a, "--gas-column-no2")) gas_col[4] = next_double(argc, argv, i, a) * 2.6868e16;
        else if (!std::strcmp(a, "--gas-column-ch4")) gas_col[5] = next_double(argc, argv, i, a);
        else if (!std::strcmp(a, "--nt-tau")) {
            if (!debug_on()) {
                std::fprintf(stderr, "error: --nt-tau requires OCRT_DEBUG=1 (companion of the research truncation "
                                     "family)\n");
                return 2;
            }
            nt_tau = true;
        } else if (!std::strcmp(a, "--aer-l-max")) { aer_L_max = next_int(argc, argv, i, a); have_aer_L = true; }
        else if (!std::strcmp(a, "--no-rayleigh")) cs.pressure_hpa = 0.0;
        else if (!std::strcmp(a, "--rayleigh")) {
            const char *v = next_str(argc, argv, i, a);
            if (!std::strcmp(v, "off")) cs.pressure_hpa = 0.0;
            else if (!std::strcmp(v, "on")) { if (cs.pressure_hpa <= 0.0) cs.pressure_hpa = 1013.25; }
            else { std::fprintf(stderr, "error: --rayleigh must be on|off\n"); return 2; }
        } else if (!std::strcmp(a, "--pressure") || !std::strcmp(a, "--surface-pressure")) {
            cs.pressure_hpa = next_double(argc, argv, i, a);
        } else if (!std::strcmp(a, "--n-mu")) {
            opts.n_mu = next_int(argc, argv, i, a);
            if (opts.n_mu != 24 && !advanced_on()) {
                std::fprintf(stderr, "error: --n-mu %d differs from the validated default 24 (convergence knob; "
                                     "requires OCRT_ADVANCED=1)\n", opts.n_mu);
                return 2;
            }
        } else if (!std::strcmp(a, "--n-layers")) { opts.n_layers = next_int(argc, argv, i, a); have_n_layers = true; }
        else if (!std::strcmp(a, "--max-orders")) {
            opts.max_orders = next_int(argc, argv, i, a);
            if (opts.max_orders != 20 && !advanced_on()) {
                std::fprintf(stderr, "error: --max-orders %d differs from the validated default 20 (convergence "
                                     "knob; requires OCRT_ADVANCED=1)\n", opts.max_orders);
                return 2;
            }
        } else if (!std::strcmp(a, "--l-max")) {
            const char *s = next_str(argc, argv, i, a);
            opts.l_max = !std::strcmp(s, "auto") ? -1 : (int)std::strtol(s, nullptr, 10);
            if (opts.l_max != -1 && !advanced_on()) {
                std::fprintf(stderr, "error: --l-max %s overrides the auto dispatch (convergence knob; requires "
                                     "OCRT_ADVANCED=1)\n", s);
                return 2;
            }
        } else if (!std::strcmp(a, "--integration-method")) {
            const char *s = next_str(argc, argv, i, a);
            if (!std::strcmp(s, "linear")) opts.integration_method = RT_INTEGRATION_METHOD_LINEAR;
            else if (!std::strcmp(s, "constant")) {
                if (!debug_on()) {
                    std::fprintf(stderr, "error: --integration-method constant requires OCRT_DEBUG=1 (production "
                                         "is linear)\n");
                    return 2;
                }
                opts.integration_method = RT_INTEGRATION_METHOD_CONSTANT;
            } else {
                std::fprintf(stderr, "error: unknown integration-method '%s' (use linear|constant)\n", s);
                return 2;
            }
        } else if (!std::strcmp(a, "--sos-max-orders")) {
            opts.sos.max_iterations = next_int(argc, argv, i, a);
            if (opts.sos.max_iterations != 20 && !advanced_on()) {
                std::fprintf(stderr, "error: --sos-max-orders %d differs from the validated default 20 "
                                     "(convergence knob; requires OCRT_ADVANCED=1)\n", opts.sos.max_iterations);
                return 2;
            }
        } else if (!std::strcmp(a, "--sos-tolerance")) {
            opts.sos.tolerance = next_double(argc, argv, i, a);
            if (opts.sos.tolerance != 1e-7 && !advanced_on()) {
                std::fprintf(stderr, "error: --sos-tolerance %g differs from the validated default 1e-7 "
                                     "(convergence knob; requires OCRT_ADVANCED=1)\n", opts.sos.tolerance);
                return 2;
            }
        } else if (!std::strcmp(a, "--sos-acceleration")) {
            const char *s = next_str(argc, argv, i, a);
            if (!std::strcmp(s, "plain")) opts.sos.acceleration = RT_SOS_ACCELERATION_PLAIN;
            else if (!std::strcmp(s, "geometric")) {
                std::fprintf(stderr, "error: --sos-acceleration geometric is NOT implemented. Use plain.\n");
                return 2;
            } else {
                std::fprintf(stderr, "error: unknown sos-acceleration '%s' (use plain)\n", s);
                return 2;
            }
        } else if (!std::strcmp(a, "--sos-save-orders")) {
            if (!debug_on()) {
                std::fprintf(stderr, "error: --sos-save-orders is a diagnostic option and requires OCRT_DEBUG=1\n");
                return 2;
            }
            opts.sos.save_orders = 1;
        } else if (!std::strcmp(a, "--m-max")) { opts.fourier_m_max = next_int(argc, argv, i, a); have_m_max = true; }
        else if (!std::strcmp(a, "--conv-tol")) {
            opts.conv_tol = next_double(argc, argv, i, a);
            if (opts.conv_tol != 1e-6 && !advanced_on()) {
                std::fprintf(stderr, "error: --conv-tol %g differs from the validated default 1e-6 (convergence "
                                     "knob; requires OCRT_ADVANCED=1)\n", opts.conv_tol);
                return 2;
            }
        } else if (!std::strcmp(a, "--verbose")) opts.verbose = 1;
        else if (!std::strcmp(a, "--batch")) batch_in = next_str(argc, argv, i, a);
        else if (!std::strcmp(a, "--output")) batch_out = next_str(argc, argv, i, a);
        else { std::fprintf(stderr, "error: unknown argument '%s' (try --help)\n", a); return 2; }
    }

    // derived switches and the aerosol gates (main.c:3255-3380)
    cs.rayleigh_on = cs.pressure_hpa > 0.0 ? 1 : 0;
    cs.aerosol_on = user_aod > 0.0 ? 1 : 0;
    if (!batch_in) {
        if (user_aod > 0.0 && !mie_path) {
            std::fprintf(stderr, "error: --aod-555/--aod-865 > 0 requires --mie <file> (aerosol phase model)\n");
            return 2;
        }
        if (user_aod <= 0.0 && (mie_path || have_aer_h || have_aer_L || trunc_fit || trunc_m || nt_tau)) {
            std::fprintf(stderr, "error: aerosol options require --aod-555/--aod-865 > 0 (aerosol is disabled at "
                                 "the default --aod-555 0)\n");
            return 2;
        }
    }
    if (!have_m_max && cs.aerosol_on) opts.fourier_m_max = 16;
    if (!have_n_layers && cs.aerosol_on) opts.n_layers = 400;
    if (!cs.aerosol_on && !advanced_on()) {
        if (have_n_layers && opts.n_layers != 40) {
            std::fprintf(stderr, "error: --n-layers %d differs from the validated default 40 (convergence knob; "
                                 "requires OCRT_ADVANCED=1)\n", opts.n_layers);
            return 2;
        }
        if (have_m_max && opts.fourier_m_max != 2) {
            std::fprintf(stderr, "error: --m-max %d differs from the Rayleigh-only default 2 (convergence knob; "
                                 "requires OCRT_ADVANCED=1)\n", opts.fourier_m_max);
            return 2;
        }
    }
    if (cs.aerosol_on && !advanced_on()) {
        if (have_aer_L && aer_L_max != 80) {
            std::fprintf(stderr, "error: --aer-l-max %d differs from the validated default 80 (convergence knob; "
                                 "requires OCRT_ADVANCED=1)\n", aer_L_max);
            return 2;
        }
        if (have_n_layers && opts.n_layers != 400) {
            std::fprintf(stderr, "error: --n-layers %d differs from the validated default 400 for aerosol runs "
                                 "(convergence knob; requires OCRT_ADVANCED=1)\n", opts.n_layers);
            return 2;
        }
        if (have_m_max && opts.fourier_m_max != 16) {
            std::fprintf(stderr, "error: --m-max %d differs from the validated default 16 for aerosol runs "
                                 "(convergence knob; requires OCRT_ADVANCED=1)\n", opts.fourier_m_max);
            return 2;
        }
    }
    opts.aerosol_aod = user_aod;
    opts.aerosol_aod_ref_nm = user_aod_ref_nm;
    opts.aerosol_l_max = aer_L_max;
    if (!have_surface) {
        std::fprintf(stderr, "error: --surface is required (black | flat | black_fresnel_ocean | ocean)\n");
        return 2;
    }
    // run mode (main.c:3486-3520)
    const bool single = have_vza && have_raa;
    const bool grid_req = lut_mode || have_vza_step || have_raa_step;
    if (have_vza != have_raa) {
        std::fprintf(stderr, "error: single-geometry mode requires BOTH --vza and --raa (got only one); omit both "
                             "for LUT grid mode\n");
        return 2;
    }
    if (single && grid_req) {
        std::fprintf(stderr, "error: conflicting mode flags: single geometry (--vza/--raa) vs LUT grid "
                             "(--output-full-grid/--lut-vza-step/--lut-raa-step). Pick one.\n");
        return 2;
    }
    if (single && batch_in) {
        std::fprintf(stderr, "error: conflicting mode flags: single geometry vs batch (--batch/--batch-full-grid). "
                             "Pick one.\n");
        return 2;
    }
    if (!single && !batch_in) {
        lut_mode = true;
        cs.vza_deg = 0.0;
        cs.raa_deg = 0.0;
    }
    if (single) opts.view_as_node = 1;
    std::string lut_name;
    if (lut_mode && !batch_in) {
        if (resolve_csv_name(lut_output, lut_name)) return 2;
        std::fprintf(stderr, "# LUT grid output: %s (vza step %.4g deg, raa step %.4g deg, vza max %.4g deg)\n",
                     lut_name.c_str(), lut_vza_step, lut_raa_step, lut_vza_max);
    }
    if (batch_in) {
        if (!batch_out) { std::fprintf(stderr, "error: --batch requires --output\n"); return 2; }
        if (lut_mode) {
            opts.lut_enable = 1;
            opts.lut_output_dir = lut_output;
            opts.lut_vza_step = lut_vza_step;
            opts.lut_raa_step = lut_raa_step;
            opts.lut_vza_max = lut_vza_max;
            if (!lut_output) { std::fprintf(stderr, "error: --lut with --batch requires --lut-output <dir>\n"); return 2; }
            if (mkdir(lut_output, 0755) != 0 && errno != EEXIST) {
                std::fprintf(stderr, "error: cannot create LUT output dir '%s'\n", lut_output);
                return 1;
            }
        }
        // the batch path never applies gas absorption (main.c:3534-3566)
        return run_batch(batch_in, batch_out, opts, cs, mie_path != nullptr);
    }
    if (!(have_sza && have_wl)) {
        std::fprintf(stderr, "error: --sza and --wavelength are required in every mode\n");
        usage(stderr, argv[0]);
        return 2;
    }
    if (!(std::isfinite(cs.wavelength_nm) && cs.wavelength_nm >= 330.0 - 1e-9 && cs.wavelength_nm <= 1100.0 + 1e-9)) {
        std::fprintf(stderr, "error: wavelength %.12g nm is outside OCRT elastic support 330-1100 nm\n",
                     cs.wavelength_nm);
        return 2;
    }
    if (cs.surface == RT_SURFACE_BLACK_FRESNEL_OCEAN && !have_wind) {
        std::fprintf(stderr, "error: --wind-speed is required for --surface ocean/black_fresnel_ocean (explicit 0 "
                             "= flat Fresnel)\n");
        return 2;
    }
    // gas absorption, always on: AFGL profile and cross sections from the
    // fixed data directories under the working directory (main.c:3592-3613);
    // six zero columns give the absorption-free atmosphere
    rt_absorption_t abs_state;
    const int arc = rt_absorption_init("inputs/afgl_atm", "inputs/xsec", (rt_afgl_profile_t)atm_profile, gas_col,
                                       &abs_state);
    if (arc < 0) {
        std::fprintf(stderr, "error: rt_absorption_init failed rc=%d\n", arc);
        return 1;
    }
    opts.abs_state = &abs_state;
    std::fprintf(stderr, "rt_absorption (SOS-integrated): wl=%.1fnm tau_abs_col=%.4e\n", cs.wavelength_nm,
                 rt_absorption_tau_total(&abs_state, cs.wavelength_nm));
    static const char *gas_names[6] = {"h2o", "o3", "o2", "co2", "no2", "ch4"};
    for (int g = 0; g < RT_N_GAS; ++g)
        std::fprintf(stderr, "  %s: column=%.3e mol/cm2%s  tau_abs_col=%.4e\n", gas_names[g],
                     abs_state.column_eff[g], abs_state.column_overridden[g] ? " (overridden)" : "",
                     rt_absorption_tau_per_gas(&abs_state, g, cs.wavelength_nm));
    if (mie_path) {
        std::fprintf(stderr, "error: --mie: the Mie-table aerosol pipeline is not served by this build; use "
                             "rt_solve_case_pol_aerosol with Greek coefficients\n");
        return 2;
    }
    rt_result_t res;
    std::memset(&res, 0, sizeof res);
    if (lut_mode) {
        if (lut_vza_step <= 0 || lut_raa_step <= 0 || lut_vza_max <= 0 || lut_vza_max >= 90.0) {
            std::fprintf(stderr, "error: invalid LUT grid step/max\n");
            return 2;
        }
        Grid g(lut_vza_max, lut_vza_step, lut_raa_step);
        const int rc = rt_solve_case_pol_lut(&cs, &opts, nullptr, &res, &g.lut);
        rt_absorption_free(&abs_state);
        if (rc != 0) { std::fprintf(stderr, "error: rt_solve_case_pol_lut rc=%d\n", rc); return 1; }
        FILE *fp = std::fopen(lut_name.c_str(), "w");
        if (!fp) { std::fprintf(stderr, "error: cannot open LUT output '%s'\n", lut_name.c_str()); return 1; }
        g.write(fp, cs);
        std::fclose(fp);
        const int n = g.lut.n_vza * g.lut.n_raa;
        std::fprintf(stderr, "rt_lut: emitted %dx%d = %d directions to %s\n", g.lut.n_vza, g.lut.n_raa, n,
                     lut_name.c_str());
        std::fprintf(stderr, "  direct_transmittance=%.6e  irradiance_transmittance=%.6e (case-level scalars, 0+ "
                             "side)\n", g.lut.T_dir_dn, g.lut.T_dir_dn + g.lut.T_diff_dn_hemi);
        return 0;
    }
    const int rc = rt_solve_case_pol(&cs, &opts, &res);
    rt_absorption_free(&abs_state);
    if (rc != 0) { std::fprintf(stderr, "error: rt_solve_case_pol rc=%d\n", rc); return 1; }
    res.aerosol_aod_ref = user_aod;
    res.aerosol_aod_ref_nm = user_aod_ref_nm;
    print_single(cs, res);
    return 0;
}
