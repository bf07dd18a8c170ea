// Reference-compatible entry points (include/ocrt_rt.h): rt_solve_case_pol,
// rt_solve_case_pol_aerosol and rt_solve_case_pol_lut of OCRT's atmosphere
// path (code/OCRT_C/src/rt_solver.c:1439-1460, driver :618-1437), served by a
// one-instance FP64 plan of the batched GPU solver, plus ocrt_rt_solve_batch,
// which puts a whole list of cases into one plan.  Host code only.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include <memory>

#include "ocrt_amd.h"
#include "ocrt_rt.h"

// gas.hip: the solver's copy of a reference absorption state (nullptr when
// the state is not initialised or malformed)
ocrt_gas *ocrt_gas_from_state(const rt_absorption_t *st);

namespace {

constexpr double kPi = 3.14159265358979323846;

// ---- Bodhaine et al. (1999) Rayleigh optical depth (rt_rayleigh.c:46-240) --
// Normal gravity of the 1967/1980 international formula with the free-air
// gradient, m s^-2.
double gravity(double lat_deg, double alt_m)
{
    const double phi = lat_deg * kPi / 180.0;
    const double s1 = std::sin(phi), s2 = std::sin(2.0 * phi);
    return 9.780327 * (1.0 + 0.0053024 * s1 * s1 - 0.0000058 * s2 * s2) - 3.086e-6 * alt_m;
}

// King factor of dry air: species factors of Bates (1984) mixed by molar
// percentage, CO2 given in ppm (Bodhaine 1999, Eqs. 5, 6, 23).
double king_factor(double wl_nm, double co2_ppm)
{
    const double inv2 = 1.0 / ((wl_nm / 1000.0) * (wl_nm / 1000.0));
    const double fn2 = 1.034 + 3.17e-4 * inv2;
    const double fo2 = 1.096 + 1.385e-3 * inv2 + 1.448e-4 * inv2 * inv2;
    const double co2 = co2_ppm / 1.0e4;
    return (78.084 * fn2 + 20.946 * fo2 + 0.934 * 1.0 + co2 * 1.15) / (78.084 + 20.946 + 0.934 + co2);
}

double rayleigh_tau(double wl_nm, double p_hpa, double lat_deg, double alt_m, double co2_ppm)
{
    // refractive index of standard air (Peck & Reeder 1972) with the CO2
    // correction (Bodhaine 1999, Eqs. 18-19)
    const double wl_um = wl_nm / 1000.0;
    const double s2 = 1.0 / (wl_um * wl_um);
    const double n300 = (8060.51 + 2480990.0 / (132.274 - s2) + 17455.7 / (39.32957 - s2)) * 1.0e-8;
    const double nm1 = n300 * (1.0 + 0.54 * (co2_ppm * 1.0e-6 - 0.0003));
    // cross section per molecule (Eq. 2) at the standard number density
    const double avogadro = 6.0221367e23;
    const double Ns = (avogadro / 22.4141) * (273.15 / 288.15) * 1.0e-3;
    // n^2 - 1 = (n - 1)(n + 1) without the cancellation of forming n^2 first
    const double a = nm1 * (2.0 + nm1), b = 3.0 + a;
    const double wl_cm = wl_nm * 1.0e-7;
    const double wl4 = wl_cm * wl_cm * wl_cm * wl_cm;
    const double sigma = 24.0 * kPi * kPi * kPi * (a * a) / (wl4 * Ns * Ns * (b * b)) * king_factor(wl_nm, co2_ppm);
    // column above the site with gravity at the mass-weighted altitude
    const double ma = 15.0556 * (co2_ppm * 1.0e-6) + 28.9595;
    const double g = 100.0 * gravity(lat_deg, 0.73737 * alt_m + 5517.56);
    return sigma * (p_hpa * 1000.0) * avogadro / (ma * g);
}

void say(const char *what)
{
    std::fprintf(stderr, "ocrt: %s is not served by this solver\n", what);
}

// One case of the reference API as an instance descriptor; the checks and
// their order follow rt_solver.c:623-661.
struct GasDeleter {
    void operator()(ocrt_gas *g) const { ocrt_gas_destroy(g); }
};
using GasPtr = std::unique_ptr<ocrt_gas, GasDeleter>;

struct Prep {
    ocrt_instance_t in{};
    std::vector<double> greek;
    std::vector<double> bsrc;   // the bottom source, [3][m + 1][n]
    double tau_R = 0.0;
    bool zero_out = false;   // refused after the reference has cleared its output
};

// gas: the converted opts->abs_state (nullptr when there is none)
int prepare(const rt_case_t *cs, const rt_options_t *opts, const rt_aerosol_input_t *aer, const ocrt_gas *gas,
            Prep &p)
{
    const double wl = cs->wavelength_nm;
    if (!(std::isfinite(wl) && wl >= 330.0 - 1e-9 && wl <= 1100.0 + 1e-9)) return OCRT_ERR_ARGS;
    if (cs->sza_deg < 0.0 || cs->sza_deg >= 90.0) return OCRT_ERR_ARGS;
    if (cs->vza_deg < 0.0 || cs->vza_deg >= 90.0) return OCRT_ERR_ARGS;
    if (opts->fourier_m_max < 0 || opts->fourier_m_max > 32) return OCRT_ERR_ARGS;
    if (opts->n_mu < 2 || opts->n_layers < 2) return OCRT_ERR_ARGS;
    if (aer) {
        if (aer->tau_a < 0.0) return OCRT_ERR_ARGS;
        if (aer->ssa_a < 0.0 || aer->ssa_a > 1.0) return OCRT_ERR_ARGS;
        if (aer->L_max < 2) return OCRT_ERR_ARGS;
        if (!aer->betal_aer || !aer->gammal_aer || !aer->alphal_aer || !aer->zetal_aer) return OCRT_ERR_ARGS;
    }
    // what this build does not serve
    if (cs->surface == RT_SURFACE_OCEAN) { say("surface ocean (coupled in-water RT)"); return OCRT_ERR_ARGS; }
    if (cs->surface < RT_SURFACE_BLACK || cs->surface > RT_SURFACE_OCEAN) return OCRT_ERR_ARGS;
    if (opts->integration_method != RT_INTEGRATION_METHOD_LINEAR) { say("integration_method CONSTANT"); return OCRT_ERR_ARGS; }
    if (!opts->view_as_node) { say("view_as_node 0"); return OCRT_ERR_ARGS; }
    if (opts->abs_state && !gas) return OCRT_ERR_ARGS;
    // the ocean-coupling bottom source (rt_solver.c:687-706, 949-977); the
    // slices in opts->sos are never read there (the driver overwrites them)
    const bool ext_any = opts->ext_bottom_per_m_I || opts->ext_bottom_per_m_Q || opts->ext_bottom_per_m_U;
    const bool ext_all = opts->ext_bottom_per_m_I && opts->ext_bottom_per_m_Q && opts->ext_bottom_per_m_U;
    if (ext_any && !ext_all) { p.zero_out = true; return OCRT_ERR_BOTTOM_SHAPE; }
    if (opts->rayleigh_model != RT_RAYLEIGH_MODEL_BODHAINE_1999) return OCRT_ERR_ARGS;
    const bool aerosol = aer && aer->tau_a > 0.0;
    // value kernel: the reference reads only the P11 angle table on this path
    // (rt_solver.c:922-925 with :190-240, rt_ipss.c:81-146)
    const bool value_kernel = aerosol && aer->use_value_kernel && aer->P11_phase;
    if (value_kernel && (aer->n_ang_phase < 2 || !aer->theta_phase)) return OCRT_ERR_ARGS;
    // Rayleigh optical depth (rt_solver.c:648-658)
    if (cs->rayleigh_on == 0) {
        p.tau_R = 0.0;
    } else if (opts->tau_r_source == RT_TAU_R_FROM_INPUT) {
        if (cs->tau_R_input <= 0.0) return OCRT_ERR_ARGS;
        p.tau_R = cs->tau_R_input;
    } else {
        p.tau_R = rayleigh_tau(wl, cs->pressure_hpa, 45.0, 0.0, 360.0);
    }
    // an atmosphere with neither Rayleigh nor aerosol is refused by the
    // reference's medium build, after the output was cleared (rt_atm.c:182)
    if (!aerosol && !(p.tau_R > 0.0)) {
        p.zero_out = true;
        return OCRT_ERR_ARGS;
    }
    ocrt_instance_t &d = p.in;
    d.sza_deg = cs->sza_deg;
    d.vza_deg = cs->vza_deg;
    d.raa_deg = cs->raa_deg;
    d.wavelength_nm = wl;
    d.surface = cs->surface;
    d.decouple_sunglint = cs->decouple_sunglint;
    d.n_water = cs->n_water;
    d.wind_speed = cs->wind_speed;
    d.sigma_type = cs->sigma_type;
    d.q_convention = cs->q_convention;
    d.tau_R = p.tau_R;
    d.depol = 0.0279;
    d.aer_h_km = opts->aer_h_km;
    d.n_mu = opts->n_mu;
    d.n_layers = opts->n_layers;
    d.m_max = opts->fourier_m_max;
    d.max_orders = opts->sos.max_iterations;
    d.tolerance = opts->sos.tolerance;
    d.ipss = opts->pssa ? 1 : 0;
    d.layers = nullptr;
    d.gas = gas;
    d.bottom_source_only = opts->bottom_source_only ? 1 : 0;
    if (ext_all) {
        // one [3][m + 1][n] block from the three per-component arrays
        const int nb = opts->ext_bottom_n_mu, mb = opts->ext_bottom_m_max;
        // the source must be shaped on the solve's ring before it is read
        if (nb < 1 || mb < 0 || nb != ocrt_ring(d.n_mu, d.vza_deg, nullptr, nullptr)) {
            p.zero_out = true;
            return OCRT_ERR_BOTTOM_SHAPE;
        }
        const size_t blk = (size_t)(mb + 1) * nb;
        p.bsrc.resize(3 * blk);
        std::memcpy(&p.bsrc[0], opts->ext_bottom_per_m_I, blk * sizeof(double));
        std::memcpy(&p.bsrc[blk], opts->ext_bottom_per_m_Q, blk * sizeof(double));
        std::memcpy(&p.bsrc[2 * blk], opts->ext_bottom_per_m_U, blk * sizeof(double));
        d.bottom_source = p.bsrc.data();
        d.bottom_n_mu = nb;
        d.bottom_m_max = mb;
    }
    if (aerosol) {
        const int Lm = aer->L_max;
        p.greek.resize(4 * (size_t)(Lm + 1));
        std::memcpy(&p.greek[0], aer->betal_aer, (Lm + 1) * sizeof(double));
        std::memcpy(&p.greek[Lm + 1], aer->gammal_aer, (Lm + 1) * sizeof(double));
        std::memcpy(&p.greek[2 * (Lm + 1)], aer->alphal_aer, (Lm + 1) * sizeof(double));
        std::memcpy(&p.greek[3 * (Lm + 1)], aer->zetal_aer, (Lm + 1) * sizeof(double));
        d.tau_a = aer->tau_a;
        d.ssa_a = aer->ssa_a;
        d.L_max = Lm;
        d.greek = p.greek.data();
        if (value_kernel) {
            d.n_ang_phase = aer->n_ang_phase;
            d.theta_phase = aer->theta_phase;
            d.p11_phase = aer->P11_phase;
        }
    } else {
        d.tau_a = 0.0;
        d.ssa_a = 1.0;
        d.L_max = 0;
        d.greek = nullptr;
    }
    return OCRT_OK;
}

// rt_result_t of one solved case (rt_solver.c:1102-1278).
// The per-mode TOA export of rt_options_t (toa_view_per_m_*, rt_types.h:745-751)
// is written only when export_modes is set (single-case calls).
void fill_result(const ocrt_result_t &r, const double *modes, const int *orders, const double *resid, int m_max,
                 const rt_options_t *opts, bool export_modes, rt_result_t *out)
{
    out->rho_I = r.rho_I; out->rho_Q = r.rho_Q; out->rho_U = r.rho_U;
    out->I_TOA = r.I_TOA; out->Q_TOA = r.Q_TOA; out->U_TOA = r.U_TOA;
    out->tau_R_total = r.tau_R_total;
    out->n_orders_used = r.n_orders_used;
    out->converged = r.converged;
    for (int m = 0; m <= m_max && m < 33; ++m) {
        out->sos_orders_per_m[m] = orders[m];
        out->sos_residual_per_m[m] = resid[m];
    }
    out->T_dir_dn = r.T_dir_dn;
    out->T_diff_dn_hemi = r.T_diff_dn_hemi;
    out->T_total_dn_hemi = r.T_total_dn_hemi;
    out->T_diff_dn_dir = r.T_diff_dn_dir;
    out->T_sg_up_dir = r.T_sg_up_dir;
    out->T_total_up_dir = r.T_total_up_dir;
    out->rho_atm_path_I = r.rho_I; out->rho_atm_path_Q = r.rho_Q; out->rho_atm_path_U = r.rho_U;
    out->rho_glint_direct_I = r.rho_glint_direct_I;
    out->rho_glint_direct_Q = r.rho_glint_direct_Q;
    out->rho_glint_direct_U = r.rho_glint_direct_U;
    out->T_dir_up_view = r.T_dir_up_view;
    out->T_diff_up_view = NAN;
    out->T_total_up_view = NAN;
    out->I_TOA_water_signal = NAN;
    out->T_up_rt_valid = 0;
    if (export_modes && opts->toa_view_per_m_I && opts->toa_view_per_m_Q && opts->toa_view_per_m_U)
        for (int m = 0; m <= m_max; ++m) {
            opts->toa_view_per_m_I[m] = modes[(long long)m * 3 + 0];
            opts->toa_view_per_m_Q[m] = modes[(long long)m * 3 + 1];
            opts->toa_view_per_m_U[m] = modes[(long long)m * 3 + 2];
        }
}

int solve(const rt_case_t *cs, const rt_options_t *opts_in, const rt_aerosol_input_t *aer, rt_result_t *out,
          rt_lut_grid_out_t *lut)
{
    if (!cs || !out) return OCRT_ERR_ARGS;
    const rt_options_t defaults = rt_options_default();
    const rt_options_t *opts = opts_in ? opts_in : &defaults;
    GasPtr gas(opts->abs_state ? ocrt_gas_from_state((const rt_absorption_t *)opts->abs_state) : nullptr);
    Prep p;
    int rc = prepare(cs, opts, aer, gas.get(), p);
    if (rc && p.zero_out) std::memset(out, 0, sizeof(*out));
    if (rc) return rc;
    if (lut) {
        if (lut->n_vza <= 0 || !lut->vza_deg || lut->n_raa <= 0 || !lut->raa_deg || !lut->rho_I || !lut->rho_Q ||
            !lut->rho_U)
            return OCRT_ERR_ARGS;
        for (int iv = 0; iv < lut->n_vza; ++iv)
            if (lut->vza_deg[iv] < 0.0 || lut->vza_deg[iv] >= 90.0) return OCRT_ERR_ARGS;
    }
    std::memset(out, 0, sizeof(*out));
    ocrt_plan *plan = nullptr;
    rc = ocrt_plan_create(&p.in, 1, OCRT_PREC_FP64, 0, &plan);
    if (rc) return rc;
    if (lut) rc = ocrt_plan_set_lut(plan, lut->n_vza, lut->vza_deg, lut->n_raa, lut->raa_deg);
    if (!rc) rc = ocrt_plan_execute(plan, nullptr);
    const int M = ocrt_plan_mmax(plan) + 1;
    ocrt_result_t r{};
    std::vector<double> modes((size_t)M * 3), resid(M);
    std::vector<int> orders(M);
    if (!rc) rc = ocrt_plan_results(plan, &r, modes.data(), orders.data());
    if (!rc) rc = ocrt_plan_residuals(plan, resid.data());
    if (!rc) rc = r.status;
    if (!rc) fill_result(r, modes.data(), orders.data(), resid.data(), p.in.m_max, opts, true, out);
    if (!rc && lut) {
        const int nv = lut->n_vza, nr = lut->n_raa;
        std::vector<double> grid((size_t)nv * nr * 6), raw((size_t)nv * nr * 3), vpm((size_t)nv * M * 3);
        double scal[2];
        rc = ocrt_plan_lut_results(plan, grid.data(), scal);
        if (!rc) rc = ocrt_plan_lut_extras(plan, raw.data(), vpm.data());
        if (!rc) {
            lut->T_dir_dn = scal[0];
            lut->T_diff_dn_hemi = scal[1];
            for (int g = 0; g < nv * nr; ++g) {
                const double *e = &grid[(size_t)g * 6];
                lut->rho_I[g] = e[0];
                lut->rho_Q[g] = e[1];
                lut->rho_U[g] = e[2];
                if (lut->T_diff_dn_dir) lut->T_diff_dn_dir[g] = e[3];
                if (lut->T_sg_up_dir) lut->T_sg_up_dir[g] = e[4];
                if (lut->T_total_up_dir) lut->T_total_up_dir[g] = e[5];
                if (lut->raw_I) {
                    lut->raw_I[g] = raw[(size_t)g * 3 + 0];
                    lut->raw_Q[g] = raw[(size_t)g * 3 + 1];
                    lut->raw_U[g] = raw[(size_t)g * 3 + 2];
                }
            }
            if (lut->view_per_m_I && lut->view_per_m_Q && lut->view_per_m_U && lut->view_m_max >= p.in.m_max) {
                for (int m = 0; m <= p.in.m_max; ++m)
                    for (int iv = 0; iv < nv; ++iv) {
                        const double *v = &vpm[((size_t)iv * M + m) * 3];
                        lut->view_per_m_I[m * nv + iv] = v[0];
                        lut->view_per_m_Q[m * nv + iv] = v[1];
                        lut->view_per_m_U[m * nv + iv] = v[2];
                    }
                lut->view_m_max_filled = p.in.m_max;
            } else {
                lut->view_m_max_filled = -1;
            }
        }
    }
    ocrt_plan_destroy(plan);
    return rc;
}

}  // namespace

extern "C" double ocrt_rayleigh_tau_bodhaine1999(double wavelength_nm, double pressure_hpa, double latitude_deg,
                                                 double altitude_m)
{
    return rayleigh_tau(wavelength_nm, pressure_hpa, latitude_deg, altitude_m, 360.0);
}

extern "C" int rt_solve_case_pol(const rt_case_t *cs, const rt_options_t *opts, rt_result_t *out)
{
    return solve(cs, opts, nullptr, out, nullptr);
}

extern "C" int rt_solve_case_pol_aerosol(const rt_case_t *cs, const rt_options_t *opts,
                                         const rt_aerosol_input_t *aer, rt_result_t *out)
{
    return solve(cs, opts, aer, out, nullptr);
}

extern "C" int rt_solve_case_pol_lut(const rt_case_t *cs, const rt_options_t *opts, const rt_aerosol_input_t *aer,
                                     rt_result_t *out, rt_lut_grid_out_t *lut_out)
{
    return solve(cs, opts, aer, out, lut_out);
}

extern "C" int ocrt_rt_solve_batch(const rt_case_t *cs, const rt_options_t *opts_in, const rt_aerosol_input_t *aer,
                                   int n, rt_result_t *out, int *rc_out)
{
    if (!cs || !out || n < 1) return OCRT_ERR_ARGS;
    const rt_options_t defaults = rt_options_default();
    const rt_options_t *opts = opts_in ? opts_in : &defaults;
    std::vector<Prep> prep(n);
    std::vector<int> rcs(n, 0), live;
    std::vector<ocrt_instance_t> inst;
    GasPtr gas(opts->abs_state ? ocrt_gas_from_state((const rt_absorption_t *)opts->abs_state) : nullptr);
    for (int i = 0; i < n; ++i) {
        rcs[i] = prepare(&cs[i], opts, aer ? &aer[i] : nullptr, gas.get(), prep[i]);
        if (!rcs[i]) live.push_back(i);
        else if (prep[i].zero_out) std::memset(&out[i], 0, sizeof(rt_result_t));
    }
    for (int i : live) {
        prep[i].in.greek = prep[i].greek.empty() ? nullptr : prep[i].greek.data();
        inst.push_back(prep[i].in);
        std::memset(&out[i], 0, sizeof(rt_result_t));
    }
    int first = 0;
    if (!live.empty()) {
        ocrt_plan *plan = nullptr;
        const int nl = (int)live.size();
        int rc = ocrt_plan_create(inst.data(), nl, OCRT_PREC_FP64, 0, &plan);
        int M = 1;
        std::vector<ocrt_result_t> res(nl);
        std::vector<double> modes, resid;
        std::vector<int> orders;
        if (!rc) {
            M = ocrt_plan_mmax(plan) + 1;
            modes.resize((size_t)nl * M * 3);
            resid.resize((size_t)nl * M);
            orders.resize((size_t)nl * M);
            rc = ocrt_plan_execute(plan, nullptr);
        }
        if (!rc) rc = ocrt_plan_results(plan, res.data(), modes.data(), orders.data());
        if (!rc) rc = ocrt_plan_residuals(plan, resid.data());
        if (plan) ocrt_plan_destroy(plan);
        for (int q = 0; q < nl; ++q) {
            const int i = live[q];
            rcs[i] = rc ? rc : res[q].status;
            if (!rcs[i])
                fill_result(res[q], &modes[(size_t)q * M * 3], &orders[(size_t)q * M], &resid[(size_t)q * M],
                            inst[q].m_max, opts, false, &out[i]);
        }
    }
    for (int i = 0; i < n; ++i) {
        if (rc_out) rc_out[i] = rcs[i];
        if (!first && rcs[i]) first = rcs[i];
    }
    return first;
}
