/*
 * ref_harness.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Thin adapter that drives the reference OCRT C solver (compiled from the
 * unmodified sources under code/OCRT_C/src by oracle/Makefile into
 * oracle/_ref/libocrt_ref.so) from the ocrt_instance_t descriptor of
 * include/ocrt_amd.h.  It is the strongest checker available: the golden
 * vectors under tests/golden are produced by it, the C restatement in
 * oracle/ocrt_oracle.c is pinned against it, and bench.py times it as the
 * "reference" CPU baseline.  It is never part of the product path.
 *
 * Entry points used (reference file:line):
 *   rt_solve_case_pol           code/OCRT_C/src/rt_solver.c:1439
 *   rt_solve_case_pol_aerosol   code/OCRT_C/src/rt_solver.c:1445
 *   rt_solve_case_pol_lut       code/OCRT_C/src/rt_solver.c:1455
 *   rt_options_default          code/OCRT_C/src/rt_types.h:431
 * The batch entry mirrors the OpenMP row loop of rt_io_run_batch
 * (code/OCRT_C/src/rt_io.c:199: parallel for schedule(dynamic, 8)).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#include "rt_solver.h"
#include "rt_types.h"
#include "rt_atm.h"
#include "rt_kernel.h"
#include "rt_sos_operator.h"
#include "rt_quadrature.h"
#include "rt_air_water.h"
#include "shared/surface.h"
#include "ocrt_amd.h"

static void fill_case(const ocrt_instance_t *in, rt_case_t *cs, rt_options_t *opts,
                      rt_aerosol_input_t *aer, int *use_aer)
{
    memset(cs, 0, sizeof(*cs));
    *opts = rt_options_default();
    cs->sza_deg = in->sza_deg;
    cs->vza_deg = in->vza_deg;
    cs->raa_deg = in->raa_deg;
    cs->wavelength_nm = in->wavelength_nm;
    cs->surface = (rt_surface_kind_t)in->surface;
    cs->rayleigh_on = (in->tau_R > 0.0) ? 1 : 0;
    cs->pressure_hpa = 1013.25;
    cs->tau_R_input = in->tau_R;
    cs->n_water = in->n_water;
    cs->wind_speed = in->wind_speed;
    cs->sigma_type = in->sigma_type;
    cs->q_convention = in->q_convention;
    cs->decouple_sunglint = in->decouple_sunglint;
    cs->F_sun = RT_F_SOLAR_PI;

    opts->n_mu = in->n_mu;
    opts->n_layers = in->n_layers;
    opts->fourier_m_max = in->m_max;
    opts->max_orders = in->max_orders;
    opts->sos.max_iterations = in->max_orders;
    opts->sos.tolerance = in->tolerance;
    opts->tau_r_source = RT_TAU_R_FROM_INPUT;
    opts->pssa = in->ipss;
    opts->view_as_node = 1;
    opts->aer_h_km = in->aer_h_km;
    if (in->bottom_source) {
        /* the coupled ocean's water-leaving field as the pass-2 bottom source
         * (rt_types.h:722-747) */
        const size_t blk = (size_t)(in->bottom_m_max + 1) * (size_t)in->bottom_n_mu;
        opts->ext_bottom_per_m_I = in->bottom_source;
        opts->ext_bottom_per_m_Q = in->bottom_source + blk;
        opts->ext_bottom_per_m_U = in->bottom_source + 2 * blk;
        opts->ext_bottom_n_mu = in->bottom_n_mu;
        opts->ext_bottom_m_max = in->bottom_m_max;
    }
    opts->bottom_source_only = in->bottom_source_only;

    memset(aer, 0, sizeof(*aer));
    *use_aer = (in->L_max > 0 && in->tau_a > 0.0 && in->greek != NULL);
    if (*use_aer) {
        const int n = in->L_max + 1;
        aer->tau_a = in->tau_a;
        aer->ssa_a = in->ssa_a;
        aer->L_max = in->L_max;
        aer->betal_aer = in->greek;
        aer->gammal_aer = in->greek + n;
        aer->alphal_aer = in->greek + 2 * n;
        aer->zetal_aer = in->greek + 3 * n;
        aer->use_value_kernel = 0;
        if (in->n_ang_phase >= 2 && in->theta_phase && in->p11_phase) {
            /* value kernel: the solver reads only the P11 table (rt_solver.c:190-240,
             * rt_ipss.c:81-146); P12 and P33 are borrowed and unused on this path */
            aer->use_value_kernel = 1;
            aer->n_ang_phase = in->n_ang_phase;
            aer->theta_phase = in->theta_phase;
            aer->P11_phase = in->p11_phase;
        }
        cs->aerosol_on = 1;
    }
}

static void copy_result(const rt_result_t *r, ocrt_result_t *out, int rc)
{
    memset(out, 0, sizeof(*out));
    out->rho_I = r->rho_I; out->rho_Q = r->rho_Q; out->rho_U = r->rho_U;
    out->I_TOA = r->I_TOA; out->Q_TOA = r->Q_TOA; out->U_TOA = r->U_TOA;
    out->tau_R_total = r->tau_R_total;
    out->T_dir_dn = r->T_dir_dn;
    out->T_diff_dn_hemi = r->T_diff_dn_hemi;
    out->T_total_dn_hemi = r->T_total_dn_hemi;
    out->T_diff_dn_dir = r->T_diff_dn_dir;
    out->T_dir_up_view = r->T_dir_up_view;
    out->T_sg_up_dir = r->T_sg_up_dir;
    out->T_total_up_dir = r->T_total_up_dir;
    out->rho_glint_direct_I = r->rho_glint_direct_I;
    out->rho_glint_direct_Q = r->rho_glint_direct_Q;
    out->rho_glint_direct_U = r->rho_glint_direct_U;
    out->kappa = 1.0;
    out->n_orders_used = r->n_orders_used;
    out->converged = r->converged;
    out->status = rc;
}

/* Solve one instance with the reference; toa_modes [m_max+1][3] and orders /
 * residuals [m_max+1] are optional.  Returns the reference return code. */
int ref_solve(const ocrt_instance_t *in, ocrt_result_t *out, double *toa_modes,
              int *orders, double *residuals)
{
    rt_case_t cs; rt_options_t opts; rt_aerosol_input_t aer; int use_aer = 0;
    fill_case(in, &cs, &opts, &aer, &use_aer);
    double pm[3][33];
    memset(pm, 0, sizeof(pm));
    opts.toa_view_per_m_I = pm[0];
    opts.toa_view_per_m_Q = pm[1];
    opts.toa_view_per_m_U = pm[2];
    rt_result_t r;
    memset(&r, 0, sizeof(r));
    const int rc = use_aer ? rt_solve_case_pol_aerosol(&cs, &opts, &aer, &r)
                           : rt_solve_case_pol(&cs, &opts, &r);
    copy_result(&r, out, rc);
    if (rc == 0 && in->m_max >= 0 && in->m_max <= 32) {
        for (int m = 0; m <= in->m_max; ++m) {
            if (toa_modes) {
                toa_modes[3 * m + 0] = pm[0][m];
                toa_modes[3 * m + 1] = pm[1][m];
                toa_modes[3 * m + 2] = pm[2][m];
            }
            if (orders) orders[m] = r.sos_orders_per_m[m];
            if (residuals) residuals[m] = r.sos_residual_per_m[m];
        }
    }
    return rc;
}

/* LUT (full-grid) solve: one SOS, reconstruction at every (vza, raa) pair.
 * Outputs are row-major [iv * n_raa + ir]; any output may be NULL except the
 * three rho arrays. */
int ref_solve_lut(const ocrt_instance_t *in, int n_vza, const double *vza,
                  int n_raa, const double *raa,
                  double *rho_I, double *rho_Q, double *rho_U,
                  double *T_diff_dn_dir, double *T_sg_up_dir, double *T_total_up_dir,
                  double *T_dir_dn, double *T_diff_dn_hemi)
{
    rt_case_t cs; rt_options_t opts; rt_aerosol_input_t aer; int use_aer = 0;
    fill_case(in, &cs, &opts, &aer, &use_aer);
    rt_lut_grid_out_t lut;
    memset(&lut, 0, sizeof(lut));
    lut.n_vza = n_vza; lut.vza_deg = vza;
    lut.n_raa = n_raa; lut.raa_deg = raa;
    lut.rho_I = rho_I; lut.rho_Q = rho_Q; lut.rho_U = rho_U;
    lut.T_diff_dn_dir = T_diff_dn_dir;
    lut.T_sg_up_dir = T_sg_up_dir;
    lut.T_total_up_dir = T_total_up_dir;
    lut.view_m_max = -1;
    rt_result_t r;
    memset(&r, 0, sizeof(r));
    const int rc = rt_solve_case_pol_lut(&cs, &opts, use_aer ? &aer : NULL, &r, &lut);
    if (T_dir_dn) *T_dir_dn = lut.T_dir_dn;
    if (T_diff_dn_hemi) *T_diff_dn_hemi = lut.T_diff_dn_hemi;
    return rc;
}

/* OpenMP batch over independent instances, one instance per thread, exactly
 * the scheduling of rt_io_run_batch (rt_io.c:199).  Returns the number of
 * instances whose solve failed. */
int ref_solve_batch(const ocrt_instance_t *in, int n, ocrt_result_t *out, int n_threads)
{
    int failed = 0;
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 8) reduction(+ : failed)
    for (int i = 0; i < n; ++i) {
        if (ref_solve(&in[i], &out[i], NULL, NULL, NULL) != 0) ++failed;
    }
    return failed;
}

int ref_max_threads(void) { return omp_get_max_threads(); }

/* In-water SOS of the reference's water body, per Fourier mode, for an
 * ocrt_instance_t in-water column (water_tau > 0, top_reflect 0/1/2, no top
 * source).  The medium is filled as the reference fills its in-water
 * atmosphere-form medium (rt_water_rt.c:2104-2236): uniform levels, fixed
 * particle / pure-water albedos, water depolarisation in the molecular slot,
 * GL ring of n_mu nodes with the refracted sun in the solar slot.  Per mode
 * the angular basis, phase tables, operator and primary source are built as
 * in the case driver (rt_solver.c:904-960) and the order loop is the
 * reference's own rt_solver_sos_pol (top_reflect 0), rt_solver_sos_pol_intrefl
 * (1, per-node rt_air_water_R_ww) or rt_solver_sos_pol_intrefl_rough (2,
 * surface_R_ww_coxmunk_fourier_kernel with 128 azimuths), as
 * rt_water_rt.c:5966-6034 calls them.  Outputs (optional):
 *   top_up  [m_max+1][3][n_mu]  upward I, Q, U at the top (k = 0)
 *   bot_dn  [m_max+1][3][n_mu]  downward I, Q, U at the bottom (k = nt)
 *   orders  [m_max+1]
 * Returns 0, or a negative code. */
int ref_water_solve(const ocrt_instance_t *in, double *top_up, double *bot_dn, int *orders)
{
    if (!in || !(in->water_tau > 0.0) || in->top_source) return -1;
    const int nt = in->n_layers, n = in->n_mu, mm = in->m_max;
    const int aer = in->L_max > 0 && in->greek && in->water_omega_p > 0.0;
    const int L = aer ? in->L_max : 0;
    int l_max = mm > 2 ? mm : 2;
    if (L > l_max) l_max = L;
    const double mus = cos(in->sza_deg * 3.14159265358979323846 / 180.0);
    rt_atm_t atm;
    memset(&atm, 0, sizeof(atm));
    rt_legendre_workspace_t ws;
    memset(&ws, 0, sizeof(ws));
    if (rt_atm_alloc(&atm, nt, n) != 0) return -3;
    const double d = in->depol;
    const double ron = 2.0 * (1.0 - d) / (2.0 + d);
    atm.tau_total = in->water_tau;
    atm.depol = d;
    atm.ssa = in->water_omega_w + in->water_omega_p;
    atm.mu_sun = mus;
    atm.rayleigh_model = RT_RAYLEIGH_MODEL_BODHAINE_1999;
    atm.beta0 = 1.0;
    atm.beta2 = 0.5 * ron;
    atm.gamma2 = -ron * sqrt(1.5);
    atm.alpha2 = 3.0 * ron;
    for (int k = 0; k <= nt; ++k) {
        const double hk = (double)k * in->water_tau / (double)nt;
        atm.h[k] = hk;
        atm.ch[k] = 0.5 * in->beam_scale * exp(-hk / mus);
        atm.xdel[k] = aer ? in->water_omega_p : 0.0;
        atm.ydel[k] = in->water_omega_w;
        if (atm.z_km_level) atm.z_km_level[k] = 1.0;
        if (k < nt && atm.tau_abs_layer) atm.tau_abs_layer[k] = 0.0;
    }
    double muq[640], wq[640];
    if (n > 639 || rt_quadrature_gauss_legendre_pos(n, muq, wq) != 0) { rt_atm_free(&atm); return -1; }
    atm.rm[0] = -mus;
    atm.gb[0] = 0.0;
    for (int j = 1; j <= n; ++j) {
        atm.rm[j] = muq[j - 1]; atm.rm[-j] = -muq[j - 1];
        atm.gb[j] = atm.gb[-j] = wq[j - 1];
    }
    if (aer) {
        const size_t q = (size_t)L + 1;
        atm.betal_aer = calloc(q, sizeof(double));
        atm.gammal_aer = calloc(q, sizeof(double));
        atm.alphal_aer = calloc(q, sizeof(double));
        atm.zetal_aer = calloc(q, sizeof(double));
        if (!atm.betal_aer || !atm.gammal_aer || !atm.alphal_aer || !atm.zetal_aer) { rt_atm_free(&atm); return -3; }
        memcpy(atm.betal_aer, in->greek, q * sizeof(double));
        memcpy(atm.gammal_aer, in->greek + q, q * sizeof(double));
        memcpy(atm.alphal_aer, in->greek + 2 * q, q * sizeof(double));
        memcpy(atm.zetal_aer, in->greek + 3 * q, q * sizeof(double));
        atm.aerosol_active = 1;
        atm.L_max = L;
    }
    if (rt_legendre_workspace_alloc(&ws, n, l_max) != 0) { rt_atm_free(&atm); return -3; }
    const int dirs = 2 * n + 1;
    const size_t cnt = (size_t)(nt + 1) * dirs;
    double *buf = calloc(9 * cnt + (size_t)n * n * 9, sizeof(double));
    if (!buf) { rt_legendre_workspace_free(&ws); rt_atm_free(&atm); return -3; }
    double *si = buf, *sq = si + cnt, *su = sq + cnt, *pi_ = su + cnt, *pq = pi_ + cnt, *pu = pq + cnt;
    double *ti = pu + cnt, *tq = ti + cnt, *tu = tq + cnt, *R = tu + cnt;
    rt_solver_sos_options_t so = rt_solver_sos_options_default();
    so.max_iterations = in->max_orders;
    so.tolerance = in->tolerance;
    int rc = 0;
    for (int m = 0; m <= mm && rc == 0; ++m) {
        if (rt_legendre_compute_pol(&ws, &atm, m) != 0 ||
            rt_kernel_phase_fourier_pol(&ws, m, aer ? atm.gammal_aer : NULL) != 0 ||
            rt_kernel_phase_fourier_aerosol_full(&ws, m, aer ? atm.alphal_aer : NULL, aer ? atm.zetal_aer : NULL) != 0 ||
            rt_kernel_phase_fourier(&ws, m, aer ? atm.betal_aer : NULL) != 0 ||
            rt_sos_operator_prepare(&atm, m, &ws) != 0 ||
            rt_solver_primary_source(&atm, m, &ws, si) != 0 ||
            rt_solver_primary_source_pol(&atm, m, &ws, sq, su) != 0) { rc = -1; break; }
        if (rt_solver_integrate(&atm, m, si, RT_INTEGRATION_METHOD_LINEAR, pi_) != 0 ||
            rt_solver_integrate(&atm, m, sq, RT_INTEGRATION_METHOD_LINEAR, pq) != 0 ||
            rt_solver_integrate(&atm, m, su, RT_INTEGRATION_METHOD_LINEAR, pu) != 0) { rc = -2; break; }
        rt_solver_sos_result_t res;
        memset(&res, 0, sizeof(res));
        if (in->top_reflect == 1) {
            for (int j = 1; j <= n; ++j) rt_air_water_R_ww(atm.rm[j], in->n_water, in->q_convention, R + (size_t)(j - 1) * 9);
            rc = rt_solver_sos_pol_intrefl(&atm, m, &ws, pi_, pq, pu, R, &so, ti, tq, tu, &res);
        } else if (in->top_reflect == 2) {
            rc = surface_R_ww_coxmunk_fourier_kernel(&atm.rm[1], n, &atm.rm[1], n, m, 128, in->wind_speed,
                                                     in->sigma_type, in->n_water, in->q_convention, R);
            if (rc == 0) rc = rt_solver_sos_pol_intrefl_rough(&atm, m, &ws, pi_, pq, pu, R, &so, ti, tq, tu, &res);
        } else {
            rc = rt_solver_sos_pol(&atm, m, &ws, pi_, pq, pu, &so, ti, tq, tu, &res);
        }
        if (rc) break;
        if (orders) orders[m] = res.n_orders_used;
        for (int j = 1; j <= n; ++j) {
            const size_t up = (size_t)(j + n), dn = (size_t)nt * dirs + (size_t)(-j + n);
            if (top_up) {
                top_up[((size_t)m * 3 + 0) * n + j - 1] = ti[up];
                top_up[((size_t)m * 3 + 1) * n + j - 1] = tq[up];
                top_up[((size_t)m * 3 + 2) * n + j - 1] = tu[up];
            }
            if (bot_dn) {
                bot_dn[((size_t)m * 3 + 0) * n + j - 1] = ti[dn];
                bot_dn[((size_t)m * 3 + 1) * n + j - 1] = tq[dn];
                bot_dn[((size_t)m * 3 + 2) * n + j - 1] = tu[dn];
            }
        }
    }
    free(buf);
    rt_legendre_workspace_free(&ws);
    rt_atm_free(&atm);
    return rc;
}
