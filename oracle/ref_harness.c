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
