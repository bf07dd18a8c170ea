/*
 * ocrt_rt.h -- reference-compatible entry points of the MI355X SOS solver.
 *
 * rt_solve_case_pol, rt_solve_case_pol_aerosol and rt_solve_case_pol_lut keep
 * the names, argument lists, return codes and struct layouts of OCRT's
 * atmospheric solve path (code/OCRT_C/src/rt_solver.h:84, :119, :207), so a
 * program built against the reference's rt_types.h / rt_solver.h (its CLI
 * main.c, the batch driver rt_io.c, the tests) links against libocrt_amd.so in
 * place of rt_solver.c's atmosphere path and gets its results from the GPU.
 *
 * The structs below are declared here field for field in the reference's
 * order and types (code/OCRT_C/src/rt_types.h:90-112 rt_solver_sos_options_t,
 * :321-482 rt_case_t, :496-651 rt_result_t, :654-754 rt_options_t,
 * :811-846 rt_lut_grid_out_t; code/OCRT_C/src/rt_solver.h:94-114
 * rt_aerosol_input_t).  Layout equality is what makes the swap possible;
 * tests/test_rt_layout.py checks every field offset against the reference.
 * Fields that belong to the coupled ocean (in-water RT, water IOPs) are
 * carried for layout only: this solver is the atmosphere path, and a case
 * that asks for the ocean is rejected with -1.
 *
 * What the atmosphere path reads, and how:
 *   - geometry, wavelength (330..1100 nm), surface, n_water, wind_speed,
 *     sigma_type, q_convention, decouple_sunglint from rt_case_t;
 *   - Rayleigh: rayleigh_on == 0 -> no Rayleigh; tau_r_source FROM_INPUT ->
 *     tau_R_input (must be > 0); AUTO -> Bodhaine et al. (1999) from the
 *     wavelength and pressure_hpa at 45 deg latitude (rt_solver.c:648-658);
 *     depolarisation 0.0279 (rt_solver.c:646);
 *   - n_mu, n_layers, fourier_m_max (0..32), sos.max_iterations,
 *     sos.tolerance, pssa (IPSS), aer_h_km from rt_options_t;
 *   - aerosol: tau_a > 0 with its Greek coefficients (L_max >= 2); with
 *     use_value_kernel and a P11 angle table (n_ang_phase >= 2, theta_phase
 *     ascending in degrees) the aerosol I -> I phase table and the IPSS
 *     aerosol phase come from that table (rt_solver.c:190-240, 922-925;
 *     rt_ipss.c:81-146), as in the reference (P12/P33 tables unread there);
 *   - the ocean-coupling bottom source: ext_bottom_per_m_I/Q/U with
 *     ext_bottom_n_mu equal to the ring size and ext_bottom_m_max >= 0, and
 *     bottom_source_only (rt_types.h:722-747; rt_solver.c:687-706, 949-977;
 *     rt_surface_boundary.c:495-567); a shape mismatch returns -7.
 * Requests this build does not serve return -1 and print one line on stderr:
 * integration_method CONSTANT (a diagnostic of the reference), view_as_node
 * 0 (its diagnostic interpolation route) and surface OCEAN.  Gas absorption
 * is served: abs_state is an rt_absorption_t (below).
 *
 * Every call solves one case in FP64 on the current HIP device, synchronously.
 * Calls from several host threads are safe; each builds its own plan.
 */
#ifndef OCRT_RT_H
#define OCRT_RT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RT_F_SOLAR_PI 3.14159265358979323846

typedef enum {
    RT_SURFACE_BLACK = 0,
    RT_SURFACE_LAMBERT = 1,
    RT_SURFACE_BLACK_FRESNEL_OCEAN = 2,
    RT_SURFACE_FLAT = 3,
    RT_SURFACE_OCEAN = 4
} rt_surface_kind_t;

typedef enum { RT_SOS_ACCELERATION_PLAIN = 0, RT_SOS_ACCELERATION_GEOMETRIC = 1 } rt_sos_acceleration_t;
typedef enum { RT_INTEGRATION_METHOD_LINEAR = 0, RT_INTEGRATION_METHOD_CONSTANT = 1 } rt_integration_method_t;
typedef enum { RT_RAYLEIGH_MODEL_BODHAINE_1999 = 0 } rt_rayleigh_model_t;
typedef enum { RT_TAU_R_AUTO = 0, RT_TAU_R_FROM_INPUT = 1 } rt_tau_r_source_t;

/* Order-loop control. */
typedef struct {
    int                   max_iterations;   /* order cap */
    double                tolerance;        /* stop when |u_o|max / |total|max < tolerance */
    rt_sos_acceleration_t acceleration;     /* only PLAIN is defined */
    int                   save_orders;
    const double         *ext_bottom_I;     /* ocean coupling only */
    const double         *ext_bottom_Q;
    const double         *ext_bottom_U;
    int                   bottom_source_only;
} rt_solver_sos_options_t;

/* One case: geometry, spectrum, surface; then the ocean-side inputs. */
typedef struct {
    double sza_deg, vza_deg, raa_deg, wavelength_nm;
    rt_surface_kind_t surface;
    int    rayleigh_on;        /* 0: no Rayleigh scattering */
    int    aerosol_on;
    double pressure_hpa;       /* scales the AUTO Rayleigh optical depth */
    double tau_R_input;        /* Rayleigh optical depth for tau_r_source FROM_INPUT */
    double n_water;            /* real refractive index of water */
    double wind_speed;         /* m/s; 0 gives the flat Fresnel interface */
    int    sigma_type;         /* slope variance: 0 Nakajima-Tanaka, 1 floor law */
    int    q_convention;       /* 0 (Rs-Rp)/2, 1 Mishchenko (Rp-Rs)/2 */
    int    decouple_sunglint;  /* 1: direct glint left out of rho */
    /* ocean side (layout only) */
    double T_water_C, S_water_gkg, F_sun;
    double a_cdom_440_m_inv, S_cdom_nm_inv, cdom_ref_lambda_nm;
    int    water_input_mode, ccrr_mode, organic_phyto_scattering, water_constituent_model;
    double ccrr_chl_mg_m3, ccrr_min_g_m3;
    int    tsm_species, organic_phyto_group;
    double detritus_a440_m_inv, detritus_slope_nm_inv;
    const char *ccrr_phase_moments_path;
    const char *ccrr_particle_phase_lut_path;
    const char *ccrr_particle_phase_case_id;
    double ccrr_particle_phase_wavelength_nm;
    int    ccrr_particle_phase_lmax, ccrr_particle_phase_nphi;
    int    water_phase_kernel;
    int    fixed_bulk_iop_mode;
    double fixed_a_total_m_inv, fixed_b_total_m_inv, fixed_bb_total_m_inv;
    int    fixed_bulk_lmax, fixed_bulk_phase_model, fixed_bulk_phase_nphi;
    double fixed_bulk_ff_n, fixed_bulk_ff_mu;
    const char *fixed_bulk_phase_lut_path;
    const char *water_mie_phase_path;
    int    water_mie_moment_mode, water_mie_truncation_mode, water_mie_ss_mode, water_mie_moment_n_mu;
    const char *fixed_bulk_phase_case_id;
    double fixed_bulk_phase_wavelength_nm;
    const char *scalar_ff_iop_path;
    double scalar_ff_bb_over_b, scalar_ff_refractive_index, scalar_ff_mu;
    int    n_layers_water_override;
    double tau_max_target_override, water_max_tau_override, water_max_z_max_m_override;
    double water_depth_bottom_tol_override, water_layer_dtau_target_override;
    int    n_mu_water_override, n_mu_water_sky, water_m_max_override;
    int    water_max_orders_override;
    int    debug_high_particle_water_max_orders;
    const double *water_view_vza_list;
    int    n_water_view_vza;
    int    water_shared_grid;
    int    water_view_as_node;
} rt_case_t;

/* Results of one case (atmosphere path fills the first blocks; the
 * ocean-side diagnostics stay zero). */
typedef struct {
    double rho_I, rho_Q, rho_U;
    int    n_orders_used;            /* max over modes */
    int    converged;                /* AND over modes */
    double I_TOA, Q_TOA, U_TOA;
    double tau_R_total;
    int    sos_orders_per_m[33];
    double sos_residual_per_m[33];
    double T_dir_dn, T_diff_dn_hemi, T_diff_dn_dir, T_sg_up_dir, T_total_up_dir;
    double rho_atm_path_I, rho_atm_path_Q, rho_atm_path_U;
    double rho_water_direct_I, rho_water_direct_Q, rho_water_direct_U;
    double rho_water_sky_I, rho_water_sky_Q, rho_water_sky_U;
    double rho_water_total_I, rho_water_total_Q, rho_water_total_U;
    double rho_glint_direct_I, rho_glint_direct_Q, rho_glint_direct_U;
    double T_total_dn_hemi, T_dir_up_view, T_diff_up_view, T_total_up_view;
    double I_TOA_water_signal;
    int    T_up_rt_valid;
    double aerosol_aod_ref, aerosol_aod_ref_nm, aerosol_aod_band, aerosol_extinction_ratio;
    /* ocean side (layout only) */
    double Lu_0minus_view, Lu_0plus_view, Qu_0minus_view, Qu_0plus_view, Uu_0minus_view, Uu_0plus_view;
    double Ed_0minus_water, Ed_0plus_air, Eu_0minus_water;
    double Ed_0plus_direct, Ed_0plus_diffuse, Ed_0minus_direct, Ed_0minus_diffuse;
    double Eu_0minus_direct, Eu_0minus_diffuse;
    double r_rs_0minus, r_rs_0minus_Q, r_rs_0minus_U, R_rs_0plus, R_rs_0plus_Q, R_rs_0plus_U;
    double Kd_0minus, Ku_0minus;
    double a_w_used, b_w_used, bb_w_used, a_cdom_used, a_pig_used, a_chl_used, b_pig_used, bb_pig_used;
    double a_min_used, b_min_used, bb_min_used, a_total_used, b_total_used, bb_total_used;
    double omega_water, tau_max_water;
    double rho_I_glint, rho_Q_glint, rho_U_glint;
} rt_result_t;

/* Solver options. */
typedef struct {
    int    n_mu;               /* Gauss-Legendre nodes per hemisphere */
    int    n_layers;
    int    max_orders;
    int    l_max;
    rt_rayleigh_model_t rayleigh_model;
    rt_integration_method_t integration_method;
    rt_solver_sos_options_t sos;
    int    fourier_m_max;
    rt_tau_r_source_t tau_r_source;
    int    pssa;               /* 1: IPSS spherical-shell correction */
    int    view_as_node;       /* 1: the view direction is a zero-weight ring node */
    int    vector_mode;
    const char *aerosol_mie_path;
    double aerosol_aod, aerosol_aod_ref_nm;
    int    aerosol_l_max;
    double aerosol_theta_cut_deg;
    int    aerosol_delta_m_N, aerosol_apply_nt_tau;
    int    output_mode_debug;
    int    lut_enable;
    const char *lut_output_dir;
    double lut_vza_step, lut_raa_step, lut_vza_max;
    const void *abs_state;     /* gas absorption: an rt_absorption_t, or NULL */
    double aer_h_km;           /* aerosol scale height */
    double conv_tol;
    int    verbose;
    const double *ext_bottom_per_m_I, *ext_bottom_per_m_Q, *ext_bottom_per_m_U;
    int    ext_bottom_n_mu, ext_bottom_m_max, bottom_source_only;
    double *toa_view_per_m_I, *toa_view_per_m_Q, *toa_view_per_m_U;   /* optional [m_max+1] exports */
} rt_options_t;

/* Aerosol of one case as Greek coefficients. */
typedef struct {
    double tau_a, ssa_a;
    int    L_max;
    const double *betal_aer, *gammal_aer, *alphal_aer, *zetal_aer;   /* [L_max+1] each */
    int    use_value_kernel, n_ang_phase;
    const double *theta_phase, *P11_phase, *P12_phase, *P33_phase;
} rt_aerosol_input_t;

/* Full-grid output: rho and transmittances at every (vza, raa), row-major
 * [iv * n_raa + ir]; every pointer but rho_* may be NULL. */
typedef struct {
    int n_vza;
    const double *vza_deg;
    int n_raa;
    const double *raa_deg;
    double *rho_I, *rho_Q, *rho_U;
    double *raw_I, *raw_Q, *raw_U;                  /* TOA Stokes before rho, IPSS and glint */
    double *view_per_m_I, *view_per_m_Q, *view_per_m_U;   /* [m * n_vza + iv] */
    int     view_m_max;                             /* capacity, inclusive; -1 unused */
    int     view_m_max_filled;                      /* set by the solver */
    double *T_diff_dn_dir, *T_sg_up_dir, *T_total_up_dir;
    double  T_dir_dn, T_diff_dn_hemi;               /* set by the solver */
} rt_lut_grid_out_t;

/* Defaults of the reference (code/OCRT_C/src/rt_types.h:760-798). */
static inline rt_options_t rt_options_default(void)
{
    rt_options_t o;
    char *b = (char *)&o;
    for (size_t i = 0; i < sizeof(o); ++i) b[i] = 0;
    o.n_mu = 24;
    o.n_layers = 40;
    o.max_orders = 20;
    o.l_max = -1;
    o.rayleigh_model = RT_RAYLEIGH_MODEL_BODHAINE_1999;
    o.integration_method = RT_INTEGRATION_METHOD_LINEAR;
    o.sos.max_iterations = 20;
    o.sos.tolerance = 1.0e-7;
    o.sos.acceleration = RT_SOS_ACCELERATION_PLAIN;
    o.fourier_m_max = 2;
    o.tau_r_source = RT_TAU_R_AUTO;
    o.view_as_node = 1;
    o.vector_mode = 1;
    o.aerosol_aod_ref_nm = 555.0;
    o.aerosol_l_max = 80;
    o.lut_vza_step = 2.5;
    o.lut_raa_step = 5.0;
    o.lut_vza_max = 85.0;
    o.aer_h_km = 2.0;
    o.ext_bottom_m_max = -1;
    o.conv_tol = 1.0e-6;
    return o;
}

/* ---- gas absorption state (code/OCRT_C/src/rt_absorption.h:24-86) ------
 * rt_options_t.abs_state points at an rt_absorption_t: an AFGL standard
 * atmosphere (levels ascending from the ground; mixing ratios in ppmv) and
 * one cross-section table per gas (sigma in cm^2 per molecule, layered
 * [n_layers][n_wl] on the first n_layers AFGL levels, or a single layer),
 * with the effective column of each gas (the AFGL default or an override in
 * molecules cm^-2).  The solver adds the layer absorption optical depth
 * sum_gas sigma(z_mid, lambda) N_gas(layer) to the medium and rescales the
 * scattering fractions (rt_atm.c:437-548). */
typedef enum {
    RT_AFGL_USSTD76 = 0, RT_AFGL_TROPICAL = 1, RT_AFGL_MLSUMM = 2, RT_AFGL_MLWINT = 3,
    RT_AFGL_SASUMM = 4, RT_AFGL_SAWINT = 5, RT_AFGL_USERDEF = 6
} rt_afgl_profile_t;

#define RT_N_GAS 6
enum { RT_GAS_H2O = 0, RT_GAS_O3, RT_GAS_O2, RT_GAS_CO2, RT_GAS_NO2, RT_GAS_CH4 };

typedef struct {
    int n_levels;
    double *z_km, *P_mbar, *T_K, *n_total_cm3;
    double *mr[RT_N_GAS];                /* ppmv per level */
    double column_default[RT_N_GAS];     /* molecules cm^-2, trapezoid over the levels */
} rt_afgl_atm_t;

typedef struct {
    int n_wl;
    int n_layers;
    double *wl_nm;                       /* [n_wl], ascending */
    double *sigma_cm2;                   /* [n_layers][n_wl] */
    double T_ref_K, P_ref_atm;
    int available;
} rt_xsec_t;

typedef struct {
    rt_afgl_atm_t atm;
    rt_xsec_t xsec[RT_N_GAS];
    double column_eff[RT_N_GAS];
    int column_overridden[RT_N_GAS];
    int initialized;
} rt_absorption_t;

/* Read inputs/afgl_atm/afgl_<profile>.dat and inputs/xsec/xsec_<gas>.dat
 * style files (the reference's formats, rt_absorption.c:40-331);
 * gas_column_override[g] < 0 keeps the AFGL column.  A missing cross-section
 * file leaves that gas out.  0 on success. */
int rt_absorption_init(const char *afgl_dir, const char *xsec_dir, rt_afgl_profile_t profile,
                       const double gas_column_override[RT_N_GAS], rt_absorption_t *abs_state);
void rt_absorption_free(rt_absorption_t *abs_state);
/* Column absorption optical depth of one gas / of all gases. */
double rt_absorption_tau_per_gas(const rt_absorption_t *abs_state, int gas_idx, double wavelength_nm);
double rt_absorption_tau_total(const rt_absorption_t *abs_state, double wavelength_nm);

int rt_solve_case_pol(const rt_case_t *cs, const rt_options_t *opts, rt_result_t *out);
int rt_solve_case_pol_aerosol(const rt_case_t *cs, const rt_options_t *opts,
                              const rt_aerosol_input_t *aer, rt_result_t *out);
int rt_solve_case_pol_lut(const rt_case_t *cs, const rt_options_t *opts,
                          const rt_aerosol_input_t *aer, rt_result_t *out,
                          rt_lut_grid_out_t *lut_out);

/* Rayleigh column optical depth of Bodhaine et al. (1999): refractive index
 * of Peck & Reeder with the CO2 correction, King factor of dry air, gravity
 * at the mass-weighted column altitude; CO2 360 ppm
 * (code/OCRT_C/src/rt_rayleigh.c:46-52, 78-91, 129-137, 198-240, 285-293). */
double ocrt_rayleigh_tau_bodhaine1999(double wavelength_nm, double pressure_hpa,
                                      double latitude_deg, double altitude_m);

/* Batch form of rt_solve_case_pol_aerosol: n cases solved together by one
 * plan (the uneven-batch path of the GPU solver) instead of one plan per
 * case.  aer may be NULL (Rayleigh-only for every case) or hold n entries
 * (an entry with tau_a == 0 is Rayleigh-only).  Returns the first non-zero
 * per-case code (rc[i] holds each one when rc is not NULL), like the
 * OpenMP row loop of rt_io_run_batch (code/OCRT_C/src/rt_io.c:199-262). */
int ocrt_rt_solve_batch(const rt_case_t *cs, const rt_options_t *opts,
                        const rt_aerosol_input_t *aer, int n, rt_result_t *out, int *rc);

/* The reference's rough-interface surface functions, same names, arguments
 * and return codes (code/OCRT_C/src/shared/surface.h:40-61, 222-281), served
 * by the GPU kernels of ocrt_interface_fourier / ocrt_interface_direct.  The
 * first call for a node set and surface computes every mode
 * 0..max(m, build cap) at once and keeps them in a per-thread cache; modes
 * above 255 are refused (-1).  surface_fkc_set_build_mmax sets the build cap
 * (0..64; other values restore 64), as in surface.c:830-838. */
int surface_coxmunk_fourier_kernel(const double *mu_o, int n_o, const double *mu_i, int n_i, int m,
                                   int n_phi_quad, double ws, int sigma_type, double n_water,
                                   int q_convention, double *R_m);
int surface_R_ww_coxmunk_fourier_kernel(const double *mu_o, int n_o, const double *mu_i, int n_i, int m,
                                        int n_phi_quad, double ws, int sigma_type, double n_water,
                                        int q_convention, double *R_m);
int surface_T_wa_coxmunk_fourier_kernel(const double *mu_o, int n_o, const double *mu_i, int n_i, int m,
                                        int n_phi_quad, double ws, int sigma_type, double n_water,
                                        int q_convention, double *T_m);
int surface_T_aw_coxmunk_fourier_kernel(const double *mu_o, int n_o, const double *mu_i, int n_i, int m,
                                        int n_phi_quad, double ws, int sigma_type, double n_water,
                                        int q_convention, double *T_m);
void surface_fkc_set_build_mmax(int m_max);
/* NaN when the GPU call fails */
double surface_T_aw_coxmunk_direct(double mu_sun_air, double n_water, double wind_speed, int sigma_type);
double surface_R_ww_coxmunk_direct(double mu_up_water, double n_water, double wind_speed, int sigma_type);

/* Atmosphere BOA export and in-water inbound field of the coupled ocean,
 * field for field as rt_atm_boa_export_t (code/OCRT_C/src/rt_types.h:874-885)
 * and rt_aw_coupled_field_t (rt_air_water_coupling.h:26-33); fields are
 * [(m_max+1) * n] with index m * n + j. */
typedef struct {
    int     m_max;
    int     n_mu;
    double *mu_quad_pos;        /* [n_mu] positive air nodes */
    double *I_per_m;            /* BOA downward diffuse I^m */
    double *Q_per_m;
    double *U_per_m;
    double  tau_atm_total;
    double  mu_sun_air;
    int     allocated;
} rt_atm_boa_export_t;

typedef struct {
    int     m_max;
    int     n_mu_water;
    double *mu_water_pos;       /* [n_mu_water] */
    double *I_inwater_per_m;    /* in-water inbound I^m */
    double *Q_inwater_per_m;
    double *U_inwater_per_m;
} rt_aw_coupled_field_t;

/* The reference's coupling entry point and helpers (rt_air_water_coupling.h:35-68),
 * served by ocrt_couple_atm_to_water; -1 bad arguments, -3 GPU failure. */
int  rt_aw_coupled_field_alloc(rt_aw_coupled_field_t *f, int m_max, int n_mu_water);
void rt_aw_coupled_field_free(rt_aw_coupled_field_t *f);
int  rt_air_water_couple_atm_to_water(const rt_atm_boa_export_t *boa_export, double n_water, int q_convention,
                                      const double *mu_water_pos, int n_mu_water, const double *w_atm_pos,
                                      double wind_speed, int sigma_type, rt_aw_coupled_field_t *coupled);
/* Reverse coupling (rt_air_water_coupling.h:70-88), served by
 * ocrt_couple_water_to_atm; the weights enter only the reference's opt-in
 * re-escape term and are not used. */
int  rt_air_water_couple_water_to_atm(const double *I_water_per_m, const double *Q_water_per_m,
                                      const double *U_water_per_m, const double *mu_water_pos,
                                      const double *w_water_pos, int n_mu_water, int m_max, double n_water,
                                      int q_convention, double wind_speed, int sigma_type,
                                      const double *mu_atm_pos, int n_mu_atm, const double *w_atm_pos,
                                      double *I_air_per_m, double *Q_air_per_m, double *U_air_per_m);

#ifdef __cplusplus
}
#endif

#endif /* OCRT_RT_H */
