/*
 * ocrt_oracle.h -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * Plain-C, double-precision restatement of the reference OCRT atmospheric
 * SOS path.  It is the checker of the HIP library: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  It is
 * never linked into, or called by, the product (libocrt_amd.so).
 *
 * Parity status of each part is documented in ocrt_oracle.c next to the
 * function, with the reference file:line it follows.
 */
#ifndef OCRT_ORACLE_H
#define OCRT_ORACLE_H

#include "ocrt_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Emulation switch for precision-budget studies: the scattering integral of
 * orders >= 2 rounds the operator coefficients to (mode >> 8) and the incoming
 * field to (mode & 0xff) significant bits before the contraction (8 =
 * bfloat16, 11 = binary16; mode 1 = both bfloat16).  Default 0 (pure double
 * precision). */
void oracle_set_bf16_emulation(int bf16);

/* Full atmospheric solve of one instance (rt_solve_case_pol_impl,
 * code/OCRT_C/src/rt_solver.c:618-1437).  toa_modes [m_max+1][3], orders and
 * residuals [m_max+1] are optional.  Unlike the reference, m_max is not
 * capped at 32. */
int oracle_solve(const ocrt_instance_t *in, ocrt_result_t *out, double *toa_modes,
                 int *orders, double *residuals);

/* LUT / full-grid solve (rt_solver.c:1031-1043, 1280-1424). */
int oracle_solve_lut(const ocrt_instance_t *in, int n_vza, const double *vza,
                     int n_raa, const double *raa,
                     double *rho_I, double *rho_Q, double *rho_U,
                     double *T_diff_dn_dir, double *T_sg_up_dir, double *T_total_up_dir,
                     double *T_dir_dn, double *T_diff_dn_hemi);

/* Reverse-mode gradient of gout . (rho_I, rho_Q, rho_U) with respect to the
 * per-level medium h[k], xdel[k], ydel[k] (k = 0..n_layers) and the wind
 * speed, IPSS off.  No counterpart in the reference: pinned by central finite
 * differences of oracle_solve. */
int oracle_adjoint(const ocrt_instance_t *in, const double *gout, double *g_h, double *g_x, double *g_y,
                   double *g_wind);

/* OpenMP batch, one instance per thread (rt_io.c:199). */
int oracle_solve_batch(const ocrt_instance_t *in, int n, ocrt_result_t *out, int n_threads);
/* Per-node Fourier fields (top upward, bottom downward, [m_max+1][3][n]) and
 * order counts of one solve; returns the ring size n or a negative code. */
int oracle_solve_nodes(const ocrt_instance_t *in, double *top_up, double *bot_dn, int *orders);

/* ---- component entry points used by the unit tests ---- */
int oracle_gauss_legendre_pos(int n, double *mu, double *w);
/* ring of positive cosines: GL(n_gauss) plus mu_view inserted (sorted,
 * dedup 1e-12) when mu_view >= 0; returns ring size, *view_index (0-based). */
int oracle_ring(int n_gauss, double mu_view, double *mu, double *w, int *view_index);
/* medium arrays [n_layers+1]; returns 0 on success */
int oracle_medium(const ocrt_instance_t *in, double *h, double *ch, double *xdel,
                  double *ydel, double *z_km);
double oracle_us62_grid_fraction_above(double z_km);
/* basis tables [(l_max+1)][n_dir] for directions mu[0..n_dir-1] */
int oracle_basis(int l_max, int m, int n_dir, const double *mu,
                 double *plm, double *rrl, double *rtl);
/* 3x3 Cox-Munk reflection Mueller matrix (surface.c:514-646) */
void oracle_coxmunk_R(double mu_out, double mu_in_dn, double cos_phi, double sin_phi,
                      double ws, int sigma_type, double n_water, int q_convention,
                      double *R);
void oracle_flat_fresnel(double mu, double n_water, int q_convention, double *MF);
/* Fourier kernels of the Cox-Munk BRDF for all m <= m_max at once:
 * out [(m_max+1)][n_o][n_i][9] (sign fold for m > 0 applied). */
int oracle_coxmunk_fourier(const double *mu_o, int n_o, const double *mu_i, int n_i,
                           int m_max, int n_phi, double ws, int sigma_type,
                           double n_water, int q_convention, double *out);
/* Rough-interface kernels of the coupled ocean, kind 0 R_aa (air-side
 * reflection), 1 R_ww (water-side reflection), 2 T_wa (water -> air), 3 T_aw
 * (air -> water); mu_o outgoing, mu_i incident zenith cosine
 * (shared/surface.c:80-136, 514-803, 1136-1249, 1476-1566). */
void oracle_interface_trig(int kind, double mu_o, double mu_i, double cos_phi, double sin_phi,
                           double ws, int sigma_type, double n_water, int q_convention, double *M);
/* their Fourier kernels for all m <= m_max: out [(m_max+1)][n_o][n_i][9];
 * -2 for T_aw on a flat surface */
int oracle_interface_fourier(int kind, const double *mu_o, int n_o, const double *mu_i, int n_i,
                             int m_max, int n_phi, double ws, int sigma_type, double n_water,
                             int q_convention, double *out);
/* slope-integrated direct terms: kind 3 T_aw(mu_sun), kind 1 R_ww(mu_up)
 * (surface.c:138-198, 229-368) */
double oracle_interface_direct(int kind, double mu, double ws, int sigma_type, double n_water);
/* IPSS kappa for an instance at (vza, raa) (rt_ipss.c:607-644) */
double oracle_ipss_kappa(const ocrt_instance_t *in, double vza_deg, double raa_deg);
double oracle_ipss_tau_dir_dn(const ocrt_instance_t *in);
/* IPSS geometry on the homogeneous 0-100 km test atmosphere of the
 * reference gates: TOA-anchored line-of-sight depth and solar depth at rB
 * (-1 when shadowed). */
double oracle_ipss_homog_tau_view(double tau_t, int nlay, double vza_deg);
double oracle_ipss_homog_tau_sun(double tau_t, int nlay, double sza_deg, const double *rB);
/* Fourier reconstruction (rt_fourier.c:9-33) */
double oracle_reconstruct_cos(const double *modes, int m_max, double raa_rad);
double oracle_reconstruct_sin(const double *modes, int m_max, double raa_rad);
/* Complete the spin-2 families from ordinary Legendre moments
 * (rt_phase_moments.c:22-64). */
int oracle_complete_spherical(int l_max, const double *beta11, const double *beta22,
                              const double *delta33, double *alpha, double *zeta);

#ifdef __cplusplus
}
#endif

#endif
