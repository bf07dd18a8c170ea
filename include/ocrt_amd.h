/*
 * ocrt_amd.h -- C ABI of the MI355X-native successive-orders-of-scattering
 * (SOS) solver for polarized (I, Q, U) top-of-atmosphere radiance.
 *
 * The boundary is a drop-in for the atmospheric solve path of OCRT:
 *   rt_solve_case_pol / rt_solve_case_pol_aerosol / rt_solve_case_pol_lut
 *   (code/OCRT_C/src/rt_solver.h:84, :119, :207; driver rt_solver.c:618-1437)
 * and for its OpenMP batch driver rt_io_run_batch (code/OCRT_C/src/rt_io.c:199-262),
 * which solves independent rows one per thread.  Here a whole batch of
 * uneven instances (different geometry, wavelength, aerosol load, layer count
 * and stream count) is solved by one plan on one GPU.
 *
 * Every signature uses plain C types and pointers.  Instance descriptors and
 * results live in host memory; all device memory is owned by the plan.
 *
 * Conventions kept from the reference:
 *   - return codes 0 ok, -1 bad arguments, -2 non-finite field, -3 allocation
 *     (code/OCRT_C/src/rt_solver.h:11 and rt_solver.c:623-640); -4 is new and
 *     means a HIP runtime error;
 *   - Convention B normalisation: rho = I_TOA / mu_sun (code/OCRT_C/src/rt_types.h:6-49);
 *   - the public relative azimuth RAA (RAA = 180 is the glint branch,
 *     code/OCRT_C/src/rt_raa_convention.h:8-37);
 *   - U reported in the fixed meridian-basis sine convention
 *     (code/OCRT_C/src/rt_fourier.c:24-33).
 */
#ifndef OCRT_AMD_H
#define OCRT_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Lower boundary kinds; numeric values equal rt_surface_kind_t
 * (code/OCRT_C/src/rt_types.h:55-61).  LAMBERT behaves like BLACK in the
 * reference atmosphere driver (rt_solver.c:979-989) and does so here. */
enum {
    OCRT_SURFACE_BLACK = 0,
    OCRT_SURFACE_LAMBERT = 1,
    OCRT_SURFACE_BLACK_FRESNEL_OCEAN = 2,
    OCRT_SURFACE_FLAT = 3
};

enum {
    OCRT_OK = 0,
    OCRT_ERR_ARGS = -1,
    OCRT_ERR_NONFINITE = -2,
    OCRT_ERR_ALLOC = -3,
    OCRT_ERR_DEVICE = -4,
    OCRT_ERR_FLAT = -5,     /* ocrt_interface_fourier: T_aw of a flat surface */
    OCRT_ERR_BOTTOM_SHAPE = -7  /* bottom source shape differs from the ring (rt_solver.h:11) */
};

/* Arithmetic forms of the solve.
 *   FP64  : every kernel in double precision (the plain parity form).
 *   MIXED : the scattering integral of orders >= 2 runs on bf16 MFMA with fp32
 *           accumulation and the layer sweeps of those orders in fp32; single
 *           scattering, the angular recurrences, the surface kernels, IPSS and
 *           the compensated fp64 order sum stay in fp64. */
enum { OCRT_PREC_FP64 = 0, OCRT_PREC_MIXED = 1 };

/* One solve.  Mirrors the fields of rt_case_t / rt_options_t /
 * rt_aerosol_input_t that the atmospheric path reads
 * (code/OCRT_C/src/rt_types.h:321-482, :654-754; rt_solver.h:94-114) with
 * tau_r_source = FROM_INPUT and view_as_node = 1 semantics. */
typedef struct {
    double sza_deg, vza_deg, raa_deg, wavelength_nm;
    int    surface;            /* OCRT_SURFACE_* */
    int    decouple_sunglint;  /* 1: no direct glint added to rho */
    double n_water;            /* real refractive index, default 1.34 */
    double wind_speed;         /* m/s; <= 0 with BLACK_FRESNEL_OCEAN = flat Fresnel */
    int    sigma_type;         /* 0 Nakajima-Tanaka, 1 OCRT floor law (default) */
    int    q_convention;       /* 1 Mishchenko (Rp-Rs)/2 (default) */
    double tau_R;              /* Rayleigh column optical depth; 0 = Rayleigh off */
    double depol;              /* Rayleigh depolarisation (the reference fixes 0.0279) */
    double tau_a;              /* aerosol column optical depth at this wavelength */
    double ssa_a;              /* aerosol single-scattering albedo */
    double aer_h_km;           /* aerosol exponential scale height, default 2 km */
    int    L_max;              /* aerosol Legendre order; 0 = no aerosol */
    int    n_mu;               /* Gauss-Legendre nodes per hemisphere */
    const double *greek;       /* host [4][L_max+1]: beta(P11), gamma(P12), alpha, zeta */
    int    n_layers;           /* plane-parallel layers (levels = n_layers + 1) */
    int    m_max;              /* highest Fourier mode (no cap at 32 here) */
    int    max_orders;         /* SOS order cap, default 20 */
    int    ipss;               /* 1: spherical-shell kappa correction */
    double tolerance;          /* order-sum stop, default 1e-7 */
    /* Optional explicit layering, host [4][n_layers], layers listed top down:
     *   [0] Rayleigh scattering optical depth, [1] aerosol extinction optical
     *   depth, [2] aerosol single-scattering albedo, [3] absorption optical
     *   depth (gas).
     * NULL selects the column generator above (tau_R, tau_a, ssa_a,
     * aer_h_km on US62 altitudes, rt_atm.c:272-400).  With layers, level k
     * sits below layer k-1 and carries that layer's scattering fractions
     * (the reference's level convention, rt_atm.c:343-360 and :521-538 for
     * absorption); tau_R and tau_a are ignored and IPSS is not available
     * (the layering carries no altitudes). */
    const double *layers;
    /* Optional gas absorption table (ocrt_gas_create); NULL: none.  Applies
     * to the column generator above (with layers, absorption is layers[3]). */
    const struct ocrt_gas *gas;
    /* Optional angle-space aerosol phase function, the value kernel of
     * rt_aerosol_input_t (code/OCRT_C/src/rt_solver.h:102-113, used at
     * rt_solver.c:190-240 and rt_ipss.c:134-146): n_ang_phase >= 2 ascending
     * scattering angles theta_phase[] in degrees with P11 values p11_phase[]
     * (normalised as beta_0 = 1; host memory).  With an aerosol it replaces
     * the Legendre sum of beta in the aerosol I -> I phase table, which is then
     * the azimuth mean of the linearly interpolated table over 720 midpoints,
     * and in the IPSS aerosol phase.  gamma, alpha and zeta still come from
     * greek.  0 / NULL: the moment kernel. */
    int n_ang_phase;
    const double *theta_phase;
    const double *p11_phase;
    /* Optional external upward source at the lower boundary: the coupled
     * ocean's water-leaving Fourier field just above the surface (the
     * ext_bottom_per_m_* options, code/OCRT_C/src/rt_types.h:722-747, applied
     * by rt_surface_boundary.c:495-527).  Host [3][bottom_m_max + 1][bottom_n_mu]
     * (I block, Q block, U block; mode m's values at the ring's upward nodes,
     * ascending mu).  bottom_n_mu must equal the instance's ring size
     * (ocrt_ring; OCRT_ERR_BOTTOM_SHAPE otherwise); modes above bottom_m_max
     * carry no source.  The source, attenuated to every level, joins the order
     * sum and the order-1 field that seeds the scattering.  It acts with the
     * FLAT and BLACK_FRESNEL_OCEAN surfaces only, as in the reference (its
     * atmosphere-only solver ignores it).  bottom_source_only = 1 drops the
     * solar single scattering and the surface's direct-beam seeds, so only the
     * bottom source drives the field.  NULL / 0: none.  Not with the adjoint. */
    const double *bottom_source;
    int bottom_n_mu, bottom_m_max;
    int bottom_source_only;
    /* In-water column of the coupled ocean (water_tau > 0 selects it): the
     * water body as the reference builds it in atmosphere form
     * (code/OCRT_C/src/rt_water_rt.c:2104-2236) -- n_layers layers uniform in
     * optical depth over water_tau, every level carrying the pure-water
     * scattering albedo water_omega_w in the molecular slot (depolarisation
     * depol, 0.039 for water) and the particle albedo water_omega_p in the
     * aerosol slot (Greek coefficients greek / L_max; value-kernel table
     * allowed).  sza_deg and vza_deg are the in-water (refracted) sun and
     * view angles; the solar single scattering is scaled by beam_scale (the
     * transmitted direct beam, 1 = the reference's F_sun = pi normalisation).
     * tau_R, tau_a, ssa_a, aer_h_km, layers, gas and ipss are not used
     * (layers, gas and ipss must be NULL / 0); surface must be BLACK (the deep
     * black bottom) and no bottom source is taken.
     *   top_reflect  0: none; 1: the flat water-side Fresnel reflection of each
     *                node (total internal reflection: identity), fed back every
     *                order as the downward boundary condition at the top
     *                (rt_solver.c:2290-2733); 2: the rough Cox-Munk water-side
     *                reflection (wind_speed, sigma_type, n_water, q_convention;
     *                128-azimuth Fourier kernels, rt_water_rt.c:5605), contracted
     *                with 2 pi mu w (m = 0) or pi mu w (rt_solver.c:2734-2922).
     *   top_source   optional host [3][top_m_max + 1][top_n_mu]: diffuse downward
     *                I, Q, U Fourier field entering below the surface (the
     *                transmitted skylight) at the ring nodes, attenuated to every
     *                level into the order sum and the order-1 field;
     *                top_n_mu must equal the ring size. */
    double water_tau, water_omega_w, water_omega_p, beam_scale;
    int top_reflect, top_n_mu, top_m_max, pad_water;
    const double *top_source;
} ocrt_instance_t;

/* Size and nodes of the ring a plan builds for n_mu Gauss nodes per hemisphere
 * and a view zenith angle: the positive half of the 2 n_mu-point
 * Gauss-Legendre rule with the view cosine inserted at its sorted position
 * with weight zero (none within 1e-12 of it; rt_angle_grid.c:21-73).  mu and w
 * (either may be NULL) receive the ascending cosines and weights, room for
 * n_mu + 1.  Returns the ring size, or OCRT_ERR_ARGS. */
int ocrt_ring(int n_mu, double vza_deg, double *mu, double *w);

/* Gas absorption of an AFGL standard atmosphere with layered HITRAN cross
 * sections (rt_absorption.c, rt_atm.c:437-548): made once on the host from
 * the reference's data files, shared by any number of instances and plans.
 *   profile           0 usstd76, 1 tropical, 2 mlsumm, 3 mlwint, 4 sasumm,
 *                     5 sawint, 6 userdef (afgl_<name>.dat under afgl_dir)
 *   column_override   [6] molecules cm^-2 for h2o, o3, o2, co2, no2, ch4;
 *                     < 0 keeps the AFGL column; NULL keeps all
 * Returns OCRT_OK, OCRT_ERR_ARGS (unreadable profile or malformed table). */
typedef struct ocrt_gas ocrt_gas;
int    ocrt_gas_create(const char *afgl_dir, const char *xsec_dir, int profile, const double *column_override,
                       ocrt_gas **gas_out);
/* Column absorption optical depth of gas g (0..5), or of all gases (g = -1). */
double ocrt_gas_tau(const ocrt_gas *gas, int g, double wavelength_nm);
void   ocrt_gas_destroy(ocrt_gas *gas);

/* Per-instance results; names follow rt_result_t (code/OCRT_C/src/rt_types.h:496-651). */
typedef struct {
    double rho_I, rho_Q, rho_U;
    double I_TOA, Q_TOA, U_TOA;
    double tau_R_total;
    double T_dir_dn, T_diff_dn_hemi, T_total_dn_hemi, T_diff_dn_dir;
    double T_dir_up_view;
    double T_sg_up_dir, T_total_up_dir;
    double rho_glint_direct_I, rho_glint_direct_Q, rho_glint_direct_U;
    double kappa;              /* IPSS factor applied (1 when IPSS is off) */
    int    n_orders_used;      /* max over m */
    int    converged;          /* AND over m */
    int    status;             /* OCRT_OK or an error code for this instance */
    int    n_ring;             /* ring size per hemisphere (GL nodes + view node) */
} ocrt_result_t;

typedef struct ocrt_plan ocrt_plan;

/* Build a plan: validate every instance (same rules as rt_solver.c:626-639
 * except the m_max <= 32 cap), build the quadrature rings on the host and
 * allocate all device memory.  One plan serves one set of instances and can be
 * executed any number of times.  one_instance_per_launch = 1 runs the
 * instances one after another, each with its own kernel launches (the plain
 * form); 0 solves the whole batch in one launch sequence. */
int ocrt_plan_create(const ocrt_instance_t *inst, int n_inst, int precision,
                     int one_instance_per_launch, ocrt_plan **plan_out);

/* Run the full solve on the given HIP stream (NULL = default stream).
 * Asynchronous with respect to the host; call ocrt_plan_results to wait. */
int ocrt_plan_execute(ocrt_plan *plan, void *hip_stream);

/* Copy results to host memory (synchronises the stream).
 *   out        [n_inst]
 *   toa_modes  optional [n_inst][mmax_all+1][3] TOA (I,Q,U)^m at the view node
 *   orders     optional [n_inst][mmax_all+1] SOS orders used per mode
 * where mmax_all is the largest m_max of the plan (ocrt_plan_mmax). */
int ocrt_plan_results(ocrt_plan *plan, ocrt_result_t *out, double *toa_modes,
                      int *orders);

int  ocrt_plan_mmax(const ocrt_plan *plan);
/* Bytes of device memory the plan holds. */
size_t ocrt_plan_device_bytes(const ocrt_plan *plan);
/* Algorithmic work of the last execute, summed over the orders actually run
 * (useful to compute roofline fractions); either pointer may be NULL:
 *   gemm_flops  : 72 n n_in (nt + 1) per item and order for the particle
 *                 operator, plus 144 n (nt + 1) for the rank <= 6 Rayleigh part
 *                 of modes m <= 2 (n_in: nodes of non-zero weight);
 *   sweep_bytes : per element (6 n (nt + 1) per item and order) the source read
 *                 (4 B mixed, 8 B fp64), the field written (2 B, 4 B for the
 *                 field pairs of modes m < 3; 8 B fp64) and the fp64 order sum
 *                 read and written (16 B), plus 1 - T read for every layer and
 *                 node of both hemispheres (2 n nt x 4 or 8 B). */
int  ocrt_plan_work(const ocrt_plan *plan, double *gemm_flops, double *sweep_bytes,
                    double *gemm_launches);
/* Per-kernel timing for roofline reports: when enabled, every scattering-
 * integral (GEMM) and layer-sweep launch of the following executes is
 * bracketed by HIP events on the execution stream; ocrt_plan_kernel_times
 * returns the summed durations (ms) and launch counts of the last execute. */
int  ocrt_plan_profile(ocrt_plan *plan, int enable);
int  ocrt_plan_kernel_times(ocrt_plan *plan, double *gemm_ms, int *gemm_launches,
                            double *sweep_ms, int *sweep_launches);
void ocrt_plan_destroy(ocrt_plan *plan);

/* Plan options for ocrt_plan_create_ex. */
enum {
    OCRT_PLAN_ONE_PER_LAUNCH = 1,  /* instances one after another (the plain form) */
    OCRT_PLAN_ADJOINT = 2,         /* keep order checkpoints for ocrt_plan_adjoint (IPSS off); FP64: fp64
                                      fields and adjoint; MIXED: the mixed order loop with every field kept
                                      as its bf16 pair, fp32 adjoint fields, bf16 MFMA products */
    OCRT_PLAN_CKPT_SPARSE = 4,     /* adjoint: checkpoint every 4th order and recompute the rest, whatever
                                      the memory allows (default: every order kept when the batch fits one
                                      chunk, else every 2nd, else every 4th) */
    OCRT_PLAN_AUTOTUNE = 8,        /* mixed form: the first execute times every GEMM tile shape per
                                      operator class, then every launch shape of the layer sweep, on this
                                      plan's work and keeps the fastest (ocrt_plan_gemm_tuning,
                                      ocrt_plan_tuning); ignored by adjoint plans */
    OCRT_PLAN_GRAPH = 16,          /* the first execute captures the whole forward launch sequence into a
                                      hipGraph; later executes replay it (not with profiling or adjoint) */
    OCRT_PLAN_DEVICE_INPUTS = 32   /* the per-layer optical depths and wind speeds may be replaced from
                                      device memory between executes (ocrt_plan_set_inputs): no two
                                      instances share their surface kernels */
};

/* Launch shape chosen for the mixed-form sweep (levels per block, wavefronts
 * per workgroup) and, after autotuning, the summed order-loop time
 * (scattering integral and sweep) of every shape tried (ms_per_shape[n_shapes],
 * may be NULL).  All shapes give the same bits. */
int ocrt_plan_tuning(const ocrt_plan *plan, int *sweep_unroll, int *sweep_waves, double *ms_per_shape,
                     int *n_shapes);
/* Selects the mixed-form sweep's launch shape by hand (unroll 4, 8 or 16 with
 * waves 2 or 4); it replaces autotuning.  No counterpart in the reference.
 * -1 for a pair that is not a shape. */
int ocrt_plan_set_tuning(ocrt_plan *plan, int sweep_unroll, int sweep_waves);

/* Tile shape of the mixed form's scattering integral per operator class
 * (0: one operator part and a single bf16 field, 1: one part and a field
 * pair, 2: two parts): shape 0 = 64 operator rows x 128 levels, 1 = 64 x 64,
 * 2 = 128 x 64, 3 = 128 x 128.  All shapes give the same bits.
 * ocrt_plan_set_gemm_shape selects one by hand; ocrt_plan_gemm_tuning reports
 * the shape of each class and, after autotuning (OCRT_PLAN_AUTOTUNE times
 * every shape on the plan's own work before the sweep's launch shapes),
 * ms[class * 4 + shape], the class's summed GEMM time per execute; it returns
 * the number of shapes timed (0 before autotuning).  No counterpart in the
 * reference. */
int ocrt_plan_set_gemm_shape(ocrt_plan *plan, int op_class, int shape);
int ocrt_plan_gemm_tuning(const ocrt_plan *plan, int *shape_of_class, double *ms);

/* Precision policy of the mixed form's scattering integral (no counterpart in
 * the reference).  The first operator products of an item use the bf16-pair
 * operator (A_hi + A_lo) and, for the modes m < 3, the bf16-pair field.  After
 * each order the item's remaining share of its order sum is estimated as
 * res r / (1 - r), with res = max |I_k| / max |order sum| (the convergence
 * test's ratio) and r = res_k / res_(k-1); once it is below tau, the next
 * order's product takes the reduced form: A_hi times the bf16 field, one MFMA
 * per block.  The reduced orders carry about tau of the item's total, so
 * their bf16 rounding (2^-9 relative) stays near tau * 2^-9 of it; slowly
 * converging items (thick water columns) wait for a proportionally smaller
 * residual.  tau = 0 keeps the full form for every order; 0 <= tau <= 1;
 * mixed plans only.  ocrt_plan_gemm_issued reports, from the last results
 * call, the MFMA flops the scattering integral launched (tile padding, low
 * halves and reduced orders counted as run), the item-orders that ran it and
 * how many of them ran reduced. */
int ocrt_plan_set_reduced_threshold(ocrt_plan *plan, double tau);
int ocrt_plan_gemm_issued(const ocrt_plan *plan, double *issued_flops, double *item_orders,
                          double *reduced_orders);

/* ocrt_plan_create with option flags.  No counterpart in the reference; the
 * adjoint is the Jacobian path of the batched solver (ocrt_adjoint_batch). */
int ocrt_plan_create_ex(const ocrt_instance_t *inst, int n_inst, int precision, int flags,
                        ocrt_plan **plan_out);

/* Reverse-mode gradient of sum_i grad_out[i] . (rho_I, rho_Q, rho_U)_i for an
 * OCRT_PLAN_ADJOINT plan, by the adjoint SOS recursion (order fields kept
 * every 1, 2 or 4 orders; the orders between checkpoints are recomputed).
 * Runs the forward pass itself unless the last ocrt_plan_execute of a
 * single-chunk plan left its checkpoints.
 *   grad_out     host [n_inst][3]
 *   grad_levels  optional host, per instance in order, 3 x (n_layers + 1):
 *                d/dh[k] (level optical depth from the top), d/dxdel[k],
 *                d/dydel[k] (aerosol and Rayleigh scattering fractions of
 *                level k, rt_atm.c:343-360)
 *   grad_wind    optional host [n_inst], d/d wind speed
 * The per-layer Jacobian of an explicitly layered instance follows by the
 * chain rule through the layer-to-level map (ocrt_amd/adjoint.py). */
int ocrt_plan_adjoint(ocrt_plan *plan, const double *grad_out, double *grad_levels, double *grad_wind,
                      void *hip_stream);

/* Device-memory inputs and outputs of a plan, for callers whose data already
 * lives on the GPU (the torch.autograd.Function of ocrt_amd/torch_ops.py): no
 * host copies, every copy and kernel on hip_stream, no synchronisation.  No
 * counterpart in the reference, whose solver takes host structs only.
 *
 * ocrt_plan_set_inputs: the layers of every instance, packed in instance
 * order as [4][n_layers] blocks (the layout of ocrt_instance_t.layers), from
 * d_layers; and the wind speeds [n_inst] from d_wind (plans made with
 * OCRT_PLAN_DEVICE_INPUTS).  Either may be NULL.  The structure of the plan
 * stays as made: an instance whose new values break it (a negative or
 * non-finite optical depth, an albedo outside [0, 1], a layer with no
 * extinction, an aerosol column appearing or vanishing, a rough surface
 * without wind or a flat one with wind) reports status OCRT_ERR_ARGS from the
 * next execute.  OCRT_ERR_ARGS when an instance is not layered (d_layers) or
 * the plan lacks OCRT_PLAN_DEVICE_INPUTS (d_wind). */
int ocrt_plan_set_inputs(ocrt_plan *plan, const double *d_layers, const double *d_wind, void *hip_stream);
/* After an execute: d_rho [n_inst][3] = (rho_I, rho_Q, rho_U), NaN for an
 * instance whose status is not OCRT_OK; d_status [n_inst] (may be NULL). */
int ocrt_plan_rho(ocrt_plan *plan, double *d_rho, int *d_status, void *hip_stream);
/* ocrt_plan_adjoint with device-memory operands and the per-layer chain rule
 * on the device: d_grad_out [n_inst][3]; d_grad_layers packed like d_layers
 * above, d/d(Rayleigh tau, aerosol tau, aerosol albedo, absorption tau) of
 * every layer (may be NULL); d_grad_wind [n_inst] (may be NULL).  Every
 * instance must be layered when d_grad_layers is given. */
int ocrt_plan_adjoint_layers(ocrt_plan *plan, const double *d_grad_out, double *d_grad_layers, double *d_grad_wind,
                             void *hip_stream);

/* One-call Jacobian form (the boundary's ocrt_adjoint_batch): plan, forward,
 * adjoint, copy results and gradients, destroy.  precision FP64 or MIXED
 * (the two adjoint forms of ocrt_plan_create_ex with OCRT_PLAN_ADJOINT). */
int ocrt_adjoint_batch(const ocrt_instance_t *inst, int n_inst, int precision, const double *grad_out,
                       ocrt_result_t *out, double *grad_levels, double *grad_wind, void *hip_stream);

/* LUT / full-grid output (rt_solve_case_pol_lut, code/OCRT_C/src/rt_solver.h:207,
 * rt_solver.c:1031-1043 and 1280-1424): after this call every execute also
 * samples each instance on the grid vza_deg[n_vza] x raa_deg[n_raa] (PCHIP in
 * the view cosine over the ring nodes, reconstruction at each azimuth, IPSS
 * factor per direction, direct glint).  Call once per plan, before execute. */
int ocrt_plan_set_lut(ocrt_plan *plan, int n_vza, const double *vza_deg, int n_raa, const double *raa_deg);

/* Copy the LUT of the last execute (synchronises):
 *   grid     [n_inst][n_vza][n_raa][6]: rho_I, rho_Q, rho_U, T_diff_dn_dir,
 *            T_sg_up_dir, T_total_up_dir (rt_lut_grid_out_t, rt_solver.h:180-205)
 *   scalars  [n_inst][2]: T_dir_dn, T_diff_dn_hemi
 * either pointer may be NULL. */
int ocrt_plan_lut_results(ocrt_plan *plan, double *grid, double *scalars);

/* The remaining LUT outputs of rt_lut_grid_out_t (rt_types.h:820-833),
 * either pointer may be NULL:
 *   raw         [n_inst][n_vza][n_raa][3]: TOA (I, Q, U) before the division
 *               by mu_sun, the IPSS factor and the glint (raw_I/Q/U)
 *   view_modes  [n_inst][n_vza][m_max_all + 1][3]: the Fourier modes sampled
 *               at each view zenith (view_per_m_I/Q/U), zero above an
 *               instance's m_max; m_max_all = ocrt_plan_mmax(plan). */
int ocrt_plan_lut_extras(ocrt_plan *plan, double *raw, double *view_modes);

/* Per-node Fourier fields, the boa_I/Q/U and per-m TOA outputs of the Python
 * batch solver (code/OCRT_Python/ocrt_py/atmos_batch.py:990-1013).
 * ocrt_plan_export_nodes enables them (call before the execute whose fields
 * are wanted; LUT plans have them already).  ocrt_plan_node_fields copies the
 * last execute's fields (synchronises), either pointer may be NULL:
 *   toa_up  [n_inst][m_max_all + 1][3][ring_pitch]: TOA upward I, Q, U at
 *           every ring node (ascending mu; rough direct-beam seed removed)
 *   boa_dn  [n_inst][m_max_all + 1][3][ring_pitch]: BOA downward I, Q, U at
 *           every ring node (ascending |mu|)
 * zero past an instance's n_ring or m_max; ring_pitch >= every n_ring. */
int ocrt_plan_export_nodes(ocrt_plan *plan);
int ocrt_plan_node_fields(ocrt_plan *plan, int ring_pitch, double *toa_up, double *boa_dn);

/* Final relative residual of every (instance, mode) of the last execute,
 * [n_inst][m_max_all + 1] (sos_residual_per_m of rt_result_t, rt_types.h:511). */
int ocrt_plan_residuals(ocrt_plan *plan, double *residuals);

/* ---------------------------------------------------------------------------
 * Rough air-water interface matrices of the coupled ocean (Cox-Munk slopes,
 * Fresnel facets, Sancer shadowing for the reflections and T_wa):
 *   OCRT_IFACE_R_AA  air-side reflection; mu_o upward (air), mu_i the
 *                    incident downward |mu|: surface_coxmunk_fourier_kernel
 *                    (code/OCRT_C/src/shared/surface.h:222, surface.c:839-1134)
 *   OCRT_IFACE_R_WW  water-side internal reflection; mu_o reflected downward,
 *                    mu_i incident upward |mu| in water:
 *                    surface_R_ww_coxmunk_fourier_kernel (surface.h:246, surface.c:1386-1447)
 *   OCRT_IFACE_T_WA  water -> air transmission; mu_o air, mu_i water:
 *                    surface_T_wa_coxmunk_fourier_kernel (surface.h:271, surface.c:1731-1811)
 *   OCRT_IFACE_T_AW  air -> water transmission; mu_o water, mu_i air:
 *                    surface_T_aw_coxmunk_fourier_kernel (surface.h:264, surface.c:1583-1721)
 */
enum { OCRT_IFACE_R_AA = 0, OCRT_IFACE_R_WW = 1, OCRT_IFACE_T_WA = 2, OCRT_IFACE_T_AW = 3 };
#define OCRT_IFACE_MAX_MODES 256     /* m_max + 1 */
#define OCRT_IFACE_MAX_PHI 65536

/* Azimuthal Fourier kernels of one interface matrix for every mode
 * m = 0..m_max in one launch: out[((m * n_o + o) * n_i + i) * 9 + 3k + l]
 * (host memory), mode m's slab equal to what the reference's per-mode
 * function writes for m.  Midpoint azimuth rule with n_phi points, cos for
 * the II IQ QI QQ UU entries and sin for the others, 1/(2 pi) at m = 0 and
 * 1/pi above; IU and QU negated for m > 0 except in R_WW.  Cosines in
 * (0, 1]; n_o * n_i < 2^23; sigma_type 1 is the OCRT slope law, any other
 * value Cox-Munk's sigma = 0.0731 sqrt(W) (surface.c:57-64).  Returns
 * OCRT_OK, OCRT_ERR_ARGS, OCRT_ERR_FLAT (T_AW with slope variance <= 1e-10,
 * where the reference returns -2), OCRT_ERR_DEVICE.  Blocks on hip_stream
 * (NULL: the default stream) until the result is in out. */
int ocrt_interface_fourier(int kind, const double *mu_o, int n_o, const double *mu_i, int n_i, int m_max,
                           int n_phi, double wind, int sigma_type, double n_water, int q_convention,
                           double *out, void *hip_stream);

/* Allocation-free form of ocrt_interface_fourier.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Inputs and output are caller-owned DEVICE memory on the current device:
 * d_mu_o[n_o], d_mu_i[n_i] (cosines in (0, 1], not checked here), d_out with
 * the layout of ocrt_interface_fourier.  The azimuth weight table lives in a
 * fixed 4 MiB device array per GPU, so (max(m_max, 1) + 1) * n_phi must not
 * exceed 262144 (m_max 255 at n_phi 1024; OCRT_ERR_ARGS and a message
 * otherwise).  The call allocates and frees nothing and does not
 * synchronise: the result is in d_out when hip_stream (NULL: default stream)
 * reaches it.  Calls that overlap on one GPU share the table and race; order
 * them on one stream.  Same bits as ocrt_interface_fourier. */
int ocrt_interface_fourier_noalloc(int kind, const double *d_mu_o, int n_o, const double *d_mu_i, int n_i,
                                   int m_max, int n_phi, double wind, int sigma_type, double n_water,
                                   int q_convention, double *d_out, void *hip_stream);

/* Slope-integrated direct irradiance terms for n cosines (host memory):
 * OCRT_IFACE_T_AW the air -> water transmittance of the solar beam,
 * surface_T_aw_coxmunk_direct (surface.h:43, surface.c:138-198);
 * OCRT_IFACE_R_WW the water-side reflectance of an upwelling beam,
 * surface_R_ww_coxmunk_direct in its default first-encounter form
 * (surface.h:61, surface.c:229-368). */
int ocrt_interface_direct(int kind, const double *mu, int n, double wind, int sigma_type, double n_water,
                          double *out, void *hip_stream);

/* Atmosphere -> water coupling of the coupled ocean
 * (code/OCRT_C/src/rt_air_water_coupling.c:244-478): the atmosphere's BOA
 * downward Fourier field I/Q/U_a [(m_max+1) * n_a] at the air nodes mu_a
 * (host arrays) mapped onto the in-water inbound field I/Q/U_w
 * [(m_max+1) * n_w] at the water nodes mu_w.
 *   w_a != NULL and a rough surface (wind > 0, slope variance > 1e-10): the
 *     Cox-Munk air -> water Fourier transmission (ocrt_interface_fourier
 *     kind T_AW, 128 azimuths) contracted with 2 pi (m = 0) or pi times
 *     mu_j w_j;
 *   otherwise: Snell refraction, the flat radiance transmission Mueller and
 *     linear interpolation in the air nodes (zero inside the total-reflection
 *     cone); with wind > 0 the m = 0 intensity is replaced by the BTDF
 *     transmission of the m = 0 skylight (192 air nodes, 96 azimuths).
 * n_water > 1, n_a >= 2, m_max <= 255.  Blocks until the outputs are in
 * host memory. */
int ocrt_couple_atm_to_water(int m_max, int n_a, const double *mu_a, const double *w_a, const double *I_a,
                             const double *Q_a, const double *U_a, double n_water, int q_convention,
                             const double *mu_w, int n_w, double wind, int sigma_type, double *I_w,
                             double *Q_w, double *U_w, void *hip_stream);

/* Water -> atmosphere coupling (rt_air_water_coupling.c:598-780): the
 * in-water upwelling Fourier field I/Q/U_w [(m_max+1) * n_w] at the water
 * nodes mu_w mapped onto the air nodes mu_a, I/Q/U_a [(m_max+1) * n_a].
 *   wind > 0: per air node, the Cox-Munk water -> air Fourier transmission
 *     (kind T_WA, 128 azimuths) over 1024 trapezoid points within 0.3 of the
 *     refracted cosine plus 96 Gauss nodes outside, contracted with the
 *     linearly interpolated water field (2 pi at m = 0, pi above); air nodes
 *     outside (0, 1] give zero;
 *   otherwise: reverse Snell, the flat water -> air radiance transmission
 *     Mueller and the interpolated water field.
 * The reference's opt-in multi-encounter re-escape term (environment switch,
 * off by default) is not added.  n_water > 1, n_w >= 2, m_max <= 255. */
int ocrt_couple_water_to_atm(int m_max, int n_w, const double *mu_w, const double *I_w, const double *Q_w,
                             const double *U_w, double n_water, int q_convention, double wind, int sigma_type,
                             int n_a, const double *mu_a, double *I_a, double *Q_a, double *U_a, void *hip_stream);

/* Batched forms of the two couplings (no counterpart in the reference, which
 * couples one case per call): n_inst instances that share the node sets, the
 * wind, the slope model, the refractive index, the Q convention and m_max.
 * `fields` holds each instance's I, Q and U blocks back to back,
 * [n_inst][3][(m_max+1) * n_in], and `out` receives
 * [n_inst][3][(m_max+1) * n_out] in the same order.  The transmission kernels
 * are computed once for the whole batch; each instance's output is bit for bit
 * what the one-instance call above gives it.  1 <= n_inst <= 65535. */
int ocrt_couple_atm_to_water_batch(int n_inst, int m_max, int n_a, const double *mu_a, const double *w_a,
                                   const double *fields, double n_water, int q_convention, const double *mu_w,
                                   int n_w, double wind, int sigma_type, double *out, void *hip_stream);
int ocrt_couple_water_to_atm_batch(int n_inst, int m_max, int n_w, const double *mu_w, const double *fields,
                                   double n_water, int q_convention, double wind, int sigma_type, int n_a,
                                   const double *mu_a, double *out, void *hip_stream);

/* One-call form: plan, execute, copy results, destroy. */
int ocrt_solve_batch(const ocrt_instance_t *inst, int n_inst, int precision,
                     ocrt_result_t *out, void *hip_stream);

/* Library identity, for the loaders' checks. */
const char *ocrt_version(void);

#ifdef __cplusplus
}
#endif

#endif /* OCRT_AMD_H */
