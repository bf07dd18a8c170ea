// Shared device/host definitions of the MI355X SOS solver.
//
// Memory layout (one "chunk" of instances is resident at a time):
//   per instance : level arrays h, ch, xdel, ydel, z [L = nt + 1] and the gas
//                  absorption of each layer [nt] (in a sixth row); ring cosines
//                  and weights [n]; Beer-Lambert factors T[layer][node]
//                  (layer-major so a wavefront reads consecutive nodes);
//                  Greek arrays [4][La+1]; rough-surface Fourier kernels
//                  R[m][o][i][9] and solar kernels Rsun[m][o][9].
//   per item     : one (instance, Fourier mode m) pair.  Angular basis
//                  plm/rrl/rtl [(lmax+1)][2n+1], six phase tables
//                  [(n+1)][2n+1], the dense scattering operator A[Mp][K]
//                  and the fields F (previous order), J (source), total
//                  (fp64 order sum) and comp (compensation, mixed form).
//   Field row k (level) holds 6n state columns ordered
//     [I_up(n) Q_up(n) U_up(n) I_dn(n) Q_dn(n) U_dn(n)], node o = 0..n-1
//   padded to the row pitch P; pad columns stay zero.
#pragma once
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#define OCRT_PI 3.14159265358979323846

namespace ocrt {

constexpr int kNPhi = 1024;      // Cox-Munk azimuth rule (rt_surface_boundary.c:111)
constexpr int kMaxRing = 640;    // ring cap (rt_angle_grid.h:11)

struct DevInst {
    double mus, muv, sza, vza, raa;
    double tauR, taua, ssa, aer_h, n_water, wind, tol;
    double beta0, beta2, gamma2, alpha2;      // Rayleigh moments
    int n, vidx, nt, mmax, La, lmax, aer, ray;
    int surf;            // 0 none, 1 flat, 2 rough
    int sigma_type, q_conv, decouple, ipss, max_orders, item0, surface_kind, surf_owner;
    int layered;         // 1: explicit per-layer optical depths (pool_layers)
    int gidx;            // index of the instance in the plan
    int gas;             // gas absorption table (Pools::gas_tab), -1: none
    long long off_lev;   // pool_lev: 7 arrays of L doubles (h, ch, xd, yd, z, layer absorption, 1 / dtau)
    long long off_ring;  // pool_ring: mu[n], w[n]
    long long off_T;     // pool_T (double) / pool_Tf (float): [nt][n]
    long long off_greek; // pool_greek: [4][La+1]
    long long off_R;     // pool_R: [(mmax+1)][n incoming][9 entries][n outgoing] (see r_index)
    long long off_Rsun;  // pool_R: [(mmax+1)][n][9]
    long long off_layers;// pool_layers: [4][nt] Rayleigh tau, aerosol tau, aerosol ssa, absorption tau
    long long off_grad;  // pool_grad: gH[L], gX[L], gY[L], gD[nt], gW[1] (adjoint)
    double wl;           // wavelength (nm)
    long long off_ang;   // pool_greek: value-kernel P11 table, theta[n_ang] (deg) then P11[n_ang]
    int n_ang;           // entries of that table; 0: the moment kernel
    int bsrc_mmax;       // external bottom source: highest mode carried, -1: none
    long long off_bsrc;  // pool_greek: bottom source [3][bsrc_mmax + 1][n] (upward I, Q, U per mode)
    int bsrc_only;       // 1: no solar single scattering or direct-beam surface seeds
    int water;           // 1: in-water column (uniform levels over wtau, fixed albedos)
    double wtau, womega_w, womega_p, beam_scale;
    int top_kind;        // water-side reflection at the top: 0 none, 1 flat Fresnel, 2 rough Cox-Munk
    int tsrc_mmax;       // diffuse top source: highest mode carried, -1: none
    long long off_tsrc;  // pool_greek: top source [3][tsrc_mmax + 1][n] (downward I, Q, U per mode)
    long long off_rtop;  // pool_rtop (top_kind 2): [(mmax+1)][n out][n in][9] water-side Fourier kernels
    float jscale, jinv;  // mixed form: the scattering source is kept in fp16 as J * jscale (a power of two)
    int aer_able;        // a phase matrix is given (aerosol scattering possible with a positive column)
    int input_bad;       // inputs replaced from device memory broke the plan's structure: status OCRT_ERR_ARGS
};

// Whether the rough surface's direct-beam seed enters the order-1 field
// (rt_surface_boundary.c:553-567): rough surface and a solar source.
__host__ __device__ inline bool rough_seed_on(const DevInst &I) { return I.surf == 2 && !I.bsrc_only; }

// One gas absorption table in the gas pool: AFGL levels z[n_lev] (km,
// ascending) and number densities n[6][n_lev] (cm^-3); per gas a
// cross-section table sigma[n_layers][n_wl] over wl[n_wl] (cm^2), layer i at
// altitude z[i].
struct DevGas {
    int n_lev, pad_;
    int n_wl[6], n_layers[6], avail[6], overridden[6];
    double col_def[6], col_eff[6];
    long long off_z, off_n;
    long long off_wl[6], off_sig[6];
};

struct DevItem {
    int inst, m, n, nt, L, P, K, parts;   // parts: bit0 particle, bit1 molecular
    int Mp, tiles_m, tiles_n, lmax;
    int fsplit;           // mixed form: field stored as a bf16 pair hi + lo (low Fourier modes)
    int op_owner;         // 1: computes its basis, phase tables and operator; 0: shares another item's
    // items of instances with identical ring, sun, phase-matrix expansion and
    // Rayleigh coefficients (the bands of one geometry and aerosol model) share
    // one basis, one set of phase tables and one operator per Fourier mode
    long long off_basis;  // pool_basis: 3 x [(lmax+1)][2n+1]
    long long off_tab;    // pool_tab: 6 x [(n+1)][2n+1]
    long long off_A;      // pool_A: [Mp][K]
    long long off_F;      // pool_F / pool_J: [L][P]
    long long off_tot;    // pool_tot: [L][P] fp64 (the fp64 form only)
    long long off_bc;     // pool_bc: [3][n] bottom boundary condition, [3][n] bottom downward field,
                          // [3][n] top boundary condition, [3][n] top upward field
    long long off_cmp;    // pool_comp: [2][P] compensation of the order sum at k = 0 and k = nt
    long long off_Flo;    // pool_Flo: [L][P] low half of the bf16 field pair (fsplit items)
    long long off_nodes;  // pool_nodes (LUT or node-export plans, global over the plan): [6][n]
    // mixed form, m <= 2 with aerosol and Rayleigh: the molecular operator
    // (rank 3) runs outside the dense GEMM.  Operator rows pbase .. pbase + 2
    // hold its input factors G (row pbase + 3 zero), so the GEMM yields
    // W = G F per level (pool_W [L][4]); the sweep adds ydel (U W) with the
    // output factors U = E C (pool_lr [3][P]).  pbase = 6 n rounded up to 4
    // when those four rows fit in the operator's padding to 64 rows, else the
    // operator gets 64 more rows from pbase = its padded row count.
    int lowrank, pbase;
    long long off_lr, off_W;
};

struct ItemState {
    int active, orders, converged, status;
    double residual;
    unsigned long long max_cur, max_tot;   // bit patterns of non-negative doubles
    int reduced;                           // mixed form: orders whose operator product ran reduced
    int rnext;                             // mixed form: the next order's operator product runs reduced
};

// Per-item outputs gathered after the order loop.
struct ItemOut {
    double mode[3];       // TOA (I, Q, U)^m at the view node
    double boa_view;      // I^m downward at BOA, view node
    double boa_hemi;      // sum_o mu_o w_o I^m downward at BOA
    int orders, converged;
    double residual;
    int reduced;          // mixed form: orders whose operator product ran reduced (Pools::lo_tau)
};

struct InstOut {
    double kappa, tau_dir_ipss;
};

struct Pools {
    double *lev, *ring, *T, *greek, *R, *basis, *tab, *tot, *bc;
    double *layers;       // explicit layering of the instances that carry one
    double *dR;           // adjoint: d R^m / d sigma^2, layout of R
    float *Rf;            // mixed form: fp32 copy of pool_R read by the per-order rough boundary condition
    void *lam;            // adjoint: [L][P] per item, adjoint of the current order's field (fp64 plans: double,
                          // mixed plans: float)
    double *sc;           // adjoint: per layer and node the scan-step coefficients (T, al, be, dT, dal, dbe),
                          // [nt][n][6] at 6 off_T
    float *scf;           // mixed adjoint: their fp32 rounding, same layout (the adjoint sweep's table)
    double *rp0, *rp1;    // fp64 adjoint: where the rebuild of an order between checkpoints puts its operator
                          // parts A_0 F, A_1 F (gemm_f64<kGemmBoth>)
    double *grad;         // adjoint: per-instance accumulators (see DevInst::off_grad)
    const double *gout;  // adjoint: upstream gradient [instance][3] of (rho_I, rho_Q, rho_U)
    float *comp;          // fp32 compensation of the mixed-form order sum on the two output rows
    double *trow;         // mixed form: fp64 order sum of the two output rows (k = 0, k = nt), [2][P] per item
                          // at off_cmp (the mixed form keeps no order sum at the interior levels)
    float *Jf;            // mixed form: fp32 source of the items that keep their field as a pair, [L][P] at off_Flo
    float *lr;            // mixed form: output factors U of the rank-3 molecular operator, [3][P] at off_lr
    float *W;             // mixed form: its projections W = G F per level, [L][4] at off_W
    float *Tf;            // fp32 1 - T for the mixed-form sweeps
    float *idt;           // fp32 1 / dtau per layer, indexed like lev (mixed-form sweeps)
    void *A, *F, *J;      // double or bf16/float depending on the form
    void *J2;             // second output of a parts-split scattering integral (adjoint)
    void *Flo;            // bf16 low half of the field pair (mixed form, fsplit items)
    double *cm_w;         // cos/sin(m phi) table [(Mall+1)][2][kNPhi]
    float *cm_wf;         // mixed form: its fp32 rounding, same layout
    const double *rtop;   // water-side reflection kernels of in-water columns (DevInst::off_rtop)
    double *nodes;        // LUT / node-export plans: per item [6][n] TOA upward I, Q, U and BOA downward
                          // I, Q, U at every node
    const DevGas *gas_tab;    // gas absorption tables of the plan
    const double *gas_pool;   // their arrays
    // mixed form: an item whose estimated remaining share of its order sum,
    // res r / (1 - r) with res = max |I_k| / max |order sum| (the convergence
    // test's ratio) and r = res_k / res_(k-1) its last decay rate, is below
    // lo_tau takes its next operator product in the reduced form: the bf16
    // operator and the bf16 field, one MFMA per block, no low halves (0: never)
    float lo_tau;
    // mixed adjoint: the adjoint source Sbar of the current order as a bf16
    // pair (two planes, [L][P] per item at off_F), the A^T product's input
    __hip_bfloat16 *sb_hi, *sb_lo;
};

// Workgroups of a launch are dealt round-robin over the 8 XCDs (block b runs
// on XCD b mod 8, each XCD with its own L2).  Launches that give every item X
// blocks decode their block index so that all blocks of an item share one
// XCD and its L2: item = b mod 8 + 8 (b / 8 / X), part = (b / 8) mod X, over
// a grid of ceil(items / 8) * 8 * X blocks (blocks past the last item idle).
constexpr int kXcdCount = 8;
__device__ inline void xcd_item_block(int X, int &item, int &part)
{
    const int b = blockIdx.x, w = b / kXcdCount;
    item = b % kXcdCount + kXcdCount * (w / X);
    part = w % X;
}
__host__ inline unsigned xcd_item_grid(int items, int X)
{
    return (unsigned)(((items + kXcdCount - 1) / kXcdCount) * kXcdCount * X);
}

// Entry kl (row-major 3 x 3) of the rough-surface kernel R^m[o][i] for
// outgoing node o and incoming node i.  The outgoing node is the fastest
// index, so threads over the outgoing nodes read consecutive addresses.
__host__ __device__ inline long long r_index(const DevInst &I, int m, int o, int i, int kl)
{
    return I.off_R + (((long long)m * I.n + i) * 9 + kl) * I.n + o;
}

// Threads of a Cox-Munk Fourier workgroup: two azimuths of the half circle each.
constexpr int kCoxMunkThreads = 256;

// Rows of the phase tables computed by one workgroup.
constexpr int kTabRows = 8;

// Azimuth midpoints of the aerosol value kernel (rt_solver.c:922-925) and the
// half of them the even integrand is evaluated at.
constexpr int kVkPhi = 720;
constexpr int kVkHalf = kVkPhi / 2;

// P11 of an ascending angle table at scattering-angle cosine cth: linear in
// the angle (degrees), clamped to the end values (rt_solver.c:190-202,
// rt_ipss.c:134-146).
__host__ __device__ inline double p11_table_at(int n, const double *th, const double *p11, double cth)
{
    cth = cth > 1.0 ? 1.0 : (cth < -1.0 ? -1.0 : cth);
    const double theta = acos(cth) * 180.0 / OCRT_PI;
    if (theta <= th[0]) return p11[0];
    if (theta >= th[n - 1]) return p11[n - 1];
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (th[mid] <= theta) lo = mid; else hi = mid;
    }
    const double t = (theta - th[lo]) / (th[hi] - th[lo]);
    return p11[lo] + t * (p11[hi] - p11[lo]);
}

struct GemmTile { int item, tm, tn, pad; };   // item < 0: idle slot
// Tile shapes (operator rows x levels) of the bf16 scattering integral; the
// plan keeps one tile list per shape and runs each operator class with the
// shape its autotuner chose (all shapes give the same bits).
struct GemmShape { int tm, tn; };
constexpr GemmShape kGemmShapes[] = {{64, 128}, {64, 64}, {128, 64}, {128, 128}};
constexpr int kGemmShapeCount = 4;
// Level width of a scattering-integral tile: bf16 form (gemm_bf16) and fp64
// form (gemm_f64); the row width is 64 in both.
constexpr int kGemmLevels = 128;
constexpr int kGemmLevelsF64 = 64;

// Epilogues of the scattering-integral kernels.
constexpr int kGemmMix = 0;     // J = xdel A_p F + ydel A_r F
constexpr int kGemmParts = 1;   // J = A_0 F, J2 = A_1 F (unscaled)
constexpr int kGemmBoth = 2;    // J as kGemmMix, and the parts to Pools::rp0 / rp1 (fp64 adjoint rebuild)

// Up to 64 consecutive state columns [col0, col_end) of one item: the unit of
// the per-column kernels (one wavefront each).  A tile never mixes upward and
// downward columns, so all its lanes walk the levels the same way.  item < 0:
// idle slot.
struct ColTile { int item, col0, col_end, pad; };
constexpr int kColTile = 64;

// State columns of an item: [I_up Q_up U_up | I_dn Q_dn U_dn], n nodes each.
__host__ __device__ inline int col_up(int n, int c, int o) { return c * n + o; }
__host__ __device__ inline int col_dn(int n, int c, int o) { return (3 + c) * n + o; }
__host__ __device__ inline void col_decode(int col, int n, int &c, int &hemi, int &o)
{
    const int t = col / n;
    hemi = t >= 3 ? 1 : 0;
    c = t - 3 * hemi;
    o = col - t * n;
}
// Type of a column in the order [I_up I_dn Q_up Q_dn U_up U_dn] of the
// operator's coefficient selectors.
__host__ __device__ inline int col_sel_type(int col, int n)
{
    const int t = col / n;
    return t < 3 ? 2 * t : 2 * (t - 3) + 1;
}

// One IPSS kappa evaluation: instance and viewing direction.
struct IpssJob { int inst; int pad; double vza, raa; };

__host__ __device__ inline int roundup(int x, int a) { return (x + a - 1) / a * a; }

// Positive half of the 2n-point Gauss-Legendre rule (host; plan.hip).
int gauss_legendre_pos(int n, double *mu, double *w);

// Fourier kernels of a rough-interface matrix for modes 0..m_max into device
// memory d_out[m][o][i][9] from device node arrays, on stream st
// (interface.hip; the arguments are checked by the caller).
int interface_fourier_device(int kind, const double *d_mo, int n_o, const double *d_mi, int n_i, int m_max,
                             int n_phi, double wind, int sigma_type, double n_water, int q_convention,
                             double *d_out, hipStream_t st);

}  // namespace ocrt
