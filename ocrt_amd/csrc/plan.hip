// Host side of the MI355X SOS solver: instance validation, quadrature rings,
// device memory layout, chunking of large batches and the launch sequence.
// Implements the C ABI declared in include/ocrt_amd.h.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include <hip/hip_bf16.h>

#include "common.h"
#include "gas.h"
#include "ocrt_amd.h"

namespace ocrt {
// kernels (setup.hip, sos.hip, ipss.hip)
__global__ void medium_levels(const DevInst *, Pools);
__global__ void medium_fractions(const DevInst *, Pools);
__global__ void medium_gas(const DevInst *, Pools);
__global__ void medium_transmit(const DevInst *, Pools);
__global__ void adj_step_table(const DevInst *, Pools);
__global__ void lowrank_factors(const DevInst *, const DevItem *, Pools);
__global__ void fourier_weights(double *, float *, int);
template <bool DERIV, typename RT, int SET>
__global__ void coxmunk_fourier(const DevInst *, Pools, const int *, const int *, int, int);
__global__ void angular_basis(const DevInst *, const DevItem *, Pools);
__global__ void phase_tables(const DevInst *, const DevItem *, Pools, int, int);
__global__ void value_weights(double *, int);
__global__ void value_phase(const DevInst *, const DevItem *, Pools, const int *, const int *, int, int,
                            const double *, int);
template <typename AT, bool TRANS>
__global__ void pack_operator(const DevInst *, const DevItem *, Pools, int, int);
template <typename FT, bool MIXED>
__global__ void order_one(const DevInst *, const DevItem *, Pools, ItemState *, const ColTile *, int, const double *,
                          const double *);
template <typename FT> __global__ void order_status(const DevInst *, const DevItem *, Pools, ItemState *, int);
template <int MODE>
__global__ void gemm_f64(const DevInst *, const DevItem *, Pools, const GemmTile *, const ItemState *, int);
template <int NP, bool SPLIT, int TM, int TN>
__global__ void gemm_bf16(const DevInst *, const DevItem *, Pools, const GemmTile *, const ItemState *, int);
template <int NP, int MODE>
__global__ void gemm_adj(const DevInst *, const DevItem *, Pools, const GemmTile *, const ItemState *, int);
template <int NP>
__global__ void gemm_rebuild(const DevInst *, const DevItem *, Pools, const GemmTile *, const ItemState *, int);
template <typename FT, typename JT, bool MIXED, int U, int W>
__global__ void sweep(const DevInst *, const DevItem *, Pools, ItemState *, const ColTile *, int, int, const double *,
                      const float *, const float *);
__global__ void extract(const DevInst *, const DevItem *, Pools, const ItemState *, ItemOut *, int);
__global__ void finalize(const DevInst *, int, const ItemOut *, const InstOut *, const Pools, ocrt_result_t *);
__global__ void ipss_kappa(const DevInst *, const IpssJob *, Pools, double *, double *);
// adjoint.hip; FV: FieldF64 (fp64 plans) or FieldPair (mixed plans), LT the
// matching adjoint storage (double or float)
struct FieldF64;
struct FieldPair;
template <typename T> __global__ void fill(T *, long long, T);
__global__ void adj_zero_grad(const DevInst *, Pools);
template <typename LT> __global__ void adj_seed(const DevInst *, const DevItem *, Pools);
template <typename FV> __global__ void adj_bottom(const DevInst *, const DevItem *, Pools, const ItemState *, int);
__global__ void adj_surface_bc(const DevInst *, const DevItem *, Pools, const ItemState *, int);
template <typename FV, typename LT>
__global__ void sweep_adj(const DevInst *, const DevItem *, Pools, const ItemState *, const ColTile *, int, int,
                          const double *);
template <typename FV, typename LT>
__global__ void surface_adj(const DevInst *, const DevItem *, Pools, const ItemState *, int);
template <typename FV, typename LT>
__global__ void order_one_adj(const DevInst *, const DevItem *, Pools, const ColTile *, int, const double *);
__global__ void glint_adj(const DevInst *, int, Pools);
// lut.hip
__global__ void lut_eval(const DevInst *, const DevItem *, Pools, const InstOut *, const int *, const double *, int,
                         const double *, int, const double *, double *, double *, double *, double *, int);

// Positive half of the 2n-point Gauss-Legendre rule, ascending, weights summing
// to one: Newton iteration from Tricomi's estimate on the three-term
// recurrence (rt_quadrature.c:20-61).
int gauss_legendre_pos(int n, double *mu, double *w)
{
    const int N = 2 * n;
    for (int k = 1; k <= n; ++k) {
        double t = cos(OCRT_PI * ((double)k - 0.25) / ((double)N + 0.5));
        double pn = 0, pm = 0, dp = 0;
        bool ok = false;
        auto eval = [&](double x) {
            double p0 = 1.0, p1 = x;
            for (int q = 1; q < N; ++q) {
                const double p2 = ((2.0 * q + 1.0) * x * p1 - (double)q * p0) / (double)(q + 1);
                p0 = p1;
                p1 = p2;
            }
            pn = p1;
            pm = p0;
        };
        for (int it = 0; it < 100; ++it) {
            eval(t);
            dp = (double)N * (t * pn - pm) / (t * t - 1.0);
            const double dt = -pn / dp;
            t += dt;
            if (fabs(dt) < 1.0e-15 * (fabs(t) + 1.0)) { ok = true; break; }
        }
        if (!ok) return -1;
        eval(t);
        dp = (double)N * (t * pn - pm) / (t * t - 1.0);
        mu[n - k] = t;
        w[n - k] = 2.0 / ((1.0 - t * t) * dp * dp);
    }
    return 0;
}

}  // namespace ocrt

using namespace ocrt;

namespace {


// Insert the view cosine at its sorted position with zero weight unless a
// node lies within 1e-12 (rt_angle_grid.c:21-73).  Returns the ring size.
int build_ring(int n_gl, double mu_view, std::vector<double> &mu, std::vector<double> &w, int &vidx)
{
    mu.assign(n_gl + 1, 0.0);
    w.assign(n_gl + 1, 0.0);
    if (gauss_legendre_pos(n_gl, mu.data(), w.data()) != 0) return -1;
    int n = n_gl;
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (mu[mid] < mu_view) lo = mid + 1; else hi = mid;
    }
    const int pos = lo;
    if (pos > 0 && fabs(mu[pos - 1] - mu_view) < 1.0e-12) { vidx = pos - 1; mu.resize(n); w.resize(n); return n; }
    if (pos < n && fabs(mu[pos] - mu_view) < 1.0e-12) { vidx = pos; mu.resize(n); w.resize(n); return n; }
    for (int i = n; i > pos; --i) { mu[i] = mu[i - 1]; w[i] = w[i - 1]; }
    mu[pos] = mu_view;
    w[pos] = 0.0;
    vidx = pos;
    return n + 1;
}

struct HostInst {
    DevInst d;           // offsets of chunk-local pools filled per chunk
    std::vector<double> mu, w;
    int status;
    long long lev_n, T_n, R_n, greek_n;
    int op_src = -1;     // an earlier instance with the same operator inputs, -1: none
};

// Window of earlier instances searched for one whose surface kernels or
// scattering operator an instance can share.
constexpr int kShareWindow = 64;

struct Chunk {
    int i0, i1, it0, it1;       // instance and item ranges
    std::vector<DevInst> hinst;
    std::vector<DevItem> hitem;
    // GEMM tiles: single-part items with a single bf16 field [0, ntiles1),
    // single-part items with a field pair [ntiles1, ntiles2), two-part items
    // [ntiles2, size)
    std::vector<GemmTile> htiles;
    int ntiles1 = 0, ntiles2 = 0;
    // mixed form: the same three classes tiled with every shape of
    // kGemmShapes, class k at [cls_off[s][k], cls_off[s][k + 1]) of list s
    std::vector<GemmTile> htiles_s[kGemmShapeCount];
    int cls_off[kGemmShapeCount][4] = {};
    GemmTile *dtiles_s[kGemmShapeCount] = {};
    std::vector<ColTile> hcols;
    std::vector<IpssJob> hjobs;
    std::vector<IpssJob> hjobs_lut;   // IPSS factors of every LUT direction
    IpssJob *djobs_lut = nullptr;
    int *djob_lut_of = nullptr;       // first LUT job of each instance, -1 without IPSS
    DevInst *dinst = nullptr;
    DevItem *ditem = nullptr;
    GemmTile *dtiles = nullptr;
    ColTile *dcols = nullptr;
    // rough-surface kernel owners of the chunk and the prefix sums of their
    // (outgoing, incoming-or-sun) pair counts: one Cox-Munk workgroup per pair
    // Cox-Munk owners and the prefix sums of their pairs: every pair (n^2 + n
    // each), and in mixed plans the solar (n) and diffuse (n^2) sets apart
    std::vector<int> hcm_owner, hcm_first, hcm_sfirst, hcm_dfirst;
    int *dcm_owner = nullptr, *dcm_first = nullptr, *dcm_sfirst = nullptr, *dcm_dfirst = nullptr;
    // instances of the chunk whose aerosol I -> I table comes from the value
    // kernel (operator owners only) and the prefix sums of their (row,
    // column) entry counts: one value_phase workgroup per entry
    std::vector<int> hvk_owner, hvk_first;
    int *dvk_owner = nullptr, *dvk_first = nullptr;
    IpssJob *djobs = nullptr;
    int *djob_of = nullptr;
    int max_L = 0, max_P = 0, max_pairs = 0, max_tab = 0, tab_groups = 1, maxit = 1, any_rough = 0, n_jobs = 0;
    int any_gas = 0;
    long long max_pack = 0;
    long long lev = 0, T = 0, R = 0, basis = 0, tab = 0, A = 0, F = 0, bc = 0, cmp = 0, flo = 0, lr = 0, wlr = 0;
};

// Mixed form: Fourier modes m < kSplitModes keep their field as a bf16 pair
// (about 16 significant bits) and pay a third MFMA (A_hi F_lo).  These modes
// carry most of the radiance; a single bf16 field there leaves errors near
// 1e-4 in I after 20 orders of a thick aerosol layer, while at m >= 3 the
// same rounding contributes ~1e-6.
constexpr int kSplitModes = 3;

// Launch shapes of plans that do not autotune: the sweep's (8 levels per
// block, 4 wavefronts per workgroup; kSweepShapes index 0) and the
// scattering integral's per operator class (128 x 128 tiles for the
// single-part classes, 64 x 128 for two-part items; kGemmShapes indices), the
// shapes the autotuner picks on the 8192-instance workload or within 2 % of
// them.
constexpr int kDefaultSweepShape = 0;
constexpr int kDefaultGemmShape[3] = {3, 3, 0};

// Default estimated remaining share below which a mixed-form item's next
// order takes the reduced operator product (ocrt_plan_set_reduced_threshold).
// On the batched workload of 8192 uneven instances (modes 0..16, 50-400
// layers, 16-64 streams) 1e-3 runs 57 % of the item-orders reduced, cuts the
// scattering integral from 161 to 140 ms per execute and moves the largest
// relative error of rho_I against the fp64 form from 2.75e-5 to 2.86e-5
// (3e-3: 3.3e-5, 1e-2: 1.1e-4).
constexpr float kReducedTau = 1e-3f;

// Workgroups of one launch are dealt round-robin over the 8 XCDs (block b runs
// on XCD b mod 8), each with its own L2.  The GEMM tiles of one item share
// operator rows and field columns, so every item's tiles are kept on one XCD:
// items go to the least-loaded of 8 groups, and slot 8j + x of the launch
// takes the j-th tile of group x (idle slots pad the shorter groups).
constexpr int kXcds = 8;

std::vector<GemmTile> xcd_tile_order(const std::vector<std::vector<GemmTile>> &per_item)
{
    std::vector<std::vector<GemmTile>> grp(kXcds);
    std::vector<size_t> load(kXcds, 0);
    std::vector<int> order(per_item.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
    // largest items first keeps the groups balanced
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return per_item[a].size() > per_item[b].size(); });
    for (int i : order) {
        const int g = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        grp[g].insert(grp[g].end(), per_item[i].begin(), per_item[i].end());
        load[g] += per_item[i].size();
    }
    const size_t rows = *std::max_element(load.begin(), load.end());
    std::vector<GemmTile> out;
    out.reserve(rows * kXcds);
    for (size_t j = 0; j < rows; ++j)
        for (int x = 0; x < kXcds; ++x)
            out.push_back(j < grp[x].size() ? grp[x][j] : GemmTile{-1, 0, 0, 0});
    return out;
}

#define HIPCHK(x)                                                                         \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) {                                                           \
            fprintf(stderr, "ocrt_amd: %s failed: %s\n", #x, hipGetErrorString(e_));      \
            return OCRT_ERR_DEVICE;                                                       \
        }                                                                                 \
    } while (0)

}  // namespace

struct ocrt_plan {
    int n_inst = 0, prec = 0, per_inst = 0, mmax_all = 0, n_items = 0;
    long long mx_F = 0;   // entries of the largest chunk's field
    std::vector<HostInst> inst;
    std::vector<Chunk> chunks;
    std::vector<int> item_of_inst;   // first global item of each instance
    // global device arrays
    double *ring = nullptr, *greek = nullptr, *cm_w = nullptr, *layers = nullptr;
    float *cm_wf = nullptr;
    double *vk_w = nullptr;   // cos(m phi_q) of the value kernel, [360][mmax_all + 1]; null without one
    double *rtop = nullptr;   // water-side Cox-Munk kernels of rough in-water columns (DevInst::off_rtop)
    std::vector<int> rtop_owner;   // instances that compute a kernel set (the others share one)
    ItemState *state = nullptr;
    ItemOut *iout = nullptr;
    InstOut *ipss = nullptr;
    ocrt_result_t *res = nullptr;
    double *kappa = nullptr, *taudir = nullptr;
    DevGas *gas_tab = nullptr;
    double *gas_pool = nullptr;
    // chunk pools (sized for the largest chunk)
    Pools pools{};
    size_t dev_bytes = 0;
    std::vector<void *> allocs;
    double gemm_flops = 0.0, gemm_flops_alg = 0.0, sweep_bytes = 0.0, gemm_launches = 0.0;
    // mixed form, from the last results: MFMA flops launched by the scattering
    // integral (padding, low halves and reduced orders included), item-orders
    // that ran it, and those that ran it reduced
    double gemm_issued = 0.0, item_orders = 0.0, reduced_orders = 0.0;
    // optional per-kernel timing: events recorded around every GEMM and sweep launch
    int profile = 0;
    // launch shape of the mixed-form sweep (kSweepShapes), optionally chosen
    // by timing every shape on the plan's own work at the first execute
    int sweep_variant = kDefaultSweepShape;
    int autotune = 0, tuned = 0;
    std::vector<double> tune_ms;
    // tile shape (kGemmShapes) of each operator class of the bf16 scattering
    // integral: single-part single-field, single-part pair, two-part; and the
    // autotuner's per-class GEMM time with each shape
    int gemm_shape[3] = {kDefaultGemmShape[0], kDefaultGemmShape[1], kDefaultGemmShape[2]};
    int gemm_tuned = 0, tune_classes = 0;
    double gemm_ms[3][kGemmShapeCount] = {};
    // replay of the captured forward launch sequence (hipGraph)
    int use_graph = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipStream_t own_stream = nullptr;
    double graph_gemm_launches = 0.0;
    std::vector<hipEvent_t> ev;       // pairs (start, stop): [gemm..., sweep...]
    std::vector<int> ev_kind;         // 0 gemm, 1 sweep, per pair
    size_t ev_used = 0;
    // adjoint plans: transposed operator, order checkpoints (u_1, u_{1+e},
    // u_{1+2e}, ... for interval e), the e - 1 recomputed fields of a segment,
    // per-instance accumulators
    int adjoint = 0;
    int adj_mixed = 0;                // mixed adjoint plan: mixed-form fields, fp32 adjoint (backward_chunk)
    void *Jadj = nullptr;             // mixed adjoint: fp32 first operator part / adjoint source, [L][P] per item
    // adjoint with sparse checkpoints: the operator parts (A_0 u, A_1 u) of
    // the orders rebuilt between two checkpoints, one pair per slot (fp64, or
    // fp32 in mixed plans)
    std::vector<void *> rpart0, rpart1;
    int dev_inputs = 0;               // OCRT_PLAN_DEVICE_INPUTS: no shared surface kernels
    int fwd_ready = 0;                // the last execute left this plan's checkpoints intact
    int ckpt_every = 1;
    void *At = nullptr;
    std::vector<void *> ckpt;
    std::vector<void *> seg;
    double *grad = nullptr, *gout = nullptr;
    long long grad_n = 0;
    // LUT output (ocrt_plan_set_lut): per-node fields of every item and the grid
    int lut = 0, lut_nv = 0, lut_nr = 0;
    long long nodes_n = 0;
    double *nodes = nullptr, *lut_vza = nullptr, *lut_raa = nullptr, *lut_grid = nullptr, *lut_scal = nullptr;
    double *lut_raw = nullptr, *lut_vpm = nullptr;
    double *kappa_lut = nullptr, *taudir_lut = nullptr;
};


// An event pair (start, stop) of kind `kind` from the plan's pool, by value:
// growing the pool must not invalidate pairs handed out before.
struct EvPair {
    hipEvent_t e[2];
    bool ok;
};
static EvPair ev_pair(ocrt_plan *p, int kind)
{
    if (p->ev_used + 2 > p->ev.size()) {
        for (int q = 0; q < 64; ++q) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return EvPair{{nullptr, nullptr}, false};
            p->ev.push_back(e);
        }
        p->ev_kind.resize(p->ev.size() / 2, 0);
    }
    const EvPair r{{p->ev[p->ev_used], p->ev[p->ev_used + 1]}, true};
    p->ev_kind[p->ev_used / 2] = kind;
    p->ev_used += 2;
    return r;
}

static int dmalloc(ocrt_plan *p, void **ptr, size_t bytes)
{
    if (bytes == 0) bytes = 256;
    bytes = (bytes + 255) & ~(size_t)255;
    const hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
        fprintf(stderr, "ocrt_amd: hipMalloc(%zu) failed: %s\n", bytes, hipGetErrorString(e));
        *ptr = nullptr;
        return OCRT_ERR_ALLOC;
    }
    p->allocs.push_back(*ptr);
    p->dev_bytes += bytes;
    return 0;
}

static double layer_sum(const ocrt_instance_t &d, int row)
{
    double s = 0.0;
    for (int l = 0; l < d.n_layers; ++l) s += d.layers[(size_t)row * d.n_layers + l];
    return s;
}

// Aerosol scattering enters only with a phase matrix and a positive column.
static bool has_aerosol(const ocrt_instance_t &d)
{
    if (d.water_tau > 0.0) return d.L_max > 0 && d.greek && d.water_omega_p > 0.0;
    if (d.layers) return d.L_max > 0 && d.greek && layer_sum(d, 1) > 0.0;
    return d.L_max > 0 && d.tau_a > 0.0 && d.greek;
}

static int validate(const ocrt_instance_t &d)
{
    if (!(d.wavelength_nm >= 330.0 - 1e-9 && d.wavelength_nm <= 1100.0 + 1e-9)) return OCRT_ERR_ARGS;
    if (!(d.sza_deg >= 0.0 && d.sza_deg < 90.0)) return OCRT_ERR_ARGS;
    if (!(d.vza_deg >= 0.0 && d.vza_deg < 90.0)) return OCRT_ERR_ARGS;
    if (d.m_max < 0 || d.m_max > 1024) return OCRT_ERR_ARGS;
    if (d.n_mu < 2 || d.n_mu > kMaxRing - 1 || d.n_layers < 2) return OCRT_ERR_ARGS;
    if (d.max_orders < 1 || !(d.tolerance > 0.0)) return OCRT_ERR_ARGS;
    const bool aer = has_aerosol(d);
    if (d.water_tau > 0.0) {
        // in-water column: albedos in [0, 1], a black bottom, no layering, gas,
        // IPSS or bottom source
        if (!(std::isfinite(d.water_tau) && d.water_omega_w >= 0.0 && d.water_omega_p >= 0.0 &&
              d.water_omega_w + d.water_omega_p <= 1.0 && d.beam_scale >= 0.0 && std::isfinite(d.beam_scale)))
            return OCRT_ERR_ARGS;
        if (d.layers || d.gas || d.ipss || d.bottom_source || d.bottom_source_only) return OCRT_ERR_ARGS;
        if (d.surface != OCRT_SURFACE_BLACK) return OCRT_ERR_ARGS;
        if (d.top_reflect < 0 || d.top_reflect > 2) return OCRT_ERR_ARGS;
        if (d.top_reflect && !(d.n_water > 1.0)) return OCRT_ERR_ARGS;
        // the rough kernels come from the all-mode interface launch (m <= 255)
        if (d.top_reflect == 2 && (!(d.wind_speed > 0.0) || d.m_max > 255)) return OCRT_ERR_ARGS;
        if (aer && d.L_max < 2) return OCRT_ERR_ARGS;
    } else if (d.top_reflect || d.top_source) {
        return OCRT_ERR_ARGS;   // top reflection and top source belong to in-water columns
    } else if (d.layers) {
        // every layer needs a positive extinction and an albedo in [0, 1]
        const int nt = d.n_layers;
        const double *R = d.layers, *A = R + nt, *W = A + nt, *G = W + nt;
        for (int l = 0; l < nt; ++l) {
            if (!(R[l] >= 0.0 && A[l] >= 0.0 && G[l] >= 0.0 && W[l] >= 0.0 && W[l] <= 1.0)) return OCRT_ERR_ARGS;
            if (!(R[l] + A[l] + G[l] > 0.0)) return OCRT_ERR_ARGS;
        }
        if (d.ipss || d.gas) return OCRT_ERR_ARGS;   // absorption of a layering is its layers[3]
        if (aer && d.L_max < 2) return OCRT_ERR_ARGS;
    } else {
        if (aer && (d.ssa_a < 0.0 || d.ssa_a > 1.0 || d.L_max < 2)) return OCRT_ERR_ARGS;
        if (d.tau_a < 0.0 || d.tau_R < 0.0) return OCRT_ERR_ARGS;
        if (!(d.tau_R + (aer ? d.tau_a : 0.0) > 0.0)) return OCRT_ERR_ARGS;
    }
    if (d.n_ang_phase != 0) {
        // the value kernel's P11 table: ascending angles in [0, 180], finite values
        if (d.n_ang_phase < 2 || !d.theta_phase || !d.p11_phase) return OCRT_ERR_ARGS;
        for (int a = 0; a < d.n_ang_phase; ++a) {
            if (!std::isfinite(d.theta_phase[a]) || !std::isfinite(d.p11_phase[a])) return OCRT_ERR_ARGS;
            if (d.theta_phase[a] < 0.0 || d.theta_phase[a] > 180.0) return OCRT_ERR_ARGS;
            if (a > 0 && d.theta_phase[a] < d.theta_phase[a - 1]) return OCRT_ERR_ARGS;
        }
    }
    if (d.surface < 0 || d.surface > 3) return OCRT_ERR_ARGS;
    if ((d.surface == OCRT_SURFACE_FLAT || d.surface == OCRT_SURFACE_BLACK_FRESNEL_OCEAN) && !(d.n_water > 0.0))
        return OCRT_ERR_ARGS;
    if (d.ipss && d.n_layers > 1536) return OCRT_ERR_ARGS;
    return 0;
}

extern "C" int ocrt_plan_create_ex(const ocrt_instance_t *in, int n_inst, int precision, int flags,
                                   ocrt_plan **out)
{
    if (!in || n_inst < 1 || !out || (precision != OCRT_PREC_FP64 && precision != OCRT_PREC_MIXED))
        return OCRT_ERR_ARGS;
    if (flags & ~(OCRT_PLAN_ONE_PER_LAUNCH | OCRT_PLAN_ADJOINT | OCRT_PLAN_CKPT_SPARSE | OCRT_PLAN_AUTOTUNE |
                  OCRT_PLAN_GRAPH | OCRT_PLAN_DEVICE_INPUTS))
        return OCRT_ERR_ARGS;
    const bool adjoint = (flags & OCRT_PLAN_ADJOINT) != 0;
    // the adjoint needs the plane-parallel solve (no IPSS).  With precision
    // MIXED it runs the mixed order loop with every field kept as its bf16 pair
    // (16 significant bits; no rank-3 molecular items, no reduced products) and
    // keeps the adjoint fields in fp32; with FP64 everything is fp64
    const bool adj_mixed = adjoint && precision == OCRT_PREC_MIXED;
    if (adjoint)
        for (int i = 0; i < n_inst; ++i)
            if (in[i].ipss || in[i].bottom_source || in[i].bottom_source_only || in[i].water_tau > 0.0)
                return OCRT_ERR_ARGS;
    *out = nullptr;
    ocrt_plan *p = new ocrt_plan();
    p->n_inst = n_inst;
    p->prec = precision;
    p->adj_mixed = adj_mixed ? 1 : 0;
    p->per_inst = (flags & OCRT_PLAN_ONE_PER_LAUNCH) ? 1 : 0;
    p->autotune = (flags & OCRT_PLAN_AUTOTUNE) && !adjoint ? 1 : 0;
    p->use_graph = (flags & OCRT_PLAN_GRAPH) && !adjoint ? 1 : 0;
    p->adjoint = adjoint ? 1 : 0;
    p->dev_inputs = (flags & OCRT_PLAN_DEVICE_INPUTS) ? 1 : 0;
    const bool mixed = p->prec == OCRT_PREC_MIXED;
    const bool gemm_bf = mixed;   // bf16-pair operator, 32-wide K slabs, 128-level tiles
    const int Palign = gemm_bf ? 32 : 16;
    p->inst.resize(n_inst);
    long long ring_n = 0, greek_n = 0, layers_n = 0;
    std::vector<const ocrt_gas *> gas_used;
    for (int i = 0; i < n_inst; ++i) {
        const ocrt_instance_t &d = in[i];
        HostInst &h = p->inst[i];
        h.status = validate(d);
        if (h.status != 0) { delete p; return h.status; }
        DevInst &I = h.d;
        memset(&I, 0, sizeof(I));
        I.sza = d.sza_deg; I.vza = d.vza_deg; I.raa = d.raa_deg;
        I.mus = cos(d.sza_deg * OCRT_PI / 180.0);
        I.muv = cos(d.vza_deg * OCRT_PI / 180.0);
        I.aer = has_aerosol(d) ? 1 : 0;
        I.aer_able = (d.L_max > 0 && d.greek) ? 1 : 0;
        I.layered = d.layers ? 1 : 0;
        I.water = d.water_tau > 0.0 ? 1 : 0;
        I.wtau = d.water_tau; I.womega_w = d.water_omega_w; I.womega_p = d.water_omega_p;
        I.beam_scale = d.beam_scale;
        I.top_kind = I.water ? d.top_reflect : 0;
        if (I.water) {
            I.tauR = d.water_omega_w * d.water_tau;
            I.taua = I.aer ? d.water_omega_p * d.water_tau : 0.0;
        } else if (d.layers) {
            I.tauR = layer_sum(d, 0);
            I.taua = I.aer ? layer_sum(d, 1) : 0.0;
        } else {
            I.tauR = d.tau_R > 0.0 ? d.tau_R : 0.0;
            I.taua = I.aer ? d.tau_a : 0.0;
        }
        I.ray = I.tauR > 0.0;
        I.ssa = d.ssa_a; I.aer_h = d.aer_h_km; I.n_water = d.n_water; I.wind = d.wind_speed;
        I.tol = d.tolerance;
        const double dp = (d.depol >= 0.0 && d.depol < 1.0) ? d.depol : 0.0279;
        const double an = 2.0 * (1.0 - dp) / (2.0 + dp);
        I.beta0 = 1.0; I.beta2 = 0.5 * an; I.gamma2 = -an * sqrt(1.5); I.alpha2 = 3.0 * an;
        int vidx = 0;
        const int n = build_ring(d.n_mu, I.muv, h.mu, h.w, vidx);
        if (n < 2) { delete p; return OCRT_ERR_ARGS; }
        I.n = n; I.vidx = vidx; I.nt = d.n_layers; I.mmax = d.m_max;
        I.La = I.aer ? d.L_max : 0;
        I.lmax = std::max(std::max(d.m_max, 2), I.aer ? d.L_max : 0);
        const bool surf_aware = d.surface == OCRT_SURFACE_FLAT || d.surface == OCRT_SURFACE_BLACK_FRESNEL_OCEAN;
        const bool flat = d.surface == OCRT_SURFACE_FLAT ||
                          (d.surface == OCRT_SURFACE_BLACK_FRESNEL_OCEAN && d.wind_speed <= 0.0);
        I.surf = surf_aware ? (flat ? 1 : 2) : 0;
        I.surface_kind = d.surface;
        I.sigma_type = d.sigma_type; I.q_conv = d.q_convention; I.decouple = d.decouple_sunglint;
        I.ipss = d.ipss ? 1 : 0; I.max_orders = d.max_orders;
        I.off_ring = ring_n;
        ring_n += 2 * n;
        I.off_greek = greek_n;
        greek_n += I.aer ? 4LL * (I.La + 1) : 0;
        I.n_ang = (I.aer && d.n_ang_phase >= 2) ? d.n_ang_phase : 0;
        I.off_ang = greek_n;
        greek_n += 2LL * I.n_ang;
        // external bottom source: shaped on this ring (rt_solver.c:687-706)
        I.bsrc_mmax = -1;
        I.bsrc_only = d.bottom_source_only ? 1 : 0;
        I.off_bsrc = greek_n;
        if (d.bottom_source) {
            if (d.bottom_n_mu != n || d.bottom_m_max < 0) { delete p; return OCRT_ERR_BOTTOM_SHAPE; }
            I.bsrc_mmax = d.bottom_m_max;
            greek_n += 3LL * (d.bottom_m_max + 1) * n;
        }
        // diffuse top source of an in-water column, shaped on this ring
        I.tsrc_mmax = -1;
        I.off_tsrc = greek_n;
        if (d.top_source) {
            if (d.top_n_mu != n || d.top_m_max < 0) { delete p; return OCRT_ERR_BOTTOM_SHAPE; }
            I.tsrc_mmax = d.top_m_max;
            greek_n += 3LL * (d.top_m_max + 1) * n;
        }
        // fp16 source scale of the mixed form: the largest source magnitude S
        // (the direct beam's pi, times the refracted beam's scale in water, and
        // the external sources) maps to [2^5, 2^6).  Fields stay below about S
        // and the scattering integral below a few times the field, so the
        // scaled source keeps clear of fp16's largest value (65504) while
        // everything above 2^-21 S stays in its normal range.
        {
            double smax = I.bsrc_only ? 0.0 : OCRT_PI * (I.water ? d.beam_scale : 1.0);
            for (long long e = 0; d.bottom_source && e < 3LL * (I.bsrc_mmax + 1) * n; ++e)
                smax = std::max(smax, std::fabs(d.bottom_source[e]));
            for (long long e = 0; d.top_source && e < 3LL * (I.tsrc_mmax + 1) * n; ++e)
                smax = std::max(smax, std::fabs(d.top_source[e]));
            int ex = 0;
            if (smax > 0.0 && std::isfinite(smax)) frexp(smax, &ex);
            const int js = std::min(std::max(6 - ex, -100), 100);
            I.jscale = ldexpf(1.0f, js);
            I.jinv = ldexpf(1.0f, -js);
        }
        I.off_layers = layers_n;
        layers_n += I.layered ? 4LL * I.nt : 0;
        I.gidx = i;
        I.wl = d.wavelength_nm;
        I.gas = -1;
        if (d.gas) {
            size_t gi = 0;
            while (gi < gas_used.size() && gas_used[gi] != d.gas) ++gi;
            if (gi == gas_used.size()) gas_used.push_back(d.gas);
            I.gas = (int)gi;
        }
        I.off_grad = p->grad_n;
        p->grad_n += 4LL * (I.nt + 1);
        h.lev_n = 7LL * (I.nt + 1);
        h.T_n = (long long)I.nt * n;
        h.R_n = (I.surf == 2) ? (long long)(I.mmax + 1) * n * n * 9 + (long long)(I.mmax + 1) * n * 9 : 0;
        p->mmax_all = std::max(p->mmax_all, I.mmax);
    }
    // Water-side reflection kernels of rough in-water columns, shared by the
    // instances with the same ring, modes and surface.
    long long rtop_n = 0;
    for (int i = 0; i < n_inst; ++i) {
        HostInst &h = p->inst[i];
        DevInst &I = h.d;
        if (I.top_kind != 2) continue;
        int owner = -1;
        for (int j : p->rtop_owner) {
            const HostInst &g = p->inst[j];
            const DevInst &J = g.d;
            if (J.n == I.n && J.mmax == I.mmax && J.wind == I.wind && J.sigma_type == I.sigma_type &&
                J.n_water == I.n_water && J.q_conv == I.q_conv && g.mu == h.mu && g.w == h.w) {
                owner = j;
                break;
            }
        }
        if (owner >= 0) {
            I.off_rtop = p->inst[owner].d.off_rtop;
        } else {
            I.off_rtop = rtop_n;
            rtop_n += (long long)(I.mmax + 1) * I.n * I.n * 9;
            p->rtop_owner.push_back(i);
        }
    }
    // Instances whose basis, phase tables and scattering operator are equal
    // bit for bit: same ring (nodes, weights, view node), sun, Fourier range,
    // phase-matrix expansion and Rayleigh coefficients.  Only their optical
    // depths and scattering fractions differ, which enter per level.
    auto same_operator = [&](int a, int b) {
        const HostInst &ha = p->inst[a], &hb = p->inst[b];
        const DevInst &A = ha.d, &B = hb.d;
        if (A.n != B.n || A.mus != B.mus || A.mmax != B.mmax || A.lmax != B.lmax || A.aer != B.aer ||
            A.ray != B.ray || A.La != B.La || A.beta0 != B.beta0 || A.beta2 != B.beta2 || A.gamma2 != B.gamma2 ||
            A.alpha2 != B.alpha2 || ha.mu != hb.mu || ha.w != hb.w)
            return false;
        if (A.aer)
            for (int e = 0; e < 4 * (A.La + 1); ++e)
                if (in[a].greek[e] != in[b].greek[e]) return false;
        if (A.n_ang != B.n_ang) return false;
        for (int e = 0; e < A.n_ang; ++e)
            if (in[a].theta_phase[e] != in[b].theta_phase[e] || in[a].p11_phase[e] != in[b].p11_phase[e])
                return false;
        return true;
    };
    for (int i = 0; i < n_inst && !p->per_inst; ++i)
        for (int j = std::max(0, i - kShareWindow); j < i; ++j)
            if (p->inst[j].op_src < 0 && same_operator(i, j)) { p->inst[i].op_src = j; break; }
    // ---- chunking: instances grouped so one chunk's working set stays bounded
    const double budget = mixed ? 48e9 : 48e9;
    const size_t fe = mixed ? 2 : 8, je = mixed ? 2 : 8, ae = gemm_bf ? 2 : 8;
    int item_global = 0;
    p->item_of_inst.resize(n_inst);
    // Adjoint plans keep the field of every ckpt-th order and recompute the
    // orders between.  The smallest interval of 1, 2, 4 that keeps the whole
    // batch in one chunk is taken (no recomputation when the fields of all
    // orders fit), unless the caller asks for the sparse form.
    auto partition = [&](int every) {
    p->chunks.clear();
    item_global = 0;
    int i = 0;
    while (i < n_inst) {
        Chunk c;
        c.i0 = i;
        c.it0 = item_global;
        double bytes = 0.0;
        while (i < n_inst) {
            HostInst &h = p->inst[i];
            const DevInst &I = h.d;
            double b = (double)(h.lev_n + 2 * h.T_n + h.R_n) * 8.0;
            const int n = I.n, L = I.nt + 1, P = roundup(6 * n, Palign), Mp = roundup(6 * n, 64);
            const bool shared_op = h.op_src >= c.i0;   // its source is in this chunk
            for (int m = 0; m <= I.mmax; ++m) {
                const int parts = (I.aer ? 1 : 0) + ((m <= 2 && I.ray) ? 1 : 0);
                const int K = std::max(parts, 1) * P;
                if (!shared_op)
                    b += 3.0 * (I.lmax + 1) * (2 * n + 1) * 8 + 6.0 * (n + 1) * (2 * n + 1) * 8 +
                         (double)Mp * K * ae * (gemm_bf ? 2 : 1) * (adjoint ? 2 : 1);
                // field and source; the fp64 form's order sum at every level;
                // the mixed form's output rows, and the bf16 residual plus the
                // fp32 source of its split modes
                b += (double)L * P * (fe + je + (mixed ? 0 : 8)) + (mixed ? 24.0 * P : 0.0) +
                     ((mixed && m < kSplitModes) ? 6.0 * L * P : 0.0) + 6.0 * n * 8;
                if (adjoint)   // checkpoints, segment fields, lambda, second parts
                    b += (double)L * P * (mixed ? 4 : 8) * ((I.max_orders - 1) / every + 1 + (every - 1) + 2);
            }
            if (adjoint) b += (double)h.R_n * 8;
            const int items_after = (item_global - c.it0) + I.mmax + 1;
            if (i > c.i0 && (bytes + b > budget || items_after > 60000 || i - c.i0 >= 60000 || p->per_inst)) break;
            bytes += b;
            p->item_of_inst[i] = item_global;
            item_global += I.mmax + 1;
            ++i;
        }
        c.i1 = i;
        c.it1 = item_global;
        p->chunks.push_back(c);
    }
    };
    p->ckpt_every = 1;
    if (adjoint && (flags & OCRT_PLAN_CKPT_SPARSE)) p->ckpt_every = 4;
    partition(p->ckpt_every);
    while (adjoint && p->chunks.size() > 1 && p->ckpt_every < 4) partition(p->ckpt_every *= 2);
    p->n_items = item_global;
    // ---- per-chunk layout
    for (Chunk &c : p->chunks) {
        long long lev = 0, T = 0, R = 0, basis = 0, tab = 0, A = 0, F = 0, bc = 0, cmp = 0, flo = 0, lr = 0, wlr = 0;
        // items grouped by the operator they use: class of each owner item and
        // the member items of each class
        std::vector<int> class_of_item;
        std::vector<std::vector<int>> class_items;
        std::vector<char> op_owner(c.i1 - c.i0, 0);
        for (int ii = c.i0; ii < c.i1; ++ii) {
            HostInst &h = p->inst[ii];
            DevInst I = h.d;
            // share the operator of an earlier owner of this chunk with equal inputs
            int op_from = -1;
            if (h.op_src >= c.i0)
                for (int jj = std::max(c.i0, ii - kShareWindow); jj < ii && op_from < 0; ++jj)
                    if (op_owner[jj - c.i0] && same_operator(ii, jj)) op_from = jj - c.i0;
            op_owner[ii - c.i0] = op_from < 0;
            I.off_lev = lev; lev += h.lev_n;
            I.off_T = T; T += h.T_n;
            if (I.surf == 2) {
                // share the surface kernels with an earlier instance of this chunk
                // that has the same ring, sun and surface (not when the wind
                // speeds may change later)
                int owner = -1;
                for (int jj = std::max(c.i0, ii - kShareWindow); jj < ii && owner < 0 && !p->dev_inputs; ++jj) {
                    const HostInst &g = p->inst[jj];
                    const DevInst &J = c.hinst[jj - c.i0];
                    if (J.surf == 2 && J.surf_owner && J.n == I.n && J.mmax == I.mmax && J.mus == I.mus &&
                        J.wind == I.wind && J.sigma_type == I.sigma_type && J.n_water == I.n_water &&
                        J.q_conv == I.q_conv && g.mu == h.mu && g.w == h.w)
                        owner = jj - c.i0;
                }
                if (owner >= 0) {
                    I.off_R = c.hinst[owner].off_R;
                    I.off_Rsun = c.hinst[owner].off_Rsun;
                    I.surf_owner = 0;
                } else {
                    I.off_R = R;
                    I.off_Rsun = R + (long long)(I.mmax + 1) * I.n * I.n * 9;
                    R += h.R_n;
                    I.surf_owner = 1;
                    if (c.hcm_first.empty()) {
                        c.hcm_first.push_back(0);
                        c.hcm_sfirst.push_back(0);
                        c.hcm_dfirst.push_back(0);
                    }
                    c.hcm_owner.push_back(ii - c.i0);
                    c.hcm_first.push_back(c.hcm_first.back() + I.n * I.n + I.n);
                    c.hcm_sfirst.push_back(c.hcm_sfirst.back() + I.n);
                    c.hcm_dfirst.push_back(c.hcm_dfirst.back() + I.n * I.n);
                }
            }
            I.item0 = p->item_of_inst[ii] - c.it0;
            const int n = I.n, L = I.nt + 1, Pp = roundup(6 * n, Palign), Mp = roundup(6 * n, 64);
            c.max_L = std::max(c.max_L, L);
            c.max_P = std::max(c.max_P, Pp);
            c.max_pairs = std::max(c.max_pairs, n * n + n);
            c.max_tab = std::max(c.max_tab, (n + 1) * (2 * n + 1));
            c.tab_groups = std::max(c.tab_groups, (n + 1 + kTabRows - 1) / kTabRows);
            c.maxit = std::max(c.maxit, I.max_orders);
            c.any_rough |= (I.surf == 2);
            c.any_gas |= (I.gas >= 0);
            for (int m = 0; m <= I.mmax; ++m) {
                DevItem it;
                memset(&it, 0, sizeof(it));
                it.inst = ii - c.i0; it.m = m; it.n = n; it.nt = I.nt; it.L = L; it.P = Pp;
                int kinds[2], np = 0;
                // the mixed form takes the molecular operator of an item with
                // aerosol out of the dense product (rank 3, see DevItem::lowrank)
                it.lowrank = (mixed && !adjoint && m <= 2 && I.ray && I.aer) ? 1 : 0;
                if (I.aer) kinds[np++] = 1;
                if (m <= 2 && I.ray && !it.lowrank) kinds[np++] = 2;
                if (np == 0) kinds[np++] = 1;
                it.parts = kinds[0] | (np > 1 ? (kinds[1] << 8) : 0);
                it.K = np * Pp;
                const int tile_lev = gemm_bf ? kGemmLevels : kGemmLevelsF64;
                // the projection rows in the operator's padding when they fit
                const int pin = roundup(6 * n, 4);
                const bool pfit = pin + 4 <= Mp;
                it.pbase = (it.lowrank && pfit) ? pin : Mp;
                const int Mop = Mp + ((it.lowrank && !pfit) ? 64 : 0);   // operator rows
                it.Mp = Mop; it.tiles_m = Mop / 64; it.tiles_n = (L + tile_lev - 1) / tile_lev;
                it.lmax = I.lmax;
                int cls;
                if (op_from >= 0) {
                    const int src = c.hinst[op_from].item0 + m;
                    const DevItem &o = c.hitem[src];
                    it.off_basis = o.off_basis; it.off_tab = o.off_tab; it.off_A = o.off_A; it.off_lr = o.off_lr;
                    it.op_owner = 0;
                    cls = class_of_item[src];
                } else {
                    it.off_basis = basis; basis += 3LL * (I.lmax + 1) * (2 * n + 1);
                    it.off_tab = tab; tab += 6LL * (n + 1) * (2 * n + 1);
                    it.off_A = A; A += (long long)Mop * it.K * (gemm_bf ? 2 : 1);
                    it.off_lr = lr; lr += it.lowrank ? 3LL * Pp : 0;
                    it.op_owner = 1;
                    cls = (int)class_items.size();
                    class_items.emplace_back();
                }
                it.off_F = F; it.off_tot = F; F += (long long)L * Pp;
                it.off_bc = bc; bc += 12LL * n;
                it.off_cmp = cmp; cmp += mixed ? 2LL * Pp : 0;
                it.off_W = wlr; wlr += it.lowrank ? 4LL * L : 0;
                it.off_nodes = p->nodes_n;
                p->nodes_n += 6LL * n;
                // the mixed adjoint keeps every field as its pair: a bf16 field
                // alone at the higher modes moves the per-layer gradients by
                // about 5e-3 of their scale
                it.fsplit = ((mixed && m < kSplitModes) || adj_mixed) ? 1 : 0;
                it.off_Flo = flo; flo += it.fsplit ? (long long)L * Pp : 0;
                c.max_pack = std::max(c.max_pack, (long long)Mp * it.K);
                const int item_local = (int)c.hitem.size();
                class_of_item.push_back(cls);
                class_items[cls].push_back(item_local);
                // direction-pure column tiles: upward half, downward half, padding
                const int bounds[4] = {0, 3 * n, 6 * n, Pp};
                for (int h = 0; h < 3; ++h)
                    for (int c0 = bounds[h]; c0 < bounds[h + 1]; c0 += kColTile)
                        c.hcols.push_back(ColTile{item_local, c0, std::min(c0 + kColTile, bounds[h + 1]), 0});
                c.hitem.push_back(it);
            }
            if (I.ipss) c.hjobs.push_back(IpssJob{ii - c.i0, 0, I.vza, I.raa});
            if (I.n_ang > 1 && op_from < 0) {
                if (c.hvk_first.empty()) c.hvk_first.push_back(0);
                c.hvk_owner.push_back(ii - c.i0);
                c.hvk_first.push_back(c.hvk_first.back() + (n + 1) * (2 * n + 1));
            }
            c.hinst.push_back(I);
        }
        // the GEMM tiles of one class stay on one XCD, ordered (row tile, item,
        // level tile) so consecutive tiles reuse the class's operator rows in L2;
        // single-part unsplit items first, then single-part split ones, then
        // two-part ones (one kernel instance each)
        std::vector<std::vector<GemmTile>> gemm_of_class[3];
        for (size_t k = 0; k < class_items.size(); ++k) {
            const DevItem &o = c.hitem[class_items[k][0]];
            std::vector<GemmTile> gt;
            for (int tm = 0; tm < o.tiles_m; ++tm)
                for (int il : class_items[k])
                    for (int tn = 0; tn < c.hitem[il].tiles_n; ++tn) gt.push_back(GemmTile{il, tm, tn, 0});
            gemm_of_class[o.K > o.P ? 2 : (o.fsplit ? 1 : 0)].push_back(std::move(gt));
        }
        if (mixed) {
            // the bf16 form's tile list for every shape
            for (int sh = 0; sh < kGemmShapeCount; ++sh) {
                const int TM = kGemmShapes[sh].tm, TN = kGemmShapes[sh].tn;
                std::vector<std::vector<GemmTile>> by_class[3];
                for (size_t k = 0; k < class_items.size(); ++k) {
                    const DevItem &o = c.hitem[class_items[k][0]];
                    std::vector<GemmTile> gt;
                    for (int tm = 0; tm < (o.Mp + TM - 1) / TM; ++tm)
                        for (int il : class_items[k])
                            for (int tn = 0; tn < (c.hitem[il].L + TN - 1) / TN; ++tn)
                                gt.push_back(GemmTile{il, tm, tn, 0});
                    by_class[o.K > o.P ? 2 : (o.fsplit ? 1 : 0)].push_back(std::move(gt));
                }
                std::vector<GemmTile> &lst = c.htiles_s[sh];
                for (int k = 0; k < 3; ++k) {
                    c.cls_off[sh][k] = (int)lst.size();
                    const std::vector<GemmTile> tk = xcd_tile_order(by_class[k]);
                    lst.insert(lst.end(), tk.begin(), tk.end());
                }
                c.cls_off[sh][3] = (int)lst.size();
            }
        }
        c.htiles = xcd_tile_order(gemm_of_class[0]);
        c.ntiles1 = (int)c.htiles.size();
        const std::vector<GemmTile> t1 = xcd_tile_order(gemm_of_class[1]);
        c.htiles.insert(c.htiles.end(), t1.begin(), t1.end());
        c.ntiles2 = (int)c.htiles.size();
        const std::vector<GemmTile> t2 = xcd_tile_order(gemm_of_class[2]);
        c.htiles.insert(c.htiles.end(), t2.begin(), t2.end());
        c.n_jobs = (int)c.hjobs.size();
        c.lev = lev; c.T = T; c.R = R; c.basis = basis; c.tab = tab; c.A = A; c.F = F; c.bc = bc;
        c.cmp = cmp; c.flo = flo; c.lr = lr; c.wlr = wlr;
    }
    // ---- device allocations
    long long mx_lev = 0, mx_T = 0, mx_R = 0, mx_basis = 0, mx_tab = 0, mx_A = 0, mx_F = 0, mx_bc = 0,
              mx_cmp = 0, mx_flo = 0, mx_lr = 0, mx_wlr = 0;
    for (const Chunk &c : p->chunks) {
        mx_lev = std::max(mx_lev, c.lev); mx_T = std::max(mx_T, c.T); mx_R = std::max(mx_R, c.R);
        mx_basis = std::max(mx_basis, c.basis); mx_tab = std::max(mx_tab, c.tab);
        mx_A = std::max(mx_A, c.A); mx_F = std::max(mx_F, c.F); mx_bc = std::max(mx_bc, c.bc);
        mx_cmp = std::max(mx_cmp, c.cmp);
        mx_flo = std::max(mx_flo, c.flo);
        mx_lr = std::max(mx_lr, c.lr);
        mx_wlr = std::max(mx_wlr, c.wlr);
    }
    if (getenv("OCRT_DEBUG_PLAN"))
        fprintf(stderr, "ocrt plan: %zu chunks items %d ring %lld greek %lld lev %lld T %lld R %lld basis %lld tab %lld "
                        "A %lld F %lld bc %lld\n", p->chunks.size(), p->n_items, ring_n, greek_n, mx_lev, mx_T,
                mx_R, mx_basis, mx_tab, mx_A, mx_F, mx_bc);
    if (getenv("OCRT_DEBUG_PLAN"))
        for (const Chunk &c : p->chunks)
            fprintf(stderr, "  chunk instances [%d, %d) items %d tiles %zu cols %zu max_pack %lld tab_groups %d\n", c.i0,
                    c.i1, c.it1 - c.it0, c.htiles.size(), c.hcols.size(), c.max_pack, c.tab_groups);
    int rc = 0;
    Pools &P = p->pools;
    P.lo_tau = (mixed && !adjoint) ? kReducedTau : 0.f;
    rc |= dmalloc(p, (void **)&p->ring, ring_n * 8);
    rc |= dmalloc(p, (void **)&p->greek, std::max(greek_n, 1LL) * 8);
    rc |= dmalloc(p, (void **)&p->layers, std::max(layers_n, 1LL) * 8);
    // gas absorption tables used by the instances, packed once
    std::vector<DevGas> hgas;
    std::vector<double> hgas_pool;
    for (const ocrt_gas *g : gas_used) hgas.push_back(pack_gas(*g, hgas_pool));
    rc |= dmalloc(p, (void **)&p->gas_tab, std::max<size_t>(hgas.size(), 1) * sizeof(DevGas));
    rc |= dmalloc(p, (void **)&p->gas_pool, std::max<size_t>(hgas_pool.size(), 1) * 8);
    // cos / sin (m phi) for m = 0 .. max(mmax, 1): the m = 1 row also gives the
    // azimuths of the Cox-Munk quadrature
    rc |= dmalloc(p, (void **)&p->cm_w, (size_t)(std::max(p->mmax_all, 1) + 1) * 2 * kNPhi * 8);
    if (mixed) rc |= dmalloc(p, (void **)&p->cm_wf, (size_t)(std::max(p->mmax_all, 1) + 1) * 2 * kNPhi * 4);
    bool any_vk = false;
    for (const Chunk &c : p->chunks) any_vk |= !c.hvk_owner.empty();
    if (any_vk) rc |= dmalloc(p, (void **)&p->vk_w, (size_t)kVkHalf * (p->mmax_all + 1) * 8);
    if (rtop_n > 0) rc |= dmalloc(p, (void **)&p->rtop, (size_t)rtop_n * 8);
    rc |= dmalloc(p, (void **)&p->state, (size_t)p->n_items * sizeof(ItemState));
    rc |= dmalloc(p, (void **)&p->iout, (size_t)p->n_items * sizeof(ItemOut));
    rc |= dmalloc(p, (void **)&p->ipss, (size_t)n_inst * sizeof(InstOut));
    rc |= dmalloc(p, (void **)&p->res, (size_t)n_inst * sizeof(ocrt_result_t));
    rc |= dmalloc(p, (void **)&p->kappa, (size_t)n_inst * 8);
    rc |= dmalloc(p, (void **)&p->taudir, (size_t)n_inst * 8);
    rc |= dmalloc(p, (void **)&P.lev, mx_lev * 8);
    rc |= dmalloc(p, (void **)&P.T, mx_T * 8);
    rc |= dmalloc(p, (void **)&P.Tf, mx_T * 4);
    rc |= dmalloc(p, (void **)&P.idt, mx_lev * 4);
    rc |= dmalloc(p, (void **)&P.R, mx_R * 8);
    // the mixed adjoint reads the fp64 rough kernels (with their derivatives)
    if (mixed && !adjoint) rc |= dmalloc(p, (void **)&P.Rf, std::max(mx_R, 1LL) * 4);
    // padded: the phase tables read the rows of a full group past the last row
    rc |= dmalloc(p, (void **)&P.basis, (mx_basis + kTabRows) * 8);
    rc |= dmalloc(p, (void **)&P.tab, mx_tab * 8);
    rc |= dmalloc(p, (void **)&P.A, mx_A * ae);
    p->mx_F = mx_F;
    rc |= dmalloc(p, (void **)&P.F, mx_F * fe);
    rc |= dmalloc(p, (void **)&P.J, mx_F * je);
    if (!mixed) rc |= dmalloc(p, (void **)&P.tot, mx_F * 8);
    if (mixed) rc |= dmalloc(p, (void **)&P.comp, std::max(mx_cmp, 1LL) * 4);
    if (mixed) rc |= dmalloc(p, (void **)&P.trow, std::max(mx_cmp, 1LL) * 8);
    if (mixed) rc |= dmalloc(p, (void **)&P.Flo, std::max(mx_flo, 1LL) * 2);
    if (mixed) rc |= dmalloc(p, (void **)&P.Jf, std::max(mx_flo, 1LL) * 4);
    if (mixed) rc |= dmalloc(p, (void **)&P.lr, std::max(mx_lr, 1LL) * 4);
    if (mixed) rc |= dmalloc(p, (void **)&P.W, std::max(mx_wlr, 1LL) * 4);
    rc |= dmalloc(p, (void **)&P.bc, mx_bc * 8);
    if (adjoint) {
        int maxit = 1;
        for (const HostInst &h : p->inst) maxit = std::max(maxit, h.d.max_orders);
        p->ckpt.assign((maxit - 1) / p->ckpt_every + 1, nullptr);
        p->seg.assign(p->ckpt_every - 1, nullptr);
        // fields: fp64, or those of the mixed plans as two bf16 planes, the
        // field (mx_F entries) then the low halves of the pairs (mx_flo <=
        // mx_F); lambda and the operator parts fp64 / fp32
        const size_t fb = mixed ? 4 : 8;
        if (mixed) rc |= dmalloc(p, &p->Jadj, mx_F * 4);
        p->rpart0.assign(p->ckpt_every - 1, nullptr);
        p->rpart1.assign(p->ckpt_every - 1, nullptr);
        for (void *&q : p->rpart0) rc |= dmalloc(p, &q, mx_F * fb);
        for (void *&q : p->rpart1) rc |= dmalloc(p, &q, mx_F * fb);
        if (mixed) rc |= dmalloc(p, (void **)&P.sb_hi, mx_F * 2);
        if (mixed) rc |= dmalloc(p, (void **)&P.sb_lo, mx_F * 2);
        for (void *&q : p->ckpt) rc |= dmalloc(p, &q, mx_F * fb);
        for (void *&q : p->seg) rc |= dmalloc(p, &q, mx_F * fb);
        // the mixed sweeps write the live columns only: the pad columns of
        // every checkpoint and segment buffer stay the zeros set here (the
        // operator products multiply them by zero entries)
        if (mixed && !rc)
            for (void *q : p->ckpt) rc |= hipMemset(q, 0, mx_F * fb) != hipSuccess;
        if (mixed && !rc)
            for (void *q : p->seg) rc |= hipMemset(q, 0, mx_F * fb) != hipSuccess;
        // and of the adjoint source's pair planes (the sweeps write the live
        // columns only)
        if (mixed && !rc)
            rc |= (hipMemset(P.sb_hi, 0, mx_F * 2) != hipSuccess) || (hipMemset(P.sb_lo, 0, mx_F * 2) != hipSuccess);
        // and of the working field, which order 1 (written to its checkpoint)
        // no longer clears
        if (mixed && !rc)
            rc |= (hipMemset(P.F, 0, mx_F * fe) != hipSuccess) ||
                  (hipMemset(P.Flo, 0, std::max(mx_flo, 1LL) * 2) != hipSuccess);
        rc |= dmalloc(p, &p->At, mx_A * ae);
        rc |= dmalloc(p, (void **)&P.lam, mx_F * fb);
        rc |= dmalloc(p, (void **)&P.sc, mx_T * 6 * 8);
        if (mixed) rc |= dmalloc(p, (void **)&P.scf, mx_T * 6 * 4);
        rc |= dmalloc(p, (void **)&P.J2, mx_F * fb);
        rc |= dmalloc(p, (void **)&P.dR, std::max(mx_R, 1LL) * 8);
        rc |= dmalloc(p, (void **)&p->grad, p->grad_n * 8);
        rc |= dmalloc(p, (void **)&p->gout, (size_t)n_inst * 3 * 8);
        P.grad = p->grad;
        P.gout = p->gout;
    }
    if (rc) { ocrt_plan_destroy(p); return OCRT_ERR_ALLOC; }
    P.ring = p->ring;
    P.greek = p->greek;
    P.layers = p->layers;
    P.cm_w = p->cm_w;
    P.cm_wf = p->cm_wf;
    P.rtop = p->rtop;
    P.nodes = nullptr;
    P.gas_tab = p->gas_tab;
    P.gas_pool = p->gas_pool;
    if ((!hgas.empty() && hipMemcpy(p->gas_tab, hgas.data(), hgas.size() * sizeof(DevGas), hipMemcpyHostToDevice) !=
                              hipSuccess) ||
        (!hgas_pool.size() ? false
                           : hipMemcpy(p->gas_pool, hgas_pool.data(), hgas_pool.size() * 8, hipMemcpyHostToDevice) !=
                                 hipSuccess)) {
        ocrt_plan_destroy(p);
        return OCRT_ERR_DEVICE;
    }
    // ---- uploads of inputs and descriptors
    std::vector<double> hring(ring_n), hgreek(std::max(greek_n, 1LL)), hlayers(std::max(layers_n, 1LL));
    for (int ii = 0; ii < n_inst; ++ii) {
        const HostInst &h = p->inst[ii];
        const int n = h.d.n;
        for (int o = 0; o < n; ++o) { hring[h.d.off_ring + o] = h.mu[o]; hring[h.d.off_ring + n + o] = h.w[o]; }
        if (h.d.aer) {
            const int q = h.d.La + 1;
            for (int e = 0; e < 4 * q; ++e) hgreek[h.d.off_greek + e] = in[ii].greek[e];
        }
        for (int e = 0; e < h.d.n_ang; ++e) {
            hgreek[h.d.off_ang + e] = in[ii].theta_phase[e];
            hgreek[h.d.off_ang + h.d.n_ang + e] = in[ii].p11_phase[e];
        }
        for (long long e = 0; e < 3LL * (h.d.bsrc_mmax + 1) * n; ++e)
            hgreek[h.d.off_bsrc + e] = in[ii].bottom_source[e];
        for (long long e = 0; e < 3LL * (h.d.tsrc_mmax + 1) * n; ++e)
            hgreek[h.d.off_tsrc + e] = in[ii].top_source[e];
        if (h.d.layered)
            for (long long e = 0; e < 4LL * h.d.nt; ++e) hlayers[h.d.off_layers + e] = in[ii].layers[e];
    }
    if (hipMemcpy(p->ring, hring.data(), ring_n * 8, hipMemcpyHostToDevice) != hipSuccess ||
        (greek_n > 0 && hipMemcpy(p->greek, hgreek.data(), greek_n * 8, hipMemcpyHostToDevice) != hipSuccess) ||
        (layers_n > 0 && hipMemcpy(p->layers, hlayers.data(), layers_n * 8, hipMemcpyHostToDevice) != hipSuccess)) {
        ocrt_plan_destroy(p);
        return OCRT_ERR_DEVICE;
    }
    // water-side Cox-Munk kernels of the rough in-water columns, every mode in
    // one launch per owner, on the ring nodes just uploaded (128 azimuths,
    // rt_water_rt.c:5605)
    for (int i : p->rtop_owner) {
        const DevInst &I = p->inst[i].d;
        const double *d_mu = p->ring + I.off_ring;
        const int rcw = interface_fourier_device(OCRT_IFACE_R_WW, d_mu, I.n, d_mu, I.n, I.mmax, 128, I.wind,
                                                 I.sigma_type, I.n_water, I.q_conv, p->rtop + I.off_rtop, 0);
        if (rcw != OCRT_OK) { ocrt_plan_destroy(p); return rcw; }
    }
    // chunk descriptors: one device buffer, sub-allocated 256-byte aligned
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    size_t desc_bytes = 0;
    for (const Chunk &c : p->chunks)
        for (int sh = 0; sh < kGemmShapeCount; ++sh)
            desc_bytes += al(std::max<size_t>(c.htiles_s[sh].size(), 1) * sizeof(GemmTile));
    for (const Chunk &c : p->chunks)
        desc_bytes += al(c.hinst.size() * sizeof(DevInst)) + al(c.hitem.size() * sizeof(DevItem)) +
                      al(std::max<size_t>(c.htiles.size(), 1) * sizeof(GemmTile)) +
                      al(std::max<size_t>(c.hcols.size(), 1) * sizeof(ColTile)) +
                      al(std::max<size_t>(c.hjobs.size(), 1) * sizeof(IpssJob)) +
                      al((size_t)(c.i1 - c.i0) * sizeof(int)) + al(std::max<size_t>(c.hcm_owner.size(), 1) * 4) +
                      al(std::max<size_t>(c.hcm_first.size(), 1) * 4) * 3 +
                      al(std::max<size_t>(c.hvk_owner.size(), 1) * 4) + al(std::max<size_t>(c.hvk_first.size(), 1) * 4);
    std::vector<unsigned char> hdesc(desc_bytes, 0);
    unsigned char *ddesc = nullptr;
    if (dmalloc(p, (void **)&ddesc, desc_bytes)) { ocrt_plan_destroy(p); return OCRT_ERR_ALLOC; }
    size_t off = 0;
    auto put = [&](const void *src, size_t bytes, size_t room) -> void * {
        if (bytes) memcpy(hdesc.data() + off, src, bytes);
        void *d = ddesc + off;
        off += al(room);
        return d;
    };
    for (Chunk &c : p->chunks) {
        std::vector<int> job_of(c.i1 - c.i0, -1);
        for (size_t j = 0; j < c.hjobs.size(); ++j) job_of[c.hjobs[j].inst] = (int)j;
        c.dinst = (DevInst *)put(c.hinst.data(), c.hinst.size() * sizeof(DevInst), c.hinst.size() * sizeof(DevInst));
        c.ditem = (DevItem *)put(c.hitem.data(), c.hitem.size() * sizeof(DevItem), c.hitem.size() * sizeof(DevItem));
        c.dtiles = (GemmTile *)put(c.htiles.data(), c.htiles.size() * sizeof(GemmTile),
                                   std::max<size_t>(c.htiles.size(), 1) * sizeof(GemmTile));
        for (int sh = 0; sh < kGemmShapeCount; ++sh)
            c.dtiles_s[sh] = (GemmTile *)put(c.htiles_s[sh].data(), c.htiles_s[sh].size() * sizeof(GemmTile),
                                             std::max<size_t>(c.htiles_s[sh].size(), 1) * sizeof(GemmTile));
        c.dcols = (ColTile *)put(c.hcols.data(), c.hcols.size() * sizeof(ColTile),
                                 std::max<size_t>(c.hcols.size(), 1) * sizeof(ColTile));
        c.djobs = (IpssJob *)put(c.hjobs.data(), c.hjobs.size() * sizeof(IpssJob),
                                 std::max<size_t>(c.hjobs.size(), 1) * sizeof(IpssJob));
        c.djob_of = (int *)put(job_of.data(), job_of.size() * sizeof(int), job_of.size() * sizeof(int));
        c.dcm_owner = (int *)put(c.hcm_owner.data(), c.hcm_owner.size() * 4, std::max<size_t>(c.hcm_owner.size(), 1) * 4);
        c.dcm_first = (int *)put(c.hcm_first.data(), c.hcm_first.size() * 4, std::max<size_t>(c.hcm_first.size(), 1) * 4);
        c.dcm_sfirst = (int *)put(c.hcm_sfirst.data(), c.hcm_sfirst.size() * 4,
                                  std::max<size_t>(c.hcm_sfirst.size(), 1) * 4);
        c.dcm_dfirst = (int *)put(c.hcm_dfirst.data(), c.hcm_dfirst.size() * 4,
                                  std::max<size_t>(c.hcm_dfirst.size(), 1) * 4);
        c.dvk_owner = (int *)put(c.hvk_owner.data(), c.hvk_owner.size() * 4, std::max<size_t>(c.hvk_owner.size(), 1) * 4);
        c.dvk_first = (int *)put(c.hvk_first.data(), c.hvk_first.size() * 4, std::max<size_t>(c.hvk_first.size(), 1) * 4);
    }
    if (hipMemcpy(ddesc, hdesc.data(), desc_bytes, hipMemcpyHostToDevice) != hipSuccess) {
        ocrt_plan_destroy(p);
        return OCRT_ERR_DEVICE;
    }
    const int cm_rows = std::max(p->mmax_all, 1);
    hipLaunchKernelGGL(fourier_weights, dim3((kNPhi + 255) / 256, cm_rows + 1), dim3(256), 0, 0, p->cm_w, p->cm_wf,
                       cm_rows);
    if (p->vk_w)
        hipLaunchKernelGGL(value_weights, dim3((kVkHalf * (p->mmax_all + 1) + 255) / 256), dim3(256), 0, 0, p->vk_w,
                           p->mmax_all + 1);
    if (hipDeviceSynchronize() != hipSuccess) { ocrt_plan_destroy(p); return OCRT_ERR_DEVICE; }
    *out = p;
    return OCRT_OK;
}

extern "C" int ocrt_plan_create(const ocrt_instance_t *in, int n_inst, int precision, int one_per_launch,
                                ocrt_plan **out)
{
    return ocrt_plan_create_ex(in, n_inst, precision, one_per_launch ? OCRT_PLAN_ONE_PER_LAUNCH : 0, out);
}

static void drop_graph(ocrt_plan *p);

// Launch shapes of the mixed-form sweep the autotuner chooses from: levels
// per block U and wavefronts per workgroup W (all give the same bits).
// Launch shapes of the mixed-form sweep: levels per load block U and
// wavefronts per workgroup W.
struct SweepShape { int U, W; };
constexpr SweepShape kSweepShapes[] = {{8, 4}, {4, 4}, {16, 4}, {8, 2}, {4, 2}, {16, 2}};
constexpr int kSweepShapeCount = sizeof(kSweepShapes) / sizeof(kSweepShapes[0]);

// The bf16 scattering integral of one order: each operator class with its
// tile shape; while the autotuner times the classes, an event pair brackets
// each class's launch (kinds 2 + class).
template <int NP, bool SP>
static void launch_gemm_class(int sh, unsigned n, const GemmTile *tiles, const Chunk &c, hipStream_t s,
                              const Pools &P, const ItemState *st, int order)
{
    switch (sh) {
    case 0: hipLaunchKernelGGL((gemm_bf16<NP, SP, 64, 128>), dim3(n), dim3(256), 0, s, c.dinst, c.ditem, P, tiles, st, order); break;
    case 1: hipLaunchKernelGGL((gemm_bf16<NP, SP, 64, 64>), dim3(n), dim3(256), 0, s, c.dinst, c.ditem, P, tiles, st, order); break;
    case 2: hipLaunchKernelGGL((gemm_bf16<NP, SP, 128, 64>), dim3(n), dim3(256), 0, s, c.dinst, c.ditem, P, tiles, st, order); break;
    default: hipLaunchKernelGGL((gemm_bf16<NP, SP, 128, 128>), dim3(n), dim3(256), 0, s, c.dinst, c.ditem, P, tiles, st, order); break;
    }
}

// order 0: the forward order loop (active items); order > 0: recomputation of
// that order (the adjoint), every item whose forward solve reached it
static void launch_gemm_mixed(ocrt_plan *p, const Chunk &c, hipStream_t s, const Pools &P, const ItemState *st,
                              int order = 0)
{
    for (int k = 0; k < 3; ++k) {
        const int sh = p->gemm_shape[k];
        const unsigned n = (unsigned)(c.cls_off[sh][k + 1] - c.cls_off[sh][k]);
        if (n == 0) continue;
        const GemmTile *tiles = c.dtiles_s[sh] + c.cls_off[sh][k];
        const EvPair e = p->tune_classes ? ev_pair(p, 2 + k) : EvPair{{nullptr, nullptr}, false};
        if (e.ok) (void)hipEventRecord(e.e[0], s);
        if (k == 0) launch_gemm_class<1, false>(sh, n, tiles, c, s, P, st, order);
        else if (k == 1) launch_gemm_class<1, true>(sh, n, tiles, c, s, P, st, order);
        else launch_gemm_class<2, true>(sh, n, tiles, c, s, P, st, order);
        if (e.ok) (void)hipEventRecord(e.e[1], s);
    }
}

// Point Q's field at a checkpoint or segment buffer of an adjoint plan: fp64,
// or (mixed plans) two bf16 planes: the field (mx_F entries), then the low
// halves of the items that keep the pair.
static void set_field(const ocrt_plan *p, Pools &Q, void *buf)
{
    Q.F = buf;
    if (p->adj_mixed) Q.Flo = reinterpret_cast<__hip_bfloat16 *>(buf) + p->mx_F;
}

// Rebuilt order of a mixed adjoint plan (gemm_rebuild) over the chunk's
// 64 x 128 tile list: the sweep's source and the order's operator parts.
static void launch_rebuild_gemm(const Chunk &c, hipStream_t s, const Pools &P, const ItemState *st, int order)
{
    const GemmTile *t = c.dtiles_s[0];
    const unsigned n1 = (unsigned)(c.cls_off[0][2] - c.cls_off[0][0]), n2 = (unsigned)(c.cls_off[0][3] - c.cls_off[0][2]);
    if (n1 > 0)
        hipLaunchKernelGGL(gemm_rebuild<1>, dim3(n1), dim3(256), 0, s, c.dinst, c.ditem, P, t + c.cls_off[0][0], st,
                           order);
    if (n2 > 0)
        hipLaunchKernelGGL(gemm_rebuild<2>, dim3(n2), dim3(256), 0, s, c.dinst, c.ditem, P, t + c.cls_off[0][2], st,
                           order);
}

// Operator products of a mixed adjoint plan over the chunk's 64 x 128 tile
// list (single-part classes 0 and 1, two-part class 2):
// MODE kGemmParts on the field in P.F / P.Flo, kGemmMix (A^T) on the adjoint
// source's bf16 pair in P.F / P.Flo.
static void launch_adj_gemm(int mode, const Chunk &c, hipStream_t s, const Pools &P, const ItemState *st, int order)
{
    const GemmTile *t = c.dtiles_s[0];
    const unsigned n1 = (unsigned)(c.cls_off[0][2] - c.cls_off[0][0]), n2 = (unsigned)(c.cls_off[0][3] - c.cls_off[0][2]);
    if (n1 > 0) {
        if (mode == kGemmParts)
            hipLaunchKernelGGL((gemm_adj<1, kGemmParts>), dim3(n1), dim3(256), 0, s, c.dinst, c.ditem, P,
                               t + c.cls_off[0][0], st, order);
        else
            hipLaunchKernelGGL((gemm_adj<1, kGemmMix>), dim3(n1), dim3(256), 0, s, c.dinst, c.ditem, P,
                               t + c.cls_off[0][0], st, order);
    }
    if (n2 > 0) {
        if (mode == kGemmParts)
            hipLaunchKernelGGL((gemm_adj<2, kGemmParts>), dim3(n2), dim3(256), 0, s, c.dinst, c.ditem, P,
                               t + c.cls_off[0][2], st, order);
        else
            hipLaunchKernelGGL((gemm_adj<2, kGemmMix>), dim3(n2), dim3(256), 0, s, c.dinst, c.ditem, P,
                               t + c.cls_off[0][2], st, order);
    }
}

static void launch_sweep_mixed(int v, int ncols, hipStream_t s, const DevInst *di, const DevItem *dt, Pools P,
                               ItemState *st, const ColTile *tiles, int order = 0)
{
    const int W = kSweepShapes[v].W;
    const dim3 grid((unsigned)((ncols + W - 1) / W)), block((unsigned)(64 * W));
#define OCRT_SWEEP_CASE(I, U_, W_)                                                                              \
    case I:                                                                                                      \
        hipLaunchKernelGGL((sweep<__hip_bfloat16, _Float16, true, U_, W_>), grid, block, 0, s, di, dt, P, st, tiles, \
                           ncols, order, (const double *)P.lev, (const float *)P.idt, (const float *)P.W);     \
        break;
    switch (v) {
        OCRT_SWEEP_CASE(0, 8, 4)
        OCRT_SWEEP_CASE(1, 4, 4)
        OCRT_SWEEP_CASE(2, 16, 4)
        OCRT_SWEEP_CASE(3, 8, 2)
        OCRT_SWEEP_CASE(4, 4, 2)
        OCRT_SWEEP_CASE(5, 16, 2)
        default: break;
    }
#undef OCRT_SWEEP_CASE
}

static int run_chunk(ocrt_plan *p, Chunk &c, hipStream_t s)
{
    const bool mixed = p->prec == OCRT_PREC_MIXED;
    Pools P = p->pools;
    const int ni = c.i1 - c.i0, nit = c.it1 - c.it0;
    ItemState *st = p->state + c.it0;
    ItemOut *io = p->iout + c.it0;
    hipLaunchKernelGGL(medium_levels, dim3((c.max_L + 127) / 128, ni), dim3(128), 0, s, c.dinst, P);
    hipLaunchKernelGGL(medium_fractions, dim3((c.max_L + 127) / 128, ni), dim3(128), 0, s, c.dinst, P);
    if (c.any_gas) hipLaunchKernelGGL(medium_gas, dim3(ni), dim3(256), 0, s, c.dinst, P);
    hipLaunchKernelGGL(medium_transmit, dim3((c.max_L + 127) / 128, ni), dim3(128), 0, s, c.dinst, P);
    if (p->adjoint)
        hipLaunchKernelGGL(adj_step_table, dim3((c.max_L + 127) / 128, ni), dim3(128), 0, s, c.dinst, P);
    if (!c.hcm_owner.empty()) {
        // one workgroup per pair, in launches of at most 2^22 workgroups (far
        // below 2^31 work-items each); mixed plans take the diffuse pairs in
        // fp32 (only their fp32 copy is read) and the solar ones in fp64
        const int no = (int)c.hcm_owner.size();
        const dim3 blk(kCoxMunkThreads);
        auto pairs = [&](int total, auto launch) {
            for (int base = 0; base < total; base += 1 << 22) launch(dim3((unsigned)std::min(total - base, 1 << 22)), base);
        };
        if (p->adj_mixed) {
            // with the derivatives: the solar pairs in fp64, the diffuse ones in fp32
            pairs(c.hcm_sfirst.back(), [&](dim3 grid, int base) {
                hipLaunchKernelGGL((coxmunk_fourier<true, double, 1>), grid, blk, 0, s, c.dinst, P, c.dcm_owner,
                                   c.dcm_sfirst, no, base);
            });
            pairs(c.hcm_dfirst.back(), [&](dim3 grid, int base) {
                hipLaunchKernelGGL((coxmunk_fourier<true, float, 2>), grid, blk, 0, s, c.dinst, P, c.dcm_owner,
                                   c.dcm_dfirst, no, base);
            });
        } else if (p->adjoint) {
            pairs(c.hcm_first.back(), [&](dim3 grid, int base) {
                hipLaunchKernelGGL((coxmunk_fourier<true, double, 0>), grid, blk, 0, s, c.dinst, P, c.dcm_owner,
                                   c.dcm_first, no, base);
            });
        } else if (mixed) {
            pairs(c.hcm_sfirst.back(), [&](dim3 grid, int base) {
                hipLaunchKernelGGL((coxmunk_fourier<false, double, 1>), grid, blk, 0, s, c.dinst, P, c.dcm_owner,
                                   c.dcm_sfirst, no, base);
            });
            pairs(c.hcm_dfirst.back(), [&](dim3 grid, int base) {
                hipLaunchKernelGGL((coxmunk_fourier<false, float, 2>), grid, blk, 0, s, c.dinst, P, c.dcm_owner,
                                   c.dcm_dfirst, no, base);
            });
        } else {
            pairs(c.hcm_first.back(), [&](dim3 grid, int base) {
                hipLaunchKernelGGL((coxmunk_fourier<false, double, 0>), grid, blk, 0, s, c.dinst, P, c.dcm_owner,
                                   c.dcm_first, no, base);
            });
        }
    }
    hipLaunchKernelGGL(angular_basis, dim3(nit), dim3(128), 0, s, c.dinst, c.ditem, P);
    hipLaunchKernelGGL(phase_tables, dim3(xcd_item_grid(nit, c.tab_groups)), dim3(64), 0, s, c.dinst, c.ditem, P,
                       nit, c.tab_groups);
    if (!c.hvk_owner.empty()) {
        // value-kernel instances replace their I -> I table, one workgroup per
        // entry, in launches of at most 2^22 workgroups
        const int total = c.hvk_first.back(), no = (int)c.hvk_owner.size();
        for (int base = 0; base < total; base += 1 << 22)
            hipLaunchKernelGGL(value_phase, dim3((unsigned)std::min(total - base, 1 << 22)), dim3(256), 0, s,
                               c.dinst, c.ditem, P, c.dvk_owner, c.dvk_first, no, base, p->vk_w, p->mmax_all + 1);
    }
    // a few blocks per item striding over its operator: an item's table reads
    // stay in one L2 and the launch stays far below 2^31 work-items
    const int pack_blocks = (int)std::min<long long>(16, (c.max_pack + 255) / 256);
    const dim3 pack_grid(xcd_item_grid(nit, pack_blocks));
    if (mixed) hipLaunchKernelGGL(lowrank_factors, dim3(nit), dim3(256), 0, s, c.dinst, c.ditem, P);
    if (mixed)
        hipLaunchKernelGGL((pack_operator<__hip_bfloat16, false>), pack_grid, dim3(256), 0, s, c.dinst, c.ditem, P,
                           nit, pack_blocks);
    else
        hipLaunchKernelGGL((pack_operator<double, false>), pack_grid, dim3(256), 0, s, c.dinst, c.ditem, P, nit,
                           pack_blocks);
    if (p->adjoint) {
        Pools Pt = P;
        Pt.A = p->At;
        if (p->adj_mixed)
            hipLaunchKernelGGL((pack_operator<__hip_bfloat16, true>), pack_grid, dim3(256), 0, s, c.dinst, c.ditem,
                               Pt, nit, pack_blocks);
        else
            hipLaunchKernelGGL((pack_operator<double, true>), pack_grid, dim3(256), 0, s, c.dinst, c.ditem, Pt, nit,
                               pack_blocks);
    }
    // adjoint plans keep u_1, u_{1+e}, u_{1+2e}, ... of every item: the sweep
    // of such an order writes straight into its checkpoint and the next order
    // reads it there (no copies); the other orders use the working buffer.
    const int every = p->ckpt_every;
    auto fields_of = [&](int order) -> Pools {
        Pools Q = P;
        if (p->adjoint && (order - 1) % every == 0) set_field(p, Q, p->ckpt[(order - 1) / every]);
        return Q;
    };
    // per-column kernels: one 64-column tile per wavefront, four per workgroup
    const int ncols = (int)c.hcols.size();
    const dim3 cols((unsigned)((ncols + 3) / 4));
    if (mixed)
        hipLaunchKernelGGL((order_one<__hip_bfloat16, true>), cols, dim3(256), 0, s, c.dinst, c.ditem, fields_of(1),
                           st, c.dcols, ncols, (const double *)P.lev, (const double *)P.T);
    else
        hipLaunchKernelGGL((order_one<double, false>), cols, dim3(256), 0, s, c.dinst, c.ditem, fields_of(1), st,
                           c.dcols, ncols, (const double *)P.lev, (const double *)P.T);
    const unsigned ntiles = (unsigned)c.htiles.size();
    for (int it = 2; it <= c.maxit; ++it) {
        const EvPair eg = p->profile ? ev_pair(p, 0) : EvPair{{nullptr, nullptr}, false};
        const EvPair es = p->profile ? ev_pair(p, 1) : EvPair{{nullptr, nullptr}, false};
        const Pools Pg = fields_of(it - 1), Ps = fields_of(it);
        if (mixed) {
            hipLaunchKernelGGL(order_status<__hip_bfloat16>, dim3(nit), dim3(64), 0, s, c.dinst, c.ditem, P, st, it);
            if (eg.ok) (void)hipEventRecord(eg.e[0], s);
            launch_gemm_mixed(p, c, s, Pg, st);
            if (eg.ok) (void)hipEventRecord(eg.e[1], s);
            if (es.ok) (void)hipEventRecord(es.e[0], s);
            launch_sweep_mixed(p->sweep_variant, ncols, s, c.dinst, c.ditem, Ps, st, c.dcols);
            if (es.ok) (void)hipEventRecord(es.e[1], s);
        } else {
            hipLaunchKernelGGL(order_status<double>, dim3(nit), dim3(64), 0, s, c.dinst, c.ditem, P, st, it);
            if (eg.ok) (void)hipEventRecord(eg.e[0], s);
            hipLaunchKernelGGL(gemm_f64<kGemmMix>, dim3(ntiles), dim3(256), 0, s, c.dinst, c.ditem, Pg, c.dtiles,
                               st, 0);
            if (eg.ok) (void)hipEventRecord(eg.e[1], s);
            if (es.ok) (void)hipEventRecord(es.e[0], s);
            hipLaunchKernelGGL((sweep<double, double, false, 8, 4>), cols, dim3(256), 0, s, c.dinst, c.ditem, Ps, st,
                               c.dcols, ncols, 0, (const double *)Ps.lev, (const float *)Ps.idt, (const float *)Ps.W);
            if (es.ok) (void)hipEventRecord(es.e[1], s);
        }
        p->gemm_launches += 1.0;
    }
    if (mixed)
        hipLaunchKernelGGL(order_status<__hip_bfloat16>, dim3(nit), dim3(64), 0, s, c.dinst, c.ditem, P, st, c.maxit + 1);
    else
        hipLaunchKernelGGL(order_status<double>, dim3(nit), dim3(64), 0, s, c.dinst, c.ditem, P, st, c.maxit + 1);
    Pools Pe = P;
    Pe.nodes = p->nodes;   // allocated only for LUT and node-export plans
    hipLaunchKernelGGL(extract, dim3(nit), dim3(64), 0, s, c.dinst, c.ditem, Pe, st, io, mixed ? 1 : 0);
    if (c.n_jobs > 0)
        hipLaunchKernelGGL(ipss_kappa, dim3(c.n_jobs), dim3(256), 0, s, c.dinst, c.djobs, P,
                           p->kappa + c.i0, p->taudir + c.i0);
    if (p->lut && !c.hjobs_lut.empty())
        hipLaunchKernelGGL(ipss_kappa, dim3((unsigned)c.hjobs_lut.size()), dim3(256), 0, s, c.dinst, c.djobs_lut, P,
                           p->kappa_lut, p->taudir_lut);
    return hipGetLastError() == hipSuccess ? 0 : OCRT_ERR_DEVICE;
}

// Backward pass of one chunk after its forward pass.  With checkpoint
// interval e, segment s holds u_{es+1} (the checkpoint) and the recomputed
// u_{es+2..es+e}; the backward step of order o needs u_o and u_{o-1}, so
// segments are rebuilt from the top down and order e(s+1)+1 (whose u_o is the
// next checkpoint) is taken with segment s.  With e = 1 nothing is recomputed.
// fp64 plans: fp64 fields, adjoint and operator products (gemm_f64).  Mixed
// plans: the fields as their bf16 pairs, rebuilt by the mixed order loop's own
// kernels (gemm_bf16, the fp32 sweep); lambda, the adjoint source and the
// operator parts in fp32, the parts and A^T products on bf16 MFMA (gemm_adj),
// the adjoint sweeps in fp64 arithmetic; gradients accumulate in fp64.
template <typename FV, typename LT>
static void backward_steps(ocrt_plan *p, Chunk &c, hipStream_t s)
{
    constexpr bool MX = sizeof(LT) == 4;
    const Pools P0 = p->pools;
    const int ni = c.i1 - c.i0, nit = c.it1 - c.it0;
    ItemState *st = p->state + c.it0;
    const int ncols = (int)c.hcols.size();
    const dim3 cols((unsigned)((ncols + 3) / 4));
    const unsigned ntiles = (unsigned)c.htiles.size();
    const int maxit = c.maxit;
    hipLaunchKernelGGL(adj_zero_grad, dim3(4, ni), dim3(256), 0, s, c.dinst, P0);
    hipLaunchKernelGGL(fill<LT>, dim3((unsigned)std::min<long long>((c.F + 255) / 256, 65536)), dim3(256), 0, s,
                       reinterpret_cast<LT *>(P0.lam), c.F, (LT)0);
    hipLaunchKernelGGL(adj_seed<LT>, dim3(nit), dim3(64), 0, s, c.dinst, c.ditem, P0);
    auto slot = [&](int seg, int j) -> void * { return j == 0 ? p->ckpt[seg] : p->seg[j - 1]; };
    auto at = [&](void *buf) {
        Pools Q = P0;
        set_field(p, Q, buf);
        return Q;
    };
    // mixed plans: the first operator part (and the adjoint source written
    // over it) in a buffer of its own, fp32
    Pools Pw = P0;
    if (MX) Pw.J = p->Jadj;
    const int every = p->ckpt_every;
    for (int sg = (maxit - 1) / every; sg >= 0; --sg) {
        const int base = every * sg + 1;   // order of the checkpoint
        // rebuild u_{base+1..base+every-1}
        for (int j = 1; j < every && base + j <= maxit; ++j) {
            const int o = base + j;
            const Pools Pin = at(slot(sg, j - 1));
            hipLaunchKernelGGL(adj_bottom<FV>, dim3(nit), dim3(64), 0, s, c.dinst, c.ditem, Pin, st, o);
            hipLaunchKernelGGL(adj_surface_bc, dim3(nit), dim3(64), 0, s, c.dinst, c.ditem, Pin, st, o);
            const Pools Pout = at(slot(sg, j));
            if constexpr (MX) {
                // the source of the rebuilt sweep, and the operator parts the
                // backward step of this order reads
                Pools Pr = Pin;
                Pr.J = p->rpart0[j - 1];
                Pr.J2 = p->rpart1[j - 1];
                launch_rebuild_gemm(c, s, Pr, st, o);
                launch_sweep_mixed(p->sweep_variant, ncols, s, c.dinst, c.ditem, Pout, st, c.dcols, o);
            } else {
                Pools Pr = Pin;
                Pr.rp0 = static_cast<double *>(p->rpart0[j - 1]);
                Pr.rp1 = static_cast<double *>(p->rpart1[j - 1]);
                hipLaunchKernelGGL(gemm_f64<kGemmBoth>, dim3(ntiles), dim3(256), 0, s, c.dinst, c.ditem, Pr,
                                   c.dtiles, st, o);
                hipLaunchKernelGGL((sweep<double, double, false, 8, 4>), cols, dim3(256), 0, s, c.dinst, c.ditem,
                                   Pout, st, c.dcols, ncols, o, (const double *)Pout.lev, (const float *)Pout.idt,
                                   (const float *)Pout.W);
            }
        }
        // backward steps
        for (int o = std::min(base + every, maxit); o >= std::max(base + 1, 2); --o) {
            void *u_o = (o == base + every) ? p->ckpt[sg + 1] : slot(sg, o - base);
            void *u_prev = slot(sg, o - 1 - base);
            Pools Pg = at(u_prev);
            Pg.J = Pw.J;
            // the operator parts A u_{o-1}: kept by the rebuild of the orders
            // between checkpoints, else one product here
            const bool kept = o - base >= 1 && o - base < every;
            if (kept) {
                Pg.J = p->rpart0[o - base - 1];
                Pg.J2 = p->rpart1[o - base - 1];
            } else if constexpr (MX) {
                launch_adj_gemm(kGemmParts, c, s, Pg, st, o);
            } else {
                hipLaunchKernelGGL(gemm_f64<kGemmParts>, dim3(ntiles), dim3(256), 0, s, c.dinst, c.ditem, Pg,
                                   c.dtiles, st, o);
            }
            Pools Ps = at(u_o);
            Ps.J = Pg.J;
            Ps.J2 = Pg.J2;
            hipLaunchKernelGGL((sweep_adj<FV, LT>), cols, dim3(256), 0, s, c.dinst, c.ditem, Ps, st, c.dcols, ncols, o,
                               (const double *)Ps.lev);
            Pools Pt = P0;
            Pt.A = p->At;
            Pt.F = MX ? (void *)P0.sb_hi : Ps.J;   // fp64: the sweep wrote Sbar over the first part
            Pt.Flo = P0.sb_lo;
            Pt.J = P0.lam;
            if constexpr (MX)
                launch_adj_gemm(kGemmMix, c, s, Pt, st, o);
            else
                hipLaunchKernelGGL(gemm_f64<kGemmMix>, dim3(ntiles), dim3(256), 0, s, c.dinst, c.ditem, Pt, c.dtiles,
                                   st, o);
            hipLaunchKernelGGL((surface_adj<FV, LT>), dim3(nit), dim3(64), 0, s, c.dinst, c.ditem, Pg, st, o);
        }
    }
    hipLaunchKernelGGL((order_one_adj<FV, LT>), cols, dim3(256), 0, s, c.dinst, c.ditem, at(p->ckpt[0]), c.dcols,
                       ncols, (const double *)p->pools.lev);
    hipLaunchKernelGGL(glint_adj, dim3((ni + 63) / 64), dim3(64), 0, s, c.dinst, ni, P0);
}

static int backward_chunk(ocrt_plan *p, Chunk &c, hipStream_t s)
{
    if (p->adj_mixed)
        backward_steps<FieldPair, float>(p, c, s);
    else
        backward_steps<FieldF64, double>(p, c, s);
    return hipGetLastError() == hipSuccess ? 0 : OCRT_ERR_DEVICE;
}

// kappa/taudir per chunk job -> InstOut per instance, then the instance results.
__global__ static void gather_ipss(const DevInst *inst, int ni, const double *kappa, const double *taudir,
                                   const int *job_of, InstOut *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ni) return;
    const int j = job_of[i];
    out[i].kappa = j >= 0 ? kappa[j] : 1.0;
    out[i].tau_dir_ipss = j >= 0 ? taudir[j] : -1.0;
}

static int forward_chunk(ocrt_plan *p, Chunk &c, hipStream_t s)
{
    if (run_chunk(p, c, s) != 0) return OCRT_ERR_DEVICE;
    const int ni = c.i1 - c.i0;
    hipLaunchKernelGGL(gather_ipss, dim3((ni + 127) / 128), dim3(128), 0, s, c.dinst, ni, p->kappa + c.i0,
                       p->taudir + c.i0, c.djob_of, p->ipss + c.i0);
    hipLaunchKernelGGL(finalize, dim3((ni + 127) / 128), dim3(128), 0, s, c.dinst, ni, p->iout + c.it0,
                       p->ipss + c.i0, p->pools, p->res + c.i0);
    if (p->lut) {
        Pools Pl = p->pools;
        Pl.nodes = p->nodes;
        hipLaunchKernelGGL(lut_eval, dim3(p->lut_nv, ni), dim3(128), 4 * (p->mmax_all + 1) * sizeof(double), s,
                           c.dinst, c.ditem, Pl, p->ipss + c.i0, c.djob_lut_of, p->kappa_lut, p->lut_nv, p->lut_vza,
                           p->lut_nr, p->lut_raa, p->lut_grid, p->lut_scal, p->lut_raw, p->lut_vpm,
                           p->mmax_all + 1);
    }
    return hipGetLastError() == hipSuccess ? OCRT_OK : OCRT_ERR_DEVICE;
}

extern "C" int ocrt_plan_set_lut(ocrt_plan *p, int n_vza, const double *vza_deg, int n_raa, const double *raa_deg)
{
    if (!p || n_vza < 1 || n_raa < 1 || !vza_deg || !raa_deg || p->lut) return OCRT_ERR_ARGS;
    drop_graph(p);   // the LUT epilogue changes the launch sequence
    for (int i = 0; i < n_vza; ++i)
        if (!(vza_deg[i] >= 0.0 && vza_deg[i] < 90.0)) return OCRT_ERR_ARGS;
    for (int i = 0; i < n_raa; ++i)
        if (!std::isfinite(raa_deg[i])) return OCRT_ERR_ARGS;
    const long long grid = (long long)n_vza * n_raa;
    int rc = 0;
    if (!p->nodes) rc |= dmalloc(p, (void **)&p->nodes, p->nodes_n * 8);
    rc |= dmalloc(p, (void **)&p->lut_vza, (size_t)n_vza * 8);
    rc |= dmalloc(p, (void **)&p->lut_raa, (size_t)n_raa * 8);
    rc |= dmalloc(p, (void **)&p->lut_grid, (size_t)p->n_inst * grid * 6 * 8);
    rc |= dmalloc(p, (void **)&p->lut_scal, (size_t)p->n_inst * 2 * 8);
    rc |= dmalloc(p, (void **)&p->lut_raw, (size_t)p->n_inst * grid * 3 * 8);
    rc |= dmalloc(p, (void **)&p->lut_vpm, (size_t)p->n_inst * n_vza * (p->mmax_all + 1) * 3 * 8);
    size_t max_jobs = 1;
    for (Chunk &c : p->chunks) {
        c.hjobs_lut.clear();
        std::vector<int> first(c.i1 - c.i0, -1);
        for (int i = c.i0; i < c.i1; ++i) {
            if (!p->inst[i].d.ipss) continue;
            first[i - c.i0] = (int)c.hjobs_lut.size();
            for (int iv = 0; iv < n_vza; ++iv)
                for (int ir = 0; ir < n_raa; ++ir) c.hjobs_lut.push_back(IpssJob{i - c.i0, 0, vza_deg[iv], raa_deg[ir]});
        }
        max_jobs = std::max(max_jobs, c.hjobs_lut.size());
        rc |= dmalloc(p, (void **)&c.djob_lut_of, first.size() * sizeof(int));
        rc |= dmalloc(p, (void **)&c.djobs_lut, std::max<size_t>(c.hjobs_lut.size(), 1) * sizeof(IpssJob));
        if (rc) return OCRT_ERR_ALLOC;
        HIPCHK(hipMemcpy(c.djob_lut_of, first.data(), first.size() * sizeof(int), hipMemcpyHostToDevice));
        if (!c.hjobs_lut.empty())
            HIPCHK(hipMemcpy(c.djobs_lut, c.hjobs_lut.data(), c.hjobs_lut.size() * sizeof(IpssJob),
                             hipMemcpyHostToDevice));
    }
    rc |= dmalloc(p, (void **)&p->kappa_lut, max_jobs * 8);
    rc |= dmalloc(p, (void **)&p->taudir_lut, max_jobs * 8);
    if (rc) return OCRT_ERR_ALLOC;
    HIPCHK(hipMemcpy(p->lut_vza, vza_deg, (size_t)n_vza * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(p->lut_raa, raa_deg, (size_t)n_raa * 8, hipMemcpyHostToDevice));
    p->lut_nv = n_vza;
    p->lut_nr = n_raa;
    p->lut = 1;
    return OCRT_OK;
}

extern "C" int ocrt_plan_lut_results(ocrt_plan *p, double *grid, double *scalars)
{
    if (!p || !p->lut) return OCRT_ERR_ARGS;
    HIPCHK(hipDeviceSynchronize());
    const size_t g = (size_t)p->n_inst * p->lut_nv * p->lut_nr * 6;
    if (grid) HIPCHK(hipMemcpy(grid, p->lut_grid, g * 8, hipMemcpyDeviceToHost));
    if (scalars) HIPCHK(hipMemcpy(scalars, p->lut_scal, (size_t)p->n_inst * 2 * 8, hipMemcpyDeviceToHost));
    return OCRT_OK;
}

extern "C" int ocrt_plan_lut_extras(ocrt_plan *p, double *raw, double *view_modes)
{
    if (!p || !p->lut) return OCRT_ERR_ARGS;
    HIPCHK(hipDeviceSynchronize());
    const size_t g = (size_t)p->n_inst * p->lut_nv * p->lut_nr * 3;
    const size_t v = (size_t)p->n_inst * p->lut_nv * (p->mmax_all + 1) * 3;
    if (raw) HIPCHK(hipMemcpy(raw, p->lut_raw, g * 8, hipMemcpyDeviceToHost));
    if (view_modes) HIPCHK(hipMemcpy(view_modes, p->lut_vpm, v * 8, hipMemcpyDeviceToHost));
    return OCRT_OK;
}

extern "C" int ocrt_plan_export_nodes(ocrt_plan *p)
{
    if (!p) return OCRT_ERR_ARGS;
    if (p->nodes) return OCRT_OK;
    drop_graph(p);   // extract's arguments change
    return dmalloc(p, (void **)&p->nodes, p->nodes_n * 8) ? OCRT_ERR_ALLOC : OCRT_OK;
}

extern "C" int ocrt_plan_node_fields(ocrt_plan *p, int ring_pitch, double *toa_up, double *boa_dn)
{
    if (!p || !p->nodes || ring_pitch < 1) return OCRT_ERR_ARGS;
    for (const Chunk &c : p->chunks)
        for (const DevItem &it : c.hitem)
            if (it.n > ring_pitch) return OCRT_ERR_ARGS;
    HIPCHK(hipDeviceSynchronize());
    std::vector<double> h((size_t)p->nodes_n);
    HIPCHK(hipMemcpy(h.data(), p->nodes, (size_t)p->nodes_n * 8, hipMemcpyDeviceToHost));
    const int M = p->mmax_all + 1;
    const size_t total = (size_t)p->n_inst * M * 3 * ring_pitch;
    if (toa_up) std::fill(toa_up, toa_up + total, 0.0);
    if (boa_dn) std::fill(boa_dn, boa_dn + total, 0.0);
    for (const Chunk &c : p->chunks)
        for (const DevItem &it : c.hitem) {
            const size_t gi = (size_t)(c.i0 + it.inst), n = (size_t)it.n;
            for (size_t s = 0; s < 3; ++s)
                for (size_t o = 0; o < n; ++o) {
                    const size_t at = ((gi * M + (size_t)it.m) * 3 + s) * (size_t)ring_pitch + o;
                    if (toa_up) toa_up[at] = h[(size_t)it.off_nodes + s * n + o];
                    if (boa_dn) boa_dn[at] = h[(size_t)it.off_nodes + (3 + s) * n + o];
                }
        }
    return OCRT_OK;
}

extern "C" int ocrt_plan_residuals(ocrt_plan *p, double *residuals)
{
    if (!p || !residuals) return OCRT_ERR_ARGS;
    HIPCHK(hipDeviceSynchronize());
    std::vector<ItemOut> io(p->n_items);
    HIPCHK(hipMemcpy(io.data(), p->iout, (size_t)p->n_items * sizeof(ItemOut), hipMemcpyDeviceToHost));
    const int M = p->mmax_all + 1;
    for (int i = 0; i < p->n_inst; ++i)
        for (int m = 0; m < M; ++m)
            residuals[(long long)i * M + m] = m <= p->inst[i].d.mmax ? io[p->item_of_inst[i] + m].residual : 0.0;
    return OCRT_OK;
}

static int execute_eager(ocrt_plan *p, hipStream_t s)
{
    p->gemm_launches = 0.0;
    p->ev_used = 0;
    // chunks follow one another in stream order (they share the pools); the
    // host does not wait
    for (Chunk &c : p->chunks)
        if (forward_chunk(p, c, s) != 0) return OCRT_ERR_DEVICE;
    // a single-chunk adjoint plan keeps the checkpoints of this forward pass
    p->fwd_ready = (p->adjoint && p->chunks.size() == 1) ? 1 : 0;
    return hipGetLastError() == hipSuccess ? OCRT_OK : OCRT_ERR_DEVICE;
}

// Autotuner: one warm-up execute, then one execute per sweep launch shape
// with the order loop's kernels bracketed by events; the shape with the least
// summed time is kept.  Every shape computes the same bits, so the
// results of the plan do not depend on the choice.
// Summed time of each operator class's GEMM launches over the last execute
// (the autotuner's per-class event pairs).
static int class_times(ocrt_plan *p, double out[3])
{
    HIPCHK(hipDeviceSynchronize());
    out[0] = out[1] = out[2] = 0.0;
    for (size_t q = 0; q + 1 < p->ev_used; q += 2) {
        const int k = p->ev_kind[q / 2];
        if (k < 2) continue;
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, p->ev[q], p->ev[q + 1]));
        out[k - 2] += ms;
    }
    return OCRT_OK;
}

static int autotune(ocrt_plan *p, hipStream_t s)
{
    const int saved = p->profile;
    p->profile = 0;
    p->sweep_variant = kDefaultSweepShape;
    int rc = execute_eager(p, s);
    p->profile = 1;
    // GEMM tile shape per operator class: one execute per shape, every class
    // timed with that shape; each class keeps its fastest
    p->tune_classes = 1;
    for (int sh = 0; sh < kGemmShapeCount && !rc; ++sh) {
        for (int k = 0; k < 3; ++k) p->gemm_shape[k] = sh;
        rc = execute_eager(p, s);
        double t[3];
        if (!rc) rc = class_times(p, t);
        for (int k = 0; k < 3; ++k) p->gemm_ms[k][sh] = t[k];
    }
    p->tune_classes = 0;
    for (int k = 0; k < 3; ++k) {
        int b = 0;
        for (int sh = 1; sh < kGemmShapeCount; ++sh)
            if (p->gemm_ms[k][sh] < p->gemm_ms[k][b]) b = sh;
        p->gemm_shape[k] = b;
    }
    p->gemm_tuned = 1;
    p->tune_ms.assign(kSweepShapeCount, 0.0);
    int best = 0;
    for (int v = 0; v < kSweepShapeCount && !rc; ++v) {
        p->sweep_variant = v;
        rc = execute_eager(p, s);
        // the whole order loop: scattering integral and sweep
        double gemm_ms = 0.0, sweep_ms = 0.0;
        if (!rc) rc = ocrt_plan_kernel_times(p, &gemm_ms, nullptr, &sweep_ms, nullptr);
        p->tune_ms[v] = gemm_ms + sweep_ms;
        if (p->tune_ms[v] < p->tune_ms[best]) best = v;
    }
    p->profile = saved;
    p->sweep_variant = best;
    p->tuned = 1;
    if (getenv("OCRT_DEBUG_PLAN"))
        for (int v = 0; v < kSweepShapeCount; ++v)
            fprintf(stderr, "ocrt autotune: order loop U %d W %d: %.3f ms%s\n", kSweepShapes[v].U, kSweepShapes[v].W,
                    p->tune_ms[v], v == best ? " (chosen)" : "");
    return rc;
}

// Capture the forward launch sequence of every chunk into one graph.
static int capture(ocrt_plan *p)
{
    if (!p->own_stream) HIPCHK(hipStreamCreateWithFlags(&p->own_stream, hipStreamNonBlocking));
    hipStream_t cs = p->own_stream;
    const int saved = p->profile;
    p->profile = 0;
    p->gemm_launches = 0.0;
    HIPCHK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (Chunk &c : p->chunks)
        if (!rc && forward_chunk(p, c, cs) != 0) rc = OCRT_ERR_DEVICE;
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(cs, &g);
    p->profile = saved;
    if (rc || ec != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        return OCRT_ERR_DEVICE;
    }
    p->graph = g;
    p->graph_gemm_launches = p->gemm_launches;
    HIPCHK(hipGraphInstantiate(&p->graph_exec, p->graph, nullptr, nullptr, 0));
    return OCRT_OK;
}

static void drop_graph(ocrt_plan *p)
{
    if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->graph_exec = nullptr;
    p->graph = nullptr;
}

extern "C" int ocrt_plan_execute(ocrt_plan *p, void *stream)
{
    if (!p) return OCRT_ERR_ARGS;
    hipStream_t s = (hipStream_t)stream;
    if (p->autotune && !p->tuned && p->prec == OCRT_PREC_MIXED) {
        const int rc = autotune(p, s);
        if (rc) return rc;
    }
    if (p->use_graph && !p->profile) {
        if (!p->graph_exec) {
            const int rc = capture(p);
            if (rc) return rc;
        }
        HIPCHK(hipGraphLaunch(p->graph_exec, s));
        p->gemm_launches = p->graph_gemm_launches;
        p->ev_used = 0;
        p->fwd_ready = 0;
        return hipGetLastError() == hipSuccess ? OCRT_OK : OCRT_ERR_DEVICE;
    }
    return execute_eager(p, s);
}

extern "C" int ocrt_plan_set_tuning(ocrt_plan *p, int sweep_unroll, int sweep_waves)
{
    if (!p) return OCRT_ERR_ARGS;
    for (int v = 0; v < kSweepShapeCount; ++v)
        if (kSweepShapes[v].U == sweep_unroll && kSweepShapes[v].W == sweep_waves) {
            drop_graph(p);   // a captured graph holds the previous shape's launches
            p->sweep_variant = v;
            p->tuned = 1;
            return OCRT_OK;
        }
    return OCRT_ERR_ARGS;
}

extern "C" int ocrt_plan_tuning(const ocrt_plan *p, int *sweep_unroll, int *sweep_waves, double *ms_per_shape,
                                int *n_shapes)
{
    if (!p) return OCRT_ERR_ARGS;
    if (sweep_unroll) *sweep_unroll = kSweepShapes[p->sweep_variant].U;
    if (sweep_waves) *sweep_waves = kSweepShapes[p->sweep_variant].W;
    if (n_shapes) *n_shapes = p->tuned ? kSweepShapeCount : 0;
    if (ms_per_shape && p->tuned)
        for (int v = 0; v < kSweepShapeCount; ++v) ms_per_shape[v] = p->tune_ms[v];
    return OCRT_OK;
}

extern "C" int ocrt_plan_set_gemm_shape(ocrt_plan *p, int op_class, int shape)
{
    if (!p || op_class < 0 || op_class > 2 || shape < 0 || shape >= kGemmShapeCount) return OCRT_ERR_ARGS;
    drop_graph(p);   // a captured graph holds the previous shape's launches
    p->gemm_shape[op_class] = shape;
    return OCRT_OK;
}

extern "C" int ocrt_plan_set_reduced_threshold(ocrt_plan *p, double tau)
{
    if (!p || p->prec != OCRT_PREC_MIXED || !(tau >= 0.0) || !(tau <= 1.0)) return OCRT_ERR_ARGS;
    drop_graph(p);   // a captured graph holds the previous threshold
    p->pools.lo_tau = (float)tau;
    return OCRT_OK;
}

extern "C" int ocrt_plan_gemm_issued(const ocrt_plan *p, double *issued_flops, double *item_orders,
                                     double *reduced_orders)
{
    if (!p) return OCRT_ERR_ARGS;
    if (issued_flops) *issued_flops = p->gemm_issued;
    if (item_orders) *item_orders = p->item_orders;
    if (reduced_orders) *reduced_orders = p->reduced_orders;
    return OCRT_OK;
}

extern "C" int ocrt_plan_gemm_tuning(const ocrt_plan *p, int *shape_of_class, double *ms)
{
    if (!p) return OCRT_ERR_ARGS;
    for (int k = 0; k < 3; ++k) {
        if (shape_of_class) shape_of_class[k] = p->gemm_shape[k];
        if (ms)
            for (int sh = 0; sh < kGemmShapeCount; ++sh)
                ms[k * kGemmShapeCount + sh] = p->gemm_tuned ? p->gemm_ms[k][sh] : 0.0;
    }
    return p->gemm_tuned ? kGemmShapeCount : 0;
}

extern "C" int ocrt_plan_adjoint(ocrt_plan *p, const double *grad_out, double *grad_levels, double *grad_wind,
                                 void *stream)
{
    if (!p || !grad_out || !p->adjoint) return OCRT_ERR_ARGS;
    hipStream_t s = (hipStream_t)stream;
    HIPCHK(hipMemcpyAsync(p->gout, grad_out, (size_t)p->n_inst * 3 * 8, hipMemcpyHostToDevice, s));
    for (Chunk &c : p->chunks) {
        if (!p->fwd_ready && forward_chunk(p, c, s) != 0) return OCRT_ERR_DEVICE;
        if (backward_chunk(p, c, s) != 0) return OCRT_ERR_DEVICE;
    }
    p->fwd_ready = (p->chunks.size() == 1) ? 1 : 0;
    HIPCHK(hipStreamSynchronize(s));   // host outputs: the call returns when they are complete
    if (!grad_levels && !grad_wind) return OCRT_OK;
    std::vector<double> g(p->grad_n);
    HIPCHK(hipMemcpy(g.data(), p->grad, (size_t)p->grad_n * 8, hipMemcpyDeviceToHost));
    long long off = 0;
    for (int i = 0; i < p->n_inst; ++i) {
        const DevInst &I = p->inst[i].d;
        const int nt = I.nt, L = nt + 1;
        const double *gH = g.data() + I.off_grad, *gX = gH + L, *gY = gX + L, *gD = gY + L;
        if (grad_levels) {
            double *dh = grad_levels + off, *dx = dh + L, *dy = dx + L;
            for (int k = 0; k <= nt; ++k) {
                // h[k] enters explicitly (gH) and through the layers above (k-1) and below (k)
                dh[k] = gH[k] + (k > 0 ? gD[k - 1] : 0.0) - (k < nt ? gD[k] : 0.0);
                dx[k] = gX[k];
                dy[k] = gY[k];
            }
        }
        if (grad_wind) grad_wind[i] = gH[4 * L - 1];
        off += 3LL * L;
    }
    return OCRT_OK;
}

// ---------------------------------------------------------------------------
// Device-memory inputs and outputs (callers whose data already lives on the
// GPU).  Every copy and kernel goes on the caller's stream.

// New layer optical depths and wind speeds of one chunk's instances: the
// column totals the medium kernels read, and the checks the host-side
// validation makes at plan creation, against the structure the plan was made
// with.
__global__ void refresh_inputs(DevInst *__restrict__ inst, int n_inst, const double *__restrict__ layers,
                               const double *__restrict__ wind, int set_layers)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_inst) return;
    DevInst &I = inst[i];
    int bad = 0;
    if (set_layers && I.layered) {
        const int nt = I.nt;
        const double *R = layers + I.off_layers, *A = R + nt, *W = A + nt, *G = W + nt;
        double sr = 0.0, sa = 0.0;
        for (int l = 0; l < nt; ++l) {
            if (!(R[l] >= 0.0 && A[l] >= 0.0 && G[l] >= 0.0 && W[l] >= 0.0 && W[l] <= 1.0)) bad = 1;
            if (!(R[l] + A[l] + G[l] > 0.0) || !isfinite(R[l] + A[l] + G[l])) bad = 1;
            sr += R[l];
            sa += A[l];
        }
        // an aerosol column appearing or vanishing changes the operator parts
        if (I.aer ? !(sa > 0.0) : (I.aer_able && sa > 0.0)) bad = 1;
        I.tauR = sr;
        I.taua = I.aer ? sa : 0.0;
    }
    if (wind) {
        const double v = wind[I.gidx];
        if (!isfinite(v)) bad = 1;
        // rough kernels need a positive wind; a flat sea was made from a calm one
        if (I.surface_kind == OCRT_SURFACE_BLACK_FRESNEL_OCEAN && (I.surf == 2) != (v > 0.0)) bad = 1;
        I.wind = v;
    }
    I.input_bad = bad;
}

extern "C" int ocrt_plan_set_inputs(ocrt_plan *p, const double *d_layers, const double *d_wind, void *stream)
{
    if (!p) return OCRT_ERR_ARGS;
    if (d_wind && !p->dev_inputs) return OCRT_ERR_ARGS;
    long long total = 0;
    for (const HostInst &h : p->inst) {
        if (d_layers && !h.d.layered) return OCRT_ERR_ARGS;
        total += h.d.layered ? 4LL * h.d.nt : 0;
    }
    hipStream_t s = (hipStream_t)stream;
    // every instance layered: the plan's layer pool is the packed concatenation
    if (d_layers && total > 0)
        HIPCHK(hipMemcpyAsync(p->layers, d_layers, (size_t)total * 8, hipMemcpyDeviceToDevice, s));
    for (Chunk &c : p->chunks) {
        const int ni = c.i1 - c.i0;
        hipLaunchKernelGGL(refresh_inputs, dim3((unsigned)((ni + 63) / 64)), dim3(64), 0, s, c.dinst, ni, p->layers,
                           d_wind, d_layers ? 1 : 0);
    }
    p->fwd_ready = 0;
    return hipGetLastError() == hipSuccess ? OCRT_OK : OCRT_ERR_DEVICE;
}

__global__ void rho_out(const ocrt_result_t *__restrict__ res, int n, double *__restrict__ rho, int *__restrict__ status)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const ocrt_result_t &r = res[i];
    const bool ok = r.status == OCRT_OK;
    const double nan = __builtin_nan("");
    rho[3 * i] = ok ? r.rho_I : nan;
    rho[3 * i + 1] = ok ? r.rho_Q : nan;
    rho[3 * i + 2] = ok ? r.rho_U : nan;
    if (status) status[i] = r.status;
}

extern "C" int ocrt_plan_rho(ocrt_plan *p, double *d_rho, int *d_status, void *stream)
{
    if (!p || !d_rho) return OCRT_ERR_ARGS;
    hipLaunchKernelGGL(rho_out, dim3((unsigned)((p->n_inst + 63) / 64)), dim3(64), 0, (hipStream_t)stream, p->res,
                       p->n_inst, d_rho, d_status);
    return hipGetLastError() == hipSuccess ? OCRT_OK : OCRT_ERR_DEVICE;
}

// Per-layer gradients of one chunk's layered instances from the adjoint's
// per-level accumulators, by the chain rule through the layer-to-level map
// (rt_atm.c:343-360):
//   h[k] = sum_{l<k} dtau_l,  dtau_l = R_l + A_l + G_l,
//   xdel[l+1] = W_l A_l / dtau_l,  ydel[l+1] = R_l / dtau_l,
//   level 0: (0, 1) with Rayleigh present, else (W_0, 0).
// The level gradient of h is the one ocrt_plan_adjoint reports, d/dh[k] =
// gH[k] + gD[k-1] - gD[k]; the sums over the levels below a layer run from
// the bottom up.  Contraction is off so the arithmetic is the plain sequence
// of roundings the host-side map (ocrt_amd/adjoint.py) performs.
__global__ void layer_grads(const DevInst *__restrict__ inst, int n_inst, const double *__restrict__ layers,
                            const double *__restrict__ grad, double *__restrict__ gl, double *__restrict__ gw)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_inst) return;
    const DevInst &I = inst[i];
    const int nt = I.nt, L = nt + 1;
    const double *gH = grad + I.off_grad, *gX = gH + L, *gY = gX + L, *gD = gY + L;
    if (gw) gw[I.gidx] = gH[4 * L - 1];
    if (!gl) return;
    const double *R = layers + I.off_layers, *A = R + nt, *W = A + nt, *G = W + nt;
    double *oR = gl + I.off_layers, *oA = oR + nt, *oW = oA + nt, *oG = oW + nt;
    const bool aer = I.aer != 0;
    double suf = 0.0, rsum = 0.0;
    for (int l = nt - 1; l >= 0; --l) {
        const int k = l + 1;
        const double dh = gH[k] + gD[k - 1] - (k < nt ? gD[k] : 0.0);
        suf = suf + dh;   // d/d dtau_l: every level below layer l
        const double dt = R[l] + A[l] + G[l];
        const double gx1 = aer ? gX[k] : 0.0, gy1 = gY[k];
        const double x1 = aer ? W[l] * A[l] / dt : 0.0, y1 = R[l] / dt;
        const double common = suf - (gx1 * x1 + gy1 * y1) / dt;
        oR[l] = common + gy1 / dt;
        oA[l] = common + (aer ? gx1 * W[l] / dt : 0.0);
        oW[l] = aer ? gx1 * A[l] / dt : 0.0;
        oG[l] = common;
        rsum += R[l];
    }
    // without Rayleigh the top level's fractions are (W_0, 0)
    if (aer && !(rsum > 0.0)) oW[0] += gX[0];
}

extern "C" int ocrt_plan_adjoint_layers(ocrt_plan *p, const double *d_grad_out, double *d_grad_layers,
                                        double *d_grad_wind, void *stream)
{
    if (!p || !d_grad_out || !p->adjoint) return OCRT_ERR_ARGS;
    if (d_grad_layers)
        for (const HostInst &h : p->inst)
            if (!h.d.layered) return OCRT_ERR_ARGS;
    hipStream_t s = (hipStream_t)stream;
    HIPCHK(hipMemcpyAsync(p->gout, d_grad_out, (size_t)p->n_inst * 3 * 8, hipMemcpyDeviceToDevice, s));
    for (Chunk &c : p->chunks) {
        if (!p->fwd_ready && forward_chunk(p, c, s) != 0) return OCRT_ERR_DEVICE;
        if (backward_chunk(p, c, s) != 0) return OCRT_ERR_DEVICE;
        const int ni = c.i1 - c.i0;
        if (d_grad_layers || d_grad_wind)
            hipLaunchKernelGGL(layer_grads, dim3((unsigned)((ni + 63) / 64)), dim3(64), 0, s, c.dinst, ni, p->layers,
                               p->grad, d_grad_layers, d_grad_wind);
    }
    p->fwd_ready = (p->chunks.size() == 1) ? 1 : 0;
    return hipGetLastError() == hipSuccess ? OCRT_OK : OCRT_ERR_DEVICE;
}

extern "C" int ocrt_plan_results(ocrt_plan *p, ocrt_result_t *out, double *toa_modes, int *orders)
{
    if (!p || !out) return OCRT_ERR_ARGS;
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(out, p->res, (size_t)p->n_inst * sizeof(ocrt_result_t), hipMemcpyDeviceToHost));
    std::vector<ItemOut> io(p->n_items);
    HIPCHK(hipMemcpy(io.data(), p->iout, (size_t)p->n_items * sizeof(ItemOut), hipMemcpyDeviceToHost));
    const int M = p->mmax_all + 1;
    double flops = 0.0, alg = 0.0, bytes = 0.0, issued = 0.0, n_ord = 0.0, n_red = 0.0;
    const bool mixed = p->prec == OCRT_PREC_MIXED;
    for (int i = 0; i < p->n_inst; ++i) {
        const DevInst &I = p->inst[i].d;
        int n_in = 0;
        for (double w : p->inst[i].w) n_in += (w != 0.0);
        for (int m = 0; m <= I.mmax; ++m) {
            const ItemOut &r = io[p->item_of_inst[i] + m];
            if (toa_modes)
                for (int c = 0; c < 3; ++c) toa_modes[((long long)i * M + m) * 3 + c] = r.mode[c];
            if (orders) orders[(long long)i * M + m] = r.orders;
            const int ord = std::max(r.orders - 1, 0);
            const double L = I.nt + 1.0;
            const bool lowrank = mixed && !p->adjoint && m <= 2 && I.ray && I.aer;
            if (mixed) {
                // the tiles of the item's class shape; per 16 x 16 x 32 block
                // and part: 2 MFMAs (A_hi, A_lo), 3 with the field pair, 1
                // reduced
                const int np = (I.aer ? 1 : 0) + ((m <= 2 && I.ray && !lowrank) ? 1 : 0);
                const bool split = m < kSplitModes;
                const int cls = np > 1 ? 2 : (split ? 1 : 0);
                const GemmShape sh = kGemmShapes[p->gemm_shape[cls]];
                const int Mp6 = roundup(6 * I.n, 64);
                const int Pp = roundup(6 * I.n, 32), Mop = Mp6 + ((lowrank && roundup(6 * I.n, 4) + 4 > Mp6) ? 64 : 0);
                const double blocks = (double)((Mop + sh.tm - 1) / sh.tm) * ((I.nt + 1 + sh.tn - 1) / sh.tn) *
                                      (sh.tm / 16) * (sh.tn / 16) * (Pp / 32) * std::max(np, 1);
                const int red = std::min(r.reduced, ord);
                issued += blocks * 16384.0 * ((ord - red) * (split ? 3.0 : 2.0) + red);
                n_ord += ord;
                n_red += red;
            }
            const int parts = (I.aer ? 1 : 0) + ((m <= 2 && I.ray && !lowrank) ? 1 : 0);
            // issued: the dense parts, and a rank-3 item's 16 projection rows
            flops += ord * 2.0 * (6.0 * I.n + (lowrank ? 16.0 : 0.0)) * (6.0 * I.n * std::max(parts, 1)) * L;
            // dense particle operator 2 (6 n)(6 n_in) per level; the molecular
            // operator has rank <= 6 and counts as 2 * 2 * (6 n) * 6 per level
            alg += ord * L * ((I.aer ? 72.0 * I.n * n_in : 0.0) + ((m <= 2 && I.ray) ? 144.0 * I.n : 0.0));
            // per element and order: read the source (mixed form: fp16, fp32
            // for the low modes), write the field (bf16, plus its residual for
            // the low modes, in the mixed form), read and write the order sum
            // (fp64; fp32 in the mixed form's interior rows); per layer and
            // node of each hemisphere: read 1 - T
            const double el = 6.0 * I.n * L;
            const double fb = mixed ? (m < kSplitModes ? 4.0 : 2.0) : 8.0;
            const double jb = mixed ? (m < kSplitModes ? 4.0 : 2.0) : 8.0;
            bytes += ord * (el * (jb + fb + (mixed ? 8.0 : 16.0)) + 2.0 * I.n * I.nt * (mixed ? 4.0 : 8.0));
        }
        if (toa_modes)
            for (int m = I.mmax + 1; m < M; ++m)
                for (int c = 0; c < 3; ++c) toa_modes[((long long)i * M + m) * 3 + c] = 0.0;
        if (orders)
            for (int m = I.mmax + 1; m < M; ++m) orders[(long long)i * M + m] = 0;
    }
    p->gemm_flops = flops;
    p->gemm_flops_alg = alg;
    p->gemm_issued = issued;
    p->item_orders = n_ord;
    p->reduced_orders = n_red;
    p->sweep_bytes = bytes;
    return OCRT_OK;
}

extern "C" int ocrt_plan_mmax(const ocrt_plan *p) { return p ? p->mmax_all : -1; }
extern "C" size_t ocrt_plan_device_bytes(const ocrt_plan *p) { return p ? p->dev_bytes : 0; }

extern "C" int ocrt_plan_work(const ocrt_plan *p, double *gemm_flops, double *sweep_bytes, double *launches)
{
    if (!p) return OCRT_ERR_ARGS;
    if (gemm_flops) *gemm_flops = p->gemm_flops_alg;
    if (sweep_bytes) *sweep_bytes = p->sweep_bytes;
    if (launches) *launches = p->gemm_launches;
    return OCRT_OK;
}

extern "C" int ocrt_plan_profile(ocrt_plan *p, int enable)
{
    if (!p) return OCRT_ERR_ARGS;
    p->profile = enable ? 1 : 0;
    return OCRT_OK;
}

extern "C" int ocrt_plan_kernel_times(ocrt_plan *p, double *gemm_ms, int *gemm_launches, double *sweep_ms,
                                      int *sweep_launches)
{
    if (!p) return OCRT_ERR_ARGS;
    HIPCHK(hipDeviceSynchronize());
    double t[2] = {0.0, 0.0};
    int cnt[2] = {0, 0};
    for (size_t q = 0; q + 1 < p->ev_used; q += 2) {
        const int k = p->ev_kind[q / 2];
        if (k > 1) continue;   // per-class GEMM pairs of the autotuner
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, p->ev[q], p->ev[q + 1]));
        t[k] += ms;
        cnt[k] += 1;
    }
    if (gemm_ms) *gemm_ms = t[0];
    if (gemm_launches) *gemm_launches = cnt[0];
    if (sweep_ms) *sweep_ms = t[1];
    if (sweep_launches) *sweep_launches = cnt[1];
    return OCRT_OK;
}

extern "C" void ocrt_plan_destroy(ocrt_plan *p)
{
    if (!p) return;
    drop_graph(p);
    if (p->own_stream) (void)hipStreamDestroy(p->own_stream);
    for (hipEvent_t e : p->ev) (void)hipEventDestroy(e);
    for (void *a : p->allocs) (void)hipFree(a);
    delete p;
}

extern "C" int ocrt_solve_batch(const ocrt_instance_t *in, int n, int precision, ocrt_result_t *out, void *stream)
{
    ocrt_plan *p = nullptr;
    int rc = ocrt_plan_create(in, n, precision, 0, &p);
    if (rc) return rc;
    rc = ocrt_plan_execute(p, stream);
    if (!rc) rc = ocrt_plan_results(p, out, nullptr, nullptr);
    ocrt_plan_destroy(p);
    return rc;
}

extern "C" int ocrt_adjoint_batch(const ocrt_instance_t *in, int n, int precision, const double *grad_out,
                                  ocrt_result_t *out, double *grad_levels, double *grad_wind, void *stream)
{
    ocrt_plan *p = nullptr;
    int rc = ocrt_plan_create_ex(in, n, precision, OCRT_PLAN_ADJOINT, &p);
    if (rc) return rc;
    rc = ocrt_plan_execute(p, stream);
    if (!rc) rc = ocrt_plan_adjoint(p, grad_out, grad_levels, grad_wind, stream);
    if (!rc && out) rc = ocrt_plan_results(p, out, nullptr, nullptr);
    ocrt_plan_destroy(p);
    return rc;
}

extern "C" int ocrt_ring(int n_mu, double vza_deg, double *mu, double *w)
{
    if (n_mu < 2 || n_mu > kMaxRing - 1 || !(vza_deg >= 0.0 && vza_deg < 90.0)) return OCRT_ERR_ARGS;
    std::vector<double> rm, rw;
    int vidx = 0;
    const int n = build_ring(n_mu, cos(vza_deg * OCRT_PI / 180.0), rm, rw, vidx);
    if (n < 2) return OCRT_ERR_ARGS;
    for (int o = 0; o < n; ++o) {
        if (mu) mu[o] = rm[o];
        if (w) w[o] = rw[o];
    }
    return n;
}

extern "C" const char *ocrt_version(void) { return "ocrt_amd 0.1 gfx950"; }
