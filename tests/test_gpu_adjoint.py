"""Adjoint (Jacobian) path on the GPU, through the C ABI.

- Level gradients of the HIP adjoint equal the oracle's reverse mode (which the
  CPU suite pins against finite differences of the oracle forward solve).
- Per-layer Jacobians of the workload's own instances (200 layers, 32 streams,
  rough ocean, modes 0..16) agree with central finite differences of the FP64
  GPU forward solve (h = 1e-4 relative) to 1e-3 relative.
- The torch.autograd.Function returns the same gradients.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import oracle_solve
from ocrt_amd.abi import make_instance
from ocrt_amd.adjoint import aod_scale_gradient, layer_gradients
from ocrt_amd import synthetic as S

pytestmark = pytest.mark.gpu

CASES = [
    dict(surface="black_fresnel_ocean", wind=7.0, sza=38.0, vza=47.0, raa=65.0, m_max=4),
    dict(surface="flat", sza=52.0, vza=23.0, raa=140.0, m_max=3),
    dict(surface="black", sza=20.0, vza=60.0, raa=10.0, m_max=3),
    dict(surface="black_fresnel_ocean", wind=4.0, decouple_sunglint=1, sza=30.0, vza=35.0, raa=170.0, m_max=3),
    dict(surface="black_fresnel_ocean", wind=11.0, sigma_type=0, sza=61.0, vza=12.0, raa=95.0, m_max=5),
]


@pytest.fixture(scope="module")
def ocrt():
    import ocrt_amd
    ocrt_amd.library()
    return ocrt_amd


def layered_cases(nt=8, seed=11):
    rng = np.random.default_rng(seed)
    gk = S.hg_greek(0.72, 24)
    out = []
    for kw in CASES:
        layers = np.stack([np.linspace(0.004, 0.02, nt), rng.uniform(0.01, 0.06, nt), rng.uniform(0.8, 0.99, nt),
                           rng.uniform(0.0, 0.01, nt)])
        out.append((layers, dict(greek=gk, n_mu=6, max_orders=30, tolerance=1e-12, **kw)))
    return out, gk


def oracle_adjoint(oracle, d, g):
    L = d.n_layers + 1
    gh, gx, gy = np.zeros(L), np.zeros(L), np.zeros(L)
    gw = C.c_double()
    gg = np.ascontiguousarray(g, dtype=np.float64)
    assert oracle.oracle_adjoint(C.byref(d), gg.ctypes.data_as(C.c_void_p), gh.ctypes.data_as(C.c_void_p),
                                 gx.ctypes.data_as(C.c_void_p), gy.ctypes.data_as(C.c_void_p), C.byref(gw)) == 0
    return gh, gx, gy, gw.value


def rel_err(a, b):
    """Largest difference relative to the reference's largest entry."""
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300)


def close(a, b, tol):
    scale = max(np.max(np.abs(b)), 1e-300)
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) <= tol * scale


def test_adjoint_matches_oracle_reverse_mode(ocrt, oracle):
    cases, gk = layered_cases()
    descs = [make_instance(layers=lay, **kw) for lay, kw in cases]
    g = np.array([[0.7, -0.4, 0.9], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.3, 0.2, -1.0], [-0.5, 0.8, 0.25]])
    print(f"[gpu] adjoint of {len(descs)} layered instances vs oracle", flush=True)
    plan = ocrt.Plan(descs, precision="fp64", adjoint=True)
    plan.execute()
    adj = plan.adjoint(g)
    plan.close()
    for i, d in enumerate(descs):
        gh, gx, gy, gw = oracle_adjoint(oracle, d, g[i])
        assert close(adj["grad_h"][i], gh, 1e-9), (i, "h")
        assert close(adj["grad_x"][i], gx, 1e-9), (i, "x")
        assert close(adj["grad_y"][i], gy, 1e-9), (i, "y")
        if d.wind_speed > 0:
            assert abs(adj["grad_wind"][i] - gw) <= 1e-6 * abs(gw) + 1e-14, (i, adj["grad_wind"][i], gw)


def test_sparse_checkpoints_equal_every_order_kept(ocrt):
    """Checkpointing every 4th order and recomputing the rest gives the
    gradients of the form that keeps every order (same kernels, same inputs)."""
    cases, gk = layered_cases(seed=7)
    descs = [make_instance(layers=lay, **kw) for lay, kw in cases]
    g = np.array([[0.7, -0.4, 0.9], [1.0, 0.2, 0.0], [0.0, 1.0, 0.3], [0.3, 0.2, -1.0], [-0.5, 0.8, 0.25]])
    print(f"[gpu] adjoint, sparse vs dense checkpoints, {len(descs)} instances", flush=True)
    out = []
    for sparse in (False, True):
        plan = ocrt.Plan(descs, precision="fp64", adjoint=True, ckpt_sparse=sparse)
        out.append(plan.adjoint(g))
        plan.close()
    for key in ("grad_h", "grad_x", "grad_y"):
        for i in range(len(descs)):
            assert close(out[1][key][i], out[0][key][i], 1e-12), (key, i)
    assert np.allclose(out[1]["grad_wind"], out[0]["grad_wind"], rtol=1e-12, atol=0.0)


def test_mixed_adjoint_matches_oracle_reverse_mode(ocrt, oracle):
    """Mixed adjoint plans (every forward field kept as its bf16 pair, fp32
    adjoint fields, bf16 MFMA operator products): level gradients within 1e-4
    of the oracle's reverse mode (relative to each gradient's largest entry),
    with dense and with sparse checkpoints."""
    cases, gk = layered_cases(seed=5)
    descs = [make_instance(layers=lay, **kw) for lay, kw in cases]
    g = np.array([[0.7, -0.4, 0.9], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.3, 0.2, -1.0], [-0.5, 0.8, 0.25]])
    ref = [oracle_adjoint(oracle, d, g[i]) for i, d in enumerate(descs)]
    for sparse in (False, True):
        print(f"[gpu] mixed adjoint of {len(descs)} layered instances vs oracle (sparse={sparse})", flush=True)
        plan = ocrt.Plan(descs, precision="mixed", adjoint=True, ckpt_sparse=sparse)
        plan.execute()
        res = plan.results()
        adj = plan.adjoint(g)
        plan.close()
        assert (res["status"] == 0).all()
        worst = max(rel_err(adj[k][i], r) for i in range(len(descs))
                    for k, r in zip(("grad_h", "grad_x", "grad_y"), ref[i][:3]))
        print(f"[gpu] mixed adjoint vs oracle reverse mode (sparse={sparse}): worst {worst:.2e} of each "
              f"gradient's largest entry", flush=True)
        for i, d in enumerate(descs):
            gh, gx, gy, gw = ref[i]
            assert close(adj["grad_h"][i], gh, 1e-4), (sparse, i, "h")
            assert close(adj["grad_x"][i], gx, 1e-4), (sparse, i, "x")
            assert close(adj["grad_y"][i], gy, 1e-4), (sparse, i, "y")
            if d.wind_speed > 0:
                assert abs(adj["grad_wind"][i] - gw) <= 1e-4 * abs(gw) + 1e-12, (sparse, i, adj["grad_wind"][i], gw)


def test_mixed_sparse_checkpoints_close_to_every_order_kept(ocrt):
    """Mixed plans rebuild the orders between checkpoints with the forward's
    own kernels; the rebuilt bottom row comes from the stored bf16 pair, so
    the gradients agree with the every-order form to 16 significant bits."""
    cases, gk = layered_cases(seed=7)
    descs = [make_instance(layers=lay, **kw) for lay, kw in cases]
    g = np.array([[0.7, -0.4, 0.9], [1.0, 0.2, 0.0], [0.0, 1.0, 0.3], [0.3, 0.2, -1.0], [-0.5, 0.8, 0.25]])
    print(f"[gpu] mixed adjoint, sparse vs dense checkpoints, {len(descs)} instances", flush=True)
    out = []
    for sparse in (False, True):
        plan = ocrt.Plan(descs, precision="mixed", adjoint=True, ckpt_sparse=sparse)
        out.append(plan.adjoint(g))
        plan.close()
    for key in ("grad_h", "grad_x", "grad_y"):
        for i in range(len(descs)):
            assert close(out[1][key][i], out[0][key][i], 1e-5), (key, i)
    assert np.allclose(out[1]["grad_wind"], out[0]["grad_wind"], rtol=1e-5, atol=1e-14)


def test_adjoint_plan_forward_equals_plain_fp64(ocrt):
    cases, gk = layered_cases()
    descs = [make_instance(layers=lay, **kw) for lay, kw in cases]
    a = ocrt.solve(descs, precision="fp64")
    p = ocrt.Plan(descs, precision="fp64", adjoint=True)
    p.execute()
    b = p.results()
    p.close()
    for f in ("rho_I", "rho_Q", "rho_U", "n_orders_used"):
        assert np.array_equal(a[f], b[f]), f


def workload_layers(oracle, d):
    """Per-layer description of a generator instance from the oracle's medium."""
    L = d.n_layers + 1
    h, ch, x, y, z = (np.zeros(L) for _ in range(5))
    assert oracle.oracle_medium(C.byref(d), *(a.ctypes.data_as(C.c_void_p) for a in (h, ch, x, y, z))) == 0
    dt = np.diff(h)
    aer = x[1:] * dt / d.ssa_a
    ray = y[1:] * dt
    return np.stack([ray, aer, np.full_like(dt, d.ssa_a), np.clip(dt - ray - aer, 0.0, None)])


def test_jacobian_matches_finite_differences_on_workload(ocrt, oracle):
    """Instances of the adjoint workload (8 bands drawn as the batched one,
    200 layers, 32 streams): d rho / d(layer aerosol tau, layer SSA), wind and
    AOD scale against central differences of the FP64 GPU forward solve."""
    gen, keep = S.config3(n_inst=4)
    base, descs = [], []
    for d in gen:
        lay = workload_layers(oracle, d)
        kw = dict(sza=d.sza_deg, vza=d.vza_deg, raa=d.raa_deg, wavelength=d.wavelength_nm,
                  surface="black_fresnel_ocean", wind=d.wind_speed, greek=np.ctypeslib.as_array(d.greek,
                  shape=(4, d.L_max + 1)).copy(), n_mu=d.n_mu, m_max=d.m_max, max_orders=d.max_orders,
                  tolerance=d.tolerance)
        base.append((lay, kw))
        descs.append(make_instance(layers=lay, **kw))
    # the layered form reproduces the generator solve
    a = ocrt.solve(gen, precision="fp64")
    b = ocrt.solve(descs, precision="fp64")
    assert np.max(np.abs(a["rho_I"] - b["rho_I"]) / np.abs(a["rho_I"])) < 1e-12
    g = np.array([1.0, 0.5, -0.5])
    plan = ocrt.Plan(descs, precision="fp64", adjoint=True)
    plan.execute()
    adj = plan.adjoint(np.tile(g, (len(descs), 1)))
    plan.close()
    rng = np.random.default_rng(5)
    worst = 0.0
    for i, (lay, kw) in enumerate(base):
        gl = layer_gradients(lay, adj["grad_h"][i], adj["grad_x"][i], adj["grad_y"][i])
        nt = lay.shape[1]
        picks = [(1, l) for l in rng.choice(nt, 3, replace=False)] + [(2, l) for l in rng.choice(nt, 3, replace=False)]
        pert = []
        for r, l in picks:
            step = 1e-4 * lay[r, l]
            lp, lm = lay.copy(), lay.copy()
            lp[r, l] += step
            lm[r, l] -= step
            pert += [make_instance(layers=lp, **kw), make_instance(layers=lm, **kw)]
        dw = 1e-4 * kw["wind"]
        pert += [make_instance(layers=lay, **dict(kw, wind=kw["wind"] + dw)),
                 make_instance(layers=lay, **dict(kw, wind=kw["wind"] - dw))]
        ls, lt = lay.copy(), lay.copy()
        ls[1] *= 1 + 1e-4
        lt[1] *= 1 - 1e-4
        pert += [make_instance(layers=ls, **kw), make_instance(layers=lt, **kw)]
        out = ocrt.solve(pert, precision="fp64")
        f = g @ np.stack([out["rho_I"], out["rho_Q"], out["rho_U"]])
        row_scale = np.max(np.abs(gl[1:3]), axis=1)
        for q, (r, l) in enumerate(picks):
            fd = (f[2 * q] - f[2 * q + 1]) / (2e-4 * lay[r, l])
            err = abs(gl[r, l] - fd) / row_scale[r - 1]
            worst = max(worst, err)
            assert err < 1e-3, (i, r, l, gl[r, l], fd)
        q = len(picks)
        fdw = (f[2 * q] - f[2 * q + 1]) / (2 * dw)
        assert abs(adj["grad_wind"][i] - fdw) <= 1e-3 * abs(fdw), (i, adj["grad_wind"][i], fdw)
        fds = (f[2 * q + 2] - f[2 * q + 3]) / 2e-4
        assert abs(aod_scale_gradient(lay, gl) - fds) <= 1e-3 * abs(fds)
    print(f"[gpu] workload Jacobian vs finite differences: worst rel {worst:.2e}", flush=True)


def test_autograd_function(ocrt, oracle):
    import torch
    from ocrt_amd.torch_ops import toa_stokes
    cases, gk = layered_cases(seed=3)
    cases = [c for c in cases if c[1].get("wind", 0.0) > 0.0]
    layers = torch.tensor(np.stack([c[0] for c in cases]), dtype=torch.float64, requires_grad=True)
    wind = torch.tensor([c[1]["wind"] for c in cases], dtype=torch.float64, requires_grad=True)
    base = [{k: v for k, v in c[1].items() if k != "wind"} for c in cases]
    rho = toa_stokes(layers, wind, base)
    w = torch.tensor([0.9, -0.2, 0.4], dtype=torch.float64)
    (rho * w).sum().backward()
    for i, (lay, kw) in enumerate(cases):
        d = make_instance(layers=lay, **kw)
        rc, r, _, _ = oracle_solve(oracle, d)
        assert abs(rho[i, 0].item() - r.rho_I) <= 1e-11 * abs(r.rho_I)
        gh, gx, gy, gw = oracle_adjoint(oracle, d, w.numpy())
        gl = layer_gradients(lay, gh, gx, gy)
        assert close(layers.grad[i].numpy(), gl, 1e-9)
        assert abs(wind.grad[i].item() - gw) <= 1e-6 * abs(gw)


def test_adjoint_forward_on_reused_nan_memory(ocrt):
    """Checkpoint buffers of an adjoint plan may start with any bit pattern:
    device memory filled with NaN and released before the plan is built does
    not reach the results (the fp64 sweeps write the pad columns of every
    order's field, which the next scattering integral reads)."""
    import torch
    cases, gk = layered_cases(seed=5)
    descs = [make_instance(layers=lay, **kw) for lay, kw in cases]
    a = ocrt.solve(descs, precision="fp64")
    for sparse in (False, True):
        junk = torch.full((1 << 28,), float("nan"), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        del junk
        torch.cuda.empty_cache()
        p = ocrt.Plan(descs, precision="fp64", adjoint=True, ckpt_sparse=sparse)
        p.execute()
        b = p.results()
        adj = p.adjoint(np.ones((len(descs), 3)))
        p.close()
        for f in ("rho_I", "rho_Q", "rho_U", "n_orders_used"):
            assert np.array_equal(a[f], b[f]), (sparse, f)
        for key in ("grad_h", "grad_x", "grad_y", "grad_wind"):
            assert np.all(np.isfinite(adj[key])), (sparse, key)
