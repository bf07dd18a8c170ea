"""The GPU adjoint against the reference's own perturbation runs: the twelve
adjoint-workload instances of tests/golden/ref_jacobians.json (555 nm, 200
layers, 32 streams, modes 0..16, rough ocean) in one adjoint plan with
checkpoints every 4 orders; d(g . rho)/d wind from the adjoint and
d(g . rho)/d(AOD scale) from its per-layer gradients through the profile's
layer Jacobian (tests/jacobian_cases.py), against central differences
(h = 1e-4 relative) of the reference solver's own solves."""
import numpy as np
import pytest

from jacobian_cases import layered, load
from ocrt_amd.adjoint import layer_gradients

pytestmark = pytest.mark.gpu


def test_gpu_adjoint_matches_reference_finite_differences(oracle):
    import ocrt_amd as ocrt
    fx = load()
    g = np.asarray(fx["g"])
    built = [layered(oracle, c["desc"]) for c in fx["cases"]]
    descs = [b[0] for b in built]
    print(f"[gpu] adjoint (checkpoints every 4 orders) of {len(descs)} workload instances vs reference FD",
          flush=True)
    plan = ocrt.Plan(descs, precision="fp64", adjoint=True, ckpt_sparse=True)
    plan.execute()
    res = plan.results()
    assert (res["status"] == 0).all()
    adj = plan.adjoint(np.tile(g, (len(descs), 1)))
    plan.close()
    worst_w = worst_a = 0.0
    for i, (c, (d, lay, dlay)) in enumerate(zip(fx["cases"], built)):
        # the layered form reproduces the reference's solve of the profile
        rho = np.array([res["rho_I"][i], res["rho_Q"][i], res["rho_U"][i]])
        assert np.max(np.abs(rho - np.array(c["rho"]))) <= 1e-10 * abs(c["rho"][0]), i
        gl = layer_gradients(lay, adj["grad_h"][i], adj["grad_x"][i], adj["grad_y"][i])
        ga = float(np.sum(gl * dlay))
        ew = abs(adj["grad_wind"][i] - c["fd_wind"]) / abs(c["fd_wind"])
        ea = abs(ga - c["fd_aod"]) / abs(c["fd_aod"])
        worst_w, worst_a = max(worst_w, ew), max(worst_a, ea)
        assert ew < 1e-5 and ea < 1e-5, (i, adj["grad_wind"][i], c["fd_wind"], ga, c["fd_aod"])
    print(f"[gpu] adjoint vs reference FD: wind worst rel {worst_w:.1e}, AOD scale worst rel {worst_a:.1e}",
          flush=True)


def test_mixed_adjoint_matches_reference_finite_differences(oracle):
    """The mixed adjoint plan (precision "mixed": the mixed order loop with
    every forward field kept as its bf16 pair, fp32 adjoint fields, bf16 MFMA
    for the operator parts and A^T): the same Jacobians within 1e-3 of the
    reference's central differences, and every per-layer gradient within 1e-3
    of the fp64 adjoint's scale."""
    import ocrt_amd as ocrt
    fx = load()
    g = np.asarray(fx["g"])
    built = [layered(oracle, c["desc"]) for c in fx["cases"]]
    descs = [b[0] for b in built]
    print(f"[gpu] mixed adjoint of {len(descs)} workload instances vs reference FD", flush=True)
    out = {}
    for prec in ("fp64", "mixed"):
        plan = ocrt.Plan(descs, precision=prec, adjoint=True, ckpt_sparse=True)
        plan.execute()
        res = plan.results()
        assert (res["status"] == 0).all()
        out[prec] = (res, plan.adjoint(np.tile(g, (len(descs), 1))))
        plan.close()
    res, adj = out["mixed"]
    _, ref = out["fp64"]
    worst = [0.0, 0.0, 0.0, 0.0]
    for i, (c, (d, lay, dlay)) in enumerate(zip(fx["cases"], built)):
        rel = abs(res["rho_I"][i] - c["rho"][0]) / abs(c["rho"][0])
        gl = layer_gradients(lay, adj["grad_h"][i], adj["grad_x"][i], adj["grad_y"][i])
        gl64 = layer_gradients(lay, ref["grad_h"][i], ref["grad_x"][i], ref["grad_y"][i])
        ga = float(np.sum(gl * dlay))
        ew = abs(adj["grad_wind"][i] - c["fd_wind"]) / abs(c["fd_wind"])
        ea = abs(ga - c["fd_aod"]) / abs(c["fd_aod"])
        el = max(np.max(np.abs(gl[r] - gl64[r])) / np.max(np.abs(gl64[r])) for r in (0, 1, 2))
        worst = [max(worst[0], rel), max(worst[1], ew), max(worst[2], ea), max(worst[3], el)]
        assert rel < 1e-4 and ew < 1e-3 and ea < 1e-3 and el < 1e-3, (i, rel, ew, ea, el)
    print(f"[gpu] mixed adjoint: rho_I rel {worst[0]:.1e}; vs reference FD wind {worst[1]:.1e}, "
          f"AOD {worst[2]:.1e}; per-layer vs fp64 adjoint {worst[3]:.1e}", flush=True)
