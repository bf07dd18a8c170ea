"""The mixed form's reduced operator products (ocrt_plan_set_reduced_threshold):
once an item's convergence ratio max |I_k| / max |order sum| falls below tau,
its following orders' scattering integrals use the bf16 operator and field
alone.

- tau = 0 runs every order in the full form (no reduced orders counted);
  the default threshold runs a share of the item-orders reduced and stays
  within the mixed form's error budget against the fp64 form (1e-4 relative
  in I, 1e-5 absolute in Q and U), as does tau = 0;
- a graph replay of a plan with a threshold equals its eager execute bit for
  bit, and changing the threshold after a capture takes effect;
- the issued-flop count falls with the reduced orders;
- fp64 plans and thresholds outside [0, 1] are refused."""
import numpy as np
import pytest

from ocrt_amd import synthetic as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case():
    import ocrt_amd
    descs, keep = S.config2(n_geom=16, seed=21, m_max=10)
    pf = ocrt_amd.Plan(descs, precision="fp64")
    pf.execute()
    ref = pf.results()
    pf.close()
    return ocrt_amd, descs, keep, ref


def _err(r, ref):
    eI = np.max(np.abs(r["rho_I"] - ref["rho_I"]) / np.abs(ref["rho_I"]))
    eQU = np.max(np.maximum(np.abs(r["rho_Q"] - ref["rho_Q"]), np.abs(r["rho_U"] - ref["rho_U"])))
    return eI, eQU


def test_threshold_full_and_default(case):
    ocrt, descs, keep, ref = case
    print(f"[gpu] reduced products: {len(descs)} instances", flush=True)
    p = ocrt.Plan(descs, precision="mixed")
    p.execute()
    r_def = p.results()
    g_def = p.gemm_issued()
    p.set_reduced_threshold(0.0)
    p.execute()
    r_full = p.results()
    g_full = p.gemm_issued()
    p.close()
    assert g_full["reduced_orders"] == 0 and g_full["item_orders"] > 0
    assert 0 < g_def["reduced_orders"] < g_def["item_orders"]
    assert g_def["issued_flops"] < g_full["issued_flops"]
    assert (r_def["status"] == 0).all() and (r_full["status"] == 0).all()
    for r in (r_def, r_full):
        eI, eQU = _err(r, ref)
        assert eI < 1e-4 and eQU < 1e-5, (eI, eQU)
    assert not np.array_equal(r_def["rho_I"], r_full["rho_I"])


def test_graph_replay_with_threshold(case):
    ocrt, descs, keep, ref = case
    e = ocrt.Plan(descs, precision="mixed")
    e.set_reduced_threshold(1e-2)
    e.execute()
    want = e.results()
    e.set_reduced_threshold(0.0)
    e.execute()
    want0 = e.results()
    e.close()
    g = ocrt.Plan(descs, precision="mixed", graph=True)
    g.set_reduced_threshold(1e-2)
    for _ in range(3):
        g.execute()
        got = g.results()
        for f in ("rho_I", "rho_Q", "rho_U", "n_orders_used"):
            assert np.array_equal(got[f], want[f]), f
    g.set_reduced_threshold(0.0)   # drops the captured graph
    g.execute()
    got = g.results()
    for f in ("rho_I", "rho_Q", "rho_U"):
        assert np.array_equal(got[f], want0[f]), f
    g.close()


def test_threshold_refused(case):
    ocrt, descs, keep, ref = case
    pf = ocrt.Plan(descs[:2], precision="fp64")
    with pytest.raises(ocrt.OcrtError):
        pf.set_reduced_threshold(1e-3)
    pf.close()
    pm = ocrt.Plan(descs[:2], precision="mixed")
    for bad in (-1e-3, 1.5, float("nan")):
        with pytest.raises(ocrt.OcrtError):
            pm.set_reduced_threshold(bad)
    pm.close()
