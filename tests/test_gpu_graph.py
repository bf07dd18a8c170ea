"""Graph replay and the autotuner (OCRT_PLAN_GRAPH, OCRT_PLAN_AUTOTUNE).

- A plan that captures its forward launch sequence into a hipGraph gives, on
  every replay, the same bits as an eager plan over the same instances; the
  replays are interleaved with eager executes and synchronisations (mixed and
  FP64 forms, and a LUT plan).
- Every launch shape of the mixed-form sweep gives the same bits, so an
  autotuned plan equals the default one; the tuner reports the shape it chose
  and the time of every shape it tried.
"""
import numpy as np
import pytest

from ocrt_amd import synthetic as S

pytestmark = pytest.mark.gpu

FIELDS = ("rho_I", "rho_Q", "rho_U", "T_dir_dn", "T_diff_dn_hemi", "n_orders_used", "converged", "status")


@pytest.fixture(scope="module")
def ocrt():
    import ocrt_amd
    ocrt_amd.library()
    return ocrt_amd


def same(a, b):
    for f in FIELDS:
        assert np.array_equal(a[f], b[f]), f


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_graph_replays_equal_eager(ocrt, precision):
    import torch
    descs, keep = S.config2(n_geom=8)
    print(f"[gpu] graph replay, {precision}, {len(descs)} instances", flush=True)
    eager = ocrt.Plan(descs, precision=precision)
    eager.execute()
    want = eager.results()
    g = ocrt.Plan(descs, precision=precision, graph=True)
    for rep in range(3):
        g.execute()
        torch.cuda.synchronize()
        same(g.results(), want)
        eager.execute()          # eager work between the replays
        torch.cuda.synchronize()
        same(eager.results(), want)
    g.close()
    eager.close()


def test_graph_lut_plan(ocrt):
    descs, keep = S.config2(n_geom=2)
    vza, raa = np.arange(0.0, 80.0, 10.0), np.arange(0.0, 360.0, 45.0)
    out = []
    for graph in (False, True):
        p = ocrt.Plan(descs, precision="fp64", graph=graph)
        p.set_lut(vza, raa)
        for _ in range(2):
            p.execute()
        out.append(p.lut())
        p.close()
    for k in out[0]:
        assert np.array_equal(out[0][k], out[1][k]), k


def test_autotuned_plan_equals_default(ocrt):
    descs, keep = S.config2(n_geom=16)
    print(f"[gpu] autotune over {len(descs)} instances", flush=True)
    a = ocrt.Plan(descs, precision="mixed")
    a.execute()
    want = a.results()
    a.close()
    t = ocrt.Plan(descs, precision="mixed", autotune=True, graph=True)
    t.execute()
    same(t.results(), want)
    info = t.tuning()
    print(f"[gpu] autotune chose {info}", flush=True)
    assert (info["unroll"], info["waves"]) in SHAPES
    assert len(info["ms_per_shape"]) == len(SHAPES) and min(info["ms_per_shape"]) > 0.0
    t.execute()
    same(t.results(), want)
    t.close()


# launch shapes of the mixed-form sweep (levels per load block, wavefronts per
# workgroup)
SHAPES = [(8, 4), (4, 4), (16, 4), (8, 2), (4, 2), (16, 2)]


@pytest.mark.parametrize("graph", [False, True])
def test_every_launch_shape_gives_the_same_bits(ocrt, graph):
    """Every sweep shape, chosen by hand, gives identical results over
    repeated executes, also replayed from a graph."""
    descs, keep = S.config2(n_geom=12, seed=5)
    a = ocrt.Plan(descs, precision="mixed")
    a.set_tuning(8, 4)
    a.execute()
    want = a.results()
    a.close()
    for u, w in SHAPES:
        p = ocrt.Plan(descs, precision="mixed", graph=graph)
        p.set_tuning(u, w)
        for _ in range(2):
            p.execute()
            same(p.results(), want)
        p.close()
