"""Graph replay and the autotuner (OCRT_PLAN_GRAPH, OCRT_PLAN_AUTOTUNE).

- A plan that captures its forward launch sequence into a hipGraph gives, on
  every replay, the same bits as an eager plan over the same instances; the
  replays are interleaved with eager executes and synchronisations (mixed and
  FP64 forms, and a LUT plan).
- Every launch shape of the mixed-form sweep and every tile shape of its
  scattering integral (per operator class) gives the same bits, so an
  autotuned plan equals the default one; the tuner reports the shapes it
  chose and the time of every shape it tried.
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
    gt = t.gemm_tuning()
    print(f"[gpu] GEMM tiles chosen {gt}", flush=True)
    assert all(sh in ocrt.Plan.GEMM_SHAPES for sh in gt["shape_of_class"])
    # every class with tiles was timed with every shape
    assert gt["ms"] is not None and any(min(row) > 0.0 for row in gt["ms"])
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


@pytest.mark.parametrize("graph", [False, True])
def test_every_gemm_tile_shape_gives_the_same_bits(ocrt, graph):
    """Every tile shape of the scattering integral (64 x 128, 64 x 64,
    128 x 64, 128 x 128), for every operator class at once and class by class, gives the
    results of the default shape bit for bit, over repeated executes, also
    replayed from a graph; the batch has all three classes (modes 0..8: field
    pairs below m = 3, Rayleigh parts up to m = 2)."""
    descs, keep = S.config2(n_geom=6, seed=9, m_max=8)
    a = ocrt.Plan(descs, precision="mixed")
    a.execute()
    want = a.results(modes=True)
    a.close()
    print(f"[gpu] GEMM tile shapes, {len(descs)} instances, graph={graph}", flush=True)
    cases = [(sh, sh, sh) for sh in range(4)] + [(1, 2, 0), (2, 3, 1), (3, 0, 2)]
    for shapes in cases:
        p = ocrt.Plan(descs, precision="mixed", graph=graph)
        for k, sh in enumerate(shapes):
            p.set_gemm_shape(k, sh)
        for _ in range(2):
            p.execute()
            got = p.results(modes=True)
            same(got, want)
            assert np.array_equal(got["toa_modes"], want["toa_modes"]), shapes
        p.close()
