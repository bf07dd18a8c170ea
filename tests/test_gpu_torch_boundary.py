"""The torch boundary with device tensors (ocrt_amd.torch_ops, device_io):
the layers and winds reach the plan through their data_ptr()s, the solve
runs on torch's current stream (here a non-default one), and the results and
gradients come back in device tensors.  The TOA Stokes vector is bit for bit
what the host path (host copies, a fresh plan) gives; the gradients agree to
1e-13 of their scale -- the adjoint sums its per-level contributions with
atomic adds, whose order (and so the last bit) varies from run to run, on
either path.  Also: the cached plan serves later calls, a backward after
another forward still gives its own forward's gradients, a repeated backward
gives them again, and an input that breaks the plan's structure comes back
as NaN."""
import numpy as np
import pytest

from test_gpu_adjoint import layered_cases

pytestmark = pytest.mark.gpu


def _same(a, b, what):
    """rho bit for bit; gradients within 1e-13 of their scale."""
    if what == "rho":
        assert np.array_equal(a, b), (what, np.max(np.abs(a - b)))
    else:
        scale = max(np.max(np.abs(b)), 1e-300)
        assert np.max(np.abs(a - b)) <= 1e-13 * scale, (what, np.max(np.abs(a - b)), scale)


def _inputs(seed):
    import torch
    cases, _ = layered_cases(seed=seed)
    cases = [c for c in cases if c[1].get("wind", 0.0) > 0.0]
    layers = np.stack([c[0] for c in cases])
    wind = np.array([c[1]["wind"] for c in cases])
    base = [{k: v for k, v in c[1].items() if k != "wind"} for c in cases]
    return layers, wind, base


def _host(layers, wind, base, w, precision="fp64"):
    import torch
    from ocrt_amd.torch_ops import toa_stokes
    lay = torch.tensor(layers, dtype=torch.float64, requires_grad=True)
    wnd = torch.tensor(wind, dtype=torch.float64, requires_grad=True)
    rho = toa_stokes(lay, wnd, base, device_io=False, precision=precision)
    (rho * torch.tensor(w)).sum().backward()
    return rho.detach().numpy(), lay.grad.numpy(), wnd.grad.numpy()


def _device(layers, wind, base, w, stream, precision="fp64"):
    import torch
    from ocrt_amd.torch_ops import toa_stokes
    with torch.cuda.stream(stream):
        lay = torch.tensor(layers, dtype=torch.float64, device="cuda", requires_grad=True)
        wnd = torch.tensor(wind, dtype=torch.float64, device="cuda", requires_grad=True)
        rho = toa_stokes(lay, wnd, base, precision=precision)
        (rho * torch.tensor(w, device="cuda")).sum().backward()
    stream.synchronize()
    return rho.detach().cpu().numpy(), lay.grad.cpu().numpy(), wnd.grad.cpu().numpy()


def test_device_io_equals_host_path_on_a_side_stream():
    import torch
    from ocrt_amd import torch_ops
    torch_ops._CACHE.clear()
    layers, wind, base = _inputs(3)
    w = [0.9, -0.2, 0.4]
    s = torch.cuda.Stream()
    print(f"[gpu] torch boundary: {len(base)} layered instances, device tensors on a side stream", flush=True)
    h = _host(layers, wind, base, w)
    d = _device(layers, wind, base, w, s)
    for a, b, name in zip(h, d, ("rho", "layer gradients", "wind gradients")):
        _same(b, a, name)
    assert len(torch_ops._CACHE) == 1
    # new values through the same cached plan
    layers2 = layers * np.array([1.1, 0.9, 1.0, 1.2])[None, :, None]
    wind2 = wind + 1.5
    h2 = _host(layers2, wind2, base, w)
    d2 = _device(layers2, wind2, base, w, s)
    for a, b, name in zip(h2, d2, ("rho", "layer gradients", "wind gradients")):
        _same(b, a, name)
    assert len(torch_ops._CACHE) == 1
    assert not np.array_equal(h[0], h2[0])


def test_backward_after_another_forward_and_repeated():
    import torch
    from ocrt_amd.torch_ops import toa_stokes
    layers, wind, base = _inputs(3)
    w = torch.tensor([0.3, 1.0, -0.7], dtype=torch.float64, device="cuda")
    ref = _host(layers, wind, base, w.cpu().numpy())
    lay_a = torch.tensor(layers, dtype=torch.float64, device="cuda", requires_grad=True)
    wnd_a = torch.tensor(wind, dtype=torch.float64, device="cuda", requires_grad=True)
    rho_a = toa_stokes(lay_a, wnd_a, base)
    # a second forward on the same cached plan, with other inputs
    lay_b = torch.tensor(layers * 1.3, dtype=torch.float64, device="cuda", requires_grad=True)
    wnd_b = torch.tensor(wind * 0.8, dtype=torch.float64, device="cuda", requires_grad=True)
    rho_b = toa_stokes(lay_b, wnd_b, base)
    loss = (rho_a * w).sum()
    loss.backward(retain_graph=True)
    g1 = lay_a.grad.cpu().numpy().copy()
    lay_a.grad = None
    loss.backward()
    g2 = lay_a.grad.cpu().numpy()
    torch.cuda.synchronize()
    _same(rho_a.detach().cpu().numpy(), ref[0], "rho")
    _same(g1, ref[1], "layer gradients")
    _same(g2, ref[1], "layer gradients")
    assert not np.array_equal(rho_a.detach().cpu().numpy(), rho_b.detach().cpu().numpy())


def test_structure_breaking_input_is_nan():
    import torch
    from ocrt_amd.torch_ops import toa_stokes
    layers, wind, base = _inputs(3)
    lay = torch.tensor(layers, dtype=torch.float64, device="cuda")
    wnd = torch.tensor(wind, dtype=torch.float64, device="cuda")
    ok = toa_stokes(lay, wnd, base)
    bad = lay.clone()
    bad[1, 2, 0] = 1.5            # an albedo above 1
    wbad = wnd.clone()
    wbad[2] = 0.0                 # a rough sea without wind
    out = toa_stokes(bad, wbad, base).cpu().numpy()
    torch.cuda.synchronize()
    assert np.all(np.isnan(out[1])) and np.all(np.isnan(out[2]))
    assert np.array_equal(out[0], ok.cpu().numpy()[0])


def test_mixed_form_through_the_autograd_function():
    """precision="mixed": the mixed adjoint form behind the same Function.
    Its TOA Stokes vector is within 1e-4 relative in I and 1e-4 absolute in
    Q, U of the FP64 form's, and every gradient within 1e-4 of its scale
    (the mixed adjoint is measured at a few 1e-6 of the scale against the
    oracle's reverse mode); the device and host paths give one answer
    (rho bit for bit, gradients within 1e-9 of their scale)."""
    import torch
    layers, wind, base = _inputs(seed=5)
    w = np.array([0.7, 0.3, -0.5])
    s = torch.cuda.Stream()
    r64, gl64, gw64 = _device(layers, wind, base, w, s)
    rmx, glmx, gwmx = _device(layers, wind, base, w, s, precision="mixed")
    assert np.all(np.abs(rmx[:, 0] - r64[:, 0]) <= 1e-4 * np.abs(r64[:, 0]))
    assert np.max(np.abs(rmx[:, 1:] - r64[:, 1:])) <= 1e-4
    for i in range(len(layers)):
        for a, b in ((glmx[i], gl64[i]), (gwmx[i:i + 1], gw64[i:i + 1])):
            scale = max(np.max(np.abs(b)), 1e-300)
            assert np.max(np.abs(a - b)) <= 1e-4 * scale, (i, np.max(np.abs(a - b)), scale)
    rh, glh, gwh = _host(layers, wind, base, w, precision="mixed")
    assert np.array_equal(rh, rmx)
    for a, b in ((glh, glmx), (gwh, gwmx)):
        scale = max(np.max(np.abs(b)), 1e-300)
        assert np.max(np.abs(a - b)) <= 1e-9 * scale
