"""The allocation-free interface-kernel entry point ocrt_interface_fourier_noalloc
through the C ABI: same bits as ocrt_interface_fourier (which allocates its own
weight table and device buffers) on two calls in a row, the second smaller,
with caller-owned device inputs and outputs on torch's current stream; and
the refusal of a weight table larger than the fixed scratch array."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

VP = C.c_void_p


@pytest.fixture(scope="module")
def lib():
    import ocrt_amd
    L = ocrt_amd.library()
    L.ocrt_interface_fourier_noalloc.argtypes = [C.c_int, VP, C.c_int, VP, C.c_int, C.c_int, C.c_int, C.c_double,
                                                 C.c_int, C.c_double, C.c_int, VP, VP]
    L.ocrt_interface_fourier_noalloc.restype = C.c_int
    return L


def _nodes(n, seed):
    rng = np.random.default_rng(seed)
    return np.ascontiguousarray(np.sort(rng.uniform(0.05, 1.0, size=n)))


def _alloc_form(kind, mu_o, mu_i, m_max, n_phi, wind):
    import ocrt_amd
    names = ["R_aa", "R_ww", "T_wa", "T_aw"]
    return ocrt_amd.interface_fourier(names[kind], mu_o, mu_i, m_max, n_phi, wind, 1, 1.34, 1).reshape(-1)


def _noalloc_form(lib, kind, mu_o, mu_i, m_max, n_phi, wind):
    import torch
    dev = torch.device("cuda:0")
    d_o = torch.from_numpy(mu_o).to(dev)
    d_i = torch.from_numpy(mu_i).to(dev)
    out = torch.full(((m_max + 1) * mu_o.size * mu_i.size * 9,), float("nan"), dtype=torch.float64, device=dev)
    st = torch.cuda.current_stream()
    rc = lib.ocrt_interface_fourier_noalloc(kind, d_o.data_ptr(), mu_o.size, d_i.data_ptr(), mu_i.size, m_max,
                                            n_phi, wind, 1, 1.34, 1, out.data_ptr(), st.cuda_stream)
    assert rc == 0
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_noalloc_equals_allocating_form_twice(lib):
    # first call: T_wa, 24 x 20 nodes, modes 0..40, 1024 azimuths
    mo, mi = _nodes(24, 1), _nodes(20, 2)
    print("[gpu] noalloc call 1: T_wa 24x20, m 0..40, 1024 azimuths", flush=True)
    a = _noalloc_form(lib, 2, mo, mi, 40, 1024, 7.0)
    b = _alloc_form(2, mo, mi, 40, 1024, 7.0)
    assert np.array_equal(a, b)
    # second call, smaller in every size and another kind: the table it leaves
    # behind is the first call's prefix region only, rewritten before use
    mo2, mi2 = _nodes(8, 3), _nodes(6, 4)
    print("[gpu] noalloc call 2: R_aa 8x6, m 0..10, 512 azimuths", flush=True)
    a2 = _noalloc_form(lib, 0, mo2, mi2, 10, 512, 4.0)
    b2 = _alloc_form(0, mo2, mi2, 10, 512, 4.0)
    assert np.array_equal(a2, b2)
    # the largest table the scratch array holds
    print("[gpu] noalloc call 3: T_aw 4x4, m 0..255, 1024 azimuths", flush=True)
    a3 = _noalloc_form(lib, 3, mo2[:4], mi2[:4], 255, 1024, 9.0)
    b3 = _alloc_form(3, mo2[:4], mi2[:4], 255, 1024, 9.0)
    assert np.array_equal(a3, b3)


def test_noalloc_refuses_oversized_table(lib):
    import torch
    x = torch.ones(16, dtype=torch.float64, device="cuda:0")
    out = torch.zeros(16 * 9 * 257, dtype=torch.float64, device="cuda:0")
    # 256 rows x 2048 azimuths > 2^18 entries: refused before any launch
    rc = lib.ocrt_interface_fourier_noalloc(0, x.data_ptr(), 2, x.data_ptr(), 2, 255, 2048, 5.0, 1, 1.34, 1,
                                            out.data_ptr(), None)
    assert rc == -1
    assert torch.count_nonzero(out).item() == 0
