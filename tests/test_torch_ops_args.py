"""Argument checks of the torch boundary that run before any solve."""
import pytest


def test_unknown_precision_is_rejected():
    torch = pytest.importorskip("torch")
    from ocrt_amd.torch_ops import toa_stokes
    with pytest.raises(ValueError, match="precision"):
        toa_stokes(torch.zeros(1, 4, 3, dtype=torch.float64), torch.zeros(1, dtype=torch.float64), [{}],
                   device_io=False, precision="bf16")
