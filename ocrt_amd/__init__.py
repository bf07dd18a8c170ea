"""MI355X-native successive-orders-of-scattering solver (OCRT atmospheric path).

The compute path is the HIP library libocrt_amd.so built by ocrt_amd.build;
this module only loads it and marshals descriptors.  There is no CPU fallback:
every entry point raises when the library or a GPU is missing.
"""
import ctypes as C
import os

import numpy as np

from .abi import (Instance, Result, RESULT_FIELDS, PREC_FP64, PREC_MIXED, make_instance,
                  instance_array, results_to_dict)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libocrt_amd.so")
_lib = None

PRECISIONS = {"fp64": PREC_FP64, "mixed": PREC_MIXED}
PLAN_ONE_PER_LAUNCH = 1
PLAN_ADJOINT = 2
PLAN_CKPT_SPARSE = 4
PLAN_AUTOTUNE = 8
PLAN_GRAPH = 16
PLAN_DEVICE_INPUTS = 32


class OcrtError(RuntimeError):
    pass


def library():
    """Load libocrt_amd.so (built in-tree by ocrt_amd.build)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OcrtError(f"{LIB_PATH} is missing: run `python -m ocrt_amd.build` first")
        # One HIP runtime per process: torch ships its own libamdhip64 with the
        # same soname (libamdhip64.so.7).  Loading torch first makes the solver
        # library bind to that runtime; loading it first would let torch open a
        # second runtime that no longer sees the device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = C.CDLL(LIB_PATH)
        lib.ocrt_plan_create.argtypes = [C.POINTER(Instance), C.c_int, C.c_int, C.c_int,
                                         C.POINTER(C.c_void_p)]
        lib.ocrt_plan_execute.argtypes = [C.c_void_p, C.c_void_p]
        lib.ocrt_plan_results.argtypes = [C.c_void_p, C.POINTER(Result), C.c_void_p, C.c_void_p]
        lib.ocrt_plan_mmax.argtypes = [C.c_void_p]
        lib.ocrt_plan_device_bytes.argtypes = [C.c_void_p]
        lib.ocrt_plan_device_bytes.restype = C.c_size_t
        lib.ocrt_plan_work.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                       C.POINTER(C.c_double)]
        lib.ocrt_plan_profile.argtypes = [C.c_void_p, C.c_int]
        lib.ocrt_plan_kernel_times.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int),
                                               C.POINTER(C.c_double), C.POINTER(C.c_int)]
        lib.ocrt_plan_destroy.argtypes = [C.c_void_p]
        lib.ocrt_solve_batch.argtypes = [C.POINTER(Instance), C.c_int, C.c_int, C.POINTER(Result),
                                         C.c_void_p]
        lib.ocrt_version.restype = C.c_char_p
        lib.ocrt_plan_set_lut.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p]
        lib.ocrt_plan_lut_results.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        lib.ocrt_plan_create_ex.argtypes = [C.POINTER(Instance), C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p)]
        lib.ocrt_plan_adjoint.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.ocrt_plan_set_inputs.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.ocrt_plan_rho.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.ocrt_plan_adjoint_layers.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.ocrt_adjoint_batch.argtypes = [C.POINTER(Instance), C.c_int, C.c_int, C.c_void_p, C.POINTER(Result),
                                           C.c_void_p, C.c_void_p, C.c_void_p]
        lib.ocrt_gas_create.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_void_p, C.POINTER(C.c_void_p)]
        lib.ocrt_gas_tau.argtypes = [C.c_void_p, C.c_int, C.c_double]
        lib.ocrt_gas_tau.restype = C.c_double
        lib.ocrt_gas_destroy.argtypes = [C.c_void_p]
        lib.ocrt_gas_destroy.restype = None
        lib.ocrt_plan_lut_extras.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        lib.ocrt_plan_tuning.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.ocrt_plan_set_tuning.argtypes = [C.c_void_p, C.c_int, C.c_int]
        lib.ocrt_plan_set_gemm_shape.argtypes = [C.c_void_p, C.c_int, C.c_int]
        lib.ocrt_plan_gemm_tuning.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        lib.ocrt_plan_set_reduced_threshold.argtypes = [C.c_void_p, C.c_double]
        lib.ocrt_plan_gemm_issued.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.ocrt_plan_residuals.argtypes = [C.c_void_p, C.c_void_p]
        lib.ocrt_plan_export_nodes.argtypes = [C.c_void_p]
        lib.ocrt_plan_node_fields.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        lib.ocrt_interface_fourier.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                               C.c_double, C.c_int, C.c_double, C.c_int, C.c_void_p, C.c_void_p]
        lib.ocrt_interface_direct.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_double,
                                              C.c_void_p, C.c_void_p]
        _lib = lib
    return _lib


def _require_gpu():
    import torch
    if not torch.cuda.is_available():
        raise OcrtError("the OCRT MI355X solver needs a ROCm GPU; none is visible")


INTERFACE_KINDS = {"R_aa": 0, "R_ww": 1, "T_wa": 2, "T_aw": 3}


def interface_fourier(kind, mu_o, mu_i, m_max, n_phi=1024, wind=5.0, sigma_type=1, n_water=1.34, q_convention=1):
    """Azimuthal Fourier kernels of a rough-interface Mueller matrix
    (ocrt_interface_fourier) for modes 0..m_max: array [m_max+1, n_o, n_i, 3, 3].
    kind: "R_aa", "R_ww", "T_wa" or "T_aw" (see include/ocrt_amd.h for the
    meaning of mu_o and mu_i in each)."""
    _require_gpu()
    lib = library()
    mo = np.ascontiguousarray(mu_o, dtype=np.float64)
    mi = np.ascontiguousarray(mu_i, dtype=np.float64)
    out = np.zeros((m_max + 1, mo.size, mi.size, 3, 3))
    rc = lib.ocrt_interface_fourier(INTERFACE_KINDS[kind], mo.ctypes.data, mo.size, mi.ctypes.data, mi.size, m_max,
                                    n_phi, wind, sigma_type, n_water, q_convention, out.ctypes.data, None)
    if rc != 0:
        raise OcrtError(f"ocrt_interface_fourier failed with code {rc}")
    return out


def interface_direct(kind, mu, wind=5.0, sigma_type=1, n_water=1.34):
    """Slope-integrated direct terms (ocrt_interface_direct): "T_aw" the air ->
    water transmittance of a beam at cosine mu, "R_ww" the water-side
    reflectance of an upwelling beam."""
    _require_gpu()
    lib = library()
    m = np.ascontiguousarray(mu, dtype=np.float64)
    out = np.zeros(m.size)
    rc = lib.ocrt_interface_direct(INTERFACE_KINDS[kind], m.ctypes.data, m.size, wind, sigma_type, n_water,
                                   out.ctypes.data, None)
    if rc != 0:
        raise OcrtError(f"ocrt_interface_direct failed with code {rc}")
    return out


PROFILES = {"usstd76": 0, "tropical": 1, "mlsumm": 2, "mlwint": 3, "sasumm": 4, "sawint": 5, "userdef": 6}


class GasTable:
    """Gas absorption table (ocrt_gas_create): an AFGL profile with layered
    cross sections read from afgl_dir / xsec_dir in the reference's formats.
    `columns` optionally overrides the total column of gases by name
    (h2o, o3, o2, co2, no2, ch4; molecules cm^-2).  Pass it as `gas=` to
    ocrt_amd.abi.make_instance."""

    def __init__(self, afgl_dir, xsec_dir, profile="usstd76", columns=None):
        self.lib = library()
        ov = (C.c_double * 6)(*[-1.0] * 6)
        for k, v in (columns or {}).items():
            ov[("h2o", "o3", "o2", "co2", "no2", "ch4").index(k)] = float(v)
        h = C.c_void_p()
        rc = self.lib.ocrt_gas_create(afgl_dir.encode(), xsec_dir.encode(), PROFILES.get(profile, profile), ov,
                                      C.byref(h))
        if rc != 0:
            raise OcrtError(f"ocrt_gas_create failed with code {rc}")
        self.handle = h.value

    def tau(self, wavelength_nm, gas=-1):
        """Column absorption optical depth of one gas (index) or of all (-1)."""
        return self.lib.ocrt_gas_tau(self.handle, gas, wavelength_nm)

    def __del__(self):
        if getattr(self, "handle", None):
            self.lib.ocrt_gas_destroy(self.handle)
            self.handle = None


class Plan:
    """A solve plan over a list of descriptors (ocrt_amd.abi.Instance).

    precision: "fp64" (plain parity form) or "mixed" (bf16 MFMA scattering
    integral for orders >= 2).  one_per_launch=True runs instances one after
    another (the plain one-instance-per-launch form)."""

    def __init__(self, descs, precision="mixed", one_per_launch=False, adjoint=False, ckpt_sparse=False,
                 autotune=False, graph=False, device_inputs=False):
        _require_gpu()
        self.lib = library()
        self.n = len(descs)
        self._arr = instance_array(descs)
        self._handle = C.c_void_p()
        self.n_layers = [d.n_layers for d in descs]
        flags = ((PLAN_ONE_PER_LAUNCH if one_per_launch else 0) | (PLAN_ADJOINT if adjoint else 0) |
                 (PLAN_CKPT_SPARSE if ckpt_sparse else 0) | (PLAN_AUTOTUNE if autotune else 0) |
                 (PLAN_GRAPH if graph else 0) | (PLAN_DEVICE_INPUTS if device_inputs else 0))
        rc = self.lib.ocrt_plan_create_ex(self._arr, self.n, PRECISIONS[precision], flags, C.byref(self._handle))
        if rc != 0:
            raise OcrtError(f"ocrt_plan_create failed with code {rc}")
        self.mmax = self.lib.ocrt_plan_mmax(self._handle)

    def set_tuning(self, unroll, waves):
        """Select the mixed-form sweep's launch shape by hand."""
        rc = self.lib.ocrt_plan_set_tuning(self._handle, int(unroll), int(waves))
        if rc != 0:
            raise OcrtError(f"no launch shape unroll {unroll} waves {waves}")

    def tuning(self):
        """(levels per block, wavefronts per workgroup) of the mixed-form sweep
        and, after autotuning, the summed order-loop ms of every shape tried."""
        u, w, n = C.c_int(), C.c_int(), C.c_int()
        ms = (C.c_double * 16)()
        self.lib.ocrt_plan_tuning(self._handle, C.byref(u), C.byref(w), ms, C.byref(n))
        return {"unroll": u.value, "waves": w.value, "ms_per_shape": list(ms[: n.value])}

    GEMM_SHAPES = ((64, 128), (64, 64), (128, 64), (128, 128))   # operator rows x levels per tile

    def set_gemm_shape(self, op_class, shape):
        """Tile shape (index into GEMM_SHAPES) of one operator class of the
        mixed form's scattering integral (0: one part, single field; 1: one
        part, field pair; 2: two parts)."""
        rc = self.lib.ocrt_plan_set_gemm_shape(self._handle, int(op_class), int(shape))
        if rc != 0:
            raise OcrtError(f"no GEMM class {op_class} / shape {shape}")

    def gemm_tuning(self):
        """Shape chosen per operator class and, after autotuning, the class's
        GEMM ms per execute with every shape ([class][shape])."""
        sh = (C.c_int * 3)()
        ns = len(self.GEMM_SHAPES)
        ms = (C.c_double * (3 * ns))()
        n = self.lib.ocrt_plan_gemm_tuning(self._handle, sh, ms)
        return {"shape_of_class": [self.GEMM_SHAPES[k] for k in sh],
                "ms": [[ms[ns * k + j] for j in range(ns)] for k in range(3)] if n else None}

    def set_reduced_threshold(self, tau):
        """Estimated remaining share of an item's order sum below which its
        next order takes the reduced (bf16 operator and field) scattering
        integral; 0: never."""
        rc = self.lib.ocrt_plan_set_reduced_threshold(self._handle, float(tau))
        if rc != 0:
            raise OcrtError(f"reduced threshold {tau} refused (mixed plans, 0 <= tau <= 1)")

    def gemm_issued(self):
        """From the last results(): MFMA flops the scattering integral
        launched, item-orders that ran it, and those that ran reduced."""
        f, o, r = C.c_double(), C.c_double(), C.c_double()
        self.lib.ocrt_plan_gemm_issued(self._handle, C.byref(f), C.byref(o), C.byref(r))
        return {"issued_flops": f.value, "item_orders": o.value, "reduced_orders": r.value}

    def set_lut(self, vza_deg, raa_deg):
        """Sample every following execute on the grid vza x raa (full-grid / LUT mode)."""
        self._lut_vza = np.ascontiguousarray(vza_deg, dtype=np.float64)
        self._lut_raa = np.ascontiguousarray(raa_deg, dtype=np.float64)
        rc = self.lib.ocrt_plan_set_lut(self._handle, len(self._lut_vza), self._lut_vza.ctypes.data_as(C.c_void_p),
                                        len(self._lut_raa), self._lut_raa.ctypes.data_as(C.c_void_p))
        if rc != 0:
            raise OcrtError(f"ocrt_plan_set_lut failed with code {rc}")

    def lut(self):
        """LUT of the last execute: arrays [n_inst, n_vza, n_raa] per column and the
        per-instance T_dir_dn, T_diff_dn_hemi."""
        nv, nr = len(self._lut_vza), len(self._lut_raa)
        grid = np.zeros((self.n, nv, nr, 6))
        scal = np.zeros((self.n, 2))
        rc = self.lib.ocrt_plan_lut_results(self._handle, grid.ctypes.data_as(C.c_void_p),
                                            scal.ctypes.data_as(C.c_void_p))
        if rc != 0:
            raise OcrtError(f"ocrt_plan_lut_results failed with code {rc}")
        names = ("rho_I", "rho_Q", "rho_U", "T_diff_dn_dir", "T_sg_up_dir", "T_total_up_dir")
        out = {k: grid[..., q] for q, k in enumerate(names)}
        out["T_dir_dn"] = scal[:, 0]
        out["T_diff_dn_hemi"] = scal[:, 1]
        return out

    def adjoint(self, grad_out, stream=None):
        """Gradient of sum_i grad_out[i] . (rho_I, rho_Q, rho_U)_i (adjoint plans).
        Returns per-instance arrays d/dh, d/dxdel, d/dydel over the levels and
        d/dwind."""
        g = np.ascontiguousarray(grad_out, dtype=np.float64).reshape(self.n, 3)
        total = sum(3 * (nt + 1) for nt in self.n_layers)
        lev = np.zeros(total)
        wind = np.zeros(self.n)
        ptr = None
        if stream is not None:
            ptr = C.c_void_p(stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream))
        rc = self.lib.ocrt_plan_adjoint(self._handle, g.ctypes.data_as(C.c_void_p), lev.ctypes.data_as(C.c_void_p),
                                        wind.ctypes.data_as(C.c_void_p), ptr)
        if rc != 0:
            raise OcrtError(f"ocrt_plan_adjoint failed with code {rc}")
        out = {"grad_h": [], "grad_x": [], "grad_y": [], "grad_wind": wind}
        off = 0
        for nt in self.n_layers:
            L = nt + 1
            out["grad_h"].append(lev[off:off + L].copy())
            out["grad_x"].append(lev[off + L:off + 2 * L].copy())
            out["grad_y"].append(lev[off + 2 * L:off + 3 * L].copy())
            off += 3 * L
        return out

    @staticmethod
    def _stream_ptr(stream):
        if stream is None:
            return None
        return C.c_void_p(stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream))

    def set_inputs(self, d_layers=0, d_wind=0, stream=None):
        """Replace the layers (device address of the packed [4][n_layers]
        blocks) and/or the wind speeds (device address of [n]) from device
        memory, on `stream` (ocrt_plan_set_inputs)."""
        rc = self.lib.ocrt_plan_set_inputs(self._handle, C.c_void_p(d_layers or None), C.c_void_p(d_wind or None),
                                           self._stream_ptr(stream))
        if rc != 0:
            raise OcrtError(f"ocrt_plan_set_inputs failed with code {rc}")

    def rho_device(self, d_rho, d_status=0, stream=None):
        """(rho_I, rho_Q, rho_U) [n][3] of the last execute into device memory
        (ocrt_plan_rho)."""
        rc = self.lib.ocrt_plan_rho(self._handle, C.c_void_p(d_rho), C.c_void_p(d_status or None),
                                    self._stream_ptr(stream))
        if rc != 0:
            raise OcrtError(f"ocrt_plan_rho failed with code {rc}")

    def adjoint_layers(self, d_grad_out, d_grad_layers=0, d_grad_wind=0, stream=None):
        """Adjoint with device-memory operands and the per-layer chain rule on
        the device (ocrt_plan_adjoint_layers)."""
        rc = self.lib.ocrt_plan_adjoint_layers(self._handle, C.c_void_p(d_grad_out),
                                               C.c_void_p(d_grad_layers or None), C.c_void_p(d_grad_wind or None),
                                               self._stream_ptr(stream))
        if rc != 0:
            raise OcrtError(f"ocrt_plan_adjoint_layers failed with code {rc}")

    def export_nodes(self):
        """Keep the per-node TOA and BOA Fourier fields of the following
        executes (ocrt_plan_export_nodes)."""
        rc = self.lib.ocrt_plan_export_nodes(self._handle)
        if rc != 0:
            raise OcrtError(f"ocrt_plan_export_nodes failed with code {rc}")

    def node_fields(self, ring_pitch):
        """(toa_up, boa_dn) of the last execute, each [n_inst, m_max_all + 1, 3,
        ring_pitch] (ocrt_plan_node_fields)."""
        shape = (self.n, self.mmax + 1, 3, ring_pitch)
        up = np.zeros(shape)
        dn = np.zeros(shape)
        rc = self.lib.ocrt_plan_node_fields(self._handle, ring_pitch, up.ctypes.data, dn.ctypes.data)
        if rc != 0:
            raise OcrtError(f"ocrt_plan_node_fields failed with code {rc}")
        return up, dn

    def execute(self, stream=None):
        ptr = None
        if stream is not None:
            ptr = C.c_void_p(stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream))
        rc = self.lib.ocrt_plan_execute(self._handle, ptr)
        if rc != 0:
            raise OcrtError(f"ocrt_plan_execute failed with code {rc}")

    def results(self, modes=False):
        res = (Result * self.n)()
        M = self.mmax + 1
        tm = np.zeros((self.n, M, 3)) if modes else None
        orders = np.zeros((self.n, M), dtype=np.int32) if modes else None
        rc = self.lib.ocrt_plan_results(
            self._handle, res,
            tm.ctypes.data_as(C.c_void_p) if modes else None,
            orders.ctypes.data_as(C.c_void_p) if modes else None)
        if rc != 0:
            raise OcrtError(f"ocrt_plan_results failed with code {rc}")
        out = results_to_dict(res)
        if modes:
            out["toa_modes"] = tm
            out["orders_per_m"] = orders
        return out

    def work(self):
        f, b, l = C.c_double(), C.c_double(), C.c_double()
        self.lib.ocrt_plan_work(self._handle, C.byref(f), C.byref(b), C.byref(l))
        return {"gemm_flops": f.value, "sweep_bytes": b.value, "order_steps": l.value}

    def profile(self, enable=True):
        self.lib.ocrt_plan_profile(self._handle, 1 if enable else 0)

    def kernel_times(self):
        gm, sm = C.c_double(), C.c_double()
        gl, sl = C.c_int(), C.c_int()
        rc = self.lib.ocrt_plan_kernel_times(self._handle, C.byref(gm), C.byref(gl), C.byref(sm), C.byref(sl))
        if rc != 0:
            raise OcrtError(f"ocrt_plan_kernel_times failed with code {rc}")
        return {"gemm_ms": gm.value, "gemm_launches": gl.value, "sweep_ms": sm.value,
                "sweep_launches": sl.value}

    def device_bytes(self):
        return self.lib.ocrt_plan_device_bytes(self._handle)

    def close(self):
        if self._handle:
            self.lib.ocrt_plan_destroy(self._handle)
            self._handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve(descs, precision="mixed", modes=False, one_per_launch=False):
    """Solve a batch of instances on the GPU; returns a dict of numpy arrays
    keyed by the result fields of include/ocrt_amd.h."""
    plan = Plan(descs, precision=precision, one_per_launch=one_per_launch)
    plan.execute()
    out = plan.results(modes=modes)
    plan.close()
    return out


__all__ = ["Plan", "solve", "library", "make_instance", "OcrtError", "PREC_FP64", "PREC_MIXED"]
