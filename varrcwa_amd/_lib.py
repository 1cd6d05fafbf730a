"""ctypes binding of ``libvarrcwa.so`` (the C ABI declared in ``include/varrcwa.h``).

The library is loaded from this package directory only; a missing or stale build
raises immediately -- there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libvarrcwa.so"

_lib = None

c_int_p = C.POINTER(C.c_int)
c_double_p = C.POINTER(C.c_double)
RULE_FN = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)


class VrcLayer(C.Structure):
    """``vrc_layer``: one Layer(x, y, eps) slab (host pointers)."""
    _fields_ = [("ncoord_x", C.c_int), ("x", c_double_p), ("ncoord_y", C.c_int), ("y", c_double_p),
                ("eps", c_double_p)]


class VrcStackInstance(C.Structure):
    """``vrc_stack_instance``: one instance of a batched layer-stack solve."""
    _fields_ = [("lam", C.c_double), ("nlayers", C.c_int), ("layers", C.POINTER(VrcLayer)),
                ("thickness", c_double_p), ("bottom", C.POINTER(VrcLayer)), ("top", C.POINTER(VrcLayer))]


class VrcRegion(C.Structure):
    """``vrc_region``: LayerSampler's Lx_, Ly_ and absorbing-region boundaries."""
    _fields_ = [(k, C.c_double) for k in ("Lx", "Ly", "b0", "b1", "u0", "u1", "l0", "l1", "r0", "r1")]


# vrc_sample_host_fn / vrc_sample_device_fn of the callback sampler
SAMPLE_HOST_FN = C.CFUNCTYPE(C.c_int, C.c_double, c_double_p, c_double_p, C.c_void_p)
SAMPLE_DEVICE_FN = C.CFUNCTYPE(C.c_int, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p)


def _sig(name, restype, *args):
    f = getattr(_lib, name)
    f.restype = restype
    f.argtypes = list(args)


def load():
    """Load the in-tree library (raises FileNotFoundError when it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise FileNotFoundError(f"{LIB_PATH} is not built; run `python -m varrcwa_amd.build`")
    # The process's HIP runtime is the one torch brings up: initialize it before the library
    # registers its kernels (loading the library first made the first context creation fail
    # with a HIP error in a process that had not touched torch.cuda yet).
    try:
        import torch
        torch.cuda.is_available()
    except ImportError:  # pragma: no cover - torch is required for device memory
        pass
    _lib = C.CDLL(str(LIB_PATH))
    vp, i, l, d = C.c_void_p, C.c_int, C.c_long, C.c_double
    _sig("vrc_version", C.c_char_p)
    _sig("vrc_ctx_create", i, i, vp, C.POINTER(vp))
    _sig("vrc_ctx_destroy", i, vp)
    _sig("vrc_last_error", C.c_char_p, vp)
    _sig("vrc_sync", i, vp)
    _sig("vrc_pool_bytes", C.c_longlong, vp)
    _sig("vrc_pool_trim", i, vp)
    _sig("vrc_timing_enable", i, vp, i)
    _sig("vrc_tune_set", i, C.c_char_p, C.c_longlong)
    _sig("vrc_timing_read", i, vp, c_double_p, c_int_p, i)
    _sig("vrc_qr_stats", i, vp, C.POINTER(C.c_longlong))
    _sig("vrc_zgemm", i, vp, i, i, i, i, i, c_double_p, vp, l, l, vp, l, l, c_double_p, vp, l, l, i)
    _sig("vrc_zgetrf", i, vp, i, vp, l, l, vp, l, vp, i)
    _sig("vrc_zgetrs", i, vp, i, i, vp, l, l, vp, l, vp, l, l, vp, l, l, i)
    _sig("vrc_zgeev", i, vp, i, vp, l, l, vp, l, vp, l, l, vp, i)
    _optional = {
        "vrc_assemble_toeplitz": (i, vp, i, i, i, C.POINTER(VrcLayer), vp, vp, vp, l),
        "vrc_assemble_pq": (i, vp, i, i, i, C.POINTER(VrcLayer), d, i, vp, vp, vp, l),
        "vrc_eig_pq": (i, vp, i, i, vp, vp, l, vp, vp, vp, vp),
        "vrc_layer_smatrix": (i, vp, i, i, c_double_p, c_double_p, vp, vp, vp, vp, vp, l, vp, vp),
        "vrc_star_product": (i, vp, i, i, vp, vp, vp, vp, vp, vp, vp, vp, l),
        "vrc_star_product_noalloc": (i, vp, i, i, vp, vp, vp, vp, vp, vp, vp, vp, l),
        "vrc_project_left": (i, vp, i, i, vp, vp, vp, vp, vp, vp, vp, vp, l),
        "vrc_redheffer_segment": (i, vp, i, d, d, d, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                  vp, vp, vp, c_double_p),
        "vrc_var_sampler_create": (i, vp, d, i, i, C.POINTER(VrcLayer), i, C.POINTER(vp)),
        "vrc_callback_sampler_create": (i, vp, d, i, i, C.POINTER(VrcRegion), i, SAMPLE_HOST_FN, SAMPLE_DEVICE_FN,
                                        vp, C.POINTER(vp)),
        "vrc_sampler_region": (i, vp, C.POINTER(VrcRegion)),
        "vrc_stack_sampler_create": (i, vp, d, i, i, i, C.POINTER(VrcLayer), c_double_p,
                                     C.POINTER(VrcLayer), C.POINTER(VrcLayer), i, C.POINTER(vp)),
        "vrc_sampler_add_rule": (i, vp, i, i, RULE_FN, vp),
        "vrc_sampler_destroy": (i, vp),
        "vrc_sampler_ndim": (i, vp),
        "vrc_sampler_sample": (i, vp, d, c_double_p, c_double_p),
        "vrc_sampler_sample_gpu": (i, vp, d, vp, vp),
        "vrc_sampler_k": (i, vp, c_double_p, c_double_p),
        "vrc_integrator_create": (i, vp, i, d, d, vp, C.POINTER(vp)),
        "vrc_integrator_destroy": (i, vp),
        "vrc_rcwa_compute": (i, vp, i),
        "vrc_rcwa_compute_edges": (i, vp, i, c_double_p),
        "vrc_redheffer_compute": (i, vp, i, i, d),
        "vrc_redheffer_compute_many": (i, i, C.POINTER(vp), i, i, d),
        "vrc_differential_compute": (i, vp, i),
        "vrc_integrator_get": (i, vp, i, c_double_p),
        "vrc_integrator_counts": (i, vp, c_int_p, c_int_p),
        "vrc_integrator_tree": (i, vp, i, c_double_p),
        "vrc_batch_stack_rcwa": (i, vp, i, i, i, C.POINTER(VrcStackInstance), i, vp, vp, vp, vp, vp, vp, vp, c_int_p),
        "vrc_batch_stack_rcwa_mixed": (i, vp, i, c_int_p, c_int_p, C.POINTER(VrcStackInstance), i, vp, vp, vp, vp,
                                       vp, vp, vp, c_int_p),
        "vrc_batch_plan_create": (i, vp, i, i, i, C.POINTER(VrcStackInstance), i, vp, vp, vp, vp, vp, vp, vp,
                                  C.POINTER(vp)),
        "vrc_batch_plan_launch": (i, vp, c_int_p),
        "vrc_autotune_stack": (i, vp, i, i, i, C.POINTER(VrcStackInstance), d, C.c_char_p, C.c_char_p, i),
        "vrc_autotune_load": (i, C.c_char_p),
        "vrc_autotune_clear": (i,),
        "vrc_batch_plan_destroy": (i, vp),
        "vrc_stack_efficiencies": (i, vp, i, i, i, C.POINTER(VrcStackInstance), vp, vp, vp, vp, vp, vp),
        "vrc_wave_eqn_derivatives": (i, vp, i, i, C.POINTER(VrcLayer), d, i, C.POINTER(VrcLayer), i, vp, vp, i, vp,
                                     vp, vp, vp, vp, vp, vp),
        "vrc_gds_sampler_create": (i, vp, d, i, i, C.c_char_p, vp, vp, vp, vp, d, C.POINTER(vp)),
        "vrc_gds_sampler_bounds": (i, vp, vp),
        "vrc_gds_polygons": (i, C.c_char_p, d, i, i, c_int_p, vp, i, vp),
        "vrc_sampler_layout": (i, vp, d, i, vp, vp, i, vp, vp, vp),
        "vrc_rcwa_solver_create": (i, vp, i, i, C.POINTER(vp)),
        "vrc_rcwa_solver_add_layer": (i, vp, C.POINTER(VrcLayer)),
        "vrc_rcwa_solver_solve": (i, vp, d, i, vp, vp, vp),
        "vrc_rcwa_solver_get": (i, vp, i, i, vp),
        "vrc_rcwa_solver_intermediate_field": (i, vp, i, c_double_p, c_double_p, c_double_p),
        "vrc_rcwa_solver_dims": (i, vp, c_int_p, c_int_p),
        "vrc_rcwa_solver_destroy": (i, vp),
        "vrc_modal_power": (i, vp, i, i, i, vp, C.c_longlong, vp, C.c_longlong, d, d, i, i, d, d, d, d, vp),
        "vrc_find_guided_modes": (i, vp, i, vp, d, d, d, d, vp, vp, vp),
        "vrc_batch_stack_forward": (i, vp, i, i, i, C.POINTER(VrcStackInstance), i, vp, vp, vp, vp, vp, vp, vp,
                                    C.POINTER(vp), c_int_p),
        "vrc_tape_cells": (C.c_longlong, vp),
        "vrc_tape_bytes": (C.c_longlong, vp),
        "vrc_batch_stack_forward_ex": (i, vp, i, i, i, C.POINTER(VrcStackInstance), i, vp, vp, vp, vp, vp, vp, vp,
                                       i, i, C.POINTER(vp), c_int_p),
        "vrc_batch_stack_vjp_shape": (i, vp, vp, c_int_p, c_int_p, c_int_p, c_double_p, c_double_p, vp, i),
        "vrc_batch_stack_vjp": (i, vp, vp, vp, i),
        "vrc_tape_destroy": (i, vp),
        "vrc_assemble_toeplitz_vjp": (i, vp, i, i, i, C.POINTER(VrcLayer), vp, vp, vp, l, vp),
        "vrc_assemble_pq_vjp": (i, vp, i, i, i, C.POINTER(VrcLayer), d, C.POINTER(VrcLayer), vp, vp, l, vp),
        "vrc_eig_pq_vjp": (i, vp, i, i, vp, vp, l, vp, vp, vp, vp, vp, vp, vp, vp),
        "vrc_layer_smatrix_vjp": (i, vp, i, i, c_double_p, c_double_p, vp, vp, vp, vp, vp, l, vp, vp, vp, vp, vp),
        "vrc_star_product_vjp": (i, vp, i, i, vp, vp, vp, vp, vp, vp, vp, vp, l, vp, vp, vp, vp, vp, vp, vp, vp),
        "vrc_project_left_vjp": (i, vp, i, i, vp, vp, vp, vp, vp, vp, vp, vp, l, vp, vp, vp, vp),
    }
    for name, (res, *args) in _optional.items():
        if hasattr(_lib, name):
            _sig(name, res, *args)
    return _lib


def exported_symbols():
    """Names of the vrc_* functions the loaded library exports."""
    lib = load()
    import re
    names = []
    hdr = (Path(__file__).resolve().parent.parent / "include" / "varrcwa.h").read_text()
    for m in re.finditer(r"\b(vrc_[a-z0-9_]+)\s*\(", hdr):
        names.append(m.group(1))
    return sorted(set(n for n in names if hasattr(lib, n))), sorted(set(names))


class VrcError(RuntimeError):
    pass


class Context:
    """Owns a ``vrc_ctx`` bound to a HIP stream (torch's current stream by default)."""

    def __init__(self, device: int = 0, stream=None):
        self.lib = load()
        import torch
        if stream is None:
            stream = torch.cuda.current_stream(device).cuda_stream
        self.torch = torch
        self.device = device
        h = C.c_void_p()
        rc = self.lib.vrc_ctx_create(device, C.c_void_p(stream), C.byref(h))
        if rc != 0:
            raise VrcError(f"vrc_ctx_create failed ({rc})")
        self.h = h

    def check(self, rc: int, what: str):
        if rc != 0:
            msg = self.lib.vrc_last_error(self.h).decode()
            raise VrcError(f"{what} failed ({rc}): {msg}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.vrc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        self.check(self.lib.vrc_sync(self.h), "vrc_sync")


def ptr(t) -> C.c_void_p:
    """Device pointer of a contiguous torch tensor."""
    assert t.is_contiguous()
    return C.c_void_p(t.data_ptr())


def cplx(z) -> "C.Array":
    z = complex(z)
    return (C.c_double * 2)(z.real, z.imag)


def to_dev_colmajor(a: np.ndarray, device="cuda"):
    """(..., m, n) numpy complex -> torch tensor whose memory is column-major per matrix."""
    import torch
    a = np.asarray(a, dtype=np.complex128)
    return torch.from_numpy(np.ascontiguousarray(np.swapaxes(a, -1, -2))).to(device)


def from_dev_colmajor(t) -> np.ndarray:
    return np.swapaxes(t.detach().cpu().numpy(), -1, -2).copy()
