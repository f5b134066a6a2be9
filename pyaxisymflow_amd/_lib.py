"""ctypes binding of libaxisym_hip.so and the array plumbing shared by the
drop-in modules.

Arrays handed to the drop-in functions may be
  * numpy float64 C-contiguous arrays (copied to the GPU, computed, and the
    outputs copied back into the same array objects), or
  * torch float64 contiguous tensors on a ROCm device (used in place; the
    work is queued on torch's current stream).
Wrong dtype or layout raises TypeError (as the reference's pybind11
`noconvert` and numba typing do); shape mismatches raise ValueError; a
missing library or a HIP failure raises RuntimeError.  There is no CPU
fallback: without the HIP library and a GPU every call fails loudly.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libaxisym_hip.so")

AXS_OK = 0
AXS_ERR_INVALID = 1
AXS_ERR_HIP = 2
AXS_ERR_UNSUPPORTED = 3
AXS_ERR_ALLOC = 4

_P = ctypes.c_void_p
_I = ctypes.c_int
_L = ctypes.c_int64
_D = ctypes.c_double

class RigidConfig(ctypes.Structure):
    """axs_rigid_config (include/axisym.h)."""

    _fields_ = [(n, ctypes.c_int) for n in (
        "nr", "nz", "kill_width", "history", "precision", "moving_body", "average_psi",
        "implicit_diffusion")] + [
        (n, ctypes.c_double) for n in (
            "dx", "nu", "brink_lam", "U0", "T_ramp", "CFL", "eps", "dt_diffusive", "osc_amp",
            "osc_omega", "drag_scale", "drag_div", "dt_cap", "period", "t_end", "body_z0",
            "body_r0", "body_aspect2", "body_radius", "body_disp", "moll_zone")]


class RigidBuffers(ctypes.Structure):
    """axs_rigid_buffers (include/axisym.h)."""

    _fields_ = [(n, ctypes.c_void_p) for n in (
        "vorticity", "psi", "u_z", "u_r", "u_pen_z", "u_pen_r", "char_func", "work0", "work1",
        "avg_psi", "scalars", "drag", "r", "z", "kill_wz", "kill_wr")]


class SoftConfig(ctypes.Structure):
    """axs_soft_config (include/axisym.h)."""

    _fields_ = [(n, ctypes.c_int) for n in (
        "nr", "nz", "kill_width", "lsq_sweeps", "reinit_iterations", "lsq_grow")] + [
        (n, ctypes.c_double) for n in (
            "dx", "nu", "G", "rho_f", "CFL", "eps", "r_ball", "z_cm", "r_cm", "moll_zone",
            "extrap_zone", "reinit_band")] + [("streaming", ctypes.c_int)] + [
        (n, ctypes.c_double) for n in (
            "brink_lam", "U0", "omega", "e_amp", "tether_rad", "freq_limit", "t_end")]


class SoftBuffers(ctypes.Structure):
    """axs_soft_buffers (include/axisym.h)."""

    _fields_ = [(n, ctypes.c_void_p) for n in (
        "vorticity", "psi", "u_z", "u_r", "eta1", "eta2", "phi", "char_func", "scalars", "r", "z",
        "kill_wz", "kill_wr")]


class SoftBuffersF32(ctypes.Structure):
    """axs_soft_buffers_f32 (include/axisym.h): the same fields, the flow
    fields (vorticity, psi, u_z, u_r, char_func) pointing at float32 data."""

    _fields_ = SoftBuffers._fields_


class SlabConfig(ctypes.Structure):
    """axs_slab_config (include/axisym.h)."""

    _fields_ = [(n, ctypes.c_int) for n in ("nr", "nz", "ghost", "kill_width")] + [
        (n, ctypes.c_double) for n in ("dx", "nu", "lam", "U0", "omega", "CFL", "eps")]


class SlabBuffers(ctypes.Structure):
    """axs_slab_buffers (include/axisym.h)."""

    _fields_ = [(n, ctypes.c_void_p) for n in (
        "vorticity", "psi", "u_z", "u_r", "scalars", "char_func", "r", "kill_wr")]


class SoftSlabConfig(ctypes.Structure):
    """axs_soft_slab_config (include/axisym.h)."""

    _fields_ = [(n, ctypes.c_int) for n in (
        "nr", "nz", "ghost", "kill_width", "sample_rows", "sample_cap",
        "max_extrap_iterations")] + [(n, ctypes.c_double) for n in (
            "dx", "nu", "G", "rho", "lam", "V_wall", "omega", "CFL", "eps", "z_max", "extrap_tol",
            "extrap_band", "freq_limit", "t_end")]


class SoftSlabBuffers(ctypes.Structure):
    """axs_soft_slab_buffers (include/axisym.h)."""

    _fields_ = [(n, ctypes.c_void_p) for n in (
        "vorticity", "psi", "u_z", "u_r", "eta1", "eta2", "solid_phi", "scalars", "samples",
        "solid_char_func", "wall_char_func", "r", "kill_wr")]


class TorusConfig(ctypes.Structure):
    """axs_torus_config (include/axisym.h)."""

    _fields_ = [(n, ctypes.c_int) for n in ("nr", "nz", "kill_width", "nmodes")] + [
        (n, ctypes.c_double) for n in (
            "dx", "nu", "lam", "moll_zone", "a", "AR", "e", "Z_cm", "R_cm", "dt")] + [
        ("amp", ctypes.c_double * 4), ("omega", ctypes.c_double * 4)]


class TorusBuffers(ctypes.Structure):
    """axs_torus_buffers (include/axisym.h)."""

    _fields_ = [(n, ctypes.c_void_p) for n in (
        "vorticity", "psi", "u_z", "u_r", "u_z_pen", "u_r_pen", "char_func", "vel_phi", "scalars",
        "r", "z", "kill_wz", "kill_wr")]


AXS_F64 = 0
AXS_F32 = 1
AXS_RIGID_SCALARS = 24


# name -> argument types (restype axs_status = int unless listed in _RESTYPE)
SIGNATURES = {
    "axs_last_error": [],
    "axs_version": [],
    "axs_kill_boundary_vorticity_sine_z_f64": [_P, _P, _I, _I, _I, _D, _P],
    "axs_kill_boundary_vorticity_sine_r_f64": [_P, _P, _I, _I, _I, _D, _P],
    "axs_compute_velocity_from_psi_unb_f64": [_P, _P, _P, _P, _I, _I, _D, _P],
    "axs_compute_vorticity_from_velocity_unb_f64": [_P, _P, _P, _I, _I, _D, _P],
    "axs_brinkmann_penalize_f64": [_D, _D, _P, _D, _D, _P, _P, _P, _P, _L, _P],
    "axs_smooth_heaviside_f64": [_P, _P, _D, _L, _P],
    "axs_advect_vorticity_eno3_f64": [_P, _P, _P, _P, _I, _I, _D, _D, _P],
    "axs_diffusion_rk2_unb_f64": [_P, _P, _P, _I, _I, _D, _D, _D, _P],
    "axs_reduce_partials_size": [_L],
    "axs_penalization_force_sum_f64": [_P, _P, _P, _D, _L, _P, _P, _P],
    "axs_penalization_force_sum_noalloc_f64": [_P, _P, _P, _D, _L, _P, _P],
    "axs_max_abs_velocity_sum_f64": [_P, _P, _L, _P, _P, _P],
    "axs_device_copy_f64": [_P, _P, _L, _P],
    "axs_solid_sigma_f64": [_P, _P, _P, _D, _D, _P, _P, _P, _P, _P, _P, _I, _I, _P],
    "axs_update_vorticity_from_solid_stress_f64": [_P, _P, _P, _P, _P, _P, _P, _I, _I, _D, _D, _P],
    "axs_advect_refmap_eno3_f64": [_P, _P, _P, _P, _P, _I, _I, _D, _D, _P],
    "axs_stokes_plan_create": [_I, _I, _D, _I, ctypes.POINTER(_P)],
    "axs_stokes_plan_destroy": [_P],
    "axs_stokes_solve_f64": [_P, _P, _P, _P],
    "axs_potential_plan_create": [_I, _I, _D, _I, ctypes.POINTER(_P)],
    "axs_potential_plan_destroy": [_P],
    "axs_potential_solve_f64": [_P, _P, _P, _P],
    "axs_velocity_from_phi_f64": [_P, _P, _P, _I, _I, _D, _P],
    "axs_periodic_ghost_comm_f64": [_P, _I, _I, _I, _I, _D, _D, _P],
    "axs_diffusion_rk2_periodic_f64": [_P, _P, _P, _I, _I, _D, _D, _D, _I, _P],
    "axs_update_vorticity_from_solid_stress_periodic_f64": [_P, _P, _P, _P, _P, _P, _P, _I, _I,
                                                            _D, _D, _I, _P],
    "axs_implicit_diffusion_plan_create": [_I, _I, _D, _D, ctypes.POINTER(_P)],
    "axs_implicit_diffusion_plan_destroy": [_P],
    "axs_implicit_diffusion_step_f64": [_P, _P, _P],
    "axs_extrapolate_lsq_first_order_f64": [_P, _P, _P, _P, _P, _P, _I, _I, _I,
                                            ctypes.POINTER(_I), _P],
    "axs_particles_to_mesh_2d_mp4_f64": [_P, _P, _P, _P, _I, _I, _D, _D, _I, ctypes.POINTER(_I),
                                         _P],
    "axs_mesh_to_particles_2d_mp4_f64": [_P, _P, _P, _P, _P, _P, _I, _I, _L, _D, _D, _I, _P],
    "axs_advect_vorticity_via_particles_f64": [_P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _I, _D, _D,
                                               _I, _P],
    "axs_rigid_stepper_create": [ctypes.POINTER(RigidConfig), ctypes.POINTER(RigidBuffers),
                                 ctypes.POINTER(_P)],
    "axs_reinit_distance_f64": [_P, _P, _P, _I, _I, _D, _D, _I, _P],
    "axs_soft_stepper_create": [ctypes.POINTER(SoftConfig), ctypes.POINTER(SoftBuffers),
                                ctypes.POINTER(_P)],
    "axs_soft_stepper_destroy": [_P],
    "axs_soft_stepper_run": [_P, _I, _I, _P],
    "axs_soft_batch_create": [_I, _P],
    "axs_soft_batch_add": [_P, ctypes.POINTER(SoftConfig), ctypes.POINTER(SoftBuffers), _P],
    "axs_soft_batch_create_f32": [_I, _P],
    "axs_soft_batch_add_f32": [_P, ctypes.POINTER(SoftConfig), ctypes.POINTER(SoftBuffersF32),
                               _P],
    "axs_soft_batch_remove": [_P, _I],
    "axs_soft_batch_run": [_P, _I, _I, _P],
    "axs_soft_batch_info": [_P, _P, _P, _P, _P],
    "axs_soft_batch_destroy": [_P],
    "axs_slab_stepper_create": [ctypes.POINTER(SlabConfig), ctypes.POINTER(SlabBuffers),
                                ctypes.POINTER(_P)],
    "axs_slab_stepper_destroy": [_P],
    "axs_slab_stepper_run": [_P, _I, _I, _P],
    "axs_soft_slab_stepper_create": [ctypes.POINTER(SoftSlabConfig),
                                     ctypes.POINTER(SoftSlabBuffers), ctypes.POINTER(_P)],
    "axs_soft_slab_stepper_destroy": [_P],
    "axs_soft_slab_stepper_run": [_P, _I, _I, _P],
    "axs_torus_stepper_create": [ctypes.POINTER(TorusConfig), ctypes.POINTER(TorusBuffers),
                                 ctypes.POINTER(_P)],
    "axs_torus_stepper_destroy": [_P],
    "axs_torus_stepper_run": [_P, _I, _I, _P],
    "axs_rigid_batch_create": [ctypes.POINTER(RigidConfig), _I, ctypes.POINTER(RigidBuffers),
                               ctypes.POINTER(_P)],
    "axs_rigid_stepper_destroy": [_P],
    "axs_rigid_stepper_run": [_P, _I, _I, _P],
    "axs_rigid_stepper_reset_scalars": [_P, _P],
    "axs_rigid_stepper_velocity": [_P, _P],
    "axs_rigid_stepper_graph_launches": [_P, ctypes.POINTER(_L), ctypes.POINTER(_L)],
    "axs_rigid_stepper_stage_count": [],
    "axs_rigid_stepper_stage_name": [_I],
    "axs_rigid_stepper_run_stage": [_P, _I, _I, _P],
}
_RESTYPE = {
    "axs_last_error": ctypes.c_char_p,
    "axs_version": ctypes.c_char_p,
    "axs_reduce_partials_size": ctypes.c_int64,
    "axs_rigid_stepper_stage_name": ctypes.c_char_p,
}

_lib = None


def load():
    """Loads libaxisym_hip.so (no GPU needed to load)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built: run `python -m pyaxisymflow_amd.csrc.build`"
            )
        lib = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = _RESTYPE.get(name, ctypes.c_int)
        _lib = lib
    return _lib


class AxisymError(RuntimeError):
    pass


def check(status, what=""):
    if status == AXS_OK:
        return
    msg = load().axs_last_error().decode(errors="replace")
    if status == AXS_ERR_INVALID:
        raise ValueError(f"{what}: {msg}")
    if status == AXS_ERR_UNSUPPORTED:
        raise NotImplementedError(f"{what}: {msg}")
    raise AxisymError(f"{what}: {msg} (status {status})")


def call(name, *args):
    check(getattr(load(), name)(*args), name)


# ---------------------------------------------------------------- devices

_torch = None


def torch():
    global _torch
    if _torch is None:
        import torch as t

        _torch = t
    return _torch


def device():
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError("pyaxisymflow_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return t.device("cuda", t.cuda.current_device())


def stream_handle():
    return ctypes.c_void_p(torch().cuda.current_stream().cuda_stream)


def is_device_tensor(a):
    t = torch()
    return isinstance(a, t.Tensor) and a.is_cuda


class Staged:
    """Device views of caller arrays for one call.

    ``arg(a, out=...)`` returns a device pointer for ``a``; numpy arrays are
    copied in (and, if ``out``, back on ``finish``)."""

    def __init__(self, dtype=np.float64):
        self.dtype = dtype
        self._copyback = []
        self._keep = []
        self.dev = None

    def _tdtype(self, dtype):
        t = torch()
        return {np.float64: t.float64, np.float32: t.float32, np.int16: t.int16,
                np.uint8: t.uint8}[dtype]

    def tensor(self, a, name, dtype=None, out=False, copy_in=True):
        dtype = dtype or self.dtype
        t = torch()
        if isinstance(a, np.ndarray):
            if a.dtype != dtype:
                raise TypeError(f"{name}: expected {np.dtype(dtype).name}, got {a.dtype}")
            if not a.flags.c_contiguous:
                raise TypeError(f"{name}: expected a C-contiguous array")
            if self.dev is None:
                self.dev = device()
            if copy_in:
                d = t.from_numpy(a).to(self.dev)
            else:
                d = t.empty(a.shape, dtype=self._tdtype(dtype), device=self.dev)
            if out:
                self._copyback.append((a, d))
            self._keep.append(d)
            return d
        if is_device_tensor(a):
            if a.dtype != self._tdtype(dtype):
                raise TypeError(f"{name}: expected {np.dtype(dtype).name}, got {a.dtype}")
            if not a.is_contiguous():
                raise TypeError(f"{name}: expected a contiguous tensor")
            self.dev = a.device
            return a
        raise TypeError(f"{name}: expected a numpy array or a ROCm torch tensor, got {type(a)}")

    def ptr(self, a, name, dtype=None, out=False, copy_in=True):
        return ctypes.c_void_p(self.tensor(a, name, dtype, out, copy_in).data_ptr())

    def scratch(self, shape, dtype=None):
        t = torch()
        if self.dev is None:
            self.dev = device()
        d = t.empty(shape, dtype=self._tdtype(dtype or self.dtype), device=self.dev)
        self._keep.append(d)
        return d

    def finish(self):
        for host, dev in self._copyback:
            host[...] = dev.cpu().numpy()
        self._copyback.clear()
        self._keep.clear()


def shape2(a):
    s = tuple(a.shape)
    if len(s) != 2:
        raise ValueError(f"expected a 2-D field, got shape {s}")
    return s


def same_shape(*arrays):
    s = shape2(arrays[0])
    for a in arrays[1:]:
        if tuple(a.shape) != s:
            raise ValueError(f"shape mismatch: {tuple(a.shape)} vs {s}")
    return s
