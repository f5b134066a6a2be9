"""Device-resident timestep of the rigid-body drivers.

``RigidBodyStepper`` runs the loop body of
examples/FlowPastSphere/flow_past_sphere.py:107-188 (and, with
``oscillation=(U_amp, omega)``, the penalisation velocity of
examples/SpheroidStreaming/spheroid_streaming.py:216-226) entirely on the
GPU: kill taper, Stokes solve, velocity and free stream, CFL dt, Brinkman
penalisation and its vorticity, drag coefficient, ENO3 advection and RK2
diffusion.  t, dt and the drag history never leave the device; one step is
captured into a hipGraph and replayed.  The fields live in torch tensors
owned by this object; the kernels are launched through the C ABI
(axs_rigid_stepper_*).
"""
import ctypes

import numpy as np

from pyaxisymflow_amd import _lib

# device scalar slots (axs_stepper.hpp)
SC_T, SC_DT, SC_DT_BY_DX, SC_LAM_DT, SC_HALF_NU_DT, SC_NU_DT, SC_PREFAC, SC_UPEN, SC_MAXU, \
    SC_DRAG_SUM, SC_IT, SC_FREQ_TIMER, SC_AVG_RESET, SC_BODY_Z, SC_FREQ_STEP, SC_PREFAC_STEP, \
    SC_UPEN_STEP = range(17)
N_SCALARS = 24  # AXS_RIGID_SCALARS


_Config = _lib.RigidConfig
_Buffers = _lib.RigidBuffers


def rigid_grid(nz, nr, z_len, r_len=None):
    """Cell-centred grid of the drivers (flow_past_sphere.py:40-60):
    dx = z_len / nz, z and r from numpy.linspace as the reference builds them."""
    dx = z_len / nz
    if r_len is None:
        r_len = nr * dx
    z = np.linspace(0 + dx / 2, z_len - dx / 2, nz)
    r = np.linspace(0 + dx / 2, r_len - dx / 2, nr)
    return dx, z, r


def kill_weights(z, r, dx, width=3):
    """Taper weights of kill_boundary_vorticity_sine.py:9-26, formed with the
    reference's own expressions (z end and r end taken as 1)."""
    wz_l = np.sin(np.pi * (z[:width] - 0.5 * dx) / 2 / (width - 1) / dx)
    wz_r = np.sin(np.pi * (1 - z[-width:] - 0.5 * dx) / 2 / (width - 1) / dx)
    wr = np.sin(np.pi * (1 - r[-width:] - 0.5 * dx) / 2 / (width - 1) / dx)
    return np.concatenate([wz_l, wz_r]), wr


class RigidBodyStepper:
    """Rigid body in an axisymmetric flow, stepped on the GPU.

    Parameters follow examples/FlowPastSphere/flow_past_sphere.py:39-56:
    grid (nz, nr) over z in [0, z_len] with dx = z_len / nz, viscosity nu,
    Brinkman lambda, body mask char_func (nr, nz), free stream U_0 with a
    sine ramp over T_ramp (None: no ramp), CFL number.  ``r_sph`` records the
    drag coefficient Cd = 4 pi dx^2 lam sum(R chi u_z) / (pi r_sph^2) every
    step; ``force_scale`` instead records force_scale * sum(R chi (u_z - U)).

    ``oscillation=(U_amp, omega)`` replaces the free stream by the
    penalisation velocity U_amp cos(omega t) (spheroid_streaming.py:216-226).
    ``moving_body=dict(z0, r0, aspect, radius, disp, moll_zone)`` rebuilds the
    mask each step from the spheroid centred at z0 + disp sin(omega t)
    (:207-212); ``dt_cap``, ``period`` and ``t_end`` add the streaming dt
    rule (:193-201) and ``average_psi`` the per-period average (:187, 204).

    ``precision="f32"`` is the precision form: fields, stencils and
    z-transforms in FP32; radial solve, mask level set and reductions FP64.
    """

    def __init__(self, nz, nr, z_len, nu, brink_lam, char_func=None, U_0=1.0, T_ramp=None,
                 r_sph=None, CFL=0.1, vorticity=None, oscillation=None, history=1 << 16,
                 kill_width=3, r_len=None, use_graph=True, precision="f64", moving_body=None,
                 average_psi=False, dt_cap=None, period=None, t_end=None, force_scale=None):
        torch = _lib.torch()
        dev = _lib.device()
        if precision not in ("f64", "f32"):
            raise ValueError("precision must be 'f64' or 'f32'")
        self.precision = precision
        self.nz, self.nr = int(nz), int(nr)
        self.dx, self.z, self.r = rigid_grid(nz, nr, z_len, r_len)
        dx = self.dx
        self.use_graph = use_graph
        self.stream = torch.cuda.Stream(device=dev)
        ft = torch.float64 if precision == "f64" else torch.float32
        self.field_dtype = ft

        def field(a=None):
            if a is None:
                return torch.zeros((nr, nz), dtype=ft, device=dev)
            a = np.asarray(a, dtype=np.float64)
            if a.shape != (nr, nz):
                raise ValueError(f"expected shape {(nr, nz)}, got {a.shape}")
            return torch.from_numpy(np.ascontiguousarray(a)).to(dev).to(ft)

        def table(a):
            return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)

        if char_func is None and moving_body is None:
            raise ValueError("need char_func or moving_body")
        self.t_vorticity = field(vorticity)
        self.t_psi = field()
        self.t_u_z = field()
        self.t_u_r = field()
        self.t_u_pen_z = field()
        self.t_u_pen_r = field()
        self.t_char_func = field(char_func)
        self.t_work0 = field()
        self.t_work1 = field()
        self.t_avg_psi = field() if average_psi else None
        self.t_scalars = torch.zeros(N_SCALARS, dtype=torch.float64, device=dev)
        self.history = int(history)
        self.t_drag = torch.zeros(max(self.history, 1), dtype=torch.float64, device=dev)
        self.t_r = table(self.r)
        self.t_z = table(self.z)
        wz, wr = kill_weights(self.z, self.r, dx, kill_width)
        self.t_wz = table(wz)
        self.t_wr = table(wr)

        cfg = _Config()
        cfg.nr, cfg.nz, cfg.kill_width, cfg.history = self.nr, self.nz, kill_width, self.history
        cfg.precision = _lib.AXS_F64 if precision == "f64" else _lib.AXS_F32
        cfg.dx, cfg.nu, cfg.brink_lam = dx, nu, brink_lam
        cfg.CFL = CFL
        cfg.eps = np.finfo(float).eps
        cfg.dt_diffusive = 0.9 * dx**2 / 4 / nu  # flow_past_sphere.py:151
        cfg.dt_cap = np.inf if dt_cap is None else dt_cap
        cfg.period = 0.0 if period is None else period
        cfg.t_end = 0.0 if t_end is None else t_end
        cfg.average_psi = 1 if average_psi else 0
        if oscillation is None:
            cfg.U0 = U_0
            cfg.T_ramp = T_ramp if T_ramp is not None else 0.0
            cfg.osc_amp, cfg.osc_omega = 0.0, 0.0
        else:
            cfg.U0, cfg.T_ramp = 0.0, 0.0
            cfg.osc_amp, cfg.osc_omega = oscillation
        if moving_body is not None:
            cfg.moving_body = 1
            cfg.body_z0 = moving_body["z0"]
            cfg.body_r0 = moving_body.get("r0", 0.0)
            cfg.body_aspect2 = moving_body.get("aspect", 1.0) ** 2
            cfg.body_radius = moving_body["radius"]
            cfg.body_disp = moving_body["disp"]
            cfg.moll_zone = moving_body["moll_zone"]
        if force_scale is not None:
            cfg.drag_scale, cfg.drag_div = force_scale, 1.0
        elif r_sph is not None:
            # flow_past_sphere.py:166-175
            cfg.drag_scale = 2 * 2 * np.pi * dx * dx * brink_lam
            cfg.drag_div = np.pi * r_sph**2
        else:
            cfg.drag_scale, cfg.drag_div = 0.0, 1.0
        self.cfg = cfg
        bufs = _Buffers()
        for name in ("vorticity", "psi", "u_z", "u_r", "u_pen_z", "u_pen_r", "char_func",
                     "work0", "work1", "scalars", "drag"):
            setattr(bufs, name, getattr(self, "t_" + name).data_ptr())
        bufs.avg_psi = self.t_avg_psi.data_ptr() if average_psi else None
        bufs.r = self.t_r.data_ptr()
        bufs.z = self.t_z.data_ptr()
        bufs.kill_wz = self.t_wz.data_ptr()
        bufs.kill_wr = self.t_wr.data_ptr()
        self._bufs = bufs
        torch.cuda.synchronize(dev)
        handle = ctypes.c_void_p()
        _lib.call("axs_rigid_stepper_create", ctypes.byref(cfg), ctypes.byref(bufs),
                  ctypes.byref(handle))
        self._h = handle

    # -------------------------------------------------------------- stepping
    def step(self, n=1):
        """Queues n timesteps on this stepper's stream (asynchronous)."""
        torch = _lib.torch()
        cur = torch.cuda.current_stream()
        self.stream.wait_stream(cur)
        _lib.call("axs_rigid_stepper_run", self._h, int(n), 1 if self.use_graph else 0,
                  ctypes.c_void_p(self.stream.cuda_stream))
        cur.wait_stream(self.stream)

    def synchronize(self):
        self.stream.synchronize()

    # -------------------------------------------------------------- state
    def _host(self, t):
        self.synchronize()
        return t.double().cpu().numpy()

    def vorticity(self):
        return self._host(self.t_vorticity)

    def psi(self):
        return self._host(self.t_psi)

    def avg_psi(self):
        return None if self.t_avg_psi is None else self._host(self.t_avg_psi)

    def char_func(self):
        return self._host(self.t_char_func)

    def velocity(self):
        """(u_z, u_r) after penalisation (the velocity ENO3 used in the last
        step), computed on demand from psi and chi."""
        _lib.call("axs_rigid_stepper_velocity", self._h, ctypes.c_void_p(self.stream.cuda_stream))
        return self._host(self.t_u_pen_z), self._host(self.t_u_pen_r)

    def scalars(self):
        return self._host(self.t_scalars)

    def time(self):
        return float(self.scalars()[SC_T])

    def steps_done(self):
        return int(self.scalars()[SC_IT])

    def dt(self):
        return float(self.scalars()[SC_DT])

    def drag_history(self):
        """Drag coefficient (or force) of every step so far, FP64."""
        n = min(self.steps_done(), self.history)
        return self._host(self.t_drag)[:n].copy()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self.synchronize()
                _lib.load().axs_rigid_stepper_destroy(h)
            except Exception:
                pass
            self._h = None


class OscillatingSpheroidStepper(RigidBodyStepper):
    """examples/SpheroidStreaming/spheroid_streaming.py:29-246 on the GPU:
    a spheroid oscillating along the axis in quiescent fluid (steady
    streaming), with the reference's parameters and derived quantities
    (:40-112).  The recorded force is 2 pi dx^2 rho lam sum(R chi (u_z - U))
    (compute_forces.py:14 with the 2 pi dx^2 of falling_rigid_sphere.py:155)."""

    def __init__(self, grid_size_z, womersley_square=100.0, radius=0.1, spheroid_AR=1.0,
                 freq=16, domain_AR=0.5, no_cycles=25, brink_lam=1e4, precision="f64",
                 history=None, **kw):
        dx = 1.0 / grid_size_z
        grid_size_r = int(domain_AR * grid_size_z)
        radius_normal = radius / spheroid_AR
        rad_eq = (radius_normal**2 * radius) ** (1 / 3)
        length_scale = spheroid_AR * rad_eq
        omega = 2 * np.pi * freq
        nond_AC = 1.0 / np.sqrt(womersley_square)
        e = 0.1
        U_0 = e * length_scale * omega
        Rs = (e / nond_AC) ** 2
        nu = e * U_0 * length_scale / Rs
        self.tEnd = no_cycles / freq
        self.freqTimer_limit = 1 / freq
        self.U_0, self.omega, self.nu_ref = U_0, omega, nu
        body = dict(z0=0.5, r0=0.0, aspect=spheroid_AR, radius=radius_normal,
                    disp=e * length_scale, moll_zone=dx * 2**0.5)
        if history is None:
            history = int(2.2 * self.tEnd / (0.9 * dx**2 / 4 / nu)) + 64
        super().__init__(grid_size_z, grid_size_r, 1.0, nu, brink_lam, char_func=None,
                         oscillation=(U_0, omega), moving_body=body, average_psi=True,
                         dt_cap=0.01 * self.freqTimer_limit, period=self.freqTimer_limit,
                         t_end=self.tEnd, force_scale=2 * np.pi * dx**2 * 1.0 * brink_lam,
                         r_len=domain_AR, precision=precision, history=history, **kw)
