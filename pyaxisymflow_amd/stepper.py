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
    SC_DRAG_SUM, SC_IT = range(11)
N_SCALARS = 16


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
    sine ramp over T_ramp (None: no ramp), CFL number.  ``r_sph`` enables the
    drag coefficient Cd = 4 pi dx^2 lam sum(R chi u_z) / (pi r_sph^2).
    ``oscillation=(U_amp, omega)`` replaces the free stream by the
    penalisation velocity U_amp cos(omega t) of spheroid_streaming.py:216-226.
    """

    def __init__(self, nz, nr, z_len, nu, brink_lam, char_func, U_0=1.0, T_ramp=None,
                 r_sph=None, CFL=0.1, vorticity=None, oscillation=None, history=1 << 16,
                 kill_width=3, r_len=None, use_graph=True):
        torch = _lib.torch()
        dev = _lib.device()
        self.nz, self.nr = int(nz), int(nr)
        self.dx, self.z, self.r = rigid_grid(nz, nr, z_len, r_len)
        dx = self.dx
        self.use_graph = use_graph
        self.stream = torch.cuda.Stream(device=dev)
        f64 = dict(dtype=torch.float64, device=dev)

        def field(a=None):
            if a is None:
                return torch.zeros((nr, nz), **f64)
            a = np.asarray(a, dtype=np.float64)
            if a.shape != (nr, nz):
                raise ValueError(f"expected shape {(nr, nz)}, got {a.shape}")
            return torch.from_numpy(np.ascontiguousarray(a)).to(dev)

        self.t_vorticity = field(vorticity)
        self.t_psi = field()
        self.t_u_z = field()
        self.t_u_r = field()
        self.t_u_pen_z = field()
        self.t_u_pen_r = field()
        self.t_char_func = field(char_func)
        self.t_work0 = field()
        self.t_work1 = field()
        self.t_scalars = torch.zeros(N_SCALARS, **f64)
        self.history = int(history)
        self.t_drag = torch.zeros(max(self.history, 1), **f64)
        self.t_r = torch.from_numpy(self.r.copy()).to(dev)
        wz, wr = kill_weights(self.z, self.r, dx, kill_width)
        self.t_wz = torch.from_numpy(wz).to(dev)
        self.t_wr = torch.from_numpy(wr).to(dev)

        cfg = _Config()
        cfg.nr, cfg.nz, cfg.kill_width, cfg.history = self.nr, self.nz, kill_width, self.history
        cfg.dx, cfg.nu, cfg.brink_lam = dx, nu, brink_lam
        cfg.CFL = CFL
        cfg.eps = np.finfo(float).eps
        cfg.dt_diffusive = 0.9 * dx**2 / 4 / nu  # flow_past_sphere.py:151
        if oscillation is None:
            cfg.U0 = U_0
            cfg.T_ramp = T_ramp if T_ramp is not None else 0.0
            cfg.osc_amp, cfg.osc_omega = 0.0, 0.0
        else:
            cfg.U0, cfg.T_ramp = 0.0, 0.0
            cfg.osc_amp, cfg.osc_omega = oscillation
        if r_sph is not None:
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
        bufs.r = self.t_r.data_ptr()
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
        return t.cpu().numpy()

    def vorticity(self):
        return self._host(self.t_vorticity)

    def psi(self):
        return self._host(self.t_psi)

    def velocity(self):
        """(u_z, u_r) after penalisation (the velocity ENO3 used)."""
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
