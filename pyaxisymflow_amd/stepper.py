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
    SC_UPEN_STEP, SC_PENDING = range(18)
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


_CASE_KEYS = ("nu", "brink_lam", "char_func", "U_0", "T_ramp", "r_sph", "CFL", "vorticity",
              "oscillation", "moving_body", "average_psi", "dt_cap", "period", "t_end",
              "force_scale", "implicit_diffusion")


def _case_config(nr, nz, dx, kill_width, history, precision, nu, brink_lam, char_func=None,
                 U_0=1.0, T_ramp=None, r_sph=None, CFL=0.1, vorticity=None, oscillation=None,
                 moving_body=None, average_psi=False, dt_cap=None, period=None, t_end=None,
                 force_scale=None, implicit_diffusion=None):
    """axs_rigid_config of one case (flow_past_sphere.py:39-56,
    spheroid_streaming.py:40-112 for the streaming terms).

    ``implicit_diffusion``: None or False for the RK2 diffusion; True for
    the implicit Euler diffusion at flow_past_sphere.py's fixed step
    dx^2 / 4 / nu (:90, 146-148); a number for that fixed step instead
    (spheroid_streaming.py's dt_limit, :104-112, 190-191)."""
    if char_func is None and moving_body is None:
        raise ValueError("need char_func or moving_body")
    cfg = _Config()
    cfg.nr, cfg.nz, cfg.kill_width, cfg.history = nr, nz, kill_width, history
    cfg.precision = _lib.AXS_F64 if precision == "f64" else _lib.AXS_F32
    cfg.dx, cfg.nu, cfg.brink_lam = dx, nu, brink_lam
    cfg.CFL = CFL
    cfg.eps = np.finfo(float).eps
    cfg.dt_diffusive = 0.9 * dx**2 / 4 / nu  # flow_past_sphere.py:151
    if implicit_diffusion is not None and implicit_diffusion is not False:
        if precision != "f64":
            raise NotImplementedError("implicit diffusion runs in the FP64 stepper")
        cfg.implicit_diffusion = 1
        # flow_past_sphere.py:90 (diffusion_dt_limit) or the caller's fixed step
        cfg.dt_diffusive = (dx**2 / 4 / nu if implicit_diffusion is True
                            else float(implicit_diffusion))
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
    return cfg


class _Checkpoint:
    """Checkpoint and resume of a stepper's device state.

    ``state()`` returns host copies of the arrays the next step depends on
    (the class's _STATE tensors: fields, scalars, histories); ``load_state``
    writes such a dict back, and a stepper built with the same parameters
    then continues bit for bit as the one it was taken from.  Everything
    else a step uses is rebuilt within the step, or cleared at its end (the
    completion of a run zeroes the velocity-maximum slots and the pending
    flag, as at creation).  ``save`` / ``load`` keep the dict in an .npz
    file (read back without pickle)."""

    _STATE = ()

    def state(self):
        self.synchronize()
        out = {}
        for n in self._STATE:
            t = getattr(self, "t_" + n, None)
            if t is not None:
                out[n] = t.cpu().numpy().copy()
        return out

    def load_state(self, state):
        torch = _lib.torch()
        self.synchronize()
        tensors = {n: getattr(self, "t_" + n, None) for n in self._STATE}
        for n, t in tensors.items():
            if t is None:
                if n in state:
                    raise ValueError(f"{n}: this stepper keeps no such array")
                continue
            if n not in state:
                raise KeyError(f"state lacks {n}")
            a = np.asarray(state[n])
            want = str(t.dtype).split(".")[-1]
            if a.shape != tuple(t.shape) or a.dtype != np.dtype(want):
                raise ValueError(f"{n}: expected {want} {tuple(t.shape)}, got {a.dtype} {a.shape}")
        for n, t in tensors.items():
            if t is not None:
                t.copy_(torch.from_numpy(np.ascontiguousarray(state[n])).to(t.device))
        torch.cuda.synchronize(_lib.device())

    def save(self, path):
        np.savez(path, **self.state())

    def load(self, path):
        with np.load(path, allow_pickle=False) as z:
            self.load_state({k: z[k] for k in z.files})


class RigidBatchStepper(_Checkpoint):
    """Many rigid-body cases of one grid, stepped together on the GPU (the
    batched form of examples/FlowPastSphere/flow_past_sphere.py:107-188).

    ``cases`` is a list of dicts with the per-case arguments of
    ``RigidBodyStepper`` (``nu``, ``brink_lam``, ``char_func``, ``U_0``,
    ``T_ramp``, ``r_sph``, ``CFL``, ``vorticity``, ``oscillation``,
    ``moving_body``, ``average_psi``, ``dt_cap``, ``period``, ``t_end``,
    ``force_scale``).  Every case keeps its own t, dt and force history;
    one launch per stage advances all of them.  Accessors take the case
    index.

    ``use_graph``: True replays captured hipGraphs, False launches the
    stage kernels one by one; "auto" (the default) takes graphs below
    EAGER_CELLS cells per step and eager launches from there on, where each
    kernel runs long enough (~8 us and more) for the host to stay ahead and
    eager launches avoid the graphs' start latency and the device's idle
    gap between consecutive graphs (~8 us each on this runtime).
    """

    EAGER_CELLS = 1 << 21
    _STATE = ("vorticity", "psi", "char_func", "avg_psi", "scalars", "drag")

    def __init__(self, nz, nr, z_len, cases, kill_width=3, r_len=None, use_graph="auto",
                 precision="f64", history=1 << 16):
        torch = _lib.torch()
        dev = _lib.device()
        if precision not in ("f64", "f32"):
            raise ValueError("precision must be 'f64' or 'f32'")
        cases = list(cases)
        if not cases:
            raise ValueError("need at least one case")
        for c in cases:
            bad = set(c) - set(_CASE_KEYS)
            if bad:
                raise TypeError(f"unknown case arguments {sorted(bad)}")
        self.batch = B = len(cases)
        self.precision = precision
        self.nz, self.nr = int(nz), int(nr)
        self.dx, self.z, self.r = rigid_grid(nz, nr, z_len, r_len)
        dx = self.dx
        if use_graph == "auto":
            use_graph = B * self.nr * self.nz < self.EAGER_CELLS
        self.use_graph = bool(use_graph)
        self.stream = torch.cuda.Stream(device=dev)
        ft = torch.float64 if precision == "f64" else torch.float32
        self.field_dtype = ft
        self.history = int(history)

        def fields(key=None):
            t = torch.zeros((B, nr, nz), dtype=ft, device=dev)
            if key is not None:
                for q, c in enumerate(cases):
                    a = c.get(key)
                    if a is None:
                        continue
                    a = np.asarray(a, dtype=np.float64)
                    if a.shape != (nr, nz):
                        raise ValueError(f"{key}: expected shape {(nr, nz)}, got {a.shape}")
                    t[q].copy_(torch.from_numpy(np.ascontiguousarray(a)).to(dev).to(ft))
            return t

        def table(a):
            return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)

        cfgs = (_Config * B)()
        for q, c in enumerate(cases):
            cfgs[q] = _case_config(self.nr, self.nz, dx, kill_width, self.history, precision, **c)
        self.cfgs = cfgs
        average_psi = bool(cfgs[0].average_psi)
        self.t_vorticity = fields("vorticity")
        self.t_psi = fields()
        self.t_u_pen_z = fields()
        self.t_u_pen_r = fields()
        self.t_char_func = fields("char_func")
        self.t_work0 = fields()
        self.t_avg_psi = fields() if average_psi else None
        self.t_scalars = torch.zeros((B, N_SCALARS), dtype=torch.float64, device=dev)
        self.t_drag = torch.zeros((B, max(self.history, 1)), dtype=torch.float64, device=dev)
        self.t_r = table(self.r)
        self.t_z = table(self.z)
        wz, wr = kill_weights(self.z, self.r, dx, kill_width)
        self.t_wz = table(wz)
        self.t_wr = table(wr)

        bufs = _Buffers()
        bufs.vorticity = self.t_vorticity.data_ptr()
        bufs.psi = self.t_psi.data_ptr()
        bufs.u_pen_z = self.t_u_pen_z.data_ptr()
        bufs.u_pen_r = self.t_u_pen_r.data_ptr()
        bufs.char_func = self.t_char_func.data_ptr()
        bufs.work0 = self.t_work0.data_ptr()
        bufs.scalars = self.t_scalars.data_ptr()
        bufs.drag = self.t_drag.data_ptr()
        bufs.avg_psi = self.t_avg_psi.data_ptr() if average_psi else None
        bufs.r = self.t_r.data_ptr()
        bufs.z = self.t_z.data_ptr()
        bufs.kill_wz = self.t_wz.data_ptr()
        bufs.kill_wr = self.t_wr.data_ptr()
        self._bufs = bufs
        torch.cuda.synchronize(dev)
        handle = ctypes.c_void_p()
        _lib.call("axs_rigid_batch_create", cfgs, B, ctypes.byref(bufs), ctypes.byref(handle))
        self._h = handle

    # -------------------------------------------------------------- stepping
    def step(self, n=1, join=True):
        """Queues n timesteps of every case on this stepper's stream
        (asynchronous).  join=True orders the steps after the work already
        queued on torch's current stream and that stream after the steps;
        join=False queues them on the stepper's stream alone (the state
        accessors below still wait for them), so several steppers driven
        from one thread run concurrently."""
        torch = _lib.torch()
        cur = torch.cuda.current_stream() if join else None
        if join:
            self.stream.wait_stream(cur)
        _lib.call("axs_rigid_stepper_run", self._h, int(n), 1 if self.use_graph else 0,
                  ctypes.c_void_p(self.stream.cuda_stream))
        if join:
            cur.wait_stream(self.stream)

    def graph_launches(self):
        """(16-step graph launches, launches of the shorter 8/4/2/1-step
        graphs) queued so far."""
        c, s = ctypes.c_int64(), ctypes.c_int64()
        _lib.call("axs_rigid_stepper_graph_launches", self._h, ctypes.byref(c), ctypes.byref(s))
        return c.value, s.value

    def synchronize(self):
        self.stream.synchronize()

    # -------------------------------------------------------------- state
    def _host(self, t):
        self.synchronize()
        return t.double().cpu().numpy()

    def vorticity(self, case=0):
        return self._host(self.t_vorticity[case])

    def psi(self, case=0):
        return self._host(self.t_psi[case])

    def avg_psi(self, case=0):
        return None if self.t_avg_psi is None else self._host(self.t_avg_psi[case])

    def char_func(self, case=0):
        return self._host(self.t_char_func[case])

    def velocity(self, case=0):
        """(u_z, u_r) after penalisation (the velocity ENO3 used in the last
        step), computed on demand from psi and chi."""
        _lib.call("axs_rigid_stepper_velocity", self._h, ctypes.c_void_p(self.stream.cuda_stream))
        return self._host(self.t_u_pen_z[case]), self._host(self.t_u_pen_r[case])

    def scalars(self, case=0):
        return self._host(self.t_scalars[case])

    def time(self, case=0):
        return float(self.scalars(case)[SC_T])

    def steps_done(self, case=0):
        return int(self.scalars(case)[SC_IT])

    def dt(self, case=0):
        return float(self.scalars(case)[SC_DT])

    def drag_history(self, case=0):
        """Drag coefficient (or force) of every step so far, FP64."""
        n = min(self.steps_done(case), self.history)
        return self._host(self.t_drag[case])[:n].copy()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self.synchronize()
                _lib.load().axs_rigid_stepper_destroy(h)
            except Exception:
                pass
            self._h = None


class UnevenBatchStepper(_Checkpoint):
    """Rigid-body cases of several grids, stepped together with one launch
    per pass over all of them (the batched form of
    examples/FlowPastSphere/flow_past_sphere.py:107-188 over uneven grids).

    ``cases`` is a list of dicts, each with the case's grid -- ``nz``,
    ``nr``, ``z_len`` and optionally ``r_len`` -- and the per-case arguments
    of ``RigidBodyStepper`` (no ``moving_body``; nz <= 1024).  Cases are
    stored grouped by grid (at most 16 grids); accessors take the caller's
    case index and return that case's (nr, nz) arrays.  Every case keeps its
    own t, dt and force history.
    """

    def __init__(self, cases, kill_width=3, use_graph=True, precision="f64", history=1 << 16):
        torch = _lib.torch()
        dev = _lib.device()
        if precision not in ("f64", "f32"):
            raise ValueError("precision must be 'f64' or 'f32'")
        cases = [dict(c) for c in cases]
        if not cases:
            raise ValueError("need at least one case")
        grids = []
        for c in cases:
            for k in ("nz", "nr", "z_len"):
                if k not in c:
                    raise TypeError(f"each case needs {k}")
            key = (int(c["nz"]), int(c["nr"]), float(c["z_len"]),
                   None if c.get("r_len") is None else float(c["r_len"]))
            c["_key"] = key
            bad = set(c) - set(_CASE_KEYS) - {"nz", "nr", "z_len", "r_len", "_key"}
            if bad:
                raise TypeError(f"unknown case arguments {sorted(bad)}")
            if key not in grids:
                grids.append(key)
        if len(grids) > 16:
            raise ValueError("at most 16 distinct grids")
        sig = {(nz, nr, z_len / nz) for nz, nr, z_len, _ in grids}
        if len(sig) != len(grids):
            raise ValueError("grids differing only in r_len cannot share one batch")
        self.batch = B = len(cases)
        self.precision = precision
        self.use_graph = use_graph
        self.history = int(history)
        self.stream = torch.cuda.Stream(device=dev)
        ft = torch.float64 if precision == "f64" else torch.float32
        self.field_dtype = ft
        # internal order: by grid, in order of first appearance
        self.order = sorted(range(B), key=lambda q: grids.index(cases[q]["_key"]))
        self.slot = {q: k for k, q in enumerate(self.order)}
        self.shape, self.offset = {}, {}
        self.grid = {}
        cfgs = (_Config * B)()
        r_tab, wz_tab, wr_tab = [], [], []
        off = 0
        for k, q in enumerate(self.order):
            c = cases[q]
            nz, nr, z_len, r_len = c["_key"]
            dx, z, r = rigid_grid(nz, nr, z_len, r_len)
            self.grid[q] = (dx, z, r)
            self.shape[q] = (nr, nz)
            self.offset[q] = off
            off += nr * nz
            args = {kk: v for kk, v in c.items() if kk in _CASE_KEYS}
            cfgs[k] = _case_config(nr, nz, dx, kill_width, self.history, precision, **args)
            wz, wr = kill_weights(z, r, dx, kill_width)
            r_tab.append(r)
            wz_tab.append(wz)
            wr_tab.append(wr)
        self.cells = off
        self.cfgs = cfgs
        average_psi = bool(cfgs[0].average_psi)

        def fields(key=None):
            t = torch.zeros(off, dtype=ft, device=dev)
            if key is not None:
                for q, c in enumerate(cases):
                    a = c.get(key)
                    if a is None:
                        continue
                    a = np.asarray(a, dtype=np.float64)
                    if a.shape != self.shape[q]:
                        raise ValueError(f"{key}: expected shape {self.shape[q]}, got {a.shape}")
                    self._view(t, q).copy_(
                        torch.from_numpy(np.ascontiguousarray(a)).to(dev).to(ft))
            return t

        def table(a):
            return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)

        self.t_vorticity = fields("vorticity")
        self.t_psi = fields()
        self.t_u_pen_z = fields()
        self.t_u_pen_r = fields()
        self.t_char_func = fields("char_func")
        self.t_work0 = fields()
        self.t_avg_psi = fields() if average_psi else None
        self.t_scalars = torch.zeros((B, N_SCALARS), dtype=torch.float64, device=dev)
        self.t_drag = torch.zeros((B, max(self.history, 1)), dtype=torch.float64, device=dev)
        self.t_r = table(np.concatenate(r_tab))
        self.t_wz = table(np.concatenate(wz_tab))
        self.t_wr = table(np.concatenate(wr_tab))
        bufs = _Buffers()
        bufs.vorticity = self.t_vorticity.data_ptr()
        bufs.psi = self.t_psi.data_ptr()
        bufs.u_pen_z = self.t_u_pen_z.data_ptr()
        bufs.u_pen_r = self.t_u_pen_r.data_ptr()
        bufs.char_func = self.t_char_func.data_ptr()
        bufs.work0 = self.t_work0.data_ptr()
        bufs.scalars = self.t_scalars.data_ptr()
        bufs.drag = self.t_drag.data_ptr()
        bufs.avg_psi = self.t_avg_psi.data_ptr() if average_psi else None
        bufs.r = self.t_r.data_ptr()
        bufs.z = None
        bufs.kill_wz = self.t_wz.data_ptr()
        bufs.kill_wr = self.t_wr.data_ptr()
        self._bufs = bufs
        torch.cuda.synchronize(dev)
        handle = ctypes.c_void_p()
        _lib.call("axs_rigid_batch_create", cfgs, B, ctypes.byref(bufs), ctypes.byref(handle))
        self._h = handle

    def _view(self, t, case):
        nr, nz = self.shape[case]
        o = self.offset[case]
        return t[o:o + nr * nz].view(nr, nz)

    _STATE = RigidBatchStepper._STATE
    step = RigidBatchStepper.step
    graph_launches = RigidBatchStepper.graph_launches
    synchronize = RigidBatchStepper.synchronize
    _host = RigidBatchStepper._host
    __del__ = RigidBatchStepper.__del__

    def vorticity(self, case=0):
        return self._host(self._view(self.t_vorticity, case))

    def psi(self, case=0):
        return self._host(self._view(self.t_psi, case))

    def avg_psi(self, case=0):
        return None if self.t_avg_psi is None else self._host(self._view(self.t_avg_psi, case))

    def char_func(self, case=0):
        return self._host(self._view(self.t_char_func, case))

    def velocity(self, case=0):
        _lib.call("axs_rigid_stepper_velocity", self._h, ctypes.c_void_p(self.stream.cuda_stream))
        return (self._host(self._view(self.t_u_pen_z, case)),
                self._host(self._view(self.t_u_pen_r, case)))

    def scalars(self, case=0):
        return self._host(self.t_scalars[self.slot[case]])

    def time(self, case=0):
        return float(self.scalars(case)[SC_T])

    def steps_done(self, case=0):
        return int(self.scalars(case)[SC_IT])

    def dt(self, case=0):
        return float(self.scalars(case)[SC_DT])

    def drag_history(self, case=0):
        n = min(self.steps_done(case), self.history)
        return self._host(self.t_drag[self.slot[case]])[:n].copy()


class RigidBodyStepper(RigidBatchStepper):
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

    ``implicit_diffusion=True`` replaces the RK2 diffusion by the implicit
    Euler step of kernels/implicit_diffusion_solver.py at the fixed
    dt = dx^2 / 4 / nu (flow_past_sphere.py:90-98, 146-148, 179-181); a
    number sets that fixed dt instead.  FP64 only.
    """

    def __init__(self, nz, nr, z_len, nu, brink_lam, char_func=None, U_0=1.0, T_ramp=None,
                 r_sph=None, CFL=0.1, vorticity=None, oscillation=None, history=1 << 16,
                 kill_width=3, r_len=None, use_graph="auto", precision="f64", moving_body=None,
                 average_psi=False, dt_cap=None, period=None, t_end=None, force_scale=None,
                 implicit_diffusion=None):
        case = dict(nu=nu, brink_lam=brink_lam, char_func=char_func, U_0=U_0, T_ramp=T_ramp,
                    r_sph=r_sph, CFL=CFL, vorticity=vorticity, oscillation=oscillation,
                    moving_body=moving_body, average_psi=average_psi, dt_cap=dt_cap,
                    period=period, t_end=t_end, force_scale=force_scale,
                    implicit_diffusion=implicit_diffusion)
        super().__init__(nz, nr, z_len, [case], kill_width=kill_width, r_len=r_len,
                         use_graph=use_graph, precision=precision, history=history)


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


# soft-step scalar slots (axs_soft.hip)
SS_T, SS_DT, SS_DT_BY_DX, SS_HALF_NU_DT, SS_NU_DT, SS_MAXU, SS_IT, SS_LSQ_SWEEPS, \
    SS_LSQ_UNCONVERGED = range(9)
# the arrays a soft step's successor depends on (checkpoints)
_SOFT_STATE = ("vorticity", "psi", "u_z", "u_r", "eta1", "eta2", "phi", "char_func", "scalars")


def _soft_case_state(obj, grid_size_z, G, nu, rho_f, r_ball, domain_AR, C, kill_width,
                     moll_cells=4, precision="f64"):
    """Device state, configuration and buffer table of one soft-sphere run
    (examples/SoftSphereInTaylorGreenVortex/soft_sphere_in_taylor_green_vortex.py:27-98),
    set as attributes of obj.  precision "f32": the flow fields (vorticity,
    psi, u_z, u_r, char_func) are float32 (the FP32-bulk soft batch); the
    reference map, level set and scalars stay float64."""
    if precision not in ("f64", "f32"):
        raise ValueError("precision must be 'f64' or 'f32'")
    torch = _lib.torch()
    dev = _lib.device()
    dx = 1.0 / grid_size_z
    nz, nr = int(grid_size_z), int(domain_AR * grid_size_z)
    obj.nz, obj.nr, obj.dx = nz, nr, dx
    moll_zone = dx * moll_cells
    extrap_zone = moll_zone + 4 * dx
    z = np.linspace(0 + dx / 2, 1 - dx / 2, nz)
    r = np.linspace(0 + dx / 2, domain_AR - dx / 2, nr)
    obj.z, obj.r = z, r
    # the driver's fields on its meshgrid (Z[i, j] = z[j], R[i, j] = r[i]);
    # built from the 1-D coordinates with the driver's operations in its
    # order, so every value is the same bits: phi on the host (numpy's
    # sqrt), the vorticity's sines on the 1-D arrays and their product on
    # the device (correctly rounded multiplies)
    phi = -np.sqrt((z[None, :] - 0.5) ** 2 + (r[:, None] - 0.0) ** 2) + r_ball
    kz = kr = 2 * np.pi
    sin_z = torch.from_numpy(C * (kz**2 + kr**2) * np.sin(kz * z)).to(dev)
    sin_r = torch.from_numpy(np.sin(kr * r)).to(dev)

    def t(a):
        return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)

    obj.precision = precision
    tdt = torch.float32 if precision == "f32" else torch.float64
    obj.t_vorticity = (sin_z[None, :] * sin_r[:, None]).to(tdt).contiguous()
    obj.t_psi = torch.zeros((nr, nz), dtype=tdt, device=dev)
    obj.t_u_z = torch.zeros((nr, nz), dtype=tdt, device=dev)
    obj.t_u_r = torch.zeros((nr, nz), dtype=tdt, device=dev)
    obj.t_eta1 = t(z)[None, :].expand(nr, nz).contiguous()
    obj.t_eta2 = t(r)[:, None].expand(nr, nz).contiguous()
    obj.t_phi = t(phi)
    from pyaxisymflow_amd.kernels.smooth_Heaviside import smooth_Heaviside

    chi = torch.zeros((nr, nz), dtype=torch.float64, device=dev)
    smooth_Heaviside(chi, obj.t_phi, moll_zone)
    obj.t_char_func = chi.to(torch.float32) if precision == "f32" else chi
    obj.t_scalars = torch.zeros(N_SCALARS, dtype=torch.float64, device=dev)
    obj.t_r, obj.t_z = t(r), t(z)
    wz, wr = kill_weights(z, r, dx, kill_width)
    obj.t_wz, obj.t_wr = t(wz), t(wr)
    cfg = _lib.SoftConfig()
    cfg.nr, cfg.nz, cfg.kill_width = nr, nz, kill_width
    cfg.lsq_sweeps, cfg.lsq_grow = 12, 1
    cfg.reinit_iterations = -1
    cfg.dx, cfg.nu, cfg.G, cfg.rho_f, cfg.CFL = dx, nu, G, rho_f, 0.1
    cfg.eps = np.finfo(float).eps
    cfg.r_ball, cfg.z_cm, cfg.r_cm = r_ball, 0.5, 0.0
    cfg.moll_zone, cfg.extrap_zone, cfg.reinit_band = moll_zone, extrap_zone, extrap_zone
    obj.cfg = cfg
    bufs = _lib.SoftBuffersF32() if precision == "f32" else _lib.SoftBuffers()
    for name in ("vorticity", "psi", "u_z", "u_r", "eta1", "eta2", "phi", "char_func", "scalars",
                 "r", "z"):
        setattr(bufs, name, getattr(obj, "t_" + name).data_ptr())
    bufs.kill_wz, bufs.kill_wr = obj.t_wz.data_ptr(), obj.t_wr.data_ptr()
    obj._bufs = bufs


class SoftSphereStepper(_Checkpoint):
    """examples/SoftSphereInTaylorGreenVortex/soft_sphere_in_taylor_green_vortex.py:27-215
    on the GPU: an elastic sphere (reference map eta, shear modulus G) in an
    axisymmetric Taylor-Green-like vortex, every stage of the loop body
    (:101-213) device resident and replayed from a hipGraph.

    Parameters and derived quantities follow the driver (:27-98): dx =
    1 / grid_size_z, nr = domain_AR * nz, moll_zone = 4 dx, extrap_zone =
    reinit_band = moll_zone + 4 dx, sphere of radius r_ball at (0.5, 0),
    vorticity C (kz^2 + kr^2) sin(kz z) sin(kr r) with kz = kr = 2 pi.
    ``lsq_sweeps`` fixes the number of extrapolation sweeps queued per step
    as separate launches; by default the queue starts at 12 and grows with
    the most sweeps a step has needed (+ 4, rounded up to 4) as graph runs
    learn it. The reference loops until a sweep fits nothing, which takes
    8-21 sweeps for stiff solids and up to ~60 for the softest as they
    deform; a loop longer than the queue is finished by one workgroup, so
    the result is always the reference loop's. ``lsq_unconverged()`` counts
    the steps the one-workgroup tail finished."""

    _STATE = _SOFT_STATE

    def __init__(self, grid_size_z=256, G=1.0, nu=1e-3, rho_f=1.0, r_ball=0.15, domain_AR=0.5,
                 C=5e-2, lsq_sweeps=None, use_graph=True, kill_width=3):
        torch = _lib.torch()
        dev = _lib.device()
        self.use_graph = use_graph
        self.stream = torch.cuda.Stream(device=dev)
        _soft_case_state(self, grid_size_z, G, nu, rho_f, r_ball, domain_AR, C, kill_width)
        # the reference's loop needs ~ band / dx sweeps at first and more as
        # the solid deforms (15-21 for G >= 1 at nz = 512, ~60 after 1000
        # steps at G = 0.1); every queued sweep costs a graph node (~5 us)
        # even after the loop has ended, a tail sweep ~50 us
        if lsq_sweeps:
            self.cfg.lsq_sweeps, self.cfg.lsq_grow = int(lsq_sweeps), 0
        torch.cuda.synchronize(dev)
        handle = ctypes.c_void_p()
        _lib.call("axs_soft_stepper_create", ctypes.byref(self.cfg), ctypes.byref(self._bufs),
                  ctypes.byref(handle))
        self._h = handle

    def step(self, n=1, join=True):
        """Queues n steps on this stepper's stream; join as in
        RigidBatchStepper.step (the server passes join=False so that the
        admitted requests' steps overlap on the device)."""
        torch = _lib.torch()
        cur = torch.cuda.current_stream() if join else None
        if join:
            self.stream.wait_stream(cur)
        _lib.call("axs_soft_stepper_run", self._h, int(n), 1 if self.use_graph else 0,
                  ctypes.c_void_p(self.stream.cuda_stream))
        if join:
            cur.wait_stream(self.stream)

    def synchronize(self):
        self.stream.synchronize()

    def _host(self, t):
        self.synchronize()
        return t.cpu().numpy()

    def vorticity(self):
        return self._host(self.t_vorticity)

    def psi(self):
        return self._host(self.t_psi)

    def velocity(self):
        return self._host(self.t_u_z), self._host(self.t_u_r)

    def eta(self):
        return self._host(self.t_eta1), self._host(self.t_eta2)

    def phi(self):
        return self._host(self.t_phi)

    def char_func(self):
        return self._host(self.t_char_func)

    def scalars(self):
        return self._host(self.t_scalars)

    def time(self):
        return float(self.scalars()[SS_T])

    def dt(self):
        return float(self.scalars()[SS_DT])

    def steps_done(self):
        return int(self.scalars()[SS_IT])

    def lsq_sweeps_last(self):
        return int(self.scalars()[SS_LSQ_SWEEPS])

    def lsq_unconverged(self):
        return int(self.scalars()[SS_LSQ_UNCONVERGED])

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self.synchronize()
                _lib.load().axs_soft_stepper_destroy(h)
            except Exception:
                pass
            self._h = None


class _SoftCaseView(_Checkpoint):
    """Host views of one case of a SoftSphereBatch (its fields live in
    tensors the batch steps in place); state() / load_state() checkpoint
    the case alone, and a state may be loaded into a case of another batch
    built with the same parameters."""

    _STATE = _SOFT_STATE

    def __init__(self, batch):
        self._batch = batch

    def synchronize(self):
        self._batch.synchronize()

    def _host(self, t):
        self._batch.synchronize()
        return t.cpu().numpy()

    def vorticity(self):
        return self._host(self.t_vorticity)

    def psi(self):
        return self._host(self.t_psi)

    def velocity(self):
        return self._host(self.t_u_z), self._host(self.t_u_r)

    def eta(self):
        return self._host(self.t_eta1), self._host(self.t_eta2)

    def phi(self):
        return self._host(self.t_phi)

    def char_func(self):
        return self._host(self.t_char_func)

    def scalars(self):
        return self._host(self.t_scalars)

    def time(self):
        return float(self.scalars()[SS_T])

    def steps_done(self):
        return int(self.scalars()[SS_IT])

    def lsq_unconverged(self):
        return int(self.scalars()[SS_LSQ_UNCONVERGED])


class SoftSphereBatch:
    """Many soft-sphere runs stepped together on one stream, each pass of the
    step one launch over all of them (axs_soft_batch_*).  ``add`` takes the
    SoftSphereStepper parameters and returns a case whose fields can be read
    like the stepper's; ``remove`` drops it.  A case's results are the same
    bits alone and in any mix; cases join and leave between ``step`` calls.

    precision "f32" is the FP32-bulk batch (axs_soft_batch_create_f32): the
    flow fields and their passes in float32, the radial elimination, the
    reference map, the level set, its reinit and the extrapolation in
    float64."""

    def __init__(self, max_cases=64, use_graph=True, precision="f64"):
        torch = _lib.torch()
        dev = _lib.device()
        if precision not in ("f64", "f32"):
            raise ValueError("precision must be 'f64' or 'f32'")
        self.precision = precision
        self.use_graph = use_graph
        self.stream = torch.cuda.Stream(device=dev)
        handle = ctypes.c_void_p()
        create = "axs_soft_batch_create_f32" if precision == "f32" else "axs_soft_batch_create"
        _lib.call(create, int(max_cases), ctypes.byref(handle))
        self._h = handle
        self.cases = {}

    def add(self, grid_size_z=256, G=1.0, nu=1e-3, rho_f=1.0, r_ball=0.15, domain_AR=0.5, C=5e-2,
            kill_width=3):
        torch = _lib.torch()
        case = _SoftCaseView(self)
        _soft_case_state(case, grid_size_z, G, nu, rho_f, r_ball, domain_AR, C, kill_width,
                         precision=self.precision)
        torch.cuda.synchronize(_lib.device())
        cid = ctypes.c_int()
        add = "axs_soft_batch_add_f32" if self.precision == "f32" else "axs_soft_batch_add"
        _lib.call(add, self._h, ctypes.byref(case.cfg), ctypes.byref(case._bufs),
                  ctypes.byref(cid))
        case.id = cid.value
        self.cases[case.id] = case
        return case

    def remove(self, case):
        _lib.call("axs_soft_batch_remove", self._h, int(case.id))
        del self.cases[case.id]

    def step(self, n=1, join=True):
        """Queues n steps of every case on the batch's stream (join: ordered
        after, and before, the current stream's work)."""
        torch = _lib.torch()
        cur = torch.cuda.current_stream() if join else None
        if join:
            self.stream.wait_stream(cur)
        _lib.call("axs_soft_batch_run", self._h, int(n), 1 if self.use_graph else 0,
                  ctypes.c_void_p(self.stream.cuda_stream))
        if join:
            cur.wait_stream(self.stream)

    def synchronize(self):
        self.stream.synchronize()

    def info(self):
        """(cases, extrapolation sweeps queued per step, graph launches,
        graph captures)."""
        n, q = ctypes.c_int(), ctypes.c_int()
        g, c = ctypes.c_int64(), ctypes.c_int64()
        _lib.call("axs_soft_batch_info", self._h, ctypes.byref(n), ctypes.byref(q),
                  ctypes.byref(g), ctypes.byref(c))
        return n.value, q.value, g.value, c.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self.synchronize()
                _lib.load().axs_soft_batch_destroy(h)
            except Exception:
                pass
            self._h = None


class PeriodicRigidSlabStepper:
    """examples/PeriodicRigidSlab/periodic_rigid_slab.py:33-158 on the GPU: a
    rigid annular wall (|r - R_wall_center| < wall_thickness, smoothed over
    dx sqrt(2)) oscillating along z at U_0 cos(omega t) in a z-periodic
    domain (ghost_size ghost columns at each end), every stage of the loop
    body (:93-158) device resident and replayed from hipGraphs
    (axs_slab_stepper_*).  Parameters and derived quantities follow the
    driver: nr = (max_r / max_z) grid_size_z, dx = max_r / nr,
    nu = 2 U_0 R_tube / Re with R_tube = R_wall_center - wall_thickness,
    omega = 2 pi freq."""

    def __init__(self, grid_size_z=16, max_z=0.05, max_r=0.5, brink_lam=1e4, wall_thickness=0.05,
                 R_wall_center=0.3, Re=25, U_0=1.0, freq=8.0, ghost_size=2, CFL=0.1,
                 use_graph=True, kill_width=3):
        torch = _lib.torch()
        dev = _lib.device()
        domain_AR = max_r / max_z
        nz, nr = int(grid_size_z), int(domain_AR * grid_size_z)
        dx = max_r / nr
        self.nz, self.nr, self.dx = nz, nr, dx
        self.use_graph = use_graph
        self.stream = torch.cuda.Stream(device=dev)
        self.z = np.linspace(0 + dx / 2, max_z - dx / 2, nz)
        self.r = np.linspace(0 + dx / 2, max_r - dx / 2, nr)
        Z, R = np.meshgrid(self.z, self.r)
        self.R_tube = R_wall_center - wall_thickness
        self.U_0 = U_0
        self.nu = U_0 * 2 * self.R_tube / Re
        self.omega = 2 * np.pi * freq
        phi0 = -np.sqrt((R - R_wall_center) ** 2) + wall_thickness

        def t(a):
            return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)

        self.t_vorticity = t(0 * Z)
        self.t_psi = t(0 * Z)
        self.t_u_z = t(0 * Z)
        self.t_u_r = t(0 * Z)
        self.t_char_func = t(0 * Z)
        self.t_phi0 = t(phi0)
        from pyaxisymflow_amd.kernels.smooth_Heaviside import smooth_Heaviside

        smooth_Heaviside(self.t_char_func, self.t_phi0, dx * 2**0.5)
        self.t_scalars = torch.zeros(N_SCALARS, dtype=torch.float64, device=dev)
        self.t_r = t(self.r)
        # kill_boundary_vorticity_sine.py:17-27 weights of the top rows, as
        # the reference forms them (its literal 1 as the r end)
        w = kill_width
        self.t_wr = t(np.sin(np.pi * (1 - self.r[-w:] - 0.5 * dx) / 2 / (w - 1) / dx))
        cfg = _lib.SlabConfig()
        cfg.nr, cfg.nz, cfg.ghost, cfg.kill_width = nr, nz, int(ghost_size), w
        cfg.dx, cfg.nu, cfg.lam, cfg.U0, cfg.omega = dx, self.nu, brink_lam, U_0, self.omega
        cfg.CFL, cfg.eps = CFL, np.finfo(float).eps
        self.cfg = cfg
        bufs = _lib.SlabBuffers()
        for name in ("vorticity", "psi", "u_z", "u_r", "scalars", "char_func", "r"):
            setattr(bufs, name, getattr(self, "t_" + name).data_ptr())
        bufs.kill_wr = self.t_wr.data_ptr()
        self._bufs = bufs
        torch.cuda.synchronize(dev)
        handle = ctypes.c_void_p()
        _lib.call("axs_slab_stepper_create", ctypes.byref(cfg), ctypes.byref(bufs),
                  ctypes.byref(handle))
        self._h = handle

    def step(self, n=1, join=True):
        torch = _lib.torch()
        cur = torch.cuda.current_stream() if join else None
        if join:
            self.stream.wait_stream(cur)
        _lib.call("axs_slab_stepper_run", self._h, int(n), 1 if self.use_graph else 0,
                  ctypes.c_void_p(self.stream.cuda_stream))
        if join:
            cur.wait_stream(self.stream)

    def synchronize(self):
        self.stream.synchronize()

    def _host(self, t):
        self.synchronize()
        return t.cpu().numpy()

    def vorticity(self):
        return self._host(self.t_vorticity)

    def psi(self):
        return self._host(self.t_psi)

    def velocity(self):
        """The penalised velocity of the last step (what it advected with)."""
        return self._host(self.t_u_z), self._host(self.t_u_r)

    def char_func(self):
        return self._host(self.t_char_func)

    def scalars(self):
        return self._host(self.t_scalars)

    def time(self):
        return float(self.scalars()[0])

    def dt(self):
        return float(self.scalars()[1])

    def steps_done(self):
        return int(self.scalars()[7])

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self.synchronize()
                _lib.load().axs_slab_stepper_destroy(h)
            except Exception:
                pass
            self._h = None


class PeriodicSoftSlabStepper:
    """examples/PeriodicSoftSlab/periodic_soft_slab.py:43-404 on the GPU: an
    elastic layer r < L_s inside fluid, driven by an annular wall at
    R_wall_center oscillating at V_wall cos(omega t), z-periodic with two
    ghost columns on each side, every stage of the loop body (:214-396)
    device resident and replayed from hipGraphs (axs_soft_slab_stepper_*).
    Parameters and derived quantities follow the driver (:55-137): max_r =
    0.5, max_z = max_r / domain_AR, nz = grid_size_r / domain_AR, dx = max_r /
    grid_size_r, R_tube = 0.2, L_f = R_tube zeta / (1 + zeta), L_s = R_tube /
    (1 + zeta), shear_rate = 2 V_wall / (omega L), nu = shear_rate omega L_f^2
    / Re, G = rho nu shear_rate omega / Er, freq = 1, t_end = 15 / freq.
    ``samples()`` returns what the driver compares with the theory
    (:229-238): the times and u_z of rows r < R_tube in column nz / 2, taken
    every 0.1 / freq."""

    def __init__(self, grid_size_r=256, Re=10.0, Er=0.25, domain_AR=32, zeta=1.0, t_end=None,
                 use_graph=True, kill_width=3, max_extrap_iterations=100000):
        torch = _lib.torch()
        dev = _lib.device()
        max_r = 0.5
        max_z = max_r / domain_AR
        nz, nr = int(grid_size_r / domain_AR), int(grid_size_r)
        dx = max_r / grid_size_r
        self.nz, self.nr, self.dx, self.max_z = nz, nr, dx, max_z
        self.use_graph = use_graph
        self.stream = torch.cuda.Stream(device=dev)
        self.z = np.linspace(0 + dx / 2, max_z - dx / 2, nz)
        self.r = np.linspace(0 + dx / 2, max_r - dx / 2, nr)
        Z, R = np.meshgrid(self.z, self.r)
        moll_zone = np.sqrt(2) * dx
        wall_half = 0.05
        self.R_wall_center = 0.5 * max_r
        self.R_tube = self.R_wall_center - wall_half
        L = 2 * self.R_tube
        self.L_f = self.R_tube * zeta / (1 + zeta)
        self.L_s = self.R_tube / (1 + zeta)
        self.freq = 1.0
        self.freq_limit = 0.1 / self.freq
        self.omega = 2 * np.pi * self.freq
        self.V_wall, self.rho_f = 1.0, 1.0
        self.shear_rate = 2 * self.V_wall / self.omega / L
        self.nu = self.shear_rate * self.omega * self.L_f**2 / Re
        self.G = self.rho_f * self.nu * self.shear_rate * self.omega / Er
        self.t_end = 15 / self.freq if t_end is None else float(t_end)
        self.sample_rows = int(nr * self.R_tube / max_r)
        self.sample_cap = int(np.ceil(self.t_end / self.freq_limit)) + 2

        def t(a):
            return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)

        from pyaxisymflow_amd.kernels.smooth_Heaviside import smooth_Heaviside

        self.t_vorticity, self.t_psi = t(0 * Z), t(0 * Z)
        self.t_u_z, self.t_u_r = t(0 * Z), t(0 * Z)
        self.t_eta1, self.t_eta2 = t(Z), t(R)
        self.t_solid_phi = t(self.L_s - R)
        self.t_solid_char_func = t(0 * Z)
        smooth_Heaviside(self.t_solid_char_func, self.t_solid_phi, moll_zone)
        self.t_wall_phi = t(-np.sqrt((R - self.R_wall_center) ** 2) + wall_half)
        self.t_wall_char_func = t(0 * Z)
        smooth_Heaviside(self.t_wall_char_func, self.t_wall_phi, moll_zone)
        sc = np.zeros(N_SCALARS)
        sc[8] = self.freq_limit  # the sample timer starts at its limit (:95)
        self.t_scalars = t(sc)
        self.t_samples = t(np.zeros((self.sample_cap, self.sample_rows + 1)))
        self.t_r = t(self.r)
        w = kill_width
        # kill_boundary_vorticity_sine.py:17-27 weights of the top rows, as
        # the reference forms them (its literal 1 as the r end)
        self.t_wr = t(np.sin(np.pi * (1 - self.r[-w:] - 0.5 * dx) / 2 / (w - 1) / dx))
        cfg = _lib.SoftSlabConfig()
        cfg.nr, cfg.nz, cfg.ghost, cfg.kill_width = nr, nz, 2, w
        cfg.sample_rows, cfg.sample_cap = self.sample_rows, self.sample_cap
        cfg.max_extrap_iterations = int(max_extrap_iterations)
        cfg.dx, cfg.nu, cfg.G, cfg.rho, cfg.lam = dx, self.nu, self.G, self.rho_f, 1e4
        cfg.V_wall, cfg.omega, cfg.CFL, cfg.eps = self.V_wall, self.omega, 0.1, np.finfo(float).eps
        cfg.z_max, cfg.extrap_tol = max_z, 1e-3
        cfg.extrap_band = moll_zone + 3 * dx
        cfg.freq_limit, cfg.t_end = self.freq_limit, self.t_end
        self.cfg = cfg
        bufs = _lib.SoftSlabBuffers()
        for name in ("vorticity", "psi", "u_z", "u_r", "eta1", "eta2", "solid_phi", "scalars",
                     "samples", "solid_char_func", "wall_char_func", "r"):
            setattr(bufs, name, getattr(self, "t_" + name).data_ptr())
        bufs.kill_wr = self.t_wr.data_ptr()
        self._bufs = bufs
        torch.cuda.synchronize(dev)
        handle = ctypes.c_void_p()
        _lib.call("axs_soft_slab_stepper_create", ctypes.byref(cfg), ctypes.byref(bufs),
                  ctypes.byref(handle))
        self._h = handle

    def step(self, n=1, join=True):
        torch = _lib.torch()
        cur = torch.cuda.current_stream() if join else None
        if join:
            self.stream.wait_stream(cur)
        _lib.call("axs_soft_slab_stepper_run", self._h, int(n), 1 if self.use_graph else 0,
                  ctypes.c_void_p(self.stream.cuda_stream))
        if join:
            cur.wait_stream(self.stream)

    def synchronize(self):
        self.stream.synchronize()

    def _host(self, t):
        self.synchronize()
        return t.cpu().numpy()

    def vorticity(self):
        return self._host(self.t_vorticity)

    def psi(self):
        return self._host(self.t_psi)

    def velocity(self):
        """The penalised velocity of the last step (what it advected with)."""
        return self._host(self.t_u_z), self._host(self.t_u_r)

    def eta(self):
        return self._host(self.t_eta1), self._host(self.t_eta2)

    def scalars(self):
        return self._host(self.t_scalars)

    def time(self):
        return float(self.scalars()[0])

    def dt(self):
        return float(self.scalars()[1])

    def steps_done(self):
        return int(self.scalars()[7])

    def jacobi_last(self):
        """Jacobi iterations of the last step's four extrapolation loops."""
        return int(self.scalars()[10])

    def jacobi_capped(self):
        """Extrapolation loops that stopped at the iteration bound."""
        return int(self.scalars()[11])

    def samples(self):
        """(times, u_z profiles) of the samples taken so far."""
        k = min(int(self.scalars()[9]), self.sample_cap)
        s = self._host(self.t_samples)[:k]
        return s[:, 0].copy(), s[:, 1:].copy()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self.synchronize()
                _lib.load().axs_soft_slab_stepper_destroy(h)
            except Exception:
                pass
            self._h = None


class TorusStreamingStepper:
    """examples/TorusStreaming/torus_streaming.py:22-233 on the GPU with
    implicit diffusion (the driver's setting): a torus (tube radius a =
    0.3415 r_core, cross-section aspect AR) whose centre oscillates in r as
    R_cm + e a sum_k A_k sin(w_k t), its dilation's source flow from the
    velocity potential and its surface velocity by Brinkman penalisation,
    every stage device resident and replayed from hipGraphs
    (axs_torus_stepper_*).  Parameters follow the driver (:22-60): grid
    256 x 128, e = 0.0675, modes 16 and 32 Hz with amplitudes 1 and 0.3,
    nu = omega (0.1124 r_core)^2, and the cycle-aligned dt (:106-112)."""

    def __init__(self, grid_size_z=256, domain_AR=0.5, use_graph=True, kill_width=3):
        torch = _lib.torch()
        dev = _lib.device()
        dx = 1.0 / grid_size_z
        nz, nr = int(grid_size_z), int(domain_AR * grid_size_z)
        self.nz, self.nr, self.dx = nz, nr, dx
        self.use_graph = use_graph
        self.stream = torch.cuda.Stream(device=dev)
        moll_zone = dx * 2**0.5
        r_core = 0.125
        a = 0.34146341463414637 * r_core
        AR = 0.4285714285714286
        e = 0.0675
        freq_modes = np.array([16, 32])
        amplitude_modes = np.array([1.0, 0.3])
        freq = freq_modes[np.argmax(amplitude_modes)]
        omega = 2 * np.pi * freq
        omega_modes = 2 * np.pi * freq_modes
        nu = omega * (0.11235582096628798 * r_core) ** 2
        U_osc = e * a * np.sum(amplitude_modes * omega_modes)
        freqTimer_limit = 1 / freq
        data_per_cycle = 100
        dataTimer_limit = 1 / freq / data_per_cycle
        diffusion_dt_limit = 0.1 / U_osc
        num_diffusion_substeps = max(int(freqTimer_limit / diffusion_dt_limit), data_per_cycle)
        diffusion_dt_limit = freqTimer_limit / num_diffusion_substeps
        self.dt_limit = min(diffusion_dt_limit, freqTimer_limit, dataTimer_limit)
        self.z = np.linspace(0 + dx / 2, 1 - dx / 2, nz)
        self.r = np.linspace(0 + dx / 2, domain_AR - dx / 2, nr)
        Z, R = np.meshgrid(self.z, self.r)
        Z_cm, R_cm = 0.5, r_core
        phi0 = -np.sqrt((Z - Z_cm) ** 2 + (R - R_cm) ** 2 / AR**2) + a

        def t(x):
            return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).to(dev)

        for name in ("vorticity", "psi", "u_z", "u_r", "u_z_pen", "u_r_pen", "char_func",
                     "vel_phi"):
            setattr(self, "t_" + name, t(0 * Z))
        self.t_phi0 = t(phi0)
        from pyaxisymflow_amd.kernels.smooth_Heaviside import smooth_Heaviside

        smooth_Heaviside(self.t_char_func, self.t_phi0, moll_zone)
        self.t_scalars = torch.zeros(N_SCALARS, dtype=torch.float64, device=dev)
        self.t_r, self.t_z = t(self.r), t(self.z)
        wz, wr = kill_weights(self.z, self.r, dx, kill_width)
        self.t_wz, self.t_wr = t(wz), t(wr)
        cfg = _lib.TorusConfig()
        cfg.nr, cfg.nz, cfg.kill_width, cfg.nmodes = nr, nz, kill_width, len(freq_modes)
        cfg.dx, cfg.nu, cfg.lam, cfg.moll_zone = dx, nu, 1e4, moll_zone
        cfg.a, cfg.AR, cfg.e, cfg.Z_cm, cfg.R_cm, cfg.dt = a, AR, e, Z_cm, R_cm, self.dt_limit
        for k in range(len(freq_modes)):
            cfg.amp[k], cfg.omega[k] = amplitude_modes[k], omega_modes[k]
        self.cfg = cfg
        bufs = _lib.TorusBuffers()
        for name in ("vorticity", "psi", "u_z", "u_r", "u_z_pen", "u_r_pen", "char_func",
                     "vel_phi", "scalars", "r", "z"):
            setattr(bufs, name, getattr(self, "t_" + name).data_ptr())
        bufs.kill_wz, bufs.kill_wr = self.t_wz.data_ptr(), self.t_wr.data_ptr()
        self._bufs = bufs
        torch.cuda.synchronize(dev)
        handle = ctypes.c_void_p()
        _lib.call("axs_torus_stepper_create", ctypes.byref(cfg), ctypes.byref(bufs),
                  ctypes.byref(handle))
        self._h = handle

    def step(self, n=1, join=True):
        torch = _lib.torch()
        cur = torch.cuda.current_stream() if join else None
        if join:
            self.stream.wait_stream(cur)
        _lib.call("axs_torus_stepper_run", self._h, int(n), 1 if self.use_graph else 0,
                  ctypes.c_void_p(self.stream.cuda_stream))
        if join:
            cur.wait_stream(self.stream)

    def synchronize(self):
        self.stream.synchronize()

    def _host(self, t):
        self.synchronize()
        return t.cpu().numpy()

    def vorticity(self):
        return self._host(self.t_vorticity)

    def psi(self):
        return self._host(self.t_psi)

    def velocity(self):
        """u from psi plus the potential flow (what the step advected with)."""
        return self._host(self.t_u_z), self._host(self.t_u_r)

    def penalised_velocity(self):
        return self._host(self.t_u_z_pen), self._host(self.t_u_r_pen)

    def char_func(self):
        return self._host(self.t_char_func)

    def scalars(self):
        return self._host(self.t_scalars)

    def time(self):
        return float(self.scalars()[0])

    def steps_done(self):
        return int(self.scalars()[5])

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self.synchronize()
                _lib.load().axs_torus_stepper_destroy(h)
            except Exception:
                pass
            self._h = None


class SoftSphereStreamingStepper(SoftSphereStepper):
    """examples/SoftSphereStreaming/soft_sphere_streaming.py:28-264 on the
    GPU: an elastic sphere (radius r_ball, shear modulus G) around a rigid
    core of radius zeta r_ball tethered to oscillate along z, U_0 cos(omega t)
    with U_0 = e r_ball omega, in fluid initially at rest (streaming flow).
    The loop body runs on the device as the soft stepper's streaming mode:
    the vorticity is advected before the solid stress acts, the core is
    Brinkman-penalised after it, and dt is clipped to the end of each cycle
    and to t_end.  Parameters and derived quantities follow the driver
    (:28-90): moll_zone = 2 dx, extrap_zone = moll_zone + 4 dx,
    nu = e U_0 r_ball / Rs with Rs = (e / nond_AC)^2, G = e rho_f
    (r_ball omega)^2 / Cauchy, brink_lam = 1e8.  The period averages and
    plots of the driver are not computed."""

    def __init__(self, grid_size_z=256, domain_AR=0.5, r_ball=0.15, freq=16, nond_AC=0.125,
                 e=0.1, Cauchy=0.1, zeta=0.25, no_cycles=30, rho_f=1.0, brink_lam=1e8,
                 lsq_sweeps=None, use_graph=True, kill_width=3):
        torch = _lib.torch()
        dev = _lib.device()
        self.use_graph = use_graph
        self.stream = torch.cuda.Stream(device=dev)
        omega = 2 * np.pi * freq
        U_0 = e * r_ball * omega
        Rs = (e / nond_AC) ** 2
        nu = e * U_0 * r_ball / Rs
        G = e * rho_f * (r_ball * omega) ** 2 / Cauchy
        self.omega, self.U_0, self.nu, self.G = omega, U_0, nu, G
        _soft_case_state(self, grid_size_z, G, nu, rho_f, r_ball, domain_AR, 0.0, kill_width,
                         moll_cells=2)
        cfg = self.cfg
        if lsq_sweeps:
            cfg.lsq_sweeps, cfg.lsq_grow = int(lsq_sweeps), 0
        cfg.streaming = 1
        cfg.brink_lam, cfg.U0, cfg.omega, cfg.e_amp = brink_lam, U_0, omega, e
        cfg.tether_rad = zeta * r_ball
        cfg.freq_limit = 1 / freq
        cfg.t_end = no_cycles / freq
        torch.cuda.synchronize(dev)
        handle = ctypes.c_void_p()
        _lib.call("axs_soft_stepper_create", ctypes.byref(cfg), ctypes.byref(self._bufs),
                  ctypes.byref(handle))
        self._h = handle
