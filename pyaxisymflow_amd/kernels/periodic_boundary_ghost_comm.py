"""Drop-in for pyaxisymflow/kernels/periodic_boundary_ghost_comm.py:4-32."""
from pyaxisymflow_amd import _lib


def _comm(field, ghost_size, eta_shift, z_max, dx):
    st = _lib.Staged()
    nr, nz = _lib.shape2(field)
    _lib.call("axs_periodic_ghost_comm_f64", st.ptr(field, "field", out=True), nr, nz,
              int(ghost_size), int(eta_shift), float(z_max), float(dx), _lib.stream_handle())
    st.finish()


def gen_periodic_boundary_ghost_comm(ghost_size):
    """Returns periodic_boundary_ghost_comm(field): field[:, :g] =
    field[:, -2g:-g], field[:, -g:] = field[:, g:2g] in place (:4-15)."""
    assert ghost_size > 0 and isinstance(ghost_size, int), "invalid ghost size"

    def periodic_boundary_ghost_comm(field):
        _comm(field, ghost_size, 0, 0.0, 0.0)

    periodic_boundary_ghost_comm.ghost_size = ghost_size
    periodic_boundary_ghost_comm.eta_shift = None
    return periodic_boundary_ghost_comm


def gen_periodic_boundary_ghost_comm_eta(ghost_size, Z_max, dx):
    """The reference map's exchange (:18-32): the left ghosts get
    - Z_max + 2 g dx, the right ones + Z_max - 2 g dx."""
    assert ghost_size > 0 and isinstance(ghost_size, int), "invalid ghost size"

    def periodic_boundary_ghost_comm(field):
        _comm(field, ghost_size, 1, Z_max, dx)

    periodic_boundary_ghost_comm.ghost_size = ghost_size
    periodic_boundary_ghost_comm.eta_shift = (Z_max, dx)
    return periodic_boundary_ghost_comm


def plain_ghost_size(per_communicator, what):
    """Ghost width of a communicator made by gen_periodic_boundary_ghost_comm
    (the fused periodic entry points exchange ghosts between their stages)."""
    g = getattr(per_communicator, "ghost_size", None)
    if g is None or getattr(per_communicator, "eta_shift", None) is not None:
        raise TypeError(f"{what} needs a communicator from gen_periodic_boundary_ghost_comm")
    return g
