"""The reference's import paths resolve to the GPU drop-ins: every function
and class that a reference driver imports from pyaxisymflow.kernels,
pyaxisymflow.elasto_kernels or pyaxisymflow.core on the path exists under
the same module path and name here, and is the pyaxisymflow_amd object (no
second implementation).  CPU only (imports, no compute)."""
import importlib

import pytest

# (module, name) as the reference defines them (file:line of the definition)
REFERENCE_NAMES = [
    ("kernels.FastDiagonalisationStokesSolver", "FastDiagonalisationStokesSolver"),  # :6
    ("kernels.FastDiagonalisationPotentialSolver", "FastDiagonalisationPotentialSolver"),  # :6
    ("kernels.implicit_diffusion_solver", "ImplicitEulerDiffusionStepper"),  # :6
    ("kernels.compute_velocity_from_psi", "compute_velocity_from_psi_unb"),  # :4
    ("kernels.compute_velocity_from_psi", "compute_velocity_from_psi_periodic"),  # :20
    ("kernels.compute_vorticity_from_velocity", "compute_vorticity_from_velocity_unb"),  # :4
    ("kernels.compute_vorticity_from_velocity", "compute_vorticity_from_velocity_periodic"),  # :16
    ("kernels.compute_velocity_from_phi", "compute_velocity_from_phi_unb"),  # :4
    ("kernels.brinkmann_penalize", "brinkmann_penalize"),  # :4
    ("kernels.smooth_Heaviside", "smooth_Heaviside"),  # :5
    ("kernels.kill_boundary_vorticity_sine", "kill_boundary_vorticity_sine_z"),  # :4
    ("kernels.kill_boundary_vorticity_sine", "kill_boundary_vorticity_sine_r"),  # :17
    ("kernels.advect_vorticity_via_eno3", "gen_advect_vorticity_via_eno3"),  # :8
    ("kernels.advect_vorticity_via_eno3", "gen_advect_vorticity_via_eno3_periodic"),  # :47
    ("kernels.diffusion_RK2", "diffusion_RK2_unb"),  # :4
    ("kernels.diffusion_RK2", "diffusion_RK2_periodic"),  # :47
    ("kernels.compute_forces", "compute_force_on_body"),  # :4
    ("kernels.periodic_boundary_ghost_comm", "gen_periodic_boundary_ghost_comm"),  # :4
    ("kernels.periodic_boundary_ghost_comm", "gen_periodic_boundary_ghost_comm_eta"),  # :18
    ("elasto_kernels.solid_sigma", "solid_sigma"),  # :4
    ("elasto_kernels.solid_sigma", "solid_sigma_periodic"),  # :32
    ("elasto_kernels.div_tau", "update_vorticity_from_solid_stress"),  # :4
    ("elasto_kernels.div_tau", "update_vorticity_from_solid_stress_periodic"),  # :37
    ("elasto_kernels.advect_refmap_via_eno3", "gen_advect_refmap_via_eno3"),  # :8
    ("elasto_kernels.advect_refmap_via_eno3", "gen_advect_refmap_via_eno3_periodic"),  # :57
    ("elasto_kernels.extrapolate_eta_using_least_squares_unb",
     "extrapolate_eta_with_least_squares"),  # :7
    ("elasto_kernels.extrapolate_using_least_squares",
     "extrapolate_eta_using_least_squares"),  # :8
    ("core.extrapolate_using_least_squares",
     "extrapolate_using_least_squares_till_first_order"),  # _bind.cpp:11-41
    ("kernels.advect_particle", "advect_vorticity_via_particles"),  # :5
    ("kernels.advect_particle", "advect_vorticity_via_particles_periodic"),  # :38
    ("core.particles_to_mesh", "particles_to_mesh_2D_mp4"),  # particles_to_mesh_bind.cpp:111
    ("core.particles_to_mesh", "particles_to_mesh_2D_unbounded_mp4"),  # :116
    ("core.mesh_to_particles", "mesh_to_particles_2D_mp4"),  # mesh_to_particles_bind.cpp:188
    ("core.mesh_to_particles", "mesh_to_particles_2D_unbounded_mp4"),  # :193
    ("utils.dump_vtk", "vtk_init"),  # :5
    ("utils.dump_vtk", "vtk_write"),  # :29
]


@pytest.mark.parametrize("module,name", REFERENCE_NAMES)
def test_reference_import_path(module, name):
    ref_path = importlib.import_module("pyaxisymflow." + module)
    amd = importlib.import_module("pyaxisymflow_amd." + module)
    assert getattr(ref_path, name) is getattr(amd, name)
