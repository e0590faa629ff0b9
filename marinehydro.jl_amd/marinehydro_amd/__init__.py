"""marinehydro_amd: MI355X (gfx950) boundary-element hot path of MarineHydro.jl.

The numerical work runs in libmhip.so (hand-written HIP kernels behind the C
ABI of include/mhip.h); this package is the host-side mirror of the
reference's Julia API for that path.
"""
from . import _lib
from .api import (SETTINGS, set_g, set_rho, GreensFunction, Rankine, RankineReflected,
                  RankineReflectedNegative, GFWu, ExactGuevelDelhommeau, TabulatedDelhommeau, gf_mask, Element, element,
                  gf_eval, greens, gradient_greens, integral, integral_gradient,
                  both_integral_and_integral_gradient, DeviceMesh, assemble_matrices,
                  assemble_matrices_batched, assemble_matrix_wu, solve, radiation_bc,
                  integrate_pressure, airy_waves_potential, airy_waves_velocity,
                  airy_waves_pressure, AiryBC, diffraction_force, FroudeKrylovForce, freq_sweep, SweepGraph,
                  calculate_radiation_forces, DiffractionForce, Tangents, shape_tangents,
                  freq_sweep_grad, froude_krylov, froude_krylov_grad,
                  assemble_matrices_vjp, rrule_assemble_matrices, solve_vjp, rrule_solve,
                  device_dof_fields, device_tangents)
from .meshes import (Mesh, Body, combine_meshes, face_geometry, face_geometry_jvp,
                     mesh_from_vertices, hemisphere, floating_sphere, sphere, vertical_cylinder, box, translated,
                     translational_dofs, rotational_dofs, rigid_dofs, axisymmetric_mesh,
                     wavebot_profile, wavebot_mesh, reference_face_geometry)
from .problems import (FloatingBody, combine_floatingbodies, LinearPotentialFlowProblem,
                       DiffractionProblem, RadiationProblem, LinearPotentialFlowResult,
                       DiffractionResult, RadiationResult, make_result, problems_from_data,
                       solve_problem, solve_all_problems, assemble_hydrodynamic_coefficients,
                       compute_hydrodynamic_coefficients, create_DimStack, DimArray,
                       compute_and_label_hydrodynamic_coefficients)
from .batches import BatchGraph, SweepItem, multi_sweep, multi_autotune, random_body, random_workload

__all__ = [n for n in dir() if not n.startswith("_")]
