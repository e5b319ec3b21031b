"""Drop-in mirror of the reference's ``core`` API (core/__init__.py:3-27),
computing on the MI355X through libkcfd.  ``AdaptiveTimeControl`` is not
on the accelerated path and is not provided."""

from .fields import ScalarField, VectorField
from .fvm_operators import (FVMSystem, apply_boundary_conditions, convection_operator_upwind,
                            diffusion_operator, linearized_source, source_term, temporal_operator,
                            under_relax)
from .gpu_solver import benchmark_solvers, detect_gpu_backend, gpu_bicgstab, gpu_spmv
from .gradient import green_gauss_gradient, least_squares_gradient
from .interpolation import compute_mass_flux, get_limiter, muscl_deferred_correction
from .linear_solver import solve_linear_system, solve_pressure_correction
from .preconditioner import create_preconditioner

__all__ = [
    "ScalarField", "VectorField",
    "FVMSystem", "diffusion_operator", "source_term", "linearized_source",
    "convection_operator_upwind", "temporal_operator",
    "apply_boundary_conditions", "under_relax",
    "green_gauss_gradient", "least_squares_gradient",
    "solve_linear_system", "solve_pressure_correction",
    "compute_mass_flux", "muscl_deferred_correction", "get_limiter",
    "create_preconditioner",
    "detect_gpu_backend", "gpu_bicgstab", "gpu_spmv", "benchmark_solvers",
]
