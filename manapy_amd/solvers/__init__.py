"""Solver classes with manapy's interface; their steps run in libmfv on the GPU."""
from .advec import AdvectionSolver
from .shallowater import ShallowWaterSolver
from .euler import EulerSolver3D, initialisation_euler_3d
from .diffusion import DiffusionSolver, AdvectionDiffusionSolver
from .ls import LinearSolver, HIPKrylovSolver, ScipySolver, MUMPSSolver, PETScKrylovSolver

__all__ = ['AdvectionSolver', 'ShallowWaterSolver', 'EulerSolver3D', 'initialisation_euler_3d', 'DiffusionSolver',
           'AdvectionDiffusionSolver', 'LinearSolver', 'HIPKrylovSolver', 'ScipySolver', 'MUMPSSolver',
           'PETScKrylovSolver']
