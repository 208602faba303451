"""Solver classes with manapy's interface; their steps run in libmfv on the GPU."""
from .advec import AdvectionSolver
from .shallowater import ShallowWaterSolver
from .euler import EulerSolver3D, initialisation_euler_3d

__all__ = ['AdvectionSolver', 'ShallowWaterSolver', 'EulerSolver3D', 'initialisation_euler_3d']
