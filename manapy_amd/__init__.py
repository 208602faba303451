"""manapy_amd: manapy's explicit finite-volume step on AMD MI355X (gfx950).

The Python classes keep manapy's interface (``Struct``, ``Domain``, ``Variable``,
solver classes with ``stepper`` / ``compute_fluxes`` / ``compute_new_val``); the
kernels run in libmfv.so (hand-written HIP, C ABI in include/mfv.h). State lives on
the GPU; nothing falls back to the CPU.
"""
from .base import Struct, make_get_conf
from .ddm import Domain
from .meshgen import jittered_triangles, kuhn_tetrahedra, gaussian_bumps

__version__ = "0.1"


def __getattr__(name):
    # GPU-side classes import torch lazily so that mesh construction works without it
    if name in ('Variable', 'Boundary'):
        from . import variable
        return getattr(variable, name)
    if name in ('AdvectionSolver', 'ShallowWaterSolver', 'EulerSolver3D', 'initialisation_euler_3d',
                'DiffusionSolver', 'AdvectionDiffusionSolver', 'LinearSolver', 'HIPKrylovSolver', 'ScipySolver',
                'MUMPSSolver', 'PETScKrylovSolver'):
        from . import solvers
        return getattr(solvers, name)
    if name == 'HIPBackend':
        from .backend import HIPBackend
        return HIPBackend
    raise AttributeError(name)


__all__ = ['Struct', 'make_get_conf', 'Domain', 'Variable', 'Boundary', 'AdvectionSolver', 'ShallowWaterSolver',
           'EulerSolver3D', 'initialisation_euler_3d', 'DiffusionSolver', 'AdvectionDiffusionSolver', 'LinearSolver',
           'HIPKrylovSolver', 'ScipySolver', 'MUMPSSolver', 'PETScKrylovSolver', 'HIPBackend', 'jittered_triangles',
           'kuhn_tetrahedra', 'gaussian_bumps']
