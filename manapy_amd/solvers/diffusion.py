"""Explicit diffusion and advection-diffusion with the diamond face gradient.

``DiffusionSolver`` (``solvers/diffusion/system.py:19-143``) and
``AdvectionDiffusionSolver`` (``solvers/advecdiff/system.py:19-175``) keep the
reference's interface (``stepper`` / ``compute_fluxes`` / ``compute_new_val``); each
call is a libmfv kernel: node interpolation (``centertovertex_2d`` / ``_3d``), face gradient
(``face_gradient_2d`` / ``_3d``), the dissipative residual (``explicitscheme_dissipative``),
the convective residual (``explicitscheme_convective_2d`` / ``_3d``) and the time steps of
``solvers/diffusion/fvm_utils.py:30-49`` / ``solvers/advecdiff/fvm_utils.py:317-340``.
"""
import ctypes

import numpy as np

from .. import _lib
from ..base import Struct, make_get_conf, process_conf
from ..device import stream_ptr
from ..variable import Variable


class DiffusionSolver(object):
    _parameters = [('Dxx', float, 0., 0., 'Diffusion in x direction'),
                   ('Dyy', float, 0., 0., 'Diffusion in y direction'),
                   ('Dzz', float, 0., 0., 'Diffusion in z direction'),
                   ('dt', float, 0., 0., 'time step'),
                   ('order', int, 1, 1, 'order of the convective scheme'),
                   ('cfl', float, .4, 0, 'cfl of the explicit scheme')]
    _convective = False

    def __init__(self, var=None, vel=(None, None, None), conf=None, **kwargs):
        if conf is None:
            conf = Struct()
        self.conf = process_conf(self._parameters, conf, kwargs)
        get = make_get_conf(self.conf, kwargs)
        if not isinstance(var, Variable):
            raise ValueError("primal var must be a Variable type")
        if not isinstance(vel[0], Variable):
            raise ValueError("u must be a Variable type")
        if not isinstance(vel[1], Variable):
            raise ValueError("v must be a Variable type")
        self.var = var
        self.comm = var.comm
        self.domain = var.domain
        self.dim = var.dim
        self.u, self.v = vel[0], vel[1]
        if len(vel) == 3 and vel[2] is not None:
            if not isinstance(vel[2], Variable):
                raise ValueError("w must be a Variable type")
            self.w = vel[2]
        else:
            self.w = Variable(domain=self.domain)
        self.Dxx = np.float64(get("Dxx"))
        self.Dyy = np.float64(get("Dyy"))
        self.Dzz = np.float64(get("Dzz"))
        self.dt = np.float64(get("dt"))
        self.order = np.int32(get("order"))
        self.cfl = np.float64(get("cfl"))
        self.diffusion = not (self.Dxx == self.Dyy == self.Dzz == 0)
        self.backend = self.domain.backend
        import torch
        dm = self.domain.device_mesh
        C = int(self.domain.nbcells)
        for t in ("convective", "dissipative", "source"):
            self.var.__dict__[t] = torch.zeros(C, dtype=torch.float64, device=dm.device)
        self._dt_dev = torch.zeros(1, dtype=torch.float64, device=dm.device)
        if self._convective and self.order == 2:
            dm.ensure_lsq()

    @property
    def _m(self):
        return self.domain.device_mesh

    def explicit_dissipative(self):
        """``diffusion/system.py:121-124``."""
        self.var.compute_face_gradient()
        v = self.var
        _lib.call('mfv_explicitscheme_dissipative', self._m.ptr, v.gradfacex.data_ptr(), v.gradfacey.data_ptr(),
                  v.gradfacez.data_ptr(), float(self.Dxx), float(self.Dyy), float(self.Dzz), v.dissipative.data_ptr(),
                  stream_ptr())

    def explicit_convective(self):
        """``advecdiff/system.py:126-137``."""
        if self.order == 2:
            self.var.compute_cell_gradient()
        v = self.var
        _lib.call('mfv_explicitscheme_convective', self._m.ptr, v.convective.data_ptr(), v.cell.data_ptr(),
                  v.ghost.data_ptr(), self.u.face.data_ptr(), self.v.face.data_ptr(), self.w.face.data_ptr(),
                  v.gradcellx.data_ptr(), v.gradcelly.data_ptr(), v.gradcellz.data_ptr(), v.psi.data_ptr(),
                  int(self.order), stream_ptr())

    def stepper(self):
        """dt (returned and stored), one device->host read as the reference's allreduce."""
        h = ctypes.c_double(0.)
        uvw = (self.u.cell.data_ptr(), self.v.cell.data_ptr(), self.w.cell.data_ptr()) if self._convective \
            else (None, None, None)
        _lib.call('mfv_time_step_diffusion', self._m.ptr, *uvw, float(self.cfl), float(self.Dxx), float(self.Dyy),
                  float(self.Dzz), self._dt_dev.data_ptr(), ctypes.byref(h), stream_ptr())
        self.dt = h.value
        return self.dt

    def compute_fluxes(self):
        self.var.update_halo_value()
        self.var.update_ghost_value()
        if self._convective:
            self.explicit_convective()
        if self.diffusion:
            self.var.interpolate_celltonode()
            self.explicit_dissipative()

    def compute_new_val(self):
        v = self.var
        _lib.call('mfv_update_new_value', self._m.ptr, v.cell.data_ptr(), v.convective.data_ptr(),
                  v.dissipative.data_ptr(), v.source.data_ptr(), float(self.dt), None, stream_ptr())


class AdvectionDiffusionSolver(DiffusionSolver):
    """``solvers/advecdiff/system.py:19-175``: upwind convection + diamond diffusion."""
    _convective = True
