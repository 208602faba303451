"""Linear advection (``solvers/advec/system.py``): upwind fluxes, orders 1 and 2."""
import ctypes

import numpy as np

from .. import _lib
from ..base import Struct, make_get_conf, process_conf
from ..device import stream_ptr
from ..variable import Variable
from ._common import DeviceClock


class AdvectionSolver(object):
    """Counterpart of ``AdvectionSolver`` (``solvers/advec/system.py:20-147``).

    ``stepper()`` returns dt as a Python float (one device->host read, as the reference's
    ``allreduce`` result); ``compute_fluxes()`` and ``compute_new_val()`` stay on the device.
    ``run(nsteps)`` advances ``nsteps`` fused steps with dt kept on the device and returns the
    simulated time (all-Neumann boundaries, the reference's default ``BC=None``)."""

    _parameters = [('dt', float, 0., 0., 'time step'),
                   ('order', int, 1, 1, 'order of the convective scheme'),
                   ('cfl', float, .4, 0, 'cfl of the explicit scheme')]

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
        self.dt = np.float64(get("dt"))
        self.order = np.int32(get("order"))
        self.cfl = np.float64(get("cfl"))
        if self.order not in (1, 2):
            raise ValueError("order must be 1 or 2")
        self.backend = self.domain.backend
        import torch
        C = int(self.domain.nbcells)
        dev = self.domain.device_mesh.device
        for t in ("convective", "dissipative", "source"):
            self.var.__dict__[t] = torch.zeros(C, dtype=torch.float64, device=dev)
        self._dt_dev = torch.zeros(1, dtype=torch.float64, device=dev)
        self._clock = DeviceClock(dev)
        if self.order == 2:
            self.domain.device_mesh.ensure_lsq()

    @property
    def _m(self):
        return self.domain.device_mesh

    def explicit_convective(self):
        """``system.py:116-126``."""
        if self.order == 2:
            self.var.compute_cell_gradient()
        v = self.var
        _lib.call('mfv_explicitscheme_convective', self._m.ptr, v.convective.data_ptr(), v.cell.data_ptr(),
                  v.ghost.data_ptr(), self.u.face.data_ptr(), self.v.face.data_ptr(), self.w.face.data_ptr(),
                  v.gradcellx.data_ptr(), v.gradcelly.data_ptr(), v.gradcellz.data_ptr(), v.psi.data_ptr(),
                  int(self.order), stream_ptr())

    def stepper(self):
        """``system.py:129-133``: dt = min_K cfl |K| / sum_f |u_K . n_f|."""
        h = ctypes.c_double(0.)
        _lib.call('mfv_time_step', self._m.ptr, self.u.cell.data_ptr(), self.v.cell.data_ptr(),
                  self.w.cell.data_ptr(), float(self.cfl), self._dt_dev.data_ptr(), ctypes.byref(h), stream_ptr())
        self.dt = h.value
        return self.dt

    def compute_fluxes(self):
        """``system.py:136-143``."""
        self.var.update_halo_value()
        self.var.update_ghost_value()
        self.explicit_convective()

    def compute_new_val(self):
        """``system.py:146-147``."""
        v = self.var
        _lib.call('mfv_update_new_value', self._m.ptr, v.cell.data_ptr(), v.convective.data_ptr(),
                  v.dissipative.data_ptr(), v.source.data_ptr(), float(self.dt), None, stream_ptr())

    def run(self, nsteps):
        """``nsteps`` x (stepper, compute_fluxes, compute_new_val) fused on the device, all-Neumann
        boundaries. Returns the simulated time; ``self.dt`` holds the last step's dt."""
        import torch
        if any(b.BCtype != "neumann" for b in self.var.BCs.values()):
            raise NotImplementedError("run() fuses the all-Neumann case; use the stepper/compute_fluxes/"
                                      "compute_new_val sequence for other boundary conditions")
        m = self._m
        s = stream_ptr()
        C = int(self.domain.nbcells)
        _lib.call('mfv_time_step', m.ptr, self.u.cell.data_ptr(), self.v.cell.data_ptr(), self.w.cell.data_ptr(),
                  float(self.cfl), self._clock.dt0.data_ptr(), None, s)
        self._clock.start()
        a = self.var.cell
        b = torch.empty_like(a)
        scratch = torch.zeros((int(self.dim) + 1) * C, dtype=torch.float64, device=a.device)
        for k in range(nsteps):
            _lib.call('mfv_advection_step', m.ptr, a.data_ptr(), b.data_ptr(), self.u.cell.data_ptr(),
                      self.v.cell.data_ptr(), self.w.cell.data_ptr(), self.u.face.data_ptr(), self.v.face.data_ptr(),
                      self.w.face.data_ptr(), float(self.cfl), int(self.order), self._clock.ring.data_ptr(), k,
                      scratch.data_ptr(), s)
            a, b = b, a
        if nsteps % 2 == 1:
            self.var.cell.copy_(a)
        self._clock.step = nsteps
        self.dt = float(self._clock.ring[(nsteps - 1) % 3]) if nsteps else self.dt
        return self._clock.time
