"""Shallow water (``solvers/shallowater/system.py``) with a Rusanov flux.

The reference's flux is SRNH (``shallowater/fvm_utils.py:152-342``); this solver computes
the second-order MUSCL scheme with least-squares gradients, the Barth-Jespersen limiter,
hydrostatic reconstruction (Audusse et al. 2004) and the Rusanov flux, with the
reference's wall treatment (h Neumann, hu/hv nonslip). The class keeps the reference's
options (``system.py:25-43``), variables and call sequence.
"""
import ctypes

import numpy as np

from .. import _lib
from ..base import Struct, make_get_conf, process_conf
from ..device import stream_ptr, SwPacked
from ..variable import Variable
from ._common import DeviceClock


class ShallowWaterSolver(object):
    """Counterpart of ``ShallowWaterSolver`` (``solvers/shallowater/system.py:23-281``)."""

    _parameters = [('Dxx', float, 0., 0., 'Diffusion in x direction'),
                   ('Dyy', float, 0., 0., 'Diffusion in y direction'),
                   ('dt', float, 0., 0., 'time step'),
                   ('order', int, 1, 1, 'order of the convective scheme'),
                   ('cfl', float, .8, 0, 'cfl of the explicit scheme'),
                   ('Mann', float, 0., 0., 'Manning number for the friction'),
                   ('fc', float, 0., 0., 'Coriolis force'),
                   ('grav', float, 9.81, 0., 'gravity constant'),
                   ('wind', bool, False, True, 'wind')]

    def __init__(self, h=None, hvel=(None, None), hc=None, Z=None, conf=None, **kwargs):
        if conf is None:
            conf = Struct()
        self.conf = process_conf(self._parameters, conf, kwargs)
        get = make_get_conf(self.conf, kwargs)
        if not isinstance(h, Variable):
            raise ValueError("h must be a Variable type")
        if not isinstance(hvel[0], Variable):
            raise ValueError("hu must be a Variable type")
        if not isinstance(hvel[1], Variable):
            raise ValueError("hv must be a Variable type")
        self.h, self.hu, self.hv = h, hvel[0], hvel[1]
        self.domain = h.domain
        self.comm = h.comm
        self.dim = h.dim
        if int(self.dim) != 2:
            raise ValueError("shallow water needs a 2D domain")
        if Z is not None and not isinstance(Z, Variable):
            raise ValueError("Z must be a Variable type")
        if hc is not None and not isinstance(hc, Variable):
            raise ValueError("hC must be a Variable type")
        self.Z = Z if Z is not None else Variable(domain=self.domain)
        self.hc = hc if hc is not None else Variable(domain=self.domain)
        self.varbs = {'h': self.h, 'hu': self.hu, 'hv': self.hv, 'hc': self.hc, 'Z': self.Z}
        self.Dxx, self.Dyy, self.Dzz = get("Dxx"), get("Dyy"), 0.
        self.dt = get("dt")
        self.order = int(get("order"))
        self.cfl = float(get("cfl"))
        self.Mann, self.fc, self.grav, self.wind = get("Mann"), get("fc"), float(get("grav")), get("wind")
        if self.order not in (1, 2):
            raise ValueError("order must be 1 or 2")
        if self.Dxx != 0 or self.Dyy != 0:
            raise NotImplementedError("the reference's shallow-water diffusion path is broken (self.diffusion "
                                      "unset, system.py:146-147); not provided")
        if self.Mann != 0 or self.fc != 0 or self.wind:
            raise NotImplementedError("friction, Coriolis and wind terms are not on the accelerated path")
        for nm, var, want in (('h', self.h, 'neumann'), ('hu', self.hu, 'nonslip'), ('hv', self.hv, 'nonslip')):
            kinds = set(b.BCtype for b in var.BCs.values())
            if kinds != {want}:
                raise NotImplementedError("walls only: %s must be %s on every boundary (got %s)"
                                          % (nm, want, sorted(kinds)))
        import torch
        m = self.domain.device_mesh
        C = int(self.domain.nbcells)
        dev = m.device
        self._rez = torch.zeros((3, C), dtype=torch.float64, device=dev)
        self._src = torch.zeros((3, C), dtype=torch.float64, device=dev)
        self._grad = torch.zeros((9, C), dtype=torch.float64, device=dev)
        self._zgrad = torch.zeros((3, C), dtype=torch.float64, device=dev)
        terms = ['source', 'dissipation', 'coriolis', 'friction', 'convective']
        for var in self.varbs.values():
            for t in terms:
                var.__dict__[t] = torch.zeros(C, dtype=torch.float64, device=dev)
        for k, var in enumerate((self.h, self.hu, self.hv)):
            var.__dict__['convective'] = self._rez[k]
            var.__dict__['source'] = self._src[k]
        self._dt_dev = torch.zeros(1, dtype=torch.float64, device=dev)
        self._clock = DeviceClock(dev)
        self._packed = None
        self._zgrad_ready = False

    @property
    def _m(self):
        return self.domain.device_mesh

    def _bathymetry_gradient(self):
        """Unlimited least-squares gradient of Z with Neumann ghosts (static)."""
        self.Z.update_ghost_value()
        _lib.call('mfv_cell_gradient', self._m.ptr, self.Z.cell.data_ptr(), self.Z.ghost.data_ptr(),
                  self._zgrad[0].data_ptr(), self._zgrad[1].data_ptr(), self._zgrad[2].data_ptr(), stream_ptr())
        self._zgrad_ready = True

    def stepper(self):
        """``system.py:186-191``: time_step_SW, stored on ``self.dt``."""
        hd = ctypes.c_double(0.)
        _lib.call('mfv_time_step_sw', self._m.ptr, self.h.cell.data_ptr(), self.hu.cell.data_ptr(),
                  self.hv.cell.data_ptr(), self.cfl, self.grav, float(self.Dxx), float(self.Dyy),
                  self._dt_dev.data_ptr(), ctypes.byref(hd), stream_ptr())
        self.dt = hd.value
        return self.dt

    def update_ghost_values(self):
        for var in self.varbs.values():
            var.update_ghost_value()

    def explicit_convective(self):
        """``system.py:162-178`` with the Rusanov flux: gradients of eta, hu, hv, then residuals."""
        if not self._zgrad_ready:
            self._bathymetry_gradient()
        m = self._m
        s = stream_ptr()
        if self.order == 2:
            _lib.call('mfv_sw_gradients', m.ptr, self.h.cell.data_ptr(), self.hu.cell.data_ptr(),
                      self.hv.cell.data_ptr(), self.Z.cell.data_ptr(), self._grad.data_ptr(), s)
        _lib.call('mfv_sw_residuals', m.ptr, self.h.cell.data_ptr(), self.hu.cell.data_ptr(), self.hv.cell.data_ptr(),
                  self.Z.cell.data_ptr(), self._grad.data_ptr(), self._zgrad.data_ptr(), self.grav, self.order,
                  self._rez.data_ptr(), self._src.data_ptr(), s)

    def compute_fluxes(self):
        """``system.py:248-281`` (halo exchange is empty on one subdomain; the second dt
        evaluation of the reference, Q4, gives the same value and is skipped)."""
        self.update_ghost_values()
        self.explicit_convective()

    def compute_new_val(self):
        """``system.py:237-242`` -> update_SW."""
        self._dt_dev.fill_(float(self.dt))
        _lib.call('mfv_update_sw', self._m.ptr, self.h.cell.data_ptr(), self.hu.cell.data_ptr(),
                  self.hv.cell.data_ptr(), self._rez.data_ptr(), self._src.data_ptr(), self._dt_dev.data_ptr(),
                  stream_ptr())

    def run(self, nsteps):
        """``nsteps`` x (stepper, compute_fluxes, compute_new_val), fused on the packed path with
        dt on the device. Returns the simulated time; ``self.dt`` holds the last dt."""
        if not self._zgrad_ready:
            self._bathymetry_gradient()
        m = self._m
        s = stream_ptr()
        if self._packed is None:
            self._packed = SwPacked(m)
        p = self._packed
        _lib.call('mfv_sw_pack', m.ptr, p.ptr, self.h.cell.data_ptr(), self.hu.cell.data_ptr(),
                  self.hv.cell.data_ptr(), self.Z.cell.data_ptr(), self._zgrad.data_ptr(), self.order, s)
        _lib.call('mfv_time_step_sw', m.ptr, self.h.cell.data_ptr(), self.hu.cell.data_ptr(), self.hv.cell.data_ptr(),
                  self.cfl, self.grav, 0., 0., self._clock.dt0.data_ptr(), None, s)
        self._clock.start()
        for k in range(nsteps):
            _lib.call('mfv_sw_recon', m.ptr, p.ptr, self.order, s)
            _lib.call('mfv_sw_flux_packed', m.ptr, p.ptr, self.grav, self.cfl, self._clock.ring.data_ptr(), k, s)
        _lib.call('mfv_sw_unpack', m.ptr, p.ptr, self.h.cell.data_ptr(), self.hu.cell.data_ptr(),
                  self.hv.cell.data_ptr(), s)
        self._clock.step = nsteps
        if nsteps:
            self.dt = float(self._clock.ring[(nsteps - 1) % 3])
        return self._clock.time
