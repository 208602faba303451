"""Shallow water (``solvers/shallowater/system.py``).

Two numerical fluxes behind the reference's class, options and call sequence
(``system.py:25-43``, ``stepper`` / ``compute_fluxes`` / ``compute_new_val``):

* ``flux='srnh'`` (default): the reference's own scheme on five fields (h, hu, hv, hc, Z):
  SRNH flux (``shallowater/fvm_utils.py:152-342``) with MUSCL on h and hc limited by
  hc's psi, the well-balanced source ``term_source_srnh_SW`` (triangles) and
  ``update_SW``. Any boundary conditions of the variables are honoured through their
  ghost values. ``quirk=True`` (default) reproduces the reference's use of the compacted
  ``cells.cellfid`` in the source term; ``quirk=False`` uses the true face neighbours.
* ``flux='rusanov'``: second-order MUSCL with least-squares gradients of eta, hu, hv, the
  Barth-Jespersen limiter, hydrostatic reconstruction (Audusse et al. 2004) and the
  Rusanov flux, with walls (h Neumann, hu/hv nonslip). ``run`` fuses it into two kernels
  per step on the packed layout.
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
                   ('wind', bool, False, True, 'wind'),
                   ('flux', str, 'srnh', 0, "numerical flux: 'srnh' (the reference's) or 'rusanov'"),
                   ('quirk', bool, True, 0, 'srnh source: read face neighbours from the compacted cellfid')]

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
        self.flux = str(get("flux")).lower()
        if self.flux not in ('srnh', 'rusanov'):
            raise ValueError("flux must be 'srnh' or 'rusanov'")
        if self.wind:
            raise NotImplementedError("wind: the reference calls term_wind_SW with arguments that do not match its "
                                      "signature (system.py:245-246) and passes zero wind to update_SW "
                                      "(system.py:237-242); the kernel itself is bound by HIPBackend.compile")
        if (self.Mann != 0 or self.fc != 0) and self.flux != 'srnh':
            raise NotImplementedError("friction and Coriolis terms run with the reference's scheme (flux='srnh')")
        self.quirk = int(bool(get("quirk")))
        import torch
        m = self.domain.device_mesh
        C = int(self.domain.nbcells)
        dev = m.device
        if self.flux == 'srnh':
            self._init_srnh(C, dev)
            return
        for nm, var, want in (('h', self.h, 'neumann'), ('hu', self.hu, 'nonslip'), ('hv', self.hv, 'nonslip')):
            kinds = set(b.BCtype for b in var.BCs.values())
            if kinds != {want}:
                raise NotImplementedError("walls only: %s must be %s on every boundary (got %s)"
                                          % (nm, want, sorted(kinds)))
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

    # --- SRNH (the reference's scheme) -----------------------------------------------------
    def _init_srnh(self, C, dev):
        import torch
        d = self.domain
        if int(d.maxcellfid) != 3 or int(d.cells.faceid[:, -1].max()) != 3:
            raise NotImplementedError("term_source_srnh_SW is defined for triangles only")
        terms = ['source', 'dissipation', 'coriolis', 'friction', 'convective']
        for var in self.varbs.values():
            for t in terms:
                var.__dict__[t] = torch.zeros(C, dtype=torch.float64, device=dev)
        t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a, dtype=dt), device=dev)
        self._nodeid = t(d.cells.nodeid, np.int32)
        self._cellfid = t(d.cells.cellfid, np.int32)
        self._nf = t(d.cells.nf, np.float64)
        self._vertex = t(d.nodes.vertex, np.float64)
        assert self._vertex.shape[1] == 4 and tuple(self._nf.shape[1:]) == (3, 3)
        self._dt_dev = torch.zeros(1, dtype=torch.float64, device=dev)
        self._zero = torch.zeros(C, dtype=torch.float64, device=dev)

    def _srnh_dt(self, dt_dev, host):
        hd = ctypes.c_double(0.) if host else None
        _lib.call('mfv_time_step_sw', self._m.ptr, self.h.cell.data_ptr(), self.hu.cell.data_ptr(),
                  self.hv.cell.data_ptr(), self.cfl, self.grav, float(self.Dxx), float(self.Dyy), dt_dev.data_ptr(),
                  ctypes.byref(hd) if host else None, stream_ptr())
        return hd.value if host else None

    def _srnh_convective(self):
        """``system.py:162-178``: gradients (+ limiter) of h and hc at order 2, then
        explicitscheme_convective_SW."""
        if self.order == 2:
            self.h.compute_cell_gradient()
            self.hc.compute_cell_gradient()
        v = (self.h, self.hu, self.hv, self.hc, self.Z)
        _lib.call('mfv_explicitscheme_convective_sw', self._m.ptr, *[x.convective.data_ptr() for x in v],
                  *[x.cell.data_ptr() for x in v], *[x.ghost.data_ptr() for x in v], self.h.gradcellx.data_ptr(),
                  self.h.gradcelly.data_ptr(), self.hc.gradcellx.data_ptr(), self.hc.gradcelly.data_ptr(),
                  self.hc.psi.data_ptr(), self.grav, self.order, stream_ptr())

    def update_term_source(self):
        """``system.py:204-215`` -> term_source_srnh_SW."""
        if self.flux != 'srnh':
            return
        v = (self.h, self.hu, self.hv, self.hc, self.Z)
        _lib.call('mfv_term_source_srnh_sw', self._m.ptr, *[x.source.data_ptr() for x in v], self.h.cell.data_ptr(),
                  self.Z.cell.data_ptr(), self.h.ghost.data_ptr(), self.Z.ghost.data_ptr(), self._nodeid.data_ptr(),
                  int(self._nodeid.shape[1]), self._cellfid.data_ptr(), int(self._cellfid.shape[1]),
                  self._nf.data_ptr(), self._vertex.data_ptr(), self.grav, self.quirk, stream_ptr())

    def _srnh_update(self, dt_dev):
        v = (self.h, self.hu, self.hv, self.hc, self.Z)
        _lib.call('mfv_update_sw5', self._m.ptr, *[x.cell.data_ptr() for x in v],
                  *[x.convective.data_ptr() for x in v], *[x.source.data_ptr() for x in v],
                  self.hc.dissipation.data_ptr(), self.hu.coriolis.data_ptr(), self.hv.coriolis.data_ptr(), 0., 0.,
                  float(self.dt), dt_dev.data_ptr() if dt_dev is not None else None, stream_ptr())

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
        if self.flux == 'srnh':
            self.dt = self._srnh_dt(self._dt_dev, True)
            return self.dt
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
        """``system.py:162-178``. Rusanov: gradients of eta, hu, hv, then residuals."""
        if self.flux == 'srnh':
            return self._srnh_convective()
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

    def update_term_friction(self):
        """``system.py:230-231`` -> term_friction_SW with the current dt."""
        _lib.call('mfv_term_friction_sw', int(self.domain.nbcells), self.h.cell.data_ptr(), self.hu.cell.data_ptr(),
                  self.hv.cell.data_ptr(), self.grav, float(self.Mann), float(self.dt), stream_ptr())

    def update_term_coriolis(self):
        """``system.py:233-234`` -> term_coriolis_SW; the reference passes the scalar fc where
        the kernel reads f_c[i], so fc is taken in every cell."""
        _lib.call('mfv_term_coriolis_sw', int(self.domain.nbcells), self.hu.cell.data_ptr(), self.hv.cell.data_ptr(),
                  self.hu.coriolis.data_ptr(), self.hv.coriolis.data_ptr(), None, float(self.fc), stream_ptr())

    def compute_fluxes(self):
        """``system.py:248-281`` (halo exchange is empty on one subdomain). With friction, the
        reference's second stepper call (Q4) follows it and its dt is the one the update uses;
        without, that call gives the driver's dt again and is skipped."""
        if self.Mann != 0:
            self.update_term_friction()
            self.stepper()
        self.update_ghost_values()
        self.explicit_convective()
        self.update_term_source()
        if self.fc != 0:
            self.update_term_coriolis()

    def compute_new_val(self):
        """``system.py:237-242`` -> update_SW."""
        if self.flux == 'srnh':
            return self._srnh_update(None)
        self._dt_dev.fill_(float(self.dt))
        _lib.call('mfv_update_sw', self._m.ptr, self.h.cell.data_ptr(), self.hu.cell.data_ptr(),
                  self.hv.cell.data_ptr(), self._rez.data_ptr(), self._src.data_ptr(), self._dt_dev.data_ptr(),
                  stream_ptr())

    def run(self, nsteps):
        """``nsteps`` x (stepper, compute_fluxes, compute_new_val), fused on the packed path with
        dt on the device. Returns the simulated time; ``self.dt`` holds the last dt."""
        if self.flux == 'srnh':
            return self._srnh_run(nsteps)
        if not self._zgrad_ready:
            self._bathymetry_gradient()
        import torch
        from ..device import DeviceMesh
        from ..renumber import renumbered
        m = self._m
        s = stream_ptr()
        if self._packed is None:
            # the packed records live in Hilbert-curve cell order (a relabelling: same bits)
            r = renumbered(self.domain)
            self._rmesh = DeviceMesh(r, lsq=True, device=m.device)
            self._perm = torch.as_tensor(r.perm, device=m.device)
            self._packed = SwPacked(self._rmesh)
        rm, p, perm = self._rmesh, self._packed, self._perm
        ins = [x.index_select(0, perm) for x in (self.h.cell, self.hu.cell, self.hv.cell, self.Z.cell)]
        zg = self._zgrad.index_select(1, perm).contiguous()
        _lib.call('mfv_sw_pack', rm.ptr, p.ptr, *[x.data_ptr() for x in ins], zg.data_ptr(), self.order, s)
        _lib.call('mfv_time_step_sw', m.ptr, self.h.cell.data_ptr(), self.hu.cell.data_ptr(), self.hv.cell.data_ptr(),
                  self.cfl, self.grav, 0., 0., self._clock.dt0.data_ptr(), None, s)
        self._clock.start()
        for k in range(nsteps):
            _lib.call('mfv_sw_recon', rm.ptr, p.ptr, self.order, s)
            _lib.call('mfv_sw_flux_packed', rm.ptr, p.ptr, self.grav, self.cfl, self._clock.ring.data_ptr(), k, s)
        _lib.call('mfv_sw_unpack', rm.ptr, p.ptr, ins[0].data_ptr(), ins[1].data_ptr(), ins[2].data_ptr(), s)
        for var, x in zip((self.h, self.hu, self.hv), ins[:3]):
            var.cell.index_copy_(0, perm, x)
        del zg
        self._clock.step = nsteps
        if nsteps:
            self.dt = float(self._clock.ring[(nsteps - 1) % 3])
        return self._clock.time

    def _srnh_run(self, nsteps):
        """The reference sequence per step with every dt kept on the device (one slot per
        step, read back once at the end)."""
        import torch
        if nsteps <= 0:
            return 0.
        if self.Mann != 0 or self.fc != 0:  # friction takes the dt on the host, as the reference does
            t = 0.
            for _ in range(nsteps):
                self.stepper()
                self.compute_fluxes()
                self.compute_new_val()
                t += float(self.dt)
            return t
        dts = torch.zeros(nsteps, dtype=torch.float64, device=self._m.device)
        for k in range(nsteps):
            dk = dts[k:k + 1]
            self._srnh_dt(dk, False)
            self.update_ghost_values()
            self._srnh_convective()
            self.update_term_source()
            self._srnh_update(dk)
        host = dts.cpu().numpy()
        self.dt = float(host[-1])
        t = 0.
        for x in host:
            t += float(x)
        return t
