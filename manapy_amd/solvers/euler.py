"""3D compressible Euler, first-order Rusanov (``solvers/euler/fvm_utils3d.py``).

The reference has kernels only, no solver class; ``EulerSolver3D`` composes them in data-
flow order: ghost_value_TubeSchok3D -> time_step_euler_3d (min over cells) ->
explicitscheme_euler_3d_rusanov -> update_euler_3d_fvc. ``quirk=True`` reproduces the
reference's two slips (Q1: left v, w in the right wave speed; Q2: slip ghost rhovg
overwritten, rhowg unset); ``quirk=False`` is the corrected scheme.
"""
import ctypes

import numpy as np

from .. import _lib
from ..base import Struct, make_get_conf, process_conf
from ..device import stream_ptr
from ._common import DeviceClock


def initialisation_euler_3d(center, rho1=1.0, rho2=0.125, P1=1.0, P2=0.1, u1=0., u2=0., xm=0.5, gamma=1.4):
    """``initialisation_euler_3d`` choix 0 (``fvm_utils3d.py:322-340``): Sod states split at x = xm.
    Returns U [6][C] = rho, P, rhou, rhov, rhow, rhoE (numpy)."""
    C = len(center)
    left = center[:, 0] < xm
    U = np.zeros((6, C))
    U[0] = np.where(left, rho1, rho2)
    U[2] = np.where(left, u1 * rho1, u2 * rho2)
    U[1] = np.where(left, P1, P2)
    U[5] = 0.5 * (U[2] ** 2 + U[3] ** 2 + U[4] ** 2) / U[0] + U[1] / (gamma - 1)
    return U


class EulerSolver3D(object):
    """Explicit first-order Rusanov Euler solver on a 3D tetrahedral Domain."""

    _parameters = [('cfl', float, 0.8, 0, 'cfl of the explicit scheme'),
                   ('gamma', float, 1.4, 0, 'ratio of specific heats'),
                   ('quirk', bool, False, 0, 'reproduce the reference slips Q1/Q2')]

    def __init__(self, domain, U0=None, conf=None, **kwargs):
        if conf is None:
            conf = Struct()
        self.conf = process_conf(self._parameters, conf, kwargs)
        get = make_get_conf(self.conf, kwargs)
        if int(domain.dim) != 3:
            raise ValueError("EulerSolver3D needs a 3D domain")
        self.domain = domain
        self.cfl = float(get('cfl'))
        self.gamma = float(get('gamma'))
        self.quirk = int(bool(get('quirk')))
        import torch
        m = domain.device_mesh
        C, F = int(domain.nbcells), int(domain.nbfaces)
        self.U = torch.zeros((6, C), dtype=torch.float64, device=m.device)
        if U0 is not None:
            self.U.copy_(torch.as_tensor(np.asarray(U0), device=m.device))
        self.G = torch.zeros((6, F), dtype=torch.float64, device=m.device)
        self.rez = torch.zeros((5, C), dtype=torch.float64, device=m.device)
        self._dt_dev = torch.zeros(1, dtype=torch.float64, device=m.device)
        self._clock = DeviceClock(m.device)
        self.dt = 0.

    rho = property(lambda self: self.U[0])
    P = property(lambda self: self.U[1])
    rhou = property(lambda self: self.U[2])
    rhov = property(lambda self: self.U[3])
    rhow = property(lambda self: self.U[4])
    rhoE = property(lambda self: self.U[5])

    @property
    def _m(self):
        return self.domain.device_mesh

    def update_ghost_values(self):
        _lib.call('mfv_euler3d_ghost', self._m.ptr, self.U.data_ptr(), self.G.data_ptr(), self.quirk, stream_ptr())

    def stepper(self):
        hd = ctypes.c_double(0.)
        _lib.call('mfv_euler3d_time_step', self._m.ptr, self.U.data_ptr(), self.cfl, self.gamma,
                  self._dt_dev.data_ptr(), ctypes.byref(hd), stream_ptr())
        self.dt = hd.value
        return self.dt

    def compute_fluxes(self):
        _lib.call('mfv_euler3d_rusanov', self._m.ptr, self.U.data_ptr(), self.G.data_ptr(), self.gamma, self.quirk,
                  self.rez.data_ptr(), stream_ptr())

    def compute_new_val(self):
        self._dt_dev.fill_(float(self.dt))
        _lib.call('mfv_euler3d_update', self._m.ptr, self.U.data_ptr(), self.rez.data_ptr(), self.gamma,
                  self._dt_dev.data_ptr(), stream_ptr())

    def step(self):
        """ghosts, dt, fluxes, update (the data-flow order of the reference kernels)."""
        self.update_ghost_values()
        self.stepper()
        self.compute_fluxes()
        self.compute_new_val()
        return self.dt

    def run(self, nsteps, packed=None):
        """``nsteps`` fused steps (ghosts computed on the fly) with dt on the device. Under
        ``quirk`` the unset rhowg of slip faces is taken as zero (zero-initialised ghosts).
        On tetrahedra (``packed`` None or True) the steps run on 64-byte cell records, on face
        tiles of 256 cells when their LDS fits (mfv_euler3d_step_packed_tiled, else
        mfv_euler3d_step_packed; same bits)."""
        import torch
        m = self._m
        s = stream_ptr()
        _lib.call('mfv_euler3d_time_step', m.ptr, self.U.data_ptr(), self.cfl, self.gamma, self._clock.dt0.data_ptr(),
                  None, s)
        self._clock.start()
        if packed is None:
            packed = m.MF == 4
        if packed:
            from ..device import EulerTet, EulerTetTiles
            et = getattr(m, '_euler_tet', None)
            if et is None:
                et = m._euler_tet = EulerTet(m)
            tt = getattr(m, '_euler_tiles', False)
            if tt is False:  # face tiles of 256 (or 128) cells when their LDS fits
                tt = None
                for T in (256, 128):
                    cand = EulerTetTiles(m, T)
                    if cand.lds_bytes() <= 64 * 1024:
                        tt = cand
                        break
                m._euler_tiles = tt
            a = torch.empty((int(m.C), 8), dtype=torch.float64, device=self.U.device)
            b = torch.empty_like(a)
            _lib.call('mfv_euler3d_pack', m.ptr, self.U.data_ptr(), a.data_ptr(), s)
            for k in range(nsteps):
                if tt is not None:
                    _lib.call('mfv_euler3d_step_packed_tiled', m.ptr, et.ptr, tt.ptr, a.data_ptr(), b.data_ptr(),
                              self.gamma, self.cfl, self.quirk, self._clock.ring.data_ptr(), k, s)
                else:
                    _lib.call('mfv_euler3d_step_packed', m.ptr, et.ptr, a.data_ptr(), b.data_ptr(), self.gamma,
                              self.cfl, self.quirk, self._clock.ring.data_ptr(), k, s)
                a, b = b, a
            _lib.call('mfv_euler3d_unpack', m.ptr, a.data_ptr(), self.U.data_ptr(), s)
        else:
            a = self.U
            b = torch.empty_like(a)
            for k in range(nsteps):
                _lib.call('mfv_euler3d_step_fused', m.ptr, a.data_ptr(), b.data_ptr(), self.gamma, self.cfl,
                          self.quirk, self._clock.ring.data_ptr(), k, s)
                a, b = b, a
            if nsteps % 2 == 1:
                self.U.copy_(a)
        if nsteps:
            self.dt = float(self._clock.ring[(nsteps - 1) % 3])
        return self._clock.time
