"""Continuous-batching service of implicit advection-diffusion solves.

A request carries a mesh (a ``Domain``), an initial field, a constant velocity and a
diffusivity D, and asks for ``nsteps`` time steps of

    u*      = u^n + dt * upwind_advection(u^n)        (explicitscheme_convective_2d, order 1)
    (I - dt D A) u^{n+1} = u*                         (diamond matrix A, get_triplet_2d)

with all-Neumann boundaries (so the Dirichlet lifting rhs0 is zero) and dt the CFL step
of the advection (constant, the velocity is constant). The implicit system is solved to
``||b - M x|| <= tol ||b||`` (default 1e-10) by the arena solver of ``krylov.EllBatch``.

``ServingEngine.run`` admits requests as they arrive while the arena has room for them
(HBM footprint: contiguous matrix slices plus the request's mesh tables, under a byte
budget), keeps every admitted request's state on the device between steps, and advances
all of them together: each engine tick launches the advection of the requests entering a
new step and one solver tick for every system in the arena, whatever stage it is at.
Requests therefore overlap on the device; each request's arithmetic is independent of the
others (per-request kernels, position-independent batched solver), so a result is the
same alone and in any mix.
"""
import ctypes
import time

import numpy as np

from . import _lib
from .device import DeviceMesh, require_gpu, stream_ptr
from .krylov import EllBatch

PHASE_WAIT, PHASE_ADVECT, PHASE_SOLVE, PHASE_DONE, PHASE_FAILED, PHASE_REJECTED = 0, 1, 2, 3, 4, 5


class AdvDiffRequest(object):
    """One solve. ``domain`` is a Domain, or the raw mesh ``(elements, nodes)`` in the partition
    layout: then the engine builds the Domain and its device tables when the request arrives
    (on its preparation threads), so the request's latency includes that preprocessing.
    ``u0`` is an array of cell values, or a callable of the cell centres."""

    def __init__(self, rid, domain, u0, velocity=(1.0, 0.5), D=1e-2, nsteps=20, arrival=0., cfl=0.8):
        self.rid = rid
        if isinstance(domain, tuple):
            self.mesh, self.domain = domain, None
            self.C = int(len(domain[0]))
        else:
            self.mesh, self.domain = None, domain
            self.C = int(domain.nbcells)
        self._u0 = u0
        self.u0 = None if callable(u0) else np.ascontiguousarray(u0, dtype=np.float64)
        self.velocity = (float(velocity[0]), float(velocity[1]))
        self.D = float(D)
        self.nsteps = int(nsteps)
        self.arrival = float(arrival)
        self.cfl = float(cfl)
        self._mesh_host = None
        self.t_prepared = None
        self.prepare_s = 0.
        # filled by the engine
        self.inst = None
        self.phase = PHASE_WAIT
        self.step = 0
        self.dt = None
        self.t_admit = self.t_done = None
        self.result = None
        self.iterations = 0
        self.error = None
        self._dev = None

    @property
    def failed(self):
        return self.phase in (PHASE_FAILED, PHASE_REJECTED)

    def prepare(self):
        """Host-side preprocessing: the Domain (when the request came as a raw mesh), then the
        tables of its GPU layout and diamond weights, so that admission only uploads them."""
        from .device import gpu_layout
        from .diamond import diamond_layout
        t0 = time.perf_counter()
        if self.domain is None:
            from .ddm import Domain
            self.domain = Domain(2, mesh=self.mesh)
            self.mesh = None
        if self.u0 is None:
            self.u0 = np.ascontiguousarray(self._u0(self.domain.cells.center), dtype=np.float64)
        self._mesh_host = gpu_layout(self.domain, lsq=False)
        self.domain._diamond_host = diamond_layout(self.domain)
        self.prepare_s = time.perf_counter() - t0
        return self

    @property
    def prepared(self):
        return self._mesh_host is not None

    def footprint(self, arena):
        """Bytes the request holds on the device while admitted."""
        slices = (self.C + 63) // 64
        F = int(self.domain.nbfaces)
        mesh = self.C * (8 * 4 + 4 * 4 + 2) + F * (4 * 3 + 8 * 8)   # mfv_mesh tables (no LSQ)
        fields = 8 * (3 * F + 3 * self.C) + 64
        return slices * arena.bytes_per_slice() + mesh + fields


class ServingEngine(object):
    def __init__(self, max_cells=4 << 20, max_inst=128, width=16, hbm_budget=None, n_iter=6, tol=1e-10,
                 tol_inner=1e-3, device=None, graph=True, prep_threads=8, time_ticks=False):
        torch = require_gpu()
        self.arena = EllBatch((int(max_cells) + 63) // 64, max_inst=max_inst, width=width, device=device)
        self.device = self.arena.device
        self.hbm_budget = int(hbm_budget) if hbm_budget else None
        self.n_iter = int(n_iter)
        self.graph = bool(graph)
        self.tol, self.tol_inner = float(tol), float(tol_inner)
        self.active = []
        self.in_use = 0
        self.ticks = 0
        self.max_active = 0
        self.prep_threads = int(prep_threads)
        self.time_ticks = bool(time_ticks)
        self._tick_events = []

    # -- admission --------------------------------------------------------------------------------
    def never_fits(self, req):
        """Why ``req`` can never be admitted by this engine (None if it can): its footprint
        exceeds the HBM budget, or its matrix exceeds the whole arena."""
        fp = req.footprint(self.arena)
        if self.hbm_budget is not None and fp > self.hbm_budget:
            return "footprint %d B exceeds the HBM budget %d B" % (fp, self.hbm_budget)
        if (req.C + 63) // 64 > self.arena.S:
            return "%d cells need %d slices, the arena has %d" % (req.C, (req.C + 63) // 64, self.arena.S)
        return None

    def can_admit(self, req):
        if req.footprint(self.arena) + self.in_use > (self.hbm_budget or 1 << 62):
            return False
        if not self.arena._free_ids:
            return False
        return self.arena._ranges.largest() >= (req.C + 63) // 64

    def admit(self, req, now):
        import torch
        inst = self.arena.alloc(req.C)
        if inst is None:
            raise RuntimeError("admit called without room")
        d = req.domain
        dm = d._device_mesh if getattr(d, '_device_mesh', None) is not None else \
            DeviceMesh(d, lsq=False, host=req._mesh_host)
        d._device_mesh = dm
        dev = self.device
        C, F = req.C, int(d.nbfaces)
        s = stream_ptr()
        uf = torch.full((F,), req.velocity[0], dtype=torch.float64, device=dev)
        vf = torch.full((F,), req.velocity[1], dtype=torch.float64, device=dev)
        wf = torch.zeros(F, dtype=torch.float64, device=dev)
        uc, vc, wc = (torch.zeros(C, dtype=torch.float64, device=dev) for _ in range(3))
        _lib.call('mfv_facetocell', dm.ptr, uf.data_ptr(), uc.data_ptr(), s)
        _lib.call('mfv_facetocell', dm.ptr, vf.data_ptr(), vc.data_ptr(), s)
        dt_dev = torch.zeros(1, dtype=torch.float64, device=dev)
        h = ctypes.c_double(0.)
        _lib.call('mfv_time_step', dm.ptr, uc.data_ptr(), vc.data_ptr(), wc.data_ptr(), req.cfl, dt_dev.data_ptr(),
                  ctypes.byref(h), s)
        req.dt = h.value
        ring = torch.zeros(4, dtype=torch.float64, device=dev)
        _lib.call('mfv_dt_ring_init', ring.data_ptr(), dt_dev.data_ptr(), s)
        # matrix of the all-Neumann diamond operator, then the geometry tables are dropped
        dd = d.device_diamond
        N = int(d.nbnodes)
        node_dir = torch.zeros(N, dtype=torch.uint8, device=dev)
        rank = torch.full((F,), -1, dtype=torch.int32, device=dev)
        self.arena.assemble_diamond(inst, dd, node_dir, rank)
        d._device_diamond = None
        row0, _, ns = self.arena.rows(inst)
        self.arena.x[row0:row0 + 64 * ns].zero_()
        self.arena.b[row0:row0 + 64 * ns].zero_()
        self.arena.x[row0:row0 + C] = torch.from_numpy(req.u0).to(dev)
        self.arena.configure(inst, 1., -req.dt * req.D, tol=self.tol, tol_in=self.tol_inner)
        req._dev = {'mesh': dm, 'uf': uf, 'vf': vf, 'wf': wf, 'uc': uc, 'vc': vc, 'wc': wc, 'ring': ring,
                    'row0': row0}
        req.inst = inst
        req.phase = PHASE_ADVECT
        req.t_admit = now
        self.in_use += req.footprint(self.arena)
        self.active.append(req)

    def _advect(self, req):
        """u* = u^n + dt * convective(u^n) into b, initial guess x = u*, status START."""
        a = self.arena
        g = req._dev
        row0, C = g['row0'], req.C
        xs, bs = a.x[row0:row0 + C], a.b[row0:row0 + C]
        _lib.call('mfv_advection_step', g['mesh'].ptr, xs.data_ptr(), bs.data_ptr(), g['uc'].data_ptr(),
                  g['vc'].data_ptr(), g['wc'].data_ptr(), g['uf'].data_ptr(), g['vf'].data_ptr(), g['wf'].data_ptr(),
                  req.cfl, 1, g['ring'].data_ptr(), req.step, None, stream_ptr())
        xs.copy_(bs)
        a.status[req.inst] = _lib.KS_START
        req.phase = PHASE_SOLVE

    def _finish(self, req, now, error=None):
        """Retire ``req``: its result (or ``error`` if its solve failed), then its arena
        instance and device tables are released so that the other requests go on."""
        row0 = req._dev['row0']
        if error is None:
            req.result = self.arena.x[row0:row0 + req.C].cpu().numpy().copy()
        req.iterations = int(self.arena.scalars(req.inst)[_lib.KS_ITTOT])
        self.arena.release(req.inst)
        self.arena.publish()
        self.in_use -= req.footprint(self.arena)
        req._dev = None
        req.domain._device_mesh = None
        req.phase = PHASE_DONE if error is None else PHASE_FAILED
        req.error = error
        req.t_done = now
        self.active.remove(req)

    # -- the loop -------------------------------------------------------------------------------
    def run(self, requests, clock=None, max_ticks=10 ** 7):
        """Serve ``requests`` (by arrival time, seconds from the start); returns them with
        t_admit / t_done set and ``result`` filled.

        A request that is not prepared yet is prepared on one of ``prep_threads`` host
        threads from its arrival on (the Domain of a raw mesh, the layout tables), while the
        GPU serves the others; it is admitted, in arrival order, once prepared and once the
        arena and the HBM budget have room. A request whose solve fails (status FAILED: no
        convergence within the outer-iteration limit, or a non-finite residual) is returned
        with ``phase == PHASE_FAILED`` and ``error`` set; its instance is released and the
        other requests go on unchanged. A request that can never be admitted (``never_fits``)
        is returned with ``phase == PHASE_REJECTED``."""
        import torch
        from concurrent.futures import ThreadPoolExecutor, wait, FIRST_COMPLETED
        pending = sorted(requests, key=lambda r: r.arrival)
        clock = clock or time.perf_counter
        pool = ThreadPoolExecutor(max(1, self.prep_threads))
        futures = {}
        t0 = clock()
        done = []

        def reject(r, why):
            r.phase, r.error = PHASE_REJECTED, "rejected: " + why
            r.t_admit = r.t_done = clock() - t0
            done.append(r)

        try:
            for _ in range(max_ticks):
                now = clock() - t0
                for r in pending:  # preparation starts at arrival
                    if r.arrival > now:
                        break
                    if id(r) not in futures and not r.prepared:
                        futures[id(r)] = pool.submit(r.prepare)
                admitted = False
                while pending and pending[0].arrival <= now:
                    r = pending[0]
                    f = futures.get(id(r))
                    if f is not None and not f.done():
                        break
                    if f is not None and f.exception() is not None:
                        pending.pop(0)
                        reject(r, "preparation failed: %s" % f.exception())
                        continue
                    r.t_prepared = clock() - t0
                    why = self.never_fits(r)
                    if why is not None:
                        pending.pop(0)
                        reject(r, why)
                        continue
                    if not self.can_admit(r):
                        if not self.active:
                            raise _lib.MfvError("request %s cannot be admitted into an empty arena" % r.rid)
                        break
                    self.admit(pending.pop(0), now)
                    admitted = True
                if admitted:
                    self.arena.publish()
                    self.max_active = max(self.max_active, len(self.active))
                if not self.active:
                    if not pending:
                        break
                    wait_s = pending[0].arrival - (clock() - t0)
                    f = futures.get(id(pending[0]))
                    if f is not None and not f.done():
                        wait([f], timeout=0.05, return_when=FIRST_COMPLETED)
                    elif wait_s > 0:
                        time.sleep(min(wait_s, 0.05))
                    continue
                for r in self.active:
                    if r.phase == PHASE_ADVECT:
                        self._advect(r)
                if self.time_ticks:
                    ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                    ev[0].record()
                self.arena.tick(self.n_iter, graph=self.graph)
                if self.time_ticks:
                    ev[1].record()
                    self._tick_events.append(ev)
                self.ticks += 1
                st = self.arena.status.cpu().numpy()
                now = clock() - t0
                for r in list(self.active):
                    s = st[r.inst]
                    if s == _lib.KS_DONE:
                        r.step += 1
                        if r.step == r.nsteps:
                            self._finish(r, now)
                            done.append(r)
                        else:
                            r.phase = PHASE_ADVECT
                    elif s == _lib.KS_FAILED:
                        self._finish(r, now, error="implicit solve failed at step %d" % r.step)
                        done.append(r)
            else:
                raise _lib.MfvError("serving loop did not finish in %d ticks" % max_ticks)
        finally:
            pool.shutdown(wait=True)
        torch.cuda.synchronize()
        return done

    def tick_seconds(self):
        """Device time spent in solver ticks (time_ticks=True), from HIP events around each."""
        import torch
        torch.cuda.synchronize()
        return sum(a.elapsed_time(b) for a, b in self._tick_events) * 1e-3


def latency_stats(done):
    """Latency from arrival to result of the completed requests (preparation included), and
    its preparation part."""
    ok = [r for r in done if r.phase == PHASE_DONE]
    lat = np.array([r.t_done - r.arrival for r in ok])
    prep = np.array([r.prepare_s for r in ok])
    return {'p50_s': float(np.percentile(lat, 50)), 'p99_s': float(np.percentile(lat, 99)),
            'mean_s': float(lat.mean()), 'max_s': float(lat.max()),
            'prepare_p50_s': float(np.percentile(prep, 50)), 'prepare_p99_s': float(np.percentile(prep, 99))}
