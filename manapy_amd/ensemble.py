"""Ensembles of independent shallow-water instances advanced together on one GPU.

The instances' meshes are concatenated into one device mesh (block-diagonal: every index
of instance i is offset by its first cell / face), each instance starting at a multiple of
64 cells so that every wavefront belongs to one instance. One launch per stage then
advances all instances (the mixed-precision reconstruction and flux kernels), with a dt
ring per instance (``mfv_sw_batch_*`` in include/mfv.h); a step sequence is captured once
into a HIP graph and replayed. ``autotune`` times the thread-block sizes of the two cell
kernels on the actual ensemble and keeps the fastest.

Each instance's arithmetic is that of the single-mesh mixed path and does not depend on
the other instances (cell-local gathers, exact per-instance min for dt), so an instance
gives the same bits alone and inside any ensemble. By default each instance's cells are
stored in Hilbert-curve order on the device (``renumber``); states go in and come out in
the instance's own numbering, and the renumbering does not change any bit.
"""
import ctypes
import json
import os
import time

import numpy as np

from . import _lib
from .device import gpu_layout, require_gpu, stream_ptr, sw_packed_layout

BLOCKS = (64, 128, 256, 512)
TILES = (256, 512)
# launch configurations the autotuner compares: the two-kernel step with its block size, the
# fused tiled step with its tile size, and the fused step with every cell state a tile reads
# staged in LDS
CONFIGS = (tuple(('split', b, b) for b in BLOCKS) + tuple(('tiled', t, t) for t in TILES) +
           (('staged', 128, 128), ('staged', 256, 256), ('staged', 512, 512)))


def _config(c):
    """(kind, block) or (kind, block, cells per tile) -> (kind, block, cells per tile)."""
    c = tuple(c)
    return c if len(c) == 3 else (c[0], int(c[1]), int(c[1]))


def _cat_cols(arrs, pads, fill, offsets=None):
    """Concatenate column-major [k][C_i] tables along cells with per-instance padding."""
    rows = max(a.shape[0] for a in arrs)
    out = []
    for i, (a, pad) in enumerate(zip(arrs, pads)):
        t = np.full((rows, a.shape[1] + pad), fill, dtype=a.dtype)
        v = a
        if offsets is not None:
            v = np.where(a >= 0, a + offsets[i], a)
        t[:a.shape[0], :a.shape[1]] = v
        out.append(t)
    return np.ascontiguousarray(np.concatenate(out, axis=1))


def _cat_rows(arrs, pads, fill):
    out = []
    for a, pad in zip(arrs, pads):
        t = np.full((a.shape[0] + pad,) + a.shape[1:], fill, dtype=a.dtype)
        t[:a.shape[0]] = a
        out.append(t)
    return np.ascontiguousarray(np.concatenate(out, axis=0))


def ensemble_layout(domains):
    """Host tables of the concatenated mesh, mixed-precision records and batch descriptor."""
    Ls = [gpu_layout(d, lsq=True) for d in domains]
    Ps = [sw_packed_layout(d) for d in domains]
    MF = Ls[0]['MF']
    if any(L['MF'] != MF for L in Ls):
        raise ValueError("all instances must have the same number of faces per cell")
    C = np.array([L['C'] for L in Ls], dtype=np.int64)
    F = np.array([L['F'] for L in Ls], dtype=np.int64)
    Cp = (C + 63) // 64 * 64
    pads = (Cp - C).tolist()
    c0 = np.concatenate([[0], np.cumsum(Cp)[:-1]])
    f0 = np.concatenate([[0], np.cumsum(F)[:-1]])
    if int(Cp.sum()) * 8 >= 2 ** 31:
        raise ValueError("ensemble too large for int32 neighbour codes")
    out = {'MF': MF, 'C': int(Cp.sum()), 'F': int(F.sum()), 'dim': 2}
    fc = [np.where(L['face_cells'] >= 0, L['face_cells'] + c0[i], -1) for i, L in enumerate(Ls)]
    out['face_cells'] = np.ascontiguousarray(np.concatenate(fc), dtype=np.int32)
    for k in ('face_name', 'face_normal', 'face_mesure', 'face_center', 'face_ghost'):
        out[k] = np.ascontiguousarray(np.concatenate([L[k] for L in Ls]))
    out['cell_center'] = _cat_rows([L['cell_center'] for L in Ls], pads, 0.)
    out['cell_volume'] = _cat_rows([L['cell_volume'] for L in Ls], pads, 1.)
    out['cell_nfaces'] = _cat_rows([L['cell_nfaces'] for L in Ls], pads, 0).astype(np.int32)
    out['lsq_ncells'] = _cat_rows([L['lsq_ncells'] for L in Ls], pads, 0).astype(np.int32)
    out['lsq_nghosts'] = _cat_rows([L['lsq_nghosts'] for L in Ls], pads, 0).astype(np.int32)
    out['cell_faces'] = _cat_cols([L['cell_faces'] for L in Ls], pads, -1, f0).astype(np.int32)
    out['order_split'] = _cat_cols([L['order_split'] for L in Ls], pads, 0)
    out['order_sorted'] = _cat_cols([L['order_sorted'] for L in Ls], pads, 0)
    out['lsq_cells'] = _cat_cols([L['lsq_cells'] for L in Ls], pads, 0, c0).astype(np.int32)
    out['lsq_ghosts'] = _cat_cols([L['lsq_ghosts'] for L in Ls], pads, 0, f0).astype(np.int32)
    out['K'] = out['lsq_cells'].shape[0]
    out['G'] = out['lsq_ghosts'].shape[0]
    out['bnd_faces'] = np.ascontiguousarray(np.concatenate([L['bnd_faces'] + f0[i] for i, L in enumerate(Ls)]),
                                            dtype=np.int32)
    # neighbour codes (other << 3 | slot << 1 | side): shift the cell part by the instance offset
    out['nbr'] = _cat_cols([np.where(P['nbr'] >= 0, P['nbr'] + (c0[i] << 3), -1).astype(np.int32)
                            for i, P in enumerate(Ps)], pads, -1)
    out['order'] = _cat_rows([P['order'] for P in Ps], pads, 0)
    out['wave_inst'] = np.repeat(np.arange(len(domains), dtype=np.int32), Cp // 64)
    out['inst_cell0'] = c0.astype(np.int32)
    out['inst_ncells'] = C.astype(np.int32)
    out['inst_faces0'] = f0.astype(np.int64)
    out['inst_nfaces'] = F
    return out


def ensemble_tiles(host, T=256, N=None):
    """Tiles of the fused step for an ensemble layout: runs of at most N (default T) cells,
    for blocks of T threads, that start at an instance's first cell and never cross into the
    next instance."""
    from .device import sw_tiles_native
    N = T if N is None else int(N)
    npad = (host['inst_ncells'].astype(np.int64) + 63) // 64 * 64
    return sw_tiles_native(host['nbr'], host['C'], T, N, host, (host['inst_cell0'], npad))


class SwEnsemble(object):
    """Device state of an ensemble; ``states`` is a list of (h, hu, hv, Z) numpy arrays."""

    def __init__(self, domains, states, grav=9.81, cfl=0.8, order=2, device=None, renumber=True):
        torch = require_gpu()
        self.device = torch.device(device if device is not None else 'cuda')
        if renumber:
            from .renumber import renumbered
            self.views = [renumbered(d) for d in domains]
            domains = self.views
            states = [tuple(v.to_new(np.asarray(a)) for a in st) for v, st in zip(self.views, states)]
        else:
            self.views = None
        host = ensemble_layout(domains)
        self.host = host
        self.n = len(domains)
        self.C, self.F, self.MF = host['C'], host['F'], host['MF']
        self.grav, self.cfl, self.order = float(grav), float(cfl), int(order)
        dev = self.device
        self.t = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in host.items()
                  if isinstance(v, np.ndarray) and k not in ('inst_faces0', 'inst_nfaces')}
        m = _lib.MfvMesh()
        m.dim, m.nb_cells, m.nb_faces, m.max_cell_faces = 2, self.C, self.F, self.MF
        m.lsq_width, m.ghost_width = host['K'], host['G']
        for k in ('face_cells', 'face_name', 'face_normal', 'face_mesure', 'face_center', 'face_ghost', 'cell_center',
                  'cell_volume', 'cell_faces', 'cell_nfaces', 'order_split', 'order_sorted', 'lsq_cells',
                  'lsq_ncells', 'lsq_ghosts', 'lsq_nghosts', 'bnd_faces'):
            setattr(m, k, self.t[k].data_ptr())
        m.nb_bnd = len(host['bnd_faces'])
        self.mesh = m
        f32 = torch.float32
        self.state = torch.zeros((self.C, 4), dtype=torch.float64, device=dev)
        self.state2 = torch.zeros((self.C, 4), dtype=torch.float64, device=dev)
        self.facerec = torch.zeros((self.C, self.MF, 4), dtype=f32, device=dev)
        self.geoA = torch.zeros((self.C, self.MF, 4), dtype=f32, device=dev)
        self.geoB = torch.zeros((self.C, self.MF, 4), dtype=f32, device=dev)
        p = _lib.MfvSwMixed()
        p.nb_cells, p.max_cell_faces = self.C, self.MF
        p.state, p.vol, p.facerec = self.state.data_ptr(), self.t['cell_volume'].data_ptr(), self.facerec.data_ptr()
        p.geoA, p.geoB = self.geoA.data_ptr(), self.geoB.data_ptr()
        p.nbr, p.order = self.t['nbr'].data_ptr(), self.t['order'].data_ptr()
        self.mixed = p
        self.rings = torch.zeros((self.n, 4), dtype=torch.float64, device=dev)
        self.sums_out = torch.zeros((self.n, 3), dtype=torch.float64, device=dev)
        b = _lib.MfvSwBatch()
        b.nb_inst, b.nb_waves = self.n, self.C // 64
        b.wave_inst, b.inst_cell0 = self.t['wave_inst'].data_ptr(), self.t['inst_cell0'].data_ptr()
        b.inst_ncells, b.rings = self.t['inst_ncells'].data_ptr(), self.rings.data_ptr()
        self.batch = b
        self.config = ('staged', 256, 256)
        self._tiles = {}
        self.step_index = 0
        self._graph = None
        self._graph_key = None
        self._pack(states)

    # -- helpers ----------------------------------------------------------------------------------
    def _cat_cells(self, arrays):
        out = np.zeros(self.C)
        for i, a in enumerate(arrays):
            c0, n = int(self.host['inst_cell0'][i]), int(self.host['inst_ncells'][i])
            out[c0:c0 + n] = a
        return out

    def _pack(self, states):
        import torch
        dev = self.device
        h, hu, hv, Z = (torch.from_numpy(self._cat_cells([s[k] for s in states])).to(dev) for k in range(4))
        zg = torch.zeros((3, self.C), dtype=torch.float64, device=dev)
        zghost = torch.zeros(self.F, dtype=torch.float64, device=dev)
        s = stream_ptr()
        mp = ctypes.byref(self.mesh)
        bf = self.t['bnd_faces']
        _lib.call('mfv_ghost_value', mp, 0, Z.data_ptr(), zghost.data_ptr(), bf.data_ptr(), int(bf.numel()), s)
        _lib.call('mfv_cell_gradient', mp, Z.data_ptr(), zghost.data_ptr(), zg[0].data_ptr(), zg[1].data_ptr(),
                  zg[2].data_ptr(), s)
        _lib.call('mfv_sw_pack_mixed', mp, ctypes.byref(self.mixed), h.data_ptr(), hu.data_ptr(), hv.data_ptr(),
                  Z.data_ptr(), zg.data_ptr(), self.order, s)
        _lib.call('mfv_sw_batch_dt_init', mp, ctypes.byref(self.mixed), ctypes.byref(self.batch), self.grav, self.cfl,
                  s)
        self._keep = (h, hu, hv, Z, zg, zghost)
        torch.cuda.synchronize()
        self._keep = None
        self.step_index = 0

    # -- stepping -------------------------------------------------------------------------------
    @property
    def block(self):
        return self.config[1]

    def tiles(self, T, stage=False, N=None):
        N = T if N is None else N
        if (T, N, stage) not in self._tiles:
            from .device import SwTiles
            lay = next((v.host for (t, n, _), v in self._tiles.items() if (t, n) == (T, N)), None)
            self._tiles[(T, N, stage)] = SwTiles(lay if lay is not None else ensemble_tiles(self.host, T, N),
                                                 self.device, stage)
        return self._tiles[(T, N, stage)]

    def _swap(self):
        self.state, self.state2 = self.state2, self.state
        self.mixed.state = self.state.data_ptr()

    def step(self, nsteps=1, config=None):
        """Eager steps with ``config`` = ('split', block), ('tiled', T[, N]) or ('staged', T[, N])
        (default: self.config)."""
        kind, size, ncell = _config(config or self.config)
        s = stream_ptr()
        for _ in range(nsteps):
            if kind == 'split':
                _lib.call('mfv_sw_batch_step', ctypes.byref(self.mesh), ctypes.byref(self.mixed),
                          ctypes.byref(self.batch), self.grav, self.cfl, self.order, self.step_index, int(size), s)
            else:
                _lib.call('mfv_sw_batch_step_tiled', ctypes.byref(self.mesh), ctypes.byref(self.mixed),
                          ctypes.byref(self.batch), self.tiles(int(size), kind == 'staged', int(ncell)).ptr,
                          self.state2.data_ptr(), self.grav,
                          self.cfl, self.order, self.step_index, s)
                self._swap()
            self.step_index += 1

    def graph(self, nsteps, config=None):
        """The captured graph of ``nsteps`` steps (a multiple of 6), built on first use. A
        replay ends in the state buffer it started from."""
        key = (int(nsteps),) + _config(config or self.config) + (self.state.data_ptr(),)
        if self._graph is None or self._graph_key != key:
            self.close_graph()
            h = ctypes.c_void_p()
            if key[1] == 'split':
                _lib.call('mfv_sw_batch_graph', ctypes.byref(self.mesh), ctypes.byref(self.mixed),
                          ctypes.byref(self.batch), self.grav, self.cfl, self.order, key[0], key[2], ctypes.byref(h))
            else:
                _lib.call('mfv_sw_batch_graph_tiled', ctypes.byref(self.mesh), ctypes.byref(self.mixed),
                          ctypes.byref(self.batch), self.tiles(key[2], key[1] == 'staged', key[3]).ptr,
                          self.state2.data_ptr(), self.grav,
                          self.cfl, self.order, key[0], ctypes.byref(h))
            self._graph, self._graph_key = h, key
            self._graph_state = self.state.data_ptr()
        return self._graph

    def run_graph(self, nsteps, chunk=30, config=None):
        """Advance ``nsteps`` steps by replaying a ``chunk``-step graph (steps must stay
        aligned with the 3-slot dt ring and the two state buffers)."""
        if self.step_index % 6 or nsteps % chunk or chunk % 6:
            raise ValueError("graph replays need step index, nsteps and chunk multiples of 6 (and of chunk)")
        g = self.graph(chunk, config)
        if self._graph_state != self.state.data_ptr():
            raise RuntimeError("the graph was captured on the other state buffer")
        s = stream_ptr()
        for _ in range(nsteps // chunk):
            _lib.call('mfv_graph_launch', g, s)
            self.step_index += chunk

    def close_graph(self):
        """Release the captured steps (after the device has finished any replay of them)."""
        if self._graph is not None:
            import torch
            torch.cuda.synchronize()
            _lib.call('mfv_graph_destroy', self._graph)
            self._graph = None

    def __del__(self):
        try:
            self.close_graph()
        except Exception:
            pass

    # -- results --------------------------------------------------------------------------------
    def conservation(self):
        _lib.call('mfv_sw_batch_conservation', ctypes.byref(self.mesh), ctypes.byref(self.mixed),
                  ctypes.byref(self.batch), self.sums_out.data_ptr(), stream_ptr())
        return self.sums_out.cpu().numpy().copy()

    def instance_state(self, i):
        c0, n = int(self.host['inst_cell0'][i]), int(self.host['inst_ncells'][i])
        st = self.state[c0:c0 + n].cpu().numpy()  # h, hu, hv, Z
        out = (st[:, 0].copy(), st[:, 1].copy(), st[:, 2].copy())
        if self.views is not None:
            out = tuple(self.views[i].to_old(a) for a in out)
        return out

    def times(self):
        return self.rings[:, 3].cpu().numpy().copy()

    # -- tuning ---------------------------------------------------------------------------------
    def autotune(self, steps=6, cache=None):
        """Times every launch configuration (CONFIGS) on ``steps`` steps of this ensemble (state
        restored after) and keeps the fastest; results are cached by (cells, instances) in
        ``cache`` (JSON)."""
        import torch
        key = "sw_ensemble:C=%d:I=%d:MF=%d" % (self.C, self.n, self.MF)
        path = cache or os.environ.get('MFV_TUNE_CACHE')
        table = {}
        if path and os.path.exists(path):
            with open(path) as fh:
                table = json.load(fh)
        if key in table:
            self.config = _config(table[key]['config'])
            return table[key]
        saved = (self.state.clone(), self.rings.clone(), self.step_index)
        st = torch.cuda.current_stream()
        res = {}
        for cfg in CONFIGS:
            if cfg[0] != 'split' and self.tiles(cfg[1], cfg[0] == 'staged', cfg[2]).lds_bytes() > 64 * 1024:
                continue  # this mesh's tiles do not fit the LDS budget at this size
            self.step(6, cfg)  # warm-up, aligned to the ring period and the state buffers
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(st)
            self.step(steps, cfg)
            e1.record(st)
            e1.synchronize()
            res[cfg] = e0.elapsed_time(e1) / steps
        self.state.copy_(saved[0])
        self.rings.copy_(saved[1])
        self.step_index = saved[2]
        best = min(res, key=res.get)
        self.config = best
        entry = {'config': list(best), 'ms_per_step': {"%s:%d:%d" % k: v for k, v in res.items()},
                 'when': time.time()}
        if path:
            table[key] = entry
            os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
            with open(path, 'w') as fh:
                json.dump(table, fh, indent=1)
        return entry


def random_instance(rng, ncells_target, lx_base=100.):
    """A jittered rectangle of about ``ncells_target`` triangles, aspect ratio U(1, 4), with a
    dam at x = U(.3, .7) Lx (h = 2 upstream, 1 downstream) over 1-10 Gaussian bumps."""
    from .meshgen import jittered_triangles, gaussian_bumps
    aspect = rng.uniform(1., 4.)
    ny = max(2, int(round(np.sqrt(ncells_target / (2. * aspect)))))
    nx = max(2, int(round(aspect * ny)))
    ly = lx_base
    lx = lx_base * nx / ny
    el, nd = jittered_triangles(nx, ny, lx, ly, 0.1, int(rng.integers(1 << 30)))
    return el, nd, lx, ly, rng


def instance_state(domain, lx, ly, rng):
    from .meshgen import gaussian_bumps
    c = domain.cells.center
    Z = gaussian_bumps(c, int(rng.integers(1, 11)), rng, [(0., lx), (0., ly)])
    xd = rng.uniform(.3, .7) * lx
    h = np.where(c[:, 0] < xd, 2.0, 1.0)
    C = len(h)
    return h, np.zeros(C), np.zeros(C), Z
