"""Arena of block-diagonal sparse systems solved on the GPU (``mfv_krylov_tick``).

An ``EllBatch`` owns one ELL matrix of ``max_slices`` 64-row slices and the vectors
of the mixed-precision solver. Independent systems ``(sigma I + tau A) x = b``
occupy contiguous slice ranges; ``alloc`` / ``release`` manage those ranges (first
fit) and ``publish`` uploads the live-slice and live-instance lists that the
persistent kernels walk. Because every device pointer and grid size stays fixed,
one tick is valid for any set of instances at any stage of their solve, and each
instance's arithmetic is independent of the others (see mfv_krylov.hip).

The solver replaces the factorisation of ``ScipySolver`` / ``MUMPSSolver``
(``solvers/ls/system.py:236-379, 624-744``): FP32 BiCGStab with Jacobi scaling
inside FP64 iterative refinement, stopped at ``||b - A x|| <= tol ||b||``.
"""
import ctypes

import numpy as np

from . import _lib
from .device import require_gpu, stream_ptr


def _pow2(n):
    return 1 << max(0, int(n) - 1).bit_length()


class RangeAllocator(object):
    """First-fit allocator of contiguous ranges of ``[0, size)`` with coalescing frees."""

    def __init__(self, size):
        self.size = int(size)
        self.free = [(0, self.size)]

    def alloc(self, n):
        for k, (s0, m) in enumerate(self.free):
            if m >= n:
                if m == n:
                    self.free.pop(k)
                else:
                    self.free[k] = (s0 + n, m - n)
                return s0
        return None

    def release(self, s0, n):
        self.free.append((s0, n))
        self.free.sort()
        merged = []
        for s, m in self.free:
            if merged and merged[-1][0] + merged[-1][1] == s:
                merged[-1] = (merged[-1][0], merged[-1][1] + m)
            else:
                merged.append((s, m))
        self.free = merged

    def available(self):
        return sum(m for _, m in self.free)

    def largest(self):
        return max((m for _, m in self.free), default=0)


class EllBatch(object):
    def __init__(self, max_slices, max_inst=64, width=32, device=None):
        torch = require_gpu()
        if width < 1 or width > 128:  # the diamond assembly holds 32 columns per row in 2D, 128 in 3D
            raise ValueError("ELL width must be in [1, 128]")
        self.S, self.I, self.W = int(max_slices), int(max_inst), int(width)
        self.device = torch.device(device if device is not None else 'cuda')
        S, I, W = self.S, self.I, self.W
        dev = self.device
        i32, f32, f64 = torch.int32, torch.float32, torch.float64
        self.slice_inst = torch.full((S,), -1, dtype=i32, device=dev)
        self.slice_w = torch.zeros(S, dtype=i32, device=dev)
        self.cols = torch.zeros(S * W * 64, dtype=i32, device=dev)
        self.cols16 = torch.zeros(S * W * 64, dtype=torch.int16, device=dev)
        self.slice_narrow = torch.zeros(S, dtype=torch.uint8, device=dev)
        self.vals = torch.zeros(S * W * 64, dtype=f64, device=dev)
        self.vals32 = torch.zeros(S * W * 64, dtype=f32, device=dev)
        self.diag = torch.zeros(S * 64, dtype=f64, device=dev)
        self.inst_slice0 = torch.zeros(I, dtype=i32, device=dev)
        self.inst_nslices = torch.zeros(I, dtype=i32, device=dev)
        self.live_slices = torch.zeros(S, dtype=i32, device=dev)
        self.n_live_slices = torch.zeros(1, dtype=i32, device=dev)
        self.live_inst = torch.zeros(I, dtype=i32, device=dev)
        self.n_live_inst = torch.zeros(1, dtype=i32, device=dev)
        R = S * 64
        self.x = torch.zeros(R, dtype=f64, device=dev)
        self.b = torch.zeros(R, dtype=f64, device=dev)
        self.f32 = {k: torch.zeros(R, dtype=f32, device=dev) for k in ('d', 'r', 'rh', 'p', 'v', 'yz', 't', 'dinv')}
        self.partial = torch.zeros(S * 4, dtype=f64, device=dev)
        self.scal = torch.zeros(I * _lib.KS_NSCAL, dtype=f64, device=dev)
        self.status = torch.zeros(I, dtype=i32, device=dev)

        e = _lib.MfvEllBatch()
        e.nb_slices, e.width, e.nb_inst = S, W, I
        for k in ('slice_inst', 'slice_w', 'cols', 'vals', 'vals32', 'diag', 'inst_slice0', 'inst_nslices',
                  'live_slices', 'n_live_slices', 'live_inst', 'n_live_inst', 'cols16', 'slice_narrow'):
            setattr(e, k, getattr(self, k).data_ptr())
        w = _lib.MfvKrylovWork()
        w.x, w.b = self.x.data_ptr(), self.b.data_ptr()
        for k, t in self.f32.items():
            setattr(w, k, t.data_ptr())
        w.partial, w.scal, w.status = self.partial.data_ptr(), self.scal.data_ptr(), self.status.data_ptr()
        self._e, self._w = e, w
        self._ranges = RangeAllocator(S)
        self._inst = {}                # inst -> (slice0, nslices, nrows)
        self._free_ids = list(range(I - 1, -1, -1))

    # -- arena ----------------------------------------------------------------------------------
    @property
    def eptr(self):
        return ctypes.byref(self._e)

    @property
    def wptr(self):
        return ctypes.byref(self._w)

    def bytes_per_slice(self):
        """HBM bytes one slice of an instance occupies (matrix slots, vectors, partials)."""
        return self.W * 64 * (4 + 2 + 8 + 4) + 64 * (8 + 8 + 8 + 8 * 4) + 4 * 8 + 8 + 1

    def free_slices(self):
        return self._ranges.available()

    def alloc(self, nrows):
        """Reserve ceil(nrows/64) contiguous slices; returns the instance id or None."""
        ns = (int(nrows) + 63) // 64
        if ns == 0 or not self._free_ids:
            return None
        s0 = self._ranges.alloc(ns)
        if s0 is None:
            return None
        inst = self._free_ids.pop()
        self._inst[inst] = (s0, ns, int(nrows))
        self.slice_inst[s0:s0 + ns] = inst
        self.slice_w[s0:s0 + ns] = 0
        self.slice_narrow[s0:s0 + ns] = 0  # int32 columns until narrow_columns has checked this matrix
        self.status[inst] = _lib.KS_IDLE
        return inst

    def release(self, inst):
        s0, ns, _ = self._inst.pop(inst)
        self.slice_inst[s0:s0 + ns] = -1
        self.status[inst] = _lib.KS_IDLE
        self._free_ids.append(inst)
        self._ranges.release(s0, ns)

    def rows(self, inst):
        s0, ns, n = self._inst[inst]
        return s0 * 64, n, ns

    def publish(self):
        """Upload the live lists (ascending slice order; instances by id)."""
        import torch
        insts = sorted(self._inst)
        slices = [np.arange(self._inst[i][0], self._inst[i][0] + self._inst[i][1], dtype=np.int32) for i in insts]
        sl = np.concatenate(slices) if slices else np.zeros(0, np.int32)
        sl.sort()
        if len(sl):
            self.live_slices[:len(sl)] = torch.from_numpy(sl).to(self.device)
        self.n_live_slices.fill_(len(sl))
        if insts:
            self.live_inst[:len(insts)] = torch.tensor(insts, dtype=torch.int32, device=self.device)
            s0 = torch.tensor([self._inst[i][0] for i in insts], dtype=torch.int32, device=self.device)
            ns = torch.tensor([self._inst[i][1] for i in insts], dtype=torch.int32, device=self.device)
            idx = torch.tensor(insts, dtype=torch.long, device=self.device)
            self.inst_slice0[idx] = s0
            self.inst_nslices[idx] = ns
        self.n_live_inst.fill_(len(insts))
        self._n_live = (len(sl), len(insts))

    # -- matrix and system ------------------------------------------------------------------------
    def assemble_diamond(self, inst, ddia, node_dir, face_dir_rank):
        """Writes the diamond matrix of a DeviceDiamond into the instance's slices
        (get_triplet_2d, ast/functions2d.py:328-580; get_triplet_3d, ast/functions3d.py:498-782)."""
        import torch
        row0, n, ns = self.rows(inst)
        if n != ddia.dmesh.C:
            raise ValueError("instance has %d rows, mesh %d cells" % (n, ddia.dmesh.C))
        s0 = row0 // 64
        info = torch.zeros(2, dtype=torch.int32, device=self.device)
        st = stream_ptr()
        sw = self.slice_w[s0:s0 + ns]
        _lib.call('mfv_diamond_widths', ddia.ptr, node_dir.data_ptr(), face_dir_rank.data_ptr(), sw.data_ptr(),
                  info.data_ptr(), st)
        ov, wmax = info.tolist()
        if ov or wmax > self.W:
            raise _lib.MfvError("diamond matrix rows need %d slots, the ELL arena has %d" % (wmax, self.W))
        off = s0 * self.W * 64
        _lib.call('mfv_diamond_assemble', ddia.ptr, node_dir.data_ptr(), face_dir_rank.data_ptr(), row0, self.W,
                  sw.data_ptr(), self.cols[off:].data_ptr(), self.vals[off:].data_ptr(), self.vals32[off:].data_ptr(),
                  self.diag[row0:].data_ptr(), st)
        self.narrow_columns(inst)
        return wmax

    def narrow_columns(self, inst, enable=True):
        """16-bit column offsets for the instance's slices whose columns stay within +-32767 of
        their rows (mfv_ell_narrow_columns); enable=False makes every slice read the int32
        columns (same results)."""
        row0, _, ns = self.rows(inst)
        s0 = row0 // 64
        if enable:
            _lib.call('mfv_ell_narrow_columns', self.eptr, s0, ns, self.cols16.data_ptr(), self.slice_narrow.data_ptr(),
                      stream_ptr())
        else:
            self.slice_narrow[s0:s0 + ns].zero_()

    def configure(self, inst, sigma, tau, tol=1e-10, tol_in=1e-4, maxit_in=400, maxit_out=60):
        """Operator coefficients and stopping parameters of one instance (counters zeroed)."""
        import torch
        sc = np.zeros(_lib.KS_NSCAL)
        sc[_lib.KS_SIGMA], sc[_lib.KS_TAU] = sigma, tau
        sc[_lib.KS_TOL_OUT], sc[_lib.KS_TOL_IN] = tol, tol_in
        sc[_lib.KS_MAXIT_IN], sc[_lib.KS_MAXIT_OUT] = maxit_in, maxit_out
        k = inst * _lib.KS_NSCAL
        self.scal[k:k + _lib.KS_NSCAL] = torch.from_numpy(sc).to(self.device)

    def set_system(self, inst, sigma, tau, b, x0=None, tol=1e-10, tol_in=1e-4, maxit_in=400, maxit_out=60):
        """Right-hand side, initial guess and parameters of one instance; status -> START."""
        row0, n, ns = self.rows(inst)
        seg = slice(row0, row0 + ns * 64)
        self.b[seg].zero_()
        self.x[seg].zero_()
        self.b[row0:row0 + n] = b
        if x0 is not None:
            self.x[row0:row0 + n] = x0
        self.configure(inst, sigma, tau, tol, tol_in, maxit_in, maxit_out)
        self.status[inst] = _lib.KS_START

    def solution(self, inst):
        row0, n, _ = self.rows(inst)
        return self.x[row0:row0 + n]

    def scalars(self, inst):
        k = inst * _lib.KS_NSCAL
        return self.scal[k:k + _lib.KS_NSCAL].cpu().numpy()

    # -- execution ----------------------------------------------------------------------------
    def grids(self):
        nsl, ni = getattr(self, '_n_live', (0, 0))
        return max(1, min(4096, (nsl + 3) // 4)), max(1, ni)

    def tick(self, n_iter=8, grid_slices=None, grid_inst=None, stream=None, graph=False):
        """One tick of every live instance. With ``graph=True`` the tick's launches are
        replayed from a HIP graph captured once per (n_iter, grid sizes rounded up to powers of
        two): the kernels read the live lists from device memory, so a graph stays valid as
        instances come and go, and the result is the same bits as the eager tick."""
        gs, gi = self.grids()
        gs, gi = int(grid_slices or gs), int(grid_inst or gi)
        if not graph:
            _lib.call('mfv_krylov_tick', self.eptr, self.wptr, int(n_iter), gs, gi, stream_ptr(stream))
            return
        key = (int(n_iter), min(4096, _pow2(gs)), min(65536, _pow2(gi)))
        graphs = self.__dict__.setdefault('_graphs', {})
        g = graphs.get(key)
        if g is None:
            out = ctypes.c_void_p()
            _lib.call('mfv_krylov_graph', self.eptr, self.wptr, key[0], key[1], key[2], ctypes.byref(out))
            g = graphs[key] = out.value
        _lib.call('mfv_graph_launch', g, stream_ptr(stream))

    def close_graphs(self):
        """Release the captured ticks (after the device has finished any replay of them)."""
        graphs = self.__dict__.pop('_graphs', {})
        if graphs:
            import torch
            torch.cuda.synchronize(self.device)
        for g in graphs.values():
            _lib.call('mfv_graph_destroy', g)

    def __del__(self):
        try:
            self.close_graphs()
        except Exception:
            pass

    def spmv(self, x, y):
        gs, _ = self.grids()
        _lib.call('mfv_ell_spmv', self.eptr, x.data_ptr(), y.data_ptr(), gs, stream_ptr())

    def run(self, insts, n_iter=8, max_ticks=20000, graph=False):
        """Tick until every listed instance is DONE or FAILED; returns their statuses."""
        idx = np.asarray(list(insts), dtype=np.int64)
        for _ in range(max_ticks):
            self.tick(n_iter, graph=graph)
            st = self.status.cpu().numpy()[idx]
            if np.all((st == _lib.KS_DONE) | (st == _lib.KS_FAILED)):
                return st
        raise _lib.MfvError("Krylov solve did not finish in %d ticks" % max_ticks)
