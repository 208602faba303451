"""Diamond-scheme linear systems (``solvers/ls/system.py``) solved on the GPU.

``LinearSolver`` mirrors the reference base class (``:36-234``): ``update_ghost_values``
(Dirichlet data -> ``Pbordnode`` / ``Pbordface``), ``assembly`` (``get_triplet_{2,3}d`` +
``get_rhs_glob_{2,3}d``) and ``compute_Sol_gradient`` (``compute_P_gradient_{2,3}d_diamond``),
all as libmfv kernels, on 2D and 3D meshes. ``HIPKrylovSolver.__call__(rhs=None)`` solves
``A x = rhs0 + rhs`` like ``ScipySolver.__call__`` (``:699-718``) and stores ``x`` in
``var.cell``; the factorisation is replaced by FP32 BiCGStab inside FP64 iterative
refinement down to ``||b - A x|| <= tol ||b||`` (default 1e-10).

``ScipySolver``, ``MUMPSSolver`` and ``PETScKrylovSolver`` accept the reference
constructors' arguments and solve the same system with the same GPU solver.
"""
import numpy as np

from .. import _lib
from ..base import Struct, make_get_conf, process_conf
from ..device import stream_ptr
from ..krylov import EllBatch
from ..variable import Variable


def compute_2dmatrix_size(domain, BCdirichlet, matrixinnerfaces, dirichletfaces):
    """Number of COO triplets of get_triplet_2d (``ast/functions2d.py:582-664``)."""
    fn = domain.faces.nodeid[:, :2]
    on = domain.nodes.oldname
    ncell = domain.nodes.cellid[:, -1].astype(np.int64)
    nghost = (domain.nodes.ghostcenter[:, :, -1] != -1).sum(axis=1).astype(np.int64)
    free = ~np.isin(on, np.asarray(BCdirichlet, dtype=np.int64))
    per_node = np.where(free, 2 * ncell + 2 * nghost, 0)
    mif = np.asarray(matrixinnerfaces, dtype=np.int64)
    return int(4 * len(mif) + per_node[fn[mif, 0]].sum() + per_node[fn[mif, 1]].sum() + 2 * len(dirichletfaces))


def compute_3dmatrix_size(domain, BCdirichlet, matrixinnerfaces, dirichletfaces):
    """Number of COO triplets of get_triplet_3d (``ast/functions3d.py:410-495``): the four node
    slots of a face (a triangle's third node twice) each count their cells and ghost centres
    twice unless the node is a Dirichlet node."""
    from ..diamond import face_nodes4
    fn = face_nodes4(domain)
    on = domain.nodes.oldname
    ncell = domain.nodes.cellid[:, -1].astype(np.int64)
    nghost = (domain.nodes.ghostcenter[:, :, -1] != -1).sum(axis=1).astype(np.int64)
    free = ~np.isin(on, np.asarray(BCdirichlet, dtype=np.int64))
    per_node = np.where(free, 2 * ncell + 2 * nghost, 0)
    mif = np.asarray(matrixinnerfaces, dtype=np.int64)
    return int(4 * len(mif) + sum(per_node[fn[mif, k]].sum() for k in range(4)) + 2 * len(dirichletfaces))


class LinearSolver(object):
    _parameters = [('scheme', 'str', 'diamond', 'fv4', 'scheme diamond or fv4.'),
                   ('reordering', 'bool', False, False, 'reordering the matrix only in serial case.')]

    def __init__(self, domain, var, conf, comm=None, **kwargs):
        if conf is None:
            conf = Struct()
        params = LinearSolver._parameters + [p for p in type(self)._parameters if p not in LinearSolver._parameters]
        self.conf = process_conf(params, conf, kwargs)
        get = make_get_conf(self.conf, kwargs)
        self.scheme = get("scheme")
        self.verbose = get("verbose")
        if self.scheme != "diamond":
            raise NotImplementedError("only the diamond scheme is implemented (the reference's fv4 gradient "
                                      "is an empty stub, ast/functions2d.py:939-940)")
        if not isinstance(var, Variable):
            raise ValueError("Variable must be a Variable type")
        self.var = var
        self.domain = domain
        self.dim = int(domain.dim)
        if self.dim not in (2, 3):
            raise ValueError("dim must be 2 or 3")
        self.comm = comm
        self.localsize = int(domain.nbcells)
        self.globalsize = self.localsize
        self.backend = domain.backend
        self.matrixinnerfaces = np.sort(np.asarray(domain.innerfaces, dtype=np.int32))
        size = compute_2dmatrix_size if self.dim == 2 else compute_3dmatrix_size
        self.dataSize = size(domain, var.BCdirichlet, self.matrixinnerfaces, var.dirichletfaces)
        import torch
        dm = domain.device_mesh
        self._dev = dm.device
        z = lambda n: torch.zeros(n, dtype=torch.float64, device=dm.device)
        self.domain.Pbordnode = z(int(domain.nbnodes))
        self.domain.Pbordface = z(int(domain.nbfaces))
        self.rhs0 = z(self.localsize)
        self._batch = None
        self._inst = None

    def update_ghost_values(self):
        """``solvers/ls/system.py:179-189``: Dirichlet face/node values into Pbordface /
        Pbordnode; Neumann nodes get 1.0."""
        for BC in self.var.BCs.values():
            if BC.BCtype == "dirichlet":
                f = torch_long(BC.BCfaces, self._dev)
                self.domain.Pbordface[f] = BC.BCvalueface[f]
                n = torch_long(np.nonzero(self.domain.nodes.oldname == BC.BCtypeindex)[0], self._dev)
                self.domain.Pbordnode[n] = BC.BCvaluenode[n]
            elif BC.BCtype == "neumann":
                n = torch_long(np.nonzero(self.domain.nodes.oldname == BC.BCtypeindex)[0], self._dev)
                self.domain.Pbordnode[n] = 1.

    def assembly(self):
        """get_triplet_{2,3}d (into the ELL arena) + get_rhs_glob_{2,3}d (``:154-171``)."""
        dd = self.domain.device_diamond
        node_dir, face_dir_rank, _ = self.var.diamond_bc()
        if self._batch is None:
            # the arena is as wide as the widest row (a triangle's stencil fits 32 columns, a
            # tetrahedron's takes up to about 110)
            W = 32 if self.dim == 2 else dd.width(node_dir, face_dir_rank)
            self._batch = EllBatch((self.localsize + 63) // 64, max_inst=1, width=W, device=self._dev)
            self._inst = self._batch.alloc(self.localsize)
            self._batch.publish()
        self.width = self._batch.assemble_diamond(self._inst, dd, node_dir, face_dir_rank)
        _lib.call('mfv_diamond_rhs', dd.ptr, node_dir.data_ptr(), face_dir_rank.data_ptr(),
                  self.domain.Pbordnode.data_ptr(), self.domain.Pbordface.data_ptr(), self.rhs0.data_ptr(),
                  stream_ptr())

    def compute_Sol_gradient(self):
        """``compute_P_gradient_{2,3}d_diamond`` into var.gradfacex / gradfacey (/ gradfacez)
        (``:191-201``); var.node must hold the node interpolation of var.cell."""
        dd = self.domain.device_diamond
        node_dir, _, face_bc = self.var.diamond_bc()
        v = self.var
        args = [dd.ptr, 1, v.cell.data_ptr(), v.ghost.data_ptr(), v.node.data_ptr(), face_bc.data_ptr(),
                node_dir.data_ptr(), self.domain.Pbordnode.data_ptr(), self.domain.Pbordface.data_ptr(),
                v.gradfacex.data_ptr(), v.gradfacey.data_ptr()]
        if self.dim == 2:
            _lib.call('mfv_face_gradient', *args, stream_ptr())
        else:
            _lib.call('mfv_face_gradient_3d', *args, v.gradfacez.data_ptr(), stream_ptr())

    def matrix_coo(self):
        """The assembled matrix as host (data, row, col), duplicates summed (for checks)."""
        b = self._batch
        row0, n, ns = b.rows(self._inst)
        s0 = row0 // 64
        W = b.W
        sw = b.slice_w[s0:s0 + ns].cpu().numpy()
        cols = b.cols[s0 * W * 64:(s0 + ns) * W * 64].cpu().numpy().reshape(ns, W, 64)
        vals = b.vals[s0 * W * 64:(s0 + ns) * W * 64].cpu().numpy().reshape(ns, W, 64)
        used = np.arange(W)[None, :, None] < sw[:, None, None]
        rows = np.broadcast_to(np.arange(ns)[:, None, None] * 64 + np.arange(64)[None, None, :], cols.shape)
        cr = cols - row0
        # unused slots repeat (row, row, 0) after the row's entries; every row has a diagonal entry
        is_diag = cr == rows
        keep = used & (rows < n) & (~is_diag | (np.cumsum(is_diag, axis=1) == 1))
        return vals[keep], rows[keep], cr[keep]


def torch_long(a, dev):
    import torch
    return torch.as_tensor(np.asarray(a, dtype=np.int64), device=dev)


class HIPKrylovSolver(LinearSolver):
    """Solve ``A x = rhs0 + rhs`` on the GPU (FP32 BiCGStab + Jacobi inside FP64 refinement)."""

    _parameters = [('reuse_mtx', 'bool', True, False, 'If True, assemble the matrix once.'),
                   ('tol', 'float', 1e-10, False, 'relative FP64 residual of the refinement loop'),
                   ('tol_inner', 'float', 1e-4, False, 'relative residual reduction of each FP32 solve'),
                   ('maxiter_inner', 'int', 400, False, 'BiCGStab iterations per refinement step'),
                   ('maxiter_outer', 'int', 60, False, 'refinement steps'),
                   ('verbose', 'bool', False, False, 'print solver statistics')]

    def __init__(self, domain=None, var=None, comm=None, conf=None, **kwargs):
        if domain is None:
            raise ValueError("domain must be given")
        if var is None:
            raise ValueError("variable must be given")
        LinearSolver.__init__(self, domain=domain, var=var, conf=conf, comm=comm, **kwargs)
        self._assembled = False
        self.stats = None

    def presolve(self, reuse_mtx=False):
        if not reuse_mtx or not self._assembled:
            self.update_ghost_values()
            self.assembly()
            self._assembled = True

    def __call__(self, rhs=None, conf=None, **kwargs):
        c = self.conf if conf is None else conf
        if not c.reuse_mtx:
            self._assembled = False
        self.presolve(reuse_mtx=c.reuse_mtx)
        import torch
        b = self.rhs0.clone()
        if rhs is not None:
            b += torch.as_tensor(rhs, dtype=torch.float64, device=self._dev)
        x0 = self.var.cell
        self._batch.set_system(self._inst, 0., 1., b, x0, tol=c.tol, tol_in=c.tol_inner, maxit_in=c.maxiter_inner,
                               maxit_out=c.maxiter_outer)
        st = self._batch.run([self._inst])
        sc = self._batch.scalars(self._inst)
        self.stats = {'status': int(st[0]), 'refinements': int(sc[_lib.KS_OUTER]),
                      'iterations': int(sc[_lib.KS_ITTOT]),
                      'relres': float(np.sqrt(sc[_lib.KS_RESN2] / sc[_lib.KS_BN2])) if sc[_lib.KS_BN2] > 0 else 0.}
        if st[0] != _lib.KS_DONE:
            raise _lib.MfvError("linear solve did not converge: %s" % self.stats)
        self.var.cell.copy_(self._batch.solution(self._inst))
        if self.verbose:
            print("HIPKrylovSolver:", self.stats)


class ScipySolver(HIPKrylovSolver):
    """``ScipySolver(domain, var, method, comm, conf)`` (``solvers/ls/system.py:624-744``)."""

    def __init__(self, domain=None, var=None, method=None, comm=None, conf=None, **kwargs):
        if method not in (None, 'superlu', 'auto', 'umfpack'):
            raise ValueError('uknown solution method! (%s)' % method)
        HIPKrylovSolver.__init__(self, domain=domain, var=var, comm=comm, conf=conf, **kwargs)


class MUMPSSolver(HIPKrylovSolver):
    """``MUMPSSolver(domain, var, comm, conf)`` (``solvers/ls/system.py:236-379``)."""


class PETScKrylovSolver(HIPKrylovSolver):
    """``PETScKrylovSolver(domain, var, comm, conf)`` (``solvers/ls/system.py:382-622``); the
    reference's method/preconditioner options are accepted and ignored."""

    _parameters = HIPKrylovSolver._parameters + [('*', None, None, False, 'PETSc options (ignored)')]
