"""Device-resident fields with manapy's ``Variable`` / ``Boundary`` interface.

``Variable(domain, terms, comm, name, BC, values)`` keeps manapy's attribute names
(``ast/core.py:24-83``) -- ``cell``, ``node``, ``face``, ``ghost``, ``halo``,
``gradcellx/y/z``, ``gradfacex/y/z``, ``psi`` ... -- but every array is a float64
tensor resident on the GPU, so solver calls never copy state to the host. Assign
with ``var.cell[:] = ...`` (tensor or numpy) and read back with ``var.cell.cpu()``.

Boundary conditions follow ``ast/core.py:105-335`` and ``boundary/bc.py``: a dict
``{"in": "neumann" | "dirichlet" | "nonslip" | "neumannNH" | "slip", ...}`` and, for
Dirichlet / non-homogeneous Neumann, ``values={"in": float | lambda x, y, z: ...}``.
The ghost fill of every BC kind runs in ``mfv_ghost_value`` on the device.
"""
import types

import numpy as np

from . import _lib
from .device import stream_ptr

_KIND = {"neumann": 0, "periodic": 0, "slip": 0, "nonslip": 1, "dirichlet": 2}


class Boundary(object):
    """One boundary location of a Variable (``boundary/bc.py:8-103``)."""

    def __init__(self, BCtype=None, BCloc=None, BCtypeindex=None, domain=None, values=None, var=None):
        if domain is None:
            raise ValueError("domain must be given")
        import torch
        self._BCtype = BCtype
        self._BCloc = BCloc
        self._BCtypeindex = BCtypeindex
        self._domain = domain
        faces = {"in": domain.infaces, "out": domain.outfaces, "bottom": domain.bottomfaces,
                 "upper": domain.upperfaces}
        if int(domain.dim) == 3:
            faces.update({"front": domain.frontfaces, "back": domain.backfaces})
        nodes = {"in": domain.innodes, "out": domain.outnodes, "bottom": domain.bottomnodes,
                 "upper": domain.uppernodes}
        if int(domain.dim) == 3:
            nodes.update({"front": domain.frontnodes, "back": domain.backnodes})
        self._BCfaces = np.asarray(faces[BCloc], dtype=np.int32)
        self.BCvaluenode = None
        self._BCnodes = np.asarray(nodes.get(BCloc, np.zeros(0, np.int32)), dtype=np.int32)
        dm = domain.device_mesh
        self._faces_dev = torch.as_tensor(self._BCfaces, device=dm.device)
        self.BCvalueface = None
        self.constNH = None
        F = int(domain.nbfaces)
        if BCtype in ("dirichlet", "neumannNH"):
            if values is None:
                raise ValueError("Value of dirichlet BC for " + str(BCloc) + " faces must be given")
            valueface = np.zeros(F)
            fc = domain.faces.center
            if isinstance(values, types.LambdaType):
                for i in self._BCfaces:
                    valueface[i] = values(fc[i][0], fc[i][1], fc[i][2])
            else:
                valueface[self._BCfaces] = float(values)
            valuenode = np.zeros(int(domain.nbnodes))
            vx = domain.nodes.vertex
            own = np.nonzero(domain.nodes.oldname == BCtypeindex)[0]
            if isinstance(values, types.LambdaType):
                for i in own:
                    valuenode[i] = values(vx[i][0], vx[i][1], vx[i][2])
            else:
                valuenode[own] = float(values)
            if BCtype == "dirichlet":
                self.BCvalueface = torch.as_tensor(valueface, device=dm.device)
                self.BCvaluenode = torch.as_tensor(valuenode, device=dm.device)
            else:
                self.constNH = torch.as_tensor(valueface, device=dm.device)

    def fill_ghost(self, var):
        """ghost_value_* of ``boundary/bc_utils.py:45-68`` for this location."""
        n = len(self._BCfaces)
        if n == 0:
            return
        dm = self._domain.device_mesh
        if self._BCtype == "neumannNH":
            # w_ghost[f] = w_c[cellid[f][0]] + cst[f] * dist[f]
            _lib.call('mfv_ghost_value', dm.ptr, 0, var.cell.data_ptr(), var.ghost.data_ptr(),
                      self._faces_dev.data_ptr(), n, stream_ptr())
            import torch
            f = self._faces_dev.long()
            var.ghost[f] = var.ghost[f] + self.constNH[f] * dm.dist_ortho[f]
            return
        kind = _KIND[self._BCtype]
        src = self.BCvalueface if kind == 2 else var.cell
        _lib.call('mfv_ghost_value', dm.ptr, kind, src.data_ptr(), var.ghost.data_ptr(), self._faces_dev.data_ptr(),
                  n, stream_ptr())

    domain = property(lambda self: self._domain)
    BCfaces = property(lambda self: self._BCfaces)
    BCnodes = property(lambda self: self._BCnodes)
    BCtypeindex = property(lambda self: self._BCtypeindex)
    BCtype = property(lambda self: self._BCtype)


class Variable(object):
    """Device counterpart of ``manapy.ast.Variable`` (``ast/core.py:24-575``)."""

    def __init__(self, domain=None, terms=None, comm=None, name=None, BC=None, values=None, *args, **kwargs):
        if domain is None:
            raise ValueError("domain must be given")
        import torch
        self._domain = domain
        self._BC = BC
        self._values = values if values is not None else {}
        self._name = name
        self._dim = domain.dim
        self._comm = comm
        dm = domain.device_mesh
        dev = dm.device
        C, N, F = int(domain.nbcells), int(domain.nbnodes), int(domain.nbfaces)
        z = lambda n: torch.zeros(n, dtype=torch.float64, device=dev)
        self.cell = z(C)
        self.node = z(N)
        self.face = z(F)
        self.ghost = z(F)
        self.halo = z(0)
        self.gradcellx, self.gradcelly, self.gradcellz = z(C), z(C), z(C)
        self.gradhalocellx, self.gradhalocelly, self.gradhalocellz = z(0), z(0), z(0)
        self.gradfacex, self.gradfacey, self.gradfacez = z(F), z(F), z(F)
        self.psi = z(C)
        self.psihalo = z(0)
        self.haloghost = z(0)
        if terms is not None:
            for t in terms:
                self.__dict__[t] = z(C)
        self._update_boundaries()

    def _update_boundaries(self):
        """``ast/core.py:105-335``: one Boundary per location, ``BC=None`` = all Neumann."""
        locs = ["in", "out", "bottom", "upper"] + (["front", "back"] if int(self._dim) == 3 else [])
        self._BCs = {}
        self.dirichletfaces, self.neumannfaces, self.neumannNHfaces = [], [], []
        self.BCdirichlet, self.BCneumann, self.BCneumannNH = [], [], []
        for loc in locs:
            bct = "neumann" if self._BC is None else self._BC.get(loc, "neumann")
            if bct == "periodic":
                raise ValueError("BC must be not periodic for " + str(loc))
            if bct not in ("neumann", "dirichlet", "nonslip", "neumannNH", "slip"):
                raise ValueError("unknown boundary condition %r for %s" % (bct, loc))
            if bct == "slip":
                # the reference assigns no ghost function to 'slip' (boundary/bc.py:59-78); the
                # generic Variable treats it as Neumann, like its BC bookkeeping (ast/core.py:283-296)
                pass
            b = Boundary(BCtype=bct, BCloc=loc, BCtypeindex=self._domain._BCs[loc][1], domain=self._domain,
                         values=self._values.get(loc) if bct in ("dirichlet", "neumannNH") else None)
            self._BCs[loc] = b
            if bct == "dirichlet":
                self.dirichletfaces.extend(b.BCfaces)
                self.BCdirichlet.append(b.BCtypeindex)
            elif bct == "neumannNH":
                self.neumannNHfaces.extend(b.BCfaces)
                self.BCneumannNH.append(b.BCtypeindex)
            else:
                self.neumannfaces.extend(b.BCfaces)
                self.BCneumann.append(b.BCtypeindex)
        self.dirichletfaces = np.sort(np.asarray(self.dirichletfaces, dtype=np.int32))
        self.neumannfaces = np.sort(np.asarray(self.neumannfaces, dtype=np.int32))
        self.neumannNHfaces = np.asarray(self.neumannNHfaces, dtype=np.int32)
        self.BCdirichlet = np.asarray(self.BCdirichlet, dtype=np.int32)
        self.BCneumann = np.asarray(self.BCneumann, dtype=np.int32)
        self.BCneumannNH = np.asarray(self.BCneumannNH, dtype=np.int32)
        self._diamond_bc = None

    def diamond_bc(self):
        """Device tables (node_dir, face_dir_rank, face_bc) of the diamond kernels."""
        if self._diamond_bc is None:
            import torch
            from .diamond import bc_tables
            dev = self._domain.device_mesh.device
            self._diamond_bc = tuple(torch.from_numpy(a).to(dev) for a in bc_tables(
                self._domain, self.BCdirichlet, self.dirichletfaces, self.neumannfaces))
        return self._diamond_bc

    # -- operators ---------------------------------------------------------------------------
    def update_halo_value(self):
        """One subdomain: no halo cells, nothing to exchange (``ast/core.py:415-420``)."""

    def update_ghost_value(self):
        """``ast/core.py:492-497``."""
        for bc in self._BCs.values():
            bc.fill_ghost(self)

    def interpolate_facetocell(self):
        """``ast/core.py:422-423`` -> ``ast_utils.facetocell``."""
        dm = self._domain.device_mesh
        _lib.call('mfv_facetocell', dm.ptr, self.face.data_ptr(), self.cell.data_ptr(), stream_ptr())

    def interpolate_celltonode(self):
        """``ast/core.py:429-440`` -> ``centertovertex_2d`` / ``centertovertex_3d``
        (pseudo-Laplacian weights)."""
        dd = self._domain.device_diamond
        _lib.call('mfv_centertovertex', dd.ptr, self.cell.data_ptr(), self.ghost.data_ptr(), self.node.data_ptr(),
                  stream_ptr())

    def compute_face_gradient(self):
        """``ast/core.py:480-490`` -> ``face_gradient_2d`` / ``face_gradient_3d`` on the inner,
        Dirichlet and Neumann faces (uses ``node``: call interpolate_celltonode first, as the
        reference does)."""
        dd = self._domain.device_diamond
        node_dir, _, face_bc = self.diamond_bc()
        args = [dd.ptr, 0, self.cell.data_ptr(), self.ghost.data_ptr(), self.node.data_ptr(), face_bc.data_ptr(),
                None, None, None, self.gradfacex.data_ptr(), self.gradfacey.data_ptr()]
        if int(self._dim) == 2:
            _lib.call('mfv_face_gradient', *args, stream_ptr())
        else:
            _lib.call('mfv_face_gradient_3d', *args, self.gradfacez.data_ptr(), stream_ptr())

    def compute_cell_gradient(self):
        """``ast/core.py:442-478``: least-squares gradient then Barth-Jespersen limiter."""
        dm = self._domain.device_mesh
        dm.ensure_lsq()
        s = stream_ptr()
        _lib.call('mfv_cell_gradient', dm.ptr, self.cell.data_ptr(), self.ghost.data_ptr(), self.gradcellx.data_ptr(),
                  self.gradcelly.data_ptr(), self.gradcellz.data_ptr(), s)
        _lib.call('mfv_barthlimiter', dm.ptr, self.cell.data_ptr(), self.ghost.data_ptr(), self.gradcellx.data_ptr(),
                  self.gradcelly.data_ptr(), self.gradcellz.data_ptr(), self.psi.data_ptr(), s)

    def norml2(self, exact, order=None):
        """``ast/core.py:500-515``: ||(w - exact) vol||_p / ||exact vol||_p (on the device)."""
        import torch
        if order is None:
            order = 1
        ex = torch.as_tensor(np.asarray(exact) if not torch.is_tensor(exact) else exact, dtype=torch.float64,
                             device=self.cell.device)
        assert self.nbcells == len(ex), 'exact solution must have length of cells'
        vol = self._domain.device_mesh.t['cell_volume']
        err = torch.linalg.vector_norm((self.cell - ex).abs() * vol, ord=order)
        ref = torch.linalg.vector_norm(ex.abs() * vol, ord=order)
        return float(err / ref)

    # -- properties (ast/core.py:517-575) -------------------------------------------------------
    domain = property(lambda self: self._domain)
    dim = property(lambda self: self._dim)
    comm = property(lambda self: self._comm)
    nbfaces = property(lambda self: int(self._domain.nbfaces))
    nbcells = property(lambda self: int(self._domain.nbcells))
    nbnodes = property(lambda self: int(self._domain.nbnodes))
    nbhalos = property(lambda self: 0)
    name = property(lambda self: self._name)
    BCs = property(lambda self: self._BCs)
    BCin = property(lambda self: self._BCs["in"])
    BCout = property(lambda self: self._BCs["out"])
    BCupper = property(lambda self: self._BCs["upper"])
    BCbottom = property(lambda self: self._BCs["bottom"])
    BCfront = property(lambda self: self._BCs.get("front", 0))
    BCback = property(lambda self: self._BCs.get("back", 0))
