"""Diamond scheme and batched sparse solves on the GPU against the oracle
(oracle/diamond.py): assembly, right-hand side, node interpolation, face gradients,
dissipative residual and time steps bit-exact; the mixed-precision solve to its
stated FP64 residual; batched instances identical alone and in a mix."""
import numpy as np
import pytest

from conftest import gpu_available, make_tri_domain

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]

BCS = [
    {"in": "dirichlet", "out": "dirichlet", "upper": "neumann", "bottom": "neumann"},
    {"in": "dirichlet", "out": "dirichlet", "upper": "dirichlet", "bottom": "dirichlet"},
    {"in": "neumann", "out": "dirichlet", "upper": "nonslip", "bottom": "dirichlet"},
]
FUN = lambda x, y, z: 1.0 + 0.5 * x - 0.25 * y + 0.3 * x * y


def _values(bcs):
    return {k: (FUN if k != "bottom" else 0.75) for k, v in bcs.items() if v == "dirichlet"}


def _oracle_system(d, var, bcs):
    from oracle import diamond as dm
    vals = {k: (lambda x, y, z, f=FUN: f(x, y, z)) if k != "bottom" else 0.75 for k in _values(bcs)}
    Pn, Pf = dm.update_ghost_values(d, {k: ("dirichlet" if v == "dirichlet" else
                                            ("neumann" if v == "neumann" else "other")) for k, v in bcs.items()},
                                    vals)
    mif = dm.matrixinnerfaces(d)
    data, row, col = dm.get_triplet_2d(d, list(var.BCdirichlet), mif, var.dirichletfaces)
    rhs0 = dm.get_rhs_glob_2d(d, Pn, Pf, list(var.BCdirichlet), mif, var.dirichletfaces)
    return data, row, col, Pn, Pf, rhs0


@pytest.mark.parametrize("ib", range(len(BCS)))
def test_assembly_rhs_bitexact(ib):
    import torch
    from oracle import diamond as dm
    from manapy_amd import Variable
    from manapy_amd.solvers.ls import HIPKrylovSolver
    bcs = BCS[ib]
    d = make_tri_domain(23, 17, 1.0, 1.0, 0.12, 40 + ib)
    P = Variable(domain=d, BC=bcs, values=_values(bcs))
    L = HIPKrylovSolver(domain=d, var=P)
    L.update_ghost_values()
    L.assembly()
    torch.cuda.synchronize()
    data, row, col, Pn, Pf, rhs0 = _oracle_system(d, P, bcs)
    assert L.dataSize == len(data)
    assert np.array_equal(d.Pbordnode.cpu().numpy(), Pn)
    assert np.array_equal(d.Pbordface.cpu().numpy(), Pf)
    ref = dm.coo_sum_in_order(data, row, col)
    gv, gr, gc = L.matrix_coo()
    got = {(int(r), int(c)): float(v) for v, r, c in zip(gv, gr, gc)}
    assert got.keys() == ref.keys()
    bad = [k for k in ref if got[k] != ref[k]]
    assert not bad, bad[:5]
    assert np.array_equal(L.rhs0.cpu().numpy(), rhs0)


def test_node_face_gradient_dissipative_bitexact():
    import torch
    from oracle import diamond as dm
    from manapy_amd import Variable, Struct
    from manapy_amd.solvers import DiffusionSolver
    bcs = BCS[2]
    d = make_tri_domain(21, 19, 1.0, 1.0, 0.1, 9)
    C, F = int(d.nbcells), int(d.nbfaces)
    rng = np.random.default_rng(3)
    w0 = rng.standard_normal(C)
    P = Variable(domain=d, BC=bcs, values=_values(bcs))
    u, v = Variable(domain=d), Variable(domain=d)
    P.cell[:] = torch.as_tensor(w0)
    P.update_ghost_value()
    P.interpolate_celltonode()
    P.compute_face_gradient()
    S = DiffusionSolver(P, vel=(u, v), conf=Struct(Dxx=0.7, Dyy=0.3, cfl=0.4))
    S.explicit_dissipative()
    torch.cuda.synchronize()
    g = P.ghost.cpu().numpy()
    wn = dm.centertovertex_2d(d, w0, g)
    assert np.array_equal(P.node.cpu().numpy(), wn)
    lists = np.concatenate([P.dirichletfaces, P.neumannfaces]).astype(np.int64)
    wx, wy = dm.face_gradient_2d(d, w0, g, wn, lists)
    assert np.array_equal(P.gradfacex.cpu().numpy(), wx)
    assert np.array_equal(P.gradfacey.cpu().numpy(), wy)
    dis = dm.explicitscheme_dissipative(d, wx, wy, 0.7, 0.3)
    assert np.array_equal(P.dissipative.cpu().numpy(), dis)
    dt = S.stepper()
    assert dt == dm.time_step_diffusion(d, 0.7, 0.3, 0.4)


def test_p_gradient_bitexact():
    import torch
    from oracle import diamond as dm
    from manapy_amd import Variable
    from manapy_amd.solvers.ls import HIPKrylovSolver
    bcs = BCS[0]
    d = make_tri_domain(19, 15, 1.0, 1.0, 0.1, 12)
    C = int(d.nbcells)
    P = Variable(domain=d, BC=bcs, values=_values(bcs))
    L = HIPKrylovSolver(domain=d, var=P)
    L.update_ghost_values()
    rng = np.random.default_rng(8)
    P.cell[:] = torch.as_tensor(rng.standard_normal(C))
    P.update_ghost_value()
    P.interpolate_celltonode()
    L.compute_Sol_gradient()
    torch.cuda.synchronize()
    Pc, Pg, Pnode = P.cell.cpu().numpy(), P.ghost.cpu().numpy(), P.node.cpu().numpy()
    px, py = dm.p_gradient_2d(d, Pc, Pg, Pnode, d.Pbordnode.cpu().numpy(), d.Pbordface.cpu().numpy(),
                              list(P.BCdirichlet), P.neumannfaces, P.dirichletfaces)
    assert np.array_equal(P.gradfacex.cpu().numpy(), px)
    assert np.array_equal(P.gradfacey.cpu().numpy(), py)


def test_advection_diffusion_steps_bitexact(oracle_lib):
    """AdvectionDiffusionSolver sequence (advecdiff/system.py:149-175) for 5 steps."""
    import torch
    from oracle import diamond as dm
    from manapy_amd import Variable, Struct
    from manapy_amd.solvers import AdvectionDiffusionSolver
    d = make_tri_domain(18, 16, 1.0, 1.0, 0.1, 21)
    C, F = int(d.nbcells), int(d.nbfaces)
    c = d.cells.center
    w0 = np.exp(-((c[:, 0] - .4) ** 2 + (c[:, 1] - .5) ** 2) / .01)
    ne, u, v = Variable(domain=d), Variable(domain=d), Variable(domain=d)
    ne.cell[:] = torch.as_tensor(w0)
    u.face[:] = 0.8
    v.face[:] = -0.3
    u.interpolate_facetocell()
    v.interpolate_facetocell()
    S = AdvectionDiffusionSolver(ne, vel=(u, v), conf=Struct(Dxx=0.01, Dyy=0.02, cfl=0.5, order=1))
    w = w0.copy()
    uf, vf, wf = np.full(F, 0.8), np.full(F, -0.3), np.zeros(F)
    uc, vc, wc = np.zeros(C), np.zeros(C), np.zeros(C)
    oracle_lib.facetocell(d, uf, uc)
    oracle_lib.facetocell(d, vf, vc)
    bf = d.boundaryfaces
    for _ in range(5):
        dt = S.stepper()
        S.compute_fluxes()
        S.compute_new_val()
        dt_ref = dm.time_step_advdiff(d, uc, vc, wc, 0.01, 0.02, 0.5)
        assert dt == dt_ref
        g = np.zeros(F)
        g[bf] = w[d.faces.cellid[bf, 0]]
        rez = np.zeros(C)
        oracle_lib.explicitscheme_convective(d, rez, w, g, uf, vf, wf, np.zeros(C), np.zeros(C), np.zeros(C),
                                             np.zeros(C), 1)
        wn = dm.centertovertex_2d(d, w, g)
        wx, wy = dm.face_gradient_2d(d, w, g, wn, np.sort(bf))
        dis = dm.explicitscheme_dissipative(d, wx, wy, 0.01, 0.02)
        oracle_lib.update_new_value(w, rez, dis, np.zeros(C), dt_ref, d.cells.volume)
        assert np.array_equal(ne.cell.cpu().numpy(), w)


def test_linear_solve_matches_superlu():
    """Poisson problem of examples/2D/laplacien2d.py (Dirichlet data) with a source: the
    GPU solution meets ||b - A x|| <= 1e-10 ||b|| for the oracle's FP64 matrix and agrees with
    SuperLU (the reference ScipySolver's factorisation) to the matrix's conditioning."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    from manapy_amd import Variable
    from manapy_amd.solvers.ls import ScipySolver
    for ib, bcs in enumerate(BCS[:2]):
        d = make_tri_domain(40, 33, 1.0, 1.0, 0.12, 70 + ib)
        C = int(d.nbcells)
        P = Variable(domain=d, BC=bcs, values=_values(bcs))
        L = ScipySolver(domain=d, var=P)
        src = np.full(C, 4.0)
        L(rhs=src)
        x = P.cell.cpu().numpy()
        data, row, col, Pn, Pf, rhs0 = _oracle_system(d, P, bcs)
        A = sp.csr_matrix((data, (row, col)), shape=(C, C))
        b = rhs0 + src
        relres = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
        assert relres <= 1e-10, (relres, L.stats)
        xs = sla.spsolve(A.tocsc(), b)
        assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 5e-7
        assert L.stats['refinements'] >= 2


def test_batched_solves_alone_equal_mix():
    """Instances of different sizes and operators give the same bits alone and batched."""
    import torch
    from manapy_amd import Variable
    from manapy_amd.krylov import EllBatch
    specs = [(14, 11, 1.0, -2e-3, 0), (33, 20, 1.0, -5e-4, 1), (9, 25, 0.0, 1.0, 2), (27, 27, 1.0, -1e-2, 3)]
    doms, bcs_used = [], []
    for nx, ny, sig, tau, seed in specs:
        d = make_tri_domain(nx, ny, 1.0, 1.0, 0.1, 100 + seed)
        bcs = BCS[seed % 2]
        P = Variable(domain=d, BC=bcs, values=_values(bcs))
        doms.append((d, P, sig, tau, seed))
    rng = np.random.default_rng(0)
    rhs = [torch.as_tensor(rng.standard_normal(int(d.nbcells)), device='cuda') for d, *_ in doms]

    def solve(which, batch, graph=False):
        insts = []
        for k in which:
            d, P, sig, tau, _ = doms[k]
            i = batch.alloc(int(d.nbcells))
            insts.append(i)
        batch.publish()
        for k, i in zip(which, insts):
            d, P, sig, tau, _ = doms[k]
            node_dir, rank, _ = P.diamond_bc()
            batch.assemble_diamond(i, d.device_diamond, node_dir, rank)
            batch.set_system(i, sig, tau, rhs[k], None, tol=1e-10)
        st = batch.run(insts, graph=graph)
        assert np.all(st == 4), st
        return [batch.solution(i).cpu().numpy().copy() for i in insts]

    alone = [solve([k], EllBatch(64, max_inst=2))[0] for k in range(len(doms))]
    # a mix with a hole in the arena and instances in another order
    big = EllBatch(200, max_inst=8)
    hole = big.alloc(64 * 7)
    big.publish()
    mixed = solve([2, 0, 3, 1], big)
    for k, x in zip([2, 0, 3, 1], mixed):
        assert np.array_equal(x, alone[k]), k
    # ticks replayed from HIP graphs: same bits as the eager ticks
    g = EllBatch(200, max_inst=8)
    graphed = solve([1, 3, 0, 2], g, graph=True)
    assert len(g._graphs) >= 1
    for k, x in zip([1, 3, 0, 2], graphed):
        assert np.array_equal(x, alone[k]), k
    g.close_graphs()


def _ref(name, module='manapy.ast.functions2d'):
    """A stand-in with the reference kernel's __name__ / __module__ (what Backend.compile sees)."""
    f = lambda *a: None
    f.__name__ = name
    f.__module__ = module
    return f


def test_backend_compile_diamond_kernels_reference_signatures():
    """HIPBackend.compile(<reference kernel>) called with the reference's positional arguments
    (numpy in, outputs written back in place), against the oracle."""
    import scipy.sparse as sp
    from oracle import diamond as dm
    from manapy_amd import Variable
    bcs = BCS[1]
    d = make_tri_domain(17, 13, 1.0, 1.0, 0.1, 31)
    C, F, N = int(d.nbcells), int(d.nbfaces), int(d.nbnodes)
    P = Variable(domain=d, BC=bcs, values=_values(bcs))
    be = d.backend
    data, row, col, Pn, Pf, rhs0 = _oracle_system(d, P, bcs)
    mif = dm.matrixinnerfaces(d)
    z = np.zeros(0, np.int32)
    BCdir = np.asarray(P.BCdirichlet, dtype=np.int32)
    size = be.compile(_ref('compute_2dmatrix_size'))(d.faces.nodeid, d.faces.halofid, d.nodes.cellid, None, None,
                                                     d.nodes.ghostcenter, None, d.nodes.oldname, BCdir, mif, z,
                                                     P.dirichletfaces)
    assert size == len(data)
    a_loc, irn, jcn = np.zeros(size), np.zeros(size, np.int32), np.zeros(size, np.int32)
    nf = d.faces
    nd = d.nodes
    be.compile(_ref('get_triplet_2d'))(nf.cellid, nf.nodeid, nd.vertex, nf.halofid, None, nd.oldname,
                                       d.cells.volume, nd.cellid, d.cells.center, None, None, None, nd.ghostcenter,
                                       None, nf.airDiamond, nd.lambda_x, nd.lambda_y, None, nd.number, nd.R_x,
                                       nd.R_y, None, nf.param1, nf.param2, nf.param3, nf.param4, None, C,
                                       np.arange(C, dtype=np.int32), BCdir, a_loc, irn, jcn, mif, z,
                                       P.dirichletfaces)
    A_gpu = sp.csr_matrix((a_loc, (irn, jcn)), shape=(C, C)).tocoo()
    got = {(int(r), int(c)): float(v) for r, c, v in zip(A_gpu.row, A_gpu.col, A_gpu.data)}
    assert got == dm.coo_sum_in_order(data, row, col)
    rhs = np.zeros(C)
    be.compile(_ref('get_rhs_glob_2d'))(nf.cellid, nf.nodeid, nd.oldname, d.cells.volume, nd.ghostcenter,
                                        np.arange(C, dtype=np.int32), nf.param1, nf.param2, nf.param3, nf.param4,
                                        Pn, Pf, rhs, BCdir, nf.ghostcenter, mif, z, P.dirichletfaces)
    assert np.array_equal(rhs, rhs0)
    rng = np.random.default_rng(2)
    wc, wg = rng.standard_normal(C), rng.standard_normal(F)
    wn = np.zeros(N)
    be.compile(_ref('centertovertex_2d'))(wc, wg, None, None, d.cells.center, None, nd.cellid, nd.ghostid, None,
                                          None, None, nd.vertex, nf.ghostcenter, None, nd.R_x, nd.R_y, None,
                                          nd.lambda_x, nd.lambda_y, None, nd.number, None, wn)
    assert np.array_equal(wn, dm.centertovertex_2d(d, wc, wg))
    wx, wy, wz = np.zeros(F), np.zeros(F), np.zeros(F)
    neu = np.zeros(0, np.int32)
    be.compile(_ref('face_gradient_2d'))(wc, wg, None, wn, nf.cellid, nf.nodeid, nf.ghostcenter, nf.halofid,
                                         d.cells.center, None, nd.vertex, nf.airDiamond, nf.normal, nf.f_1, nf.f_2,
                                         nf.f_3, nf.f_4, None, wx, wy, wz, d.innerfaces, z, P.dirichletfaces, neu, z)
    rx, ry = dm.face_gradient_2d(d, wc, wg, wn, np.concatenate([P.dirichletfaces, neu]))
    assert np.array_equal(wx, rx) and np.array_equal(wy, ry)
    px, py, pz = np.zeros(F), np.zeros(F), np.zeros(F)
    be.compile(_ref('compute_P_gradient_2d_diamond'))(wc, wg, None, wn, nf.cellid, nf.nodeid, nf.ghostcenter,
                                                      nf.halofid, d.cells.center, None, nd.oldname, nf.airDiamond,
                                                      nf.f_1, nf.f_2, nf.f_3, nf.f_4, nf.normal, None, Pn, Pf, px, py,
                                                      pz, BCdir, d.innerfaces, z, neu, P.dirichletfaces, z)
    qx, qy = dm.p_gradient_2d(d, wc, wg, wn, Pn, Pf, list(BCdir), neu, P.dirichletfaces)
    assert np.array_equal(px, qx) and np.array_equal(py, qy)
    dis = np.zeros(C)
    be.compile(_ref('explicitscheme_dissipative', 'manapy.solvers.diffusion.fvm_utils'))(
        rx, ry, np.zeros(F), nf.cellid, nf.normal, nf.name, dis, 0.3, 0.2, 0.)
    assert np.array_equal(dis, dm.explicitscheme_dissipative(d, rx, ry, 0.3, 0.2))
    u, v, w = np.full(C, .6), np.full(C, -.2), np.zeros(C)
    args = (u, v, w, 0.7, nf.normal, nf.mesure, d.cells.volume, d.cells.faceid, 2, 0.3, 0.2, 0.)
    dt_d = be.compile(_ref('time_step', 'manapy.solvers.diffusion.fvm_utils'))(*args)
    dt_a = be.compile(_ref('time_step', 'manapy.solvers.advecdiff.fvm_utils'))(*args)
    assert dt_d == dm.time_step_diffusion(d, 0.3, 0.2, 0.7)
    assert dt_a == dm.time_step_advdiff(d, u, v, w, 0.3, 0.2, 0.7)


# ---------------------------------------------------------------------------------------------
# 3D diamond scheme (ast/functions3d.py) on jittered Kuhn tetrahedra, the reference's cube.msh and
# its mixed tetrahedra / pyramids / hexahedra cube_bis.msh (quadrangular faces: D is a fourth node)

LOCS3 = ("in", "out", "upper", "bottom", "front", "back")
BCS3 = [
    {"in": "dirichlet", "out": "dirichlet", "upper": "neumann", "bottom": "neumann", "front": "neumann",
     "back": "neumann"},
    {k: "dirichlet" for k in LOCS3},
    {"in": "neumann", "out": "dirichlet", "upper": "nonslip", "bottom": "dirichlet", "front": "dirichlet",
     "back": "neumann"},
]
FUN3 = lambda x, y, z: 1.0 + 0.5 * x - 0.25 * y + 0.4 * z + 0.3 * x * y * z


def _values3(bcs):
    return {k: (FUN3 if k != "bottom" else 0.75) for k, v in bcs.items() if v == "dirichlet"}


def _mesh3(kind, seed=0):
    import os
    from conftest import make_tet_domain
    if kind == "kuhn":
        return make_tet_domain(4, 0.08, 60 + seed)
    from manapy_amd.ddm import Domain
    from manapy_amd.gmsh import read_gmsh
    here = os.path.dirname(os.path.abspath(__file__))
    return Domain(3, mesh=read_gmsh(os.path.join(here, 'golden', 'meshes', kind + '.msh.gz'), 3))


def _oracle_system3(d, var, bcs):
    from oracle import diamond as dm
    vals = {k: (lambda x, y, z, f=FUN3: f(x, y, z)) if k != "bottom" else 0.75 for k in _values3(bcs)}
    Pn, Pf = dm.update_ghost_values(d, {k: ("dirichlet" if v == "dirichlet" else
                                            ("neumann" if v == "neumann" else "other")) for k, v in bcs.items()},
                                    vals)
    mif = dm.matrixinnerfaces(d)
    data, row, col = dm.get_triplet_3d(d, list(var.BCdirichlet), mif, var.dirichletfaces)
    rhs0 = dm.get_rhs_glob_3d(d, Pn, Pf, list(var.BCdirichlet), mif, var.dirichletfaces)
    return data, row, col, Pn, Pf, rhs0


@pytest.mark.parametrize("ib,mesh", [(0, "kuhn"), (1, "kuhn"), (2, "kuhn"), (0, "cube"), (0, "cube_bis"),
                                     (2, "cube_bis")])
def test_assembly_rhs_bitexact_3d(ib, mesh):
    """get_triplet_3d (duplicates summed in insertion order) and get_rhs_glob_3d, bit for bit;
    the matrix size count of compute_3dmatrix_size."""
    import torch
    from oracle import diamond as dm
    from manapy_amd import Variable
    from manapy_amd.solvers.ls import HIPKrylovSolver
    bcs = BCS3[ib]
    d = _mesh3(mesh, ib)
    P = Variable(domain=d, BC=bcs, values=_values3(bcs))
    L = HIPKrylovSolver(domain=d, var=P)
    L.update_ghost_values()
    L.assembly()
    torch.cuda.synchronize()
    data, row, col, Pn, Pf, rhs0 = _oracle_system3(d, P, bcs)
    assert L.dataSize == len(data)
    assert np.array_equal(d.Pbordnode.cpu().numpy(), Pn)
    assert np.array_equal(d.Pbordface.cpu().numpy(), Pf)
    ref = dm.coo_sum_in_order(data, row, col)
    gv, gr, gc = L.matrix_coo()
    got = {(int(r), int(c)): float(v) for v, r, c in zip(gv, gr, gc)}
    assert got.keys() == ref.keys()
    bad = [k for k in ref if got[k] != ref[k]]
    assert not bad, bad[:5]
    assert np.array_equal(L.rhs0.cpu().numpy(), rhs0)


@pytest.mark.parametrize("mesh", ["kuhn", "cube", "cube_bis"])
def test_node_face_gradient_dissipative_bitexact_3d(mesh):
    """centertovertex_3d, face_gradient_3d (x, y, z), the dissipative residual with Dzz and the
    diffusion time step in 3D, bit for bit."""
    import torch
    from oracle import diamond as dm
    from manapy_amd import Variable, Struct
    from manapy_amd.solvers import DiffusionSolver
    bcs = BCS3[2]
    d = _mesh3(mesh, 5)
    C = int(d.nbcells)
    w0 = np.random.default_rng(3).standard_normal(C)
    P = Variable(domain=d, BC=bcs, values=_values3(bcs))
    u, v, w = Variable(domain=d), Variable(domain=d), Variable(domain=d)
    P.cell[:] = torch.as_tensor(w0)
    P.update_ghost_value()
    P.interpolate_celltonode()
    P.compute_face_gradient()
    S = DiffusionSolver(P, vel=(u, v, w), conf=Struct(Dxx=0.7, Dyy=0.3, Dzz=0.5, cfl=0.4))
    S.explicit_dissipative()
    torch.cuda.synchronize()
    g = P.ghost.cpu().numpy()
    wn = dm.centertovertex_3d(d, w0, g)
    assert np.array_equal(P.node.cpu().numpy(), wn)
    lists = np.concatenate([P.dirichletfaces, P.neumannfaces]).astype(np.int64)
    wx, wy, wz = dm.face_gradient_3d(d, w0, g, wn, lists)
    assert np.array_equal(P.gradfacex.cpu().numpy(), wx)
    assert np.array_equal(P.gradfacey.cpu().numpy(), wy)
    assert np.array_equal(P.gradfacez.cpu().numpy(), wz)
    dis = dm.explicitscheme_dissipative(d, wx, wy, 0.7, 0.3, wz, 0.5)
    assert np.array_equal(P.dissipative.cpu().numpy(), dis)
    assert S.stepper() == dm.time_step_diffusion(d, 0.7, 0.3, 0.4, 0.5)


def test_p_gradient_bitexact_3d():
    import torch
    from oracle import diamond as dm
    from manapy_amd import Variable
    from manapy_amd.solvers.ls import HIPKrylovSolver
    bcs = BCS3[2]
    d = _mesh3("kuhn", 7)
    C = int(d.nbcells)
    P = Variable(domain=d, BC=bcs, values=_values3(bcs))
    L = HIPKrylovSolver(domain=d, var=P)
    L.update_ghost_values()
    P.cell[:] = torch.as_tensor(np.random.default_rng(8).standard_normal(C))
    P.update_ghost_value()
    P.interpolate_celltonode()
    L.compute_Sol_gradient()
    torch.cuda.synchronize()
    Pc, Pg, Pnode = P.cell.cpu().numpy(), P.ghost.cpu().numpy(), P.node.cpu().numpy()
    px, py, pz = dm.p_gradient_3d(d, Pc, Pg, Pnode, d.Pbordnode.cpu().numpy(), d.Pbordface.cpu().numpy(),
                                  list(P.BCdirichlet), P.neumannfaces, P.dirichletfaces)
    assert np.array_equal(P.gradfacex.cpu().numpy(), px)
    assert np.array_equal(P.gradfacey.cpu().numpy(), py)
    assert np.array_equal(P.gradfacez.cpu().numpy(), pz)


@pytest.mark.parametrize("mesh", ["kuhn", "cube", "cube_bis"])
def test_linear_solve_3d_matches_superlu(mesh):
    """A 3D Poisson problem with Dirichlet data on two sides and a source: the GPU solution meets
    ||b - A x|| <= 1e-10 ||b|| for the oracle's FP64 matrix and agrees with SuperLU."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    from manapy_amd import Variable
    from manapy_amd.solvers.ls import ScipySolver
    bcs = BCS3[0]
    d = _mesh3(mesh, 9)
    C = int(d.nbcells)
    P = Variable(domain=d, BC=bcs, values=_values3(bcs))
    L = ScipySolver(domain=d, var=P)
    src = np.full(C, 4.0)
    L(rhs=src)
    x = P.cell.cpu().numpy()
    data, row, col, Pn, Pf, rhs0 = _oracle_system3(d, P, bcs)
    A = sp.csr_matrix((data, (row, col)), shape=(C, C))
    b = rhs0 + src
    relres = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    assert relres <= 1e-10, (relres, L.stats)
    xs = sla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 5e-7


def test_narrow_columns_give_the_same_products():
    """The 16-bit column offsets (mfv_ell_narrow_columns) are used on the slices of a banded
    numbering and refused where a column lies farther than 32767 rows away; the FP64 product
    and a whole solve are bit-identical with and without them."""
    import torch
    from manapy_amd import _lib
    from manapy_amd.device import stream_ptr
    from manapy_amd.ddm import Domain
    from manapy_amd.krylov import EllBatch
    from manapy_amd.meshgen import jittered_triangles
    el, nd = jittered_triangles(200, 200, 1.0, 1.0, 0.1, 77)     # 80 000 cells, row-major order
    perm = np.random.default_rng(5).permutation(len(el))
    for d, expect_narrow in ((Domain(2, mesh=(el, nd)), True), (Domain(2, mesh=(el[perm], nd)), False)):
        C = int(d.nbcells)
        dd = d.device_diamond
        node_dir = torch.zeros(int(d.nbnodes), dtype=torch.uint8, device='cuda')
        rank = torch.full((int(d.nbfaces),), -1, dtype=torch.int32, device='cuda')
        b = EllBatch((C + 63) // 64, max_inst=1, width=32)
        inst = b.alloc(C)
        b.publish()
        b.assemble_diamond(inst, dd, node_dir, rank)
        ns = (C + 63) // 64
        frac = float(b.slice_narrow[:ns].float().mean())
        assert (frac == 1.0) if expect_narrow else (frac < 0.5), frac
        x = torch.as_tensor(np.random.default_rng(1).standard_normal(ns * 64), device='cuda')
        y1, y2 = torch.zeros_like(x), torch.zeros_like(x)
        _lib.call('mfv_ell_spmv', b.eptr, x.data_ptr(), y1.data_ptr(), 256, stream_ptr())
        b.narrow_columns(inst, enable=False)
        _lib.call('mfv_ell_spmv', b.eptr, x.data_ptr(), y2.data_ptr(), 256, stream_ptr())
        assert torch.equal(y1, y2)
        sols = []
        rhs = torch.as_tensor(np.random.default_rng(2).standard_normal(C), device='cuda')
        for enable in (True, False):
            b.narrow_columns(inst, enable=enable)
            b.set_system(inst, 1.0, -2e-4, rhs)
            for _ in range(200):
                b.tick(8)
                if int(b.status[inst]) in (_lib.KS_DONE, _lib.KS_FAILED):
                    break
            assert int(b.status[inst]) == _lib.KS_DONE
            sols.append(b.x[:C].clone())
        assert torch.equal(sols[0], sols[1])
