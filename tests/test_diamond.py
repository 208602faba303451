"""Pins for the diamond-scheme oracle (oracle/diamond.py): the properties that define the
scheme, checked independently of its formulas."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from conftest import make_tri_domain
from oracle import diamond as dm

DIR_ALL = {"in": "dirichlet", "out": "dirichlet", "upper": "dirichlet", "bottom": "dirichlet"}
NEU_ALL = {"in": "neumann", "out": "neumann", "upper": "neumann", "bottom": "neumann"}


@pytest.fixture(scope="module")
def dd():
    return make_tri_domain(18, 14, 1.0, 1.0, 0.1, 3)


def _dirichlet_lists(d, bcs):
    idx = {"in": 1, "out": 2, "upper": 3, "bottom": 4}
    faces = {"in": d.infaces, "out": d.outfaces, "upper": d.upperfaces, "bottom": d.bottomfaces}
    BCdir = [idx[k] for k, v in bcs.items() if v == "dirichlet"]
    dirfaces = np.sort(np.concatenate([faces[k] for k, v in bcs.items() if v == "dirichlet"] or
                                      [np.zeros(0, np.int32)])).astype(np.int32)
    return BCdir, dirfaces


def _matrix(d, bcs):
    BCdir, dirfaces = _dirichlet_lists(d, bcs)
    mif = dm.matrixinnerfaces(d)
    data, row, col = dm.get_triplet_2d(d, BCdir, mif, dirfaces)
    assert len(data) == dm.compute_2dmatrix_size(d, BCdir, mif, dirfaces)
    C = int(d.nbcells)
    return sp.csc_matrix((data, (row, col)), shape=(C, C)), BCdir, dirfaces, mif


def test_diamond_affine_is_exact(dd):
    """The diamond scheme reproduces affine solutions of the Laplace problem whenever its
    boundary treatment is exact: u = a + b x with Dirichlet data on the x = const sides (the
    reference's mirror ghost uses the face-centre value, functions2d.py:736-737, exact when u
    has no tangential gradient there) and Neumann on the y = const sides (ghost = owner value,
    exact for u independent of y). Every row of A u - rhs0 vanishes and SuperLU (the
    reference ScipySolver) returns u."""
    d = dd
    bcs = {"in": "dirichlet", "out": "dirichlet", "upper": "neumann", "bottom": "neumann"}
    f = lambda x, y, z: 0.3 + 1.7 * x
    A, BCdir, dirfaces, mif = _matrix(d, bcs)
    Pn, Pf = dm.update_ghost_values(d, bcs, {"in": f, "out": f})
    rhs0 = dm.get_rhs_glob_2d(d, Pn, Pf, BCdir, mif, dirfaces)
    u = f(d.cells.center[:, 0], d.cells.center[:, 1], 0.)
    assert np.abs(A @ u - rhs0).max() < 1e-9 * np.abs(rhs0).max()
    x = sla.spsolve(A, rhs0)
    assert np.abs(x - u).max() < 1e-11


def test_diamond_dirichlet_tangential_gradient_is_not_exact(dd):
    """...and with a tangential gradient on Dirichlet faces the boundary rows are not exact
    (interior rows still are): the reference's property, kept."""
    d = dd
    f = lambda x, y, z: 0.3 + 1.7 * x - 0.9 * y
    A, BCdir, dirfaces, mif = _matrix(d, DIR_ALL)
    Pn, Pf = dm.update_ghost_values(d, DIR_ALL, {k: f for k in DIR_ALL})
    rhs0 = dm.get_rhs_glob_2d(d, Pn, Pf, BCdir, mif, dirfaces)
    u = f(d.cells.center[:, 0], d.cells.center[:, 1], 0.)
    r = np.abs(A @ u - rhs0)
    bcells = np.unique(d.faces.cellid[d.boundaryfaces, 0])
    far = np.setdiff1d(np.arange(int(d.nbcells)), bcells)
    assert r[bcells].max() > 1.0
    near_nodes = np.unique(d.cells.nodeid[bcells, :3])
    touch = np.isin(d.cells.nodeid[:, :3], near_nodes).any(axis=1)
    interior = np.nonzero(~touch)[0]
    assert len(interior) > 100 and r[interior].max() < 1e-9 * np.abs(rhs0).max()


def test_diamond_neumann_rows_sum_to_zero(dd):
    A, _, _, _ = _matrix(dd, NEU_ALL)
    rs = np.asarray(A.sum(axis=1)).ravel()
    assert np.abs(rs).max() < 1e-9 * abs(A).max()
    # and the operator is not trivially zero, with negative diagonal (a discrete Laplacian)
    assert (A.diagonal() < 0).all()


def test_diamond_mixed_bc_converges_to_quadratic(dd):
    """Poisson with Dirichlet data of u = x^2 + y^2 and source 4 converges (first order in the
    max norm: the boundary treatment above is first order)."""
    errs = []
    for n in (10, 20):
        d = make_tri_domain(n, n, 1.0, 1.0, 0.1, 1)
        f = lambda x, y, z: x * x + y * y
        A, BCdir, dirfaces, mif = _matrix(d, DIR_ALL)
        Pn, Pf = dm.update_ghost_values(d, DIR_ALL, {k: f for k in DIR_ALL})
        rhs0 = dm.get_rhs_glob_2d(d, Pn, Pf, BCdir, mif, dirfaces)
        x = sla.spsolve(A, rhs0 + 4.0)
        u = f(d.cells.center[:, 0], d.cells.center[:, 1], 0.)
        errs.append(np.abs(x - u).max())
    assert errs[1] < errs[0] / 1.6


def test_node_interpolation_and_face_gradient_affine(dd):
    d = dd
    g = lambda x, y: 0.5 - 2.0 * x + 3.0 * y
    wc = g(d.cells.center[:, 0], d.cells.center[:, 1])
    wg = np.zeros(int(d.nbfaces))
    bf = d.boundaryfaces
    wg[bf] = g(d.faces.ghostcenter[bf, 0], d.faces.ghostcenter[bf, 1])
    wn = dm.centertovertex_2d(d, wc, wg)
    assert np.abs(wn - g(d.nodes.vertex[:, 0], d.nodes.vertex[:, 1])).max() < 1e-12
    wx, wy = dm.face_gradient_2d(d, wc, wg, wn, bf)
    assert np.abs(wx + 2.0).max() < 1e-10
    assert np.abs(wy - 3.0).max() < 1e-10
    dis = dm.explicitscheme_dissipative(d, wx, wy, 0.7, 0.7)
    # affine field: the discrete divergence of a constant flux vanishes cell by cell
    # (boundary faces carry the same constant gradient)
    assert np.abs(dis).max() < 1e-11


def test_diffusion_time_step_closed_form(dd):
    d = dd
    dt = dm.time_step_diffusion(d, 0.1, 0.1, 0.8)
    lam = np.zeros(int(d.nbcells))
    for j in range(3):
        m = d.faces.mesure[d.cells.faceid[:, j]]
        lam += 0.2 * m * m / d.cells.volume
    assert np.isclose(dt, (0.8 * d.cells.volume / lam).min(), rtol=1e-13)


def test_p_gradient_affine_with_dirichlet_mirror(dd):
    """compute_P_gradient_2d_diamond returns minus the gradient (Px = -dP/dx, Py = -dP/dy);
    exact for P = a + b x with Dirichlet data on x = const faces (mirror value 2 P_f - P_c is
    exact without a tangential gradient) and Neumann ghosts on y = const faces."""
    d = dd
    bcs = {"in": "dirichlet", "out": "dirichlet", "upper": "neumann", "bottom": "neumann"}
    f = lambda x, y, z: 0.3 + 1.7 * x
    Pn, Pf = dm.update_ghost_values(d, bcs, {"in": f, "out": f})
    Pc = f(d.cells.center[:, 0], d.cells.center[:, 1], 0.)
    F = int(d.nbfaces)
    Pg = np.zeros(F)
    bf = d.boundaryfaces
    Pg[bf] = Pc[d.faces.cellid[bf, 0]]
    # exact node values (the pseudo-Laplacian interpolation is exact for affine data)
    Pnode = dm.centertovertex_2d(d, Pc, np.where(np.isin(np.arange(F), bf),
                                                 f(d.faces.ghostcenter[:, 0], d.faces.ghostcenter[:, 1], 0.), 0.))
    BCdir, dirfaces = _dirichlet_lists(d, bcs)
    neu = np.sort(np.concatenate([d.upperfaces, d.bottomfaces])).astype(np.int32)
    px, py = dm.p_gradient_2d(d, Pc, Pg, Pnode, Pn, Pf, BCdir, neu, dirfaces)
    faces = np.concatenate([d.innerfaces, neu, dirfaces]).astype(np.int64)
    assert np.abs(px[faces] + 1.7).max() < 1e-10
    assert np.abs(py[faces]).max() < 1e-10


def test_advdiff_time_step_reduces_to_pinned_steps(dd, oracle_lib):
    d = dd
    C = int(d.nbcells)
    z = np.zeros(C)
    assert dm.time_step_advdiff(d, z, z, z, 0.1, 0.2, 0.7) == dm.time_step_diffusion(d, 0.1, 0.2, 0.7)
    u, v = np.full(C, 0.9), np.full(C, -0.4)
    assert dm.time_step_advdiff(d, u, v, z, 0., 0., 0.7) == oracle_lib.time_step(d, u, v, z, 0.7)


def test_coo_insertion_sum_matches_scipy(dd):
    A, BCdir, dirfaces, mif = _matrix(dd, DIR_ALL)
    data, row, col = dm.get_triplet_2d(dd, BCdir, mif, dirfaces)
    ref = dm.coo_sum_in_order(data, row, col)
    Acoo = A.tocoo()
    got = {(int(r), int(c)): float(v) for r, c, v in zip(Acoo.row, Acoo.col, Acoo.data)}
    assert got.keys() == ref.keys()
    err = max(abs(got[k] - ref[k]) for k in ref)
    assert err <= 1e-12 * max(abs(v) for v in ref.values())
    # integer-valued triplets sum exactly in any order
    ints = dm.coo_sum_in_order(np.array([1., 2., 3., 4.]), np.array([0, 1, 0, 0]), np.array([0, 1, 0, 1]))
    assert ints == {(0, 0): 4., (1, 1): 2., (0, 1): 4.}


def test_host_diamond_layout_matches_reference_expressions(dd):
    """The per-incidence weights of the GPU layout equal the reference's inline expression
    (evaluated by the oracle) bit for bit; the triplet count equals compute_2dmatrix_size."""
    from manapy_amd.diamond import diamond_layout
    from manapy_amd.solvers.ls import compute_2dmatrix_size
    d = dd
    L = diamond_layout(d)
    ptr, cells, alpha = L['node_cell_ptr'], L['node_cell'], L['node_cell_alpha']
    for nod in (0, 5, 77, int(d.nbnodes) - 1):
        for q in range(ptr[nod], ptr[nod + 1]):
            c = cells[q]
            assert alpha[q] == dm._alpha(d, nod, d.cells.center[c][0], d.cells.center[c][1])
    gptr, gface, gal, gown = L['node_ghost_ptr'], L['node_ghost_face'], L['node_ghost_alpha'], L['node_ghost_owner']
    for nod in np.nonzero(np.diff(gptr))[0][:20]:
        for q in range(gptr[nod], gptr[nod + 1]):
            f = gface[q]
            gc = d.faces.ghostcenter[f]
            assert gal[q] == dm._alpha(d, nod, gc[0], gc[1])
            assert gown[q] == d.faces.cellid[f, 0]
    for bcs in (DIR_ALL, NEU_ALL, {"in": "dirichlet", "out": "neumann", "upper": "neumann", "bottom": "dirichlet"}):
        BCdir, dirfaces = _dirichlet_lists(d, bcs)
        mif = dm.matrixinnerfaces(d)
        assert compute_2dmatrix_size(d, BCdir, mif, dirfaces) == dm.compute_2dmatrix_size(d, BCdir, mif, dirfaces)


# ---------------------------------------------------------------------------------------------
# 3D (functions3d.py): the same defining properties on jittered Kuhn tetrahedra

IDX3 = {"in": 1, "out": 2, "upper": 3, "bottom": 4, "front": 5, "back": 6}


@pytest.fixture(scope="module")
def dd3():
    from conftest import make_tet_domain
    return make_tet_domain(5, 0.08, 3)


def _lists3(d, bcs):
    faces = {"in": d.infaces, "out": d.outfaces, "upper": d.upperfaces, "bottom": d.bottomfaces,
             "front": d.frontfaces, "back": d.backfaces}
    BCdir = [IDX3[k] for k, v in bcs.items() if v == "dirichlet"]
    dirfaces = np.sort(np.concatenate([faces[k] for k, v in bcs.items() if v == "dirichlet"] or
                                      [np.zeros(0, np.int32)])).astype(np.int32)
    neufaces = np.sort(np.concatenate([faces[k] for k, v in bcs.items() if v == "neumann"] or
                                      [np.zeros(0, np.int32)])).astype(np.int32)
    return BCdir, dirfaces, neufaces


def _matrix3(d, bcs):
    BCdir, dirfaces, _ = _lists3(d, bcs)
    mif = dm.matrixinnerfaces(d)
    data, row, col = dm.get_triplet_3d(d, BCdir, mif, dirfaces)
    assert len(data) == dm.compute_3dmatrix_size(d, BCdir, mif, dirfaces)
    C = int(d.nbcells)
    return sp.csc_matrix((data, (row, col)), shape=(C, C)), BCdir, dirfaces, mif


X_DIR = {"in": "dirichlet", "out": "dirichlet", "upper": "neumann", "bottom": "neumann", "front": "neumann",
         "back": "neumann"}


def test_diamond_3d_affine_is_exact(dd3):
    """get_triplet_3d + get_rhs_glob_3d reproduce u = a + b x with Dirichlet data on the x
    sides (the mirror 2 V_K - V_L is exact without a tangential gradient) and Neumann on the
    others (zero flux, ghost = owner value, exact for u independent of y and z): every row of
    A u - rhs0 vanishes and SuperLU returns u. A single Dirichlet diagonal term (the reference
    adds -param3/|K| twice), a parameter in the wrong node slot or a dropped z term of the node
    weights breaks it."""
    d = dd3
    f = lambda x, y, z: 0.3 + 1.7 * x
    A, BCdir, dirfaces, mif = _matrix3(d, X_DIR)
    Pn, Pf = dm.update_ghost_values(d, X_DIR, {"in": f, "out": f})
    rhs0 = dm.get_rhs_glob_3d(d, Pn, Pf, BCdir, mif, dirfaces)
    c = d.cells.center
    u = f(c[:, 0], c[:, 1], c[:, 2])
    assert np.abs(A @ u - rhs0).max() < 1e-12 * np.abs(rhs0).max()
    x = sla.spsolve(A, rhs0)
    assert np.abs(x - u).max() < 1e-11


def test_diamond_3d_neumann_rows_sum_to_zero(dd3):
    """Pure Neumann: constants are in the kernel of the operator and the lifting is zero."""
    d = dd3
    C = int(d.nbcells)
    A, BCdir, dirfaces, mif = _matrix3(d, {k: "neumann" for k in IDX3})
    assert np.abs(A @ np.ones(C)).max() < 1e-12 * np.abs(A.diagonal()).max()
    Pn, Pf = dm.update_ghost_values(d, {k: "neumann" for k in IDX3}, {})
    assert not np.any(dm.get_rhs_glob_3d(d, Pn, Pf, BCdir, mif, dirfaces))


def test_diamond_3d_tangential_dirichlet_is_not_exact(dd3):
    """With a tangential gradient on the Dirichlet faces the rows of the cells on those faces
    are not exact (the mirror value is the point reflection through the face centre), while
    the rows away from them stay exact."""
    d = dd3
    f = lambda x, y, z: 0.3 + 1.7 * x + 0.8 * y
    bcs = {k: "dirichlet" for k in IDX3}
    A, BCdir, dirfaces, mif = _matrix3(d, bcs)
    Pn, Pf = dm.update_ghost_values(d, bcs, {k: f for k in IDX3})
    rhs0 = dm.get_rhs_glob_3d(d, Pn, Pf, BCdir, mif, dirfaces)
    c = d.cells.center
    r = np.abs(A @ f(c[:, 0], c[:, 1], c[:, 2]) - rhs0)
    wall = np.zeros(int(d.nbcells), bool)
    wall[d.faces.cellid[dirfaces, 0]] = True
    assert r[~wall].max() < 1e-12 * np.abs(rhs0).max()
    assert r[wall].max() > 1e-6 * np.abs(rhs0).max()


def test_centertovertex_and_face_gradient_3d_are_exact_on_affine_fields(dd3):
    """The pseudo-Laplacian node weights reproduce affine fields (cell and mirror centres) and
    the diamond face gradient returns the field's gradient, in x, y and z, on inner and
    boundary faces."""
    d = dd3
    g = lambda p: 0.3 + 1.7 * p[:, 0] - 0.6 * p[:, 1] + 0.9 * p[:, 2]
    wc, wg = g(d.cells.center), g(d.faces.ghostcenter[:, :3])
    wn = dm.centertovertex_3d(d, wc, wg)
    assert np.abs(wn - g(d.nodes.vertex[:, :3])).max() < 1e-12
    bnd = np.asarray(d.boundaryfaces)
    grads = dm.face_gradient_3d(d, wc, wg, wn, bnd)
    fl = np.concatenate([np.asarray(d.innerfaces), bnd])
    for k, b in enumerate((1.7, -0.6, 0.9)):
        assert np.abs(grads[k][fl] - b).max() < 1e-11


def test_p_gradient_3d_is_minus_the_gradient(dd3):
    """compute_P_gradient_3d_diamond returns -grad P for P = a + b x with Dirichlet data on the
    x sides: inner faces, Neumann faces (ghost = owner value) and Dirichlet faces (the
    Dirichlet node values and the mirror 2 V_K - V_L)."""
    d = dd3
    f = lambda x, y, z: 0.3 + 1.7 * x
    BCdir, dirfaces, neufaces = _lists3(d, X_DIR)
    Pn, Pf = dm.update_ghost_values(d, X_DIR, {"in": f, "out": f})
    c = d.cells.center
    Pc = f(c[:, 0], c[:, 1], c[:, 2])
    Pg = np.zeros(int(d.nbfaces))
    Pg[neufaces] = Pc[d.faces.cellid[neufaces, 0]]
    v = d.nodes.vertex
    Pnode = f(v[:, 0], v[:, 1], v[:, 2])
    px, py, pz = dm.p_gradient_3d(d, Pc, Pg, Pnode, Pn, Pf, BCdir, neufaces, dirfaces)
    fl = np.concatenate([np.asarray(d.innerfaces), neufaces, dirfaces])
    assert np.abs(px[fl] + 1.7).max() < 1e-11
    assert np.abs(py[fl]).max() < 1e-11 and np.abs(pz[fl]).max() < 1e-11
