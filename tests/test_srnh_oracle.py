"""Pins of the CPU restatement of the reference's shallow-water scheme (oracle/srnh.c):
the SRNH eigenvalues and sign matrix against the rotated Jacobian, consistency and
left/right antisymmetry of the flux, exact lake-at-rest balance of flux + source at first
order, the source's closed form for a uniform depth, the reference's compacted-cellfid
reading, and update_SW's conservation identity."""
import numpy as np
import pytest

from conftest import make_tri_domain

G = 9.81


def _jacobian(h, u, v, c):
    """A of the rotated system W = (h, q_n, q_t, hc, Z) with the bathymetry term g h dZ
    (non-conservative) and a passive tracer: dW/dt + A dW/dx = 0."""
    s2 = G * h
    return np.array([[0., 1., 0., 0., 0.],
                     [s2 - u * u, 2. * u, 0., 0., s2],
                     [-u * v, v, u, 0., 0.],
                     [-u * c, c, 0., u, 0.],
                     [0., 0., 0., 0., 0.]])


def test_srnh_eigenvalues_and_sign_matrix(oracle_lib):
    rng = np.random.default_rng(0)
    for _ in range(20):
        h, u, v, c = rng.uniform(0.2, 4.), rng.uniform(-2., 2.), rng.uniform(-2., 2.), rng.uniform(0., 3.)
        lam, r = oracle_lib.srnh_roe(h, u, v, c, G)
        cs = np.sqrt(G * h)
        # the trigonometric cubic's roots are the eigenvalues 0, u - c, u + c
        assert np.allclose(np.sort(lam[:3]), np.sort([0., u - cs, u + cs]), atol=1e-12, rtol=1e-12)
        assert lam[3] == u and lam[4] == u
        w, V = np.linalg.eig(_jacobian(h, u, v, c))
        w = w.real
        sgn = np.where(np.abs(w) < 1e-10, 0., np.sign(w))
        S = (V @ np.diag(sgn) @ np.linalg.inv(V)).real
        assert np.abs(S - r).max() <= 1e-12 * max(1., np.abs(S).max())


def test_srnh_flux_consistency_and_antisymmetry(oracle_lib):
    rng = np.random.default_rng(1)
    for _ in range(20):
        h, hu, hv, hc, z = rng.uniform(0.3, 3.), rng.uniform(-2, 2), rng.uniform(-2, 2), rng.uniform(0, 1), 0.2
        n = np.array([rng.normal(), rng.normal(), 0.])
        mes = np.hypot(n[0], n[1])
        f = oracle_lib.srnh_flux(hu, hu, hv, hv, h, h, hc, hc, z, z, n, mes, G)
        qn = hu * n[0] + hv * n[1]
        want = [qn, qn * hu / h + 0.5 * G * h * h * n[0], qn * hv / h + 0.5 * G * h * h * n[1], qn * hc / h, 0.]
        assert np.allclose(f, want, rtol=1e-12, atol=1e-12)
        # F(L, R, n) = -F(R, L, -n)
        hl, hr = rng.uniform(0.5, 2., 2)
        ul, ur, vl, vr = rng.uniform(-1, 1, 4)
        cl, cr = rng.uniform(0, 1, 2)
        zl, zr = rng.uniform(0, .3, 2)
        f1 = oracle_lib.srnh_flux(ul, ur, vl, vr, hl, hr, cl, cr, zl, zr, n, mes, G)
        f2 = oracle_lib.srnh_flux(ur, ul, vr, vl, hr, hl, cr, cl, zr, zl, -n, mes, G)
        assert np.abs(f1 + f2).max() <= 1e-13 * max(1., np.abs(f1).max())


def _domain():
    return make_tri_domain(14, 11, 1., 1., 0.1, 3)


def test_lake_at_rest_is_balanced_at_first_order(oracle_lib):
    """h + Z = const, u = 0: the SRNH flux and the source cancel in every cell (with the
    true face neighbours). A wrong sign-matrix column or source term breaks this."""
    d = _domain()
    C = int(d.nbcells)
    x, y = d.cells.center[:, 0], d.cells.center[:, 1]
    for Z in (0.1 * np.exp(-((x - .5) ** 2 + (y - .4) ** 2) / 0.05), 0.1 + 0.05 * x - 0.03 * y):
        S = np.zeros((5, C))
        S[4] = Z
        S[0] = 1.0 - Z
        S[3] = 0.3 * S[0]
        Gh = oracle_lib.srnh_ghosts(d, S)
        zero = tuple(np.zeros(C) for _ in range(5))
        rez = oracle_lib.srnh_convective(d, S, Gh, zero, G, 1)
        src = oracle_lib.srnh_source(d, S, Gh, G, quirk=False)
        scale = np.abs(src[1:3]).max()
        assert scale > 1e-3
        assert np.abs(rez[1:3] + src[1:3]).max() <= 1e-12 * scale
        assert np.abs(rez[0]).max() <= 1e-14 and np.abs(rez[3]).max() <= 1e-14


def test_source_closed_form_uniform_depth(oracle_lib):
    """Uniform h: the face depths solve to h and src = -g/2 h sum_j (Z_j - Z_i) nf_j, with
    Z_j the value across face j (true neighbour, or with quirk the compacted cellfid
    entry, -1 wrapping to the last cell)."""
    d = _domain()
    C = int(d.nbcells)
    rng = np.random.default_rng(4)
    S = np.zeros((5, C))
    S[0] = 1.3
    S[4] = rng.uniform(0., .2, C)
    Gh = oracle_lib.srnh_ghosts(d, S)
    fid, cfid, nf, name, cellid = d.cells.faceid, d.cells.cellfid, d.cells.nf, d.faces.name, d.faces.cellid
    for quirk in (False, True):
        src = oracle_lib.srnh_source(d, S, Gh, G, quirk=quirk)
        want = np.zeros((2, C))
        for i in range(C):
            acc = np.zeros(2)
            for j in range(3):
                f = fid[i, j]
                if name[f] == 0:
                    nb = cfid[i, j] if quirk else (cellid[f, 1] if cellid[f, 0] == i else cellid[f, 0])
                    zj = S[4][nb]  # numpy wraps -1 like the reference
                else:
                    zj = Gh[4][f]
                acc += (zj - S[4][i]) * nf[i, j, :2]
            want[:, i] = -G * 0.5 * 1.3 * acc
        assert np.allclose(src[1:3], want, rtol=1e-10, atol=1e-12), quirk
        assert not np.any(src[[0, 3, 4]])
    # the quirk changes only cells with a boundary face
    a = oracle_lib.srnh_source(d, S, Gh, G, quirk=False)
    b = oracle_lib.srnh_source(d, S, Gh, G, quirk=True)
    bnd_cells = np.unique(cellid[d.boundaryfaces, 0])
    inner = np.setdiff1d(np.arange(C), bnd_cells)
    assert np.array_equal(a[:, inner], b[:, inner])
    assert np.any(a[:, bnd_cells] != b[:, bnd_cells])


def test_cellfid_is_compacted_like_the_reference():
    """cells.cellfid lists neighbours across faces in face order, skipping boundary faces
    (ddm_utils2d.py:32-53) -- the layout the source term's quirk depends on."""
    d = _domain()
    fid, cfid, cellid = d.cells.faceid, d.cells.cellfid, d.faces.cellid
    for i in range(int(d.nbcells)):
        nbs = [cellid[f, 1] if cellid[f, 0] == i else cellid[f, 0] for f in fid[i, :fid[i, -1]]]
        nbs = [c for c in nbs if c >= 0]
        assert list(cfid[i, :len(nbs)]) == nbs and cfid[i, -1] == len(nbs)
        assert np.all(cfid[i, len(nbs):-1] == -1)


def test_update_sw5_conserves_and_adds_terms(oracle_lib):
    d = _domain()
    C = int(d.nbcells)
    rng = np.random.default_rng(5)
    S = rng.uniform(0.5, 1.5, (5, C))
    rez = rng.normal(size=(5, C)) * 1e-3
    src = rng.normal(size=(5, C)) * 1e-3
    vol = d.cells.volume
    S0 = S.copy()
    oracle_lib.srnh_update(d, S, rez, src, 0.0)
    assert np.array_equal(S, S0)
    dt = 0.01
    cor = rng.normal(size=(2, C))
    oracle_lib.srnh_update(d, S, rez, src, dt, corio_hu=cor[0], corio_hv=cor[1], wind=(0.5, -0.25))
    # mass changes by dt * sum(rez_h + src_h) (to rounding)
    dm = np.sum((S[0] - S0[0]) * vol)
    assert abs(dm - dt * np.sum(rez[0] + src[0])) <= 1e-12
    # with zero residuals the momentum moves by dt * (coriolis + wind) exactly per cell
    S1 = S0.copy()
    z = np.zeros((5, C))
    oracle_lib.srnh_update(d, S1, z, z, dt, corio_hu=cor[0], corio_hv=cor[1], wind=(0.5, -0.25))
    assert np.allclose(S1[1] - S0[1], dt * (cor[0] + 0.5), rtol=1e-12, atol=1e-15)
    assert np.allclose(S1[2] - S0[2], dt * (cor[1] - 0.25), rtol=1e-12, atol=1e-15)
    assert np.array_equal(S1[[0, 3, 4]], S0[[0, 3, 4]])


@pytest.mark.parametrize("order", [1, 2])
def test_srnh_step_conserves_mass_with_walls_interior(oracle_lib, order):
    """Inner faces cancel: the change of total mass equals the wall-face mass fluxes only."""
    d = _domain()
    C = int(d.nbcells)
    x = d.cells.center[:, 0]
    S = np.zeros((5, C))
    S[0] = np.where(x < .5, 2., 1.)
    S[4] = 0.05 * np.sin(3 * x)
    S[3] = 0.1 * S[0]
    Gh = oracle_lib.srnh_ghosts(d, S)
    grads = oracle_lib.srnh_gradients(d, S, Gh) if order == 2 else tuple(np.zeros(C) for _ in range(5))
    rez = oracle_lib.srnh_convective(d, S, Gh, grads, G, order)
    # boundary-face fluxes recomputed face by face
    wall = 0.
    for f in d.boundaryfaces:
        L = d.faces.cellid[f, 0]
        fl = oracle_lib.srnh_flux(S[1][L], Gh[1][f], S[2][L], Gh[2][f],
                                  S[0][L] + (order - 1) * grads[4][L] * (
                                      grads[0][L] * (d.faces.center[f, 0] - d.cells.center[L, 0]) +
                                      grads[1][L] * (d.faces.center[f, 1] - d.cells.center[L, 1])),
                                  Gh[0][f],
                                  S[3][L] + (order - 1) * grads[4][L] * (
                                      grads[2][L] * (d.faces.center[f, 0] - d.cells.center[L, 0]) +
                                      grads[3][L] * (d.faces.center[f, 1] - d.cells.center[L, 1])),
                                  Gh[3][f], S[4][L], Gh[4][f], d.faces.normal[f], d.faces.mesure[f], G)
        wall += fl[0]
    assert abs(np.sum(rez[0]) + wall) <= 1e-12 * max(1., abs(wall))


def test_term_friction_closed_form(oracle_lib):
    """term_friction_SW (fvm_utils.py:512-529): with g time eta^2 = 1, h = 8 (h^(4/3) = 16) and
    |u| = 3, A = 1 + 3/16, so the momenta shrink by 16/19 and keep their direction; a cell at
    rest is unchanged."""
    h = np.array([8., 8., 2.])
    hu = np.array([1.8 * 8, -3. * 8, 0.])
    hv = np.array([2.4 * 8, 0., 0.])
    hu0, hv0 = hu.copy(), hv.copy()
    oracle_lib.term_friction_sw(h, hu, hv, 2.0, 0.5, 2.0)
    assert np.allclose(hu[:2], hu0[:2] * 16. / 19., rtol=1e-14)
    assert np.allclose(hv[:2], hv0[:2] * 16. / 19., rtol=1e-14)
    assert hu[2] == 0. and hv[2] == 0.


def test_term_coriolis_and_wind_closed_forms(oracle_lib):
    """term_coriolis_SW keeps the reference's signs (corio_hu = f hu, corio_hv = -f hv);
    term_wind_SW gives 1.25 (0.75 + 0.067 |w|) 1e-3 |w| w: for w = (3, 4), 0.0203438 and 0.027125."""
    hu, hv, f = np.array([1., -2.]), np.array([3., 4.]), np.array([0.5, 0.25])
    cu, cv = np.zeros(2), np.zeros(2)
    oracle_lib.term_coriolis_sw(hu, hv, cu, cv, f)
    assert np.array_equal(cu, [0.5, -0.5]) and np.array_equal(cv, [-1.5, -1.0])
    tx, ty = np.zeros(2), np.zeros(2)
    oracle_lib.term_wind_sw(tx, ty, np.array([3., 0.]), np.array([4., 0.]))
    c = 1.25 * (0.75 + 0.067 * 5.) * 1e-3
    assert np.allclose([tx[0], ty[0]], [c * 15., c * 20.], rtol=1e-15) and tx[1] == 0. and ty[1] == 0.
