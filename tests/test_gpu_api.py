"""The manapy-compatible API on the GPU: solver classes driven like manapy's example
drivers, and HIPBackend.compile with the reference kernels' positional signatures,
against the CPU restatement (oracle) -- bit-exact."""
import numpy as np
import pytest

from conftest import gpu_available, make_tri_domain, make_tet_domain

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]


def test_advection_solver_reference_sequence_and_run(oracle_lib):
    """examples/2D/advection2d.py loop: u.face/v.face -> interpolate_facetocell -> stepper ->
    compute_fluxes -> compute_new_val, all-Neumann, order 2."""
    from manapy_amd import Struct, Variable, AdvectionSolver
    d = make_tri_domain(30, 24, 1.0, 1.0, 0.1, 5)
    C, F = int(d.nbcells), int(d.nbfaces)
    c = d.cells.center
    w0 = np.exp(-((c[:, 0] - .3) ** 2 + (c[:, 1] - .3) ** 2) / .005)
    for order in (1, 2):
        ne, u, v = Variable(domain=d), Variable(domain=d), Variable(domain=d)
        ne.cell[:] = __import__('torch').as_tensor(w0)
        S = AdvectionSolver(ne, vel=(u, v), conf=Struct(order=order, cfl=0.8))
        t = 0.
        for _ in range(25):
            u.face[:] = 1.0
            v.face[:] = 0.5
            u.interpolate_facetocell()
            v.interpolate_facetocell()
            t += S.stepper()
            S.compute_fluxes()
            S.compute_new_val()
        w = w0.copy()
        uf, vf, wf = np.full(F, 1.0), np.full(F, 0.5), np.zeros(F)
        uc, vc, wc = np.zeros(C), np.zeros(C), np.zeros(C)
        oracle_lib.facetocell(d, uf, uc)
        oracle_lib.facetocell(d, vf, vc)
        t_ref = sum(oracle_lib.advection_step(d, w, uc, vc, wc, uf, vf, wf, 0.8, order) for _ in range(25))
        assert np.array_equal(ne.cell.cpu().numpy(), w)
        assert t == t_ref
        # fused run continues from the same state
        ne2 = Variable(domain=d)
        ne2.cell[:] = ne.cell
        S2 = AdvectionSolver(ne2, vel=(u, v), conf=Struct(order=order, cfl=0.8))
        t2 = S2.run(15)
        t_ref2 = sum(oracle_lib.advection_step(d, w, uc, vc, wc, uf, vf, wf, 0.8, order) for _ in range(15))
        assert np.array_equal(ne2.cell.cpu().numpy(), w)
        assert abs(t2 - t_ref2) <= 1e-15 * t_ref2


def test_dirichlet_advection_matches_oracle(oracle_lib):
    """Dirichlet in/out, Neumann elsewhere (examples/2D/advection2d.py:76-83) through update_ghost_value."""
    import torch
    from manapy_amd import Struct, Variable, AdvectionSolver
    d = make_tri_domain(20, 16, 1.0, 1.0, 0.1, 6)
    C, F = int(d.nbcells), int(d.nbfaces)
    bcs = {"in": "dirichlet", "out": "dirichlet", "upper": "neumann", "bottom": "neumann"}
    vals = {"in": 2.0, "out": 0.0}
    ne, u, v = Variable(domain=d, BC=bcs, values=vals), Variable(domain=d), Variable(domain=d)
    u.face[:] = 2.0
    u.interpolate_facetocell()
    v.interpolate_facetocell()
    S = AdvectionSolver(ne, vel=(u, v), conf=Struct(order=1, cfl=0.8))
    for _ in range(20):
        S.stepper()
        S.compute_fluxes()
        S.compute_new_val()
    w = np.zeros(C)
    uf, vf, wf = np.full(F, 2.0), np.zeros(F), np.zeros(F)
    uc, vc, wc = np.zeros(C), np.zeros(C), np.zeros(C)
    oracle_lib.facetocell(d, uf, uc)
    val = np.zeros(F)
    val[d.infaces] = 2.0
    z = np.zeros(C)
    for _ in range(20):
        dt = oracle_lib.time_step(d, uc, vc, wc, 0.8)
        g = np.zeros(F)
        oracle_lib.ghost_value(d, 2, val, g, np.concatenate([d.infaces, d.outfaces]))
        oracle_lib.ghost_value(d, 0, w, g, np.concatenate([d.upperfaces, d.bottomfaces]))
        rez = np.zeros(C)
        oracle_lib.explicitscheme_convective(d, rez, w, g, uf, vf, wf, z, z, z, z, 1)
        oracle_lib.update_new_value(w, rez, z, z, dt, d.cells.volume)
    assert np.array_equal(ne.cell.cpu().numpy(), w)
    assert w.max() > 0.5


def test_shallow_water_solver(oracle_lib):
    import torch
    from manapy_amd import Struct, Variable, ShallowWaterSolver, gaussian_bumps
    d = make_tri_domain(30, 26, 100., 100., 0.1, 7)
    C = int(d.nbcells)
    c = d.cells.center
    Zc = gaussian_bumps(c, 8, np.random.default_rng(2), [(0., 100.), (0., 100.)])
    h0 = np.where(c[:, 0] < 50., 2.0, 1.0)
    walls = {"in": "nonslip", "out": "nonslip", "upper": "nonslip", "bottom": "nonslip"}
    res = {}
    for mode in ("sequence", "run"):
        h, hu, hv, Z = Variable(domain=d), Variable(domain=d, BC=walls), Variable(domain=d, BC=walls), \
            Variable(domain=d)
        h.cell[:] = torch.as_tensor(h0)
        Z.cell[:] = torch.as_tensor(Zc)
        S = ShallowWaterSolver(h=h, hvel=(hu, hv), Z=Z, conf=Struct(order=2, cfl=0.8, flux='rusanov'))
        if mode == "sequence":
            t = 0.
            for _ in range(30):
                S.stepper()
                t += S.dt
                S.compute_fluxes()
                S.compute_new_val()
        else:
            t = S.run(30)
        res[mode] = (h.cell.cpu().numpy(), hu.cell.cpu().numpy(), hv.cell.cpu().numpy(), t)
    hh, uu, vv = h0.copy(), np.zeros(C), np.zeros(C)
    zg = oracle_lib.z_gradient(d, Zc)
    t_ref = sum(oracle_lib.sw_step(d, hh, uu, vv, Zc, zg, 9.81, 0.8, 2) for _ in range(30))
    for mode, (a, b, cc, t) in res.items():
        assert np.array_equal(a, hh), mode
        assert np.array_equal(b, uu), mode
        assert np.array_equal(cc, vv), mode
    assert res["sequence"][3] == t_ref
    assert abs(res["run"][3] - t_ref) <= 1e-14 * t_ref


@pytest.mark.parametrize("quirk", [False, True])
def test_euler_solver(oracle_lib, quirk):
    from manapy_amd import Struct, EulerSolver3D, initialisation_euler_3d
    d = make_tet_domain(5, 0.08, 3)
    U0 = initialisation_euler_3d(d.cells.center)
    a = EulerSolver3D(d, U0=U0, conf=Struct(cfl=0.8, gamma=1.4, quirk=quirk))
    for _ in range(15):
        a.step()
    b = EulerSolver3D(d, U0=U0, conf=Struct(cfl=0.8, gamma=1.4, quirk=quirk))
    b.run(15)                   # packed records on face tiles
    c = EulerSolver3D(d, U0=U0, conf=Struct(cfl=0.8, gamma=1.4, quirk=quirk))
    c.run(15, packed=False)     # the reference's [6][C] arrays
    U = U0.copy()
    G = np.zeros((6, int(d.nbfaces)))
    for _ in range(15):
        oracle_lib.euler_step(d, U, G, 0.8, 1.4, int(quirk))
    assert np.array_equal(a.U.cpu().numpy(), U)
    assert np.array_equal(b.U.cpu().numpy(), U)
    assert np.array_equal(c.U.cpu().numpy(), U)
    assert b.dt == c.dt


def test_backend_compile_reference_signatures(oracle_lib):
    """A reference driver binding kernels through domain.backend.compile runs unchanged with
    numpy arrays (mirrored to the device) and gets the oracle's bits."""
    d = make_tri_domain(16, 12, 1.0, 1.0, 0.1, 8)
    be = d.backend
    C, F = int(d.nbcells), int(d.nbfaces)
    rng = np.random.default_rng(1)
    w = rng.random(C)
    g = np.zeros(F)
    ghost = be.compile('ghost_value_neumann')
    ghost(w, g, d.faces.cellid, d.boundaryfaces, np.zeros(1), np.zeros(1))
    g_ref = np.zeros(F)
    oracle_lib.ghost_value(d, 0, w, g_ref, d.boundaryfaces)
    assert np.array_equal(g, g_ref)
    wx, wy, wz, psi = np.zeros(C), np.zeros(C), np.zeros(C), np.zeros(C)
    be.compile('cell_gradient_2d')(w, g, None, None, d.cells.center, d.cells.cellnid, d.cells.ghostnid,
                                   d.cells.haloghostnid, d.cells.halonid, d.cells.nodeid, d.cells.periodicnid,
                                   d.nodes.periodicid, d.faces.ghostcenter, None, d.nodes.vertex, None, d.cells.shift,
                                   wx, wy, wz)
    be.compile('barthlimiter_2d')(w, g, None, wx, wy, wz, psi, d.faces.cellid, d.cells.faceid, d.faces.name,
                                  d.faces.halofid, d.cells.center, d.faces.center)
    r = [np.zeros(C) for _ in range(4)]
    oracle_lib.cell_gradient(d, w, g_ref, r[0], r[1], r[2])
    oracle_lib.barthlimiter(d, w, g_ref, r[0], r[1], r[2], r[3])
    for a, b in zip((wx, wy, wz, psi), r):
        assert np.array_equal(a, b)
    uf, vf, wf = rng.uniform(-1, 1, F), rng.uniform(-1, 1, F), np.zeros(F)
    rez = np.zeros(C)
    e = np.zeros(0)
    be.compile('explicitscheme_convective_2d')(rez, w, g, e, uf, vf, wf, wx, wy, wz, e, e, e, psi, e,
                                               d.cells.center, d.faces.center, d.halos.centvol, d.faces.ghostcenter,
                                               d.faces.cellid, d.faces.normal, d.faces.halofid, d.faces.name,
                                               d.innerfaces, d.halofaces, d.boundaryfaces, d.periodicboundaryfaces,
                                               d.cells.shift, 2)
    rez_ref = np.zeros(C)
    oracle_lib.explicitscheme_convective(d, rez_ref, w, g_ref, uf, vf, wf, wx, wy, wz, psi, 2)
    assert np.array_equal(rez, rez_ref)
    uc, vc = rng.uniform(-1, 1, C), rng.uniform(-1, 1, C)
    dt = be.compile('time_step')(uc, vc, np.zeros(C), 0.8, d.faces.normal, d.faces.mesure, d.cells.volume,
                                 d.cells.faceid, 2)
    assert dt == oracle_lib.time_step(d, uc, vc, np.zeros(C), 0.8)
    with pytest.raises(NotImplementedError):
        be.compile('srnh_scheme')


@pytest.mark.parametrize("quirk", [0, 1])
def test_backend_compile_euler_driver(oracle_lib, quirk):
    """A reference Euler driver (ghost_value_TubeSchok3D -> time_step_euler_3d with the min over
    dt_c -> explicitscheme_euler_3d_rusanov -> update_euler_3d_fvc) binding the kernels through
    domain.backend.compile with the reference's separate NumPy arrays gets the oracle's bits."""
    from manapy_amd import initialisation_euler_3d
    d = make_tet_domain(5, 0.08, 3)
    be = d.backend
    C, F = int(d.nbcells), int(d.nbfaces)
    U = initialisation_euler_3d(d.cells.center)
    U[3] = 0.05 * np.sin(7 * d.cells.center[:, 1])  # some v, w, so the slip quirk shows
    U[4] = -0.03 * np.cos(5 * d.cells.center[:, 2])
    U[5] = 0.5 * (U[2] ** 2 + U[3] ** 2 + U[4] ** 2) / U[0] + U[1] / 0.4
    cell = [U[k].copy() for k in range(6)]
    ghost = [np.zeros(F) for _ in range(6)]
    vel = [np.zeros(F) for _ in range(3)]
    rez = [np.zeros(C) for _ in range(5)]
    dtc = np.zeros(C)
    e = np.zeros(0)
    ghost_k = be.compile('ghost_value_TubeSchok3D')
    dt_k = be.compile('time_step_euler_3d')
    flux_k = be.compile('explicitscheme_euler_3d_rusanov')
    upd_k = be.compile('update_euler_3d_fvc')
    Ur, Gr = U.copy(), np.zeros((6, F))
    for _ in range(6):
        rho, P, ru, rv, rw, rE = cell
        ghost_k(ghost[0], ghost[1], ghost[2], ghost[3], ghost[4], vel[0], vel[1], vel[2], ghost[5], rho, P, ru, rv,
                rw, rE, d.faces.cellid, d.faces.name, d.faces.normal, d.faces.mesure, quirk=quirk)
        dt_k(rho, P, ru, rv, rw, 0.8, d.faces.normal, d.faces.mesure, d.cells.volume, d.cells.faceid, 1.4, dtc)
        dt = min(1e6, float(dtc.min()))
        flux_k(*rez, rho, P, ru, rv, rw, rE, *ghost, e, e, e, e, e, e, d.faces.cellid, d.faces.halofid,
               d.faces.normal, d.faces.mesure, d.faces.name, 1.4, e, e, quirk=quirk)
        upd_k(rho, P, ru, rv, rw, rE, *rez, 1.4, dt, d.cells.volume)
        oracle_lib.euler_ghost(d, Ur, Gr, quirk)
        dtc_ref = oracle_lib.euler_dt(d, Ur, 0.8, 1.4)
        assert np.array_equal(dtc, dtc_ref)
        assert np.array_equal(np.stack(ghost), Gr)
        bf = d.boundaryfaces
        assert np.array_equal(vel[0][bf], Gr[2, bf] / Gr[0, bf])
        assert np.array_equal(vel[2][bf], Gr[4, bf] / Gr[0, bf])
        rez_ref = oracle_lib.euler_rusanov(d, Ur, Gr, 1.4, quirk)
        assert np.array_equal(np.stack(rez), rez_ref)
        oracle_lib.euler_update(d, Ur, rez_ref, 1.4, dt)
        assert np.array_equal(np.stack(cell), Ur)
