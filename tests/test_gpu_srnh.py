"""The reference's shallow-water scheme on the GPU (SRNH flux, srnh source, update_SW on
five fields) against the CPU restatement: kernels one by one and 20-step runs through
ShallowWaterSolver's reference call sequence and its device-resident run().

Tolerance: the flux calls acos, cos and pow, whose device and host implementations may
differ in the last bit, so the comparison is relative to each field's maximum: 1e-13 for
single kernels and 1e-12 after 20 steps (the parity bar for the float64 path)."""
import numpy as np
import pytest

from conftest import gpu_available, make_tri_domain

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]

G = 9.81


def _close(a, b, tol):
    a, b = np.asarray(a), np.asarray(b)
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() <= tol * scale


def _case(seed=3):
    from manapy_amd import gaussian_bumps
    d = make_tri_domain(26, 22, 100., 100., 0.1, seed)
    C = int(d.nbcells)
    c = d.cells.center
    S = np.zeros((5, C))
    S[4] = gaussian_bumps(c, 8, np.random.default_rng(seed), [(0., 100.), (0., 100.)])
    S[0] = np.where(c[:, 0] < 50., 2.0, 1.0)
    S[3] = 0.1 * S[0]
    return d, S


def _dev(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device='cuda')


@pytest.mark.parametrize("order", [1, 2])
def test_srnh_kernels_match_oracle(oracle_lib, order):
    import torch
    from manapy_amd import _lib
    from manapy_amd.device import stream_ptr
    d, S = _case()
    C = int(d.nbcells)
    rng = np.random.default_rng(7)
    S[1] = rng.uniform(-0.3, 0.3, C)
    S[2] = rng.uniform(-0.3, 0.3, C)
    Gh = oracle_lib.srnh_ghosts(d, S)
    grads = oracle_lib.srnh_gradients(d, S, Gh) if order == 2 else tuple(np.zeros(C) for _ in range(5))
    rez_ref = oracle_lib.srnh_convective(d, S, Gh, grads, G, order)
    m = d.device_mesh
    s = stream_ptr()
    tS, tG = _dev(S), _dev(Gh)
    tg = [_dev(x) for x in grads]
    rez = torch.full((5, C), np.nan, dtype=torch.float64, device='cuda')
    _lib.call('mfv_explicitscheme_convective_sw', m.ptr, *[rez[k].data_ptr() for k in range(5)],
              *[tS[k].data_ptr() for k in range(5)], *[tG[k].data_ptr() for k in range(5)],
              *[x.data_ptr() for x in tg], G, order, s)
    r = rez.cpu().numpy()
    for k in range(5):
        assert _close(r[k], rez_ref[k], 1e-13) or np.abs(rez_ref[k]).max() == 0. and np.abs(r[k]).max() == 0., k
    for quirk in (0, 1):
        src_ref = oracle_lib.srnh_source(d, S, Gh, G, quirk=bool(quirk))
        src = torch.full((5, C), np.nan, dtype=torch.float64, device='cuda')
        nid = torch.as_tensor(d.cells.nodeid, device='cuda')
        cfid = torch.as_tensor(d.cells.cellfid, device='cuda')
        nf = _dev(d.cells.nf)
        vx = _dev(d.nodes.vertex)
        _lib.call('mfv_term_source_srnh_sw', m.ptr, *[src[k].data_ptr() for k in range(5)], tS[0].data_ptr(),
                  tS[4].data_ptr(), tG[0].data_ptr(), tG[4].data_ptr(), nid.data_ptr(), int(nid.shape[1]),
                  cfid.data_ptr(), int(cfid.shape[1]), nf.data_ptr(), vx.data_ptr(), G, quirk, s)
        # no transcendental functions in the source: bit-exact
        assert np.array_equal(src.cpu().numpy(), src_ref), quirk
    # update_SW with all terms
    dt = 0.0123
    cor = rng.normal(size=(3, C))
    Sref = S.copy()
    oracle_lib.srnh_update(d, Sref, rez_ref, src_ref, dt, dissip_hc=cor[2], corio_hu=cor[0], corio_hv=cor[1],
                           wind=(0.3, -0.1))
    tS2 = _dev(S)
    trez, tsrc, tcor = _dev(rez_ref), _dev(src_ref), _dev(cor)
    _lib.call('mfv_update_sw5', m.ptr, *[tS2[k].data_ptr() for k in range(5)], *[trez[k].data_ptr() for k in range(5)],
              *[tsrc[k].data_ptr() for k in range(5)], tcor[2].data_ptr(), tcor[0].data_ptr(), tcor[1].data_ptr(),
              0.3, -0.1, dt, None, s)
    assert np.array_equal(tS2.cpu().numpy(), Sref)


@pytest.mark.parametrize("order,quirk", [(1, True), (2, True), (2, False)])
def test_srnh_solver_20_steps(oracle_lib, order, quirk):
    import torch
    from manapy_amd import Struct, Variable, ShallowWaterSolver
    d, S0 = _case(5)
    walls = {"in": "nonslip", "out": "nonslip", "upper": "nonslip", "bottom": "nonslip"}
    res = {}
    for mode in ("sequence", "run"):
        h, hu, hv = Variable(domain=d), Variable(domain=d, BC=walls), Variable(domain=d, BC=walls)
        hc, Z = Variable(domain=d), Variable(domain=d)
        for var, k in ((h, 0), (hc, 3), (Z, 4)):
            var.cell[:] = torch.as_tensor(S0[k])
        sol = ShallowWaterSolver(h=h, hvel=(hu, hv), hc=hc, Z=Z,
                                 conf=Struct(order=order, cfl=0.8, flux='srnh', quirk=quirk))
        if mode == "sequence":
            t = 0.
            for _ in range(20):
                sol.stepper()
                t += sol.dt
                sol.compute_fluxes()
                sol.compute_new_val()
        else:
            t = sol.run(20)
        res[mode] = ([v.cell.cpu().numpy() for v in (h, hu, hv, hc, Z)], t)
    S = S0.copy()
    t_ref = 0.
    for _ in range(20):
        t_ref += oracle_lib.srnh_step(d, S, G, 0.8, order, quirk=quirk)
    for mode, (fields, t) in res.items():
        for k in range(5):
            assert _close(fields[k], S[k], 1e-12), (mode, k)
        assert abs(t - t_ref) <= 1e-12 * t_ref, mode
    # the two host paths run the same kernels in the same order
    for k in range(5):
        assert np.array_equal(res["sequence"][0][k], res["run"][0][k]), k


def test_backend_compile_srnh_reference_signatures(oracle_lib):
    """The reference's own call (shallowater/system.py:162-178, 204-215, 237-242) through
    domain.backend.compile with numpy arrays."""
    d, S = _case(9)
    be = d.backend
    C, F = int(d.nbcells), int(d.nbfaces)
    rng = np.random.default_rng(2)
    S[1] = rng.uniform(-0.2, 0.2, C)
    S[2] = rng.uniform(-0.2, 0.2, C)
    Gh = oracle_lib.srnh_ghosts(d, S)
    grads = oracle_lib.srnh_gradients(d, S, Gh)
    e = np.zeros(0)
    rez = [np.zeros(C) for _ in range(5)]
    be.compile('explicitscheme_convective_SW')(
        *rez, *[S[k].copy() for k in range(5)], *[Gh[k].copy() for k in range(5)], e, e, e, e, e,
        grads[0], grads[1], e, e, grads[2], grads[3], e, e, grads[4], e, d.cells.center, d.faces.center,
        d.halos.centvol, d.faces.ghostcenter, d.faces.cellid, d.faces.mesure, d.faces.normal, d.faces.halofid,
        d.innerfaces, d.halofaces, d.boundaryfaces, G, 2)
    ref = oracle_lib.srnh_convective(d, S, Gh, grads, G, 2)
    for k in range(4):
        assert _close(rez[k], ref[k], 1e-13), k
    src = [np.full(C, np.nan) for _ in range(5)]
    be.compile('term_source_srnh_SW')(
        *src, *[S[k] for k in range(5)], *[Gh[k] for k in range(5)], e, e, e, e, e, grads[0], grads[1], grads[4],
        e, e, e, d.cells.nodeid, d.cells.faceid, d.cells.cellfid, d.faces.cellid, d.cells.center, d.cells.nf,
        d.faces.name, d.faces.center, d.halos.centvol, d.nodes.vertex, d.faces.halofid, G, 2)
    sref = oracle_lib.srnh_source(d, S, Gh, G, quirk=True)
    assert np.array_equal(np.array(src), sref)
    S2 = [S[k].copy() for k in range(5)]
    z = np.zeros(C)
    be.compile('update_SW')(*S2, *ref, *sref, z, z, z, 0., 0., 0.01, d.cells.volume)
    Sref = S.copy()
    oracle_lib.srnh_update(d, Sref, ref, sref, 0.01)
    assert np.array_equal(np.array(S2), Sref)


def test_friction_coriolis_wind_kernels(oracle_lib):
    """term_friction_SW (pow on the device and the host may differ in the last bit: 1e-15 of
    the field), term_coriolis_SW and term_wind_SW (bit for bit) through HIPBackend.compile with
    the kernels' own signatures."""
    d, S = _case(7)
    C = int(d.nbcells)
    rng = np.random.default_rng(2)
    h, hu, hv = S[0] + 0.1 * rng.random(C), rng.standard_normal(C), rng.standard_normal(C)
    be = d.backend
    hu_g, hv_g = hu.copy(), hv.copy()
    be.compile('term_friction_SW')(h, hu_g, hv_g, G, 0.03, 0.7)
    hu_r, hv_r = hu.copy(), hv.copy()
    oracle_lib.term_friction_sw(h, hu_r, hv_r, G, 0.03, 0.7)
    assert _close(hu_g, hu_r, 1e-15) and _close(hv_g, hv_r, 1e-15)
    assert not np.allclose(hu_r, hu)
    f = rng.random(C)
    cu, cv, cu_r, cv_r = (np.zeros(C) for _ in range(4))
    be.compile('term_coriolis_SW')(hu, hv, cu, cv, f)
    oracle_lib.term_coriolis_sw(hu, hv, cu_r, cv_r, f)
    assert np.array_equal(cu, cu_r) and np.array_equal(cv, cv_r)
    be.compile('term_coriolis_SW')(hu, hv, cu, cv, 0.3)
    oracle_lib.term_coriolis_sw(hu, hv, cu_r, cv_r, np.full(C, 0.3))
    assert np.array_equal(cu, cu_r) and np.array_equal(cv, cv_r)
    wx, wy = rng.standard_normal(C), rng.standard_normal(C)
    tx, ty, tx_r, ty_r = (np.zeros(C) for _ in range(4))
    be.compile('term_wind_SW')(tx, ty, wx, wy)
    oracle_lib.term_wind_sw(tx_r, ty_r, wx, wy)
    assert np.array_equal(tx, tx_r) and np.array_equal(ty, ty_r)


def test_srnh_solver_with_friction_and_coriolis(oracle_lib):
    """ShallowWaterSolver with Mann and fc: friction with the driver's dt, the second stepper
    call of compute_fluxes (whose dt the update uses), Coriolis into update_SW; 15 steps of the
    reference sequence and of run() against the oracle's sequence."""
    import torch
    from manapy_amd import Struct, Variable, ShallowWaterSolver
    d, S0 = _case(9)
    walls = {"in": "nonslip", "out": "nonslip", "upper": "nonslip", "bottom": "nonslip"}
    res = {}
    for mode in ("sequence", "run"):
        h, hu, hv = Variable(domain=d), Variable(domain=d, BC=walls), Variable(domain=d, BC=walls)
        hc, Z = Variable(domain=d), Variable(domain=d)
        for var, k in ((h, 0), (hc, 3), (Z, 4)):
            var.cell[:] = torch.as_tensor(S0[k])
        sol = ShallowWaterSolver(h=h, hvel=(hu, hv), hc=hc, Z=Z,
                                 conf=Struct(order=2, cfl=0.8, flux='srnh', Mann=0.03, fc=1e-2))
        if mode == "sequence":
            t = 0.
            for _ in range(15):
                sol.stepper()
                sol.compute_fluxes()
                t += sol.dt
                sol.compute_new_val()
        else:
            t = sol.run(15)
        res[mode] = ([v.cell.cpu().numpy() for v in (h, hu, hv, hc, Z)], t)
    S = S0.copy()
    t_ref = 0.
    for _ in range(15):
        t_ref += oracle_lib.srnh_step(d, S, G, 0.8, 2, Mann=0.03, fc=1e-2)
    for mode, (fields, t) in res.items():
        for k in range(5):
            assert _close(fields[k], S[k], 1e-12), (mode, k)
        assert abs(t - t_ref) <= 1e-12 * t_ref, mode
    assert np.abs(S[2]).max() > 0.
