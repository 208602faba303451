"""GPU parity: every libmfv entry point, called through the C ABI, against the CPU
restatement (oracle) on the same seeded inputs. The bar is bit-exact float64 for
each kernel and after N fused steps: the device kernels evaluate the reference's
expressions in the reference's order with no contraction."""
import ctypes

import numpy as np
import pytest

from conftest import gpu_available, make_tri_domain, make_tet_domain

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]


@pytest.fixture(scope="module")
def torch():
    import torch as t
    return t


@pytest.fixture(scope="module")
def d2():
    return make_tri_domain(40, 30, 1.0, 1.0, 0.1, 21)


@pytest.fixture(scope="module")
def d3():
    return make_tet_domain(6, 0.08, 9)


@pytest.fixture(scope="module")
def m2(d2):
    from manapy_amd.device import DeviceMesh
    return DeviceMesh(d2)


@pytest.fixture(scope="module")
def m3(d3):
    from manapy_amd.device import DeviceMesh
    return DeviceMesh(d3)


def dev(torch, a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def host(t):
    return t.detach().cpu().numpy()


def call(name, *args):
    from manapy_amd import _lib
    from manapy_amd.device import stream_ptr
    _lib.call(name, *args, stream_ptr())


def P(t):
    """Device address of a tensor the caller keeps alive until the call has run: a temporary
    would go back to the caching allocator and its block could be handed to the next one."""
    return t.data_ptr()


def same(a, b, what=""):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, what
    diff = np.nonzero(~((a == b) | (np.isnan(a) & np.isnan(b))))[0] if a.ndim == 1 else None
    assert np.array_equal(a, b, equal_nan=True), "%s: %d differ, max |diff| %g" % (
        what, int(np.sum(a != b)), float(np.nanmax(np.abs(a - b))))
    del diff


def field(d, seed, scale=1.0, offset=0.0):
    rng = np.random.default_rng(seed)
    c = d.cells.center
    return offset + scale * (np.sin(5 * c[:, 0] + 1) * np.cos(3 * c[:, 1]) + 0.05 * rng.standard_normal(len(c)))


# ------------------------------------------------------------------------------- generic operators
def test_ghost_and_facetocell(torch, d2, m2, oracle_lib):
    C, F = int(d2.nbcells), int(d2.nbfaces)
    w = field(d2, 1)
    bf = d2.boundaryfaces
    tbf = torch.as_tensor(bf).cuda()
    for kind, src in ((0, w), (1, w), (2, np.random.default_rng(2).random(F))):
        g_ref = np.full(F, 3.0)
        oracle_lib.ghost_value(d2, kind, src, g_ref, bf)
        g = dev(torch, np.full(F, 3.0))
        tsrc = dev(torch, src)
        call('mfv_ghost_value', m2.ptr, kind, P(tsrc), P(g), P(tbf), len(bf))
        same(host(g), g_ref, "ghost kind %d" % kind)
    uf = np.random.default_rng(3).random(F)
    uc_ref = np.zeros(C)
    oracle_lib.facetocell(d2, uf, uc_ref)
    uc = dev(torch, np.zeros(C))
    tuf = dev(torch, uf)
    call('mfv_facetocell', m2.ptr, P(tuf), P(uc))
    same(host(uc), uc_ref, "facetocell")


@pytest.mark.parametrize("dim", [2, 3])
def test_gradient_and_limiter(torch, dim, d2, m2, d3, m3, oracle_lib):
    d, m = (d2, m2) if dim == 2 else (d3, m3)
    C, F = int(d.nbcells), int(d.nbfaces)
    w = field(d, 4)
    g = np.zeros(F)
    oracle_lib.ghost_value(d, 0, w, g, d.boundaryfaces)
    ref = [np.zeros(C) for _ in range(4)]
    oracle_lib.cell_gradient(d, w, g, ref[0], ref[1], ref[2])
    oracle_lib.barthlimiter(d, w, g, ref[0], ref[1], ref[2], ref[3])
    out = [dev(torch, np.zeros(C)) for _ in range(4)]
    tw, tg = dev(torch, w), dev(torch, g)
    call('mfv_cell_gradient', m.ptr, P(tw), P(tg), P(out[0]), P(out[1]), P(out[2]))
    call('mfv_barthlimiter', m.ptr, P(tw), P(tg), P(out[0]), P(out[1]), P(out[2]), P(out[3]))
    for k, nm in enumerate(('w_x', 'w_y', 'w_z', 'psi')):
        same(host(out[k]), ref[k], nm)
    assert ref[3].min() < 1.0


# ------------------------------------------------------------------------------- advection
@pytest.mark.parametrize("dim,order", [(2, 1), (2, 2), (3, 1), (3, 2)])
def test_advection_kernels(torch, dim, order, d2, m2, d3, m3, oracle_lib):
    d, m = (d2, m2) if dim == 2 else (d3, m3)
    C, F = int(d.nbcells), int(d.nbfaces)
    w = field(d, 5)
    rng = np.random.default_rng(6)
    uf, vf, wf = rng.uniform(-1, 1, F), rng.uniform(-1, 1, F), (rng.uniform(-1, 1, F) if dim == 3 else np.zeros(F))
    g = np.zeros(F)
    oracle_lib.ghost_value(d, 0, w, g, d.boundaryfaces)
    wx, wy, wz, psi = (np.zeros(C) for _ in range(4))
    oracle_lib.cell_gradient(d, w, g, wx, wy, wz)
    oracle_lib.barthlimiter(d, w, g, wx, wy, wz, psi)
    rez_ref = np.zeros(C)
    oracle_lib.explicitscheme_convective(d, rez_ref, w, g, uf, vf, wf, wx, wy, wz, psi, order)
    rez = dev(torch, np.zeros(C))
    args = [dev(torch, a) for a in (w, g, uf, vf, wf, wx, wy, wz, psi)]
    call('mfv_explicitscheme_convective', m.ptr, P(rez), *[P(a) for a in args], order)
    same(host(rez), rez_ref, "rez")
    # time step (cell velocity) and update
    uc, vc, wc = rng.uniform(-1, 1, C), rng.uniform(-1, 1, C), rng.uniform(-1, 1, C)
    dt_ref = oracle_lib.time_step(d, uc, vc, wc, 0.8)
    dtd = dev(torch, np.zeros(1))
    dth = ctypes.c_double(0.)
    tuc, tvc, twc = dev(torch, uc), dev(torch, vc), dev(torch, wc)
    call('mfv_time_step', m.ptr, P(tuc), P(tvc), P(twc), 0.8, P(dtd), ctypes.byref(dth))
    assert dth.value == dt_ref and host(dtd)[0] == dt_ref
    dis, src = rng.random(C), rng.random(C)
    w_ref = w.copy()
    oracle_lib.update_new_value(w_ref, rez_ref, dis, src, dt_ref, d.cells.volume)
    tw = dev(torch, w)
    tdis, tsrc = dev(torch, dis), dev(torch, src)
    call('mfv_update_new_value', m.ptr, P(tw), P(rez), P(tdis), P(tsrc), 0.0, P(dtd))
    same(host(tw), w_ref, "update")


@pytest.mark.parametrize("order", [1, 2])
def test_advection_fused_100_steps(torch, order, d2, m2, oracle_lib):
    """Config-1 style run (Gaussian blob, u = (1, 0.5), all-Neumann, CFL 0.8) for 100 steps:
    the fused device step equals the reference step sequence bit for bit."""
    d, m = d2, m2
    C, F = int(d.nbcells), int(d.nbfaces)
    c = d.cells.center
    w0 = np.exp(-((c[:, 0] - .3) ** 2 + (c[:, 1] - .3) ** 2) / .005)
    uf, vf, wf = np.full(F, 1.0), np.full(F, 0.5), np.zeros(F)
    uc, vc, wc = np.zeros(C), np.zeros(C), np.zeros(C)
    oracle_lib.facetocell(d, uf, uc)
    oracle_lib.facetocell(d, vf, vc)
    w_ref = w0.copy()
    dts = [oracle_lib.advection_step(d, w_ref, uc, vc, wc, uf, vf, wf, 0.8, order) for _ in range(100)]
    a = dev(torch, w0)
    b = dev(torch, np.zeros(C))
    scratch = dev(torch, np.zeros(3 * C))
    dt0 = dev(torch, np.zeros(1))
    ring = dev(torch, np.zeros(4))
    tv = [dev(torch, x) for x in (uc, vc, wc, uf, vf, wf)]
    call('mfv_time_step', m.ptr, P(tv[0]), P(tv[1]), P(tv[2]), 0.8, P(dt0), None)
    call('mfv_dt_ring_init', P(ring), P(dt0))
    for s in range(100):
        call('mfv_advection_step', m.ptr, P(a), P(b), *[P(x) for x in tv], 0.8, order, P(ring), s, P(scratch))
        a, b = b, a
    assert host(ring)[99 % 3] == dts[-1]
    same(host(a), w_ref, "w after 100 steps")
    assert abs(w_ref.max() - w0.max()) > 1e-3  # the blob moved


# ------------------------------------------------------------------------------- shallow water
@pytest.fixture(scope="module")
def sw_case():
    d = make_tri_domain(36, 30, 100., 100., 0.1, 8)
    from manapy_amd.device import DeviceMesh
    from manapy_amd.meshgen import gaussian_bumps
    c = d.cells.center
    Z = gaussian_bumps(c, 8, np.random.default_rng(3), [(0., 100.), (0., 100.)])
    h = np.where(c[:, 0] < 50., 2.0, 1.0)
    return d, DeviceMesh(d), Z, h


def test_sw_kernels(torch, sw_case, oracle_lib):
    d, m, Z, h0 = sw_case
    C = int(d.nbcells)
    rng = np.random.default_rng(9)
    h = h0 + 0.1 * rng.random(C)
    hu, hv = 0.3 * rng.standard_normal(C), 0.3 * rng.standard_normal(C)
    G = 9.81
    dt_ref = oracle_lib.time_step_sw(d, h, hu, hv, 0.8, G)
    th, thu, thv, tZ = (dev(torch, x) for x in (h, hu, hv, Z))
    dtd = dev(torch, np.zeros(1))
    dth = ctypes.c_double()
    call('mfv_time_step_sw', m.ptr, P(th), P(thu), P(thv), 0.8, G, 0., 0., P(dtd), ctypes.byref(dth))
    assert dth.value == dt_ref
    zg_ref = oracle_lib.z_gradient(d, Z)
    grad_ref = oracle_lib.sw_gradients(d, h, hu, hv, Z)
    grad = dev(torch, np.zeros((9, C)))
    call('mfv_sw_gradients', m.ptr, P(th), P(thu), P(thv), P(tZ), P(grad))
    same(host(grad), grad_ref, "sw grad")
    zg = dev(torch, zg_ref)
    rez_ref, src_ref = oracle_lib.sw_residuals(d, h, hu, hv, Z, grad_ref, zg_ref, G, 2)
    rez, src = dev(torch, np.zeros((3, C))), dev(torch, np.zeros((3, C)))
    call('mfv_sw_residuals', m.ptr, P(th), P(thu), P(thv), P(tZ), P(grad), P(zg), G, 2, P(rez), P(src))
    same(host(rez), rez_ref, "sw rez")
    same(host(src), src_ref, "sw src")
    hr, hur, hvr = h.copy(), hu.copy(), hv.copy()
    oracle_lib.update_sw(hr, hur, hvr, rez_ref, src_ref, dt_ref, d.cells.volume)
    call('mfv_update_sw', m.ptr, P(th), P(thu), P(thv), P(rez), P(src), P(dtd))
    same(host(th), hr, "h")
    same(host(thu), hur, "hu")
    same(host(thv), hvr, "hv")


def test_sw_fused_60_steps(torch, sw_case, oracle_lib):
    """Dam break over bumps, MUSCL + LSQ + Barth-Jespersen + Rusanov, walls, CFL 0.8: 60 fused
    device steps equal 60 reference-sequence oracle steps bit for bit."""
    d, m, Z, h0 = sw_case
    C = int(d.nbcells)
    G = 9.81
    h, hu, hv = h0.copy(), np.zeros(C), np.zeros(C)
    zg_ref = oracle_lib.z_gradient(d, Z)
    for _ in range(60):
        dt_last = oracle_lib.sw_step(d, h, hu, hv, Z, zg_ref, G, 0.8, 2)
    S = [dev(torch, np.stack([h0, np.zeros(C), np.zeros(C)])), dev(torch, np.zeros((3, C)))]
    tZ, zg = dev(torch, Z), dev(torch, zg_ref)
    grad = dev(torch, np.zeros((9, C)))
    dt0 = dev(torch, np.zeros(1))
    ring = dev(torch, np.zeros(4))
    call('mfv_time_step_sw', m.ptr, P(S[0][0]), P(S[0][1]), P(S[0][2]), 0.8, G, 0., 0., P(dt0), None)
    call('mfv_dt_ring_init', P(ring), P(dt0))
    for s in range(60):
        a, b = S[s % 2], S[(s + 1) % 2]
        call('mfv_sw_gradients', m.ptr, P(a[0]), P(a[1]), P(a[2]), P(tZ), P(grad))
        call('mfv_sw_step_fused', m.ptr, P(a), P(b), P(tZ), P(grad), P(zg), G, 0.8, 2, P(ring), s)
    assert host(ring)[59 % 3] == dt_last
    out = host(S[0])
    same(out[0], h, "h")
    same(out[1], hu, "hu")
    same(out[2], hv, "hv")
    assert np.abs(hu).max() > 0.1


# ------------------------------------------------------------------------------- Euler 3D
def sod(d, gamma=1.4):
    C = int(d.nbcells)
    left = d.cells.center[:, 0] < 0.5
    U = np.zeros((6, C))
    U[0] = np.where(left, 1.0, 0.125)
    U[1] = np.where(left, 1.0, 0.1)
    U[5] = 0.5 * (U[2] ** 2 + U[3] ** 2 + U[4] ** 2) / U[0] + U[1] / (gamma - 1)
    return U


@pytest.mark.parametrize("quirk", [0, 1])
def test_euler_kernels(torch, quirk, d3, m3, oracle_lib):
    d, m = d3, m3
    C, F = int(d.nbcells), int(d.nbfaces)
    rng = np.random.default_rng(12)
    U = sod(d)
    U[2:5] = 0.2 * rng.standard_normal((3, C))
    U[5] = 0.5 * (U[2] ** 2 + U[3] ** 2 + U[4] ** 2) / U[0] + U[1] / 0.4
    G_ref = np.zeros((6, F))
    oracle_lib.euler_ghost(d, U, G_ref, quirk)
    tU, tG = dev(torch, U), dev(torch, np.zeros((6, F)))
    call('mfv_euler3d_ghost', m.ptr, P(tU), P(tG), quirk)
    same(host(tG), G_ref, "ghost")
    dtc = oracle_lib.euler_dt(d, U, 0.8, 1.4)
    dtd = dev(torch, np.zeros(1))
    dth = ctypes.c_double()
    call('mfv_euler3d_time_step', m.ptr, P(tU), 0.8, 1.4, P(dtd), ctypes.byref(dth))
    assert dth.value == min(1e6, dtc.min())
    rez_ref = oracle_lib.euler_rusanov(d, U, G_ref, 1.4, quirk)
    rez = dev(torch, np.zeros((5, C)))
    call('mfv_euler3d_rusanov', m.ptr, P(tU), P(tG), 1.4, quirk, P(rez))
    same(host(rez), rez_ref, "rez")
    U_ref = U.copy()
    oracle_lib.euler_update(d, U_ref, rez_ref, 1.4, dth.value)
    call('mfv_euler3d_update', m.ptr, P(tU), P(rez), 1.4, P(dtd))
    same(host(tU), U_ref, "update")


@pytest.mark.parametrize("quirk", [0, 1])
def test_euler_fused_40_steps(torch, quirk, d3, m3, oracle_lib):
    d, m = d3, m3
    C, F = int(d.nbcells), int(d.nbfaces)
    U0 = sod(d)
    U = U0.copy()
    G_ = np.zeros((6, F))
    for _ in range(40):
        oracle_lib.euler_step(d, U, G_, 0.8, 1.4, quirk)
    S = [dev(torch, U0), dev(torch, np.zeros((6, C)))]
    dt0 = dev(torch, np.zeros(1))
    ring = dev(torch, np.zeros(4))
    call('mfv_euler3d_time_step', m.ptr, P(S[0]), 0.8, 1.4, P(dt0), None)
    call('mfv_dt_ring_init', P(ring), P(dt0))
    for s in range(40):
        call('mfv_euler3d_step_fused', m.ptr, P(S[s % 2]), P(S[(s + 1) % 2]), 1.4, 0.8, quirk, P(ring), s)
    same(host(S[0]), U, "U after 40 steps")


@pytest.mark.parametrize("quirk,T,n", [(0, 128, 6), (1, 256, 6), (1, 128, 11), (0, 256, 11)])
def test_euler_tiled_40_steps(torch, quirk, T, n, d3, m3, oracle_lib):
    """The face-tiled Euler step (each face evaluated once per tile in LDS) gives the oracle's
    bits after 40 steps, and those of the per-cell fused step; n = 11 has a partial last tile."""
    from manapy_amd.device import DeviceMesh, FaceTiles
    d, m = (d3, m3) if n == 6 else (make_tet_domain(n, 0.08, 4), None)
    m = m if m is not None else DeviceMesh(d, lsq=False)
    C, F = int(d.nbcells), int(d.nbfaces)
    U0 = sod(d)
    U = U0.copy()
    G_ = np.zeros((6, F))
    for _ in range(40):
        oracle_lib.euler_step(d, U, G_, 0.8, 1.4, quirk)
    ft = FaceTiles(m, T)
    assert ft.lds_bytes() <= 64 * 1024
    res = []
    for tiled in (True, False):
        S = [dev(torch, U0), dev(torch, np.zeros((6, C)))]
        dt0 = dev(torch, np.zeros(1))
        ring = dev(torch, np.zeros(4))
        call('mfv_euler3d_time_step', m.ptr, P(S[0]), 0.8, 1.4, P(dt0), None)
        call('mfv_dt_ring_init', P(ring), P(dt0))
        for s in range(40):
            if tiled:
                call('mfv_euler3d_step_tiled', m.ptr, ft.ptr, P(S[s % 2]), P(S[(s + 1) % 2]), 1.4, 0.8, quirk, P(ring),
                     s)
            else:
                call('mfv_euler3d_step_fused', m.ptr, P(S[s % 2]), P(S[(s + 1) % 2]), 1.4, 0.8, quirk, P(ring), s)
        res.append((host(S[0]), host(ring)))
    same(res[0][0], U, "tiled U after 40 steps")
    same(res[0][0], res[1][0], "tiled = fused")
    same(res[0][1], res[1][1], "dt ring")


@pytest.mark.parametrize("quirk,n,renumber", [(0, 6, False), (1, 6, False), (1, 11, True), (0, 7, True)])
def test_euler_packed_40_steps(torch, quirk, n, renumber, d3, m3, oracle_lib):
    """The packed Euler step (64-byte cell records, faces renumbered by first use) gives the
    oracle's bits after 40 steps, and the per-cell fused step's dt ring; pack then unpack is
    the identity and the records carry the cell volume."""
    from manapy_amd.device import DeviceMesh, EulerTet
    if n == 6 and not renumber:
        d, m = d3, m3
    else:
        d = make_tet_domain(n, 0.08, 5)
        if renumber:
            from manapy_amd.renumber import renumbered
            d = renumbered(d)
        m = DeviceMesh(d, lsq=False)
    C, F = int(d.nbcells), int(d.nbfaces)
    U0 = sod(d)
    U = U0.copy()
    G_ = np.zeros((6, F))
    for _ in range(40):
        oracle_lib.euler_step(d, U, G_, 0.8, 1.4, quirk)
    et = EulerTet(m)
    tU = dev(torch, U0)
    R = [torch.full((C, 8), np.nan, dtype=torch.float64, device='cuda') for _ in range(2)]
    call('mfv_euler3d_pack', m.ptr, P(tU), P(R[0]))
    rec = host(R[0])
    same(rec[:, :6].T, U0, "packed state")
    same(rec[:, 6], d.cells.volume, "packed volume")
    assert np.all(rec[:, 7] == 0.)
    back = dev(torch, np.zeros((6, C)))
    call('mfv_euler3d_unpack', m.ptr, P(R[0]), P(back))
    same(host(back), U0, "unpack(pack(U))")
    dt0 = dev(torch, np.zeros(1))
    rings = [dev(torch, np.zeros(4)) for _ in range(2)]
    for ring in rings:
        call('mfv_euler3d_time_step', m.ptr, P(tU), 0.8, 1.4, P(dt0), None)
        call('mfv_dt_ring_init', P(ring), P(dt0))
    S = [dev(torch, U0), dev(torch, np.zeros((6, C)))]
    for s in range(40):
        call('mfv_euler3d_step_packed', m.ptr, et.ptr, P(R[s % 2]), P(R[(s + 1) % 2]), 1.4, 0.8, quirk, P(rings[0]),
             s)
        call('mfv_euler3d_step_fused', m.ptr, P(S[s % 2]), P(S[(s + 1) % 2]), 1.4, 0.8, quirk, P(rings[1]), s)
    call('mfv_euler3d_unpack', m.ptr, P(R[0]), P(back))
    same(host(back), U, "packed U after 40 steps")
    same(host(R[0])[:, 6], d.cells.volume, "volume carried")
    same(host(rings[0]), host(rings[1]), "dt ring")


@pytest.mark.parametrize("quirk,T,n", [(0, 128, 6), (1, 256, 6), (1, 128, 11), (0, 256, 11), (1, 512, 11),
                                        (0, 512, 6)])
def test_euler_packed_tiled_40_steps(torch, quirk, T, n, oracle_lib):
    """The packed Euler step on face tiles (each face once per tile in LDS) and its single-phase
    variant (each cell evaluates its own faces) give the oracle's bits after 40 steps and the
    per-cell packed step's dt ring; n = 11 has a partial last tile. Tiles of 512 cells run the
    single-phase step only (the face fluxes of such a tile do not fit the LDS)."""
    from manapy_amd.device import DeviceMesh, EulerTet, EulerTetTiles
    from manapy_amd.renumber import renumbered
    d = renumbered(make_tet_domain(n, 0.08, 6))
    m = DeviceMesh(d, lsq=False)
    C, F = int(d.nbcells), int(d.nbfaces)
    U0 = sod(d)
    U = U0.copy()
    G_ = np.zeros((6, F))
    for _ in range(40):
        oracle_lib.euler_step(d, U, G_, 0.8, 1.4, quirk)
    et, tt = EulerTet(m), EulerTetTiles(m, T)
    assert tt.lds_bytes(cells=(T == 512)) <= 64 * 1024
    tU = dev(torch, U0)
    res = {}
    kinds = ('cells', 'packed') if T == 512 else ('tiled', 'cells', 'packed')
    for kind in kinds:
        R = [torch.zeros((C, 8), dtype=torch.float64, device='cuda') for _ in range(2)]
        dt0 = dev(torch, np.zeros(1))
        ring = dev(torch, np.zeros(4))
        call('mfv_euler3d_time_step', m.ptr, P(tU), 0.8, 1.4, P(dt0), None)
        call('mfv_dt_ring_init', P(ring), P(dt0))
        call('mfv_euler3d_pack', m.ptr, P(tU), P(R[0]))
        for s in range(40):
            if kind == 'tiled':
                call('mfv_euler3d_step_packed_tiled', m.ptr, et.ptr, tt.ptr, P(R[s % 2]), P(R[(s + 1) % 2]), 1.4, 0.8,
                     quirk, P(ring), s)
            elif kind == 'cells':
                call('mfv_euler3d_step_cells_tiled', m.ptr, et.ptr, tt.ptr, P(R[s % 2]), P(R[(s + 1) % 2]), 1.4, 0.8,
                     quirk, P(ring), s)
            else:
                call('mfv_euler3d_step_packed', m.ptr, et.ptr, P(R[s % 2]), P(R[(s + 1) % 2]), 1.4, 0.8, quirk,
                     P(ring), s)
        back = dev(torch, np.zeros((6, C)))
        call('mfv_euler3d_unpack', m.ptr, P(R[0]), P(back))
        res[kind] = (host(back), host(ring))
    for kind in kinds:
        same(res[kind][0], U, "%s U after 40 steps" % kind)
        same(res[kind][1], res['packed'][1], "%s dt ring" % kind)


@pytest.mark.parametrize("packed", [False, True])
def test_euler_tiled_dt_ignores_stale_partials(torch, packed, oracle_lib):
    """The tiled Euler steps fold exactly the dt partials their waves write. A finer mesh's dt
    kernel first fills the stream's partials buffer with smaller values; the tiled step on a
    7986-cell mesh (125 waves, 3 slots short of a 256-thread grid) must still give the oracle's
    next dt and state."""
    from manapy_amd.device import DeviceMesh, EulerTet, EulerTetTiles, FaceTiles
    from manapy_amd.renumber import renumbered
    big = make_tet_domain(14, 0.08, 3)
    mb = DeviceMesh(big, lsq=False)
    Ub = dev(torch, sod(big))
    dtb = dev(torch, np.zeros(1))
    call('mfv_euler3d_time_step', mb.ptr, P(Ub), 0.8, 1.4, P(dtb), None)
    d = renumbered(make_tet_domain(11, 0.08, 6)) if packed else make_tet_domain(11, 0.08, 6)
    m = DeviceMesh(d, lsq=False)
    C, F = int(d.nbcells), int(d.nbfaces)
    assert (C + 63) // 64 == 125
    U0 = sod(d)
    U = U0.copy()
    G_ = np.zeros((6, F))
    dt_first = float(np.min(oracle_lib.euler_dt(d, U, 0.8, 1.4)))
    oracle_lib.euler_step(d, U, G_, 0.8, 1.4, 1)
    dt_next = float(np.min(oracle_lib.euler_dt(d, U, 0.8, 1.4)))
    assert float(host(dtb)[0]) < dt_next  # the stale slots would win the min
    ring = dev(torch, np.array([dt_first, 1e6, 1e6, 0.]))
    if packed:
        et, tt = EulerTet(m), EulerTetTiles(m, 256)
        R = [torch.zeros((C, 8), dtype=torch.float64, device='cuda') for _ in range(2)]
        tU = dev(torch, U0)
        call('mfv_euler3d_pack', m.ptr, P(tU), P(R[0]))
        call('mfv_euler3d_step_packed_tiled', m.ptr, et.ptr, tt.ptr, P(R[0]), P(R[1]), 1.4, 0.8, 1, P(ring), 0)
        out = dev(torch, np.zeros((6, C)))
        call('mfv_euler3d_unpack', m.ptr, P(R[1]), P(out))
    else:
        ft = FaceTiles(m, 256)
        S = [dev(torch, U0), dev(torch, np.zeros((6, C)))]
        call('mfv_euler3d_step_tiled', m.ptr, ft.ptr, P(S[0]), P(S[1]), 1.4, 0.8, 1, P(ring), 0)
        out = S[1]
    same(host(out), U, "state after one tiled step")
    assert float(host(ring)[1]) == dt_next


@pytest.mark.parametrize("quirk,T,n", [(0, 256, 8), (1, 128, 8), (0, 128, 11), (1, 256, 11), (1, 512, 11)])
def test_euler32_tiled_matches_model_and_fp64(torch, quirk, T, n, oracle_lib):
    """FP32-bulk Euler step on face tiles: the float32 records and the dt ring equal its CPU
    model (oracle/euler32.c) bit for bit after 200 steps; every field within 1e-5 relative L1 of
    the float64 oracle; the float64 conservation sums agree with the model's and track the
    float64 run's (relative drift printed)."""
    from manapy_amd.device import DeviceMesh, Euler32
    from manapy_amd.renumber import renumbered
    d = renumbered(make_tet_domain(n, 0.08, 6))
    m = DeviceMesh(d, lsq=False)
    C, F = int(d.nbcells), int(d.nbfaces)
    U0 = sod(d)
    U = U0.copy()
    G_ = np.zeros((6, F))
    model = oracle_lib.Euler32Model(d, U0, quirk)
    dt = float(np.min(oracle_lib.euler_dt(d, U0, 0.8, 1.4)))
    e = Euler32(m, T)
    assert e.lds_bytes() <= 64 * 1024
    tU = dev(torch, U0)
    e.pack(tU)
    same(host(e.rec[0]), model.state, "packed float32 records")
    ring = dev(torch, np.array([dt, 1e6, 1e6, 0.]))
    dt_m = dt
    vol = d.cells.volume
    for s in range(200):
        oracle_lib.euler_step(d, U, G_, 0.8, 1.4, quirk)
        dt_m = model.step(0.8, dt_m)
        e.step(1.4, 0.8, quirk, ring, s)
    same(host(e.rec[e.k % 2]), model.state, "float32 records after 200 steps")
    assert host(ring)[200 % 3] == dt_m
    cons = host(e.cons)
    assert np.allclose(cons, model.cons, rtol=1e-12, atol=1e-15)
    cons64 = np.array([np.sum(U[k] * vol) for k in (0, 2, 3, 4, 5)])
    print("FP32-bulk conservation sums relative to float64:", (cons - cons64) / np.abs(cons64))
    out = dev(torch, np.zeros((6, C)))
    e.unpack(out)
    got = host(out)
    for k in range(6):
        err = np.sum(np.abs(got[k] - U[k]) * vol) / np.sum(np.abs(U[k]) * vol)
        assert err <= 1e-5, (k, err)


def test_sw_packed_60_steps(torch, sw_case, oracle_lib):
    """The packed path (one 64-byte record per cell, MUSCL values per face slot) gives the
    same bits as the reference step sequence."""
    from manapy_amd.device import SwPacked
    d, m, Z, h0 = sw_case
    C = int(d.nbcells)
    G = 9.81
    h, hu, hv = h0.copy(), np.zeros(C), np.zeros(C)
    zg_ref = oracle_lib.z_gradient(d, Z)
    dts = [oracle_lib.sw_step(d, h, hu, hv, Z, zg_ref, G, 0.8, 2) for _ in range(60)]
    p = SwPacked(m)
    th, thu, thv, tZ, zg = (dev(torch, x) for x in (h0, np.zeros(C), np.zeros(C), Z, zg_ref))
    call('mfv_sw_pack', m.ptr, p.ptr, P(th), P(thu), P(thv), P(tZ), P(zg), 2)
    dt0 = dev(torch, np.zeros(1))
    ring = dev(torch, np.zeros(4))
    call('mfv_time_step_sw', m.ptr, P(th), P(thu), P(thv), 0.8, G, 0., 0., P(dt0), None)
    call('mfv_dt_ring_init', P(ring), P(dt0))
    for s in range(60):
        call('mfv_sw_recon', m.ptr, p.ptr, 2)
        call('mfv_sw_flux_packed', m.ptr, p.ptr, G, 0.8, P(ring), s)
    oh, ohu, ohv = (dev(torch, np.zeros(C)) for _ in range(3))
    call('mfv_sw_unpack', m.ptr, p.ptr, P(oh), P(ohu), P(ohv))
    same(host(oh), h, "h")
    same(host(ohu), hu, "hu")
    same(host(ohv), hv, "hv")
    assert host(ring)[59 % 3] == dts[-1]
    # eta stored in the record is h + Z of the new state
    rec = host(p.cellrec)
    same(rec[:, 2], h + Z, "eta")


def test_sw_packed_first_order_and_quads(torch, oracle_lib):
    """Order 1 on a mixed triangle/quad mesh (max 4 faces per cell)."""
    from manapy_amd.ddm import Domain
    from manapy_amd.device import DeviceMesh, SwPacked
    from manapy_amd.meshgen import jittered_triangles
    el, nd = jittered_triangles(20, 14, 100., 60., 0.0, 4)
    # merge every other triangle pair of the first rows into a quad (a, b, c, d)
    quads = []
    keep = np.ones(len(el), dtype=bool)
    for q in range(0, 40, 2):
        t1, t2 = el[2 * q], el[2 * q + 1]
        quads.append([t1[0], t1[1], t1[2], t2[2], 4])
        keep[2 * q] = keep[2 * q + 1] = False
    el4 = np.concatenate([np.concatenate([el[keep][:, :3], -np.ones((keep.sum(), 1), np.int32),
                                          el[keep][:, 3:]], axis=1), np.array(quads, dtype=np.int32)])
    d = Domain(2, mesh=(el4, nd))
    assert d.cells.faceid.shape[1] == 5
    C = int(d.nbcells)
    c = d.cells.center
    Z = 0.1 * np.sin(c[:, 0] / 10.)
    h0 = np.where(c[:, 0] < 50., 2.0, 1.0)
    h, hu, hv = h0.copy(), np.zeros(C), np.zeros(C)
    zg_ref = oracle_lib.z_gradient(d, Z)
    for _ in range(30):
        oracle_lib.sw_step(d, h, hu, hv, Z, zg_ref, 9.81, 0.8, 1)
    m = DeviceMesh(d)
    p = SwPacked(m)
    th, thu, thv, tZ, zg = (dev(torch, x) for x in (h0, np.zeros(C), np.zeros(C), Z, zg_ref))
    call('mfv_sw_pack', m.ptr, p.ptr, P(th), P(thu), P(thv), P(tZ), P(zg), 1)
    dt0 = dev(torch, np.zeros(1))
    ring = dev(torch, np.zeros(4))
    call('mfv_time_step_sw', m.ptr, P(th), P(thu), P(thv), 0.8, 9.81, 0., 0., P(dt0), None)
    call('mfv_dt_ring_init', P(ring), P(dt0))
    for s in range(30):
        call('mfv_sw_recon', m.ptr, p.ptr, 1)
        call('mfv_sw_flux_packed', m.ptr, p.ptr, 9.81, 0.8, P(ring), s)
    oh, ohu, ohv = (dev(torch, np.zeros(C)) for _ in range(3))
    call('mfv_sw_unpack', m.ptr, p.ptr, P(oh), P(ohu), P(ohv))
    same(host(oh), h, "h")
    same(host(ohu), hu, "hu")
    same(host(ohv), hv, "hv")


def test_sw_mixed_precision_error_and_conservation(torch, oracle_lib):
    """Mixed-precision path: bit-identical to its CPU float32 model (oracle/mixed32.c) after
    5, 25 and 200 steps of the dam break over bumps; relative L1 error (Variable.norml2, order
    1) of every field against the float64 oracle <= 1e-5 at each of them (the early steps,
    where the tangential momentum is still small, are the hard case: <= 1e-6 is asserted to
    keep the well-balanced deviation form of the bathymetry term); the float64 mass total
    conserved."""
    from manapy_amd.device import DeviceMesh, SwMixed
    from manapy_amd.meshgen import gaussian_bumps
    d = make_tri_domain(60, 60, 100., 100., 0.1, 13)
    C = int(d.nbcells)
    c = d.cells.center
    Z = gaussian_bumps(c, 8, np.random.default_rng(5), [(0., 100.), (0., 100.)])
    h0 = np.where(c[:, 0] < 50., 2.0, 1.0)
    G = 9.81
    checks = (5, 25, 200)
    h, hu, hv = h0.copy(), np.zeros(C), np.zeros(C)
    zg_ref = oracle_lib.z_gradient(d, Z)
    model = oracle_lib.MixedModel(d, h0, np.zeros(C), np.zeros(C), Z, zg_ref)
    dt_m = oracle_lib.time_step_sw(d, h0, np.zeros(C), np.zeros(C), 0.8, G)
    ref, mod, dts = {}, {}, {}
    for k in range(1, checks[-1] + 1):
        oracle_lib.sw_step(d, h, hu, hv, Z, zg_ref, G, 0.8, 2)
        dt_m = model.step(G, 0.8, dt_m)
        if k in checks:
            ref[k] = (h.copy(), hu.copy(), hv.copy())
            mod[k] = model.state.copy()
            dts[k] = dt_m
    m = DeviceMesh(d)
    p = SwMixed(m)
    th, thu, thv, tZ, zg = (dev(torch, x) for x in (h0, np.zeros(C), np.zeros(C), Z, zg_ref))
    call('mfv_sw_pack_mixed', m.ptr, p.ptr, P(th), P(thu), P(thv), P(tZ), P(zg), 2)
    assert np.array_equal(host(p.geoA), model.geoA) and np.array_equal(host(p.geoB), model.geoB)
    call('mfv_sw_conservation', m.ptr, p.ptr, P(p.partial), p.partial.numel(), P(p.sums))
    mass0 = host(p.sums)[0]
    dt0 = dev(torch, np.zeros(1))
    ring = dev(torch, np.zeros(4))
    call('mfv_time_step_sw', m.ptr, P(th), P(thu), P(thv), 0.8, G, 0., 0., P(dt0), None)
    call('mfv_dt_ring_init', P(ring), P(dt0))
    vol = d.cells.volume
    oh, ohu, ohv = (dev(torch, np.zeros(C)) for _ in range(3))
    for s in range(checks[-1]):
        call('mfv_sw_recon_mixed', m.ptr, p.ptr, 2)
        call('mfv_sw_flux_mixed', m.ptr, p.ptr, G, 0.8, P(ring), s)
        k = s + 1
        if k not in checks:
            continue
        same(host(p.state), mod[k], "mixed state vs its float32 model after %d steps" % k)
        assert host(ring)[k % 3] == dts[k]
        call('mfv_sw_unpack_mixed', m.ptr, p.ptr, P(oh), P(ohu), P(ohv))
        errs = {}
        for nm, a, b in (('h', host(oh), ref[k][0]), ('hu', host(ohu), ref[k][1]), ('hv', host(ohv), ref[k][2])):
            errs[nm] = float(np.sum(np.abs(a - b) * vol) / np.sum(np.abs(b) * vol))
        print("mixed-precision relative L1 after %d steps:" % k, errs)
        assert max(errs.values()) <= 1e-5, (k, errs)
        assert max(errs.values()) <= 1e-6, (k, errs)
    call('mfv_sw_conservation', m.ptr, p.ptr, P(p.partial), p.partial.numel(), P(p.sums))
    assert abs(host(p.sums)[0] - mass0) <= 1e-12 * mass0
    assert abs(mass0 - np.sum(h0 * vol)) <= 1e-12 * mass0


def test_sw_packed_and_mixed_on_hilbert_order(torch, sw_case, oracle_lib):
    """The packed float64 path and the mixed path on the Hilbert-renumbered device mesh give
    the original-numbering results permuted, bit for bit."""
    from manapy_amd.device import DeviceMesh, SwPacked, SwMixed
    from manapy_amd.renumber import renumbered
    d, _, Z, h0 = sw_case
    r = renumbered(d)
    rm = DeviceMesh(r)
    C = int(d.nbcells)
    G = 9.81
    h, hu, hv = h0.copy(), np.zeros(C), np.zeros(C)
    zg_ref = oracle_lib.z_gradient(d, Z)
    model = oracle_lib.MixedModel(d, h0, np.zeros(C), np.zeros(C), Z, zg_ref)
    dts = [oracle_lib.sw_step(d, h, hu, hv, Z, zg_ref, G, 0.8, 2) for _ in range(30)]
    th, tz, tZ = dev(torch, r.to_new(h0)), dev(torch, np.zeros(C)), dev(torch, r.to_new(Z))
    zg = dev(torch, np.ascontiguousarray(zg_ref[:, r.perm]))
    p = SwPacked(rm)
    call('mfv_sw_pack', rm.ptr, p.ptr, P(th), P(tz), P(tz), P(tZ), P(zg), 2)
    dt0 = dev(torch, np.zeros(1))
    ring = dev(torch, np.zeros(4))
    call('mfv_time_step_sw', rm.ptr, P(th), P(tz), P(tz), 0.8, G, 0., 0., P(dt0), None)
    call('mfv_dt_ring_init', P(ring), P(dt0))
    for s in range(30):
        call('mfv_sw_recon', rm.ptr, p.ptr, 2)
        call('mfv_sw_flux_packed', rm.ptr, p.ptr, G, 0.8, P(ring), s)
    oh, ohu, ohv = (dev(torch, np.zeros(C)) for _ in range(3))
    call('mfv_sw_unpack', rm.ptr, p.ptr, P(oh), P(ohu), P(ohv))
    same(r.to_old(host(oh)), h, "h")
    same(r.to_old(host(ohu)), hu, "hu")
    same(r.to_old(host(ohv)), hv, "hv")
    assert host(ring)[29 % 3] == dts[-1]
    # mixed path
    q = SwMixed(rm)
    zg3 = dev(torch, np.vstack([zg_ref[:, r.perm], np.zeros((1, C))]))
    call('mfv_sw_pack_mixed', rm.ptr, q.ptr, P(th), P(tz), P(tz), P(tZ), P(zg3), 2)
    dt = model.dt(G, 0.8)
    ring2 = dev(torch, np.zeros(4))
    dt0[0] = dt
    call('mfv_dt_ring_init', P(ring2), P(dt0))
    for s in range(30):
        dt = model.step(G, 0.8, dt)
        call('mfv_sw_recon_mixed', rm.ptr, q.ptr, 2)
        call('mfv_sw_flux_mixed', rm.ptr, q.ptr, G, 0.8, P(ring2), s)
    st = host(q.state)
    for k in range(3):
        same(r.to_old(st[:, k]), model.state[:, k], "mixed %d" % k)


@pytest.mark.parametrize("T,N,hilbert,stage", [(256, 256, True, False), (512, 512, True, False),
                                               (256, 256, False, False), (128, 128, True, True),
                                               (256, 256, True, True), (128, 128, False, True),
                                               (256, 192, True, True), (128, 64, True, True)])
def test_sw_mixed_tiled_step_equals_two_kernel_step(torch, oracle_lib, T, N, hilbert, stage):
    """The fused tiled step (face records in LDS, double-buffered state) gives the bits of
    reconstruction + flux kernels, step after step, including the dt ring."""
    from manapy_amd.device import DeviceMesh, SwMixed
    from manapy_amd.meshgen import gaussian_bumps
    from manapy_amd.renumber import renumbered
    d0 = make_tri_domain(45, 37, 100., 80., 0.1, 21)
    d = renumbered(d0) if hilbert else d0
    C = int(d.nbcells)
    c = d.cells.center
    Z = gaussian_bumps(c, 6, np.random.default_rng(3), [(0., 100.), (0., 80.)])
    h0 = np.where(c[:, 0] < 40., 2.0, 1.0)
    zg = oracle_lib.z_gradient(d, Z)
    m = DeviceMesh(d)
    runs = []
    for tiled in (False, True):
        p = SwMixed(m)
        th, tz, tZ, tg = (dev(torch, x) for x in (h0, np.zeros(C), Z, zg))
        call('mfv_sw_pack_mixed', m.ptr, p.ptr, P(th), P(tz), P(tz), P(tZ), P(tg), 2)
        dt0 = dev(torch, np.zeros(1))
        ring = dev(torch, np.zeros(4))
        call('mfv_time_step_sw', m.ptr, P(th), P(tz), P(tz), 0.8, 9.81, 0., 0., P(dt0), None)
        call('mfv_dt_ring_init', P(ring), P(dt0))
        rings = []
        for s in range(25):
            if tiled:
                p.step_tiled(m, 9.81, 0.8, 2, ring, s, T=T, stage=stage, N=N)
            else:
                call('mfv_sw_recon_mixed', m.ptr, p.ptr, 2)
                call('mfv_sw_flux_mixed', m.ptr, p.ptr, 9.81, 0.8, P(ring), s)
            rings.append(host(ring).copy())
        runs.append((host(p.state), rings))
    same(runs[1][0], runs[0][0], "tiled state")
    for a, b in zip(runs[0][1], runs[1][1]):
        same(b, a, "dt ring")


@pytest.mark.parametrize("T,hilbert", [(128, True), (256, True), (128, False)])
def test_sw_packed_tiled_step_equals_oracle(torch, oracle_lib, T, hilbert):
    """The fused float64 step on tiles (records and face values in LDS) gives the reference
    sequence's bits, like the two-kernel packed path."""
    from manapy_amd.device import DeviceMesh, SwPacked
    from manapy_amd.meshgen import gaussian_bumps
    from manapy_amd.renumber import renumbered
    d0 = make_tri_domain(43, 35, 100., 80., 0.1, 17)
    d = renumbered(d0) if hilbert else d0
    C = int(d.nbcells)
    c = d.cells.center
    Z = gaussian_bumps(c, 6, np.random.default_rng(4), [(0., 100.), (0., 80.)])
    h0 = np.where(c[:, 0] < 40., 2.0, 1.0)
    G = 9.81
    h, hu, hv = h0.copy(), np.zeros(C), np.zeros(C)
    zg = oracle_lib.z_gradient(d, Z)
    dts = [oracle_lib.sw_step(d, h, hu, hv, Z, zg, G, 0.8, 2) for _ in range(25)]
    m = DeviceMesh(d)
    p = SwPacked(m)
    th, tz, tZ, tg = (dev(torch, x) for x in (h0, np.zeros(C), Z, zg))
    p.pack(m, th, tz, tz, tZ, tg, 2)
    dt0 = dev(torch, np.zeros(1))
    ring = dev(torch, np.zeros(4))
    call('mfv_time_step_sw', m.ptr, P(th), P(tz), P(tz), 0.8, G, 0., 0., P(dt0), None)
    call('mfv_dt_ring_init', P(ring), P(dt0))
    for s in range(25):
        p.step_tiled(m, G, 0.8, 2, ring, s, T=T)
    oh, ohu, ohv = (dev(torch, np.zeros(C)) for _ in range(3))
    call('mfv_sw_unpack', m.ptr, p.ptr, P(oh), P(ohu), P(ohv))
    same(host(oh), h, "h")
    same(host(ohu), hu, "hu")
    same(host(ohv), hv, "hv")
    assert host(ring)[24 % 3] == dts[-1]
