"""GPU parity at sizes where the tile and halo layouts of the fast paths take their general
shape (hundreds of tiles, halos and ext lists near their maxima, a partial last tile, the
LDS budget of the larger tiles): the float64 fused SW step and the mixed step on 180k
triangles, the float64 and FP32-bulk Euler steps on 83k tetrahedra, each against the oracle
(or its float32 model) after 10 steps, bit for bit."""
import numpy as np
import pytest

from conftest import gpu_available, make_tri_domain, make_tet_domain

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]


def dev(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device='cuda')


@pytest.fixture(scope="module")
def sw_big():
    from manapy_amd.meshgen import gaussian_bumps
    from manapy_amd.renumber import renumbered
    d = renumbered(make_tri_domain(300, 300, 100., 100., 0.1, 17))
    c = d.cells.center
    Z = gaussian_bumps(c, 8, np.random.default_rng(4), [(0., 100.), (0., 100.)])
    h0 = np.where(c[:, 0] < 50., 2.0, 1.0)
    return d, Z, h0


@pytest.mark.parametrize("T", [128, 256])
def test_sw_fused_float64_180k(sw_big, T, oracle_lib):
    from manapy_amd import _lib
    from manapy_amd.device import DeviceMesh, SwPacked, stream_ptr
    d, Z, h0 = sw_big
    C = int(d.nbcells)
    zg_ref = oracle_lib.z_gradient(d, Z)
    h, hu, hv = h0.copy(), np.zeros(C), np.zeros(C)
    for _ in range(10):
        dt_last = oracle_lib.sw_step(d, h, hu, hv, Z, zg_ref, 9.81, 0.8, 2)
    m = DeviceMesh(d)
    s = stream_ptr()
    th, tz, tZ, zg = dev(h0), dev(np.zeros(C)), dev(Z), dev(zg_ref)
    pk = SwPacked(m)
    assert pk.lds_bytes(T) <= 64 * 1024
    pk.pack(m, th, tz, tz, tZ, zg, 2, s)
    dt0, ring = dev(np.zeros(1)), dev(np.zeros(4))
    _lib.call('mfv_time_step_sw', m.ptr, th.data_ptr(), tz.data_ptr(), tz.data_ptr(), 0.8, 9.81, 0., 0.,
              dt0.data_ptr(), None, s)
    _lib.call('mfv_dt_ring_init', ring.data_ptr(), dt0.data_ptr(), s)
    for k in range(10):
        pk.step_tiled(m, 9.81, 0.8, 2, ring, k, T=T, stream=s)
    o = [dev(np.zeros(C)) for _ in range(3)]
    _lib.call('mfv_sw_unpack', m.ptr, pk.ptr, o[0].data_ptr(), o[1].data_ptr(), o[2].data_ptr(), s)
    for a, b in zip(o, (h, hu, hv)):
        assert np.array_equal(a.cpu().numpy(), b)
    assert ring.cpu().numpy()[9 % 3] == dt_last


@pytest.mark.parametrize("T,stage", [(256, True), (512, True), (256, False)])
def test_sw_mixed_tiled_180k(sw_big, T, stage, oracle_lib):
    from manapy_amd import _lib
    from manapy_amd.device import DeviceMesh, SwMixed, stream_ptr
    d, Z, h0 = sw_big
    C = int(d.nbcells)
    zg_ref = oracle_lib.z_gradient(d, Z)
    model = oracle_lib.MixedModel(d, h0, np.zeros(C), np.zeros(C), Z, zg_ref)
    dt_m = oracle_lib.time_step_sw(d, h0, np.zeros(C), np.zeros(C), 0.8, 9.81)
    dt0 = dt_m
    for _ in range(10):
        dt_m = model.step(9.81, 0.8, dt_m)
    m = DeviceMesh(d)
    s = stream_ptr()
    p = SwMixed(m)
    th, tz, tZ, zg = dev(h0), dev(np.zeros(C)), dev(Z), dev(zg_ref)
    _lib.call('mfv_sw_pack_mixed', m.ptr, p.ptr, th.data_ptr(), tz.data_ptr(), tz.data_ptr(), tZ.data_ptr(),
              zg.data_ptr(), 2, s)
    ring = dev(np.array([dt0, 1e6, 1e6, 0.]))
    tiles = p.tiles(T, stage)
    if tiles.lds_bytes() > 64 * 1024:
        pytest.skip("tiles of %d need %d B of LDS" % (T, tiles.lds_bytes()))
    for k in range(10):
        p.step_tiled(m, 9.81, 0.8, 2, ring, k, T=T, stage=stage, stream=s)
    assert np.array_equal(p.state.cpu().numpy(), model.state)
    assert ring.cpu().numpy()[10 % 3] == dt_m


@pytest.mark.parametrize("T", [128, 256])
def test_euler_float64_and_fp32_bulk_83k(T, oracle_lib):
    import torch
    from manapy_amd import _lib
    from manapy_amd.device import DeviceMesh, EulerTet, EulerTetTiles, Euler32, stream_ptr
    from manapy_amd.renumber import renumbered
    from manapy_amd.solvers.euler import initialisation_euler_3d
    d = renumbered(make_tet_domain(24, 0.08, 12))
    C, F = int(d.nbcells), int(d.nbfaces)
    U0 = initialisation_euler_3d(d.cells.center)
    U = U0.copy()
    G_ = np.zeros((6, F))
    model = oracle_lib.Euler32Model(d, U0, 0)
    dt = float(np.min(oracle_lib.euler_dt(d, U0, 0.8, 1.4)))
    dt_m = dt
    for _ in range(10):
        dt_last = oracle_lib.euler_step(d, U, G_, 0.8, 1.4, 0)
        dt_m = model.step(0.8, dt_m)
    m = DeviceMesh(d, lsq=False)
    s = stream_ptr()
    et, tt = EulerTet(m), EulerTetTiles(m, T)
    assert tt.lds_bytes() <= 64 * 1024
    tU = dev(U0)
    R = [torch.zeros((C, 8), dtype=torch.float64, device='cuda') for _ in range(2)]
    _lib.call('mfv_euler3d_pack', m.ptr, tU.data_ptr(), R[0].data_ptr(), s)
    ring = dev(np.array([dt, 1e6, 1e6, 0.]))
    for k in range(10):
        _lib.call('mfv_euler3d_step_packed_tiled', m.ptr, et.ptr, tt.ptr, R[k % 2].data_ptr(),
                  R[(k + 1) % 2].data_ptr(), 1.4, 0.8, 0, ring.data_ptr(), k, s)
    back = dev(np.zeros((6, C)))
    _lib.call('mfv_euler3d_unpack', m.ptr, R[0].data_ptr(), back.data_ptr(), s)
    assert np.array_equal(back.cpu().numpy(), U)
    assert ring.cpu().numpy()[9 % 3] == dt_last
    e = Euler32(m, et=et, tt=tt)
    e.pack(tU, s)
    ring = dev(np.array([dt, 1e6, 1e6, 0.]))
    for k in range(10):
        e.step(1.4, 0.8, 0, ring, k, stream=s)
    assert np.array_equal(e.rec[e.k % 2].cpu().numpy(), model.state)
    assert ring.cpu().numpy()[10 % 3] == dt_m
