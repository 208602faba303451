"""Meshes with cells of different face counts (triangles and quadrilaterals, max 4 faces per
cell): the shallow-water kernels load every slot of their [MF]-wide tables whatever the cell's
face count, so the unused slots of triangles (code -1, zero geometry) must never reach the
arithmetic. Float64 paths against the oracle bit for bit; mixed paths against each other."""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]

G = 9.81


def tri_quad_domain(nx, ny, lx, ly, every=3, seed=4):
    """Jittered triangles with every ``every``-th triangle pair merged into a quadrilateral."""
    from manapy_amd.ddm import Domain
    from manapy_amd.meshgen import jittered_triangles
    el, nd = jittered_triangles(nx, ny, lx, ly, 0.05, seed)
    quads = []
    keep = np.ones(len(el), dtype=bool)
    for q in range(0, len(el) // 2, every):
        t1, t2 = el[2 * q], el[2 * q + 1]
        quads.append([t1[0], t1[1], t1[2], t2[2], 4])
        keep[2 * q] = keep[2 * q + 1] = False
    el4 = np.concatenate([np.concatenate([el[keep][:, :3], -np.ones((keep.sum(), 1), np.int32), el[keep][:, 3:]],
                                         axis=1), np.array(quads, dtype=np.int32)])
    return Domain(2, mesh=(el4, nd))


def _dev(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device='cuda')


def _case(hilbert):
    from manapy_amd.meshgen import gaussian_bumps
    from manapy_amd.renumber import renumbered
    d0 = tri_quad_domain(31, 23, 100., 70.)
    d = renumbered(d0) if hilbert else d0
    assert d.cells.faceid.shape[1] == 5 and set(np.unique(d.cells.faceid[:, -1])) == {3, 4}
    c = d.cells.center
    Z = gaussian_bumps(c, 5, np.random.default_rng(2), [(0., 100.), (0., 70.)])
    h0 = np.where(c[:, 0] < 45., 2.0, 1.0)
    return d, Z, h0


@pytest.mark.parametrize("order,path,hilbert", [(2, 'packed', True), (2, 'fused128', True), (2, 'fused256', True),
                                                (1, 'fused256', True), (2, 'fused128', False)])
def test_float64_paths_on_triangles_and_quads(oracle_lib, order, path, hilbert):
    import torch
    from manapy_amd import _lib
    from manapy_amd.device import DeviceMesh, SwPacked, stream_ptr
    d, Z, h0 = _case(hilbert)
    C = int(d.nbcells)
    h, hu, hv = h0.copy(), np.zeros(C), np.zeros(C)
    zg = oracle_lib.z_gradient(d, Z)
    dts = [oracle_lib.sw_step(d, h, hu, hv, Z, zg, G, 0.8, order) for _ in range(20)]
    m = DeviceMesh(d)
    p = SwPacked(m)
    s = stream_ptr()
    th, tz, tZ, tg = (_dev(x) for x in (h0, np.zeros(C), Z, zg))
    p.pack(m, th, tz, tz, tZ, tg, order)
    dt0, ring = _dev(np.zeros(1)), _dev(np.zeros(4))
    _lib.call('mfv_time_step_sw', m.ptr, th.data_ptr(), tz.data_ptr(), tz.data_ptr(), 0.8, G, 0., 0.,
              dt0.data_ptr(), None, s)
    _lib.call('mfv_dt_ring_init', ring.data_ptr(), dt0.data_ptr(), s)
    for k in range(20):
        if path == 'packed':
            _lib.call('mfv_sw_recon', m.ptr, p.ptr, order, s)
            _lib.call('mfv_sw_flux_packed', m.ptr, p.ptr, G, 0.8, ring.data_ptr(), k, s)
        else:
            p.step_tiled(m, G, 0.8, order, ring, k, T=int(path[5:]), stream=s)
    out = [torch.zeros(C, dtype=torch.float64, device='cuda') for _ in range(3)]
    _lib.call('mfv_sw_unpack', m.ptr, p.ptr, *[o.data_ptr() for o in out], s)
    for o, ref, name in zip(out, (h, hu, hv), ('h', 'hu', 'hv')):
        assert np.array_equal(o.cpu().numpy(), ref), name
    assert float(ring[19 % 3]) == dts[-1]


@pytest.mark.parametrize("T,stage", [(128, True), (256, True), (256, False)])
def test_mixed_paths_on_triangles_and_quads(oracle_lib, T, stage):
    """Fused tiled mixed step = two-kernel mixed step (bits and dt ring) = the CPU float32
    model, on a triangle/quadrilateral mesh."""
    from manapy_amd import _lib
    from manapy_amd.device import DeviceMesh, SwMixed, stream_ptr
    d, Z, h0 = _case(True)
    C = int(d.nbcells)
    zg = oracle_lib.z_gradient(d, Z)
    m = DeviceMesh(d)
    s = stream_ptr()
    runs = []
    for tiled in (False, True):
        p = SwMixed(m)
        th, tz, tZ, tg = (_dev(x) for x in (h0, np.zeros(C), Z, zg))
        _lib.call('mfv_sw_pack_mixed', m.ptr, p.ptr, th.data_ptr(), tz.data_ptr(), tz.data_ptr(), tZ.data_ptr(),
                  tg.data_ptr(), 2, s)
        dt0, ring = _dev(np.zeros(1)), _dev(np.zeros(4))
        _lib.call('mfv_time_step_sw', m.ptr, th.data_ptr(), tz.data_ptr(), tz.data_ptr(), 0.8, G, 0., 0.,
                  dt0.data_ptr(), None, s)
        _lib.call('mfv_dt_ring_init', ring.data_ptr(), dt0.data_ptr(), s)
        rings = []
        for k in range(20):
            if tiled:
                p.step_tiled(m, G, 0.8, 2, ring, k, T=T, stage=stage, stream=s)
            else:
                _lib.call('mfv_sw_recon_mixed', m.ptr, p.ptr, 2, s)
                _lib.call('mfv_sw_flux_mixed', m.ptr, p.ptr, G, 0.8, ring.data_ptr(), k, s)
            rings.append(ring.cpu().numpy().copy())
        runs.append((p.state.cpu().numpy(), rings))
    assert np.array_equal(runs[0][0], runs[1][0])
    for a, b in zip(runs[0][1], runs[1][1]):
        assert np.array_equal(a, b)
    model = oracle_lib.MixedModel(d, h0, np.zeros(C), np.zeros(C), Z, zg)
    dt = oracle_lib.time_step_sw(d, h0, np.zeros(C), np.zeros(C), 0.8, G)  # the ring's first dt (float64)
    for _ in range(20):
        dt = model.step(G, 0.8, dt)
    for k in range(3):
        assert np.array_equal(runs[0][0][:, k], model.state[:, k]), k
