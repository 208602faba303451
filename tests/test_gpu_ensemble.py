"""Shallow-water ensembles (many meshes in one launch per stage): every instance gives the
same bits alone and in the mix, graph replays equal eager steps, each instance equals the
CPU float32 model of the mixed kernels bit for bit and stays within 1e-5 relative L1 of the
float64 oracle, and mass is conserved."""
import numpy as np
import pytest

from conftest import gpu_available, make_tri_domain

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]


def _instances():
    from manapy_amd.ensemble import instance_state
    specs = [(23, 11, 100., 48., 1), (9, 30, 40., 100., 2), (31, 17, 100., 55., 3), (14, 14, 60., 60., 4)]
    doms, states = [], []
    for nx, ny, lx, ly, seed in specs:
        d = make_tri_domain(nx, ny, lx, ly, 0.1, 500 + seed)
        doms.append(d)
        states.append(instance_state(d, lx, ly, np.random.default_rng(seed)))
    return doms, states


@pytest.mark.parametrize("config", [("tiled", 256), ("split", 128), ("tiled", 512), ("staged", 128),
                                    ("staged", 256, 192), ("staged", 256)])
def test_alone_equals_mix_and_graph_equals_eager(config):
    """Fused tiled and two-kernel steps, eager and in graphs, alone and mixed: all the same bits."""
    from manapy_amd.ensemble import SwEnsemble
    doms, states = _instances()
    alone = []
    for d, s in zip(doms, states):
        e = SwEnsemble([d], [s])
        e.step(36)
        alone.append((e.instance_state(0), e.times()[0]))
    order = [2, 0, 3, 1]
    mix = SwEnsemble([doms[k] for k in order], [states[k] for k in order])
    mix.config = config
    eager = SwEnsemble([doms[k] for k in order], [states[k] for k in order])
    # graph: two replays of 18 steps, eager steps in between on another ensemble
    for rep in range(2):
        mix.run_graph(18, chunk=18)
        eager.step(18, ("split", 64) if config[0] == "tiled" else ("tiled", 256))
        import torch
        torch.cuda.synchronize()
        for j in range(len(order)):
            a, b = mix.instance_state(j), eager.instance_state(j)
            for x, y in zip(a, b):
                assert np.array_equal(x, y), (rep, j)
        assert np.array_equal(mix.times(), eager.times())
    for j, k in enumerate(order):
        st, t = alone[k]
        for x, y in zip(mix.instance_state(j), st):
            assert np.array_equal(x, y), k
        assert mix.times()[j] == t
    # the Hilbert relabelling of the device cells changes no bit
    plain = SwEnsemble([doms[1]], [states[1]], renumber=False)
    plain.step(36, ("split", 256))
    for x, y in zip(plain.instance_state(0), alone[1][0]):
        assert np.array_equal(x, y)
    assert plain.times()[0] == alone[1][1]


def test_ensemble_error_vs_fp64_oracle_and_conservation(oracle_lib):
    from manapy_amd.ensemble import SwEnsemble
    doms, states = _instances()
    e = SwEnsemble(doms, states)
    m0 = e.conservation()[:, 0]
    e.step(60)
    m1 = e.conservation()[:, 0]
    assert np.all(np.abs(m1 - m0) <= 1e-12 * m0)
    for i, (d, (h0, hu0, hv0, Z)) in enumerate(zip(doms, states)):
        vol = d.cells.volume
        assert abs(m0[i] - np.sum(h0 * vol)) <= 1e-12 * m0[i]
        h, hu, hv = h0.copy(), hu0.copy(), hv0.copy()
        zg = oracle_lib.z_gradient(d, Z)
        for _ in range(60):
            oracle_lib.sw_step(d, h, hu, hv, Z, zg, 9.81, 0.8, 2)
        a = e.instance_state(i)
        model = oracle_lib.MixedModel(d, h0, hu0, hv0, Z, zg)
        dt = model.dt(9.81, 0.8)
        for _ in range(60):
            dt = model.step(9.81, 0.8, dt)
        for k in range(3):
            assert np.array_equal(a[k], model.state[:, k]), (i, k)
        for nm, x, y in (('h', a[0], h), ('hu', a[1], hu), ('hv', a[2], hv)):
            err = np.sum(np.abs(x - y) * vol) / np.sum(np.abs(y) * vol)
            assert err <= 1e-5, (i, nm, err)


def test_autotune_restores_state(tmp_path):
    from manapy_amd.ensemble import SwEnsemble
    doms, states = _instances()
    e = SwEnsemble(doms, states)
    e.step(3)
    before = [e.instance_state(i) for i in range(e.n)]
    entry = e.autotune(steps=3, cache=str(tmp_path / "tune.json"))
    from manapy_amd.ensemble import CONFIGS
    assert tuple(entry['config']) in CONFIGS
    for i in range(e.n):
        for x, y in zip(e.instance_state(i), before[i]):
            assert np.array_equal(x, y)
    e2 = SwEnsemble(doms, states)
    assert e2.autotune(cache=str(tmp_path / "tune.json"))['config'] == entry['config']
