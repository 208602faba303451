"""VTU output (ddm/domain.py:624-849 save_on_cell / save_on_node / *_multi): the files parse as
VTK XML, carry every cell with its vertices, offsets and VTK type, the points, and the fields
bit for bit; the .pvtu index names the piece."""
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from conftest import make_tri_domain, make_tet_domain, gpu_available


def _check_mesh(d, arr):
    nodeid = d.cells.nodeid
    off = np.concatenate([[0], arr['offsets']])
    for i in range(int(d.nbcells)):
        n = nodeid[i, -1]
        assert np.array_equal(arr['connectivity'][off[i]:off[i + 1]], nodeid[i, :n])
    assert np.array_equal(arr['Points'], d.nodes.vertex[:, :3])


def test_cell_and_node_fields_round_trip(tmp_path, monkeypatch):
    from manapy_amd.vtu import read_vtu
    monkeypatch.chdir(tmp_path)
    d = make_tri_domain(9, 7, 1.0, 1.0, 0.1, 3)
    rng = np.random.default_rng(0)
    w = rng.standard_normal(int(d.nbcells))
    d.save_on_cell(0.1, 1.0, 5, 2, value=w)
    wn = rng.standard_normal(int(d.nbnodes))
    d.save_on_node_multi(0.1, 1.0, 5, 3, variables=['a', 'b'], values=[wn, 2 * wn])
    d.wait_outputs()
    arr = read_vtu(os.path.join('results', 'visu0-2.vtu'))
    assert np.array_equal(arr['w'], w)
    assert np.all(arr['types'] == 5)
    _check_mesh(d, arr)
    arr = read_vtu(os.path.join('results', 'visu0-3.vtu'))
    assert np.array_equal(arr['a'], wn) and np.array_equal(arr['b'], 2 * wn)
    root = ET.parse(os.path.join('results', 'visu3.pvtu')).getroot()
    assert [p.get('Source') for p in root.iter('Piece')] == ['visu0-3.vtu']
    assert [a.get('Name') for a in root.iter('PDataArray')][-2:] == ['a', 'b']
    with pytest.raises(ValueError):
        d.save_on_cell(value=np.zeros(3))


def test_tetrahedra_and_mixed_cells(tmp_path):
    from manapy_amd.vtu import mesh_arrays, write_vtu, read_vtu
    from manapy_amd.gmsh import domain_from_gmsh
    here = os.path.dirname(os.path.abspath(__file__))
    for d, types in ((make_tet_domain(3, 0.05, 1), {10}),
                     (domain_from_gmsh(os.path.join(here, 'golden', 'meshes', 'carre_hybrid.msh.gz'), 2), {5, 9})):
        p = str(tmp_path / 'x.vtu')
        write_vtu(p, *mesh_arrays(d), cell_data={'vol': d.cells.volume})
        arr = read_vtu(p)
        assert set(np.unique(arr['types'])) == types
        assert np.array_equal(arr['vol'], d.cells.volume)
        _check_mesh(d, arr)


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")
def test_device_fields_written_in_the_background(tmp_path):
    """A device field is copied on a side stream after the work queued before the call and
    written by a worker thread; later work on the field does not change the file."""
    import torch
    from manapy_amd.vtu import VtuWriter, read_vtu
    d = make_tri_domain(30, 20, 1.0, 1.0, 0.1, 4)
    C = int(d.nbcells)
    t = torch.arange(C, dtype=torch.float64, device='cuda')
    t.mul_(3.0)  # queued before the save
    wr = VtuWriter(d, str(tmp_path))
    wr.save(7, ['w'], [t])
    t.fill_(-1.0)  # queued after: must not reach the file
    wr.wait()
    arr = read_vtu(str(tmp_path / 'visu0-7.vtu'))
    assert np.array_equal(arr['w'], 3.0 * np.arange(C))
