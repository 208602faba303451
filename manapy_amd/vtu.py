"""VTU (VTK XML unstructured grid) output of cell and node fields, written off the critical path.

The reference writes ``results/visu{rank}-{miter}.vtu`` through meshio and a
``results/visu{miter}.pvtu`` index naming each rank's piece (``ddm/domain.py:624-849``:
``save_on_cell``, ``save_on_node``, ``save_on_cell_multi``, ``save_on_node_multi``). Here the
same files are written by this module: one piece (one GPU = one subdomain), cells in the
Domain's own order with a per-cell VTK type (5 triangle, 9 quadrangle, 10 tetrahedron, 12
hexahedron, 14 pyramid), arrays as base64 binary with a UInt64 byte-count header.

Fields may be device tensors: ``VtuWriter.save`` starts their copy into pinned host memory on a
side stream and hands the wait and the file write to a worker thread, so the solver loop goes
on while the output is produced; ``VtuWriter.wait()`` joins every pending write.
"""
import base64
import os
import struct
import threading

import numpy as np

VTK_TYPES = {(2, 3): 5, (2, 4): 9, (3, 4): 10, (3, 8): 12, (3, 5): 14}


def _b64(a):
    raw = np.ascontiguousarray(a).tobytes()
    return base64.b64encode(struct.pack('<Q', len(raw)) + raw).decode('ascii')


def _array(name, a, ncomp=1):
    a = np.asarray(a)
    vt = {np.dtype(np.float64): 'Float64', np.dtype(np.float32): 'Float32', np.dtype(np.int64): 'Int64',
          np.dtype(np.int32): 'Int32', np.dtype(np.uint8): 'UInt8'}[a.dtype]
    return ('<DataArray type="%s" Name="%s" NumberOfComponents="%d" format="binary">\n%s\n</DataArray>\n'
            % (vt, name, ncomp, _b64(a)))


def mesh_arrays(domain):
    """Points [N, 3], connectivity, offsets and VTK types of a Domain's cells (Domain order)."""
    nodeid = np.asarray(domain.cells.nodeid)
    count = nodeid[:, -1].astype(np.int64)
    dim = int(domain.dim)
    mask = np.arange(nodeid.shape[1] - 1)[None, :] < count[:, None]
    conn = nodeid[:, :-1][mask].astype(np.int64)
    offsets = np.cumsum(count).astype(np.int64)
    types = np.array([VTK_TYPES[(dim, int(c))] for c in count], dtype=np.uint8) if len(count) else \
        np.zeros(0, np.uint8)
    points = np.ascontiguousarray(np.asarray(domain.nodes.vertex)[:, :3], dtype=np.float64)
    return points, conn, offsets, types


def write_vtu(path, points, conn, offsets, types, cell_data=None, point_data=None):
    """One .vtu piece with the given cell and point scalars ({name: [C] or [N] array})."""
    out = ['<?xml version="1.0"?>\n',
           '<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" header_type="UInt64">\n',
           '<UnstructuredGrid>\n',
           '<Piece NumberOfPoints="%d" NumberOfCells="%d">\n' % (len(points), len(offsets)),
           '<Points>\n', _array('Points', points, 3), '</Points>\n',
           '<Cells>\n', _array('connectivity', conn), _array('offsets', offsets), _array('types', types),
           '</Cells>\n']
    if cell_data:
        out.append('<CellData>\n')
        out += [_array(k, np.asarray(v, dtype=np.float64)) for k, v in cell_data.items()]
        out.append('</CellData>\n')
    if point_data:
        out.append('<PointData>\n')
        out += [_array(k, np.asarray(v, dtype=np.float64)) for k, v in point_data.items()]
        out.append('</PointData>\n')
    out += ['</Piece>\n', '</UnstructuredGrid>\n', '</VTKFile>\n']
    tmp = path + '.tmp'
    with open(tmp, 'w') as fh:
        fh.write(''.join(out))
    os.replace(tmp, path)


def write_pvtu(path, pieces, names, on_cells, precision='Float64'):
    """The parallel index the reference writes next to its pieces (ddm/domain.py:654-675)."""
    kind = 'PCellData' if on_cells else 'PPointData'
    lines = ['<?xml version="1.0"?>\n',
             '<VTKFile type="PUnstructuredGrid" version="1.0" byte_order="LittleEndian" header_type="UInt64">\n',
             '<PUnstructuredGrid GhostLevel="0">\n', '<PPoints>\n',
             '<PDataArray type="Float64" Name="Points" NumberOfComponents="3" format="binary"/>\n', '</PPoints>\n',
             '<PCells>\n', '<PDataArray type="Int64" Name="connectivity" format="binary"/>\n',
             '<PDataArray type="Int64" Name="offsets" format="binary"/>\n',
             '<PDataArray type="UInt8" Name="types" format="binary"/>\n', '</PCells>\n', '<%s>\n' % kind]
    lines += ['<PDataArray type="%s" Name="%s" format="binary"/>\n' % (precision, n) for n in names]
    lines += ['</%s>\n' % kind] + ['<Piece Source="%s"/>\n' % p for p in pieces]
    lines += ['</PUnstructuredGrid>\n', '</VTKFile>\n']
    with open(path, 'w') as fh:
        fh.write(''.join(lines))


def read_vtu(path):
    """Arrays of a .vtu written by write_vtu: {name: numpy array} (checks and tools)."""
    import xml.etree.ElementTree as ET
    dt = {'Float64': np.float64, 'Float32': np.float32, 'Int64': np.int64, 'Int32': np.int32, 'UInt8': np.uint8}
    out = {}
    for el in ET.parse(path).iter('DataArray'):
        raw = base64.b64decode(el.text.strip())
        n = struct.unpack('<Q', raw[:8])[0]
        a = np.frombuffer(raw[8:8 + n], dtype=dt[el.get('type')])
        nc = int(el.get('NumberOfComponents', '1'))
        out[el.get('Name')] = a.reshape(-1, nc) if nc > 1 else a
    return out


class VtuWriter(object):
    """Asynchronous writer of a Domain's fields (cell or node arrays, NumPy or device tensors)."""

    def __init__(self, domain, path='results'):
        self.domain = domain
        self.path = path
        os.makedirs(path, exist_ok=True)
        self._mesh = None
        self._threads = []
        self._stream = None

    def _host(self, values):
        """Host copies of ``values``: device tensors go to pinned memory on a side stream (the
        copies wait for the work already queued on the current stream, and work queued later
        on it waits for the copies); returns the arrays and the event the worker must wait for
        (None if nothing was on the device)."""
        import torch
        hosts, ev = [], None
        dev = [v for v in values if torch.is_tensor(v) and v.is_cuda]
        if dev:
            if self._stream is None:
                self._stream = torch.cuda.Stream()
            self._stream.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(self._stream):
                for v in values:
                    if torch.is_tensor(v) and v.is_cuda:
                        h = torch.empty(v.shape, dtype=v.dtype, pin_memory=True)
                        h.copy_(v, non_blocking=True)
                        v.record_stream(self._stream)
                        hosts.append(h)
                    else:
                        hosts.append(v)
                ev = torch.cuda.Event()
                ev.record(self._stream)
            # later work on the current stream (which may overwrite the fields) waits for the
            # copies; the host does not
            torch.cuda.current_stream().wait_stream(self._stream)
        else:
            hosts = list(values)
        return hosts, ev

    def save(self, miter, names, values, on_cells=True, rank=0):
        """Write visu{rank}-{miter}.vtu and visu{miter}.pvtu in the background."""
        if self._mesh is None:
            self._mesh = mesh_arrays(self.domain)
        n = int(self.domain.nbcells if on_cells else self.domain.nbnodes)
        for v in values:
            if len(v) != n:
                raise ValueError("value size %d != number of %s %d" % (len(v), 'cells' if on_cells else 'nodes', n))
        hosts, ev = self._host(values)
        piece = 'visu%d-%d.vtu' % (rank, miter)

        def work():
            if ev is not None:
                ev.synchronize()
            arrays = {k: np.asarray(v.numpy() if hasattr(v, 'numpy') else v) for k, v in zip(names, hosts)}
            write_vtu(os.path.join(self.path, piece), *self._mesh, cell_data=arrays if on_cells else None,
                      point_data=None if on_cells else arrays)
            write_pvtu(os.path.join(self.path, 'visu%d.pvtu' % miter), [piece], list(names), on_cells)
        t = threading.Thread(target=work, daemon=True)
        t.start()
        self._threads.append(t)
        return t

    def wait(self):
        for t in self._threads:
            t.join()
        self._threads = []
