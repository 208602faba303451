"""The C ABI: libmfv.so loads (no GPU needed) and exports exactly the entry points
declared in include/mfv.h; the ctypes table binds each of them."""
import os
import re

from manapy_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, 'include', 'mfv.h')).read()
    src = re.sub(r'/\*.*?\*/', '', src, flags=re.S)
    return sorted(set(re.findall(r'\b(mfv_[a-z0-9_]+)\s*\(', src)))


def test_header_and_binding_agree():
    assert declared() == _lib.exported_symbols()


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in declared():
        assert hasattr(lib, name), name
    assert lib.mfv_version().decode().startswith('mfv')


def test_errors_are_reported_without_a_gpu():
    """Argument validation runs on the host: a NULL descriptor is rejected with a message."""
    lib = _lib.load()
    rc = lib.mfv_facetocell(None, None, None, None)
    assert rc == -1
    assert 'NULL' in lib.mfv_last_error().decode()


def test_mesh_struct_layout():
    import ctypes
    # 6 int32 + 17 pointers + 2 int32, natural alignment
    assert ctypes.sizeof(_lib.MfvMesh) == 6 * 4 + 17 * 8 + 2 * 4
