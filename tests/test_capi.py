"""The C ABI: libmfv.so loads (no GPU needed) and exports exactly the entry points
declared in include/mfv.h; the ctypes table binds each of them."""
import os
import re

import pytest

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


def test_every_mirrored_struct_matches_the_header(tmp_path):
    """Each ctypes mirror in the binding has the size and field offsets that a C compiler gives
    the struct of the same name in include/mfv.h."""
    import ctypes
    import re as _re
    import shutil
    import subprocess
    if shutil.which('gcc') is None:
        pytest.skip("no C compiler")
    mirrors = []
    for name in dir(_lib):
        cls = getattr(_lib, name)
        if isinstance(cls, type) and issubclass(cls, ctypes.Structure) and cls.__doc__:
            m = _re.search(r'struct (mfv_[a-z0-9_]+)', cls.__doc__)
            if m:
                mirrors.append((m.group(1), cls))
    assert len(mirrors) >= 10
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mfv.h"', 'int main(void) {']
    for cname, cls in mirrors:
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for f in cls._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f[0], cname, f[0]))
    lines += ['return 0;', '}']
    src = tmp_path / 'layout.c'
    src.write_text('\n'.join(lines))
    exe = tmp_path / 'layout'
    subprocess.run(['gcc', '-std=c11', '-I', os.path.join(ROOT, 'include'), str(src), '-o', str(exe)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                  text=True).stdout.splitlines())
    for cname, cls in mirrors:
        assert int(got[cname]) == ctypes.sizeof(cls), cname
        for f in cls._fields_:
            assert int(got['%s.%s' % (cname, f[0])]) == getattr(cls, f[0]).offset, (cname, f[0])


def test_tile_stencil_row_stride_matches_the_header(tmp_path):
    """The Python builders pad the tiles' stencil rows to the stride MFV_SW_TILES_KP gives."""
    import shutil
    import subprocess
    from manapy_amd.device import sw_tiles_kp
    if shutil.which('gcc') is None:
        pytest.skip("no C compiler")
    ks = list(range(0, 40))
    src = tmp_path / 'kp.c'
    src.write_text('#include <stdio.h>\n#include "mfv.h"\nint main(void) {\n' +
                   ''.join('printf("%%d\\n", MFV_SW_TILES_KP(%d));\n' % k for k in ks) + 'return 0;\n}\n')
    exe = tmp_path / 'kp'
    subprocess.run(['gcc', '-std=c11', '-I', os.path.join(ROOT, 'include'), str(src), '-o', str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert got == [sw_tiles_kp(k) for k in ks]
    assert all(kp % 8 == 0 and kp >= max(k, 8) for k, kp in zip(ks, got))


def test_noalloc_step_rejects_meshes_beyond_its_static_scratch():
    """The allocation-free step checks the mesh size against its static partials array before
    touching the device."""
    import ctypes
    lib = _lib.load()
    m = _lib.MfvMesh()
    m.nb_cells = 64 * (1 << 20) + 1
    rc = lib.mfv_sw_step_packed_tiled_noalloc(ctypes.byref(m), None, None, None, 9.81, 0.8, 2, None, 0, None)
    assert rc == -1
    assert 'allocation-free' in lib.mfv_last_error().decode()
