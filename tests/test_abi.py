"""CPU checks of the C ABI: liblmgpu.so loads and exports every function
include/lmgpu.h declares; host-side argument errors surface as LmError
without a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

from lm_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, 'include', 'lmgpu.h')).read()
    src = re.sub(r'/\*.*?\*/', '', src, flags=re.S)
    return sorted(set(re.findall(r'\b(lmgpu_[a-z_0-9]+)\s*\(', src)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(_lib.EXPORTED)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip('liblmgpu.so not built')
    L = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(L, name), name


def test_host_side_errors_without_gpu():
    L = _lib.lib()
    assert L.lmgpu_version().startswith(b'lmgpu')
    # NULL out-pointers are rejected before any device call
    assert L.lmgpu_accel_create(None) == _lib.ERR_INVALID_ARGUMENT
    assert b'NULL' in L.lmgpu_last_error()
    assert L.lmgpu_accel_build(None, None, 0, 0, None, None) == _lib.ERR_INVALID_ARGUMENT
    assert L.lmgpu_scene_create(None, None, None) == _lib.ERR_INVALID_ARGUMENT
    assert L.lmgpu_pt_render(None, None, None, None, 0, None, None) == _lib.ERR_INVALID_ARGUMENT
    assert L.lmgpu_pt_render_guided(None, None, None, None, None, None, 0, None, None) == _lib.ERR_INVALID_ARGUMENT
    assert L.lmgpu_guide_stats(None, None, None) == _lib.ERR_INVALID_ARGUMENT
    with pytest.raises(_lib.LmError):
        _lib.call('lmgpu_accel_info_get', None, None)


def test_structure_sizes_match_header():
    # lmgpu_ray 32 B, lmgpu_hit 16 B as the kernels assume
    assert ctypes.sizeof(_lib.BuildOpts) == 32
    assert ctypes.sizeof(_lib.PTParams) == 4 * 4 + 8 + 4 + 4 + 8 + 32
    assert ctypes.sizeof(_lib.PTStats) == 4 * 8 + 8 + 4 * 8
    assert ctypes.sizeof(_lib.GuideOpts) == 4 * 12
    assert ctypes.sizeof(_lib.LTParams) == 8 + 4 + 4 + 8 + 8 + 32


def test_host_components_built():
    """The C++ host components (liblmhost.so over the C ABI) and their test
    program are built and load (their GPU run is in test_raycast_gpu.py)."""
    pkg = os.path.join(ROOT, 'lightmetrica-v3_amd')
    lib = os.path.join(pkg, 'liblmhost.so')
    exe = os.path.join(pkg, 'bin', 'test_host')
    if not os.path.exists(lib):
        pytest.skip('liblmhost.so not built')
    ctypes.CDLL(lib)
    assert os.access(exe, os.X_OK)


def _serial(width, nodes, tris, num_leaves=1, magic=b'LMGPUBVH', version=1):
    """Serialized accel bytes: 128-byte header + nodes + triangles."""
    import struct
    nn, nt = len(nodes), len(tris)
    h = magic + struct.pack('<IIII', version, width, 2, 1) + struct.pack('<qqq', nt, nn, num_leaves)
    h += struct.pack('<d', 0.0) + struct.pack('<6f', 0, 0, 0, 1, 1, 1)
    h += b'\0' * (128 - len(h))
    body = b''
    for child, count in nodes:   # one node: 6*W floats of bounds, W children, W counts
        body += struct.pack(f'<{6 * width}f', *([0.0] * 3 * width + [1.0] * 3 * width))
        body += struct.pack(f'<{width}i', *child) + struct.pack(f'<{width}i', *count)
    for tid in tris:
        body += struct.pack('<3f', 0, 0, 0) + struct.pack('<i', tid) + struct.pack('<8f', *([0.0] * 8))
    return h + body


def test_deserialize_rejects_malformed_bvh_on_the_host():
    """lmgpu_accel_deserialize validates everything before any device work:
    magic, version, sizes, child ranges, tree structure (no cycles / shared
    children), leaf ranges and triangle ids."""
    from lm_amd.accel import Accel
    a = Accel(4)
    good = _serial(4, [([0, 0, 0, 0], [1, -1, -1, -1])], [0])
    bad = {
        'magic': _serial(4, [([0, 0, 0, 0], [1, -1, -1, -1])], [0], magic=b'NOTABVH!'),
        'version': _serial(4, [([0, 0, 0, 0], [1, -1, -1, -1])], [0], version=9),
        'size': good[:-4],
        'cycle': _serial(4, [([1, 0, 0, 0], [0, -1, -1, -1]), ([1, 0, 0, 0], [0, -1, -1, -1])], [0]),
        'child range': _serial(4, [([7, 0, 0, 0], [0, -1, -1, -1])], [0]),
        'leaf range': _serial(4, [([0, 0, 0, 0], [3, -1, -1, -1])], [0]),
        'triangle id': _serial(4, [([0, 0, 0, 0], [1, -1, -1, -1])], [5]),
    }
    for name, data in bad.items():
        with pytest.raises(_lib.LmError):
            a.deserialize(data)
        print(name, _lib.lib().lmgpu_last_error().decode())
