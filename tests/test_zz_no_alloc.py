"""lmgpu_accel_intersect_host_noalloc, the allocation-free host ray cast,
against lmgpu_accel_intersect_host (same kernels, buffers the library
allocates per call): identical hit bytes on two calls in a row, the second
with fewer rays, and the size limit rejected before any device work."""
import ctypes

import numpy as np
import pytest

from lm_amd import _lib, scenes
from lm_amd.accel import HIT_FLOATS, Accel, rays_from_dict

MAX_RAYS = 65536


def cast(acc, rays, fn):
    r = np.ascontiguousarray(rays, dtype=np.float32)
    h = np.full((r.shape[0], HIT_FLOATS), np.nan, dtype=np.float32)
    _lib.call(fn, acc.handle, r.ctypes.data_as(ctypes.c_void_p), h.ctypes.data_as(ctypes.c_void_p), r.shape[0])
    return h


@pytest.mark.gpu
def test_noalloc_equals_host_intersect(device):
    sph, cl = scenes.config1_mesh(levels=5, clutter_count=2000)
    m = scenes.merge_meshes([sph, cl])
    acc = Accel(4, 2)
    acc.build(m.triangles())
    rays = rays_from_dict(scenes.surface_rays(sph, MAX_RAYS, 3))
    for n in (MAX_RAYS, 1000):
        sub = rays[:n] if n == MAX_RAYS else rays[-n:]
        ref = cast(acc, sub, 'lmgpu_accel_intersect_host')
        got = cast(acc, sub, 'lmgpu_accel_intersect_host_noalloc')
        prim = got[:, 0].view(np.int32)
        print(f'{n} rays: {np.count_nonzero(prim >= 0)} hits')
        assert np.count_nonzero(prim >= 0) > n // 20   # most rays leave the sphere outward
        assert got.tobytes() == ref.tobytes(), f'{n} rays: hits differ from lmgpu_accel_intersect_host'


@pytest.mark.gpu
def test_noalloc_rejects_oversize_batch(device):
    sph, _ = scenes.config1_mesh(levels=2, clutter_count=10)
    acc = Accel(4, 2)
    acc.build(sph.triangles())
    r = np.zeros((MAX_RAYS + 1, 8), dtype=np.float32)
    h = np.zeros((MAX_RAYS + 1, HIT_FLOATS), dtype=np.float32)
    with pytest.raises(_lib.LmError, match='at most 65536 rays'):
        _lib.call('lmgpu_accel_intersect_host_noalloc', acc.handle, r.ctypes.data_as(ctypes.c_void_p),
                  h.ctypes.data_as(ctypes.c_void_p), MAX_RAYS + 1)
