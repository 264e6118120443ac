"""``accel::mi355x`` — host-side drop-in for the reference ``accel::sahbvh``.

Mirrors ``lm::Accel`` (include/lm/accel.h:26-71): ``build(scene)`` and
``intersect(ray, tmin, tmax) -> Hit | None``, plus the batched device entry
points the GPU path tracer uses. All work runs in liblmgpu.so (HIP, gfx950);
torch provides device memory and streams.
"""
from __future__ import annotations

import ctypes
import dataclasses
from typing import Optional

import numpy as np

from . import _lib

RAY_FLOATS = 8   # org[3], tmin, dir[3], tmax  (lmgpu_ray, 32 B)
HIT_FLOATS = 4   # prim (int32 bits), t, u, v (lmgpu_hit, 16 B)


@dataclasses.dataclass
class Hit:
    """lm::Accel::Hit (include/lm/accel.h:49-55) with identity transforms."""
    t: float
    uv: tuple
    primitive: int   # primitive (scene node) index
    face: int        # face index within the primitive's mesh
    tri: int         # global (flattened) triangle index


def pack_rays(o: np.ndarray, d: np.ndarray, tmin, tmax) -> np.ndarray:
    """[n,8] float32 array in lmgpu_ray layout."""
    o = np.asarray(o, dtype=np.float32).reshape(-1, 3)
    d = np.asarray(d, dtype=np.float32).reshape(-1, 3)
    n = o.shape[0]
    r = np.empty((n, RAY_FLOATS), dtype=np.float32)
    r[:, 0:3] = o
    r[:, 3] = tmin
    r[:, 4:7] = d
    r[:, 7] = tmax
    return r


def rays_from_dict(rays: dict) -> np.ndarray:
    return pack_rays(rays['o'], rays['d'], rays['tmin'], rays['tmax'])


def unpack_hits(h: np.ndarray) -> dict:
    h = np.ascontiguousarray(h, dtype=np.float32).reshape(-1, HIT_FLOATS)
    return {'prim': h[:, 0].view(np.int32).copy(), 't': h[:, 1].astype(np.float64),
            'u': h[:, 2].astype(np.float64), 'v': h[:, 3].astype(np.float64)}


def _stream_ptr(stream) -> Optional[int]:
    if stream is None:
        try:
            import torch
            if torch.cuda.is_available():
                return torch.cuda.current_stream().cuda_stream
        except Exception:
            return None
        return None
    return int(stream)


class Accel:
    """accel::mi355x: GPU LBVH collapsed to a W-wide BVH + ordered traversal.
    refill 0 takes the library's default traversal refill threshold (8)."""

    def __init__(self, width: int = 4, leaf_size: int = 2, refill: int = 0, builder: int = 1, radius: int = 16,
                 collapse: int = 0, sah_tri_cost: float = 0.0):
        self.width = width
        self.leaf_size = leaf_size
        self.refill = refill
        self.builder = builder
        self.radius = radius
        self.collapse = collapse
        self.sah_tri_cost = sah_tri_cost
        h = ctypes.c_void_p()
        _lib.call('lmgpu_accel_create', ctypes.byref(h))
        self._h = h
        self.tri_prim = None
        self.tri_face = None

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value and _lib._lib is not None:
            _lib._lib.lmgpu_accel_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def build(self, scene_or_verts, stream=None):
        """Accel::build (include/lm/accel.h:38). Accepts a FlatScene (keeps its
        triangle -> (primitive, face) tables), a [N,9] float32 numpy array or a
        [N,9] float32 CUDA tensor."""
        verts = scene_or_verts
        if hasattr(scene_or_verts, 'verts'):
            verts = scene_or_verts.verts
            self.tri_prim = np.asarray(scene_or_verts.tri_prim)
            self.tri_face = np.asarray(scene_or_verts.tri_face)
        opts = _lib.BuildOpts(self.width, self.leaf_size, self.refill, self.builder, self.radius, self.collapse,
                              self.sah_tri_cost)
        try:
            import torch
            is_tensor = isinstance(verts, torch.Tensor)
        except Exception:
            is_tensor = False
        if is_tensor:
            vp = _lib.device_tensor(verts, 'verts', torch.float32, cols=9)
            n = verts.shape[0]
            _lib.call('lmgpu_accel_build', self._h, ctypes.c_void_p(vp), n, 1,
                      ctypes.byref(opts), ctypes.c_void_p(_stream_ptr(stream)))
        else:
            v = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 9)
            _lib.call('lmgpu_accel_build', self._h, v.ctypes.data_as(ctypes.c_void_p), v.shape[0], 0,
                      ctypes.byref(opts), ctypes.c_void_p(_stream_ptr(stream)))

    def serialize(self) -> bytes:
        """The built BVH as bytes (lmgpu_accel_serialize)."""
        n = ctypes.c_size_t(0)
        _lib.call('lmgpu_accel_serialize', self._h, None, ctypes.byref(n))
        buf = ctypes.create_string_buffer(n.value)
        _lib.call('lmgpu_accel_serialize', self._h, ctypes.cast(buf, ctypes.c_void_p), ctypes.byref(n))
        return buf.raw[:n.value]

    def deserialize(self, data: bytes):
        """Replace the BVH with a serialized one (validated before upload)."""
        buf = ctypes.create_string_buffer(bytes(data), len(data))
        opts = _lib.BuildOpts(0, 0, self.refill, 0, 0)
        _lib.call('lmgpu_accel_deserialize', self._h, ctypes.cast(buf, ctypes.c_void_p), len(data), ctypes.byref(opts))

    def save(self, path: str):
        with open(path, 'wb') as f:
            f.write(self.serialize())

    def load(self, path: str):
        with open(path, 'rb') as f:
            self.deserialize(f.read())

    def info(self) -> dict:
        i = _lib.AccelInfo()
        _lib.call('lmgpu_accel_info_get', self._h, ctypes.byref(i))
        return {f: getattr(i, f) for f, _ in _lib.AccelInfo._fields_}

    # -- single ray (lm::Accel::intersect) --------------------------------------
    def intersect(self, o, d, tmin: float, tmax: float) -> Optional[Hit]:
        """Accel::intersect (include/lm/accel.h:70): closest hit in [tmin, tmax]
        or None. The ray is converted to FP32 at the boundary."""
        r = pack_rays(np.asarray(o, dtype=np.float64)[None], np.asarray(d, dtype=np.float64)[None], tmin, tmax)
        h = np.zeros((1, HIT_FLOATS), dtype=np.float32)
        _lib.call('lmgpu_accel_intersect_host', self._h, r.ctypes.data_as(ctypes.c_void_p),
                  h.ctypes.data_as(ctypes.c_void_p), 1)
        res = unpack_hits(h)
        tri = int(res['prim'][0])
        if tri < 0:
            return None
        prim = int(self.tri_prim[tri]) if self.tri_prim is not None else -1
        face = int(self.tri_face[tri]) if self.tri_face is not None else tri
        return Hit(float(res['t'][0]), (float(res['u'][0]), float(res['v'][0])), prim, face, tri)

    def intersect_host(self, rays: np.ndarray) -> dict:
        """Batch from host memory ([n,8] lmgpu_ray rows); returns hit dict."""
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, RAY_FLOATS)
        h = np.zeros((r.shape[0], HIT_FLOATS), dtype=np.float32)
        _lib.call('lmgpu_accel_intersect_host', self._h, r.ctypes.data_as(ctypes.c_void_p),
                  h.ctypes.data_as(ctypes.c_void_p), r.shape[0])
        return unpack_hits(h)

    # -- batched device entry points ----------------------------------------------
    @staticmethod
    def _buffers(rays, out, name):
        """Checks rays ([n,8] float32) and the output buffer (hits [>=n,4]
        float32 or occl [>=n] uint8), all before any call; returns (n, ray
        pointer, output pointer)."""
        import torch
        rp = _lib.device_tensor(rays, 'rays', torch.float32, cols=RAY_FLOATS, cuda=False)
        n = rays.shape[0]
        if name == 'occl':
            op = _lib.device_tensor(out, 'occl', torch.uint8, min_rows=n, cuda=False)
        else:
            op = _lib.device_tensor(out, 'hits', torch.float32, cols=HIT_FLOATS, min_rows=n, cuda=False)
        _lib.on_device(('rays', rays), (name, out))
        return n, rp, op

    def intersect_device(self, rays, hits, stream=None):
        """rays: [n,8] float32 CUDA tensor; hits: [>=n,4] float32 CUDA tensor
        (column 0 holds the int32 triangle index bits). Asynchronous. Shapes,
        dtypes, contiguity and the device are checked before the call."""
        n, rp, hp = self._buffers(rays, hits, 'hits')
        _lib.call('lmgpu_accel_intersect', self._h, ctypes.c_void_p(rp), ctypes.c_void_p(hp), n,
                  ctypes.c_void_p(_stream_ptr(stream)))

    def occluded_device(self, rays, occl, stream=None):
        """rays: [n,8] float32 CUDA tensor; occl: [>=n] uint8 CUDA tensor."""
        n, rp, op = self._buffers(rays, occl, 'occl')
        _lib.call('lmgpu_accel_occluded', self._h, ctypes.c_void_p(rp), ctypes.c_void_p(op), n,
                  ctypes.c_void_p(_stream_ptr(stream)))

    def intersect_counted(self, rays, hits, stream=None):
        """As intersect_device; returns (node visits, triangle tests) summed
        over the rays. Synchronous."""
        n, rp, hp = self._buffers(rays, hits, 'hits')
        nv = ctypes.c_int64(0)
        nt = ctypes.c_int64(0)
        _lib.call('lmgpu_accel_intersect_counted', self._h, ctypes.c_void_p(rp), ctypes.c_void_p(hp), n,
                  ctypes.byref(nv), ctypes.byref(nt), ctypes.c_void_p(_stream_ptr(stream)))
        return nv.value, nt.value

    def status(self, stream=None):
        _lib.call('lmgpu_accel_status', self._h, ctypes.c_void_p(_stream_ptr(stream)))
