"""``renderer::pt_mi355x`` and ``film::bitmap`` host mirrors over the C ABI.

``RendererPT`` takes the construction properties of the reference
``renderer::pt`` (src/renderer/renderer_pt.cpp:51-77 and the spp schedulers,
src/scheduler.cpp:32-35,72-75): ``scene``, ``output``, ``max_verts``,
``seed``, ``sampling_mode`` (naive|nee|mis), ``scheduler`` (sample|time),
``spp`` / ``render_time``; plus ``rr_depth`` (Russian roulette when
num_verts > rr_depth; the reference hard-codes 5, renderer_pt.cpp:281).
``render()`` returns ``{"processed", "elapsed"}`` as the reference does
(:314) and leaves the mean image (sum / spp) in the film.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import _lib
from .accel import _stream_ptr

_MODES = {'naive': 0, 'nee': 1, 'mis': 2}
_PRIMARY = {'pixel': 0, 'image': 1}   # primary_ray_sampling_mode


class Film:
    """film::bitmap: FP64 RGB, [h][w][3], row 0 = bottom (film_bitmap.cpp:123-275).

    Besides the mean image it keeps the per-pixel sum and sum of squares of
    the per-sample estimates of the last render (for Welch tests)."""

    def __init__(self, w: int, h: int):
        self.w = int(w)
        self.h = int(h)
        self.data = np.zeros((self.h, self.w, 3), dtype=np.float64)
        self.sum = None
        self.sumsq = None
        self.spp = 0

    def size(self):
        return self.w, self.h

    def num_pixels(self) -> int:
        return self.w * self.h

    def aspect(self) -> float:
        return self.w / self.h

    def clear(self):
        self.data[:] = 0

    def rescale(self, s: float):
        self.data *= s

    def buffer(self) -> np.ndarray:
        return self.data

    def set_pixel(self, x: int, y: int, v):
        self.data[y, x] = v

    def splat_pixel(self, x: int, y: int, v):
        self.data[y, x] += v

    def raster_to_pixel(self, rp):
        """include/lm/film.h:166-171 (clamped)."""
        x = min(max(int(rp[0] * self.w), 0), self.w - 1)
        y = min(max(int(rp[1] * self.h), 0), self.h - 1)
        return x, y

    def splat(self, rp, v):
        self.splat_pixel(*self.raster_to_pixel(rp), v)


def scene_desc(flat):
    """lmgpu_scene_desc of a lm_amd.scenes.FlatScene; returns (keepalive
    arrays, descriptor). The descriptor points into the arrays."""
    keep = {
        'verts': np.ascontiguousarray(flat.verts, dtype=np.float32),
        'normals': np.ascontiguousarray(flat.normals, dtype=np.float32),
        'tri_prim': np.ascontiguousarray(flat.tri_prim, dtype=np.int32),
        'mat_type': np.ascontiguousarray(flat.mat_type, dtype=np.int32),
        'mat_param': np.ascontiguousarray(flat.mat_param, dtype=np.float32),
        'prim_Ke': np.ascontiguousarray(flat.prim_Ke, dtype=np.float32),
        'light_prims': np.ascontiguousarray(flat.light_prims, dtype=np.int32),
    }
    k = keep
    d = _lib.SceneDesc()
    d.num_tris = flat.num_triangles
    d.verts = k['verts'].ctypes.data
    d.normals = k['normals'].ctypes.data
    d.tri_prim = k['tri_prim'].ctypes.data
    d.num_prims = int(k['mat_type'].shape[0])
    d.mat_type = k['mat_type'].ctypes.data
    d.mat_param = k['mat_param'].ctypes.data
    d.prim_Ke = k['prim_Ke'].ctypes.data
    d.num_area_lights = int(k['light_prims'].shape[0])
    d.light_prims = k['light_prims'].ctypes.data if k['light_prims'].size else None
    d.has_env = 1 if flat.has_env else 0
    for i in range(3):
        d.env_Le[i] = float(flat.env_Le[i])
        d.cam_pos[i] = float(flat.camera.position[i])
        d.cam_center[i] = float(flat.camera.center[i])
        d.cam_up[i] = float(flat.camera.up[i])
    d.vfov = float(flat.camera.vfov)
    d.aspect = float(flat.camera.aspect)
    d.w = int(flat.w)
    d.h = int(flat.h)
    nprim = d.num_prims
    mat_ks = flat.mat_ks if flat.mat_ks is not None else np.zeros((nprim, 3), dtype=np.float32)
    mat_tex = flat.mat_tex if flat.mat_tex is not None else np.full(nprim, -1, dtype=np.int32)
    k['mat_ks'] = np.ascontiguousarray(mat_ks, dtype=np.float32)
    k['mat_tex'] = np.ascontiguousarray(mat_tex, dtype=np.int32)
    d.mat_ks = k['mat_ks'].ctypes.data
    d.mat_tex = k['mat_tex'].ctypes.data
    if flat.textures:
        k['uvs'] = np.ascontiguousarray(flat.uvs, dtype=np.float32)
        k['tex_info'], k['tex_data'] = flat.texture_tables()
        d.uvs = k['uvs'].ctypes.data
        d.num_textures = len(flat.textures)
        d.tex_info = k['tex_info'].ctypes.data
        d.tex_data = k['tex_data'].ctypes.data
        d.tex_data_len = int(k['tex_data'].size)
    return keep, d


class GpuScene:
    """Device scene (geometry, accel, materials, lights, camera) built from a
    lm_amd.scenes.FlatScene through lmgpu_scene_create."""

    def __init__(self, flat, width: int = 4, leaf_size: int = 2, refill: int = 32, builder: int = 1,
                 radius: int = 16, collapse: int = 0, sah_tri_cost: float = 0.0):
        self.flat = flat
        self._keep, d = scene_desc(flat)
        opts = _lib.BuildOpts(width, leaf_size, refill, builder, radius, collapse, sah_tri_cost)
        h = ctypes.c_void_p()
        _lib.call('lmgpu_scene_create', ctypes.byref(h), ctypes.byref(d), ctypes.byref(opts))
        self._h = h
        self.w, self.h = int(flat.w), int(flat.h)
        v = np.asarray(flat.verts, dtype=np.float32).reshape(-1, 3)
        self._lo, self._hi = v.min(axis=0), v.max(axis=0)

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value and _lib._lib is not None:
            _lib._lib.lmgpu_scene_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def accel_info(self) -> dict:
        a = ctypes.c_void_p()
        _lib.call('lmgpu_scene_accel', self._h, ctypes.byref(a))
        i = _lib.AccelInfo()
        _lib.call('lmgpu_accel_info_get', a, ctypes.byref(i))
        return {f: getattr(i, f) for f, _ in _lib.AccelInfo._fields_}

    def bounds(self):
        """Scene bounding box (lo, hi) of the triangle vertices."""
        return self._lo.copy(), self._hi.copy()

    def _film(self, t, name, device=None):
        """Device pointer of a [h, w, 3] float64 contiguous CUDA film tensor
        of this scene's size (ValueError otherwise, before any call)."""
        import torch
        return _lib.device_tensor(t, name, torch.float64, shape=(self.h, self.w, 3), device=device)

    def render_raw(self, spp: int = 1, max_verts: int = 8, sampling_mode: int = 2, rr_depth: Optional[int] = None,
                   seed: int = 0, sample_offset: int = 0, render_time: float = 0.0,
                   film_sum=None, film_sumsq=None, stream=None, guide=None, guide_opts: Optional[dict] = None,
                   primary_mode: str = 'pixel', num_samples: int = 0):
        """lmgpu_pt_render (lmgpu_pt_render_guided with a `guide`). With
        torch CUDA tensors for film_sum / film_sumsq ([h,w,3] float64) the
        film stays on the device and is accumulated into; otherwise host
        arrays are returned (fresh sums). primary_mode 'image' is
        primary_ray_sampling_mode = "image" with scheduler::spi::sample over
        `num_samples` samples (or spi::time with render_time); spp is then
        ignored, the film value is sum * w*h / stats['samples'] and there are
        no sums of squares (q is None; film_sumsq must be None). guide_opts keys: alpha,
        learning_rate, batch, steps_per_pass, record_depth, weight_clamp.
        rr_depth: Russian roulette once a path has more vertices than this;
        default 5 (renderer_pt.cpp:281) without a guide and max_verts (no
        roulette) with one, because roulette on throughput alone cuts the
        low-throughput paths that the guide steers toward the light."""
        p = _lib.PTParams()
        p.spp = int(spp)
        p.max_verts = int(max_verts)
        p.sampling_mode = int(sampling_mode)
        if rr_depth is None:
            rr_depth = 5 if guide is None else max(int(max_verts), 5)
        p.rr_depth = int(rr_depth)
        p.seed = int(seed)
        p.sample_offset = int(sample_offset)
        p.guiding = 0
        p.render_time = float(render_time)
        if primary_mode not in _PRIMARY:
            raise ValueError(f'invalid primary_mode {primary_mode!r}')
        p.primary_mode = _PRIMARY[primary_mode]
        p.num_samples = int(num_samples)
        image = primary_mode == 'image'
        if image and film_sumsq is not None:
            raise ValueError('image sampling mode has no sums of squares: pass film_sumsq=None')
        st = _lib.PTStats()
        if guide is not None:
            o = _lib.GuideOpts(*GUIDE_DEFAULTS)
            for k, v in (guide_opts or {}).items():
                if k not in ('alpha', 'learning_rate', 'batch', 'steps_per_pass', 'record_depth', 'weight_clamp'):
                    raise ValueError(f'unknown guide option {k!r}')
                setattr(o, k, v)
            fn, pre = 'lmgpu_pt_render_guided', (self._h, ctypes.byref(p), guide.handle, ctypes.byref(o))
        else:
            fn, pre = 'lmgpu_pt_render', (self._h, ctypes.byref(p))
        on_device = film_sum is not None and hasattr(film_sum, 'data_ptr')
        if on_device:
            sp = self._film(film_sum, 'film_sum')
            sq = ctypes.c_void_p(self._film(film_sumsq, 'film_sumsq', film_sum.device)) \
                if film_sumsq is not None else None
            _lib.call(fn, *pre, ctypes.c_void_p(sp), sq, 1, ctypes.byref(st),
                      ctypes.c_void_p(_stream_ptr(stream)))
            s, q = film_sum, film_sumsq
        else:
            s = np.zeros((self.h, self.w, 3), dtype=np.float64)
            q = None if image else np.zeros((self.h, self.w, 3), dtype=np.float64)
            _lib.call(fn, *pre, s.ctypes.data_as(ctypes.c_void_p),
                      q.ctypes.data_as(ctypes.c_void_p) if q is not None else None, 0,
                      ctypes.byref(st), ctypes.c_void_p(_stream_ptr(stream)))
        stats = {'samples': st.samples, 'spp': st.spp, 'extension_rays': st.extension_rays,
                 'shadow_rays': st.shadow_rays, 'elapsed': st.elapsed_s, 'guide_records': st.guide_records}
        return s, q, stats

    def render_lt(self, num_samples: int = 0, max_verts: int = 8, seed: int = 0, pass_offset: int = 0,
                  render_time: float = 0.0, film=None, stream=None):
        """lmgpu_lt_render: splat sums of `num_samples` light paths (or of
        chunks of whole w*h-path passes for `render_time` s), both in the
        regenerating loop. A torch CUDA tensor `film`
        ([h,w,3] float64) is accumulated into on the device; otherwise a fresh
        host array is returned. The film value is sum * w*h / samples."""
        p = _lib.LTParams()
        p.num_samples = int(num_samples)
        p.max_verts = int(max_verts)
        p.pass_offset = int(pass_offset)
        p.seed = int(seed)
        p.render_time = float(render_time)
        st = _lib.PTStats()
        if film is not None and not hasattr(film, 'data_ptr'):
            raise ValueError('film must be a torch CUDA tensor ([h,w,3] float64) or None')
        if film is not None:
            _lib.call('lmgpu_lt_render', self._h, ctypes.byref(p), ctypes.c_void_p(self._film(film, 'film')), 1,
                      ctypes.byref(st), ctypes.c_void_p(_stream_ptr(stream)))
            s = film
        else:
            s = np.zeros((self.h, self.w, 3), dtype=np.float64)
            _lib.call('lmgpu_lt_render', self._h, ctypes.byref(p), s.ctypes.data_as(ctypes.c_void_p), 0,
                      ctypes.byref(st), ctypes.c_void_p(_stream_ptr(stream)))
        stats = {'samples': st.samples, 'passes': st.spp, 'extension_rays': st.extension_rays,
                 'shadow_rays': st.shadow_rays, 'elapsed': st.elapsed_s}
        return s, stats

    def render_bdpt(self, num_samples: int = 0, max_verts: int = 8, seed: int = 0, pass_offset: int = 0,
                    render_time: float = 0.0, min_verts: int = 2, film=None, stream=None):
        """lmgpu_bdpt_render: splat sums of `num_samples` bidirectional
        samples (passes of min(w*h, 2^k) samples, 2^k from 2^21 down to 2^18 as max_verts
        grows); same film contract as render_lt."""
        p = _lib.LTParams()
        p.num_samples = int(num_samples)
        p.max_verts = int(max_verts)
        p.pass_offset = int(pass_offset)
        p.seed = int(seed)
        p.render_time = float(render_time)
        p.min_verts = int(min_verts)
        st = _lib.PTStats()
        if film is not None and not hasattr(film, 'data_ptr'):
            raise ValueError('film must be a torch CUDA tensor ([h,w,3] float64) or None')
        if film is not None:
            _lib.call('lmgpu_bdpt_render', self._h, ctypes.byref(p), ctypes.c_void_p(self._film(film, 'film')), 1,
                      ctypes.byref(st), ctypes.c_void_p(_stream_ptr(stream)))
            s = film
        else:
            s = np.zeros((self.h, self.w, 3), dtype=np.float64)
            _lib.call('lmgpu_bdpt_render', self._h, ctypes.byref(p), s.ctypes.data_as(ctypes.c_void_p), 0,
                      ctypes.byref(st), ctypes.c_void_p(_stream_ptr(stream)))
        stats = {'samples': st.samples, 'passes': st.spp, 'extension_rays': st.extension_rays,
                 'shadow_rays': st.shadow_rays, 'elapsed': st.elapsed_s}
        return s, stats


class RendererRaycast:
    """renderer::raycast_mi355x: drop-in for renderer::raycast
    (src/renderer/renderer_raycast.cpp:48-97) on the GPU accel."""

    def __init__(self, scene: GpuScene, output: Film, bg_color=(0.0, 0.0, 0.0), use_constant_color: bool = False,
                 visualize_normal: bool = False, color=None):
        if (output.w, output.h) != (scene.w, scene.h):
            raise ValueError('film size differs from the scene film size')
        self.scene = scene
        self.film = output
        p = _lib.RaycastParams()
        for i in range(3):
            p.bg_color[i] = float(bg_color[i])
        if color is not None:
            p.use_color = 1
            for i in range(3):
                p.color[i] = float(color[i])
        p.use_constant_color = 1 if use_constant_color else 0
        p.visualize_normal = 1 if visualize_normal else 0
        self._p = p

    def render(self, stream=None) -> dict:
        import time
        t0 = time.perf_counter()
        img = np.zeros((self.scene.h, self.scene.w, 3), dtype=np.float64)
        _lib.call('lmgpu_raycast_render', self.scene.handle, ctypes.byref(self._p),
                  img.ctypes.data_as(ctypes.c_void_p), 0, ctypes.c_void_p(_stream_ptr(stream)))
        self.film.data[:] = img
        return {'elapsed': time.perf_counter() - t0}


# lmgpu_guide_opts defaults: alpha, learning_rate, batch, steps_per_pass, record_depth
GUIDE_DEFAULTS = (0.65, 1e-2, 1 << 14, 4, 8)


class RendererPT:
    """renderer::pt_mi355x: drop-in for renderer::pt (renderer_pt.cpp:22-315)
    with both primary_ray_sampling_mode values ('pixel', 'image')."""

    def __init__(self, scene: GpuScene, output: Film, max_verts: int, seed: Optional[int] = None,
                 sampling_mode: str = 'mis', scheduler: str = 'sample', spp: Optional[int] = None,
                 render_time: Optional[float] = None, rr_depth: int = 5,
                 primary_ray_sampling_mode: str = 'pixel', num_samples: Optional[int] = None):
        if sampling_mode not in _MODES:
            raise ValueError(f'invalid sampling_mode {sampling_mode!r}')
        if primary_ray_sampling_mode not in _PRIMARY:
            raise ValueError(f'invalid primary_ray_sampling_mode {primary_ray_sampling_mode!r}')
        image = primary_ray_sampling_mode == 'image'
        # pixel mode runs scheduler::spp::*, image mode scheduler::spi::* (renderer_pt.cpp:60-76)
        if scheduler == 'sample' and not image and not spp:
            raise ValueError('scheduler "sample" needs spp')
        if scheduler == 'sample' and image and not num_samples:
            raise ValueError('scheduler "sample" in image mode needs num_samples')
        if scheduler == 'time' and not render_time:
            raise ValueError('scheduler "time" needs render_time')
        if (output.w, output.h) != (scene.w, scene.h):
            raise ValueError('film size differs from the scene film size')
        self.scene = scene
        self.film = output
        self.max_verts = int(max_verts)
        self.seed = int(seed) if seed is not None else int(np.random.SeedSequence().generate_state(1)[0])
        self.mode = _MODES[sampling_mode]
        self.scheduler = scheduler
        self.spp = int(spp or 0)
        self.render_time = float(render_time or 0.0)
        self.rr_depth = int(rr_depth)
        self.primary_mode = primary_ray_sampling_mode
        self.num_samples = int(num_samples or 0)

    def render(self) -> dict:
        film = self.film
        film.clear()
        s, q, st = self.scene.render_raw(
            spp=self.spp if self.scheduler == 'sample' else 0, max_verts=self.max_verts,
            sampling_mode=self.mode, rr_depth=self.rr_depth, seed=self.seed,
            render_time=self.render_time if self.scheduler == 'time' else 0.0, primary_mode=self.primary_mode,
            num_samples=self.num_samples if self.scheduler == 'sample' else 0)
        if self.primary_mode == 'image':
            # Film::rescale(w*h / processed) (renderer_pt.cpp:310-312)
            processed = st['samples']
            film.sum, film.sumsq, film.spp = s, None, st['spp']
            film.data[:] = s * (film.num_pixels() / processed)
        else:
            processed = st['spp']
            film.sum, film.sumsq, film.spp = s, q, processed
            film.data[:] = s / processed
        return {'processed': processed, 'elapsed': st['elapsed'], 'stats': st}


class RendererLT:
    """renderer::lt_mi355x: drop-in for renderer::lt (light tracing with NEE to
    the camera, src/renderer/renderer_lt.cpp:22-170) with the image-space
    schedulers scheduler::spi::sample (`num_samples` light paths) and
    scheduler::spi::time (`render_time` seconds, in passes of w*h paths)."""

    def __init__(self, scene: GpuScene, output: Film, max_verts: int, seed: Optional[int] = None,
                 scheduler: str = 'sample', num_samples: Optional[int] = None, render_time: Optional[float] = None):
        if scheduler == 'sample' and not num_samples:
            raise ValueError('scheduler "sample" needs num_samples')
        if scheduler == 'time' and not render_time:
            raise ValueError('scheduler "time" needs render_time')
        if scheduler not in ('sample', 'time'):
            raise ValueError(f'invalid scheduler {scheduler!r}')
        if (output.w, output.h) != (scene.w, scene.h):
            raise ValueError('film size differs from the scene film size')
        self.scene = scene
        self.film = output
        self.max_verts = int(max_verts)
        self.seed = int(seed) if seed is not None else int(np.random.SeedSequence().generate_state(1)[0])
        self.scheduler = scheduler
        self.num_samples = int(num_samples or 0)
        self.render_time = float(render_time or 0.0)

    def render(self) -> dict:
        film = self.film
        film.clear()
        s, st = self.scene.render_lt(
            num_samples=self.num_samples if self.scheduler == 'sample' else 0, max_verts=self.max_verts,
            seed=self.seed, render_time=self.render_time if self.scheduler == 'time' else 0.0)
        processed = st['samples']
        # Film::rescale(w*h / processed) of the splatted sums (renderer_lt.cpp:166)
        film.data[:] = s * (film.num_pixels() / processed)
        return {'processed': processed, 'elapsed': st['elapsed'], 'stats': st}


class RendererBDPT:
    """renderer::bdpt_mi355x: drop-in for renderer::bdpt (bidirectional path
    tracing with power-heuristic MIS, src/renderer/renderer_bdpt.cpp:339-411)
    with the keys of Renderer_Path_Base::construct (:43-55): max_verts,
    min_verts (2), seed, scheduler (sample: num_samples | time: render_time)."""

    def __init__(self, scene: GpuScene, output: Film, max_verts: int, seed: Optional[int] = None,
                 scheduler: str = 'sample', num_samples: Optional[int] = None, render_time: Optional[float] = None,
                 min_verts: int = 2):
        if scheduler == 'sample' and not num_samples:
            raise ValueError('scheduler "sample" needs num_samples')
        if scheduler == 'time' and not render_time:
            raise ValueError('scheduler "time" needs render_time')
        if scheduler not in ('sample', 'time'):
            raise ValueError(f'invalid scheduler {scheduler!r}')
        if (output.w, output.h) != (scene.w, scene.h):
            raise ValueError('film size differs from the scene film size')
        self.scene = scene
        self.film = output
        self.max_verts = int(max_verts)
        self.min_verts = int(min_verts)
        self.seed = int(seed) if seed is not None else int(np.random.SeedSequence().generate_state(1)[0])
        self.scheduler = scheduler
        self.num_samples = int(num_samples or 0)
        self.render_time = float(render_time or 0.0)

    def render(self) -> dict:
        film = self.film
        film.clear()
        s, st = self.scene.render_bdpt(
            num_samples=self.num_samples if self.scheduler == 'sample' else 0, max_verts=self.max_verts,
            seed=self.seed, render_time=self.render_time if self.scheduler == 'time' else 0.0,
            min_verts=self.min_verts)
        processed = st['samples']
        film.data[:] = s * (film.num_pixels() / processed)   # renderer_bdpt.cpp:398-399
        return {'processed': processed, 'elapsed': st['elapsed'], 'stats': st}
