"""ctypes wrapper of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() (as
the checker) and by bench.py's cpu_baseline leg (as the timed CPU baseline).
The product package never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, 'liboracle.so')
_lib = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags='C_CONTIGUOUS')
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags='C_CONTIGUOUS')
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags='C_CONTIGUOUS')
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags='C_CONTIGUOUS')
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags='C_CONTIGUOUS')


class OrcScene(ctypes.Structure):
    _fields_ = [
        ('num_tris', ctypes.c_int64),
        ('verts', ctypes.c_void_p),
        ('normals', ctypes.c_void_p),
        ('tri_prim', ctypes.c_void_p),
        ('num_prims', ctypes.c_int32),
        ('mat_type', ctypes.c_void_p),
        ('mat_param', ctypes.c_void_p),
        ('prim_Ke', ctypes.c_void_p),
        ('num_area_lights', ctypes.c_int32),
        ('light_prims', ctypes.c_void_p),
        ('has_env', ctypes.c_int32),
        ('env_Le', ctypes.c_float * 3),
        ('cam_pos', ctypes.c_float * 3),
        ('cam_center', ctypes.c_float * 3),
        ('cam_up', ctypes.c_float * 3),
        ('vfov', ctypes.c_float),
        ('aspect', ctypes.c_float),
        ('w', ctypes.c_int32),
        ('h', ctypes.c_int32),
        ('mat_ks', ctypes.c_void_p),
        ('mat_tex', ctypes.c_void_p),
        ('uvs', ctypes.c_void_p),
        ('num_textures', ctypes.c_int32),
        ('tex_info', ctypes.c_void_p),
        ('tex_data', ctypes.c_void_p),
        ('tex_data_len', ctypes.c_int64),
    ]


class OrcPTParams(ctypes.Structure):
    _fields_ = [
        ('spp', ctypes.c_int32),
        ('max_verts', ctypes.c_int32),
        ('sampling_mode', ctypes.c_int32),
        ('rr_depth', ctypes.c_int32),
        ('seed', ctypes.c_uint64),
        ('nthreads', ctypes.c_int32),
        ('sample_offset', ctypes.c_int32),
        ('primary_mode', ctypes.c_int32),
        ('pad', ctypes.c_int32),
        ('num_samples', ctypes.c_int64),
    ]


class OrcLTParams(ctypes.Structure):
    _fields_ = [
        ('num_samples', ctypes.c_int64),
        ('max_verts', ctypes.c_int32),
        ('pass_offset', ctypes.c_int32),
        ('seed', ctypes.c_uint64),
        ('nthreads', ctypes.c_int32),
        ('min_verts', ctypes.c_int32),
    ]


class OrcPTStats(ctypes.Structure):
    _fields_ = [
        ('samples', ctypes.c_int64),
        ('extension_rays', ctypes.c_int64),
        ('shadow_rays', ctypes.c_int64),
        ('path_vertices', ctypes.c_int64),
        ('extension_node_visits', ctypes.c_int64),
        ('extension_tri_tests', ctypes.c_int64),
        ('shadow_node_visits', ctypes.c_int64),
        ('shadow_tri_tests', ctypes.c_int64),
    ]


def build_library(force: bool = False) -> str:
    """Compile liboracle.so with its Makefile (g++)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(['make', '-s', '-C', _HERE, '-B' if force else 'all'])
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build_library()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_accel_build.restype = ctypes.c_void_p
        L.orc_accel_build.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int]
        L.orc_accel_free.argtypes = [ctypes.c_void_p]
        L.orc_accel_num_nodes.restype = ctypes.c_int64
        L.orc_accel_num_nodes.argtypes = [ctypes.c_void_p]
        L.orc_accel_export.argtypes = [ctypes.c_void_p, _f64p, _i32p, _i32p]
        L.orc_accel_intersect.argtypes = [ctypes.c_void_p, _f32p, _f32p, _f32p, _f32p, ctypes.c_int64,
                                          _i32p, _f64p, _f64p, _f64p,
                                          ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                          ctypes.c_int]
        L.orc_brute_force.argtypes = [_f32p, ctypes.c_int64, _f32p, _f32p, _f32p, _f32p, ctypes.c_int64,
                                      _i32p, _f64p, _f64p, _f64p, ctypes.c_int]
        L.orc_accel_check.restype = ctypes.c_int
        L.orc_accel_check.argtypes = [ctypes.c_void_p]
        L.orc_classify.argtypes = [ctypes.c_void_p, _f32p, _f32p, _f32p, _f32p, ctypes.c_int64,
                                   _i32p, _f64p, ctypes.c_double, _u8p, ctypes.c_int]
        L.orc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.orc_mat_eval.argtypes = [ctypes.c_int, _f64p, _f64p, ctypes.c_int64, _f64p, _f64p, _f64p, _f64p]
        L.orc_mat_sample.argtypes = [ctypes.c_int, _f64p, _f64p, ctypes.c_int64, _f64p, _f64p, _f64p, _f64p,
                                     _i32p, _i32p]
        L.orc_pt_render.restype = ctypes.c_int
        L.orc_pt_render.argtypes = [ctypes.POINTER(OrcScene), ctypes.c_void_p, ctypes.POINTER(OrcPTParams),
                                    _f64p, ctypes.c_void_p, ctypes.POINTER(OrcPTStats)]
        L.orc_bdpt_render.restype = ctypes.c_int
        L.orc_bdpt_render.argtypes = [ctypes.POINTER(OrcScene), ctypes.c_void_p, ctypes.POINTER(OrcLTParams),
                                      _f64p, ctypes.POINTER(OrcPTStats)]
        L.orc_lt_render.restype = ctypes.c_int
        L.orc_lt_render.argtypes = [ctypes.POINTER(OrcScene), ctypes.c_void_p, ctypes.POINTER(OrcLTParams),
                                    _f64p, ctypes.POINTER(OrcPTStats)]
        _lib = L
    return _lib


class Accel:
    """accel::sahbvh restated in FP64 (binary SAH BVH + Moller-Trumbore)."""

    def __init__(self, verts: np.ndarray, nthreads: int = 0):
        self.verts = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 9)
        self.h = lib().orc_accel_build(self.verts, self.verts.shape[0], nthreads)

    def __del__(self):
        if getattr(self, 'h', None) and _lib is not None:
            _lib.orc_accel_free(self.h)
            self.h = None

    @property
    def num_nodes(self) -> int:
        return lib().orc_accel_num_nodes(self.h)

    def export(self):
        n = self.num_nodes
        b = np.zeros((n, 6), dtype=np.float64)
        info = np.zeros((n, 5), dtype=np.int32)
        idx = np.zeros(self.verts.shape[0], dtype=np.int32)
        lib().orc_accel_export(self.h, b, info, idx)
        return b, info, idx

    def check(self) -> int:
        return lib().orc_accel_check(self.h)

    def intersect(self, rays: dict, nthreads: int = 0, stats: bool = False):
        o, d, tmin, tmax = _ray_arrays(rays)
        n = tmin.shape[0]
        prim = np.empty(n, dtype=np.int32)
        t = np.empty(n)
        u = np.empty(n)
        v = np.empty(n)
        nv = ctypes.c_int64(0)
        nt = ctypes.c_int64(0)
        lib().orc_accel_intersect(self.h, o, d, tmin, tmax, n, prim, t, u, v,
                                  ctypes.byref(nv), ctypes.byref(nt), nthreads)
        out = {'prim': prim, 't': t, 'u': u, 'v': v}
        if stats:
            out['node_visits'] = nv.value
            out['tri_tests'] = nt.value
        return out

    def classify(self, rays: dict, ref: dict, kappa: float = 64.0, nthreads: int = 0) -> np.ndarray:
        o, d, tmin, tmax = _ray_arrays(rays)
        flags = np.zeros(tmin.shape[0], dtype=np.uint8)
        lib().orc_classify(self.h, o, d, tmin, tmax, tmin.shape[0],
                           np.ascontiguousarray(ref['prim'], dtype=np.int32),
                           np.ascontiguousarray(ref['t'], dtype=np.float64), kappa, flags, nthreads)
        return flags


def _ray_arrays(rays):
    o = np.ascontiguousarray(rays['o'], dtype=np.float32).reshape(-1)
    d = np.ascontiguousarray(rays['d'], dtype=np.float32).reshape(-1)
    tmin = np.ascontiguousarray(rays['tmin'], dtype=np.float32)
    tmax = np.ascontiguousarray(rays['tmax'], dtype=np.float32)
    return o, d, tmin, tmax


def brute_force(verts: np.ndarray, rays: dict, nthreads: int = 0):
    verts = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 9)
    o, d, tmin, tmax = _ray_arrays(rays)
    n = tmin.shape[0]
    prim = np.empty(n, dtype=np.int32)
    t = np.empty(n)
    u = np.empty(n)
    v = np.empty(n)
    lib().orc_brute_force(verts, verts.shape[0], o, d, tmin, tmax, n, prim, t, u, v, nthreads)
    return {'prim': prim, 't': t, 'u': u, 'v': v}


def philox4x32_10(ctr, key) -> np.ndarray:
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(c, k, out)
    return out


def make_orc_scene(flat) -> tuple:
    """Build the OrcScene struct from a lm_amd.scenes.FlatScene; returns
    (struct, keepalive) where keepalive holds the arrays."""
    keep = {
        'verts': np.ascontiguousarray(flat.verts, dtype=np.float32),
        'normals': np.ascontiguousarray(flat.normals, dtype=np.float32),
        'tri_prim': np.ascontiguousarray(flat.tri_prim, dtype=np.int32),
        'mat_type': np.ascontiguousarray(flat.mat_type, dtype=np.int32),
        'mat_param': np.ascontiguousarray(flat.mat_param, dtype=np.float32),
        'prim_Ke': np.ascontiguousarray(flat.prim_Ke, dtype=np.float32),
        'light_prims': np.ascontiguousarray(flat.light_prims, dtype=np.int32),
    }
    s = OrcScene()
    s.num_tris = flat.num_triangles
    s.verts = keep['verts'].ctypes.data
    s.normals = keep['normals'].ctypes.data
    s.tri_prim = keep['tri_prim'].ctypes.data
    s.num_prims = int(flat.mat_type.shape[0])
    s.mat_type = keep['mat_type'].ctypes.data
    s.mat_param = keep['mat_param'].ctypes.data
    s.prim_Ke = keep['prim_Ke'].ctypes.data
    s.num_area_lights = int(flat.light_prims.shape[0])
    s.light_prims = keep['light_prims'].ctypes.data if flat.light_prims.size else None
    s.has_env = 1 if flat.has_env else 0
    for i in range(3):
        s.env_Le[i] = float(flat.env_Le[i])
        s.cam_pos[i] = float(flat.camera.position[i])
        s.cam_center[i] = float(flat.camera.center[i])
        s.cam_up[i] = float(flat.camera.up[i])
    s.vfov = float(flat.camera.vfov)
    s.aspect = float(flat.camera.aspect)
    s.w = int(flat.w)
    s.h = int(flat.h)
    nprim = s.num_prims
    mat_ks = getattr(flat, 'mat_ks', None)
    mat_tex = getattr(flat, 'mat_tex', None)
    keep['mat_ks'] = np.ascontiguousarray(mat_ks if mat_ks is not None else np.zeros((nprim, 3)), dtype=np.float32)
    keep['mat_tex'] = np.ascontiguousarray(mat_tex if mat_tex is not None else np.full(nprim, -1), dtype=np.int32)
    s.mat_ks = keep['mat_ks'].ctypes.data
    s.mat_tex = keep['mat_tex'].ctypes.data
    if getattr(flat, 'textures', None):
        keep['uvs'] = np.ascontiguousarray(flat.uvs, dtype=np.float32)
        keep['tex_info'], keep['tex_data'] = flat.texture_tables()
        s.uvs = keep['uvs'].ctypes.data
        s.num_textures = len(flat.textures)
        s.tex_info = keep['tex_info'].ctypes.data
        s.tex_data = keep['tex_data'].ctypes.data
        s.tex_data_len = int(keep['tex_data'].size)
    return s, keep


def pt_render(flat, spp: int, max_verts: int, sampling_mode: int = 2, rr_depth: int = 5,
              seed: int = 0, nthreads: int = 0, accel: Accel = None, sample_offset: int = 0,
              sumsq: bool = True, primary_mode: str = 'pixel', num_samples: int = None):
    """renderer::pt restated in FP64. Returns dict with 'sum', 'sumsq'
    ([h,w,3] float64 unscaled), 'image' (the film value) and 'stats'.
    primary_mode is the reference's primary_ray_sampling_mode, 'pixel' or
    'image'. Pixel mode renders spp samples per pixel (image = sum / spp).
    Image mode renders num_samples samples over the image (default
    spp * w * h; image = sum * w*h / num_samples) and has no 'sumsq' (a
    sample splats into two pixels)."""
    if primary_mode not in ('pixel', 'image'):
        raise ValueError(f'primary_mode {primary_mode!r}')
    image = primary_mode == 'image'
    sumsq = sumsq and not image
    if accel is None:
        accel = Accel(flat.verts, nthreads)
    s, keep = make_orc_scene(flat)
    npix = flat.w * flat.h
    if num_samples is None:
        num_samples = spp * npix
    P = OrcPTParams(spp, max_verts, sampling_mode, rr_depth, seed, nthreads, sample_offset, int(image), 0,
                    int(num_samples) if image else 0)
    npix = flat.w * flat.h
    fsum = np.zeros(npix * 3, dtype=np.float64)
    fsq = np.zeros(npix * 3, dtype=np.float64) if sumsq else None
    st = OrcPTStats()
    rc = lib().orc_pt_render(ctypes.byref(s), accel.h, ctypes.byref(P), fsum,
                             fsq.ctypes.data if fsq is not None else None, ctypes.byref(st))
    if rc != 0:
        raise RuntimeError(f'orc_pt_render failed ({rc})')
    scale = npix / num_samples if image else 1.0 / spp
    out = {'sum': fsum.reshape(flat.h, flat.w, 3), 'image': fsum.reshape(flat.h, flat.w, 3) * scale,
           'stats': {'samples': st.samples, 'extension_rays': st.extension_rays,
                     'shadow_rays': st.shadow_rays, 'path_vertices': st.path_vertices,
                     'extension_node_visits': st.extension_node_visits,
                     'extension_tri_tests': st.extension_tri_tests,
                     'shadow_node_visits': st.shadow_node_visits, 'shadow_tri_tests': st.shadow_tri_tests}}
    if fsq is not None:
        out['sumsq'] = fsq.reshape(flat.h, flat.w, 3)
    return out


def lt_render(flat, num_samples: int, max_verts: int, seed: int = 0, pass_offset: int = 0, nthreads: int = 0,
              accel: Accel = None):
    """renderer::lt restated in FP64 (orc_lt_render). Returns dict with 'sum'
    ([h,w,3] float64, unscaled splat sums), 'image' (= sum * w*h /
    num_samples, the reference's film) and 'stats'."""
    if accel is None:
        accel = Accel(flat.verts, nthreads)
    s, keep = make_orc_scene(flat)
    P = OrcLTParams(int(num_samples), int(max_verts), int(pass_offset), int(seed), int(nthreads))
    film = np.zeros(flat.w * flat.h * 3, dtype=np.float64)
    st = OrcPTStats()
    rc = lib().orc_lt_render(ctypes.byref(s), accel.h, ctypes.byref(P), film, ctypes.byref(st))
    if rc != 0:
        raise RuntimeError(f'orc_lt_render failed ({rc})')
    f = film.reshape(flat.h, flat.w, 3)
    return {'sum': f, 'image': f * (flat.w * flat.h / num_samples),
            'stats': {'samples': st.samples, 'extension_rays': st.extension_rays,
                      'shadow_rays': st.shadow_rays, 'path_vertices': st.path_vertices}}


def bdpt_render(flat, num_samples: int, max_verts: int, seed: int = 0, pass_offset: int = 0, nthreads: int = 0,
                accel: Accel = None, min_verts: int = 2):
    """renderer::bdpt restated in FP64 (orc_bdpt_render; area lights only).
    Returns dict with 'sum' ([h,w,3] unscaled splat sums), 'image' (= sum *
    w*h / num_samples, the reference's film) and 'stats'."""
    if accel is None:
        accel = Accel(flat.verts, nthreads)
    s, keep = make_orc_scene(flat)
    P = OrcLTParams(int(num_samples), int(max_verts), int(pass_offset), int(seed), int(nthreads), int(min_verts))
    film = np.zeros(flat.w * flat.h * 3, dtype=np.float64)
    st = OrcPTStats()
    rc = lib().orc_bdpt_render(ctypes.byref(s), accel.h, ctypes.byref(P), film, ctypes.byref(st))
    if rc != 0:
        raise RuntimeError(f'orc_bdpt_render failed ({rc})')
    f = film.reshape(flat.h, flat.w, 3)
    return {'sum': f, 'image': f * (flat.w * flat.h / num_samples),
            'stats': {'samples': st.samples, 'extension_rays': st.extension_rays,
                      'shadow_rays': st.shadow_rays, 'path_vertices': st.path_vertices}}


def _mat_params(color=(1, 1, 1), ax=0.0, ay=0.0, ni=1.0, ks=(0, 0, 0), alpha=1.0):
    """Probe parameters: color rgb (Kd / Ks), ax, ay, ni, mixture Ks rgb, mixture alpha."""
    return np.array([color[0], color[1], color[2], ax, ay, ni, ks[0], ks[1], ks[2], alpha, 0, 0],
                    dtype=np.float64)


def mat_eval(mtype, params, normal, wi, wo):
    wi = np.ascontiguousarray(wi, dtype=np.float64).reshape(-1, 3)
    wo = np.ascontiguousarray(wo, dtype=np.float64).reshape(-1, 3)
    n = wi.shape[0]
    f = np.zeros((n, 3))
    pdf = np.zeros(n)
    lib().orc_mat_eval(int(mtype), np.ascontiguousarray(params, dtype=np.float64),
                       np.ascontiguousarray(normal, dtype=np.float64), n, wi, wo, f, pdf)
    return f, pdf


def mat_sample(mtype, params, normal, wi, u):
    wi = np.ascontiguousarray(wi, dtype=np.float64).reshape(-1, 3)
    u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1, 4)
    n = wi.shape[0]
    wo = np.zeros((n, 3))
    w = np.zeros((n, 3))
    comp = np.zeros(n, dtype=np.int32)
    ok = np.zeros(n, dtype=np.int32)
    lib().orc_mat_sample(int(mtype), np.ascontiguousarray(params, dtype=np.float64),
                         np.ascontiguousarray(normal, dtype=np.float64), n, wi, u, wo, w, comp, ok)
    return wo, w, comp, ok.astype(bool)
