"""Wavefront OBJ / MTL ingestion into a lm_amd.scenes.Scene (model::wavefrontobj,
src/model/model_wavefrontobj.cpp:84-227, with the OBJ / MTL parsing of
src/objloader/objloader_simple.cpp).

Faces are grouped into one primitive per (object / group, material) run as
the reference's mesh groups; polygons are triangulated as fans; position,
texture and normal indices may be negative (relative). Materials follow the
reference's mapping (model_wavefrontobj.cpp:155-215):
  illum 7          -> material::glass (Ni)
  illum 5          -> material::mirror
  (both black material::diffuse with skip_specular_mat)
  otherwise        -> material::mixture_wavefrontobj: diffuse Kd (map_Kd)
                      + glossy Ks with ax = r / as, ay = r * as,
                      r = 2 / (2 + Ns), as = sqrt(1 - 0.9 an), both >= 1e-3;
                      the map's alpha channel masks it unless
                      skip_specular_mat
With no_mixture (the reference's NO_MIXTURE_MATERIAL build switch) the last
case is material::diffuse when Ks == 0, else material::glossy. Ke > 0 makes
the group an area light (light::area).

map_Kd loads as texture::bitmap does through stbi_loadf
(texture_bitmap.cpp:50-66): 8-bit images become floats (c / 255)^2.2 on the
colour channels and c / 255 on alpha, rows flipped vertically (flip = true);
grey images are expanded to RGB (the reference reads 3 channels from every
texel, which for 1- and 2-channel images reads neighbouring texels). Loading
needs Pillow; a texture path that does not load is an error unless
`ignore_textures` is set (then Kd is used).
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import scenes


@dataclass
class MtlParams:
    name: str
    Kd: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    Ks: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    Ke: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    Ns: float = 0.0
    Ni: float = 1.0
    an: float = 0.0
    illum: int = 2
    map_Kd: str = ''


def _floats(tok, n):
    v = [float(x) for x in tok[:n]]
    while len(v) < n:
        v.append(v[-1] if v else 0.0)
    return tuple(v)


def load_mtl(path: str) -> Dict[str, MtlParams]:
    mats: Dict[str, MtlParams] = {}
    cur: Optional[MtlParams] = None
    with open(path) as f:
        for line in f:
            t = line.split('#', 1)[0].split()
            if not t:
                continue
            k, a = t[0], t[1:]
            if k == 'newmtl':
                cur = MtlParams(' '.join(a))
                mats[cur.name] = cur
            elif cur is None:
                continue
            elif k in ('Kd', 'Ks', 'Ke'):
                setattr(cur, k, _floats(a, 3))
            elif k == 'Ns':
                cur.Ns = float(a[0])
            elif k == 'Ni':
                cur.Ni = float(a[0])
            elif k in ('an', 'aniso'):
                cur.an = float(a[0])
            elif k == 'illum':
                cur.illum = int(float(a[0]))
            elif k == 'map_Kd':
                cur.map_Kd = a[-1]
    return mats


def load_texture(path: str, flip: bool = True) -> scenes.Texture:
    """texture::bitmap from an image file, with stbi_loadf's conversion of
    8-bit data (colour (c/255)^2.2, alpha c/255) and its vertical flip."""
    try:
        from PIL import Image
    except ImportError as e:   # pragma: no cover - Pillow is part of the environment
        raise NotImplementedError('loading texture images needs Pillow') from e
    with Image.open(path) as im:
        has_alpha = im.mode in ('RGBA', 'LA', 'PA') or (im.mode == 'P' and 'transparency' in im.info)
        im = im.convert('RGBA' if has_alpha else 'RGB')
        a = np.asarray(im, dtype=np.float64) / 255.0
    a = a.copy()
    a[..., :3] = a[..., :3] ** 2.2
    if flip:
        a = a[::-1]
    return scenes.Texture(np.ascontiguousarray(a, dtype=np.float32))


def material_of(m: MtlParams, ignore_textures: bool = False, texture: Optional[scenes.Texture] = None,
                no_mixture: bool = False, skip_specular_mat: bool = False) -> scenes.Material:
    """Reference mapping of MTL parameters to a device material
    (model_wavefrontobj.cpp:155-215); `texture` is the loaded map_Kd."""
    if m.map_Kd and texture is None and not ignore_textures:
        raise NotImplementedError(f'material {m.name!r}: map_Kd {m.map_Kd!r} was not loaded')
    if m.illum in (5, 7):
        if skip_specular_mat:
            return scenes.Material(scenes.MAT_DIFFUSE, (0.0, 0.0, 0.0))
        if m.illum == 7:
            return scenes.Material(scenes.MAT_GLASS, ni=float(m.Ni))
        return scenes.Material(scenes.MAT_MIRROR)
    r = 2.0 / (2.0 + m.Ns)
    a_s = math.sqrt(max(0.0, 1.0 - m.an * 0.9))
    ax, ay = max(1e-3, r / a_s), max(1e-3, r * a_s)
    if no_mixture:
        if max(m.Ks) == 0.0:
            return scenes.Material(scenes.MAT_DIFFUSE, tuple(m.Kd), texture=texture)
        return scenes.Material(scenes.MAT_GLOSSY, tuple(m.Ks), ax, ay)
    return scenes.Material(scenes.MAT_MIXTURE, tuple(m.Kd), ax, ay, ks=tuple(m.Ks), texture=texture,
                           alpha_mask=not skip_specular_mat)


@dataclass
class _Group:
    material: str
    # (position index, normal index or -1, texture coordinate index or -1)
    corners: List[Tuple[int, int, int]] = field(default_factory=list)


def load_obj(path: str, ignore_textures: bool = False, no_mixture: bool = False,
             skip_specular_mat: bool = False) -> List[scenes.Primitive]:
    """Primitives of an OBJ file (one per face group of one material). The
    options are model::wavefrontobj's skip_specular_mat and the
    NO_MIXTURE_MATERIAL switch; textures are shared per file name."""
    base = os.path.dirname(os.path.abspath(path))
    ps: List[Tuple[float, float, float]] = []
    ns: List[Tuple[float, float, float]] = []
    ts: List[Tuple[float, float]] = []
    mats: Dict[str, MtlParams] = {}
    groups: List[_Group] = []
    cur_mtl = ''
    cur: Optional[_Group] = None

    def index(s, count):
        i = int(s)
        return i - 1 if i > 0 else count + i

    with open(path) as f:
        for line in f:
            t = line.split('#', 1)[0].split()
            if not t:
                continue
            k, a = t[0], t[1:]
            if k == 'v':
                ps.append(_floats(a, 3))
            elif k == 'vn':
                ns.append(_floats(a, 3))
            elif k == 'vt':
                ts.append(_floats(a, 2))
            elif k == 'mtllib':
                mats.update(load_mtl(os.path.join(base, ' '.join(a))))
            elif k == 'usemtl':
                cur_mtl = ' '.join(a)
                cur = None
            elif k in ('o', 'g'):
                cur = None
            elif k == 'f':
                if len(a) < 3:
                    raise ValueError(f'{path}: face with fewer than 3 vertices')
                if cur is None:
                    cur = _Group(cur_mtl)
                    groups.append(cur)
                corner = []
                for v in a:
                    parts = v.split('/')
                    pi = index(parts[0], len(ps))
                    # a texture index in a file without vt lines is ignored
                    ti = index(parts[1], len(ts)) if len(parts) > 1 and parts[1] and ts else -1
                    ni = index(parts[2], len(ns)) if len(parts) > 2 and parts[2] else -1
                    if (not 0 <= pi < len(ps) or (ni != -1 and not 0 <= ni < len(ns))
                            or (ti != -1 and not 0 <= ti < len(ts))):
                        raise ValueError(f'{path}: face index out of range')
                    corner.append((pi, ni, ti))
                for j in range(1, len(corner) - 1):   # fan
                    cur.corners += [corner[0], corner[j], corner[j + 1]]
    P = np.asarray(ps, dtype=np.float64).reshape(-1, 3)
    N = np.asarray(ns, dtype=np.float64).reshape(-1, 3)
    TC = np.asarray(ts, dtype=np.float64).reshape(-1, 2)
    textures: Dict[str, scenes.Texture] = {}
    prims: List[scenes.Primitive] = []
    for g in groups:
        if not g.corners:
            continue
        c = np.asarray(g.corners, dtype=np.int64)
        tri = P[c[:, 0]].reshape(-1, 3, 3)
        gn = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
        ln = np.linalg.norm(gn, axis=1, keepdims=True)
        gn = np.where(ln > 0, gn / np.where(ln > 0, ln, 1), 0)
        nrm = np.repeat(gn, 3, axis=0)
        has = c[:, 1] >= 0
        if has.any():
            vn = N[c[has, 1]]
            vn = vn / np.maximum(np.linalg.norm(vn, axis=1, keepdims=True), 1e-30)
            nrm[has] = vn
        uv = np.zeros((c.shape[0], 2))
        has_t = c[:, 2] >= 0
        if has_t.any():
            uv[has_t] = TC[c[has_t, 2]]
        mesh = scenes.Mesh(tri.reshape(-1, 3).astype(np.float32), nrm.astype(np.float32),
                           np.arange(tri.shape[0] * 3, dtype=np.int32).reshape(-1, 3), uv.astype(np.float32))
        m = mats.get(g.material, MtlParams(g.material or 'default', Kd=(0.8, 0.8, 0.8)))
        tex = None
        if m.map_Kd and m.illum not in (5, 7):
            # texture_<stem>: one texture per file name (model_wavefrontobj.cpp:133-153)
            key = os.path.splitext(os.path.basename(m.map_Kd))[0]
            if key not in textures:
                try:
                    textures[key] = load_texture(os.path.join(base, m.map_Kd))
                except (OSError, NotImplementedError):
                    if not ignore_textures:
                        raise
            tex = textures.get(key)
        Ke = tuple(m.Ke) if max(m.Ke) > 0 else None
        prims.append(scenes.Primitive(mesh, material_of(m, ignore_textures, tex, no_mixture, skip_specular_mat),
                                      Ke))
    return prims


def scene_from_obj(path: str, camera: scenes.Camera, width: int, height: int,
                   env_Le: Optional[tuple] = None, ignore_textures: bool = False, no_mixture: bool = False,
                   skip_specular_mat: bool = False) -> scenes.Scene:
    return scenes.Scene(load_obj(path, ignore_textures, no_mixture, skip_specular_mat), camera, env_Le,
                        width, height)
