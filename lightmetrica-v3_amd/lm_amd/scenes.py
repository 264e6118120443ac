"""Seeded synthetic scenes and ray sets for the path-tracing hot path.

Every workload is generated here, from a master seed, as FP32 geometry with
identity transforms, so the GPU path and the FP64 CPU oracle see exactly the
same vertices (the oracle promotes them to double).

Scene semantics follow the reference's scene model:

* A scene is a list of primitives, each a mesh (positions, vertex normals,
  faces) with one material and optionally an area light that emits from the
  same mesh (``src/scene.cpp:72-129`` requires mesh+material together; an
  emitter is a primitive with mesh, material and ``light::area``).
* Triangles are flattened in primitive order, then face order, which is the
  order ``Accel_SAHBVH::build`` enumerates them
  (``src/accel/accel_sahbvh.cpp:132-151`` via ``Scene::traverse_primitive_nodes``
  and ``Mesh::foreach_triangle``). The global triangle index is the index in
  that order and maps back to (primitive, face).
* The light list is the area lights in primitive order, then the constant
  environment light (uniform light selection, ``src/scene.cpp:374-387``).
* Camera parameters are those of ``camera::pinhole``
  (``src/camera/camera_pinhole.cpp:76-97``): position, center, up, vertical
  FoV in degrees, aspect = width / height.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

MASTER_SEED = 0x5EED

# Material type codes shared with the C ABI (include/lmgpu.h).
MAT_DIFFUSE = 0   # material::diffuse  (Kd)
MAT_GLOSSY = 1    # material::glossy   (Ks, ax, ay)
MAT_MIRROR = 2    # material::mirror
MAT_GLASS = 3     # material::glass    (Ni)
MAT_MIXTURE = 4   # material::mixture_wavefrontobj (Kd [+ map_Kd] diffuse + Ks/ax/ay glossy, alpha mask)


def subseed(config: int, seed: int = MASTER_SEED) -> int:
    """Deterministic per-config seed derived from the master seed."""
    return int(np.random.SeedSequence([int(seed), int(config)]).generate_state(1)[0])


# ----------------------------------------------------------------------------
# Meshes
# ----------------------------------------------------------------------------

@dataclasses.dataclass
class Mesh:
    ps: np.ndarray   # [V,3] float32 positions
    ns: np.ndarray   # [V,3] float32 vertex normals
    fs: np.ndarray   # [F,3] int32 vertex indices (counter-clockwise = front)
    ts: Optional[np.ndarray] = None   # [V,2] float32 texture coordinates (Mesh::surface_point's t), or None

    @property
    def num_triangles(self) -> int:
        return int(self.fs.shape[0])

    def triangles(self) -> np.ndarray:
        """[F,9] float32 world-space vertices (p1,p2,p3) per face."""
        return self.ps[self.fs].reshape(-1, 9).astype(np.float32, copy=False)

    def corner_normals(self) -> np.ndarray:
        """[F,9] float32 vertex normals (n1,n2,n3) per face."""
        return self.ns[self.fs].reshape(-1, 9).astype(np.float32, copy=False)

    def corner_uvs(self) -> np.ndarray:
        """[F,6] float32 texture coordinates (t1,t2,t3) per face (zeros without ts)."""
        if self.ts is None:
            return np.zeros((self.num_triangles, 6), dtype=np.float32)
        return self.ts[self.fs].reshape(-1, 6).astype(np.float32, copy=False)


def _normalize(v: np.ndarray) -> np.ndarray:
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def flat_mesh(tris: np.ndarray) -> Mesh:
    """Mesh with one vertex per corner and vertex normals = geometric normal."""
    tris = np.asarray(tris, dtype=np.float64).reshape(-1, 3, 3)
    n = _normalize(np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0]))
    ps = tris.reshape(-1, 3).astype(np.float32)
    ns = np.repeat(n, 3, axis=0).astype(np.float32)
    fs = np.arange(ps.shape[0], dtype=np.int32).reshape(-1, 3)
    return Mesh(ps, ns, fs)


def quad(p0, p1, p2, p3, normal) -> np.ndarray:
    """Two triangles (p0,p1,p2),(p0,p2,p3), wound so their normal is `normal`."""
    p0, p1, p2, p3 = (np.asarray(p, dtype=np.float64) for p in (p0, p1, p2, p3))
    t = np.array([[p0, p1, p2], [p0, p2, p3]])
    if np.dot(np.cross(p1 - p0, p2 - p0), np.asarray(normal, dtype=np.float64)) < 0:
        t = np.array([[p0, p2, p1], [p0, p3, p2]])
    return t


def merge_meshes(meshes: List[Mesh]) -> Mesh:
    ps, ns, fs, off = [], [], [], 0
    for m in meshes:
        ps.append(m.ps)
        ns.append(m.ns)
        fs.append(m.fs + off)
        off += m.ps.shape[0]
    return Mesh(np.concatenate(ps), np.concatenate(ns), np.concatenate(fs).astype(np.int32))


def icosphere(levels: int):
    """Unit icosphere: 20*4^levels faces, counter-clockwise seen from outside."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0),
                  (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
                  (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)], dtype=np.float64)
    v = _normalize(v)
    f = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
                  (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
                  (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
                  (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)], dtype=np.int64)
    # Orient every base face outward.
    c = v[f].mean(axis=1)
    n = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    flip = np.einsum('ij,ij->i', n, c) < 0
    f[flip] = f[flip][:, [0, 2, 1]]
    for _ in range(levels):
        nv = v.shape[0]
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        key = e.min(axis=1) * nv + e.max(axis=1)
        uniq, inv = np.unique(key, return_inverse=True)
        mid = _normalize(v[uniq // nv] + v[uniq % nv])
        v = np.concatenate([v, mid])
        m = inv.reshape(3, -1) + nv
        a, b, cc = f[:, 0], f[:, 1], f[:, 2]
        ab, bc, ca = m[0], m[1], m[2]
        f = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1),
                            np.stack([cc, ca, bc], 1), np.stack([ab, bc, ca], 1)])
    return v, f


class GradientNoise:
    """3-D gradient noise (Perlin's improved noise: quintic fade, 12 edge
    gradients) over a seeded permutation table; fBm sums octaves with
    lacunarity 2 and gain 0.5, normalised by the amplitude sum to [-1,1]."""

    _GRAD = np.array([(1, 1, 0), (-1, 1, 0), (1, -1, 0), (-1, -1, 0),
                      (1, 0, 1), (-1, 0, 1), (1, 0, -1), (-1, 0, -1),
                      (0, 1, 1), (0, -1, 1), (0, 1, -1), (0, -1, -1)], dtype=np.float64)

    def __init__(self, rng: np.random.Generator):
        p = rng.permutation(256)
        self.perm = np.concatenate([p, p]).astype(np.int64)

    @staticmethod
    def _fade(t):
        return t * t * t * (t * (t * 6 - 15) + 10)

    def noise(self, p: np.ndarray) -> np.ndarray:
        pi = np.floor(p)
        pf = p - pi
        ii = pi.astype(np.int64) & 255
        u = self._fade(pf)
        perm = self.perm
        out = np.zeros(p.shape[0])
        acc = []
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    h = perm[perm[perm[ii[:, 0] + dx] + ii[:, 1] + dy] + ii[:, 2] + dz] % 12
                    g = self._GRAD[h]
                    d = pf - np.array([dx, dy, dz], dtype=np.float64)
                    acc.append(np.einsum('ij,ij->i', g, d))
        # trilinear blend with the quintic fade weights
        x0 = acc[0] + u[:, 0] * (acc[1] - acc[0])
        x1 = acc[2] + u[:, 0] * (acc[3] - acc[2])
        x2 = acc[4] + u[:, 0] * (acc[5] - acc[4])
        x3 = acc[6] + u[:, 0] * (acc[7] - acc[6])
        y0 = x0 + u[:, 1] * (x1 - x0)
        y1 = x2 + u[:, 1] * (x3 - x2)
        out = y0 + u[:, 2] * (y1 - y0)
        return out

    def fbm(self, p: np.ndarray, octaves: int = 6, base_freq: float = 2.0) -> np.ndarray:
        s = np.zeros(p.shape[0])
        amp, freq, norm = 1.0, base_freq, 0.0
        for _ in range(octaves):
            s += amp * self.noise(p * freq)
            norm += amp
            amp *= 0.5
            freq *= 2.0
        return s / norm


def vertex_normals(v: np.ndarray, f: np.ndarray) -> np.ndarray:
    """Area-weighted vertex normals (the face cross product summed per vertex)."""
    fn = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    n = np.zeros_like(v)
    for k in range(3):
        for c in range(3):
            n[:, c] += np.bincount(f[:, k], weights=fn[:, c], minlength=v.shape[0])
    return _normalize(n)


def displaced_icosphere(seed: int, levels: int = 8, amplitude: float = 0.1) -> Mesh:
    """Icosphere subdivided `levels` times, radius 1 + amplitude * fBm6(p)."""
    rng = np.random.default_rng(seed)
    v, f = icosphere(levels)
    noise = GradientNoise(rng)
    r = 1.0 + amplitude * noise.fbm(v)
    v = v * r[:, None]
    v32 = v.astype(np.float32)
    n = vertex_normals(v32.astype(np.float64), f)
    return Mesh(v32, n.astype(np.float32), f.astype(np.int32))


def clutter(seed: int, count: int, lo: np.ndarray, hi: np.ndarray,
            mu: float = -4.0, sigma: float = 1.0, max_edge: float = 1.0) -> Mesh:
    """`count` equilateral triangles with centroids uniform in [lo,hi], edge
    length LogNormal(mu, sigma) clamped at `max_edge`, uniformly random
    orientation; flat shaded."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(lo, hi, size=(count, 3))
    L = np.minimum(np.exp(rng.normal(mu, sigma, size=count)), max_edge)
    n = _normalize(rng.normal(size=(count, 3)))
    a = np.where(np.abs(n[:, :1]) < 0.9, np.array([[1.0, 0, 0]]), np.array([[0, 1.0, 0]]))
    t1 = _normalize(np.cross(n, a))
    t2 = np.cross(n, t1)
    phi = rng.uniform(0, 2 * np.pi, size=count)
    R = L / math.sqrt(3.0)
    tris = np.empty((count, 3, 3))
    for k in range(3):
        ang = phi + 2 * np.pi * k / 3
        tris[:, k] = c + R[:, None] * (np.cos(ang)[:, None] * t1 + np.sin(ang)[:, None] * t2)
    return flat_mesh(tris)


def box_room(open_front: bool = True, slit: Optional[float] = None) -> Mesh:
    """[-1,1]^3 room, faces wound to face inward; front (+z) face open.
    With `slit`, the left wall (x=-1) has a vertical slit of that width
    centred at z=0 through its full height."""
    tris = []
    tris.append(quad((-1, -1, -1), (1, -1, -1), (1, -1, 1), (-1, -1, 1), (0, 1, 0)))   # floor
    tris.append(quad((-1, 1, -1), (1, 1, -1), (1, 1, 1), (-1, 1, 1), (0, -1, 0)))      # ceiling
    tris.append(quad((-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), (0, 0, 1)))   # back
    tris.append(quad((1, -1, -1), (1, 1, -1), (1, 1, 1), (1, -1, 1), (-1, 0, 0)))      # right
    if slit is None:
        tris.append(quad((-1, -1, -1), (-1, 1, -1), (-1, 1, 1), (-1, -1, 1), (1, 0, 0)))  # left
    else:
        h = slit / 2
        tris.append(quad((-1, -1, -1), (-1, 1, -1), (-1, 1, -h), (-1, -1, -h), (1, 0, 0)))
        tris.append(quad((-1, -1, h), (-1, 1, h), (-1, 1, 1), (-1, -1, 1), (1, 0, 0)))
    if not open_front:
        tris.append(quad((-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1), (0, 0, -1)))
    return flat_mesh(np.concatenate(tris))


def cube(size: float, angle: float, center) -> Mesh:
    """Axis-aligned cube of edge `size` rotated by `angle` about +y, outward faces."""
    h = 0.5
    faces = [
        quad((h, -h, -h), (h, h, -h), (h, h, h), (h, -h, h), (1, 0, 0)),
        quad((-h, -h, -h), (-h, h, -h), (-h, h, h), (-h, -h, h), (-1, 0, 0)),
        quad((-h, h, -h), (h, h, -h), (h, h, h), (-h, h, h), (0, 1, 0)),
        quad((-h, -h, -h), (h, -h, -h), (h, -h, h), (-h, -h, h), (0, -1, 0)),
        quad((-h, -h, h), (h, -h, h), (h, h, h), (-h, h, h), (0, 0, 1)),
        quad((-h, -h, -h), (h, -h, -h), (h, h, -h), (-h, h, -h), (0, 0, -1)),
    ]
    t = np.concatenate(faces) * size
    ca, sa = math.cos(angle), math.sin(angle)
    R = np.array([[ca, 0, sa], [0, 1, 0], [-sa, 0, ca]])
    t = t @ R.T + np.asarray(center, dtype=np.float64)
    return flat_mesh(t)


# ----------------------------------------------------------------------------
# Scene
# ----------------------------------------------------------------------------

@dataclasses.dataclass(eq=False)
class Texture:
    """texture::bitmap: texels [h, w, c] (c = 3 or 4) as the reference holds
    them after loading (linear floats, row 0 = the first row after its
    vertical flip); looked up nearest-texel with wrapping."""
    data: np.ndarray


@dataclasses.dataclass
class Material:
    type: int
    color: tuple = (1.0, 1.0, 1.0)   # Kd (diffuse, mixture) or Ks (glossy)
    ax: float = 0.0
    ay: float = 0.0
    ni: float = 1.0
    ks: tuple = (0.0, 0.0, 0.0)      # mixture: Ks of the glossy lobe
    texture: Optional[Texture] = None   # map_Kd of diffuse / mixture (takes priority over color)
    alpha_mask: bool = True          # mixture: use the texture's alpha channel as a mask


@dataclasses.dataclass
class Primitive:
    mesh: Mesh
    material: Material
    Ke: Optional[tuple] = None       # light::area radiance, or None


@dataclasses.dataclass
class Camera:
    position: tuple
    center: tuple
    up: tuple
    vfov: float
    aspect: float


@dataclasses.dataclass
class Scene:
    primitives: List[Primitive]
    camera: Camera
    env_Le: Optional[tuple] = None   # light::envconst radiance, or None
    w: int = 256
    h: int = 256

    def flatten(self) -> "FlatScene":
        tris, norms, prim, face, uvs = [], [], [], [], []
        for pi, p in enumerate(self.primitives):
            tris.append(p.mesh.triangles())
            norms.append(p.mesh.corner_normals())
            uvs.append(p.mesh.corner_uvs())
            nf = p.mesh.num_triangles
            prim.append(np.full(nf, pi, dtype=np.int32))
            face.append(np.arange(nf, dtype=np.int32))
        nprim = len(self.primitives)
        mat_type = np.array([p.material.type for p in self.primitives], dtype=np.int32)
        mat_param = np.zeros((nprim, 8), dtype=np.float32)
        mat_ks = np.zeros((nprim, 3), dtype=np.float32)
        mat_tex = np.full(nprim, -1, dtype=np.int32)
        textures, tex_index = [], {}
        for i, p in enumerate(self.primitives):
            m = p.material
            mat_param[i, 0:3] = m.color
            mat_param[i, 3] = m.ax
            mat_param[i, 4] = m.ay
            mat_param[i, 5] = m.ni
            mat_param[i, 6] = 1.0 if m.alpha_mask else 0.0
            mat_ks[i] = m.ks
            if m.texture is not None:
                if id(m.texture) not in tex_index:
                    tex_index[id(m.texture)] = len(textures)
                    textures.append(np.ascontiguousarray(m.texture.data, dtype=np.float32))
                mat_tex[i] = tex_index[id(m.texture)]
        prim_Ke = np.zeros((nprim, 3), dtype=np.float32)
        light_prims = []
        for i, p in enumerate(self.primitives):
            if p.Ke is not None:
                prim_Ke[i] = p.Ke
                light_prims.append(i)
        return FlatScene(
            verts=np.ascontiguousarray(np.concatenate(tris), dtype=np.float32),
            normals=np.ascontiguousarray(np.concatenate(norms), dtype=np.float32),
            tri_prim=np.ascontiguousarray(np.concatenate(prim)),
            tri_face=np.ascontiguousarray(np.concatenate(face)),
            mat_type=mat_type, mat_param=mat_param, prim_Ke=prim_Ke,
            light_prims=np.array(light_prims, dtype=np.int32),
            env_Le=np.array(self.env_Le if self.env_Le is not None else (0, 0, 0), dtype=np.float32),
            has_env=self.env_Le is not None,
            camera=self.camera, w=self.w, h=self.h,
            mat_ks=mat_ks, mat_tex=mat_tex,
            uvs=np.ascontiguousarray(np.concatenate(uvs), dtype=np.float32) if textures else None,
            textures=textures)


@dataclasses.dataclass
class FlatScene:
    verts: np.ndarray       # [N,9] float32
    normals: np.ndarray     # [N,9] float32
    tri_prim: np.ndarray    # [N] int32 primitive (scene node) index
    tri_face: np.ndarray    # [N] int32 face index within the primitive's mesh
    mat_type: np.ndarray    # [P] int32
    mat_param: np.ndarray   # [P,8] float32: color rgb, ax, ay, ni, alpha mask, pad
    prim_Ke: np.ndarray     # [P,3] float32
    light_prims: np.ndarray # [L] int32 primitives carrying light::area, in order
    env_Le: np.ndarray      # [3] float32
    has_env: bool
    camera: Camera
    w: int
    h: int
    mat_ks: Optional[np.ndarray] = None    # [P,3] float32 mixture Ks
    mat_tex: Optional[np.ndarray] = None   # [P] int32 texture of Kd, -1 none
    uvs: Optional[np.ndarray] = None       # [N,6] float32 texture coordinates (textured scenes)
    textures: list = dataclasses.field(default_factory=list)   # [h,w,c] float32 texel arrays

    def texture_tables(self):
        """(tex_info [T,4] int32: width, height, channels, offset; tex_data
        float32) of the scene's textures."""
        info = np.zeros((len(self.textures), 4), dtype=np.int32)
        off = 0
        for t, a in enumerate(self.textures):
            if a.ndim != 3 or a.shape[2] not in (3, 4):
                raise ValueError('a texture must be [h, w, 3] or [h, w, 4]')
            info[t] = (a.shape[1], a.shape[0], a.shape[2], off)
            off += a.size
        data = (np.concatenate([a.reshape(-1) for a in self.textures]).astype(np.float32)
                if self.textures else np.zeros(1, dtype=np.float32))
        return info, data

    @property
    def num_triangles(self) -> int:
        return int(self.verts.shape[0])

    @property
    def num_lights(self) -> int:
        return int(self.light_prims.shape[0]) + (1 if self.has_env else 0)


# ----------------------------------------------------------------------------
# Camera rays (camera::pinhole::primary_ray, src/camera/camera_pinhole.cpp:128-132)
# ----------------------------------------------------------------------------

def camera_basis(cam: Camera):
    pos = np.asarray(cam.position, dtype=np.float64)
    w = _normalize(pos - np.asarray(cam.center, dtype=np.float64))
    u = _normalize(np.cross(np.asarray(cam.up, dtype=np.float64), w))
    v = np.cross(w, u)
    tf = math.tan(cam.vfov * math.pi / 180.0 * 0.5)
    return pos, u, v, w, tf


def primary_rays(cam: Camera, w: int, h: int):
    """Pixel-centre primaries in FP64, returned as FP32 (o,d) arrays [w*h,3],
    row-major with y=0 the bottom row (renderer::raycast, renderer_raycast.cpp:70-73)."""
    pos, u, v, wv, tf = camera_basis(cam)
    ys, xs = np.meshgrid(np.arange(h), np.arange(w), indexing='ij')
    rx = 2.0 * (xs.reshape(-1) + 0.5) / w - 1.0
    ry = 2.0 * (ys.reshape(-1) + 0.5) / h - 1.0
    d = np.stack([cam.aspect * tf * rx, tf * ry, -np.ones_like(rx)], 1)
    d = _normalize(d)
    dw = d[:, :1] * u + d[:, 1:2] * v + d[:, 2:3] * wv
    o = np.broadcast_to(pos, dw.shape)
    return o.astype(np.float32), dw.astype(np.float32)


# ----------------------------------------------------------------------------
# Configurations
# ----------------------------------------------------------------------------

def _albedo(rng):
    return tuple(float(x) for x in rng.uniform(0.2, 0.8, size=3))


def _two_cubes(rng):
    cubes = []
    for sx in (-1.0, 1.0):
        s = float(rng.uniform(0.3, 0.7))
        ang = float(rng.uniform(0.0, 2 * np.pi))
        z = float(rng.uniform(-0.45, 0.45))
        cubes.append(cube(s, ang, (0.5 * sx, -1.0 + s / 2, z)))
    return cubes


def config0(seed: int = MASTER_SEED) -> Scene:
    """Box room (10 tris, front open) + two seeded cubes, Lambertian albedo
    U[0.2,0.8]^3 per object, one 0.5x0.5 ceiling quad emitter of radiance 17,
    pinhole camera fov 40 at 256x256."""
    rng = np.random.default_rng(subseed(0, seed))
    room = box_room()
    cubes = _two_cubes(rng)
    prims = [Primitive(room, Material(MAT_DIFFUSE, _albedo(rng)))]
    for c in cubes:
        prims.append(Primitive(c, Material(MAT_DIFFUSE, _albedo(rng))))
    y = 1.0 - 1e-3
    emitter = flat_mesh(quad((-.25, y, -.25), (.25, y, -.25), (.25, y, .25), (-.25, y, .25), (0, -1, 0)))
    prims.append(Primitive(emitter, Material(MAT_DIFFUSE, (0.78, 0.78, 0.78)), Ke=(17.0, 17.0, 17.0)))
    cam = Camera((0, 0, 3.5), (0, 0, 0), (0, 1, 0), 40.0, 1.0)
    return Scene(prims, cam, None, 256, 256)


def config1_mesh(seed: int = MASTER_SEED, levels: int = 8, clutter_count: int = 10000):
    """Displaced icosphere + lognormal clutter in its bounding box."""
    s = subseed(1, seed)
    sphere = displaced_icosphere(s, levels)
    lo = sphere.ps.min(axis=0).astype(np.float64)
    hi = sphere.ps.max(axis=0).astype(np.float64)
    cl = clutter(s + 1, clutter_count, lo, hi)
    return sphere, cl


def config1(seed: int = MASTER_SEED, levels: int = 8, clutter_count: int = 10000,
            width: int = 1920, height: int = 1080, n_incoherent: int = 2_000_000):
    """Closest-hit ray-cast microbench: returns (scene, rays) where rays has
    'coherent' and 'incoherent' entries of FP32 (o, d, tmin, tmax)."""
    sphere, cl = config1_mesh(seed, levels, clutter_count)
    prims = [Primitive(sphere, Material(MAT_DIFFUSE, (0.5, 0.5, 0.5))),
             Primitive(cl, Material(MAT_DIFFUSE, (0.5, 0.5, 0.5)))]
    cam = Camera((0, 0, 3), (0, 0, 0), (0, 1, 0), 40.0, width / height)
    scene = Scene(prims, cam, None, width, height)
    o, d = primary_rays(cam, width, height)
    coherent = make_rays(o, d)
    incoherent = surface_rays(sphere, n_incoherent, subseed(101, seed))
    return scene, {'coherent': coherent, 'incoherent': incoherent}


def make_rays(o, d, tmin=1e-3, tmax=1e10):
    n = o.shape[0]
    return {'o': np.ascontiguousarray(o, dtype=np.float32),
            'd': np.ascontiguousarray(d, dtype=np.float32),
            'tmin': np.full(n, tmin, dtype=np.float32),
            'tmax': np.full(n, tmax, dtype=np.float32)}


def surface_rays(mesh: Mesh, n: int, seed: int):
    """Origins uniform by area on `mesh`, directions uniform over the
    hemisphere of the face's geometric normal; tmin = 1e-3 (the reference Eps)."""
    rng = np.random.default_rng(seed)
    t = mesh.ps[mesh.fs].astype(np.float64)
    cr = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    area = 0.5 * np.linalg.norm(cr, axis=1)
    cdf = np.cumsum(area)
    cdf /= cdf[-1]
    fi = np.minimum(np.searchsorted(cdf, rng.random(n), side='right'), len(cdf) - 1)
    u0 = rng.random(n)
    u1 = rng.random(n)
    s = np.sqrt(u0)
    # barycentric (1-s, u1*s) as in light_area.cpp:49-54
    wb = (1 - s)[:, None]
    wc = (u1 * s)[:, None]
    p = t[fi, 0] * (1 - wb - wc) + t[fi, 1] * wb + t[fi, 2] * wc
    gn = _normalize(cr[fi])
    dd = _normalize(rng.normal(size=(n, 3)))
    flip = np.einsum('ij,ij->i', dd, gn) < 0
    dd[flip] = -dd[flip]
    return make_rays(p.astype(np.float32), dd.astype(np.float32))


_MATERIAL_SET = (MAT_DIFFUSE, MAT_GLOSSY, MAT_MIRROR, MAT_GLASS)


def _material_from(rng, mtype):
    if mtype == MAT_DIFFUSE:
        return Material(MAT_DIFFUSE, _albedo(rng))
    if mtype == MAT_GLOSSY:
        a = float(rng.uniform(0.05, 0.5))
        return Material(MAT_GLOSSY, _albedo(rng), a, a)
    if mtype == MAT_MIRROR:
        return Material(MAT_MIRROR)
    return Material(MAT_GLASS, ni=1.5)


def scaled(mesh: Mesh, scale: float, offset) -> Mesh:
    ps = (mesh.ps.astype(np.float64) * scale + np.asarray(offset, dtype=np.float64)).astype(np.float32)
    return Mesh(ps, mesh.ns.copy(), mesh.fs.copy())


def config2(seed: int = MASTER_SEED, levels: int = 8, clutter_count: int = 10000,
            width: int = 1920, height: int = 1080) -> Scene:
    """The config-1 mesh (scaled to radius 0.5, lifted to y=0.2) inside the
    config-0 room with its two cubes. Room walls are Lambertian; the sphere,
    the clutter and the two cubes take a seeded permutation of {Lambertian,
    GGX alpha~U[0.05,0.5], mirror, dielectric Ni=1.5} so each appears once.
    Four area lights (two ceiling quads, one on each side wall) of radiance
    U[5,20] plus a constant environment of 0.1."""
    rng = np.random.default_rng(subseed(2, seed))
    room = box_room()
    cubes = _two_cubes(rng)
    sphere, cl = config1_mesh(seed, levels, clutter_count)
    off = (0.0, 0.2, 0.0)
    objs = [scaled(sphere, 0.5, off), scaled(cl, 0.5, off)] + cubes
    perm = rng.permutation(len(_MATERIAL_SET))
    prims = [Primitive(room, Material(MAT_DIFFUSE, _albedo(rng)))]
    for o, k in zip(objs, perm):
        prims.append(Primitive(o, _material_from(rng, _MATERIAL_SET[k])))
    y = 1.0 - 1e-3
    x = 1.0 - 1e-3
    quads = [
        quad((-.6, y, -.6), (-.2, y, -.6), (-.2, y, -.2), (-.6, y, -.2), (0, -1, 0)),
        quad((.2, y, .2), (.6, y, .2), (.6, y, .6), (.2, y, .6), (0, -1, 0)),
        quad((-x, .2, -.5), (-x, .6, -.5), (-x, .6, -.1), (-x, .2, -.1), (1, 0, 0)),
        quad((x, .2, .1), (x, .6, .1), (x, .6, .5), (x, .2, .5), (-1, 0, 0)),
    ]
    for q in quads:
        ke = float(rng.uniform(5.0, 20.0))
        prims.append(Primitive(flat_mesh(q), Material(MAT_DIFFUSE, (0.78, 0.78, 0.78)), Ke=(ke, ke, ke)))
    cam = Camera((0, 0, 3.5), (0, 0, 0), (0, 1, 0), 40.0, width / height)
    return Scene(prims, cam, (0.1, 0.1, 0.1), width, height)


def config4(seed: int = MASTER_SEED, clutter_count: int = 200_000,
            width: int = 1920, height: int = 1080) -> Scene:
    """Room lit only through a 0.3-wide slit in the left wall by an outside
    area light of radiance 50 facing the slit; GGX floor (alpha 0.1);
    seeded clutter inside the room; no environment light."""
    rng = np.random.default_rng(subseed(4, seed))
    walls = box_room(slit=0.3)
    wt = walls.triangles().reshape(-1, 3, 3)
    floor_mask = np.all(np.abs(wt[:, :, 1] + 1.0) < 1e-6, axis=1)
    floor = flat_mesh(wt[floor_mask].astype(np.float64))
    rest = flat_mesh(wt[~floor_mask].astype(np.float64))
    cl = clutter(subseed(41, seed), clutter_count, np.array([-0.9, -0.95, -0.9]),
                 np.array([0.9, 0.3, 0.6]), mu=-4.0, sigma=1.0, max_edge=0.2)
    prims = [
        Primitive(floor, Material(MAT_GLOSSY, (0.8, 0.8, 0.8), 0.1, 0.1)),
        Primitive(rest, Material(MAT_DIFFUSE, _albedo(rng))),
        Primitive(cl, Material(MAT_DIFFUSE, _albedo(rng))),
    ]
    xl = -1.3
    light = flat_mesh(quad((xl, -.4, -.3), (xl, .8, -.3), (xl, .8, .3), (xl, -.4, .3), (1, 0, 0)))
    prims.append(Primitive(light, Material(MAT_DIFFUSE, (0.0, 0.0, 0.0)), Ke=(50.0, 50.0, 50.0)))
    cam = Camera((0, 0, 3.5), (0, 0, 0), (0, 1, 0), 40.0, width / height)
    return Scene(prims, cam, None, width, height)


def config3_jobs(seed: int = MASTER_SEED, n_jobs: int = 64, scale: float = 1.0):
    """Batched-render job list: per job, triangle count log-uniform in
    [1e4, 2e6] * scale, square resolution log-uniform in [256, 1024] * sqrt(scale),
    spp log-uniform in [16, 256]. Returns a list of dicts (the scene of each
    job is built by `config3_scene`)."""
    rng = np.random.default_rng(subseed(3, seed))
    jobs = []
    for j in range(n_jobs):
        tris = int(math.exp(rng.uniform(math.log(1e4), math.log(2e6))) * scale)
        res = int(math.exp(rng.uniform(math.log(256), math.log(1024))) * math.sqrt(scale))
        spp = int(round(math.exp(rng.uniform(math.log(16), math.log(256)))))
        jobs.append({'job': j, 'tris': max(tris, 64), 'res': max(res, 8), 'spp': max(spp, 1),
                     'seed': subseed(300 + j, seed)})
    return jobs


def config3_scene(job: dict) -> Scene:
    """Scene of one batched job: displaced icosphere with clutter sized to
    `job['tris']` (subdivision level chosen so the sphere holds ~90% of the
    budget), in the config-2 room/material/light mixture."""
    target = job['tris']
    level = max(1, min(8, int(round(math.log(max(target * 0.9, 80) / 20.0, 4)))))
    nsphere = 20 * 4 ** level
    ncl = max(16, target - nsphere)
    rng = np.random.default_rng(job['seed'])
    sphere = displaced_icosphere(job['seed'], level)
    lo = sphere.ps.min(axis=0).astype(np.float64)
    hi = sphere.ps.max(axis=0).astype(np.float64)
    cl = clutter(job['seed'] + 1, ncl, lo, hi)
    off = (0.0, 0.2, 0.0)
    room = box_room()
    perm = rng.permutation(len(_MATERIAL_SET))
    prims = [Primitive(room, Material(MAT_DIFFUSE, _albedo(rng))),
             Primitive(scaled(sphere, 0.5, off), _material_from(rng, _MATERIAL_SET[perm[0]])),
             Primitive(scaled(cl, 0.5, off), _material_from(rng, _MATERIAL_SET[perm[1]]))]
    y = 1.0 - 1e-3
    for q in (quad((-.6, y, -.6), (-.2, y, -.6), (-.2, y, -.2), (-.6, y, -.2), (0, -1, 0)),
              quad((.2, y, .2), (.6, y, .2), (.6, y, .6), (.2, y, .6), (0, -1, 0))):
        ke = float(rng.uniform(5.0, 20.0))
        prims.append(Primitive(flat_mesh(q), Material(MAT_DIFFUSE, (0.78, 0.78, 0.78)), Ke=(ke, ke, ke)))
    res = job['res']
    cam = Camera((0, 0, 3.5), (0, 0, 0), (0, 1, 0), 40.0, 1.0)
    return Scene(prims, cam, (0.1, 0.1, 0.1), res, res)
