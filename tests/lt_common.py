"""Scenes and statistics shared by the renderer::lt tests.

Light tracing with NEE to the camera (renderer_lt.cpp:56-170) estimates every
path of 3 to max_verts vertices whose camera-side vertex is not specular; the
two-vertex paths (an emitter seen directly) are not sampled. For a scene whose
camera sees only non-specular surfaces, the expected LT image therefore equals
PT(max_verts) - PT(2) pixel by pixel; with the same seed, PT(2) holds exactly
the direct-emission part of each PT sample, so the difference carries no
extra noise from it."""
import math

import numpy as np

from lm_amd import scenes


def furnace_scene(rho=0.5, Le=1.0, w=32, h=32):
    room = scenes.box_room(open_front=False)
    prim = scenes.Primitive(room, scenes.Material(scenes.MAT_DIFFUSE, (rho, rho, rho)), Ke=(Le, Le, Le))
    cam = scenes.Camera((0, 0, 0.5), (0, 0, -1), (0, 1, 0), 60.0, w / h)
    return scenes.Scene([prim], cam, None, w, h)


def furnace_expected(rho, Le, max_verts):
    """LT furnace: light paths of k = 1 .. max_verts - 2 reflections."""
    return Le * sum(rho ** k for k in range(1, max_verts - 1))


def env_cube_scene(rho=(0.6, 0.4, 0.2), Le=(1.0, 1.0, 1.0), w=48, h=48):
    c = scenes.cube(1.6, 0.0, (0, 0, 0))
    cam = scenes.Camera((0, 0, 4.0), (0, 0, 0), (0, 1, 0), 30.0, w / h)
    return scenes.Scene([scenes.Primitive(c, scenes.Material(scenes.MAT_DIFFUSE, rho))], cam, Le, w, h)


def glass_lamp_scene(w=24, h=24):
    """A diffuse floor lit by an emitter sealed inside a glass cube: every
    light path leaves the glass through one refraction, so a wrong eta^2 on
    importance-transport refraction scales the image by 2.25. The camera
    sees only the floor."""
    floor = scenes.flat_mesh(scenes.quad((-3, -1, -3), (3, -1, -3), (3, -1, 3), (-3, -1, 3), (0, 1, 0)))
    box = scenes.cube(0.6, 0.3, (0, 0.2, 0))
    y = 0.1
    lamp = scenes.flat_mesh(scenes.quad((-.2, y, -.2), (.2, y, -.2), (.2, y, .2), (-.2, y, .2), (0, -1, 0)))
    prims = [scenes.Primitive(floor, scenes.Material(scenes.MAT_DIFFUSE, (0.7, 0.6, 0.5))),
             scenes.Primitive(box, scenes.Material(scenes.MAT_GLASS, ni=1.5)),
             scenes.Primitive(lamp, scenes.Material(scenes.MAT_DIFFUSE, (0.0, 0.0, 0.0)), Ke=(10.0, 10.0, 10.0))]
    cam = scenes.Camera((0, -0.6, 2.0), (0, -1.0, 0.0), (0, 1, 0), 40.0, w / h)
    return scenes.Scene(prims, cam, None, w, h)


def smooth_sphere_scene(w=32, h=32):
    """Box room with a ceiling emitter and a coarse icosphere with
    interpolated vertex normals: shading and geometric normals differ, so the
    LE shading-normal correction matters."""
    room = scenes.box_room(open_front=True)
    v, f = scenes.icosphere(1)
    v = v * 0.5 + np.array([0.0, -0.45, 0.0])
    mesh = scenes.Mesh(v.astype(np.float32), scenes.vertex_normals(v, f).astype(np.float32), f.astype(np.int32))
    y = 1.0 - 1e-3
    lamp = scenes.flat_mesh(scenes.quad((-.3, y, -.3), (.3, y, -.3), (.3, y, .3), (-.3, y, .3), (0, -1, 0)))
    prims = [scenes.Primitive(room, scenes.Material(scenes.MAT_DIFFUSE, (0.6, 0.6, 0.6))),
             scenes.Primitive(mesh, scenes.Material(scenes.MAT_DIFFUSE, (0.8, 0.5, 0.3))),
             scenes.Primitive(lamp, scenes.Material(scenes.MAT_DIFFUSE, (0.5, 0.5, 0.5)), Ke=(12.0, 12.0, 12.0))]
    cam = scenes.Camera((0, 0, 3.5), (0, -0.3, 0), (0, 1, 0), 35.0, w / h)
    return scenes.Scene(prims, cam, None, w, h)


def block_z(lt_batches, pt_img, pt_sumsq, pt_spp, direct_img, block=4):
    """z-scores of block means of the LT image (mean over independent batches
    `lt_batches`, [K,h,w,3] reference-scaled films) against PT(M) - PT(2)
    (`pt_img` - `direct_img`), over pixels without direct emission. Returns
    (z of blocks with a nonzero standard error, number of such blocks)."""
    lt_batches = np.asarray(lt_batches)
    K = lt_batches.shape[0]
    ref = pt_img - direct_img
    pt_se2 = np.maximum(pt_sumsq / pt_spp - pt_img ** 2, 0) / pt_spp
    mask = (direct_img == 0).all(-1)[..., None] * np.ones(3)
    h, w = ref.shape[:2]
    bh, bw = h // block, w // block

    def bsum(a):
        return (a * mask)[..., :bh * block, :bw * block, :].reshape(a.shape[:-3] + (bh, block, bw, block, 3)).sum(
            (-4, -2))

    # one light path splats into several pixels, so the pixels of an LT batch
    # are correlated: the block sums' variance comes from the batches
    lt_blocks = bsum(lt_batches)
    cnt = bsum(np.ones_like(ref))
    diff = lt_blocks.mean(0) - bsum(ref)
    se = np.sqrt(lt_blocks.var(0, ddof=1) / K + bsum(pt_se2))
    ok = (cnt > 0) & (se > 0)
    z = diff[ok] / se[ok]
    return z, int(ok.sum())


def batch_mean_z(values, expected):
    v = np.asarray(values, dtype=np.float64)
    se = v.std(ddof=1) / math.sqrt(len(v))
    return (v.mean() - expected) / se


def checker_texture(n=4, alpha=False, seed=0):
    """n x n texels of seeded colours in [0.1, 0.9]; with `alpha`, a fourth
    channel of alternating 0 / 1 (a binary mask)."""
    rng = np.random.default_rng(seed)
    rgb = rng.uniform(0.1, 0.9, size=(n, n, 3))
    if not alpha:
        return scenes.Texture(rgb.astype(np.float32))
    a = ((np.arange(n)[:, None] + np.arange(n)[None, :]) % 2).astype(np.float64)
    return scenes.Texture(np.concatenate([rgb, a[..., None]], -1).astype(np.float32))


def textured_quad(tex, mtype=scenes.MAT_MIXTURE, uv_scale=1.0, z=0.0):
    """Unit square [-1,1]^2 at depth z facing +z, texture coordinates
    (x+1)/2 * uv_scale, (y+1)/2 * uv_scale."""
    ps = np.array([[-1, -1, z], [1, -1, z], [1, 1, z], [-1, 1, z]], dtype=np.float32)
    ns = np.tile(np.array([[0, 0, 1]], dtype=np.float32), (4, 1))
    fs = np.array([[0, 1, 2], [0, 2, 3]], dtype=np.int32)
    ts = ((ps[:, :2] + 1) / 2 * uv_scale).astype(np.float32)
    mat = scenes.Material(mtype, (1.0, 1.0, 1.0), 0.3, 0.3, ks=(0.0, 0.0, 0.0), texture=tex)
    return scenes.Primitive(scenes.Mesh(ps, ns, fs, ts), mat)


def masked_quad_scene(w=32, h=32, uv_scale=1.0):
    """A textured mixture quad (Ks = 0, binary alpha mask) filling the view,
    under a constant environment Le = 1: where alpha = 1 the pixel is the
    texel's Kd (a one-sided diffuse plane under a constant environment
    reflects rho Le), where alpha = 0 the environment behind (1)."""
    tex = checker_texture(4, alpha=True, seed=5)
    cam = scenes.Camera((0, 0, 2.0), (0, 0, 0), (0, 1, 0), 40.0, w / h)
    return scenes.Scene([textured_quad(tex, uv_scale=uv_scale)], cam, (1.0, 1.0, 1.0), w, h), tex


def mixture_room_scene(w=24, h=24):
    """Box room with mixture walls (Kd + Ks), a textured mixture floor panel,
    a glossy-heavy mixture cube and a ceiling emitter."""
    room = scenes.box_room(open_front=True)
    tex = checker_texture(8, seed=9)
    floor = textured_quad(tex, uv_scale=2.0)
    # lay the panel on the floor (y = -1 + 1e-3), facing +y
    fl = floor.mesh
    ps = np.stack([fl.ps[:, 0] * 0.8, np.full(4, -1 + 1e-3, dtype=np.float32), -fl.ps[:, 1] * 0.8], -1)
    floor = scenes.Primitive(scenes.Mesh(ps.astype(np.float32), np.tile(np.array([[0, 1, 0]], np.float32), (4, 1)),
                                         fl.fs, fl.ts),
                             scenes.Material(scenes.MAT_MIXTURE, (1, 1, 1), 0.2, 0.2, ks=(0.2, 0.2, 0.2), texture=tex))
    cube = scenes.cube(0.6, 0.5, (0.35, -0.7, -0.2))
    y = 1.0 - 1e-3
    lamp = scenes.flat_mesh(scenes.quad((-.3, y, -.3), (.3, y, -.3), (.3, y, .3), (-.3, y, .3), (0, -1, 0)))
    prims = [scenes.Primitive(room, scenes.Material(scenes.MAT_MIXTURE, (0.5, 0.45, 0.4), 0.3, 0.3,
                                                    ks=(0.15, 0.15, 0.15))),
             floor,
             scenes.Primitive(cube, scenes.Material(scenes.MAT_MIXTURE, (0.1, 0.2, 0.3), 0.08, 0.15,
                                                    ks=(0.6, 0.5, 0.4))),
             scenes.Primitive(lamp, scenes.Material(scenes.MAT_DIFFUSE, (0.5, 0.5, 0.5)), Ke=(12.0, 12.0, 12.0))]
    cam = scenes.Camera((0, 0, 3.5), (0, -0.2, 0), (0, 1, 0), 38.0, w / h)
    return scenes.Scene(prims, cam, None, w, h)


def masked_quad_expected(flat, tex, uv_scale=1.0):
    """Per-pixel expected radiance of masked_quad_scene and a mask of the
    pixels whose whole footprint maps to one texel."""
    w, h = flat.w, flat.h
    tf = math.tan(math.radians(flat.camera.vfov) / 2)
    asp = flat.camera.aspect
    n = tex.data.shape[0]
    rx = np.arange(w + 1) / w
    ry = np.arange(h + 1) / h
    # camera at (0, 0, 2) looking down -z: the hit on z = 0 of raster (rx, ry)
    X = 2 * asp * tf * (2 * rx - 1)
    Y = 2 * tf * (2 * ry - 1)

    def texel(c):
        t = (c + 1) / 2 * uv_scale
        return np.clip(((t - np.floor(t)) * n).astype(int), 0, n - 1)

    tx, ty = texel(X), texel(Y)
    same = (tx[:-1] == tx[1:])[None, :] & (ty[:-1] == ty[1:])[:, None]
    T = tex.data[ty[:-1]][:, tx[:-1]]          # [h, w, 4]
    exp = np.where(T[..., 3:4] > 0.5, T[..., :3], 1.0)
    return exp, same
