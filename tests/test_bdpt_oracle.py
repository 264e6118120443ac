"""The FP64 oracle of renderer::bdpt (renderer_bdpt.cpp:339-411, bidir.h),
restated literally in path space (explicit subpath weights, connection
terms and pdf products for every strategy of every full path), pinned on the
CPU by the furnace closed form and by agreement with the separately pinned
path-tracing oracle: E[BDPT(M)] = E[PT(M)], and with min_verts = 3,
E[BDPT] = E[PT(M)] - E[PT(2)]."""
import numpy as np
import pytest

from lm_amd import scenes
from oracle import oracle as orc

from lt_common import (batch_mean_z, furnace_scene, glass_lamp_scene, mixture_room_scene, smooth_sphere_scene)


def bdpt_batches(flat, K, per, M, seed, acc, min_verts=2):
    npix = flat.w * flat.h
    return np.asarray([orc.bdpt_render(flat, per * npix, M, seed=seed, pass_offset=per * k, accel=acc,
                                       min_verts=min_verts)['image'] for k in range(K)])


def block_z_vs(batches, ref_img, ref_sumsq, spp, block=4):
    K = batches.shape[0]
    h, w = ref_img.shape[:2]

    def bs(a):
        return a.reshape(a.shape[:-3] + (h // block, block, w // block, block, 3)).sum((-4, -2))

    gb = bs(batches)
    se = np.sqrt(gb.var(0, ddof=1) / K + bs(np.maximum(ref_sumsq / spp - ref_img ** 2, 0) / spp))
    ok = se > 0
    return (gb.mean(0) - bs(ref_img))[ok] / se[ok]


def test_bdpt_furnace_closed_form():
    """All path lengths 2..M: E = Le sum_{k=0}^{M-2} rho^k, as path tracing."""
    rho, Le, M = 0.5, 1.0, 8
    flat = furnace_scene(rho, Le).flatten()
    acc = orc.Accel(flat.verts, 0)
    means = bdpt_batches(flat, 16, 2, M, 3, acc)[..., 0].mean((1, 2))
    exp = Le * sum(rho ** k for k in range(M - 1))
    z = batch_mean_z(means, exp)
    print(f'BDPT oracle furnace {means.mean():.5f} expected {exp:.5f} z {z:.2f}')
    assert abs(z) < 5


SCENES = {'config0': scenes.config0, 'glass_lamp': glass_lamp_scene, 'smooth_sphere': smooth_sphere_scene,
          'mixture_room': mixture_room_scene}


@pytest.mark.parametrize('name', list(SCENES))
def test_bdpt_equals_pt(name):
    flat = SCENES[name]().flatten()
    flat.w = flat.h = 24
    flat.camera.aspect = 1.0
    acc = orc.Accel(flat.verts, 0)
    # path tracing reaches the lamp sealed in glass only through BSDF sampling:
    # its heavy-tailed estimate needs more samples for a reliable variance
    spp, M = (8192 if name == 'glass_lamp' else 2048), 6
    pt = orc.pt_render(flat, spp, M, seed=1, accel=acc)
    b = bdpt_batches(flat, 32, 24, M, 7, acc)
    z = block_z_vs(b, pt['image'], pt['sumsq'], spp)
    zt = block_z_vs(b, pt['image'], pt['sumsq'], spp, block=24)
    print(f'BDPT oracle {name}: mean {b.mean():.5f} vs PT {pt["image"].mean():.5f} (z {zt}); '
          f'max|z| {np.abs(z).max():.2f}, mean z^2 {np.mean(z ** 2):.2f}')
    assert np.all(np.abs(zt) < 4)
    assert np.abs(z).max() < 5 and np.mean(z ** 2) < 2.0


def test_bdpt_min_verts_drops_short_paths():
    """min_verts = 3 removes the two-vertex paths (emitters seen directly):
    E[BDPT] = E[PT(M)] - E[PT(2)] (same seed: PT(2) is exactly the directly
    seen emission of each PT sample)."""
    flat = scenes.config0().flatten()
    flat.w = flat.h = 24
    acc = orc.Accel(flat.verts, 0)
    spp, M = 2048, 6
    pt = orc.pt_render(flat, spp, M, seed=1, accel=acc)
    pt2 = orc.pt_render(flat, spp, 2, seed=1, accel=acc)
    b = bdpt_batches(flat, 24, 16, M, 9, acc, min_verts=3)
    ref = pt['image'] - pt2['image']
    z = block_z_vs(b, ref, pt['sumsq'] - pt2['sumsq'], spp)
    print(f'BDPT oracle min_verts 3: max|z| {np.abs(z).max():.2f}, mean z^2 {np.mean(z ** 2):.2f}')
    assert np.abs(z).max() < 5 and np.mean(z ** 2) < 2.0


def env_scenes():
    from lt_common import env_cube_scene

    def room_env():
        sc = scenes.config0()
        sc.env_Le = (0.6, 0.7, 0.8)   # the open front lets the environment in
        return sc

    def mixture_env():
        sc = mixture_room_scene()
        sc.env_Le = (0.5, 0.5, 0.5)
        return sc
    return {'env_cube': lambda: env_cube_scene(), 'room_env': room_env, 'mixture_env': mixture_env}


@pytest.mark.parametrize('name', ['env_cube', 'room_env', 'mixture_env'])
def test_bdpt_env_equals_pt(name):
    """light::envconst in renderer::bdpt (bidir.h:245-258 and 399-411: the
    environment's joint direction and disk-position sample as a
    non-connectable light endpoint; eye subpaths ending on the environment,
    strategy s = 0): E[BDPT(M)] = E[PT(M)] against the pinned path-tracing
    oracle, which samples the environment by NEE and BSDF hits."""
    flat = env_scenes()[name]().flatten()
    flat.w = flat.h = 24
    flat.camera.aspect = 1.0
    acc = orc.Accel(flat.verts, 0)
    spp, M = 2048, 6
    pt = orc.pt_render(flat, spp, M, seed=1, accel=acc)
    b = bdpt_batches(flat, 32, 24, M, 7, acc)
    z = block_z_vs(b, pt['image'], pt['sumsq'], spp)
    zt = block_z_vs(b, pt['image'], pt['sumsq'], spp, block=24)
    print(f'BDPT oracle {name}: mean {b.mean():.5f} vs PT {pt["image"].mean():.5f} (z {zt}); '
          f'max|z| {np.abs(z).max():.2f}, mean z^2 {np.mean(z ** 2):.2f}')
    assert np.all(np.abs(zt) < 4)
    assert np.abs(z).max() < 5 and np.mean(z ** 2) < 2.0


def test_bdpt_env_cube_closed_form():
    """A convex diffuse cube under a constant environment (no other light):
    every pixel that sees only the cube is rho * Le, reached by eye paths
    ending on the environment and by light paths that start on the
    bounding-sphere disk (z-tests over batches; renderer::bdpt splats each
    sample at its own raster position, so pixel values also carry the count
    noise of the uniform raster sampling)."""
    from lt_common import env_cube_scene
    rho, Le = (0.6, 0.4, 0.2), (1.0, 1.0, 1.0)
    flat = env_cube_scene(rho, Le, 24, 24).flatten()
    acc = orc.Accel(flat.verts, 0)
    b = bdpt_batches(flat, 16, 16, 4, 5, acc)
    img = b.mean(0)
    cover = orc.pt_render(flat, 64, 2, seed=3, accel=acc)['image'][..., 0]   # 0: every sample sees the cube
    cube = cover == 0.0
    se = b.std(0, ddof=1) / np.sqrt(b.shape[0])
    exp_cube = np.asarray(rho) * np.asarray(Le)
    z_cube = (img[cube] - exp_cube) / np.maximum(se[cube], 1e-12)
    zm = (b[:, cube].mean(1).mean(0) - exp_cube) / (b[:, cube].mean(1).std(0, ddof=1) / np.sqrt(b.shape[0]))
    print(f'BDPT env cube: {cube.sum()} cube pixels mean {img[cube].mean(0)} expected {exp_cube} '
          f'(z {zm}, per-pixel max|z| {np.abs(z_cube).max():.2f})')
    assert cube.sum() > 200
    assert np.abs(z_cube).max() < 5.5
    assert np.all(np.abs(zm) < 4)
