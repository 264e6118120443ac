"""GPU material::mixture_wavefrontobj and texture::bitmap (the OBJ material
path of model_wavefrontobj.cpp:155-215, 340-513) in renderer::pt,
renderer::lt and the batched renderer: closed form with a binary alpha mask,
same-stream parity with the FP64 oracle, and the LT = PT - direct identity."""
import numpy as np
import pytest

from lm_amd import _lib
from lm_amd.batch import Batch
from lm_amd.renderer import Film, GpuScene, RendererRaycast
from oracle import oracle as orc

from lt_common import block_z, masked_quad_expected, masked_quad_scene, mixture_room_scene

pytestmark = pytest.mark.gpu


def test_masked_texture_quad_closed_form(device):
    s, tex = masked_quad_scene()
    flat = s.flatten()
    spp = 64
    sum_, sq, _ = GpuScene(flat).render_raw(spp=spp, max_verts=8, seed=3)
    img = sum_ / spp
    exp, same = masked_quad_expected(flat, tex)
    se = np.sqrt(np.maximum(sq / spp - img ** 2, 0) / spp)
    hole = same & (exp == 1.0).all(-1)
    assert hole.sum() > 100 and (same & ~hole).sum() > 100
    assert np.allclose(img[hole], 1.0, rtol=1e-6)
    z = (img - exp)[same & ~hole] / se[same & ~hole]
    print(f'masked quad GPU: max|z| {np.abs(z).max():.2f}, mean z^2 {np.mean(z ** 2):.2f}')
    assert np.abs(z).max() < 5 and np.mean(z ** 2) < 1.5


@pytest.mark.parametrize('name', ['mixture_room', 'masked_quad'])
def test_pt_same_stream_matches_oracle(device, name):
    flat = mixture_room_scene(48, 48).flatten() if name == 'mixture_room' else masked_quad_scene()[0].flatten()
    spp, M, seed = 16, 8, 99
    s, _, _ = GpuScene(flat).render_raw(spp=spp, max_verts=M, seed=seed)
    ref = orc.pt_render(flat, spp, M, seed=seed)
    b = ref['sum']
    close = np.abs(s - b) <= 1e-3 * (1.0 + np.abs(b))
    frac = close.all(-1).mean()
    rel = abs(s.sum() - b.sum()) / b.sum()
    print(f'{name} PT same-stream: pixels within 1e-3 {frac:.4f}, total rel diff {rel:.2e}')
    assert frac >= 0.98 and rel < 1e-3


def test_lt_same_stream_matches_oracle_mixture(device):
    flat = mixture_room_scene(48, 48).flatten()
    npix = flat.w * flat.h
    n, M, seed = 16 * npix, 8, 7
    s, _ = GpuScene(flat).render_lt(num_samples=n, max_verts=M, seed=seed)
    ref = orc.lt_render(flat, n, M, seed=seed)
    b = ref['sum']
    close = np.abs(s - b) <= 1e-3 * (1.0 + np.abs(b))
    frac = close.all(-1).mean()
    rel = abs(s.sum() - b.sum()) / b.sum()
    print(f'mixture room LT same-stream: pixels within 1e-3 {frac:.4f}, total rel diff {rel:.2e}')
    assert frac >= 0.97 and rel < 2e-3


def test_lt_equals_pt_minus_direct_mixture(device):
    flat = mixture_room_scene(64, 64).flatten()
    npix = flat.w * flat.h
    M, spp = 8, 4096
    sc = GpuScene(flat)
    s, q, _ = sc.render_raw(spp=spp, max_verts=M, seed=1)
    s2, _, _ = sc.render_raw(spp=spp, max_verts=2, seed=1)
    lts = []
    for k in range(32):
        f, st = sc.render_lt(num_samples=64 * npix, max_verts=M, seed=101, pass_offset=64 * k)
        lts.append(f * (npix / st['samples']))
    z, nb = block_z(lts, s / spp, q, spp, s2 / spp, block=8)
    zt, _ = block_z(lts, s / spp, q, spp, s2 / spp, block=flat.w)
    print(f'mixture room LT vs PT-PT2: whole-image z {zt}; {nb} blocks, max|z| {np.abs(z).max():.2f}, '
          f'mean z^2 {np.mean(z ** 2):.2f}')
    assert nb >= 100 and np.all(np.abs(zt) < 4.0)
    assert np.abs(z).max() < 5.5 and np.mean(z ** 2) < 2.0


def test_raycast_shows_texture_reflectance(device):
    """renderer::raycast uses Material::reflectance: the texel's Kd of a
    textured mixture (use_constant_color: no shading factor)."""
    s, tex = masked_quad_scene()
    flat = s.flatten()
    film = Film(flat.w, flat.h)
    RendererRaycast(GpuScene(flat), film, use_constant_color=True).render()
    exp, same = masked_quad_expected(flat, tex)
    # reflectance ignores the alpha mask: the texel colour everywhere
    tx = np.broadcast_to(exp, film.data.shape)
    solid = same & ~(exp == 1.0).all(-1)
    assert np.allclose(film.data[solid], tx[solid], rtol=1e-5)


def test_batch_textured_jobs_equal_single_renders(device):
    """Textured jobs in one batch (texture::bitmap Kd and an alpha mask):
    each job's texture tables sit in the batch's arenas behind the job's own
    base pointers, so every job renders as lmgpu_pt_render of its scene
    alone (the same Philox streams and per-vertex code: equal films). An
    untextured job between the two checks the offsets."""
    from lm_amd import scenes
    jobs = [mixture_room_scene(40, 32).flatten(), scenes.config0(seed=3).flatten(),
            masked_quad_scene(24, 24)[0].flatten()]
    jobs[1].w, jobs[1].h = 32, 32
    jobs[1].camera.aspect = 1.0
    params = [dict(spp=6, max_verts=8, seed=4), dict(spp=4, max_verts=6, seed=5), dict(spp=5, max_verts=6, seed=6)]
    b = Batch()
    for flat, p in zip(jobs, params):
        b.add_job(flat, **p)
    b.run(width=4, path_capacity=4096, block=256)
    for j, (flat, p) in enumerate(zip(jobs, params)):
        ref, _, _ = GpuScene(flat, width=4).render_raw(spp=p['spp'], max_verts=p['max_verts'], seed=p['seed'])
        got = b.film(j)
        got = got[0] if isinstance(got, tuple) else got
        err = np.abs(got - ref).max() / max(1.0, np.abs(ref).max())
        print(f'textured batch job {j}: max rel diff {err:.2e}')
        assert np.allclose(got, ref, rtol=1e-9, atol=1e-12)
    # a texture index beyond the job's table is an error when the arena is built
    bad = masked_quad_scene()[0].flatten()
    bad.mat_tex[:] = 3
    b2 = Batch()
    b2.add_job(bad, spp=1, max_verts=4, seed=1)
    with pytest.raises(_lib.LmError):
        b2.run(width=4)
