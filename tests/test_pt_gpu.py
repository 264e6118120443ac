"""GPU renderer::pt_mi355x against known answers and the FP64 oracle of
renderer::pt (statistical parity: per-pixel Welch t-test at alpha = 0.01)."""
import math

import numpy as np
import pytest

from conftest import cached

from lm_amd import scenes, stats
from lm_amd.renderer import Film, GpuScene, RendererPT
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def env_only_scene(w=64, h=48, Le=(0.3, 0.5, 0.7)):
    # one tiny triangle behind the camera so the accel is not empty
    tri = np.array([[[0, 0, 10], [0.01, 0, 10], [0, 0.01, 10]]], dtype=np.float64)
    prim = scenes.Primitive(scenes.flat_mesh(tri), scenes.Material(scenes.MAT_DIFFUSE, (0.5, 0.5, 0.5)))
    cam = scenes.Camera((0, 0, 3.5), (0, 0, 0), (0, 1, 0), 40.0, w / h)
    return scenes.Scene([prim], cam, Le, w, h)


def furnace_scene(rho=0.5, Le=1.0, w=32, h=32):
    room = scenes.box_room(open_front=False)
    prim = scenes.Primitive(room, scenes.Material(scenes.MAT_DIFFUSE, (rho, rho, rho)), Ke=(Le, Le, Le))
    cam = scenes.Camera((0, 0, 0.5), (0, 0, -1), (0, 1, 0), 60.0, w / h)
    return scenes.Scene([prim], cam, None, w, h)


def cube_env_scene(rho=(0.6, 0.4, 0.2), Le=(1.0, 1.0, 1.0), mtype=scenes.MAT_DIFFUSE, w=48, h=48):
    c = scenes.cube(1.6, 0.0, (0, 0, 0))
    mat = scenes.Material(mtype, rho) if mtype == scenes.MAT_DIFFUSE else scenes.Material(mtype)
    cam = scenes.Camera((0, 0, 4.0), (0, 0, 0), (0, 1, 0), 30.0, w / h)
    return scenes.Scene([scenes.Primitive(c, mat)], cam, Le, w, h)


def render_gpu(flat, spp, max_verts, mode='mis', seed=1, width=4):
    sc = GpuScene(flat, width=width)
    film = Film(flat.w, flat.h)
    r = RendererPT(sc, film, max_verts=max_verts, seed=seed, sampling_mode=mode, spp=spp)
    out = r.render()
    assert out['processed'] == spp
    return film, out


def test_env_only_exact(device):
    flat = env_only_scene().flatten()
    film, _ = render_gpu(flat, 4, 8)
    assert np.allclose(film.data, np.array([0.3, 0.5, 0.7]), rtol=1e-6, atol=0)


@pytest.mark.parametrize('mode', ['naive', 'nee', 'mis'])
def test_furnace_known_answer(device, mode):
    """Closed box, every wall emits Le and reflects rho (diffuse):
    E[L] = Le * sum_{k=0}^{M-2} rho^k for max_verts = M, in every sampling
    mode (MIS weights must sum to one; RR must stay unbiased)."""
    rho, Le, M, spp = 0.5, 1.0, 8, 64
    flat = furnace_scene(rho, Le).flatten()
    film, _ = render_gpu(flat, spp, M, mode=mode)
    expected = Le * sum(rho ** k for k in range(M - 1))
    n = flat.w * flat.h * spp
    s = film.sum[..., 0].sum()
    q = film.sumsq[..., 0].sum()
    mean = s / n
    var = q / n - mean ** 2
    z = (mean - expected) / math.sqrt(var / n)
    print(f'furnace {mode}: mean {mean:.6f} expected {expected:.6f} z {z:.2f}')
    assert abs(z) < 5


def test_convex_diffuse_under_env(device):
    """A convex diffuse object under a constant environment reflects
    L = rho * Le (reflected rays never re-hit the object)."""
    rho = np.array([0.6, 0.4, 0.2])
    flat = cube_env_scene(tuple(rho)).flatten()
    spp = 64
    film, _ = render_gpu(flat, spp, 8)
    c = slice(flat.h // 2 - 6, flat.h // 2 + 6)
    s = film.sum[c, c].reshape(-1, 3).sum(0)
    q = film.sumsq[c, c].reshape(-1, 3).sum(0)
    n = 12 * 12 * spp
    mean = s / n
    se = np.sqrt((q / n - mean ** 2) / n)
    z = (mean - rho) / se
    print('convex cube', mean, z)
    assert np.all(np.abs(z) < 5)
    # corners see the environment only
    assert np.allclose(film.data[0, 0], 1.0) and np.allclose(film.data[-1, -1], 1.0)


def test_mirror_convex_under_env(device):
    flat = cube_env_scene(mtype=scenes.MAT_MIRROR).flatten()
    film, _ = render_gpu(flat, 8, 8)
    assert np.allclose(film.data, 1.0, rtol=1e-5)


def orc_render(flat, spp, max_verts, seed, mode=2):
    return orc.pt_render(flat, spp=spp, max_verts=max_verts, sampling_mode=mode, seed=seed)


def test_c0_same_stream_matches_oracle(device):
    """Same Philox streams: the FP32 GPU walk and the FP64 oracle walk take
    the same decisions almost everywhere, so per-pixel sums agree closely."""
    flat = scenes.config0().flatten()
    spp, M, seed = 16, 8, 1234
    film, out = render_gpu(flat, spp, M, seed=seed)
    ref = orc_render(flat, spp, M, seed)
    a, b = film.sum, ref['sum']
    close = np.abs(a - b) <= 1e-3 * (1.0 + np.abs(b))
    frac = close.all(-1).mean()
    rel = np.abs(a.sum() - b.sum()) / b.sum()
    print(f'C0 same-stream: pixels within 1e-3: {frac:.4f}; total rel diff {rel:.2e}; '
          f'gpu rays {out["stats"]}; oracle {ref["stats"]}')
    assert frac >= 0.98
    assert rel < 1e-3


@pytest.mark.parametrize('mode', ['mis', 'nee', 'naive'])
def test_c0_welch_vs_oracle(device, mode):
    """Independent seeds: per-pixel Welch t-test on luminance at alpha=0.01;
    the failing fraction must stay within the binomial bound of 1%."""
    flat = scenes.config0().flatten()
    flat.w = flat.h = 128
    spp, M = 64, 8
    m = {'naive': 0, 'nee': 1, 'mis': 2}[mode]
    film, _ = render_gpu(flat, spp, M, mode=mode, seed=77)
    ref = orc_render(flat, spp, M, seed=991, mode=m)
    frac, p, a = stats.welch(film.sum, film.sumsq, spp, ref['sum'], ref['sumsq'], spp)
    bound = stats.binomial_bound(0.01, flat.w * flat.h)
    print(f'C0 {mode} Welch fail fraction {frac:.4f} (bound {bound:.4f}); '
          f'rMSE gpu-vs-oracle {stats.rel_mse(film.sum / spp, ref["image"]):.3e}')
    assert frac <= bound


def c2_flat(levels=8, w=160, h=90):
    def make():
        s = scenes.config2(levels=levels, width=w, height=h)
        return s.flatten()
    return cached(('c2', levels, w, h), make)


def test_c2_welch_vs_oracle(device):
    """Config 2 (all four materials, 4 area lights + env, 1.32M triangles) at
    160x90, 128 spp, max 16 vertices, independent seeds."""
    flat = c2_flat()
    spp, M = 128, 16
    film, out = render_gpu(flat, spp, M, seed=5)
    oacc = cached('c2_oracle_accel', lambda: orc.Accel(flat.verts))
    ref = orc.pt_render(flat, spp=spp, max_verts=M, sampling_mode=2, seed=6, accel=oacc)
    frac, p, a = stats.welch(film.sum, film.sumsq, spp, ref['sum'], ref['sumsq'], spp)
    bound = stats.binomial_bound(0.01, flat.w * flat.h)
    print(f'C2 Welch fail fraction {frac:.4f} (bound {bound:.4f}) stats {out["stats"]}')
    assert frac <= bound


def test_c2_same_stream(device):
    flat = c2_flat()
    spp, M, seed = 8, 16, 42
    film, _ = render_gpu(flat, spp, M, seed=seed)
    oacc = cached('c2_oracle_accel', lambda: orc.Accel(flat.verts))
    ref = orc.pt_render(flat, spp=spp, max_verts=M, sampling_mode=2, seed=seed, accel=oacc)
    a, b = film.sum, ref['sum']
    close = (np.abs(a - b) <= 1e-3 * (1.0 + np.abs(b))).all(-1)
    print(f'C2 same-stream: pixels within 1e-3: {close.mean():.4f}')
    assert close.mean() >= 0.90


def test_time_scheduler_and_device_film(device):
    import torch
    flat = scenes.config0().flatten()
    sc = GpuScene(flat)
    s = torch.zeros((flat.h, flat.w, 3), dtype=torch.float64, device=device)
    q = torch.zeros_like(s)
    _, _, st = sc.render_raw(spp=2, max_verts=8, seed=3, film_sum=s, film_sumsq=q)
    assert st['spp'] == 2
    _, _, st2 = sc.render_raw(spp=3, max_verts=8, seed=3, sample_offset=2, film_sum=s, film_sumsq=q)
    torch.cuda.synchronize()
    # five samples accumulated on the device equal a 5-spp host render
    h, hq, _ = sc.render_raw(spp=5, max_verts=8, seed=3)
    assert np.allclose(s.cpu().numpy(), h, rtol=1e-12, atol=1e-12)
    film = Film(flat.w, flat.h)
    r = RendererPT(sc, film, max_verts=8, seed=1, scheduler='time', render_time=0.2)
    out = r.render()
    assert out['processed'] >= 1 and out['elapsed'] >= 0.2
    # the time scheduler renders chunks of whole samples per pixel on consecutive
    # sample indices: the film equals a fixed-spp render of as many samples
    t, tq, st = sc.render_raw(max_verts=8, seed=4, render_time=0.3)
    assert st['spp'] >= 2 and st['samples'] == st['spp'] * flat.w * flat.h and st['elapsed'] >= 0.3
    f, fq, _ = sc.render_raw(spp=st['spp'], max_verts=8, seed=4)
    assert np.allclose(t, f, rtol=1e-9, atol=1e-12) and np.allclose(tq, fq, rtol=1e-9, atol=1e-12)


def test_invalid_render_args(device):
    from lm_amd._lib import LmError
    flat = scenes.config0().flatten()
    sc = GpuScene(flat)
    with pytest.raises(LmError):
        sc.render_raw(spp=1, max_verts=1)
    with pytest.raises(LmError):
        sc.render_raw(spp=0, max_verts=8)
    with pytest.raises(LmError):
        sc.render_raw(spp=1, max_verts=8, sampling_mode=7)
    with pytest.raises(LmError, match='beyond'):
        sc.render_raw(spp=4, max_verts=8, sample_offset=2**31 - 2)
