"""CPU pins of the FP64 oracle: every function is checked against golden
vectors, brute force, closed forms or mathematical invariants (none of these
retype the oracle's formulas)."""
import math

import numpy as np
import pytest

from lm_amd import scenes
from oracle import oracle as orc

# Random123 kat_vectors, philox4x32 with 10 rounds: (ctr, key, expected)
PHILOX_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


def philox_spec(ctr, key, rounds=10):
    """Independent from-spec Philox4x32 (Salmon et al. SC'11, sec. 3):
    S-box (L0,R0,L1,R1) -> (hi(M1*R1)^L1^k0... ) written in the paper's
    bijection form with Python integers."""
    M0, M1 = 0xD2511F53, 0xCD9E8D57
    W0, W1 = 0x9E3779B9, 0xBB67AE85
    x = list(ctr)
    k0, k1 = key
    mask = 0xffffffff
    for _ in range(rounds):
        p0 = M0 * x[0]
        p1 = M1 * x[2]
        x = [((p1 >> 32) ^ x[1] ^ k0) & mask, p1 & mask, ((p0 >> 32) ^ x[3] ^ k1) & mask, p0 & mask]
        k0 = (k0 + W0) & mask
        k1 = (k1 + W1) & mask
    return x


@pytest.mark.parametrize('ctr,key,exp', PHILOX_KAT)
def test_philox_kat(ctr, key, exp):
    assert list(orc.philox4x32_10(ctr, key)) == exp
    assert philox_spec(ctr, key) == exp


def test_philox_random_counters_match_spec():
    rng = np.random.default_rng(0)
    for _ in range(200):
        c = [int(x) for x in rng.integers(0, 2 ** 32, 4)]
        k = [int(x) for x in rng.integers(0, 2 ** 32, 2)]
        assert list(orc.philox4x32_10(c, k)) == philox_spec(c, k)


def small_scene():
    sph, cl = scenes.config1_mesh(levels=3, clutter_count=300)
    return scenes.merge_meshes([sph, cl]), sph


def test_bvh_equals_brute_force():
    m, sph = small_scene()
    verts = m.triangles()
    acc = orc.Accel(verts)
    assert acc.check() == 0
    rays = scenes.surface_rays(sph, 5000, 3)
    cam = scenes.Camera((0, 0, 3), (0, 0, 0), (0, 1, 0), 40.0, 1.0)
    o, d = scenes.primary_rays(cam, 64, 64)
    for r in (rays, scenes.make_rays(o, d)):
        a = acc.intersect(r)
        b = orc.brute_force(verts, r)
        assert np.array_equal(a['prim'], b['prim'])
        assert np.array_equal(a['t'], b['t'])


def test_moller_trumbore_closed_form():
    """Known answers: closed edges, closed [tmin,tmax], |det| < 1e-8 rejected."""
    tri = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], dtype=np.float32)
    def cast(o, d, tmin=1e-3, tmax=1e10):
        r = {'o': np.array([o], np.float32), 'd': np.array([d], np.float32),
             'tmin': np.array([tmin], np.float32), 'tmax': np.array([tmax], np.float32)}
        return orc.brute_force(tri, r)
    h = cast((0.25, 0.5, 2.0), (0, 0, -1))
    assert h['prim'][0] == 0 and h['t'][0] == 2.0 and h['u'][0] == 0.25 and h['v'][0] == 0.5
    assert cast((0.5, 0.5, 2.0), (0, 0, -1))['prim'][0] == 0          # on the hypotenuse
    assert cast((0.0, 0.0, 2.0), (0, 0, -1))['prim'][0] == 0          # on a vertex
    assert cast((0.5, 0.51, 2.0), (0, 0, -1))['prim'][0] == -1
    assert cast((0.2, 0.2, 2.0), (0, 0, -1), tmax=2.0)['prim'][0] == 0    # t == tmax accepted
    assert cast((0.2, 0.2, 2.0), (0, 0, -1), tmin=2.0)['prim'][0] == 0    # t == tmin accepted
    assert cast((0.2, 0.2, 2.0), (0, 0, -1), tmax=1.999)['prim'][0] == -1
    tiny = np.array([[0, 0, 0, 1e-4, 0, 0, 0, 1e-4, 0]], dtype=np.float32)   # det = 1e-8 * |d|
    r = {'o': np.array([[1e-5, 1e-5, 1.0]], np.float32), 'd': np.array([[0, 0, -1]], np.float32),
         'tmin': np.array([1e-3], np.float32), 'tmax': np.array([1e10], np.float32)}
    tiny2 = tiny * 0.5
    assert orc.brute_force(tiny2, r)['prim'][0] == -1     # det 2.5e-9 < 1e-8
    big = tiny * 2
    assert orc.brute_force(big, r)['prim'][0] == 0        # det 4e-8


def test_equal_t_later_triangle_wins():
    """Two coincident triangles: the reference keeps the later one in
    traversal order (t <= tmax accepted), brute force = index order."""
    t = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]] * 2, dtype=np.float32)
    r = {'o': np.array([[0.2, 0.2, 1.0]], np.float32), 'd': np.array([[0, 0, -1]], np.float32),
         'tmin': np.array([1e-3], np.float32), 'tmax': np.array([1e10], np.float32)}
    assert orc.brute_force(t, r)['prim'][0] == 1


def test_sah_toy_split():
    """Two pairs of coincident unit triangles 10 apart: the root splits the
    pairs (cost 1 + small), each pair stays a leaf (splitting costs 3 > 2)."""
    a = [0, 0, 0, 1, 0, 0, 0, 1, 0]
    b = [10, 0, 0, 11, 0, 0, 10, 1, 0]
    verts = np.array([a, b, a, b], dtype=np.float32)
    acc = orc.Accel(verts)
    bounds, info, idx = acc.export()
    assert acc.num_nodes == 3
    assert info[0, 0] == 0                     # root internal
    leaves = [info[i] for i in (info[0, 3], info[0, 4])]
    assert all(l[0] == 1 and l[2] - l[1] == 2 for l in leaves)
    sets = [sorted(idx[l[1]:l[2]]) for l in leaves]
    assert sorted(sets) == [[0, 2], [1, 3]]


def _bound(tris):
    t = tris.reshape(-1, 3, 3)
    return t.min(axis=(0, 1)), t.max(axis=(0, 1))


def _sa(lo, hi):
    d = hi - lo
    return 2 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0])


def test_sah_root_split_is_brute_force_optimum():
    """The root partition minimises 1 + (A_L n_L + A_R n_R)/A over all axes
    and all split positions of the centroid order (accel_sahbvh.cpp:225-250),
    recomputed here by brute force in numpy."""
    rng = np.random.default_rng(4)
    c = rng.uniform(-1, 1, size=(40, 3))
    e = rng.normal(scale=0.1, size=(40, 3, 3))
    verts = (c[:, None, :] + e).reshape(-1, 9).astype(np.float32)
    acc = orc.Accel(verts)
    bounds, info, idx = acc.export()
    v = verts.astype(np.float64).reshape(-1, 3, 3)
    lo, hi = v.min(axis=1), v.max(axis=1)
    cen = (lo + hi) / 2
    A = _sa(lo.min(0), hi.max(0))
    best = None
    for ax in range(3):
        order = np.argsort(cen[:, ax], kind='stable')
        for i in range(1, len(order)):
            L, R = order[:i], order[i:]
            cost = 1 + (_sa(lo[L].min(0), hi[L].max(0)) * len(L) + _sa(lo[R].min(0), hi[R].max(0)) * len(R)) / A
            if best is None or cost < best[0] - 1e-12:
                best = (cost, set(L.tolist()))
    left = info[info[0, 3]]
    # the left child's triangle set equals the brute-force optimum's left set
    stack, got = [int(info[0, 3])], set()
    while stack:
        k = stack.pop()
        if info[k, 0]:
            got |= set(idx[info[k, 1]:info[k, 2]].tolist())
        else:
            stack += [int(info[k, 3]), int(info[k, 4])]
    assert got == best[1] or got == set(range(40)) - best[1]


# ---- materials -----------------------------------------------------------------

N = np.array([0.0, 0.0, 1.0])


def _hemi_quadrature(n_theta=256, n_phi=512):
    th = (np.arange(n_theta) + 0.5) / n_theta * (np.pi / 2)
    ph = (np.arange(n_phi) + 0.5) / n_phi * 2 * np.pi
    T, P = np.meshgrid(th, ph, indexing='ij')
    w = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    dA = (np.sin(T) * (np.pi / 2 / n_theta) * (2 * np.pi / n_phi)).reshape(-1)
    return w, dA


@pytest.mark.parametrize('mtype,params', [
    (scenes.MAT_DIFFUSE, orc._mat_params((0.7, 0.5, 0.3))),
    (scenes.MAT_GLOSSY, orc._mat_params((0.9, 0.9, 0.9), 0.3, 0.3)),
    (scenes.MAT_GLOSSY, orc._mat_params((0.8, 0.8, 0.8), 0.1, 0.4)),
    (scenes.MAT_MIXTURE, orc._mat_params((0.5, 0.4, 0.3), 0.2, 0.3, ks=(0.3, 0.35, 0.4))),
])
def test_pdf_integrates_to_one_and_energy(mtype, params):
    """Projected-solid-angle pdf integrates to 1 (diffuse exactly, GGX up to
    the below-horizon mass of the visible half vectors); white furnace
    int f cos = Kd for diffuse; energy <= 1 for GGX."""
    w, dA = _hemi_quadrature()
    wi = np.array([0.3, 0.2, math.sqrt(1 - 0.13)])
    f, pdf = orc.mat_eval(mtype, params, N, np.repeat(wi[None], w.shape[0], 0), w)
    cos = w[:, 2]
    integ_pdf = np.sum(pdf * cos * dA)
    albedo = np.sum(f * cos[:, None] * dA[:, None], axis=0)
    if mtype == scenes.MAT_DIFFUSE:
        assert abs(integ_pdf - 1) < 1e-4
        assert np.allclose(albedo, params[:3], atol=1e-4)
    else:
        assert 0.85 < integ_pdf <= 1.0 + 1e-3
        assert np.all(albedo <= 1.0 + 1e-3)


@pytest.mark.parametrize('mtype,params', [
    (scenes.MAT_DIFFUSE, orc._mat_params((0.7, 0.5, 0.3))),
    (scenes.MAT_GLOSSY, orc._mat_params((0.9, 0.6, 0.3), 0.25, 0.25)),
    (scenes.MAT_MIXTURE, orc._mat_params((0.5, 0.4, 0.3), 0.2, 0.3, ks=(0.3, 0.35, 0.4))),
])
def test_reciprocity(mtype, params):
    rng = np.random.default_rng(1)
    a = rng.normal(size=(500, 3)); a[:, 2] = np.abs(a[:, 2]); a /= np.linalg.norm(a, axis=1, keepdims=True)
    b = rng.normal(size=(500, 3)); b[:, 2] = np.abs(b[:, 2]); b /= np.linalg.norm(b, axis=1, keepdims=True)
    f1, _ = orc.mat_eval(mtype, params, N, a, b)
    f2, _ = orc.mat_eval(mtype, params, N, b, a)
    assert np.allclose(f1, f2, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize('mtype,params', [
    (scenes.MAT_DIFFUSE, orc._mat_params((0.7, 0.5, 0.3))),
    (scenes.MAT_GLOSSY, orc._mat_params((0.9, 0.6, 0.3), 0.25, 0.25)),
    (scenes.MAT_MIXTURE, orc._mat_params((0.5, 0.4, 0.3), 0.2, 0.3, ks=(0.3, 0.35, 0.4))),
])
def test_sample_weight_equals_f_over_pdf_and_chi2(mtype, params):
    """Sampled directions: weight = f / pdf (projected measure), and the
    sample histogram over (cos theta) bins follows the pdf (chi-square)."""
    rng = np.random.default_rng(2)
    n = 200000
    wi = np.repeat(np.array([[0.3, 0.2, math.sqrt(1 - 0.13)]]), n, 0)
    u = rng.random((n, 4))
    wo, w, comp, ok = orc.mat_sample(mtype, params, N, wi, u)
    f, pdf = orc.mat_eval(mtype, params, N, wi[ok], wo[ok])
    good = pdf > 1e-6
    assert np.allclose(w[ok][good], f[good] / pdf[good][:, None], rtol=1e-6)
    # chi-square on cos(theta) bins; expected masses by midpoint quadrature in
    # (mu = cos theta, phi), nodes aligned with the bins (d omega = d mu d phi)
    edges = np.linspace(0, 1, 21)
    hist, _ = np.histogram(wo[ok][:, 2], bins=edges)
    nmu, nphi = 64, 1024
    mass = []
    for a, b in zip(edges[:-1], edges[1:]):
        mu = a + (np.arange(nmu) + 0.5) / nmu * (b - a)
        ph = (np.arange(nphi) + 0.5) / nphi * 2 * np.pi
        M_, P_ = np.meshgrid(mu, ph, indexing='ij')
        s_ = np.sqrt(1 - M_ ** 2)
        qw = np.stack([s_ * np.cos(P_), s_ * np.sin(P_), M_], -1).reshape(-1, 3)
        _, qpdf = orc.mat_eval(mtype, params, N, np.repeat(wi[:1], qw.shape[0], 0), qw)
        mass.append(np.sum(qpdf * qw[:, 2]) * (b - a) / nmu * 2 * np.pi / nphi)
    mass = np.array(mass)
    expected = mass / mass.sum() * hist.sum()
    m = expected > 20
    chi2 = np.sum((hist[m] - expected[m]) ** 2 / expected[m])
    from scipy.stats import chi2 as chi2d
    assert chi2d.sf(chi2, m.sum() - 1) > 1e-4, chi2


def test_mixture_is_the_sum_of_its_lobes():
    """material::mixture_wavefrontobj (model_wavefrontobj.cpp:340-513): f =
    f_diffuse(Kd) + f_glossy(Ks), pdf = w p_diffuse + (1-w) p_glossy with
    w = max(Kd) / (max(Kd) + max(Ks)), both evaluated by the separately
    pinned diffuse and glossy materials."""
    rng = np.random.default_rng(3)
    a = rng.normal(size=(400, 3)); a[:, 2] = np.abs(a[:, 2]); a /= np.linalg.norm(a, axis=1, keepdims=True)
    b = rng.normal(size=(400, 3)); b[:, 2] = np.abs(b[:, 2]); b /= np.linalg.norm(b, axis=1, keepdims=True)
    kd, ks, ax, ay = (0.5, 0.4, 0.3), (0.2, 0.6, 0.1), 0.15, 0.35
    fm, pm = orc.mat_eval(scenes.MAT_MIXTURE, orc._mat_params(kd, ax, ay, ks=ks), N, a, b)
    fd, pd = orc.mat_eval(scenes.MAT_DIFFUSE, orc._mat_params(kd), N, a, b)
    fg, pg = orc.mat_eval(scenes.MAT_GLOSSY, orc._mat_params(ks, ax, ay), N, a, b)
    w = 0.5 / (0.5 + 0.6)
    assert np.allclose(fm, fd + fg, rtol=1e-12, atol=1e-14)
    assert np.allclose(pm, w * pd + (1 - w) * pg, rtol=1e-12, atol=1e-14)
    # black Kd and Ks: the diffuse lobe is chosen (w = 1), f = 0
    f0, p0 = orc.mat_eval(scenes.MAT_MIXTURE, orc._mat_params((0, 0, 0), ax, ay), N, a, b)
    assert np.all(f0 == 0) and np.allclose(p0, 1 / np.pi)


def test_mixture_alpha_mask_components():
    """Alpha mask: component 0 (diffuse+glossy) with probability alpha and
    weight 1/alpha, component 1 (pass-through, specular) otherwise with
    weight 1/(1-alpha) and direction -wi; the reference's pass-through
    direction weight is 1 (model_wavefrontobj.cpp:466-473), so comp 0's
    total weight is f_sum / pdf and comp 1's is 1/(1-alpha)."""
    rng = np.random.default_rng(4)
    n, alpha = 50000, 0.3
    kd, ks = (0.5, 0.4, 0.3), (0.2, 0.2, 0.2)
    p = orc._mat_params(kd, 0.2, 0.2, ks=ks, alpha=alpha)
    wi = np.repeat(np.array([[0.3, 0.2, math.sqrt(1 - 0.13)]]), n, 0)
    wo, w, comp, ok = orc.mat_sample(scenes.MAT_MIXTURE, p, N, wi, rng.random((n, 4)))
    frac = (comp == 1).mean()
    assert abs(frac - (1 - alpha)) < 5 * math.sqrt(alpha * (1 - alpha) / n)
    c1 = comp == 1
    assert np.allclose(wo[c1], -wi[c1]) and np.allclose(w[c1], 1 / (1 - alpha))
    c0 = (comp == 0) & ok
    f, pdf = orc.mat_eval(scenes.MAT_MIXTURE, p, N, wi[c0], wo[c0])   # includes the alpha factor
    good = pdf > 1e-6
    assert np.allclose(w[c0][good], f[good] / alpha / pdf[good][:, None], rtol=1e-6)


def test_mirror_and_glass():
    """Mirror: reflection about n, weight 1. Glass: component probabilities
    Fr / 1-Fr with weight 1/p; reflection weight Fr, refraction
    (1-Fr) eta^2 (importance transport), so comp weight * dir weight is 1 for
    reflection and eta^2 for refraction; refracted ray obeys Snell's law."""
    wi = np.array([[0.6, 0.0, 0.8]])
    wo, w, comp, ok = orc.mat_sample(scenes.MAT_MIRROR, orc._mat_params(), N, wi, np.zeros((1, 4)))
    assert np.allclose(wo, [[-0.6, 0, 0.8]]) and np.allclose(w, 1)
    ni = 1.5
    p = orc._mat_params(ni=ni)
    us = np.array([[0.0, 0, 0, 0], [0.999, 0, 0, 0]])
    wo, w, comp, ok = orc.mat_sample(scenes.MAT_GLASS, p, N, np.repeat(wi, 2, 0), us)
    assert comp.tolist() == [0, 1]
    assert np.allclose(wo[0], [-0.6, 0, 0.8]) and np.allclose(w[0], 1.0)
    eta = 1 / ni
    assert np.allclose(w[1], eta * eta)
    sin_t = math.hypot(wo[1, 0], wo[1, 1])
    assert abs(sin_t - 0.6 / ni) < 1e-12 and wo[1, 2] < 0
    # Schlick Fresnel at normal incidence is ((1-n)/(1+n))^2: fraction of reflections
    rng = np.random.default_rng(0)
    n = 100000
    wo, w, comp, ok = orc.mat_sample(scenes.MAT_GLASS, p, N, np.repeat([[0, 0, 1.0]], n, 0), rng.random((n, 4)))
    r0 = ((1 - ni) / (1 + ni)) ** 2
    assert abs((comp == 0).mean() - r0) < 5 * math.sqrt(r0 * (1 - r0) / n)


# ---- renderer::pt known answers --------------------------------------------------

def test_env_only_exact_oracle():
    tri = np.array([[[0, 0, 10], [0.01, 0, 10], [0, 0.01, 10]]], dtype=np.float64)
    prim = scenes.Primitive(scenes.flat_mesh(tri), scenes.Material(scenes.MAT_DIFFUSE, (0.5, 0.5, 0.5)))
    cam = scenes.Camera((0, 0, 3.5), (0, 0, 0), (0, 1, 0), 40.0, 1.0)
    flat = scenes.Scene([prim], cam, (0.3, 0.5, 0.7), 16, 16).flatten()
    r = orc.pt_render(flat, spp=4, max_verts=8, seed=1)
    assert np.array_equal(r['image'], np.broadcast_to(np.array([0.3, 0.5, 0.7], np.float32).astype(np.float64), r['image'].shape))


@pytest.mark.parametrize('mode', [0, 1, 2])
def test_furnace_oracle(mode):
    rho, Le, M, spp = 0.5, 1.0, 8, 64
    room = scenes.box_room(open_front=False)
    prim = scenes.Primitive(room, scenes.Material(scenes.MAT_DIFFUSE, (rho,) * 3), Ke=(Le,) * 3)
    cam = scenes.Camera((0, 0, 0.5), (0, 0, -1), (0, 1, 0), 60.0, 1.0)
    flat = scenes.Scene([prim], cam, None, 24, 24).flatten()
    r = orc.pt_render(flat, spp=spp, max_verts=M, sampling_mode=mode, seed=3)
    n = 24 * 24 * spp
    mean = r['sum'][..., 0].sum() / n
    var = r['sumsq'][..., 0].sum() / n - mean ** 2
    expected = Le * sum(rho ** k for k in range(M - 1))
    assert abs(mean - expected) / math.sqrt(var / n) < 5


def _polygon_irradiance(x, n, poly, Le):
    """Lambert's closed form: E = Le/2 * sum_i angle(r_i, r_i+1) * n . unit(r_i x r_i+1)
    (sign chosen so the result is positive for a polygon facing x)."""
    E = 0.0
    for i in range(len(poly)):
        a = np.asarray(poly[i], float) - x
        b = np.asarray(poly[(i + 1) % len(poly)], float) - x
        a /= np.linalg.norm(a)
        b /= np.linalg.norm(b)
        c = np.cross(a, b)
        E += math.acos(np.clip(np.dot(a, b), -1, 1)) * np.dot(n, c / np.linalg.norm(c))
    return abs(E) * Le / 2


@pytest.mark.parametrize('mode', [0, 1, 2])
def test_area_light_irradiance_closed_form(mode):
    """Diffuse floor under a square emitter, camera looking straight down at
    one point with a tiny field of view and max_verts = 3: L = rho/pi * E
    with E from Lambert's polygon formula (pins light::area sampling, the
    geometry term, diffuse eval and the MIS weights)."""
    rho, Le, hgt, a = 0.8, 5.0, 1.0, 0.4
    floor = scenes.flat_mesh(scenes.quad((-3, 0, -3), (3, 0, -3), (3, 0, 3), (-3, 0, 3), (0, 1, 0)))
    poly = [(-a, hgt, -a + 0.3), (a, hgt, -a + 0.3), (a, hgt, a + 0.3), (-a, hgt, a + 0.3)]
    light = scenes.flat_mesh(scenes.quad(*poly, (0, -1, 0)))
    prims = [scenes.Primitive(floor, scenes.Material(scenes.MAT_DIFFUSE, (rho,) * 3)),
             scenes.Primitive(light, scenes.Material(scenes.MAT_DIFFUSE, (0, 0, 0)), Ke=(Le,) * 3)]
    cam = scenes.Camera((0, 0.5, 0), (0, 0, 0), (0, 0, -1), 0.001, 1.0)
    flat = scenes.Scene(prims, cam, None, 1, 1).flatten()
    spp = 400000
    r = orc.pt_render(flat, spp=spp, max_verts=3, sampling_mode=mode, seed=9)
    mean = r['sum'][0, 0, 0] / spp
    se = math.sqrt((r['sumsq'][0, 0, 0] / spp - mean ** 2) / spp)
    E = _polygon_irradiance(np.zeros(3), np.array([0, 1.0, 0]), poly, Le)
    expected = rho / math.pi * E
    assert abs(mean - expected) < 5 * se + 1e-9, (mean, expected, se)


# ---- closed-form pins of material::glossy and material::glass -----------------------
#
# material::glossy (src/material/material_glossy.cpp:43-58,102-109) is a GGX
# microfacet BRDF f = Ks * F(wo.h) * D(h) * G1(wi) G1(wo) / (4 cos_i cos_o)
# with Schlick F = Ks + (1 - Ks)(1 - wo.h)^5, i.e. Ks enters twice. The
# checks below write D and G1 in the spherical-angle form of Walter et al.
# (EGSR 2007, eq. 33-34) and Heitz (JCGT 3(2), 2014, eq. 85-86) for the
# anisotropic case, not in the oracle's dot-product form. With N = (0,0,1) the
# reference's two-sided basis (include/lm/math.h:539-546, surface.h:192-195)
# is u = x, v = y, so ax belongs to the x axis.

def _ggx_walter(ks, ax, ay, wi, wo):
    """GGX BRDF and projected-solid-angle pdf from angles (N = z)."""
    ks = np.asarray(ks, dtype=np.float64)
    h = (wi + wo) / np.linalg.norm(wi + wo)
    th = math.acos(h[2])
    ph = math.atan2(h[1], h[0])
    t2 = math.tan(th) ** 2
    D = 1.0 / (math.pi * ax * ay * math.cos(th) ** 4 *
               (1.0 + t2 * (math.cos(ph) ** 2 / ax ** 2 + math.sin(ph) ** 2 / ay ** 2)) ** 2)

    def G1(w):
        tw = math.acos(w[2])
        pw = math.atan2(w[1], w[0])
        a2 = math.cos(pw) ** 2 * ax ** 2 + math.sin(pw) ** 2 * ay ** 2
        return 2.0 / (1.0 + math.sqrt(1.0 + a2 * math.tan(tw) ** 2))

    F = ks + (1 - ks) * (1 - float(np.dot(wo, h))) ** 5
    f = ks * F * D * G1(wi) * G1(wo) / (4 * wi[2] * wo[2])
    pdf = D * h[2] / (4 * float(np.dot(wo, h)) * wo[2])
    return f, pdf


def test_glossy_normal_incidence_closed_form():
    """wi = wo = n: h = n, D = 1/(pi ax ay), G1(n) = 1, F = Ks, so
    f = Ks^2 / (4 pi ax ay) per channel (a single Ks would give Ks / ...)."""
    ks, ax, ay = np.array([0.9, 0.6, 0.3]), 0.3, 0.2
    f, pdf = orc.mat_eval(scenes.MAT_GLOSSY, orc._mat_params(ks, ax, ay), N, N[None], N[None])
    assert np.allclose(f[0], ks ** 2 / (4 * math.pi * ax * ay), rtol=1e-12)
    assert abs(pdf[0] - 1 / (4 * math.pi * ax * ay)) < 1e-12 * pdf[0]


@pytest.mark.parametrize('theta_deg,alpha', [(50.0, 0.25), (80.0, 0.05), (20.0, 0.5)])
def test_glossy_specular_direction_isotropic(theta_deg, alpha):
    """Mirror configuration (h = n) at an oblique angle, isotropic alpha:
    f = Ks (Ks + (1-Ks)(1-cos)^5) G1^2 / (4 pi alpha^2 cos^2) with the Smith
    G1 = 2 / (1 + sqrt(1 + alpha^2 tan^2)) of Walter et al.; the max(Eps, sin)
    clamp of material_glossy.cpp:52 does not bite at these angles."""
    th = math.radians(theta_deg)
    wi = np.array([math.sin(th), 0.0, math.cos(th)])
    wo = np.array([-math.sin(th), 0.0, math.cos(th)])
    ks = np.array([0.8, 0.5, 0.2])
    g1 = 2 / (1 + math.sqrt(1 + alpha ** 2 * math.tan(th) ** 2))
    fr = ks + (1 - ks) * (1 - math.cos(th)) ** 5
    exp = ks * fr * g1 * g1 / (4 * math.pi * alpha ** 2 * math.cos(th) ** 2)
    f, pdf = orc.mat_eval(scenes.MAT_GLOSSY, orc._mat_params(ks, alpha, alpha), N, wi[None], wo[None])
    assert np.allclose(f[0], exp, rtol=1e-10), (f[0], exp)
    assert abs(pdf[0] - 1 / (4 * math.pi * alpha ** 2 * math.cos(th) ** 2)) < 1e-10 * pdf[0]


def test_glossy_anisotropic_off_specular_closed_form():
    """Random off-specular pairs against the spherical-angle anisotropic GGX
    (Heitz 2014), which pins the axis each roughness belongs to (swapping ax
    and ay changes f by up to 3x here) and the uncorrelated Smith product."""
    rng = np.random.default_rng(11)
    ks, ax, ay = (0.7, 0.4, 0.9), 0.15, 0.45
    for _ in range(200):
        a = rng.normal(size=3); a[2] = abs(a[2]) + 0.05; a /= np.linalg.norm(a)
        b = rng.normal(size=3); b[2] = abs(b[2]) + 0.05; b /= np.linalg.norm(b)
        f, pdf = orc.mat_eval(scenes.MAT_GLOSSY, orc._mat_params(ks, ax, ay), N, a[None], b[None])
        ef, ep = _ggx_walter(ks, ax, ay, a, b)
        assert np.allclose(f[0], ef, rtol=1e-9), (a, b, f[0], ef)
        assert abs(pdf[0] - ep) <= 1e-9 * ep
    # the roughness axes matter: swapped ax / ay gives another value
    a = np.array([0.6, 0.1, 0.79]); a /= np.linalg.norm(a)
    b = np.array([-0.2, 0.5, 0.84]); b /= np.linalg.norm(b)
    f1, _ = orc.mat_eval(scenes.MAT_GLOSSY, orc._mat_params(ks, ax, ay), N, a[None], b[None])
    f2, _ = _ggx_walter(ks, ay, ax, a, b)
    assert not np.allclose(f1[0], f2, rtol=1e-3)


def _glass_reflect_fraction_brackets(wi, fr, ni=1.5):
    """material::glass picks reflection when uc < Fr (material_glass.cpp:
    124-141): components drawn just below and just above the expected Fr."""
    us = np.array([[fr * (1 - 1e-9), 0, 0, 0], [min(fr * (1 + 1e-9), 1 - 1e-12), 0, 0, 0]])
    wo, w, comp, ok = orc.mat_sample(scenes.MAT_GLASS, orc._mat_params(ni=ni), N, np.repeat(wi[None], 2, 0), us)
    return comp, wo, w


def test_glass_fresnel_entering_oblique():
    """Entering at 60 deg: Schlick with cos = wi.n = 0.5, r0 = 0.04:
    Fr = 0.04 + 0.96 * 0.5^5 = 0.07."""
    th = math.radians(60.0)
    wi = np.array([math.sin(th), 0.0, math.cos(th)])
    comp, wo, w = _glass_reflect_fraction_brackets(wi, 0.07)
    assert comp.tolist() == [0, 1]


def test_glass_fresnel_leaving_uses_transmitted_cosine():
    """Leaving the glass (wi.n < 0) at 30 deg inside: Snell gives
    sin(theta_t) = 1.5 * 0.5 = 0.75 outside, and the reference's Schlick uses
    cos = wt.n = sqrt(1 - 0.5625) (material_glass.cpp:90-98):
    Fr = 0.04 + 0.96 (1 - 0.661437828)^5 = 0.044270... Using |wi.n| = cos 30
    instead would give 0.040041, which the brackets exclude."""
    th = math.radians(30.0)
    wi = np.array([math.sin(th), 0.0, -math.cos(th)])
    cos_t = math.sqrt(1 - (1.5 * math.sin(th)) ** 2)
    fr = 0.04 + 0.96 * (1 - cos_t) ** 5
    assert abs(fr - 0.0442704) < 1e-6
    comp, wo, w = _glass_reflect_fraction_brackets(wi, fr)
    assert comp.tolist() == [0, 1]
    wrong = 0.04 + 0.96 * (1 - math.cos(th)) ** 5
    comp_w, _, _ = _glass_reflect_fraction_brackets(wi, wrong)
    assert comp_w.tolist() == [0, 0]           # both draws lie below the true Fr: reflection
    # the refracted direction leaves into the upper side with sin = 0.75, and
    # the refraction weight is eta^2 = Ni^2 for importance transport
    t = wo[1]
    assert t[2] > 0 and abs(math.hypot(t[0], t[1]) - 0.75) < 1e-12
    assert np.allclose(w[1], 1.5 ** 2)


def test_glass_total_internal_reflection():
    """Leaving at 60 deg inside: 1.5 sin 60 > 1, total internal reflection,
    Fr = 1: every component draw reflects, with weight 1 about the normal."""
    th = math.radians(60.0)
    wi = np.array([math.sin(th), 0.0, -math.cos(th)])
    us = np.array([[0.0, 0, 0, 0], [0.5, 0, 0, 0], [1 - 1e-12, 0, 0, 0]])
    wo, w, comp, ok = orc.mat_sample(scenes.MAT_GLASS, orc._mat_params(ni=1.5), N, np.repeat(wi[None], 3, 0), us)
    assert comp.tolist() == [0, 0, 0] and ok.all()
    assert np.allclose(wo, [[-math.sin(th), 0.0, -math.cos(th)]] * 3)
    assert np.allclose(w, 1.0)
