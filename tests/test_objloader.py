"""Wavefront OBJ / MTL ingestion (model::wavefrontobj): parsing, material
mapping of model_wavefrontobj.cpp:155-215, area lights from Ke, and a GPU
render of an OBJ scene against its closed form."""
import math

import numpy as np
import pytest

from lm_amd import objloader, scenes

MTL = """
newmtl white
Kd 0.5 0.5 0.5
Ks 0 0 0
newmtl shiny
Kd 0.1 0.1 0.1
Ks 0.7 0.6 0.5
Ns 98
an 0.5
newmtl glass
illum 7
Ni 1.45
newmtl mirror
illum 5
newmtl lamp
Kd 0 0 0
Ke 4 5 6
newmtl tex
Kd 0.3 0.3 0.3
map_Kd wood.png
"""

OBJ = """
mtllib scene.mtl
v -1 -1 0
v 1 -1 0
v 1 1 0
v -1 1 0
vn 0 0 1
o floor
usemtl white
f 1//1 2//1 3//1 4//1
o panel
usemtl shiny
f -4 -3 -2
usemtl glass
f 1 3 4
usemtl mirror
f 2 3 4
usemtl lamp
f 1/1/1 2/1/1 3/1/1
"""


def write(tmp_path, obj=OBJ, mtl=MTL):
    (tmp_path / 'scene.mtl').write_text(mtl)
    p = tmp_path / 'scene.obj'
    p.write_text(obj)
    return str(p)


def test_parse_groups_and_triangulation(tmp_path):
    prims = objloader.load_obj(write(tmp_path))
    assert len(prims) == 5
    # quad triangulated as a fan with the given vertex normals
    q = prims[0].mesh
    assert q.num_triangles == 2
    assert np.allclose(q.ns, [0, 0, 1])
    assert np.allclose(q.triangles()[1], [-1, -1, 0, 1, 1, 0, -1, 1, 0])
    # negative indices, no normals -> geometric normal
    t = prims[1].mesh
    assert np.allclose(t.triangles()[0], [-1, -1, 0, 1, -1, 0, 1, 1, 0])
    assert np.allclose(t.ns, [0, 0, 1])


def test_material_mapping_follows_the_reference(tmp_path):
    prims = objloader.load_obj(write(tmp_path))
    white, shiny, glass, mirror, lamp = (p.material for p in prims)
    assert white.type == scenes.MAT_DIFFUSE and white.color == (0.5, 0.5, 0.5)
    r = 2.0 / (2.0 + 98.0)
    a_s = math.sqrt(1.0 - 0.5 * 0.9)
    assert shiny.type == scenes.MAT_GLOSSY and shiny.color == (0.7, 0.6, 0.5)
    assert shiny.ax == pytest.approx(r / a_s) and shiny.ay == pytest.approx(r * a_s)
    assert glass.type == scenes.MAT_GLASS and glass.ni == pytest.approx(1.45)
    assert mirror.type == scenes.MAT_MIRROR
    assert prims[4].Ke == (4.0, 5.0, 6.0) and prims[0].Ke is None
    # roughness floor of 1e-3 for very high Ns
    m = objloader.material_of(objloader.MtlParams('x', Ks=(1, 1, 1), Ns=1e9))
    assert m.ax == pytest.approx(1e-3) and m.ay == pytest.approx(1e-3)


def test_textures_and_bad_indices_are_errors(tmp_path):
    p = write(tmp_path, OBJ + "usemtl tex\nf 1 2 3\n")
    with pytest.raises(NotImplementedError):
        objloader.load_obj(p)
    prims = objloader.load_obj(p, ignore_textures=True)
    assert prims[-1].material.color == (0.3, 0.3, 0.3)
    with pytest.raises(ValueError):
        objloader.load_obj(write(tmp_path, "v 0 0 0\nv 1 0 0\nf 1 2 5\n"))


BOX_OBJ = """
mtllib box.mtl
usemtl wall
v -1 -1 -1
v 1 -1 -1
v -1 1 -1
v 1 1 -1
v -1 -1 1
v 1 -1 1
v -1 1 1
v 1 1 1
f 1 2 4 3
f 5 7 8 6
f 1 5 6 2
f 3 4 8 7
f 1 3 7 5
f 2 6 8 4
"""


@pytest.mark.gpu
def test_obj_furnace_renders_closed_form(device, tmp_path):
    """A closed OBJ box whose walls emit Ke = 1 and reflect Kd = 0.5 renders
    E[L] = sum_k 0.5^k (max 8 vertices) on the GPU."""
    (tmp_path / 'box.mtl').write_text("newmtl wall\nKd 0.5 0.5 0.5\nKe 1 1 1\n")
    p = tmp_path / 'box.obj'
    p.write_text(BOX_OBJ)
    cam = scenes.Camera((0, 0, 0.5), (0, 0, -1), (0, 1, 0), 60.0, 1.0)
    flat = objloader.scene_from_obj(str(p), cam, 32, 32).flatten()
    from lm_amd.renderer import GpuScene
    spp, M = 64, 8
    s, q, _ = GpuScene(flat).render_raw(spp=spp, max_verts=M, seed=9)
    n = 32 * 32 * spp
    mean = s[..., 0].sum() / n
    var = q[..., 0].sum() / n - mean ** 2
    expected = sum(0.5 ** k for k in range(M - 1))
    z = (mean - expected) / math.sqrt(var / n)
    print(f'OBJ furnace mean {mean:.6f} expected {expected:.6f} z {z:.2f}')
    assert abs(z) < 5
