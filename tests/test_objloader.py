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
    """Default build: every non-specular MTL material is
    material::mixture_wavefrontobj (Kd, Ks, ax, ay); NO_MIXTURE_MATERIAL:
    diffuse when Ks == 0, else glossy; skip_specular_mat: black diffuse for
    illum 5 / 7 and no alpha mask."""
    r = 2.0 / (2.0 + 98.0)
    a_s = math.sqrt(1.0 - 0.5 * 0.9)
    prims = objloader.load_obj(write(tmp_path))
    white, shiny, glass, mirror, lamp = (p.material for p in prims)
    assert white.type == scenes.MAT_MIXTURE and white.color == (0.5, 0.5, 0.5) and white.ks == (0.0, 0.0, 0.0)
    assert shiny.type == scenes.MAT_MIXTURE and shiny.color == (0.1, 0.1, 0.1) and shiny.ks == (0.7, 0.6, 0.5)
    assert shiny.ax == pytest.approx(r / a_s) and shiny.ay == pytest.approx(r * a_s) and shiny.alpha_mask
    assert glass.type == scenes.MAT_GLASS and glass.ni == pytest.approx(1.45)
    assert mirror.type == scenes.MAT_MIRROR
    assert prims[4].Ke == (4.0, 5.0, 6.0) and prims[0].Ke is None
    white, shiny, glass, mirror, lamp = (p.material for p in objloader.load_obj(write(tmp_path), no_mixture=True))
    assert white.type == scenes.MAT_DIFFUSE and white.color == (0.5, 0.5, 0.5)
    assert shiny.type == scenes.MAT_GLOSSY and shiny.color == (0.7, 0.6, 0.5)
    assert shiny.ax == pytest.approx(r / a_s) and shiny.ay == pytest.approx(r * a_s)
    white, shiny, glass, mirror, lamp = (p.material for p in
                                         objloader.load_obj(write(tmp_path), skip_specular_mat=True))
    assert glass.type == scenes.MAT_DIFFUSE and glass.color == (0.0, 0.0, 0.0) and mirror.color == (0.0, 0.0, 0.0)
    assert shiny.type == scenes.MAT_MIXTURE and not shiny.alpha_mask
    # roughness floor of 1e-3 for very high Ns
    m = objloader.material_of(objloader.MtlParams('x', Ks=(1, 1, 1), Ns=1e9))
    assert m.ax == pytest.approx(1e-3) and m.ay == pytest.approx(1e-3)


def test_missing_textures_and_bad_indices_are_errors(tmp_path):
    p = write(tmp_path, OBJ + "usemtl tex\nf 1 2 3\n")
    with pytest.raises(OSError):
        objloader.load_obj(p)   # wood.png does not exist
    prims = objloader.load_obj(p, ignore_textures=True)
    assert prims[-1].material.color == (0.3, 0.3, 0.3) and prims[-1].material.texture is None
    with pytest.raises(ValueError):
        objloader.load_obj(write(tmp_path, "v 0 0 0\nv 1 0 0\nf 1 2 5\n"))
    with pytest.raises(ValueError):
        objloader.load_obj(write(tmp_path, "v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nf 1/1 2/2 3/1\n"))


def test_texture_loading_follows_stbi_loadf(tmp_path):
    """map_Kd through texture::bitmap's stbi_loadf: colour (c/255)^2.2, alpha
    c/255, rows flipped; texture coordinates from vt per corner; one texture
    per file name."""
    from PIL import Image
    img = np.zeros((2, 3, 4), dtype=np.uint8)
    img[0, 0] = (255, 0, 0, 255)      # top-left in the file
    img[1, 2] = (10, 128, 200, 64)    # bottom-right in the file
    Image.fromarray(img, 'RGBA').save(tmp_path / 'leaf.png')
    mtl = "newmtl leaf\nKd 1 1 1\nKs 0 0 0\nmap_Kd leaf.png\nnewmtl leaf2\nKd 1 1 1\nmap_Kd leaf.png\n"
    obj = ("mtllib scene.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 0.25 0.75\n"
           "usemtl leaf\nf 1/1 2/2 3/3\nusemtl leaf2\nf 1/3 2/2 3/1\n")
    prims = objloader.load_obj(write(tmp_path, obj, mtl))
    t = prims[0].material.texture
    assert t is not None and t is prims[1].material.texture
    a = t.data
    assert a.shape == (2, 3, 4)
    # flipped: file row 1 is texture row 0
    assert np.allclose(a[0, 2, :3], (np.array([10, 128, 200]) / 255.0) ** 2.2, rtol=1e-6)
    assert a[0, 2, 3] == pytest.approx(64 / 255.0)
    assert np.allclose(a[1, 0], (1, 0, 0, 1))
    assert np.allclose(prims[0].mesh.corner_uvs()[0], [0, 0, 1, 0, 0.25, 0.75])
    assert np.allclose(prims[1].mesh.corner_uvs()[0], [0.25, 0.75, 1, 0, 0, 0])
    flat = scenes.Scene(prims, scenes.Camera((0, 0, 3), (0, 0, 0), (0, 1, 0), 40, 1), (1, 1, 1), 8, 8).flatten()
    info, data = flat.texture_tables()
    assert info.tolist() == [[3, 2, 4, 0]] and data.size == 24 and flat.mat_tex.tolist() == [0, 0]


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
