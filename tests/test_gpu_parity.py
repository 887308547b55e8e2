"""HIP render path vs the CPU oracle, on the same scenes, cameras and RNG
seeds.  Every call goes through the C ABI of libasuna.so.

Bars (stated per test):
  first hit     primitive / instance id identical for >= 99.99 % of pixels;
                t and shading normal within 1e-5 relative (observed: bit-exact)
  radiance      per-pixel |gpu - oracle| <= 3 sigma of the per-pixel Monte
                Carlo standard error for >= 99 % of pixels at equal spp, and
                image means within 0.5 %
  AOV / weights frame-0 channels equal to 1e-5; weight image 1e-6 relative
"""
import os

import numpy as np
import pytest

import asuna_amd as A
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def renderer():
    r = A.Renderer(0)
    yield r
    r.close()


def _load(r, path, branching=4, node_format=0):
    s = A.HostScene(path)
    r.upload(s)
    info = r.build(branching=branching, node_format=node_format)
    w, h = s.film
    r.set_film(w, h)
    return s, info


def _first_hit_parity(r, s, osc, cam):
    w, h = s.film
    g = r.first_hit(cam)
    o = osc.first_hit(cam, w, h)
    prim = np.mean((g["prim_id"] == o["prim_id"]) & (g["inst_id"] == o["inst_id"]))
    hit = o["inst_id"] >= 0
    rel_t = np.abs(g["t"][hit] - o["t"][hit]) / np.maximum(np.abs(o["t"][hit]), 1e-30)
    same = hit & (g["prim_id"] == o["prim_id"]) & (g["inst_id"] == o["inst_id"])
    dn = np.abs(g["normal"][same] - o["normal"][same]).max(initial=0.0)
    log(f"first hit: id match {prim:.6f}, max rel t {rel_t.max(initial=0):.3g}, max |dN| {dn:.3g}, "
        f"bit-exact t {np.mean(g['t'] == o['t']):.6f}")
    return prim, rel_t, dn, same, g, o


def test_first_hit_box_bit_exact(renderer, box_scene):
    s, _ = _load(renderer, box_scene)
    cam, _ = s.shot(0)
    osc = O.OracleScene(s.desc)
    prim, rel_t, dn, same, g, o = _first_hit_parity(renderer, s, osc, cam)
    assert prim >= 0.9999
    assert rel_t.max(initial=0) <= 1e-5
    assert dn <= 1e-5
    np.testing.assert_allclose(g["pos"][same], o["pos"][same], rtol=1e-5, atol=1e-6)


def test_first_hit_spheres_vs_oracle_bvh(renderer, spheres_small):
    s, info = _load(renderer, spheres_small)
    cam, _ = s.shot(0)
    osc = O.OracleScene(s.desc)
    prim, rel_t, dn, same, g, o = _first_hit_parity(renderer, s, osc, cam)
    assert prim >= 0.9999
    assert rel_t.max(initial=0) <= 1e-5
    assert dn <= 1e-5


@pytest.mark.parametrize("branching,node_format", [(2, 0), (4, 1), (4, 2)])
def test_first_hit_oblique_cameras(renderer, spheres_small, branching, node_format):
    """Random poses (rays at every angle to the BVH slabs): binary, float
    4-wide and quantized 4-wide trees."""
    s, _ = _load(renderer, spheres_small, branching, node_format)
    osc = O.OracleScene(s.desc)
    rng = np.random.default_rng(11)
    w, h = s.film
    for _ in range(3):
        eye = rng.normal(size=3)
        eye = eye / np.linalg.norm(eye) * rng.uniform(1.5, 4.0)
        eye[1] = abs(eye[1]) + 0.2
        cam = A.camera_lookat(w, h, float(rng.uniform(30, 80)), tuple(eye), (0.0, 0.4, 0.0))
        prim, rel_t, dn, *_ = _first_hit_parity(renderer, s, osc, cam)
        assert prim >= 0.9999 and rel_t.max(initial=0) <= 1e-5 and dn <= 1e-5


@pytest.mark.parametrize("node_format", [1, 2])
def test_wide_and_binary_trees_render_identically(renderer, spheres_small, node_format):
    """The 4-wide walk (nearest-first, distance-culled stack) finds the same
    closest hit as the binary walk, so whole renders are bit-identical; it
    visits fewer nodes.  Quantized nodes (boxes rounded outwards) too."""
    s, i2 = _load(renderer, spheres_small, 2)
    cam, pc = s.shot(0)
    w, h = s.film
    renderer.set_film(w, h)
    st2 = renderer.render(cam, pc, 0, 4, count_traversal=True)
    img2 = [renderer.read(i) for i in (0, 1, 2, 8)]
    hit2 = renderer.first_hit(cam)
    s, i4 = _load(renderer, spheres_small, 4, node_format)
    renderer.set_film(w, h)
    st4 = renderer.render(cam, pc, 0, 4, count_traversal=True)
    img4 = [renderer.read(i) for i in (0, 1, 2, 8)]
    hit4 = renderer.first_hit(cam)
    log(f"binary {i2}\nwide {i4}\nnodes/ray {st2['ext_node_visits'] / st2['extension_rays']:.2f} -> "
        f"{st4['ext_node_visits'] / st4['extension_rays']:.2f}")
    assert i2["branching"] == 2 and i2["num_wide_nodes"] == 0
    assert i4["branching"] == 4 and 0 < i4["num_wide_nodes"] < i4["num_nodes"]
    assert i4["node_bytes"] == (128 if node_format == 1 else 64) and i2["node_bytes"] == 64
    assert i4["wide_depth"] < i4["max_depth"]
    for a, b in zip(img2, img4):
        np.testing.assert_array_equal(a, b)
    for k in ("t", "prim_id", "inst_id", "normal"):
        np.testing.assert_array_equal(hit2[k], hit4[k])
    assert st2["extension_rays"] == st4["extension_rays"] and st2["shadow_rays"] == st4["shadow_rays"]
    assert st4["ext_node_visits"] < st2["ext_node_visits"]


@pytest.mark.parametrize("branching", [2, 4])
def test_ploc_and_lbvh_trees_render_identically(renderer, spheres_small, branching):
    """The clustered (PLOC) tree and the LBVH hold the same triangles under
    different boxes; the conservative walks find the same closest and any
    hits, so renders and first hits are bit-identical, and the clustered
    tree (lower SAH cost) is walked with fewer node visits."""
    s = A.HostScene(spheres_small)
    cam, pc = s.shot(0)
    w, h = s.film
    out = {}
    for q in (1, 2):
        renderer.upload(s)
        info = renderer.build(quality=q, branching=branching)
        renderer.set_film(w, h)
        st = renderer.render(cam, pc, 0, 4, count_traversal=True)
        out[q] = (info, st, [renderer.read(i) for i in (0, 1, 2, 8)], renderer.first_hit(cam))
        log(f"branching {branching} quality {q}: sah {info['sah_cost']:.3f} nodes/ray "
            f"{st['ext_node_visits'] / st['extension_rays']:.2f} shadow nodes/ray "
            f"{st['shadow_node_visits'] / max(1, st['shadow_rays']):.2f}")
    for a, b in zip(out[1][2], out[2][2]):
        np.testing.assert_array_equal(a, b)
    for k in ("t", "prim_id", "inst_id", "normal"):
        np.testing.assert_array_equal(out[1][3][k], out[2][3][k])
    assert out[2][0]["sah_cost"] < out[1][0]["sah_cost"]
    assert out[2][1]["ext_node_visits"] < out[1][1]["ext_node_visits"]


def _radiance_parity(r, s, spp, crop=None, precision=0):
    cam, pc = s.shot(0)
    w, h = s.film
    r.set_film(w, h)
    dch = s.state.depth_out_channel
    st = r.render(cam, pc, 0, spp, depth_channel=dch)
    gimg = np.stack([r.read(i) for i in range(9)])
    osc = O.OracleScene(s.desc)
    sq = np.zeros((h, w, 4), np.float32)
    oimg, ost = osc.render(cam, pc, w, h, 0, spp, sq=sq, depth_channel=dch, crop=crop)
    return st, ost, gimg, oimg, sq


def _three_sigma(g, o, sq, spp, mask=None):
    mean = o[..., :3]
    var = np.maximum(sq[..., :3] / spp - mean ** 2, 0.0)
    se = np.sqrt(2 * var / spp) + 1e-6  # difference of two independent means
    bad = np.abs(g[..., :3] - mean) > 3 * se + 1e-5
    if mask is not None:
        bad = bad[mask]
    frac = bad.any(-1).mean()
    return frac


def test_radiance_box_matches_oracle(renderer, box_scene):
    """The 512x512 box workload at its stated 64 spp, depth 8."""
    s, _ = _load(renderer, box_scene)
    spp = 64
    assert s.shot(0)[1].maxPathDepth == 8
    st, ost, g, o, sq = _radiance_parity(renderer, s, spp)
    frac_bad = _three_sigma(g[0], o[0], sq, spp)
    mg, mo = g[0][..., :3].mean((0, 1)), o[0][..., :3].mean((0, 1))
    exact = np.mean((g[0][..., :3] == o[0][..., :3]).all(-1))
    log(f"box radiance: outside 3 sigma {frac_bad:.5f}, bit-exact pixels {exact:.3f}, means {mg} vs {mo}")
    assert frac_bad <= 0.01
    np.testing.assert_allclose(mg, mo, rtol=5e-3)
    for c in range(1, 1 + 3):  # diffuse, normal, position AOVs (frame 0)
        np.testing.assert_allclose(g[c], o[c], rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(g[8], o[8], rtol=1e-6)
    assert abs(st["extension_rays"] - ost["extension_rays"]) / ost["extension_rays"] < 1e-3


def test_radiance_materials_scene_matches_oracle(renderer, spheres_small):
    """pbr metal/dielectric spheres, lambertian ground, env map with importance
    sampling, 4 triangle mesh lights, depth 6, depth extension channel."""
    s, _ = _load(renderer, spheres_small)
    spp = 16
    st, ost, g, o, sq = _radiance_parity(renderer, s, spp)
    frac_bad = _three_sigma(g[0], o[0], sq, spp)
    mg, mo = g[0][..., :3].mean((0, 1)), o[0][..., :3].mean((0, 1))
    log(f"spheres radiance: outside 3 sigma {frac_bad:.5f}, means {mg} vs {mo}; rays gpu {st['extension_rays']} "
        f"oracle {ost['extension_rays']}, dielectric shadow {st['dielectric_shadow_rays']}")
    assert frac_bad <= 0.01
    np.testing.assert_allclose(mg, mo, rtol=1e-2)
    for c in (1, 2, 3):  # diffuse, normal, depth
        np.testing.assert_allclose(g[c], o[c], rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(g[8], o[8], rtol=1e-6)


def test_rough_plastic_and_emissive_scene(renderer, tmp_path):
    """Rough plastic (the reference's permuted eval/pdf arguments), emissive
    material, point + rect lights, textured lambertian floor."""
    d = tmp_path / "plastic"
    d.mkdir()
    sp = A.generate_scene(0, str(d))
    tex = np.zeros((16, 16, 4), np.float32)
    tex[..., 3] = 1
    tex[::2, ::2, :3] = 0.8
    tex[1::2, 1::2, :3] = 0.3
    A.write_image(str(d / "checker.exr"), tex)
    import json
    js = json.load(open(sp))
    js["textures"] = [{"name": "checker", "path": "checker.exr"}]
    js["materials"][0] = {"name": js["materials"][0]["name"], "type": "brdf_lambertian",
                          "diffuse_texture": "checker"}
    js["materials"][1] = {"name": js["materials"][1]["name"], "type": "brdf_rough_plastic",
                          "diffuse_reflectance": [0.6, 0.3, 0.2], "alpha": [0.15, 0.4], "ior": 1.6}
    js["materials"][5] = {"name": js["materials"][5]["name"], "type": "brdf_rough_plastic",
                          "diffuse_reflectance": [0.2, 0.5, 0.7], "alpha": [0.05, 0.05]}
    js["materials"][6] = {"name": js["materials"][6]["name"], "type": "brdf_emissive", "radiance": [2, 1, 0.5]}
    js["lights"].append({"type": "point", "radiance": [1.5, 1.5, 1.5], "position": [0.5, 0.5, 0.5]})
    js["lights"].append({"type": "distant", "radiance": [0.3, 0.3, 0.3], "direction": [0.2, 1.0, 0.4]})
    json.dump(js, open(sp, "w"))
    s, _ = _load(renderer, sp)
    spp = 16
    st, ost, g, o, sq = _radiance_parity(renderer, s, spp)
    frac_bad = _three_sigma(g[0], o[0], sq, spp)
    mg, mo = g[0][..., :3].mean((0, 1)), o[0][..., :3].mean((0, 1))
    log(f"plastic radiance: outside 3 sigma {frac_bad:.5f}, means {mg} vs {mo}")
    assert frac_bad <= 0.01
    np.testing.assert_allclose(mg, mo, rtol=5e-3)


def test_normal_mapped_scene_matches_oracle(renderer, tmp_path):
    """Normal maps (brdf_lambertian.rchit:83-90 and the same block in the
    other plugins): a checker of two tilted tangent-space normals on a
    lambertian, a rough plastic and a pbr surface of the box scene."""
    d = tmp_path / "nmap"
    d.mkdir()
    sp = A.generate_scene(0, str(d))
    tex = np.ones((16, 16, 4), np.float32)
    tex[..., :3] = (0.5, 0.5, 1.0)
    tex[::2, ::2, :3] = (0.8125, 0.5, 0.875)
    tex[1::2, 1::2, :3] = (0.5, 0.25, 0.875)
    A.write_image(str(d / "bumps.exr"), tex)
    import json
    js = json.load(open(sp))
    js["textures"] = [{"name": "bumps", "path": "bumps.exr"}]
    mats = js["materials"]
    mats[0] = {"name": mats[0]["name"], "type": "brdf_lambertian", "diffuse_reflectance": [0.6, 0.6, 0.5],
               "normal_texture": "bumps"}
    mats[1] = {"name": mats[1]["name"], "type": "brdf_rough_plastic", "diffuse_reflectance": [0.5, 0.3, 0.2],
               "alpha": [0.2, 0.2], "normal_texture": "bumps"}
    mats[2] = {"name": mats[2]["name"], "type": "brdf_pbr_metalness_roughness", "diffuse_reflectance": [0.7, 0.5, 0.3],
               "metalness": 0.5, "roughness": 0.4, "normal_texture": "bumps"}
    json.dump(js, open(sp, "w"))
    s, _ = _load(renderer, sp)
    spp = 16
    st, ost, g, o, sq = _radiance_parity(renderer, s, spp)
    frac_bad = _three_sigma(g[0], o[0], sq, spp)
    mg, mo = g[0][..., :3].mean((0, 1)), o[0][..., :3].mean((0, 1))
    log(f"normal-mapped radiance: outside 3 sigma {frac_bad:.5f}, means {mg} vs {mo}")
    assert frac_bad <= 0.01
    np.testing.assert_allclose(mg, mo, rtol=5e-3)


def test_more_plugins_scene_matches_oracle(renderer, tmp_path):
    """Mirror, conductor (Au), rough conductor (Cu, anisotropic), smooth
    plastic, phong and kang18 surfaces in the box, lit by its rect light plus
    a point light (delta light paths through every plugin's pdf)."""
    d = tmp_path / "more"
    d.mkdir()
    sp = A.generate_scene(0, str(d))
    import json
    js = json.load(open(sp))
    mats = js["materials"]
    mats[0] = {"name": mats[0]["name"], "type": "brdf_mirror", "diffuse_reflectance": [0.9, 0.85, 0.8]}
    mats[1] = {"name": mats[1]["name"], "type": "brdf_conductor", "material": "Au"}
    mats[2] = {"name": mats[2]["name"], "type": "brdf_rough_conductor", "material": "Cu", "alpha": [0.2, 0.35],
               "diffuse_reflectance": [1, 1, 1]}
    mats[3] = {"name": mats[3]["name"], "type": "brdf_plastic", "diffuse_reflectance": [0.2, 0.4, 0.7], "ior": 1.6}
    mats[4] = {"name": mats[4]["name"], "type": "brdf_phong", "diffuse_reflectance": [0.5, 0.3, 0.2],
               "specular_reflectance": [0.3, 0.3, 0.3], "shininess": 30}
    mats[5] = {"name": mats[5]["name"], "type": "brdf_kang18", "diffuse_reflectance": [0.3, 0.5, 0.3],
               "specular_reflectance": [0.5, 0.5, 0.5], "alpha": [0.2, 0.4]}
    for m in mats:
        if m["type"] in ("brdf_conductor",):
            m["diffuse_reflectance"] = [1, 1, 1]
    js["lights"].append({"type": "point", "radiance": [1.0, 1.0, 1.0], "position": [0.3, 0.6, 0.2]})
    json.dump(js, open(sp, "w"))
    s, _ = _load(renderer, sp)
    spp = 16
    st, ost, g, o, sq = _radiance_parity(renderer, s, spp)
    frac_bad = _three_sigma(g[0], o[0], sq, spp)
    mg, mo = g[0][..., :3].mean((0, 1)), o[0][..., :3].mean((0, 1))
    log(f"more plugins: outside 3 sigma {frac_bad:.5f}, means {mg} vs {mo}; rays gpu {st['extension_rays']} "
        f"oracle {ost['extension_rays']}, zero-NEE shadow {st['dielectric_shadow_rays']} vs "
        f"{ost['dielectric_shadow_rays']}")
    assert frac_bad <= 0.01
    np.testing.assert_allclose(mg, mo, rtol=1e-2)
    assert abs(st["extension_rays"] - ost["extension_rays"]) / ost["extension_rays"] < 1e-3
    assert st["dielectric_shadow_rays"] > 0


def test_disney_scene_matches_oracle(renderer, tmp_path):
    """Disney BRDF surfaces: diffuse with retro-reflection, metal, partial
    opacity (pass-through draws) and -- through the material records, since
    the JSON loader reads only base colour, metallic, roughness and opacity --
    subsurface, sheen, tints, anisotropy and clearcoat; rect + point light."""
    d = tmp_path / "disney"
    d.mkdir()
    sp = A.generate_scene(0, str(d))
    import json
    js = json.load(open(sp))
    mats = js["materials"]
    for k, (base, metal, rough, opac) in enumerate([((0.8, 0.7, 0.6), 0.0, 0.6, 0.0), ((0.9, 0.6, 0.3), 1.0, 0.3, 0.0),
                                                    ((0.3, 0.5, 0.8), 0.2, 0.5, 0.0), ((0.7, 0.7, 0.2), 0.0, 0.2, 0.0),
                                                    ((0.5, 0.3, 0.6), 0.5, 0.7, 0.3), ((0.2, 0.6, 0.4), 0.0, 0.4, 0.0)]):
        mats[k] = {"name": mats[k]["name"], "type": "brdf_disney", "diffuse_reflectance": list(base),
                   "metallic": metal, "roughness": rough, "opacity": opac}
    js["lights"].append({"type": "point", "radiance": [1.0, 1.0, 1.0], "position": [0.3, 0.6, 0.2]})
    json.dump(js, open(sp, "w"))
    s = A.HostScene(sp)
    m = s.desc.materials  # record 0 is the loader's default material: JSON material k is record k + 1
    m[3].subsurface, m[3].sheen, m[3].sheenTint, m[3].specularTint = 0.8, 1.0, 0.5, 0.7
    m[4].clearcoat, m[4].clearcoatGloss, m[4].anisotropic = 1.0, 0.7, 0.8
    m[6].clearcoat, m[6].clearcoatGloss, m[6].ior = 0.5, 0.2, 1.8
    assert all(m[k + 1].type == 9 for k in range(6))
    renderer.upload(s)
    renderer.build()
    spp = 16
    st, ost, g, o, sq = _radiance_parity(renderer, s, spp)
    frac_bad = _three_sigma(g[0], o[0], sq, spp)
    mg, mo = g[0][..., :3].mean((0, 1)), o[0][..., :3].mean((0, 1))
    log(f"disney: outside 3 sigma {frac_bad:.5f}, means {mg} vs {mo}; rays gpu {st['extension_rays']} "
        f"oracle {ost['extension_rays']}")
    assert np.isfinite(g[0]).all()
    assert frac_bad <= 0.01
    np.testing.assert_allclose(mg, mo, rtol=1e-2)
    assert abs(st["extension_rays"] - ost["extension_rays"]) / ost["extension_rays"] < 2e-3


def test_sun_and_sky_scene_matches_oracle(renderer, tmp_path):
    """The procedural sun & sky in place of the background: miss radiance
    (sky, sun disk and glow, ground below the horizon) and uniform-sphere NEE
    toward it, in the open box with a pbr floor; the record resets to
    in_use 0 on the next upload."""
    d = tmp_path / "sky"
    d.mkdir()
    sp = A.generate_scene(0, str(d))
    s = A.HostScene(sp)
    renderer.upload(s)
    renderer.build()
    sky = A.sun_and_sky_defaults()
    sky.in_use = 1
    sky.sun_direction[:] = (0.4, 0.5, -0.6)
    sky.haze = 1.0
    sky.horizon_blur = 0.5
    sky.multiplier = 2e-5
    renderer.set_sun_and_sky(sky)
    cam, pc = s.shot(0)
    w, h = s.film
    renderer.set_film(w, h)
    spp = 16
    st = renderer.render(cam, pc, 0, spp)
    g = renderer.read(0)
    osc = O.OracleScene(s.desc)
    osc.set_sun_and_sky(sky)
    sq = np.zeros((h, w, 4), np.float32)
    o, ost = osc.render(cam, pc, w, h, 0, spp, sq=sq)
    frac_bad = _three_sigma(g, o[0], sq, spp)
    mg, mo = g[..., :3].mean((0, 1)), o[0][..., :3].mean((0, 1))
    log(f"sun & sky: outside 3 sigma {frac_bad:.5f}, means {mg} vs {mo}")
    assert frac_bad <= 0.01
    np.testing.assert_allclose(mg, mo, rtol=1e-2)
    assert abs(st["shadow_rays"] - ost["shadow_rays"]) / ost["shadow_rays"] < 2e-3
    # a new upload starts from the defaults again (in_use 0): the background colour is back
    renderer.upload(s)
    renderer.build()
    renderer.set_film(w, h)
    renderer.render(cam, pc, 0, 2)
    base = renderer.read(0)
    osc2 = O.OracleScene(s.desc)
    o2, _ = osc2.render(cam, pc, w, h, 0, 2)
    np.testing.assert_allclose(base[..., :3].mean((0, 1)), o2[0][..., :3].mean((0, 1)), rtol=2e-2)


def test_env_tables_on_device_equal_oracle(renderer, spheres_small):
    s, _ = _load(renderer, spheres_small)
    env = np.ctypeslib.as_array(s.desc.envmap.rgba, (s.desc.envmap.height, s.desc.envmap.width, 4)).copy()
    m, c = renderer.env_tables(s.desc.envmap.width, s.desc.envmap.height)
    import ctypes as C
    H, W = env.shape[:2]
    om = (C.c_float * (H * 2))()
    oc = (C.c_float * (W * H * 2))()
    O.lib().orc_env_tables(W, H, env.ctypes.data_as(C.POINTER(C.c_float)), om, oc)
    np.testing.assert_array_equal(m, np.array(om, np.float32).reshape(H, 2))
    np.testing.assert_array_equal(c, np.array(oc, np.float32).reshape(H, W, 2))


def test_progressive_resume_is_bit_identical(renderer, box_scene):
    """Frames split over calls (first_frame > 0) accumulate exactly as one
    call: the RNG depends only on (x, y, frame)."""
    s, _ = _load(renderer, box_scene)
    cam, pc = s.shot(0)
    w, h = s.film
    renderer.set_film(w, h)
    renderer.render(cam, pc, 0, 8)
    a = renderer.read(0)
    renderer.set_film(w, h)
    renderer.render(cam, pc, 0, 3)
    renderer.render(cam, pc, 3, 5)
    b = renderer.read(0)
    np.testing.assert_array_equal(a, b)
    renderer.set_film(w, h)
    renderer.render(cam, pc, 0, 8, queue_paths=1 << 18)  # one frame per batch
    np.testing.assert_array_equal(a, renderer.read(0))


def test_render_is_deterministic(renderer, spheres_small):
    s, _ = _load(renderer, spheres_small)
    cam, pc = s.shot(0)
    w, h = s.film
    out = []
    for _ in range(2):
        renderer.set_film(w, h)
        renderer.render(cam, pc, 0, 4)
        out.append(renderer.read(0))
    np.testing.assert_array_equal(out[0], out[1])


def test_multishot_batch_equals_single_shots(renderer, multishot_scene):
    """64 shots in one batched wavefront == each shot rendered alone."""
    s, _ = _load(renderer, multishot_scene)
    n = s.state.num_shots
    assert n == 64
    cams, pcs = zip(*[s.shot(i) for i in range(n)])
    w, h = s.film
    renderer.set_film(w, h)
    st = renderer.render_shots(list(cams), list(pcs), 2)
    assert st["paths"] == n * w * h * 2
    for i in (0, 17, 63):
        alone = A.Renderer(0)
        alone.upload(s)
        alone.build()
        alone.set_film(w, h)
        alone.render(cams[i], pcs[i], 0, 2)
        np.testing.assert_array_equal(renderer.read_shot(i, 0), alone.read(0))
        np.testing.assert_array_equal(renderer.read_shot(i, 8), alone.read(8))
        alone.close()


def test_offline_cli_outputs(tmp_path, box_scene):
    """run_offline writes the reference's file names (tracer.cpp:266-291)."""
    out = str(tmp_path / "render")
    st = A.run_offline(box_scene, out, spp_override=4)
    assert st["paths"] == 512 * 512 * 4
    for name in ["render_shot_0000.exr", "render_shot_0000_channel_0000.exr", "render_shot_0000_channel_0001.exr",
                 "render_shot_0000_channel_0002.exr"]:
        assert os.path.exists(tmp_path / name), name
    img = A.read_image(str(tmp_path / "render_shot_0000.exr"))
    assert img.shape == (512, 512, 4) and np.isfinite(img).all() and img[..., :3].mean() > 0


def test_error_paths_do_not_exit(renderer):
    with pytest.raises(A.AsunaError):
        r = A.Renderer(0)
        cam = A.camera_lookat(8, 8, 45, (0, 0, 1), (0, 0, 0))
        r.render(cam, A.GpuPushConstantRaytrace(), 0, 1)  # render before upload
