"""CPU tests of libasuna.so's host side: exported C ABI, POD layouts, the
scene loader (src/loader + src/scene semantics), OBJ/EXR/HDR/PNG/NPY IO and
the procedural scene generator.  No GPU is touched."""
import ctypes as C
import filecmp
import json
import math
import os
import re

import numpy as np
import pytest

import asuna_amd as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    L = A.lib()
    names = set()
    for hdr in sorted(os.listdir(os.path.join(ROOT, "include"))):
        text = open(os.path.join(ROOT, "include", hdr)).read()
        names |= set(re.findall(r"^\s*(?:[\w\*]+\s+)+\**(ash_\w+)\s*\(", text, re.M))
    assert len(names) >= 50 and "ash_set_sun_and_sky" in names and "ash_server_submit" in names
    for n in sorted(names):
        assert hasattr(L, n), n


def test_pod_layouts_match_the_reference_structs():
    # sizeof / offset of the last field: GpuVertex 44 (vertex.h), GpuMaterial 132
    # (material.h), GpuLight 76 (light.h), GpuCamera 224 (camera.h), push
    # constants 84 (pushconstant.h)
    assert A.pod_layout() == [44, 32, 132, 128, 76, 72, 224, 220, 84, 80]
    assert C.sizeof(A.GpuMaterial) == 132 and C.sizeof(A.GpuLight) == 76 and C.sizeof(A.GpuCamera) == 224
    assert C.sizeof(A.GpuPushConstantRaytrace) == 84 and C.sizeof(A.GpuVertex) == 44
    assert C.sizeof(A.GpuSunAndSky) == 96 and A.GpuSunAndSky.in_use.offset == 92  # sun_and_sky.h


def test_server_opts_layout():
    # ash_server_opts (asuna_server.h): u64 budget, six u32 fields
    assert C.sizeof(A.ServerOpts) == 32 and A.ServerOpts.policy.offset == 24


def test_version_string():
    assert b"gfx950" in A.lib().ash_version()


def test_loader_box_scene(box_scene):
    s = A.HostScene(box_scene)
    d = s.desc
    assert s.film == (512, 512) and s.state.num_shots == 1
    # Scene::reset adds dummy material 0 and dummy light 0 (scene.cpp:93-111)
    assert d.num_materials == 8 and d.num_lights == 2
    l0, l1 = d.lights[0], d.lights[1]
    assert l0.type == 0 and l0.doubleSide == 1 and list(l0.radiance) == [0, 0, 0]
    assert l1.type == 1 and list(l1.radiance) == [17, 12, 4] and l1.doubleSide == 0
    assert abs(l1.area - 0.0625) < 1e-7
    np.testing.assert_allclose(np.cross(list(l1.u), list(l1.v)), [0, -0.0625, 0], atol=1e-7)
    # rect light -> 4-vertex mesh + emitter instance, created before the
    # scene's own meshes and instances (scene.cpp:220-244)
    assert d.num_meshes == 7 and d.num_instances == 8
    assert d.instances[0].light_id == 1 and d.meshes[d.instances[0].mesh_id].num_vertices == 4
    assert [d.instances[i].light_id for i in range(1, 8)] == [-1] * 7
    assert [d.instances[i].material_id for i in range(1, 8)] == list(range(1, 8))
    cam, pc = s.shot(0)
    assert pc.spp == 64 and pc.maxPathDepth == 8 and pc.numLights == 1 and pc.nMultiChannel == 3
    assert (pc.diffuseOutChannel, pc.normalOutChannel, pc.positionOutChannel) == (0, 1, 2)
    assert pc.uvOutChannel == -1 and pc.hasEnvMap == 0
    assert s.state.output_hdr == 1 and s.state.tone_mapping == 4


def test_loader_toworld_stack_left_multiplies(box_scene):
    js = json.load(open(box_scene))
    inst = [i for i in js["instances"] if i["mesh"] == "box"][0]
    ops = inst["toworld"]
    S = np.diag(ops[0]["value"] + [1.0])
    a = math.radians(ops[1]["value"])
    Ry = np.array([[math.cos(a), 0, math.sin(a), 0], [0, 1, 0, 0], [-math.sin(a), 0, math.cos(a), 0], [0, 0, 0, 1]])
    T = np.eye(4)
    T[:3, 3] = ops[2]["value"]
    want = T @ Ry @ S
    s = A.HostScene(box_scene)
    idx = [i for i in range(s.desc.num_instances) if s.desc.instances[i].material_id == 6][0]
    got = np.array(s.desc.instances[idx].transform).reshape(4, 4).T  # column-major storage
    np.testing.assert_allclose(got, want, atol=1e-6)


def test_loader_mesh_lights_and_env(spheres_small):
    s = A.HostScene(spheres_small)
    d = s.desc
    tri = [d.lights[i] for i in range(1, d.num_lights)]
    assert len(tri) == 4 and all(l.type == 2 for l in tri)  # "mesh" -> triangle lights
    for l in tri:
        assert abs(l.area - 0.5 * np.linalg.norm(np.cross(list(l.u), list(l.v)))) < 1e-6
        assert np.cross(list(l.u), list(l.v))[1] < 0  # emits downwards
        assert l.doubleSide == 0
    assert d.envmap.width == 256 and d.envmap.height == 128
    cam, pc = s.shot(0)
    assert pc.hasEnvMap == 1 and list(pc.envMapResolution) == [256.0, 128.0] and pc.numLights == 4
    assert s.state.depth_out_channel == 2


def _write_scene(tmp_path, body, obj="v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n"):
    (tmp_path / "m.obj").write_text(obj)
    p = tmp_path / "s.json"
    p.write_text(json.dumps(body))
    return str(p)


def _base():
    return {"state": {"path_tracing": {"spp": 2}}, "camera": {"type": "perspective", "film": {"resolution": [8, 6]}},
            "materials": [{"name": "m", "type": "brdf_lambertian"}], "meshes": [{"name": "a", "path": "m.obj"}],
            "instances": [{"mesh": "a", "material": "m"}],
            "shots": [{"type": "lookat", "eye": [0, 0, 2], "lookat": [0, 0, 0], "up": [0, 1, 0]}]}


def test_loader_triangle_light_is_a_half_area_rect(tmp_path):
    b = _base()
    b["lights"] = [{"type": "triangle", "radiance": [1, 1, 1], "position": [0, 1, 0], "v1": [1, 1, 0],
                    "v2": [0, 1, 1]}]
    s = A.HostScene(_write_scene(tmp_path, b))
    l = s.desc.lights[1]
    assert l.type == 1 and abs(l.area - 0.5) < 1e-7  # loader.cpp:284-296


def test_loader_shot_state_override_and_material_defaults(tmp_path):
    b = _base()
    b["shots"][0]["state"] = {"path_tracing": {"spp": 7, "max_path_depth": 5, "background_color": [1, 2, 3]}}
    b["materials"].append({"name": "g", "type": "bsdf_dielectric"})
    b["materials"].append({"name": "rp", "type": "brdf_rough_plastic", "ior": 1.5})
    s = A.HostScene(_write_scene(tmp_path, b))
    cam, pc = s.shot(0)
    assert pc.spp == 7 and pc.maxPathDepth == 5 and list(pc.bgColor) == [1, 2, 3]
    m = s.desc.materials
    assert m[2].type == 10 and abs(m[2].ior - 1.5) < 1e-7  # default ior (material.h:9-35)
    assert m[1].type == 0 and abs(m[1].roughness - 0.5) < 1e-7 and list(m[1].radianceFactor) == [1, 1, 1]
    assert abs(m[3].radiance[0] - 0.5963) < 2e-3  # fdrInt, internal diffuse Fresnel of ior 1.5


def test_loader_disney_reads_four_inputs(tmp_path):
    """brdf_disney reads base colour, metallic, roughness and opacity (held
    in rhoSpec.x, default 0); its other lobe inputs keep the material
    defaults (material.h:9-35, loader/material.cpp:359-389)."""
    b = _base()
    b["materials"].append({"name": "d", "type": "brdf_disney", "diffuse_reflectance": [0.1, 0.2, 0.3],
                           "metallic": 0.4, "roughness": 0.3, "opacity": 0.25, "sheen": 9.0})
    b["materials"].append({"name": "d2", "type": "brdf_disney"})
    s = A.HostScene(_write_scene(tmp_path, b))
    m = s.desc.materials
    d, d2 = m[2], m[3]
    assert d.type == 9 and d2.type == 9
    np.testing.assert_allclose(list(d.diffuse), [0.1, 0.2, 0.3], rtol=1e-7)
    assert abs(d.metalness - 0.4) < 1e-7 and abs(d.roughness - 0.3) < 1e-7 and abs(d.rhoSpec[0] - 0.25) < 1e-7
    assert d.sheen == 0 and d.clearcoat == 0 and d.subsurface == 0 and d.anisotropic == 0 and abs(d.ior - 1.5) < 1e-7
    assert d2.rhoSpec[0] == 0 and abs(d2.roughness - 0.5) < 1e-7 and d2.metalness == 0


@pytest.mark.parametrize("mutate,code", [
    (lambda b: b.pop("instances"), "malformed"),
    (lambda b: b["instances"][0].pop("material"), "malformed"),
    (lambda b: b["instances"][0].update(material="nope"), "malformed"),
    (lambda b: b["meshes"][0].update(path="missing.obj"), "i/o"),
    (lambda b: b["materials"][0].update(type="brdf_unknown"), "malformed"),
    (lambda b: b["camera"].update(type="fisheye"), "malformed"),
])
def test_loader_errors_are_codes_not_exits(tmp_path, mutate, code):
    b = _base()
    mutate(b)
    with pytest.raises(A.AsunaError, match=code):
        A.HostScene(_write_scene(tmp_path, b))


def test_loader_rejects_broken_json(tmp_path):
    p = tmp_path / "bad.json"
    p.write_text('{"state": {,}')
    with pytest.raises(A.AsunaError, match="malformed"):
        A.HostScene(str(p))


def test_obj_semantics_deindex_uvflip_fan(tmp_path):
    """loadMesh (mesh.cpp:112-145): one vertex per face corner, index =
    running count, uv.y flipped, missing normals zero, polygons fanned."""
    obj = ("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvt 0.25 0.1\nvt 0.75 0.9\nvn 0 0 1\n"
           "f 1/1/1 2/2/1 3/1/1 4/2/1\nf -4 -3 -2\n")
    b = _base()
    s = A.HostScene(_write_scene(tmp_path, b, obj))
    (v, idx), = s.meshes()
    assert len(idx) == 9 and list(idx) == list(range(9))  # quad -> 2 tris + 1 tri
    np.testing.assert_allclose(v[:3, 0:3], [[0, 0, 0], [1, 0, 0], [1, 1, 0]])
    np.testing.assert_allclose(v[3:6, 0:3], [[0, 0, 0], [1, 1, 0], [0, 1, 0]])
    np.testing.assert_allclose(v[0, 3:5], [0.25, 0.9], atol=1e-7)  # (u, 1 - v)
    np.testing.assert_allclose(v[1, 3:5], [0.75, 0.1], atol=1e-7)
    np.testing.assert_allclose(v[0, 5:8], [0, 0, 1])
    np.testing.assert_allclose(v[6:9, 5:8], 0)  # no normal index -> zero normal


def test_image_io_round_trips(tmp_path):
    rng = np.random.default_rng(0)
    img = rng.random((13, 17, 4)).astype(np.float32) * 4
    img[..., 3] = 1
    # EXR (half): values that are exact halves survive bit-exactly
    half = img.astype(np.float16).astype(np.float32)
    A.write_image(str(tmp_path / "a.exr"), half)
    back = A.read_image(str(tmp_path / "a.exr"))
    np.testing.assert_array_equal(back[..., :3], half[..., :3])
    # HDR (RGBE): 8-bit shared-exponent mantissas
    A.write_image(str(tmp_path / "a.hdr"), img)
    back = A.read_image(str(tmp_path / "a.hdr"))
    # shared exponent: every channel is quantised to 1/256 of the pixel's largest one
    tol = img[..., :3].max(-1, keepdims=True) / 128
    assert (np.abs(back[..., :3] - img[..., :3]) <= tol).all()
    assert (back[..., :3] <= img[..., :3] + 1e-7).all()  # mantissas truncate
    # PNG: v -> int(clamp(255 v + 0.5)), read back as v/255
    ldr = rng.random((5, 7, 4)).astype(np.float32)
    A.write_image(str(tmp_path / "a.png"), ldr)
    back = A.read_image(str(tmp_path / "a.png"))
    np.testing.assert_allclose(back, np.floor(np.clip(ldr * 255 + 0.5, 0, 255)) / 255, atol=1e-7)
    # NPY
    A.write_image(str(tmp_path / "a.npy"), img)
    np.testing.assert_array_equal(np.load(str(tmp_path / "a.npy")), img)


def test_tonemap_filmic_closed_form():
    x = np.array([[0.0, 0.5, 2.0, 1.0]], np.float32)
    y = x.copy()
    A.lib().ash_tonemap(4, 1, 1, A.fptr(y))
    v = np.maximum(x[0, :3] - 0.004, 0)
    np.testing.assert_allclose(y[0, :3], (v * (6.2 * v + 0.5)) / (v * (6.2 * v + 1.7) + 0.06), rtol=1e-6)


def _tonemap(kind, x):
    y = np.zeros((1, len(x), 4), np.float32)
    y[0, :, :3] = np.asarray(x, np.float32)[:, None]
    A.lib().ash_tonemap(kind, len(x), 1, A.fptr(y))
    return y[0, :, 0].astype(np.float64)


def test_tonemap_gamma_reinhard_aces_and_pbrt_closed_forms():
    """post.idle.frag:75-96 with tonemapping.glsl:32-81:
    gamma    (x / (1 + x/1.5))^(1/2.2): 0 at 0, tends to 1.5^(1/2.2) for large x;
    reinhard Hejl & Burgess-Dawson's filmic fit, gamma built in: 0 up to the
             0.004 toe, (v(6.2v + .5)) / (v(6.2v + 1.7) + .06), tends to 1;
    aces     Narkowicz's fit (x(2.51x + .03)) / (x(2.43x + .59) + .14),
             clamped, then the 1/2.2 gamma: 0 at 0, 1 once the fit passes 1;
    pbrt     the sRGB transfer function: 12.92 x below 0.0031308, else
             1.055 x^(1/2.4) - 0.055 (continuous there; 0.5 -> 0.735357,
             1 -> 1)."""
    g = _tonemap(1, [0.0, 1.5, 1e6])
    np.testing.assert_allclose(g, [0.0, 0.75 ** (1 / 2.2), 1.5 ** (1 / 2.2)], rtol=1e-5)
    r = _tonemap(2, [0.0, 0.004, 0.5, 1e4])
    v = 0.496
    assert r[0] == 0.0 and r[1] == 0.0 and abs(r[3] - 1.0) < 1e-4
    np.testing.assert_allclose(r[2], v * (6.2 * v + 0.5) / (v * (6.2 * v + 1.7) + 0.06), rtol=1e-6)
    a = _tonemap(3, [0.0, 1.0, 100.0])
    fit1 = 2.54 / 3.16
    np.testing.assert_allclose(a, [0.0, fit1 ** (1 / 2.2), 1.0], rtol=1e-6, atol=1e-7)
    p = _tonemap(5, [0.001, 0.0031308, 0.5, 1.0])
    np.testing.assert_allclose(p, [0.01292, 1.055 * 0.0031308 ** (1 / 2.4) - 0.055, 0.735357, 1.0], rtol=2e-6)
    assert abs(12.92 * 0.0031308 - p[1]) < 1e-6  # the two pieces meet
    n = _tonemap(0, [0.25, 7.0])
    np.testing.assert_array_equal(n, [0.25, 7.0])


def _pcg3d(v):
    v = [(x * 1664525 + 1013904223) & 0xFFFFFFFF for x in v]
    m = 0xFFFFFFFF
    v[0] = (v[0] + v[1] * v[2]) & m
    v[1] = (v[1] + v[2] * v[0]) & m
    v[2] = (v[2] + v[0] * v[1]) & m
    v = [x ^ (x >> 16) for x in v]
    v[0] = (v[0] + v[1] * v[2]) & m
    v[1] = (v[1] + v[2] * v[0]) & m
    v[2] = (v[2] + v[0] * v[1]) & m
    return v


def test_tonemap_custom_is_dithered_uncharted2():
    """Custom operator: Uncharted 2 (Hable 2010, exposure bias 2, white 11.2)
    then gamma 2.2, dithered to one of the two neighbouring 1/255 levels,
    chosen by pcg3d noise of (x, y, 0) (Jarzynski & Olano 2020)."""
    w, h = 7, 5
    rng = np.random.default_rng(3)
    img = np.concatenate([rng.uniform(0, 4, (h, w, 3)), np.ones((h, w, 1))], -1).astype(np.float32)
    out = img.copy()
    A.lib().ash_tonemap(6, w, h, A.fptr(out))

    def u2(x):
        a, b, c, d, e, f = 0.15, 0.50, 0.10, 0.20, 0.02, 0.30
        return (x * (a * x + c * b) + d * e) / (x * (a * x + b) + d * f) - e / f

    t = (u2(2.0 * img[..., :3].astype(np.float64)) / u2(11.2)) ** (1 / 2.2)
    lo = np.floor(t * 255 + 1e-9) / 255
    got = out[..., :3].astype(np.float64)
    # every value is one of the two 1/255 levels around the tone-mapped value
    assert np.all((np.abs(got - lo) < 2e-6) | (np.abs(got - (lo + 1 / 255)) < 2e-6) | (np.abs(got - t) < 2e-6))
    # and the upper level is taken exactly when the noise falls below the
    # linear-space position of the value between the two levels
    for y in range(h):
        for x in range(w):
            r = _pcg3d([x, y, 0])
            for k in range(3):
                noise = np.frombuffer(np.uint32(0x3F800000 | (r[k] >> 9)).tobytes(), np.float32)[0] - 1.0
                l0, l1 = lo[y, x, k], lo[y, x, k] + 1 / 255
                lin = t[y, x, k] ** 2.2
                frac = (lin - l0 ** 2.2) / (l1 ** 2.2 - l0 ** 2.2)
                if abs(frac - noise) > 1e-4:  # away from the decision boundary
                    assert abs(got[y, x, k] - (l1 if noise < frac else l0)) < 2e-6
    np.testing.assert_array_equal(out[..., 3], 1.0)


def test_scene_generator_is_deterministic(tmp_path):
    a = A.generate_scene(1, str(tmp_path / "a"), scale=1)
    b = A.generate_scene(1, str(tmp_path / "b"), scale=1)
    for f in os.listdir(os.path.dirname(a)):
        assert filecmp.cmp(os.path.join(os.path.dirname(a), f), os.path.join(os.path.dirname(b), f), shallow=False), f


def test_training_scenes_follow_the_denoiser_data_spec(tmp_path):
    """Generator config 3 (the denoiser's training set): per seed 1-20
    displaced meshes with materials from the plugin set, a ground, 1-4 rect
    lights facing down and an env map; the same seed writes the same files,
    different seeds different scenes."""
    import json
    kinds, counts = set(), []
    for k in range(12):
        p = A.generate_scene(3, str(tmp_path / f"s{k}"), seed=1000 + k)
        js = json.load(open(p))
        n = len(js["instances"]) - 1  # the ground is the last instance
        assert 1 <= n <= 20 and js["instances"][-1]["mesh"] == "ground"
        assert 1 <= len(js["lights"]) <= 4
        for l in js["lights"]:
            assert l["type"] == "rect"
            u = np.subtract(l["v1"], l["position"])
            v = np.subtract(l["v2"], l["position"])
            assert np.cross(u, v)[1] < 0  # emits downwards
        assert js["envmap"]["path"] == "sky.exr"
        kinds |= {m["type"] for m in js["materials"][:-1]}
        counts.append(n)
        s = A.HostScene(p)
        assert s.film == (256, 256) and s.state.num_triangles > 0
    assert kinds <= {"brdf_lambertian", "brdf_pbr_metalness_roughness", "brdf_rough_plastic", "bsdf_dielectric"}
    assert len(kinds) >= 3 and len(set(counts)) > 3
    a = A.generate_scene(3, str(tmp_path / "again"), seed=1000)
    for f in os.listdir(os.path.dirname(a)):
        assert filecmp.cmp(os.path.join(os.path.dirname(a), f), os.path.join(str(tmp_path / "s0"), f), shallow=False), f


def test_multishot_scene_has_64_shots_with_own_fov(multishot_scene):
    s = A.HostScene(multishot_scene)
    assert s.state.num_shots == 64
    cams = [s.shot(i)[0] for i in range(64)]
    fx = {round(c.rasterToCamera[0], 6) for c in cams}
    assert len(fx) > 32  # per-shot fov extension reaches the camera matrices
    eyes = np.array([[c.cameraToWorld[12], c.cameraToWorld[13], c.cameraToWorld[14]] for c in cams])
    r = np.linalg.norm(eyes, axis=1)
    assert r.min() >= 5 - 1e-4 and r.max() <= 9 + 1e-4


def test_jpeg_textures_decode_like_libjpeg(tmp_path):
    """.jpg textures (texture.cpp:308-327): baseline and progressive files,
    4:4:4 / 4:2:2 / 4:2:0 chroma, grey, restart markers and sizes that are
    not whole MCUs decode to within 3/255 of libjpeg's islow decoder (via
    Pillow) and 0.25/255 on average; channels are (v/255)^gamma, alpha 1."""
    Image = pytest.importorskip("PIL.Image")
    rng = np.random.default_rng(8)
    cases = [dict(quality=90, subsampling=2), dict(quality=95, subsampling=0), dict(quality=80, subsampling=1),
             dict(quality=90, subsampling=2, progressive=True), dict(quality=97, subsampling=0, progressive=True),
             dict(quality=75, subsampling=2, restart_marker_blocks=3)]
    for (w, h) in [(123, 77), (17, 9)]:
        y, x = np.mgrid[0:h, 0:w]
        base = np.stack([128 + 100 * np.sin(x / 7.0 + y / 11.0), 128 + 90 * np.cos(x / 5.0 - y / 9.0),
                         (x * 3 + y * 2) % 256], -1) + rng.normal(0, 12, (h, w, 3))
        rgb = np.clip(base, 0, 255).astype(np.uint8)
        for mode, arr in (("RGB", rgb), ("L", rgb[..., 0])):
            for kw in cases:
                p = str(tmp_path / "t.jpg")
                Image.fromarray(arr, mode).save(p, "JPEG", **kw)
                ref = np.asarray(Image.open(p).convert("RGB")).astype(np.float64)
                got = A.read_image(p)
                assert got.shape == (h, w, 4) and np.all(got[..., 3] == 1.0)
                d = np.abs(got[..., :3] * 255.0 - ref)
                assert d.max() <= 3.0 and d.mean() <= 0.25, (w, h, mode, kw, d.max(), d.mean())
    # gamma: the same 8-bit values raised to the texture's gamma
    g = A.read_image(p, gamma=2.2)
    lin = A.read_image(p)
    np.testing.assert_allclose(g[..., :3], np.round(lin[..., :3] * 255) ** 2.2 / 255 ** 2.2, rtol=1e-5, atol=1e-7)
    # a flat colour survives the round trip within one level
    Image.fromarray(np.full((24, 40, 3), (200, 40, 90), np.uint8), "RGB").save(p, "JPEG", quality=100, subsampling=0)
    np.testing.assert_allclose(A.read_image(p)[..., :3] * 255.0, np.broadcast_to([200, 40, 90], (24, 40, 3)), atol=1.0)


def test_loader_reads_a_jpg_texture(tmp_path):
    Image = pytest.importorskip("PIL.Image")
    Image.fromarray(np.full((8, 8, 3), (255, 128, 0), np.uint8), "RGB").save(str(tmp_path / "t.jpg"), "JPEG",
                                                                              quality=100, subsampling=0)
    (tmp_path / "tri.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 0 1\nvn 0 0 1\n"
                                      "f 1/1/1 2/2/1 3/3/1\n")
    js = {"state": {"path_tracing": {"spp": 1}}, "camera": {"type": "perspective", "film": {"resolution": [8, 8]}},
          "textures": [{"name": "t", "path": "t.jpg"}],
          "materials": [{"name": "m", "type": "brdf_lambertian", "diffuse_texture": "t"}],
          "meshes": [{"name": "tri", "path": "tri.obj"}], "instances": [{"mesh": "tri", "material": "m"}],
          "shots": [{"type": "lookat", "eye": [0, 0, 2], "lookat": [0, 0, 0], "up": [0, 1, 0]}]}
    import json
    (tmp_path / "s.json").write_text(json.dumps(js))
    s = A.HostScene(str(tmp_path / "s.json"))
    tid = [s.desc.materials[i].diffuseTextureId for i in range(s.desc.num_materials)
           if s.desc.materials[i].diffuseTextureId >= 0]
    assert len(tid) == 1
    t = s.desc.textures[tid[0]]
    assert (t.width, t.height) == (8, 8)
    px = np.ctypeslib.as_array(t.rgba, (t.height, t.width, 4))
    np.testing.assert_allclose(px[0, 0, :3], np.array([255, 128, 0]) / 255.0, atol=1.5 / 255)
