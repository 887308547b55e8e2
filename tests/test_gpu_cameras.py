"""Camera models, per-shot flags, the accumulate kernel's ragged waves and
the `--output_scanline` writer on the GPU, each against the CPU oracle or a
closed form, through the C ABI.

Bars:
  first hit   primitive / instance ids identical on >= 99.99 % of pixels; t and
              normal within 1e-5 relative (frame 0: the thin lens draws its two
              lens numbers from the pixel's seed, raytrace.projective.rgen:56-69)
  radiance    per-pixel |gpu - oracle| <= 3 sigma of the Monte Carlo standard
              error on >= 99 % of pixels, image means within 0.5 %
  bit-exact   batching and file contents (integer-exact float comparison)
"""
import json
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


def _edit_box(tmp_path, name, edit):
    d = tmp_path / name
    d.mkdir()
    sp = A.generate_scene(0, str(d))
    js = json.load(open(sp))
    edit(js)
    json.dump(js, open(sp, "w"))
    return A.HostScene(sp)


def _check_first_hit(r, s, osc, cam):
    w, h = s.film
    g = r.first_hit(cam)
    o = osc.first_hit(cam, w, h)
    same = (g["prim_id"] == o["prim_id"]) & (g["inst_id"] == o["inst_id"])
    hit = same & (o["inst_id"] >= 0)
    rel_t = np.abs(g["t"][hit] - o["t"][hit]) / np.abs(o["t"][hit])
    dn = np.abs(g["normal"][hit] - o["normal"][hit]).max(initial=0.0)
    log(f"first hit: id match {same.mean():.6f}, hits {hit.mean():.3f}, max rel t {rel_t.max(initial=0):.3g}, "
        f"max |dN| {dn:.3g}")
    assert same.mean() >= 0.9999
    assert hit.mean() > 0.5
    assert rel_t.max(initial=0) <= 1e-5 and dn <= 1e-5
    return g, o


def _check_radiance(r, s, osc, cam, pc, spp):
    w, h = s.film
    r.set_film(w, h)
    r.render(cam, pc, 0, spp)
    g = np.stack([r.read(i) for i in range(9)])
    sq = np.zeros((h, w, 4), np.float32)
    o, _ = osc.render(cam, pc, w, h, 0, spp, sq=sq)
    mean = o[0][..., :3]
    var = np.maximum(sq[..., :3] / spp - mean ** 2, 0.0)
    se = np.sqrt(2 * var / spp) + 1e-6
    bad = (np.abs(g[0][..., :3] - mean) > 3 * se + 1e-5).any(-1).mean()
    mg, mo = g[0][..., :3].mean((0, 1)), mean.mean((0, 1))
    log(f"radiance: outside 3 sigma {bad:.5f}, means {mg} vs {mo}")
    assert np.isfinite(g[0]).all()
    assert bad <= 0.01
    np.testing.assert_allclose(mg, mo, rtol=5e-3)
    for c in range(1, 1 + pc.nMultiChannel):  # frame-0 AOVs
        np.testing.assert_allclose(g[c], o[c], rtol=1e-5, atol=1e-5)
    return g, o


def test_thin_lens_camera_matches_oracle(renderer, tmp_path):
    """Perspective camera with aperture 0.12 focused at 3.4 (the box's
    centre plane): first hits and 16-spp radiance vs the oracle."""
    def edit(js):
        js["camera"]["aperture"] = 0.12
        js["camera"]["focal_distance"] = 3.4
        js["shots"][0]["eye"] = [0.35, 0.2, 3.4]
    s = _edit_box(tmp_path, "lens", edit)
    cam, pc = s.shot(0)
    assert cam.aperture == pytest.approx(0.12) and cam.focalDistance == pytest.approx(3.4)
    renderer.upload(s)
    renderer.build()
    w, h = s.film
    renderer.set_film(w, h)
    osc = O.OracleScene(s.desc)
    g, o = _check_first_hit(renderer, s, osc, cam)
    # the lens moved the frame-0 rays: first hits differ from the pinhole's
    pin = A.camera_lookat(w, h, 40.0, (0.35, 0.2, 3.4), (0.0, 0.0, 0.0))
    gp = renderer.first_hit(pin)
    assert np.mean(gp["t"] != g["t"]) > 0.5
    _check_radiance(renderer, s, osc, cam, pc, 16)


def _opencv_extrinsic(eye, look):
    """World -> camera [R | t] of a camera at eye looking at look (x right,
    y down, z forward)."""
    eye, look = np.asarray(eye, float), np.asarray(look, float)
    z = look - eye
    z /= np.linalg.norm(z)
    x = np.cross(z, [0.0, -1.0, 0.0])
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    R = np.stack([x, y, z])
    ext = np.eye(4)
    ext[:3, :3], ext[:3, 3] = R, -R @ eye
    return ext


def test_opencv_camera_matches_oracle(renderer, tmp_path):
    """`opencv` camera (fx, fy, cx, cy; rgen:75-87) with an `opencv` shot
    (extrinsic matrix, loader.cpp:430-437): first hits and radiance."""
    ext = _opencv_extrinsic((0.4, -0.25, 3.2), (0.0, 0.05, 0.0))
    def edit(js):
        js["camera"] = {"type": "opencv", "film": {"resolution": [384, 288]},
                        "fx": 420.0, "fy": 436.0, "cx": 199.5, "cy": 140.25}
        js["shots"] = [{"type": "opencv", "matrix": [float(v) for v in ext.ravel()]}]
    s = _edit_box(tmp_path, "opencv", edit)
    cam, pc = s.shot(0)
    assert cam.type == 1 and s.film == (384, 288)
    renderer.upload(s)
    renderer.build()
    renderer.set_film(*s.film)
    osc = O.OracleScene(s.desc)
    _check_first_hit(renderer, s, osc, cam)
    _check_radiance(renderer, s, osc, cam, pc, 16)


def _emissive_instance(scene):
    """Index of the one instance whose material is brdf_emissive (material
    type 2 of the reference's MaterialType enum)."""
    d = scene.desc
    idx = [i for i in range(d.num_instances) if d.materials[d.instances[i].material_id].type == 2]
    assert len(idx) == 1
    return idx[0]


def test_face_normal_and_ignore_emissive_flags(renderer, tmp_path):
    """use_face_normal reads the never-written geometric normal
    (rchit_layouts.glsl:62), so the normal AOV is exactly zero; with
    ignore_emissive an emissive surface adds nothing (brdf_emissive.rchit:
    12-26) and a camera ray that meets it ends with radiance exactly 0;
    without the flag such a pixel holds the emission itself on every frame.
    Radiance with both flags against the oracle."""
    def edit(js, flags):
        js["state"]["path_tracing"]["use_face_normal"] = flags
        js["state"]["path_tracing"]["ignore_emissive"] = flags
        js["materials"][6] = {"name": js["materials"][6]["name"], "type": "brdf_emissive",
                              "radiance": [2.0, 1.0, 0.5]}
    s = _edit_box(tmp_path, "flags", lambda js: edit(js, True))
    s0 = _edit_box(tmp_path, "noflags", lambda js: edit(js, False))
    cam, pc = s.shot(0)
    assert pc.useFaceNormal == 1 and pc.ignoreEmissive == 1
    renderer.upload(s)
    renderer.build()
    renderer.set_film(*s.film)
    osc = O.OracleScene(s.desc)
    hits = renderer.first_hit(cam)
    g, o = _check_radiance(renderer, s, osc, cam, pc, 16)
    hit = hits["inst_id"] >= 0
    assert np.all(g[pc.normalOutChannel + 1][hit][:, :3] == 0.0)
    # pixels whose 3x3 neighbourhood of frame-0 rays all meet the emissive
    # box first: every jittered sample of such a pixel meets it too
    inst = _emissive_instance(s)
    e = np.pad(hits["inst_id"] == inst, 1)
    emissive = np.ones_like(hits["inst_id"], bool)
    for dy in (0, 1, 2):
        for dx in (0, 1, 2):
            emissive &= e[dy:dy + e.shape[0] - 2, dx:dx + e.shape[1] - 2]
    assert emissive.mean() > 0.005
    assert np.all(g[0][emissive][:, :3] == 0.0)
    cam0, pc0 = s0.shot(0)
    assert pc0.ignoreEmissive == 0
    renderer.upload(s0)
    renderer.build()
    renderer.set_film(*s0.film)
    renderer.render(cam0, pc0, 0, 8)
    img = renderer.read(0)
    np.testing.assert_array_equal(img[emissive][:, :3], np.broadcast_to(np.float32([2.0, 1.0, 0.5]),
                                                                       (int(emissive.sum()), 3)))


def test_accumulate_ragged_waves_and_long_batches(renderer, tmp_path):
    """kAccumulate on a 100x37 film (3700 pixels: the last wave of each shot
    is partial and, with three shots in one batch, a wave spans two work
    units) with 40 frames per batch (several LDS chunks per pixel): the
    radiance and weight images of every shot equal those of a render with
    one frame per batch, bit for bit."""
    sc = A.HostScene(A.generate_scene(0, str(tmp_path / "acc")))
    _, pc = sc.shot(0)
    w, h = 100, 37
    cams = [A.camera_lookat(w, h, fov, eye, (0.0, 0.0, 0.0))
            for fov, eye in ((40.0, (0.0, 0.0, 3.4)), (55.0, (0.6, 0.3, 3.0)), (35.0, (-0.5, -0.2, 3.6)))]
    renderer.upload(sc)
    renderer.build()
    renderer.set_film(w, h)
    out = {}
    for q in (0, len(cams) * w * h):
        st = renderer.render_shots(cams, [pc] * len(cams), 40, queue_paths=q)
        out[q] = [[renderer.read_shot(k, i) for i in (0, 8)] for k in range(len(cams))]
        log(f"queue_paths {q}: {st['batches']} batches")
        assert st["batches"] == (1 if q == 0 else 40)
    for k in range(len(cams)):
        for a, b in zip(out[0][k], out[len(cams) * w * h][k]):
            np.testing.assert_array_equal(a, b)
        assert out[0][k][0][..., :3].mean() > 0


def test_output_scanline_npy_contents(tmp_path):
    """--output_scanline (tracer.cpp:365-391): the first AOV channel is the
    uv channel (u, v, 1 on a hit); its .npy holds (u, v, pixel index) of
    every pixel whose third component is 1, in pixel order, and each later
    channel's .npy holds that channel's xyz at the same pixels.  As in the
    reference, the .npy data goes to the channel's image name
    (`<out>_shot_%04d_channel_%04d.exr`, tracer.cpp:281-287)."""
    d = tmp_path / "scan"
    d.mkdir()
    sp = A.generate_scene(0, str(d))
    js = json.load(open(sp))
    js["state"]["path_tracing"]["multi_channel"] = ["uv", "normal"]
    js["camera"]["film"]["resolution"] = [96, 64]
    js["camera"]["fov"] = 70.0
    js["shots"][0]["eye"] = [0.0, 0.0, 5.5]  # the box covers part of the film: some rays miss
    json.dump(js, open(sp, "w"))
    out = str(d / "render")
    A.run_offline(sp, out, output_scanline=True, spp_override=2)
    s = A.HostScene(sp)
    cam, pc = s.shot(0)
    r = A.Renderer(0)
    try:
        r.upload(s)
        r.build()
        r.set_film(96, 64)
        r.render(cam, pc, 0, 2)
        uv = r.read(1).reshape(-1, 4)
        nrm = r.read(2).reshape(-1, 4)
    finally:
        r.close()
    valid = np.nonzero(uv[:, 2] == 1.0)[0]
    assert 0 < valid.size < uv.shape[0]
    a0 = np.load(os.path.join(d, "render_shot_0000_channel_0000.exr"), allow_pickle=False)
    a1 = np.load(os.path.join(d, "render_shot_0000_channel_0001.exr"), allow_pickle=False)
    assert a0.dtype == np.float32 and a0.shape == (valid.size, 3) and a1.shape == (valid.size, 3)
    np.testing.assert_array_equal(a0[:, :2], uv[valid, :2])
    np.testing.assert_array_equal(a0[:, 2], valid.astype(np.float32))
    np.testing.assert_array_equal(a1, nrm[valid, :3])
    assert os.path.exists(os.path.join(d, "render_shot_0000.exr"))
