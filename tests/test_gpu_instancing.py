"""Instancing without duplication (the two-level build) on the GPU.

Scene: four fBm-displaced icospheres of 20,480 triangles each, instanced
256 times on a 16 x 16 grid with seeded rotations and scales (5.24 M
triangles in the world, 81,920 distinct), on the ground quad under four
triangle lights and an env map.

Bars, with the flat build of the same scene beside it:
  memory      the two-level tree stores each distinct mesh once: its device
              bytes are below 1/20 of the flat tree's
  first hit   against the oracle (world-space triangles) on all 960x540
              pixels: primitive / instance ids identical on >= 99.99 %, t and
              shading normal within 1e-5 relative
  3 sigma     on 8 fixed 32x32 crops at 16 spp: |gpu - oracle FP32| <= 3 sigma
              of the Monte Carlo standard error on >= 99 % of the pixels
  scale       48 x 48 instances (47.2 M triangles in the world): below 1/400
              of the flat tree's bytes, first hits equal to the flat tree's
"""
import json
import os
import time

import numpy as np
import pytest

import asuna_amd as A
from oracle import oracle as O

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)
W, H = 960, 540


def log(*a):
    print(*a, flush=True)


def instanced_scene(d, grid, eye):
    """The four spheres of a generated 2 x 2 scene instanced on a grid x grid
    lattice (seeded mesh choice, scale and rotation) on its ground and
    lights; returns the scene file."""
    p = A.generate_scene(4, d, scale=81920)  # 2 x 2 spheres at 20,480 triangles
    j = json.load(open(p))
    rng = np.random.default_rng(20261203)
    spheres = [m["name"] for m in j["meshes"] if m["name"].startswith("sphere")]
    assert len(spheres) == 4
    inst = [i for i in j["instances"] if not i["mesh"].startswith("sphere")]
    h = (grid - 1) / 2
    for gz in range(grid):
        for gx in range(grid):
            k = int(rng.integers(0, 4))
            s = float(rng.uniform(0.3, 0.42))
            rot = [float(v) for v in rng.uniform(-180, 180, 3)]
            inst.append({"mesh": spheres[k], "material": f"m_{spheres[k]}",
                         "toworld": [{"type": "scale", "value": [s, s, s]},
                                     {"type": "rotate", "value": rot},
                                     {"type": "translate", "value": [gx - h, s * 1.1, gz - h]}]})
    j["instances"] = inst
    j["camera"]["film"]["resolution"] = [W, H]
    j["shots"] = [{"type": "lookat", "eye": list(eye), "lookat": [0.0, 0.3, 0.0], "up": [0, 1, 0]}]
    q = os.path.join(d, "instanced.json")
    with open(q, "w") as f:
        json.dump(j, f)
    return q


@pytest.fixture(scope="module")
def inst_scene(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("instanced"))
    s = A.HostScene(instanced_scene(d, 16, (0.0, 5.5, 11.0)))
    log(f"instanced scene: {s.state.num_triangles} triangles, film {s.film}")
    flat, two = A.Renderer(0), A.Renderer(0)
    flat.upload(s)
    two.upload(s)
    fi = flat.build()
    ti = two.build(two_level=True)
    for r in (flat, two):
        r.set_film(W, H)
    log(f"flat: {fi}")
    log(f"two-level: {ti}")
    osc = O.OracleScene(s.desc)
    yield s, flat, two, fi, ti, osc
    flat.close()
    two.close()


def test_two_level_stores_each_mesh_once(inst_scene):
    s, flat, two, fi, ti, _ = inst_scene
    assert fi["num_triangles"] == ti["num_triangles"] == 256 * 20480 + 2 + 4
    assert fi["blas_count"] == 0
    assert ti["blas_count"] == 4 + 1 + 4  # the spheres, the ground quad, the four lights
    assert ti["blas_triangles"] == 4 * 20480 + 2 + 4
    assert ti["device_bytes"] * 20 < fi["device_bytes"]
    with pytest.raises(A.AsunaError):
        two.build(two_level=True, branching=2)
    ti2 = two.build(two_level=True)  # rebuild after the refused one
    assert ti2["blas_triangles"] == ti["blas_triangles"]


def test_first_hit_matches_oracle(inst_scene):
    s, flat, two, _, _, osc = inst_scene
    cam, _ = s.shot(0)
    g = two.first_hit(cam)
    f = flat.first_hit(cam)
    o = osc.first_hit(cam, W, H, threads=THREADS)
    same = (g["prim_id"] == o["prim_id"]) & (g["inst_id"] == o["inst_id"])
    hit = same & (o["inst_id"] >= 0)
    rel_t = np.abs(g["t"][hit] - o["t"][hit]) / np.abs(o["t"][hit])
    dn = np.abs(g["normal"][hit] - o["normal"][hit]).max(initial=0.0)
    flat_same = (f["prim_id"] == g["prim_id"]) & (f["inst_id"] == g["inst_id"])
    log(f"two-level first hit: id match {same.mean():.7f} ({int((~same).sum())} differ), hits {hit.mean():.4f}, "
        f"max rel t {rel_t.max(initial=0):.3g}, max |dN| {dn:.3g}; same ids as flat {flat_same.mean():.7f}")
    assert same.mean() >= 0.9999
    assert hit.mean() > 0.5
    assert rel_t.max(initial=0) <= 1e-5
    assert dn <= 1e-5


def test_radiance_within_three_sigma_of_oracle(inst_scene):
    s, flat, two, _, _, osc = inst_scene
    cam, pc = s.shot(0)
    c = 32
    origins = [(int(x), int(y)) for x, y in zip(np.linspace(100, 830, 8), np.linspace(180, 470, 8))]
    two.render(cam, pc, 0, 16)
    img = two.read(0)
    got = np.stack([img[y:y + c, x:x + c, :3] for x, y in origins])
    images = np.zeros((9, H, W, 4), np.float32)
    sq = np.zeros((H, W, 4), np.float32)
    for x, y in origins:
        osc.render(cam, pc, W, H, 0, 16, images, crop=(x, y, x + c, y + c), threads=THREADS, sq=sq)
    mean = np.stack([images[0][y:y + c, x:x + c, :3] for x, y in origins])
    sq16 = np.stack([sq[y:y + c, x:x + c, :3] for x, y in origins])
    var = np.maximum(sq16 / 16 - mean ** 2, 0.0)
    se = np.sqrt(2 * var / 16) + 1e-6
    bad = (np.abs(got - mean) > 3 * se + 1e-5).any(-1).mean()
    log(f"two-level crops at 16 spp: outside 3 sigma {bad:.5f}; means {got.mean((0, 1, 2))} vs {mean.mean((0, 1, 2))}")
    assert bad <= 0.01
    # throughput of both trees on the full film (logged for the record)
    for name, r in (("flat", flat), ("two-level", two)):
        r.render(cam, pc, 0, 1)
        t0 = time.perf_counter()
        st = r.render(cam, pc, 0, 4)
        dt = time.perf_counter() - t0
        log(f"{name}: 4 spp of {W}x{H} in {dt * 1e3:.1f} ms ({4 * W * H / dt / 1e6:.1f} Mpaths/s), {st}")


def test_large_instance_grid_matches_flat(tmp_path):
    """48 x 48 instances (47.2 M triangles in the world, the same 81,926
    stored): the two-level tree takes under 1/400 of the flat tree's
    device bytes and finds the flat tree's first hits on >= 99.95 % of the
    960x540 pixels, bit for bit in t and normal where they agree.  The
    two-level walk tests the ray against object-space triangles, so on a
    shared edge (sub-pixel at this distance) it may take the neighbour of the
    flat tree's triangle: a differing pixel counts as such a tie when both
    hit within 1e-4 relative in t; other differences (silhouettes) stay
    below 2e-4 of the pixels."""
    s = A.HostScene(instanced_scene(str(tmp_path), 48, (0.0, 14.0, 30.0)))
    assert s.state.num_triangles == 48 * 48 * 20480 + 2 + 4
    cam, _ = s.shot(0)
    flat, two = A.Renderer(0), A.Renderer(0)
    try:
        flat.upload(s)
        two.upload(s)
        fi = flat.build()
        ti = two.build(two_level=True)
        log(f"48 x 48 instances: flat {fi['device_bytes'] / 1e9:.2f} GB in {fi['build_ms']:.1f} ms, "
            f"two-level {ti['device_bytes'] / 1e6:.1f} MB in {ti['build_ms']:.1f} ms")
        assert ti["blas_triangles"] == 4 * 20480 + 2 + 4
        assert ti["device_bytes"] * 400 < fi["device_bytes"]
        for r in (flat, two):
            r.set_film(W, H)
        f = flat.first_hit(cam)
        g = two.first_hit(cam)
        same = (f["prim_id"] == g["prim_id"]) & (f["inst_id"] == g["inst_id"])
        hit = same & (f["inst_id"] >= 0)
        both = ~same & (f["inst_id"] >= 0) & (g["inst_id"] >= 0)
        tie = both & (np.abs(f["t"] - g["t"]) <= 1e-4 * np.abs(f["t"]))
        other = ~same & ~tie
        log(f"two-level vs flat: same ids {same.mean():.7f}, hits {hit.mean():.4f}, "
            f"t equal {np.mean(f['t'][hit] == g['t'][hit]):.7f}; differing: {int(tie.sum())} edge ties, "
            f"{int(other.sum())} other")
        assert same.mean() >= 0.9995
        assert other.mean() <= 2e-4
        assert hit.mean() > 0.5
        np.testing.assert_array_equal(f["t"][hit], g["t"][hit])
        np.testing.assert_array_equal(f["normal"][hit], g["normal"][hit])
    finally:
        flat.close()
        two.close()
