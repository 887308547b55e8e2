"""Parity at the headline scale: the 1080p procedural scene (5.24 M
triangles, env map + 4 triangle lights, depth 6) on the default tree the
bench times (clustered build, 4-wide nodes).

Bars:
  first hit   primitive / instance ids identical to the oracle's on >= 99.99 %
              of the 2,073,600 pixels; t and shading normal within 1e-5 relative
  3 sigma     on 16 fixed 32x32 crops of shot 0 at 16 spp: per-pixel
              |gpu - oracle FP32| <= 3 sigma of the Monte Carlo standard error on
              >= 99 % of pixels
  relMSE      vs the FP64 CPU oracle at 4096 spp on the same crops
              (tests/golden/c2_reference_crops.npz), relMSE = mean over pixels
              and RGB of (x - ref)^2 / (ref^2 + 1e-2): at 16 spp the HIP
              render's equals the FP32 oracle's at 16 spp within 5 % (same
              estimator, same samples); at 256 spp it is below 1/8 of the 16-spp
              value (Monte Carlo convergence towards the oracle)
"""
import os

import numpy as np
import pytest

import asuna_amd as A
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "c2_reference_crops.npz")
THREADS = min(16, os.cpu_count() or 1)


def log(*a):
    print(*a, flush=True)


def relmse(x, ref):
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + 1e-2)))


@pytest.fixture(scope="module")
def c2(tmp_path_factory):
    d = tmp_path_factory.mktemp("c2")
    s = A.HostScene(A.generate_scene(1, str(d)))
    r = A.Renderer(0)
    r.upload(s)
    info = r.build()
    r.set_film(*s.film)
    log(f"C2 tree: {info}")
    osc = O.OracleScene(s.desc)
    yield s, r, osc, info
    r.close()


def test_first_hit_full_1080p_scene(c2):
    s, r, osc, info = c2
    assert info["num_triangles"] == 5242886 and info["branching"] == 4
    cam, _ = s.shot(0)
    w, h = s.film
    g = r.first_hit(cam)
    o = osc.first_hit(cam, w, h, threads=THREADS)
    same = (g["prim_id"] == o["prim_id"]) & (g["inst_id"] == o["inst_id"])
    hit = same & (o["inst_id"] >= 0)
    rel_t = np.abs(g["t"][hit] - o["t"][hit]) / np.abs(o["t"][hit])
    dn = np.abs(g["normal"][hit] - o["normal"][hit]).max(initial=0.0)
    log(f"C2 first hit: id match {same.mean():.7f} ({int((~same).sum())} pixels differ), hits {hit.mean():.4f}, "
        f"max rel t {rel_t.max(initial=0):.3g}, max |dN| {dn:.3g}, bit-exact t {np.mean(g['t'] == o['t']):.7f}")
    assert same.mean() >= 0.9999
    assert hit.mean() > 0.5
    assert rel_t.max(initial=0) <= 1e-5
    assert dn <= 1e-5


def test_crops_three_sigma_and_relmse_vs_fp64_oracle(c2):
    s, r, osc, _ = c2
    g = np.load(GOLD)
    c = int(g["crop"])
    origins = [tuple(int(v) for v in o) for o in g["origins"]]
    ref = g["ref"]
    w, h = s.film
    assert (w, h) == (int(g["width"]), int(g["height"]))
    assert s.state.num_triangles == int(g["triangles"])
    cam, pc = s.shot(0)

    def crops(img):
        return np.stack([img[y:y + c, x:x + c, :3] for x, y in origins])

    out = {}
    for spp in (16, 256):
        r.set_film(w, h)
        r.render(cam, pc, 0, spp)
        out[spp] = crops(r.read(0))
    images = np.zeros((9, h, w, 4), np.float32)
    sq = np.zeros((h, w, 4), np.float32)
    for x, y in origins:
        osc.render(cam, pc, w, h, 0, 16, images, crop=(x, y, x + c, y + c), threads=THREADS, sq=sq)
    mean, sq16 = crops(images[0]), crops(sq)
    var = np.maximum(sq16 / 16 - mean ** 2, 0.0)
    se = np.sqrt(2 * var / 16) + 1e-6  # difference of two independent means
    bad = (np.abs(out[16] - mean) > 3 * se + 1e-5).any(-1).mean()
    e16, e256, o16 = relmse(out[16], ref), relmse(out[256], ref), relmse(mean, ref)
    log(f"C2 crops at 16 spp: outside 3 sigma {bad:.5f}; means {out[16].mean((0, 1, 2))} vs {mean.mean((0, 1, 2))}")
    log(f"relMSE: gpu 16 spp {e16:.5f}, oracle FP32 16 spp {o16:.5f}, gpu 256 spp {e256:.6f}")
    assert bad <= 0.01
    assert abs(e16 - o16) <= 0.05 * o16
    assert e256 < e16 / 8
