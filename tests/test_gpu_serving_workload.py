"""The serving workload on the GPU: 8 resident procedural scenes (about
10k, 100k, 1M and 5.2M triangles, two seeds each), a Poisson stream of 64
jobs at 200 jobs/s drawn from the serving mix (resolution 256^2 / 512^2 /
1024^2 / 1920x1080 with weights .4/.3/.2/.1, spp 1/4/16/64, denoise flag on
half), admitted by an HBM budget and co-batched by the server while later
jobs arrive.  Two jobs are pinned to the largest case: 1920x1080 at 64 spp on
each 5.2M-triangle scene, one with the denoise flag.

Bar: every job's radiance and weight images, and for flagged jobs the
denoised image, equal a single render of the same job (and the denoiser run
on that render's planes) bit for bit."""
import os
import tempfile
import time

import numpy as np
import pytest

import asuna_amd as A

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def scenes(tmp_path_factory):
    t0 = time.time()
    out = []
    for tris in (10_000, 100_000, 1_000_000, 5_000_000):
        for seed in (1, 2):
            d = os.path.join(tempfile.gettempdir(), f"asuna_serve_{tris}_{seed}")
            p = os.path.join(d, "spheres.json")
            if not os.path.exists(p):
                p = A.generate_scene(4, d, scale=tris, seed=seed)
            out.append(A.HostScene(p))
    log(f"8 scenes in {time.time() - t0:.1f}s: {[s.state.num_triangles for s in out]}")
    return out


def _jobs(scenes, rng, n):
    res = [(256, 256), (512, 512), (1024, 1024), (1920, 1080)]
    jobs = []
    for k in range(n):
        sid = int(rng.integers(0, len(scenes)))
        w, h = res[int(rng.choice(4, p=[0.4, 0.3, 0.2, 0.1]))]
        spp = int(rng.choice([1, 4, 16, 64]))
        dn = bool(rng.integers(0, 2))
        if k < 2:  # the largest case on both 5.2M-triangle scenes
            sid, w, h, spp, dn = 6 + k, 1920, 1080, 64, k == 0
        eye = rng.normal(size=3)
        eye = eye / np.linalg.norm(eye) * rng.uniform(3.0, 6.0)
        eye[1] = abs(eye[1]) + 0.5
        cam = A.camera_lookat(w, h, float(rng.uniform(35, 70)), tuple(eye), (0.0, 0.3, 0.0))
        _, pc = scenes[sid].shot(0)
        jobs.append((sid, cam, pc, w, h, spp, dn))
    return jobs


def test_poisson_stream_over_eight_resident_scenes(scenes):
    assert max(s.state.num_triangles for s in scenes) > 5_000_000
    rng = np.random.default_rng(0xA5A52026)
    jobs = _jobs(scenes, rng, 64)
    den = A.Denoiser(0, seed=7)
    srv = A.Server(0, hbm_budget_bytes=8 << 30, queue_paths=1 << 25)
    srv.set_denoiser(den)
    for s in scenes:
        srv.add_scene(s)
    arrivals = np.cumsum(rng.exponential(1.0 / 200.0, size=len(jobs)))
    base = time.perf_counter()
    ids = []
    for (sid, cam, pc, w, h, spp, dn), a in zip(jobs, arrivals):
        dt = a - (time.perf_counter() - base)
        if dt > 0:
            time.sleep(dt)
        ids.append(srv.submit(sid, cam, pc, w, h, spp, depth_channel=scenes[sid].state.depth_out_channel,
                              denoise=dn))
    status = [srv.wait(j, 120000) for j in ids]
    span = time.perf_counter() - base
    st = srv.stats()
    lat = [s["done_ms"] - s["submit_ms"] for s in status]
    log(f"64 jobs in {span:.2f}s ({len(jobs) / span:.1f} jobs/s), p50 {np.percentile(lat, 50):.1f} ms, "
        f"p99 {np.percentile(lat, 99):.1f} ms; server {st}")
    assert st["jobs_done"] == len(jobs)
    # jobs overlapped on the device: some job started before an earlier-submitted one finished
    assert any(status[k]["start_ms"] < status[k - 1]["done_ms"] for k in range(1, len(jobs)))
    rs = {}
    try:
        for k, ((sid, cam, pc, w, h, spp, dn), jid) in enumerate(zip(jobs, ids)):
            if sid not in rs:
                r = A.Renderer(0)
                r.upload(scenes[sid])
                r.build()
                rs[sid] = r
            r = rs[sid]
            dch = scenes[sid].state.depth_out_channel
            r.set_film(w, h)
            r.render(cam, pc, 0, spp, depth_channel=dch)
            for img in (0, 8):
                np.testing.assert_array_equal(srv.read(jid, img), r.read(img), err_msg=f"job {k} image {img}")
            if dn:
                den.denoise_film(r, pc.diffuseOutChannel, pc.normalOutChannel, dch, 7)
                np.testing.assert_array_equal(srv.read(jid, 9), r.read(7), err_msg=f"job {k} denoised")
            srv.release(jid)
            log(f"job {k}: scene {sid} {w}x{h} {spp} spp denoise {dn}: latency {lat[k]:.1f} ms, equal")
        assert srv.stats()["arena_used_bytes"] == 0
    finally:
        for r in rs.values():
            r.close()
        srv.close()
        den.close()
