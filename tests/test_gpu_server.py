"""Render server: resident scenes, jobs of mixed resolution / spp / depth
co-batched in shared wavefronts (several jobs per batch, several scenes on
their own streams at once).  Every job's images must equal a plain
ash_render of the same camera bit for bit: the RNG is keyed by (pixel, frame)
and every pixel folds its frames in order whatever it was batched with."""
import numpy as np
import pytest

import asuna_amd as A

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


def _jobs(scenes, rng):
    out = []
    for k in range(14):
        sid = k % len(scenes)
        s = scenes[sid]
        w, h = [(64, 48), (128, 128), (200, 120), (320, 180), (96, 256)][k % 5]
        spp = [1, 3, 4, 9, 16][(k * 3) % 5]
        eye = rng.normal(size=3)
        eye = eye / np.linalg.norm(eye) * rng.uniform(2.0, 4.0)
        eye[1] = abs(eye[1]) + 0.3
        cam = A.camera_lookat(w, h, float(rng.uniform(35, 70)), tuple(eye), (0.0, 0.3, 0.0))
        _, pc = s.shot(0)
        pc.maxPathDepth = [2, 4, 6][k % 3]
        out.append((sid, cam, pc, w, h, spp))
    return out


@pytest.mark.parametrize("policy", [A.SERVE_ROUND_ROBIN, A.SERVE_OLDEST_FIRST, A.SERVE_OLDEST_GLOBAL])
def test_server_jobs_equal_single_renders(box_scene, spheres_small, policy):
    scenes = [A.HostScene(box_scene), A.HostScene(spheres_small)]
    # a small batch so that jobs share batches and large jobs span several
    srv = A.Server(0, hbm_budget_bytes=64 << 20, queue_paths=1 << 16, batches_ahead=2, policy=policy)
    ids = [srv.add_scene(s) for s in scenes]
    assert ids == [0, 1]
    jobs = _jobs(scenes, np.random.default_rng(5))
    jid = [srv.submit(*j) for j in jobs]
    status = [srv.wait(j, 60000) for j in jid]
    st = srv.stats()
    log(f"server: {st}")
    assert st["jobs_done"] == len(jobs)
    assert st["batches"] < sum(w * h * spp for _, _, _, w, h, spp in jobs) / (1 << 16) + len(jobs)
    got = [[srv.read(j, img) for img in (0, 1, 8)] for j in jid]
    for j in jid:
        srv.release(j)
    assert srv.stats()["arena_used_bytes"] == 0
    srv.close()
    rs = []
    for s in scenes:
        r = A.Renderer(0)
        r.upload(s)
        r.build()
        rs.append(r)
    for k, (sid, cam, pc, w, h, spp) in enumerate(jobs):
        r = rs[sid]
        r.set_film(w, h)
        r.render(cam, pc, 0, spp)
        for img, g in zip((0, 1, 8), got[k]):
            np.testing.assert_array_equal(r.read(img), g, err_msg=f"job {k} image {img}")
        lat = status[k]["done_ms"] - status[k]["submit_ms"]
        log(f"job {k}: scene {sid} {w}x{h} spp {spp} depth {pc.maxPathDepth}: latency {lat:.1f} ms")
    for r in rs:
        r.close()


def test_server_admission_waits_for_budget(box_scene):
    """A budget that holds one 256x256 job at a time still completes all jobs."""
    s = A.HostScene(box_scene)
    one = 256 * 256 * 9 * 16
    srv = A.Server(0, hbm_budget_bytes=one + 4096, queue_paths=1 << 18)
    sid = srv.add_scene(s)
    cam, pc = s.shot(0)
    cam = A.camera_lookat(256, 256, 45.0, (0.0, 0.0, 3.0), (0.0, 0.0, 0.0))
    jid = [srv.submit(sid, cam, pc, 256, 256, 2) for _ in range(3)]
    st0 = srv.wait(jid[0], 60000)
    a = srv.read(jid[0], 0)
    srv.release(jid[0])
    st1 = srv.wait(jid[1], 60000)
    assert st1["start_ms"] >= st0["done_ms"]  # admitted only after job 0's images were released
    b = srv.read(jid[1], 0)
    srv.release(jid[1])
    srv.wait(jid[2], 60000)
    np.testing.assert_array_equal(a, b)
    srv.release(jid[2])
    with pytest.raises(A.AsunaError):
        srv.submit(sid, cam, pc, 1024, 1024, 1)  # larger than the whole budget
    srv.close()


def test_server_denoise_flag_equals_film_denoise(spheres_small):
    """A job with the denoise flag gets image 9 = the denoiser run on its own
    radiance / albedo / normal / depth planes, identical to denoising the
    same render in a plain context."""
    s = A.HostScene(spheres_small)
    st = s.state
    cam, pc = s.shot(0)
    w, h = s.film
    den = A.Denoiser(0, seed=3)
    srv = A.Server(0, hbm_budget_bytes=256 << 20, queue_paths=1 << 17)
    srv.set_denoiser(den)
    sid = srv.add_scene(s)
    jid = srv.submit(sid, cam, pc, w, h, 4, depth_channel=st.depth_out_channel, denoise=True)
    status = srv.wait(jid, 60000)
    got = srv.read(jid, 9)
    log(f"server denoise: {status['denoise_ms']:.2f} ms")
    assert status["denoise_ms"] > 0
    srv.release(jid)
    srv.close()
    r = A.Renderer(0)
    r.upload(s)
    r.build()
    r.set_film(w, h)
    r.render(cam, pc, 0, 4, depth_channel=st.depth_out_channel)
    den.denoise_film(r, pc.diffuseOutChannel, pc.normalOutChannel, st.depth_out_channel, 7)
    np.testing.assert_array_equal(got, r.read(7))
    r.close()
    den.close()
