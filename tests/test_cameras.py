"""Camera models and shot types against their closed forms (CPU):
the thin lens of raytrace.projective.rgen:55-69, the OpenCV intrinsics of
rgen:75-87 with the `opencv` shot of loader.cpp:430-437, and the `toworld`
shot of loader.cpp:421-429.  The product loader builds the cameras; the
oracle's camera ray is the one the GPU parity tests compare against."""
import ctypes as C
import json

import numpy as np

import asuna_amd as A
from oracle import oracle as O


def _ray(cam, px, py, l0=0.5, l1=0.5):
    o, d = (C.c_float * 3)(), (C.c_float * 3)()
    O.lib().orc_camera_ray_lens(C.byref(cam), px, py, l0, l1, o, d)
    return np.array(o, float), np.array(d, float)


def _unit(v):
    v = np.asarray(v, float)
    return v / np.linalg.norm(v)


def test_thin_lens_rays_of_a_pixel_meet_on_the_plane_of_focus():
    """pLens = aperture * concentric(u); pFocus = (f / r.z) r; the ray runs
    from pLens towards pFocus (its direction normalised by
    transformDirection, math.glsl:63-66), so every lens sample of a pixel
    passes through one world point: the point of the pinhole ray at depth
    f along the view axis.  The origins lie in the lens plane within the
    aperture, and the lens centre (u = (1/2, 1/2)) is the pinhole ray."""
    eye, look, f, ap = np.array([1.0, 2.0, 5.0]), np.array([0.2, -0.1, 0.0]), 3.5, 0.25
    cam = A.camera_lookat(320, 180, 55.0, tuple(eye), tuple(look), aperture=ap, focal=f)
    cam0 = A.camera_lookat(320, 180, 55.0, tuple(eye), tuple(look))
    fwd = _unit(look - eye)
    rng = np.random.default_rng(5)
    for px, py in [(160.0, 90.0), (20.5, 170.25), (300.0, 10.0)]:
        _, dp = _ray(cam0, px, py)
        o0, d0 = _ray(cam, px, py)
        np.testing.assert_allclose(o0, eye, atol=2e-6)
        np.testing.assert_allclose(d0, dp, atol=2e-6)
        P = eye + dp * f / np.dot(dp, fwd)  # the pinhole ray at depth f
        spread = 0.0
        for l0, l1 in rng.random((8, 2)):
            o, d = _ray(cam, px, py, float(l0), float(l1))
            assert abs(np.dot(o - eye, fwd)) < 2e-6
            assert np.linalg.norm(o - eye) <= ap * (1 + 1e-5)
            assert abs(np.linalg.norm(d) - 1) < 1e-6
            v = P - o
            miss = np.linalg.norm(v - np.dot(v, d) * d)
            assert miss < 3e-5 * np.linalg.norm(v), (px, py, l0, l1, miss)
            spread = max(spread, np.linalg.norm(o - eye))
        assert spread > 0.1 * ap  # the samples did move over the lens


def _scene_with_camera(tmp_path, camera, shot):
    (tmp_path / "tri.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nf 1//1 2//1 3//1\n")
    js = {"state": {"path_tracing": {"spp": 1}}, "camera": camera,
          "materials": [{"name": "m", "type": "brdf_lambertian", "diffuse_reflectance": [0.5, 0.5, 0.5]}],
          "meshes": [{"name": "tri", "path": "tri.obj"}], "instances": [{"mesh": "tri", "material": "m"}],
          "shots": [shot]}
    p = tmp_path / "scene.json"
    p.write_text(json.dumps(js))
    return A.HostScene(str(p))


def _rotation(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def test_opencv_camera_inverts_the_pinhole_projection(tmp_path):
    """An `opencv` shot carries the extrinsic [R | t] (world -> camera, x
    right, y down, z forward) and the `opencv` camera the intrinsics; the
    ray of pixel (u, v) must start at the camera centre -R^T t and run along
    R^T ((u - cx)/fx, (v - cy)/fy, 1): the inverse of x = K (R X + t)."""
    rng = np.random.default_rng(17)
    R = _rotation(rng)
    t = rng.normal(size=3) * 2
    ext = np.eye(4)
    ext[:3, :3], ext[:3, 3] = R, t
    fx, fy, cx, cy = 410.0, 395.0, 170.5, 88.25
    s = _scene_with_camera(tmp_path, {"type": "opencv", "film": {"resolution": [320, 180]},
                                      "fx": fx, "fy": fy, "cx": cx, "cy": cy},
                           {"type": "opencv", "matrix": [float(v) for v in ext.ravel()]})
    cam, _ = s.shot(0)
    assert cam.type == 1 and list(cam.fxfycxcy) == [fx, fy, cx, cy]
    centre = -R.T @ t
    for u, v in [(cx, cy), (0.5, 0.5), (319.5, 179.5), (40.25, 150.0)]:
        o, d = _ray(cam, u, v)
        np.testing.assert_allclose(o, centre, atol=3e-5 * (1 + np.abs(centre).max()))
        want = R.T @ np.array([(u - cx) / fx, (v - cy) / fy, 1.0])
        np.testing.assert_allclose(_unit(d), _unit(want), atol=3e-6)
    # a world point projects to the pixel whose ray passes through it
    X = centre + R.T @ np.array([0.3, -0.2, 4.0])
    xc = R @ X + t
    u, v = fx * xc[0] / xc[2] + cx, fy * xc[1] / xc[2] + cy
    o, d = _ray(cam, u, v)
    np.testing.assert_allclose(_unit(X - o), _unit(d), atol=3e-6)


def test_toworld_shot_looks_along_the_matrix_axes(tmp_path):
    """A `toworld` shot is a Mitsuba camera-to-world matrix M (x left, y up,
    z forward): eye = M's translation, the view axis M z, up M y.  The ray
    of the film centre runs along M z; a pixel right of the centre leans
    towards -M x, one below it towards -M y (raster y points down)."""
    rng = np.random.default_rng(23)
    M = np.eye(4)
    M[:3, :3] = _rotation(rng)
    M[:3, 3] = rng.normal(size=3)
    s = _scene_with_camera(tmp_path, {"type": "perspective", "film": {"resolution": [200, 100]}, "fov": 40},
                           {"type": "toworld", "matrix": [float(v) for v in M.ravel()]})
    cam, _ = s.shot(0)
    x, y, z = M[:3, 0], M[:3, 1], M[:3, 2]
    o, d = _ray(cam, 100.0, 50.0)
    np.testing.assert_allclose(o, M[:3, 3], atol=2e-6)
    np.testing.assert_allclose(_unit(d), z, atol=2e-6)
    _, dr = _ray(cam, 150.0, 50.0)
    assert np.dot(_unit(dr), -x) > 0.1 and abs(np.dot(_unit(dr), y)) < 2e-6
    _, db = _ray(cam, 100.0, 80.0)
    assert np.dot(_unit(db), -y) > 0.1 and abs(np.dot(_unit(db), x)) < 2e-6
    # horizontal fov: the film's left edge is 20 degrees off the axis
    _, de = _ray(cam, 0.0, 50.0)
    assert abs(np.degrees(np.arccos(np.clip(np.dot(_unit(de), z), -1, 1))) - 20.0) < 1e-3
