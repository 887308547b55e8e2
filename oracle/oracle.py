"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product never
imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    import asuna_amd as A  # struct definitions only (shared POD layout)
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    vp = C.c_void_p
    F = P(C.c_float)
    sigs = {
        "orc_xxhash32_seed": (C.c_uint32, [C.c_uint32] * 3),
        "orc_pcg": (C.c_uint32, [P(C.c_uint32)]),
        "orc_rand_stream": (None, [C.c_uint32, C.c_int, F]),
        "orc_camera_perspective": (None, [C.c_uint32, C.c_uint32, C.c_float, F, F, F, C.c_float, C.c_float, F,
                                          P(A.GpuCamera)]),
        "orc_camera_ray": (None, [P(A.GpuCamera), C.c_float, C.c_float, F, F]),
        "orc_camera_ray_lens": (None, [P(A.GpuCamera), C.c_float, C.c_float, C.c_float, C.c_float, F, F]),
        "orc_env_tables": (None, [C.c_uint32, C.c_uint32, F, F, F]),
        "orc_texture_bilinear": (None, [C.c_uint32, C.c_uint32, F, C.c_float, C.c_float, F]),
        "orc_scene_create": (C.c_int, [P(A.SceneDesc), P(vp)]),
        "orc_scene_destroy": (None, [vp]),
        "orc_scene_num_triangles": (C.c_uint64, [vp]),
        "orc_trace_closest": (C.c_int, [vp, F, F, C.c_float, C.c_float, P(A.Hit), C.c_int]),
        "orc_trace_any": (C.c_int, [vp, F, F, C.c_float, C.c_float, C.c_int]),
        "orc_first_hit": (C.c_int, [vp, P(A.GpuCamera), C.c_uint32, C.c_uint32, vp, C.c_int]),
        "orc_render": (C.c_int, [vp, P(A.GpuCamera), P(A.GpuPushConstantRaytrace), C.c_uint32, C.c_uint32,
                                 C.c_int32, C.c_int32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int,
                                 C.c_int, F, F, P(C.c_uint64), C.c_int32]),
        "orc_concentric_disk": (None, [C.c_float, C.c_float, F]),
        "orc_cosine_hemisphere": (None, [C.c_float, C.c_float, F]),
        "orc_offset_ray": (None, [F, F, F]),
        "orc_basis": (None, [F, F, F]),
        "orc_power_heuristic": (C.c_float, [C.c_float, C.c_float]),
        "orc_compute_diffuse_fresnel": (C.c_float, [C.c_float, C.c_int]),
        "orc_bsdf_eval": (None, [P(A.GpuMaterial), F, F, C.c_uint32, F]),
        "orc_bsdf_pdf": (C.c_float, [P(A.GpuMaterial), F, F, C.c_uint32]),
        "orc_conductor_reflectance": (C.c_float, [C.c_float, C.c_float, C.c_float]),
        "orc_rough_plastic_light_terms": (None, [P(A.GpuMaterial), F, F, C.c_uint32, F, F]),
        "orc_bsdf_sample": (None, [P(A.GpuMaterial), F, P(C.c_uint32), F, F, F, P(C.c_uint32)]),
        "orc_sample_envmap": (None, [vp, F, C.c_float, C.c_float, C.c_float, F, F, F]),
        "orc_pdf_envmap": (C.c_float, [vp, F, F]),
        "orc_eval_envmap": (None, [vp, F, C.c_float, F, F]),
        "orc_set_sun_and_sky": (None, [vp, P(A.GpuSunAndSky)]),
        "orc_sample_light": (None, [P(A.GpuLight), C.c_float, C.c_float, F, F, F, F, F, P(C.c_uint32)]),
        "orc_sun_and_sky": (None, [P(A.GpuSunAndSky), F, C.c_int, F]),
        "orc_sky_physical_scale": (None, [C.c_float, C.c_float, C.c_float, F]),
    }
    for n, (r, a) in sigs.items():
        f = getattr(L, n)
        f.restype = r
        f.argtypes = a
    _lib = L
    return L


def f3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


def arr(ptr_arr, n):
    return np.array([ptr_arr[i] for i in range(n)], np.float32)


class OracleScene:
    """The oracle's own flattened scene + binned-SAH BVH."""

    def __init__(self, desc):
        self._s = C.c_void_p()
        rc = lib().orc_scene_create(C.byref(desc), C.byref(self._s))
        if rc != 0:
            raise RuntimeError(f"orc_scene_create failed: {rc}")
        self.desc = desc

    def __del__(self):
        if getattr(self, "_s", None):
            lib().orc_scene_destroy(self._s)
            self._s = None

    @property
    def num_triangles(self):
        return lib().orc_scene_num_triangles(self._s)

    def first_hit(self, cam, w, h, threads=8):
        import asuna_amd as A
        out = np.zeros(w * h, A.HIT_DTYPE)
        lib().orc_first_hit(self._s, C.byref(cam), w, h, out.ctypes.data_as(C.c_void_p), threads)
        return out.reshape(h, w)

    def trace_closest(self, o, d, tmin=1e-5, tmax=1e10, brute=False):
        import asuna_amd as A
        hit = A.Hit()
        r = lib().orc_trace_closest(self._s, f3(o), f3(d), tmin, tmax, C.byref(hit), int(brute))
        return r, hit

    def set_sun_and_sky(self, sky):
        lib().orc_set_sun_and_sky(self._s, C.byref(sky))

    def trace_any(self, o, d, tmin, tmax, brute=False):
        return lib().orc_trace_any(self._s, f3(o), f3(d), tmin, tmax, int(brute))

    def render(self, cam, pc, w, h, first_frame, n_frames, images=None, crop=None, precision=0, threads=8,
               sq=None, depth_channel=-1):
        """Progressive render into images (9, h, w, 4) float32 (created when None)."""
        if images is None:
            images = np.zeros((9, h, w, 4), np.float32)
        x0, y0, x1, y1 = crop if crop else (0, 0, w, h)
        stats = (C.c_uint64 * 4)()
        sqp = sq.ctypes.data_as(C.POINTER(C.c_float)) if sq is not None else None
        rc = lib().orc_render(self._s, C.byref(cam), C.byref(pc), w, h, first_frame, n_frames, x0, y0, x1, y1,
                              precision, threads, images.ctypes.data_as(C.POINTER(C.c_float)), sqp, stats,
                              depth_channel)
        if rc != 0:
            raise RuntimeError(f"orc_render failed: {rc}")
        return images, {"paths": stats[0], "extension_rays": stats[1], "shadow_rays": stats[2],
                        "dielectric_shadow_rays": stats[3]}
