"""asuna_amd -- MI355X-native render path with Asuna's scene format and estimator.

Thin ctypes layer over ``asuna_amd/lib/libasuna.so`` (C ABI in
``include/asuna_hip.h`` and ``include/asuna_host.h``).  The library holds
the gfx950 HIP kernels and the C++ host code (loader, scene, offline loop);
this module only marshals arguments.  It raises when the library is missing:
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libasuna.so")

ASH_OK = 0
ASH_NUM_OUTPUT_IMAGES = 9
ERRORS = {-1: "invalid argument", -2: "HIP runtime failure", -3: "call out of order",
          -4: "device allocation failed", -5: "i/o failure", -6: "malformed scene file",
          -7: "unsupported feature"}


# ----------------------------------------------------------------- PODs
class GpuVertex(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("uv", C.c_float * 2), ("normal", C.c_float * 3),
                ("tangent", C.c_float * 3)]


class GpuMaterial(C.Structure):
    _fields_ = [("diffuse", C.c_float * 3), ("rhoSpec", C.c_float * 3), ("anisoAlpha", C.c_float * 2),
                ("ior", C.c_float), ("roughness", C.c_float), ("subsurface", C.c_float),
                ("specular", C.c_float), ("specularTint", C.c_float), ("anisotropic", C.c_float),
                ("sheen", C.c_float), ("sheenTint", C.c_float), ("clearcoat", C.c_float),
                ("clearcoatGloss", C.c_float), ("radiance", C.c_float * 3), ("metalness", C.c_float),
                ("radianceFactor", C.c_float * 3), ("diffuseTextureId", C.c_int32),
                ("roughnessTextureId", C.c_int32), ("metalnessTextureId", C.c_int32),
                ("radianceTextureId", C.c_int32), ("normalTextureId", C.c_int32),
                ("tangentTextureId", C.c_int32), ("opacityTextureId", C.c_int32), ("type", C.c_uint32)]


class GpuSunAndSky(C.Structure):
    _fields_ = [("rgb_unit_conversion", C.c_float * 3), ("multiplier", C.c_float), ("haze", C.c_float),
                ("redblueshift", C.c_float), ("saturation", C.c_float), ("horizon_height", C.c_float),
                ("ground_color", C.c_float * 3), ("horizon_blur", C.c_float), ("night_color", C.c_float * 3),
                ("sun_disk_intensity", C.c_float), ("sun_direction", C.c_float * 3), ("sun_disk_scale", C.c_float),
                ("sun_glow_intensity", C.c_float), ("y_is_up", C.c_int32), ("physically_scaled_sun", C.c_int32),
                ("in_use", C.c_int32)]


def sun_and_sky_defaults():
    """The sun & sky record a scene starts with (scene.cpp:112-129; in_use 0)."""
    s = GpuSunAndSky()
    lib().ash_sun_and_sky_defaults(C.byref(s))
    return s


class GpuLight(C.Structure):
    _fields_ = [("type", C.c_int32), ("position", C.c_float * 3), ("direction", C.c_float * 3),
                ("radiance", C.c_float * 3), ("u", C.c_float * 3), ("v", C.c_float * 3),
                ("radius", C.c_float), ("area", C.c_float), ("doubleSide", C.c_uint32)]


class GpuCamera(C.Structure):
    _fields_ = [("rasterToCamera", C.c_float * 16), ("cameraToWorld", C.c_float * 16),
                ("envTransform", C.c_float * 16), ("fxfycxcy", C.c_float * 4), ("type", C.c_uint32),
                ("aperture", C.c_float), ("focalDistance", C.c_float), ("padding", C.c_float)]


class GpuPushConstantRaytrace(C.Structure):
    _fields_ = [("spp", C.c_int32), ("curFrame", C.c_int32), ("maxPathDepth", C.c_int32),
                ("numLights", C.c_int32), ("bgColor", C.c_float * 3), ("useFaceNormal", C.c_uint32),
                ("ignoreEmissive", C.c_uint32), ("hasEnvMap", C.c_uint32),
                ("envMapResolution", C.c_float * 2), ("envMapIntensity", C.c_float),
                ("nMultiChannel", C.c_uint32), ("diffuseOutChannel", C.c_int32),
                ("specularOutChannel", C.c_int32), ("roughnessOutChannel", C.c_int32),
                ("normalOutChannel", C.c_int32), ("positionOutChannel", C.c_int32),
                ("tangentOutChannel", C.c_int32), ("uvOutChannel", C.c_int32)]


class MeshDesc(C.Structure):
    _fields_ = [("num_vertices", C.c_uint32), ("num_indices", C.c_uint32),
                ("vertices", C.POINTER(GpuVertex)), ("indices", C.POINTER(C.c_uint32))]


class InstanceDesc(C.Structure):
    _fields_ = [("transform", C.c_float * 16), ("mesh_id", C.c_uint32), ("material_id", C.c_uint32),
                ("light_id", C.c_int32), ("pad", C.c_uint32)]


class ImageDesc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgba", C.POINTER(C.c_float))]


class SceneDesc(C.Structure):
    _fields_ = [("num_meshes", C.c_uint32), ("meshes", C.POINTER(MeshDesc)),
                ("num_instances", C.c_uint32), ("instances", C.POINTER(InstanceDesc)),
                ("num_materials", C.c_uint32), ("materials", C.POINTER(GpuMaterial)),
                ("num_lights", C.c_uint32), ("lights", C.POINTER(GpuLight)),
                ("num_textures", C.c_uint32), ("textures", C.POINTER(ImageDesc)),
                ("envmap", ImageDesc)]


class Hit(C.Structure):
    _fields_ = [("t", C.c_float), ("u", C.c_float), ("v", C.c_float), ("inst_id", C.c_int32),
                ("prim_id", C.c_int32), ("normal", C.c_float * 3), ("pos", C.c_float * 3)]


HIT_DTYPE = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("inst_id", "<i4"), ("prim_id", "<i4"),
                      ("normal", "<f4", 3), ("pos", "<f4", 3)])
assert HIT_DTYPE.itemsize == C.sizeof(Hit)


class Stats(C.Structure):
    _fields_ = [("paths", C.c_uint64), ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("dielectric_shadow_rays", C.c_uint64), ("render_ms", C.c_double), ("extend_ms", C.c_double),
                ("shadow_ms", C.c_double), ("shade_ms", C.c_double), ("extend_launches", C.c_uint64),
                ("shadow_launches", C.c_uint64), ("ext_node_visits", C.c_uint64), ("ext_tri_tests", C.c_uint64),
                ("shadow_node_visits", C.c_uint64), ("shadow_tri_tests", C.c_uint64), ("batches", C.c_uint64),
                ("graph_launches", C.c_uint64), ("trace_launches", C.c_uint64), ("trace_ms", C.c_double),
                ("trace_ext_rays", C.c_uint64), ("trace_shadow_rays", C.c_uint64),
                ("trace_ext_node_visits", C.c_uint64), ("trace_ext_tri_tests", C.c_uint64),
                ("trace_shadow_node_visits", C.c_uint64), ("trace_shadow_tri_tests", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BuildOpts(C.Structure):
    _fields_ = [("leaf_max_tris", C.c_uint32), ("quality", C.c_uint32), ("branching", C.c_uint32),
                ("node_format", C.c_uint32), ("two_level", C.c_uint32)]


class BuildInfo(C.Structure):
    _fields_ = [("num_triangles", C.c_uint64), ("num_nodes", C.c_uint64), ("max_depth", C.c_uint32),
                ("node_bytes", C.c_uint32), ("build_ms", C.c_double), ("sah_cost", C.c_double),
                ("device_bytes", C.c_uint64), ("num_wide_nodes", C.c_uint64), ("wide_depth", C.c_uint32),
                ("branching", C.c_uint32), ("blas_count", C.c_uint32), ("pad", C.c_uint32),
                ("blas_triangles", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


class RenderOpts(C.Structure):
    _fields_ = [("queue_paths", C.c_uint32), ("blocks_per_cu", C.c_uint32), ("timing", C.c_uint32),
                ("depthOutChannel", C.c_int32), ("max_bounce_iters", C.c_uint32), ("count_traversal", C.c_uint32),
                ("no_graphs", C.c_uint32), ("lds_stack", C.c_uint32), ("traversal", C.c_uint32)]

    @staticmethod
    def make(queue_paths=0, blocks_per_cu=0, timing=False, depth_channel=-1, max_bounce_iters=0,
             count_traversal=False, graphs=True, lds_stack=0, traversal=0):
        return RenderOpts(queue_paths, blocks_per_cu, int(bool(timing)), depth_channel, max_bounce_iters,
                          int(bool(count_traversal)), 0 if graphs else 1, lds_stack, traversal)


class TuneResult(C.Structure):
    _fields_ = [("queue_paths", C.c_uint32), ("blocks_per_cu", C.c_uint32), ("candidates", C.c_uint32),
                ("lds_stack", C.c_uint32), ("traversal", C.c_uint32), ("pad", C.c_uint32),
                ("best_ms", C.c_double), ("worst_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerOpts(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("queue_paths", C.c_uint32), ("blocks_per_cu", C.c_uint32),
                ("max_jobs_per_batch", C.c_uint32), ("batches_ahead", C.c_uint32), ("policy", C.c_uint32),
                ("pad", C.c_uint32)]


class JobDesc(C.Structure):
    _fields_ = [("scene", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_int32),
                ("camera", GpuCamera), ("pc", GpuPushConstantRaytrace), ("depth_channel", C.c_int32),
                ("flags", C.c_uint32)]


class JobStatus(C.Structure):
    _fields_ = [("state", C.c_int32), ("pad", C.c_uint32), ("submit_ms", C.c_double), ("start_ms", C.c_double),
                ("done_ms", C.c_double), ("paths", C.c_uint64), ("denoise_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


class ServerCounters(C.Structure):
    _fields_ = [("jobs_submitted", C.c_uint64), ("jobs_done", C.c_uint64), ("batches", C.c_uint64),
                ("paths", C.c_uint64), ("arena_bytes", C.c_uint64), ("arena_used_bytes", C.c_uint64),
                ("busy_ms", C.c_double), ("uptime_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SceneState(C.Structure):
    _fields_ = [("film_width", C.c_uint32), ("film_height", C.c_uint32), ("num_shots", C.c_uint32),
                ("tone_mapping", C.c_uint32), ("output_hdr", C.c_uint32),
                ("output_render_result", C.c_uint32), ("channel_output_ldr", C.c_uint32),
                ("depth_out_channel", C.c_int32), ("num_triangles", C.c_uint64)]


class OfflineOpts(C.Structure):
    _fields_ = [("gpu_id", C.c_int32), ("output_scanline", C.c_int32), ("spp_override", C.c_int32),
                ("quiet", C.c_int32), ("render", RenderOpts)]


class AsunaError(RuntimeError):
    pass


_lib = None


def lib():
    """The loaded libasuna.so (raises AsunaError when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise AsunaError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                         "(make -C asuna_amd/csrc)")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    vp = C.c_void_p
    sigs = {
        "ash_create": (C.c_int, [C.c_int, P(vp)]),
        "ash_destroy": (None, [vp]),
        "ash_last_error": (C.c_char_p, [vp]),
        "ash_upload_scene": (C.c_int, [vp, P(SceneDesc)]),
        "ash_build_accel": (C.c_int, [vp, P(BuildOpts), P(BuildInfo)]),
        "ash_set_film": (C.c_int, [vp, C.c_uint32, C.c_uint32]),
        "ash_sun_and_sky_defaults": (None, [P(GpuSunAndSky)]),
        "ash_set_sun_and_sky": (C.c_int, [vp, P(GpuSunAndSky)]),
        "ash_render": (C.c_int, [vp, P(GpuCamera), P(GpuPushConstantRaytrace), C.c_int32, C.c_int32,
                                 P(RenderOpts), P(Stats)]),
        "ash_render_noalloc": (C.c_int, [vp, P(GpuCamera), P(GpuPushConstantRaytrace), C.c_int32, C.c_int32,
                                         C.c_uint32, C.c_uint32, vp, P(Stats)]),
        "ash_render_shots": (C.c_int, [vp, C.c_uint32, P(GpuCamera), P(GpuPushConstantRaytrace), C.c_int32,
                                       P(RenderOpts), P(Stats)]),
        "ash_autotune": (C.c_int, [vp, C.c_uint32, P(GpuCamera), P(GpuPushConstantRaytrace), C.c_int32,
                                   P(RenderOpts), P(TuneResult)]),
        "ash_read_shot_output": (C.c_int, [vp, C.c_uint32, C.c_int32, P(C.c_float)]),
        "ash_denoiser_create": (C.c_int, [C.c_int, C.c_uint32, P(vp)]),
        "ash_denoiser_destroy": (None, [vp]),
        "ash_denoiser_last_error": (C.c_char_p, [vp]),
        "ash_denoiser_num_params": (C.c_int, [vp, P(C.c_uint64)]),
        "ash_denoiser_get_params": (C.c_int, [vp, P(C.c_float)]),
        "ash_denoiser_set_params": (C.c_int, [vp, P(C.c_float)]),
        "ash_denoiser_forward": (C.c_int, [vp, C.c_uint32, C.c_uint32, C.c_uint32, P(C.c_float), P(C.c_float)]),
        "ash_denoiser_gradient": (C.c_int, [vp, C.c_uint32, C.c_uint32, C.c_uint32, P(C.c_float), P(C.c_float),
                                            P(C.c_float), P(C.c_float)]),
        "ash_denoiser_train_step": (C.c_int, [vp, C.c_uint32, C.c_uint32, C.c_uint32, P(C.c_float), P(C.c_float),
                                              C.c_float, P(C.c_float)]),
        "ash_denoise_film": (C.c_int, [vp, vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, P(C.c_double)]),
        "ash_server_create": (C.c_int, [C.c_int, P(ServerOpts), P(vp)]),
        "ash_server_destroy": (None, [vp]),
        "ash_server_last_error": (C.c_char_p, [vp]),
        "ash_server_add_scene": (C.c_int, [vp, P(SceneDesc), P(BuildOpts), P(C.c_uint32)]),
        "ash_server_submit": (C.c_int, [vp, P(JobDesc), P(C.c_uint64)]),
        "ash_server_set_denoiser": (C.c_int, [vp, vp]),
        "ash_denoise_planes": (C.c_int, [vp, vp, vp, vp, vp, C.c_uint32, C.c_uint32, vp, P(C.c_double)]),
        "ash_server_wait": (C.c_int, [vp, C.c_uint64, C.c_int32, P(JobStatus)]),
        "ash_server_read": (C.c_int, [vp, C.c_uint64, C.c_int32, P(C.c_float)]),
        "ash_server_release": (C.c_int, [vp, C.c_uint64]),
        "ash_server_stats": (C.c_int, [vp, P(ServerCounters)]),
        "ash_read_output": (C.c_int, [vp, C.c_int32, P(C.c_float)]),
        "ash_output_device_ptr": (C.c_int, [vp, C.c_int32, P(vp)]),
        "ash_first_hit": (C.c_int, [vp, P(GpuCamera), vp]),
        "ash_env_tables": (C.c_int, [vp, P(C.c_float), P(C.c_float)]),
        "ash_pod_layout": (None, [P(C.c_uint32)]),
        "ash_version": (C.c_char_p, []),
        "ash_scene_load": (C.c_int, [C.c_char_p, P(vp), C.c_char_p, C.c_size_t]),
        "ash_scene_free": (None, [vp]),
        "ash_scene_get_desc": (C.c_int, [vp, P(SceneDesc)]),
        "ash_scene_get_state": (C.c_int, [vp, P(SceneState)]),
        "ash_scene_shot": (C.c_int, [vp, C.c_uint32, P(GpuCamera), P(GpuPushConstantRaytrace)]),
        "ash_camera_lookat": (None, [C.c_uint32, C.c_uint32, C.c_float, P(C.c_float), P(C.c_float),
                                     P(C.c_float), C.c_float, C.c_float, P(C.c_float), P(GpuCamera)]),
        "ash_run_offline": (C.c_int, [C.c_char_p, C.c_char_p, P(OfflineOpts), P(Stats), C.c_char_p,
                                      C.c_size_t]),
        "ash_write_image": (C.c_int, [C.c_char_p, C.c_uint32, C.c_uint32, P(C.c_float)]),
        "ash_read_image": (C.c_int, [C.c_char_p, C.c_float, P(C.c_uint32), P(C.c_uint32),
                                     P(P(C.c_float))]),
        "ash_free_image": (None, [P(C.c_float)]),
        "ash_tonemap": (None, [C.c_uint32, C.c_uint32, C.c_uint32, P(C.c_float)]),
        "ash_env_tables_host": (None, [C.c_uint32, C.c_uint32, P(C.c_float), P(C.c_float), P(C.c_float)]),
        "ash_generate_scene": (C.c_int, [C.c_int32, C.c_uint64, C.c_int32, C.c_char_p, C.c_char_p,
                                         C.c_size_t]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_float))


# ------------------------------------------------------------ host scene
class HostScene:
    """A scene file loaded by the library's Loader (JSON + OBJ + images)."""

    def __init__(self, path: str):
        L = lib()
        self._h = C.c_void_p()
        err = C.create_string_buffer(4096)
        rc = L.ash_scene_load(path.encode(), C.byref(self._h), err, len(err))
        if rc != ASH_OK:
            self._h = None
            raise AsunaError(f"scene load failed ({ERRORS.get(rc, rc)}): {err.value.decode()}")
        self.path = path
        self.desc = SceneDesc()
        L.ash_scene_get_desc(self._h, C.byref(self.desc))
        self.state = SceneState()
        L.ash_scene_get_state(self._h, C.byref(self.state))

    def close(self):
        if self._h:
            lib().ash_scene_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def film(self):
        return self.state.film_width, self.state.film_height

    def shot(self, i: int = 0):
        cam, pc = GpuCamera(), GpuPushConstantRaytrace()
        rc = lib().ash_scene_shot(self._h, i, C.byref(cam), C.byref(pc))
        if rc != ASH_OK:
            raise AsunaError(f"bad shot index {i}")
        return cam, pc

    def meshes(self):
        out = []
        for m in range(self.desc.num_meshes):
            md = self.desc.meshes[m]
            v = np.ctypeslib.as_array(C.cast(md.vertices, C.POINTER(C.c_float)), (md.num_vertices * 11,))
            i = np.ctypeslib.as_array(md.indices, (md.num_indices,))
            out.append((v.reshape(-1, 11).copy(), i.copy()))
        return out


def camera_lookat(width, height, fov, eye, lookat, up=(0, 1, 0), aperture=0.0, focal=0.1, env=None):
    cam = GpuCamera()
    e = (C.c_float * 3)(*eye)
    l_ = (C.c_float * 3)(*lookat)
    u = (C.c_float * 3)(*up)
    envp = (C.c_float * 16)(*env) if env is not None else None
    lib().ash_camera_lookat(width, height, fov, e, l_, u, aperture, focal, envp, C.byref(cam))
    return cam


def generate_scene(config: int, directory: str, seed: int = 0xA5A52026, scale: int = 0) -> str:
    os.makedirs(directory, exist_ok=True)
    buf = C.create_string_buffer(4096)
    rc = lib().ash_generate_scene(config, seed, scale, directory.encode(), buf, len(buf))
    if rc != ASH_OK:
        raise AsunaError(f"scene generation failed: {ERRORS.get(rc, rc)}")
    return buf.value.decode()


def write_image(path: str, rgba: np.ndarray):
    rgba = np.ascontiguousarray(rgba, dtype=np.float32)
    h, w = rgba.shape[:2]
    if lib().ash_write_image(path.encode(), w, h, fptr(rgba)) != ASH_OK:
        raise AsunaError(f"cannot write {path}")


def read_image(path: str, gamma: float = 1.0) -> np.ndarray:
    w, h = C.c_uint32(), C.c_uint32()
    p = C.POINTER(C.c_float)()
    if lib().ash_read_image(path.encode(), gamma, C.byref(w), C.byref(h), C.byref(p)) != ASH_OK:
        raise AsunaError(f"cannot read {path}")
    a = np.ctypeslib.as_array(p, (h.value, w.value, 4)).copy()
    lib().ash_free_image(p)
    return a


def env_tables_host(rgba: np.ndarray):
    """Env importance tables as the library builds them at upload: (marginal (H,2), conditional (H,W,2))."""
    rgba = np.ascontiguousarray(rgba, dtype=np.float32)
    h, w = rgba.shape[:2]
    m = np.zeros((h, 2), np.float32)
    c = np.zeros((h, w, 2), np.float32)
    lib().ash_env_tables_host(w, h, fptr(rgba), fptr(m), fptr(c))
    return m, c


def pod_layout():
    out = (C.c_uint32 * 10)()
    lib().ash_pod_layout(out)
    return list(out)


# ----------------------------------------------------------- render ctx
class Renderer:
    """A device context of the render library (one GPU)."""

    def __init__(self, device: int = 0):
        L = lib()
        self._c = C.c_void_p()
        rc = L.ash_create(device, C.byref(self._c))
        if rc != ASH_OK:
            self._c = None
            raise AsunaError(f"ash_create({device}) failed: {ERRORS.get(rc, rc)}")
        self.width = self.height = 0

    def _check(self, rc, what):
        if rc != ASH_OK:
            msg = lib().ash_last_error(self._c).decode()
            raise AsunaError(f"{what} failed ({ERRORS.get(rc, rc)}): {msg}")

    def close(self):
        if self._c:
            lib().ash_destroy(self._c)
            self._c = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, scene: HostScene):
        self._check(lib().ash_upload_scene(self._c, C.byref(scene.desc)), "upload_scene")

    def upload_desc(self, desc: SceneDesc):
        self._check(lib().ash_upload_scene(self._c, C.byref(desc)), "upload_scene")

    def build(self, leaf_max_tris=4, quality=2, branching=4, node_format=0, two_level=False) -> dict:
        info = BuildInfo()
        opts = BuildOpts(leaf_max_tris, quality, branching, node_format, int(bool(two_level)))
        self._check(lib().ash_build_accel(self._c, C.byref(opts), C.byref(info)),
                    "build_accel")
        return info.as_dict()

    def set_sun_and_sky(self, sky):
        self._check(lib().ash_set_sun_and_sky(self._c, C.byref(sky)), "set_sun_and_sky")

    def set_film(self, w, h):
        self._check(lib().ash_set_film(self._c, w, h), "set_film")
        self.width, self.height = w, h

    def render(self, cam, pc, first_frame=0, n_frames=1, queue_paths=0, timing=False, depth_channel=-1,
               max_bounce_iters=0, count_traversal=False, blocks_per_cu=0, graphs=True, lds_stack=0,
               traversal=0) -> dict:
        opts = RenderOpts.make(queue_paths, blocks_per_cu, timing, depth_channel, max_bounce_iters, count_traversal,
                               graphs, lds_stack, traversal)
        st = Stats()
        self._check(lib().ash_render(self._c, C.byref(cam), C.byref(pc), first_frame, n_frames, C.byref(opts),
                                     C.byref(st)), "render")
        return st.as_dict()

    def render_noalloc(self, cam, pc, first_frame, n_frames, width, height, images) -> dict:
        """ash_render_noalloc into the device buffer `images` (an integer
        device address of 9 planes of width x height RGBA32F)."""
        st = Stats()
        self._check(lib().ash_render_noalloc(self._c, C.byref(cam), C.byref(pc), first_frame, n_frames, width, height,
                                             C.c_void_p(images), C.byref(st)), "render_noalloc")
        return st.as_dict()

    def output_device_ptr(self, image: int = 0) -> int:
        """Device address of output image `image` of the film (planes are
        contiguous, so image 0's address is the whole film's)."""
        p = C.c_void_p()
        self._check(lib().ash_output_device_ptr(self._c, image, C.byref(p)), "output_device_ptr")
        return int(p.value)

    def render_shots(self, cams, pcs, n_frames, queue_paths=0, timing=False, blocks_per_cu=0, graphs=True,
                     count_traversal=False, lds_stack=0, traversal=0, depth_channel=-1) -> dict:
        n = len(cams)
        ca = (GpuCamera * n)(*cams)
        pa = (GpuPushConstantRaytrace * n)(*pcs)
        opts = RenderOpts.make(queue_paths, blocks_per_cu, timing, depth_channel, 0, count_traversal, graphs, lds_stack,
                               traversal)
        st = Stats()
        self._check(lib().ash_render_shots(self._c, n, ca, pa, n_frames, C.byref(opts), C.byref(st)),
                    "render_shots")
        return st.as_dict()

    def autotune(self, cams, pcs, probe_frames=1, **opts) -> tuple:
        """Returns (RenderOpts with the fastest queue_paths / blocks_per_cu, report dict)."""
        n = len(cams)
        ca = (GpuCamera * n)(*cams)
        pa = (GpuPushConstantRaytrace * n)(*pcs)
        o = RenderOpts.make(**opts)
        res = TuneResult()
        self._check(lib().ash_autotune(self._c, n, ca, pa, probe_frames, C.byref(o), C.byref(res)), "autotune")
        return o, res.as_dict()

    def read_shot(self, shot, image) -> np.ndarray:
        out = np.zeros((self.height, self.width, 4), np.float32)
        self._check(lib().ash_read_shot_output(self._c, shot, image, fptr(out)), "read_shot_output")
        return out

    def read(self, image: int) -> np.ndarray:
        out = np.zeros((self.height, self.width, 4), np.float32)
        self._check(lib().ash_read_output(self._c, image, fptr(out)), "read_output")
        return out

    def first_hit(self, cam) -> np.ndarray:
        out = np.zeros(self.width * self.height, HIT_DTYPE)
        self._check(lib().ash_first_hit(self._c, C.byref(cam), out.ctypes.data_as(C.c_void_p)), "first_hit")
        return out.reshape(self.height, self.width)

    def env_tables(self, w, h):
        m = np.zeros((h, 2), np.float32)
        c = np.zeros((h, w, 2), np.float32)
        self._check(lib().ash_env_tables(self._c, fptr(m), fptr(c)), "env_tables")
        return m, c


def run_offline(scene: str, out: str, gpu_id: int = 0, output_scanline: bool = False, spp_override: int = 0,
                quiet: bool = True) -> dict:
    o = OfflineOpts(gpu_id, int(output_scanline), spp_override, int(quiet), RenderOpts.make())
    st = Stats()
    err = C.create_string_buffer(4096)
    rc = lib().ash_run_offline(scene.encode(), out.encode(), C.byref(o), C.byref(st), err, len(err))
    if rc != ASH_OK:
        raise AsunaError(f"run_offline failed ({ERRORS.get(rc, rc)}): {err.value.decode()}")
    return st.as_dict()


# ------------------------------------------------------------ render server
SERVE_OLDEST_GLOBAL, SERVE_ROUND_ROBIN, SERVE_OLDEST_FIRST = 0, 1, 2  # ash_server_opts.policy


class Server:
    """Resident render server (include/asuna_server.h): scenes stay in HBM,
    jobs are admitted by an HBM budget and co-batched per scene.  policy:
    SERVE_OLDEST_GLOBAL (default), SERVE_ROUND_ROBIN or SERVE_OLDEST_FIRST
    (batch fill order, and for the first the order in which scenes get the
    device)."""

    def __init__(self, device=0, hbm_budget_bytes=0, queue_paths=0, blocks_per_cu=0, max_jobs_per_batch=0,
                 batches_ahead=0, policy=0):
        self._s = C.c_void_p()
        o = ServerOpts(hbm_budget_bytes, queue_paths, blocks_per_cu, max_jobs_per_batch, batches_ahead, policy, 0)
        rc = lib().ash_server_create(device, C.byref(o), C.byref(self._s))
        if rc != ASH_OK:
            self._s = None
            raise AsunaError(f"ash_server_create failed: {ERRORS.get(rc, rc)}")
        self._sizes = {}
        self._scenes = []

    def _check(self, rc, what):
        if rc != ASH_OK:
            msg = lib().ash_server_last_error(self._s).decode()
            raise AsunaError(f"{what} failed ({ERRORS.get(rc, rc)}): {msg}")

    def add_scene(self, scene: "HostScene", leaf_max_tris=4, quality=2, branching=4) -> int:
        sid = C.c_uint32()
        self._check(lib().ash_server_add_scene(self._s, C.byref(scene.desc),
                                               C.byref(BuildOpts(leaf_max_tris, quality, branching, 0)),
                                               C.byref(sid)), "add_scene")
        self._scenes.append(scene)  # keeps the host arrays alive
        return sid.value

    def set_denoiser(self, denoiser):
        self._den = denoiser  # keeps it alive
        self._check(lib().ash_server_set_denoiser(self._s, denoiser._d), "set_denoiser")

    def submit(self, scene_id, cam, pc, width, height, spp, depth_channel=-1, denoise=False) -> int:
        d = JobDesc(scene_id, width, height, spp, cam, pc, depth_channel, 1 if denoise else 0)
        jid = C.c_uint64()
        self._check(lib().ash_server_submit(self._s, C.byref(d), C.byref(jid)), "submit")
        self._sizes[jid.value] = (height, width)
        return jid.value

    def wait(self, job_id, timeout_ms=-1) -> dict:
        st = JobStatus()
        self._check(lib().ash_server_wait(self._s, job_id, timeout_ms, C.byref(st)), "wait")
        return st.as_dict()

    def read(self, job_id, image=0) -> np.ndarray:
        h, w = self._sizes[job_id]
        out = np.zeros((h, w, 4), np.float32)
        self._check(lib().ash_server_read(self._s, job_id, image, fptr(out)), "read")
        return out

    def release(self, job_id):
        self._check(lib().ash_server_release(self._s, job_id), "release")
        self._sizes.pop(job_id, None)

    def stats(self) -> dict:
        c = ServerCounters()
        self._check(lib().ash_server_stats(self._s, C.byref(c)), "stats")
        return c.as_dict()

    def close(self):
        if self._s:
            lib().ash_server_destroy(self._s)
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------ denoiser
class Denoiser:
    """Learned denoiser (include/asuna_denoise.h): U-Net trained and run by the
    library's BF16 MFMA kernels.  Batches are NHWC float32: input (N, H, W, 10)
    = radiance RGB, albedo RGB, normal XYZ, depth; target (N, H, W, 3)."""

    def __init__(self, device=0, seed=1):
        self._d = C.c_void_p()
        rc = lib().ash_denoiser_create(device, seed, C.byref(self._d))
        if rc != ASH_OK:
            self._d = None
            raise AsunaError(f"ash_denoiser_create failed: {ERRORS.get(rc, rc)}")

    def _check(self, rc, what):
        if rc != ASH_OK:
            msg = lib().ash_denoiser_last_error(self._d).decode()
            raise AsunaError(f"{what} failed ({ERRORS.get(rc, rc)}): {msg}")

    def num_params(self) -> int:
        n = C.c_uint64()
        self._check(lib().ash_denoiser_num_params(self._d, C.byref(n)), "num_params")
        return n.value

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.num_params(), np.float32)
        self._check(lib().ash_denoiser_get_params(self._d, fptr(out)), "get_params")
        return out

    def set_params(self, theta):
        t = np.ascontiguousarray(theta, np.float32)
        assert t.size == self.num_params()
        self._check(lib().ash_denoiser_set_params(self._d, fptr(t)), "set_params")

    @staticmethod
    def _batch(x, c):
        x = np.ascontiguousarray(x, np.float32)
        assert x.ndim == 4 and x.shape[-1] == c
        return x

    def forward(self, x) -> np.ndarray:
        x = self._batch(x, 10)
        n, h, w, _ = x.shape
        out = np.zeros((n, h, w, 3), np.float32)
        self._check(lib().ash_denoiser_forward(self._d, n, h, w, fptr(x), fptr(out)), "forward")
        return out

    def gradient(self, x, target):
        x, t = self._batch(x, 10), self._batch(target, 3)
        n, h, w, _ = x.shape
        g = np.zeros(self.num_params(), np.float32)
        loss = C.c_float()
        self._check(lib().ash_denoiser_gradient(self._d, n, h, w, fptr(x), fptr(t), C.byref(loss), fptr(g)),
                    "gradient")
        return loss.value, g

    def train_step(self, x, target, lr=1e-3) -> float:
        x, t = self._batch(x, 10), self._batch(target, 3)
        n, h, w, _ = x.shape
        loss = C.c_float()
        self._check(lib().ash_denoiser_train_step(self._d, n, h, w, fptr(x), fptr(t), lr, C.byref(loss)),
                    "train_step")
        return loss.value

    def denoise_film(self, renderer, albedo_channel, normal_channel, depth_channel, out_image) -> float:
        ms = C.c_double()
        self._check(lib().ash_denoise_film(self._d, renderer._c, albedo_channel, normal_channel, depth_channel,
                                           out_image, C.byref(ms)), "denoise_film")
        return ms.value

    def close(self):
        if self._d:
            lib().ash_denoiser_destroy(self._d)
            self._d = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
