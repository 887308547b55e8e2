// Host C ABI (include/asuna_host.h): scene loading, shots, the offline render
// loop of Tracer::runOffline (src/tracer/tracer.cpp:177-291) and image IO.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "json.h"
#include "scene.h"

namespace {

void setErr(char* err, size_t len, const std::string& m) {
  if (err && len) {
    std::snprintf(err, len, "%s", m.c_str());
  }
}

// Scene::setShot (scene.cpp:439-453): per-shot state fields override the
// pipeline state; the offline loop then forces spp = 1 per dispatch.
void shotState(const ash_host_scene& s, uint32_t shot, GpuCamera& cam, GpuPushConstantRaytrace& pc) {
  const ash::Shot& sh = s.shots[shot];
  pc = s.state.rtx;
  pc.spp = sh.state.rtx.spp;
  pc.maxPathDepth = sh.state.rtx.maxPathDepth;
  pc.useFaceNormal = sh.state.rtx.useFaceNormal;
  pc.ignoreEmissive = sh.state.rtx.ignoreEmissive;
  pc.envMapIntensity = sh.state.rtx.envMapIntensity;
  std::memcpy(pc.bgColor, sh.state.rtx.bgColor, sizeof pc.bgColor);
  pc.curFrame = 0;
  ash::CameraModel cm = s.camera;
  if (sh.fov > 0.f) cm.fov = sh.fov;  // per-shot fov extension
  cam = ash::cameraToGpu(cm, sh.eye, sh.lookat, sh.up, sh.envTransform);
}

}  // namespace

extern "C" {

int ash_scene_load(const char* path, ash_host_scene** out, char* err, size_t errLen) {
  if (!path || !out) return ASH_E_INVALID;
  *out = nullptr;
  ash_host_scene* s = new ash_host_scene;
  try {
    ash::loadSceneFile(path, *s);
  } catch (const ash::JsonError& e) {
    setErr(err, errLen, e.what());
    delete s;
    return ASH_E_PARSE;
  } catch (const std::exception& e) {
    setErr(err, errLen, e.what());
    delete s;
    std::string m = e.what();
    if (m.find("failed to load") != std::string::npos || m.find("cannot open") != std::string::npos ||
        m.find("failed to read") != std::string::npos)
      return ASH_E_IO;
    return ASH_E_PARSE;
  }
  *out = s;
  return ASH_OK;
}

void ash_scene_free(ash_host_scene* s) { delete s; }

int ash_scene_get_desc(const ash_host_scene* s, ash_scene_desc* d) {
  if (!s || !d) return ASH_E_INVALID;
  std::memset(d, 0, sizeof *d);
  d->num_meshes = uint32_t(s->meshDescs.size());
  d->meshes = s->meshDescs.data();
  d->num_instances = uint32_t(s->instDescs.size());
  d->instances = s->instDescs.data();
  d->num_materials = uint32_t(s->materials.size());
  d->materials = s->materials.data();
  d->num_lights = uint32_t(s->lights.size());
  d->lights = s->lights.data();
  d->num_textures = uint32_t(s->texDescs.size());
  d->textures = s->texDescs.data();
  if (s->hasEnv) {
    d->envmap.width = s->env.w;
    d->envmap.height = s->env.h;
    d->envmap.rgba = s->env.rgba.data();
  }
  return ASH_OK;
}

int ash_scene_get_state(const ash_host_scene* s, ash_scene_state* o) {
  if (!s || !o) return ASH_E_INVALID;
  std::memset(o, 0, sizeof *o);
  o->film_width = s->camera.width;
  o->film_height = s->camera.height;
  o->num_shots = uint32_t(s->shots.size());
  o->tone_mapping = s->state.tmType;
  o->output_hdr = s->state.outputHdr;
  o->output_render_result = s->state.outputRenderResult;
  for (size_t c = 0; c < s->state.channelOutputLdr.size(); ++c)
    if (s->state.channelOutputLdr[c]) o->channel_output_ldr |= 1u << c;
  o->depth_out_channel = s->state.depthOutChannel;
  for (const auto& in : s->instances) o->num_triangles += s->meshes[in.mesh].indices.size() / 3;
  return ASH_OK;
}

int ash_scene_shot(const ash_host_scene* s, uint32_t shot, GpuCamera* cam, GpuPushConstantRaytrace* pc) {
  if (!s || !cam || !pc || shot >= s->shots.size()) return ASH_E_INVALID;
  shotState(*s, shot, *cam, *pc);
  return ASH_OK;
}

void ash_camera_lookat(uint32_t w, uint32_t h, float fov, const float eye[3], const float lookat[3], const float up[3],
                       float aperture, float focal, const float env[16], GpuCamera* out) {
  ash::CameraModel cm;
  cm.type = CameraTypePerspective;
  cm.width = w;
  cm.height = h;
  cm.fov = fov;
  cm.aperture = aperture;
  cm.focal = focal;
  ash::Mat4 e;
  if (env) std::memcpy(e.m, env, 64);
  *out = ash::cameraToGpu(cm, ash::Vec3f(eye[0], eye[1], eye[2]), ash::Vec3f(lookat[0], lookat[1], lookat[2]),
                          ash::Vec3f(up[0], up[1], up[2]), e);
}

int ash_write_image(const char* path, uint32_t w, uint32_t h, const float* rgba) {
  std::string err;
  return ash::writeImage(path, w, h, rgba, err) ? ASH_OK : ASH_E_IO;
}

int ash_read_image(const char* path, float gamma, uint32_t* w, uint32_t* h, float** rgba) {
  ash::Image im;
  std::string err;
  if (!ash::readImage(path, gamma, im, err)) return ASH_E_IO;
  *w = im.w;
  *h = im.h;
  *rgba = new float[im.rgba.size()];
  std::memcpy(*rgba, im.rgba.data(), im.rgba.size() * 4);
  return ASH_OK;
}

void ash_free_image(float* p) { delete[] p; }

void ash_tonemap(uint32_t type, uint32_t w, uint32_t h, float* rgba) { ash::toneMap(type, w, h, rgba); }

void ash_env_tables_host(uint32_t w, uint32_t h, const float* rgba, float* marg, float* cond) {
  std::vector<float> m, c;
  ash::buildEnvTables(w, h, rgba, m, c);
  std::memcpy(marg, m.data(), m.size() * sizeof(float));
  std::memcpy(cond, c.data(), c.size() * sizeof(float));
}

int ash_run_offline(const char* scenePath, const char* outName, const ash_offline_opts* o, ash_stats* total,
                    char* err, size_t errLen) {
  ash_host_scene* s = nullptr;
  int rc = ash_scene_load(scenePath, &s, err, errLen);
  if (rc) return rc;
  ash_offline_opts opts{};
  if (o) opts = *o;
  ash_ctx* ctx = nullptr;
  rc = ash_create(opts.gpu_id, &ctx);
  if (rc) {
    setErr(err, errLen, "cannot open GPU " + std::to_string(opts.gpu_id));
    ash_scene_free(s);
    return rc;
  }
  auto bail = [&](int code) {
    setErr(err, errLen, ash_last_error(ctx));
    ash_destroy(ctx);
    ash_scene_free(s);
    return code;
  };
  ash_scene_desc d;
  ash_scene_get_desc(s, &d);
  if ((rc = ash_upload_scene(ctx, &d))) return bail(rc);
  ash_build_info bi;
  ash_build_opts bo{4, 2, 4, 0};
  if ((rc = ash_build_accel(ctx, &bo, &bi))) return bail(rc);
  uint32_t W = s->camera.width, H = s->camera.height;
  std::vector<float> px(size_t(W) * H * 4);
  if (total) std::memset(total, 0, sizeof *total);
  std::string out = outName && *outName ? outName : "asuna_out.hdr";
  ash_render_opts ro = opts.render;
  ro.depthOutChannel = s->state.depthOutChannel;
  for (uint32_t shot = 0; shot < s->shots.size(); ++shot) {
    GpuCamera cam;
    GpuPushConstantRaytrace pc;
    shotState(*s, shot, cam, pc);
    int spp = opts.spp_override > 0 ? opts.spp_override : pc.spp;
    if ((rc = ash_set_film(ctx, W, H))) return bail(rc);
    ash_stats st{};
    if (spp > 0 && (rc = ash_render(ctx, &cam, &pc, 0, spp, &ro, &st))) return bail(rc);
    if (total) {
      total->paths += st.paths;
      total->extension_rays += st.extension_rays;
      total->shadow_rays += st.shadow_rays;
      total->dielectric_shadow_rays += st.dielectric_shadow_rays;
      total->render_ms += st.render_ms;
    }
    if (!opts.quiet)
      std::fprintf(stderr, "[asuna] shot %u: %d spp, %.1f ms, %.2f Mpaths/s\n", shot, spp, st.render_ms,
                   st.render_ms > 0 ? double(st.paths) / st.render_ms / 1e3 : 0.0);
    // callSavingImage, tracer.cpp:266-291
    const ash::State& state = s->state;
    if (!state.outputRenderResult) continue;
    char name[4096];
    std::string e;
    if ((rc = ash_read_output(ctx, 0, px.data()))) return bail(rc);
    if (state.outputHdr) {
      std::snprintf(name, sizeof name, "%s_shot_%04u.exr", out.c_str(), shot);
      if (!ash::writeImage(name, W, H, px.data(), e)) return bail(ASH_E_IO);
    } else {
      std::snprintf(name, sizeof name, "%s_shot_%04u.png", out.c_str(), shot);
      ash::toneMap(state.tmType, W, H, px.data());
      if (!ash::writeImage(name, W, H, px.data(), e)) return bail(ASH_E_IO);
    }
    std::vector<int> valid;  // tracer.cpp:344-392 (--output_scanline)
    for (uint32_t cid = 0; cid < pc.nMultiChannel; ++cid) {
      bool ldr = cid < state.channelOutputLdr.size() && state.channelOutputLdr[cid];
      std::snprintf(name, sizeof name, "%s_shot_%04u_channel_%04u.%s", out.c_str(), shot, cid, ldr ? "png" : "exr");
      if ((rc = ash_read_output(ctx, int32_t(cid) + 1, px.data()))) return bail(rc);
      if (!opts.output_scanline) {
        if (!ash::writeImage(name, W, H, px.data(), e)) return bail(ASH_E_IO);
        continue;
      }
      if (cid == 0) {
        valid.clear();
        for (uint32_t i = 0; i < W * H; ++i)
          if (px[size_t(i) * 4 + 2] == 1.0f) {
            valid.push_back(int(i));
            px[size_t(i) * 4 + 2] = float(i);
          }
      }
      std::vector<float> vd;
      vd.reserve(valid.size() * 3);
      for (int i : valid)
        for (int k = 0; k < 3; ++k) vd.push_back(px[size_t(i) * 4 + k]);
      if (!ash::writeNpy(name, {valid.size(), 3}, vd.data(), e)) return bail(ASH_E_IO);
    }
  }
  ash_destroy(ctx);
  ash_scene_free(s);
  return ASH_OK;
}

}  // extern "C"
