// Host scene model: what the reference's Loader fills into Scene
// (src/loader/loader.cpp, src/scene/scene.cpp) before Scene::submit.
#pragma once

#include <map>
#include <string>
#include <vector>

#include "../../../include/asuna_host.h"
#include "hmath.h"

namespace ash {

// State (src/core/state.h:6-63): the push constants plus output options.
struct State {
  GpuPushConstantRaytrace rtx;
  uint32_t tmType = 4;  // ToneMappingTypeFilmic
  bool outputHdr = false;
  bool outputRenderResult = true;
  std::vector<bool> channelOutputLdr;
  int32_t depthOutChannel = -1;  // extension channel "depth"
  State();
};

struct Shot {  // CameraShot, src/core/camera.h:8-15
  Vec3f lookat, eye, up;
  Mat4 envTransform;
  State state;
  float fov = -1.f;  // extension: per-shot horizontal fov (degrees), <= 0 = camera fov
};

struct CameraModel {  // Camera / CameraPerspective / CameraOpencv
  uint32_t type = CameraTypeUndefined;
  uint32_t width = 0, height = 0;
  float fov = 45.f, aperture = 0.f, focal = 0.1f;
  float fxfycxcy[4] = {0, 0, 0, 0};
};

struct Image {
  uint32_t w = 1, h = 1;
  std::vector<float> rgba = std::vector<float>(4, 0.f);
};

struct MeshData {
  std::vector<GpuVertex> vertices;
  std::vector<uint32_t> indices;
  Vec3f posMin{1e30f, 1e30f, 1e30f}, posMax{-1e30f, -1e30f, -1e30f};
};

struct InstanceRec {
  Mat4 transform;
  uint32_t mesh = 0, material = 0;
  int32_t light = -1;
};

}  // namespace ash

struct ash_host_scene {
  std::string dir;
  ash::State state;
  ash::CameraModel camera;
  std::vector<ash::Image> textures;
  std::map<std::string, uint32_t> textureIds;
  std::vector<GpuMaterial> materials;
  std::map<std::string, uint32_t> materialIds;
  std::vector<GpuLight> lights;
  std::vector<ash::MeshData> meshes;
  std::map<std::string, uint32_t> meshIds;
  std::vector<ash::InstanceRec> instances;
  bool hasEnv = false;
  ash::Image env;
  std::vector<ash::Shot> shots;
  // Flat views handed to ash_upload_scene.
  std::vector<ash_mesh_desc> meshDescs;
  std::vector<ash_instance_desc> instDescs;
  std::vector<ash_image_desc> texDescs;
  void buildDescs();
};

namespace ash {
// Loader::parse + Scene::submit host part; throws std::runtime_error.
void loadSceneFile(const std::string& path, ash_host_scene& s);
// OBJ loader with loadMesh semantics (src/core/mesh.cpp:112-145).
void loadObj(const std::string& path, std::vector<GpuVertex>& v, std::vector<uint32_t>& idx);
// Camera UBO for the camera model at a look-at pose.
GpuCamera cameraToGpu(const CameraModel& cam, Vec3f eye, Vec3f lookat, Vec3f up, const Mat4& env);
// Env-map importance tables (src/core/texture.cpp:144-226): marginal H x 2,
// conditional W*H x 2 (x = inverse-CDF position, y = pdf).
void buildEnvTables(uint32_t w, uint32_t h, const float* rgba, std::vector<float>& marg,
                    std::vector<float>& cond);
// Image IO (src/core/texture.cpp:16-75, 306-372).
bool readImage(const std::string& path, float gamma, Image& out, std::string& err);
// Baseline / progressive Huffman JPEG (jpeg.cpp); channels (v/255)^gamma, alpha 1.
bool readJpeg(const std::vector<unsigned char>& bytes, float gamma, Image& out, std::string& err);
bool writeImage(const std::string& path, uint32_t w, uint32_t h, const float* rgba, std::string& err);
bool writeNpy(const std::string& path, const std::vector<unsigned long>& shape, const float* data,
              std::string& err);
void toneMap(uint32_t type, uint32_t width, uint32_t height, float* rgba);
}  // namespace ash
