// Scene-file loading: the reference's Loader (src/loader/loader.cpp,
// src/loader/material.cpp) and the host half of Scene (src/scene/scene.cpp).
// Errors throw std::runtime_error with the reference's message wording; the C
// ABI turns them into return codes (the reference exits instead).
#include "scene.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "json.h"

namespace ash {

static const float kPi = 3.14159265358979323846f;
static const float kToRad = kPi / 180.f;

State::State() {
  // src/core/state.h:15-62
  std::memset(&rtx, 0, sizeof rtx);
  rtx.curFrame = -1;
  rtx.spp = 1;
  rtx.maxPathDepth = 3;
  rtx.numLights = 0;
  rtx.useFaceNormal = 0;
  rtx.ignoreEmissive = 0;
  rtx.hasEnvMap = 0;
  rtx.envMapIntensity = 1.f;
  rtx.nMultiChannel = 0;
  rtx.diffuseOutChannel = -1;
  rtx.specularOutChannel = -1;
  rtx.roughnessOutChannel = -1;
  rtx.normalOutChannel = -1;
  rtx.positionOutChannel = -1;
  rtx.tangentOutChannel = -1;
  rtx.uvOutChannel = -1;
}

namespace {

[[noreturn]] void fail(const std::string& m) { throw std::runtime_error(m); }

void checkKeys(const Json& j, std::initializer_list<const char*> keys) {
  // JsonCheckKeys, loader/utils.h:58-67
  for (const char* k : keys)
    if (!j.contains(k)) fail(std::string("Json: missing key [\"") + k + "\"] in " + j.dump());
}
Vec3f vec3(const Json& j) {
  if (j.size() < 3) fail("Json: failed to extract vector3f from " + j.dump());
  return {j[0].f(), j[1].f(), j[2].f()};
}
void vec2(const Json& j, float& a, float& b) {
  if (j.size() < 2) fail("Json: failed to extract vector2f from " + j.dump());
  a = j[0].f();
  b = j[1].f();
}
Mat4 mat4RowMajor(const Json& j) {  // Json2Mat4: JSON is row-major
  if (j.size() < 16) fail("Json: failed to extract mat4f from " + j.dump());
  Mat4 m;
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) m.at(r, c) = j[size_t(r * 4 + c)].f();
  return m;
}
void set3(float* d, Vec3f v) { d[0] = v.x, d[1] = v.y, d[2] = v.z; }

bool isAbsolute(const std::string& p) { return !p.empty() && p[0] == '/'; }
bool fileExists(const std::string& p) {
  std::ifstream f(p);
  return f.good();
}
std::string findFile(const std::string& p, const std::string& dir) {
  // nvh::findFile(path, {sceneDir}): absolute paths as given, else relative
  // to the scene directory, then to the working directory.
  if (isAbsolute(p)) return fileExists(p) ? p : std::string();
  std::string c = dir.empty() ? p : dir + "/" + p;
  if (fileExists(c)) return c;
  if (fileExists(p)) return p;
  return std::string();
}

// loader/material.cpp:6-37
float dielectricReflectance(float eta, float cosThetaI, float& cosThetaT) {
  if (cosThetaI < 0.0f) {
    eta = 1.0f / eta;
    cosThetaI = -cosThetaI;
  }
  float sinThetaTSq = eta * eta * (1.0f - cosThetaI * cosThetaI);
  if (sinThetaTSq > 1.0f) {
    cosThetaT = 0.0f;
    return 1.0f;
  }
  cosThetaT = std::sqrt(std::max(1.0f - sinThetaTSq, 0.0f));
  float Rs = (eta * cosThetaI - cosThetaT) / (eta * cosThetaI + cosThetaT);
  float Rp = (eta * cosThetaT - cosThetaI) / (eta * cosThetaT + cosThetaI);
  return (Rs * Rs + Rp * Rp) * 0.5f;
}
float diffuseFresnel(float ior, int n) {
  double acc = 0.0;
  float ct;
  float fb = dielectricReflectance(ior, 0.0f, ct);
  for (int i = 1; i <= n; ++i) {
    float c2 = float(i) / float(n);
    float fa = dielectricReflectance(ior, std::min(std::sqrt(c2), 1.0f), ct);
    acc += double(fa + fb) * (0.5 / n);
    fb = fa;
  }
  return float(acc);
}

// parseState, loader.cpp:148-234
void parseState(const Json& sj, State& st) {
  if (sj.contains("path_tracing")) {
    const Json& pt = sj["path_tracing"];
    auto& r = st.rtx;
    if (pt.contains("spp")) r.spp = pt["spp"].i();
    if (pt.contains("max_path_depth")) r.maxPathDepth = pt["max_path_depth"].i();
    if (pt.contains("use_face_normal")) r.useFaceNormal = pt["use_face_normal"].b() ? 1 : 0;
    if (pt.contains("ignore_emissive")) r.ignoreEmissive = pt["ignore_emissive"].b() ? 1 : 0;
    if (pt.contains("background_color")) set3(r.bgColor, vec3(pt["background_color"]));
    if (pt.contains("envmap_intensity")) r.envMapIntensity = pt["envmap_intensity"].f();
    if (pt.contains("multi_channel")) {
      const Json& mc = pt["multi_channel"];
      uint32_t n = uint32_t(mc.size());
      if (n > ASH_NUM_OUTPUT_IMAGES - 1) fail("Loader: channel numbers can not exceed [8]");
      r.nMultiChannel = n;
      for (uint32_t cid = 0; cid < n; ++cid) {
        const std::string& t = mc[cid].str();
        if (t == "diffuse") r.diffuseOutChannel = int32_t(cid);
        if (t == "normal") r.normalOutChannel = int32_t(cid);
        if (t == "specular") r.specularOutChannel = int32_t(cid);
        if (t == "tangent") r.tangentOutChannel = int32_t(cid);
        if (t == "roughness") r.roughnessOutChannel = int32_t(cid);
        if (t == "position") r.positionOutChannel = int32_t(cid);
        if (t == "uv") r.uvOutChannel = int32_t(cid);
        if (t == "depth") st.depthOutChannel = int32_t(cid);  // extension
      }
    }
    st.channelOutputLdr.assign(r.nMultiChannel, false);
    if (pt.contains("multi_channel_ldr")) {
      const Json& ml = pt["multi_channel_ldr"];
      for (uint32_t cid = 0; cid < std::min<uint32_t>(r.nMultiChannel, uint32_t(ml.size())); ++cid)
        st.channelOutputLdr[cid] = ml[cid].b();
    }
  }
  if (sj.contains("post_processing")) {
    const Json& pp = sj["post_processing"];
    std::string s = pp.contains("tone_mapping") ? pp["tone_mapping"].str() : std::string();
    uint32_t tm = 0;  // unknown names fall back to none with a warning
    if (s == "none") tm = 0;
    else if (s == "gamma") tm = 1;
    else if (s == "reinhard") tm = 2;
    else if (s == "Aces") tm = 3;
    else if (s == "filmic") tm = 4;
    else if (s == "pbrt") tm = 5;
    else if (s == "custom") tm = 6;
    else std::fprintf(stderr, "[asuna] Loader: no matching tone mapper for [%s], use default\n", s.c_str());
    st.tmType = tm;
  }
  if (sj.contains("output_render_result")) st.outputRenderResult = sj["output_render_result"].b();
  if (sj.contains("output_hdr")) st.outputHdr = sj["output_hdr"].b();
}

// parseToWorld, loader.cpp:366-396: each op left-multiplies.
Mat4 parseToWorld(const Json& tj, bool banTranslation) {
  Mat4 xf;
  for (const Json& op : tj.arr()) {
    checkKeys(op, {"type", "value"});
    const std::string& t = op["type"].str();
    Mat4 m;
    if (t == "matrix") {
      m = mat4RowMajor(op["value"]);
    } else if (t == "translate" && !banTranslation) {
      Vec3f v = vec3(op["value"]);
      m = translateMat(v.x, v.y, v.z);
    } else if (t == "translate") {
      // banned for env_toworld: identity
    } else if (t == "scale") {
      Vec3f v = vec3(op["value"]);
      m = scaleMat(v.x, v.y, v.z);
    } else if (t == "rotx") {
      m = rotX(kToRad * op["value"].f());
    } else if (t == "roty") {
      m = rotY(kToRad * op["value"].f());
    } else if (t == "rotz") {
      m = rotZ(kToRad * op["value"].f());
    } else if (t == "rotate") {
      Vec3f v = vec3(op["value"]);
      m = mul(mul(rotZ(kToRad * v.z), rotY(kToRad * v.y)), rotX(kToRad * v.x));
    } else {
      fail("Loader: unrecognized toworld singleton type [" + t + "]");
    }
    xf = mul(m, xf);
  }
  return xf;
}

// Spectral complex indices of refraction (eta, k) at the RGB primaries that
// brdf_conductor / brdf_rough_conductor look up by name
// (loader/material.cpp:245-290; the measured values are data and are kept).
struct ComplexIor {
  const char* name;
  float eta[3], k[3];
};
const ComplexIor kComplexIor[] = {
    {"a-C", {2.9440999183f, 2.2271502925f, 1.9681668794f}, {0.8874329109f, 0.7993216383f, 0.8152862927f}},
    {"Ag", {0.1552646489f, 0.1167232965f, 0.1383806959f}, {4.8283433224f, 3.1222459278f, 2.1469504455f}},
    {"Al", {1.6574599595f, 0.8803689579f, 0.5212287346f}, {9.2238691996f, 6.2695232477f, 4.8370012281f}},
    {"AlAs", {3.6051023902f, 3.2329365777f, 2.2175611545f}, {0.0006670247f, -0.0004999400f, 0.0074261204f}},
    {"AlSb", {-0.0485225705f, 4.1427547893f, 4.6697691348f}, {-0.0363741915f, 0.0937665154f, 1.3007390124f}},
    {"Au", {0.1431189557f, 0.3749570432f, 1.4424785571f}, {3.9831604247f, 2.3857207478f, 1.6032152899f}},
    {"Be", {4.1850592788f, 3.1850604423f, 2.7840913457f}, {3.8354398268f, 3.0101260162f, 2.8690088743f}},
    {"Cr", {4.3696828663f, 2.9167024892f, 1.6547005413f}, {5.2064337956f, 4.2313645277f, 3.7549467933f}},
    {"CsI", {2.1449030413f, 1.7023164587f, 1.6624194173f}, {0.0f, 0.0f, 0.0f}},
    {"Cu", {0.2004376970f, 0.9240334304f, 1.1022119527f}, {3.9129485033f, 2.4528477015f, 2.1421879552f}},
    {"Cu2O", {3.5492833755f, 2.9520622449f, 2.7369202137f}, {0.1132179294f, 0.1946659670f, 0.6001681264f}},
    {"CuO", {3.2453822204f, 2.4496293965f, 2.1974114493f}, {0.5202739621f, 0.5707372756f, 0.7172250613f}},
    {"d-C", {2.7112524747f, 2.3185812849f, 2.2288565009f}, {0.0f, 0.0f, 0.0f}},
    {"Hg", {2.3989314904f, 1.4400254917f, 0.9095512090f}, {6.3276269444f, 4.3719414152f, 3.4217899270f}},
    {"HgTe", {4.7795267752f, 3.2309984581f, 2.6600252401f}, {1.6319827058f, 1.5808189339f, 1.7295753852f}},
    {"Ir", {3.0864098394f, 2.0821938440f, 1.6178866805f}, {5.5921510077f, 4.0671757150f, 3.2672611269f}},
    {"K", {0.0640493070f, 0.0464100621f, 0.0381842017f}, {2.1042155920f, 1.3489364357f, 0.9132113889f}},
    {"Li", {0.2657871942f, 0.1956102432f, 0.2209198538f}, {3.5401743407f, 2.3111306542f, 1.6685930000f}},
    {"MgO", {2.0895885542f, 1.6507224525f, 1.5948759692f}, {0.0f, -0.0f, 0.0f}},
    {"Mo", {4.4837010280f, 3.5254578255f, 2.7760769438f}, {4.1111307988f, 3.4208716252f, 3.1506031404f}},
    {"Na", {0.0602665320f, 0.0561412435f, 0.0619909494f}, {3.1792906496f, 2.1124800781f, 1.5790940266f}},
    {"Nb", {3.4201353595f, 2.7901921379f, 2.3955856658f}, {3.4413817900f, 2.7376437930f, 2.5799132708f}},
    {"Ni", {2.3672753521f, 1.6633583302f, 1.4670554172f}, {4.4988329911f, 3.0501643957f, 2.3454274399f}},
    {"Rh", {2.5857954933f, 1.8601866068f, 1.5544279524f}, {6.7822927110f, 4.7029501026f, 3.9760892461f}},
    {"Se-e", {5.7242724833f, 4.1653992967f, 4.0816099264f}, {0.8713747439f, 1.1052845009f, 1.5647788766f}},
    {"Se", {4.0592611085f, 2.8426947380f, 2.8207582835f}, {0.7543791750f, 0.6385150558f, 0.5215872029f}},
    {"SiC", {3.1723450205f, 2.5259677964f, 2.4793623897f}, {0.0000007284f, -0.0000006859f, 0.0000100150f}},
    {"SnTe", {4.5251865890f, 1.9811525984f, 1.2816819226f}, {0.0f, 0.0f, 0.0f}},
    {"Ta", {2.0625846607f, 2.3930915569f, 2.6280684948f}, {2.4080467973f, 1.7413705864f, 1.9470377016f}},
    {"Te-e", {7.5090397678f, 4.2964603080f, 2.3698732430f}, {5.5842076830f, 4.9476231084f, 3.9975145063f}},
    {"Te", {7.3908396088f, 4.4821028985f, 2.6370708478f}, {3.2561412892f, 3.5273908133f, 3.2921683116f}},
    {"ThF4", {1.8307187117f, 1.4422274283f, 1.3876488528f}, {0.0f, 0.0f, 0.0f}},
    {"TiC", {3.7004673762f, 2.8374356509f, 2.5823030278f}, {3.2656905818f, 2.3515586388f, 2.1727857800f}},
    {"TiN", {1.6484691607f, 1.1504482522f, 1.3797795097f}, {3.3684596226f, 1.9434888540f, 1.1020123347f}},
    {"TiO2-e", {3.1065574823f, 2.5131551146f, 2.5823844157f}, {0.0000289537f, -0.0000251484f, 0.0001775555f}},
    {"TiO2", {3.4566203131f, 2.8017076558f, 2.9051485020f}, {0.0001026662f, -0.0000897534f, 0.0006356902f}},
    {"VC", {3.6575665991f, 2.7527298065f, 2.5326814570f}, {3.0683516659f, 2.1986687713f, 1.9631816252f}},
    {"VN", {2.8656011588f, 2.1191817791f, 1.9400767149f}, {3.0323264950f, 2.0561075580f, 1.6162930914f}},
    {"V", {4.2775126218f, 3.5131538236f, 2.7611257461f}, {3.4911844504f, 2.8893580874f, 3.1116965117f}},
    {"W", {4.3707029924f, 3.3002972445f, 2.9982666528f}, {3.5006778591f, 2.6048652781f, 2.2731930614f}},
};

GpuMaterial defaultMaterial() {
  // Material::Material, src/core/material.h:9-35
  GpuMaterial m;
  std::memset(&m, 0, sizeof m);
  m.ior = 1.5f;
  m.roughness = 0.5f;
  m.radianceFactor[0] = m.radianceFactor[1] = m.radianceFactor[2] = 1.f;
  m.diffuseTextureId = m.radianceTextureId = m.metalnessTextureId = m.normalTextureId = -1;
  m.roughnessTextureId = m.tangentTextureId = m.opacityTextureId = -1;
  m.type = MaterialTypeBrdfLambertian;
  return m;
}

void updateAabb(MeshData& md, const GpuVertex& v) {
  md.posMin.x = std::min(md.posMin.x, v.pos[0]);
  md.posMin.y = std::min(md.posMin.y, v.pos[1]);
  md.posMin.z = std::min(md.posMin.z, v.pos[2]);
  md.posMax.x = std::max(md.posMax.x, v.pos[0]);
  md.posMax.y = std::max(md.posMax.y, v.pos[1]);
  md.posMax.z = std::max(md.posMax.z, v.pos[2]);
}

class Loader {
 public:
  explicit Loader(ash_host_scene& s) : s_(s) {}

  void load(const std::string& path) {
    std::ifstream f(path);
    if (!f) fail("Loader: failed to load scene from file [" + path + "]");
    std::stringstream ss;
    ss << f.rdbuf();
    size_t slash = path.find_last_of('/');
    s_.dir = slash == std::string::npos ? std::string(".") : path.substr(0, slash);
    Json j = Json::parse(ss.str());
    reset();
    parse(j);
    submit();
  }

 private:
  ash_host_scene& s_;

  void reset() {
    // Scene::reset, scene.cpp:87-130: dummy texture 0, dummy material 0,
    // dummy directional light 0.
    s_.textures.assign(1, Image());
    s_.textureIds.clear();
    s_.textureIds["add_by_default_dummy_texture"] = 0;
    s_.materials.assign(1, defaultMaterial());
    s_.materialIds.clear();
    s_.materialIds["add_by_default_dummy_material"] = 0;
    s_.lights.clear();
    GpuLight dummy;
    std::memset(&dummy, 0, sizeof dummy);
    dummy.type = LightTypeDirectional;
    dummy.doubleSide = 1;
    addLight(dummy);
  }

  void parse(const Json& j) {
    checkKeys(j, {"state", "camera", "meshes", "instances"});
    addState(j["state"]);
    addCamera(j["camera"]);
    if (j.contains("textures"))
      for (const Json& t : j["textures"].arr()) addTexture(t);
    if (j.contains("materials"))
      for (const Json& m : j["materials"].arr()) addMaterial(m);
    if (j.contains("lights"))
      for (const Json& l : j["lights"].arr()) addLightJson(l);
    if (j.contains("envmap")) addEnvMap(j["envmap"]);
    for (const Json& m : j["meshes"].arr()) addMesh(m);
    for (const Json& i : j["instances"].arr()) addInstance(i);
    if (j.contains("shots"))
      for (const Json& sh : j["shots"].arr()) addShot(sh);
  }

  void addState(const Json& sj) {
    State st;
    parseState(sj, st);
    s_.state = st;
  }

  void addCamera(const Json& cj) {
    // loader.cpp:242-268
    checkKeys(cj, {"type", "film"});
    const Json& film = cj["film"];
    checkKeys(film, {"resolution"});
    float rw, rh;
    vec2(film["resolution"], rw, rh);
    CameraModel c;
    c.width = uint32_t(rw);
    c.height = uint32_t(rh);
    const std::string& t = cj["type"].str();
    if (t == "perspective") {
      c.type = CameraTypePerspective;
      if (cj.contains("fov")) c.fov = cj["fov"].f();
      if (cj.contains("aperture")) c.aperture = cj["aperture"].f();
      if (cj.contains("focal_distance")) c.focal = cj["focal_distance"].f();
    } else if (t == "opencv") {
      checkKeys(cj, {"fx", "fy", "cx", "cy"});
      c.type = CameraTypeOpencv;
      c.fxfycxcy[0] = cj["fx"].f();
      c.fxfycxcy[1] = cj["fy"].f();
      c.fxfycxcy[2] = cj["cx"].f();
      c.fxfycxcy[3] = cj["cy"].f();
    } else {
      fail("Loader: unrecognized camera type [" + t + "]");
    }
    if (c.width == 0 || c.height == 0) fail("Loader: film resolution must be positive");
    s_.camera = c;
  }

  void addTexture(const Json& tj) {
    checkKeys(tj, {"name", "path"});
    std::string name = tj["name"].str();
    std::string p = findFile(tj["path"].str(), s_.dir);
    if (p.empty()) fail("Loader: failed to load texture from file [" + tj["path"].str() + "]");
    float gamma = tj.contains("gamma") ? tj["gamma"].f() : 1.0f;
    if (s_.textureIds.count(name)) fail("Scene: texture [\"" + name + "\"] defined twice");
    Image im;
    std::string err;
    if (!readImage(p, gamma, im, err)) fail(err);
    s_.textureIds[name] = uint32_t(s_.textures.size());
    s_.textures.push_back(std::move(im));
  }

  int32_t textureId(const Json& v) {
    const std::string& n = v.str();
    auto it = s_.textureIds.find(n);
    if (it == s_.textureIds.end()) fail("Scene: texture [\"" + n + "\"] does not exist");
    return int32_t(it->second);
  }

  void addMaterial(const Json& mj) {
    // Loader::addMaterial, loader/material.cpp:54-231 (in-scope types plus
    // the parameter parsing of the others so their scenes still load).
    checkKeys(mj, {"type", "name"});
    std::string name = mj["name"].str();
    GpuMaterial m = defaultMaterial();
    const std::string& t = mj["type"].str();
    auto opt3 = [&](const char* k, float* d) {
      if (mj.contains(k)) set3(d, vec3(mj[k]));
    };
    auto optTex = [&](const char* k, int32_t& id) {
      if (mj.contains(k)) id = textureId(mj[k]);
    };
    if (t == "brdf_lambertian") {
      m.type = MaterialTypeBrdfLambertian;
      opt3("diffuse_reflectance", m.diffuse);
      optTex("diffuse_texture", m.diffuseTextureId);
      optTex("normal_texture", m.normalTextureId);
    } else if (t == "brdf_pbr_metalness_roughness") {
      m.type = MaterialTypeBrdfPbrMetalnessRoughness;
      optTex("normal_texture", m.normalTextureId);
      opt3("diffuse_reflectance", m.diffuse);
      optTex("diffuse_texture", m.diffuseTextureId);
      if (mj.contains("metalness")) m.metalness = mj["metalness"].f();
      optTex("metalness_texture", m.metalnessTextureId);
      if (mj.contains("roughness")) m.roughness = mj["roughness"].f();
      optTex("roughness_texture", m.roughnessTextureId);
      m.specular = 0.f;  // used as opacity (material.cpp:137-138)
      optTex("opacity_texture", m.opacityTextureId);
    } else if (t == "brdf_emissive") {
      m.type = MaterialTypeBrdfEmissive;
      opt3("radiance", m.radiance);
      opt3("radiance_factor", m.radianceFactor);
      optTex("radiance_texture", m.radianceTextureId);
    } else if (t == "bsdf_dielectric") {
      m.type = MaterialTypeBsdfDielectric;
      optTex("normal_texture", m.normalTextureId);
      if (mj.contains("ior")) m.ior = mj["ior"].f();
    } else if (t == "brdf_rough_plastic" || t == "brdf_plastic") {
      m.type = t == "brdf_plastic" ? MaterialTypeBrdfPlastic : MaterialTypeBrdfRoughPlastic;
      if (mj.contains("ior")) m.ior = mj["ior"].f();
      opt3("diffuse_reflectance", m.diffuse);
      optTex("diffuse_texture", m.diffuseTextureId);
      optTex("normal_texture", m.normalTextureId);
      if (t == "brdf_rough_plastic") {
        if (mj.contains("alpha")) vec2(mj["alpha"], m.anisoAlpha[0], m.anisoAlpha[1]);
        optTex("alpha_texture", m.roughnessTextureId);
      }
      m.radiance[0] = diffuseFresnel(m.ior, 1000);  // fdrInt
    } else if (t == "brdf_kang18") {  // material.cpp:156-186
      m.type = MaterialTypeBrdfKang18;
      optTex("normal_texture", m.normalTextureId);
      optTex("tangent_texture", m.tangentTextureId);
      if (mj.contains("diffuse_texture")) optTex("diffuse_texture", m.diffuseTextureId);
      else set3(m.diffuse, vec3(mj["diffuse_reflectance"]));
      if (mj.contains("specular_texture")) optTex("specular_texture", m.metalnessTextureId);
      else set3(m.rhoSpec, vec3(mj["specular_reflectance"]));
      if (mj.contains("alpha_texture")) optTex("alpha_texture", m.roughnessTextureId);
      else vec2(mj["alpha"], m.anisoAlpha[0], m.anisoAlpha[1]);
      m.metalness = 0.f;  // used as opacity
      optTex("opacity_texture", m.opacityTextureId);
    } else if (t == "brdf_mirror") {  // material.cpp:232-243
      m.type = MaterialTypeBrdfMirror;
      opt3("diffuse_reflectance", m.diffuse);
      optTex("diffuse_texture", m.diffuseTextureId);
      optTex("normal_texture", m.normalTextureId);
    } else if (t == "brdf_conductor" || t == "brdf_rough_conductor") {  // material.cpp:303-356
      m.type = t == "brdf_conductor" ? MaterialTypeBrdfConductor : MaterialTypeBrdfRoughConductor;
      std::string metal = mj.contains("material") ? mj["material"].str() : std::string("Cu");
      const ComplexIor* ci = nullptr;
      for (const ComplexIor& c : kComplexIor)
        if (metal == c.name) ci = &c;
      if (!ci) fail("Loader: unrecognized material name in brdf_conductor [" + metal + "]");
      for (int k = 0; k < 3; ++k) m.radiance[k] = ci->eta[k], m.radianceFactor[k] = ci->k[k];
      opt3("diffuse_reflectance", m.diffuse);
      optTex("diffuse_texture", m.diffuseTextureId);
      optTex("normal_texture", m.normalTextureId);
      if (t == "brdf_rough_conductor") {
        if (mj.contains("alpha")) vec2(mj["alpha"], m.anisoAlpha[0], m.anisoAlpha[1]);
        optTex("alpha_texture", m.roughnessTextureId);
      }
    } else if (t == "brdf_disney") {  // material.cpp:358-388
      m.type = MaterialTypeBrdfDisney;
      optTex("normal_texture", m.normalTextureId);
      opt3("diffuse_reflectance", m.diffuse);
      optTex("diffuse_texture", m.diffuseTextureId);
      if (mj.contains("metallic")) m.metalness = mj["metallic"].f();
      optTex("metallic_texture", m.metalnessTextureId);
      if (mj.contains("roughness")) m.roughness = mj["roughness"].f();
      optTex("roughness_texture", m.roughnessTextureId);
      m.rhoSpec[0] = mj.contains("opacity") ? mj["opacity"].f() : 0.f;
      optTex("opacity_texture", m.opacityTextureId);
    } else if (t == "brdf_phong") {  // material.cpp:391-405
      m.type = MaterialTypeBrdfPhong;
      optTex("normal_texture", m.normalTextureId);
      opt3("diffuse_reflectance", m.diffuse);
      optTex("diffuse_texture", m.diffuseTextureId);
      opt3("specular_reflectance", m.rhoSpec);
      if (mj.contains("shininess")) m.specular = mj["shininess"].f();
    } else {
      fail("Loader: unrecognized material type [" + t + "]");
    }
    if (s_.materialIds.count(name)) fail("Scene: material [\"" + name + "\"] defined twice");
    s_.materialIds[name] = uint32_t(s_.materials.size());
    s_.materials.push_back(m);
  }

  uint32_t addMeshData(const std::string& name, MeshData&& md) {
    if (s_.meshIds.count(name)) fail("Scene: mesh [\"" + name + "\"] defined twice");
    uint32_t id = uint32_t(s_.meshes.size());
    s_.meshIds[name] = id;
    s_.meshes.push_back(std::move(md));
    return id;
  }

  // Scene::addLight(const GpuLight&), scene.cpp:220-244: rect lights become a
  // 4-vertex / 2-triangle mesh (Mesh::Mesh(Primitive), mesh.cpp:18-49) and an
  // emitter instance.
  void addLight(const GpuLight& light) {
    int32_t lightId = int32_t(s_.lights.size());
    if (light.type == LightTypeRect) {
      MeshData md;
      md.vertices.assign(4, GpuVertex{});
      Vec3f p(light.position[0], light.position[1], light.position[2]);
      Vec3f u(light.u[0], light.u[1], light.u[2]), v(light.v[0], light.v[1], light.v[2]);
      set3(md.vertices[0].pos, p);
      set3(md.vertices[1].pos, p + u);
      set3(md.vertices[2].pos, p + u + v);
      set3(md.vertices[3].pos, p + v);
      md.indices = {0, 1, 2, 0, 2, 3};
      for (auto& vx : md.vertices) updateAabb(md, vx);
      char name[40];
      std::snprintf(name, sizeof name, "__rectLight:%d", lightId);
      uint32_t meshId = addMeshData(name, std::move(md));
      InstanceRec in;
      in.mesh = meshId;
      in.light = lightId;
      s_.instances.push_back(in);
    }
    s_.lights.push_back(light);
  }

  // Scene::addLight(light, meshPath), scene.cpp:246-283: one triangle light
  // (+ mesh + instance) per facet.
  void addMeshLight(GpuLight light, const std::string& path) {
    std::vector<GpuVertex> v;
    std::vector<uint32_t> idx;
    loadObj(path, v, idx);
    light.type = LightTypeTriangle;
    for (size_t i = 0; i + 2 < idx.size(); i += 3) {
      const GpuVertex &v0 = v[idx[i]], &v1 = v[idx[i + 1]], &v2 = v[idx[i + 2]];
      Vec3f p0(v0.pos[0], v0.pos[1], v0.pos[2]);
      Vec3f u = Vec3f(v1.pos[0], v1.pos[1], v1.pos[2]) - p0;
      Vec3f w = Vec3f(v2.pos[0], v2.pos[1], v2.pos[2]) - p0;
      set3(light.position, p0);
      set3(light.u, u);
      set3(light.v, w);
      light.area = len3(cross3(u, w)) * 0.5f;
      light.doubleSide = 0;
      MeshData md;
      md.vertices.assign(3, GpuVertex{});
      set3(md.vertices[0].pos, p0);
      set3(md.vertices[1].pos, p0 + u);
      set3(md.vertices[2].pos, p0 + w);
      md.indices = {0, 1, 2};
      for (auto& vx : md.vertices) updateAabb(md, vx);
      int32_t lightId = int32_t(s_.lights.size());
      char name[40];
      std::snprintf(name, sizeof name, "__meshLight:%d", lightId);
      uint32_t meshId = addMeshData(name, std::move(md));
      InstanceRec in;
      in.mesh = meshId;
      in.light = lightId;
      s_.instances.push_back(in);
      s_.lights.push_back(light);
    }
  }

  void addLightJson(const Json& lj) {
    // Loader::addLight, loader.cpp:270-332
    checkKeys(lj, {"type", "radiance"});
    GpuLight l;
    std::memset(&l, 0, sizeof l);
    l.type = LightTypeUndefined;
    set3(l.radiance, vec3(lj["radiance"]));
    const std::string& t = lj["type"].str();
    if (t == "rect" || t == "triangle") {
      checkKeys(lj, {"position", "v1", "v2"});
      Vec3f p = vec3(lj["position"]), a = vec3(lj["v1"]), b = vec3(lj["v2"]);
      Vec3f u = a - p, v = b - p;
      set3(l.position, p);
      set3(l.u, u);
      set3(l.v, v);
      l.area = len3(cross3(u, v));
      l.type = LightTypeRect;
      if (lj.contains("double_side")) l.doubleSide = lj["double_side"].b() ? 1 : 0;
      if (t == "triangle") l.area *= 0.5f;  // still a rect light, with half the area
    } else if (t == "point") {
      checkKeys(lj, {"position"});
      set3(l.position, vec3(lj["position"]));
      l.type = LightTypePoint;
    } else if (t == "distant") {
      checkKeys(lj, {"direction"});
      set3(l.direction, vec3(lj["direction"]));
      l.type = LightTypeDirectional;
    } else if (t == "mesh") {
      checkKeys(lj, {"path"});
      std::string p = findFile(lj["path"].str(), s_.dir);
      if (p.empty()) fail("Loader: failed to load mesh light from [" + lj["path"].str() + "]");
      addMeshLight(l, p);
      return;
    } else {
      fail("Loader: unrecognized light type [" + t + "]");
    }
    addLight(l);
  }

  void addEnvMap(const Json& ej) {
    checkKeys(ej, {"path"});
    std::string p = findFile(ej["path"].str(), s_.dir);
    if (p.empty()) fail("Loader: failed to load envmap from file [" + ej["path"].str() + "]");
    std::string err;
    Image im;
    if (!readImage(p, 1.0f, im, err)) fail(err);
    s_.env = std::move(im);
    s_.hasEnv = true;
    s_.state.rtx.hasEnvMap = 1;
    s_.state.rtx.envMapResolution[0] = float(s_.env.w);
    s_.state.rtx.envMapResolution[1] = float(s_.env.h);
  }

  void addMesh(const Json& mj) {
    // Loader::addMesh + Mesh::Mesh(path, recomputeNormal, uvScale),
    // loader.cpp:349-364, mesh.cpp:51-77
    checkKeys(mj, {"name", "path"});
    std::string name = mj["name"].str();
    std::string p = findFile(mj["path"].str(), s_.dir);
    if (p.empty()) fail("Loader: failed to load mesh from file [" + mj["path"].str() + "]");
    bool recompute = mj.contains("recompute_normal") ? mj["recompute_normal"].b() : false;
    float us = 1.f, vs = 1.f;
    if (mj.contains("uv_scale")) vec2(mj["uv_scale"], us, vs);
    MeshData md;
    loadObj(p, md.vertices, md.indices);
    for (size_t i = 0; i + 2 < md.indices.size(); i += 3) {
      GpuVertex& v0 = md.vertices[md.indices[i]];
      GpuVertex& v1 = md.vertices[md.indices[i + 1]];
      GpuVertex& v2 = md.vertices[md.indices[i + 2]];
      if (recompute) {
        Vec3f a(v0.pos[0], v0.pos[1], v0.pos[2]), b(v1.pos[0], v1.pos[1], v1.pos[2]),
            c(v2.pos[0], v2.pos[1], v2.pos[2]);
        Vec3f n = norm3(cross3(b - a, c - a));
        set3(v0.normal, n);
        set3(v1.normal, n);
        set3(v2.normal, n);
      }
      for (GpuVertex* vx : {&v0, &v1, &v2}) {
        vx->uv[0] = us * vx->uv[0];
        vx->uv[1] = vs * vx->uv[1];
        updateAabb(md, *vx);
      }
    }
    addMeshData(name, std::move(md));
  }

  void addInstance(const Json& ij) {
    // loader.cpp:398-409: a missing material reads a missing default key
    // in the reference and throws; here it is a load error too.
    checkKeys(ij, {"mesh"});
    const std::string& mesh = ij["mesh"].str();
    if (!ij.contains("material")) fail("Loader: instance of mesh [" + mesh + "] names no material");
    const std::string& mat = ij["material"].str();
    auto mi = s_.meshIds.find(mesh);
    if (mi == s_.meshIds.end()) fail("Scene: mesh [\"" + mesh + "\"] does not exist");
    auto ma = s_.materialIds.find(mat);
    if (ma == s_.materialIds.end()) fail("Scene: material [\"" + mat + "\"] does not exist");
    InstanceRec in;
    if (ij.contains("toworld")) in.transform = parseToWorld(ij["toworld"], false);
    in.mesh = mi->second;
    in.material = ma->second;
    s_.instances.push_back(in);
  }

  void addShot(const Json& sj) {
    // Loader::addShot, loader.cpp:411-457
    checkKeys(sj, {"type"});
    const std::string& t = sj["type"].str();
    Shot sh;
    if (t == "lookat") {
      checkKeys(sj, {"eye", "lookat", "up"});
      sh.eye = vec3(sj["eye"]);
      sh.lookat = vec3(sj["lookat"]);
      sh.up = vec3(sj["up"]);
    } else if (t == "toworld" || t == "opencv") {
      checkKeys(sj, {"matrix"});
      Mat4 m = mat4RowMajor(sj["matrix"]);
      Mat4 c2w = t == "toworld" ? m : invertRotTrans(m);
      sh.eye = Vec3f(c2w.at(0, 3), c2w.at(1, 3), c2w.at(2, 3));
      float uy = t == "toworld" ? 1.f : -1.f;
      sh.up = Vec3f(c2w.at(0, 1) * uy, c2w.at(1, 1) * uy, c2w.at(2, 1) * uy);
      sh.lookat = Vec3f(c2w.at(0, 2) + c2w.at(0, 3), c2w.at(1, 2) + c2w.at(1, 3), c2w.at(2, 2) + c2w.at(2, 3));
    } else {
      fail("Loader: unrecognized shot type [" + t + "]");
    }
    sh.state = s_.state;
    if (sj.contains("state")) parseState(sj["state"], sh.state);
    if (sj.contains("fov")) sh.fov = sj["fov"].f();  // extension
    if (sj.contains("env_toworld")) sh.envTransform = parseToWorld(sj["env_toworld"], true);
    s_.shots.push_back(sh);
  }

  void submit() {
    // Scene::submit host part, scene.cpp:26-85
    s_.state.rtx.numLights = int32_t(s_.lights.size()) - 1;
    for (auto& sh : s_.shots) sh.state.rtx.numLights = s_.state.rtx.numLights;
    if (s_.shots.empty()) fitCamera();
    s_.buildDescs();
  }

  void fitCamera() {
    // Scene::computeSceneDimensions + fitCamera (scene.cpp:520-553).  The
    // reference delegates to nvh::CameraManipulator::fit, which is not in the
    // source tree; this frames the bounding sphere along +z with the film's
    // horizontal fov.
    Vec3f lo(1e30f, 1e30f, 1e30f), hi(-1e30f, -1e30f, -1e30f);
    for (const InstanceRec& in : s_.instances) {
      const MeshData& md = s_.meshes[in.mesh];
      for (int k = 0; k < 8; ++k) {
        Vec3f p((k & 1) ? md.posMax.x : md.posMin.x, (k & 2) ? md.posMax.y : md.posMin.y,
                (k & 4) ? md.posMax.z : md.posMin.z);
        const Mat4& m = in.transform;
        Vec3f w(m.at(0, 0) * p.x + m.at(0, 1) * p.y + m.at(0, 2) * p.z + m.at(0, 3),
                m.at(1, 0) * p.x + m.at(1, 1) * p.y + m.at(1, 2) * p.z + m.at(1, 3),
                m.at(2, 0) * p.x + m.at(2, 1) * p.y + m.at(2, 2) * p.z + m.at(2, 3));
        lo = Vec3f(std::min(lo.x, w.x), std::min(lo.y, w.y), std::min(lo.z, w.z));
        hi = Vec3f(std::max(hi.x, w.x), std::max(hi.y, w.y), std::max(hi.z, w.z));
      }
    }
    if (!(hi.x > lo.x) || !(hi.y > lo.y) || !(hi.z > lo.z)) lo = Vec3f(-1, -1, -1), hi = Vec3f(1, 1, 1);
    Vec3f ctr = (lo + hi) * 0.5f;
    float radius = len3(hi - lo) * 0.5f;
    float fov = s_.camera.fov * kToRad;
    float dist = radius / std::sin(0.5f * fov);
    Shot sh;
    sh.lookat = ctr;
    sh.eye = ctr + Vec3f(0, 0, dist);
    sh.up = Vec3f(0, 1, 0);
    sh.state = s_.state;
    s_.shots.push_back(sh);
  }
};

}  // namespace

GpuCamera cameraToGpu(const CameraModel& cam, Vec3f eye, Vec3f lookat, Vec3f up, const Mat4& env) {
  // Camera::getView (camera.h:26-28) then toGpuStruct (camera.cpp:77-98).
  GpuCamera g;
  std::memset(&g, 0, sizeof g);
  Mat4 view = mul(scaleMat(1.f, -1.f, -1.f), lookAt(eye, lookat, up));
  Mat4 c2w = invertRotTrans(view);
  std::memcpy(g.cameraToWorld, c2w.m, 64);
  std::memcpy(g.envTransform, env.m, 64);
  g.type = cam.type;
  if (cam.type == CameraTypePerspective) {
    // cameraToRasterTransform (camera.cpp:28-62) and its inverse
    float aspect = float(cam.width) / float(cam.height);
    Mat4 persp = Mat4::zero();
    float nearz = 0.1f, farz = 100.0f;
    float recip = 1.0f / (farz - nearz);
    float ctot = float(1.0 / double(std::tan(float(double(cam.fov * kToRad) * 0.5))));
    persp.at(0, 0) = ctot;
    persp.at(1, 1) = ctot;
    persp.at(2, 2) = farz * recip;
    persp.at(2, 3) = -nearz * farz * recip;
    persp.at(3, 2) = 1.f;
    Mat4 c2s = mul(scaleMat(1.f, aspect, 1.f), persp);
    c2s = mul(translateMat(1.f, 1.f, 0.f), c2s);
    Mat4 c2n = mul(scaleMat(0.5f, 0.5f, 1.f), c2s);
    Mat4 c2r = mul(scaleMat(float(cam.width), float(cam.height), 1.f), c2n);
    Mat4 r2c;
    invertGeneral(c2r, r2c);
    std::memcpy(g.rasterToCamera, r2c.m, 64);
    g.aperture = cam.aperture;
    g.focalDistance = cam.focal;
  } else {
    std::memcpy(g.fxfycxcy, cam.fxfycxcy, 16);
  }
  return g;
}

void buildEnvTables(uint32_t width, uint32_t height, const float* px, std::vector<float>& marg,
                    std::vector<float>& cond) {
  // EnvMap::EnvMap, texture.cpp:144-226.  The weight literals and the 1e-7
  // guard are double in the reference, so those steps round through double.
  const int W = int(width), H = int(height);
  std::vector<float> pdf2(size_t(W) * H), cdf2(size_t(W) * H), pdf1(H), cdf1(H);
  float total = 0.f;
  for (int j = 0; j < H; ++j) {
    float row = 0.f;
    size_t base = size_t(j) * W;
    for (int i = 0; i < W; ++i) {
      const float* c = px + (base + i) * 4;
      float w = float(0.3 * c[0] + 0.6 * c[1] + 0.1 * c[2]);
      row += w;
      pdf2[base + i] = w;
      cdf2[base + i] = row;
    }
    const double denom = double(row) + 1e-7;
    for (int i = 0; i < W; ++i) {
      pdf2[base + i] = float(double(pdf2[base + i]) / denom);
      cdf2[base + i] = float(double(cdf2[base + i]) / denom);
    }
    total += row;
    pdf1[j] = row;
    cdf1[j] = total;
  }
  const double tden = double(total) + 1e-7;
  for (int j = 0; j < H; ++j) {
    cdf1[j] = float(double(cdf1[j]) / tden);
    pdf1[j] = float(double(pdf1[j]) / tden);
  }
  // First index whose cdf is >= value (std::lower_bound on a sorted range).
  auto firstNotLess = [](const float* a, int n, float value) {
    return int(std::lower_bound(a, a + n, value) - a);
  };
  marg.assign(size_t(H) * 2, 0.f);
  cond.assign(size_t(W) * H * 2, 0.f);
  for (int j = 0; j < H; ++j) {
    float target = float(j + 1) / float(H);
    marg[size_t(j) * 2] = float(firstNotLess(cdf1.data(), H, target)) / float(H);
    marg[size_t(j) * 2 + 1] = pdf1[j];
  }
  for (int j = 0; j < H; ++j) {
    const float* rowCdf = cdf2.data() + size_t(j) * W;
    for (int i = 0; i < W; ++i) {
      float target = float(i + 1) / float(W);
      size_t o = (size_t(j) * W + i) * 2;
      cond[o] = float(firstNotLess(rowCdf, W, target)) / float(W);
      cond[o + 1] = pdf2[size_t(j) * W + i];
    }
  }
}

void loadSceneFile(const std::string& path, ash_host_scene& s) { Loader(s).load(path); }

}  // namespace ash

void ash_host_scene::buildDescs() {
  meshDescs.resize(meshes.size());
  for (size_t i = 0; i < meshes.size(); ++i) {
    meshDescs[i].num_vertices = uint32_t(meshes[i].vertices.size());
    meshDescs[i].num_indices = uint32_t(meshes[i].indices.size());
    meshDescs[i].vertices = meshes[i].vertices.data();
    meshDescs[i].indices = meshes[i].indices.data();
  }
  instDescs.resize(instances.size());
  for (size_t i = 0; i < instances.size(); ++i) {
    std::memcpy(instDescs[i].transform, instances[i].transform.m, 64);
    instDescs[i].mesh_id = instances[i].mesh;
    instDescs[i].material_id = instances[i].material;
    instDescs[i].light_id = instances[i].light;
    instDescs[i].pad = 0;
  }
  texDescs.resize(textures.size());
  for (size_t i = 0; i < textures.size(); ++i) {
    texDescs[i].width = textures[i].w;
    texDescs[i].height = textures[i].h;
    texDescs[i].rgba = textures[i].rgba.data();
  }
}
