// asuna: command-line renderer (src/main.cpp:6-30).
//   asuna --scene <file.json> --offline [--out name] [--gpu_id N]
//         [--output_scanline] [--spp N]
// --spp is accepted like the reference; it only takes effect with
// --spp_override (the reference ignores it, scene.cpp:373).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/asuna_host.h"

int main(int argc, char** argv) {
  std::string scene = "PLEASE_SET_SCENE_PATH", out = "asuna_out.hdr";
  ash_offline_opts o;
  std::memset(&o, 0, sizeof o);
  o.render.depthOutChannel = -1;
  bool offline = false, override = false, missing = false;
  int spp = 0;
  for (int i = 1; i < argc && !missing; ++i) {
    std::string a = argv[i];
    // the value after a flag; a flag at the end of the line is reported below
    auto val = [&](const char* name) -> const char* {
      if (i + 1 >= argc) {
        std::fprintf(stderr, "missing value for %s\n", name);
        missing = true;
        return "";
      }
      return argv[++i];
    };
    if (a == "--help") {
      std::printf("usage: asuna --scene file.json --offline [--out name] [--gpu_id N] [--output_scanline] [--spp N]\n");
      return 0;
    } else if (a == "--offline") offline = true;
    else if (a == "--output_scanline") o.output_scanline = 1;
    else if (a == "--gpu_id") o.gpu_id = std::atoi(val("--gpu_id"));
    else if (a == "--out") out = val("--out");
    else if (a == "--scene") scene = val("--scene");
    else if (a == "--spp") spp = std::atoi(val("--spp"));
    else if (a == "--spp_override") override = true;
    else if (a == "--quiet") o.quiet = 1;
  }
  if (missing) return 2;
  if (!offline) {
    std::fprintf(stderr, "asuna: only --offline rendering is available (the interactive window is not part of this build)\n");
    return 2;
  }
  if (override) o.spp_override = spp;
  ash_stats st;
  char err[2048] = {0};
  int rc = ash_run_offline(scene.c_str(), out.c_str(), &o, &st, err, sizeof err);
  if (rc) {
    std::fprintf(stderr, "asuna: error %d: %s\n", rc, err);
    return 1;
  }
  std::printf("{\"paths\": %llu, \"extension_rays\": %llu, \"shadow_rays\": %llu, \"render_ms\": %.3f}\n",
              (unsigned long long)st.paths, (unsigned long long)st.extension_rays,
              (unsigned long long)st.shadow_rays, st.render_ms);
  return 0;
}
