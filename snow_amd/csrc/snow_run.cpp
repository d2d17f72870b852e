// snow_run.cpp - command-line driver of the snow step through the C++ step interface
// (include/snow_time_update.hpp), the way the reference's launchSnow drives ExplicitTimeUpdate
// (src/snow/utils/launchHelper/explicit.cpp:2-116), without a window or renderer.
//
//   snow_run SCENE_FILE STEPS OUT_FILE [flags] [--ply PREFIX [--every K]] [--semi-implicit BETA]
//   snow_run --scene NAME [--asset FILE] STEPS OUT_FILE [flags] [--ply PREFIX [--every K]]
//   snow_run --scene NAME [--asset FILE] --dump-scene SCENE_FILE
//
// --semi-implicit BETA steps with SnowSemiImplicitTimeUpdate (the reference's SemiImplicitTimeUpdate,
// empty there) at that beta, tolerance 1e-5 and at most 100 solver iterations per step.
// --scene builds one of the reference's scene programs (src/exe/<NAME>.cpp, include/snow_scenes.hpp)
// instead of reading a scene file; --dump-scene writes that scene as a scene file and exits.
// --ply writes PREFIX_<step>.ply frames (x, y, z, mass) every K steps (default: the last step only).
//
// SCENE_FILE (little endian): "SNOWSCN1", int64 n, int32 dim[3], float origin[3], float h,
//   snow_material (10 floats), int32 ncol, ncol x snow_collider (12 floats), float dt,
//   float x[3n], float v[3n], float mass[n], float fe[9n], float fp[9n] (column-major 3x3)
// OUT_FILE: "SNOWOUT1", int64 n, float x[3n], v[3n], fe[9n], fp[9n] after init + STEPS steps.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "snow_scenes.hpp"
#include "snow_time_update.hpp"

namespace {

using snow_amd::reference_scenes::Scene;

template <class T>
bool read_into(FILE* f, T* dst, size_t count) {
  return std::fread(dst, sizeof(T), count, f) == count;
}

template <class T>
bool write_from(FILE* f, const T* src, size_t count) {
  return std::fwrite(src, sizeof(T), count, f) == count;
}

bool read_scene_file(const char* path, Scene& s) {
  FILE* f = std::fopen(path, "rb");
  if (!f) {
    std::perror(path);
    return false;
  }
  char magic[8];
  int64_t n = 0;
  int32_t ncol = 0;
  bool ok = read_into(f, magic, 8) && std::memcmp(magic, "SNOWSCN1", 8) == 0 && read_into(f, &n, 1) && n >= 0 &&
            n <= (int64_t(1) << 31) && read_into(f, s.grid.dim, 3) && read_into(f, s.grid.origin, 3) &&
            read_into(f, &s.grid.spacing, 1) && read_into(f, &s.material, 1) && read_into(f, &ncol, 1) &&
            ncol >= 0 && ncol <= 4096;
  const size_t nn = ok ? static_cast<size_t>(n) : 0;
  s.colliders.resize(ok ? size_t(ncol) : 0);
  s.x.resize(3 * nn);
  s.v.resize(3 * nn);
  s.mass.resize(nn);
  s.fe.resize(9 * nn);
  s.fp.resize(9 * nn);
  ok = ok && read_into(f, s.colliders.data(), s.colliders.size()) && read_into(f, &s.dt, 1) &&
       read_into(f, s.x.data(), s.x.size()) && read_into(f, s.v.data(), s.v.size()) &&
       read_into(f, s.mass.data(), s.mass.size()) && read_into(f, s.fe.data(), s.fe.size()) &&
       read_into(f, s.fp.data(), s.fp.size());
  std::fclose(f);
  if (!ok) std::fprintf(stderr, "%s: not a SNOWSCN1 scene file\n", path);
  return ok;
}

bool write_scene_file(const char* path, const Scene& s) {
  FILE* f = std::fopen(path, "wb");
  if (!f) {
    std::perror(path);
    return false;
  }
  const int64_t n = s.size();
  const int32_t ncol = int32_t(s.colliders.size());
  const bool ok = write_from(f, "SNOWSCN1", 8) && write_from(f, &n, 1) && write_from(f, s.grid.dim, 3) &&
                  write_from(f, s.grid.origin, 3) && write_from(f, &s.grid.spacing, 1) &&
                  write_from(f, &s.material, 1) && write_from(f, &ncol, 1) &&
                  write_from(f, s.colliders.data(), s.colliders.size()) && write_from(f, &s.dt, 1) &&
                  write_from(f, s.x.data(), s.x.size()) && write_from(f, s.v.data(), s.v.size()) &&
                  write_from(f, s.mass.data(), s.mass.size()) && write_from(f, s.fe.data(), s.fe.size()) &&
                  write_from(f, s.fp.data(), s.fp.size());
  std::fclose(f);
  if (!ok) std::fprintf(stderr, "%s: write failed\n", path);
  return ok;
}

int usage(const char* prog) {
  std::fprintf(stderr,
               "usage: %s SCENE_FILE STEPS OUT_FILE [flags] [--ply PREFIX [--every K]] [--semi-implicit BETA]\n"
               "       %s --scene NAME [--asset FILE] STEPS OUT_FILE [flags] [--ply PREFIX [--every K]]\n"
               "       %s --scene NAME [--asset FILE] --dump-scene SCENE_FILE\n"
               "scenes:",
               prog, prog, prog);
  for (const auto& name : snow_amd::reference_scenes::scene_names()) std::fprintf(stderr, " %s", name.c_str());
  std::fprintf(stderr, "\n");
  return 2;
}

}  // namespace

int main(int argc, char** argv) {
  std::string scene_name, asset, dump, ply;
  long every = 0;
  double beta = -1.0;  // < 0: the explicit step
  std::vector<const char*> pos;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    const bool has_value = i + 1 < argc;
    if (a == "--scene" && has_value) scene_name = argv[++i];
    else if (a == "--asset" && has_value) asset = argv[++i];
    else if (a == "--dump-scene" && has_value) dump = argv[++i];
    else if (a == "--ply" && has_value) ply = argv[++i];
    else if (a == "--every" && has_value) every = std::strtol(argv[++i], nullptr, 10);
    else if (a == "--semi-implicit" && has_value) beta = std::strtod(argv[++i], nullptr);
    else if (a.rfind("--", 0) == 0) return usage(argv[0]);
    else pos.push_back(argv[i]);
  }

  Scene s;
  if (!scene_name.empty()) {
    try {
      s = snow_amd::reference_scenes::make_scene(scene_name, asset);
    } catch (const std::exception& e) {
      std::fprintf(stderr, "%s\n", e.what());
      return 2;
    }
    if (!dump.empty()) {
      if (!write_scene_file(dump.c_str(), s)) return 2;
      std::printf("snow_run: scene %s, %lld particles, %zu colliders -> %s\n", scene_name.c_str(),
                  static_cast<long long>(s.size()), s.colliders.size(), dump.c_str());
      return 0;
    }
  } else {
    if (pos.empty()) return usage(argv[0]);
    if (!read_scene_file(pos[0], s)) return 2;
    pos.erase(pos.begin());
  }
  if (pos.size() < 2) return usage(argv[0]);
  const long steps = std::strtol(pos[0], nullptr, 10);
  const char* out_path = pos[1];
  const uint32_t flags = pos.size() > 2 ? uint32_t(std::strtoul(pos[2], nullptr, 0)) : 0u;
  const snow_scene_desc d = s.desc();
  const int64_t n = s.size();

  try {
    std::unique_ptr<snow_amd::SnowExplicitTimeUpdate> owner;
    if (beta >= 0.0) owner.reset(new snow_amd::SnowSemiImplicitTimeUpdate(d, beta, 1e-5, 100, flags));
    else owner.reset(new snow_amd::SnowExplicitTimeUpdate(d, flags));
    snow_amd::SnowExplicitTimeUpdate& sim = *owner;  // TimeUpdate-shaped: init() and update(dt)
    sim.init();
    char frame[4096];
    for (long step = 1; step <= steps; ++step) {
      sim.update(d.dt);
      if (!ply.empty() && ((every > 0 && step % every == 0) || step == steps)) {
        std::snprintf(frame, sizeof frame, "%s_%06ld.ply", ply.c_str(), step);
        sim.write_ply(0, frame);
      }
    }
    std::vector<float> ox, ov, ofe, ofp;
    sim.read_particles(0, &ox, &ov, &ofe, &ofp);
    FILE* o = std::fopen(out_path, "wb");
    if (!o) {
      std::perror(out_path);
      return 2;
    }
    const bool wok = write_from(o, "SNOWOUT1", 8) && write_from(o, &n, 1) && write_from(o, ox.data(), ox.size()) &&
                     write_from(o, ov.data(), ov.size()) && write_from(o, ofe.data(), ofe.size()) &&
                     write_from(o, ofp.data(), ofp.size());
    std::fclose(o);
    if (!wok) {
      std::fprintf(stderr, "%s: write failed\n", out_path);
      return 2;
    }
    std::printf("snow_run: %lld particles, %ld steps\n", static_cast<long long>(n), steps);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
