/*
 * snow_scenes.hpp - the reference's scene programs as scene data for the snow C ABI
 * (header only, C++17).
 *
 * The reference ships its scenes as programs (src/exe/*.cpp) that fill a ParticleSystem and
 * call launchSnow (src/snow/utils/launchHelper/explicit.cpp:2-116). launchSnow adds the grid of
 * defines.hpp:14-21, six boundary half-planes after the scene's own colliders, and steps an
 * ExplicitTimeUpdate with PHYSIC_DT (defines.hpp:23). make_scene() builds the same scenes, with
 * the particle sets the reference creates:
 *  - The samplers draw C rand() scaled by 1/32727 (e.g. snowBallsCrash.cpp:12-15). That only
 *    works with a 15-bit rand(), the one of the Microsoft C runtime the reference is built with
 *    (src/snow/CMakeLists.txt:14 links opengl32.lib; with a 31-bit rand() the rejection loops
 *    would practically never accept). MsvcRand is that runtime's generator: a linear
 *    congruential x <- 214013 x + 2531011 (mod 2^32) that returns bits 16..30, seeded with 1 (the
 *    scene programs sample before launchSnow calls srand).
 *  - The sampling expressions keep their float arithmetic, including the double promotions of
 *    the mass rules, so positions and masses are the same floats.
 *  - Velocities are handed to the reference as Vector3i(v * 1e8) and decoded as float(iv) * 1e-8
 *    (p2gTransfer.glsl:67); Scene stores the decoded floats.
 * particles_from_file() reads "x,y,z" lines like ParticleSystem::initParticlesFromFile
 * (particlesystem.cpp:19-38); unlike the reference, a missing file is an error (std::runtime_error)
 * instead of a printed message and an empty scene. The bunny asset (model/bunny.voxel) is not in
 * the reference repository, so snowBunnyScene needs the file's path.
 */
#ifndef SNOW_SCENES_HPP
#define SNOW_SCENES_HPP

#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "snow.h"

namespace snow_amd {
namespace reference_scenes {

// defines.hpp:14-24 and particle.hpp:6 (default particle mass)
constexpr double kGridPos[3] = {0.5125, 0.5125, 1.5125};
constexpr int kGridDim = 201;
constexpr double kGridSpacing = 0.05;
constexpr int kPlaneOffset = 4;  // GRID_COLLISION_PLANE_OFFSET
constexpr double kPhysicDt = 1e-3;
constexpr float kDefaultMass = 1.25e-3f;

// The Microsoft C runtime's rand(): 15-bit outputs of a 32-bit linear congruential generator.
struct MsvcRand {
  uint32_t state = 1u;
  int next() {
    state = state * 214013u + 2531011u;
    return int((state >> 16) & 0x7fffu);
  }
  // float(rand()) / 32727.0f * extent, the reference's sample expression
  float sample(float extent) { return (float(next()) / 32727.0f) * extent; }
};

struct Scene {
  std::string name;
  std::vector<float> x, v, mass, fe, fp;  // [n][3], [n][3], [n], [n][9], [n][9] (column-major 3x3)
  std::vector<snow_collider> colliders;
  snow_grid grid{};
  snow_material material{};
  float dt = float(kPhysicDt);

  int64_t size() const { return int64_t(mass.size()); }

  // a particle with F_E = F_P = I and the reference's fixed-point velocity iv = v * 1e8
  void add(float px, float py, float pz, int32_t ivx, int32_t ivy, int32_t ivz, float m) {
    x.insert(x.end(), {px, py, pz});
    v.insert(v.end(), {float(ivx) * 1e-8f, float(ivy) * 1e-8f, float(ivz) * 1e-8f});
    mass.push_back(m);
    for (auto* f : {&fe, &fp})
      f->insert(f->end(), {1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 1.f});
  }

  // The descriptor points into this object's vectors: keep the Scene alive while it is used.
  snow_scene_desc desc() const {
    snow_scene_desc d{};
    d.num_particles = size();
    d.x = x.data();
    d.v = v.data();
    d.mass = mass.data();
    d.fe = fe.data();
    d.fp = fp.data();
    d.grid = grid;
    d.material = material;
    d.num_colliders = int32_t(colliders.size());
    d.colliders = colliders.data();
    d.dt = dt;
    return d;
  }
};

inline snow_collider make_collider(float px, float py, float pz, int type, float vx, float vy, float vz,
                                   float radius, float nx, float ny, float nz, float friction = 0.2f) {
  snow_collider c{};
  c.pos[0] = px; c.pos[1] = py; c.pos[2] = pz;
  c.type = float(type);
  c.vel[0] = vx; c.vel[1] = vy; c.vel[2] = vz;
  c.radius = radius;
  c.normal[0] = nx; c.normal[1] = ny; c.normal[2] = nz;
  c.friction = friction;  // collider.hpp:14-17 default
  return c;
}

// launchSnow's part of every scene (explicit.cpp:6-67): the defines.hpp grid and material, and the
// six inward half-planes GRID_COLLISION_PLANE_OFFSET cells inside the grid, appended after the
// scene's own colliders.
inline void finish_like_launch_snow(Scene& s) {
  snow_grid_defaults(&s.grid);
  snow_material_defaults(&s.material);
  float lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = float(kGridPos[a] + kPlaneOffset * kGridSpacing);
    hi[a] = float(kGridPos[a] + (kGridDim - kPlaneOffset) * kGridSpacing);
  }
  for (int a = 0; a < 3; ++a)
    s.colliders.push_back(make_collider(lo[0], lo[1], lo[2], 0, 0.f, 0.f, 0.f, 0.f, a == 0 ? 1.f : 0.f,
                                        a == 1 ? 1.f : 0.f, a == 2 ? 1.f : 0.f));
  for (int a = 0; a < 3; ++a)
    s.colliders.push_back(make_collider(hi[0], hi[1], hi[2], 0, 0.f, 0.f, 0.f, 0.f, a == 0 ? -1.f : 0.f,
                                        a == 1 ? -1.f : 0.f, a == 2 ? -1.f : 0.f));
  s.dt = float(kPhysicDt);
}

// Mass of a ball particle. Rule "fixed" draws no extra sample. The drop and crash rules draw a
// fourth sample r per candidate and keep (r > 0.5) ? ((r - 0.5) * 10 + 1) * scale : other,
// evaluated in double as the C++ literal 0.5 makes it (snowBallDrop.cpp:24-25,
// snowBallsCrash.cpp:19-20).
struct MassRule {
  bool random = false;
  float scale = 0.f, other = 0.f, fixed = kDefaultMass;
};

// n particles by rejection from the cube [corner, corner + width)^3, kept when inside the inscribed
// sphere (the sampling loops of snowBallDrop.cpp, snowBallSmash.cpp, snowBallsCrash.cpp,
// snowSingleParticle.cpp).
inline void sample_ball(MsvcRand& rng, Scene& s, uint32_t n, float cx, float cy, float cz, float width,
                        const int32_t iv[3], const MassRule& rule) {
  const float radius = width / 2.0f;
  for (uint32_t kept = 0; kept < n;) {
    const float r1 = rng.sample(width), r2 = rng.sample(width), r3 = rng.sample(width);
    const float r4 = rule.random ? rng.sample(width) : 0.f;
    const float d2 = (r1 - radius) * (r1 - radius) + (r2 - radius) * (r2 - radius) + (r3 - radius) * (r3 - radius);
    if (!(d2 < radius * radius)) continue;
    float m = rule.fixed;
    if (rule.random)
      m = (r4 > 0.5) ? float((((double(r4) - 0.5) * double(10.0f)) + double(1.0f)) * double(rule.scale)) : rule.other;
    s.add(cx + r1, cy + r2, cz + r3, iv[0], iv[1], iv[2], m);
    ++kept;
  }
}

// n particles uniformly in the box [corner, corner + (ex, ey, ez)) at rest with the default mass
// (snowPlowTrack.cpp:11-25, snowPlowField.cpp:11-24).
inline void sample_box(MsvcRand& rng, Scene& s, uint32_t n, float cx, float cy, float cz, float ex, float ey,
                       float ez) {
  for (uint32_t i = 0; i < n; ++i) {
    const float r1 = rng.sample(ex), r2 = rng.sample(ey), r3 = rng.sample(ez);
    s.add(cx + r1, cy + r2, cz + r3, 0, 0, 0, kDefaultMass);
  }
}

// Particles from a text file of "x,y,z" lines, placed at scale * p + pos with mass 6.25e-3
// (particlesystem.cpp:19-38). Returns the number read.
inline int64_t particles_from_file(Scene& s, const std::string& path, const float pos[3], const float scale[3]) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("snow_scenes: cannot open particle file " + path);
  std::string line;
  int64_t count = 0;
  while (std::getline(in, line)) {
    float p[3];
    std::string rest = line;
    for (int k = 0; k < 3; ++k) {
      const size_t comma = k < 2 ? rest.find(',') : std::string::npos;
      const std::string field = rest.substr(0, comma);
      char* endp = nullptr;
      p[k] = std::strtof(field.c_str(), &endp);
      if (endp == field.c_str()) throw std::runtime_error("snow_scenes: bad particle line '" + line + "' in " + path);
      if (k < 2) {
        if (comma == std::string::npos)
          throw std::runtime_error("snow_scenes: bad particle line '" + line + "' in " + path);
        rest.erase(0, comma + 1);
      }
    }
    s.add(scale[0] * p[0] + pos[0], scale[1] * p[1] + pos[1], scale[2] * p[2] + pos[2], 0, 0, 0, 6.25e-3f);
    ++count;
  }
  return count;
}

inline const std::vector<std::string>& scene_names() {
  static const std::vector<std::string> names = {"snowBallDrop",  "snowBallSmash", "snowBallsCrash",
                                                 "snowSingleParticle", "snowSquare", "snowPlowTrack",
                                                 "snowPlowField", "snowBunnyScene"};
  return names;
}

// The scene of src/exe/<name>.cpp. asset: the particle file of snowBunnyScene.
inline Scene make_scene(const std::string& name, const std::string& asset = "") {
  Scene s;
  s.name = name;
  MsvcRand rng;
  const double plane_y = kGridPos[1] + kPlaneOffset * kGridSpacing;  // GRID_POS_Y + offset * h
  if (name == "snowBallDrop") {  // snowBallDrop.cpp:14-35
    const int32_t iv[3] = {0, 0, 0};
    MassRule rule;
    rule.random = true; rule.scale = 1.0e-4f; rule.other = 5e-3f;
    sample_ball(rng, s, 32 * 32 * 32, 5.0f, 1.7125f, 5.0f, float(1.0), iv, rule);
  } else if (name == "snowBallSmash") {  // snowBallSmash.cpp:6-21
    const int32_t iv[3] = {int32_t(-10 * 1e8), 0, 0};
    MassRule rule;
    rule.fixed = 6.25e-3f;
    sample_ball(rng, s, 32 * 32 * 32, 2.0f, 1.525f, 5.0f, float(1.0), iv, rule);
  } else if (name == "snowBallsCrash") {  // snowBallsCrash.cpp:6-49
    MassRule rule;
    rule.random = true; rule.scale = 8.0e-4f; rule.other = 10e-3f;
    const int32_t iv1[3] = {int32_t(10 * 1e8), int32_t(0 * 1e8), int32_t(1 * 1e8)};
    sample_ball(rng, s, 32 * 32 * 16, 1.0f, 1.7125f, 5.0f, float(0.8), iv1, rule);
    const int32_t iv2[3] = {int32_t(-12 * 1e8), int32_t(1 * 1e8), int32_t(-5 * 1e8)};
    sample_ball(rng, s, 32 * 32 * 16, 4.5f, 1.7125f, 5.5f, float(0.8), iv2, rule);
  } else if (name == "snowSingleParticle") {  // snowSingleParticle.cpp:6-22
    const int32_t iv[3] = {0, 0, 0};
    MassRule rule;
    rule.fixed = 1.25e-3f;
    sample_ball(rng, s, 1, 5.0f, 1.7125f, 5.0f, float(1.0), iv, rule);
  } else if (name == "snowSquare") {  // snowSquare.cpp:5-25: a 32^3 lattice, y slowest, x fastest
    const float px = 0.9625f, py = 0.72f, pz = float(0.9625);
    for (int y = 0; y < 32; ++y)
      for (int z = 0; z < 32; ++z)
        for (int xi = 0; xi < 32; ++xi)
          s.add(px + float(xi) * 0.025f, py + float(y) * 0.025f, pz + float(z) * 0.025f, 0, 0, 0, kDefaultMass);
  } else if (name == "snowPlowTrack") {  // snowPlowTrack.cpp:5-37
    sample_box(rng, s, 32 * 32 * 64, float(0.8125), float(plane_y), float(2.4), float(2.0), float(0.5), float(0.8));
    s.colliders.push_back(make_collider(0.2125f, 1.1125f, 2.8f, 1, 5.0f, 0.f, 0.f, 0.3f, 0.f, 0.f, 0.f));
  } else if (name == "snowPlowField") {  // snowPlowField.cpp:5-58
    const float height = float(kGridDim * kGridSpacing * 0.1);
    const float width = float(kGridDim * kGridSpacing * 0.2);
    const float length = float(kGridDim * kGridSpacing * 0.2);
    sample_box(rng, s, 2, float(0.8125), float(plane_y), float(2.4), width, height, length);
    s.colliders.push_back(make_collider(0.4125f, 1.0125f, 2.8f, 1, 10.0f, 0.f, 0.f, 0.2f, 0.f, 0.f, 0.f));
    s.colliders.push_back(make_collider(4.0125f, 1.0125f, 1.2f, 1, -10.0f, 0.f, 0.f, 0.2f, 0.f, 0.f, 0.f));
    s.colliders.push_back(make_collider(1.2f, 1.0125f, 0.4125f, 1, 0.f, 0.f, 10.0f, 0.2f, 0.f, 0.f, 0.f));
  } else if (name == "snowBunnyScene") {  // snowBunnyScene.cpp:5-14
    if (asset.empty()) throw std::runtime_error("snow_scenes: snowBunnyScene needs its particle file (model/bunny.voxel)");
    const float pos[3] = {5.0f, 0.0f, 5.0f};
    const float sc = 12.5f * 1.259f;
    const float scale[3] = {sc, sc, sc};
    particles_from_file(s, asset, pos, scale);
  } else {
    throw std::runtime_error("snow_scenes: unknown scene " + name);
  }
  finish_like_launch_snow(s);
  return s;
}

}  // namespace reference_scenes
}  // namespace snow_amd

#endif  // SNOW_SCENES_HPP
