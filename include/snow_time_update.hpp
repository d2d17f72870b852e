/*
 * snow_time_update.hpp - C++ step interface over the snow C ABI (header only).
 *
 * GL-free restatement of the reference's simulation-step classes:
 *   TimeUpdate { virtual void init() = 0; virtual void update(double dt) = 0; }
 *       (src/snow/simulation/timeUpdate.hpp:10-18)
 *   ExplicitTimeUpdate::init / update
 *       (src/snow/simulation/timeUpdate/explicittimeupdate.cpp:4-63)
 *   physicEngine / MPMPhysicEngine { bool init(); void update(double); }
 *       (src/snow/simulation/physicEngine.hpp:7-16, MPMPhysicEngine.cpp:3-7)
 * The reference binds its buffers implicitly through OpenGL SSBO binding points
 * (defines.hpp:30-48); here the scene is handed over explicitly once, as a snow_scene_desc,
 * and the state stays in device memory between steps. Errors are thrown as
 * std::runtime_error with the library's message (the reference has no error path).
 */
#ifndef SNOW_TIME_UPDATE_HPP
#define SNOW_TIME_UPDATE_HPP

#include <cstdint>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "snow.h"

namespace snow_amd {

class TimeUpdate {
 public:
  virtual ~TimeUpdate() = default;
  virtual void init() = 0;
  virtual void update(double dt) = 0;
};

// Explicit MPM step of one scene, or of a batch of independent scenes sharing one launch
// per stage (each scene keeps its own material, grid, colliders and dt).
class SnowExplicitTimeUpdate : public TimeUpdate {
 public:
  explicit SnowExplicitTimeUpdate(const snow_scene_desc& scene, uint32_t flags = 0)
      : SnowExplicitTimeUpdate(std::vector<snow_scene_desc>{scene}, flags) {}
  SnowExplicitTimeUpdate(const std::vector<snow_scene_desc>& scenes, uint32_t flags = 0) {
    check(snow_create_batch(scenes.data(), int32_t(scenes.size()), flags, &ctx_), nullptr);
  }
  ~SnowExplicitTimeUpdate() override {
    if (ctx_) snow_destroy(ctx_);
  }
  SnowExplicitTimeUpdate(const SnowExplicitTimeUpdate&) = delete;
  SnowExplicitTimeUpdate& operator=(const SnowExplicitTimeUpdate&) = delete;

  // ExplicitTimeUpdate::init: grid mass and particle rest density (explicittimeupdate.cpp:4-34).
  void init() override { check(snow_init(ctx_)); }
  // ExplicitTimeUpdate::update(dt): one step of every scene with this dt (explicittimeupdate.cpp:36-63).
  void update(double dt) override { check(snow_step(ctx_, dt)); }
  // One step of every scene with the scene's own dt.
  void update_batch() { check(snow_step_batch(ctx_)); }
  // `steps` steps as replays of the captured step graph (each scene's own dt).
  void run(int32_t steps) { check(snow_run_graph(ctx_, steps)); }
  void sync() { check(snow_sync(ctx_)); }

  int num_scenes() const { return snow_num_scenes(ctx_); }
  int64_t num_particles(int32_t scene = 0) const { return snow_num_particles(ctx_, scene); }

  // Particle state of one scene in creation order ([N][3] x, v; [N][9] F_E, F_P column-major).
  void read_particles(int32_t scene, std::vector<float>* x, std::vector<float>* v, std::vector<float>* fe,
                      std::vector<float>* fp) {
    const size_t n = size_t(num_particles(scene));
    if (x) x->resize(3 * n);
    if (v) v->resize(3 * n);
    if (fe) fe->resize(9 * n);
    if (fp) fp->resize(9 * n);
    check(snow_read_particles(ctx_, scene, x ? x->data() : nullptr, v ? v->data() : nullptr,
                              fe ? fe->data() : nullptr, fp ? fp->data() : nullptr, nullptr));
  }

  // One frame of one scene as a binary little-endian PLY point set with the properties x, y, z and
  // mass: the vec4 (x, m) buffer the reference renders as GL points (particlesystem.cpp:4-13,
  // pPoints.vert:2-12), in creation order. The host must be little-endian.
  void write_ply(int32_t scene, const std::string& path) {
    const size_t n = size_t(num_particles(scene));
    std::vector<float> pxm(4 * n);
    check(snow_export_reference_particles(ctx_, scene, pxm.data(), nullptr, nullptr, nullptr));
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("snow: cannot write " + path);
    f << "ply\nformat binary_little_endian 1.0\nelement vertex " << n
      << "\nproperty float x\nproperty float y\nproperty float z\nproperty float mass\nend_header\n";
    f.write(reinterpret_cast<const char*>(pxm.data()), std::streamsize(pxm.size() * sizeof(float)));
    if (!f) throw std::runtime_error("snow: write failed: " + path);
  }

  snow_ctx* handle() { return ctx_; }

 protected:
  void check(int rc) { check(rc, ctx_); }
  static void check(int rc, const snow_ctx* c) {
    if (rc != SNOW_OK) throw std::runtime_error("snow: " + std::string(snow_last_error(c)));
  }

 private:
  snow_ctx* ctx_ = nullptr;
};

// SemiImplicitTimeUpdate (src/snow/simulation/timeUpdate/semiimplicittimeupdate.hpp:6-13, whose
// init/update are empty in the reference, semiimplicittimeupdate.cpp:3-5): the explicit step with
// the semi-implicit grid velocity solve of snow_step_semi_implicit (Stomakhin et al. 2013, eq. 9).
// init() and the state readback are the explicit class's.
class SnowSemiImplicitTimeUpdate : public SnowExplicitTimeUpdate {
 public:
  explicit SnowSemiImplicitTimeUpdate(const snow_scene_desc& scene, double beta = 1.0, double tol = 1e-5,
                                      int32_t max_iter = 100, uint32_t flags = 0)
      : SnowExplicitTimeUpdate(scene, flags), beta_(beta), tol_(tol), max_iter_(max_iter) {}

  // SemiImplicitTimeUpdate::update(dt): one step of every scene with the semi-implicit solve.
  void update(double dt) override { check(snow_step_semi_implicit(handle(), dt, beta_, tol_, max_iter_)); }
  // iterations and relative residual of a scene's last solve
  void stats(int32_t scene, int32_t* iterations, double* rel_residual) {
    check(snow_implicit_stats(handle(), scene, iterations, rel_residual));
  }

 private:
  double beta_, tol_;
  int32_t max_iter_;
};

}  // namespace snow_amd

#endif  // SNOW_TIME_UPDATE_HPP
