// snow_capi.hip - host side of the MI355X snow MPM step behind the C ABI of include/snow.h.
//
// Replaces ExplicitTimeUpdate (src/snow/simulation/timeUpdate/explicittimeupdate.cpp:4-63), its
// eight Technique wrappers (src/snow/simulation/computingTechnique/*.cpp) and the SSBO set-up of
// ParticleSystem / Grid / CollisionObjects (particlesystem.cpp:101-228, grid.cpp:74-148,
// collisionObjects.cpp:25-66). One context owns one HIP stream and all device memory of a
// batch of scenes; every launch of a step covers all scenes at once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/snow.h"
#include "snow_implicit.hpp"
#include "snow_kernels.hpp"
#include "snow_sort.hpp"

using namespace snow;

// Staging of snow_read_particles_noalloc: a fixed device array instead of memory the library
// allocates. Each GPU the library runs on holds its own copy until the process ends; every call on
// that GPU shares it (the window it reads is written in full by each call, so nothing is cleared).
constexpr int64_t NOALLOC_MAX_PARTICLES = 65536;
constexpr int64_t NOALLOC_FLOATS = 25;  // x 3, v 3, F_E 9, F_P 9, rho0 1 per particle
static __device__ __align__(256) float scratch[NOALLOC_MAX_PARTICLES * NOALLOC_FLOATS];

namespace {

enum Stage {
  ST_RIGID = 0,   // #Sim:Advance Rigids
  ST_RESET,       // #Sim:Scratch Grid
  ST_P2G,         // #Sim:Transfer to Grid
  ST_GRID,        // #Sim:Compute Derivatives
  ST_G2P,         // #Sim:Transfer to Particles (fused with the advance in snow_step)
  ST_ADVANCE,     // #Sim:Advance Particles (split stage only)
  ST_SORT,        // #Sim:Sort Particles (keys, radix sort, cell bounds, rigid advance)
  ST_COUNT
};
const char* kStageNames[ST_COUNT] = {"#Sim:Advance Rigids",        "#Sim:Scratch Grid",
                                     "#Sim:Transfer to Grid",      "#Sim:Compute Derivatives",
                                     "#Sim:Transfer to Particles", "#Sim:Advance Particles",
                                     "#Sim:Sort Particles"};

struct SceneHost {
  int64_t n;
  int D[3];
  int ncol;
};

thread_local std::string g_create_error;

template <class T>
inline unsigned grid_for(T n, unsigned block) {
  return unsigned((int64_t(n) + block - 1) / block);
}

}  // namespace

struct snow_ctx {
  hipStream_t stream = nullptr;
  int device = 0;
  uint32_t flags = 0;
  std::string err;

  int nscenes = 0;
  std::vector<SceneDev> scenes;
  std::vector<SceneHost> sh;
  std::vector<Collider> cols;
  uint32_t N = 0;
  int64_t total_nodes = 0;
  uint32_t total_ctiles = 0, total_ntiles = 0;
  int key_bits = 1;
  int sort_bits = 11;  // largest radix digit width (sort-bin width), tuned by snow_autotune

  SceneDev* d_scenes = nullptr;
  uint16_t *d_pscene = nullptr, *d_tile_scene = nullptr, *d_ntile_scene = nullptr;
  float4* d_rec[2] = {nullptr, nullptr};  // 2 float4 per particle: (x, m), (v, rho0)
  float* d_fe[2] = {nullptr, nullptr};
  float* d_fp[2] = {nullptr, nullptr};
  uint32_t* d_pid[2] = {nullptr, nullptr};
  float* d_fa = nullptr;  // force matrices, 9 planes, in the slot order of the current buffer
  uint32_t* d_keys[2] = {nullptr, nullptr};
  uint32_t* d_slot = nullptr;  // slot of the particle at each sorted position (the sort's permutation)
  void* d_sort_tmp = nullptr;
  size_t sort_tmp_bytes = 0;
  // The cell sort's first-pass histogram (rows of 2^hist_bits digits per 4096-key tile). The fused
  // G2P counts the keys it writes into it, and the sort's first scatter clears it again.
  uint32_t* d_hist0 = nullptr;
  int hist_bits = 0;        // digit bits of its row layout (0: none yet)
  bool hist_ready = false;  // it holds the counts of d_keys[0]; if not, the next sort recounts them
  uint32_t *d_tile_c0 = nullptr, *d_tile_c1 = nullptr;  // first and last P2G chunk of each tile's particles
  float* d_tpart = nullptr;
  float* d_hpart = nullptr;
  float* d_zero_part = nullptr;  // NQ * HALO_NODES zeros
  uint32_t* d_list_ctr = nullptr;  // TILE_LISTS append counters
  uint32_t* d_tile_list = nullptr; // TILE_LISTS sublists of total_ctiles entries
  StepCtl* d_ctl = nullptr;
  uint32_t* d_pbound = nullptr;    // [nscenes][2 parities][PB_SLOTS][PB_STRIDE] fixed-point bounds of the next P2G
  uint32_t* d_tile_stamp = nullptr;
  int num_cu = 256;
  float4 *d_g_mmv = nullptr, *d_g_f = nullptr;
  float2* d_g_node = nullptr;  // 3 float2 per node
  float* d_g_mass = nullptr;   // node mass of the last grid update
  Collider* d_cols = nullptr;
  // [nscenes][ACTIVE_SLOTS] active-node slots (ACTIVE_STRIDE words apart), then [1] fp64 count
  unsigned long long* d_counts = nullptr;
  int* d_bad = nullptr;
  float* d_aux = nullptr;
  void* d_stage = nullptr;
  size_t stage_bytes = 0;

  int cur = 0;
  bool keys_valid = false;
  bool force_valid = false;
  bool bound_valid = false;  // the next P2G's bound slots hold the bound of the current records
  bool initialized = false;
  int grid_state = 0;   // 0 nothing, 1 tiled (p2g/step), 2 dense (host write / dense update)
  bool keep_valid = false;
  bool aux_valid = false;

  bool timers = false;
  hipEvent_t ev[2] = {nullptr, nullptr};
  std::vector<hipEvent_t> ev_free;
  struct Pending { int st; hipEvent_t a, b; };
  std::vector<Pending> pending;
  double stage_ms[ST_COUNT] = {0};
  int64_t stage_calls[ST_COUNT] = {0};
  double stage_min_ms[ST_COUNT] = {0}, stage_max_ms[ST_COUNT] = {0}, stage_sq_ms[ST_COUNT] = {0};

  float* noalloc_stage = nullptr;  // device address of `scratch` on this context's GPU

  hipGraph_t graph[2] = {nullptr, nullptr};      // batched step starting from buffer parity 0 / 1
  hipGraphExec_t gexec[2] = {nullptr, nullptr};

  // launch shapes (autotuned)
  int g2p_block = 256;
  uint32_t total_chunks = 0;       // P2G workgroups (chunks of every scene)
  uint16_t* d_chunk_scene = nullptr;
  uint32_t* d_chunk_lo = nullptr;  // first sorted position of each P2G chunk, as the P2G chose it
  uint32_t* d_chunk_w = nullptr;   // planned window start of each P2G chunk (before tile alignment)
  uint32_t* d_chunk_runs = nullptr;  // runs (tiles) of each chunk in the last P2G
  uint32_t* d_fold_job = nullptr;    // per chunk: the split tile it folds in the last P2G (or ~0)
  unsigned long long* d_plan_pref = nullptr;  // cost prefix scratch of the chunk plan
  // G2P waves: each scene's sorted range cut into waves of 64 positions, (first position, scene)
  // per wave; only for batches of several scenes (one scene: wave w starts at 64 w)
  uint2* d_waves = nullptr;
  uint32_t nwaves = 0;

  // semi-implicit solve (allocated on the first snow_step_semi_implicit)
  float4* d_imp_vec = nullptr;           // IMP_VECS dense node vectors
  float* d_imp_mass = nullptr;           // node masses of the step (0 where unused)
  long long* d_imp_acc = nullptr;        // fixed-point df, 3 per node
  float4* d_imp_pd = nullptr;            // IMP_PLANES float4 planes of prepared particle data
  float* d_imp_G = nullptr;              // G_p: float4 planes G0..G3, G4..G7, then a float plane G8
  double* d_imp_part = nullptr;          // 3 partial sums per node block
  ImpBlock* d_imp_blocks = nullptr;      // node blocks (scene, first node)
  uint32_t* d_imp_scene_blocks = nullptr;  // first node block of every scene, then the total
  ImpScal* d_imp_scal = nullptr;         // per-scene solver state
  uint32_t* d_imp_gslots = nullptr;      // IMP_GSLOTS bound slots per scene, IMP_GSTRIDE words apart
  ImpScal* h_imp_scal = nullptr;         // pinned host copy for the convergence polls
  uint32_t imp_nblocks = 0;
  int imp_last_iters = 0;                // iterations of the last solve (for the first poll)
  uint8_t* d_imp_nlive = nullptr;        // per node tile: written by this step's grid update
  uint8_t* d_imp_blive = nullptr;        // per node-vector block: holds a node of such a tile
};

namespace {

// the count of particles that took the FP64 path, after the active-node slots of d_counts
unsigned long long* fp64_counter(snow_ctx* c) {
  return c->d_counts + size_t(c->nscenes) * ACTIVE_SLOTS * ACTIVE_STRIDE;
}

int fail(snow_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg; else g_create_error = msg;
  return code;
}

#define HIPCHK(ctx, expr)                                                                        \
  do {                                                                                           \
    hipError_t _e = (expr);                                                                      \
    if (_e != hipSuccess)                                                                        \
      return fail(ctx, SNOW_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));           \
  } while (0)

#define LAUNCHCHK(ctx)                                                                           \
  do {                                                                                           \
    hipError_t _e = hipGetLastError();                                                           \
    if (_e != hipSuccess) return fail(ctx, SNOW_E_HIP, std::string("launch: ") + hipGetErrorString(_e)); \
  } while (0)

// The device address of `scratch` on GPU `dev` (the current device), looked up once per device.
std::mutex g_scratch_mu;
float* g_scratch_addr[256] = {};

hipError_t scratch_address(int dev, float** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (dev < 0 || dev >= 256) return hipErrorInvalidDevice;
  if (!g_scratch_addr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_addr[dev] = static_cast<float*>(p);
  }
  *out = g_scratch_addr[dev];
  return hipSuccess;
}

template <class T>
hipError_t dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  return hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
}

// Stage timing: an event pair per timed launch, recorded on the context stream without any
// host synchronisation; pairs are resolved (and their events recycled) when times are read.
struct StageTimer {
  snow_ctx* c;
  int st;
  hipEvent_t a = nullptr;
  StageTimer(snow_ctx* c_, int st_);
  int done();
};

hipEvent_t take_event(snow_ctx* c) {
  if (!c->ev_free.empty()) {
    hipEvent_t e = c->ev_free.back();
    c->ev_free.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  // timing-only events: no system-scope fence (cache writeback) between the timed kernels
  if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
  return e;
}

StageTimer::StageTimer(snow_ctx* c_, int st_) : c(c_), st(st_) {
  if (!c->timers) return;
  a = take_event(c);
  if (a) (void)hipEventRecord(a, c->stream);
}

int StageTimer::done() {
  if (!c->timers || !a) return 0;
  hipEvent_t b = take_event(c);
  if (!b) return fail(c, SNOW_E_HIP, "hipEventCreate failed");
  if (hipEventRecord(b, c->stream) != hipSuccess) return fail(c, SNOW_E_HIP, "hipEventRecord failed");
  c->pending.push_back({st, a, b});
  return 0;
}

int resolve_timers(snow_ctx* c) {
  if (c->pending.empty()) return 0;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (auto& p : c->pending) {
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, p.a, p.b));
    const double d = ms;
    if (c->stage_calls[p.st] == 0 || d < c->stage_min_ms[p.st]) c->stage_min_ms[p.st] = d;
    if (c->stage_calls[p.st] == 0 || d > c->stage_max_ms[p.st]) c->stage_max_ms[p.st] = d;
    c->stage_ms[p.st] += d;
    c->stage_sq_ms[p.st] += d * d;
    c->stage_calls[p.st] += 1;
    c->ev_free.push_back(p.a);
    c->ev_free.push_back(p.b);
  }
  c->pending.clear();
  return 0;
}

int validate_scene(const snow_scene_desc& d, int idx, std::string& msg) {
  char buf[256];
  if (d.num_particles < 0 || d.num_particles > (int64_t(1) << 31)) {
    snprintf(buf, sizeof buf, "scene %d: num_particles out of range", idx);
    msg = buf; return -1;
  }
  if (d.num_particles > 0 && (!d.x || !d.v || !d.mass)) {
    snprintf(buf, sizeof buf, "scene %d: x, v and mass are required", idx);
    msg = buf; return -1;
  }
  for (int k = 0; k < 3; ++k)
    if (d.grid.dim[k] <= 0 || d.grid.dim[k] > 4096) {
      snprintf(buf, sizeof buf, "scene %d: grid dim %d must be in [1, 4096]", idx, k);
      msg = buf; return -1;
    }
  if (!(d.grid.spacing > 0.0f) || !std::isfinite(d.grid.spacing)) {
    snprintf(buf, sizeof buf, "scene %d: grid spacing must be > 0", idx);
    msg = buf; return -1;
  }
  if (d.num_colliders < 0 || (d.num_colliders > 0 && !d.colliders)) {
    snprintf(buf, sizeof buf, "scene %d: bad collider list", idx);
    msg = buf; return -1;
  }
  for (int c = 0; c < d.num_colliders; ++c) {
    const float t = d.colliders[c].type;
    if (!(t >= 0.0f) || !std::isfinite(t)) {
      snprintf(buf, sizeof buf, "scene %d: collider %d has invalid type", idx, c);
      msg = buf; return -1;
    }
  }
  if (!(d.material.poisson < 0.5f) || !(d.material.poisson > -1.0f)) {
    snprintf(buf, sizeof buf, "scene %d: poisson ratio must be in (-1, 0.5)", idx);
    msg = buf; return -1;
  }
  if (d.dt < 0.0f || !std::isfinite(d.dt)) {
    snprintf(buf, sizeof buf, "scene %d: dt must be >= 0", idx);
    msg = buf; return -1;
  }
  return 0;
}

int ensure_stage(snow_ctx* c, size_t bytes) {
  if (c->stage_bytes >= bytes) return 0;
  if (c->d_stage) (void)hipFree(c->d_stage);
  c->d_stage = nullptr;
  c->stage_bytes = 0;
  HIPCHK(c, hipMalloc(&c->d_stage, bytes));
  c->stage_bytes = bytes;
  return 0;
}

// Clears every bound slot (a writer outside the fused step then publishes the bound of the current
// records alone, whatever earlier steps left there).
int clear_bounds(snow_ctx* c) {
  const uint32_t w = uint32_t(c->nscenes) * 2u * PB_SLOTS * PB_STRIDE;
  k_fill_u32<<<grid_for(w, 256), 256, 0, c->stream>>>(c->d_pbound, 0u, int64_t(w));
  LAUNCHCHK(c);
  return 0;
}

// Bound of the current records into the slots of step parity (step + ahead) & 1.
int publish_bounds(snow_ctx* c, uint32_t ahead, int skip_heavy) {
  int rc;
  if ((rc = clear_bounds(c))) return rc;
  if (c->N) {
    k_particle_bounds<<<grid_for(c->N, 256), 256, 0, c->stream>>>(c->d_scenes, c->d_pscene, c->d_rec[c->cur], c->d_fa,
                                                                  c->d_ctl, c->d_pbound, ahead, skip_heavy, c->N);
    LAUNCHCHK(c);
  }
  c->bound_valid = true;
  return 0;
}

int run_keys(snow_ctx* c) {
  if (c->keys_valid) return 0;
  if (c->N) {
    k_keys<<<grid_for(c->N, 256), 256, 0, c->stream>>>(c->d_scenes, c->d_pscene, c->d_rec[c->cur], c->d_keys[0],
                                                       c->N);
    LAUNCHCHK(c);
  }
  c->keys_valid = true;
  c->hist_ready = false;
  return 0;
}

// The first-pass histogram of the current keys in d_hist0 (counted here unless the G2P that wrote
// them did), in the row layout of the current digit width.
int prepare_hist0(snow_ctx* c) {
  if (!c->N) return 0;
  const int bits0 = pass_bits(c->key_bits, c->sort_bits);
  if (bits0 != c->hist_bits) {  // new row layout: start from zeros everywhere
    const int64_t words = int64_t(hist0_words(c->N));
    k_fill_u32<<<grid_for(words, 256), 256, 0, c->stream>>>(c->d_hist0, 0u, words);
    LAUNCHCHK(c);
    c->hist_bits = bits0;
    c->hist_ready = false;
  }
  if (!c->hist_ready) {
    HIPCHK(c, radix_count_first_pass(c->d_keys[0], c->N, c->key_bits, c->sort_bits, c->d_hist0, c->stream,
                                     c->num_cu));
    c->hist_ready = true;
  }
  return 0;
}

int run_force_prepass(snow_ctx* c) {
  if (c->force_valid) return 0;
  int rc;
  if ((rc = clear_bounds(c))) return rc;
  if (c->N) {
    const int k = c->cur;
    k_force_prepass<<<grid_for(c->N, 256), 256, 0, c->stream>>>(c->d_scenes, c->d_pscene, c->d_rec[k], c->d_fe[k],
                                                                 c->d_fp[k], c->d_fa, c->N, c->d_ctl, c->d_pbound);
    LAUNCHCHK(c);
  }
  c->force_valid = true;
  c->bound_valid = true;
  return 0;
}

// Step start (stamp, counters, rigid advance when dt_rigid >= 0) + sort. The step start rides on
// the sort's first launch; it runs on its own only when there is nothing to sort.
int run_sort(snow_ctx* c, float dt_rigid) {
  const StepBegin begin{c->d_ctl,    c->d_counts, c->nscenes * ACTIVE_SLOTS, fp64_counter(c),
                        c->nscenes, c->d_scenes, c->d_cols,                c->d_list_ctr, dt_rigid};
  int rc = run_keys(c);
  if (rc) return rc;
  StageTimer tm(c, ST_SORT);
  if (c->N) {
    if ((rc = prepare_hist0(c))) return rc;
    size_t tb = c->sort_tmp_bytes;
    HIPCHK(c, radix_sort_pairs(c->d_sort_tmp, &tb, c->d_keys[0], c->d_keys[1], nullptr, c->d_slot, c->N,
                               c->key_bits, c->sort_bits, c->stream, c->num_cu, &begin, c->d_hist0));
    c->hist_ready = false;  // consumed: the scatter cleared the rows
  } else {
    k_step_begin<<<1, 256, 0, c->stream>>>(begin);
    LAUNCHCHK(c);
  }
  return tm.done();
}

int run_p2g(snow_ctx* c, int mass_only) {
  if (!c->bound_valid) {  // records no writer published (after the step start: this step's parity)
    const int rc = publish_bounds(c, 0u, mass_only ? 0 : 1);
    if (rc) return rc;
  }
  StageTimer tm(c, ST_P2G);
  if (c->total_chunks) {
    k_p2g_chunk<<<c->total_chunks, 64 * P2G_WAVES, 0, c->stream>>>(
        c->d_scenes, c->d_tile_scene, c->d_keys[1], c->d_slot, c->d_tile_c0, c->d_tile_c1, c->d_tile_stamp,
        c->d_ctl, c->d_list_ctr, c->d_tile_list, c->total_ctiles, c->d_rec[c->cur], c->d_fa, c->N, c->d_tpart,
        c->d_hpart, c->d_chunk_scene, c->d_pbound, c->d_chunk_lo, c->d_chunk_w, c->d_chunk_runs,
        mass_only ? 0 : 1, c->total_chunks <= uint32_t(P2G_BLOCKS_PER_CU * c->num_cu) ? 1 : 0, c->d_fold_job);
    LAUNCHCHK(c);
    k_fold_partials<<<c->total_chunks, 1024, 0, c->stream>>>(c->d_fold_job, c->d_tile_c0, c->d_chunk_lo, c->d_tpart,
                                                            c->d_hpart);
    LAUNCHCHK(c);
  }
  c->bound_valid = false;  // consumed; the slots of the other parity are cleared for the writers
  return tm.done();
}

// plan: the launch also plans the P2G chunk windows of the next step from the runs of the P2G
// just done (the fused step and init do; the per-stage API keeps the current plan, so its calls
// leave later steps unchanged)
int run_grid(snow_ctx* c, bool keep, bool update, float dt, bool plan) {
  StageTimer tm(c, ST_GRID);
  unsigned long long* cnt = c->d_counts;
  if (c->total_chunks == 0) return tm.done();
  const unsigned nplan = plan ? unsigned(c->nscenes) : 0u;
  const unsigned blocks = nplan + std::max(1u, std::min<unsigned>(c->total_ctiles, unsigned(c->num_cu * 4)));
#define GRID_ARGS                                                                                          \
  c->d_scenes, c->nscenes == 1 ? nullptr : c->d_tile_scene, c->d_list_ctr, c->d_tile_list, c->total_ctiles,    \
      c->d_ctl, c->d_tile_stamp,                                                                              \
      c->d_tpart, c->d_zero_part, c->d_g_mmv, c->d_g_f, c->d_g_node, c->d_g_mass, c->d_cols, dt, cnt, nplan,  \
      c->d_chunk_lo, c->d_chunk_runs, c->d_plan_pref, c->d_chunk_w
  if (keep && update)
    k_grid<true, true><<<blocks, GRID_THREADS, 0, c->stream>>>(GRID_ARGS);
  else if (update)
    k_grid<false, true><<<blocks, GRID_THREADS, 0, c->stream>>>(GRID_ARGS);
  else
    k_grid<true, false><<<blocks, GRID_THREADS, 0, c->stream>>>(GRID_ARGS);
#undef GRID_ARGS
  LAUNCHCHK(c);
  return tm.done();
}

int run_g2p_advance(snow_ctx* c, float dt) {
  StageTimer tm(c, ST_G2P);
  const int a = c->cur, b = 1 - c->cur;
  unsigned long long* f64 = fp64_counter(c);
  const bool mixed = (c->flags & SNOW_FLAG_MIXED) != 0;
  const bool comp = (c->flags & SNOW_FLAG_COMPENSATED_G2P) != 0;
  if (c->N) {
#define G2P_ARGS                                                                                          \
  c->d_scenes, c->nscenes > 1 ? c->d_waves : nullptr, c->nwaves, c->d_slot, c->d_rec[a], c->d_fe[a], c->d_fp[a], \
      c->d_pid[a], c->d_rec[b], c->d_fe[b], c->d_fp[b], c->d_pid[b], c->d_fa, c->d_keys[0], c->d_g_node,       \
      c->d_cols, dt, c->N, f64, c->d_ctl, c->d_pbound, c->hist_bits ? c->d_hist0 : nullptr, c->hist_bits
    const unsigned bs = unsigned(c->g2p_block);
    const unsigned g = grid_for(c->nwaves, bs / 64u);
    if (mixed && comp) k_g2p_advance<true, true><<<g, bs, 0, c->stream>>>(G2P_ARGS);
    else if (mixed) k_g2p_advance<true, false><<<g, bs, 0, c->stream>>>(G2P_ARGS);
    else if (comp) k_g2p_advance<false, true><<<g, bs, 0, c->stream>>>(G2P_ARGS);
    else k_g2p_advance<false, false><<<g, bs, 0, c->stream>>>(G2P_ARGS);
#undef G2P_ARGS
    LAUNCHCHK(c);
  }
  c->cur = b;
  c->keys_valid = true;
  c->hist_ready = c->N && c->hist_bits;  // counted into the zeros the sort's scatter left
  c->force_valid = true;
  c->bound_valid = true;
  c->aux_valid = false;
  return tm.done();
}

// One fused step of every scene. dt > 0 overrides the per-scene dt.
// ---- semi-implicit solve ------------------------------------------------------------------
int ensure_implicit(snow_ctx* c) {
  if (c->d_imp_scal) return 0;
  const size_t nn = size_t(std::max<int64_t>(1, c->total_nodes)), np = std::max<size_t>(1, c->N);
  std::vector<ImpBlock> blocks;
  std::vector<uint32_t> first(size_t(c->nscenes) + 1, 0u);
  for (int s = 0; s < c->nscenes; ++s) {
    const SceneDev& S = c->scenes[s];
    first[s] = uint32_t(blocks.size());
    const int64_t nodes = int64_t(S.D[0]) * S.D[1] * S.D[2];
    for (int64_t q = 0; q < nodes; q += IMP_NODES_PER_BLOCK) blocks.push_back(ImpBlock{uint32_t(s), 0u, S.node_off + q});
  }
  first[c->nscenes] = uint32_t(blocks.size());
  c->imp_nblocks = uint32_t(blocks.size());
  auto bad = [&](hipError_t e) {
    if (e == hipSuccess) return false;
    void* ptrs[] = {c->d_imp_vec, c->d_imp_mass, c->d_imp_acc, c->d_imp_pd, c->d_imp_G, c->d_imp_part,
                    c->d_imp_blocks, c->d_imp_scene_blocks, c->d_imp_scal, c->d_imp_gslots,
                    c->d_imp_nlive, c->d_imp_blive};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    if (c->h_imp_scal) (void)hipHostFree(c->h_imp_scal);
    c->d_imp_vec = nullptr; c->d_imp_mass = nullptr; c->d_imp_acc = nullptr; c->d_imp_pd = nullptr;
    c->d_imp_G = nullptr; c->d_imp_part = nullptr; c->d_imp_blocks = nullptr; c->d_imp_scene_blocks = nullptr;
    c->d_imp_scal = nullptr; c->h_imp_scal = nullptr; c->d_imp_gslots = nullptr;
    c->d_imp_nlive = nullptr; c->d_imp_blive = nullptr;
    return true;
  };
  if (bad(dalloc(&c->d_imp_vec, nn * IMP_VECS)) || bad(dalloc(&c->d_imp_mass, nn)) ||
      bad(dalloc(&c->d_imp_acc, 3 * nn)) || bad(dalloc(&c->d_imp_pd, np * IMP_PLANES)) ||
      bad(dalloc(&c->d_imp_G, np * 9)) || bad(dalloc(&c->d_imp_part, 3 * std::max<size_t>(1, blocks.size()))) ||
      bad(dalloc(&c->d_imp_blocks, std::max<size_t>(1, blocks.size()))) ||
      bad(dalloc(&c->d_imp_scene_blocks, first.size())) || bad(dalloc(&c->d_imp_scal, size_t(c->nscenes))) ||
      bad(dalloc(&c->d_imp_gslots, size_t(c->nscenes) * IMP_GSLOTS * IMP_GSTRIDE)) ||
      bad(dalloc(&c->d_imp_nlive, std::max<size_t>(1, c->total_ntiles))) ||
      bad(dalloc(&c->d_imp_blive, std::max<size_t>(1, blocks.size()))) ||
      bad(hipHostMalloc(reinterpret_cast<void**>(&c->h_imp_scal), sizeof(ImpScal) * c->nscenes)))
    return fail(c, SNOW_E_CAPACITY, "semi-implicit solve: no device memory for its node vectors");
  HIPCHK(c, hipMemsetAsync(c->d_imp_acc, 0, sizeof(long long) * 3 * nn, c->stream));
  // vectors and masses start at zero: nodes outside the written tiles are never written by the solve
  HIPCHK(c, hipMemsetAsync(c->d_imp_vec, 0, sizeof(float4) * nn * IMP_VECS, c->stream));
  HIPCHK(c, hipMemsetAsync(c->d_imp_mass, 0, sizeof(float) * nn, c->stream));
  if (!blocks.empty())
    HIPCHK(c, hipMemcpyAsync(c->d_imp_blocks, blocks.data(), sizeof(ImpBlock) * blocks.size(),
                             hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_imp_scene_blocks, first.data(), sizeof(uint32_t) * first.size(),
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

// Conjugate residuals preconditioned by the lumped mass for K v = M v* on the grid the grid update
// just wrote; the node records then hold v instead of v*. Convergence is polled every IMP_POLL
// iterations (one small copy to the host).
int run_implicit(snow_ctx* c, float dt, float beta, float tol, int max_iter) {
  int rc;
  if ((rc = ensure_implicit(c))) return rc;
  constexpr int IMP_POLL = 4;
  const uint32_t nb = c->imp_nblocks;
  const int64_t nn = c->total_nodes;
  const uint2* waves = c->nscenes > 1 ? c->d_waves : nullptr;
  const unsigned pg = grid_for(c->nwaves, 4u);
  const unsigned sg = grid_for(c->nwaves, 4u * IMP_WPW);
  auto scalars = [&](int phase) -> int {
    k_imp_scalars<<<c->nscenes, IMP_THREADS, 0, c->stream>>>(phase, c->d_imp_scene_blocks, nb, c->d_imp_part,
                                                             c->d_imp_scal, c->d_scenes, dt, beta, tol, max_iter,
                                                             c->d_imp_gslots);
    LAUNCHCHK(c);
    return 0;
  };
  // K applied to vector `v` (0 = x, 1 = r) and the node phase that consumes it
  auto apply = [&](int v) -> int {
    if (c->N) {
      k_imp_apply<<<pg, 256, 0, c->stream>>>(c->d_scenes, waves, c->nwaves, c->d_imp_pd, c->N,
                                             c->d_imp_vec + size_t(v) * nn, c->d_imp_G, c->d_imp_scal,
                                             c->d_imp_gslots);
      LAUNCHCHK(c);
      k_imp_scatter<IMP_WPW><<<sg, 256, 0, c->stream>>>(c->d_scenes, waves, c->nwaves, c->d_imp_pd, c->d_imp_G, c->N,
                                               reinterpret_cast<unsigned long long*>(c->d_imp_acc), c->d_imp_scal,
                                               c->d_imp_gslots);
      LAUNCHCHK(c);
    }
    return 0;
  };
#define NODE_ARGS c->d_scenes, c->d_imp_blocks, nb, c->d_imp_mass, c->d_imp_vec, c->d_imp_acc, c->d_imp_scal, nn, \
                  c->d_imp_part, c->d_imp_gslots, c->d_imp_blive
  if ((rc = scalars(-1))) return rc;
  if (c->N) {
    k_imp_prep<<<pg, 256, 0, c->stream>>>(c->d_scenes, waves, c->nwaves, c->d_slot, c->d_rec[c->cur],
                                          c->d_fe[c->cur], c->d_fp[c->cur], c->N, c->d_imp_pd, 1);
    LAUNCHCHK(c);
  }
  if (nb == 0) return 0;
  k_imp_ntiles<<<grid_for(c->total_ntiles, 256), 256, 0, c->stream>>>(
      c->d_scenes, c->nscenes > 1 ? c->d_ntile_scene : nullptr, c->total_ntiles, c->d_ctl, c->d_tile_stamp,
      c->d_imp_nlive);
  LAUNCHCHK(c);
  k_imp_setup<<<nb, IMP_THREADS, 0, c->stream>>>(c->d_scenes, c->d_imp_blocks, c->d_imp_nlive, c->d_g_node,
                                                 c->d_g_mass, c->d_imp_mass, c->d_imp_vec, c->d_imp_acc,
                                                 c->d_imp_blive);
  LAUNCHCHK(c);
  if ((rc = apply(0))) return rc;
  k_imp_node<IMP_INIT><<<nb, IMP_THREADS, 0, c->stream>>>(NODE_ARGS);
  LAUNCHCHK(c);
  if ((rc = scalars(IMP_INIT))) return rc;
  // Polls: before the first update, then where the context's last solve stopped (solves of
  // consecutive steps take about as many iterations), then every IMP_POLL. Updates launched after
  // a scene is done return at once (its flag), so a late poll costs only those launches.
  int next_poll = 0, polled_iters = 0;
  bool finished = false;
  for (int it = 0; it < max_iter; ++it) {
    if (it == next_poll) {
      HIPCHK(c, hipMemcpyAsync(c->h_imp_scal, c->d_imp_scal, sizeof(ImpScal) * c->nscenes, hipMemcpyDeviceToHost,
                               c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      bool all = true;
      polled_iters = 0;
      for (int s = 0; s < c->nscenes; ++s) {
        all = all && c->h_imp_scal[s].done;
        polled_iters = std::max(polled_iters, int(c->h_imp_scal[s].iters));
      }
      if (all) {
        finished = true;
        break;
      }
      next_poll = c->imp_last_iters > it ? c->imp_last_iters : it + IMP_POLL;
    }
    if ((rc = apply(5))) return rc;  // K z
    k_imp_node<IMP_KR><<<nb, IMP_THREADS, 0, c->stream>>>(NODE_ARGS);
    LAUNCHCHK(c);
    if ((rc = scalars(IMP_KR))) return rc;
    k_imp_node<IMP_PX><<<nb, IMP_THREADS, 0, c->stream>>>(NODE_ARGS);
    LAUNCHCHK(c);
    if ((rc = scalars(IMP_PX))) return rc;
  }
#undef NODE_ARGS
  if (!finished) {  // the loop ran out: read where the scenes stopped
    HIPCHK(c, hipMemcpyAsync(c->h_imp_scal, c->d_imp_scal, sizeof(ImpScal) * c->nscenes, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    polled_iters = 0;
    for (int s = 0; s < c->nscenes; ++s) polled_iters = std::max(polled_iters, int(c->h_imp_scal[s].iters));
  }
  c->imp_last_iters = polled_iters;
  k_imp_finish<<<nb, IMP_THREADS, 0, c->stream>>>(c->d_scenes, c->d_imp_blocks, c->d_imp_mass, c->d_imp_vec,
                                                  c->d_g_node, c->d_g_mass, c->d_imp_blive);
  LAUNCHCHK(c);
  return 0;
}

// The explicit step with the semi-implicit grid velocity solve between the grid update and G2P.
int step_semi_implicit_impl(snow_ctx* c, float dt, float beta, float tol, int max_iter) {
  if (!c->initialized) return fail(c, SNOW_E_STATE, "snow_init must run before stepping");
  int rc;
  // the solver's memory first, so that a failed allocation leaves the step undone
  if ((rc = ensure_implicit(c))) return rc;
  if ((rc = run_force_prepass(c))) return rc;
  if ((rc = run_sort(c, dt))) return rc;
  if ((rc = run_p2g(c, 0))) return rc;
  const bool keep = (c->flags & SNOW_FLAG_KEEP_GRID) != 0;
  if ((rc = run_grid(c, keep, true, dt, true))) return rc;
  if ((rc = run_implicit(c, dt, beta, tol, max_iter))) return rc;
  if ((rc = run_g2p_advance(c, dt))) return rc;
  c->grid_state = 1;
  c->keep_valid = keep;
  return 0;
}

int step_impl(snow_ctx* c, float dt) {
  if (!c->initialized) return fail(c, SNOW_E_STATE, "snow_init must run before stepping");
  int rc;
  if ((rc = run_force_prepass(c))) return rc;
  if ((rc = run_sort(c, dt))) return rc;
  if ((rc = run_p2g(c, 0))) return rc;
  const bool keep = (c->flags & SNOW_FLAG_KEEP_GRID) != 0;
  if ((rc = run_grid(c, keep, true, dt, true))) return rc;
  if ((rc = run_g2p_advance(c, dt))) return rc;
  c->grid_state = 1;
  c->keep_valid = keep;
  return 0;
}

void release(snow_ctx* c) {
  if (!c) return;
  for (int p = 0; p < 2; ++p) {
    if (c->gexec[p]) (void)hipGraphExecDestroy(c->gexec[p]);
    if (c->graph[p]) (void)hipGraphDestroy(c->graph[p]);
  }
  void* ptrs[] = {c->d_scenes, c->d_pscene, c->d_tile_scene, c->d_ntile_scene, c->d_rec[0], c->d_rec[1],
                  c->d_fe[0], c->d_fe[1], c->d_fp[0], c->d_fp[1], c->d_fa, c->d_pid[0],
                  c->d_pid[1], c->d_keys[0], c->d_keys[1], c->d_slot, c->d_sort_tmp, c->d_hist0,
                  c->d_tile_c0, c->d_tile_c1, c->d_tpart, c->d_hpart, c->d_zero_part, c->d_list_ctr, c->d_tile_list, c->d_chunk_scene, c->d_chunk_lo, c->d_chunk_w, c->d_chunk_runs, c->d_fold_job, c->d_plan_pref, c->d_waves, c->d_ctl, c->d_pbound, c->d_tile_stamp, c->d_g_mmv, c->d_g_f, c->d_g_node, c->d_g_mass,
                  c->d_cols, c->d_counts, c->d_bad, c->d_aux, c->d_stage,
                  c->d_imp_vec, c->d_imp_mass, c->d_imp_acc, c->d_imp_pd, c->d_imp_G, c->d_imp_part,
                  c->d_imp_blocks, c->d_imp_scene_blocks, c->d_imp_scal, c->d_imp_gslots, c->d_imp_nlive,
                  c->d_imp_blive};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (c->h_imp_scal) (void)hipHostFree(c->h_imp_scal);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& p : c->pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : c->ev_free) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

}  // namespace

// =========================================================================================
extern "C" {

void snow_material_defaults(snow_material* m) {
  if (!m) return;
  m->young = 6e6f;        // YOUNG_MODULUS (defines.hpp:50)
  m->poisson = 0.2f;      // POISSON
  m->hardening = 30.0f;   // HARDENING
  m->crit_comp = 1.0e-2f; // CRIT_COMPRESSION
  m->crit_stretch = 1e-3f;// CRIT_STRETCH
  m->flip_alpha = 0.95f;  // g2pTransfer.glsl:2
  m->gravity[0] = 0.0f; m->gravity[1] = -9.81f; m->gravity[2] = 0.0f;  // gForceCompute.glsl:39
  m->grad_scale = 1.0f;   // cubic.include.glsl:25-40 has no 1/h
}

void snow_grid_defaults(snow_grid* g) {
  if (!g) return;
  g->dim[0] = g->dim[1] = g->dim[2] = 201;  // GRID_DIM_* (defines.hpp:17-19)
  g->origin[0] = 0.5125f; g->origin[1] = 0.5125f; g->origin[2] = 1.5125f;  // GRID_POS_*
  g->spacing = 0.05f;  // GRID_SPACING
}

int snow_create(const snow_scene_desc* scene, uint32_t flags, snow_ctx** out) {
  return snow_create_batch(scene, 1, flags, out);
}

int snow_create_batch(const snow_scene_desc* scenes, int32_t num_scenes, uint32_t flags, snow_ctx** out) {
  if (!out) return fail(nullptr, SNOW_E_ARG, "out is NULL");
  *out = nullptr;
  if (!scenes || num_scenes <= 0 || num_scenes > 65535)
    return fail(nullptr, SNOW_E_ARG, "need 1..65535 scene descriptors");
  std::string msg;
  int64_t ntot = 0;
  for (int s = 0; s < num_scenes; ++s) {
    if (validate_scene(scenes[s], s, msg)) return fail(nullptr, SNOW_E_ARG, msg);
    ntot += scenes[s].num_particles;
  }
  // SoA planes are addressed with 32-bit byte offsets in the kernels
  if (ntot >= (int64_t(1) << 30)) return fail(nullptr, SNOW_E_CAPACITY, "2^30 or more particles in a batch");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(nullptr, SNOW_E_NODEVICE, "no HIP device available");

  snow_ctx* c = new snow_ctx();
  c->flags = flags;
  c->nscenes = num_scenes;
  c->N = uint32_t(ntot);
  (void)hipGetDevice(&c->device);
  {
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device) == hipSuccess && ncu > 0)
      c->num_cu = ncu;
  }

  // scene table, tile index space
  uint64_t ct_off = 0, nt_off = 0, chunk_off = 0;
  int64_t node_off = 0, p_off = 0;
  int col_off = 0;
  for (int s = 0; s < num_scenes; ++s) {
    const snow_scene_desc& d = scenes[s];
    SceneDev S;
    std::memset(&S, 0, sizeof S);
    for (int k = 0; k < 3; ++k) {
      S.D[k] = d.grid.dim[k];
      S.origin[k] = d.grid.origin[k];
      S.ct[k] = (d.grid.dim[k] + 3 + TILE - 1) / TILE;
      S.nt[k] = (d.grid.dim[k] + TILE - 1) / TILE;
      S.g[k] = d.material.gravity[k];
    }
    S.h = d.grid.spacing;
    S.young = d.material.young; S.poisson = d.material.poisson; S.hardening = d.material.hardening;
    S.crit_comp = d.material.crit_comp; S.crit_stretch = d.material.crit_stretch;
    S.alpha = d.material.flip_alpha;
    S.gscale = d.material.grad_scale > 0.0f ? d.material.grad_scale : 1.0f / d.grid.spacing;
    S.dt = d.dt;
    S.ncol = d.num_colliders;
    S.col_off = col_off;
    S.ct_off = uint32_t(ct_off);
    const uint64_t nct = uint64_t(S.ct[0]) * S.ct[1] * S.ct[2];
    S.dead_tile = uint32_t(ct_off + nct);
    S.nt_off = uint32_t(nt_off);
    S.node_off = node_off;
    S.p_off = p_off;
    S.p_cnt = d.num_particles;
    S.inv_h = 1.0f / d.grid.spacing;
    for (int k = 0; k < 3; ++k) S.uo[k] = -d.grid.origin[k] / d.grid.spacing;
    // P2G chunks: the mean chunk is a power of two from 256 to P2G_SEG_MEAN giving at least 64
    // chunks where it can (small scenes alone still fill many workgroups, big ones keep few tile
    // splits); the windows are planned from the previous P2G's runs, up to P2G_SEG_MAX - align_w
    // long. The chunking depends on the scene's own particles only.
    {
      uint32_t seg = 256;
      while (seg < uint32_t(P2G_SEG_MEAN) && uint64_t(seg) * 2 * 64 <= uint64_t(d.num_particles)) seg *= 2;
      S.seg = seg;
      // bounds move to the first tile start within align_w positions of the planned window
      S.align_w = std::min(seg, 256u);
      S.chunk_off = uint32_t(chunk_off);
      chunk_off += (uint64_t(d.num_particles) + seg - 1) / seg;
    }
    S.mu0 = d.material.young / (2.0f * (1.0f + d.material.poisson));
    S.lambda0 = d.material.young * d.material.poisson / ((1.0f + d.material.poisson) * (1.0f - 2.0f * d.material.poisson));
    ct_off += nct + 1;
    nt_off += uint64_t(S.nt[0]) * S.nt[1] * S.nt[2];
    node_off += int64_t(S.D[0]) * S.D[1] * S.D[2];
    p_off += d.num_particles;
    col_off += d.num_colliders;
    c->scenes.push_back(S);
    c->sh.push_back(SceneHost{d.num_particles, {S.D[0], S.D[1], S.D[2]}, d.num_colliders});
    for (int i = 0; i < d.num_colliders; ++i) {
      Collider C;
      std::memcpy(&C, &d.colliders[i], sizeof C);
      c->cols.push_back(C);
    }
  }
  if (ct_off * TILE_CELLS >= (uint64_t(1) << 32)) {
    release(c);
    return fail(nullptr, SNOW_E_CAPACITY, "batch grid too large for 32-bit cell keys");
  }
  c->total_ctiles = uint32_t(ct_off);
  c->total_chunks = uint32_t(chunk_off);
  c->total_ntiles = uint32_t(nt_off);
  c->total_nodes = node_off;
  {
    uint64_t maxkey = ct_off * TILE_CELLS - 1;
    int bits = 1;
    while (bits < 32 && (maxkey >> bits)) ++bits;
    c->key_bits = bits;
  }

#define CREATECHK(expr)                                                                          \
  do {                                                                                           \
    hipError_t _e = (expr);                                                                      \
    if (_e != hipSuccess) {                                                                      \
      std::string m_ = std::string(#expr) + ": " + hipGetErrorString(_e);                       \
      release(c);                                                                                \
      return fail(nullptr, _e == hipErrorOutOfMemory ? SNOW_E_CAPACITY : SNOW_E_HIP, m_);        \
    }                                                                                            \
  } while (0)

  CREATECHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  CREATECHK(scratch_address(c->device, &c->noalloc_stage));
  CREATECHK(hipEventCreate(&c->ev[0]));
  CREATECHK(hipEventCreate(&c->ev[1]));
  const uint32_t N = c->N;
  CREATECHK(dalloc(&c->d_scenes, c->scenes.size()));
  CREATECHK(dalloc(&c->d_pscene, N));
  CREATECHK(dalloc(&c->d_tile_scene, c->total_ctiles));
  CREATECHK(dalloc(&c->d_ntile_scene, std::max<uint32_t>(1, c->total_ntiles)));
  for (int b = 0; b < 2; ++b) {
    CREATECHK(dalloc(&c->d_rec[b], size_t(2) * N));
    CREATECHK(dalloc(&c->d_fe[b], size_t(9) * N));
    CREATECHK(dalloc(&c->d_fp[b], size_t(9) * N));
    CREATECHK(dalloc(&c->d_pid[b], N));
    CREATECHK(dalloc(&c->d_keys[b], N));
  }
  CREATECHK(dalloc(&c->d_fa, size_t(9) * N));
  CREATECHK(dalloc(&c->d_slot, N));
  {
    size_t tb = 0;
    CREATECHK(radix_sort_pairs(nullptr, &tb, c->d_keys[0], c->d_keys[1], nullptr, c->d_slot,
                               std::max<uint32_t>(N, 1), c->key_bits, 11, c->stream, c->num_cu));
    c->sort_tmp_bytes = tb;
    CREATECHK(hipMalloc(&c->d_sort_tmp, std::max<size_t>(tb, 16)));
    CREATECHK(dalloc(&c->d_hist0, hist0_words(std::max<uint32_t>(N, 1))));
    CREATECHK(hipMemset(c->d_hist0, 0, sizeof(uint32_t) * hist0_words(std::max<uint32_t>(N, 1))));
  }
  CREATECHK(dalloc(&c->d_tile_c0, c->total_ctiles));
  CREATECHK(dalloc(&c->d_tile_c1, c->total_ctiles));
  CREATECHK(dalloc(&c->d_tpart, size_t(c->total_ctiles) * NQ * HALO_NODES));
  CREATECHK(dalloc(&c->d_hpart, size_t(std::max<uint32_t>(1, c->total_chunks)) * NQ * HALO_NODES));
  CREATECHK(dalloc(&c->d_chunk_scene, std::max<uint32_t>(1, c->total_chunks)));
  CREATECHK(dalloc(&c->d_chunk_lo, std::max<uint32_t>(1, c->total_chunks)));
  CREATECHK(dalloc(&c->d_chunk_w, std::max<uint32_t>(1, c->total_chunks)));
  CREATECHK(dalloc(&c->d_chunk_runs, std::max<uint32_t>(1, c->total_chunks)));
  CREATECHK(dalloc(&c->d_fold_job, std::max<uint32_t>(1, c->total_chunks)));
  CREATECHK(dalloc(&c->d_plan_pref, std::max<uint32_t>(1, c->total_chunks)));
  {
    std::vector<uint2> waves;
    for (int s = 0; s < num_scenes; ++s) {
      const SceneDev& S = c->scenes[s];
      for (int64_t q = 0; q < S.p_cnt; q += 64) waves.push_back(make_uint2(uint32_t(S.p_off + q), uint32_t(s)));
    }
    c->nwaves = uint32_t(waves.size());
    CREATECHK(dalloc(&c->d_waves, std::max<size_t>(1, waves.size())));
    if (!waves.empty())
      CREATECHK(hipMemcpy(c->d_waves, waves.data(), sizeof(uint2) * waves.size(), hipMemcpyHostToDevice));
  }
  CREATECHK(dalloc(&c->d_zero_part, size_t(NQ) * HALO_NODES));
  CREATECHK(dalloc(&c->d_list_ctr, size_t(TILE_LISTS) * TILE_LIST_STRIDE));
  CREATECHK(hipMemset(c->d_list_ctr, 0, sizeof(uint32_t) * TILE_LISTS * TILE_LIST_STRIDE));
  CREATECHK(dalloc(&c->d_tile_list, size_t(TILE_LISTS) * c->total_ctiles));
  CREATECHK(hipMemset(c->d_zero_part, 0, sizeof(float) * NQ * HALO_NODES));
  CREATECHK(dalloc(&c->d_ctl, 1));
  CREATECHK(dalloc(&c->d_pbound, size_t(c->nscenes) * 2 * PB_SLOTS * PB_STRIDE));
  CREATECHK(dalloc(&c->d_tile_stamp, c->total_ctiles));
  CREATECHK(dalloc(&c->d_g_mmv, size_t(c->total_nodes)));
  CREATECHK(dalloc(&c->d_g_f, size_t(c->total_nodes)));
  CREATECHK(dalloc(&c->d_g_node, 3 * size_t(c->total_nodes)));
  CREATECHK(dalloc(&c->d_g_mass, size_t(c->total_nodes)));
  CREATECHK(dalloc(&c->d_cols, std::max<size_t>(1, c->cols.size())));
  CREATECHK(dalloc(&c->d_counts, size_t(c->nscenes) * ACTIVE_SLOTS * ACTIVE_STRIDE + 1));
  CREATECHK(dalloc(&c->d_bad, 1));

  // tables
  {
    std::vector<uint16_t> ps(N), ts(c->total_ctiles), nts(std::max<uint32_t>(1, c->total_ntiles));
    for (int s = 0; s < num_scenes; ++s) {
      const SceneDev& S = c->scenes[s];
      for (int64_t i = 0; i < S.p_cnt; ++i) ps[size_t(S.p_off + i)] = uint16_t(s);
      const uint32_t nct = uint32_t(S.ct[0] * S.ct[1] * S.ct[2]) + 1;
      for (uint32_t t = 0; t < nct; ++t) ts[S.ct_off + t] = uint16_t(s);
      const uint32_t nnt = uint32_t(S.nt[0] * S.nt[1] * S.nt[2]);
      for (uint32_t t = 0; t < nnt; ++t) nts[S.nt_off + t] = uint16_t(s);
    }
    CREATECHK(hipMemcpy(c->d_scenes, c->scenes.data(), sizeof(SceneDev) * c->scenes.size(), hipMemcpyHostToDevice));
    CREATECHK(hipMemcpy(c->d_pscene, ps.data(), sizeof(uint16_t) * N, hipMemcpyHostToDevice));
    CREATECHK(hipMemcpy(c->d_tile_scene, ts.data(), sizeof(uint16_t) * ts.size(), hipMemcpyHostToDevice));
    std::vector<uint16_t> cs(std::max<uint32_t>(1, c->total_chunks), 0);
    std::vector<uint32_t> cw(std::max<uint32_t>(1, c->total_chunks), 0);
    for (int s = 0; s < num_scenes; ++s) {
      const SceneDev& S = c->scenes[s];
      const uint32_t nch = uint32_t((uint64_t(S.p_cnt) + S.seg - 1) / S.seg);
      for (uint32_t k = 0; k < nch; ++k) {
        cs[S.chunk_off + k] = uint16_t(s);
        cw[S.chunk_off + k] = uint32_t(S.p_off) + k * S.seg;  // until the first plan: even windows
      }
    }
    CREATECHK(hipMemcpy(c->d_chunk_scene, cs.data(), sizeof(uint16_t) * cs.size(), hipMemcpyHostToDevice));
    CREATECHK(hipMemcpy(c->d_chunk_w, cw.data(), sizeof(uint32_t) * cw.size(), hipMemcpyHostToDevice));
    CREATECHK(hipMemcpy(c->d_ntile_scene, nts.data(), sizeof(uint16_t) * nts.size(), hipMemcpyHostToDevice));
    if (!c->cols.empty())
      CREATECHK(hipMemcpy(c->d_cols, c->cols.data(), sizeof(Collider) * c->cols.size(), hipMemcpyHostToDevice));
  }
  // particle upload (creation order = slot order, pid = identity; rest density 0 until snow_init)
  {
    std::vector<float4> rec(size_t(2) * N);
    std::vector<float> fe(size_t(9) * N), fp(size_t(9) * N);
    std::vector<uint32_t> pid(N);
    for (int s = 0; s < num_scenes; ++s) {
      const snow_scene_desc& d = scenes[s];
      const int64_t off = c->scenes[s].p_off;
      for (int64_t k = 0; k < d.num_particles; ++k) {
        const size_t i = size_t(off + k);
        rec[2 * i] = make_float4(d.x[3 * k], d.x[3 * k + 1], d.x[3 * k + 2], d.mass[k]);
        rec[2 * i + 1] = make_float4(d.v[3 * k], d.v[3 * k + 1], d.v[3 * k + 2], 0.f);
        for (int e = 0; e < 9; ++e) {
          const float id = (e == 0 || e == 4 || e == 8) ? 1.0f : 0.0f;
          fe[size_t(e) * N + i] = d.fe ? d.fe[9 * k + e] : id;
          fp[size_t(e) * N + i] = d.fp ? d.fp[9 * k + e] : id;
        }
        pid[i] = uint32_t(i);
      }
    }
    if (N) {
      CREATECHK(hipMemcpy(c->d_rec[0], rec.data(), sizeof(float4) * rec.size(), hipMemcpyHostToDevice));
      CREATECHK(hipMemcpy(c->d_fe[0], fe.data(), sizeof(float) * fe.size(), hipMemcpyHostToDevice));
      CREATECHK(hipMemcpy(c->d_fp[0], fp.data(), sizeof(float) * fp.size(), hipMemcpyHostToDevice));
      CREATECHK(hipMemcpy(c->d_pid[0], pid.data(), sizeof(uint32_t) * N, hipMemcpyHostToDevice));
      CREATECHK(hipMemset(c->d_fa, 0, sizeof(float) * 9 * size_t(N)));
    }
  }
  CREATECHK(hipMemset(c->d_tile_c0, 0, sizeof(uint32_t) * c->total_ctiles));
  CREATECHK(hipMemset(c->d_tile_c1, 0, sizeof(uint32_t) * c->total_ctiles));
  CREATECHK(hipMemset(c->d_ctl, 0, sizeof(StepCtl)));
  CREATECHK(hipMemset(c->d_pbound, 0, sizeof(uint32_t) * c->nscenes * 2 * PB_SLOTS * PB_STRIDE));
  CREATECHK(hipMemset(c->d_tile_stamp, 0, sizeof(uint32_t) * c->total_ctiles));
  CREATECHK(hipMemset(c->d_g_mmv, 0, sizeof(float4) * std::max<int64_t>(1, c->total_nodes)));
  CREATECHK(hipMemset(c->d_g_f, 0, sizeof(float4) * std::max<int64_t>(1, c->total_nodes)));
  CREATECHK(hipMemset(c->d_g_node, 0, 3 * sizeof(float2) * std::max<int64_t>(1, c->total_nodes)));
  CREATECHK(hipMemset(c->d_g_mass, 0, sizeof(float) * std::max<int64_t>(1, c->total_nodes)));
  CREATECHK(hipMemset(c->d_counts, 0, sizeof(unsigned long long) * (size_t(c->nscenes) * ACTIVE_SLOTS * ACTIVE_STRIDE + 1)));
  CREATECHK(hipDeviceSynchronize());
#undef CREATECHK
  c->cur = 0;
  c->keys_valid = false;
  c->force_valid = false;  // rho0 not known yet; A = 0 until snow_init
  *out = c;
  return SNOW_OK;
}

int snow_destroy(snow_ctx* c) {
  if (!c) return SNOW_E_ARG;
  (void)hipStreamSynchronize(c->stream);
  release(c);
  return SNOW_OK;
}

const char* snow_last_error(const snow_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }
int snow_num_scenes(const snow_ctx* c) { return c ? c->nscenes : SNOW_E_ARG; }
int64_t snow_num_particles(const snow_ctx* c, int32_t s) {
  if (!c || s < 0 || s >= c->nscenes) return SNOW_E_ARG;
  return c->sh[s].n;
}

int snow_stage_grid_mass(snow_ctx* c) {
  if (!c) return SNOW_E_ARG;
  int rc;
  // A = 0 during the mass pass: zero the force part of the records (gMassCompute only adds m)
  if ((rc = clear_bounds(c))) return rc;
  if (c->N) {
    k_zero_force<<<grid_for(c->N, 256), 256, 0, c->stream>>>(c->d_scenes, c->d_pscene, c->d_rec[c->cur], c->d_fa,
                                                             c->N, c->d_ctl, c->d_pbound);
    LAUNCHCHK(c);
  }
  c->bound_valid = true;
  c->force_valid = false;
  if ((rc = run_sort(c, -1.0f))) return rc;
  if ((rc = run_p2g(c, 1))) return rc;
  if ((rc = run_grid(c, true, false, 0.0f, true))) return rc;
  c->grid_state = 1;
  c->keep_valid = true;
  return 0;
}

int snow_stage_particle_volume(snow_ctx* c) {
  if (!c) return SNOW_E_ARG;
  if (!(c->grid_state == 1 && c->keep_valid) && c->grid_state != 2)
    return fail(c, SNOW_E_STATE, "particle volume needs grid masses (run snow_stage_grid_mass)");
  if (c->N) {
    k_particle_volume<<<grid_for(c->N, 256), 256, 0, c->stream>>>(c->d_scenes, c->d_pscene, c->d_rec[c->cur],
                                                                   c->d_g_mmv, c->N);
    LAUNCHCHK(c);
  }
  c->force_valid = false;
  return 0;
}

int snow_init(snow_ctx* c) {
  if (!c) return SNOW_E_ARG;
  int rc;
  if ((rc = snow_stage_grid_mass(c))) return rc;
  if ((rc = snow_stage_particle_volume(c))) return rc;
  if ((rc = run_force_prepass(c))) return rc;
  c->initialized = true;
  return 0;
}

int snow_step(snow_ctx* c, double dt) {
  if (!c) return SNOW_E_ARG;
  const float fdt = float(dt);  // technique.cpp:122-125 narrows the uniform to float
  if (!(fdt > 0.0f) || !std::isfinite(fdt)) return fail(c, SNOW_E_ARG, "dt must be > 0");
  return step_impl(c, fdt);
}

int snow_step_semi_implicit(snow_ctx* c, double dt, double beta, double tol, int32_t max_iter) {
  if (!c) return SNOW_E_ARG;
  const float fdt = float(dt);
  if (!(fdt >= 0.0f) || !std::isfinite(fdt)) return fail(c, SNOW_E_ARG, "dt must be >= 0");
  if (fdt == 0.0f)
    for (const auto& S : c->scenes)
      if (!(S.dt > 0.0f)) return fail(c, SNOW_E_ARG, "dt = 0 takes each scene's dt, which must be > 0");
  if (!(beta >= 0.0) || !(beta <= 1.0)) return fail(c, SNOW_E_ARG, "beta must lie in [0, 1]");
  if (!(tol > 0.0) || !std::isfinite(tol)) return fail(c, SNOW_E_ARG, "tol must be > 0");
  if (max_iter < 0) return fail(c, SNOW_E_ARG, "max_iter must be >= 0");
  for (const auto& S : c->scenes)
    if (S.D[0] > IMP_MAX_DIM || S.D[1] > IMP_MAX_DIM || S.D[2] > IMP_MAX_DIM)
      return fail(c, SNOW_E_ARG, "the semi-implicit step takes grids of at most 1021 nodes per axis");
  return step_semi_implicit_impl(c, fdt, float(beta), float(tol), int(max_iter));
}

int snow_implicit_stats(snow_ctx* c, int32_t scene, int32_t* iterations, double* rel_residual) {
  if (!c || scene < 0 || scene >= c->nscenes) return SNOW_E_ARG;
  if (!c->d_imp_scal) return fail(c, SNOW_E_STATE, "no semi-implicit step has run on this context");
  ImpScal z;
  HIPCHK(c, hipMemcpyAsync(&z, c->d_imp_scal + scene, sizeof(ImpScal), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (iterations) *iterations = z.iters;
  if (rel_residual) *rel_residual = z.bb > 0.0 ? std::sqrt(z.rr / z.bb) : 0.0;
  return 0;
}

int snow_step_batch(snow_ctx* c) {
  if (!c) return SNOW_E_ARG;
  for (const auto& S : c->scenes)
    if (!(S.dt > 0.0f)) return fail(c, SNOW_E_ARG, "every scene needs dt > 0 for snow_step_batch");
  return step_impl(c, 0.0f);
}

int snow_run_graph(snow_ctx* c, int32_t steps) {
  if (!c || steps < 0) return SNOW_E_ARG;
  if (!c->initialized) return fail(c, SNOW_E_STATE, "snow_init must run before stepping");
  for (const auto& S : c->scenes)
    if (!(S.dt > 0.0f)) return fail(c, SNOW_E_ARG, "every scene needs dt > 0 for snow_run_graph");
  int rc;
  // prerequisites of the captured step, run eagerly: force matrices, cell keys and the P2G bound
  // of the state (the next step start makes parity step + 1 the current one)
  if ((rc = run_force_prepass(c))) return rc;
  if ((rc = run_keys(c))) return rc;
  if ((rc = prepare_hist0(c))) return rc;  // the captured sorts take counted histograms
  if (!c->bound_valid && (rc = publish_bounds(c, 1u, 1))) return rc;
  if (!c->gexec[0]) {
    // One graph per buffer parity: the step reads buffers `cur` and writes `1 - cur`. Every launch
    // in it has fixed arguments (per-step counters live in the device control block), and the
    // captured kernels include no memset node.
    const bool timers = c->timers;
    const int cur0 = c->cur;
    c->timers = false;
    for (int p = 0; p < 2; ++p) {
      c->cur = p;
      HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      rc = run_sort(c, 0.0f);
      if (!rc) rc = run_p2g(c, 0);
      if (!rc) rc = run_grid(c, (c->flags & SNOW_FLAG_KEEP_GRID) != 0, true, 0.0f, true);
      if (!rc) rc = run_g2p_advance(c, 0.0f);
      hipGraph_t g = nullptr;
      const hipError_t ec = hipStreamEndCapture(c->stream, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        c->timers = timers;
        c->cur = cur0;
        return rc;
      }
      if (ec != hipSuccess) {
        c->timers = timers;
        c->cur = cur0;
        return fail(c, SNOW_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
      }
      c->graph[p] = g;
      HIPCHK(c, hipGraphInstantiate(&c->gexec[p], g, nullptr, nullptr, 0));
    }
    c->timers = timers;
    c->cur = cur0;
    c->keys_valid = true;
    c->force_valid = true;
  }
  for (int32_t i = 0; i < steps; ++i) {
    HIPCHK(c, hipGraphLaunch(c->gexec[c->cur], c->stream));
    c->cur = 1 - c->cur;
  }
  if (steps > 0) {
    c->keys_valid = true;
    c->force_valid = true;
    c->bound_valid = true;
    c->aux_valid = false;
    c->grid_state = 1;
    c->keep_valid = (c->flags & SNOW_FLAG_KEEP_GRID) != 0;
  }
  return 0;
}

int snow_autotune(snow_ctx* c, int32_t steps, int32_t* g2p_block, int32_t* p2g_chunk, int32_t* grid_blocks,
                  int32_t* sort_bits) {
  if (!c || steps <= 0) return SNOW_E_ARG;
  if (!c->initialized) return fail(c, SNOW_E_STATE, "snow_init must run before autotuning");
  for (const auto& S : c->scenes)
    if (!(S.dt > 0.0f)) return fail(c, SNOW_E_ARG, "every scene needs dt > 0 for snow_autotune");
  int rc;
  if ((rc = run_force_prepass(c))) return rc;
  if ((rc = run_keys(c))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  // snapshot of everything a step reads and writes
  const size_t N = std::max<size_t>(1, c->N);
  const int k = c->cur;
  struct Buf { void* dst; const void* src; size_t bytes; };
  std::vector<Buf> bufs = {
      {nullptr, c->d_rec[k], sizeof(float4) * 2 * N}, {nullptr, c->d_fe[k], sizeof(float) * 9 * N},
      {nullptr, c->d_fp[k], sizeof(float) * 9 * N},  {nullptr, c->d_fa, sizeof(float) * 9 * N},
      {nullptr, c->d_pid[k], sizeof(uint32_t) * N},   {nullptr, c->d_keys[0], sizeof(uint32_t) * N},
      {nullptr, c->d_ctl, sizeof(StepCtl)},
      {nullptr, c->d_chunk_w, sizeof(uint32_t) * std::max<uint32_t>(1, c->total_chunks)},
      {nullptr, c->d_cols, sizeof(Collider) * std::max<size_t>(1, c->cols.size())}};
  auto free_all = [&]() {
    for (auto& b : bufs)
      if (b.dst) (void)hipFree(b.dst);
  };
  for (auto& b : bufs) {
    if (hipMalloc(&b.dst, b.bytes) != hipSuccess) {
      free_all();
      return fail(c, SNOW_E_CAPACITY, "autotune: no device memory for the state snapshot");
    }
    HIPCHK(c, hipMemcpyAsync(b.dst, b.src, b.bytes, hipMemcpyDeviceToDevice, c->stream));
  }
  auto restore = [&]() -> int {
    for (auto& b : bufs) HIPCHK(c, hipMemcpyAsync(const_cast<void*>(b.src), b.dst, b.bytes, hipMemcpyDeviceToDevice, c->stream));
    // The restored control block rewinds the step stamp, so a later step can carry the number a
    // trial already stamped tiles with: clear the stamps (0 is never a step's stamp).
    k_fill_u32<<<grid_for(c->total_ctiles, 256), 256, 0, c->stream>>>(c->d_tile_stamp, 0u, int64_t(c->total_ctiles));
    LAUNCHCHK(c);
    c->cur = k;
    c->keys_valid = true;
    c->hist_ready = false;  // the trials' G2P counted their own keys
    c->force_valid = true;
    c->bound_valid = false;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
  };
  const bool timers = c->timers;
  c->timers = false;
  auto trial = [&](double& ms) -> int {
    int r = restore();
    if (r) return r;
    HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    for (int32_t i = 0; i < steps; ++i)
      if ((r = step_impl(c, 0.0f))) return r;
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev[1]));
    float f = 0.f;
    HIPCHK(c, hipEventElapsedTime(&f, c->ev[0], c->ev[1]));
    ms = f;
    return 0;
  };
  // the G2P block size is the launch shape that leaves the results unchanged; the P2G and grid
  // widths follow from the particle counts
  const int g2p_cands[] = {64, 128, 256};
  int best_g = c->g2p_block;
  double best = 1e30, t = 0.0;
  for (int g : g2p_cands) {
    c->g2p_block = g;
    if ((rc = trial(t))) break;
    if (t < best) { best = t; best_g = g; }
  }
  c->g2p_block = best_g;
  // the sort-bin width: the largest radix digit (the stable sort's result does not depend on it)
  const int sort_cands[] = {8, 9, 10, 11};
  int best_s = c->sort_bits;
  best = 1e30;
  for (int sb : sort_cands) {
    if (rc) break;
    c->sort_bits = sb;
    if ((rc = trial(t))) break;
    if (t < best) { best = t; best_s = sb; }
  }
  c->sort_bits = best_s;
  c->timers = timers;
  const int rr = restore();
  free_all();
  if (rc) return rc;
  if (rr) return rr;
  // captured graphs hold the old launch shapes
  for (int p = 0; p < 2; ++p) {
    if (c->gexec[p]) (void)hipGraphExecDestroy(c->gexec[p]);
    if (c->graph[p]) (void)hipGraphDestroy(c->graph[p]);
    c->gexec[p] = nullptr;
    c->graph[p] = nullptr;
  }
  if (g2p_block) *g2p_block = c->g2p_block;
  if (p2g_chunk) *p2g_chunk = c->scenes.empty() ? 0 : int32_t(c->scenes[0].seg);
  if (grid_blocks) *grid_blocks = int32_t(c->total_chunks);
  if (sort_bits) *sort_bits = c->sort_bits;
  return 0;
}

int snow_stage_rigid_advance(snow_ctx* c, double dt) {
  if (!c) return SNOW_E_ARG;
  StageTimer tm(c, ST_RIGID);
  k_rigid_advance<<<1, 256, 0, c->stream>>>(c->d_scenes, c->nscenes, c->d_cols, float(dt));
  LAUNCHCHK(c);
  return tm.done();
}

int snow_stage_grid_reset(snow_ctx* c) {
  if (!c) return SNOW_E_ARG;
  StageTimer tm(c, ST_RESET);
  // g_mmv, g_f: 4 words per node; g_node: 6 words per node; g_mass: 1 word per node
  const int64_t nodes = c->total_nodes;
  for (int k = 0; k < 4; ++k) {
    uint32_t* p = k == 0 ? reinterpret_cast<uint32_t*>(c->d_g_mmv)
                         : k == 1 ? reinterpret_cast<uint32_t*>(c->d_g_f)
                                  : k == 2 ? reinterpret_cast<uint32_t*>(c->d_g_node)
                                           : reinterpret_cast<uint32_t*>(c->d_g_mass);
    const int64_t w = nodes * (k < 2 ? 4 : (k == 2 ? 6 : 1));
    if (w) {
      k_fill_u32<<<grid_for(w, 256), 256, 0, c->stream>>>(p, 0u, w);
      LAUNCHCHK(c);
    }
  }
  c->grid_state = 2;
  c->keep_valid = true;
  return tm.done();
}

int snow_stage_p2g(snow_ctx* c) {
  if (!c) return SNOW_E_ARG;
  int rc;
  if ((rc = run_force_prepass(c))) return rc;
  if ((rc = run_sort(c, -1.0f))) return rc;
  if ((rc = run_p2g(c, 0))) return rc;
  if ((rc = run_grid(c, true, false, 0.0f, false))) return rc;
  c->grid_state = 1;
  c->keep_valid = true;
  return 0;
}

int snow_stage_grid_update(snow_ctx* c, double dt) {
  if (!c) return SNOW_E_ARG;
  const float fdt = float(dt);
  if (!(fdt > 0.0f)) return fail(c, SNOW_E_ARG, "dt must be > 0");
  // after snow_stage_p2g: the tile merge + update of snow_step (the dense node fields the P2G stage
  // kept stay valid)
  if (c->grid_state == 1) return run_grid(c, false, true, fdt, false);
  if (c->grid_state != 2) return fail(c, SNOW_E_STATE, "grid update needs a P2G result or a written grid");
  StageTimer tm(c, ST_GRID);
  if (c->total_nodes) {
    k_grid_update_dense<<<grid_for(c->total_nodes, 256), 256, 0, c->stream>>>(
        c->d_scenes, c->nscenes, c->d_g_mmv, c->d_g_f, c->d_g_node, c->d_g_mass, c->d_cols, fdt, c->total_nodes);
    LAUNCHCHK(c);
  }
  return tm.done();
}

int snow_stage_g2p(snow_ctx* c) {
  if (!c) return SNOW_E_ARG;
  if (c->grid_state == 0) return fail(c, SNOW_E_STATE, "G2P needs grid velocities");
  if (!c->d_aux) HIPCHK(c, dalloc(&c->d_aux, size_t(12) * std::max<uint32_t>(1, c->N)));
  StageTimer tm(c, ST_G2P);
  if (c->N) {
    k_g2p_only<<<grid_for(c->N, 256), 256, 0, c->stream>>>(c->d_scenes, c->d_pscene, c->d_rec[c->cur], c->d_g_node,
                                                            c->d_aux, c->N);
    LAUNCHCHK(c);
  }
  c->aux_valid = true;
  return tm.done();
}

int snow_stage_particle_advance(snow_ctx* c, double dt) {
  if (!c) return SNOW_E_ARG;
  const float fdt = float(dt);
  if (!(fdt > 0.0f)) return fail(c, SNOW_E_ARG, "dt must be > 0");
  if (!c->aux_valid) return fail(c, SNOW_E_STATE, "particle advance needs a G2P result (snow_stage_g2p)");
  int rc;
  if ((rc = clear_bounds(c))) return rc;
  StageTimer tm(c, ST_ADVANCE);
  const int k = c->cur;
  if (c->N) {
    k_advance_only<<<grid_for(c->N, 256), 256, 0, c->stream>>>(c->d_scenes, c->d_pscene, c->d_rec[k], c->d_fe[k],
                                                                c->d_fp[k], c->d_fa, c->d_aux, c->d_cols, fdt,
                                                                c->d_keys[0], c->N, c->d_ctl, c->d_pbound);
    LAUNCHCHK(c);
  }
  c->keys_valid = true;
  c->hist_ready = false;
  c->force_valid = true;
  c->bound_valid = true;
  c->aux_valid = false;
  return tm.done();
}

int snow_stage_g2p_advance(snow_ctx* c, double dt) {
  if (!c) return SNOW_E_ARG;
  const float fdt = float(dt);
  if (!(fdt > 0.0f) || !std::isfinite(fdt)) return fail(c, SNOW_E_ARG, "dt must be > 0");
  if (c->grid_state == 0) return fail(c, SNOW_E_STATE, "G2P needs grid velocities");
  int rc;
  if ((rc = run_keys(c))) return rc;
  // the cell sort of the current positions gives the slot permutation the fused kernel applies;
  // it runs without the step start, so the grid of the last P2G stays readable
  if (c->N) {
    if ((rc = prepare_hist0(c))) return rc;
    size_t tb = c->sort_tmp_bytes;
    HIPCHK(c, radix_sort_pairs(c->d_sort_tmp, &tb, c->d_keys[0], c->d_keys[1], nullptr, c->d_slot, c->N,
                               c->key_bits, c->sort_bits, c->stream, c->num_cu, nullptr, c->d_hist0));
    c->hist_ready = false;
  }
  // the FP64-path counter of this call; the next P2G's bound from this call's records alone
  k_fill_u32<<<1, 64, 0, c->stream>>>(reinterpret_cast<uint32_t*>(fp64_counter(c)), 0u, 2);
  LAUNCHCHK(c);
  if ((rc = clear_bounds(c))) return rc;
  return run_g2p_advance(c, fdt);
}

int snow_read_particles(snow_ctx* c, int32_t s, float* x, float* v, float* fe, float* fp, float* rho0) {
  if (!c || s < 0 || s >= c->nscenes) return SNOW_E_ARG;
  const int64_t n = c->sh[s].n;
  if (n == 0) return 0;
  const size_t fl = size_t(n) * (3 + 3 + 9 + 9 + 1 + 1);
  int rc = ensure_stage(c, fl * sizeof(float));
  if (rc) return rc;
  float* st = static_cast<float*>(c->d_stage);
  float *ox = st, *ov = ox + 3 * n, *ofe = ov + 3 * n, *ofp = ofe + 9 * n, *orho = ofp + 9 * n, *om = orho + n;
  const int k = c->cur;
  k_gather_out<<<grid_for(n, 256), 256, 0, c->stream>>>(c->d_rec[k], c->d_fe[k], c->d_fp[k], c->d_pid[k],
                                                         c->scenes[s].p_off, n, c->N, ox, ov, ofe, ofp, orho, om);
  LAUNCHCHK(c);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (x) HIPCHK(c, hipMemcpy(x, ox, sizeof(float) * 3 * n, hipMemcpyDeviceToHost));
  if (v) HIPCHK(c, hipMemcpy(v, ov, sizeof(float) * 3 * n, hipMemcpyDeviceToHost));
  if (fe) HIPCHK(c, hipMemcpy(fe, ofe, sizeof(float) * 9 * n, hipMemcpyDeviceToHost));
  if (fp) HIPCHK(c, hipMemcpy(fp, ofp, sizeof(float) * 9 * n, hipMemcpyDeviceToHost));
  if (rho0) HIPCHK(c, hipMemcpy(rho0, orho, sizeof(float) * n, hipMemcpyDeviceToHost));
  return 0;
}

int snow_read_particles_noalloc(snow_ctx* c, int32_t s, int64_t first, int64_t count, float* x, float* v,
                                float* fe, float* fp, float* rho0) {
  if (!c || s < 0 || s >= c->nscenes) return SNOW_E_ARG;
  const int64_t n = c->sh[s].n;
  if (first < 0 || count < 0 || first > n || count > n - first)
    return fail(c, SNOW_E_ARG, "particle window outside the scene");
  if (count > NOALLOC_MAX_PARTICLES)
    return fail(c, SNOW_E_CAPACITY, "snow_read_particles_noalloc reads at most 65536 particles per call");
  if (count == 0) return 0;
  float* st = c->noalloc_stage;
  float *ox = st, *ov = ox + 3 * count, *ofe = ov + 3 * count, *ofp = ofe + 9 * count, *orho = ofp + 9 * count;
  const int k = c->cur;
  k_gather_window<<<grid_for(n, 256), 256, 0, c->stream>>>(c->d_rec[k], c->d_fe[k], c->d_fp[k], c->d_pid[k],
                                                            c->scenes[s].p_off, n, c->N, first, count, ox, ov, ofe,
                                                            ofp, orho);
  LAUNCHCHK(c);
  if (x) HIPCHK(c, hipMemcpyAsync(x, ox, sizeof(float) * 3 * count, hipMemcpyDeviceToHost, c->stream));
  if (v) HIPCHK(c, hipMemcpyAsync(v, ov, sizeof(float) * 3 * count, hipMemcpyDeviceToHost, c->stream));
  if (fe) HIPCHK(c, hipMemcpyAsync(fe, ofe, sizeof(float) * 9 * count, hipMemcpyDeviceToHost, c->stream));
  if (fp) HIPCHK(c, hipMemcpyAsync(fp, ofp, sizeof(float) * 9 * count, hipMemcpyDeviceToHost, c->stream));
  if (rho0) HIPCHK(c, hipMemcpyAsync(rho0, orho, sizeof(float) * count, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int snow_write_particles(snow_ctx* c, int32_t s, const float* x, const float* v, const float* fe, const float* fp,
                         const float* rho0) {
  if (!c || s < 0 || s >= c->nscenes) return SNOW_E_ARG;
  const int64_t n = c->sh[s].n;
  if (n == 0) return 0;
  const size_t fl = size_t(n) * (3 + 3 + 9 + 9 + 1);
  int rc = ensure_stage(c, fl * sizeof(float));
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  float* st = static_cast<float*>(c->d_stage);
  float *ix = st, *iv = ix + 3 * n, *ife = iv + 3 * n, *ifp = ife + 9 * n, *irho = ifp + 9 * n;
  int mask = 0;
  if (x) { HIPCHK(c, hipMemcpy(ix, x, sizeof(float) * 3 * n, hipMemcpyHostToDevice)); mask |= 1; }
  if (v) { HIPCHK(c, hipMemcpy(iv, v, sizeof(float) * 3 * n, hipMemcpyHostToDevice)); mask |= 2; }
  if (fe) { HIPCHK(c, hipMemcpy(ife, fe, sizeof(float) * 9 * n, hipMemcpyHostToDevice)); mask |= 4; }
  if (fp) { HIPCHK(c, hipMemcpy(ifp, fp, sizeof(float) * 9 * n, hipMemcpyHostToDevice)); mask |= 8; }
  if (rho0) { HIPCHK(c, hipMemcpy(irho, rho0, sizeof(float) * n, hipMemcpyHostToDevice)); mask |= 16; }
  const int k = c->cur;
  k_scatter_in<<<grid_for(n, 256), 256, 0, c->stream>>>(c->d_rec[k], c->d_fe[k], c->d_fp[k], c->d_pid[k],
                                                         c->scenes[s].p_off, n, c->N, ix, iv, ife, ifp, irho, nullptr,
                                                         mask);
  LAUNCHCHK(c);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (mask & 1) {
    c->keys_valid = false;
    c->hist_ready = false;
  }
  if (mask & (4 | 8 | 16)) c->force_valid = false;
  c->bound_valid = false;
  if (rho0) c->initialized = true;
  return 0;
}

int snow_read_grid(snow_ctx* c, int32_t s, float* m, float* mv, float* f, float* vstar) {
  if (!c || s < 0 || s >= c->nscenes) return SNOW_E_ARG;
  if ((mv || f) && !c->keep_valid)
    return fail(c, SNOW_E_STATE, "grid momentum/force are kept only with SNOW_FLAG_KEEP_GRID or the stage API");
  const SceneHost& H = c->sh[s];
  const int64_t nn = int64_t(H.D[0]) * H.D[1] * H.D[2];
  int rc = ensure_stage(c, size_t(nn) * 10 * sizeof(float));
  if (rc) return rc;
  float* st = static_cast<float*>(c->d_stage);
  float *om = st, *omv = om + nn, *of = omv + 3 * nn, *ovs = of + 3 * nn;
  // With keep off, the mass is read from the node records.
  k_grid_export<<<grid_for(nn, 256), 256, 0, c->stream>>>(c->d_scenes, s, c->d_tile_stamp, c->d_ctl,
                                                          c->grid_state == 2,
                                                          c->d_g_mmv, c->d_g_f, c->d_g_node, c->d_g_mass,
                                                          c->keep_valid ? 0 : 1,
                                                          om, omv, of, ovs);
  LAUNCHCHK(c);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (m) HIPCHK(c, hipMemcpy(m, om, sizeof(float) * nn, hipMemcpyDeviceToHost));
  if (mv) HIPCHK(c, hipMemcpy(mv, omv, sizeof(float) * 3 * nn, hipMemcpyDeviceToHost));
  if (f) HIPCHK(c, hipMemcpy(f, of, sizeof(float) * 3 * nn, hipMemcpyDeviceToHost));
  if (vstar) HIPCHK(c, hipMemcpy(vstar, ovs, sizeof(float) * 3 * nn, hipMemcpyDeviceToHost));
  return 0;
}

int snow_write_grid(snow_ctx* c, int32_t s, const float* m, const float* mv, const float* f, const float* vstar) {
  if (!c || s < 0 || s >= c->nscenes) return SNOW_E_ARG;
  if (!m) return fail(c, SNOW_E_ARG, "grid mass is required");
  if (c->grid_state != 2) {
    // switch to the dense representation: every other node starts at zero
    int rc = snow_stage_grid_reset(c);
    if (rc) return rc;
  }
  const SceneHost& H = c->sh[s];
  const int64_t nn = int64_t(H.D[0]) * H.D[1] * H.D[2];
  std::vector<float4> mm(nn), ff(nn), vs(nn);
  for (int64_t i = 0; i < nn; ++i) {
    mm[i] = make_float4(mv ? mv[3 * i] : 0.f, mv ? mv[3 * i + 1] : 0.f, mv ? mv[3 * i + 2] : 0.f, m[i]);
    ff[i] = make_float4(f ? f[3 * i] : 0.f, f ? f[3 * i + 1] : 0.f, f ? f[3 * i + 2] : 0.f, 0.f);
    vs[i] = make_float4(vstar ? vstar[3 * i] : 0.f, vstar ? vstar[3 * i + 1] : 0.f, vstar ? vstar[3 * i + 2] : 0.f, 0.f);
  }
  int rc = ensure_stage(c, sizeof(float4) * size_t(nn));
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const int64_t off = c->scenes[s].node_off;
  HIPCHK(c, hipMemcpy(c->d_g_mmv + off, mm.data(), sizeof(float4) * nn, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->d_g_f + off, ff.data(), sizeof(float4) * nn, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->d_stage, vs.data(), sizeof(float4) * nn, hipMemcpyHostToDevice));
  if (nn) {
    k_grid_import<<<grid_for(nn, 256), 256, 0, c->stream>>>(c->d_scenes, s, c->d_g_mmv,
                                                            static_cast<const float4*>(c->d_stage), c->d_g_node,
                                                            c->d_g_mass);
    LAUNCHCHK(c);
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->grid_state = 2;
  c->keep_valid = true;
  return 0;
}

int snow_read_colliders(snow_ctx* c, int32_t s, snow_collider* out) {
  if (!c || s < 0 || s >= c->nscenes || !out) return SNOW_E_ARG;
  const int nc = c->sh[s].ncol;
  if (!nc) return 0;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(out, c->d_cols + c->scenes[s].col_off, sizeof(Collider) * nc, hipMemcpyDeviceToHost));
  return 0;
}

int snow_import_reference_particles(int64_t n, const float* pxm, const int32_t* pv, const float* pfe,
                                    const float* pfp, float* x, float* v, float* mass, float* fe, float* fp) {
  if (n < 0 || (n > 0 && (!pxm || !pv))) return SNOW_E_ARG;
  for (int64_t i = 0; i < n; ++i) {
    if (x) { x[3 * i] = pxm[4 * i]; x[3 * i + 1] = pxm[4 * i + 1]; x[3 * i + 2] = pxm[4 * i + 2]; }
    if (mass) mass[i] = pxm[4 * i + 3];
    // pv.xyz = v * 1e8 (particleAdvance.glsl:206); decode as p2gTransfer.glsl:67 does
    if (v) for (int k = 0; k < 3; ++k) v[3 * i + k] = float(pv[4 * i + k]) * 1e-8f;
    for (int cc = 0; cc < 3; ++cc)
      for (int r = 0; r < 3; ++r) {
        const float id = cc == r ? 1.0f : 0.0f;
        if (fe) fe[9 * i + cc * 3 + r] = pfe ? pfe[16 * i + cc * 4 + r] : id;
        if (fp) fp[9 * i + cc * 3 + r] = pfp ? pfp[16 * i + cc * 4 + r] : id;
      }
  }
  return 0;
}

int snow_export_reference_particles(snow_ctx* c, int32_t s, float* pxm, int32_t* pv, float* pfe, float* pfp) {
  if (!c || s < 0 || s >= c->nscenes) return SNOW_E_ARG;
  const int64_t n = c->sh[s].n;
  std::vector<float> x(3 * n), v(3 * n), fe(9 * n), fp(9 * n), rho(n);
  int rc = snow_read_particles(c, s, x.data(), v.data(), fe.data(), fp.data(), rho.data());
  if (rc) return rc;
  // snow_read_particles leaves the gathered masses in the last plane of the staging buffer
  std::vector<float> mass(n);
  HIPCHK(c, hipMemcpy(mass.data(), static_cast<float*>(c->d_stage) + 25 * n, sizeof(float) * n,
                      hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < n; ++i) {
    if (pxm) { pxm[4 * i] = x[3 * i]; pxm[4 * i + 1] = x[3 * i + 1]; pxm[4 * i + 2] = x[3 * i + 2]; pxm[4 * i + 3] = mass[i]; }
    if (pv) {
      for (int k = 0; k < 3; ++k) pv[4 * i + k] = int32_t(v[3 * i + k] * 1e8f);  // ivec3(v*1e8)
      pv[4 * i + 3] = int32_t(rho[i] * 1e6f);                                       // rho0*1e6
    }
    for (int cc = 0; cc < 4; ++cc)
      for (int r = 0; r < 4; ++r) {
        const bool in = cc < 3 && r < 3;
        if (pfe) pfe[16 * i + cc * 4 + r] = in ? fe[9 * i + cc * 3 + r] : (cc == r ? 1.0f : 0.0f);
        if (pfp) pfp[16 * i + cc * 4 + r] = in ? fp[9 * i + cc * 3 + r] : (cc == r ? 1.0f : 0.0f);
      }
  }
  return 0;
}

int snow_sync(snow_ctx* c) {
  if (!c) return SNOW_E_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int snow_check_finite(snow_ctx* c) {
  if (!c) return SNOW_E_ARG;
  if (!c->N) return 0;
  HIPCHK(c, hipMemsetAsync(c->d_bad, 0, sizeof(int), c->stream));
  const int k = c->cur;
  k_check_finite<<<grid_for(c->N, 256), 256, 0, c->stream>>>(c->d_rec[k], c->d_fe[k], c->d_fp[k], c->d_fa, c->N,
                                                              c->d_bad);
  LAUNCHCHK(c);
  int bad = 0;
  HIPCHK(c, hipMemcpyAsync(&bad, c->d_bad, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (bad) return fail(c, SNOW_E_NONFINITE, "non-finite particle state");
  return 0;
}

int snow_enable_stage_timers(snow_ctx* c, int32_t enable) {
  if (!c) return SNOW_E_ARG;
  c->timers = enable != 0;
  return 0;
}
int snow_num_stages(void) { return ST_COUNT; }
const char* snow_stage_name(int32_t st) { return (st >= 0 && st < ST_COUNT) ? kStageNames[st] : nullptr; }
int snow_stage_times(snow_ctx* c, double* total_ms, int64_t* calls) {
  if (!c) return SNOW_E_ARG;
  int rc = resolve_timers(c);
  if (rc) return rc;
  for (int i = 0; i < ST_COUNT; ++i) {
    if (total_ms) total_ms[i] = c->stage_ms[i];
    if (calls) calls[i] = c->stage_calls[i];
  }
  return 0;
}
int snow_reset_stage_times(snow_ctx* c) {
  if (!c) return SNOW_E_ARG;
  int rc = resolve_timers(c);
  if (rc) return rc;
  for (int i = 0; i < ST_COUNT; ++i) {
    c->stage_ms[i] = c->stage_sq_ms[i] = c->stage_min_ms[i] = c->stage_max_ms[i] = 0;
    c->stage_calls[i] = 0;
  }
  return 0;
}
int snow_stage_stats(snow_ctx* c, double* max_ns, double* min_ns, double* avg_ns, double* stdev_ns,
                     int64_t* calls) {
  if (!c) return SNOW_E_ARG;
  int rc = resolve_timers(c);
  if (rc) return rc;
  for (int i = 0; i < ST_COUNT; ++i) {
    const int64_t n = c->stage_calls[i];
    const double avg = n ? c->stage_ms[i] / double(n) : 0.0;
    // sample standard deviation, as stats::stdev (statistics.hpp:9-15)
    const double var = n > 1 ? std::max(0.0, (c->stage_sq_ms[i] - double(n) * avg * avg) / double(n - 1)) : 0.0;
    if (max_ns) max_ns[i] = c->stage_max_ms[i] * 1e6;
    if (min_ns) min_ns[i] = c->stage_min_ms[i] * 1e6;
    if (avg_ns) avg_ns[i] = avg * 1e6;
    if (stdev_ns) stdev_ns[i] = std::sqrt(var) * 1e6;
    if (calls) calls[i] = n;
  }
  return 0;
}
int snow_write_benchstats(snow_ctx* c, const char* path) {
  if (!c || !path) return SNOW_E_ARG;
  double mx[ST_COUNT], mn[ST_COUNT], av[ST_COUNT], sd[ST_COUNT];
  int64_t n[ST_COUNT];
  int rc = snow_stage_stats(c, mx, mn, av, sd, n);
  if (rc) return rc;
  FILE* f = std::fopen(path, "w");
  if (!f) return fail(c, SNOW_E_ARG, std::string("cannot open ") + path);
  // Benchmarker::write layout (benchmarker.cpp:97-120), times in ns
  std::fprintf(f, "BENCHMARKER STATS: (all times in ns)\n\n");
  std::fprintf(f, "Name:;Max:;Min:;Avg:;Stdev:;#Calls:\n");
  for (int i = 0; i < ST_COUNT; ++i) {
    if (!n[i]) continue;
    std::fprintf(f, "%s;%llu;%llu;%g;%g;%lld;\n", kStageNames[i], static_cast<unsigned long long>(std::llround(mx[i])),
                 static_cast<unsigned long long>(std::llround(mn[i])), av[i], sd[i], static_cast<long long>(n[i]));
  }
  const bool ok = std::fclose(f) == 0;
  return ok ? 0 : fail(c, SNOW_E_ARG, std::string("write failed: ") + path);
}
int snow_positions_view(snow_ctx* c, int32_t scene, const float** device_ptr, int64_t* stride_bytes,
                        int64_t* count) {
  if (!c || !device_ptr || scene < 0 || scene >= c->nscenes) return SNOW_E_ARG;
  const SceneDev& S = c->scenes[scene];
  *device_ptr = reinterpret_cast<const float*>(c->d_rec[c->cur] + 2 * size_t(S.p_off));
  if (stride_bytes) *stride_bytes = int64_t(2 * sizeof(float4));
  if (count) *count = S.p_cnt;
  return 0;
}
int snow_active_nodes(snow_ctx* c, int32_t s, int64_t* out) {
  if (!c || s < 0 || s >= c->nscenes || !out) return SNOW_E_ARG;
  std::vector<unsigned long long> v(size_t(ACTIVE_SLOTS) * ACTIVE_STRIDE);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(v.data(), c->d_counts + size_t(s) * ACTIVE_SLOTS * ACTIVE_STRIDE,
                      sizeof(unsigned long long) * v.size(), hipMemcpyDeviceToHost));
  unsigned long long sum = 0;
  for (int k = 0; k < ACTIVE_SLOTS; ++k) sum += v[size_t(k) * ACTIVE_STRIDE];
  *out = int64_t(sum);
  return 0;
}

int snow_fp64_particles(snow_ctx* c, int64_t* out) {
  if (!c || !out) return SNOW_E_ARG;
  unsigned long long v = 0;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(&v, fp64_counter(c), sizeof v, hipMemcpyDeviceToHost));
  *out = int64_t(v);
  return 0;
}

}  // extern "C"
