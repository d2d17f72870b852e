/*
 * snow.h - C ABI of the MI355X snow MPM timestep.
 *
 * This library replaces the GLSL compute-shader dispatch chain of the explicit
 * snow MPM step (reference: src/snow/simulation/timeUpdate/explicittimeupdate.cpp:4-63)
 * and the SSBO buffer interface it reads and writes (src/snow/object/particlesystem.cpp:101-228,
 * src/snow/object/grid.cpp:74-148, src/snow/object/collisionObjects.cpp:25-88).
 *
 * All functions return 0 on success or a negative SNOW_E* code; snow_last_error()
 * returns the message of the last failure on a context (or of a failed snow_create*).
 * Nothing here calls exit(). One context = one HIP stream; use a context from one
 * host thread at a time. Calls are asynchronous on the context's stream except
 * snow_read_*, snow_sync and snow_create*.
 *
 * Matrices are 3x3, column-major (element (row r, col c) at index c*3+r), the
 * order GLSL uses for mat3/mat4 (particle.hpp:8-11, particleAdvance.glsl:82-85).
 */
#ifndef SNOW_H
#define SNOW_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SNOW_ABI_VERSION 1

/* ---- status codes -------------------------------------------------------- */
#define SNOW_OK 0
#define SNOW_E_ARG (-1)        /* invalid argument or descriptor */
#define SNOW_E_HIP (-2)        /* HIP runtime error (message has the HIP string) */
#define SNOW_E_NONFINITE (-3)  /* snow_check_finite found NaN/Inf in particle state */
#define SNOW_E_CAPACITY (-4)   /* size beyond what the context was created for */
#define SNOW_E_NODEVICE (-5)   /* no usable gfx950 device / code object */
#define SNOW_E_STATE (-6)      /* call not valid in the context's current state */

/* ---- parameter blocks ------------------------------------------------------ */

/* Material constants. Defaults (snow_material_defaults) are the values of
 * src/snow/utils/defines.hpp:50-54 plus the shader literals alpha=0.95
 * (g2pTransfer.glsl:2) and g=(0,-9.81,0) (gForceCompute.glsl:39). */
typedef struct snow_material {
  float young;        /* E0, YOUNG_MODULUS (defines.hpp:50)            */
  float poisson;      /* nu, POISSON (defines.hpp:51)                  */
  float hardening;    /* xi, HARDENING (defines.hpp:52)                */
  float crit_comp;    /* theta_c, CRIT_COMPRESSION (defines.hpp:53)    */
  float crit_stretch; /* theta_s, CRIT_STRETCH (defines.hpp:54)        */
  float flip_alpha;   /* FLIP blend alpha (g2pTransfer.glsl:2)         */
  float gravity[3];   /* body acceleration (gForceCompute.glsl:39)     */
  /* Scale applied to the cubic B-spline weight gradient. The reference omits
   * the 1/h factor (cubic.include.glsl:25-40), i.e. grad_scale = 1. A value of
   * 0 selects the physical gradient 1/h. */
  float grad_scale;
} snow_material;

/* Dense node grid: node (i,j,k) sits at origin + spacing*(i,j,k), index
 * i + j*dim[0] + k*dim[0]*dim[1] (grid.cpp:98-108, gridIndex.include.glsl:7-9). */
typedef struct snow_grid {
  int32_t dim[3];
  float origin[3];
  float spacing;
} snow_grid;

/* One collision object, laid out as the reference's three vec4 SSBO entries
 * cx/cv/cn (collisionObjects.cpp:25-66, gForceCompute.glsl:40-53):
 *   pos,type      = cx  (type 0 half-plane, 1 sphere, 2 axis-aligned box [extension])
 *   vel,radius    = cv  (radius used by spheres)
 *   normal,friction = cn (half-plane normal as stored; box half-extents for type 2)
 * Any other type never collides (gForceCompute.glsl:49-53). */
typedef struct snow_collider {
  float pos[3];
  float type;
  float vel[3];
  float radius;
  float normal[3];
  float friction;
} snow_collider;

/* One scene: particles + grid + material + colliders. Host arrays are copied at
 * snow_create* and never retained. fe/fp may be NULL for identity. */
typedef struct snow_scene_desc {
  int64_t num_particles;
  const float* x;     /* [N][3] positions                                   */
  const float* v;     /* [N][3] velocities in m/s                           */
  const float* mass;  /* [N]                                                */
  const float* fe;    /* [N][9] elastic deformation gradient or NULL (=I)   */
  const float* fp;    /* [N][9] plastic deformation gradient or NULL (=I)   */
  snow_grid grid;
  snow_material material;
  int32_t num_colliders;
  const snow_collider* colliders;
  float dt;           /* step used by snow_step_batch (PHYSIC_DT, defines.hpp:23) */
} snow_scene_desc;

/* Context options (flags for snow_create_batch). */
#define SNOW_FLAG_COMPENSATED_G2P (1u << 0) /* compensated sums in the G2P gather     */
#define SNOW_FLAG_MIXED (1u << 1)           /* FP64 SVD fallback near the clamp       */
#define SNOW_FLAG_KEEP_GRID (1u << 2)       /* snow_step also stores m, mv, f per node */

typedef struct snow_ctx snow_ctx;

/* ---- lifetime --------------------------------------------------------------- */

/* Fill defaults: defines.hpp:14-54 (grid 201^3, h=0.05, origin (0.5125,0.5125,1.5125),
 * E=6e6, nu=0.2, xi=30, theta_c=1e-2, theta_s=1e-3, dt=1e-3). */
void snow_material_defaults(snow_material* out);
void snow_grid_defaults(snow_grid* out);

/* Replaces ExplicitTimeUpdate's constructor + ParticleSystem/Grid/CollisionObjects::initSSBO
 * (explicittimeupdate.hpp:14-25, particlesystem.cpp:101-228, grid.cpp:74-148,
 * collisionObjects.cpp:25-66). A single scene is a batch of one. */
int snow_create(const snow_scene_desc* scene, uint32_t flags, snow_ctx** out);
int snow_create_batch(const snow_scene_desc* scenes, int32_t num_scenes, uint32_t flags,
                      snow_ctx** out);
int snow_destroy(snow_ctx* ctx);
const char* snow_last_error(const snow_ctx* ctx);
int snow_num_scenes(const snow_ctx* ctx);
int64_t snow_num_particles(const snow_ctx* ctx, int32_t scene);

/* ---- step interface (TimeUpdate::init / update, timeUpdate.hpp:10-18) ---------- */

/* ExplicitTimeUpdate::init (explicittimeupdate.cpp:4-34): grid reset, grid mass,
 * particle rest density (gMassCompute.glsl, pVolumeCompute.glsl). */
int snow_init(snow_ctx* ctx);
/* ExplicitTimeUpdate::update(dt) (explicittimeupdate.cpp:36-63): rigidAdvance,
 * gReset, p2gTransfer, gForceCompute, g2pTransfer, particleAdvance, for every
 * scene, with the same dt (narrowed to float as technique.cpp:122-125 does). */
int snow_step(snow_ctx* ctx, double dt);
/* Same step with each scene's own snow_scene_desc::dt. */
int snow_step_batch(snow_ctx* ctx);
/* Run `steps` batched steps as replays of a captured HIP graph of the step (one graph per
 * buffer parity, captured on first use). No reference counterpart: the reference dispatches
 * each compute shader from the host every step (explicittimeupdate.cpp:36-63). */
int snow_run_graph(snow_ctx* ctx, int32_t steps);
/* SemiImplicitTimeUpdate::update(dt) (semiimplicittimeupdate.hpp:6-13; its body is empty in the
 * reference, semiimplicittimeupdate.cpp:3-5): the explicit step with the grid velocities v* replaced,
 * before the transfer back to the particles, by the solution v of
 *   M v - beta dt df(dt v) = M v*   (Stomakhin et al. 2013, eq. 9),
 * df the Hessian action of the energy whose gradient is the explicit force. Solved per scene by
 * conjugate residuals preconditioned by the lumped mass, from v*, until |r| <= tol |M v*| or
 * max_iter direction updates. beta in [0, 1] (0: the explicit step, bit for bit; 1: backward Euler
 * in the velocities); dt > 0 for every scene, or 0 for each scene's own dt. Allocates the solver's
 * node vectors on first use (about 130 bytes per grid node). */
int snow_step_semi_implicit(snow_ctx* ctx, double dt, double beta, double tol, int32_t max_iter);
/* Iterations and relative residual |r| / |M v*| of a scene's last semi-implicit solve. */
int snow_implicit_stats(snow_ctx* ctx, int32_t scene, int32_t* iterations, double* rel_residual);
/* Time the launch shapes that do not change the results, `steps` steps per candidate on a
 * snapshot of the current state, keep the fastest and restore the state: the G2P workgroup size
 * (64, 128, 256; *g2p_block) and the sort-bin width, the largest radix digit of the stable sort
 * (8..11 bits; *sort_bits). *p2g_chunk receives the P2G chunk size (particles) of scene 0 and
 * *grid_blocks the number of P2G / grid-merge workgroups; both follow from the particle counts at
 * creation (not tuned: the P2G rounding depends on the chunking). Any of the out pointers may be
 * NULL. No reference counterpart (the reference's work-group sizes are fixed at 1024,
 * defines.hpp:28). */
int snow_autotune(snow_ctx* ctx, int32_t steps, int32_t* g2p_block, int32_t* p2g_chunk,
                  int32_t* grid_blocks, int32_t* sort_bits);

/* ---- per-stage entry points (the Technique::dispatch of each compute shader) ---- */
int snow_stage_grid_mass(snow_ctx* ctx);        /* gMassCompute.glsl:33-84 (init)      */
int snow_stage_particle_volume(snow_ctx* ctx);  /* pVolumeCompute.glsl:32-83 (init)    */
int snow_stage_rigid_advance(snow_ctx* ctx, double dt); /* rigidAdvance.glsl:14-18    */
int snow_stage_grid_reset(snow_ctx* ctx);       /* gReset.glsl:11-16                   */
int snow_stage_p2g(snow_ctx* ctx);              /* p2gTransfer.glsl:44-151             */
int snow_stage_grid_update(snow_ctx* ctx, double dt); /* gForceCompute.glsl:57-103     */
int snow_stage_g2p(snow_ctx* ctx);              /* g2pTransfer.glsl:53-143 (+ advance) */
int snow_stage_particle_advance(snow_ctx* ctx, double dt); /* particleAdvance.glsl:75-217 */
/* g2pTransfer.glsl:53-143 and particleAdvance.glsl:75-217 as the one fused launch snow_step runs
 * (it also computes the next P2G's force matrices and cell keys), on the current particles and the
 * grid velocities of the last grid stage or snow_write_grid. The reference dispatches the two
 * shaders separately (explicittimeupdate.cpp:55-62); this entry point exists so the fused kernel
 * can be run from a given state. It also sets the count of snow_fp64_particles. */
int snow_stage_g2p_advance(snow_ctx* ctx, double dt);

/* ---- state transfer --------------------------------------------------------- */

/* Particle state of one scene in its original (creation) order. Any pointer may be
 * NULL to skip that field. rho0 is the rest density from snow_init (pv.w*1e-6). */
int snow_read_particles(snow_ctx* ctx, int32_t scene, float* x, float* v, float* fe, float* fp,
                        float* rho0);
/* The particles with creation index in [first, first + count) of one scene, as snow_read_particles
 * gives them, staged through a fixed device array of the library instead of device memory it
 * allocates: this call allocates nothing (no hipMalloc). No counterpart in the reference; it is for
 * callers that cannot allocate device memory (serving loops, code that forbids allocation).
 * count <= 65536 (SNOW_E_CAPACITY above). The staging array is shared by every context on a GPU:
 * calls that overlap on one GPU race. */
int snow_read_particles_noalloc(snow_ctx* ctx, int32_t scene, int64_t first, int64_t count, float* x, float* v,
                                float* fe, float* fp, float* rho0);
/* Overwrite particle state of one scene (original order); NULL fields are kept. */
int snow_write_particles(snow_ctx* ctx, int32_t scene, const float* x, const float* v,
                         const float* fe, const float* fp, const float* rho0);
/* Dense grid fields of one scene ([D^3] m, [D^3][3] mv, f, vstar), as left by the last
 * P2G / grid update (gv, gf, gvn buffers, grid.cpp:119-147). NULL fields are skipped. */
int snow_read_grid(snow_ctx* ctx, int32_t scene, float* m, float* mv, float* f, float* vstar);
/* Overwrite the dense grid fields read by the grid update and G2P stages. */
int snow_write_grid(snow_ctx* ctx, int32_t scene, const float* m, const float* mv,
                    const float* f, const float* vstar);
/* Collider state (positions advance with rigidAdvance). */
int snow_read_colliders(snow_ctx* ctx, int32_t scene, snow_collider* out);

/* Reference SSBO encodings (Appendix of particlesystem.cpp:101-228): pxm vec4 (x, m),
 * pv ivec4 (v*1e8, rho0*1e6), pFE / pFP mat4 (column-major, upper-left 3x3).
 * Converts to / from the SoA arrays of snow_scene_desc. */
int snow_import_reference_particles(int64_t n, const float* pxm, const int32_t* pv,
                                    const float* pfe, const float* pfp, float* x, float* v,
                                    float* mass, float* fe, float* fp);
int snow_export_reference_particles(snow_ctx* ctx, int32_t scene, float* pxm, int32_t* pv,
                                    float* pfe, float* pfp);

/* ---- sync, checks, timing --------------------------------------------------- */
int snow_sync(snow_ctx* ctx);
/* Returns SNOW_E_NONFINITE if any particle of any scene has a non-finite x, v, F_E, F_P. */
int snow_check_finite(snow_ctx* ctx);
/* Per-stage timing with hipEvents on the context stream (BenchmarkerGPU::time,
 * benchmarker.hpp:72-78). Stage names follow explicittimeupdate.cpp:39-62. */
int snow_enable_stage_timers(snow_ctx* ctx, int32_t enable);
int snow_num_stages(void);
const char* snow_stage_name(int32_t stage);
/* Accumulated milliseconds and call counts per stage since the last reset. */
int snow_stage_times(snow_ctx* ctx, double* total_ms, int64_t* calls);
int snow_reset_stage_times(snow_ctx* ctx);
/* Per-stage statistics of the timed calls since the last reset, in ns: max, min, mean, sample
 * standard deviation, count (stats::max/min/avg/stdev, statistics.hpp:8-31). Arrays of
 * snow_num_stages() entries; any pointer may be NULL. */
int snow_stage_stats(snow_ctx* ctx, double* max_ns, double* min_ns, double* avg_ns, double* stdev_ns,
                     int64_t* calls);
/* Write the statistics in the reference's benchstats.csv layout (Benchmarker::write,
 * benchmarker.cpp:97-120): "Name:;Max:;Min:;Avg:;Stdev:;#Calls:" rows, stages with calls only. */
int snow_write_benchstats(snow_ctx* ctx, const char* path);
/* Zero-copy view of one scene's particles for rendering: *device_ptr points at a device array of
 * *count float4 (x, y, z, m), *stride_bytes apart - the vec4 (x, m) the reference's point renderer
 * reads from posB (particlesystem.cpp:4-13, pPoints.vert:2-12). Particles are in the order of the
 * last sort, not creation order. Valid until the next step or state write on the context. */
int snow_positions_view(snow_ctx* ctx, int32_t scene, const float** device_ptr, int64_t* stride_bytes,
                        int64_t* count);
/* Number of grid nodes of a scene with mass > 0 after the last P2G (for byte counts). */
int snow_active_nodes(snow_ctx* ctx, int32_t scene, int64_t* out);
/* Particles (all scenes) that took the FP64 SVD path in the last step (SNOW_FLAG_MIXED). */
int snow_fp64_particles(snow_ctx* ctx, int64_t* out);

#ifdef __cplusplus
}
#endif
#endif /* SNOW_H */
