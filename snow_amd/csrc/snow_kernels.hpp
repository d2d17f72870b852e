// snow_kernels.hpp - HIP kernels of the snow MPM step for gfx950 (CDNA4).
//
// Data layout (all scenes of a batch concatenated; see snow_capi.hip for allocation):
//  particles, ping-pong buffers in "sorted" order after each step:
//    rec[i]   : 2 x float4 = { x y z m | vx vy vz rho0 } (rho0: rest density, pv.w of the reference)
//    fe/fp    : 9 SoA float planes, plane e = c*3+r (column-major 3x3)
//    pid      : original (creation) index
//  one buffer, in the slot order of the current particle buffer (written by G2P, read by P2G only):
//    fa       : 9 SoA float planes of the force matrix A, plane r*3+c = A(r, c), f_node = -A grad w
//  grid, dense per scene (node i + j*Dx + k*Dx*Dy):
//    g_node   : 3 x float2 per node: (v*.x v*.y | v*.z u.x | u.y u.z), u = v* - alpha p/m,
//               the gForceCompute output gvn and the FLIP/PIC blend term read together by G2P
//    g_mmv,g_f: float4 (p, m), (f, 0)    - gv / gf of the reference (stage API, debug)
//  cell tiles: 4x4x4 cells per tile over the shifted cell range c = base + 2, base in [-2, D]
//    (a cubic stencil of base cell b touches nodes b-1..b+2); a tile's 7^3-node halo starts
//    at node 4k-3; tpart[tile][q][7^3] holds the tile's partial sums of the 7 P2G quantities
//    (exact fixed-point sums, rounded once to float).
#pragma once
#include "snow_device.hpp"

namespace snow {

constexpr int TILE = 4;              // cells per tile edge
constexpr int TILE_CELLS = 64;       // 4^3
constexpr int HALO = 7;              // TILE + 3 nodes per edge
constexpr int HALO_NODES = 343;      // 7^3
constexpr int NQ = 7;                // m, p.xyz, f.xyz
constexpr int P2G_WAVES = 8;         // waves per P2G workgroup
constexpr int P2G_SEG_MAX = 3072;    // largest P2G chunk in particles (sizes the LDS run lists)
constexpr int P2G_BLOCKS_PER_CU = 2; // P2G workgroups resident per CU (about 70 KB of LDS each)
constexpr int P2G_SEG_MEAN = 2048;   // largest mean chunk of a scene (its particles / its chunks)
constexpr int SORT_TILE_KEYS = 4096;  // keys per tile of the cell sort (one histogram row each)
constexpr uint32_t P2G_RUN_COST = 96; // planning cost of a run (tile) in particles: its barriers,
                                      // partial write and short batches (measured: ~2.6 us per
                                      // run against ~0.03 us per particle)

struct SceneDev {
  int D[3];
  int ncol;
  float origin[3];
  float h;
  float young, poisson, hardening, crit_comp;
  float crit_stretch, alpha, gscale, dt;
  float g[3];
  int col_off;
  int ct[3];          // cell tiles per axis
  uint32_t ct_off;    // global index of the scene's first cell tile
  int nt[3];          // node tiles per axis
  uint32_t nt_off;    // global index of the scene's first node tile
  uint32_t dead_tile; // global index of the scene's out-of-grid tile
  uint32_t seg;       // mean P2G chunk of the scene, in particles: ceil(p_cnt / seg) chunks
  int64_t node_off;   // offset of node 0 in the dense node arrays
  int64_t p_off;      // first particle
  int64_t p_cnt;      // particle count
  float mu0, lambda0; // E/(2(1+nu)), E nu/((1+nu)(1-2nu))
  float inv_h;        // 1/h
  uint32_t chunk_off; // global index of the scene's first P2G chunk
  float uo[3];        // -origin/h: grid coordinate u = x/h + uo
  uint32_t align_w;   // a P2G chunk bound moves to the nearest cell-tile start within
                      // align_w / 2 positions of its planned position (0: it stays there)
};

// Element i of an SoA plane addressed by the byte offset 4*i in 32 bits (a batch holds fewer than
// 2^30 particles, checked at creation): with the plane base uniform, every plane access is a
// scalar base plus one per-lane offset register, instead of a 64-bit address per plane.
__device__ __forceinline__ float ld_plane(const float* __restrict__ base, uint32_t i) {
  return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + (i << 2));
}
__device__ __forceinline__ void st_plane(float* __restrict__ base, uint32_t i, float v) {
  *reinterpret_cast<float*>(reinterpret_cast<char*>(base) + (i << 2)) = v;
}

// Grid coordinate of a position: u = (x - origin)/h evaluated as x*(1/h) - origin/h (one FMA).
__device__ __forceinline__ float grid_u(const SceneDev& S, float x, int k) { return fmaf(x, S.inv_h, S.uo[k]); }

// Node coordinate as the reference stores it in the gxm buffer: origin + h * (float)i computed on
// the host in float with two roundings (grid.cpp:98-102); kept unfused, because a node can lie
// exactly on a collider plane and the rounded value decides whether it collides.
__device__ __forceinline__ float node_coord(const SceneDev& S, int k, int i) {
  return __fadd_rn(S.origin[k], __fmul_rn(S.h, float(i)));
}

// Cubic B-spline weights of the four stencil nodes of one axis, for base = floor(u) and
// f = u - base in [0,1). Node o (o = 0..3) sits at base + o - 1, so its distance d = o - 1 - f
// always falls in the same polynomial piece of cubic.include.glsl:7-16:
//   w0 = (1-f)^3/6, w1 = (3f^3 - 6f^2 + 4)/6, w2 = (-3f^3 + 3f^2 + 3f + 1)/6, w3 = f^3/6,
// and weightingGradient(-d) (cubic.include.glsl:25-34) is dw_o/df:
//   q0 = -(1-f)^2/2, q1 = 3f^2/2 - 2f, q2 = -3f^2/2 + f + 1/2, q3 = f^2/2.
// Flooring (instead of the reference's truncation, Q6) only changes which zero-weight or
// out-of-grid nodes are visited: in-grid contributions are identical.
__device__ __forceinline__ float spline_w(int o, float f) {
  const float c3 = o == 0 ? -1.f / 6.f : (o == 1 ? 0.5f : (o == 2 ? -0.5f : 1.f / 6.f));
  const float c2 = o == 0 ? 0.5f : (o == 1 ? -1.0f : (o == 2 ? 0.5f : 0.f));
  const float c1 = o == 0 ? -0.5f : (o == 2 ? 0.5f : 0.f);
  const float c0 = o == 0 ? 1.f / 6.f : (o == 1 ? 2.f / 3.f : (o == 2 ? 1.f / 6.f : 0.f));
  return fmaf(fmaf(fmaf(c3, f, c2), f, c1), f, c0);
}
__device__ __forceinline__ float spline_q(int o, float f) {
  const float b2 = o == 0 ? -0.5f : (o == 1 ? 1.5f : (o == 2 ? -1.5f : 0.5f));
  const float b1 = o == 0 ? 1.0f : (o == 1 ? -2.0f : (o == 2 ? 1.0f : 0.f));
  const float b0 = o == 0 ? -0.5f : (o == 2 ? 0.5f : 0.f);
  return fmaf(fmaf(b2, f, b1), f, b0);
}

// Global cell key of a position: (tile) * 64 + (cell within tile) for the cell floor(u) + 2;
// particles whose base cell lies outside [-2, D] (no in-grid stencil node) go to the scene's
// dead tile.
__device__ __forceinline__ bool in_stencil_range(const SceneDev& S, float ux, float uy, float uz) {
  // floor(u) in [-2, D]: at least one in-grid node within the cubic support
  return ux >= -2.0f && uy >= -2.0f && uz >= -2.0f && ux < float(S.D[0] + 1) && uy < float(S.D[1] + 1) &&
         uz < float(S.D[2] + 1);
}

__device__ __forceinline__ uint32_t cell_key(const SceneDev& S, float x, float y, float z) {
  const float ux = grid_u(S, x, 0), uy = grid_u(S, y, 1), uz = grid_u(S, z, 2);
  if (!in_stencil_range(S, ux, uy, uz)) return S.dead_tile * TILE_CELLS;
  const int cx = int(floorf(ux)) + 2, cy = int(floorf(uy)) + 2, cz = int(floorf(uz)) + 2;
  const uint32_t tile = S.ct_off + uint32_t((cx >> 2) + S.ct[0] * ((cy >> 2) + S.ct[1] * (cz >> 2)));
  return tile * TILE_CELLS + uint32_t((cx & 3) + 4 * (cy & 3) + 16 * (cz & 3));
}

// ---------------------------------------------------------------------------------------
// Keys of the current particle order (after upload or host writes).
__global__ void __launch_bounds__(256) k_keys(const SceneDev* __restrict__ scenes,
                                              const uint16_t* __restrict__ pscene,
                                              const float4* __restrict__ rec, uint32_t* __restrict__ keys,
                                              uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const SceneDev& S = scenes[pscene[i]];
  const float4 xm = rec[2 * size_t(i)];
  keys[i] = cell_key(S, xm.x, xm.y, xm.z);
}


// Node record written by the grid update and read by G2P: v* (gvn of gForceCompute.glsl) and the
// FLIP/PIC term u = v* - alpha p/m of g2pTransfer.glsl:100, packed as three float2 (24 bytes;
// a row of 4 nodes along x is 96 contiguous bytes). The node mass is kept apart in g_mass.
__device__ __forceinline__ void store_node(float2* __restrict__ g, float* __restrict__ g_mass, int64_t n, float vx,
                                           float vy, float vz, float ux, float uy, float uz, float m) {
  g[3 * n] = make_float2(vx, vy);
  g[3 * n + 1] = make_float2(vz, ux);
  g[3 * n + 2] = make_float2(uy, uz);
  g_mass[n] = m;
}

// ---------------------------------------------------------------------------------------
// Per-step control block in device memory, so every launch of a step has fixed arguments
// (graph-capturable) while the set of active tiles changes from step to step.
struct StepCtl {
  uint32_t step;      // step stamp, incremented at the start of every step (step_begin_block)
  uint32_t n_active;  // unused
  uint32_t spare;
  uint32_t pad;
};

// The list of cell tiles with particles is kept in TILE_LISTS sublists, each with its own
// append counter (TILE_LIST_STRIDE words apart, so the counters sit in different cache lines):
// appends from many workgroups then do not serialise on one address.
constexpr int TILE_LISTS = 64;
constexpr int TILE_LIST_STRIDE = 32;

// Active-node counts: ACTIVE_SLOTS slots per scene, one 128-byte line apart, so that the adds of
// many waves spread over lines; a scene's count is the sum of its slots.
constexpr int ACTIVE_SLOTS = 64;
constexpr int ACTIVE_STRIDE = 16;  // 64-bit words from one slot to the next

// Start of a step: new stamp, cleared counters, and the collider advance of rigidAdvance.glsl:14-18
// (cx += dt*cv) when dt_rigid >= 0 (0 selects each scene's own dt). Nothing the sort reads is
// touched, so the first sort launch carries this work (workgroup 0 of k_radix_hist, or an extra
// workgroup of k_radix_colscan when the first histogram comes counted); the kernel k_step_begin
// runs it alone when there is nothing to sort.
struct StepBegin {
  StepCtl* ctl;
  unsigned long long* counts;  // active-node slots: ncounts slot heads, ACTIVE_STRIDE words apart
  int ncounts;
  unsigned long long* fp64;    // particles that took the FP64 path this step
  int nscenes;
  const SceneDev* scenes;
  Collider* cols;
  uint32_t* list_ctr;
  float dt_rigid;
};

__device__ __forceinline__ void step_begin_block(const StepBegin& a, int t, int nt) {
  if (t < TILE_LISTS) a.list_ctr[t * TILE_LIST_STRIDE] = 0u;
  if (t == 0) {
    a.ctl->step += 1u;
    a.ctl->n_active = 0u;
    a.ctl->spare = 0u;
  }
  for (int i = t; i < a.ncounts; i += nt) a.counts[size_t(i) * ACTIVE_STRIDE] = 0ull;
  if (t == 0) *a.fp64 = 0ull;
  if (a.dt_rigid >= 0.0f) {
    for (int s = t; s < a.nscenes; s += nt) {
      const SceneDev& S = a.scenes[s];
      const float dt = a.dt_rigid > 0.0f ? a.dt_rigid : S.dt;
      for (int c = 0; c < S.ncol; ++c) {
        Collider& C = a.cols[S.col_off + c];
        C.px += dt * C.vx; C.py += dt * C.vy; C.pz += dt * C.vz;
      }
    }
  }
}

__global__ void __launch_bounds__(256) k_step_begin(StepBegin a) { step_begin_block(a, threadIdx.x, blockDim.x); }

// Collider advance alone (rigidAdvance.glsl:14-18) for the per-stage API.
__global__ void __launch_bounds__(256) k_rigid_advance(const SceneDev* __restrict__ scenes, int nscenes,
                                                       Collider* __restrict__ cols, float dt_rigid) {
  for (int s = 0; s < nscenes; ++s) {
    const SceneDev& S = scenes[s];
    const float dt = dt_rigid > 0.0f ? dt_rigid : S.dt;
    for (int c = threadIdx.x; c < S.ncol; c += blockDim.x) {
      Collider& C = cols[S.col_off + c];
      C.px += dt * C.vx; C.py += dt * C.vy; C.pz += dt * C.vz;
    }
  }
}

// ---------------------------------------------------------------------------------------
// P2G (p2gTransfer.glsl:44-151 / gMassCompute.glsl:33-84) over chunks of the sorted particles.
// A scene's particles are contiguous in sorted order; its chunk c (global index chunk_off + c)
// is a sorted range of about seg particles whose window the chunk plan (plan_chunks) sets from
// the previous fused step, so that every workgroup gets about the same cost (particles plus a
// fixed cost per run of one tile) however unevenly the particles fill the cell tiles; a scene's
// chunks are the same alone or in a batch. Both window bounds move to the nearest cell-tile
// start within align_w / 2 positions, so that a tile is split between chunks only when it is
// longer than that: no continuation partial, no fold.
// The workgroup first reads the chunk's keys once, all loads in flight together: it keeps the
// local cells in LDS and marks where the cell tile changes (the chunk falls
// apart into runs of one tile, contiguous and cell by cell); the fixed-point bound is the
// scene's, published by the writers of the records (below). Then, run by run, the 8 waves take
// equal contiguous shares and stage P2G_CHUNK particles at a time into their LDS slot: one lane
// per particle evaluates the particle's 12 weights and 12 gradients once, from records loaded one
// batch ahead. Then lane l of the wave is the stencil offset o = (l&3, (l>>2)&3, l>>4) (getIJK
// order, neighborIndex.include.glsl:6-12): for each staged particle the 64 lanes compute the 64
// node contributions from broadcast LDS reads and sum them in registers. When the cell changes,
// the register sums are added to the run's LDS accumulator as fixed point: rounded to integers
// held in doubles, whose sums stay below 2^53 and are therefore exact, so the run sums do not
// depend on the order of the commits and the result is bit-reproducible. The fixed-point scale
// is a power of two per chunk and quantity, 2^FIX_BITS / (count * max |per-particle term| of the scene), so no
// sum can reach 2^FIX_BITS and the resolution is ~2^-39 of the scene's largest term for a
// 2048-particle chunk.
// A run that starts its tile writes the tile partial tpart[tile]; a run that continues a tile
// begun in an earlier chunk is the chunk's first run and writes hpart[chunk]. The fold adds the
// hpart of the tile's later chunks to tpart[tile] in chunk order, before the grid merge.
struct P2GBound { float m, p, f; };

// The largest per-particle terms of the P2G sums: |m|, |m v_k| and the absolute row sums of A
// (a row-major 3x3), which bound |A grad w| for |grad w| components <= 1 times grad_scale.
__device__ __forceinline__ P2GBound particle_bound(float m, float vx, float vy, float vz, const float Arm[9],
                                                   int skip_heavy) {
  P2GBound b;
  const bool skip = skip_heavy && fabsf(m) > 100.0f;
  b.m = skip ? 0.f : fabsf(m);
  b.p = skip ? 0.f : fmaxf(fabsf(vx * m), fmaxf(fabsf(vy * m), fabsf(vz * m)));
  const float f0 = fabsf(Arm[0]) + fabsf(Arm[1]) + fabsf(Arm[2]);
  const float f1 = fabsf(Arm[3]) + fabsf(Arm[4]) + fabsf(Arm[5]);
  const float f2 = fabsf(Arm[6]) + fabsf(Arm[7]) + fabsf(Arm[8]);
  b.f = skip ? 0.f : fmaxf(f0, fmaxf(f1, f2));
  return b;
}

constexpr int FIX_BITS = 50;  // |fixed-point sums| < 2^FIX_BITS < 2^53: exact in double

// exponent e with count*maxval*2^e <= 2^FIX_BITS (clamped to the float range)
__device__ __forceinline__ int fixed_exponent(float maxval, uint32_t count) {
  if (!(maxval > 0.0f)) return 0;
  const float bound = maxval * float(count);
  int ex = 0;
  frexpf(bound, &ex);  // bound < 2^ex
  const int e = FIX_BITS - ex;
  return e > 126 ? 126 : (e < -126 ? -126 : e);
}

// Fixed-point bounds of the next P2G. Every kernel that writes particle records for a P2G (the
// fused G2P, the stage-API advance, the init force pass, the grid-mass pass) also publishes the
// per-scene maxima of the P2G terms of what it wrote: an atomic max of the non-negative float
// bits into PB_SLOTS slots per scene and step parity, so the P2G needs no pass over the records
// of its own. The P2G of step s reads parity s & 1 and clears parity (s+1) & 1, which the
// writers of step s then fill. Out-of-grid particles are left out (the P2G skips them). Each
// slot has a 128-byte line of its own (words m, p, f), so the atomics of different waves do not
// queue on one line.
constexpr int PB_SLOTS = 64;
constexpr int PB_STRIDE = 32;  // words per slot

__device__ __forceinline__ uint32_t* bound_slots(uint32_t* pb, int sid, uint32_t parity) {
  return pb + (size_t(sid) * 2 + parity) * (PB_SLOTS * PB_STRIDE);
}

// Wave-uniform minimum / maximum by DPP (no LDS traffic): row shifts by 1, 2, 4, 8 leave each
// row's extreme in its lane 15, the row-15 and row-31 broadcasts carry them to lane 63 (lanes
// without a source keep their own value), which is read as a scalar.
template <bool MAX>
__device__ __forceinline__ int wave_ext_i(int v) {
  auto op = [](int a, int b) { return MAX ? max(a, b) : min(a, b); };
  v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x111, 0xf, 0xf, false));  // row_shr:1
  v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x112, 0xf, 0xf, false));  // row_shr:2
  v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x114, 0xf, 0xf, false));  // row_shr:4
  v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x118, 0xf, 0xf, false));  // row_shr:8
  v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x142, 0xa, 0xf, false));  // row_bcast:15 to rows 1, 3
  v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x143, 0xc, 0xf, false));  // row_bcast:31 to rows 2, 3
  return __builtin_amdgcn_readlane(v, 63);
}
__device__ __forceinline__ int wave_min_i(int v) { return wave_ext_i<false>(v); }
__device__ __forceinline__ int wave_max_i(int v) { return wave_ext_i<true>(v); }

// Coordinates of the box index idx = ix + ex (iy + ey iz) for extents uniform across the caller,
// by float reciprocals instead of integer division (an integer division by a variable is ~20
// VALU). With r = 1/ex in float (correctly rounded, or within 1 ulp), fma(idx, r, r/2) is within
// a relative 2^-22 of (idx + 1/2)/ex, whose distance to the nearest integer is at least 1/(2 ex);
// so truncation gives idx / ex exactly for 0 <= idx < 2^21, whatever ex >= 1.
struct BoxSplit {
  int ex, ey;
  float rex, rey;
};
__device__ __forceinline__ BoxSplit box_split(int ex, int ey) {
  return BoxSplit{ex, ey, 1.0f / float(ex), 1.0f / float(ey)};
}
__device__ __forceinline__ void box_coords(const BoxSplit& b, int idx, int& ix, int& iy, int& iz) {
  const int t = int(fmaf(float(idx), b.rex, 0.5f * b.rex));  // idx / ex
  ix = idx - t * b.ex;
  iz = int(fmaf(float(t), b.rey, 0.5f * b.rey));  // t / ey
  iy = t - iz * b.ey;
}

// A wave whose 64 lanes are all active and of one scene publishes its maxima once (DPP
// reductions); otherwise each active lane publishes its own bound. NaN and infinite terms count
// as the largest float.
__device__ __forceinline__ void publish_bound(uint32_t* __restrict__ pb, uint32_t parity, int sid, P2GBound b) {
  float m = b.m <= __FLT_MAX__ ? b.m : __FLT_MAX__;
  float p = b.p <= __FLT_MAX__ ? b.p : __FLT_MAX__;
  float f = b.f <= __FLT_MAX__ ? b.f : __FLT_MAX__;
  const uint32_t slot = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) % PB_SLOTS;
  if (__ballot(true) == ~0ull) {
    const int sid0 = __builtin_amdgcn_readfirstlane(sid);
    if (__ballot(sid != sid0) == 0ull) {
      // values in [0, FLT_MAX]: their bit patterns order as the floats do
      m = __int_as_float(wave_max_i(__float_as_int(m)));
      p = __int_as_float(wave_max_i(__float_as_int(p)));
      f = __int_as_float(wave_max_i(__float_as_int(f)));
      if ((threadIdx.x & 63) == 0) {
        uint32_t* sl = bound_slots(pb, sid0, parity) + slot * PB_STRIDE;
        if (m > 0.f) atomicMax(sl, __float_as_uint(m));
        if (p > 0.f) atomicMax(sl + 1, __float_as_uint(p));
        if (f > 0.f) atomicMax(sl + 2, __float_as_uint(f));
      }
      return;
    }
  }
  uint32_t* sl = bound_slots(pb, sid, parity) + slot * PB_STRIDE;
  if (m > 0.f) atomicMax(sl, __float_as_uint(m));
  if (p > 0.f) atomicMax(sl + 1, __float_as_uint(p));
  if (f > 0.f) atomicMax(sl + 2, __float_as_uint(f));
}

// The same, reduced over the workgroup first (4x fewer global atomics for 256-thread blocks) when
// all its waves hold one scene (one_scene, uniform): s_pb (4 words, zeroed before a barrier at
// kernel start, before any thread returns) collects the block maxima and counts the waves done,
// and the last of the block's `waves` live waves publishes them. Otherwise each wave publishes
// on its own.
__device__ __forceinline__ void publish_bound_block(uint32_t* __restrict__ pb, uint32_t parity, int sid, P2GBound b,
                                                    uint32_t* s_pb, bool one_scene, uint32_t waves) {
  if (!one_scene) {
    publish_bound(pb, parity, sid, b);
    return;
  }
  float m = b.m <= __FLT_MAX__ ? b.m : __FLT_MAX__;
  float p = b.p <= __FLT_MAX__ ? b.p : __FLT_MAX__;
  float f = b.f <= __FLT_MAX__ ? b.f : __FLT_MAX__;
  const int lane = threadIdx.x & 63;
  if (__ballot(true) == ~0ull) {
    // values in [0, FLT_MAX]: their bit patterns order as the floats do
    m = __int_as_float(wave_max_i(__float_as_int(m)));
    p = __int_as_float(wave_max_i(__float_as_int(p)));
    f = __int_as_float(wave_max_i(__float_as_int(f)));
    if (lane == 0) {
      atomicMax(&s_pb[0], __float_as_uint(m));
      atomicMax(&s_pb[1], __float_as_uint(p));
      atomicMax(&s_pb[2], __float_as_uint(f));
    }
  } else {  // a partial wave (its live lanes are its low lanes)
    atomicMax(&s_pb[0], __float_as_uint(m));
    atomicMax(&s_pb[1], __float_as_uint(p));
    atomicMax(&s_pb[2], __float_as_uint(f));
  }
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): this wave's maxima are in LDS
  if (lane == 0 && atomicAdd(&s_pb[3], 1u) == waves - 1u) {
    const uint32_t slot = blockIdx.x % PB_SLOTS;
    uint32_t* sl = bound_slots(pb, sid, parity) + slot * PB_STRIDE;
    const uint32_t bm = s_pb[0], bp = s_pb[1], bf = s_pb[2];
    if (bm) atomicMax(sl, bm);
    if (bp) atomicMax(sl + 1, bp);
    if (bf) atomicMax(sl + 2, bf);
  }
}

// Bound of a particle (mass m, velocity v, row-major force matrix Arm) whose cell key is key;
// zero when the particle is out of the grid.
__device__ __forceinline__ P2GBound record_bound(const SceneDev& S, uint32_t key, float m, float vx, float vy,
                                                 float vz, const float Arm[9], int skip_heavy) {
  if (key / TILE_CELLS == S.dead_tile) return P2GBound{0.f, 0.f, 0.f};
  return particle_bound(m, vx, vy, vz, Arm, skip_heavy);
}

// Force matrix of particle slot i from the planes of fa (row-major order).
__device__ __forceinline__ void load_force(const float* __restrict__ fa, size_t n, uint32_t i, float Arm[9]) {
#pragma unroll
  for (int e = 0; e < 9; ++e) Arm[e] = ld_plane(fa + e * n, i);
}
// Store of a column-major force matrix A into the row-major planes of fa.
__device__ __forceinline__ void store_force(float* __restrict__ fa, size_t n, uint32_t i, const float A[9]) {
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) st_plane(fa + (r * 3 + c) * n, i, A[c * 3 + r]);
}

// Fallback publisher for records no writer has published (host uploads, restored snapshots):
// parity (step + ahead) & 1.
__global__ void __launch_bounds__(256) k_particle_bounds(const SceneDev* __restrict__ scenes,
                                                         const uint16_t* __restrict__ pscene,
                                                         const float4* __restrict__ rec, const float* __restrict__ fa,
                                                         const StepCtl* __restrict__ ctl, uint32_t* __restrict__ pb,
                                                         uint32_t ahead, int skip_heavy, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int sid = pscene[i];
  const SceneDev& S = scenes[sid];
  const float4 xm = rec[2 * size_t(i)], vr = rec[2 * size_t(i) + 1];
  float Arm[9];
  load_force(fa, n, i, Arm);
  const P2GBound b = record_bound(S, cell_key(S, xm.x, xm.y, xm.z), xm.w, vr.x, vr.y, vr.z, Arm, skip_heavy);
  publish_bound(pb, (ctl->step + ahead) & 1u, sid, b);
}

constexpr int P2G_CHUNK = 32;        // particles staged per wave at a time
constexpr int P2G_STAGE = 40;        // floats per staged particle
constexpr uint8_t P2G_DEAD = 0xff;   // local-cell mark of out-of-grid particles

__global__ void __launch_bounds__(64 * P2G_WAVES) k_p2g_chunk(
    const SceneDev* __restrict__ scenes, const uint16_t* __restrict__ tile_scene,
    const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
    uint32_t* __restrict__ tile_c0, uint32_t* __restrict__ tile_c1, uint32_t* __restrict__ tile_stamp,
    const StepCtl* __restrict__ ctl, uint32_t* __restrict__ list_ctr, uint32_t* __restrict__ tile_list,
    uint32_t list_cap, const float4* __restrict__ rec, const float* __restrict__ fa, uint32_t np,
    float* __restrict__ tpart, float* __restrict__ hpart, const uint16_t* __restrict__ chunk_scene,
    uint32_t* __restrict__ pbound, uint32_t* __restrict__ chunk_lo, const uint32_t* __restrict__ chunk_w,
    uint32_t* __restrict__ chunk_runs, int skip_heavy, int by_progress, uint32_t* __restrict__ fold_job) {
  constexpr int NT = 64 * P2G_WAVES;
  constexpr int PER = P2G_SEG_MAX / NT;  // chunk positions per thread in the prologue
  __shared__ double s_acc[NQ * HALO_NODES];  // integer-valued fixed-point sums
  // per-wave staging of P2G_CHUNK particles: for each axis the interleaved (w_o, q_o), o = 0..3
  // (24 floats), then m, m*v (4), then A row-major (9), 3 pad
  __shared__ __attribute__((aligned(16))) float s_stage[P2G_WAVES][P2G_CHUNK * P2G_STAGE];
  __shared__ uint8_t s_cell[P2G_SEG_MAX];        // local cells of the chunk's particles (P2G_DEAD: out of grid)
  __shared__ uint16_t s_run[P2G_SEG_MAX + 1];    // run starts (chunk positions), then the end
  __shared__ uint32_t s_flag[P2G_SEG_MAX / 32];  // run-start bits
  __shared__ float s_bound[3];
  __shared__ uint32_t s_nrun;
  __shared__ uint32_t s_edge[2];
  __shared__ uint32_t s_bnd[2];

  const int t = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int lane = t & 63;
  // lane l is the stencil offset o = (l&3, (l>>2)&3, l>>4) (getIJK order, neighborIndex.include.glsl:6-12)
  const int ox = lane & 3, oy = (lane >> 2) & 3, oz = lane >> 4;
  float* stage = s_stage[wave];
  const uint32_t chunk = blockIdx.x;
  const uint32_t step = ctl->step;
  const int sid = chunk_scene[chunk];
  const SceneDev& S = scenes[sid];
  const uint32_t end = uint32_t(S.p_off + S.p_cnt);
  // the planned window [w0, w1) of the chunk (k_plan_chunks): w1 - w0 <= P2G_SEG_MAX - align_w
  const uint32_t nch = uint32_t((S.p_cnt + S.seg - 1) / S.seg);
  const uint32_t w0 = min(end, chunk_w[chunk]);
  const uint32_t w1 = chunk + 1u < S.chunk_off + nch ? min(end, chunk_w[chunk + 1]) : end;
  uint32_t lo = w0, hi = max(w0, w1);
  if (S.align_w) {  // uniform: both bounds to the nearest tile start within align_w / 2, if any
    const uint32_t A = S.align_w >> 1, pb = uint32_t(S.p_off) + 1u;
    // the scene's start and end stay; a searched bound w lies in (p_off, end), and its window
    // [qlo(w), qhi(w)) in [p_off + 1, end], so keys[q - 1] is the scene's
    auto qlo = [&](uint32_t w) { return w >= pb + A ? w - A : pb; };
    if (t < 2) s_bnd[t] = 0xffffffffu;
    __syncthreads();
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const uint32_t w = side ? w1 : w0;
      if (w > uint32_t(S.p_off) && w < end) {
        const uint32_t q0 = qlo(w), q1 = min(end, w + A);
        // key: distance to w, then position (ties go to the earlier start)
        for (uint32_t q = q0 + uint32_t(t); q < q1; q += NT)
          if (keys[q] / TILE_CELLS != keys[q - 1] / TILE_CELLS)
            atomicMin(&s_bnd[side], ((q >= w ? q - w : w - q) << 20) | (q - q0));
      }
    }
    __syncthreads();
    if (s_bnd[0] != 0xffffffffu) lo = qlo(w0) + (s_bnd[0] & 0xfffffu);
    if (s_bnd[1] != 0xffffffffu) hi = qlo(w1) + (s_bnd[1] & 0xfffffu);
    hi = max(hi, lo);  // nearest-start rounding keeps the bounds in order; kept explicit
  }
  const uint32_t len = min(hi - lo, uint32_t(P2G_SEG_MAX));  // hi - lo < w1 - w0 + align_w
  if (t == 0) {
    chunk_lo[chunk] = lo;  // for the fold of continuation partials
    // tiles of the particles just before and after the chunk (same scene only)
    s_edge[0] = lo > uint32_t(S.p_off) ? keys[lo - 1] / TILE_CELLS : 0xffffffffu;
    s_edge[1] = hi < end ? keys[hi] / TILE_CELLS : 0xffffffffu;
  }

  // ---- prologue: slots, cells, run starts; the scene's fixed-point bound
  {
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const uint32_t o = uint32_t(t + NT * j);
      bool flag = false;
      if (o < len) {
        const uint32_t key = keys[lo + o];
        const uint32_t tl = key / TILE_CELLS;
        flag = o == 0 || tl != keys[lo + o - 1] / TILE_CELLS;
        s_cell[o] = tl == S.dead_tile ? P2G_DEAD : uint8_t(key % TILE_CELLS);
      }
      const unsigned long long bits = __ballot(flag);
      if (lane == 0) {
        const uint32_t w = uint32_t(NT * j + 64 * wave) / 32u;
        s_flag[w] = uint32_t(bits);
        s_flag[w + 1] = uint32_t(bits >> 32);
      }
    }
    const uint32_t par = step & 1u;
    if (wave == P2G_WAVES - 1) {  // max over the scene's bound slots of this step
      // slots hold the bits of non-negative floats: the integer maximum is the float maximum
      const uint32_t* sl = bound_slots(pbound, sid, par) + lane * PB_STRIDE;
      const int bm = wave_max_i(int(sl[0])), bp = wave_max_i(int(sl[1])), bf = wave_max_i(int(sl[2]));
      if (lane == 0) {
        s_bound[0] = __int_as_float(bm);
        s_bound[1] = __int_as_float(bp);
        s_bound[2] = __int_as_float(bf);
      }
    }
    // the scene's first chunk clears the slots the writers of this step fill for the next one
    if (chunk == S.chunk_off && t < PB_SLOTS * 4)
      bound_slots(pbound, sid, par ^ 1u)[(t >> 2) * PB_STRIDE + (t & 3)] = 0u;
    for (int i = t; i < NQ * HALO_NODES; i += NT) s_acc[i] = 0.0;
  }
  __syncthreads();
  if (wave == 0) {  // ordered run starts from the bit mask, two 32-position words per lane
    constexpr int WORDS = P2G_SEG_MAX / 32;
    uint32_t wd[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) wd[h] = 2 * lane + h < WORDS ? s_flag[2 * lane + h] : 0u;
    const uint32_t cnt = __popc(wd[0]) + __popc(wd[1]);
    uint32_t incl = cnt;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = __shfl_up(incl, off, 64);
      if (lane >= off) incl += y;
    }
    uint32_t o = incl - cnt;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      uint32_t word = wd[h];
      while (word) {
        const int b = __ffs(word) - 1;
        word &= word - 1u;
        s_run[o++] = uint16_t(32 * (2 * lane + h) + b);
      }
    }
    if (lane == 63) {
      s_nrun = incl;
      s_run[incl] = uint16_t(len);
    }
  }
  const float bm = s_bound[0], bp = s_bound[1], bf = s_bound[2];
  const int em = fixed_exponent(bm, len), ep = fixed_exponent(bp, len), ef = fixed_exponent(bf, len);
  const float sm = ldexpf(1.0f, em), sp = ldexpf(1.0f, ep), sf = ldexpf(1.0f, ef);
  const float inv_h = S.inv_h, uo0 = S.uo[0], uo1 = S.uo[1], uo2 = S.uo[2];
  __syncthreads();
  const uint32_t nrun = s_nrun;
  if (t == 0) {
    chunk_runs[chunk] = nrun;  // for the plan of the next P2G's windows
    // the fold of this chunk: the tile its first run continues from an earlier chunk, if that run
    // also ends here (then this chunk is the tile's last and adds the continuation partials)
    uint32_t job = 0xffffffffu;
    if (nrun > 0 && s_cell[s_run[0]] != P2G_DEAD) {
      const uint32_t tf = keys[lo + s_run[0]] / TILE_CELLS;
      if (s_edge[0] == tf && (nrun > 1 || s_edge[1] != tf)) job = tf;
    }
    fold_job[chunk] = job;
  }

  // ---- the wave's batches over all runs, with the records of the next batch loaded ahead
  // cursor: run nr, batch [nc, min(nc + P2G_CHUNK, ne)) of the wave's share [., ne) of run nr
  uint32_t nr = 0, nc = 0, ne = 0;
  auto seek = [&](uint32_t r) {  // first run >= r, not out of grid, with a non-empty share
    for (; r < nrun; ++r) {
      const uint32_t r_lo = s_run[r], r_hi = s_run[r + 1];
      if (s_cell[r_lo] == P2G_DEAD) continue;
      const uint32_t sh = (r_hi - r_lo + P2G_WAVES - 1) / P2G_WAVES;
      nc = min(r_hi, r_lo + uint32_t(wave) * sh);
      ne = min(r_hi, nc + sh);
      if (nc < ne) break;
    }
    nr = __builtin_amdgcn_readfirstlane(r);
    nc = __builtin_amdgcn_readfirstlane(nc);
    ne = __builtin_amdgcn_readfirstlane(ne);
  };
  float4 q0 = make_float4(0.f, 0.f, 0.f, 0.f), q1 = q0;  // (x, m), (v, rho0)
  float qa[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};  // A, row-major
  uint32_t qcell = 0u;
  auto fetch = [&]() {
    if (nr < nrun && uint32_t(lane) < min(uint32_t(P2G_CHUNK), ne - nc)) {
      const uint32_t p = vals[lo + nc + lane];
      qcell = s_cell[nc + lane];
      q0 = rec[2 * size_t(p)];
      q1 = rec[2 * size_t(p) + 1];
      load_force(fa, np, p, qa);
    }
  };
  seek(0);
  fetch();
  // Issue priority by progress, when every workgroup of the launch is resident at once (by_progress):
  // a CU then holds two of them to the end, and its arbiter prefers older waves, so without this
  // the first workgroup runs ahead and the second finishes alone at half the occupancy. A wave's
  // priority falls by one at each quarter of its share, so the waves of a CU stay within a quarter
  // of each other. With more workgroups than slots the oldest-first order frees slots soonest and
  // is kept.
  const uint32_t share = len / P2G_WAVES + 1u;
  uint32_t done = 0u;
  int prio = 0;
  if (by_progress) {
    prio = 3;
    __builtin_amdgcn_s_setprio(3);
  }

  for (uint32_t r = 0; r < nrun; ++r) {
    const uint32_t r_lo = s_run[r];
    if (s_cell[r_lo] == P2G_DEAD) continue;  // out-of-grid particles (uniform)
    const uint32_t tile = keys[lo + r_lo] / TILE_CELLS;
    const bool starts_tile = r > 0 || s_edge[0] != tile;

    // Lane l sums the contributions of the staged particles to its stencil node o in registers
    // (mass, momentum and A grad(w), p2gTransfer.glsl:84-146) while the cell stays the same;
    // all particles of a cell share the 64 stencil nodes.
    int cur = -1;  // local cell of the register sums
    float am = 0.f, apx = 0.f, apy = 0.f, apz = 0.f, afx = 0.f, afy = 0.f, afz = 0.f;
    auto commit = [&]() {
      const int cx = cur & 3, cy = (cur >> 2) & 3, cz = cur >> 4;
      double* a = s_acc + (cx + ox) + HALO * (cy + oy) + HALO * HALO * (cz + oz);
      atomicAdd(a + 0 * HALO_NODES, double(rintf(am * sm)));
      atomicAdd(a + 1 * HALO_NODES, double(rintf(apx * sp)));
      atomicAdd(a + 2 * HALO_NODES, double(rintf(apy * sp)));
      atomicAdd(a + 3 * HALO_NODES, double(rintf(apz * sp)));
      // the sums hold +A grad(w); the node force is -A grad(w)
      atomicAdd(a + 4 * HALO_NODES, double(rintf(-afx * sf)));
      atomicAdd(a + 5 * HALO_NODES, double(rintf(-afy * sf)));
      atomicAdd(a + 6 * HALO_NODES, double(rintf(-afz * sf)));
      am = apx = apy = apz = afx = afy = afz = 0.f;
    };
    while (nr == r) {  // the wave's batches of this run
      const uint32_t cnt = min(uint32_t(P2G_CHUNK), ne - nc);
      if (by_progress) {
        const int want = 3 - int(min(3u, (4u * done) / share));
        if (want != prio) {
          prio = want;
          if (want == 2) __builtin_amdgcn_s_setprio(2);
          else if (want == 1) __builtin_amdgcn_s_setprio(1);
          else __builtin_amdgcn_s_setprio(0);
        }
      }
      done += cnt;
      // stage: lane k < cnt evaluates the weights of its prefetched particle once
      const uint32_t cellv = uint32_t(lane) < cnt ? qcell : 0xffffffffu;
      if (uint32_t(lane) < cnt) {
        float4* d4 = reinterpret_cast<float4*>(stage + lane * P2G_STAGE);
        {
          // p2gTransfer.glsl:64-66 skips |m| > 100 (gMassCompute does not)
          const float use = (skip_heavy && fabsf(q0.w) > 100.0f) ? 0.0f : 1.0f;
          const float m = q0.w * use;
          // every particle of a cell has base = floor(u) = that cell; f = u - base
          const float ux = fmaf(q0.x, inv_h, uo0), uy = fmaf(q0.y, inv_h, uo1), uz = fmaf(q0.z, inv_h, uo2);
          const float fx = ux - floorf(ux), fy = uy - floorf(uy), fz = uz - floorf(uz);
          d4[0] = make_float4(spline_w(0, fx), spline_q(0, fx), spline_w(1, fx), spline_q(1, fx));
          d4[1] = make_float4(spline_w(2, fx), spline_q(2, fx), spline_w(3, fx), spline_q(3, fx));
          d4[2] = make_float4(spline_w(0, fy), spline_q(0, fy), spline_w(1, fy), spline_q(1, fy));
          d4[3] = make_float4(spline_w(2, fy), spline_q(2, fy), spline_w(3, fy), spline_q(3, fy));
          d4[4] = make_float4(spline_w(0, fz), spline_q(0, fz), spline_w(1, fz), spline_q(1, fz));
          d4[5] = make_float4(spline_w(2, fz), spline_q(2, fz), spline_w(3, fz), spline_q(3, fz));
          d4[6] = make_float4(m, q1.x * m, q1.y * m, q1.z * m);
          d4[7] = make_float4(qa[0] * use, qa[1] * use, qa[2] * use, qa[3] * use);
          d4[8] = make_float4(qa[4] * use, qa[5] * use, qa[6] * use, qa[7] * use);
          d4[9] = make_float4(qa[8] * use, 0.f, 0.f, 0.f);
        }
      }
      // advance the cursor and load the next batch's records while this one is summed
      if (nc + P2G_CHUNK < ne) nc += P2G_CHUNK;
      else seek(nr + 1);
      fetch();
      // the wave's own LDS writes complete before its reads (in-order LDS per wave)
      __builtin_amdgcn_wave_barrier();
      // bit k: staged particle k starts a new cell (bit 0: against the cell of the register sums),
      // so the loop tests a scalar bit and reads the cell only where it changes
      const uint32_t prevc = uint32_t(__shfl_up(int(cellv), 1, 64));
      const uint64_t chg = __ballot(uint32_t(lane) < cnt && (lane == 0 ? int(cellv) != cur : cellv != prevc));
      for (uint32_t k = 0; k < cnt; ++k) {
        if ((chg >> k) & 1ull) {
          if (cur >= 0) commit();
          cur = __builtin_amdgcn_readlane(int(cellv), int(k));
        }
        const float* d = stage + k * P2G_STAGE;
        const float2 wqx = reinterpret_cast<const float2*>(d)[ox];
        const float2 wqy = reinterpret_cast<const float2*>(d + 8)[oy];
        const float2 wqz = reinterpret_cast<const float2*>(d + 16)[oz];
        const float4 u0 = reinterpret_cast<const float4*>(d)[6];
        const float4 u1 = reinterpret_cast<const float4*>(d)[7];
        const float4 u2 = reinterpret_cast<const float4*>(d)[8];
        const float a22 = d[36];
        // w = wx wy wz and grad(w) = (qx wy wz, wx qy wz, wx wy qz) (cubic.include.glsl:36-40)
        const float wyz = wqy.x * wqz.x;
        const float w = wqx.x * wyz;
        const float g0 = wqx.y * wyz, g1 = wqy.y * (wqx.x * wqz.x), g2 = wqz.y * (wqx.x * wqy.x);
        am = fmaf(u0.x, w, am);
        apx = fmaf(u0.y, w, apx);
        apy = fmaf(u0.z, w, apy);
        apz = fmaf(u0.w, w, apz);
        afx = fmaf(u1.x, g0, fmaf(u1.y, g1, fmaf(u1.z, g2, afx)));
        afy = fmaf(u1.w, g0, fmaf(u2.x, g1, fmaf(u2.y, g2, afy)));
        afz = fmaf(u2.z, g0, fmaf(u2.w, g1, fmaf(a22, g2, afz)));
      }
      __builtin_amdgcn_wave_barrier();
    }
    if (cur >= 0) commit();
    __syncthreads();
    float* out = starts_tile ? tpart + size_t(tile) * (NQ * HALO_NODES) : hpart + size_t(chunk) * (NQ * HALO_NODES);
    // write the run sums out and clear the accumulator for the next run
    for (int i = t; i < NQ * HALO_NODES; i += NT) {
      const int q = i / HALO_NODES;
      const int e = q == 0 ? em : (q < 4 ? ep : ef);
      const double v = s_acc[i];
      s_acc[i] = 0.0;
      out[i] = float(ldexp(v, -e));  // one rounding to float
    }
    __syncthreads();
  }

  // ---- the tiles of the chunk's runs, once for the whole chunk, so that no run waits on the list
  // atomic: the tiles a run starts get this chunk as their first and this step's stamp and join a
  // tile sublist (one atomic per 64 runs), the tiles a run ends this chunk as their last
  if (wave == 0) {
    const uint32_t sub = chunk % TILE_LISTS;
    for (uint32_t r0 = 0; r0 < nrun; r0 += 64) {
      const uint32_t r = r0 + uint32_t(lane);
      bool starts = false;
      uint32_t tile = 0u;
      if (r < nrun) {
        const uint32_t r_lo = s_run[r];
        if (s_cell[r_lo] != P2G_DEAD) {
          tile = keys[lo + r_lo] / TILE_CELLS;
          starts = r > 0 || s_edge[0] != tile;
          if (starts) {
            tile_c0[tile] = chunk;  // the chunks of the tile's first and last runs, for the fold
            tile_stamp[tile] = step;
          }
          if (r + 1 < nrun || s_edge[1] != tile) tile_c1[tile] = chunk;
        }
      }
      const uint64_t mask = __ballot(starts);
      if (mask) {
        uint32_t base = 0u;
        if (lane == 0) base = atomicAdd(&list_ctr[sub * TILE_LIST_STRIDE], uint32_t(__popcll(mask)));
        base = uint32_t(__builtin_amdgcn_readfirstlane(int(base)));
        if (starts) tile_list[size_t(sub) * list_cap + base + uint32_t(__popcll(mask & ((1ull << lane) - 1ull)))] = tile;
      }
    }
  }
}

// ---------------------------------------------------------------------------------------
// A tile whose particles span several P2G chunks has its first run in tpart[tile] and one
// continuation partial per later chunk. The block of the tile's last chunk (fold_job, recorded by
// that chunk in the P2G) adds the continuations to tpart[tile] in chunk order (double sums of the
// float partials).
__global__ void __launch_bounds__(1024) k_fold_partials(const uint32_t* __restrict__ fold_job,
                                                       const uint32_t* __restrict__ tile_c0,
                                                       const uint32_t* __restrict__ chunk_lo,
                                                       float* __restrict__ tpart,
                                                       const float* __restrict__ hpart) {
  const uint32_t chunk = blockIdx.x;
  const uint32_t tile = fold_job[chunk];
  if (tile == 0xffffffffu) return;  // no tile ends here that began in an earlier chunk
  const uint32_t c0 = tile_c0[tile];
  float* dst = tpart + size_t(tile) * (NQ * HALO_NODES);
  // the chunks after the first that hold particles (each wrote a continuation partial; a chunk
  // whose planned window was empty wrote nothing; this chunk holds the tile's last run)
  for (int i = threadIdx.x; i < NQ * HALO_NODES; i += blockDim.x) {
    double acc = double(dst[i]);
    // eight loads in flight at a time; the sum stays in chunk order (adding 0.0 is exact)
    for (uint32_t c = c0 + 1u; c <= chunk; c += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u)
        v[u] = (c + u <= chunk && (c + u == chunk || chunk_lo[c + u] < chunk_lo[c + u + 1]))
                   ? double(hpart[size_t(c + u) * (NQ * HALO_NODES) + i]) : 0.0;
#pragma unroll
      for (int u = 0; u < 8; ++u) acc += v[u];
    }
    dst[i] = float(acc);
  }
}

// Chunk windows of scene S for the next P2G, from the runs its chunks had in this one. Chunk c
// costs its particles plus k runs (k = P2G_RUN_COST, lowered where needed so that no window can
// exceed P2G_SEG_MAX - align_w - 2 positions); the windows cut the cost, taken as uniform along
// each chunk, into equal parts: window j starts where the cost prefix reaches j total / nch. A
// window is at most total / nch + 2 <= that bound long, because a chunk's cost is at least its
// length. The plan depends on the scene's own chunks only (the same alone or in a batch).
__device__ void plan_chunks(const SceneDev& S, const uint32_t* __restrict__ chunk_lo,
                            const uint32_t* __restrict__ chunk_runs, unsigned long long* __restrict__ pref,
                            uint32_t* __restrict__ chunk_w) {
  constexpr uint32_t LDS_CHUNKS = 4096;  // cost prefixes kept in LDS up to this many chunks
  __shared__ unsigned long long s_red[32];
  __shared__ unsigned long long s_carry;
  __shared__ unsigned long long s_pref[LDS_CHUNKS];
  const uint32_t n = uint32_t(S.p_cnt);
  const uint32_t nch = (n + S.seg - 1) / S.seg;
  if (nch == 0) return;
  const uint32_t off = S.chunk_off, pbase = uint32_t(S.p_off), end = pbase + n;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6, nw = blockDim.x >> 6;
  unsigned long long* pf = nch <= LDS_CHUNKS ? s_pref : pref + off;  // pf[c]: cost before chunk c
  auto len_of = [&](uint32_t c) {  // positions of chunk c (scene-relative index)
    return (c + 1 < nch ? chunk_lo[off + c + 1] : end) - chunk_lo[off + c];
  };
  // block sum of 64-bit values (all threads call it)
  auto block_sum = [&](unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    __syncthreads();
    if (lane == 0) s_red[wave] = v;
    __syncthreads();
    unsigned long long r = 0;
    for (int w = 0; w < nw; ++w) r += s_red[w];
    return r;
  };
  unsigned long long runs = 0;
  for (uint32_t c = t; c < nch; c += blockDim.x) runs += chunk_runs[off + c];
  const unsigned long long R = block_sum(runs);
  const unsigned long long wmax = uint64_t(P2G_SEG_MAX) - S.align_w - 2u;
  const unsigned long long room = uint64_t(nch) * wmax > n ? uint64_t(nch) * wmax - n : 0u;
  const unsigned long long k = R ? (room / R < P2G_RUN_COST ? room / R : P2G_RUN_COST) : 0u;
  // exclusive prefix of cost(c) = len(c) + k runs(c), 1024 chunks at a time
  if (t == 0) s_carry = 0;
  for (uint32_t c0 = 0; c0 < nch; c0 += blockDim.x) {
    const uint32_t c = c0 + t;
    const unsigned long long v = c < nch ? len_of(c) + k * chunk_runs[off + c] : 0u;
    unsigned long long incl = v;
    for (int o = 1; o < 64; o <<= 1) {
      const unsigned long long y = __shfl_up(incl, o, 64);
      if (lane >= o) incl += y;
    }
    __syncthreads();  // s_red and s_carry of the previous pass are read
    if (lane == 63) s_red[wave] = incl;
    __syncthreads();
    unsigned long long before = s_carry;
    for (int w = 0; w < wave; ++w) before += s_red[w];
    if (c < nch) pf[c] = before + incl - v;
    __syncthreads();
    if (t == blockDim.x - 1) s_carry = before + incl;
  }
  __syncthreads();
  const unsigned long long total = s_carry;
  if (t == 0) chunk_w[off] = pbase;
  for (uint32_t j = 1 + t; j < nch; j += blockDim.x) {
    const unsigned long long target = total * j / nch;
    uint32_t a = 0, b = nch;  // pref[a] <= target < pref[b] (pref[nch] = total)
    while (b - a > 1u) {
      const uint32_t m = (a + b) >> 1;
      if (pf[m] <= target) a = m; else b = m;
    }
    const unsigned long long cost = (a + 1 < nch ? pf[a + 1] : total) - pf[a];
    const uint32_t len = len_of(a);
    const unsigned long long into = target - pf[a];
    const uint32_t w = chunk_lo[off + a] + (cost ? uint32_t(into * len / cost) : 0u);
    chunk_w[off + j] = w < end ? w : end;
  }
}

// ---------------------------------------------------------------------------------------
// Merge of the tile partials into node values + grid update (gForceCompute.glsl:57-103).
// A node n = 4K + r (per axis) receives from cell tiles K (halo index r + 3) and, when r >= 1,
// K+1 (halo index r - 1): at most 8 tiles, summed in a fixed order. Workgroups take the listed
// cell tiles k (those with particles) in turn; wave a of the workgroup handles node tile
// K = k - a, a in {0,1}^3, if k is the first stamped tile among K + b (b in the fixed order
// 000,100,010,...), so every node tile reached by a particle stencil is processed exactly once.
__device__ __forceinline__ bool tile_live(const SceneDev& S, const uint32_t* __restrict__ stamp, uint32_t step,
                                          int kx, int ky, int kz) {
  if (kx < 0 || ky < 0 || kz < 0 || kx >= S.ct[0] || ky >= S.ct[1] || kz >= S.ct[2]) return false;
  return stamp[S.ct_off + uint32_t(kx + S.ct[0] * (ky + S.ct[1] * kz))] == step;
}

constexpr int GRID_THREADS = 512;  // 8 waves: one per candidate node tile

// tile_scene (scene of every cell tile) may be null for a context of one scene.
template <bool KEEP, bool UPDATE>
__global__ void __launch_bounds__(GRID_THREADS) k_grid(const SceneDev* __restrict__ scenes,
                                                       const uint16_t* __restrict__ tile_scene,
                                                       const uint32_t* __restrict__ list_ctr,
                                                       const uint32_t* __restrict__ tile_list, uint32_t list_cap,
                                                       const StepCtl* __restrict__ ctl,
                                                       const uint32_t* __restrict__ tile_stamp,
                                                       const float* __restrict__ tpart,
                                                       const float* __restrict__ zero_part,
                                                       float4* __restrict__ g_mmv, float4* __restrict__ g_f,
                                                       float2* __restrict__ g_node, float* __restrict__ g_mass,
                                                       const Collider* __restrict__ cols, float dt_override,
                                                       unsigned long long* __restrict__ active_count,
                                                       uint32_t nplan, const uint32_t* __restrict__ chunk_lo,
                                                       const uint32_t* __restrict__ chunk_runs,
                                                       unsigned long long* __restrict__ plan_pref,
                                                       uint32_t* __restrict__ chunk_w) {
  // blocks 0 .. nplan-1 plan the P2G chunk windows of scene blockIdx.x for the next step (first,
  // so that they run beside the merge blocks rather than after them)
  if (blockIdx.x < nplan) {
    plan_chunks(scenes[blockIdx.x], chunk_lo, chunk_runs, plan_pref, chunk_w);
    return;
  }
  const uint32_t bid = blockIdx.x - nplan, nblk = gridDim.x - nplan;
  const uint32_t step = ctl->step;
  const int t = threadIdx.x;
  const int a = __builtin_amdgcn_readfirstlane(t >> 6);
  const int lane = t & 63;
  const int rx = lane & 3, ry = (lane >> 2) & 3, rz = lane >> 4;
  // sublist sizes (lane i: sublist i) and their inclusive prefix
  const uint32_t cnt = lane < TILE_LISTS ? list_ctr[lane * TILE_LIST_STRIDE] : 0u;
  uint32_t incl = cnt;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(incl, off, 64);
    if (lane >= off) incl += y;
  }
  const uint32_t total = uint32_t(__builtin_amdgcn_readlane(int(incl), 63));
  // cell tile of list entry e (uniform values: the loads become scalar loads)
  auto entry = [&](uint32_t e) -> uint32_t {
    const int sub = __builtin_amdgcn_readfirstlane(__ffsll((long long)__ballot(incl > e)) - 1);
    const uint32_t first = uint32_t(__builtin_amdgcn_readlane(int(incl - cnt), sub));
    return tile_list[size_t(sub) * list_cap + (e - first)];
  };
  // the wave's active-node count, added to slot cnt_slot of its scene when the scene changes and
  // at the end (one atomic per wave and scene instead of one per node tile)
  int cnt_sid = -1;
  const int cnt_slot = int((bid * uint32_t(GRID_THREADS / 64) + uint32_t(a)) % uint32_t(ACTIVE_SLOTS));
  unsigned long long cnt_sum = 0ull;
  uint32_t k_next = bid < total ? entry(bid) : 0u;
  for (uint32_t e = bid; e < total; e += nblk) {
    const uint32_t k = k_next;
    // the next entry's list slot loads while this tile is merged
    if (e + nblk < total) k_next = entry(e + nblk);
    const int sid = tile_scene ? tile_scene[k] : 0;  // no table: one scene, whose record loads at once
    const SceneDev& S = scenes[sid];
    if (sid != cnt_sid) {  // the wave's count so far belongs to another scene
      if (lane == 0 && cnt_sum && active_count)
        atomicAdd(active_count + size_t(cnt_sid * ACTIVE_SLOTS + cnt_slot) * ACTIVE_STRIDE, cnt_sum);
      cnt_sid = sid;
      cnt_sum = 0ull;
    }
    const uint32_t kl = k - S.ct_off;
    const int kx = int(kl % uint32_t(S.ct[0])), ky = int((kl / uint32_t(S.ct[0])) % uint32_t(S.ct[1]));
    const int kz = int(kl / uint32_t(S.ct[0] * S.ct[1]));
    const int Kx = kx - (a & 1), Ky = ky - ((a >> 1) & 1), Kz = kz - (a >> 2);
    if (Kx < 0 || Ky < 0 || Kz < 0 || Kx >= S.nt[0] || Ky >= S.nt[1] || Kz >= S.nt[2]) continue;
    // owner: first live tile among K + b, b = 0..7 in order
    bool live[8];
    int owner = -1;
#pragma unroll
    for (int bb = 0; bb < 8; ++bb) {
      live[bb] = tile_live(S, tile_stamp, step, Kx + (bb & 1), Ky + ((bb >> 1) & 1), Kz + (bb >> 2));
      if (owner < 0 && live[bb]) owner = bb;
    }
    if (owner != a) continue;

    const int nx = TILE * Kx + rx, ny = TILE * Ky + ry, nz = TILE * Kz + rz;
    const bool inside = nx < S.D[0] && ny < S.D[1] && nz < S.D[2];
    if (!inside) continue;
    // the tile partials are exact fixed-point sums rounded once to float; they are summed in
    // double, in a fixed order, before the one rounding of the node value to float.
    // Candidates that do not contribute read a zero block, so all loads issue together.
    const float* src[8];
#pragma unroll
    for (int bb = 0; bb < 8; ++bb) {
      const int bx = bb & 1, by = (bb >> 1) & 1, bz = bb >> 2;
      const bool use = live[bb] && !((bx && rx < 1) || (by && ry < 1) || (bz && rz < 1));
      const uint32_t ctile = S.ct_off + uint32_t((Kx + bx) + S.ct[0] * ((Ky + by) + S.ct[1] * (Kz + bz)));
      const int j = (bx ? rx - 1 : rx + 3) + HALO * (by ? ry - 1 : ry + 3) + HALO * HALO * (bz ? rz - 1 : rz + 3);
      src[bb] = use ? tpart + size_t(ctile) * (NQ * HALO_NODES) + j : zero_part;
    }
    double acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      acc[q] = double(src[0][q * HALO_NODES]);
#pragma unroll
      for (int bb = 1; bb < 8; ++bb) acc[q] += double(src[bb][q * HALO_NODES]);
    }
    const float m = float(acc[0]);
    const float px = float(acc[1]), py = float(acc[2]), pz = float(acc[3]);
    const float fx = float(acc[4]), fy = float(acc[5]), fz = float(acc[6]);
    const int64_t node = S.node_off + nx + int64_t(S.D[0]) * (ny + int64_t(S.D[1]) * nz);
    if (KEEP) {
      g_mmv[node] = make_float4(px, py, pz, m);
      g_f[node] = make_float4(fx, fy, fz, 0.0f);
    }
    if (UPDATE) {
      const float dt = dt_override > 0.0f ? dt_override : S.dt;
      float vx = 0.f, vy = 0.f, vz = 0.f, bx = 0.f, by = 0.f, bz = 0.f;
      if (m > 0.0f) {
        vx = px / m + dt * (fx / m + S.g[0]);
        vy = py / m + dt * (fy / m + S.g[1]);
        vz = pz / m + dt * (fz / m + S.g[2]);
        const float xi = node_coord(S, 0, nx), yi = node_coord(S, 1, ny), zi = node_coord(S, 2, nz);
        collide(xi, yi, zi, vx, vy, vz, cols + S.col_off, S.ncol);
        bx = vx - S.alpha * (px / m);
        by = vy - S.alpha * (py / m);
        bz = vz - S.alpha * (pz / m);
      }
      store_node(g_node, g_mass, node, vx, vy, vz, bx, by, bz, m);
    }
    // active nodes of the scene, counted by lane 0 (node (4K) of the tile: always inside)
    const unsigned long long act = __ballot(m > 0.0f);
    if (lane == 0) cnt_sum += (unsigned long long)__popcll(act);
  }
  if (lane == 0 && cnt_sum && active_count)
    atomicAdd(active_count + size_t(cnt_sid * ACTIVE_SLOTS + cnt_slot) * ACTIVE_STRIDE, cnt_sum);
}

// ---------------------------------------------------------------------------------------
// Dense grid update for the per-stage API (reads g_mmv/g_f of every node of every scene).
__global__ void __launch_bounds__(256) k_grid_update_dense(const SceneDev* __restrict__ scenes, int nscenes,
                                                           const float4* __restrict__ g_mmv,
                                                           const float4* __restrict__ g_f,
                                                           float2* __restrict__ g_node, float* __restrict__ g_mass,
                                                           const Collider* __restrict__ cols, float dt_override,
                                                           int64_t total_nodes) {
  const int64_t node = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (node >= total_nodes) return;
  int sid = 0;
  while (sid + 1 < nscenes && scenes[sid + 1].node_off <= node) ++sid;
  const SceneDev& S = scenes[sid];
  const int64_t l = node - S.node_off;
  const int nx = int(l % S.D[0]), ny = int((l / S.D[0]) % S.D[1]), nz = int(l / (int64_t(S.D[0]) * S.D[1]));
  const float4 mm = g_mmv[node], ff = g_f[node];
  const float m = mm.w;
  const float dt = dt_override > 0.0f ? dt_override : S.dt;
  float vx = 0.f, vy = 0.f, vz = 0.f, bx = 0.f, by = 0.f, bz = 0.f;
  if (m > 0.0f) {
    vx = mm.x / m + dt * (ff.x / m + S.g[0]);
    vy = mm.y / m + dt * (ff.y / m + S.g[1]);
    vz = mm.z / m + dt * (ff.z / m + S.g[2]);
    collide(node_coord(S, 0, nx), node_coord(S, 1, ny), node_coord(S, 2, nz), vx, vy, vz, cols + S.col_off, S.ncol);
    bx = vx - S.alpha * (mm.x / m);
    by = vy - S.alpha * (mm.y / m);
    bz = vz - S.alpha * (mm.z / m);
  }
  store_node(g_node, g_mass, node, vx, vy, vz, bx, by, bz, m);
}

// From dense (m, p) and a host-written v* (float4 (v*, -) per node of one scene): the G2P inputs.
__global__ void __launch_bounds__(256) k_grid_import(const SceneDev* __restrict__ scenes, int sid,
                                                     const float4* __restrict__ g_mmv,
                                                     const float4* __restrict__ vstar, float2* __restrict__ g_node,
                                                     float* __restrict__ g_mass) {
  const SceneDev& S = scenes[sid];
  const int64_t nn = int64_t(S.D[0]) * S.D[1] * S.D[2];
  const int64_t l = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (l >= nn) return;
  const int64_t node = S.node_off + l;
  const float4 mm = g_mmv[node];
  const float m = mm.w;
  if (m > 0.0f) {
    const float4 vs = vstar[l];
    store_node(g_node, g_mass, node, vs.x, vs.y, vs.z, vs.x - S.alpha * (mm.x / m), vs.y - S.alpha * (mm.y / m),
               vs.z - S.alpha * (mm.z / m), m);
  } else {
    store_node(g_node, g_mass, node, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, m);  // gvn = 0 where m <= 0 (gForceCompute.glsl:100-102)
  }
}

// Dense export of one scene's grid (zeros where the tiled path did not touch the node).
__global__ void __launch_bounds__(256) k_grid_export(const SceneDev* __restrict__ scenes, int sid,
                                                     const uint32_t* __restrict__ tile_stamp,
                                                     const StepCtl* __restrict__ ctl, int dense_valid,
                                                     const float4* __restrict__ g_mmv, const float4* __restrict__ g_f,
                                                     const float2* __restrict__ g_node,
                                                     const float* __restrict__ g_mass, int mass_from_node,
                                                     float* __restrict__ out_m,
                                                     float* __restrict__ out_mv, float* __restrict__ out_f,
                                                     float* __restrict__ out_vs) {
  const SceneDev& S = scenes[sid];
  const int64_t nn = int64_t(S.D[0]) * S.D[1] * S.D[2];
  const int64_t l = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (l >= nn) return;
  const int nx = int(l % S.D[0]), ny = int((l / S.D[0]) % S.D[1]), nz = int(l / (int64_t(S.D[0]) * S.D[1]));
  bool live = dense_valid != 0;
  if (!live) {
    const uint32_t step = ctl->step;
    const int Kx = nx >> 2, Ky = ny >> 2, Kz = nz >> 2, rx = nx & 3, ry = ny & 3, rz = nz & 3;
    for (int a = 0; a < 8 && !live; ++a) {
      const int ax = a & 1, ay = (a >> 1) & 1, az = a >> 2;
      if ((ax && rx < 1) || (ay && ry < 1) || (az && rz < 1)) continue;
      live = tile_live(S, tile_stamp, step, Kx + ax, Ky + ay, Kz + az);
    }
  }
  const int64_t node = S.node_off + l;
  const float4 mm = live ? g_mmv[node] : make_float4(0.f, 0.f, 0.f, 0.f);
  const float4 ff = live ? g_f[node] : make_float4(0.f, 0.f, 0.f, 0.f);
  const float2 v01 = live ? g_node[3 * node] : make_float2(0.f, 0.f);
  const float2 v2u = live ? g_node[3 * node + 1] : make_float2(0.f, 0.f);
  const float4 vs = make_float4(v01.x, v01.y, v2u.x, 0.f);
  const float mass = mass_from_node ? (live ? g_mass[node] : 0.f) : mm.w;
  if (out_m) out_m[l] = mass;
  if (out_mv) { out_mv[3 * l] = mm.x; out_mv[3 * l + 1] = mm.y; out_mv[3 * l + 2] = mm.z; }
  if (out_f) { out_f[3 * l] = ff.x; out_f[3 * l + 1] = ff.y; out_f[3 * l + 2] = ff.z; }
  if (out_vs) { out_vs[3 * l] = vs.x; out_vs[3 * l + 1] = vs.y; out_vs[3 * l + 2] = vs.z; }
}

// ---------------------------------------------------------------------------------------
// Per-particle stencil weights of one axis: base = floor(u), offsets o = 0..3 -> node base+o-1,
// w[o] = N(node - u), q[o] = N'(u - node) (cubic.include.glsl:21-40 with the distance vector of
// g2pTransfer.glsl:85-90), from the fixed polynomial pieces in f = u - base.
__device__ __forceinline__ int axis_weights(float u, float w[4], float q[4]) {
  const float fl = floorf(u);
  const float f = u - fl;
#pragma unroll
  for (int o = 0; o < 4; ++o) {
    w[o] = spline_w(o, f);
    q[o] = spline_q(o, f);
  }
  return int(fl);
}

// Gather of one particle from the node records (g2pTransfer.glsl:53-143):
// vpn = sum_i w (v*_i - alpha p_i/m_i), dvp = sum_i v*_i (x) grad w (column k = v* * dw/dx_k).
// Nodes with m <= 0 hold v* = 0 and u = 0 (written so by the grid update), so they add nothing
// and the m_i > 0 test of g2pTransfer.glsl:99 needs no branch. Each (y,z) row of the stencil is
// reduced along x first (separable weights); interior particles (whole stencil in the grid)
// take an unchecked path with one row address per (y,z) and immediate x offsets.
__device__ __forceinline__ void g2p_row(const float2* __restrict__ rowp, const float wx[4], const float qx[4],
                                        float wyz, float qyz, float wqz, float vpn[3], float dvp[9], int xlo,
                                        int xhi) {
  float sq0 = 0.f, sq1 = 0.f, sq2 = 0.f, sw0 = 0.f, sw1 = 0.f, sw2 = 0.f, fw0 = 0.f, fw1 = 0.f, fw2 = 0.f;
#pragma unroll
  for (int ox = 0; ox < 4; ++ox) {
    if (ox < xlo || ox > xhi) continue;
    const float2 n0 = rowp[3 * ox], n1 = rowp[3 * ox + 1], n2 = rowp[3 * ox + 2];
    // v* = (n0.x, n0.y, n1.x), u = (n1.y, n2.x, n2.y)
    sq0 = fmaf(n0.x, qx[ox], sq0); sq1 = fmaf(n0.y, qx[ox], sq1); sq2 = fmaf(n1.x, qx[ox], sq2);
    sw0 = fmaf(n0.x, wx[ox], sw0); sw1 = fmaf(n0.y, wx[ox], sw1); sw2 = fmaf(n1.x, wx[ox], sw2);
    fw0 = fmaf(n1.y, wx[ox], fw0); fw1 = fmaf(n2.x, wx[ox], fw1); fw2 = fmaf(n2.y, wx[ox], fw2);
  }
  vpn[0] = fmaf(fw0, wyz, vpn[0]); vpn[1] = fmaf(fw1, wyz, vpn[1]); vpn[2] = fmaf(fw2, wyz, vpn[2]);
  dvp[0] = fmaf(sq0, wyz, dvp[0]); dvp[1] = fmaf(sq1, wyz, dvp[1]); dvp[2] = fmaf(sq2, wyz, dvp[2]);
  dvp[3] = fmaf(sw0, qyz, dvp[3]); dvp[4] = fmaf(sw1, qyz, dvp[4]); dvp[5] = fmaf(sw2, qyz, dvp[5]);
  dvp[6] = fmaf(sw0, wqz, dvp[6]); dvp[7] = fmaf(sw1, wqz, dvp[7]); dvp[8] = fmaf(sw2, wqz, dvp[8]);
}

__device__ __forceinline__ bool g2p_gather(const SceneDev& S, float x, float y, float z,
                                           const float2* __restrict__ g_node, float vpn[3], float dvp[9],
                                           bool compensated) {
  vpn[0] = vpn[1] = vpn[2] = 0.f;
#pragma unroll
  for (int i = 0; i < 9; ++i) dvp[i] = 0.f;
  const float ux = grid_u(S, x, 0), uy = grid_u(S, y, 1), uz = grid_u(S, z, 2);
  if (!in_stencil_range(S, ux, uy, uz)) return false;
  float wx[4], wy[4], wz[4], qx[4], qy[4], qz[4];
  const int bx = axis_weights(ux, wx, qx);
  const int by = axis_weights(uy, wy, qy);
  const int bz = axis_weights(uz, wz, qz);
  const int D0 = S.D[0], D1 = S.D[1], D2 = S.D[2];
  const bool interior = bx >= 1 && by >= 1 && bz >= 1 && bx + 2 < D0 && by + 2 < D1 && bz + 2 < D2;
  if (compensated) {
    float lv[3] = {0.f, 0.f, 0.f}, ld[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) ld[i] = 0.f;
    for (int oz = 0; oz < 4; ++oz) {
      const int iz = bz + oz - 1;
      if (iz < 0 || iz >= D2) continue;
      for (int oy = 0; oy < 4; ++oy) {
        const int iy = by + oy - 1;
        if (iy < 0 || iy >= D1) continue;
        const int64_t row = S.node_off + int64_t(D0) * (iy + int64_t(D1) * iz);
        const float wyz = wy[oy] * wz[oz], qyz = qy[oy] * wz[oz], wqz = wy[oy] * qz[oz];
        for (int ox = 0; ox < 4; ++ox) {
          const int ix = bx + ox - 1;
          if (ix < 0 || ix >= D0) continue;
          const float2 n0 = g_node[3 * (row + ix)], n1 = g_node[3 * (row + ix) + 1], n2 = g_node[3 * (row + ix) + 2];
          const float4 a = make_float4(n0.x, n0.y, n1.x, n1.y);
          const float4 b = make_float4(n2.x, n2.y, 0.f, 0.f);
          const float w = wx[ox] * wyz;
          const float g0 = qx[ox] * wyz, g1 = wx[ox] * qyz, g2 = wx[ox] * wqz;
          const float c[12] = {a.w * w, b.x * w, b.y * w, a.x * g0, a.y * g0, a.z * g0,
                               a.x * g1, a.y * g1, a.z * g1, a.x * g2, a.y * g2, a.z * g2};
#pragma unroll
          for (int k = 0; k < 3; ++k) two_sum_acc(vpn[k], lv[k], c[k]);
#pragma unroll
          for (int k = 0; k < 9; ++k) two_sum_acc(dvp[k], ld[k], c[3 + k]);
        }
      }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) vpn[k] += lv[k];
#pragma unroll
    for (int k = 0; k < 9; ++k) dvp[k] += ld[k];
  } else if (interior) {
    const float2* p0 = g_node + 3 * (S.node_off + (bx - 1) + int64_t(D0) * ((by - 1) + int64_t(D1) * (bz - 1)));
    const int64_t sy = 3 * int64_t(D0), sz = 3 * int64_t(D0) * D1;
    // two rows in flight (a fully unrolled row loop holds 4 more VGPRs than the fused kernel's 128)
#pragma unroll 1
    for (int oz = 0; oz < 4; ++oz)
#pragma unroll 2
      for (int oy = 0; oy < 4; ++oy)
        g2p_row(p0 + oz * sz + oy * sy, wx, qx, wy[oy] * wz[oz], qy[oy] * wz[oz], wy[oy] * qz[oz], vpn, dvp, 0, 3);
  } else {
    const int xlo = max(0, 1 - bx), xhi = min(3, D0 - bx);  // ox with 0 <= bx + ox - 1 < D0
    for (int oz = 0; oz < 4; ++oz) {
      const int iz = bz + oz - 1;
      if (iz < 0 || iz >= D2) continue;
      for (int oy = 0; oy < 4; ++oy) {
        const int iy = by + oy - 1;
        if (iy < 0 || iy >= D1) continue;
        const float2* rowp = g_node + 3 * (S.node_off + (bx - 1) + int64_t(D0) * (iy + int64_t(D1) * iz));
        g2p_row(rowp, wx, qx, wy[oy] * wz[oz], qy[oy] * wz[oz], wy[oy] * qz[oz], vpn, dvp, xlo, xhi);
      }
    }
  }
  const float gs = S.gscale;
#pragma unroll
  for (int k = 0; k < 9; ++k) dvp[k] *= gs;  // weight gradient scale (1 in the reference)
  return true;
}

// Wave-cooperative form of the interior gather. The 64 particles of a wave are consecutive in
// sorted order, so their stencils cover a small box of nodes: the wave copies that box (at most
// G2P_BOX nodes) from the node records into its LDS slot once, and every lane then gathers its
// 64 nodes from LDS. This moves the 64 x 24 bytes per particle off the vector L1 / address path.
// Waves that are not full, mix scenes, have a particle near the grid boundary or need a larger
// box use g2p_gather. The sums are the same as g2p_gather's, in the same order.
constexpr int G2P_BOX = 384;

__device__ __forceinline__ void g2p_gather_wave(const SceneDev& S, int sid, float x, float y, float z,
                                                const float2* __restrict__ g_node, float2* box, float vpn[3],
                                                float dvp[9]) {
  const float ux = grid_u(S, x, 0), uy = grid_u(S, y, 1), uz = grid_u(S, z, 2);
  const int bx = int(floorf(ux)), by = int(floorf(uy)), bz = int(floorf(uz));
  const int D0 = S.D[0], D1 = S.D[1], D2 = S.D[2];
  const bool interior = in_stencil_range(S, ux, uy, uz) && bx >= 1 && by >= 1 && bz >= 1 && bx + 2 < D0 &&
                        by + 2 < D1 && bz + 2 < D2;
  const uint64_t act = __ballot(true);
  const int sid0 = __builtin_amdgcn_readfirstlane(sid);
  const bool full = act == ~0ull;
  if (!full || __ballot(interior && sid == sid0) != act) {
    g2p_gather(S, x, y, z, g_node, vpn, dvp, false);
    return;
  }
  // node box [xmin, xmin+ex) x [ymin, ymin+ey) x [zmin, zmin+ez)
  const int xmin = wave_min_i(bx) - 1, ymin = wave_min_i(by) - 1, zmin = wave_min_i(bz) - 1;
  const int ex = wave_max_i(bx) + 3 - xmin, ey = wave_max_i(by) + 3 - ymin, ez = wave_max_i(bz) + 3 - zmin;
  const int vol = ex * ey * ez;
  if (vol > G2P_BOX) {
    g2p_gather(S, x, y, z, g_node, vpn, dvp, false);
    return;
  }
  const int lane = threadIdx.x & 63;
  // the box's first node; node offsets within the box fit 32 bits (D0 D1 D2 < 2^31)
  const float2* g0 = g_node + 3 * (S.node_off + xmin + int64_t(D0) * (ymin + int64_t(D1) * zmin));
  const BoxSplit bs = box_split(ex, ey);
  // every load of the box in flight before the first LDS write
  constexpr int BU = G2P_BOX / 64;
  float2 ta[BU], tb[BU], tc[BU];
#pragma unroll
  for (int u = 0; u < BU; ++u) {
    const int idx = lane + 64 * u;
    if (idx < vol) {
      int ix, iy, iz;
      box_coords(bs, idx, ix, iy, iz);
      const float2* src = g0 + 3u * uint32_t(ix + D0 * (iy + D1 * iz));
      ta[u] = src[0];
      tb[u] = src[1];
      tc[u] = src[2];
    }
  }
#pragma unroll
  for (int u = 0; u < BU; ++u) {
    const int idx = lane + 64 * u;
    if (idx < vol) {
      box[3 * idx] = ta[u];
      box[3 * idx + 1] = tb[u];
      box[3 * idx + 2] = tc[u];
    }
  }
  // the wave's own LDS writes complete before its reads (in-order LDS per wave)
  __builtin_amdgcn_wave_barrier();
  float wx[4], wy[4], wz[4], qx[4], qy[4], qz[4];
  axis_weights(ux, wx, qx);
  axis_weights(uy, wy, qy);
  axis_weights(uz, wz, qz);
  vpn[0] = vpn[1] = vpn[2] = 0.f;
#pragma unroll
  for (int k = 0; k < 9; ++k) dvp[k] = 0.f;
  const float2* p0 = box + 3 * ((bx - 1 - xmin) + ex * ((by - 1 - ymin) + ey * (bz - 1 - zmin)));
  const int sy = 3 * ex, sz = 3 * ex * ey;
  // two rows of LDS reads in flight: enough for the LDS latency, and it leaves the registers that
  // four rows would hold to the rest of the kernel
#pragma unroll 1
  for (int oz = 0; oz < 4; ++oz)
#pragma unroll 2
    for (int oy = 0; oy < 4; ++oy)
      g2p_row(p0 + oz * sz + oy * sy, wx, qx, wy[oy] * wz[oz], qy[oy] * wz[oz], wy[oy] * qz[oz], vpn, dvp, 0, 3);
  const float gs = S.gscale;
#pragma unroll
  for (int k = 0; k < 9; ++k) dvp[k] *= gs;  // weight gradient scale (1 in the reference)
}

// particleAdvance.glsl:95-140 in precision R: F_E, F_P update with the singular-value clamp.
// Wo, Vo receive the SVD factors of F_E' = (I + dt L) F_E and so[3] its diagonal factor before the
// clamp (S[0][0], S[1][1], S[2][2]).
template <class R>
__device__ __forceinline__ void deform_update(const float dvp[9], float dt, const float FEf[9], const float FPf[9],
                                              float crit_c, float crit_s, float FEo[9], float FPo[9], R Wo[9],
                                              R Vo[9], R so[3]) {
  R L[9], FE[9], FP[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) { L[i] = R(dt) * R(dvp[i]); FE[i] = R(FEf[i]); FP[i] = R(FPf[i]); }
  L[0] += R(1); L[4] += R(1); L[8] += R(1);
  // F_E' = (I + dt L) F_E and F' = (I + dt L) F_E F_P = F_E' F_P (particleAdvance.glsl:95-96)
  R FEn[9], Fn[9];
  mul(L, FE, FEn);
  mul(FEn, FP, Fn);
  R W[9], S[9], V[9];
  svd3(FEn, W, S, V);
  so[0] = S[0]; so[1] = S[4]; so[2] = S[8];
  const R lb = R(1) - R(crit_c), ub = R(1) + R(crit_s);
  S[0] = S[0] < lb ? lb : (S[0] > ub ? ub : S[0]);
  S[4] = S[4] < lb ? lb : (S[4] > ub ? ub : S[4]);
  S[8] = S[8] < lb ? lb : (S[8] > ub ? ub : S[8]);
  R WS[9], Fe2[9];
  mul(W, S, WS);
  mul_bt(WS, V, Fe2);
  // F_P = V diag(1/S) W^T F  (particleAdvance.glsl:128-132)
  R VSi[9];
  const R i0 = recip(S[0]), i1 = recip(S[4]), i2 = recip(S[8]);
#pragma unroll
  for (int r = 0; r < 3; ++r) { VSi[r] = V[r] * i0; VSi[3 + r] = V[3 + r] * i1; VSi[6 + r] = V[6 + r] * i2; }
  R VSiWt[9], Fp2[9];
  mul_bt(VSi, W, VSiWt);
  mul(VSiWt, Fn, Fp2);
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    FEo[i] = float(round_to(Fe2[i], R(1e9), R(1e-9)));
    FPo[i] = float(round_to(Fp2[i], R(1e9), R(1e-9)));
    Wo[i] = W[i];
    Vo[i] = V[i];
  }
}

// Mixed precision: a particle whose F_E' has a singular value (the diagonal d of the FP32 update's
// own SVD, before the clamp) within 2e-4 of 1 - theta_c or 1 + theta_s, or whose F_E' is ill-
// conditioned (min |d| <= 1e-3 max |d|), is updated again on the FP64 path.
__device__ __forceinline__ bool needs_fp64(const float d[3], float crit_c, float crit_s) {
  const float tol = 2e-4f;
  const float lb = 1.f - crit_c, ub = 1.f + crit_s;
  bool near = false;
#pragma unroll
  for (int k = 0; k < 3; ++k) near |= fabsf(fabsf(d[k]) - lb) < tol || fabsf(fabsf(d[k]) - ub) < tol;
  const float smax = fmaxf(fabsf(d[0]), fmaxf(fabsf(d[1]), fabsf(d[2])));
  const float smin = fminf(fabsf(d[0]), fminf(fabsf(d[1]), fabsf(d[2])));
  near |= !(smin > 1e-3f * smax);
  return near;
}

// The next sort's first-pass histogram (rows of 2^hbits digits per tile of SORT_TILE_KEYS keys, zero
// on entry): the wave's active lanes are a prefix of consecutive positions, and each run of equal
// (tile, digit) among them adds its length with one atomic from its first lane. Sorted order makes
// the runs long, so a wave adds a few counts instead of 64.
__device__ __forceinline__ void count_first_digit(uint32_t* __restrict__ hist0, int hbits, uint32_t i,
                                                  uint32_t key) {
  const int lane = int(threadIdx.x & 63);
  const uint32_t tile = i / uint32_t(SORT_TILE_KEYS), d = key & ((1u << hbits) - 1u);
  const uint32_t ptile = uint32_t(__shfl_up(int(tile), 1, 64)), pd = uint32_t(__shfl_up(int(d), 1, 64));
  const bool head = lane == 0 || tile != ptile || d != pd;
  const uint64_t heads = __ballot(head);
  const int nact = __popcll(__ballot(true));
  if (head) {
    const uint64_t later = lane == 63 ? 0ull : heads & (~0ull << (lane + 1));
    const int end = later ? __ffsll((long long)later) - 1 : nact;
    atomicAdd(hist0 + size_t(tile) * (size_t(1) << hbits) + d, uint32_t(end - lane));
  }
}

// ---------------------------------------------------------------------------------------
// Least waves per SIMD the G2P register allocation must keep: 4 for the FP32 kernel (at most 128
// VGPRs), 3 with the FP64 path compiled in (at most 168 VGPRs, so that the FP64 branch never
// spills; the compiler currently fits it in 128).
template <bool MIXED> struct G2PWaves { static constexpr int value = MIXED ? 3 : 4; };

// Fused G2P + particle advance (g2pTransfer.glsl:53-143, particleAdvance.glsl:75-217).
// Thread i handles sorted position i: it reads the particle from its previous slot
// vals[i] (buffers *_in), writes the new state to slot i (buffers *_out), computes the next
// step's force matrix (from the SVD of the F_E update) into slot i of fa and the next step's cell
// key. The permutation of the sort is thereby applied for free by the writes of this kernel.
// The force matrix is not read here: fa is the next P2G's input only.
// Every wave holds particles of one scene, so the scene record is wave-uniform (scalar loads):
// waves[w] = (first sorted position, scene) of wave w, each scene's sorted range cut into waves of
// 64 positions (its last one partial); null for a context of one scene, whose wave w starts at 64w.
template <bool MIXED, bool COMP>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(G2PWaves<MIXED>::value))) k_g2p_advance(
    const SceneDev* __restrict__ scenes, const uint2* __restrict__ waves, uint32_t nwaves,
    const uint32_t* __restrict__ vals,
    const float4* __restrict__ rec_in, const float* __restrict__ fe_in, const float* __restrict__ fp_in,
    const uint32_t* __restrict__ pid_in, float4* __restrict__ rec_out, float* __restrict__ fe_out,
    float* __restrict__ fp_out, uint32_t* __restrict__ pid_out, float* __restrict__ fa_out,
    uint32_t* __restrict__ keys_next, const float2* __restrict__ g_node, const Collider* __restrict__ cols,
    float dt_override, uint32_t n, unsigned long long* __restrict__ fp64_count, const StepCtl* __restrict__ ctl,
    uint32_t* __restrict__ pbound, uint32_t* __restrict__ hist0, int hbits) {
  __shared__ float2 s_box[4][3 * G2P_BOX];  // one node box per wave (blocks of up to 256 threads)
  __shared__ uint32_t s_pb[4];               // block maxima of the P2G bound, waves done
  const uint32_t wpb = blockDim.x >> 6;      // waves per block
  const uint32_t w0 = blockIdx.x * wpb;
  const uint32_t wb = w0 + wpb <= nwaves ? wpb : nwaves - w0;  // live waves of the block
  if (threadIdx.x < 4) s_pb[threadIdx.x] = 0u;
  // all waves of the block in one scene (the scene of a wave is constant along the table)
  const bool one_scene = !waves || waves[w0].y == waves[w0 + wb - 1].y;
  const uint32_t next_parity = (ctl->step + 1u) & 1u;
  __syncthreads();
  const uint32_t w = w0 + uint32_t(__builtin_amdgcn_readfirstlane(threadIdx.x >> 6));
  if (w >= nwaves) return;
  const uint2 ws = waves ? waves[w] : make_uint2(64u * w, 0u);
  const int s0 = int(ws.y);
  const uint32_t i = ws.x + (threadIdx.x & 63);
  if (int64_t(i) >= scenes[s0].p_off + scenes[s0].p_cnt) return;
  {
    const SceneDev& S = scenes[s0];
    const float dt = dt_override > 0.0f ? dt_override : S.dt;
    const uint32_t p = vals[i];
    const float4 xm = rec_in[2 * size_t(p)];      // x, m
    const float4 vr = rec_in[2 * size_t(p) + 1];  // v, rho0 (in flight through the gather)
    const uint32_t id = pid_in[p];

    float vpn[3], dvp[9];
    if (COMP)
      g2p_gather(S, xm.x, xm.y, xm.z, g_node, vpn, dvp, true);
    else
      g2p_gather_wave(S, s0, xm.x, xm.y, xm.z, g_node, s_box[threadIdx.x >> 6], vpn, dvp);
    // F_E, F_P are loaded after the gather: not live through it, and in flight behind the
    // velocity update below
    float FE[9], FP[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) { FE[e] = ld_plane(fe_in + size_t(e) * n, p); FP[e] = ld_plane(fp_in + size_t(e) * n, p); }

    // velocity, collisions at the pre-advection position, advection (particleAdvance.glsl:180-210);
    // independent of the F_E/F_P update below, so done and stored first (fewer live registers)
    float vx = vr.x * S.alpha + vpn[0], vy = vr.y * S.alpha + vpn[1], vz = vr.z * S.alpha + vpn[2];
    collide(xm.x, xm.y, xm.z, vx, vy, vz, cols + S.col_off, S.ncol);
    const float x = xm.x + dt * vx, y = xm.y + dt * vy, z = xm.z + dt * vz;
    rec_out[2 * size_t(i)] = make_float4(x, y, z, xm.w);
    rec_out[2 * size_t(i) + 1] = make_float4(vx, vy, vz, vr.w);
    pid_out[i] = id;
    const uint32_t key = cell_key(S, x, y, z);
    keys_next[i] = key;
    if (hist0) count_first_digit(hist0, hbits, i, key);

    // F_E, F_P update; the next step's force matrix (p2gTransfer.glsl:109-141, once per particle)
    // takes R_E = W V^T from the factors of this update's SVD: the stored F_E is W S' V^T with
    // S' the clamped triangular factor, so its own SVD gives the same W V^T.
    float FEo[9], FPo[9], A[9];
    float W[9], V[9], sv[3];
    deform_update<float>(dvp, dt, FE, FP, S.crit_comp, S.crit_stretch, FEo, FPo, W, V, sv);
    const bool use64 = MIXED && needs_fp64(sv, S.crit_comp, S.crit_stretch);
    if (use64) {
      double Wd[9], Vd[9], svd[3], FEd[9], FPd[9], Ad[9];
      deform_update<double>(dvp, dt, FE, FP, S.crit_comp, S.crit_stretch, FEo, FPo, Wd, Vd, svd);
#pragma unroll
      for (int e = 0; e < 9; ++e) { FEd[e] = FEo[e]; FPd[e] = FPo[e]; }
      force_matrix_wv<double>(FEd, FPd, Wd, Vd, double(xm.w), double(vr.w), double(S.mu0), double(S.lambda0),
                              double(S.hardening), double(S.gscale), Ad);
#pragma unroll
      for (int e = 0; e < 9; ++e) A[e] = float(Ad[e]);
    } else {
      force_matrix_wv<float>(FEo, FPo, W, V, xm.w, vr.w, S.mu0, S.lambda0, S.hardening, S.gscale, A);
    }
    if (MIXED && fp64_count) {  // one add per wave: its lanes that took the FP64 path
      const unsigned long long b64 = __ballot(use64);
      if (b64 && int(threadIdx.x & 63) == __ffsll((long long)__ballot(true)) - 1)
        atomicAdd(fp64_count, (unsigned long long)__popcll(b64));
    }
#pragma unroll
    for (int e = 0; e < 9; ++e) { st_plane(fe_out + size_t(e) * n, i, FEo[e]); st_plane(fp_out + size_t(e) * n, i, FPo[e]); }
    store_force(fa_out, n, i, A);
    const float Arm[9] = {A[0], A[3], A[6], A[1], A[4], A[7], A[2], A[5], A[8]};
    publish_bound_block(pbound, next_parity, s0, record_bound(S, key, xm.w, vx, vy, vz, Arm, 1), s_pb, one_scene,
                        wb);
  }
}

// Split stages for the per-stage API: G2P alone writes (vpn, dvp) per particle slot;
// the advance alone consumes them in place (g2pTransfer.glsl / particleAdvance.glsl).
__global__ void __launch_bounds__(256) k_g2p_only(const SceneDev* __restrict__ scenes,
                                                  const uint16_t* __restrict__ pscene,
                                                  const float4* __restrict__ rec, const float2* __restrict__ g_node,
                                                  float* __restrict__ aux,
                                                  uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const SceneDev& S = scenes[pscene[i]];
  const float4 r0 = rec[2 * size_t(i)];
  float vpn[3], dvp[9];
  g2p_gather(S, r0.x, r0.y, r0.z, g_node, vpn, dvp, false);
#pragma unroll
  for (int k = 0; k < 3; ++k) aux[size_t(k) * n + i] = vpn[k];
#pragma unroll
  for (int k = 0; k < 9; ++k) aux[size_t(3 + k) * n + i] = dvp[k];
}

__global__ void __launch_bounds__(256) k_advance_only(const SceneDev* __restrict__ scenes,
                                                      const uint16_t* __restrict__ pscene, float4* __restrict__ rec,
                                                      float* __restrict__ fe, float* __restrict__ fp,
                                                      float* __restrict__ fa, const float* __restrict__ aux,
                                                      const Collider* __restrict__ cols, float dt_override,
                                                      uint32_t* __restrict__ keys,
                                                      uint32_t n, const StepCtl* __restrict__ ctl,
                                                      uint32_t* __restrict__ pbound) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int sid = pscene[i];
  const SceneDev& S = scenes[sid];
  const float dt = dt_override > 0.0f ? dt_override : S.dt;
  const float4 xm = rec[2 * size_t(i)];
  const float4 vr = rec[2 * size_t(i) + 1];
  float FE[9], FP[9], vpn[3], dvp[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) { FE[e] = fe[size_t(e) * n + i]; FP[e] = fp[size_t(e) * n + i]; }
#pragma unroll
  for (int k = 0; k < 3; ++k) vpn[k] = aux[size_t(k) * n + i];
#pragma unroll
  for (int k = 0; k < 9; ++k) dvp[k] = aux[size_t(3 + k) * n + i];
  float FEo[9], FPo[9], W[9], V[9], sv[3];
  deform_update<float>(dvp, dt, FE, FP, S.crit_comp, S.crit_stretch, FEo, FPo, W, V, sv);
  float vx = vr.x * S.alpha + vpn[0], vy = vr.y * S.alpha + vpn[1], vz = vr.z * S.alpha + vpn[2];
  collide(xm.x, xm.y, xm.z, vx, vy, vz, cols + S.col_off, S.ncol);
  const float x = xm.x + dt * vx, y = xm.y + dt * vy, z = xm.z + dt * vz;
  float A[9];
  force_matrix_wv<float>(FEo, FPo, W, V, xm.w, vr.w, S.mu0, S.lambda0, S.hardening, S.gscale, A);
  rec[2 * size_t(i)] = make_float4(x, y, z, xm.w);
  rec[2 * size_t(i) + 1] = make_float4(vx, vy, vz, vr.w);
#pragma unroll
  for (int e = 0; e < 9; ++e) { fe[size_t(e) * n + i] = FEo[e]; fp[size_t(e) * n + i] = FPo[e]; }
  store_force(fa, n, i, A);
  const uint32_t key = cell_key(S, x, y, z);
  keys[i] = key;
  const float Arm[9] = {A[0], A[3], A[6], A[1], A[4], A[7], A[2], A[5], A[8]};
  publish_bound(pbound, (ctl->step + 1u) & 1u, sid, record_bound(S, key, xm.w, vx, vy, vz, Arm, 1));
}

// ---------------------------------------------------------------------------------------
// Init: rest density rho_p0 = sum_i m_i w_ip / h^3 (pVolumeCompute.glsl:32-83), gathered per
// particle (the reference scatters with a 64-way per-particle atomic); stored in rec[2i+1].w.
__global__ void __launch_bounds__(256) k_particle_volume(const SceneDev* __restrict__ scenes,
                                                         const uint16_t* __restrict__ pscene,
                                                         float4* __restrict__ rec,
                                                         const float4* __restrict__ g_mmv, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const SceneDev& S = scenes[pscene[i]];
  const float4 r0 = rec[2 * size_t(i)];
  const float ux = grid_u(S, r0.x, 0), uy = grid_u(S, r0.y, 1), uz = grid_u(S, r0.z, 2);
  float acc = 0.f;
  if (in_stencil_range(S, ux, uy, uz)) {
    float wx[4], wy[4], wz[4], q[4];
    const int bx = axis_weights(ux, wx, q);
    const int by = axis_weights(uy, wy, q);
    const int bz = axis_weights(uz, wz, q);
    const float cell = S.h * S.h * S.h;
    for (int oz = 0; oz < 4; ++oz) {
      const int iz = bz + oz - 1;
      if (iz < 0 || iz >= S.D[2]) continue;
      for (int oy = 0; oy < 4; ++oy) {
        const int iy = by + oy - 1;
        if (iy < 0 || iy >= S.D[1]) continue;
        for (int ox = 0; ox < 4; ++ox) {
          const int ix = bx + ox - 1;
          if (ix < 0 || ix >= S.D[0]) continue;
          const float mi = g_mmv[S.node_off + ix + int64_t(S.D[0]) * (iy + int64_t(S.D[1]) * iz)].w;
          acc += mi * (wx[ox] * wy[oy] * wz[oz]) / cell;
        }
      }
    }
  }
  reinterpret_cast<float*>(rec)[8 * size_t(i) + 7] = acc;
}

// Force matrices of the current state (after init or host writes).
__global__ void __launch_bounds__(256) k_force_prepass(const SceneDev* __restrict__ scenes,
                                                       const uint16_t* __restrict__ pscene,
                                                       const float4* __restrict__ rec, const float* __restrict__ fe,
                                                       const float* __restrict__ fp, float* __restrict__ fa,
                                                       uint32_t n, const StepCtl* __restrict__ ctl,
                                                       uint32_t* __restrict__ pbound) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int sid = pscene[i];
  const SceneDev& S = scenes[sid];
  float FE[9], FP[9], A[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) { FE[e] = fe[size_t(e) * n + i]; FP[e] = fp[size_t(e) * n + i]; }
  const float4 xm = rec[2 * size_t(i)], vr = rec[2 * size_t(i) + 1];
  force_matrix<float>(FE, FP, xm.w, vr.w, S.mu0, S.lambda0, S.hardening, S.gscale, A);
  store_force(fa, n, i, A);
  const float Arm[9] = {A[0], A[3], A[6], A[1], A[4], A[7], A[2], A[5], A[8]};
  publish_bound(pbound, (ctl->step + 1u) & 1u, sid,
                record_bound(S, cell_key(S, xm.x, xm.y, xm.z), xm.w, vr.x, vr.y, vr.z, Arm, 1));
}

// Zero the force matrices (init: the grid-mass pass needs m only).
__global__ void __launch_bounds__(256) k_zero_force(const SceneDev* __restrict__ scenes,
                                                    const uint16_t* __restrict__ pscene,
                                                    const float4* __restrict__ rec, float* __restrict__ fa,
                                                    uint32_t n, const StepCtl* __restrict__ ctl,
                                                    uint32_t* __restrict__ pbound) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int sid = pscene[i];
  const SceneDev& S = scenes[sid];
  const float4 xm = rec[2 * size_t(i)], vr = rec[2 * size_t(i) + 1];
  const float Arm[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  store_force(fa, n, i, Arm);
  // the grid-mass P2G keeps particles of any mass (gMassCompute.glsl has no |m| > 100 skip)
  publish_bound(pbound, (ctl->step + 1u) & 1u, sid,
                record_bound(S, cell_key(S, xm.x, xm.y, xm.z), xm.w, vr.x, vr.y, vr.z, Arm, 0));
}

// ---------------------------------------------------------------------------------------
// Host transfer helpers: original order <-> current slot order through pid.
__global__ void __launch_bounds__(256) k_gather_out(const float4* __restrict__ rec, const float* __restrict__ fe,
                                                    const float* __restrict__ fp, const uint32_t* __restrict__ pid,
                                                    int64_t p_off, int64_t cnt, uint32_t n, float* __restrict__ ox,
                                                    float* __restrict__ ov, float* __restrict__ ofe,
                                                    float* __restrict__ ofp, float* __restrict__ orho,
                                                    float* __restrict__ om) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= cnt) return;
  const uint32_t i = uint32_t(p_off + k);
  const int64_t o = int64_t(pid[i]) - p_off;
  const float4 xm = rec[2 * size_t(i)], vr = rec[2 * size_t(i) + 1];
  ox[3 * o] = xm.x; ox[3 * o + 1] = xm.y; ox[3 * o + 2] = xm.z;
  om[o] = xm.w;
  ov[3 * o] = vr.x; ov[3 * o + 1] = vr.y; ov[3 * o + 2] = vr.z;
#pragma unroll
  for (int e = 0; e < 9; ++e) { ofe[9 * o + e] = fe[size_t(e) * n + i]; ofp[9 * o + e] = fp[size_t(e) * n + i]; }
  orho[o] = vr.w;
}

// The particles with creation index in [first, first + count) of one scene (p_off, cnt), from
// the current slot order, into window order: o = pid - p_off - first.
__global__ void __launch_bounds__(256) k_gather_window(const float4* __restrict__ rec, const float* __restrict__ fe,
                                                       const float* __restrict__ fp, const uint32_t* __restrict__ pid,
                                                       int64_t p_off, int64_t cnt, uint32_t n, int64_t first,
                                                       int64_t count, float* __restrict__ ox, float* __restrict__ ov,
                                                       float* __restrict__ ofe, float* __restrict__ ofp,
                                                       float* __restrict__ orho) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= cnt) return;
  const uint32_t i = uint32_t(p_off + k);
  const int64_t o = int64_t(pid[i]) - p_off - first;
  if (o < 0 || o >= count) return;
  const float4 xm = rec[2 * size_t(i)], vr = rec[2 * size_t(i) + 1];
  ox[3 * o] = xm.x; ox[3 * o + 1] = xm.y; ox[3 * o + 2] = xm.z;
  ov[3 * o] = vr.x; ov[3 * o + 1] = vr.y; ov[3 * o + 2] = vr.z;
#pragma unroll
  for (int e = 0; e < 9; ++e) { ofe[9 * o + e] = fe[size_t(e) * n + i]; ofp[9 * o + e] = fp[size_t(e) * n + i]; }
  orho[o] = vr.w;
}

// mask bits: 1 x, 2 v, 4 fe, 8 fp, 16 rho, 32 mass
__global__ void __launch_bounds__(256) k_scatter_in(float4* __restrict__ rec, float* __restrict__ fe,
                                                    float* __restrict__ fp, const uint32_t* __restrict__ pid,
                                                    int64_t p_off, int64_t cnt, uint32_t n,
                                                    const float* __restrict__ ix, const float* __restrict__ iv,
                                                    const float* __restrict__ ife, const float* __restrict__ ifp,
                                                    const float* __restrict__ irho, const float* __restrict__ im,
                                                    int mask) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= cnt) return;
  const uint32_t i = uint32_t(p_off + k);
  const int64_t o = int64_t(pid[i]) - p_off;
  float4 xm = rec[2 * size_t(i)], vr = rec[2 * size_t(i) + 1];
  if (mask & 1) { xm.x = ix[3 * o]; xm.y = ix[3 * o + 1]; xm.z = ix[3 * o + 2]; }
  if (mask & 32) xm.w = im[o];
  if (mask & 2) { vr.x = iv[3 * o]; vr.y = iv[3 * o + 1]; vr.z = iv[3 * o + 2]; }
  if (mask & 16) vr.w = irho[o];
  rec[2 * size_t(i)] = xm;
  rec[2 * size_t(i) + 1] = vr;
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    if (mask & 4) fe[size_t(e) * n + i] = ife[9 * o + e];
    if (mask & 8) fp[size_t(e) * n + i] = ifp[9 * o + e];
  }
}

// Non-finite scan of the particle state (record, F_E, F_P and the force matrix).
__global__ void __launch_bounds__(256) k_check_finite(const float4* __restrict__ rec, const float* __restrict__ fe,
                                                      const float* __restrict__ fp, const float* __restrict__ fa,
                                                      uint32_t n, int* __restrict__ bad) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  bool ok = true;
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const float4 v = rec[2 * size_t(i) + r];
    ok &= isfinite(v.x) && isfinite(v.y) && isfinite(v.z) && isfinite(v.w);
  }
#pragma unroll
  for (int e = 0; e < 9; ++e)
    ok &= isfinite(fe[size_t(e) * n + i]) && isfinite(fp[size_t(e) * n + i]) && isfinite(fa[size_t(e) * n + i]);
  if (!ok) atomicOr(bad, 1);
}

// Fill kernel (used instead of hipMemsetAsync so the step stays graph-capturable).
__global__ void __launch_bounds__(256) k_fill_u32(uint32_t* __restrict__ p, uint32_t v, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace snow
