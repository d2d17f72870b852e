// snow_implicit.hpp - semi-implicit grid velocity solve of the snow MPM step for gfx950 (CDNA4).
//
// The reference declares SemiImplicitTimeUpdate (src/snow/simulation/timeUpdate/
// semiimplicittimeupdate.hpp:6-13) with empty init/update bodies (semiimplicittimeupdate.cpp:3-5).
// This is the published method the reference's explicit step follows: Stomakhin et al. 2013,
// "A material point method for snow simulation", eq. (9) with its Hessian action, for the energy
// whose gradient is the explicit force of this code (fixed corotated with hardening, the same
// grad w and grad_scale):
//   K v = M v - beta dt df(dt v) = M v*,        df_i(du) = -sum_p V_p0 A_p F_Ep^T grad w_ip,
//   A_p = d2Psi/dF2 (F_Ep) : dF_p,              dF_p = sum_j du_j grad w_jp^T F_Ep,
//   d2Psi/dF2 : dF = 2 mu (dF - dR) + lambda (cof F : dF) cof F + lambda (J - 1) d(cof F).
// R is the rotation the explicit force takes (W V^T of the reference SVD, rounded to 1e-9);
// R^T dR = [w]x with (tr S I - S) w = axial(R^T dF - dF^T R), S the symmetric part of R^T F.
// K is symmetric; conjugate residuals (the paper's solver), preconditioned by the lumped mass
// (z = M^-1 r), run on the grid from v = v*, then the transfer back to the particles takes v instead
// of v*. The node masses span orders of magnitude at the surface of the material: without the
// preconditioner the residual criterion is met before the light nodes move from v*.
//
// Layout: dense node vectors over the context's nodes (float4 per node: x, r, p, Kp, Kz, z), the node
// masses of this step (0 outside the node tiles the grid update wrote or where m <= 0), a dense
// fixed-point accumulator (3 x int64 per node) for df, and per sorted particle 6 float4 planes of
// data prepared once per step (x, gs V_p0, F_E, R, 2 mu, lambda) plus G_p = gs V_p0 A_p F_Ep^T per
// application (two float4 planes and one float plane).
// Sums of df are exact: each wave reduces its particles per cell in registers and adds the
// rounded fixed-point values with integer atomics, so the result does not depend on their order.
// Dot products are block partials summed in a fixed order, so a solve is bit-reproducible.
#pragma once
#include "snow_kernels.hpp"

namespace snow {

// Per-scene solver state in device memory.
struct ImpScal {
  double rho;    // z . K z of the current update
  double alpha;  // step length of the current update
  double beta;   // weight of the previous direction in the current update
  double bb;     // |M v*|^2
  double rr;     // |r|^2
  double tol2;   // squared relative residual at which the solve stops
  double pp;     // Kp . M^-1 Kp of the current direction
  uint32_t spare;
  int32_t iters; // direction updates done
  int32_t done;  // converged, broke down or out of iterations
  int32_t max_iter;
  float beta_dt; // beta * dt
  float dt;
  uint32_t pad[2];
};

enum ImpPhase { IMP_INIT = 0, IMP_KR = 1, IMP_PX = 2 };

constexpr int IMP_THREADS = 1024;
constexpr int IMP_NODES_PER_THREAD = 4;
constexpr int IMP_NODES_PER_BLOCK = IMP_THREADS * IMP_NODES_PER_THREAD;
constexpr int IMP_VECS = 6;       // x, r, p, Kp, Kz, z
// float4 planes of prepared particle data: (x, y, z, gs V_p0), (F0..F3), (F4..F7), (F8, R0, R1, R2),
// (R3..R6), (R7, R8, 2 mu, lambda); F and R column-major
constexpr int IMP_PLANES = 6;
constexpr int IMP_STAGE = 36;     // LDS floats per staged particle in the scatter
constexpr int IMP_BATCH = 16;     // particles a scatter wave stages at a time
constexpr int IMP_MAX_DIM = 1021; // largest grid dimension: base cells + 2 fit 10 bits per axis

// A node-vector block: scene and first node (the blocks of a scene are consecutive).
struct ImpBlock {
  uint32_t scene;
  uint32_t pad;
  int64_t first;
};

__device__ __forceinline__ void cross_add(const float* a, const float* b, float* o) {
  o[0] += a[1] * b[2] - a[2] * b[1];
  o[1] += a[2] * b[0] - a[0] * b[2];
  o[2] += a[0] * b[1] - a[1] * b[0];
}

// Fixed-point exponent of a scene's df sums: every node sums at most N_s terms of magnitude at most
// gmax * 8/27 (|N| <= 2/3, |N'| <= 2/3, so each weight-gradient component is at most 8/27), so with
// 2^e * N_s * gmax * 8/27 < 2^62 no sum leaves the int64 range.
__device__ __forceinline__ int imp_exponent(uint32_t gmax_bits, int64_t count) {
  const double bound = double(__uint_as_float(gmax_bits)) * (8.0 / 27.0) * double(count > 0 ? count : 1);
  if (!(bound > 0.0)) return 0;
  int ex;
  (void)frexp(bound, &ex);  // bound < 2^ex
  return 62 - ex;
}

// (tr S I - S)^-1 as (m00, m11, m22, m01, m02, m12), S the symmetric part of R^T F, from the
// adjugate (tr S I - S is symmetric, about 2 I for F near I); 0 where it is singular.
__device__ __forceinline__ void imp_spin_inverse(const float F[9], const float Rot[9], float mi[6]) {
  float Sm[9];
  mul_at(Rot, F, Sm);  // R^T F
  const float s00 = Sm[0], s11 = Sm[4], s22 = Sm[8];
  const float s01 = 0.5f * (Sm[1] + Sm[3]), s02 = 0.5f * (Sm[2] + Sm[6]), s12 = 0.5f * (Sm[5] + Sm[7]);
  const float t00 = s11 + s22, t11 = s00 + s22, t22 = s00 + s11, t01 = -s01, t02 = -s02, t12 = -s12;
  const float a00 = t11 * t22 - t12 * t12, a11 = t00 * t22 - t02 * t02, a22 = t00 * t11 - t01 * t01;
  const float a01 = t02 * t12 - t01 * t22, a02 = t01 * t12 - t02 * t11, a12 = t01 * t02 - t00 * t12;
  const float det = t00 * a00 + t01 * a01 + t02 * a02;
  const float id = fabsf(det) > 1e-20f ? 1.0f / det : 0.0f;
  mi[0] = a00 * id; mi[1] = a11 * id; mi[2] = a22 * id;
  mi[3] = a01 * id; mi[4] = a02 * id; mi[5] = a12 * id;
}

// ---------------------------------------------------------------------------------------
// Once per step, per sorted position: x, gs V_p0, F_E, R, 2 mu and lambda (the planes of pd). Waves
// hold particles of one scene (the G2P wave table).
__global__ void __launch_bounds__(256) k_imp_prep(const SceneDev* __restrict__ scenes, const uint2* __restrict__ waves,
                                                  uint32_t nwaves, const uint32_t* __restrict__ vals,
                                                  const float4* __restrict__ rec, const float* __restrict__ fe,
                                                  const float* __restrict__ fp, uint32_t n,
                                                  float4* __restrict__ pd, int skip_heavy) {
  const uint32_t w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (w >= nwaves) return;
  const uint2 ws = waves ? waves[w] : make_uint2(64u * w, 0u);
  const SceneDev& S = scenes[ws.y];
  const uint32_t i = ws.x + (threadIdx.x & 63);
  if (int64_t(i) >= S.p_off + S.p_cnt) return;
  const uint32_t p = vals[i];
  const float4 xm = rec[2 * size_t(p)], vr = rec[2 * size_t(p) + 1];
  float F[9], FP[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) { F[e] = ld_plane(fe + size_t(e) * n, p); FP[e] = ld_plane(fp + size_t(e) * n, p); }
  const float hf = hardening_factor(S.hardening, det3(FP));
  const float mu = S.mu0 * hf, la = S.lambda0 * hf;
  float W[9], Sg[9], V[9], Rot[9];
  svd3(F, W, Sg, V);
  mul_bt(W, V, Rot);
#pragma unroll
  for (int e = 0; e < 9; ++e) Rot[e] = round_to(Rot[e], 1e9f, 1e-9f);
  const float m = xm.w, r0 = vr.w;
  const bool use = r0 > 0.0f && !(skip_heavy && fabsf(m) > 100.0f);  // as the explicit force
  const float v0 = use ? S.gscale * (m / r0) : 0.0f;
  const size_t N = n;
  pd[i] = make_float4(xm.x, xm.y, xm.z, v0);
  pd[N + i] = make_float4(F[0], F[1], F[2], F[3]);
  pd[2 * N + i] = make_float4(F[4], F[5], F[6], F[7]);
  pd[3 * N + i] = make_float4(F[8], Rot[0], Rot[1], Rot[2]);
  pd[4 * N + i] = make_float4(Rot[3], Rot[4], Rot[5], Rot[6]);
  pd[5 * N + i] = make_float4(Rot[7], Rot[8], 2.0f * mu, la);
}

// ---------------------------------------------------------------------------------------
// Node tiles the grid update wrote this step: node tile K is written when one of the cell tiles
// K + {0,1}^3 holds particles (the rule of k_grid). One thread per node tile of the context.
__global__ void __launch_bounds__(256) k_imp_ntiles(const SceneDev* __restrict__ scenes,
                                                    const uint16_t* __restrict__ ntile_scene, uint32_t total_ntiles,
                                                    const StepCtl* __restrict__ ctl,
                                                    const uint32_t* __restrict__ tile_stamp,
                                                    uint8_t* __restrict__ nlive) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total_ntiles) return;
  const SceneDev& S = scenes[ntile_scene ? ntile_scene[t] : 0];
  const uint32_t l = t - S.nt_off;
  const int Kx = int(l % uint32_t(S.nt[0])), Ky = int((l / uint32_t(S.nt[0])) % uint32_t(S.nt[1]));
  const int Kz = int(l / uint32_t(S.nt[0] * S.nt[1]));
  const uint32_t step = ctl->step;
  bool live = false;
#pragma unroll
  for (int b = 0; b < 8; ++b) live |= tile_live(S, tile_stamp, step, Kx + (b & 1), Ky + ((b >> 1) & 1), Kz + (b >> 2));
  nlive[t] = live ? 1 : 0;
}

// Node setup: the mass of every node of a node tile the grid update wrote this step (0 elsewhere),
// x = v* where m > 0 (0 elsewhere), cleared accumulator; and the block's flag: a block with no such
// node is skipped by every later kernel of the solve, which reads nothing of it (the gather and the
// scatter reach only nodes of written tiles).
__global__ void __launch_bounds__(IMP_THREADS) k_imp_setup(const SceneDev* __restrict__ scenes,
                                                           const ImpBlock* __restrict__ blocks,
                                                           const uint8_t* __restrict__ nlive,
                                                           const float2* __restrict__ g_node,
                                                           const float* __restrict__ g_mass,
                                                           float* __restrict__ mass, float4* __restrict__ vec,
                                                           long long* __restrict__ acc, uint8_t* __restrict__ blive) {
  const ImpBlock B = blocks[blockIdx.x];
  const SceneDev& S = scenes[B.scene];
  const int64_t end = S.node_off + int64_t(S.D[0]) * S.D[1] * S.D[2];
  bool live[IMP_NODES_PER_THREAD];
  int any = 0;
#pragma unroll
  for (int j = 0; j < IMP_NODES_PER_THREAD; ++j) {
    const int64_t n = B.first + j * IMP_THREADS + threadIdx.x;
    live[j] = false;
    if (n < end) {
      const int64_t l = n - S.node_off;
      const int nx = int(l % S.D[0]), ny = int((l / S.D[0]) % S.D[1]), nz = int(l / (int64_t(S.D[0]) * S.D[1]));
      live[j] = nlive[S.nt_off + uint32_t(nx / TILE + S.nt[0] * (ny / TILE + S.nt[1] * (nz / TILE)))] != 0;
      any |= live[j] ? 1 : 0;
    }
  }
  const int blk = __syncthreads_or(any);
  if (threadIdx.x == 0) blive[blockIdx.x] = blk ? 1 : 0;
  if (!blk) return;
#pragma unroll
  for (int j = 0; j < IMP_NODES_PER_THREAD; ++j) {
    const int64_t n = B.first + j * IMP_THREADS + threadIdx.x;
    if (n >= end) break;
    float m = live[j] ? g_mass[n] : 0.0f;
    m = m > 0.0f ? m : 0.0f;
    float4 x = make_float4(0.f, 0.f, 0.f, 0.f);
    if (m > 0.0f) {
      const float2 a = g_node[3 * n], b = g_node[3 * n + 1];
      x = make_float4(a.x, a.y, b.x, 0.f);
    }
    mass[n] = m;
    vec[n] = x;  // vector 0: x
    acc[3 * n] = 0;
    acc[3 * n + 1] = 0;
    acc[3 * n + 2] = 0;
  }
}

// ---------------------------------------------------------------------------------------
// The largest row sum of |G| of an application is published as atomic maxima into IMP_GSLOTS slots
// per scene, one 128-byte line apart (workgroup b into slot b mod IMP_GSLOTS), so that the maxima of
// many workgroups do not serialise on one line; readers take the maximum of the slots.
constexpr int IMP_GSLOTS = 64;
constexpr int IMP_GSTRIDE = 32;  // words from one slot to the next
__device__ __forceinline__ uint32_t* imp_slot(uint32_t* gslots, int sid, int k) {
  return gslots + size_t(sid * IMP_GSLOTS + k) * IMP_GSTRIDE;
}
__device__ __forceinline__ uint32_t imp_gmax(const uint32_t* __restrict__ gslots, int sid) {
  return uint32_t(wave_max_i(int(gslots[size_t(sid * IMP_GSLOTS + (threadIdx.x & 63)) * IMP_GSTRIDE])));
}

constexpr int IMP_BOX = 384;       // nodes of a wave's LDS gather box (as the G2P box)
constexpr int IMP_ACC_BOX = 768;   // nodes of a workgroup's LDS accumulation box
constexpr int IMP_WPW = 4;         // wave-table entries (64 sorted positions each) per scatter wave

// One row of 4 nodes along x of the gather L += sum v grad(w)^T (column c = sum v dw/dx_c): the row
// is reduced with the x weights first (sx = sum qx v, sw = sum wx v), then
// grad(w) = (qx wy wz, wx qy wz, wx wy qz) takes the row's y/z factors.
__device__ __forceinline__ void imp_gather_row(const float4 v[4], const float wx[4], const float qx[4], float wyz,
                                               float qyz, float wqz, float L[9]) {
  float sx0 = 0.f, sx1 = 0.f, sx2 = 0.f, sw0 = 0.f, sw1 = 0.f, sw2 = 0.f;
#pragma unroll
  for (int ox = 0; ox < 4; ++ox) {
    sx0 = fmaf(qx[ox], v[ox].x, sx0); sx1 = fmaf(qx[ox], v[ox].y, sx1); sx2 = fmaf(qx[ox], v[ox].z, sx2);
    sw0 = fmaf(wx[ox], v[ox].x, sw0); sw1 = fmaf(wx[ox], v[ox].y, sw1); sw2 = fmaf(wx[ox], v[ox].z, sw2);
  }
  L[0] = fmaf(sx0, wyz, L[0]); L[1] = fmaf(sx1, wyz, L[1]); L[2] = fmaf(sx2, wyz, L[2]);
  L[3] = fmaf(sw0, qyz, L[3]); L[4] = fmaf(sw1, qyz, L[4]); L[5] = fmaf(sw2, qyz, L[5]);
  L[6] = fmaf(sw0, wqz, L[6]); L[7] = fmaf(sw1, wqz, L[7]); L[8] = fmaf(sw2, wqz, L[8]);
}

// First half of K applied to the node vector vin: per particle dF = dt (sum_j vin_j grad w_jp^T) F,
// A = d2Psi/dF2 : dF and G = gs V0 A F^T (planes G0..G3, G4..G7 as float4, G8 as float; G
// column-major), and the scene's largest row sum of |G| for the fixed-point scale. The particle's
// prepared data is loaded first, so that it arrives while the wave gathers. A full wave of interior
// particles gathers from an LDS copy of the node box its stencils cover (sorted particles span a
// few cells; all of its loads are in flight together), else from global memory.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) k_imp_apply(const SceneDev* __restrict__ scenes, const uint2* __restrict__ waves,
                                                   uint32_t nwaves, const float4* __restrict__ pd, uint32_t n,
                                                   const float4* __restrict__ vin, float* __restrict__ G,
                                                   const ImpScal* __restrict__ scal, uint32_t* __restrict__ gslots) {
  __shared__ float4 s_box[4][IMP_BOX];
  __shared__ uint32_t s_gm[2];  // the workgroup's maximum, waves that have added theirs
  if (threadIdx.x == 0) {
    s_gm[0] = 0u;
    s_gm[1] = 0u;
  }
  __syncthreads();
  const uint32_t w0 = blockIdx.x * 4u, w = w0 + uint32_t(__builtin_amdgcn_readfirstlane(threadIdx.x >> 6));
  if (w >= nwaves) return;
  const uint32_t wlast = min(w0 + 3u, nwaves - 1u);
  // the workgroup's waves hold one scene (then the done flag, too, is the same for all of them)
  const bool one_scene = !waves || waves[wlast].y == waves[w0].y;
  const uint2 ws = waves ? waves[w] : make_uint2(64u * w, 0u);
  const int sid = int(ws.y);
  if (scal[sid].done) return;
  const SceneDev& S = scenes[sid];
  const int lane = threadIdx.x & 63;
  const uint32_t i = ws.x + uint32_t(lane);
  const bool valid = int64_t(i) < S.p_off + S.p_cnt;
  const size_t N = n;
  float4 q[IMP_PLANES];
#pragma unroll
  for (int k = 0; k < IMP_PLANES; ++k) q[k] = make_float4(0.f, 0.f, 0.f, 0.f);
  if (valid) {
#pragma unroll
    for (int k = 0; k < IMP_PLANES; ++k) q[k] = pd[size_t(k) * N + i];
  }
  const float v0 = q[0].w;
  const float ux = grid_u(S, q[0].x, 0), uy = grid_u(S, q[0].y, 1), uz = grid_u(S, q[0].z, 2);
  const bool inr = valid && in_stencil_range(S, ux, uy, uz);
  const bool work = inr && v0 != 0.0f;
  const int D0 = S.D[0], D1 = S.D[1], D2 = S.D[2];
  float wx[4], wy[4], wz[4], qx[4], qy[4], qz[4];
  const int bx = axis_weights(ux, wx, qx), by = axis_weights(uy, wy, qy), bz = axis_weights(uz, wz, qz);
  const bool interior = inr && bx >= 1 && by >= 1 && bz >= 1 && bx + 2 < D0 && by + 2 < D1 && bz + 2 < D2;
  float L[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};  // column c = sum v dw/dx_c
  bool boxed = false;
  if (__ballot(interior) == ~0ull) {
    const int xmin = wave_min_i(bx) - 1, ymin = wave_min_i(by) - 1, zmin = wave_min_i(bz) - 1;
    const int ex = wave_max_i(bx) + 3 - xmin, ey = wave_max_i(by) + 3 - ymin, ez = wave_max_i(bz) + 3 - zmin;
    const int vol = ex * ey * ez;
    if (vol <= IMP_BOX) {
      boxed = true;
      float4* box = s_box[threadIdx.x >> 6];
      constexpr int BU = IMP_BOX / 64;
      float4 tmp[BU];
      const float4* v0 = vin + (S.node_off + xmin + int64_t(D0) * (ymin + int64_t(D1) * zmin));
      const BoxSplit bs = box_split(ex, ey);
#pragma unroll
      for (int u = 0; u < BU; ++u) {
        const int idx = lane + 64 * u;
        tmp[u] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (idx < vol) {
          int ix, iy, iz;
          box_coords(bs, idx, ix, iy, iz);
          tmp[u] = v0[uint32_t(ix + D0 * (iy + D1 * iz))];
        }
      }
#pragma unroll
      for (int u = 0; u < BU; ++u)
        if (lane + 64 * u < vol) box[lane + 64 * u] = tmp[u];
      __builtin_amdgcn_wave_barrier();  // the wave's own LDS writes complete before its reads
      if (work) {
        const float4* p0 = box + ((bx - 1 - xmin) + ex * ((by - 1 - ymin) + ey * (bz - 1 - zmin)));
        for (int oz = 0; oz < 4; ++oz)
#pragma unroll
          for (int oy = 0; oy < 4; ++oy) {
            const float4* row = p0 + ex * (oy + ey * oz);
            float4 v4[4];
#pragma unroll
            for (int ox = 0; ox < 4; ++ox) v4[ox] = row[ox];
            imp_gather_row(v4, wx, qx, wy[oy] * wz[oz], qy[oy] * wz[oz], wy[oy] * qz[oz], L);
          }
      }
    }
  }
  if (!boxed && work) {
    for (int oz = 0; oz < 4; ++oz) {
      const int iz = bz + oz - 1;
      if (iz < 0 || iz >= D2) continue;
      for (int oy = 0; oy < 4; ++oy) {
        const int iy = by + oy - 1;
        if (iy < 0 || iy >= D1) continue;
        const int64_t row = S.node_off + int64_t(D0) * (iy + int64_t(D1) * iz);
        float4 v4[4];
#pragma unroll
        for (int ox = 0; ox < 4; ++ox) {
          const int ix = bx + ox - 1;
          v4[ox] = (ix >= 0 && ix < D0) ? vin[row + ix] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        imp_gather_row(v4, wx, qx, wy[oy] * wz[oz], qy[oy] * wz[oz], wy[oy] * qz[oz], L);
      }
    }
  }
  float Gm[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (work) {
    const float k = scal[sid].dt * S.gscale;
#pragma unroll
    for (int e = 0; e < 9; ++e) L[e] *= k;
    const float F[9] = {q[1].x, q[1].y, q[1].z, q[1].w, q[2].x, q[2].y, q[2].z, q[2].w, q[3].x};
    const float Rot[9] = {q[3].y, q[3].z, q[3].w, q[4].x, q[4].y, q[4].z, q[4].w, q[5].x, q[5].y};
    const float mu2 = q[5].z, la = q[5].w;
    float dF[9], D[9], mi[6];
    mul(L, F, dF);       // dF = dt L F
    mul_at(Rot, dF, D);  // R^T dF
    imp_spin_inverse(F, Rot, mi);
    // axial vector of D - D^T: (D(2,1) - D(1,2), D(0,2) - D(2,0), D(1,0) - D(0,1)), D(r,c) = D[c*3+r]
    const float a0 = D[5] - D[7], a1 = D[6] - D[2], a2 = D[1] - D[3];
    const float w0 = mi[0] * a0 + mi[3] * a1 + mi[4] * a2;
    const float w1 = mi[3] * a0 + mi[1] * a1 + mi[5] * a2;
    const float w2 = mi[4] * a0 + mi[5] * a1 + mi[2] * a2;
    // [w]x column-major: (1,0) = w2, (2,0) = -w1, (0,1) = -w2, (2,1) = w0, (0,2) = w1, (1,2) = -w0
    const float Om[9] = {0.f, w2, -w1, -w2, 0.f, w0, w1, -w0, 0.f};
    float dR[9];
    mul(Rot, Om, dR);
    // cof F (columns: cross products of the other two) and its derivative along dF
    float C[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    cross_add(F + 3, F + 6, C);
    cross_add(F + 6, F, C + 3);
    cross_add(F, F + 3, C + 6);
    const float laj = la * (det3(F) - 1.0f);
    float dJ = 0.f;
#pragma unroll
    for (int e = 0; e < 9; ++e) dJ = fmaf(C[e], dF[e], dJ);
    float dC[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    cross_add(dF + 3, F + 6, dC);
    cross_add(F + 3, dF + 6, dC);
    cross_add(dF + 6, F, dC + 3);
    cross_add(F + 6, dF, dC + 3);
    cross_add(dF, F + 3, dC + 6);
    cross_add(F, dF + 3, dC + 6);
    float A[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) A[e] = mu2 * (dF[e] - dR[e]) + (la * dJ) * C[e] + laj * dC[e];
    mul_bt(A, F, Gm);  // A F^T
#pragma unroll
    for (int e = 0; e < 9; ++e) Gm[e] *= v0;
  }
  if (valid) {
    float4* G4 = reinterpret_cast<float4*>(G);
    G4[i] = make_float4(Gm[0], Gm[1], Gm[2], Gm[3]);
    G4[N + i] = make_float4(Gm[4], Gm[5], Gm[6], Gm[7]);
    G[8 * N + i] = Gm[8];
  }
  // the scene's largest row sum of |G| (non-negative floats compare as their bits)
  float rs = 0.f;
#pragma unroll
  for (int r = 0; r < 3; ++r) rs = fmaxf(rs, fabsf(Gm[r]) + fabsf(Gm[3 + r]) + fabsf(Gm[6 + r]));
  const int mx = wave_max_i(int(__float_as_uint(rs)));
  if (!one_scene) {
    if (lane == 0 && mx > 0) atomicMax(imp_slot(gslots, sid, int(w % IMP_GSLOTS)), uint32_t(mx));
    return;
  }
  // one atomic per workgroup: the last of its live waves publishes the workgroup's maximum
  if (lane == 0) {
    if (mx > 0) atomicMax(&s_gm[0], uint32_t(mx));
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): this wave's maximum is in LDS
    if (atomicAdd(&s_gm[1], 1u) == wlast - w0) {
      const uint32_t m = s_gm[0];
      if (m) atomicMax(imp_slot(gslots, sid, int(blockIdx.x % IMP_GSLOTS)), m);
    }
  }
}

// ---------------------------------------------------------------------------------------
// Second half: acc_i += sum_p G_p grad w_ip (unscaled grad w: G holds gs) in fixed point. A workgroup
// takes 4 * WPW consecutive entries of the wave table, WPW per wave. For each entry (64 sorted
// positions of one scene) the wave stages IMP_BATCH particles at a time in LDS (lane k: particle k's
// weights and G), then lane l is the stencil node getIJK(l) and sums the staged particles in
// registers while the base cell stays the same, adding the rounded sum where it changes (a scalar
// mask of the changes, taken before the loop). The positions of all the wave's entries are loaded
// first, G entry by entry (fewer live registers).
// When the workgroup's entries are of one scene and their stencils cover at most IMP_ACC_BOX nodes,
// the sums go into an LDS box first (integer LDS atomics) and the box is added to acc once;
// otherwise straight to acc.
template <int WPW>
__global__ void __launch_bounds__(256) k_imp_scatter(const SceneDev* __restrict__ scenes,
                                                     const uint2* __restrict__ waves, uint32_t nwaves,
                                                     const float4* __restrict__ pd, const float* __restrict__ G,
                                                     uint32_t n, unsigned long long* __restrict__ acc,
                                                     const ImpScal* __restrict__ scal,
                                                     const uint32_t* __restrict__ gslots) {
  __shared__ __attribute__((aligned(16))) float s_stage[4][IMP_BATCH * IMP_STAGE];
  __shared__ unsigned long long s_acc[3 * IMP_ACC_BOX];
  __shared__ int s_ext[6];  // box min (x, y, z), max (x, y, z) of the base cells + 2
  // wave index, scene and loop bounds are wave-uniform (scalar registers, scalar loads and branches)
  const int t = threadIdx.x, lane = t & 63, wv = __builtin_amdgcn_readfirstlane(t >> 6);
  const uint32_t e0 = blockIdx.x * uint32_t(4 * WPW);
  const uint32_t elast = min(e0 + uint32_t(4 * WPW - 1), nwaves - 1u);
  const int sid0 = waves ? int(waves[e0].y) : 0;
  // one scene for the whole workgroup (the scene of an entry is constant along the table)
  const bool one_scene = !waves || int(waves[elast].y) == sid0;
  if (t < 3) s_ext[t] = 0x7fffffff;
  else if (t < 6) s_ext[t] = -1;
  const size_t N = n;
  // this lane's particle of each of the wave's entries: grid coordinate and base cell + 2 (10 bits
  // per axis; -1: none, out of the grid or the scene's solve is done)
  // (the positions of all entries load together, before any wait)
  float ux[WPW], uy[WPW], uz[WPW];
  int cell[WPW];
  float4 xq[WPW];
  bool live[WPW];
#pragma unroll
  for (int j = 0; j < WPW; ++j) {
    const uint32_t e = e0 + uint32_t(wv * WPW + j);
    xq[j] = make_float4(0.f, 0.f, 0.f, 0.f);
    live[j] = false;
    if (e < nwaves) {
      const uint2 ws = waves ? waves[e] : make_uint2(64u * e, 0u);
      const SceneDev& S = scenes[ws.y];
      const uint32_t i = ws.x + uint32_t(lane);
      live[j] = !scal[ws.y].done && int64_t(i) < S.p_off + S.p_cnt;
      if (live[j]) xq[j] = pd[i];
    }
  }
  int cmin[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff}, cmax[3] = {-1, -1, -1};
#pragma unroll
  for (int j = 0; j < WPW; ++j) {
    ux[j] = uy[j] = uz[j] = 0.f;
    cell[j] = -1;
    if (live[j]) {
      const uint32_t e = e0 + uint32_t(wv * WPW + j);
      const SceneDev& S = scenes[waves ? waves[e].y : 0u];
      ux[j] = grid_u(S, xq[j].x, 0);
      uy[j] = grid_u(S, xq[j].y, 1);
      uz[j] = grid_u(S, xq[j].z, 2);
      if (in_stencil_range(S, ux[j], uy[j], uz[j])) {  // base cell in [-2, D] per axis
        const int cx = int(floorf(ux[j])) + 2, cy = int(floorf(uy[j])) + 2, cz = int(floorf(uz[j])) + 2;
        cell[j] = cx | (cy << 10) | (cz << 20);
        cmin[0] = min(cmin[0], cx); cmin[1] = min(cmin[1], cy); cmin[2] = min(cmin[2], cz);
        cmax[0] = max(cmax[0], cx); cmax[1] = max(cmax[1], cy); cmax[2] = max(cmax[2], cz);
      }
    }
  }
  __syncthreads();  // s_ext initialised
  if (one_scene) {
    const int lx = wave_min_i(cmin[0]), ly = wave_min_i(cmin[1]), lz = wave_min_i(cmin[2]);
    const int hx = wave_max_i(cmax[0]), hy = wave_max_i(cmax[1]), hz = wave_max_i(cmax[2]);
    if (lane == 0 && hx >= 0) {
      atomicMin(&s_ext[0], lx); atomicMin(&s_ext[1], ly); atomicMin(&s_ext[2], lz);
      atomicMax(&s_ext[3], hx); atomicMax(&s_ext[4], hy); atomicMax(&s_ext[5], hz);
    }
  }
  __syncthreads();
  // the accumulation box: nodes (c - 2) - 1 .. (c - 2) + 2 of the base cells, clipped to the grid
  const SceneDev& S0 = scenes[sid0];
  int bx0 = 0, by0 = 0, bz0 = 0, ex = 0, ey = 0, ez = 0;
  bool boxed = false;
  if (one_scene && s_ext[3] >= 0) {
    bx0 = max(0, s_ext[0] - 3); by0 = max(0, s_ext[1] - 3); bz0 = max(0, s_ext[2] - 3);
    ex = min(S0.D[0], s_ext[3] + 1) - bx0; ey = min(S0.D[1], s_ext[4] + 1) - by0; ez = min(S0.D[2], s_ext[5] + 1) - bz0;
    boxed = ex > 0 && ey > 0 && ez > 0 && ex * ey * ez <= IMP_ACC_BOX;
  }
  if (boxed)
    for (int k = t; k < 3 * ex * ey * ez; k += 256) s_acc[k] = 0ull;
  __syncthreads();
  // lane = (ox, oy) of the stencil and a subgroup sg of the particles: the lane sums the nodes
  // (ox, oy, 0..3) for the particles k = 4i + sg of the current cell, so one LDS read of a staged
  // particle serves four nodes; at a change of cell the four subgroups' sums are reduced and lane
  // (ox, oy, sg) adds node (ox, oy, sg)
  const int ox = lane & 3, oy = (lane >> 2) & 3, sg = lane >> 4;
  // register sums of the nodes (ox, oy, oz) of cell cur, in fixed point with 2^e = scale
  int cur = -1;
  float sx[4] = {0.f, 0.f, 0.f, 0.f}, sy[4] = {0.f, 0.f, 0.f, 0.f}, sz[4] = {0.f, 0.f, 0.f, 0.f};
  int csid = -1, cD0 = 0, cD1 = 0, cD2 = 0;  // scene of the sums and its grid
  int64_t cnode = 0;
  double scale = 0.0;
  auto commit = [&]() {
    float s0 = 0.f, s1 = 0.f, s2 = 0.f;
#pragma unroll
    for (int z = 0; z < 4; ++z) {
      // sum over the subgroups (lanes l, l ^ 16, l ^ 32, l ^ 48), the same order in every lane
      float a = sx[z], b = sy[z], c = sz[z];
      a += __shfl_xor(a, 16, 64); b += __shfl_xor(b, 16, 64); c += __shfl_xor(c, 16, 64);
      a += __shfl_xor(a, 32, 64); b += __shfl_xor(b, 32, 64); c += __shfl_xor(c, 32, 64);
      if (sg == z) { s0 = a; s1 = b; s2 = c; }
      sx[z] = sy[z] = sz[z] = 0.f;
    }
    const int nx = (cur & 1023) - 3 + ox, ny = ((cur >> 10) & 1023) - 3 + oy, nz = (cur >> 20) - 3 + sg;
    if (nx >= 0 && ny >= 0 && nz >= 0 && nx < cD0 && ny < cD1 && nz < cD2) {
      const long long q0 = llrint(double(s0) * scale), q1 = llrint(double(s1) * scale), q2 = llrint(double(s2) * scale);
      if (boxed) {
        unsigned long long* a = s_acc + 3 * ((nx - bx0) + ex * ((ny - by0) + ey * (nz - bz0)));
        if (q0) atomicAdd(a, (unsigned long long)q0);
        if (q1) atomicAdd(a + 1, (unsigned long long)q1);
        if (q2) atomicAdd(a + 2, (unsigned long long)q2);
      } else {
        const int64_t node = cnode + nx + int64_t(cD0) * (ny + int64_t(cD1) * nz);
        if (q0) atomicAdd(acc + 3 * node, (unsigned long long)q0);
        if (q1) atomicAdd(acc + 3 * node + 1, (unsigned long long)q1);
        if (q2) atomicAdd(acc + 3 * node + 2, (unsigned long long)q2);
      }
    }
  };
  float* st = s_stage[wv];
  // this lane's read offsets in a staged particle (floats): (w, q) of ox and of oy
  const int lx = 2 * ox + IMP_STAGE * sg, ly = 8 + 2 * oy + IMP_STAGE * sg, lz = IMP_STAGE * sg;
#pragma unroll
  for (int j = 0; j < WPW; ++j) {
    const uint32_t e = e0 + uint32_t(wv * WPW + j);
    if (e >= nwaves) break;
    const uint2 ws = waves ? waves[e] : make_uint2(64u * e, 0u);
    const int sid = __builtin_amdgcn_readfirstlane(int(ws.y));
    if (scal[sid].done) continue;
    const SceneDev& Sj = scenes[sid];
    if (sid != csid) {  // the sums so far belong to another scene
      if (cur >= 0) commit();
      cur = -1;
      csid = sid;
      cD0 = Sj.D[0];
      cD1 = Sj.D[1];
      cD2 = Sj.D[2];
      cnode = Sj.node_off;
      scale = ldexp(1.0, imp_exponent(imp_gmax(gslots, sid), Sj.p_cnt));
    }
    const uint32_t cnt = uint32_t(min<int64_t>(64, Sj.p_off + Sj.p_cnt - int64_t(ws.x)));
    const int cj = cell[j];
    const int prevc = __shfl_up(cj, 1, 64);
    float4 gaj = make_float4(0.f, 0.f, 0.f, 0.f), gbj = gaj;
    float gcj = 0.f;
    if (cj >= 0) {
      const uint32_t i = ws.x + uint32_t(lane);
      gaj = reinterpret_cast<const float4*>(G)[i];
      gbj = reinterpret_cast<const float4*>(G)[N + i];
      gcj = G[8 * N + i];
    }
    for (uint32_t b0 = 0; b0 < cnt; b0 += uint32_t(IMP_BATCH)) {
      const uint32_t bc = min(uint32_t(IMP_BATCH), cnt - b0);
      const bool mine = uint32_t(lane) >= b0 && uint32_t(lane) < b0 + bc;
      if (mine && cj >= 0) {
        // (w_o, q_o) interleaved for x and y, then w_0..3 and q_0..3 for z, then G
        float4* d4 = reinterpret_cast<float4*>(st + (lane - int(b0)) * IMP_STAGE);
        const float fx = ux[j] - floorf(ux[j]), fy = uy[j] - floorf(uy[j]), fz = uz[j] - floorf(uz[j]);
        d4[0] = make_float4(spline_w(0, fx), spline_q(0, fx), spline_w(1, fx), spline_q(1, fx));
        d4[1] = make_float4(spline_w(2, fx), spline_q(2, fx), spline_w(3, fx), spline_q(3, fx));
        d4[2] = make_float4(spline_w(0, fy), spline_q(0, fy), spline_w(1, fy), spline_q(1, fy));
        d4[3] = make_float4(spline_w(2, fy), spline_q(2, fy), spline_w(3, fy), spline_q(3, fy));
        d4[4] = make_float4(spline_w(0, fz), spline_w(1, fz), spline_w(2, fz), spline_w(3, fz));
        d4[5] = make_float4(spline_q(0, fz), spline_q(1, fz), spline_q(2, fz), spline_q(3, fz));
        d4[6] = gaj;
        d4[7] = gbj;
        d4[8] = make_float4(gcj, 0.f, 0.f, 0.f);
      }
      __builtin_amdgcn_wave_barrier();  // the wave's own staging writes complete before its reads
      // bit k: staged particle k has another cell than the one before it (bit 0: than cur)
      const uint32_t chg =
          uint32_t(__ballot(mine && (uint32_t(lane) == b0 ? cj != cur : cj != prevc)) >> b0);
      uint32_t k = 0;
      while (k < bc) {  // segments of one cell: [k, kend)
        if ((chg >> k) & 1u) {
          if (cur >= 0) commit();
          cur = __builtin_amdgcn_readlane(cj, int(b0 + k));
        }
        const uint32_t rest = k + 1u < uint32_t(IMP_BATCH) ? chg >> (k + 1u) : 0u;
        const uint32_t kend = rest ? min(bc, k + 1u + uint32_t(__builtin_ctz(rest))) : bc;
        if (cur >= 0) {
          for (uint32_t kk = k; kk < kend; kk += 4u) {
            if (kk + uint32_t(sg) < kend) {
              const float* d = st + kk * IMP_STAGE;
              const float2 wqx = *reinterpret_cast<const float2*>(d + lx);
              const float2 wqy = *reinterpret_cast<const float2*>(d + ly);
              const float4 wz = *reinterpret_cast<const float4*>(d + lz + 16);
              const float4 qz = *reinterpret_cast<const float4*>(d + lz + 20);
              const float4 u0 = *reinterpret_cast<const float4*>(d + lz + 24);
              const float4 u1 = *reinterpret_cast<const float4*>(d + lz + 28);
              const float u2 = d[lz + 32];
              const float wxy = wqx.x * wqy.x, qxwy = wqx.y * wqy.x, wxqy = wqx.x * wqy.y;
              // G grad(w) of node (ox, oy, z) = wz[z] a + qz[z] b with a = G(:,0) qx wy + G(:,1) wx qy
              // and b = G(:,2) wx wy (grad(w) = (qx wy wz, wx qy wz, wx wy qz)); G(r,c) = G[c*3 + r]
              const float a0 = fmaf(u0.x, qxwy, u0.w * wxqy), a1 = fmaf(u0.y, qxwy, u1.x * wxqy),
                          a2 = fmaf(u0.z, qxwy, u1.y * wxqy);
              const float b0 = u1.z * wxy, b1 = u1.w * wxy, b2 = u2 * wxy;
              const float wzv[4] = {wz.x, wz.y, wz.z, wz.w}, qzv[4] = {qz.x, qz.y, qz.z, qz.w};
#pragma unroll
              for (int z = 0; z < 4; ++z) {
                sx[z] = fmaf(wzv[z], a0, fmaf(qzv[z], b0, sx[z]));
                sy[z] = fmaf(wzv[z], a1, fmaf(qzv[z], b1, sy[z]));
                sz[z] = fmaf(wzv[z], a2, fmaf(qzv[z], b2, sz[z]));
              }
            }
          }
        }
        k = kend;
      }
      __builtin_amdgcn_wave_barrier();  // all reads of the batch done before the next one is staged
    }
  }
  if (cur >= 0) commit();
  if (!boxed) return;
  __syncthreads();
  const int vol = ex * ey * ez;
  const BoxSplit bs = box_split(ex, ey);
  for (int k = t; k < vol; k += 256) {
    int ix, iy, iz;
    box_coords(bs, k, ix, iy, iz);
    const int64_t node = S0.node_off + (bx0 + ix) + int64_t(S0.D[0]) * ((by0 + iy) + int64_t(S0.D[1]) * (bz0 + iz));
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const unsigned long long v = s_acc[3 * k + q];
      if (v) atomicAdd(acc + 3 * node + q, v);
    }
  }
}

// ---------------------------------------------------------------------------------------
// Block sum of up to three doubles in a fixed order (waves by shuffles, then wave 0 over the waves).
__device__ __forceinline__ void imp_block_sum(double v[3], int nq, double* __restrict__ part, uint32_t nblocks) {
  __shared__ double s_red[3][IMP_THREADS / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int q = 0; q < nq; ++q) {
    double x = v[q];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) x += __shfl_down(x, off, 64);
    if (lane == 0) s_red[q][wv] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int q = 0; q < nq; ++q) {
      double s = 0.0;
      for (int k = 0; k < IMP_THREADS / 64; ++k) s += s_red[q][k];
      part[size_t(q) * nblocks + blockIdx.x] = s;
    }
}

// Node phases of the solve (K = M v - beta dt df, df = -acc 2^-e; z = r / m):
//   IMP_INIT: vin = x = v*: r = M v* - K x, z = r / m, p = Kp = 0; partials |M v*|^2, |r|^2
//   IMP_KR:   vin = z:      Kz = K z; partials z . Kz, Kz . Kz / m, Kz . Kp / m (Kp of the last
//                           direction), from which one reduction gives beta and the new direction's
//                           Kp . M^-1 Kp = a + 2 beta b + beta^2 pp (Kp = Kz + beta Kp), hence alpha
//   IMP_PX:   p = z + beta p, Kp = Kz + beta Kp, x += alpha p, r -= alpha Kp, z -= alpha Kp / m;
//             partial |r|^2
template <int PHASE>
__global__ void __launch_bounds__(IMP_THREADS) k_imp_node(const SceneDev* __restrict__ scenes,
                                                          const ImpBlock* __restrict__ blocks, uint32_t nblocks,
                                                          const float* __restrict__ mass, float4* __restrict__ vec,
                                                          long long* __restrict__ acc,
                                                          const ImpScal* __restrict__ scal, int64_t total_nodes,
                                                          double* __restrict__ part, const uint32_t* __restrict__ gslots,
                                                          const uint8_t* __restrict__ blive) {
  const ImpBlock B = blocks[blockIdx.x];
  const ImpScal& Z = scal[B.scene];
  if (PHASE != IMP_INIT && Z.done) return;  // the scene's partials are not read again
  constexpr int NQ = PHASE == IMP_INIT ? 2 : (PHASE == IMP_KR ? 3 : 1);
  if (!blive[blockIdx.x]) {  // no node of a written tile: zero partials
    if (threadIdx.x < NQ) part[size_t(threadIdx.x) * nblocks + blockIdx.x] = 0.0;
    return;
  }
  const SceneDev& S = scenes[B.scene];
  const int64_t end = S.node_off + int64_t(S.D[0]) * S.D[1] * S.D[2];
  float4* X = vec;
  float4* Rv = vec + total_nodes;
  float4* P = vec + 2 * total_nodes;
  float4* KP = vec + 3 * total_nodes;
  float4* KZ = vec + 4 * total_nodes;
  float4* ZV = vec + 5 * total_nodes;
  const double inv = (PHASE == IMP_INIT || PHASE == IMP_KR) ? ldexp(1.0, -imp_exponent(imp_gmax(gslots, B.scene), S.p_cnt)) : 0.0;
  const float bdt = Z.beta_dt;
  const float al = float(Z.alpha), be = float(Z.beta);
  double sum[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int j = 0; j < IMP_NODES_PER_THREAD; ++j) {
    const int64_t nd = B.first + j * IMP_THREADS + threadIdx.x;
    if (nd >= end) break;
    const float m = mass[nd];
    // vectors are 0 where m <= 0 (set so by IMP_INIT) and the accumulator too (no particle with
    // mass has a weight gradient there), so the later phases touch only nodes with m > 0
    if (PHASE != IMP_INIT && !(m > 0.0f)) continue;
    if (PHASE == IMP_INIT || PHASE == IMP_KR) {
      // df = -acc 2^-e; K v = m v - beta dt df = m v + beta dt acc 2^-e
      const float a0 = float(double(acc[3 * nd]) * inv), a1 = float(double(acc[3 * nd + 1]) * inv),
                  a2 = float(double(acc[3 * nd + 2]) * inv);
      acc[3 * nd] = 0;
      acc[3 * nd + 1] = 0;
      acc[3 * nd + 2] = 0;
      if (PHASE == IMP_INIT) {
        float4 r = make_float4(0.f, 0.f, 0.f, 0.f), b = r, z = r;
        if (m > 0.0f) {
          const float4 x = X[nd];
          b = make_float4(m * x.x, m * x.y, m * x.z, 0.f);
          const float kx = fmaf(bdt, a0, b.x), ky = fmaf(bdt, a1, b.y), kz = fmaf(bdt, a2, b.z);
          r = make_float4(b.x - kx, b.y - ky, b.z - kz, 0.f);
          z = make_float4(r.x / m, r.y / m, r.z / m, 0.f);
        }
        Rv[nd] = r;
        ZV[nd] = z;
        P[nd] = make_float4(0.f, 0.f, 0.f, 0.f);
        KP[nd] = make_float4(0.f, 0.f, 0.f, 0.f);
        KZ[nd] = make_float4(0.f, 0.f, 0.f, 0.f);
        sum[0] += double(b.x) * b.x + double(b.y) * b.y + double(b.z) * b.z;
        sum[1] += double(r.x) * r.x + double(r.y) * r.y + double(r.z) * r.z;
      } else {
        const float4 z = ZV[nd], kp = KP[nd];
        const float4 kz = make_float4(fmaf(bdt, a0, m * z.x), fmaf(bdt, a1, m * z.y), fmaf(bdt, a2, m * z.z), 0.f);
        KZ[nd] = kz;
        const double im = 1.0 / double(m);
        sum[0] += double(z.x) * kz.x + double(z.y) * kz.y + double(z.z) * kz.z;
        sum[1] += (double(kz.x) * kz.x + double(kz.y) * kz.y + double(kz.z) * kz.z) * im;
        sum[2] += (double(kz.x) * kp.x + double(kz.y) * kp.y + double(kz.z) * kp.z) * im;
      }
    } else {
      const float4 z = ZV[nd], kz = KZ[nd], p = P[nd], kp = KP[nd], x = X[nd], r = Rv[nd];
      const float4 pn = make_float4(fmaf(be, p.x, z.x), fmaf(be, p.y, z.y), fmaf(be, p.z, z.z), 0.f);
      const float4 kpn = make_float4(fmaf(be, kp.x, kz.x), fmaf(be, kp.y, kz.y), fmaf(be, kp.z, kz.z), 0.f);
      P[nd] = pn;
      KP[nd] = kpn;
      X[nd] = make_float4(fmaf(al, pn.x, x.x), fmaf(al, pn.y, x.y), fmaf(al, pn.z, x.z), 0.f);
      const float4 rn = make_float4(fmaf(-al, kpn.x, r.x), fmaf(-al, kpn.y, r.y), fmaf(-al, kpn.z, r.z), 0.f);
      const float am = al / m;
      ZV[nd] = make_float4(fmaf(-am, kpn.x, z.x), fmaf(-am, kpn.y, z.y), fmaf(-am, kpn.z, z.z), 0.f);
      Rv[nd] = rn;
      sum[0] += double(rn.x) * rn.x + double(rn.y) * rn.y + double(rn.z) * rn.z;
    }
  }
  imp_block_sum(sum, NQ, part, nblocks);
}

// One workgroup per scene: the scene's block partials summed in a fixed order, then the scalars
// of the phase (after IMP_KR: beta, the direction's Kp . M^-1 Kp and alpha; after IMP_PX: the
// residual and the stop test). START (phase < 0) resets the state for a new solve.
__global__ void __launch_bounds__(IMP_THREADS) k_imp_scalars(int phase, const uint32_t* __restrict__ scene_blocks,
                                                             uint32_t nblocks, const double* __restrict__ part,
                                                             ImpScal* __restrict__ scal, const SceneDev* __restrict__ scenes,
                                                             float dt_override, float beta, float tol, int max_iter,
                                                             uint32_t* __restrict__ gslots) {
  const int sid = blockIdx.x;
  ImpScal& Z = scal[sid];
  // the bound slots of the next application start at 0 (after the start, and after each phase
  // that consumed them)
  if ((phase < 0 || phase == IMP_INIT || phase == IMP_KR) && threadIdx.x < IMP_GSLOTS)
    *imp_slot(gslots, sid, int(threadIdx.x)) = 0u;
  if (phase < 0) {
    if (threadIdx.x == 0) {
      const float dt = dt_override > 0.0f ? dt_override : scenes[sid].dt;
      Z.rho = Z.alpha = Z.beta = Z.bb = Z.rr = Z.pp = 0.0;
      Z.tol2 = double(tol) * double(tol);
      Z.iters = 0;
      Z.done = 0;
      Z.max_iter = max_iter;
      Z.dt = dt;
      Z.beta_dt = beta * dt;
    }
    return;
  }
  if (phase != IMP_INIT && Z.done) return;
  const uint32_t b0 = scene_blocks[sid], b1 = scene_blocks[sid + 1];
  const int nq = phase == IMP_INIT ? 2 : (phase == IMP_KR ? 3 : 1);
  __shared__ double s_v[3][IMP_THREADS];
  for (int q = 0; q < nq; ++q) {
    double s = 0.0;
    for (uint32_t b = b0 + threadIdx.x; b < b1; b += IMP_THREADS) s += part[size_t(q) * nblocks + b];
    s_v[q][threadIdx.x] = s;
  }
  __syncthreads();
  for (int h = IMP_THREADS / 2; h >= 1; h >>= 1) {
    if (int(threadIdx.x) < h)
      for (int q = 0; q < nq; ++q) s_v[q][threadIdx.x] += s_v[q][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const double t0 = s_v[0][0], t1 = s_v[1][0], t2 = s_v[2][0];
  if (phase == IMP_INIT) {
    Z.bb = t0;
    Z.rr = t1;
    Z.pp = 0.0;
    Z.done = (t1 <= Z.tol2 * t0 || Z.max_iter <= 0) ? 1 : 0;
  } else if (phase == IMP_KR) {
    if (Z.iters > 0 && !(Z.rho != 0.0)) {  // z . K z vanished (K is not definite): no next direction
      Z.done = 1;
      return;
    }
    const double be = Z.iters == 0 ? 0.0 : t0 / Z.rho;
    const double pp = t1 + 2.0 * be * t2 + be * be * Z.pp;  // Kp . M^-1 Kp of p = z + beta p
    Z.beta = be;
    Z.rho = t0;
    Z.pp = pp;
    if (pp > 0.0) Z.alpha = t0 / pp;
    else Z.done = 1;  // K p = 0: no further progress
  } else {
    Z.rr = t0;
    Z.iters += 1;
    Z.done = (t0 <= Z.tol2 * Z.bb || Z.iters >= Z.max_iter || !(t0 == t0)) ? 1 : 0;
  }
}

// v = x where m > 0: the grid update's record becomes (v, u + v - v*), u = v* - alpha p/m.
__global__ void __launch_bounds__(IMP_THREADS) k_imp_finish(const SceneDev* __restrict__ scenes,
                                                            const ImpBlock* __restrict__ blocks,
                                                            const float* __restrict__ mass,
                                                            const float4* __restrict__ vec,
                                                            float2* __restrict__ g_node, float* __restrict__ g_mass,
                                                            const uint8_t* __restrict__ blive) {
  if (!blive[blockIdx.x]) return;
  const ImpBlock B = blocks[blockIdx.x];
  const SceneDev& S = scenes[B.scene];
  const int64_t end = S.node_off + int64_t(S.D[0]) * S.D[1] * S.D[2];
#pragma unroll
  for (int j = 0; j < IMP_NODES_PER_THREAD; ++j) {
    const int64_t nd = B.first + j * IMP_THREADS + threadIdx.x;
    if (nd >= end) break;
    const float m = mass[nd];
    if (!(m > 0.0f)) continue;
    const float4 x = vec[nd];
    const float2 a = g_node[3 * nd], b = g_node[3 * nd + 1], c = g_node[3 * nd + 2];
    if (x.x == a.x && x.y == a.y && x.z == b.x) continue;  // unchanged (beta = 0, no iteration): keep the bits
    store_node(g_node, g_mass, nd, x.x, x.y, x.z, b.y + (x.x - a.x), c.x + (x.y - a.y), c.y + (x.z - b.x), m);
  }
}

}  // namespace snow
