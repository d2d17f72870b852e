// snow_device.hpp - device math of the snow MPM step for gfx950.
//
// Restates, for registers on CDNA4, the GLSL include files of the reference:
//   (cubic B-spline weights, cubic.include.glsl:7-40, are in snow_kernels.hpp: spline_w/q)
//   Lame parameters with hardening  shader/compute/lame/lame.include.glsl:11-27
//   3x3 SVD / polar decomposition .. shader/compute/svd/SVD.include.glsl:24-385
//                                   (McAdams et al. 2011, UW-Madison TR1690: quaternion
//                                   Jacobi eigenanalysis of A^T A, 4 sweeps, then a Givens QR)
//   collision response ............ shader/compute/mpm/gForceCompute.glsl:40-93
// Matrices are held as 9 scalars in registers, m[c*3+r] = (row r, column c), the
// column-major order of GLSL mat3. Everything is branch-free selects except the
// collider loop. The SVD is templated on the scalar type so the mixed-precision
// path can rerun it in FP64 for particles near the plastic clamp.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace snow {

template <class R> struct M3 { R m[9]; };  // m[c*3+r]

// SVD square roots: the hardware v_rsq_f32 / v_sqrt_f32 (1 ulp) without the library's rescaling of
// denormal inputs, which the SVD never passes (the Jacobi step takes its exact rotation only for a
// normal ch^2; the Givens pivot is at least EPS^2). FP64 uses the library functions.
template <class R> __device__ __forceinline__ R rsq(R x);
template <> __device__ __forceinline__ float rsq<float>(float x) { return __builtin_amdgcn_rsqf(x); }
template <> __device__ __forceinline__ double rsq<double>(double x) { return rsqrt(x); }
template <class R> __device__ __forceinline__ R sq_root(R x);
template <> __device__ __forceinline__ float sq_root<float>(float x) { return __builtin_amdgcn_sqrtf(x); }
template <> __device__ __forceinline__ double sq_root<double>(double x) { return sqrt(x); }
// GLSL round(): nearest integer; the direction of exact .5 ties is left to the implementation,
// here the hardware's round-to-nearest-even (one instruction).
template <class R> __device__ __forceinline__ R rnd(R x);
template <> __device__ __forceinline__ float rnd<float>(float x) { return rintf(x); }
template <> __device__ __forceinline__ double rnd<double>(double x) { return rint(x); }
template <class R> __device__ __forceinline__ R absv(R x) { return x < R(0) ? -x : x; }
// reciprocal: v_rcp_f32 (1 ulp) for float, exact division for double
template <class R> __device__ __forceinline__ R recip(R x);
template <> __device__ __forceinline__ float recip<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <> __device__ __forceinline__ double recip<double>(double x) { return 1.0 / x; }
// round(scale*x)/scale of the reference shaders (SVD.include.glsl:356, p2gTransfer.glsl:114,
// particleAdvance.glsl:122/138), with the division by the constant done as a multiplication.
template <class R> __device__ __forceinline__ R round_to(R x, R scale, R inv_scale) { return rnd(scale * x) * inv_scale; }

// C = A * B (column-major 3x3)
template <class R>
__device__ __forceinline__ void mul(const R* A, const R* B, R* C) {
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) C[c * 3 + r] = A[r] * B[c * 3] + A[3 + r] * B[c * 3 + 1] + A[6 + r] * B[c * 3 + 2];
}
// C = A * B^T
template <class R>
__device__ __forceinline__ void mul_bt(const R* A, const R* B, R* C) {
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) C[c * 3 + r] = A[r] * B[c] + A[3 + r] * B[3 + c] + A[6 + r] * B[6 + c];
}
// C = A^T * B
template <class R>
__device__ __forceinline__ void mul_at(const R* A, const R* B, R* C) {
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r)
      C[c * 3 + r] = A[r * 3] * B[c * 3] + A[r * 3 + 1] * B[c * 3 + 1] + A[r * 3 + 2] * B[c * 3 + 2];
}
template <class R>
__device__ __forceinline__ R det3(const R* m) {
  // m[c*3+r]: expansion along the first column, as GLSL determinant()
  return m[0] * (m[4] * m[8] - m[7] * m[5]) - m[3] * (m[1] * m[8] - m[7] * m[2]) +
         m[6] * (m[1] * m[5] - m[4] * m[2]);
}

// ---- Lame parameters (lame.include.glsl:11-27) -----------------------------------------
template <class R>
__device__ __forceinline__ R hardening_factor(R xi, R JP);
// FP32: the hardware exp2 (__expf, a few ulp for the small arguments xi (1 - J_P) here; GLSL's
// exp allows 3 + 2|x| ulp)
template <> __device__ __forceinline__ float hardening_factor<float>(float xi, float JP) {
  return __expf(xi * (1.0f - JP));
}
template <> __device__ __forceinline__ double hardening_factor<double>(double xi, double JP) {
  return exp(xi * (1.0 - JP));
}

// ---- SVD (SVD.include.glsl) -----------------------------------------------------------
template <class R>
__device__ __forceinline__ void quat_to_mat(R q0, R q1, R q2, R q3, R* M) {
  const R qxx = q1 * q1, qyy = q2 * q2, qzz = q3 * q3;
  const R qxz = q1 * q3, qxy = q1 * q2, qyz = q2 * q3;
  const R qwx = q0 * q1, qwy = q0 * q2, qwz = q0 * q3;
  M[0] = R(1) - R(2) * (qyy + qzz); M[1] = R(2) * (qxy + qwz); M[2] = R(2) * (qxz - qwy);
  M[3] = R(2) * (qxy - qwz); M[4] = R(1) - R(2) * (qxx + qzz); M[5] = R(2) * (qyz + qwx);
  M[6] = R(2) * (qxz + qwy); M[7] = R(2) * (qyz - qwx); M[8] = R(1) - R(2) * (qxx + qyy);
}

// One Jacobi conjugation on the symmetric 3x3 held as its 6 distinct entries
// (a00 a11 a22 a01 a02 a12). The reference rotates the (0,1) pair and relabels the
// matrix cyclically (SVD.include.glsl:110-112); the relabelling is carried by the
// caller through the (x,y,z) argument exactly as the quaternion update does.
template <class R>
__device__ __forceinline__ void jacobi_conj(int x, int y, int z, R& s00, R& s11, R& s22, R& s01,
                                            R& s02, R& s12, R q[4]) {
  const R GAMMA = R(5.828427124), CSTAR = R(0.923879532), SSTAR = R(0.3826834323);
  const R ch = R(2) * (s00 - s11), sh = s01;
  const R ch2 = ch * ch, sh2 = sh * sh;
  // exact rotation when ch^2 > gamma sh^2 (SVD.include.glsl:83-92), else the fixed pi/8 one; a
  // denormal ch^2 (never reached by matrices with entries of order one) also takes the fixed one
  const bool flag = GAMMA * sh2 < ch2 && ch2 >= R(1.17549435e-38);
  const R w = rsq(ch2 + sh2);
  const R chn = flag ? w * ch : CSTAR;
  const R shn = flag ? w * sh : SSTAR;
  // cos(2t), sin(2t) of the normalised half-angle pair; the reference also divides both by
  // chn^2 + shn^2, which is 1 up to rounding in either branch
  const R a = chn * chn - shn * shn;
  const R b = R(2) * (chn * shn);
  const R a2 = a * a, b2 = b * b, ab = a * b;
  // S[1][0]=s01, S[2][0]=s02, S[2][1]=s12, S[1][2]=s12, S[0][2]=s02
  const R n0 = a2 * s00 + R(2) * ab * s01 + b2 * s11;
  const R n2 = a * s02 + b * s12;
  const R n3 = (a2 - b2) * s01 + ab * (s11 - s00);
  const R n4 = (s00 + s11) - n0;  // the rotation keeps the trace of the (0,1) block
  const R n5 = a * s12 - b * s02;
  const R n8 = s22;
  // S = mat3(n4, n5, n3, n5, n8, n2, n3, n2, n0)
  s00 = n4; s11 = n8; s22 = n0; s01 = n5; s02 = n3; s12 = n2;
  const R t1 = shn * q[1], t2 = shn * q[2], t3 = shn * q[3];
  const R tmp[3] = {t1, t2, t3};
  const R shq = shn * q[0];
  q[0] *= chn; q[1] *= chn; q[2] *= chn; q[3] *= chn;
  q[z + 1] += shq;
  q[0] -= tmp[z];
  q[x + 1] += tmp[y];
  q[y + 1] -= tmp[x];
}

template <class R>
__device__ __forceinline__ void givens_qr(R a1, R a2, R& ch, R& sh) {
  const R EPS = R(1e-6);
  const R rho = sq_root(a1 * a1 + a2 * a2);
  R s = rho > EPS ? a2 : R(0);
  R c = absv(a1) + (rho > EPS ? rho : EPS);
  const bool b = a1 < R(0);
  const R t = s;
  s = b ? c : s;
  c = b ? t : c;
  const R w = rsq(c * c + s * s);
  ch = c * w;
  sh = s * w;
}

// W, S (upper-triangular R of the QR, as in the reference), V; all column-major.
template <class R>
__device__ __forceinline__ void svd3(const R* A, R* W, R* S, R* V) {
  // A^T A
  R s00 = A[0] * A[0] + A[1] * A[1] + A[2] * A[2];
  R s11 = A[3] * A[3] + A[4] * A[4] + A[5] * A[5];
  R s22 = A[6] * A[6] + A[7] * A[7] + A[8] * A[8];
  R s01 = A[0] * A[3] + A[1] * A[4] + A[2] * A[5];
  R s02 = A[0] * A[6] + A[1] * A[7] + A[2] * A[8];
  R s12 = A[3] * A[6] + A[4] * A[7] + A[5] * A[8];
  R q[4] = {R(1), R(0), R(0), R(0)};
#pragma unroll
  for (int sweep = 0; sweep < 4; ++sweep) {
    jacobi_conj(0, 1, 2, s00, s11, s22, s01, s02, s12, q);
    jacobi_conj(1, 2, 0, s00, s11, s22, s01, s02, s12, q);
    jacobi_conj(2, 0, 1, s00, s11, s22, s01, s02, s12, q);
  }
  quat_to_mat(q[0], q[1], q[2], q[3], V);
#pragma unroll
  for (int i = 0; i < 9; ++i) V[i] = round_to(V[i], R(1e5), R(1e-5));
  R B[9];
  mul(A, V, B);
  // sort columns by squared norm, descending, with negating swaps (SVD.include.glsl:191-228)
  R rho1 = B[0] * B[0] + B[1] * B[1] + B[2] * B[2];
  R rho2 = B[3] * B[3] + B[4] * B[4] + B[5] * B[5];
  R rho3 = B[6] * B[6] + B[7] * B[7] + B[8] * B[8];
  {
    const bool c = rho1 < rho2;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      R tb = -B[r], tv = -V[r];
      B[r] = c ? B[3 + r] : B[r]; B[3 + r] = c ? tb : B[3 + r];
      V[r] = c ? V[3 + r] : V[r]; V[3 + r] = c ? tv : V[3 + r];
    }
    R t = rho1; rho1 = c ? rho2 : rho1; rho2 = c ? t : rho2;
  }
  {
    const bool c = rho1 < rho3;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      R tb = -B[r], tv = -V[r];
      B[r] = c ? B[6 + r] : B[r]; B[6 + r] = c ? tb : B[6 + r];
      V[r] = c ? V[6 + r] : V[r]; V[6 + r] = c ? tv : V[6 + r];
    }
    R t = rho1; rho1 = c ? rho3 : rho1; rho3 = c ? t : rho3;
  }
  {
    const bool c = rho2 < rho3;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      R tb = -B[3 + r], tv = -V[3 + r];
      B[3 + r] = c ? B[6 + r] : B[3 + r]; B[6 + r] = c ? tb : B[6 + r];
      V[3 + r] = c ? V[6 + r] : V[3 + r]; V[6 + r] = c ? tv : V[6 + r];
    }
  }
  // Givens QR (SVD.include.glsl:253-332); R accumulates in S, Q in quaternion form.
#pragma unroll
  for (int i = 0; i < 9; ++i) S[i] = B[i];
  R ch, sh, s0, s1;
  R Q0, Q1, Q2, Q3;
  // 1st rotation eliminates (row1, col0) using rows 0/1
  givens_qr(S[0], S[1], ch, sh);
  s0 = R(1) - R(2) * sh * sh; s1 = R(2) * sh * ch;
#pragma unroll
  for (int c = 0; c < 3; ++c) {  // R = U^T R, U = [[s0,-s1,0],[s1,s0,0],[0,0,1]]
    const R r0 = S[c * 3], r1 = S[c * 3 + 1];
    S[c * 3] = s0 * r0 + s1 * r1;
    S[c * 3 + 1] = -s1 * r0 + s0 * r1;
  }
  Q0 = ch; Q1 = R(0); Q2 = R(0); Q3 = sh;
  // 2nd rotation eliminates (row2, col0) using rows 0/2
  givens_qr(S[0], S[2], ch, sh);
  s0 = R(1) - R(2) * sh * sh; s1 = R(2) * sh * ch;
#pragma unroll
  for (int c = 0; c < 3; ++c) {  // U = [[s0,0,-s1],[0,1,0],[s1,0,s0]]
    const R r0 = S[c * 3], r2 = S[c * 3 + 2];
    S[c * 3] = s0 * r0 + s1 * r2;
    S[c * 3 + 2] = -s1 * r0 + s0 * r2;
  }
  {
    const R q0 = ch * Q0 + sh * Q2, q1 = ch * Q1 + sh * Q3, q2 = ch * Q2 - sh * Q0, q3 = ch * Q3 - sh * Q1;
    Q0 = q0; Q1 = q1; Q2 = q2; Q3 = q3;
  }
  // 3rd rotation eliminates (row2, col1) using rows 1/2
  givens_qr(S[4], S[5], ch, sh);
  s0 = R(1) - R(2) * sh * sh; s1 = R(2) * sh * ch;
#pragma unroll
  for (int c = 0; c < 3; ++c) {  // U = [[1,0,0],[0,s0,-s1],[0,s1,s0]]
    const R r1 = S[c * 3 + 1], r2 = S[c * 3 + 2];
    S[c * 3 + 1] = s0 * r1 + s1 * r2;
    S[c * 3 + 2] = -s1 * r1 + s0 * r2;
  }
  {
    const R q0 = ch * Q0 - sh * Q1, q1 = sh * Q0 + ch * Q1, q2 = ch * Q2 + sh * Q3, q3 = ch * Q3 - sh * Q2;
    Q0 = q0; Q1 = q1; Q2 = q2; Q3 = q3;
  }
  quat_to_mat(Q0, Q1, Q2, Q3, W);
}

// Force matrix A with f_i = -A grad(w_ip) (p2gTransfer.glsl:109-141), times grad_scale:
// A = gs * (m/rho0) * (2 mu(J_P) (F_E - R_E) F_E^T + lambda(J_P) (J_E - 1) J_E I),
// R_E = W V^T rounded to 1e-9 (p2gTransfer.glsl:112-116), from given SVD factors W, V of F_E.
// Zero when rho0 <= 0.
template <class R>
__device__ __forceinline__ void force_matrix_wv(const R* FE, const R* FP, const R* W, const R* V, R m, R rho0,
                                                R mu0, R lambda0, R xi, R gs, R* Aout) {
  R Rot[9];
  mul_bt(W, V, Rot);
#pragma unroll
  for (int i = 0; i < 9; ++i) Rot[i] = round_to(Rot[i], R(1e9), R(1e-9));
  const R JP = det3(FP), JE = det3(FE);
  const R hf = hardening_factor(xi, JP);
  const R mu = mu0 * hf, la = lambda0 * hf;
  R D[9], P[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) D[i] = FE[i] - Rot[i];
  mul_bt(D, FE, P);  // (F_E - R_E) F_E^T
  const R lterm = la * (JE - R(1)) * JE;
  const R k = rho0 > R(0) ? gs * (m * recip(rho0)) : R(0);  // m / rho0 within 2 ulp (GLSL: 2.5)
#pragma unroll
  for (int i = 0; i < 9; ++i) Aout[i] = k * (R(2) * mu * P[i] + ((i == 0 || i == 4 || i == 8) ? lterm : R(0)));
}

// The same with the SVD of F_E computed here (SVD.include.glsl via p2gTransfer.glsl:109-111).
template <class R>
__device__ __forceinline__ void force_matrix(const R* FE, const R* FP, R m, R rho0, R mu0, R lambda0,
                                             R xi, R gs, R* Aout) {
  R W[9], S[9], V[9];
  svd3(FE, W, S, V);
  force_matrix_wv(FE, FP, W, V, m, rho0, mu0, lambda0, xi, gs, Aout);
}

// ---- colliders (gForceCompute.glsl:40-93, particleAdvance.glsl:60-73/182-204) ----------------
struct Collider {
  float px, py, pz, type;
  float vx, vy, vz, radius;
  float nx, ny, nz, friction;
};

__device__ __forceinline__ void collide(float x, float y, float z, float& vx, float& vy, float& vz,
                                        const Collider* cs, int nc) {
  for (int i = 0; i < nc; ++i) {
    const Collider c = cs[i];
    const unsigned t = unsigned(c.type);
    float nx = c.nx, ny = c.ny, nz = c.nz;
    const float dx = x - c.px, dy = y - c.py, dz = z - c.pz;
    bool hit = false;
    if (t == 0u) {
      hit = dx * nx + dy * ny + dz * nz <= 0.0f;
    } else if (t == 1u) {
      const float len = sqrtf(dx * dx + dy * dy + dz * dz);
      nx = dx / len; ny = dy / len; nz = dz / len;
      hit = len < c.radius;
    } else if (t == 2u) {  // extension: axis-aligned box, half-extents in the normal slot
      const float px = c.nx - fabsf(dx), py = c.ny - fabsf(dy), pz = c.nz - fabsf(dz);
      hit = px > 0.0f && py > 0.0f && pz > 0.0f;
      int a = 0;
      float pm = px;
      if (py < pm) { a = 1; pm = py; }
      if (pz < pm) { a = 2; }
      nx = a == 0 ? (dx < 0.0f ? -1.0f : 1.0f) : 0.0f;
      ny = a == 1 ? (dy < 0.0f ? -1.0f : 1.0f) : 0.0f;
      nz = a == 2 ? (dz < 0.0f ? -1.0f : 1.0f) : 0.0f;
    }
    if (hit) {
      const float rx = vx - c.vx, ry = vy - c.vy, rz = vz - c.vz;
      const float vn = rx * nx + ry * ny + rz * nz;
      if (vn < 0.0f) {
        float tx = rx - nx * vn, ty = ry - ny * vn, tz = rz - nz * vn;
        const float muvn = c.friction * vn;
        const float lt = sqrtf(tx * tx + ty * ty + tz * tz);
        if (lt <= -muvn) {
          tx = 0.0f; ty = 0.0f; tz = 0.0f;
        } else {
          const float k = muvn / lt;
          tx += tx * k; ty += ty * k; tz += tz * k;
        }
        vx = tx + c.vx; vy = ty + c.vy; vz = tz + c.vz;
      }
    }
  }
}

// Knuth TwoSum accumulation of x into the pair (hi, lo).
__device__ __forceinline__ void two_sum_acc(float& hi, float& lo, float x) {
  const float s = hi + x;
  const float bb = s - hi;
  const float e = (hi - (s - bb)) + (x - bb);
  hi = s;
  lo += e;
}

}  // namespace snow
