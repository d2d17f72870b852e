// mpm_oracle.hpp - CPU restatement of the reference snow MPM timestep.
//
// TEST INFRASTRUCTURE ONLY. This is the correctness oracle (Real = double) and the
// CPU baseline (Real = float, multithreaded) for the MI355X snow step. Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only as
// the checker; the product (snow_amd/) never links or calls it.
//
// Each function restates one GLSL function of the reference line by line (GLSL
// mat3 is column-major, M[col][row]; kept literally here as m[col][row]):
//   weighting / weightingGradient ..... shader/compute/interpolation/cubic.include.glsl:7-40
//   getIJK / getIndex .................. shader/compute/indexing/neighborIndex.include.glsl:6-12,
//                                        shader/compute/indexing/gridIndex.include.glsl:7-9
//   mu / lambda ........................ shader/compute/lame/lame.include.glsl:11-27
//   SVD (McAdams et al. 2011) ........... shader/compute/svd/SVD.include.glsl:24-385
//   grid mass / particle volume ........ shader/compute/mpm/gMassCompute.glsl:33-84,
//                                        shader/compute/mpm/pVolumeCompute.glsl:32-83
//   P2G ................................ shader/compute/mpm/p2gTransfer.glsl:44-151
//   grid update + collisions ........... shader/compute/mpm/gForceCompute.glsl:40-103
//   G2P ................................ shader/compute/mpm/g2pTransfer.glsl:53-143
//   particle advance ................... shader/compute/mpm/particleAdvance.glsl:56-217
//   rigid advance ...................... shader/compute/mpm/rigidAdvance.glsl:14-18
//   stage order ........................ src/snow/simulation/timeUpdate/explicittimeupdate.cpp:4-63
//   semi-implicit step ................. semiimplicittimeupdate.cpp:3-5 is empty; restated from the
//                                        published method (Stomakhin et al. 2013, eq. 9), see below
//
// Deliberate differences from the reference, all recorded as quirk decisions:
//  * accumulation is in Real (not int32 fixed point x1e8 / x1e6, p2gTransfer.glsl:98-146);
//  * the polar decomposition of F_E is computed once per particle instead of once per
//    (particle, stencil node) (p2gTransfer.glsl:109-110); the result is identical;
//  * grid reset / update are bounds-checked (gReset.glsl:13 and gForceCompute.glsl:58 are not);
//  * the weight gradient is multiplied by grad_scale (1 = reference, which omits 1/h);
//  * collider type 2 (axis-aligned box) is an extension; types 0/1 are as in the reference.
// Pinning: tests/test_oracle.py checks every function here against closed forms,
// invariants and brute force (the reference ships no test vectors).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>
#include <algorithm>
#include <functional>

#include "../include/snow.h"

namespace oracle {

template <class R>
struct Mat3 {
  R m[3][3];  // m[col][row]
  static Mat3 zero() { Mat3 a; std::memset(a.m, 0, sizeof(a.m)); return a; }
  static Mat3 ident(R d = R(1)) { Mat3 a = zero(); a.m[0][0] = a.m[1][1] = a.m[2][2] = d; return a; }
  // GLSL mat3(c0r0, c0r1, c0r2, c1r0, ...)
  static Mat3 cols(R a0, R a1, R a2, R a3, R a4, R a5, R a6, R a7, R a8) {
    Mat3 a; a.m[0][0]=a0; a.m[0][1]=a1; a.m[0][2]=a2; a.m[1][0]=a3; a.m[1][1]=a4; a.m[1][2]=a5;
    a.m[2][0]=a6; a.m[2][1]=a7; a.m[2][2]=a8; return a;
  }
};
template <class R> struct Vec3 { R x, y, z; R& operator[](int i) { return (&x)[i]; } R operator[](int i) const { return (&x)[i]; } };

template <class R> inline Vec3<R> v3(R a, R b, R c) { return Vec3<R>{a, b, c}; }
template <class R> inline Vec3<R> operator+(Vec3<R> a, Vec3<R> b) { return v3(a.x+b.x, a.y+b.y, a.z+b.z); }
template <class R> inline Vec3<R> operator-(Vec3<R> a, Vec3<R> b) { return v3(a.x-b.x, a.y-b.y, a.z-b.z); }
template <class R> inline Vec3<R> operator*(Vec3<R> a, R s) { return v3(a.x*s, a.y*s, a.z*s); }
template <class R> inline Vec3<R> operator/(Vec3<R> a, R s) { return v3(a.x/s, a.y/s, a.z/s); }
template <class R> inline R dot(Vec3<R> a, Vec3<R> b) { return a.x*b.x + a.y*b.y + a.z*b.z; }
template <class R> inline R length(Vec3<R> a) { return std::sqrt(dot(a, a)); }
template <class R> inline Vec3<R> normalize(Vec3<R> a) { return a / length(a); }

template <class R> inline Mat3<R> operator*(const Mat3<R>& A, const Mat3<R>& B) {
  Mat3<R> C;
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) C.m[c][r] = A.m[0][r]*B.m[c][0] + A.m[1][r]*B.m[c][1] + A.m[2][r]*B.m[c][2];
  return C;
}
template <class R> inline Vec3<R> operator*(const Mat3<R>& A, Vec3<R> v) {
  Vec3<R> o;
  for (int r = 0; r < 3; ++r) o[r] = A.m[0][r]*v.x + A.m[1][r]*v.y + A.m[2][r]*v.z;
  return o;
}
template <class R> inline Mat3<R> operator+(const Mat3<R>& A, const Mat3<R>& B) {
  Mat3<R> C; for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) C.m[c][r] = A.m[c][r] + B.m[c][r]; return C;
}
template <class R> inline Mat3<R> operator-(const Mat3<R>& A, const Mat3<R>& B) {
  Mat3<R> C; for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) C.m[c][r] = A.m[c][r] - B.m[c][r]; return C;
}
template <class R> inline Mat3<R> operator*(const Mat3<R>& A, R s) {
  Mat3<R> C; for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) C.m[c][r] = A.m[c][r] * s; return C;
}
template <class R> inline Mat3<R> operator*(R s, const Mat3<R>& A) { return A * s; }
template <class R> inline Mat3<R> transpose(const Mat3<R>& A) {
  Mat3<R> C; for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) C.m[c][r] = A.m[r][c]; return C;
}
template <class R> inline R determinant(const Mat3<R>& A) {
  const auto& m = A.m;
  return m[0][0]*(m[1][1]*m[2][2] - m[2][1]*m[1][2]) - m[1][0]*(m[0][1]*m[2][2] - m[2][1]*m[0][2]) +
         m[2][0]*(m[0][1]*m[1][2] - m[1][1]*m[0][2]);
}
template <class R> inline Vec3<R> column(int i, const Mat3<R>& A) { return v3(A.m[i][0], A.m[i][1], A.m[i][2]); }
template <class R> inline Mat3<R> from_cols(Vec3<R> a, Vec3<R> b, Vec3<R> c) {
  return Mat3<R>::cols(a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z);
}
template <class R> inline R rsqrt(R x) { return R(1) / std::sqrt(x); }
// GLSL round(): nearest, halves away from zero here (the GLSL spec leaves halves open).
template <class R> inline R round_to(R x, R scale) { return std::round(scale * x) / scale; }

// ---- cubic.include.glsl:7-40 --------------------------------------------------
template <class R> inline R weighting(R x) {
  const R absX = std::fabs(x);
  if (absX < R(1)) return R(0.5) * absX * absX * absX - x * x + R(2) / R(3);
  if (absX < R(2)) return -R(1) / R(6) * absX * absX * absX + x * x - R(2) * absX + R(4) / R(3);
  return R(0);
}
template <class R> inline R weightingGradient(R x) {
  const R absX = std::fabs(x);
  if (absX < R(1)) return R(1.5) * x * absX - R(2) * x;
  if (absX < R(2)) return -R(1) / R(2) * absX * x + R(2) * x - R(2) * x / absX;
  return R(0);
}
template <class R> inline R weighting3(Vec3<R> d) { return weighting(d.x) * weighting(d.y) * weighting(d.z); }
template <class R> inline Vec3<R> weightingGradient3(Vec3<R> d) {
  return v3(weightingGradient(d.x) * weighting(d.y) * weighting(d.z),
            weighting(d.x) * weightingGradient(d.y) * weighting(d.z),
            weighting(d.x) * weighting(d.y) * weightingGradient(d.z));
}

// ---- neighborIndex.include.glsl:6-12: s in [0,64) -> offset in {-1..2}^3 --------
inline void getIJK(int index, int ijk[3]) {
  const int width = 4;
  int temp = index % (width * width);
  ijk[0] = temp % width - 1;
  ijk[1] = temp / width - 1;
  ijk[2] = index / (width * width) - 1;
}

// ---- lame.include.glsl:11-27 ----------------------------------------------------
template <class R> inline R mu(R young, R poisson, R hardening, R JP) {
  return young / (R(2) * (R(1) + poisson)) * std::exp(hardening * (R(1) - JP));
}
template <class R> inline R lambda(R young, R poisson, R hardening, R JP) {
  return young * poisson / ((R(1) + poisson) * (R(1) - R(2) * poisson)) * std::exp(hardening * (R(1) - JP));
}

// ---- SVD.include.glsl:24-385 (McAdams, Selle, Tamstorf, Teran, Sifakis 2011) -------
template <class R> struct SVDConst {
  static constexpr double GAMMA = 5.828427124;      // 3 + sqrt(8)
  static constexpr double CSTAR = 0.923879532;      // cos(pi/8)
  static constexpr double SSTAR = 0.3826834323;     // sin(pi/8)
  static constexpr double EPSILON = 1e-6;
};

template <class R> inline void fromQuat(const R q[4], Mat3<R>& M) {
  R qxx = q[1]*q[1], qyy = q[2]*q[2], qzz = q[3]*q[3];
  R qxz = q[1]*q[3], qxy = q[1]*q[2], qyz = q[2]*q[3];
  R qwx = q[0]*q[1], qwy = q[0]*q[2], qwz = q[0]*q[3];
  M.m[0][0] = R(1) - R(2)*(qyy+qzz); M.m[0][1] = R(2)*(qxy+qwz); M.m[0][2] = R(2)*(qxz-qwy);
  M.m[1][0] = R(2)*(qxy-qwz); M.m[1][1] = R(1) - R(2)*(qxx+qzz); M.m[1][2] = R(2)*(qyz+qwx);
  M.m[2][0] = R(2)*(qxz+qwy); M.m[2][1] = R(2)*(qyz-qwx); M.m[2][2] = R(1) - R(2)*(qxx+qyy);
}

// SVD.include.glsl:71-128
template <class R> inline void jacobiConjugation(int x, int y, int z, Mat3<R>& S, R qV[4]) {
  const R GAMMA = R(SVDConst<R>::GAMMA), CSTAR = R(SVDConst<R>::CSTAR), SSTAR = R(SVDConst<R>::SSTAR);
  R ch = R(2) * (S.m[0][0] - S.m[1][1]), ch2 = ch * ch;
  R sh = S.m[0][1], sh2 = sh * sh;
  bool flag = (GAMMA * sh2 < ch2);
  R w = rsqrt(ch2 + sh2);
  ch = flag ? w * ch : CSTAR; ch2 = ch * ch;
  sh = flag ? w * sh : SSTAR; sh2 = sh * sh;
  R scale = R(1) / (ch2 + sh2);
  R a = (ch2 - sh2) * scale;
  R b = (R(2) * sh * ch) * scale;
  R a2 = a * a, b2 = b * b, ab = a * b;
  R s0 = a2 * S.m[0][0] + R(2) * ab * S.m[1][0] + b2 * S.m[1][1];
  R s2 = a * S.m[2][0] + b * S.m[2][1];
  R s3 = (a2 - b2) * S.m[1][0] + ab * (S.m[1][1] - S.m[0][0]);
  R s4 = b2 * S.m[0][0] - R(2) * ab * S.m[1][0] + a2 * S.m[1][1];
  R s5 = a * S.m[1][2] - b * S.m[0][2];
  R s8 = S.m[2][2];
  S = Mat3<R>::cols(s4, s5, s3, s5, s8, s2, s3, s2, s0);
  R tmp[3] = {sh * qV[1], sh * qV[2], sh * qV[3]};
  sh *= qV[0];
  for (int i = 0; i < 4; ++i) qV[i] *= ch;
  qV[z + 1] += sh;
  qV[0] -= tmp[z];
  qV[x + 1] += tmp[y];
  qV[y + 1] -= tmp[x];
}

// SVD.include.glsl:130-162: 4 sweeps of the 3 cyclic conjugations.
template <class R> inline void jacobiEigenanalysis(Mat3<R>& S, R qV[4]) {
  qV[0] = R(1); qV[1] = qV[2] = qV[3] = R(0);
  for (int sweep = 0; sweep < 4; ++sweep) {
    jacobiConjugation(0, 1, 2, S, qV);
    jacobiConjugation(1, 2, 0, S, qV);
    jacobiConjugation(2, 0, 1, S, qV);
  }
}

template <class R> inline void condSwap(bool c, R& x, R& y) { R t = x; x = c ? y : x; y = c ? t : y; }
template <class R> inline void condNegSwap(bool c, Vec3<R>& x, Vec3<R>& y) {
  Vec3<R> t = v3(-x.x, -x.y, -x.z); x = c ? y : x; y = c ? t : y;
}

// SVD.include.glsl:191-228
template <class R> inline void sortSingularValues(Mat3<R>& B, Mat3<R>& V) {
  Vec3<R> b1 = column(0, B), v1 = column(0, V);
  Vec3<R> b2 = column(1, B), v2 = column(1, V);
  Vec3<R> b3 = column(2, B), v3c = column(2, V);
  R rho1 = dot(b1, b1), rho2 = dot(b2, b2), rho3 = dot(b3, b3);
  bool c;
  c = rho1 < rho2; condNegSwap(c, b1, b2); condNegSwap(c, v1, v2); condSwap(c, rho1, rho2);
  c = rho1 < rho3; condNegSwap(c, b1, b3); condNegSwap(c, v1, v3c); condSwap(c, rho1, rho3);
  c = rho2 < rho3; condNegSwap(c, b2, b3); condNegSwap(c, v2, v3c);
  B = from_cols(b1, b2, b3);
  V = from_cols(v1, v2, v3c);
}

// SVD.include.glsl:231-248
template <class R> inline void QRGivensQuaternion(R a1, R a2, R& ch, R& sh) {
  const R EPSILON = R(SVDConst<R>::EPSILON);
  R rho = std::sqrt(a1 * a1 + a2 * a2);
  sh = rho > EPSILON ? a2 : R(0);
  ch = std::fabs(a1) + std::max(rho, EPSILON);
  bool b = a1 < R(0);
  condSwap(b, sh, ch);
  R w = rsqrt(ch * ch + sh * sh);
  ch *= w;
  sh *= w;
}

// SVD.include.glsl:253-332
template <class R> inline void QRDecomposition(const Mat3<R>& B, Mat3<R>& Q, Mat3<R>& Rm) {
  Rm = B;
  R qQ[4];
  R ch, sh, s0, s1;
  QRGivensQuaternion(Rm.m[0][0], Rm.m[0][1], ch, sh);
  s0 = R(1) - R(2) * sh * sh; s1 = R(2) * sh * ch;
  Mat3<R> U = Mat3<R>::cols(s0, s1, R(0), -s1, s0, R(0), R(0), R(0), R(1));
  Rm = transpose(U) * Rm;
  qQ[0] = R(1); qQ[1] = qQ[2] = qQ[3] = R(0);
  R q0 = ch * qQ[0] - sh * qQ[3], q1 = ch * qQ[1] + sh * qQ[2];
  R q2 = ch * qQ[2] - sh * qQ[1], q3 = sh * qQ[0] + ch * qQ[3];
  qQ[0] = q0; qQ[1] = q1; qQ[2] = q2; qQ[3] = q3;

  QRGivensQuaternion(Rm.m[0][0], Rm.m[0][2], ch, sh);
  s0 = R(1) - R(2) * sh * sh; s1 = R(2) * sh * ch;
  U = Mat3<R>::cols(s0, R(0), s1, R(0), R(1), R(0), -s1, R(0), s0);
  Rm = transpose(U) * Rm;
  q0 = ch * qQ[0] + sh * qQ[2]; q1 = ch * qQ[1] + sh * qQ[3];
  q2 = ch * qQ[2] - sh * qQ[0]; q3 = ch * qQ[3] - sh * qQ[1];
  qQ[0] = q0; qQ[1] = q1; qQ[2] = q2; qQ[3] = q3;

  QRGivensQuaternion(Rm.m[1][1], Rm.m[1][2], ch, sh);
  s0 = R(1) - R(2) * sh * sh; s1 = R(2) * sh * ch;
  U = Mat3<R>::cols(R(1), R(0), R(0), R(0), s0, s1, R(0), -s1, s0);
  Rm = transpose(U) * Rm;
  q0 = ch * qQ[0] - sh * qQ[1]; q1 = sh * qQ[0] + ch * qQ[1];
  q2 = ch * qQ[2] + sh * qQ[3]; q3 = ch * qQ[3] - sh * qQ[2];
  qQ[0] = q0; qQ[1] = q1; qQ[2] = q2; qQ[3] = q3;
  fromQuat(qQ, Q);
}

// SVD.include.glsl:334-371 (V rounded to 1e-5 at 354-358)
template <class R> inline void computeSVD(const Mat3<R>& A, Mat3<R>& W, Mat3<R>& S, Mat3<R>& V) {
  Mat3<R> ATA = transpose(A) * A;
  R qV[4];
  jacobiEigenanalysis(ATA, qV);
  fromQuat(qV, V);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) V.m[i][j] = round_to(V.m[i][j], R(1e5));
  Mat3<R> B = A * V;
  sortSingularValues(B, V);
  QRDecomposition(B, W, S);
}

// SVD.include.glsl:374-385 (P is unused by every caller)
template <class R> inline void computePD(const Mat3<R>& A, Mat3<R>& Rot) {
  Mat3<R> W, S, V;
  computeSVD(A, W, S, V);
  Rot = W * transpose(V);
}

// ---- collisions: gForceCompute.glsl:40-53, particleAdvance.glsl:60-73 -------------
template <class R> struct Collider { Vec3<R> pos; R type; Vec3<R> vel; R radius; Vec3<R> normal; R friction; };

// Returns whether x lies in collider c; sets n to the contact normal.
template <class R> inline bool collides(Vec3<R> x, const Collider<R>& c, Vec3<R>& n) {
  unsigned t = unsigned(c.type);
  if (t == 0u) return dot(x - c.pos, c.normal) <= R(0);
  if (t == 1u) {
    n = normalize(x - c.pos);
    return length(x - c.pos) < c.radius;
  }
  if (t == 2u) {  // extension: axis-aligned box, centre pos, half-extents normal
    Vec3<R> d = x - c.pos;
    R pen[3];
    for (int k = 0; k < 3; ++k) {
      pen[k] = c.normal[k] - std::fabs(d[k]);
      if (!(pen[k] > R(0))) return false;
    }
    int a = 0;
    if (pen[1] < pen[a]) a = 1;
    if (pen[2] < pen[a]) a = 2;
    n = v3(R(0), R(0), R(0));
    n[a] = d[a] < R(0) ? R(-1) : R(1);
    return true;
  }
  return false;
}

// Sequential Coulomb response of gForceCompute.glsl:72-93 / particleAdvance.glsl:182-204.
template <class R> inline Vec3<R> collide_velocity(Vec3<R> x, Vec3<R> vin, const Collider<R>* cs, int nc) {
  for (int i = 0; i < nc; ++i) {
    Vec3<R> n = cs[i].normal;
    if (collides(x, cs[i], n)) {
      Vec3<R> vco = cs[i].vel;
      Vec3<R> vrel = vin - vco;
      R vn = dot(vrel, n);
      if (vn < R(0)) {
        Vec3<R> vt = vrel - n * vn;
        R muvn = cs[i].friction * vn;
        Vec3<R> vrelt = vt;
        R lengthvt = length(vt);
        if (lengthvt <= -muvn) vrelt = v3(R(0), R(0), R(0));
        else vrelt = vrelt + vt * (muvn / lengthvt);
        vin = vrelt + vco;
      }
    }
  }
  return vin;
}

// ---- scene state -----------------------------------------------------------------
template <class R> struct Scene {
  int64_t n = 0;
  R *x = nullptr, *v = nullptr, *m = nullptr, *fe = nullptr, *fp = nullptr, *rho0 = nullptr;
  int D[3] = {0, 0, 0};
  R origin[3] = {0, 0, 0};
  R h = 1;
  // material
  R young = 0, poisson = 0, hardening = 0, crit_comp = 0, crit_stretch = 0, alpha = 0.95;
  Vec3<R> g{0, 0, 0};
  R gscale = 1;
  int nc = 0;
  Collider<R>* col = nullptr;
  // dense grid fields
  R *gm = nullptr, *gmv = nullptr, *gf = nullptr, *gvstar = nullptr;
  int nthreads = 1;
  int64_t nodes() const { return int64_t(D[0]) * D[1] * D[2]; }
  Vec3<R> xp(int64_t p) const { return v3(x[3*p], x[3*p+1], x[3*p+2]); }
  Mat3<R> FE(int64_t p) const { Mat3<R> a; std::memcpy(a.m, fe + 9*p, sizeof(a.m)); return a; }
  Mat3<R> FP(int64_t p) const { Mat3<R> a; std::memcpy(a.m, fp + 9*p, sizeof(a.m)); return a; }
};

inline void parallel_for(int nthreads, int64_t n, const std::function<void(int, int64_t, int64_t)>& fn) {
  if (nthreads <= 1 || n < 2) { fn(0, 0, n); return; }
  std::vector<std::thread> th;
  int64_t chunk = (n + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; ++t) {
    int64_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    th.emplace_back(fn, t, b, e);
  }
  for (auto& t : th) t.join();
}

// ParticleInGrid = (xp - gGridPos)/gridSpacing; base = ivec3(ParticleInGrid) (truncation).
template <class R> inline void particle_in_grid(const Scene<R>& s, Vec3<R> x, Vec3<R>& u, int base[3]) {
  u = v3((x.x - s.origin[0]) / s.h, (x.y - s.origin[1]) / s.h, (x.z - s.origin[2]) / s.h);
  for (int k = 0; k < 3; ++k) base[k] = int(u[k]);  // C++ int() truncates toward zero like GLSL
}

// The factors of weighting3(d) and weightingGradient3(-d) for the 64 stencil nodes of one
// particle, tabulated per axis: d[k] = (base[k] + o) - u[k] for the offsets o = -1..2 of getIJK.
// weighting() is even in its argument (|x| and x*x only), so weighting(-d) = weighting(d) bit
// for bit, and the per-node products below are the same products, in the same order, as the
// vec3 overloads (cubic.include.glsl:36-40) evaluate: w = (wx*wy)*wz, grad.x = (gx*wy)*wz,
// grad.y = (wx*gy)*wz, grad.z = (wx*wy)*gz.
template <class R> struct StencilWeights {
  R w[3][4], g[3][4];
  StencilWeights(Vec3<R> u, const int base[3]) {
    for (int k = 0; k < 3; ++k)
      for (int o = 0; o < 4; ++o) {
        const R d = R(base[k] + o - 1) - u[k];
        w[k][o] = weighting(d);
        g[k][o] = weightingGradient(-d);
      }
  }
  // o: getIJK offsets in {-1..2}
  R weight(const int o[3]) const { return w[0][o[0] + 1] * w[1][o[1] + 1] * w[2][o[2] + 1]; }
  Vec3<R> gradient(const int o[3]) const {
    const R wx = w[0][o[0] + 1], wy = w[1][o[1] + 1], wz = w[2][o[2] + 1];
    return v3(g[0][o[0] + 1] * wy * wz, wx * g[1][o[1] + 1] * wz, wx * wy * g[2][o[2] + 1]);
  }
};

// gReset.glsl:11-16 (bounds-checked)
template <class R> void grid_reset(Scene<R>& s) {
  int64_t nn = s.nodes();
  std::fill(s.gm, s.gm + nn, R(0));
  std::fill(s.gmv, s.gmv + 3*nn, R(0));
  std::fill(s.gf, s.gf + 3*nn, R(0));
}

// Per-particle force matrix A_p with f_i = -A_p * grad w_ip (p2gTransfer.glsl:109-141):
// A_p = (m_p/rho_p0) * (2 mu(J_P) (F_E - R_E) F_E^T + lambda(J_P) (J_E - 1) J_E I).
template <class R> inline Mat3<R> p2g_force_matrix(const Scene<R>& s, int64_t p) {
  Mat3<R> FEp = s.FE(p), FPp = s.FP(p);
  Mat3<R> REp;
  computePD(FEp, REp);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) REp.m[i][j] = round_to(REp.m[i][j], R(1e9));
  R JPp = determinant(FPp), JEp = determinant(FEp);
  R mp = s.m[p], pp0 = s.rho0[p];
  return (mp / pp0) * (R(2) * mu(s.young, s.poisson, s.hardening, JPp) * ((FEp - REp) * transpose(FEp)) +
                       lambda(s.young, s.poisson, s.hardening, JPp) * (JEp - R(1)) * JEp * Mat3<R>::ident());
}

// P2G scatter into the nodes whose z lies in [k0,k1): every particle visits its 64 stencil
// offsets (p2gTransfer.glsl:70-149) in getIJK order. Particles are visited bucketed by base z,
// so each node sums its contributions in the same order whatever the thread count.
template <class R> void p2g_scatter(Scene<R>& s, bool mass_only, const std::vector<Mat3<R>>& A,
                                    const std::vector<std::vector<int64_t>>& zb, int k0, int k1) {
  const int Dx = s.D[0], Dy = s.D[1], Dz = s.D[2];
  for (int bz = k0 - 2; bz <= k1 + 1; ++bz) {
    if (bz < -2 || bz > Dz) continue;
    for (int64_t p : zb[bz + 2]) {
      R mp = s.m[p];
      if (!mass_only && std::fabs(mp) > R(100)) continue;  // p2gTransfer.glsl:64-66
      Vec3<R> vp = v3(s.v[3*p], s.v[3*p+1], s.v[3*p+2]);
      Vec3<R> u; int base[3];
      particle_in_grid(s, s.xp(p), u, base);
      const StencilWeights<R> sw(u, base);
      for (int sidx = 0; sidx < 64; ++sidx) {
        int o[3]; getIJK(sidx, o);
        int gi[3] = {base[0] + o[0], base[1] + o[1], base[2] + o[2]};
        if (!(gi[0] >= 0 && gi[1] >= 0 && gi[2] >= 0 && gi[0] < Dx && gi[1] < Dy && gi[2] < Dz)) continue;
        if (gi[2] < k0 || gi[2] >= k1) continue;
        R wip = sw.weight(o);  // weighting3(gi - u)
        int64_t gI = gi[0] + int64_t(gi[1]) * Dx + int64_t(gi[2]) * Dx * Dy;
        s.gm[gI] += mp * wip;
        if (mass_only) continue;
        Vec3<R> vel = vp * mp * wip;
        s.gmv[3*gI] += vel.x; s.gmv[3*gI+1] += vel.y; s.gmv[3*gI+2] += vel.z;
        if (s.rho0[p] > R(0)) {  // p2gTransfer.glsl:126
          Vec3<R> wipg = sw.gradient(o) * s.gscale;  // weightingGradient3(u - gi)
          Vec3<R> force = A[p] * wipg;
          s.gf[3*gI] -= force.x; s.gf[3*gI+1] -= force.y; s.gf[3*gI+2] -= force.z;
        }
      }
    }
  }
}

template <class R> std::vector<std::vector<int64_t>> z_buckets(const Scene<R>& s) {
  std::vector<std::vector<int64_t>> zb(s.D[2] + 3);
  for (int64_t p = 0; p < s.n; ++p) {
    Vec3<R> u; int base[3];
    particle_in_grid(s, s.xp(p), u, base);
    if (base[2] >= -2 && base[2] <= s.D[2]) zb[base[2] + 2].push_back(p);
  }
  return zb;
}

template <class R> void p2g(Scene<R>& s, bool mass_only) {
  grid_reset(s);
  std::vector<Mat3<R>> A;
  if (!mass_only) {
    A.resize(s.n);
    parallel_for(s.nthreads, s.n, [&](int, int64_t b, int64_t e) {
      for (int64_t p = b; p < e; ++p)
        A[p] = s.rho0[p] > R(0) ? p2g_force_matrix(s, p) : Mat3<R>::zero();
    });
  }
  auto zb = z_buckets(s);
  // Thread t owns the node layers [kb[t], kb[t+1]), placed so that every thread visits about the
  // same number of particles. Each node is summed by one thread, bucket by bucket in increasing
  // base z and particle index order, so the sums do not depend on where the boundaries fall.
  const int T = std::max(1, std::min(s.nthreads, s.D[2]));
  std::vector<int> kb(T + 1, s.D[2]);
  {
    std::vector<int64_t> cum(s.D[2] + 1, 0);  // particles with base z < k (layers -2..D)
    for (int k = 0; k < s.D[2]; ++k) cum[k + 1] = cum[k] + int64_t(zb[k + 2].size()) + (k == 0 ? int64_t(zb[0].size() + zb[1].size()) : 0);
    const int64_t tot = cum[s.D[2]];
    kb[0] = 0;
    int k = 0;
    for (int t = 1; t < T; ++t) {
      const int64_t want = tot * t / T;
      while (k < s.D[2] && cum[k] < want) ++k;
      kb[t] = std::max(kb[t - 1], k);
    }
  }
  parallel_for(T, T, [&](int, int64_t tb, int64_t te) {
    for (int64_t t = tb; t < te; ++t) p2g_scatter(s, mass_only, A, zb, kb[t], kb[t + 1]);
  });
}

// pVolumeCompute.glsl:32-83: rho_p0 = sum_i m_i w_ip / h^3.
template <class R> void particle_volume(Scene<R>& s) {
  const int Dx = s.D[0], Dy = s.D[1], Dz = s.D[2];
  const R cell = s.h * s.h * s.h;
  parallel_for(s.nthreads, s.n, [&](int, int64_t b, int64_t e) {
    for (int64_t p = b; p < e; ++p) {
      Vec3<R> u; int base[3];
      particle_in_grid(s, s.xp(p), u, base);
      const StencilWeights<R> sw(u, base);
      R acc = 0;
      for (int sidx = 0; sidx < 64; ++sidx) {
        int o[3]; getIJK(sidx, o);
        int gi[3] = {base[0] + o[0], base[1] + o[1], base[2] + o[2]};
        if (!(gi[0] >= 0 && gi[1] >= 0 && gi[2] >= 0 && gi[0] < Dx && gi[1] < Dy && gi[2] < Dz)) continue;
        R wip = sw.weight(o);  // weighting3(gi - u)
        int64_t gI = gi[0] + int64_t(gi[1]) * Dx + int64_t(gi[2]) * Dx * Dy;
        acc += s.gm[gI] * wip / cell;
      }
      s.rho0[p] = acc;
    }
  });
}

// ExplicitTimeUpdate::init (explicittimeupdate.cpp:29-33): reset, grid mass, volumes.
template <class R> void init(Scene<R>& s) {
  p2g(s, /*mass_only=*/true);
  particle_volume(s);
}

// rigidAdvance.glsl:14-18
template <class R> void rigid_advance(Scene<R>& s, R dt) {
  for (int i = 0; i < s.nc; ++i) s.col[i].pos = s.col[i].pos + s.col[i].vel * dt;
}

// Node coordinate as the reference stores it in the gxm buffer: computed on the host in float,
// x_off + h * (float)i with two roundings (grid.cpp:98-102), and read by the shader as data. (In
// exact arithmetic a node can lie exactly on a collider plane; the float value decides.)
inline float node_coord(float x_off, float h, int64_t i) { return x_off + h * float(i); }

// gForceCompute.glsl:57-103 (bounds-checked).
template <class R> void grid_update(Scene<R>& s, R dt) {
  const int Dx = s.D[0], Dy = s.D[1];
  parallel_for(s.nthreads, s.nodes(), [&](int, int64_t b, int64_t e) {
    for (int64_t gI = b; gI < e; ++gI) {
      int64_t i = gI % Dx, j = (gI / Dx) % Dy, k = gI / (int64_t(Dx) * Dy);
      const float hf = float(s.h);
      Vec3<R> xi = v3(R(node_coord(float(s.origin[0]), hf, i)), R(node_coord(float(s.origin[1]), hf, j)),
                      R(node_coord(float(s.origin[2]), hf, k)));
      R mi = s.gm[gI];
      Vec3<R> vi = v3(s.gmv[3*gI], s.gmv[3*gI+1], s.gmv[3*gI+2]);
      Vec3<R> fi = v3(s.gf[3*gI], s.gf[3*gI+1], s.gf[3*gI+2]);
      Vec3<R> vin = v3(R(0), R(0), R(0));
      if (mi > R(0)) {
        vin = vi / mi + (fi / mi + s.g) * dt;
        vin = collide_velocity(xi, vin, s.col, s.nc);
      }
      s.gvstar[3*gI] = vin.x; s.gvstar[3*gI+1] = vin.y; s.gvstar[3*gI+2] = vin.z;
    }
  });
}

// g2pTransfer.glsl:53-143 followed by particleAdvance.glsl:75-217 for one particle.
template <class R> void g2p_advance_particle(Scene<R>& s, int64_t p, R dt) {
  const int Dx = s.D[0], Dy = s.D[1], Dz = s.D[2];
  const R alpha = s.alpha;
  Vec3<R> xp = s.xp(p);
  Vec3<R> u; int base[3];
  particle_in_grid(s, xp, u, base);
  const StencilWeights<R> sw(u, base);
  Vec3<R> pvn = v3(R(0), R(0), R(0));
  Mat3<R> dvp = Mat3<R>::zero();
  for (int sidx = 0; sidx < 64; ++sidx) {
    int o[3]; getIJK(sidx, o);
    int gi[3] = {base[0] + o[0], base[1] + o[1], base[2] + o[2]};
    if (!(gi[0] >= 0 && gi[1] >= 0 && gi[2] >= 0 && gi[0] < Dx && gi[1] < Dy && gi[2] < Dz)) continue;
    R wip = sw.weight(o);                           // weighting3(gi - u)
    Vec3<R> gwip = sw.gradient(o) * s.gscale;       // weightingGradient3(u - gi)
    int64_t gI = gi[0] + int64_t(gi[1]) * Dx + int64_t(gi[2]) * Dx * Dy;
    Vec3<R> vin = v3(s.gvstar[3*gI], s.gvstar[3*gI+1], s.gvstar[3*gI+2]);
    Vec3<R> vi = v3(s.gmv[3*gI], s.gmv[3*gI+1], s.gmv[3*gI+2]);
    R mi = s.gm[gI];
    if (mi > R(0)) {
      Vec3<R> vpn = (vin * ((R(1) - alpha) * wip)) + ((vin - vi / mi) * (alpha * wip));
      pvn = pvn + vpn;
      for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) dvp.m[c][r] += vin[r] * gwip[c];
    }
  }
  // particleAdvance.glsl:95-140
  Mat3<R> FEp = s.FE(p), FPp = s.FP(p);
  Mat3<R> L = Mat3<R>::ident() + dt * dvp;
  Mat3<R> FEpn = L * FEp;
  Mat3<R> Fpn = L * (FEp * FPp);
  Mat3<R> W, S, V;
  computeSVD(FEpn, W, S, V);
  auto sclamp = [](R& f, R lb, R ub) { f = (f < lb) ? lb : (f > ub) ? ub : f; };
  sclamp(S.m[0][0], R(1) - s.crit_comp, R(1) + s.crit_stretch);
  sclamp(S.m[1][1], R(1) - s.crit_comp, R(1) + s.crit_stretch);
  sclamp(S.m[2][2], R(1) - s.crit_comp, R(1) + s.crit_stretch);
  FEpn = W * S * transpose(V);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) FEpn.m[i][j] = round_to(FEpn.m[i][j], R(1e9));
  Mat3<R> S_I = Mat3<R>::zero();
  S_I.m[0][0] = R(1) / S.m[0][0]; S_I.m[1][1] = R(1) / S.m[1][1]; S_I.m[2][2] = R(1) / S.m[2][2];
  Mat3<R> FPpn = V * S_I * transpose(W) * Fpn;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) FPpn.m[i][j] = round_to(FPpn.m[i][j], R(1e9));
  std::memcpy(s.fe + 9*p, FEpn.m, sizeof(FEpn.m));
  std::memcpy(s.fp + 9*p, FPpn.m, sizeof(FPpn.m));
  // particleAdvance.glsl:173-210
  Vec3<R> vp = v3(s.v[3*p], s.v[3*p+1], s.v[3*p+2]);
  Vec3<R> vpn1 = vp * alpha + pvn;
  vpn1 = collide_velocity(xp, vpn1, s.col, s.nc);  // pre-advection position (Q11)
  s.v[3*p] = vpn1.x; s.v[3*p+1] = vpn1.y; s.v[3*p+2] = vpn1.z;
  s.x[3*p] += dt * vpn1.x; s.x[3*p+1] += dt * vpn1.y; s.x[3*p+2] += dt * vpn1.z;
}

template <class R> void g2p_advance(Scene<R>& s, R dt) {
  parallel_for(s.nthreads, s.n, [&](int, int64_t b, int64_t e) {
    for (int64_t p = b; p < e; ++p) g2p_advance_particle(s, p, dt);
  });
}

// ExplicitTimeUpdate::update (explicittimeupdate.cpp:36-63)
template <class R> void step(Scene<R>& s, R dt) {
  rigid_advance(s, dt);
  p2g(s, /*mass_only=*/false);  // gReset + p2gTransfer
  grid_update(s, dt);
  g2p_advance(s, dt);
}

// ---- Semi-implicit velocity update ----------------------------------------------------------
// The reference declares SemiImplicitTimeUpdate (src/snow/simulation/timeUpdate/
// semiimplicittimeupdate.hpp:6-13) with empty init/update bodies (semiimplicittimeupdate.cpp:3-5).
// This restates the published method the reference's explicit step follows: Stomakhin et al.
// 2013, "A material point method for snow simulation" (SIGGRAPH), eq. (9) and the Hessian action
// of its implicit-update section, for the energy whose gradient is the explicit force above
// (fixed corotated with hardening, the same grad w with its grad_scale):
//   K v = M v - beta dt df(dt v) = M v*,     df_i(du) = -sum_p V_p0 A_p F_Ep^T grad w_ip,
//   A_p = d2Psi/dF2 (F_Ep) : dF_p,           dF_p = sum_j du_j grad w_jp^T F_Ep,
//   d2Psi/dF2 : dF = 2 mu (dF - dR) + lambda (cof F : dF) cof F + lambda (J - 1) d(cof F),
// with mu, lambda hardened by J_P as in the explicit force and V_p0 = m_p / rho_p0 (0 where the
// explicit force skips the particle: rho0 <= 0 or |m| > 100, p2gTransfer.glsl:64-66/126).
// K is symmetric, so it is solved by conjugate residuals (the paper's solver) from v*, preconditioned
// by the lumped mass.
// This oracle takes the exact polar factor (scaled Newton iteration) and dR = R [omega]x with
// (tr S I - S) omega = axial(R^T dF - dF^T R), S = R^T F; the device takes the factors of the
// SVD the explicit force uses.

// cof F = J F^-T: column k is the cross product of the other two columns (cyclic order).
template <class R> inline Mat3<R> cofactor(const Mat3<R>& F) {
  auto cr = [](const R* a, const R* b, R* o) {
    o[0] = a[1] * b[2] - a[2] * b[1]; o[1] = a[2] * b[0] - a[0] * b[2]; o[2] = a[0] * b[1] - a[1] * b[0];
  };
  Mat3<R> C;
  cr(F.m[1], F.m[2], C.m[0]);
  cr(F.m[2], F.m[0], C.m[1]);
  cr(F.m[0], F.m[1], C.m[2]);
  return C;
}
// d(cof F)[dF]: each cross product differentiated by the product rule.
template <class R> inline Mat3<R> cofactor_diff(const Mat3<R>& F, const Mat3<R>& dF) {
  auto cr = [](const R* a, const R* b, R* o) {
    o[0] += a[1] * b[2] - a[2] * b[1]; o[1] += a[2] * b[0] - a[0] * b[2]; o[2] += a[0] * b[1] - a[1] * b[0];
  };
  Mat3<R> C = Mat3<R>::zero();
  const int nx[3][2] = {{1, 2}, {2, 0}, {0, 1}};
  for (int k = 0; k < 3; ++k) {
    cr(dF.m[nx[k][0]], F.m[nx[k][1]], C.m[k]);
    cr(F.m[nx[k][0]], dF.m[nx[k][1]], C.m[k]);
  }
  return C;
}
template <class R> inline R ddot(const Mat3<R>& A, const Mat3<R>& B) {
  R s = 0;
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) s += A.m[c][r] * B.m[c][r];
  return s;
}
template <class R> inline Mat3<R> inverse(const Mat3<R>& A) { return transpose(cofactor(A)) * (R(1) / determinant(A)); }

// Rotation factor of F = R S (det F > 0) by the scaled Newton iteration R <- (g R + R^-T / g) / 2.
template <class R> inline Mat3<R> polar_newton(const Mat3<R>& F) {
  Mat3<R> X = F;
  for (int it = 0; it < 100; ++it) {
    const Mat3<R> Xit = transpose(inverse(X));
    const R g = std::sqrt(std::sqrt(ddot(Xit, Xit) / ddot(X, X)));
    const Mat3<R> Y = (X * g + Xit * (R(1) / g)) * R(0.5);
    const Mat3<R> D = Y - X;
    X = Y;
    if (ddot(D, D) < R(1e-30) * ddot(X, X)) break;
  }
  return X;
}

// First Piola-Kirchhoff stress of the energy: P = 2 mu (F - R) + lambda (J - 1) cof F. The explicit
// force matrix is V0 P F^T (p2g_force_matrix, with the reference's rounded R).
template <class R> inline Mat3<R> fc_stress(const Mat3<R>& F, R mu_, R la) {
  const Mat3<R> Rot = polar_newton(F);
  return R(2) * mu_ * (F - Rot) + la * (determinant(F) - R(1)) * cofactor(F);
}

// d2Psi/dF2 (F) : dF.
template <class R> inline Mat3<R> fc_hessian_action(const Mat3<R>& F, R mu_, R la, const Mat3<R>& dF) {
  const Mat3<R> Rot = polar_newton(F);
  const Mat3<R> S = transpose(Rot) * F;
  const Mat3<R> Ss = (S + transpose(S)) * R(0.5);
  const R tr = Ss.m[0][0] + Ss.m[1][1] + Ss.m[2][2];
  const Mat3<R> T = Mat3<R>::ident(tr) - Ss;
  const Mat3<R> D = transpose(Rot) * dF;
  // axial vector of the skew D - D^T (m[col][row]): a0 = K(2,1), a1 = K(0,2), a2 = K(1,0)
  const Vec3<R> a = v3(D.m[1][2] - D.m[2][1], D.m[2][0] - D.m[0][2], D.m[0][1] - D.m[1][0]);
  const Vec3<R> w = inverse(T) * a;
  Mat3<R> Om = Mat3<R>::zero();  // [w]x: (1,0) = w.z, (0,2) = w.y, (2,1) = w.x, skew
  Om.m[0][1] = w.z; Om.m[1][0] = -w.z;
  Om.m[2][0] = w.y; Om.m[0][2] = -w.y;
  Om.m[1][2] = w.x; Om.m[2][1] = -w.x;
  const Mat3<R> dR = Rot * Om;
  const Mat3<R> C = cofactor(F);
  const R J = determinant(F);
  return R(2) * mu_ * (dF - dR) + la * ddot(C, dF) * C + la * (J - R(1)) * cofactor_diff(F, dF);
}

// mu, lambda (hardened) and V0 of particle p as the explicit force takes them.
template <class R> inline void implicit_coefficients(const Scene<R>& s, int64_t p, R& mu_, R& la, R& V0) {
  const R JP = determinant(s.FP(p));
  mu_ = mu(s.young, s.poisson, s.hardening, JP);
  la = lambda(s.young, s.poisson, s.hardening, JP);
  const R mp = s.m[p], r0 = s.rho0[p];
  V0 = (r0 > R(0) && std::fabs(mp) <= R(100)) ? mp / r0 : R(0);
}

// sum_j du_j grad w_jp^T over the in-grid stencil nodes (grad w with grad_scale).
template <class R> inline Mat3<R> implicit_gather(const Scene<R>& s, int64_t p, const R* du) {
  const int Dx = s.D[0], Dy = s.D[1], Dz = s.D[2];
  Vec3<R> u; int base[3];
  particle_in_grid(s, s.xp(p), u, base);
  const StencilWeights<R> sw(u, base);
  Mat3<R> L = Mat3<R>::zero();
  for (int sidx = 0; sidx < 64; ++sidx) {
    int o[3]; getIJK(sidx, o);
    int gi[3] = {base[0] + o[0], base[1] + o[1], base[2] + o[2]};
    if (!(gi[0] >= 0 && gi[1] >= 0 && gi[2] >= 0 && gi[0] < Dx && gi[1] < Dy && gi[2] < Dz)) continue;
    const Vec3<R> g = sw.gradient(o) * s.gscale;
    const int64_t gI = gi[0] + int64_t(gi[1]) * Dx + int64_t(gi[2]) * Dx * Dy;
    for (int c = 0; c < 3; ++c)
      for (int r = 0; r < 3; ++r) L.m[c][r] += du[3 * gI + r] * g[c];
  }
  return L;
}

// out_i -= sum_p G_p grad w_ip over the nodes with z in [k0, k1), in the bucket order of p2g.
template <class R> void implicit_scatter(const Scene<R>& s, const std::vector<Mat3<R>>& G,
                                         const std::vector<std::vector<int64_t>>& zb, int k0, int k1, R* out) {
  const int Dx = s.D[0], Dy = s.D[1], Dz = s.D[2];
  for (int bz = k0 - 2; bz <= k1 + 1; ++bz) {
    if (bz < -2 || bz > Dz) continue;
    for (int64_t p : zb[bz + 2]) {
      Vec3<R> u; int base[3];
      particle_in_grid(s, s.xp(p), u, base);
      const StencilWeights<R> sw(u, base);
      for (int sidx = 0; sidx < 64; ++sidx) {
        int o[3]; getIJK(sidx, o);
        int gi[3] = {base[0] + o[0], base[1] + o[1], base[2] + o[2]};
        if (!(gi[0] >= 0 && gi[1] >= 0 && gi[2] >= 0 && gi[0] < Dx && gi[1] < Dy && gi[2] < Dz)) continue;
        if (gi[2] < k0 || gi[2] >= k1) continue;
        const Vec3<R> f = G[p] * (sw.gradient(o) * s.gscale);
        const int64_t gI = gi[0] + int64_t(gi[1]) * Dx + int64_t(gi[2]) * Dx * Dy;
        out[3 * gI] -= f.x; out[3 * gI + 1] -= f.y; out[3 * gI + 2] -= f.z;
      }
    }
  }
}

template <class R> void implicit_scatter_all(const Scene<R>& s, const std::vector<Mat3<R>>& G, R* out) {
  std::fill(out, out + 3 * s.nodes(), R(0));
  auto zb = z_buckets(s);
  const int T = std::max(1, std::min(s.nthreads, s.D[2]));
  parallel_for(T, s.D[2], [&](int, int64_t kb, int64_t ke) { implicit_scatter(s, G, zb, int(kb), int(ke), out); });
}

// df(du) at every node (the Hessian action with the sign of a force, no dt).
template <class R> void implicit_df(const Scene<R>& s, const R* du, R* out) {
  std::vector<Mat3<R>> G(s.n);
  parallel_for(s.nthreads, s.n, [&](int, int64_t b, int64_t e) {
    for (int64_t p = b; p < e; ++p) {
      R mu_, la, V0;
      implicit_coefficients(s, p, mu_, la, V0);
      if (V0 == R(0)) { G[p] = Mat3<R>::zero(); continue; }
      const Mat3<R> F = s.FE(p);
      const Mat3<R> dF = implicit_gather(s, p, du) * F;
      G[p] = V0 * (fc_hessian_action(F, mu_, la, dF) * transpose(F));
    }
  });
  implicit_scatter_all(s, G, out);
}

// The node force of the energy at displaced nodes x_j + du_j (F_p -> (I + sum_j du_j grad w^T) F_p),
// f_i = -sum_p V0 P(F_p') F_p^T grad w_ip; implicit_df is its derivative at du = 0.
template <class R> void implicit_force(const Scene<R>& s, const R* du, R* out) {
  std::vector<Mat3<R>> G(s.n);
  parallel_for(s.nthreads, s.n, [&](int, int64_t b, int64_t e) {
    for (int64_t p = b; p < e; ++p) {
      R mu_, la, V0;
      implicit_coefficients(s, p, mu_, la, V0);
      if (V0 == R(0)) { G[p] = Mat3<R>::zero(); continue; }
      const Mat3<R> F = s.FE(p);
      const Mat3<R> Fd = (Mat3<R>::ident() + implicit_gather(s, p, du)) * F;
      G[p] = V0 * (fc_stress(Fd, mu_, la) * transpose(F));
    }
  });
  implicit_scatter_all(s, G, out);
}

// Kv = M v - beta dt df(dt v) on the nodes with m > 0 (0 elsewhere); v is read only there.
template <class R> void implicit_apply(const Scene<R>& s, R dt, R beta, const R* v, R* out) {
  const int64_t nn = s.nodes();
  std::vector<R> du(3 * nn, R(0)), df(3 * nn);
  for (int64_t i = 0; i < nn; ++i)
    if (s.gm[i] > R(0))
      for (int k = 0; k < 3; ++k) du[3 * i + k] = dt * v[3 * i + k];
  implicit_df(s, du.data(), df.data());
  for (int64_t i = 0; i < nn; ++i)
    for (int k = 0; k < 3; ++k)
      out[3 * i + k] = s.gm[i] > R(0) ? s.gm[i] * v[3 * i + k] - beta * dt * df[3 * i + k] : R(0);
}

// Conjugate residuals for K v = M v*, preconditioned by the lumped mass (z = M^-1 r), from v = v*,
// on the nodes with m > 0; v* (s.gvstar) is replaced by the solution. Stops when the residual
// |r| <= tol |M v*| or after max_iter updates. The node masses span orders of magnitude (light
// nodes at the surface of the material); without the preconditioner the residual criterion is met
// before those nodes move from v*.
template <class R> void implicit_solve(Scene<R>& s, R dt, R beta, R tol, int max_iter, int* iters, R* relres) {
  const int64_t n3 = 3 * s.nodes();
  std::vector<R> x(s.gvstar, s.gvstar + n3), b(n3), r(n3), z(n3), p(n3), Kp(n3), Kz(n3), minv(n3);
  for (int64_t i = 0; i < s.nodes(); ++i)
    for (int k = 0; k < 3; ++k) {
      const bool live = s.gm[i] > R(0);
      x[3 * i + k] = live ? x[3 * i + k] : R(0);
      b[3 * i + k] = live ? s.gm[i] * x[3 * i + k] : R(0);
      minv[3 * i + k] = live ? R(1) / s.gm[i] : R(0);
    }
  auto dotv = [&](const std::vector<R>& a, const std::vector<R>& c) {
    R acc = 0;
    for (int64_t i = 0; i < n3; ++i) acc += a[i] * c[i];
    return acc;
  };
  implicit_apply(s, dt, beta, x.data(), Kz.data());
  for (int64_t i = 0; i < n3; ++i) {
    r[i] = b[i] - Kz[i];
    z[i] = minv[i] * r[i];
  }
  const R bb = dotv(b, b);
  R rr = dotv(r, r);
  int it = 0;
  R rho = 0;
  std::fill(p.begin(), p.end(), R(0));
  std::fill(Kp.begin(), Kp.end(), R(0));
  while (rr > tol * tol * bb && it < max_iter) {
    if (it > 0 && !(rho != R(0))) break;  // z . K z vanished (K is not definite): no next direction
    implicit_apply(s, dt, beta, z.data(), Kz.data());
    const R rho_new = dotv(z, Kz);
    const R bc = it == 0 ? R(0) : rho_new / rho;
    rho = rho_new;
    R pp = 0;  // Kp . M^-1 Kp
    for (int64_t i = 0; i < n3; ++i) {
      p[i] = z[i] + bc * p[i];
      Kp[i] = Kz[i] + bc * Kp[i];
      pp += Kp[i] * (minv[i] * Kp[i]);
    }
    if (!(pp > R(0))) break;
    const R alpha = rho / pp;
    for (int64_t i = 0; i < n3; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * Kp[i];
      z[i] -= alpha * (minv[i] * Kp[i]);
    }
    rr = dotv(r, r);
    ++it;
  }
  std::copy(x.begin(), x.end(), s.gvstar);
  if (iters) *iters = it;
  if (relres) *relres = bb > R(0) ? std::sqrt(rr / bb) : R(0);
}

// One semi-implicit step: the explicit step with the grid velocities v* replaced by the solution
// of K v = M v* before the transfer back to the particles (the paper's steps 1-6, 7, 8-10).
template <class R> void step_semi_implicit(Scene<R>& s, R dt, R beta, R tol, int max_iter, int* iters, R* relres) {
  rigid_advance(s, dt);
  p2g(s, /*mass_only=*/false);
  grid_update(s, dt);
  implicit_solve(s, dt, beta, tol, max_iter, iters, relres);
  g2p_advance(s, dt);
}

}  // namespace oracle
