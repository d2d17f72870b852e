// mpm_oracle.cpp - C entry points of the CPU oracle (Real=double) and of the CPU
// baseline (Real=float, std::thread pool). TEST INFRASTRUCTURE ONLY: loaded by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker, never by the
// product. Built by oracle/Makefile into oracle/liboracle.so.
#include "mpm_oracle.hpp"

extern "C" {

#define ORACLE_SCENE(T, NAME)                                                             \
  typedef struct NAME {                                                                   \
    int64_t n;                                                                            \
    T *x, *v, *m, *fe, *fp, *rho0;                                                        \
    snow_grid grid;                                                                       \
    snow_material material;                                                               \
    int32_t num_colliders;                                                                \
    T* colliders; /* [nc][12]: pos, type, vel, radius, normal, friction */                \
    T *gm, *gmv, *gf, *gvstar; /* dense [D^3], [D^3][3], [D^3][3], [D^3][3] */            \
    int32_t nthreads;                                                                     \
  } NAME;

ORACLE_SCENE(double, oracle_scene64)
ORACLE_SCENE(float, oracle_scene32)
}

namespace {

template <class R, class C>
oracle::Scene<R> wrap(const C* c) {
  oracle::Scene<R> s;
  s.n = c->n;
  s.x = c->x; s.v = c->v; s.m = c->m; s.fe = c->fe; s.fp = c->fp; s.rho0 = c->rho0;
  for (int k = 0; k < 3; ++k) { s.D[k] = c->grid.dim[k]; s.origin[k] = R(c->grid.origin[k]); }
  s.h = R(c->grid.spacing);
  const snow_material& mt = c->material;
  s.young = R(mt.young); s.poisson = R(mt.poisson); s.hardening = R(mt.hardening);
  s.crit_comp = R(mt.crit_comp); s.crit_stretch = R(mt.crit_stretch); s.alpha = R(mt.flip_alpha);
  s.g = oracle::v3(R(mt.gravity[0]), R(mt.gravity[1]), R(mt.gravity[2]));
  s.gscale = mt.grad_scale > 0.0f ? R(mt.grad_scale) : R(1) / s.h;
  s.nc = c->num_colliders;
  static_assert(sizeof(oracle::Collider<R>) == 12 * sizeof(R), "collider layout");
  s.col = reinterpret_cast<oracle::Collider<R>*>(c->colliders);
  s.gm = c->gm; s.gmv = c->gmv; s.gf = c->gf; s.gvstar = c->gvstar;
  s.nthreads = c->nthreads > 0 ? c->nthreads : 1;
  return s;
}

template <class C>
int check(const C* c) {
  if (!c || c->n < 0) return -1;
  for (int k = 0; k < 3; ++k)
    if (c->grid.dim[k] <= 0) return -1;
  if (!(c->grid.spacing > 0.0f)) return -1;
  if (c->n > 0 && (!c->x || !c->v || !c->m || !c->fe || !c->fp || !c->rho0)) return -1;
  if (!c->gm || !c->gmv || !c->gf || !c->gvstar) return -1;
  if (c->num_colliders > 0 && !c->colliders) return -1;
  return 0;
}

}  // namespace

#define ORACLE_API(T, SC, P)                                                               \
  extern "C" int P##_init(SC* c) {                                                         \
    if (check(c)) return -1;                                                               \
    auto s = wrap<T>(c);                                                                   \
    oracle::init(s);                                                                       \
    return 0;                                                                              \
  }                                                                                        \
  extern "C" int P##_grid_mass(SC* c) {                                                    \
    if (check(c)) return -1;                                                               \
    auto s = wrap<T>(c);                                                                   \
    oracle::p2g(s, true);                                                                  \
    return 0;                                                                              \
  }                                                                                        \
  extern "C" int P##_particle_volume(SC* c) {                                              \
    if (check(c)) return -1;                                                               \
    auto s = wrap<T>(c);                                                                   \
    oracle::particle_volume(s);                                                            \
    return 0;                                                                              \
  }                                                                                        \
  extern "C" int P##_rigid_advance(SC* c, double dt) {                                     \
    if (check(c)) return -1;                                                               \
    auto s = wrap<T>(c);                                                                   \
    oracle::rigid_advance(s, T(dt));                                                       \
    return 0;                                                                              \
  }                                                                                        \
  extern "C" int P##_p2g(SC* c) {                                                          \
    if (check(c)) return -1;                                                               \
    auto s = wrap<T>(c);                                                                   \
    oracle::p2g(s, false);                                                                 \
    return 0;                                                                              \
  }                                                                                        \
  extern "C" int P##_grid_update(SC* c, double dt) {                                       \
    if (check(c)) return -1;                                                               \
    auto s = wrap<T>(c);                                                                   \
    oracle::grid_update(s, T(dt));                                                         \
    return 0;                                                                              \
  }                                                                                        \
  extern "C" int P##_g2p_advance(SC* c, double dt) {                                       \
    if (check(c)) return -1;                                                               \
    auto s = wrap<T>(c);                                                                   \
    oracle::g2p_advance(s, T(dt));                                                         \
    return 0;                                                                              \
  }                                                                                        \
  extern "C" int P##_step(SC* c, double dt, int32_t nsteps) {                              \
    if (check(c)) return -1;                                                               \
    auto s = wrap<T>(c);                                                                   \
    for (int32_t i = 0; i < nsteps; ++i) oracle::step(s, T(dt));                           \
    return 0;                                                                              \
  }                                                                                        \
  extern "C" T P##_weighting(T x) { return oracle::weighting(x); }                         \
  extern "C" T P##_weighting_gradient(T x) { return oracle::weightingGradient(x); }        \
  extern "C" T P##_mu(T E, T nu, T xi, T JP) { return oracle::mu(E, nu, xi, JP); }          \
  extern "C" T P##_lambda(T E, T nu, T xi, T JP) { return oracle::lambda(E, nu, xi, JP); }  \
  extern "C" void P##_svd(const T* A, T* W, T* S, T* V) {                                  \
    oracle::Mat3<T> a, w, sm, vm;                                                          \
    std::memcpy(a.m, A, sizeof(a.m));                                                      \
    oracle::computeSVD(a, w, sm, vm);                                                      \
    std::memcpy(W, w.m, sizeof(w.m));                                                      \
    std::memcpy(S, sm.m, sizeof(sm.m));                                                    \
    std::memcpy(V, vm.m, sizeof(vm.m));                                                    \
  }                                                                                        \
  extern "C" void P##_polar(const T* A, T* R) {                                            \
    oracle::Mat3<T> a, r;                                                                  \
    std::memcpy(a.m, A, sizeof(a.m));                                                      \
    oracle::computePD(a, r);                                                               \
    std::memcpy(R, r.m, sizeof(r.m));                                                      \
  }                                                                                        \
  extern "C" void P##_getijk(int32_t s, int32_t* out) {                                    \
    int o[3];                                                                              \
    oracle::getIJK(s, o);                                                                  \
    out[0] = o[0]; out[1] = o[1]; out[2] = o[2];                                           \
  }                                                                                        \
  extern "C" void P##_collide(const T* x, const T* v, const T* colliders, int32_t nc,      \
                              T* out) {                                                    \
    auto cs = reinterpret_cast<const oracle::Collider<T>*>(colliders);                     \
    auto r = oracle::collide_velocity(oracle::v3(x[0], x[1], x[2]),                        \
                                      oracle::v3(v[0], v[1], v[2]), cs, nc);               \
    out[0] = r.x; out[1] = r.y; out[2] = r.z;                                              \
  }

ORACLE_API(double, oracle_scene64, oracle64)
ORACLE_API(float, oracle_scene32, oracle32)

// ---- semi-implicit update (FP64 only) ---------------------------------------------------------
namespace {
oracle::Mat3<double> mat_in(const double* a) { oracle::Mat3<double> m; std::memcpy(m.m, a, sizeof(m.m)); return m; }
}
extern "C" void oracle64_fc_stress(const double* F, double mu, double la, double* P) {
  const auto r = oracle::fc_stress(mat_in(F), mu, la);
  std::memcpy(P, r.m, sizeof(r.m));
}
extern "C" void oracle64_fc_hessian(const double* F, double mu, double la, const double* dF, double* A) {
  const auto r = oracle::fc_hessian_action(mat_in(F), mu, la, mat_in(dF));
  std::memcpy(A, r.m, sizeof(r.m));
}
extern "C" void oracle64_polar_exact(const double* F, double* R) {
  const auto r = oracle::polar_newton(mat_in(F));
  std::memcpy(R, r.m, sizeof(r.m));
}
extern "C" int oracle64_implicit_df(oracle_scene64* c, const double* du, double* out) {
  if (check(c) || !du || !out) return -1;
  auto s = wrap<double>(c);
  oracle::implicit_df(s, du, out);
  return 0;
}
extern "C" int oracle64_implicit_force(oracle_scene64* c, const double* du, double* out) {
  if (check(c) || !du || !out) return -1;
  auto s = wrap<double>(c);
  oracle::implicit_force(s, du, out);
  return 0;
}
extern "C" int oracle64_implicit_apply(oracle_scene64* c, double dt, double beta, const double* v, double* out) {
  if (check(c) || !v || !out) return -1;
  auto s = wrap<double>(c);
  oracle::implicit_apply(s, dt, beta, v, out);
  return 0;
}
extern "C" int oracle64_implicit_solve(oracle_scene64* c, double dt, double beta, double tol, int32_t max_iter,
                                       int32_t* iters, double* relres) {
  if (check(c)) return -1;
  auto s = wrap<double>(c);
  int it = 0;
  oracle::implicit_solve(s, dt, beta, tol, int(max_iter), &it, relres);
  if (iters) *iters = it;
  return 0;
}
extern "C" int oracle64_step_semi_implicit(oracle_scene64* c, double dt, double beta, double tol, int32_t max_iter,
                                           int32_t nsteps, int32_t* iters, double* relres) {
  if (check(c)) return -1;
  auto s = wrap<double>(c);
  for (int32_t i = 0; i < nsteps; ++i) {
    int it = 0;
    oracle::step_semi_implicit(s, dt, beta, tol, int(max_iter), &it, relres);
    if (iters) *iters = it;
  }
  return 0;
}
