// adjoint_check.cc -- host-side C wrappers of the step's vector-Jacobian
// products (grains3dgpu_amd/csrc/gr_adjoint.hpp), the same code the adjoint
// kernels run, for tests/test_adjoint_cpu.py: each product is compared with
// central finite differences of the FP64 oracle's forward functions.
// Built with hipcc as a host library (no device code is launched).
#include "../../grains3dgpu_amd/csrc/gr_adjoint.hpp"
#include "../../grains3dgpu_amd/csrc/gr_move.hpp"

using namespace gr;
using namespace gr::adj;

namespace {
V v3p(const double* p) { return {p[0], p[1], p[2]}; }
void put(double* o, V v) {
  o[0] = v.x;
  o[1] = v.y;
  o[2] = v.z;
}
Pose pose12(const double* t) {
  Pose p;
  for (int k = 0; k < 9; ++k) p.R.m[k] = t[k];
  p.x = {t[9], t[10], t[11]};
  return p;
}
void put_posebar(double* o, const PoseBar& b) {  // 12 doubles: R (9), x (3)
  for (int k = 0; k < 9; ++k) o[k] = b.R.m[k];
  put(o + 9, b.x);
}
}  // namespace

extern "C" {

// the forward: R = rot(q) of the kernels (row major)
void adjc_rot(const double* q, double* R) {
  const M r = rot(Q{q[0], q[1], q[2], q[3]});
  for (int k = 0; k < 9; ++k) R[k] = r.m[k];
}

// q (x,y,z,w), Rb (9) -> qb (4)
void adjc_rot_vjp(const double* q, const double* Rb, double* qb) {
  M R;
  for (int k = 0; k < 9; ++k) R.m[k] = Rb[k];
  Q g = rot_vjp(Q{q[0], q[1], q[2], q[3]}, R);
  qb[0] = g.x;
  qb[1] = g.y;
  qb[2] = g.z;
  qb[3] = g.w;
}

// separated crust cores: poses as 12-double transforms, n and d of the cores,
// (cv_bar, cp_bar) -> adjoints of the two transforms (12 each)
void adjc_gjk_geo_vjp(int kindA, const double* coreA, const double* tA, int kindB, const double* coreB,
                      const double* tB, const double* n, double d, double ctSum, const double* cvb,
                      const double* cpb, double* tAb, double* tBb) {
  PoseBar a = posebar_zero(), b = posebar_zero();
  gjk_geo_vjp(kindA, coreA, pose12(tA), kindB, coreB, pose12(tB), v3p(n), d, ctSum, v3p(cvb), v3p(cpb), a, b);
  put_posebar(tAb, a);
  put_posebar(tBb, b);
}

void adjc_deep_geo_vjp(const double* xA, const double* xB, double ctSum, const double* cvb, const double* cpb,
                       double* xAb, double* xBb) {
  V a = vzero(), b = vzero();
  deep_geo_vjp(v3p(xA), v3p(xB), ctSum, v3p(cvb), v3p(cpb), a, b);
  put(xAb, a);
  put(xBb, b);
}

void adjc_sphere_geo_vjp(const double* xA, const double* xB, double rA, double rB, const double* cvb,
                         const double* cpb, double* xAb, double* xBb) {
  V a = vzero(), b = vzero();
  sphere_geo_vjp(v3p(xA), v3p(xB), rA, rB, v3p(cvb), v3p(cpb), a, b);
  put(xAb, a);
  put(xBb, b);
}

void adjc_wall_geo_vjp(int kindA, const double* dimsA, const double* tA, const double* rNorm, const double* cvb,
                       const double* cpb, double* tAb) {
  PoseBar a = posebar_zero();
  wall_geo_vjp(kindA, dimsA, pose12(tA), v3p(rNorm), v3p(cvb), v3p(cpb), a);
  put_posebar(tAb, a);
}

// prm = kn, muen, etat, muc; in = cv, cp, vA, wA, vB, wB, xA, xB (24);
// out = the same 24 adjoints + 4 parameter adjoints; F (3) the force on A
void adjc_hooke_vjp(const double* prm, const double* in, double m1, double m2, const double* Fb,
                    const double* Mb, double* out, double* F) {
  HookeIn h = {prm[0], prm[1], prm[2], prm[3]};
  HookeBar o = hookebar_zero();
  V f;
  hooke_vjp(h, v3p(in), v3p(in + 3), v3p(in + 6), v3p(in + 9), v3p(in + 12), v3p(in + 15), v3p(in + 18),
            v3p(in + 21), m1, m2, v3p(Fb), v3p(Mb), o, &f);
  const V vs[8] = {o.cv, o.cp, o.vA, o.wA, o.vB, o.wB, o.xA, o.xB};
  for (int k = 0; k < 8; ++k) put(out + 3 * k, vs[k]);
  out[24] = o.kn;
  out[25] = o.muen;
  out[26] = o.etat;
  out[27] = o.muc;
  put(F, f);
}

double adjc_dmuen_den(double en) { return dmuen_den(en); }

// the forward of one particle's step (the integration kernel's code):
// state (13) in, next state (13) out
void adjc_move(double mass, const double* I, const double* Ii, double dt, const double* state, const double* F,
               const double* Mt, double* out) {
  V x = v3p(state), v = v3p(state + 7), w = v3p(state + 10);
  Q q = {state[3], state[4], state[5], state[6]};
  move_particle(mass, I, Ii, dt, v3p(F), v3p(Mt), x, q, v, w);
  put(out, x);
  out[3] = q.x;
  out[4] = q.y;
  out[5] = q.z;
  out[6] = q.w;
  put(out + 7, v);
  put(out + 10, w);
}

// state = x (3), q (4), v (3), w (3); sb1 likewise; out sb (13), Fb (3), Mb (3)
void adjc_integrate_vjp(double mass, const double* I, const double* Ii, double dt, const double* state,
                        const double* Mt, const double* sb1, double* sb, double* Fb, double* Mb) {
  Body b = {mass, I, Ii};
  StateBar s1, s0;
  s1.x = v3p(sb1);
  s1.q = Q{sb1[3], sb1[4], sb1[5], sb1[6]};
  s1.v = v3p(sb1 + 7);
  s1.w = v3p(sb1 + 10);
  V fb, mb;
  integrate_vjp(b, dt, v3p(state), Q{state[3], state[4], state[5], state[6]}, v3p(state + 7), v3p(state + 10),
                v3p(Mt), s1, s0, fb, mb);
  put(sb, s0.x);
  sb[3] = s0.q.x;
  sb[4] = s0.q.y;
  sb[5] = s0.q.z;
  sb[6] = s0.q.w;
  put(sb + 7, s0.v);
  put(sb + 10, s0.w);
  put(Fb, fb);
  put(Mb, mb);
}

}  // extern "C"
