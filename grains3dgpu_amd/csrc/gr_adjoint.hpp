// gr_adjoint.hpp -- reverse-mode derivatives (vector-Jacobian products) of
// one DEM time step in FP64, physics semantics, for a fixed contact set:
// contact geometry, Hooke force and explicit integration.  Every function is
// __host__ __device__: the adjoint kernels (gr_adjoint.hip) and a host check
// harness (tests/cpp/adjoint_check.cc) call the same code.
//
// Shapes with a smooth support map: spheres and ellipsoids (superquadrics
// with n1 = n2 = 2); walls are Rectangles.
//
// Contact geometry between two separated crust cores is differentiated
// implicitly.  With n the unit normal from A to B, the closest points are
// pa = S_A(n) and pb = S_B(-n), S(u) = x + R s(R^T u) the world support map,
// and w = pb - pa = d n.  Perturbing the poses by dtheta, the condition
// P w = 0 (P = I - n n^T) gives
//     (d P + J_A + J_B + n n^T) dn = P dw_theta,      dd = n . dw_theta,
// with J = R J_loc R^T the (symmetric, n-annihilating) Jacobian of the support
// map and dw_theta the change of S_B(-n) - S_A(n) at fixed n (the distance's
// derivative is the envelope theorem).  The deep-core branch (centre line),
// sphere-sphere contacts and walls are closed forms.
#pragma once

#include "../../include/grains_hip.h"
#include "gr_math.hpp"

namespace gr {
namespace adj {

using V = V3<double>;
using M = M3<double>;
using Q = Qt<double>;

constexpr double kMid = 1.e-10;   // EPS of Vector3::round (Basic.hh:114-128)
constexpr double kHigh = 1.e-15;  // HIGHEPS

// ---- small linear algebra ------------------------------------------------------
GR_HD V vzero() { return {0.0, 0.0, 0.0}; }
GR_HD M mzero() { return M{{0, 0, 0, 0, 0, 0, 0, 0, 0}}; }
GR_HD M outer(V a, V b) {
  return M{{a.x * b.x, a.x * b.y, a.x * b.z, a.y * b.x, a.y * b.y, a.y * b.z, a.z * b.x, a.z * b.y, a.z * b.z}};
}
GR_HD M madd(const M& a, const M& b) {
  M r;
  for (int k = 0; k < 9; ++k) r.m[k] = a.m[k] + b.m[k];
  return r;
}
GR_HD M mscale(double s, const M& a) {
  M r;
  for (int k = 0; k < 9; ++k) r.m[k] = s * a.m[k];
  return r;
}
GR_HD M eye() { return M{{1, 0, 0, 0, 1, 0, 0, 0, 1}}; }
// R J R^T
GR_HD M conj_rot(const M& R, const M& J) { return mmul(mmul(R, J), transpose(R)); }
// solve A y = b (A invertible 3x3) by the adjugate
GR_HD V solve3(const M& A, V b) {
  const double* a = A.m;
  const double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8], c02 = a[3] * a[7] - a[4] * a[6];
  const double c10 = a[2] * a[7] - a[1] * a[8], c11 = a[0] * a[8] - a[2] * a[6], c12 = a[1] * a[6] - a[0] * a[7];
  const double c20 = a[1] * a[5] - a[2] * a[4], c21 = a[2] * a[3] - a[0] * a[5], c22 = a[0] * a[4] - a[1] * a[3];
  const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
  // inverse = adj / det, adj = cofactor^T
  return V{(c00 * b.x + c10 * b.y + c20 * b.z) / det, (c01 * b.x + c11 * b.y + c21 * b.z) / det,
           (c02 * b.x + c12 * b.y + c22 * b.z) / det};
}
// Vector3::round mask: components kept (|c| >= tol)
GR_HD V keep_mask(V a, double tol) {
  return {fabs(a.x) < tol ? 0.0 : 1.0, fabs(a.y) < tol ? 0.0 : 1.0, fabs(a.z) < tol ? 0.0 : 1.0};
}
GR_HD V hadamard(V a, V b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }

// ---- quaternions as 4-vectors ----------------------------------------------------
GR_HD Q pure(V v) { return {v.x, v.y, v.z, 0.0}; }
GR_HD V vec(Q q) { return {q.x, q.y, q.z}; }
GR_HD Q qadd(Q a, Q b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
GR_HD Q qscale(double s, Q a) { return {s * a.x, s * a.y, s * a.z, s * a.w}; }

// R = rot(q) (gr_math.hpp): qbar += dR/dq^T Rbar
GR_HD Q rot_vjp(Q q, const M& Rb) {
  const double x = q.x, y = q.y, z = q.z, w = q.w;
  const double* b = Rb.m;
  Q g;
  g.x = 2 * y * b[1] + 2 * z * b[2] + 2 * y * b[3] - 4 * x * b[4] - 2 * w * b[5] + 2 * z * b[6] + 2 * w * b[7] -
        4 * x * b[8];
  g.y = -4 * y * b[0] + 2 * x * b[1] + 2 * w * b[2] + 2 * x * b[3] + 2 * z * b[5] - 2 * w * b[6] + 2 * z * b[7] -
        4 * y * b[8];
  g.z = -4 * z * b[0] - 2 * w * b[1] + 2 * x * b[2] + 2 * w * b[3] - 4 * z * b[4] + 2 * y * b[5] + 2 * x * b[6] +
        2 * y * b[7];
  g.w = -2 * z * b[1] + 2 * y * b[2] + 2 * z * b[3] - 2 * x * b[5] - 2 * y * b[6] + 2 * x * b[7];
  return g;
}

// ---- support maps with their Jacobian (body frame) ---------------------------
struct SupJ {
  V s;  // support point
  M J;  // d s / d u (symmetric, J u = 0)
};
// kind GRAINS_SPHERE: dims[0] = radius; GRAINS_SUPERQUADRIC with n1 = n2 = 2:
// the ellipsoid of semi-axes dims[0..2], s = A^2 u / |A u|
GR_HD SupJ sup_jac(int kind, const double* dims, V u) {
  SupJ r;
  if (kind == GRAINS_SPHERE) {
    const double nu = norm(u);
    const V uh = u / nu;
    r.s = dims[0] * uh;
    r.J = mscale(dims[0] / nu, madd(eye(), mscale(-1.0, outer(uh, uh))));
  } else {
    const double a2 = dims[0] * dims[0], b2 = dims[1] * dims[1], c2 = dims[2] * dims[2];
    const V A2u = {a2 * u.x, b2 * u.y, c2 * u.z};
    const double h = sqrt(dot(u, A2u));
    r.s = A2u / h;
    M D = M{{a2, 0, 0, 0, b2, 0, 0, 0, c2}};
    r.J = mscale(1.0 / h, madd(D, mscale(-1.0, outer(r.s, r.s))));
  }
  return r;
}

// world pose of a body
struct Pose {
  V x;
  M R;
};
struct PoseBar {
  V x;
  M R;
};
GR_HD PoseBar posebar_zero() { return {vzero(), mzero()}; }

// y = x + R s(R^T u) at fixed u: accumulate x_bar, R_bar for y_bar
GR_HD void sup_world_vjp(const M& R, V sloc, const M& Jloc, V u, V yb, PoseBar& pb) {
  pb.x = pb.x + yb;
  pb.R = madd(pb.R, madd(outer(yb, sloc), outer(u, mv(Jloc, vm(yb, R)))));
}

// ---- contact geometry ----------------------------------------------------------
// Output of the contact dispatch as seen by the force: the overlap vector cv
// (Hooke's `pen`) and the contact point cp.
struct Geo {
  V cv, cp;
};

// Separated crust cores (GJK branch of closestPointsRigidBodies with physics
// semantics): cv = (d - ctSum) round(n), cp = (pa + pb) / 2, n and d from the
// converged search.  Given cv_bar, cp_bar, accumulate the pose adjoints.
GR_HD void gjk_geo_vjp(int kindA, const double* coreA, const Pose& A, int kindB, const double* coreB,
                       const Pose& B, V n, double d, double ctSum, V cvb, V cpb, PoseBar& Ab, PoseBar& Bb) {
  const V uA = vm(n, A.R), uB = vm(-n, B.R);     // R^T n, R^T (-n)
  const SupJ sA = sup_jac(kindA, coreA, uA), sB = sup_jac(kindB, coreB, uB);
  const M JA = conj_rot(A.R, sA.J), JB = conj_rot(B.R, sB.J);
  const V m = keep_mask(n, kMid);
  const double db = dot(cvb, hadamard(m, n));
  const V nb = (d - ctSum) * hadamard(m, cvb);
  const V pab = 0.5 * cpb, pbb = 0.5 * cpb;
  const V nbt = (nb + mv(JA, pab)) - mv(JB, pbb);
  const M P = madd(eye(), mscale(-1.0, outer(n, n)));
  const M H = madd(madd(mscale(d, P), madd(JA, JB)), outer(n, n));
  const V h = solve3(H, nbt);
  const V gb = mv(P, h) + db * n;
  sup_world_vjp(A.R, sA.s, sA.J, n, pab - gb, Ab);
  sup_world_vjp(B.R, sB.s, sB.J, -n, pbb + gb, Bb);
}

// Overlapping crust cores: the centre line, cv = ctSum round(u),
// u = (xA - xB) / |xA - xB|, cp = (xA + xB) / 2.
GR_HD void deep_geo_vjp(V xA, V xB, double ctSum, V cvb, V cpb, V& xAb, V& xBb) {
  const V e = xA - xB;
  const double L = norm(e);
  const V u = e / L;
  const V ub = ctSum * hadamard(keep_mask(u, kMid), cvb);
  const V eb = (ub - dot(ub, u) * u) / L;
  xAb = xAb + eb + 0.5 * cpb;
  xBb = xBb - eb + 0.5 * cpb;
}

// closestPointsSpheres (CollisionDetection.cpp:27-48): overlap = |xB - xA| -
// rA - rB, u = (xB - xA) / |.|, cp = xA + (rA + overlap / 2) u, cv = overlap u.
GR_HD void sphere_geo_vjp(V xA, V xB, double rA, double rB, V cvb, V cpb, V& xAb, V& xBb) {
  const V e = xB - xA;
  const double L = norm(e);
  const V u = e / L;
  const double ov = L - rA - rB;
  const V ub = (rA + 0.5 * ov) * cpb + ov * cvb;
  const double ovb = 0.5 * dot(cpb, u) + dot(cvb, u);
  const V eb = ovb * u + (ub - dot(ub, u) * u) / L;
  xAb = xAb + cpb - eb;
  xBb = xBb + eb;
}

// closestPointsRectangle (CollisionDetection.cpp:54-90) with A penetrating
// the wall plane (point rPt, unit normal rNorm facing A): pointA = S_A(-rNorm)
// of A's full shape, t = (rPt - pointA) . rNorm, cv = t rNorm,
// cp = pointA + t/2 rNorm.
GR_HD void wall_geo_vjp(int kindA, const double* dimsA, const Pose& A, V rNorm, V cvb, V cpb, PoseBar& Ab) {
  const V u = -rNorm;
  const SupJ sA = sup_jac(kindA, dimsA, vm(u, A.R));
  const double tb = dot(cvb, rNorm) + 0.5 * dot(cpb, rNorm);
  const V pAb = cpb - tb * rNorm;
  sup_world_vjp(A.R, sA.s, sA.J, u, pAb, Ab);
}

// ---- Hooke (HookeContactForceModel.cpp:91-186), physics semantics ----------------
struct HookeIn {
  double kn, muen, etat, muc;  // muen = ln e / sqrt(pi^2 + ln^2 e) (HookeContactForceModel.cpp:56)
};
struct HookeBar {
  V cv, cp, vA, wA, vB, wB, xA, xB;
  double kn, muen, etat, muc;
};
GR_HD HookeBar hookebar_zero() {
  return {vzero(), vzero(), vzero(), vzero(), vzero(), vzero(), vzero(), vzero(), 0.0, 0.0, 0.0, 0.0};
}
// d muen / d e
GR_HD double dmuen_den(double en) {
  const double L = log(en), pi2 = 9.869604401089358;
  return pi2 / (pow(pi2 + L * L, 1.5) * en);
}

// Force F on A and torque M = (cp - xA) x F of one contact, and the adjoint of
// (cv, cp, vA, wA, vB, wB, xA, xB, parameters) for (F_bar, M_bar).  The
// relative velocity uses the levers cp - xA and cp - xB; m2 = 0 (a wall) takes
// the reference's avmass branch.  Rolling resistance (kr) is not covered.
GR_HD void hooke_vjp(const HookeIn& h, V cv, V cp, V vA, V wA, V vB, V wB, V xA, V xB, double m1, double m2,
                     V Fb_in, V Mb, HookeBar& o, V* F_out = nullptr) {
  // forward, in the order of the force model
  const V lA = cp - xA, lB = cp - xB;
  const V rel = (vA + cross(wA, lA)) - (vB + cross(wB, lB));
  const double L = norm(cv);
  const V ph = cv / L;
  const V mk = keep_mask(ph, kMid);
  const V p = hadamard(mk, ph);
  const double vns = dot(rel, p);
  const V vn = vns * p;
  const V vt = rel - vn;
  const double nt = norm(vt);
  const bool tan_on = nt > kMid;
  const V tng = tan_on ? vt / nt : vzero();
  double av = m1 * m2 / (m1 + m2);
  double om0 = sqrt(h.kn / av);
  if (av == 0.0) {
    av = m2 == 0.0 ? 0.5 * m1 : 0.5 * m2;
    om0 = 2.0 * sqrt(h.kn / av);
  }
  const double c = 2.0 * om0 * h.muen * av;  // delFN = kn cv + c v_n
  const V dFN = h.kn * cv + c * vn;
  const double nFN = norm(dFN);
  const V dFTv = (-h.etat * 2.0 * av) * vt;
  const double fn = h.muc * nFN, ft = norm(dFTv);
  const bool coul = fn < ft;
  const V F = dFN + (coul ? (-fn) * tng : dFTv);
  if (F_out) *F_out = F;
  // backward
  const V Fb = Fb_in + cross(Mb, lA);
  V lAb = cross(F, Mb);
  V vtb = vzero();
  double nFNb = 0.0;
  if (coul) {
    const double fnb = -dot(Fb, tng);
    const V tb = (-fn) * Fb;
    nFNb += h.muc * fnb;
    o.muc += nFN * fnb;
    if (tan_on) vtb = vtb + (tb - dot(tb, tng) * tng) / nt;
  } else {
    vtb = vtb + (-h.etat * 2.0 * av) * Fb;
    o.etat += -2.0 * av * dot(vt, Fb);
  }
  V dFNb = Fb;
  if (nFN > 0.0) dFNb = dFNb + (nFNb / nFN) * dFN;
  V cvb = h.kn * dFNb;
  o.kn += dot(cv, dFNb) + (c / (2.0 * h.kn)) * dot(vn, dFNb);
  o.muen += 2.0 * om0 * av * dot(vn, dFNb);
  const V vnb = c * dFNb - vtb;  // v_t = rel - v_n
  const V relb = vtb + dot(p, vnb) * p;
  const V pb = vns * vnb + dot(p, vnb) * rel;
  const V qb = hadamard(mk, pb);
  cvb = cvb + (qb - dot(qb, ph) * ph) / L;
  o.cv = o.cv + cvb;
  o.vA = o.vA + relb;
  o.vB = o.vB - relb;
  o.wA = o.wA + cross(lA, relb);
  lAb = lAb + cross(relb, wA);
  o.wB = o.wB - cross(lB, relb);
  const V lBb = -cross(relb, wB);
  o.cp = o.cp + lAb + lBb;
  o.xA = o.xA - lAb;
  o.xB = o.xB - lBb;
}

// ---- explicit integration (the moveParticles kernel) ------------------------------
struct Body {
  double mass;
  const double* I;   // inertia, upper triangle xx xy xz yy yz zz
  const double* Ii;  // inverse inertia, same layout
};
GR_HD V symv(const double* S, V a) {
  return {S[0] * a.x + S[1] * a.y + S[2] * a.z, S[1] * a.x + S[3] * a.y + S[4] * a.z,
          S[2] * a.x + S[4] * a.y + S[5] * a.z};
}
struct StateBar {
  V x;
  Q q;
  V v, w;
};

// Adjoint of (x, q, v, w, F, M) -> (x', q', v', w') for the next state's
// adjoint sb1; returns the state adjoint through the integrator alone and the
// force and torque adjoints.  Follows RigidBody::computeMomentum on the
// quaternion state (RigidBody.cpp:267-307), FirstOrderExplicit::Move
// (FirstOrderExplicit.cpp:46-60), computeQuaternionChange
// (TimeIntegrator.cpp:28-44) and q' = dq x q.  The sandwich products
// vec(q* v q) and vec(q a q*) are differentiated as quaternion products
// (adjoint of left multiplication by c is left multiplication by c*):
//   d<A, q* v q> = <-2 v q A, dq> + <q A q*, dv>,
//   d<A, q a q*> = <-2 A q a, dq> + <q* A q, da>.
GR_HD void integrate_vjp(const Body& b, double dt, V x, Q q, V v, V w, V Mt, const StateBar& sb1, StateBar& sb,
                         V& Fb, V& Mb) {
  (void)x;
  const Q qc = conj(q);
  const V av = mult_to_v3(qc, vq(w, q));
  const V am0 = mult_to_v3(qc, vq(Mt, q));
  const V Iav = symv(b.I, av);
  const V am = am0 + cross(Iav, av);
  const V ab = symv(b.Ii, am);
  const double nO = norm(w);
  Q dq = {0.0, 0.0, 0.0, 1.0};
  double th = 0.0, sn = 0.0;
  if (nO > kHigh) {
    th = nO * dt / 2.0;
    sn = sin(th);
    dq = qmake((sn / nO) * w, cos(th));
  }
  // x' = x + dt v, v' = v + dt F / m, w' = w + dt alpha, q' = dq q
  sb.x = sb1.x;
  sb.v = sb1.v + dt * sb1.x;
  Fb = (dt / b.mass) * sb1.v;
  sb.w = sb1.w;
  const V alb = dt * sb1.w;
  sb.q = qmul(conj(dq), sb1.q);
  const Q dqb = qmul(sb1.q, qc);
  if (nO > kHigh) {
    const double s = sn / nO;
    const double ds = (cos(th) * (dt / 2.0) * nO - sn) / (nO * nO);
    const V dqvb = vec(dqb);
    sb.w = sb.w + s * dqvb + (dot(w, dqvb) * ds / nO) * w - (dqb.w * sn * (dt / 2.0) / nO) * w;
  }
  // alpha = vec(q ab q*)
  const Q A = pure(alb);
  sb.q = qadd(sb.q, qscale(-2.0, qmul(qmul(A, q), pure(ab))));
  const V abb = vec(qmul(qmul(qc, A), q));
  const V amb = symv(b.Ii, abb);
  // am = am0 + Iav x av
  const V Iavb = cross(av, amb);
  const V avb = cross(amb, Iav) + symv(b.I, Iavb);
  // am0 = vec(q* M q)
  sb.q = qadd(sb.q, qscale(-2.0, qmul(qmul(pure(Mt), q), pure(amb))));
  Mb = vec(qmul(qmul(q, pure(amb)), qc));
  // av = vec(q* w q)
  sb.q = qadd(sb.q, qscale(-2.0, qmul(qmul(pure(w), q), pure(avb))));
  sb.w = sb.w + vec(qmul(qmul(q, pure(avb)), qc));
}

}  // namespace adj
}  // namespace gr
