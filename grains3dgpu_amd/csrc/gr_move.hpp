// gr_move.hpp -- one particle's explicit time step on the quaternion state,
// FP64: RigidBody::computeMomentum (RigidBody.cpp:267-307), FirstOrderExplicit
// ::Move (FirstOrderExplicit.cpp:46-60) or a second-order scheme,
// TimeIntegrator::computeQuaternionChange
// (TimeIntegrator.cpp:28-44) and Transform3::updateTransform
// (Transform3.cpp:323-330) as q <- dq x q.  __host__ __device__: the
// integration kernel and the adjoint's host check call the same code.
#pragma once

#include "../../include/grains_hip.h"
#include "gr_math.hpp"

namespace gr {

// x, q, v, w: state in, next state out; F, M: total force and torque;
// I, Ii: inertia and its inverse, upper triangle xx xy xz yy yz zz.
// scheme: GRAINS_FIRST_ORDER_EXPLICIT (FirstOrderExplicit::Move) or one of
// the second-order schemes of grains_hip.h.  hist: ADAMS_BASHFORTH, in, the
// previous step's v, w, a, alpha (hist[2].x NaN: no history, this step is
// first order), out, this step's; LEAPFROG, hist[0].x NaN: the first step,
// a half kick (set to 0 after it).
GR_HD void move_particle_scheme(int scheme, double mass, const double* I, const double* Ii, double dt,
                                V3<double> F, V3<double> M, V3<double>& x, Qt<double>& q, V3<double>& v,
                                V3<double>& w, V3<double>* hist) {
  const Qt<double> qi = q;
  const V3<double> om = w, vel = v;
  const Qt<double> qc = conj(qi);
  const V3<double> av = mult_to_v3(qc, vq(om, qi));
  V3<double> am = mult_to_v3(qc, vq(M, qi));
  V3<double> tmp = {I[0] * av.x + I[1] * av.y + I[2] * av.z, I[1] * av.x + I[3] * av.y + I[4] * av.z,
                    I[2] * av.x + I[4] * av.y + I[5] * av.z};
  am = am + cross(tmp, av);
  tmp = {Ii[0] * am.x + Ii[1] * am.y + Ii[2] * am.z, Ii[1] * am.x + Ii[3] * am.y + Ii[4] * am.z,
         Ii[2] * am.x + Ii[4] * am.y + Ii[5] * am.z};
  const V3<double> alpha = mult_to_v3(qi, vq(tmp, qc));
  const V3<double> accel = F / mass;
  V3<double> trans, wr;  // translation over the step, angular velocity of the rotation
  if (scheme == GRAINS_SECOND_ORDER_EXPLICIT) {
    trans = dt * vel + (0.5 * dt * dt) * accel;
    wr = om + (0.5 * dt) * alpha;
    v = vel + dt * accel;
    w = om + dt * alpha;
  } else if (scheme == GRAINS_SECOND_ORDER_LEAPFROG) {
    const double kick = hist[0].x == hist[0].x ? dt : 0.5 * dt;
    v = vel + kick * accel;
    w = om + kick * alpha;
    trans = dt * v;
    wr = w;
    hist[0].x = 0.0;
  } else if (scheme == GRAINS_SECOND_ORDER_ADAMS_BASHFORTH && hist[2].x == hist[2].x) {
    trans = dt * (1.5 * vel - 0.5 * hist[0]);
    wr = 1.5 * om - 0.5 * hist[1];
    v = vel + dt * (1.5 * accel - 0.5 * hist[2]);
    w = om + dt * (1.5 * alpha - 0.5 * hist[3]);
  } else {
    trans = dt * vel;
    wr = om;
    v = vel + dt * accel;
    w = om + dt * alpha;
  }
  if (scheme == GRAINS_SECOND_ORDER_ADAMS_BASHFORTH) {
    hist[0] = vel;
    hist[1] = om;
    hist[2] = accel;
    hist[3] = alpha;
  }
  Qt<double> dq = {0.0, 0.0, 0.0, 1.0};
  const double nO = norm(wr);
  if (nO > Eps<double>::HIGH) {
    const double c = cos(nO * dt / 2.0);
    const double sn = sin(nO * dt / 2.0);
    dq = qmake((sn / nO) * wr, c);
  }
  x = x + trans;
  q = qmul(dq, qi);
}

// FirstOrderExplicit::Move, the reference's scheme
GR_HD void move_particle(double mass, const double* I, const double* Ii, double dt, V3<double> F,
                         V3<double> M, V3<double>& x, Qt<double>& q, V3<double>& v, V3<double>& w) {
  move_particle_scheme(GRAINS_FIRST_ORDER_EXPLICIT, mass, I, Ii, dt, F, M, x, q, v, w, nullptr);
}

}  // namespace gr
