// gr_move.hpp -- one particle's explicit time step on the quaternion state,
// FP64: RigidBody::computeMomentum (RigidBody.cpp:267-307), FirstOrderExplicit
// ::Move (FirstOrderExplicit.cpp:46-60), TimeIntegrator::computeQuaternionChange
// (TimeIntegrator.cpp:28-44) and Transform3::updateTransform
// (Transform3.cpp:323-330) as q <- dq x q.  __host__ __device__: the
// integration kernel and the adjoint's host check call the same code.
#pragma once

#include "gr_math.hpp"

namespace gr {

// x, q, v, w: state in, next state out; F, M: total force and torque;
// I, Ii: inertia and its inverse, upper triangle xx xy xz yy yz zz
GR_HD void move_particle(double mass, const double* I, const double* Ii, double dt, V3<double> F,
                         V3<double> M, V3<double>& x, Qt<double>& q, V3<double>& v, V3<double>& w) {
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
  const V3<double> trans = dt * vel;
  v = vel + dt * accel;
  Qt<double> dq = {0.0, 0.0, 0.0, 1.0};
  const double nO = norm(om);
  if (nO > Eps<double>::HIGH) {
    const double c = cos(nO * dt / 2.0);
    const double sn = sin(nO * dt / 2.0);
    dq = qmake((sn / nO) * om, c);
  }
  w = om + dt * alpha;
  x = x + trans;
  q = qmul(dq, qi);
}

}  // namespace gr
