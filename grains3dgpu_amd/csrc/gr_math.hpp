// gr_math.hpp -- small vector / matrix / quaternion types for the DEM kernels.
//
// Operation order follows the reference's header-only operators so that the
// FP64 path rounds like the reference CPU code (built without FMA
// contraction): dot products start from 0 + x*x' (VectorMath.hh operator*),
// matrix rows sum left to right (MatrixMath.hh:128-175), quaternion products
// as QuaternionMath.hh and Quaternion.cpp:241-270.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define GR_HD __host__ __device__ __forceinline__

namespace gr {

template <class T>
struct V3 {
  T x, y, z;
};
template <class T>
struct M3 {
  T m[9];  // row major: XX XY XZ YX YY YZ ZX ZY ZZ
};
template <class T>
struct Qt {
  T x, y, z, w;
};
template <class T>
struct Tr {
  M3<T> R;
  V3<T> o;
};

template <class T> GR_HD V3<T> v3(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> GR_HD V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> GR_HD V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> GR_HD V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> GR_HD V3<T> operator*(T d, V3<T> a) { return {d * a.x, d * a.y, d * a.z}; }
template <class T> GR_HD V3<T> operator/(V3<T> a, T d) { return {a.x / d, a.y / d, a.z / d}; }
template <class T> GR_HD T dot(V3<T> a, V3<T> b) {
  T o = T(0);
  o += a.x * b.x;
  o += a.y * b.y;
  o += a.z * b.z;
  return o;
}
template <class T> GR_HD V3<T> cross(V3<T> a, V3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> GR_HD T norm2(V3<T> a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
template <class T> GR_HD T norm(V3<T> a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
// Vector3::round (Vector3.cpp:121-126)
template <class T> GR_HD V3<T> round_tol(V3<T> a, T tol) {
  return {fabs(a.x) < tol ? T(0) : a.x, fabs(a.y) < tol ? T(0) : a.y, fabs(a.z) < tol ? T(0) : a.z};
}
template <class T> GR_HD bool same(V3<T> a, V3<T> b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

// m * v: row . v
template <class T> GR_HD V3<T> mv(const M3<T>& m, V3<T> v) {
  const T* b = m.m;
  return {b[0] * v.x + b[1] * v.y + b[2] * v.z, b[3] * v.x + b[4] * v.y + b[5] * v.z,
          b[6] * v.x + b[7] * v.y + b[8] * v.z};
}
// v * m: m^T v
template <class T> GR_HD V3<T> vm(V3<T> v, const M3<T>& m) {
  const T* b = m.m;
  return {b[0] * v.x + b[3] * v.y + b[6] * v.z, b[1] * v.x + b[4] * v.y + b[7] * v.z,
          b[2] * v.x + b[5] * v.y + b[8] * v.z};
}
template <class T> GR_HD M3<T> transpose(const M3<T>& a) {
  const T* b = a.m;
  return M3<T>{{b[0], b[3], b[6], b[1], b[4], b[7], b[2], b[5], b[8]}};
}
template <class T> GR_HD M3<T> mmul(const M3<T>& A, const M3<T>& B) {
  const T* a = A.m;
  const T* b = B.m;
  M3<T> r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      r.m[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
  return r;
}
// Transform3::operator() (Transform3.cpp:334-337)
template <class T> GR_HD V3<T> apply(const Tr<T>& t, V3<T> v) { return mv(t.R, v) + t.o; }
// b2a = Transform3(a2w, b2w) (Transform3.cpp:34-39, 313-318)
template <class T> GR_HD Tr<T> relative(const Tr<T>& a2w, const Tr<T>& b2w) {
  Tr<T> r;
  M3<T> inv = transpose(a2w.R);
  r.R = mmul(inv, b2w.R);
  r.o = mv(inv, b2w.o - a2w.o);
  return r;
}

// ---- quaternions (x, y, z, w) ------------------------------------------------
template <class T> GR_HD V3<T> qv(Qt<T> q) { return {q.x, q.y, q.z}; }
template <class T> GR_HD Qt<T> qmake(V3<T> v, T w) { return {v.x, v.y, v.z, w}; }
template <class T> GR_HD Qt<T> conj(Qt<T> q) { return {-q.x, -q.y, -q.z, q.w}; }
// [0, v] x q
template <class T> GR_HD Qt<T> vq(V3<T> v, Qt<T> q) {
  T tmp = dot(-v, qv(q));
  V3<T> t = cross(v, qv(q)) + q.w * v;
  return qmake(t, tmp);
}
// vector part of a x q
template <class T> GR_HD V3<T> mult_to_v3(Qt<T> a, Qt<T> q) {
  return cross(qv(a), qv(q)) + a.w * qv(q) + q.w * qv(a);
}
// a x b
template <class T> GR_HD Qt<T> qmul(Qt<T> a, Qt<T> b) {
  T tmp = (a.w * b.w) - dot(qv(a), qv(b));
  V3<T> t = cross(qv(a), qv(b)) + a.w * qv(b) + b.w * qv(a);
  return qmake(t, tmp);
}
// rotation matrix (body -> world) of a unit quaternion
template <class T> GR_HD M3<T> rot(Qt<T> q) {
  T x = q.x, y = q.y, z = q.z, w = q.w;
  return M3<T>{{T(1) - T(2) * (y * y + z * z), T(2) * (x * y - w * z), T(2) * (x * z + w * y),
                T(2) * (x * y + w * z), T(1) - T(2) * (x * x + z * z), T(2) * (y * z - w * x),
                T(2) * (x * z - w * y), T(2) * (y * z + w * x), T(1) - T(2) * (x * x + y * y)}};
}

template <class T, class U> GR_HD V3<T> cvt(V3<U> a) { return {T(a.x), T(a.y), T(a.z)}; }
template <class T, class U> GR_HD M3<T> cvt(const M3<U>& a) {
  M3<T> r;
#pragma unroll
  for (int i = 0; i < 9; ++i) r.m[i] = T(a.m[i]);
  return r;
}

// constants (Grains/Base/include/Basic.hh:114-128)
template <class T> struct Eps;
template <> struct Eps<double> {
  static constexpr double HIGH = 1.e-15, MID = 1.e-10, LOW = 1.e-05;
};
template <> struct Eps<float> {
  static constexpr float HIGH = 1.e-08f, MID = 1.e-05f, LOW = 1.e-03f;
};

}  // namespace gr
