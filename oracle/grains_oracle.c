/*
 * grains_oracle.c -- FP64 CPU oracle of the DEM time step (TEST INFRASTRUCTURE).
 *
 * A plain-C restatement of the reference's CPU step, written from the
 * reference's behaviour, not copied.  Every routine cites the reference
 * file:line it restates (paths relative to the reference root).  Arithmetic
 * is written in the reference's operation order (dot products start from
 * 0 + x*x', matrix rows sum left to right, no FMA contraction: build with
 * -ffp-contract=off) so that integer outputs and sqrt/div-only paths match
 * bit for bit.  The reference's own CPU build cannot be compiled here (its
 * headers need CUDA, curand and xerces, none present), so this restatement is
 * pinned by the checks in tests/test_oracle_*.py:
 *
 *   cell counts / hash / neighbours  brute force over points, exact cell
 *                                    boundaries, the survey's 12^3 example
 *   bucket order                     numpy stable argsort
 *   candidate pairs                  O(N^2) brute force
 *   sphere contact                   closed form
 *   rectangle contact                closed form (plane projection)
 *   superquadric support             ellipsoid closed form (n1=n2=2),
 *                                    maximality over sampled surface points
 *   box / polyhedron support         brute-force argmax over vertices
 *   OBB                              separating-axis brute force over
 *                                    sampled directions (never rejects an
 *                                    overlapping pair)
 *   GJK distance                     sphere and box closed forms, symmetry,
 *                                    sampled-support lower bound
 *   Hooke force                      binary collision restitution = en
 *                                    (exact for the linear spring-dashpot),
 *                                    Coulomb cap
 *   integration                      free fall closed form of explicit Euler,
 *                                    free rotation about a principal axis
 *   mass properties                  sphere/box closed forms, superquadric
 *                                    volume and inertia by quadrature
 */
#include "grains_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- constants, Grains/Base/include/Basic.hh:103-128 (double instances) --- */
#define G_PI 3.1415926535897932385
#define HIGHEPS 1.e-15
#define EPS 1.e-10
#define LOWEPS 1.e-05
/* Candidate prefilter slack: pairs whose centre distance is within this
 * relative margin of the bounding-sphere sum also reach the narrow phase, so
 * the narrow phase's own (reference) tests decide every borderline pair. */
#define PREFILTER_SLACK 1.e-10

typedef struct { double x, y, z; } v3;
typedef struct { double m[9]; } m3; /* row major: XX XY XZ YX YY YZ ZX ZY ZZ */
typedef struct { double x, y, z, w; } qt;

/* ---- Vector3 operators, Grains/Base/include/VectorMath.hh ----------------- */
static inline v3 V(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vneg(v3 a) { return V(-a.x, -a.y, -a.z); }
static inline v3 vscale(double d, v3 a) { return V(d * a.x, d * a.y, d * a.z); }
static inline v3 vdiv(v3 a, double d) { return V(a.x / d, a.y / d, a.z / d); }
/* dot: out = 0; out += b1[i]*b2[i] (VectorMath.hh operator*) */
static inline double vdot(v3 a, v3 b) {
  double o = 0.0;
  o += a.x * b.x;
  o += a.y * b.y;
  o += a.z * b.z;
  return o;
}
static inline v3 vcross(v3 a, v3 b) {
  return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline double vnorm(v3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
static inline double vnorm2(v3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
/* Vector3::round (Vector3.cpp:121-126) */
static inline v3 vround(v3 a, double tol) {
  return V(fabs(a.x) < tol ? 0.0 : a.x, fabs(a.y) < tol ? 0.0 : a.y,
           fabs(a.z) < tol ? 0.0 : a.z);
}
static inline double vget(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

/* ---- Matrix3 operators, Grains/Base/include/MatrixMath.hh:32-175 ---------- */
static inline v3 mv(const m3* m, v3 v) { /* m * v: row . v */
  const double* b = m->m;
  return V(b[0] * v.x + b[1] * v.y + b[2] * v.z, b[3] * v.x + b[4] * v.y + b[5] * v.z,
           b[6] * v.x + b[7] * v.y + b[8] * v.z);
}
static inline v3 vm(v3 v, const m3* m) { /* v * m: column . v = m^T v */
  const double* b = m->m;
  return V(b[0] * v.x + b[3] * v.y + b[6] * v.z, b[1] * v.x + b[4] * v.y + b[7] * v.z,
           b[2] * v.x + b[5] * v.y + b[8] * v.z);
}
static inline m3 mtrans(const m3* a) {
  const double* b = a->m;
  m3 r = {{b[0], b[3], b[6], b[1], b[4], b[7], b[2], b[5], b[8]}};
  return r;
}
static inline m3 mmul(const m3* A, const m3* B) {
  const double* a = A->m;
  const double* b = B->m;
  m3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      r.m[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
  return r;
}

/* ---- Transform3 (Grains/Geometry/src/Transform3.cpp) --------------------- */
typedef struct { m3 R; v3 o; } tr3;
/* operator(): basis * v + origin (Transform3.cpp:334-337) */
static inline v3 tapply(const tr3* t, v3 v) { return vadd(mv(&t->R, v), t->o); }
/* b2a = Transform3(a2w, b2w) (Transform3.cpp:34-39, 313-318) */
static inline tr3 trelative(const tr3* a2w, const tr3* b2w) {
  tr3 r;
  m3 inv = mtrans(&a2w->R);
  r.R = mmul(&inv, &b2w->R);
  r.o = mv(&inv, vsub(b2w->o, a2w->o));
  return r;
}
static tr3 tr_from12(const double* t) {
  tr3 r;
  memcpy(r.R.m, t, 9 * sizeof(double));
  r.o = V(t[9], t[10], t[11]);
  return r;
}

/* ---- Quaternion (Quaternion.cpp, QuaternionMath.hh) ---------------------- */
static inline qt Q(v3 v, double w) { qt q = {v.x, v.y, v.z, w}; return q; }
static inline v3 qv(qt q) { return V(q.x, q.y, q.z); }
static inline qt qconj(qt q) { return Q(vneg(qv(q)), q.w); }
/* [0, v] x q  (QuaternionMath.hh, operator*(Vector3, Quaternion)) */
static inline qt vq(v3 v, qt q) {
  double tmp = vdot(vneg(v), qv(q));
  v3 vt = vadd(vcross(v, qv(q)), vscale(q.w, v));
  return Q(vt, tmp);
}
/* vector part of this x q (Quaternion.cpp:253-258) */
static inline v3 qmult_to_v3(qt a, qt q) {
  return vadd(vadd(vcross(qv(a), qv(q)), vscale(a.w, qv(q))), vscale(q.w, qv(a)));
}
/* q1 x q2 (QuaternionMath.hh operator*(Quaternion, Quaternion)) */
static inline qt qmul(qt a, qt b) {
  double tmp = (a.w * b.w) - vdot(qv(a), qv(b));
  v3 vt = vadd(vadd(vcross(qv(a), qv(b)), vscale(a.w, qv(b))), vscale(b.w, qv(a)));
  return Q(vt, tmp);
}
/* Quaternion::setQuaternion(Matrix3) (Quaternion.cpp:125-169) */
static qt quat_from_matrix(const m3* rot) {
  const double* b = rot->m;
  enum { XX, XY, XZ, YX, YY, YZ, ZX, ZY, ZZ };
  qt q = {0.0, 0.0, 0.0, 1.0};
  double den;
  if (b[YY] > -b[ZZ] && b[XX] > -b[YY] && b[XX] > -b[ZZ]) {
    den = pow(1.0 + b[XX] + b[YY] + b[ZZ], 0.5);
    q.w = 0.5 * den;
    q.x = 0.5 * (b[ZY] - b[YZ]) / den;
    q.y = 0.5 * (b[XZ] - b[ZX]) / den;
    q.z = 0.5 * (b[YX] - b[XY]) / den;
  } else if (b[YY] < -b[ZZ] && b[XX] > b[YY] && b[XX] > b[ZZ]) {
    den = pow(1.0 + b[XX] - b[YY] - b[ZZ], 0.5);
    q.w = 0.5 * (b[ZY] - b[YZ]) / den;
    q.x = 0.5 * den;
    q.y = 0.5 * (b[XY] + b[YX]) / den;
    q.z = 0.5 * (b[ZX] + b[XZ]) / den;
  } else if (b[YY] > b[ZZ] && b[XX] < b[YY] && b[XX] < -b[ZZ]) {
    den = pow(1.0 - b[XX] + b[YY] - b[ZZ], 0.5);
    q.w = 0.5 * (b[XZ] - b[ZX]) / den;
    q.x = 0.5 * (b[XY] + b[YX]) / den;
    q.y = 0.5 * den;
    q.z = 0.5 * (b[YZ] + b[ZY]) / den;
  } else if (b[YY] < b[ZZ] && b[XX] < -b[YY] && b[XX] < b[ZZ]) {
    den = pow(1.0 - b[XX] - b[YY] + b[ZZ], 0.5);
    q.w = 0.5 * (b[YX] - b[XY]) / den;
    q.x = 0.5 * (b[ZX] + b[XZ]) / den;
    q.y = 0.5 * (b[YZ] + b[ZY]) / den;
    q.z = 0.5 * den;
  }
  /* else: the reference prints a warning and leaves q unset; identity here */
  return q;
}
/* Transform3::composeLeftByRotation(Quaternion) (Transform3.cpp:229-246):
 * basis <- M(q) * basis, column by column. */
static void compose_left_by_rotation(m3* basis, qt q) {
  double* b = basis->m;
  double qx = q.x, qy = q.y, qz = q.z, qw = q.w;
  for (int i = 0; i < 3; ++i) {
    double px = qy * b[6 + i] - qz * b[3 + i] + qw * b[i];
    double py = qz * b[i] - qx * b[6 + i] + qw * b[3 + i];
    double pz = qx * b[3 + i] - qy * b[i] + qw * b[6 + i];
    double pw = -qx * b[i] - qy * b[3 + i] - qz * b[6 + i];
    b[i] = qy * pz - qz * py - pw * qx + qw * px;
    b[3 + i] = qz * px - qx * pz - pw * qy + qw * py;
    b[6 + i] = qx * py - qy * px - pw * qz + qw * pz;
  }
}
static m3 matrix_from_quat(qt q) {
  double x = q.x, y = q.y, z = q.z, w = q.w;
  m3 r = {{1.0 - 2.0 * (y * y + z * z), 2.0 * (x * y - w * z), 2.0 * (x * z + w * y),
           2.0 * (x * y + w * z), 1.0 - 2.0 * (x * x + z * z), 2.0 * (y * z - w * x),
           2.0 * (x * z - w * y), 2.0 * (y * z + w * x), 1.0 - 2.0 * (x * x + y * y)}};
  return r;
}

/* ---- shapes --------------------------------------------------------------- */
typedef struct {
  grains_shape_t s;
  double volume, mass, inertia[6], inv[6], rc, bbox[3], scaling[3];
  double rc_exact; /* true circumscribed radius (physics semantics) */
  double bbox_exact[3]; /* OBB half-extents containing the shape (physics semantics) */
  double gd[5];         /* dims the support map reads (CYLINDER r, h/2; CONE r, h/4, sin) */
  double core_dims[5];
  const double* verts; /* full shape vertices (polyhedron)          */
  double* core_verts;  /* vertices scaled to the crust core          */
  int nv;
} oshape;

/* std::beta as libstdc++ evaluates it (exp of log-gammas); used by the
 * superquadric volume and inertia (Grains/Base/include/MiscMath.hh:82-90). */
static double grains_beta(double x, double y) {
  return exp(lgamma(x) + lgamma(y) - lgamma(x + y));
}

/* Superquadric::computeCircumscribedRadius (Superquadric.cpp:122-155) */
static double sq_circumscribed(double a, double b, double c, double n1, double n2) {
  if ((n1 == 2.0) && (n2 == 2.0))
    return fmax(a, fmax(b, c));
  else if (n1 == 2.0) {
    double alpha = pow(b / a, 2.0 / (n2 - 2.0));
    double xt = 1.0 / pow(1.0 + pow(alpha, n2), 1.0 / n2);
    return fmax(c, sqrt(a * xt * a * xt + alpha * b * xt * alpha * b * xt));
  } else if (n2 == 2.0) {
    double m = fmax(a, b);
    double beta = pow(c * c / (m * m), 1.0 / (n1 - 2.0));
    double xt = 1.0 / pow(1.0 + pow(beta, n1), 1.0 / n1);
    return sqrt(m * xt * m * xt + beta * c * xt * beta * c * xt);
  } else {
    double alpha = pow(b / a, 2.0 / (n2 - 2.0));
    double gamma = pow(1.0 + pow(alpha, n2), n1 / n2 - 1.0);
    double beta = pow(gamma * c * c / (a * a), 1.0 / (n1 - 1.0));
    double xt = 1.0 / pow(pow(1.0 + pow(alpha, n2), n1 / n2) + pow(beta, n1), 1.0 / n1);
    return sqrt(a * xt * a * xt + alpha * b * xt * alpha * b * xt +
                beta * c * xt * beta * c * xt);
  }
}

/* Largest |p| on the superquadric (|x/a|^n2 + |y/b|^n2)^(n1/n2) + |z/c|^n1 = 1,
 * the true circumscribed radius: the maximum over the critical points of |p|^2
 * of every coordinate family (axes, the three coordinate planes, and the
 * interior point p = (a xt, alpha b xt, beta c xt) with
 * alpha = (b/a)^(2/(n2-2)), beta = ((c/a)^2 (1+alpha^n2)^(n1/n2-1))^(1/(n1-2))).
 * The reference's general branch (Superquadric.cpp:143-153) uses the exponent
 * 1/(n1-1) for beta, which can underestimate the radius; the physics
 * semantics use this exact value. */
static double superellipse_max(double a, double b, double n) {
  double r = fmax(a, b);
  if (n > 2.0) {
    double t = pow(b / a, n / (n - 2.0));
    double x = a * pow(1.0 + pow(t * a / b, n), -1.0 / n);
    double c = sqrt(x * x + t * x * t * x);
    if (isfinite(c)) r = fmax(r, c); /* overflow near n = 2: the axis value is the limit */
  }
  return r;
}
static double sq_circumscribed_exact(double a, double b, double c, double n1, double n2) {
  double r = fmax(superellipse_max(a, b, n2), fmax(superellipse_max(a, c, n1), superellipse_max(b, c, n1)));
  if (n1 > 2.0 && n2 > 2.0) {
    double alpha = pow(b / a, 2.0 / (n2 - 2.0));
    double T = 1.0 + pow(alpha, n2);
    double beta = pow(c * c / (a * a) * pow(T, n1 / n2 - 1.0), 1.0 / (n1 - 2.0));
    double xt = pow(pow(T, n1 / n2) + pow(beta, n1), -1.0 / n1);
    double x = a * xt, y = alpha * b * xt, z = beta * c * xt;
    double q = sqrt(x * x + y * y + z * z);
    if (isfinite(q)) r = fmax(r, q);
  }
  return r;
}

/* Polyhedron (not in the reference): vertices are the hull points.  Volume,
 * centroid and inertia of the convex hull: every supporting plane through
 * three of the points is a facet; its coplanar points are ordered by angle
 * and fan-triangulated, and each triangle closes a tetrahedron with the
 * origin (divergence theorem). */
static void poly_mass(const double* vtx, int nv, double* vol, double* cen,
                      double* I6 /* about centroid, unit density */) {
  double V0 = 0, cx = 0, cy = 0, cz = 0;
  double acc[6] = {0, 0, 0, 0, 0, 0}; /* xx yy zz xy xz yz second moments about origin */
  int on[64];
  double ang[64];
  for (int i = 0; i < nv; ++i)
    for (int j = i + 1; j < nv; ++j)
      for (int k = j + 1; k < nv; ++k) {
        v3 a = V(vtx[3 * i], vtx[3 * i + 1], vtx[3 * i + 2]);
        v3 b = V(vtx[3 * j], vtx[3 * j + 1], vtx[3 * j + 2]);
        v3 c = V(vtx[3 * k], vtx[3 * k + 1], vtx[3 * k + 2]);
        v3 n = vcross(vsub(b, a), vsub(c, a));
        double nn = vnorm(n);
        if (nn == 0.0) continue;
        int pos = 0, neg = 0, m = 0, first_ok = 1;
        for (int l = 0; l < nv; ++l) {
          double sd = vdot(n, vsub(V(vtx[3 * l], vtx[3 * l + 1], vtx[3 * l + 2]), a));
          if (sd > 1e-12 * nn) ++pos;
          else if (sd < -1e-12 * nn) ++neg;
          else {
            if (m < 64) on[m++] = l;
            /* the plane is handled by its three lowest-index points only */
            if (l < k && l != i && l != j) first_ok = 0;
          }
        }
        if ((pos && neg) || !first_ok) continue;
        v3 nrm = pos ? vneg(n) : n; /* outward */
        v3 ctr = V(0, 0, 0);
        for (int t = 0; t < m; ++t)
          ctr = vadd(ctr, V(vtx[3 * on[t]], vtx[3 * on[t] + 1], vtx[3 * on[t] + 2]));
        ctr = vscale(1.0 / m, ctr);
        v3 e1 = vsub(a, ctr);
        e1 = vscale(1.0 / vnorm(e1), e1);
        v3 e2 = vcross(vscale(1.0 / nn, nrm), e1);
        for (int t = 0; t < m; ++t) {
          v3 d = vsub(V(vtx[3 * on[t]], vtx[3 * on[t] + 1], vtx[3 * on[t] + 2]), ctr);
          ang[t] = atan2(vdot(d, e2), vdot(d, e1));
        }
        for (int t = 1; t < m; ++t) /* insertion sort by angle (counter-clockwise about nrm) */
          for (int u = t; u > 0 && ang[u] < ang[u - 1]; --u) {
            double ta = ang[u]; ang[u] = ang[u - 1]; ang[u - 1] = ta;
            int ti = on[u]; on[u] = on[u - 1]; on[u - 1] = ti;
          }
        for (int t = 1; t + 1 < m; ++t) {
          v3 p0 = V(vtx[3 * on[0]], vtx[3 * on[0] + 1], vtx[3 * on[0] + 2]);
          v3 p1 = V(vtx[3 * on[t]], vtx[3 * on[t] + 1], vtx[3 * on[t] + 2]);
          v3 p2 = V(vtx[3 * on[t + 1]], vtx[3 * on[t + 1] + 1], vtx[3 * on[t + 1] + 2]);
          /* signed tetra (0, p0, p1, p2) */
          double d = vdot(p0, vcross(p1, p2)) / 6.0;
          V0 += d;
          cx += d * (p0.x + p1.x + p2.x) / 4.0;
          cy += d * (p0.y + p1.y + p2.y) / 4.0;
          cz += d * (p0.z + p1.z + p2.z) / 4.0;
          /* second moments of a tetra with one vertex at the origin:
           * int x_i x_j = d/20 (sum_r P_ri P_rj + (sum_r P_ri)(sum_r P_rj)) */
          double P[3][3] = {{p0.x, p0.y, p0.z}, {p1.x, p1.y, p1.z}, {p2.x, p2.y, p2.z}};
          int ii[6] = {0, 1, 2, 0, 0, 1}, jj[6] = {0, 1, 2, 1, 2, 2};
          for (int q = 0; q < 6; ++q) {
            double s1 = 0, si = 0, sj = 0;
            for (int r = 0; r < 3; ++r) {
              s1 += P[r][ii[q]] * P[r][jj[q]];
              si += P[r][ii[q]];
              sj += P[r][jj[q]];
            }
            acc[q] += d / 20.0 * (s1 + si * sj);
          }
        }
      }
  *vol = V0;
  cen[0] = cx / V0;
  cen[1] = cy / V0;
  cen[2] = cz / V0;
  /* shift second moments to the centroid */
  double Sxx = acc[0] - V0 * cen[0] * cen[0], Syy = acc[1] - V0 * cen[1] * cen[1];
  double Szz = acc[2] - V0 * cen[2] * cen[2], Sxy = acc[3] - V0 * cen[0] * cen[1];
  double Sxz = acc[4] - V0 * cen[0] * cen[2], Syz = acc[5] - V0 * cen[1] * cen[2];
  /* inertia tensor layout of RigidBody: [xx, xy, xz, yy, yz, zz] */
  I6[0] = Syy + Szz;
  I6[1] = -Sxy;
  I6[2] = -Sxz;
  I6[3] = Sxx + Szz;
  I6[4] = -Syz;
  I6[5] = Sxx + Syy;
}

static void sym_inverse(const double* I, double* inv) {
  double a = I[0], b = I[1], c = I[2], d = I[3], e = I[4], f = I[5];
  double A = d * f - e * e, B = -(b * f - c * e), C = b * e - c * d;
  double det = a * A + b * B + c * C;
  inv[0] = A / det;
  inv[1] = B / det;
  inv[2] = C / det;
  inv[3] = (a * f - c * c) / det;
  inv[4] = -(a * e - b * c) / det;
  inv[5] = (a * d - b * b) / det;
}

/* Cone(r, h) members (Cone.cpp:9-14): r, h/4 and the sine of the half angle */
static void cone_dims(double* g, double r, double q) {
  g[0] = r;
  g[1] = q;
  g[2] = r / sqrt(r * r + 16.0 * q * q);
}

/* RigidBody ctor (RigidBody.cpp:19-59) with the Convex virtuals of each shape:
 * Sphere.cpp:72-105, Box.cpp:81-123, Superquadric.cpp:81-163, Rectangle.cpp:72-109,
 * Cylinder.cpp:9-12,72-111, Cone.cpp:9-14,76-122. */
static int shape_derive(oshape* o, const grains_shape_t* s, const double* verts) {
  memset(o, 0, sizeof(*o));
  o->s = *s;
  const double* d = s->dims;
  double I[6] = {0}, Iinv[6] = {0};
  switch (s->kind) {
  case GRAINS_SPHERE: {
    double r = d[0];
    o->volume = 4.0 * G_PI * r * r * r / 3.0;
    double r5 = r * r * r * r * r;
    I[5] = I[3] = I[0] = 8.0 * G_PI / 15.0 * r5;
    Iinv[5] = Iinv[3] = Iinv[0] = 1.0 / I[0];
    o->rc = r;
    o->bbox[0] = o->bbox[1] = o->bbox[2] = r;
    break;
  }
  case GRAINS_BOX: {
    double ex = d[0], ey = d[1], ez = d[2];
    o->volume = 8.0 * ex * ey * ez;
    I[0] = 8.0 * ex * ey * ez * (ey * ey + ez * ez) / 3.0;
    I[3] = 8.0 * ex * ey * ez * (ex * ex + ez * ez) / 3.0;
    I[5] = 8.0 * ex * ey * ez * (ey * ey + ex * ex) / 3.0;
    Iinv[0] = 1.0 / I[0];
    Iinv[3] = 1.0 / I[3];
    Iinv[5] = 1.0 / I[5];
    o->rc = vnorm(V(ex, ey, ez));
    o->bbox[0] = ex; o->bbox[1] = ey; o->bbox[2] = ez;
    break;
  }
  case GRAINS_SUPERQUADRIC: {
    double a = d[0], b = d[1], c = d[2], n1 = d[3], n2 = d[4];
    double C = 2.0 / 3.0 * a * b * c, e1 = 2.0 / n1, e2 = 2.0 / n2;
    o->volume = C * e1 * e2 * grains_beta(e1, 0.5 * e1) * grains_beta(0.5 * e2, 0.5 * e2);
    double Ci = 0.4 * a * b * c * e1 * e2;
    double p1 = grains_beta(1.5 * e2, 0.5 * e2) * grains_beta(2.0 * e1, 0.5 * e1);
    double p2 = c * c * grains_beta(0.5 * e2, 0.5 * e2) * grains_beta(1.5 * e1, e1);
    I[0] = Ci * (b * b * p1 + p2);
    I[3] = Ci * (a * a * p1 + p2);
    I[5] = Ci * (a * a + b * b) * p1;
    Iinv[0] = 1.0 / I[0];
    Iinv[3] = 1.0 / I[3];
    Iinv[5] = 1.0 / I[5];
    o->rc = sq_circumscribed(a, b, c, n1, n2);
    o->rc_exact = sq_circumscribed_exact(a, b, c, n1, n2);
    o->bbox[0] = a; o->bbox[1] = b; o->bbox[2] = c;
    break;
  }
  case GRAINS_CYLINDER: {
    /* dims: radius, height (Cylinder(r, h)); the support reads r, h/2 */
    double r = d[0], hh = d[1] / 2.0;
    o->volume = 2.0 * G_PI * hh * r * r;
    double r2 = r * r, c = 0.5 * G_PI * hh * r2;
    I[0] = I[5] = c * (4.0 * hh * hh / 3.0 + r2);
    I[3] = 2.0 * c * r2;
    Iinv[5] = Iinv[0] = 1.0 / I[0];
    Iinv[3] = 1.0 / I[3];
    o->rc = sqrt(r * r + hh * hh);
    o->bbox[0] = r; o->bbox[1] = hh; o->bbox[2] = r;
    break;
  }
  case GRAINS_CONE: {
    /* dims: base radius, height (Cone(r, h)); origin at the centroid, apex at
     * +3h/4 on y, base at -h/4; the support reads r, h/4, sin(half angle) */
    double r = d[0], q = d[1] / 4.0;
    o->volume = 4.0 * q * G_PI * r * r / 3.0;
    double r2 = r * r, c = 0.2 * G_PI * q * r2;
    I[0] = I[5] = c * (4.0 * q * q + r2);
    I[3] = 2.0 * c * r2;
    Iinv[5] = Iinv[0] = 1.0 / I[0];
    Iinv[3] = 1.0 / I[3];
    o->rc = fmax(sqrt(r * r + q * q), 3.0 * q);
    /* Cone.cpp:116-122: the reference box is centred on the centroid and
     * leaves out the apex region y in (h/2, 3h/4] */
    o->bbox[0] = r; o->bbox[1] = 2.0 * q; o->bbox[2] = r;
    break;
  }
  case GRAINS_RECTANGLE: {
    double LX = d[0], LY = d[1];
    o->volume = 4.0 * LX * LY;
    I[0] = 4.0 / 3.0 * LX * LY * LY * LY;
    I[3] = 4.0 / 3.0 * LX * LX * LX * LY;
    I[5] = 2.0 / 3.0 * LX * LY * (LX * LX + LY * LY);
    Iinv[0] = 1.0 / I[0];
    Iinv[3] = 1.0 / I[3];
    Iinv[5] = 1.0 / I[5];
    o->rc = sqrt(LX * LX + LY * LY);
    o->bbox[0] = LX; o->bbox[1] = LY; o->bbox[2] = 0.0;
    break;
  }
  case GRAINS_POLYHEDRON: {
    if (!verts || s->n_vertices < 4) return GRAINS_EINVAL;
    o->verts = verts + 3 * (size_t)s->vertex_offset;
    o->nv = s->n_vertices;
    double cen[3];
    poly_mass(o->verts, o->nv, &o->volume, cen, I);
    sym_inverse(I, Iinv);
    double rc = 0;
    for (int k = 0; k < 3; ++k) o->bbox[k] = 0;
    for (int i = 0; i < o->nv; ++i) {
      v3 p = V(o->verts[3 * i], o->verts[3 * i + 1], o->verts[3 * i + 2]);
      rc = fmax(rc, vnorm(p));
      o->bbox[0] = fmax(o->bbox[0], fabs(p.x));
      o->bbox[1] = fmax(o->bbox[1], fabs(p.y));
      o->bbox[2] = fmax(o->bbox[2], fabs(p.z));
    }
    o->rc = rc;
    break;
  }
  default:
    return GRAINS_EUNSUPPORTED;
  }
  if (s->kind != GRAINS_SUPERQUADRIC) o->rc_exact = o->rc;
  for (int k = 0; k < 3; ++k) o->bbox_exact[k] = o->bbox[k];
  if (s->kind == GRAINS_CONE) o->bbox_exact[1] = 3.0 * d[1] / 4.0;
  memcpy(o->gd, s->dims, sizeof(o->gd));
  if (s->kind == GRAINS_CYLINDER) o->gd[1] = d[1] / 2.0;
  if (s->kind == GRAINS_CONE) cone_dims(o->gd, d[0], d[1] / 4.0);
  o->mass = s->density * o->volume;
  if (s->density == 0) {
    for (int i = 0; i < 6; ++i) o->inertia[i] = o->inv[i] = 0;
  } else {
    for (int i = 0; i < 6; ++i) {
      o->inertia[i] = I[i] * s->density;
      o->inv[i] = Iinv[i] / s->density;
    }
  }
  /* m_scaling (RigidBody.cpp:27-31) and the crust core it defines */
  for (int k = 0; k < 3; ++k)
    o->scaling[k] = o->bbox[k] > 0 ? (o->bbox[k] - s->crust) / o->bbox[k] : 1.0;
  memcpy(o->core_dims, o->gd, sizeof(o->core_dims));
  switch (s->kind) {
  case GRAINS_SPHERE: o->core_dims[0] = d[0] - s->crust; break;
  case GRAINS_CYLINDER: /* Cylinder(r - ct, h - 2 ct): the box scaling of m_scaling */
    o->core_dims[0] = o->gd[0] - s->crust;
    o->core_dims[1] = o->gd[1] - s->crust;
    break;
  case GRAINS_CONE: /* Cone(r - ct, h - 2 ct): y scaled by (h/2 - ct)/(h/2) about the centroid */
    cone_dims(o->core_dims, o->gd[0] - s->crust, o->gd[1] - 0.5 * s->crust);
    break;
  case GRAINS_BOX:
  case GRAINS_SUPERQUADRIC:
    for (int k = 0; k < 3; ++k) o->core_dims[k] = d[k] - s->crust;
    break;
  case GRAINS_POLYHEDRON:
    o->core_verts = (double*)malloc(sizeof(double) * 3 * o->nv);
    for (int i = 0; i < o->nv; ++i)
      for (int k = 0; k < 3; ++k) o->core_verts[3 * i + k] = o->verts[3 * i + k] * o->scaling[k];
    break;
  default: break;
  }
  return GRAINS_OK;
}

/* sgn (MiscMath.hh:68-72) */
static inline double sgn(double x) { return (double)((0.0 < x) - (x < 0.0)); }

/* Superquadric::support (Superquadric.cpp:169-205) */
static v3 sq_support(const double* d, v3 v) {
  double a = d[0], b = d[1], c = d[2], n1 = d[3], n2 = d[4];
  double ax = fabs(v.x), ay = fabs(v.y), az = fabs(v.z);
  double sx = sgn(v.x), sy = sgn(v.y), sz = sgn(v.z);
  if (ax == 0.0) {
    if (ay == 0.0) return V(0.0, 0.0, sz * c);
    double alpha = pow(c / b * az / ay, 1.0 / (n1 - 1.0));
    double yt = 1.0 / pow(1.0 + pow(alpha, n1), 1.0 / n1);
    return V(0.0, sy * b * yt, sz * alpha * c * yt);
  }
  double alpha = pow(b / a * ay / ax, 1.0 / (n2 - 1.0));
  double temp = 1.0 + pow(alpha, n2);
  double gamma = pow(temp, (n1 - n2) / (n2 * (n1 - 1.0)));
  double beta = gamma * pow(c / a * az / ax, 1.0 / (n1 - 1.0));
  double xt = 1.0 / pow(pow(temp, n1 / n2) + pow(beta, n1), 1.0 / n1);
  return V(sx * a * xt, sy * alpha * b * xt, sz * beta * c * xt);
}

/* Support of shape `o` (full shape when core == 0, crust core otherwise).
 * Sphere.cpp:111-114 (r * v/|v|, NaN at v = 0), Box.cpp:129-134,
 * Rectangle.cpp:115-120, Superquadric.cpp:169-205, Cylinder.cpp:118-131,
 * Cone.cpp:126-141; polyhedron: first vertex maximising v.x. */
static v3 support(const oshape* o, v3 v, int core) {
  const double* d = core ? o->core_dims : o->gd;
  switch (o->s.kind) {
  case GRAINS_SPHERE: {
    double n = vnorm(v);
    v3 u = vdiv(v, n);
    return vscale(d[0], u);
  }
  case GRAINS_BOX:
    return V(v.x < 0.0 ? -d[0] : d[0], v.y < 0.0 ? -d[1] : d[1], v.z < 0.0 ? -d[2] : d[2]);
  case GRAINS_SUPERQUADRIC:
    return sq_support(d, v);
  case GRAINS_RECTANGLE:
    return V(v.x < 0.0 ? -d[0] : d[0], v.y < 0.0 ? -d[1] : d[1], 0.0);
  case GRAINS_CYLINDER: {
    double s = sqrt(v.x * v.x + v.z * v.z);
    double hy = v.y < 0.0 ? -d[1] : d[1];
    if (s > EPS) {
      double k = d[0] / s;
      return V(v.x * k, hy, v.z * k);
    }
    return V(0.0, hy, 0.0);
  }
  case GRAINS_CONE: {
    if (v.y > vnorm(v) * d[2]) return V(0.0, 3.0 * d[1], 0.0);
    double s = sqrt(v.x * v.x + v.z * v.z);
    if (s > EPS) {
      double k = d[0] / s;
      return V(v.x * k, -d[1], v.z * k);
    }
    return V(0.0, -d[1], 0.0);
  }
  case GRAINS_POLYHEDRON: {
    const double* P = core ? o->core_verts : o->verts;
    int best = 0;
    double bd = vdot(v, V(P[0], P[1], P[2]));
    for (int i = 1; i < o->nv; ++i) {
      double dd = vdot(v, V(P[3 * i], P[3 * i + 1], P[3 * i + 2]));
      if (dd > bd) { bd = dd; best = i; }
    }
    return V(P[3 * best], P[3 * best + 1], P[3 * best + 2]);
  }
  }
  return V(0, 0, 0);
}

/* ---- OBB, Grains/CollisionDetection/src/OBB.cpp:32-94 --------------------- */
static int obb_intersect(const double* a, const double* b, const tr3* A, const tr3* B) {
  m3 ori = mtrans(&A->R);
  v3 c3 = mv(&ori, vsub(B->o, A->o));
  ori = mmul(&ori, &B->R);
  double cen[3] = {c3.x, c3.y, c3.z};
  double o[3][3], oa[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      o[i][j] = ori.m[3 * i + j];
      oa[i][j] = fabs(o[i][j]) + LOWEPS;
    }
  for (int i = 0; i < 3; ++i)
    if (fabs(cen[i]) > (a[i] + b[0] * oa[i][0] + b[1] * oa[i][1] + b[2] * oa[i][2])) return 0;
  for (int i = 0; i < 3; ++i)
    if (fabs(cen[0] * o[0][i] + cen[1] * o[1][i] + cen[2] * o[2][i]) >
        (b[i] + a[0] * oa[0][i] + a[1] * oa[1][i] + a[2] * oa[2][i]))
      return 0;
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) {
      int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
      if (fabs(cen[i2] * o[i1][j] - cen[i1] * o[i2][j]) >
          (a[i1] * oa[i2][j] + a[i2] * oa[i1][j] + b[j1] * oa[i][j2] + b[j2] * oa[i][j1]))
        return 0;
    }
  return 1;
}

/* ---- GJK with Johnson's distance sub-algorithm ----------------------------
 * Grains/CollisionDetection/src/GJK_JH.cpp:9-207 (computeDet, valid,
 * computeVector, computePoints, proper, closest, degenerate) and 310-391
 * (computeClosestPoints_GJK_JH). */
typedef struct {
  v3 y[4], p[4], q[4];
  double det[16][4], dp[4][4];
} jh_state;

static void jh_compute_det(jh_state* S, unsigned bits, unsigned last, unsigned last_bit,
                           unsigned all_bits) {
  v3* y = S->y;
  for (unsigned i = 0, bit = 1; i < 4; ++i, bit <<= 1)
    if (bits & bit) S->dp[i][last] = S->dp[last][i] = vdot(y[i], y[last]);
  S->dp[last][last] = vdot(y[last], y[last]);
  double(*det)[4] = S->det;
  double(*dp)[4] = S->dp;
  det[last_bit][last] = 1.0;
  for (unsigned j = 0, sj = 1; j < 4; ++j, sj <<= 1) {
    if (bits & sj) {
      unsigned s2 = sj | last_bit;
      det[s2][j] = dp[last][last] - dp[last][j];
      det[s2][last] = dp[j][j] - dp[j][last];
      for (unsigned k = 0, sk = 1; k < j; ++k, sk <<= 1) {
        if (bits & sk) {
          unsigned s3 = sk | s2;
          det[s3][k] = det[s2][j] * (dp[j][j] - dp[j][k]) + det[s2][last] * (dp[last][j] - dp[last][k]);
          det[s3][j] = det[sk | last_bit][k] * (dp[k][k] - dp[k][j]) +
                       det[sk | last_bit][last] * (dp[last][k] - dp[last][j]);
          det[s3][last] = det[sk | sj][k] * (dp[k][k] - dp[k][last]) +
                          det[sk | sj][j] * (dp[j][k] - dp[j][last]);
        }
      }
    }
  }
  if (all_bits == 15) {
    det[15][0] = det[14][1] * (dp[1][1] - dp[1][0]) + det[14][2] * (dp[2][1] - dp[2][0]) +
                 det[14][3] * (dp[3][1] - dp[3][0]);
    det[15][1] = det[13][0] * (dp[0][0] - dp[0][1]) + det[13][2] * (dp[2][0] - dp[2][1]) +
                 det[13][3] * (dp[3][0] - dp[3][1]);
    det[15][2] = det[11][0] * (dp[0][0] - dp[0][2]) + det[11][1] * (dp[1][0] - dp[1][2]) +
                 det[11][3] * (dp[3][0] - dp[3][2]);
    det[15][3] = det[7][0] * (dp[0][0] - dp[0][3]) + det[7][1] * (dp[1][0] - dp[1][3]) +
                 det[7][2] * (dp[2][0] - dp[2][3]);
  }
}

static int jh_valid(unsigned s, unsigned all_bits, double (*det)[4]) {
  for (unsigned i = 0, bit = 1; i < 4; ++i, bit <<= 1) {
    if (all_bits & bit) {
      if (s & bit) {
        if (det[s][i] <= EPS) return 0;
      } else if (det[s | bit][i] > 0.0)
        return 0;
    }
  }
  return 1;
}

static v3 jh_vector(unsigned bits, const v3* y, double (*det)[4]) {
  double sum = 0.0;
  v3 v = V(0, 0, 0);
  for (unsigned i = 0, bit = 1; i < 4; ++i, bit <<= 1)
    if (bits & bit) {
      sum += det[bits][i];
      v = vadd(v, vscale(det[bits][i], y[i]));
    }
  return vscale(1.0 / sum, v);
}

static void jh_points(unsigned bits, double det[16][4], const v3* p, const v3* q, v3* p1, v3* p2) {
  double sum = 0.0;
  *p1 = V(0, 0, 0);
  *p2 = V(0, 0, 0);
  for (unsigned i = 0, bit = 1; i < 4; ++i, bit <<= 1)
    if (bits & bit) {
      sum += det[bits][i];
      *p1 = vadd(*p1, vscale(det[bits][i], p[i]));
      *p2 = vadd(*p2, vscale(det[bits][i], q[i]));
    }
  double s = 1.0 / sum;
  *p1 = vscale(s, *p1);
  *p2 = vscale(s, *p2);
}

static int jh_proper(unsigned s, double (*det)[4]) {
  for (unsigned i = 0, bit = 1; i < 4; ++i, bit <<= 1)
    if ((s & bit) && det[s][i] <= EPS) return 0;
  return 1;
}

static int jh_closest(jh_state* S, unsigned* bits, unsigned last, unsigned last_bit,
                      unsigned all_bits, v3* v) {
  unsigned s;
  jh_compute_det(S, *bits, last, last_bit, all_bits);
  for (s = *bits; s; --s) {
    if ((s & *bits) == s) {
      if (jh_valid(s | last_bit, all_bits, S->det)) {
        *bits = s | last_bit;
        *v = jh_vector(*bits, S->y, S->det);
        return 1;
      }
    }
  }
  if (jh_valid(last_bit, all_bits, S->det)) {
    *bits = last_bit;
    *v = S->y[last];
    return 1;
  }
  double min_dist2 = INFINITY;
  for (s = all_bits; s; --s) {
    if ((s & all_bits) == s) {
      if (jh_proper(s, S->det)) {
        v3 u = jh_vector(s, S->y, S->det);
        double d2 = vnorm2(u);
        if (d2 < min_dist2) {
          min_dist2 = d2;
          *bits = s;
          *v = u;
        }
      }
    }
  }
  return 0;
}

static int jh_degenerate(unsigned all_bits, const v3* y, v3 w) {
  for (unsigned i = 0, bit = 1; i < 4; ++i, bit <<= 1)
    if ((all_bits & bit) && y[i].x == w.x && y[i].y == w.y && y[i].z == w.z) return 1;
  return 0;
}

/* ---- signed-volume distance sub-algorithm (physics semantics) -------------
 * The closest point of a simplex of up to four points to the origin, from
 * the signs of the origin's barycentric coordinates (ratios of signed
 * volumes / areas), after the method of Montanari, Petrinic and Barbieri
 * (ACM TOG 2017), of which the reference's unused GJK_SV.cpp is another
 * rendition.  No absolute thresholds: a degenerate feature (zero area or
 * volume) hands over to its sub-features.  y[0] is the newest point; only
 * features containing it are candidates, as GJK guarantees.
 * Every ratio is taken relative to the vertex a = y[0]: edges e = y - a are
 * exact for nearby points, and the coordinates come from products of a with
 * edges (n.(e2 x a), a.(ac x ad), ...), never from cross products of two
 * far-away vertices such as b x c, whose cancellation costs eps |a|^2 / |n|
 * when the simplex has shrunk to the contact (|e| ~ 1e-6 |a| near
 * convergence).  The first coordinate is 1 minus the others, and the closest
 * point is a + sum lambda_k e_k. */
typedef struct {
  v3 y[4], p[4], q[4];
  double lam[4];
  int n;
} sv_simplex;
typedef struct {
  v3 v;
  double lam[4];
  unsigned mask;
} sv_res;

static sv_res sv_seg(const sv_simplex* S, int ia, int ib) {
  sv_res r;
  memset(&r, 0, sizeof(r));
  v3 a = S->y[ia], b = S->y[ib];
  v3 t = vsub(b, a);
  double tt = vdot(t, t);
  if (tt > 0.0) {
    double l = -vdot(a, t) / tt;
    if (l > 0.0 && l < 1.0) {
      r.v = vadd(a, vscale(l, t));
      r.lam[ia] = 1.0 - l;
      r.lam[ib] = l;
      r.mask = (1u << ia) | (1u << ib);
      return r;
    }
    if (l >= 1.0) {
      r.v = b;
      r.lam[ib] = 1.0;
      r.mask = 1u << ib;
      return r;
    }
  }
  r.v = a;
  r.lam[ia] = 1.0;
  r.mask = 1u << ia;
  return r;
}

static sv_res sv_better(sv_res a, sv_res b) { return vnorm2(b.v) < vnorm2(a.v) ? b : a; }

static sv_res sv_tri(const sv_simplex* S, int ia, int ib, int ic) {
  v3 a = S->y[ia];
  v3 e1 = vsub(S->y[ib], a), e2 = vsub(S->y[ic], a);
  v3 n = vcross(e1, e2);
  double nn = vdot(n, n);
  double la = 0.0, lb = 0.0, lc = 0.0;
  if (nn > 0.0) {
    lb = vdot(n, vcross(e2, a)) / nn;
    lc = vdot(n, vcross(a, e1)) / nn;
    la = (1.0 - lb) - lc;
    if (la > 0.0 && lb > 0.0 && lc > 0.0) {
      sv_res r;
      memset(&r, 0, sizeof(r));
      r.v = vadd(vadd(a, vscale(lb, e1)), vscale(lc, e2));
      r.lam[ia] = la;
      r.lam[ib] = lb;
      r.lam[ic] = lc;
      r.mask = (1u << ia) | (1u << ib) | (1u << ic);
      return r;
    }
  }
  int ab = !(nn > 0.0) || lc <= 0.0, ac = !(nn > 0.0) || lb <= 0.0;
  if (!ab && !ac) ab = ac = 1; /* only la <= 0: rounding; keep a's edges */
  if (ab && ac) return sv_better(sv_seg(S, ia, ib), sv_seg(S, ia, ic));
  return ab ? sv_seg(S, ia, ib) : sv_seg(S, ia, ic);
}

static sv_res sv_tet(const sv_simplex* S) {
  v3 a = S->y[0];
  v3 ab = vsub(S->y[1], a), ac = vsub(S->y[2], a), ad = vsub(S->y[3], a);
  double D = vdot(ab, vcross(ac, ad));
  double l[4] = {0, 0, 0, 0};
  if (D != 0.0) {
    l[1] = -vdot(a, vcross(ac, ad)) / D;
    l[2] = vdot(ab, vcross(ad, a)) / D;
    l[3] = vdot(ab, vcross(a, ac)) / D;
    l[0] = ((1.0 - l[1]) - l[2]) - l[3];
    if (l[0] > 0.0 && l[1] > 0.0 && l[2] > 0.0 && l[3] > 0.0) {
      sv_res r;
      memset(&r, 0, sizeof(r));
      r.v = V(0.0, 0.0, 0.0);
      for (int i = 0; i < 4; ++i) r.lam[i] = l[i];
      r.mask = 15u;
      return r;
    }
  }
  int fb = D == 0.0 || l[1] <= 0.0, fc = D == 0.0 || l[2] <= 0.0, fd = D == 0.0 || l[3] <= 0.0;
  if (!fb && !fc && !fd) fb = fc = fd = 1;
  sv_res best;
  int have = 0;
  if (fb) { best = sv_tri(S, 0, 2, 3); have = 1; }
  if (fc) { sv_res r = sv_tri(S, 0, 1, 3); best = have ? sv_better(best, r) : r; have = 1; }
  if (fd) { sv_res r = sv_tri(S, 0, 1, 2); best = have ? sv_better(best, r) : r; }
  return best;
}

/* keep the points of r.mask, newest first, with their weights */
static void sv_compact(sv_simplex* S, const sv_res* r) {
  sv_simplex N;
  memset(&N, 0, sizeof(N));
  for (int i = 0; i < 4; ++i)
    if (r->mask & (1u << i)) {
      N.y[N.n] = S->y[i];
      N.p[N.n] = S->p[i];
      N.q[N.n] = S->q[i];
      N.lam[N.n] = r->lam[i];
      ++N.n;
    }
  *S = N;
}

/* Momentum is applied to this many support points, then plain GJK: on the
 * GJK distance test set of superquadric pairs this takes 25 % fewer
 * iterations than plain GJK at gap 1e-6 (19 % at 1e-10), with a worst case no
 * worse than plain GJK; uncapped, the momentum zig-zags near the optimum and
 * can reach the iteration cap. */
#define SV_NESTEROV_ITERS 8

/* GJK with the signed-volume sub-algorithm, physics semantics: crust cores,
 * the pair's unit frame (power-of-two scale), v0 = centre difference, the
 * van den Bergen gap test, NESTEROV momentum on the support direction.
 * Separation early-out: once the lower bound mu exceeds `sep` (the crust
 * sum, in the caller's unit) the cores are farther apart than the crusts can
 * bridge, the pair has no contact whatever the exact distance, and the
 * search stops with *separated = 1 (pass INFINITY for a plain distance
 * query).  Every contact decision equals that of the full search: the full
 * search reports no contact exactly when its distance, an upper bound, is
 * above sep, and mu > sep implies it. */
static double gjk_sv(const oshape* A, const oshape* B, const tr3* a2w, const tr3* b2w,
                     uint32_t flags, double rel_tol, double sep, v3* pa, v3* pb, int* nb_iter,
                     int* deep, int* separated) {
  const int nesterov = (flags & GRAINS_NESTEROV) != 0;
  const double rel = rel_tol > 0 ? rel_tol : LOWEPS;
  tr3 b2a = trelative(a2w, b2w);
  int e;
  frexp(A->rc_exact + B->rc_exact, &e);
  const double sc = ldexp(1.0, -e);
  b2a.o = vscale(sc, b2a.o);
  const double sep_s = sc * sep;
  sv_simplex S;
  memset(&S, 0, sizeof(S));
  v3 v = vneg(b2a.o), w = V(0, 0, 0), pl = V(0, 0, 0), ql = V(0, 0, 0);
  double dist = vnorm(v), mu = 0.0;
  v3 dir = v, wprev = v;
  int accel = nesterov, num_iter = 0, inside = 0, ran = 0;
  *separated = 0;
  while (dist > HIGHEPS && num_iter < 1000) {
    v3 d = v;
    /* momentum on the first SV_NESTEROV_ITERS support points only: near the
     * optimum it zig-zags (sublinear on superquadrics) */
    if (accel && num_iter >= SV_NESTEROV_ITERS) accel = 0;
    if (accel) {
      double delta = (double)(num_iter + 1) / (double)(num_iter + 3);
      v3 yk = vadd(vscale(delta, v), vscale(1.0 - delta, wprev));
      double ny = vnorm(yk), nd = vnorm(dir);
      if (num_iter != 0 && ny > HIGHEPS && nd > HIGHEPS)
        d = vadd(vscale(delta / nd, dir), vscale((1.0 - delta) / ny, yk));
      dir = d;
    }
    pl = vscale(sc, support(A, vneg(d), 1));
    ql = vscale(sc, support(B, vm(d, &b2a.R), 1));
    w = vsub(pl, tapply(&b2a, ql));
    ran = 1;
    double lb = accel ? vdot(d, w) / vnorm(d) : vdot(v, w) / dist;
    if (mu < lb) mu = lb;
    if (mu > sep_s) {
      *separated = 1;
      break;
    }
    if (dist - mu <= dist * rel) break;
    int degen = 0;
    for (int i = 0; i < S.n; ++i)
      if (S.y[i].x == w.x && S.y[i].y == w.y && S.y[i].z == w.z) degen = 1;
    if (degen) {
      if (accel) { accel = 0; continue; }
      break;
    }
    /* the simplex holds at most 3 points here (a tetrahedron either
     * encloses the origin or is reduced), so the shift loses nothing and can
     * be undone */
    for (int i = 3; i > 0; --i) {
      S.y[i] = S.y[i - 1];
      S.p[i] = S.p[i - 1];
      S.q[i] = S.q[i - 1];
      S.lam[i] = S.lam[i - 1];
    }
    S.y[0] = w;
    S.p[0] = pl;
    S.q[0] = ql;
    S.lam[0] = 0.0;
    const int n_prev = S.n;
    S.n = S.n + 1;
    ++num_iter;
    sv_res r;
    if (S.n == 1) {
      memset(&r, 0, sizeof(r));
      r.v = w;
      r.lam[0] = 1.0;
      r.mask = 1u;
    } else if (S.n == 2) {
      r = sv_seg(&S, 0, 1);
    } else if (S.n == 3) {
      r = sv_tri(&S, 0, 1, 2);
    } else {
      r = sv_tet(&S);
    }
    if (r.mask == 15u) {
      sv_compact(&S, &r);
      inside = 1;
      dist = 0.0;
      break;
    }
    double nd = vnorm(r.v);
    /* No progress: rounding has the last word (a thin triangle's projection
     * can come out worse than its edge).  Drop w and restore the previous
     * simplex, whose |v| is the best bound found; an accelerated search
     * continues unaccelerated, a plain one stops.  The first support point
     * is exempt: v0 (the centre difference) is not a point of the simplex. */
    if (num_iter > 1 && !(nd < dist)) {
      for (int i = 0; i < 3; ++i) {
        S.y[i] = S.y[i + 1];
        S.p[i] = S.p[i + 1];
        S.q[i] = S.q[i + 1];
        S.lam[i] = S.lam[i + 1];
      }
      S.n = n_prev;
      if (accel) { accel = 0; continue; }
      break;
    }
    sv_compact(&S, &r);
    wprev = w;
    v = r.v;
    dist = nd;
  }
  v3 p = V(0, 0, 0), q = V(0, 0, 0);
  if (S.n == 0) {
    if (ran) { p = pl; q = ql; }
  } else {
    for (int i = 0; i < S.n; ++i) {
      p = vadd(p, vscale(S.lam[i], S.p[i]));
      q = vadd(q, vscale(S.lam[i], S.q[i]));
    }
  }
  *pa = vscale(1.0 / sc, p);
  *pb = vscale(1.0 / sc, q);
  if (deep) *deep = inside || dist <= HIGHEPS;
  *nb_iter = num_iter;
  return dist / sc;
}

/* computeClosestPoints_GJK_JH (GJK_JH.cpp:310-391), exactly the reference
 * (the REF_QUIRKS semantics): full shapes, v0 = s_A(0) - b2a(s_B(0)), the
 * break on mu < 1e-4*relError, the stale |v| kept when Johnson's backup
 * procedure is taken.  The physics semantics use gjk_sv. */
static double gjk_jh(const oshape* A, const oshape* B, const tr3* a2w, const tr3* b2w,
                     double rel_tol, v3* pa, v3* pb, int* nb_iter) {
  unsigned bits = 0, last = 0, last_bit = 0, all_bits = 0;
  jh_state S;
  memset(&S, 0, sizeof(S));
  double mu = 0.0;
  int num_iter = 0;
  const double rel = rel_tol > 0 ? rel_tol : LOWEPS;
  const double abs_err = 1.e-4 * rel;
  tr3 b2a = trelative(a2w, b2w);
  v3 v = vsub(support(A, V(0, 0, 0), 0), tapply(&b2a, support(B, V(0, 0, 0), 0)));
  v3 w = V(0, 0, 0);
  double dist = vnorm(v);
  while (bits < 15 && dist > HIGHEPS && num_iter < 1000) {
    last = 0;
    last_bit = 1;
    while (bits & last_bit) {
      ++last;
      last_bit <<= 1;
    }
    S.p[last] = support(A, vneg(v), 0);
    S.q[last] = support(B, vm(v, &b2a.R), 0);
    w = vsub(S.p[last], tapply(&b2a, S.q[last]));
    double lb = vdot(v, w) / dist;
    if (mu < lb) mu = lb;
    if (dist - mu <= dist * rel || mu < abs_err) break;
    if (jh_degenerate(all_bits, S.y, w)) break;
    S.y[last] = w;
    all_bits = bits | last_bit;
    ++num_iter;
    if (!jh_closest(&S, &bits, last, last_bit, all_bits, &v)) break;
    dist = vnorm(v);
  }
  jh_points(bits, S.det, S.p, S.q, pa, pb);
  *nb_iter = num_iter;
  return dist;
}

/* ---- contact info ---------------------------------------------------------- */
typedef struct { v3 cp, cv; double overlap; } cinfo;
static const cinfo NO_CONTACT = {{0, 0, 0}, {0, 0, 0}, 0.0};

/* closestPointsSpheres (CollisionDetection.cpp:27-48) */
static cinfo cp_spheres(const oshape* A, const oshape* B, const tr3* a2w, const tr3* b2w) {
  double rA = A->rc, rB = B->rc;
  v3 cenA = a2w->o;
  v3 vecBA = vsub(b2w->o, cenA);
  double overlap = vnorm(vecBA) - rA - rB;
  vecBA = vdiv(vecBA, vnorm(vecBA));
  if (overlap < 0.0) {
    cinfo c;
    c.cp = vadd(cenA, vscale(rA + 0.5 * overlap, vecBA));
    c.cv = vscale(overlap, vecBA);
    c.overlap = overlap;
    return c;
  }
  return NO_CONTACT;
}

/* closestPointsRectangle (CollisionDetection.cpp:54-90) */
static cinfo cp_rectangle(const oshape* A, const tr3* a2w, const tr3* b2w) {
  v3 rPt = b2w->o;
  v3 rNorm = V(b2w->R.m[2], b2w->R.m[5], b2w->R.m[8]);
  rNorm = vscale(copysign(1.0, vdot(rNorm, vsub(a2w->o, rPt))), rNorm);
  v3 pointA = tapply(a2w, support(A, vm(vneg(rNorm), &a2w->R), 0));
  if (vdot(rNorm, vsub(pointA, rPt)) < 0.0) {
    v3 pointB = vadd(vscale(vdot(vsub(rPt, pointA), rNorm), rNorm), pointA);
    cinfo c;
    c.cp = vscale(0.5, vadd(pointA, pointB));
    c.cv = vsub(pointB, pointA);
    c.overlap = -vnorm(c.cv);
    return c;
  }
  return NO_CONTACT;
}

typedef struct {
  uint64_t gjk_calls, gjk_iters, deep;
} np_counters;

/* closestPointsRigidBodies (CollisionDetection.cpp:181-284).  With physics
 * semantics GJK runs on the crust cores (the commented-out intent of lines
 * 231-236) with the signed-volume GJK (gjk_sv); REF_QUIRKS runs the
 * reference's Johnson GJK on the full shapes. */
static cinfo closest_points(const oshape* A, const oshape* B, const tr3* a2w, const tr3* b2w,
                            uint32_t flags, double rel_tol, np_counters* cnt) {
  if (A->s.kind == GRAINS_SPHERE && B->s.kind == GRAINS_SPHERE)
    return cp_spheres(A, B, a2w, b2w);
  if (B->s.kind == GRAINS_RECTANGLE) return cp_rectangle(A, a2w, b2w);
  const int quirks = (flags & GRAINS_REF_QUIRKS) != 0;
  double overlap = vnorm(vsub(b2w->o, a2w->o)) -
                   (quirks ? A->rc + B->rc : A->rc_exact + B->rc_exact);
  if (overlap < 0.0) {
    if (quirks ? obb_intersect(A->bbox, B->bbox, a2w, b2w)
               : obb_intersect(A->bbox_exact, B->bbox_exact, a2w, b2w)) {
      v3 ptA, ptB;
      int it = 0, deep = 0, separated = 0;
      double ctSum = A->s.crust + B->s.crust;
      double distance = (flags & GRAINS_REF_QUIRKS)
                            ? gjk_jh(A, B, a2w, b2w, rel_tol, &ptA, &ptB, &it)
                            : gjk_sv(A, B, a2w, b2w, flags, rel_tol, ctSum, &ptA, &ptB, &it, &deep,
                                     &separated);
      if (cnt) {
        cnt->gjk_calls++;
        cnt->gjk_iters += (uint64_t)it;
      }
      if (separated) return NO_CONTACT;
      distance -= ctSum;
      if (deep) distance = -ctSum; /* cores overlap: deepest depth the crusts allow */
      if (distance > 0.0) return NO_CONTACT;
      ptA = tapply(a2w, ptA);
      ptB = tapply(b2w, ptB);
      cinfo c;
      c.cp = vdiv(vadd(ptA, ptB), 2.0);
      v3 cv = vsub(ptA, ptB);
      if (deep) {
        /* crust cores overlap: direction of the centre line, depth ctSum */
        if (cnt) cnt->deep++;
        cv = vsub(a2w->o, b2w->o);
        c.cp = vdiv(vadd(a2w->o, b2w->o), 2.0);
      }
      cv = vdiv(cv, vnorm(cv));
      cv = vround(cv, EPS);
      cv = vscale(-distance, cv);
      c.cv = cv;
      c.overlap = distance;
      return c;
    }
  }
  return NO_CONTACT;
}

/* ---- Hooke, HookeContactForceModel.cpp:91-186 ----------------------------- */
typedef struct { double kn, en, etat, muc, kr, muen; } hooke_t;

static void hooke_forces(const hooke_t* h, const cinfo* ci, v3 relVel, v3 relAngVel, double m1,
                         double m2, v3 origin, v3* force, v3* torque) {
  v3 pen = ci->cv;
  v3 delFN = vscale(h->kn, pen);
  pen = vdiv(pen, vnorm(pen));
  pen = vround(pen, EPS);
  v3 v_n = vscale(vdot(relVel, pen), pen);
  v3 v_t = vsub(relVel, v_n);
  double normv_t = vnorm(v_t);
  v3 tangent = V(0, 0, 0);
  if (normv_t > EPS) tangent = vdiv(v_t, normv_t);
  double avmass = m1 * m2 / (m1 + m2);
  double omega0 = sqrt(h->kn / avmass);
  if (avmass == 0.0) {
    avmass = m2 == 0.0 ? 0.5 * m1 : 0.5 * m2;
    omega0 = 2.0 * sqrt(h->kn / avmass);
  }
  double muen = -omega0 * h->muen;
  delFN = vadd(delFN, vscale(-2.0 * muen * avmass, v_n));
  double normFN = vnorm(delFN);
  v3 delFT = vscale(-h->etat * 2.0 * avmass, v_t);
  double fn = h->muc * normFN;
  double ft = vnorm(delFT);
  if (fn < ft) delFT = vscale(-fn, tangent);
  v3 delM = V(0, 0, 0);
  if (h->kr != 0.0) {
    v3 wn = vscale(vdot(relAngVel, pen), pen);
    v3 wt = vsub(relAngVel, wn);
    double normwt = vnorm(wt);
    delM = vscale(-h->kr * normFN * 0.001, wn);
    if (normwt > EPS) delM = vsub(delM, vscale(h->kr * normFN, wt));
  }
  /* computeForces: torce.addForce(F, cp - origin) (+ rolling moment) */
  v3 F = vadd(delFN, delFT);
  *force = F;
  *torque = vcross(vsub(ci->cp, origin), F);
  if (h->kr != 0.0) *torque = vadd(*torque, delM);
}

/* ---- world ------------------------------------------------------------------ */
typedef struct {
  grains_sim_t s;
  double h;
  uint32_t n3[3], ncells, base;
  uint32_t* nbr; /* 27 per cell */
} osim;

struct go_world {
  grains_config_t cfg;
  char err[256];
  oshape* shapes;
  uint32_t nshapes;
  double* verts;
  hooke_t* tables;
  uint32_t nmat, ntables;
  osim* sims;
  uint32_t nsims;
  /* obstacles */
  uint32_t nobs;
  uint32_t *obs_shape, *obs_sim;
  tr3* obs_tr;
  /* particles */
  uint64_t n;
  uint32_t *shape, *sim;
  tr3* tr;
  v3 *vel, *omega, *force, *torque;
  v3 *last_force, *last_torque;
  v3* hist; /* 4 per particle: ADAMS_BASHFORTH, the previous v, w, a, alpha (NaN: none);
              LEAPFROG, hist[0].x NaN until the first (half) kick */
  int32_t* result;
  /* broad phase of the last step */
  uint32_t* keys;
  uint32_t* sorted;
  uint32_t* cell_start; /* per global cell: first index in `sorted`  */
  uint32_t* cell_count;
  uint32_t ncells_total;
  /* per-particle pair lists of the last step */
  uint32_t** cand;
  uint32_t* ncand;
  uint32_t* capcand;
  uint32_t** cont;
  uint32_t* ncont;
  uint32_t* capcont;
  grains_stats_t stats;
  uint64_t *pp_gjk, *pp_iter, *pp_deep, *pp_wall, *pp_ood;
};

static int set_err(go_world* w, int code, const char* msg) {
  snprintf(w->err, sizeof(w->err), "%s", msg);
  return code;
}

go_world* go_create(const grains_config_t* cfg) {
  go_world* w = (go_world*)calloc(1, sizeof(go_world));
  if (!w) return NULL;
  w->cfg = *cfg;
  return w;
}

static void free_particles(go_world* w) {
  free(w->shape); free(w->sim); free(w->tr); free(w->vel); free(w->omega);
  free(w->force); free(w->torque); free(w->last_force); free(w->last_torque);
  free(w->result); free(w->keys); free(w->sorted); free(w->hist); w->hist = NULL;
  if (w->cand)
    for (uint64_t i = 0; i < w->n; ++i) { free(w->cand[i]); free(w->cont[i]); }
  free(w->cand); free(w->ncand); free(w->capcand);
  free(w->cont); free(w->ncont); free(w->capcont);
  free(w->pp_gjk); free(w->pp_iter); free(w->pp_deep); free(w->pp_wall); free(w->pp_ood);
  w->shape = w->sim = NULL; w->tr = NULL; w->cand = w->cont = NULL;
  w->n = 0;
}

void go_destroy(go_world* w) {
  if (!w) return;
  free_particles(w);
  for (uint32_t i = 0; i < w->nshapes; ++i) free(w->shapes[i].core_verts);
  free(w->shapes); free(w->verts); free(w->tables);
  for (uint32_t i = 0; i < w->nsims; ++i) free(w->sims[i].nbr);
  free(w->sims);
  free(w->obs_shape); free(w->obs_sim); free(w->obs_tr);
  free(w->cell_start); free(w->cell_count);
  free(w);
}

const char* go_last_error(const go_world* w) { return w ? w->err : "null world"; }

int go_set_shapes(go_world* w, const grains_shape_t* shapes, uint32_t n_shapes,
                  const double* vertices, uint32_t n_vertices) {
  for (uint32_t i = 0; i < w->nshapes; ++i) free(w->shapes[i].core_verts);
  free(w->shapes);
  free(w->verts);
  w->verts = NULL;
  if (n_vertices) {
    w->verts = (double*)malloc(sizeof(double) * 3 * n_vertices);
    memcpy(w->verts, vertices, sizeof(double) * 3 * n_vertices);
  }
  w->shapes = (oshape*)calloc(n_shapes, sizeof(oshape));
  w->nshapes = n_shapes;
  for (uint32_t i = 0; i < n_shapes; ++i) {
    int rc = shape_derive(&w->shapes[i], &shapes[i], w->verts);
    if (rc) return set_err(w, rc, "unsupported or invalid shape");
  }
  return GRAINS_OK;
}

int go_set_contact_tables(go_world* w, const grains_hooke_t* t, uint32_t n_materials,
                          uint32_t n_tables) {
  uint32_t per = n_materials * (n_materials + 1) / 2;
  free(w->tables);
  w->tables = (hooke_t*)malloc(sizeof(hooke_t) * per * n_tables);
  for (uint32_t i = 0; i < per * n_tables; ++i) {
    hooke_t* h = &w->tables[i];
    h->kn = t[i].kn; h->en = t[i].en; h->etat = t[i].etat; h->muc = t[i].muc; h->kr = t[i].kr;
    /* m_muen (HookeContactForceModel.cpp:56) */
    h->muen = log(h->en) / sqrt(G_PI * G_PI + log(h->en) * log(h->en));
  }
  w->nmat = n_materials;
  w->ntables = n_tables;
  return GRAINS_OK;
}

/* LinkedCell ctor cell counts (LinkedCell.cpp:23-29) */
void go_cell_counts(const double* min, const double* max, double h, uint32_t* n3) {
  for (int k = 0; k < 3; ++k) {
    double c = EPS;
    c += max[k] - min[k];
    c /= h;
    n3[k] = (uint32_t)c;
  }
}

/* neighbour table (LinkedCell.cpp:35-72); periodic directions wrap */
static void neighbor_table_p(const uint32_t* n3, int periodic, uint32_t* table) {
  uint32_t nc = n3[0] * n3[1] * n3[2];
  for (uint32_t c = 0; c < nc; ++c) {
    uint32_t off = c * 27;
    int cx = (int)(c % n3[0]), cy = (int)((c / n3[0]) % n3[1]), cz = (int)(c / (n3[0] * n3[1]));
    for (int k = -1; k < 2; ++k)
      for (int j = -1; j < 2; ++j)
        for (int i = -1; i < 2; ++i) {
          int nx = cx + i, ny = cy + j, nz = cz + k;
          if (periodic & 1) nx = (nx + (int)n3[0]) % (int)n3[0];
          if (periodic & 2) ny = (ny + (int)n3[1]) % (int)n3[1];
          if (periodic & 4) nz = (nz + (int)n3[2]) % (int)n3[2];
          if (nx >= 0 && nx < (int)n3[0] && ny >= 0 && ny < (int)n3[1] && nz >= 0 && nz < (int)n3[2])
            table[off++] = (uint32_t)nx + (uint32_t)ny * n3[0] + (uint32_t)nz * n3[0] * n3[1];
          else
            table[off++] = UINT32_MAX;
        }
  }
}
void go_neighbor_table(const uint32_t* n3, uint32_t* table) { neighbor_table_p(n3, 0, table); }

/* computeCellId + computeLinearCellHash (LinkedCell.cpp:117-149).  The
 * reference converts floor() to uint directly (undefined below 0 and reads
 * out of bounds above n); here each index is clamped into the grid and the
 * point counted as out of domain. */
static inline uint32_t cell_hash1(const double* min, const double* max, double h,
                                  const uint32_t* n3, v3 p, int* ood) {
  double pp[3] = {p.x, p.y, p.z};
  double f[3] = {floor((p.x - min[0]) / h), floor((p.y - min[1]) / h), floor((p.z - min[2]) / h)};
  uint32_t c[3];
  for (int k = 0; k < 3; ++k) {
    /* centres outside [min, max] are counted; the partial last cell beyond
     * n*h folds into cell n-1 */
    if (!(pp[k] >= min[k] && pp[k] <= max[k])) *ood = 1;
    if (!(f[k] >= 0.0)) c[k] = 0;
    else if (f[k] >= (double)n3[k]) c[k] = n3[k] - 1;
    else c[k] = (uint32_t)f[k];
  }
  return (c[2] * n3[1] + c[1]) * n3[0] + c[0];
}

uint64_t go_cell_hash(const double* min, double h, const uint32_t* n3, const double* pos,
                      uint64_t n, uint32_t* keys) {
  uint64_t bad = 0;
  const double max[3] = {min[0] + n3[0] * h, min[1] + n3[1] * h, min[2] + n3[2] * h};
  for (uint64_t i = 0; i < n; ++i) {
    int ood = 0;
    keys[i] = cell_hash1(min, max, h, n3, V(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]), &ood);
    bad += (uint64_t)ood;
  }
  return bad;
}

static int finish_sims(go_world* w) {
  uint32_t base = 0;
  for (uint32_t s = 0; s < w->nsims; ++s) {
    osim* S = &w->sims[s];
    double h = S->s.cell_size;
    if (!(h > 0)) {
      /* Grains.cpp:226-242: max circumscribed radius over the sim's particles */
      const int quirks = (w->cfg.flags & GRAINS_REF_QUIRKS) != 0;
      double lc = 0;
      for (uint64_t i = 0; i < w->n; ++i) {
        if (w->sim[i] != s) continue;
        double r = quirks ? w->shapes[w->shape[i]].rc : w->shapes[w->shape[i]].rc_exact;
        if (r > lc) lc = r;
      }
      double coeff = S->s.cell_size_factor < 1.0 ? 1.0 : S->s.cell_size_factor;
      h = coeff * 2.0 * lc;
    }
    if (!(h > 0)) return set_err(w, GRAINS_EINVAL, "cell size is zero");
    S->h = h;
    go_cell_counts(S->s.domain_min, S->s.domain_max, h, S->n3);
    for (int k = 0; k < 3; ++k) {
      if (S->n3[k] == 0) return set_err(w, GRAINS_EINVAL, "domain smaller than one cell");
      if ((S->s.periodic >> k & 1) && S->n3[k] < 3)
        return set_err(w, GRAINS_EINVAL, "periodic direction needs >= 3 cells");
    }
    S->ncells = S->n3[0] * S->n3[1] * S->n3[2];
    S->base = base;
    base += S->ncells;
    free(S->nbr);
    S->nbr = (uint32_t*)malloc(sizeof(uint32_t) * 27 * S->ncells);
    neighbor_table_p(S->n3, S->s.periodic, S->nbr);
  }
  w->ncells_total = base;
  free(w->cell_start);
  free(w->cell_count);
  w->cell_start = (uint32_t*)calloc(base, sizeof(uint32_t));
  w->cell_count = (uint32_t*)calloc(base, sizeof(uint32_t));
  return GRAINS_OK;
}

int go_set_sims(go_world* w, const grains_sim_t* sims, uint32_t n_sims) {
  for (uint32_t i = 0; i < w->nsims; ++i) free(w->sims[i].nbr);
  free(w->sims);
  w->sims = (osim*)calloc(n_sims, sizeof(osim));
  w->nsims = n_sims;
  for (uint32_t i = 0; i < n_sims; ++i) w->sims[i].s = sims[i];
  return w->n ? finish_sims(w) : GRAINS_OK;
}

int go_set_obstacles(go_world* w, const uint32_t* shape, const uint32_t* sim,
                     const double* transforms, uint32_t n) {
  free(w->obs_shape); free(w->obs_sim); free(w->obs_tr);
  w->nobs = n;
  w->obs_shape = (uint32_t*)malloc(sizeof(uint32_t) * (n + 1));
  w->obs_sim = (uint32_t*)malloc(sizeof(uint32_t) * (n + 1));
  w->obs_tr = (tr3*)malloc(sizeof(tr3) * (n + 1));
  for (uint32_t i = 0; i < n; ++i) {
    w->obs_shape[i] = shape[i];
    w->obs_sim[i] = sim ? sim[i] : 0;
    w->obs_tr[i] = tr_from12(transforms + 12 * (size_t)i);
  }
  return GRAINS_OK;
}

/* no history: the next two-step update is a first-order one */
static void hist_reset(go_world* w) {
  for (uint64_t i = 0; i < 4 * w->n; ++i) w->hist[i] = V(NAN, NAN, NAN);
}

int go_set_particles(go_world* w, const uint32_t* shape, const uint32_t* sim, uint64_t n) {
  free_particles(w);
  w->n = n;
  w->shape = (uint32_t*)malloc(sizeof(uint32_t) * n);
  w->sim = (uint32_t*)malloc(sizeof(uint32_t) * n);
  w->tr = (tr3*)calloc(n, sizeof(tr3));
  w->vel = (v3*)calloc(n, sizeof(v3));
  w->omega = (v3*)calloc(n, sizeof(v3));
  w->force = (v3*)calloc(n, sizeof(v3));
  w->torque = (v3*)calloc(n, sizeof(v3));
  w->last_force = (v3*)calloc(n, sizeof(v3));
  w->last_torque = (v3*)calloc(n, sizeof(v3));
  w->result = (int32_t*)calloc(n, sizeof(int32_t));
  w->keys = (uint32_t*)calloc(n, sizeof(uint32_t));
  w->sorted = (uint32_t*)calloc(n, sizeof(uint32_t));
  w->cand = (uint32_t**)calloc(n, sizeof(uint32_t*));
  w->ncand = (uint32_t*)calloc(n, sizeof(uint32_t));
  w->capcand = (uint32_t*)calloc(n, sizeof(uint32_t));
  w->cont = (uint32_t**)calloc(n, sizeof(uint32_t*));
  w->ncont = (uint32_t*)calloc(n, sizeof(uint32_t));
  w->capcont = (uint32_t*)calloc(n, sizeof(uint32_t));
  w->pp_gjk = (uint64_t*)calloc(n, sizeof(uint64_t));
  w->pp_iter = (uint64_t*)calloc(n, sizeof(uint64_t));
  w->pp_deep = (uint64_t*)calloc(n, sizeof(uint64_t));
  w->pp_wall = (uint64_t*)calloc(n, sizeof(uint64_t));
  w->pp_ood = (uint64_t*)calloc(n, sizeof(uint64_t));
  w->hist = (v3*)malloc(sizeof(v3) * 4 * n);
  hist_reset(w);
  for (uint64_t i = 0; i < n; ++i) {
    w->shape[i] = shape[i];
    if (shape[i] >= w->nshapes) return set_err(w, GRAINS_EINVAL, "shape index out of range");
    w->sim[i] = sim ? sim[i] : 0;
    if (w->sim[i] >= w->nsims) return set_err(w, GRAINS_EINVAL, "sim index out of range");
    m3 I = {{1, 0, 0, 0, 1, 0, 0, 0, 1}};
    w->tr[i].R = I;
  }
  return finish_sims(w);
}

int go_set_state(go_world* w, const double* pos, const double* quat, const double* vel,
                 const double* omega, uint64_t n) {
  if (n != w->n) return set_err(w, GRAINS_EINVAL, "particle count mismatch");
  for (uint64_t i = 0; i < n; ++i) {
    if (pos) w->tr[i].o = V(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    if (quat) {
      qt q = {quat[4 * i], quat[4 * i + 1], quat[4 * i + 2], quat[4 * i + 3]};
      w->tr[i].R = matrix_from_quat(q);
    }
    if (vel) w->vel[i] = V(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
    if (omega) w->omega[i] = V(omega[3 * i], omega[3 * i + 1], omega[3 * i + 2]);
  }
  hist_reset(w);
  return GRAINS_OK;
}

int go_get_state(go_world* w, double* pos, double* quat, double* vel, double* omega, uint64_t n) {
  if (n != w->n) return set_err(w, GRAINS_EINVAL, "particle count mismatch");
  for (uint64_t i = 0; i < n; ++i) {
    if (pos) { pos[3 * i] = w->tr[i].o.x; pos[3 * i + 1] = w->tr[i].o.y; pos[3 * i + 2] = w->tr[i].o.z; }
    if (quat) {
      qt q = quat_from_matrix(&w->tr[i].R);
      quat[4 * i] = q.x; quat[4 * i + 1] = q.y; quat[4 * i + 2] = q.z; quat[4 * i + 3] = q.w;
    }
    if (vel) { vel[3 * i] = w->vel[i].x; vel[3 * i + 1] = w->vel[i].y; vel[3 * i + 2] = w->vel[i].z; }
    if (omega) { omega[3 * i] = w->omega[i].x; omega[3 * i + 1] = w->omega[i].y; omega[3 * i + 2] = w->omega[i].z; }
  }
  return GRAINS_OK;
}

int go_set_transforms(go_world* w, const double* t, uint64_t n) {
  if (n != w->n) return set_err(w, GRAINS_EINVAL, "particle count mismatch");
  for (uint64_t i = 0; i < n; ++i) w->tr[i] = tr_from12(t + 12 * i);
  hist_reset(w);
  return GRAINS_OK;
}

int go_get_transforms(go_world* w, double* t, uint64_t n) {
  if (n != w->n) return set_err(w, GRAINS_EINVAL, "particle count mismatch");
  for (uint64_t i = 0; i < n; ++i) {
    memcpy(t + 12 * i, w->tr[i].R.m, 9 * sizeof(double));
    t[12 * i + 9] = w->tr[i].o.x;
    t[12 * i + 10] = w->tr[i].o.y;
    t[12 * i + 11] = w->tr[i].o.z;
  }
  return GRAINS_OK;
}

static void push(uint32_t** arr, uint32_t* n, uint32_t* cap, uint32_t v) {
  if (*n == *cap) {
    *cap = *cap ? 2 * *cap : 16;
    *arr = (uint32_t*)realloc(*arr, sizeof(uint32_t) * *cap);
  }
  (*arr)[(*n)++] = v;
}

/* minimum-image offset of B relative to A along periodic directions */
static v3 min_image(const osim* S, v3 d) {
  double L[3] = {S->s.domain_max[0] - S->s.domain_min[0], S->s.domain_max[1] - S->s.domain_min[1],
                 S->s.domain_max[2] - S->s.domain_min[2]};
  if (S->s.periodic & 1) d.x -= L[0] * rint(d.x / L[0]);
  if (S->s.periodic & 2) d.y -= L[1] * rint(d.y / L[1]);
  if (S->s.periodic & 4) d.z -= L[2] * rint(d.z / L[2]);
  return d;
}

/* updateLinks (ComponentManagerCPU.cpp:278-296): keys, then residents of each
 * cell in ascending particle index (a stable bucket order). */
static void update_links(go_world* w) {
  memset(w->cell_count, 0, sizeof(uint32_t) * w->ncells_total);
  for (uint64_t i = 0; i < w->n; ++i) {
    const osim* S = &w->sims[w->sim[i]];
    int ood = 0;
    w->keys[i] = S->base + cell_hash1(S->s.domain_min, S->s.domain_max, S->h, S->n3, w->tr[i].o, &ood);
    w->pp_ood[i] = (uint64_t)ood;
    w->cell_count[w->keys[i]]++;
  }
  uint32_t acc = 0;
  for (uint32_t c = 0; c < w->ncells_total; ++c) {
    w->cell_start[c] = acc;
    acc += w->cell_count[c];
  }
  uint32_t* fill = (uint32_t*)calloc(w->ncells_total ? w->ncells_total : 1, sizeof(uint32_t));
  for (uint64_t i = 0; i < w->n; ++i) {
    uint32_t k = w->keys[i];
    w->sorted[w->cell_start[k] + fill[k]++] = (uint32_t)i;
  }
  free(fill);
}

/* detectCollisionAndComputeContactForcesParticles (ComponentManagerCPU.cpp:356-423)
 * for particle i, then ...Obstacles (301-351) for the same particle. */
static void contacts_of(go_world* w, uint64_t i) {
  const uint32_t flags = w->cfg.flags;
  const int quirks = (flags & GRAINS_REF_QUIRKS) != 0;
  const osim* S = &w->sims[w->sim[i]];
  const oshape* A = &w->shapes[w->shape[i]];
  const tr3* trA = &w->tr[i];
  const double massA = A->mass;
  const uint32_t matA = (uint32_t)A->s.material;
  const hooke_t* tab = w->tables + (size_t)S->s.contact_table * (w->nmat * (w->nmat + 1) / 2);
  const uint32_t local = w->keys[i] - S->base;
  const uint32_t* nb = S->nbr + 27 * (size_t)local;
  np_counters cnt = {0, 0, 0};
  w->ncand[i] = 0;
  w->ncont[i] = 0;
  v3 force = w->force[i], torque = w->torque[i];
  for (int c = 0; c < 27; ++c) {
    uint32_t cell = nb[c];
    if (cell == UINT32_MAX) continue;
    cell += S->base;
    for (uint32_t k = 0; k < w->cell_count[cell]; ++k) {
      uint32_t j = w->sorted[w->cell_start[cell] + k];
      if (j == i) continue;
      const oshape* B = &w->shapes[w->shape[j]];
      tr3 trB = w->tr[j];
      v3 d = vsub(trB.o, trA->o);
      if (S->s.periodic) {
        v3 dm = min_image(S, d);
        if (dm.x != d.x || dm.y != d.y || dm.z != d.z) {
          trB.o = vadd(trA->o, dm);
          d = dm;
        }
      }
      /* candidate prefilter: bounding spheres (with a tiny slack) */
      double rs = quirks ? A->rc + B->rc : A->rc_exact + B->rc_exact;
      if (!(vnorm2(d) < rs * rs * (1.0 + PREFILTER_SLACK))) continue;
      push(&w->cand[i], &w->ncand[i], &w->capcand[i], j);
      cinfo ci = closest_points(A, B, trA, &trB, flags, w->cfg.gjk_rel_tol, &cnt);
      if (ci.overlap < 0.0) {
        uint32_t mb = (uint32_t)B->s.material;
        uint32_t s_ = matA < mb ? matA : mb, l_ = matA < mb ? mb : matA;
        const hooke_t* h = &tab[l_ * (l_ + 1) / 2 + s_]; /* computeHash */
        v3 relVel;
        if (quirks)
          relVel = vsub(vadd(w->vel[i], vcross(w->omega[i], ci.cp)),
                        vadd(w->vel[j], vcross(w->omega[j], ci.cp)));
        else
          relVel = vsub(vadd(w->vel[i], vcross(w->omega[i], vsub(ci.cp, trA->o))),
                        vadd(w->vel[j], vcross(w->omega[j], vsub(ci.cp, trB.o))));
        v3 relAng = vsub(w->omega[i], w->omega[j]);
        v3 f, t;
        hooke_forces(h, &ci, relVel, relAng, massA, B->mass, trA->o, &f, &t);
        force = vadd(force, f);
        torque = vadd(torque, t);
        push(&w->cont[i], &w->ncont[i], &w->capcont[i], j);
      }
      w->result[i] += (ci.overlap < 0.0);
    }
  }
  uint64_t walls = 0;
  for (uint32_t o = 0; o < w->nobs; ++o) {
    if (w->obs_sim[o] != w->sim[i]) continue;
    const oshape* B = &w->shapes[w->obs_shape[o]];
    const tr3* trB = &w->obs_tr[o];
    cinfo ci = closest_points(A, B, trA, trB, flags, w->cfg.gjk_rel_tol, &cnt);
    if (ci.overlap < 0.0) {
      uint32_t mb = (uint32_t)B->s.material;
      uint32_t s_ = matA < mb ? matA : mb, l_ = matA < mb ? mb : matA;
      const hooke_t* h = &tab[l_ * (l_ + 1) / 2 + s_];
      v3 relVel = quirks ? vadd(w->vel[i], vcross(w->omega[i], ci.cp))
                         : vadd(w->vel[i], vcross(w->omega[i], vsub(ci.cp, trA->o)));
      /* obstacle kinematics are zero (ComponentManagerCPU.cpp:330-332) */
      relVel = vsub(relVel, V(0, 0, 0));
      v3 relAng = vsub(w->omega[i], V(0, 0, 0));
      v3 f, t;
      hooke_forces(h, &ci, relVel, relAng, massA, B->mass, trA->o, &f, &t);
      force = vadd(force, f);
      torque = vadd(torque, t);
      ++walls;
    }
  }
  /* addExternalForces (ComponentManagerCPU.cpp:448-460) */
  force = vadd(force, vscale(massA, V(S->s.gravity[0], S->s.gravity[1], S->s.gravity[2])));
  w->force[i] = force;
  w->torque[i] = torque;
  w->pp_gjk[i] = cnt.gjk_calls;
  w->pp_iter[i] = cnt.gjk_iters;
  w->pp_deep[i] = cnt.deep;
  w->pp_wall[i] = walls;
}

/* moveParticles (ComponentManagerCPU.cpp:465-500) with
 * RigidBody::computeMomentum (RigidBody.cpp:267-307),
 * FirstOrderExplicit::Move (FirstOrderExplicit.cpp:46-60) or one of the
 * second-order schemes named in TimeIntegratorBuilderFactory.cpp:45-49 (the
 * reference implements none; their updates are stated in grains_hip.h),
 * TimeIntegrator::computeQuaternionChange (TimeIntegrator.cpp:28-44),
 * Transform3::updateTransform (Transform3.cpp:323-330). */
static void move_particle(go_world* w, uint64_t i) {
  const oshape* rb = &w->shapes[w->shape[i]];
  const double dt = w->cfg.dt;
  qt q = quat_from_matrix(&w->tr[i].R);
  qt qc = qconj(q);
  v3 om = w->omega[i];
  v3 angVel = qmult_to_v3(qc, vq(om, q));
  v3 angMom = qmult_to_v3(qc, vq(w->torque[i], q));
  const double* I = rb->inertia;
  const double* Ii = rb->inv;
  v3 tmp = V(I[0] * angVel.x + I[1] * angVel.y + I[2] * angVel.z,
             I[1] * angVel.x + I[3] * angVel.y + I[4] * angVel.z,
             I[2] * angVel.x + I[4] * angVel.y + I[5] * angVel.z);
  angMom = vadd(angMom, vcross(tmp, angVel));
  tmp = V(Ii[0] * angMom.x + Ii[1] * angMom.y + Ii[2] * angMom.z,
          Ii[1] * angMom.x + Ii[3] * angMom.y + Ii[4] * angMom.z,
          Ii[2] * angMom.x + Ii[4] * angMom.y + Ii[5] * angMom.z);
  v3 alpha = qmult_to_v3(q, vq(tmp, qc));
  v3 accel = vdiv(w->force[i], rb->mass);
  w->last_force[i] = w->force[i];
  w->last_torque[i] = w->torque[i];
  w->force[i] = V(0, 0, 0);
  w->torque[i] = V(0, 0, 0);
  /* Move: translation over the step, the angular velocity of the rotation */
  const v3 vel = w->vel[i];
  const int scheme = w->cfg.integrator;
  v3 transMotion, wr;
  v3* h = w->hist + 4 * i;
  if (scheme == GRAINS_SECOND_ORDER_EXPLICIT) {
    transMotion = vadd(vscale(dt, vel), vscale(0.5 * dt * dt, accel));
    wr = vadd(om, vscale(0.5 * dt, alpha));
    w->vel[i] = vadd(vel, vscale(dt, accel));
    w->omega[i] = vadd(om, vscale(dt, alpha));
  } else if (scheme == GRAINS_SECOND_ORDER_LEAPFROG) {
    const double kick = isnan(h[0].x) ? 0.5 * dt : dt; /* the first step: a half kick */
    w->vel[i] = vadd(vel, vscale(kick, accel));
    w->omega[i] = vadd(om, vscale(kick, alpha));
    transMotion = vscale(dt, w->vel[i]);
    wr = w->omega[i];
    h[0].x = 0.0;
  } else if (scheme == GRAINS_SECOND_ORDER_ADAMS_BASHFORTH && !isnan(h[2].x)) {
    transMotion = vscale(dt, vsub(vscale(1.5, vel), vscale(0.5, h[0])));
    wr = vsub(vscale(1.5, om), vscale(0.5, h[1]));
    w->vel[i] = vadd(vel, vscale(dt, vsub(vscale(1.5, accel), vscale(0.5, h[2]))));
    w->omega[i] = vadd(om, vscale(dt, vsub(vscale(1.5, alpha), vscale(0.5, h[3]))));
  } else {
    transMotion = vscale(dt, vel);
    wr = om;
    w->vel[i] = vadd(vel, vscale(dt, accel));
    w->omega[i] = vadd(om, vscale(dt, alpha));
  }
  if (scheme == GRAINS_SECOND_ORDER_ADAMS_BASHFORTH) {
    h[0] = vel;
    h[1] = om;
    h[2] = accel;
    h[3] = alpha;
  }
  qt dq = {0.0, 0.0, 0.0, 1.0};
  double nOmega = vnorm(wr);
  if (nOmega > HIGHEPS) {
    double c = cos(nOmega * dt / 2.0);
    double s = sin(nOmega * dt / 2.0);
    dq = Q(vscale(s / nOmega, wr), c);
  }
  /* updateTransform */
  w->tr[i].o = vadd(w->tr[i].o, transMotion);
  compose_left_by_rotation(&w->tr[i].R, dq);
  /* periodic wrap of the centre (not in the reference) */
  const osim* S = &w->sims[w->sim[i]];
  if (S->s.periodic) {
    double* o = &w->tr[i].o.x;
    for (int k = 0; k < 3; ++k) {
      if (!(S->s.periodic >> k & 1)) continue;
      double L = S->s.domain_max[k] - S->s.domain_min[k];
      if (o[k] < S->s.domain_min[k]) o[k] += L;
      else if (o[k] >= S->s.domain_max[k]) o[k] -= L;
    }
  }
}

int go_step(go_world* w, uint32_t n_steps, int n_threads) {
  if (!w->n || !w->shapes || !w->tables || !w->sims) return set_err(w, GRAINS_ESTATE, "scene not set");
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#else
  (void)n_threads;
#endif
  memset(&w->stats, 0, sizeof(w->stats));
  for (uint32_t step = 0; step < n_steps; ++step) {
    update_links(w);
    const int64_t n = (int64_t)w->n;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) contacts_of(w, (uint64_t)i);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) move_particle(w, (uint64_t)i);
    for (uint64_t i = 0; i < w->n; ++i) {
      w->stats.candidate_pairs += w->ncand[i];
      w->stats.contacts += w->ncont[i];
      w->stats.gjk_calls += w->pp_gjk[i];
      w->stats.gjk_iterations += w->pp_iter[i];
      w->stats.deep_contacts += w->pp_deep[i];
      w->stats.wall_contacts += w->pp_wall[i];
      w->stats.out_of_domain += w->pp_ood[i];
      if (w->ncand[i] > w->stats.max_candidates) w->stats.max_candidates = w->ncand[i];
    }
    w->stats.steps++;
  }
  return GRAINS_OK;
}

int go_get_cell_keys(go_world* w, uint32_t* keys, uint32_t* sorted, uint64_t n) {
  if (n != w->n) return set_err(w, GRAINS_EINVAL, "particle count mismatch");
  if (keys) memcpy(keys, w->keys, sizeof(uint32_t) * n);
  if (sorted) memcpy(sorted, w->sorted, sizeof(uint32_t) * n);
  return GRAINS_OK;
}

int go_get_cell_layout(go_world* w, uint32_t sim, uint32_t* ncell3, double* cell_size,
                       uint32_t* cell_base) {
  if (sim >= w->nsims) return set_err(w, GRAINS_EINVAL, "sim out of range");
  for (int k = 0; k < 3; ++k) ncell3[k] = w->sims[sim].n3[k];
  *cell_size = w->sims[sim].h;
  *cell_base = w->sims[sim].base;
  return GRAINS_OK;
}

int go_get_pairs(go_world* w, int which, uint32_t* ij, uint64_t cap, uint64_t* m) {
  uint64_t k = 0;
  for (uint64_t i = 0; i < w->n; ++i) {
    uint32_t cnt = which == GRAINS_PAIRS_CONTACT ? w->ncont[i] : w->ncand[i];
    const uint32_t* L = which == GRAINS_PAIRS_CONTACT ? w->cont[i] : w->cand[i];
    for (uint32_t a = 0; a < cnt; ++a, ++k)
      if (k < cap) {
        ij[2 * k] = (uint32_t)i;
        ij[2 * k + 1] = L[a];
      }
  }
  *m = k;
  return GRAINS_OK;
}

int go_get_contact_count(go_world* w, int32_t* counts, uint64_t n) {
  if (n != w->n) return set_err(w, GRAINS_EINVAL, "particle count mismatch");
  memcpy(counts, w->result, sizeof(int32_t) * n);
  return GRAINS_OK;
}

int go_get_torce(go_world* w, double* force, double* torque, uint64_t n) {
  if (n != w->n) return set_err(w, GRAINS_EINVAL, "particle count mismatch");
  for (uint64_t i = 0; i < n; ++i) {
    if (force) { force[3 * i] = w->last_force[i].x; force[3 * i + 1] = w->last_force[i].y; force[3 * i + 2] = w->last_force[i].z; }
    if (torque) { torque[3 * i] = w->last_torque[i].x; torque[3 * i + 1] = w->last_torque[i].y; torque[3 * i + 2] = w->last_torque[i].z; }
  }
  return GRAINS_OK;
}

int go_get_stats(go_world* w, grains_stats_t* st) {
  *st = w->stats;
  return GRAINS_OK;
}

/* ---- primitives --------------------------------------------------------------- */
int go_shape_props(const grains_shape_t* s, const double* vertices, double* out) {
  oshape o;
  int rc = shape_derive(&o, s, vertices);
  if (rc) return rc;
  out[0] = o.volume;
  out[1] = o.mass;
  for (int i = 0; i < 6; ++i) { out[2 + i] = o.inertia[i]; out[8 + i] = o.inv[i]; }
  out[14] = o.rc;
  for (int k = 0; k < 3; ++k) out[15 + k] = o.bbox[k];
  out[18] = o.rc_exact;
  free(o.core_verts);
  return GRAINS_OK;
}

int go_support(const grains_shape_t* s, const double* vertices, const double* dir, double* out) {
  oshape o;
  int rc = shape_derive(&o, s, vertices);
  if (rc) return rc;
  v3 p = support(&o, V(dir[0], dir[1], dir[2]), 0);
  out[0] = p.x; out[1] = p.y; out[2] = p.z;
  free(o.core_verts);
  return GRAINS_OK;
}

double go_gjk(const grains_shape_t* a, const grains_shape_t* b, const double* vertices,
              const double* ta, const double* tb, uint32_t flags, double rel_tol, double* pa,
              double* pb, int* iters) {
  oshape A, B;
  if (shape_derive(&A, a, vertices) || shape_derive(&B, b, vertices)) return NAN;
  /* distance of the full shapes: run with core supports equal to the shape */
  memcpy(A.core_dims, A.gd, sizeof(A.core_dims));
  memcpy(B.core_dims, B.gd, sizeof(B.core_dims));
  if (A.core_verts) memcpy(A.core_verts, A.verts, sizeof(double) * 3 * A.nv);
  if (B.core_verts) memcpy(B.core_verts, B.verts, sizeof(double) * 3 * B.nv);
  tr3 TA = tr_from12(ta), TB = tr_from12(tb);
  v3 p, q;
  int it = 0, sep = 0;
  double d = (flags & GRAINS_REF_QUIRKS) ? gjk_jh(&A, &B, &TA, &TB, rel_tol, &p, &q, &it)
                                         : gjk_sv(&A, &B, &TA, &TB, flags, rel_tol, INFINITY, &p, &q, &it,
                                                  NULL, &sep);
  p = tapply(&TA, p);
  q = tapply(&TB, q);
  if (pa) { pa[0] = p.x; pa[1] = p.y; pa[2] = p.z; }
  if (pb) { pb[0] = q.x; pb[1] = q.y; pb[2] = q.z; }
  if (iters) *iters = it;
  free(A.core_verts);
  free(B.core_verts);
  return d;
}

int go_closest_points(const grains_shape_t* a, const grains_shape_t* b, const double* vertices,
                      const double* ta, const double* tb, uint32_t flags, double rel_tol, double* out) {
  oshape A, B;
  if (shape_derive(&A, a, vertices) || shape_derive(&B, b, vertices)) return -1;
  tr3 TA = tr_from12(ta), TB = tr_from12(tb);
  cinfo c = closest_points(&A, &B, &TA, &TB, flags, rel_tol, NULL);
  out[0] = c.cp.x; out[1] = c.cp.y; out[2] = c.cp.z;
  out[3] = c.cv.x; out[4] = c.cv.y; out[5] = c.cv.z;
  out[6] = c.overlap;
  free(A.core_verts);
  free(B.core_verts);
  return c.overlap < 0.0;
}

void go_hooke(const grains_hooke_t* p, const double* ci, const double* rel_vel,
              const double* rel_ang_vel, double m1, double m2, const double* origin_a,
              double* out) {
  hooke_t h = {p->kn, p->en, p->etat, p->muc, p->kr, 0};
  h.muen = log(h.en) / sqrt(G_PI * G_PI + log(h.en) * log(h.en));
  cinfo c = {V(ci[0], ci[1], ci[2]), V(ci[3], ci[4], ci[5]), ci[6]};
  v3 f, t;
  hooke_forces(&h, &c, V(rel_vel[0], rel_vel[1], rel_vel[2]),
               V(rel_ang_vel[0], rel_ang_vel[1], rel_ang_vel[2]), m1, m2,
               V(origin_a[0], origin_a[1], origin_a[2]), &f, &t);
  out[0] = f.x; out[1] = f.y; out[2] = f.z;
  out[3] = t.x; out[4] = t.y; out[5] = t.z;
}

void go_quat_from_matrix(const double* m9, double* q) {
  m3 m;
  memcpy(m.m, m9, 9 * sizeof(double));
  qt r = quat_from_matrix(&m);
  q[0] = r.x; q[1] = r.y; q[2] = r.z; q[3] = r.w;
}

void go_matrix_from_quat(const double* q, double* m9) {
  qt r = {q[0], q[1], q[2], q[3]};
  m3 m = matrix_from_quat(r);
  memcpy(m9, m.m, 9 * sizeof(double));
}

/* Transform3::setBasis(aX, aY, aZ) (Transform3.cpp:151-162) */
void go_basis_from_angles(double aX, double aY, double aZ, double* m9) {
  double v[9] = {cos(aZ) * cos(aY),
                 cos(aZ) * sin(aY) * sin(aX) - sin(aZ) * cos(aX),
                 cos(aZ) * sin(aY) * cos(aX) + sin(aZ) * sin(aX),
                 sin(aZ) * cos(aY),
                 sin(aZ) * sin(aY) * sin(aX) + cos(aZ) * cos(aX),
                 sin(aZ) * sin(aY) * cos(aX) - cos(aZ) * sin(aX),
                 -sin(aY),
                 cos(aY) * sin(aX),
                 cos(aY) * cos(aX)};
  memcpy(m9, v, sizeof(v));
}
