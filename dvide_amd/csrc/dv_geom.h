// Geometry of the contact step shared by the host and the gfx950 kernels:
// length/normalize in the two rounding modes, the narrowphase keep test and
// contact record (CollisionDetector.cu:284-328, 432-498), the contact tangent
// frame (System.cu:753-762) and the Coulomb-cone projection (APGD.cu:47-78).
//
// Everything here is written so that the same source compiled for the host
// (g++) and for gfx950 (hipcc), both with -ffp-contract=off, performs the
// same IEEE operations in the same order: no fused multiply-adds, no
// approximate reciprocals, sqrt and division correctly rounded.  The contact
// set, normals, gaps and frames are therefore bit-identical on both sides.
#pragma once

#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define DV_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define DV_HD inline
#endif

namespace dv {

struct d3 {
  double x, y, z;
};
DV_HD d3 mk3(double x, double y, double z) {
  d3 r;
  r.x = x; r.y = y; r.z = z;
  return r;
}
DV_HD d3 add3(d3 a, d3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
DV_HD d3 sub3(d3 a, d3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
DV_HD d3 neg3(d3 a) { return mk3(-a.x, -a.y, -a.z); }
DV_HD d3 scl3(d3 a, double s) { return mk3(a.x * s, a.y * s, a.z * s); }
DV_HD double dot3(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
DV_HD d3 cross3(d3 a, d3 b) {
  return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

DV_HD double dsqrt(double x) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_sqrt(x);
#else
  return std::sqrt(x);
#endif
}

// Geometry rounding mode.  1: the reference's double3 length/normalize call the
// binary32 sqrtf/rsqrtf under nvcc (helper_math_double.h:576-598), so lengths
// and unit normals carry binary32 rounding; restated with correctly rounded
// binary32 sqrt and reciprocal, each obtained from a double operation whose
// rounding to binary32 is exact-rounding-safe (53 >= 2*24+2).  0: plain FP64.
DV_HD double lengthOf(d3 v, int mode) {
  double d = dot3(v, v);
  if (mode == 1) return (double)(float)dsqrt((double)(float)d);
  return dsqrt(d);
}
DV_HD double invLengthOf(d3 v, int mode) {
  double d = dot3(v, v);
  if (mode == 1) {
    float s = (float)dsqrt((double)(float)d);
    return (double)(float)(1.0 / (double)s);
  }
  return 1.0 / dsqrt(d);
}
DV_HD d3 normalize3(d3 v, int mode) { return scl3(v, invLengthOf(v, mode)); }

// Narrowphase for the ordered pair (A, B) with A < B.  gA/gB are the stored
// geometries ((r,0,0) sphere, half extents for a box).  Returns 1 and fills
// the record when the pair is kept (penetration >= -envelope); box-box pairs
// are never kept (CollisionDetector.cu:287-328).  The stored normal points
// from B to A and w = -penetration is the gap (CollisionDetector.cu:497).
DV_HD int narrowphase(d3 posA, d3 gA, d3 posB, d3 gB, double envelope, int mode, d3* nOut,
                      double* wOut, d3* locOut) {
  d3 normal = mk3(1, 0, 0), location = mk3(0, 0, 0);
  double penetration;
  if (gA.y == 0 && gB.y == 0) {
    penetration = (gA.x + gB.x) - lengthOf(sub3(posB, posA), mode);
    if (!(penetration >= -envelope)) return 0;
    if (nOut == 0) return 1;
    normal = normalize3(sub3(posB, posA), mode);
    location = add3(posA, scl3(normal, gA.x));
  } else if ((gA.y != 0 && gB.y == 0) || (gA.y == 0 && gB.y != 0)) {
    double dmin = 0, rr = gB.x;
    d3 c = posB, bmin = sub3(posA, gA), bmax = add3(posA, gA);
    const bool sphereA = (gA.y == 0 && gB.y != 0);
    if (sphereA) {
      rr = gA.x;
      c = posA;
      bmin = sub3(posB, gB);
      bmax = add3(posB, gB);
    }
    d3 nn = mk3(0, 0, 0);
    if (c.x < bmin.x) { dmin += (c.x - bmin.x) * (c.x - bmin.x); nn.x += c.x - bmin.x; }
    else if (c.x > bmax.x) { dmin += (c.x - bmax.x) * (c.x - bmax.x); nn.x += c.x - bmax.x; }
    if (c.y < bmin.y) { dmin += (c.y - bmin.y) * (c.y - bmin.y); nn.y += c.y - bmin.y; }
    else if (c.y > bmax.y) { dmin += (c.y - bmax.y) * (c.y - bmax.y); nn.y += c.y - bmax.y; }
    if (c.z < bmin.z) { dmin += (c.z - bmin.z) * (c.z - bmin.z); nn.z += c.z - bmin.z; }
    else if (c.z > bmax.z) { dmin += (c.z - bmax.z) * (c.z - bmax.z); nn.z += c.z - bmax.z; }
    penetration = rr - dsqrt(dmin);
    if (!(penetration >= -envelope)) return 0;
    if (nOut == 0) return 1;
    normal = normalize3(nn, mode);
    location = sub3(c, scl3(normal, rr));
    if (sphereA) normal = neg3(normal);
  } else {
    return 0;
  }
  *nOut = neg3(normal);
  *wOut = -penetration;
  *locOut = location;
  return 1;
}

// Tangent frame (u, v) of a contact normal n (System.cu:753-762).
DV_HD void contactFrame(d3 n, int mode, d3* u, d3* v) {
  d3 uu;
  if (n.z != 0) uu = normalize3(mk3(1, 0, -n.x / n.z), mode);
  else if (n.x != 0) uu = normalize3(mk3(-n.z / n.x, 0, 1), mode);
  else uu = normalize3(mk3(1, -n.x / n.y, 0), mode);
  *u = uu;
  *v = normalize3(cross3(n, uu), mode);
}

// Projection of (gn, gu, gv) onto the Coulomb cone of friction mu, in place
// (APGD.cu:47-78, including its special cases: mu = 0, the polar cone, and
// |gamma_n| < 1e-14).  Arithmetic in T; the FP64 instantiation rounds like
// the reference.
template <typename T>
DV_HD void projectCone(T& g0, T& g1, T& g2, T mu) {
  T gn = g0;
#ifdef __HIP_DEVICE_COMPILE__
  T gt = (T)__builtin_sqrt((double)(g1 * g1 + g2 * g2));
#else
  T gt = (T)std::sqrt((double)(g1 * g1 + g2 * g2));
#endif
  if (mu == 0) {
    g1 = 0; g2 = 0;
    if (gn < 0) g0 = 0;
  } else if (gt < mu * gn) {
    // inside the cone
  } else if ((gt < -((T)1 / mu) * gn) || ((gn < 0 ? -gn : gn) < (T)10e-15)) {
    g0 = 0; g1 = 0; g2 = 0;
  } else {
    T gnp = (gt * mu + gn) / (mu * mu + (T)1);
    T gtp = gnp * mu;
    T q = gtp / gt;
    g0 = gnp;
    g1 = q * g1;
    g2 = q * g2;
  }
}

// Shape functions of the 4-node ANCF shell element (System.cu:904-961,
// CollisionDetector.cu:50-69) at (xi, eta) in [0,1]^2 for sides a (along xi)
// and b (along eta).  S[3q], S[3q+1], S[3q+2] weight node q's position,
// x-gradient and y-gradient; node q = 0..3 sits at corner (0,0), (1,0),
// (1,1), (0,1).  With u = |xi - xi_q| and w = |eta - eta_q| they are
// (1-u)(1-w)(1+u+w-2u^2-2w^2), +-a u(1-u)^2 (1-w) and +-b w(1-w)^2 (1-u), the
// sign negative where u or w runs against xi or eta.
DV_HD void shellShape(double xi, double eta, double a, double b, double S[12]) {
  for (int q = 0; q < 4; ++q) {
    const bool fx = q == 1 || q == 2, fy = q >= 2;
    const double u = fx ? 1.0 - xi : xi, w = fy ? 1.0 - eta : eta;
    const double ou = 1.0 - u, ow = 1.0 - w;
    S[3 * q] = (ou * ow) * (((1.0 + u) + w) - 2.0 * (u * u) - 2.0 * (w * w));
    S[3 * q + 1] = (fx ? -a : a) * ((u * (ou * ou)) * ow);
    S[3 * q + 2] = (fy ? -b : b) * ((w * (ow * ow)) * ou);
  }
}

}  // namespace dv
