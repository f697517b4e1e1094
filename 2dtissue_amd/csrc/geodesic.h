// Front updates of the geodesic distance table (local fast marching on the triangle mesh),
// shared by the host march (mesh.cpp, built by the host compiler) and the GPU march
// (mesh_gpu.hip): one source text, compiled without contraction on both sides, so both round
// every operation alike (IEEE add, multiply, divide and square root).
#pragma once
#include <cmath>

#if defined(__HIPCC__)
#define TISSUE_GEO_HD __host__ __device__
#else
#define TISSUE_GEO_HD
#endif

namespace tissue {
namespace geo {

TISSUE_GEO_HD inline double inf() { return __builtin_inf(); }
TISSUE_GEO_HD inline void sub3(const double* a, const double* b, double* o) {
  o[0] = a[0] - b[0];
  o[1] = a[1] - b[1];
  o[2] = a[2] - b[2];
}
TISSUE_GEO_HD inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
TISSUE_GEO_HD inline double norm3(const double* a) { return sqrt(dot3(a, a)); }

// first-order fast-marching update of vertex C from triangle (A, B, C) with known arrival
// times tA, tB (Kimmel & Sethian 1998, written for 3D edge vectors): the planar front
// T(P) = tC + g.(P - C), |g| = 1, through A and B; accepted only when the characteristic
// reaching C enters the triangle through edge AB.
TISSUE_GEO_HD inline double tri_update(const double* A, double tA, const double* B, double tB, const double* C) {
  double e1[3], e2[3];
  sub3(A, C, e1);
  sub3(B, C, e2);
  const double m00 = dot3(e1, e1), m01 = dot3(e1, e2), m11 = dot3(e2, e2);
  const double det = m00 * m11 - m01 * m01;
  if (!(det > 1e-300)) return inf();
  const double i00 = m11 / det, i01 = -m01 / det, i11 = m00 / det;
  const double a = i00 + 2 * i01 + i11;
  const double b = (i00 + i01) * tA + (i01 + i11) * tB;
  const double c = i00 * tA * tA + 2 * i01 * tA * tB + i11 * tB * tB - 1.0;
  const double disc = b * b - a * c;
  if (!(a > 0) || disc < 0) return inf();
  const double T = (b + sqrt(disc)) / a;
  const double r0 = tA - T, r1 = tB - T;
  const double c0 = i00 * r0 + i01 * r1, c1 = i01 * r0 + i11 * r1;
  const double tol = 1e-12;
  const double tmax = tA < tB ? tB : tA;
  if (c0 > tol || c1 > tol || T < tmax) return inf();
  return T;
}

// length of the straight path from V to X across edge (A, B) once the two faces (V,A,B) and
// (A,B,X) are unfolded into one plane; infinity when that segment misses the edge
TISSUE_GEO_HD inline double unfold_across(const double* V, const double* A, const double* B, const double* X) {
  double ab[3], av[3], ax[3];
  sub3(B, A, ab);
  sub3(V, A, av);
  sub3(X, A, ax);
  const double L = norm3(ab);
  if (!(L > 0)) return inf();
  const double vu = dot3(av, ab) / L, xu = dot3(ax, ab) / L;
  const double hv = dot3(av, av) - vu * vu, hx = dot3(ax, ax) - xu * xu;
  const double vh = sqrt(0.0 < hv ? hv : 0.0);
  const double xh = sqrt(0.0 < hx ? hx : 0.0);
  if (!(vh + xh > 0)) return inf();
  const double t = vu + (xu - vu) * vh / (vh + xh);  // crossing point along AB
  if (t < 0 || t > L) return inf();
  return sqrt((xu - vu) * (xu - vu) + (xh + vh) * (xh + vh));
}

}  // namespace geo
}  // namespace tissue
