// Q9 element geometry and PML stretch shared by the assembly (K1) and adjoint (K12) kernels.
//
// Restates SawSim saw2d/piezo_fem.py: 3x3 Gauss rule (:48-53), biquadratic shapes in Gmsh
// QUADS9 order (:56-103), signed detJ and physical gradients (:248-272), the centre-node
// PML stretch frozen at omega_top (:222-242) and the multiplier codes of T9 (:127-142).
#pragma once
#include "common.h"

namespace q9 {

__constant__ static double kGaussPt[3] = {-0.774596669241483, 0.0, 0.774596669241483};
__constant__ static double kGaussW[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
// node -> 1D Lagrange index in xi and eta (0: -1, 1: 0, 2: +1)
__constant__ static int kIa[9] = {0, 2, 2, 0, 1, 2, 1, 0, 1};
__constant__ static int kJa[9] = {0, 0, 2, 2, 0, 1, 2, 1, 1};
// strain row of displacement component A for derivative direction d (9-component strain)
__constant__ static int kStrainRow[3][2] = {{0, 5}, {8, 1}, {7, 6}};
// PML multiplier codes: 0 A = a2 a3/a1, 1 b = a3, 2 c = a2, 3 D = a1 a3/a2, 4 e = a1, 5 F = a1 a2/a3
__constant__ static int kT9Code[9][9] = {
    {0, 1, 2, 2, 2, 1, 1, 0, 0}, {1, 3, 4, 4, 4, 3, 3, 1, 1}, {2, 4, 5, 5, 5, 4, 4, 2, 2},
    {2, 4, 5, 5, 5, 4, 4, 2, 2}, {2, 4, 5, 5, 5, 4, 4, 2, 2}, {1, 3, 4, 4, 4, 3, 3, 1, 1},
    {1, 3, 4, 4, 4, 3, 3, 1, 1}, {0, 1, 2, 2, 2, 1, 1, 0, 0}, {0, 1, 2, 2, 2, 1, 1, 0, 0}};
// dielectric multipliers [[A, b, c], [b, D, e], [c, e, F]] restricted to x, y
__constant__ static int kTepsCode[2][2] = {{0, 1}, {1, 3}};

struct PmlRegion {
  int tag, kind;   // kind bit 0: x stretch, bit 1: y stretch
  double dmax, n, xa, xp, ya, yp;
};
struct PmlSet { PmlRegion r[8]; };

__device__ inline double lag(int i, double t) {
  return i == 0 ? 0.5 * t * (t - 1.0) : (i == 1 ? 1.0 - t * t : 0.5 * t * (t + 1.0));
}
__device__ inline double dlag(int i, double t) { return i == 0 ? t - 0.5 : (i == 1 ? -2.0 * t : t + 0.5); }

// Gauss point q: shape values N[a], physical gradients dNx[a][d], weight w * detJ (signed).
__device__ inline void gauss_point(int q, const double* sx, const double* sy, double* N, double (*dNx)[2],
                                   double* dV) {
  const double xi = kGaussPt[q % 3], eta = kGaussPt[q / 3];
  const double w = kGaussW[q % 3] * kGaussW[q / 3];
  double dr[9][2];
  double J00 = 0, J01 = 0, J10 = 0, J11 = 0;
  for (int a = 0; a < 9; ++a) {
    const int ia = kIa[a], ja = kJa[a];
    N[a] = lag(ia, xi) * lag(ja, eta);
    dr[a][0] = dlag(ia, xi) * lag(ja, eta);
    dr[a][1] = lag(ia, xi) * dlag(ja, eta);
    J00 += sx[a] * dr[a][0]; J01 += sx[a] * dr[a][1];
    J10 += sy[a] * dr[a][0]; J11 += sy[a] * dr[a][1];
  }
  const double det = J00 * J11 - J01 * J10;
  const double i00 = J11 / det, i01 = -J01 / det, i10 = -J10 / det, i11 = J00 / det;
  for (int a = 0; a < 9; ++a) {
    dNx[a][0] = dr[a][0] * i00 + dr[a][1] * i10;
    dNx[a][1] = dr[a][0] * i01 + dr[a][1] * i11;
  }
  *dV = w * det;
}

// Stretch factors a1, a2, a3 of an element from its centre node coordinates.
__device__ inline void pml_alpha(const PmlSet& ps, int npml, int tag, double xc, double yc, double omega_top,
                                 double2* a) {
  a[0] = make_double2(1, 0); a[1] = make_double2(1, 0); a[2] = make_double2(1, 0);
  for (int r = 0; r < npml; ++r) {
    const PmlRegion& R = ps.r[r];
    if (R.tag != tag) continue;
    if (R.kind & 1) {
      double ratio = (xc - R.xp) * (xc - R.xp) / ((R.xa - R.xp) * (R.xa - R.xp));
      ratio = fmin(fmax(ratio, 0.0), 1.0);
      a[0] = make_double2(1.0, -R.dmax * pow(1.0 - ratio, R.n) / omega_top);   // 1 + d / (i w)
    }
    if (R.kind & 2) {
      double ratio = (yc - R.yp) * (yc - R.yp) / ((R.ya - R.yp) * (R.ya - R.yp));
      ratio = fmin(fmax(ratio, 0.0), 1.0);
      a[1] = make_double2(1.0, -R.dmax * pow(1.0 - ratio, R.n) / omega_top);
    }
  }
}

__device__ inline void pml_codes(const double2* a, double2* code) {
  code[0] = cdiv(cmul(a[1], a[2]), a[0]);
  code[1] = a[2];
  code[2] = a[1];
  code[3] = cdiv(cmul(a[0], a[2]), a[1]);
  code[4] = a[0];
  code[5] = cdiv(cmul(a[0], a[1]), a[2]);
}

inline PmlSet make_pml_set(int npml, const int* tag, const int* kind, const double* par) {
  PmlSet ps;
  for (int r = 0; r < npml; ++r) {
    ps.r[r].tag = tag[r]; ps.r[r].kind = kind[r];
    ps.r[r].dmax = par[6 * r + 0]; ps.r[r].n = par[6 * r + 1];
    ps.r[r].xa = par[6 * r + 2]; ps.r[r].xp = par[6 * r + 3];
    ps.r[r].ya = par[6 * r + 4]; ps.r[r].yp = par[6 * r + 5];
  }
  return ps;
}

}  // namespace q9
