// K1' Hex27 element assembly and the 27-node K/M/D gather.
//
// Restates the element operators of SawSim saw2d/saw3d/assembly.py (form_stiffness_mass_3d):
//   * 27-point Gauss rule with the rounded points/weights 0.774596669241483,
//     0.555555555555556 and 0.888888888888889 (elements.py:22-25, 63-76), xi1 fastest;
//   * tensor-product quadratic shapes in the reference H27 node order (elements.py:29-42,
//     79-105): node i sits at natural coordinates kH27[i];
//   * Jacobian, its inverse and |detJ| weights (assembly.py:31-70);
//   * strain rows of B with gamma_yz = duy/dz only (assembly.py:81-101, the quirk at :92);
//   * PML stretch factors from the element's centre node (local node 26) frozen at omega_top
//     and the per-Voigt multipliers of C, e and eps, rho a1 a2 a3 (assembly.py:127-213);
//   * blocks K_uu = KM_SCALE B^T C B, K_uphi = B^T e^T G, K_phiu = KM_SCALE conj(K_uphi)^T,
//     K_phiphi = -G^T eps G, M = KM_SCALE rho N N^T on each displacement component
//     (assembly.py:288-308, 311-327).
// The material of an element is constant over it (the stretch is evaluated once per element),
// so every block is a contraction of nine real Gram matrices
//     G_dd'(a, b) = sum_q w_q |detJ_q| dN_a/dx_d dN_b/dx_d'
// with complex 3x3 (K_uu), 3-vector (K_uphi) and scalar (K_phiphi) material coefficients.
// One workgroup per element; thread t owns the node pairs t, t + 256, t + 512 (< 729).
// Output: Ke[e][A * 4 + B][a * 27 + b] (complex128, A, B = ux, uy, uz, phi), Me[e][a * 27 + b].
#include "common.h"

namespace {

__constant__ double kG3[3] = {-0.774596669241483, 0.0, 0.774596669241483};
__constant__ double kW3[3] = {0.555555555555556, 0.888888888888889, 0.555555555555556};
// natural coordinates of the 27 local nodes (reference H27 order)
__constant__ signed char kH27[27][3] = {
    {1, 1, 1}, {1, -1, 1}, {-1, -1, 1}, {-1, 1, 1}, {1, 1, -1}, {1, -1, -1}, {-1, -1, -1}, {-1, 1, -1},
    {1, 0, 1}, {0, 1, 1}, {1, 1, 0}, {0, -1, 1}, {1, -1, 0}, {-1, 0, 1}, {-1, -1, 0}, {-1, 1, 0},
    {1, 0, -1}, {0, 1, -1}, {0, -1, -1}, {-1, 0, -1}, {0, 0, 1}, {1, 0, 0}, {0, 1, 0}, {0, -1, 0},
    {-1, 0, 0}, {0, 0, -1}, {0, 0, 0}};
// Voigt strain row fed by derivative d of displacement component A (-1: none; the uz, y entry
// is the dropped duz/dy of gamma_yz)
__constant__ int kRow[3][3] = {{0, 5, 4}, {5, 1, 3}, {4, -1, 2}};
// stretch multiplier codes: 0 a2a3/a1, 1 a1a3/a2, 2 a1a2/a3, 3 a1, 4 a2, 5 a3
__constant__ int kSC[6][6] = {{0, 5, 4, 5, 0, 0}, {5, 1, 3, 1, 5, 5}, {4, 3, 2, 3, 4, 4},
                              {4, 3, 2, 3, 4, 4}, {4, 3, 2, 3, 4, 4}, {5, 1, 3, 1, 5, 5}};
__constant__ int kSE[3][6] = {{0, 5, 4, 4, 4, 5}, {5, 1, 3, 3, 3, 1}, {4, 3, 2, 2, 2, 3}};
__constant__ int kSP[3][3] = {{0, 5, 4}, {5, 1, 3}, {4, 3, 2}};

struct Pml3 {
  int kind;              // bit 0: stretch x (a1), bit 1: stretch z (a3)
  double dmax, n, xa, xp, za, zp;
};
struct Pml3Set { Pml3 r[8]; };

// 1D quadratic Lagrange factor on nodes -1, 0, 1 and its derivative, node coordinate c
__device__ inline double lag1(int c, double t) {
  return c < 0 ? 0.5 * t * (t - 1.0) : (c == 0 ? 1.0 - t * t : 0.5 * t * (t + 1.0));
}
__device__ inline double dlag1(int c, double t) { return c < 0 ? t - 0.5 : (c == 0 ? -2.0 * t : t + 0.5); }

// Natural derivatives of shape a at t (dn[d] = dN_a / dxi_d) and its value.
__device__ inline double h27_dn(int a, const double* t, double* dn) {
  const int c0 = kH27[a][0], c1 = kH27[a][1], c2 = kH27[a][2];
  const double l0 = lag1(c0, t[0]), l1 = lag1(c1, t[1]), l2 = lag1(c2, t[2]);
  dn[0] = dlag1(c0, t[0]) * l1 * l2;
  dn[1] = l0 * dlag1(c1, t[1]) * l2;
  dn[2] = l0 * l1 * dlag1(c2, t[2]);
  return l0 * l1 * l2;
}

// Gauss point q: shapes N[a], physical gradients g[a][d], weight |detJ| w.  The natural
// derivatives are evaluated twice (Jacobian pass, gradient pass) instead of being kept.
__device__ inline void h27_point(int q, const double (*X)[3], double* N, double (*g)[3], double* wq) {
  const double t[3] = {kG3[q % 3], kG3[(q / 3) % 3], kG3[q / 9]};
  const double w = kW3[q % 3] * kW3[(q / 3) % 3] * kW3[q / 9];
  double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll 1
  for (int a = 0; a < 27; ++a) {
    double dn[3];
    N[a] = h27_dn(a, t, dn);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) J[i][j] += dn[i] * X[a][j];     // J[i][j] = d x_j / d xi_i
  }
  const double det = J[0][0] * (J[1][1] * J[2][2] - J[2][1] * J[1][2]) -
                     J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                     J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  double Ji[3][3];
  Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
  Ji[0][1] = -(J[0][1] * J[2][2] - J[0][2] * J[2][1]) / det;
  Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
  Ji[1][0] = -(J[1][0] * J[2][2] - J[1][2] * J[2][0]) / det;
  Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
  Ji[1][2] = -(J[0][0] * J[1][2] - J[0][2] * J[1][0]) / det;
  Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
  Ji[2][1] = -(J[0][0] * J[2][1] - J[0][1] * J[2][0]) / det;
  Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
#pragma unroll 1
  for (int a = 0; a < 27; ++a) {
    double dn[3];
    h27_dn(a, t, dn);
    for (int d = 0; d < 3; ++d) g[a][d] = Ji[d][0] * dn[0] + Ji[d][1] * dn[1] + Ji[d][2] * dn[2];
  }
  *wq = fabs(det) * w;
}

// Stretch codes of an element: 0 a2a3/a1, 1 a1a3/a2, 2 a1a2/a3, 3 a1, 4 a2, 5 a3, 6 a1a2a3.
__device__ inline void h27_codes(const Pml3Set& ps, int r, double xc, double zc, double omega_top, double2* code) {
  double2 a1 = make_double2(1, 0), a2 = make_double2(1, 0), a3 = make_double2(1, 0);
  if (r >= 0) {
    const Pml3& P = ps.r[r];
    if (P.kind & 1) {
      const double d = P.dmax * pow(1.0 - (xc - P.xp) * (xc - P.xp) / ((P.xa - P.xp) * (P.xa - P.xp)), P.n);
      a1 = make_double2(1.0, -d / omega_top);             // 1 + d / (i w)
    }
    if (P.kind & 2) {
      const double d = P.dmax * pow(1.0 - (zc - P.zp) * (zc - P.zp) / ((P.za - P.zp) * (P.za - P.zp)), P.n);
      a3 = make_double2(1.0, -d / omega_top);
    }
  }
  code[0] = cdiv(cmul(a2, a3), a1);
  code[1] = cdiv(cmul(a1, a3), a2);
  code[2] = cdiv(cmul(a1, a2), a3);
  code[3] = a1;
  code[4] = a2;
  code[5] = a3;
  code[6] = cmul(cmul(a1, a2), a3);
}

__device__ inline double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }

__global__ void __launch_bounds__(256) k_elem_h27(
    int ne, const int* __restrict__ conn, const double* __restrict__ xyz, const int* __restrict__ mat_idx,
    const int* __restrict__ pml_idx, const double2* __restrict__ Cm, const double2* __restrict__ Em,
    const double2* __restrict__ Pm, const double* __restrict__ RHO, const Pml3Set pset, double omega_top,
    double scale, double2* __restrict__ Ke, double2* __restrict__ Me) {
  const int e = blockIdx.x;
  if (e >= ne) return;
  const int t = threadIdx.x;
  __shared__ double X[27][3];
  __shared__ double sg[27][27][3];        // [q][a][d]
  __shared__ double sN[27][27];           // [q][a]
  __shared__ double sw[27];
  __shared__ double2 code[7];
  __shared__ double2 Cs[6][6], Es[3][6], Ps[3][3];
  __shared__ double2 rho_s;
  if (t < 81) {
    const int a = t / 3, d = t % 3;
    X[a][d] = xyz[3 * (int64_t)conn[(int64_t)e * 27 + a] + d];
  }
  __syncthreads();
  if (t < 27) h27_point(t, X, sN[t], sg[t], &sw[t]);
  if (t == 64) h27_codes(pset, pml_idx[e], X[26][0], X[26][2], omega_top, code);
  __syncthreads();
  const int m = mat_idx[e];
  if (t < 36) {
    const int s = t / 6, u = t % 6;
    Cs[s][u] = cmul(code[kSC[s][u]], Cm[m * 36 + t]);
  } else if (t < 54) {
    const int k = (t - 36) / 6, s = (t - 36) % 6;
    Es[k][s] = cmul(code[kSE[k][s]], Em[m * 18 + (t - 36)]);
  } else if (t < 63) {
    const int j = (t - 54) / 3, k = (t - 54) % 3;
    Ps[j][k] = cmul(code[kSP[j][k]], Pm[m * 9 + (t - 54)]);
  } else if (t == 63) {
    rho_s = cscale(code[6], RHO[m]);
  }
  __syncthreads();
  double2* Kout = Ke + (size_t)e * 16 * 729;
  for (int p = t; p < 729; p += 256) {
    const int a = p / 27, b = p % 27;
    double G[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double GM = 0;
#pragma unroll 3
    for (int q = 0; q < 27; ++q) {
      const double w = sw[q];
      const double ga0 = w * sg[q][a][0], ga1 = w * sg[q][a][1], ga2 = w * sg[q][a][2];
      const double gb0 = sg[q][b][0], gb1 = sg[q][b][1], gb2 = sg[q][b][2];
      G[0][0] += ga0 * gb0; G[0][1] += ga0 * gb1; G[0][2] += ga0 * gb2;
      G[1][0] += ga1 * gb0; G[1][1] += ga1 * gb1; G[1][2] += ga1 * gb2;
      G[2][0] += ga2 * gb0; G[2][1] += ga2 * gb1; G[2][2] += ga2 * gb2;
      GM += w * sN[q][a] * sN[q][b];
    }
    // K_uu[A][B] = SCALE sum_{d, d'} C[row(A, d)][row(B, d')] G[d][d']
    for (int A = 0; A < 3; ++A)
      for (int B = 0; B < 3; ++B) {
        double2 acc = make_double2(0, 0);
        for (int d = 0; d < 3; ++d) {
          const int s = kRow[A][d];
          if (s < 0) continue;
          for (int d2 = 0; d2 < 3; ++d2) {
            const int u = kRow[B][d2];
            if (u < 0) continue;
            acc.x += Cs[s][u].x * G[d][d2];
            acc.y += Cs[s][u].y * G[d][d2];
          }
        }
        Kout[(A * 4 + B) * 729 + p] = cscale(acc, scale);
      }
    // K_uphi[A](a, b) = sum_{d, k} e[k][row(A, d)] G[d][k];  K_phiu[A](a, b) = SCALE conj(K_uphi[A](b, a))
    for (int A = 0; A < 3; ++A) {
      double2 up = make_double2(0, 0), pu = make_double2(0, 0);
      for (int d = 0; d < 3; ++d) {
        const int s = kRow[A][d];
        if (s < 0) continue;
        for (int k = 0; k < 3; ++k) {
          up.x += Es[k][s].x * G[d][k]; up.y += Es[k][s].y * G[d][k];
          pu.x += Es[k][s].x * G[k][d]; pu.y += Es[k][s].y * G[k][d];
        }
      }
      Kout[(A * 4 + 3) * 729 + p] = up;
      Kout[(3 * 4 + A) * 729 + p] = make_double2(scale * pu.x, -scale * pu.y);
    }
    double2 pp = make_double2(0, 0);
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) { pp.x += Ps[j][k].x * G[j][k]; pp.y += Ps[j][k].y * G[j][k]; }
    Kout[15 * 729 + p] = make_double2(-pp.x, -pp.y);
    Me[(size_t)e * 729 + p] = cscale(rho_s, GM * scale);
  }
}

// Global K, M, D of the 27-node pattern, gathered per node pair k and block (A, B):
// slot = A*4*nnzn + 4*node_ptr[i] + B*deg[i] + (k - node_ptr[i]),  i = pair_row[k];
// gather codes e * 729 + i * 27 + j in list order.  D collects the K_uu part of the elements
// flagged in dmask; M fills the diagonal displacement blocks.
__global__ void k_gather_h27(int nnzn, const int* __restrict__ pair_row, const int64_t* __restrict__ node_ptr,
                             const int64_t* __restrict__ gptr, const int64_t* __restrict__ gidx,
                             const unsigned char* __restrict__ dmask, const double2* __restrict__ Ke,
                             const double2* __restrict__ Me, double2* __restrict__ K, double2* __restrict__ M,
                             double2* __restrict__ D) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnzn) return;
  const int AB = blockIdx.y;
  const int A = AB / 4, B = AB % 4;
  const int i = pair_row[k];
  const int64_t np = node_ptr[i];
  const int64_t deg = node_ptr[i + 1] - np;
  const int64_t slot = (int64_t)A * 4 * nnzn + 4 * np + (int64_t)B * deg + (k - np);
  double2 kv = make_double2(0, 0), dv = make_double2(0, 0), mv = make_double2(0, 0);
  const bool uu = (A < 3) && (B < 3);
  const bool diag = uu && (A == B);
  for (int64_t c = gptr[k]; c < gptr[k + 1]; ++c) {
    const int64_t code = gidx[c];
    const int64_t e = code / 729;
    const int64_t ij = code - e * 729;
    const double2 v = Ke[(e * 16 + AB) * 729 + ij];
    kv.x += v.x; kv.y += v.y;
    if (uu && dmask[e]) { dv.x += v.x; dv.y += v.y; }
    if (diag) { const double2 w = Me[e * 729 + ij]; mv.x += w.x; mv.y += w.y; }
  }
  K[slot] = kv;
  M[slot] = mv;
  if (D) D[slot] = dv;
}

// K12' (Hex27 backward): sensitivity sums of the pair (l, d) per element and frequency.
// With Q = s^T K d and A_aa x = b, the adjoint lambda = A_aa^-T kts_a gives
//   dQ = l^T dK d + i w beta l^T dD d - w^2 l^T dM d,   l = s - T P_a^T lambda,
// d the full displacement/potential vector (x on active and slave dofs through the fold,
// V on the signal electrode, 0 on fixed and ground dofs).  K is linear in each element's
// C, e, eps and M in rho, so per element the kernel forms, already multiplied by the
// element's stretch codes (assembly order of the forward kernel):
//   Hc[s][t] = code_C(s,t) sum_q w (B l)_s (B d)_t                                 (pass 0)
//   He[k][s] = code_e(k,s) sum_q w (B l)_s (G d)_k + SCALE conj(code_e(k,s)) sum_q w (G l)_k (B d)_s
//   Hp[j][k] = code_eps(j,k) sum_q w (G l)_j (G d)_k
//   Hm       = a1 a2 a3 sum_q w sum_c (N l_c)(N d_c)                                (pass 1)
// B is the forward kernel's strain operator (gamma_yz = duy/dz), G the potential gradient.
// The host contracts the element sums with the tensor derivatives (SCALE on K_uu and M, the
// sign of K_phiphi and the damping factor applied there).  x and lambda come transposed,
// [dof][Bp], so the 64 frequencies of a wave read contiguous values.
template <int PASS>
__global__ void __launch_bounds__(64) k_adj_h27(int ne0, const int* __restrict__ elems, const int* __restrict__ conn,
                                                const double* __restrict__ xyz, const int* __restrict__ pml_idx,
                                                const Pml3Set pset, double omega_top, int N,
                                                const int* __restrict__ pos_rep, const unsigned char* __restrict__ zmask,
                                                double V, double scale, int B, int Bp, const double2* __restrict__ xt,
                                                const double2* __restrict__ lt, double2* __restrict__ out) {
  const int ei = blockIdx.x;
  const int e = elems[ei];
  const int t = threadIdx.x;
  __shared__ double X[27][3];
  __shared__ double sg[27][27][3];
  __shared__ double sN[27][27];
  __shared__ double sw[27];
  __shared__ double2 code[7];
  __shared__ int sp[27][4];
  __shared__ unsigned char sz[27][4];
  if (t < 27) {
    const int nd = conn[(int64_t)e * 27 + t];
    for (int d = 0; d < 3; ++d) X[t][d] = xyz[3 * (int64_t)nd + d];
    for (int c = 0; c < 4; ++c) {
      const int64_t dof = c < 3 ? 3 * (int64_t)nd + c : 3 * (int64_t)N + nd;
      sp[t][c] = pos_rep[dof];
      sz[t][c] = zmask[dof];
    }
  }
  __syncthreads();
  if (t < 27) h27_point(t, X, sN[t], sg[t], &sw[t]);
  if (t == 32) h27_codes(pset, pml_idx[e], X[26][0], X[26][2], omega_top, code);
  __syncthreads();
  const int b = blockIdx.y * 64 + t;
  if (b >= B) return;
  constexpr int NA = PASS == 0 ? 36 : 46;     // pass 0: Hc; pass 1: He (two parts), Hp, Hm
  double2 H[NA];
  for (int i = 0; i < NA; ++i) H[i] = make_double2(0, 0);
  for (int q = 0; q < 27; ++q) {
    double2 el[6], ed[6], gl[3], gd[3], nl[3], nd[3];
    for (int i = 0; i < 6; ++i) el[i] = ed[i] = make_double2(0, 0);
    for (int i = 0; i < 3; ++i) gl[i] = gd[i] = nl[i] = nd[i] = make_double2(0, 0);
    for (int a = 0; a < 27; ++a) {
      double2 lv[4], dv[4];
      for (int c = 0; c < 4; ++c) {
        const int p = sp[a][c];
        const bool z = sz[a][c] != 0;
        const double2 xv = p >= 0 ? xt[(int64_t)p * Bp + b] : make_double2(0, 0);
        const double2 lm = p >= 0 ? lt[(int64_t)p * Bp + b] : make_double2(0, 0);
        dv[c] = z ? make_double2(V, 0) : xv;
        lv[c] = make_double2((z ? 1.0 : 0.0) - lm.x, -lm.y);
      }
      const double g0 = sg[q][a][0], g1 = sg[q][a][1], g2 = sg[q][a][2], Nn = sN[q][a];
      // strains (Voigt xx, yy, zz, yz = duy/dz, xz, xy)
      el[0] = cadd(el[0], cscale(lv[0], g0)); ed[0] = cadd(ed[0], cscale(dv[0], g0));
      el[1] = cadd(el[1], cscale(lv[1], g1)); ed[1] = cadd(ed[1], cscale(dv[1], g1));
      el[2] = cadd(el[2], cscale(lv[2], g2)); ed[2] = cadd(ed[2], cscale(dv[2], g2));
      el[3] = cadd(el[3], cscale(lv[1], g2)); ed[3] = cadd(ed[3], cscale(dv[1], g2));
      el[4] = cadd(el[4], cadd(cscale(lv[0], g2), cscale(lv[2], g0)));
      ed[4] = cadd(ed[4], cadd(cscale(dv[0], g2), cscale(dv[2], g0)));
      el[5] = cadd(el[5], cadd(cscale(lv[0], g1), cscale(lv[1], g0)));
      ed[5] = cadd(ed[5], cadd(cscale(dv[0], g1), cscale(dv[1], g0)));
      if (PASS == 1) {
        gl[0] = cadd(gl[0], cscale(lv[3], g0)); gd[0] = cadd(gd[0], cscale(dv[3], g0));
        gl[1] = cadd(gl[1], cscale(lv[3], g1)); gd[1] = cadd(gd[1], cscale(dv[3], g1));
        gl[2] = cadd(gl[2], cscale(lv[3], g2)); gd[2] = cadd(gd[2], cscale(dv[3], g2));
        for (int c = 0; c < 3; ++c) { nl[c] = cadd(nl[c], cscale(lv[c], Nn)); nd[c] = cadd(nd[c], cscale(dv[c], Nn)); }
      }
    }
    const double w = sw[q];
    if (PASS == 0) {
      for (int s = 0; s < 6; ++s) {
        const double2 ls = cscale(el[s], w);
        for (int u = 0; u < 6; ++u) H[s * 6 + u] = cfma(H[s * 6 + u], ls, ed[u]);
      }
    } else {
      for (int k = 0; k < 3; ++k) {
        const double2 gdk = cscale(gd[k], w), glk = cscale(gl[k], w);
        for (int s = 0; s < 6; ++s) {
          H[k * 6 + s] = cfma(H[k * 6 + s], el[s], gdk);          // u-phi part
          H[18 + k * 6 + s] = cfma(H[18 + k * 6 + s], glk, ed[s]);  // phi-u part
        }
        for (int j = 0; j < 3; ++j) H[36 + k * 3 + j] = cfma(H[36 + k * 3 + j], glk, gd[j]);   // Hp[k][j]
      }
      for (int c = 0; c < 3; ++c) H[45] = cfma(H[45], cscale(nl[c], w), nd[c]);
    }
  }
  double2* o = out + ((int64_t)b * ne0 + ei) * 64;
  if (PASS == 0) {
    for (int s = 0; s < 6; ++s)
      for (int u = 0; u < 6; ++u) o[s * 6 + u] = cmul(code[kSC[s][u]], H[s * 6 + u]);
  } else {
    for (int k = 0; k < 3; ++k)
      for (int s = 0; s < 6; ++s) {
        const double2 c = code[kSE[k][s]];
        const double2 cc = make_double2(scale * c.x, -scale * c.y);
        o[36 + k * 6 + s] = cadd(cmul(c, H[k * 6 + s]), cmul(cc, H[18 + k * 6 + s]));
      }
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) o[54 + j * 3 + k] = cmul(code[kSP[j][k]], H[36 + j * 3 + k]);
    o[63] = cmul(code[6], H[45]);
  }
}

// W[b][k] = sum_e out[b][e][k], elements in list order (deterministic).
__global__ void k_adj_reduce_h27(int ne0, int B, const double2* __restrict__ out, double2* __restrict__ W) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)B * 64) return;
  const int b = (int)(g / 64), k = (int)(g % 64);
  double2 s = make_double2(0, 0);
  const double2* o = out + (int64_t)b * ne0 * 64 + k;
  for (int e = 0; e < ne0; ++e) s = cadd(s, o[(int64_t)e * 64]);
  W[g] = s;
}

}  // namespace

extern "C" {

int sawfem_elem_h27(void* stream, int ne, const int* conn, const double* xyz, const int* mat_idx, const int* pml_idx,
                    const void* C, const void* E, const void* EPS, const double* rho, int npml, const int* pml_kind,
                    const double* pml_par, double omega_top, double scale, void* Ke, void* Me) {
  if (ne < 0 || npml < 0 || npml > 8 || !conn || !xyz || !mat_idx || !pml_idx || !Ke || !Me) {
    sawfem_set_error("sawfem_elem_h27: invalid arguments");
    return 1;
  }
  if (ne == 0) return 0;
  Pml3Set ps;
  for (int r = 0; r < npml; ++r) {
    ps.r[r].kind = pml_kind[r];
    ps.r[r].dmax = pml_par[6 * r + 0]; ps.r[r].n = pml_par[6 * r + 1];
    ps.r[r].xa = pml_par[6 * r + 2]; ps.r[r].xp = pml_par[6 * r + 3];
    ps.r[r].za = pml_par[6 * r + 4]; ps.r[r].zp = pml_par[6 * r + 5];
  }
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_elem_h27, dim3(ne), dim3(256), 0, s, ne, conn, xyz, mat_idx, pml_idx, (const double2*)C,
                     (const double2*)E, (const double2*)EPS, rho, ps, omega_top, scale, (double2*)Ke, (double2*)Me);
  SAW_HIP(hipGetLastError());
  return 0;
}

int sawfem_gather_h27(void* stream, int nnzn, const int* pair_row, const int64_t* node_ptr, const int64_t* gptr,
                      const int64_t* gidx, const unsigned char* dmask, const void* Ke, const void* Me, void* K, void* M,
                      void* D) {
  if (nnzn < 0 || !pair_row || !node_ptr || !gptr || !gidx || !dmask || !Ke || !Me || !K || !M) {
    sawfem_set_error("sawfem_gather_h27: invalid arguments");
    return 1;
  }
  if (nnzn == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_gather_h27, dim3((nnzn + 255) / 256, 16), dim3(256), 0, s, nnzn, pair_row, node_ptr, gptr,
                     gidx, dmask, (const double2*)Ke, (const double2*)Me, (double2*)K, (double2*)M, (double2*)D);
  SAW_HIP(hipGetLastError());
  return 0;
}

// Element sensitivities of the pair (l, d) for B frequencies (see k_adj_h27).  xt, lt: x and
// lambda transposed to [n][Bp] (Bp >= B); work: B * ne0 * 64 complex128; W: B * 64 complex128,
// W[b] = (Hc 6x6, He 3x6, Hp 3x3, Hm).
int sawfem_adjoint_h27(void* stream, int ne0, const int* elems, const int* conn, const double* xyz, const int* pml_idx,
                       int npml, const int* pml_kind, const double* pml_par, double omega_top, int N,
                       const int* pos_rep, const unsigned char* zmask, double V, double scale, int B, int Bp,
                       const void* xt, const void* lt, void* work, void* W) {
  if (ne0 < 0 || B <= 0 || Bp < B || npml < 0 || npml > 8 || !pos_rep || !zmask || !xt || !lt || !work || !W) {
    sawfem_set_error("sawfem_adjoint_h27: invalid arguments");
    return 1;
  }
  Pml3Set ps;
  for (int r = 0; r < npml; ++r) {
    ps.r[r].kind = pml_kind[r];
    ps.r[r].dmax = pml_par[6 * r + 0]; ps.r[r].n = pml_par[6 * r + 1];
    ps.r[r].xa = pml_par[6 * r + 2]; ps.r[r].xp = pml_par[6 * r + 3];
    ps.r[r].za = pml_par[6 * r + 4]; ps.r[r].zp = pml_par[6 * r + 5];
  }
  hipStream_t s = (hipStream_t)stream;
  if (ne0 > 0) {
    dim3 grid(ne0, (B + 63) / 64);
    hipLaunchKernelGGL(k_adj_h27<0>, grid, dim3(64), 0, s, ne0, elems, conn, xyz, pml_idx, ps, omega_top, N, pos_rep,
                       zmask, V, scale, B, Bp, (const double2*)xt, (const double2*)lt, (double2*)work);
    hipLaunchKernelGGL(k_adj_h27<1>, grid, dim3(64), 0, s, ne0, elems, conn, xyz, pml_idx, ps, omega_top, N, pos_rep,
                       zmask, V, scale, B, Bp, (const double2*)xt, (const double2*)lt, (double2*)work);
    SAW_HIP(hipGetLastError());
  }
  const int64_t tot = (int64_t)B * 64;
  hipLaunchKernelGGL(k_adj_reduce_h27, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, ne0, B,
                     (const double2*)work, (double2*)W);
  SAW_HIP(hipGetLastError());
  return 0;
}

}  // extern "C"
