// K1 element assembly (Q9, plane) and K2/K3 deterministic gathers.
//
// K1 restates the element contractions of SawSim saw2d/piezo_fem.py:248-365:
//   * 3x3 Gauss rule and biquadratic Q9 shapes in Gmsh QUADS9 order (:48-103),
//   * signed detJ (:264, 272), physical gradients, Gram matrices G[d1][d2] and GM,
//   * PML stretch evaluated at the element centre node, frozen at omega_top (:222-242),
//     multipliers T9 (:127-142) and Teps (:145-155), rho_eff = rho a1 a2 a3 (:315),
//   * blocks K_uu = SCALE sum C_eff G (:328-334), K_uphi = sum E_eff G (:337-342),
//     K_phiu = SCALE K_uphi^T (:343), K_phiphi = -sum P_eff G (:346-350),
//     M = SCALE rho_eff GM on each displacement block (:357-362).
// One workgroup per element; thread t < 81 owns the node pair (a = t / 9, b = t % 9).
// Output layout: Ke[e][A][B][a][b] (complex128), Me[e][a][b] (complex128).
#include "q9.h"
#include <cstdio>

namespace {

using namespace q9;

__global__ void __launch_bounds__(128) k_elem_q9(
    int ne, int ndu, const int* __restrict__ conn, const double* __restrict__ xy,
    const int* __restrict__ mat_idx, const int* __restrict__ elem_tag,
    const double2* __restrict__ C9, const double2* __restrict__ E9, const double2* __restrict__ EPS,
    const double* __restrict__ RHO, int npml, const PmlSet pmlset, double omega_top,
    double scale, double2* __restrict__ Ke, double2* __restrict__ Me) {
  const int e = blockIdx.x;
  if (e >= ne) return;
  const int t = threadIdx.x;
  __shared__ double sx[9], sy[9];
  __shared__ double dNx[9][9][2];   // [q][a][d]
  __shared__ double Nq[9][9];       // [q][a]
  __shared__ double dV[9];
  __shared__ double2 al[3];
  // multiplier codes in LDS: they are indexed by table lookups, which a private array would
  // have to serve from scratch memory
  __shared__ double2 code[6];
  if (t < 9) {
    const int nd = conn[e * 9 + t];
    sx[t] = xy[2 * nd];
    sy[t] = xy[2 * nd + 1];
  }
  __syncthreads();
  if (t < 9) gauss_point(t, sx, sy, Nq[t], dNx[t], &dV[t]);
  if (t == 32) {
    pml_alpha(pmlset, npml, elem_tag[e], sx[8], sy[8], omega_top, al);
    pml_codes(al, code);
  }
  __syncthreads();
  if (t >= 81) return;
  const int a = t / 9, b = t % 9;
  double G[2][2] = {{0, 0}, {0, 0}}, GM = 0;
  for (int q = 0; q < 9; ++q) {
    const double v = dV[q];
    G[0][0] += v * dNx[q][a][0] * dNx[q][b][0];
    G[0][1] += v * dNx[q][a][0] * dNx[q][b][1];
    G[1][0] += v * dNx[q][a][1] * dNx[q][b][0];
    G[1][1] += v * dNx[q][a][1] * dNx[q][b][1];
    GM += v * Nq[q][a] * Nq[q][b];
  }
  const int mat = mat_idx[e];
  const double2* Cm = C9 + mat * 81;
  const double2* Em = E9 + mat * 27;
  const double2* Pm = EPS + mat * 9;
  const int nd = ndu + 1;
  double2* Kout = Ke + (size_t)e * nd * nd * 81;
  // K_uu: sum over direction pairs of C_eff[r(A,dv)][r(B,du)] G[dv][du]
  for (int A = 0; A < ndu; ++A) {
    for (int B = 0; B < ndu; ++B) {
      double2 acc = make_double2(0, 0);
      for (int dv = 0; dv < 2; ++dv)
        for (int du = 0; du < 2; ++du) {
          const int rv = kStrainRow[A][dv], ru = kStrainRow[B][du];
          const double2 ceff = cmul(code[kT9Code[rv][ru]], Cm[rv * 9 + ru]);
          acc.x += ceff.x * G[dv][du];
          acc.y += ceff.y * G[dv][du];
        }
      acc.x *= scale; acc.y *= scale;
      Kout[(A * nd + B) * 81 + t] = acc;
    }
  }
  // K_uphi[A](a, b) = sum_{dv, k} E_eff[k][rv] G[dv][k](a, b); K_phiu[A] = SCALE * K_uphi[A]^T,
  // so entry (a, b) of K_phiu needs G[dv][k](b, a) = G[k][dv](a, b).
  for (int A = 0; A < ndu; ++A) {
    double2 up = make_double2(0, 0), pu = make_double2(0, 0);
    for (int dv = 0; dv < 2; ++dv)
      for (int k = 0; k < 2; ++k) {
        const int rv = kStrainRow[A][dv];
        const double2 eeff = cmul(code[kT9Code[k][rv]], Em[k * 9 + rv]);
        up.x += eeff.x * G[dv][k]; up.y += eeff.y * G[dv][k];
        pu.x += eeff.x * G[k][dv]; pu.y += eeff.y * G[k][dv];
      }
    Kout[(A * nd + ndu) * 81 + t] = up;
    Kout[(ndu * nd + A) * 81 + t] = make_double2(scale * pu.x, scale * pu.y);
  }
  {
    double2 pp = make_double2(0, 0);
    for (int j = 0; j < 2; ++j)
      for (int k = 0; k < 2; ++k) {
        const double2 peff = cmul(code[kTepsCode[j][k]], Pm[j * 3 + k]);
        pp.x += peff.x * G[j][k]; pp.y += peff.y * G[j][k];
      }
    Kout[(ndu * nd + ndu) * 81 + t] = make_double2(-pp.x, -pp.y);
  }
  const double2 a123 = cmul(cmul(al[0], al[1]), al[2]);
  const double rs = RHO[mat] * GM * scale;
  Me[(size_t)e * 81 + t] = make_double2(a123.x * rs, a123.y * rs);
}

// K2: global K, M, D on the structural pattern, gathered per node pair k and block (A, B).
// slot = A*nd*nnzn + nd*node_ptr[i] + B*deg[i] + (k - node_ptr[i]),  i = pair_row[k]
__global__ void k_gather_kmd(int nnzn, int nd, int ndu, const int* __restrict__ pair_row,
                             const int64_t* __restrict__ node_ptr, const int64_t* __restrict__ gptr,
                             const int64_t* __restrict__ gidx, const unsigned char* __restrict__ dmask,
                             const double2* __restrict__ Ke, const double2* __restrict__ Me,
                             double2* __restrict__ K, double2* __restrict__ M, double2* __restrict__ D) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnzn) return;
  const int AB = blockIdx.y;
  const int A = AB / nd, B = AB % nd;
  const int i = pair_row[k];
  const int64_t np = node_ptr[i];
  const int64_t deg = node_ptr[i + 1] - np;
  const int64_t slot = (int64_t)A * nd * nnzn + (int64_t)nd * np + (int64_t)B * deg + (k - np);
  double2 kv = make_double2(0, 0), dv = make_double2(0, 0), mv = make_double2(0, 0);
  const bool uu = (A < ndu) && (B < ndu);
  const bool diag = uu && (A == B);
  for (int64_t c = gptr[k]; c < gptr[k + 1]; ++c) {
    const int64_t code = gidx[c];
    const int64_t e = code / 81;
    const int64_t ij = code - e * 81;
    const double2 v = Ke[(e * nd * nd + AB) * 81 + ij];
    kv.x += v.x; kv.y += v.y;
    if (uu && dmask[e]) { dv.x += v.x; dv.y += v.y; }
    if (diag) { const double2 w = Me[e * 81 + ij]; mv.x += w.x; mv.y += w.y; }
  }
  K[slot] = kv;
  M[slot] = mv;
  if (D) D[slot] = dv;
}

// K3: out[j] = sum_{c in [ptr[j], ptr[j+1])} src[idx[c]] for up to three sources sharing the list.
__global__ void k_gather3(int64_t nout, const int64_t* __restrict__ ptr, const int64_t* __restrict__ idx,
                          const double2* __restrict__ s0, const double2* __restrict__ s1,
                          const double2* __restrict__ s2, double2* __restrict__ d0,
                          double2* __restrict__ d1, double2* __restrict__ d2) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nout) return;
  double2 a0 = make_double2(0, 0), a1 = a0, a2 = a0;
  for (int64_t c = ptr[j]; c < ptr[j + 1]; ++c) {
    const int64_t s = idx[c];
    const double2 v0 = s0[s]; a0.x += v0.x; a0.y += v0.y;
    if (s1) { const double2 v1 = s1[s]; a1.x += v1.x; a1.y += v1.y; }
    if (s2) { const double2 v2 = s2[s]; a2.x += v2.x; a2.y += v2.y; }
  }
  d0[j] = a0;
  if (d1) d1[j] = a1;
  if (d2) d2[j] = a2;
}

}  // namespace

static thread_local char g_err[512] = "";
const char* sawfem_set_error(const char* msg) {
  snprintf(g_err, sizeof(g_err), "%s", msg ? msg : "");
  return g_err;
}

extern "C" {

const char* sawfem_last_error(void) { return g_err; }

int sawfem_elem_q9(void* stream, int ne, int ndu, const int* conn, const double* xy, const int* mat_idx,
                   const int* elem_tag, const void* C9, const void* E9, const void* EPS, const double* rho,
                   int npml, const int* pml_tag, const int* pml_kind, const double* pml_par,
                   double omega_top, double scale, void* Ke, void* Me) {
  if (ne < 0 || (ndu != 2 && ndu != 3) || npml < 0 || npml > 8) {
    sawfem_set_error("sawfem_elem_q9: invalid arguments");
    return 1;
  }
  if (ne == 0) return 0;
  const q9::PmlSet ps = q9::make_pml_set(npml, pml_tag, pml_kind, pml_par);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_elem_q9, dim3(ne), dim3(128), 0, s, ne, ndu, conn, xy, mat_idx, elem_tag,
                     (const double2*)C9, (const double2*)E9, (const double2*)EPS, rho, npml, ps,
                     omega_top, scale, (double2*)Ke, (double2*)Me);
  SAW_HIP(hipGetLastError());
  return 0;
}

int sawfem_gather_kmd(void* stream, int nnzn, int nd, int ndu, const int* pair_row, const int64_t* node_ptr,
                      const int64_t* gptr, const int64_t* gidx, const unsigned char* dmask, const void* Ke,
                      const void* Me, void* K, void* M, void* D) {
  if (nnzn < 0 || nd < 3 || nd > 4 || ndu != nd - 1) {
    sawfem_set_error("sawfem_gather_kmd: invalid arguments");
    return 1;
  }
  if (nnzn == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_gather_kmd, dim3((nnzn + 255) / 256, nd * nd), dim3(256), 0, s, nnzn, nd, ndu, pair_row,
                     node_ptr, gptr, gidx, dmask, (const double2*)Ke, (const double2*)Me, (double2*)K,
                     (double2*)M, (double2*)D);
  SAW_HIP(hipGetLastError());
  return 0;
}

int sawfem_gather3(void* stream, int64_t nout, const int64_t* ptr, const int64_t* idx, const void* s0,
                   const void* s1, const void* s2, void* d0, void* d1, void* d2) {
  if (nout < 0 || !s0 || !d0 || (!s1) != (!d1) || (!s2) != (!d2)) {
    sawfem_set_error("sawfem_gather3: invalid arguments");
    return 1;
  }
  if (nout == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_gather3, dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, s, nout, ptr, idx,
                     (const double2*)s0, (const double2*)s1, (const double2*)s2, (double2*)d0, (double2*)d1,
                     (double2*)d2);
  SAW_HIP(hipGetLastError());
  return 0;
}

}  // extern "C"
