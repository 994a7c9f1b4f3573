// K12: backward element kernel for dY11/dp.
//
// With Q = s^T K d (s: indicator of the signal-electrode potentials, d: full displacement /
// potential vector) and A_aa x = b, the adjoint lambda = A_aa^-T kts_a gives
//     dQ = l^T dK d + i w beta l^T dD d + i w dbeta l^T D d,     l = s - T P_a^T lambda,
// (the u part of s vanishes, so -lambda_u = l_u).  K, D are linear in the lifted material
// tensors of each element, so per element this kernel forms the bilinear "sensitivity"
// tensors of the pair (l, d), already multiplied by the element's PML codes:
//     Hc[i][j] = code(Ri, Rj) * sum_q dV_q epsl_q[Ri] epsd_q[Rj]                  (C9 entries)
//     He[k][i] = code(k, Ri)  * sum_q dV_q (epsl_q[Ri] gd_q[k] + SCALE gl_q[k] epsd_q[Ri])
//     Hp[j][k] = teps(j, k)   * sum_q dV_q gl_q[j] gd_q[k]                        (eps entries)
// where Ri = 9-component strain row of displacement component A = i / 2 and direction
// d = i % 2, and gl/gd are potential gradients.  The host contracts the element sums with
// the derivatives of the lifted tensors (SCALE and the sign of K_phiphi applied there).
// Element matrices follow SawSim saw2d/piezo_fem.py:312-350; the element loop runs over the
// elements of the parameterised material only.
#include "q9.h"

#include <mutex>

// Scratch of sawfem_adjoint_elem_noalloc: the per-element sensitivity sums (B * ne0 * NT
// complex128) live in this file-scope device array instead of a caller workspace.  Each GPU the
// library runs on holds its own copy from module load until the process ends.
constexpr size_t kAdjScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kAdjScratchBytes];

namespace {

// Device address of `scratch` per device, looked up once per device with that device current;
// host code uses these addresses, never the symbol itself.
constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

int scratch_address(void** out) {
  int dev = -1;
  SAW_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    sawfem_set_error("sawfem_adjoint_elem_noalloc: device ordinal out of range");
    return 1;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    SAW_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
      sawfem_set_error("sawfem_adjoint_elem_noalloc: no device address for the scratch array");
      return 3;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

using namespace q9;

// Gauss-point gradients of component c in direction d of the node values v(c, a).
template <typename Val>
__device__ __forceinline__ double2 grad_at(const double (*dNq)[2], int c, int d, Val v) {
  double2 s = make_double2(0, 0);
  for (int a = 0; a < 9; ++a) {
    const double g = dNq[a][d];
    const double2 x = v(c, a);
    s.x += g * x.x; s.y += g * x.y;
  }
  return s;
}

// Sensitivity sums of one pass: C rows [I0, I0 + NI) of Hc and, with EP, He and Hp; written to o
// already multiplied by the element's codes.  dval/lval give the node values of d and l.
template <int NDU, int I0, int NI, bool EP, typename DV, typename LV>
__device__ __forceinline__ void adj_pass(const double (*dNx)[9][2], const double* dV, const double2* code,
                                         double scale, DV dval, LV lval, double2* o) {
  constexpr int ND = NDU + 1;
  constexpr int NR = 2 * NDU;
  double2 Hc[NI][NR], He[2][NR], Hp[2][2];
  for (int i = 0; i < NI; ++i)
    for (int j = 0; j < NR; ++j) Hc[i][j] = make_double2(0, 0);
  for (int i = 0; i < NR; ++i) He[0][i] = He[1][i] = make_double2(0, 0);
  Hp[0][0] = Hp[0][1] = Hp[1][0] = Hp[1][1] = make_double2(0, 0);
  // one Gauss point at a time (unrolled, the gradients of all nine would be live together)
#pragma unroll 1
  for (int q = 0; q < 9; ++q) {
    // compiler-only fence: keeps LDS-held node values from being hoisted out of the loop into
    // registers (that would bring back the register demand the LDS copy avoids)
    asm volatile("" ::: "memory");
    double2 gd[ND][2], gl[ND][2];
    for (int c = 0; c < ND; ++c)
      for (int d = 0; d < 2; ++d) {
        gd[c][d] = grad_at(dNx[q], c, d, dval);
        gl[c][d] = grad_at(dNx[q], c, d, lval);
      }
    const double w = dV[q];
    for (int ii = 0; ii < NI; ++ii) {
      const int i = I0 + ii;
      const double2 li = gl[i / 2][i % 2];
      for (int j = 0; j < NR; ++j) {
        const double2 pr = cmul(li, gd[j / 2][j % 2]);
        Hc[ii][j].x += w * pr.x; Hc[ii][j].y += w * pr.y;
      }
    }
    if (EP) {
      for (int i = 0; i < NR; ++i) {
        const double2 li = gl[i / 2][i % 2];
        const double2 di = gd[i / 2][i % 2];
        for (int k = 0; k < 2; ++k) {
          const double2 a1 = cmul(li, gd[NDU][k]);
          const double2 a2 = cmul(gl[NDU][k], di);
          He[k][i].x += w * (a1.x + scale * a2.x);
          He[k][i].y += w * (a1.y + scale * a2.y);
        }
      }
      for (int j = 0; j < 2; ++j)
        for (int k = 0; k < 2; ++k) {
          const double2 pr = cmul(gl[NDU][j], gd[NDU][k]);
          Hp[j][k].x += w * pr.x; Hp[j][k].y += w * pr.y;
        }
    }
  }
  for (int ii = 0; ii < NI; ++ii)
    for (int j = 0; j < NR; ++j) {
      const int i = I0 + ii;
      const int ri = kStrainRow[i / 2][i % 2], rj = kStrainRow[j / 2][j % 2];
      o[i * NR + j] = cmul(code[kT9Code[ri][rj]], Hc[ii][j]);
    }
  if (EP) {
    for (int k = 0; k < 2; ++k)
      for (int i = 0; i < NR; ++i) {
        const int ri = kStrainRow[i / 2][i % 2];
        o[NR * NR + k * NR + i] = cmul(code[kT9Code[k][ri]], He[k][i]);
      }
    for (int j = 0; j < 2; ++j)
      for (int k = 0; k < 2; ++k) o[NR * NR + 2 * NR + j * 2 + k] = cmul(code[kTepsCode[j][k]], Hp[j][k]);
  }
}

// One workgroup of 64 threads per element and 64 frequencies (thread = frequency).
// NDU = 2 keeps the node values of d and l in registers and forms every sum in one pass.
// NDU = 3 has 4/3 more node values and 9/4 more C sums than fit the register file, so its node
// values sit in (dynamic) LDS, [2][ND * 9][64], each thread reading only its own column, and
// the C sums are formed in two passes of three rows: no scratch memory in either case.
template <int NDU>
__global__ void __launch_bounds__(64) k_adj_elem(int ne0, const int* __restrict__ elems, const int* __restrict__ conn,
                                                 const double* __restrict__ xy, const int* __restrict__ elem_tag,
                                                 int npml, const PmlSet pmlset, double omega_top, int N,
                                                 const int* __restrict__ pos_rep, const unsigned char* __restrict__ zmask,
                                                 double V, double scale, int B, int n, const double2* __restrict__ x,
                                                 const double2* __restrict__ lam, double2* __restrict__ out) {
  constexpr int ND = NDU + 1;
  constexpr int NR = 2 * NDU;
  constexpr int NT = NR * NR + 2 * NR + 4;
  const int ei = blockIdx.x;
  const int e = elems[ei];
  const int t = threadIdx.x;
  __shared__ double sx[9], sy[9];
  __shared__ double dNx[9][9][2];
  __shared__ double Nq[9][9];
  __shared__ double dV[9];
  __shared__ double2 al[3];
  __shared__ double2 code[6];
  __shared__ int dofs[9 * ND];
  extern __shared__ double2 snode[];
  if (t < 9) {
    const int nd = conn[e * 9 + t];
    sx[t] = xy[2 * nd];
    sy[t] = xy[2 * nd + 1];
    for (int c = 0; c < ND; ++c) dofs[c * 9 + t] = c * N + nd;
  }
  __syncthreads();
  if (t < 9) gauss_point(t, sx, sy, Nq[t], dNx[t], &dV[t]);
  if (t == 32) {
    pml_alpha(pmlset, npml, elem_tag[e], sx[8], sy[8], omega_top, al);
    pml_codes(al, code);
  }
  __syncthreads();
  const int b = blockIdx.y * 64 + t;
  if (b >= B) return;
  double2* o = out + ((int64_t)b * ne0 + ei) * NT;
  auto node = [&](int c, int a, double2& dv, double2& lv) {
    const int dof = dofs[c * 9 + a];
    const int p = pos_rep[dof];
    const bool z = zmask[dof] != 0;
    const double2 xv = p >= 0 ? x[(int64_t)b * n + p] : make_double2(0, 0);
    const double2 l = p >= 0 ? lam[(int64_t)b * n + p] : make_double2(0, 0);
    dv = z ? make_double2(V, 0) : xv;
    lv = make_double2((z ? 1.0 : 0.0) - l.x, -l.y);
  };
  if constexpr (NDU == 2) {
    double2 dvl[ND][9], lvl[ND][9];
    for (int c = 0; c < ND; ++c)
      for (int a = 0; a < 9; ++a) node(c, a, dvl[c][a], lvl[c][a]);
    auto dval = [&](int c, int a) { return dvl[c][a]; };
    auto lval = [&](int c, int a) { return lvl[c][a]; };
    adj_pass<NDU, 0, NR, true>(dNx, dV, code, scale, dval, lval, o);
  } else {
    double2* sd = snode;                  // [ND * 9][64]
    double2* sl = snode + ND * 9 * 64;
    for (int c = 0; c < ND; ++c)
      for (int a = 0; a < 9; ++a) node(c, a, sd[(c * 9 + a) * 64 + t], sl[(c * 9 + a) * 64 + t]);
    auto dval = [&](int c, int a) { return sd[(c * 9 + a) * 64 + t]; };
    auto lval = [&](int c, int a) { return sl[(c * 9 + a) * 64 + t]; };
    adj_pass<NDU, 0, NR / 2, true>(dNx, dV, code, scale, dval, lval, o);
    adj_pass<NDU, NR / 2, NR / 2, false>(dNx, dV, code, scale, dval, lval, o);
  }
}

// W[b][t] = sum_e out[b][e][t], elements in list order (deterministic).
__global__ void k_adj_reduce(int ne0, int NT, int B, const double2* __restrict__ out, double2* __restrict__ W) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)B * NT) return;
  const int b = (int)(g / NT), tt = (int)(g % NT);
  double2 s = make_double2(0, 0);
  const double2* o = out + (int64_t)b * ne0 * NT + tt;
  for (int e = 0; e < ne0; ++e) {
    const double2 v = o[(int64_t)e * NT];
    s.x += v.x; s.y += v.y;
  }
  W[g] = s;
}

}  // namespace

extern "C" {

// Element sensitivities of the pair (l, d) for B frequencies; see the file header.
// work: B * ne0 * NT complex128 scratch, W: B * NT complex128 output,
// NT = (2 ndu)^2 + 4 ndu + 4.
int sawfem_adjoint_elem(void* stream, int ne0, const int* elems, const int* conn, const double* xy,
                        const int* elem_tag, int npml, const int* pml_tag, const int* pml_kind, const double* pml_par,
                        double omega_top, int ndu, int N, const int* pos_rep, const unsigned char* zmask, double V,
                        double scale, int B, int n, const void* x, const void* lam, void* work, void* W) {
  if (ne0 < 0 || B <= 0 || n <= 0 || (ndu != 2 && ndu != 3) || npml < 0 || npml > 8) {
    sawfem_set_error("sawfem_adjoint_elem: invalid arguments");
    return 1;
  }
  const int NR = 2 * ndu, NT = NR * NR + 2 * NR + 4;
  hipStream_t s = (hipStream_t)stream;
  if (ne0 > 0) {
    const q9::PmlSet ps = q9::make_pml_set(npml, pml_tag, pml_kind, pml_par);
    dim3 grid(ne0, (B + 63) / 64);
    if (ndu == 2)
      hipLaunchKernelGGL(k_adj_elem<2>, grid, dim3(64), 0, s, ne0, elems, conn, xy, elem_tag, npml, ps, omega_top, N,
                         pos_rep, zmask, V, scale, B, n, (const double2*)x, (const double2*)lam, (double2*)work);
    else {
      // node values of d and l for 64 frequencies: 2 * 36 * 64 complex128 = 72 KB of LDS
      const int lds = 2 * 4 * 9 * 64 * (int)sizeof(double2);
      SAW_HIP(hipFuncSetAttribute((const void*)k_adj_elem<3>, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
      hipLaunchKernelGGL(k_adj_elem<3>, grid, dim3(64), lds, s, ne0, elems, conn, xy, elem_tag, npml, ps, omega_top, N,
                         pos_rep, zmask, V, scale, B, n, (const double2*)x, (const double2*)lam, (double2*)work);
    }
    SAW_HIP(hipGetLastError());
  }
  const int64_t tot = (int64_t)B * NT;
  hipLaunchKernelGGL(k_adj_reduce, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, ne0, NT, B,
                     (const double2*)work, (double2*)W);
  SAW_HIP(hipGetLastError());
  return 0;
}

// The same computation with the scratch array above in place of `work` (no device memory is
// allocated).  B * ne0 * NT * 16 bytes must fit the array (64 MiB); the kernels write every
// scratch entry they read, so nothing needs clearing.  Calls that overlap on one GPU race on
// the array.
int sawfem_adjoint_elem_noalloc(void* stream, int ne0, const int* elems, const int* conn, const double* xy,
                                const int* elem_tag, int npml, const int* pml_tag, const int* pml_kind,
                                const double* pml_par, double omega_top, int ndu, int N, const int* pos_rep,
                                const unsigned char* zmask, double V, double scale, int B, int n, const void* x,
                                const void* lam, void* W) {
  if (ne0 < 0 || B <= 0 || (ndu != 2 && ndu != 3)) {
    sawfem_set_error("sawfem_adjoint_elem_noalloc: invalid arguments");
    return 1;
  }
  const int NR = 2 * ndu, NT = NR * NR + 2 * NR + 4;
  if ((size_t)B * (size_t)ne0 * (size_t)NT * sizeof(double2) > kAdjScratchBytes) {
    sawfem_set_error("sawfem_adjoint_elem_noalloc: B * ne0 * NT complex128 sums exceed the 64 MiB scratch array");
    return 1;
  }
  void* work = nullptr;
  const int rc = scratch_address(&work);
  if (rc) return rc;
  return sawfem_adjoint_elem(stream, ne0, elems, conn, xy, elem_tag, npml, pml_tag, pml_kind, pml_par, omega_top,
                             ndu, N, pos_rep, zmask, V, scale, B, n, x, lam, work, W);
}

}  // extern "C"
