// Refined chunk sweep: Q(w_b) of B frequencies from the batched multifrontal factorization,
// refined against the FP64 operator, with the complex128 refactorization of the frequencies
// whose complex64 refinement does not converge.  This is the whole per-chunk solve behind the
// reference's FrequencySolver.solve (SawSim saw2d/sweep.py:335-364: right-hand side, backend
// solve of A(w) x = b, Q = kts . disp) for a caller that only has the C ABI; it drives the
// factor/solve/residual/charge entry points of mf.hip on one stream and decides the refinement
// steps on the host (one Q read-back per step).
//
// Refinement (the same rule as the Python engine): after each correction, dq = |dQ| / |Q| per
// frequency.  A frequency leaves the refinement when dq <= ir_tol; from the second correction
// on, one whose correction does not shrink by half while above ir_tol has stalled.  Stalled and
// unconverged frequencies of a complex64 chunk are refactored and refined in complex128.  The
// reported error estimate extrapolates the observed rate r = dq / dq_prev: dq r / (1 - r) for
// r < 1/2 (the corrections still to come), dq otherwise.
#include "common.h"
#include "../../include/sawfem.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace {

__global__ void k_zero_rows(int n, int B, const unsigned char* __restrict__ keep, double2* __restrict__ r) {
  const int b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && !keep[b]) r[(int64_t)b * n + i] = make_double2(0, 0);
}

__global__ void k_copy(int64_t n, const double2* __restrict__ a, double2* __restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

struct Layout {
  size_t arena, uarena, vec, tvec, qd, om, sel, keep, total;
};

// share: consecutive frequencies per complex64 factorization (1: one each)
Layout layout_of(void* plan, int n, int B, int prec, int share = 1) {
  Layout L{};
  const int64_t ae = sawfem_mf_arena_elems(plan), ue = sawfem_mf_uarena_elems(plan);
  const int Bp = (B + 63) / 64 * 64;
  const int Bf = prec == 0 ? (B + share - 1) / share : B;
  // the arena also holds the complex128 refactorization of at least one frequency
  L.arena = align256(std::max<size_t>((size_t)ae * Bf * (prec == 0 ? 8 : 16), (size_t)ae * 16));
  L.uarena = align256((size_t)std::max<int64_t>(1, ue) * B * 16);
  L.vec = align256((size_t)n * B * 16);         // rhs, x, r: [B][n]
  L.tvec = align256((size_t)n * Bp * 16);       // rhst, xt, rt: [n][Bp]
  L.qd = align256((size_t)B * 16);
  L.om = align256((size_t)2 * B * 8);          // the frequencies, then the group means
  L.sel = align256((size_t)B * 4);
  L.keep = align256((size_t)B);
  L.total = L.arena + L.uarena + 3 * L.vec + 3 * L.tvec + L.qd + L.om + L.sel + L.keep;
  return L;
}

size_t hsel_count(const std::vector<unsigned char>& a) {
  size_t c = 0;
  for (unsigned char v : a) c += v ? 1 : 0;
  return c;
}

double ir_error(double dq, double dq_prev) {
  const double r = std::isfinite(dq_prev) ? dq / std::max(dq_prev, 1e-300) : 1.0;
  return r < 0.5 ? dq * r / (1.0 - r) : dq;
}

#define SAW_CALL(call)             \
  do {                             \
    const int rc_ = (call);        \
    if (rc_ != 0) return rc_;      \
  } while (0)

// One precision pass over B frequencies (omega on the host).  Q (2B doubles), err (B) and
// stalled/unconverged flags (B) out.
// With share G > 1 (complex64): groups of G consecutive frequencies are factored once, at the mean
// of their angular frequencies, and each frequency is refined against its own FP64 operator with
// those factors; the solves take the vectors of a group together (sawfem_mf_solve_groups).
int refine_pass(void* plan, hipStream_t s, const sawfem_operator* op, int B, const double* omega, double V, int prec,
                double ir_tol, int max_ir, unsigned char* ws, int64_t ws_bytes, double* Q, double* err,
                unsigned char* bad, int* steps_out, int share = 1) {
  const int n = op->n;
  const int G = prec == 0 ? std::max(1, std::min(share, B)) : 1;
  const Layout L = layout_of(plan, n, B, prec, G);
  if ((int64_t)L.total > ws_bytes) {
    sawfem_set_error("sawfem_sweep_chunk: workspace smaller than sawfem_sweep_chunk_bytes");
    return 4;
  }
  unsigned char* p = ws;
  void* arena = p; p += L.arena;
  void* uarena = p; p += L.uarena;
  double2* rhs = (double2*)p; p += L.vec;
  double2* x = (double2*)p; p += L.vec;
  double2* r = (double2*)p; p += L.vec;
  double2* rhst = (double2*)p; p += L.tvec;
  double2* xt = (double2*)p; p += L.tvec;
  double2* rt = (double2*)p; p += L.tvec;
  double2* Qd = (double2*)p; p += L.qd;
  double* om = (double*)p; p += L.om;
  int* sel = (int*)p; p += L.sel;
  unsigned char* keep = p;
  const int Bp = (B + 63) / 64 * 64;
  const int nb = prec == 0 ? 64 : 32;
  // frequencies, then (shared factors) the group means summed in frequency order
  const int Bf = (B + G - 1) / G;
  std::vector<double> hom(omega, omega + B);
  for (int g = 0; g < Bf && G > 1; ++g) {
    double sum = 0;
    int cnt = 0;
    for (int b = g * G; b < std::min(B, (g + 1) * G); ++b, ++cnt) sum += omega[b];
    hom.push_back(sum / cnt);
  }
  SAW_HIP(hipMemcpyAsync(om, hom.data(), sizeof(double) * hom.size(), hipMemcpyHostToDevice, s));
  const double* omf = G > 1 ? om + B : om;
  SAW_CALL(sawfem_mf_factor(plan, s, prec, Bf, omf, op->beta, op->Kaa, op->Maa, op->Daa, arena, nb));
  SAW_CALL(sawfem_rhs(s, n, B, om, op->beta, V, op->rK, op->rM, op->rD, rhs));
  hipLaunchKernelGGL(k_copy, dim3((unsigned)(((int64_t)n * B + 255) / 256)), dim3(256), 0, s, (int64_t)n * B, rhs, x);
  SAW_HIP(hipGetLastError());
  // solve of the selected frequencies (G > 1: every group with one of them; the others' right-hand
  // sides are zero)
  std::vector<int> hgsel;
  auto solve = [&](double2* v, const std::vector<unsigned char>* active) -> int {
    if (G == 1) {
      if (!active) return sawfem_mf_solve(plan, s, prec, B, 1, arena, uarena, v, 0, nb, nullptr, 0);
      return sawfem_mf_solve(plan, s, prec, B, 1, arena, uarena, v, 0, nb, sel, (int)hsel_count(*active));
    }
    const int nfull = B / G;
    int tail = B % G ? 1 : 0;
    if (!active) return sawfem_mf_solve_groups(plan, s, B, 1, arena, G, uarena, v, 0, nb, nullptr, 0, tail);
    hgsel.clear();
    for (int g = 0; g < nfull; ++g) {
      bool any = false;
      for (int b = g * G; b < (g + 1) * G; ++b) any = any || (*active)[b];
      if (any) hgsel.push_back(g);
    }
    bool tany = false;
    for (int b = nfull * G; b < B; ++b) tany = tany || (*active)[b];
    tail = tany ? 1 : 0;
    if (!hgsel.empty())
      SAW_HIP(hipMemcpyAsync(sel, hgsel.data(), sizeof(int) * hgsel.size(), hipMemcpyHostToDevice, s));
    if (hgsel.empty() && !tail) return 0;
    return sawfem_mf_solve_groups(plan, s, B, 1, arena, G, uarena, v, 0, nb, hgsel.empty() ? nullptr : sel,
                                  (int)hgsel.size(), tail);
  };
  SAW_CALL(solve(x, nullptr));
  SAW_CALL(sawfem_transpose(s, B, n, n, Bp, rhs, rhst));
  std::vector<double> q(2 * B), qn(2 * B), dqp(B, INFINITY), e(B, INFINITY);
  std::vector<unsigned char> act(B, 1), stalled(B, 0);
  std::vector<int> hsel;
  auto charge = [&](double* out) -> int {
    SAW_CALL(sawfem_charge(s, n, B, op->kts_a, op->kts_z_re, op->kts_z_im, V, x, Qd));
    SAW_HIP(hipMemcpyAsync(out, Qd, sizeof(double2) * B, hipMemcpyDeviceToHost, s));
    SAW_HIP(hipStreamSynchronize(s));
    return 0;
  };
  SAW_CALL(charge(q.data()));
  int steps = 0;
  for (steps = 1; steps <= max_ir; ++steps) {
    SAW_CALL(sawfem_transpose(s, B, n, n, Bp, x, xt));
    SAW_CALL(sawfem_residual_t(s, n, B, Bp, op->indptr, op->indices, nullptr, op->Kaa, op->Maa, op->Daa, om,
                               op->beta, xt, rhst, rt));
    SAW_CALL(sawfem_transpose(s, n, B, Bp, n, rt, r));
    hsel.clear();
    for (int b = 0; b < B; ++b)
      if (act[b]) hsel.push_back(b);
    const bool all = (int)hsel.size() == B;
    if (!all) {
      SAW_HIP(hipMemcpyAsync(keep, act.data(), B, hipMemcpyHostToDevice, s));
      if (G == 1) SAW_HIP(hipMemcpyAsync(sel, hsel.data(), sizeof(int) * hsel.size(), hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL(k_zero_rows, dim3((n + 255) / 256, B), dim3(256), 0, s, n, B, keep, r);
      SAW_HIP(hipGetLastError());
    }
    SAW_CALL(solve(r, all ? nullptr : &act));
    SAW_CALL(sawfem_axpy(s, (int64_t)n * B, x, r));
    SAW_CALL(charge(qn.data()));
    bool any = false;
    for (int b = 0; b < B; ++b) {
      const double dr = qn[2 * b] - q[2 * b], di = qn[2 * b + 1] - q[2 * b + 1];
      const double dq = std::hypot(dr, di) / std::max(std::hypot(qn[2 * b], qn[2 * b + 1]), 1e-300);
      const double eb = ir_error(dq, dqp[b]);
      if (act[b]) {
        if (steps >= 2 && dq > 0.5 * dqp[b] && dq > ir_tol) stalled[b] = 1;
        e[b] = eb;
        dqp[b] = dq;
      }
      q[2 * b] = qn[2 * b];
      q[2 * b + 1] = qn[2 * b + 1];
      act[b] = act[b] && dqp[b] > ir_tol && !stalled[b];
      any = any || act[b];
    }
    if (!any) break;
  }
  for (int b = 0; b < B; ++b) {
    Q[2 * b] = q[2 * b];
    Q[2 * b + 1] = q[2 * b + 1];
    err[b] = e[b];
    bad[b] = !(dqp[b] <= ir_tol);
  }
  if (steps_out) *steps_out = std::min(steps, max_ir);
  return 0;
}

}  // namespace

extern "C" {

int64_t sawfem_sweep_chunk_bytes(void* plan, int n, int B, int prec) {
  if (!plan || n <= 0 || B <= 0 || (prec != 0 && prec != 1)) return -1;
  return (int64_t)layout_of(plan, n, B, prec).total;
}

int64_t sawfem_sweep_chunk_share_bytes(void* plan, int n, int B, int prec, int share) {
  if (!plan || n <= 0 || B <= 0 || (prec != 0 && prec != 1) || share < 1) return -1;
  return (int64_t)layout_of(plan, n, B, prec, prec == 0 ? std::min(share, B) : 1).total;
}

int sawfem_sweep_chunk(void* plan, void* stream, const sawfem_operator* op, int B, const double* omega, double V,
                       int prec, double ir_tol, int max_ir, int fallback, void* workspace, int64_t workspace_bytes,
                       double* Q, double* err, int* flags) {
  return sawfem_sweep_chunk_share(plan, stream, op, B, omega, V, prec, 1, ir_tol, max_ir, fallback, workspace,
                                  workspace_bytes, Q, err, flags);
}

int sawfem_sweep_chunk_share(void* plan, void* stream, const sawfem_operator* op, int B, const double* omega,
                             double V, int prec, int share, double ir_tol, int max_ir, int fallback, void* workspace,
                             int64_t workspace_bytes, double* Q, double* err, int* flags) {
  if (!plan || !op || !omega || !workspace || !Q || !err || !flags || B <= 0 || op->n <= 0 || !op->indptr ||
      !op->indices || !op->Kaa || !op->Maa || !op->rK || !op->rM || !op->kts_a || (prec != 0 && prec != 1) ||
      max_ir < 1 || !(ir_tol > 0) || !(V > 0) || share < 1) {
    sawfem_set_error("sawfem_sweep_chunk: invalid arguments");
    return 1;
  }
  if ((op->Daa == nullptr) != (op->rD == nullptr)) {
    sawfem_set_error("sawfem_sweep_chunk: Daa and rD must be given together");
    return 1;
  }
  hipStream_t s = (hipStream_t)stream;
  std::vector<unsigned char> bad(B, 0);
  int steps = 0;
  const int G = prec == 0 ? std::min(share, B) : 1;
  SAW_CALL(refine_pass(plan, s, op, B, omega, V, prec, ir_tol, max_ir, (unsigned char*)workspace, workspace_bytes, Q,
                       err, bad.data(), &steps, G));
  for (int b = 0; b < B; ++b) flags[b] = bad[b] ? 2 : 0;       // bit 1: not converged in this precision
  bool refactored = false;
  if (prec == 0 && fallback && G > 1) {
    // shared factors that did not converge: the frequency's own complex64 factors first
    std::vector<int> todo;
    for (int b = 0; b < B; ++b)
      if (bad[b]) todo.push_back(b);
    int nfit = 0;
    while (nfit < (int)todo.size() && sawfem_sweep_chunk_bytes(plan, op->n, nfit + 1, 0) <= workspace_bytes) ++nfit;
    for (size_t c0 = 0; nfit > 0 && c0 < todo.size(); c0 += nfit) {
      const int k = (int)std::min<size_t>(nfit, todo.size() - c0);
      std::vector<double> om(k), q(2 * k), e(k);
      std::vector<unsigned char> bb(k);
      for (int j = 0; j < k; ++j) om[j] = omega[todo[c0 + j]];
      SAW_CALL(refine_pass(plan, s, op, k, om.data(), V, 0, ir_tol, max_ir, (unsigned char*)workspace,
                           workspace_bytes, q.data(), e.data(), bb.data(), nullptr, 1));
      for (int j = 0; j < k; ++j) {
        const int b = todo[c0 + j];
        Q[2 * b] = q[2 * j];
        Q[2 * b + 1] = q[2 * j + 1];
        err[b] = e[j];
        bad[b] = bb[j];
        flags[b] = 4 | (bb[j] ? 2 : 0);                          // bit 2: own complex64 factors
      }
    }
  }
  if (prec == 0 && fallback) {
    std::vector<int> todo;
    for (int b = 0; b < B; ++b)
      if (bad[b]) todo.push_back(b);
    // complex128 chunks that fit the workspace
    int nfit = 0;
    while (nfit < (int)todo.size() &&
           sawfem_sweep_chunk_bytes(plan, op->n, nfit + 1, 1) <= workspace_bytes) ++nfit;
    if (!todo.empty() && nfit == 0) {
      sawfem_set_error("sawfem_sweep_chunk: workspace too small for the complex128 refactorization");
      return 4;
    }
    for (size_t c0 = 0; c0 < todo.size(); c0 += nfit) {
      const int k = (int)std::min<size_t>(nfit, todo.size() - c0);
      std::vector<double> om(k), q(2 * k), e(k);
      std::vector<unsigned char> bb(k);
      for (int j = 0; j < k; ++j) om[j] = omega[todo[c0 + j]];
      SAW_CALL(refine_pass(plan, s, op, k, om.data(), V, 1, ir_tol, max_ir, (unsigned char*)workspace,
                           workspace_bytes, q.data(), e.data(), bb.data(), nullptr));
      for (int j = 0; j < k; ++j) {
        const int b = todo[c0 + j];
        Q[2 * b] = q[2 * j];
        Q[2 * b + 1] = q[2 * j + 1];
        err[b] = e[j];
        flags[b] = (flags[b] & 4) | 1 | (bb[j] ? 2 : 0);         // bit 0: refactored in complex128
      }
    }
    refactored = !todo.empty();
  }
  for (int b = 0; b < 2 * B; ++b)
    if (!std::isfinite(Q[b])) {
      sawfem_set_error("sawfem_sweep_chunk: non-finite charge");
      return 5;
    }
  return refactored ? 6 : 0;
}

}  // extern "C"
