// Multifrontal numeric LU of the Bloch-reduced system A(w) = K + i w beta D - w^2 M,
// batched over frequencies, in complex FP32 (mixed path) or complex FP64 (plain path).
//
// Replaces the per-frequency sparse direct solve behind FrequencySolver._backend()
// (SawSim saw2d/sweep.py:326-333, 349-353: PARDISO or SuperLU on the real-block form).
// The symbolic analysis (ordering, fronts, index maps) is built once on the host.
//
// Storage: the numeric arena holds every front of every frequency of the batch as a
// dense column-major m x m matrix; front f, frequency b lives at
//     arena + off[f] * B + b * m_f^2.
// After factorization the first p columns hold L (unit lower, below the diagonal) and
// U (upper), rows/columns p.. of the trailing block hold the Schur complement that the
// parent assembles (extend-add), and U12 = F[0:p, p:m], L21 = F[p:m, 0:p].
//
// Per level (leaves first), per front-size bucket:
//   k_leaf_front    : childless fronts: zeros and their A(w) entries in one pass
//   k_front_init    : F = sum of children's contribution blocks (gather, no atomics)
//   k_front_entries : F += A(w) entries owned by the other fronts (unique targets)
//   for each pivot panel [k0, k0 + kb) (left-looking):
//     k_gemm mode 4 : the panel's column and row parts -= the factored panels before it (MFMA)
//     k_diag        : LU of the kb x kb diagonal block (no pivoting, tiny-pivot guard), replaced
//                     in place by the inverses of its triangles
//     k_gemm 2 / 3  : L panel = A U_D^-1, U panel = L_D^-1 A  (MFMA)
//   k_gemm mode 1   : Schur complement F22 -= L21 U12 with K = p  (MFMA)
// Launches walk host-built work lists (one entry per workgroup); the buckets of a level run
// their panel phases concurrently on side streams between a fork and a join.
#include "common.h"
#include "../../include/sawfem.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <type_traits>
#include <vector>

namespace {

constexpr int TILE = 64;   // C tile edge of the GEMM workgroup (complex entries)
constexpr int KC = 16;     // K chunk staged in LDS
constexpr int SHORT_K = 128;   // GEMM reductions up to this length prefetch their C tile

struct FrontsDev {
  const int* m;
  const int* p;
  const int64_t* off;      // front offset (entries) for one frequency
  const int64_t* uoff;     // update-vector offset for one frequency
  const int64_t* rows_ptr;
  const int* rows;
  const int* child_ptr;    // children of f: child_ids[child_ptr[f] .. child_ptr[f+1])
  const int* child_ids;
  const int64_t* inv_ptr;  // per child slot: parent-row -> child CB index map (length m_parent)
  const int* inv;
  // solves only.  A solve launch works on blocks of R vectors (R = the kernels' template
  // argument); block b is vectors b R .. b R + R - 1 of the vector array.  Block b reads the
  // factors of arena slot fslot[b] of an arena of Bf factorizations, or slot b / vdiv when
  // fslot is null and vdiv > 0 (uniform groups), or slot b of an arena of B (one factorization
  // per block).  Grid row y is block sel[y / cpg] * cpg + y % cpg with a selection list (cpg
  // blocks per selected entry), else block boff + y.  nvec: vectors of the whole call (the
  // update vectors of front f start at uoff[f] * nvec).  Vector layout (xat, uat): vector-major
  // with xld = 0 (x[v][n], a front's update vectors [v][q]); with xld > 0 row-major (x[n][xld],
  // a front's update vectors [q][nvec]), in which the vectors of one row are contiguous.
  const int* fslot = nullptr;
  int Bf = 0;
  int vdiv = 0;
  int cpg = 1;
  int boff = 0;
  int64_t nvec = 0;
  int64_t xld = 0;
};

// Index of entry (vector v, row i) of the solve vectors x (n rows) and of entry (vector v,
// contribution-block row j) of a front's update vectors (q rows, from uarena + uoff[f] nvec).
__device__ __forceinline__ int64_t xat(const FrontsDev& F, int n, int64_t v, int64_t i) {
  return F.xld ? i * F.xld + v : v * n + i;
}
__device__ __forceinline__ int64_t uat(const FrontsDev& F, int q, int64_t v, int64_t j) {
  return F.xld ? j * F.nvec + v : v * q + j;
}
// The same with the layout fixed at compile time (XR: row-major).
template <bool XR>
__device__ __forceinline__ int64_t xat_t(const FrontsDev& F, int n, int64_t v, int64_t i) {
  return XR ? i * F.xld + v : v * n + i;
}
template <bool XR>
__device__ __forceinline__ int64_t uat_t(const FrontsDev& F, int q, int64_t v, int64_t j) {
  return XR ? j * F.nvec + v : v * q + j;
}

template <typename T>
__device__ __forceinline__ typename cplx<T>::type* front_ptr(typename cplx<T>::type* arena, const FrontsDev& F,
                                                             int f, int b, int B) {
  const int64_t m = F.m[f];
  return arena + F.off[f] * (int64_t)B + (int64_t)b * m * m;
}

// The factors a solve of right-hand side b reads (see FrontsDev::fslot).
template <typename T>
__device__ __forceinline__ typename cplx<T>::type* factor_ptr(typename cplx<T>::type* arena, const FrontsDev& F,
                                                              int f, int b, int B) {
  if (F.fslot) return front_ptr<T>(arena, F, f, F.fslot[b], F.Bf);
  if (F.vdiv > 0) return front_ptr<T>(arena, F, f, b / F.vdiv, F.Bf);
  return front_ptr<T>(arena, F, f, b, B);
}

// Vector block of grid row y of a solve launch (see FrontsDev).
__device__ __forceinline__ int solve_block(const FrontsDev& F, const int* sel, int y) {
  return sel ? sel[y / F.cpg] * F.cpg + y % F.cpg : F.boff + y;
}

// Strided loop over idx = first, first + step, ... < count with its global loads issued U at a
// time: load(idx) for U indices (clamped to count - 1, so every address is one the loop
// visits), then use(idx, value) for those in range.  A plain loop waits for each load before
// the next iteration; here U loads are in flight together.  count must be positive when
// first < count.
template <int U, typename Load, typename Use>
__device__ __forceinline__ void batched(int first, int count, int step, Load load, Use use) {
  for (int base = first; base < count; base += U * step) {
    decltype(load(0)) a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) a[u] = load(min(base + u * step, count - 1));
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (base + u * step < count) use(base + u * step, a[u]);
  }
}

// ---------------------------------------------------------------- front assembly
// Front f = sum of its children's contribution blocks (zero elsewhere), columns [c0, c1) of
// one frequency per workgroup (lvl: work list front << 12 | column chunk; frequency
// blockIdx.y).  Threads cover rows in slabs of RS (64, 128 or 256) and CG = 256 / RS columns at
// a time.  Two children (nested dissection): the parent-row -> child-CB
// maps of a thread's row stay in registers and the values of CJ columns are loaded
// unconditionally from clamped in-front addresses before any is selected or stored, so the
// loads are in flight together; other child counts take a generic loop.
template <typename T>
__device__ __forceinline__ void front_init_body(FrontsDev F, const int* __restrict__ lvl, int B,
                                                typename cplx<T>::type* __restrict__ arena, int cols_per_wg, int b) {
  using V = typename cplx<T>::type;
  constexpr int CJ = 8;   // columns per thread whose loads are issued before any store
  const int f = lvl[blockIdx.x] >> 12;
  const int m = F.m[f];
  const int c0 = (lvl[blockIdx.x] & 4095) * cols_per_wg;
  if (c0 >= m) return;
  const int c1 = min(m, c0 + cols_per_wg);
  V* Fb = front_ptr<T>(arena, F, f, b, B);
  const int cp0 = F.child_ptr[f];
  const int nch = F.child_ptr[f + 1] - cp0;
  const int RS = m <= 64 ? 64 : (m <= 128 ? 128 : 256);
  const int CG = 256 / RS;
  const int rsub = threadIdx.x % RS, cg = threadIdx.x / RS;
  if (nch == 0) {                       // leaf: zeros
    for (int j = c0 + cg; j < c1; j += CG)
      for (int i = rsub; i < m; i += RS) Fb[i + (int64_t)j * m] = czero<V>();
    return;
  }
  if (nch == 2) {                       // nested dissection: two children, loads batched
    const V* cb[2];
    const int* cinv[2];
    int cmc[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int ch = F.child_ids[cp0 + c];
      const int mc = F.m[ch], pc = F.p[ch];
      cb[c] = front_ptr<T>(arena, F, ch, b, B) + pc + (int64_t)pc * mc;
      cinv[c] = F.inv + F.inv_ptr[cp0 + c];
      cmc[c] = mc;
    }
    for (int i0 = 0; i0 < m; i0 += RS) {
      const int i = i0 + rsub;
      const bool row_ok = i < m;
      const int ir = min(i, m - 1);
      const int ci0 = cinv[0][ir], ci1 = cinv[1][ir];
      for (int jb = c0 + cg; jb < c1; jb += CG * CJ) {
        int cj0[CJ], cj1[CJ];
        V w0[CJ], w1[CJ];
#pragma unroll
        for (int q = 0; q < CJ; ++q) {
          const int j = min(jb + q * CG, c1 - 1);
          cj0[q] = cinv[0][j];
          cj1[q] = cinv[1][j];
        }
#pragma unroll
        for (int q = 0; q < CJ; ++q) {
          w0[q] = cb[0][max(ci0, 0) + (int64_t)max(cj0[q], 0) * cmc[0]];
          w1[q] = cb[1][max(ci1, 0) + (int64_t)max(cj1[q], 0) * cmc[1]];
        }
#pragma unroll
        for (int q = 0; q < CJ; ++q) {
          const int j = jb + q * CG;
          V v = czero<V>();
          if (ci0 >= 0 && cj0[q] >= 0) { v.x += w0[q].x; v.y += w0[q].y; }
          if (ci1 >= 0 && cj1[q] >= 0) { v.x += w1[q].x; v.y += w1[q].y; }
          if (row_ok && j < c1) Fb[i + (int64_t)j * m] = v;
        }
      }
    }
    return;
  }
  // any other number of children: one entry at a time
  for (int j = c0 + cg; j < c1; j += CG)
    for (int i = rsub; i < m; i += RS) {
      V v = czero<V>();
      for (int c = 0; c < nch; ++c) {
        const int slot = cp0 + c;
        const int* inv = F.inv + F.inv_ptr[slot];
        const int ci = inv[i], cj = inv[j];
        if (ci >= 0 && cj >= 0) {
          const int ch = F.child_ids[slot];
          const int64_t mc = F.m[ch], pc = F.p[ch];
          const V u = front_ptr<T>(arena, F, ch, b, B)[(pc + ci) + (pc + cj) * mc];
          v.x += u.x; v.y += u.y;
        }
      }
      Fb[i + (int64_t)j * m] = v;
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_front_init(FrontsDev F, const int* __restrict__ lvl, int B,
                                                    typename cplx<T>::type* __restrict__ arena, int cols_per_wg) {
  front_init_body<T>(F, lvl, B, arena, cols_per_wg, blockIdx.y);

}

// Entries of the fronts of one level that go through the panel path (skip[f] != 0: the front
// is assembled by k_small_front).  One entry per thread and ENT_FB frequencies per thread:
// the FP64 K, M, D values are read once for those frequencies; consecutive entries of a front
// are consecutive column-major offsets, so the front updates are coalesced.
constexpr int ENT_FB = 16;

template <typename T>
__global__ void __launch_bounds__(256) k_front_entries(FrontsDev F, int64_t e0, int64_t e1, int B,
                                const double* __restrict__ omega, double beta, const int* __restrict__ ent_front,
                                const int64_t* __restrict__ ent_slot, const int* __restrict__ ent_loc,
                                const double2* __restrict__ Kaa, const double2* __restrict__ Maa,
                                const double2* __restrict__ Daa, typename cplx<T>::type* __restrict__ arena,
                                const unsigned char* __restrict__ skip) {
  using V = typename cplx<T>::type;
  const int64_t e = e0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= e1) return;
  const int f = ent_front[e];
  if (skip[f]) return;
  const int64_t sl = ent_slot[e];
  const double2 k = Kaa[sl], mm = Maa[sl];
  const double2 d = Daa ? Daa[sl] : make_double2(0, 0);
  const int64_t loc = ent_loc[e];
  const int64_t msq = (int64_t)F.m[f] * F.m[f];
  V* base = arena + F.off[f] * (int64_t)B + loc;
  const int b0 = blockIdx.y * ENT_FB, b1 = min(B, b0 + ENT_FB);
  for (int b = b0; b < b1; ++b) {
    const double w = omega[b];
    const double ax = k.x - w * w * mm.x - w * beta * d.y;
    const double ay = k.y - w * w * mm.y + w * beta * d.x;
    V* c = base + (int64_t)b * msq;
    V cur = *c;
    cur.x += (T)ax;
    cur.y += (T)ay;
    *c = cur;
  }
}

constexpr int LEAF_MAXPOS = 8192;   // largest column chunk (positions) k_leaf_front maps in LDS
constexpr int LEAF_FB = 16;         // frequencies per k_leaf_front workgroup

// Childless fronts (nested-dissection leaves) assembled in one pass: a workgroup takes columns
// [c0, c1) of one front for LEAF_FB frequencies (grid y).  The front's A_aa entries in those columns
// (ent[elo, ehi): entries of a front are sorted by column-major offset) are placed in an LDS map
// position -> entry, then each thread writes its positions for all B frequencies: the entry
// value K - w^2 M + i w beta D (rounded to T once) or zero.  This replaces the zero fill and the
// entries' read-modify-write pass for these fronts (k_front_entries skips them).
template <typename T>
__global__ void __launch_bounds__(256) k_leaf_front(FrontsDev F, const int* __restrict__ items,
                                                    const int64_t* __restrict__ erange, int cols_per_wg, int B,
                                                    const double* __restrict__ omega, double beta,
                                                    const int64_t* __restrict__ ent_slot, const int* __restrict__ ent_loc,
                                                    const double2* __restrict__ Kaa, const double2* __restrict__ Maa,
                                                    const double2* __restrict__ Daa,
                                                    typename cplx<T>::type* __restrict__ arena) {
  using V = typename cplx<T>::type;
  extern __shared__ int emap[];                  // [positions of the column chunk]
  const int b0 = blockIdx.y * LEAF_FB, b1 = min(B, b0 + LEAF_FB);
  const int f = items[blockIdx.x] >> 12;
  const int m = F.m[f];
  const int c0 = (items[blockIdx.x] & 4095) * cols_per_wg;
  const int c1 = min(m, c0 + cols_per_wg);
  const int base = c0 * m, npos = (c1 - c0) * m;
  const int64_t elo = erange[2 * blockIdx.x], ehi = erange[2 * blockIdx.x + 1];
  const int t = threadIdx.x;
  for (int i = t; i < npos; i += blockDim.x) emap[i] = -1;
  __syncthreads();
  for (int64_t e = elo + t; e < ehi; e += blockDim.x) emap[ent_loc[e] - base] = (int)(e - elo);
  __syncthreads();
  const int64_t msq = (int64_t)m * m;
  V* Fb = arena + F.off[f] * (int64_t)B + base;
  // four positions per thread at a time: their K, M, D stay in registers over the frequencies
  for (int g = 0; g < npos; g += 4 * 256) {
    double2 kv[4], mv[4], dv[4];
    bool has[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int pos = g + t + 256 * u;
      const int idx = pos < npos ? emap[pos] : -1;
      has[u] = idx >= 0;
      kv[u] = mv[u] = dv[u] = make_double2(0, 0);
      if (has[u]) {
        const int64_t sl = ent_slot[elo + idx];
        kv[u] = Kaa[sl];
        mv[u] = Maa[sl];
        if (Daa) dv[u] = Daa[sl];
      }
    }
    for (int b = b0; b < b1; ++b) {
      const double w = omega[b];
      V* Fw = Fb + (int64_t)b * msq;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int pos = g + t + 256 * u;
        if (pos < npos) {
          V v;
          v.x = (T)(kv[u].x - w * w * mv[u].x - w * beta * dv[u].y);
          v.y = (T)(kv[u].y - w * w * mv[u].y + w * beta * dv[u].x);
          Fw[pos] = v;
        }
      }
    }
  }
}

// ---------------------------------------------------------------- static pivoting
// Pivot guard (static pivoting with perturbation, Li & Demmel, SuperLU_DIST): no row exchanges;
// a pivot whose magnitude is below tau = sqrt(u_T) * s_g(w) (u_T the unit roundoff of the front
// precision) is replaced by tau with the pivot's own phase (tau itself for a zero or non-finite
// pivot) and counted.  sqrt(u_T) balances the perturbation against the element growth it
// allows (a smaller tau leaves factors whose rounding errors exceed the perturbation).  s_g(w) = |K|_g + w^2 |M|_g +
// w beta |D|_g bounds the largest entry of row g of A(w) (|X|_g: the largest magnitude in row g
// of X, computed per factorization by k_row_scale).  The scale is per row because the rows of
// this system span many orders of magnitude (displacement rows carry the stiffness scaling,
// potential rows the permittivities): a matrix- or front-wide norm would flag every potential
// pivot.  Refinement against the FP64 operator (or the complex128 refactorization) absorbs
// the perturbation.
struct PivotGuard {
  const double* rs;          // [3][n]: row maxima of |K|, |M|, |D| (D: zero when absent)
  const double* omega;
  double beta, eps;
  int n;
  int* npert;                // replaced pivots
  unsigned* pmin;            // smallest |pivot| / s_g seen, as float bits (non-negative floats
                             // order like their bit patterns)
};

__device__ __forceinline__ double row_scale_at(const PivotGuard& G, int g, double w) {
  return G.rs[g] + w * w * G.rs[G.n + g] + w * fabs(G.beta) * G.rs[2 * G.n + g];
}

// Returns the pivot to use; perturbed is set when it differs from piv.
template <typename V>
__device__ __forceinline__ V guard_pivot(V piv, double scale, double eps, double* ratio, bool* perturbed) {
  const double mag = sqrt((double)piv.x * piv.x + (double)piv.y * piv.y);
  const double tau = eps * scale;
  *ratio = scale > 0 ? mag / scale : 0.0;
  *perturbed = !(mag >= tau) || !(mag < 1e300);
  if (!*perturbed) return piv;
  if (mag > 0 && mag < 1e300) {
    const double f = tau / mag;
    V r; r.x = (decltype(piv.x))(piv.x * f); r.y = (decltype(piv.y))(piv.y * f); return r;
  }
  V r; r.x = (decltype(piv.x))tau; r.y = 0; return r;
}

__device__ __forceinline__ void note_pivots(const PivotGuard& G, double rmin, int nrep) {
  if (nrep) atomicAdd(G.npert, nrep);
  const float r = (float)rmin;
  // most blocks find a smaller minimum already recorded and skip the atomic
  if (r < __uint_as_float(*(volatile unsigned*)G.pmin)) atomicMin(G.pmin, __float_as_uint(r));
}

__global__ void k_fill_f64(int64_t n, double* __restrict__ x, double v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

// Row maxima of |K|, |M|, |D| over the A_aa entries (each entry is listed once, with the
// front and the local position that give its row); max of non-negative doubles through
// their bit patterns, so the result does not depend on the order of the atomics.
__global__ void k_row_scale(int64_t nent, FrontsDev F, const int* __restrict__ ent_front,
                            const int64_t* __restrict__ ent_slot, const int* __restrict__ ent_loc,
                            const double2* __restrict__ Kaa, const double2* __restrict__ Maa,
                            const double2* __restrict__ Daa, int n, double* __restrict__ rs) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nent) return;
  const int f = ent_front[e];
  const int m = F.m[f];
  const int g = F.rows[F.rows_ptr[f] + ent_loc[e] % m];
  const int64_t sl = ent_slot[e];
  auto put = [&](double* dst, double2 v) {
    const double a = sqrt(v.x * v.x + v.y * v.y);
    if (a > 0) atomicMax((unsigned long long*)dst, (unsigned long long)__double_as_longlong(a));
  };
  put(rs + g, Kaa[sl]);
  put(rs + n + g, Maa[sl]);
  if (Daa) put(rs + 2 * n + g, Daa[sl]);
}

// ---------------------------------------------------------------- panel factorization
// In-place inversion of the unit-lower and upper triangles of an n x n block D (column-major,
// leading dimension ld, in LDS) by row operations (Gauss-Jordan):
//   L^{-1}: step k, rows i > k:  X[i][j] -= L[i][k] X[k][j] (j < k), X[i][k] = -L[i][k]
//   U^{-1}: step k = n-1..0: row k scaled by d = 1/U[k][k] (phase a), then rows i < k:
//     Y[i][j] -= U[i][k] Y[k][j] (j > k), Y[i][k] = -U[i][k] d (phase b); the diagonal
//     entries become 1/U[k][k] at the end
// n <= 64 and at least 64 threads when n > 32; all threads of the block call it (it synchronizes).
// Thread t owns row t % RW (RW = 32 or 64, a power of two >= n) and every (nt / RW)-th column of
// a step's update, so no index needs a division.
template <typename V>
__device__ void invert_lu_block(V* D, int ld, int n) {
  // In-place inverses of the unit-lower L and upper U held in D (LDS): afterwards the strict
  // lower triangle holds L^{-1} (unit diagonal implied) and the upper triangle U^{-1}.  Step st
  // eliminates column st of L and column n-1-st of U (Gauss-Jordan); the two halves touch
  // disjoint triangles.
  const int t = threadIdx.x, nt = blockDim.x;
  const int rs = n <= 32 ? 5 : 6;
  const int i = t & ((1 << rs) - 1), tc = t >> rs, ncg = nt >> rs;
  const V one{1, 0};
  for (int st = 0; st < n; ++st) {
    const int kL = st, kU = n - 1 - st;
    const V dU = cdiv(one, D[kU + kU * ld]);          // U[kU][kU] stays in place until the end
    // a: row kU of U right of the diagonal scaled by 1 / U[kU][kU]
    for (int j = kU + 1 + t; j < n; j += nt) D[kU + j * ld] = cmul(D[kU + j * ld], dU);
    __syncthreads();
    // b: L rows i > kL, columns j < kL and U rows i < kU, columns j > kU
    if (i > kL && i < n) {
      const V l = D[i + kL * ld];
      for (int j = tc; j < kL; j += ncg) D[i + j * ld] = cfms(D[i + j * ld], l, D[kL + j * ld]);
    }
    if (i < kU) {
      const V u = D[i + kU * ld];
      for (int j = kU + 1 + tc; j < n; j += ncg) D[i + j * ld] = cfms(D[i + j * ld], u, D[kU + j * ld]);
    }
    __syncthreads();
    // c: the eliminated columns
    if (tc == 0 && i > kL && i < n) {
      const V v = D[i + kL * ld];
      D[i + kL * ld] = V{-v.x, -v.y};
    }
    if (tc == ncg - 1 && i < kU) {
      const V y = cmul(D[i + kU * ld], dU);
      D[i + kU * ld] = V{-y.x, -y.y};
    }
    __syncthreads();
  }
  for (int k = t; k < n; k += nt) D[k * (ld + 1)] = cdiv(one, D[k * (ld + 1)]);
  __syncthreads();
}

// LU of the kb x kb diagonal block of pivot panel [k0, k0 + kb) (no pivoting; small pivots
// perturbed by the pivot guard), then the inverses of its unit-lower and upper
// factors stored IN PLACE of the factors: the strictly lower triangle of the block holds
// L_D^{-1} (unit diagonal implied) and the upper triangle U_D^{-1}.  Neither the panel solves
// (GEMMs with these inverses) nor the triangular solves need L_D, U_D themselves.  Everything
// happens in LDS; the block is written back once.
template <typename T>
__device__ __forceinline__ void diag_body(FrontsDev F, const int* __restrict__ lvl, int B,
                                          typename cplx<T>::type* __restrict__ arena, int k0, int nb,
                                          const PivotGuard& PG, int b) {
  using V = typename cplx<T>::type;
  extern __shared__ unsigned char smem_raw[];
  V* Ds = reinterpret_cast<V*>(smem_raw);   // [nb][nb+1], column-major with padding
  double* rsc = reinterpret_cast<double*>(Ds + nb * (nb + 1));   // [nb] row scales of the pivots
  const int f = lvl[blockIdx.x];
  const int p = F.p[f];
  if (k0 >= p) return;
  const int m = F.m[f];
  const int kb = min(nb, p - k0);
  const int ld = nb + 1;
  V* Fb = front_ptr<T>(arena, F, f, b, B);
  V* G = Fb + (int64_t)k0 + (int64_t)k0 * m;
  // thread t: row i = t % RW of the block (RW = 32 or 64 >= kb), columns tc, tc + ncg, ...
  const int t = threadIdx.x, nt = blockDim.x;
  const int rs = nb <= 32 ? 5 : 6;
  const int i = t & ((1 << rs) - 1), tc = t >> rs, ncg = nt >> rs;
  const int ic = min(i, kb - 1);
  batched<8>(tc, kb, ncg, [&](int j) { return G[ic + (int64_t)j * m]; },
             [&](int j, V v) { if (i < kb) Ds[i + j * ld] = v; });
  if (tc == 0 && i < kb) rsc[i] = row_scale_at(PG, F.rows[F.rows_ptr[f] + k0 + i], PG.omega[b]);
  __syncthreads();
  // right-looking LU: multipliers of column k, then the rank-1 update of the trailing block
  double rmin = 3.0e38;
  int nrep = 0;
  for (int k = 0; k < kb; ++k) {
    double ratio;
    bool pert;
    const V piv = guard_pivot(Ds[k + k * ld], rsc[k], PG.eps, &ratio, &pert);
    rmin = fmin(rmin, ratio);
    if (pert) {
      ++nrep;
      __syncthreads();           // every thread has read the old pivot
      if (t == 0) Ds[k + k * ld] = piv;
    }
    if (tc == 0 && i > k && i < kb) Ds[i + k * ld] = cdiv(Ds[i + k * ld], piv);
    __syncthreads();
    if (i > k && i < kb) {
      const V l = Ds[i + k * ld];
      for (int j = k + 1 + tc; j < kb; j += ncg) Ds[i + j * ld] = cfms(Ds[i + j * ld], l, Ds[k + j * ld]);
    }
    __syncthreads();
  }
  if (t == 0) note_pivots(PG, rmin, nrep);
  invert_lu_block(Ds, ld, kb);
  if (i < kb)
    for (int j = tc; j < kb; j += ncg) G[i + (int64_t)j * m] = Ds[i + j * ld];
}

template <typename T>
__global__ void __launch_bounds__(512) k_diag(FrontsDev F, const int* __restrict__ lvl, int B,
                                              typename cplx<T>::type* __restrict__ arena, int k0, int nb,
                                              const PivotGuard PG) {
  diag_body<T>(F, lvl, B, arena, k0, nb, PG, blockIdx.y);
}

// Diagonal block and both panel solves of a 32-wide pivot panel in one wave per front and
// frequency (fronts of moderate order, where the three separate launches are latency-bound):
// diag_body factors and inverts the kb x kb block in LDS, then each lane takes one row of the
// column panel, L21[i, :] = A21[i, :] U_D^{-1}, or one column of the row panel,
// U12[:, j] = L_D^{-1} A12[:, j]: the kb inputs are loaded into registers (the row's loads are
// coalesced across lanes, a column is contiguous), the triangular products read the inverse
// from LDS (the same address on every lane), and the results overwrite the inputs.
constexpr int PANEL_LU_NB = 32;

template <typename T>
__global__ void __launch_bounds__(64) k_panel_lu(FrontsDev F, const int* __restrict__ lvl, int B,
                                                 typename cplx<T>::type* __restrict__ arena, int k0,
                                                 const PivotGuard PG) {
  using V = typename cplx<T>::type;
  constexpr int NB = PANEL_LU_NB;
  const int b = blockIdx.y;
  diag_body<T>(F, lvl, B, arena, k0, NB, PG, b);
  const int f = lvl[blockIdx.x];
  const int p = F.p[f];
  if (k0 >= p) return;
  const int m = F.m[f];
  const int kb = min(NB, p - k0), k1 = k0 + kb;
  const int q = m - k1;
  extern __shared__ unsigned char smem_raw[];
  const V* Ds = reinterpret_cast<const V*>(smem_raw);   // L_D^{-1} below, U_D^{-1} on/above
  constexpr int ld = NB + 1;
  V* Fb = front_ptr<T>(arena, F, f, b, B);
  for (int it = threadIdx.x; it < 2 * q; it += 64) {
    V a[NB];
    if (it < q) {                 // row i of the column panel
      const int i = k1 + it;
#pragma unroll
      for (int k = 0; k < NB; ++k) a[k] = k < kb ? Fb[(int64_t)i + (int64_t)(k0 + k) * m] : czero<V>();
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        if (c < kb) {
          V acc = czero<V>();
#pragma unroll
          for (int k = 0; k <= c; ++k) acc = cfma(acc, a[k], Ds[k + c * ld]);
          Fb[(int64_t)i + (int64_t)(k0 + c) * m] = acc;
        }
      }
    } else {                      // column j of the row panel
      const int j = k1 + (it - q);
      const V* col = Fb + (int64_t)k0 + (int64_t)j * m;
#pragma unroll
      for (int r = 0; r < NB; ++r) a[r] = r < kb ? col[r] : czero<V>();
#pragma unroll
      for (int r = 0; r < NB; ++r) {
        if (r < kb) {
          V acc = a[r];
#pragma unroll
          for (int k = 0; k < r; ++k) acc = cfma(acc, Ds[r + k * ld], a[k]);
          Fb[(int64_t)(k0 + r) + (int64_t)j * m] = acc;
        }
      }
    }
  }
}

// ---------------------------------------------------------------- small fronts, fused
// One workgroup of 512 threads per (front, frequency) for fronts whose m x m block fits in
// LDS (m <= 128): the whole factorization of the front happens there.
//  1. assembly: zero, extend-add of each child's contribution block (scatter through the
//     child-CB -> parent-row map rebuilt in LDS, children one after the other), then the
//     A = K - w^2 M + i w beta D entries (FP64, rounded to T once);
//  2. partial LU of the p pivots, right-looking with rank-1 updates of the whole trailing
//     front (the Schur complement included); pivot guard as in k_diag;
//  3. write-back of everything but the nbl x nbl diagonal blocks of the pivot triangle;
//     those are inverted in place (L_D^{-1} strictly below, U_D^{-1} on/above the diagonal,
//     the storage the panel path and the solves use) and written last.
constexpr int SMALL_NT = 512;
constexpr int SMALL_MMAX = 128;

template <typename T>
__global__ void __launch_bounds__(SMALL_NT) k_small_front(
    FrontsDev F, const int* __restrict__ lvl, int B, const double* __restrict__ omega, double beta,
    const int64_t* __restrict__ fent_lo, const int64_t* __restrict__ fent_hi, const int64_t* __restrict__ ent_slot,
    const int* __restrict__ ent_loc, const double2* __restrict__ Kaa, const double2* __restrict__ Maa,
    const double2* __restrict__ Daa, typename cplx<T>::type* __restrict__ arena, int nbl,
    const PivotGuard PG) {
  using V = typename cplx<T>::type;
  extern __shared__ unsigned char smem_raw[];
  const int f = lvl[blockIdx.x];
  const int b = blockIdx.y;
  const int m = F.m[f], p = F.p[f];
  const int ld = m + 1;
  V* S = reinterpret_cast<V*>(smem_raw);                                // [m][ld] column-major
  int* rel = reinterpret_cast<int*>(smem_raw + sizeof(V) * (size_t)m * ld);   // [m]
  const int t = threadIdx.x;
  for (int idx = t; idx < m * ld; idx += SMALL_NT) S[idx] = czero<V>();
  __syncthreads();
  // 1. assembly
  for (int slot = F.child_ptr[f]; slot < F.child_ptr[f + 1]; ++slot) {
    const int ch = F.child_ids[slot];
    const int mc = F.m[ch], pc = F.p[ch], qc = mc - pc;
    const int* inv = F.inv + F.inv_ptr[slot];
    for (int i = t; i < m; i += SMALL_NT) {
      const int ci = inv[i];
      if (ci >= 0) rel[ci] = i;
    }
    __syncthreads();
    const V* Fc = front_ptr<T>(arena, F, ch, b, B) + pc + (int64_t)pc * mc;
    for (int idx = t; idx < qc * qc; idx += SMALL_NT) {
      const int a = idx % qc, c = idx / qc;
      const V w = Fc[a + (int64_t)c * mc];
      V* d = S + rel[a] + rel[c] * ld;
      d->x += w.x;
      d->y += w.y;
    }
    __syncthreads();
  }
  {
    const double w = omega[b];
    for (int64_t e = fent_lo[f] + t; e < fent_hi[f]; e += SMALL_NT) {
      const int64_t sl = ent_slot[e];
      const double2 k = Kaa[sl], mm = Maa[sl];
      double ax = k.x - w * w * mm.x, ay = k.y - w * w * mm.y;
      if (Daa) {
        const double2 d = Daa[sl];
        ax -= w * beta * d.y;
        ay += w * beta * d.x;
      }
      const int loc = ent_loc[e];
      V* d = S + (loc % m) + (loc / m) * ld;
      d->x += (T)ax;
      d->y += (T)ay;
    }
  }
  __syncthreads();
  // 2. partial LU: wave w updates columns j = w (mod 8) four at a time (all loads issued before
  //    the stores), lane owns rows k+1+lane and k+65+lane
  const int lane = t & 63, wave = t >> 6;
  constexpr int NW = SMALL_NT / 64;
  const double wb = omega[b];
  double rmin = 3.0e38;
  int nrep = 0;
  for (int k = 0; k < p; ++k) {
    double ratio;
    bool pert;
    const V piv = guard_pivot(S[k + k * ld], row_scale_at(PG, F.rows[F.rows_ptr[f] + k], wb), PG.eps, &ratio,
                              &pert);
    rmin = fmin(rmin, ratio);
    if (pert) {
      ++nrep;
      __syncthreads();           // every thread has read the old pivot
      if (t == 0) S[k + k * ld] = piv;
    }
    const V one{1, 0};
    const V ipiv = cdiv(one, piv);
    for (int i = k + 1 + t; i < m; i += SMALL_NT) S[i + k * ld] = cmul(S[i + k * ld], ipiv);
    __syncthreads();
    const int i0 = k + 1 + lane, i1 = i0 + 64;
    const bool r0 = i0 < m, r1 = i1 < m;
    const V l0 = r0 ? S[i0 + k * ld] : czero<V>();
    const V l1 = r1 ? S[i1 + k * ld] : czero<V>();
    for (int j = k + 1 + wave; j < m; j += 4 * NW) {
      V u[4], a0[4], a1[4];
      int jj[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        jj[c] = j + c * NW;
        const bool ok = jj[c] < m;
        u[c] = ok ? S[k + jj[c] * ld] : czero<V>();
        a0[c] = (ok && r0) ? S[i0 + jj[c] * ld] : czero<V>();
        a1[c] = (ok && r1) ? S[i1 + jj[c] * ld] : czero<V>();
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        if (jj[c] < m) {
          if (r0) S[i0 + jj[c] * ld] = cfms(a0[c], l0, u[c]);
          if (r1) S[i1 + jj[c] * ld] = cfms(a1[c], l1, u[c]);
        }
      }
    }
    __syncthreads();
  }
  if (t == 0) note_pivots(PG, rmin, nrep);
  // 3. write-back (diagonal blocks of the pivot triangle last)
  V* G = front_ptr<T>(arena, F, f, b, B);
  for (int idx = t; idx < m * m; idx += SMALL_NT) {
    const int i = idx % m, j = idx / m;
    if (!(i < p && j < p && i / nbl == j / nbl)) G[idx] = S[i + j * ld];
  }
  // in-place inversion of each diagonal block, then its write-back
  for (int J0 = 0; J0 < p; J0 += nbl) {
    const int nbk = min(nbl, p - J0);
    V* D = S + J0 + J0 * ld;
    invert_lu_block(D, ld, nbk);
    for (int idx = t; idx < nbk * nbk; idx += SMALL_NT) {
      const int i = idx % nbk, j = idx / nbk;
      G[(int64_t)(J0 + i) + (int64_t)(J0 + j) * m] = D[i + j * ld];
    }
  }
}

// ---------------------------------------------------------------- complex GEMM on MFMA
// C[r, c] -= sum_k A[r, k] B[k, c] inside one front, for the C tile (r0.., c0..) of 64x64.
// The MFMA computes D = B^T A^T so that the lane-contiguous output index is the C row,
// which is contiguous in the column-major front (coalesced epilogue).
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

template <typename T> struct WaveAcc;

template <> struct WaveAcc<float> {
  // one 32x32 complex block per wave: D[i = col][j = row].  Three real products per complex
  // product (Gauss): P1 = Ar Br, P2 = Ai Bi, P3 = (Ar + Ai)(Br + Bi); then Re = P1 - P2 and
  // Im = P3 - P1 - P2.  Normwise as accurate as the four-product form; FP64 refinement of the
  // solves absorbs the difference.
  floatx16 p1, p2, p3;
  __device__ void zero() {
    for (int v = 0; v < 16; ++v) { p1[v] = 0.f; p2[v] = 0.f; p3[v] = 0.f; }
  }
  // As: [KC][TR] (k-major, rows contiguous); Bs: [TC][KC+1] (col-major over k)
  template <int TR>
  __device__ void step(const float2* As, const float2* Bs, int wrow, int wcol, int lane) {
#pragma unroll
    for (int kk = 0; kk < KC; kk += 2) {
      const int k = kk + (lane >> 5);
      const float2 bv = Bs[(wcol * 32 + (lane & 31)) * (KC + 1) + k];   // B[k][col]
      const float2 av = As[k * TR + wrow * 32 + (lane & 31)];          // A[row][k]
      p1 = __builtin_amdgcn_mfma_f32_32x32x2f32(bv.x, av.x, p1, 0, 0, 0);
      p2 = __builtin_amdgcn_mfma_f32_32x32x2f32(bv.y, av.y, p2, 0, 0, 0);
      p3 = __builtin_amdgcn_mfma_f32_32x32x2f32(bv.x + bv.y, av.x + av.y, p3, 0, 0, 0);
    }
  }
  template <typename Fn> __device__ void each(int lane, Fn fn) {
    // D[i][j]: j = lane % 32 (row), i = 8 (v / 4) + 4 (lane / 32) + v % 4 (col)
    for (int v = 0; v < 16; ++v) {
      const int i = 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
      fn(v, lane & 31, i, p1[v] - p2[v], p3[v] - p1[v] - p2[v]);
    }
  }
  // (row, col) of the lane's 16 outputs, in the order of each()
  template <typename Fn> static __device__ void coords(int lane, Fn fn) {
    for (int v = 0; v < 16; ++v) fn(v, lane & 31, 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3));
  }
};

template <> struct WaveAcc<double> {
  // 2x2 blocks of 16x16: blk[ci][rj]
  doublex4 re[2][2], im[2][2];
  __device__ void zero() {
    for (int a = 0; a < 2; ++a)
      for (int c = 0; c < 2; ++c)
        for (int v = 0; v < 4; ++v) { re[a][c][v] = 0.0; im[a][c][v] = 0.0; }
  }
  template <int TR>
  __device__ void step(const double2* As, const double2* Bs, int wrow, int wcol, int lane) {
#pragma unroll
    for (int kk = 0; kk < KC; kk += 4) {
      const int k = kk + (lane >> 4);
      double2 bv[2], av[2];
      for (int ci = 0; ci < 2; ++ci) bv[ci] = Bs[(wcol * 32 + ci * 16 + (lane & 15)) * (KC + 1) + k];
      for (int rj = 0; rj < 2; ++rj) av[rj] = As[k * TR + wrow * 32 + rj * 16 + (lane & 15)];
      for (int ci = 0; ci < 2; ++ci)
        for (int rj = 0; rj < 2; ++rj) {
          re[ci][rj] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[ci].x, av[rj].x, re[ci][rj], 0, 0, 0);
          re[ci][rj] = __builtin_amdgcn_mfma_f64_16x16x4f64(-bv[ci].y, av[rj].y, re[ci][rj], 0, 0, 0);
          im[ci][rj] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[ci].x, av[rj].y, im[ci][rj], 0, 0, 0);
          im[ci][rj] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[ci].y, av[rj].x, im[ci][rj], 0, 0, 0);
        }
    }
  }
  template <typename Fn> __device__ void each(int lane, Fn fn) {
    // f64 16x16x4 output: D[i][j] with j = lane % 16 (row of C), i = 4 v + lane / 16 (column of C)
    for (int ci = 0; ci < 2; ++ci)
      for (int rj = 0; rj < 2; ++rj)
        for (int v = 0; v < 4; ++v)
          fn((ci * 2 + rj) * 4 + v, rj * 16 + (lane & 15), ci * 16 + 4 * v + (lane >> 4), re[ci][rj][v], im[ci][rj][v]);
  }
  template <typename Fn> static __device__ void coords(int lane, Fn fn) {
    for (int ci = 0; ci < 2; ++ci)
      for (int rj = 0; rj < 2; ++rj)
        for (int v = 0; v < 4; ++v) fn((ci * 2 + rj) * 4 + v, rj * 16 + (lane & 15), ci * 16 + 4 * v + (lane >> 4));
  }
};

// mode 1: Schur complement rows/cols [p, m), K = [0, p)
// mode 2: L panel  F[k1:m, k0:k1] = F[k1:m, k0:k1] * Uinv      (in place, overwrite)
// mode 3: U panel  F[k0:k1, k1:m] = Linv * F[k0:k1, k1:m]      (in place, overwrite)
// mode 4: left-looking panel update before the diagonal LU of panel [k0, k1), K = [0, k0):
//         column panel F[k0:m, k0:k1] -= F[k0:m, 0:k0] F[0:k0, k0:k1]  (part 0 tiles)
//         row panel    F[k0:k1, k1:m] -= F[k0:k1, 0:k0] F[0:k0, k1:m]  (part 1 tiles)
// The inverses are the ones k_diag stored in the diagonal block.  Each workgroup takes one
// entry of a tile list (lvl): the front and the C tile.
// C tiles are TR x TC (4 waves of 32 x 32): 64 x 64 in general; 128 x 32 and 32 x 128 for the
// 32-wide panels of the small-front levels, so no wave works on columns outside the panel.
template <typename T, int mode, int TR = TILE, int TC = TILE, bool PRE = false>
__device__ __forceinline__ void gemm_body(FrontsDev F, const int* __restrict__ lvl, int B,
                                          typename cplx<T>::type* __restrict__ arena, int k0, int nb, int b) {
  using V = typename cplx<T>::type;
  static_assert((TR / 32) * (TC / 32) == 4, "four 32x32 waves per tile");
  __shared__ V As[2][KC * TR];
  __shared__ V Bs[2][TC * (KC + 1)];
  // lvl is a tile list, one entry per workgroup: front << 12 | tile row << 6 | tile column
  // (mode 1), front << 12 | tile (modes 2, 3), front << 12 | part << 11 | tile (mode 4)
  const int e = lvl[blockIdx.x];
  const int f = e >> 12;
  const int p = F.p[f];
  const int m = F.m[f];
  // C rows [r_lo, r_hi) x cols [c_lo, c_hi); reduction length K
  int r_lo, r_hi, c_lo, c_hi, K, kb = 0;
  int ti, tj;
  if (mode == 1) {
    if (p == 0) return;
    r_lo = c_lo = p; r_hi = c_hi = m; K = p;
  } else if (mode == 4) {
    if (k0 >= p) return;
    kb = min(nb, p - k0);
    const int k1 = k0 + kb;
    K = k0;
    if (!(e & 2048)) { r_lo = k0; r_hi = m; c_lo = k0; c_hi = k1; ti = e & 2047; tj = 0; }
    else { r_lo = k0; r_hi = k1; c_lo = k1; c_hi = m; ti = 0; tj = e & 2047; }
  } else {
    if (k0 >= p) return;
    kb = min(nb, p - k0);
    const int k1 = k0 + kb;
    if (mode == 2) { r_lo = k1; r_hi = m; c_lo = k0; c_hi = k1; K = kb; }
    else { r_lo = k0; r_hi = k1; c_lo = k1; c_hi = m; K = kb; }
  }
  if (r_lo >= r_hi || c_lo >= c_hi) return;
  const int ntr = (r_hi - r_lo + TR - 1) / TR, ntc = (c_hi - c_lo + TC - 1) / TC;
  if (mode == 1) { ti = (e >> 6) & 63; tj = e & 63; }
  else if (mode == 2) { ti = e & 4095; tj = 0; }
  else if (mode == 3) { ti = 0; tj = e & 4095; }
  if (ti >= ntr || tj >= ntc) return;
  const int r0 = r_lo + ti * TR, c0 = c_lo + tj * TC;
  V* Fb = front_ptr<T>(arena, F, f, b, B);
  const int t = threadIdx.x;
  const int wave = t >> 6, lane = t & 63;
  constexpr int WR = TR / 32;
  const int wrow = wave % WR, wcol = wave / WR;
  // Loads are issued unconditionally from clamped in-front addresses and selected when stored
  // to LDS, so a chunk's loads are all in flight together; the next chunk is loaded into
  // registers while the current one feeds the MFMAs (double-buffered LDS, one barrier a chunk).
  constexpr int NEA = KC * TR / 256, NEB = KC * TC / 256;
  V ra[NEA], rb[NEB];
  unsigned char sa[NEA], sb[NEB];   // 0: zero, 1: loaded value, 2: one (unit diagonal of Linv)
  auto load = [&](int kc) {
#pragma unroll
    for (int e = 0; e < NEA; ++e) {
      const int idx = t + 256 * e;
      const int row = idx % TR, k = idx / TR;
      const int gr = r0 + row, gk = kc + k;
      const bool ok = gr < r_hi && gk < K;
      const int grc = min(gr, r_hi - 1), gkc = min(gk, K - 1);
      int64_t addr;
      if (mode == 2) addr = (int64_t)grc + (int64_t)(k0 + gkc) * m;
      else if (mode == 1 || mode == 4) addr = (int64_t)grc + (int64_t)gkc * m;
      else addr = (int64_t)grc + (int64_t)(k0 + gkc) * m;                  // Linv[row][k]
      ra[e] = Fb[addr];
      if (mode == 3) {
        const int rr = gr - k0;
        sa[e] = !ok ? 0 : (rr > gk ? 1 : (rr == gk ? 2 : 0));
      } else {
        sa[e] = ok ? 1 : 0;
      }
    }
#pragma unroll
    for (int e = 0; e < NEB; ++e) {
      const int idx = t + 256 * e;
      const int k = idx % KC, col = idx / KC;
      const int gk = kc + k, gc = c0 + col;
      const bool ok = gc < c_hi && gk < K;
      const int gcc = min(gc, c_hi - 1), gkc = min(gk, K - 1);
      int64_t addr;
      if (mode == 1 || mode == 4) addr = (int64_t)gkc + (int64_t)gcc * m;
      else addr = (int64_t)(k0 + gkc) + (int64_t)gcc * m;                     // mode 2: Uinv[k][col]
      rb[e] = Fb[addr];
      if (mode == 2) sb[e] = (ok && gk <= gc - k0) ? 1 : 0;
      else sb[e] = ok ? 1 : 0;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int e = 0; e < NEA; ++e) {
      const int idx = t + 256 * e;
      const int row = idx % TR, k = idx / TR;
      As[buf][k * TR + row] = sa[e] == 1 ? ra[e] : (sa[e] == 2 ? V{1, 0} : czero<V>());
    }
#pragma unroll
    for (int e = 0; e < NEB; ++e) {
      const int idx = t + 256 * e;
      const int k = idx % KC, col = idx / KC;
      Bs[buf][col * (KC + 1) + k] = sb[e] ? rb[e] : czero<V>();
    }
  };
  WaveAcc<T> acc;
  acc.zero();
  const int nch = (K + KC - 1) / KC;
  load(0);
  // read-modify-write modes: with PRE (short K) the C tile is fetched with the first chunk
  // (clamped in-front addresses) so its latency overlaps the K loop instead of following it
  constexpr bool RMW = mode == 1 || mode == 4;
  V cpre[RMW && PRE ? 16 : 1];
  if (RMW && PRE)
    WaveAcc<T>::coords(lane, [&](int v, int j, int i) {
      const int gr = min(r0 + wrow * 32 + j, r_hi - 1), gc = min(c0 + wcol * 32 + i, c_hi - 1);
      cpre[v] = Fb[(int64_t)gr + (int64_t)gc * m];
    });
  store(0);
  __syncthreads();
  for (int c = 0; c < nch; ++c) {
    const int buf = c & 1;
    if (c + 1 < nch) load((c + 1) * KC);
    acc.template step<TR>(As[buf], Bs[buf], wrow, wcol, lane);
    if (c + 1 < nch) store(buf ^ 1);
    __syncthreads();
  }
  acc.each(lane, [&](int v, int j, int i, T re, T im) {
    const int gr = r0 + wrow * 32 + j;
    const int gc = c0 + wcol * 32 + i;
    if (gr < r_hi && gc < c_hi) {
      V* c = Fb + (int64_t)gr + (int64_t)gc * m;
      V o;
      if (RMW && PRE) {
        o = cpre[v];
        o.x -= re; o.y -= im;
      } else if (RMW) {
        o = *c;
        o.x -= re; o.y -= im;
      } else {
        o.x = re; o.y = im;
      }
      *c = o;
    }
  });
}

template <typename T, int mode, int TR = TILE, int TC = TILE, bool PRE = false>
__global__ void __launch_bounds__(256, (PRE && sizeof(T) == 4 && TC != 128) ? 3 : 1) k_gemm(FrontsDev F, const int* __restrict__ lvl, int B,
                                              typename cplx<T>::type* __restrict__ arena, int k0, int nb) {
  gemm_body<T, mode, TR, TC, PRE>(F, lvl, B, arena, k0, nb, blockIdx.z);

}

// ---------------------------------------------------------------- triangular solves
// Per level, each direction runs two kernels:
//   * a pivot-block kernel, one workgroup per (front, frequency): the p x p triangle walked in
//     the factorization's nbl-wide diagonal blocks, staged in LDS.  A diagonal block holds the
//     inverses of its triangular factors, so its solve is a mat-vec (all 256 threads, four
//     partial sums per row); the off-diagonal part of the triangle then updates the remaining
//     pivot rows;
//   * a panel kernel over 64-wide tiles of the off-diagonal panel (L21, U12 or their
//     transposes), many workgroups per front, coalesced loads (transposed panels go through
//     an LDS tile).
// A workgroup solves one block of R right-hand sides that share one factorization (the one or
// two vectors of a frequency, or those of a group of frequencies factored together): every
// factor entry is read once and applied to all R vectors.
// Vectors: x [nvec][n] complex128, block b = rows b R .. b R + R - 1 (pivot rows of each front
// are owned by that front); the update vectors of front f (its contribution to the parent's
// rows) live in uarena at uoff[f] * nvec + (b * R + r) * (m - p).
template <typename T> struct SolveTile;
template <> struct SolveTile<float> { static constexpr int SB = 64; };    // widest panel, complex64
template <> struct SolveTile<double> { static constexpr int SB = 32; };   // widest panel, complex128

__device__ __forceinline__ double2 d2add(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 d2sub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

// w[r][i] (i < p) = x_r[rows[i]] + children's update entries; uf_r[i - p] (i >= p) = children's
// entries.  w is [R][p] in LDS, or [R][m] holding the CB rows too when ALL is set.
template <int R, bool ALL>
__device__ __forceinline__ void solve_gather(const FrontsDev& F, int f, int b, int B, int n, int m, int p,
                                             const int* rows, const double2* x, double2* uarena, double2* w) {
  const int nch = F.child_ptr[f + 1] - F.child_ptr[f];
  const int q = m - p;
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    double2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = (i < p) ? x[xat(F, n, (int64_t)b * R + r, rows[i])] : make_double2(0, 0);
    for (int c = 0; c < nch; ++c) {
      const int slot = F.child_ptr[f] + c;
      const int ci = F.inv[F.inv_ptr[slot] + i];
      if (ci >= 0) {
        const int ch = F.child_ids[slot];
        const int qc = F.m[ch] - F.p[ch];
        const double2* u = uarena + F.uoff[ch] * F.nvec;
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = d2add(v[r], u[uat(F, qc, (int64_t)b * R + r, ci)]);
      }
    }
    if (ALL) {
#pragma unroll
      for (int r = 0; r < R; ++r) w[r * m + i] = v[r];
    } else if (i < p) {
#pragma unroll
      for (int r = 0; r < R; ++r) w[r * p + i] = v[r];
    } else {
      double2* uf = uarena + F.uoff[f] * F.nvec;
#pragma unroll
      for (int r = 0; r < R; ++r) uf[uat(F, q, (int64_t)b * R + r, i - p)] = v[r];
    }
  }
}

// Bytes of k_tri's tile region: the diagonal block (kinds 0/1) or a 64-row panel tile (kinds
// 2/3) of factor entries, and with FULL kind 1 at least the R x threads partial sums of the
// CB-to-pivot product, which use the region before the diagonal blocks do.  A multiple of 16.
__host__ __device__ inline size_t tri_tile_bytes(size_t vsize, int kind, int R, bool full, int nbl, int threads) {
  const size_t t = vsize * (size_t)(kind >= 2 ? 64 : nbl) * (size_t)(nbl + 1);
  const size_t red = (full && kind == 1) ? sizeof(double2) * (size_t)R * (size_t)threads : 0;
  return ((t > red ? t : red) + 15) / 16 * 16;
}

// Pivot triangle.  kind 0: L11 (unit lower, forward)   1: U11 (upper, backward)
//                  2: U11^T (lower, forward)            3: L11^T (unit upper, backward)
// Forward kinds gather the right-hand sides and the children's updates first.  nbl is the
// panel width the front was factored with (the diagonal blocks hold L_D^{-1} below and
// U_D^{-1} on/above the diagonal).  With FULL the kernel also does the front's off-diagonal
// panel (the k_panel work) so one workgroup handles the whole front (small fronts): the
// vectors then cover all m rows, the forward column/row-panel updates run down to row m and
// the backward kinds start with the CB-to-pivot product.
template <typename T, int KIND, int R, bool FULL>
__device__ __forceinline__ void tri_body(FrontsDev F, const int* __restrict__ lvl, int B, int n, int nbl,
                                         const typename cplx<T>::type* __restrict__ arena,
                                         double2* __restrict__ uarena, double2* __restrict__ x, int b) {
  using V = typename cplx<T>::type;
  const int LD = nbl + 1;
  constexpr bool ROWTILE = KIND >= 2;          // the tile also stages 64 columns of a row panel
  extern __shared__ unsigned char smem_raw[];
  V* tile = reinterpret_cast<V*>(smem_raw);                                   // [nbl or 64][LD]
  double2* part = reinterpret_cast<double2*>(smem_raw + tri_tile_bytes(sizeof(V), KIND, R, FULL, nbl, blockDim.x));  // [R][4][PS]
  const int PS = nbl <= 32 ? 32 : 64;                                         // rows of a block
  double2* w = part + R * 4 * PS;                                             // [R][W]
  const int f = lvl[blockIdx.x];
  const int m = F.m[f], p = F.p[f];
  const int W = FULL ? m : p;
  const V* Fb = factor_ptr<T>(const_cast<V*>(arena), F, f, b, B);
  const int* rows = F.rows + F.rows_ptr[f];
  const int t = threadIdx.x;
  constexpr bool fwd = (KIND == 0 || KIND == 2);
  if (fwd) {
    solve_gather<R, FULL>(F, f, b, B, n, m, p, rows, x, uarena, w);
  } else {
    for (int i = t; i < W; i += blockDim.x)
#pragma unroll
      for (int r = 0; r < R; ++r) w[r * W + i] = x[xat(F, n, (int64_t)b * R + r, rows[i])];
  }
  __syncthreads();
  if (FULL && !fwd && p < m) {
    // CB -> pivot rows: kind 1: w_i -= sum_j U[i][p + j] w_{p+j};  kind 3: w_k -= sum_j L[p + j][k] w_{p+j}
    const int q = m - p;
    if (KIND == 1) {
      // 64 pivot rows at a time, the q CB columns split over G = blockDim / 64 thread groups so
      // that every thread has loads in flight (one thread per row left 3/4 of the workgroup idle
      // and each row a chain of q / 8 dependent load batches); the G partial sums meet in the
      // tile region, unused until the diagonal blocks (sized for R * blockDim double2 by
      // tri_tile_bytes)
      double2* red = reinterpret_cast<double2*>(tile);
      const int G = blockDim.x / 64;
      const int i = t & 63, g = t >> 6;
      for (int i0 = 0; i0 < p; i0 += 64) {
        double2 s[R];
#pragma unroll
        for (int r = 0; r < R; ++r) s[r] = make_double2(0, 0);
        if (i0 + i < p) {
          const V* row = Fb + (int64_t)(i0 + i) + (int64_t)p * m;
          batched<8>(g, q, G, [&](int j) { return row[(int64_t)j * m]; }, [&](int j, V v) {
            const double2 a = to_d2(v);
#pragma unroll
            for (int r = 0; r < R; ++r) s[r] = cfma(s[r], a, w[r * W + p + j]);
          });
        }
#pragma unroll
        for (int r = 0; r < R; ++r) red[(r * G + g) * 64 + i] = s[r];
        __syncthreads();
        if (t < 64 && i0 + t < p) {
#pragma unroll
          for (int r = 0; r < R; ++r) {
            double2 tot = red[r * G * 64 + t];
            for (int gg = 1; gg < G; ++gg) tot = d2add(tot, red[(r * G + gg) * 64 + t]);
            w[r * W + i0 + t] = d2sub(w[r * W + i0 + t], tot);
          }
        }
        __syncthreads();
      }
    } else {
      for (int c0 = 0; c0 < p; c0 += 64) {
        const int nc = min(64, p - c0);
        double2 s[R];
        if (t < nc) {
#pragma unroll
          for (int r = 0; r < R; ++r) s[r] = w[r * W + c0 + t];
        }
        for (int j0 = 0; j0 < q; j0 += nbl) {
          const int nj = min(nbl, q - j0);
          __syncthreads();
          batched<8>(t, nj * nc, blockDim.x,
                     [&](int idx) { return Fb[(int64_t)(p + j0 + idx % nj) + (int64_t)(c0 + idx / nj) * m]; },
                     [&](int idx, V v) { tile[(idx / nj) * LD + idx % nj] = v; });
          __syncthreads();
          if (t < nc)
            for (int jj = 0; jj < nj; ++jj) {
              const double2 a = to_d2(tile[t * LD + jj]);
#pragma unroll
              for (int r = 0; r < R; ++r) s[r] = cfms(s[r], a, w[r * W + p + j0 + jj]);
            }
        }
        if (t < nc) {
#pragma unroll
          for (int r = 0; r < R; ++r) w[r * W + c0 + t] = s[r];
        }
      }
    }
    __syncthreads();
  }
  const int nblk = (p + nbl - 1) / nbl;
  for (int bi = 0; bi < nblk; ++bi) {
    const int J0 = fwd ? bi * nbl : (nblk - 1 - bi) * nbl;
    const int nbk = min(nbl, p - J0);
    // diagonal block to LDS: tile[r + c * LD] = F[J0 + r][J0 + c]
    batched<8>(t, nbk * nbk, blockDim.x,
               [&](int idx) { return Fb[(int64_t)(J0 + idx % nbk) + (int64_t)(J0 + idx / nbk) * m]; },
               [&](int idx, V v) { tile[idx % nbk + (idx / nbk) * LD] = v; });
    __syncthreads();
    {
      const int i = t % PS, g = t / PS;   // output row, quarter of the reduction (g < 4)
      double2 s[R];
#pragma unroll
      for (int r = 0; r < R; ++r) s[r] = make_double2(0, 0);
      if (g < 4 && i < nbk) {
        if (KIND == 0) {          // L_D^{-1} (unit lower): w_i + sum_{k<i} Li[i][k] w_k
          for (int k = g; k < i; k += 4) {
            const double2 a = to_d2(tile[i + k * LD]);
#pragma unroll
            for (int r = 0; r < R; ++r) s[r] = cfma(s[r], a, w[r * W + J0 + k]);
          }
        } else if (KIND == 1) {   // U_D^{-1}: sum_{k>=i} Ui[i][k] w_k
          for (int k = i + g; k < nbk; k += 4) {
            const double2 a = to_d2(tile[i + k * LD]);
#pragma unroll
            for (int r = 0; r < R; ++r) s[r] = cfma(s[r], a, w[r * W + J0 + k]);
          }
        } else if (KIND == 2) {   // U_D^{-T}: sum_{k<=i} Ui[k][i] w_k
          for (int k = g; k <= i; k += 4) {
            const double2 a = to_d2(tile[k + i * LD]);
#pragma unroll
            for (int r = 0; r < R; ++r) s[r] = cfma(s[r], a, w[r * W + J0 + k]);
          }
        } else {                  // L_D^{-T} (unit upper): w_i + sum_{k>i} Li[k][i] w_k
          for (int k = i + 1 + g; k < nbk; k += 4) {
            const double2 a = to_d2(tile[k + i * LD]);
#pragma unroll
            for (int r = 0; r < R; ++r) s[r] = cfma(s[r], a, w[r * W + J0 + k]);
          }
        }
      }
      if (g < 4) {
#pragma unroll
        for (int r = 0; r < R; ++r) part[r * 4 * PS + g * PS + i] = s[r];
      }
    }
    __syncthreads();
    if (t < nbk) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        double2 tot = part[r * 4 * PS + t];
        for (int g = 1; g < 4; ++g) tot = d2add(tot, part[r * 4 * PS + g * PS + t]);
        if (KIND == 0 || KIND == 3) tot = d2add(tot, w[r * W + J0 + t]);
        w[r * W + J0 + t] = tot;
      }
    }
    __syncthreads();
    // remaining rows: forward kinds update [J0 + nbk, W) (W = m with FULL: the CB rows too),
    // backward kinds [0, J0)
    const int lo = fwd ? J0 + nbk : 0, hi = fwd ? W : J0;
    if (KIND == 0 || KIND == 1) {
      // column panel F[lo:hi, J0:J0+nbk] (contiguous rows): thread per row
      for (int i = lo + t; i < hi; i += blockDim.x) {
        double2 s[R];
#pragma unroll
        for (int r = 0; r < R; ++r) s[r] = w[r * W + i];
        const V* row = Fb + (int64_t)i + (int64_t)J0 * m;
        batched<8>(0, nbk, 1, [&](int k) { return row[(int64_t)k * m]; }, [&](int k, V v) {
          const double2 a = to_d2(v);
#pragma unroll
          for (int r = 0; r < R; ++r) s[r] = cfms(s[r], a, w[r * W + J0 + k]);
        });
#pragma unroll
        for (int r = 0; r < R; ++r) w[r * W + i] = s[r];
      }
    } else {
      // row panel F[J0:J0+nbk, lo:hi] read as its transpose through a 64-column LDS tile
      for (int c0 = lo; c0 < hi; c0 += 64) {
        const int nc = min(64, hi - c0);
        __syncthreads();
        batched<8>(t, nbk * nc, blockDim.x,
                   [&](int idx) { return Fb[(int64_t)(J0 + idx % nbk) + (int64_t)(c0 + idx / nbk) * m]; },
                   [&](int idx, V v) { tile[(idx / nbk) * LD + idx % nbk] = v; });
        __syncthreads();
        if (t < nc) {
          double2 s[R];
#pragma unroll
          for (int r = 0; r < R; ++r) s[r] = w[r * W + c0 + t];
          for (int k = 0; k < nbk; ++k) {
            const double2 a = to_d2(tile[t * LD + k]);
#pragma unroll
            for (int r = 0; r < R; ++r) s[r] = cfms(s[r], a, w[r * W + J0 + k]);
          }
#pragma unroll
          for (int r = 0; r < R; ++r) w[r * W + c0 + t] = s[r];
        }
      }
    }
    __syncthreads();
  }
  for (int i = t; i < W; i += blockDim.x) {
    if (i < p) {
#pragma unroll
      for (int r = 0; r < R; ++r) x[xat(F, n, (int64_t)b * R + r, rows[i])] = w[r * W + i];
    } else if (fwd) {
      double2* uf = uarena + F.uoff[f] * F.nvec;
#pragma unroll
      for (int r = 0; r < R; ++r) uf[uat(F, m - p, (int64_t)b * R + r, i - p)] = w[r * W + i];
    }
  }
}

template <typename T, int KIND, int R, bool FULL>
__global__ void __launch_bounds__(512) k_tri(FrontsDev F, const int* __restrict__ lvl, int B, int n, int nbl,
                                             const typename cplx<T>::type* __restrict__ arena,
                                             double2* __restrict__ uarena, double2* __restrict__ x,
                                             const int* __restrict__ bsel) {
  tri_body<T, KIND, R, FULL>(F, lvl, B, n, nbl, arena, uarena, x, solve_block(F, bsel, blockIdx.y));

}

// Whole-front solves of small fronts with the front's factor block staged in LDS, so the solve
// itself touches no global memory between the gather and the write-back (the pivot-triangle +
// panel kernels wait on a global load per diagonal block and per panel row chunk).  Persistent
// 256-thread workgroups walk the (front, frequency) items of a bucket; the next item's block is
// loaded into registers while the current one is solved from LDS, so loads and solves overlap.
//  kind 0 (forward):  the first p columns of the front (L11 and L21, contiguous in the
//                     column-major front) -> Pl[k m + i] = F[i][k];
//                     w = gathered right-hand sides and children's updates over all m rows;
//                     per diagonal block: w_blk = L_D^{-1} w_blk, then w[below] -= L[below][blk] w_blk;
//                     pivot rows to x, the other rows to the front's update vector
//  kind 1 (backward): the first p rows of every column (U11 and U12) -> Pl[j p + i] = F[i][j];
//                     w = x over all m rows (the contribution-block rows are final);
//                     w[0:p] -= U12 w[p:m], then per diagonal block, last first:
//                     w_blk = U_D^{-1} w_blk, w[above] -= U[above][blk] w_blk; pivot rows to x
// The diagonal blocks hold L_D^{-1} strictly below and U_D^{-1} on/above the diagonal (nbl wide).
// block entries a thread prefetches: blocks up to 4096 complex64 or 2048 complex128 entries
template <typename T> constexpr int tri_lds_npf() { return sizeof(T) == 4 ? 16 : 8; }

// Loads of a front's factor block (kind 0: first p columns; kind 1: first p rows of every column)
// into registers, NPF entries per thread, clamped to the block so every address is valid.
template <typename T, int KIND, int NPF>
__device__ __forceinline__ void tri_lds_fetch(const FrontsDev& F, const typename cplx<T>::type* arena, int f, int b,
                                              int B, T (&px)[NPF], T (&py)[NPF]) {
  using V = typename cplx<T>::type;
  const int t = threadIdx.x;
  const int m = F.m[f], p = F.p[f], mp = m * p;
  const V* Fb = factor_ptr<T>(const_cast<V*>(arena), F, f, b, B);
#pragma unroll
  for (int u = 0; u < NPF; ++u) {
    const int idx = min(t + 256 * u, mp - 1);
    V v;
    if (KIND == 0) {
      v = Fb[idx];
    } else {
      const int j = idx / p;
      v = Fb[(int64_t)(idx - j * p) + (int64_t)j * m];
    }
    px[u] = v.x;         // components kept as scalars: arrays of vector types went to scratch
    py[u] = v.y;
  }
}

template <typename T, int KIND, int R>
__global__ void __launch_bounds__(256) k_tri_lds(FrontsDev F, const int* __restrict__ lvl, int nf, int B, int n,
                                                 int nbl, const typename cplx<T>::type* __restrict__ arena,
                                                 double2* __restrict__ uarena, double2* __restrict__ x,
                                                 const int* __restrict__ bsel, int nsel) {
  using V = typename cplx<T>::type;
  static_assert(KIND == 0 || KIND == 1, "forward L or backward U");
  constexpr int NPF = tri_lds_npf<T>();
  const int t = threadIdx.x;
  const int nitems = nf * nsel;
  extern __shared__ unsigned char smem_raw[];
  double2* w = reinterpret_cast<double2*>(smem_raw);            // [R][m]
  double2* part = w + R * 256;                                   // [R][4][64] (m <= 256)
  V* Pl = reinterpret_cast<V*>(part + R * 4 * 64);               // [p][m] (kind 0) or [m][p] (kind 1)
  T px[NPF], py[NPF];
  auto front_of = [&](int it) { return lvl[it % nf]; };
  auto freq_of = [&](int it) { return solve_block(F, bsel, it / nf); };
  int item = blockIdx.x;
  if (item < nitems) tri_lds_fetch<T, KIND, NPF>(F, arena, front_of(item), freq_of(item), B, px, py);
  for (; item < nitems; item += gridDim.x) {
    const int f = front_of(item), b = freq_of(item);
    const int m = F.m[f], p = F.p[f], mp = m * p;
    const int* rows = F.rows + F.rows_ptr[f];
#pragma unroll
    for (int u = 0; u < NPF; ++u)
      if (t + 256 * u < mp) {
        V v;
        v.x = px[u];
        v.y = py[u];
        Pl[t + 256 * u] = v;
      }
    if (KIND == 0) {
      solve_gather<R, true>(F, f, b, B, n, m, p, rows, x, uarena, w);
    } else {
      for (int i = t; i < m; i += 256)
#pragma unroll
        for (int r = 0; r < R; ++r) w[r * m + i] = x[xat(F, n, (int64_t)b * R + r, rows[i])];
    }
    __syncthreads();
    if (item + (int)gridDim.x < nitems)     // in flight during the solve below
      tri_lds_fetch<T, KIND, NPF>(F, arena, front_of(item + gridDim.x), freq_of(item + gridDim.x), B, px, py);
    const int i64 = t & 63, g = t >> 6;     // row of a 64-row slab, quarter of a reduction
    if (KIND == 1 && p < m) {
      // w[i] -= sum_j U[i][p + j] w[p + j]: four partial sums per pivot row
      const int q = m - p;
      for (int i0 = 0; i0 < p; i0 += 64) {
        const int i = i0 + i64;
        double2 s[R];
#pragma unroll
        for (int r = 0; r < R; ++r) s[r] = make_double2(0, 0);
        if (i < p)
          for (int j = g; j < q; j += 4) {
            const double2 a = to_d2(Pl[(p + j) * p + i]);
#pragma unroll
            for (int r = 0; r < R; ++r) s[r] = cfma(s[r], a, w[r * m + p + j]);
          }
#pragma unroll
        for (int r = 0; r < R; ++r) part[(r * 4 + g) * 64 + i64] = s[r];
        __syncthreads();
        if (t < 64 && i < p)
#pragma unroll
          for (int r = 0; r < R; ++r) {
            double2 tot = part[r * 256 + t];
            for (int gg = 1; gg < 4; ++gg) tot = d2add(tot, part[(r * 4 + gg) * 64 + t]);
            w[r * m + i] = make_double2(w[r * m + i].x - tot.x, w[r * m + i].y - tot.y);
          }
        __syncthreads();
      }
    }
    const int nblk = (p + nbl - 1) / nbl;
    for (int bi = 0; bi < nblk; ++bi) {
      const int J0 = KIND == 0 ? bi * nbl : (nblk - 1 - bi) * nbl;
      const int nbk = min(nbl, p - J0);
      {   // diagonal block: four partial sums per row
        double2 s[R];
#pragma unroll
        for (int r = 0; r < R; ++r) s[r] = make_double2(0, 0);
        if (i64 < nbk) {
          if (KIND == 0) {        // L_D^{-1} (unit lower): w_i + sum_{k<i} Li[i][k] w_k
            for (int k = g; k < i64; k += 4) {
              const double2 a = to_d2(Pl[(J0 + k) * m + J0 + i64]);
#pragma unroll
              for (int r = 0; r < R; ++r) s[r] = cfma(s[r], a, w[r * m + J0 + k]);
            }
          } else {                // U_D^{-1}: sum_{k>=i} Ui[i][k] w_k
            for (int k = i64 + g; k < nbk; k += 4) {
              const double2 a = to_d2(Pl[(J0 + k) * p + J0 + i64]);
#pragma unroll
              for (int r = 0; r < R; ++r) s[r] = cfma(s[r], a, w[r * m + J0 + k]);
            }
          }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) part[(r * 4 + g) * 64 + i64] = s[r];
      }
      __syncthreads();
      if (t < nbk)
#pragma unroll
        for (int r = 0; r < R; ++r) {
          double2 tot = part[r * 256 + t];
          for (int gg = 1; gg < 4; ++gg) tot = d2add(tot, part[(r * 4 + gg) * 64 + t]);
          if (KIND == 0) tot = d2add(tot, w[r * m + J0 + t]);
          w[r * m + J0 + t] = tot;
        }
      __syncthreads();
      // the other rows of the block's columns: forward below the block (all m rows), backward above
      const int lo = KIND == 0 ? J0 + nbk : 0, hi = KIND == 0 ? m : J0;
      for (int i = lo + t; i < hi; i += 256) {
        double2 s[R];
#pragma unroll
        for (int r = 0; r < R; ++r) s[r] = w[r * m + i];
        for (int k = 0; k < nbk; ++k) {
          const double2 a = to_d2(KIND == 0 ? Pl[(J0 + k) * m + i] : Pl[(J0 + k) * p + i]);
#pragma unroll
          for (int r = 0; r < R; ++r) s[r] = cfms(s[r], a, w[r * m + J0 + k]);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) w[r * m + i] = s[r];
      }
      __syncthreads();
    }
    for (int i = t; i < m; i += 256) {
      if (i < p) {
#pragma unroll
        for (int r = 0; r < R; ++r) x[xat(F, n, (int64_t)b * R + r, rows[i])] = w[r * m + i];
      } else if (KIND == 0) {
        double2* uf = uarena + F.uoff[f] * F.nvec;
#pragma unroll
        for (int r = 0; r < R; ++r) uf[uat(F, m - p, (int64_t)b * R + r, i - p)] = w[r * m + i];
      }
    }
    __syncthreads();          // w and Pl are rewritten by the next item
  }
}

// Whole-front solves of one block of RC vectors (RC = 4, 8 or 16: the vectors of a group of
// frequencies that share one complex64 factorization), the vectors staged in LDS.  Every step
// is a small product whose (row, vector) outputs are spread over the 256 threads: TPR = RC / 4
// threads per row, four vectors each, so one factor entry serves four vectors of a thread.  The
// TPR threads of a row split its factor entries (every TPR-th, eight loads of a thread in
// flight, coalesced over the rows of a column) and pass them to each other by DPP, so a row of
// a 32-wide block arrives in one round trip; only the vectors use LDS (m (RC + 1) entries), so
// several workgroups share a CU and hide each other's loads.
//  kind 0 (forward):  w = the block's vectors at the pivot rows plus the children's update
//                     vectors, over all m rows; per diagonal block w_blk = L_D^{-1} w_blk, then
//                     w[below] -= L[below][blk] w_blk; pivot rows to x, the other rows to the
//                     front's update vectors
//  kind 1 (backward): w = x over all m rows (the contribution-block rows are final);
//                     w[0:p] -= U12 w[p:m], then per diagonal block, last first,
//                     w_blk = U_D^{-1} w_blk, w[above] -= U[above][blk] w_blk; pivot rows to x
// The diagonal blocks hold L_D^{-1} strictly below and U_D^{-1} on/above the diagonal (nbl wide).
// w rows are padded to RC + 1 entries (the gather's row-strided LDS writes then spread over the
// banks).
constexpr int GRP_NB = 8;      // factor loads of a thread in flight together

// Value v of lane S of this lane's TPR-lane row group (TPR consecutive lanes: 4 = a quad, 2 =
// a pair within the quad, 1 = itself), by a DPP quad permutation (no LDS).  All lanes of a
// group must be active.
template <int TPR, int S>
__device__ __forceinline__ float grp_bcast(float v) {
  if constexpr (TPR == 1) {
    return v;
  } else {
    constexpr int ctrl = TPR == 4 ? S * 0x55 : (S | (S << 2) | ((S + 2) << 4) | ((S + 2) << 6));
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), ctrl, 0xF, 0xF, false));
  }
}

// s[v] op= sum_{k in [k0, k1)} col[k m] w[(wrow0 + k) WS + r0 + v] for the TPR threads of one row:
// each loads every TPR-th entry (GRP_NB loads in flight), the entries reach the others of the
// group by grp_bcast, so a batch covers TPR * GRP_NB entries of the row in one round trip.
// SUB: subtract (update) or add (diagonal-block product).  k0, k1 must be equal for the TPR
// threads of a row.
// One batch of it: this thread's GRP_NB entries of [kb, kb + TPR GRP_NB) (clamped to k1 - 1;
// kb < k1), then their use.
template <int TPR>
__device__ __forceinline__ void grp_load(const float2* __restrict__ col, int64_t m, int kb, int k1, int q,
                                         float (&ex)[GRP_NB], float (&ey)[GRP_NB]) {
#pragma unroll
  for (int u = 0; u < GRP_NB; ++u) {
    const int kc = min(kb + u * TPR + q, k1 - 1);
    const float2 e = col[(int64_t)kc * m];
    ex[u] = e.x;
    ey[u] = e.y;
  }
}

template <int TPR, int VPT, int WS, bool SUB>
__device__ __forceinline__ void grp_use(const float (&ex)[GRP_NB], const float (&ey)[GRP_NB], int kb, int k1,
                                        const double2* __restrict__ w, int wrow0, int r0, double2 (&s)[VPT]) {
#pragma unroll
  for (int u = 0; u < GRP_NB; ++u) {
    auto step = [&](auto sc) {
      constexpr int S = decltype(sc)::value;
      const int k = kb + u * TPR + S;
      const double2 a = make_double2((double)grp_bcast<TPR, S>(ex[u]), (double)grp_bcast<TPR, S>(ey[u]));
      if (k < k1) {
#pragma unroll
        for (int v = 0; v < VPT; ++v)
          s[v] = SUB ? cfms(s[v], a, w[(wrow0 + k) * WS + r0 + v]) : cfma(s[v], a, w[(wrow0 + k) * WS + r0 + v]);
      }
    };
    step(std::integral_constant<int, 0>{});
    if constexpr (TPR > 1) step(std::integral_constant<int, 1>{});
    if constexpr (TPR > 2) {
      step(std::integral_constant<int, 2>{});
      step(std::integral_constant<int, 3>{});
    }
  }
}

template <int TPR, int VPT, int WS, bool SUB>
__device__ __forceinline__ void grp_row_product(const float2* __restrict__ col, int64_t m, int k0, int k1, int q,
                                                const double2* __restrict__ w, int wrow0, int r0, double2 (&s)[VPT]) {
  for (int kb = k0; kb < k1; kb += TPR * GRP_NB) {
    float ex[GRP_NB], ey[GRP_NB];
    grp_load<TPR>(col, m, kb, k1, q, ex, ey);
    grp_use<TPR, VPT, WS, SUB>(ex, ey, kb, k1, w, wrow0, r0, s);
  }
}

// grp_row_product whose first batch (at k0) grp_load has already brought into ex, ey: the same
// operations in the same order.
template <int TPR, int VPT, int WS, bool SUB>
__device__ __forceinline__ void grp_row_product_pre(const float2* __restrict__ col, int64_t m, int k0, int k1, int q,
                                                    const double2* __restrict__ w, int wrow0, int r0,
                                                    double2 (&s)[VPT], const float (&ex)[GRP_NB],
                                                    const float (&ey)[GRP_NB]) {
  if (k0 >= k1) return;
  grp_use<TPR, VPT, WS, SUB>(ex, ey, k0, k1, w, wrow0, r0, s);
  grp_row_product<TPR, VPT, WS, SUB>(col, m, k0 + TPR * GRP_NB, k1, q, w, wrow0, r0, s);
}

// PIV: the pivot rows only (fronts whose off-diagonal panels k_grp_panel handles): w holds the
// p pivot rows, kind 0 adds the children's updates of the pivot rows only (k_grp_panel forms
// the update vectors), kind 1 starts from x at the pivot rows (k_grp_panel has already
// subtracted U12 times the contribution-block rows).
template <int KIND, int RC, bool PIV = false, bool XR = false>
__global__ void __launch_bounds__(256) k_grp_lds(FrontsDev F, const int* __restrict__ lvl, int B, int n, int nbl,
                                                 const float2* __restrict__ arena, double2* __restrict__ uarena,
                                                 double2* __restrict__ x, const int* __restrict__ sel) {
  static_assert(KIND == 0 || KIND == 1, "forward L or backward U");
  constexpr int VPT = 4, TPR = RC / VPT, RPP = 256 / TPR, WS = RC + 1;
  static_assert(RC % VPT == 0 && RC <= 16, "blocks of 4, 8 or 16 vectors");
  extern __shared__ unsigned char smem_raw[];
  const int f = lvl[blockIdx.x];
  const int b = solve_block(F, sel, blockIdx.y);
  const int m = F.m[f], p = F.p[f], q = m - p;
  const int W = PIV ? p : m;                    // rows held in w
  double2* w = reinterpret_cast<double2*>(smem_raw);                  // [W][WS]
  const float2* Fb = factor_ptr<float>(const_cast<float2*>(arena), F, f, b, B);
  const int* rows = F.rows + F.rows_ptr[f];
  const int t = threadIdx.x;
  const int64_t v0 = (int64_t)b * RC;           // first vector of the block
  // vectors, row index fastest over the threads (runs of x rows coalesce); kind 0 keeps zero
  // at the contribution-block rows before the children's updates
  const int tr = t / TPR, r0 = (t % TPR) * VPT;     // row slot, first vector of the thread
  const int q4 = t % TPR;                           // thread within the row's group
  const int nblk = (p + nbl - 1) / nbl;
  // The factor entries of the first product do not depend on the vectors: this thread's first
  // batch of them is loaded before the gather, so its round trip overlaps the gather's.  First
  // product: forward, row tr of the first diagonal block (k in [0, tr)); backward with
  // contribution-block rows, row tr of U12 (k in [0, q)).  Not in the row-major forward kernel:
  // there the batch held across the gather costs 19 registers (80 -> 99, 6 -> 4 waves a SIMD),
  // more than the round trip it saves.
  constexpr bool pre_fwd = KIND == 0 && !XR;
  const bool cb_first = KIND == 1 && !PIV && q > 0;
  const int nbf = min(nbl, p);
  float pex[GRP_NB], pey[GRP_NB];
  if (pre_fwd && tr > 0 && tr < nbf) grp_load<TPR>(Fb + tr, m, 0, tr, q4, pex, pey);
  if (KIND == 1 && cb_first && tr < p) grp_load<TPR>(Fb + tr + (int64_t)p * m, m, 0, q, q4, pex, pey);
  const int mr = W * RC;
  // entry idx of the block's W x RC vector entries: the row fastest over the threads in the
  // vector-major layout, the vector fastest in the row-major one (coalesced either way)
  auto split = [W](int idx) -> int2 {    // (vector r, row i)
    if (XR) {
      const int i = idx / RC;
      return make_int2(idx - i * RC, i);
    }
    const int r = idx / W;
    return make_int2(r, idx - r * W);
  };
  const int c0 = F.child_ptr[f], nch = KIND == 0 ? F.child_ptr[f + 1] - c0 : 0;
  int q2[2] = {0, 0};
  if (nch == 2) {
    q2[0] = F.m[F.child_ids[c0]] - F.p[F.child_ids[c0]];
    q2[1] = F.m[F.child_ids[c0 + 1]] - F.p[F.child_ids[c0 + 1]];
  }
  if (nch == 2 && q2[0] > 0 && q2[1] > 0) {
    // two children (nested dissection): x and both update vectors in one pass, the three loads of
    // an entry in flight together (a row outside a child's update reads that vector's first
    // entry, a line in use anyway); sums in the order of the general loop below
    const int* inv0 = F.inv + F.inv_ptr[c0];
    const int* inv1 = F.inv + F.inv_ptr[c0 + 1];
    const double2* u0 = uarena + F.uoff[F.child_ids[c0]] * F.nvec;
    const double2* u1 = uarena + F.uoff[F.child_ids[c0 + 1]] * F.nvec;
    constexpr int U = 4;
    for (int base = t; base < mr; base += U * 256) {
      double2 a[U], b0[U], b1[U];
      int k0[U], k1[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int idx = min(base + u * 256, mr - 1);
        const int2 ri = split(idx);
        const int r = ri.x, i = ri.y;
        k0[u] = inv0[i];
        k1[u] = inv1[i];
        a[u] = x[xat_t<XR>(F, n, v0 + r, rows[min(i, p - 1)])];
        b0[u] = u0[uat_t<XR>(F, q2[0], v0 + r, max(k0[u], 0))];
        b1[u] = u1[uat_t<XR>(F, q2[1], v0 + r, max(k1[u], 0))];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int idx = base + u * 256;
        if (idx < mr) {
          const int2 ri = split(idx);
          const int r = ri.x, i = ri.y;
          double2 v = i < p ? a[u] : make_double2(0, 0);
          if (k0[u] >= 0) v = d2add(v, b0[u]);
          if (k1[u] >= 0) v = d2add(v, b1[u]);
          w[i * WS + r] = v;
        }
      }
    }
  } else {
    batched<8>(t, mr, 256,
               [&](int idx) {
                 // kind 0 uses x at the pivot rows only: the other lanes re-read the last pivot row
                 // (a line the wave loads anyway) instead of a contribution-block row
                 const int2 ri = split(idx);
                 const int r = ri.x, i = ri.y;
                 return x[xat_t<XR>(F, n, v0 + r, rows[KIND == 0 ? min(i, p - 1) : i])];
               },
               [&](int idx, double2 v) {
                 const int2 ri = split(idx);
                 const int r = ri.x, i = ri.y;
                 w[i * WS + r] = (KIND == 1 || i < p) ? v : make_double2(0, 0);
               });
    for (int c = 0; c < nch; ++c) {
      const int* inv = F.inv + F.inv_ptr[c0 + c];
      const int ch = F.child_ids[c0 + c];
      const int qc = F.m[ch] - F.p[ch];
      const double2* u = uarena + F.uoff[ch] * F.nvec;
      // each thread adds to the w entries it wrote above (same index -> thread map): no barrier
      batched<8>(t, mr, 256,
                 [&](int idx) {
                   const int2 ri = split(idx);
                   const int r = ri.x, i = ri.y;
                   return u[uat_t<XR>(F, qc, v0 + r, max(inv[i], 0))];
                 },
                 [&](int idx, double2 v) {
                   const int2 ri = split(idx);
                   const int r = ri.x, i = ri.y;
                   if (inv[i] >= 0) w[i * WS + r] = d2add(w[i * WS + r], v);
                 });
    }
  }
  __syncthreads();
  double2 s[VPT];
  // s -= sum_{k in [k0, k1)} F[i][c0 + k] w[wrow0 + k] (entries of row i of the column-major front)
  auto row_update = [&](int i, int c0, int k0, int k1, int wrow0) {
    grp_row_product<TPR, VPT, WS, true>(Fb + i + (int64_t)c0 * m, m, k0, k1, q4, w, wrow0, r0, s);
  };
  if (cb_first) {
    // contribution-block rows -> pivot rows: w[i] -= sum_j U[i][p + j] w[p + j] (row tr from the
    // batch loaded before the gather)
    if (tr < p) {
#pragma unroll
      for (int v = 0; v < VPT; ++v) s[v] = w[tr * WS + r0 + v];
      grp_row_product_pre<TPR, VPT, WS, true>(Fb + tr + (int64_t)p * m, m, 0, q, q4, w, p, r0, s, pex, pey);
#pragma unroll
      for (int v = 0; v < VPT; ++v) w[tr * WS + r0 + v] = s[v];
    }
    for (int i = tr + RPP; i < p; i += RPP) {
#pragma unroll
      for (int v = 0; v < VPT; ++v) s[v] = w[i * WS + r0 + v];
      row_update(i, p, 0, q, p);
#pragma unroll
      for (int v = 0; v < VPT; ++v) w[i * WS + r0 + v] = s[v];
    }
    __syncthreads();
  }
  // diagonal block (nbk <= 64 <= RPP rows: one row per thread slot), from the old values:
  // s = L_D^{-1} w (forward, unit lower: w_i + sum_{k<i} Li[i][k] w_k) or U_D^{-1} w (backward:
  // sum_{k>=i} Ui[i][k] w_k) for row slot tr < nbk
  auto diag = [&](int J0, int nbk, auto pre) {
    const int i = tr;
    const float2* col = Fb + (J0 + i) + (int64_t)J0 * m;
#pragma unroll
    for (int v = 0; v < VPT; ++v) s[v] = KIND == 0 ? w[(J0 + i) * WS + r0 + v] : make_double2(0, 0);
    const int k0 = KIND == 0 ? 0 : i, k1 = KIND == 0 ? i : nbk;
    if constexpr (decltype(pre)::value)
      grp_row_product_pre<TPR, VPT, WS, false>(col, m, k0, k1, q4, w, J0, r0, s, pex, pey);
    else
      grp_row_product<TPR, VPT, WS, false>(col, m, k0, k1, q4, w, J0, r0, s);
  };
  // forward: the first block's product from the batch loaded before the gather, ahead of the loop
  if (pre_fwd && tr < nbf) diag(0, nbf, std::true_type{});
  for (int bi = 0; bi < nblk; ++bi) {
    const int J0 = KIND == 0 ? bi * nbl : (nblk - 1 - bi) * nbl;
    const int nbk = min(nbl, p - J0);
    const bool own = tr < nbk;
    if (own && (!pre_fwd || bi > 0)) diag(J0, nbk, std::false_type{});
    __syncthreads();
    if (own) {
#pragma unroll
      for (int v = 0; v < VPT; ++v) w[(J0 + tr) * WS + r0 + v] = s[v];
    }
    __syncthreads();
    // the other rows of the block's columns: forward below the block (all m rows), backward above
    const int lo = KIND == 0 ? J0 + nbk : 0, hi = KIND == 0 ? W : J0;
    for (int i = lo + tr; i < hi; i += RPP) {
#pragma unroll
      for (int v = 0; v < VPT; ++v) s[v] = w[i * WS + r0 + v];
      row_update(i, J0, 0, nbk, J0);
#pragma unroll
      for (int v = 0; v < VPT; ++v) w[i * WS + r0 + v] = s[v];
    }
    __syncthreads();
  }
  double2* uf = uarena + F.uoff[f] * F.nvec;
  for (int idx = t; idx < mr; idx += 256) {
    const int2 ri = split(idx);
    const int r = ri.x, i = ri.y;
    if (i < p) x[xat_t<XR>(F, n, v0 + r, rows[i])] = w[i * WS + r];
    else if (KIND == 0) uf[uat_t<XR>(F, q, v0 + r, i - p)] = w[i * WS + r];
  }
}

// Off-diagonal panel of one front for one block of RC vectors sharing its factors, 64 output
// rows per workgroup (lvl: tile list front << 12 | tile, as k_panel), 64 TPR threads: TPR =
// RC / 4 threads per row, four vectors each, the row's factor entries split over them and
// passed on by DPP (grp_row_product); the input vector entries are staged in LDS in chunks of
// GRP_VC, [k][WS] as k_grp_lds's w.
//  kind 0 (forward,  L21): uf[o] = sum_c child_c[o] - sum_{k<p} L[p + o][k] y[k]   o: the front's
//                          contribution-block rows; y = x at the pivot rows (solved); the
//                          children's update vectors are gathered here (k_grp_lds with PIV
//                          takes only the pivot rows)
//  kind 1 (backward, U12): x[o] -= sum_{j<q} U[o][p + j] x[p + j]                   o: pivot rows
constexpr int GRP_VC = 128;
template <int KIND, int RC, bool XR = false>
__global__ void __launch_bounds__(256) k_grp_panel(FrontsDev F, const int* __restrict__ lvl, int B, int n,
                                                   const float2* __restrict__ arena, double2* __restrict__ uarena,
                                                   double2* __restrict__ x, const int* __restrict__ sel) {
  static_assert(KIND == 0 || KIND == 1, "forward L21 or backward U12");
  constexpr int VPT = 4, TPR = RC / VPT, WS = RC + 1, NT = 64 * TPR;
  static_assert(RC % VPT == 0 && RC <= 16, "blocks of 4, 8 or 16 vectors");
  extern __shared__ unsigned char smem_raw[];
  double2* vin = reinterpret_cast<double2*>(smem_raw);               // [GRP_VC][WS]
  const int f = lvl[blockIdx.x] >> 12;
  const int o0 = (lvl[blockIdx.x] & 4095) * 64;
  const int m = F.m[f], p = F.p[f], q = m - p;
  const int nout = KIND == 0 ? q : p, nin = KIND == 0 ? p : q;
  if (o0 >= nout || nin == 0) return;
  const int b = solve_block(F, sel, blockIdx.y);
  const float2* Fb = factor_ptr<float>(const_cast<float2*>(arena), F, f, b, B);
  const int* rows = F.rows + F.rows_ptr[f];
  const int t = threadIdx.x;
  const int o = t / TPR, q4 = t % TPR, r0 = q4 * VPT;
  const bool act = o0 + o < nout;
  const int64_t v0 = (int64_t)b * RC;
  const int in0 = KIND == 0 ? 0 : p;                                  // first input row of the front
  const int frow = (KIND == 0 ? p : 0) + o0 + o;                      // the thread's factor row
  double2 s[VPT];
#pragma unroll
  for (int v = 0; v < VPT; ++v) s[v] = make_double2(0, 0);
  for (int c0 = 0; c0 < nin; c0 += GRP_VC) {
    const int nc = min(GRP_VC, nin - c0);
    __syncthreads();
    batched<8>(t, nc * RC, NT,
               [&](int idx) {
                 // the row (vector-major layout) or the vector (row-major) fastest over the threads
                 const int r = XR ? idx % RC : idx / nc, k = XR ? idx / RC : idx - r * nc;
                 return x[xat_t<XR>(F, n, v0 + r, rows[in0 + c0 + k])];
               },
               [&](int idx, double2 v) {
                 const int r = XR ? idx % RC : idx / nc, k = XR ? idx / RC : idx - r * nc;
                 vin[k * WS + r] = v;
               });
    __syncthreads();
    if (act)
      grp_row_product<TPR, VPT, WS, true>(Fb + frow + (int64_t)(in0 + c0) * m, m, 0, nc, q4, vin, 0, r0, s);
  }
  if (!act) return;
  if (KIND == 0) {
    const int i = p + o0 + o;                                         // front row
    const int cp0 = F.child_ptr[f], nch = F.child_ptr[f + 1] - cp0;
    if (nch == 2) {
      // two children: both row maps, then both update vectors, in flight together (the sums in
      // the order of the general loop below)
      const int ci0 = F.inv[F.inv_ptr[cp0] + i], ci1 = F.inv[F.inv_ptr[cp0 + 1] + i];
      const int ch0 = F.child_ids[cp0], ch1 = F.child_ids[cp0 + 1];
      const int qc0 = F.m[ch0] - F.p[ch0], qc1 = F.m[ch1] - F.p[ch1];
      const double2* u0 = uarena + F.uoff[ch0] * F.nvec;
      const double2* u1 = uarena + F.uoff[ch1] * F.nvec;
      double2 a0[VPT], a1[VPT];
#pragma unroll
      for (int v = 0; v < VPT; ++v) {
        if (ci0 >= 0) a0[v] = u0[uat_t<XR>(F, qc0, v0 + r0 + v, ci0)];
        if (ci1 >= 0) a1[v] = u1[uat_t<XR>(F, qc1, v0 + r0 + v, ci1)];
      }
#pragma unroll
      for (int v = 0; v < VPT; ++v) {
        if (ci0 >= 0) s[v] = d2add(s[v], a0[v]);
        if (ci1 >= 0) s[v] = d2add(s[v], a1[v]);
      }
    }
    for (int c = 0; c < (nch == 2 ? 0 : nch); ++c) {
      const int ci = F.inv[F.inv_ptr[cp0 + c] + i];
      if (ci >= 0) {
        const int ch = F.child_ids[cp0 + c];
        const int qc = F.m[ch] - F.p[ch];
        const double2* u = uarena + F.uoff[ch] * F.nvec;
#pragma unroll
        for (int v = 0; v < VPT; ++v) s[v] = d2add(s[v], u[uat_t<XR>(F, qc, v0 + r0 + v, ci)]);
      }
    }
    double2* uf = uarena + F.uoff[f] * F.nvec;
#pragma unroll
    for (int v = 0; v < VPT; ++v) uf[uat_t<XR>(F, q, v0 + r0 + v, o0 + o)] = s[v];
  } else {
    const int row = rows[o0 + o];
#pragma unroll
    for (int v = 0; v < VPT; ++v) {
      double2* d = x + xat_t<XR>(F, n, v0 + r0 + v, row);
      *d = d2add(*d, s[v]);
    }
  }
}

// Off-diagonal panels, one workgroup per 64-wide output tile of one front and frequency; lvl is
// a tile list (front << 12 | tile), the frequency comes from blockIdx.y.
//  kind 0 (forward,  L21):   uf[r]  -= sum_k L[p + r][k] y[k]        r in CB rows
//  kind 1 (backward, U12):   x_p[r] -= sum_j U[r][p + j] x_cb[j]      r in pivot rows
//  kind 2 (forward,  U12^T): uf[j]  -= sum_k U[k][p + j] w[k]         j in CB rows
//  kind 3 (backward, L21^T): x_p[k] -= sum_j L[p + j][k] x_cb[j]      k in pivot rows
template <typename T, int KIND, int R>
__device__ __forceinline__ void panel_body(FrontsDev F, const int* __restrict__ lvl, int B, int n,
                                           const typename cplx<T>::type* __restrict__ arena,
                                           double2* __restrict__ uarena, double2* __restrict__ x, int b) {
  using V = typename cplx<T>::type;
  extern __shared__ unsigned char smem_raw[];
  const int f = lvl[blockIdx.x] >> 12;
  const int m = F.m[f], p = F.p[f];
  const int q = m - p;
  const bool out_cb = (KIND == 0 || KIND == 2);
  const int nout = out_cb ? q : p;     // outputs
  const int nin = out_cb ? p : q;      // reduction length
  const int o0 = (lvl[blockIdx.x] & 4095) * 64;
  if (o0 >= nout || nin == 0) return;
  const int no = min(64, nout - o0);
  const V* Fb = factor_ptr<T>(const_cast<V*>(arena), F, f, b, B);
  const int* rows = F.rows + F.rows_ptr[f];
  double2* uf = uarena + F.uoff[f] * F.nvec;                  // + uat(F, q, b R + r, j)
  constexpr int VC = 256;                                     // input entries staged per chunk
  double2* v = reinterpret_cast<double2*>(smem_raw);        // [R][VC] input vector chunk
  double2* part = v + R * VC;                                 // [R][4][64] partial sums
  V* tile = reinterpret_cast<V*>(part + R * 256);             // [64][KC2 + 1] (transposed kinds)
  const int t = threadIdx.x;
  const int o = t & 63, g = t >> 6;   // output within tile, reduction quarter
  double2 s[R];
#pragma unroll
  for (int r = 0; r < R; ++r) s[r] = make_double2(0, 0);
  for (int v0 = 0; v0 < nin; v0 += VC) {
    const int nv = min(VC, nin - v0);
    __syncthreads();
    for (int i = t; i < nv; i += blockDim.x)
#pragma unroll
      for (int r = 0; r < R; ++r) v[r * VC + i] = x[xat(F, n, (int64_t)b * R + r, rows[(out_cb ? 0 : p) + v0 + i])];
    __syncthreads();
    if (KIND == 0 || KIND == 1) {
      // direct: output o is a row of the panel; coalesced across o
      const int row = (KIND == 0 ? p : 0) + o0 + o;
      const int cbase = (KIND == 0 ? 0 : p) + v0;
      if (o < no)
        batched<8>(g, nv, 4, [&](int k) { return Fb[(int64_t)row + (int64_t)(cbase + k) * m]; }, [&](int k, V e) {
          const double2 a = to_d2(e);
#pragma unroll
          for (int r = 0; r < R; ++r) s[r] = cfma(s[r], a, v[r * VC + k]);
        });
    } else {
      // transposed: output o is a column of the stored panel; stage 64 x KC2 tiles through LDS
      constexpr int KC2 = SolveTile<T>::SB;                    // reduction sub-chunk
      const int rbase = (KIND == 2 ? 0 : p) + v0;              // stored rows (reduction index)
      const int cbase = (KIND == 2 ? p : 0) + o0;              // stored columns (outputs)
      for (int k0 = 0; k0 < nv; k0 += KC2) {
        const int nk = min(KC2, nv - k0);
        __syncthreads();
        batched<8>(t, nk * no, blockDim.x,
                   [&](int idx) { return Fb[(int64_t)(rbase + k0 + idx % nk) + (int64_t)(cbase + idx / nk) * m]; },
                   [&](int idx, V e) { tile[(idx / nk) * (KC2 + 1) + idx % nk] = e; });
        __syncthreads();
        if (o < no)
          for (int kk = g; kk < nk; kk += 4) {
            const double2 a = to_d2(tile[o * (KC2 + 1) + kk]);
#pragma unroll
            for (int r = 0; r < R; ++r) s[r] = cfma(s[r], a, v[r * VC + k0 + kk]);
          }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r) part[r * 256 + g * 64 + o] = s[r];
  __syncthreads();
  if (t < no) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      double2 tot = part[r * 256 + t];
      for (int gg = 1; gg < 4; ++gg) tot = d2add(tot, part[r * 256 + gg * 64 + t]);
      if (out_cb) {
        double2* d = uf + uat(F, q, (int64_t)b * R + r, o0 + t);
        d->x -= tot.x; d->y -= tot.y;
      } else {
        double2* d = x + xat(F, n, (int64_t)b * R + r, rows[o0 + t]);
        d->x -= tot.x; d->y -= tot.y;
      }
    }
  }
}

template <typename T, int KIND, int R>
__global__ void __launch_bounds__(256) k_panel(FrontsDev F, const int* __restrict__ lvl, int B, int n,
                                               const typename cplx<T>::type* __restrict__ arena,
                                               double2* __restrict__ uarena, double2* __restrict__ x,
                                               const int* __restrict__ bsel) {
  panel_body<T, KIND, R>(F, lvl, B, n, arena, uarena, x, solve_block(F, bsel, blockIdx.y));

}

__global__ void k_rhs(int n, int B, const double* __restrict__ omega, double beta, double V,
                      const double2* __restrict__ rK, const double2* __restrict__ rM,
                      const double2* __restrict__ rD, double2* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = blockIdx.y;
  const double w = omega[b];
  const double2 k = rK[i], mm = rM[i];
  double2 a = make_double2(k.x - w * w * mm.x, k.y - w * w * mm.y);
  if (rD) { const double2 d = rD[i]; a.x -= w * beta * d.y; a.y += w * beta * d.x; }
  out[(int64_t)b * n + i] = make_double2(-V * a.x, -V * a.y);
}

// r = rhs - A(w) x for all frequencies; one thread per row, 8 frequencies per pass.
// With tslot != NULL the operator is A^T: the value of slot s is taken from slot tslot[s]
// (the pattern is structurally symmetric).
__global__ void k_residual(int n, int B, const int64_t* __restrict__ indptr, const int* __restrict__ indices,
                           const int64_t* __restrict__ tslot,
                           const double2* __restrict__ Kaa, const double2* __restrict__ Maa,
                           const double2* __restrict__ Daa, const double* __restrict__ omega, double beta,
                           const double2* __restrict__ x, const double2* __restrict__ rhs, double2* __restrict__ r) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b0 = blockIdx.y * 8;
  const int nb = min(8, B - b0);
  // the eight frequencies' accumulators stay in registers (fully unrolled, guarded by nb)
  double2 acc[8];
  double w[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) { acc[q] = make_double2(0, 0); w[q] = q < nb ? omega[b0 + q] : 0.0; }
  for (int64_t s0 = indptr[i]; s0 < indptr[i + 1]; ++s0) {
    const int j = indices[s0];
    const int64_t s = tslot ? tslot[s0] : s0;
    const double2 k = Kaa[s], mm = Maa[s];
    const double2 d = Daa ? Daa[s] : make_double2(0, 0);
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (q < nb) {
        double2 a = make_double2(k.x - w[q] * w[q] * mm.x - w[q] * beta * d.y,
                                 k.y - w[q] * w[q] * mm.y + w[q] * beta * d.x);
        acc[q] = cfma(acc[q], a, x[(int64_t)(b0 + q) * n + j]);
      }
  }
#pragma unroll
  for (int q = 0; q < 8; ++q)
    if (q < nb) {
      const double2 bb = rhs[(int64_t)(b0 + q) * n + i];
      r[(int64_t)(b0 + q) * n + i] = make_double2(bb.x - acc[q].x, bb.y - acc[q].y);
    }
}

// Frequency-contiguous residual: xt, rhst, rt are [n][Bp] (row-major, Bp = frequencies padded to
// a multiple of 64).  One wave per (row, 64 frequencies); matrix values are uniform across the wave,
// the ω-dependent coefficient and the x entries are per lane (coalesced).
__global__ void __launch_bounds__(256) k_residual_t(int n, int B, int Bp, const int64_t* __restrict__ indptr,
                                                    const int* __restrict__ indices, const int64_t* __restrict__ tslot,
                                                    const double2* __restrict__ Kaa, const double2* __restrict__ Maa,
                                                    const double2* __restrict__ Daa, const double* __restrict__ omega,
                                                    double beta, const double2* __restrict__ xt,
                                                    const double2* __restrict__ rhst, double2* __restrict__ rt) {
  const int lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= n) return;
  const int b = blockIdx.y * 64 + lane;
  const double w = b < B ? omega[b] : 0.0;
  const double w2 = w * w, wb = w * beta;
  double2 acc = make_double2(0, 0);
  for (int64_t s0 = indptr[i]; s0 < indptr[i + 1]; ++s0) {
    const int j = indices[s0];
    const int64_t s = tslot ? tslot[s0] : s0;
    const double2 k = Kaa[s], mm = Maa[s];
    const double2 d = Daa ? Daa[s] : make_double2(0, 0);
    const double2 a = make_double2(k.x - w2 * mm.x - wb * d.y, k.y - w2 * mm.y + wb * d.x);
    acc = cfma(acc, a, xt[(int64_t)j * Bp + b]);
  }
  const double2 r = rhst[i * Bp + b];
  rt[i * Bp + b] = make_double2(r.x - acc.x, r.y - acc.y);
}

// The same residual over row groups: the rows of a group (at most RES_G, listed in
// grows[gptr[g] .. gptr[g + 1])) have identical column lists, as the active unknowns of one
// merged mesh node do, so one wave per group reads each x row once for all the group's rows.
// The wave covers FT tiles of 64 frequencies (ft <= FT of them in the last column of the grid):
// the group's column indices and K, M, D values are the same for every lane, and with the group
// index made wave-uniform (readfirstlane) they arrive by scalar loads, each used for the wave's
// FT x 64 frequencies (per-lane vector loads of them bound the kernel by address processing).
// Per row the products are summed in the order of k_residual_t, bit for bit the same result.
// Columns: B logical columns c (tiles of 64) stored at column cmap[c] (or c) of the [n][Bp]
// arrays; omega is indexed by the stored column.  Lanes past B read and write nothing.
constexpr int RES_G = 4;
template <int FT>
__global__ void __launch_bounds__(256) k_residual_sv(int ng, int B, int Bp, const int* __restrict__ cmap,
                                                     const int* __restrict__ gptr,
                                                     const int* __restrict__ grows, const int64_t* __restrict__ indptr,
                                                     const int* __restrict__ indices, const int64_t* __restrict__ tslot,
                                                     const double2* __restrict__ Kaa, const double2* __restrict__ Maa,
                                                     const double2* __restrict__ Daa,
                                                     const double* __restrict__ omega, double beta,
                                                     const double2* __restrict__ xt, const double2* __restrict__ rhst,
                                                     double2* __restrict__ rt) {
  const int lane = threadIdx.x & 63;
  const int g = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
  if (g >= ng) return;
  const int t0 = blockIdx.y * FT;
  const int ft = min(FT, (B + 63) / 64 - t0);
  double w2[FT], wb[FT];
  int col[FT];          // stored column of this lane in tile f, or -1
#pragma unroll
  for (int f = 0; f < FT; ++f) {
    const int c = (t0 + f) * 64 + lane;
    col[f] = (f < ft && c < B) ? (cmap ? cmap[c] : c) : -1;
    const double w = col[f] >= 0 ? omega[col[f]] : 0.0;
    w2[f] = w * w;
    wb[f] = w * beta;
  }
  const int g0 = gptr[g], cnt = gptr[g + 1] - g0;
  int64_t base[RES_G];
#pragma unroll
  for (int t = 0; t < RES_G; ++t) base[t] = indptr[grows[g0 + min(t, cnt - 1)]];
  const int64_t len = indptr[grows[g0] + 1] - base[0];
  const int* cols = indices + base[0];
  double2 acc[RES_G][FT];
#pragma unroll
  for (int t = 0; t < RES_G; ++t)
#pragma unroll
    for (int f = 0; f < FT; ++f) acc[t][f] = make_double2(0, 0);
  for (int64_t k = 0; k < len; ++k) {
    const int64_t j = cols[k];
    double2 xv[FT];
#pragma unroll
    for (int f = 0; f < FT; ++f) xv[f] = col[f] >= 0 ? xt[j * Bp + col[f]] : make_double2(0, 0);
#pragma unroll
    for (int t = 0; t < RES_G; ++t) {
      if (t < cnt) {
        const int64_t s = tslot ? tslot[base[t] + k] : base[t] + k;
        const double2 kk = Kaa[s], mm = Maa[s];
        const double2 d = Daa ? Daa[s] : make_double2(0, 0);
#pragma unroll
        for (int f = 0; f < FT; ++f) {
          const double2 a = make_double2(kk.x - w2[f] * mm.x - wb[f] * d.y, kk.y - w2[f] * mm.y + wb[f] * d.x);
          acc[t][f] = cfma(acc[t][f], a, xv[f]);
        }
      }
    }
  }
#pragma unroll
  for (int t = 0; t < RES_G; ++t) {
    if (t < cnt) {
      const int64_t i = grows[g0 + t];
#pragma unroll
      for (int f = 0; f < FT; ++f) {
        if (col[f] >= 0) {
          const int64_t o = i * Bp + col[f];
          const double2 r = rhst ? rhst[o] : make_double2(0, 0);
          rt[o] = make_double2(r.x - acc[t][f].x, r.y - acc[t][f].y);
        }
      }
    }
  }
}

// Mapped transpose (tiled through LDS): out row omap[c] (or c), column r =
// add[omap[c]][r] (or 0) + in[imap[r]][c] (input row imap[r], or r), for r < R, c < C.  Gathers the
// refinement's active vectors into the frequency-major layout and scatters the residuals back.
__global__ void __launch_bounds__(256) k_transpose_map(int R, int C, int ld_in, int ld_out,
                                                       const double2* __restrict__ in, double2* __restrict__ out,
                                                       const int* __restrict__ imap, const int* __restrict__ omap,
                                                       const double2* __restrict__ add) {
  __shared__ double2 tile[32][33];
  const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int k = ty; k < 32; k += 8) {
    const int r = r0 + k, c = c0 + tx;
    if (r < R && c < C) tile[k][tx] = in[(int64_t)(imap ? imap[r] : r) * ld_in + c];
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int c = c0 + k, r = r0 + tx;
    if (r < R && c < C) {
      const int64_t o = (int64_t)(omap ? omap[c] : c) * ld_out + r;
      double2 v = tile[tx][k];
      if (add) v = d2add(v, add[o]);
      out[o] = v;
    }
  }
}

// out[j][i] = in[i][j] for an R x C complex128 matrix (tiled through LDS).
__global__ void __launch_bounds__(256) k_transpose(int R, int C, int ld_in, int ld_out, const double2* __restrict__ in,
                                                   double2* __restrict__ out) {
  __shared__ double2 tile[32][33];
  const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int k = ty; k < 32; k += 8) {
    const int r = r0 + k, c = c0 + tx;
    if (r < R && c < C) tile[k][tx] = in[(int64_t)r * ld_in + c];
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int c = c0 + k, r = r0 + tx;
    if (r < R && c < C) out[(int64_t)c * ld_out + r] = tile[tx][k];
  }
}

// s + a b with every product and sum rounded on its own (no contraction into FMAs): the charge
// and duality products below use it, so that their forms over either vector layout agree bit for
// bit whatever fusing the compiler would choose in each loop.
__device__ __forceinline__ double2 cfma_nc(double2 s, double2 a, double2 b) {
#pragma clang fp contract(off)
  return make_double2(s.x + (a.x * b.x - a.y * b.y), s.y + (a.x * b.y + a.y * b.x));
}
__device__ __forceinline__ double2 charge_out(double2 sum, double V, double2 kts_z) {
#pragma clang fp contract(off)
  return make_double2(sum.x + V * kts_z.x, sum.y + V * kts_z.y);
}

// Q[b] = sum_p kts_a[p] x[b][p] + V kts_z: one 1024-thread block per frequency (many loads in
// flight for the long dot product), fixed-order tree reduction (deterministic)
constexpr int CHARGE_NT = 1024;
__global__ void __launch_bounds__(CHARGE_NT) k_charge(int n, const double2* __restrict__ kts_a, double2 kts_z,
                                                      double V, const double2* __restrict__ x,
                                                      double2* __restrict__ Q) {
  __shared__ double2 red[CHARGE_NT];
  const int b = blockIdx.x;
  double2 s = make_double2(0, 0);
#pragma unroll 4
  for (int i = threadIdx.x; i < n; i += CHARGE_NT) s = cfma_nc(s, kts_a[i], x[(int64_t)b * n + i]);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = CHARGE_NT / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      red[threadIdx.x].x += red[threadIdx.x + h].x;
      red[threadIdx.x].y += red[threadIdx.x + h].y;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) Q[b] = charge_out(red[0], V, kts_z);
}

// out[k] = sum_i a_b[i] w[i] c_b[i] (no conjugation; w real, or 1 when null) for b = sel[k] (or
// k): the refinement's duality products and dual-weighted residuals in one pass over the two
// vectors, one 1024-thread block per product, fixed-order tree reduction (deterministic).
constexpr int WDOT_NT = 1024;
__global__ void __launch_bounds__(WDOT_NT) k_wdot(int n, const double2* __restrict__ a, int64_t sa,
                                                  const double* __restrict__ w, const double2* __restrict__ c,
                                                  int64_t sc, const int* __restrict__ sel, double2* __restrict__ out) {
  __shared__ double2 red[WDOT_NT];
  const int k = blockIdx.x;
  const int64_t b = sel ? sel[k] : k;
  const double2* ab = a + b * sa;
  const double2* cb = c + b * sc;
  double2 s = make_double2(0, 0);
#pragma unroll 4
  for (int i = threadIdx.x; i < n; i += WDOT_NT) {
    double2 u = ab[i];
    const double wi = w ? w[i] : 1.0;
    u.x *= wi;
    u.y *= wi;
    s = cfma_nc(s, u, cb[i]);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = WDOT_NT / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      red[threadIdx.x].x += red[threadIdx.x + h].x;
      red[threadIdx.x].y += red[threadIdx.x + h].y;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[k] = red[0];
}

// The same products on vectors stored row-major (entry i of column c at [i ld + c]), in two
// stages that keep k_charge's / k_wdot's summation order exactly: stage 1 forms, for every
// column, the 1024 strided partial sums (thread t of the column-major kernels: entries t, t +
// 1024, ... in order), 64 columns x 16 partial sums per workgroup so that the row loads coalesce
// over the columns; stage 2 reduces each column's 1024 partial sums by the same tree.
//   charge: a = kts_a (a vector), c = x at column ccol[k] (or k)
//   wdot:   a = a at column acol[k] times w (or 1), c = c at column ccol[k]
template <bool CHARGE>
__global__ void __launch_bounds__(1024) k_rows_partial(int n, int nc, int64_t ld, const double2* __restrict__ a,
                                                       const int* __restrict__ acol, const double* __restrict__ w,
                                                       const double2* __restrict__ c, const int* __restrict__ ccol,
                                                       double2* __restrict__ part) {
  const int lane = threadIdx.x & 63;
  const int k = blockIdx.x * 64 + lane;
  const int t = blockIdx.y * 16 + (threadIdx.x >> 6);
  if (k >= nc) return;
  const int64_t ca = CHARGE ? 0 : (acol ? acol[k] : k);
  const int64_t cc = ccol ? ccol[k] : k;
  double2 s = make_double2(0, 0);
  for (int i = t; i < n; i += 1024) {
    if (CHARGE) {
      s = cfma_nc(s, a[i], c[(int64_t)i * ld + cc]);
    } else {
      double2 u = a[(int64_t)i * ld + ca];
      const double wi = w ? w[i] : 1.0;
      u.x *= wi;
      u.y *= wi;
      s = cfma_nc(s, u, c[(int64_t)i * ld + cc]);
    }
  }
  part[(int64_t)t * nc + k] = s;
}

// Both duality products of the paired refinement in one pass over two row-major arrays whose
// columns pair up per frequency (2b: forward vector, 2b + 1: adjoint): out1[b] = sum_i X[i][2b +
// 1] w[i] C[i][2b], out2[b] = sum_i X[i][2b] w[i] C[i][2b + 1], stage 1 as k_rows_partial (each a
// k_wdot order), a lane reading its frequency's two adjacent columns of each array.
__global__ void __launch_bounds__(1024) k_rows_partial_pair(int n, int nb, int64_t ld, const double2* __restrict__ X,
                                                            const double* __restrict__ w,
                                                            const double2* __restrict__ C, double2* __restrict__ part) {
  const int lane = threadIdx.x & 63;
  const int b = blockIdx.x * 64 + lane;
  const int t = blockIdx.y * 16 + (threadIdx.x >> 6);
  if (b >= nb) return;
  double2 s1 = make_double2(0, 0), s2 = make_double2(0, 0);
  for (int i = t; i < n; i += 1024) {
    const double2* xr = X + (int64_t)i * ld + 2 * b;
    const double2* cr = C + (int64_t)i * ld + 2 * b;
    const double2 x0 = xr[0], x1 = xr[1], c0 = cr[0], c1 = cr[1];
    const double wi = w ? w[i] : 1.0;
    double2 u = x1;
    u.x *= wi;
    u.y *= wi;
    s1 = cfma_nc(s1, u, c0);
    u = x0;
    u.x *= wi;
    u.y *= wi;
    s2 = cfma_nc(s2, u, c1);
  }
  part[(int64_t)t * nb + b] = s1;
  part[(int64_t)(1024 + t) * nb + b] = s2;
}

// One refinement update of the paired row-major vectors and its measurements in one pass: X +=
// R on the column pairs (2b forward, 2b + 1 adjoint: k_axpy's sums), then the charges of both
// updated columns (k_charge's products) and mu[b] = sum_i X[i][2b + 1] w[i] rhs[i][2b] (k_wdot's),
// stage 1 as k_rows_partial (pq: [1024][2 nb], pm: [1024][nb]).
__global__ void __launch_bounds__(1024) k_rows_update_measure(int n, int nb, int64_t ld, double2* __restrict__ X,
                                                              const double2* __restrict__ R,
                                                              const double2* __restrict__ kts_a,
                                                              const double* __restrict__ w,
                                                              const double2* __restrict__ rhs,
                                                              double2* __restrict__ pq, double2* __restrict__ pm) {
  const int lane = threadIdx.x & 63;
  const int b = blockIdx.x * 64 + lane;
  const int t = blockIdx.y * 16 + (threadIdx.x >> 6);
  if (b >= nb) return;
  double2 q0 = make_double2(0, 0), q1 = make_double2(0, 0), mu = make_double2(0, 0);
  for (int i = t; i < n; i += 1024) {
    double2* xr = X + (int64_t)i * ld + 2 * b;
    const double2* rr = R + (int64_t)i * ld + 2 * b;
    double2 x0 = xr[0], x1 = xr[1];
    const double2 r0 = rr[0], r1 = rr[1];
    x0.x += r0.x;
    x0.y += r0.y;
    x1.x += r1.x;
    x1.y += r1.y;
    xr[0] = x0;
    xr[1] = x1;
    const double2 k = kts_a[i];
    q0 = cfma_nc(q0, k, x0);
    q1 = cfma_nc(q1, k, x1);
    const double wi = w[i];
    double2 u = x1;
    u.x *= wi;
    u.y *= wi;
    mu = cfma_nc(mu, u, rhs[(int64_t)i * ld + 2 * b]);
  }
  pq[(int64_t)t * (2 * nb) + 2 * b] = q0;
  pq[(int64_t)t * (2 * nb) + 2 * b + 1] = q1;
  pm[(int64_t)t * nb + b] = mu;
}

template <bool CHARGE>
__global__ void __launch_bounds__(1024) k_rows_reduce(int nc, const double2* __restrict__ part, double2 kts_z,
                                                      double V, double2* __restrict__ out) {
  __shared__ double2 red[1024];
  const int k = blockIdx.x;
  red[threadIdx.x] = part[(int64_t)threadIdx.x * nc + k];
  __syncthreads();
  for (int h = 512; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      red[threadIdx.x].x += red[threadIdx.x + h].x;
      red[threadIdx.x].y += red[threadIdx.x + h].y;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (CHARGE) out[k] = charge_out(red[0], V, kts_z);
    else out[k] = red[0];
  }
}

__global__ void k_axpy(int64_t n, double2* __restrict__ x, const double2* __restrict__ d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i].x += d[i].x;
  x[i].y += d[i].y;
}

// ---------------------------------------------------------------- plan
// Fronts of one level that take one factorization path.
// Work lists of one pivot panel [k0, k0 + nbl) over the fronts of a size bucket (entries
// front << 12 | tile), so the panel launches have no empty workgroups
struct PanelTiles {
  const int* diag = nullptr;     // fronts with p > k0 (diagonal-block LU)
  int ndiag = 0;
  const int* t23 = nullptr;      // panel solves: row tiles of L21 = column tiles of U12
  int n23 = 0;
  const int* t4 = nullptr;       // left-looking update: n4c column-part tiles, then n4r row-part
  int n4c = 0, n4r = 0;          //   tiles (those carry 2048 in the tile field)
};

struct Sub {
  int nf = 0, mmax = 0, pmax = 0, cbmax = 0;
  int64_t mpmax = 0;               // largest m * p (factor block of the LDS-staged solves)
  const int* fronts = nullptr;   // device list
  std::vector<int> hfronts;      // host copy
  std::vector<PanelTiles> panels;  // per pivot panel, for the panel width of panel_nb
  // Schur-complement tiles of these fronts, one entry per TILE x TILE tile of a contribution
  // block: front << 12 | tile row << 6 | tile column (so a launch has no empty workgroups)
  const int* tiles = nullptr;
  int ntiles = 0;
  // solve panel tiles (front << 12 | 64-row tile) over the contribution rows (forward kinds)
  // and over the pivot rows (backward kinds), for fronts with both
  const int* stiles_cb = nullptr;
  const int* stiles_p = nullptr;
  int nstiles_cb = 0, nstiles_p = 0;
  // front assembly: columns per workgroup (about 4096 entries of the bucket's largest front,
  // in whole batches of 8) and the work list front << 12 | column chunk
  int init_cpw = 8;
  const int* itiles = nullptr;
  int nitiles = 0;
  // childless fronts assembled with their entries by k_leaf_front: work list (as itiles), the
  // entry range of each item (pairs) and the LDS of its position map
  const int* litiles = nullptr;
  const int64_t* lranges = nullptr;
  int nlitiles = 0;
  size_t leaf_lds = 0;
};

struct Level {
  int nf, mmax, pmax, cbmax;     // all fronts of the level (solves, panel width)
  int64_t e0, e1;
  const int* fronts;             // device pointer into lvl_fronts
  // per precision (0 complex64, 1 complex128): the fused set and the panel-path fronts, the
  // latter in size buckets so that each launch's grid fits its fronts
  Sub small[2];
  std::vector<Sub> big[2];
  // solves: one-workgroup (pivot triangle + panel) set and the two-kernel fronts, bucketed
  std::vector<Sub> ssmall[2];   // one-workgroup forward solves, in size buckets
  std::vector<Sub> sbig[2];
};

// Optional per-kernel timing with HIP events on the launch stream (kind ids below).
enum Kind { K_INIT = 0, K_ENTRIES, K_DIAG, K_TRSM, K_GEMM_L, K_GEMM_S, K_FWD, K_BWD, K_SMALL, K_NKIND };

struct Plan {
  int n = 0, nf = 0;
  int profile = 0;               // 0 off, 1 every launch (factorizations eager), 2 grouped solves only
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  struct Mark { int kind, level; size_t i0; };
  std::vector<Mark> ev_marks;                      // (kind, level, index of the start event)
  int cur_level = 0;                               // level of the launches being recorded
  std::vector<double> level_ms;                    // [level][kind] totals of the last read
  int64_t arena_elems = 0, uarena_elems = 0;
  std::vector<Level> levels;
  FrontsDev F{};
  std::vector<void*> owned;
  int* lvl_fronts = nullptr;
  int* ent_front = nullptr;
  int64_t* ent_slot = nullptr;
  int* ent_loc = nullptr;
  int* npert = nullptr;
  unsigned* pmin = nullptr;      // smallest |pivot| / row scale since the last reset (float bits)
  double* rscale = nullptr;      // [3][n] row maxima of |K|, |M|, |D| of the last factorization
  int64_t nent = 0;
  int mmax = 0;
  int pmax = 0;
  // fused small-front path: per-front entry ranges, per-precision front flags and the
  // largest front order it takes (0: off); host copies to rebuild the level partitions
  int64_t* fent_lo = nullptr;
  int64_t* fent_hi = nullptr;
  unsigned char* fused[2] = {nullptr, nullptr};
  int small_limit[2] = {0, 0};   // device LDS bound per precision
  int lds_max = 64 * 1024;       // LDS one workgroup may allocate on this device
  int lds_per_cu = 64 * 1024;    // LDS of one CU
  int ncu = 1;                   // compute units
  int small_mmax[2] = {0, 0};    // threshold in use
  int solve_full_m = 0;          // fronts up to this order are solved by one workgroup each
  int panel_lu_m = 0;            // 32-wide panels of buckets up to this order: k_panel_lu
  int tri_lds_bytes = 0;         // one-workgroup solve buckets whose factor block fits: k_tri_lds
  int group_rc = 16;             // grouped solves: vectors per workgroup of k_grp_lds at most
  int group_rc_tri = 4;          // grouped solves: vectors per workgroup of the other kernels
  int grp_lds_bytes = 48 * 1024;   // one-workgroup buckets whose k_grp_lds LDS fits: k_grp_lds
  int grp_big_pmax = 1024;         // two-kernel buckets with pivot blocks up to this: k_grp_lds + k_grp_panel
  std::vector<void*> part_owned[2];
  std::vector<void*> panel_owned[2];   // the buckets' panel work lists (rebuilt per panel width)
  int panel_nb[2] = {0, 0};      // panel width the buckets' panel lists were built for (0: none)
  // independent size buckets of a level run on side streams (default priority) next to the
  // launch stream: fork after the level's assembly, join before the next level
  static constexpr int NSIDE = 2;
  // grouped solves: the partial last group's launches, concurrent with the full groups'
  hipStream_t tail_stream = nullptr;
  hipEvent_t tail_fork = nullptr, tail_join = nullptr;
  bool overlap = true;
  hipStream_t side[NSIDE] = {};
  hipEvent_t fork_ev = nullptr, join_ev[NSIDE] = {};
  std::vector<int> hm, hp, hlvl_ptr, hlvl_fronts;
  std::vector<int> hchild_ptr, hent_loc;       // host copies for the leaf-front work lists
  std::vector<int> hchild_cb;                  // per front: sum of its children's contribution-block rows
  // k_grp_lds launches under profiling: HIP-event pair and algorithmic bytes of each
  struct GrpMark { size_t i0; double bytes; };
  std::vector<GrpMark> grp_marks;
  std::vector<int64_t> hfent_lo, hfent_hi;
  // captured factorizations (sawfem_mf_factor_graph): the launch sequence of one argument set,
  // replayed by hipGraphLaunch; dropped when a kernel-selection knob changes
  struct FactorGraph {
    int prec, B, nb;
    const void *omega, *Kaa, *Maa, *Daa, *arena;
    double beta;
    hipGraph_t graph;
    hipGraphExec_t exec;
    size_t nodes;
    uint64_t used;
  };
  static constexpr int NGRAPH = 4;
  std::vector<FactorGraph> graphs;
  uint64_t graph_clock = 0;
  size_t graph_nodes = 0;          // nodes of the graph the last sawfem_mf_factor_graph ran
  hipStream_t capture = nullptr;   // stream the factorizations are captured on
};

void drop_graphs(Plan* P) {
  for (auto& g : P->graphs) {
    (void)hipGraphExecDestroy(g.exec);
    (void)hipGraphDestroy(g.graph);
  }
  P->graphs.clear();
}

// Record a start/stop event pair around one launch when every launch is profiled (mode 1).
struct Timed {
  Plan* P; hipStream_t s; int kind; size_t i0; bool on;
  Timed(Plan* p, hipStream_t st, int k) : P(p), s(st), kind(k), i0(0), on(p->profile == 1) {
    if (!on) return;
    while (P->ev_pool.size() < P->ev_used + 2) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) { P->profile = 0; on = false; return; }
      P->ev_pool.push_back(e);
    }
    i0 = P->ev_used;
    P->ev_used += 2;
    (void)hipEventRecord(P->ev_pool[i0], s);
  }
  ~Timed() {
    if (!on) return;
    (void)hipEventRecord(P->ev_pool[i0 + 1], s);
    P->ev_marks.push_back({kind, P->cur_level, i0});
  }
};

template <typename X>
int upload(Plan* P, X** dst, const X* src, int64_t count) {
  if (count <= 0) { *dst = nullptr; return 0; }
  void* d = nullptr;
  SAW_HIP(hipMalloc(&d, sizeof(X) * count));
  P->owned.push_back(d);
  *dst = (X*)d;
  SAW_HIP(hipMemcpy(d, src, sizeof(X) * count, hipMemcpyHostToDevice));
  return 0;
}

// Panel width of a level: narrow panels for levels of small fronts.  The factorization and the
// solves must agree on it (the diagonal blocks hold inverses of nbl x nbl triangles).
int panel_width(int nb, const Level& L) { return (nb == 64 && L.pmax <= 128) ? 32 : nb; }

// Side streams and events of the bucket overlap, created on first use.
// A failure part-way leaves what was created for the next call (or the plan's destroy).
int ensure_side(Plan* P) {
  if (P->fork_ev) return 0;
  for (int k = 0; k < Plan::NSIDE; ++k) {
    if (!P->side[k]) SAW_HIP(hipStreamCreateWithFlags(&P->side[k], hipStreamNonBlocking));
    if (!P->join_ev[k]) SAW_HIP(hipEventCreateWithFlags(&P->join_ev[k], hipEventDisableTiming));
  }
  SAW_HIP(hipEventCreateWithFlags(&P->fork_ev, hipEventDisableTiming));
  return 0;
}

// Fork: the side streams wait for everything issued on s so far.  Join: s waits for the side
// streams.
int fork_side(Plan* P, hipStream_t s) {
  SAW_HIP(hipEventRecord(P->fork_ev, s));
  for (int k = 0; k < Plan::NSIDE; ++k) SAW_HIP(hipStreamWaitEvent(P->side[k], P->fork_ev, 0));
  return 0;
}
int join_side(Plan* P, hipStream_t s) {
  for (int k = 0; k < Plan::NSIDE; ++k) {
    SAW_HIP(hipEventRecord(P->join_ev[k], P->side[k]));
    SAW_HIP(hipStreamWaitEvent(s, P->join_ev[k], 0));
  }
  return 0;
}
int ensure_tail(Plan* P) {
  if (!P->tail_stream) SAW_HIP(hipStreamCreateWithFlags(&P->tail_stream, hipStreamNonBlocking));
  if (!P->tail_fork) SAW_HIP(hipEventCreateWithFlags(&P->tail_fork, hipEventDisableTiming));
  if (!P->tail_join) SAW_HIP(hipEventCreateWithFlags(&P->tail_join, hipEventDisableTiming));
  return 0;
}

// Split every level into the fronts the fused small-front kernel takes (m <= max_m) and the
// rest, for precision index pi; max_m <= 0 sends every front through the panel path.
int build_partition(Plan* P, int pi, int max_m) {
  for (void* d : P->part_owned[pi]) (void)hipFree(d);
  P->part_owned[pi].clear();
  P->small_mmax[pi] = std::max(0, std::min(max_m, P->small_limit[pi]));
  const int nf = P->nf;
  // Schur tile lists pack front (19 bits) and tile row / column (6 bits each)
  if (nf >= (1 << 19)) { sawfem_set_error("sawfem_mf: more than 524287 fronts"); return 2; }
  for (int f = 0; f < nf; ++f) {
    if (P->hm[f] - P->hp[f] > 64 * TILE) { sawfem_set_error("sawfem_mf: contribution block above 4096 rows"); return 2; }
    // assembly work lists pack the column chunk (>= 8 columns) in 12 bits
    if (P->hm[f] > 8 * 4095) { sawfem_set_error("sawfem_mf: front order above 32760"); return 2; }
  }
  std::vector<unsigned char> flag(nf, 0);
  auto put = [&](const std::vector<int>& v, const int** dst) -> int {
    *dst = nullptr;
    if (v.empty()) return 0;
    void* d = nullptr;
    SAW_HIP(hipMalloc(&d, sizeof(int) * v.size()));
    P->part_owned[pi].push_back(d);
    *dst = (const int*)d;
    SAW_HIP(hipMemcpy(d, v.data(), sizeof(int) * v.size(), hipMemcpyHostToDevice));
    return 0;
  };
  // the solves' one-workgroup set must fit the pivot-triangle kernel's LDS with the full front
  const int full_m = std::min(std::max(P->solve_full_m, P->small_mmax[pi]), 1024);
  for (size_t l = 0; l < P->levels.size(); ++l) {
    Level& L = P->levels[l];
    std::vector<int> sm, bg, ssm, sbg;
    for (int k = P->hlvl_ptr[l]; k < P->hlvl_ptr[l + 1]; ++k) {
      const int f = P->hlvl_fronts[k];
      (P->hm[f] <= P->small_mmax[pi] ? sm : bg).push_back(f);
      (P->hm[f] <= full_m ? ssm : sbg).push_back(f);
    }
    auto stats = [&](const std::vector<int>& v, Sub& S) {
      S = Sub{};
      S.nf = (int)v.size();
      for (int f : v) {
        S.mmax = std::max(S.mmax, P->hm[f]);
        S.pmax = std::max(S.pmax, P->hp[f]);
        S.cbmax = std::max(S.cbmax, P->hm[f] - P->hp[f]);
        S.mpmax = std::max(S.mpmax, (int64_t)P->hm[f] * P->hp[f]);
      }
    };
    // size buckets by front order (edges roughly 1.7x apart): fronts of one bucket need grids
    // of similar size, so few workgroups of a launch find no tile.  The one-workgroup solves
    // reserve LDS for their bucket's largest front (16 vectors: 272 B a row), so up to 180 (the
    // largest 16-vector block within 48 KB) their edges follow the LDS occupancy steps of a
    // 160 KB CU: 7, 6, 5, 4, 3 workgroups.
    auto buckets = [&](std::vector<int> v, std::vector<Sub>& out, bool factor, bool solve1 = false) -> int {
      out.clear();
      std::stable_sort(v.begin(), v.end(), [&](int a, int b) { return P->hm[a] < P->hm[b]; });
      static const int edges_f[] = {64, 128, 192, 320, 576, 1088, 1 << 30};
      static const int edges_s[] = {64, 84, 98, 117, 147, 180, 320, 576, 1088, 1 << 30};
      const int* e0 = solve1 ? edges_s : edges_f;
      const int ne = solve1 ? (int)(sizeof(edges_s) / sizeof(int)) : (int)(sizeof(edges_f) / sizeof(int));
      size_t i = 0;
      for (int ei = 0; ei < ne; ++ei) {
        const int e = e0[ei];
        std::vector<int> part;
        while (i < v.size() && P->hm[v[i]] <= e) part.push_back(v[i++]);
        if (part.empty()) continue;
        Sub S;
        stats(part, S);
        if (put(part, &S.fronts)) return 3;
        S.hfronts = part;
        std::vector<int> tl;
        for (int f : part) {
          const int q = P->hm[f] - P->hp[f];
          if (q <= 0 || P->hp[f] <= 0) continue;
          const int nt = (q + TILE - 1) / TILE;
          for (int tj = 0; tj < nt; ++tj)
            for (int ti = 0; ti < nt; ++ti) tl.push_back(f << 12 | ti << 6 | tj);
        }
        S.ntiles = (int)tl.size();
        if (put(tl, &S.tiles)) return 3;
        std::vector<int> scb, sp;
        for (int f : part) {
          const int q = P->hm[f] - P->hp[f], pp = P->hp[f];
          if (q <= 0 || pp <= 0) continue;
          for (int t = 0; t < (q + 63) / 64; ++t) scb.push_back(f << 12 | t);
          for (int t = 0; t < (pp + 63) / 64; ++t) sp.push_back(f << 12 | t);
        }
        S.nstiles_cb = (int)scb.size();
        S.nstiles_p = (int)sp.size();
        if (put(scb, &S.stiles_cb) || put(sp, &S.stiles_p)) return 3;
        S.init_cpw = 8 * std::max(1, std::min(8, (4096 / std::max(1, S.mmax) + 4) / 8));
        std::vector<int> it, lt;
        std::vector<int64_t> lr;
        for (int f : part) {
          const int mf = P->hm[f], nch = P->hchild_ptr[f + 1] - P->hchild_ptr[f];
          const int64_t elo = P->hfent_lo[f], ehi = P->hfent_hi[f];
          const int* loc = P->hent_loc.data();
          // leaves whose entries are in column-major order and whose chunk map fits LDS
          const bool leaf = factor && nch == 0 && (int64_t)S.init_cpw * S.mmax <= LEAF_MAXPOS &&
                            std::is_sorted(loc + elo, loc + ehi);
          for (int c = 0; c < (mf + S.init_cpw - 1) / S.init_cpw; ++c) {
            if (!leaf) { it.push_back(f << 12 | c); continue; }
            const int c0 = c * S.init_cpw, c1 = std::min(mf, c0 + S.init_cpw);
            lt.push_back(f << 12 | c);
            lr.push_back(std::lower_bound(loc + elo, loc + ehi, c0 * mf) - loc);
            lr.push_back(std::lower_bound(loc + elo, loc + ehi, c1 * mf) - loc);
          }
          if (leaf) flag[f] = 1;
        }
        S.nitiles = (int)it.size();
        S.nlitiles = (int)lt.size();
        S.leaf_lds = sizeof(int) * (size_t)S.init_cpw * S.mmax;
        if (put(it, &S.itiles) || put(lt, &S.litiles)) return 3;
        S.lranges = nullptr;
        if (!lr.empty()) {
          void* d = nullptr;
          SAW_HIP(hipMalloc(&d, sizeof(int64_t) * lr.size()));
          P->part_owned[pi].push_back(d);
          SAW_HIP(hipMemcpy(d, lr.data(), sizeof(int64_t) * lr.size(), hipMemcpyHostToDevice));
          S.lranges = (const int64_t*)d;
        }
        out.push_back(S);
      }
      return 0;
    };
    stats(sm, L.small[pi]);
    if (put(sm, &L.small[pi].fronts) || buckets(ssm, L.ssmall[pi], false, true) || buckets(bg, L.big[pi], true) ||
        buckets(sbg, L.sbig[pi], false))
      return 3;
    for (int f : sm) flag[f] = 1;
  }
  void* d = nullptr;
  SAW_HIP(hipMalloc(&d, nf));
  P->part_owned[pi].push_back(d);
  P->fused[pi] = (unsigned char*)d;
  SAW_HIP(hipMemcpy(d, flag.data(), nf, hipMemcpyHostToDevice));
  P->panel_nb[pi] = 0;
  return 0;
}

// Panel work lists of every size bucket for panel width nb (the factorization's; a level's
// panels are panel_width(nb, level) wide).  32-wide panels of buckets with m > 64 ("narrow")
// run in 128 x 32 / 32 x 128 tiles, everything else in TILE x TILE tiles.
int build_panel_lists(Plan* P, int pi, int nb) {
  for (void* d : P->panel_owned[pi]) (void)hipFree(d);
  P->panel_owned[pi].clear();
  P->panel_nb[pi] = 0;
  for (Level& L : P->levels) {
    const int nbl = panel_width(nb, L);
    for (Sub& G : L.big[pi]) {
      const int R = (nbl == 32 && G.mmax > 64) ? 128 : TILE;
      G.panels.clear();
      for (int k0 = 0; k0 < G.pmax; k0 += nbl) {
        std::vector<int> dg, t23, t4c, t4r;
        for (int f : G.hfronts) {
          const int m = P->hm[f], p = P->hp[f];
          if (p <= k0) continue;
          if ((m + R - 1) / R >= 2048) { sawfem_set_error("sawfem_mf: front too large for the panel tile lists"); return 2; }
          const int k1 = std::min(p, k0 + nbl);
          dg.push_back(f);
          for (int t = 0; t < (m - k1 + R - 1) / R; ++t) t23.push_back(f << 12 | t);
          if (k0 > 0) {
            for (int t = 0; t < (m - k0 + R - 1) / R; ++t) t4c.push_back(f << 12 | t);
            for (int t = 0; t < (m - k1 + R - 1) / R; ++t) t4r.push_back(f << 12 | 2048 | t);
          }
        }
        PanelTiles PT;
        PT.ndiag = (int)dg.size();
        PT.n23 = (int)t23.size();
        PT.n4c = (int)t4c.size();
        PT.n4r = (int)t4r.size();
        t4c.insert(t4c.end(), t4r.begin(), t4r.end());
        for (auto [v, dst] : {std::pair<std::vector<int>*, const int**>{&dg, &PT.diag}, {&t23, &PT.t23}, {&t4c, &PT.t4}}) {
          *dst = nullptr;
          if (v->empty()) continue;
          void* d = nullptr;
          SAW_HIP(hipMalloc(&d, sizeof(int) * v->size()));
          P->panel_owned[pi].push_back(d);
          SAW_HIP(hipMemcpy(d, v->data(), sizeof(int) * v->size(), hipMemcpyHostToDevice));
          *dst = (const int*)d;
        }
        G.panels.push_back(PT);
      }
    }
  }
  P->panel_nb[pi] = nb;
  return 0;
}

template <typename T>
int factor_impl(Plan* P, hipStream_t s, int B, const double* omega, double beta, const void* Kaa, const void* Maa,
                const void* Daa, void* arena_v, int nb) {
  using V = typename cplx<T>::type;
  V* arena = (V*)arena_v;
  const int pi = sizeof(T) == 4 ? 0 : 1;
  if (P->panel_nb[pi] != nb) {
    const int rc = build_panel_lists(P, pi, nb);
    if (rc) return rc;
  }
  // pivot guard: row scales of this operator, threshold sqrt(u_T) times the row scale
  const int64_t n3 = 3 * (int64_t)P->n;
  hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, s, n3, P->rscale, 0.0);
  if (P->nent > 0)
    hipLaunchKernelGGL(k_row_scale, dim3((unsigned)((P->nent + 255) / 256)), dim3(256), 0, s, P->nent, P->F,
                       P->ent_front, P->ent_slot, P->ent_loc, (const double2*)Kaa, (const double2*)Maa,
                       (const double2*)Daa, P->n, P->rscale);
  const PivotGuard PG{P->rscale, omega, beta, sizeof(T) == 4 ? 0x1p-12 : 0x1.6a09e667f3bcdp-27, P->n, P->npert, P->pmin};
  for (size_t li = 0; li < P->levels.size(); ++li) {
    const Level& L = P->levels[li];
    P->cur_level = (int)li;
    const int nbl = panel_width(nb, L);
    const Sub& S = L.small[pi];
    if (S.nf > 0) {
      const size_t lds = sizeof(V) * S.mmax * (S.mmax + 1) + sizeof(int) * S.mmax;
      Timed tm(P, s, K_SMALL);
      hipLaunchKernelGGL(k_small_front<T>, dim3(S.nf, B), dim3(SMALL_NT), lds, s, P->F, S.fronts, B, omega, beta,
                         P->fent_lo, P->fent_hi, P->ent_slot, P->ent_loc, (const double2*)Kaa, (const double2*)Maa,
                         (const double2*)Daa, arena, nbl, PG);
    }
    if (L.big[pi].empty()) continue;
    for (const Sub& G : L.big[pi]) {
      if (G.nlitiles > 0) {
        Timed tm(P, s, K_INIT);
        hipLaunchKernelGGL(k_leaf_front<T>, dim3(G.nlitiles, (B + LEAF_FB - 1) / LEAF_FB), dim3(256), G.leaf_lds, s,
                           P->F, G.litiles, G.lranges, G.init_cpw, B, omega, beta, P->ent_slot, P->ent_loc,
                           (const double2*)Kaa, (const double2*)Maa, (const double2*)Daa, arena);
      }
      if (G.nitiles > 0) {
        Timed tm(P, s, K_INIT);
        hipLaunchKernelGGL(k_front_init<T>, dim3(G.nitiles, B), dim3(256), 0, s, P->F, G.itiles, B, arena, G.init_cpw);
      }
    }
    if (L.e1 > L.e0) { Timed tm(P, s, K_ENTRIES);
      hipLaunchKernelGGL(k_front_entries<T>, dim3((unsigned)((L.e1 - L.e0 + 255) / 256), (B + ENT_FB - 1) / ENT_FB),
                         dim3(256), 0, s, P->F,
                         L.e0, L.e1, B, omega, beta, P->ent_front, P->ent_slot, P->ent_loc, (const double2*)Kaa,
                         (const double2*)Maa, (const double2*)Daa, arena, P->fused[pi]); }
    const bool fork = P->overlap && L.big[pi].size() > 1;
    if (fork && (ensure_side(P) || fork_side(P, s))) return 3;
    for (size_t gi = 0; gi < L.big[pi].size(); ++gi) {
      const Sub& G = L.big[pi][gi];
      const hipStream_t sg = (!fork || gi == 0) ? s : P->side[(gi - 1) % Plan::NSIDE];
      // left-looking panels: update panel k0 with the factored panels before it (one GEMM with
      // K = k0 for its column and row parts), then its diagonal LU + inverses, then the two
      // panel solves; the contribution block waits for the final Schur complement
      // 32-wide panels use 128 x 32 / 32 x 128 tiles (no wave on columns outside the panel);
      // the bucket of fronts with m <= 64 keeps 64 x 64 tiles (measured equal or better there)
      const bool narrow = nbl == 32 && G.mmax > 64;
      // diagonal block + panel solves in one wave per front (k_panel_lu) for these buckets
      const bool panel_lu = nbl == PANEL_LU_NB && G.mmax <= P->panel_lu_m;
      for (size_t pk = 0; pk < G.panels.size(); ++pk) {
        const PanelTiles& PT = G.panels[pk];
        const int k0 = (int)pk * nbl;
        if (PT.n4c + PT.n4r > 0) {
          Timed tm(P, sg, K_GEMM_L);
          const bool pre = k0 <= SHORT_K;   // short reduction (K = k0): prefetch the C tile
          const int* t4r = PT.t4 + PT.n4c;
          if (narrow) {         // column part in 128 x 32 tiles, row part in 32 x 128 tiles
            if (pre)
              hipLaunchKernelGGL((k_gemm<T, 4, 128, 32, true>), dim3(PT.n4c, 1, B), dim3(256), 0, sg, P->F, PT.t4, B,
                                 arena, k0, nbl);
            else
              hipLaunchKernelGGL((k_gemm<T, 4, 128, 32>), dim3(PT.n4c, 1, B), dim3(256), 0, sg, P->F, PT.t4, B, arena,
                                 k0, nbl);
            if (PT.n4r > 0) {
              if (pre)
                hipLaunchKernelGGL((k_gemm<T, 4, 32, 128, true>), dim3(PT.n4r, 1, B), dim3(256), 0, sg, P->F, t4r, B,
                                   arena, k0, nbl);
              else
                hipLaunchKernelGGL((k_gemm<T, 4, 32, 128>), dim3(PT.n4r, 1, B), dim3(256), 0, sg, P->F, t4r, B, arena,
                                   k0, nbl);
            }
          } else if (pre) {
            hipLaunchKernelGGL((k_gemm<T, 4, TILE, TILE, true>), dim3(PT.n4c + PT.n4r, 1, B), dim3(256), 0, sg, P->F,
                               PT.t4, B, arena, k0, nbl);
          } else {
            hipLaunchKernelGGL((k_gemm<T, 4>), dim3(PT.n4c + PT.n4r, 1, B), dim3(256), 0, sg, P->F, PT.t4, B, arena,
                               k0, nbl);
          }
        }
        const size_t lds_diag = sizeof(V) * nbl * (nbl + 1) + sizeof(double) * nbl;
        // 32-wide blocks: one wave per block (its barriers cost next to nothing and more blocks
        // share a CU); 64-wide blocks (levels of few large fronts, few workgroups): eight waves
        // share each step's updates (config 2: 7.2 -> 3.1 ms per top level)
        if (PT.ndiag > 0 && panel_lu) {
          Timed tm(P, sg, K_DIAG);
          hipLaunchKernelGGL(k_panel_lu<T>, dim3(PT.ndiag, B), dim3(64), lds_diag, sg, P->F, PT.diag, B, arena, k0,
                             PG);
        } else if (PT.ndiag > 0) {
          Timed tm(P, sg, K_DIAG);
          hipLaunchKernelGGL(k_diag<T>, dim3(PT.ndiag, B), dim3(nbl <= 32 ? 64 : 512), lds_diag, sg, P->F, PT.diag, B,
                             arena, k0, nbl, PG);
        }
        if (PT.n23 > 0 && !panel_lu) {
          Timed tm(P, sg, K_TRSM);
          if (narrow) {
            hipLaunchKernelGGL((k_gemm<T, 2, 128, 32>), dim3(PT.n23, 1, B), dim3(256), 0, sg, P->F, PT.t23, B, arena,
                               k0, nbl);
            hipLaunchKernelGGL((k_gemm<T, 3, 32, 128>), dim3(PT.n23, 1, B), dim3(256), 0, sg, P->F, PT.t23, B, arena,
                               k0, nbl);
          } else {
            hipLaunchKernelGGL((k_gemm<T, 2>), dim3(PT.n23, 1, B), dim3(256), 0, sg, P->F, PT.t23, B, arena, k0, nbl);
            hipLaunchKernelGGL((k_gemm<T, 3>), dim3(PT.n23, 1, B), dim3(256), 0, sg, P->F, PT.t23, B, arena, k0, nbl);
          }
        }
      }
    }
    if (fork && join_side(P, s)) return 3;
    // the Schur complements after the join, one bucket after the other on the launch stream
    for (const Sub& G : L.big[pi]) {
      if (G.ntiles == 0) continue;
      Timed tm(P, s, K_GEMM_S);
      if (G.pmax <= SHORT_K)   // short reduction: the contribution-block read would follow the K loop
        hipLaunchKernelGGL((k_gemm<T, 1, TILE, TILE, true>), dim3(G.ntiles, 1, B), dim3(256), 0, s, P->F, G.tiles,
                           B, arena, 0, nb);
      else
        hipLaunchKernelGGL((k_gemm<T, 1>), dim3(G.ntiles, 1, B), dim3(256), 0, s, P->F, G.tiles, B, arena, 0, nb);
    }
  }
  SAW_HIP(hipGetLastError());
  return 0;
}

// One solve call: B frequencies of R vectors each ([B][R][n]).  Factor slots: uniform groups of
// G consecutive frequencies share slot b / G (G > 0; sel lists full groups, tail says whether
// the last, partial group is solved), or frequency b reads slot fslot[b] of Bf (G == 0; sel
// lists frequencies).  rc_max bounds the vectors one workgroup takes (G > 0).
struct SolveCall {
  int B, R, G;
  const int* fslot;
  int Bf;
  const int* sel;
  int nsel;
  bool tail;
  int rc_max;       // k_grp_lds
  int rc_tri;       // the pivot-triangle, panel and k_tri_lds kernels
  int64_t ldx = 0;  // 0: x as [B][R][n]; else x as [n][ldx] (see FrontsDev::xld)
};

// One launch of a solve kernel over a set of vector blocks: block size rc (the kernels' R), the
// fronts with the set's block -> factor slot and grid row -> block mapping, the selection and
// the grid rows.
struct SolveSet {
  int rc;
  FrontsDev F;
  const int* sel;
  int ny;
  bool tail = false;   // the partial last group of a grouped call
};

// The launch sets of a kernel whose LDS need for blocks of rc vectors is lds_of(rc): groups
// of G frequencies go in blocks of rc = R 2^k vectors (2^k dividing G, rc <= rc_max, the
// largest whose LDS fits), cpg = G R / rc blocks per group; the frequencies of a partial last
// group in blocks of R.  Without groups: one set of blocks of R.  Returns the number of sets.
template <typename LdsOf>
int solve_sets(const Plan* P, const SolveCall& C, int rc_max, LdsOf&& lds_of, SolveSet (&out)[2]) {
  FrontsDev F = P->F;
  F.nvec = (int64_t)C.B * C.R;
  F.xld = C.ldx;
  if (C.G <= 0) {
    F.fslot = C.fslot;
    F.Bf = C.Bf;
    out[0] = SolveSet{C.R, F, C.sel, C.nsel};
    return C.nsel > 0 ? 1 : 0;
  }
  const int nfull = C.B / C.G, ntail = C.B - nfull * C.G;
  F.Bf = nfull + (ntail > 0 ? 1 : 0);
  int rc = C.R;
  while (rc * 2 <= std::min(rc_max, 16) && (C.G * C.R) % (rc * 2) == 0 && lds_of(rc * 2) <= (size_t)P->lds_max)
    rc *= 2;
  int k = 0;
  const int ng = C.sel ? C.nsel : nfull;
  if (ng > 0) {
    FrontsDev Fm = F;
    Fm.cpg = C.G * C.R / rc;
    Fm.vdiv = Fm.cpg;
    out[k++] = SolveSet{rc, Fm, C.sel, ng * Fm.cpg};
  }
  if (ntail > 0 && C.tail) {
    FrontsDev Ft = F;
    Ft.vdiv = C.G;
    Ft.boff = nfull * C.G;
    out[k++] = SolveSet{C.R, Ft, nullptr, ntail, true};
  }
  return k;
}

// Calls fn(std::integral_constant<int, rc>) for the block sizes the solve kernels are built for.
template <typename T, typename Fn>
bool with_rc(int rc, Fn&& fn) {
  switch (rc) {
    case 1: fn(std::integral_constant<int, 1>{}); return true;
    case 2: fn(std::integral_constant<int, 2>{}); return true;
    default: break;
  }
  if constexpr (sizeof(T) == 4) {     // shared (grouped) factors exist in complex64 only
    switch (rc) {
      case 4: fn(std::integral_constant<int, 4>{}); return true;
      case 8: fn(std::integral_constant<int, 8>{}); return true;
      case 16: fn(std::integral_constant<int, 16>{}); return true;
      default: break;
    }
  }
  return false;
}

template <typename T>
int solve_impl(Plan* P, hipStream_t s, const SolveCall& C, const void* arena, void* uarena, void* x, int transpose,
               int nb) {
  using V = typename cplx<T>::type;
  constexpr int LD = SolveTile<T>::SB + 1;   // panel kernels' transposition tile
  const int pi = sizeof(T) == 4 ? 0 : 1;
  const V* A = (const V*)arena;
  double2* U = (double2*)uarena;
  double2* X = (double2*)x;
  const int B = C.B;
  bool bad = false;           // a block size without kernels (not reached: rc is R or R 2^k <= 16)
  hipStream_t tail_s = nullptr;   // the partial last group's stream (grouped calls), see below
  // fronts of the fused set (m <= the small-front threshold): one workgroup does the whole
  // front; the others: pivot-triangle kernel + panel kernel
  // forward and backward passes take the one-workgroup set of the solve partition (the
  // transposed passes keep the factorization's partition)
  // pivot-triangle kernel (with FULL the whole front) on one launch set
  auto tri_launch = [&](const Level& L, const Sub& S, int kind, bool full, hipStream_t st, const SolveSet& Q) {
    if (Q.tail && tail_s) st = tail_s;
    const int nbl = panel_width(nb, L);
    const int ps = nbl <= 32 ? 32 : 64;     // partial sums: 4 x ps per vector
    // 64-wide panels (levels of few large fronts): 512 threads share the row-panel updates
    const dim3 blk(!full && nbl == 64 ? 512 : 256);
    auto lds_of = [&](int rc) {
      return tri_tile_bytes(sizeof(V), kind, rc, full, nbl, (int)blk.x) +
             sizeof(double2) * rc * (4 * ps + (full ? S.mmax : S.pmax));
    };
    {
      const size_t lds = lds_of(Q.rc);
      const dim3 grid(S.nf, Q.ny);
      bad |= !with_rc<T>(Q.rc, [&](auto c) {
        constexpr int RC = decltype(c)::value;
        // above 64 KB the kernel must be allowed the larger dynamic LDS (within the device
        // limit: sawfem_mf_solve checks blocks of R, larger blocks are chosen to fit)
#define SAW_TRI(K, FULL)                                                                                   \
  do {                                                                                                    \
    if (lds > 64 * 1024)                                                                                  \
      (void)hipFuncSetAttribute((const void*)k_tri<T, K, RC, FULL>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                                (int)lds);                                                                \
    hipLaunchKernelGGL((k_tri<T, K, RC, FULL>), grid, blk, lds, st, Q.F, S.fronts, B, P->n, nbl, A, U, X, Q.sel); \
  } while (0)
        if (full) {
          switch (kind) { case 0: SAW_TRI(0, true); break; case 1: SAW_TRI(1, true); break;
                          case 2: SAW_TRI(2, true); break; default: SAW_TRI(3, true); }
        } else {
          switch (kind) { case 0: SAW_TRI(0, false); break; case 1: SAW_TRI(1, false); break;
                          case 2: SAW_TRI(2, false); break; default: SAW_TRI(3, false); }
        }
#undef SAW_TRI
      });
    }
  };
  auto tri_sub = [&](const Level& L, const Sub& S, int kind, bool full, hipStream_t st) {
    if (S.nf == 0) return;
    Timed tm(P, st, (kind == 0 || kind == 2) ? K_FWD : K_BWD);
    const int nbl = panel_width(nb, L);
    const int ps = nbl <= 32 ? 32 : 64;
    const int threads = !full && nbl == 64 ? 512 : 256;
    auto lds_of = [&](int rc) {
      return tri_tile_bytes(sizeof(V), kind, rc, full, nbl, threads) +
             sizeof(double2) * rc * (4 * ps + (full ? S.mmax : S.pmax));
    };
    SolveSet sets[2];
    const int ns = solve_sets(P, C, C.rc_tri, lds_of, sets);
    for (int k = 0; k < ns; ++k) tri_launch(L, S, kind, full, st, sets[k]);
  };
  // grouped calls, one-workgroup buckets (kinds 0/1): k_grp_lds on blocks of 4 or more vectors
  // whose LDS (the padded vectors and the factor block) fits the budget; smaller blocks (a
  // partial last group) through the whole-front pivot-triangle kernel
  auto grp_lds_of = [&](const Sub& S, int rc) {
    return sizeof(double2) * (size_t)S.mmax * (rc + 1);
  };
  auto grp_ok = [&](const Sub& S) {
    return sizeof(T) == 4 && C.G > 0 && !transpose && C.rc_max >= 4 && (C.G * C.R) % 4 == 0 &&
           grp_lds_of(S, 4) <= (size_t)std::min(P->grp_lds_bytes, P->lds_max);
  };
  // algorithmic bytes of one block of rc vectors over the fronts of S (k_grp_lds, whole fronts):
  // the factor entries the kind reads once (kind 0: the unit-lower triangle's stored inverse blocks
  // below their diagonals and the column panels below them; kind 1: U12, the diagonal blocks on and
  // above their diagonals and the column panels above them) at 8 B, the vectors gathered and
  // written at 16 B per entry (kind 0: the pivot rows in and out, the children's update vectors
  // in, the front's update vector out; kind 1: all m rows in, the pivot rows out) and the rows list
  auto grp_bytes = [&](const Sub& S, int kind, int rc, int nbl) {
    double b = 0;
    for (int f : S.hfronts) {
      const double m = P->hm[f], p = P->hp[f], q = m - p;
      double ent = kind == 0 ? 0 : p * q;
      for (int J0 = 0; J0 < (int)p; J0 += nbl) {
        const double nbk = std::min<double>(nbl, p - J0);
        ent += kind == 0 ? nbk * (nbk - 1) / 2 + (m - J0 - nbk) * nbk : nbk * (nbk + 1) / 2 + J0 * nbk;
      }
      const double vec = kind == 0 ? p + P->hchild_cb[f] + p + q : m + p;
      b += 8 * ent + 16.0 * rc * vec + 4 * m;
    }
    return b;
  };
  auto grp_sub = [&](const Level& L, const Sub& S, int kind, hipStream_t st) {
    if (S.nf == 0) return;
    Timed tm(P, st, kind == 0 ? K_FWD : K_BWD);
    const int nbl = panel_width(nb, L);
    const size_t budget = (size_t)std::min(P->grp_lds_bytes, P->lds_max);
    SolveSet sets[2];
    const int ns = solve_sets(P, C, C.rc_max, [&](int rc) { return grp_lds_of(S, rc) <= budget ? grp_lds_of(S, rc)
                                                                                              : (size_t)-1; }, sets);
    for (int k = 0; k < ns; ++k) {
      const SolveSet& Q = sets[k];
      if (Q.rc < 4) {
        // blocks of one frequency (the partial last group): the pivot-triangle kernel's block size
        const int threads = 256;
        const int ps = nbl <= 32 ? 32 : 64;
        const size_t lds_t = tri_tile_bytes(sizeof(V), kind, Q.rc, true, nbl, threads) +
                             sizeof(double2) * Q.rc * (4 * ps + S.mmax);
        if (lds_t > (size_t)P->lds_max) { bad = true; continue; }
        tri_launch(L, S, kind, true, st, Q);
        continue;
      }
      const size_t lds = grp_lds_of(S, Q.rc);
      const dim3 grid(S.nf, Q.ny);
      // profiling: an event pair around this launch alone and its algorithmic bytes (the
      // roofline input of the grouped solve kernel)
      size_t gi0 = 0;
      if (P->profile) {
        while (P->ev_pool.size() < P->ev_used + 2) {
          hipEvent_t e;
          if (hipEventCreate(&e) != hipSuccess) break;
          P->ev_pool.push_back(e);
        }
        if (P->ev_pool.size() >= P->ev_used + 2) {
          gi0 = P->ev_used;
          P->ev_used += 2;
          (void)hipEventRecord(P->ev_pool[gi0], st);
          P->grp_marks.push_back({gi0, (double)Q.ny * grp_bytes(S, kind, Q.rc, nbl)});
        }
      }
      if constexpr (sizeof(T) == 4) {
#define SAW_GRP_X(K, RCV, XRV)                                                                              \
  do {                                                                                                      \
    if (lds > 64 * 1024)                                                                                    \
      (void)hipFuncSetAttribute((const void*)k_grp_lds<K, RCV, false, XRV>,                                 \
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);                      \
    hipLaunchKernelGGL((k_grp_lds<K, RCV, false, XRV>), grid, dim3(256), lds, st, Q.F, S.fronts, B, P->n, nbl, \
                       A, U, X, Q.sel);                                                                     \
  } while (0)
#define SAW_GRP(K, RCV) do { if (Q.F.xld) SAW_GRP_X(K, RCV, true); else SAW_GRP_X(K, RCV, false); } while (0)
        if (kind == 0) {
          if (Q.rc == 4) SAW_GRP(0, 4); else if (Q.rc == 8) SAW_GRP(0, 8); else SAW_GRP(0, 16);
        } else {
          if (Q.rc == 4) SAW_GRP(1, 4); else if (Q.rc == 8) SAW_GRP(1, 8); else SAW_GRP(1, 16);
        }
#undef SAW_GRP
#undef SAW_GRP_X
      }
      if (P->profile && !P->grp_marks.empty() && P->grp_marks.back().i0 == gi0 && gi0 + 1 < P->ev_pool.size())
        (void)hipEventRecord(P->ev_pool[gi0 + 1], st);
    }
  };
  auto panel_lds_of = [&](int kind, int rc) {
    return sizeof(double2) * rc * (256 + 256) + ((kind >= 2) ? sizeof(V) * 64 * LD : 0);
  };
  // off-diagonal panel kernel on one launch set
  auto panel_launch = [&](const Sub& S, int kind, hipStream_t st, const SolveSet& Q) {
    if (Q.tail && tail_s) st = tail_s;
    const bool out_cb = (kind == 0 || kind == 2);
    const int nt = out_cb ? S.nstiles_cb : S.nstiles_p;
    const int* tl = out_cb ? S.stiles_cb : S.stiles_p;
    if (nt == 0) return;
    {
      const size_t lds = panel_lds_of(kind, Q.rc);
      const dim3 grid(nt, Q.ny);
      bad |= !with_rc<T>(Q.rc, [&](auto c) {
        constexpr int RC = decltype(c)::value;
#define SAW_PANEL(K)                                                                                      \
  do {                                                                                                    \
    if (lds > 64 * 1024)                                                                                  \
      (void)hipFuncSetAttribute((const void*)k_panel<T, K, RC>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                                (int)lds);                                                                \
    hipLaunchKernelGGL((k_panel<T, K, RC>), grid, dim3(256), lds, st, Q.F, tl, B, P->n, A, U, X, Q.sel); \
  } while (0)
        switch (kind) { case 0: SAW_PANEL(0); break; case 1: SAW_PANEL(1); break;
                        case 2: SAW_PANEL(2); break; default: SAW_PANEL(3); }
#undef SAW_PANEL
      });
    }
  };
  auto panel_sub = [&](const Sub& S, int kind, hipStream_t st) {
    const bool out_cb = (kind == 0 || kind == 2);
    if ((out_cb ? S.nstiles_cb : S.nstiles_p) == 0) return;
    Timed tm(P, st, out_cb ? K_FWD : K_BWD);
    SolveSet sets[2];
    const int ns = solve_sets(P, C, C.rc_tri, [&](int rc) { return panel_lds_of(kind, rc); }, sets);
    for (int k = 0; k < ns; ++k) panel_launch(S, kind, st, sets[k]);
  };
  // grouped calls, two-kernel buckets (kinds 0/1): the pivot rows by k_grp_lds (PIV) and the
  // off-diagonal panel by k_grp_panel, on blocks of 4 or more vectors whose pivot rows fit the
  // LDS; forward: pivot rows, then the panel (which also gathers the children's updates of the
  // contribution-block rows); backward: the panel, then the pivot rows.  A partial last group
  // goes through the pivot-triangle + panel pair.
  auto grp_big_lds_of = [&](const Sub& S, int rc) { return sizeof(double2) * (size_t)S.pmax * (rc + 1); };
  auto grp_big_ok = [&](const Sub& S) {
    return sizeof(T) == 4 && C.G > 0 && !transpose && C.rc_max >= 4 && (C.G * C.R) % 4 == 0 &&
           P->grp_lds_bytes > 0 && S.pmax <= P->grp_big_pmax && grp_big_lds_of(S, 4) <= (size_t)P->lds_max;
  };
  auto grp_big = [&](const Level& L, const Sub& S, int kind, hipStream_t st) {
    if (S.nf == 0) return;
    Timed tm(P, st, kind == 0 ? K_FWD : K_BWD);
    const int nbl = panel_width(nb, L);
    SolveSet sets[2];
    const int ns = solve_sets(P, C, C.rc_max, [&](int rc) { return grp_big_lds_of(S, rc); }, sets);
    for (int k = 0; k < ns; ++k) {
      const SolveSet& Q = sets[k];
      if (Q.rc < 4) {
        if (kind == 0) { tri_launch(L, S, 0, false, st, Q); panel_launch(S, 0, st, Q); }
        else { panel_launch(S, 1, st, Q); tri_launch(L, S, 1, false, st, Q); }
        continue;
      }
      if constexpr (sizeof(T) == 4) {
        const size_t lds = grp_big_lds_of(S, Q.rc);
        const int ntile = kind == 0 ? S.nstiles_cb : S.nstiles_p;
        const int* tl = kind == 0 ? S.stiles_cb : S.stiles_p;
#define SAW_GRP_PIV_X(K, RCV, XRV)                                                                                \
  do {                                                                                                            \
    if (lds > 64 * 1024)                                                                                          \
      (void)hipFuncSetAttribute((const void*)k_grp_lds<K, RCV, true, XRV>,                                        \
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);                            \
    hipLaunchKernelGGL((k_grp_lds<K, RCV, true, XRV>), dim3(S.nf, Q.ny), dim3(256), lds, st, Q.F, S.fronts, B,    \
                       P->n, nbl, A, U, X, Q.sel);                                                                \
  } while (0)
#define SAW_GRP_PAN_X(K, RCV, XRV)                                                                                \
  do {                                                                                                            \
    if (ntile > 0)                                                                                                \
      hipLaunchKernelGGL((k_grp_panel<K, RCV, XRV>), dim3(ntile, Q.ny), dim3(16 * RCV), sizeof(double2) * GRP_VC * \
                         (RCV + 1), st, Q.F, tl, B, P->n, A, U, X, Q.sel);                                        \
  } while (0)
#define SAW_GRP_PIV(K, RCV) do { if (Q.F.xld) SAW_GRP_PIV_X(K, RCV, true); else SAW_GRP_PIV_X(K, RCV, false); } while (0)
#define SAW_GRP_PAN(K, RCV) do { if (Q.F.xld) SAW_GRP_PAN_X(K, RCV, true); else SAW_GRP_PAN_X(K, RCV, false); } while (0)
        if (kind == 0) {
          if (Q.rc == 4) { SAW_GRP_PIV(0, 4); SAW_GRP_PAN(0, 4); }
          else if (Q.rc == 8) { SAW_GRP_PIV(0, 8); SAW_GRP_PAN(0, 8); }
          else { SAW_GRP_PIV(0, 16); SAW_GRP_PAN(0, 16); }
        } else {
          if (Q.rc == 4) { SAW_GRP_PAN(1, 4); SAW_GRP_PIV(1, 4); }
          else if (Q.rc == 8) { SAW_GRP_PAN(1, 8); SAW_GRP_PIV(1, 8); }
          else { SAW_GRP_PAN(1, 16); SAW_GRP_PIV(1, 16); }
        }
#undef SAW_GRP_PIV
#undef SAW_GRP_PAN
#undef SAW_GRP_PIV_X
#undef SAW_GRP_PAN_X
      }
    }
  };
  // The one-workgroup set and the size buckets of a level are independent: with more than one,
  // they run on the launch stream and the side streams between a fork and a join.
  auto stream_of = [&](bool fork, size_t unit) { return (!fork || unit == 0) ? s : P->side[(unit - 1) % Plan::NSIDE]; };
  const int kf = transpose ? 2 : 0, kb = transpose ? 3 : 1;
  // one-workgroup buckets whose factor block fits the LDS budget: whole-front solves from LDS
  auto lds_bytes = [&](const Sub& S, int rc) {
    return sizeof(V) * (size_t)S.mpmax + sizeof(double2) * rc * (256 + 4 * 64);
  };
  auto lds_ok = [&](const Sub& S) {
    return !transpose && P->tri_lds_bytes > 0 && S.mmax <= 256 && S.mpmax <= 256 * tri_lds_npf<T>() &&
           sizeof(V) * (size_t)S.mpmax <= (size_t)P->tri_lds_bytes && lds_bytes(S, C.R) <= (size_t)P->lds_max;
  };
  auto tri_lds = [&](const Level& L, const Sub& S, int kind, hipStream_t st) {
    if (S.nf == 0) return;
    Timed tm(P, st, kind == 0 ? K_FWD : K_BWD);
    const int nbl = panel_width(nb, L);
    SolveSet sets[2];
    const int ns = solve_sets(P, C, C.rc_tri, [&](int rc) { return lds_bytes(S, rc); }, sets);
    for (int k = 0; k < ns; ++k) {
      const SolveSet& Q = sets[k];
      const hipStream_t sq = (Q.tail && tail_s) ? tail_s : st;
      const size_t lds = lds_bytes(S, Q.rc);
      // persistent grid: as many workgroups as the CUs hold at this LDS size (at most 8 a CU)
      const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(8, (size_t)P->lds_per_cu / lds));
      const int64_t nitems = (int64_t)S.nf * Q.ny;
      const unsigned grid = (unsigned)std::min<int64_t>(nitems, (int64_t)P->ncu * per_cu);
      bad |= !with_rc<T>(Q.rc, [&](auto c) {
        constexpr int RC = decltype(c)::value;
        if (kind == 0) {
          if (lds > 64 * 1024)
            (void)hipFuncSetAttribute((const void*)k_tri_lds<T, 0, RC>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds);
          hipLaunchKernelGGL((k_tri_lds<T, 0, RC>), dim3(grid), dim3(256), lds, sq, Q.F, S.fronts, S.nf, B, P->n,
                             nbl, A, U, X, Q.sel, Q.ny);
        } else {
          if (lds > 64 * 1024)
            (void)hipFuncSetAttribute((const void*)k_tri_lds<T, 1, RC>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds);
          hipLaunchKernelGGL((k_tri_lds<T, 1, RC>), dim3(grid), dim3(256), lds, sq, Q.F, S.fronts, S.nf, B, P->n,
                             nbl, A, U, X, Q.sel, Q.ny);
        }
      });
    }
  };
  // A grouped call's partial last group (its own vectors, update vectors and factor slot) runs on
  // a stream of its own, concurrently with the full groups: forked here, joined at the end.  Its
  // launches keep their order on that stream, which is all they depend on.
  if (C.G > 0 && C.tail && B % C.G != 0 && P->overlap) {
    if (ensure_tail(P)) return 3;
    SAW_HIP(hipEventRecord(P->tail_fork, s));
    SAW_HIP(hipStreamWaitEvent(P->tail_stream, P->tail_fork, 0));
    tail_s = P->tail_stream;
  }
  for (size_t li = 0; li < P->levels.size(); ++li) {
    const Level& L = P->levels[li];
    P->cur_level = (int)li;
    const size_t nfull = L.ssmall[pi].size();
    const bool fork = P->overlap && L.sbig[pi].size() + nfull > 1;
    if (fork && (ensure_side(P) || fork_side(P, s))) return 3;
    for (size_t gi = 0; gi < nfull; ++gi) {
      const Sub& S = L.ssmall[pi][gi];
      if (grp_ok(S)) grp_sub(L, S, 0, stream_of(fork, gi));
      else if (lds_ok(S)) tri_lds(L, S, 0, stream_of(fork, gi));
      else tri_sub(L, S, kf, true, stream_of(fork, gi));
    }
    for (size_t gi = 0; gi < L.sbig[pi].size(); ++gi) {
      const hipStream_t st = stream_of(fork, gi + nfull);
      if (grp_big_ok(L.sbig[pi][gi])) {
        grp_big(L, L.sbig[pi][gi], 0, st);
      } else {
        tri_sub(L, L.sbig[pi][gi], kf, false, st);
        panel_sub(L.sbig[pi][gi], kf, st);
      }
    }
    if (fork && join_side(P, s)) return 3;
  }
  for (int l = (int)P->levels.size() - 1; l >= 0; --l) {
    const Level& L = P->levels[l];
    P->cur_level = l;
    if (!transpose) {
      // the forward partition: one-workgroup buckets (LDS-staged where their block fits, else
      // the whole-front pivot-triangle kernel, whose CB-to-pivot product splits the CB columns
      // over the workgroup), the rest through the panel + triangle pair
      const size_t nfull = L.ssmall[pi].size();
      const bool fork = P->overlap && L.sbig[pi].size() + nfull > 1;
      if (fork && (ensure_side(P) || fork_side(P, s))) return 3;
      for (size_t gi = 0; gi < L.sbig[pi].size(); ++gi) {
        const hipStream_t st = stream_of(fork, gi);
        if (grp_big_ok(L.sbig[pi][gi])) {
          grp_big(L, L.sbig[pi][gi], 1, st);
        } else {
          panel_sub(L.sbig[pi][gi], kb, st);
          tri_sub(L, L.sbig[pi][gi], kb, false, st);
        }
      }
      for (size_t gi = 0; gi < nfull; ++gi) {
        const Sub& S = L.ssmall[pi][gi];
        const hipStream_t st = stream_of(fork, gi + L.sbig[pi].size());
        if (grp_ok(S)) grp_sub(L, S, 1, st);
        else if (lds_ok(S)) tri_lds(L, S, 1, st);
        else tri_sub(L, S, kb, true, st);
      }
      if (fork && join_side(P, s)) return 3;
      continue;
    }
    const bool full = L.small[pi].nf > 0;
    const bool fork = P->overlap && L.big[pi].size() + (full ? 1 : 0) > 1;
    if (fork && (ensure_side(P) || fork_side(P, s))) return 3;
    for (size_t gi = 0; gi < L.big[pi].size(); ++gi) {
      const hipStream_t st = stream_of(fork, gi + (full ? 1 : 0));
      panel_sub(L.big[pi][gi], kb, st);
      tri_sub(L, L.big[pi][gi], kb, false, st);
    }
    if (full) tri_sub(L, L.small[pi], kb, true, s);
    if (fork && join_side(P, s)) return 3;
  }
  if (tail_s) {
    SAW_HIP(hipEventRecord(P->tail_join, tail_s));
    SAW_HIP(hipStreamWaitEvent(s, P->tail_join, 0));
  }
  if (bad) { sawfem_set_error("sawfem_mf_solve: no kernels for the block size"); return 1; }
  SAW_HIP(hipGetLastError());
  return 0;
}

// LDS of the pivot-block kernel for blocks of nrhs vectors: the widest tile (64 x 65 complex64
// or 64 x 33 complex128, both <= 64 * 33 * 16 bytes), partial sums and the pivot entries of the
// vectors.  Grouped solves take larger blocks only where they fit.
int solve_lds_check(const Plan* P, int nrhs) {
  const size_t tile = 64 * 33 * 16;
  if ((size_t)nrhs * (P->pmax + 256) * sizeof(double2) + tile > (size_t)P->lds_max) {
    sawfem_set_error("sawfem_mf_solve: front too large for the LDS solve");
    return 1;
  }
  return 0;
}

}  // namespace

extern "C" {

int sawfem_mf_destroy(void* plan);

// Build a numeric plan from the host symbolic analysis.  All arrays are host pointers.
//   m, p, rows_ptr(nf+1), rows: front orders, pivot counts and row lists (active ids)
//   child_ptr(nf+1), child_ids: children per front; inv_ptr(nchild+1), inv: per child
//     slot, for each parent row, the child contribution-block index or -1
//   lvl_ptr(nlev+1), lvl_fronts: fronts per level, leaves first
//   ent_*: A_aa slot -> (front, column-major local offset), sorted by level then front;
//     ent_lvl_ptr(nlev+1) delimits the entries of each level
int sawfem_mf_create(void** out, int n, int nf, const int* m, const int* p, const int64_t* rows_ptr, const int* rows,
                     const int* child_ptr, const int* child_ids, const int64_t* inv_ptr, const int* inv, int nlev,
                     const int* lvl_ptr, const int* lvl_fronts, int64_t nent, const int* ent_front,
                     const int64_t* ent_slot, const int* ent_loc, const int64_t* ent_lvl_ptr) {
  if (!out || n <= 0 || nf <= 0 || nlev <= 0) {
    sawfem_set_error("sawfem_mf_create: invalid arguments");
    return 1;
  }
  Plan* P = new Plan();
  P->n = n; P->nf = nf;
  std::vector<int64_t> off(nf), uoff(nf);
  int64_t acc = 0, uacc = 0;
  for (int f = 0; f < nf; ++f) {
    if (m[f] <= 0 || p[f] < 0 || p[f] > m[f] || rows_ptr[f + 1] - rows_ptr[f] != m[f]) {
      sawfem_mf_destroy(P);
      sawfem_set_error("sawfem_mf_create: inconsistent front");
      return 1;
    }
    off[f] = acc; acc += (int64_t)m[f] * m[f];
    uoff[f] = uacc; uacc += m[f] - p[f];
    P->mmax = std::max(P->mmax, m[f]);
    P->pmax = std::max(P->pmax, p[f]);
  }
  P->arena_elems = acc;
  P->uarena_elems = uacc;
  const int nchild = child_ptr[nf];
  int* dm; int* dp; int64_t* doff; int64_t* duoff; int64_t* drp; int* drows; int* dcp; int* dci; int64_t* dip; int* dinv;
  int rc = 0;
  rc |= upload(P, &dm, m, nf);
  rc |= upload(P, &dp, p, nf);
  rc |= upload(P, &doff, off.data(), nf);
  rc |= upload(P, &duoff, uoff.data(), nf);
  rc |= upload(P, &drp, rows_ptr, nf + 1);
  rc |= upload(P, &drows, rows, rows_ptr[nf]);
  rc |= upload(P, &dcp, child_ptr, nf + 1);
  rc |= upload(P, &dci, child_ids, nchild);
  rc |= upload(P, &dip, inv_ptr, nchild + 1);
  rc |= upload(P, &dinv, inv, inv_ptr[nchild]);
  rc |= upload(P, &P->lvl_fronts, lvl_fronts, lvl_ptr[nlev]);
  rc |= upload(P, &P->ent_front, ent_front, nent);
  rc |= upload(P, &P->ent_slot, ent_slot, nent);
  rc |= upload(P, &P->ent_loc, ent_loc, nent);
  int zero = 0;
  rc |= upload(P, &P->npert, &zero, 1);
  const unsigned big = 0x7f7fffffu;   // FLT_MAX
  rc |= upload(P, &P->pmin, &big, 1);
  {
    std::vector<double> rs0(3 * (size_t)n, 0.0);
    rc |= upload(P, &P->rscale, rs0.data(), 3 * (int64_t)n);
  }
  P->nent = nent;
  if (rc) { sawfem_mf_destroy(P); return 3; }
  // per-front entry ranges (entries are grouped by front within each level)
  {
    std::vector<int64_t> lo(nf, 0), hi(nf, 0);
    for (int64_t e = nent - 1; e >= 0; --e) lo[ent_front[e]] = e;
    for (int64_t e = 0; e < nent; ++e) hi[ent_front[e]] = e + 1;
    std::vector<char> seen(nf, 0);
    for (int64_t e = 0; e < nent; ++e) seen[ent_front[e]] = 1;
    for (int f = 0; f < nf; ++f)
      if (!seen[f]) lo[f] = hi[f] = 0;
    for (int64_t e = 0; e < nent; ++e)
      if (e < lo[ent_front[e]] || e >= hi[ent_front[e]] || (e > lo[ent_front[e]] && ent_front[e - 1] != ent_front[e])) {
        sawfem_mf_destroy(P);
        sawfem_set_error("sawfem_mf_create: entries of a front are not contiguous");
        return 1;
      }
    rc |= upload(P, &P->fent_lo, lo.data(), nf);
    rc |= upload(P, &P->fent_hi, hi.data(), nf);
    if (rc) { sawfem_mf_destroy(P); return 3; }
    P->hfent_lo = lo;
    P->hfent_hi = hi;
  }
  P->hchild_ptr.assign(child_ptr, child_ptr + nf + 1);
  P->hchild_cb.assign(nf, 0);
  for (int f = 0; f < nf; ++f)
    for (int c = child_ptr[f]; c < child_ptr[f + 1]; ++c) P->hchild_cb[f] += m[child_ids[c]] - p[child_ids[c]];
  P->hent_loc.assign(ent_loc, ent_loc + nent);
  P->hm.assign(m, m + nf);
  P->hp.assign(p, p + nf);
  P->hlvl_ptr.assign(lvl_ptr, lvl_ptr + nlev + 1);
  P->hlvl_fronts.assign(lvl_fronts, lvl_fronts + lvl_ptr[nlev]);
  {
    int dev = 0, lds = 64 * 1024;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
      lds = 64 * 1024;
    P->lds_max = lds;
    int ldscu = lds, ncu = 1;
    if (hipDeviceGetAttribute(&ldscu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess)
      ldscu = lds;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1) ncu = 1;
    P->lds_per_cu = std::max(ldscu, lds);
    P->ncu = ncu;
    for (int pi = 0; pi < 2; ++pi) {
      const size_t sz = pi == 0 ? sizeof(c64) : sizeof(c128);
      int mm = 0;
      while (mm < SMALL_MMAX && sz * (size_t)(mm + 1) * (mm + 2) + sizeof(int) * (mm + 1) <= (size_t)lds) ++mm;
      P->small_limit[pi] = mm;
    }
    const void* fk[2] = {(const void*)k_small_front<float>, (const void*)k_small_front<double>};
    for (int pi = 0; pi < 2; ++pi) {
      const size_t need = (pi == 0 ? sizeof(c64) : sizeof(c128)) * (size_t)P->small_limit[pi] * (P->small_limit[pi] + 1)
                          + sizeof(int) * P->small_limit[pi];
      if (need > 64 * 1024 && hipFuncSetAttribute(fk[pi], hipFuncAttributeMaxDynamicSharedMemorySize, (int)need) != hipSuccess) {
        (void)hipGetLastError();
        int mm = 0;
        const size_t sz = pi == 0 ? sizeof(c64) : sizeof(c128);
        while (sz * (size_t)(mm + 1) * (mm + 2) + sizeof(int) * (mm + 1) <= 64 * 1024) ++mm;
        P->small_limit[pi] = mm;
      }
    }
  }
  P->F = FrontsDev{dm, dp, doff, duoff, drp, drows, dcp, dci, dip, dinv};
  for (int l = 0; l < nlev; ++l) {
    Level L{};
    L.nf = lvl_ptr[l + 1] - lvl_ptr[l];
    L.fronts = P->lvl_fronts + lvl_ptr[l];
    for (int k = lvl_ptr[l]; k < lvl_ptr[l + 1]; ++k) {
      const int f = lvl_fronts[k];
      L.mmax = std::max(L.mmax, m[f]);
      L.pmax = std::max(L.pmax, p[f]);
      L.cbmax = std::max(L.cbmax, m[f] - p[f]);
    }
    L.e0 = ent_lvl_ptr[l];
    L.e1 = ent_lvl_ptr[l + 1];
    P->levels.push_back(L);
  }
  // the fused small-front path is opt-in (sawfem_mf_set_small): on the config-2 sweep it
  // measured slower than the panel kernels (one 512-thread workgroup per CU, latency-bound)
  for (int pi = 0; pi < 2; ++pi)
    if (build_partition(P, pi, 0)) { sawfem_mf_destroy(P); return 3; }
  *out = P;
  return 0;
}

int sawfem_mf_destroy(void* plan) {
  Plan* P = (Plan*)plan;
  if (!P) return 0;
  if (P->fork_ev || P->side[0] || P->tail_stream) (void)hipDeviceSynchronize();
  for (int k = 0; k < Plan::NSIDE; ++k) {
    if (P->side[k]) (void)hipStreamDestroy(P->side[k]);
    if (P->join_ev[k]) (void)hipEventDestroy(P->join_ev[k]);
  }
  if (P->fork_ev) (void)hipEventDestroy(P->fork_ev);
  if (P->tail_stream) (void)hipStreamDestroy(P->tail_stream);
  if (P->tail_fork) (void)hipEventDestroy(P->tail_fork);
  if (P->tail_join) (void)hipEventDestroy(P->tail_join);
  for (void* d : P->owned) (void)hipFree(d);
  for (int pi = 0; pi < 2; ++pi) {
    for (void* d : P->part_owned[pi]) (void)hipFree(d);
    for (void* d : P->panel_owned[pi]) (void)hipFree(d);
  }
  for (hipEvent_t e : P->ev_pool) (void)hipEventDestroy(e);
  drop_graphs(P);
  if (P->capture) (void)hipStreamDestroy(P->capture);
  delete P;
  return 0;
}

// Largest front order factored by the fused small-front kernel (prec 0: complex64, 1:
// complex128); max_m <= 0 (the default) sends every front through the panel path.  The value
// in use (the request clipped to what the device LDS holds, at most 128) is returned, or -1.
int sawfem_mf_set_small(void* plan, int prec, int max_m) {
  Plan* P = (Plan*)plan;
  if (!P || (prec != 0 && prec != 1)) { sawfem_set_error("sawfem_mf_set_small: invalid arguments"); return -1; }
  drop_graphs(P);
  if (build_partition(P, prec, max_m)) return -1;
  return P->small_mmax[prec];
}

// In the forward passes, fronts of order <= max_m are solved by one workgroup each (pivot
// triangle and off-diagonal panel together, k_tri with FULL) instead of the pivot-triangle +
// panel kernel pair; 0 keeps the pair for every front that the factorization did not fuse.
// Returns the order in use.
// One-workgroup solve buckets (fronts up to the solve_full order) whose largest factor block
// (m p entries of the front precision) fits max_bytes of LDS are solved with that block staged in
// LDS, forward and backward; 0 keeps the pivot-triangle kernels.  Returns the bytes in use.
int sawfem_mf_set_tri_lds(void* plan, int max_bytes) {
  Plan* P = (Plan*)plan;
  if (!P || max_bytes < 0) { sawfem_set_error("sawfem_mf_set_tri_lds: invalid arguments"); return -1; }
  P->tri_lds_bytes = std::min(max_bytes, P->lds_max);
  return P->tri_lds_bytes;
}

// 32-wide pivot panels of size buckets whose largest front has order <= max_m are factored by
// one kernel per panel (diagonal block and both panel solves, one wave per front and frequency)
// instead of the diagonal-block kernel and two MFMA panel-solve GEMMs; 0 keeps the latter.
int sawfem_mf_set_panel_lu(void* plan, int max_m) {
  Plan* P = (Plan*)plan;
  if (!P || max_m < 0) { sawfem_set_error("sawfem_mf_set_panel_lu: invalid arguments"); return -1; }
  drop_graphs(P);
  P->panel_lu_m = max_m;
  return P->panel_lu_m;
}

int sawfem_mf_set_solve_full(void* plan, int max_m) {
  Plan* P = (Plan*)plan;
  if (!P || max_m < 0) { sawfem_set_error("sawfem_mf_set_solve_full: invalid arguments"); return -1; }
  drop_graphs(P);
  P->solve_full_m = std::min(max_m, 1024);
  for (int pi = 0; pi < 2; ++pi)
    if (build_partition(P, pi, P->small_mmax[pi])) return -1;
  return P->solve_full_m;
}

int64_t sawfem_mf_arena_elems(void* plan) { return plan ? ((Plan*)plan)->arena_elems : -1; }
int64_t sawfem_mf_uarena_elems(void* plan) { return plan ? ((Plan*)plan)->uarena_elems : -1; }

// Number of pivots the guard replaced since the plan was created (or the counter last reset).
int sawfem_mf_perturbed(void* plan, int reset) {
  Plan* P = (Plan*)plan;
  int h = 0;
  if (!P) return -1;
  if (hipMemcpy(&h, P->npert, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) return -1;
  if (reset) { int z = 0; (void)hipMemcpy(P->npert, &z, sizeof(int), hipMemcpyHostToDevice); }
  return h;
}

// Pivot statistics since the last reset: pivots replaced and the smallest |pivot| / row scale.
int sawfem_mf_pivot_stats(void* plan, int reset, int64_t* perturbed, double* min_ratio) {
  Plan* P = (Plan*)plan;
  if (!P || !perturbed || !min_ratio) { sawfem_set_error("sawfem_mf_pivot_stats: invalid arguments"); return 1; }
  int h = 0;
  unsigned u = 0;
  SAW_HIP(hipMemcpy(&h, P->npert, sizeof(int), hipMemcpyDeviceToHost));
  SAW_HIP(hipMemcpy(&u, P->pmin, sizeof(unsigned), hipMemcpyDeviceToHost));
  float r;
  std::memcpy(&r, &u, sizeof(r));
  *perturbed = h;
  *min_ratio = r;
  if (reset) {
    const int z = 0;
    const unsigned big = 0x7f7fffffu;
    SAW_HIP(hipMemcpy(P->npert, &z, sizeof(int), hipMemcpyHostToDevice));
    SAW_HIP(hipMemcpy(P->pmin, &big, sizeof(unsigned), hipMemcpyHostToDevice));
  }
  return 0;
}

// prec: 0 = complex64 fronts (mixed path), 1 = complex128 fronts (plain path).
int sawfem_mf_factor(void* plan, void* stream, int prec, int B, const double* omega, double beta, const void* Kaa,
                     const void* Maa, const void* Daa, void* arena, int nb) {
  Plan* P = (Plan*)plan;
  if (!P || B <= 0 || !omega || !Kaa || !Maa || !arena || (nb != 32 && nb != 64) || (prec != 0 && prec != 1)) {
    sawfem_set_error("sawfem_mf_factor: invalid arguments");
    return 1;
  }
  if (prec == 1 && nb > 32) { sawfem_set_error("sawfem_mf_factor: nb > 32 with complex128"); return 1; }
  hipStream_t s = (hipStream_t)stream;
  return prec == 0 ? factor_impl<float>(P, s, B, omega, beta, Kaa, Maa, Daa, arena, nb)
                   : factor_impl<double>(P, s, B, omega, beta, Kaa, Maa, Daa, arena, nb);
}

int sawfem_mf_factor_graph(void* plan, void* stream, int prec, int B, const double* omega, double beta,
                           const void* Kaa, const void* Maa, const void* Daa, void* arena, int nb) {
  Plan* P = (Plan*)plan;
  if (!P || B <= 0 || !omega || !Kaa || !Maa || !arena || (nb != 32 && nb != 64) || (prec != 0 && prec != 1) ||
      (prec == 1 && nb > 32)) {
    sawfem_set_error("sawfem_mf_factor_graph: invalid arguments");
    return 1;
  }
  hipStream_t s = (hipStream_t)stream;
  // per-launch event timing does not go into graphs: with every launch profiled the factorization
  // runs eagerly
  if (P->profile == 1) return sawfem_mf_factor(plan, stream, prec, B, omega, beta, Kaa, Maa, Daa, arena, nb);
  Plan::FactorGraph* hit = nullptr;
  for (auto& g : P->graphs)
    if (g.prec == prec && g.B == B && g.nb == nb && g.omega == omega && g.Kaa == Kaa && g.Maa == Maa &&
        g.Daa == Daa && g.arena == arena && g.beta == beta)
      hit = &g;
  if (!hit) {
    // everything the launches need from the host side first (work lists, side streams), so the
    // capture holds launches and stream dependencies only
    const int pi = prec == 0 ? 0 : 1;
    if (P->panel_nb[pi] != nb && build_panel_lists(P, pi, nb)) return 1;
    if (P->overlap && ensure_side(P)) return 3;
    if (!P->capture) SAW_HIP(hipStreamCreateWithFlags(&P->capture, hipStreamNonBlocking));
    SAW_HIP(hipStreamBeginCapture(P->capture, hipStreamCaptureModeThreadLocal));
    const int rc = prec == 0 ? factor_impl<float>(P, P->capture, B, omega, beta, Kaa, Maa, Daa, arena, nb)
                             : factor_impl<double>(P, P->capture, B, omega, beta, Kaa, Maa, Daa, arena, nb);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(P->capture, &graph);
    if (rc || ec != hipSuccess || !graph) {
      if (graph) (void)hipGraphDestroy(graph);
      if (!rc) sawfem_set_error("sawfem_mf_factor_graph: capture failed");
      return rc ? rc : 3;
    }
    hipGraphExec_t exec = nullptr;
    if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
      (void)hipGraphDestroy(graph);
      sawfem_set_error("sawfem_mf_factor_graph: graph instantiation failed");
      return 3;
    }
    size_t nodes = 0;
    (void)hipGraphGetNodes(graph, nullptr, &nodes);
    if ((int)P->graphs.size() >= Plan::NGRAPH) {      // replace the least recently used
      auto lru = std::min_element(P->graphs.begin(), P->graphs.end(),
                                  [](const Plan::FactorGraph& a, const Plan::FactorGraph& b) { return a.used < b.used; });
      (void)hipGraphExecDestroy(lru->exec);
      (void)hipGraphDestroy(lru->graph);
      P->graphs.erase(lru);
    }
    P->graphs.push_back(Plan::FactorGraph{prec, B, nb, omega, Kaa, Maa, Daa, arena, beta, graph, exec, nodes, 0});
    hit = &P->graphs.back();
  }
  hit->used = ++P->graph_clock;
  P->graph_nodes = hit->nodes;
  SAW_HIP(hipGraphLaunch(hit->exec, s));
  return 0;
}

int64_t sawfem_mf_graph_nodes(void* plan) { return plan ? (int64_t)((Plan*)plan)->graph_nodes : -1; }

// x: [B][nrhs][n] complex128 right-hand sides (nrhs = 1 or 2 per frequency, sharing that
// frequency's factors), overwritten with the solutions.  nb must be the value the arena was
// factored with.  uarena: uarena_elems * B * nrhs complex128.  bsel (device int[nsel], or
// NULL for all B): solve only these frequencies of the chunk; the others' x is untouched.
int sawfem_mf_solve(void* plan, void* stream, int prec, int B, int nrhs, const void* arena, void* uarena, void* x,
                    int transpose, int nb, const int* bsel, int nsel) {
  return sawfem_mf_solve_shared(plan, stream, prec, B, nrhs, arena, B, nullptr, uarena, x, transpose, nb, bsel, nsel);
}

int sawfem_mf_solve_shared(void* plan, void* stream, int prec, int B, int nrhs, const void* arena, int Bf,
                           const int* fslot, void* uarena, void* x, int transpose, int nb, const int* bsel, int nsel) {
  Plan* P = (Plan*)plan;
  if (!bsel) nsel = B;
  if (!P || B <= 0 || !arena || !uarena || !x || (prec != 0 && prec != 1) || (nb != 32 && nb != 64) ||
      (prec == 1 && nb > 32) || (nrhs != 1 && nrhs != 2) || nsel < 0 || nsel > B || Bf <= 0 || (!fslot && Bf != B)) {
    sawfem_set_error("sawfem_mf_solve: invalid arguments");
    return 1;
  }
  if (solve_lds_check(P, nrhs)) return 1;
  if (nsel == 0) return 0;
  const SolveCall C{B, nrhs, 0, fslot, Bf, bsel, nsel, false, nrhs, nrhs};
  hipStream_t s = (hipStream_t)stream;
  return prec == 0 ? solve_impl<float>(P, s, C, arena, uarena, x, transpose, nb)
                   : solve_impl<double>(P, s, C, arena, uarena, x, transpose, nb);
}

int sawfem_mf_solve_groups(void* plan, void* stream, int B, int nrhs, const void* arena, int G, void* uarena,
                           void* x, int transpose, int nb, const int* gsel, int ngsel, int tail) {
  Plan* P = (Plan*)plan;
  const int nfull = G > 0 ? B / G : 0;
  if (!gsel) ngsel = nfull;
  if (!P || B <= 0 || G <= 0 || !arena || !uarena || !x || (nb != 32 && nb != 64) || (nrhs != 1 && nrhs != 2) ||
      ngsel < 0 || ngsel > nfull || (tail != 0 && tail != 1)) {
    sawfem_set_error("sawfem_mf_solve_groups: invalid arguments");
    return 1;
  }
  if (solve_lds_check(P, nrhs)) return 1;
  const bool do_tail = tail && B % G != 0;
  if (ngsel == 0 && !do_tail) return 0;
  const SolveCall C{B, nrhs, G, nullptr, 0, gsel, ngsel, do_tail, P->group_rc, P->group_rc_tri};
  return solve_impl<float>(P, (hipStream_t)stream, C, arena, uarena, x, transpose, nb);
}

// sawfem_mf_solve_groups on vectors stored row-major: entry (row i, vector b nrhs + r) of x at
// x[i ldx + b nrhs + r], ldx >= B nrhs; forward + backward solves.  The update vectors in uarena
// are stored row-major too, so a call's uarena is not shared with a vector-major call's.
int sawfem_mf_solve_groups_rows(void* plan, void* stream, int B, int nrhs, const void* arena, int G, void* uarena,
                                void* x, int64_t ldx, int nb, const int* gsel, int ngsel, int tail) {
  Plan* P = (Plan*)plan;
  const int nfull = G > 0 ? B / G : 0;
  if (!gsel) ngsel = nfull;
  if (!P || B <= 0 || G <= 0 || !arena || !uarena || !x || (nb != 32 && nb != 64) || (nrhs != 1 && nrhs != 2) ||
      ngsel < 0 || ngsel > nfull || (tail != 0 && tail != 1) || ldx < (int64_t)B * nrhs) {
    sawfem_set_error("sawfem_mf_solve_groups_rows: invalid arguments");
    return 1;
  }
  if (solve_lds_check(P, nrhs)) return 1;
  const bool do_tail = tail && B % G != 0;
  if (ngsel == 0 && !do_tail) return 0;
  SolveCall C{B, nrhs, G, nullptr, 0, gsel, ngsel, do_tail, P->group_rc, P->group_rc_tri};
  C.ldx = ldx;
  return solve_impl<float>(P, (hipStream_t)stream, C, arena, uarena, x, 0, nb);
}

int sawfem_mf_set_group_rc(void* plan, int rc_max) {
  Plan* P = (Plan*)plan;
  if (!P || rc_max < 1) { sawfem_set_error("sawfem_mf_set_group_rc: invalid arguments"); return -1; }
  P->group_rc = std::min(rc_max, 16);
  return P->group_rc;
}

int sawfem_mf_set_group_big(void* plan, int max_p) {
  Plan* P = (Plan*)plan;
  if (!P || max_p < 0) { sawfem_set_error("sawfem_mf_set_group_big: invalid arguments"); return -1; }
  P->grp_big_pmax = max_p;
  return P->grp_big_pmax;
}

int sawfem_mf_set_group_tri(void* plan, int rc_max, int lds_bytes) {
  Plan* P = (Plan*)plan;
  if (!P || rc_max < 1 || lds_bytes < 0) { sawfem_set_error("sawfem_mf_set_group_tri: invalid arguments"); return -1; }
  P->group_rc_tri = std::min(rc_max, 16);
  P->grp_lds_bytes = std::min(lds_bytes, P->lds_max);
  return P->group_rc_tri;
}

// Per-kernel event timing: 1 every launch, 2 the grouped solve kernel only (factorizations may run
// as graphs), 0 off; turning it on clears what was recorded.
int sawfem_mf_profile(void* plan, int enable) {
  Plan* P = (Plan*)plan;
  if (!P || enable < 0 || enable > 2) { sawfem_set_error("sawfem_mf_profile: invalid arguments"); return 1; }
  P->profile = enable;
  P->ev_used = 0;
  P->ev_marks.clear();
  P->grp_marks.clear();
  return 0;
}

// The grouped block-solve kernel (k_grp_lds on whole fronts) under profiling: summed HIP-event
// time (ms), algorithmic bytes and launches since profiling was turned on; waits for the events.
int sawfem_mf_profile_grp(void* plan, double* ms, double* bytes, int64_t* launches) {
  Plan* P = (Plan*)plan;
  if (!P || !ms || !bytes || !launches) { sawfem_set_error("sawfem_mf_profile_grp: invalid arguments"); return 1; }
  *ms = 0;
  *bytes = 0;
  *launches = 0;
  for (auto& g : P->grp_marks) {
    SAW_HIP(hipEventSynchronize(P->ev_pool[g.i0 + 1]));
    float t = 0;
    SAW_HIP(hipEventElapsedTime(&t, P->ev_pool[g.i0], P->ev_pool[g.i0 + 1]));
    *ms += t;
    *bytes += g.bytes;
    *launches += 1;
  }
  P->grp_marks.clear();
  return 0;
}

// Waits for the recorded launches and writes, per kind (init, entries, diag, trsm, gemm-L,
// gemm-S, fwd, bwd, small), total milliseconds to ms[kind] and launch counts to count[kind]
// (arrays of 9).
int sawfem_mf_profile_read(void* plan, double* ms, int64_t* count) {
  Plan* P = (Plan*)plan;
  if (!P || !ms || !count) { sawfem_set_error("sawfem_mf_profile_read: invalid arguments"); return 1; }
  for (int k = 0; k < K_NKIND; ++k) { ms[k] = 0; count[k] = 0; }
  P->level_ms.assign(P->levels.size() * K_NKIND, 0.0);
  for (auto& mk : P->ev_marks) {
    SAW_HIP(hipEventSynchronize(P->ev_pool[mk.i0 + 1]));
    float t = 0;
    SAW_HIP(hipEventElapsedTime(&t, P->ev_pool[mk.i0], P->ev_pool[mk.i0 + 1]));
    ms[mk.kind] += t;
    count[mk.kind] += 1;
    if (mk.level >= 0 && mk.level < (int)P->levels.size()) P->level_ms[(size_t)mk.level * K_NKIND + mk.kind] += t;
  }
  P->ev_used = 0;
  P->ev_marks.clear();
  P->grp_marks.clear();   // their events are released with the pool
  return 0;
}

// Per-level totals of the last sawfem_mf_profile_read: ms[level * 9 + kind] for every level
// (leaves first).  Returns the number of levels written (at most max_levels), or -1.
int sawfem_mf_profile_levels(void* plan, double* ms, int max_levels) {
  Plan* P = (Plan*)plan;
  if (!P || !ms || max_levels < 0) { sawfem_set_error("sawfem_mf_profile_levels: invalid arguments"); return -1; }
  const int nl = std::min(max_levels, (int)P->levels.size());
  for (int l = 0; l < nl; ++l)
    for (int k = 0; k < K_NKIND; ++k)
      ms[l * K_NKIND + k] = P->level_ms.empty() ? 0.0 : P->level_ms[(size_t)l * K_NKIND + k];
  return nl;
}

int sawfem_rhs(void* stream, int n, int B, const double* omega, double beta, double V, const void* rK,
               const void* rM, const void* rD, void* out) {
  if (n <= 0 || B <= 0) { sawfem_set_error("sawfem_rhs: invalid arguments"); return 1; }
  hipLaunchKernelGGL(k_rhs, dim3((n + 255) / 256, B), dim3(256), 0, (hipStream_t)stream, n, B, omega, beta, V,
                     (const double2*)rK, (const double2*)rM, (const double2*)rD, (double2*)out);
  SAW_HIP(hipGetLastError());
  return 0;
}

int sawfem_residual(void* stream, int n, int B, const int64_t* indptr, const int* indices, const int64_t* tslot,
                    const void* Kaa, const void* Maa, const void* Daa, const double* omega, double beta,
                    const void* x, const void* rhs, void* r) {
  if (n <= 0 || B <= 0) { sawfem_set_error("sawfem_residual: invalid arguments"); return 1; }
  hipLaunchKernelGGL(k_residual, dim3((n + 255) / 256, (B + 7) / 8), dim3(256), 0, (hipStream_t)stream, n, B, indptr,
                     indices, tslot, (const double2*)Kaa, (const double2*)Maa, (const double2*)Daa, omega, beta,
                     (const double2*)x, (const double2*)rhs, (double2*)r);
  SAW_HIP(hipGetLastError());
  return 0;
}

// Residual in the frequency-contiguous layout: xt/rhst/rt are n x Bp (Bp >= B, multiple of 64).
int sawfem_residual_t(void* stream, int n, int B, int Bp, const int64_t* indptr, const int* indices,
                      const int64_t* tslot, const void* Kaa, const void* Maa, const void* Daa, const double* omega,
                      double beta, const void* xt, const void* rhst, void* rt) {
  if (n <= 0 || B <= 0 || Bp < B || Bp % 64) { sawfem_set_error("sawfem_residual_t: invalid arguments"); return 1; }
  hipLaunchKernelGGL(k_residual_t, dim3((n + 3) / 4, Bp / 64), dim3(256), 0, (hipStream_t)stream, n, B, Bp, indptr,
                     indices, tslot, (const double2*)Kaa, (const double2*)Maa, (const double2*)Daa, omega, beta,
                     (const double2*)xt, (const double2*)rhst, (double2*)rt);
  SAW_HIP(hipGetLastError());
  return 0;
}

int sawfem_residual_sv(void* stream, int ng, int B, int Bp, const int* cmap, const int* gptr, const int* grows,
                       const int64_t* indptr, const int* indices, const int64_t* tslot, const void* Kaa,
                       const void* Maa, const void* Daa, const double* omega, double beta, const void* xt,
                       const void* rhst, void* rt) {
  if (ng <= 0 || B <= 0 || Bp < 1 || (!cmap && Bp < B) || !gptr || !grows) {
    sawfem_set_error("sawfem_residual_sv: invalid arguments");
    return 1;
  }
  // tiles of 64 columns per wave: all of them up to 8, else columns of 8 (the last one shorter)
  const int nt = (B + 63) / 64;
  const int ft = nt >= 8 ? 8 : (nt >= 4 ? 4 : (nt >= 2 ? 2 : 1));
  const dim3 grid((ng + 3) / 4, (nt + ft - 1) / ft);
#define SAW_RES(F)                                                                                                 \
  hipLaunchKernelGGL(k_residual_sv<F>, grid, dim3(256), 0, (hipStream_t)stream, ng, B, Bp, cmap, gptr, grows,      \
                     indptr, indices, tslot, (const double2*)Kaa, (const double2*)Maa, (const double2*)Daa, omega,  \
                     beta, (const double2*)xt, (const double2*)rhst, (double2*)rt)
  switch (ft) { case 8: SAW_RES(8); break; case 4: SAW_RES(4); break; case 2: SAW_RES(2); break; default: SAW_RES(1); }
#undef SAW_RES
  SAW_HIP(hipGetLastError());
  return 0;
}

// out (C x ld_out) = transpose of in (R x ld_in), complex128.
int sawfem_transpose(void* stream, int R, int C, int ld_in, int ld_out, const void* in, void* out) {
  if (R < 0 || C < 0 || ld_in < C || ld_out < R) { sawfem_set_error("sawfem_transpose: invalid arguments"); return 1; }
  if (R == 0 || C == 0) return 0;
  hipLaunchKernelGGL(k_transpose, dim3((C + 31) / 32, (R + 31) / 32), dim3(256), 0, (hipStream_t)stream, R, C, ld_in,
                     ld_out, (const double2*)in, (double2*)out);
  SAW_HIP(hipGetLastError());
  return 0;
}

int sawfem_transpose_map(void* stream, int R, int C, int ld_in, int ld_out, const void* in, void* out, const int* imap,
                         const int* omap, const void* add) {
  if (R < 0 || C < 0 || (!imap && ld_in < C) || (!omap && ld_out < R) || !in || !out) {
    sawfem_set_error("sawfem_transpose_map: invalid arguments");
    return 1;
  }
  if (R == 0 || C == 0) return 0;
  hipLaunchKernelGGL(k_transpose_map, dim3((C + 31) / 32, (R + 31) / 32), dim3(256), 0, (hipStream_t)stream, R, C,
                     ld_in, ld_out, (const double2*)in, (double2*)out, imap, omap, (const double2*)add);
  SAW_HIP(hipGetLastError());
  return 0;
}

int sawfem_charge(void* stream, int n, int B, const void* kts_a, double kz_re, double kz_im, double V, const void* x,
                  void* Q) {
  if (n <= 0 || B <= 0) { sawfem_set_error("sawfem_charge: invalid arguments"); return 1; }
  hipLaunchKernelGGL(k_charge, dim3(B), dim3(CHARGE_NT), 0, (hipStream_t)stream, n, (const double2*)kts_a,
                     make_double2(kz_re, kz_im), V, (const double2*)x, (double2*)Q);
  SAW_HIP(hipGetLastError());
  return 0;
}

int sawfem_wdot(void* stream, int n, int nb, const void* a, int64_t sa, const double* w, const void* c, int64_t sc,
                const int* sel, void* out) {
  if (n <= 0 || nb < 0 || !a || !c || !out || sa < 0 || sc < 0) {
    sawfem_set_error("sawfem_wdot: invalid arguments");
    return 1;
  }
  if (nb == 0) return 0;
  hipLaunchKernelGGL(k_wdot, dim3(nb), dim3(WDOT_NT), 0, (hipStream_t)stream, n, (const double2*)a, sa, w,
                     (const double2*)c, sc, sel, (double2*)out);
  SAW_HIP(hipGetLastError());
  return 0;
}

int sawfem_charge_rows(void* stream, int n, int nc, int64_t ld, const void* kts_a, double kz_re, double kz_im, double V,
                       const void* x, const int* cols, void* work, void* Q) {
  if (n <= 0 || nc <= 0 || ld < 1 || (!cols && ld < nc) || !kts_a || !x || !work || !Q) {
    sawfem_set_error("sawfem_charge_rows: invalid arguments");
    return 1;
  }
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_rows_partial<true>, dim3((nc + 63) / 64, 64), dim3(1024), 0, s, n, nc, ld,
                     (const double2*)kts_a, nullptr, nullptr, (const double2*)x, cols, (double2*)work);
  hipLaunchKernelGGL(k_rows_reduce<true>, dim3(nc), dim3(1024), 0, s, nc, (const double2*)work,
                     make_double2(kz_re, kz_im), V, (double2*)Q);
  SAW_HIP(hipGetLastError());
  return 0;
}

int sawfem_wdot_rows(void* stream, int n, int nc, int64_t ld, const void* a, const int* acol, const double* w,
                     const void* c, const int* ccol, void* work, void* out) {
  if (n <= 0 || nc <= 0 || ld < 1 || ((!acol || !ccol) && ld < nc) || !a || !c || !work || !out) {
    sawfem_set_error("sawfem_wdot_rows: invalid arguments");
    return 1;
  }
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_rows_partial<false>, dim3((nc + 63) / 64, 64), dim3(1024), 0, s, n, nc, ld,
                     (const double2*)a, acol, w, (const double2*)c, ccol, (double2*)work);
  hipLaunchKernelGGL(k_rows_reduce<false>, dim3(nc), dim3(1024), 0, s, nc, (const double2*)work,
                     make_double2(0, 0), 0.0, (double2*)out);
  SAW_HIP(hipGetLastError());
  return 0;
}

int sawfem_wdot_pairs_rows(void* stream, int n, int nb, int64_t ld, const void* X, const double* w, const void* C,
                           void* work, void* out1, void* out2) {
  if (n <= 0 || nb <= 0 || ld < 2 * (int64_t)nb || !X || !C || !work || !out1 || !out2) {
    sawfem_set_error("sawfem_wdot_pairs_rows: invalid arguments");
    return 1;
  }
  hipStream_t s = (hipStream_t)stream;
  double2* part = (double2*)work;
  hipLaunchKernelGGL(k_rows_partial_pair, dim3((nb + 63) / 64, 64), dim3(1024), 0, s, n, nb, ld, (const double2*)X, w,
                     (const double2*)C, part);
  hipLaunchKernelGGL(k_rows_reduce<false>, dim3(nb), dim3(1024), 0, s, nb, part, make_double2(0, 0), 0.0,
                     (double2*)out1);
  hipLaunchKernelGGL(k_rows_reduce<false>, dim3(nb), dim3(1024), 0, s, nb, part + (int64_t)1024 * nb,
                     make_double2(0, 0), 0.0, (double2*)out2);
  SAW_HIP(hipGetLastError());
  return 0;
}

int sawfem_rows_update_measure(void* stream, int n, int nb, int64_t ld, void* X, const void* R, const void* kts_a,
                               double kz_re, double kz_im, double V, const double* w, const void* rhs, void* work,
                               void* Q, void* mu) {
  if (n <= 0 || nb <= 0 || ld < 2 * (int64_t)nb || !X || !R || !kts_a || !w || !rhs || !work || !Q || !mu) {
    sawfem_set_error("sawfem_rows_update_measure: invalid arguments");
    return 1;
  }
  hipStream_t s = (hipStream_t)stream;
  double2* pq = (double2*)work;
  double2* pm = pq + (int64_t)1024 * 2 * nb;
  hipLaunchKernelGGL(k_rows_update_measure, dim3((nb + 63) / 64, 64), dim3(1024), 0, s, n, nb, ld, (double2*)X,
                     (const double2*)R, (const double2*)kts_a, w, (const double2*)rhs, pq, pm);
  hipLaunchKernelGGL(k_rows_reduce<true>, dim3(2 * nb), dim3(1024), 0, s, 2 * nb, pq, make_double2(kz_re, kz_im), V,
                     (double2*)Q);
  hipLaunchKernelGGL(k_rows_reduce<false>, dim3(nb), dim3(1024), 0, s, nb, pm, make_double2(0, 0), 0.0,
                     (double2*)mu);
  SAW_HIP(hipGetLastError());
  return 0;
}

int sawfem_axpy(void* stream, int64_t n, void* x, const void* d) {
  if (n < 0) { sawfem_set_error("sawfem_axpy: invalid arguments"); return 1; }
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_axpy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, n, (double2*)x,
                     (const double2*)d);
  SAW_HIP(hipGetLastError());
  return 0;
}

}  // extern "C"
