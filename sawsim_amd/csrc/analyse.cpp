// Host symbolic analysis of the Bloch-reduced active system and the multifrontal plan arrays.
//
// Plain C++ (no GPU call), built with floating-point contraction off so that every cut
// coordinate is computed exactly as written.  One call replaces the analysis phase that the
// reference leaves to PARDISO / SuperLU at every frequency (sweep.py:326-333, 349-353): it is
// done once per cell and shared by all frequencies.
//
// 1. Supervariables: the active unknowns of one merged mesh node (or one column of a slab)
//    share their sparsity pattern; the ordering works on the supervariable graph.
// 2. Ordering: geometric nested dissection on supervariable coordinates, periodic in x.  A
//    sub-domain is split at the weighted median of y, or of x (a sub-domain that wraps around
//    the periodic cell is cut at x mod P = P/2), the vertex separator is the lighter of the two
//    one-sided boundary layers, and the cut with the lower separator weight per balance wins.
//    Each separator and each leaf sub-domain becomes one front; fronts are numbered in post
//    order, which is the elimination order.
// 3. Fronts: a front's rows are its pivots, the pattern neighbours of its pivots and its
//    children's contribution rows, restricted to unknowns eliminated later, in elimination
//    order; supervariables are expanded to scalar unknowns (component order within one).
// 4. Plan arrays: the child-row -> parent-row maps, the level lists, and every A_aa entry's
//    (owner front, column-major offset in that front) sorted by level, front and offset.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <new>
#include <numeric>
#include <vector>

#include "../../include/sawfem.h"

const char* sawfem_set_error(const char* msg);

namespace {

using i64 = int64_t;
using std::vector;

int fail(int code, const char* msg) {
  sawfem_set_error(msg);
  return code;
}

// numpy.remainder for a positive divisor: fmod, shifted into [0, P), +0.0 for an exact zero
double py_mod(double a, double P) {
  double r = std::fmod(a, P);
  if (r != 0.0) {
    if (r < 0.0) r += P;
  } else {
    r = 0.0;
  }
  return r;
}

// an array of a host result: data, element count, element size in bytes
struct Item {
  const void* data;
  int64_t count;
  int elem;
};
template <typename T>
Item of(const vector<T>& v) {
  return {v.data(), (int64_t)v.size(), (int)sizeof(T)};
}

// a host computation's output arrays, read through sawfem_host_size / sawfem_host_copy
struct HostResult {
  virtual ~HostResult() = default;
  virtual Item item(int what) const = 0;
};

struct Analysis : HostResult {
  // supervariable graph
  i64 n = 0, nsv = 0;
  vector<i64> sv_ids, sv_of, gptr, gidx, w;
  vector<double> x, y;
  double period = 0.0;
  bool periodic = false;
  i64 leaf = 0;
  // dissection state
  vector<i64> label, stamp;
  i64 tag = 0, stamp_id = 0;
  // fronts (post order)
  vector<vector<i64>> piv;
  vector<vector<int>> kids;
  vector<int> parent;
  // results
  vector<int> m, p, level, child_ptr, child_ids, lvl_ptr, lvl_fronts, rows, inv, ent_front, ent_loc;
  vector<i64> rows_ptr, inv_ptr, ent_slot, ent_lvl_ptr, perm, piv_ptr, piv_sv, rows_sv_ptr, rows_sv;

  i64 wsum(const vector<i64>& v) const {
    i64 s = 0;
    for (i64 a : v) s += w[a];
    return s;
  }

  double weighted_median(const vector<double>& key, const vector<i64>& verts) const {
    const size_t k = verts.size();
    vector<i64> o(k);
    std::iota(o.begin(), o.end(), 0);
    std::stable_sort(o.begin(), o.end(), [&](i64 a, i64 b) { return key[a] < key[b]; });
    vector<i64> cw(k);
    i64 c = 0;
    for (size_t i = 0; i < k; ++i) cw[i] = c += w[verts[o[i]]];
    const double half = (double)cw[k - 1] / 2.0;
    size_t j = std::lower_bound(cw.begin(), cw.end(), half, [](i64 v, double h) { return (double)v < h; }) -
               cw.begin();
    return key[o[std::min(j, k - 1)]];
  }

  // the cut candidates of a sub-domain, in order: y at its weighted median, then x
  vector<vector<char>> candidates(const vector<i64>& verts) const {
    const size_t k = verts.size();
    vector<double> xv(k), yv(k);
    for (size_t i = 0; i < k; ++i) xv[i] = x[verts[i]], yv[i] = y[verts[i]];
    vector<vector<char>> out;
    vector<char> mk(k);
    const double ym = weighted_median(yv, verts);
    for (size_t i = 0; i < k; ++i) mk[i] = yv[i] < ym;
    out.push_back(mk);
    if (periodic) {
      const double P = period;
      vector<double> r(k);
      for (size_t i = 0; i < k; ++i) r[i] = std::nearbyint(xv[i] / P * 1e9);
      std::sort(r.begin(), r.end());
      r.erase(std::unique(r.begin(), r.end()), r.end());
      const size_t nx = r.size();
      vector<double> xs(nx);
      for (size_t i = 0; i < nx; ++i) xs[i] = r[i] / 1e9 * P;
      std::sort(xs.begin(), xs.end());
      vector<double> gaps(nx);
      for (size_t i = 0; i + 1 < nx; ++i) gaps[i] = xs[i + 1] - xs[i];
      gaps[nx - 1] = (xs[0] + P) - xs[nx - 1];
      const size_t g = std::max_element(gaps.begin(), gaps.end()) - gaps.begin();
      const bool wraps = nx > 1 ? gaps[g] < 0.26 * P : false;
      vector<double> xu(k);
      if (wraps) {
        const double half = P / 2 - 1e-15 * P;
        for (size_t i = 0; i < k; ++i) {
          xu[i] = py_mod(xv[i], P);
          mk[i] = xu[i] < half;
        }
      } else {
        const double start = xs[(g + 1) % nx];
        const double eps = 1e-12 * P;
        for (size_t i = 0; i < k; ++i) {
          const double d = xv[i] - start;
          xu[i] = py_mod(d + eps, P);
        }
        const double xm = weighted_median(xu, verts);
        for (size_t i = 0; i < k; ++i) mk[i] = xu[i] < xm;
      }
    } else {
      const double xm = weighted_median(xv, verts);
      for (size_t i = 0; i < k; ++i) mk[i] = xv[i] < xm;
    }
    out.push_back(mk);
    return out;
  }

  // the lighter one-sided boundary layer between the two sides of a cut (ascending ids)
  vector<i64> separator(const vector<i64>& verts, const vector<char>& side_a) {
    tag += 2;
    const i64 ta = tag, tb = tag + 1;
    for (size_t i = 0; i < verts.size(); ++i) label[verts[i]] = side_a[i] ? ta : tb;
    vector<i64> sa, sb;
    for (i64 v : verts) {
      const i64 lv = label[v];
      const i64 other = lv == ta ? tb : ta;
      for (i64 e = gptr[v]; e < gptr[v + 1]; ++e) {
        if (label[gidx[e]] == other) {
          (lv == ta ? sa : sb).push_back(v);
          break;
        }
      }
    }
    return wsum(sa) <= wsum(sb) ? sa : sb;
  }

  int make_front(vector<i64> pivots, vector<int> cids) {
    const int fid = (int)piv.size();
    piv.push_back(std::move(pivots));
    for (int c : cids) parent[c] = fid;
    kids.push_back(std::move(cids));
    parent.push_back(-1);
    return fid;
  }

  // dissect a sub-domain (ascending, non-empty) and number its fronts in post order;
  // returns the sub-tree's root front, or -1 for a sub-domain that yields none
  int run(const vector<i64>& verts) {
    const i64 W = wsum(verts);
    if (W <= leaf || verts.size() <= 2) return make_front(verts, {});
    bool have = false;
    double best = 0.0;
    vector<i64> bsep, bpa, bpb;
    for (const vector<char>& side_a : candidates(verts)) {
      size_t na = 0;
      for (char c : side_a) na += c != 0;
      if (na == 0 || na == side_a.size()) continue;
      vector<i64> sep = separator(verts, side_a);
      ++stamp_id;
      for (i64 v : sep) stamp[v] = stamp_id;
      vector<i64> pa, pb;
      for (size_t i = 0; i < verts.size(); ++i) {
        if (stamp[verts[i]] == stamp_id) continue;
        (side_a[i] ? pa : pb).push_back(verts[i]);
      }
      const double bal = (double)std::min(wsum(pa), wsum(pb)) / (double)std::max<i64>(1, W);
      const double score = (double)wsum(sep) / std::max(bal, 1e-3);
      if (!have || score < best) {
        have = true;
        best = score;
        bsep.swap(sep);
        bpa.swap(pa);
        bpb.swap(pb);
      }
    }
    if (!have) return make_front(verts, {});
    vector<int> cids;
    for (const vector<i64>* part : {&bpa, &bpb}) {
      if (part->empty()) continue;
      const int c = run(*part);
      if (c >= 0) cids.push_back(c);
    }
    if (bsep.empty()) {
      if (cids.empty()) return -1;
      if (cids.size() == 1) return cids[0];
    }
    return make_front(std::move(bsep), std::move(cids));
  }

  int build(i64 n_, const i64* indptr, const i64* indices, const i64* sv_in, const i64* comp_of, i64 nxy,
            const double* xy) {
    n = n_;
    // supervariables: sorted distinct ids of sv_in
    i64 vmax = -1;
    for (i64 i = 0; i < n; ++i) {
      if (sv_in[i] < 0 || sv_in[i] >= nxy) return fail(1, "sawfem_analyse: supervariable id outside the coordinates");
      vmax = std::max(vmax, sv_in[i]);
    }
    vector<i64> map(vmax + 1, -1);
    for (i64 i = 0; i < n; ++i) map[sv_in[i]] = 1;
    for (i64 v = 0; v <= vmax; ++v)
      if (map[v] >= 0) map[v] = (i64)sv_ids.size(), sv_ids.push_back(v);
    nsv = (i64)sv_ids.size();
    sv_of.resize(n);
    for (i64 i = 0; i < n; ++i) sv_of[i] = map[sv_in[i]];
    w.assign(nsv, 0);
    for (i64 i = 0; i < n; ++i) ++w[sv_of[i]];
    x.resize(nsv);
    y.resize(nsv);
    for (i64 s = 0; s < nsv; ++s) x[s] = xy[2 * sv_ids[s]], y[s] = xy[2 * sv_ids[s] + 1];
    // supervariable adjacency: distinct neighbours in ascending order, no self loops
    if (indptr[0] != 0) return fail(1, "sawfem_analyse: indptr[0] must be 0");
    for (i64 r = 0; r < n; ++r) {
      if (indptr[r + 1] < indptr[r]) return fail(1, "sawfem_analyse: indptr not ascending");
      for (i64 e = indptr[r]; e < indptr[r + 1]; ++e)
        if (indices[e] < 0 || indices[e] >= n) return fail(1, "sawfem_analyse: column index out of range");
    }
    vector<i64> sv_rows_ptr(nsv + 1, 0), sv_rows(n);
    for (i64 i = 0; i < n; ++i) ++sv_rows_ptr[sv_of[i] + 1];
    for (i64 s = 0; s < nsv; ++s) sv_rows_ptr[s + 1] += sv_rows_ptr[s];
    {
      vector<i64> fill(sv_rows_ptr.begin(), sv_rows_ptr.end() - 1);
      for (i64 i = 0; i < n; ++i) sv_rows[fill[sv_of[i]]++] = i;
    }
    vector<i64> seen(nsv, -1);
    gptr.assign(nsv + 1, 0);
    for (i64 a = 0; a < nsv; ++a) {
      const size_t g0 = gidx.size();
      for (i64 j = sv_rows_ptr[a]; j < sv_rows_ptr[a + 1]; ++j) {
        const i64 r = sv_rows[j];
        for (i64 e = indptr[r]; e < indptr[r + 1]; ++e) {
          const i64 b = sv_of[indices[e]];
          if (b != a && seen[b] != a) seen[b] = a, gidx.push_back(b);
        }
      }
      std::sort(gidx.begin() + g0, gidx.end());
      gptr[a + 1] = (i64)gidx.size();
    }
    // dissection
    label.assign(nsv, -1);
    stamp.assign(std::max<i64>(nsv, n), 0);
    vector<i64> all(nsv);
    std::iota(all.begin(), all.end(), 0);
    if (nsv == 0) return fail(1, "sawfem_analyse: empty system");
    run(all);
    const int nf = (int)piv.size();
    // elimination order of supervariables
    vector<i64> pos_sv(nsv), start(nf + 1, 0);
    piv_ptr.assign(nf + 1, 0);
    for (int f = 0; f < nf; ++f) {
      for (i64 s : piv[f]) pos_sv[s] = (i64)piv_sv.size(), piv_sv.push_back(s);
      piv_ptr[f + 1] = start[f + 1] = (i64)piv_sv.size();
    }
    if ((i64)piv_sv.size() != nsv) return fail(1, "sawfem_analyse: dissection lost supervariables");
    // active unknowns in elimination order: by supervariable position, then component, then id
    vector<i64> bucket(nsv + 1, 0);
    for (i64 i = 0; i < n; ++i) ++bucket[pos_sv[sv_of[i]] + 1];
    for (i64 s = 0; s < nsv; ++s) bucket[s + 1] += bucket[s];
    const vector<i64> sv_dof_ptr = bucket;     // start of each position's unknowns
    perm.assign(n, 0);
    for (i64 i = 0; i < n; ++i) perm[bucket[pos_sv[sv_of[i]]]++] = i;
    for (i64 s = 0; s < nsv; ++s)
      std::stable_sort(perm.begin() + sv_dof_ptr[s], perm.begin() + sv_dof_ptr[s + 1],
                       [&](i64 a, i64 b) { return comp_of[a] < comp_of[b]; });
    // row structure, bottom-up (children precede their parent)
    rows_sv_ptr.assign(nf + 1, 0);
    vector<i64> cand;
    for (int f = 0; f < nf; ++f) {
      ++stamp_id;
      cand.clear();
      auto add = [&](i64 s) {
        if (stamp[s] != stamp_id && pos_sv[s] >= start[f]) stamp[s] = stamp_id, cand.push_back(s);
      };
      for (i64 s : piv[f]) {
        add(s);
        for (i64 e = gptr[s]; e < gptr[s + 1]; ++e) add(gidx[e]);
      }
      for (int c : kids[f]) {
        const i64 c0 = rows_sv_ptr[c] + (i64)piv[c].size();
        for (i64 j = c0; j < rows_sv_ptr[c + 1]; ++j) add(rows_sv[j]);
      }
      std::sort(cand.begin(), cand.end(), [&](i64 a, i64 b) { return pos_sv[a] < pos_sv[b]; });
      rows_sv.insert(rows_sv.end(), cand.begin(), cand.end());
      rows_sv_ptr[f + 1] = (i64)rows_sv.size();
    }
    // scalar rows
    m.assign(nf, 0);
    p.assign(nf, 0);
    rows_ptr.assign(nf + 1, 0);
    for (int f = 0; f < nf; ++f) {
      i64 mm = 0, pp = 0;
      for (i64 j = rows_sv_ptr[f]; j < rows_sv_ptr[f + 1]; ++j) {
        const i64 ps = pos_sv[rows_sv[j]];
        for (i64 k = sv_dof_ptr[ps]; k < sv_dof_ptr[ps + 1]; ++k) {
          if (rows.size() >= (size_t)INT32_MAX) return fail(2, "sawfem_analyse: front rows exceed int32");
          rows.push_back((int)perm[k]);
        }
        mm += sv_dof_ptr[ps + 1] - sv_dof_ptr[ps];
        if (j - rows_sv_ptr[f] < (i64)piv[f].size()) pp = mm;
      }
      // an entry's offset li + lj*m is stored as int32
      if (mm > 46340) return fail(2, "sawfem_analyse: front order above 46340");
      m[f] = (int)mm;
      p[f] = (int)pp;
      rows_ptr[f + 1] = (i64)rows.size();
    }
    // levels (leaves 0), child lists
    level.assign(nf, 0);
    child_ptr.assign(nf + 1, 0);
    int nlev = 0;
    for (int f = 0; f < nf; ++f) {
      int L = 0;
      for (int c : kids[f]) L = std::max(L, level[c] + 1);
      level[f] = L;
      nlev = std::max(nlev, L + 1);
      child_ids.insert(child_ids.end(), kids[f].begin(), kids[f].end());
      child_ptr[f + 1] = (int)child_ids.size();
    }
    lvl_ptr.assign(nlev + 1, 0);
    for (int f = 0; f < nf; ++f) ++lvl_ptr[level[f] + 1];
    for (int L = 0; L < nlev; ++L) lvl_ptr[L + 1] += lvl_ptr[L];
    lvl_fronts.assign(nf, 0);
    {
      vector<int> fill_l(lvl_ptr.begin(), lvl_ptr.end() - 1);
      for (int f = 0; f < nf; ++f) lvl_fronts[fill_l[level[f]]++] = f;
    }
    // parent row -> child contribution index maps
    vector<int> scratch(n, -1);
    inv_ptr.assign(1, 0);
    for (int f = 0; f < nf; ++f) {
      for (int c : kids[f]) {
        const i64 c0 = rows_ptr[c] + p[c], c1 = rows_ptr[c + 1];
        for (i64 j = c0; j < c1; ++j) scratch[rows[j]] = (int)(j - c0);
        i64 found = 0;
        for (i64 j = rows_ptr[f]; j < rows_ptr[f + 1]; ++j) {
          const int v = scratch[rows[j]];
          found += v >= 0;
          inv.push_back(v);
        }
        if (found != c1 - c0) return fail(1, "sawfem_analyse: child contribution rows missing from parent front");
        for (i64 j = c0; j < c1; ++j) scratch[rows[j]] = -1;
        inv_ptr.push_back((i64)inv.size());
      }
    }
    // A_aa entries: owner front (the front of the earlier-eliminated of row and column)
    vector<int> owner(n, -1);
    for (int f = 0; f < nf; ++f)
      for (i64 j = rows_ptr[f]; j < rows_ptr[f] + p[f]; ++j) owner[rows[j]] = f;
    vector<i64> rank(n);
    for (i64 i = 0; i < n; ++i) rank[perm[i]] = i;
    const i64 nnz = indptr[n];
    vector<int> fpos(nf);
    for (int i = 0; i < nf; ++i) fpos[lvl_fronts[i]] = i;
    vector<i64> fcnt(nf + 1, 0);
    vector<int> own(nnz);
    for (i64 r = 0; r < n; ++r)
      for (i64 e = indptr[r]; e < indptr[r + 1]; ++e) {
        const i64 c = indices[e];
        own[e] = rank[r] <= rank[c] ? owner[r] : owner[c];
        ++fcnt[fpos[own[e]] + 1];
      }
    for (int i = 0; i < nf; ++i) fcnt[i + 1] += fcnt[i];
    vector<i64> ent(nnz), ent_row(nnz);
    {
      vector<i64> fill_e(fcnt.begin(), fcnt.end() - 1);
      for (i64 r = 0; r < n; ++r)
        for (i64 e = indptr[r]; e < indptr[r + 1]; ++e) {
          const i64 k = fill_e[fpos[own[e]]]++;
          ent[k] = e;
          ent_row[k] = r;
        }
    }
    own.clear();
    own.shrink_to_fit();
    ent_front.assign(nnz, 0);
    ent_loc.assign(nnz, 0);
    ent_slot.assign(nnz, 0);
    // within a front, entries in column-major order: counting sort by row, then (stably) by
    // column; (row, column) pairs are distinct, so this is the order of the offsets
    vector<int> li(nnz ? 1 : 0), lj, cnt;
    vector<i64> tmp;
    for (int i = 0; i < nf; ++i) {
      const int f = lvl_fronts[i];
      const i64 k0 = fcnt[i], E = fcnt[i + 1] - k0;
      const int mf = m[f];
      for (i64 j = rows_ptr[f]; j < rows_ptr[f + 1]; ++j) scratch[rows[j]] = (int)(j - rows_ptr[f]);
      li.resize(E);
      lj.resize(E);
      tmp.resize(E);
      for (i64 t = 0; t < E; ++t) {
        li[t] = scratch[ent_row[k0 + t]];
        lj[t] = scratch[indices[ent[k0 + t]]];
        if (li[t] < 0 || lj[t] < 0) return fail(1, "sawfem_analyse: A_aa entry outside its owner front");
      }
      // pass 1: by row into tmp (positions t)
      cnt.assign(mf + 1, 0);
      for (i64 t = 0; t < E; ++t) ++cnt[li[t] + 1];
      for (int r = 0; r < mf; ++r) cnt[r + 1] += cnt[r];
      for (i64 t = 0; t < E; ++t) tmp[cnt[li[t]]++] = t;
      // pass 2: by column, stable
      cnt.assign(mf + 1, 0);
      for (i64 t = 0; t < E; ++t) ++cnt[lj[t] + 1];
      for (int c = 0; c < mf; ++c) cnt[c + 1] += cnt[c];
      for (i64 u = 0; u < E; ++u) {
        const i64 t = tmp[u];
        const i64 k = k0 + cnt[lj[t]]++;
        ent_front[k] = f;
        ent_loc[k] = li[t] + lj[t] * mf;
        ent_slot[k] = ent[k0 + t];
      }
      for (i64 j = rows_ptr[f]; j < rows_ptr[f + 1]; ++j) scratch[rows[j]] = -1;
    }
    ent_lvl_ptr.assign(nlev + 1, 0);
    for (int L = 0; L <= nlev; ++L) ent_lvl_ptr[L] = fcnt[lvl_ptr[L]];
    return 0;
  }

  Item item(int what) const override {
    switch (what) {
      case SAWFEM_SYM_M: return of(m);
      case SAWFEM_SYM_P: return of(p);
      case SAWFEM_SYM_ROWS_PTR: return of(rows_ptr);
      case SAWFEM_SYM_ROWS: return of(rows);
      case SAWFEM_SYM_CHILD_PTR: return of(child_ptr);
      case SAWFEM_SYM_CHILD_IDS: return of(child_ids);
      case SAWFEM_SYM_PARENT: return of(parent);
      case SAWFEM_SYM_LEVEL: return of(level);
      case SAWFEM_SYM_INV_PTR: return of(inv_ptr);
      case SAWFEM_SYM_INV: return of(inv);
      case SAWFEM_SYM_LVL_PTR: return of(lvl_ptr);
      case SAWFEM_SYM_LVL_FRONTS: return of(lvl_fronts);
      case SAWFEM_SYM_ENT_FRONT: return of(ent_front);
      case SAWFEM_SYM_ENT_SLOT: return of(ent_slot);
      case SAWFEM_SYM_ENT_LOC: return of(ent_loc);
      case SAWFEM_SYM_ENT_LVL_PTR: return of(ent_lvl_ptr);
      case SAWFEM_SYM_PERM: return of(perm);
      case SAWFEM_SYM_PIV_SV_PTR: return of(piv_ptr);
      case SAWFEM_SYM_PIV_SV: return of(piv_sv);
      case SAWFEM_SYM_ROWS_SV_PTR: return of(rows_sv_ptr);
      case SAWFEM_SYM_ROWS_SV: return of(rows_sv);
      case SAWFEM_SYM_SV_IDS: return of(sv_ids);
      default: return {nullptr, -1, 0};
    }
  }
};

// Bloch reduction of the structural K pattern to the active system (sweep.py:283-314): the
// A_aa pattern of (T^T K T)[active, active] and three gather lists over K slots (see
// sawfem_bloch_layout), every list in ascending slot order within its target.
struct Bloch : HostResult {
  vector<i64> aa_indptr, aa_indices, aa_ptr, aa_idx, rhs_ptr, rhs_idx, kts_ptr, kts_idx;

  int build(int nd, i64 N, bool interleaved, const i64* node_ptr, const i64* node_idx, const i64* rep,
            const i64* pos, i64 n, const unsigned char* is_z, const unsigned char* is_f) {
    const i64 gdof = nd * N;
    for (i64 d = 0; d < gdof; ++d)
      if (rep[d] < 0 || rep[d] >= gdof || pos[d] < -1 || pos[d] >= n)
        return fail(1, "sawfem_bloch_layout: rep or pos out of range");
    for (i64 i = 0; i < N; ++i)
      if (node_ptr[i + 1] < node_ptr[i]) return fail(1, "sawfem_bloch_layout: node_ptr not ascending");
    for (i64 k = 0; k < node_ptr[N]; ++k)
      if (node_idx[k] < 0 || node_idx[k] >= N) return fail(1, "sawfem_bloch_layout: node index out of range");
    const i64 nnz_node = node_ptr[N];
    // K row d = (component a, node i); its slots are one ascending range: block (a, b) of node
    // pair k at a*nd*nnz_node + nd*node_ptr[i] + b*deg[i] + (k - node_ptr[i])
    auto comp = [&](i64 d) -> int {
      if (!interleaved) return (int)(d / N);
      return d < 3 * N ? (int)(d % 3) : 3;
    };
    auto node = [&](i64 d) -> i64 {
      if (!interleaved) return d % N;
      return d < 3 * N ? d / 3 : d - 3 * N;
    };
    auto dof = [&](int a, i64 nd_) -> i64 {
      if (!interleaved) return a * N + nd_;
      return a < 3 ? 3 * nd_ + a : 3 * N + nd_;
    };
    auto start = [&](i64 d) { return comp(d) * nd * nnz_node + nd * node_ptr[node(d)]; };
    // calls f(slot, column dof) for the slots of K row d, ascending
    auto row_slots = [&](i64 d, auto&& f) {
      const i64 i = node(d), k0 = node_ptr[i], deg = node_ptr[i + 1] - k0, s0 = start(d);
      for (int b = 0; b < nd; ++b)
        for (i64 k = 0; k < deg; ++k) f(s0 + b * deg + k, dof(b, node_idx[k0 + k]));
    };
    // the K rows merged into each active row, by ascending slot range
    vector<i64> rptr(n + 1, 0), rrows;
    for (i64 d = 0; d < gdof; ++d)
      if (pos[rep[d]] >= 0) ++rptr[pos[rep[d]] + 1];
    for (i64 j = 0; j < n; ++j) rptr[j + 1] += rptr[j];
    rrows.resize(rptr[n]);
    {
      vector<i64> fill(rptr.begin(), rptr.end() - 1);
      for (i64 d = 0; d < gdof; ++d)
        if (pos[rep[d]] >= 0) rrows[fill[pos[rep[d]]]++] = d;
    }
    // merged active position and electrode flag of every column dof
    vector<i64> prep(gdof);
    vector<unsigned char> zrep(gdof);
    for (i64 d = 0; d < gdof; ++d) prep[d] = pos[rep[d]], zrep[d] = is_z[rep[d]];
    // per active row: its slots ascending with their merged columns; the distinct columns
    // ascending, each with its slots; the electrode-column slots
    vector<i64> mark(n, -1), rk(n, 0), ents, entc, cols, cnt;
    aa_indptr.assign(n + 1, 0);
    aa_ptr.assign(1, 0);
    rhs_ptr.assign(n + 1, 0);
    for (i64 pr = 0; pr < n; ++pr) {
      std::sort(rrows.begin() + rptr[pr], rrows.begin() + rptr[pr + 1],
                [&](i64 a, i64 b) { return start(a) < start(b); });
      ents.clear();
      entc.clear();
      cols.clear();
      for (i64 t = rptr[pr]; t < rptr[pr + 1]; ++t)
        row_slots(rrows[t], [&](i64 s, i64 c) {
          const i64 pc = prep[c];
          if (pc >= 0) {
            ents.push_back(s);
            entc.push_back(pc);
            if (mark[pc] != pr) mark[pc] = pr, cols.push_back(pc);
          }
          if (zrep[c]) rhs_idx.push_back(s);
        });
      rhs_ptr[pr + 1] = (i64)rhs_idx.size();
      std::sort(cols.begin(), cols.end());
      cnt.assign(cols.size() + 1, 0);
      for (size_t q = 0; q < cols.size(); ++q) rk[cols[q]] = (i64)q;
      for (i64 c : entc) ++cnt[rk[c] + 1];
      const i64 base = aa_ptr.back();
      for (size_t q = 0; q < cols.size(); ++q) {
        cnt[q + 1] += cnt[q];
        aa_ptr.push_back(base + cnt[q + 1]);
      }
      aa_idx.resize(base + (i64)ents.size());
      for (size_t t = 0; t < ents.size(); ++t) aa_idx[base + cnt[rk[entc[t]]]++] = ents[t];
      aa_indices.insert(aa_indices.end(), cols.begin(), cols.end());
      aa_indptr[pr + 1] = (i64)aa_indices.size();
    }
    // charge weights: slots of the (unmerged) electrode rows, by target then slot
    vector<std::pair<i64, i64>> kt;
    for (i64 d = 0; d < gdof; ++d) {
      if (!is_z[d]) continue;
      row_slots(d, [&](i64 s, i64 c) {
        const i64 pc = pos[rep[c]];
        if (pc >= 0 && !is_z[c] && !is_f[c]) kt.emplace_back(pc, s);
        else if (is_z[c]) kt.emplace_back(n, s);
      });
    }
    std::sort(kt.begin(), kt.end());
    kts_ptr.assign(n + 2, 0);
    kts_idx.resize(kt.size());
    for (size_t t = 0; t < kt.size(); ++t) ++kts_ptr[kt[t].first + 1], kts_idx[t] = kt[t].second;
    for (i64 j = 0; j <= n; ++j) kts_ptr[j + 1] += kts_ptr[j];
    return 0;
  }

  Item item(int what) const override {
    switch (what) {
      case SAWFEM_BLOCH_AA_INDPTR: return of(aa_indptr);
      case SAWFEM_BLOCH_AA_INDICES: return of(aa_indices);
      case SAWFEM_BLOCH_AA_PTR: return of(aa_ptr);
      case SAWFEM_BLOCH_AA_IDX: return of(aa_idx);
      case SAWFEM_BLOCH_RHS_PTR: return of(rhs_ptr);
      case SAWFEM_BLOCH_RHS_IDX: return of(rhs_idx);
      case SAWFEM_BLOCH_KTS_PTR: return of(kts_ptr);
      case SAWFEM_BLOCH_KTS_IDX: return of(kts_idx);
      default: return {nullptr, -1, 0};
    }
  }
};

}  // namespace

extern "C" {

int sawfem_analyse(void** sym, int64_t n, const int64_t* aa_indptr, const int64_t* aa_indices, const int64_t* sv_of,
                   const int64_t* comp_of, int64_t nxy, const double* xy, double period, int64_t leaf_weight) {
  if (!sym || n <= 0 || !aa_indptr || !aa_indices || !sv_of || !comp_of || !xy || nxy <= 0 || leaf_weight < 0)
    return fail(1, "sawfem_analyse: invalid argument");
  if (n > INT32_MAX) return fail(2, "sawfem_analyse: more than 2^31 unknowns");
  if (!(period >= 0.0) || std::isinf(period)) return fail(1, "sawfem_analyse: period must be finite and >= 0");
  *sym = nullptr;
  Analysis* a = new (std::nothrow) Analysis();
  if (!a) return fail(4, "sawfem_analyse: out of host memory");
  a->periodic = period > 0.0;
  a->period = period;
  a->leaf = leaf_weight;
  int rc;
  try {
    rc = a->build(n, aa_indptr, aa_indices, sv_of, comp_of, nxy, xy);
  } catch (const std::bad_alloc&) {
    rc = fail(4, "sawfem_analyse: out of host memory");
  }
  if (rc) {
    delete a;
    return rc;
  }
  *sym = a;
  return 0;
}

int sawfem_bloch_layout(void** out, int nd, int64_t N, int interleaved, const int64_t* node_ptr,
                        const int64_t* node_idx, const int64_t* rep, const int64_t* pos, int64_t n_active,
                        const unsigned char* is_zheng, const unsigned char* is_fu) {
  if (!out || nd < 1 || nd > 4 || N <= 0 || !node_ptr || !node_idx || !rep || !pos || n_active <= 0 || !is_zheng ||
      !is_fu || (interleaved && nd != 4))
    return fail(1, "sawfem_bloch_layout: invalid argument");
  *out = nullptr;
  Bloch* b = new (std::nothrow) Bloch();
  if (!b) return fail(4, "sawfem_bloch_layout: out of host memory");
  int rc;
  try {
    rc = b->build(nd, N, interleaved != 0, node_ptr, node_idx, rep, pos, n_active, is_zheng, is_fu);
  } catch (const std::bad_alloc&) {
    rc = fail(4, "sawfem_bloch_layout: out of host memory");
  }
  if (rc) {
    delete b;
    return rc;
  }
  *out = b;
  return 0;
}

int sawfem_transpose_slots(int64_t n, const int64_t* indptr, const int64_t* indices, int64_t* tslot) {
  if (n <= 0 || !indptr || !indices || !tslot) return fail(1, "sawfem_transpose_slots: invalid argument");
  // rows are visited in ascending order, so the visits to column c meet row c's entries in
  // their (ascending) column order: a cursor per row finds each transposed entry
  vector<i64> cur(indptr, indptr + n);
  for (i64 r = 0; r < n; ++r)
    for (i64 e = indptr[r]; e < indptr[r + 1]; ++e) {
      const i64 c = indices[e];
      if (c < 0 || c >= n) return fail(1, "sawfem_transpose_slots: column index out of range");
      const i64 q = cur[c]++;
      if (q >= indptr[c + 1] || indices[q] != r)
        return fail(1, "sawfem_transpose_slots: pattern is not structurally symmetric with ascending columns");
      tslot[e] = q;
    }
  return 0;
}

int64_t sawfem_host_size(const void* res, int what) {
  if (!res) return -1;
  return static_cast<const HostResult*>(res)->item(what).count;
}

int sawfem_host_copy(const void* res, int what, void* dst) {
  if (!res) return fail(1, "sawfem_host_copy: null result");
  const Item it = static_cast<const HostResult*>(res)->item(what);
  if (it.count < 0) return fail(1, "sawfem_host_copy: unknown array id");
  if (it.count && !dst) return fail(1, "sawfem_host_copy: null destination");
  if (it.count) std::copy_n(static_cast<const char*>(it.data), it.count * it.elem, static_cast<char*>(dst));
  return 0;
}

int sawfem_host_destroy(void* res) {
  delete static_cast<HostResult*>(res);
  return 0;
}

int sawfem_mf_create_sym(void** plan, const void* sym) {
  if (!sym) return fail(1, "sawfem_mf_create_sym: null analysis");
  const Analysis* ap = dynamic_cast<const Analysis*>(static_cast<const HostResult*>(sym));
  if (!ap) return fail(1, "sawfem_mf_create_sym: not a sawfem_analyse result");
  const Analysis& a = *ap;
  return sawfem_mf_create(plan, (int)a.n, (int)a.m.size(), a.m.data(), a.p.data(), a.rows_ptr.data(), a.rows.data(),
                          a.child_ptr.data(), a.child_ids.data(), a.inv_ptr.data(), a.inv.data(),
                          (int)a.lvl_ptr.size() - 1, a.lvl_ptr.data(), a.lvl_fronts.data(), (int64_t)a.ent_front.size(),
                          a.ent_front.data(), a.ent_slot.data(), a.ent_loc.data(), a.ent_lvl_ptr.data());
}

}  // extern "C"
