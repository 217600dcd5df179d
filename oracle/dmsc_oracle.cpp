// CPU oracle for the compute_dmsc path: TEST INFRASTRUCTURE ONLY.
//
// A from-scratch, single-threaded C++17 restatement of the reference CPU (TBB) algorithm in
// canonical single-thread order. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg load it, and only as the checker. The product path (dmsc_amd) never links or calls it.
//
// Every stage cites the reference function it restates (paths relative to the reference repo).
// Ordering conventions that the reference leaves to thread scheduling (critical-cell lists) are
// fixed to ascending cell id, which is what the reference produces with one thread.
//
// Parity pinning (see tests/test_oracle.py):
//   * all 4 single-thread fixtures (primal/dual x raw/filtered) compared field by field;
//   * multi-thread and GPU fixtures compared up to relabelling;
//   * gradient_queue vs a closed-form/LUT formulation on every 3x3 order pattern (test_gradient);
//   * raw persistence vs an independent Kruskal + elder-rule implementation in numpy;
//   * labels vs brute-force V-path walks; invariants (Euler characteristic, acyclicity).
// The thresholded-UF corner where the reference reads crit_mins[-1] cannot be reached with
// monotone float arithmetic; it is guarded here (t_idx -1 cancels are skipped, as the reference
// skips them in simplify_arcs_geometry).

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// ---------------------------------------------------------------- cell codec
// reference: dmsc/csrc/cell_complex.hxx:13-51
static inline int cid(int t, int y, int x, int Nx) { return t + 4 * (y * Nx + x); }
static inline int ctype(int id) { return id % 4; }
static inline int cyy(int id, int Nx) { return (id / 4) / Nx; }
static inline int cxx(int id, int Nx) { return (id / 4) % Nx; }
static inline bool ok_v(int y, int x, int H, int W) { return y >= 0 && y < H && x >= 0 && x < W; }
static inline bool ok_h(int y, int x, int H, int W) { return y >= 0 && y < H && x >= 0 && x <= W; }
static inline bool ok_e(int y, int x, int H, int W) { return y >= 0 && y <= H && x >= 0 && x < W; }
static inline bool ok_f(int y, int x, int H, int W) { return y >= 0 && y <= H && x >= 0 && x <= W; }

// ---------------------------------------------------------------- comparisons
// reference: dmsc/csrc/cpu/cell_compare.hxx:9-31. std::tie ordering: a float pair that is
// neither < nor > (equal, or NaN involved) falls through to the integer id.
template <bool D>
static inline bool val_less(float a, int ia, float b, int ib) {
  if (!D) return (a < b) || (!(b < a) && ia < ib);
  return (b < a) || (!(a < b) && ib < ia);
}
template <bool D>
static inline bool val_greater(float a, int ia, float b, int ib) { return val_less<!D>(a, ia, b, ib); }

struct Ctx {
  int H, W, Nx;
  const float* f;
  template <bool D>
  bool pix_greater(int y1, int x1, int y2, int x2) const {
    int i1 = y1 * W + x1, i2 = y2 * W + x2;
    return val_greater<D>(f[i1], i1, f[i2], i2);
  }
  // reference: cell_compare.hxx:33-63 (candidate order matters when NaN breaks transitivity)
  template <bool D>
  std::pair<int, int> top_vertex(int t, int y, int x) const {
    int by = -1, bx = -1;
    auto consider = [&](int vy, int vx) {
      if (by == -1 || pix_greater<D>(vy, vx, by, bx)) { by = vy; bx = vx; }
    };
    if (t == 0) {
      consider(y, x);
    } else if (t == 1) {
      if (x - 1 >= 0) consider(y, x - 1);
      if (x < W) consider(y, x);
    } else if (t == 2) {
      if (y - 1 >= 0) consider(y - 1, x);
      if (y < H) consider(y, x);
    } else {
      if (y - 1 >= 0 && x - 1 >= 0) consider(y - 1, x - 1);
      if (y - 1 >= 0 && x < W) consider(y - 1, x);
      if (y < H && x - 1 >= 0) consider(y, x - 1);
      if (y < H && x < W) consider(y, x);
    }
    return {by, bx};
  }
  // reference: cell_compare.hxx:82-87
  template <bool D>
  float cell_val(int t, int y, int x) const {
    auto v = top_vertex<D>(t, y, x);
    return f[v.first * W + v.second];
  }
};

// ---------------------------------------------------------------- gradient
// reference: dmsc/csrc/cpu/gradient.hxx:19-196 (lower-star pairing per vertex, FIFO face queue)
template <bool D>
static void gradient(const Ctx& c, std::vector<int>& pw, int nthreads) {
  const int H = c.H, W = c.W, Nx = c.Nx;
  pw.assign((size_t)4 * (H + 1) * (W + 1), -1);
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
  for (int y = 0; y < H; ++y) {
    int le[4], lf[4], fe[4];
    std::vector<int> q;
    q.reserve(16);
    for (int x = 0; x < W; ++x) {
      int ne = 0, nf = 0;
      const std::pair<int, int> me{y, x};
      if (c.top_vertex<D>(1, y, x) == me) le[ne++] = cid(1, y, x, Nx);
      if (c.top_vertex<D>(1, y, x + 1) == me) le[ne++] = cid(1, y, x + 1, Nx);
      if (c.top_vertex<D>(2, y, x) == me) le[ne++] = cid(2, y, x, Nx);
      if (c.top_vertex<D>(2, y + 1, x) == me) le[ne++] = cid(2, y + 1, x, Nx);
      if (c.top_vertex<D>(3, y, x) == me) lf[nf++] = cid(3, y, x, Nx);
      if (c.top_vertex<D>(3, y, x + 1) == me) lf[nf++] = cid(3, y, x + 1, Nx);
      if (c.top_vertex<D>(3, y + 1, x) == me) lf[nf++] = cid(3, y + 1, x, Nx);
      if (c.top_vertex<D>(3, y + 1, x + 1) == me) lf[nf++] = cid(3, y + 1, x + 1, Nx);

      // vertex <-> edge: the lower-star edge whose other end is lowest (gradient.hxx:76-112)
      int best = -1, by = -1, bx = -1;
      for (int i = 0; i < ne; ++i) {
        int e = le[i], ey = cyy(e, Nx), ex = cxx(e, Nx);
        int oy = -1, ox = -1;
        if (ctype(e) == 1) {
          if (ex - 1 == x && ex < W) { oy = ey; ox = ex; }
          else if (ex == x && ex - 1 >= 0) { oy = ey; ox = ex - 1; }
        } else {
          if (ey - 1 == y && ey < H) { oy = ey; ox = ex; }
          else if (ey == y && ey - 1 >= 0) { oy = ey - 1; ox = ex; }
        }
        if (oy != -1 && (best == -1 || c.pix_greater<D>(by, bx, oy, ox))) { best = e; by = oy; bx = ox; }
      }
      if (best != -1) {
        int v = cid(0, y, x, Nx);
        pw[v] = best;
        pw[best] = v;
      }

      // face <-> edge FIFO (gradient.hxx:115-191)
      q.clear();
      for (int i = 0; i < nf; ++i) q.push_back(lf[i]);
      size_t head = 0;
      while (head < q.size()) {
        int fc = q[head++];
        if (pw[fc] != -1) continue;
        int fy = cyy(fc, Nx), fx = cxx(fc, Nx), nfe = 0;
        if (ok_h(fy - 1, fx, H, W)) fe[nfe++] = cid(1, fy - 1, fx, Nx);
        if (ok_h(fy, fx, H, W)) fe[nfe++] = cid(1, fy, fx, Nx);
        if (ok_e(fy, fx - 1, H, W)) fe[nfe++] = cid(2, fy, fx - 1, Nx);
        if (ok_e(fy, fx, H, W)) fe[nfe++] = cid(2, fy, fx, Nx);
        int free_e = -1, nfree = 0;
        for (int k = 0; k < nfe; ++k) {
          bool inl = false;
          for (int j = 0; j < ne; ++j) inl |= (le[j] == fe[k]);
          if (inl && pw[fe[k]] == -1) { ++nfree; free_e = fe[k]; }
        }
        if (nfree != 1) continue;
        pw[fc] = free_e;
        pw[free_e] = fc;
        int ey = cyy(free_e, Nx), ex = cxx(free_e, Nx);
        auto push = [&](int nf2) {
          if (nf2 == fc) return;
          bool inl = false;
          for (int j = 0; j < nf; ++j) inl |= (lf[j] == nf2);
          if (!inl || pw[nf2] != -1) return;
          for (size_t k = head; k < q.size(); ++k)
            if (q[k] == nf2) return;
          q.push_back(nf2);
        };
        if (ctype(free_e) == 1) {
          if (ok_f(ey, ex, H, W)) push(cid(3, ey, ex, Nx));
          if (ok_f(ey + 1, ex, H, W)) push(cid(3, ey + 1, ex, Nx));
        } else {
          if (ok_f(ey, ex, H, W)) push(cid(3, ey, ex, Nx));
          if (ok_f(ey, ex + 1, H, W)) push(cid(3, ey, ex + 1, Nx));
        }
      }
    }
  }
}

// ---------------------------------------------------------------- union-find with void (-1)
// reference: dmsc/csrc/union_find.hxx:6-49
struct UF {
  std::vector<int> p;
  void reset(int n) { p.resize(n); std::iota(p.begin(), p.end(), 0); }
  int find(int i) {
    if (i == -1) return -1;
    int r = i;
    while (r != -1 && p[r] != r) r = p[r];
    int cur = i;
    while (cur != r && cur != -1) { int nx = p[cur]; p[cur] = r; cur = nx; }
    return r;
  }
  void link_roots(int a, int b) {
    if (a != b && a != -1) p[a] = b;
  }
};

struct SadEvent {
  int saddle_id, c1, c2;
  float s_val;
  int pair_id = -1;
  float pers = 0.0f;
};
struct Cancel {
  int saddle_id, s_idx, target_id, t_idx;
  float pers;
  // reference: dmsc/csrc/cpu/persistence_struct.hxx:15-19
  bool operator<(const Cancel& o) const {
    if (pers != o.pers) return pers < o.pers;
    if (saddle_id != o.saddle_id) return saddle_id < o.saddle_id;
    return target_id < o.target_id;
  }
};
struct Arc { int target, offset, length; };
struct SNode { int alive; Arc mx[2], mn[2]; };
struct DagNode { int l, r; bool lf, rf; int len; };
struct PRef { int id; bool fwd; };

struct Result {
  int status = 0;  // 0 ok, 1 = "cancellation tree is too deep" (reference TORCH_CHECK)
  int32_t shape[2];
  std::vector<int32_t> i32[19];
  std::vector<float> f32[19];
  int ndim[19];
  int64_t dims[19][2];
};

template <bool D>
struct Engine {
  Ctx c;
  int H, W, Nx, ncell, nthreads;
  std::vector<int> pw, mins, sads, maxs, fmap;
  std::vector<float> minv, maxv;
  std::vector<int> max_len, min_len;
  std::vector<SadEvent> smax, smin;
  std::vector<SNode> nodes;
  std::vector<int> gmax, gmin;  // raw arc arenas
  std::vector<uint8_t> max_alive, min_alive;
  UF uf_max, uf_min;
  std::vector<Cancel> cmax, cmin;
  std::vector<int> face_groups, vert_groups;
  int status = 0;

  Engine(const float* f, int H_, int W_, int nt) : H(H_), W(W_), Nx(W_ + 1), nthreads(nt) {
    c = Ctx{H, W, Nx, f};
    ncell = 4 * (H + 1) * (W + 1);
  }

  // reference: critical_points.hxx:14-65 (single-thread order = ascending id) and
  // gradient.hxx:198-231 (fast_crit_map, extremum values)
  void critical_points() {
    mins.clear(); sads.clear(); maxs.clear();
    for (int id = 0; id < ncell; ++id) {
      if (pw[id] != -1) continue;
      int t = ctype(id), y = cyy(id, Nx), x = cxx(id, Nx);
      if (t == 0) { if (ok_v(y, x, H, W)) mins.push_back(id); }
      else if (t == 1) { if (ok_h(y, x, H, W)) sads.push_back(id); }
      else if (t == 2) { if (ok_e(y, x, H, W)) sads.push_back(id); }
      else { if (ok_f(y, x, H, W)) maxs.push_back(id); }
    }
    fmap.assign(ncell, -1);
    maxv.resize(maxs.size());
    minv.resize(mins.size());
    for (size_t i = 0; i < maxs.size(); ++i) {
      fmap[maxs[i]] = (int)i;
      maxv[i] = c.cell_val<D>(3, cyy(maxs[i], Nx), cxx(maxs[i], Nx));
    }
    for (size_t i = 0; i < mins.size(); ++i) {
      fmap[mins[i]] = (int)i;
      minv[i] = c.cell_val<D>(0, cyy(mins[i], Nx), cxx(mins[i], Nx));
    }
    for (size_t i = 0; i < sads.size(); ++i) fmap[sads[i]] = (int)i;
  }

  // the two cofaces of edge e (always valid for a valid edge)
  void edge_faces(int e, int& f1, int& f2) const {
    int ey = cyy(e, Nx), ex = cxx(e, Nx);
    f1 = f2 = -1;
    if (ctype(e) == 1) {
      if (ok_f(ey, ex, H, W)) f1 = cid(3, ey, ex, Nx);
      if (ok_f(ey + 1, ex, H, W)) f2 = cid(3, ey + 1, ex, Nx);
    } else {
      if (ok_f(ey, ex, H, W)) f1 = cid(3, ey, ex, Nx);
      if (ok_f(ey, ex + 1, H, W)) f2 = cid(3, ey, ex + 1, Nx);
    }
  }
  // the (up to two) vertices of edge e, -1 where the edge lies on the image border
  void edge_verts(int e, int& v1, int& v2) const {
    int ey = cyy(e, Nx), ex = cxx(e, Nx);
    v1 = v2 = -1;
    if (ctype(e) == 1) {
      if (ex - 1 >= 0) v1 = cid(0, ey, ex - 1, Nx);
      if (ex < W) v2 = cid(0, ey, ex, Nx);
    } else {
      if (ey - 1 >= 0) v1 = cid(0, ey - 1, ex, Nx);
      if (ey < H) v2 = cid(0, ey, ex, Nx);
    }
  }
  // reference: arcs_topology.hxx:19-65 (V-path walks with step counts)
  std::pair<int, int> walk_faces(int s) const {
    if (s == -1) return {-1, 0};
    int cur = s, n = 0;
    while (cur != -1 && pw[cur] != -1) {
      ++n;
      int a, b;
      edge_faces(pw[cur], a, b);
      cur = (a == cur) ? b : a;
    }
    return {cur, n};
  }
  std::pair<int, int> walk_verts(int s) const {
    if (s == -1) return {-1, 0};
    int cur = s, n = 0;
    while (cur != -1 && pw[cur] != -1) {
      ++n;
      int a, b;
      edge_verts(pw[cur], a, b);
      cur = (a == cur) ? b : a;
    }
    return {cur, n};
  }

  // reference: arcs_topology.hxx:67-172
  void saddle_records() {
    int ns = (int)sads.size();
    smax.assign(ns, SadEvent{});
    smin.assign(ns, SadEvent{});
    max_len.assign(2 * ns, 0);
    min_len.assign(2 * ns, 0);
    std::vector<float> sv(ns);
    for (int j = 0; j < ns; ++j) {
      int s = sads[j];
      sv[j] = c.cell_val<D>(ctype(s), cyy(s, Nx), cxx(s, Nx));
    }
    std::vector<int> ord(ns);
    std::iota(ord.begin(), ord.end(), 0);
    // keys are unique (ids), so a stable sort equals any correct sort; stable_sort never reads
    // out of bounds even if NaN makes the comparator inconsistent.
    std::stable_sort(ord.begin(), ord.end(),
                     [&](int i, int j) { return val_greater<D>(sv[i], sads[i], sv[j], sads[j]); });
    for (int j = 0; j < ns; ++j) {
      int i = ord[j], s = sads[i];
      int f1, f2;
      edge_faces(s, f1, f2);
      auto a = f1 != -1 ? walk_faces(f1) : std::pair<int, int>{-1, 0};
      auto b = f2 != -1 ? walk_faces(f2) : std::pair<int, int>{-1, 0};
      SadEvent ev{-999, -1, -1, 0.0f};
      if (a.first != -1 && b.first != -1) {
        ev = SadEvent{s, fmap[a.first], fmap[b.first], sv[i]};
        max_len[2 * i] = a.second;
        max_len[2 * i + 1] = b.second;
      } else if (a.first != -1 || b.first != -1) {
        auto r = a.first != -1 ? a : b;
        ev = SadEvent{s, fmap[r.first], -1, sv[i]};
        max_len[2 * i] = r.second;
      }
      smax[j] = ev;
    }
    for (int j = 0; j < ns; ++j) {
      int i = ord[ns - j - 1], s = sads[i];
      int v1, v2;
      edge_verts(s, v1, v2);
      auto a = v1 != -1 ? walk_verts(v1) : std::pair<int, int>{-1, 0};
      auto b = v2 != -1 ? walk_verts(v2) : std::pair<int, int>{-1, 0};
      SadEvent ev{-999, -1, -1, 0.0f};
      if (a.first != -1 && b.first != -1) {
        ev = SadEvent{s, fmap[a.first], fmap[b.first], sv[i]};
        min_len[2 * i] = a.second;
        min_len[2 * i + 1] = b.second;
      } else if (a.first != -1 || b.first != -1) {
        auto r = a.first != -1 ? a : b;
        ev = SadEvent{s, fmap[r.first], -1, sv[i]};
        min_len[2 * i] = r.second;
      }
      smin[j] = ev;
    }
  }

  // reference: arcs_geometry.hxx:14-169 (flat arenas of raw arcs; arc k = k-th existing start)
  void raw_arcs(bool tmax, bool tmin) {
    int ns = (int)sads.size();
    nodes.assign(ns, SNode{});
    if (ns == 0) return;
    std::vector<int> moff(2 * ns + 1, 0), noff(2 * ns + 1, 0);
    for (int i = 0; i < 2 * ns; ++i) {
      moff[i + 1] = tmax ? moff[i] + max_len[i] + 1 : 0;
      noff[i + 1] = tmin ? noff[i] + min_len[i] + 1 : 0;
    }
    gmax.assign(moff.back(), 0);
    gmin.assign(noff.back(), 0);
    for (int i = 0; i < ns; ++i) {
      SNode& nd = nodes[i];
      nd.alive = 1;
      int s = sads[i];
      nd.mx[0] = nd.mx[1] = Arc{-1, 0, 0};
      nd.mn[0] = nd.mn[1] = Arc{-1, 0, 0};
      if (tmax) {
        int f[2];
        edge_faces(s, f[0], f[1]);
        int k = 0;
        for (int a = 0; a < 2; ++a) {
          if (f[a] == -1) continue;
          int cur = f[a], base = moff[2 * i + k], n = 0;
          while (cur != -1) {
            if (ctype(cur) == 3) gmax[base + n++] = cur;
            if (pw[cur] == -1) break;
            int p, q2;
            edge_faces(pw[cur], p, q2);
            cur = (p == cur) ? q2 : p;
          }
          nd.mx[k] = Arc{cur != -1 ? fmap[cur] : -1, base, n};
          ++k;
        }
      }
      if (tmin) {
        int v[2];
        edge_verts(s, v[0], v[1]);
        int k = 0;
        for (int a = 0; a < 2; ++a) {
          if (v[a] == -1) continue;
          int cur = v[a], base = noff[2 * i + k], n = 0;
          while (cur != -1) {
            if (ctype(cur) == 0) gmin[base + n++] = cur;
            if (pw[cur] == -1) break;
            int p, q2;
            edge_verts(pw[cur], p, q2);
            cur = (p == cur) ? q2 : p;
          }
          nd.mn[k] = Arc{cur != -1 ? fmap[cur] : -1, base, n};
          ++k;
        }
      }
    }
  }

  // reference: persistence.hxx:16-182 (elder rule with void, thresholded UF, cancellations)
  template <bool MAXSIDE>
  void persistence_side(float thr, bool record, std::vector<int>& removed) {
    auto& evs = MAXSIDE ? smax : smin;
    auto& ext = MAXSIDE ? maxs : mins;
    auto& vals = MAXSIDE ? maxv : minv;
    auto& alive = MAXSIDE ? max_alive : min_alive;
    auto& uft = MAXSIDE ? uf_max : uf_min;
    auto& canc = MAXSIDE ? cmax : cmin;
    UF ufg = uft;
    for (auto& ev : evs) {
      int r1 = ufg.find(ev.c1), r2 = ufg.find(ev.c2);
      if (r1 == r2) {
        ev.pair_id = -1;
        ev.pers = -1.0f;
        continue;
      }
      int surv = -1, dying = -1;
      if (r1 == -1) dying = r2;
      else if (r2 == -1) dying = r1;
      else {
        bool r1_dies = MAXSIDE ? val_less<D>(vals[r1], ext[r1], vals[r2], ext[r2])
                               : val_greater<D>(vals[r1], ext[r1], vals[r2], ext[r2]);
        if (r1_dies) { surv = r2; dying = r1; }
        else { surv = r1; dying = r2; }
      }
      if (!MAXSIDE) ev.pers = D ? vals[dying] - ev.s_val : ev.s_val - vals[dying];
      else ev.pers = D ? ev.s_val - vals[dying] : vals[dying] - ev.s_val;
      ev.pair_id = ext[dying];
      ufg.link_roots(dying, surv);
      if (ev.pers <= thr) {
        removed.push_back(ev.saddle_id);
        alive[dying] = 0;
        int d = uft.find(ev.c1), sv = uft.find(ev.c2);
        if (dying == r2) std::swap(d, sv);
        uft.link_roots(d, sv);
        if (record) canc.push_back(Cancel{ev.saddle_id, fmap[ev.saddle_id], d >= 0 ? ext[d] : -1, d, ev.pers});
      }
    }
  }

  void persistence(float thr, bool tmax, bool tmin) {
    min_alive.assign(mins.size(), 1);
    max_alive.assign(maxs.size(), 1);
    uf_min.reset((int)mins.size());
    uf_max.reset((int)maxs.size());
    cmin.clear();
    cmax.clear();
    std::vector<int> rmin, rmax;
    persistence_side<false>(thr, tmin, rmin);
    persistence_side<true>(thr, tmax, rmax);
    for (int s : rmin) fmap[s] = -1;
    for (int s : rmax) fmap[s] = -1;
    auto gone = [&](const SadEvent& e) { return fmap[e.saddle_id] < 0; };
    smin.erase(std::remove_if(smin.begin(), smin.end(), gone), smin.end());
    smax.erase(std::remove_if(smax.begin(), smax.end(), gone), smax.end());
  }

  // reference: arcs_simplification.hxx:310-504 and 37-308 (path DAG contraction, flat assembly)
  struct GeoSide {
    std::vector<int> parent, tinit, blen, boff;
    std::vector<PRef> w;
    std::vector<DagNode> dag;
    int dsz = 0;
    std::vector<uint8_t> salive;
  };

  PRef merge(GeoSide& g, PRef a, PRef b, int N2) {
    if (a.id == -1) return b;
    if (b.id == -1) return a;
    if ((size_t)g.dsz >= g.dag.size()) { status = 1; return a; }
    int la = a.id < N2 ? g.blen[a.id] : g.dag[a.id - N2].len;
    int lb = b.id < N2 ? g.blen[b.id] : g.dag[b.id - N2].len;
    int nid = g.dsz + N2;
    g.dag[g.dsz++] = DagNode{a.id, b.id, a.fwd, b.fwd, la + lb};
    return PRef{nid, true};
  }
  int gfind(GeoSide& g, int i, int N2) {
    if (i == -1) return -1;
    if (g.parent[i] == i) return i;
    int p = g.parent[i];
    int r = gfind(g, p, N2);
    if (p != -1 && g.w[p].id != -1) g.w[i] = merge(g, g.w[i], g.w[p], N2);
    g.parent[i] = r;
    return r;
  }
  int plen(GeoSide& g, PRef p, int N2) { return p.id < N2 ? g.blen[p.id] : g.dag[p.id - N2].len; }

  void contract(GeoSide& g, std::vector<Cancel>& canc, int next, int N2) {
    std::stable_sort(canc.begin(), canc.end());
    for (const auto& cc : canc) {
      int s = cc.s_idx;
      if (!g.salive[s]) continue;
      int dead = cc.t_idx;
      if (dead < 0 || dead >= next) continue;
      int T0 = g.tinit[2 * s], T1 = g.tinit[2 * s + 1];
      int R0 = gfind(g, T0, N2), R1 = gfind(g, T1, N2);
      PRef w0 = T0 == -1 ? PRef{-1, true} : g.w[T0];
      PRef w1 = T1 == -1 ? PRef{-1, true} : g.w[T1];
      PRef F0 = merge(g, PRef{2 * s, true}, w0, N2);
      PRef F1 = merge(g, PRef{2 * s + 1, true}, w1, N2);
      bool merged = false;
      if (R0 == dead && R0 != R1) {
        g.parent[R0] = R1;
        g.w[R0] = merge(g, PRef{F0.id, !F0.fwd}, F1, N2);
        merged = true;
      } else if (R1 == dead && R0 != R1) {
        g.parent[R1] = R0;
        g.w[R1] = merge(g, PRef{F1.id, !F1.fwd}, F0, N2);
        merged = true;
      }
      if (merged) g.salive[s] = 0;
    }
    for (int i = 0; i < next; ++i) gfind(g, i, N2);
  }

  void expand(const GeoSide& g, PRef p, int* out, const std::vector<int>& arena, int N2) const {
    if (p.id == -1) return;
    std::vector<PRef> st{p};
    int h = 0;
    while (!st.empty()) {
      PRef cur = st.back();
      st.pop_back();
      if (cur.id < N2) {
        int len = g.blen[cur.id], off = g.boff[cur.id];
        for (int k = 0; k < len; ++k) out[h++] = cur.fwd ? arena[off + k] : arena[off + len - 1 - k];
      } else {
        const DagNode& n = g.dag[cur.id - N2];
        if (cur.fwd) { st.push_back({n.r, n.rf}); st.push_back({n.l, n.lf}); }
        else { st.push_back({n.l, !n.lf}); st.push_back({n.r, !n.rf}); }
      }
    }
  }

  void simplify_geometry(bool tmax, bool tmin) {
    int ns = (int)sads.size(), N2 = 2 * ns;
    GeoSide gx, gn;
    for (GeoSide* g : {&gx, &gn}) {
      g->tinit.assign(N2, -1);
      g->blen.assign(N2, 0);
      g->boff.assign(N2, 0);
      g->salive.assign(ns, 1);
    }
    gx.parent.resize(maxs.size());
    std::iota(gx.parent.begin(), gx.parent.end(), 0);
    gn.parent.resize(mins.size());
    std::iota(gn.parent.begin(), gn.parent.end(), 0);
    gx.w.assign(maxs.size(), PRef{-1, true});
    gn.w.assign(mins.size(), PRef{-1, true});
    // reference capacity heuristic: arcs_simplification.hxx:16-19,336-339
    gx.dag.resize(std::max<size_t>((size_t)1 << 18, cmax.size() * 5));
    gn.dag.resize(std::max<size_t>((size_t)1 << 18, cmin.size() * 5));
    for (int i = 0; i < ns; ++i) {
      if (tmax)
        for (int k = 0; k < 2; ++k) {
          gx.tinit[2 * i + k] = nodes[i].mx[k].target;
          gx.blen[2 * i + k] = nodes[i].mx[k].length;
          gx.boff[2 * i + k] = nodes[i].mx[k].offset;
        }
      if (tmin)
        for (int k = 0; k < 2; ++k) {
          gn.tinit[2 * i + k] = nodes[i].mn[k].target;
          gn.blen[2 * i + k] = nodes[i].mn[k].length;
          gn.boff[2 * i + k] = nodes[i].mn[k].offset;
        }
    }
    if (tmax) contract(gx, cmax, (int)maxs.size(), N2);
    if (tmin) contract(gn, cmin, (int)mins.size(), N2);

    // assemble_simplified_geometry (arcs_simplification.hxx:81-142): final path per arc of each
    // surviving saddle, then prefix offsets, then DAG expansion into fresh arenas.
    std::vector<PRef> fx(N2, PRef{-1, true}), fn(N2, PRef{-1, true});
    std::vector<int> lx(N2, 0), ln(N2, 0);
    for (int i = 0; i < ns; ++i) {
      if (fmap[sads[i]] == -1) continue;
      for (int k = 0; k < 2; ++k) {
        int a = 2 * i + k;
        if (gx.tinit[a] != -1) { fx[a] = merge(gx, PRef{a, true}, gx.w[gx.tinit[a]], N2); lx[a] = plen(gx, fx[a], N2); }
      }
    }
    for (int i = 0; i < ns; ++i) {
      if (fmap[sads[i]] == -1) continue;
      for (int k = 0; k < 2; ++k) {
        int a = 2 * i + k;
        if (gn.tinit[a] != -1) { fn[a] = merge(gn, PRef{a, true}, gn.w[gn.tinit[a]], N2); ln[a] = plen(gn, fn[a], N2); }
      }
    }
    if (status) return;
    std::vector<int64_t> ox(N2 + 1, 0), on(N2 + 1, 0);
    for (int a = 0; a < N2; ++a) { ox[a + 1] = ox[a] + lx[a]; on[a + 1] = on[a] + ln[a]; }
    std::vector<int> nx(ox.back()), nn(on.back());
    for (int i = 0; i < ns; ++i) {
      bool al = gx.salive[i] && gn.salive[i];
      nodes[i].alive = al;
      if (!al) {
        for (int k = 0; k < 2; ++k) nodes[i].mx[k].length = nodes[i].mn[k].length = 0;
        continue;
      }
      for (int k = 0; k < 2; ++k) {
        int a = 2 * i + k;
        if (gx.tinit[a] != -1) {
          nodes[i].mx[k] = Arc{gx.parent[gx.tinit[a]], (int)ox[a], lx[a]};
          expand(gx, fx[a], nx.data() + ox[a], gmax, N2);
        }
        if (gn.tinit[a] != -1) {
          nodes[i].mn[k] = Arc{gn.parent[gn.tinit[a]], (int)on[a], ln[a]};
          expand(gn, fn[a], nn.data() + on[a], gmin, N2);
        }
      }
    }
    gmax.swap(nx);
    gmin.swap(nn);
  }

  // reference: cell_groups.hxx:96-229. The reference's opportunistic label cache only changes
  // how the V-path terminal is reached, not the label, so each cell is walked directly here.
  void cell_groups(bool faces, bool verts) {
    for (int i = 0; i < (int)uf_max.p.size(); ++i) uf_max.find(i);
    for (int i = 0; i < (int)uf_min.p.size(); ++i) uf_min.find(i);
    std::vector<int> reg_max(maxs.size(), -1), reg_min(mins.size(), -1);
    for (int i = 0, r = 0; i < (int)maxs.size(); ++i) if (max_alive[i]) reg_max[i] = r++;
    for (int i = 0, r = 0; i < (int)mins.size(); ++i) if (min_alive[i]) reg_min[i] = r++;
    if (faces) {
      face_groups.assign((size_t)(H + 1) * (W + 1), -1);
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
      for (int y = 0; y <= H; ++y)
        for (int x = 0; x <= W; ++x) {
          int t = walk_faces(cid(3, y, x, Nx)).first, lab = -1;
          if (t != -1 && fmap[t] != -1) {
            int r = uf_max.p[fmap[t]];
            if (r != -1) lab = reg_max[r];
          }
          face_groups[(size_t)y * Nx + x] = lab;
        }
    }
    if (verts) {
      vert_groups.assign((size_t)H * W, -1);
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
      for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
          int t = walk_verts(cid(0, y, x, Nx)).first, lab = -1;
          if (t != -1 && fmap[t] != -1) {
            int r = uf_min.p[fmap[t]];
            if (r != -1) lab = reg_min[r];
          }
          vert_groups[(size_t)y * W + x] = lab;
        }
    }
  }

  // reference: cpu/workspace.hxx:86-365 (output assembly, dual swap)
  void complex(Result& R, bool grad, bool tmax, bool tmin, bool gface, bool gvert) {
    int nsurv = (int)smin.size();
    std::vector<int> max_to(maxs.size(), -1), min_to(mins.size(), -1), sad_to(sads.size(), -1);
    std::vector<int32_t> omax, omin, osad(nsurv), ppmax(nsurv), ppmin(nsurv), emax, emin;
    std::vector<float> pmax(nsurv), pmin(nsurv);
    for (size_t i = 0; i < maxs.size(); ++i) if (max_alive[i]) { max_to[i] = (int)omax.size(); omax.push_back(maxs[i]); }
    for (size_t i = 0; i < mins.size(); ++i) if (min_alive[i]) { min_to[i] = (int)omin.size(); omin.push_back(mins[i]); }
    for (int i = 0; i < nsurv; ++i) {
      auto& ev = smin[i];
      sad_to[fmap[ev.saddle_id]] = i;
      osad[i] = ev.saddle_id;
      ev.saddle_id = i;
      ev.pair_id = ev.pair_id != -1 ? fmap[ev.pair_id] : -1;
    }
    for (int i = 0; i < nsurv; ++i) {
      auto& ev = smax[i];
      ev.saddle_id = fmap[ev.saddle_id];
      ev.pair_id = ev.pair_id != -1 ? fmap[ev.pair_id] : -1;
    }
    auto out_of = [](const std::vector<int>& to, int r) { return r >= 0 ? to[r] : -1; };
    for (int i = 0; i < nsurv; ++i) {
      const auto& ev = smin[i];
      int si = ev.saddle_id;
      pmin[si] = ev.pers;
      ppmin[si] = ev.pair_id != -1 ? out_of(min_to, uf_min.find(ev.pair_id)) : -1;
      int r1 = ev.c1 != -1 ? uf_min.find(ev.c1) : -1, r2 = ev.c2 != -1 ? uf_min.find(ev.c2) : -1;
      if (r1 != -1 && min_to[r1] != -1) { emin.push_back(si); emin.push_back(min_to[r1]); }
      if (r1 != r2 && r2 != -1 && min_to[r2] != -1) { emin.push_back(si); emin.push_back(min_to[r2]); }
    }
    for (int i = 0; i < nsurv; ++i) {
      const auto& ev = smax[i];
      int si = sad_to[ev.saddle_id];
      pmax[si] = ev.pers;
      ppmax[si] = ev.pair_id != -1 ? out_of(max_to, uf_max.find(ev.pair_id)) : -1;
      int r1 = ev.c1 != -1 ? uf_max.find(ev.c1) : -1, r2 = ev.c2 != -1 ? uf_max.find(ev.c2) : -1;
      if (r1 != -1 && max_to[r1] != -1) { emax.push_back(si); emax.push_back(max_to[r1]); }
      if (r1 != r2 && r2 != -1 && max_to[r2] != -1) { emax.push_back(si); emax.push_back(max_to[r2]); }
    }
    std::vector<int32_t> rf, rfo, afo, rv, rvo, avo;
    if (tmax || tmin) {
      if (tmax) rfo.push_back(0);
      if (tmin) rvo.push_back(0);
      for (const auto& ev : smin) {
        int si = fmap[osad[ev.saddle_id]];
        const SNode& nd = nodes[si];
        if (!nd.alive) continue;
        if (tmax) {
          for (int k = 0; k < 2; ++k) {
            rf.insert(rf.end(), gmax.begin() + nd.mx[k].offset, gmax.begin() + nd.mx[k].offset + nd.mx[k].length);
            if (k == 0) afo.push_back((int32_t)rf.size());
          }
          rfo.push_back((int32_t)rf.size());
        }
        if (tmin) {
          for (int k = 0; k < 2; ++k) {
            rv.insert(rv.end(), gmin.begin() + nd.mn[k].offset, gmin.begin() + nd.mn[k].offset + nd.mn[k].length);
            if (k == 0) avo.push_back((int32_t)rv.size());
          }
          rvo.push_back((int32_t)rv.size());
        }
      }
    }
    R.shape[0] = H;
    R.shape[1] = W;
    auto put = [&](int slot, std::vector<int32_t> v, int64_t d0, int64_t d1, int nd) {
      R.i32[slot] = std::move(v);
      R.ndim[slot] = nd;
      R.dims[slot][0] = d0;
      R.dims[slot][1] = d1;
    };
    auto putf = [&](int slot, std::vector<float> v) {
      int64_t n = (int64_t)v.size();
      R.f32[slot] = std::move(v);
      R.ndim[slot] = 1;
      R.dims[slot][0] = n;
      R.dims[slot][1] = 0;
    };
    auto put1 = [&](int slot, std::vector<int32_t> v) { int64_t n = (int64_t)v.size(); put(slot, std::move(v), n, 0, 1); };
    put1(0, {H, W});
    std::vector<int32_t> gi;
    if (grad) gi.assign(pw.begin(), pw.end());
    put1(10, std::move(gi));
    int64_t nemax = (int64_t)emax.size() / 2, nemin = (int64_t)emin.size() / 2;
    std::vector<int32_t> fg, vg;
    int64_t fh = 0, fw = 0, vh = 0, vw = 0;
    if (gface) { fg.assign(face_groups.begin(), face_groups.end()); fh = H + 1; fw = W + 1; }
    if (gvert) { vg.assign(vert_groups.begin(), vert_groups.end()); vh = H; vw = W; }
    put1(3, std::move(osad));
    if (!D) {
      put1(1, std::move(omax)); put1(2, std::move(omin));
      put(4, std::move(emax), nemax, 2, 2); put(5, std::move(emin), nemin, 2, 2);
      putf(6, std::move(pmax)); putf(7, std::move(pmin));
      put1(8, std::move(ppmax)); put1(9, std::move(ppmin));
      put(11, std::move(fg), fh, fw, 2);
      put1(12, tmax ? rf : std::vector<int32_t>{}); put1(13, tmax ? rfo : std::vector<int32_t>{}); put1(14, tmax ? afo : std::vector<int32_t>{});
      put(15, std::move(vg), vh, vw, 2);
      put1(16, tmin ? rv : std::vector<int32_t>{}); put1(17, tmin ? rvo : std::vector<int32_t>{}); put1(18, tmin ? avo : std::vector<int32_t>{});
    } else {
      put1(1, std::move(omin)); put1(2, std::move(omax));
      put(4, std::move(emin), nemin, 2, 2); put(5, std::move(emax), nemax, 2, 2);
      putf(6, std::move(pmin)); putf(7, std::move(pmax));
      put1(8, std::move(ppmin)); put1(9, std::move(ppmax));
      put(11, std::move(vg), vh, vw, 2);
      put1(12, tmin ? rv : std::vector<int32_t>{}); put1(13, tmin ? rvo : std::vector<int32_t>{}); put1(14, tmin ? avo : std::vector<int32_t>{});
      put(15, std::move(fg), fh, fw, 2);
      put1(16, tmax ? rf : std::vector<int32_t>{}); put1(17, tmax ? rfo : std::vector<int32_t>{}); put1(18, tmax ? afo : std::vector<int32_t>{});
    }
  }

  // reference: dmsc_cpu.cpp:29-51 (stage order, flag swap in dual)
  void run(Result& R, float thr, bool grad, bool tmax, bool tmin, bool gmax_, bool gmin_) {
    if (D) { std::swap(tmax, tmin); std::swap(gmax_, gmin_); }
    gradient<D>(c, pw, nthreads);
    critical_points();
    saddle_records();
    bool any = tmax || tmin;
    if (any) raw_arcs(tmax, tmin);
    persistence(thr, tmax, tmin);
    if (any) simplify_geometry(tmax, tmin);
    if (status) { R.status = status; return; }
    if (gmax_ || gmin_) cell_groups(gmax_, gmin_);
    complex(R, grad, tmax, tmin, gmax_, gmin_);
    R.status = 0;
  }
};

}  // namespace orc

extern "C" {

enum {
  ORC_RETURN_GRADIENT = 1,
  ORC_DUAL = 2,
  ORC_TRACE_MAX_ARCS = 4,
  ORC_TRACE_MIN_ARCS = 8,
  ORC_TRACE_MAX_GROUPS = 16,
  ORC_TRACE_MIN_GROUPS = 32,
};

// Runs the full restated pipeline on a contiguous HxW float32 field.
// flags use the reference's extract_dmsc argument meaning (before the dual swap).
void* orc_extract(const float* field, int H, int W, float thr, unsigned flags, int nthreads, int* status) {
  auto* R = new orc::Result();
  bool grad = flags & ORC_RETURN_GRADIENT, tmax = flags & ORC_TRACE_MAX_ARCS, tmin = flags & ORC_TRACE_MIN_ARCS;
  bool gmx = flags & ORC_TRACE_MAX_GROUPS, gmn = flags & ORC_TRACE_MIN_GROUPS;
  if (nthreads < 1) nthreads = 1;
  if (flags & ORC_DUAL) {
    orc::Engine<true> e(field, H, W, nthreads);
    e.run(*R, thr, grad, tmax, tmin, gmx, gmn);
  } else {
    orc::Engine<false> e(field, H, W, nthreads);
    e.run(*R, thr, grad, tmax, tmin, gmx, gmn);
  }
  *status = R->status;
  return R;
}

// Output slot access (slot order = reference DMSComplex.__getitem__). Returns ndim; is_float tells
// whether data points to float32 (p_max, p_min) or int32.
int orc_get(void* h, int slot, const void** data, int64_t* dims, int* is_float) {
  auto* R = static_cast<orc::Result*>(h);
  *is_float = (slot == 6 || slot == 7);
  *data = *is_float ? (const void*)R->f32[slot].data() : (const void*)R->i32[slot].data();
  dims[0] = R->dims[slot][0];
  dims[1] = R->dims[slot][1];
  return R->ndim[slot];
}

void orc_free(void* h) { delete static_cast<orc::Result*>(h); }

// Gradient stage only (reference gradient.hxx:19-196): writes 4*(H+1)*(W+1) ints into pw.
void orc_gradient(const float* field, int H, int W, int dual, int nthreads, int* pw) {
  std::vector<int> v;
  orc::Ctx c{H, W, W + 1, field};
  if (dual) orc::gradient<true>(c, v, nthreads < 1 ? 1 : nthreads);
  else orc::gradient<false>(c, v, nthreads < 1 ? 1 : nthreads);
  std::memcpy(pw, v.data(), v.size() * sizeof(int));
}

}  // extern "C"
