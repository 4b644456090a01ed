// gg_host.cpp -- host-side graph construction and the plain simulated
// annealing driver of libggeval.so (C ABI in include/gg_host.h).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

#include "gg_eval.h"
#include "gg_incr.h"
#include "gg_host.h"
#include "gg_internal.h"
#include "gg_sa_common.h"

using gg::EdgeSet;
using gg::ekey;
using gg::now_s;
using gg::Rng;

namespace {

// Fixed-degree adjacency under construction.
struct Adj {
  int n, d;
  std::vector<int32_t> nb;
  std::vector<int32_t> deg;
  Adj(int n_, int d_) : n(n_), d(d_), nb((size_t)n_ * d_, -1), deg(n_, 0) {}
  bool has(int u, int v) const {
    const int32_t *r = &nb[(size_t)u * d];
    for (int k = 0; k < deg[u]; ++k)
      if (r[k] == v) return true;
    return false;
  }
  void add(int u, int v) {
    nb[(size_t)u * d + deg[u]++] = v;
    nb[(size_t)v * d + deg[v]++] = u;
  }
  void remove_one(int u, int v) {
    int32_t *r = &nb[(size_t)u * d];
    for (int k = 0; k < deg[u]; ++k)
      if (r[k] == v) {
        r[k] = r[--deg[u]];
        r[deg[u]] = -1;
        return;
      }
  }
  void remove(int u, int v) {
    remove_one(u, v);
    remove_one(v, u);
  }
};

bool connected_adj(int n, int d, const int32_t *adj) {
  std::vector<char> seen(n, 0);
  std::vector<int> q;
  q.reserve(n);
  q.push_back(0);
  seen[0] = 1;
  for (size_t h = 0; h < q.size(); ++h) {
    int v = q[h];
    for (int j = 0; j < d; ++j) {
      int u = adj[(size_t)v * d + j];
      if (u >= 0 && !seen[u]) {
        seen[u] = 1;
        q.push_back(u);
      }
    }
  }
  return (int)q.size() == n;
}

// One attempt of stub pairing.  `ok(u,v)` decides admissibility; returns
// false when the remaining stubs admit no pair (caller restarts).
template <typename Ok, typename Add>
bool pair_stubs(std::vector<int32_t> &pts, Rng &rng, Ok ok, Add add) {
  size_t rem = pts.size();
  size_t fails = 0;
  while (rem > 0) {
    if (rem == 1) return false;
    uint32_t i = rng.below((uint32_t)rem), j = rng.below((uint32_t)rem);
    if (i == j) continue;
    int u = pts[i], v = pts[j];
    if (!ok(u, v)) {
      if (++fails > 64 * rem + 256) {
        // Is any admissible pair left at all?
        bool any = false;
        for (size_t a = 0; a < rem && !any; ++a)
          for (size_t b = a + 1; b < rem && !any; ++b) any = ok(pts[a], pts[b]);
        if (!any) return false;
        fails = 0;
      }
      continue;
    }
    add(u, v);
    fails = 0;
    if (i < j) std::swap(i, j);  // remove the larger index first
    pts[i] = pts[--rem];
    pts[j] = pts[--rem];
  }
  return true;
}

}  // namespace

using gg::set_error;

extern "C" {

int gg_edges_to_adj(int nodes, int lines, int degree, const int32_t *edges, int32_t *adj) {
  // general/sa.c:61-74: each vertex's row filled in edge order.  Rows with
  // fewer than `degree` edges are padded with the vertex itself (a
  // self-loop changes no distance); a vertex with more edges is an error.
  std::vector<int32_t> fill(nodes, 0);
  for (int e = 0; e < lines; ++e) {
    int a = edges[2 * e], b = edges[2 * e + 1];
    if (a < 0 || a >= nodes || b < 0 || b >= nodes)
      return set_error(GG_ERR_INVALID, "edge %d (%d,%d) outside [0,%d)", e, a, b, nodes);
    if (fill[a] >= degree || fill[b] >= degree)
      return set_error(GG_ERR_INVALID, "edge %d exceeds degree %d", e, degree);
    adj[(size_t)a * degree + fill[a]++] = b;
    adj[(size_t)b * degree + fill[b]++] = a;
  }
  for (int v = 0; v < nodes; ++v)
    for (int k = fill[v]; k < degree; ++k) adj[(size_t)v * degree + k] = v;
  return GG_OK;
}

int gg_random_regular(int nodes, int degree, uint64_t seed, int32_t *edges) {
  if (nodes < 2 || degree < 1 || degree >= nodes || ((int64_t)nodes * degree) % 2)
    return set_error(GG_ERR_INVALID, "no simple %d-regular graph on %d vertices", degree, nodes);
  Rng rng(seed);
  for (int attempt = 0; attempt < 1000; ++attempt) {
    Adj A(nodes, degree);
    std::vector<int32_t> pts((size_t)nodes * degree);
    for (size_t k = 0; k < pts.size(); ++k) pts[k] = (int32_t)(k / degree);
    size_t line = 0;
    bool ok = pair_stubs(
        pts, rng, [&](int u, int v) { return u != v && !A.has(u, v); },
        [&](int u, int v) {
          A.add(u, v);
          edges[2 * line] = u;
          edges[2 * line + 1] = v;
          ++line;
        });
    if (!ok) continue;
    if (!connected_adj(nodes, degree, A.nb.data())) continue;
    return GG_OK;
  }
  return set_error(GG_ERR_INVALID, "random regular graph construction did not converge");
}

int gg_random_symmetric(int nodes, int degree, int groups, uint64_t seed, int32_t *edges) {
  if (groups < 1 || nodes % groups)
    return set_error(GG_ERR_INVALID, "nodes %d not divisible by groups %d", nodes, groups);
  if (groups == 1) return gg_random_regular(nodes, degree, seed, edges);
  const int B = nodes / groups;
  if (((int64_t)B * degree) % 2 || degree >= nodes)
    return set_error(GG_ERR_INVALID, "no symmetric %d-regular graph with B=%d", degree, B);
  Rng rng(seed);
  for (int attempt = 0; attempt < 1000; ++attempt) {
    Adj A(nodes, degree);
    std::vector<int32_t> pts((size_t)B * degree);
    for (size_t k = 0; k < pts.size(); ++k) pts[k] = (int32_t)(k / degree);
    int orbit = 0;
    uint32_t volt = 0;
    // The voltage of the pair under test is drawn before the admissibility
    // test and reused by the add callback.
    bool ok = pair_stubs(
        pts, rng,
        [&](int a, int b) {
          if (a == b) return false;
          volt = rng.below((uint32_t)groups);
          return !A.has(a, b + (int)volt * B);
        },
        [&](int a, int b) {
          for (int i = 0; i < groups; ++i) {
            int x = a + i * B, y = b + ((i + (int)volt) % groups) * B;
            A.add(x, y);
            int line = orbit * groups + i;
            edges[2 * line] = x;
            edges[2 * line + 1] = y;
          }
          ++orbit;
        });
    if (!ok) continue;
    if (!connected_adj(nodes, degree, A.nb.data())) continue;
    return GG_OK;
  }
  return set_error(GG_ERR_INVALID, "symmetric graph construction did not converge");
}

int gg_random_grid(int width, int height, int degree, int length, uint64_t seed, int32_t *edges) {
  const int n = width * height;
  if (width < 1 || height < 1 || degree < 1 || length < 1 || ((int64_t)n * degree) % 2)
    return set_error(GG_ERR_INVALID, "bad grid parameters");
  // Candidate partners within Manhattan distance `length` (grid/common.c:44-51).
  std::vector<std::vector<int32_t>> cand(n);
  for (int v = 0; v < n; ++v) {
    int x = v / height, y = v % height;
    for (int dx = -length; dx <= length; ++dx)
      for (int dy = -length; dy <= length; ++dy) {
        if (!dx && !dy) continue;
        if (std::abs(dx) + std::abs(dy) > length) continue;
        int X = x + dx, Y = y + dy;
        if (X < 0 || X >= width || Y < 0 || Y >= height) continue;
        cand[v].push_back(X * height + Y);
      }
    if ((int)cand[v].size() < degree)
      return set_error(GG_ERR_INVALID, "vertex %d has fewer than %d partners within %d", v, degree, length);
  }
  const int lines = n * degree / 2;
  Rng rng(seed);
  for (int attempt = 0; attempt < 200; ++attempt) {
    Adj A(n, degree);
    // Most-constrained-first greedy: slack(v) = admissible partners - free stubs.
    auto admissible = [&](int v, int u) { return A.deg[u] < degree && !A.has(v, u); };
    auto slack = [&](int v) {
      int c = 0;
      for (int u : cand[v]) c += admissible(v, u);
      return c - (degree - A.deg[v]);
    };
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    bool stuck = false;
    int placed = 0;
    while (placed < lines) {
      int best = -1, bs = 1 << 30;
      int ties = 0;
      for (int v = 0; v < n; ++v) {
        if (A.deg[v] == degree) continue;
        int s = slack(v);
        if (s < bs) {
          bs = s;
          best = v;
          ties = 1;
        } else if (s == bs && rng.below(++ties) == 0) {
          best = v;
        }
      }
      if (best < 0 || bs < 0) {
        stuck = true;
        break;
      }
      int pick = -1, ps = 1 << 30;
      ties = 0;
      for (int u : cand[best]) {
        if (!admissible(best, u)) continue;
        int s = slack(u);
        if (s < ps) {
          ps = s;
          pick = u;
          ties = 1;
        } else if (s == ps && rng.below(++ties) == 0) {
          pick = u;
        }
      }
      if (pick < 0) {
        stuck = true;
        break;
      }
      A.add(best, pick);
      edges[2 * placed] = best;
      edges[2 * placed + 1] = pick;
      ++placed;
    }
    if (stuck) continue;
    // Randomise with length-respecting 2-opt switches, then require connectivity.
    auto dist = [&](int a, int b) {
      return std::abs(a / height - b / height) + std::abs(a % height - b % height);
    };
    auto try_switch = [&]() {
      int e0 = (int)rng.below(lines), e1 = (int)rng.below(lines);
      if (e0 == e1) return;
      int a = edges[2 * e0], b = edges[2 * e0 + 1], c = edges[2 * e1], d = edges[2 * e1 + 1];
      if (a == c || a == d || b == c || b == d) return;
      if (rng.below(2)) std::swap(c, d);
      // (a,b),(c,d) -> (a,c),(b,d)
      if (dist(a, c) > length || dist(b, d) > length) return;
      if (A.has(a, c) || A.has(b, d)) return;
      A.remove(a, b);
      A.remove(c, d);
      A.add(a, c);
      A.add(b, d);
      edges[2 * e0] = a;
      edges[2 * e0 + 1] = c;
      edges[2 * e1] = b;
      edges[2 * e1 + 1] = d;
    };
    for (int k = 0; k < 4 * lines; ++k) try_switch();
    bool conn = false;
    for (int round = 0; round < 200 && !conn; ++round) {
      std::vector<int32_t> adj((size_t)n * degree);
      if (gg_edges_to_adj(n, lines, degree, edges, adj.data()) != GG_OK) return GG_ERR_INVALID;
      conn = connected_adj(n, degree, adj.data());
      if (!conn)
        for (int k = 0; k < lines; ++k) try_switch();
    }
    if (conn) return GG_OK;
  }
  return set_error(GG_ERR_INVALID, "grid graph construction did not converge");
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Plain simulated annealing for general graphs (groups == 1).
//
// Follows general/sa.c:413-520 step for step -- the same libc random() draws
// in the same order (getRandom general/common.c:44-47, uniform_rand
// general/sa.c:36-39), the same 2-opt proposal (general/sa.c:217-369), edge
// orientation rule (order(), general/common.c:72-87), accept rule
// (general/sa.c:371-411) and restore-on-reject -- so that, given the same
// seed and start graph, it visits the same graphs as the reference.  Only the
// evaluator differs (the device kernels, or any plugged-in function).

namespace {

enum { kRight = 0, kLeft = 1, kMiddle = 2 };

using gg::ref_random_below;
using gg::ref_uniform;

// Orientation of an edge on the ring of vertex ids (no added centres).
int ring_orientation(int n, int a, int b) {
  if (n % 2 == 0 && (a - b) % (n / 2) == 0) return kMiddle;
  const double half = n / 2.0;
  if (a < half) {
    if (a > b) return kLeft;
    return (a + half > b) ? kRight : kLeft;
  }
  if (a < b) return kRight;
  return (a - half > b) ? kRight : kLeft;
}

// general/main.c:244-265 (Moore bound).
void moore(int nodes, int degree, int *low_diam, double *low_aspl) {
  gg::moore_bound(nodes, degree, low_diam, low_aspl);
}

struct Move {  // what a 2-opt changed, for restore
  int line[2];
  int old_edge[4];
  int ay[4], ax[4], aval[4];
};

struct GeneralState {
  int n, d, lines;
  int32_t *edges;
  std::vector<int32_t> adj;
  EdgeSet set;
  GeneralState(int n_, int d_, int32_t *e)
      : n(n_), d(d_), lines(n_ * d_ / 2), edges(e), adj((size_t)n_ * d_), set((size_t)n_ * d_ / 2) {
    gg_edges_to_adj(n, lines, d, edges, adj.data());
    for (int i = 0; i < lines; ++i) set.add(ekey(edges[2 * i], edges[2 * i + 1]));
  }
  int find(int y, int v) const {
    for (int x = 0; x < d; ++x)
      if (adj[(size_t)y * d + x] == v) return x;
    return -1;
  }
  // exchange_edge_2opt with groups == 1 (general/sa.c:217-369).
  int propose(Move &m) {
    int t0, t1, te[2][2], r;
    while (true) {
      while (true) {
        do {
          t0 = ref_random_below(lines);
          t1 = ref_random_below(lines);
        } while (t0 == t1);
        int a = edges[2 * t0], b = edges[2 * t0 + 1], c = edges[2 * t1], e = edges[2 * t1 + 1];
        if (a == c || b == e || a == e || b == c) continue;  // has_duplicated_vertex
        break;
      }
      te[0][0] = edges[2 * t0];
      te[0][1] = edges[2 * t0 + 1];
      te[1][0] = edges[2 * t1];
      te[1][1] = edges[2 * t1 + 1];
      r = ref_random_below(2);
      if (r == 0) std::swap(te[0][1], te[1][1]);
      else std::swap(te[0][1], te[1][0]);
      // check_duplicate_current_edge: the new edges must not exist elsewhere.
      if (set.has(ekey(te[0][0], te[0][1])) || set.has(ekey(te[1][0], te[1][1]))) continue;
      break;
    }
    for (int i = 0; i < 2; ++i)
      if (ring_orientation(n, te[i][0], te[i][1]) == kRight) std::swap(te[i][0], te[i][1]);
    const int y0 = edges[2 * t0], y1 = edges[2 * t0 + 1], y2 = edges[2 * t1], y3 = edges[2 * t1 + 1];
    const int x0 = find(y0, y1), x1 = find(y1, y0), x2 = find(y2, y3), x3 = find(y3, y2);
    if (x0 < 0 || x1 < 0 || x2 < 0 || x3 < 0)
      return set_error(GG_ERR_STATE, "adjacency and edge list disagree");
    const int ys[4] = {y0, y1, y2, y3}, xs[4] = {x0, x1, x2, x3};
    for (int k = 0; k < 4; ++k) {
      m.ay[k] = ys[k];
      m.ax[k] = xs[k];
      m.aval[k] = adj[(size_t)ys[k] * d + xs[k]];
    }
    m.line[0] = t0;
    m.line[1] = t1;
    m.old_edge[0] = y0;
    m.old_edge[1] = y1;
    m.old_edge[2] = y2;
    m.old_edge[3] = y3;
    if (r == 0) {
      adj[(size_t)y0 * d + x0] = y3;
      adj[(size_t)y1 * d + x1] = y2;
      adj[(size_t)y2 * d + x2] = y1;
      adj[(size_t)y3 * d + x3] = y0;
    } else {
      adj[(size_t)y0 * d + x0] = y2;
      adj[(size_t)y1 * d + x1] = y3;
      adj[(size_t)y2 * d + x2] = y0;
      adj[(size_t)y3 * d + x3] = y1;
    }
    set.remove(ekey(y0, y1));
    set.remove(ekey(y2, y3));
    set.add(ekey(te[0][0], te[0][1]));
    set.add(ekey(te[1][0], te[1][1]));
    edges[2 * t0] = te[0][0];
    edges[2 * t0 + 1] = te[0][1];
    edges[2 * t1] = te[1][0];
    edges[2 * t1 + 1] = te[1][1];
    return GG_OK;
  }
  void restore(const Move &m) {
    for (int k = 0; k < 4; ++k) adj[(size_t)m.ay[k] * d + m.ax[k]] = m.aval[k];
    for (int i = 0; i < 2; ++i) {
      int t = m.line[i];
      set.remove(ekey(edges[2 * t], edges[2 * t + 1]));
      edges[2 * t] = m.old_edge[2 * i];
      edges[2 * t + 1] = m.old_edge[2 * i + 1];
      set.add(ekey(edges[2 * t], edges[2 * t + 1]));
    }
  }
};

int device_eval(void *, int nodes, int degree, const int32_t *adj, int32_t *diameter, uint64_t *sum,
                int32_t *connected) {
  gg_graph g = {nodes, degree, nodes, 1, 0, GG_GRAPH_SYMMETRIC, adj, nullptr};
  gg_result r;
  int rc = gg_apsp(&g, &r);
  if (rc) return rc;
  *diameter = r.diameter;
  *sum = r.sum;
  *connected = r.connected;
  return GG_OK;
}

}  // namespace

namespace {
// sa() (general/sa.c:413-520); seed_rng = false continues the libc random()
// stream where the caller left it.
int sa_general_run(int nodes, int degree, int32_t *edges, const gg_sa_params *P, gg_eval_fn eval,
                   void *ctx, gg_sa_stats *stats, int32_t *log_diam, uint64_t *log_sum,
                   bool seed_rng) {
  if (!edges || !P || nodes < 3 || degree < 1 || ((int64_t)nodes * degree) % 2)
    return set_error(GG_ERR_INVALID, "bad SA arguments");
  if (P->incremental && eval)
    return set_error(GG_ERR_INVALID, "incremental scoring runs on the device evaluator only");
  if (!eval) {
    // Bring the device up before srandom(): its runtime's start-up may draw
    // from the same libc generator.
    const int rc = gg_init(-1);
    if (rc) return rc;
    eval = device_eval;
  }
  const double t_start = now_s();
  double t_eval = 0;
  GeneralState S(nodes, degree, edges);
  const int lines = S.lines;
  std::vector<int32_t> best(edges, edges + 2 * (size_t)lines);
  int64_t logn = 0, evals = 0;
  // Incremental scorer: the same exact (diameter, S) per proposal, from the
  // sources the move can affect (include/gg_incr.h).
  struct IncrGuard {
    gg_incr *p = nullptr;
    ~IncrGuard() { gg_incr_destroy(p); }
  } inc;
  if (P->incremental) {
    inc.p = gg_incr_create(nodes, degree, S.adj.data());
    if (!inc.p) return GG_ERR_INVALID;
  }
  // m: the proposal just applied to S (nullptr: the current graph).
  auto evaluate = [&](int *diam, double *aspl, const Move *m) -> int {
    int32_t dm = 0, conn = 0;
    uint64_t sum = 0;
    double t0 = now_s();
    int rc;
    if (inc.p) {
      gg_result r{};
      if (!m) {
        rc = gg_incr_state(inc.p, &r);
      } else {
        const int32_t rem[4] = {m->old_edge[0], m->old_edge[1], m->old_edge[2], m->old_edge[3]};
        const int32_t add[4] = {edges[2 * m->line[0]], edges[2 * m->line[0] + 1],
                                edges[2 * m->line[1]], edges[2 * m->line[1] + 1]};
        rc = gg_incr_eval(inc.p, rem, add, &r, nullptr);
      }
      dm = r.diameter;
      sum = r.sum;
      conn = r.connected;
    } else {
      rc = eval(ctx, nodes, degree, S.adj.data(), &dm, &sum, &conn);
    }
    t_eval += now_s() - t0;
    if (rc) return rc;
    ++evals;
    if (logn < P->log_capacity) {
      if (log_diam) log_diam[logn] = dm;
      if (log_sum) log_sum[logn] = sum;
      ++logn;
    }
    const bool ok = conn && dm < nodes;  // general/evaluation.c:254
    *diam = ok ? dm : 2147483647;
    *aspl = ok ? (double)sum / (((double)nodes - 1) * nodes) : 1.7976931348623157e308;
    return ok ? 1 : 0;
  };
  int low_diam;
  double low_aspl;
  moore(nodes, degree, &low_diam, &low_aspl);
  const double max_t = P->hill_climbing ? 0.0 : P->max_temp;
  const double min_t = P->hill_climbing ? 0.0 : P->min_temp;
  const int cycle = P->cooling_cycle > 0 ? P->cooling_cycle : 1;
  const double rate = P->hill_climbing ? 1.0
                      : (max_t != min_t) ? std::pow(min_t / max_t, (double)cycle / P->ncalcs)
                                         : 1.0;  // general/main.c:436-442
  double temp = max_t;
  if (seed_rng) srandom(P->seed);
  int cur_diam = 0, best_diam = 0;
  double cur_aspl = 0, best_aspl = 0;
  int rc = evaluate(&cur_diam, &cur_aspl, nullptr);
  if (rc < 0) return rc;
  best_diam = cur_diam;
  best_aspl = cur_aspl;
  int64_t accepts = 0, ii;
  bool restore = false;
  Move mv{};
  for (ii = 0; ii < P->ncalcs; ++ii) {
    int nd;
    double na;
    while (true) {
      if (restore) S.restore(mv);
      if ((rc = S.propose(mv)) != GG_OK) return rc;
      rc = evaluate(&nd, &na, &mv);
      if (rc < 0) return rc;
      if (rc == 1) break;
      restore = true;  // disconnected: undo and propose again
      if (inc.p && (rc = gg_incr_reject(inc.p))) return rc;
    }
    bool acc;
    if (nd < cur_diam) acc = true;
    else if (nd > cur_diam) acc = false;
    else if (na <= cur_aspl) acc = true;
    else if (P->hill_climbing) acc = false;
    else {
      const double diff = ((cur_aspl - na) * nodes * (nodes - 1)) / 1;  // groups == 1
      acc = std::exp(diff / temp) > ref_uniform();
    }
    if (inc.p && (rc = acc ? gg_incr_accept(inc.p) : gg_incr_reject(inc.p))) return rc;
    if (!acc) {
      restore = true;
    } else {
      restore = false;
      ++accepts;
      cur_aspl = na;
      cur_diam = nd;
      if (best_diam > cur_diam || (best_diam == cur_diam && best_aspl > cur_aspl)) {
        std::copy(edges, edges + 2 * (size_t)lines, best.begin());
        best_aspl = cur_aspl;
        best_diam = cur_diam;
      }
      if (best_diam == cur_diam && best_aspl == low_aspl) break;  // general/sa.c:498
    }
    if ((ii + 1) % cycle == 0) temp *= rate;
  }
  std::copy(best.begin(), best.end(), edges);
  if (stats) {
    stats->steps = ii;
    stats->accepts = accepts;
    stats->evaluations = evals;
    stats->best_diameter = best_diam;
    stats->best_aspl = best_aspl;
    stats->seconds_total = now_s() - t_start;
    stats->seconds_eval = t_eval;
  }
  return GG_OK;
}

// A simple d-regular edge list: endpoints in range, full degree everywhere,
// no self-loop, no repeated edge.
int check_simple_regular(int nodes, int degree, const int32_t *edges) {
  const int lines = nodes * degree / 2;
  std::vector<int> deg(nodes, 0);
  EdgeSet set((size_t)lines);
  for (int i = 0; i < lines; ++i) {
    const int a = edges[2 * i], b = edges[2 * i + 1];
    if (a < 0 || a >= nodes || b < 0 || b >= nodes)
      return set_error(GG_ERR_INVALID, "edge %d endpoint out of range", i);
    if (a == b) return set_error(GG_ERR_INVALID, "edge %d is a self-loop", i);
    if (set.has(ekey(a, b))) return set_error(GG_ERR_INVALID, "edge %d repeats an earlier edge", i);
    set.add(ekey(a, b));
    ++deg[a];
    ++deg[b];
  }
  for (int v = 0; v < nodes; ++v)
    if (deg[v] != degree)
      return set_error(GG_ERR_INVALID, "vertex %d has degree %d, not %d", v, deg[v], degree);
  return GG_OK;
}

}  // namespace

extern "C" int gg_sa_general(int nodes, int degree, int32_t *edges, const gg_sa_params *P,
                             gg_eval_fn eval, void *ctx, gg_sa_stats *stats, int32_t *log_diam,
                             uint64_t *log_sum) {
  return sa_general_run(nodes, degree, edges, P, eval, ctx, stats, log_diam, log_sum, true);
}

extern "C" int gg_general_program(int nodes, int degree, int32_t *edges, int halfway,
                                  const gg_sa_params *P, gg_eval_fn eval, void *ctx,
                                  gg_sa_stats *stats, int32_t *log_diam, uint64_t *log_sum) {
  if (!edges || !P || nodes < 3 || degree < 1 || nodes <= degree || ((int64_t)nodes * degree) % 2)
    return set_error(GG_ERR_INVALID, "bad program arguments (nodes %d degree %d)", nodes, degree);
  int rc = check_simple_regular(nodes, degree, edges);
  if (rc) return rc;
  const gg_eval_fn user_eval = eval;
  if (!eval) {
    if ((rc = gg_init(-1))) return rc;  // before srandom(), as in sa_general_run
    eval = device_eval;
  }
  const int lines = nodes * degree / 2;
  auto connected = [&](const int32_t *adj, int *diam, double *aspl) -> int {
    int32_t dm = 0, conn = 0;
    uint64_t sum = 0;
    const int r = eval(ctx, nodes, degree, adj, &dm, &sum, &conn);
    if (r) return r;
    const bool ok = conn && dm < nodes;  // general/evaluation.c:254
    if (diam) *diam = dm;
    if (aspl) *aspl = ok ? (double)sum / (((double)nodes - 1) * nodes) : 0.0;
    return ok ? 1 : 0;
  };
  srandom(P->seed);  // general/main.c:360
  Move mv{};
  if (!halfway) {  // create_symmetric_edge with groups = 1 (general/main.c:175-210)
    GeneralState S(nodes, degree, edges);
    for (int i = 0; i < 100; ++i)
      if ((rc = S.propose(mv))) return rc;
    while (true) {
      if ((rc = S.propose(mv))) return rc;
      if ((rc = connected(S.adj.data(), nullptr, nullptr)) < 0) return rc;
      if (rc == 1) break;
    }
  }
  {  // check_current_edge (general/sa.c:557-572)
    std::vector<int32_t> adj((size_t)nodes * degree);
    if ((rc = gg_edges_to_adj(nodes, lines, degree, edges, adj.data()))) return rc;
    int diam = 0;
    double aspl = 0;
    if ((rc = connected(adj.data(), &diam, &aspl)) < 0) return rc;
    if (rc == 0)
      return set_error(GG_ERR_INVALID, "The input file has a node which is never reached by another node.");
    int low_diam;
    double low_aspl;
    moore(nodes, degree, &low_diam, &low_aspl);
    if (aspl == low_aspl) {
      if (stats) {
        *stats = gg_sa_stats{};
        stats->best_diameter = diam;
        stats->best_aspl = aspl;
      }
      return GG_ALREADY_OPTIMAL;
    }
  }
  {  // estimate_elapse_time's five moves on a copy (general/sa.c:523-554):
     // only their random() draws carry over; their timing evaluations are skipped.
    std::vector<int32_t> copy(edges, edges + 2 * (size_t)lines);
    GeneralState T(nodes, degree, copy.data());
    for (int i = 0; i < 5; ++i)
      if ((rc = T.propose(mv))) return rc;
  }
  return sa_general_run(nodes, degree, edges, P, user_eval, ctx, stats, log_diam, log_sum, false);
}


// ---------------------------------------------------------------------------
// Grid SA (grid/sa.c:283-361): same RNG draws and moves as the reference.

namespace {

// Vertex v of a width x height grid sits at column v / height, row v % height
// (WIDTH/HEIGHT, grid/common.c:53-61).
struct GridGeom {
  int height, width, groups;
  int col(int v) const { return v / height; }
  int row(int v) const { return v % height; }
  int dist(int v, int w) const { return std::abs(col(v) - col(w)) + std::abs(row(v) - row(w)); }
  // ROTATE (grid/common.c:63-86); groups == 4 assumes a square grid.
  int rotate(int v, int deg) const {
    const int w = col(v), h = row(v);
    if (groups == 2) return (width - w - 1) * height + (height - h - 1);
    if (deg == 90) return h * height + (height - w - 1);
    if (deg == 180) return (height - w - 1) * height + (height - h - 1);
    return (height - h - 1) * height + w;
  }
};

bool same_edge(int a0, int a1, int b0, int b1) { return (a0 == b0 && a1 == b1) || (a0 == b1 && a1 == b0); }

struct GridMover {
  int lines, bl;  // lines, lines per group
  GridGeom G;
  int length;
  // check_duplicate_current_edge (grid/common.c:262-307): lines of the base
  // group (or, for a centre 1g move, every line outside the move's orbit)
  // must not already hold one of the new edges.
  bool fresh(const int32_t *e, int count, const int (*te)[2], int skip0, int skip1, bool all_lines,
             int orbit) const {
    const int end = all_lines ? lines : bl;
    for (int i = 0; i < end; ++i) {
      if (all_lines ? (i % bl == orbit) : (i == skip0 || i == skip1)) continue;
      for (int j = 0; j < count; ++j)
        if (same_edge(e[2 * i], e[2 * i + 1], te[j][0], te[j][1])) return false;
    }
    return true;
  }
  // edge_1g_opt (grid/sa.c:34-135).
  bool one_group(int32_t *e, int start) const {
    const int g = G.groups;
    int tl[4], te[4][2], pattern;
    for (int i = 0; i < g; ++i) {
      const int t = start + bl * i;
      tl[i] = t < lines ? t : t - lines;
    }
    if (e[2 * tl[0]] == e[2 * tl[g - 1] + 1] || e[2 * tl[0] + 1] == e[2 * tl[g - 1]]) return false;
    const int s = e[2 * tl[0]], t = e[2 * tl[g / 2] + 1];
    while (true) {
      pattern = ref_random_below(g + 1);
      if (g == 2) {
        if (pattern != g) {
          te[0][0] = s;
          te[1][0] = G.rotate(s, 180);
          te[0][1] = pattern == 0 ? t : G.rotate(t, 180);
          te[1][1] = pattern == 0 ? G.rotate(t, 180) : t;
        } else {
          te[0][0] = s;
          te[0][1] = G.rotate(s, 180);
          te[1][0] = G.rotate(t, 180);
          te[1][1] = t;
        }
      } else {  // g == 4
        if (pattern != g) {
          te[0][0] = s;
          te[1][0] = G.rotate(s, 90);
          te[2][0] = G.rotate(s, 180);
          te[3][0] = G.rotate(s, 270);
          // pattern k pairs rotation r of s with rotation (r - k) of t
          const int rot[4] = {0, 90, 180, 270};
          for (int i = 0; i < 4; ++i) {
            const int k = ((i - pattern) % 4 + 4) % 4;
            te[i][1] = k == 0 ? t : G.rotate(t, rot[k]);
          }
        } else {
          te[0][0] = s;
          te[0][1] = G.rotate(s, 180);
          te[1][0] = G.rotate(s, 90);
          te[1][1] = G.rotate(s, 270);
          te[2][0] = t;
          te[2][1] = G.rotate(t, 180);
          te[3][0] = G.rotate(t, 90);
          te[3][1] = G.rotate(t, 270);
        }
      }
      if (t != te[g / 2][1]) break;
    }
    for (int i = 0; i < g; i += 2)
      if (G.dist(te[i][0], te[i][1]) > length) return false;
    if (!fresh(e, g, te, tl[0] % bl, -1, pattern == g, tl[0] % bl)) return false;
    for (int i = 0; i < g; ++i) {
      e[2 * tl[i]] = te[i][0];
      e[2 * tl[i] + 1] = te[i][1];
    }
    return true;
  }
  // exchange_edge (grid/sa.c:137-221).
  void exchange(int32_t *e) const {
    const int g = G.groups;
    int tl[8];
    while (true) {
      bool done = false, retry = false;
      while (true) {
        tl[0] = ref_random_below(lines);
        tl[1] = ref_random_below(lines);
        if (tl[0] == tl[1]) continue;
        const int a = e[2 * tl[0]], b = e[2 * tl[0] + 1], c = e[2 * tl[1]], d = e[2 * tl[1] + 1];
        if (a == c || b == d || a == d || b == c) continue;  // has_duplicated_vertex
        if ((tl[0] - tl[1]) % bl == 0) {
          if (one_group(e, tl[0])) {
            done = true;
            break;
          }
          continue;
        }
        break;
      }
      if (done) return;
      auto corner_pair = [&](int t) {  // the edge joins a vertex and its 180-degree image
        const int u = e[2 * t], v = e[2 * t + 1];
        return G.col(u) + G.col(v) == G.width - 1 && G.row(u) + G.row(v) == G.height - 1;
      };
      if ((corner_pair(tl[0]) || corner_pair(tl[1])) && g % 2 == 0) {
        if (one_group(e, tl[0])) return;
        retry = true;
      }
      if (retry) continue;
      for (int i = 1; i < g; ++i) {
        tl[2 * i] = tl[0] + bl * i;
        tl[2 * i + 1] = tl[1] + bl * i;
        if (tl[2 * i] >= lines) tl[2 * i] -= lines;
        if (tl[2 * i + 1] >= lines) tl[2 * i + 1] -= lines;
      }
      const int r = ref_random_below(2);
      int te[8][2];
      for (int i = 0; i < g; ++i) {
        for (int j = 0; j < 2; ++j) {
          te[2 * i][j] = e[2 * tl[2 * i] + j];
          te[2 * i + 1][j] = e[2 * tl[2 * i + 1] + j];
        }
        std::swap(te[2 * i][1], te[2 * i + 1][r]);
      }
      if (G.dist(te[0][0], te[0][1]) > length || G.dist(te[1][0], te[1][1]) > length) continue;
      bool ok = true;  // check_duplicate_tmp_edge (grid/common.c:246-260)
      for (int i = 0; i < 2; ++i)
        for (int j = 2; j < 2 * g; ++j)
          if (same_edge(te[i][0], te[i][1], te[j][0], te[j][1])) ok = false;
      if (!ok) continue;
      if (!fresh(e, 2 * g, te, tl[0] % bl, tl[1] % bl, false, -1)) continue;
      for (int i = 0; i < g; ++i)
        for (int j = 0; j < 2; ++j) {
          e[2 * tl[2 * i] + j] = te[2 * i][j];
          e[2 * tl[2 * i + 1] + j] = te[2 * i + 1][j];
        }
      return;
    }
  }
};

int device_eval_grid(void *, int nodes, int degree, int groups, int based_nodes, const int32_t *adj,
                     const int32_t *rotate_hash, int32_t *diameter, uint64_t *sum, int32_t *connected) {
  gg_graph g = {nodes, degree, based_nodes, groups, 0, GG_GRAPH_SYMMETRIC, adj, rotate_hash};
  gg_result r;
  int rc = gg_apsp(&g, &r);
  if (rc) return rc;
  *diameter = r.diameter;
  *sum = r.sum;
  *connected = r.connected;
  return GG_OK;
}

}  // namespace

// lower_bound_of_diam_aspl (grid/main.c:480-529): per vertex, the number of
// vertices within k hops is at most min(Moore count, grid vertices within
// Manhattan distance k * length).  The diameter bound is read from the last
// vertex's profile, as the reference does.
extern "C" int gg_grid_lower_bound(int width, int height, int degree, int length, int *low_diam,
                                   double *low_aspl) {
  if (width < 1 || height < 1 || degree < 2 || length < 1 || !low_diam || !low_aspl)
    return set_error(GG_ERR_INVALID, "bad grid lower-bound arguments");
  const int m = width, n = height, mn = m * n;
  std::vector<long long> moore(2, 0);
  moore[0] = 1;
  moore[1] = degree + 1;
  long long current = degree;
  int ii;
  for (ii = 2;; ++ii) {
    current *= degree - 1;
    moore.push_back(moore[ii - 1] + current);
    if (moore[ii] >= mn) {
      moore[ii] = mn;
      break;
    }
  }
  const int maxhop = std::max((m + n - 2 + (length - 1)) / length, ii);
  moore.resize(maxhop + 1, mn);
  for (int i = ii + 1; i <= maxhop; ++i) moore[i] = mn;
  std::vector<long long> hist(maxhop + 1), mh(maxhop + 1);
  double sum = 0;
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < n; ++j) {
      std::fill(hist.begin(), hist.end(), 0);
      for (int i2 = 0; i2 < m; ++i2)
        for (int j2 = 0; j2 < n; ++j2) hist[(std::abs(i - i2) + std::abs(j - j2) + length - 1) / length]++;
      for (int k = 1; k <= maxhop; ++k) hist[k] += hist[k - 1];
      for (int k = 0; k <= maxhop; ++k) mh[k] = std::min(hist[k], moore[k]);
      for (int k = 1; k <= maxhop; ++k) sum += (double)(mh[k] - mh[k - 1]) * k;
    }
  int dboth = 0;
  while (mh[dboth] != mn) ++dboth;
  *low_diam = dboth;
  *low_aspl = sum / ((double)mn * (mn - 1));
  return GG_OK;
}

namespace {
// sa() (grid/sa.c:283-361); seed_rng = false continues the libc random()
// stream where the caller left it.
int sa_grid_run(const gg_grid_shape *shape, int degree, int32_t *edges, const int32_t *rotate_hash,
                const gg_sa_params *P, double fixed_temp, gg_eval_grid_fn eval, void *ctx,
                gg_sa_stats *stats, int32_t *log_diam, uint64_t *log_sum, bool seed_rng) {
  if (!shape || !edges || !P || degree < 1)
    return set_error(GG_ERR_INVALID, "bad grid SA arguments");
  const int width = shape->width, height = shape->height, groups = shape->groups;
  if (width < 1 || height < 1 || (groups != 1 && groups != 2 && groups != 4) || shape->length < 1)
    return set_error(GG_ERR_INVALID, "bad grid shape %dx%d g=%d r=%d", width, height, groups,
                     shape->length);
  if ((groups == 4 && width != height) || (groups == 2 && height % 2))
    return set_error(GG_ERR_INVALID, "grid %dx%d cannot carry %d-fold symmetry", width, height, groups);
  const int nodes = width * height;
  if (((int64_t)nodes * degree) % 2 || nodes % groups || (nodes * degree / 2) % groups)
    return set_error(GG_ERR_INVALID, "nodes*degree/2 must be a whole multiple of groups");
  const int lines = nodes * degree / 2;
  std::vector<int32_t> ident;
  if (!rotate_hash) {
    if (groups != 1) return set_error(GG_ERR_INVALID, "groups > 1 needs rotate_hash");
    ident.resize(nodes);
    std::iota(ident.begin(), ident.end(), 0);
    rotate_hash = ident.data();
  }
  for (int i = 0; i < 2 * lines; ++i)
    if (edges[i] < 0 || edges[i] >= nodes) return set_error(GG_ERR_INVALID, "edge endpoint out of range");
  if (!eval) {
    const int rc0 = gg_init(-1);  // before srandom(): the runtime's start-up may draw from it
    if (rc0) return rc0;
    eval = device_eval_grid;
  }
  const double t_start = now_s();
  double t_eval = 0;
  GridMover M{lines, lines / groups, GridGeom{height, width, groups}, shape->length};
  std::vector<int32_t> tmp(edges, edges + 2 * (size_t)lines), best(tmp), adj((size_t)nodes * degree);
  int64_t logn = 0, evals = 0;
  auto evaluate = [&](const int32_t *e, int *diam, double *aspl) -> int {
    int rc = gg_edges_to_adj(nodes, lines, degree, e, adj.data());
    if (rc) return rc;
    int32_t dm = 0, conn = 0;
    uint64_t sum = 0;
    const double t0 = now_s();
    rc = eval(ctx, nodes, degree, groups, nodes / groups, adj.data(), rotate_hash, &dm, &sum, &conn);
    t_eval += now_s() - t0;
    if (rc) return rc;
    ++evals;
    if (logn < P->log_capacity) {
      if (log_diam) log_diam[logn] = dm;
      if (log_sum) log_sum[logn] = sum;
      ++logn;
    }
    const bool ok = conn && !(dm > nodes);  // grid/evaluation.c:182-185
    *diam = dm;
    *aspl = ok ? (double)sum / (((double)nodes - 1) * nodes) : 0.0;
    return ok ? 1 : 0;
  };
  int low_diam;
  double low_aspl;
  int rc = gg_grid_lower_bound(width, height, degree, shape->length, &low_diam, &low_aspl);
  if (rc) return rc;
  const bool fixed = fixed_temp > 0 && !P->hill_climbing;
  const double max_t = P->hill_climbing ? 0.0 : P->max_temp;
  const double min_t = P->hill_climbing ? 0.0 : P->min_temp;
  const int cycle = P->cooling_cycle > 0 ? P->cooling_cycle : 1;
  const double rate = P->hill_climbing ? 1.0 : std::pow(min_t / max_t, (double)cycle / P->ncalcs);
  double temp = max_t;
  if (seed_rng) srandom(P->seed);
  int cur_diam = 0, best_diam = 0;
  double cur_aspl = 0, best_aspl = 0;
  rc = evaluate(edges, &cur_diam, &cur_aspl);
  if (rc < 0) return rc;
  if (rc == 0) return set_error(GG_ERR_INVALID, "the start graph is not connected");
  best_diam = cur_diam;
  best_aspl = cur_aspl;
  int64_t accepts = 0, ii;
  double max_diff = -1;  // NOT_N_DEFINED (grid/main.c:597)
  for (ii = 0; ii < P->ncalcs; ++ii) {
    int nd = 0;
    double na = 0;
    while (true) {
      std::copy(edges, edges + 2 * (size_t)lines, tmp.begin());
      M.exchange(tmp.data());
      rc = evaluate(tmp.data(), &nd, &na);
      if (rc < 0) return rc;
      if (rc == 1) break;
    }
    // accept (grid/sa.c:223-268)
    const double diff = ((cur_aspl - na) * nodes * (nodes - 1)) / groups;
    bool acc;
    if (P->detect_temp) {  // every proposal faces the Metropolis test (grid/sa.c:231-233)
      max_diff = std::max(max_diff, -1.0 * diff);
      acc = std::exp(diff / (fixed ? fixed_temp : temp)) > ref_uniform();
    } else if (nd < cur_diam) acc = true;
    else if (nd > cur_diam) acc = false;
    else if (diff >= 0) acc = true;
    else if (P->hill_climbing) acc = false;
    else acc = std::exp(diff / (fixed ? fixed_temp : temp)) > ref_uniform();
    if (acc) {
      ++accepts;
      cur_aspl = na;
      cur_diam = nd;
      std::copy(tmp.begin(), tmp.end(), edges);
      if (best_diam > nd || (best_diam == nd && best_aspl > na)) {
        std::copy(tmp.begin(), tmp.end(), best.begin());
        best_diam = nd;
        best_aspl = na;
      }
      if (best_diam == low_diam && best_aspl == low_aspl) break;
    }
    if ((ii + 1) % cycle == 0) temp *= rate;
  }
  std::copy(best.begin(), best.end(), edges);
  if (stats) {
    stats->steps = ii;
    stats->accepts = accepts;
    stats->evaluations = evals;
    stats->best_diameter = best_diam;
    stats->best_aspl = best_aspl;
    stats->seconds_total = now_s() - t_start;
    stats->seconds_eval = t_eval;
    stats->max_diff_energy = max_diff;
  }
  return GG_OK;
}
}  // namespace

extern "C" int gg_sa_grid(const gg_grid_shape *shape, int degree, int32_t *edges,
                          const int32_t *rotate_hash, const gg_sa_params *P, double fixed_temp,
                          gg_eval_grid_fn eval, void *ctx, gg_sa_stats *stats, int32_t *log_diam,
                          uint64_t *log_sum) {
  return sa_grid_run(shape, degree, edges, rotate_hash, P, fixed_temp, eval, ctx, stats, log_diam,
                     log_sum, true);
}

namespace {
// Vertices not reached from vertex 0 (simple_bfs, grid/main.c:197-230).
int unreached_from_zero(int n, int d, const std::vector<int32_t> &adj) {
  std::vector<char> seen(n, 0);
  std::vector<int> q{0};
  seen[0] = 1;
  for (size_t h = 0; h < q.size(); ++h)
    for (int j = 0; j < d; ++j) {
      const int u = adj[(size_t)q[h] * d + j];
      if (!seen[u]) {
        seen[u] = 1;
        q.push_back(u);
      }
    }
  return n - (int)q.size();
}
}  // namespace

// create_symmetric_edge (grid/main.c:393-445) over a seeded base lattice.
extern "C" int gg_random_grid_symmetric(int width, int height, int degree, int length, int groups,
                                        uint64_t seed, int32_t *edges) {
  if (!edges || (groups != 1 && groups != 2 && groups != 4))
    return set_error(GG_ERR_INVALID, "groups must be 1, 2 or 4");
  if (groups == 1) return gg_random_grid(width, height, degree, length, seed, edges);
  if ((groups == 4 && (width != height || width % 2)) || height % 2)
    return set_error(GG_ERR_INVALID, "grid %dx%d cannot carry %d-fold symmetry", width, height, groups);
  const int bw = groups == 4 ? width / 2 : width, bh = height / 2;
  const int nodes = width * height, lines = nodes * degree / 2, bl = lines / groups;
  if (((int64_t)nodes * degree) % 2 || lines % groups)
    return set_error(GG_ERR_INVALID, "nodes*degree/2 must be a whole multiple of groups");
  std::vector<int32_t> base((size_t)2 * bl);
  int rc = gg_random_grid(bw, bh, degree, length, seed, base.data());
  if (rc) return rc;
  return gg_grid_symmetric_from_base(width, height, degree, length, groups, (uint32_t)seed,
                                     base.data(), edges);
}

namespace {
// create_symmetric_edge (grid/main.c:393-445) on the current libc random()
// stream; base may alias the first lines/groups edges of `edges`.
int symmetric_from_base_run(int width, int height, int degree, int length, int groups,
                            const int32_t *base_in, int32_t *edges) {
  if (!edges || (groups != 2 && groups != 4))
    return set_error(GG_ERR_INVALID, "groups must be 2 or 4");
  if ((groups == 4 && (width != height || width % 2)) || height % 2)
    return set_error(GG_ERR_INVALID, "grid %dx%d cannot carry %d-fold symmetry", width, height, groups);
  const int bw = groups == 4 ? width / 2 : width, bh = height / 2;
  const int nodes = width * height, lines = nodes * degree / 2, bl = lines / groups;
  if (((int64_t)nodes * degree) % 2 || lines % groups || length < 1)
    return set_error(GG_ERR_INVALID, "nodes*degree/2 must be a whole multiple of groups");
  if (!base_in) return set_error(GG_ERR_INVALID, "no base edges");
  const std::vector<int32_t> base(base_in, base_in + 2 * (size_t)bl);
  for (int i = 0; i < 2 * bl; ++i)
    if (base[i] < 0 || base[i] >= bw * bh)
      return set_error(GG_ERR_INVALID, "base edge endpoint %d outside the %dx%d base grid", base[i], bw, bh);
  GridGeom G{height, width, groups};
  for (int i = 0; i < 2 * bl; ++i) edges[i] = (base[i] / bh) * height + base[i] % bh;
  for (int i = 0; i < bl; ++i)
    for (int j = 0; j < 2; ++j) {
      const int v = edges[2 * i + j];
      if (groups == 2) {
        edges[2 * (bl + i) + j] = G.rotate(v, 180);
      } else {
        edges[2 * (bl + i) + j] = G.rotate(v, 90);
        edges[2 * (2 * bl + i) + j] = G.rotate(v, 180);
        edges[2 * (3 * bl + i) + j] = G.rotate(v, 270);
      }
    }
  int rc = GG_OK;
  GridMover M{lines, bl, G, length};
  for (int64_t i = 0; i < (int64_t)lines * 10; ++i) M.exchange(edges);  // INITIAL_TIMES
  std::vector<int32_t> adj((size_t)nodes * degree), tmp(edges, edges + 2 * (size_t)lines);
  if ((rc = gg_edges_to_adj(nodes, lines, degree, edges, adj.data()))) return rc;
  int min_num = unreached_from_zero(nodes, degree, adj);
  for (int64_t tries = 0; min_num != 0; ++tries) {
    if (tries > 1000000) return set_error(GG_ERR_INVALID, "could not connect the symmetric grid graph");
    std::copy(edges, edges + 2 * (size_t)lines, tmp.begin());
    M.exchange(tmp.data());
    if ((rc = gg_edges_to_adj(nodes, lines, degree, tmp.data(), adj.data()))) return rc;
    const int num = unreached_from_zero(nodes, degree, adj);
    if (num <= min_num) {
      min_num = num;
      std::copy(tmp.begin(), tmp.end(), edges);
    }
  }
  return GG_OK;
}

// create_lattice (grid/main.c:232-330) on a width x height grid: `degree`
// parallel edges between columns 2x and 2x+1 (an odd last column pairs rows
// 2y and 2y+1; an odd last vertex takes degree/2 self-loops), lines*10
// length-limited swaps (simple_exchange_edge, grid/main.c:120-143), then swaps
// on a copy kept while they do not raise the number of vertices unreached
// from vertex 0 (always at least one swap, until none is unreached), then of
// self-loops, then of repeated edges.  Draws random() % lines as the
// reference does.
int lattice_run(int width, int height, int degree, int length, int32_t *e) {
  const int nodes = width * height, lines = nodes * degree / 2;
  int i = 0;
  for (int x = 0; x < width / 2; ++x)
    for (int y = 0; y < height; ++y)
      for (int k = 0; k < degree; ++k, ++i) {
        e[2 * i] = y + 2 * x * height;
        e[2 * i + 1] = e[2 * i] + height;
      }
  if (width % 2) {
    for (int y = 0; y < height / 2; ++y)
      for (int k = 0; k < degree; ++k, ++i) {
        e[2 * i] = (width - 1) * height + 2 * y;
        e[2 * i + 1] = e[2 * i] + 1;
      }
    if (height % 2)
      for (int k = 0; k < degree / 2; ++k, ++i) e[2 * i] = e[2 * i + 1] = nodes - 1;
  }
  if (i != lines) return set_error(GG_ERR_INVALID, "a %dx%d lattice of degree %d does not close", width, height, degree);
  GridGeom G{height, width, 1};
  auto swap = [&](int32_t *f) {
    while (true) {
      int e1, e2;
      do {
        e1 = (int)(random() % lines);
        e2 = (int)(random() % lines);
      } while (e1 == e2);
      const int v1 = f[2 * e1], w1 = f[2 * e1 + 1], v2 = f[2 * e2], w2 = f[2 * e2 + 1];
      if (G.dist(v1, v2) <= length && G.dist(w1, w2) <= length) {
        f[2 * e1 + 1] = v2;
        f[2 * e2] = w1;
      } else if (G.dist(v1, w2) <= length && G.dist(w1, v2) <= length) {
        f[2 * e1 + 1] = w2;
        f[2 * e2] = w1;
        f[2 * e2 + 1] = v2;
      } else {
        continue;
      }
      return;
    }
  };
  for (int64_t t = 0; t < (int64_t)lines * 10; ++t) swap(e);  // INITIAL_TIMES
  std::vector<int32_t> adj((size_t)nodes * degree), tmp(e, e + 2 * (size_t)lines);
  int rc;
  auto unreached = [&](const int32_t *f, int *num) -> int {
    int r = gg_edges_to_adj(nodes, lines, degree, f, adj.data());
    if (r) return r;
    *num = unreached_from_zero(nodes, degree, adj);
    return GG_OK;
  };
  auto loops = [&](const int32_t *f, int *num) -> int {
    int c = 0;
    for (int l = 0; l < lines; ++l) c += f[2 * l] == f[2 * l + 1];
    *num = c;
    return GG_OK;
  };
  auto repeats = [&](const int32_t *f, int *num) -> int {  // pairs of equal lines
    std::vector<uint64_t> keys;
    keys.reserve(lines);
    for (int l = 0; l < lines; ++l) keys.push_back(ekey(f[2 * l], f[2 * l + 1]));
    std::sort(keys.begin(), keys.end());
    long long c = 0;
    for (size_t a = 0; a < keys.size();) {
      size_t b = a;
      while (b < keys.size() && keys[b] == keys[a]) ++b;
      c += (long long)(b - a) * (b - a - 1) / 2;
      a = b;
    }
    *num = (int)std::min<long long>(c, 1 << 30);
    return GG_OK;
  };
  // The connectivity pass runs at least once; the other two only when needed.
  auto reduce = [&](auto count, bool always) -> int {
    int min_num;
    int r = count(e, &min_num);
    if (r) return r;
    if (!always && min_num == 0) return GG_OK;
    for (int64_t tries = 0;; ++tries) {
      if (tries > 10000000) return set_error(GG_ERR_INVALID, "lattice construction did not converge");
      std::copy(e, e + 2 * (size_t)lines, tmp.begin());
      swap(tmp.data());
      int num;
      if ((r = count(tmp.data(), &num))) return r;
      if (num == 0) {
        std::copy(tmp.begin(), tmp.end(), e);
        return GG_OK;
      }
      if (num <= min_num) {
        min_num = num;
        std::copy(tmp.begin(), tmp.end(), e);
      }
    }
  };
  if ((rc = reduce(unreached, true))) return rc;
  if ((rc = reduce(loops, false))) return rc;
  return reduce(repeats, false);
}
}  // namespace

extern "C" int gg_grid_symmetric_from_base(int width, int height, int degree, int length,
                                           int groups, uint32_t seed, const int32_t *base,
                                           int32_t *edges) {
  srandom(seed);
  return symmetric_from_base_run(width, height, degree, length, groups, base, edges);
}

extern "C" int gg_grid_lattice(int width, int height, int degree, int length, uint32_t seed,
                               int32_t *edges) {
  if (!edges || width < 1 || height < 1 || degree < 1 || length < 1 ||
      ((int64_t)width * height * degree) % 2)
    return set_error(GG_ERR_INVALID, "bad lattice arguments");
  srandom(seed);
  return lattice_run(width, height, degree, length, edges);
}

extern "C" int gg_grid_program(const gg_grid_shape *shape, int degree, int32_t *edges, int source,
                               const int32_t *rotate_hash, const gg_sa_params *P, double fixed_temp,
                               gg_eval_grid_fn eval, void *ctx, gg_sa_stats *stats,
                               int32_t *log_diam, uint64_t *log_sum) {
  if (!shape || !edges || !P || degree < 1 || source < GG_GRID_LATTICE || source > GG_GRID_FULL_FILE)
    return set_error(GG_ERR_INVALID, "bad grid program arguments");
  const int width = shape->width, height = shape->height, g = shape->groups, length = shape->length;
  if (width < 1 || height < 1 || (g != 1 && g != 2 && g != 4) || length < 1)
    return set_error(GG_ERR_INVALID, "bad grid shape %dx%d g=%d r=%d", width, height, g, length);
  const int bw = g == 4 ? width / 2 : width, bh = g == 1 ? height : height / 2;
  const int nodes = width * height;
  // grid/main.c:684-695
  if (g == 4 && bw != bh)
    return set_error(GG_ERR_INVALID, "When g = 4, width(%d) must be equal to height(%d).", bw, bh);
  if (g == 4 && width % 2 && height % 2)
    return set_error(GG_ERR_INVALID, "When g = 4, width(%d) and height(%d) are divisible by 2.", width, height);
  if (g == 2 && height % 2)
    return set_error(GG_ERR_INVALID, "When g = 2, height(%d) os divisible by 2.", height);
  if (((int64_t)nodes * degree) % 2 || nodes % g || (nodes * degree / 2) % g)
    return set_error(GG_ERR_INVALID, "(nodes*degree/2) must be divisible by groups(%d)", g);
  const int lines = nodes * degree / 2;
  const gg_eval_grid_fn user_eval = eval;
  if (!eval) {
    const int rc0 = gg_init(-1);  // before srandom(), as in sa_grid_run
    if (rc0) return rc0;
    eval = device_eval_grid;
  }
  std::vector<int32_t> ident;
  if (!rotate_hash) {
    if (g != 1) return set_error(GG_ERR_INVALID, "groups > 1 needs rotate_hash");
    ident.resize(nodes);
    std::iota(ident.begin(), ident.end(), 0);
    rotate_hash = ident.data();
  }
  srandom(P->seed);  // grid/main.c:630
  int rc;
  if (source == GG_GRID_LATTICE && (rc = lattice_run(bw, bh, degree, length, edges))) return rc;
  if (source != GG_GRID_FULL_FILE && g != 1 &&
      (rc = symmetric_from_base_run(width, height, degree, length, g, edges, edges)))
    return rc;
  // verfy_graph (grid/main.c:446-472)
  std::vector<int> deg(nodes, 0);
  GridGeom G{height, width, g};
  for (int i = 0; i < lines; ++i) {
    const int a = edges[2 * i], b = edges[2 * i + 1];
    if (a < 0 || a >= nodes || b < 0 || b >= nodes)
      return set_error(GG_ERR_INVALID, "edge %d endpoint out of range", i);
    ++deg[a];
    ++deg[b];
    if (G.dist(a, b) > length)
      return set_error(GG_ERR_INVALID, "Over length in line %d: length = %d, distance = %d", i + 1,
                       length, G.dist(a, b));
  }
  for (int v = 0; v < nodes; ++v)
    if (deg[v] != degree)
      return set_error(GG_ERR_INVALID, "Not regular graph. degree = %d n[%d] = %d.", degree, v, deg[v]);
  {  // check_current_edge (grid/sa.c:378-398)
    std::vector<int32_t> adj((size_t)nodes * degree);
    if ((rc = gg_edges_to_adj(nodes, lines, degree, edges, adj.data()))) return rc;
    int32_t dm = 0, conn = 0;
    uint64_t sum = 0;
    if ((rc = eval(ctx, nodes, degree, g, nodes / g, adj.data(), rotate_hash, &dm, &sum, &conn)))
      return rc;
    if (!conn || dm > nodes)
      return set_error(GG_ERR_INVALID, "The input file has a node which is never reached by another node.");
    const double aspl = (double)sum / (((double)nodes - 1) * nodes);
    int low_diam;
    double low_aspl;
    if ((rc = gg_grid_lower_bound(width, height, degree, length, &low_diam, &low_aspl))) return rc;
    if (dm == low_diam && aspl == low_aspl) {
      if (stats) {
        *stats = gg_sa_stats{};
        stats->best_diameter = dm;
        stats->best_aspl = aspl;
      }
      return GG_ALREADY_OPTIMAL;
    }
  }
  {  // estimated_elapse_time's five moves, each on a fresh copy (grid/sa.c:350-376):
     // only their random() draws carry over; their timing evaluations are skipped.
    GridMover M{lines, lines / g, G, length};
    std::vector<int32_t> tmp(2 * (size_t)lines);
    for (int i = 0; i < 5; ++i) {
      std::copy(edges, edges + 2 * (size_t)lines, tmp.begin());
      M.exchange(tmp.data());
    }
  }
  return sa_grid_run(shape, degree, edges, rotate_hash, P, fixed_temp, user_eval, ctx, stats, log_diam,
                     log_sum, false);
}
