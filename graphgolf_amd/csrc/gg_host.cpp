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

int ref_random_below(int max) { return (int)(random() * ((double)max) / (1.0 + RAND_MAX)); }
double ref_uniform() { return ((double)random() + 1.0) / ((double)RAND_MAX + 2.0); }

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
  int diam = -1, n = 1, r = 1;
  double aspl = 0.0;
  while (true) {
    int tmp = n + degree * (int)std::pow(degree - 1, r - 1);
    if (tmp >= nodes) break;
    n = tmp;
    aspl += r * degree * std::pow(degree - 1, r - 1);
    diam = r++;
  }
  diam++;
  aspl += diam * (nodes - n);
  aspl /= (nodes - 1);
  *low_diam = diam;
  *low_aspl = aspl;
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

extern "C" int gg_sa_general(int nodes, int degree, int32_t *edges, const gg_sa_params *P,
                             gg_eval_fn eval, void *ctx, gg_sa_stats *stats, int32_t *log_diam,
                             uint64_t *log_sum) {
  if (!edges || !P || nodes < 3 || degree < 1 || ((int64_t)nodes * degree) % 2)
    return set_error(GG_ERR_INVALID, "bad SA arguments");
  if (!eval) eval = device_eval;
  const double t_start = now_s();
  double t_eval = 0;
  GeneralState S(nodes, degree, edges);
  const int lines = S.lines;
  std::vector<int32_t> best(edges, edges + 2 * (size_t)lines);
  int64_t logn = 0, evals = 0;
  auto evaluate = [&](int *diam, double *aspl) -> int {
    int32_t dm = 0, conn = 0;
    uint64_t sum = 0;
    double t0 = now_s();
    int rc = eval(ctx, nodes, degree, S.adj.data(), &dm, &sum, &conn);
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
  srandom(P->seed);
  int cur_diam, best_diam;
  double cur_aspl, best_aspl;
  int rc = evaluate(&cur_diam, &cur_aspl);
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
      rc = evaluate(&nd, &na);
      if (rc < 0) return rc;
      if (rc == 1) break;
      restore = true;  // disconnected: undo and propose again
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

