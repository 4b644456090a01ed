// gg_sa_sym.cpp -- the general SA and program for rotationally symmetric
// graphs (-g groups) with added centres (-v), C ABI in include/gg_host.h.
//
// Follows general/sa.c:217-520 and general/common.c:44-276 step for step,
// with libc random() after srandom(seed), so that a run visits the same
// graphs as the reference's:
//   * a proposal draws two lines; lines of one orbit (t0 - t1 a multiple of
//     the base line count) or a line of length (n - c)/2 with even groups
//     take the 1g move (edge_1g_opt): the orbit of one edge is re-paired
//     inside itself by a random rotation pattern; otherwise the 2g move
//     swaps endpoints of the two lines in every group (exchange_edge_2opt);
//   * the duplicate checks scan the base lines only (all lines for the
//     length-(n-c)/2 pattern), as the reference does;
//   * rejected and disconnected proposals are undone from the recorded
//     adjacency entries and lines (restore_adj, restore_edge);
//   * accept: smaller diameter wins; equal diameter accepts a non-increasing
//     ASPL, else exp(dS/g / T) against uniform_rand().
// The program entry adds what general/main.c runs before sa(): the symmetric
// copies of the base file (create_symmetric_edge, with 100 random moves and
// moves until connected), or the centre edges (create_symmetric_edge_with_
// vertexes), and the draws of the five timing moves.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <vector>

#include "gg_eval.h"
#include "gg_host.h"
#include "gg_internal.h"
#include "gg_sa_common.h"

using gg::now_s;
using gg::ref_random_below;
using gg::ref_uniform;
using gg::set_error;

namespace {

constexpr int kRight = 0, kLeft = 1, kMiddle = 2;  // general/common.h:31-33

bool same_edge(int a0, int a1, int b0, int b1) { return (a0 == b0 && a1 == b1) || (a0 == b1 && a1 == b0); }

struct SymState {
  int n, d, g, c, lines, bl, bn;
  bool simple = true;  // false (-N): moves leave the adjacency to be rebuilt
  int32_t *edges;
  std::vector<int32_t> adj;
  // What the last move changed (restore_adj / restore_edge).
  int kind = 0;  // 1: 1g move, 2: 2g move
  std::vector<int> r_line, r_edge, r_y, r_x, r_val;

  SymState(int n_, int d_, int g_, int c_, int32_t *e)
      : n(n_), d(d_), g(g_), c(c_), lines(n_ * d_ / 2), bl(n_ * d_ / 2 / g_), bn((n_ - c_) / g_),
        edges(e), adj((size_t)n_ * d_), r_line(2 * g_), r_edge(4 * g_), r_y(4 * g_), r_x(4 * g_),
        r_val(4 * g_) {}
  int init() { return gg_edges_to_adj(n, lines, d, edges, adj.data()); }

  // distance / order (general/sa.c:77-84, general/common.c:64-79).
  int distance(int a, int b) const {
    if (a >= n - c || b >= n - c) return -1;
    const int v = std::max(a, b) - std::min(a, b);
    const int m = c ? n - c : n;
    return (v < m / 2.0) ? v : m - v;
  }
  int order(int a, int b) const {
    int m = n;
    if (!c && m % 2 == 0 && (a - b) % (m / 2) == 0) return kMiddle;
    if (a >= m - c || b >= m - c) return kMiddle;
    if (c && (m - c) % 2 == 0 && (a - b) % ((m - c) / 2) == 0) return kMiddle;
    if (c) m -= c;
    if (a < m / 2.0) {
      if (a > b) return kLeft;
      return (a + m / 2.0 > b) ? kRight : kLeft;
    }
    if (a < b) return kRight;
    return (a - m / 2.0 > b) ? kRight : kLeft;
  }
  int find(int y, int v) const {
    for (int x = 0; x < d; ++x)
      if (adj[(size_t)y * d + x] == v) return x;
    return -1;
  }
  // check_duplicate_current_edge (general/common.c:119-160).
  bool fresh(const int *tl, int ntl, const int (*te)[2], int g_opt, bool is_center) const {
    if (g_opt == 2) {
      const int l0 = tl[0] % bl, l1 = tl[1] % bl;
      for (int i = 0; i < bl; ++i) {
        if (i == l0 || i == l1) continue;
        for (int j = 0; j < ntl; ++j)
          if (same_edge(edges[2 * i], edges[2 * i + 1], te[j][0], te[j][1])) return false;
      }
      return true;
    }
    const int l0 = tl[0] % bl;
    const int end = is_center ? lines : bl;
    for (int i = 0; i < end; ++i) {
      if (i % bl == l0) continue;
      for (int j = 0; j < ntl; ++j)
        if (same_edge(edges[2 * i], edges[2 * i + 1], te[j][0], te[j][1])) return false;
    }
    return true;
  }
  // edge_1g_opt (general/common.c:162-276); false = draw again.
  int one_g(int start, bool *done) {
    *done = false;
    if (g == 1) {  // the reference reports success without a move
      *done = true;
      return GG_OK;
    }
    if (edges[2 * start] >= n - c || edges[2 * start + 1] >= n - c) return GG_OK;
    std::vector<int> tl(g);
    std::vector<int> te(2 * g);
    for (int i = 0; i < g; ++i) tl[i] = start % bl + i * bl;
    const int s0 = edges[2 * tl[0]] % bn;
    int e0 = s0;  // get_end_edge (general/common.c:49-57)
    for (int i = 0; i < g && e0 == s0; ++i)
      for (int j = 0; j < 2; ++j)
        if (edges[2 * tl[i] + j] % bn != s0) {
          e0 = edges[2 * tl[i] + j] % bn;
          break;
        }
    if (e0 == s0) return GG_OK;
    const int diff = edges[2 * tl[0]] - edges[2 * tl[0] + 1];
    const int m = n - c;
    int pattern;
    while (true) {
      pattern = (g % 2 == 0) ? ref_random_below(g + 1) : ref_random_below(g);
      if (pattern == g) {
        for (int i = 0; i < g / 2; ++i) {
          te[2 * i] = s0 + bn * i;
          te[2 * i + 1] = te[2 * i] + m / 2;
        }
        for (int i = g / 2; i < g; ++i) {
          te[2 * i] = e0 + bn * (i - g / 2);
          te[2 * i + 1] = te[2 * i] + m / 2;
        }
      } else {
        for (int i = 0; i < g; ++i) te[2 * i] = s0 + bn * i;
        te[1] = e0 + bn * pattern;
        for (int i = 1; i < g; ++i) {
          const int t = te[1] + bn * i;
          te[2 * i + 1] = t < m ? t : t - m;
        }
      }
      if (diff != te[0] - te[1]) break;
    }
    if (!fresh(tl.data(), g, (const int(*)[2])te.data(), 1, pattern == g)) return GG_OK;
    for (int i = 0; i < g; ++i)
      if (order(te[2 * i], te[2 * i + 1]) == kRight) std::swap(te[2 * i], te[2 * i + 1]);
    if (!simple) {
      for (int i = 0; i < g; ++i) {
        edges[2 * tl[i]] = te[2 * i];
        edges[2 * tl[i] + 1] = te[2 * i + 1];
      }
      kind = 1;
      *done = true;
      return GG_OK;
    }
    // The row slot of every endpoint's old edge (the hash of the reference).
    std::vector<int> slot(n, -1);
    for (int i = 0; i < g; ++i) {
      const int y0 = edges[2 * tl[i]], y1 = edges[2 * tl[i] + 1];
      const int x0 = find(y0, y1), x1 = find(y1, y0);
      if (x0 < 0 || x1 < 0) return set_error(GG_ERR_STATE, "adjacency and edge list disagree");
      slot[y0] = x0;
      slot[y1] = x1;
    }
    for (int i = 0; i < g; ++i) {
      const int t0 = te[2 * i], t1 = te[2 * i + 1];
      if (slot[t0] < 0 || slot[t1] < 0) return set_error(GG_ERR_STATE, "1g move outside its orbit");
      r_y[2 * i] = t0;
      r_x[2 * i] = slot[t0];
      r_y[2 * i + 1] = t1;
      r_x[2 * i + 1] = slot[t1];
      r_val[2 * i] = adj[(size_t)t0 * d + slot[t0]];
      r_val[2 * i + 1] = adj[(size_t)t1 * d + slot[t1]];
      r_line[i] = tl[i];
      r_edge[2 * i] = edges[2 * tl[i]];
      r_edge[2 * i + 1] = edges[2 * tl[i] + 1];
      adj[(size_t)t0 * d + slot[t0]] = t1;
      adj[(size_t)t1 * d + slot[t1]] = t0;
    }
    for (int i = 0; i < g; ++i) {
      edges[2 * tl[i]] = te[2 * i];
      edges[2 * tl[i] + 1] = te[2 * i + 1];
    }
    kind = 1;
    *done = true;
    return GG_OK;
  }
  // exchange_edge_2opt (general/sa.c:217-369).
  int exchange() {
    std::vector<int> tl(2 * g), te(4 * g);
    int r = 0;
    while (true) {
      int t0, t1;
      while (true) {
        do {
          t0 = ref_random_below(lines);
          t1 = ref_random_below(lines);
        } while (t0 == t1);
        const int a = edges[2 * t0], b = edges[2 * t0 + 1], cc = edges[2 * t1], e = edges[2 * t1 + 1];
        if (a == cc || b == e || a == e || b == cc) continue;  // has_duplicated_vertex
        if ((t0 - t1) % bl == 0) {
          bool done;
          const int rc = one_g(t0, &done);
          if (rc) return rc;
          if (done) return GG_OK;
          continue;
        }
        break;
      }
      const bool f0 = distance(edges[2 * t0], edges[2 * t0 + 1]) == (n - c) / 2;
      const bool f1 = distance(edges[2 * t1], edges[2 * t1 + 1]) == (n - c) / 2;
      if ((f0 || f1) && g % 2 == 0) {
        bool done;
        const int rc = one_g(t0, &done);
        if (rc) return rc;
        if (done) return GG_OK;
        continue;
      }
      tl[0] = t0;
      tl[1] = t1;
      for (int i = 1; i < g; ++i) {
        const int a = t0 + bl * i, b = t1 + bl * i;
        tl[2 * i] = a >= lines ? a - lines : a;
        tl[2 * i + 1] = b >= lines ? b - lines : b;
      }
      for (int i = 0; i < 2 * g; ++i) {
        te[2 * i] = edges[2 * tl[i]];
        te[2 * i + 1] = edges[2 * tl[i] + 1];
      }
      r = ref_random_below(2);
      for (int i = 0; i < g; ++i) {
        if (r == 0) std::swap(te[2 * (2 * i) + 1], te[2 * (2 * i + 1) + 1]);
        else std::swap(te[2 * (2 * i) + 1], te[2 * (2 * i + 1)]);
      }
      bool ok = true;  // check_duplicate_tmp_edge (general/common.c:103-117)
      for (int i = 0; i < 2 && ok; ++i)
        for (int j = 2; j < 2 * g; ++j)
          if (same_edge(te[2 * i], te[2 * i + 1], te[2 * j], te[2 * j + 1])) ok = false;
      if (!ok) continue;
      if (!fresh(tl.data(), 2 * g, (const int(*)[2])te.data(), 2, false)) continue;
      break;
    }
    for (int i = 0; i < 2 * g; ++i)
      if (order(te[2 * i], te[2 * i + 1]) == kRight) std::swap(te[2 * i], te[2 * i + 1]);
    if (!simple) {
      for (int i = 0; i < g; ++i)
        for (int k = 0; k < 2; ++k) {
          edges[2 * tl[2 * i + k]] = te[2 * (2 * i + k)];
          edges[2 * tl[2 * i + k] + 1] = te[2 * (2 * i + k) + 1];
        }
      kind = 2;
      return GG_OK;
    }
    // Slots and restore values from the unchanged adjacency, then the changes.
    std::vector<int> ys(4 * g), xs(4 * g);
    for (int i = 0; i < g; ++i) {
      const int y[4] = {edges[2 * tl[2 * i]], edges[2 * tl[2 * i] + 1], edges[2 * tl[2 * i + 1]],
                        edges[2 * tl[2 * i + 1] + 1]};
      const int x[4] = {find(y[0], y[1]), find(y[1], y[0]), find(y[2], y[3]), find(y[3], y[2])};
      for (int k = 0; k < 4; ++k) {
        if (x[k] < 0) return set_error(GG_ERR_STATE, "adjacency and edge list disagree");
        ys[4 * i + k] = y[k];
        xs[4 * i + k] = x[k];
        r_y[4 * i + k] = y[k];
        r_x[4 * i + k] = x[k];
        r_val[4 * i + k] = adj[(size_t)y[k] * d + x[k]];
      }
      r_line[2 * i] = tl[2 * i];
      r_line[2 * i + 1] = tl[2 * i + 1];
      r_edge[4 * i] = y[0];
      r_edge[4 * i + 1] = y[1];
      r_edge[4 * i + 2] = y[2];
      r_edge[4 * i + 3] = y[3];
    }
    for (int i = 0; i < g; ++i) {
      const int *y = &ys[4 * i], *x = &xs[4 * i];
      const int nv[4] = {r == 0 ? y[3] : y[2], r == 0 ? y[2] : y[3], r == 0 ? y[1] : y[0],
                         r == 0 ? y[0] : y[1]};
      for (int k = 0; k < 4; ++k) adj[(size_t)y[k] * d + x[k]] = nv[k];
    }
    for (int i = 0; i < g; ++i)
      for (int k = 0; k < 2; ++k) {
        edges[2 * tl[2 * i + k]] = te[2 * (2 * i + k)];
        edges[2 * tl[2 * i + k] + 1] = te[2 * (2 * i + k) + 1];
      }
    kind = 2;
    return GG_OK;
  }
  void restore() {
    for (int i = 0; i < kind * g * 2; ++i) adj[(size_t)r_y[i] * d + r_x[i]] = r_val[i];
    for (int i = 0; i < g * kind; ++i) {
      edges[2 * r_line[i]] = r_edge[2 * i];
      edges[2 * r_line[i] + 1] = r_edge[2 * i + 1];
    }
  }
};

int device_eval_sym(void *, int nodes, int degree, int based_nodes, int groups, int added_centers,
                    const int32_t *adj, int32_t *diameter, uint64_t *sum, int32_t *connected) {
  gg_graph gr = {nodes, degree, based_nodes, groups, added_centers, GG_GRAPH_SYMMETRIC, adj, nullptr};
  gg_result r;
  const int rc = gg_apsp(&gr, &r);
  if (rc) return rc;
  *diameter = r.diameter;
  *sum = r.sum;
  *connected = r.connected;
  return GG_OK;
}

int check_shape(const gg_general_shape *s) {
  if (!s || s->nodes < 3 || s->degree < 1 || s->groups < 1 || s->added_centers < 0 ||
      ((int64_t)s->nodes * s->degree) % 2 || (s->nodes - s->added_centers) % s->groups ||
      (s->nodes * s->degree / 2) % s->groups)
    return set_error(GG_ERR_INVALID, "bad general shape");
  return GG_OK;
}

// sa() (general/sa.c:413-520) for any groups / centres.
int sa_sym_run(const gg_general_shape *sh, int32_t *edges, const gg_sa_params *P,
               gg_eval_sym_fn eval, void *ctx, gg_sa_stats *stats, int32_t *log_diam,
               uint64_t *log_sum, bool seed_rng) {
  const int nodes = sh->nodes, degree = sh->degree, groups = sh->groups, cen = sh->added_centers;
  const int based = (nodes - cen) / groups;
  if (!eval) {
    // Bring the device up before srandom(): its runtime's start-up may draw
    // from the same libc generator.
    const int rc0 = gg_init(-1);
    if (rc0) return rc0;
    eval = device_eval_sym;
  }
  const double t_start = now_s();
  double t_eval = 0;
  SymState S(nodes, degree, groups, cen, edges);
  S.simple = !sh->not_simple;
  int rc = S.init();
  if (rc) return rc;
  const int lines = S.lines;
  std::vector<int32_t> best(edges, edges + 2 * (size_t)lines);
  std::vector<int32_t> saved;  // -N: the edges before the current move
  double max_diff = 0;         // general/main.c:327
  int64_t logn = 0, evals = 0;
  auto evaluate = [&](int *diam, double *aspl) -> int {
    int32_t dm = 0, conn = 0;
    uint64_t sum = 0;
    const double t0 = now_s();
    const int r = eval(ctx, nodes, degree, based, groups, cen, S.adj.data(), &dm, &sum, &conn);
    t_eval += now_s() - t0;
    if (r) return r;
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
  gg::moore_bound(nodes, degree, &low_diam, &low_aspl);
  const double max_t = P->hill_climbing ? 0.0 : P->max_temp;
  const double min_t = P->hill_climbing ? 0.0 : P->min_temp;
  const int cycle = P->cooling_cycle > 0 ? P->cooling_cycle : 1;
  const double rate = P->hill_climbing ? 1.0
                      : (max_t != min_t) ? std::pow(min_t / max_t, (double)cycle / P->ncalcs)
                                         : 1.0;
  double temp = max_t;
  if (seed_rng) srandom(P->seed);
  int cur_diam = 0, best_diam = 0;
  double cur_aspl = 0, best_aspl = 0;
  rc = evaluate(&cur_diam, &cur_aspl);
  if (rc < 0) return rc;
  best_diam = cur_diam;
  best_aspl = cur_aspl;
  int64_t accepts = 0, ii;
  bool restore = false;
  for (ii = 0; ii < P->ncalcs; ++ii) {
    int nd;
    double na;
    while (true) {
      if (S.simple) {
        if (restore) S.restore();
      } else {
        saved.assign(edges, edges + 2 * (size_t)lines);
      }
      if ((rc = S.exchange())) return rc;
      if (!S.simple && (rc = S.init())) return rc;  // create_adj (general/sa.c:463-464)
      rc = evaluate(&nd, &na);
      if (rc < 0) return rc;
      if (rc == 1) break;
      if (S.simple) restore = true;
      else std::copy(saved.begin(), saved.end(), edges);
    }
    bool acc;  // accept (general/sa.c:371-411)
    if (nd < cur_diam) acc = true;
    else if (nd > cur_diam) acc = false;
    else if (na <= cur_aspl) acc = true;
    else if (P->hill_climbing) acc = false;
    else {
      const double diff = ((cur_aspl - na) * nodes * (nodes - 1)) / groups;
      if (P->detect_temp) max_diff = std::max(max_diff, -1.0 * diff);
      acc = std::exp(diff / temp) > ref_uniform();
    }
    if (!acc) {
      if (S.simple) restore = true;
      else std::copy(saved.begin(), saved.end(), edges);
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
    stats->max_diff_energy = max_diff;
  }
  return GG_OK;
}

}  // namespace

extern "C" int gg_sa_general_sym(const gg_general_shape *shape, int32_t *edges,
                                 const gg_sa_params *params, gg_eval_sym_fn eval, void *ctx,
                                 gg_sa_stats *stats, int32_t *log_diam, uint64_t *log_sum) {
  int rc = check_shape(shape);
  if (rc) return rc;
  if (!edges || !params) return set_error(GG_ERR_INVALID, "null argument");
  return sa_sym_run(shape, edges, params, eval, ctx, stats, log_diam, log_sum, true);
}

extern "C" int gg_general_program_sym(const gg_general_shape *shape, int32_t *edges, int source,
                                      int file_lines, const gg_sa_params *P, gg_eval_sym_fn eval,
                                      void *ctx, gg_sa_stats *stats, int32_t *log_diam,
                                      uint64_t *log_sum) {
  int rc = check_shape(shape);
  if (rc) return rc;
  if (!edges || !P || source < GG_GENERAL_FULL_FILE || source > GG_GENERAL_BASE_CENTRES)
    return set_error(GG_ERR_INVALID, "bad program arguments");
  const int nodes = shape->nodes, degree = shape->degree, groups = shape->groups;
  const int cen = shape->added_centers, based = (nodes - cen) / groups;
  const int lines = nodes * degree / 2, bl = lines / groups;
  gg_eval_sym_fn ev = eval ? eval : device_eval_sym;
  if (!eval && (rc = gg_init(-1))) return rc;  // before srandom(), as in sa_sym_run
  srandom(P->seed);  // general/main.c:360
  if (source == GG_GENERAL_BASE_CENTRES) {
    // create_symmetric_edge_with_vertexes (general/main.c:128-173): the file
    // holds org = file_lines base lines; centre edges and the remaining base
    // lines are laid down, then every group is a rotated copy.
    const int org = file_lines, per = degree / groups, org_nodes = nodes - cen;
    if (org < 0 || org + cen * per > bl) return set_error(GG_ERR_INVALID, "centre layout does not fit");
    for (int i = 0; i < cen; ++i)
      for (int j = 0; j < per; ++j) {
        const int line = org + i * per + j;
        edges[2 * line] = i * per + j;
        edges[2 * line + 1] = org_nodes + i;
      }
    int jj = 1;
    for (int i = cen * per; i < bl - org; ++i) {
      const int line = org + i;
      edges[2 * line] = i;
      edges[2 * line + 1] = 2 * based - jj++;
    }
    for (int j = 1; j < groups; ++j) {
      for (int i = 0; i < org; ++i) {
        edges[2 * (bl * j + i)] = edges[2 * i] + based * j;
        edges[2 * (bl * j + i) + 1] = edges[2 * i + 1] + based * j;
      }
      for (int i = org; i < org + cen * per; ++i) {
        edges[2 * (bl * j + i)] = edges[2 * i] + based * j;
        edges[2 * (bl * j + i) + 1] = edges[2 * i + 1];
      }
      for (int i = org + cen * per; i < bl; ++i) {
        edges[2 * (bl * j + i)] = edges[2 * i] + based * j;
        const int t = edges[2 * i + 1] + based * j;
        edges[2 * (bl * j + i) + 1] = t < org_nodes ? t : t - org_nodes;
      }
    }
  } else if (source == GG_GENERAL_BASE_FILE) {
    // create_symmetric_edge (general/main.c:175-210), centres absent.
    for (int j = 1; j < groups; ++j)
      for (int i = 0; i < bl; ++i) {
        edges[2 * (bl * j + i)] = edges[2 * i] + based * j;
        edges[2 * (bl * j + i) + 1] = edges[2 * i + 1] + based * j;
      }
    SymState S(nodes, degree, groups, 0, edges);
    S.simple = !shape->not_simple;
    if ((rc = S.init())) return rc;
    for (int i = 0; i < 100; ++i)
      if ((rc = S.exchange())) return rc;
    while (true) {
      if ((rc = S.exchange())) return rc;
      if (!S.simple && (rc = S.init())) return rc;
      int32_t dm = 0, conn = 0;
      uint64_t sum = 0;
      if ((rc = ev(ctx, nodes, degree, based, groups, 0, S.adj.data(), &dm, &sum, &conn))) return rc;
      if (conn && dm < nodes) break;
    }
  }
  // verfy_graph's degree and self-loop checks (general/main.c:212-241).
  std::vector<int> deg(nodes, 0);
  for (int i = 0; i < lines; ++i) {
    const int a = edges[2 * i], b = edges[2 * i + 1];
    if (a < 0 || a >= nodes || b < 0 || b >= nodes)
      return set_error(GG_ERR_INVALID, "edge %d endpoint out of range", i);
    if (a == b) return set_error(GG_ERR_INVALID, "NG. The same node in the edge.");
    ++deg[a];
    ++deg[b];
  }
  for (int v = 0; v < nodes; ++v)
    if (deg[v] != degree)
      return set_error(GG_ERR_INVALID, "NG. Not regular graph. degree = %d n[%d] = %d", degree, v, deg[v]);
  if (groups == 1) {  // check_current_edge (general/sa.c:557-572)
    std::vector<int32_t> adj((size_t)nodes * degree);
    if ((rc = gg_edges_to_adj(nodes, lines, degree, edges, adj.data()))) return rc;
    int32_t dm = 0, conn = 0;
    uint64_t sum = 0;
    if ((rc = ev(ctx, nodes, degree, based, groups, cen, adj.data(), &dm, &sum, &conn))) return rc;
    if (!conn || dm >= nodes)
      return set_error(GG_ERR_INVALID, "The input file has a node which is never reached by another node.");
    int low_diam;
    double low_aspl;
    gg::moore_bound(nodes, degree, &low_diam, &low_aspl);
    const double aspl = (double)sum / (((double)nodes - 1) * nodes);
    if (aspl == low_aspl) {
      if (stats) {
        *stats = gg_sa_stats{};
        stats->best_diameter = dm;
        stats->best_aspl = aspl;
      }
      return GG_ALREADY_OPTIMAL;
    }
  }
  {  // estimate_elapse_time's five moves on a copy (general/sa.c:523-554)
    std::vector<int32_t> copy(edges, edges + 2 * (size_t)lines);
    SymState T(nodes, degree, groups, cen, copy.data());
    T.simple = !shape->not_simple;
    if ((rc = T.init())) return rc;
    for (int i = 0; i < 5; ++i) {
      if ((rc = T.exchange())) return rc;
      if (!T.simple && (rc = T.init())) return rc;
    }
  }
  return sa_sym_run(shape, edges, P, eval, ctx, stats, log_diam, log_sum, false);
}
