// gg_layout.cpp -- bank-balanced vertex layout for the tile BFS (host code of
// libggeval.so; C ABI gg_bank_layout in include/gg_eval.h).  No counterpart
// in the reference, which gathers from host memory.
//
// A pull level reads entry slot s of 32 consecutive rows with one LDS read
// per half-wave; words that share a bank but not a 4-byte (8-byte for
// 64-bit words) LDS word are served one after another.  The device row
// order (gg_order_rows_kernel) permutes entries within rows, but a group
// whose 32 rows hold 25 entries on one bank (the typical busiest bank of a
// random 16-regular graph) cannot be served in fewer than 25 bank cycles
// over its 16 slots.  This layout changes the numbering as well:
//   1. relabel: vertices trade places only with vertices of the same bank
//      key, so every reach word keeps its bank while the groups' compositions
//      change; simulated annealing on sum over groups and banks of (entries
//      of the group on the bank)^2 flattens each group's bank histogram;
//   2. order: per group and slot, rows are matched to banks by augmenting
//      paths (banks with the most remaining entries first, capacities raised
//      only where the remaining slots could not take a bank's entries
//      otherwise), so most slots read every bank at most once.
// The result is an isomorphic numbering plus an order of each row's
// entries: every evaluation of it gives the same diameter and S.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "gg_eval.h"
#include "gg_internal.h"

namespace {

// LDS word index of id w for `shift` (0: 8-bit, 1: 16-bit, 2: 32- and
// 64-bit words, as gg_row_order defines it); bank = word & 31.
inline uint32_t lds_word(uint32_t w, int shift) { return (w << shift) >> 2; }

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x2545F4914F6CDD1Dull) {
    if (!s) s = 1;
  }
  uint64_t next() {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return s;
  }
  double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// Cost of a row table rows[n][dp] (new ids): sum over groups and slots of
// the most distinct LDS words on one bank, over `nt` threads.
uint64_t table_cost(const std::vector<uint32_t> &rows, int n, int dp, int shift, int nt) {
  const int ngroups = (n + 31) / 32;
  std::vector<uint64_t> part(nt, 0);
  auto work = [&](int t) {
    uint64_t total = 0;
    for (int g = t; g < ngroups; g += nt) {
      const int g0 = g * 32, rows_in = std::min(32, n - g0);
      for (int s = 0; s < dp; ++s) {
        uint32_t words[32];
        for (int r = 0; r < rows_in; ++r) words[r] = lds_word(rows[(size_t)(g0 + r) * dp + s], shift);
        std::sort(words, words + rows_in);
        int cnt[32] = {0};
        for (int r = 0; r < rows_in; ++r)
          if (r == 0 || words[r] != words[r - 1]) ++cnt[words[r] & 31];
        total += (uint64_t)*std::max_element(cnt, cnt + 32);
      }
    }
    part[t] = total;
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(work, t);
  work(0);
  for (auto &th : pool) th.join();
  uint64_t total = 0;
  for (uint64_t x : part) total += x;
  return total;
}

// Slot order of one group by per-slot bipartite matching (rows to banks).
// In slot s, a bank with more entries left than the slots after this one
// must take the excess now (its lower bound lb); every other bank takes at
// most one row.  The lower bounds are met first by augmenting paths from
// the banks' side, then the remaining rows are matched from the rows' side
// (banks with the most remaining entries tried first); augmenting never
// lowers a bank's load, so the bounds stay met.  A row left without a bank
// takes its entry on the slot's least-loaded bank.
struct GroupOrder {
  int d, shift;
  int rows_in;
  uint32_t rem[32][32];  // remaining entries of each row (new ids)
  int nrem[32];
  int cnt[32][32];      // cnt[r][b]: remaining entries of row r on bank b
  uint32_t bmask[32];   // rows with a remaining entry on bank b
  int remb[32];         // remaining entries per bank
  int load[32], cap[32], who[32][32], nwho[32], assigned[32];
  int inv[32];          // banks by priority (most remaining entries first)
  uint32_t rmask[32];   // per row: its banks as priority bits
  uint32_t vis_b, vis_r;

  // Row-side augmenting path from row r (Kuhn), bank loads bounded by cap;
  // banks tried in priority order.
  bool try_row(int r) {
    uint32_t m = rmask[r] & ~vis_b;
    while (m) {
      const int k = __builtin_ctz(m);
      m &= m - 1;
      if (vis_b >> k & 1u) continue;
      vis_b |= 1u << k;
      const int b = inv[k];
      if (load[b] < cap[b]) {
        who[b][nwho[b]++] = r;
        ++load[b];
        assigned[r] = b;
        return true;
      }
      for (int i = 0; i < nwho[b]; ++i) {
        const int r2 = who[b][i];
        if (try_row(r2)) {
          who[b][i] = r;
          assigned[r] = b;
          return true;
        }
      }
    }
    return false;
  }
  // Bank-side augmenting path: one more row for bank b.  A free row is taken
  // directly; an assigned row is taken if its bank can get another row.
  bool try_bank(int b) {
    uint32_t m = bmask[b] & ~vis_r;
    while (m) {
      const int r = __builtin_ctz(m);
      m &= m - 1;
      if ((vis_r >> r & 1u) || assigned[r] == b) continue;
      vis_r |= 1u << r;
      const int b2 = assigned[r];
      if (b2 < 0 || try_bank(b2)) {
        if (b2 >= 0) {  // r leaves b2 (whose load was restored by the recursion)
          for (int i = 0; i < nwho[b2]; ++i)
            if (who[b2][i] == r) {
              who[b2][i] = who[b2][--nwho[b2]];
              break;
            }
          --load[b2];
        }
        who[b][nwho[b]++] = r;
        ++load[b];
        assigned[r] = b;
        return true;
      }
    }
    return false;
  }
  // rows_new[r][0..d) in, reordered in place.
  void run(uint32_t *rows_new, int dp) {
    std::memset(cnt, 0, sizeof cnt);
    std::memset(bmask, 0, sizeof bmask);
    std::memset(remb, 0, sizeof remb);
    for (int r = 0; r < rows_in; ++r) {
      nrem[r] = d;
      for (int j = 0; j < d; ++j) {
        rem[r][j] = rows_new[(size_t)r * dp + j];
        const int b = (int)(lds_word(rem[r][j], shift) & 31u);
        ++cnt[r][b];
        ++remb[b];
        bmask[b] |= 1u << r;
      }
    }
    for (int s = 0; s < d; ++s) {
      const int left = d - s - 1;  // slots after this one
      int lb[32], rank[32];
      for (int b = 0; b < 32; ++b) {
        inv[b] = b;
        lb[b] = std::max(0, remb[b] - left);
        load[b] = 0;
        nwho[b] = 0;
      }
      std::stable_sort(inv, inv + 32, [&](int a, int b) { return remb[a] > remb[b]; });
      for (int k = 0; k < 32; ++k) rank[inv[k]] = k;
      for (int r = 0; r < rows_in; ++r) {
        assigned[r] = -1;
        uint32_t m = 0;
        for (int b = 0; b < 32; ++b)
          if (cnt[r][b]) m |= 1u << rank[b];
        rmask[r] = m;
      }
      // Lower bounds first (most constrained banks first).
      for (int k = 0; k < 32; ++k) {
        const int b = inv[k];
        for (int i = 0; i < lb[b]; ++i) {
          vis_r = 0;
          if (!try_bank(b)) break;
        }
      }
      for (int b = 0; b < 32; ++b) cap[b] = std::max(1, lb[b]);
      for (int r = 0; r < rows_in; ++r) {
        if (assigned[r] >= 0) continue;
        vis_b = 0;
        try_row(r);
      }
      for (int r = 0; r < rows_in; ++r) {
        int b = assigned[r];
        if (b < 0) {  // no bank within bounds: the slot's least-loaded bank
          int best = 1 << 30;
          for (int bb = 0; bb < 32; ++bb)
            if (cnt[r][bb] && load[bb] < best) {
              best = load[bb];
              b = bb;
            }
          ++load[b];
        }
        int pick = -1;
        for (int i = 0; i < nrem[r] && pick < 0; ++i)
          if ((int)(lds_word(rem[r][i], shift) & 31u) == b) pick = i;
        rows_new[(size_t)r * dp + s] = rem[r][pick];
        if (--cnt[r][b] == 0) bmask[b] &= ~(1u << r);
        --remb[b];
        rem[r][pick] = rem[r][--nrem[r]];
      }
    }
  }
};

}  // namespace

extern "C" int gg_bank_layout(int nodes, int degree, const int32_t *adj, int word_bits,
                              int64_t iterations, uint32_t seed, int threads, int32_t *relabel,
                              int32_t *adj_out, uint64_t cost[2]) {
  using gg::set_error;
  if (nodes < 2 || nodes > 65536 || degree < 1 || degree > 32 || !adj || !relabel || !adj_out ||
      !cost)
    return set_error(GG_ERR_INVALID, "bad layout arguments");
  if (word_bits != 8 && word_bits != 16 && word_bits != 32 && word_bits != 64)
    return set_error(GG_ERR_INVALID, "word bits must be 8/16/32/64");
  const int n = nodes, d = degree, dp = (d + 7) / 8 * 8;
  for (size_t k = 0; k < (size_t)n * d; ++k)
    if ((uint32_t)adj[k] >= (uint32_t)n)
      return set_error(GG_ERR_INVALID, "adjacency entry %d outside [0,%d)", adj[k], n);
  const int shift = word_bits == 8 ? 0 : word_bits == 16 ? 1 : 2;
  const int ngroups = (n + 31) / 32;
  auto key = [&](uint32_t v) { return (int)(lds_word(v, shift) & 31u); };

  int nt = threads > 0 ? threads : (int)std::thread::hardware_concurrency();
  nt = std::max(1, std::min(nt, std::max(1, ngroups / 16)));

  // Given table (identity numbering, padding = own id): its cost.
  std::vector<uint32_t> rows((size_t)n * dp);
  for (int v = 0; v < n; ++v)
    for (int j = 0; j < dp; ++j) rows[(size_t)v * dp + j] = j < d ? (uint32_t)adj[(size_t)v * d + j] : (uint32_t)v;
  cost[0] = table_cost(rows, n, dp, shift, nt);

  // 1. Relabel.  slot_of[v] = new id of v; vert_at[i] = vertex at new id i.
  std::vector<int32_t> slot_of(n), vert_at(n);
  for (int v = 0; v < n; ++v) slot_of[v] = vert_at[v] = v;
  std::vector<uint8_t> hv((size_t)n * 32, 0);  // bank histogram of each row
  for (int v = 0; v < n; ++v)
    for (int j = 0; j < d; ++j) ++hv[(size_t)v * 32 + key((uint32_t)adj[(size_t)v * d + j])];
  std::vector<int32_t> H((size_t)ngroups * 32, 0);
  for (int i = 0; i < n; ++i)
    for (int b = 0; b < 32; ++b) H[(size_t)(i >> 5) * 32 + b] += hv[(size_t)i * 32 + b];
  if (iterations <= 0) iterations = (int64_t)n * 80;
  const int epochs = 40;
  const int64_t per_thread = std::max<int64_t>(1, iterations / ((int64_t)epochs * nt));
  const double t0 = 4.0, t1 = 0.1;
  Rng master(seed);
  std::vector<int> gperm(ngroups);
  for (int g = 0; g < ngroups; ++g) gperm[g] = g;
  for (int ep = 0; ep < epochs && ngroups >= 2; ++ep) {
    // Groups dealt out to the threads at random; a thread swaps only within
    // its own groups, so no two threads touch one group.
    for (int g = ngroups - 1; g > 0; --g) std::swap(gperm[g], gperm[master.next() % (uint64_t)(g + 1)]);
    const uint64_t ep_seed = master.next();
    auto work = [&](int t) {
      Rng rng(ep_seed + 7919ull * (uint64_t)t);
      // This thread's slots per bank key.
      std::vector<std::vector<int32_t>> by_key(32);
      for (int k = t; k < ngroups; k += nt) {
        const int g = gperm[k];
        for (int i = g * 32; i < std::min(n, g * 32 + 32); ++i) by_key[key((uint32_t)i)].push_back(i);
      }
      std::vector<int32_t> all;
      for (auto &lst : by_key) all.insert(all.end(), lst.begin(), lst.end());
      if (all.size() < 2) return;
      double T = t0;
      for (int64_t it = 0; it < per_thread; ++it) {
        if ((it & 1023) == 0) T = t0 * std::pow(t1 / t0, ((double)ep + (double)it / (double)per_thread) / epochs);
        const int32_t i1 = all[rng.next() % all.size()];
        const auto &lst = by_key[key((uint32_t)i1)];
        const int32_t i2 = lst[rng.next() % lst.size()];
        const int g1 = i1 >> 5, g2 = i2 >> 5;
        if (g1 == g2) continue;
        const int u = vert_at[i1], v = vert_at[i2];
        const uint8_t *hu = &hv[(size_t)u * 32], *hw = &hv[(size_t)v * 32];
        int32_t *h1 = &H[(size_t)g1 * 32], *h2 = &H[(size_t)g2 * 32];
        int q = 0;
        for (int b = 0; b < 32; ++b) {
          const int dd = (int)hw[b] - (int)hu[b];
          q += 2 * dd * (h1[b] - h2[b]) + 2 * dd * dd;
        }
        if (q <= 0 || rng.unit() < std::exp(-(double)q / T)) {
          for (int b = 0; b < 32; ++b) {
            const int dd = (int)hw[b] - (int)hu[b];
            h1[b] += dd;
            h2[b] -= dd;
          }
          vert_at[i1] = v;
          vert_at[i2] = u;
          slot_of[v] = i1;
          slot_of[u] = i2;
        }
      }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto &th : pool) th.join();
  }

  // 2. Order: the relabelled rows (new ids), matched slot by slot per group.
  for (int i = 0; i < n; ++i) {
    const int v = vert_at[i];
    for (int j = 0; j < dp; ++j)
      rows[(size_t)i * dp + j] = j < d ? (uint32_t)slot_of[adj[(size_t)v * d + j]] : (uint32_t)i;
  }
  {
    auto work = [&](int t) {
      GroupOrder go;
      go.d = d;
      go.shift = shift;
      for (int g = t; g < ngroups; g += nt) {
        go.rows_in = std::min(32, n - g * 32);
        go.run(&rows[(size_t)g * 32 * dp], dp);
      }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto &th : pool) th.join();
  }
  cost[1] = table_cost(rows, n, dp, shift, nt);

  // Outputs in the caller's numbering: row v lists its neighbours in the
  // slot order of its new row slot_of[v].
  for (int v = 0; v < n; ++v) {
    relabel[v] = slot_of[v];
    const uint32_t *r = &rows[(size_t)slot_of[v] * dp];
    for (int j = 0; j < d; ++j) adj_out[(size_t)v * d + j] = vert_at[r[j]];
  }
  return GG_OK;
}
