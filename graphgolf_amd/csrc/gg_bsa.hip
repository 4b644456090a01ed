// gg_bsa.hip -- batched, surrogate-ranked simulated annealing over the
// patched-plan evaluator (gg_sa_batched, include/gg_host.h).
//
// No reference counterpart as a whole: the reference draws one 2-opt move at
// a time (general/sa.c:217-369) and scores it with evaluation().  Here every
// step scores `exact` candidate graphs (the current graph plus one or more
// vertex-disjoint moves) in one launch of a patched plan; a device-side
// feature kernel and the BF16 MLP (gg_mlp.hip) decide which moves go into
// the candidates.  The accept rule is the reference's (general/sa.c:371-411)
// applied to the best candidate, so every accepted graph is scored exactly.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <unordered_map>
#include <vector>

#include "gg_device.h"
#include "gg_eval.h"
#include "gg_host.h"
#include "gg_incr.h"
#include "gg_internal.h"
#include "gg_mlp.h"
#include "gg_sa_common.h"

using gg::set_error;

namespace {
constexpr int kFeat = 64;        // MLP input width
constexpr int kVStat = 10;       // per-vertex statistics (see VertexStats)
constexpr int kMaxCombine = 128;  // moves per candidate (4 patch entries per copy)
constexpr int kMaxGroups = 8;     // rotational copies of a move (groups)
constexpr int kMaxSlots = 4 * kMaxGroups;
constexpr int kBlock = 256;
}  // namespace

__device__ __forceinline__ int bsa_block_sum(int v, int *red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  int t = 0;
  for (int w = 0; w < kBlock / 64; ++w) t += red[w];
  __syncthreads();
  return t;
}

// Features of proposal p = (a,b),(c,d) -> variant r, one workgroup each.
// Removed pairs (a,b), (c,d); added pairs (p1,q1), (p2,q2) with
// r = 0: (a,d), (c,b) and r = 1: (a,c), (b,d).  Bx / B2x are the closed 1- and
// 2-hop balls of x in the current graph `adj` (row-major n x d).
//   f[10e + k]   vstat[endpoint e][k]            e over a, b, c, d
//   f[40 + 4q]   |B1x & B1y| / d                 q over the four pairs
//   f[41 + 4q]   |B1x & B2y| / d
//   f[42 + 4q]   |B2x & B2y| / d^2
//   f[43 + 4q]   (vstat[x][7] + vstat[y][7]) / 2 (mean relative distance sum)
//   f[56]        r;  the rest 0
// LDS: two n-bit bitmaps (B2x, B2y), reused for the four pairs.
__global__ void __launch_bounds__(kBlock)
gg_bsa_features(const int32_t *__restrict__ adj, int n, int d, const int32_t *__restrict__ props,
                const float *__restrict__ vstat, float *__restrict__ feats) {
  extern __shared__ uint32_t bm[];
  __shared__ int red[kBlock / 64];
  __shared__ int cnt[4][3];
  const int words = (n + 31) >> 5;
  const int tid = threadIdx.x;
  const int32_t *pr = props + 5 * (size_t)blockIdx.x;
  const int a = pr[0], b = pr[1], c = pr[2], e = pr[3], r = pr[4];
  const int ball = 1 + d + d * d;
  for (int q = 0; q < 4; ++q) {
    const int x = q == 0 ? a : q == 1 ? c : q == 2 ? a : (r ? b : c);
    const int y = q == 0 ? b : q == 1 ? e : q == 2 ? (r ? c : e) : (r ? e : b);
    for (int i = tid; i < 2 * words; i += kBlock) bm[i] = 0;
    __syncthreads();
    for (int i = tid; i < 2 * ball; i += kBlock) {
      const int side = i >= ball;
      const int k = i - side * ball;
      const int root = side ? y : x;
      int v;
      if (k == 0) {
        v = root;
      } else if (k <= d) {
        v = adj[(size_t)root * d + k - 1];
      } else {
        const int j = (k - 1 - d) / d, l = (k - 1 - d) - j * d;
        v = adj[(size_t)adj[(size_t)root * d + j] * d + l];
      }
      atomicOr(&bm[side * words + (v >> 5)], 1u << (v & 31));
    }
    __syncthreads();
    int c0 = 0, c1 = 0, c2 = 0;
    for (int k = tid; k <= d; k += kBlock) {
      const int u = k == 0 ? x : adj[(size_t)x * d + k - 1];
      c1 += (bm[words + (u >> 5)] >> (u & 31)) & 1;
      int in1 = u == y;
      for (int j = 0; j < d; ++j) in1 |= adj[(size_t)y * d + j] == u;
      c0 += in1;
    }
    for (int w = tid; w < words; w += kBlock) c2 += __popc(bm[w] & bm[words + w]);
    c0 = bsa_block_sum(c0, red);
    c1 = bsa_block_sum(c1, red);
    c2 = bsa_block_sum(c2, red);
    if (tid == 0) {
      cnt[q][0] = c0;
      cnt[q][1] = c1;
      cnt[q][2] = c2;
    }
  }
  __syncthreads();
  float *f = feats + (size_t)blockIdx.x * kFeat;
  for (int i = tid; i < kFeat; i += kBlock) {
    float v = 0.0f;
    if (i < 40) {
      const int end = i / 10;
      const int vx = end == 0 ? a : end == 1 ? b : end == 2 ? c : e;
      v = vstat[(size_t)vx * kVStat + (i - 10 * end)];
    } else if (i < 56) {
      const int q = (i - 40) >> 2, k = (i - 40) & 3;
      const int x = q == 0 ? a : q == 1 ? c : q == 2 ? a : (r ? b : c);
      const int y = q == 0 ? b : q == 1 ? e : q == 2 ? (r ? c : e) : (r ? e : b);
      if (k == 0) v = (float)cnt[q][0] / d;
      else if (k == 1) v = (float)cnt[q][1] / d;
      else if (k == 2) v = (float)cnt[q][2] / ((float)d * d);
      else v = 0.5f * (vstat[(size_t)x * kVStat + 7] + vstat[(size_t)y * kVStat + 7]);
    } else if (i == 56) {
      v = (float)r;
    }
    f[i] = v;
  }
}

// dst[dst_row[i]] = src[src_row[i]] for i < rows (rows of kFeat floats; a
// null index array means the identity).
__global__ void gg_bsa_copy_rows(const float *__restrict__ src, const int32_t *__restrict__ src_row,
                                 float *__restrict__ dst, const int32_t *__restrict__ dst_row,
                                 int rows) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * kFeat) return;
  const int row = i / kFeat, k = i - row * kFeat;
  const int s = src_row ? src_row[row] : row, t = dst_row ? dst_row[row] : row;
  dst[(size_t)t * kFeat + k] = src[(size_t)s * kFeat + k];
}

namespace {

using gg::EdgeSet;
using gg::ekey;
using gg::now_s;
using gg::Rng;

#define BSA_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      return set_error(GG_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
  } while (0)

// A 2-opt move on edge-list lines t0 = (a,b), t1 = (c,d):
// r = 0 -> (a,d), (c,b);  r = 1 -> (a,c), (b,d).
// ---------------------------------------------------------------------------
// Move generation on the device (ranked steps of general graphs).

// splitmix64 finaliser: a counter-based stream, one value per (seed, step,
// proposal, draw).
__device__ __forceinline__ uint64_t bsa_mix(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__device__ __forceinline__ bool row_has(const int32_t *__restrict__ adj, int d, int v, int u) {
  bool hit = false;
  for (int j = 0; j < d; ++j) hit |= adj[(size_t)v * d + j] == u;
  return hit;
}

// Thread i draws lines t0 != t1 and a variant r (r = 0 adds (a,d),(c,b);
// r = 1 adds (a,c),(b,d), as exchange_edge_2opt's two choices,
// general/sa.c:254-266), redrawing while the two lines share a vertex
// (has_duplicated_vertex, general/common.c:119-122), and marks the move
// valid when neither added edge exists in the current adjacency (the
// duplicate check, general/common.c:184-276, as d-entry row scans).
// out[i] = (a, b, c, d, r) for the feature kernel; lines_out[i] = (t0, t1,
// valid).
__global__ void __launch_bounds__(256)
gg_bsa_propose(int count, uint64_t seed, uint64_t step, int lines, const int32_t *__restrict__ edges,
               const int32_t *__restrict__ adj, int d, int32_t *__restrict__ out,
               int32_t *__restrict__ lines_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint64_t base = bsa_mix(seed ^ bsa_mix(step * 0x100000001B3ull + (uint64_t)i));
  int t0 = 0, t1 = 0, r = 0, a = 0, b = 0, c = 0, e = 0;
  bool ok = false;
  for (int k = 0; k < 64 && !ok; ++k) {
    const uint64_t x = bsa_mix(base + (uint64_t)k);
    t0 = (int)(((x & 0xffffffffull) * (uint64_t)lines) >> 32);
    t1 = (int)(((x >> 32) * (uint64_t)lines) >> 32);
    r = (int)((bsa_mix(x) >> 63) & 1u);
    a = edges[2 * t0];
    b = edges[2 * t0 + 1];
    c = edges[2 * t1];
    e = edges[2 * t1 + 1];
    ok = t0 != t1 && a != c && a != e && b != c && b != e;
  }
  bool valid = ok;
  if (ok) {
    const int q1 = r ? c : e, p2 = r ? b : c, q2 = r ? e : b;
    valid = !row_has(adj, d, a, q1) && !row_has(adj, d, p2, q2);
  }
  int32_t *o = out + 5 * (size_t)i;
  o[0] = a;
  o[1] = b;
  o[2] = c;
  o[3] = e;
  o[4] = r;
  int32_t *l = lines_out + 3 * (size_t)i;
  l[0] = t0;
  l[1] = t1;
  l[2] = valid ? 1 : 0;
}

// Grid graphs (vertex v at column v / height, row v % height, every edge at
// most `length` Manhattan steps): line t0 as above, then a vertex c within
// the length of t0's first endpoint (offsets: packed (dx, dy) of the length
// ball) and t1 one of c's lines (lineat[c*d + j]), as the host's local draw
// does, so that most draws respect the limit; valid also needs both added
// edges within the length (grid/sa.c:196-201).
__global__ void __launch_bounds__(256)
gg_bsa_propose_grid(int count, uint64_t seed, uint64_t step, int lines, const int32_t *__restrict__ edges,
                    const int32_t *__restrict__ adj, const int32_t *__restrict__ lineat, int d, int n,
                    int height, int length, const int32_t *__restrict__ offsets, int noff,
                    int32_t *__restrict__ out, int32_t *__restrict__ lines_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint64_t base = bsa_mix(seed ^ bsa_mix(step * 0x100000001B3ull + (uint64_t)i));
  const int width = n / height;
  auto near = [&](int u, int v) {
    const int dx = u / height - v / height, dy = u % height - v % height;
    return (dx < 0 ? -dx : dx) + (dy < 0 ? -dy : dy) <= length;
  };
  int t0 = 0, t1 = 0, r = 0, a = 0, b = 0, c = 0, e = 0;
  bool ok = false;
  for (int k = 0; k < 64 && !ok; ++k) {
    const uint64_t x = bsa_mix(base + (uint64_t)k), y = bsa_mix(x);
    t0 = (int)(((x & 0xffffffffull) * (uint64_t)lines) >> 32);
    a = edges[2 * t0];
    b = edges[2 * t0 + 1];
    const int o = offsets[(int)(((x >> 32) * (uint64_t)noff) >> 32)];
    const int cx = a / height + (o / 4096 - 1024), cy = a % height + (o % 4096 - 1024);
    if (cx < 0 || cy < 0 || cy >= height || cx >= width) continue;
    const int cv = cx * height + cy;
    t1 = lineat[(size_t)cv * d + (int)(((y & 0xffffffffull) * (uint64_t)d) >> 32)];
    r = (int)((y >> 63) & 1u);
    c = edges[2 * t1];
    e = edges[2 * t1 + 1];
    ok = t0 != t1 && a != c && a != e && b != c && b != e;
  }
  bool valid = ok;
  if (ok) {
    const int q1 = r ? c : e, p2 = r ? b : c, q2 = r ? e : b;
    valid = near(a, q1) && near(p2, q2) && !row_has(adj, d, a, q1) && !row_has(adj, d, p2, q2);
  }
  int32_t *o = out + 5 * (size_t)i;
  o[0] = a;
  o[1] = b;
  o[2] = c;
  o[3] = e;
  o[4] = r;
  int32_t *l = lines_out + 3 * (size_t)i;
  l[0] = t0;
  l[1] = t1;
  l[2] = valid ? 1 : 0;
}

// Writes changed lines (line, u, v) into the device edge list, and records
// (-1 - slot, line, 0) into the grid line table.
__global__ void __launch_bounds__(256)
gg_bsa_set_lines(const int32_t *__restrict__ upd, int count, int32_t *__restrict__ edges,
                 int32_t *__restrict__ lineat) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int t = upd[3 * i];
  if (t >= 0) {
    edges[2 * t] = upd[3 * i + 1];
    edges[2 * t + 1] = upd[3 * i + 2];
  } else if (lineat) {
    lineat[-1 - t] = upd[3 * i + 1];
  }
}

// Symmetric graphs (groups g > 1): a move is the 2g-edge 2-opt of
// exchange_edge_2opt (general/sa.c:270-300, grid/sa.c:183-206) -- lines t0,
// t1 and their g-1 rotated copies, every copy re-paired with the same
// variant.  rotk[i*n + v] is vertex v rotated i times (rotk[0] = identity),
// rotline[i*lines + t] the line holding line t's edge rotated i times.
// Drawn like gg_bsa_propose (grids: t1 near t0 as gg_bsa_propose_grid) and
// redrawn while the lines share a vertex, lie in one orbit, or one of them
// maps to itself under a rotation (the reference re-pairs such lines with
// edge_1g_opt; this draw leaves them to other moves).  Valid when the 2g
// added edges are no loops, pairwise distinct, absent from the current
// adjacency and, on a grid, within the length.
__global__ void __launch_bounds__(256)
gg_bsa_propose_sym(int count, uint64_t seed, uint64_t step, int lines, const int32_t *__restrict__ edges,
                   const int32_t *__restrict__ adj, int d, int n, int groups,
                   const int32_t *__restrict__ rotk, const int32_t *__restrict__ rotline,
                   const int32_t *__restrict__ lineat, int height, int length,
                   const int32_t *__restrict__ offsets, int noff, int32_t *__restrict__ out,
                   int32_t *__restrict__ lines_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint64_t base = bsa_mix(seed ^ bsa_mix(step * 0x100000001B3ull + (uint64_t)i));
  auto near = [&](int u, int v) {
    if (height <= 0) return true;
    const int dx = u / height - v / height, dy = u % height - v % height;
    return (dx < 0 ? -dx : dx) + (dy < 0 ? -dy : dy) <= length;
  };
  int t0 = 0, t1 = 0, r = 0, a = 0, b = 0, c = 0, e = 0;
  bool ok = false;
  for (int k = 0; k < 64 && !ok; ++k) {
    const uint64_t x = bsa_mix(base + (uint64_t)k), y = bsa_mix(x);
    t0 = (int)(((x & 0xffffffffull) * (uint64_t)lines) >> 32);
    a = edges[2 * t0];
    b = edges[2 * t0 + 1];
    if (height > 0) {
      const int o = offsets[(int)(((x >> 32) * (uint64_t)noff) >> 32)];
      const int cx = a / height + (o / 4096 - 1024), cy = a % height + (o % 4096 - 1024);
      if (cx < 0 || cy < 0 || cy >= height || cx >= n / height) continue;
      t1 = lineat[(size_t)(cx * height + cy) * d + (int)(((y & 0xffffffffull) * (uint64_t)d) >> 32)];
    } else {
      t1 = (int)(((x >> 32) * (uint64_t)lines) >> 32);
    }
    r = (int)((y >> 63) & 1u);
    c = edges[2 * t1];
    e = edges[2 * t1 + 1];
    ok = t0 != t1 && a != c && a != e && b != c && b != e;
    for (int q = 1; q < groups && ok; ++q) {
      const int r0 = rotline[(size_t)q * lines + t0], r1 = rotline[(size_t)q * lines + t1];
      ok = r0 != t0 && r1 != t1 && r0 != t1;
    }
  }
  bool valid = ok;
  if (ok) {
    const int q1 = r ? c : e, p2 = r ? b : c, q2 = r ? e : b;
    valid = near(a, q1) && near(p2, q2);
    int nu[2 * kMaxGroups], nv[2 * kMaxGroups];
    for (int q = 0; q < groups && valid; ++q) {
      const int32_t *rk = rotk + (size_t)q * n;
      nu[2 * q] = rk[a];
      nv[2 * q] = rk[q1];
      nu[2 * q + 1] = rk[p2];
      nv[2 * q + 1] = rk[q2];
      for (int h = 2 * q; h < 2 * q + 2 && valid; ++h)
        valid = nu[h] != nv[h] && !row_has(adj, d, nu[h], nv[h]);
    }
    for (int h = 0; h < 2 * groups && valid; ++h)
      for (int m = h + 1; m < 2 * groups && valid; ++m)
        valid = !((nu[h] == nu[m] && nv[h] == nv[m]) || (nu[h] == nv[m] && nv[h] == nu[m]));
  }
  int32_t *o = out + 5 * (size_t)i;
  o[0] = a;
  o[1] = b;
  o[2] = c;
  o[3] = e;
  o[4] = r;
  int32_t *l = lines_out + 3 * (size_t)i;
  l[0] = t0;
  l[1] = t1;
  l[2] = valid ? 1 : 0;
}

struct Prop {
  int t0, t1, r, a, b, c, d;
};

// Adjacency entries a move rewrites: 4 per rotated copy.
struct Slots {
  int n = 0;
  int at[kMaxSlots], val[kMaxSlots];
};

struct SaGraph {
  int n, d, lines;
  int32_t *edges;
  std::vector<int32_t> adj;  // current graph, row-major n x d
  EdgeSet set;
  SaGraph(int n_, int d_, int32_t *e)
      : n(n_), d(d_), lines(n_ * d_ / 2), edges(e), adj((size_t)n_ * d_), set((size_t)n_ * d_ / 2) {}
  int init() {
    int rc = gg_edges_to_adj(n, lines, d, edges, adj.data());
    if (rc) return rc;
    for (int i = 0; i < lines; ++i) {
      const int u = edges[2 * i], v = edges[2 * i + 1];
      if (u == v || set.has(ekey(u, v)))
        return set_error(GG_ERR_INVALID, "batched SA needs a simple graph (line %d: %d-%d)", i, u, v);
      set.add(ekey(u, v));
    }
    for (int v = 0; v < n; ++v)
      for (int j = 0; j < d; ++j)
        if (adj[(size_t)v * d + j] == v)
          return set_error(GG_ERR_INVALID, "batched SA needs a %d-regular graph", d);
    return GG_OK;
  }
  // Reads the endpoints; false when the two lines share a vertex.
  bool load(Prop &p) const {
    p.a = edges[2 * p.t0];
    p.b = edges[2 * p.t0 + 1];
    p.c = edges[2 * p.t1];
    p.d = edges[2 * p.t1 + 1];
    return p.t0 != p.t1 && p.a != p.c && p.a != p.d && p.b != p.c && p.b != p.d;
  }
  void draw(Rng &rng, Prop &p) const {
    do {
      p.t0 = (int)rng.below((uint32_t)lines);
      p.t1 = (int)rng.below((uint32_t)lines);
      p.r = (int)rng.below(2);
    } while (!load(p) || !sym_lines_ok(p));
  }

  // ---- Symmetric graphs (groups > 1): rotk[i*n + v] = v rotated i times,
  // rotline[i*lines + t] = the line holding line t's edge rotated i times.
  int groups = 1;
  std::vector<int32_t> rotk, rotline;
  // General graphs rotate by v -> v + n/groups (mod n); grids of the given
  // height by ROTATE (grid/common.c:63-86): 180 degrees for 2 groups, 90 per
  // copy for 4.  Fails unless the edge set is invariant under the rotation.
  int sym_setup(int g_, int height) {
    groups = g_;
    const int B = n / groups;
    rotk.assign((size_t)groups * n, 0);
    const int width = height > 0 ? n / height : 0;
    auto grid_rot = [&](int v, int deg) {
      const int w = v / height, h = v % height;
      if (groups == 2) return (width - w - 1) * height + (height - h - 1);
      if (deg == 90) return h * height + (height - w - 1);
      if (deg == 180) return (height - w - 1) * height + (height - h - 1);
      return (height - h - 1) * height + w;
    };
    for (int q = 0; q < groups; ++q)
      for (int v = 0; v < n; ++v)
        rotk[(size_t)q * n + v] = q == 0 ? v : height > 0 ? grid_rot(v, 90 * q) : (v + q * B) % n;
    std::unordered_map<uint64_t, int> line_of;
    line_of.reserve((size_t)lines * 2);
    for (int t = 0; t < lines; ++t) line_of[ekey(edges[2 * t], edges[2 * t + 1])] = t;
    rotline.assign((size_t)groups * lines, 0);
    for (int q = 0; q < groups; ++q)
      for (int t = 0; t < lines; ++t) {
        const int32_t *rk = rotk.data() + (size_t)q * n;
        auto it = line_of.find(ekey(rk[edges[2 * t]], rk[edges[2 * t + 1]]));
        if (it == line_of.end())
          return set_error(GG_ERR_INVALID, "the graph is not invariant under its %d-fold rotation (line %d)",
                           groups, t);
        rotline[(size_t)q * lines + t] = it->second;
      }
    return GG_OK;
  }
  int rot(int q, int v) const { return rotk[(size_t)q * n + v]; }
  int rline(int q, int t) const { return rotline[(size_t)q * lines + t]; }
  // Lines in different orbits, neither mapped to itself by a rotation.
  bool sym_lines_ok(const Prop &p) const {
    for (int q = 1; q < groups; ++q)
      if (rline(q, p.t0) == p.t0 || rline(q, p.t1) == p.t1 || rline(q, p.t0) == p.t1) return false;
    return true;
  }
  // The vertices a move touches (4 per rotated copy).
  int move_vertices(const Prop &p, int *out) const {
    if (groups == 1) {
      out[0] = p.a;
      out[1] = p.b;
      out[2] = p.c;
      out[3] = p.d;
      return 4;
    }
    for (int q = 0; q < groups; ++q) {
      out[4 * q] = rot(q, p.a);
      out[4 * q + 1] = rot(q, p.b);
      out[4 * q + 2] = rot(q, p.c);
      out[4 * q + 3] = rot(q, p.d);
    }
    return 4 * groups;
  }
  // The move's added edges, copy q: (nu[2q], nv[2q]), (nu[2q+1], nv[2q+1]).
  void sym_added(const Prop &p, int *nu, int *nv) const {
    const int q1 = p.r ? p.c : p.d, p2 = p.r ? p.b : p.c, q2 = p.r ? p.d : p.b;
    for (int q = 0; q < groups; ++q) {
      nu[2 * q] = rot(q, p.a);
      nv[2 * q] = rot(q, q1);
      nu[2 * q + 1] = rot(q, p2);
      nv[2 * q + 1] = rot(q, q2);
    }
  }
  bool sym_fresh(const Prop &p) const {
    if (!sym_lines_ok(p)) return false;
    int nu[2 * kMaxGroups], nv[2 * kMaxGroups];
    sym_added(p, nu, nv);
    if (!short_enough(nu[0], nv[0]) || !short_enough(nu[1], nv[1])) return false;
    for (int h = 0; h < 2 * groups; ++h) {
      if (nu[h] == nv[h] || set.has(ekey(nu[h], nv[h]))) return false;
      for (int m = h + 1; m < 2 * groups; ++m)
        if (ekey(nu[h], nv[h]) == ekey(nu[m], nv[m])) return false;
    }
    return true;
  }
  // The 4g adjacency entries: per copy q, a_q's b_q, b_q's a_q, c_q's d_q,
  // d_q's c_q (the rows of distinct copies may be the same vertex; their
  // entries differ since the removed edges do).
  bool sym_slots(const Prop &p, Slots &s) const {
    s.n = 0;
    for (int q = 0; q < groups; ++q) {
      const int a = rot(q, p.a), b = rot(q, p.b), c = rot(q, p.c), d_ = rot(q, p.d);
      const int x0 = find(a, b), x1 = find(b, a), x2 = find(c, d_), x3 = find(d_, c);
      if (x0 < 0 || x1 < 0 || x2 < 0 || x3 < 0) return false;
      const int at[4] = {a * d + x0, b * d + x1, c * d + x2, d_ * d + x3};
      const int val[4] = {p.r ? c : d_, p.r ? d_ : c, p.r ? a : b, p.r ? b : a};
      for (int k = 0; k < 4; ++k) {
        s.at[s.n] = at[k];
        s.val[s.n] = val[k];
        ++s.n;
      }
    }
    return true;
  }
  // Lines written by a move (line, u, v): rotated copies of t0's and t1's
  // new edges, oriented as the rotation of copy 0.
  int sym_lines(const Prop &p, int32_t *out) const {
    int nu[2 * kMaxGroups], nv[2 * kMaxGroups];
    sym_added(p, nu, nv);
    for (int q = 0; q < groups; ++q) {
      const int l0 = rline(q, p.t0), l1 = rline(q, p.t1);
      out[6 * q] = l0;
      out[6 * q + 1] = nu[2 * q];
      out[6 * q + 2] = nv[2 * q];
      out[6 * q + 3] = l1;
      out[6 * q + 4] = nu[2 * q + 1];
      out[6 * q + 5] = nv[2 * q + 1];
    }
    return 2 * groups;
  }
  void sym_apply(const Prop &p, const Slots &s) {
    for (int k = 0; k < s.n; ++k) adj[s.at[k]] = s.val[k];
    if (!lineat.empty())
      for (int q = 0; q < groups; ++q) {
        const int l0 = rline(q, p.t0), l1 = rline(q, p.t1);
        lineat[s.at[4 * q]] = l0;
        lineat[s.at[4 * q + 1]] = l1;
        lineat[s.at[4 * q + 2]] = p.r ? l0 : l1;
        lineat[s.at[4 * q + 3]] = p.r ? l1 : l0;
      }
    int32_t upd[6 * kMaxGroups];
    sym_lines(p, upd);
    for (int q = 0; q < groups; ++q) {
      for (int t : {rline(q, p.t0), rline(q, p.t1)}) set.remove(ekey(edges[2 * t], edges[2 * t + 1]));
    }
    for (int h = 0; h < 2 * groups; ++h) {
      const int t = upd[3 * h];
      edges[2 * t] = upd[3 * h + 1];
      edges[2 * t + 1] = upd[3 * h + 2];
      set.add(ekey(upd[3 * h + 1], upd[3 * h + 2]));
    }
  }
  // Grid graphs: line t1 is drawn among the lines at a vertex within `length`
  // steps of line t0's first endpoint, so that most draws respect the limit.
  std::vector<int32_t> lineat;  // lineat[v*d + j]: the line holding edge (v, adj[v][j])
  std::vector<int> offsets;     // packed (dx, dy) of the length ball
  void grid_setup() {
    lineat.assign((size_t)n * d, -1);
    for (int i = 0; i < lines; ++i)
      for (int side = 0; side < 2; ++side) {
        const int u = edges[2 * i + side], v = edges[2 * i + 1 - side];
        for (int j = 0; j < d; ++j)
          if (adj[(size_t)u * d + j] == v && lineat[(size_t)u * d + j] < 0) {
            lineat[(size_t)u * d + j] = i;
            break;
          }
      }
    offsets.clear();
    for (int dx = -grid_length; dx <= grid_length; ++dx)
      for (int dy = -grid_length; dy <= grid_length; ++dy)
        if ((dx || dy) && std::abs(dx) + std::abs(dy) <= grid_length)
          offsets.push_back((dx + 1024) * 4096 + (dy + 1024));
  }
  bool draw_local(Rng &rng, Prop &p) const {
    p.t0 = (int)rng.below((uint32_t)lines);
    const int a = edges[2 * p.t0];
    const int o = offsets[rng.below((uint32_t)offsets.size())];
    const int cx = a / grid_height + (o / 4096 - 1024), cy = a % grid_height + (o % 4096 - 1024);
    if (cx < 0 || cy < 0 || cy >= grid_height || cx >= n / grid_height) return false;
    const int c = cx * grid_height + cy;
    p.t1 = lineat[(size_t)c * d + rng.below((uint32_t)d)];
    p.r = (int)rng.below(2);
    return load(p) && sym_lines_ok(p);
  }
  // A valid move (simple graph kept, grid lengths respected); false after
  // `tries` failed draws.
  bool draw_valid(Rng &rng, Prop &p, int tries) const {
    for (int k = 0; k < tries; ++k) {
      if (grid_height > 0 ? draw_local(rng, p) : (draw(rng, p), true))
        if (fresh(p)) return true;
    }
    return false;
  }
  // Grid graphs (height > 0): vertex v at column v / height, row v % height,
  // every edge at most `length` Manhattan steps long (grid/sa.c:196-201).
  int grid_height = 0, grid_length = 0;
  bool short_enough(int u, int v) const {
    if (grid_height <= 0) return true;
    return std::abs(u / grid_height - v / grid_height) + std::abs(u % grid_height - v % grid_height) <=
           grid_length;
  }
  // True when neither added edge exists yet (the move keeps the graph simple)
  // and, on a grid, both respect the length limit.
  bool fresh(const Prop &p) const {
    if (groups > 1) return sym_fresh(p);
    const int q1 = p.r ? p.c : p.d, p2 = p.r ? p.b : p.c, q2 = p.r ? p.d : p.b;
    return short_enough(p.a, q1) && short_enough(p2, q2) && !set.has(ekey(p.a, q1)) &&
           !set.has(ekey(p2, q2));
  }
  int find(int y, int v) const {
    for (int x = 0; x < d; ++x)
      if (adj[(size_t)y * d + x] == v) return x;
    return -1;
  }
  // Adjacency entries the move rewrites: a's b, b's a, c's d, d's c.
  bool slots(const Prop &p, Slots &s) const {
    if (groups > 1) return sym_slots(p, s);
    const int x0 = find(p.a, p.b), x1 = find(p.b, p.a), x2 = find(p.c, p.d), x3 = find(p.d, p.c);
    if (x0 < 0 || x1 < 0 || x2 < 0 || x3 < 0) return false;
    s.n = 4;
    s.at[0] = p.a * d + x0;
    s.at[1] = p.b * d + x1;
    s.at[2] = p.c * d + x2;
    s.at[3] = p.d * d + x3;
    s.val[0] = p.r ? p.c : p.d;
    s.val[1] = p.r ? p.d : p.c;
    s.val[2] = p.r ? p.a : p.b;
    s.val[3] = p.r ? p.b : p.a;
    return true;
  }
  void apply(const Prop &p, const Slots &s) {
    if (groups > 1) {
      sym_apply(p, s);
      return;
    }
    for (int k = 0; k < 4; ++k) adj[s.at[k]] = s.val[k];
    if (!lineat.empty()) {  // a's and d's (r=0) or a's and c's (r=1) new edges are line t0
      lineat[s.at[0]] = p.t0;
      lineat[s.at[1]] = p.t1;
      lineat[s.at[2]] = p.r ? p.t0 : p.t1;
      lineat[s.at[3]] = p.r ? p.t1 : p.t0;
    }
    set.remove(ekey(p.a, p.b));
    set.remove(ekey(p.c, p.d));
    const int e0a = p.a, e0b = p.r ? p.c : p.d, e1a = p.r ? p.b : p.c, e1b = p.r ? p.d : p.b;
    set.add(ekey(e0a, e0b));
    set.add(ekey(e1a, e1b));
    edges[2 * p.t0] = e0a;
    edges[2 * p.t0 + 1] = e0b;
    edges[2 * p.t1] = e1a;
    edges[2 * p.t1 + 1] = e1b;
  }
};

// Per-vertex statistics from distance histograms (refreshed every few
// accepts): h_1..h_7 / n, S_v / mean(S) - 1, ecc_v / max ecc, and the mean of
// S_v / mean(S) - 1 over v's neighbours.
struct VertexStats {
  int maxd = 8;
  std::vector<int32_t> hist;
  std::vector<float> host;  // [n][kVStat]
  int refresh(const SaGraph &g) {
    const int n = g.n;
    hist.assign((size_t)n * (maxd + 1), 0);
    int rc = gg_distance_histogram(n, g.d, g.adj.data(), maxd, hist.data());
    if (rc) return rc;
    std::vector<double> S(n);
    std::vector<int> ecc(n);
    double tot = 0;
    int diam = 1;
    for (int v = 0; v < n; ++v) {
      const int32_t *h = &hist[(size_t)v * (maxd + 1)];
      double s = 0;
      int e = 0;
      for (int k = 1; k <= maxd; ++k) {
        s += (double)k * h[k];
        if (h[k]) e = k;
      }
      S[v] = s;
      ecc[v] = e;
      tot += s;
      diam = std::max(diam, e);
    }
    const double mean = std::max(tot / n, 1.0);
    host.assign((size_t)n * kVStat, 0.0f);
    for (int v = 0; v < n; ++v) {
      float *o = &host[(size_t)v * kVStat];
      const int32_t *h = &hist[(size_t)v * (maxd + 1)];
      for (int k = 1; k <= 7; ++k) o[k - 1] = k <= maxd ? (float)h[k] / n : 0.0f;
      o[7] = (float)(S[v] / mean - 1.0);
      o[8] = (float)ecc[v] / diam;
    }
    for (int v = 0; v < n; ++v) {
      double acc = 0;
      for (int j = 0; j < g.d; ++j) acc += host[(size_t)g.adj[(size_t)v * g.d + j] * kVStat + 7];
      host[(size_t)v * kVStat + 9] = (float)(acc / g.d);
    }
    return GG_OK;
  }
};

double spearman(const std::vector<double> &x, const std::vector<double> &y) {
  const size_t n = x.size();
  if (n < 3) return 0;
  auto ranks = [&](const std::vector<double> &v) {
    std::vector<size_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](size_t i, size_t j) { return v[i] < v[j]; });
    std::vector<double> r(n);
    for (size_t i = 0; i < n; ++i) r[idx[i]] = (double)i;
    return r;
  };
  std::vector<double> rx = ranks(x), ry = ranks(y);
  double mx = (n - 1) / 2.0, sxy = 0, sxx = 0, syy = 0;
  for (size_t i = 0; i < n; ++i) {
    sxy += (rx[i] - mx) * (ry[i] - mx);
    sxx += (rx[i] - mx) * (rx[i] - mx);
    syy += (ry[i] - mx) * (ry[i] - mx);
  }
  return (sxx > 0 && syy > 0) ? sxy / std::sqrt(sxx * syy) : 0;
}

// A candidate graph: the current graph plus nm vertex-disjoint moves.
struct Cand {
  int nm = 0;
  int prop[kMaxCombine];
  Slots sl[kMaxCombine];
};

template <class T>
struct DevArr {
  T *p = nullptr;
  ~DevArr() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t count) {
    if (hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess)
      return set_error(GG_ERR_HIP, "device allocation of %zu bytes failed", count * sizeof(T));
    return GG_OK;
  }
};

struct PlanHolder {
  gg_plan *p = nullptr;
  ~PlanHolder() {
    if (p) gg_plan_destroy(p);
  }
};

struct MlpHolder {
  gg_mlp *p = nullptr;
  ~MlpHolder() {
    if (p) gg_mlp_destroy(p);
  }
};

}  // namespace

extern "C" int gg_apply_move(int lines, int32_t *edges, int line0, int line1, int variant) {
  if (!edges || line0 < 0 || line1 < 0 || line0 >= lines || line1 >= lines || line0 == line1)
    return set_error(GG_ERR_INVALID, "bad move");
  const int a = edges[2 * line0], b = edges[2 * line0 + 1], c = edges[2 * line1], d = edges[2 * line1 + 1];
  edges[2 * line0] = a;
  edges[2 * line0 + 1] = variant ? c : d;
  edges[2 * line1] = variant ? b : c;
  edges[2 * line1 + 1] = variant ? d : b;
  return GG_OK;
}

extern "C" int gg_apply_move_sym(int nodes, int degree, int groups, int grid_height, int32_t *edges,
                                 int line0, int line1, int variant) {
  if (!edges || nodes < 4 || degree < 1) return set_error(GG_ERR_INVALID, "bad move arguments");
  const int lines = (int)((int64_t)nodes * degree / 2);
  if (groups <= 1) return gg_apply_move(lines, edges, line0, line1, variant);
  if (groups > kMaxGroups || nodes % groups || line0 < 0 || line1 < 0 || line0 >= lines ||
      line1 >= lines || (variant != 0 && variant != 1))
    return set_error(GG_ERR_INVALID, "bad symmetric move");
  SaGraph g(nodes, degree, edges);
  int rc = g.init();
  if (!rc) rc = g.sym_setup(groups, grid_height > 0 ? grid_height : 0);
  if (rc) return rc;
  Prop p{line0, line1, variant, 0, 0, 0, 0};
  Slots s;
  if (!g.load(p) || !g.sym_lines_ok(p) || !g.slots(p, s))
    return set_error(GG_ERR_INVALID, "lines %d, %d do not make a symmetric move", line0, line1);
  g.apply(p, s);
  return GG_OK;
}

// The device proposal draw of ranked steps, for tests: `count` proposals of
// the graph (edges [lines][2], adjacency [n][d]) for (seed, step).
extern "C" int gg_sa_proposals(int nodes, int degree, const int32_t *edges, const int32_t *adj,
                               int count, uint64_t seed, uint64_t step, int grid_height,
                               int grid_length, int32_t *props, int32_t *lines_out) {
  if (!edges || !adj || !props || !lines_out || nodes < 4 || degree < 1 || count < 1)
    return set_error(GG_ERR_INVALID, "bad proposal arguments");
  if (grid_height > 0 && (nodes % grid_height || grid_length < 1))
    return set_error(GG_ERR_INVALID, "bad grid shape");
  int rc = gg_init(-1);
  if (rc) return rc;
  const int lines = nodes * degree / 2;
  DevArr<int32_t> e_d, a_d, p_d, l_d, la_d, of_d;
  if ((rc = e_d.alloc(2 * (size_t)lines)) || (rc = a_d.alloc((size_t)nodes * degree)) ||
      (rc = p_d.alloc((size_t)count * 5)) || (rc = l_d.alloc((size_t)count * 3)))
    return rc;
  hipStream_t es = gg::engine_stream();
  BSA_HIP(hipMemcpyAsync(e_d.p, edges, 2 * (size_t)lines * 4, hipMemcpyHostToDevice, es));
  BSA_HIP(hipMemcpyAsync(a_d.p, adj, (size_t)nodes * degree * 4, hipMemcpyHostToDevice, es));
  if (grid_height > 0) {
    std::vector<int32_t> e(edges, edges + 2 * (size_t)lines);
    SaGraph g(nodes, degree, e.data());
    if ((rc = g.init())) return rc;
    g.grid_height = grid_height;
    g.grid_length = grid_length;
    g.grid_setup();
    if ((rc = la_d.alloc(g.lineat.size())) || (rc = of_d.alloc(g.offsets.size()))) return rc;
    BSA_HIP(hipMemcpy(la_d.p, g.lineat.data(), g.lineat.size() * 4, hipMemcpyHostToDevice));
    BSA_HIP(hipMemcpy(of_d.p, g.offsets.data(), g.offsets.size() * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(gg_bsa_propose_grid, dim3((count + 255) / 256), dim3(256), 0, es, count, seed,
                       step, lines, (const int32_t *)e_d.p, (const int32_t *)a_d.p,
                       (const int32_t *)la_d.p, degree, nodes, grid_height, grid_length,
                       (const int32_t *)of_d.p, (int)g.offsets.size(), p_d.p, l_d.p);
  } else {
    hipLaunchKernelGGL(gg_bsa_propose, dim3((count + 255) / 256), dim3(256), 0, es, count, seed, step,
                       lines, (const int32_t *)e_d.p, (const int32_t *)a_d.p, degree, p_d.p, l_d.p);
  }
  BSA_HIP(hipGetLastError());
  BSA_HIP(hipMemcpyAsync(props, p_d.p, (size_t)count * 5 * 4, hipMemcpyDeviceToHost, es));
  BSA_HIP(hipMemcpyAsync(lines_out, l_d.p, (size_t)count * 3 * 4, hipMemcpyDeviceToHost, es));
  BSA_HIP(hipStreamSynchronize(es));
  return GG_OK;
}

extern "C" int gg_sa_features(int nodes, int degree, const int32_t *adj, int count,
                              const int32_t *props, const float *vstat, float *feats) {
  if (!adj || !props || !feats || nodes < 2 || degree < 1 || count < 1)
    return set_error(GG_ERR_INVALID, "bad feature arguments");
  for (size_t i = 0; i < (size_t)nodes * degree; ++i)
    if (adj[i] < 0 || adj[i] >= nodes) return set_error(GG_ERR_INVALID, "adjacency entry out of range");
  for (int i = 0; i < count; ++i)
    for (int k = 0; k < 5; ++k) {
      const int v = props[5 * i + k];
      if (k < 4 ? (v < 0 || v >= nodes) : (v != 0 && v != 1))
        return set_error(GG_ERR_INVALID, "proposal %d out of range", i);
    }
  int rc = gg_init(-1);
  if (rc) return rc;
  hipStream_t es = gg::engine_stream();
  DevArr<int32_t> adj_d, props_d;
  DevArr<float> vstat_d, feats_d;
  if ((rc = adj_d.alloc((size_t)nodes * degree)) || (rc = props_d.alloc((size_t)count * 5)) ||
      (rc = vstat_d.alloc((size_t)nodes * kVStat)) || (rc = feats_d.alloc((size_t)count * kFeat)))
    return rc;
  BSA_HIP(hipMemcpyAsync(adj_d.p, adj, (size_t)nodes * degree * 4, hipMemcpyHostToDevice, es));
  BSA_HIP(hipMemcpyAsync(props_d.p, props, (size_t)count * 5 * 4, hipMemcpyHostToDevice, es));
  if (vstat)
    BSA_HIP(hipMemcpyAsync(vstat_d.p, vstat, (size_t)nodes * kVStat * 4, hipMemcpyHostToDevice, es));
  else
    BSA_HIP(hipMemsetAsync(vstat_d.p, 0, (size_t)nodes * kVStat * 4, es));
  hipLaunchKernelGGL(gg_bsa_features, dim3(count), dim3(kBlock), 2 * (size_t)((nodes + 31) / 32) * 4, es,
                     (const int32_t *)adj_d.p, nodes, degree, (const int32_t *)props_d.p,
                     (const float *)vstat_d.p, feats_d.p);
  BSA_HIP(hipGetLastError());
  BSA_HIP(hipMemcpyAsync(feats, feats_d.p, (size_t)count * kFeat * 4, hipMemcpyDeviceToHost, es));
  BSA_HIP(hipStreamSynchronize(es));
  return GG_OK;
}

extern "C" int gg_sa_batched(int nodes, int degree, int32_t *edges, const gg_bsa_params *P,
                             gg_bsa_stats *stats, double *log_time, double *log_best_aspl,
                             int32_t *log_best_diam, int32_t *accepted, int32_t *accepted_diam,
                             uint64_t *accepted_sum, int accepted_capacity) {
  if (!edges || !P || nodes < 8 || degree < 2 || ((int64_t)nodes * degree) % 2)
    return set_error(GG_ERR_INVALID, "bad batched SA arguments");
  if (P->combine < 1 || P->combine > kMaxCombine)
    return set_error(GG_ERR_INVALID, "combine must be in 1..%d", kMaxCombine);
  const int E = std::max(1, P->exact);
  const bool use_sur = P->use_surrogate != 0;
  const int NP = use_sur ? std::max(P->proposals, 1) : 0;
  const int npmax = std::max(NP, use_sur ? std::max(E, std::max(64, 8 * E)) : E);
  const double t0 = now_s();
  double t_exact = 0, t_sur = 0, t_feat = 0, t_train = 0;
  int rc = gg_init(-1);
  if (rc) return rc;
  SaGraph g(nodes, degree, edges);
  if ((rc = g.init())) return rc;
  const int groups = std::max(1, P->groups);
  if (groups > kMaxGroups || nodes % groups || (nodes * (int64_t)degree / 2) % groups)
    return set_error(GG_ERR_INVALID, "groups %d must divide nodes %d and lines (at most %d)", groups,
                     nodes, kMaxGroups);
  if (groups > 1 && P->grid_height > 0 && groups != 2 && groups != 4)
    return set_error(GG_ERR_INVALID, "grid graphs take 2 or 4 groups");
  if (groups > 1 && P->incremental)
    return set_error(GG_ERR_INVALID, "the incremental scorer takes groups = 1");
  if (groups > 1 && (rc = g.sym_setup(groups, P->grid_height > 0 ? P->grid_height : 0))) return rc;
  if (P->grid_height > 0) {
    if (nodes % P->grid_height || P->grid_length < 1)
      return set_error(GG_ERR_INVALID, "grid height %d / length %d do not fit %d nodes", P->grid_height,
                       P->grid_length, nodes);
    g.grid_height = P->grid_height;
    g.grid_length = P->grid_length;
    for (int i = 0; i < g.lines; ++i)
      if (!g.short_enough(edges[2 * i], edges[2 * i + 1]))
        return set_error(GG_ERR_INVALID, "line %d is longer than %d", i, P->grid_length);
    g.grid_setup();
  }
  std::vector<int32_t> best_edges(edges, edges + 2 * (size_t)g.lines);
  Rng rng(P->seed);
  hipStream_t es = gg::engine_stream();
  // Incremental scorer for single-move candidates (plain runs only).
  struct IncrGuard {
    gg_incr *p = nullptr;
    ~IncrGuard() { gg_incr_destroy(p); }
  } incr;
  bool incr_pending = false;
  // With the surrogate, the scorer serves the explore steps until the first
  // ranked step (merged candidates would flag nearly every source).
  if (P->incremental) {
    incr.p = gg::incr_create_unchecked(nodes, degree, g.adj.data());
    if (!incr.p) return GG_ERR_INVALID;
  }
  // The move of a single-move candidate as (removed, added) edges.
  auto move_edges = [](const Prop &p, int32_t rem[4], int32_t add[4]) {
    rem[0] = p.a;
    rem[1] = p.b;
    rem[2] = p.c;
    rem[3] = p.d;
    add[0] = p.a;
    add[1] = p.r ? p.c : p.d;
    add[2] = p.r ? p.b : p.c;
    add[3] = p.r ? p.d : p.b;
  };

  // Symmetric graphs are scored from their base sources with the groups
  // weights; a grid's base quadrant goes first through its rotate_hash
  // (create_rotate_hash, grid/common.c:312-341).
  std::vector<int32_t> rhash;
  if (groups > 1 && g.grid_height > 0) {
    const int h = g.grid_height, w = nodes / h, B = nodes / groups, bh = h / 2;
    rhash.assign(nodes, -1);
    for (int i = 0; i < B; ++i) {
      const int j = (i / bh) * h + (i % bh);
      rhash[j] = i;
      for (int q = 1; q < groups; ++q) rhash[g.rot(q, j)] = i + q * B;
    }
    for (int v = 0; v < nodes; ++v)
      if (rhash[v] < 0) return set_error(GG_ERR_INVALID, "grid %dx%d has no %d-fold base quadrant", w, h, groups);
  }
  // A general graph the row order would order is numbered by the
  // bank-balanced layout instead (gg_bank_layout), once: the plan relabels
  // every candidate with it, and the driver's adjacency takes the layout's
  // order of each row's entries (the same graph), so candidates' patches
  // keep every row in that order.  The plan then orders nothing itself.
  std::vector<int32_t> lay;
  const int dp = (degree + 7) / 8 * 8;
  if (groups == 1 && rhash.empty() && gg::order_rows_applies(nodes, dp)) {
    int32_t cfg[4];
    if ((rc = gg_tune_lookup(nodes, degree, nodes, 1, 0, E > 1, cfg))) return rc;
    std::vector<int32_t> ordered(g.adj.size());
    uint64_t cost[2];
    lay.resize(nodes);
    // A light annealing (20 swaps per vertex): the run pays for it once.
    if ((rc = gg_bank_layout(nodes, degree, g.adj.data(), cfg[0], (int64_t)nodes * 20, 1, 0, lay.data(),
                             ordered.data(), cost)))
      return rc;
    g.adj = ordered;
  }
  gg_graph shape = {nodes, degree, nodes / groups, groups, 0, GG_GRAPH_SYMMETRIC, g.adj.data(),
                    !lay.empty() ? lay.data() : rhash.empty() ? nullptr : rhash.data()};
  const int max_patch = 4 * groups * P->combine;
  PlanHolder plan;
  const int prev_order = lay.empty() ? -1 : gg_set_row_order(0);
  plan.p = gg_plan_create_patched(E, &shape, max_patch);
  if (prev_order >= 0) gg_set_row_order(prev_order);
  if (!plan.p) return gg_last_error_code() ? gg_last_error_code() : GG_ERR_HIP;
  std::vector<gg_result> res(E);
  std::vector<int32_t> patches((size_t)E * max_patch * 2, -1);
  rc = gg_plan_set_base(plan.p, g.adj.data());
  if (!rc) rc = gg_plan_set_patches(plan.p, patches.data());
  if (!rc) rc = gg_plan_run(plan.p, res.data());
  if (rc) return rc;
  const double npairs = ((double)nodes - 1) * nodes;
  int cur_diam = res[0].diameter;
  uint64_t cur_S = res[0].sum;
  int best_diam = cur_diam;
  uint64_t best_S = cur_S;
  int64_t evals = 1, accepts = 0, step = 0, n_logged = 0, moves_acc = 0;
  double to_target = -1;

  // surrogate state
  MlpHolder mlp;
  VertexStats vs;
  vs.maxd = std::min(std::max(P->hist_maxd, 2), 15);
  const int cap = 16384;
  const int bt = std::max(P->batch_train, 32);
  DevArr<int32_t> props_d, idx_d, tidx_d, edges_d, plines_d, upd_d, lineat_d, offsets_d, rotk_d,
      rotline_d;
  // Lines (2 per copy) and grid line-table entries (4 per copy) of an
  // accepted candidate's moves.
  const size_t upd_rows = (size_t)6 * groups * kMaxCombine;
  // Explore steps of surrogate runs draw on the device too: this many
  // proposals, of which the first valid ones become the singles.
  const int nexp = std::max(64, 8 * E);
  struct Pinned {
    int32_t *props = nullptr, *lines = nullptr, *upd = nullptr;
    ~Pinned() {
      if (props) (void)hipHostFree(props);
      if (lines) (void)hipHostFree(lines);
      if (upd) (void)hipHostFree(upd);
    }
  } pin;
  DevArr<float> vstat_d, feats_d, replay_d, batch_d;
  std::vector<double> by(cap);  // raw labels of the replay rows
  int nbuf = 0, head = 0;
  double scale = 0;
  std::vector<double> corr_pred, corr_true;
  if (use_sur) {
    const int dims[5] = {kFeat, 256, 256, 256, 1};
    mlp.p = gg_mlp_create(4, dims, P->seed ^ 0x5eedull);
    if (!mlp.p) return GG_ERR_HIP;
    if ((rc = props_d.alloc((size_t)npmax * 5)) || (rc = feats_d.alloc((size_t)npmax * kFeat)) ||
        (rc = replay_d.alloc((size_t)cap * kFeat)) || (rc = batch_d.alloc((size_t)bt * kFeat)) ||
        (rc = vstat_d.alloc((size_t)nodes * kVStat)) ||
        (rc = idx_d.alloc(2 * (size_t)std::max(npmax, bt))) || (rc = tidx_d.alloc((size_t)bt)) ||
        (rc = edges_d.alloc(2 * (size_t)g.lines)) || (rc = plines_d.alloc((size_t)npmax * 3)) ||
        (rc = upd_d.alloc(3 * upd_rows)))
      return rc;
    if (hipHostMalloc((void **)&pin.props, (size_t)npmax * 5 * 4) != hipSuccess ||
        hipHostMalloc((void **)&pin.lines, (size_t)npmax * 3 * 4) != hipSuccess ||
        hipHostMalloc((void **)&pin.upd, 3 * upd_rows * 4) != hipSuccess)
      return set_error(GG_ERR_HIP, "pinned allocation failed");
    if (groups > 1) {  // the rotation tables of the symmetric draws
      if ((rc = rotk_d.alloc(g.rotk.size())) || (rc = rotline_d.alloc(g.rotline.size()))) return rc;
      if (hipMemcpy(rotk_d.p, g.rotk.data(), g.rotk.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
          hipMemcpy(rotline_d.p, g.rotline.data(), g.rotline.size() * 4, hipMemcpyHostToDevice) !=
              hipSuccess)
        return set_error(GG_ERR_HIP, "rotation table upload failed");
    }
    if (hipMemcpy(edges_d.p, edges, 2 * (size_t)g.lines * 4, hipMemcpyHostToDevice) != hipSuccess)
      return set_error(GG_ERR_HIP, "edge list upload failed");
    if (g.grid_height > 0) {  // the line table and length ball of the grid draws
      if ((rc = lineat_d.alloc(g.lineat.size())) || (rc = offsets_d.alloc(g.offsets.size())))
        return rc;
      if (hipMemcpy(lineat_d.p, g.lineat.data(), g.lineat.size() * 4, hipMemcpyHostToDevice) !=
              hipSuccess ||
          hipMemcpy(offsets_d.p, g.offsets.data(), g.offsets.size() * 4, hipMemcpyHostToDevice) !=
              hipSuccess)
        return set_error(GG_ERR_HIP, "grid table upload failed");
    }
    const double tf = now_s();
    rc = vs.refresh(g);
    if (!rc && hipMemcpy(vstat_d.p, vs.host.data(), vs.host.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
      rc = set_error(GG_ERR_HIP, "vertex statistics upload failed");
    t_feat += now_s() - tf;
    if (rc) return rc;
  }
  const size_t feat_lds = 2 * (size_t)((nodes + 31) / 32) * 4;

  std::vector<Prop> props(npmax);
  std::vector<int32_t> props_h((size_t)npmax * 5);
  std::vector<float> pred(npmax, 0.0f), noise(npmax, 0.0f);
  std::vector<int> order(npmax);
  std::vector<Cand> cand(E);
  std::vector<int> cand_verts;
  std::vector<int32_t> src_rows, dst_rows;
  std::vector<int32_t> idx_h(2 * (size_t)std::max(npmax, bt)), tidx_h(bt);
  std::vector<float> yb(bt);
  int m_cur = P->combine;
  int since_refresh = 0;
  const double maxT = P->max_temp, minT = P->min_temp > 0 ? P->min_temp : 1e-12;

  for (step = 0; step < P->max_steps; ++step) {
    if (P->time_budget_s > 0 && now_s() - t0 > P->time_budget_s) break;
    const double temp =
        P->max_steps > 1 ? maxT * std::pow(minT / maxT, (double)step / (P->max_steps - 1)) : maxT;
    const bool explore = !use_sur || step < P->warmup ||
                         (P->explore_every > 0 && (step - P->warmup) % P->explore_every == P->explore_every - 1);
    // --- proposals
    double tp = now_s();
    // Surrogate runs draw every proposal on the device (ranked steps: NP to
    // rank; explore steps: nexp, the first valid ones becoming the singles);
    // plain runs draw their singles on the host.
    const bool dev_props = use_sur;
    const int np = use_sur ? (explore ? nexp : NP) : E;
    for (int i = 0; i < (dev_props ? 0 : np); ++i) {
      Prop &p = props[i];
      if (explore) {
        Slots s;
        if (!g.draw_valid(rng, p, 1000000) || !g.slots(p, s))
          return set_error(GG_ERR_INVALID, "no valid 2-opt move found");
        cand[i].nm = 1;
        cand[i].prop[0] = i;
        cand[i].sl[0] = s;
      } else if (g.grid_height > 0) {
        if (!g.draw_valid(rng, p, 1000000)) return set_error(GG_ERR_INVALID, "no valid 2-opt move found");
      } else {
        g.draw(rng, p);
      }
      int32_t *o = &props_h[(size_t)i * 5];
      o[0] = p.a;
      o[1] = p.b;
      o[2] = p.c;
      o[3] = p.d;
      o[4] = p.r;
    }
    t_feat += now_s() - tp;
    // --- features + ranking on the device
    if (use_sur) {
      const double ts = now_s();
      if (dev_props && groups > 1) {
        hipLaunchKernelGGL(gg_bsa_propose_sym, dim3((np + 255) / 256), dim3(256), 0, es, np, P->seed,
                           (uint64_t)step, g.lines, (const int32_t *)edges_d.p,
                           (const int32_t *)gg_plan_base_device(plan.p), degree, nodes, groups,
                           (const int32_t *)rotk_d.p, (const int32_t *)rotline_d.p,
                           (const int32_t *)lineat_d.p, g.grid_height, g.grid_length,
                           (const int32_t *)offsets_d.p, (int)g.offsets.size(), props_d.p, plines_d.p);
        BSA_HIP(hipGetLastError());
      } else if (dev_props && g.grid_height > 0) {
        hipLaunchKernelGGL(gg_bsa_propose_grid, dim3((np + 255) / 256), dim3(256), 0, es, np,
                           P->seed, (uint64_t)step, g.lines, (const int32_t *)edges_d.p,
                           (const int32_t *)gg_plan_base_device(plan.p), (const int32_t *)lineat_d.p,
                           degree, nodes, g.grid_height, g.grid_length, (const int32_t *)offsets_d.p,
                           (int)g.offsets.size(), props_d.p, plines_d.p);
        BSA_HIP(hipGetLastError());
      } else if (dev_props) {
        hipLaunchKernelGGL(gg_bsa_propose, dim3((np + 255) / 256), dim3(256), 0, es, np, P->seed,
                           (uint64_t)step, g.lines, (const int32_t *)edges_d.p,
                           (const int32_t *)gg_plan_base_device(plan.p), degree, props_d.p,
                           plines_d.p);
        BSA_HIP(hipGetLastError());
      } else {
        BSA_HIP(hipMemcpyAsync(props_d.p, props_h.data(), (size_t)np * 5 * 4, hipMemcpyHostToDevice, es));
      }
      hipLaunchKernelGGL(gg_bsa_features, dim3(np), dim3(kBlock), feat_lds, es,
                         (const int32_t *)gg_plan_base_device(plan.p), nodes, degree,
                         (const int32_t *)props_d.p, (const float *)vstat_d.p, feats_d.p);
      BSA_HIP(hipGetLastError());
      BSA_HIP(hipStreamSynchronize(es));
      if ((rc = gg::mlp_forward_device(mlp.p, feats_d.p, np, pred.data()))) return rc;
      if (dev_props) {
        // The drawn moves come back for candidate assembly; invalid ones
        // (no vertex-disjoint pair in 64 draws, or an added edge that
        // exists) rank last.
        BSA_HIP(hipMemcpyAsync(pin.props, props_d.p, (size_t)np * 5 * 4, hipMemcpyDeviceToHost, es));
        BSA_HIP(hipMemcpyAsync(pin.lines, plines_d.p, (size_t)np * 3 * 4, hipMemcpyDeviceToHost, es));
        BSA_HIP(hipStreamSynchronize(es));
        for (int i = 0; i < np; ++i) {
          Prop &p = props[i];
          const int32_t *o = pin.props + 5 * (size_t)i, *l = pin.lines + 3 * (size_t)i;
          p.a = o[0];
          p.b = o[1];
          p.c = o[2];
          p.d = o[3];
          p.r = o[4];
          p.t0 = l[0];
          p.t1 = l[1];
          if (!l[2]) pred[i] = 1e30f;
        }
      }
      t_sur += now_s() - ts;
    }
    std::vector<float> &rank = (use_sur && P->random_ranking && !explore) ? noise : pred;
    if (&rank == &noise)
      for (int i = 0; i < np; ++i) noise[i] = -1.0f - (float)rng.below(1u << 24) * (1.0f / (1u << 24));
    // --- candidates
    int ncand = explore ? E : 0;
    if (explore && dev_props) {
      // The first valid device draws become this step's singles.
      ncand = 0;
      for (int i = 0; i < np && ncand < E; ++i) {
        Slots sl;
        if (pred[i] >= 1e29f || !g.fresh(props[i]) || !g.slots(props[i], sl)) continue;
        cand[ncand].nm = 1;
        cand[ncand].prop[0] = i;
        cand[ncand].sl[0] = sl;
        ++ncand;
      }
    }
    if (!explore) {
      const double tc = now_s();
      std::iota(order.begin(), order.begin() + np, 0);
      const int want = std::min(np, E * m_cur * 4 + 64);
      std::partial_sort(order.begin(), order.begin() + want, order.begin() + np,
                        [&](int x, int y) { return rank[x] < rank[y]; });
      for (int c = 0; c < E; ++c) cand[c].nm = 0;
      // Greedy fill in rank order: candidate c takes up to m_cur valid moves
      // that share no vertex; a move predicted to raise S closes a non-empty
      // candidate (and may then start the next one alone).
      int c = 0;
      cand_verts.clear();
      for (int oi = 0; oi < want && c < E; ++oi) {
        const int i = order[oi];
        const Prop &p = props[i];
        if (cand[c].nm > 0 && rank[i] >= 0.0f) {
          cand_verts.clear();
          if (++c >= E) break;
        }
        int mv[4 * kMaxGroups];
        const int nmv = g.move_vertices(p, mv);
        bool clash = false;
        for (int v : cand_verts)
          for (int k = 0; k < nmv; ++k) clash |= v == mv[k];
        Slots s;
        if (clash || !g.fresh(p) || !g.slots(p, s)) continue;
        Cand &C = cand[c];
        C.prop[C.nm] = i;
        C.sl[C.nm] = s;
        ++C.nm;
        cand_verts.insert(cand_verts.end(), mv, mv + nmv);
        if (C.nm >= m_cur) {
          cand_verts.clear();
          ++c;
        }
      }
      ncand = std::min(E, c + (c < E && cand[c].nm > 0 ? 1 : 0));
      t_feat += now_s() - tc;
    }
    // --- exact scoring of all candidates in one plan launch
    std::fill(patches.begin(), patches.end(), -1);
    for (int c = 0; c < ncand; ++c) {
      size_t at = (size_t)c * max_patch;
      for (int k = 0; k < cand[c].nm; ++k)
        for (int j = 0; j < cand[c].sl[k].n; ++j, ++at) {
          patches[at * 2] = cand[c].sl[k].at[j];
          patches[at * 2 + 1] = cand[c].sl[k].val[j];
        }
    }
    const double te = now_s();
    if (incr.p && !explore) {  // the first ranked step retires the scorer
      gg_incr_destroy(incr.p);
      incr.p = nullptr;
    }
    if (incr.p) {
      // One candidate stays pending for the accept decision; with several,
      // each is scored and dropped, and the accepted one is scored again.
      for (int c = 0; c < ncand; ++c) {
        int32_t rem[4], add[4];
        move_edges(props[cand[c].prop[0]], rem, add);
        if ((rc = gg_incr_eval(incr.p, rem, add, &res[c], nullptr))) return rc;
        if (ncand > 1 && (rc = gg_incr_reject(incr.p))) return rc;
      }
      incr_pending = ncand == 1;
    } else {
      rc = gg_plan_set_patches(plan.p, patches.data());
      if (!rc) rc = gg_plan_launch(plan.p);
      if (rc) return rc;
    }
    // While the candidates are scored, train on the rows stored so far (the
    // network's own stream; this step's rows join the buffer afterwards).
    if (use_sur && nbuf >= 32 && P->train_steps > 0) {
      const double tt = now_s();
      const int b = std::min(bt, nbuf);
      hipStream_t ms = gg::mlp_stream(mlp.p);
      for (int s = 0; s < P->train_steps; ++s) {
        for (int i = 0; i < b; ++i) {
          const int j = (int)rng.below((uint32_t)nbuf);
          tidx_h[i] = j;
          yb[i] = (float)(by[j] / scale);
        }
        BSA_HIP(hipMemcpyAsync(tidx_d.p, tidx_h.data(), (size_t)b * 4, hipMemcpyHostToDevice, ms));
        hipLaunchKernelGGL(gg_bsa_copy_rows, dim3((b * kFeat + 255) / 256), dim3(256), 0, ms,
                           (const float *)replay_d.p, (const int32_t *)tidx_d.p, batch_d.p,
                           (const int32_t *)nullptr, b);
        BSA_HIP(hipGetLastError());
        float loss = 0;
        if ((rc = gg::mlp_train_device(mlp.p, batch_d.p, yb.data(), b, P->lr, &loss))) return rc;
      }
      t_train += now_s() - tt;
    }
    if (!incr.p) rc = gg_plan_wait(plan.p, res.data());
    t_exact += now_s() - te;
    if (rc) return rc;
    evals += ncand;
    int bi = -1;
    for (int c = 0; c < ncand; ++c) {
      if (cand[c].nm == 0 || !res[c].connected || res[c].diameter >= nodes) continue;
      if (bi < 0 || res[c].diameter < res[bi].diameter ||
          (res[c].diameter == res[bi].diameter && res[c].sum < res[bi].sum))
        bi = c;
    }
    // --- training rows: singles keep their exact dS; a merged candidate's
    // members share its prediction error equally.
    if (use_sur) {
      const double ts = now_s();
      src_rows.clear();
      dst_rows.clear();
      for (int c = 0; c < ncand; ++c) {
        const Cand &C = cand[c];
        if (C.nm == 0 || !res[c].connected || res[c].diameter >= nodes) continue;
        const double dS = (double)res[c].sum - (double)cur_S;
        double psum = 0;
        for (int k = 0; k < C.nm; ++k) psum += pred[C.prop[k]];
        const double resid = (dS - scale * psum) / C.nm;
        for (int k = 0; k < C.nm; ++k) {
          src_rows.push_back(C.prop[k]);
          dst_rows.push_back(head);
          by[head] = C.nm == 1 ? dS : scale * pred[C.prop[k]] + resid;
          head = (head + 1) % cap;
          nbuf = std::min(nbuf + 1, cap);
        }
        const double mag = std::fabs(dS) / C.nm + 1.0;
        scale = scale == 0 ? mag : 0.995 * scale + 0.005 * mag;
        if (explore && step >= P->warmup && C.nm == 1) {
          corr_pred.push_back(pred[C.prop[0]]);
          corr_true.push_back(dS);
        }
      }
      const int rows = (int)src_rows.size();
      if (rows) {
        std::copy(src_rows.begin(), src_rows.end(), idx_h.begin());
        std::copy(dst_rows.begin(), dst_rows.end(), idx_h.begin() + rows);
        BSA_HIP(hipMemcpyAsync(idx_d.p, idx_h.data(), (size_t)2 * rows * 4, hipMemcpyHostToDevice, es));
        hipLaunchKernelGGL(gg_bsa_copy_rows, dim3((rows * kFeat + 255) / 256), dim3(256), 0, es,
                           (const float *)feats_d.p, (const int32_t *)idx_d.p, replay_d.p,
                           (const int32_t *)idx_d.p + rows, rows);
        BSA_HIP(hipGetLastError());
      }
      // (Stored on the evaluator's stream, ahead of the next step's feature
      // kernel that overwrites feats_d.)
      t_sur += now_s() - ts;
    }
    // --- the reference's accept rule (general/sa.c:371-411) on the best candidate
    bool improved = false;
    if (bi >= 0) {
      const int nd = res[bi].diameter;
      const uint64_t nS = res[bi].sum;
      bool acc;
      if (nd < cur_diam) {
        acc = true;
      } else if (nd > cur_diam) {
        acc = false;
      } else if (nS <= cur_S) {
        acc = true;
      } else {
        // general/sa.c:395: (ASPL - ASPL') n (n-1) / groups = -dS / groups
        const double diff = -((double)nS - (double)cur_S) / groups;
        acc = std::exp(diff / temp) > (rng.below(1u << 30) + 1.0) / ((double)(1u << 30) + 2.0);
      }
      improved = nd < cur_diam || (nd == cur_diam && nS < cur_S);
      if (acc) {
        const Cand &C = cand[bi];
        for (int k = 0; k < C.nm; ++k) {
          const Prop &p = props[C.prop[k]];
          g.apply(p, C.sl[k]);
          if (accepted && n_logged < accepted_capacity) {
            accepted[3 * n_logged] = p.t0;
            accepted[3 * n_logged + 1] = p.t1;
            accepted[3 * n_logged + 2] = p.r;
            const bool last = k + 1 == C.nm;
            if (accepted_diam) accepted_diam[n_logged] = last ? nd : -1;
            if (accepted_sum) accepted_sum[n_logged] = last ? nS : 0;
            ++n_logged;
          }
        }
        moves_acc += C.nm;
        if (use_sur) {  // the device edge list follows the accepted moves
          int nu = 0;
          for (int k = 0; k < C.nm; ++k) {
            const Prop &p = props[C.prop[k]];
            for (int q = 0; q < groups; ++q)
              for (int t : {groups > 1 ? g.rline(q, p.t0) : p.t0, groups > 1 ? g.rline(q, p.t1) : p.t1}) {
                pin.upd[3 * nu] = t;
                pin.upd[3 * nu + 1] = edges[2 * t];
                pin.upd[3 * nu + 2] = edges[2 * t + 1];
                ++nu;
              }
          }
          // Grid graphs: the line table entries the moves rewrote, as
          // (-1 - slot, line, 0) records after the lines.
          int nl = 0;
          if (g.grid_height > 0)
            for (int k = 0; k < C.nm; ++k)
              for (int j = 0; j < C.sl[k].n; ++j) {
                const int at = C.sl[k].at[j];
                pin.upd[3 * (nu + nl)] = -1 - at;
                pin.upd[3 * (nu + nl) + 1] = g.lineat[at];
                pin.upd[3 * (nu + nl) + 2] = 0;
                ++nl;
              }
          BSA_HIP(hipMemcpyAsync(upd_d.p, pin.upd, (size_t)(nu + nl) * 3 * 4, hipMemcpyHostToDevice, es));
          hipLaunchKernelGGL(gg_bsa_set_lines, dim3((nu + nl + 255) / 256), dim3(256), 0, es,
                             (const int32_t *)upd_d.p, nu + nl, edges_d.p, lineat_d.p);
          BSA_HIP(hipGetLastError());
        }
        if (incr.p) {
          const double ti = now_s();
          if (!incr_pending) {
            int32_t rem[4], add[4];
            gg_result r2;
            move_edges(props[C.prop[0]], rem, add);
            if ((rc = gg_incr_eval(incr.p, rem, add, &r2, nullptr))) return rc;
          }
          if ((rc = gg_incr_accept(incr.p))) return rc;
          incr_pending = false;
          t_exact += now_s() - ti;
        }
        if ((use_sur || !incr.p) && (rc = gg_plan_set_base(plan.p, g.adj.data()))) return rc;
        cur_diam = nd;
        cur_S = nS;
        ++accepts;
        if (cur_diam < best_diam || (cur_diam == best_diam && cur_S < best_S)) {
          best_diam = cur_diam;
          best_S = cur_S;
          std::copy(edges, edges + 2 * (size_t)g.lines, best_edges.begin());
        }
        if (use_sur && ++since_refresh >= std::max(1, P->feature_refresh)) {
          since_refresh = 0;
          const double tf = now_s();
          rc = vs.refresh(g);
          if (!rc && hipMemcpy(vstat_d.p, vs.host.data(), vs.host.size() * 4, hipMemcpyHostToDevice) !=
                         hipSuccess)
            rc = set_error(GG_ERR_HIP, "vertex statistics upload failed");
          t_feat += now_s() - tf;
          if (rc) return rc;
        }
      }
    }
    if (incr_pending) {  // not accepted
      if ((rc = gg_incr_reject(incr.p))) return rc;
      incr_pending = false;
    }
    if (!explore) m_cur = improved ? std::min(P->combine, m_cur + 1) : std::max(1, m_cur / 2);
    const double best_aspl = (double)best_S / npairs;
    if (log_time && step < P->log_capacity) {
      log_time[step] = now_s() - t0;
      if (log_best_aspl) log_best_aspl[step] = best_aspl;
      if (log_best_diam) log_best_diam[step] = best_diam;
    }
    if (P->target_aspl > 0 && best_aspl <= P->target_aspl) {
      to_target = now_s() - t0;
      ++step;
      break;
    }
  }
  std::copy(best_edges.begin(), best_edges.end(), edges);
  if (stats) {
    stats->steps = step;
    stats->accepts = accepts;
    stats->evaluations = evals;
    stats->best_diameter = best_diam;
    stats->best_aspl = (double)best_S / npairs;
    stats->seconds_total = now_s() - t0;
    stats->seconds_exact = t_exact;
    stats->seconds_surrogate = t_sur;
    stats->seconds_features = t_feat;
    stats->seconds_to_target = to_target;
    stats->surrogate_rank_corr = spearman(corr_pred, corr_true);
    stats->moves_accepted = moves_acc;
    stats->seconds_train = t_train;
  }
  return GG_OK;
}
