// gg_wide.hip -- the bit-parallel BFS for what the LDS-resident tile kernel
// does not serve: graphs above 65,536 vertices (uint32 ids) and the
// reference's bfs() semantics (algo = BFS) on any table.
//
// Reference:
//   * general/evaluation.c:153-263 / :265-373 (matrix_op and its tiled
//     variant): B[v] |= A[adj[v][j]] per level, i.e. bit s of v's word means
//     "v reaches s along adj"; per level the unreached (source, vertex)
//     pairs are added, g-weighted for base rows and 2g for centre rows;
//   * general/evaluation.c:55-150 (bfs): one BFS per source s along adj
//     (out-edges), summing (distance[i]) * (g - i/B) over targets i > s
//     (centres: * g), centre sources over i > s with weight 1, the diameter
//     being the largest eccentricity; "reached" looks only at i > s; ASPL =
//     sum / (n(n-1)/2).
//
// How it is laid out:
//   * One 1024-thread workgroup per tile of 64 sources, persistent over the
//     tiles.  Each workgroup owns two reach arrays of n uint64 words in HBM
//     (current and next level): at these sizes one tile's words (1 MiB for
//     n = 131072) do not fit the 160 KiB LDS.  A level reads the current
//     word of every vertex, ORs in its row's words until it is full, writes
//     the next word and counts; one block reduction per level.
//   * Rows: matrix_op semantics pull along adj, so the rows are adj's own
//     (relabelled, uint32).  bfs() semantics follow out-edges from the
//     source: "s reaches i" grows by the words of i's in-neighbours, so for
//     a table that is not symmetric the rows are the transposed table (CSR
//     built on the device: count, scan, scatter).  A symmetric table's rows
//     serve both.
//   * bfs() semantics count, per vertex v of a tile [f, f+l), only the
//     sources s < v (mask of the v-f lowest bits) with weight (g - v/B) for
//     v < B*g, g for centre targets, 1 in centre tiles, and run until no bit
//     changes (the eccentricities give the diameter).
//   * A symmetric, g-rotation-invariant table with n <= 65536 has bfs() sum
//     = S/2 exactly (every unordered pair once), so algo = BFS is then served
//     by the LDS tile kernels; a check kernel decides on the device.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "gg_device.h"
#include "gg_eval.h"
#include "gg_internal.h"
#include "gg_tile.h"

namespace {

constexpr int kWideBlock = 1024;
constexpr int kWideWaves = kWideBlock / 64;
constexpr int kSemMatrix = 0;  // matrix_op: S over ordered pairs, symmetry weighted
constexpr int kSemBfs = 1;     // bfs(): weighted sum over pairs s < i

struct WideDesc {
  const uint32_t *off;  // CSR offsets [n+1], or null: row v is col[v*d .. v*d+d)
  const uint32_t *col;  // neighbour ids
  uint64_t *scratch;    // [gridDim.x][2][n] reach words
  const int32_t *flags; // [0]: invalid entries (the tiles then report an error)
  int n, d, split, c, B, g, units_base, units;
};

struct WideOut {
  uint64_t sum;
  int32_t connected, diameter, max_ecc, error;
};

// Sum over the workgroup, returned to every thread.  `red` holds
// [2][kWideWaves] per counter; `par` alternates so one barrier per call
// suffices.
template <int K>
__device__ __forceinline__ void block_sum(uint64_t (&v)[K], uint64_t *red, int par) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    uint64_t x = v[k];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    if (lane == 0) red[(par * K + k) * kWideWaves + wave] = x;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    uint64_t s = 0;
#pragma unroll
    for (int w = 0; w < kWideWaves; ++w) s += red[(par * K + k) * kWideWaves + w];
    v[k] = s;
  }
}

template <bool CSR>
__device__ __forceinline__ uint64_t gather_row(const WideDesc &P, const uint64_t *__restrict__ Ra,
                                               int v, uint64_t val, uint64_t full) {
  uint32_t e0, e1;
  if constexpr (CSR) {
    e0 = P.off[v];
    e1 = P.off[v + 1];
  } else {
    e0 = (uint32_t)v * (uint32_t)P.d;
    e1 = e0 + (uint32_t)P.d;
  }
  uint32_t e = e0;
  for (; e + 4 <= e1 && val != full; e += 4) {
    const uint32_t a = P.col[e], b = P.col[e + 1], c = P.col[e + 2], d = P.col[e + 3];
    val |= Ra[a] | Ra[b] | Ra[c] | Ra[d];
  }
  for (; e < e1 && val != full; ++e) val |= Ra[P.col[e]];
  return val;
}

// All levels of tiles blockIdx.x, blockIdx.x + gridDim.x, ...  Tile u <
// units_base takes base sources [64u, min(64u+64, B)); the rest take centre
// sources from n - c.
template <bool CSR, int SEM>
__global__ void __launch_bounds__(kWideBlock)
gg_wide_bfs_kernel(const WideDesc P, UnitOut *__restrict__ outs) {
  __shared__ uint64_t red[2 * 4 * kWideWaves];
  const int tid = threadIdx.x;
  const int n = P.n, split = P.split;
  uint64_t *Ra0 = P.scratch + (size_t)blockIdx.x * 2 * (size_t)n;
  const bool bad = P.flags[0] != 0;
  for (int u = blockIdx.x; u < P.units; u += gridDim.x) {
    const bool centre = u >= P.units_base;
    const int first = centre ? n - P.c + 64 * (u - P.units_base) : 64 * u;
    const int l = centre ? min(64, P.c - 64 * (u - P.units_base)) : min(64, P.B - 64 * u);
    if (bad) {
      if (tid == 0) outs[u] = UnitOut{0, 0, n + 1, 0, 1, 1};
      continue;
    }
    const uint64_t full = l >= 64 ? ~0ull : ((1ull << l) - 1ull);
    uint64_t *Ra = Ra0, *Rb = Ra0 + n;
    __syncthreads();  // the previous tile's last reads of its arrays are done
    // Level 0: each source reaches itself.  bfs(): M = the weighted number
    // of counted pairs (source s < target v).
    uint64_t m[1] = {0};
    for (int v = tid; v < n; v += kWideBlock) {
      const int rel = v - first;
      Ra[v] = (rel >= 0 && rel < l) ? (1ull << rel) : 0ull;
      if constexpr (SEM == kSemBfs) {
        const uint64_t cnt = (uint64_t)min(max(rel, 0), l);
        const uint64_t w = centre ? 1ull : (v < split ? (uint64_t)(P.g - v / P.B) : (uint64_t)P.g);
        m[0] += w * cnt;
      }
    }
    int par = 0;
    block_sum<1>(m, red, par);  // also the barrier after the level-0 words
    par ^= 1;
    const uint64_t M = m[0];
    const uint64_t lo_full = (uint64_t)split * (uint64_t)l, hi_full = (uint64_t)P.c * (uint64_t)l;
    const uint64_t all_full = (uint64_t)n * (uint64_t)l;
    uint64_t acc_lo = 0, acc_hi = 0;
    if constexpr (SEM == kSemBfs) acc_lo = M;  // every counted pair is at distance >= 1
    int level = 0, done_level = n + 1, last_change = 0, disconnected = 0;
    while (true) {
      ++level;
      // k[0]: reached pairs on base rows (matrix) / weighted counted pairs
      // reached (bfs); k[1]: centre rows (matrix) / all pairs reached (bfs);
      // k[2]: changed words.
      uint64_t k[3] = {0, 0, 0};
      for (int v = tid; v < n; v += kWideBlock) {
        const uint64_t old = Ra[v];
        const uint64_t val = old == full ? old : gather_row<CSR>(P, Ra, v, old, full);
        Rb[v] = val;
        const uint64_t pc = (uint64_t)__popcll(val);
        k[2] += val != old;
        if constexpr (SEM == kSemMatrix) {
          if (v < split) k[0] += pc;
          else k[1] += pc;
        } else {
          const int rel = v - first;
          if (rel > 0) {
            const uint64_t mask = rel >= l ? full : ((1ull << rel) - 1ull);
            const uint64_t w = centre ? 1ull : (v < split ? (uint64_t)(P.g - v / P.B) : (uint64_t)P.g);
            k[0] += w * (uint64_t)__popcll(val & mask);
          }
          k[1] += pc;
        }
      }
      block_sum<3>(k, red, par);  // also the barrier: Rb complete, Ra reads done
      par ^= 1;
      uint64_t *t = Ra;
      Ra = Rb;
      Rb = t;
      const bool chg = k[2] != 0;
      if (chg) last_change = level;
      if constexpr (SEM == kSemMatrix) {
        const bool done = centre ? (k[1] == hi_full) : (k[0] + k[1] == lo_full + hi_full);
        if (done) {
          done_level = level;
          break;
        }
        if (!chg || level >= n) {
          disconnected = 1;
          done_level = n + 1;
          break;
        }
        if (!centre) acc_lo += lo_full - k[0];
        acc_hi += hi_full - k[1];
      } else {
        if (chg) acc_lo += M - k[0];
        if (!chg || k[1] == all_full || level >= n) {
          disconnected = k[0] != M;
          done_level = last_change;
          break;
        }
      }
    }
    if (tid == 0) outs[u] = UnitOut{acc_lo, acc_hi, done_level, last_change, disconnected, 0};
  }
}

// Per-graph result from the tiles (one workgroup).
template <int SEM>
__global__ void __launch_bounds__(256)
gg_wide_combine_kernel(const UnitOut *__restrict__ outs, int units, int units_base, int n, int g,
                       WideOut *__restrict__ res) {
  __shared__ uint64_t s64[3][256];
  __shared__ int s32[4][256];
  const int tid = threadIdx.x;
  uint64_t lo = 0, hi = 0, chi = 0;
  int diam = SEM == kSemMatrix ? 1 : 0, ecc = 0, disc = 0, err = 0;
  for (int u = tid; u < units; u += 256) {
    const UnitOut o = outs[u];
    if (u >= units_base) chi += o.acc_hi + o.acc_lo;
    else {
      lo += o.acc_lo;
      hi += o.acc_hi;
    }
    diam = max(diam, o.done_level);
    ecc = max(ecc, o.last_change);
    disc |= o.disconnected;
    err |= o.error;
  }
  s64[0][tid] = lo;
  s64[1][tid] = hi;
  s64[2][tid] = chi;
  s32[0][tid] = diam;
  s32[1][tid] = ecc;
  s32[2][tid] = disc;
  s32[3][tid] = err;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) {
      s64[0][tid] += s64[0][tid + w];
      s64[1][tid] += s64[1][tid + w];
      s64[2][tid] += s64[2][tid + w];
      s32[0][tid] = max(s32[0][tid], s32[0][tid + w]);
      s32[1][tid] = max(s32[1][tid], s32[1][tid + w]);
      s32[2][tid] |= s32[2][tid + w];
      s32[3][tid] |= s32[3][tid + w];
    }
    __syncthreads();
  }
  if (tid == 0) {
    WideOut r;
    const uint64_t G = (uint64_t)g, N = (uint64_t)n;
    if constexpr (SEM == kSemMatrix) {
      r.sum = s64[0][0] * G + s64[1][0] * 2ull * G + s64[2][0] + N * (N - 1ull);
      r.diameter = s32[2][0] ? n + 1 : s32[0][0];
    } else {
      r.sum = s64[0][0] + s64[2][0];
      r.diameter = s32[0][0];
    }
    r.connected = s32[2][0] ? 0 : 1;
    r.max_ecc = s32[1][0];
    r.error = s32[3][0];
    *res = r;
  }
}

// adj int32 [n][d] (+ relabel) -> uint32 rows [n][d] in relabelled ids.
// Invalid entries (ids or relabel entries outside [0,n)) are counted in
// flags[0] and replaced by the row's own id; a row whose relabel entry is
// out of range is not stored.
__global__ void __launch_bounds__(256)
gg_wide_rows_kernel(const int32_t *__restrict__ adj, const int32_t *__restrict__ relabel, int n,
                    int d, uint32_t *__restrict__ rows, int32_t *__restrict__ flags) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)n * (size_t)d) return;
  const int i = (int)(e / (size_t)d), j = (int)(e - (size_t)i * d);
  const int ri = relabel ? relabel[i] : i;
  if ((unsigned)ri >= (unsigned)n) {
    atomicAdd(&flags[0], 1);
    return;
  }
  const int u = adj[e];
  int ru = (unsigned)u < (unsigned)n ? (relabel ? relabel[u] : u) : -1;
  if ((unsigned)ru >= (unsigned)n) {
    atomicAdd(&flags[0], 1);
    ru = ri;
  }
  rows[(size_t)ri * d + j] = (uint32_t)ru;
}

// Symmetry checks of adj int32 [n][d] for the bfs() semantics: flags[0]
// invalid entries, flags[1] entries (i -> u) whose multiplicity in row i
// differs from that of i in row u, flags[2] entries whose multiplicity
// differs from that of rot(u) in row rot(i), rot(x) = the same vertex in the
// next group (x < B*g), centres fixed.
__global__ void __launch_bounds__(256)
gg_wide_check_kernel(const int32_t *__restrict__ adj, int n, int d, int B, int g,
                     int32_t *__restrict__ flags) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)n * (size_t)d) return;
  const int i = (int)(e / (size_t)d);
  const int u = adj[e];
  if ((unsigned)u >= (unsigned)n) {
    atomicAdd(&flags[0], 1);
    return;
  }
  const int split = B * g;
  auto rot = [&](int x) { return x < split ? ((x / B + 1) % g) * B + x % B : x; };
  const int32_t *ri = adj + (size_t)i * d;
  const int32_t *ru = adj + (size_t)u * d;
  const int i2 = rot(i), u2 = rot(u);
  const int32_t *r2 = adj + (size_t)i2 * d;
  int c1 = 0, c2 = 0, c3 = 0;
  for (int k = 0; k < d; ++k) {
    c1 += ri[k] == u;
    c2 += ru[k] == i;
    c3 += r2[k] == u2;
  }
  if (c1 != c2) atomicAdd(&flags[1], 1);
  if (c1 != c3) atomicAdd(&flags[2], 1);
}

// Transposed table (in-neighbour lists) as CSR: counts, one-workgroup scan,
// scatter.  The order inside a list depends on atomics; the OR over a row
// does not depend on it.
__global__ void __launch_bounds__(256)
gg_wide_indeg_kernel(const int32_t *__restrict__ adj, int n, int d, uint32_t *__restrict__ cnt) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)n * (size_t)d) return;
  const int u = adj[e];
  if ((unsigned)u < (unsigned)n) atomicAdd(&cnt[u], 1u);
}

__global__ void __launch_bounds__(1024)
gg_wide_scan_kernel(const uint32_t *__restrict__ cnt, int n, uint32_t *__restrict__ off,
                    uint32_t *__restrict__ pos) {
  __shared__ uint32_t part[1024];
  const int tid = threadIdx.x;
  const int per = (n + 1023) / 1024;
  const int b = min(n, tid * per), e = min(n, b + per);
  uint32_t s = 0;
  for (int v = b; v < e; ++v) s += cnt[v];
  part[tid] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {  // inclusive scan of the parts
    const uint32_t x = tid >= o ? part[tid - o] : 0u;
    __syncthreads();
    part[tid] += x;
    __syncthreads();
  }
  uint32_t run = tid ? part[tid - 1] : 0u;
  for (int v = b; v < e; ++v) {
    off[v] = run;
    pos[v] = run;
    run += cnt[v];
  }
  if (tid == 1023) off[n] = part[1023];
}

__global__ void __launch_bounds__(256)
gg_wide_scatter_kernel(const int32_t *__restrict__ adj, int n, int d, uint32_t *__restrict__ pos,
                       uint32_t *__restrict__ col) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)n * (size_t)d) return;
  const int u = adj[e];
  if ((unsigned)u < (unsigned)n) col[atomicAdd(&pos[u], 1u)] = (uint32_t)(e / (size_t)d);
}

__global__ void gg_wide_clear_kernel(uint32_t *p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

struct Buf {
  void *p = nullptr;
  size_t cap = 0;
  int ensure(size_t bytes) {
    if (bytes <= cap) return GG_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    const size_t want = std::max(bytes, (size_t)256);
    const hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess)
      return gg::set_error(GG_ERR_HIP, "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    cap = want;
    return GG_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

struct WideSet {
  Buf adj, relabel, rows, off, cnt, pos, flags, scratch, outs, res;
};
WideSet W;

#define WIDE_HIP(call)                                                                      \
  do {                                                                                      \
    const hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                                   \
      return gg::set_error(GG_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                           __FILE__, __LINE__);                                             \
  } while (0)

unsigned blocks_for(size_t items, int per) { return (unsigned)((items + per - 1) / per); }

}  // namespace

namespace gg {

int wide_release() {
  for (Buf *b : {&W.adj, &W.relabel, &W.rows, &W.off, &W.cnt, &W.pos, &W.flags, &W.scratch, &W.outs,
                 &W.res})
    b->release();
  return GG_OK;
}

// One evaluation through the wide kernels.  `sem` 0: matrix_op (any n, with
// relabel); 1: bfs() semantics (no relabel), out->sum = the weighted sum over
// pairs s < i.  Adjacency (and relabel) are host pointers unless device_ptrs.
int wide_eval(const gg_graph &g, int sem, bool device_ptrs, gg_result *out) {
  int rc;
  if ((rc = init(-1))) return rc;
  const hipStream_t st = engine_stream();
  const int n = g.nodes, d = g.degree;
  const size_t entries = (size_t)n * (size_t)d;
  if (entries >= 0xffffffffull) return set_error(GG_ERR_TOO_LARGE, "n*d = %zu entries", entries);
  if (sem == kSemBfs && g.relabel) return set_error(GG_ERR_INVALID, "bfs() semantics take no relabel");
  const int32_t *adj_d = g.adj, *rel_d = g.relabel;
  if (!device_ptrs) {
    if ((rc = W.adj.ensure(entries * 4))) return rc;
    WIDE_HIP(hipMemcpyAsync(W.adj.p, g.adj, entries * 4, hipMemcpyHostToDevice, st));
    adj_d = (const int32_t *)W.adj.p;
    if (g.relabel) {
      if ((rc = W.relabel.ensure((size_t)n * 4))) return rc;
      WIDE_HIP(hipMemcpyAsync(W.relabel.p, g.relabel, (size_t)n * 4, hipMemcpyHostToDevice, st));
      rel_d = (const int32_t *)W.relabel.p;
    }
  }
  if ((rc = W.flags.ensure(16))) return rc;
  int32_t *flags = (int32_t *)W.flags.p;
  hipLaunchKernelGGL(gg_wide_clear_kernel, dim3(1), dim3(64), 0, st, (uint32_t *)flags, (size_t)4);
  const unsigned eb = blocks_for(entries, 256);
  bool csr = false;
  if (sem == kSemBfs) {
    int32_t hf[4] = {0, 0, 0, 0};
    hipLaunchKernelGGL(gg_wide_check_kernel, dim3(eb), dim3(256), 0, st, adj_d, n, d, g.based_nodes,
                       g.groups, flags);
    WIDE_HIP(hipGetLastError());
    WIDE_HIP(hipMemcpyAsync(hf, flags, 16, hipMemcpyDeviceToHost, st));
    WIDE_HIP(hipStreamSynchronize(st));
    if (hf[0]) return set_error(GG_ERR_INVALID, "adjacency has %d entries outside [0,%d)", hf[0], n);
    if (hf[1] == 0 && hf[2] == 0 && n <= 65536) {
      // Symmetric and invariant under the rotation: sum = S/2 exactly.
      gg_graph g2 = g;
      g2.adj = adj_d;
      g2.flags = (g.flags | GG_GRAPH_SYMMETRIC) & ~GG_GRAPH_BFS;
      if ((rc = eval_host(1, &g2, out, true, 0, nullptr))) return rc;
      if (out->connected) {
        if (out->sum & 1ull) return set_error(GG_ERR_STATE, "odd S on a symmetric graph");
        out->sum >>= 1;
      }
      out->diameter = out->connected ? out->diameter : n + 1;
      return GG_OK;
    }
    csr = hf[1] != 0;  // directed: pull along in-neighbours
  }
  WideDesc P = {};
  P.n = n;
  P.d = d;
  P.B = g.based_nodes;
  P.g = g.groups;
  P.c = g.added_centers;
  P.split = g.based_nodes * g.groups;
  P.flags = flags;
  if (csr) {
    if ((rc = W.off.ensure(((size_t)n + 1) * 4)) || (rc = W.cnt.ensure((size_t)n * 4)) ||
        (rc = W.pos.ensure((size_t)n * 4)) || (rc = W.rows.ensure(entries * 4)))
      return rc;
    hipLaunchKernelGGL(gg_wide_clear_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, st,
                       (uint32_t *)W.cnt.p, (size_t)n);
    hipLaunchKernelGGL(gg_wide_indeg_kernel, dim3(eb), dim3(256), 0, st, adj_d, n, d,
                       (uint32_t *)W.cnt.p);
    hipLaunchKernelGGL(gg_wide_scan_kernel, dim3(1), dim3(1024), 0, st, (const uint32_t *)W.cnt.p, n,
                       (uint32_t *)W.off.p, (uint32_t *)W.pos.p);
    hipLaunchKernelGGL(gg_wide_scatter_kernel, dim3(eb), dim3(256), 0, st, adj_d, n, d,
                       (uint32_t *)W.pos.p, (uint32_t *)W.rows.p);
    P.off = (const uint32_t *)W.off.p;
  } else {
    if ((rc = W.rows.ensure(entries * 4))) return rc;
    hipLaunchKernelGGL(gg_wide_rows_kernel, dim3(eb), dim3(256), 0, st, adj_d, rel_d, n, d,
                       (uint32_t *)W.rows.p, flags);
  }
  P.col = (const uint32_t *)W.rows.p;
  P.units_base = (g.based_nodes + 63) / 64;
  P.units = P.units_base + (g.added_centers + 63) / 64;
  int cus = 256;
  {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
      cus = prop.multiProcessorCount;
  }
  // Two workgroups per CU, and at most 8 GiB of reach words.
  const size_t per_block = 2 * (size_t)n * 8;
  const size_t cap_blocks = std::max<size_t>(1, ((size_t)8 << 30) / per_block);
  const int blocks = (int)std::min<size_t>({(size_t)P.units, (size_t)2 * cus, cap_blocks});
  if ((rc = W.scratch.ensure(per_block * blocks)) || (rc = W.outs.ensure((size_t)P.units * sizeof(UnitOut))) ||
      (rc = W.res.ensure(sizeof(WideOut))))
    return rc;
  P.scratch = (uint64_t *)W.scratch.p;
  UnitOut *outs = (UnitOut *)W.outs.p;
  if (sem == kSemBfs) {
    if (csr) hipLaunchKernelGGL((gg_wide_bfs_kernel<true, kSemBfs>), dim3(blocks), dim3(kWideBlock), 0, st, P, outs);
    else hipLaunchKernelGGL((gg_wide_bfs_kernel<false, kSemBfs>), dim3(blocks), dim3(kWideBlock), 0, st, P, outs);
    hipLaunchKernelGGL(gg_wide_combine_kernel<kSemBfs>, dim3(1), dim3(256), 0, st, outs, P.units,
                       P.units_base, n, g.groups, (WideOut *)W.res.p);
  } else {
    hipLaunchKernelGGL((gg_wide_bfs_kernel<false, kSemMatrix>), dim3(blocks), dim3(kWideBlock), 0, st, P, outs);
    hipLaunchKernelGGL(gg_wide_combine_kernel<kSemMatrix>, dim3(1), dim3(256), 0, st, outs, P.units,
                       P.units_base, n, g.groups, (WideOut *)W.res.p);
  }
  WIDE_HIP(hipGetLastError());
  WideOut r;
  WIDE_HIP(hipMemcpyAsync(&r, W.res.p, sizeof r, hipMemcpyDeviceToHost, st));
  WIDE_HIP(hipStreamSynchronize(st));
  if (r.error) return set_error(GG_ERR_INVALID, "adjacency entry outside [0,%d)", n);
  out->connected = r.connected;
  out->diameter = r.connected ? r.diameter : n + 1;
  out->max_ecc = r.max_ecc;
  out->word_bits = 64;
  out->sum = r.sum;
  return GG_OK;
}

}  // namespace gg
