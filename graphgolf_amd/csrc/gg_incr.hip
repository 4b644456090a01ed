// gg_incr.hip -- incremental re-scoring of 2-opt moves (C ABI in
// include/gg_incr.h).
//
// Per move, on the evaluator's stream, with no host round trip before the
// result copy:
//   1. probe units (gg_tile_bfs_kernel, kTileProbe) on the current graph,
//      per endpoint p ceil((d-1)/15) units, each with sources p and up to 15
//      of p's neighbours other than its removed partner: per vertex s, the
//      level L0 at which p reaches s and, per unit, the first level L1 at
//      which one of its other sources does.  Since |d(s,w) - d(s,p)| <= 1 for
//      a neighbour w, "p has a parent other than its partner at the level
//      before it, seen from s" is (the least L1 of p's units) < L0.
//   2. the adjacency patch (four entries) that makes the moved graph;
//   3. gg_incr_flag_kernel: the sources whose distances can change (the
//      test in include/gg_incr.h), appended to a list, with the sum of their
//      current distance sums and a histogram of their current eccentricities;
//   4. the list units (gg_tile_list_kernel, kTileList, persistent) on the
//      moved graph: exact per-unit sums and completion levels;
//   5. gg_incr_finalize_kernel: S' = S - (old sums of the list) + (new sums),
//      diameter' = max(list completion levels, largest eccentricity still
//      held by a source outside the list).
// Accept re-runs the list with per-source outputs (kTileSrc) and commits
// them; reject restores the four adjacency entries.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

#include "gg_device.h"
#include "gg_eval.h"
#include "gg_incr.h"
#include "gg_internal.h"
#include "gg_tile.h"

using gg::set_error;

#define GG_HIP(call)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return set_error(GG_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),     \
                       __FILE__, __LINE__);                                                   \
  } while (0)

namespace {

constexpr int kProbes = 4;        // endpoints of a move
constexpr int kWord = 16;         // sources per unit (uint16 reach words)
constexpr int kMaxDegree = 32;    // rows of at most four 16-byte chunks
// Probe units per endpoint: the endpoint plus 15 of its other neighbours each.
constexpr int probe_units(int d) { return (d - 1 + kWord - 2) / (kWord - 1); }
constexpr int kLocalHist = 64;    // eccentricities histogrammed in LDS by the flag kernel
constexpr double kInlineSrc = 0.3;

// The moved graph's result, written by the finalize kernel.
struct IncrResult {
  uint64_t sum;
  int32_t connected, diameter, affected, pad;
};

// What the flag kernel needs of a move: the probe index of each endpoint of
// the two added edges (probes 0..3 are u1, v1, u2, v2), and the probe units
// per endpoint (endpoint p's units are p*pu .. p*pu + pu - 1).
struct MoveProbes {
  int32_t add[4];
  int32_t pu;
};

struct Patch {
  int32_t at[4];   // adjR element index (row * dp + column)
  int32_t val[4];  // id to store
};

__global__ void __launch_bounds__(64) gg_incr_patch_kernel(uint16_t *adjR, Patch p) {
  const int t = threadIdx.x;
  if (t < 4) adjR[p.at[t]] = (uint16_t)p.val[t];
}

__global__ void __launch_bounds__(256)
gg_incr_clear_kernel(int32_t *aff_count, unsigned long long *aff_sum_old, int32_t *removed_hist,
                     int bins) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) {
    *aff_count = 0;
    *aff_sum_old = 0ull;
  }
  if (t < bins) removed_hist[t] = 0;
}

// One thread per source s.
__global__ void __launch_bounds__(256)
gg_incr_flag_kernel(int n, const uint16_t *__restrict__ l0, const uint16_t *__restrict__ l1,
                    MoveProbes mv, const uint32_t *__restrict__ src_sum,
                    const uint16_t *__restrict__ src_ecc, uint16_t *__restrict__ aff,
                    int32_t *__restrict__ aff_count, unsigned long long *__restrict__ aff_sum_old,
                    int32_t *__restrict__ removed_hist) {
  __shared__ int32_t lh[kLocalHist];
  const int tid = threadIdx.x;
  if (tid < kLocalHist) lh[tid] = 0;
  __syncthreads();
  const int s = blockIdx.x * blockDim.x + tid;
  bool hit = false;
  if (s < n) {
    int L0[kProbes];
    bool par[kProbes];
#pragma unroll
    for (int p = 0; p < kProbes; ++p) {
      const size_t u0 = (size_t)(p * mv.pu);
      L0[p] = l0[u0 * n + s];
      int L1 = l1[u0 * n + s];
      for (int k = 1; k < mv.pu; ++k) L1 = min(L1, (int)l1[(u0 + k) * n + s]);
      par[p] = L1 < L0[p];
    }
    // Removed edges (u1,v1) = probes (0,1) and (u2,v2) = probes (2,3).
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int a = 2 * e, b = 2 * e + 1;
      hit |= (L0[a] < L0[b] && !par[b]) || (L0[b] < L0[a] && !par[a]);
    }
    // Added edges.
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int dx = L0[mv.add[2 * e]] - L0[mv.add[2 * e + 1]];
      hit |= dx >= 2 || dx <= -2;
    }
  }
  // Append the wave's hits with one reservation.
  const uint64_t m = __builtin_amdgcn_ballot_w64(hit);
  const int lane = tid & 63;
  if (m) {
    const int cnt = __popcll(m);
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(aff_count, cnt);
    base = __shfl(base, leader);
    const uint64_t below = lane ? (m & ((1ull << lane) - 1ull)) : 0ull;
    uint32_t old = 0;
    if (hit) {
      aff[base + __popcll(below)] = (uint16_t)s;
      old = src_sum[s];
      const int e = src_ecc[s];
      if (e < kLocalHist) atomicAdd(&lh[e], 1);
      else atomicAdd(&removed_hist[e], 1);
    }
    // Wave sum of the old distance sums (at most 64 * n * levels: fits 64 bits).
    unsigned long long v = old;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (lane == leader) atomicAdd(aff_sum_old, v);
  }
  __syncthreads();
  if (tid < kLocalHist && lh[tid]) atomicAdd(&removed_hist[tid], lh[tid]);
}

__global__ void __launch_bounds__(256)
gg_incr_finalize_kernel(int n, const int32_t *__restrict__ aff_count,
                        const UnitOut *__restrict__ outs,
                        const unsigned long long *__restrict__ aff_sum_old,
                        const int32_t *__restrict__ hist, const int32_t *__restrict__ removed_hist,
                        unsigned long long sum_cur, int diam_cur, IncrResult *__restrict__ res) {
  __shared__ unsigned long long ssum[256];
  __shared__ int smax[256], sdisc[256];
  const int tid = threadIdx.x;
  const int k = *aff_count;
  const int nu = (k + kWord - 1) / kWord;
  unsigned long long acc = 0;
  int mx = 0, disc = 0;
  for (int u = tid; u < nu; u += blockDim.x) {
    const UnitOut o = outs[u];
    const int l = min(kWord, k - u * kWord);
    // A unit's distance sum: its per-level unreached counts plus the level-0
    // term l*(n-1) the unit loop starts from.
    acc += o.acc_lo + (unsigned long long)l * (unsigned long long)(n - 1);
    mx = max(mx, o.done_level);
    disc |= o.disconnected | (o.error != 0);
  }
  ssum[tid] = acc;
  smax[tid] = mx;
  sdisc[tid] = disc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) {
      ssum[tid] += ssum[tid + w];
      smax[tid] = max(smax[tid], smax[tid + w]);
      sdisc[tid] |= sdisc[tid + w];
    }
    __syncthreads();
  }
  if (tid == 0) {
    IncrResult r;
    r.affected = k;
    r.pad = 0;
    if (sdisc[0]) {
      r.connected = 0;
      r.diameter = n + 1;
      r.sum = 0;
    } else {
      int keep = 0;  // largest eccentricity held outside the list
      for (int e = diam_cur; e >= 0; --e)
        if (hist[e] - removed_hist[e] > 0) {
          keep = e;
          break;
        }
      r.connected = 1;
      r.diameter = max(keep, smax[0]);
      r.sum = sum_cur - *aff_sum_old + ssum[0];
    }
    *res = r;
  }
}

// Writes the list's new per-source values into the per-vertex arrays.
__global__ void __launch_bounds__(256)
gg_incr_commit_kernel(const int32_t *__restrict__ aff_count, const uint16_t *__restrict__ aff,
                      const uint32_t *__restrict__ new_sum, const uint16_t *__restrict__ new_ecc,
                      uint32_t *__restrict__ src_sum, uint16_t *__restrict__ src_ecc,
                      int32_t *__restrict__ hist) {
  // Histogram changes are gathered per block in LDS (most eccentricities
  // share a few values), then added with one atomic per touched bin.
  __shared__ int32_t dh[kLocalHist];
  const int tid = threadIdx.x;
  if (tid < kLocalHist) dh[tid] = 0;
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + tid;
  if (i < *aff_count) {
    const int s = aff[i];
    const int eo = src_ecc[s], en = new_ecc[i];
    if (eo != en) {
      if (eo < kLocalHist) atomicSub(&dh[eo], 1);
      else atomicSub(&hist[eo], 1);
      if (en < kLocalHist) atomicAdd(&dh[en], 1);
      else atomicAdd(&hist[en], 1);
    }
    src_ecc[s] = (uint16_t)en;
    src_sum[s] = new_sum[i];
  }
  __syncthreads();
  if (tid < kLocalHist && dh[tid]) atomicAdd(&hist[tid], dh[tid]);
}

typedef void (*UnitKernel)(const GraphDev *, const UnitDesc *, UnitOut *, int, int, int);
typedef void (*ListKernel)(const GraphDev *, UnitOut *, int, int);

template <int B, int V> struct Kernels {
  static UnitKernel probe() { return (UnitKernel)&gg_tile_bfs_kernel<uint16_t, B, V, kTileProbe>; }
  static ListKernel list() { return (ListKernel)&gg_tile_list_kernel<uint16_t, B, V, kTileList>; }
  static ListKernel src() { return (ListKernel)&gg_tile_list_kernel<uint16_t, B, V, kTileSrc>; }
};

struct KernelSet {
  UnitKernel probe = nullptr;
  ListKernel list = nullptr, src = nullptr;
};

template <int B, int V> KernelSet kernel_set() {
  return {Kernels<B, V>::probe(), Kernels<B, V>::list(), Kernels<B, V>::src()};
}

KernelSet kernels_for(int block, int vpt) {
  if (block == 256) {
    switch (vpt) {
      case 1: return kernel_set<256, 1>();
      case 2: return kernel_set<256, 2>();
      case 4: return kernel_set<256, 4>();
      case 8: return kernel_set<256, 8>();
    }
  } else if (block == 1024) {
    switch (vpt) {
      case 4: return kernel_set<1024, 4>();
      case 8: return kernel_set<1024, 8>();
      case 16: return kernel_set<1024, 16>();
      case 32: return kernel_set<1024, 32>();
      case 64: return kernel_set<1024, 64>();
    }
  }
  return {};
}

template <typename T> int dev_alloc(T **p, size_t count) {
  *p = nullptr;
  GG_HIP(hipMalloc((void **)p, std::max<size_t>(count * sizeof(T), 256)));
  return GG_OK;
}

}  // namespace

struct gg_incr {
  int n = 0, d = 0, dp = 0;
  int block = 0, vpt = 0, n_cap = 0, qcap = 0, grid = 0;
  size_t lds = 0;
  KernelSet K;
  hipStream_t stream = nullptr;
  std::vector<int32_t> adj;  // current graph, host copy [n][d]
  // Device id of vertex v: graphs the row-order mode orders are numbered by
  // the bank-balanced layout (gg_bank_layout) once; the device state (rows,
  // per-source values, lists) uses these ids throughout, the caller's ids
  // are mapped at every move.
  std::vector<int32_t> sig;  // [n]
  // Device-row column of adjacency entry (v, j): the rows are put in a
  // bank-aware order once, and patches keep every position.
  std::vector<uint16_t> slot;  // [n][d]
  // Device state.
  uint16_t *adjR = nullptr;          // [n][dp] + 64 bytes (current graph; patched while pending)
  int32_t *flags = nullptr;          // [2]
  GraphDev *gdesc = nullptr;         // [0] probe, [1] list, [2] src (list), [3] src (all vertices)
  int pu = 1;                        // probe units per endpoint
  UnitDesc *pdesc = nullptr;         // [4 * pu] probe units
  uint16_t *probe_src = nullptr;     // [4 * pu][16]
  uint16_t *l0 = nullptr, *l1 = nullptr;  // [4 * pu][n]
  uint16_t *aff = nullptr;           // [n]
  int32_t *aff_count = nullptr;      // [1] (+ [1] = n, the all-vertices count)
  unsigned long long *aff_sum_old = nullptr;
  int32_t *hist = nullptr, *removed_hist = nullptr;  // [n + 2]
  uint32_t *src_sum = nullptr, *new_sum = nullptr;    // [n]
  uint16_t *src_ecc = nullptr, *new_ecc = nullptr;    // [n]
  UnitOut *outs = nullptr;           // [ceil(n/16)]
  IncrResult *res = nullptr;
  // Pinned host staging.
  IncrResult *res_h = nullptr;
  uint16_t *probe_src_h = nullptr;
  UnitDesc *pdesc_h = nullptr;
  // Current graph.
  uint64_t sum = 0;
  int diameter = 0;
  // Accepted fraction of scored moves (moving average).  Above kInlineSrc
  // the per-source values are computed while scoring (a slower list run)
  // instead of by a second run on accept.
  double accept_rate = 0.5;
  // Pending move.
  bool pending = false;
  bool pend_has_src = false;
  Patch undo{}, redo{};
  int32_t pend_rows[4], pend_cols[4], pend_vals[4], pend_old[4];
  IncrResult pend{};
};

namespace {

int free_all(gg_incr *s) {
  void *ptrs[] = {s->adjR, s->flags, s->gdesc, s->pdesc, s->probe_src, s->l0, s->l1, s->aff,
                  s->aff_count, s->aff_sum_old, s->hist, s->removed_hist, s->src_sum,
                  s->new_sum, s->src_ecc, s->new_ecc, s->outs, s->res};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (s->res_h) (void)hipHostFree(s->res_h);
  if (s->probe_src_h) (void)hipHostFree(s->probe_src_h);
  if (s->pdesc_h) (void)hipHostFree(s->pdesc_h);
  return GG_OK;
}

int col_of(const gg_incr *s, int row, int id) {
  const int32_t *r = s->adj.data() + (size_t)row * s->d;
  for (int j = 0; j < s->d; ++j)
    if (r[j] == id) return j;
  return -1;
}

// Simple, regular, symmetric adjacency with ids in range.
int check_graph(int n, int d, const int32_t *adj) {
  std::vector<int32_t> row(d);
  for (int v = 0; v < n; ++v) {
    for (int j = 0; j < d; ++j) {
      const int u = adj[(size_t)v * d + j];
      if (u < 0 || u >= n) return set_error(GG_ERR_INVALID, "adjacency entry %d of row %d out of range", u, v);
      if (u == v) return set_error(GG_ERR_INVALID, "vertex %d has a self-loop", v);
      row[j] = u;
    }
    std::sort(row.begin(), row.end());
    for (int j = 1; j < d; ++j)
      if (row[j] == row[j - 1]) return set_error(GG_ERR_INVALID, "vertex %d repeats neighbour %d", v, row[j]);
    for (int j = 0; j < d; ++j) {
      const int u = row[j];
      bool back = false;
      for (int k = 0; k < d; ++k) back |= adj[(size_t)u * d + k] == v;
      if (!back) return set_error(GG_ERR_INVALID, "edge %d-%d is not symmetric", v, u);
    }
  }
  return GG_OK;
}

// Runs the per-source kernel over the list described by gdesc[idx].
int run_src(gg_incr *s, int idx) {
  hipLaunchKernelGGL(s->K.src, dim3(s->grid), dim3(s->block), s->lds, s->stream,
                     (const GraphDev *)(s->gdesc + idx), s->outs, s->n_cap, s->qcap);
  GG_HIP(hipGetLastError());
  return GG_OK;
}

}  // namespace

extern "C" {

gg_incr *gg_incr_create(int nodes, int degree, const int32_t *adj) {
  if (!adj || nodes < 3 || nodes > 65536 || degree < 2 || degree > kMaxDegree) {
    set_error(GG_ERR_INVALID, "incremental scoring needs 3 <= nodes <= 65536, 2 <= degree <= 32");
    return nullptr;
  }
  if (check_graph(nodes, degree, adj)) return nullptr;
  return gg::incr_create_unchecked(nodes, degree, adj);
}

}  // extern "C"

gg_incr *gg::incr_create_unchecked(int nodes, int degree, const int32_t *adj) {
  if (!adj || nodes < 3 || nodes > 65536 || degree < 2 || degree > kMaxDegree) {
    set_error(GG_ERR_INVALID, "incremental scoring needs 3 <= nodes <= 65536, 2 <= degree <= 32");
    return nullptr;
  }
  if (gg_init(-1)) return nullptr;
  gg_incr *s = new gg_incr();
  auto fail = [&]() -> gg_incr * {
    free_all(s);
    delete s;
    return nullptr;
  };
  const int n = nodes, d = degree, dp = (d + 7) & ~7;
  s->n = n;
  s->d = d;
  s->dp = dp;
  s->pu = probe_units(d);
  const int nprobe = kProbes * s->pu;
  s->stream = gg::engine_stream();
  s->adj.assign(adj, adj + (size_t)n * d);
  if (gg::tile_shape(kWord, n, &s->block, &s->vpt, &s->n_cap, &s->qcap, &s->lds)) return fail();
  s->K = kernels_for(s->block, s->vpt);
  if (!s->K.probe) {
    set_error(GG_ERR_STATE, "no kernel for block %d vpt %d", s->block, s->vpt);
    return fail();
  }
  const void *kf[3] = {(const void *)s->K.probe, (const void *)s->K.list, (const void *)s->K.src};
  for (const void *f : kf)
    if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)s->lds) != hipSuccess) {
      set_error(GG_ERR_HIP, "hipFuncSetAttribute failed");
      return fail();
    }
  int per_cu = 0, dev = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)s->K.list, s->block,
                                                   s->lds) != hipSuccess ||
      hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
    set_error(GG_ERR_HIP, "occupancy query failed");
    return fail();
  }
  s->grid = std::max(1, per_cu) * std::max(1, cus);
  const int units = (n + kWord - 1) / kWord;
  s->grid = std::min(s->grid, units);
  if (dev_alloc(&s->adjR, (size_t)n * dp + 32) || dev_alloc(&s->flags, 2) ||
      dev_alloc(&s->gdesc, 4) || dev_alloc(&s->pdesc, nprobe) ||
      dev_alloc(&s->probe_src, nprobe * kWord) || dev_alloc(&s->l0, (size_t)nprobe * n) ||
      dev_alloc(&s->l1, (size_t)nprobe * n) || dev_alloc(&s->aff, (size_t)n) ||
      dev_alloc(&s->aff_count, 2) || dev_alloc(&s->aff_sum_old, 1) ||
      dev_alloc(&s->hist, (size_t)n + 2) || dev_alloc(&s->removed_hist, (size_t)n + 2) ||
      dev_alloc(&s->src_sum, (size_t)n) || dev_alloc(&s->new_sum, (size_t)n) ||
      dev_alloc(&s->src_ecc, (size_t)n) || dev_alloc(&s->new_ecc, (size_t)n) ||
      dev_alloc(&s->outs, (size_t)std::max(units, nprobe)) || dev_alloc(&s->res, 1))
    return fail();
  if (hipHostMalloc((void **)&s->res_h, sizeof(IncrResult)) != hipSuccess ||
      hipHostMalloc((void **)&s->probe_src_h, nprobe * kWord * sizeof(uint16_t)) != hipSuccess ||
      hipHostMalloc((void **)&s->pdesc_h, nprobe * sizeof(UnitDesc)) != hipSuccess) {
    set_error(GG_ERR_HIP, "hipHostMalloc failed");
    return fail();
  }
  // Row-major uint16 rows padded with the row's own id (a self-loop adds
  // nothing to a reach set), as the evaluator's prep kernel writes them:
  // row sig[v] holds sig of v's neighbours, in the layout's slot order.
  const bool layout = gg::order_rows_applies(n, dp);
  std::vector<int32_t> ordered(adj, adj + (size_t)n * d);
  s->sig.resize(n);
  if (layout) {
    uint64_t cost[2];
    if (gg_bank_layout(n, d, adj, kWord, 0, 1, 0, s->sig.data(), ordered.data(), cost)) return fail();
  } else {
    std::iota(s->sig.begin(), s->sig.end(), 0);
  }
  const int32_t *sig = s->sig.data();
  std::vector<uint16_t> rows((size_t)n * dp + 32, 0);
  for (int v = 0; v < n; ++v) {
    const int r = sig[v];
    for (int j = 0; j < dp; ++j)
      rows[(size_t)r * dp + j] = (uint16_t)(j < d ? sig[ordered[(size_t)v * d + j]] : r);
  }
  std::vector<uint16_t> all(n);
  std::iota(all.begin(), all.end(), 0);
  GraphDev g[4];
  for (int i = 0; i < 4; ++i) {
    GraphDev &x = g[i];
    x.adjR = s->adjR;
    x.flags = s->flags;
    x.n = n;
    x.dp = dp;
    x.split = n;
    x.c = 0;
    x.push_ok = 1;
    x.groups = 1;
    x.hist_maxd = 0;
    x.trace_units = 0;
    x.hist = nullptr;
    x.trace = nullptr;
  }
  g[0].srcs = s->probe_src;
  g[0].probe_l0 = s->l0;
  g[0].probe_l1 = s->l1;
  g[1].srcs = s->aff;
  g[1].src_count = s->aff_count;
  g[2].srcs = s->aff;
  g[2].src_count = s->aff_count;
  g[2].src_sum = s->new_sum;
  g[2].src_ecc = s->new_ecc;
  g[3].srcs = s->aff;  // holds 0..n-1 while the start graph is measured
  g[3].src_count = s->aff_count + 1;
  g[3].src_sum = s->src_sum;
  g[3].src_ecc = s->src_ecc;
  const int32_t counts[2] = {0, n};
  const int32_t zero_flags[2] = {0, 0};
  std::vector<uint16_t> ecc(n);
  std::vector<uint32_t> sums(n);
  auto upload = [&]() -> int {
    GG_HIP(hipMemcpy(s->adjR, rows.data(), rows.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    if (!layout) {
      int orc = gg::order_rows_sync(s->adjR, n, dp, kWord, s->stream);
      if (orc) return orc;
      GG_HIP(hipMemcpy(rows.data(), s->adjR, (size_t)n * dp * sizeof(uint16_t), hipMemcpyDeviceToHost));
    }
    s->slot.assign((size_t)n * d, 0);
    for (int v = 0; v < n; ++v) {
      uint32_t taken = 0;  // dp <= 32 whenever the rows were reordered
      const size_t r = (size_t)sig[v];
      for (int j = 0; j < d; ++j) {
        const uint16_t want = (uint16_t)sig[adj[(size_t)v * d + j]];
        int k = 0;
        while (k < dp && ((k < 32 && ((taken >> k) & 1u)) || rows[r * dp + k] != want)) ++k;
        if (k == dp) return set_error(GG_ERR_STATE, "row %d lost neighbour %d", v, (int)want);
        if (k < 32) taken |= 1u << k;
        s->slot[(size_t)v * d + j] = (uint16_t)k;
      }
    }
    GG_HIP(hipMemcpy(s->flags, zero_flags, sizeof(zero_flags), hipMemcpyHostToDevice));
    GG_HIP(hipMemcpy(s->gdesc, g, sizeof(g), hipMemcpyHostToDevice));
    GG_HIP(hipMemcpy(s->aff, all.data(), (size_t)n * sizeof(uint16_t), hipMemcpyHostToDevice));
    GG_HIP(hipMemcpy(s->aff_count, counts, sizeof(counts), hipMemcpyHostToDevice));
    int rc = run_src(s, 3);
    if (rc) return rc;
    GG_HIP(hipStreamSynchronize(s->stream));
    GG_HIP(hipMemcpy(ecc.data(), s->src_ecc, (size_t)n * sizeof(uint16_t), hipMemcpyDeviceToHost));
    GG_HIP(hipMemcpy(sums.data(), s->src_sum, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return GG_OK;
  };
  if (upload()) return fail();
  std::vector<int32_t> hist((size_t)n + 2, 0);
  uint64_t total = 0;
  int diam = 0;
  for (int v = 0; v < n; ++v) {
    if (ecc[v] == 0xffff || ecc[v] > n) {
      set_error(GG_ERR_INVALID, "the start graph is not connected");
      return fail();
    }
    hist[ecc[v]]++;
    total += sums[v];
    diam = std::max(diam, (int)ecc[v]);
  }
  if (hipMemcpy(s->hist, hist.data(), hist.size() * sizeof(int32_t), hipMemcpyHostToDevice) !=
      hipSuccess) {
    set_error(GG_ERR_HIP, "hist upload failed");
    return fail();
  }
  s->sum = total;
  s->diameter = diam;
  return s;
}

extern "C" {

int gg_incr_state(const gg_incr *s, gg_result *out) {
  if (!s || !out) return set_error(GG_ERR_INVALID, "null argument");
  out->connected = 1;
  out->diameter = s->diameter;
  out->max_ecc = s->diameter;
  out->word_bits = kWord;
  out->sum = s->sum;
  return GG_OK;
}

int gg_incr_eval(gg_incr *s, const int32_t rem[4], const int32_t add[4], gg_result *out,
                 int32_t *affected) {
  if (!s || !rem || !add || !out) return set_error(GG_ERR_INVALID, "null argument");
  if (s->pending) return set_error(GG_ERR_STATE, "a move is pending (accept or reject it first)");
  const int n = s->n, d = s->d, dp = s->dp;
  // Endpoints: probes 0..3 = u1, v1, u2, v2.
  const int P[4] = {rem[0], rem[1], rem[2], rem[3]};
  for (int i = 0; i < 4; ++i) {
    if (P[i] < 0 || P[i] >= n) return set_error(GG_ERR_INVALID, "move endpoint %d out of range", P[i]);
    for (int j = 0; j < i; ++j)
      if (P[i] == P[j]) return set_error(GG_ERR_INVALID, "move endpoints must be distinct");
  }
  auto probe_of = [&](int v) {
    for (int i = 0; i < 4; ++i)
      if (P[i] == v) return i;
    return -1;
  };
  MoveProbes mv;
  mv.pu = s->pu;
  const int nprobe = kProbes * s->pu;
  int seen[4] = {0, 0, 0, 0};
  for (int i = 0; i < 4; ++i) {
    const int p = probe_of(add[i]);
    if (p < 0) return set_error(GG_ERR_INVALID, "added edges must join the removed edges' endpoints");
    mv.add[i] = p;
    seen[p]++;
  }
  for (int i = 0; i < 4; ++i)
    if (seen[i] != 1) return set_error(GG_ERR_INVALID, "each endpoint must appear once in the added edges");
  for (int e = 0; e < 2; ++e) {
    const int x = add[2 * e], y = add[2 * e + 1];
    if (mv.add[2 * e] / 2 == mv.add[2 * e + 1] / 2)
      return set_error(GG_ERR_INVALID, "an added edge re-joins a removed edge's endpoints");
    if (col_of(s, x, y) >= 0) return set_error(GG_ERR_INVALID, "added edge %d-%d already exists", x, y);
  }
  // Adjacency positions: endpoint P[i] loses its partner and gains its new neighbour.
  const int partner[4] = {P[1], P[0], P[3], P[2]};
  Patch redo, undo;
  for (int i = 0; i < 4; ++i) {
    const int col = col_of(s, P[i], partner[i]);
    if (col < 0) return set_error(GG_ERR_INVALID, "edge %d-%d does not exist", P[i], partner[i]);
    int nb = -1;
    for (int e = 0; e < 2; ++e) {
      if (add[2 * e] == P[i]) nb = add[2 * e + 1];
      if (add[2 * e + 1] == P[i]) nb = add[2 * e];
    }
    s->pend_rows[i] = P[i];
    s->pend_cols[i] = col;
    s->pend_vals[i] = nb;
    s->pend_old[i] = partner[i];
    redo.at[i] = undo.at[i] = s->sig[P[i]] * dp + s->slot[(size_t)P[i] * d + col];
    redo.val[i] = s->sig[nb];
    undo.val[i] = s->sig[partner[i]];
  }
  // Probe sources: per unit the endpoint, then up to 15 of its neighbours
  // other than its partner.
  for (int i = 0; i < 4; ++i) {
    const int32_t *row = s->adj.data() + (size_t)P[i] * d;
    int j = 0;
    for (int u = 0; u < s->pu; ++u) {
      const int unit = i * s->pu + u;
      uint16_t *ps = s->probe_src_h + unit * kWord;
      int k = 0;
      ps[k++] = (uint16_t)s->sig[P[i]];
      for (; j < d && k < kWord; ++j)
        if (row[j] != partner[i]) ps[k++] = (uint16_t)s->sig[row[j]];
      s->pdesc_h[unit] = UnitDesc{0, unit * kWord, k, 0};
    }
  }
  hipStream_t st = s->stream;
  GG_HIP(hipMemcpyAsync(s->probe_src, s->probe_src_h, nprobe * kWord * sizeof(uint16_t),
                        hipMemcpyHostToDevice, st));
  GG_HIP(hipMemcpyAsync(s->pdesc, s->pdesc_h, nprobe * sizeof(UnitDesc), hipMemcpyHostToDevice, st));
  const int bins = s->diameter + 2;
  hipLaunchKernelGGL(gg_incr_clear_kernel, dim3((bins + 255) / 256), dim3(256), 0, st,
                     s->aff_count, s->aff_sum_old, s->removed_hist, bins);
  hipLaunchKernelGGL(s->K.probe, dim3(nprobe), dim3(s->block), s->lds, st,
                     (const GraphDev *)s->gdesc, (const UnitDesc *)s->pdesc, s->outs, 0, s->n_cap,
                     s->qcap);
  hipLaunchKernelGGL(gg_incr_patch_kernel, dim3(1), dim3(64), 0, st, s->adjR, redo);
  hipLaunchKernelGGL(gg_incr_flag_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n,
                     (const uint16_t *)s->l0, (const uint16_t *)s->l1, mv,
                     (const uint32_t *)s->src_sum, (const uint16_t *)s->src_ecc, s->aff,
                     s->aff_count, s->aff_sum_old, s->removed_hist);
  const bool inline_src = s->accept_rate > kInlineSrc;
  hipLaunchKernelGGL(inline_src ? s->K.src : s->K.list, dim3(s->grid), dim3(s->block), s->lds, st,
                     (const GraphDev *)(s->gdesc + (inline_src ? 2 : 1)), s->outs, s->n_cap,
                     s->qcap);
  hipLaunchKernelGGL(gg_incr_finalize_kernel, dim3(1), dim3(256), 0, st, n,
                     (const int32_t *)s->aff_count, (const UnitOut *)s->outs,
                     (const unsigned long long *)s->aff_sum_old, (const int32_t *)s->hist,
                     (const int32_t *)s->removed_hist, (unsigned long long)s->sum, s->diameter,
                     s->res);
  GG_HIP(hipGetLastError());
  GG_HIP(hipMemcpyAsync(s->res_h, s->res, sizeof(IncrResult), hipMemcpyDeviceToHost, st));
  GG_HIP(hipStreamSynchronize(st));
  s->pend = *s->res_h;
  s->pend_has_src = inline_src;
  s->redo = redo;
  s->undo = undo;
  s->pending = true;
  out->connected = s->pend.connected;
  out->diameter = s->pend.diameter;
  out->max_ecc = s->pend.connected ? s->pend.diameter : 0;
  out->word_bits = kWord;
  out->sum = s->pend.sum;
  if (affected) *affected = s->pend.affected;
  return GG_OK;
}

int gg_incr_accept(gg_incr *s) {
  if (!s) return set_error(GG_ERR_INVALID, "null argument");
  if (!s->pending) return set_error(GG_ERR_STATE, "no pending move");
  if (!s->pend.connected) return set_error(GG_ERR_STATE, "the pending move disconnects the graph");
  hipStream_t st = s->stream;
  if (!s->pend_has_src) {
    const int rc = run_src(s, 2);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(gg_incr_commit_kernel, dim3((s->n + 255) / 256), dim3(256), 0, st,
                     (const int32_t *)s->aff_count, (const uint16_t *)s->aff,
                     (const uint32_t *)s->new_sum, (const uint16_t *)s->new_ecc, s->src_sum,
                     s->src_ecc, s->hist);
  GG_HIP(hipGetLastError());
  GG_HIP(hipStreamSynchronize(st));
  for (int i = 0; i < 4; ++i) s->adj[(size_t)s->pend_rows[i] * s->d + s->pend_cols[i]] = s->pend_vals[i];
  s->sum = s->pend.sum;
  s->diameter = s->pend.diameter;
  s->pending = false;
  s->accept_rate = 0.95 * s->accept_rate + 0.05;
  return GG_OK;
}

int gg_incr_reject(gg_incr *s) {
  if (!s) return set_error(GG_ERR_INVALID, "null argument");
  if (!s->pending) return set_error(GG_ERR_STATE, "no pending move");
  hipLaunchKernelGGL(gg_incr_patch_kernel, dim3(1), dim3(64), 0, s->stream, s->adjR, s->undo);
  GG_HIP(hipGetLastError());
  s->pending = false;
  s->accept_rate *= 0.95;
  return GG_OK;
}

int gg_incr_adjacency(const gg_incr *s, int32_t *adj) {
  if (!s || !adj) return set_error(GG_ERR_INVALID, "null argument");
  std::copy(s->adj.begin(), s->adj.end(), adj);
  return GG_OK;
}

void gg_incr_destroy(gg_incr *s) {
  if (!s) return;
  (void)hipStreamSynchronize(s->stream);
  free_all(s);
  delete s;
}

}  // extern "C"
