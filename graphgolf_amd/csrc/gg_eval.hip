// gg_eval.hip -- gfx950 bit-parallel APSP evaluator (kernels + host engine).
//
// What it computes (bit-exact with the reference's evaluation()):
//   general/evaluation.c:153-263 (matrix_op), :265-373 (tiled variant) and
//   grid/evaluation.c:122-190 (rotate_hash relabelling).  For every source
//   tile the reference runs a level-synchronous BFS over bitsets
//   ("bit s of row v = source s reaches v"), counts reached (source, vertex)
//   pairs per level, and accumulates the unreached count until every pair is
//   reached.  S and the diameter are integer functions of those per-level
//   counts, so any algorithm that produces the same reach sets per level
//   produces identical outputs.
//
// How it is laid out on MI355X:
//   * One workgroup owns one (graph, source tile) unit and runs ALL of its
//     BFS levels with the tile's reach bitsets resident in LDS: R[v] is one
//     W-bit word per vertex (W = 8/16/32/64 sources).  n=65536 with W=16
//     is 128 KiB, inside gfx950's 160 KiB LDS, so the level loop never
//     touches HBM except for the adjacency stream.
//   * The adjacency is converted once per call into a row-major uint16 table
//     padded to whole 16-byte chunks (8 neighbour ids per load); it is read
//     from L2 by every tile, with the next vertex's row prefetched while the
//     current one gathers.
//   * Direction-optimising levels.  While the frontier of the tile is small
//     it is kept as an LDS queue of (vertex, new bits) and pushed to the
//     neighbours with LDS atomic ORs (top-down, only for symmetric
//     adjacency, which the prep kernel verifies).  Once the queue overflows
//     the tile switches to pull levels (bottom-up): every vertex ORs its
//     neighbours' words, stopping as soon as its own word is full; the new
//     words are staged in registers and written back after a barrier.
//   * Counts: wave64 shuffle reductions, two barriers per level; per-unit
//     integer partial sums combined per graph by a tiny kernel.
//   * No MFMA: the work is integer bit logic and popcounts.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cinttypes>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "gg_eval.h"
#include "gg_host.h"
#include "gg_internal.h"
#include "gg_device.h"
#include "gg_tile.h"

struct GraphMeta {
  int32_t unit_begin, unit_count, n, groups;
};

struct GraphOut {
  uint64_t sum;
  int32_t connected, diameter, max_ecc, error;
};


// ---------------------------------------------------------------------------
// Adjacency preparation: row-major int32 [n][d] (+ optional relabel) ->
// row-major uint16 [n][dp], dp = d rounded up to 8 so that a row is a whole
// number of 16-byte chunks (8 neighbour ids each).  Padding entries are the
// row's own id (a self-loop adds nothing to a reach set).  Invalid entries
// (neighbour ids or relabel entries outside [0,n)) are replaced by the row's
// own id, or their row is not stored, and counted in flags[0]; no store
// leaves the table.  flags[1] counts asymmetric entries when check_sym is
// set.

struct PrepDesc {  // one graph of a prep launch
  const int32_t *adj, *relabel;
  uint16_t *adjR;
  int32_t *flags;
  int n, d, dp, check_sym;
};

// One adjacency entry (row i, column j < dp) of a graph: the relabelled id
// to store at out-row relabel(i), a validity count and a symmetry-mismatch
// count (0 or 1 each).
__device__ __forceinline__ uint32_t prep_entry(const int32_t *__restrict__ adj,
                                               const int32_t *__restrict__ relabel, int n, int d,
                                               int i, int j, int ri, int check_sym, int *bad,
                                               int *asym) {
  if (j >= d) return (uint32_t)ri;  // padding: the row's own id
  const int32_t *row = adj + (size_t)i * d;
  const int u = row[j];
  if (u < 0 || u >= n) {
    *bad = 1;
    return (uint32_t)ri;
  }
  if (check_sym) {
    int c1 = 0, c2 = 0;
    const int32_t *urow = adj + (size_t)u * d;
    for (int k = 0; k < d; ++k) {
      c1 += (row[k] == u);
      c2 += (urow[k] == i);
    }
    *asym = (c1 != c2);
  }
  const int ru = relabel ? relabel[u] : u;
  if ((unsigned)ru >= (unsigned)n) {  // a relabel entry out of range
    *bad = 1;
    return (uint32_t)ri;
  }
  return (uint32_t)ru;
}

// Symmetry of the (up to) 8 entries j0..j0+7 of row i: the entries u whose
// multiplicity in row i differs from that of i in row u.  Row i is read
// once for all 8 counts, then each row u once; with rows of whole 16-byte
// chunks (d % 4 == 0, aligned table) every read is a 16-byte load.
__device__ __forceinline__ int chunk_asymmetry(const int32_t *__restrict__ adj, int n, int d, int i,
                                               int j0) {
  const int32_t *row = adj + (size_t)i * d;
  int u[8], c1[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    u[q] = j0 + q < d ? row[j0 + q] : -1;
    c1[q] = 0;
  }
  const bool vec = (d & 3) == 0 && ((uintptr_t)adj & 15) == 0;  // uniform
  if (vec) {
    for (int k = 0; k < d; k += 4) {
      const int4 r = *(const int4 *)(row + k);
#pragma unroll
      for (int q = 0; q < 8; ++q)
        c1[q] += (r.x == u[q]) + (r.y == u[q]) + (r.z == u[q]) + (r.w == u[q]);
    }
  } else {
    for (int k = 0; k < d; ++k) {
      const int r = row[k];
#pragma unroll
      for (int q = 0; q < 8; ++q) c1[q] += r == u[q];
    }
  }
  int asym = 0;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    if ((unsigned)u[q] >= (unsigned)n) continue;  // padding (ids were validated)
    const int32_t *urow = adj + (size_t)u[q] * d;
    int c2 = 0;
    if (vec) {
      for (int k = 0; k < d; k += 4) {
        const int4 r = *(const int4 *)(urow + k);
        c2 += (r.x == i) + (r.y == i) + (r.z == i) + (r.w == i);
      }
    } else {
      for (int k = 0; k < d; ++k) c2 += urow[k] == i;
    }
    asym += c1[q] != c2;
  }
  return asym;
}

// Adds this wave's counts to a graph's flags with one atomic per counter.
__device__ __forceinline__ void flag_counts(int32_t *flags, int bad, int asym) {
  const uint64_t mb = __builtin_amdgcn_ballot_w64(bad != 0);
  const uint64_t ma = __builtin_amdgcn_ballot_w64(asym != 0);
  const int lane = threadIdx.x & 63;
  const int first = __builtin_amdgcn_readfirstlane(lane);
  if (lane == first) {
    if (mb) atomicAdd(&flags[0], (int)__popcll(mb));
    if (ma) atomicAdd(&flags[1], (int)__popcll(ma));
  }
}

// All graphs of a launch set in one launch, one thread per 16-byte chunk (8
// entries) of the padded row-major table, written with one vector store;
// blockIdx.y selects the graph.  Invalid and asymmetric entries are rare, so
// a thread that finds some adds its counts with its own atomics.
__global__ void __launch_bounds__(256) gg_prep_kernel(const PrepDesc *__restrict__ descs) {
  const PrepDesc P = descs[blockIdx.y];  // descs: this launch's first graph
  const uint32_t nch = (uint32_t)P.dp >> 3;
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (uint32_t)P.n * nch) return;
  const int i = (int)(k / nch), j0 = 8 * (int)(k - (uint32_t)i * nch);
  const int ri = P.relabel ? P.relabel[i] : i;
  if ((unsigned)ri >= (unsigned)P.n) {  // a relabel entry out of range: no store
    atomicAdd(&P.flags[0], 1);
    return;
  }
  int bad = 0, asym = 0;
  uint32_t w[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    int b0 = 0, a0 = 0, b1 = 0, a1 = 0;
    const uint32_t lo = prep_entry(P.adj, P.relabel, P.n, P.d, i, j0 + 2 * q, ri, 0, &b0, &a0);
    const uint32_t hi = prep_entry(P.adj, P.relabel, P.n, P.d, i, j0 + 2 * q + 1, ri, 0, &b1, &a1);
    w[q] = (lo & 0xffffu) | (hi << 16);
    bad += b0 + b1;
  }
  *(Chunk *)(P.adjR + (size_t)ri * P.dp + j0) = Chunk{w[0], w[1], w[2], w[3]};
  if (P.check_sym && !bad) asym = chunk_asymmetry(P.adj, P.n, P.d, i, j0);
  if (bad) atomicAdd(&P.flags[0], bad);
  if (asym) atomicAdd(&P.flags[1], asym);
}

// Patched variant for batches of proposals that differ from one base graph
// in a few adjacency entries: slot y = blockIdx.y converts the shared base
// table (entry-parallel, as gg_prep_kernel); gg_patch_kernel then overrides
// the entries listed in patches[y][k] = (v*d + j, new id) (index -1 =
// unused).  The caller guarantees symmetry (SA moves keep it).
__global__ void __launch_bounds__(256)
gg_prep_patched_kernel(const int32_t *__restrict__ base, const int32_t *__restrict__ relabel, int n,
                       int d, int dp, uint16_t *__restrict__ adjR0, size_t slot_stride,
                       int32_t *__restrict__ flags0, int slot0) {
  const int slot = slot0 + (int)blockIdx.y;
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = k < (uint32_t)n * (uint32_t)dp;
  int bad = 0, asym = 0;
  if (live) {
    const int i = (int)(k / (uint32_t)dp), j = (int)(k - (uint32_t)i * (uint32_t)dp);
    const int ri = relabel ? relabel[i] : i;
    if ((unsigned)ri < (unsigned)n)
      adjR0[slot * slot_stride + (size_t)ri * dp + j] =
          (uint16_t)prep_entry(base, relabel, n, d, i, j, ri, 0, &bad, &asym);
    else
      bad = 1;
  }
  if (__builtin_amdgcn_ballot_w64(live) != 0ull) flag_counts(flags0 + 2 * slot, bad, asym);
}

__global__ void __launch_bounds__(256)
gg_patch_kernel(const int32_t *__restrict__ patches, int max_patch, int count,
                const int32_t *__restrict__ relabel, int n, int d, int dp,
                uint16_t *__restrict__ adjR0, size_t slot_stride, int32_t *__restrict__ flags0) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count * max_patch) return;
  const int slot = t / max_patch;
  const int at = patches[2 * t], val = patches[2 * t + 1];
  if (at < 0 || at >= n * d) return;  // unused entry
  const int i = at / d, j = at - i * d;
  const int ri = relabel ? relabel[i] : i;
  if ((unsigned)ri >= (unsigned)n) {
    atomicAdd(&flags0[2 * slot], 1);
    return;
  }
  int rv = ri;
  if (val < 0 || val >= n) atomicAdd(&flags0[2 * slot], 1);
  else rv = relabel ? relabel[val] : val;
  if ((unsigned)rv >= (unsigned)n) {
    atomicAdd(&flags0[2 * slot], 1);
    rv = ri;
  }
  adjR0[slot * slot_stride + (size_t)ri * dp + j] = (uint16_t)rv;
}


// ---------------------------------------------------------------------------
// Bank-aware order of the rows.  A pull level of the tile BFS reads entry
// slot j of 32 consecutive rows with one LDS read per half-wave (vertex
// r*BLOCK + t, BLOCK a multiple of 64), and the LDS serves a half-wave from
// 32 banks of 4 bytes: ids whose reach words lie in one bank but in
// different 4-byte words are served one after another.  Random ids put
// about 3.5 words on the busiest bank of a slot; a row's entries can be
// permuted freely (the OR over a row does not depend on its order), so this
// kernel orders every row of a 32-row group slot by slot to spread each
// slot over the banks:
//   * capacity plan: with D the largest number of group entries on one bank
//     and dp slots, kh = D - clo*dp slots allow clo+1 rows per bank and the
//     rest clo, clo = max(1, D/dp) -- enough for every entry;
//   * per slot, every row takes its unused entry on the bank with the most
//     unused group entries, LDS tickets deciding between rows on one bank;
//     a row turned away takes the bank with room and the most unused
//     entries, and after kOrderRounds its least-loaded bank.
// The given order is then costed exactly (sum over slots of the largest
// number of distinct words on one bank) and kept per group unless the new
// order's slot loads (an upper bound of its cost) are lower, so structured
// graphs whose given order is already conflict-free keep it.
// Each wave orders two groups (its half-waves); a block holds four waves.
struct OrderDesc {
  uint16_t *adjR;
  unsigned long long *cost;  // [2]: original, kept (nullable)
  int n, dp, shift;          // word index of id u = (u << shift) >> 2
};

// Per-slot bank counters: bits 0-7 tickets handed out, 8-15 rows that
// overflowed onto the bank (at most 32 per round and half-wave).  Bank 32
// of the arrays is a sink for used entries: no unused entries, no room and
// a load no real bank reaches.
constexpr int kOrderRounds = 3;
constexpr int kOrderWaves = 4;  // one wave per SIMD; the waves share nothing
constexpr uint32_t kTicket = 1u, kOverflow = 1u << 8, kSink = 0xffffu;

template <int DP> struct OrderLds {  // one wave's share
  uint32_t degr[2][33];    // unused group entries per bank
  union {
    uint32_t tk[DP][2][33];  // greedy: tickets | overflows << 8, per slot
    struct {                 // costs (after the greedy's slot loads are read)
      __attribute__((aligned(16))) uint32_t cnt[DP][2][32];
      uint32_t bm[2048];  // one bit per 4-byte word (ids < 65536, words up to 8 bytes)
    } c;
  } u;
  uint16_t rowL[64][DP];  // the given row
  uint16_t nr[64][DP];    // the new row
};

template <int NCH>
__global__ void __launch_bounds__(64 * kOrderWaves)
gg_order_rows_kernel(const OrderDesc *__restrict__ descs) {
  constexpr int DP = 8 * NCH;
  const OrderDesc P = descs[blockIdx.y];
  if (P.dp != DP || P.shift < 0) return;  // uniform
  __shared__ OrderLds<DP> lds_all[kOrderWaves];
  const int lane = threadIdx.x & 63, h = lane >> 5, L = lane & 31;
  OrderLds<DP> &S = lds_all[threadIdx.x >> 6];
  const int v = ((int)blockIdx.x * kOrderWaves + (int)(threadIdx.x >> 6)) * 64 + lane;
  if (__builtin_amdgcn_readfirstlane(v - lane) >= P.n) return;  // whole wave past n
  const bool live = v < P.n;
  const int sh = P.shift;
  // Groups never cross waves, so a wave barrier (LDS operations of one wave
  // complete in order) is all the synchronisation needed.
  auto wsync = [] {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  uint32_t wd[DP];  // 4-byte LDS word of each entry's reach word; bank = wd & 31
  {
    const Chunk *row = (const Chunk *)(P.adjR + (size_t)(live ? v : 0) * DP);
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      const Chunk x = live ? row[c] : Chunk{0u, 0u, 0u, 0u};
      const uint32_t q[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        wd[8 * c + 2 * k] = ((q[k] & 0xffffu) << sh) >> 2;
        wd[8 * c + 2 * k + 1] = ((q[k] >> 16) << sh) >> 2;
        S.rowL[lane][8 * c + 2 * k] = (uint16_t)q[k];
        S.rowL[lane][8 * c + 2 * k + 1] = (uint16_t)(q[k] >> 16);
      }
    }
  }
  for (int k = lane; k < DP * 2 * 33; k += 64) (&S.u.tk[0][0][0])[k] = k % 33 == 32 ? kSink : 0u;
  if (lane < 2) S.degr[lane][32] = 0u;
  S.degr[h][L] = 0u;
  wsync();
  uint32_t bk[DP];  // bank of entry i; 32 once used
#pragma unroll
  for (int i = 0; i < DP; ++i) {
    bk[i] = live ? wd[i] & 31u : 32u;
    if (live) atomicAdd(&S.degr[h][bk[i]], 1u);
  }
  wsync();
  int dmax = (int)S.degr[h][L];
#pragma unroll
  for (int off = 16; off >= 1; off >>= 1) dmax = max(dmax, __shfl_xor(dmax, off, 32));
  const int clo = max(1, dmax / DP), kh = max(0, dmax - clo * DP);

  for (int s = 0; s < DP; ++s) {
    const uint32_t cap = (uint32_t)(clo + (s < kh ? 1 : 0));
    // First ticket pass: every bank has room; the entry on the bank with
    // the most unused entries (key = degree << 5 | entry).
    uint32_t kd[DP];
    uint32_t best = 0;
#pragma unroll
    for (int i = 0; i < DP; ++i) {
      kd[i] = (S.degr[h][bk[i]] << 5) | (uint32_t)i;
      best = max(best, kd[i]);
    }
    uint32_t bi = best & 31u;
    uint32_t pick = S.rowL[lane][bi];
    uint32_t bw = (pick << sh) >> 2;
    bool placed = !live;
    if (live) placed = (atomicAdd(&S.u.tk[s][h][bw & 31u], kTicket) & 255u) < cap;
    wsync();
    // Later ticket passes for the rest: the bank with room and the most
    // unused entries, else (and in the last pass always) the least-loaded bank.
    for (int round = 1; round <= kOrderRounds; ++round) {
      if (__builtin_amdgcn_ballot_w64(!placed) == 0ull) break;
      if (!placed) {
        // All counters first, then branch-free keys: a branch per entry
        // would wait for each read in turn.
        const uint32_t not_last = round == kOrderRounds ? 0u : ~0u;
        uint32_t tt[DP];
#pragma unroll
        for (int i = 0; i < DP; ++i) tt[i] = S.u.tk[s][h][bk[i]];
        best = 0;
#pragma unroll
        for (int i = 0; i < DP; ++i) {
          const uint32_t tick = tt[i] & 255u, load = min(min(tick, cap) + (tt[i] >> 8), kSink);
          const uint32_t k_full = (1u << 30) | ((kSink - load) << 5) | (uint32_t)i;
          const uint32_t k_room = (2u << 30) | kd[i];
          const uint32_t room = not_last & (0u - (uint32_t)(tick < cap));
          best = max(best, k_full ^ ((k_full ^ k_room) & room));
        }
        bi = best & 31u;
        pick = S.rowL[lane][bi];
        bw = (pick << sh) >> 2;
        if (best >= (2u << 30)) {
          placed = (atomicAdd(&S.u.tk[s][h][bw & 31u], kTicket) & 255u) < cap;
        } else {
          atomicAdd(&S.u.tk[s][h][bw & 31u], kOverflow);
          placed = true;
        }
      }
      wsync();
    }
    if (live) {
#pragma unroll
      for (int i = 0; i < DP; ++i) bk[i] = (uint32_t)i == bi ? 32u : bk[i];
      atomicSub(&S.degr[h][bw & 31u], 1u);
      S.nr[lane][s] = (uint16_t)pick;
    }
    wsync();
  }

  // Exact cost of an order: per slot, a lane counts its word if it is the
  // first of its half to set the word's bit (LDS atomic OR, one half-wave
  // at a time; the bit is cleared again straight after); the slot's cost
  // is the most counted words on one bank.  Uniform per half-wave.
  auto order_cost = [&](const uint16_t (*rw)[DP]) -> int {
#pragma unroll
    for (int s = 0; s < DP; ++s) {
      const uint32_t x = ((uint32_t)rw[lane][s] << sh) >> 2;
#pragma unroll
      for (int half = 0; half < 2; ++half) {
        if (live && h == half) {
          const uint32_t old = atomicOr(&S.u.c.bm[x >> 5], 1u << (x & 31u));
          if (!(old & (1u << (x & 31u)))) atomicAdd(&S.u.c.cnt[s][h][x & 31u], 1u);
          S.u.c.bm[x >> 5] = 0u;  // every lane of this half with this word has set it by now
        }
        wsync();
      }
    }
    int m = 0;
    if (L < DP) {
#pragma unroll
      for (int j = 0; j < 32; j += 4) {
        const Chunk c = *(const Chunk *)&S.u.c.cnt[L][h][j];
        m = max(m, (int)max(max(c.x, c.y), max(c.z, c.w)));
      }
    }
#pragma unroll
    for (int off = 16; off >= 1; off >>= 1) m += __shfl_xor(m, off, 32);
    wsync();
    for (int k = lane; k < DP * 64; k += 64) (&S.u.c.cnt[0][0][0])[k] = 0u;
    wsync();
    return m;
  };
  // The new order's slot loads (rows per bank, so an upper bound of its
  // cost) decide; its exact cost is only needed when it is reported.
  int bound = 0;
  if (L < DP) {
    const uint32_t cap = (uint32_t)(clo + (L < kh ? 1 : 0));
#pragma unroll
    for (int j = 0; j < 32; ++j) {
      const uint32_t t = S.u.tk[L][h][j];
      bound = max(bound, (int)(min(t & 255u, cap) + (t >> 8)));
    }
  }
#pragma unroll
  for (int off = 16; off >= 1; off >>= 1) bound += __shfl_xor(bound, off, 32);
  wsync();  // every lane has read the tickets: their space now holds the costs' counters
  for (int k = lane; k < DP * 64; k += 64) (&S.u.c.cnt[0][0][0])[k] = 0u;
  for (int k = lane; k < 2048; k += 64) S.u.c.bm[k] = 0u;
  wsync();
  const int cost_old = order_cost(S.rowL);
  const bool keep_new = bound < cost_old;  // uniform within the half-wave
  const int cost_new = P.cost ? order_cost(S.nr) : bound;
  if (live && keep_new) {
    Chunk *row = (Chunk *)(P.adjR + (size_t)v * DP);
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      Chunk x;
      x.x = (uint32_t)S.nr[lane][8 * c + 0] | ((uint32_t)S.nr[lane][8 * c + 1] << 16);
      x.y = (uint32_t)S.nr[lane][8 * c + 2] | ((uint32_t)S.nr[lane][8 * c + 3] << 16);
      x.z = (uint32_t)S.nr[lane][8 * c + 4] | ((uint32_t)S.nr[lane][8 * c + 5] << 16);
      x.w = (uint32_t)S.nr[lane][8 * c + 6] | ((uint32_t)S.nr[lane][8 * c + 7] << 16);
      row[c] = x;
    }
  }
  if (P.cost && L == 0) {
    atomicAdd(&P.cost[0], (unsigned long long)cost_old);
    atomicAdd(&P.cost[1], (unsigned long long)(keep_new ? cost_new : cost_old));
  }
}

// Per-graph combination of the unit partial sums
// (general/evaluation.c:199-200, :241, :247-249 as integers).  One
// workgroup per graph; its units are contiguous.  `res` may be pinned host
// memory (the result then needs no copy).  With `flags` set, the graph's
// two prep flags (and, in workgroup 0, the queue counter after them) are
// zeroed for the next run: every kernel that reads them has finished.
__global__ void __launch_bounds__(256)
gg_combine_kernel(const GraphMeta *__restrict__ meta, const UnitDesc *__restrict__ units,
                  const UnitOut *__restrict__ outs, GraphOut *__restrict__ res, int ngraphs,
                  int32_t *__restrict__ flags) {
  __shared__ uint64_t s64[3][256];
  __shared__ int s32[4][256];
  const int gidx = blockIdx.x;
  const int tid = threadIdx.x;
  if (gidx >= ngraphs) return;
  if (flags && tid < 2) flags[2 * gidx + tid] = 0;
  if (flags && gidx == 0 && tid == 2) flags[2 * ngraphs] = 0;
  const GraphMeta m = meta[gidx];
  uint64_t lo = 0, hi = 0, chi = 0;
  int diam = 1, ecc = 0, disc = 0, err = 0;
  // Eight records per thread in flight at a time (a single workgroup reads
  // every unit of the graph; one load at a time is latency-bound).
  constexpr int kBatch = 8;
  for (int k0 = 0; k0 < m.unit_count; k0 += 256 * kBatch) {
    UnitOut o[kBatch];
    int cen[kBatch];
#pragma unroll
    for (int j = 0; j < kBatch; ++j) {
      const int k = k0 + j * 256 + tid;
      const bool ok = k < m.unit_count;
      const int u = m.unit_begin + (ok ? k : 0);
      o[j] = outs[u];
      cen[j] = ok ? (int)units[u].centre : -1;
    }
#pragma unroll
    for (int j = 0; j < kBatch; ++j) {
      if (cen[j] < 0) continue;
      if (cen[j]) chi += o[j].acc_hi;
      else {
        lo += o[j].acc_lo;
        hi += o[j].acc_hi;
      }
      diam = max(diam, o[j].done_level);
      ecc = max(ecc, o[j].last_change);
      disc |= o[j].disconnected;
      err |= o[j].error;
    }
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
    GraphOut r;
    const uint64_t g = (uint64_t)m.groups, n = (uint64_t)m.n;
    r.sum = s64[0][0] * g + s64[1][0] * 2ull * g + s64[2][0] + n * (n - 1ull);
    r.connected = s32[2][0] ? 0 : 1;
    r.diameter = s32[2][0] ? m.n + 1 : s32[0][0];
    r.max_ecc = s32[1][0];
    r.error = s32[3][0];
    res[gidx] = r;
  }
}


// ---------------------------------------------------------------------------
// Scratch of the allocation-free entry point (gg_apsp_device_noalloc): the
// converted row table, the descriptors, the unit list, the unit partials,
// the result record and a relabel bitmap of one evaluation, for graphs of
// at most kNoAllocMaxNodes vertices and kNoAllocMaxDegree neighbours.  Each
// GPU holds its own copy for the life of the process; host code reaches it
// only through the address hipGetSymbolAddress returns at init.
constexpr int kNoAllocMaxNodes = 65536;
constexpr int kNoAllocMaxDegree = 32;
constexpr size_t kNoAllocAdjBytes = (size_t)kNoAllocMaxNodes * kNoAllocMaxDegree * 2 + 256;
constexpr int kNoAllocMaxUnits = kNoAllocMaxNodes / 8 + 2;
constexpr size_t kNoAllocScratchBytes = (size_t)5 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// Writes one evaluation's descriptors (passed by value: they travel in the
// kernel arguments, not in an upload), its unit list, and clears its flags
// and relabel bitmap.
__global__ void __launch_bounds__(256)
gg_noalloc_setup_kernel(GraphDev gd, PrepDesc pd, OrderDesc od, GraphMeta gm, GraphDev *gd_out,
                        PrepDesc *pd_out, OrderDesc *od_out, GraphMeta *gm_out,
                        UnitDesc *__restrict__ units, int bits, int B, int c, int n,
                        int32_t *__restrict__ flags, uint32_t *__restrict__ bitmap, int bm_words) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) {
    *gd_out = gd;
    *pd_out = pd;
    *od_out = od;
    *gm_out = gm;
    flags[0] = 0;
    flags[1] = 0;
  }
  const int ub = (B + bits - 1) / bits, uc = (c + bits - 1) / bits;
  for (int u = t; u < ub + uc; u += gridDim.x * blockDim.x) {
    if (u < ub) units[u] = UnitDesc{0, u * bits, min(bits, B - u * bits), 0};
    else units[u] = UnitDesc{0, n - c + (u - ub) * bits, min(bits, c - (u - ub) * bits), 1};
  }
  for (int w = t; w < bm_words; w += gridDim.x * blockDim.x) bitmap[w] = 0u;
}

// Counts relabel entries outside [0,n) and repeated ones (a bitmap of the
// values seen) into flags[0], so a table that is not a permutation ends the
// evaluation with an error.
__global__ void __launch_bounds__(256)
gg_relabel_check_kernel(const int32_t *__restrict__ relabel, int n, uint32_t *__restrict__ bitmap,
                        int32_t *__restrict__ flags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int r = relabel[i];
  if ((unsigned)r >= (unsigned)n) {
    atomicAdd(&flags[0], 1);
    return;
  }
  const uint32_t bit = 1u << (r & 31);
  if (atomicOr(&bitmap[r >> 5], bit) & bit) atomicAdd(&flags[0], 1);
}

// ---------------------------------------------------------------------------
// Host engine.

namespace gg {

static thread_local char g_err[512] = "";
static thread_local int g_err_code = GG_OK;

int set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  g_err_code = code;
  return code;
}
const char *last_error() { return g_err; }
int last_error_code() { return g_err_code; }

#define GG_HIP(call)                                                                  \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      return set_error(GG_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                       __FILE__, __LINE__);                                           \
  } while (0)

// Dense LDS limit per workgroup on gfx950.
static constexpr int kLdsBytes = 160 * 1024;
static constexpr int kMaxNodes = 65536;

struct DevBuf {
  void *p = nullptr;
  size_t cap = 0;
  int ensure(size_t bytes) {
    if (bytes <= cap) return GG_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    size_t want = std::max(bytes, (size_t)256);
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess)
      return set_error(GG_ERR_HIP, "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    cap = want;
    return GG_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

struct PinBuf {
  void *p = nullptr;
  size_t cap = 0;
  // coherent: kernels write it directly (results), visible to the host once
  // the stream has been synchronized.
  int ensure(size_t bytes, bool coherent = false) {
    if (bytes <= cap) return GG_OK;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    size_t want = std::max(bytes, (size_t)256);
    hipError_t e = hipHostMalloc(&p, want, coherent ? hipHostMallocCoherent : hipHostMallocDefault);
    if (e != hipSuccess)
      return set_error(GG_ERR_HIP, "hipHostMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    cap = want;
    return GG_OK;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
  }
};

// Kernel instance key: word width, block size, vertices per thread, histogram.
struct KernelKey {
  int bits, block, vpt, hist;
  bool operator<(const KernelKey &o) const {
    return std::tie(bits, block, vpt, hist) < std::tie(o.bits, o.block, o.vpt, o.hist);
  }
};

typedef void (*TileKernel)(const GraphDev *, const UnitDesc *, UnitOut *, int, int, int);

template <typename W, int B, int V, int H>
static TileKernel kptr() {
  return (TileKernel)&gg_tile_bfs_kernel<W, B, V, H>;
}

template <typename W, int H>
static TileKernel pick_vpt(int block, int vpt) {
  if (block == 256) {
    switch (vpt) {
      case 1: return kptr<W, 256, 1, H>();
      case 2: return kptr<W, 256, 2, H>();
      case 4: return kptr<W, 256, 4, H>();
      case 8: return kptr<W, 256, 8, H>();
    }
  } else {
    switch (vpt) {
      case 4: return kptr<W, 1024, 4, H>();
      case 8: return kptr<W, 1024, 8, H>();
      case 16: return kptr<W, 1024, 16, H>();
      case 32: return kptr<W, 1024, 32, H>();
      case 64:
        if constexpr (sizeof(W) <= 4) return kptr<W, 1024, 64, H>();
        break;
    }
  }
  return nullptr;
}

// k.hist: kTileNormal or kTileHist (the source-list modes are instantiated
// by the incremental scorer).
static TileKernel kernel_for(const KernelKey &k) {
  const bool h = k.hist != 0;
  switch (k.bits) {
    case 8: return h ? pick_vpt<uint8_t, kTileHist>(k.block, k.vpt) : pick_vpt<uint8_t, kTileNormal>(k.block, k.vpt);
    case 16: return h ? pick_vpt<uint16_t, kTileHist>(k.block, k.vpt) : pick_vpt<uint16_t, kTileNormal>(k.block, k.vpt);
    case 32: return h ? pick_vpt<uint32_t, kTileHist>(k.block, k.vpt) : pick_vpt<uint32_t, kTileNormal>(k.block, k.vpt);
    case 64: return h ? pick_vpt<uint64_t, kTileHist>(k.block, k.vpt) : pick_vpt<uint64_t, kTileNormal>(k.block, k.vpt);
  }
  return nullptr;
}

// The persistent (queue-driven) instances, kTileNormal only.
typedef void (*QueueKernel)(const GraphDev *, const UnitDesc *, UnitOut *, int, int, int, int,
                            int *);

template <typename W, int B, int V>
static QueueKernel qptr() {
  return (QueueKernel)&gg_tile_queue_kernel<W, B, V, kTileNormal>;
}

template <typename W>
static QueueKernel pick_qvpt(int block, int vpt) {
  if (block == 256) {
    switch (vpt) {
      case 1: return qptr<W, 256, 1>();
      case 2: return qptr<W, 256, 2>();
      case 4: return qptr<W, 256, 4>();
      case 8: return qptr<W, 256, 8>();
    }
  } else {
    switch (vpt) {
      case 4: return qptr<W, 1024, 4>();
      case 8: return qptr<W, 1024, 8>();
      case 16: return qptr<W, 1024, 16>();
      case 32: return qptr<W, 1024, 32>();
      case 64:
        if constexpr (sizeof(W) <= 4) return qptr<W, 1024, 64>();
        break;
    }
  }
  return nullptr;
}

static QueueKernel queue_kernel_for(const KernelKey &k) {
  if (k.hist) return nullptr;
  switch (k.bits) {
    case 8: return pick_qvpt<uint8_t>(k.block, k.vpt);
    case 16: return pick_qvpt<uint16_t>(k.block, k.vpt);
    case 32: return pick_qvpt<uint32_t>(k.block, k.vpt);
    case 64: return pick_qvpt<uint64_t>(k.block, k.vpt);
  }
  return nullptr;
}

// Threads per block and vertices per thread for an n-vertex graph: blocks
// of 256 up to 2048 vertices unless `want_block` asks for 1024.
static void block_vpt_pref(int n, int want_block, int *block, int *vpt) {
  if (n <= 2048 && want_block != 1024) {
    *block = 256;
    int v = (n + 255) / 256;
    *vpt = v <= 1 ? 1 : v <= 2 ? 2 : v <= 4 ? 4 : 8;
  } else {
    *block = 1024;
    int v = (n + 1023) / 1024;
    *vpt = v <= 4 ? 4 : v <= 8 ? 8 : v <= 16 ? 16 : v <= 32 ? 32 : 64;
  }
}

static void block_vpt_for(int n, int *block, int *vpt) { block_vpt_pref(n, 0, block, vpt); }


static size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

static size_t lds_bytes_for(int bits, int n_cap, int qcap) {
  return tile_lds_bytes((size_t)bits / 8, n_cap, qcap);
}

// The largest queue that fits the LDS budget, capped at what push levels
// can use (a push level needs qn*4 <= n).
static int qcap_for(int bits, int n, size_t budget) {
  size_t base = lds_bytes_for(bits, n, 0);
  if (base >= budget) return -1;
  size_t per = 2 * (size_t)(bits / 8) + 2 * 2;
  size_t fit = (budget - base - 64) / per;  // slack for the alignment rounding
  int want = std::max(64, n / 4 + 64);
  return (int)std::min(fit, (size_t)want);
}

// Launch configuration of one graph, or of a whole batch: reach-word width,
// threads per block (0: from n), row-order mode (-1: the global mode) and,
// for batches, whether one queue-driven launch drains every graph.
struct Config {
  int bits, block, order, persistent;
};

// Tuning table key: a shape bucket.  nb = ceil(log2 n) (of the largest graph
// for a batch), degree, groups, whether there are centres, batch or single.
struct TuneKey {
  int nb, d, g, c, batched;
  bool operator<(const TuneKey &o) const {
    return std::tie(nb, d, g, c, batched) < std::tie(o.nb, o.d, o.g, o.c, o.batched);
  }
};

static int bucket_of(int n) {
  int b = 0;
  while ((1 << b) < n) ++b;
  return b;
}

constexpr int kMaxDevices = 64;

struct Engine {
  bool ready = false;
  int device = 0;
  int cus = 256;
  hipStream_t stream = nullptr;
  int forced_bits = 0;
  std::map<TuneKey, Config> tune;
  bool forced_cfg = false;  // the tuner's candidate runs
  Config forced{0, 0, -1, 0};
  unsigned char *scratch_dev[kMaxDevices] = {};  // `scratch`'s address on each device
};

static Engine E;

Engine &engine() { return E; }

hipStream_t engine_stream() { return E.stream; }

int init(int device) {
  if (E.ready) return GG_OK;
  if (device < 0) {
    const char *s = getenv("GG_DEVICE");
    device = s ? atoi(s) : 0;
  }
  int cnt = 0;
  GG_HIP(hipGetDeviceCount(&cnt));
  if (cnt <= 0) return set_error(GG_ERR_HIP, "no HIP device visible");
  if (device >= cnt) device = device % cnt;  // like general/evaluation.cu:184
  GG_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  GG_HIP(hipGetDeviceProperties(&prop, device));
  E.cus = prop.multiProcessorCount;
  if (device < kMaxDevices && !E.scratch_dev[device]) {
    // With the device current: the address of this device's copy of the
    // allocation-free entry point's scratch array.
    void *ptr = nullptr;
    GG_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return set_error(GG_ERR_HIP, "no device address for the scratch array");
    E.scratch_dev[device] = (unsigned char *)ptr;
  }
  GG_HIP(hipStreamCreateWithFlags(&E.stream, hipStreamNonBlocking));
  // Allow the large dynamic LDS allocation on every instance we may launch.
  for (int bits : {8, 16, 32, 64})
    for (int block : {256, 1024})
      for (int vpt : {1, 2, 4, 8, 16, 32, 64})
        for (int h : {0, 1}) {
          TileKernel k = kernel_for({bits, block, vpt, h});
          if (!k) continue;
          GG_HIP(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     kLdsBytes));
          if (QueueKernel q = queue_kernel_for({bits, block, vpt, h}))
            GG_HIP(hipFuncSetAttribute((const void *)q, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       kLdsBytes));
        }
  E.device = device;
  E.ready = true;
  return GG_OK;
}

void finalize() {
  if (!E.ready) return;
  (void)hipStreamSynchronize(E.stream);
  (void)hipStreamDestroy(E.stream);
  E.stream = nullptr;
  E.ready = false;
}

int tile_shape(int bits, int n, int *block, int *vpt, int *n_cap, int *qcap, size_t *lds) {
  block_vpt_for(n, block, vpt);
  *n_cap = (n + 7) & ~7;
  const size_t rbytes = (size_t)*n_cap * (bits / 8);
  const size_t budget =
      rbytes * 2 + 4096 <= (size_t)kLdsBytes / 2 ? rbytes * 2 + 4096 : (size_t)kLdsBytes;
  *qcap = qcap_for(bits, *n_cap, budget);
  if (*qcap < 64) return set_error(GG_ERR_TOO_LARGE, "LDS budget too small for n = %d", n);
  *lds = lds_bytes_for(bits, *n_cap, *qcap);
  return GG_OK;
}

int max_bits_for(int n) {
  for (int bits : {64, 32, 16, 8})
    if (qcap_for(bits, n, kLdsBytes) >= 64) return bits;
  return 0;
}

// Word width for one graph.  Single evaluations favour narrow words when the
// graph has few source tiles (more workgroups in flight); batches favour the
// widest word (fewest adjacency passes).
// The launch configuration of one graph (batched = false) or of a batch
// whose largest graph has this shape: the tuning table's entry for the
// shape bucket when there is one, else the built-in choice -- single
// graphs narrow the word until the graph has >= CUs/2 source tiles (more
// workgroups in flight), batches take the widest word that fits LDS and one
// queue-driven launch.  gg_set_word_bits overrides the width.
Config config_for(int n, int d, int b, int g, int c, bool batched) {
  Config cfg{0, 0, -1, batched ? 1 : 0};
  const int maxb = max_bits_for(n);
  if (maxb == 0) return cfg;
  if (batched) {
    cfg.bits = maxb;
  } else {
    int bits = maxb;
    const int srcs = b + c;
    while (bits > 8 && (srcs + bits - 1) / bits < E.cus / 2) bits /= 2;
    cfg.bits = bits;
  }
  auto it = E.tune.find({bucket_of(n), d, g, c > 0 ? 1 : 0, batched ? 1 : 0});
  if (it != E.tune.end()) cfg = it->second;
  if (E.forced_cfg) cfg = E.forced;
  if (E.forced_bits) cfg.bits = E.forced_bits;
  if (cfg.bits < 8 || cfg.bits > maxb) cfg.bits = maxb;
  if (cfg.block == 256 && n > 2048) cfg.block = 0;
  return cfg;
}

int choose_bits(int n, int d, int b, int g, int c, int batch) {
  return config_for(n, d, b, g, c, batch > 1).bits;
}

// The row-order mode a configuration runs with: an explicit gg_set_row_order
// (0 or 2) wins over the table's choice; the default mode (1) takes it.
extern int g_row_order;
int effective_order(const Config &cfg) {
  if (g_row_order != 1 || cfg.order < 0) return g_row_order;
  return cfg.order;
}

// ---- Batch assembly -------------------------------------------------------

struct Slot {  // one graph in an assembled launch set
  int n, d, dp, b, g, c, bits, block, vpt, flags;
  size_t adj_off, adjT_off, rel_off;  // element offsets in the device buffers
  bool has_rel;
};

struct LaunchGroup {
  KernelKey key;
  int n_cap, qcap, unit_begin, unit_count;
  size_t lds;
  int grid;  // persistent launches: workgroups (resident per CU x CUs, at most the units)
};

struct Assembly {
  std::vector<Slot> slots;
  std::vector<GraphDev> gdev;
  std::vector<UnitDesc> units;
  std::vector<GraphMeta> meta;
  std::vector<LaunchGroup> groups;
  size_t adj_elems = 0, adjT_elems = 0, rel_elems = 0;
  int order_nch_mask = 0, order_max_n = 0;  // set by upload_order
  bool persistent = false;  // one queue-driven launch for every graph
  int order_mode = 1;       // row-order mode of this set
  Config cfg{0, 0, -1, 0};  // the configuration it was assembled with
};

int validate(const gg_graph &g, bool any_size = false) {
  if (g.nodes < 2 || g.degree < 1 || g.based_nodes < 1 || g.groups < 1 || g.added_centers < 0)
    return set_error(GG_ERR_INVALID, "invalid graph sizes (n=%d d=%d B=%d g=%d c=%d)", g.nodes,
                     g.degree, g.based_nodes, g.groups, g.added_centers);
  if ((int64_t)g.based_nodes * g.groups + g.added_centers != g.nodes)
    return set_error(GG_ERR_INVALID, "based_nodes*groups+added_centers (%d*%d+%d) != nodes (%d)",
                     g.based_nodes, g.groups, g.added_centers, g.nodes);
  if (g.added_centers > g.based_nodes)
    return set_error(GG_ERR_INVALID, "added_centers > based_nodes");
  if (g.nodes > kMaxNodes && !any_size)
    return set_error(GG_ERR_TOO_LARGE, "nodes=%d above the tile kernels' limit %d", g.nodes,
                     kMaxNodes);
  if (g.degree > 4096) return set_error(GG_ERR_INVALID, "degree %d too large", g.degree);
  if ((int64_t)g.nodes * g.degree >= ((int64_t)1 << 32))
    return set_error(GG_ERR_TOO_LARGE, "nodes*degree = %lld entries (at most 2^32 - 1)",
                     (long long)g.nodes * g.degree);
  return GG_OK;
}

int assemble(int count, const gg_graph *gs, int hist_maxd, Assembly &A) {
  A = Assembly();
  A.slots.resize(count);
  const bool batched = count > 1;
  // A batch is configured by its largest graph's shape bucket.
  int big = 0;
  for (int i = 1; i < count; ++i)
    if (gs[i].nodes > gs[big].nodes) big = i;
  int rc = validate(gs[big]);
  if (rc) return rc;
  const Config bc = config_for(gs[big].nodes, gs[big].degree, gs[big].based_nodes, gs[big].groups,
                               gs[big].added_centers, batched);
  A.cfg = bc;
  A.persistent = bc.persistent && hist_maxd == 0;
  A.order_mode = effective_order(bc);
  const int n_max = gs[big].nodes;
  for (int i = 0; i < count; ++i) {
    const gg_graph &g = gs[i];
    if ((rc = validate(g))) return rc;
    Slot &s = A.slots[i];
    s.n = g.nodes;
    s.d = g.degree;
    s.dp = (g.degree + 7) & ~7;
    s.b = g.based_nodes;
    s.g = g.groups;
    s.c = g.added_centers;
    s.flags = g.flags;
    // Persistent batches run every graph on the instance of the largest.
    const int sized = A.persistent ? n_max : s.n;
    s.bits = std::min(bc.bits, max_bits_for(sized));
    if (s.bits == 0) return set_error(GG_ERR_TOO_LARGE, "n=%d does not fit LDS", s.n);
    block_vpt_pref(sized, bc.block, &s.block, &s.vpt);
    if (!kernel_for({s.bits, s.block, s.vpt, 0}))
      return set_error(GG_ERR_TOO_LARGE, "no %d-bit instance for n=%d", s.bits, s.n);
    s.adj_off = A.adj_elems;
    A.adj_elems += (size_t)s.n * s.d;
    s.adjT_off = A.adjT_elems;
    A.adjT_elems += (size_t)s.n * s.dp;
    s.has_rel = g.relabel != nullptr;
    s.rel_off = A.rel_elems;
    if (s.has_rel) A.rel_elems += s.n;
  }
  // Group graphs by kernel instance; inside a group, largest graphs first.
  std::map<KernelKey, std::vector<int>> by_key;
  for (int i = 0; i < count; ++i) {
    const Slot &s = A.slots[i];
    by_key[{s.bits, s.block, s.vpt, hist_maxd > 0}].push_back(i);
  }
  A.meta.resize(count);
  A.gdev.resize(count);
  for (auto &kv : by_key) {
    std::vector<int> &ids = kv.second;
    std::stable_sort(ids.begin(), ids.end(), [&](int x, int y) {
      return (int64_t)A.slots[x].n * A.slots[x].d > (int64_t)A.slots[y].n * A.slots[y].d;
    });
    LaunchGroup lg;
    lg.key = kv.first;
    lg.n_cap = 0;
    for (int i : ids) lg.n_cap = std::max(lg.n_cap, A.slots[i].n);
    lg.n_cap = (lg.n_cap + 7) & ~7;
    // Budget: when the reach array is small, cap LDS so several workgroups
    // share a CU; large graphs take the whole 160 KiB.
    size_t rbytes = (size_t)lg.n_cap * (kv.first.bits / 8);
    size_t budget = rbytes * 2 + 4096 <= (size_t)kLdsBytes / 2 ? rbytes * 2 + 4096 : (size_t)kLdsBytes;
    lg.qcap = qcap_for(kv.first.bits, lg.n_cap, budget);
    if (lg.qcap < 64) return set_error(GG_ERR_TOO_LARGE, "LDS budget too small");
    lg.lds = lds_bytes_for(kv.first.bits, lg.n_cap, lg.qcap);
    lg.grid = 0;
    lg.unit_begin = (int)A.units.size();
    for (int i : ids) {
      const Slot &s = A.slots[i];
      GraphMeta &m = A.meta[i];
      m.unit_begin = (int)A.units.size();
      m.n = s.n;
      m.groups = s.g;
      const int bits = s.bits;
      for (int f = 0; f < s.b; f += bits) A.units.push_back({i, f, std::min(bits, s.b - f), 0});
      for (int f = 0; f < s.c; f += bits)
        A.units.push_back({i, s.n - s.c + f, std::min(bits, s.c - f), 1});
      m.unit_count = (int)A.units.size() - m.unit_begin;
    }
    lg.unit_count = (int)A.units.size() - lg.unit_begin;
    if (A.persistent) {
      const QueueKernel qk = queue_kernel_for(lg.key);
      if (!qk) return set_error(GG_ERR_STATE, "no persistent instance for %d-bit words", lg.key.bits);
      int per_cu = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)qk, lg.key.block,
                                                       lg.lds + 16) != hipSuccess || per_cu < 1)
        per_cu = 1;
      lg.grid = std::max(1, std::min(lg.unit_count, per_cu * E.cus));
    }
    A.groups.push_back(lg);
  }
  return GG_OK;
}

bool order_applies_mode(int mode, int size, int min_size, int dp);
constexpr int kOrderMinTiles = 1024;

// Device buffers of one launch set (either the engine's single scratch or a plan's own).
struct DevSet {
  DevBuf adj_in, relabel_in, adjT, flags, gdesc, units, outs, meta, hist, prep, order;
  PinBuf res_h;  // per-graph results, written by the combine kernel
  void release() {
    for (DevBuf *b : {&adj_in, &relabel_in, &adjT, &flags, &gdesc, &units, &outs, &meta, &hist,
                      &prep, &order})
      b->release();
    res_h.release();
  }
};

int prepare_buffers(Assembly &A, DevSet &D, int hist_maxd, int32_t *hist_dev,
                    uint64_t *trace_dev = nullptr, int trace_units = 0) {
  const int count = (int)A.slots.size();
  int rc;
  if ((rc = D.adj_in.ensure(A.adj_elems * 4))) return rc;
  if ((rc = D.relabel_in.ensure(std::max<size_t>(A.rel_elems, 1) * 4))) return rc;
  // +64 bytes: the pull levels always load a row's first two 16-byte chunks,
  // so a one-chunk last row reads 16 bytes past the table.
  if ((rc = D.adjT.ensure(A.adjT_elems * 2 + 64))) return rc;
  if ((rc = D.flags.ensure(((size_t)count * 2 + 1) * 4))) return rc;  // + the queue counter
  // Zero here; after that every run's combine re-zeroes them.
  GG_HIP(hipMemsetAsync(D.flags.p, 0, ((size_t)count * 2 + 1) * 4, E.stream));
  if ((rc = D.gdesc.ensure((size_t)count * sizeof(GraphDev)))) return rc;
  if ((rc = D.units.ensure(A.units.size() * sizeof(UnitDesc)))) return rc;
  if ((rc = D.outs.ensure(A.units.size() * sizeof(UnitOut)))) return rc;
  if ((rc = D.meta.ensure((size_t)count * sizeof(GraphMeta)))) return rc;
  if ((rc = D.res_h.ensure((size_t)count * sizeof(GraphOut), true))) return rc;
  for (int i = 0; i < count; ++i) {
    const Slot &s = A.slots[i];
    GraphDev &g = A.gdev[i];
    g.adjR = (const uint16_t *)D.adjT.p + s.adjT_off;
    g.flags = (int32_t *)D.flags.p + 2 * i;
    g.n = s.n;
    g.dp = s.dp;
    // The device row order permutes padding into the rows: no skipping then.
    g.pad4 = s.d <= s.dp - 4 &&
                     !order_applies_mode(A.order_mode, A.meta[i].unit_count, kOrderMinTiles, s.dp)
                 ? 1 : 0;
    g.split = s.b * s.g;
    g.c = s.c;
    g.push_ok = 1;
    g.groups = s.g;
    g.hist_maxd = hist_maxd;
    g.trace_units = trace_units;
    g.hist = hist_dev;
    g.trace = trace_dev;
  }
  GG_HIP(hipMemcpyAsync(D.gdesc.p, A.gdev.data(), count * sizeof(GraphDev), hipMemcpyHostToDevice, E.stream));
  GG_HIP(hipMemcpyAsync(D.units.p, A.units.data(), A.units.size() * sizeof(UnitDesc),
                        hipMemcpyHostToDevice, E.stream));
  GG_HIP(hipMemcpyAsync(D.meta.p, A.meta.data(), count * sizeof(GraphMeta), hipMemcpyHostToDevice,
                        E.stream));
  return GG_OK;
}

int upload_order(Assembly &A, DevSet &D);

// Per-graph prep descriptors for gg_prep_kernel (uploaded before enqueue, so
// that a captured sequence holds no host-to-device copy of them).
int upload_prep(Assembly &A, DevSet &D, const int32_t *const *adj_src, const int32_t *const *rel_src,
                const int *check_sym) {
  const int count = (int)A.slots.size();
  std::vector<PrepDesc> h(count);
  for (int i = 0; i < count; ++i) {
    const Slot &s = A.slots[i];
    h[i] = {adj_src[i], rel_src[i], (uint16_t *)D.adjT.p + s.adjT_off, (int32_t *)D.flags.p + 2 * i,
            s.n, s.d, s.dp, check_sym[i]};
  }
  int rc = D.prep.ensure((size_t)count * sizeof(PrepDesc));
  if (rc) return rc;
  GG_HIP(hipMemcpyAsync(D.prep.p, h.data(), (size_t)count * sizeof(PrepDesc), hipMemcpyHostToDevice,
                        E.stream));
  GG_HIP(hipStreamSynchronize(E.stream));  // h is a local
  return upload_order(A, D);
}

// Row-order descriptors (gg_order_rows_kernel), uploaded with the prep
// descriptors.  A graph is ordered when its rows are at most four chunks
// and (mode 1) it has enough source tiles for the ordering to pay, or
// (mode 2) always.  The tile BFS's gathers grow with rows x tiles and the
// ordering with rows alone: on C5 (n = 65536, u16) the ordering costs 37 us
// and saves 15 % of a 1.86 ms BFS over 4096 tiles, so it breaks even near
// 540 tiles; mode 1 asks for kOrderMinTiles.  A row table that is ordered
// once and then patched for many evaluations (the incremental scorer) is
// ordered from kOrderMinNodes vertices.  Banks: 8/16/32-bit words are served from 32 banks
// of 4 bytes per half-wave (bank = 4-byte word index mod 32); 64-bit words
// from 32 slots of 8 bytes per half-wave, keyed by the word index mod 32
// (measured: 3.5 cycles per wave-read spread over the 32 slots, 6.7 at
// random), so their key is the id itself, as for 32-bit words.
int g_row_order = 1;
constexpr int kOrderMinNodes = 8192;

int order_shift(int bits) { return bits == 8 ? 0 : bits == 16 ? 1 : 2; }

// `size` is the graph's source-tile count, or its vertex count for a table
// ordered once (min_size says which threshold applies).
bool order_applies_mode(int mode, int size, int min_size, int dp) {
  if (dp > 32 || dp % 8) return false;
  return mode == 2 || (mode == 1 && size >= min_size);
}

bool order_applies(int size, int min_size, int dp) {
  return order_applies_mode(g_row_order, size, min_size, dp);
}

int upload_order(Assembly &A, DevSet &D) {
  const int count = (int)A.slots.size();
  std::vector<OrderDesc> h(count);
  A.order_nch_mask = 0;
  A.order_max_n = 0;
  for (int i = 0; i < count; ++i) {
    const Slot &s = A.slots[i];
    const int shift = order_applies_mode(A.order_mode, A.meta[i].unit_count, kOrderMinTiles, s.dp)
                          ? order_shift(s.bits) : -1;
    h[i] = {(uint16_t *)D.adjT.p + s.adjT_off, nullptr, s.n, s.dp, shift};
    if (shift >= 0) {
      A.order_nch_mask |= 1 << (s.dp / 8);
      A.order_max_n = std::max(A.order_max_n, s.n);
    }
  }
  if (!A.order_nch_mask) return GG_OK;
  int rc = D.order.ensure((size_t)count * sizeof(OrderDesc));
  if (rc) return rc;
  GG_HIP(hipMemcpyAsync(D.order.p, h.data(), (size_t)count * sizeof(OrderDesc), hipMemcpyHostToDevice,
                        E.stream));
  GG_HIP(hipStreamSynchronize(E.stream));  // h is a local
  return GG_OK;
}

void launch_order(const OrderDesc *descs, int count, int max_n, int nch_mask, hipStream_t st) {
  const unsigned gx = (unsigned)((max_n + 64 * kOrderWaves - 1) / (64 * kOrderWaves));
  for (int g0 = 0; g0 < count; g0 += 65535) {  // gridDim.y is at most 65535
    const dim3 grid(gx, std::min(65535, count - g0));
    const dim3 block(64 * kOrderWaves);
    if (nch_mask & 2) hipLaunchKernelGGL(gg_order_rows_kernel<1>, grid, block, 0, st, descs + g0);
    if (nch_mask & 4) hipLaunchKernelGGL(gg_order_rows_kernel<2>, grid, block, 0, st, descs + g0);
    if (nch_mask & 8) hipLaunchKernelGGL(gg_order_rows_kernel<3>, grid, block, 0, st, descs + g0);
    if (nch_mask & 16) hipLaunchKernelGGL(gg_order_rows_kernel<4>, grid, block, 0, st, descs + g0);
  }
}

bool order_rows_applies(int n, int dp) { return order_applies(n, kOrderMinNodes, dp); }

int order_rows_sync(uint16_t *adjR, int n, int dp, int bits, hipStream_t st) {
  if (!order_applies(n, kOrderMinNodes, dp)) return GG_OK;
  const int shift = order_shift(bits);
  DevBuf desc;
  int rc = desc.ensure(sizeof(OrderDesc));
  if (rc) return rc;
  const OrderDesc od = {adjR, nullptr, n, dp, shift};
  if (hipMemcpyAsync(desc.p, &od, sizeof od, hipMemcpyHostToDevice, st) != hipSuccess) {
    desc.release();
    return set_error(GG_ERR_HIP, "row-order upload failed");
  }
  launch_order((const OrderDesc *)desc.p, 1, n, 1 << (dp / 8), st);
  const hipError_t e = hipGetLastError();
  const hipError_t w = hipStreamSynchronize(st);
  desc.release();
  if (e != hipSuccess || w != hipSuccess)
    return set_error(GG_ERR_HIP, "row order failed: %s", hipGetErrorString(e != hipSuccess ? e : w));
  return GG_OK;
}

// Enqueue prep + tile BFS + combine for an assembled set whose adjacency
// already sits in `adj_src[i]` (device pointers) -- graph-capturable (no
// memset nodes, no host sync).
struct PatchSrc {
  const int32_t *base;
  const int32_t *patches;
  int max_patch;
};

int enqueue(Assembly &A, DevSet &D, const int32_t *const *adj_src, const int32_t *const *rel_src,
            const int *check_sym, hipEvent_t ev_begin = nullptr, hipEvent_t ev_end = nullptr,
            const PatchSrc *patched = nullptr) {
  const int count = (int)A.slots.size();
  // The flags and the queue counter are zero on entry (zeroed by
  // prepare_buffers, then by every run's combine), so the sequence needs
  // neither a memset node nor a fill kernel.
  if (patched) {
    // All slots share one shape (checked at plan creation) and one base table.
    const Slot &s = A.slots[0];
    const size_t stride = count > 1 ? A.slots[1].adjT_off - A.slots[0].adjT_off : 0;
    const uint32_t entries = (uint32_t)s.n * (uint32_t)s.dp;
    for (int g0 = 0; g0 < count; g0 += 65535)  // gridDim.y is at most 65535
      hipLaunchKernelGGL(gg_prep_patched_kernel, dim3((entries + 255) / 256, std::min(65535, count - g0)),
                         dim3(256), 0, E.stream, patched->base, rel_src[0], s.n, s.d, s.dp,
                         (uint16_t *)D.adjT.p, stride, (int32_t *)D.flags.p, g0);
    const int np = count * patched->max_patch;
    hipLaunchKernelGGL(gg_patch_kernel, dim3((np + 255) / 256), dim3(256), 0, E.stream,
                       patched->patches, patched->max_patch, count, rel_src[0], s.n, s.d, s.dp,
                       (uint16_t *)D.adjT.p, stride, (int32_t *)D.flags.p);
  } else {
    // The per-graph descriptors were uploaded by upload_prep (outside any
    // capture); one launch covers every graph.
    (void)adj_src;
    (void)rel_src;
    (void)check_sym;
    uint32_t max_chunks = 1;  // one prep thread per 16-byte row chunk
    for (const Slot &s : A.slots) max_chunks = std::max(max_chunks, (uint32_t)s.n * (uint32_t)(s.dp / 8));
    for (int g0 = 0; g0 < count; g0 += 65535)  // gridDim.y is at most 65535
      hipLaunchKernelGGL(gg_prep_kernel, dim3((max_chunks + 255) / 256, std::min(65535, count - g0)),
                         dim3(256), 0, E.stream, (const PrepDesc *)D.prep.p + g0);
  }
  if (A.order_nch_mask)
    launch_order((const OrderDesc *)D.order.p, count, A.order_max_n, A.order_nch_mask, E.stream);
  if (ev_begin) GG_HIP(hipEventRecord(ev_begin, E.stream));
  for (const LaunchGroup &lg : A.groups) {
    if (A.persistent) {
      const QueueKernel qk = queue_kernel_for(lg.key);
      if (!qk) return set_error(GG_ERR_STATE, "no persistent instance bits=%d", lg.key.bits);
      hipLaunchKernelGGL(qk, dim3(lg.grid), dim3(lg.key.block), lg.lds + 16, E.stream,
                         (const GraphDev *)D.gdesc.p, (const UnitDesc *)D.units.p, (UnitOut *)D.outs.p,
                         lg.unit_begin, lg.unit_count, lg.n_cap, lg.qcap,
                         (int *)D.flags.p + 2 * count);
      continue;
    }
    TileKernel k = kernel_for(lg.key);
    if (!k) return set_error(GG_ERR_STATE, "no kernel instance bits=%d block=%d vpt=%d", lg.key.bits,
                             lg.key.block, lg.key.vpt);
    hipLaunchKernelGGL(k, dim3(lg.unit_count), dim3(lg.key.block), lg.lds, E.stream,
                       (const GraphDev *)D.gdesc.p, (const UnitDesc *)D.units.p, (UnitOut *)D.outs.p,
                       lg.unit_begin, lg.n_cap, lg.qcap);
  }
  if (ev_end) GG_HIP(hipEventRecord(ev_end, E.stream));
  // Results go straight to the pinned host buffer.
  hipLaunchKernelGGL(gg_combine_kernel, dim3(count), dim3(256), 0, E.stream,
                     (const GraphMeta *)D.meta.p, (const UnitDesc *)D.units.p,
                     (const UnitOut *)D.outs.p, (GraphOut *)D.res_h.p, count, (int32_t *)D.flags.p);
  GG_HIP(hipGetLastError());
  return GG_OK;
}

int collect(Assembly &A, DevSet &D, gg_result *out) {
  GG_HIP(hipStreamSynchronize(E.stream));
  const GraphOut *r = (const GraphOut *)D.res_h.p;
  int rc = GG_OK;
  for (size_t i = 0; i < A.slots.size(); ++i) {
    out[i].connected = r[i].connected;
    out[i].diameter = r[i].diameter;
    out[i].max_ecc = r[i].max_ecc;
    out[i].word_bits = A.slots[i].bits;
    out[i].sum = r[i].sum;
    if (r[i].error & 2)
      rc = set_error(GG_ERR_STATE, "graph %zu: the reach array is not at LDS address 0", i);
    else if (r[i].error)
      rc = set_error(GG_ERR_INVALID, "graph %zu: adjacency entry outside [0,%d)", i, A.slots[i].n);
  }
  return rc;
}

static DevSet g_single;

// A host relabel table must be a permutation of 0..n-1: every row of the
// transposed table is then written exactly once.
int check_permutation(const int32_t *p, int n) {
  std::vector<unsigned char> seen(n, 0);
  for (int i = 0; i < n; ++i) {
    if (p[i] < 0 || p[i] >= n || seen[p[i]])
      return set_error(GG_ERR_INVALID, "relabel is not a permutation (entry %d = %d)", i, p[i]);
    seen[p[i]] = 1;
  }
  return GG_OK;
}

static int g_trace_units = 0;      // >0: next eval_host records per-level traces
static uint64_t *g_trace_host = nullptr;

// One graph from host memory (the reference ABI's call, repeated with the
// same shape while the caller changes the table): the assembly, the device
// descriptors and a captured hipGraph of prep + row order + tile BFS +
// combine (results written to pinned host memory) are kept for the last
// shape, so a repeated call is
// one adjacency upload, one graph launch and one wait.
struct SingleCache {
  bool valid = false;
  int n = 0, d = 0, b = 0, g = 0, c = 0, flags = 0, has_rel = 0, bits = 0, order = 0, block = 0;
  Assembly A;
  DevSet D;
  std::vector<const int32_t *> src, rel;
  std::vector<int> sym;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  PinBuf stage;  // pinned staging of the table's upload
  void drop() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    valid = false;
  }
  void release() {
    drop();
    D.release();
    stage.release();
  }
};
static SingleCache g_cache;
constexpr size_t kStageChunks = 4;

int eval_single_cached(const gg_graph &g, gg_result *out) {
  SingleCache &C = g_cache;
  const Config cfg = config_for(g.nodes, g.degree, g.based_nodes, g.groups, g.added_centers, false);
  const int bits = cfg.bits, order = effective_order(cfg);
  const int has_rel = g.relabel != nullptr;
  const bool same = C.valid && C.n == g.nodes && C.d == g.degree && C.b == g.based_nodes &&
                    C.g == g.groups && C.c == g.added_centers && C.flags == g.flags &&
                    C.has_rel == has_rel && C.bits == bits && C.order == order &&
                    C.block == cfg.block;
  int rc;
  if (!same) {
    C.drop();
    if ((rc = assemble(1, &g, 0, C.A)) || (rc = prepare_buffers(C.A, C.D, 0, nullptr))) return rc;
    C.src.assign(1, (const int32_t *)C.D.adj_in.p);
    C.rel.assign(1, has_rel ? (const int32_t *)C.D.relabel_in.p : nullptr);
    C.sym.assign(1, (g.flags & GG_GRAPH_SYMMETRIC) ? 0 : 1);
    if ((rc = upload_prep(C.A, C.D, C.src.data(), C.rel.data(), C.sym.data()))) return rc;
    GG_HIP(hipStreamSynchronize(E.stream));
    GG_HIP(hipStreamBeginCapture(E.stream, hipStreamCaptureModeThreadLocal));
    rc = enqueue(C.A, C.D, C.src.data(), C.rel.data(), C.sym.data());
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(E.stream, &graph);
    if (rc || e != hipSuccess || !graph ||
        hipGraphInstantiate(&C.exec, graph, nullptr, nullptr, 0) != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      C.exec = nullptr;
      return rc ? rc : set_error(GG_ERR_HIP, "single-graph capture failed: %s", hipGetErrorString(e));
    }
    C.graph = graph;
    C.n = g.nodes;
    C.d = g.degree;
    C.b = g.based_nodes;
    C.g = g.groups;
    C.c = g.added_centers;
    C.flags = g.flags;
    C.has_rel = has_rel;
    C.bits = bits;
    C.order = order;
    C.block = cfg.block;
    C.valid = true;
  }
  const Slot &s = C.A.slots[0];
  // The table goes through pinned staging in chunks of at least 1 MiB: the
  // host copies chunk k+1 while chunk k's DMA runs (the previous call's DMAs
  // completed before it returned, so the staging is free).  Smaller tables
  // go in one chunk: each DMA call costs more than the overlap saves there
  // (n = 4096, d = 16: 100 us per call in 64 KiB chunks, 73 us in one).
  const size_t bytes = (size_t)s.n * s.d * 4, rbytes = has_rel ? (size_t)s.n * 4 : 0;
  if ((rc = C.stage.ensure(bytes + rbytes))) return rc;
  const size_t chunk = std::max<size_t>((bytes + kStageChunks - 1) / kStageChunks, 1024 * 1024);
  for (size_t o = 0; o < bytes; o += chunk) {
    const size_t len = std::min(chunk, bytes - o);
    memcpy((char *)C.stage.p + o, (const char *)g.adj + o, len);
    GG_HIP(hipMemcpyAsync((char *)C.D.adj_in.p + o, (const char *)C.stage.p + o, len,
                          hipMemcpyHostToDevice, E.stream));
  }
  if (has_rel) {
    memcpy((char *)C.stage.p + bytes, g.relabel, rbytes);
    GG_HIP(hipMemcpyAsync(C.D.relabel_in.p, (const char *)C.stage.p + bytes, rbytes,
                          hipMemcpyHostToDevice, E.stream));
  }
  GG_HIP(hipGraphLaunch(C.exec, E.stream));
  return collect(C.A, C.D, out);
}

int eval_host(int count, const gg_graph *gs, gg_result *out, bool device_ptrs, int hist_maxd,
              int32_t *hist_host) {
  // Argument checks first: they need no device.
  int rc;
  for (int i = 0; i < count; ++i) {
    if ((rc = validate(gs[i]))) return rc;
    if (!device_ptrs && gs[i].relabel && (rc = check_permutation(gs[i].relabel, gs[i].nodes)))
      return rc;
  }
  if ((rc = init(-1))) return rc;
  if (count == 1 && !device_ptrs && hist_maxd == 0 && g_trace_units == 0)
    return eval_single_cached(gs[0], out);
  if (device_ptrs) {
    // A device relabel table is checked on the host too (n int32 read back):
    // with duplicates some rows of the converted table would not be written.
    for (int i = 0; i < count; ++i) {
      if (!gs[i].relabel) continue;
      std::vector<int32_t> h((size_t)gs[i].nodes);
      GG_HIP(hipMemcpy(h.data(), gs[i].relabel, h.size() * 4, hipMemcpyDeviceToHost));
      if ((rc = check_permutation(h.data(), gs[i].nodes))) return rc;
    }
  }
  Assembly A;
  if ((rc = assemble(count, gs, hist_maxd, A))) return rc;
  DevSet &D = g_single;
  int32_t *hist_dev = nullptr;
  size_t hist_bytes = 0;
  if (hist_maxd > 0) {
    hist_bytes = (size_t)gs[0].nodes * (hist_maxd + 1) * 4;
    if ((rc = D.hist.ensure(hist_bytes))) return rc;
    hist_dev = (int32_t *)D.hist.p;
    GG_HIP(hipMemsetAsync(hist_dev, 0, hist_bytes, E.stream));
  }
  DevBuf trace_buf;
  const int trace_units = std::min<int>(g_trace_units, (int)A.units.size());
  if (trace_units > 0) {
    size_t tb = (size_t)A.units.size() * kTraceLevels * kTraceFields * sizeof(uint64_t);
    if ((rc = trace_buf.ensure(tb))) return rc;
    GG_HIP(hipMemsetAsync(trace_buf.p, 0, tb, E.stream));
  }
  if ((rc = prepare_buffers(A, D, hist_maxd, hist_dev, (uint64_t *)trace_buf.p, trace_units)))
    return rc;
  std::vector<const int32_t *> src(count), rel(count);
  std::vector<int> sym(count);
  for (int i = 0; i < count; ++i) {
    const Slot &s = A.slots[i];
    sym[i] = (gs[i].flags & GG_GRAPH_SYMMETRIC) ? 0 : 1;
    if (device_ptrs) {
      src[i] = gs[i].adj;
      rel[i] = gs[i].relabel;
    } else {
      int32_t *dst = (int32_t *)D.adj_in.p + s.adj_off;
      GG_HIP(hipMemcpyAsync(dst, gs[i].adj, (size_t)s.n * s.d * 4, hipMemcpyHostToDevice, E.stream));
      src[i] = dst;
      rel[i] = nullptr;
      if (s.has_rel) {
        int32_t *rd = (int32_t *)D.relabel_in.p + s.rel_off;
        GG_HIP(hipMemcpyAsync(rd, gs[i].relabel, (size_t)s.n * 4, hipMemcpyHostToDevice, E.stream));
        rel[i] = rd;
      }
    }
  }
  if ((rc = upload_prep(A, D, src.data(), rel.data(), sym.data()))) return rc;
  if ((rc = enqueue(A, D, src.data(), rel.data(), sym.data()))) return rc;
  if (hist_maxd > 0)
    GG_HIP(hipMemcpyAsync(hist_host, hist_dev, hist_bytes, hipMemcpyDeviceToHost, E.stream));
  if (trace_units > 0)
    GG_HIP(hipMemcpyAsync(g_trace_host, trace_buf.p,
                          (size_t)trace_units * kTraceLevels * kTraceFields * sizeof(uint64_t),
                          hipMemcpyDeviceToHost, E.stream));
  rc = collect(A, D, out);
  trace_buf.release();
  return rc;
}

// ---- The allocation-free evaluation ----------------------------------------
// gg_apsp_device's kernels on one graph whose buffers all live in this
// device's `scratch` array (the descriptors are written by a setup kernel
// from its arguments), so nothing is allocated: no device memory, no pinned
// or heap host memory.  Calls on one device share the array, so calls that
// overlap on one device race; every call rewrites all it reads.
static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

int eval_noalloc(const gg_graph &g, gg_result *out) {
  int rc = validate(g);
  if (rc) return rc;
  if (g.nodes > kNoAllocMaxNodes || g.degree > kNoAllocMaxDegree)
    return set_error(GG_ERR_TOO_LARGE,
                     "allocation-free evaluation: n=%d d=%d beyond its scratch (n <= %d, d <= %d)",
                     g.nodes, g.degree, kNoAllocMaxNodes, kNoAllocMaxDegree);
  if (g.flags & GG_GRAPH_BFS)
    return set_error(GG_ERR_INVALID, "allocation-free evaluation: GG_GRAPH_BFS is not served");
  if ((rc = init(-1))) return rc;
  unsigned char *base = E.device < kMaxDevices ? E.scratch_dev[E.device] : nullptr;
  if (!base) return set_error(GG_ERR_STATE, "scratch array address unknown on device %d", E.device);
  const int n = g.nodes, d = g.degree, dp = (d + 7) & ~7;
  const int B = g.based_nodes, c = g.added_centers;
  const Config cfg = config_for(n, d, B, g.groups, c, false);
  const int bits = cfg.bits;
  if (bits == 0) return set_error(GG_ERR_TOO_LARGE, "n=%d does not fit LDS", n);
  int block, vpt;
  block_vpt_pref(n, cfg.block, &block, &vpt);
  const TileKernel k = kernel_for({bits, block, vpt, 0});
  if (!k) return set_error(GG_ERR_TOO_LARGE, "no %d-bit instance for n=%d", bits, n);
  const int n_cap = (n + 7) & ~7;
  const size_t rbytes = (size_t)n_cap * (bits / 8);
  const size_t budget = rbytes * 2 + 4096 <= (size_t)kLdsBytes / 2 ? rbytes * 2 + 4096 : (size_t)kLdsBytes;
  const int qcap = qcap_for(bits, n_cap, budget);
  if (qcap < 64) return set_error(GG_ERR_TOO_LARGE, "LDS budget too small for n = %d", n);
  const size_t lds = lds_bytes_for(bits, n_cap, qcap);
  const int units = (B + bits - 1) / bits + (c + bits - 1) / bits;
  if (units > kNoAllocMaxUnits) return set_error(GG_ERR_TOO_LARGE, "%d source tiles", units);
  // Layout of the scratch array.
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off = align256(off + bytes);
    return base + at;
  };
  uint16_t *adjR = (uint16_t *)take(kNoAllocAdjBytes);
  int32_t *flags = (int32_t *)take(2 * sizeof(int32_t));
  GraphDev *gd_dev = (GraphDev *)take(sizeof(GraphDev));
  PrepDesc *pd_dev = (PrepDesc *)take(sizeof(PrepDesc));
  OrderDesc *od_dev = (OrderDesc *)take(sizeof(OrderDesc));
  GraphMeta *gm_dev = (GraphMeta *)take(sizeof(GraphMeta));
  GraphOut *gout_dev = (GraphOut *)take(sizeof(GraphOut));
  UnitDesc *units_dev = (UnitDesc *)take((size_t)kNoAllocMaxUnits * sizeof(UnitDesc));
  UnitOut *outs_dev = (UnitOut *)take((size_t)kNoAllocMaxUnits * sizeof(UnitOut));
  const int bm_words = (kNoAllocMaxNodes + 31) / 32;
  uint32_t *bitmap = (uint32_t *)take((size_t)bm_words * 4);
  if (off > kNoAllocScratchBytes) return set_error(GG_ERR_STATE, "scratch layout exceeds the array");

  GraphDev gd;
  gd.adjR = adjR;
  gd.flags = flags;
  gd.n = n;
  gd.dp = dp;
  gd.split = B * g.groups;
  gd.c = c;
  gd.push_ok = 1;
  gd.groups = g.groups;
  gd.hist_maxd = 0;
  gd.trace_units = 0;
  gd.hist = nullptr;
  gd.trace = nullptr;
  const PrepDesc pd = {g.adj, g.relabel, adjR, flags, n, d, dp, (g.flags & GG_GRAPH_SYMMETRIC) ? 0 : 1};
  const int shift =
      order_applies_mode(effective_order(cfg), units, kOrderMinTiles, dp) ? order_shift(bits) : -1;
  const OrderDesc od = {adjR, nullptr, n, dp, shift};
  const GraphMeta gm = {0, units, n, g.groups};
  const hipStream_t st = E.stream;
  const int setup_items = std::max(units, bm_words);
  hipLaunchKernelGGL(gg_noalloc_setup_kernel, dim3((setup_items + 255) / 256), dim3(256), 0, st, gd, pd,
                     od, gm, gd_dev, pd_dev, od_dev, gm_dev, units_dev, bits, B, c, n, flags, bitmap,
                     bm_words);
  if (g.relabel)
    hipLaunchKernelGGL(gg_relabel_check_kernel, dim3((n + 255) / 256), dim3(256), 0, st, g.relabel, n,
                       bitmap, flags);
  const uint32_t chunks = (uint32_t)n * (uint32_t)(dp / 8);
  hipLaunchKernelGGL(gg_prep_kernel, dim3((chunks + 255) / 256, 1), dim3(256), 0, st,
                     (const PrepDesc *)pd_dev);
  if (shift >= 0) launch_order(od_dev, 1, n, 1 << (dp / 8), st);
  hipLaunchKernelGGL(k, dim3(units), dim3(block), lds, st, (const GraphDev *)gd_dev,
                     (const UnitDesc *)units_dev, outs_dev, 0, n_cap, qcap);
  hipLaunchKernelGGL(gg_combine_kernel, dim3(1), dim3(256), 0, st, (const GraphMeta *)gm_dev,
                     (const UnitDesc *)units_dev, (const UnitOut *)outs_dev, gout_dev, 1, nullptr);
  GG_HIP(hipGetLastError());
  GraphOut r;
  GG_HIP(hipMemcpyAsync(&r, gout_dev, sizeof r, hipMemcpyDeviceToHost, st));
  GG_HIP(hipStreamSynchronize(st));
  if (r.error & 2) return set_error(GG_ERR_STATE, "the reach array is not at LDS address 0");
  if (r.error)
    return set_error(GG_ERR_INVALID, "adjacency or relabel entry outside [0,%d), or relabel not a permutation", n);
  out->connected = r.connected;
  out->diameter = r.diameter;
  out->max_ecc = r.max_ecc;
  out->word_bits = bits;
  out->sum = r.sum;
  return GG_OK;
}

// ---- Plans ---------------------------------------------------------------

struct Plan {
  Assembly A;
  DevSet D;
  // Patched plans: every slot is the base graph plus up to max_patch entries.
  int max_patch = 0;
  DevBuf base, patches;
  PatchSrc psrc{nullptr, nullptr, 0};
  std::vector<const int32_t *> src, rel;
  std::vector<int> sym;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t tev_begin = nullptr, tev_end = nullptr;  // around the tile-BFS launches (eager, timed)
  bool launched = false, timed = false;
};

}  // namespace gg

using namespace gg;

struct gg_plan : Plan {};

extern "C" {

int gg_init(int device) { return gg::init(device); }
void gg_finalize(void) {
  g_single.release();
  g_cache.release();
  gg::wide_release();
  gg::finalize();
}
const char *gg_last_error(void) { return gg::last_error(); }
int gg_last_error_code(void) { return gg::last_error_code(); }
int gg_device_count(void) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) return 0;
  return c;
}

// Graphs above the tile kernels' limit, and the bfs() semantics, go to the
// HBM-resident kernels (gg_wide.hip).
static int apsp_any(const gg_graph *g, gg_result *out, bool device_ptrs) {
  if (!g || !out || !g->adj) return set_error(GG_ERR_INVALID, "null argument");
  const bool bfs = (g->flags & GG_GRAPH_BFS) != 0;
  if (!bfs && g->nodes <= kMaxNodes) return eval_host(1, g, out, device_ptrs, 0, nullptr);
  int rc = validate(*g, true);
  if (rc) return rc;
  if (bfs && g->relabel) return set_error(GG_ERR_INVALID, "GG_GRAPH_BFS takes no relabel");
  if (!device_ptrs && g->relabel && (rc = check_permutation(g->relabel, g->nodes))) return rc;
  if (device_ptrs && g->relabel) {
    if ((rc = init(-1))) return rc;
    std::vector<int32_t> h((size_t)g->nodes);
    GG_HIP(hipMemcpy(h.data(), g->relabel, h.size() * 4, hipMemcpyDeviceToHost));
    if ((rc = check_permutation(h.data(), g->nodes))) return rc;
  }
  return wide_eval(*g, bfs ? 1 : 0, device_ptrs, out);
}

int gg_apsp(const gg_graph *g, gg_result *out) { return apsp_any(g, out, false); }

int gg_apsp_device_noalloc(const gg_graph *g, gg_result *out) {
  if (!g || !out || !g->adj) return set_error(GG_ERR_INVALID, "null argument");
  return eval_noalloc(*g, out);
}

int gg_apsp_device(const gg_graph *g, gg_result *out) { return apsp_any(g, out, true); }

int gg_eval_batch(int count, const gg_graph *graphs, gg_result *out) {
  if (count <= 0 || !graphs || !out) return set_error(GG_ERR_INVALID, "bad batch arguments");
  for (int i = 0; i < count; ++i)
    if (!graphs[i].adj) return set_error(GG_ERR_INVALID, "graph %d has no adjacency", i);
  return eval_host(count, graphs, out, false, 0, nullptr);
}

int gg_distance_histogram(int nodes, int degree, const int32_t *adj, int maxd, int32_t *hist) {
  if (!adj || !hist || maxd < 1) return set_error(GG_ERR_INVALID, "bad histogram arguments");
  gg_graph g = {nodes, degree, nodes, 1, 0, 0, adj, nullptr};
  gg_result r;
  return eval_host(1, &g, &r, false, maxd, hist);
}

int gg_eval_features(int nodes, int degree, const int32_t *adj, int maxd, int32_t *ecc,
                     int32_t *hist, gg_result *out) {
  if (!adj || maxd < 1 || nodes < 2) return set_error(GG_ERR_INVALID, "bad feature arguments");
  std::vector<int32_t> own;
  int32_t *h = hist;
  if (!h) {
    own.assign((size_t)nodes * (size_t)(maxd + 1), 0);
    h = own.data();
  }
  gg_graph g = {nodes, degree, nodes, 1, 0, 0, adj, nullptr};
  gg_result r;
  const int rc = eval_host(1, &g, &r, false, maxd, h);
  if (rc) return rc;
  if (out) *out = r;
  if (ecc) {
    for (int v = 0; v < nodes; ++v) {
      const int32_t *row = h + (size_t)v * (size_t)(maxd + 1);
      int64_t reached = 0;
      int far = 0;
      for (int k = 0; k <= maxd; ++k) {
        reached += row[k];
        if (row[k] > 0) far = k;
      }
      ecc[v] = reached < nodes ? -1 : far;  // far == maxd: at least maxd
    }
  }
  return GG_OK;
}

int gg_trace_levels(const gg_graph *g, int units, uint64_t *trace, gg_result *out) {
  if (!g || !trace || !out || units < 1) return set_error(GG_ERR_INVALID, "bad trace arguments");
  g_trace_units = units;
  g_trace_host = trace;
  int rc = eval_host(1, g, out, false, 0, nullptr);
  g_trace_units = 0;
  g_trace_host = nullptr;
  return rc;
}

int gg_set_row_order(int mode) {
  if (mode < 0 || mode > 2) return set_error(GG_ERR_INVALID, "row-order mode must be 0, 1 or 2");
  const int prev = g_row_order;
  g_row_order = mode;
  return prev;
}

int gg_row_order(int nodes, int degree, const int32_t *adj, int word_bits, int32_t *rows,
                 uint64_t *cost) {
  if (nodes < 2 || nodes > 65536 || degree < 1 || !adj || !rows || !cost)
    return set_error(GG_ERR_INVALID, "bad row-order arguments");
  if (word_bits != 8 && word_bits != 16 && word_bits != 32 && word_bits != 64)
    return set_error(GG_ERR_INVALID, "word bits must be 8/16/32/64");
  if (degree > 32) return set_error(GG_ERR_INVALID, "rows of more than 32 entries are not ordered");
  int rc = gg::init(-1);
  if (rc) return rc;
  const int dp = (degree + 7) / 8 * 8;
  const size_t entries = (size_t)nodes * dp;
  DevBuf adj_d, rows_d, flags_d, desc_d, cost_d;
  auto done = [&](int r) {
    for (DevBuf *b : {&adj_d, &rows_d, &flags_d, &desc_d, &cost_d}) b->release();
    return r;
  };
  if ((rc = adj_d.ensure((size_t)nodes * degree * 4)) || (rc = rows_d.ensure(entries * 2 + 64)) ||
      (rc = flags_d.ensure(8)) || (rc = desc_d.ensure(sizeof(PrepDesc) + sizeof(OrderDesc))) ||
      (rc = cost_d.ensure(16)))
    return done(rc);
  hipStream_t st = engine().stream;
  const int shift = order_shift(word_bits);
  const PrepDesc pd = {(const int32_t *)adj_d.p, nullptr, (uint16_t *)rows_d.p, (int32_t *)flags_d.p,
                       nodes, degree, dp, 0};
  const OrderDesc od = {(uint16_t *)rows_d.p, (unsigned long long *)cost_d.p, nodes, dp, shift};
  const uint64_t zero[2] = {0, 0};
  std::vector<uint16_t> h(entries);
  int32_t flags[2] = {0, 0};
  if (hipMemcpyAsync(adj_d.p, adj, (size_t)nodes * degree * 4, hipMemcpyHostToDevice, st) ||
      hipMemcpyAsync(flags_d.p, zero, 8, hipMemcpyHostToDevice, st) ||
      hipMemcpyAsync(cost_d.p, zero, 16, hipMemcpyHostToDevice, st) ||
      hipMemcpyAsync(desc_d.p, &pd, sizeof pd, hipMemcpyHostToDevice, st) ||
      hipMemcpyAsync((char *)desc_d.p + sizeof(PrepDesc), &od, sizeof od, hipMemcpyHostToDevice, st))
    return done(set_error(GG_ERR_HIP, "row-order upload failed"));
  hipLaunchKernelGGL(gg_prep_kernel, dim3((unsigned)((entries / 8 + 255) / 256), 1), dim3(256), 0, st,
                     (const PrepDesc *)desc_d.p);
  launch_order((const OrderDesc *)((const char *)desc_d.p + sizeof(PrepDesc)), 1, nodes,
               1 << (dp / 8), st);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(h.data(), rows_d.p, entries * 2, hipMemcpyDeviceToHost, st) ||
      hipMemcpyAsync(flags, flags_d.p, 8, hipMemcpyDeviceToHost, st) ||
      hipMemcpyAsync(cost, cost_d.p, 16, hipMemcpyDeviceToHost, st) || hipStreamSynchronize(st))
    return done(set_error(GG_ERR_HIP, "row-order launch failed"));
  if (flags[0]) return done(set_error(GG_ERR_INVALID, "adjacency has %d out-of-range entries", flags[0]));
  for (size_t k = 0; k < entries; ++k) rows[k] = h[k];
  return done(GG_OK);
}

int gg_set_word_bits(int bits) {
  if (bits != 0 && bits != 8 && bits != 16 && bits != 32 && bits != 64)
    return set_error(GG_ERR_INVALID, "word bits must be 0/8/16/32/64");
  engine().forced_bits = bits;
  return GG_OK;
}

static gg_plan *plan_create(int count, const gg_graph *shapes, int max_patch);

// Times every candidate configuration of `count` graphs (host adjacency as
// given) through a plan and keeps the fastest in the tuning table under the
// shape bucket of the largest graph.  Candidates: every word width that
// fits, 256- or 1024-thread blocks (up to 2048 vertices), row order off or
// on, and one queue-driven launch or one launch with a workgroup per unit
// (per kernel instance, for batches).  Every candidate must return the same (connected, diameter, S)
// for every graph, else the tuner fails with GG_ERR_STATE.
static int tune_run(int count, const gg_graph *graphs, int reps, Config *best_cfg, double *best_ms) {
  if (count < 1 || !graphs || reps < 1) return set_error(GG_ERR_INVALID, "bad tuning arguments");
  int rc;
  if ((rc = init(-1))) return rc;
  int big = 0;
  for (int i = 0; i < count; ++i) {
    if ((rc = validate(graphs[i]))) return rc;
    if (!graphs[i].adj) return set_error(GG_ERR_INVALID, "graph %d has no adjacency", i);
    if (graphs[i].nodes > graphs[big].nodes) big = i;
  }
  const bool batched = count > 1;
  const int n_max = graphs[big].nodes;
  const int maxb = max_bits_for(n_max);
  if (maxb == 0) return set_error(GG_ERR_TOO_LARGE, "n=%d does not fit LDS", n_max);
  std::vector<Config> cands;
  for (int bits : {8, 16, 32, 64}) {
    if (bits > maxb) continue;
    for (int block : {0, 1024}) {
      if (block == 1024 && n_max > 2048) continue;  // 1024 is the only choice there
      for (int order : {0, 2})
        for (int pers : {1, 0}) cands.push_back({bits, block, order, pers});
    }
  }
  const bool saved_on = E.forced_cfg;
  const Config saved = E.forced;
  const int saved_bits = E.forced_bits;
  E.forced_bits = 0;
  std::vector<gg_result> ref(count), res(count);
  bool have_ref = false;
  double best_t = 1e30;
  Config best{0, 0, -1, 0};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  rc = GG_OK;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    rc = set_error(GG_ERR_HIP, "event creation failed");
  for (size_t k = 0; k < cands.size() && rc == GG_OK; ++k) {
    E.forced_cfg = true;
    E.forced = cands[k];
    gg_plan *p = plan_create(count, graphs, 0);
    E.forced_cfg = saved_on;
    E.forced = saved;
    if (!p) {
      rc = last_error_code() ? last_error_code() : GG_ERR_STATE;
      break;
    }
    for (int i = 0; i < count && rc == GG_OK; ++i) rc = gg_plan_set_adj(p, i, graphs[i].adj);
    if (rc == GG_OK) rc = gg_plan_run(p, res.data());  // warm-up
    if (rc == GG_OK) {
      if (hipEventRecord(e0, E.stream) != hipSuccess) rc = set_error(GG_ERR_HIP, "tuner event record failed");
      for (int r = 0; r < reps && rc == GG_OK; ++r) rc = gg_plan_launch(p);
      if (rc == GG_OK && hipEventRecord(e1, E.stream) != hipSuccess)
        rc = set_error(GG_ERR_HIP, "tuner event record failed");
      if (rc == GG_OK) rc = gg_plan_wait(p, res.data());
    }
    float ms = 0;
    if (rc == GG_OK && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
      rc = set_error(GG_ERR_HIP, "timing failed");
    gg_plan_destroy(p);
    if (rc) break;
    for (int i = 0; i < count; ++i) {
      if (!have_ref) continue;
      if (res[i].connected != ref[i].connected || res[i].diameter != ref[i].diameter ||
          (res[i].connected && res[i].sum != ref[i].sum))
        rc = set_error(GG_ERR_STATE, "tuning candidate %d-bit/%d/%d/%d disagrees on graph %d", cands[k].bits,
                       cands[k].block, cands[k].order, cands[k].persistent, i);
    }
    if (!have_ref) {
      ref = res;
      have_ref = true;
    }
    const double t = ms / reps;
    if (getenv("GG_TUNE_VERBOSE"))
      fprintf(stderr, "gg tune: %d graphs, bits %d block %d order %d persistent %d: %.4f ms\n", count,
              cands[k].bits, cands[k].block, cands[k].order, cands[k].persistent, t);
    if (rc == GG_OK && t < best_t) {
      best_t = t;
      best = cands[k];
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  E.forced_bits = saved_bits;
  if (rc) return rc;
  const gg_graph &g = graphs[big];
  E.tune[{bucket_of(g.nodes), g.degree, g.groups, g.added_centers > 0 ? 1 : 0, batched ? 1 : 0}] = best;
  if (best_cfg) *best_cfg = best;
  if (best_ms) *best_ms = best_t;
  return GG_OK;
}

// A seeded graph of the given shape (`batch` copies of it) through tune_run.
int gg_autotune(int nodes, int degree, int based_nodes, int groups, int batch, int reps,
                int *best_bits, double *best_ms) {
  if (nodes < 2 || degree < 1 || batch < 1 || reps < 1 || groups < 1 || based_nodes < 1)
    return set_error(GG_ERR_INVALID, "bad autotune arguments");
  const int c = nodes - based_nodes * groups;
  std::vector<int32_t> edges((size_t)nodes * degree);
  int rc = groups > 1 && c == 0 ? gg_random_symmetric(nodes, degree, groups, 12345, edges.data())
                                : gg_random_regular(nodes, degree, 12345, edges.data());
  if (rc) return rc;
  std::vector<int32_t> adj((size_t)nodes * degree);
  if ((rc = gg_edges_to_adj(nodes, nodes * degree / 2, degree, edges.data(), adj.data()))) return rc;
  gg_graph g = {nodes, degree, based_nodes, groups, c, GG_GRAPH_SYMMETRIC, adj.data(), nullptr};
  std::vector<gg_graph> shapes(batch, g);
  Config best;
  if ((rc = tune_run(batch, shapes.data(), reps, &best, best_ms))) return rc;
  if (best_bits) *best_bits = best.bits;
  return GG_OK;
}

int gg_autotune_batch(int count, const gg_graph *graphs, int reps, int32_t *config, double *best_ms) {
  Config best;
  const int rc = tune_run(count, graphs, reps, &best, best_ms);
  if (rc) return rc;
  if (config) {
    config[0] = best.bits;
    config[1] = best.block;
    config[2] = best.order;
    config[3] = best.persistent;
  }
  return GG_OK;
}

int gg_tune_lookup(int nodes, int degree, int based_nodes, int groups, int added_centers, int batched,
                   int32_t *config) {
  if (!config || nodes < 2 || degree < 1 || groups < 1 || based_nodes < 1 || added_centers < 0)
    return set_error(GG_ERR_INVALID, "bad lookup arguments");
  const Config c = config_for(nodes, degree, based_nodes, groups, added_centers, batched != 0);
  if (c.bits == 0) return set_error(GG_ERR_TOO_LARGE, "n=%d does not fit LDS", nodes);
  int block, vpt;
  block_vpt_pref(nodes, c.block, &block, &vpt);
  config[0] = c.bits;
  config[1] = block;
  config[2] = effective_order(c);
  config[3] = c.persistent;
  return GG_OK;
}

int gg_tune_table_clear(void) {
  engine().tune.clear();
  return GG_OK;
}

int gg_tune_table_save(const char *path) {
  FILE *f = fopen(path, "w");
  if (!f) return set_error(GG_ERR_INVALID, "cannot write %s", path);
  fprintf(f, "# n_bucket degree groups centres batched word_bits block row_order persistent\n");
  for (auto &kv : engine().tune)
    fprintf(f, "%d %d %d %d %d %d %d %d %d\n", kv.first.nb, kv.first.d, kv.first.g, kv.first.c,
            kv.first.batched, kv.second.bits, kv.second.block, kv.second.order, kv.second.persistent);
  fclose(f);
  return GG_OK;
}

int gg_tune_table_load(const char *path) {
  FILE *f = fopen(path, "r");
  if (!f) return set_error(GG_ERR_INVALID, "cannot read %s", path);
  char line[256];
  int loaded = 0;
  while (fgets(line, sizeof line, f)) {
    if (line[0] == '#') continue;
    int nb, d, g, c, bt, bits, block, order, pers;
    if (sscanf(line, "%d %d %d %d %d %d %d %d %d", &nb, &d, &g, &c, &bt, &bits, &block, &order, &pers) == 9 &&
        (bits == 8 || bits == 16 || bits == 32 || bits == 64) && (block == 0 || block == 256 || block == 1024) &&
        order >= -1 && order <= 2 && (pers == 0 || pers == 1)) {
      engine().tune[{nb, d, g, c ? 1 : 0, bt ? 1 : 0}] = Config{bits, block, order, pers};
      ++loaded;
    }
  }
  fclose(f);
  return loaded;
}

static gg_plan *plan_fail(gg_plan *p) {
  for (hipEvent_t ev : {p->tev_begin, p->tev_end})
    if (ev) (void)hipEventDestroy(ev);
  p->D.release();
  p->base.release();
  p->patches.release();
  delete p;
  return nullptr;
}

static gg_plan *plan_create(int count, const gg_graph *shapes, int max_patch) {
  if (gg::init(-1)) return nullptr;
  if (count <= 0 || !shapes || max_patch < 0 || max_patch > 4096) {
    set_error(GG_ERR_INVALID, "bad plan arguments");
    return nullptr;
  }
  for (int i = 0; i < count; ++i)
    if (shapes[i].relabel && check_permutation(shapes[i].relabel, shapes[i].nodes)) return nullptr;
  gg_plan *p = new gg_plan();
  p->max_patch = max_patch;
  if (assemble(count, shapes, 0, p->A) || prepare_buffers(p->A, p->D, 0, nullptr)) return plan_fail(p);
  p->src.resize(count);
  p->rel.assign(count, nullptr);
  p->sym.resize(count);
  for (int i = 0; i < count; ++i) {
    p->src[i] = (const int32_t *)p->D.adj_in.p + p->A.slots[i].adj_off;
    p->sym[i] = (shapes[i].flags & GG_GRAPH_SYMMETRIC) ? 0 : 1;
    if (shapes[i].relabel) {
      int32_t *rd = (int32_t *)p->D.relabel_in.p + p->A.slots[i].rel_off;
      if (hipMemcpyAsync(rd, shapes[i].relabel, (size_t)shapes[i].nodes * 4, hipMemcpyHostToDevice,
                         engine().stream) != hipSuccess) {
        set_error(GG_ERR_HIP, "relabel upload failed");
        return plan_fail(p);
      }
      p->rel[i] = rd;
    }
  }
  if (max_patch > 0) {
    const gg_graph &g = shapes[0];
    std::vector<int32_t> none((size_t)count * max_patch * 2, -1);
    if (p->base.ensure((size_t)g.nodes * g.degree * 4) ||
        p->patches.ensure(none.size() * 4) ||
        hipMemcpyAsync(p->patches.p, none.data(), none.size() * 4, hipMemcpyHostToDevice,
                       engine().stream) != hipSuccess)
      return plan_fail(p);
    if (hipStreamSynchronize(engine().stream) != hipSuccess) return plan_fail(p);
    p->psrc = {(const int32_t *)p->base.p, (const int32_t *)p->patches.p, max_patch};
  }
  if (hipEventCreate(&p->tev_begin) != hipSuccess || hipEventCreate(&p->tev_end) != hipSuccess) {
    set_error(GG_ERR_HIP, "event creation failed");
    return plan_fail(p);
  }
  // Capture prep + BFS + combine once; replays reuse it.
  if (max_patch == 0 && upload_prep(p->A, p->D, p->src.data(), p->rel.data(), p->sym.data()))
    return plan_fail(p);
  if (max_patch > 0 && upload_order(p->A, p->D)) return plan_fail(p);
  hipStream_t s = engine().stream;
  if (hipStreamSynchronize(s) != hipSuccess ||
      hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    set_error(GG_ERR_HIP, "stream capture could not start");
    return plan_fail(p);
  }
  int rc = enqueue(p->A, p->D, p->src.data(), p->rel.data(), p->sym.data(), nullptr, nullptr,
                   max_patch > 0 ? &p->psrc : nullptr);
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(s, &graph);
  if (rc || e != hipSuccess || !graph ||
      hipGraphInstantiate(&p->exec, graph, nullptr, nullptr, 0) != hipSuccess) {
    if (!rc) set_error(GG_ERR_HIP, "graph capture/instantiate failed: %s", hipGetErrorString(e));
    if (graph) (void)hipGraphDestroy(graph);
    return plan_fail(p);
  }
  p->graph = graph;
  return p;
}

gg_plan *gg_plan_create(int count, const gg_graph *shapes) { return plan_create(count, shapes, 0); }

gg_plan *gg_plan_create_patched(int count, const gg_graph *shape, int max_patch) {
  if (!shape || count <= 0 || max_patch < 1) {
    set_error(GG_ERR_INVALID, "bad patched plan arguments");
    return nullptr;
  }
  if (!(shape->flags & GG_GRAPH_SYMMETRIC)) {
    set_error(GG_ERR_INVALID, "patched plans need GG_GRAPH_SYMMETRIC (moves keep symmetry)");
    return nullptr;
  }
  std::vector<gg_graph> shapes(count, *shape);
  return plan_create(count, shapes.data(), max_patch);
}

int gg_plan_set_base(gg_plan *p, const int32_t *adj_host) {
  if (!p || !p->max_patch || !adj_host) return set_error(GG_ERR_INVALID, "not a patched plan");
  const Slot &s = p->A.slots[0];
  GG_HIP(hipMemcpyAsync(p->base.p, adj_host, (size_t)s.n * s.d * 4, hipMemcpyHostToDevice,
                        engine().stream));
  return GG_OK;
}

int32_t *gg_plan_base_device(gg_plan *p) {
  return (p && p->max_patch) ? (int32_t *)p->base.p : nullptr;
}

int gg_plan_set_patches(gg_plan *p, const int32_t *patches_host) {
  if (!p || !p->max_patch || !patches_host) return set_error(GG_ERR_INVALID, "not a patched plan");
  const Slot &s = p->A.slots[0];
  const size_t cnt = p->A.slots.size() * (size_t)p->max_patch;
  for (size_t k = 0; k < cnt; ++k) {
    const int32_t at = patches_host[2 * k], val = patches_host[2 * k + 1];
    if (at != -1 && (at < 0 || at >= s.n * s.d || val < 0 || val >= s.n))
      return set_error(GG_ERR_INVALID, "patch %zu (%d -> %d) out of range", k, at, val);
  }
  GG_HIP(hipMemcpyAsync(p->patches.p, patches_host, cnt * 2 * 4, hipMemcpyHostToDevice,
                        engine().stream));
  return GG_OK;
}

int gg_plan_set_adj(gg_plan *p, int index, const int32_t *adj_host) {
  if (!p || index < 0 || index >= (int)p->A.slots.size() || !adj_host)
    return set_error(GG_ERR_INVALID, "bad plan slot");
  const Slot &s = p->A.slots[index];
  GG_HIP(hipMemcpyAsync((int32_t *)p->D.adj_in.p + s.adj_off, adj_host, (size_t)s.n * s.d * 4,
                        hipMemcpyHostToDevice, engine().stream));
  return GG_OK;
}

int32_t *gg_plan_adj_device(gg_plan *p, int index) {
  if (!p || index < 0 || index >= (int)p->A.slots.size()) return nullptr;
  return (int32_t *)p->D.adj_in.p + p->A.slots[index].adj_off;
}

int gg_plan_launch(gg_plan *p) {
  if (!p || !p->exec) return set_error(GG_ERR_STATE, "plan not ready");
  GG_HIP(hipGraphLaunch(p->exec, engine().stream));
  p->launched = true;
  p->timed = false;
  return GG_OK;
}

int gg_plan_wait(gg_plan *p, gg_result *out) {
  if (!p || !p->launched) return set_error(GG_ERR_STATE, "plan not launched");
  p->launched = false;
  return collect(p->A, p->D, out);
}

int gg_plan_run(gg_plan *p, gg_result *out) {
  int rc = gg_plan_launch(p);
  if (rc) return rc;
  return gg_plan_wait(p, out);
}

// The plan's kernels enqueued directly on the stream (not through the
// captured graph), with HIP events around the tile-BFS launches so their
// duration can be read by gg_plan_kernel_ms() after gg_plan_wait().
int gg_plan_launch_timed(gg_plan *p) {
  if (!p || !p->exec) return set_error(GG_ERR_STATE, "plan not ready");
  int rc = enqueue(p->A, p->D, p->src.data(), p->rel.data(), p->sym.data(), p->tev_begin,
                   p->tev_end, p->max_patch > 0 ? &p->psrc : nullptr);
  if (rc) return rc;
  p->launched = true;
  p->timed = true;
  return GG_OK;
}

float gg_plan_kernel_ms(gg_plan *p) {
  // Only eager timed launches carry events (the captured graph has none).
  if (!p || !p->timed || !p->tev_begin) return -1.0f;
  float ms = -1.0f;
  if (hipEventElapsedTime(&ms, p->tev_begin, p->tev_end) != hipSuccess) {
    (void)hipGetLastError();  // leave no error behind for the next launch's check
    return -1.0f;
  }
  return ms;
}

int gg_plan_config(gg_plan *p, int32_t *config) {
  if (!p || !config) return set_error(GG_ERR_INVALID, "bad plan config arguments");
  const Slot &s = p->A.slots[0];
  config[0] = s.bits;
  config[1] = s.block;
  config[2] = p->A.order_mode;
  config[3] = p->A.persistent ? 1 : 0;
  return GG_OK;
}

int gg_synchronize(void) {
  if (!engine().ready) return GG_OK;
  GG_HIP(hipStreamSynchronize(engine().stream));
  GG_HIP(hipDeviceSynchronize());
  return GG_OK;
}

void gg_plan_destroy(gg_plan *p) {
  if (!p) return;
  (void)hipStreamSynchronize(engine().stream);
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  p->base.release();
  p->patches.release();
  for (hipEvent_t ev : {p->tev_begin, p->tev_end})
    if (ev) (void)hipEventDestroy(ev);
  p->D.release();
  delete p;
}

}  // extern "C"
