// gg_tile.h -- the tile BFS kernel template and its device helpers, shared
// by the evaluator (gg_eval.hip) and the incremental scorer (gg_incr.hip).
// Everything here is a template or an inline device function.
#ifndef GG_TILE_H
#define GG_TILE_H

#include <hip/hip_runtime.h>

#include <cstdint>

// ---------------------------------------------------------------------------
// Device-side descriptors.

struct GraphDev {
  const uint16_t *adjR;  // [n][dp] row-major, padded with the row's own id
  int32_t *flags;        // [0] = bad adjacency entries, [1] = asymmetric
  int32_t n, dp, split, c;
  int32_t push_ok, groups;
  int32_t hist_maxd;     // >0: accumulate per-vertex distance histogram
  int32_t trace_units;   // units whose per-level timings are recorded
  int32_t *hist;         // [n][hist_maxd+1] (histogram mode only)
  uint64_t *trace;       // [trace_units][kTraceLevels][2] or null
  // Source-list modes (kTileList, kTileProbe, kTileSrc; groups = 1, no
  // centres): a unit's sources are srcs[first_row .. first_row + l).  With
  // src_count set the launch is persistent: block b runs units b,
  // b + gridDim.x, ... below ceil(*src_count / W), unit u taking list
  // entries [W*u, min(W*u + W, *src_count)) and writing outs[u].
  const uint16_t *srcs = nullptr;
  const int32_t *src_count = nullptr;
  // kTileProbe, per unit [unit][n]: the level at which the unit's first
  // source reaches v, and the first level at which any other source does.
  uint16_t *probe_l0 = nullptr, *probe_l1 = nullptr;
  // kTileSrc, per list entry: distance sum and eccentricity (0xffff when
  // some vertex is unreached).
  uint32_t *src_sum = nullptr;
  uint16_t *src_ecc = nullptr;
  // 1: the last 4 entries of every row are padding (degree <= dp - 4); the
  // pull levels of one-chunk rows then skip them.
  int32_t pad4 = 0;
};

// Kernel modes.
constexpr int kTileNormal = 0;  // (graph, tile) units, sums per unit
constexpr int kTileHist = 1;    // + per-vertex distance histograms
constexpr int kTileList = 2;    // sources from a list, persistent over *src_count
constexpr int kTileProbe = 3;   // sources from a list (UnitDesc units) + probe levels
constexpr int kTileSrc = 4;     // kTileList + per-source distance sums and eccentricities

// Per-level trace record: cycles since the unit started at the level's end,
// mode word, and (wave 0's view) cycles when its reads of the level were
// issued and when the barrier after them released.
constexpr int kTraceLevels = 64;
constexpr int kTraceFields = 4;

struct UnitDesc {
  int32_t graph;
  int32_t first_row;
  int32_t l;
  int32_t centre;
};

struct UnitOut {
  uint64_t acc_lo, acc_hi;
  int32_t done_level, last_change, disconnected, error;
};


// ---------------------------------------------------------------------------
// Word helpers.

template <typename W> struct WordBits { static constexpr int value = 8 * sizeof(W); };

// Word 3 of a raw buffer resource on gfx9-family parts (32-bit data format).
constexpr int kBufferWord3 = 0x00020000;

// Pointer into the global address space.
template <typename T> using gptr = __attribute__((address_space(1))) T *;

template <typename W> __device__ __forceinline__ int wpopc(W x) {
  if constexpr (sizeof(W) == 8) return __popcll((unsigned long long)x);
  else return __popc((unsigned)x);
}

// LDS atomic OR of `f` into element x of R; returns the element's old value.
template <typename W> __device__ __forceinline__ W lds_or(W *R, uint32_t x, W f) {
  if constexpr (sizeof(W) == 8) {
    return (W)atomicOr((unsigned long long *)(R + x), (unsigned long long)f);
  } else if constexpr (sizeof(W) == 4) {
    return (W)atomicOr((unsigned *)(R + x), (unsigned)f);
  } else if constexpr (sizeof(W) == 2) {
    unsigned *w = (unsigned *)R + (x >> 1);
    unsigned sh = (x & 1u) * 16u;
    return (W)(atomicOr(w, (unsigned)f << sh) >> sh);
  } else {
    unsigned *w = (unsigned *)R + (x >> 2);
    unsigned sh = (x & 3u) * 8u;
    return (W)(atomicOr(w, (unsigned)f << sh) >> sh);
  }
}

// LDS atomic OR of `f` into element x of R, without the old value (the
// non-returning form: the wave does not wait for it).
template <typename W> __device__ __forceinline__ void lds_or_nr(W *R, uint32_t x, W f) {
  if constexpr (sizeof(W) == 8) {
    atomicOr((unsigned long long *)(R + x), (unsigned long long)f);
  } else if constexpr (sizeof(W) == 4) {
    atomicOr((unsigned *)(R + x), (unsigned)f);
  } else if constexpr (sizeof(W) == 2) {
    atomicOr((unsigned *)R + (x >> 1), (unsigned)f << ((x & 1u) * 16u));
  } else {
    atomicOr((unsigned *)R + (x >> 2), (unsigned)f << ((x & 3u) * 8u));
  }
}

// Register staging of VPT words per thread, packed into 32-bit registers.
template <typename W, int VPT> struct StageRegs {
  static constexpr int value = (VPT * (int)sizeof(W) + 3) / 4;
};

template <typename W> __device__ __forceinline__ void stage_put(uint32_t *st, int r, W val) {
  if constexpr (sizeof(W) == 8) {
    st[2 * r] = (uint32_t)val;
    st[2 * r + 1] = (uint32_t)((uint64_t)val >> 32);
  } else if constexpr (sizeof(W) == 4) {
    st[r] = (uint32_t)val;
  } else if constexpr (sizeof(W) == 2) {
    if (r & 1) st[r >> 1] |= (uint32_t)val << 16;
    else st[r >> 1] = (uint32_t)val;
  } else {
    if (r & 3) st[r >> 2] |= (uint32_t)val << ((r & 3) * 8);
    else st[r >> 2] = (uint32_t)val;
  }
}

template <typename W> __device__ __forceinline__ W stage_get(const uint32_t *st, int r) {
  if constexpr (sizeof(W) == 8) return (W)(((uint64_t)st[2 * r + 1] << 32) | st[2 * r]);
  else if constexpr (sizeof(W) == 4) return (W)st[r];
  else if constexpr (sizeof(W) == 2) return (W)(st[r >> 1] >> ((r & 1) * 16));
  else return (W)(st[r >> 2] >> ((r & 3) * 8));
}

// Cross-lane sums with DPP (VALU lane moves) instead of ds_bpermute, which
// would queue behind the gathers in the LDS pipe.  DPP controls (GFX9):
// quad_perm [1,0,3,2] = 0xB1, [2,3,0,1] = 0x4E; row_shr:k = 0x110 + k;
// row_ror:k = 0x120 + k; row_bcast:15 = 0x142; row_bcast:31 = 0x143.
#define GG_DPP(src, ctrl, rm, bm) __builtin_amdgcn_update_dpp(0u, (src), (ctrl), (rm), (bm), true)

// Sum over the 64 lanes, returned in every lane (uniform).
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
  v += GG_DPP(v, 0xB1, 0xf, 0xf);
  v += GG_DPP(v, 0x4E, 0xf, 0xf);
  v += GG_DPP(v, 0x124, 0xf, 0xf);
  v += GG_DPP(v, 0x128, 0xf, 0xf);  // every lane: its 16-lane row's sum
  v += GG_DPP(v, 0x142, 0xa, 0xf);  // rows 1, 3 += rows 0, 2
  v += GG_DPP(v, 0x143, 0xc, 0xf);  // rows 2, 3 += rows 0-1
  return __builtin_amdgcn_readlane(v, 63);
}

// Inclusive prefix sum over the lanes (lane i: sum of lanes 0..i).
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
  uint32_t t = x + GG_DPP(x, 0x111, 0xf, 0xf);
  t += GG_DPP(x, 0x112, 0xf, 0xf);
  t += GG_DPP(x, 0x113, 0xf, 0xf);  // lanes i-3..i of the row
  t += GG_DPP(t, 0x114, 0xf, 0xe);  // lanes 4-15 of a row: += 4 back
  t += GG_DPP(t, 0x118, 0xf, 0xc);  // lanes 8-15: += 8 back (row prefix done)
  t += GG_DPP(t, 0x142, 0xa, 0xf);  // rows 1, 3 += last lane of rows 0, 2
  t += GG_DPP(t, 0x143, 0xc, 0xf);  // rows 2, 3 += lane 31
  return t;
}


// ---------------------------------------------------------------------------
// The tile BFS.  One workgroup = one UnitDesc.  LDS layout (dynamic):
//   R[n_cap] (W) | qb[2][qcap] (W) | qv[2][qcap] (u16) | red[3][16][4] |
//   qcount[4] | sacc[2][64]
// tile_lds_bytes gives its size for `wb`-byte words.
__host__ __device__ constexpr size_t tile_align16(size_t x) { return (x + 15) & ~(size_t)15; }
__host__ __device__ constexpr size_t tile_lds_bytes(size_t wb, int n_cap, int qcap) {
  return tile_align16((size_t)n_cap * wb) + tile_align16((size_t)2 * qcap * wb) +
         tile_align16((size_t)2 * qcap * 2) + 3 * 16 * 4 * 4 + 16 + 2 * 64 * 4;
}

// Sum of the popcounts of a 16-byte chunk.
__device__ __forceinline__ uint32_t popc_chunk(const uint32_t x, const uint32_t y, const uint32_t z,
                                               const uint32_t w) {
  return (uint32_t)(__popc(x) + __popc(y) + __popc(z) + __popc(w));
}

// 16 bytes of an adjacency row: 8 uint16 neighbour ids.
typedef uint32_t Chunk __attribute__((ext_vector_type(4)));

// Byte offset into a W array of the id held in the low (H = 0) or high
// (H = 1) 16 bits of x: one SDWA shift per id instead of an extract and a
// shift.
template <typename W, int H> __device__ __forceinline__ uint32_t id_off(uint32_t x) {
  constexpr int S = sizeof(W) == 1 ? 0 : sizeof(W) == 2 ? 1 : sizeof(W) == 4 ? 2 : 3;
  uint32_t r;
  if constexpr (H == 0)
    asm("v_lshlrev_b32_sdwa %0, %2, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD "
        "src1_sel:WORD_0"
        : "=v"(r)
        : "v"(x), "n"(S));
  else
    asm("v_lshlrev_b32_sdwa %0, %2, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD "
        "src1_sel:WORD_1"
        : "=v"(r)
        : "v"(x), "n"(S));
  return r;
}

// The reach array R starts at LDS address 0 (the kernel checks this on
// entry), so an id's byte offset is its LDS address and the gathers need no
// base add.
template <typename W> __device__ __forceinline__ W word_at(uint32_t lds_addr) {
  return *(const __attribute__((address_space(3))) W *)(size_t)lds_addr;
}

// OR of the reach words of 4 neighbour ids (two packed uint32 pairs).
template <typename W> __device__ __forceinline__ W gather4(const W *, uint32_t a, uint32_t b) {
  return (W)(word_at<W>(id_off<W, 0>(a)) | word_at<W>(id_off<W, 1>(a)) |
             word_at<W>(id_off<W, 0>(b)) | word_at<W>(id_off<W, 1>(b)));
}

// ORs a chunk's 8 neighbours into val, in two halves, stopping once full.
template <typename W> __device__ __forceinline__ W gather_chunk(const W *R, Chunk c, W val, W full) {
  val = (W)(val | gather4<W>(R, c.x, c.y));
  if (val != full) val = (W)(val | gather4<W>(R, c.z, c.w));
  return val;
}

template <typename W, int BLOCK, int VPT, int MODE>
__global__ void __launch_bounds__(BLOCK)
gg_tile_bfs_kernel(const GraphDev *__restrict__ graphs, const UnitDesc *__restrict__ units,
                   UnitOut *__restrict__ outs, int unit_base, int n_cap, int qcap) {
  const int uidx = unit_base + blockIdx.x;
  const UnitDesc U = units[uidx];
  const GraphDev G = graphs[U.graph];
#define GG_UNIT_EXIT return
#include "gg_tile_body.inc"
#undef GG_UNIT_EXIT
}

// Persistent variant for batches: every workgroup takes (graph, tile) units
// from a device-side counter until `unit_count` are taken, so one launch
// drains the whole batch (units are listed largest graph first, so the
// last units to start are the cheapest).  *counter must be 0 at launch.
// The unit index is handed to the workgroup through the 16 bytes after the
// body's LDS layout (a static __shared__ variable would move the dynamic
// LDS, whose reach array must sit at address 0); the launch adds them.
template <typename W, int BLOCK, int VPT, int MODE>
__global__ void __launch_bounds__(BLOCK)
gg_tile_queue_kernel(const GraphDev *__restrict__ graphs, const UnitDesc *__restrict__ units,
                     UnitOut *__restrict__ outs, int unit_base, int unit_count, int n_cap, int qcap,
                     int *__restrict__ counter) {
  extern __shared__ __align__(16) unsigned char smem_q[];
  volatile int *s_next = (volatile int *)(smem_q + tile_lds_bytes(sizeof(W), n_cap, qcap));
  for (;;) {
    __syncthreads();  // LDS (and the slot) is reused from the previous unit
    if (threadIdx.x == 0) *s_next = atomicAdd(counter, 1);
    __syncthreads();
    // Uniform (one value for the workgroup): readfirstlane keeps the unit's
    // descriptors in scalar registers, as in the one-unit kernels.
    const int k = __builtin_amdgcn_readfirstlane(*s_next);
    if (k >= unit_count) break;
    const int uidx = unit_base + k;
    const UnitDesc U = units[uidx];
    const GraphDev G = graphs[U.graph];
#define GG_UNIT_EXIT continue
#include "gg_tile_body.inc"
#undef GG_UNIT_EXIT
  }
}

// Persistent variant for the source-list modes over graphs[0]: units are
// derived from the device-side list length, so the launch needs no host
// round trip.  Block b runs units b, b + gridDim.x, ...
template <typename W, int BLOCK, int VPT, int MODE>
__global__ void __launch_bounds__(BLOCK)
gg_tile_list_kernel(const GraphDev *__restrict__ graphs, UnitOut *__restrict__ outs, int n_cap,
                    int qcap) {
  static_assert(MODE == kTileList || MODE == kTileSrc, "list kernel modes");
  const int k_list = *graphs[0].src_count;
  const int n_units = (k_list + WordBits<W>::value - 1) / WordBits<W>::value;
  for (int uu = blockIdx.x; uu < n_units; uu += gridDim.x) {
    __syncthreads();  // LDS is reused from the previous unit
    const int uidx = uu;
    const UnitDesc U = {0, uu * WordBits<W>::value, min(WordBits<W>::value, k_list - uu * WordBits<W>::value), 0};
    const GraphDev G = graphs[0];
#define GG_UNIT_EXIT continue
#include "gg_tile_body.inc"
#undef GG_UNIT_EXIT
  }
}

#endif  // GG_TILE_H
