// gg_lds_probe.hip -- measures the LDS random-gather throughput that bounds
// the tile BFS's pull levels on this GPU (a standalone executable).
//
// One 1024-thread workgroup per CU holds 65536 uint16 words (128 KiB) in LDS
// and ORs random words into a register, like a pull level's gathers.  Each
// block is timed from a barrier to a barrier (timing one wave alone
// overstates the rate: the LDS arbiter favours the oldest wave).  Modes:
//   random   16 independent random ids per lane per iteration
//   rows     ids read from 32-byte adjacency rows in global memory, each row
//            loaded one iteration ahead (the kernel's situation)
// and, with each lane's 16 ids fixed in registers (XORed every iteration
// with a multiple of 64 dwords, which keeps every id's bank), bank patterns:
//   fixed    random ids
//   spread   the 32 lanes of each half-wave on 32 distinct 4-byte banks
//   2-way    lane pairs on one bank (different words)
//   one-bank every lane of a half-wave on one bank
// Output: cycles per 64-lane gather instruction per CU (shader clock).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <vector>

typedef uint32_t Chunk __attribute__((ext_vector_type(4)));

template <int MODE>
__global__ void __launch_bounds__(1024)
lds_probe(const Chunk *rows, int nrows, uint32_t *out, int iters, unsigned long long *cyc) {
  extern __shared__ unsigned char sm[];
  uint16_t *R = (uint16_t *)sm;
  for (int i = threadIdx.x; i < 65536; i += 1024) R[i] = (uint16_t)(i * 2654435761u);
  __syncthreads();
  uint32_t x = (threadIdx.x * 0x9E3779B9u) ^ (blockIdx.x * 0x85ebca6bu) ^ 12345u;
  uint32_t acc = 0;
  int v = (blockIdx.x * 1024 + threadIdx.x) % nrows;
  Chunk na = rows[2 * v], nb = rows[2 * v + 1];
  __syncthreads();
  const unsigned long long t0 = clock64();
  for (int it = 0; it < iters; ++it) {
    uint32_t id[16];
    if (MODE == 1) {
      const Chunk a = na, b = nb;
      v = (v + 1024 * 7 + 13) % nrows;
      na = rows[2 * v];
      nb = rows[2 * v + 1];
      const uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        id[2 * k] = w[k] & 0xffffu;
        id[2 * k + 1] = w[k] >> 16;
      }
    } else {
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        x = x * 1664525u + 1013904223u;
        id[k] = x >> 16;
      }
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) acc |= (uint32_t)R[id[k]];
  }
  __syncthreads();
  const unsigned long long t1 = clock64();
  out[blockIdx.x * 1024 + threadIdx.x] = acc;
  if (threadIdx.x == 0) cyc[blockIdx.x] = t1 - t0;
}

// Bank patterns: the 4-byte word (dword) each lane's id lies in.
template <int P>
__device__ __forceinline__ uint32_t bank_id(uint32_t lane, uint32_t rnd) {
  uint32_t dw;
  if (P == 0) return rnd & 0xffffu;                            // fixed random ids
  if (P == 1) dw = (lane & 31u) + 32u * (rnd & 1023u);          // distinct banks per half-wave
  if (P == 2) dw = ((lane >> 1) & 31u) + 32u * (rnd & 1023u);   // lane pairs share a bank
  if (P == 3) dw = 32u * (rnd & 1023u);                         // one bank
  return ((dw & 32767u) << 1) | ((rnd >> 20) & 1u);
}

template <int P>
__global__ void __launch_bounds__(1024)
lds_bank_probe(const Chunk *, int, uint32_t *out, int iters, unsigned long long *cyc) {
  extern __shared__ unsigned char sm[];
  uint16_t *R = (uint16_t *)sm;
  for (int i = threadIdx.x; i < 65536; i += 1024) R[i] = (uint16_t)(i * 2654435761u);
  const uint32_t lane = threadIdx.x & 63u;
  uint32_t x = (threadIdx.x * 0x9E3779B9u) ^ (blockIdx.x * 0x85ebca6bu) ^ 12345u;
  uint32_t id[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    x = x * 1664525u + 1013904223u;
    id[k] = bank_id<P>(lane, x >> 8) & 0xffffu;
  }
  uint32_t acc = 0;
  __syncthreads();
  const unsigned long long t0 = clock64();
  for (int it = 0; it < iters; ++it) {
    const uint32_t xo = ((uint32_t)it & 255u) << 7;  // 64 dwords: the bank stays
#pragma unroll
    for (int k = 0; k < 16; ++k) acc |= (uint32_t)R[id[k] ^ xo];
  }
  __syncthreads();
  const unsigned long long t1 = clock64();
  out[blockIdx.x * 1024 + threadIdx.x] = acc;
  if (threadIdx.x == 0) cyc[blockIdx.x] = t1 - t0;
}

#define CHECK(x)                                                         \
  do {                                                                   \
    hipError_t e_ = (x);                                                 \
    if (e_ != hipSuccess) {                                              \
      fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));            \
      return 1;                                                          \
    }                                                                    \
  } while (0)

int main() {
  const int nrows = 65536, iters = 256;
  int dev = 0, cus = 0;
  CHECK(hipGetDevice(&dev));
  CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int blocks = cus;
  std::vector<uint32_t> h((size_t)nrows * 8);
  uint32_t s = 7;
  for (auto &w : h) w = (s = s * 1664525u + 1013904223u);
  Chunk *rows;
  uint32_t *out;
  unsigned long long *cyc;
  CHECK(hipMalloc(&rows, (size_t)nrows * 32));
  CHECK(hipMalloc(&out, (size_t)blocks * 1024 * 4));
  CHECK(hipMalloc(&cyc, (size_t)blocks * 8));
  CHECK(hipMemcpy(rows, h.data(), (size_t)nrows * 32, hipMemcpyHostToDevice));
  std::vector<unsigned long long> hc(blocks);
  auto run = [&](const char *name, auto k) -> int {
    CHECK(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, 131072));
    for (int rep = 0; rep < 2; ++rep)
      hipLaunchKernelGGL(k, dim3(blocks), dim3(1024), 131072, 0, rows, nrows, out, iters, cyc);
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    CHECK(hipMemcpy(hc.data(), cyc, (size_t)blocks * 8, hipMemcpyDeviceToHost));
    double avg = 0;
    for (auto c : hc) avg += (double)c;
    avg /= blocks;
    const double wave_reads = 1024.0 * iters * 16 / 64;
    printf("%-8s %.3f cycles per 64-lane gather instruction per CU (%d blocks)\n", name,
           avg / wave_reads, blocks);
    return 0;
  };
  if (run("random", lds_probe<0>) || run("rows", lds_probe<1>) || run("fixed", lds_bank_probe<0>) ||
      run("spread", lds_bank_probe<1>) || run("2-way", lds_bank_probe<2>) ||
      run("one-bank", lds_bank_probe<3>))
    return 1;
  return 0;
}
