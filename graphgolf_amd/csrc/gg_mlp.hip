// gg_mlp.hip -- the swap-ranking surrogate: a small MLP trained and run on
// the GPU with the project's own BF16 MFMA kernels (FP32 accumulate).
//
// Network: dims[0] -> dims[1] -> ... -> dims[L] with ReLU between layers and
// a linear output (the predicted change of the distance sum for a swap).
// Master parameters are FP32; every GEMM reads BF16 copies of its operands
// and accumulates in FP32 with gfx950's v_mfma_f32_32x32x16_bf16 (64x64
// output tiles per workgroup, operands staged in LDS).  Training minimises
// mean squared error with Adam.
//
// No reference counterpart: the reference proposes swaps uniformly at random
// (general/sa.c:217-369); this network only re-orders proposals, and every
// accepted move is still scored by the exact evaluator.

#include <hip/hip_runtime.h>
#include <cassert>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "gg_eval.h"
#include "gg_internal.h"
#include "gg_device.h"
#include "gg_mlp.h"

using gg::set_error;

typedef float f32x16 __attribute__((ext_vector_type(16)));

// FP32 -> BF16 with round-to-nearest-even (NaN kept quiet).
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__global__ void gg_mlp_to_bf16(const float *__restrict__ in, uint16_t *__restrict__ out, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = f2bf(in[i]);
}

// Rows of `cols` fp32 -> rows of `cols_p` BF16, the padding columns zero.
__global__ void gg_mlp_to_bf16_rows(const float *__restrict__ in, int cols, uint16_t *__restrict__ out,
                                    int cols_p, int rows) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * cols_p) return;
  const int r = i / cols_p, c = i - r * cols_p;
  out[i] = c < cols ? f2bf(in[(size_t)r * cols + c]) : (uint16_t)0;
}

// C[M x N] (fp32, row-major, ldc) = op(A)[M x K] * op(B)[K x N] (+ bias[N])
// (then ReLU when relu != 0); optionally also writes a BF16 copy (ldh).
// op(A)(i,k) = transA ? A[k*lda + i] : A[i*lda + k]
// op(B)(k,j) = transB ? B[j*ldb + k] : B[k*ldb + j]
// M, K, lda, ldb multiples of 8 (16-byte rows of BF16), N of 32.
//
// One 256-thread workgroup per 64x64 tile of C, one 32x32 quarter per wave,
// with gfx950's v_mfma_f32_32x32x16_bf16 (K = 16 per instruction).  Each
// K step of 32 stages op(A)'s 64x32 block as As[i][k] and op(B)'s 32x64
// block as Bs[j][k] in LDS (rows padded by 8 elements against bank
// conflicts), read from global memory as 16-byte vectors in each operand's
// own memory order (transposed operands are scattered into LDS), so that
// every lane reads its 8 consecutive k of one row or column with one
// 16-byte LDS read.  Operand layout of the MFMA: lane l holds A(l%32,
// 8*(l/32)+e) and B(8*(l/32)+e, l%32), e = 0..7; accumulator register r of
// lane l is D(8*(r/4) + 4*(l/32) + r%4, l%32).  Shapes the callers keep:
// N a multiple of 32 (a wave's quarter is whole or wholly past N), K a
// multiple of 8 and, for a transposed A, M a multiple of 8 (each 16-byte
// load is whole or wholly out of range); rows past M are not stored.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
constexpr int kGemmTile = 64, kGemmK = 32, kGemmLd = kGemmK + 8;

__global__ void __launch_bounds__(256)
gg_mlp_gemm(const uint16_t *__restrict__ A, int lda, int transA, const uint16_t *__restrict__ B,
            int ldb, int transB, float *__restrict__ C, int ldc, uint16_t *__restrict__ Cb, int ldh,
            const float *__restrict__ bias, int relu, int M, int N, int K) {
  __shared__ __attribute__((aligned(16))) uint16_t As[kGemmTile][kGemmLd];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[kGemmTile][kGemmLd];
  const int tid = threadIdx.x, l = tid & 63, w = tid >> 6;
  const int m0 = blockIdx.x * kGemmTile, n0 = blockIdx.y * kGemmTile;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;  // this wave's quarter of the tile
  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
  const u32x4 zero = {0u, 0u, 0u, 0u};
  for (int k0 = 0; k0 < K; k0 += kGemmK) {
    // Stage op(A) rows m0.., k0.. : 256 loads of 8 BF16.
    if (!transA) {
      const int i = tid >> 2, kk = (tid & 3) * 8;  // row i, k kk..kk+7
      const int gi = m0 + i, gk = k0 + kk;
      u32x4 v = zero;
      if (gi < M && gk < K) v = *(const u32x4 *)(A + (size_t)gi * lda + gk);
      *(u32x4 *)&As[i][kk] = v;
    } else {
      const int kk = tid >> 3, i = (tid & 7) * 8;  // k kk, rows i..i+7 (contiguous in memory)
      const int gk = k0 + kk, gi = m0 + i;
      u32x4 v = zero;
      if (gk < K && gi < M) v = *(const u32x4 *)(A + (size_t)gk * lda + gi);
      const uint16_t *e = (const uint16_t *)&v;
#pragma unroll
      for (int q = 0; q < 8; ++q) As[i + q][kk] = e[q];
    }
    // Stage op(B) columns n0.., k0.. as Bs[j][k].
    if (transB) {
      const int j = tid >> 2, kk = (tid & 3) * 8;  // column j, k kk..kk+7 (contiguous)
      const int gj = n0 + j, gk = k0 + kk;
      u32x4 v = zero;
      if (gj < N && gk < K) v = *(const u32x4 *)(B + (size_t)gj * ldb + gk);
      *(u32x4 *)&Bs[j][kk] = v;
    } else {
      const int kk = tid >> 3, j = (tid & 7) * 8;  // k kk, columns j..j+7 (contiguous)
      const int gk = k0 + kk, gj = n0 + j;
      u32x4 v = zero;
      if (gk < K && gj < N) v = *(const u32x4 *)(B + (size_t)gk * ldb + gj);
      const uint16_t *e = (const uint16_t *)&v;
#pragma unroll
      for (int q = 0; q < 8; ++q) Bs[j + q][kk] = e[q];
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < kGemmK; ks += 16) {
      const int kq = ks + 8 * (l >> 5);
      const bf16x8 a = *(const bf16x8 *)&As[wm + (l & 31)][kq];
      const bf16x8 b = *(const bf16x8 *)&Bs[wn + (l & 31)][kq];
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
  if (m0 + wm >= M || n0 + wn >= N) return;  // a quarter past the matrix
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int i = m0 + wm + 8 * (r >> 2) + 4 * (l >> 5) + (r & 3);
    const int j = n0 + wn + (l & 31);
    if (i >= M) continue;
    float v = acc[r];
    if (bias) v += bias[j];
    if (relu) v = fmaxf(v, 0.0f);
    C[(size_t)i * ldc + j] = v;
    if (Cb) Cb[(size_t)i * ldh + j] = f2bf(v);
  }
}

// dy = 2 (y - t) / m for real rows (0 for padding), written into column 0 of
// a [Mp x ldd] gradient (other columns 0); the loss sum goes to *loss.
__global__ void gg_mlp_mse(const float *__restrict__ y, int ldy, const float *__restrict__ t, int m,
                           int Mp, float *__restrict__ dy, int ldd, float *__restrict__ loss) {
  __shared__ float part[256];
  float s = 0.0f;
  for (int i = threadIdx.x; i < Mp; i += 256) {
    float g = 0.0f;
    if (i < m) {
      const float e = y[(size_t)i * ldy] - t[i];
      s += e * e;
      g = 2.0f * e / (float)m;
    }
    for (int j = 0; j < ldd; ++j) dy[(size_t)i * ldd + j] = (j == 0) ? g : 0.0f;
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = part[0] / (float)m;
}

// dZ = dH * (Z > 0), and its BF16 copy.
__global__ void gg_mlp_relu_back(const float *__restrict__ dH, const float *__restrict__ Z,
                                 float *__restrict__ dZ, uint16_t *__restrict__ dZb, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const float g = Z[i] > 0.0f ? dH[i] : 0.0f;
    dZ[i] = g;
    dZb[i] = f2bf(g);
  }
}

// db[j] = sum_i dZ[i][j]: a block per 64 columns, 16 row groups of one
// wave each (a wave reads 64 consecutive columns of a row), the groups'
// partial sums added in a fixed order (the same result every run).
__global__ void __launch_bounds__(1024)
gg_mlp_colsum(const float *__restrict__ dZ, int M, int N, float *__restrict__ db) {
  __shared__ float part[16][64];
  const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int j = blockIdx.x * 64 + c;
  float s = 0.0f;
  if (j < N)
    for (int i = g; i < M; i += 16) s += dZ[(size_t)i * N + j];
  part[g][c] = s;
  __syncthreads();
  if (g == 0 && j < N) {
    float t = 0.0f;
#pragma unroll
    for (int q = 0; q < 16; ++q) t += part[q][c];
    db[j] = t;
  }
}

__global__ void gg_mlp_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m1,
                            float *__restrict__ m2, int n, float lr, float b1, float b2, float eps,
                            float c1, float c2) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.0f - b1) * gi;
  const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  p[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// y[i] = Z[i * ld] for i < m (the network's single output column).
__global__ void gg_mlp_col0(const float *__restrict__ Z, int ld, int m, float *__restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) y[i] = Z[(size_t)i * ld];
}

// ---------------------------------------------------------------------------

namespace {

int pad32(int x) { return (x + 31) & ~31; }

struct Layer {
  int in, out, out_p;       // out_p: output width padded to 32
  size_t w_off, b_off;      // offsets into the flat parameter arrays
};

}  // namespace

struct gg_mlp {
  std::vector<Layer> layers;
  int in_dim = 0, in_p = 0, n_params = 0, cap_m = 0;  // in_p: in_dim rounded up to 8
  int step = 0;
  hipStream_t stream = nullptr;
  // flat parameter buffers (fp32 master, grads, Adam moments, bf16 weights)
  float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
  uint16_t *pb = nullptr;
  // activations for a batch of cap_m rows
  uint16_t *x_b = nullptr;           // [Mp][in]
  std::vector<float *> Z;            // [Mp][out_p] pre-activation per layer
  std::vector<uint16_t *> Hb;        // [Mp][out_p] bf16 post-activation per layer
  float *dA = nullptr, *dB = nullptr;  // [Mp][maxw] gradient scratch
  uint16_t *dAb = nullptr;
  float *x_f = nullptr, *t_f = nullptr, *loss = nullptr;
  float *y_f = nullptr;              // [Mp] contiguous predictions
  int maxw = 0;
};

#define MLP_HIP(call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return set_error(GG_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));     \
  } while (0)

static int mlp_reserve(gg_mlp *M, int m) {
  const int Mp = pad32(m);
  if (Mp <= M->cap_m) return GG_OK;
  // Grow: release the row buffers (pointers cleared, so that a failed
  // allocation below leaves nothing for gg_mlp_destroy to free twice).
  for (float *z : M->Z) (void)hipFree(z);
  for (uint16_t *h : M->Hb) (void)hipFree(h);
  M->Z.clear();
  M->Hb.clear();
  for (void **q : {(void **)&M->x_b, (void **)&M->dA, (void **)&M->dB, (void **)&M->dAb,
                   (void **)&M->x_f, (void **)&M->t_f, (void **)&M->y_f}) {
    if (*q) (void)hipFree(*q);
    *q = nullptr;
  }
  M->cap_m = 0;
  MLP_HIP(hipMalloc(&M->x_b, (size_t)Mp * M->in_p * 2));
  MLP_HIP(hipMalloc(&M->x_f, (size_t)Mp * M->in_dim * 4));
  MLP_HIP(hipMalloc(&M->t_f, (size_t)Mp * 4));
  MLP_HIP(hipMalloc(&M->y_f, (size_t)Mp * 4));
  for (const Layer &L : M->layers) {
    float *z;
    uint16_t *h;
    MLP_HIP(hipMalloc(&z, (size_t)Mp * L.out_p * 4));
    MLP_HIP(hipMalloc(&h, (size_t)Mp * L.out_p * 2));
    M->Z.push_back(z);
    M->Hb.push_back(h);
  }
  MLP_HIP(hipMalloc(&M->dA, (size_t)Mp * M->maxw * 4));
  MLP_HIP(hipMalloc(&M->dB, (size_t)Mp * M->maxw * 4));
  MLP_HIP(hipMalloc(&M->dAb, (size_t)Mp * M->maxw * 2));
  M->cap_m = Mp;
  return GG_OK;
}

static void launch_gemm(hipStream_t s, const uint16_t *A, int lda, int tA, const uint16_t *B, int ldb,
                        int tB, float *C, int ldc, uint16_t *Cb, int ldh, const float *bias, int relu,
                        int Mr, int Nr, int K) {
  // Widths are padded at creation (outputs to 32, the input to 8).
  assert(Nr % 32 == 0 && K % 8 == 0 && (!tA || Mr % 8 == 0));
  hipLaunchKernelGGL(gg_mlp_gemm, dim3((Mr + kGemmTile - 1) / kGemmTile, (Nr + kGemmTile - 1) / kGemmTile),
                     dim3(256), 0, s, A, lda, tA, B, ldb, tB, C, ldc, Cb, ldh, bias, relu, Mr, Nr, K);
}

// Uploads x (m rows, fp32) and runs the forward pass; output in Z.back() col 0.
static int mlp_forward(gg_mlp *M, const float *x_host, int m, bool device_x) {
  int rc = mlp_reserve(M, m);
  if (rc) return rc;
  const int Mp = pad32(m), in = M->in_dim, in_p = M->in_p;
  MLP_HIP(hipMemsetAsync(M->x_f, 0, (size_t)Mp * in * 4, M->stream));
  MLP_HIP(hipMemcpyAsync(M->x_f, x_host, (size_t)m * in * 4,
                         device_x ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, M->stream));
  const int nx = Mp * in_p;
  hipLaunchKernelGGL(gg_mlp_to_bf16_rows, dim3((nx + 255) / 256), dim3(256), 0, M->stream, M->x_f, in,
                     M->x_b, in_p, Mp);
  MLP_HIP(hipGetLastError());
  const uint16_t *h = M->x_b;
  int hw = in_p;
  for (size_t li = 0; li < M->layers.size(); ++li) {
    const Layer &L = M->layers[li];
    const bool last = li + 1 == M->layers.size();
    launch_gemm(M->stream, h, hw, 0, M->pb + L.w_off, L.out_p, 0, M->Z[li], L.out_p, M->Hb[li],
                L.out_p, M->p + L.b_off, last ? 0 : 1, Mp, L.out_p, L.in);
    h = M->Hb[li];
    hw = L.out_p;
  }
  MLP_HIP(hipGetLastError());
  return GG_OK;
}

// Backward pass from targets t (device t_f); fills M->g.
static int mlp_backward(gg_mlp *M, int m) {
  const int Mp = pad32(m);
  const int nl = (int)M->layers.size();
  MLP_HIP(hipMemsetAsync(M->g, 0, (size_t)M->n_params * 4, M->stream));
  const Layer &Lo = M->layers[nl - 1];
  // dZ_last into dA (fp32) + bf16 copy
  hipLaunchKernelGGL(gg_mlp_mse, dim3(1), dim3(256), 0, M->stream, M->Z[nl - 1], Lo.out_p, M->t_f, m,
                     Mp, M->dA, Lo.out_p, M->loss);
  int n = Mp * Lo.out_p;
  hipLaunchKernelGGL(gg_mlp_to_bf16, dim3((n + 255) / 256), dim3(256), 0, M->stream, M->dA, M->dAb, n);
  for (int li = nl - 1; li >= 0; --li) {
    const Layer &L = M->layers[li];
    const uint16_t *hin = li ? M->Hb[li - 1] : M->x_b;
    const int win = li ? M->layers[li - 1].out_p : M->in_p;
    // dW = H_in^T * dZ   ([win x out_p]; only the first L.in rows are real)
    launch_gemm(M->stream, hin, win, 1, M->dAb, L.out_p, 0, M->g + L.w_off, L.out_p, nullptr, 0,
                nullptr, 0, L.in, L.out_p, Mp);
    hipLaunchKernelGGL(gg_mlp_colsum, dim3((L.out_p + 63) / 64), dim3(1024), 0, M->stream, M->dA, Mp,
                       L.out_p, M->g + L.b_off);
    if (li == 0) break;
    // dH_in = dZ * W^T   ([Mp x win])
    launch_gemm(M->stream, M->dAb, L.out_p, 0, M->pb + L.w_off, L.out_p, 1, M->dB, win, nullptr, 0,
                nullptr, 0, Mp, win, L.out_p);
    n = Mp * win;
    hipLaunchKernelGGL(gg_mlp_relu_back, dim3((n + 255) / 256), dim3(256), 0, M->stream, M->dB,
                       M->Z[li - 1], M->dA, M->dAb, n);
  }
  MLP_HIP(hipGetLastError());
  return GG_OK;
}

static int refresh_bf16(gg_mlp *M) {
  hipLaunchKernelGGL(gg_mlp_to_bf16, dim3((M->n_params + 255) / 256), dim3(256), 0, M->stream, M->p,
                     M->pb, M->n_params);
  MLP_HIP(hipGetLastError());
  return GG_OK;
}

// Output column 0 of the last forward pass -> y (host), contiguous.
static int copy_predictions(gg_mlp *M, int m, float *y) {
  const Layer &Lo = M->layers.back();
  hipLaunchKernelGGL(gg_mlp_col0, dim3((m + 255) / 256), dim3(256), 0, M->stream, M->Z.back(), Lo.out_p,
                     m, M->y_f);
  MLP_HIP(hipGetLastError());
  MLP_HIP(hipMemcpyAsync(y, M->y_f, (size_t)m * 4, hipMemcpyDeviceToHost, M->stream));
  MLP_HIP(hipStreamSynchronize(M->stream));
  return GG_OK;
}

// Forward + backward for rows x (host or device) and host targets t.
static int mlp_gradient(gg_mlp *M, const float *x, bool device_x, const float *t, int m, float *loss,
                        float *grad) {
  int rc = mlp_forward(M, x, m, device_x);
  if (rc) return rc;
  MLP_HIP(hipMemcpyAsync(M->t_f, t, (size_t)m * 4, hipMemcpyHostToDevice, M->stream));
  if ((rc = mlp_backward(M, m))) return rc;
  if (grad) MLP_HIP(hipMemcpyAsync(grad, M->g, (size_t)M->n_params * 4, hipMemcpyDeviceToHost, M->stream));
  if (loss) MLP_HIP(hipMemcpyAsync(loss, M->loss, 4, hipMemcpyDeviceToHost, M->stream));
  MLP_HIP(hipStreamSynchronize(M->stream));
  return GG_OK;
}

static int mlp_train(gg_mlp *M, const float *x, bool device_x, const float *t, int m, float lr,
                     float *loss) {
  int rc = mlp_gradient(M, x, device_x, t, m, loss, nullptr);
  if (rc) return rc;
  ++M->step;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - std::pow(b1, (float)M->step), c2 = 1.0f - std::pow(b2, (float)M->step);
  hipLaunchKernelGGL(gg_mlp_adam, dim3((M->n_params + 255) / 256), dim3(256), 0, M->stream, M->p, M->g,
                     M->m1, M->m2, M->n_params, lr, b1, b2, 1e-8f, c1, c2);
  MLP_HIP(hipGetLastError());
  if ((rc = refresh_bf16(M))) return rc;
  MLP_HIP(hipStreamSynchronize(M->stream));
  return GG_OK;
}

extern "C" {

gg_mlp *gg_mlp_create(int n_layers, const int *dims, uint64_t seed) {
  if (n_layers < 1 || !dims) {
    set_error(GG_ERR_INVALID, "bad MLP shape");
    return nullptr;
  }
  for (int i = 0; i <= n_layers; ++i)
    if (dims[i] < 1) {
      set_error(GG_ERR_INVALID, "layer widths must be positive");
      return nullptr;
    }
  if (gg_init(-1) != GG_OK) return nullptr;
  gg_mlp *M = new gg_mlp();
  M->in_dim = dims[0];
  M->in_p = (dims[0] + 7) & ~7;  // 16-byte BF16 rows for the GEMM's vector loads
  size_t off = 0;
  int prev_p = M->in_p;
  M->maxw = M->in_p;
  for (int i = 0; i < n_layers; ++i) {
    Layer L;
    L.in = prev_p;  // the previous layer's padded width (padding columns stay zero)
    L.out = dims[i + 1];
    L.out_p = pad32(dims[i + 1]);
    L.w_off = off;
    off += (size_t)L.in * L.out_p;
    L.b_off = off;
    off += L.out_p;
    M->layers.push_back(L);
    M->maxw = std::max(M->maxw, L.out_p);
    prev_p = L.out_p;
  }
  M->n_params = (int)off;
  std::vector<float> host(off, 0.0f);
  // He-uniform initialisation from a splitmix64 stream; padded rows/columns zero.
  uint64_t st = seed;
  auto rnd = [&]() {
    st += 0x9E3779B97F4A7C15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) / 9007199254740992.0;
  };
  int real_in = dims[0];
  for (int i = 0; i < n_layers; ++i) {
    const Layer &L = M->layers[i];
    const double lim = std::sqrt(6.0 / real_in);
    for (int r = 0; r < real_in; ++r)
      for (int c = 0; c < L.out; ++c) host[L.w_off + (size_t)r * L.out_p + c] = (float)((2 * rnd() - 1) * lim);
    real_in = L.out;
  }
  bool ok = hipStreamCreateWithFlags(&M->stream, hipStreamNonBlocking) == hipSuccess &&
            hipMalloc(&M->p, off * 4) == hipSuccess && hipMalloc(&M->g, off * 4) == hipSuccess &&
            hipMalloc(&M->m1, off * 4) == hipSuccess && hipMalloc(&M->m2, off * 4) == hipSuccess &&
            hipMalloc(&M->pb, off * 2) == hipSuccess && hipMalloc(&M->loss, 4) == hipSuccess;
  if (!ok) {
    set_error(GG_ERR_HIP, "MLP allocation failed");
    gg_mlp_destroy(M);
    return nullptr;
  }
  if (hipMemcpy(M->p, host.data(), off * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(M->m1, 0, off * 4) != hipSuccess || hipMemset(M->m2, 0, off * 4) != hipSuccess ||
      hipMemset(M->g, 0, off * 4) != hipSuccess) {
    set_error(GG_ERR_HIP, "MLP parameter upload failed");
    gg_mlp_destroy(M);
    return nullptr;
  }
  if (refresh_bf16(M) || hipStreamSynchronize(M->stream) != hipSuccess) {
    gg_mlp_destroy(M);
    return nullptr;
  }
  return M;
}

void gg_mlp_destroy(gg_mlp *M) {
  if (!M) return;
  // Release path: errors are not reported.
  if (M->stream) (void)hipStreamSynchronize(M->stream);
  for (float *z : M->Z) (void)hipFree(z);
  for (uint16_t *h : M->Hb) (void)hipFree(h);
  for (void *q : {(void *)M->p, (void *)M->g, (void *)M->m1, (void *)M->m2, (void *)M->pb,
                  (void *)M->x_b, (void *)M->dA, (void *)M->dB, (void *)M->dAb, (void *)M->x_f,
                  (void *)M->t_f, (void *)M->loss, (void *)M->y_f})
    if (q) (void)hipFree(q);
  if (M->stream) (void)hipStreamDestroy(M->stream);
  delete M;
}

int gg_mlp_param_count(const gg_mlp *M) { return M ? M->n_params : 0; }

int gg_mlp_layout(const gg_mlp *M, int *rows, int *cols_padded, int *w_off, int *b_off) {
  if (!M) return set_error(GG_ERR_INVALID, "null MLP");
  for (size_t i = 0; i < M->layers.size(); ++i) {
    rows[i] = M->layers[i].in;
    cols_padded[i] = M->layers[i].out_p;
    w_off[i] = (int)M->layers[i].w_off;
    b_off[i] = (int)M->layers[i].b_off;
  }
  return (int)M->layers.size();
}

int gg_mlp_get_params(gg_mlp *M, float *out) {
  if (!M || !out) return set_error(GG_ERR_INVALID, "bad arguments");
  MLP_HIP(hipMemcpyAsync(out, M->p, (size_t)M->n_params * 4, hipMemcpyDeviceToHost, M->stream));
  MLP_HIP(hipStreamSynchronize(M->stream));
  return GG_OK;
}

int gg_mlp_set_params(gg_mlp *M, const float *in) {
  if (!M || !in) return set_error(GG_ERR_INVALID, "bad arguments");
  MLP_HIP(hipMemcpyAsync(M->p, in, (size_t)M->n_params * 4, hipMemcpyHostToDevice, M->stream));
  int rc = refresh_bf16(M);
  if (rc) return rc;
  MLP_HIP(hipStreamSynchronize(M->stream));
  return GG_OK;
}

int gg_mlp_forward(gg_mlp *M, const float *x, int m, float *y) {
  if (!M || !x || !y || m < 1) return set_error(GG_ERR_INVALID, "bad arguments");
  int rc = mlp_forward(M, x, m, false);
  if (rc) return rc;
  return copy_predictions(M, m, y);
}

int gg_mlp_gradient(gg_mlp *M, const float *x, const float *t, int m, float *loss, float *grad) {
  if (!M || !x || !t || m < 1) return set_error(GG_ERR_INVALID, "bad arguments");
  return mlp_gradient(M, x, false, t, m, loss, grad);
}

int gg_mlp_train_step(gg_mlp *M, const float *x, const float *t, int m, float lr, float *loss) {
  if (!M || !x || !t || m < 1) return set_error(GG_ERR_INVALID, "bad arguments");
  return mlp_train(M, x, false, t, m, lr, loss);
}

}  // extern "C"

namespace gg {

hipStream_t mlp_stream(gg_mlp *M) { return M ? M->stream : nullptr; }

int mlp_forward_device(gg_mlp *M, const float *x_dev, int m, float *y_host) {
  if (!M || !x_dev || !y_host || m < 1) return set_error(GG_ERR_INVALID, "bad arguments");
  int rc = mlp_forward(M, x_dev, m, true);
  if (rc) return rc;
  return copy_predictions(M, m, y_host);
}

int mlp_train_device(gg_mlp *M, const float *x_dev, const float *t_host, int m, float lr,
                     float *loss) {
  if (!M || !x_dev || !t_host || m < 1) return set_error(GG_ERR_INVALID, "bad arguments");
  return mlp_train(M, x_dev, true, t_host, m, lr, loss);
}

}  // namespace gg

