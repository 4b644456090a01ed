// gg_device.h -- HIP-side internals shared by the .hip translation units
// (streams and device-pointer entry points that the C ABI does not expose).
#ifndef GG_DEVICE_H
#define GG_DEVICE_H
#include <hip/hip_runtime.h>

#include <cstdint>

struct gg_mlp;
struct gg_incr;

struct gg_graph;
struct gg_result;

namespace gg {
// Selects the device and creates the evaluator's stream (idempotent).
int init(int device);
// The LDS tile path: `count` graphs (host or device adjacency pointers),
// optionally with per-vertex distance histograms.
int eval_host(int count, const gg_graph *gs, gg_result *out, bool device_ptrs, int hist_maxd,
              int32_t *hist_host);
// The HBM-resident path (gg_wide.hip): any n; sem 0 = matrix_op, 1 = the
// reference's bfs() semantics (out->sum = its weighted sum over pairs s < i).
int wide_eval(const gg_graph &g, int sem, bool device_ptrs, gg_result *out);
int wide_release();
// The evaluator's stream (adjacency uploads, plans, histograms run on it).
hipStream_t engine_stream();
// Launch shape of one tile-BFS launch of `bits`-wide words over an n-vertex
// graph: threads per block, vertices per thread, reach-array capacity, queue
// capacity and dynamic LDS bytes (the evaluator's own choice for that n).
int tile_shape(int bits, int n, int *block, int *vpt, int *n_cap, int *qcap, size_t *lds);
// Bank-aware order (the evaluator's gg_order_rows_kernel) of one converted
// row table [n][dp] on `st`, for reach words of `bits` bits; a no-op where the
// evaluator would not order (64-bit words, n below its threshold, rows of
// more than four chunks, or ordering turned off).  Waits for completion.
int order_rows_sync(uint16_t *adjR, int n, int dp, int bits, hipStream_t st);
// Whether the row-order mode orders a table of n rows of dp entries.
bool order_rows_applies(int n, int dp);
// gg_incr_create without the graph checks, for callers that hold a simple
// regular graph already (the batched SA validates its graph on entry).
gg_incr *incr_create_unchecked(int nodes, int degree, const int32_t *adj);
// Forward pass on m rows already in device memory (row-major, in_dim floats
// each); copies the m predictions to y_host and waits for them.
int mlp_forward_device(gg_mlp *M, const float *x_dev, int m, float *y_host);
// The stream the network's kernels run on.
hipStream_t mlp_stream(gg_mlp *M);
// One Adam step on m device-resident rows with host targets t_host.
int mlp_train_device(gg_mlp *M, const float *x_dev, const float *t_host, int m, float lr,
                     float *loss);
}  // namespace gg
#endif
