// Training and inference kernels of the variant selector's MLP (gfx950).
//
// The selector is a small fully connected network (features -> 64 -> 64 ->
// one predicted log-time per SpMM variant).  Its batches are a few thousand
// rows by at most a few hundred columns, so each kernel maps one output
// element to one thread and loops over the reduction dimension; weights are
// row-major [in_dim x out_dim] and y = x W + b.
#include <string>

#include "dtc_internal.h"

namespace dtc {
namespace {

// y[b][o] = act(sum_i x[b][i] w[i][o] + bias[o])
__global__ __launch_bounds__(256) void k_linear_fwd(const float *x, const float *w, const float *bias,
                                                    float *y, int batch, int in_dim, int out_dim,
                                                    int relu) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)batch * out_dim) return;
  const int b = (int)(t / out_dim);
  const int o = (int)(t % out_dim);
  const float *xr = x + (int64_t)b * in_dim;
  float acc = bias ? bias[o] : 0.f;
  for (int i = 0; i < in_dim; ++i) acc = fmaf(xr[i], w[(int64_t)i * out_dim + o], acc);
  y[t] = (relu && acc < 0.f) ? 0.f : acc;
}

// dz = dy * act'(y): for ReLU the derivative is (y > 0).
__device__ __forceinline__ float dz_at(const float *dy, const float *y, int64_t idx, int relu) {
  const float d = dy[idx];
  return (relu && !(y[idx] > 0.f)) ? 0.f : d;
}

// dw[i][o] = sum_b x[b][i] dz[b][o]
__global__ __launch_bounds__(256) void k_linear_bwd_w(const float *x, const float *y, const float *dy,
                                                      float *dw, int batch, int in_dim, int out_dim,
                                                      int relu) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)in_dim * out_dim) return;
  const int i = (int)(t / out_dim);
  const int o = (int)(t % out_dim);
  float acc = 0.f;
  for (int b = 0; b < batch; ++b)
    acc = fmaf(x[(int64_t)b * in_dim + i], dz_at(dy, y, (int64_t)b * out_dim + o, relu), acc);
  dw[t] = acc;
}

// db[o] = sum_b dz[b][o]
__global__ __launch_bounds__(256) void k_linear_bwd_b(const float *y, const float *dy, float *db,
                                                      int batch, int out_dim, int relu) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= out_dim) return;
  float acc = 0.f;
  for (int b = 0; b < batch; ++b) acc += dz_at(dy, y, (int64_t)b * out_dim + o, relu);
  db[o] = acc;
}

// dx[b][i] = sum_o dz[b][o] w[i][o]
__global__ __launch_bounds__(256) void k_linear_bwd_x(const float *w, const float *y, const float *dy,
                                                      float *dx, int batch, int in_dim, int out_dim,
                                                      int relu) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)batch * in_dim) return;
  const int b = (int)(t / in_dim);
  const int i = (int)(t % in_dim);
  float acc = 0.f;
  for (int o = 0; o < out_dim; ++o)
    acc = fmaf(dz_at(dy, y, (int64_t)b * out_dim + o, relu), w[(int64_t)i * out_dim + o], acc);
  dx[t] = acc;
}

__global__ void k_zero1(float *p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0.f;
}

// Masked mean squared error: loss = inv_count * sum mask (pred - target)^2,
// dpred = 2 inv_count mask (pred - target).
__global__ __launch_bounds__(256) void k_mse(const float *pred, const float *target, const float *mask,
                                             int64_t n, float inv_count, float *dpred, float *loss) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (t < n) {
    const float d = (pred[t] - target[t]) * mask[t];
    dpred[t] = 2.f * inv_count * d;
    l = inv_count * d * d;
  }
#pragma unroll
  for (int s = 32; s >= 1; s >>= 1) l += __shfl_xor(l, s, 64);
  if ((threadIdx.x & 63) == 0 && l != 0.f) atomicAdd(loss, l);
}

// Adam (Kingma & Ba, 2015) with bias correction.
__global__ __launch_bounds__(256) void k_adam(float *p, const float *g, float *m, float *v, int64_t n,
                                              float lr, float beta1, float beta2, float eps,
                                              float bc1, float bc2) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float gt = g[t];
  const float mt = beta1 * m[t] + (1.f - beta1) * gt;
  const float vt = beta2 * v[t] + (1.f - beta2) * gt * gt;
  m[t] = mt;
  v[t] = vt;
  p[t] -= lr * (mt / bc1) / (sqrtf(vt / bc2) + eps);
}

unsigned blocks_for(int64_t n) { return (unsigned)std::max<int64_t>(1, ceil_div(n, 256)); }

dtc_status_t check_dims(int batch, int in_dim, int out_dim) {
  if (batch < 0 || in_dim <= 0 || out_dim <= 0)
    return fail(DTC_ERROR_INVALID_VALUE, "batch >= 0, in_dim >= 1 and out_dim >= 1 required");
  return DTC_SUCCESS;
}

}  // namespace
}  // namespace dtc

using namespace dtc;

extern "C" {

dtc_status_t dtc_mlp_linear_forward(const float *x, const float *w, const float *bias, float *y,
                                    int32_t batch, int32_t in_dim, int32_t out_dim, int32_t relu,
                                    dtc_stream_t stream_) {
  clear_error();
  dtc_status_t st = check_dims(batch, in_dim, out_dim);
  if (st != DTC_SUCCESS) return st;
  if (batch == 0) return DTC_SUCCESS;
  if (!x || !w || !y) return fail(DTC_ERROR_INVALID_VALUE, "NULL argument");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream_);
  k_linear_fwd<<<blocks_for((int64_t)batch * out_dim), 256, 0, s>>>(x, w, bias, y, batch, in_dim,
                                                                    out_dim, relu);
  DTC_LAUNCH_CHECK("mlp_linear_forward");
  return DTC_SUCCESS;
}

dtc_status_t dtc_mlp_linear_backward(const float *x, const float *w, const float *y, const float *dy,
                                     float *dx, float *dw, float *db, int32_t batch, int32_t in_dim,
                                     int32_t out_dim, int32_t relu, dtc_stream_t stream_) {
  clear_error();
  dtc_status_t st = check_dims(batch, in_dim, out_dim);
  if (st != DTC_SUCCESS) return st;
  if (!x || !w || !y || !dy || !dw || !db) return fail(DTC_ERROR_INVALID_VALUE, "NULL argument");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream_);
  k_linear_bwd_w<<<blocks_for((int64_t)in_dim * out_dim), 256, 0, s>>>(x, y, dy, dw, batch, in_dim,
                                                                       out_dim, relu);
  DTC_LAUNCH_CHECK("mlp_linear_backward_w");
  k_linear_bwd_b<<<blocks_for(out_dim), 256, 0, s>>>(y, dy, db, batch, out_dim, relu);
  DTC_LAUNCH_CHECK("mlp_linear_backward_b");
  if (dx && batch > 0) {
    k_linear_bwd_x<<<blocks_for((int64_t)batch * in_dim), 256, 0, s>>>(w, y, dy, dx, batch, in_dim,
                                                                       out_dim, relu);
    DTC_LAUNCH_CHECK("mlp_linear_backward_x");
  }
  return DTC_SUCCESS;
}

dtc_status_t dtc_mlp_mse(const float *pred, const float *target, const float *mask, int64_t n,
                         float inv_count, float *dpred, float *loss, dtc_stream_t stream_) {
  clear_error();
  if (n < 0) return fail(DTC_ERROR_INVALID_VALUE, "n must be >= 0");
  if (!loss) return fail(DTC_ERROR_INVALID_VALUE, "loss is NULL");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream_);
  k_zero1<<<1, 64, 0, s>>>(loss);
  DTC_LAUNCH_CHECK("mlp_mse_zero");
  if (n == 0) return DTC_SUCCESS;
  if (!pred || !target || !mask || !dpred) return fail(DTC_ERROR_INVALID_VALUE, "NULL argument");
  k_mse<<<blocks_for(n), 256, 0, s>>>(pred, target, mask, n, inv_count, dpred, loss);
  DTC_LAUNCH_CHECK("mlp_mse");
  return DTC_SUCCESS;
}

dtc_status_t dtc_adam_step(float *param, const float *grad, float *m, float *v, int64_t n, float lr,
                           float beta1, float beta2, float eps, int32_t step, dtc_stream_t stream_) {
  clear_error();
  if (n < 0 || step < 1) return fail(DTC_ERROR_INVALID_VALUE, "n >= 0 and step >= 1 required");
  if (n == 0) return DTC_SUCCESS;
  if (!param || !grad || !m || !v) return fail(DTC_ERROR_INVALID_VALUE, "NULL argument");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream_);
  const float bc1 = 1.f - powf(beta1, (float)step);
  const float bc2 = 1.f - powf(beta2, (float)step);
  k_adam<<<blocks_for(n), 256, 0, s>>>(param, grad, m, v, n, lr, beta1, beta2, eps, bc1, bc2);
  DTC_LAUNCH_CHECK("adam_step");
  return DTC_SUCCESS;
}

}  // extern "C"
