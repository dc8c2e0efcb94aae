// Hand-written HIP kernels of the learned cost model: forward, backward
// (mean-squared-error) and Adam for a small fully connected ReLU network.
// Shapes are tiny (batch of a few thousand schedules, widths <= 128), so
// each kernel maps one thread to one output element and sums in a fixed
// order: results are deterministic run to run.
#include <hip/hip_runtime.h>
#include <vector>

#include "pmx_rt.h"

namespace {

__global__ void k_dense_fwd(const float* __restrict__ W, const float* __restrict__ b,
                            const float* __restrict__ X, float* __restrict__ Y,
                            int batch, int din, int dout, int relu) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= batch * dout) return;
  int r = t / dout, o = t - r * dout;
  const float* x = X + (size_t)r * din;
  const float* w = W + (size_t)o * din;
  float acc = b[o];
  for (int i = 0; i < din; ++i) acc += w[i] * x[i];
  if (relu && acc < 0.f) acc = 0.f;
  Y[t] = acc;
}

// delta = (y - t) / batch
__global__ void k_out_delta(const float* __restrict__ Y, const float* __restrict__ T,
                            float* __restrict__ D, int n, float inv_batch) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  D[t] = (Y[t] - T[t]) * inv_batch;
}

// loss = sum((y-t)^2)/2/batch, one block, fixed-order tree reduction.
__global__ void k_loss(const float* __restrict__ Y, const float* __restrict__ T, int n,
                       float inv_batch, float* __restrict__ loss) {
  __shared__ float s[256];
  float acc = 0.f;
  for (int i = threadIdx.x; i < n; i += 256) {
    float d = Y[i] - T[i];
    acc += 0.5f * d * d;
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[0] = s[0] * inv_batch;
}

// dW[o,i] = sum_b D[b,o] * A[b,i]
__global__ void k_grad_w(const float* __restrict__ D, const float* __restrict__ A,
                         float* __restrict__ dW, int batch, int din, int dout) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= din * dout) return;
  int o = t / din, i = t - o * din;
  float acc = 0.f;
  for (int r = 0; r < batch; ++r) acc += D[(size_t)r * dout + o] * A[(size_t)r * din + i];
  dW[t] = acc;
}

__global__ void k_grad_b(const float* __restrict__ D, float* __restrict__ db, int batch, int dout) {
  int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= dout) return;
  float acc = 0.f;
  for (int r = 0; r < batch; ++r) acc += D[(size_t)r * dout + o];
  db[o] = acc;
}

// Dp[b,i] = sum_o W[o,i] * D[b,o], masked by ReLU'(A[b,i])
__global__ void k_back_delta(const float* __restrict__ W, const float* __restrict__ D,
                             const float* __restrict__ A, float* __restrict__ Dp,
                             int batch, int din, int dout) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= batch * din) return;
  int r = t / din, i = t - r * din;
  float acc = 0.f;
  for (int o = 0; o < dout; ++o) acc += W[(size_t)o * din + i] * D[(size_t)r * dout + o];
  Dp[t] = A[t] > 0.f ? acc : 0.f;
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  float gi = g[t];
  float mi = b1 * m[t] + (1.f - b1) * gi;
  float vi = b2 * v[t] + (1.f - b2) * gi * gi;
  m[t] = mi;
  v[t] = vi;
  float mh = mi / c1, vh = vi / c2;
  p[t] -= lr * mh / (sqrtf(vh) + eps);
}

inline unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

int hip_ok(hipError_t e, const char* what) {
  if (e == hipSuccess) return 0;
  pmx_set_error(what);
  return -(int)e;
}

}  // namespace

extern "C" {

int pmx_mlp_param_count(const int* dims, int nlayers, int* count) {
  long n = 0;
  for (int l = 0; l < nlayers; ++l) n += (long)dims[l + 1] * dims[l] + dims[l + 1];
  *count = (int)n;
  return 0;
}

static long act_offset(const int* dims, int l, int batch) {
  long off = 0;
  for (int k = 0; k < l; ++k) off += (long)batch * dims[k];
  return off;
}

int pmx_mlp_forward(const float* params, const int* dims, int nlayers, const float* x,
                    int batch, float* acts, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  int rc = hip_ok(hipMemcpyAsync(acts, x, sizeof(float) * (size_t)batch * dims[0],
                                 hipMemcpyDeviceToDevice, s), "mlp: copy input");
  if (rc) return rc;
  const float* p = params;
  for (int l = 0; l < nlayers; ++l) {
    int din = dims[l], dout = dims[l + 1];
    const float* W = p;
    const float* b = p + (long)din * dout;
    p = b + dout;
    float* X = acts + act_offset(dims, l, batch);
    float* Y = acts + act_offset(dims, l + 1, batch);
    hipLaunchKernelGGL(k_dense_fwd, dim3(nblk((long)batch * dout)), dim3(256), 0, s, W, b, X, Y,
                       batch, din, dout, l + 1 < nlayers ? 1 : 0);
  }
  return hip_ok(hipGetLastError(), "mlp: forward launch");
}

int pmx_mlp_backward(const float* params, const int* dims, int nlayers, const float* acts,
                     const float* target, int batch, float* grads, float* loss, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  int maxw = 0;
  for (int l = 0; l <= nlayers; ++l) maxw = dims[l] > maxw ? dims[l] : maxw;
  void* scratch = nullptr;
  int rc = pmx_arena_get("pmx_mlp_delta", sizeof(float) * (size_t)batch * maxw * 2, &scratch);
  if (rc) return rc;
  float* D = (float*)scratch;
  float* Dp = D + (size_t)batch * maxw;
  // parameter offsets per layer
  std::vector<long> woff(nlayers), boff(nlayers);
  long off = 0;
  for (int l = 0; l < nlayers; ++l) {
    woff[l] = off;
    off += (long)dims[l] * dims[l + 1];
    boff[l] = off;
    off += dims[l + 1];
  }
  int L = nlayers;
  const float* Y = acts + act_offset(dims, L, batch);
  long nout = (long)batch * dims[L];
  float inv = 1.f / (float)batch;
  hipLaunchKernelGGL(k_loss, dim3(1), dim3(256), 0, s, Y, target, (int)nout, inv, loss);
  hipLaunchKernelGGL(k_out_delta, dim3(nblk(nout)), dim3(256), 0, s, Y, target, D, (int)nout, inv);
  for (int l = L - 1; l >= 0; --l) {
    int din = dims[l], dout = dims[l + 1];
    const float* A = acts + act_offset(dims, l, batch);
    hipLaunchKernelGGL(k_grad_w, dim3(nblk((long)din * dout)), dim3(256), 0, s, D, A,
                       grads + woff[l], batch, din, dout);
    hipLaunchKernelGGL(k_grad_b, dim3(nblk(dout)), dim3(256), 0, s, D, grads + boff[l], batch, dout);
    if (l > 0) {
      hipLaunchKernelGGL(k_back_delta, dim3(nblk((long)batch * din)), dim3(256), 0, s,
                         params + woff[l], D, A, Dp, batch, din, dout);
      float* t = D;
      D = Dp;
      Dp = t;
    }
  }
  return hip_ok(hipGetLastError(), "mlp: backward launch");
}

int pmx_mlp_adam(float* params, const float* grads, float* m, float* v, int count, float lr,
                 float beta1, float beta2, float eps, int step, void* stream) {
  float c1 = 1.f - powf(beta1, (float)step);
  float c2 = 1.f - powf(beta2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(nblk(count)), dim3(256), 0, (hipStream_t)stream, params, grads,
                     m, v, count, lr, beta1, beta2, eps, c1, c2);
  return hip_ok(hipGetLastError(), "mlp: adam launch");
}

// Full-batch training loop in one call: forward, MSE backward and Adam for
// `epochs` epochs, launched back to back on `stream` (no host round trip per
// epoch).  loss[e] receives epoch e's loss; Adam's step count starts at
// step0 + 1.
int pmx_mlp_train(float* params, const int* dims, int nlayers, const float* x,
                  const float* target, int batch, int epochs, float lr, float beta1,
                  float beta2, float eps, float* m, float* v, float* acts, float* grads,
                  float* loss, int step0, void* stream) {
  int count = 0;
  int rc = pmx_mlp_param_count(dims, nlayers, &count);
  if (rc) return rc;
  for (int e = 0; e < epochs; ++e) {
    rc = pmx_mlp_forward(params, dims, nlayers, x, batch, acts, stream);
    if (rc) return rc;
    rc = pmx_mlp_backward(params, dims, nlayers, acts, target, batch, grads, loss + e, stream);
    if (rc) return rc;
    rc = pmx_mlp_adam(params, grads, m, v, count, lr, beta1, beta2, eps, step0 + e + 1, stream);
    if (rc) return rc;
  }
  return 0;
}

}  // extern "C"
