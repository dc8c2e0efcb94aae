// Hand-written HIP kernels of the learned cost model: forward, backward
// (mean-squared-error) and Adam for a small fully connected ReLU network,
// for one network or an ensemble of networks trained side by side (the
// model index is blockIdx.y; every model has its own parameters, Adam
// moments, activations and gradients, and all read the same inputs).
// Shapes are tiny (a batch of a few thousand schedules, widths <= 128).
// Every sum runs in a fixed order, so results are deterministic run to run
// and a model's arithmetic does not depend on how many are trained with it.
#include <hip/hip_runtime.h>
#include <vector>

#include "pmx_rt.h"

namespace {

// batch rows are split into this many chunks for the weight gradients;
// the chunk partial sums are added in chunk order
constexpr int kSplit = 8;

__global__ void k_dense_fwd(const float* __restrict__ W, const float* __restrict__ b,
                            const float* __restrict__ X, float* __restrict__ Y,
                            int batch, int din, int dout, int relu,
                            long sP, long sX, long sY) {
  const long mi = blockIdx.y;
  W += mi * sP; b += mi * sP; X += mi * sX; Y += mi * sY;
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
                            float* __restrict__ D, int n, float inv_batch, long sY, long sD) {
  const long mi = blockIdx.y;
  Y += mi * sY; D += mi * sD;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  D[t] = (Y[t] - T[t]) * inv_batch;
}

// loss = sum((y-t)^2)/2/batch, one block per model, fixed-order tree.
__global__ void k_loss(const float* __restrict__ Y, const float* __restrict__ T, int n,
                       float inv_batch, float* __restrict__ loss, long sY, long sL) {
  __shared__ float s[256];
  const long mi = blockIdx.y;
  Y += mi * sY;
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
  if (threadIdx.x == 0) loss[mi * sL] = s[0] * inv_batch;
}

// Partial weight and bias gradients of one output unit o over one chunk c
// of the batch rows: P[c][o][i] = sum_{r in chunk c} D[r,o] * A[r,i] for
// i < din and sum D[r,o] for i == din (the bias).  256 threads: 64 lanes
// over i times 4 row groups (rows r0+g, r0+g+4, ...), the 4 group sums
// added in group order.  Grid (dout * kSplit, nmodels).
__global__ void k_grad_part(const float* __restrict__ D, const float* __restrict__ A,
                            float* __restrict__ P, int batch, int din, int dout,
                            long sD, long sA, long sPart) {
  __shared__ float s[4][64];
  const long mi = blockIdx.y;
  D += mi * sD; A += mi * sA; P += mi * sPart;
  const int o = blockIdx.x / kSplit, c = blockIdx.x - o * kSplit;
  const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int r0 = (int)((long)batch * c / kSplit), r1 = (int)((long)batch * (c + 1) / kSplit);
  float* out = P + ((long)c * dout + o) * (din + 1);
  for (int i0 = 0; i0 <= din; i0 += 64) {
    const int i = i0 + lane;
    float acc = 0.f;
    if (i <= din) {
      for (int r = r0 + g; r < r1; r += 4) {
        const float d = D[(size_t)r * dout + o];
        acc += d * (i < din ? A[(size_t)r * din + i] : 1.f);
      }
    }
    s[g][lane] = acc;
    __syncthreads();
    if (g == 0 && i <= din) out[i] = ((s[0][lane] + s[1][lane]) + s[2][lane]) + s[3][lane];
    __syncthreads();
  }
}

// dW[o,i] / db[o] = chunk partials added in chunk order.
__global__ void k_grad_sum(const float* __restrict__ P, float* __restrict__ dW,
                           float* __restrict__ db, int din, int dout, long sPart, long sG) {
  const long mi = blockIdx.y;
  P += mi * sPart; dW += mi * sG; db += mi * sG;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = (din + 1) * dout;
  if (t >= n) return;
  float acc = 0.f;
  for (int c = 0; c < kSplit; ++c) acc += P[(long)c * n + t];
  const int o = t / (din + 1), i = t - o * (din + 1);
  if (i < din) dW[o * din + i] = acc;
  else db[o] = acc;
}

// Dp[b,i] = sum_o W[o,i] * D[b,o], masked by ReLU'(A[b,i])
__global__ void k_back_delta(const float* __restrict__ W, const float* __restrict__ D,
                             const float* __restrict__ A, float* __restrict__ Dp,
                             int batch, int din, int dout, long sP, long sD, long sA) {
  const long mi = blockIdx.y;
  W += mi * sP; D += mi * sD; A += mi * sA; Dp += mi * sD;
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
  const long off = (long)blockIdx.y * n;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  p += off; g += off; m += off; v += off;
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

int param_count(const int* dims, int nlayers) {
  long n = 0;
  for (int l = 0; l < nlayers; ++l) n += (long)dims[l + 1] * dims[l] + dims[l + 1];
  return (int)n;
}

// activations of layers 1..nlayers of one model, [batch, dims[l]] each
long hidden_size(const int* dims, int nlayers, int batch) {
  long n = 0;
  for (int l = 1; l <= nlayers; ++l) n += (long)batch * dims[l];
  return n;
}

long hidden_offset(const int* dims, int l, int batch) {  // l >= 1
  long off = 0;
  for (int k = 1; k < l; ++k) off += (long)batch * dims[k];
  return off;
}

// Layer-l input: x0 (layer 0, shared by all models: stride sX0) or the
// models' hidden activations.
struct Acts {
  const float* x0; long sX0; float* h; long sH; int batch; const int* dims;
  const float* in(int l) const { return l == 0 ? x0 : h + hidden_offset(dims, l, batch); }
  long stride(int l) const { return l == 0 ? sX0 : sH; }
  float* out(int l) const { return h + hidden_offset(dims, l, batch); }  // l >= 1
};

int forward_impl(const float* params, const int* dims, int nlayers, int nm, const Acts& a,
                 hipStream_t s) {
  const long sP = param_count(dims, nlayers);
  const float* p = params;
  for (int l = 0; l < nlayers; ++l) {
    int din = dims[l], dout = dims[l + 1];
    const float* W = p;
    const float* b = p + (long)din * dout;
    p = b + dout;
    hipLaunchKernelGGL(k_dense_fwd, dim3(nblk((long)a.batch * dout), nm), dim3(256), 0, s, W, b,
                       a.in(l), a.out(l + 1), a.batch, din, dout, l + 1 < nlayers ? 1 : 0, sP,
                       a.stride(l), a.sH);
  }
  return hip_ok(hipGetLastError(), "mlp: forward launch");
}

int backward_impl(const float* params, const int* dims, int nlayers, int nm, const Acts& a,
                  const float* target, float* grads, float* loss, long sLoss, hipStream_t s) {
  const int batch = a.batch;
  int maxw = 0;
  for (int l = 0; l <= nlayers; ++l) maxw = dims[l] > maxw ? dims[l] : maxw;
  const long sP = param_count(dims, nlayers);
  long spart = 0;
  for (int l = 0; l < nlayers; ++l) {
    long n = (long)kSplit * (dims[l] + 1) * dims[l + 1];
    spart = n > spart ? n : spart;
  }
  const long sD = (long)batch * maxw;
  void* scratch = nullptr;
  int rc = pmx_arena_get("pmx_mlp_delta", sizeof(float) * (size_t)nm * (2 * sD + spart), &scratch);
  if (rc) return rc;
  float* D = (float*)scratch;
  float* Dp = D + (size_t)nm * sD;
  float* P = Dp + (size_t)nm * sD;
  std::vector<long> woff(nlayers), boff(nlayers);
  long off = 0;
  for (int l = 0; l < nlayers; ++l) {
    woff[l] = off;
    off += (long)dims[l] * dims[l + 1];
    boff[l] = off;
    off += dims[l + 1];
  }
  const int L = nlayers;
  const float* Y = a.out(L);
  const long nout = (long)batch * dims[L];
  const float inv = 1.f / (float)batch;
  hipLaunchKernelGGL(k_loss, dim3(1, nm), dim3(256), 0, s, Y, target, (int)nout, inv, loss, a.sH,
                     sLoss);
  hipLaunchKernelGGL(k_out_delta, dim3(nblk(nout), nm), dim3(256), 0, s, Y, target, D, (int)nout,
                     inv, a.sH, sD);
  for (int l = L - 1; l >= 0; --l) {
    const int din = dims[l], dout = dims[l + 1];
    const float* A = a.in(l);
    hipLaunchKernelGGL(k_grad_part, dim3(dout * kSplit, nm), dim3(256), 0, s, D, A, P, batch, din,
                       dout, sD, a.stride(l), spart);
    hipLaunchKernelGGL(k_grad_sum, dim3(nblk((long)(din + 1) * dout), nm), dim3(256), 0, s, P,
                       grads + woff[l], grads + boff[l], din, dout, spart, sP);
    if (l > 0) {
      hipLaunchKernelGGL(k_back_delta, dim3(nblk((long)batch * din), nm), dim3(256), 0, s,
                         params + woff[l], D, A, Dp, batch, din, dout, sP, sD, a.stride(l));
      float* t = D;
      D = Dp;
      Dp = t;
    }
  }
  return hip_ok(hipGetLastError(), "mlp: backward launch");
}

int adam_impl(float* params, const float* grads, float* m, float* v, int count, int nm, float lr,
              float beta1, float beta2, float eps, int step, hipStream_t s) {
  float c1 = 1.f - powf(beta1, (float)step);
  float c2 = 1.f - powf(beta2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(nblk(count), nm), dim3(256), 0, s, params, grads, m, v, count,
                     lr, beta1, beta2, eps, c1, c2);
  return hip_ok(hipGetLastError(), "mlp: adam launch");
}

int train_impl(float* params, const int* dims, int nlayers, int nm, const Acts& a,
               const float* target, int epochs, float lr, float beta1, float beta2, float eps,
               float* m, float* v, float* grads, float* loss, long sLoss, int step0,
               hipStream_t s) {
  const int count = param_count(dims, nlayers);
  for (int e = 0; e < epochs; ++e) {
    int rc = forward_impl(params, dims, nlayers, nm, a, s);
    if (rc) return rc;
    rc = backward_impl(params, dims, nlayers, nm, a, target, grads, loss + e, sLoss, s);
    if (rc) return rc;
    rc = adam_impl(params, grads, m, v, count, nm, lr, beta1, beta2, eps, step0 + e + 1, s);
    if (rc) return rc;
  }
  return 0;
}

Acts single_acts(const int* dims, float* acts, int batch) {
  // public single-model layout: layer 0 (a copy of x) then layers 1..L
  Acts a;
  a.x0 = acts;
  a.sX0 = 0;
  a.h = acts + (long)batch * dims[0];
  a.sH = 0;
  a.batch = batch;
  a.dims = dims;
  return a;
}

}  // namespace

extern "C" {

int pmx_mlp_param_count(const int* dims, int nlayers, int* count) {
  *count = param_count(dims, nlayers);
  return 0;
}

int pmx_mlp_forward(const float* params, const int* dims, int nlayers, const float* x,
                    int batch, float* acts, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  int rc = hip_ok(hipMemcpyAsync(acts, x, sizeof(float) * (size_t)batch * dims[0],
                                 hipMemcpyDeviceToDevice, s), "mlp: copy input");
  if (rc) return rc;
  return forward_impl(params, dims, nlayers, 1, single_acts(dims, acts, batch), s);
}

int pmx_mlp_backward(const float* params, const int* dims, int nlayers, const float* acts,
                     const float* target, int batch, float* grads, float* loss, void* stream) {
  return backward_impl(params, dims, nlayers, 1, single_acts(dims, (float*)acts, batch), target,
                       grads, loss, 0, (hipStream_t)stream);
}

int pmx_mlp_adam(float* params, const float* grads, float* m, float* v, int count, float lr,
                 float beta1, float beta2, float eps, int step, void* stream) {
  return adam_impl(params, grads, m, v, count, 1, lr, beta1, beta2, eps, step,
                   (hipStream_t)stream);
}

// Full-batch training loop in one call: forward, MSE backward and Adam for
// `epochs` epochs, launched back to back on `stream` (no host round trip per
// epoch).  loss[e] receives epoch e's loss; Adam's step count starts at
// step0 + 1.
int pmx_mlp_train(float* params, const int* dims, int nlayers, const float* x,
                  const float* target, int batch, int epochs, float lr, float beta1,
                  float beta2, float eps, float* m, float* v, float* acts, float* grads,
                  float* loss, int step0, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  int rc = hip_ok(hipMemcpyAsync(acts, x, sizeof(float) * (size_t)batch * dims[0],
                                 hipMemcpyDeviceToDevice, s), "mlp: copy input");
  if (rc) return rc;
  return train_impl(params, dims, nlayers, 1, single_acts(dims, acts, batch), target, epochs, lr,
                    beta1, beta2, eps, m, v, grads, loss, 0, step0, s);
}

int pmx_mlp_train_ensemble(float* params, const int* dims, int nlayers, int nmodels,
                           const float* x, const float* target, int batch, int epochs, float lr,
                           float beta1, float beta2, float eps, float* m, float* v, float* hidden,
                           float* grads, float* loss, int step0, void* stream) {
  if (nmodels <= 0 || batch <= 0 || epochs < 0) {
    pmx_set_error("mlp: bad ensemble shape");
    return -22;
  }
  Acts a;
  a.x0 = x;
  a.sX0 = 0;
  a.h = hidden;
  a.sH = hidden_size(dims, nlayers, batch);
  a.batch = batch;
  a.dims = dims;
  return train_impl(params, dims, nlayers, nmodels, a, target, epochs, lr, beta1, beta2, eps, m,
                    v, grads, loss, epochs, step0, (hipStream_t)stream);
}

}  // extern "C"
