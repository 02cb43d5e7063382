// Learned log-delta proposal: spectral features of a rotated column and a small FP32 MLP, with the forward,
// backward and Adam kernels that train it on the GPU.
//
// Features (34 per column, on the projected basis whose eigenvalues s are ascending):
//   f_b  (b < 32) = log(mean of y_k^2 over the k whose eigenvalue falls in bin b) - log(mean of y_k^2 over all k),
//                   bins log-spaced in s over [edges[0], edges[32]) (s outside is clamped into the end bins);
//                   an empty bin contributes 0;
//   f_32 = log(N) / 10, f_33 = C / 10.
// The features are invariant to the scale of y, as the optimal delta is.
//
// Network: x (34) -> tanh(W1 x + b1) (128) -> tanh(W2 h1 + b2) (128) -> w3 . h2 + b3, the proposed log delta.
// Parameters live in one flat fp32 buffer: W1 (128 x 34), b1 (128), W2 (128 x 128), b2 (128), w3 (128), b3 (1).
// Every layer is one kernel of its own (thread per output element); a batch is rows of a row-major matrix.
#include "kernels.h"

namespace flmm {

constexpr int NBIN = 32;

// one wave per column: bin sums of y^2 in a fixed order, then the 34 features.  y^2 is staged through LDS 1,024
// elements at a time and lane b < 32 adds the chunk's elements of bin b in index order, so the features are
// bit-reproducible from run to run.
__global__ __launch_bounds__(256) void mlp_features_kernel(const float* __restrict__ uy, int64_t P, int64_t p_pad,
                                                           int64_t neff, const int* __restrict__ bin_of_k,
                                                           float logN, float C, float* __restrict__ feat) {
  constexpr int CH = 1024;
  __shared__ float acc[4][NBIN];
  __shared__ int cnt[4][NBIN];
  __shared__ float y2s[4][CH];
  __shared__ int bins[CH];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t p = (int64_t)blockIdx.x * 4 + w;
  float tot = 0.f, my_acc = 0.f;
  int my_cnt = 0;
  for (int64_t k0 = 0; k0 < neff; k0 += CH) {
    const int n = neff - k0 < CH ? (int)(neff - k0) : CH;
    __syncthreads();  // the previous chunk has been consumed
    for (int i = threadIdx.x; i < n; i += blockDim.x) bins[i] = bin_of_k[k0 + i];
    if (p < P) {
      for (int i = lane; i < n; i += 64) {
        const int64_t k = k0 + i;
        const float y = uy[((k >> 5) * p_pad + p) * 32 + (k & 31)];
        y2s[w][i] = y * y;
        tot += y * y;
      }
    }
    __syncthreads();
    if (p < P && lane < NBIN) {
      for (int i = 0; i < n; ++i) {
        if (bins[i] == lane) {
          my_acc += y2s[w][i];
          ++my_cnt;
        }
      }
    }
  }
  if (lane < NBIN) { acc[w][lane] = my_acc; cnt[w][lane] = my_cnt; }
  tot = wave_sum(tot);
  __syncthreads();
  // per-bin log(mean y^2 / overall mean); empty bins interpolated linearly (in bin index) between the nearest
  // occupied bins, held flat beyond the occupied range
  __shared__ float val[4][NBIN];
  if (p < P && lane < NBIN) {
    const float mean_all = fmaxf(tot / (float)neff, 1e-30f);
    const int c = cnt[w][lane];
    val[w][lane] = c > 0 ? __logf(fmaxf(acc[w][lane] / (float)c, 1e-30f) / mean_all) : 0.f;
  }
  __syncthreads();
  if (p < P && lane < NBIN) {
    // occupancy depends only on the spectrum, so every wave sees the same pattern
    int lo = -1, hi = -1;
    for (int b = lane; b >= 0; --b)
      if (cnt[w][b] > 0) { lo = b; break; }
    for (int b = lane; b < NBIN; ++b)
      if (cnt[w][b] > 0) { hi = b; break; }
    float v;
    if (lo == lane) v = val[w][lane];
    else if (lo >= 0 && hi >= 0) {
      const float f = (float)(lane - lo) / (float)(hi - lo);
      v = (1.f - f) * val[w][lo] + f * val[w][hi];
    } else if (lo >= 0) v = val[w][lo];
    else if (hi >= 0) v = val[w][hi];
    else v = 0.f;
    feat[p * 34 + lane] = v;
  }
  if (p < P && lane == 0) {
    feat[p * 34 + 32] = logN / 10.f;
    feat[p * 34 + 33] = C / 10.f;
  }
}

// out[b][j] = act(bias[j] + sum_i W[j][i] in[b][i]); act = tanh (act == 1) or identity (act == 0)
__global__ void mlp_dense_fwd_kernel(const float* __restrict__ in, int64_t B, int n_in, const float* __restrict__ W,
                                     const float* __restrict__ bias, int n_out, int act, float* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= B * n_out) return;
  const int64_t b = e / n_out;
  const int j = (int)(e % n_out);
  const float* x = in + b * n_in;
  const float* wr = W + (int64_t)j * n_in;
  float z = bias[j];
  for (int i = 0; i < n_in; ++i) z = fmaf(wr[i], x[i], z);
  out[e] = act ? tanhf(z) : z;
}

// dz[b][j] = dout[b][j] * act'(out[b][j])  (in place on dout)
__global__ void mlp_act_bwd_kernel(float* __restrict__ dout, const float* __restrict__ out, int64_t n, int act) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n || !act) return;
  const float y = out[e];
  dout[e] *= 1.f - y * y;
}

// din[b][i] = sum_j dz[b][j] W[j][i]
__global__ void mlp_dense_bwd_in_kernel(const float* __restrict__ dz, int64_t B, int n_out,
                                        const float* __restrict__ W, int n_in, float* __restrict__ din) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= B * n_in) return;
  const int64_t b = e / n_in;
  const int i = (int)(e % n_in);
  float s = 0.f;
  for (int j = 0; j < n_out; ++j) s = fmaf(dz[b * n_out + j], W[(int64_t)j * n_in + i], s);
  din[e] = s;
}

// dW[j][i] = sum_b dz[b][j] in[b][i] (i < n_in), db[j] = sum_b dz[b][j] (i == n_in); one block per (j, i)
__global__ __launch_bounds__(256) void mlp_dense_bwd_w_kernel(const float* __restrict__ dz, const float* __restrict__ in,
                                                              int64_t B, int n_in, int n_out, float* __restrict__ dW,
                                                              float* __restrict__ db) {
  __shared__ float red[256];
  const int64_t e = blockIdx.x;
  const int j = (int)(e / (n_in + 1));
  const int i = (int)(e % (n_in + 1));
  float s = 0.f;
  for (int64_t b = threadIdx.x; b < B; b += 256) {
    const float g = dz[b * n_out + j];
    s += (i < n_in) ? g * in[b * n_in + i] : g;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (i < n_in) dW[(int64_t)j * n_in + i] = red[0];
    else db[j] = red[0];
  }
}

// dpred[b] = 2 (pred[b] - target[b]) / B  (gradient of the mean squared error)
__global__ void mlp_mse_grad_kernel(const float* __restrict__ pred, const float* __restrict__ target, int64_t B,
                                    float* __restrict__ dpred) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b < B) dpred[b] = 2.f * (pred[b] - target[b]) / (float)B;
}

// Adam (Kingma & Ba 2015) on a flat parameter buffer; step t >= 1
__global__ void adam_kernel(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, int64_t n, float lr, float beta1, float beta2, float eps, int t) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float gi = g[e];
  const float mi = beta1 * m[e] + (1.f - beta1) * gi;
  const float vi = beta2 * v[e] + (1.f - beta2) * gi * gi;
  m[e] = mi;
  v[e] = vi;
  const float mh = mi / (1.f - powf(beta1, (float)t));
  const float vh = vi / (1.f - powf(beta2, (float)t));
  w[e] -= lr * mh / (sqrtf(vh) + eps);
}

static inline dim3 blocks_for(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

hipError_t launch_mlp_features(const float* uy, int64_t P, int64_t p_pad, int64_t neff, const int* bin_of_k,
                               float logN, float C, float* feat, hipStream_t stream) {
  hipLaunchKernelGGL(mlp_features_kernel, dim3((unsigned)((P + 3) / 4)), dim3(256), 0, stream, uy, P, p_pad, neff,
                     bin_of_k, logN, C, feat);
  return hipGetLastError();
}

hipError_t launch_mlp_dense_fwd(const float* in, int64_t B, int n_in, const float* W, const float* bias, int n_out,
                                int act, float* out, hipStream_t stream) {
  hipLaunchKernelGGL(mlp_dense_fwd_kernel, blocks_for(B * n_out), dim3(256), 0, stream, in, B, n_in, W, bias, n_out,
                     act, out);
  return hipGetLastError();
}

hipError_t launch_mlp_dense_bwd(const float* dout_in, float* dz, const float* out, const float* in, int64_t B,
                                int n_in, int n_out, int act, const float* W, float* din, float* dW, float* db,
                                hipStream_t stream) {
  const int64_t n = B * n_out;
  if (dz != dout_in) {
    hipError_t e = hipMemcpyAsync(dz, dout_in, n * sizeof(float), hipMemcpyDeviceToDevice, stream);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(mlp_act_bwd_kernel, blocks_for(n), dim3(256), 0, stream, dz, out, n, act);
  hipLaunchKernelGGL(mlp_dense_bwd_w_kernel, dim3((unsigned)(n_out * (n_in + 1))), dim3(256), 0, stream, dz, in, B,
                     n_in, n_out, dW, db);
  if (din)
    hipLaunchKernelGGL(mlp_dense_bwd_in_kernel, blocks_for(B * n_in), dim3(256), 0, stream, dz, B, n_out, W, n_in,
                       din);
  return hipGetLastError();
}

hipError_t launch_mlp_mse_grad(const float* pred, const float* target, int64_t B, float* dpred, hipStream_t stream) {
  hipLaunchKernelGGL(mlp_mse_grad_kernel, blocks_for(B), dim3(256), 0, stream, pred, target, B, dpred);
  return hipGetLastError();
}

hipError_t launch_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, float beta1, float beta2,
                       float eps, int t, hipStream_t stream) {
  hipLaunchKernelGGL(adam_kernel, blocks_for(n), dim3(256), 0, stream, w, g, m, v, n, lr, beta1, beta2, eps, t);
  return hipGetLastError();
}

}  // namespace flmm
