// Caller-side statistics of the scan output on the GPU (reference fasterlmm/cli.py:128-132, 405-412):
//   -log10 p with p = F(1, df2).sf(F)   (cli.py:405-406)
//   EffectSize = beta^2 * var(genotype) / var(phenotype)   (cli.py:132)
//
// The F(1, v) survival function is the regularised incomplete beta I_x(v/2, 1/2) at x = v / (v + F).  It is
// evaluated in log space so that p far below the FP64 range still gives a finite -log10 p:
//   log I_x(a, b) = a log x + b log(1 - x) - log a - log B(a, b) + log CF(a, b, x)
// with CF the continued fraction of I_x (modified Lentz evaluation, as published in Press et al., Numerical
// Recipes, sect. 6.4), used when x < (a + 1) / (a + b + 2); otherwise p = 1 - I_{1-x}(b, a) through the same
// continued fraction.
#include <cmath>

#include "kernels.h"

namespace flmm {

__device__ inline double betacf(double a, double b, double x) {
  const double TINY = 1e-300, EPS = 1e-16;
  const double qab = a + b, qap = a + 1.0, qam = a - 1.0;
  double c = 1.0;
  double d = 1.0 - qab * x / qap;
  if (fabs(d) < TINY) d = TINY;
  d = 1.0 / d;
  double h = d;
  for (int m = 1; m <= 10000; ++m) {
    const double m2 = 2.0 * m;
    double aa = m * (b - m) * x / ((qam + m2) * (a + m2));
    d = 1.0 + aa * d;
    if (fabs(d) < TINY) d = TINY;
    c = 1.0 + aa / c;
    if (fabs(c) < TINY) c = TINY;
    d = 1.0 / d;
    h *= d * c;
    aa = -(a + m) * (qab + m) * x / ((a + m2) * (qap + m2));
    d = 1.0 + aa * d;
    if (fabs(d) < TINY) d = TINY;
    c = 1.0 + aa / c;
    if (fabs(c) < TINY) c = TINY;
    d = 1.0 / d;
    const double del = d * c;
    h *= del;
    if (fabs(del - 1.0) < EPS) break;
  }
  return h;
}

// natural log of I_x(a, b); lnorm_a = log a + log B(a, b), lnorm_b = log b + log B(a, b) (constants of a call)
__device__ inline double log_ibeta(double a, double b, double x, double lnorm_a, double lnorm_b) {
  if (x <= 0.0) return -INFINITY;
  if (x >= 1.0) return 0.0;
  if (x < (a + 1.0) / (a + b + 2.0)) {
    return a * log(x) + b * log1p(-x) - lnorm_a + log(betacf(a, b, x));
  }
  const double y = 1.0 - x;
  const double comp = exp(b * log(y) + a * log1p(-y) - lnorm_b) * betacf(b, a, y);
  return log1p(-comp);
}

__device__ inline double neglog10_f1_sf(double F, double df2, double lnorm_a, double lnorm_b) {
  if (isnan(F)) return NAN;
  if (F <= 0.0) return 0.0;
  if (isinf(F)) return INFINITY;
  const double x = df2 / (df2 + F);
  return -log_ibeta(0.5 * df2, 0.5, x, lnorm_a, lnorm_b) / 2.302585092994045684;
}

// out[i] = -log10 F(1, df2).sf(F[i])
__global__ void neglog10p_kernel(const double* __restrict__ F, int64_t n, double df2, double lnorm_a,
                                 double lnorm_b, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = neglog10_f1_sf(F[i], df2, lnorm_a, lnorm_b);
}

// (M, n_real) row-major beta -> effect[m][j] = beta^2 * g_var[m] / p_var[j]
__global__ void effect_size_kernel(const double* __restrict__ beta, int64_t M, int64_t n_real,
                                   const double* __restrict__ g_var, const double* __restrict__ p_var,
                                   double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M * n_real) return;
  const int64_t m = i / n_real, j = i % n_real;
  const double b = beta[i];
  out[i] = b * b * g_var[m] / p_var[j];
}

hipError_t launch_neglog10p(const double* F, int64_t n, double df2, double* out, hipStream_t stream) {
  // the call's constants on the host: log B(a, b) = lgamma(a) + lgamma(b) - lgamma(a + b), a = df2 / 2, b = 1/2
  const double a = 0.5 * df2, b = 0.5;
  const double lbeta = std::lgamma(a) + std::lgamma(b) - std::lgamma(a + b);
  hipLaunchKernelGGL(neglog10p_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, F, n, df2,
                     std::log(a) + lbeta, std::log(b) + lbeta, out);
  return hipGetLastError();
}

hipError_t launch_effect_size(const double* beta, int64_t M, int64_t n_real, const double* g_var,
                              const double* p_var, double* out, hipStream_t stream) {
  const int64_t n = M * n_real;
  hipLaunchKernelGGL(effect_size_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, beta, M, n_real,
                     g_var, p_var, out);
  return hipGetLastError();
}

}  // namespace flmm
