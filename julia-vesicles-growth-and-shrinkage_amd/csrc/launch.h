// Generic element-wise and reduction launches built on device lambdas.
#pragma once
#include "common.h"

namespace sav {

// Workgroups of the Krylov vector kernels per system.
constexpr int KRY_VEC_BLOCKS = 2048;
// Workgroups of a reduce_n / max stage: 8 per CU of 256, so that the streaming loads
// of a reduction have enough wavefronts in flight.
constexpr int RED_STAGE_BLOCKS = 2048;
// Doubles in the solver's partials buffer.
constexpr long long RED_PARTIALS = 3 * 65536;

template <class F>
__global__ __launch_bounds__(256) void k_foreach(long long n, F f) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) f(i);
}

template <class F>
inline void foreach_n(hipStream_t s, long long n, F f) {
  if (n <= 0) return;
  long long nb = (n + 255) / 256;
  if (nb > 16384) nb = 16384;
  hipLaunchKernelGGL(k_foreach<F>, dim3((unsigned)nb), dim3(256), 0, s, n, f);
  SAV_LAUNCH_CHECK();
}

template <class F>
__global__ void k_single(F f) {
  if (threadIdx.x == 0 && blockIdx.x == 0) f();
}

template <class F>
inline void single(hipStream_t s, F f) {
  hipLaunchKernelGGL(k_single<F>, dim3(1), dim3(64), 0, s, f);
  SAV_LAUNCH_CHECK();
}

// Per-group reduction: group g = blockIdx.y covers elements [0, n) of that group;
// f(i, g, acc) accumulates K doubles.  Partials land in partials[(g*nb + b)*K + k].
template <int K, class F>
__global__ __launch_bounds__(256) void k_reduce(long long n, F f, double* __restrict__ partials) {
  __shared__ double scratch[K * 4];
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0.0;
  const int g = blockIdx.y;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) f(i, g, acc);
  block_sum<K>(acc, scratch);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) partials[((long long)g * gridDim.x + blockIdx.x) * K + k] = acc[k];
  }
}

// out[g*K + k] = sum_b partials[(g*nb + b)*K + k], fixed order -> deterministic.
__global__ __launch_bounds__(256) void k_red_fin(const double* __restrict__ partials, int nb, int K,
                                                 double* __restrict__ out);

template <int K, class F>
inline void reduce_n(hipStream_t s, long long n, int ngroups, F f, double* partials, double* out) {
  long long nb = (n + 1023) / 1024;
  if (nb < 1) nb = 1;
  if (nb > RED_STAGE_BLOCKS) nb = RED_STAGE_BLOCKS;
  const long long cap = RED_PARTIALS / ((long long)K * (ngroups > 0 ? ngroups : 1));
  if (nb > cap) nb = cap;
  if (nb < 1) throw HipError("reduce_n: too many groups for the partials buffer", -1);
  hipLaunchKernelGGL((k_reduce<K, F>), dim3((unsigned)nb, ngroups), dim3(256), 0, s, n, f, partials);
  SAV_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_red_fin, dim3(K, ngroups), dim3(256), 0, s, partials, (int)nb, K, out);
  SAV_LAUNCH_CHECK();
}

}  // namespace sav
