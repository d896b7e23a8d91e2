// Common device/host helpers for libsav_hip (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace sav {

// ---------------------------------------------------------------------------
// Complex numbers in interleaved (re, im) layout -- the memory layout of Julia's
// ComplexF64 / ComplexF32 and of FFTW's fftw_complex.
// ---------------------------------------------------------------------------
template <typename R>
struct alignas(2 * sizeof(R)) cplx {
  R x, y;
  cplx() = default;
  __host__ __device__ constexpr cplx(R a, R b) : x(a), y(b) {}
  template <typename S>
  __host__ __device__ constexpr cplx(const cplx<S>& o) : x((R)o.x), y((R)o.y) {}
};

template <typename R> __host__ __device__ __forceinline__ cplx<R> mk(R a, R b) { return cplx<R>{a, b}; }
template <typename R> __host__ __device__ __forceinline__ cplx<R> operator+(cplx<R> a, cplx<R> b) { return {a.x + b.x, a.y + b.y}; }
template <typename R> __host__ __device__ __forceinline__ cplx<R> operator-(cplx<R> a, cplx<R> b) { return {a.x - b.x, a.y - b.y}; }
template <typename R> __host__ __device__ __forceinline__ cplx<R> operator*(cplx<R> a, cplx<R> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename R> __host__ __device__ __forceinline__ cplx<R> operator*(R s, cplx<R> a) { return {s * a.x, s * a.y}; }
template <typename R> __host__ __device__ __forceinline__ cplx<R> conj(cplx<R> a) { return {a.x, -a.y}; }
// multiply by +i or -i
template <typename R> __host__ __device__ __forceinline__ cplx<R> mul_i(cplx<R> a) { return {-a.y, a.x}; }
template <typename R> __host__ __device__ __forceinline__ cplx<R> mul_mi(cplx<R> a) { return {a.y, -a.x}; }

// complex scalar held in double precision (Krylov scalars)
struct zd {
  double x, y;
};
__host__ __device__ __forceinline__ zd zmul(zd a, zd b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__host__ __device__ __forceinline__ zd zdiv(zd a, zd b) {
  // Smith's algorithm
  if (fabs(b.x) >= fabs(b.y)) {
    double r = b.y / b.x, d = b.x + b.y * r;
    return {(a.x + a.y * r) / d, (a.y - a.x * r) / d};
  } else {
    double r = b.x / b.y, d = b.x * r + b.y;
    return {(a.x * r + a.y) / d, (a.y * r - a.x) / d};
  }
}

// ---------------------------------------------------------------------------
// Errors
// ---------------------------------------------------------------------------
struct HipError : std::runtime_error {
  int code;
  HipError(const std::string& m, int c) : std::runtime_error(m), code(c) {}
};

#define SAV_HIP(call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      throw ::sav::HipError(std::string(#call) + ": " + hipGetErrorString(e_), (int)e_); \
  } while (0)

#define SAV_LAUNCH_CHECK() SAV_HIP(hipGetLastError())

// ---------------------------------------------------------------------------
// Wave / block reductions in double precision (wave64).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  return v;
}

// Sum K values across the block; result valid in thread 0.  `scratch` must hold
// K * (blockDim.x / 64) doubles.
template <int K>
__device__ __forceinline__ void block_sum(double (&acc)[K], double* scratch) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double v = wave_sum(acc[k]);
    if (lane == 0) scratch[k * nw + wid] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double s = 0.0;
      for (int w = 0; w < nw; ++w) s += scratch[k * nw + w];
      acc[k] = s;
    }
  }
}

template <typename R> __device__ __forceinline__ double to_d(R v) { return (double)v; }

// Strided partial sums of a single-workgroup finish kernel: for k < K, acc[k] += p[i * K + k]
// over i = t, t + NT, t + 2 NT, ... < nb, added in that order.  The loads are issued in
// groups of 8 before any of them is added, so that they are in flight together: the
// partials were written by the previous kernel on every XCD and come from beyond this
// XCD's L2, and one load at a time made such a kernel latency-bound.
template <int K, int NT>
__device__ __forceinline__ void strided_sums(const double* __restrict__ p, int nb, int t, double (&acc)[K]) {
  constexpr int U = 8;
  for (int i0 = t; i0 < nb; i0 += U * NT) {
    double v[U][K];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = i0 + u * NT;
#pragma unroll
      for (int k = 0; k < K; ++k) v[u][k] = i < nb ? p[(long long)i * K + k] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (i0 + u * NT < nb) {
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] += v[u][k];
      }
  }
}

// ---------------------------------------------------------------------------
// XCD-aware workgroup numbering.  The workgroups of a launch reach the 8 XCDs of the
// chip round-robin by linear id, and each XCD has its own L2.  Tiles that share cache
// lines (e.g. two 64-byte column segments of one 128-byte line) should run on one XCD
// at about the same time, so that the second read of a line hits that XCD's L2 instead
// of fetching the line from HBM again.
// ---------------------------------------------------------------------------
constexpr unsigned kNumXcd = 8;

// Bijection on [0, total): the ids one XCD receives (lin = x, x + 8, x + 16, ...) map to a
// contiguous range of logical ids, in dispatch order.
__host__ __device__ __forceinline__ unsigned xcd_spread(unsigned lin, unsigned total) {
  const unsigned q = total / kNumXcd, r = total % kNumXcd, x = lin % kNumXcd, k = lin / kNumXcd;
  return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

struct BlockIdx3 {
  int x, y, z;
};

// Logical block index of block (bx, by, bz) in a (gx, gy, gz) grid.  With a multiple of 8
// blocks along x the renumbering stays within each (y, z) row of tiles (every row keeps all
// 8 XCDs, so rows that exit early, such as converged systems of a batched solve, idle no
// XCD); otherwise it runs over the whole grid.
__host__ __device__ __forceinline__ BlockIdx3 xcd_map(unsigned bx, unsigned by, unsigned bz, unsigned gx, unsigned gy,
                                                      unsigned gz) {
  if (gx % kNumXcd == 0) return BlockIdx3{(int)xcd_spread(bx, gx), (int)by, (int)bz};
  const unsigned l = xcd_spread(bx + gx * (by + gy * bz), gx * gy * gz);
  return BlockIdx3{(int)(l % gx), (int)((l / gx) % gy), (int)(l / (gx * gy))};
}

// Logical block index of the calling workgroup.
__device__ __forceinline__ BlockIdx3 xcd_block() {
  return xcd_map(blockIdx.x, blockIdx.y, blockIdx.z, gridDim.x, gridDim.y, gridDim.z);
}

}  // namespace sav
