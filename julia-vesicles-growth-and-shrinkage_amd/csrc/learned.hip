// BF16 MFMA kernels of the learned psi preconditioner (inference and training).
//
// A periodic 3x3 convolution is an implicit GEMM per 16-pixel strip:
//   D[o][p] = sum_k Wm[o][k] * B[k][p],  k = tap*Cin + c,  B[k][p] = a[p + d(tap)][c]
// with one v_mfma_f32_16x16x32_bf16 per 32 values of k (A = weights, B = the
// activations of 16 consecutive pixels).  Activations are BF16, channel-fastest
// (NHWC), so a lane's 8 consecutive k of one tap are one 16-byte load; FP32
// accumulation; bias, GELU and its derivative in the FP32 epilogue.  The
// backward-data pass is the same kernel with the flipped, transposed weights
// (Wt[c][tap'*Cout + o] = W[o][(8 - tap')*Cin + c]); the weight gradient is an
// MFMA GEMM over pixels (A = gradient, B = activations) with per-block partials.
#include <cmath>

#include "learned.h"

namespace sav {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int CTX = 32, CTY = 16;  // output tile of one workgroup (4 waves)

__device__ __forceinline__ float gelu_f(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * expf(-0.5f * z * z) * 0.39894228040143268f;
}

template <int CIN>
__device__ __forceinline__ void load_halo(bf16_t* tile, const bf16_t* __restrict__ in, long long ibase, int H, int W,
                                          int y0, int x0) {
  for (int e = threadIdx.x; e < (CTY + 2) * (CTX + 2); e += blockDim.x) {
    const int ly = e / (CTX + 2), lx = e % (CTX + 2);
    const int gy = (y0 - 1 + ly + H) % H, gx = (x0 - 1 + lx + W) % W;
    const bf16_t* src = in + (ibase + (long long)gy * W + gx) * CIN;
    bf16_t* dst = tile + e * CIN;
    if constexpr (CIN == 16) {
      reinterpret_cast<uint4*>(dst)[0] = reinterpret_cast<const uint4*>(src)[0];
      reinterpret_cast<uint4*>(dst)[1] = reinterpret_cast<const uint4*>(src)[1];
    } else {
#pragma unroll
      for (int c = 0; c < CIN; ++c) dst[c] = src[c];
    }
  }
}

// B fragment: 8 consecutive k of this lane for output pixel (ty, tx) of the tile.
template <int CIN>
__device__ __forceinline__ bf16x8 frag_b(const bf16_t* tile, int kb, int g, int ty, int tx) {
  bf16x8 B;
  if constexpr (CIN == 16) {
    const int k0 = kb * 32 + 8 * g;
    const int tap = k0 >> 4, c0 = k0 & 15;
    if (tap < 9) {
      const int di = tap / 3, dj = tap % 3;
      B = *reinterpret_cast<const bf16x8*>(tile + ((ty + di) * (CTX + 2) + tx + dj) * 16 + c0);
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j) B[j] = (__bf16)0.0f;
    }
  } else {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int k = kb * 32 + 8 * g + j;
      const int tap = k / CIN, c = k % CIN;
      uint16_t v = 0;
      if (tap < 9) {
        const int di = tap / 3, dj = tap % 3;
        v = tile[((ty + di) * (CTX + 2) + tx + dj) * CIN + c];
      }
      B[j] = __builtin_bit_cast(__bf16, v);
    }
  }
  return B;
}

// EPI 0: hidden layer  (z -> zout, gelu(z) -> aout), EPI 1: linear output -> yout (fp32),
// EPI 2: backward data (D * gelu'(zprev) -> aout)
template <int CIN, int COUT, int EPI>
__global__ __launch_bounds__(256) void k_conv(const bf16_t* __restrict__ in, int H, int W,
                                              const bf16_t* __restrict__ Wm, const float* __restrict__ bias,
                                              bf16_t* __restrict__ zout, bf16_t* __restrict__ aout,
                                              float* __restrict__ yout, const bf16_t* __restrict__ zprev) {
  constexpr int KP = ((9 * CIN + 31) / 32) * 32;
  constexpr int KB = KP / 32;
  __shared__ __attribute__((aligned(16))) bf16_t tile[(CTY + 2) * (CTX + 2) * CIN];
  const long long ibase = (long long)blockIdx.z * H * W;
  const int y0 = blockIdx.y * CTY, x0 = blockIdx.x * CTX;
  load_halo<CIN>(tile, in, ibase, H, W, y0, x0);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, n = lane & 15;
  bf16x8 A[KB];
#pragma unroll
  for (int kb = 0; kb < KB; ++kb) A[kb] = *reinterpret_cast<const bf16x8*>(Wm + n * KP + kb * 32 + 8 * g);
  float bo[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) bo[i] = (EPI != 2 && 4 * g + i < COUT) ? bias[4 * g + i] : 0.0f;
#pragma unroll 2
  for (int t = 0; t < 8; ++t) {
    const int ty = wave * 4 + (t >> 1);
    const int tx = (t & 1) * 16 + n;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
      const bf16x8 B = frag_b<CIN>(tile, kb, g, ty, tx);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[kb], B, acc, 0, 0, 0);
    }
    const long long pix = ibase + (long long)(y0 + ty) * W + (x0 + tx);
    if constexpr (EPI == 0) {
      uint16_t zz[4], aa[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float z = acc[i] + bo[i];
        zz[i] = f2bf(z);
        aa[i] = f2bf(gelu_f(z));
      }
      const uint2 zv = make_uint2(zz[0] | ((uint32_t)zz[1] << 16), zz[2] | ((uint32_t)zz[3] << 16));
      const uint2 av = make_uint2(aa[0] | ((uint32_t)aa[1] << 16), aa[2] | ((uint32_t)aa[3] << 16));
      if (zout) *reinterpret_cast<uint2*>(zout + pix * 16 + 4 * g) = zv;
      *reinterpret_cast<uint2*>(aout + pix * 16 + 4 * g) = av;
    } else if constexpr (EPI == 1) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (4 * g + i < COUT) yout[pix * COUT + 4 * g + i] = acc[i] + bo[i];
    } else {
      const uint2 zp = *reinterpret_cast<const uint2*>(zprev + pix * 16 + 4 * g);
      const uint16_t zs[4] = {(uint16_t)(zp.x & 0xFFFF), (uint16_t)(zp.x >> 16), (uint16_t)(zp.y & 0xFFFF),
                              (uint16_t)(zp.y >> 16)};
      uint16_t o[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) o[i] = f2bf(acc[i] * gelu_d(bf2f(zs[i])));
      *reinterpret_cast<uint2*>(aout + pix * 16 + 4 * g) =
          make_uint2(o[0] | ((uint32_t)o[1] << 16), o[2] | ((uint32_t)o[3] << 16));
    }
  }
}

// Weight gradient partials: part[blk][o][n] = sum_{p in tile} g[p][o] * a[p + d(tap)][c], n = tap*CIN + c;
// part[blk][16*NKP + o] = sum_p g[p][o].
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_wgrad(const bf16_t* __restrict__ a, const bf16_t* __restrict__ gr, int H,
                                               int W, float* __restrict__ part) {
  constexpr int NK = 9 * CIN;
  constexpr int NKP = ((NK + 15) / 16) * 16;
  constexpr int NT = NKP / 16;
  constexpr int NPIX = CTY * CTX;
  __shared__ __attribute__((aligned(16))) bf16_t tile[(CTY + 2) * (CTX + 2) * CIN];
  __shared__ __attribute__((aligned(16))) bf16_t gT[16 * NPIX];
  __shared__ float bsum[16][4];
  const long long ibase = (long long)blockIdx.z * H * W;
  const int y0 = blockIdx.y * CTY, x0 = blockIdx.x * CTX;
  load_halo<CIN>(tile, a, ibase, H, W, y0, x0);
  float bacc[COUT];
#pragma unroll
  for (int o = 0; o < COUT; ++o) bacc[o] = 0.f;
  for (int p = threadIdx.x; p < NPIX; p += 256) {
    const int py = p / CTX, px = p % CTX;
    const long long pix = ibase + (long long)(y0 + py) * W + (x0 + px);
#pragma unroll
    for (int o = 0; o < 16; ++o) {
      const bf16_t v = o < COUT ? gr[pix * COUT + o] : (bf16_t)0;
      gT[o * NPIX + p] = v;
      if (o < COUT) bacc[o < COUT ? o : 0] += bf2f(v);
    }
  }
  // bias partials
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int o = 0; o < COUT; ++o) {
    float v = bacc[o];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) bsum[o][wave] = v;
  }
  __syncthreads();
  const long long blk = ((long long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
  float* out = part + blk * (16 * NKP + 16);
  if (threadIdx.x < 16) {
    const int o = threadIdx.x;
    out[16 * NKP + o] = o < COUT ? (bsum[o][0] + bsum[o][1] + bsum[o][2] + bsum[o][3]) : 0.f;
  }
  const int g = lane >> 4, n = lane & 15;
  for (int nt = wave; nt < NT; nt += 4) {
    const int col = nt * 16 + n;
    const int tap = col / CIN, c = col % CIN;
    const bool colok = col < NK;
    const int di = colok ? tap / 3 : 0, dj = colok ? tap % 3 : 0;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int ks = 0; ks < NPIX / 32; ++ks) {
      const int p0 = ks * 32 + 8 * g;
      const bf16x8 A = *reinterpret_cast<const bf16x8*>(gT + n * NPIX + p0);
      bf16x8 B;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int p = p0 + j;
        const int py = p / CTX, px = p % CTX;
        const uint16_t v = colok ? tile[((py + di) * (CTX + 2) + px + dj) * CIN + c] : (uint16_t)0;
        B[j] = __builtin_bit_cast(__bf16, v);
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) out[(4 * g + i) * NKP + col] = acc[i];
  }
}

// Sum weight-gradient partials over blocks (fixed order).
__global__ __launch_bounds__(256) void k_wgrad_reduce(const float* __restrict__ part, int nblk, int cout, int nk,
                                                      int nkp, float* __restrict__ gW, float* __restrict__ gb) {
  const int idx = blockIdx.x * 256 + threadIdx.x;
  const int nw = cout * nk;
  if (idx < nw) {
    const int o = idx / nk, k = idx % nk;
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += part[(long long)b * (16 * nkp + 16) + o * nkp + k];
    gW[idx] = (float)s;
  } else if (idx < nw + cout) {
    const int o = idx - nw;
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += part[(long long)b * (16 * nkp + 16) + 16 * nkp + o];
    gb[o] = (float)s;
  }
}

// Masked MSE over C channels: loss = sum_interior (y - t)^2 / n_int, n_int = interior
// pixels * C; g = 2 (y - t) / n_int on the interior, 0 outside.
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ y, const float* __restrict__ t, int nimg,
                                             int H, int W, int C, int border, bf16_t* __restrict__ g,
                                             double* __restrict__ lp) {
  __shared__ double sc[4];
  const long long n = (long long)nimg * H * W * C;
  const double nint = (double)nimg * (H - 2 * border) * (W - 2 * border) * C;
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const long long p = i / C;
    const int x = (int)(p % W), yy = (int)((p / W) % H);
    const bool in = x >= border && x < W - border && yy >= border && yy < H - border;
    const float d = y[i] - t[i];
    acc += in ? (double)d * (double)d : 0.0;
    g[i] = f2bf(in ? (float)(2.0 * d / nint) : 0.f);
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) sc[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) lp[blockIdx.x] = (sc[0] + sc[1] + sc[2] + sc[3]) / nint;
}

// Stencil loss in divergence form.  The network's first output channel is a coefficient
// c(p); the prediction is the 5-point form of -div(c grad v) with face averages,
//   pred[p] = rx/2 [(c_p + c_E)(v_p - v_E) + (c_p + c_W)(v_p - v_W)]
//           + ry/2 [(c_p + c_N)(v_p - v_N) + (c_p + c_S)(v_p - v_S)]
// (E/W = x -/+ 1 neighbours, N/S = y +/- 1, periodic in each image; rx = dy/dx, ry = dx/dy),
// which is the analytic first-order stencil for c = 1 - M / M_max and, for any c, annihilates
// constants and is symmetric, as the exact correction (a divergence) is.
// k_div_pred: d[p] = pred[p] - t[p] on the interior (0 outside), loss partials.
__global__ __launch_bounds__(256) void k_div_pred(const float* __restrict__ w, const float* __restrict__ v,
                                                  const float* __restrict__ t, int nimg, int H, int W, int border,
                                                  double rx, double ry, float* __restrict__ dres,
                                                  double* __restrict__ lp) {
  __shared__ double sc[4];
  const long long n = (long long)nimg * H * W;
  const double nint = (double)nimg * (H - 2 * border) * (W - 2 * border);
  double acc = 0.0;
  for (long long p = (long long)blockIdx.x * 256 + threadIdx.x; p < n; p += (long long)gridDim.x * 256) {
    const int xc = (int)(p % W), yy = (int)((p / W) % H);
    const long long b = p - (long long)yy * W - xc;  // first pixel of the image
    const bool in = xc >= border && xc < W - border && yy >= border && yy < H - border;
    const long long pe = b + (long long)yy * W + (xc + 1) % W, pw = b + (long long)yy * W + (xc + W - 1) % W;
    const long long pn = b + (long long)((yy + 1) % H) * W + xc, ps = b + (long long)((yy + H - 1) % H) * W + xc;
    const double c = w[p * 9], vp = v[p];
    const double pred = 0.5 * rx * ((c + w[pe * 9]) * (vp - v[pe]) + (c + w[pw * 9]) * (vp - v[pw])) +
                        0.5 * ry * ((c + w[pn * 9]) * (vp - v[pn]) + (c + w[ps * 9]) * (vp - v[ps]));
    const double d = in ? pred - (double)t[p] : 0.0;
    dres[p] = (float)d;
    acc += d * d;
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) sc[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) lp[blockIdx.x] = (sc[0] + sc[1] + sc[2] + sc[3]) / nint;
}

// k_div_grad: output gradient of the divergence-form loss, d loss / d c(q) collects the
// five predictions c(q) enters (q and its four neighbours); the other channels get 0.
__global__ __launch_bounds__(256) void k_div_grad(const float* __restrict__ v, const float* __restrict__ dres, int nimg,
                                                  int H, int W, int border, double rx, double ry,
                                                  bf16_t* __restrict__ g) {
  const long long n = (long long)nimg * H * W;
  const double nint = (double)nimg * (H - 2 * border) * (W - 2 * border);
  for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long long)gridDim.x * 256) {
    const int xc = (int)(q % W), yy = (int)((q / W) % H);
    const long long b = q - (long long)yy * W - xc;
    const long long pe = b + (long long)yy * W + (xc + 1) % W, pw = b + (long long)yy * W + (xc + W - 1) % W;
    const long long pn = b + (long long)((yy + 1) % H) * W + xc, ps = b + (long long)((yy + H - 1) % H) * W + xc;
    const double vq = v[q], ve = v[pe], vw = v[pw], vn = v[pn], vs = v[ps];
    const double gq = (double)dres[q] * 0.5 * (rx * (2.0 * vq - ve - vw) + ry * (2.0 * vq - vn - vs)) +
                      0.5 * rx * ((double)dres[pw] * (vw - vq) + (double)dres[pe] * (ve - vq)) +
                      0.5 * ry * ((double)dres[ps] * (vs - vq) + (double)dres[pn] * (vn - vq));
    g[q * 9] = f2bf((float)(2.0 * gq / nint));
#pragma unroll
    for (int ch = 1; ch < 9; ++ch) g[q * 9 + ch] = f2bf(0.f);
  }
}

__global__ void k_sum_d(const double* __restrict__ p, int n, double* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += p[i];
    *out = s;
  }
}

__global__ __launch_bounds__(256) void k_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                                              const float* __restrict__ g, int n, float lr, float b1, float b2,
                                              float eps, float c1, float c2) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Operand packing from the master weights.
__global__ __launch_bounds__(256) void k_pack(const float* __restrict__ W, int cin, int cout, int kp, int kpt,
                                              bf16_t* __restrict__ Wm, bf16_t* __restrict__ Wt) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < 16 * kp) {
    const int o = i / kp, k = i % kp;
    Wm[i] = (o < cout && k < 9 * cin) ? f2bf(W[o * 9 * cin + k]) : (bf16_t)0;
  }
  if (i < 16 * kpt) {
    const int c = i / kpt, kk = i % kpt;
    const int tap = kk / cout, o = kk % cout;
    Wt[i] = (c < cin && tap < 9) ? f2bf(W[o * 9 * cin + (8 - tap) * cin + c]) : (bf16_t)0;
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
constexpr int Net::CH[Net::NL + 1];

template <class T>
T* Net::alloc(size_t n) {
  void* p = nullptr;
  SAV_HIP(hipMalloc(&p, n * sizeof(T) + 64));
  SAV_HIP(hipMemsetAsync(p, 0, n * sizeof(T) + 64, st));
  allocs.push_back(p);
  return reinterpret_cast<T*>(p);
}

Net::Net(hipStream_t s) : st(s) {
  for (int l = 0; l < NL; ++l) {
    NetLayerDev& d = L[l];
    d.cin = CH[l];
    d.cout = CH[l + 1];
    d.kp = ((9 * d.cin + 31) / 32) * 32;
    d.kpt = ((9 * d.cout + 31) / 32) * 32;
    const int nw = d.cout * 9 * d.cin;
    d.W = alloc<float>(nw);
    d.b = alloc<float>(d.cout);
    d.mW = alloc<float>(nw);
    d.vW = alloc<float>(nw);
    d.mb = alloc<float>(d.cout);
    d.vb = alloc<float>(d.cout);
    d.gW = alloc<float>(nw);
    d.gb = alloc<float>(d.cout);
    d.Wm = alloc<bf16_t>(16 * d.kp);
    d.Wt = alloc<bf16_t>(16 * d.kpt);
  }
  dloss = alloc<double>(4096 + 1);
}

Net::~Net() {
  // a destructor cannot report: errors here are dropped on purpose
  if (st) (void)hipStreamSynchronize(st);
  for (void* p : allocs) (void)hipFree(p);
}

void Net::ensure(long long npix) {
  if (npix <= cap) return;
  // buffers are only ever grown; old ones stay allocated until the Net is destroyed
  cap = npix;
  for (int l = 1; l < NL; ++l) {
    act[l] = alloc<bf16_t>(npix * 16);
    zb[l] = alloc<bf16_t>(npix * 16);
  }
  g16a = alloc<bf16_t>(npix * 16);
  g16b = alloc<bf16_t>(npix * 16);
  glast = alloc<bf16_t>(npix * COUT);
  yb = alloc<float>(npix * COUT);
  dres = alloc<float>(npix);
  const long long nblk = (npix + CTX * CTY - 1) / (CTX * CTY);
  partials = alloc<float>(nblk * (16 * 144 + 16));
}

void Net::pack_operands() {
  for (int l = 0; l < NL; ++l) {
    NetLayerDev& d = L[l];
    const int n = 16 * (d.kp > d.kpt ? d.kp : d.kpt);
    hipLaunchKernelGGL(k_pack, dim3((n + 255) / 256), dim3(256), 0, st, d.W, d.cin, d.cout, d.kp, d.kpt, d.Wm, d.Wt);
  }
  SAV_LAUNCH_CHECK();
}

void Net::set_params(const float* h) {
  size_t off = 0;
  for (int l = 0; l < NL; ++l) {
    NetLayerDev& d = L[l];
    // host layout W[o][c][di][dj]; device layout W[o][tap*cin + c]
    std::vector<float> w(d.cout * 9 * d.cin);
    for (int o = 0; o < d.cout; ++o)
      for (int c = 0; c < d.cin; ++c)
        for (int tap = 0; tap < 9; ++tap) w[o * 9 * d.cin + tap * d.cin + c] = h[off + (o * d.cin + c) * 9 + tap];
    off += w.size();
    SAV_HIP(hipMemcpyAsync(d.W, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice, st));
    SAV_HIP(hipMemcpyAsync(d.b, h + off, d.cout * sizeof(float), hipMemcpyHostToDevice, st));
    off += d.cout;
    SAV_HIP(hipMemsetAsync(d.mW, 0, w.size() * sizeof(float), st));
    SAV_HIP(hipMemsetAsync(d.vW, 0, w.size() * sizeof(float), st));
    SAV_HIP(hipMemsetAsync(d.mb, 0, d.cout * sizeof(float), st));
    SAV_HIP(hipMemsetAsync(d.vb, 0, d.cout * sizeof(float), st));
    SAV_HIP(hipStreamSynchronize(st));
  }
  t_adam = 0;
  pack_operands();
  SAV_HIP(hipStreamSynchronize(st));
}

void Net::init_random(uint64_t seed) {
  uint64_t z = seed;
  auto next = [&z]() {  // SplitMix64
    z += 0x9E3779B97F4A7C15ULL;
    uint64_t x = z;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
  };
  std::vector<float> h(NPARAM, 0.f);
  size_t off = 0;
  for (int l = 0; l < NL; ++l) {
    const int cin = CH[l], cout = CH[l + 1], nw = cout * cin * 9;
    const double a = (l == NL - 1) ? 0.0 : std::sqrt(6.0 / (9.0 * cin));
    for (int i = 0; i < nw; ++i) {
      const double u = (double)(next() >> 11) * (1.0 / 9007199254740992.0);
      h[off + i] = (float)(a * (2.0 * u - 1.0));
    }
    off += nw + cout;
  }
  set_params(h.data());
}

static void to_host_layout(const NetLayerDev& d, const float* dev, float* h) {
  std::vector<float> w(d.cout * 9 * d.cin);
  SAV_HIP(hipMemcpy(w.data(), dev, w.size() * sizeof(float), hipMemcpyDeviceToHost));
  for (int o = 0; o < d.cout; ++o)
    for (int c = 0; c < d.cin; ++c)
      for (int tap = 0; tap < 9; ++tap) h[(o * d.cin + c) * 9 + tap] = w[o * 9 * d.cin + tap * d.cin + c];
}

void Net::get_params(float* h) {
  SAV_HIP(hipStreamSynchronize(st));
  size_t off = 0;
  for (int l = 0; l < NL; ++l) {
    const NetLayerDev& d = L[l];
    to_host_layout(d, d.W, h + off);
    off += d.cout * 9 * d.cin;
    SAV_HIP(hipMemcpy(h + off, d.b, d.cout * sizeof(float), hipMemcpyDeviceToHost));
    off += d.cout;
  }
}

void Net::get_grads(float* h) {
  SAV_HIP(hipStreamSynchronize(st));
  size_t off = 0;
  for (int l = 0; l < NL; ++l) {
    const NetLayerDev& d = L[l];
    to_host_layout(d, d.gW, h + off);
    off += d.cout * 9 * d.cin;
    SAV_HIP(hipMemcpy(h + off, d.gb, d.cout * sizeof(float), hipMemcpyDeviceToHost));
    off += d.cout;
  }
}

static void check_dims(int H, int W) {
  if (H % CTY != 0 || W % CTX != 0)
    throw HipError("learned preconditioner images must be multiples of 16 x 32", -1);
}

void Net::forward(const bf16_t* x, int nimg, int H, int W, float* y, bool keep) {
  check_dims(H, W);
  const long long npix = (long long)nimg * H * W;
  ensure(npix);
  dim3 grid(W / CTX, H / CTY, nimg);
  hipLaunchKernelGGL((k_conv<2, 16, 0>), grid, dim3(256), 0, st, x, H, W, L[0].Wm, L[0].b, keep ? zb[1] : nullptr,
                     act[1], nullptr, nullptr);
  hipLaunchKernelGGL((k_conv<16, 16, 0>), grid, dim3(256), 0, st, act[1], H, W, L[1].Wm, L[1].b,
                     keep ? zb[2] : nullptr, act[2], nullptr, nullptr);
  hipLaunchKernelGGL((k_conv<16, 16, 0>), grid, dim3(256), 0, st, act[2], H, W, L[2].Wm, L[2].b,
                     keep ? zb[3] : nullptr, act[3], nullptr, nullptr);
  hipLaunchKernelGGL((k_conv<16, COUT, 1>), grid, dim3(256), 0, st, act[3], H, W, L[3].Wm, L[3].b, nullptr,
                     nullptr, y ? y : yb, nullptr);
  SAV_LAUNCH_CHECK();
}

double Net::loss_backward(const bf16_t* x, const float* target, int nimg, int H, int W, int border) {
  forward(x, nimg, H, W, nullptr, true);
  const long long n = (long long)nimg * H * W * COUT;
  int nb = (int)((n + 1023) / 1024);
  if (nb > 4096) nb = 4096;
  hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, st, yb, target, nimg, H, W, COUT, border, glast, dloss);
  nb_loss = nb;
  return backward(x, nimg, H, W);
}

double Net::loss_backward_stencil(const bf16_t* x, const float* v, const float* target, int nimg, int H, int W,
                                  int border, double rx, double ry) {
  forward(x, nimg, H, W, nullptr, true);
  const long long n = (long long)nimg * H * W;
  int nb = (int)((n + 1023) / 1024);
  if (nb > 4096) nb = 4096;
  hipLaunchKernelGGL(k_div_pred, dim3(nb), dim3(256), 0, st, yb, v, target, nimg, H, W, border, rx, ry, dres, dloss);
  hipLaunchKernelGGL(k_div_grad, dim3(nb), dim3(256), 0, st, v, dres, nimg, H, W, border, rx, ry, glast);
  SAV_LAUNCH_CHECK();
  nb_loss = nb;
  return backward(x, nimg, H, W);
}

// Back-propagates the output gradient in `glast` (loss partials in dloss[0 .. nb_loss)).
double Net::backward(const bf16_t* x, int nimg, int H, int W) {
  hipLaunchKernelGGL(k_sum_d, dim3(1), dim3(64), 0, st, dloss, nb_loss, dloss + 4096);
  dim3 grid(W / CTX, H / CTY, nimg);
  const int nblk = (int)(grid.x * grid.y * grid.z);
  auto reduce = [&](int l) {
    const NetLayerDev& d = L[l];
    const int nkp = ((9 * d.cin + 15) / 16) * 16;
    const int n = d.cout * 9 * d.cin + d.cout;
    hipLaunchKernelGGL(k_wgrad_reduce, dim3((n + 255) / 256), dim3(256), 0, st, partials, nblk, d.cout, 9 * d.cin,
                       nkp, d.gW, d.gb);
  };
  // layer 4
  hipLaunchKernelGGL((k_wgrad<16, COUT>), grid, dim3(256), 0, st, act[3], glast, H, W, partials);
  reduce(3);
  hipLaunchKernelGGL((k_conv<COUT, 16, 2>), grid, dim3(256), 0, st, glast, H, W, L[3].Wt, nullptr, nullptr, g16a,
                     nullptr, zb[3]);
  // layer 3
  hipLaunchKernelGGL((k_wgrad<16, 16>), grid, dim3(256), 0, st, act[2], g16a, H, W, partials);
  reduce(2);
  hipLaunchKernelGGL((k_conv<16, 16, 2>), grid, dim3(256), 0, st, g16a, H, W, L[2].Wt, nullptr, nullptr, g16b,
                     nullptr, zb[2]);
  // layer 2
  hipLaunchKernelGGL((k_wgrad<16, 16>), grid, dim3(256), 0, st, act[1], g16b, H, W, partials);
  reduce(1);
  hipLaunchKernelGGL((k_conv<16, 16, 2>), grid, dim3(256), 0, st, g16b, H, W, L[1].Wt, nullptr, nullptr, g16a,
                     nullptr, zb[1]);
  // layer 1
  hipLaunchKernelGGL((k_wgrad<2, 16>), grid, dim3(256), 0, st, x, g16a, H, W, partials);
  reduce(0);
  SAV_LAUNCH_CHECK();
  double loss = 0.0;
  SAV_HIP(hipMemcpyAsync(&loss, dloss + 4096, sizeof(double), hipMemcpyDeviceToHost, st));
  SAV_HIP(hipStreamSynchronize(st));
  return loss;
}

void Net::adam_step(float lr, float b1, float b2, float eps) {
  ++t_adam;
  const float c1 = 1.f - powf(b1, (float)t_adam), c2 = 1.f - powf(b2, (float)t_adam);
  for (int l = 0; l < NL; ++l) {
    NetLayerDev& d = L[l];
    const int nw = d.cout * 9 * d.cin;
    hipLaunchKernelGGL(k_adam, dim3((nw + 255) / 256), dim3(256), 0, st, d.W, d.mW, d.vW, d.gW, nw, lr, b1, b2, eps,
                       c1, c2);
    hipLaunchKernelGGL(k_adam, dim3(1), dim3(256), 0, st, d.b, d.mb, d.vb, d.gb, d.cout, lr, b1, b2, eps, c1, c2);
  }
  SAV_LAUNCH_CHECK();
  pack_operands();
}

}  // namespace sav
