// Batched real/complex FFT kernels over the spatial axes of (Nx, Ny[, Nz]) fields
// stored x-fastest (the Julia column-major layout).  See fft.h for the team FFT.
#include "fft.h"

namespace sav {

// Columns handled per workgroup for a column (strided) transform of length L: 8, or 4 for FP64
// columns of 512 points or more, whose 8-column tile (128 KB of LDS for L = 1024) allows one
// workgroup per CU; with 4 columns two workgroups share a CU and overlap their load, transform
// and store phases (1024^2: 60.8 -> 58.9 us inverse, 28.6 -> 24.5 us forward, per launch).
// FP32 keeps 8 (4 columns: 35.7 -> 39.0 us inverse).
// The width is also tunable at run time (FftGeom::colw, 4 / 8 / 16; ColCfgW), since each
// column's transform is the same whichever tile it shares: the results are bit-identical.
template <typename R, int L, int CW = 0>
using ColCfgW = ColTile<R, L, CW != 0 ? CW : ((sizeof(R) == 8 && L >= 512) ? 4 : 8), cols_maxr(L, sizeof(R) == 4)>;
template <typename R, int L>
using ColCfg = ColCfgW<R, L, 0>;

int col_tile_columns(int L, int real_bytes, int cw) {
  switch (L) {
#define X(LL)                                                                                              \
  case LL:                                                                                                 \
    if (real_bytes == 8)                                                                                   \
      return cw == 4 ? ColCfgW<double, LL, 4>::C : cw == 16 ? ColCfgW<double, LL, 16>::C : cw == 8 ? ColCfgW<double, LL, 8>::C : ColCfgW<double, LL, 0>::C; \
    return cw == 4 ? ColCfgW<float, LL, 4>::C : cw == 16 ? ColCfgW<float, LL, 16>::C : cw == 8 ? ColCfgW<float, LL, 8>::C : ColCfgW<float, LL, 0>::C;
    X(2) X(3) X(4) X(6) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024)
#undef X
    default: return 0;
  }
}

// Rows handled per workgroup for a contiguous transform of length L: one wavefront per
// workgroup (measured faster than four rows per 256-thread workgroup), or one row when
// a row's team is larger than a wavefront.
template <typename R, int L>
struct RowCfg {
  using F = TeamFFT<R, L, false>;
  static constexpr int TT = F::TT;
  static constexpr int ROWS = TT >= 64 ? 1 : 64 / TT;
  static constexpr int NT = ROWS * TT;
  static constexpr int LDSN = ROWS * (L + 1);
};

struct ColGeom {
  int inner;            // columns per sub-batch
  int stride;           // elements between consecutive positions
  int nsub;             // sub-batches per field
  long long sub_stride; // elements between sub-batches
  long long field_stride;
  int geom;             // 0: kx = col, ky = pos ; 1: kx = col % Nxh, ky = col / Nxh, kz = pos ; 2: none
  int Nxh;
  int use_src;          // 1: read job.src(/src2) with multiplier, 0: in place on job.dst
  long long toff_unit;  // unused by the kernel (host bookkeeping)
};

// ---------------------------------------------------------------------------
// Column transform (strided axis), several adjacent columns per workgroup.
// ---------------------------------------------------------------------------
template <typename R, int L, bool INV, int CW>
__global__ __launch_bounds__((ColCfgW<R, L, CW>::NT)) void k_cols(SpecJobs jobs, ColGeom cg,
                                                               const cplx<R>* __restrict__ tw, KTables kt) {
  using F = TeamFFT<R, L, INV, (sizeof(R) == 8 ? 128 : 64), 0, default_twm<R>(), 0, cols_maxr(L, sizeof(R) == 4)>;
  constexpr int C = ColCfgW<R, L, CW>::C;
  constexpr int E = F::E;
  __shared__ cplx<R> lds[C * L];
  const BlockIdx3 bk = xcd_block();
  const SpecJob& jb = jobs.j[bk.z];
  const int fidx = bk.y / cg.nsub;
  const int sub = bk.y % cg.nsub;
  if (fidx >= jb.count) return;
  const int c = threadIdx.x % C;
  const int t = threadIdx.x / C;
  const int col = bk.x * C + c;
  const bool act = col < cg.inner;
  const long long base = (long long)fidx * cg.field_stride + (long long)sub * cg.sub_stride + col;
  cplx<R>* dst = reinterpret_cast<cplx<R>*>(jb.dst);
  cplx<R> v[E];
  if (cg.use_src) {
    const cplx<R>* src = reinterpret_cast<const cplx<R>*>(jb.src);
    const cplx<R>* src2 = reinterpret_cast<const cplx<R>*>(jb.src2);
    double kx = 0, ky = 0;
    if (act) {
      if (cg.geom == 1) {
        kx = kt.kx[col % cg.Nxh];
        ky = kt.ky[col / cg.Nxh];
      } else {
        kx = kt.kx[col];
      }
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int pos = F::in_pos(t, e);
      cplx<R> a = cplx<R>{R(0), R(0)};
      if (act) {
        const long long off = base + (long long)pos * cg.stride;
        a = src[off];
        if (src2) {
          cplx<R> b = src2[off];
          a = cplx<R>{R(2.0) * a.x - b.x, R(2.0) * a.y - b.y};
        }
        if (jb.mul) {
          double kyy = ky, kz = 0.0;
          if (cg.geom == 0) kyy = kt.ky[pos];
          else kz = kt.kz[pos];
          a = apply_mul<R>(a, jb.mul, kx, kyy, kz);
        }
        if (jb.scale != 1.0) a = (R)jb.scale * a;
      }
      v[e] = a;
    }
  } else {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int pos = F::in_pos(t, e);
      v[e] = act ? dst[base + (long long)pos * cg.stride] : cplx<R>{R(0), R(0)};
    }
  }
  F::run(v, lds + c, C, t, tw);
  if (act) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int pos = F::out_pos(t, e);
      dst[base + (long long)pos * cg.stride] = v[e];
    }
  }
}

// ---------------------------------------------------------------------------
// Rows: real -> half spectrum along x (Nx = 2L), via a length-L complex FFT of
// z[j] = x[2j] + i x[2j+1] and the split  X[k] = E[k] + W^k O[k].
// ---------------------------------------------------------------------------
template <typename R, int L>
__global__ __launch_bounds__((RowCfg<R, L>::NT)) void k_rows_r2c(SpecJobs jobs, int rows_per_field,
                                                              const cplx<R>* __restrict__ tw,
                                                              const cplx<R>* __restrict__ tw2) {
  using F = TeamFFT<R, L, false>;
  using CF = RowCfg<R, L>;
  constexpr int E = F::E, TT = F::TT;
  __shared__ cplx<R> lds[CF::LDSN];
  const SpecJob& jb = jobs.j[blockIdx.y];
  const int rloc = threadIdx.x / TT, t = threadIdx.x % TT;
  const long long row = (long long)blockIdx.x * CF::ROWS + rloc;
  const long long nrows = (long long)jb.count * rows_per_field;
  if ((long long)blockIdx.x * CF::ROWS >= nrows) return;
  const bool act = row < nrows;
  const cplx<R>* src = reinterpret_cast<const cplx<R>*>(reinterpret_cast<const R*>(jb.src) + row * (2 * L));
  cplx<R>* dst = reinterpret_cast<cplx<R>*>(jb.dst) + row * (L + 1);
  cplx<R>* my = lds + rloc * (L + 1);
  cplx<R> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = act ? src[F::in_pos(t, e)] : cplx<R>{R(0), R(0)};
  F::run(v, my, 1, t, tw);
  __syncthreads();
#pragma unroll
  for (int e = 0; e < E; ++e) my[F::out_pos(t, e)] = v[e];
  __syncthreads();
  for (int k = t; k <= L; k += TT) {
    const cplx<R> zk = my[k == L ? 0 : k];
    const cplx<R> zm = conj(my[k == 0 ? 0 : L - k]);
    const cplx<R> ev = R(0.5) * (zk + zm);
    const cplx<R> od = mul_mi(R(0.5) * (zk - zm));  // (zk - zm) / (2i)
    const cplx<R> w = tw2[k];                        // exp(-2 pi i k / (2L))
    if (act) dst[k] = ev + w * od;
  }
}

// ---------------------------------------------------------------------------
// Rows: half spectrum -> real along x.  Imag parts of the DC and Nyquist bins
// are ignored (FFTW c2r semantics).  Output scaled by `scale`.
// ---------------------------------------------------------------------------
template <typename R, int L>
__global__ __launch_bounds__((RowCfg<R, L>::NT)) void k_rows_c2r(SpecJobs jobs, int rows_per_field, const cplx<R>* tmp,
                                                              double scale,
                                                              const cplx<R>* __restrict__ tw,
                                                              const cplx<R>* __restrict__ tw2) {
  using F = TeamFFT<R, L, true>;
  using CF = RowCfg<R, L>;
  constexpr int E = F::E, TT = F::TT;
  __shared__ cplx<R> lds[CF::LDSN];
  const SpecJob& jb = jobs.j[blockIdx.y];
  const int rloc = threadIdx.x / TT, t = threadIdx.x % TT;
  const long long row = (long long)blockIdx.x * CF::ROWS + rloc;
  const long long nrows = (long long)jb.count * rows_per_field;
  if ((long long)blockIdx.x * CF::ROWS >= nrows) return;
  const bool act = row < nrows;
  const cplx<R>* src = tmp + jb.toff + row * (L + 1);
  cplx<R>* my = lds + rloc * (L + 1);
  for (int k = t; k <= L; k += TT) my[k] = act ? src[k] : cplx<R>{R(0), R(0)};
  __syncthreads();
  cplx<R> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = F::in_pos(t, e);
    cplx<R> xk = my[k];
    cplx<R> xm = my[L - k];
    if (k == 0) {
      xk.y = R(0);
      xm.y = R(0);
    }
    const cplx<R> ev = xk + conj(xm);
    const cplx<R> od = (xk - conj(xm)) * conj(tw2[k]);
    v[e] = ev + mul_i(od);
  }
  F::run(v, my, 1, t, tw);
  R* out = reinterpret_cast<R*>(jb.dst) + row * (2 * L);
  cplx<R>* outc = reinterpret_cast<cplx<R>*>(out);
  const R sc = (R)scale;
  if (act) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const cplx<R> z = v[e];
      outc[F::out_pos(t, e)] = cplx<R>{sc * z.x, sc * z.y};
    }
  }
}

// ---------------------------------------------------------------------------
// Rows: full complex c2c along x (length L), in place, optional scale.
// ---------------------------------------------------------------------------
template <typename R, int L, bool INV>
__global__ __launch_bounds__((RowCfg<R, L>::NT)) void k_rows_c2c(cplx<R>* data, long long nrows, double scale,
                                                              const cplx<R>* __restrict__ tw) {
  using F = TeamFFT<R, L, INV>;
  using CF = RowCfg<R, L>;
  constexpr int E = F::E, TT = F::TT;
  __shared__ cplx<R> lds[CF::LDSN];
  const int rloc = threadIdx.x / TT, t = threadIdx.x % TT;
  const long long row = (long long)blockIdx.x * CF::ROWS + rloc;
  const bool act = row < nrows;
  cplx<R>* p = data + row * L;
  cplx<R> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = act ? p[F::in_pos(t, e)] : cplx<R>{R(0), R(0)};
  F::run(v, lds + rloc * (L + 1), 1, t, tw);
  const R sc = (R)scale;
  if (act) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      cplx<R> z = v[e];
      if (scale != 1.0) z = sc * z;
      p[F::out_pos(t, e)] = z;
    }
  }
}

// ---------------------------------------------------------------------------
// 3D planes: one workgroup holds a whole (x, y) plane of a half spectrum,
// (Nx/2+1) x Ny complex values (x fastest, contiguous in HBM), in LDS, and runs the
// y transforms and the x real<->half transforms of that plane in one pass over HBM.
// Square planes (Nx == Ny == N) whose half spectrum fits the 160 KB of a CU's LDS.
// Teams of both phases lie within one wavefront and exchange through their own
// column or row of the plane with wavefront-level barriers.  When a phase has more
// teams than columns or rows, the wavefronts past the last one skip the phase and the
// spare teams of the last wavefront repeat its last column or row (identical values
// written to identical addresses), so no barrier is ever divergent.
// ---------------------------------------------------------------------------
template <typename R, int N>
struct PlaneCfg {
  static constexpr int LX = N / 2, NXH = N / 2 + 1;
  template <bool INV>
  using FY = TeamFFT<R, N, INV, 64, 0, default_twm<R>(), 1>;
  template <bool INV>
  using FX = TeamFFT<R, LX, INV, 64, 0, default_twm<R>(), 1>;
  static constexpr int TTY = FY<false>::TT, TTX = FX<false>::TT;
  static constexpr int need = NXH * TTY > N * TTX ? NXH * TTY : N * TTX;
  static constexpr int NT0 = ((need + 63) / 64) * 64;
  static constexpr int NT = NT0 > 1024 ? 1024 : NT0;
  static constexpr int LDSN = N * NXH;
  static constexpr bool fits = (long long)LDSN * sizeof(cplx<R>) <= 163840;
};

// Team slot of the calling thread in round r of a phase with `count` items handled by
// teams of TT lanes: returns false when the caller's whole wavefront has nothing to do
// in this round; otherwise sets `item` (clamped to count - 1 for spare teams).
template <int TT>
__device__ __forceinline__ bool plane_slot(int r, int teams_per_round, int count, int& item) {
  const int team = threadIdx.x / TT;
  const int wave_first = (threadIdx.x / 64) * (64 / TT);
  if (r * teams_per_round + wave_first >= count) return false;
  const int i = r * teams_per_round + team;
  item = i < count ? i : count - 1;
  return true;
}

// y transforms of the NXH columns of a plane held in LDS (element (x, y) at x + y NXH).
template <typename R, int N, bool INV>
__device__ __forceinline__ void plane_cols(cplx<R>* pl, const cplx<R>* __restrict__ tw) {
  using PC = PlaneCfg<R, N>;
  using F = typename PC::template FY<INV>;
  constexpr int TT = F::TT, E = F::E, NXH = PC::NXH, TPR = PC::NT / TT;
  const int t = threadIdx.x % TT;
  for (int r = 0; r * TPR < NXH; ++r) {
    int col;
    if (!plane_slot<TT>(r, TPR, NXH, col)) continue;
    cplx<R> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = pl[F::in_pos(t, e) * NXH + col];
    F::run(v, pl + col, NXH, t, tw);
    F::sync();
#pragma unroll
    for (int e = 0; e < E; ++e) pl[F::out_pos(t, e) * NXH + col] = v[e];
  }
}

// Forward: rfft over (x, y) of whole planes, out = half spectrum planes (the z pass
// follows).  grid = (planes per field, fields, jobs).
template <typename R, int N>
__global__ __launch_bounds__((PlaneCfg<R, N>::NT)) void k_plane_r2c(SpecJobs jobs, const cplx<R>* __restrict__ tw_x,
                                                                 const cplx<R>* __restrict__ tw_x2,
                                                                 const cplx<R>* __restrict__ tw_y) {
  using PC = PlaneCfg<R, N>;
  using F = typename PC::template FX<false>;
  constexpr int L = PC::LX, NXH = PC::NXH, TT = F::TT, E = F::E, TPR = PC::NT / TT;
  constexpr int KN = (L + TT) / TT;  // split outputs k = t + j TT <= L per lane
  __shared__ cplx<R> pl[PC::LDSN];
  const SpecJob& jb = jobs.j[blockIdx.z];
  if ((int)blockIdx.y >= jb.count) return;
  const long long plane = (long long)blockIdx.y * gridDim.x + blockIdx.x;
  const R* src = reinterpret_cast<const R*>(jb.src) + plane * (long long)(N * N);
  cplx<R>* dst = reinterpret_cast<cplx<R>*>(jb.dst) + plane * (long long)PC::LDSN;
  const int t = threadIdx.x % TT;
  // rows: length-L complex FFT of z[j] = x[2j] + i x[2j+1], then the split
  // X[k] = (Z_k + conj Z_{L-k}) / 2 - i W^k (Z_k - conj Z_{L-k}) / 2, W = e^{-2 pi i / 2L}
  for (int r = 0; r * TPR < N; ++r) {
    int row;
    if (!plane_slot<TT>(r, TPR, N, row)) continue;
    const cplx<R>* s = reinterpret_cast<const cplx<R>*>(src + (long long)row * N);
    cplx<R>* my = pl + row * NXH;
    cplx<R> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = s[F::in_pos(t, e)];
    F::run(v, my, 1, t, tw_x);
    F::sync();
#pragma unroll
    for (int e = 0; e < E; ++e) my[F::out_pos(t, e)] = v[e];
    F::sync();
    cplx<R> xs[KN];
#pragma unroll
    for (int j = 0; j < KN; ++j) {
      const int k = t + j * TT;
      if (k <= L) {
        const cplx<R> zk = my[k == L ? 0 : k];
        const cplx<R> zm = conj(my[k == 0 ? 0 : L - k]);
        const cplx<R> ev = R(0.5) * (zk + zm);
        const cplx<R> od = mul_mi(R(0.5) * (zk - zm));
        xs[j] = ev + tw_x2[k] * od;
      }
    }
    F::sync();
#pragma unroll
    for (int j = 0; j < KN; ++j) {
      const int k = t + j * TT;
      if (k <= L) my[k] = xs[j];
    }
  }
  __syncthreads();
  plane_cols<R, N, false>(pl, tw_y);
  __syncthreads();
  for (int i = threadIdx.x; i < PC::LDSN; i += PC::NT) dst[i] = pl[i];
}

// Inverse: irfft over (x, y) of whole half-spectrum planes (after the z pass):
// y transforms, then the c2r rows that ignore the imaginary parts of the x DC and
// Nyquist bins (FFTW's c2r semantics); output scaled by `scale`.
template <typename R, int N>
__global__ __launch_bounds__((PlaneCfg<R, N>::NT)) void k_plane_c2r(SpecJobs jobs, const cplx<R>* __restrict__ tmp,
                                                                 double scale, const cplx<R>* __restrict__ tw_x,
                                                                 const cplx<R>* __restrict__ tw_x2,
                                                                 const cplx<R>* __restrict__ tw_y) {
  using PC = PlaneCfg<R, N>;
  using F = typename PC::template FX<true>;
  constexpr int L = PC::LX, NXH = PC::NXH, TT = F::TT, E = F::E, TPR = PC::NT / TT;
  __shared__ cplx<R> pl[PC::LDSN];
  const SpecJob& jb = jobs.j[blockIdx.z];
  if ((int)blockIdx.y >= jb.count) return;
  const long long plane = (long long)blockIdx.y * gridDim.x + blockIdx.x;
  const cplx<R>* src = tmp + jb.toff + plane * (long long)PC::LDSN;
  R* dst = reinterpret_cast<R*>(jb.dst) + plane * (long long)(N * N);
  for (int i = threadIdx.x; i < PC::LDSN; i += PC::NT) pl[i] = src[i];
  __syncthreads();
  plane_cols<R, N, true>(pl, tw_y);
  __syncthreads();
  const R sc = (R)scale;
  const int t = threadIdx.x % TT;
  for (int r = 0; r * TPR < N; ++r) {
    int row;
    if (!plane_slot<TT>(r, TPR, N, row)) continue;
    cplx<R>* my = pl + row * NXH;
    cplx<R> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = F::in_pos(t, e);
      cplx<R> xk = my[k];
      cplx<R> xm = my[L - k];
      if (k == 0) {
        xk.y = R(0);
        xm.y = R(0);
      }
      const cplx<R> ev = xk + conj(xm);
      const cplx<R> od = (xk - conj(xm)) * conj(tw_x2[k]);
      v[e] = ev + mul_i(od);
    }
    F::run(v, my, 1, t, tw_x);
    cplx<R>* o = reinterpret_cast<cplx<R>*>(dst + (long long)row * N);
#pragma unroll
    for (int e = 0; e < E; ++e) o[F::out_pos(t, e)] = cplx<R>{sc * v[e].x, sc * v[e].y};
  }
}

// ---------------------------------------------------------------------------
// Host dispatch
// ---------------------------------------------------------------------------
#define SAV_FFT_SIZES(X) \
  X(2) X(3) X(4) X(6) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024)

bool fft_size_supported(int n) {
  switch (n) {
#define X(L) case L:
    SAV_FFT_SIZES(X)
#undef X
    return true;
    default: return false;
  }
}

static int max_count(const SpecJobs& jobs) {
  int m = 0;
  for (int i = 0; i < jobs.n; ++i) m = jobs.j[i].count > m ? jobs.j[i].count : m;
  return m;
}

template <typename R, bool INV, int CW>
static void launch_cols_w(int L, const SpecJobs& jobs, const ColGeom& cg, const cplx<R>* tw, const KTables& kt,
                          hipStream_t s) {
  const int mc = max_count(jobs);
  switch (L) {
#define X(LL)                                                                                 \
  case LL: {                                                                                  \
    using CC = ColCfgW<R, LL, CW>;                                                            \
    dim3 grid((cg.inner + CC::C - 1) / CC::C, mc * cg.nsub, jobs.n);                          \
    hipLaunchKernelGGL((k_cols<R, LL, INV, CW>), grid, dim3(CC::NT), 0, s, jobs, cg, tw, kt); \
    break;                                                                                    \
  }
    SAV_FFT_SIZES(X)
#undef X
    default: throw HipError("unsupported FFT length " + std::to_string(L), -1);
  }
  SAV_LAUNCH_CHECK();
}

// Column transforms with the tile width chosen for the handle (0: the default rule).
template <typename R, bool INV>
static void launch_cols(int L, const SpecJobs& jobs, const ColGeom& cg, const cplx<R>* tw, const KTables& kt,
                        hipStream_t s, int colw) {
  if (colw == 4)
    launch_cols_w<R, INV, 4>(L, jobs, cg, tw, kt, s);
  else if (colw == 8)
    launch_cols_w<R, INV, 8>(L, jobs, cg, tw, kt, s);
  else if (colw == 16)
    launch_cols_w<R, INV, 16>(L, jobs, cg, tw, kt, s);
  else
    launch_cols_w<R, INV, 0>(L, jobs, cg, tw, kt, s);
}

template <typename R>
static void launch_r2c(int L, const SpecJobs& jobs, int rows_per_field, const cplx<R>* tw, const cplx<R>* tw2,
                       hipStream_t s) {
  const long long mr = (long long)max_count(jobs) * rows_per_field;
  switch (L) {
#define X(LL)                                                                                       \
  case LL: {                                                                                        \
    using CC = RowCfg<R, LL>;                                                                       \
    dim3 grid((unsigned)((mr + CC::ROWS - 1) / CC::ROWS), jobs.n);                                  \
    hipLaunchKernelGGL((k_rows_r2c<R, LL>), grid, dim3(CC::NT), 0, s, jobs, rows_per_field, tw, tw2); \
    break;                                                                                          \
  }
    SAV_FFT_SIZES(X)
#undef X
    default: throw HipError("unsupported FFT length " + std::to_string(L), -1);
  }
  SAV_LAUNCH_CHECK();
}

template <typename R>
static void launch_c2r(int L, const SpecJobs& jobs, int rows_per_field, const cplx<R>* tmp,
                       double scale, const cplx<R>* tw, const cplx<R>* tw2, hipStream_t s) {
  const long long mr = (long long)max_count(jobs) * rows_per_field;
  switch (L) {
#define X(LL)                                                                                        \
  case LL: {                                                                                         \
    using CC = RowCfg<R, LL>;                                                                        \
    dim3 grid((unsigned)((mr + CC::ROWS - 1) / CC::ROWS), jobs.n);                                   \
    hipLaunchKernelGGL((k_rows_c2r<R, LL>), grid, dim3(CC::NT), 0, s, jobs, rows_per_field, tmp,       \
                       scale, tw, tw2);                                                              \
    break;                                                                                           \
  }
    SAV_FFT_SIZES(X)
#undef X
    default: throw HipError("unsupported FFT length " + std::to_string(L), -1);
  }
  SAV_LAUNCH_CHECK();
}

template <typename R, bool INV>
static void launch_rows_c2c(int L, cplx<R>* data, long long nrows, double scale, const cplx<R>* tw, hipStream_t s) {
  switch (L) {
#define X(LL)                                                                                     \
  case LL: {                                                                                      \
    using CC = RowCfg<R, LL>;                                                                     \
    dim3 grid((unsigned)((nrows + CC::ROWS - 1) / CC::ROWS));                                     \
    hipLaunchKernelGGL((k_rows_c2c<R, LL, INV>), grid, dim3(CC::NT), 0, s, data, nrows, scale, tw); \
    break;                                                                                        \
  }
    SAV_FFT_SIZES(X)
#undef X
    default: throw HipError("unsupported FFT length " + std::to_string(L), -1);
  }
  SAV_LAUNCH_CHECK();
}

// Column geometry of the y and z passes over half spectra.
template <typename R>
static ColGeom geom_y(const FftGeom<R>& g) {
  ColGeom c{};
  c.inner = g.Nxh;
  c.stride = g.Nxh;
  c.nsub = g.ndim == 3 ? g.Nz : 1;
  c.sub_stride = (long long)g.Nxh * g.Ny;
  c.field_stride = (long long)g.Nxh * g.Ny * (g.ndim == 3 ? g.Nz : 1);
  c.geom = g.ndim == 3 ? 2 : 0;
  c.Nxh = g.Nxh;
  return c;
}

template <typename R>
static ColGeom geom_z(const FftGeom<R>& g) {
  ColGeom c{};
  c.inner = g.Nxh * g.Ny;
  c.stride = g.Nxh * g.Ny;
  c.nsub = 1;
  c.sub_stride = 0;
  c.field_stride = (long long)g.Nxh * g.Ny * g.Nz;
  c.geom = 1;
  c.Nxh = g.Nxh;
  return c;
}

// Plane sizes with fused (x, y) kernels: N x N planes of 3D fields.
#define SAV_PLANE_SIZES(X) X(16) X(32) X(64) X(96) X(128)

template <typename R>
static bool plane_path(const FftGeom<R>& g) {
  if (g.ndim != 3 || g.Nx != g.Ny) return false;
  switch (g.Nx) {
#define X(NN) \
  case NN: return PlaneCfg<R, NN>::fits;
    SAV_PLANE_SIZES(X)
#undef X
    default: return false;
  }
}

template <typename R>
static void launch_plane(const FftGeom<R>& g, const SpecJobs& jobs, bool inverse, const cplx<R>* tmp, double scale,
                         hipStream_t s) {
  const dim3 grid(g.Nz, max_count(jobs), jobs.n);
  switch (g.Nx) {
#define X(NN)                                                                                              \
  case NN: {                                                                                               \
    if constexpr (PlaneCfg<R, NN>::fits) {                                                                 \
      if (inverse)                                                                                         \
        hipLaunchKernelGGL((k_plane_c2r<R, NN>), grid, dim3(PlaneCfg<R, NN>::NT), 0, s, jobs, tmp, scale,    \
                           g.tw_xh, g.tw_x2, g.tw_y);                                                      \
      else                                                                                                 \
        hipLaunchKernelGGL((k_plane_r2c<R, NN>), grid, dim3(PlaneCfg<R, NN>::NT), 0, s, jobs, g.tw_xh,       \
                           g.tw_x2, g.tw_y);                                                               \
    }                                                                                                      \
    break;                                                                                                 \
  }
    SAV_PLANE_SIZES(X)
#undef X
    default: throw HipError("no plane kernel for N = " + std::to_string(g.Nx), -1);
  }
  SAV_LAUNCH_CHECK();
}

template <typename R>
void rfft_jobs(const FftGeom<R>& g, const SpecJobs& jobs, hipStream_t s) {
  if (jobs.n == 0) return;
  if (plane_path(g)) {
    launch_plane<R>(g, jobs, false, nullptr, 1.0, s);
    ColGeom cz = geom_z(g);
    cz.use_src = 0;
    launch_cols<R, false>(g.Nz, jobs, cz, g.tw_z, g.k, s, g.colw);
    return;
  }
  const int rows = g.Ny * (g.ndim == 3 ? g.Nz : 1);
  launch_r2c<R>(g.Nx / 2, jobs, rows, g.tw_xh, g.tw_x2, s);
  ColGeom cy = geom_y(g);
  cy.use_src = 0;
  launch_cols<R, false>(g.Ny, jobs, cy, g.tw_y, g.k, s, g.colw);
  if (g.ndim == 3) {
    ColGeom cz = geom_z(g);
    cz.use_src = 0;
    launch_cols<R, false>(g.Nz, jobs, cz, g.tw_z, g.k, s, g.colw);
  }
}

template <typename R>
void irfft_jobs(const FftGeom<R>& g, const SpecJobs& jobs_in, cplx<R>* tmp, hipStream_t s) {
  if (jobs_in.n == 0) return;
  const long long nh = (long long)g.Nxh * g.Ny * (g.ndim == 3 ? g.Nz : 1);
  // column passes write into tmp; the row c2r pass reads tmp + toff
  SpecJobs jobs = jobs_in;
  SpecJobs tj = jobs_in;
  long long off = 0;
  for (int i = 0; i < jobs.n; ++i) {
    jobs.j[i].toff = off;
    tj.j[i].dst = tmp + off;
    off += nh * jobs.j[i].count;
  }
  if (g.ndim == 2) {
    ColGeom cy = geom_y(g);
    cy.use_src = 1;
    launch_cols<R, true>(g.Ny, tj, cy, g.tw_y, g.k, s, g.colw);
  } else {
    ColGeom cz = geom_z(g);
    cz.use_src = 1;
    launch_cols<R, true>(g.Nz, tj, cz, g.tw_z, g.k, s, g.colw);
    if (plane_path(g)) {
      launch_plane<R>(g, jobs, true, tmp, 1.0 / ((double)g.Nx * g.Ny * g.Nz), s);
      return;
    }
    ColGeom cy = geom_y(g);
    cy.use_src = 0;
    launch_cols<R, true>(g.Ny, tj, cy, g.tw_y, g.k, s, g.colw);
  }
  const int rows = g.Ny * (g.ndim == 3 ? g.Nz : 1);
  const double scale = 1.0 / ((double)g.Nx * g.Ny * (g.ndim == 3 ? g.Nz : 1));
  launch_c2r<R>(g.Nx / 2, jobs, rows, tmp, scale, g.tw_xh, g.tw_x2, s);
}

template <typename R>
void fft_full_2d(const FftGeom<R>& g, cplx<R>* data, int count, bool inverse, hipStream_t s) {
  const long long nrows = (long long)count * g.Ny;
  SpecJobs jobs{};
  jobs.n = 1;
  jobs.j[0].dst = data;
  jobs.j[0].count = count;
  jobs.j[0].scale = 1.0;
  ColGeom c{};
  c.inner = g.Nx;
  c.stride = g.Nx;
  c.nsub = 1;
  c.field_stride = (long long)g.Nx * g.Ny;
  c.geom = 2;
  c.use_src = 0;
  if (!inverse) {
    launch_rows_c2c<R, false>(g.Nx, data, nrows, 1.0, g.tw_x, s);
    launch_cols<R, false>(g.Ny, jobs, c, g.tw_y, g.k, s, g.colw);
  } else {
    launch_cols<R, true>(g.Ny, jobs, c, g.tw_y, g.k, s, g.colw);
    launch_rows_c2c<R, true>(g.Nx, data, nrows, 1.0 / ((double)g.Nx * g.Ny), g.tw_x, s);
  }
}

template void rfft_jobs<double>(const FftGeom<double>&, const SpecJobs&, hipStream_t);
template void rfft_jobs<float>(const FftGeom<float>&, const SpecJobs&, hipStream_t);
template void irfft_jobs<double>(const FftGeom<double>&, const SpecJobs&, cplx<double>*, hipStream_t);
template void irfft_jobs<float>(const FftGeom<float>&, const SpecJobs&, cplx<float>*, hipStream_t);
template void fft_full_2d<double>(const FftGeom<double>&, cplx<double>*, int, bool, hipStream_t);
template void fft_full_2d<float>(const FftGeom<float>&, cplx<float>*, int, bool, hipStream_t);

}  // namespace sav
