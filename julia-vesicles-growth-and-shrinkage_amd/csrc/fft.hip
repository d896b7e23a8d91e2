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
// k_cols is instantiated with CW = the resolved column count (ColCfgW<R, L, CW>::C is a
// fixed point), so widths that resolve to the same tile share one kernel.
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
    X(2) X(3) X(4) X(5) X(6) X(8) X(10) X(12) X(16) X(20) X(24) X(32) X(40) X(48) X(64) X(80) X(96) X(128) X(160) X(192) X(256) X(320) X(384) X(512) X(640) X(768) X(1024) X(1280) X(2048)
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
// Register-column (x, y) transforms of 128 x 128 planes.
//
// The LDS-plane kernels above hold a whole half-spectrum plane (65 x 128 complex: 133 KB
// in FP64), so one FP64 workgroup fills a CU and its load, transform and store phases
// cannot overlap with another's.  These kernels keep the columns in registers and use
// LDS only as a 64 KB half plane (64 rows x 64 columns), so two FP64 workgroups (four
// FP32 ones) share a CU.
//
// 512 threads = 8 wavefronts; lane l of wavefront w is (c, t) = (l & 7, l >> 3) in the
// column phase and (s, rl) = (l & 7, l >> 3) in the row phase.
//  * columns: thread (w, c, t) holds column x = 8w + c at rows y = t + 8m (m < 16) and
//    runs the length-128 y transform as 16 x 8: a 16-point DFT over m in registers, the
//    twiddles W_128^{t k1}, then for each k1 an 8-point DFT over t after an exchange
//    through the wavefront's own 8 KB of LDS (two rounds of 8 k1 values), leaving the
//    rows ky = 8h + t + 16 k2 (h < 2, k2 < 8) in registers.
//  * the x DC and Nyquist columns (x = 0 and x = 64) travel packed in column 0 as
//    c0 + i c64: both are real-valued along x after the y transform (c2r: the imaginary
//    parts of those two bins are ignored; r2c: a real row has real DC and Nyquist bins),
//    so one complex y transform carries both.  c2r packs the Hermitian parts
//    (c(ky) + conj c(-ky)) / 2 on load; r2c separates F(ky) and conj F(-ky) on store.
//  * rows, one half plane at a time: row rho of the half (rho = 8w + rl) is carried by
//    the 8 lanes s of (w, rl) as 64 = 8 x 8: an 8-point DFT over m of the elements
//    s + 8m, twiddles W_64^{s k1}, an exchange through the row's own LDS slots (XOR
//    swizzled: conflict-free both ways), and an 8-point DFT; the real r2c / c2r split
//    of length 128 over the length-64 complex transform as in k_rows_r2c / k_rows_c2r.
//  * every 64-slot LDS row goes through hpi(), which moves odd rows by half a bank
//    window: gfx950 serves a 16-byte LDS access in groups of 16 lanes (two teams, or two
//    rows), and without it those pairs collide on the same banks.
// Every LDS exchange inside a phase stays within one wavefront (wavefront barriers);
// the half-plane hand-overs between phases use workgroup barriers.
// ---------------------------------------------------------------------------
constexpr int kHPlaneNT = 512;

// LDS slot of element `col` (< 64) of row r of a 64-slot-row region: bit 3 of the column
// flips with the row's parity, so two adjacent rows accessed by one 16-lane group (8
// consecutive or XOR-permuted slots each) fall on opposite halves of the 64 banks.
__device__ __forceinline__ int hpi(int r, int col) { return r * 64 + (col ^ ((r & 1) << 3)); }

// Twiddles W_128^{t k1} (conjugated for INV) on v[8h + j], k1 = 8h + j, then the 8-point
// DFT over t of k1 = 8h + t after the exchange through the wavefront's region ex (value
// k1 = 8h + j of team t of column c in row j, slot 8t + c): out r[k2] = y[8h + t + 16 k2].  The twiddles are applied here, on
// the way to LDS, so that they are not live across the 16-point DFT.
template <typename R, bool INV>
__device__ __forceinline__ void hplane_ycombine(const cplx<R> (&v)[16], int h, cplx<R>* ex, int c, int t,
                                                const cplx<R>* tws, cplx<R> (&r)[8]) {
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int k1 = 8 * h + j;
    cplx<R> a = v[k1];
    if (k1 > 0) {
      const cplx<R> w = tws[t * 24 + k1];
      a = a * (INV ? conj(w) : w);
    }
    ex[hpi(j, 8 * t + c)] = a;
  }
  wave_lds_sync();
#pragma unroll
  for (int tp = 0; tp < 8; ++tp) r[tp] = ex[hpi(t, 8 * tp + c)];
  bfly<R, INV, 8>(r);
  wave_lds_sync();
}

// Forward: rfft over (x, y) of 128 x 128 real planes into half-spectrum planes (the z
// pass follows).  grid = (planes per field, fields, jobs).
template <typename R>
__global__ __launch_bounds__(kHPlaneNT, 4) void k_hplane_r2c(SpecJobs jobs, const cplx<R>* __restrict__ twp) {
  __shared__ cplx<R> lds[64 * 64];
  __shared__ cplx<R> tws[kHPlaneTw];
  __shared__ cplx<R> pk[128];  // the packed DC / Nyquist column after the y transform
  const SpecJob& jb = jobs.j[blockIdx.z];
  if ((int)blockIdx.y >= jb.count) return;
  const long long plane = (long long)blockIdx.y * gridDim.x + blockIdx.x;
  const R* src = reinterpret_cast<const R*>(jb.src) + plane * (128LL * 128);
  cplx<R>* dst = reinterpret_cast<cplx<R>*>(jb.dst) + plane * (65LL * 128);
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int c = lane & 7, t = lane >> 3;  // column phase
  const int s = c, rho = 8 * w + t;       // row phase: lane s of row rho of the half plane
  if (tid < kHPlaneTw) tws[tid] = twp[tid];
  cplx<R> v[16];
  // z[n] = x[2n] + i x[2n+1] of rows y = rho and 64 + rho, n = s + 8m: both halves are
  // loaded up front, so the second half's loads overlap the first half's transforms
  cplx<R> zz[2][8];
#pragma unroll
  for (int hh = 0; hh < 2; ++hh) {
    const cplx<R>* zr = reinterpret_cast<const cplx<R>*>(src + (64 * hh + rho) * 128);
#pragma unroll
    for (int m = 0; m < 8; ++m) zz[hh][m] = zr[s + 8 * m];
  }
#pragma unroll
  for (int hh = 0; hh < 2; ++hh) {
    cplx<R> z[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) z[m] = zz[hh][m];
    __syncthreads();  // twiddles loaded; the previous half's column reads are done
    bfly<R, false, 8>(z);
#pragma unroll
    for (int k1 = 1; k1 < 8; ++k1) z[k1] = z[k1] * tws[kHPlaneTX + k1 * 8 + s];
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1) lds[hpi(rho, k1 * 8 + (s ^ k1))] = z[k1];
    wave_lds_sync();
    cplx<R> u[8];
#pragma unroll
    for (int sp = 0; sp < 8; ++sp) u[sp] = lds[hpi(rho, s * 8 + (sp ^ s))];
    bfly<R, false, 8>(u);  // u[k2] = Z[s + 8 k2]
    wave_lds_sync();
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) lds[hpi(rho, s + 8 * k2)] = u[k2];
    wave_lds_sync();
    // X[k] = (Z_k + conj Z_{64-k}) / 2 - i W^k (Z_k - conj Z_{64-k}) / 2, and slot 0
    // holds X[0] + i X[64] = (Re Z0 + Im Z0) + i (Re Z0 - Im Z0).  Z_k is read back from
    // LDS rather than kept in u: the first half's columns are live here as well.
    cplx<R> xs[8];
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) {
      const int k = s + 8 * k2;
      const cplx<R> zk = lds[hpi(rho, k)];
      const cplx<R> zm = conj(lds[hpi(rho, (64 - k) & 63)]);
      const cplx<R> ev = R(0.5) * (zk + zm);
      const cplx<R> od = mul_mi(R(0.5) * (zk - zm));
      const cplx<R> gen = ev + tws[kHPlaneT2 + k] * od;
      xs[k2] = k == 0 ? cplx<R>{zk.x + zk.y, zk.x - zk.y} : gen;
    }
    wave_lds_sync();
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) lds[hpi(rho, s + 8 * k2)] = xs[k2];
    __syncthreads();
    // column x = 8w + c at rows y = t + 8m, m = 8 hh + mp
    const int x = 8 * w + c;
#pragma unroll
    for (int mp = 0; mp < 8; ++mp) v[8 * hh + mp] = lds[hpi(t + 8 * mp, x)];
  }
  bfly<R, false, 16>(v);
  __syncthreads();  // every column read of the half plane is done: the exchange regions reuse it
  cplx<R>* ex = lds + w * 512;
  const int x = 8 * w + c;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    cplx<R> r[8];
    hplane_ycombine<R, false>(v, h, ex, c, t, tws, r);
    if (x != 0) {
#pragma unroll
      for (int k2 = 0; k2 < 8; ++k2) dst[x + 65 * (8 * h + t + 16 * k2)] = r[k2];
    } else {
#pragma unroll
      for (int k2 = 0; k2 < 8; ++k2) pk[8 * h + t + 16 * k2] = r[k2];
    }
  }
  if (w == 0) {
    // packed column: F(ky) = A(ky) + i B(ky) with A, B the spectra of the real DC and
    // Nyquist columns; A = (F(ky) + conj F(-ky)) / 2, B = -i (F(ky) - conj F(-ky)) / 2
    wave_lds_sync();
    if (c == 0) {
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int k2 = 0; k2 < 8; ++k2) {
          const int ky = 8 * h + t + 16 * k2;
          const cplx<R> fk = pk[ky];
          const cplx<R> fm = conj(pk[(128 - ky) & 127]);
          dst[65 * ky] = R(0.5) * (fk + fm);
          dst[64 + 65 * ky] = mul_mi(R(0.5) * (fk - fm));
        }
    }
  }
}

// Flat plane index q over the jobs (count_j fields of nz planes each) -> job and plane.
__device__ __forceinline__ bool hplane_locate(const SpecJobs& jobs, int nz, long long q, int& j, long long& plane) {
  for (int i = 0; i < jobs.n; ++i) {
    const long long np = (long long)jobs.j[i].count * nz;
    if (q < np) {
      j = i;
      plane = q;
      return true;
    }
    q -= np;
  }
  return false;
}

// c2r column loads v[M0 .. M0 + 7] of column x (rows y = t + 8m) of the half-spectrum plane
// at src; column 0 takes c0 + i c64 from the Hermitian parts (c(y) + conj c(-y)) / 2.
// With pm = MUL_DX / MUL_DY the plane is that of a derivative, i kx X or i ky X, whose z pass
// is the plain field's (kx = x dkx, ky = (y < 64 ? y : y - 128) dky): the other columns are
// multiplied after the loads (hplane_c2r_mul); on the packed column, i ky commutes with the
// Hermitian parts except at the Nyquist row y = 64, set here to -ky(64) (Im c0 + i Im c64),
// and i kx (0 at x = 0, kN at x = 64) leaves -kN times the anti-Hermitian part of c64.
template <typename R, int M0>
__device__ __forceinline__ void hplane_c2r_load(const cplx<R>* __restrict__ src, int x, int t, cplx<R> (&v)[16], int pm,
                                                double dkx, double dky) {
  if (x != 0) {
#pragma unroll
    for (int mm = 0; mm < 8; ++mm) v[M0 + mm] = src[x + 65 * (t + 8 * (M0 + mm))];
  } else {
#pragma unroll
    for (int mm = 0; mm < 8; ++mm) {
      const int y = t + 8 * (M0 + mm), ym = (128 - y) & 127;
      const cplx<R> a = src[65 * y], am = src[65 * ym];
      const cplx<R> b = src[64 + 65 * y], bm = src[64 + 65 * ym];
      if (pm == MUL_DX) {
        const R kn = (R)(dkx * 64.0);
        v[M0 + mm] = cplx<R>{-kn * (R(0.5) * (b.x - bm.x)), -kn * (R(0.5) * (b.y + bm.y))};
      } else if (pm == MUL_DY && y == 64) {
        const R kq = (R)(dky * -64.0);
        v[M0 + mm] = cplx<R>{-kq * a.y, -kq * b.y};
      } else {
        v[M0 + mm] = cplx<R>{R(0.5) * (a.x + am.x) - R(0.5) * (b.y - bm.y), R(0.5) * (a.y - am.y) + R(0.5) * (b.x + bm.x)};
      }
    }
  }
}

template <typename R>
__device__ __forceinline__ void hplane_c2r_mul(cplx<R> (&v)[16], int x, int t, int pm, double dkx, double dky) {
  if (pm == MUL_NONE || (x == 0 && pm == MUL_DX)) return;
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const int y = t + 8 * e;
    if (x == 0 && y == 64) continue;  // set by the load
    const R k = (R)(pm == MUL_DX ? dkx * (double)x : dky * (double)(y < 64 ? y : y - 128));
    v[e] = cplx<R>{-k * v[e].y, k * v[e].x};
  }
}

// Inverse: irfft over (x, y) of 128 x 128 half-spectrum planes (after the z pass), with
// FFTW's c2r semantics (the imaginary parts of the x DC and Nyquist bins are ignored
// after the y transform); output scaled by `scale`.  Workgroup b transforms plane b of the
// `total` planes of the jobs.  (A persistent form that loads the next plane into registers
// during the row phase does not fit: the row phase already holds ~118 VGPRs, and two
// workgroups per CU allow 128.)
template <typename R>
__global__ __launch_bounds__(kHPlaneNT, 4) void k_hplane_c2r(SpecJobs jobs, int nz, long long total,
                                                             const cplx<R>* __restrict__ tmp, double scale,
                                                             const cplx<R>* __restrict__ twp, double dkx,
                                                             double dky) {
  __shared__ cplx<R> lds[64 * 64];
  __shared__ cplx<R> tws[kHPlaneTw];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int c = lane & 7, t = lane >> 3;
  const int x = 8 * w + c;
  const int s = c, rho = 8 * w + t;  // row phase: lane s of row rho of the half plane
  cplx<R>* ex = lds + w * 512;
  if (tid < kHPlaneTw) tws[tid] = twp[tid];
  const R sc = (R)scale;
  {
    const long long q = blockIdx.x;
    int j = 0;
    long long plane = 0;
    if (q >= total || !hplane_locate(jobs, nz, q, j, plane)) return;
    const cplx<R>* src = tmp + jobs.j[j].toff + plane * (65LL * 128);
    R* dst = reinterpret_cast<R*>(jobs.j[j].dst) + plane * (128LL * 128);
    cplx<R> v[16];
    const int pm = jobs.j[j].pmul;
    hplane_c2r_load<R, 0>(src, x, t, v, pm, dkx, dky);
    hplane_c2r_load<R, 8>(src, x, t, v, pm, dkx, dky);
    hplane_c2r_mul<R>(v, x, t, pm, dkx, dky);
    __syncthreads();  // twiddles
    bfly<R, true, 16>(v);
    cplx<R> r0[8], r1[8];
    hplane_ycombine<R, true>(v, 0, ex, c, t, tws, r0);
    hplane_ycombine<R, true>(v, 1, ex, c, t, tws, r1);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      __syncthreads();  // the exchange regions / the previous half are no longer read
#pragma unroll
      for (int k2 = 0; k2 < 8; ++k2) lds[hpi(k2 * 8 + t, x)] = h == 0 ? r0[k2] : r1[k2];
      __syncthreads();
      const int ky = 8 * h + (rho & 7) + 16 * (rho >> 3);
      cplx<R> z[8];
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        const int k = s + 8 * m;
        const cplx<R> pk = lds[hpi(rho, k)];
        const cplx<R> pm = lds[hpi(rho, (64 - k) & 63)];
        const cplx<R> xk = k == 0 ? cplx<R>{pk.x, R(0)} : pk;
        const cplx<R> xm = k == 0 ? cplx<R>{pk.y, R(0)} : pm;
        const cplx<R> ev = xk + conj(xm);
        const cplx<R> od = (xk - conj(xm)) * conj(tws[kHPlaneT2 + k]);
        z[m] = ev + mul_i(od);
      }
      bfly<R, true, 8>(z);
#pragma unroll
      for (int k1 = 1; k1 < 8; ++k1) z[k1] = z[k1] * conj(tws[kHPlaneTX + k1 * 8 + s]);
      wave_lds_sync();  // the wavefront's rows are read
#pragma unroll
      for (int k1 = 0; k1 < 8; ++k1) lds[hpi(rho, k1 * 8 + (s ^ k1))] = z[k1];
      wave_lds_sync();
      cplx<R> u[8];
#pragma unroll
      for (int sp = 0; sp < 8; ++sp) u[sp] = lds[hpi(rho, s * 8 + (sp ^ s))];
      bfly<R, true, 8>(u);  // u[k2] = z[s + 8 k2] = x[2k] + i x[2k+1], k = s + 8 k2
      cplx<R>* o = reinterpret_cast<cplx<R>*>(dst + ky * 128);
#pragma unroll
      for (int k2 = 0; k2 < 8; ++k2) o[s + 8 * k2] = cplx<R>{sc * u[k2].x, sc * u[k2].y};
    }
    wave_lds_sync();
  }
}

// ---------------------------------------------------------------------------
// Host dispatch
// ---------------------------------------------------------------------------
#define SAV_FFT_SIZES(X) \
  X(2) X(3) X(4) X(5) X(6) X(8) X(10) X(12) X(16) X(20) X(24) X(32) X(40) X(48) X(64) X(80) X(96) X(128) X(160) X(192) X(256) X(320) X(384) X(512) X(640) X(768) X(1024) X(1280) X(2048)

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
    hipLaunchKernelGGL((k_cols<R, LL, INV, CC::C>), grid, dim3(CC::NT), 0, s, jobs, cg, tw, kt); \
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
  if (g.Nx == 128 && g.tw_hp && g.plane_kind == 0) {
    if (inverse) {
      long long total = 0;
      for (int i = 0; i < jobs.n; ++i) total += (long long)jobs.j[i].count * g.Nz;
      if (total == 0) return;
      const dim3 pgrid((unsigned)total);
      hipLaunchKernelGGL((k_hplane_c2r<R>), pgrid, dim3(kHPlaneNT), 0, s, jobs, g.Nz, total, tmp, scale, g.tw_hp,
                         g.dkx, g.dky);
    } else
      hipLaunchKernelGGL((k_hplane_r2c<R>), grid, dim3(kHPlaneNT), 0, s, jobs, g.tw_hp);
    SAV_LAUNCH_CHECK();
    return;
  }
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
    launch_cols<R, false>(g.Nz, jobs, cz, g.tw_z, g.k, s, g.colw_fwd);
    return;
  }
  const int rows = g.Ny * (g.ndim == 3 ? g.Nz : 1);
  launch_r2c<R>(g.Nx / 2, jobs, rows, g.tw_xh, g.tw_x2, s);
  ColGeom cy = geom_y(g);
  cy.use_src = 0;
  launch_cols<R, false>(g.Ny, jobs, cy, g.tw_y, g.k, s, g.colw_fwd);
  if (g.ndim == 3) {
    ColGeom cz = geom_z(g);
    cz.use_src = 0;
    launch_cols<R, false>(g.Nz, jobs, cz, g.tw_z, g.k, s, g.colw_fwd);
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
  if (g.ndim == 3 && plane_path(g) && g.Nx == 128 && g.tw_hp && g.plane_kind == 0) {
    // i kx and i ky are constant along z: such a job's z pass is the plain field's, and jobs
    // on the same source share it; the plane pass applies the multiplier (pmul)
    tj.n = 0;
    for (int i = 0; i < jobs.n; ++i) {
      SpecJob& jb = jobs.j[i];
      const bool zc = jb.mul == MUL_NONE || jb.mul == MUL_DX || jb.mul == MUL_DY;
      jb.pmul = zc ? jb.mul : MUL_NONE;
      int share = -1;
      if (zc)
        for (int k = 0; k < i; ++k) {
          const SpecJob& o = jobs.j[k];
          if ((o.mul == MUL_NONE || o.mul == MUL_DX || o.mul == MUL_DY) && o.src == jb.src && o.src2 == jb.src2 &&
              o.count == jb.count && o.scale == jb.scale) {
            share = k;
            break;
          }
        }
      if (share >= 0) {
        jb.toff = jobs.j[share].toff;
        continue;
      }
      jb.toff = off;
      SpecJob z = jb;
      z.mul = zc ? MUL_NONE : jb.mul;
      z.dst = tmp + off;
      tj.j[tj.n++] = z;
      off += nh * jb.count;
    }
    ColGeom cz = geom_z(g);
    cz.use_src = 1;
    launch_cols<R, true>(g.Nz, tj, cz, g.tw_z, g.k, s, g.colw);
    launch_plane<R>(g, jobs, true, tmp, 1.0 / ((double)g.Nx * g.Ny * g.Nz), s);
    return;
  }
  for (int i = 0; i < jobs.n; ++i) {
    jobs.j[i].toff = off;
    jobs.j[i].pmul = MUL_NONE;
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
    launch_cols<R, false>(g.Ny, jobs, c, g.tw_y, g.k, s, g.colw_fwd);
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
