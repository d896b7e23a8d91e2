// Batched real/complex FFT kernels over the spatial axes of (Nx, Ny[, Nz]) fields
// stored x-fastest (the Julia column-major layout).  See fft.h for the team FFT.
#include "fft.h"

namespace sav {

// Columns handled per workgroup for a column (strided) transform of length L.
template <typename R, int L>
using ColCfg = ColTile<R, L>;

// Rows handled per workgroup for a contiguous transform of length L.
template <typename R, int L>
struct RowCfg {
  using F = TeamFFT<R, L, false>;
  static constexpr int TT = F::TT;
  static constexpr int ROWS = TT >= 256 ? 1 : 256 / TT;
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
template <typename R, int L, bool INV>
__global__ __launch_bounds__((ColCfg<R, L>::NT)) void k_cols(SpecJobs jobs, ColGeom cg,
                                                          const cplx<R>* __restrict__ tw, KTables kt) {
  using F = TeamFFT<R, L, INV>;
  constexpr int C = ColCfg<R, L>::C;
  constexpr int E = F::E;
  __shared__ cplx<R> lds[C * L];
  const SpecJob& jb = jobs.j[blockIdx.z];
  const int fidx = blockIdx.y / cg.nsub;
  const int sub = blockIdx.y % cg.nsub;
  if (fidx >= jb.count) return;
  const int c = threadIdx.x % C;
  const int t = threadIdx.x / C;
  const int col = blockIdx.x * C + c;
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

template <typename R, bool INV>
static void launch_cols(int L, const SpecJobs& jobs, const ColGeom& cg, const cplx<R>* tw, const KTables& kt,
                        hipStream_t s) {
  const int mc = max_count(jobs);
  switch (L) {
#define X(LL)                                                                                 \
  case LL: {                                                                                  \
    using CC = ColCfg<R, LL>;                                                                 \
    dim3 grid((cg.inner + CC::C - 1) / CC::C, mc * cg.nsub, jobs.n);                          \
    hipLaunchKernelGGL((k_cols<R, LL, INV>), grid, dim3(CC::NT), 0, s, jobs, cg, tw, kt);     \
    break;                                                                                    \
  }
    SAV_FFT_SIZES(X)
#undef X
    default: throw HipError("unsupported FFT length " + std::to_string(L), -1);
  }
  SAV_LAUNCH_CHECK();
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

template <typename R>
void rfft_jobs(const FftGeom<R>& g, const SpecJobs& jobs, hipStream_t s) {
  if (jobs.n == 0) return;
  const int rows = g.Ny * (g.ndim == 3 ? g.Nz : 1);
  launch_r2c<R>(g.Nx / 2, jobs, rows, g.tw_xh, g.tw_x2, s);
  ColGeom cy = geom_y(g);
  cy.use_src = 0;
  launch_cols<R, false>(g.Ny, jobs, cy, g.tw_y, g.k, s);
  if (g.ndim == 3) {
    ColGeom cz = geom_z(g);
    cz.use_src = 0;
    launch_cols<R, false>(g.Nz, jobs, cz, g.tw_z, g.k, s);
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
    launch_cols<R, true>(g.Ny, tj, cy, g.tw_y, g.k, s);
  } else {
    ColGeom cz = geom_z(g);
    cz.use_src = 1;
    launch_cols<R, true>(g.Nz, tj, cz, g.tw_z, g.k, s);
    ColGeom cy = geom_y(g);
    cy.use_src = 0;
    launch_cols<R, true>(g.Ny, tj, cy, g.tw_y, g.k, s);
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
    launch_cols<R, false>(g.Ny, jobs, c, g.tw_y, g.k, s);
  } else {
    launch_cols<R, true>(g.Ny, jobs, c, g.tw_y, g.k, s);
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
