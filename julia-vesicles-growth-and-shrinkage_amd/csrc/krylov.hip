// BiCGSTAB for the variable-mobility psi operator (2D), restated from the
// algorithm of Krylov.jl 0.10.6 `bicgstab` as called at src/Solvers.jl:73-75:
// left preconditioner M = P^{-1} applied by multiplication, shadow vector c = b,
// x0 = 0, eps = atol + rtol*||M b||, stop on ||r|| <= eps, iter >= itmax, alpha == 0.
//
// Operator (src/Solvers.jl:39-53), on full complex spectra of N stacked fields:
//   A x = (a/dt) x - sum_d D1f_d .* FFT( M_psi .* IFFT( D1f_d .* (S4 x) ) )
// Preconditioner (src/Solvers.jl:57): P^{-1} = 1 / (a/dt - Mmax * Lap_full * S4).
//
// A matvec is three kernels: column inverse FFTs of D1f_x S4 x and D1f_y S4 x,
// one row kernel doing inverse FFT -> *M_psi -> forward FFT, and column forward
// FFTs fused with the combination, the preconditioner and the next dot products.
// All Krylov scalars stay on the device; every kernel of an iteration returns at
// once when the solve has finished, so iterations can be queued ahead of the
// host's convergence check.
#include <algorithm>
#include <cmath>
#include <vector>

#include <hip/hip_ext.h>

#include "krylov.h"
#include "launch.h"

namespace sav {

// Column kernels of the matvec: CW adjacent columns per workgroup (4, 8 or 16; chosen per
// handle by timing them at the first solve, see autotune_columns), the column plan of
// col_plan_maxr / col_plan_ttmax (FWD: the forward kernel's plan) with k-major twiddles
// (KrylovCtx::tw_ci, tw_cf).
// The kernels are instantiated with CW = the tile's resolved column count (KColCfg<R, L, CW>::C
// is a fixed point), so widths that resolve to the same tile share one kernel.
template <typename R, int L, int CW = 8, bool FWD = false>
using KColCfg = ColTile<R, L, CW, col_plan_maxr<R>(L, FWD), col_plan_ttmax<R>(L, FWD)>;
template <typename R, int L, bool INV>
using KColFFT = TeamFFT<R, L, INV, col_plan_ttmax<R>(L, !INV), 0, 2, 0, col_plan_maxr<R>(L, !INV)>;
// Registers of the column-inverse kernel: radix-32 teams hold 32 elements per lane (one
// wavefront per SIMD); FP64 64-lane teams hold S4 x and a transform of 16 complex doubles
// per lane, at two wavefronts per SIMD (the LDS holds one workgroup per CU anyway).
template <typename R, int L>
constexpr int kcol_min_waves() {
  return col_plan_maxr<R>(L, false) == 32 ? 1 : (sizeof(R) == 8 ? 2 : 4);
}

// Row kernel: see KRowCfg; the LDS of each row is padded by one slot every 16 positions
// against bank conflicts.
// In FP64, where a team is exactly one wavefront, its LDS exchanges synchronise the
// wavefront only (no workgroup barrier; measured faster in FP64 only).
template <typename R, int L>
struct KRowCfg {
  // Radix-16 passes for power-of-two rows (fewer LDS exchanges), k-major twiddle tables.
  // The forward transform runs the inverse's plan reversed, so its first pass reads
  // exactly the elements the inverse's last pass left in each lane: the M_psi product is
  // taken in registers, with no LDS exchange between the two transforms.  One wavefront
  // per field (64 / TT rows), synchronised at wavefront level.
  static constexpr int MAXR = row_plan_maxr(L);
  static constexpr int PADS = 4;
  using FI = TeamFFT<R, L, true, 64, PADS, 2, 1, MAXR, false>;
  using FF = TeamFFT<R, L, false, 64, PADS, 2, 1, MAXR, true>;
  static_assert(FF::R0 == FI::RL && FF::TT == FI::TT && FF::E == FI::E, "reversed plan");
  static constexpr int TT = FI::TT;
  static constexpr int ROWS = 64 / TT;
  static constexpr int NT = 128;  // two wavefronts: the Bx rows and the By rows
  static constexpr int RS = FI::LDS_LEN | 1;   // odd row stride
};

__device__ __forceinline__ double pinv_diag(double a_dt, double Mmax, double S4, double kx0, double ky) {
  const double lapf = -(kx0 * kx0 + ky * ky);
  return 1.0 / (a_dt - Mmax * lapf * S4);
}

// ---------------------------------------------------------------------------
// K1: column inverse FFTs of D1f_x*(S4 x) and D1f_y*(S4 x)
// ---------------------------------------------------------------------------
// UPD selects the vector x the matvec is applied to, formed here from BiCGSTAB vectors as
// each element is loaded (the vector update then needs no pass of its own):
//   0: x = X;
//   1: x = s = r - alpha v   (X = r, V = v; s is written to Xo);
//   2: x = p = r + beta (p - omega v)   (X = r, V = v, Xo = p, updated in place).
// The updates' arithmetic is that of k_kry_upd1s / k_kry_upd3.
template <typename R, int L, int CW, int UPD>
__global__ __launch_bounds__((KColCfg<R, L, CW>::NT), (kcol_min_waves<R, L>())) void k_mv_cols_inv(const cplx<R>* __restrict__ X,
                                                                    const cplx<R>* __restrict__ V,
                                                                    cplx<R>* Xo,
                                                                    cplx<R>* __restrict__ Bx,
                                                                    cplx<R>* __restrict__ By, int Nx, double S4,
                                                                    const double* __restrict__ kx0,
                                                                    const double* __restrict__ ky,
                                                                    const cplx<R>* __restrict__ tw,
                                                                    const KScal* __restrict__ ks, long long vs) {
  // block z: system of a batched solve (vectors vs apart, one KScal each)
  const BlockIdx3 bk = xcd_block();
  const int z = bk.z;
  ks += z;
  if (ks->done) return;
  X += z * vs;
  Bx += z * vs;
  By += z * vs;
  if (UPD != 0) {
    V += z * vs;
    Xo += z * vs;
  }
  // UPD 1: -alpha; UPD 2: -omega and beta (rounded to R as in the separate update kernels)
  const cplx<R> c1 = UPD == 1 ? cplx<R>{(R)-ks->ar, (R)-ks->ai} : cplx<R>{(R)-ks->wr, (R)-ks->wi};
  const cplx<R> c2 = cplx<R>{(R)ks->br, (R)ks->bi};
  const cplx<R> one = cplx<R>{R(1), R(0)};
  using F = KColFFT<R, L, true>;
  constexpr int C = KColCfg<R, L, CW>::C;
  constexpr int E = F::E;
  __shared__ cplx<R> lds[C * L];
  const int c = threadIdx.x % C, t = threadIdx.x / C;
  const int col = bk.x * C + c;
  const bool act = col < Nx;
  const long long base = (long long)bk.y * Nx * L + col;
  const R kxv = act ? (R)kx0[col] : R(0);
  const R s4 = (R)S4;
  // b1 = S4 x is kept; the two derivative transforms run one after the other
  cplx<R> b1[E], v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int pos = F::in_pos(t, e);
    const long long idx = base + (long long)pos * Nx;
    cplx<R> a = cplx<R>{R(0), R(0)};
    if (act) {
      if (UPD == 0) {
        a = X[idx];
      } else if (UPD == 1) {
        a = X[idx] + c1 * V[idx];
        Xo[idx] = a;
      } else {
        const cplx<R> pa = Xo[idx] + c1 * V[idx];
        a = one * X[idx] + c2 * pa;
        Xo[idx] = a;
      }
    }
    b1[e] = cplx<R>{s4 * a.x, s4 * a.y};
    v[e] = cplx<R>{-kxv * b1[e].y, kxv * b1[e].x};
  }
  F::run(v, lds + c, C, t, tw);
  if (act) {
#pragma unroll
    for (int e = 0; e < E; ++e) Bx[base + (long long)F::out_pos(t, e) * Nx] = v[e];
  }
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const R kyv = (R)ky[F::in_pos(t, e)];
    v[e] = cplx<R>{-kyv * b1[e].y, kyv * b1[e].x};
  }
  F::run(v, lds + c, C, t, tw);
  if (act) {
#pragma unroll
    for (int e = 0; e < E; ++e) By[base + (long long)F::out_pos(t, e) * Nx] = v[e];
  }
}

// ---------------------------------------------------------------------------
// K2: rows: inverse FFT (scaled 1/(Nx Ny)) -> * M_psi -> forward FFT, in place
// ---------------------------------------------------------------------------
// A workgroup is two wavefronts: the first transforms rows q of Bx, the second the same
// rows of By, so the two reads of each M_psi row come from one CU at the same time (with
// the two rows in different workgroups the second read of an FP64 M row mostly missed the
// L2: 1.11x the algorithmic bytes).  Each wavefront keeps its own LDS rows and
// synchronises at wavefront level.
template <typename R, int L>
__global__ __launch_bounds__((KRowCfg<R, L>::NT)) void k_mv_rows(cplx<R>* __restrict__ Bx, cplx<R>* __restrict__ By,
                                                                const R* __restrict__ M, long long nrows_half,
                                                                double scale, const cplx<R>* __restrict__ twi,
                                                                const cplx<R>* __restrict__ twf,
                                                                const KScal* __restrict__ ks, long long vs,
                                                                long long ms) {
  const int z = blockIdx.y;
  ks += z;
  if (ks->done) return;
  Bx += z * vs;
  By += z * vs;
  M += (z / 3) * ms;  // the mobility of the member the system belongs to
  using FI = typename KRowCfg<R, L>::FI;
  using FF = typename KRowCfg<R, L>::FF;
  constexpr int TT = FI::TT, E = FI::E, ROWS = KRowCfg<R, L>::ROWS, RS = KRowCfg<R, L>::RS;
  __shared__ cplx<R> lds[2 * ROWS * RS];
  const int wv = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int rloc = lane / TT, t = lane % TT;
  const long long row = (long long)blockIdx.x * ROWS + rloc;
  const bool act = row < nrows_half;
  const long long rr = act ? row : 0;
  cplx<R>* data = (wv == 0 ? Bx : By) + rr * L;
  const R* m = M + rr * L;
  cplx<R>* my = lds + (wv * ROWS + rloc) * RS;
  // FP64: M_psi is loaded with the row rather than after the first transform (measured
  // faster in FP64, slower in FP32)
  constexpr bool MPRE = sizeof(R) == 8;
  cplx<R> v[E];
  R mv[MPRE ? E : 1];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    v[e] = act ? data[FI::in_pos(t, e)] : cplx<R>{R(0), R(0)};
    if constexpr (MPRE) mv[e] = act ? m[FI::out_pos(t, e)] : R(0);
  }
  FI::run(v, my, 1, t, twi);
  const R sc = (R)scale;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    R mval;
    if constexpr (MPRE)
      mval = mv[e];
    else
      mval = act ? m[FI::out_pos(t, e)] : R(0);
    const cplx<R> q = cplx<R>{sc * v[e].x, sc * v[e].y};
    v[e] = cplx<R>{mval * q.x, mval * q.y};
  }
  // (the forward transform's first LDS write is preceded by its own barrier)
  FF::run(v, my, 1, t, twf);
  if (act) {
#pragma unroll
    for (int e = 0; e < E; ++e) data[FF::out_pos(t, e)] = v[e];
  }
}

// ---------------------------------------------------------------------------
// K3: column forward FFTs of Bx, By, combine, precondition, partial dots.
//   MODE 1: out = P^{-1} A y and partial <c, out>
//   MODE 2: out = P^{-1} A y and partials <out, y>, <out, out>
//   MODE 0: out only
// ---------------------------------------------------------------------------
template <typename R, int L, int MODE, int CW>
__global__ __launch_bounds__((KColCfg<R, L, CW, true>::NT)) void k_mv_cols_fwd(
    const cplx<R>* __restrict__ Bx, const cplx<R>* __restrict__ By, const cplx<R>* __restrict__ Y,
    const cplx<R>* __restrict__ Cv, cplx<R>* __restrict__ out, int Nx, double S4, const double* __restrict__ kx0,
    const double* __restrict__ ky, const cplx<R>* __restrict__ tw, const DScal* __restrict__ ds,
    const KScal* __restrict__ ks, double* __restrict__ partials, long long vs, long long ps) {
  const BlockIdx3 bk = xcd_block();
  const int z = bk.z;
  ks += z;
  if (ks->done) return;
  Bx += z * vs;
  By += z * vs;
  Y += z * vs;
  if (Cv) Cv += z * vs;
  out += z * vs;
  partials += z * ps;
  ds += z / 3;  // the step scalars of the member the system belongs to
  using F = KColFFT<R, L, false>;
  constexpr int C = KColCfg<R, L, CW, true>::C;
  constexpr int E = F::E;
  __shared__ cplx<R> lds[C * L];
  const int c = threadIdx.x % C, t = threadIdx.x / C;
  const int col = bk.x * C + c;
  const bool act = col < Nx;
  const long long base = (long long)bk.y * Nx * L + col;
  const double a_dt = ds->a / ds->dt;
  const double Mmax = ds->Mmax;
  const double kxd = act ? kx0[col] : 0.0;
  const R kxv = (R)kxd;
  const R adt = (R)a_dt;
  // Both fields stay in registers through their transforms (in FP64 248 VGPRs, which the
  // 128 KB of LDS per workgroup allows anyway: one workgroup per CU).  Parking the
  // x-derivative term in `out` and reading it back halves the registers but adds two
  // vector passes; FP64 ran 129.7 -> 135.8 SAV steps/s at 1024^2 without it.
  cplx<R> vx[E], vy[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const long long off = base + (long long)F::in_pos(t, e) * Nx;
    vx[e] = act ? Bx[off] : cplx<R>{R(0), R(0)};
    vy[e] = act ? By[off] : cplx<R>{R(0), R(0)};
  }
  F::run(vx, lds + c, C, t, tw);
  F::run(vy, lds + c, C, t, tw);
  double acc[3] = {0.0, 0.0, 0.0};
  if (act) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int pos = F::out_pos(t, e);
      const long long off = base + (long long)pos * Nx;
      const double kyd = ky[pos];
      const R kyv = (R)kyd;
      const cplx<R> y = Y[off];
      const cplx<R> dx = cplx<R>{-kxv * vx[e].y, kxv * vx[e].x};
      const cplx<R> dy = cplx<R>{-kyv * vy[e].y, kyv * vy[e].x};
      const cplx<R> sum = dx + dy;
      const cplx<R> q = cplx<R>{adt * y.x - sum.x, adt * y.y - sum.y};
      cplx<R> o = q;
      if constexpr ((MODE & 4) == 0) {
        const R pv = (R)pinv_diag(a_dt, Mmax, S4, kxd, kyd);
        o = cplx<R>{q.x * pv, q.y * pv};
      }
      out[off] = o;
      if constexpr ((MODE & 3) == 1) {
        const cplx<R> cc = Cv[off];
        acc[0] += (double)cc.x * (double)o.x + (double)cc.y * (double)o.y;
        acc[1] += (double)cc.x * (double)o.y - (double)cc.y * (double)o.x;
      } else if constexpr ((MODE & 3) == 2) {
        const cplx<R> dd = Cv[off];
        acc[0] += (double)o.x * (double)dd.x + (double)o.y * (double)dd.y;
        acc[1] += (double)o.x * (double)dd.y - (double)o.y * (double)dd.x;
        acc[2] += (double)o.x * (double)o.x + (double)o.y * (double)o.y;
      }
    }
  }
  if constexpr ((MODE & 3) != 0) {
    // one partial per column, summed over the column's team in a fixed order: the dot
    // products do not depend on how many columns a workgroup holds
    constexpr int TT = F::TT;
    static_assert(3 * C * TT * (int)sizeof(double) <= C * L * (int)sizeof(cplx<R>), "LDS reuse");
    __syncthreads();  // the transforms are done with the LDS
    double* red = reinterpret_cast<double*>(lds);
    red[3 * threadIdx.x + 0] = acc[0];
    red[3 * threadIdx.x + 1] = acc[1];
    red[3 * threadIdx.x + 2] = acc[2];
    __syncthreads();
    if (threadIdx.x < C) {
      const int cc = threadIdx.x;
      const int colc = bk.x * C + cc;
      double s0 = 0.0, s1 = 0.0, s2 = 0.0;
      for (int tt = 0; tt < TT; ++tt) {
        const int th = tt * C + cc;
        s0 += red[3 * th + 0];
        s1 += red[3 * th + 1];
        s2 += red[3 * th + 2];
      }
      if (colc < Nx) {
        const long long b = (long long)bk.y * Nx + colc;
        partials[b * 3 + 0] = s0;
        partials[b * 3 + 1] = s1;
        partials[b * 3 + 2] = s2;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Vector kernels
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void k_kry_init(const cplx<R>* __restrict__ b, cplx<R>* __restrict__ r,
                                                  cplx<R>* __restrict__ p, cplx<R>* __restrict__ x, long long n,
                                                  int Nx, int Ny, double S4, const double* __restrict__ kx0,
                                                  const double* __restrict__ ky, const DScal* __restrict__ ds,
                                                  double* __restrict__ partials, long long vs, long long ps) {
  __shared__ double scratch[3 * 4];
  // blockIdx.y: system of a batched solve
  const int z = blockIdx.y;
  b += z * vs;
  r += z * vs;
  p += z * vs;
  x += z * vs;
  partials += z * ps;
  ds += z / 3;  // the step scalars of the member the system belongs to
  double acc[3] = {0.0, 0.0, 0.0};
  const double a_dt = ds->a / ds->dt, Mmax = ds->Mmax;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const int ix = (int)(i % Nx);
    const int iy = (int)((i / Nx) % Ny);
    const R pv = (R)pinv_diag(a_dt, Mmax, S4, kx0[ix], ky[iy]);
    const cplx<R> bb = b[i];
    const cplx<R> rr = cplx<R>{bb.x * pv, bb.y * pv};
    r[i] = rr;
    p[i] = rr;
    x[i] = cplx<R>{R(0), R(0)};
    acc[0] += (double)bb.x * (double)rr.x + (double)bb.y * (double)rr.y;
    acc[1] += (double)bb.x * (double)rr.y - (double)bb.y * (double)rr.x;
    acc[2] += (double)rr.x * (double)rr.x + (double)rr.y * (double)rr.y;
  }
  block_sum<3>(acc, scratch);
  if (threadIdx.x == 0) {
    partials[blockIdx.x * 3 + 0] = acc[0];
    partials[blockIdx.x * 3 + 1] = acc[1];
    partials[blockIdx.x * 3 + 2] = acc[2];
  }
}

// Finalise a 3-wide partial array with one block (deterministic order).
__global__ __launch_bounds__(256) void k_kry_fin(const double* __restrict__ partials, int nb, KScal* __restrict__ ks,
                                                 int stage, double atol, double rtol, int itmax, long long ps,
                                                 const KScal* __restrict__ outer) {
  __shared__ double scratch[3 * 4];
  // one block per system of a batched solve
  ks += blockIdx.x;
  partials += blockIdx.x * ps;
  if (stage != 0 && ks->done) return;
  if (stage == 0 && outer != nullptr && outer[blockIdx.x].done) {
    // inner solve of a mixed-precision system whose outer solve has converged: nothing to do
    if (threadIdx.x == 0) {
      KScal k{};
      k.done = 1;
      k.solved = 1;
      k.status = 4;
      *ks = k;
    }
    return;
  }
  double acc[3] = {0.0, 0.0, 0.0};
  strided_sums<3, 256>(partials, nb, threadIdx.x, acc);
  block_sum<3>(acc, scratch);
  if (threadIdx.x != 0) return;
  KScal k = *ks;
  if (stage == 0) {  // after init: <c, r0>, ||r0||^2
    k.iter = 0;
    k.done = 0;
    k.solved = 0;
    k.status = 0;
    k.rnorm = sqrt(acc[2]);
    k.ar = k.ai = 0;
    k.wr = k.wi = 0;
    if (k.rnorm == 0.0) {
      k.done = 1;
      k.solved = 1;
      k.status = 4;
    } else {
      k.eps = atol + rtol * k.rnorm;
      // inner solve of a mixed-precision system: no need to go below half the outer
      // threshold (the round that reaches it is the last one)
      if (outer != nullptr) k.eps = fmax(k.eps, 0.5 * outer[blockIdx.x].eps);
      k.nr = acc[0];
      k.ni = acc[1];
      if (k.nr == 0.0 && k.ni == 0.0) {
        k.done = 1;
        k.status = 5;
      } else if (k.rnorm <= k.eps) {
        k.done = 1;
        k.solved = 1;
        k.status = 1;
      } else if (itmax <= 0) {
        k.done = 1;
        k.status = 2;
      }
    }
  } else if (stage == 1) {  // alpha = rho / <c, v>
    k.rr = k.nr;
    k.ri = k.ni;
    const zd al = zdiv(zd{k.rr, k.ri}, zd{acc[0], acc[1]});
    k.ar = al.x;
    k.ai = al.y;
  } else if (stage == 2) {  // omega = <t, s> / <t, t>
    const zd om = zdiv(zd{acc[0], acc[1]}, zd{acc[2], 0.0});
    k.wr = om.x;
    k.wi = om.y;
  } else {  // stage 3: next rho, ||r||, beta, stopping tests
    k.nr = acc[0];
    k.ni = acc[1];
    k.rnorm = sqrt(acc[2]);
    const zd q1 = zdiv(zd{k.nr, k.ni}, zd{k.rr, k.ri});
    const zd q2 = zdiv(zd{k.ar, k.ai}, zd{k.wr, k.wi});
    const zd be = zmul(q1, q2);
    k.br = be.x;
    k.bi = be.y;
    k.iter += 1;
    const bool solved = (k.rnorm <= k.eps) || (k.rnorm + 1.0 <= 1.0);
    const bool tired = k.iter >= itmax;
    const bool breakdown = (k.ar == 0.0 && k.ai == 0.0) || (k.ar != k.ar) || (k.ai != k.ai);
    if (solved || tired || breakdown) {
      k.done = 1;
      k.solved = solved ? 1 : 0;
      k.status = solved ? 1 : (breakdown ? 3 : 2);
    }
  }
  *ks = k;
}

template <typename R>
__global__ __launch_bounds__(256) void k_kry_upd1(const cplx<R>* __restrict__ r, const cplx<R>* __restrict__ v,
                                                  const cplx<R>* __restrict__ p, cplx<R>* __restrict__ s,
                                                  cplx<R>* __restrict__ x, long long n, const KScal* __restrict__ ks,
                                                  long long vs) {
  const int z = blockIdx.y;
  ks += z;
  if (ks->done) return;
  r += z * vs;
  v += z * vs;
  p += z * vs;
  s += z * vs;
  x += z * vs;
  const R ar = (R)ks->ar, ai = (R)ks->ai;
  const cplx<R> ma = cplx<R>{-ar, -ai};
  const cplx<R> al = cplx<R>{ar, ai};
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    s[i] = r[i] + ma * v[i];
    x[i] = x[i] + al * p[i];
  }
}

// s = r - alpha v, in its own pass (FP32; FP64 forms s inside the matvec, see kry_iteration).
// The iterate's update x += alpha p waits for k_kry_upd2, which reads x once per iteration;
// `done` changes only after upd2, so both run or neither.
template <typename R>
__global__ __launch_bounds__(256) void k_kry_upd1s(const cplx<R>* __restrict__ r, const cplx<R>* __restrict__ v,
                                                   cplx<R>* __restrict__ s, long long n,
                                                   const KScal* __restrict__ ks, long long vs) {
  const int z = blockIdx.y;
  ks += z;
  if (ks->done) return;
  r += z * vs;
  v += z * vs;
  s += z * vs;
  const cplx<R> ma = cplx<R>{(R)-ks->ar, (R)-ks->ai};
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    s[i] = r[i] + ma * v[i];
}

// x = (x + alpha p) + omega s (the two updates of Krylov.jl's iteration, rounded as two
// steps); r = s - omega t; partial <c, r>, ||r||^2.
template <typename R>
__global__ __launch_bounds__(256) void k_kry_upd2(const cplx<R>* __restrict__ s, const cplx<R>* __restrict__ t,
                                                  const cplx<R>* __restrict__ c, const cplx<R>* __restrict__ p,
                                                  cplx<R>* __restrict__ x, cplx<R>* __restrict__ r, long long n,
                                                  const KScal* __restrict__ ks, double* __restrict__ partials,
                                                  long long vs, long long ps) {
  __shared__ double scratch[3 * 4];
  const int z = blockIdx.y;
  ks += z;
  if (ks->done) return;
  s += z * vs;
  t += z * vs;
  c += z * vs;
  p += z * vs;
  x += z * vs;
  r += z * vs;
  partials += z * ps;
  const R wr = (R)ks->wr, wi = (R)ks->wi;
  const cplx<R> om = cplx<R>{wr, wi};
  const cplx<R> mom = cplx<R>{-wr, -wi};
  const cplx<R> al = cplx<R>{(R)ks->ar, (R)ks->ai};
  double acc[3] = {0.0, 0.0, 0.0};
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const cplx<R> si = s[i];
    const cplx<R> x1 = x[i] + al * p[i];
    x[i] = x1 + om * si;
    const cplx<R> ri = si + mom * t[i];
    r[i] = ri;
    const cplx<R> ci = c[i];
    acc[0] += (double)ci.x * (double)ri.x + (double)ci.y * (double)ri.y;
    acc[1] += (double)ci.x * (double)ri.y - (double)ci.y * (double)ri.x;
    acc[2] += (double)ri.x * (double)ri.x + (double)ri.y * (double)ri.y;
  }
  block_sum<3>(acc, scratch);
  if (threadIdx.x == 0) {
    partials[blockIdx.x * 3 + 0] = acc[0];
    partials[blockIdx.x * 3 + 1] = acc[1];
    partials[blockIdx.x * 3 + 2] = acc[2];
  }
}

template <typename R>
__global__ __launch_bounds__(256) void k_kry_upd3(const cplx<R>* __restrict__ r, const cplx<R>* __restrict__ v,
                                                  cplx<R>* __restrict__ p, long long n, const KScal* __restrict__ ks,
                                                  long long vs) {
  const int z = blockIdx.y;
  ks += z;
  if (ks->done) return;
  r += z * vs;
  v += z * vs;
  p += z * vs;
  const cplx<R> mom = cplx<R>{(R)-ks->wr, (R)-ks->wi};
  const cplx<R> be = cplx<R>{(R)ks->br, (R)ks->bi};
  const cplx<R> one = cplx<R>{R(1), R(0)};
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const cplx<R> pa = p[i] + mom * v[i];
    p[i] = one * r[i] + be * pa;
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
// Workgroups of a Krylov vector kernel over n elements per system: at most one per 1024
// elements, hence at most N Nx of them (Ny <= 1024), which is what the per-system
// partial slots (3 max(N Nx, 512) doubles) hold.
static int vec_blocks(long long n) {
  long long nb = (n + 1023) / 1024;
  if (nb < 1) nb = 1;
  if (nb > KRY_VEC_BLOCKS) nb = KRY_VEC_BLOCKS;
  return (int)nb;
}

#define SAV_KRY_SIZES(X)                                                                                         \
  X(4) X(6) X(8) X(10) X(12) X(16) X(20) X(24) X(32) X(40) X(48) X(64) X(80) X(96) X(128) X(160) X(192) X(256) X(320) \
  X(384) X(512) X(640) X(768) X(1024) X(1280) X(2048)

// Grid lengths the psi solve's matvec kernels are built for (Nx and Ny of a 2D grid).
bool kry_size_supported(int n) {
  switch (n) {
#define X(L) case L:
    SAV_KRY_SIZES(X)
#undef X
    return true;
    default: return false;
  }
}

// Start/stop events of the next launch of kernel `id` when it is the timed one and the
// launch is sampled (null otherwise). They are passed to hipExtLaunchKernel, which stamps
// them from the dispatch itself, so their interval is the kernel's own duration. The
// events delay the following dispatch by a few microseconds, so only every ev_stride-th
// launch carries them (an odd stride samples both matvecs of an iteration alike).
template <typename R>
static void tev_pre(KrylovCtx<R>& k, int id, hipEvent_t& e0, hipEvent_t& e1) {
  e0 = e1 = nullptr;
  if (k.timed == id && k.ev_used < k.ev_cap && (k.ev_seq++ % k.ev_stride) == 0) {
    e0 = k.ev[2 * k.ev_used];
    e1 = k.ev[2 * k.ev_used + 1];
  }
}
template <typename R>
static void tev_post(KrylovCtx<R>& k, hipEvent_t e0, double bytes, int sys = -1) {
  if (e0) {
    k.ev_iter[k.ev_used] = k.cur_iter;
    k.ev_bytes[k.ev_used] = bytes;
    k.ev_sys[k.ev_used] = sys;
    k.ev_used++;
  }
}

// Accumulates the timed launches that did work: a launch queued for iteration i
// counts the bytes of the systems whose solve ran at least i iterations.
template <typename R>
static void tev_collect(KrylovCtx<R>& k, const int* iters, int nsys) {
  for (int i = 0; i < k.ev_used; ++i) {
    int active = 0;
    if (k.ev_sys[i] >= 0)  // a launch of one system (k.seq)
      active = k.ev_sys[i] < nsys && k.ev_iter[i] <= iters[k.ev_sys[i]] ? 1 : 0;
    else
      for (int z = 0; z < nsys; ++z) active += k.ev_iter[i] <= iters[z] ? 1 : 0;
    if (active == 0) continue;
    float ms = 0.f;
    SAV_HIP(hipEventElapsedTime(&ms, k.ev[2 * i], k.ev[2 * i + 1]));
    k.tk_ms += ms;
    k.tk_bytes += k.ev_bytes[i] * active;
    k.tk_count += 1;
  }
  k.ev_used = 0;
}

// Operands of one matvec launch sequence: the three kernels of a matvec over `nsys` systems
// whose vectors (and Krylov scalars, partials) start at these pointers.
template <typename R>
struct MvArgs {
  const cplx<R>* xin;      // column-inverse input: y, or r when an update is formed on load
  const cplx<R>* V;        // v (updates only)
  cplx<R>* xo;             // s or p (updates only)
  cplx<R>* Bx;             // the two intermediate fields
  cplx<R>* By;
  const cplx<R>* y;        // the vector A is applied to (a/dt y term of the forward kernel)
  const cplx<R>* partner;  // second operand of the fused dot products
  cplx<R>* out;
  const KScal* ks;
  double* partials;
  const R* M;              // mobility of the batch's first member (system z's: M + (z / 3) nfull)
  const DScal* ds;         // its step scalars (system z's: ds + z / 3)
  int nsys;
};

// upd: the vector update the kernel forms on load (k_mv_cols_inv's UPD, FP64 only); 1 and 2
// take r and v and write s (1) or p (2).
template <typename R, int CW>
static void launch_cols_inv(KrylovCtx<R>& k, const MvArgs<R>& a, int upd, hipEvent_t e0, hipEvent_t e1) {
  if (upd != 0 && sizeof(R) != 8) throw HipError("launch_cols_inv: the fused vector update is built for FP64 only", -1);
  const long long vs = k.vs;
#define KL(L, U)                                                                                                   \
  hipExtLaunchKernelGGL((k_mv_cols_inv<R, L, CC::C, U>), grid, dim3(CC::NT), 0, k.st, e0, e1, 0, a.xin, a.V, a.xo, a.Bx, \
                        a.By, k.Nx, k.S4, k.kx0, k.ky, k.tw_ci, a.ks, vs)
  switch (k.Ny) {
#define X(L)                                                  \
  case L: {                                                   \
    using CC = KColCfg<R, L, CW>;                             \
    dim3 grid((k.Nx + CC::C - 1) / CC::C, k.N, a.nsys);       \
    if constexpr (sizeof(R) == 8) {                           \
      if (upd == 1)                                           \
        KL(L, 1);                                             \
      else if (upd == 2)                                      \
        KL(L, 2);                                             \
      else                                                    \
        KL(L, 0);                                             \
    } else {                                                  \
      KL(L, 0);                                               \
    }                                                         \
    break;                                                    \
  }
    SAV_KRY_SIZES(X)
#undef X
#undef KL
    default: throw HipError("unsupported Ny for the psi solve: " + std::to_string(k.Ny), -1);
  }
  SAV_LAUNCH_CHECK();
}

template <typename R, int CW>
static int launch_cols_fwd(KrylovCtx<R>& k, const MvArgs<R>& a, int mode, hipEvent_t e0, hipEvent_t e1) {
  const long long vs = k.vs, ps = k.ps;
  int nb3 = 0;
#define KL(L, M)                                                                                                 \
  hipExtLaunchKernelGGL((k_mv_cols_fwd<R, L, M, CC::C>), grid, dim3(CC::NT), 0, k.st, e0, e1, 0, a.Bx, a.By, a.y,   \
                        a.partner, a.out, k.Nx, k.S4, k.kx0, k.ky, k.tw_cf, a.ds, a.ks, a.partials, vs, ps)
  switch (k.Ny) {
#define X(L)                                            \
  case L: {                                             \
    using CC = KColCfg<R, L, CW, true>;                 \
    dim3 grid((k.Nx + CC::C - 1) / CC::C, k.N, a.nsys); \
    nb3 = k.Nx * k.N;  /* one partial per column */      \
    switch (mode) {                                     \
      case 1: KL(L, 1); break;                          \
      case 2: KL(L, 2); break;                          \
      case 4: KL(L, 4); break;                          \
      case 5: KL(L, 5); break;                          \
      case 6: KL(L, 6); break;                          \
      default: KL(L, 0); break;                         \
    }                                                   \
    break;                                              \
  }
    SAV_KRY_SIZES(X)
#undef X
    default: break;
  }
#undef KL
  SAV_LAUNCH_CHECK();
  return nb3;
}

template <typename R>
static void launch_rows(KrylovCtx<R>& k, const MvArgs<R>& a, hipEvent_t e0, hipEvent_t e1) {
  const long long nrh = (long long)k.N * k.Ny;
  const double scale = 1.0 / ((double)k.Nx * k.Ny);
  switch (k.Nx) {
#define X(L)                                                                                                  \
  case L: {                                                                                                   \
    using CC = KRowCfg<R, L>;                                                                                 \
    dim3 grid((unsigned)((nrh + CC::ROWS - 1) / CC::ROWS), a.nsys);                                           \
    hipExtLaunchKernelGGL((k_mv_rows<R, L>), grid, dim3(CC::NT), 0, k.st, e0, e1, 0, a.Bx, a.By, a.M, nrh,    \
                          scale, k.tw_ri, k.tw_rf, a.ks, k.vs, k.nfull);                                      \
    break;                                                                                                    \
  }
    SAV_KRY_SIZES(X)
#undef X
    default: throw HipError("unsupported Nx for the psi solve: " + std::to_string(k.Nx), -1);
  }
  SAV_LAUNCH_CHECK();
}

// The three kernels of a matvec on the operands `a` (sys: the system a single-system
// sequence belongs to, for the kernel timing; -1 when `a` covers all systems).
template <typename R>
static int matvec_kernels(KrylovCtx<R>& k, const MvArgs<R>& a, int mode, int upd, int sys) {
  const double sn = (double)sizeof(R) * (double)k.nfull;
  hipEvent_t e0, e1;
  tev_pre(k, 1, e0, e1);
  if (k.colw == 16)
    launch_cols_inv<R, 16>(k, a, upd, e0, e1);
  else if (k.colw == 4)
    launch_cols_inv<R, 4>(k, a, upd, e0, e1);
  else
    launch_cols_inv<R, 8>(k, a, upd, e0, e1);
  tev_post(k, e0, (upd == 0 ? 6.0 : upd == 1 ? 10.0 : 12.0) * sn, sys);
  tev_pre(k, 2, e0, e1);
  launch_rows(k, a, e0, e1);
  tev_post(k, e0, 9.0 * sn, sys);
  tev_pre(k, 3, e0, e1);
  const int nb3 = k.colw == 16  ? launch_cols_fwd<R, 16>(k, a, mode, e0, e1)
                  : k.colw == 4 ? launch_cols_fwd<R, 4>(k, a, mode, e0, e1)
                                : launch_cols_fwd<R, 8>(k, a, mode, e0, e1);
  tev_post(k, e0, ((mode & 3) == 1 ? 10.0 : (mode & 3) == 2 ? 10.0 : 6.0) * sn, sys);
  return nb3;
}

template <typename T>
static T* shifted(T* p, long long d) {
  return p ? p + d : p;
}

// One matvec on nsys systems (vectors k.vs apart, KScal slots slot .. slot+nsys-1).
// mode bits: &3 = 0 no dots, 1 <partner, out>, 2 <out, partner> and <out, out>; &4 = no preconditioner
// upd: 0 = apply to y; 1 = y is s = r - alpha v, 2 = y is p = r + beta (p - omega v), formed
// (and stored) by the first kernel (launch_cols_inv).
// With k.seq the systems run one after the other, each through all three kernels, over one
// shared pair of intermediate fields: the intermediate written by one kernel is then read
// by the next while it is still in the 256 MB infinity cache (at 1024^2, N = 2 in FP64 one
// system's pair is 64 MB, three systems' 192 MB), and the pair is overwritten in that
// cache rather than written back.  Each system's arithmetic is unchanged.
template <typename R>
static int matvec_launch(KrylovCtx<R>& k, const cplx<R>* y, cplx<R>* out, int mode, int slot,
                         const cplx<R>* partner = nullptr, int nsys = 1, int upd = 0) {
  MvArgs<R> a;
  a.xin = upd == 0 ? y : k.r;
  a.V = upd == 0 ? nullptr : k.v;
  a.xo = upd == 1 ? k.s : upd == 2 ? k.p : nullptr;
  a.Bx = k.Bx;
  a.By = k.By;
  a.y = y;
  a.partner = partner;
  a.out = out;
  a.ks = k.ks + slot;
  a.partials = k.partials;
  a.M = k.M;
  a.ds = k.ds;
  a.nsys = nsys;
  if (!k.seq || nsys == 1) return matvec_kernels(k, a, mode, upd, -1);
  int nb3 = 0;
  for (int z = 0; z < nsys; ++z) {
    const long long d = z * k.vs;
    MvArgs<R> b = a;
    b.xin = shifted(a.xin, d);
    b.V = shifted(a.V, d);
    b.xo = shifted(a.xo, d);
    b.y = shifted(a.y, d);
    b.partner = shifted(a.partner, d);
    b.out = shifted(a.out, d);
    b.ks = a.ks + z;
    b.partials = a.partials + z * k.ps;
    b.M = k.M + (z / 3) * k.nfull;
    b.ds = k.ds + z / 3;
    b.nsys = 1;
    nb3 = matvec_kernels(k, b, mode, upd, z);
  }
  return nb3;
}

// Columns per workgroup that width CW resolves to for column length Ny, inverse kernel
// x 1000 + forward kernel (the LDS and thread limits can cap 16 down to 8).
template <typename R, int CW>
static int effective_columns(int Ny) {
  switch (Ny) {
#define X(L) \
  case L: return 1000 * KColCfg<R, L, CW>::C + KColCfg<R, L, CW, true>::C;
    SAV_KRY_SIZES(X)
#undef X
    default: return 0;
  }
}

// Milliseconds per matvec of three pairs of matvecs in the shape of a step's batched solve
// (k.tune_nsys systems, KScal slots from 0, on the right-hand sides in k.b, output into k.t;
// in FP64 with the s and p updates formed on load, as an iteration runs them), after one
// warm-up.  The slots are reinitialised by the solve that follows.
template <typename R>
static double time_matvec_pairs(KrylovCtx<R>& k, hipEvent_t a, hipEvent_t b) {
  const int nsys = k.tune_nsys;
  constexpr int u1 = sizeof(R) == 8 ? 1 : 0, u2 = sizeof(R) == 8 ? 2 : 0;
  const std::vector<KScal> zs(nsys, KScal{});
  SAV_HIP(hipMemcpyAsync(k.ks, zs.data(), nsys * sizeof(KScal), hipMemcpyHostToDevice, k.st));
  matvec_launch(k, k.b, k.t, 1, 0, k.b, nsys, u1);  // warm-up
  SAV_HIP(hipEventRecord(a, k.st));
  for (int r = 0; r < 3; ++r) {
    matvec_launch(k, k.b, k.t, 1, 0, k.b, nsys, u1);
    matvec_launch(k, k.b, k.t, 2, 0, k.b, nsys, u2);
  }
  SAV_HIP(hipEventRecord(b, k.st));
  SAV_HIP(hipEventSynchronize(b));
  float ms = 0.f;
  SAV_HIP(hipEventElapsedTime(&ms, a, b));
  return ms / 6.0;
}

// Autotuner of the matvec at the first solve of a handle.  Column width: 8, 16 and 4 columns
// per workgroup (widths that resolve to the kernel shape of a width already timed are
// skipped), all systems per launch; the fastest is kept.  One system alone favours narrow
// tiles (1024^2 FP64: 4 columns 0.085 against 0.089 ms per matvec) that are slower over the
// three systems of a step, hence the three-system timing.  Order: with that width, all
// systems per launch against system by system over one shared intermediate pair (k.seq, see
// matvec_launch); the faster is kept.  Every choice computes the same numbers: the same
// transform per column, the Krylov dot products summed per column in a fixed order.
template <typename R>
static void autotune_columns(KrylovCtx<R>& k) {
  if (k.colw != 0 && k.order != 0) {
    k.seq = k.order == 2;
    return;
  }
  hipEvent_t a, b;
  SAV_HIP(hipEventCreate(&a));
  SAV_HIP(hipEventCreate(&b));
  const int saved_timed = k.timed;
  k.timed = 0;
  if (k.colw == 0) {
    const int e8 = effective_columns<R, 8>(k.Ny), e16 = effective_columns<R, 16>(k.Ny),
              e4 = effective_columns<R, 4>(k.Ny);
    const int widths[3] = {8, 16, 4};
    const bool timed_w[3] = {true, e16 != e8, e4 != e8 && e4 != e16};
    k.seq = 0;
    if (e8 == e16 && e8 == e4) {
      k.colw = 8;  // one kernel shape whatever the width: nothing to choose
    } else {
      for (int w = 0; w < 3; ++w) {
        if (!timed_w[w]) continue;
        k.colw = widths[w];
        k.tune_ms[w] = time_matvec_pairs(k, a, b);
      }
      int best = 0;
      for (int w = 1; w < 3; ++w)
        if (timed_w[w] && k.tune_ms[w] < k.tune_ms[best]) best = w;
      k.colw = widths[best];
    }
  }
  if (k.order == 0) {
    for (int o = 0; o < 2; ++o) {
      k.seq = o;
      k.tune_order_ms[o] = time_matvec_pairs(k, a, b);
    }
    k.order = k.tune_order_ms[1] < k.tune_order_ms[0] ? 2 : 1;
  }
  k.seq = k.order == 2;
  k.timed = saved_timed;
  SAV_HIP(hipEventDestroy(a));
  SAV_HIP(hipEventDestroy(b));
}

template <typename R>
void matvec_prec_2d(KrylovCtx<R>& k, const cplx<R>* y, cplx<R>* out, int slot) {
  // make sure the slot is "running" so the kernels do not skip
  KScal z{};
  SAV_HIP(hipMemcpyAsync(k.ks + slot, &z, sizeof(KScal), hipMemcpyHostToDevice, k.st));
  matvec_launch(k, y, out, 0, slot);
  SAV_HIP(hipStreamSynchronize(k.st));
}

template <typename R>
void matvec_plain_2d(KrylovCtx<R>& k, const cplx<R>* y, cplx<R>* out, int slot) {
  KScal z{};
  SAV_HIP(hipMemcpyAsync(k.ks + slot, &z, sizeof(KScal), hipMemcpyHostToDevice, k.st));
  matvec_launch(k, y, out, 4, slot);
}

template <typename R>
static void fin_launch(KrylovCtx<R>& k, int nb, int slot, int stage, int nsys) {
  hipLaunchKernelGGL(k_kry_fin, dim3(nsys), dim3(256), 0, k.st, k.partials, nb, k.ks + slot, stage, k.atol, k.rtol,
                     k.itmax, k.ps, k.outer ? k.outer + slot : nullptr);
}

// x = 0, r = p = P^{-1} b, rho, the stopping threshold: KScal slots slot .. slot + nsys - 1
// start "running".
template <typename R>
static void kry_start(KrylovCtx<R>& k, int slot, int nsys) {
  const long long n = k.nfull;
  const int nbv = vec_blocks(n);
  hipLaunchKernelGGL((k_kry_init<R>), dim3(nbv, nsys), dim3(256), 0, k.st, k.b, k.r, k.p, k.x, n, k.Nx, k.Ny, k.S4,
                     k.kx0, k.ky, k.ds, k.partials, k.vs, k.ps);
  SAV_LAUNCH_CHECK();
  fin_launch(k, nbv, slot, 0, nsys);
  SAV_LAUNCH_CHECK();
}

// One BiCGSTAB iteration of every system still running (a system whose `done` flag is set
// skips every kernel), queued on k.st with no host synchronisation.
// In FP64 the updates of s and of p run inside the first kernel of the matvec that consumes
// them (launch_cols_inv's upd), so an iteration is two matvecs, one vector kernel and three
// scalar finishes; `first` marks the iteration after kry_start, whose p is r as initialised.
// In FP32 that kernel holds 32 elements per lane at two wavefronts per SIMD, too few to hide
// the update's extra loads, and the separate update passes are faster (1024^2, two vesicles,
// per launch over 2-3 systems: s 66 us fused against 32 + 24 us, p 88 against 32 + 27; in
// FP64 99 against 62 + 48 and 113 against 62 + 55).
template <typename R>
static void kry_iteration(KrylovCtx<R>& k, int slot, int nsys, bool first) {
  constexpr bool fuse = sizeof(R) == 8;
  const long long n = k.nfull;
  const int nbv = vec_blocks(n);
  const long long vs = k.vs, ps = k.ps;
  KScal* ks = k.ks + slot;
  // [p = r + beta (p - omega v)] ; v = P^{-1} A p ; alpha = rho / <c, v>
  int nb3 = matvec_launch(k, k.p, k.v, 1, slot, k.b, nsys, (fuse && !first) ? 2 : 0);
  fin_launch(k, nb3, slot, 1, nsys);
  // s = r - alpha v ; t = P^{-1} A s ; omega = <t, s> / <t, t>
  if (!fuse) hipLaunchKernelGGL((k_kry_upd1s<R>), dim3(nbv, nsys), dim3(256), 0, k.st, k.r, k.v, k.s, n, ks, vs);
  nb3 = matvec_launch(k, k.s, k.t, 2, slot, k.s, nsys, fuse ? 1 : 0);
  fin_launch(k, nb3, slot, 2, nsys);
  // x += alpha p, x += omega s ; r = s - omega t ; <c, r>, ||r||
  hipLaunchKernelGGL((k_kry_upd2<R>), dim3(nbv, nsys), dim3(256), 0, k.st, k.s, k.t, k.b, k.p, k.x, k.r, n, ks,
                     k.partials, vs, ps);
  fin_launch(k, nbv, slot, 3, nsys);
  // p = r + beta (p - omega v), for the next iteration's matvec
  if (!fuse) hipLaunchKernelGGL((k_kry_upd3<R>), dim3(nbv, nsys), dim3(256), 0, k.st, k.r, k.v, k.p, n, ks, vs);
  SAV_LAUNCH_CHECK();
}

template <typename R>
void bicgstab_queue(KrylovCtx<R>& k, int slot, int nsys, int iters) {
  if (k.colw == 0 || k.order == 0) throw HipError("bicgstab_queue: the matvec's tile and order are not chosen yet", -1);
  kry_start(k, slot, nsys);
  for (int it = 0; it < iters && it < k.itmax; ++it) kry_iteration(k, slot, nsys, it == 0);
}

template <typename R>
int bicgstab_batch(KrylovCtx<R>& k, int slot, int nsys) {
  autotune_columns(k);
  const long long n = k.nfull;
  KScal* ks = k.ks + slot;
  kry_start(k, slot, nsys);
  int queued = 0;
  const bool logging = k.snap != nullptr && nsys == 1;
  int guess = 0;
  for (int z = 0; z < nsys; ++z) guess = std::max(guess, k.guess[slot + z]);
  int batch = logging ? 1 : (guess > 0 ? guess : 8);
  k.snap_n = 0;
  while (true) {
    const int nb = batch < k.itmax - queued ? batch : k.itmax - queued;
    for (int it = 0; it < nb; ++it) {
      k.cur_iter = queued + it + 1;
      kry_iteration(k, slot, nsys, queued + it == 0);
    }
    queued += nb;
    SAV_HIP(hipMemcpyAsync(k.hks + slot, ks, nsys * sizeof(KScal), hipMemcpyDeviceToHost, k.st));
    SAV_HIP(hipStreamSynchronize(k.st));
    if (logging && k.snap_n < k.snap_cap && k.hks[slot].iter > k.snap_n) {
      // iterate x_k after iteration k (only iterations that ran)
      SAV_HIP(hipMemcpyAsync(k.snap + (long long)k.snap_n * n, k.x, n * sizeof(cplx<R>), hipMemcpyDeviceToDevice,
                             k.st));
      k.snap_n++;
    }
    bool all_done = true;
    for (int z = 0; z < nsys; ++z) all_done = all_done && k.hks[slot + z].done;
    if (all_done || queued >= k.itmax) break;
    if (!logging) batch = 2;
  }
  std::vector<int> iters(nsys, 0);
  int maxit = 0;
  for (int z = 0; z < nsys; ++z) {
    iters[z] = k.hks[slot + z].iter;
    k.guess[slot + z] = iters[z] > 0 ? iters[z] : 1;
    maxit = std::max(maxit, iters[z]);
  }
  if (k.timed) tev_collect(k, iters.data(), nsys);
  return maxit;
}

template <typename R>
int bicgstab_2d(KrylovCtx<R>& k, int slot) {
  return bicgstab_batch(k, slot, 1);
}

// ---------------------------------------------------------------------------
// Mixed-precision solve: FP64 defect correction around an FP32 inner BiCGSTAB
// ---------------------------------------------------------------------------
// Round j: the inner solve reduces the preconditioned residual of A e = r_j by its own
// rtol in FP32, then x += e and r_{j+1} = b - A x in FP64.  Only e, whose size is that of
// the residual, is ever rounded to FP32, so the solution's accuracy is set by the FP64
// residual test, not by FP32 rounding of x.

// x = 0, r32 = b (FP32), partial ||P^{-1} b||^2.
__global__ __launch_bounds__(256) void k_ref_start(const cplx<double>* __restrict__ b, cplx<double>* __restrict__ x,
                                                   cplx<float>* __restrict__ r32, long long n, int Nx, int Ny,
                                                   double S4, const double* __restrict__ kx0,
                                                   const double* __restrict__ ky, const DScal* __restrict__ ds,
                                                   double* __restrict__ partials, long long vs, long long ps) {
  __shared__ double scratch[3 * 4];
  const int z = blockIdx.y;
  b += z * vs;
  x += z * vs;
  r32 += z * vs;
  partials += z * ps;
  const double a_dt = ds->a / ds->dt, Mmax = ds->Mmax;
  double acc[3] = {0.0, 0.0, 0.0};
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const int ix = (int)(i % Nx);
    const int iy = (int)((i / Nx) % Ny);
    const double pv = pinv_diag(a_dt, Mmax, S4, kx0[ix], ky[iy]);
    const cplx<double> bb = b[i];
    x[i] = cplx<double>{0.0, 0.0};
    r32[i] = cplx<float>{(float)bb.x, (float)bb.y};
    acc[0] += pv * pv * (bb.x * bb.x + bb.y * bb.y);
  }
  block_sum<3>(acc, scratch);
  if (threadIdx.x == 0) {
    partials[blockIdx.x * 3 + 0] = acc[0];
    partials[blockIdx.x * 3 + 1] = acc[1];
    partials[blockIdx.x * 3 + 2] = acc[2];
  }
}

// x += e (the inner solution) for the systems still iterating.
__global__ __launch_bounds__(256) void k_ref_update(cplx<double>* __restrict__ x, const cplx<float>* __restrict__ e,
                                                    long long n, const KScal* __restrict__ ks, long long vs) {
  const int z = blockIdx.y;
  if (ks[z].done) return;
  x += z * vs;
  e += z * vs;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const cplx<float> d = e[i];
    const cplx<double> xi = x[i];
    x[i] = cplx<double>{xi.x + (double)d.x, xi.y + (double)d.y};
  }
}

// r = b - Ax (ax = A x, unpreconditioned), r32 = r in FP32, partial ||P^{-1} r||^2.
__global__ __launch_bounds__(256) void k_ref_resid(const cplx<double>* __restrict__ b,
                                                   const cplx<double>* __restrict__ ax, cplx<float>* __restrict__ r32,
                                                   long long n, int Nx, int Ny, double S4,
                                                   const double* __restrict__ kx0, const double* __restrict__ ky,
                                                   const DScal* __restrict__ ds, const KScal* __restrict__ ks,
                                                   double* __restrict__ partials, long long vs, long long ps) {
  __shared__ double scratch[3 * 4];
  const int z = blockIdx.y;
  if (ks[z].done) return;
  b += z * vs;
  ax += z * vs;
  r32 += z * vs;
  partials += z * ps;
  const double a_dt = ds->a / ds->dt, Mmax = ds->Mmax;
  double acc[3] = {0.0, 0.0, 0.0};
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const int ix = (int)(i % Nx);
    const int iy = (int)((i / Nx) % Ny);
    const double pv = pinv_diag(a_dt, Mmax, S4, kx0[ix], ky[iy]);
    const cplx<double> bb = b[i], aa = ax[i];
    const double rx = bb.x - aa.x, ry = bb.y - aa.y;
    r32[i] = cplx<float>{(float)rx, (float)ry};
    acc[0] += pv * pv * (rx * rx + ry * ry);
    acc[1] += pv * pv * (bb.x * bb.x + bb.y * bb.y);
  }
  block_sum<3>(acc, scratch);
  if (threadIdx.x == 0) {
    partials[blockIdx.x * 3 + 0] = acc[0];
    partials[blockIdx.x * 3 + 1] = acc[1];
    partials[blockIdx.x * 3 + 2] = acc[2];
  }
}

// Outer scalars, one block per system.  stage 0: ||P^{-1} b||, the threshold
// atol + rtol ||P^{-1} b||; stage 1 (after a round): the true residual norm, the inner
// iterations of the round, the stopping tests; stage 2 (warm start, after k_ref_resid on
// x0): ||P^{-1}(b - A x0)|| and the threshold from ||P^{-1} b||; stage -1: reset.
__global__ __launch_bounds__(256) void k_ref_fin(const double* __restrict__ partials, int nb, KScal* __restrict__ ks,
                                                 const KScal* __restrict__ inner, int stage, double atol, double rtol,
                                                 int max_rounds, long long ps) {
  __shared__ double scratch[3 * 4];
  ks += blockIdx.x;
  partials += blockIdx.x * ps;
  if (stage < 0) {  // reset: every system iterates (before a warm-started residual)
    if (threadIdx.x == 0) *ks = KScal{};
    return;
  }
  if (stage == 1 && ks->done) return;
  double acc[3] = {0.0, 0.0, 0.0};
  strided_sums<3, 256>(partials, nb, threadIdx.x, acc);
  block_sum<3>(acc, scratch);
  if (threadIdx.x != 0) return;
  KScal k = *ks;
  const double rn = sqrt(acc[0]);
  if (stage == 2) {  // warm start: rn = ||P (b - A x0)||, the threshold from ||P b||
    k = KScal{};
    k.rnorm = rn;
    k.eps = atol + rtol * sqrt(acc[1]);
    if (rn <= k.eps) {
      k.done = 1;
      k.solved = 1;
      k.status = rn == 0.0 ? 4 : 1;
    }
  } else if (stage == 0) {
    k = KScal{};
    k.rnorm = rn;
    k.eps = atol + rtol * rn;
    if (rn == 0.0) {
      k.done = 1;
      k.solved = 1;
      k.status = 4;
    } else if (rn <= k.eps) {
      k.done = 1;
      k.solved = 1;
      k.status = 1;
    }
  } else {
    k.rounds += 1;
    k.iter += inner[blockIdx.x].iter;
    k.rnorm = rn;
    if (rn <= k.eps) {
      k.done = 1;
      k.solved = 1;
      k.status = 1;
    } else if (k.rounds >= max_rounds || rn != rn) {
      k.done = 1;
      k.status = 2;
    }
  }
  *ks = k;
}

int bicgstab_refined(KrylovCtx<double>& k, KrylovCtx<float>& k32, int slot, int nsys, int max_rounds, bool warm) {
  if (k32.vs != k.vs || k32.nfull != k.nfull || k32.st != k.st)
    throw HipError("bicgstab_refined: the FP32 context does not match the FP64 one", -1);
  autotune_columns(k);
  autotune_columns(k32);
  const long long n = k.nfull, vs = k.vs, ps = k.ps;
  const int nbv = vec_blocks(n);
  KScal* ks = k.ks + slot;
  k32.outer = k.ks;
  if (warm) {
    // x0 = the iterate k.x holds (the previous solve's solution): r = b - A x0 in FP64
    hipLaunchKernelGGL(k_ref_fin, dim3(nsys), dim3(256), 0, k.st, k.partials, 0, ks, k32.ks + slot, -1, k.atol,
                       k.rtol, max_rounds, ps);
    matvec_launch<double>(k, k.x, k.t, 4, slot, nullptr, nsys);
    hipLaunchKernelGGL(k_ref_resid, dim3(nbv, nsys), dim3(256), 0, k.st, k.b, k.t, k32.b, n, k.Nx, k.Ny, k.S4,
                       k.kx0, k.ky, k.ds, ks, k.partials, vs, ps);
    hipLaunchKernelGGL(k_ref_fin, dim3(nsys), dim3(256), 0, k.st, k.partials, nbv, ks, k32.ks + slot, 2, k.atol,
                       k.rtol, max_rounds, ps);
  } else {
    hipLaunchKernelGGL(k_ref_start, dim3(nbv, nsys), dim3(256), 0, k.st, k.b, k.x, k32.b, n, k.Nx, k.Ny, k.S4,
                       k.kx0, k.ky, k.ds, k.partials, vs, ps);
    hipLaunchKernelGGL(k_ref_fin, dim3(nsys), dim3(256), 0, k.st, k.partials, nbv, ks, k32.ks + slot, 0, k.atol,
                       k.rtol, max_rounds, ps);
  }
  SAV_LAUNCH_CHECK();
  for (int round = 0; round < max_rounds; ++round) {
    // inner FP32 solve of A e = r on the systems still iterating (the others start done)
    bicgstab_batch<float>(k32, slot, nsys);
    hipLaunchKernelGGL(k_ref_update, dim3(nbv, nsys), dim3(256), 0, k.st, k.x, k32.x, n, ks, vs);
    // r = b - A x in FP64: A x into k.t (unpreconditioned, no dots), skipping converged systems
    matvec_launch<double>(k, k.x, k.t, 4, slot, nullptr, nsys);
    hipLaunchKernelGGL(k_ref_resid, dim3(nbv, nsys), dim3(256), 0, k.st, k.b, k.t, k32.b, n, k.Nx, k.Ny, k.S4,
                       k.kx0, k.ky, k.ds, ks, k.partials, vs, ps);
    hipLaunchKernelGGL(k_ref_fin, dim3(nsys), dim3(256), 0, k.st, k.partials, nbv, ks, k32.ks + slot, 1, k.atol,
                       k.rtol, max_rounds, ps);
    SAV_LAUNCH_CHECK();
    SAV_HIP(hipMemcpyAsync(k.hks + slot, ks, nsys * sizeof(KScal), hipMemcpyDeviceToHost, k.st));
    SAV_HIP(hipStreamSynchronize(k.st));
    bool all_done = true;
    for (int z = 0; z < nsys; ++z) all_done = all_done && k.hks[slot + z].done;
    if (all_done) break;
  }
  k32.outer = nullptr;
  int maxit = 0;
  for (int z = 0; z < nsys; ++z) maxit = std::max(maxit, k.hks[slot + z].iter);
  return maxit;
}

// Flexible (right-preconditioned) BiCGSTAB with a possibly nonlinear preconditioner
// z = prec(r).  x0 = 0, shadow vector c = b.  Stops when the diagonally
// preconditioned TRUE residual ||P^{-1}(b - A x)|| <= atol + rtol ||P^{-1} b||,
// the quantity the reference's left-preconditioned solve tests.
template <typename R>
__global__ __launch_bounds__(256) void k_fkry_init(const cplx<R>* __restrict__ b, cplx<R>* __restrict__ r,
                                                   cplx<R>* __restrict__ p, cplx<R>* __restrict__ x, long long n,
                                                   int Nx, int Ny, double S4, const double* __restrict__ kx0,
                                                   const double* __restrict__ ky, const DScal* __restrict__ ds,
                                                   double* __restrict__ partials) {
  __shared__ double scratch[3 * 4];
  double acc[3] = {0.0, 0.0, 0.0};
  const double a_dt = ds->a / ds->dt, Mmax = ds->Mmax;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const int ix = (int)(i % Nx);
    const int iy = (int)((i / Nx) % Ny);
    const double pv = pinv_diag(a_dt, Mmax, S4, kx0[ix], ky[iy]);
    const cplx<R> bb = b[i];
    r[i] = bb;
    p[i] = bb;
    x[i] = cplx<R>{R(0), R(0)};
    acc[0] += (double)bb.x * (double)bb.x + (double)bb.y * (double)bb.y;
    acc[2] += pv * pv * ((double)bb.x * (double)bb.x + (double)bb.y * (double)bb.y);
  }
  block_sum<3>(acc, scratch);
  if (threadIdx.x == 0) {
    partials[blockIdx.x * 3 + 0] = acc[0];
    partials[blockIdx.x * 3 + 1] = acc[1];
    partials[blockIdx.x * 3 + 2] = acc[2];
  }
}

// x += omega s_hat ; r = s - omega t ; partials <c, r>, ||P^{-1} r||^2
template <typename R>
__global__ __launch_bounds__(256) void k_fkry_upd2(const cplx<R>* __restrict__ s, const cplx<R>* __restrict__ shat,
                                                   const cplx<R>* __restrict__ t, const cplx<R>* __restrict__ c,
                                                   cplx<R>* __restrict__ x, cplx<R>* __restrict__ r, long long n,
                                                   int Nx, int Ny, double S4, const double* __restrict__ kx0,
                                                   const double* __restrict__ ky, const DScal* __restrict__ ds,
                                                   const KScal* __restrict__ ks, double* __restrict__ partials) {
  __shared__ double scratch[3 * 4];
  if (ks->done) return;
  const cplx<R> om = cplx<R>{(R)ks->wr, (R)ks->wi};
  const cplx<R> mom = cplx<R>{(R)-ks->wr, (R)-ks->wi};
  const double a_dt = ds->a / ds->dt, Mmax = ds->Mmax;
  double acc[3] = {0.0, 0.0, 0.0};
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    x[i] = x[i] + om * shat[i];
    const cplx<R> ri = s[i] + mom * t[i];
    r[i] = ri;
    const cplx<R> ci = c[i];
    const int ix = (int)(i % Nx);
    const int iy = (int)((i / Nx) % Ny);
    const double pv = pinv_diag(a_dt, Mmax, S4, kx0[ix], ky[iy]);
    acc[0] += (double)ci.x * (double)ri.x + (double)ci.y * (double)ri.y;
    acc[1] += (double)ci.x * (double)ri.y - (double)ci.y * (double)ri.x;
    acc[2] += pv * pv * ((double)ri.x * (double)ri.x + (double)ri.y * (double)ri.y);
  }
  block_sum<3>(acc, scratch);
  if (threadIdx.x == 0) {
    partials[blockIdx.x * 3 + 0] = acc[0];
    partials[blockIdx.x * 3 + 1] = acc[1];
    partials[blockIdx.x * 3 + 2] = acc[2];
  }
}

template <typename R>
int fbicgstab_2d(KrylovCtx<R>& k, int slot, const std::function<void(const cplx<R>*, cplx<R>*)>& prec) {
  autotune_columns(k);
  const long long n = k.nfull;
  const int nbv = vec_blocks(n);
  KScal* ks = k.ks + slot;
  hipLaunchKernelGGL((k_fkry_init<R>), dim3(nbv), dim3(256), 0, k.st, k.b, k.r, k.p, k.x, n, k.Nx, k.Ny, k.S4,
                     k.kx0, k.ky, k.ds, k.partials);
  fin_launch(k, nbv, slot, 0, 1);
  SAV_LAUNCH_CHECK();
  SAV_HIP(hipMemcpyAsync(k.hks + slot, ks, sizeof(KScal), hipMemcpyDeviceToHost, k.st));
  SAV_HIP(hipStreamSynchronize(k.st));
  for (int it = 0; it < k.itmax && !k.hks[slot].done; ++it) {
    k.cur_iter = it + 1;
    prec(k.p, k.Bp);                                        // p_hat
    int nb3 = matvec_launch(k, k.Bp, k.v, 5, slot, k.b);    // v = A p_hat, <c, v>
    fin_launch(k, nb3, slot, 1, 1);
    hipLaunchKernelGGL((k_kry_upd1<R>), dim3(nbv), dim3(256), 0, k.st, k.r, k.v, k.Bp, k.s, k.x, n, ks, 0LL);
    prec(k.s, k.Bs);                                        // s_hat
    nb3 = matvec_launch(k, k.Bs, k.t, 6, slot, k.s);        // t = A s_hat, <t, s>, <t, t>
    fin_launch(k, nb3, slot, 2, 1);
    hipLaunchKernelGGL((k_fkry_upd2<R>), dim3(nbv), dim3(256), 0, k.st, k.s, k.Bs, k.t, k.b, k.x, k.r, n, k.Nx,
                       k.Ny, k.S4, k.kx0, k.ky, k.ds, ks, k.partials);
    fin_launch(k, nbv, slot, 3, 1);
    hipLaunchKernelGGL((k_kry_upd3<R>), dim3(nbv), dim3(256), 0, k.st, k.r, k.v, k.p, n, ks, 0LL);
    SAV_LAUNCH_CHECK();
    SAV_HIP(hipMemcpyAsync(k.hks + slot, ks, sizeof(KScal), hipMemcpyDeviceToHost, k.st));
    SAV_HIP(hipStreamSynchronize(k.st));
  }
  return k.hks[slot].iter;
}

template int fbicgstab_2d<double>(KrylovCtx<double>&, int,
                                  const std::function<void(const cplx<double>*, cplx<double>*)>&);
template int fbicgstab_2d<float>(KrylovCtx<float>&, int, const std::function<void(const cplx<float>*, cplx<float>*)>&);
template void matvec_plain_2d<double>(KrylovCtx<double>&, const cplx<double>*, cplx<double>*, int);
template void matvec_plain_2d<float>(KrylovCtx<float>&, const cplx<float>*, cplx<float>*, int);
template int bicgstab_2d<double>(KrylovCtx<double>&, int);
template int bicgstab_2d<float>(KrylovCtx<float>&, int);
template int bicgstab_batch<double>(KrylovCtx<double>&, int, int);
template int bicgstab_batch<float>(KrylovCtx<float>&, int, int);
template void bicgstab_queue<double>(KrylovCtx<double>&, int, int, int);
template void bicgstab_queue<float>(KrylovCtx<float>&, int, int, int);
template void matvec_prec_2d<double>(KrylovCtx<double>&, const cplx<double>*, cplx<double>*, int);
template void matvec_prec_2d<float>(KrylovCtx<float>&, const cplx<float>*, cplx<float>*, int);

}  // namespace sav
