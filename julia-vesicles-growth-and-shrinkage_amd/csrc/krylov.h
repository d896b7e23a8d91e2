// psi BiCGSTAB of src/Solvers.jl:39-81 on full complex spectra (2D).
#pragma once
#include <functional>

#include "fft.h"
#include "solver.h"

namespace sav {

// Systems one batched solve can cover: the three solves of a step for each of up to 64
// members of a size-class group (sav_ensemble_set_batched).  System z of a batch belongs
// to member z / 3: the kernels take its mobility at M + (z / 3) * nfull and its step scalars
// at ds + z / 3 (a handle's own solves have z < 3).
constexpr int KRY_MAX_SYS = 3 * 64;

template <typename R>
struct KrylovCtx {
  int Nx, Ny, N;
  long long nfull;            // N * Nx * Ny
  const double* kx0;          // fftfreq wavenumbers over Nx (Nyquist -)
  const double* ky;
  const cplx<R>* tw_x;        // length-Nx table
  const cplx<R>* tw_y;        // length-Ny table
  const cplx<R>* tw_ri;       // row plan of the matvec (row_plan_maxr), inverse: k-major table
  const cplx<R>* tw_rf;       // the same plan reversed (forward)
  const cplx<R>* tw_ci;       // column plans of the matvec (col_plan_maxr), inverse kernel
  const cplx<R>* tw_cf;       // forward kernel (k-major tables)
  const R* M;                 // M_psi(phi*), physical (N, Ny, Nx)
  const DScal* ds;            // a, dt, Mmax
  double S4, atol, rtol;
  int itmax;
  cplx<R> *x, *r, *p, *v, *s, *t, *b, *Bx, *By;
  double* partials;           // >= nsys * ps doubles
  long long vs = 0;           // elements between the vectors of consecutive systems of a batch
  long long ps = 0;           // partials between consecutive systems
  KScal* ks;                  // device, one per solve slot
  KScal* hks;                 // pinned host mirror
  hipStream_t st;
  int guess[KRY_MAX_SYS];     // expected iteration count per slot (batching of convergence checks)
  // optional event timing of one matvec kernel (1: column inverse, 2: rows, 3: column forward):
  // every ev_stride-th launch of it carries start/stop events (ev_seq counts its launches)
  int timed = 0;
  hipEvent_t* ev = nullptr;   // 2 * ev_cap events
  int ev_cap = 0, ev_used = 0;
  int ev_stride = 9;
  long long ev_seq = 0;
  int cur_iter = 0;
  int ev_iter[512];
  int ev_sys[512];             // system of a single-system launch, -1: all systems
  double ev_bytes[512];
  long long tk_count = 0;
  double tk_ms = 0.0, tk_bytes = 0.0;
  // iterate snapshots for logging training pairs (nullptr: off)
  cplx<R>* snap = nullptr;
  int snap_cap = 0, snap_n = 0;
  // columns per workgroup of the matvec's column kernels: 0 = not tuned yet, 8 or 16
  int colw = 0;
  // order of a matvec's kernels over the systems of a batch: 0 = not chosen yet, 1 = all
  // systems per launch, 2 = system by system over one shared intermediate pair (infinity-
  // cache resident); chosen after the column width at the first solve.  seq = (order == 2).
  int order = 0;
  int seq = 0;
  int tune_nsys = 3;          // systems the matvec autotuner times (a step's batch; a group's)
  double tune_order_ms[2] = {0.0, 0.0};  // measured matvec time in order 1 and 2 (0: not timed)
  double tune_ms[3] = {0.0, 0.0, 0.0};  // measured matvec time with 8, 16 and 4 columns (0: not timed)
  // preconditioned directions of the flexible solver
  cplx<R>* Bp = nullptr;
  cplx<R>* Bs = nullptr;
  // inner solve of a mixed-precision solve: system z is skipped (done, zero right-hand
  // side) when outer[slot + z] is done
  const KScal* outer = nullptr;
};

// Grid lengths (Nx, Ny) the 2D psi solve is built for.
bool kry_size_supported(int n);

// Solves A x = b with b already in ctx.b; x in ctx.x on return.  Returns the
// number of iterations; ctx.hks[slot] holds the final Krylov scalars.
template <typename R>
int bicgstab_2d(KrylovCtx<R>& k, int slot);
// nsys independent systems solved together (every launch covers all of them): system z
// uses the vectors at offset z * ctx.vs (b, x, r, p, v, s, t, Bx, By) and KScal slot + z;
// each system's iterates are exactly those of its own solve.  Returns the largest
// iteration count; ctx.hks[slot + z] holds each system's final scalars.
template <typename R>
int bicgstab_batch(KrylovCtx<R>& k, int slot, int nsys);
// The same solve queued with exactly `iters` iterations (capped at itmax) and no host
// synchronisation, for graph capture: a system that converges earlier skips the
// remaining kernels.  ctx.colw must be chosen already.  The caller reads ctx.ks.
template <typename R>
void bicgstab_queue(KrylovCtx<R>& k, int slot, int nsys, int iters);

// A*y then the diagonal preconditioner: out = P^{-1} (a/dt y - sum_d D1f_d F(M F^{-1}(D1f_d S4 y)))
// (one matvec + preconditioner, used by tests through the solver).
template <typename R>
void matvec_prec_2d(KrylovCtx<R>& k, const cplx<R>* y, cplx<R>* out, int slot);

// out = A y without the preconditioner (queued, no synchronisation).
template <typename R>
void matvec_plain_2d(KrylovCtx<R>& k, const cplx<R>* y, cplx<R>* out, int slot);

// Mixed-precision solve of nsys systems (SAV_PREC_MIXED): FP64 iterate k.x, right-hand
// side k.b and residual, FP32 inner BiCGSTAB (k32, its own rtol and itmax) on the residual
// equation A e = b - A x, x += e, until the diagonally preconditioned FP64 residual
// ||P^{-1}(b - A x)|| <= k.atol + k.rtol ||P^{-1} b|| (the reference's stopping test,
// Solvers.jl:73-75, taken on the true residual) or max_rounds rounds.  k32 shares the
// operator tables, k.ds and the partials buffer; k32.M holds M_psi in FP32.  k.hks[slot+z]
// returns: iter = inner iterations of all rounds, rounds, rnorm = final true residual,
// solved.  With `warm` the solve starts from the iterate k.x already holds (the previous
// solve's solution) instead of 0: one more FP64 matvec, fewer inner iterations; the stopping
// test is the same.  Returns the largest inner iteration total.
int bicgstab_refined(KrylovCtx<double>& k, KrylovCtx<float>& k32, int slot, int nsys, int max_rounds, bool warm);

// Flexible right-preconditioned BiCGSTAB with preconditioner `prec` (see krylov.hip).
template <typename R>
int fbicgstab_2d(KrylovCtx<R>& k, int slot, const std::function<void(const cplx<R>*, cplx<R>*)>& prec);

}  // namespace sav
