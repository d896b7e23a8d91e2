/*
 * sav_abi.h -- C ABI of libsav_hip.so, an MI355X (gfx950) implementation of the
 * seven-step pseudo-spectral SAV timestep of the Julia code
 * Forstylime/Julia-Vesicles-growth-and-shrinkage.
 *
 * Conventions
 *   - Plain C types only; every array argument is a caller-owned HOST pointer.
 *   - Arrays are column-major with x fastest, exactly the Julia layout:
 *       phi  : (Nx, Ny, N)        3D: (Nx, Ny, Nz, N)     double
 *       phi_hat: (Nx/2+1, Ny, N)  3D: (Nx/2+1, Ny, Nz, N) complex (interleaved re, im)
 *       u    : (Nx, Ny, 2)        3D: (Nx, Ny, Nz, 3)
 *       p    : (Nx, Ny)           3D: (Nx, Ny, Nz)
 *     Complex values are interleaved double[2] -- layout-compatible with ComplexF64.
 *   - The library copies during the call and never retains host pointers.
 *   - Every entry point returns an int status: 0 = OK, >0 = warning, <0 = error.
 *   - A handle is single-threaded; it owns one HIP stream and all device state.
 */
#ifndef SAV_ABI_H
#define SAV_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SAV_ABI_VERSION 1

/* status codes */
#define SAV_OK 0
#define SAV_W_KRYLOV_NOT_CONVERGED 1 /* @warn "BiCGSTAB 未收敛" (src/Solvers.jl:76); step committed */
#define SAV_W_THETA2_NONNEG 2        /* @warn when theta_2 >= 0 (src/Solvers_3d.jl:388); step committed */
#define SAV_E_ARG (-1)
#define SAV_E_HIP (-2)
#define SAV_E_WC_NEGATIVE (-3) /* error("FATAL: W + C < 0") (src/Utils.jl:162,180,188); step not committed */
#define SAV_E_NONFINITE (-4)   /* new: NaN/Inf in the committed scalars */
#define SAV_E_UNSUPPORTED (-5)

#define SAV_PREC_FP64 0
#define SAV_PREC_FP32 1 /* FP32 fields/FFTs, FP64 reductions, 3x3 solve and Krylov scalars */
/* FP64 fields and every stage in FP64 except the 2D psi BiCGSTAB's inner iterations: an
 * FP32 BiCGSTAB on the residual equation inside an FP64 defect correction that stops on the
 * reference's preconditioned residual test (||P(b - A x)|| <= atol + tol ||P b||,
 * src/Solvers.jl:73-75) taken on the FP64 true residual.  3D steps are those of FP64. */
#define SAV_PREC_MIXED 2

#define SAV_PRECOND_DIAG 0
#define SAV_PRECOND_LEARNED 1
#define SAV_PRECOND_STENCIL 2 /* first-order Neumann stencil in place of the network (see below) */

#define SAV_PRESENT 0
#define SAV_OLD 1

typedef struct sav_handle sav_handle_t;

/*
 * Mirrors `Config` field for field (src/Types.jl:27-74, kw-ctor :101-134;
 * src/Types_3d.jl adds Nz, Lz).  The Int fields of Config are int64 here.
 * psi_in_v / psi_out_v point to N doubles each (copied at sav_create).
 */
typedef struct sav_config {
  int32_t ndim; /* 2 or 3 */
  int32_t N;
  int32_t Nx, Ny, Nz; /* Nz ignored in 2D */
  double Lx, Ly, Lz;
  double epsilon, M_phi, M0_psi, eta, gamma_surf;
  int64_t gamma_area;
  double gamma_bend;
  int64_t gamma_in, beta_in, gamma_out, beta_out, lamda;
  const double* psi_in_v;
  const double* psi_out_v;
  double S1, S2, S3, S4, C1, C2, C3;
  double dt, T, tol;
  /* additions of this library */
  int32_t precision; /* SAV_PREC_* */
  int32_t precond;   /* SAV_PRECOND_* */
  int32_t itmax;     /* BiCGSTAB iteration cap; 20 in src/Solvers.jl:75 */
  double atol;       /* BiCGSTAB absolute tolerance; 1e-14 in src/Solvers.jl:74 */
  int32_t device;
} sav_config_t;

typedef struct sav_step_stats {
  int32_t krylov_iters[3]; /* psi_11, psi_12, psi_21 solves (0 in 3D) */
  int32_t krylov_solved[3];
  double krylov_rnorm[3];  /* final preconditioned residual norm */
  double energy;           /* compute_modified_energy after the step (src/Utils.jl:96-141) */
  double area_ratio;       /* sum|A - A0| / sum A0 (src/simulation.jl:147) */
  double Q, R1, R2, R3;
  int32_t status;
} sav_step_stats_t;

/* ABI version (SAV_ABI_VERSION). */
int sav_abi_version(void);

/* Builds the handle: the `Operators` bundle of build_operators (src/SpectralUtils.jl:36-118),
 * device state, twiddles and work buffers.  Replaces Config(...)+build_operators(conf).
 * Grid lengths (FFTW's plans accept any, SpectralUtils.jl:61-71): Nx even, and Nx, Nx/2, Ny
 * (and Nz) each 2^a, 3 2^a or 5 2^a up to 1280, or 2048; in 2D Nx, Ny >= 4.  Other lengths
 * return SAV_E_ARG with the rule in sav_last_error. */
int sav_create(const sav_config_t* cfg, sav_handle_t** out);

/* Frees the handle and its device memory. */
int sav_destroy(sav_handle_t* h);

/* Last error message of the handle (or the last failed sav_create when h == NULL). */
const char* sav_last_error(const sav_handle_t* h);

/*
 * Copies a FieldState in or out (src/Types.jl:167-201).  which = SAV_PRESENT or SAV_OLD.
 * scalars = {Q, R1, R2, R3}; A0 has N entries (present only; ignored for OLD); dt is
 * FieldState.dt.  Any array pointer may be NULL on get (skipped); on set all are required.
 * As in the reference's loop (simulation.jl:99-104), the physical fields of a state set here
 * are expected to be the back-transforms of its spectra (up to round-off, e.g. phi_hat =
 * rfft(phi) of an initial condition): a step takes its inner products and the functionals'
 * derivatives from the spectra.  mu, nu and p of a stepped state are back-transformed when
 * it is read.
 */
int sav_set_state(sav_handle_t* h, int which, const double* phi, const double* phi_hat, const double* psi,
                  const double* psi_hat, const double* u, const double* u_hat, const double* p,
                  const double* p_hat, const double* mu, const double* mu_hat, const double* nu,
                  const double* nu_hat, const double* scalars, const double* A0, double dt);
int sav_get_state(sav_handle_t* h, int which, double* phi, double* phi_hat, double* psi, double* psi_hat,
                  double* u, double* u_hat, double* p, double* p_hat, double* mu, double* mu_hat, double* nu,
                  double* nu_hat, double* scalars, double* A0, double* dt);

/* SAV initialisation of src/simulation.jl:24-53 on the PRESENT state (phi, psi and their
 * spectra must be set): A0, R1..R3, Q = 1, mu_hat, nu_hat, mu, nu; then old = present. */
int sav_init_sav(sav_handle_t* h);

/* One time-loop body (src/simulation.jl:73-110, monitor :146-148): solve_step1..7 of
 * src/Solvers.jl (src/Solvers_3d.jl), state rotation, back-transforms and the monitor.
 * n is the 1-based loop index (BDF1 for n == 1, bdf_coeff, src/Types.jl:426). */
int sav_step(sav_handle_t* h, double dt, int n, sav_step_stats_t* stats);

/* Runs n_nodes-1 steps over the time nodes t_nodes (dt = t[k+1]-t[k]; loop index
 * starts at n0), writing energy/area histories (length n_nodes-1, may be NULL).
 * The run_simulation loop of src/simulation.jl:71-152 without plots or .mat I/O. */
int sav_run(sav_handle_t* h, const double* t_nodes, int n_nodes, int n0, double* energy_hist,
            double* area_hist, int32_t* krylov_iters_total);

/* Transforms on host buffers (test hooks for to_spectral!/to_physical!,
 * src/SpectralUtils.jl:132-148, and the full-complex plans :69-71).
 * count = number of scalar fields (e.g. N, or 2 for a 2D velocity). */
int sav_rfft(sav_handle_t* h, const double* in, double* out_hat, int count);
int sav_irfft(sav_handle_t* h, const double* in_hat, double* out, int count);
int sav_fft_full(sav_handle_t* h, const double* in, double* out, int count, int inverse);

/* Allocation-free irfft on DEVICE buffers.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The same transform as sav_irfft (to_physical!,
 * src/SpectralUtils.jl:144-148, input untouched), on the handle's stream, returning when
 * out_dev is written.  in_hat_dev: count half spectra, out_dev: count fields, both in the
 * handle's storage type (double for SAV_PREC_FP64 / SAV_PREC_MIXED, float for SAV_PREC_FP32).
 * The column pass's intermediate lives in a file-scope device array of
 * SAV_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends); the call makes no
 * hipMalloc/hipFree.  SAV_E_ARG when count * (Nx/2+1) * Ny [* Nz] complex values exceed it.
 * Calls that overlap on one GPU (several handles or threads) share that array and race. */
#define SAV_NOALLOC_SCRATCH_BYTES (64u << 20)
int sav_irfft_noalloc(sav_handle_t* h, const void* in_hat_dev, void* out_dev, int count);

/* Functionals at (phi, psi): H1, H2, H3, G (each N fields), W[3], area[N]
 * (get_H1/get_H2/get_H3/get_MG, get_W1..3, calculate_area; src/Utils.jl:66-249).
 * A0 is taken from the present state.  Returns SAV_E_WC_NEGATIVE like the reference. */
int sav_eval_functionals(sav_handle_t* h, const double* phi, const double* psi, double* H1, double* H2, double* H3,
                         double* G, double* W, double* area);

/* 2D psi solve of src/Solvers.jl:39-81 with M_psi built from phi_star and the
 * BDF coefficient a and dt: out_hat = solve_psi!(rhs_hat).  iters/solved/rnorm may be NULL. */
int sav_psi_solve(sav_handle_t* h, const double* phi_star, double a, double dt, const double* rhs_hat,
                  double* out_hat, int32_t* iters, int32_t* solved, double* rnorm);

/* compute_modified_energy(present, old) (src/Utils.jl:96-141) and the area ratio. */
int sav_modified_energy(sav_handle_t* h, double* energy, double* area_ratio);

/* Autotuner of the 2D psi matvec (no reference counterpart): at the first psi solve of a
 * handle the column-FFT kernels are timed with 8, 16 and 4 columns per workgroup (three
 * matvecs each on the solve's right-hand side; a width whose kernels resolve to those of a
 * width already timed is skipped) and the fastest is kept; all compute the same numbers
 * (the dot products are summed per column).  sav_set_column_tile(h, 4 | 8 | 16) fixes the
 * width, 0 retunes at the next solve.  sav_autotune_info returns the width in use (0 before
 * the first solve) and the measured milliseconds per matvec with 8 and 16 columns;
 * sav_column_tile_info adds the 4-column time (0: not timed). */
int sav_set_column_tile(sav_handle_t* h, int columns);
int sav_autotune_info(sav_handle_t* h, int32_t* columns, double* ms_8, double* ms_16);
int sav_column_tile_info(sav_handle_t* h, int32_t* columns, double* ms_4, double* ms_8, double* ms_16);

/* Order of the 2D psi matvec's three kernels over the systems of a batched solve (no
 * reference counterpart): 1 = every kernel covers all systems; 2 = system by system, each
 * through all three kernels over one shared pair of intermediate fields, which then stays
 * in the GPU's infinity cache between the kernels instead of going to HBM.  0 (the default)
 * times both at the first psi solve, with the chosen column width, and keeps the faster.
 * Both orders compute the same numbers.  sav_matvec_order_info returns the order in use
 * (0 before the first solve) and the measured milliseconds per matvec in order 1 and 2
 * (0: not timed). */
int sav_set_matvec_order(sav_handle_t* h, int order);
int sav_matvec_order_info(sav_handle_t* h, int32_t* order, double* ms_batched, double* ms_sequenced);

/* Autotuner of every other column (strided-axis) transform (no reference counterpart): at a
 * handle's first step inverse transforms (irfft of N and of 4N fields of the present state)
 * and forward transforms (rfft of the same) are timed with 4, 8 and 16 columns per workgroup
 * (widths that resolve to the same tile for this grid are timed once) and the fastest width
 * of each direction is kept; all widths compute bit-identical results.
 * sav_set_transform_tile(h, 4 | 8 | 16) fixes both widths, 0 retunes at the next step.
 * sav_transform_tile_info (inverse) and sav_transform_tile_fwd_info (forward) return the
 * width in use (0 before the first step) and the measured milliseconds of each width (-1:
 * same tile as a smaller width, 0: not timed yet). */
int sav_set_transform_tile(sav_handle_t* h, int columns);
int sav_transform_tile_info(sav_handle_t* h, int32_t* columns, double* ms_4, double* ms_8, double* ms_16);
int sav_transform_tile_fwd_info(sav_handle_t* h, int32_t* columns, double* ms_4, double* ms_8, double* ms_16);

/* (x, y) plane kernels of 3D grids with 128 x 128 planes (no reference counterpart; they
 * carry SpectralUtils_3d.jl's rfft/irfft over x and y).  SAV_PLANE_REGISTER keeps the plane's
 * columns in registers and stages a 64 KB half plane (FP64) through LDS; SAV_PLANE_LDS holds
 * the whole half-spectrum plane in LDS.  Both compute the same transform (results agree to
 * FP round-off, not bit for bit).  With SAV_PLANE_AUTO (the default) the next step times an
 * irfft and an rfft of the present phi with each and keeps SAV_PLANE_REGISTER unless
 * SAV_PLANE_LDS is more than 3 % faster.
 * sav_plane_kernel_info returns the kind in use (SAV_PLANE_AUTO before it is chosen) and
 * the measured milliseconds of each (0: not timed). Other grids use SAV_PLANE_LDS where the
 * plane fits in LDS; sav_set_plane_kernel(h, SAV_PLANE_REGISTER) is SAV_E_UNSUPPORTED there. */
#define SAV_PLANE_AUTO 0
#define SAV_PLANE_REGISTER 1
#define SAV_PLANE_LDS 2
int sav_set_plane_kernel(sav_handle_t* h, int kind);
int sav_plane_kernel_info(sav_handle_t* h, int32_t* kind, double* ms_register, double* ms_lds);

/* Graph replay of steps (no reference counterpart; the loop body of src/simulation.jl:73-110
 * as one HIP graph).  With on = 1, sav_step and sav_run capture a whole step (steps 1-7,
 * the back-transform of the new state, the monitor) once per buffer rotation and psi
 * iteration budget, and replay it: one graph launch and one host synchronisation per
 * step.  The graph queues one psi iteration more than the previous step needed; a
 * system still iterating after that budget makes the step run again without the graph
 * (present and old are untouched by the replay).  Results are bit-identical to on = 0
 * (the default).  Steps run without the graph while a kernel timing hook is on, with the
 * learned or stencil preconditioner, and before the first 2D psi solve.  on = 2 decides by
 * measurement: from the third step on, once steps can be replayed, 5 eager and 5 replayed
 * steps (not counting those that capture a graph) are timed on the host, and replay is kept
 * if its median step is more than 3 % shorter (it is for launch-bound small grids).
 * sav_graph_info: steps replayed, steps repeated without the graph, graphs held.
 * sav_graph_choice: -1 while on = 2 is still measuring, else 0 (eager) or 1 (replay). */
int sav_set_graph(sav_handle_t* h, int on);
int sav_graph_choice(sav_handle_t* h, int32_t* choice);
int sav_graph_info(sav_handle_t* h, long long* replays, long long* repeats, long long* graphs);

/* Milliseconds of device time of the last sav_run / sav_step (HIP events on the handle's stream). */
int sav_last_elapsed_ms(sav_handle_t* h, double* ms);

/* Measurement hook (no reference counterpart): brackets every launch of one 2D
 * psi-matvec kernel with HIP events on the handle's stream during the following
 * steps.  kernel_id: 1 column inverse FFTs, 2 row FFT*M*FFT, 3 column forward FFTs
 * + preconditioner, 0 off.  Launches queued past convergence are not counted.
 * _get returns the launch count, their summed duration and the algorithmic HBM
 * bytes per launch (reads + writes of whole vectors). */
int sav_kernel_timing(sav_handle_t* h, int kernel_id);
int sav_kernel_timing_get(sav_handle_t* h, int64_t* launches, double* total_ms, double* bytes_per_launch);

/*
 * Learned psi preconditioner (no reference counterpart; it replaces the
 * `prec_func!`/`P_inv_diag` plug-in of src/Solvers.jl:57-65 in the 2D psi solve).
 *   z = P ( r + F( g ) ),   g(x) = kappa sum_tap w_tap(x) Re F^{-1}(P r)(x + d(tap))
 * P the diagonal preconditioner, F the full 2D FFT, kappa = S4 max(M_psi) / (dx dy),
 * and w(x) the weights at pixel x (tap = di*3 + dj, offset (di-1, dj-1), periodic) of the
 * 5-point stencil of -div(c grad .) with face averages of a coefficient c(x) (the analytic
 * first-order stencil is c = 1 - M_psi / max M_psi), c being the first output channel of a
 * periodic 3x3 convolutional network 2 -> 16 -> 16 -> 16 -> 9 with GELU, run once per
 * mobility field on the inputs
 * (4 (M_psi / max M_psi - 0.75), 1): BF16 weights and activations on the MFMA units,
 * FP32 accumulation.  The preconditioner is linear in r, as BiCGSTAB needs.  The
 * psi solve is then a flexible right-preconditioned BiCGSTAB that stops on the same
 * diagonally preconditioned residual as the reference's left-preconditioned solve
 * (||P r|| <= atol + rtol ||P b||).  2D only; needs Nx % 32 == 0 and Ny % 16 == 0
 * (else SAV_E_UNSUPPORTED).
 *
 * Parameters: a flat float vector of SAV_NET_NPARAM values, for each layer
 * W[cout][cin][3][3] (cross-correlation, periodic) then b[cout].
 */
#define SAV_NET_NPARAM 6249

/* Selects the psi preconditioner of the following solves: SAV_PRECOND_DIAG,
 * SAV_PRECOND_LEARNED, or SAV_PRECOND_STENCIL, the same form with the analytic
 * first-order stencil g = -S4 div((max M - M) grad v) (5-point, face-averaged
 * coefficients) in place of the network: z = P (2 r - A P r) up to discretisation. */
int sav_set_precond(sav_handle_t* h, int precond);
/* He-uniform hidden layers from a SplitMix64 seed, zero output layer (the network then adds nothing). */
int sav_precond_init(sav_handle_t* h, uint64_t seed);
/* Copies the network parameters in (load) or out (save); n must be SAV_NET_NPARAM. */
int sav_precond_load(sav_handle_t* h, const float* params, int n);
int sav_precond_save(sav_handle_t* h, float* params, int n);
/* on != 0: the following psi solves run with the diagonal preconditioner and log training
 * pairs: for every iterate x_k of a solve, e = x* - x_k, r = A e, and patches_per_snapshot
 * random 64 x 64 windows of the network input, v = Re F^{-1}(P r) / s and two targets:
 * the correction Re F^{-1}(P^{-1} e - r) / (kappa s) of the solver's error, and the
 * analytic first-order stencil's output on v (SAV_PRECOND_STENCIL's weights);
 * s = rms of F^{-1}(P r).
 * The store is cleared when collection starts and holds up to 24576 windows.
 * on == 0 stops logging (the store is kept for sav_precond_train). */
int sav_precond_collect(sav_handle_t* h, int on, int patches_per_snapshot, uint64_t seed);
/* Windows in the store, and the number of solves with the learned (or stencil)
 * preconditioner that did not reach the tolerance within itmax and were therefore
 * repeated from x0 = 0 with the diagonal preconditioner (either pointer may be NULL). */
int sav_precond_samples(sav_handle_t* h, int64_t* n_windows, int64_t* fallbacks);
/* Adam (beta 0.9/0.999) on the masked MSE (4-pixel border excluded) of the stencil
 * prediction -div(c grad v) (the form above) against the target, over random batches of the
 * store, learning rate decaying from lr to lr/10 (cosine).  loss_first/loss_last are the
 * mean losses of the first and last 10 iterations. */
int sav_precond_train(sav_handle_t* h, int iters, int batch, double lr, uint64_t seed, double* loss_first,
                      double* loss_last);
/* Targets of the following sav_precond_train calls: 0 (the default) the solver's error
 * corrections, 1 the analytic first-order stencil's output on the same windows
 * (distillation: the network learns that stencil, which a 3x3 network of the mobility can
 * represent, and can then be trained further on the error corrections). */
int sav_precond_train_target(sav_handle_t* h, int target);
/* Test hooks: the network on host images x[nimg][H][W][2] (channel fastest, rounded to
 * BF16) -> y[nimg][H][W][9]; and the masked MSE of y against t[nimg][H][W][9] with the
 * parameter gradient (flat layout).  H % 16 == 0, W % 32 == 0. */
int sav_net_forward(sav_handle_t* h, const float* x, int nimg, int H, int W, float* y);
int sav_net_grad(sav_handle_t* h, const float* x, const float* t, int nimg, int H, int W, int border,
                 float* grads, double* loss);

/*
 * Ensemble of independent instances (no reference counterpart: the reference runs
 * one simulation per process, scripts/Main.jl:86-114).  Each member is an ordinary
 * handle with its own stream; sav_ensemble_run drives all members over the same time
 * nodes from n_threads host threads, so one member's Krylov convergence waits overlap
 * the others' kernels.  Members are set up through sav_ensemble_member (borrowed
 * handles: sav_set_state, sav_init_sav, sav_get_state ...; never sav_destroy them).
 * energy_hist / area_hist: [n][n_nodes-1] (may be NULL); krylov_iters, status: [n]
 * (may be NULL).  Returns the most severe member status.
 */
typedef struct sav_ensemble sav_ensemble_t;
int sav_ensemble_create(const sav_config_t* cfgs, int n, sav_ensemble_t** out);
int sav_ensemble_destroy(sav_ensemble_t* e);
const char* sav_ensemble_last_error(const sav_ensemble_t* e);
int sav_ensemble_size(const sav_ensemble_t* e);
sav_handle_t* sav_ensemble_member(sav_ensemble_t* e, int i);
int sav_ensemble_run(sav_ensemble_t* e, const double* t_nodes, int n_nodes, int n0, int n_threads,
                     double* energy_hist, double* area_hist, int32_t* krylov_iters, int32_t* status);

/* Size classes (no reference counterpart): with on = 1, members with the same 2D grid, vesicle
 * count, storage precision (FP64 or FP32 fields; not mixed), S4 and psi-solve settings
 * (diagonal preconditioner, tol, atol, itmax) form groups of 2 to 64 whose three psi solves per
 * step run as ONE batched BiCGSTAB over all their systems (every Krylov and matvec launch
 * covers the size class, the system on a grid dimension).  Each member's iterates, iteration
 * counts and states stay bit-identical to its own solves.  sav_ensemble_run then steps every
 * member in a host thread of its own (n_threads is not used); members meet at each step's
 * solve, which the last to arrive runs.  Outside sav_ensemble_run, and with on = 0, members
 * solve on their own (a member may be stepped alone through its handle).  The groups' pools
 * replace the members' own psi-solve buffers and stay in use.  n_groups
 * returns the number of groups.  sav_ensemble_batched_info: each member's group (-1: none)
 * and the number of batched solves its group ran (either array may be NULL). */
int sav_ensemble_set_batched(sav_ensemble_t* e, int on, int32_t* n_groups);
int sav_ensemble_batched_info(sav_ensemble_t* e, int32_t* group_of, int64_t* solves);

#ifdef __cplusplus
}
#endif
#endif
