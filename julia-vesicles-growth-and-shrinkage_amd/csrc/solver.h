// Device-resident SAV solver (one grid, 2D or 3D, FP64 or FP32 fields).
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "../../include/sav_abi.h"
#include "fft.h"

namespace sav {

constexpr int MAXN = 16;  // vesicles per instance

// Scalars of one FieldState (Q, R1, R2, R3 of src/Types.jl:193-196).
struct DStateScal {
  double Q, R1, R2, R3;
};

// Device scalars of one step.  All reductions and scalar solves land here, so a
// step runs without host round trips except for Krylov convergence checks.
struct DScal {
  double a, b, c, dt, dt5;  // BDF coefficients, dt of the step, dt used by step 5
  double W1, W2, W3, R1s, R2s, R3s;
  double area[MAXN];
  double Mmax;
  double Rs[6];  // R11 R21 R31 R12 R22 R32
  double lhs[9], rhs1[3], rhs2[3];
  double theta1, theta2, Q;
  double energy, area_ratio;
  double area_mon[MAXN];
  int flags;  // 1: W1+C1<0, 2: W2+C2<0, 4: W3+C3<0, 8: theta2 >= 0 (3D), 16: non-finite
  int pad;
};

// Krylov scalars of one BiCGSTAB solve (Krylov.jl bicgstab, restated).
struct KScal {
  double ar, ai;  // alpha
  double wr, wi;  // omega
  double rr, ri;  // rho
  double nr, ni;  // next rho
  double br, bi;  // beta
  double rnorm, eps;
  int iter, done, solved, status;  // status: 0 running, 1 solved, 2 itmax, 3 breakdown, 4 zero rhs, 5 bc=0
  int rounds;                      // defect-correction rounds of a mixed-precision solve
  int pad;
};

struct HostCfg {
  sav_config_t c;
  std::vector<double> psi_in, psi_out;
};

class ISolver {
 public:
  virtual ~ISolver() = default;
  virtual void set_state(int which, const double* const* arrays, const double* scal, const double* A0, double dt) = 0;
  virtual void get_state(int which, double* const* arrays, double* scal, double* A0, double* dt) = 0;
  virtual int init_sav() = 0;
  virtual int step(double dt, int n, sav_step_stats_t* st) = 0;
  virtual int run(const double* t, int nn, int n0, double* eh, double* ah, int32_t* kit) = 0;
  virtual void rfft_host(const double* in, double* out, int count) = 0;
  virtual void irfft_host(const double* in, double* out, int count) = 0;
  virtual void fft_full_host(const double* in, double* out, int count, int inverse) = 0;
  virtual int eval_functionals(const double* phi, const double* psi, double* H1, double* H2, double* H3, double* G,
                               double* W, double* area) = 0;
  virtual int psi_solve_host(const double* phi_star, double a, double dt, const double* rhs, double* out,
                             int32_t* iters, int32_t* solved, double* rnorm) = 0;
  virtual int modified_energy(double* E, double* ar) = 0;
  // irfft of device buffers with the file-scope scratch array as the column-pass
  // intermediate (no allocation); returns SAV_E_ARG when count fields exceed the scratch
  virtual int irfft_noalloc(const void* in_hat, void* out, int count) = 0;
  virtual double last_elapsed_ms() = 0;
  virtual void kernel_timing(int id) = 0;
  virtual void kernel_timing_get(long long* n, double* ms, double* bytes) = 0;
  // learned psi preconditioner (2D)
  virtual int set_precond(int mode) = 0;
  virtual int precond_init(uint64_t seed) = 0;
  virtual int precond_params(bool set, float* p, int n) = 0;
  virtual int precond_collect(int on, int patches_per_snapshot, uint64_t seed) = 0;
  virtual long long precond_samples() = 0;
  virtual long long precond_fallbacks() = 0;
  // psi matvec column-tile autotuner (2D)
  virtual int set_column_tile(int cw) = 0;
  virtual void column_tile(int* cw, double* ms8, double* ms16, double* ms4) = 0;
  // order of the psi matvec's kernels over a batch of systems (2D): 0 = autotuned, 1 = all
  // systems per launch, 2 = system by system over one shared intermediate pair
  virtual int set_matvec_order(int order) = 0;
  virtual void matvec_order(int* order, double* ms_batched, double* ms_sequenced) = 0;
  // column-transform tile autotuner (every transform outside the psi matvec)
  virtual int set_transform_tile(int cw) = 0;
  virtual void transform_tile(int* cw, double* ms4, double* ms8, double* ms16) = 0;      // inverse
  virtual void transform_tile_fwd(int* cw, double* ms4, double* ms8, double* ms16) = 0;  // forward
  // (x, y) plane kernels of 128 x 128 x Nz grids: 0 = autotuned, 1 = register columns with a
  // half plane in LDS, 2 = the whole plane in LDS
  virtual int set_plane_kernel(int kind) = 0;
  virtual void plane_kernel(int* kind, double* ms_register, double* ms_lds) = 0;
  virtual int set_graph(int on) = 0;
  // Eager steps run step 1's advection terms and step 4 on a second stream (on by default);
  // ensemble members, which overlap with each other, keep one stream each.
  virtual void set_step_fork(bool on) = 0;
  virtual void graph_info(long long* replays, long long* repeats, long long* graphs) = 0;
  // -1 while graph mode 2 is still measuring, else 0 (eager) or 1 (replay)
  virtual int graph_choice() const = 0;
  // targets of the following precond_train calls: 0 the solver's error corrections, 1 the
  // analytic first-order stencil's output on the same windows
  virtual int precond_train_target(int kind) = 0;
  virtual int precond_train(int iters, int batch, double lr, uint64_t seed, double* loss_first,
                            double* loss_last) = 0;
  virtual int net_forward(const float* x, int nimg, int Hh, int Ww, float* y) = 0;
  virtual int net_grad(const float* x, const float* t, int nimg, int Hh, int Ww, int border, float* grads,
                       double* loss) = 0;
};

std::unique_ptr<ISolver> make_solver(const HostCfg& cfg);

// Library-internal: the step fork of a handle's solver (see ISolver::set_step_fork).
void set_step_fork(sav_handle_t* h, bool on);
// Library-internal: the solver behind a handle.
ISolver* handle_solver(sav_handle_t* h);

// A size-class group of ensemble members (no reference counterpart).  Members with the same
// 2D grid, vesicle count, storage precision (FP64 or FP32 fields; not mixed), S4 and
// psi-solve settings (diagonal preconditioner, tol, atol, itmax) run the three psi solves of
// a step as ONE batched BiCGSTAB over all their systems: every Krylov and matvec launch
// covers the whole size class, with the system (3 per member) on a grid dimension.  Their
// Krylov vectors, mobility fields and step scalars live in pools the group owns, one slice
// per member; each system's iterates are those of its member's own solve (bit-identical).
// Members step in their own host threads and meet at the solve: the last to arrive runs it.
class ISolveGroup {
 public:
  virtual ~ISolveGroup() = default;
  // Starts a run: every member takes part again.
  virtual void begin_run() = 0;
  // Member m's run has ended (or failed): the others no longer wait for it.
  virtual void leave(int m) = 0;
  // Couples the members' solves (only while all members are stepped together, or a member
  // stepped alone would wait for the others) or decouples them; the pools stay in use.
  virtual void set_coupled(bool on) = 0;
  virtual int size() const = 0;
  // batched solves run so far, and the matvec tile width and order the group tuned
  virtual void info(long long* solves, int* columns, int* order) const = 0;
};
// Builds the group of `members` (at most 64, compatible as above); throws
// std::invalid_argument if they do not qualify.  The group must outlive the members' use of
// their pools: destroy the members first.
std::unique_ptr<ISolveGroup> make_solve_group(const std::vector<ISolver*>& members);

}  // namespace sav
