// C ABI of libsav_hip.so (declared in include/sav_abi.h).
#include <cstring>
#include <string>

#include "krylov.h"
#include "solver.h"

struct sav_handle {
  std::unique_ptr<sav::ISolver> s;
  std::string err;
  int ndim = 2, N = 1;
};

static thread_local std::string g_create_error;

namespace sav {
void set_step_fork(sav_handle_t* h, bool on) {
  if (h && h->s) h->s->set_step_fork(on);
}

ISolver* handle_solver(sav_handle_t* h) { return h ? h->s.get() : nullptr; }
}  // namespace sav

#define SAV_GUARD(h, body)                                  \
  try {                                                     \
    body;                                                   \
  } catch (const sav::HipError& e) {                        \
    (h)->err = e.what();                                    \
    return e.code == SAV_E_UNSUPPORTED ? SAV_E_UNSUPPORTED : SAV_E_HIP; \
  } catch (const std::exception& e) {                       \
    (h)->err = e.what();                                    \
    return SAV_E_ARG;                                       \
  }

extern "C" {

int sav_abi_version(void) { return SAV_ABI_VERSION; }

static bool size_ok(int n) { return sav::fft_size_supported(n) && n % 2 == 0 && sav::fft_size_supported(n / 2); }

int sav_create(const sav_config_t* cfg, sav_handle_t** out) {
  if (!cfg || !out) return SAV_E_ARG;
  *out = nullptr;
  try {
    if (cfg->ndim != 2 && cfg->ndim != 3) throw std::invalid_argument("ndim must be 2 or 3");
    if (cfg->N < 1 || cfg->N > sav::MAXN) throw std::invalid_argument("N out of range");
    if (!size_ok(cfg->Nx) || !sav::fft_size_supported(cfg->Ny) || (cfg->ndim == 3 && !sav::fft_size_supported(cfg->Nz)))
      throw std::invalid_argument(
          "grid sizes: Nx even with Nx, Nx/2, Ny (and Nz) in {2^a, 3 2^a, 5 2^a up to 1280, 2048}");
    if (cfg->ndim == 2 && (!sav::kry_size_supported(cfg->Nx) || !sav::kry_size_supported(cfg->Ny)))
      throw std::invalid_argument("2D grids need Nx, Ny in {4 .. 2048} of the form 2^a, 3 2^a or 5 2^a (psi solve)");
    if (cfg->ndim == 2 && cfg->Ny < 4) throw std::invalid_argument("Ny must be >= 4");
    if (!cfg->psi_in_v || !cfg->psi_out_v) throw std::invalid_argument("psi_in_v/psi_out_v required");
    if (cfg->precision != SAV_PREC_FP64 && cfg->precision != SAV_PREC_FP32 && cfg->precision != SAV_PREC_MIXED)
      throw std::invalid_argument("unknown precision");
    if (cfg->precond != SAV_PRECOND_DIAG && cfg->precond != SAV_PRECOND_LEARNED)
      throw std::invalid_argument("unknown preconditioner");
    if (cfg->precond == SAV_PRECOND_LEARNED && (cfg->ndim != 2 || cfg->Nx % 32 != 0 || cfg->Ny % 16 != 0))
      throw std::invalid_argument("the learned preconditioner needs 2D, Nx % 32 == 0 and Ny % 16 == 0");
    sav::HostCfg hc;
    hc.c = *cfg;
    hc.psi_in.assign(cfg->psi_in_v, cfg->psi_in_v + cfg->N);
    hc.psi_out.assign(cfg->psi_out_v, cfg->psi_out_v + cfg->N);
    if (hc.c.itmax <= 0) hc.c.itmax = 20;
    auto h = new sav_handle();
    try {
      h->s = sav::make_solver(hc);
      if (cfg->precond == SAV_PRECOND_LEARNED) h->s->set_precond(SAV_PRECOND_LEARNED);
    } catch (...) {
      delete h;
      throw;
    }
    h->ndim = cfg->ndim;
    h->N = cfg->N;
    *out = h;
    return SAV_OK;
  } catch (const sav::HipError& e) {
    g_create_error = e.what();
    return SAV_E_HIP;
  } catch (const std::exception& e) {
    g_create_error = e.what();
    return SAV_E_ARG;
  }
}

int sav_destroy(sav_handle_t* h) {
  delete h;
  return SAV_OK;
}

const char* sav_last_error(const sav_handle_t* h) { return h ? h->err.c_str() : g_create_error.c_str(); }

int sav_set_state(sav_handle_t* h, int which, const double* phi, const double* phi_hat, const double* psi,
                  const double* psi_hat, const double* u, const double* u_hat, const double* p, const double* p_hat,
                  const double* mu, const double* mu_hat, const double* nu, const double* nu_hat, const double* scalars,
                  const double* A0, double dt) {
  if (!h) return SAV_E_ARG;
  const double* a[12] = {phi, phi_hat, psi, psi_hat, u, u_hat, p, p_hat, mu, mu_hat, nu, nu_hat};
  for (auto x : a)
    if (!x) {
      h->err = "sav_set_state: all arrays are required";
      return SAV_E_ARG;
    }
  if (!scalars || (which != SAV_PRESENT && which != SAV_OLD)) return SAV_E_ARG;
  SAV_GUARD(h, h->s->set_state(which, a, scalars, A0, dt));
  return SAV_OK;
}

int sav_get_state(sav_handle_t* h, int which, double* phi, double* phi_hat, double* psi, double* psi_hat, double* u,
                  double* u_hat, double* p, double* p_hat, double* mu, double* mu_hat, double* nu, double* nu_hat,
                  double* scalars, double* A0, double* dt) {
  if (!h || (which != SAV_PRESENT && which != SAV_OLD)) return SAV_E_ARG;
  double* a[12] = {phi, phi_hat, psi, psi_hat, u, u_hat, p, p_hat, mu, mu_hat, nu, nu_hat};
  SAV_GUARD(h, h->s->get_state(which, a, scalars, A0, dt));
  return SAV_OK;
}

int sav_init_sav(sav_handle_t* h) {
  if (!h) return SAV_E_ARG;
  int rc = 0;
  SAV_GUARD(h, rc = h->s->init_sav());
  if (rc == SAV_E_WC_NEGATIVE) h->err = "FATAL: W + C < 0 at initialisation";
  return rc;
}

// Error text of a step status shared by sav_step and sav_run (and so the ensemble).
static void set_step_error(sav_handle_t* h, int rc) {
  if (rc == SAV_E_WC_NEGATIVE) h->err = "FATAL: W + C < 0; step not committed";
  if (rc == SAV_E_NONFINITE) h->err = "non-finite SAV scalars; step not committed";
}

int sav_step(sav_handle_t* h, double dt, int n, sav_step_stats_t* stats) {
  if (!h || n < 1 || !(dt > 0)) return SAV_E_ARG;
  int rc = 0;
  SAV_GUARD(h, rc = h->s->step(dt, n, stats));
  set_step_error(h, rc);
  return rc;
}

int sav_run(sav_handle_t* h, const double* t_nodes, int n_nodes, int n0, double* energy_hist, double* area_hist,
            int32_t* krylov_iters_total) {
  if (!h || !t_nodes || n_nodes < 1 || n0 < 1) return SAV_E_ARG;
  int rc = 0;
  SAV_GUARD(h, rc = h->s->run(t_nodes, n_nodes, n0, energy_hist, area_hist, krylov_iters_total));
  set_step_error(h, rc);
  return rc;
}

int sav_rfft(sav_handle_t* h, const double* in, double* out_hat, int count) {
  if (!h || !in || !out_hat || count < 1) return SAV_E_ARG;
  SAV_GUARD(h, h->s->rfft_host(in, out_hat, count));
  return SAV_OK;
}

int sav_irfft(sav_handle_t* h, const double* in_hat, double* out, int count) {
  if (!h || !in_hat || !out || count < 1) return SAV_E_ARG;
  SAV_GUARD(h, h->s->irfft_host(in_hat, out, count));
  return SAV_OK;
}

int sav_irfft_noalloc(sav_handle_t* h, const void* in_hat_dev, void* out_dev, int count) {
  if (!h) return SAV_E_ARG;
  if (!in_hat_dev || !out_dev || count < 1) {
    h->err = "sav_irfft_noalloc: null buffer or count < 1";
    return SAV_E_ARG;
  }
  int rc = 0;
  SAV_GUARD(h, rc = h->s->irfft_noalloc(in_hat_dev, out_dev, count));
  if (rc == SAV_E_ARG)
    h->err = "sav_irfft_noalloc: count half spectra exceed the " + std::to_string(SAV_NOALLOC_SCRATCH_BYTES >> 20) +
             " MiB scratch (count * (Nx/2+1) * Ny [* Nz] * complex size)";
  return rc;
}

int sav_fft_full(sav_handle_t* h, const double* in, double* out, int count, int inverse) {
  if (!h || !in || !out || count < 1) return SAV_E_ARG;
  if (h->ndim != 2) {
    h->err = "sav_fft_full: 2D only";
    return SAV_E_UNSUPPORTED;
  }
  SAV_GUARD(h, h->s->fft_full_host(in, out, count, inverse));
  return SAV_OK;
}

int sav_eval_functionals(sav_handle_t* h, const double* phi, const double* psi, double* H1, double* H2, double* H3,
                         double* G, double* W, double* area) {
  if (!h || !phi || !psi) return SAV_E_ARG;
  int rc = 0;
  SAV_GUARD(h, rc = h->s->eval_functionals(phi, psi, H1, H2, H3, G, W, area));
  if (rc == SAV_E_WC_NEGATIVE) h->err = "FATAL: W + C < 0";
  return rc;
}

int sav_psi_solve(sav_handle_t* h, const double* phi_star, double a, double dt, const double* rhs_hat,
                  double* out_hat, int32_t* iters, int32_t* solved, double* rnorm) {
  if (!h || !phi_star || !rhs_hat || !out_hat) return SAV_E_ARG;
  if (h->ndim != 2) {
    h->err = "sav_psi_solve: the Krylov psi solve exists in 2D only";
    return SAV_E_UNSUPPORTED;
  }
  int rc = 0;
  SAV_GUARD(h, rc = h->s->psi_solve_host(phi_star, a, dt, rhs_hat, out_hat, iters, solved, rnorm));
  return rc;
}

int sav_modified_energy(sav_handle_t* h, double* energy, double* area_ratio) {
  if (!h) return SAV_E_ARG;
  SAV_GUARD(h, h->s->modified_energy(energy, area_ratio));
  return SAV_OK;
}

int sav_kernel_timing(sav_handle_t* h, int kernel_id) {
  if (!h || kernel_id < 0 || kernel_id > 3) return SAV_E_ARG;
  SAV_GUARD(h, h->s->kernel_timing(kernel_id));
  return SAV_OK;
}

int sav_kernel_timing_get(sav_handle_t* h, int64_t* launches, double* total_ms, double* bytes_per_launch) {
  if (!h) return SAV_E_ARG;
  long long n = 0;
  SAV_GUARD(h, h->s->kernel_timing_get(&n, total_ms, bytes_per_launch));
  if (launches) *launches = n;
  return SAV_OK;
}

int sav_last_elapsed_ms(sav_handle_t* h, double* ms) {
  if (!h || !ms) return SAV_E_ARG;
  *ms = h->s->last_elapsed_ms();
  return SAV_OK;
}


int sav_set_column_tile(sav_handle_t* h, int columns) {
  if (!h) return SAV_E_ARG;
  SAV_GUARD(h, return h->s->set_column_tile(columns));
}

int sav_autotune_info(sav_handle_t* h, int32_t* columns, double* ms_8, double* ms_16) {
  if (!h) return SAV_E_ARG;
  int cw = 0;
  SAV_GUARD(h, h->s->column_tile(&cw, ms_8, ms_16, nullptr); if (columns) *columns = cw; return SAV_OK);
}

int sav_column_tile_info(sav_handle_t* h, int32_t* columns, double* ms_4, double* ms_8, double* ms_16) {
  if (!h) return SAV_E_ARG;
  int cw = 0;
  SAV_GUARD(h, h->s->column_tile(&cw, ms_8, ms_16, ms_4); if (columns) *columns = cw; return SAV_OK);
}

int sav_set_matvec_order(sav_handle_t* h, int order) {
  if (!h) return SAV_E_ARG;
  SAV_GUARD(h, return h->s->set_matvec_order(order));
}

int sav_matvec_order_info(sav_handle_t* h, int32_t* order, double* ms_batched, double* ms_sequenced) {
  if (!h) return SAV_E_ARG;
  int o = 0;
  SAV_GUARD(h, h->s->matvec_order(&o, ms_batched, ms_sequenced); if (order) *order = o; return SAV_OK);
}

int sav_set_transform_tile(sav_handle_t* h, int columns) {
  if (!h) return SAV_E_ARG;
  int rc = 0;
  SAV_GUARD(h, rc = h->s->set_transform_tile(columns));
  return rc;
}

int sav_transform_tile_info(sav_handle_t* h, int32_t* columns, double* ms_4, double* ms_8, double* ms_16) {
  if (!h) return SAV_E_ARG;
  int cw = 0;
  SAV_GUARD(h, h->s->transform_tile(&cw, ms_4, ms_8, ms_16));
  if (columns) *columns = cw;
  return SAV_OK;
}

int sav_transform_tile_fwd_info(sav_handle_t* h, int32_t* columns, double* ms_4, double* ms_8, double* ms_16) {
  if (!h) return SAV_E_ARG;
  int cw = 0;
  SAV_GUARD(h, h->s->transform_tile_fwd(&cw, ms_4, ms_8, ms_16));
  if (columns) *columns = cw;
  return SAV_OK;
}

int sav_set_plane_kernel(sav_handle_t* h, int kind) {
  if (!h) return SAV_E_ARG;
  int rc = 0;
  SAV_GUARD(h, rc = h->s->set_plane_kernel(kind));
  return rc;
}

int sav_plane_kernel_info(sav_handle_t* h, int32_t* kind, double* ms_register, double* ms_lds) {
  if (!h) return SAV_E_ARG;
  int k = 0;
  SAV_GUARD(h, h->s->plane_kernel(&k, ms_register, ms_lds));
  if (kind) *kind = k;
  return SAV_OK;
}

int sav_set_graph(sav_handle_t* h, int on) {
  if (!h) return SAV_E_ARG;
  SAV_GUARD(h, return h->s->set_graph(on));
}

int sav_graph_choice(sav_handle_t* h, int32_t* choice) {
  if (!h || !choice) return SAV_E_ARG;
  SAV_GUARD(h, *choice = h->s->graph_choice(); return SAV_OK);
}

int sav_graph_info(sav_handle_t* h, long long* replays, long long* repeats, long long* graphs) {
  if (!h) return SAV_E_ARG;
  SAV_GUARD(h, h->s->graph_info(replays, repeats, graphs); return SAV_OK);
}

int sav_set_precond(sav_handle_t* h, int precond) {
  if (!h) return SAV_E_ARG;
  SAV_GUARD(h, return h->s->set_precond(precond));
}

int sav_precond_init(sav_handle_t* h, uint64_t seed) {
  if (!h) return SAV_E_ARG;
  SAV_GUARD(h, return h->s->precond_init(seed));
}

int sav_precond_load(sav_handle_t* h, const float* params, int n) {
  if (!h || !params) return SAV_E_ARG;
  SAV_GUARD(h, return h->s->precond_params(true, const_cast<float*>(params), n));
}

int sav_precond_save(sav_handle_t* h, float* params, int n) {
  if (!h || !params) return SAV_E_ARG;
  SAV_GUARD(h, return h->s->precond_params(false, params, n));
}

int sav_precond_collect(sav_handle_t* h, int on, int patches_per_snapshot, uint64_t seed) {
  if (!h) return SAV_E_ARG;
  SAV_GUARD(h, return h->s->precond_collect(on, patches_per_snapshot, seed));
}

int sav_precond_samples(sav_handle_t* h, int64_t* n_windows, int64_t* fallbacks) {
  if (!h) return SAV_E_ARG;
  SAV_GUARD(h, if (n_windows) *n_windows = h->s->precond_samples();
            if (fallbacks) *fallbacks = h->s->precond_fallbacks(); return SAV_OK);
}

int sav_precond_train_target(sav_handle_t* h, int target) {
  if (!h) return SAV_E_ARG;
  SAV_GUARD(h, return h->s->precond_train_target(target));
}

int sav_precond_train(sav_handle_t* h, int iters, int batch, double lr, uint64_t seed, double* loss_first,
                      double* loss_last) {
  if (!h) return SAV_E_ARG;
  SAV_GUARD(h, return h->s->precond_train(iters, batch, lr, seed, loss_first, loss_last));
}

int sav_net_forward(sav_handle_t* h, const float* x, int nimg, int H, int W, float* y) {
  if (!h || !x || !y || nimg < 1 || H < 16 || W < 32 || H % 16 || W % 32) return SAV_E_ARG;
  SAV_GUARD(h, return h->s->net_forward(x, nimg, H, W, y));
}

int sav_net_grad(sav_handle_t* h, const float* x, const float* t, int nimg, int H, int W, int border, float* grads,
                 double* loss) {
  if (!h || !x || !t || !grads || nimg < 1 || H < 16 || W < 32 || H % 16 || W % 32 || border < 0 ||
      2 * border >= H || 2 * border >= W)
    return SAV_E_ARG;
  SAV_GUARD(h, return h->s->net_grad(x, t, nimg, H, W, border, grads, loss));
}

}  // extern "C"
