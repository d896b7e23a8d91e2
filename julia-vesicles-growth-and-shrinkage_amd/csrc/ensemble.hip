// Ensemble of independent SAV instances (no reference counterpart: the reference
// runs one simulation per process, scripts/Main.jl:86-114).
//
// Each member is an ordinary handle with its own HIP stream.  sav_ensemble_run
// drives the members' time loops from a pool of host threads: while one member's
// thread waits on a Krylov convergence check (a host synchronisation per batch of
// BiCGSTAB iterations), the kernels of the other members keep the GPU busy.  Small
// grids are latency-bound on their own, so overlapping instances is what fills the
// 256 CUs.  Members are handed out largest grid first so the long runs start early.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "solver.h"

struct sav_ensemble {
  std::vector<sav_handle_t*> members;
  std::vector<long long> work;  // grid points x vesicles, for the hand-out order
  std::string err;
};

static thread_local std::string g_ens_error;

extern "C" {

int sav_ensemble_create(const sav_config_t* cfgs, int n, sav_ensemble_t** out) {
  if (!cfgs || n < 1 || !out) return SAV_E_ARG;
  *out = nullptr;
  auto* e = new sav_ensemble();
  for (int i = 0; i < n; ++i) {
    sav_handle_t* h = nullptr;
    const int rc = sav_create(&cfgs[i], &h);
    if (rc != SAV_OK) {
      g_ens_error = "member " + std::to_string(i) + ": " + sav_last_error(nullptr);
      for (auto* m : e->members) sav_destroy(m);
      delete e;
      return rc;
    }
    sav::set_step_fork(h, false);  // members overlap with each other: one stream each
    e->members.push_back(h);
    const long long pts = (long long)cfgs[i].Nx * cfgs[i].Ny * (cfgs[i].ndim == 3 ? cfgs[i].Nz : 1);
    e->work.push_back(pts * cfgs[i].N);
  }
  *out = e;
  return SAV_OK;
}

int sav_ensemble_destroy(sav_ensemble_t* e) {
  if (!e) return SAV_OK;
  for (auto* m : e->members) sav_destroy(m);
  delete e;
  return SAV_OK;
}

const char* sav_ensemble_last_error(const sav_ensemble_t* e) { return e ? e->err.c_str() : g_ens_error.c_str(); }

int sav_ensemble_size(const sav_ensemble_t* e) { return e ? (int)e->members.size() : 0; }

sav_handle_t* sav_ensemble_member(sav_ensemble_t* e, int i) {
  if (!e || i < 0 || i >= (int)e->members.size()) return nullptr;
  return e->members[i];
}

int sav_ensemble_run(sav_ensemble_t* e, const double* t_nodes, int n_nodes, int n0, int n_threads,
                     double* energy_hist, double* area_hist, int32_t* krylov_iters, int32_t* status) {
  if (!e || !t_nodes || n_nodes < 1) return SAV_E_ARG;
  const int n = (int)e->members.size();
  const int m = n_nodes - 1;
  if (n_threads < 1) n_threads = 1;
  if (n_threads > n) n_threads = n;
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return e->work[a] > e->work[b]; });
  std::vector<int> rcs(n, SAV_OK);
  std::atomic<int> next{0};
  int device = 0;
  (void)hipGetDevice(&device);
  auto worker = [&]() {
    (void)hipSetDevice(device);  // the current device is per host thread
    for (int k = next.fetch_add(1); k < n; k = next.fetch_add(1)) {
      const int i = order[k];
      int32_t kit = 0;
      rcs[i] = sav_run(e->members[i], t_nodes, n_nodes, n0, energy_hist ? energy_hist + (long long)i * m : nullptr,
                       area_hist ? area_hist + (long long)i * m : nullptr, &kit);
      if (krylov_iters) krylov_iters[i] = kit;
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < n_threads; ++t) pool.emplace_back(worker);
  worker();
  for (auto& th : pool) th.join();
  int worst = SAV_OK;
  e->err.clear();
  for (int i = 0; i < n; ++i) {
    if (status) status[i] = rcs[i];
    if (rcs[i] < 0) {
      if (worst >= 0 || rcs[i] < worst) worst = rcs[i];
      if (e->err.empty())
        e->err = "member " + std::to_string(i) + ": " + sav_last_error(e->members[i]);
    } else if (rcs[i] > 0 && worst == SAV_OK) {
      worst = rcs[i];
    }
  }
  return worst;
}

}  // extern "C"
