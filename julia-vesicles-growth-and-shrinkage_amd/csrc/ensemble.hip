// Ensemble of independent SAV instances (no reference counterpart: the reference
// runs one simulation per process, scripts/Main.jl:86-114).
//
// Each member is an ordinary handle with its own HIP stream.  sav_ensemble_run
// drives the members' time loops from a pool of host threads: while one member's
// thread waits on a Krylov convergence check (a host synchronisation per batch of
// BiCGSTAB iterations), the kernels of the other members keep the GPU busy.  Small
// grids are latency-bound on their own, so overlapping instances is what fills the
// 256 CUs.  Members are handed out largest grid first so the long runs start early.
// With sav_ensemble_set_batched the members of each size class (same grid, vesicle
// count and solve settings) also share their psi solves: one batched BiCGSTAB per step
// over all their systems (sav::ISolveGroup), each member stepping in a thread of its own.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "krylov.h"
#include "solver.h"

struct sav_ensemble {
  std::vector<sav_handle_t*> members;
  std::vector<sav_config_t> cfgs;
  std::vector<long long> work;  // grid points x vesicles, for the hand-out order
  std::string err;
  // size-class groups (sav_ensemble_set_batched): group_of[i] = group index or -1, slot_of[i]
  // = the member's index in its group
  std::vector<std::unique_ptr<sav::ISolveGroup>> groups;
  std::vector<int> group_of, slot_of;
  bool batched = false;
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
    sav_set_graph(h, 0);           // eager unless set per member: the members' timings interleave
    e->members.push_back(h);
    e->cfgs.push_back(cfgs[i]);
    const long long pts = (long long)cfgs[i].Nx * cfgs[i].Ny * (cfgs[i].ndim == 3 ? cfgs[i].Nz : 1);
    e->work.push_back(pts * cfgs[i].N);
  }
  *out = e;
  return SAV_OK;
}

int sav_ensemble_destroy(sav_ensemble_t* e) {
  if (!e) return SAV_OK;
  for (auto* m : e->members) sav_destroy(m);  // the members first: they use the groups' pools
  e->groups.clear();
  delete e;
  return SAV_OK;
}

int sav_ensemble_set_batched(sav_ensemble_t* e, int on, int32_t* n_groups) {
  if (!e || (on != 0 && on != 1)) return SAV_E_ARG;
  try {
    const int n = (int)e->members.size();
    if (on && e->groups.empty()) {
      // size classes: members whose psi solves can share launches (same 2D grid, vesicle
      // count, storage precision, S4 and solve settings); at most 64 members per group
      e->group_of.assign(n, -1);
      e->slot_of.assign(n, -1);
      std::vector<bool> taken(n, false);
      for (int i = 0; i < n; ++i) {
        if (taken[i]) continue;
        const sav_config_t& a = e->cfgs[i];
        if (a.ndim != 2 || a.precision == SAV_PREC_MIXED || a.precond != SAV_PRECOND_DIAG) continue;
        std::vector<int> cls;
        for (int j = i; j < n; ++j) {
          const sav_config_t& b = e->cfgs[j];
          if (taken[j] || b.ndim != 2 || b.precision != a.precision || b.precond != SAV_PRECOND_DIAG) continue;
          if (b.Nx != a.Nx || b.Ny != a.Ny || b.N != a.N || b.S4 != a.S4 || b.tol != a.tol || b.atol != a.atol ||
              b.itmax != a.itmax || b.device != a.device)
            continue;
          cls.push_back(j);
        }
        for (size_t c0 = 0; c0 < cls.size(); c0 += sav::KRY_MAX_SYS / 3) {
          const size_t c1 = std::min(cls.size(), c0 + (size_t)sav::KRY_MAX_SYS / 3);
          if (c1 - c0 < 2) continue;
          std::vector<sav::ISolver*> ss;
          for (size_t k = c0; k < c1; ++k) ss.push_back(sav::handle_solver(e->members[cls[k]]));
          auto grp = sav::make_solve_group(ss);
          for (size_t k = c0; k < c1; ++k) {
            e->group_of[cls[k]] = (int)e->groups.size();
            e->slot_of[cls[k]] = (int)(k - c0);
            taken[cls[k]] = true;
          }
          e->groups.push_back(std::move(grp));
        }
      }
    }
    e->batched = on != 0;  // the groups' solves are coupled during sav_ensemble_run only
    if (n_groups) *n_groups = (int32_t)e->groups.size();
    return SAV_OK;
  } catch (const std::exception& ex) {
    e->err = ex.what();
    return SAV_E_ARG;
  }
}

int sav_ensemble_batched_info(sav_ensemble_t* e, int32_t* group_of, int64_t* solves) {
  if (!e) return SAV_E_ARG;
  const int n = (int)e->members.size();
  for (int i = 0; i < n; ++i) {
    const int gi = e->group_of.empty() ? -1 : e->group_of[i];
    if (group_of) group_of[i] = gi;
    if (solves) {
      long long ns = 0;
      if (gi >= 0) e->groups[gi]->info(&ns, nullptr, nullptr);
      solves[i] = ns;
    }
  }
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
  if (e->batched && !e->groups.empty()) {
    // one host thread per member: the members of a size-class group meet at every step's
    // psi solve, which the last to arrive runs for all of them; a member whose run ends
    // (or fails) leaves its group, so the others never wait for it
    for (auto& g : e->groups) {
      g->begin_run();
      g->set_coupled(true);
    }
    auto run_one = [&](int i) {
      (void)hipSetDevice(device);
      int32_t kit = 0;
      rcs[i] = sav_run(e->members[i], t_nodes, n_nodes, n0, energy_hist ? energy_hist + (long long)i * m : nullptr,
                       area_hist ? area_hist + (long long)i * m : nullptr, &kit);
      if (krylov_iters) krylov_iters[i] = kit;
      if (e->group_of[i] >= 0) e->groups[e->group_of[i]]->leave(e->slot_of[i]);
    };
    std::vector<std::thread> ts;
    for (int k = 1; k < n; ++k) ts.emplace_back(run_one, order[k]);
    run_one(order[0]);
    for (auto& th : ts) th.join();
    for (auto& g : e->groups) g->set_coupled(false);  // members may be stepped alone again
  } else {
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
  }
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
