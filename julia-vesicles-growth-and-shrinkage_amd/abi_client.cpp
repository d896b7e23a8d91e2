// A plain C++ client of libsav_hip.so: the C ABI used the way a Julia `ccall` binding
// uses it (caller-owned host arrays in Julia's column-major layout, interleaved
// complex), with no Python or torch in the process.
//
// One vesicle on a 64 x 64 grid: SAV initialisation, a few time steps, checks that
// every step succeeds, the psi solves converge and the modified energy does not
// increase; then a two-member ensemble over the same steps must reproduce the single
// handle's energies exactly.  Exit status 0 on success.
#include <cmath>
#include <complex>
#include <cstdio>
#include <vector>

#include "../include/sav_abi.h"

namespace {

sav_config_t make_config(int n, const double* pin, const double* pout) {
  sav_config_t c{};
  c.ndim = 2;
  c.N = 1;
  c.Nx = c.Ny = n;
  c.Nz = 1;
  c.Lx = c.Ly = c.Lz = 2.0 * M_PI;
  c.epsilon = 0.2;
  c.M_phi = 1.0;
  c.M0_psi = 0.5;
  c.eta = 1.0;
  c.gamma_surf = 1.0;
  c.gamma_area = 1000;
  c.gamma_bend = 0.1;
  c.gamma_in = 1000;
  c.beta_in = 0;
  c.gamma_out = 1000;
  c.beta_out = 0;
  c.lamda = 1;
  c.psi_in_v = pin;
  c.psi_out_v = pout;
  // stabilisers S1 = 3 sqrt(2)/8 eps gamma_b, S2 = 0, S3 = gamma_b / eps^3, S4 = min(gamma_in, gamma_out)
  c.S1 = 3.0 * std::sqrt(2.0) / 8.0 * c.epsilon * c.gamma_bend;
  c.S2 = 0.0;
  c.S3 = c.gamma_bend / (c.epsilon * c.epsilon * c.epsilon);
  c.S4 = 1000.0;
  c.C1 = 1.0;
  c.C2 = 4e6;
  c.C3 = 8e6;
  c.dt = 1e-4;
  c.T = 1e-3;
  c.tol = 1e-10;
  c.precision = SAV_PREC_FP64;
  c.precond = SAV_PRECOND_DIAG;
  c.itmax = 20;
  c.atol = 1e-14;
  c.device = 0;
  return c;
}

struct State {  // a FieldState in Julia layout
  int n, nh;
  std::vector<double> phi, psi, mu, nu, u, p;
  std::vector<std::complex<double>> phi_hat, psi_hat, mu_hat, nu_hat, u_hat, p_hat;
  explicit State(int n_) : n(n_), nh(n_ / 2 + 1) {
    const size_t P = (size_t)n * n, H = (size_t)nh * n;
    phi.assign(P, 0.0);
    psi.assign(P, 0.0);
    mu.assign(P, 0.0);
    nu.assign(P, 0.0);
    u.assign(2 * P, 0.0);
    p.assign(P, 0.0);
    phi_hat.assign(H, 0.0);
    psi_hat.assign(H, 0.0);
    mu_hat.assign(H, 0.0);
    nu_hat.assign(H, 0.0);
    u_hat.assign(2 * H, 0.0);
    p_hat.assign(H, 0.0);
  }
};

double* d(std::vector<std::complex<double>>& v) { return reinterpret_cast<double*>(v.data()); }

int set_initial(sav_handle_t* h, State& s) {
  const int n = s.n;
  const double L = 2.0 * M_PI, eps = 0.2;
  for (int j = 0; j < n; ++j)      // y (second index)
    for (int i = 0; i < n; ++i) {  // x (fastest)
      const double x = i * L / n - L / 2, y = j * L / n - L / 2;
      const double r = std::sqrt(x * x / (1.2 * 1.2) + y * y / (0.8 * 0.8));
      const double ph = std::tanh((1.0 - r) * 0.9 / (std::sqrt(2.0) * eps));
      s.phi[(size_t)j * n + i] = ph;
      s.psi[(size_t)j * n + i] = 0.5 * (1 + ph) * 1.2 + 0.5 * (1 - ph) * 1.0;
    }
  int rc = sav_rfft(h, s.phi.data(), d(s.phi_hat), 1);
  if (rc == SAV_OK) rc = sav_rfft(h, s.psi.data(), d(s.psi_hat), 1);
  const double scal[4] = {1.0, 1.0, 1.0, 1.0};
  const double A0[1] = {0.0};
  if (rc == SAV_OK)
    rc = sav_set_state(h, SAV_PRESENT, s.phi.data(), d(s.phi_hat), s.psi.data(), d(s.psi_hat), s.u.data(),
                       d(s.u_hat), s.p.data(), d(s.p_hat), s.mu.data(), d(s.mu_hat), s.nu.data(), d(s.nu_hat),
                       scal, A0, 1e-4);
  if (rc == SAV_OK) rc = sav_init_sav(h);
  return rc;
}

}  // namespace

int main() {
  std::printf("sav_abi_version %d\n", sav_abi_version());
  const int n = 64, steps = 4;
  const double pin[1] = {1.2}, pout[1] = {1.0};
  sav_config_t cfg = make_config(n, pin, pout);
  sav_handle_t* h = nullptr;
  if (sav_create(&cfg, &h) != SAV_OK) {
    std::printf("sav_create failed: %s\n", sav_last_error(nullptr));
    return 1;
  }
  State s(n);
  if (set_initial(h, s) != SAV_OK) {
    std::printf("set-up failed: %s\n", sav_last_error(h));
    return 1;
  }
  // the time nodes of the ensemble run below, so both see the same dt = t[k] - t[k-1]
  std::vector<double> t(steps + 1);
  for (int k = 0; k <= steps; ++k) t[k] = k * 1e-4;
  std::vector<double> E(steps);
  for (int k = 1; k <= steps; ++k) {
    sav_step_stats_t st{};
    const int rc = sav_step(h, t[k] - t[k - 1], k, &st);
    std::printf("step %d rc %d energy %.12e area_ratio %.3e psi_iters %d %d %d\n", k, rc, st.energy, st.area_ratio,
                st.krylov_iters[0], st.krylov_iters[1], st.krylov_iters[2]);
    if (rc != SAV_OK || !st.krylov_solved[0] || !st.krylov_solved[1] || !st.krylov_solved[2]) return 2;
    if (k > 1 && st.energy > E[k - 2] * (1 + 1e-12)) return 3;
    E[k - 1] = st.energy;
  }
  sav_destroy(h);

  // the same run as a two-member ensemble (one thread per member)
  sav_config_t cfgs[2] = {cfg, cfg};
  sav_ensemble_t* e = nullptr;
  if (sav_ensemble_create(cfgs, 2, &e) != SAV_OK) {
    std::printf("sav_ensemble_create failed: %s\n", sav_ensemble_last_error(nullptr));
    return 4;
  }
  State s0(n), s1(n);
  if (set_initial(sav_ensemble_member(e, 0), s0) != SAV_OK || set_initial(sav_ensemble_member(e, 1), s1) != SAV_OK)
    return 5;
  std::vector<double> EE(2 * steps);
  std::vector<int32_t> kit(2), status(2);
  const int rc = sav_ensemble_run(e, t.data(), steps + 1, 1, 2, EE.data(), nullptr, kit.data(), status.data());
  sav_ensemble_destroy(e);
  if (rc != SAV_OK) return 6;
  for (int m = 0; m < 2; ++m)
    for (int k = 0; k < steps; ++k)
      if (EE[m * steps + k] != E[k]) {
        std::printf("ensemble member %d step %d energy %.17e != %.17e\n", m, k + 1, EE[m * steps + k], E[k]);
        return 7;
      }
  std::printf("abi client ok\n");
  return 0;
}
