// Device-resident seven-step SAV timestep (src/Solvers.jl, src/Solvers_3d.jl),
// functionals (src/Utils.jl), SAV initialisation and monitor (src/simulation.jl).
//
// Layout in HBM: every field is x-fastest, one scalar field = Nx*Ny*Nz values
// ("P"), one half spectrum = (Nx/2+1)*Ny*Nz complex values ("H").  A FieldState
// is two contiguous blocks:
//   physical: [phi (N) | psi (N) | mu (N) | nu (N) | u (ndim) | p (1)]
//   spectral: [phi_hat | psi_hat | mu_hat | nu_hat | u_hat | p_hat]
// Three such buffers rotate (present, old, next) so update_state! is a pointer
// swap and the back-transform of a new state is a single batched irfft.
#include <algorithm>
#include <cmath>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "krylov.h"
#include "launch.h"
#include "learned.h"
#include "solver.h"

namespace sav {

__global__ __launch_bounds__(256) void k_red_fin(const double* __restrict__ partials, int nb, int K,
                                                 double* __restrict__ out) {
  __shared__ double sc[4];
  const int k = blockIdx.x, g = blockIdx.y;
  double s = 0.0;
  {
    // partials of group g: [b * K + k], b < nb; summed in the order b = t, t + 256, ...
    const double* pg = partials + (long long)g * nb * K + k;
    constexpr int U = 8;
    for (int b0 = threadIdx.x; b0 < nb; b0 += U * 256) {
      double v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = b0 + u * 256 < nb ? pg[(long long)(b0 + u * 256) * K] : 0.0;
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (b0 + u * 256 < nb) s += v[u];
    }
  }
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) sc[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[g * K + k] = sc[0] + sc[1] + sc[2] + sc[3];
}

namespace {

constexpr double SQ2 = 1.41421356237309504880168872420969808;

// Twiddle table exp(-2 pi i n / L), n in [0, L), accurate to long double.
template <typename R>
std::vector<cplx<R>> twiddles(int L) {
  std::vector<cplx<R>> w(L);
  for (int n = 0; n < L; ++n) {
    long double ang = -2.0L * 3.141592653589793238462643383279502884L * (long double)n / (long double)L;
    w[n] = cplx<R>{(R)cosl(ang), (R)sinl(ang)};
  }
  return w;
}

std::vector<double> rfftfreq_k(int n, double L) {  // 2 pi * rfftfreq(n, n/L)
  std::vector<double> k(n / 2 + 1);
  for (int i = 0; i <= n / 2; ++i) k[i] = 2.0 * M_PI * ((double)i * ((double)n / L) / (double)n);
  return k;
}
std::vector<double> fftfreq_k(int n, double L) {  // 2 pi * fftfreq(n, n/L)
  std::vector<double> k(n);
  const double fs = (double)n / L;
  for (int i = 0; i < n; ++i) {
    int j = (i < (n + 1) / 2) ? i : i - n;
    k[i] = 2.0 * M_PI * ((double)j * fs / (double)n);
  }
  return k;
}

struct KIdx {
  int Nxh, Ny, Nz;
  __device__ __forceinline__ void operator()(long long i, int& ix, int& iy, int& iz, int& f) const {
    ix = (int)(i % Nxh);
    long long r = i / Nxh;
    iy = (int)(r % Ny);
    r /= Ny;
    iz = (int)(r % Nz);
    f = (int)(r / Nz);
  }
};

}  // namespace

enum { F_PHI = 0, F_PSI = 1, F_MU = 2, F_NU = 3 };

// Scratch of the allocation-free entry point (sav_irfft_noalloc): a file-scope device array,
// one copy per GPU, held until the process ends.  Host code never passes the array's name
// to a kernel (on the host it is only a stand-in symbol): it takes the per-device address
// from hipGetSymbolAddress, looked up once per device when a handle is created there.
static __device__ __align__(256) unsigned char noalloc_scratch[SAV_NOALLOC_SCRATCH_BYTES];
static std::mutex g_noalloc_mu;
static void* g_noalloc_ptr[64] = {};

// Device address of noalloc_scratch on `device` (which must be current).
static void* noalloc_scratch_address(int device) {
  if (device < 0 || device >= 64) throw HipError("device index out of range for the scratch table", SAV_E_ARG);
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[device]) {
    void* p = nullptr;
    SAV_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch)));
    if (!p) throw HipError("hipGetSymbolAddress returned a null scratch address", SAV_E_HIP);
    g_noalloc_ptr[device] = p;
  }
  return g_noalloc_ptr[device];
}

template <typename R>
class SolveGroup;

template <typename R>
class Solver final : public ISolver {
 public:
  using C = cplx<R>;
  // pointwise and spectral arithmetic runs in FP64 registers; storage is R
  using T = double;
  using CD = cplx<double>;
  HostCfg hc;
  sav_config_t c;
  int nd, N, Nx, Ny, Nz, Nxh;
  long long P, H;
  double dv;  // dx*dy(*dz) in the reference's multiplication order is applied explicitly
  hipStream_t s = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // Step 1's advection terms and step 4 read only the present/old state and the
  // extrapolations of step 1's first kernels, so an eager step runs them on a second
  // stream beside the functionals and the psi solves (they fill the solves'
  // single-workgroup finishes and launch gaps); their transforms use their own scratch
  // (tmp4, w_grad).  The psi solves wait for the advection terms, step 5 for step 4.
  hipStream_t s4 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_gphi = nullptr, ev_adv = nullptr, ev_join = nullptr;
  bool fork_step4 = false, fork_enabled = true;
  void set_step_fork(bool on) override { fork_enabled = on; }
  // The second stream and its scratch are made at the first step that forks: HIP maps
  // streams to hardware queues round-robin as they are created, so a stream made with
  // every handle would leave the handles' main streams on half of the queues (ensemble
  // members, which do not fork, would then serialise on them).
  void ensure_fork() {
    if (s4) return;
    SAV_HIP(hipStreamCreateWithFlags(&s4, hipStreamNonBlocking));
    SAV_HIP(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
    SAV_HIP(hipEventCreateWithFlags(&ev_gphi, hipEventDisableTiming));
    SAV_HIP(hipEventCreateWithFlags(&ev_adv, hipEventDisableTiming));
    SAV_HIP(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    // step 4's transforms: nd^2 velocity derivatives + 2 nd N derivatives of mu, nu; the
    // advection terms: nd N gradients of psi*
    tmp4 = dalloc<C>((long long)(nd * nd + 2 * nd * N) * H);
    w_grad = dalloc<R>((long long)nd * N * P);
    SAV_HIP(hipStreamSynchronize(s));
  }
  FftGeom<R> g{};
  std::vector<void*> allocs;
  // tables
  double *d_kx = nullptr, *d_ky = nullptr, *d_kz = nullptr, *d_kx0 = nullptr;
  // states
  R* sphys[3];
  C* shalf[3];
  int ip = 0, io = 1, inew = 2;
  DStateScal* dstate = nullptr;  // [3]
  double* dA0 = nullptr;
  double *d_psi_in = nullptr, *d_psi_out = nullptr;
  DScal* ds = nullptr;
  DScal* hs = nullptr;  // pinned
  double* partials = nullptr;
  double* red = nullptr;  // reduction outputs (512 doubles)
  // work
  R *w_phis, *w_psis, *w_M, *w_us, *w_d4, *w_gphi, *w_om, *w_H, *w_w, *w_adv, *w_du, *w_dmn,
      *w_rhs2;
  C *w_phat, *w_omh, *w_Hh, *cache, *w_wh, *w_rhs, *w_advh, *w_rhs2h, *w_uth, *tmp;
  // scratch of the forked branch (made by ensure_fork)
  R* w_grad = nullptr;
  C* tmp4 = nullptr;
  // Krylov
  KrylovCtx<R> kc{};
  void* scratch = nullptr;  // this device's noalloc_scratch (sav_irfft_noalloc)
  // SAV_PREC_MIXED (2D; FP64 storage): every stage in FP64 except the psi BiCGSTAB's inner
  // iterations, which run in FP32 on the residual equation of an FP64 defect correction
  // (bicgstab_refined); the FP64 context kc holds the iterate, right-hand side and residual.
  bool mixed = false;
  KrylovCtx<float> kc32{};
  KScal* hks32 = nullptr;
  float* d_M32 = nullptr;  // M_psi of the inner iterations
  static constexpr int MIXED_ROUNDS = 8;
  // inner relative tolerance: at 1024^2 (two vesicles, 100 steps) 1e-5 gave 175.8 SAV steps/s,
  // 1e-4 159.4, 3e-4 152.7, 1e-6 163.7 (fewer rounds against more inner iterations)
  static constexpr double MIXED_INNER_RTOL = 1e-5;
  KScal* hks = nullptr;       // this handle's Krylov scalars on the host (a group's slice once grouped)
  KScal* hks_own = nullptr;   // its own pinned allocation
  DStateScal* hst = nullptr;  // pinned: scalars of the state a step produced
  // size-class group (ISolveGroup): the psi solves of a step run as one batched solve with
  // the other members' while `group` is set; the pools stay in use after decoupling
  SolveGroup<R>* group = nullptr;
  int group_m = -1;
  C* w_bphys_c = nullptr;
  float last_ms = 0.f;

  template <class T>
  T* dalloc(size_t n) {
    void* p = nullptr;
    size_t bytes = n * sizeof(T) + 256;
    SAV_HIP(hipMalloc(&p, bytes));
    SAV_HIP(hipMemsetAsync(p, 0, bytes, s));
    allocs.push_back(p);
    return reinterpret_cast<T*>(p);
  }
  void dfree(void* p) {
    auto it = std::find(allocs.begin(), allocs.end(), p);
    if (it == allocs.end()) return;
    allocs.erase(it);
    SAV_HIP(hipFree(p));
  }

  // Moves this handle's psi-solve buffers into the slices of a size-class group's pools:
  // Krylov vectors (vec: x, r, p, v, s, t, b, Bx, By of its three systems), mobility, step
  // scalars (copied over) and Krylov scalars; its own Krylov vectors are freed.
  void adopt_group(SolveGroup<R>* grp, int m, C* const* vec, R* Mp, DScal* dsp, KScal* ksp, KScal* hksp) {
    SAV_HIP(hipStreamSynchronize(s));
    drop_graphs();  // captured steps hold the previous addresses
    SAV_HIP(hipMemcpyAsync(dsp, ds, sizeof(DScal), hipMemcpyDeviceToDevice, s));
    SAV_HIP(hipStreamSynchronize(s));
    ds = dsp;
    kc.ds = dsp;
    kc32.ds = dsp;
    w_M = Mp;
    kc.M = Mp;
    C** own[9] = {&kc.x, &kc.r, &kc.p, &kc.v, &kc.s, &kc.t, &kc.b, &kc.Bx, &kc.By};
    for (int i = 0; i < 9; ++i) {
      dfree(*own[i]);
      *own[i] = vec[i];
    }
    kc.ks = ksp;
    kc.hks = hksp;
    hks = hksp;
    group = grp;
    group_m = m;
  }
  void group_solve();
  bool group_ready() const {
    return nd == 2 && !mixed && !use_learned && !collecting && precond_kind == SAV_PRECOND_DIAG;
  }

  explicit Solver(const HostCfg& h) : hc(h), c(h.c) {
    nd = c.ndim;
    N = c.N;
    Nx = c.Nx;
    Ny = c.Ny;
    Nz = nd == 3 ? c.Nz : 1;
    Nxh = Nx / 2 + 1;
    P = (long long)Nx * Ny * Nz;
    H = (long long)Nxh * Ny * Nz;
    c.psi_in_v = hc.psi_in.data();
    c.psi_out_v = hc.psi_out.data();
    SAV_HIP(hipSetDevice(c.device));
    scratch = noalloc_scratch_address(c.device);
    SAV_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    SAV_HIP(hipEventCreate(&ev0));
    SAV_HIP(hipEventCreate(&ev1));
    // wavenumbers (SpectralUtils.jl:42-44; SpectralUtils_3d.jl)
    auto kx = rfftfreq_k(Nx, c.Lx), ky = fftfreq_k(Ny, c.Ly), kx0 = fftfreq_k(Nx, c.Lx);
    std::vector<double> kz = nd == 3 ? fftfreq_k(Nz, c.Lz) : std::vector<double>{0.0};
    d_kx = upload(kx);
    d_ky = upload(ky);
    d_kz = upload(kz);
    d_kx0 = upload(kx0);
    g.ndim = nd;
    g.Nx = Nx;
    g.Ny = Ny;
    g.Nz = Nz;
    g.Nxh = Nxh;
    // FFT pass tables in the layout TeamFFT<R> reads; tw_x2 is the plain length-Nx
    // table of the r2c/c2r split (X[k] = E[k] + W^k O[k]).
    g.tw_x = upload(default_twiddles<R>(Nx));
    g.tw_x2 = upload(twiddles<R>(Nx));
    g.tw_xh = upload(default_twiddles<R>(Nx / 2));
    g.tw_y = upload(column_twiddles<R>(Ny));
    g.tw_z = nd == 3 ? upload(column_twiddles<R>(Nz)) : g.tw_y;
    if (nd == 3 && Nx == 128 && Ny == 128) g.tw_hp = upload(hplane_twiddles<R>());
    g.k = KTables{d_kx, d_ky, d_kz, d_kx0};
    g.dkx = 2.0 * M_PI / c.Lx;
    g.dky = 2.0 * M_PI / c.Ly;
    // states
    const long long nphys = (4LL * N + nd + 1) * P, nhalf = (4LL * N + nd + 1) * H;
    for (int b = 0; b < 3; ++b) {
      sphys[b] = dalloc<R>(nphys);
      shalf[b] = dalloc<C>(nhalf);
    }
    dstate = dalloc<DStateScal>(3);
    dA0 = dalloc<double>(MAXN);
    d_psi_in = upload(hc.psi_in);
    d_psi_out = upload(hc.psi_out);
    ds = dalloc<DScal>(1);
    SAV_HIP(hipHostMalloc(reinterpret_cast<void**>(&hs), sizeof(DScal), hipHostMallocDefault));
    partials = dalloc<double>(RED_PARTIALS);
    red = dalloc<double>(1024);
    // work buffers
    const long long NP = (long long)N * P, NH = (long long)N * H;
    w_phis = dalloc<R>(NP);
    w_psis = dalloc<R>(NP);
    w_M = dalloc<R>(NP);
    w_us = dalloc<R>(nd * P);
    w_d4 = dalloc<R>(6 * NP);
    w_gphi = dalloc<R>((long long)nd * NP);
    w_om = dalloc<R>(NP);
    w_H = dalloc<R>(4 * NP);
    w_w = dalloc<R>(3 * NP);
    w_adv = dalloc<R>(2 * NP);
    w_du = dalloc<R>(9 * P);
    w_dmn = dalloc<R>(6 * NP);
    w_rhs2 = dalloc<R>(3 * P);
    w_phat = dalloc<C>(NH);
    w_omh = dalloc<C>(NH);
    w_Hh = dalloc<C>(4 * NH);
    cache = dalloc<C>(16 * NH);
    w_wh = dalloc<C>(3 * NH);
    w_rhs = dalloc<C>(3 * NH);
    w_advh = dalloc<C>(2 * NH);
    w_rhs2h = dalloc<C>(3 * H);
    w_uth = dalloc<C>(6 * H);
    tmp = dalloc<C>((16LL * N + 16) * H);
    // Krylov workspace (2D)
    SAV_HIP(hipHostMalloc(reinterpret_cast<void**>(&hks), 4 * sizeof(KScal), hipHostMallocDefault));
    hks_own = hks;
    SAV_HIP(hipHostMalloc(reinterpret_cast<void**>(&hst), sizeof(DStateScal), hipHostMallocDefault));
    std::memset(hks, 0, 4 * sizeof(KScal));
    if (nd == 2) {
      const long long nf = (long long)N * Nx * Ny;
      kc.Nx = Nx;
      kc.Ny = Ny;
      kc.N = N;
      kc.nfull = nf;
      kc.kx0 = d_kx0;
      kc.ky = d_ky;
      kc.tw_x = g.tw_x;
      kc.tw_ri = upload(kmajor_twiddles<R>(Nx, 64, row_plan_maxr(Nx), false));
      kc.tw_rf = upload(kmajor_twiddles<R>(Nx, 64, row_plan_maxr(Nx), true));
      kc.tw_ci = upload(kmajor_twiddles<R>(Ny, col_plan_ttmax<R>(Ny, false), col_plan_maxr<R>(Ny, false), false));
      kc.tw_cf = upload(kmajor_twiddles<R>(Ny, col_plan_ttmax<R>(Ny, true), col_plan_maxr<R>(Ny, true), false));
      kc.tw_y = g.tw_y;
      kc.M = w_M;
      kc.ds = ds;
      kc.S4 = c.S4;
      kc.atol = c.atol;
      kc.rtol = c.tol;
      kc.itmax = c.itmax;
      // three systems (the batched step-1 solves), nf apart
      kc.vs = nf;
      // partials per system: 3 x (one per column of the column-forward kernel, or one per
      // block of a vector kernel, at most N Nx)
      kc.ps = 3 * std::max<long long>((long long)Nx * N, 512);
      kc.x = dalloc<C>(3 * nf);
      kc.r = dalloc<C>(3 * nf);
      kc.p = dalloc<C>(3 * nf);
      kc.v = dalloc<C>(3 * nf);
      kc.s = dalloc<C>(3 * nf);
      kc.t = dalloc<C>(3 * nf);
      kc.b = dalloc<C>(3 * nf);
      kc.Bx = dalloc<C>(3 * nf);
      kc.By = dalloc<C>(3 * nf);
      kc.partials = partials;
      kc.ks = dalloc<KScal>(4);
      kc.hks = hks;
      kc.st = s;
      kc.guess[0] = kc.guess[1] = kc.guess[2] = 8;
      if constexpr (std::is_same<R, double>::value) {
        mixed = c.precision == SAV_PREC_MIXED;
        if (mixed) {
          kc32.Nx = Nx;
          kc32.Ny = Ny;
          kc32.N = N;
          kc32.nfull = nf;
          kc32.kx0 = d_kx0;
          kc32.ky = d_ky;
          kc32.tw_ri = upload(kmajor_twiddles<float>(Nx, 64, row_plan_maxr(Nx), false));
          kc32.tw_rf = upload(kmajor_twiddles<float>(Nx, 64, row_plan_maxr(Nx), true));
          kc32.tw_ci = upload(
              kmajor_twiddles<float>(Ny, col_plan_ttmax<float>(Ny, false), col_plan_maxr<float>(Ny, false), false));
          kc32.tw_cf = upload(
              kmajor_twiddles<float>(Ny, col_plan_ttmax<float>(Ny, true), col_plan_maxr<float>(Ny, true), false));
          d_M32 = dalloc<float>(nf);
          x_prev = dalloc<C>(3 * nf);
          x_prev2 = dalloc<C>(3 * nf);
          kc32.M = d_M32;
          kc32.ds = ds;
          kc32.S4 = c.S4;
          kc32.atol = 0.0;
          kc32.rtol = MIXED_INNER_RTOL;
          kc32.itmax = c.itmax;
          kc32.vs = nf;
          kc32.ps = kc.ps;
          kc32.x = dalloc<cplx<float>>(3 * nf);
          kc32.r = dalloc<cplx<float>>(3 * nf);
          kc32.p = dalloc<cplx<float>>(3 * nf);
          kc32.v = dalloc<cplx<float>>(3 * nf);
          kc32.s = dalloc<cplx<float>>(3 * nf);
          kc32.t = dalloc<cplx<float>>(3 * nf);
          kc32.b = dalloc<cplx<float>>(3 * nf);
          kc32.Bx = dalloc<cplx<float>>(3 * nf);
          kc32.By = dalloc<cplx<float>>(3 * nf);
          kc32.partials = partials;
          kc32.ks = dalloc<KScal>(4);
          SAV_HIP(hipHostMalloc(reinterpret_cast<void**>(&hks32), 4 * sizeof(KScal), hipHostMallocDefault));
          std::memset(hks32, 0, 4 * sizeof(KScal));
          kc32.hks = hks32;
          kc32.st = s;
          kc32.guess[0] = kc32.guess[1] = kc32.guess[2] = 4;
        }
      }
    }
    SAV_HIP(hipStreamSynchronize(s));
  }

  ~Solver() override {
    if (s) hipStreamSynchronize(s);
    for (auto e : tev) hipEventDestroy(e);
    for (void* p : allocs) hipFree(p);
    if (hs) hipHostFree(hs);
    if (hks_own) hipHostFree(hks_own);
    if (hks32) hipHostFree(hks32);
    if (hst) hipHostFree(hst);
    for (auto& gr : graphs) hipGraphExecDestroy(gr.exec);
    if (ev0) hipEventDestroy(ev0);
    if (ev1) hipEventDestroy(ev1);
    if (ev_fork) hipEventDestroy(ev_fork);
    if (ev_gphi) hipEventDestroy(ev_gphi);
    if (ev_adv) hipEventDestroy(ev_adv);
    if (ev_join) hipEventDestroy(ev_join);
    if (s4) {
      hipStreamSynchronize(s4);
      hipStreamDestroy(s4);
    }
    if (s) hipStreamDestroy(s);
  }

  template <class T>
  T* upload(const std::vector<T>& v) {
    T* p = dalloc<T>(v.size());
    SAV_HIP(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    SAV_HIP(hipStreamSynchronize(s));
    return p;
  }

  // ---- field accessors ---------------------------------------------------
  R* fphys(int b, int f) { return sphys[b] + (long long)f * N * P; }
  C* fhalf(int b, int f) { return shalf[b] + (long long)f * N * H; }
  R* uphys(int b) { return sphys[b] + 4LL * N * P; }
  C* uhalf(int b) { return shalf[b] + 4LL * N * H; }
  R* pphys(int b) { return sphys[b] + (4LL * N + nd) * P; }
  C* phalf(int b) { return shalf[b] + (4LL * N + nd) * H; }
  C* cf(int k) { return cache + (long long)k * N * H; }  // Step1Cache field k
  enum {
    PHI11 = 0, MU11, PSI11, NU11, PHI12, MU12, PSI12, NU12, PHI13, MU13, PHI14, MU14, PHI21, MU21, PSI21, NU21
  };

  // ---- transform helpers ---------------------------------------------------
  static SpecJob job(const void* src, void* dst, int count, int mul = MUL_NONE, const void* src2 = nullptr,
                     double scale = 1.0) {
    SpecJob j{};
    j.src = src;
    j.src2 = src2;
    j.dst = dst;
    j.mul = mul;
    j.count = count;
    j.scale = scale;
    return j;
  }
  void rfft(std::initializer_list<SpecJob> l) {
    SpecJobs js{};
    for (auto& j : l) js.j[js.n++] = j;
    rfft_jobs<R>(g, js, s);
  }
  void irfft(std::initializer_list<SpecJob> l) {
    SpecJobs js{};
    for (auto& j : l) {
      if (js.n == MAXJ) {
        irfft_jobs<R>(g, js, tmp, s);
        js.n = 0;
      }
      js.j[js.n++] = j;
    }
    irfft_jobs<R>(g, js, tmp, s);
  }
  void irfft_v(const std::vector<SpecJob>& l) {
    SpecJobs js{};
    for (auto& j : l) {
      if (js.n == MAXJ) {
        irfft_jobs<R>(g, js, tmp, s);
        js.n = 0;
      }
      js.j[js.n++] = j;
    }
    irfft_jobs<R>(g, js, tmp, s);
  }
  int DX(int d) const { return MUL_DX + d; }

  // ---- device-side constants ------------------------------------------------
  struct PhysC {
    double eps, S1, S2, S3, S4, gs, gb, gin, gout, bin, bout, M0, C1, C2, C3, garea, dx, dy, dz;
  };
  PhysC pc() const {
    return PhysC{c.epsilon, c.S1, c.S2, c.S3, c.S4, c.gamma_surf, c.gamma_bend, (double)c.gamma_in,
                 (double)c.gamma_out, (double)c.beta_in, (double)c.beta_out, c.M0_psi, c.C1, c.C2, c.C3,
                 (double)c.gamma_area, c.Lx / c.Nx, c.Ly / c.Ny, nd == 3 ? c.Lz / c.Nz : 1.0};
  }
  // integrate(...) = sum * dx * dy [* dz]
  static __host__ __device__ double integ(double sum, const PhysC& q, int nd) {
    double v = sum * q.dx * q.dy;
    if (nd == 3) v = v * q.dz;
    return v;
  }

  // out = the part of X = 2 a - b (X = a when b is null) that a c2r keeps: X itself off the
  // kx = 0 and kx = Nx/2 planes, (X(k) + conj X(-k)) / 2 on them; `count` fields of H bins.
  // rfft(irfft(X)) equals it up to the transforms' round-off.
  void hermitian_part(const C* a, const C* b, C* out, int count) {
    const int Nx_ = Nx, Nxh_ = Nxh, Ny_ = Ny, Nz_ = Nz;
    const long long Hl = H;
    const KIdx ki{Nxh, Ny, Nz};
    foreach_n(s, (long long)count * H, [=] __device__(long long i) {
      int ix, iy, iz, f;
      ki(i, ix, iy, iz, f);
      auto ld = [&](long long j) {
        const CD x = a[j];
        if (!b) return x;
        const CD y = b[j];
        return CD{T(2) * x.x - y.x, T(2) * x.y - y.y};
      };
      CD v = ld(i);
      if (ix == 0 || 2 * ix == Nx_) {
        const int iy2 = iy ? Ny_ - iy : 0, iz2 = iz ? Nz_ - iz : 0;
        const CD m = ld((long long)f * Hl + ((long long)iz2 * Ny_ + iy2) * Nxh_ + ix);
        v = CD{T(0.5) * (v.x + m.x), T(0.5) * (v.y - m.y)};
      }
      out[i] = v;
    });
  }

  // =========================================================================
  // Functionals at (phi, psi): fills w_Hh = rfft of [H1 | G | H2 | H3] (and, with
  // phys_H, w_H with the physical fields), the scalars W1..W3, R1s..R3s, area[n]
  // (Utils.jl:36-249).  When set_A0 is true (SAV initialisation) A0 := area first, as
  // simulation.jl:25-26 does.
  // =========================================================================
  // grad phi (nd components -> w_gphi) and Lap phi (-> w_d4 slot 2) from rfft(phi)
  // (Utils.jl:37, 149, 167, 209; the functionals use the x and y components, step 1's
  // advection term all of them).  In a
    // step (`spec` given) phi = 2 phi^n - phi^{n-1} is the extrapolation of two states whose
    // physical fields are the c2r of their spectra, so rfft(phi) is the part of
    // 2 phi_hat^n - phi_hat^{n-1} that a c2r keeps (its Hermitian part on the kx = 0 and
    // kx = Nx/2 planes), formed here in one pass instead of a transform.  In FP32 this also
    // keeps the FFT roundoff floor of rfft(phi) (~1e-7 of the largest coefficient at every k)
    // out of the S1 k^4 biharmonic term.
  void functionals_grad(const R* phi, const C* spec, const C* spec_old) {
    const long long NP = (long long)N * P;
    if (spec) {
      hermitian_part(spec, spec_old, w_phat, N);
    } else {
      rfft({job(phi, w_phat, N)});
    }
    std::vector<SpecJob> jl;
    for (int d = 0; d < nd; ++d) jl.push_back(job(w_phat, w_gphi + d * NP, N, DX(d)));
    jl.push_back(job(w_phat, w_d4 + 2 * NP, N, MUL_LAP));
    irfft_v(jl);
  }

  // grads_ready: functionals_grad has run for this phi
  void functionals(const R* phi, const R* psi, bool set_A0, const C* spec = nullptr, const C* spec_old = nullptr,
                   bool phys_H = false, bool grads_ready = false) {
    const long long NP = (long long)N * P, NH = (long long)N * H;
    const PhysC q = pc();
    const int ndl = nd;
    if (!grads_ready) functionals_grad(phi, spec, spec_old);
    const R* gx = w_gphi;
    const R* gy = w_gphi + NP;
    const R* lap = w_d4 + 2 * NP;
    // per-vesicle sums: f_surf, f_bend, E2 integrand, f_osm, psi^2 ; omega -> w_om
    R* om = w_om;
    const double* pin = d_psi_in;
    const double* pout = d_psi_out;
    const long long Pl = P;
    reduce_n<5>(s, P, N,
                [=] __device__(long long i, int n, double(&acc)[5]) {
                  const long long k = (long long)n * Pl + i;
                  const T ph = phi[k], ps = psi[k], gxx = gx[k], gyy = gy[k], lp = lap[k];
                  const T eps = (T)q.eps;
                  // f_surf (Utils.jl:41)
                  const T gg = T(0.25) * (ph * ph - T(1)) * (ph * ph - T(1));
                  const T c1 = (T)(3.0 * SQ2 / 4.0);
                  const T fs = c1 * (gg / eps) + eps * (gxx * gxx + gyy * gyy) / T(2);
                  // f_bend (Utils.jl:49)
                  const T gp = ph * (ph * ph - T(1));
                  const T c2 = (T)(3.0 * SQ2 / (16.0 * q.eps));
                  const T tb = gp / eps - eps * lp;
                  const T fb = c2 * (tb * tb);
                  // E2 integrand (Utils.jl:175-177)
                  const T e2 = T(0.5) * (T)q.S1 * (lp * lp) + T(0.5) * (T)q.S2 * (gxx * gxx + gyy * gyy) +
                               T(0.5) * (T)q.S3 * (ph * ph);
                  // f_osmotic (Utils.jl:53-60)
                  const T P_ = T(-0.5) * (ph * ph * ph) + T(1.5) * ph;
                  const T di = ps - (T)pin[n], dou = ps - (T)pout[n];
                  const T fin = T(0.5) * (T)q.gin * (di * di) + (T)q.bin;
                  const T fou = T(0.5) * (T)q.gout * (dou * dou) + (T)q.bout;
                  const T fo = ((T(1) + P_) / T(2)) * fin + ((T(1) - P_) / T(2)) * fou;
                  // omega (Utils.jl:152)
                  om[k] = (ph * ph * ph - ph) / eps - eps * lp;
                  acc[0] += (double)fs;
                  acc[1] += (double)fb;
                  acc[2] += (double)e2;
                  acc[3] += (double)fo;
                  acc[4] += (double)ps * (double)ps;
                },
                partials, red);
    // scalars (Utils.jl:159-190) -- red[n*5 + k]
    DScal* dsl = ds;
    double* A0 = dA0;
    const double* rd = red;
    const int Nl = N;
    single(s, [=] __device__() {
      double A[MAXN];
      double s_fs = 0, s_fb = 0, s_e2 = 0, s_fo = 0, s_p2 = 0;
      for (int n = 0; n < Nl; ++n) {
        A[n] = integ(rd[n * 5 + 0], q, ndl);
        s_fs += rd[n * 5 + 0];
        s_fb += rd[n * 5 + 1];
        s_e2 += rd[n * 5 + 2];
        s_fo += rd[n * 5 + 3];
        s_p2 += rd[n * 5 + 4];
      }
      if (set_A0)
        for (int n = 0; n < Nl; ++n) A0[n] = A[n];
      double W1 = 0.0;
      for (int n = 0; n < Nl; ++n) {
        const double d = A[n] - A0[n];
        W1 += 0.5 * q.garea * (d * d);
        dsl->area[n] = A[n];
      }
      const double Fs = q.gs * integ(s_fs, q, ndl);
      const double Fb = q.gb * integ(s_fb, q, ndl);
      const double E2 = integ(s_e2, q, ndl);
      const double W2 = Fs + Fb - E2;
      const double Wo = integ(s_fo, q, ndl);
      const double E3 = 0.5 * q.S4 * integ(s_p2, q, ndl);
      const double W3 = Wo - E3;
      dsl->W1 = W1;
      dsl->W2 = W2;
      dsl->W3 = W3;
      int fl = dsl->flags;
      if (W1 + q.C1 < 0) fl |= 1;
      if (W2 + q.C2 < 0) fl |= 2;
      if (W3 + q.C3 < 0) fl |= 4;
      dsl->flags = fl;
      dsl->R1s = sqrt(W1 + q.C1);
      dsl->R2s = sqrt(W2 + q.C2);
      dsl->R3s = sqrt(W3 + q.C3);
    });
    // H1, G, H2, H3 (Utils.jl:198-249).  The terms of H1 and H2 that are linear in omega and phi
    // are added in spectral space: H1 = fac_n omega / R1 is rfft'd as fac_n omega_hat / R1,
    // and in H2 -eps Lap omega and S1 Lap^2 phi - S2 Lap phi + S3 phi become
    // eps k^2 omega_hat and (S1 k^4 + S2 k^2 + S3) phi_hat (the reference transforms Lap omega
    // and Lap^2 phi back, then H1 and H2 forward).  G, H3 and the rest of H2 are transformed.
    rfft({job(om, w_omh, N)});
    R* G = w_H + NP;
    R* H2n = w_H + 2 * NP;  // the part of H2 not linear in omega or phi
    R* H3 = w_H + 3 * NP;
    foreach_n(s, NP, [=] __device__(long long k) {
      const int n = (int)(k / Pl);
      const T ph = phi[k], ps = psi[k], w = om[k];
      const T eps = (T)q.eps;
      const T sb = (T)q.gs * (T)(3.0 * SQ2 / 4.0) * w + (T)q.gb * (T)(3.0 * SQ2 / (8.0 * q.eps)) * (w * (T(3) * (ph * ph) - T(1)) / eps);
      H2n[k] = sb / (T)dsl->R2s;
      const T pp = T(-1.5) * (ph * ph) + T(1.5);
      const T di = ps - (T)pin[n], dou = ps - (T)pout[n];
      const T fin = T(0.5) * (T)q.gin * (di * di) + (T)q.bin;
      const T fou = T(0.5) * (T)q.gout * (dou * dou) + (T)q.bout;
      H3[k] = (T(0.5) * pp * (fin - fou)) / (T)dsl->R3s;
      const T P_ = T(-0.5) * (ph * ph * ph) + T(1.5) * ph;
      const T G1 = T(0.5) * (T(1) + P_) * ((T)q.gin * di) + T(0.5) * (T(1) - P_) * ((T)q.gout * dou);
      const T G2 = (T)q.S4 * ps;
      G[k] = (G1 - G2) / (T)dsl->R3s;
    });
    rfft({job(G, w_Hh + NH, 3 * N)});
    C* Hh = w_Hh;
    const C* omh = w_omh;
    const C* phh = w_phat;
    const SpecC sc = spc();
    const KIdx ki{Nxh, Ny, Nz};
    const long long Hl = H;
    foreach_n(s, NH, [=] __device__(long long i) {
      int ix, iy, iz, f;
      ki(i, ix, iy, iz, f);
      const int n = (int)(i / Hl);
      const double kx = sc.kx[ix], ky = sc.ky[iy], kz = sc.kz[iz];
      const double k2 = kx * kx + ky * ky + kz * kz;
      const CD wv = omh[i], pv = phh[i];
      const double fac = q.garea * (dsl->area[n] - A0[n]) * (3.0 * SQ2 / 4.0) / dsl->R1s;
      Hh[i] = CD{fac * wv.x, fac * wv.y};
      const double cw = q.gb * (3.0 * SQ2 / (8.0 * q.eps)) * q.eps * k2;
      const double cp = sc.S1 * (k2 * k2) + sc.S2 * k2 + sc.S3;
      const double r2 = 1.0 / dsl->R2s;
      const CD h2 = Hh[2 * NH + i];
      Hh[2 * NH + i] = CD{h2.x + (cw * wv.x - cp * pv.x) * r2, h2.y + (cw * wv.y - cp * pv.y) * r2};
    });
    if (phys_H) irfft({job(w_Hh, w_H, 4 * N)});
  }

  // ---- spectral operator helpers (device) ------------------------------------
  struct SpecC {
    const double *kx, *ky, *kz;
    double S1, S2, S3, S4, Mphi, eta, M0;
  };
  SpecC spc() const { return SpecC{d_kx, d_ky, d_kz, c.S1, c.S2, c.S3, c.S4, c.M_phi, c.eta, c.M0_psi}; }

  // Right-hand sides of the psi solves: b_z = fft(irfft(rhs_z)) (Solvers.jl:70-71) for nsys
  // systems of N half spectra (rhs_z at z N H, b_z at z vs), written directly: the full
  // spectrum of the real field irfft(X) is, for kx <= Nx/2, the part of X that a c2r keeps
  // (the Hermitian part (X(k) + conj X(-k)) / 2 on the kx = 0 and kx = Nx/2 columns, X
  // itself elsewhere) and beyond that its conjugate at (Nx - kx, -ky); equal to the two
  // transforms up to their round-off, in one pass.
  void full_rhs(const C* rhs, C* b, int nsys, long long vs) {
    const long long NP = (long long)N * P, nh = (long long)N * H, Hl = H;
    const int Nx_ = Nx, Ny_ = Ny, Nxh_ = Nxh;
    foreach_n(s, nsys * NP, [=] __device__(long long i) {
      const long long z = i / NP, j = i % NP;
      const int ix = (int)(j % Nx_);
      const long long r = j / Nx_;
      const int iy = (int)(r % Ny_);
      const long long f = r / Ny_;
      const int my = iy ? Ny_ - iy : 0;
      const C* X = rhs + z * nh + f * Hl;
      CD v;
      if (ix < Nxh_) {
        v = X[(long long)iy * Nxh_ + ix];
        if (ix == 0 || 2 * ix == Nx_) {
          const CD m = X[(long long)my * Nxh_ + ix];
          v = CD{T(0.5) * (v.x + m.x), T(0.5) * (v.y - m.y)};
        }
      } else {
        const CD m = X[(long long)my * Nxh_ + (Nx_ - ix)];
        v = CD{m.x, -m.y};
      }
      b[z * vs + j] = v;
    });
  }

  // ==========================================================================
  // psi solve (2D): out = solve_psi!(rhs) -- Solvers.jl:68-81
  // ==========================================================================
  int psi_solve(const C* rhs, C* out, int slot) {
    full_rhs(rhs, kc.b, 1, (long long)N * P);
    int it;
    if (mixed) {
      mobility_fp32();
      it = refined_solve(slot, 1);
    } else if (use_learned && !collecting) {
      if (stencil_stale) refresh_stencil();
      it = fbicgstab_2d<R>(kc, slot, [this](const C* in, C* o) { learned_apply(in, o); });
      if (!hks[slot].solved) {
        // a preconditioner that failed to reach the tolerance is not trusted: the solve
        // is repeated from x0 = 0 with the reference's diagonal preconditioner; the
        // iteration count reports both attempts
        ++learned_fallbacks;
        it += bicgstab_2d<R>(kc, slot);
      }
    } else {
      if (collecting) {
        kc.snap = snapbuf;
        kc.snap_cap = c.itmax;
      }
      it = bicgstab_2d<R>(kc, slot);
      kc.snap = nullptr;
    }
    const C* x = kc.x;
    const int Nxl = Nx, Nxhl = Nxh;
    foreach_n(s, (long long)N * H, [=] __device__(long long i) {
      const long long row = i / Nxhl;
      const int ix = (int)(i % Nxhl);
      out[i] = x[row * Nxl + ix];
    });
    if (collecting) harvest(kc.snap_n);
    return it;
  }

  // M_psi in FP32 for the inner iterations of the mixed-precision solve.
  void mobility_fp32() {
    const R* M = w_M;
    float* m32 = d_M32;
    foreach_n(s, (long long)N * P, [=] __device__(long long i) { m32[i] = (float)M[i]; });
  }
  // The step's three solves start from the solutions of the previous steps (kc.x holds the
  // last ones x_n, x_prev x_{n-1}, x_prev2 x_{n-2}): from x_n after one step, from 2 x_n - x_{n-1}
  // after two, from 3 x_n - 3 x_{n-1} + x_{n-2} after three.  A single-system solve
  // (sav_psi_solve) starts from 0 and drops the history.
  int mixed_hist = 0;
  C* x_prev = nullptr;
  C* x_prev2 = nullptr;
  int refined_solve(int slot, int nsys) {
    if constexpr (std::is_same<R, double>::value) {
      const long long n3 = 3 * kc.nfull;
      if (nsys == 3 && mixed_hist >= 3) {
        C* x2 = x_prev2;
        const C *xc = kc.x, *x1 = x_prev;
        foreach_n(s, n3, [=] __device__(long long i) {
          const CD a = xc[i], b = x1[i], c0 = x2[i];
          x2[i] = CD{3.0 * (a.x - b.x) + c0.x, 3.0 * (a.y - b.y) + c0.y};
        });
        // kc.x: the extrapolation; x_prev: x_n; x_prev2: x_{n-1}
        C* e = x_prev2;
        x_prev2 = x_prev;
        x_prev = kc.x;
        kc.x = e;
      } else if (nsys == 3 && mixed_hist == 2) {
        C* x2 = x_prev2;
        const C *xc = kc.x, *x1 = x_prev;
        foreach_n(s, n3, [=] __device__(long long i) {
          const CD a = xc[i], b = x1[i];
          x2[i] = CD{2.0 * a.x - b.x, 2.0 * a.y - b.y};
        });
        C* e = x_prev2;
        x_prev2 = x_prev;
        x_prev = kc.x;
        kc.x = e;
      } else if (nsys == 3 && mixed_hist == 1) {
        SAV_HIP(hipMemcpyAsync(x_prev, kc.x, n3 * sizeof(C), hipMemcpyDeviceToDevice, s));
      }
      const bool warm = nsys == 3 && mixed_hist >= 1;
      const int it = bicgstab_refined(kc, kc32, slot, nsys, MIXED_ROUNDS, warm);
      mixed_hist = nsys == 3 ? std::min(mixed_hist + 1, 3) : 0;
      return it;
    } else {
      throw HipError("the mixed-precision solve needs FP64 storage", SAV_E_UNSUPPORTED);
    }
  }

  // The three psi solves of step 1 (rhs: 3 consecutive half spectra of N fields) as one
  // batched BiCGSTAB: systems z = 0, 1, 2 use KScal slots 0, 1, 2 and the vectors at
  // offset z * kc.vs.  Each system's iterates equal those of its own solve.
  void psi_solve3(const C* rhs, C* o0, C* o1, C* o2) {
    const long long nf = kc.nfull;
    const long long vs = kc.vs;
    full_rhs(rhs, kc.b, 3, vs);
    if (mixed) {
      mobility_fp32();
      refined_solve(0, 3);
    } else if (capture_iters > 0) {
      bicgstab_queue<R>(kc, 0, 3, capture_iters);
    } else if (group) {
      group_solve();  // one batched solve with the group's other members
    } else {
      bicgstab_batch<R>(kc, 0, 3);
    }
    const C* x = kc.x;
    const int Nxl = Nx, Nxhl = Nxh;
    const long long nh = (long long)N * H;
    foreach_n(s, 3 * nh, [=] __device__(long long i) {
      const long long z = i / nh, j = i % nh;
      const long long row = j / Nxhl;
      const int ix = (int)(j % Nxhl);
      C* o = z == 0 ? o0 : (z == 1 ? o1 : o2);
      o[j] = x[z * vs + row * Nxl + ix];
    });
    (void)nf;
  }

  // ==========================================================================
  // Learned psi preconditioner (no reference counterpart).
  //   z = P ( r + F( g ) ),  g(x) = kappa * sum_tap w_tap(x) * Re F^{-1}(P r)(x + d(tap))
  // P = the diagonal preconditioner of Solvers.jl:57, F = the full 2D FFT,
  // kappa = S4 * M_max / (dx dy), and w(x) the 5-point stencil of -div(c grad .) with face
  // averages of a coefficient c(x) = net(4 (M_psi / M_max - 0.75), 1) (the first output
  // channel of the network), computed once per M_psi (once per step).  The form is that of
  // the analytic first-order stencil (c = 1 - M_psi / M_max): for any c it annihilates
  // constants and is symmetric, as the exact correction (a divergence) is, so an error of
  // the network perturbs the differences of v, not v itself.  The
  // preconditioner is therefore a fixed LINEAR operator during a solve, which BiCGSTAB
  // needs.  Its target: A^{-1} = P (I + (P^{-1} - A) A^{-1}) and
  // (P^{-1} - A) e = S4 div((M_max - M_psi) grad e), a local operator of e ~ P r.  The
  // network is trained on windows logged from diagonal solves, e = x* - x_k (error of
  // iterate k), r = A e, v = Re F^{-1}(P r), against one of two targets: the correction
  // Re F^{-1}(P^{-1} e - r) / kappa of the solver's own error, or the analytic
  // first-order stencil's output on the same v (distillation: that stencil, a function of
  // M_psi on a 3x3 neighbourhood, is within the network's reach).  The outer P damps the
  // BF16 noise of the stencil at high k.  The flexible BiCGSTAB (fbicgstab_2d) that uses
  // it stops on the same diagonally preconditioned residual as the reference's solve.
  // ==========================================================================
  std::unique_ptr<Net> net;
  bool use_learned = false, collecting = false, stencil_stale = true;
  int precond_kind = SAV_PRECOND_DIAG;
  long long learned_fallbacks = 0;  // learned/stencil solves repeated with the diagonal preconditioner
  bf16_t* lp_in = nullptr;   // [N][Ny][Nx][2]: mobility channel, constant channel
  float* lp_W = nullptr;     // [N][Ny][Nx][9]: stencil weights
  C* lp_tmp = nullptr;       // full complex work vectors
  C* lp_g = nullptr;
  C* snapbuf = nullptr;      // itmax iterates of one diagonal solve
  int ppsnap = 0;
  uint64_t rng = 0;
  long long store_cap = 0, store_n = 0;
  bf16_t* st_in = nullptr;   // [store_cap][PATCH][PATCH][2]
  float* st_v = nullptr;     // [store_cap][PATCH][PATCH]
  float* st_tg = nullptr;    // [store_cap][PATCH][PATCH]: correction of the solver's error
  float* st_ta = nullptr;    // [store_cap][PATCH][PATCH]: the analytic first-order stencil's
  int train_target = 0;      // targets of precond_train: 0 st_tg, 1 st_ta
  int* d_pos = nullptr;      // window origins (image, y0, x0)
  std::vector<int> h_pos;
  bf16_t* bt_in = nullptr;   // training batch
  float* bt_v = nullptr;
  float* bt_tg = nullptr;
  int* d_idx = nullptr;
  int bcap = 0;
  std::vector<int> h_idx;

  uint64_t next_rand() {  // SplitMix64
    rng += 0x9E3779B97F4A7C15ULL;
    uint64_t z = rng;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }

  void require_learned_grid() const {
    if (nd != 2) throw HipError("the learned preconditioner is built for the 2D psi solve", SAV_E_UNSUPPORTED);
    if (Nx % 32 != 0 || Ny % 16 != 0)
      throw HipError("the learned preconditioner needs Nx % 32 == 0 and Ny % 16 == 0", SAV_E_UNSUPPORTED);
  }

  void make_net() {
    if (net) return;
    net.reset(new Net(s));
    net->init_random(0x5EED5EEDULL);
  }

  void ensure_net() {
    require_learned_grid();
    make_net();
    if (lp_in) return;
    const long long nf = kc.nfull;
    lp_in = dalloc<bf16_t>(2 * nf);
    lp_W = dalloc<float>(9 * nf);
    lp_tmp = dalloc<C>(nf);
    lp_g = dalloc<C>(nf);
    kc.Bp = dalloc<C>(nf);
    kc.Bs = dalloc<C>(nf);
  }

  struct PD {  // diagonal preconditioner of Solvers.jl:57 on the full grid
    const double *kx0, *ky;
    int Nx, Ny;
    double S4;
    const DScal* ds;
    __device__ __forceinline__ double operator()(long long i) const {
      const int ix = (int)(i % Nx), iy = (int)((i / Nx) % Ny);
      const double lapf = -(kx0[ix] * kx0[ix] + ky[iy] * ky[iy]);
      return 1.0 / (ds->a / ds->dt - ds->Mmax * lapf * S4);
    }
  };
  PD pdiag() const { return PD{d_kx0, d_ky, Nx, Ny, c.S4, ds}; }
  double kappa_over_mmax() const { return c.S4 * (double)Nx * Ny / (c.Lx * c.Ly); }

  // network input of every pixel: (4 (M / M_max - 0.75), 1)
  void mobility_input(bf16_t* xin) {
    const R* M = w_M;
    const DScal* dsl = ds;
    foreach_n(s, kc.nfull, [=] __device__(long long i) {
      xin[2 * i] = f2bf((float)(4.0 * ((double)M[i] / dsl->Mmax - 0.75)));
      xin[2 * i + 1] = f2bf(1.0f);
    });
  }

  void refresh_stencil() {
    if (precond_kind == SAV_PRECOND_STENCIL) {
      // first-order Neumann stencil: g = -S4 div((M_max - M) grad v), 5-point form with
      // face averages, in units of kappa = S4 M_max / (dx dy)
      const R* M = w_M;
      const DScal* dsl = ds;
      float* w = lp_W;
      const int Nxl = Nx, Nyl = Ny;
      const double rx = (c.Ly / Ny) / (c.Lx / Nx), ry = 1.0 / rx;  // dy/dx, dx/dy
      foreach_n(s, kc.nfull, [=] __device__(long long i) {
        const int ix = (int)(i % Nxl);
        const long long row = i / Nxl;
        const int iy = (int)(row % Nyl);
        const long long base = (row - iy) * Nxl;
        const double mm = dsl->Mmax;
        auto D = [&](int y, int x) { return 1.0 - (double)M[base + (long long)((y + Nyl) % Nyl) * Nxl + (x + Nxl) % Nxl] / mm; };
        const double dc = D(iy, ix);
        const double de = 0.5 * (dc + D(iy, ix + 1)) * rx, dw = 0.5 * (dc + D(iy, ix - 1)) * rx;
        const double dn = 0.5 * (dc + D(iy + 1, ix)) * ry, dsv = 0.5 * (dc + D(iy - 1, ix)) * ry;
        float* wi = w + i * 9;
        wi[0] = 0.f;
        wi[1] = (float)(-dsv);
        wi[2] = 0.f;
        wi[3] = (float)(-dw);
        wi[4] = (float)(de + dw + dn + dsv);
        wi[5] = (float)(-de);
        wi[6] = 0.f;
        wi[7] = (float)(-dn);
        wi[8] = 0.f;
      });
    } else {
      // the network's coefficient c (first output channel) in the same divergence form:
      // g = -S4 div(c grad v), the analytic stencil being c = 1 - M / M_max
      mobility_input(lp_in);
      net->forward(lp_in, N, Ny, Nx, lp_W, false);
      float* w = lp_W;
      float* cf = reinterpret_cast<float*>(lp_tmp);  // c of every pixel (lp_tmp is free here)
      foreach_n(s, kc.nfull, [=] __device__(long long i) { cf[i] = w[i * 9]; });
      const int Nxl = Nx, Nyl = Ny;
      const double rx = (c.Ly / Ny) / (c.Lx / Nx), ry = 1.0 / rx;  // dy/dx, dx/dy
      foreach_n(s, kc.nfull, [=] __device__(long long i) {
        const int ix = (int)(i % Nxl);
        const long long row = i / Nxl;
        const int iy = (int)(row % Nyl);
        const long long base = (row - iy) * Nxl;
        auto C_ = [&](int y, int x) { return (double)cf[base + (long long)((y + Nyl) % Nyl) * Nxl + (x + Nxl) % Nxl]; };
        const double cc = C_(iy, ix);
        const double de = 0.5 * (cc + C_(iy, ix + 1)) * rx, dw = 0.5 * (cc + C_(iy, ix - 1)) * rx;
        const double dn = 0.5 * (cc + C_(iy + 1, ix)) * ry, dsv = 0.5 * (cc + C_(iy - 1, ix)) * ry;
        float* wi = w + i * 9;
        wi[0] = 0.f;
        wi[1] = (float)(-dsv);
        wi[2] = 0.f;
        wi[3] = (float)(-dw);
        wi[4] = (float)(de + dw + dn + dsv);
        wi[5] = (float)(-de);
        wi[6] = 0.f;
        wi[7] = (float)(-dn);
        wi[8] = 0.f;
      });
    }
    stencil_stale = false;
  }

  void learned_apply(const C* r, C* z) {
    const long long nf = kc.nfull;
    const PD pd = pdiag();
    C* y = lp_tmp;
    foreach_n(s, nf, [=] __device__(long long i) {
      const double pv = pd(i);
      const C ri = r[i];
      y[i] = CD{pv * (double)ri.x, pv * (double)ri.y};
    });
    fft_full_2d<R>(g, lp_tmp, N, true, s);
    const float* w = lp_W;
    const C* yc = lp_tmp;
    C* gg = lp_g;
    const int Nxl = Nx, Nyl = Ny;
    const double kom = kappa_over_mmax();
    const DScal* dsl = ds;
    foreach_n(s, nf, [=] __device__(long long i) {
      const int ix = (int)(i % Nxl);
      const long long row = i / Nxl;
      const int iy = (int)(row % Nyl);
      const long long base = (row - iy) * Nxl;  // first pixel of this field
      double acc = 0.0;
#pragma unroll
      for (int tap = 0; tap < 9; ++tap) {
        const int sy = (iy + tap / 3 - 1 + Nyl) % Nyl, sx = (ix + tap % 3 - 1 + Nxl) % Nxl;
        acc += (double)w[i * 9 + tap] * (double)yc[base + (long long)sy * Nxl + sx].x;
      }
      gg[i] = CD{kom * dsl->Mmax * acc, 0.0};
    });
    fft_full_2d<R>(g, lp_g, N, false, s);
    const C* gc = lp_g;
    foreach_n(s, nf, [=] __device__(long long i) {
      const double pv = pd(i);
      const C ri = r[i], gi = gc[i];
      z[i] = CD{pv * ((double)ri.x + (double)gi.x), pv * ((double)ri.y + (double)gi.y)};
    });
  }

  // Training pairs from the iterates of the diagonal solve that just finished
  // (kc.x = x*, snapbuf[k-1] = x_k): e_k = x* - x_k for k = 0 .. nsnap-1 (x_0 = 0).
  void harvest(int nsnap) {
    const long long nf = kc.nfull;
    const double norm = (double)Nx * Ny * (double)nf;  // sum|x|^2 = sum|X|^2 / (Nx Ny)
    mobility_input(lp_in);
    const PD pd = pdiag();
    const double kom = kappa_over_mmax();
    for (int k = 0; k < nsnap && store_n < store_cap; ++k) {
      C* e = kc.s;
      C* rr = kc.t;
      C* dl = kc.v;
      const C* xs = kc.x;
      const C* xk = k == 0 ? nullptr : snapbuf + (long long)(k - 1) * nf;
      foreach_n(s, nf, [=] __device__(long long i) {
        const C a = xs[i];
        e[i] = xk ? CD{(double)a.x - (double)xk[i].x, (double)a.y - (double)xk[i].y} : CD{(double)a.x, (double)a.y};
      });
      matvec_plain_2d<R>(kc, e, rr, 3);
      C* y = lp_tmp;
      const C* rc = rr;
      foreach_n(s, nf, [=] __device__(long long i) {
        const double pv = pd(i);
        const C ri = rc[i], ei = e[i];
        y[i] = CD{pv * (double)ri.x, pv * (double)ri.y};
        dl[i] = CD{(double)ei.x / pv - (double)ri.x, (double)ei.y / pv - (double)ri.y};
      });
      const C* yc = lp_tmp;
      reduce_n<1>(s, nf, 1, [=] __device__(long long i, int, double(&acc)[1]) {
        const C v = yc[i];
        acc[0] += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
      }, partials, red);
      fft_full_2d<R>(g, lp_tmp, N, true, s);
      fft_full_2d<R>(g, dl, N, true, s);
      const long long np = std::min<long long>(ppsnap, store_cap - store_n);
      h_pos.resize(3 * np);
      for (long long p = 0; p < np; ++p) {
        h_pos[3 * p] = (int)(next_rand() % (uint64_t)N);
        h_pos[3 * p + 1] = (int)(next_rand() % (uint64_t)Ny);
        h_pos[3 * p + 2] = (int)(next_rand() % (uint64_t)Nx);
      }
      SAV_HIP(hipMemcpyAsync(d_pos, h_pos.data(), 3 * np * sizeof(int), hipMemcpyHostToDevice, s));
      const double* rd = red;
      const C* d = dl;
      const DScal* dsl = ds;
      const bf16_t* xin = lp_in;
      const int* pos = d_pos;
      bf16_t* tin = st_in + store_n * (long long)(2 * PATCH * PATCH);
      float* tv = st_v + store_n * (long long)(PATCH * PATCH);
      float* ttg = st_tg + store_n * (long long)(PATCH * PATCH);
      float* tta = st_ta + store_n * (long long)(PATCH * PATCH);
      const R* M = w_M;
      const int Nxl = Nx, Nyl = Ny;
      const double rx = (c.Ly / Ny) / (c.Lx / Nx), ry = 1.0 / rx;  // dy/dx, dx/dy
      foreach_n(s, np * PATCH * PATCH, [=] __device__(long long i) {
        const long long p = i / (PATCH * PATCH);
        const int py = (int)((i / PATCH) % PATCH), px = (int)(i % PATCH);
        const int img = pos[3 * p], yy = (pos[3 * p + 1] + py) % Nyl, xx = (pos[3 * p + 2] + px) % Nxl;
        const long long src = ((long long)img * Nyl + yy) * Nxl + xx;
        const double sy = sqrt(rd[0] / norm);
        const double iy = sy > 0.0 ? 1.0 / sy : 0.0;
        tin[2 * i] = xin[2 * src];
        tin[2 * i + 1] = xin[2 * src + 1];
        tv[i] = (float)((double)yc[src].x * iy);
        ttg[i] = (float)((double)d[src].x * iy / (kom * dsl->Mmax));
        // the analytic first-order stencil's prediction at this pixel (refresh_stencil's
        // weights on the same v): the distillation target
        const long long fb = (long long)img * Nyl * Nxl;
        auto at = [&](int y2, int x2) { return fb + (long long)((y2 + Nyl) % Nyl) * Nxl + (x2 + Nxl) % Nxl; };
        auto Dm = [&](int y2, int x2) { return 1.0 - (double)M[at(y2, x2)] / dsl->Mmax; };
        auto V = [&](int y2, int x2) { return (double)yc[at(y2, x2)].x; };
        const double dc = Dm(yy, xx);
        const double de = 0.5 * (dc + Dm(yy, xx + 1)) * rx, dw = 0.5 * (dc + Dm(yy, xx - 1)) * rx;
        const double dn = 0.5 * (dc + Dm(yy + 1, xx)) * ry, dsv = 0.5 * (dc + Dm(yy - 1, xx)) * ry;
        const double an = (de + dw + dn + dsv) * V(yy, xx) - de * V(yy, xx + 1) - dw * V(yy, xx - 1) -
                          dn * V(yy + 1, xx) - dsv * V(yy - 1, xx);
        tta[i] = (float)(an * iy);
      });
      store_n += np;
    }
  }

  int set_precond(int mode) override {
    if (mixed && mode != SAV_PRECOND_DIAG) return SAV_E_UNSUPPORTED;  // diagonal inner solves only
    if (mode != SAV_PRECOND_DIAG && mode != SAV_PRECOND_LEARNED && mode != SAV_PRECOND_STENCIL) return SAV_E_ARG;
    if (mode != SAV_PRECOND_DIAG) ensure_net();
    use_learned = mode != SAV_PRECOND_DIAG;
    precond_kind = mode;
    stencil_stale = true;
    return SAV_OK;
  }

  int precond_init(uint64_t seed) override {
    make_net();
    net->init_random(seed);
    stencil_stale = true;
    return SAV_OK;
  }

  int precond_params(bool set, float* p, int n) override {
    make_net();
    if (n != Net::NPARAM) return SAV_E_ARG;
    if (set) {
      net->set_params(p);
      stencil_stale = true;
    } else {
      net->get_params(p);
    }
    return SAV_OK;
  }

  int precond_collect(int on, int per_snapshot, uint64_t seed) override {
    if (mixed) return SAV_E_UNSUPPORTED;
    ensure_net();
    if (!on) {
      collecting = false;
      return SAV_OK;
    }
    if (per_snapshot <= 0) return SAV_E_ARG;
    const long long nf = kc.nfull;
    if (!snapbuf) snapbuf = dalloc<C>((long long)c.itmax * nf);
    if (!st_in) {
      store_cap = 24576;  // windows of 64 x 64: 805 MB
      st_in = dalloc<bf16_t>(store_cap * 2 * PATCH * PATCH);
      st_v = dalloc<float>(store_cap * PATCH * PATCH);
      st_tg = dalloc<float>(store_cap * PATCH * PATCH);
      st_ta = dalloc<float>(store_cap * PATCH * PATCH);
    }
    if (!d_pos || per_snapshot > ppsnap) d_pos = dalloc<int>(3LL * per_snapshot);
    ppsnap = per_snapshot;
    rng = seed;
    store_n = 0;
    collecting = true;
    return SAV_OK;
  }

  long long precond_samples() override { return store_n; }
  long long precond_fallbacks() override { return learned_fallbacks; }

  int set_column_tile(int cw) override {
    if (nd != 2) return SAV_E_UNSUPPORTED;
    if (cw != 0 && cw != 4 && cw != 8 && cw != 16) return SAV_E_ARG;
    if (cw != kc.colw) drop_graphs();  // captured matvecs use the previous width
    kc.colw = cw;
    return SAV_OK;
  }

  // Autotuner of the column transforms outside the psi matvec (steps 1-7, back-transforms,
  // monitor): at a handle's first step inverse transforms (with the Laplacian multiplier) and
  // forward transforms of the present state are timed with 4, 8 and 16 columns per workgroup,
  // and the fastest width of each direction is kept for every later transform of the handle.
  // Widths that resolve to the same tile for this grid's column lengths are timed once.  Each
  // column is transformed by the same team code whatever tile it shares, so every width gives
  // bit-identical results.
  int tcolw = 0, tcolw_f = 0;                 // inverse and forward widths
  double tune_t_ms[3] = {0.0, 0.0, 0.0};      // inverse transforms with 4, 8, 16 columns
  double tune_f_ms[3] = {0.0, 0.0, 0.0};      // forward transforms
  int pkind = 0;  // plane kernels: 0 = tune at the next step, 1 = register columns, 2 = LDS plane
  double tune_p_ms[2] = {0.0, 0.0};
  // Milliseconds of one irfft and one rfft of the present phi (three repeats after a warm-up).
  double time_transforms() {
    auto once = [&]() {
      irfft({job(fhalf(ip, F_PHI), w_d4, N, MUL_LAP)});
      rfft({job(fphys(ip, F_PHI), w_phat, N)});
    };
    hipEvent_t a, b;
    SAV_HIP(hipEventCreate(&a));
    SAV_HIP(hipEventCreate(&b));
    once();  // warm-up
    SAV_HIP(hipEventRecord(a, s));
    for (int r = 0; r < 3; ++r) once();
    SAV_HIP(hipEventRecord(b, s));
    SAV_HIP(hipEventSynchronize(b));
    float ms = 0.f;
    SAV_HIP(hipEventElapsedTime(&ms, a, b));
    SAV_HIP(hipEventDestroy(a));
    SAV_HIP(hipEventDestroy(b));
    return ms / 3.0;
  }
  void autotune_transforms() {
    autotune_columns();
    if (pkind == 0) {
      if (!g.tw_hp) {
        pkind = 2;  // no register-column kernels for this grid
      } else {
        for (int k = 0; k < 2; ++k) {
          g.plane_kind = k;
          tune_p_ms[k] = time_transforms();
        }
        // the timed pair (one field set, largely cache-resident) favours neither family
        // much; the register kernels move the same bytes with less LDS traffic, so the
        // LDS kernels are kept only when clearly faster
        pkind = tune_p_ms[1] < 0.97 * tune_p_ms[0] ? 2 : 1;
      }
    }
    g.plane_kind = pkind == 1 ? 0 : 1;
  }
  int set_plane_kernel(int kind) override {
    if (kind < 0 || kind > 2) return SAV_E_ARG;
    if (kind == 1 && !g.tw_hp) return SAV_E_UNSUPPORTED;
    if (kind != pkind) drop_graphs();  // captured transforms use the previous kernels
    pkind = kind;
    if (kind) g.plane_kind = kind == 1 ? 0 : 1;
    return SAV_OK;
  }
  void plane_kernel(int* kind, double* ms_register, double* ms_lds) override {
    if (kind) *kind = pkind;
    if (ms_register) *ms_register = tune_p_ms[0];
    if (ms_lds) *ms_lds = tune_p_ms[1];
  }
  void autotune_columns() {
    if (tcolw != 0) return;
    const int widths[3] = {4, 8, 16};
    const int rb = (int)sizeof(R);
    auto key = [&](int w) {
      return col_tile_columns(Ny, rb, w) * 10000 + (nd == 3 ? col_tile_columns(Nz, rb, w) : 0);
    };
    hipEvent_t a, b;
    SAV_HIP(hipEventCreate(&a));
    SAV_HIP(hipEventCreate(&b));
    // the inverse and the forward transforms are timed apart (their best widths differ: at
    // 1024^2 in FP64 the inverse is faster with 4 columns at every field count, the forward
    // one with 8), each over a launch of N fields and one of 4N (the step's launches carry
    // 2 to 11 N-field jobs); w_Hh is recomputed by step 1
    double best_i = 1e300, best_f = 1e300;
    int wi = 0, wf = 0;
    std::vector<int> seen;
    for (int i = 0; i < 3; ++i) {
      const int w = widths[i];
      const int kk = key(w);
      if (std::find(seen.begin(), seen.end(), kk) != seen.end()) {
        tune_t_ms[i] = tune_f_ms[i] = -1.0;  // the same tile as a width already timed
        continue;
      }
      seen.push_back(kk);
      g.colw = g.colw_fwd = w;
      auto inv = [&]() {
        irfft({job(fhalf(ip, F_PHI), w_d4, N, MUL_LAP)});
        irfft({job(fhalf(ip, F_PHI), w_d4, 4 * N, MUL_LAP)});
      };
      auto fwd = [&]() {
        rfft({job(fphys(ip, F_PHI), w_phat, N)});
        rfft({job(fphys(ip, F_PHI), w_Hh, 4 * N)});
      };
      for (int dir = 0; dir < 2; ++dir) {
        if (dir == 0) inv(); else fwd();  // warm-up
        SAV_HIP(hipEventRecord(a, s));
        for (int r = 0; r < 3; ++r) {
          if (dir == 0) inv(); else fwd();
        }
        SAV_HIP(hipEventRecord(b, s));
        SAV_HIP(hipEventSynchronize(b));
        float ms = 0.f;
        SAV_HIP(hipEventElapsedTime(&ms, a, b));
        const double t = ms / 3.0;
        if (dir == 0) {
          tune_t_ms[i] = t;
          if (t < best_i) { best_i = t; wi = w; }
        } else {
          tune_f_ms[i] = t;
          if (t < best_f) { best_f = t; wf = w; }
        }
      }
    }
    SAV_HIP(hipEventDestroy(a));
    SAV_HIP(hipEventDestroy(b));
    tcolw = wi;
    tcolw_f = wf;
    g.colw = tcolw;
    g.colw_fwd = tcolw_f;
  }
  int set_transform_tile(int cw) override {
    if (cw != 0 && cw != 4 && cw != 8 && cw != 16) return SAV_E_ARG;
    if (cw != tcolw || cw != tcolw_f) drop_graphs();  // captured transforms use the previous widths
    tcolw = tcolw_f = cw;
    g.colw = g.colw_fwd = cw;
    return SAV_OK;
  }
  void transform_tile(int* cw, double* ms4, double* ms8, double* ms16) override {
    if (cw) *cw = tcolw;
    if (ms4) *ms4 = tune_t_ms[0];
    if (ms8) *ms8 = tune_t_ms[1];
    if (ms16) *ms16 = tune_t_ms[2];
  }
  void transform_tile_fwd(int* cw, double* ms4, double* ms8, double* ms16) override {
    if (cw) *cw = tcolw_f;
    if (ms4) *ms4 = tune_f_ms[0];
    if (ms8) *ms8 = tune_f_ms[1];
    if (ms16) *ms16 = tune_f_ms[2];
  }

  void column_tile(int* cw, double* ms8, double* ms16, double* ms4) override {
    if (cw) *cw = kc.colw;
    if (ms8) *ms8 = kc.tune_ms[0];
    if (ms16) *ms16 = kc.tune_ms[1];
    if (ms4) *ms4 = kc.tune_ms[2];
  }
  int set_matvec_order(int order) override {
    if (nd != 2) return SAV_E_UNSUPPORTED;
    if (order < 0 || order > 2) return SAV_E_ARG;
    if (order != kc.order) drop_graphs();  // captured matvecs use the previous order
    kc.order = order;
    kc32.order = order;
    return SAV_OK;
  }
  void matvec_order(int* order, double* ms_batched, double* ms_sequenced) override {
    if (order) *order = kc.order;
    if (ms_batched) *ms_batched = kc.tune_order_ms[0];
    if (ms_sequenced) *ms_sequenced = kc.tune_order_ms[1];
  }

  int precond_train(int iters, int batch, double lr, uint64_t seed, double* l0, double* l1) override {
    ensure_net();
    if (store_n < 1 || iters < 1 || batch < 1) return SAV_E_ARG;
    if (batch > bcap) {
      bt_in = dalloc<bf16_t>((long long)batch * 2 * PATCH * PATCH);
      bt_v = dalloc<float>((long long)batch * PATCH * PATCH);
      bt_tg = dalloc<float>((long long)batch * PATCH * PATCH);
      d_idx = dalloc<int>(batch);
      bcap = batch;
    }
    uint64_t save = rng;
    rng = seed;
    double s0 = 0.0, s1 = 0.0;
    int n0 = 0, n1 = 0;
    h_idx.resize(batch);
    for (int it = 0; it < iters; ++it) {
      for (int b = 0; b < batch; ++b) h_idx[b] = (int)(next_rand() % (uint64_t)store_n);
      SAV_HIP(hipMemcpyAsync(d_idx, h_idx.data(), batch * sizeof(int), hipMemcpyHostToDevice, s));
      const int* idx = d_idx;
      const bf16_t* sin = st_in;
      const float *sv = st_v, *stg = train_target == 1 ? st_ta : st_tg;
      bf16_t* bin = bt_in;
      float *bv = bt_v, *btg = bt_tg;
      foreach_n(s, (long long)batch * PATCH * PATCH, [=] __device__(long long i) {
        const long long b = i / (PATCH * PATCH), q = i % (PATCH * PATCH);
        const long long src = (long long)idx[b] * PATCH * PATCH + q;
        bin[2 * i] = sin[2 * src];
        bin[2 * i + 1] = sin[2 * src + 1];
        bv[i] = sv[src];
        btg[i] = stg[src];
      });
      const double rx = (c.Ly / Ny) / (c.Lx / Nx);
      const double loss =
          net->loss_backward_stencil(bt_in, bt_v, bt_tg, batch, PATCH, PATCH, PATCH_BORDER, rx, 1.0 / rx);
      if (it < 10) {
        s0 += loss;
        ++n0;
      }
      if (it >= iters - 10) {
        s1 += loss;
        ++n1;
      }
      // cosine decay from lr to lr / 10
      const double lrt = lr * (0.1 + 0.45 * (1.0 + std::cos(M_PI * (double)it / (double)iters)));
      net->adam_step((float)lrt, 0.9f, 0.999f, 1e-8f);
    }
    SAV_HIP(hipStreamSynchronize(s));
    rng = save;
    stencil_stale = true;
    if (l0) *l0 = s0 / (n0 > 0 ? n0 : 1);
    if (l1) *l1 = s1 / (n1 > 0 ? n1 : 1);
    return SAV_OK;
  }

  int precond_train_target(int kind) override {
    if (kind != 0 && kind != 1) return SAV_E_ARG;
    train_target = kind;
    return SAV_OK;
  }

  int net_forward(const float* x, int nimg, int Hh, int Ww, float* y) override {
    make_net();
    const long long np = (long long)nimg * Hh * Ww;
    std::vector<bf16_t> xb(2 * np);
    for (long long i = 0; i < 2 * np; ++i) xb[i] = f2bf(x[i]);
    bf16_t* dx = nullptr;
    float* dy = nullptr;
    SAV_HIP(hipMalloc(&dx, xb.size() * sizeof(bf16_t) + 64));
    SAV_HIP(hipMalloc(&dy, np * Net::COUT * sizeof(float) + 64));
    SAV_HIP(hipMemcpyAsync(dx, xb.data(), xb.size() * sizeof(bf16_t), hipMemcpyHostToDevice, s));
    net->forward(dx, nimg, Hh, Ww, dy, false);
    SAV_HIP(hipMemcpyAsync(y, dy, np * Net::COUT * sizeof(float), hipMemcpyDeviceToHost, s));
    SAV_HIP(hipStreamSynchronize(s));
    SAV_HIP(hipFree(dx));
    SAV_HIP(hipFree(dy));
    return SAV_OK;
  }

  int net_grad(const float* x, const float* t, int nimg, int Hh, int Ww, int border, float* grads,
               double* loss) override {
    make_net();
    const long long np = (long long)nimg * Hh * Ww;
    std::vector<bf16_t> xb(2 * np);
    for (long long i = 0; i < 2 * np; ++i) xb[i] = f2bf(x[i]);
    bf16_t* dx = nullptr;
    float* dt = nullptr;
    SAV_HIP(hipMalloc(&dx, xb.size() * sizeof(bf16_t) + 64));
    SAV_HIP(hipMalloc(&dt, np * Net::COUT * sizeof(float) + 64));
    SAV_HIP(hipMemcpyAsync(dx, xb.data(), xb.size() * sizeof(bf16_t), hipMemcpyHostToDevice, s));
    SAV_HIP(hipMemcpyAsync(dt, t, np * Net::COUT * sizeof(float), hipMemcpyHostToDevice, s));
    const double l = net->loss_backward(dx, dt, nimg, Hh, Ww, border);
    net->get_grads(grads);
    SAV_HIP(hipFree(dx));
    SAV_HIP(hipFree(dt));
    if (loss) *loss = l;
    return SAV_OK;
  }

  void set_bdf(double a, double b, double cc, double dt) {
    DScal* dsl = ds;
    const double dt5 = nd == 3 ? c.dt : dt;
    single(s, [=] __device__() {
      dsl->a = a;
      dsl->b = b;
      dsl->c = cc;
      dsl->dt = dt;
      dsl->dt5 = dt5;
      dsl->flags = 0;
    });
  }

  // Step 1's advection terms (Solvers.jl:117-131): phi_21 = -rfft(u* . grad phi*) / lhs_phi,
  // mu_21, and the right-hand side of psi_21, with grad phi*, grad psi* from the state spectra
  // 2 X^n - X^{n-1}.  Reads only the state and u* (step 1's extrapolation); writes w_adv,
  // w_advh, the 21 slots of the cache and the third psi right-hand side.  `grad` holds the
  // 2 nd N gradient fields (w_d4 in the serial order, where it is free at this point).
  // grad psi* (nd N-fields) into gpsi
  void step1_advection_grad(R* gpsi) {
    const long long NP = (long long)N * P;
    const C *pshN = fhalf(ip, F_PSI), *pshO = fhalf(io, F_PSI);
    std::vector<SpecJob> jl;
    for (int d = 0; d < nd; ++d) jl.push_back(job(pshN, gpsi + d * NP, N, DX(d), pshO));
    irfft_v(jl);
  }
  // The advection terms from grad psi* (gpsi) and grad phi* (w_gphi, computed by
  // functionals_grad from the Hermitian part of 2 phi_hat^n - phi_hat^{n-1}: the reference
  // differentiates 2 phi_hat^n - phi_hat^{n-1} itself, whose c2r is the same field up to
  // round-off, since phi_hat is the rfft of a real field).
  void step1_advection(const R* gpsi) {
    const long long NP = (long long)N * P, NH = (long long)N * H;
    const R* us = w_us;
    {
      R* adv = w_adv;
      const R* gph = w_gphi;
      const R* gps = gpsi;
      const int ndd = nd;
      const long long Pl = P;
      foreach_n(s, NP, [=] __device__(long long i) {
        const long long sp = i % Pl;
        R a = us[sp] * gph[i];
        R b = us[sp] * gps[i];
        for (int d = 1; d < ndd; ++d) {
          a = a + us[d * Pl + sp] * gph[d * NP + i];
          b = b + us[d * Pl + sp] * gps[d * NP + i];
        }
        adv[i] = a;
        adv[NP + i] = b;
      });
      rfft({job(w_adv, w_advh, 2 * N)});
      // phi_21 = -adv_phi_hat / lhs_phi and mu_21 = L_phi phi_21 are formed where they are used
      // (Splits); here the right-hand side of psi_21
      const C* ah = w_advh;
      C* rhs21 = w_rhs + 2 * NH;
      foreach_n(s, NH, [=] __device__(long long i) {
        const CD t2 = ah[NH + i];
        rhs21[i] = CD{-t2.x, -t2.y};
      });
    }
  }

  // ==========================================================================
  // Step 1 (Solvers.jl:14-137; Solvers_3d.jl:14-92)
  // ==========================================================================
  void step1() {
    const long long NP = (long long)N * P, NH = (long long)N * H;
    const R *phN = fphys(ip, F_PHI), *phO = fphys(io, F_PHI), *psN = fphys(ip, F_PSI), *psO = fphys(io, F_PSI);
    const R *uN = uphys(ip), *uO = uphys(io);
    R *phs = w_phis, *pss = w_psis, *M = w_M, *us = w_us;
    const double M0 = c.M0_psi;
    foreach_n(s, NP, [=] __device__(long long i) {
      const T a = T(2.0) * phN[i] - phO[i];
      phs[i] = a;
      pss[i] = T(2.0) * psN[i] - psO[i];
      const T t = a * a - T(1.0);
      M[i] = T(0.1) * (T(1.0) - (T)M0 * (t * t));
    });
    foreach_n(s, (long long)nd * P, [=] __device__(long long i) { us[i] = T(2.0) * uN[i] - uO[i]; });
    // on the second stream: grad psi*, then (once grad phi* is there) the advection terms and step 4
    auto on_s4 = [&](auto&& f) {
      std::swap(s, s4);
      std::swap(tmp, tmp4);
      try {
        f();
      } catch (...) {
        std::swap(s, s4);
        std::swap(tmp, tmp4);
        throw;
      }
      std::swap(s, s4);
      std::swap(tmp, tmp4);
    };
    if (fork_step4) {
      SAV_HIP(hipEventRecord(ev_fork, s));
      SAV_HIP(hipStreamWaitEvent(s4, ev_fork, 0));
      on_s4([&] { step1_advection_grad(w_grad); });
    }
    if (nd == 2) {
      // M_max = maximum(M_psi) (Solvers.jl:27)
      max_reduce(M, NP, &ds->Mmax);
      stencil_stale = true;
    }
    functionals_grad(phs, fhalf(ip, F_PHI), fhalf(io, F_PHI));
    if (fork_step4) {
      SAV_HIP(hipEventRecord(ev_gphi, s));
      SAV_HIP(hipStreamWaitEvent(s4, ev_gphi, 0));
      on_s4([&] {
        step1_advection(w_grad);
        SAV_HIP(hipEventRecord(ev_adv, s));
        step4();
      });
      SAV_HIP(hipEventRecord(ev_join, s4));
    }
    functionals(phs, pss, false, fhalf(ip, F_PHI), fhalf(io, F_PHI), false, true);  // -> w_Hh
    // 11, 12, 13, 14 spectral divides (Solvers.jl:86-115): phi_1j and mu_1j are pointwise in the
    // state spectra and H_hat and are formed where they are used (Splits); here only the right-hand
    // side of psi_11, -(b psi_hat^n + c psi_hat^{n-1}) / dt
    const C *pshN = fhalf(ip, F_PSI), *pshO = fhalf(io, F_PSI);
    C* rhs = w_rhs;
    const SpecC sc = spc();
    const KIdx ki{Nxh, Ny, Nz};
    const DScal* dsl = ds;
    foreach_n(s, NH, [=] __device__(long long i) {
      const T b = (T)dsl->b, cc = (T)dsl->c, dt = (T)dsl->dt;
      const CD yN = pshN[i], yO = pshO[i];
      rhs[i] = CD{-(b * yN.x + cc * yO.x) / dt, -(b * yN.y + cc * yO.y) / dt};
    });
    // grad G -> M * grad G -> rfft -> rhs12 = sum_d D1_d * (.)   (Solvers.jl:99-105)
    const C* Gh = w_Hh + NH;
    if (nd == 3) {
      // constant mobility M0_psi (Solvers_3d.jl:48-57): rfft(M0 irfft(X)) of X = D1_d G_hat is
      // M0 times the part of X that a c2r keeps, the Hermitian part (X(k) + conj X(-k)) / 2 on
      // the kx = 0 and kx = Nx/2 planes and X itself elsewhere, so the 2 nd N-field transforms
      // of the reference reduce to one spectral pass (equal up to the transforms' round-off)
      C* rhs12 = w_rhs + NH;
      const double m0 = c.M0_psi;
      const int Nx_ = Nx, Nxh_ = Nxh, Ny_ = Ny, Nz_ = Nz;
      foreach_n(s, NH, [=] __device__(long long i) {
        int ix, iy, iz, f;
        ki(i, ix, iy, iz, f);
        const double kk[3] = {sc.kx[ix], sc.ky[iy], sc.kz[iz]};
        const CD gk = Gh[i];
        const bool plane = ix == 0 || 2 * ix == Nx_;
        CD gm = CD{0.0, 0.0};
        double km[3] = {0.0, 0.0, 0.0};
        if (plane) {
          const int iy2 = iy ? Ny_ - iy : 0, iz2 = iz ? Nz_ - iz : 0;
          gm = Gh[(((long long)f * Nz_ + iz2) * Ny_ + iy2) * Nxh_ + ix];
          km[0] = sc.kx[ix];
          km[1] = sc.ky[iy2];
          km[2] = sc.kz[iz2];
        }
        CD acc = CD{0.0, 0.0};
        for (int d = 0; d < 3; ++d) {
          CD x = CD{-kk[d] * gk.y, kk[d] * gk.x};  // X(k) = i k_d G(k)
          if (plane) {
            const CD xm = CD{-km[d] * gm.y, km[d] * gm.x};  // X(-k)
            x = CD{0.5 * (x.x + xm.x), 0.5 * (x.y - xm.y)};
          }
          x = CD{m0 * x.x, m0 * x.y};
          const CD t = CD{-kk[d] * x.y, kk[d] * x.x};
          acc = d == 0 ? t : acc + t;
        }
        rhs12[i] = acc;
      });
    } else {
      std::vector<SpecJob> jl;
      for (int d = 0; d < nd; ++d) jl.push_back(job(Gh, w_d4 + d * NP, N, DX(d)));
      irfft_v(jl);
      R* w = w_w;
      const R* gG = w_d4;
      const int ndl = nd;
      const bool var = nd == 2;
      const T m0 = (T)c.M0_psi;
      foreach_n(s, (long long)nd * NP, [=] __device__(long long i) {
        const T m = var ? M[i % NP] : m0;
        w[i] = m * gG[i];
      });
      (void)ndl;
      rfft({job(w_w, w_wh, nd * N)});
      const C* wh = w_wh;
      C* rhs12 = w_rhs + NH;
      const int ndd = nd;
      foreach_n(s, NH, [=] __device__(long long i) {
        int ix, iy, iz, f;
        ki(i, ix, iy, iz, f);
        const T kk[3] = {(T)sc.kx[ix], (T)sc.ky[iy], (T)sc.kz[iz]};
        CD acc = CD{T(0), T(0)};
        for (int d = 0; d < ndd; ++d) {
          const CD a = wh[d * NH + i];
          const CD t = CD{-kk[d] * a.y, kk[d] * a.x};
          acc = d == 0 ? t : acc + t;
        }
        rhs12[i] = acc;
      });
    }
    // 21: u* . grad(phi*), u* . grad(psi*) (forked with step 4 in an eager step)
    if (fork_step4) {
      SAV_HIP(hipStreamWaitEvent(s, ev_adv, 0));
    } else {
      step1_advection_grad(w_d4 + (long long)nd * NP);
      step1_advection(w_d4 + (long long)nd * NP);
    }
    // psi solves
    if (nd == 2) {
      if (!use_learned && !collecting) {
        // the three solves share the operator: one batched BiCGSTAB (Solvers.jl:73 x 3)
        psi_solve3(w_rhs, cf(PSI11), cf(PSI12), cf(PSI21));
        if (capture_iters == 0)
          for (int z = 0; z < 3; ++z) kit[z] = hks[z].iter;
      } else {
        kit[0] = psi_solve(w_rhs, cf(PSI11), 0);
        kit[1] = psi_solve(w_rhs + NH, cf(PSI12), 1);
        kit[2] = psi_solve(w_rhs + 2 * NH, cf(PSI21), 2);
      }
    } else {
      // spectral divide by a/dt - M0*S4*Lap (Solvers_3d.jl:30,40,57,88)
      const C* r = w_rhs;
      C* p11 = cf(PSI11);
      C* p12 = cf(PSI12);
      C* p21 = cf(PSI21);
      foreach_n(s, NH, [=] __device__(long long i) {
        int ix, iy, iz, f;
        ki(i, ix, iy, iz, f);
        const double kx = sc.kx[ix], ky = sc.ky[iy], kz = sc.kz[iz];
        const double lapd = -(kx * kx + ky * ky + kz * kz);
        const T lh = (T)((dsl->a / dsl->dt) - sc.M0 * sc.S4 * lapd);
        const CD a = r[i], b = r[NH + i], cc = r[2 * NH + i];
        p11[i] = CD{a.x / lh, a.y / lh};
        p12[i] = CD{b.x / lh, b.y / lh};
        p21[i] = CD{cc.x / lh, cc.y / lh};
      });
      kit[0] = kit[1] = kit[2] = 0;
    }
    // nu_11, nu_12, nu_21 = S4 psi (+ G_hat for nu_12) are written by step 2's pass
  }

  int kit[3] = {0, 0, 0};
  int capture_iters = 0;  // > 0 while a step is captured: psi iterations queued without host checks

  void max_reduce(const R* x, long long n, double* out) {
    reduce_max_impl(x, n, out);
  }
  void reduce_max_impl(const R* x, long long n, double* out);

  // ==========================================================================
  // Step 2 (Solvers.jl:144-234): 22 inner products, two 3x3 LU solves
  // ==========================================================================
  void step2() {
    const long long NH = (long long)N * H;
    // The 22 inner products (Solvers.jl:144-234) by Parseval over the half spectra; the
    // reference transforms the eight split solutions back and sums in physical space.  H_hat
    // and G_hat are rffts of real fields, so sum_x h f = (1/P) sum_k w(kx) Re(conj(h_hat) f_hat)
    // with w = 1 on the kx = 0 and kx = Nx/2 planes and 2 elsewhere, for any f_hat whose c2r is
    // f: the split solutions, and b phi^n + c phi^{n-1} (b psi^n + c psi^{n-1}) from the state
    // spectra, whose c2r are the states' physical fields.
    const C* Hh = w_Hh;  // [H1 | G | H2 | H3]
    const C *phN = fhalf(ip, F_PHI), *phO = fhalf(io, F_PHI), *psN = fhalf(ip, F_PSI), *psO = fhalf(io, F_PSI);
    const DScal* dsl = ds;
    const int Nxh_ = Nxh, Nx_ = Nx;
    const Splits sp = splits();
    reduce_n<22>(s, NH, 1,
                 [=] __device__(long long i, int, double(&acc)[22]) {
                   const int ix = (int)(i % Nxh_);
                   const double w = (ix == 0 || 2 * ix == Nx_) ? 1.0 : 2.0;
                   const CD h1 = Hh[i], gg = Hh[NH + i], h2 = Hh[2 * NH + i], h3 = Hh[3 * NH + i];
                   auto re = [&](const CD& a, const CD& f) {
                     return w * ((double)a.x * (double)f.x + (double)a.y * (double)f.y);
                   };
                   const T b = (T)dsl->b, cc = (T)dsl->c;
                   const CD xn = phN[i], xo = phO[i], yn = psN[i], yo = psO[i];
                   const CD tph = CD{b * xn.x + cc * xo.x, b * xn.y + cc * xo.y};
                   const CD tps = CD{b * yn.x + cc * yo.x, b * yn.y + cc * yo.y};
                   acc[0] += re(h1, tph);
                   acc[1] += re(h2, tph);
                   acc[2] += re(h3, tph);
                   acc[3] += re(gg, tps);
                   CD ph[5], mu[5];
                   sp.phi_mu(i, ph, mu);
                   const CD p11 = ph[0], p12 = ph[1], p13 = ph[2], p14 = ph[3], p21 = ph[4];
                   const CD s11 = sp.at(PSI11, i), s12 = sp.at(PSI12, i), s21 = sp.at(PSI21, i);
                   acc[4] += re(h1, p12);
                   acc[5] += re(h1, p13);
                   acc[6] += re(h1, p14);
                   acc[7] += re(h2, p12);
                   acc[8] += re(h2, p13);
                   acc[9] += re(h2, p14);
                   acc[10] += re(h3, p12);
                   acc[11] += re(h3, p13);
                   acc[12] += re(h3, p14);
                   acc[13] += re(gg, s12);
                   acc[14] += re(h1, p11);
                   acc[15] += re(h2, p11);
                   acc[16] += re(h3, p11);
                   acc[17] += re(gg, s11);
                   acc[18] += re(h1, p21);
                   acc[19] += re(h2, p21);
                   acc[20] += re(h3, p21);
                   acc[21] += re(gg, s21);
                 },
                 partials, red);
    const PhysC q = pc();
    const int ndl = nd;
    const double* rd = red;
    DScal* dsw = ds;
    const DStateScal* stN = dstate + ip;
    const DStateScal* stO = dstate + io;
    const double invP = 1.0 / (double)P;
    single(s, [=] __device__() {
      auto idot = [&](int k) { return integ(rd[k] * invP, q, ndl); };
      const double a = dsw->a, b = dsw->b, cc = dsw->c;
      const double U1 = -(b * stN->R1 + cc * stO->R1) / a + idot(0) / (2.0 * a);
      const double U2 = -(b * stN->R2 + cc * stO->R2) / a + idot(1) / (2.0 * a);
      const double U3 = -(b * stN->R3 + cc * stO->R3) / a + (idot(2) + idot(3)) / (2.0 * a);
      double L[9];
      L[0] = 1.0 - 0.5 * idot(4);
      L[1] = -0.5 * idot(5);
      L[2] = -0.5 * idot(6);
      L[3] = -0.5 * idot(7);
      L[4] = 1.0 - 0.5 * idot(8);
      L[5] = -0.5 * idot(9);
      L[6] = -0.5 * idot(10);
      L[7] = -0.5 * idot(11);
      L[8] = 1.0 - 0.5 * (idot(12) + idot(13));
      double r1[3], r2[3];
      r1[0] = 0.5 * idot(14) + U1;
      r1[1] = 0.5 * idot(15) + U2;
      r1[2] = 0.5 * (idot(16) + idot(17)) + U3;
      r2[0] = 0.5 * idot(18);
      r2[1] = 0.5 * idot(19);
      r2[2] = 0.5 * (idot(20) + idot(21));
      for (int k = 0; k < 9; ++k) dsw->lhs[k] = L[k];
      for (int k = 0; k < 3; ++k) {
        dsw->rhs1[k] = r1[k];
        dsw->rhs2[k] = r2[k];
      }
      // LU with partial pivoting (LAPACK getrf/getrs), both right-hand sides
      int piv[3] = {0, 1, 2};
      double A[9];
      for (int k = 0; k < 9; ++k) A[k] = L[k];
      for (int k = 0; k < 3; ++k) {
        int pr = k;
        double mx = fabs(A[k * 3 + k]);
        for (int r = k + 1; r < 3; ++r)
          if (fabs(A[r * 3 + k]) > mx) {
            mx = fabs(A[r * 3 + k]);
            pr = r;
          }
        if (pr != k) {
          for (int j = 0; j < 3; ++j) {
            double t = A[k * 3 + j];
            A[k * 3 + j] = A[pr * 3 + j];
            A[pr * 3 + j] = t;
          }
          int t = piv[k];
          piv[k] = piv[pr];
          piv[pr] = t;
          double u = r1[k];
          r1[k] = r1[pr];
          r1[pr] = u;
          u = r2[k];
          r2[k] = r2[pr];
          r2[pr] = u;
        }
        for (int r = k + 1; r < 3; ++r) {
          const double l = A[r * 3 + k] / A[k * 3 + k];
          A[r * 3 + k] = l;
          for (int j = k + 1; j < 3; ++j) A[r * 3 + j] -= l * A[k * 3 + j];
          r1[r] -= l * r1[k];
          r2[r] -= l * r2[k];
        }
      }
      for (int k = 2; k >= 0; --k) {
        double a1 = r1[k], a2 = r2[k];
        for (int j = k + 1; j < 3; ++j) {
          a1 -= A[k * 3 + j] * r1[j];
          a2 -= A[k * 3 + j] * r2[j];
        }
        r1[k] = a1 / A[k * 3 + k];
        r2[k] = a2 / A[k * 3 + k];
      }
      dsw->Rs[0] = r1[0];
      dsw->Rs[1] = r1[1];
      dsw->Rs[2] = r1[2];
      dsw->Rs[3] = r2[0];
      dsw->Rs[4] = r2[1];
      dsw->Rs[5] = r2[2];
    });
  }

  // ==========================================================================
  // Step 3 (Solvers.jl:242-259)
  // ==========================================================================
  // Step 1's split solutions where they are used.  phi_1j, mu_1j (j = 11, 12, 13, 14;
  // Solvers.jl:86-115) are pointwise in the state spectra phi_hat^n, phi_hat^{n-1} and in
  // H1_hat, H2_hat, H3_hat, and nu_1j = S4 psi_1j (+ G_hat for nu_12; Solvers.jl:92, 107, 134),
  // so they are formed in registers by the passes that read them (steps 2, 5 and 6) instead of
  // stored, and so are phi_21 = -rfft(u* . grad phi*) / lhs_phi and mu_21 = L_phi phi_21
  // (Solvers.jl:117-131); psi_11, psi_12, psi_21 (the psi solves) are stored in the cache.
  // comb() forms the combinations of step 3 (Solvers.jl:242-259): x_1 = x_11 + R11 x_12 +
  // R21 x_13 + R31 x_14 for phi and mu (x_11 + R31 x_12 for psi and nu) and x_2 alike with
  // x_21, R12, R22, R32.
  struct Splits {
    const C* k;                // cache
    const C *xN, *xO, *Hh;     // phi_hat^n, phi_hat^{n-1}, [H1 | G | H2 | H3]_hat
    const C* ah;               // rfft(u* . grad phi*) (step 1's advection term)
    const double *kx, *ky, *kz;
    double S1, S2, S3, S4, Mphi;
    const DScal* ds;
    KIdx ki;
    long long NH;
    __device__ __forceinline__ CD at(int f, long long i) const { return k[f * NH + i]; }
    // ph[j], mu[j] = phi_1(j+1), mu_1(j+1) for j < 4; ph[4], mu[4] = phi_21, mu_21
    __device__ __forceinline__ void phi_mu(long long i, CD (&ph)[5], CD (&mu)[5]) const {
      int ix, iy, iz, f;
      ki(i, ix, iy, iz, f);
      const double kxv = kx[ix], kyv = ky[iy], kzv = kz[iz];
      const double k2 = kxv * kxv + kyv * kyv + kzv * kzv;
      const double lapd = -k2, bihd = k2 * k2;
      const double Lphi = S1 * bihd - S2 * lapd + S3;
      const double lhs = (ds->a / ds->dt) + Mphi * Lphi;
      const T Lp = (T)Lphi, lh = (T)lhs;
      const T b = (T)ds->b, cc = (T)ds->c, dt = (T)ds->dt;
      const CD a = xN[i], o = xO[i];
      const CD t = CD{-(b * a.x + cc * o.x) / dt, -(b * a.y + cc * o.y) / dt};
      ph[0] = CD{t.x / lh, t.y / lh};
      mu[0] = CD{Lp * ph[0].x, Lp * ph[0].y};
      const T mM = -(T)Mphi;
      const int hs[3] = {0, 2, 3};
      for (int j = 0; j < 3; ++j) {
        const CD h = Hh[hs[j] * NH + i];
        ph[j + 1] = CD{(mM * h.x) / lh, (mM * h.y) / lh};
        mu[j + 1] = CD{Lp * ph[j + 1].x + h.x, Lp * ph[j + 1].y + h.y};
      }
      const CD t2 = ah[i];  // Solvers.jl:117-131
      ph[4] = CD{(-t2.x) / lh, (-t2.y) / lh};
      mu[4] = CD{Lp * ph[4].x, Lp * ph[4].y};
    }
    __device__ __forceinline__ void psi_nu(long long i, CD (&ps)[3], CD (&nu)[3]) const {  // 11, 12, 21
      ps[0] = at(PSI11, i);
      ps[1] = at(PSI12, i);
      ps[2] = at(PSI21, i);
      const CD g = Hh[NH + i];
      const T s4 = (T)S4;
      nu[0] = CD{s4 * ps[0].x, s4 * ps[0].y};
      nu[1] = CD{s4 * ps[1].x + g.x, s4 * ps[1].y + g.y};
      nu[2] = CD{s4 * ps[2].x, s4 * ps[2].y};
    }
    static __device__ __forceinline__ CD ax(const CD& a, T r, const CD& b) { return CD{a.x + r * b.x, a.y + r * b.y}; }
    // x1[which], x2[which] for which = 0 phi, 1 mu, 2 psi, 3 nu (phi and psi only with want_phi_psi)
    __device__ __forceinline__ void comb(long long i, bool want_phi_psi, CD (&x1)[4], CD (&x2)[4]) const {
      const double* Rs = ds->Rs;
      CD ph[5], mu[5], ps[3], nu[3];
      phi_mu(i, ph, mu);
      psi_nu(i, ps, nu);
      for (int part = 0; part < 2; ++part) {
        const T Ra = (T)Rs[3 * part + 0], Rb = (T)Rs[3 * part + 1], Rc = (T)Rs[3 * part + 2];
        CD(&o)[4] = part ? x2 : x1;
        o[1] = ax(ax(ax(part ? mu[4] : mu[0], Ra, mu[1]), Rb, mu[2]), Rc, mu[3]);
        o[3] = ax(part ? nu[2] : nu[0], Rc, nu[1]);
        if (want_phi_psi) {
          o[0] = ax(ax(ax(part ? ph[4] : ph[0], Ra, ph[1]), Rb, ph[2]), Rc, ph[3]);
          o[2] = ax(part ? ps[2] : ps[0], Rc, ps[1]);
        }
      }
    }
  };
  Splits splits() const {
    return Splits{cache, shalf[ip], shalf[io], w_Hh, w_advh, d_kx, d_ky, d_kz, c.S1, c.S2, c.S3, c.S4, c.M_phi, ds,
                  KIdx{Nxh, Ny, Nz}, (long long)N * H};
  }

  // ==========================================================================
  // Step 4 (Solvers.jl:267-337; Solvers_3d.jl:221-292)
  // ==========================================================================
  void step4() {
    const long long NP = (long long)N * P;
    const C *uhN = uhalf(ip), *uhO = uhalf(io);
    // du_c/dx_d -> w_du[c*nd + d]
    {
      std::vector<SpecJob> jl;
      for (int cc = 0; cc < nd; ++cc)
        for (int d = 0; d < nd; ++d) jl.push_back(job(uhN + cc * H, w_du + (cc * nd + d) * P, 1, DX(d), uhO + cc * H));
      for (int d = 0; d < nd; ++d)
        jl.push_back(job(fhalf(ip, F_MU), w_dmn + d * NP, N, DX(d), fhalf(io, F_MU)));
      for (int d = 0; d < nd; ++d)
        jl.push_back(job(fhalf(ip, F_NU), w_dmn + (nd + d) * NP, N, DX(d), fhalf(io, F_NU)));
      irfft_v(jl);
    }
    const R *us = w_us, *du = w_du, *dmn = w_dmn, *phs = w_phis, *pss = w_psis;
    R* rhs2 = w_rhs2;
    const int ndd = nd, Nl = N;
    const long long Pl = P;
    const T lam = (T)c.lamda;
    // rhs2 = -(u* . grad) u* - lamda sum_n (phi* grad mu* + psi* grad nu*); step 5 needs only its
    // spectrum, so the convection and coupling fields are not kept
    foreach_n(s, P, [=] __device__(long long i) {
      R conv[3];
      for (int cc = 0; cc < ndd; ++cc) {
        R a = us[i] * du[(cc * ndd) * Pl + i];
        for (int d = 1; d < ndd; ++d) a = a + us[d * Pl + i] * du[(cc * ndd + d) * Pl + i];
        conv[cc] = a;
      }
      for (int d = 0; d < ndd; ++d) {
        R sum = T(0);
        for (int n = 0; n < Nl; ++n) {
          const long long k = (long long)n * Pl + i;
          const T t = phs[k] * dmn[d * NP + k] + pss[k] * dmn[(ndd + d) * NP + k];
          sum = n == 0 ? t : sum + t;
        }
        rhs2[d * Pl + i] = -conv[d] - lam * sum;
      }
    });
    rfft({job(w_rhs2, w_rhs2h, nd)});
    const C* r2h = w_rhs2h;
    const C* phat = phalf(ip);
    C* uth = w_uth;  // [ut1h (nd) | ut2h (nd)]
    const SpecC sc = spc();
    const KIdx ki{Nxh, Ny, Nz};
    const DScal* dsl = ds;
    const long long Hl = H;
    foreach_n(s, (long long)nd * H, [=] __device__(long long i) {
      int ix, iy, iz, d;
      ki(i, ix, iy, iz, d);
      const long long sp = i % Hl;
      const double kk[3] = {sc.kx[ix], sc.ky[iy], sc.kz[iz]};
      const double k2 = kk[0] * kk[0] + kk[1] * kk[1] + kk[2] * kk[2];
      const T lh = (T)((dsl->a / dsl->dt) - sc.eta * (-k2));
      const T b = (T)dsl->b, cc = (T)dsl->c, dt = (T)dsl->dt;
      const CD xN = uhN[i], xO = uhO[i];
      const CD pp = phat[sp];
      const T kd = (T)kk[d];
      const CD dp = CD{-kd * pp.y, kd * pp.x};
      const CD r1 = CD{-(b * xN.x + cc * xO.x) / dt - dp.x, -(b * xN.y + cc * xO.y) / dt - dp.y};
      uth[i] = CD{r1.x / lh, r1.y / lh};
      const CD r2 = r2h[i];
      uth[ndd * Hl + i] = CD{r2.x / lh, r2.y / lh};
    });
  }

  // ==========================================================================
  // Step 5 (Solvers.jl:344-423; Solvers_3d.jl:299-392)
  // ==========================================================================
  void step5() {
    const long long NH = (long long)N * H;
    // The theta sums (Solvers.jl:344-423) by Parseval over the half spectra; the reference
    // transforms mu_j, nu_j and u~_j back and sums in physical space.  w_advh holds the rffts of
    // the two advection fields u* . grad phi*, u* . grad psi* (step 1), the step-3 combinations
    // give the spectra mu_1, nu_1, mu_2, nu_2, w_rhs2h is the rfft of step 4's rhs2 =
    // -(u* . grad) u* - lamda sum_n (phi* grad mu* + psi* grad nu*) and w_uth the spectra of
    // u~_1 and u~_2, so lamda sum (coupling . u~_j) + sum (convection . u~_j) = -sum (rhs2 . u~_j).
    const C* ah = w_advh;
    const Splits sp = splits();
    const C* r2h = w_rhs2h;
    const C* uth = w_uth;
    const int Nxh_ = Nxh, Nx_ = Nx, ndd = nd;
    const long long Hl = H;
    reduce_n<6>(s, NH, 1,
                [=] __device__(long long i, int, double(&acc)[6]) {
                  const int ix = (int)(i % Nxh_);
                  const double w = (ix == 0 || 2 * ix == Nx_) ? 1.0 : 2.0;
                  auto re = [&](const CD& a, const CD& b) { return w * ((double)a.x * (double)b.x + (double)a.y * (double)b.y); };
                  const CD ap = ah[i], as = ah[NH + i];
                  CD x1[4], x2[4];
                  sp.comb(i, false, x1, x2);
                  acc[0] += re(ap, x1[1]);  // t11: mu_1
                  acc[1] += re(as, x1[3]);  // t12: nu_1
                  acc[2] += re(ap, x2[1]);  // t21: mu_2
                  acc[3] += re(as, x2[3]);  // t22: nu_2
                  if (i < Hl)
                    for (int d = 0; d < ndd; ++d) {
                      const CD r = r2h[d * Hl + i];
                      acc[4] += re(r, uth[d * Hl + i]);
                      acc[5] += re(r, uth[(ndd + d) * Hl + i]);
                    }
                },
                partials, red);
    const double invP = 1.0 / (double)P;
    const PhysC q = pc();
    const double* rd = red;
    DScal* dsw = ds;
    const DStateScal* stN = dstate + ip;
    const DStateScal* stO = dstate + io;
    const double lam = (double)c.lamda;
    single(s, [=] __device__() {
      double dvol = q.dx * q.dy;
      if (ndd == 3) dvol = dvol * q.dz;
      // spectral sums: x 1/P
      const double th1 = dvol * (lam * (rd[0] * invP + rd[1] * invP) - rd[4] * invP);
      const double th2 = dvol * (lam * (rd[2] * invP + rd[3] * invP) - rd[5] * invP);
      const double a = dsw->a, b = dsw->b, cc = dsw->c, dt = dsw->dt5;
      const double rhsQ = -(b * stN->Q + cc * stO->Q) / dt + th1;
      const double lhsQ = (a / dt) - th2;
      dsw->theta1 = th1;
      dsw->theta2 = th2;
      dsw->Q = rhsQ / lhsQ;
      if (ndd == 3 && !(th2 < 0)) dsw->flags |= 8;
    });
  }

  // ==========================================================================
  // Steps 6 and 7 (Solvers.jl:431-498) written into the next state buffer
  // ==========================================================================
  void step67() {
    const Splits sp = splits();
    C *phN = fhalf(inew, F_PHI), *psN = fhalf(inew, F_PSI), *muN = fhalf(inew, F_MU), *nuN = fhalf(inew, F_NU);
    C* uN = uhalf(inew);
    C* pN = phalf(inew);
    const C* pOld = phalf(ip);
    const C* uth = w_uth;
    const SpecC sc = spc();
    const KIdx ki{Nxh, Ny, Nz};
    const DScal* dsl = ds;
    const int ndd = nd, Nl = N;
    const long long Hl = H;
    foreach_n(s, H, [=] __device__(long long i) {
      const T Q = (T)dsl->Q;
      for (int n = 0; n < Nl; ++n) {
        const long long k = (long long)n * Hl + i;
        CD x1[4], x2[4];
        sp.comb(k, true, x1, x2);
        auto cmb = [&](int f) {  // x^{n+1} = x_1 + Q x_2 (Solvers.jl:431-453)
          return CD{x1[f].x + Q * x2[f].x, x1[f].y + Q * x2[f].y};
        };
        phN[k] = cmb(0);
        muN[k] = cmb(1);
        psN[k] = cmb(2);
        nuN[k] = cmb(3);
      }
      int ix, iy, iz, f;
      ki(i, ix, iy, iz, f);
      const double kk[3] = {sc.kx[ix], sc.ky[iy], sc.kz[iz]};
      C ut[3];
      CD div = CD{T(0), T(0)};
      for (int d = 0; d < ndd; ++d) {
        const CD a = uth[d * Hl + i], b = uth[(ndd + d) * Hl + i];
        ut[d] = CD{a.x + Q * b.x, a.y + Q * b.y};
        const T kd = (T)kk[d];
        const CD t = CD{-kd * ut[d].y, kd * ut[d].x};
        div = d == 0 ? t : div + t;
      }
      const double k2 = kk[0] * kk[0] + kk[1] * kk[1] + kk[2] * kk[2];
      const T lapv = (T)(-k2);
      const T fac = (T)(dsl->a / dsl->dt);
      CD dp = CD{(fac * div.x) / lapv, (fac * div.y) / lapv};
      if (i == 0) dp = CD{T(0), T(0)};
      const double dta = dsl->dt / dsl->a;
      for (int d = 0; d < ndd; ++d) {
        const T kd = (T)(dta * kk[d]);
        uN[d * Hl + i] = CD{ut[d].x - (-kd * dp.y), ut[d].y - kd * dp.x};
      }
      const T eta = (T)sc.eta;
      const CD po = pOld[i];
      pN[i] = CD{(po.x + dp.x) - eta * div.x, (po.y + dp.y) - eta * div.y};
    });
    DStateScal* stw = dstate + inew;
    DScal* dsw = ds;
    single(s, [=] __device__() {
      const double Q = dsw->Q;
      stw->Q = Q;
      stw->R1 = dsw->Rs[0] + Q * dsw->Rs[3];
      stw->R2 = dsw->Rs[1] + Q * dsw->Rs[4];
      stw->R3 = dsw->Rs[2] + Q * dsw->Rs[5];
      if (!isfinite(Q) || !isfinite(stw->R1) || !isfinite(stw->R2) || !isfinite(stw->R3)) dsw->flags |= 16;
    });
  }

  // ==========================================================================
  // Monitor: compute_modified_energy (Utils.jl:96-141) and the area ratio
  // (simulation.jl:147) for the present/old pair.
  // ==========================================================================
  void monitor() {
    const C *phh = fhalf(ip, F_PHI), *phhO = fhalf(io, F_PHI), *ph = phalf(ip);
    const bool s2 = c.S2 != 0.0;
    // area: f_surf of rfft(phi) (calculate_area); phi is the c2r of phi_hat, so rfft(phi) is the
    // Hermitian part of phi_hat on the kx = 0 and kx = Nx/2 planes, phi_hat elsewhere
    const R* phi = fphys(ip, F_PHI);
    const PhysC q = pc();
    const R *u = uphys(ip), *uo = uphys(io), *phO = fphys(io, F_PHI), *ps = fphys(ip, F_PSI), *psO = fphys(io, F_PSI);
    const long long Pl = P, Hl = H;
    const int ndl = nd;
    // physical sums, group n = vesicle n: the g(phi) part of f_surf, phi^2, (2 phi - phi_old)^2,
    // psi^2, (2 psi - psi_old)^2; group 0 also |u|^2 and |2u - u_old|^2 -> red[n*7 + k]
    double* redP = red;
    reduce_n<7>(s, P, N,
                [=] __device__(long long i, int n, double(&acc)[7]) {
                  const long long k = (long long)n * Pl + i;
                  const T p = phi[k];
                  const T eps = (T)q.eps;
                  const T gg = T(0.25) * (p * p - T(1)) * (p * p - T(1));
                  acc[0] += (double)((T)(3.0 * SQ2 / 4.0) * (gg / eps));
                  const double pd = p, pst = (double)(T(2) * phi[k] - phO[k]);
                  const double y = ps[k], yst = (double)(T(2) * ps[k] - psO[k]);
                  acc[1] += pd * pd;
                  acc[2] += pst * pst;
                  acc[3] += y * y;
                  acc[4] += yst * yst;
                  if (n == 0)
                    for (int d = 0; d < ndl; ++d) {
                      const double a = u[d * Pl + i];
                      const double st = (double)(T(2) * u[d * Pl + i] - uo[d * Pl + i]);
                      acc[5] += a * a;
                      acc[6] += st * st;
                    }
                },
                partials, redP);
    // the squared norms of derivative fields by Parseval over the half spectra (the reference
    // transforms each back): sum_x c2r(X)^2 = (1/P) [sum off the kx = 0, Nx/2 planes of 2 |X|^2
    // + sum on them of |(X(k) + conj X(-k)) / 2|^2], the part of X that a c2r keeps.  Group n:
    // |grad phi_n|^2 of rfft(phi) (x and y, as f_surf), (Lap phi)^2, (Lap phi*)^2, and with S2
    // (dx phi)^2, (dy phi)^2, (dx phi*)^2, (dy phi*)^2 (phi* = 2 phi - phi_old); group 0 also
    // (dx p)^2, (dy p)^2 -> red[256 + n*9 + k]
    double* redS = red + 256;
    const SpecC sc = spc();
    const int Nx_ = Nx, Nxh_ = Nxh, Ny_ = Ny, Nz_ = Nz;
    const KIdx ki{Nxh, Ny, Nz};
    reduce_n<9>(s, H, N,
                [=] __device__(long long i, int n, double(&acc)[9]) {
                  int ix, iy, iz, f;
                  ki(i, ix, iy, iz, f);
                  const bool plane = ix == 0 || 2 * ix == Nx_;
                  int iy2 = iy, iz2 = iz;
                  long long im = i;
                  if (plane) {
                    iy2 = iy ? Ny_ - iy : 0;
                    iz2 = iz ? Nz_ - iz : 0;
                    im = ((long long)iz2 * Ny_ + iy2) * Nxh_ + ix;
                  }
                  const long long k = (long long)n * Hl + i, km = (long long)n * Hl + im;
                  const double kx = sc.kx[ix], ky = sc.ky[iy], kz = sc.kz[iz], kym = sc.ky[iy2];
                  const double l = -(kx * kx + ky * ky + kz * kz);  // -k^2 is even in k
                  auto sq = [&](const CD& x, const CD& xm) {
                    if (!plane) return 2.0 * (x.x * x.x + x.y * x.y);
                    const double re = 0.5 * (x.x + xm.x), imv = 0.5 * (x.y - xm.y);
                    return re * re + imv * imv;
                  };
                  auto dk = [](double kk, const CD& a) { return CD{-kk * a.y, kk * a.x}; };
                  auto sc2 = [](double v, const CD& a) { return CD{v * a.x, v * a.y}; };
                  const CD x = phh[k], xo = phhO[k], xm = phh[km], xom = phhO[km];
                  CD a = x, am = xm;  // rfft(phi) at k and -k
                  if (plane) {
                    a = CD{0.5 * (x.x + xm.x), 0.5 * (x.y - xm.y)};
                    am = CD{a.x, -a.y};
                  }
                  acc[0] += sq(dk(kx, a), dk(kx, am)) + sq(dk(ky, a), dk(kym, am));
                  const CD st = CD{2.0 * x.x - xo.x, 2.0 * x.y - xo.y};
                  const CD stm = CD{2.0 * xm.x - xom.x, 2.0 * xm.y - xom.y};
                  acc[1] += sq(sc2(l, x), sc2(l, xm));
                  acc[2] += sq(sc2(l, st), sc2(l, stm));
                  if (s2) {
                    acc[3] += sq(dk(kx, x), dk(kx, xm));
                    acc[4] += sq(dk(ky, x), dk(kym, xm));
                    acc[5] += sq(dk(kx, st), dk(kx, stm));
                    acc[6] += sq(dk(ky, st), dk(kym, stm));
                  }
                  if (n == 0) {
                    const CD pp = ph[i], ppm = ph[im];
                    acc[7] += sq(dk(kx, pp), dk(kx, ppm));
                    acc[8] += sq(dk(ky, pp), dk(kym, ppm));
                  }
                },
                partials, redS);
    const double invP = 1.0 / (double)P;
    const double* rd = red;
    DScal* dsw = ds;
    const DStateScal* stN = dstate + ip;
    const DStateScal* stO = dstate + io;
    const double lam = (double)c.lamda;
    const double* A0 = dA0;
    const int Nl = N;
    single(s, [=] __device__() {
      auto sumP = [&](int k) {
        double t = rd[k];
        for (int n = 1; n < Nl; ++n) t += rd[n * 7 + k];
        return integ(t, q, ndl);
      };
      auto sumS = [&](int k) {
        double t = rd[256 + k];
        for (int n = 1; n < Nl; ++n) t += rd[256 + n * 9 + k];
        return integ(t * invP, q, ndl);
      };
      const double kinetic = 0.25 * integ(rd[5], q, ndl) + 0.25 * integ(rd[6], q, ndl);
      const double dt = dsw->dt;
      const double pressure = (dt * dt / 3.0) * (integ(rd[256 + 7] * invP, q, ndl) + integ(rd[256 + 8] * invP, q, ndl));
      auto se = [](double rn, double ro) { return 0.5 * rn * rn + 0.5 * (2 * rn - ro) * (2 * rn - ro); };
      const double Re = se(stN->R1, stO->R1) + se(stN->R2, stO->R2) + se(stN->R3, stO->R3);
      const double Qe = 0.5 * se(stN->Q, stO->Q);
      const double S1t = 0.25 * q.S1 * (sumS(1) + sumS(2));
      const double S2t = 0.25 * q.S2 * (sumS(3) + sumS(4) + sumS(5) + sumS(6));
      const double S3t = 0.25 * q.S3 * (sumP(1) + sumP(2));
      const double S4t = 0.25 * q.S4 * (sumP(3) + sumP(4));
      dsw->energy = (kinetic + pressure + lam * Re + Qe + lam * (S1t + S2t + S3t + S4t) -
                     lam * (q.C1 + q.C2 + q.C3) - 0.5);
      double num = 0.0, den = 0.0;
      for (int n = 0; n < Nl; ++n) {
        const double A = integ(rd[n * 7] + (0.5 * q.eps) * (rd[256 + n * 9] * invP), q, ndl);
        dsw->area_mon[n] = A;
        num += fabs(A - A0[n]);
        den += A0[n];
      }
      dsw->area_ratio = num / den;
    });
  }

  // ==========================================================================
  // ISolver
  // ==========================================================================
  void set_state(int which, const double* const* a, const double* scal, const double* A0, double dt) override {
    const int b = which == SAV_PRESENT ? ip : io;
    const long long NP = (long long)N * P, NH = (long long)N * H;
    lazy_phys[b] = false;
    // order: phi, phi_hat, psi, psi_hat, u, u_hat, p, p_hat, mu, mu_hat, nu, nu_hat
    put_real(a[0], fphys(b, F_PHI), NP);
    put_cplx(a[1], fhalf(b, F_PHI), NH);
    put_real(a[2], fphys(b, F_PSI), NP);
    put_cplx(a[3], fhalf(b, F_PSI), NH);
    put_real(a[4], uphys(b), nd * P);
    put_cplx(a[5], uhalf(b), nd * H);
    put_real(a[6], pphys(b), P);
    put_cplx(a[7], phalf(b), H);
    put_real(a[8], fphys(b, F_MU), NP);
    put_cplx(a[9], fhalf(b, F_MU), NH);
    put_real(a[10], fphys(b, F_NU), NP);
    put_cplx(a[11], fhalf(b, F_NU), NH);
    DStateScal st{scal[0], scal[1], scal[2], scal[3]};
    SAV_HIP(hipMemcpyAsync(dstate + b, &st, sizeof(st), hipMemcpyHostToDevice, s));
    if (which == SAV_PRESENT && A0) {
      SAV_HIP(hipMemcpyAsync(dA0, A0, sizeof(double) * N, hipMemcpyHostToDevice, s));
      present_dt = dt;
    }
    SAV_HIP(hipStreamSynchronize(s));
  }
  double present_dt = 0.0;

  void get_state(int which, double* const* a, double* scal, double* A0, double* dt) override {
    const int b = which == SAV_PRESENT ? ip : io;
    const long long NP = (long long)N * P, NH = (long long)N * H;
    materialize(b);
    get_real(a[0], fphys(b, F_PHI), NP);
    get_cplx(a[1], fhalf(b, F_PHI), NH);
    get_real(a[2], fphys(b, F_PSI), NP);
    get_cplx(a[3], fhalf(b, F_PSI), NH);
    get_real(a[4], uphys(b), nd * P);
    get_cplx(a[5], uhalf(b), nd * H);
    get_real(a[6], pphys(b), P);
    get_cplx(a[7], phalf(b), H);
    get_real(a[8], fphys(b, F_MU), NP);
    get_cplx(a[9], fhalf(b, F_MU), NH);
    get_real(a[10], fphys(b, F_NU), NP);
    get_cplx(a[11], fhalf(b, F_NU), NH);
    if (scal) {
      DStateScal st;
      SAV_HIP(hipMemcpyAsync(&st, dstate + b, sizeof(st), hipMemcpyDeviceToHost, s));
      SAV_HIP(hipStreamSynchronize(s));
      scal[0] = st.Q;
      scal[1] = st.R1;
      scal[2] = st.R2;
      scal[3] = st.R3;
    }
    if (A0) {
      SAV_HIP(hipMemcpyAsync(A0, dA0, sizeof(double) * N, hipMemcpyDeviceToHost, s));
      SAV_HIP(hipStreamSynchronize(s));
    }
    if (dt) *dt = present_dt;
  }

  // host <-> device with precision conversion
  void put_real(const double* h, R* d, long long n) {
    if (!h) return;
    if constexpr (sizeof(R) == 8) {
      SAV_HIP(hipMemcpyAsync(d, h, n * sizeof(double), hipMemcpyHostToDevice, s));
      SAV_HIP(hipStreamSynchronize(s));
    } else {
      std::vector<R> t(n);
      for (long long i = 0; i < n; ++i) t[i] = (R)h[i];
      SAV_HIP(hipMemcpyAsync(d, t.data(), n * sizeof(R), hipMemcpyHostToDevice, s));
      SAV_HIP(hipStreamSynchronize(s));
    }
  }
  void put_cplx(const double* h, C* d, long long n) { put_real(h, reinterpret_cast<R*>(d), 2 * n); }
  void get_real(double* h, const R* d, long long n) {
    if (!h) return;
    if constexpr (sizeof(R) == 8) {
      SAV_HIP(hipMemcpyAsync(h, d, n * sizeof(double), hipMemcpyDeviceToHost, s));
      SAV_HIP(hipStreamSynchronize(s));
    } else {
      std::vector<R> t(n);
      SAV_HIP(hipMemcpyAsync(t.data(), d, n * sizeof(R), hipMemcpyDeviceToHost, s));
      SAV_HIP(hipStreamSynchronize(s));
      for (long long i = 0; i < n; ++i) h[i] = (double)t[i];
    }
  }
  void get_cplx(double* h, const C* d, long long n) { get_real(h, reinterpret_cast<const R*>(d), 2 * n); }

  int check_flags() {
    SAV_HIP(hipMemcpyAsync(hs, ds, sizeof(DScal), hipMemcpyDeviceToHost, s));
    SAV_HIP(hipStreamSynchronize(s));
    if (hs->flags & 7) return SAV_E_WC_NEGATIVE;
    return 0;
  }

  int init_sav() override {
    const long long NH = (long long)N * H, NP = (long long)N * P;
    materialize(ip);  // old := present copies the physical fields too
    set_bdf(1.0, -1.0, 0.0, present_dt > 0 ? present_dt : c.dt);
    functionals(fphys(ip, F_PHI), fphys(ip, F_PSI), true, sizeof(R) == 8 ? nullptr : fhalf(ip, F_PHI));
    int rc = check_flags();
    if (rc) return rc;
    // mu_hat = L_phi phi_hat + R1 H1^ + R2 H2^ + R3 H3^ ; nu_hat = S4 psi_hat + R3 G^ (simulation.jl:35-47)
    const C* Hh = w_Hh;
    C* muh = fhalf(ip, F_MU);
    C* nuh = fhalf(ip, F_NU);
    const C* phh = fhalf(ip, F_PHI);
    const C* psh = fhalf(ip, F_PSI);
    const SpecC sc = spc();
    const KIdx ki{Nxh, Ny, Nz};
    const DScal* dsl = ds;
    foreach_n(s, NH, [=] __device__(long long i) {
      int ix, iy, iz, f;
      ki(i, ix, iy, iz, f);
      const double kx = sc.kx[ix], ky = sc.ky[iy], kz = sc.kz[iz];
      const double k2 = kx * kx + ky * ky + kz * kz;
      const T Lp = (T)(sc.S1 * (k2 * k2) - sc.S2 * (-k2) + sc.S3);
      const T r1 = (T)dsl->R1s, r2 = (T)dsl->R2s, r3 = (T)dsl->R3s;
      const CD p = phh[i], h1 = Hh[i], g = Hh[NH + i], h2 = Hh[2 * NH + i], h3 = Hh[3 * NH + i];
      muh[i] = CD{Lp * p.x + r1 * h1.x + r2 * h2.x + r3 * h3.x, Lp * p.y + r1 * h1.y + r2 * h2.y + r3 * h3.y};
      const CD y = psh[i];
      const T S4 = (T)sc.S4;
      nuh[i] = CD{S4 * y.x + r3 * g.x, S4 * y.y + r3 * g.y};
    });
    irfft({job(muh, fphys(ip, F_MU), N), job(nuh, fphys(ip, F_NU), N)});
    DStateScal* stw = dstate + ip;
    DScal* dsw = ds;
    single(s, [=] __device__() {
      stw->Q = 1.0;
      stw->R1 = dsw->R1s;
      stw->R2 = dsw->R2s;
      stw->R3 = dsw->R3s;
    });
    // old = deepcopy(present)
    const long long nphys = (4LL * N + nd + 1) * P, nhalf = (4LL * N + nd + 1) * H;
    SAV_HIP(hipMemcpyAsync(sphys[io], sphys[ip], nphys * sizeof(R), hipMemcpyDeviceToDevice, s));
    SAV_HIP(hipMemcpyAsync(shalf[io], shalf[ip], nhalf * sizeof(C), hipMemcpyDeviceToDevice, s));
    SAV_HIP(hipMemcpyAsync(dstate + io, dstate + ip, sizeof(DStateScal), hipMemcpyDeviceToDevice, s));
    SAV_HIP(hipStreamSynchronize(s));
    lazy_phys[io] = false;
    (void)NP;
    return 0;
  }

  void rotate() {  // old <- present, present <- new (update_state!, Types.jl:243-265)
    const int oldp = ip, oldo = io;
    ip = inew;
    io = oldp;
    inew = oldo;
  }
  void unrotate() {  // undoes rotate() of a step that is not committed
    const int oldp = io, oldo = inew;
    inew = ip;
    ip = oldp;
    io = oldo;
  }
  // After rotate(): the back-transform of the new state (simulation.jl:99-104) and the
  // energy/area monitor (simulation.jl:146-148).
  // phi, psi and u, which the next step and the monitor read, are transformed here; mu, nu and p
  // (read only by the host) when the state is read (materialize).
  void step_tail(bool monitor_on) {
    irfft({job(shalf[ip], sphys[ip], 2 * N), job(uhalf(ip), uphys(ip), nd)});
    if (monitor_on) monitor();
  }
  // lazy_phys[b]: the physical mu, nu and p of buffer b are not yet back-transformed from its spectra
  bool lazy_phys[3] = {false, false, false};
  void materialize(int b) {
    if (!lazy_phys[b]) return;
    irfft({job(fhalf(b, F_MU), fphys(b, F_MU), 2 * N), job(phalf(b), pphys(b), 1)});
    lazy_phys[b] = false;
  }

  // Step statistics and status from hs, hks, hst (copied from the device by the caller).
  int finish_step(sav_step_stats_t* st) {
    int rc = 0;
    bool notconv = false;
    if (st) {
      std::memset(st, 0, sizeof(*st));
      for (int k = 0; k < 3; ++k) {
        st->krylov_iters[k] = nd == 2 ? kit[k] : 0;  // all iterations spent (a fallback solve included)
        st->krylov_solved[k] = nd == 2 ? hks[k].solved : 1;
        st->krylov_rnorm[k] = nd == 2 ? hks[k].rnorm : 0.0;
      }
      st->energy = hs->energy;
      st->area_ratio = hs->area_ratio;
      st->Q = hs->Q;
      st->R1 = hst->R1;
      st->R2 = hst->R2;
      st->R3 = hst->R3;
    }
    if (nd == 2)
      for (int k = 0; k < 3; ++k)
        if (!hks[k].solved) notconv = true;
    if (notconv) rc = SAV_W_KRYLOV_NOT_CONVERGED;
    else if (hs->flags & 8) rc = SAV_W_THETA2_NONNEG;
    if (st) st->status = rc;
    return rc;
  }

  // ==========================================================================
  // Graph replay of whole steps (no reference counterpart).  A step (steps 1-7, the
  // back-transform of the new state and the monitor) is captured once per buffer
  // rotation, psi iteration budget K and monitor flag into a HIP graph and replayed: one
  // graph launch and one host synchronisation per step instead of ~160 kernel launches
  // and a host check per batch of Krylov iterations.  The psi solve's iteration count is
  // data dependent: the graph queues K = (largest count of the previous step) + 1
  // batched iterations and every Krylov kernel of a converged system returns at once.  If
  // a system still iterates after K < itmax iterations, the step is repeated eagerly: the
  // graph wrote only the next-state buffer and work arrays, present and old are intact.
  // Results are bit-identical to the eager step.  set_bdf runs outside the graph (its
  // coefficients and dt are kernel arguments); every kernel in the graph reads them from
  // the device scalars.
  // ==========================================================================
  struct StepGraph {
    int ip, io, K, mon;
    hipGraphExec_t exec;
  };
  std::vector<StepGraph> graphs;
  long long graph_replays = 0, graph_repeats = 0;
  // Replay mode: 0 off, 1 on, 2 measured.  In mode 2, once steps can be replayed (the
  // handle's tuning done) and after its first two steps, the wall times of GAUTO_N eager
  // steps and of GAUTO_N replayed steps are taken (steps that capture a graph are not
  // counted), and replay is kept when its median step is more than 3 % shorter: launch-bound
  // small grids gain, large grids and forked 3D steps lose a little.
  static constexpr int GAUTO_N = 5;
  int graph_mode = 0;
  int gauto_phase = 0;  // 0 timing eager steps, 1 timing replayed steps, 2 decided
  int gauto_tries = 0;
  long long steps_run = 0;
  bool gauto_replay = false;
  std::vector<double> gauto_eager, gauto_graph;

  bool graph_base_usable() const {
    // a mixed-precision step decides its rounds on the host: eager only
    return kc.timed == 0 && group == nullptr &&
           (nd == 3 || (!mixed && !use_learned && !collecting && kc.colw != 0 && kc.order != 0));
  }
  bool graph_usable() const {
    const bool on = graph_mode == 1 || (graph_mode == 2 && (gauto_phase == 1 || (gauto_phase == 2 && gauto_replay)));
    return on && graph_base_usable();
  }
  static double median(std::vector<double> v) {
    std::sort(v.begin(), v.end());
    return v[v.size() / 2];
  }
  void gauto_record(double ms, bool captured) {
    if (gauto_phase == 0) {
      gauto_eager.push_back(ms);
      if ((int)gauto_eager.size() == GAUTO_N) gauto_phase = 1;
    } else if (gauto_phase == 1) {
      if (!captured) gauto_graph.push_back(ms);
      if ((int)gauto_graph.size() == GAUTO_N || ++gauto_tries >= 3 * GAUTO_N) {
        gauto_replay = !gauto_graph.empty() && median(gauto_graph) < 0.97 * median(gauto_eager);
        gauto_phase = 2;
        if (!gauto_replay) drop_graphs();
      }
    }
  }
  void drop_graphs() {
    if (!graphs.empty()) SAV_HIP(hipStreamSynchronize(s));
    for (auto& gr : graphs) (void)hipGraphExecDestroy(gr.exec);
    graphs.clear();
  }

  hipGraphExec_t capture_step(int K, bool monitor_on) {
    const int sip = ip, sio = io, snew = inew;
    hipGraph_t gph = nullptr;
    SAV_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
      capture_iters = K > 0 ? K : 1;
      step1();
      step2();
      step4();
      step5();
      step67();
      rotate();
      step_tail(monitor_on);
    } catch (...) {
      capture_iters = 0;
      ip = sip;
      io = sio;
      inew = snew;
      (void)hipStreamEndCapture(s, &gph);
      if (gph) (void)hipGraphDestroy(gph);
      throw;
    }
    capture_iters = 0;
    ip = sip;
    io = sio;
    inew = snew;
    SAV_HIP(hipStreamEndCapture(s, &gph));
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, gph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gph);
    SAV_HIP(e);
    return ex;
  }

  // Replays the step's graph (set_bdf already queued).  Returns false when the step
  // must be repeated eagerly.
  bool step_graph(bool monitor_on, sav_step_stats_t* st, int& rc) {
    int K = 0;
    if (nd == 2) {
      for (int z = 0; z < 3; ++z) K = std::max(K, kc.guess[z]);
      K = std::min(K + 1, kc.itmax);
    }
    hipGraphExec_t ex = nullptr;
    for (auto& gr : graphs)
      if (gr.ip == ip && gr.io == io && gr.K == K && gr.mon == (int)monitor_on) ex = gr.exec;
    if (!ex) {
      ex = capture_step(K, monitor_on);
      graphs.push_back(StepGraph{ip, io, K, (int)monitor_on, ex});
    }
    SAV_HIP(hipGraphLaunch(ex, s));
    SAV_HIP(hipMemcpyAsync(hs, ds, sizeof(DScal), hipMemcpyDeviceToHost, s));
    if (nd == 2) SAV_HIP(hipMemcpyAsync(hks, kc.ks, 3 * sizeof(KScal), hipMemcpyDeviceToHost, s));
    SAV_HIP(hipMemcpyAsync(hst, dstate + inew, sizeof(DStateScal), hipMemcpyDeviceToHost, s));
    SAV_HIP(hipStreamSynchronize(s));
    ++graph_replays;
    // fatal W + C < 0 or non-finite: the step is not committed
    if (hs->flags & 7) {
      rc = SAV_E_WC_NEGATIVE;
      return true;
    }
    if (hs->flags & 16) {
      rc = SAV_E_NONFINITE;
      return true;
    }
    if (nd == 2 && K < kc.itmax)
      for (int z = 0; z < 3; ++z)
        if (!hks[z].done) {
          ++graph_repeats;
          return false;
        }
    if (nd == 2)
      for (int z = 0; z < 3; ++z) {
        kit[z] = hks[z].iter;
        kc.guess[z] = kit[z] > 0 ? kit[z] : 1;
      }
    rotate();
    lazy_phys[ip] = true;
    rc = finish_step(st);
    return true;
  }

  int set_graph(int on) override {
    if (on < 0 || on > 2) return SAV_E_ARG;
    graph_mode = on;
    gauto_phase = 0;
    gauto_tries = 0;
    gauto_replay = false;
    gauto_eager.clear();
    gauto_graph.clear();
    if (on != 1) drop_graphs();
    return 0;
  }
  void graph_info(long long* replays, long long* repeats, long long* ngraphs) override {
    if (replays) *replays = graph_replays;
    if (repeats) *repeats = graph_repeats;
    if (ngraphs) *ngraphs = (long long)graphs.size();
  }
  int graph_choice() const override {
    if (graph_mode != 2) return graph_mode;
    if (gauto_phase == 2) return gauto_replay ? 1 : 0;
    return graph_base_usable() ? -1 : 0;  // a form that never replays (mixed, learned) stays eager
  }

  int step_impl(double dt, int n, sav_step_stats_t* st, bool monitor_on) {
    ++steps_run;
    if (graph_mode != 2 || gauto_phase == 2 || steps_run <= 2 || !graph_base_usable())
      return step_body(dt, n, st, monitor_on);
    const size_t ng0 = graphs.size();
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = step_body(dt, n, st, monitor_on);
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (rc >= 0) gauto_record(ms, graphs.size() != ng0);
    return rc;
  }

  int step_body(double dt, int n, sav_step_stats_t* st, bool monitor_on) {
    autotune_transforms();
    const double a = n == 1 ? 1.0 : 1.5, b = n == 1 ? -1.0 : -2.0, cc = n == 1 ? 0.0 : 0.5;
    present_dt = dt;
    set_bdf(a, b, cc, dt);
    if (graph_usable()) {
      int rc = 0;
      if (step_graph(monitor_on, st, rc)) return rc;
      set_bdf(a, b, cc, dt);  // psi iteration budget exceeded: repeat the step eagerly
    }
    const bool forked = fork_enabled;
    if (forked) ensure_fork();
    fork_step4 = forked;
    try {
      step1();
    } catch (...) {
      fork_step4 = false;
      throw;
    }
    fork_step4 = false;
    step2();
    if (forked)
      SAV_HIP(hipStreamWaitEvent(s, ev_join, 0));
    else
      step4();
    step5();
    step67();
    // The back-transform and the monitor write only the new state's buffers, so they are
    // queued before the status is known (one host synchronisation per step); a fatal
    // W + C < 0 or non-finite step is then not committed: present and old are intact.
    rotate();
    step_tail(monitor_on);
    lazy_phys[ip] = true;  // (a step that is not committed leaves it on the discarded buffer)
    SAV_HIP(hipMemcpyAsync(hs, ds, sizeof(DScal), hipMemcpyDeviceToHost, s));
    SAV_HIP(hipMemcpyAsync(hst, dstate + ip, sizeof(DStateScal), hipMemcpyDeviceToHost, s));
    SAV_HIP(hipStreamSynchronize(s));
    if (hs->flags & 23) {
      unrotate();
      return (hs->flags & 7) ? SAV_E_WC_NEGATIVE : SAV_E_NONFINITE;
    }
    return finish_step(st);
  }

  int step(double dt, int n, sav_step_stats_t* st) override {
    SAV_HIP(hipEventRecord(ev0, s));
    int rc = step_impl(dt, n, st, true);
    SAV_HIP(hipEventRecord(ev1, s));
    SAV_HIP(hipEventSynchronize(ev1));
    SAV_HIP(hipEventElapsedTime(&last_ms, ev0, ev1));
    return rc;
  }

  int run(const double* t, int nn, int n0, double* eh, double* ah, int32_t* kit_total) override {
    int worst = 0;
    long long ktot = 0;
    SAV_HIP(hipEventRecord(ev0, s));
    for (int k = 0; k + 1 < nn; ++k) {
      sav_step_stats_t st;
      const int rc = step_impl(t[k + 1] - t[k], n0 + k, &st, true);
      if (rc < 0) return rc;
      if (rc > worst) worst = rc;
      if (eh) eh[k] = st.energy;
      if (ah) ah[k] = st.area_ratio;
      ktot += st.krylov_iters[0] + st.krylov_iters[1] + st.krylov_iters[2];
    }
    SAV_HIP(hipEventRecord(ev1, s));
    SAV_HIP(hipEventSynchronize(ev1));
    SAV_HIP(hipEventElapsedTime(&last_ms, ev0, ev1));
    if (kit_total) *kit_total = (int32_t)ktot;
    return worst;
  }

  double last_elapsed_ms() override { return last_ms; }

  // Event timing of one psi-matvec kernel over the following steps.
  std::vector<hipEvent_t> tev;
  void kernel_timing(int id) override {
    if (nd != 2) throw HipError("kernel timing covers the 2D psi matvec kernels", SAV_E_UNSUPPORTED);
    if (tev.empty()) {
      tev.resize(2 * 256);
      for (auto& e : tev) SAV_HIP(hipEventCreate(&e));
      kc.ev = tev.data();
      kc.ev_cap = 256;
    }
    kc.timed = id;
    kc.ev_used = 0;
    kc.tk_count = 0;
    kc.tk_ms = 0.0;
    kc.tk_bytes = 0.0;
    if (mixed) {  // the FP32 inner iterations carry the matvec kernels of a mixed solve
      kc32.ev = tev.data();
      kc32.ev_cap = 256;
      kc32.timed = id;
      kc32.ev_used = 0;
      kc32.tk_count = 0;
      kc32.tk_ms = 0.0;
      kc32.tk_bytes = 0.0;
      kc.timed = 0;
    }
  }
  void kernel_timing_get(long long* n, double* ms, double* bytes) override {
    const long long cnt = mixed ? kc32.tk_count : kc.tk_count;
    const double tms = mixed ? kc32.tk_ms : kc.tk_ms, tb = mixed ? kc32.tk_bytes : kc.tk_bytes;
    if (n) *n = cnt;
    if (ms) *ms = tms;
    if (bytes) *bytes = cnt ? tb / (double)cnt : 0.0;
  }

  void rfft_host(const double* in, double* out, int count) override {
    R* dphys = dalloc_tmp_phys(count);
    C* dhalf = dalloc_tmp_half(count);
    put_real(in, dphys, count * P);
    rfft({job(dphys, dhalf, count)});
    get_cplx(out, dhalf, count * H);
  }
  void irfft_host(const double* in, double* out, int count) override {
    R* dphys = dalloc_tmp_phys(count);
    C* dhalf = dalloc_tmp_half(count);
    put_cplx(in, dhalf, count * H);
    irfft({job(dhalf, dphys, count)});
    get_real(out, dphys, count * P);
  }
  void fft_full_host(const double* in, double* out, int count, int inverse) override {
    if (nd != 2) throw HipError("full complex transforms are 2D only", SAV_E_UNSUPPORTED);
    C* d = reinterpret_cast<C*>(dalloc_tmp_phys(4 * count));  // 2*count*P complex -> 4*count*P reals
    put_cplx(in, d, count * P);
    fft_full_2d<R>(g, d, count, inverse != 0, s);
    get_cplx(out, d, count * P);
  }
  R* tphys = nullptr;
  long long tphys_n = 0;
  C* thalf = nullptr;
  long long thalf_n = 0;
  R* dalloc_tmp_phys(int count) {
    if ((long long)count * P > tphys_n) {
      tphys_n = (long long)count * P;
      tphys = dalloc<R>(tphys_n);
    }
    return tphys;
  }
  C* dalloc_tmp_half(int count) {
    if ((long long)count * H > thalf_n) {
      thalf_n = (long long)count * H;
      thalf = dalloc<C>(thalf_n);
    }
    return thalf;
  }

  int eval_functionals(const double* phi, const double* psi, double* H1, double* H2, double* H3, double* G,
                       double* W, double* area) override {
    const long long NP = (long long)N * P;
    put_real(phi, w_phis, NP);
    put_real(psi, w_psis, NP);
    SAV_HIP(hipMemsetAsync(&ds->flags, 0, sizeof(int), s));
    functionals(w_phis, w_psis, false, nullptr, nullptr, true);
    int rc = check_flags();
    get_real(H1, w_H, NP);
    get_real(G, w_H + NP, NP);
    get_real(H2, w_H + 2 * NP, NP);
    get_real(H3, w_H + 3 * NP, NP);
    if (W) {
      W[0] = hs->W1;
      W[1] = hs->W2;
      W[2] = hs->W3;
    }
    if (area)
      for (int n = 0; n < N; ++n) area[n] = hs->area[n];
    return rc;
  }

  int psi_solve_host(const double* phi_star, double a, double dt, const double* rhs, double* out, int32_t* iters,
                     int32_t* solved, double* rnorm) override {
    if (nd != 2) return SAV_E_UNSUPPORTED;
    const long long NP = (long long)N * P, NH = (long long)N * H;
    put_real(phi_star, w_phis, NP);
    set_bdf(a, 0.0, 0.0, dt);
    R* M = w_M;
    const R* phs = w_phis;
    const double M0 = c.M0_psi;
    foreach_n(s, NP, [=] __device__(long long i) {
      const T t = phs[i] * phs[i] - T(1.0);
      M[i] = T(0.1) * (T(1.0) - (T)M0 * (t * t));
    });
    max_reduce(M, NP, &ds->Mmax);
    stencil_stale = true;
    C* dr = w_rhs;
    put_cplx(rhs, dr, NH);
    C* dout = w_rhs + NH;
    int it = psi_solve(dr, dout, 0);
    get_cplx(out, dout, NH);
    if (iters) *iters = it;
    if (solved) *solved = hks[0].solved;
    if (rnorm) *rnorm = hks[0].rnorm;
    return hks[0].solved ? 0 : SAV_W_KRYLOV_NOT_CONVERGED;
  }

  // irfft of `count` device half spectra (storage type R) into device fields, with the
  // intermediate of the column pass in this device's noalloc_scratch: launches only, no
  // allocation.  Every element of the intermediate is written before it is read, so the
  // scratch needs no clearing; calls that overlap on one GPU share it and race.
  int irfft_noalloc(const void* in_hat, void* out, int count) override {
    const long long need = (long long)count * H * (long long)sizeof(C);
    if (count < 1 || need > (long long)SAV_NOALLOC_SCRATCH_BYTES) return SAV_E_ARG;
    SpecJobs js{};
    js.j[js.n++] = job(in_hat, out, count);
    irfft_jobs<R>(g, js, reinterpret_cast<C*>(scratch), s);
    SAV_HIP(hipStreamSynchronize(s));
    return 0;
  }

  int modified_energy(double* E, double* ar) override {
    monitor();
    SAV_HIP(hipMemcpyAsync(hs, ds, sizeof(DScal), hipMemcpyDeviceToHost, s));
    SAV_HIP(hipStreamSynchronize(s));
    if (E) *E = hs->energy;
    if (ar) *ar = hs->area_ratio;
    return 0;
  }
};

// ---------------------------------------------------------------------------
// Size-class group of ensemble members (see ISolveGroup in solver.h).
// ---------------------------------------------------------------------------
template <typename R>
class SolveGroup final : public ISolveGroup {
 public:
  using C = cplx<R>;
  std::vector<Solver<R>*> mem;
  int K = 0;
  KrylovCtx<R> gk{};  // one batched solve over 3 K systems, system 3 m + z = member m's solve z
  hipStream_t st = nullptr;
  std::vector<hipEvent_t> ready;  // member m's right-hand sides are queued
  hipEvent_t solved = nullptr;
  std::vector<void*> allocs;
  KScal* hks = nullptr;  // pinned, 3 K + 1
  // rendezvous of the members' host threads
  std::mutex mu;
  std::condition_variable cv;
  int parties = 0, arrived = 0;
  long long gen = 0, solves = 0;
  std::vector<char> active;
  std::vector<int> left;  // members that left since the last solve (their systems are cleared)
  bool coupled = true;
  std::string fail;

  template <class T>
  T* alloc(size_t n) {
    void* q = nullptr;
    SAV_HIP(hipMalloc(&q, n * sizeof(T) + 256));
    SAV_HIP(hipMemsetAsync(q, 0, n * sizeof(T) + 256, st));
    allocs.push_back(q);
    return reinterpret_cast<T*>(q);
  }

  explicit SolveGroup(const std::vector<Solver<R>*>& members) : mem(members), K((int)members.size()) {
    const Solver<R>& m0 = *mem[0];
    SAV_HIP(hipSetDevice(m0.c.device));
    SAV_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    const long long nf = m0.kc.nfull;
    gk = m0.kc;  // tables, sizes, tolerances
    gk.vs = nf;
    gk.ps = m0.kc.ps;
    C** vec[9] = {&gk.x, &gk.r, &gk.p, &gk.v, &gk.s, &gk.t, &gk.b, &gk.Bx, &gk.By};
    for (auto v : vec) *v = alloc<C>(3LL * K * nf);
    R* Mp = alloc<R>((long long)K * nf);
    DScal* dsp = alloc<DScal>(K);
    gk.M = Mp;
    gk.ds = dsp;
    gk.partials = alloc<double>(3LL * K * gk.ps);
    gk.ks = alloc<KScal>(3 * K + 1);
    SAV_HIP(hipHostMalloc(reinterpret_cast<void**>(&hks), (3 * K + 1) * sizeof(KScal), hipHostMallocDefault));
    std::memset(hks, 0, (3 * K + 1) * sizeof(KScal));
    gk.hks = hks;
    gk.st = st;
    gk.colw = 0;
    gk.order = 0;
    gk.seq = 0;
    gk.tune_nsys = 3 * K;
    gk.timed = 0;
    gk.ev = nullptr;
    gk.ev_cap = gk.ev_used = 0;
    gk.snap = nullptr;
    gk.outer = nullptr;
    for (int z = 0; z < KRY_MAX_SYS; ++z) gk.guess[z] = 8;
    SAV_HIP(hipStreamSynchronize(st));
    ready.resize(K);
    for (auto& e : ready) SAV_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    SAV_HIP(hipEventCreateWithFlags(&solved, hipEventDisableTiming));
    for (int m = 0; m < K; ++m) {
      C* v[9];
      for (int i = 0; i < 9; ++i) v[i] = *vec[i] + 3LL * m * nf;
      mem[m]->adopt_group(this, m, v, Mp + (long long)m * nf, dsp + m, gk.ks + 3 * m, hks + 3 * m);
    }
    active.assign(K, 1);
    parties = K;
    set_coupled(false);  // coupled only while sav_ensemble_run steps all members together
  }

  ~SolveGroup() override {
    if (st) hipStreamSynchronize(st);
    for (auto e : ready) hipEventDestroy(e);
    if (solved) hipEventDestroy(solved);
    for (void* q : allocs) hipFree(q);
    if (hks) hipHostFree(hks);
    if (st) hipStreamDestroy(st);
  }

  void begin_run() override {
    std::lock_guard<std::mutex> lk(mu);
    active.assign(K, 1);
    left.clear();
    parties = K;
    arrived = 0;
    fail.clear();
  }

  void set_coupled(bool on) override {
    coupled = on;
    for (int m = 0; m < K; ++m) mem[m]->group = on ? this : nullptr;
  }

  int size() const override { return K; }
  void info(long long* n, int* columns, int* order) const override {
    if (n) *n = solves;
    if (columns) *columns = gk.colw;
    if (order) *order = gk.order;
  }

  // The batched solve of every member still taking part (lock not held; the others wait).
  void run_solve() {
    for (int m : left) {
      // a member that left: zero right-hand sides, its systems finish at once
      SAV_HIP(hipMemsetAsync(gk.b + 3LL * m * gk.vs, 0, 3 * gk.vs * sizeof(C), st));
    }
    left.clear();
    for (int m = 0; m < K; ++m)
      if (active[m]) SAV_HIP(hipStreamWaitEvent(st, ready[m], 0));
    bicgstab_batch<R>(gk, 0, 3 * K);
    SAV_HIP(hipEventRecord(solved, st));
    ++solves;
  }

  // With `lk` held and every member still taking part arrived: run the solve, wake them.
  void release(std::unique_lock<std::mutex>& lk) {
    arrived = 0;
    lk.unlock();
    try {
      run_solve();
    } catch (const std::exception& e) {
      fail = e.what();
    }
    lk.lock();
    ++gen;
    cv.notify_all();
  }

  // Member m's right-hand sides are queued on its stream: wait for the group's solve, after
  // which its stream waits for the solution (its Krylov scalars are then in hks).
  void solve_point(int m) {
    Solver<R>* me = mem[m];
    SAV_HIP(hipEventRecord(ready[m], me->s));
    std::unique_lock<std::mutex> lk(mu);
    const long long g0 = gen;
    if (++arrived == parties)
      release(lk);
    else
      cv.wait(lk, [&] { return gen != g0; });
    if (!fail.empty()) throw HipError("size-class group solve: " + fail, SAV_E_HIP);
    lk.unlock();
    SAV_HIP(hipStreamWaitEvent(me->s, solved, 0));
  }

  void leave(int m) override {
    std::unique_lock<std::mutex> lk(mu);
    if (m < 0 || m >= K || !active[m]) return;
    active[m] = 0;
    left.push_back(m);
    --parties;
    if (parties > 0 && arrived == parties) release(lk);
  }
};

template <typename R>
void Solver<R>::group_solve() {
  group->solve_point(group_m);
}

template <typename R>
static std::unique_ptr<ISolveGroup> make_group_t(const std::vector<ISolver*>& members) {
  std::vector<Solver<R>*> ms;
  for (ISolver* i : members) {
    auto* q = dynamic_cast<Solver<R>*>(i);
    if (!q) throw std::invalid_argument("size-class group: members of one storage precision");
    ms.push_back(q);
  }
  const Solver<R>& a = *ms[0];
  for (Solver<R>* q : ms) {
    if (!q->group_ready() || q->group)
      throw std::invalid_argument("size-class group: 2D, not mixed, diagonal preconditioner, not grouped yet");
    const sav_config_t &x = a.c, &y = q->c;
    if (y.Nx != x.Nx || y.Ny != x.Ny || y.N != x.N || y.S4 != x.S4 || y.tol != x.tol || y.atol != x.atol ||
        y.itmax != x.itmax || y.device != x.device)
      throw std::invalid_argument("size-class group: members differ in grid, N, S4 or the psi-solve settings");
  }
  return std::unique_ptr<ISolveGroup>(new SolveGroup<R>(ms));
}

std::unique_ptr<ISolveGroup> make_solve_group(const std::vector<ISolver*>& members) {
  if (members.size() < 2 || members.size() * 3 > (size_t)KRY_MAX_SYS)
    throw std::invalid_argument("size-class group: 2 to 64 members");
  if (dynamic_cast<Solver<double>*>(members[0])) return make_group_t<double>(members);
  return make_group_t<float>(members);
}

// Global maximum of a field into a device double.
template <typename R>
__global__ __launch_bounds__(256) void k_max_stage(const R* __restrict__ x, long long n, double* __restrict__ partials) {
  __shared__ double sc[4];
  double m = -INFINITY;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    m = fmax(m, (double)x[i]);
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) sc[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) partials[blockIdx.x] = fmax(fmax(sc[0], sc[1]), fmax(sc[2], sc[3]));
}
__global__ __launch_bounds__(256) void k_max_fin(const double* __restrict__ partials, int nb, double* __restrict__ out) {
  __shared__ double sc[4];
  double m = -INFINITY;
  for (int i = threadIdx.x; i < nb; i += 256) m = fmax(m, partials[i]);
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) sc[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) *out = fmax(fmax(sc[0], sc[1]), fmax(sc[2], sc[3]));
}

template <typename R>
void Solver<R>::reduce_max_impl(const R* x, long long n, double* out) {
  long long nb = (n + 1023) / 1024;
  if (nb < 1) nb = 1;
  if (nb > RED_STAGE_BLOCKS) nb = RED_STAGE_BLOCKS;
  hipLaunchKernelGGL((k_max_stage<R>), dim3((unsigned)nb), dim3(256), 0, s, x, n, partials);
  hipLaunchKernelGGL(k_max_fin, dim3(1), dim3(256), 0, s, partials, (int)nb, out);
  SAV_LAUNCH_CHECK();
}

std::unique_ptr<ISolver> make_solver(const HostCfg& cfg) {
  if (cfg.c.precision == SAV_PREC_FP32) return std::unique_ptr<ISolver>(new Solver<float>(cfg));
  // SAV_PREC_FP64 and SAV_PREC_MIXED (FP64 storage; FP32 inner psi iterations in 2D)
  return std::unique_ptr<ISolver>(new Solver<double>(cfg));
}

}  // namespace sav
