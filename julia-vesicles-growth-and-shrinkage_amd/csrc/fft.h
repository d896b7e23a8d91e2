// Batched mixed-radix (2/3/4/5/8, optionally 16/32) Stockham FFTs for gfx950.
//
// One transform of length L is carried by a "team" of TT = L/E threads, each
// holding E elements in registers.  A radix-R pass makes every thread do E/R
// butterflies in registers; passes exchange data through LDS.  The first pass
// reads its inputs straight from registers that the caller loaded from HBM in
// the order  pos = t + q*TT + k*(L/R0)  (consecutive threads -> consecutive
// addresses), and the last pass leaves its outputs in registers in the order
// pos = t + q*TT + m*(L/RL), again coalesced for the caller's store.
//
// Stockham autosort pass with radix R after passes of total length Ls:
//   butterfly j in [0, L/R): inputs x[j + k*L/R], twiddles W_{Ls*R}^{(j mod Ls)*k},
//   outputs at (j / Ls)*Ls*R + (j mod Ls) + m*Ls.
#pragma once
#include <cmath>
#include <vector>

#include "common.h"

namespace sav {

struct FftPlanData {
  int np;
  int rad[12];
  int E;
};

// ttmax: largest team (lanes per transform).  FP64 uses 128-lane teams (E = 8 for
// L = 1024: low VGPR count, 4 waves/SIMD); FP32 uses 64-lane teams so that a
// 512-thread column workgroup still spans 8 adjacent columns (64-byte segments).
// maxr: largest power-of-two radix of a pass (8, 16 or 32): fewer passes, hence fewer
// LDS exchanges, for more registers per lane.  rev: the passes in reverse order (the
// first pass of a reversed plan has the radix of the last pass of the plain one).
constexpr FftPlanData make_fft_plan(int L, int ttmax = 64, int maxr = 8, bool rev = false) {
  FftPlanData p{};
  int n = L, np = 0, has3 = 0, has5 = 0;
  while (n % 3 == 0) { p.rad[np++] = 3; n /= 3; has3 = 1; }
  while (n % 5 == 0) { p.rad[np++] = 5; n /= 5; has5 = 1; }
  int a = 0;
  while (n % 2 == 0) { n /= 2; ++a; }
  // with a radix-5 pass the power-of-two passes are at most radix 4, so that a lane's
  // elements (a multiple of every radix) stay at 20
  const int lg = has5 ? 2 : maxr >= 32 ? 5 : maxr >= 16 ? 4 : 3;
  int used = 1;
  const int nbig = a / lg, rem = a % lg;
  for (int i = 0; i < nbig; ++i) { p.rad[np++] = 1 << lg; used = 1 << lg; }
  if (rem > 0) { p.rad[np++] = 1 << rem; if (used < (1 << rem)) used = 1 << rem; }
  p.np = np;
  int lcm = (has3 ? 3 : 1) * (has5 ? 5 : 1) * used;
  if (lcm < 1) lcm = 1;
  int E = lcm;
  while (L / E > ttmax && (L % (2 * E)) == 0 && 2 * E <= 32) E *= 2;
  p.E = E;
  if (rev)
    for (int i = 0; i < np / 2; ++i) {
      const int t = p.rad[i];
      p.rad[i] = p.rad[np - 1 - i];
      p.rad[np - 1 - i] = t;
    }
  return p;
}

// Largest radix of the psi matvec's row plan: 16 for powers of two, else 8.
constexpr int row_plan_maxr(int L) { return (L & (L - 1)) == 0 ? 16 : 8; }

// The psi matvec's column plans for power-of-two columns of at least 512: the inverse
// kernel (one input field) runs radix-32 passes in teams of 32 lanes (E = 32) in FP32 and
// radix-16 passes in teams of 64 in FP64 (at 128 KB of LDS per workgroup, one workgroup
// per CU: 64-lane teams keep two wavefronts per SIMD instead of one); the forward kernel
// (two input fields) radix-16 passes in teams of 64 (E = 16); other lengths use the
// default plan.  FWD selects the forward kernel's plan.
template <typename R>
constexpr int col_plan_maxr(int L, bool fwd) {
  return ((L & (L - 1)) == 0 && L >= 512) ? ((fwd || sizeof(R) == 8) ? 16 : 32) : 8;
}
// (FP64 columns of 2048 points run 128-lane teams, E = 16: with 64-lane teams a lane would
// hold 32 complex doubles per field.)
template <typename R>
constexpr int col_plan_ttmax(int L, bool fwd) {
  return col_plan_maxr<R>(L, fwd) == 32                                ? 32
         : col_plan_maxr<R>(L, fwd) == 16 ? ((sizeof(R) == 8 && L >= 2048) ? 128 : 64)
                                           : (sizeof(R) == 8 ? 128 : 64);
}

// Largest radix of the general column transforms (k_cols: y and z passes, the full
// complex 2D transforms): 16 in FP32 for power-of-two lengths of at least 512 (16 16 4 at
// 1024: two LDS exchanges instead of three), else the default 8.
constexpr int cols_maxr(int L, bool fp32) { return (fp32 && (L & (L - 1)) == 0 && L >= 512) ? 16 : 8; }


// Lengths with a plan: 2^a 3^b or 2^a 5^c (3 and 5 are not mixed) up to 2048.
constexpr bool fft_len_supported(int L) {
  int n = L, f3 = 0, f5 = 0;
  while (n % 3 == 0) { n /= 3; f3 = 1; }
  while (n % 5 == 0) { n /= 5; f5 = 1; }
  while (n % 2 == 0) n /= 2;
  return n == 1 && !(f3 && f5) && L >= 2 && L <= 2048;
}

// ---------------------------------------------------------------------------
// Butterflies.  INV selects e^{+i} kernels.
// ---------------------------------------------------------------------------
template <typename R, bool INV>
__device__ __forceinline__ void bfly2(cplx<R>* a) {
  cplx<R> t = a[0];
  a[0] = t + a[1];
  a[1] = t - a[1];
}

template <typename R, bool INV>
__device__ __forceinline__ void bfly3(cplx<R>* a) {
  const R s = R(0.866025403784438646763723170752936183);
  cplx<R> b = a[1] + a[2];
  cplx<R> d = s * (a[1] - a[2]);
  cplx<R> t1 = a[0] - R(0.5) * b;
  a[0] = a[0] + b;
  cplx<R> t2 = INV ? mul_i(d) : mul_mi(d);
  a[1] = t1 + t2;
  a[2] = t1 - t2;
}

// 5-point DFT from its definition: with b1 = a1 + a4, b2 = a2 + a3, d1 = a1 - a4,
// d2 = a2 - a3, c_m = cos(2 pi m / 5), s_m = sin(2 pi m / 5):
//   X0 = a0 + b1 + b2,  X1,4 = (a0 + c1 b1 + c2 b2) -/+ i (s1 d1 + s2 d2),
//   X2,3 = (a0 + c2 b1 + c1 b2) -/+ i (s2 d1 - s1 d2)   (signs of i flipped for INV).
template <typename R, bool INV>
__device__ __forceinline__ void bfly5(cplx<R>* a) {
  const R c1 = R(0.309016994374947424102293417182819058860);
  const R c2 = R(-0.809016994374947424102293417182819058860);
  const R s1 = R(0.951056516295153572116439333379382143406);
  const R s2 = R(0.587785252292473129168705954639072768598);
  const cplx<R> b1 = a[1] + a[4], b2 = a[2] + a[3], d1 = a[1] - a[4], d2 = a[2] - a[3];
  const cplx<R> t1 = a[0] + c1 * b1 + c2 * b2;
  const cplx<R> t2 = a[0] + c2 * b1 + c1 * b2;
  const cplx<R> u1 = s1 * d1 + s2 * d2;
  const cplx<R> u2 = s2 * d1 - s1 * d2;
  const cplx<R> iu1 = INV ? mul_i(u1) : mul_mi(u1);
  const cplx<R> iu2 = INV ? mul_i(u2) : mul_mi(u2);
  a[0] = a[0] + b1 + b2;
  a[1] = t1 + iu1;
  a[4] = t1 - iu1;
  a[2] = t2 + iu2;
  a[3] = t2 - iu2;
}

template <typename R, bool INV>
__device__ __forceinline__ void bfly4(cplx<R>& a0, cplx<R>& a1, cplx<R>& a2, cplx<R>& a3) {
  cplx<R> t0 = a0 + a2, t1 = a0 - a2, t2 = a1 + a3;
  cplx<R> t3 = INV ? mul_i(a1 - a3) : mul_mi(a1 - a3);
  a0 = t0 + t2;
  a2 = t0 - t2;
  a1 = t1 + t3;
  a3 = t1 - t3;
}

template <typename R, bool INV>
__device__ __forceinline__ void bfly8(cplx<R>* a) {
  const R h = R(0.707106781186547524400844362104849039);
  cplx<R> e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
  cplx<R> o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
  bfly4<R, INV>(e0, e1, e2, e3);
  bfly4<R, INV>(o0, o1, o2, o3);
  // W8^m * O_m
  cplx<R> w1, w2, w3;
  if (!INV) {
    w1 = {h * (o1.x + o1.y), h * (o1.y - o1.x)};
    w2 = mul_mi(o2);
    w3 = {h * (o3.y - o3.x), -h * (o3.x + o3.y)};
  } else {
    w1 = {h * (o1.x - o1.y), h * (o1.x + o1.y)};
    w2 = mul_i(o2);
    w3 = {-h * (o3.x + o3.y), h * (o3.x - o3.y)};
  }
  a[0] = e0 + o0;
  a[4] = e0 - o0;
  a[1] = e1 + w1;
  a[5] = e1 - w1;
  a[2] = e2 + w2;
  a[6] = e2 - w2;
  a[3] = e3 + w3;
  a[7] = e3 - w3;
}

template <typename R, bool INV, int RAD>
__device__ __forceinline__ void bfly(cplx<R>* a);

// cos(2 pi m / 32), m = 0..8 (correctly rounded); the other angles by symmetry.
constexpr double kCos32[9] = {1.0,
                              0.9807852804032304491261822361342390369739,
                              0.9238795325112867561281831893967882868224,
                              0.8314696123025452370787883776179057567386,
                              0.7071067811865475244008443621048490392848,
                              0.5555702330196022247428308139485328743749,
                              0.3826834323650897717284599840303988667613,
                              0.1950903220161282678482848684770222409277,
                              0.0};

// a * W_RAD^m, W = e^{-2 pi i / RAD} (e^{+} for INV); m is a constant after unrolling.
template <typename R, bool INV, int RAD>
__device__ __forceinline__ cplx<R> rot_const(cplx<R> a, int m) {
  m %= RAD;
  if (m == 0) return a;
  if (2 * m == RAD) return cplx<R>{-a.x, -a.y};
  if (4 * m == RAD) return INV ? mul_i(a) : mul_mi(a);
  if (4 * m == 3 * RAD) return INV ? mul_mi(a) : mul_i(a);
  const int q = m * (32 / RAD);  // angle 2 pi q / 32
  // cos and sin of 2 pi q / 32 from the first octant
  const int o = q & 7, quad = q >> 3;
  const double c0 = kCos32[o], s0 = kCos32[8 - o];
  const double cq = quad == 0 ? c0 : quad == 1 ? -s0 : quad == 2 ? -c0 : s0;
  const double sq = quad == 0 ? s0 : quad == 1 ? c0 : quad == 2 ? -s0 : -c0;
  const R c = (R)cq, sn = (R)(INV ? sq : -sq);
  return cplx<R>{a.x * c - a.y * sn, a.x * sn + a.y * c};
}

// Radix R1*R2 butterfly from R1- and R2-point ones (Cooley-Tukey in registers):
// X[k1 + R1 k2] = sum_n2 W_R2^{n2 k2} W_RAD^{n2 k1} sum_n1 a[R2 n1 + n2] W_R1^{n1 k1}.
template <typename R, bool INV, int R1, int R2>
__device__ __forceinline__ void bfly_comp(cplx<R>* a) {
  constexpr int RAD = R1 * R2;
  cplx<R> y[RAD];
#pragma unroll
  for (int n2 = 0; n2 < R2; ++n2) {
    cplx<R> t[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) t[n1] = a[R2 * n1 + n2];
    bfly<R, INV, R1>(t);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) y[n2 * R1 + k1] = rot_const<R, INV, RAD>(t[k1], n2 * k1);
  }
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    cplx<R> u[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) u[n2] = y[n2 * R1 + k1];
    bfly<R, INV, R2>(u);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) a[k1 + R1 * k2] = u[k2];
  }
}

template <typename R, bool INV, int RAD>
__device__ __forceinline__ void bfly(cplx<R>* a) {
  if constexpr (RAD == 2) bfly2<R, INV>(a);
  else if constexpr (RAD == 3) bfly3<R, INV>(a);
  else if constexpr (RAD == 4) bfly4<R, INV>(a[0], a[1], a[2], a[3]);
  else if constexpr (RAD == 5) bfly5<R, INV>(a);
  else if constexpr (RAD == 8) bfly8<R, INV>(a);
  else if constexpr (RAD == 16) bfly_comp<R, INV, 4, 4>(a);
  else if constexpr (RAD == 32) bfly_comp<R, INV, 8, 4>(a);
  else static_assert(RAD == 2, "unsupported radix");
}

// ---------------------------------------------------------------------------
// Team FFT.  `lds` points at this team's element 0, element p lives at
// lds[p * ls] (ls = 1 for private row storage, = team count for interleaved
// column storage).  `tw` = exp(-2 pi i n / L), n in [0, L).
// ---------------------------------------------------------------------------
// Pass-ordered twiddle table of a length-L plan: for pass p >= 1 (stride LS_p,
// radix R_p) the row of butterfly residue jm holds W_{LS R}^{jm k}, k = 1..R-1, at
// offset twoff(p) + jm*8 + (k-1): one butterfly reads 8 consecutive entries.
constexpr int twiddle_table_len(int L) {
  FftPlanData P = make_fft_plan(L);
  int ls = 1, n = 0;
  for (int p = 0; p < P.np; ++p) {
    if (p > 0) n += ls * 8;
    ls *= P.rad[p];
  }
  return n > 0 ? n : 1;
}

// TTMAX: largest team; PADS: LDS padding, one slot every 2^PADS positions (0 = none),
// which spreads the stride-R Stockham writes of the early passes over the banks.
// TWM: 0 = plain table exp(-2 pi i n / L), 1 = pass-ordered table (twiddle_table_len),
// 2 = k-major pass table of the plan (kmajor_twiddles).
// WSYNC: the workgroup is one wavefront holding whole teams (TT divides 64): LDS
// exchanges between passes need only a wavefront-level barrier.
// MAXR, REV: the plan's largest radix and pass order (make_fft_plan); the pass-ordered
// twiddle table (TWM = 1) exists for the default plan only.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Default twiddle layout: the k-major pass table, in which the lanes of a team read
// consecutive entries for each twiddle of a butterfly (one or two cache lines per load
// instruction; the plain and pass-ordered tables scatter a wavefront's loads over many).
template <typename R>
constexpr int default_twm() {
  return 2;
}

template <typename R, int L, bool INV, int TTMAX = (sizeof(R) == 8 ? 128 : 64), int PADS = 0,
          int TWM = default_twm<R>(), int WSYNC = 0, int MAXR = 8, bool REV = false>
struct TeamFFT {
  static constexpr FftPlanData P = make_fft_plan(L, TTMAX, MAXR, REV);
  static_assert(TWM != 1 || (MAXR == 8 && !REV), "pass-ordered twiddles exist for the default plan only");
  static constexpr int E = P.E;
  static constexpr int TT = L / E;
  static constexpr int R0 = P.rad[0];
  static constexpr int RL = P.rad[P.np - 1];
  static constexpr int LDS_LEN = PADS ? L + (L >> PADS) + 1 : L;   // LDS slots per transform
  __device__ static __forceinline__ int lidx(int p) { return PADS ? p + (p >> PADS) : p; }
  static_assert(!WSYNC || 64 % (L / P.E) == 0, "wavefront-synchronised teams must lie within one wavefront");
  __device__ static __forceinline__ void sync() {
    if constexpr (WSYNC) wave_lds_sync();
    else __syncthreads();
  }

  __device__ static __forceinline__ int in_pos(int t, int e) { return t + (e / R0) * TT + (e % R0) * (L / R0); }
  __device__ static __forceinline__ int out_pos(int t, int e) { return t + (e / RL) * TT + (e % RL) * (L / RL); }

  template <int PI, int LS, int TWO>
  __device__ static __forceinline__ void passes(cplx<R> (&v)[E], cplx<R>* lds, int ls, int t,
                                                const cplx<R>* __restrict__ tw) {
    if constexpr (PI < P.np) {
      constexpr int RAD = P.rad[PI];
      constexpr int NB = E / RAD;
      constexpr int LR = L / RAD;
      constexpr int TWS = L / (LS * RAD);
      if constexpr (PI > 0) {
#pragma unroll
        for (int q = 0; q < NB; ++q) {
          const int j = t + q * TT;
#pragma unroll
          for (int k = 0; k < RAD; ++k) v[q * RAD + k] = lds[lidx(j + k * LR) * ls];
        }
      }
#pragma unroll
      for (int q = 0; q < NB; ++q) {
        const int j = t + q * TT;
        if constexpr (LS > 1) {
          const int jm = j % LS;
          const cplx<R>* twr = tw + TWO + (TWM == 2 ? jm : jm * 8);
#pragma unroll
          for (int k = 1; k < RAD; ++k) {
            cplx<R> w = TWM == 1 ? twr[k - 1] : TWM == 2 ? twr[(k - 1) * LS] : tw[jm * k * TWS];
            if (INV) w.y = -w.y;
            v[q * RAD + k] = v[q * RAD + k] * w;
          }
        }
        bfly<R, INV, RAD>(&v[q * RAD]);
      }
      if constexpr (PI + 1 < P.np) {
        sync();
#pragma unroll
        for (int q = 0; q < NB; ++q) {
          const int j = t + q * TT;
          const int base = (j / LS) * LS * RAD + (j % LS);
#pragma unroll
          for (int m = 0; m < RAD; ++m) lds[lidx(base + m * LS) * ls] = v[q * RAD + m];
        }
        sync();
        passes<PI + 1, LS * RAD, (PI > 0 ? TWO + LS * (TWM == 2 ? RAD - 1 : 8) : TWO)>(v, lds, ls, t, tw);
      }
    }
  }

  __device__ static __forceinline__ void run(cplx<R> (&v)[E], cplx<R>* lds, int ls, int t,
                                             const cplx<R>* __restrict__ tw) {
    passes<0, 1, 0>(v, lds, ls, t, tw);
  }
};

// Host: pass-ordered twiddle table (see twiddle_table_len), computed in long double.
template <typename R>
inline void build_pass_twiddles(int L, cplx<R>* out) {
  const FftPlanData P = make_fft_plan(L);
  int ls = 1, off = 0;
  const long double two_pi = 6.283185307179586476925286766559005768L;
  for (int p = 0; p < P.np; ++p) {
    const int rad = P.rad[p];
    if (p > 0) {
      for (int jm = 0; jm < ls; ++jm)
        for (int k = 0; k < 8; ++k) {
          cplx<R> w{R(0), R(0)};
          if (k + 1 < rad) {
            const long double a = -two_pi * (long double)(jm * (k + 1)) / (long double)(ls * rad);
            w = cplx<R>{(R)cosl(a), (R)sinl(a)};
          }
          out[off + jm * 8 + k] = w;
        }
      off += ls * 8;
    }
    ls *= rad;
  }
  if (off == 0) out[0] = cplx<R>{R(1), R(0)};
}

// Host: k-major pass table (TWM = 2) of the plan make_fft_plan(L, ttmax, maxr, rev): for
// pass p >= 1 (stride LS_p, radix R_p) W_{LS R}^{jm k} sits at off_p + (k - 1) LS_p + jm,
// so the lanes of a team read consecutive entries for each k.
template <typename R>
inline std::vector<cplx<R>> kmajor_twiddles(int L, int ttmax, int maxr, bool rev) {
  const FftPlanData P = make_fft_plan(L, ttmax, maxr, rev);
  const long double two_pi = 6.283185307179586476925286766559005768L;
  std::vector<cplx<R>> out;
  int ls = 1;
  for (int p = 0; p < P.np; ++p) {
    const int rad = P.rad[p];
    if (p > 0)
      for (int k = 1; k < rad; ++k)
        for (int jm = 0; jm < ls; ++jm) {
          const long double a = -two_pi * (long double)(jm * k) / (long double)(ls * rad);
          out.push_back(cplx<R>{(R)cosl(a), (R)sinl(a)});
        }
    ls *= rad;
  }
  if (out.empty()) out.push_back(cplx<R>{R(1), R(0)});
  return out;
}

// Host: the table the default TeamFFT<R, L, ...> expects for length L.
template <typename R>
inline std::vector<cplx<R>> default_twiddles(int L) {
  if (default_twm<R>() == 2) return kmajor_twiddles<R>(L, 64, 8, false);
  if (default_twm<R>() == 0) {
    std::vector<cplx<R>> w(L);
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (int n = 0; n < L; ++n) {
      const long double a = -two_pi * (long double)n / (long double)L;
      w[n] = cplx<R>{(R)cosl(a), (R)sinl(a)};
    }
    return w;
  }
  std::vector<cplx<R>> w(twiddle_table_len(L));
  build_pass_twiddles<R>(L, w.data());
  return w;
}

// Host: the twiddle table k_cols<R, L> reads.
template <typename R>
inline std::vector<cplx<R>> column_twiddles(int L) {
  return kmajor_twiddles<R>(L, 64, cols_maxr(L, sizeof(R) == 4), false);
}

// Host: the twiddle table of the register-column plane kernels for 128 x 128 planes
// (k_hplane_r2c / k_hplane_c2r), forward sign, in three parts laid out so that the lanes
// of a 16-lane LDS access group (two teams) read opposite halves of the 64 banks:
//   [0, 192)    W_128^{t k1}, t < 8, k1 < 16, at t * 24 + k1    (y pass, 128 = 16 x 8)
//   [192, 256)  W_64^{s k1},  s < 8, k1 < 8,  at 192 + k1 * 8 + s (x pass, 64 = 8 x 8)
//   [256, 321)  W_128^k,      k <= 64,        at 256 + k          (the r2c / c2r split)
constexpr int kHPlaneTw = 321;
constexpr int kHPlaneTX = 192, kHPlaneT2 = 256;
template <typename R>
inline std::vector<cplx<R>> hplane_twiddles() {
  const long double two_pi = 6.283185307179586476925286766559005768L;
  std::vector<cplx<R>> w(kHPlaneTw, cplx<R>{R(0), R(0)});
  auto put = [&](int i, int num, int den) {
    const long double a = -two_pi * (long double)num / (long double)den;
    w[i] = cplx<R>{(R)cosl(a), (R)sinl(a)};
  };
  for (int t = 0; t < 8; ++t)
    for (int k1 = 0; k1 < 16; ++k1) put(t * 24 + k1, t * k1, 128);
  for (int s = 0; s < 8; ++s)
    for (int k1 = 0; k1 < 8; ++k1) put(kHPlaneTX + k1 * 8 + s, s * k1, 64);
  for (int k = 0; k <= 64; ++k) put(kHPlaneT2 + k, k, 128);
  return w;
}

// Columns per workgroup of a strided (column) transform: 8 adjacent columns (128-byte
// row segments in FP64, 64-byte in FP32), at least 256 threads per workgroup, at most
// 1024 threads and the 160 KB of LDS of a gfx950 CU.
template <typename R, int L, int WANT = 8, int MAXR = 8, int TTMAX = (sizeof(R) == 8 ? 128 : 64)>
struct ColTile {
  using F = TeamFFT<R, L, false, TTMAX, 0, default_twm<R>(), 0, MAXR>;
  static constexpr int TT = F::TT;
  static constexpr int ES = (int)sizeof(cplx<R>);
  static constexpr int want = WANT > (256 / TT) ? WANT : (256 / TT);
  static constexpr int cap0 = 163840 / (L * ES);
  static constexpr int capp2 = cap0 >= 256 ? 256 : cap0 >= 128 ? 128 : cap0 >= 64 ? 64 : cap0 >= 32 ? 32 : cap0 >= 16 ? 16 : cap0 >= 8 ? 8 : cap0 >= 4 ? 4 : cap0 >= 2 ? 2 : 1;
  static constexpr int C0 = want < capp2 ? want : capp2;
  static constexpr int C = (C0 * TT > 1024) ? (1024 / TT) : C0;
  static constexpr int NT = C * TT;
};

// Multipliers applied while loading the first inverse pass.
enum SpecMul : int {
  MUL_NONE = 0,
  MUL_DX = 1,   // i*kx
  MUL_DY = 2,   // i*ky
  MUL_DZ = 3,   // i*kz
  MUL_LAP = 4,  // -(kx^2+ky^2+kz^2)
  MUL_BIH = 5,  // (kx^2+ky^2+kz^2)^2
};

// Device wavenumber tables (FP64): kx rfftfreq (Nyquist +), ky/kz fftfreq (Nyquist -),
// kx0 fftfreq over Nx (Nyquist -).
struct KTables {
  const double* kx;   // Nxh
  const double* ky;   // Ny
  const double* kz;   // Nz (3D) or nullptr
  const double* kx0;  // Nx
};

template <typename R>
__device__ __forceinline__ cplx<R> apply_mul(cplx<R> a, int mul, double kx, double ky, double kz) {
  switch (mul) {
    // (0 + i k)(a.x + i a.y) = (-k a.y) + i (k a.x)
    case MUL_DX: { R k = (R)kx; return cplx<R>{-k * a.y, k * a.x}; }
    case MUL_DY: { R k = (R)ky; return cplx<R>{-k * a.y, k * a.x}; }
    case MUL_DZ: { R k = (R)kz; return cplx<R>{-k * a.y, k * a.x}; }
    case MUL_LAP: {
      R k2 = (R)(kx * kx + ky * ky + kz * kz);
      R l = -k2;
      return cplx<R>{l * a.x, l * a.y};
    }
    case MUL_BIH: {
      R k2 = (R)(kx * kx + ky * ky + kz * kz);
      R b = k2 * k2;
      return cplx<R>{b * a.x, b * a.y};
    }
    default: return a;
  }
}

// ---------------------------------------------------------------------------
// Job descriptors for batched launches (up to 8 jobs per launch).
// ---------------------------------------------------------------------------
constexpr int MAXJ = 12;

struct SpecJob {
  const void* src;   // half spectrum (or physical for r2c)
  const void* src2;  // optional: load 2*src - src2
  void* dst;         // destination
  int mul;           // SpecMul applied on load (inverse only)
  int count;         // number of scalar fields in this job
  double scale;      // extra real factor applied on load
  long long toff;    // offset (complex elements) of this job inside the irfft scratch
  int pmul;          // inverse 3D plane path: multiplier constant along z (MUL_DX / MUL_DY)
                     // applied by the plane pass to a z pass it may share with other jobs
};

struct SpecJobs {
  SpecJob j[MAXJ];
  int n;
};

// Host API -------------------------------------------------------------------
template <typename R>
struct FftGeom {
  int ndim, Nx, Ny, Nz, Nxh;
  const cplx<R>* tw_x2;   // table for L = Nx (length Nx) -- r2c/c2r post-twiddles
  const cplx<R>* tw_xh;   // table for L = Nx/2
  const cplx<R>* tw_x;    // table for L = Nx (full c2c rows)
  const cplx<R>* tw_y;    // table for L = Ny
  const cplx<R>* tw_z;    // table for L = Nz
  KTables k;
  int colw = 0;           // columns per workgroup of the inverse column transforms: 0 = default rule, 4 / 8 / 16
  int colw_fwd = 0;       // the same for the forward column transforms
  const cplx<R>* tw_hp = nullptr;  // hplane_twiddles() for 128 x 128 planes of 3D fields
  int plane_kind = 0;     // 128 x 128 planes: 0 = register-column kernels, 1 = whole plane in LDS
  double dkx = 0.0, dky = 0.0;  // wavenumber spacings 2 pi / Lx, 2 pi / Ly
};

// Columns per workgroup that width cw (0, 4, 8, 16) resolves to for a column transform of
// length L with real_bytes-byte values (the LDS and thread limits can cap it).
int col_tile_columns(int L, int real_bytes, int cw);

// rfft over the spatial axes: dst_j = rfft(src_j), count_j fields each.
template <typename R>
void rfft_jobs(const FftGeom<R>& g, const SpecJobs& jobs, hipStream_t s);

// irfft over the spatial axes with an optional spectral multiplier on load:
// dst_j = irfft(mul_j * scale_j * (src_j or 2 src_j - src2_j)).  `tmp` must hold
// sum(count_j) half-spectrum fields.
template <typename R>
void irfft_jobs(const FftGeom<R>& g, const SpecJobs& jobs, cplx<R>* tmp, hipStream_t s);

// Full complex 2D transforms over (x, y) of `count` fields, in place.
template <typename R>
void fft_full_2d(const FftGeom<R>& g, cplx<R>* data, int count, bool inverse, hipStream_t s);

bool fft_size_supported(int n);

}  // namespace sav
