"""Config / BDF types mirroring the reference's ``src/Types.jl`` and
``src/Types_3d.jl``, the reference's parameter defaults (``src/Init.jl:8-69``)
and the seeded benchmark workloads (single vesicle 128^2, four vesicles 512^2,
two vesicles 1024^2, two ellipsoids 128^3, and the uneven 2D ensemble).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, replace
from typing import List, Optional

import numpy as np

SQ2 = math.sqrt(2.0)


class InexactError(ValueError):
    """Julia raises InexactError when a non-integral value is stored in one of
    Config's Int fields (Types.jl:35-43)."""


def _as_int(name, v):
    fv = float(v)
    if not fv.is_integer():
        raise InexactError(f"Config.{name} is Int in the reference; got {v!r}")
    return int(fv)


@dataclass(frozen=True)
class Config:
    """Immutable parameters (Types.jl:27-74; Types_3d.jl adds Nz, Lz, dz).

    Derived fields follow the keyword constructor (Types.jl:101-134):
    ``Nt = round(T/dt)``, ``dx = Lx/Nx``, ``dy = Ly/Ny`` (``dz = Lz/Nz``),
    ``spe_len = (Nx/2+1)*Ny*N`` (times Nz in 3D).
    """
    N: int
    epsilon: float
    M_phi: float
    M0_psi: float
    eta: float
    gamma_surf: float
    gamma_area: int
    gamma_bend: float
    gamma_in: int
    beta_in: int
    psi_in_v: tuple
    gamma_out: int
    beta_out: int
    psi_out_v: tuple
    lamda: int
    S1: float
    S2: float
    S3: float
    S4: float
    C1: float
    C2: float
    C3: float
    dt: float
    T: float
    Nt: int
    Nx: int
    Ny: int
    Lx: float
    Ly: float
    dx: float
    dy: float
    tol: float
    goal: str
    spe_len: int
    ndim: int = 2
    Nz: int = 1
    Lz: float = 1.0
    dz: float = 1.0

    @staticmethod
    def make(*, N, epsilon, M_phi, M0_psi, eta, gamma_surf, gamma_area, gamma_bend,
             gamma_in, beta_in, psi_in_v, gamma_out, beta_out, psi_out_v, lamda,
             S1, S2, S3, S4, C1, C2, C3, dt, T, Nx, Ny, Lx, Ly, tol, goal="s",
             Nz=None, Lz=None) -> "Config":
        ndim = 2 if Nz is None else 3
        Nz_ = 1 if Nz is None else int(Nz)
        Lz_ = 1.0 if Lz is None else float(Lz)
        Nt = int(round(T / dt))
        dx, dy = Lx / Nx, Ly / Ny
        dz = Lz_ / Nz_ if ndim == 3 else 1.0
        spe_len = (Nx // 2 + 1) * Ny * Nz_ * N
        return Config(N=int(N), epsilon=float(epsilon), M_phi=float(M_phi), M0_psi=float(M0_psi),
                      eta=float(eta), gamma_surf=float(gamma_surf),
                      gamma_area=_as_int("gamma_area", gamma_area), gamma_bend=float(gamma_bend),
                      gamma_in=_as_int("gamma_in", gamma_in), beta_in=_as_int("beta_in", beta_in),
                      psi_in_v=tuple(float(v) for v in psi_in_v),
                      gamma_out=_as_int("gamma_out", gamma_out), beta_out=_as_int("beta_out", beta_out),
                      psi_out_v=tuple(float(v) for v in psi_out_v), lamda=_as_int("lamda", lamda),
                      S1=float(S1), S2=float(S2), S3=float(S3), S4=float(S4),
                      C1=float(C1), C2=float(C2), C3=float(C3), dt=float(dt), T=float(T), Nt=Nt,
                      Nx=int(Nx), Ny=int(Ny), Lx=float(Lx), Ly=float(Ly), dx=dx, dy=dy,
                      tol=float(tol), goal=str(goal), spe_len=spe_len, ndim=ndim, Nz=Nz_, Lz=Lz_, dz=dz)

    def with_(self, **kw) -> "Config":
        return replace(self, **kw)

    @property
    def Nxh(self) -> int:
        return self.Nx // 2 + 1


@dataclass(frozen=True)
class BDFCoeff:
    """Types.jl:407-417."""
    a: float
    b: float
    c: float


def bdf_coeff(n: int) -> BDFCoeff:
    """Types.jl:426: BDF1 for n == 1, BDF2 otherwise."""
    return BDFCoeff(1.0, -1.0, 0.0) if n == 1 else BDFCoeff(1.5, -2.0, 0.5)


def stabilisers(epsilon, gamma_bend, gamma_in, gamma_out):
    """S1..S4 as set_para_base defines them (Init.jl:40-43)."""
    S1 = (3 * SQ2 / 8) * epsilon * gamma_bend
    S2 = 0.0
    S3 = 1.0 * gamma_bend / epsilon ** 3
    S4 = 1.0 * min(gamma_in, gamma_out)
    return S1, S2, S3, S4


def set_para_base(dt: float, T: float) -> Config:
    """The reference's hard-coded parameters (Init.jl:8-69): N=1, 256^2 on the
    unit square, eps=0.01, gamma_in=gamma_out=1e5, goal :s."""
    eps, gb, gin, gout = 0.01, 0.1, 1e5, 1e5
    S1, S2, S3, S4 = stabilisers(eps, gb, gin, gout)
    return Config.make(N=1, epsilon=eps, M_phi=1.0, M0_psi=0.5, eta=1.0, gamma_surf=1.0,
                       gamma_area=5e4, gamma_bend=gb, gamma_in=gin, beta_in=0.0, psi_in_v=[0.1],
                       gamma_out=gout, beta_out=0.0, psi_out_v=[0.8], lamda=1.0,
                       S1=S1, S2=S2, S3=S3, S4=S4, C1=1.0, C2=5.0e5, C3=1.0e6,
                       dt=dt, T=T, Nx=256, Ny=256, Lx=1.0, Ly=1.0, tol=1e-12, goal="s")


# ---------------------------------------------------------------------------
# Seeded RNG: SplitMix64 (Steele, Lea, Flood 2014).  The reference never seeds
# (Init.jl:151-152), so the workloads fix their own counter-based stream.
# ---------------------------------------------------------------------------
class SplitMix64:
    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.state = seed & self.MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & self.MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))


# ---------------------------------------------------------------------------
# Workloads
# ---------------------------------------------------------------------------
@dataclass
class Vesicle:
    centre: tuple
    semi_axes: tuple       # (ax, ay[, az])
    c_in: float


@dataclass
class Workload:
    name: str
    conf: Config
    vesicles: List[Vesicle]
    c_out: float
    steps: int
    t_nodes: Optional[np.ndarray] = None
    notes: str = ""

    def initial_fields(self):
        return initial_fields(self.conf, self.vesicles, self.c_out)


def _grid(conf):
    x = np.arange(conf.Nx) * conf.dx
    y = np.arange(conf.Ny) * conf.dy
    if conf.ndim == 2:
        return np.meshgrid(x, y, indexing="xy")            # (Ny, Nx)
    z = np.arange(conf.Nz) * conf.dz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")         # (Nz, Ny, Nx)
    return X, Y, Z


def _min_image(d, L):
    return d - L * np.round(d / L)


def initial_fields(conf: Config, vesicles: List[Vesicle], c_out: float):
    """phi_n = tanh((rho_n(theta) - r) / (sqrt(2) eps)), +1 inside (Init.jl:113),
    with rho the polar radius of an axis-aligned ellipse (ellipsoid) and r the
    periodic minimum-image distance to the centre.  psi_n = (1+phi_n)/2*c_in +
    (1-phi_n)/2*c_out.  Arrays are in memory order: (N, Ny, Nx) / (N, Nz, Ny, Nx)."""
    eps = conf.epsilon
    g = _grid(conf)
    Ls = (conf.Lx, conf.Ly, conf.Lz)
    phis, psis = [], []
    for v in vesicles:
        d = [_min_image(g[i] - v.centre[i], Ls[i]) for i in range(conf.ndim)]
        r = np.sqrt(sum(di ** 2 for di in d))
        # rho(dir) = 1 / sqrt(sum (n_i / a_i)^2), n = d / r
        with np.errstate(divide="ignore", invalid="ignore"):
            s = np.sqrt(sum((d[i] / v.semi_axes[i]) ** 2 for i in range(conf.ndim)))
            rho = np.where(r > 0, r / np.where(s > 0, s, 1.0), v.semi_axes[0])
        phi = np.tanh((rho - r) / (SQ2 * eps))
        phis.append(phi)
        psis.append(0.5 * (1 + phi) * v.c_in + 0.5 * (1 - phi) * c_out)
    return np.ascontiguousarray(np.stack(phis)), np.ascontiguousarray(np.stack(psis))


# Parameters the benchmark workloads leave open are fixed here once and used
# identically by the oracle and the GPU: M_phi=1, M0_psi=0.5, eta=1, lamda=1,
# beta=0 (the reference defaults, Init.jl:46-62); gamma_in=gamma_out=1000 so that
# S4 = 1000; C1..C3 sized so that W_i + C_i > 0 on [0, 2pi]^d; S2 from s2_for
# (0, the reference default, unless the explicit bending term needs it).
GAMMA_IO = 1000


def bending_ratio(eps, gamma_bend, dt, a=1.5):
    """rho = c2 / (2 sqrt(S1 a / dt)): the explicit k^2 coefficient of the linearised
    bending derivative, c2 = gamma_b (3 sqrt 2 / (8 eps)) * 2 (3 phi^2 - 1) / eps at
    phi = +-1 = (3 sqrt 2 / 2) gamma_b / eps (get_H2, Utils.jl:208-226), against the
    largest damping the implicit S1 k^4 + a/dt part gives to a k^2 term (S1 of
    set_para_base, Init.jl:40).  Measured on the GPU (FP64): two-vesicle 1024^2
    (rho 0.91) and the 512^2 ensemble members with rho >= 0.53 let mu grow ~1.6x per step,
    R1 and Q collapse and a 1e-10 perturbation reach 1e-1; every instance with
    rho <= 0.512 stays bounded."""
    S1 = (3 * SQ2 / 8) * eps * gamma_bend
    c2 = (3 * SQ2 / 2) * gamma_bend / eps
    return c2 / (2 * math.sqrt(S1 * a / dt))


def s2_for(eps, gamma_bend, dt):
    """S2 stabiliser of a workload: the reference's S2 = 0 (Init.jl:42) where
    bending_ratio < 0.5, else S2 = c2 / 2 = (3 sqrt 2 / 4) gamma_b / eps, which
    moves half of the explicit k^2 bending term to the implicit side (the stabilised
    BDF2 bound) and keeps those workloads bounded (1024^2: Q 0.93 and a 1e-10
    perturbation at 1.3e-10 after 500 steps)."""
    if bending_ratio(eps, gamma_bend, dt) < 0.5:
        return 0.0
    return (3 * SQ2 / 4) * gamma_bend / eps


def _conf2d(N, n, eps, dt, steps, gamma_bend, c_in, c_out, gamma_area, C2, C3, tol=1e-10):
    S1, S2, S3, S4 = stabilisers(eps, gamma_bend, GAMMA_IO, GAMMA_IO)
    S2 = s2_for(eps, gamma_bend, dt)
    return Config.make(N=N, epsilon=eps, M_phi=1.0, M0_psi=0.5, eta=1.0, gamma_surf=1.0,
                       gamma_area=gamma_area, gamma_bend=gamma_bend, gamma_in=GAMMA_IO, beta_in=0,
                       psi_in_v=c_in, gamma_out=GAMMA_IO, beta_out=0, psi_out_v=[c_out] * N,
                       lamda=1, S1=S1, S2=S2, S3=S3, S4=S4, C1=1.0, C2=C2, C3=C3,
                       dt=dt, T=dt * steps, Nx=n, Ny=n, Lx=2 * math.pi, Ly=2 * math.pi, tol=tol)


def _c2_for(eps, N, L=2 * math.pi, dim=2, gamma_bend=0.1):
    """C2 > N * (1/2) S3 |Omega| bounds -W2 (the S3 phi^2 term dominates, phi^2 <= 1)."""
    S3 = gamma_bend / eps ** 3
    vol = L ** dim
    return float(10 ** math.ceil(math.log10(0.5 * S3 * vol * N * 1.5)))


def _c3_for(cmax, N, L=2 * math.pi, dim=2):
    """C3 > N * (1/2) S4 |Omega| max(psi)^2 bounds -W3."""
    vol = L ** dim
    return float(10 ** math.ceil(math.log10(0.5 * GAMMA_IO * vol * cmax ** 2 * N * 1.5)))


def _random_vesicles(rng, n, L, dim, rmin, rmax, amax, cin_lo, cin_hi, eps, max_tries=10000):
    ves = []
    tries = 0
    while len(ves) < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("could not place non-overlapping vesicles")
        c = tuple(rng.uniform(0, L) for _ in range(dim))
        r = rng.uniform(rmin, rmax)
        asp = rng.uniform(1.0, amax)
        axes = (r * asp,) + (r,) * (dim - 1)
        cin = rng.uniform(cin_lo, cin_hi)
        ok = True
        for v in ves:
            dd = math.sqrt(sum(_min_image(np.array(c[i] - v.centre[i]), L) ** 2 for i in range(dim)))
            if dd <= max(axes) + max(v.semi_axes) + 4 * eps:
                ok = False
                break
        if ok:
            ves.append(Vesicle(c, axes, cin))
    return ves


def workload_c1(n=128, steps=200) -> Workload:
    """Single ellipse, semi-axes 1.2 (x) and 0.8 (y) centred at (pi, pi), eps=0.08,
    gamma_s=1, gamma_b=0.1, gamma_area=1000, dt=1e-4, BDF1 then BDF2, FP64,
    BiCGSTAB rtol 1e-10."""
    eps = 0.08
    ves = [Vesicle((math.pi, math.pi), (1.2, 0.8), 1.2)]
    conf = _conf2d(1, n, eps, 1e-4, steps, 0.1, [1.2], 1.0, 1000, _c2_for(eps, 1), _c3_for(1.2, 1))
    return Workload("single_vesicle_128", conf, ves, 1.0, steps)


def workload_c2(n=512, steps=1000, seed=2001) -> Workload:
    """Four seeded vesicles on 512^2, eps=0.03, dt=5e-5."""
    eps = 0.03
    rng = SplitMix64(seed)
    ves = _random_vesicles(rng, 4, 2 * math.pi, 2, 0.4, 0.7, 1.6, 0.8, 1.4, eps)
    conf = _conf2d(4, n, eps, 5e-5, steps, 0.1, [v.c_in for v in ves], 1.0, 1000,
                   _c2_for(eps, 4), _c3_for(1.4, 4))
    return Workload("four_vesicles_512", conf, ves, 1.0, steps)


def workload_c3(n=1024, steps=100, seed=3001) -> Workload:
    """Two seeded vesicles on 1024^2, eps=0.015, dt=2e-5 (learned-preconditioner case)."""
    eps = 0.015
    rng = SplitMix64(seed)
    ves = _random_vesicles(rng, 2, 2 * math.pi, 2, 0.4, 0.7, 1.6, 0.8, 1.4, eps)
    conf = _conf2d(2, n, eps, 2e-5, steps, 0.1, [v.c_in for v in ves], 1.0, 1000,
                   _c2_for(eps, 2), _c3_for(1.4, 2))
    return Workload("two_vesicles_1024", conf, ves, 1.0, steps)


def workload_c4(n=128, steps=500, seed=4001) -> Workload:
    """Two seeded ellipsoids on 128^3 over [0, 2pi]^3, eps=0.06, dt=1e-4,
    spectral psi solve with constant mobility M0_psi."""
    eps = 0.06
    rng = SplitMix64(seed)
    L = 2 * math.pi
    ves = []
    tries = 0
    while len(ves) < 2:
        tries += 1
        c = tuple(rng.uniform(0, L) for _ in range(3))
        axes = tuple(rng.uniform(0.6, 1.0) for _ in range(3))
        cin = rng.uniform(1.1, 1.3)
        ok = all(math.sqrt(sum(_min_image(np.array(c[i] - v.centre[i]), L) ** 2 for i in range(3)))
                 > max(axes) + max(v.semi_axes) + 4 * eps for v in ves)
        if ok:
            ves.append(Vesicle(c, axes, cin))
    S1, S2, S3, S4 = stabilisers(eps, 0.1, GAMMA_IO, GAMMA_IO)
    conf = Config.make(N=2, epsilon=eps, M_phi=1.0, M0_psi=0.5, eta=1.0, gamma_surf=1.0,
                       gamma_area=1000, gamma_bend=0.1, gamma_in=GAMMA_IO, beta_in=0,
                       psi_in_v=[v.c_in for v in ves], gamma_out=GAMMA_IO, beta_out=0,
                       psi_out_v=[1.0, 1.0], lamda=1, S1=S1, S2=S2, S3=S3, S4=S4,
                       C1=1.0, C2=_c2_for(eps, 2, dim=3), C3=_c3_for(1.3, 2, dim=3),
                       dt=1e-4, T=1e-4 * steps, Nx=n, Ny=n, Lx=L, Ly=L, tol=1e-10, Nz=n, Lz=L)
    return Workload("two_ellipsoids_128cubed", conf, ves, 1.0, steps)


def workload_c5(steps=500, seed=5001, sizes=(128, 256, 384, 512), per_size=16) -> List[Workload]:
    """64 independent 2D instances (16 each of 128^2, 256^2, 384^2, 512^2), each
    with 1-3 seeded vesicles, c_in ~ U[0.8,1.4], gamma_bend log-uniform on
    [0.02, 0.5]; eps = 3 dx, dt = 5e-5."""
    out = []
    k = 0
    for n in sizes:
        for i in range(per_size):
            rng = SplitMix64(seed * 1000 + k)
            k += 1
            nv = 1 + int(rng.uniform(0, 3) // 1)
            nv = min(nv, 3)
            gb = math.exp(rng.uniform(math.log(0.02), math.log(0.5)))
            eps = 3 * (2 * math.pi / n)
            ves = _random_vesicles(rng, nv, 2 * math.pi, 2, 0.4, 0.7, 1.6, 0.8, 1.4, eps)
            conf = _conf2d(nv, n, eps, 5e-5, steps, gb, [v.c_in for v in ves], 1.0, 1000,
                           _c2_for(eps, nv, gamma_bend=gb), _c3_for(1.4, nv))
            out.append(Workload(f"ensemble_{n}_{i}", conf, ves, 1.0, steps))
    return out


WORKLOADS = {
    "single_vesicle_128": workload_c1,
    "four_vesicles_512": workload_c2,
    "two_vesicles_1024": workload_c3,
    "two_ellipsoids_128cubed": workload_c4,
}
