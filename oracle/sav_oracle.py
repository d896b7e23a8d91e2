"""FP64 CPU oracle for the seven-step pseudo-spectral SAV timestep.

TEST INFRASTRUCTURE ONLY.  This module is the checker against which the HIP
library is compared.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product path never does.

It restates, function by function, the reference Julia code
(Forstylime/Julia-Vesicles-growth-and-shrinkage) in numpy/scipy:

* ``src/SpectralUtils.jl`` / ``src/SpectralUtils_3d.jl`` -> :class:`Ops`
* ``src/Utils.jl`` / ``src/Utils_3d.jl``                 -> ``f_surf`` ... ``get_MG``
* ``src/Solvers.jl`` / ``src/Solvers_3d.jl``             -> ``solve_step1`` ... ``solve_step7``
* Krylov.jl 0.10.6 ``bicgstab`` (third-party, not vendored in the reference;
  restated from its published algorithm)                 -> :func:`bicgstab`
* ``src/simulation.jl:22-110,146-148``                     -> :func:`init_sav`, :func:`step`

Array layout.  Julia stores ``phi`` as ``(Nx, Ny, N)`` column-major (x fastest).
The oracle stores the same memory as a C-order array of shape ``(N, Ny, Nx)``
(3D: ``(N, Nz, Ny, Nx)``), so ``julia_array == oracle_array.T``.  Half spectra
are ``(N, Ny, Nx//2+1)``; the x axis is the halved one, exactly as FFTW's
``rfft(u, (1,2))`` halves Julia's first dimension.

Transforms are scipy.fft (pocketfft).  Its multi-dimensional c2r performs the
inverse c2c over the non-halved axes first and then a 1-D c2r along x that
reads only the real part of the DC and Nyquist bins -- the same semantics as
FFTW's ``irfft`` on non-Hermitian half spectra.  ``tests/test_oracle_kat.py``
pins this against a brute-force sum.

What pins each function (tests/test_oracle_kat.py, tests/test_golden.py):

* ``Ops`` wavenumbers and D1/Lap/Bih: the reference's own known-answer tests
  (test/test_Types_Ops.jl:71-136, Nyquist signs included).
* ``rfft``/``irfft``: the reference's round-trip and derivative tests
  (test_Types_Ops.jl:142-314), a brute-force DFT, and the brute-force c2r definition
  on non-Hermitian input.
* ``bdf_coeff``: test_Types_Ops.jl:369-395.
* ``f_surf``/``calculate_area``, ``f_bend``/``get_omega``: closed forms (tanh stripe,
  constant phi).
* ``get_W2``/``get_H2``, ``get_W3``/``get_H3``/``get_MG``: Gateaux derivatives
  (H2 R2 = dW2/dphi exactly when gamma_surf = 0; H3, G exact derivatives of W3), and
  the whole of H2 and W2 (surface and bending terms) in closed form on a tanh stripe of
  width delta != eps.
* ``get_H1``/``get_W1``: H1 is not an exact derivative (Utils.jl:198-206); it is pinned by
  its closed form on the same stripe with A0 != A (omega = (phi^3 - phi)(1/eps -
  eps/delta^2), area per interface delta/(2 eps) + sqrt2 eps/(3 delta)), and by the
  constant-phi case.
* ``bicgstab``: dense ``numpy.linalg.solve`` of the assembled operator on 8x8 grids,
  the constant-mobility case (a spectral divide), a zero right-hand side, itmax
  reporting, and at every iteration (also when stopped at itmax) the identities that
  define the recurrence: r_k = M(b - A x_k), <c, s_k> = 0, <t_k, r_k> = 0,
  rho_{k+1} = <c, r_k>.  Krylov.jl's own floating-point operation order is not
  reproduced bit for bit (the Julia package is absent).
* ``AltOps``: the same transforms by a second route, for the round-off floor between
  two FFT implementations (tests/test_oracle_fft_floor.py).
* ``solve_step1`` .. ``solve_step7``, ``step``: the coupled BDF-SAV scheme that the
  seven steps decompose, equation by equation at the new state (phi, mu, psi, nu, R1-R3,
  the intermediate velocity, Q, the projection), to round-off in 2D and 3D; planting a
  swapped R index, a dropped term or a wrong sign in any step leaves an O(1e-6..1)
  residual there; and invariants (spectral div u = 0, modified energy non-increasing,
  mean psi conserved, BDF1 first step).  A trajectory of the Julia code itself is not
  available to compare against (no Julia here, no golden data in the reference).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.fft as sfft

SQ2 = math.sqrt(2.0)


class SavFatal(RuntimeError):
    """Mirrors ``error("FATAL: W + C < 0 ...")`` in src/Utils.jl:162,180,188."""


# ---------------------------------------------------------------------------
# Operators  (src/SpectralUtils.jl:36-118, src/SpectralUtils_3d.jl:36-124)
# ---------------------------------------------------------------------------
class Ops:
    """Wavenumbers, spectral multipliers and transforms for one grid.

    kx  = 2*pi*rfftfreq(Nx, Nx/Lx)  -> Nyquist +pi*Nx/Lx   (SpectralUtils.jl:42)
    kx0 = 2*pi*fftfreq(Nx, Nx/Lx)   -> Nyquist -pi*Nx/Lx   (SpectralUtils.jl:43)
    ky  = 2*pi*fftfreq(Ny, Ny/Ly)   -> Nyquist -pi*Ny/Ly   (SpectralUtils.jl:44)
    """

    def __init__(self, conf, workers: int = 1):
        self.ndim = int(getattr(conf, "ndim", 2))
        self.Nx, self.Ny = int(conf.Nx), int(conf.Ny)
        self.Nz = int(getattr(conf, "Nz", 1)) if self.ndim == 3 else 1
        self.workers = workers
        Nx, Ny, Nz = self.Nx, self.Ny, self.Nz
        kx = 2 * np.pi * np.fft.rfftfreq(Nx, d=conf.Lx / Nx)
        kx0 = 2 * np.pi * np.fft.fftfreq(Nx, d=conf.Lx / Nx)
        ky = 2 * np.pi * np.fft.fftfreq(Ny, d=conf.Ly / Ny)
        if self.ndim == 2:
            self.kx = kx[None, :]
            self.kx0 = kx0[None, :]
            self.ky = ky[:, None]
            self.K = (self.kx, self.ky)
            self.K_full = (self.kx0, self.ky)
            self.axes = (-2, -1)
            self.shape = (Ny, Nx)
        else:
            kz = 2 * np.pi * np.fft.fftfreq(Nz, d=conf.Lz / Nz)
            self.kx = kx[None, None, :]
            self.kx0 = kx0[None, None, :]
            self.ky = ky[None, :, None]
            self.kz = kz[:, None, None]
            self.K = (self.kx, self.ky, self.kz)
            self.K_full = (self.kx0, self.ky, self.kz)
            self.axes = (-3, -2, -1)
            self.shape = (Nz, Ny, Nx)
        # D1 = i*k ; D1_full uses kx0 (SpectralUtils.jl:46-47)
        self.D1 = tuple(1j * k for k in self.K)
        self.D1_full = tuple(1j * k for k in self.K_full)
        k2 = sum(k ** 2 for k in self.K)
        self.Laplacian = -k2                                   # :49
        self.Laplacian_full = -sum(k ** 2 for k in self.K_full)  # :50
        self.Biharmonic = k2 ** 2                              # :51
        self.inv_s = 1.0 / (Nx * Ny * Nz)

    # to_spectral!  (SpectralUtils.jl:132-135): unnormalised r2c
    def rfft(self, u):
        return sfft.rfftn(u, axes=self.axes, workers=self.workers)

    # to_physical!  (SpectralUtils.jl:144-148): c2r scaled by 1/(Nx*Ny[*Nz])
    def irfft(self, u_hat):
        return sfft.irfftn(u_hat, s=self.shape, axes=self.axes, workers=self.workers)

    # fft_plan_full / ifft_plan_full (SpectralUtils.jl:69-71)
    def fft(self, u):
        return sfft.fftn(u, axes=self.axes, workers=self.workers)

    def ifft(self, u_hat):
        return sfft.ifftn(u_hat, axes=self.axes, workers=self.workers)

    def dvol(self, conf):
        v = conf.dx * conf.dy
        if self.ndim == 3:
            v *= conf.dz
        return v


class AltOps(Ops):
    """The same transforms by a second route (2D), for measuring the round-off floor that
    separates two correct FFT implementations: r2c through the full complex FFT, c2r
    through a c2c over y and a Hermitian extension along x inverted by a complex FFT (Im of
    the DC and Nyquist bins dropped, the c2r semantics), full complex transforms through
    the conjugation identity."""

    def rfft(self, u):
        return sfft.fftn(np.asarray(u, dtype=complex), axes=self.axes, workers=self.workers)[..., : self.Nx // 2 + 1]

    def irfft(self, u_hat):
        nx = self.Nx
        y = sfft.ifftn(u_hat, axes=self.axes[:-1], workers=self.workers).copy()
        y[..., 0] = y[..., 0].real
        y[..., nx // 2] = y[..., nx // 2].real
        full = np.concatenate([y, np.conj(y[..., 1: nx // 2][..., ::-1])], axis=-1)
        return sfft.ifft(full, axis=-1, workers=self.workers).real

    def fft(self, u):
        return np.conj(sfft.ifftn(np.conj(u), axes=self.axes, workers=self.workers, norm="forward"))

    def ifft(self, u_hat):
        return np.conj(sfft.fftn(np.conj(u_hat), axes=self.axes, workers=self.workers, norm="forward"))


def _bcast_n(v, ndim):
    """reshape(v, 1, 1, :) in Julia == v[:, None, None] in the oracle layout."""
    v = np.asarray(v, dtype=np.float64)
    return v.reshape((-1,) + (1,) * ndim)


# ---------------------------------------------------------------------------
# Pointwise helpers  (src/Utils.jl:8-29)
# ---------------------------------------------------------------------------
def g(phi):
    return 0.25 * (phi ** 2 - 1.0) ** 2


def g_prime(phi):
    return phi * (phi ** 2 - 1.0)


def p_phi(phi):
    return -0.5 * phi ** 3 + 1.5 * phi


def p_phi_prime(phi):
    return -1.5 * phi ** 2 + 1.5


def get_M_psi(phi, conf):
    return 0.1 * (1.0 - conf.M0_psi * (phi ** 2 - 1.0) ** 2)


def integrate(f, conf, ops):
    return float(np.sum(f)) * ops.dvol(conf)


def integrate_sq(f, conf, ops):
    return float(np.sum(np.abs(f) ** 2)) * ops.dvol(conf)


# ---------------------------------------------------------------------------
# Energy densities and functionals  (src/Utils.jl:36-249)
# ---------------------------------------------------------------------------
def f_surf(phi, ops, conf):
    """Utils.jl:36-42.  Only d/dx and d/dy, also in 3D (Utils_3d.jl:36-42)."""
    phi_hat = ops.rfft(phi)
    gx = ops.irfft(ops.D1[0] * phi_hat)
    gy = ops.irfft(ops.D1[1] * phi_hat)
    eps = conf.epsilon
    return (3 * SQ2 / 4) * (g(phi) / eps) + eps * (gx ** 2 + gy ** 2) / 2


def f_bend(phi, ops, conf):
    """Utils.jl:45-50."""
    phi_hat = ops.rfft(phi)
    eps = conf.epsilon
    lap = ops.irfft(ops.Laplacian * phi_hat)
    return (3 * SQ2 / (16 * eps)) * (g_prime(phi) / eps - eps * lap) ** 2


def f_osmotic(phi, psi, conf, ops):
    """Utils.jl:53-60."""
    P = p_phi(phi)
    pin = _bcast_n(conf.psi_in_v[: phi.shape[0]], ops.ndim)
    pout = _bcast_n(conf.psi_out_v[: phi.shape[0]], ops.ndim)
    f_in = 0.5 * conf.gamma_in * (psi - pin) ** 2 + conf.beta_in
    f_out = 0.5 * conf.gamma_out * (psi - pout) ** 2 + conf.beta_out
    return ((1 + P) / 2) * f_in + ((1 - P) / 2) * f_out


def calculate_area(phi, ops, conf):
    """Utils.jl:66-73: per-vesicle integral of f_surf (a surface functional)."""
    surf = f_surf(phi, ops, conf)
    return np.array([integrate(surf[n], conf, ops) for n in range(phi.shape[0])])


def get_omega(phi, ops, conf):
    """Utils.jl:148-153."""
    phi_hat = ops.rfft(phi)
    eps = conf.epsilon
    lap = ops.irfft(ops.Laplacian * phi_hat)
    return (phi ** 3 - phi) / eps - eps * lap


def get_W1(phi, ops, conf, A0):
    """Utils.jl:159-164."""
    A = calculate_area(phi, ops, conf)
    W1 = float(np.sum(0.5 * conf.gamma_area * (A - A0) ** 2))
    if W1 + conf.C1 < 0:
        raise SavFatal(f"FATAL: W1 + C1 < 0, W1 = {W1}")
    return W1


def get_W2(phi, ops, conf):
    """Utils.jl:166-182 (3D: only d/dx, d/dy in the S2 term, Utils_3d.jl:166-182)."""
    phi_hat = ops.rfft(phi)
    F_surf = conf.gamma_surf * integrate(f_surf(phi, ops, conf), conf, ops)
    F_bend = conf.gamma_bend * integrate(f_bend(phi, ops, conf), conf, ops)
    lap = ops.irfft(ops.Laplacian * phi_hat)
    gx = ops.irfft(ops.D1[0] * phi_hat)
    gy = ops.irfft(ops.D1[1] * phi_hat)
    E2 = integrate(0.5 * conf.S1 * lap ** 2 + 0.5 * conf.S2 * (gx ** 2 + gy ** 2)
                   + 0.5 * conf.S3 * phi ** 2, conf, ops)
    W2 = F_surf + F_bend - E2
    if W2 + conf.C2 < 0:
        raise SavFatal(f"FATAL: W2 + C2 < 0, W2 = {W2}")
    return W2


def get_W3(phi, psi, conf, ops):
    """Utils.jl:184-190."""
    W_osm = integrate(f_osmotic(phi, psi, conf, ops), conf, ops)
    E3 = 0.5 * conf.S4 * integrate_sq(psi, conf, ops)
    W3 = W_osm - E3
    if W3 + conf.C3 < 0:
        raise SavFatal(f"FATAL: W3 + C3 < 0, W3 = {W3}")
    return W3


def get_H1(phi, ops, conf, A0):
    """Utils.jl:198-206."""
    W1 = get_W1(phi, ops, conf, A0)
    R1 = math.sqrt(W1 + conf.C1)
    A = calculate_area(phi, ops, conf)
    omega = get_omega(phi, ops, conf)
    factor = conf.gamma_area * (A - A0) * (3 * SQ2 / 4)
    return (_bcast_n(factor, ops.ndim) * omega) / R1


def get_H2(phi, ops, conf):
    """Utils.jl:208-226."""
    phi_hat = ops.rfft(phi)
    W2 = get_W2(phi, ops, conf)
    R2 = math.sqrt(W2 + conf.C2)
    eps = conf.epsilon
    omega = get_omega(phi, ops, conf)
    omega_hat = ops.rfft(omega)
    lap_omega = ops.irfft(ops.Laplacian * omega_hat)
    H2_sb = (conf.gamma_surf * (3 * SQ2 / 4) * omega
             + conf.gamma_bend * (3 * SQ2 / (8 * eps))
             * (omega * (3 * phi ** 2 - 1) / eps - eps * lap_omega))
    bih = ops.irfft(ops.Biharmonic * phi_hat)
    lap = ops.irfft(ops.Laplacian * phi_hat)
    H2_lin = conf.S1 * bih - conf.S2 * lap + conf.S3 * phi
    return (H2_sb - H2_lin) / R2


def get_H3(phi, psi, conf, ops):
    """Utils.jl:228-236."""
    W3 = get_W3(phi, psi, conf, ops)
    R3 = math.sqrt(W3 + conf.C3)
    pin = _bcast_n(conf.psi_in_v[: phi.shape[0]], ops.ndim)
    pout = _bcast_n(conf.psi_out_v[: phi.shape[0]], ops.ndim)
    f_in = 0.5 * conf.gamma_in * (psi - pin) ** 2 + conf.beta_in
    f_out = 0.5 * conf.gamma_out * (psi - pout) ** 2 + conf.beta_out
    return (0.5 * p_phi_prime(phi) * (f_in - f_out)) / R3


def get_MG(phi, psi, conf, ops):
    """Utils.jl:238-249."""
    W3 = get_W3(phi, psi, conf, ops)
    R3 = math.sqrt(W3 + conf.C3)
    P = p_phi(phi)
    pin = _bcast_n(conf.psi_in_v[: phi.shape[0]], ops.ndim)
    pout = _bcast_n(conf.psi_out_v[: phi.shape[0]], ops.ndim)
    G1 = 0.5 * (1 + P) * conf.gamma_in * (psi - pin) + 0.5 * (1 - P) * conf.gamma_out * (psi - pout)
    G2 = conf.S4 * psi
    return (G1 - G2) / R3


# ---------------------------------------------------------------------------
# State  (src/Types.jl:167-265)
# ---------------------------------------------------------------------------
@dataclass
class State:
    phi: np.ndarray
    phi_hat: np.ndarray
    psi: np.ndarray
    psi_hat: np.ndarray
    u: np.ndarray
    u_hat: np.ndarray
    p: np.ndarray
    p_hat: np.ndarray
    mu: np.ndarray
    mu_hat: np.ndarray
    nu: np.ndarray
    nu_hat: np.ndarray
    Q: float = 1.0
    R1: float = 1.0
    R2: float = 1.0
    R3: float = 1.0
    A0: np.ndarray = field(default_factory=lambda: np.zeros(1))
    dt: float = 0.0

    def copy(self):
        return State(**{k: (v.copy() if isinstance(v, np.ndarray) else v)
                        for k, v in self.__dict__.items()})


def update_state(dest: State, src: State):
    """Types.jl:243-265: copies the 12 arrays and Q, R1..R3; not A0, not dt."""
    for k in ("phi", "phi_hat", "psi", "psi_hat", "u", "u_hat", "p", "p_hat",
              "mu", "mu_hat", "nu", "nu_hat"):
        getattr(dest, k)[...] = getattr(src, k)
    dest.Q, dest.R1, dest.R2, dest.R3 = src.Q, src.R1, src.R2, src.R3
    return dest


def bdf_coeff(n: int):
    """Types.jl:414-426: BDF1 (1,-1,0) for n == 1, else BDF2 (1.5,-2,0.5)."""
    return (1.0, -1.0, 0.0) if n == 1 else (1.5, -2.0, 0.5)


def state_from_physical(phi, psi, ops, ndim):
    """Builds a FieldState the way generate_initial_condition does (Init.jl:200-216):
    spectra are rfft of the physical fields, u = p = mu = nu = 0, Q=R1=R2=R3=1."""
    sp = phi.shape[1:]
    ncomp = 2 if ndim == 2 else 3
    u = np.zeros((ncomp,) + sp)
    p = np.zeros(sp)
    z = np.zeros_like(phi)
    return State(phi=phi.copy(), phi_hat=ops.rfft(phi), psi=psi.copy(), psi_hat=ops.rfft(psi),
                 u=u, u_hat=ops.rfft(u), p=p, p_hat=ops.rfft(p),
                 mu=z.copy(), mu_hat=ops.rfft(z), nu=z.copy(), nu_hat=ops.rfft(z),
                 Q=1.0, R1=1.0, R2=1.0, R3=1.0, A0=np.zeros(phi.shape[0]), dt=0.0)


# ---------------------------------------------------------------------------
# Krylov.jl 0.10.6 bicgstab restated (left preconditioner M applied by mul!,
# shadow vector c = b, x0 = 0).  Called from src/Solvers.jl:73-75.
# ---------------------------------------------------------------------------
@dataclass
class KrylovStats:
    niter: int = 0
    solved: bool = False
    status: str = ""
    residuals: list = field(default_factory=list)
    alphas: list = field(default_factory=list)
    omegas: list = field(default_factory=list)


def _kdot(x, y):
    # @kdot(n, x, y) = dot(x, y) = sum(conj(x) .* y)
    return complex(np.vdot(x, y))


def _knrm2(x):
    return float(np.linalg.norm(x))


def bicgstab(A, b, M, atol, rtol, itmax, history=True, callback=None):
    """BiCGSTAB of van der Vorst (1992) in the form Krylov.jl 0.10.6 implements.

    r0 = M b, p = r0, c = b, rho = <c, r>, eps = atol + rtol*||r0||.
    Stops on ||r|| <= eps (or ||r|| + 1 <= 1), iter >= itmax, or alpha == 0/NaN.
    ``callback(k, dict)``, if given, sees every iteration's vectors (x, r, s, t, v, p) and
    scalars (alpha, omega, rho, next_rho) after the iteration, for the identity tests.
    """
    x = np.zeros_like(b)
    r = M(b)
    p = r.copy()
    c = b
    stats = KrylovStats()
    rnorm = _knrm2(r)
    if history:
        stats.residuals.append(rnorm)
    if rnorm == 0.0:
        stats.niter, stats.solved, stats.status = 0, True, "x = 0 is a zero-residual solution"
        return x, stats
    it = 0
    eps = atol + rtol * rnorm
    next_rho = _kdot(c, r)
    if next_rho == 0:
        stats.niter, stats.solved, stats.status = 0, False, "Breakdown bᴴc = 0"
        return x, stats
    solved = rnorm <= eps
    tired = it >= itmax
    breakdown = False
    while not (solved or tired or breakdown):
        it += 1
        rho = next_rho
        q = A(p)
        v = M(q)
        alpha = rho / _kdot(c, v)
        s = r - alpha * v
        x = x + alpha * p
        d = A(s)
        t = M(d)
        omega = _kdot(t, s) / _kdot(t, t)
        x = x + omega * s
        r = s - omega * t
        next_rho = _kdot(c, r)
        beta = (next_rho / rho) * (alpha / omega)
        p = p - omega * v
        p = r + beta * p
        rnorm = _knrm2(r)
        if callback is not None:
            callback(it, dict(x=x, r=r, s=s, t=t, v=v, p=p, c=c, alpha=alpha, omega=omega, rho=rho,
                              next_rho=next_rho))
        if history:
            stats.residuals.append(rnorm)
            stats.alphas.append(alpha)
            stats.omegas.append(omega)
        solved = (rnorm <= eps) or (rnorm + 1.0 <= 1.0)
        tired = it >= itmax
        breakdown = (alpha == 0) or (alpha != alpha)
    stats.niter = it
    stats.solved = bool(solved)
    stats.status = ("solution good enough given atol and rtol" if solved else
                    "breakdown αₖ == 0" if breakdown else
                    "maximum number of iterations exceeded")
    return x, stats


# ---------------------------------------------------------------------------
# The seven steps  (src/Solvers.jl, src/Solvers_3d.jl)
# ---------------------------------------------------------------------------
class Step1Cache:
    """Types.jl:358-385: 16 half spectra."""
    names = ("phi_11", "mu_11", "psi_11", "nu_11", "phi_12", "mu_12", "psi_12", "nu_12",
             "phi_13", "mu_13", "phi_14", "mu_14", "phi_21", "mu_21", "psi_21", "nu_21")


class Oracle:
    """The reference timestep on one grid.  ``atol``/``itmax`` default to the
    values hard-coded at Solvers.jl:73-75."""

    def __init__(self, conf, workers: int = 1, itmax: int = 20, atol: float = 1e-14):
        self.conf = conf
        self.ops = Ops(conf, workers)
        self.ndim = self.ops.ndim
        self.itmax = itmax
        self.atol = atol
        self.krylov_log = []     # KrylovStats of every psi solve (2D)
        self.warnings = []

    # -- helpers ---------------------------------------------------------
    def _L_phi(self):
        c, o = self.conf, self.ops
        return c.S1 * o.Biharmonic - c.S2 * o.Laplacian + c.S3

    def solve_psi(self, rhs_rhat, M_psi, M_max, a, dt):
        """Solvers.jl:39-81: full-complex matrix-free operator, diagonal
        preconditioner, Krylov bicgstab, truncation to the half spectrum."""
        o, c = self.ops, self.conf
        L_psi = c.S4
        a_dt = a / dt
        D1f = o.D1_full
        shape = (rhs_rhat.shape[0],) + o.shape

        def A(xv):
            xh = xv.reshape(shape)
            b1 = L_psi * xh
            acc = 0
            for d in range(o.ndim if o.ndim == 2 else 2):
                bd = D1f[d] * b1
                w = o.ifft(bd)
                w = M_psi * w
                acc = acc + D1f[d] * o.fft(w)
            return (a_dt * xh - acc).reshape(-1)

        P_inv_diag = 1.0 / (a_dt - M_max * o.Laplacian_full * L_psi)

        def Mp(rv):
            return (rv.reshape(shape) * P_inv_diag).reshape(-1)

        rhs_phys = o.irfft(rhs_rhat)
        rhs_full = o.fft(rhs_phys)
        sol, stats = bicgstab(A, rhs_full.reshape(-1), Mp, self.atol, c.tol, self.itmax)
        self.krylov_log.append(stats)
        if not stats.solved:
            self.warnings.append(f"BiCGSTAB not converged: {stats.status}")
        sol = sol.reshape(shape)
        nxh = o.Nx // 2 + 1
        return sol[..., :nxh].copy()

    def functionals(self, phi_star, psi_star, A0):
        """H1, H2, H3 and G at the extrapolated state (Utils.jl:198-249).
        solve_step2 recomputes exactly the same four fields (Solvers.jl:160-163)."""
        o, c = self.ops, self.conf
        H1 = get_H1(phi_star, o, c, A0)
        G = get_MG(phi_star, psi_star, c, o)
        H2 = get_H2(phi_star, o, c)
        H3 = get_H3(phi_star, psi_star, c, o)
        return H1, H2, H3, G

    # -- step 1 ------------------------------------------------------------
    def solve_step1(self, present, old, bdf):
        """Solvers.jl:14-137 (2D), Solvers_3d.jl:14-92 (3D)."""
        o, c = self.ops, self.conf
        dt = present.dt
        a, b, cc = bdf
        A0 = present.A0
        phi_star = 2.0 * present.phi - old.phi
        psi_star = 2.0 * present.psi - old.psi
        phi_star_hat = 2.0 * present.phi_hat - old.phi_hat
        psi_star_hat = 2.0 * present.psi_hat - old.psi_hat
        L_phi = self._L_phi()
        L_psi = c.S4
        lhs_phi = (a / dt) + c.M_phi * L_phi
        cache = {}
        if self.ndim == 2:
            M_psi = get_M_psi(phi_star, c)
            M_max = float(np.max(M_psi))

            def solve_psi(rhs):
                return self.solve_psi(rhs, M_psi, M_max, a, dt)
            M_mult = M_psi
        else:
            lhs_psi = (a / dt) - c.M0_psi * L_psi * o.Laplacian

            def solve_psi(rhs):
                return rhs / lhs_psi
            M_mult = c.M0_psi

        # 11
        rhs = -(b * present.phi_hat + cc * old.phi_hat) / dt
        cache["phi_11"] = rhs / lhs_phi
        cache["mu_11"] = L_phi * cache["phi_11"]
        rhs = -(b * present.psi_hat + cc * old.psi_hat) / dt
        cache["psi_11"] = solve_psi(rhs)
        cache["nu_11"] = L_psi * cache["psi_11"]
        # 12
        H1, H2, H3, G = self.functionals(phi_star, psi_star, A0)
        H1h = o.rfft(H1)
        cache["phi_12"] = (-c.M_phi * H1h) / lhs_phi
        cache["mu_12"] = L_phi * cache["phi_12"] + H1h
        Gh = o.rfft(G)
        rhs = 0
        for d in range(o.ndim):
            gd = o.irfft(o.D1[d] * Gh)
            rhs = rhs + o.D1[d] * o.rfft(M_mult * gd)
        cache["psi_12"] = solve_psi(rhs)
        cache["nu_12"] = L_psi * cache["psi_12"] + Gh
        # 13 & 14
        H2h = o.rfft(H2)
        H3h = o.rfft(H3)
        cache["phi_13"] = (-c.M_phi * H2h) / lhs_phi
        cache["phi_14"] = (-c.M_phi * H3h) / lhs_phi
        cache["mu_13"] = L_phi * cache["phi_13"] + H2h
        cache["mu_14"] = L_phi * cache["phi_14"] + H3h
        # 21 : gradients from the state spectra (Solvers.jl:118-131)
        u_star = 2.0 * present.u - old.u
        adv = sum(u_star[d] * o.irfft(o.D1[d] * phi_star_hat) for d in range(o.ndim))
        t = o.rfft(adv)
        cache["phi_21"] = -t / lhs_phi
        cache["mu_21"] = L_phi * cache["phi_21"]
        adv = sum(u_star[d] * o.irfft(o.D1[d] * psi_star_hat) for d in range(o.ndim))
        t = o.rfft(adv)
        cache["psi_21"] = solve_psi(-t)
        cache["nu_21"] = L_psi * cache["psi_21"]
        self._fields = (H1, H2, H3, G)
        return cache

    # -- step 2 ------------------------------------------------------------
    def solve_step2(self, present, old, cache, bdf, fields=None):
        """Solvers.jl:144-234.  ``fields`` lets the caller pass the H1..G of
        step 1, which step 2 recomputes from identical inputs."""
        o, c = self.ops, self.conf
        a, b, cc = bdf
        phi_star = 2.0 * present.phi - old.phi
        psi_star = 2.0 * present.psi - old.psi
        if fields is None:
            fields = self.functionals(phi_star, psi_star, present.A0)
        H1, H2, H3, G = fields
        dv = o.dvol(c)

        def int_dot(A, B):
            return float(np.dot(A.reshape(-1), B.reshape(-1))) * dv
        tphi = b * present.phi + cc * old.phi
        tpsi = b * present.psi + cc * old.psi
        U1 = -(b * present.R1 + cc * old.R1) / a + int_dot(H1, tphi) / (2.0 * a)
        U2 = -(b * present.R2 + cc * old.R2) / a + int_dot(H2, tphi) / (2.0 * a)
        U3 = -(b * present.R3 + cc * old.R3) / a + (int_dot(H3, tphi) + int_dot(G, tpsi)) / (2.0 * a)
        ph = {k: o.irfft(cache[k]) for k in ("phi_12", "phi_13", "phi_14", "psi_12",
                                             "phi_11", "psi_11", "phi_21", "psi_21")}
        L = np.zeros((3, 3))
        R1v = np.zeros(3)
        Hs = (H1, H2, H3)
        for i in range(3):
            for j, k in enumerate(("phi_12", "phi_13", "phi_14")):
                L[i, j] = -0.5 * int_dot(Hs[i], ph[k])
            L[i, i] += 1.0
        L[2, 2] = 1.0 - 0.5 * (int_dot(H3, ph["phi_14"]) + int_dot(G, ph["psi_12"]))
        R1v[0] = 0.5 * int_dot(H1, ph["phi_11"]) + U1
        R1v[1] = 0.5 * int_dot(H2, ph["phi_11"]) + U2
        R1v[2] = 0.5 * (int_dot(H3, ph["phi_11"]) + int_dot(G, ph["psi_11"])) + U3
        R2v = np.zeros(3)
        R2v[0] = 0.5 * int_dot(H1, ph["phi_21"])
        R2v[1] = 0.5 * int_dot(H2, ph["phi_21"])
        R2v[2] = 0.5 * (int_dot(H3, ph["phi_21"]) + int_dot(G, ph["psi_21"]))
        sol1 = np.linalg.solve(L, R1v)      # LAPACK LU with partial pivoting
        sol2 = np.linalg.solve(L.copy(), R2v)
        self._step2_system = (L.copy(), R1v.copy(), R2v.copy())
        return dict(R_11=sol1[0], R_21=sol1[1], R_31=sol1[2],
                    R_12=sol2[0], R_22=sol2[1], R_32=sol2[2])

    # -- step 3 ------------------------------------------------------------
    @staticmethod
    def solve_step3(cache, s2):
        """Solvers.jl:242-259."""
        R11, R21, R31 = s2["R_11"], s2["R_21"], s2["R_31"]
        R12, R22, R32 = s2["R_12"], s2["R_22"], s2["R_32"]
        k = cache
        return dict(
            phi_1=k["phi_11"] + R11 * k["phi_12"] + R21 * k["phi_13"] + R31 * k["phi_14"],
            mu_1=k["mu_11"] + R11 * k["mu_12"] + R21 * k["mu_13"] + R31 * k["mu_14"],
            psi_1=k["psi_11"] + R31 * k["psi_12"],
            nu_1=k["nu_11"] + R31 * k["nu_12"],
            phi_2=k["phi_21"] + R12 * k["phi_12"] + R22 * k["phi_13"] + R32 * k["phi_14"],
            mu_2=k["mu_21"] + R12 * k["mu_12"] + R22 * k["mu_13"] + R32 * k["mu_14"],
            psi_2=k["psi_21"] + R32 * k["psi_12"],
            nu_2=k["nu_21"] + R32 * k["nu_12"])

    # -- step 4 ------------------------------------------------------------
    def solve_step4(self, present, old, bdf):
        """Solvers.jl:267-337 (2D); Solvers_3d.jl:221-292 (3D).

        3D deviation: Solvers_3d.jl:266 assigns an (Nx,Ny,Nz,N) broadcast into an
        (Nx,Ny,Nz) slice, which Julia rejects for N >= 2.  The coupling force is
        summed over vesicles here, exactly as the 2D code does (Solvers.jl:313-318)
        and as 3D step 5 does (Solvers_3d.jl:369); for N == 1 both agree."""
        o, c = self.ops, self.conf
        dt, eta, lam = present.dt, c.eta, c.lamda
        a, b, cc = bdf
        nd = o.ndim
        phi_star = 2.0 * present.phi - old.phi
        psi_star = 2.0 * present.psi - old.psi
        u_star = 2.0 * present.u - old.u
        u_star_hat = 2.0 * present.u_hat - old.u_hat
        mu_star_hat = 2.0 * present.mu_hat - old.mu_hat
        nu_star_hat = 2.0 * present.nu_hat - old.nu_hat
        lhs = (a / dt) - eta * o.Laplacian
        rhs1 = np.stack([-(b * present.u_hat[d] + cc * old.u_hat[d]) / dt - o.D1[d] * present.p_hat
                         for d in range(nd)])
        conv = np.zeros((nd,) + o.shape)
        for comp in range(nd):
            for d in range(nd):
                conv[comp] += u_star[d] * o.irfft(o.D1[d] * u_star_hat[comp])
        coupl = np.zeros((nd,) + o.shape)
        for d in range(nd):
            dmu = o.irfft(o.D1[d] * mu_star_hat)
            dnu = o.irfft(o.D1[d] * nu_star_hat)
            coupl[d] = np.sum(phi_star * dmu + psi_star * dnu, axis=0)
        rhs2 = np.stack([o.rfft(-conv[d] - lam * coupl[d]) for d in range(nd)])
        ut1h = rhs1 / lhs
        ut2h = rhs2 / lhs
        return dict(u_tilde_1_hat=ut1h, u_tilde_2_hat=ut2h,
                    u_tilde_1=o.irfft(ut1h), u_tilde_2=o.irfft(ut2h), conv=conv, coupl=coupl)

    # -- step 5 ------------------------------------------------------------
    def solve_step5(self, present, old, s3, s4, bdf):
        """Solvers.jl:344-423; 3D uses conf.dt (Solvers_3d.jl:386-388)."""
        o, c = self.ops, self.conf
        a, b, cc = bdf
        lam = c.lamda
        nd = o.ndim
        dt = present.dt if nd == 2 else c.dt
        phi_star_hat = 2.0 * present.phi_hat - old.phi_hat
        psi_star_hat = 2.0 * present.psi_hat - old.psi_hat
        u_star = 2.0 * present.u - old.u
        gphi = [o.irfft(o.D1[d] * phi_star_hat) for d in range(nd)]
        gpsi = [o.irfft(o.D1[d] * psi_star_hat) for d in range(nd)]
        conv, coupl = s4["conv"], s4["coupl"]
        mu_1 = o.irfft(s3["mu_1"])
        nu_1 = o.irfft(s3["nu_1"])
        mu_2 = o.irfft(s3["mu_2"])
        nu_2 = o.irfft(s3["nu_2"])
        ut1, ut2 = s4["u_tilde_1"], s4["u_tilde_2"]
        adv_phi = sum(u_star[d] * gphi[d] for d in range(nd))
        adv_psi = sum(u_star[d] * gpsi[d] for d in range(nd))

        def dot(x, y):
            return float(np.dot(x.reshape(-1), y.reshape(-1)))
        t11 = dot(adv_phi, mu_1)
        t12 = dot(adv_psi, nu_1)
        t13 = sum(dot(coupl[d], ut1[d]) for d in range(nd))
        t14 = sum(dot(conv[d], ut1[d]) for d in range(nd))
        dv = o.dvol(c)
        theta_1 = dv * (lam * (t11 + t12 + t13) + t14)
        t21 = dot(adv_phi, mu_2)
        t22 = dot(adv_psi, nu_2)
        t23 = sum(dot(coupl[d], ut2[d]) for d in range(nd))
        t24 = sum(dot(conv[d], ut2[d]) for d in range(nd))
        theta_2 = dv * (lam * (t21 + t22 + t23) + t24)
        if nd == 3 and not theta_2 < 0:
            self.warnings.append(f"theta_2 = {theta_2} >= 0")
        rhs_Q = -(b * present.Q + cc * old.Q) / dt + theta_1
        lhs_Q = (a / dt) - theta_2
        self._theta = (theta_1, theta_2)
        return rhs_Q / lhs_Q

    # -- step 6 ------------------------------------------------------------
    @staticmethod
    def solve_step6(s2, s3, s4, Q):
        """Solvers.jl:431-453."""
        return dict(phi_hat=s3["phi_1"] + Q * s3["phi_2"], psi_hat=s3["psi_1"] + Q * s3["psi_2"],
                    mu_hat=s3["mu_1"] + Q * s3["mu_2"], nu_hat=s3["nu_1"] + Q * s3["nu_2"],
                    R1=s2["R_11"] + Q * s2["R_12"], R2=s2["R_21"] + Q * s2["R_22"],
                    R3=s2["R_31"] + Q * s2["R_32"],
                    u_tilde_hat=s4["u_tilde_1_hat"] + Q * s4["u_tilde_2_hat"])

    # -- step 7 ------------------------------------------------------------
    def solve_step7(self, present, s6, bdf):
        """Solvers.jl:459-498 (3D: Solvers_3d.jl:428-468)."""
        o, c = self.ops, self.conf
        a, dt, eta = bdf[0], present.dt, c.eta
        ut = s6["u_tilde_hat"]
        div = sum(o.D1[d] * ut[d] for d in range(o.ndim))
        with np.errstate(divide="ignore", invalid="ignore"):
            dp = (a / dt) * div / o.Laplacian
        dp[(0,) * o.ndim] = 0.0
        u_hat = np.stack([ut[d] - (dt / a) * o.D1[d] * dp for d in range(o.ndim)])
        p_hat = present.p_hat + dp - eta * div
        return dict(u_hat=u_hat, p_hat=p_hat, delta_p_hat=dp)

    # -- modified energy (Utils.jl:96-141) ----------------------------------
    def modified_energy(self, present, old):
        o, c = self.ops, self.conf
        kinetic = 0.25 * integrate_sq(present.u, c, o) + 0.25 * integrate_sq(2 * present.u - old.u, c, o)
        gpx = o.irfft(o.D1[0] * present.p_hat)
        gpy = o.irfft(o.D1[1] * present.p_hat)
        pressure = (present.dt ** 2 / 3.0) * (integrate_sq(gpx, c, o) + integrate_sq(gpy, c, o))

        def se(rn, ro):
            return 0.5 * rn ** 2 + 0.5 * (2 * rn - ro) ** 2
        R_energy = se(present.R1, old.R1) + se(present.R2, old.R2) + se(present.R3, old.R3)
        Q_energy = 0.5 * se(present.Q, old.Q)
        phi_star = 2 * present.phi - old.phi
        psi_star = 2 * present.psi - old.psi
        phi_star_hat = 2 * present.phi_hat - old.phi_hat
        lap_new = o.irfft(o.Laplacian * present.phi_hat)
        lap_star = o.irfft(o.Laplacian * phi_star_hat)
        gx_new = o.irfft(o.D1[0] * present.phi_hat)
        gy_new = o.irfft(o.D1[1] * present.phi_hat)
        gx_star = o.irfft(o.D1[0] * phi_star_hat)
        gy_star = o.irfft(o.D1[1] * phi_star_hat)
        S1_term = 0.25 * c.S1 * (integrate_sq(lap_new, c, o) + integrate_sq(lap_star, c, o))
        S2_term = 0.25 * c.S2 * (integrate_sq(gx_new, c, o) + integrate_sq(gy_new, c, o)
                                 + integrate_sq(gx_star, c, o) + integrate_sq(gy_star, c, o))
        S3_term = 0.25 * c.S3 * (integrate_sq(present.phi, c, o) + integrate_sq(phi_star, c, o))
        S4_term = 0.25 * c.S4 * (integrate_sq(present.psi, c, o) + integrate_sq(psi_star, c, o))
        lam = c.lamda
        return (kinetic + pressure + lam * R_energy + Q_energy
                + lam * (S1_term + S2_term + S3_term + S4_term)
                - lam * (c.C1 + c.C2 + c.C3) - 0.5)

    # -- SAV initialisation (simulation.jl:24-53) ----------------------------
    def init_sav(self, present: State):
        o, c = self.ops, self.conf
        A_0 = calculate_area(present.phi, o, c)
        present.A0 = A_0
        present.R1 = math.sqrt(get_W1(present.phi, o, c, A_0) + c.C1)
        present.R2 = math.sqrt(get_W2(present.phi, o, c) + c.C2)
        present.R3 = math.sqrt(get_W3(present.phi, present.psi, c, o) + c.C3)
        present.Q = 1.0
        L_phi = self._L_phi()
        H1 = get_H1(present.phi, o, c, A_0)
        H2 = get_H2(present.phi, o, c)
        H3 = get_H3(present.phi, present.psi, c, o)
        G = get_MG(present.phi, present.psi, c, o)
        present.mu_hat[...] = (L_phi * present.phi_hat + present.R1 * o.rfft(H1)
                               + present.R2 * o.rfft(H2) + present.R3 * o.rfft(H3))
        present.nu_hat[...] = c.S4 * present.psi_hat + present.R3 * o.rfft(G)
        present.mu[...] = o.irfft(present.mu_hat)
        present.nu[...] = o.irfft(present.nu_hat)
        old = present.copy()
        return present, old

    # -- one loop body (simulation.jl:73-110, 146-148) -----------------------
    def step(self, present: State, old: State, dt: float, n: int, monitor: bool = True):
        o = self.ops
        present.dt = dt
        bdf = bdf_coeff(n)
        self.krylov_log = []
        cache = self.solve_step1(present, old, bdf)
        s2 = self.solve_step2(present, old, cache, bdf, fields=self._fields)
        s3 = self.solve_step3(cache, s2)
        s4 = self.solve_step4(present, old, bdf)
        Q = self.solve_step5(present, old, s3, s4, bdf)
        s6 = self.solve_step6(s2, s3, s4, Q)
        s7 = self.solve_step7(present, s6, bdf)
        update_state(old, present)
        present.phi_hat[...] = s6["phi_hat"]
        present.psi_hat[...] = s6["psi_hat"]
        present.mu_hat[...] = s6["mu_hat"]
        present.nu_hat[...] = s6["nu_hat"]
        present.u_hat[...] = s7["u_hat"]
        present.p_hat[...] = s7["p_hat"]
        present.phi[...] = o.irfft(present.phi_hat)
        present.psi[...] = o.irfft(present.psi_hat)
        present.mu[...] = o.irfft(present.mu_hat)
        present.nu[...] = o.irfft(present.nu_hat)
        present.p[...] = o.irfft(present.p_hat)
        present.u[...] = o.irfft(present.u_hat)
        present.R1, present.R2, present.R3 = s6["R1"], s6["R2"], s6["R3"]
        present.Q = Q
        self.last = dict(cache=cache, s2=s2, s3=s3, s4=s4, Q=Q, s6=s6, s7=s7)
        out = dict(krylov=list(self.krylov_log))
        if monitor:
            out["energy"] = self.modified_energy(present, old)
            A = calculate_area(present.phi, o, self.conf)
            out["area_ratio"] = float(np.sum(np.abs(A - present.A0)) / np.sum(present.A0))
        return out
