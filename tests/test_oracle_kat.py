"""Pins the CPU oracle before it is trusted as the checker.

* Known-answer tests of the reference's own suite, test/test_Types_Ops.jl
  (fixture: 128x64 on [0, 2pi]^2, N=1): G2 shapes/signs, G3 wavenumbers incl.
  Nyquist signs, G4 round trips, G5 spectral derivatives of sin(3x)cos(4y),
  G6 pseudo-spectral products, G8 BDF coefficients.
* Closed forms and brute force for what those leave open: the c2r definition
  on non-Hermitian half spectra, the 1-D interface energy of the tanh profile,
  Gateaux derivatives of W2/W3, a dense solve of the psi operator, invariants
  of the full step (spectral incompressibility, energy decay, mean of psi).
"""
import math

import numpy as np
import pytest


class Fixture:
    """The reference test fixture (test_Types_Ops.jl:28-41)."""
    N = 1
    Nx, Ny = 128, 64
    Lx = Ly = 2 * math.pi
    dx, dy = Lx / Nx, Ly / Ny
    ndim = 2


@pytest.fixture(scope="module")
def ops(so):
    return so.Ops(Fixture)


# --- G2 --------------------------------------------------------------------
def test_g2_operator_shapes_and_signs(ops):
    Nkx = Fixture.Nx // 2 + 1
    assert ops.kx.shape == (1, Nkx) and ops.ky.shape == (Fixture.Ny, 1)
    assert np.all(ops.D1[0].real == 0) and np.all(ops.D1[1].real == 0)
    assert np.allclose(ops.D1[0].imag, ops.kx) and np.allclose(ops.D1[1].imag, ops.ky)
    assert ops.Laplacian.shape == (Fixture.Ny, Nkx) and ops.Biharmonic.shape == (Fixture.Ny, Nkx)
    assert np.all(ops.Laplacian <= 0) and np.all(ops.Biharmonic >= 0)
    assert np.allclose(ops.Biharmonic, ops.Laplacian ** 2)


# --- G3 --------------------------------------------------------------------
def test_g3_wavenumber_values(ops):
    Nx, Ny, Lx, Ly = Fixture.Nx, Fixture.Ny, Fixture.Lx, Fixture.Ly
    Nkx = Nx // 2 + 1
    kx, ky = ops.kx[0], ops.ky[:, 0]
    assert kx[0] == pytest.approx(0.0, abs=1e-15)
    assert kx[1] == pytest.approx(2 * math.pi / Lx)
    assert kx[Nkx - 1] == pytest.approx(math.pi * Nx / Lx)          # Nyquist x is +
    assert ky[0] == pytest.approx(0.0, abs=1e-15)
    assert ky[1] == pytest.approx(2 * math.pi / Ly)
    assert ky[Ny // 2] == pytest.approx(-math.pi * Ny / Ly)         # Nyquist y is -
    assert ops.kx0[0, Nx // 2] == pytest.approx(-math.pi * Nx / Lx)  # full-spectrum Nyquist x is -


# --- G4 --------------------------------------------------------------------
def test_g4_round_trips(ops):
    rng = np.random.default_rng(4)
    u = rng.standard_normal((1, Fixture.Ny, Fixture.Nx))
    assert np.max(np.abs(ops.irfft(ops.rfft(u)) - u)) < 1e-13
    c = np.full((1, Fixture.Ny, Fixture.Nx), 3.7)
    assert np.max(np.abs(ops.irfft(ops.rfft(c)) - c)) < 1e-15


# --- G5 --------------------------------------------------------------------
def test_g5_spectral_derivatives(ops):
    Nx, Ny, Lx, Ly = Fixture.Nx, Fixture.Ny, Fixture.Lx, Fixture.Ly
    km, kn = 3 * 2 * math.pi / Lx, 4 * 2 * math.pi / Ly
    x = Lx * np.arange(Nx) / Nx
    y = Ly * np.arange(Ny) / Ny
    Y, X = np.meshgrid(y, x, indexing="ij")
    scale = Nx * Ny / 4
    f = (np.sin(km * X) * np.cos(kn * Y) / scale)[None]
    fh = ops.rfft(f)
    assert np.sort(np.abs(fh).ravel())[::-1][0] == pytest.approx(1.0, abs=1e-10)
    dfx = (km * np.cos(km * X) * np.cos(kn * Y) / scale)[None]
    dfy = (-kn * np.sin(km * X) * np.sin(kn * Y) / scale)[None]
    lap = (-(km ** 2 + kn ** 2) * np.sin(km * X) * np.cos(kn * Y) / scale)[None]
    bih = ((km ** 2 + kn ** 2) ** 2 * np.sin(km * X) * np.cos(kn * Y) / scale)[None]
    assert np.max(np.abs(ops.irfft(ops.D1[0] * fh) - dfx)) < 1e-15
    assert np.max(np.abs(ops.irfft(ops.D1[1] * fh) - dfy)) < 1e-15
    assert np.max(np.abs(ops.irfft(ops.Laplacian * fh) - lap)) < 1e-14
    assert np.max(np.abs(ops.irfft(ops.Biharmonic * fh) - bih)) < 1e-11


# --- G6 --------------------------------------------------------------------
def _mult(ops, uh, vh):
    """mult! (SpectralUtils.jl:158-168)."""
    return ops.rfft(ops.irfft(uh) * ops.irfft(vh))


def test_g6_products(ops):
    Nx, Ny = Fixture.Nx, Fixture.Ny
    x = 2 * math.pi * np.arange(Nx) / Nx
    y = 2 * math.pi * np.arange(Ny) / Ny
    Y, X = np.meshgrid(y, x, indexing="ij")
    f, gg = np.sin(X)[None], np.cos(Y)[None]
    fh, gh = ops.rfft(f), ops.rfft(gg)
    # reference bound 1e-15 (test_Types_Ops.jl:277) is FFTW's roundoff; pocketfft's sits within 2 ulp of it
    assert np.max(np.abs(ops.irfft(_mult(ops, fh, gh)) - (np.sin(X) * np.cos(Y))[None])) < 4e-15
    h = (np.sin(3 * X) + np.sin(4 * X))[None]
    hh = ops.rfft(h)
    assert np.max(np.abs(ops.irfft(_mult(ops, hh, hh)) - h ** 2)) < 1e-14
    assert np.max(np.abs(_mult(ops, fh, gh) - _mult(ops, gh, fh))) <= 1e-15


# --- G8 --------------------------------------------------------------------
def test_g8_bdf(so):
    assert so.bdf_coeff(1) == (1.0, -1.0, 0.0)
    assert so.bdf_coeff(2) == (1.5, -2.0, 0.5)
    a, b, c = so.bdf_coeff(2)
    dt = 0.01
    approx = (a * 1.0 + b * (1 - dt) ** 2 + c * (1 - 2 * dt) ** 2) / dt
    assert abs(approx - 2.0) < 1e-2


# --- c2r on non-Hermitian input: brute force of the definition --------------
def _c2r_bruteforce(X, Nx, Ny):
    """irfft(X)[x, y] = 1/(Nx Ny) sum_kx w_kx Re( e^{2 pi i kx x/Nx} sum_ky X[kx, ky] e^{2 pi i ky y/Ny} ),
    w = 1 at kx = 0 and kx = Nx/2, else 2 (FFTW c2r: c2c over y, then c2r over x)."""
    out = np.zeros((Ny, Nx))
    Nkx = Nx // 2 + 1
    for yy in range(Ny):
        inner = np.array([np.sum(X[:, kx] * np.exp(2j * np.pi * np.arange(Ny) * yy / Ny)) for kx in range(Nkx)])
        for xx in range(Nx):
            s = 0.0
            for kx in range(Nkx):
                w = 1.0 if kx in (0, Nx // 2) else 2.0
                s += w * (inner[kx] * np.exp(2j * np.pi * kx * xx / Nx)).real
            out[yy, xx] = s / (Nx * Ny)
    return out


def test_c2r_definition_on_nonhermitian(so):
    class F:
        N, Nx, Ny, Lx, Ly, ndim = 1, 8, 6, 2 * math.pi, 2 * math.pi, 2
    o = so.Ops(F)
    rng = np.random.default_rng(8)
    X = rng.standard_normal((6, 5)) + 1j * rng.standard_normal((6, 5))
    got = o.irfft(X[None])[0]
    assert np.max(np.abs(got - _c2r_bruteforce(X, 8, 6))) < 1e-14
    # the derivative at the Nyquist planes makes non-Hermitian data: d/dx drops Nyquist-x
    Xh = o.rfft(rng.standard_normal((1, 6, 8)))
    assert np.max(np.abs(o.irfft(o.D1[0] * Xh)[0] - _c2r_bruteforce((o.D1[0] * Xh)[0], 8, 6))) < 1e-13


def test_rfft_bruteforce_dft(so):
    class F:
        N, Nx, Ny, Lx, Ly, ndim = 1, 6, 4, 1.0, 1.0, 2
    o = so.Ops(F)
    rng = np.random.default_rng(1)
    u = rng.standard_normal((4, 6))
    X = np.zeros((4, 4), complex)
    for ky in range(4):
        for kx in range(4):
            X[ky, kx] = sum(u[y, x] * np.exp(-2j * np.pi * (kx * x / 6 + ky * y / 4)) for y in range(4) for x in range(6))
    assert np.max(np.abs(o.rfft(u[None])[0] - X)) < 1e-13


# --- functionals: closed forms ---------------------------------------------
def _conf(jv, n=256, eps=0.08, **kw):
    c = jv.workload_c1(n=n, steps=1).conf
    S = jv.stabilisers(eps, c.gamma_bend, c.gamma_in, c.gamma_out)
    c = c.with_(epsilon=eps, S1=S[0], S3=S[2])
    return c.with_(**kw) if kw else c


def test_interface_energy_of_tanh_stripe(jv, so):
    """Two straight interfaces of the equilibrium profile: per unit length
    int f_surf = (3 sqrt2/4) int g/eps + int eps phi'^2/2 = 1/2 + sqrt2/3."""
    conf = _conf(jv, n=256, eps=0.08)
    o = so.Ops(conf)
    x = np.arange(conf.Nx) * conf.dx
    s2e = math.sqrt(2) * conf.epsilon
    x1, x2 = math.pi / 2, 3 * math.pi / 2
    prof = np.tanh((x - x1) / s2e) - np.tanh((x - x2) / s2e) - 1.0
    phi = np.broadcast_to(prof[None, :], (conf.Ny, conf.Nx))[None].copy()
    area = so.calculate_area(phi, o, conf)[0]
    assert area == pytest.approx(2 * conf.Ly * (0.5 + math.sqrt(2) / 3), rel=1e-6)


def test_constant_phase_functionals(jv, so):
    conf = _conf(jv, n=32)
    o = so.Ops(conf)
    c0 = 0.3
    phi = np.full((1, conf.Ny, conf.Nx), c0)
    psi = np.full_like(phi, 1.1)
    vol = conf.Lx * conf.Ly
    g = 0.25 * (c0 ** 2 - 1) ** 2
    assert so.calculate_area(phi, o, conf)[0] == pytest.approx((3 * math.sqrt(2) / 4) * g / conf.epsilon * vol, rel=1e-12)
    fb = 3 * math.sqrt(2) / (16 * conf.epsilon) * (c0 * (c0 ** 2 - 1) / conf.epsilon) ** 2
    assert so.integrate(so.f_bend(phi, o, conf), conf, o) == pytest.approx(fb * vol, rel=1e-12)
    A0 = so.calculate_area(phi, o, conf)
    assert np.max(np.abs(so.get_H1(phi, o, conf, A0))) == 0.0


def _gateaux(fun, x, eta, h=1e-4):
    return (fun(x + h * eta) - fun(x - h * eta)) / (2 * h)


def test_gateaux_H3_G_are_derivatives_of_W3(jv, so):
    conf = _conf(jv, n=64, eps=0.2)
    o = so.Ops(conf)
    wl = jv.workload_c1(n=64, steps=1)
    phi, psi = jv.initial_fields(conf, wl.vesicles, wl.c_out)
    rng = np.random.default_rng(2)
    eta = o.irfft(o.rfft(rng.standard_normal(phi.shape)) * (np.abs(o.Laplacian) < 20))
    dv = conf.dx * conf.dy
    R3 = math.sqrt(so.get_W3(phi, psi, conf, o) + conf.C3)
    d_phi = _gateaux(lambda p: so.get_W3(p, psi, conf, o), phi, eta)
    assert d_phi == pytest.approx(np.sum(so.get_H3(phi, psi, conf, o) * R3 * eta) * dv, rel=1e-6)
    d_psi = _gateaux(lambda q: so.get_W3(phi, q, conf, o), psi, eta)
    assert d_psi == pytest.approx(np.sum(so.get_MG(phi, psi, conf, o) * R3 * eta) * dv, rel=1e-6)


def test_gateaux_H2_bending_and_linear_parts(jv, so):
    """With gamma_surf = 0, H2 * R2 is the exact variational derivative of W2."""
    conf = _conf(jv, n=64, eps=0.2).with_(gamma_surf=0.0, S2=0.7)
    o = so.Ops(conf)
    wl = jv.workload_c1(n=64, steps=1)
    phi, _ = jv.initial_fields(conf, wl.vesicles, wl.c_out)
    rng = np.random.default_rng(3)
    eta = o.irfft(o.rfft(rng.standard_normal(phi.shape)) * (np.abs(o.Laplacian) < 20))
    R2 = math.sqrt(so.get_W2(phi, o, conf) + conf.C2)
    d = _gateaux(lambda p: so.get_W2(p, o, conf), phi, eta, h=1e-5)
    ref = np.sum(so.get_H2(phi, o, conf) * R2 * eta) * conf.dx * conf.dy
    assert d == pytest.approx(ref, rel=1e-6)


def test_wc_negative_raises(jv, so):
    conf = _conf(jv, n=32).with_(C3=1.0)
    o = so.Ops(conf)
    phi = np.full((1, conf.Ny, conf.Nx), 0.5)
    with pytest.raises(so.SavFatal):
        so.get_W3(phi, np.full_like(phi, 1.0), conf, o)


# --- Krylov restatement ------------------------------------------------------
def _psi_setup(jv, so, n=8):
    conf = _conf(jv, n=n, eps=0.5)
    orc = so.Oracle(conf)
    wl = jv.workload_c1(n=n, steps=1)
    phi, _ = jv.initial_fields(conf, wl.vesicles, wl.c_out)
    return conf, orc, phi


def test_bicgstab_matches_dense_solve(jv, so):
    conf, orc, phi = _psi_setup(jv, so, n=8)
    o = orc.ops
    M = so.get_M_psi(phi, conf)
    a, dt = 1.5, conf.dt
    rng = np.random.default_rng(11)
    rhs = o.rfft(rng.standard_normal(phi.shape))
    x = orc.solve_psi(rhs, M, float(M.max()), a, dt)
    # assemble the operator of Solvers.jl:39-53 column by column
    n = phi.size
    shape = phi.shape
    A = np.zeros((n, n), complex)
    for j in range(n):
        e = np.zeros(n, complex)
        e[j] = 1.0
        xh = e.reshape(shape)
        acc = 0
        for d in range(2):
            acc = acc + o.D1_full[d] * o.fft(M * o.ifft(o.D1_full[d] * (conf.S4 * xh)))
        A[:, j] = ((a / dt) * xh - acc).reshape(-1)
    b = o.fft(o.irfft(rhs)).reshape(-1)
    xd = np.linalg.solve(A, b).reshape(shape)[..., : conf.Nx // 2 + 1]
    assert np.max(np.abs(x - xd)) < 1e-8 * np.max(np.abs(xd))
    ks = orc.krylov_log[-1]
    assert ks.solved and ks.residuals[-1] <= 1e-14 + conf.tol * ks.residuals[0]


def test_bicgstab_constant_mobility_is_the_spectral_solve(jv, so):
    """M0_psi = 0 makes M_psi = 0.1 everywhere: the operator is diagonal and equal
    to the preconditioner, so the solve is the 3D-style spectral divide."""
    conf, orc, phi = _psi_setup(jv, so, n=16)
    conf = conf.with_(M0_psi=0.0)
    orc = so.Oracle(conf)
    o = orc.ops
    M = so.get_M_psi(phi, conf)
    rng = np.random.default_rng(12)
    rhs = o.rfft(rng.standard_normal(phi.shape))
    x = orc.solve_psi(rhs, M, float(M.max()), 1.0, conf.dt)
    ref = o.rfft(o.irfft(rhs)) / ((1.0 / conf.dt) - 0.1 * conf.S4 * o.Laplacian)
    assert orc.krylov_log[-1].niter <= 1
    assert np.max(np.abs(x - ref)) < 1e-9 * np.max(np.abs(ref))


def test_bicgstab_zero_rhs_returns_zero(jv, so):
    conf, orc, phi = _psi_setup(jv, so, n=8)
    M = so.get_M_psi(phi, conf)
    x = orc.solve_psi(np.zeros((1, conf.Ny, conf.Nx // 2 + 1), complex), M, float(M.max()), 1.0, conf.dt)
    ks = orc.krylov_log[-1]
    assert ks.niter == 0 and ks.solved and np.all(x == 0)


def test_bicgstab_itmax_reported(jv, so):
    conf, orc, phi = _psi_setup(jv, so, n=16)
    orc.itmax = 1
    M = so.get_M_psi(phi, conf)
    rng = np.random.default_rng(13)
    orc.solve_psi(orc.ops.rfft(rng.standard_normal(phi.shape)), M, float(M.max()), 1.0, conf.dt)
    ks = orc.krylov_log[-1]
    assert ks.niter == 1 and not ks.solved and "maximum" in ks.status
    assert orc.warnings


# --- full-step invariants ----------------------------------------------------
@pytest.fixture(scope="module")
def short_run(jv, so):
    wl = jv.workload_c1(n=64, steps=12)
    conf = wl.conf
    phi, psi = wl.initial_fields()
    orc = so.Oracle(conf)
    pres, old = orc.init_sav(so.state_from_physical(phi, psi, orc.ops, 2))
    m0 = pres.psi.mean()
    E, divs, means = [], [], []
    for n in range(1, 13):
        out = orc.step(pres, old, conf.dt, n)
        E.append(out["energy"])
        o = orc.ops
        div = o.D1[0] * pres.u_hat[0] + o.D1[1] * pres.u_hat[1]
        div[0, 0] = 0
        divs.append(np.max(np.abs(div)) / max(np.max(np.abs(pres.u_hat)), 1e-300))
        means.append(pres.psi.mean())
    return dict(E=E, divs=divs, means=means, m0=m0, u=np.max(np.abs(pres.u)))


def test_step_velocity_is_spectrally_divergence_free(short_run):
    assert short_run["u"] > 0
    assert max(short_run["divs"]) < 1e-12


def test_step_modified_energy_decays(short_run):
    E = short_run["E"]
    assert all(E[i + 1] <= E[i] for i in range(len(E) - 1)), E


def test_step_mean_psi_conserved(short_run):
    assert max(abs(m - short_run["m0"]) for m in short_run["means"]) < 1e-9 * abs(short_run["m0"])


def test_first_step_is_bdf1_with_old_equal_present(jv, so):
    """At n = 1 old = present, so the extrapolations are the identity (simulation.jl:53)."""
    wl = jv.workload_c1(n=32, steps=1)
    phi, psi = wl.initial_fields()
    orc = so.Oracle(wl.conf)
    pres, old = orc.init_sav(so.state_from_physical(phi, psi, orc.ops, 2))
    assert np.array_equal(2.0 * pres.phi - old.phi, pres.phi)


# --- learned-preconditioner network reference --------------------------------
def test_conv_oracle_gradients_by_finite_differences():
    from oracle import conv_oracle as co
    rng = np.random.default_rng(21)
    params = co.init_params(3)
    params = [(W, rng.standard_normal(b.shape) * 0.1) for W, b in params]
    x = rng.standard_normal((2, 2, 8, 6))
    t = rng.standard_normal((2, 1, 8, 6))
    loss, grads = co.mse_loss(params, x, t)
    for l in range(4):
        W, b = params[l]
        for idx in [(0, 0, 0, 0), (W.shape[0] - 1, W.shape[1] - 1, 2, 1), (0, W.shape[1] - 1, 1, 2)]:
            h = 1e-6
            Wp = W.copy(); Wp[idx] += h
            Wm = W.copy(); Wm[idx] -= h
            pp = list(params); pp[l] = (Wp, b)
            pm = list(params); pm[l] = (Wm, b)
            fd = (co.mse_loss(pp, x, t)[0] - co.mse_loss(pm, x, t)[0]) / (2 * h)
            assert grads[l][0][idx] == pytest.approx(fd, rel=1e-5, abs=1e-9)
        h = 1e-6
        bp = b.copy(); bp[0] += h
        bm = b.copy(); bm[0] -= h
        pp = list(params); pp[l] = (W, bp)
        pm = list(params); pm[l] = (W, bm)
        fd = (co.mse_loss(pp, x, t)[0] - co.mse_loss(pm, x, t)[0]) / (2 * h)
        assert grads[l][1][0] == pytest.approx(fd, rel=1e-5, abs=1e-9)


def test_conv_oracle_conv_is_a_periodic_correlation():
    from oracle import conv_oracle as co
    rng = np.random.default_rng(22)
    a = rng.standard_normal((1, 1, 5, 4))
    W = rng.standard_normal((1, 1, 3, 3))
    out = co.conv3(W, a)
    i, j = 0, 3
    ref = sum(W[0, 0, di, dj] * a[0, 0, (i + di - 1) % 5, (j + dj - 1) % 4] for di in range(3) for dj in range(3))
    assert out[0, 0, i, j] == pytest.approx(ref, rel=1e-14)
    assert co.gelu(np.array([0.0]))[0] == 0.0
    assert co.gelu(np.array([30.0]))[0] == pytest.approx(30.0)


# --- stripe closed forms for H1 and the whole of H2 -------------------------
def _stripe(jv, so, delta, **kw):
    """phi = tanh((x - x1)/(sqrt2 delta)) - tanh((x - x2)/(sqrt2 delta)) - 1: two straight
    interfaces pi apart, each within exp(-pi/(sqrt2 delta)) of the single-interface profile
    (delta small enough that the slope mismatch of the tails at x = 0 = 2 pi is ~1e-16)."""
    conf = _conf(jv, n=512, eps=0.08, **kw)
    o = so.Ops(conf)
    x = np.arange(conf.Nx) * conf.dx
    s2d = math.sqrt(2) * delta
    x1, x2 = math.pi / 2, 3 * math.pi / 2
    prof = np.tanh((x - x1) / s2d) - np.tanh((x - x2) / s2d) - 1.0
    phi = np.broadcast_to(prof[None, :], (conf.Ny, conf.Nx))[None].copy()
    return conf, o, phi


def test_H1_closed_form_on_a_stripe_with_A0_not_A(jv, so):
    """On phi = tanh(x/(sqrt2 delta)), delta != eps: phi'' = (phi^3 - phi)/delta^2, so
    omega = (phi^3 - phi)(1/eps - eps/delta^2), and the surface functional per interface
    and unit length is delta/(2 eps) + sqrt2 eps/(3 delta).  H1 (Utils.jl:198-206) is then
    gamma_area (A - A0)(3 sqrt2/4) omega / sqrt(W1 + C1), W1 = gamma_area (A - A0)^2 / 2."""
    delta = 0.06
    conf, o, phi = _stripe(jv, so, delta, gamma_area=1000, C1=3.0)
    eps = conf.epsilon
    A = 2 * conf.Ly * (delta / (2 * eps) + math.sqrt(2) * eps / (3 * delta))
    assert so.calculate_area(phi, o, conf)[0] == pytest.approx(A, rel=1e-9)
    A0 = np.array([0.97 * A])
    W1 = 0.5 * conf.gamma_area * (A - A0[0]) ** 2
    assert so.get_W1(phi, o, conf, A0) == pytest.approx(W1, rel=1e-8)
    omega = (phi ** 3 - phi) * (1 / eps - eps / delta ** 2)
    np.testing.assert_allclose(so.get_omega(phi, o, conf), omega, rtol=0, atol=1e-9 * np.max(np.abs(omega)))
    H1 = conf.gamma_area * (A - A0[0]) * (3 * math.sqrt(2) / 4) * omega / math.sqrt(W1 + conf.C1)
    got = so.get_H1(phi, o, conf, A0)
    assert np.max(np.abs(got - H1)) <= 1e-8 * np.max(np.abs(H1))
    # the sign follows A - A0: an area above its target gives H1 of omega's sign
    assert np.sum(got * omega) > 0


def test_H2_surface_and_bending_closed_form_on_a_stripe(jv, so):
    """H2 (Utils.jl:208-226) with gamma_s, gamma_b != 0 and S1..S3 = 0 on the stripe:
    lap omega = c phi (1 - phi^2)(4 - 6 phi^2)/delta^2 with c = 1/eps - eps/delta^2, and
    int omega^2 dx per interface = c^2 sqrt2 delta 4/15, which fixes W2 too."""
    delta = 0.06
    conf, o, phi = _stripe(jv, so, delta, gamma_surf=1.3, gamma_bend=0.2, S1=0.0, S2=0.0, S3=0.0, C2=50.0)
    eps = conf.epsilon
    c = 1 / eps - eps / delta ** 2
    A = 2 * conf.Ly * (delta / (2 * eps) + math.sqrt(2) * eps / (3 * delta))
    Fb = (3 * math.sqrt(2) / (16 * eps)) * 2 * conf.Ly * c ** 2 * math.sqrt(2) * delta * 4 / 15
    W2 = conf.gamma_surf * A + conf.gamma_bend * Fb
    assert so.get_W2(phi, o, conf) == pytest.approx(W2, rel=1e-8)
    omega = c * (phi ** 3 - phi)
    lap_omega = c * phi * (1 - phi ** 2) * (4 - 6 * phi ** 2) / delta ** 2
    H2 = (conf.gamma_surf * (3 * math.sqrt(2) / 4) * omega
          + conf.gamma_bend * (3 * math.sqrt(2) / (8 * eps)) * (omega * (3 * phi ** 2 - 1) / eps - eps * lap_omega)
          ) / math.sqrt(W2 + conf.C2)
    got = so.get_H2(phi, o, conf)
    assert np.max(np.abs(got - H2)) <= 1e-8 * np.max(np.abs(H2))
    # the surface term alone (gamma_b = 0): a wrong 3 sqrt2/4 or sign would fail here
    conf_s = conf.with_(gamma_bend=0.0)
    H2s = conf.gamma_surf * (3 * math.sqrt(2) / 4) * omega / math.sqrt(conf.gamma_surf * A + conf.C2)
    assert np.max(np.abs(so.get_H2(phi, o, conf_s) - H2s)) <= 1e-8 * np.max(np.abs(H2s))


# --- BiCGSTAB identities at every iteration ---------------------------------
def test_bicgstab_identities_every_iteration(jv, so):
    """On the 8x8 psi operator, after every iteration k (also at the itmax stop):
    r_k = M (b - A x_k) (the recurrence residual is the preconditioned true residual),
    <c, s_k> = 0 (alpha = rho / <c, v>), <t_k, r_k> = 0 (omega = <t, s>/<t, t>, first
    argument conjugated), rho_{k+1} = <c, r_k>."""
    conf, orc, phi = _psi_setup(jv, so, n=8)
    o = orc.ops
    Mpsi = so.get_M_psi(phi, conf)
    a, dt = 1.5, conf.dt
    shape = phi.shape
    D1f = o.D1_full

    def A(xv):
        xh = xv.reshape(shape)
        acc = sum(D1f[d] * o.fft(Mpsi * o.ifft(D1f[d] * (conf.S4 * xh))) for d in range(2))
        return ((a / dt) * xh - acc).reshape(-1)
    Pd = (1.0 / (a / dt - float(Mpsi.max()) * o.Laplacian_full * conf.S4)).reshape(-1)
    M = lambda v: Pd * v
    rng = np.random.default_rng(17)
    b = o.fft(rng.standard_normal(shape)).reshape(-1)
    seen = []

    def cb(k, d):
        x, r, s, t, c = d["x"], d["r"], d["s"], d["t"], d["c"]
        scale = np.linalg.norm(M(b))
        assert np.linalg.norm(r - M(b - A(x))) <= 1e-12 * scale, k
        assert abs(np.vdot(c, s)) <= 1e-12 * np.linalg.norm(c) * np.linalg.norm(s), k
        assert abs(np.vdot(t, r)) <= 1e-12 * np.linalg.norm(t) * max(np.linalg.norm(r), 1e-300), k
        assert abs(np.vdot(c, r) - d["next_rho"]) <= 1e-12 * abs(d["next_rho"]) + 1e-300, k
        seen.append(k)

    x, st = so.bicgstab(A, b, M, 1e-14, 1e-10, 20, callback=cb)
    assert st.solved and seen == list(range(1, st.niter + 1)) and st.niter >= 3
    # stopped at itmax = 2: the same identities hold for the committed iterate
    seen.clear()
    x2, st2 = so.bicgstab(A, b, M, 1e-14, 1e-10, 2, callback=cb)
    assert st2.niter == 2 and not st2.solved and seen == [1, 2]
    # and that iterate is the first two iterates of the converging run
    x_a, _ = so.bicgstab(A, b, M, 1e-14, 1e-10, 2)
    assert np.array_equal(x_a, x2)


# --- the split against the coupled scheme it decomposes -----------------------
@pytest.mark.parametrize("case", ["2d", "3d"])
def test_step_solves_the_coupled_scheme(jv, so, case):
    """Three oracle steps (BDF1, then BDF2 with a non-zero velocity): the state that the
    seven-step split produces satisfies every equation of the coupled scheme it decomposes
    (helpers.scheme_residuals) to round-off.  In 2D the psi mobility is made constant
    (M0_psi = 0), where the full-complex psi solve keeps Hermitian symmetry and its half
    spectrum is the whole solution (to the solve's rtol); with a variable mobility the
    solution carries an anti-Hermitian part that the half spectrum drops, as the
    reference's does.  Planting a swapped R index in step 3, a dropped G term in the R
    system, a wrong sign of phi_12, or a dropped coupling force in u~_2 or in theta_1 leaves
    residuals of 1e-6 to 1 here."""
    from helpers import scheme_residuals
    if case == "2d":
        wl = jv.workload_c1(n=32, steps=3)
        conf = wl.conf.with_(M0_psi=0.0)
    else:
        wl = jv.workload_c4(n=16, steps=3)
        conf = wl.conf
    phi, psi = wl.initial_fields()
    orc = so.Oracle(conf)
    pres, old = orc.init_sav(so.state_from_physical(phi, psi, orc.ops, conf.ndim))
    for n in (1, 2, 3):
        p0, o0 = pres.copy(), old.copy()
        p0.dt = conf.dt
        orc.step(pres, old, conf.dt, n)
        res = scheme_residuals(so, conf, orc.ops, p0, o0, pres, n)
        # round-off of the transforms against terms up to S1 |k|^4 phi; the 2D psi solve
        # stops at rtol 1e-10
        tol = {"psi": 1e-8 if case == "2d" else 1e-10}
        for k, v in res.items():
            assert v < tol.get(k, 1e-10), (n, k, v)
    assert np.max(np.abs(pres.u)) > 0 and pres.Q != 1.0
