"""Shared helpers for the parity tests: build identical oracle / device runs."""
import numpy as np


def relerr(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    d = np.max(np.abs(a - b))
    s = np.max(np.abs(b))
    return d / s if s > 0 else d


def device_state(jv, ops, conf, phi, psi):
    """FieldState from memory-order (N, [Nz,] Ny, Nx) fields, spectra by the device rfft."""
    st = jv.FieldState.zeros(conf)
    st.phi[...] = phi.T
    st.psi[...] = psi.T
    st.phi_hat[...] = ops.rfft_accurate(st.phi)
    st.psi_hat[...] = ops.rfft_accurate(st.psi)
    st.Q = st.R1 = st.R2 = st.R3 = 1.0
    return st


def paired_runs(jv, so, conf, phi, psi, nsteps, precision=0, check_every=1, tol=None):
    """Runs the oracle and the device side by side; yields (n, oracle_state, device_state, stats)."""
    ops = jv.Operators(conf, precision=precision)
    ops.set_state(jv.SAV_PRESENT, device_state(jv, ops, conf, phi, psi))
    ops.init_sav()
    orc = so.Oracle(conf)
    pres, old = orc.init_sav(so.state_from_physical(phi, psi, orc.ops, conf.ndim))
    try:
        for n in range(1, nsteps + 1):
            out = orc.step(pres, old, conf.dt, n)
            rc, st = ops.step(conf.dt, n)
            if n % check_every == 0 or n == nsteps:
                yield n, pres, old, out, ops.get_state(jv.SAV_PRESENT), st, rc
    finally:
        ops.close()


def oracle_view(so, st):
    """A device FieldState (Julia shapes) as an oracle State (memory-order arrays)."""
    arrs = {k: np.ascontiguousarray(np.asarray(getattr(st, k)).T)
            for k in ("phi", "phi_hat", "psi", "psi_hat", "u", "u_hat", "p", "p_hat", "mu", "mu_hat", "nu", "nu_hat")}
    return so.State(**arrs, Q=st.Q, R1=st.R1, R2=st.R2, R3=st.R3, A0=np.array(st.A0, dtype=float), dt=st.dt)


def scheme_residuals(so, conf, o, pres, old, new, n):
    """Residual of each equation of the coupled BDF-SAV scheme (Solvers.jl:14-498) that one
    step from (pres, old) to `new` must satisfy, relative to the size of that equation's
    terms (oracle States; `o` the oracle's Ops; pres.dt is the step's dt):

      phi:  (a phi' + b phi^n + c phi^{n-1})/dt = -M_phi mu' - Q' (u*.grad phi*)
      mu:   mu' = L_phi phi' + R1' H1 + R2' H2 + R3' H3        (H_i, G at phi*, psi*)
      psi:  (a psi' + b psi^n + c psi^{n-1})/dt = div(M_psi grad nu') - Q' (u*.grad psi*)
            (S4 psi' inside the psi solve's operator, R3' G through physical space; the
            psi mobility must be constant: the full-complex 2D solve then keeps Hermitian
            symmetry, and its half spectrum is the whole solution)
      nu:   nu' = S4 psi' + R3' G
      R_i:  a R_i' + b R_i^n + c R_i^{n-1} = 1/2 int H_i (a phi' + b phi^n + c phi^{n-1})
            (+ 1/2 int G (a psi' + ...) for R3)
      div:  div u' = 0
      u~:   (a u~ + b u^n + c u^{n-1})/dt - eta Lap u~ = -grad p^n - Q' (conv* + lambda coupl*),
            u~ recovered from u' = u~ - (dt/a) grad dp, p' = p^n + dp - eta div u~,
            Lap dp = (a/dt) div u~
      Q:    (a Q' + b Q^n + c Q^{n-1})/dt = dV [lambda (<adv_phi, mu'> + <adv_psi, nu'>
                                               + <coupl*, u~>) + <conv*, u~>]

    Every arrangement the seven steps make of these equations (split solutions, the R
    system, the Q equation, combinations, projection) is undone here, so a wrong sign,
    coefficient or index in any step leaves a residual far above round-off."""
    c = conf
    dt = pres.dt
    a, b, cc = so.bdf_coeff(n)
    nd = o.ndim
    rf, irf = o.rfft, o.irfft
    phi_s = 2 * pres.phi - old.phi
    psi_s = 2 * pres.psi - old.psi
    phi_s_hat = 2 * pres.phi_hat - old.phi_hat
    psi_s_hat = 2 * pres.psi_hat - old.psi_hat
    u_s = 2 * pres.u - old.u
    u_s_hat = 2 * pres.u_hat - old.u_hat
    mu_s_hat = 2 * pres.mu_hat - old.mu_hat
    nu_s_hat = 2 * pres.nu_hat - old.nu_hat
    adv_phi = sum(u_s[d] * irf(o.D1[d] * phi_s_hat) for d in range(nd))
    adv_psi = sum(u_s[d] * irf(o.D1[d] * psi_s_hat) for d in range(nd))
    H1 = so.get_H1(phi_s, o, c, pres.A0)
    H2 = so.get_H2(phi_s, o, c)
    H3 = so.get_H3(phi_s, psi_s, c, o)
    G = so.get_MG(phi_s, psi_s, c, o)
    Hh = [rf(H1), rf(H2), rf(H3)]
    Gh = rf(G)
    L_phi = c.S1 * o.Biharmonic - c.S2 * o.Laplacian + c.S3
    ph, ps, mu, nu = new.phi_hat, new.psi_hat, new.mu_hat, new.nu_hat
    R = (new.R1, new.R2, new.R3)
    Q = new.Q

    def rel(lhs, rhs):
        return float(np.max(np.abs(lhs - rhs)) / max(np.max(np.abs(lhs)), np.max(np.abs(rhs)), 1e-300))

    out = {}
    out["phi"] = rel((a * ph + b * pres.phi_hat + cc * old.phi_hat) / dt, -c.M_phi * mu - Q * rf(adv_phi))
    out["mu"] = rel(mu, L_phi * ph + R[0] * Hh[0] + R[1] * Hh[1] + R[2] * Hh[2])
    Mpsi = so.get_M_psi(phi_s, c) if nd == 2 else c.M0_psi
    Mc = float(np.max(Mpsi))
    flux_g = sum(o.D1[d] * rf(Mpsi * irf(o.D1[d] * Gh)) for d in range(nd))
    out["psi"] = rel((a * ps + b * pres.psi_hat + cc * old.psi_hat) / dt,
                     Mc * c.S4 * o.Laplacian * ps + R[2] * flux_g - Q * rf(adv_psi))
    out["nu"] = rel(nu, c.S4 * ps + R[2] * Gh)
    dv = o.dvol(c)

    def integ(f, g):
        return float(np.sum(f * g)) * dv
    phi_t = irf(a * ph + b * pres.phi_hat + cc * old.phi_hat)
    psi_t = irf(a * ps + b * pres.psi_hat + cc * old.psi_hat)
    Rn, Ro = (pres.R1, pres.R2, pres.R3), (old.R1, old.R2, old.R3)
    for i, Hi in enumerate((H1, H2, H3)):
        lhs = a * R[i] + b * Rn[i] + cc * Ro[i]
        rhs = 0.5 * integ(Hi, phi_t) + (0.5 * integ(G, psi_t) if i == 2 else 0.0)
        out[f"R{i + 1}"] = abs(lhs - rhs) / max(abs(lhs), abs(rhs), abs(b * Rn[i]), 1e-300)
    div_new = sum(o.D1[d] * new.u_hat[d] for d in range(nd))
    out["div"] = float(np.max(np.abs(div_new.reshape(-1)[1:]))) / max(float(np.max(np.abs(new.u_hat))), 1e-300) \
        / max(float(np.max(np.abs(o.D1[0]))), 1e-300)
    # recover u~ and dp from u', p' and p^n
    dp = (new.p_hat - pres.p_hat) / (1.0 - c.eta * (dt / a) * o.Laplacian)
    ut = np.stack([new.u_hat[d] + (dt / a) * o.D1[d] * dp for d in range(nd)])
    conv = np.stack([sum(u_s[d] * irf(o.D1[d] * u_s_hat[comp]) for d in range(nd)) for comp in range(nd)])
    coupl = np.stack([np.sum(phi_s * irf(o.D1[d] * mu_s_hat) + psi_s * irf(o.D1[d] * nu_s_hat), axis=0)
                      for d in range(nd)])
    lhs = np.stack([(a * ut[d] + b * pres.u_hat[d] + cc * old.u_hat[d]) / dt - c.eta * o.Laplacian * ut[d]
                    for d in range(nd)])
    rhs = np.stack([-o.D1[d] * pres.p_hat - Q * rf(conv[d] + c.lamda * coupl[d]) for d in range(nd)])
    out["u_tilde"] = rel(lhs, rhs)
    ut_phys = [irf(ut[d]) for d in range(nd)]
    theta = dv * (c.lamda * (float(np.sum(adv_phi * irf(mu))) + float(np.sum(adv_psi * irf(nu)))
                             + sum(float(np.sum(coupl[d] * ut_phys[d])) for d in range(nd)))
                  + sum(float(np.sum(conv[d] * ut_phys[d])) for d in range(nd)))
    lhs_q = (a * Q + b * pres.Q + cc * old.Q) / dt
    out["Q"] = abs(lhs_q - theta) / max(abs(lhs_q), abs(theta), abs(b * pres.Q / dt), 1e-300)
    return out
