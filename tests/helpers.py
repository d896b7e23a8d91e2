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
