"""Stability of the two-vesicle 1024^2 workload (FP64) with and without the S2 stabiliser
(config.s2_for): with the reference's S2 = 0 the explicit k^2 bending term is undamped
(bending ratio 0.92 > 1/2), mu grows and R1 and Q collapse within 50 steps; with
S2 = (3 sqrt 2 / 4) gamma_b / eps the run stays bounded and a 1e-10 perturbation of phi does
not grow."""
import numpy as np
import pytest

from tests.helpers import relerr, device_state

pytestmark = pytest.mark.gpu


def _run(jv, conf, wl, steps, scale=1.0):
    phi, psi = wl.initial_fields()
    ops = jv.Operators(conf)
    try:
        ops.set_state(jv.SAV_PRESENT, device_state(jv, ops, conf, phi * scale, psi))
        ops.init_sav()
        Q, R1 = [], []
        for n in range(1, steps + 1):
            rc, st = ops.step(conf.dt, n)
            assert rc >= 0
            Q.append(st.Q)
            R1.append(st.R1)
        return np.array(Q), np.array(R1), ops.get_state(jv.SAV_PRESENT)
    finally:
        ops.close()


def test_without_S2_the_bending_term_is_unstable(jv):
    wl = jv.workload_c3(steps=50)
    Q, R1, st = _run(jv, wl.conf.with_(S2=0.0), wl, 50)
    print(f"S2 = 0: Q {Q[0]:.4f} -> {Q[-1]:.3g}, min R1 {R1.min():.3g}, max|mu| {np.max(np.abs(st.mu)):.3g}")
    assert Q[-1] < 1e-3 and R1.min() < 0.5 and np.max(np.abs(st.mu)) > 100


def test_with_S2_the_workload_is_bounded_and_not_chaotic(jv):
    wl = jv.workload_c3(steps=200)
    assert wl.conf.S2 > 0
    Q, R1, st = _run(jv, wl.conf, wl, 200)
    Qp, R1p, stp = _run(jv, wl.conf, wl, 200, scale=1.0 + 1e-10)
    growth = relerr(stp.phi, st.phi)
    print(f"S2 = {wl.conf.S2:.3f}: Q {Q[0]:.4f} -> {Q[-1]:.4f}, min R1 {R1.min():.6f}, "
          f"max|mu| {np.max(np.abs(st.mu)):.3g}, 1e-10 perturbation of phi -> {growth:.2e}")
    assert Q.min() > 0.9 and R1.min() > 0.999 and np.max(np.abs(st.mu)) < 10
    assert growth < 1e-8
