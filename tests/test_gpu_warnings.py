"""The reference's warning paths on the device, against the oracle run the same way.

* BiCGSTAB stopping at itmax (Solvers.jl:73-76): the step is committed with a warning.
  With itmax = 2 every solve stops early; the committed iterate of the batched device
  solve must equal the oracle's Krylov.jl recurrence after exactly two iterations, in
  eager steps and in replayed graphs (whose iteration budget then equals itmax).
* theta_2 >= 0 in 3D (Solvers_3d.jl:388): with lamda = 0 and u = 0 every theta term
  vanishes, theta_2 = 0, and each step warns and is committed."""
import numpy as np
import pytest

from tests.helpers import relerr, device_state

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10


def _pair(jv, so, conf, phi, psi, itmax=20, graph=False):
    ops = jv.Operators(conf, precision=jv.SAV_PREC_FP64, itmax=itmax)
    ops.set_state(jv.SAV_PRESENT, device_state(jv, ops, conf, phi, psi))
    ops.init_sav()
    ops.set_graph(graph)
    orc = so.Oracle(conf, itmax=itmax)
    pres, old = orc.init_sav(so.state_from_physical(phi, psi, orc.ops, conf.ndim))
    return ops, orc, pres, old


@pytest.mark.parametrize("graph", [False, True])
def test_itmax_commits_the_iterate_with_a_warning(jv, so, graph):
    wl = jv.workload_c1(n=64, steps=4)
    conf = wl.conf
    phi, psi = wl.initial_fields()
    ops, orc, pres, old = _pair(jv, so, conf, phi, psi, itmax=2, graph=graph)
    try:
        for n in range(1, 5):
            nw = len(orc.warnings)
            out = orc.step(pres, old, conf.dt, n)
            rc, st = ops.step(conf.dt, n)
            print(f"graph={graph} step {n}: rc {rc} iters {list(st.krylov_iters)} "
                  f"solved {list(st.krylov_solved)}", flush=True)
            assert rc == jv.SAV_W_KRYLOV_NOT_CONVERGED
            # at step 1 u = 0, so the advection solve (psi_21) has a zero right-hand side and
            # returns at once; every solve with a non-zero right-hand side stops at itmax
            assert list(st.krylov_iters) == [k.niter for k in out["krylov"]]
            assert list(st.krylov_solved) == [int(k.solved) for k in out["krylov"]]
            unsolved = 3 - sum(st.krylov_solved)
            assert unsolved >= 2 and st.krylov_iters[0] == st.krylov_iters[1] == 2
            assert len(orc.warnings) - nw == unsolved
            g = ops.get_state(jv.SAV_PRESENT)
            for name in ("phi", "psi", "mu", "nu", "phi_hat", "psi_hat"):
                e = relerr(getattr(g, name).T, getattr(pres, name))
                assert e < FP64_TOL, (n, name, e)
            assert relerr(g.u.T, pres.u) < 1e-8
            assert abs(st.energy - out["energy"]) <= FP64_TOL * abs(out["energy"])
        if graph:
            assert ops.graph_info()["replays"] >= 2
    finally:
        ops.close()


def test_theta2_nonnegative_warns_in_3d(jv, so):
    wl = jv.workload_c4(n=16, steps=3)
    S = jv.stabilisers(0.5, 0.1, 1000, 1000)
    conf = wl.conf.with_(epsilon=0.5, S1=S[0], S3=S[2], lamda=0)
    phi, psi = jv.initial_fields(conf, wl.vesicles, wl.c_out)
    ops, orc, pres, old = _pair(jv, so, conf, phi, psi)
    try:
        for n in range(1, 4):
            nw = len(orc.warnings)
            out = orc.step(pres, old, conf.dt, n)
            rc, st = ops.step(conf.dt, n)
            assert rc == jv.SAV_W_THETA2_NONNEG
            assert any("theta_2" in w for w in orc.warnings[nw:])
            g = ops.get_state(jv.SAV_PRESENT)
            for name in ("phi", "psi", "mu", "nu"):
                assert relerr(getattr(g, name).T, getattr(pres, name)) < FP64_TOL, name
            assert g.Q == pytest.approx(pres.Q, abs=1e-14)
    finally:
        ops.close()


def test_theta2_negative_does_not_warn_in_3d(jv, so):
    wl = jv.workload_c4(n=16, steps=2)
    S = jv.stabilisers(0.5, 0.1, 1000, 1000)
    conf = wl.conf.with_(epsilon=0.5, S1=S[0], S3=S[2])
    phi, psi = jv.initial_fields(conf, wl.vesicles, wl.c_out)
    ops, orc, pres, old = _pair(jv, so, conf, phi, psi)
    try:
        for n in range(1, 3):
            orc.step(pres, old, conf.dt, n)
            rc, st = ops.step(conf.dt, n)
            assert rc == 0
        assert not orc.warnings
    finally:
        ops.close()
