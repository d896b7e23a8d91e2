"""Device SAV path against the FP64 oracle on identical seeded inputs."""
import numpy as np
import pytest

from tests.helpers import relerr, device_state, paired_runs

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10   # relative max-norm, FP64 mode


def test_functionals_match_oracle(jv, so):
    wl = jv.workload_c1(n=64, steps=1)
    conf = wl.conf
    phi, psi = wl.initial_fields()
    ops = jv.Operators(conf)
    orc = so.Oracle(conf)
    A0 = so.calculate_area(phi, orc.ops, conf) * 0.99
    st = device_state(jv, ops, conf, phi, psi)
    st.A0 = A0
    ops.set_state(jv.SAV_PRESENT, st)
    H1, H2, H3, G, W, area = ops.eval_functionals(phi.T, psi.T)
    assert relerr(area, so.calculate_area(phi, orc.ops, conf)) < 1e-13
    assert abs(W[0] - so.get_W1(phi, orc.ops, conf, A0)) < 1e-10 * abs(W[0])
    assert abs(W[1] - so.get_W2(phi, orc.ops, conf)) < 1e-12 * abs(W[1])
    assert abs(W[2] - so.get_W3(phi, psi, conf, orc.ops)) < 1e-12 * abs(W[2])
    assert relerr(H1.T, so.get_H1(phi, orc.ops, conf, A0)) < 1e-11
    assert relerr(H2.T, so.get_H2(phi, orc.ops, conf)) < 1e-11
    assert relerr(H3.T, so.get_H3(phi, psi, conf, orc.ops)) < 1e-12
    assert relerr(G.T, so.get_MG(phi, psi, conf, orc.ops)) < 1e-12
    ops.close()


def test_psi_solve_matches_oracle(jv, so):
    wl = jv.workload_c1(n=64, steps=1)
    conf = wl.conf
    phi, psi = wl.initial_fields()
    orc = so.Oracle(conf)
    M = so.get_M_psi(phi, conf)
    rng = np.random.default_rng(5)
    rhs = orc.ops.rfft(rng.standard_normal(phi.shape))
    ref = orc.solve_psi(rhs, M, float(M.max()), 1.5, conf.dt)
    ks = orc.krylov_log[-1]
    ops = jv.Operators(conf)
    got, info = ops.psi_solve(phi.T, 1.5, conf.dt, rhs.T)
    assert info["iters"] == ks.niter
    assert info["solved"] == ks.solved
    assert relerr(got.T, ref) < 1e-11
    ops.close()


@pytest.mark.parametrize("n,nsteps", [(32, 3), (128, 4)])
def test_single_vesicle_steps_match_oracle(jv, so, n, nsteps):
    wl = jv.workload_c1(n=n, steps=nsteps)
    conf = wl.conf
    phi, psi = wl.initial_fields()
    for k, pres, old, out, g, st, rc in paired_runs(jv, so, conf, phi, psi, nsteps):
        assert rc == 0
        for name in ("phi", "psi", "mu", "nu", "phi_hat", "psi_hat"):
            e = relerr(getattr(g, name).T, getattr(pres, name))
            assert e < FP64_TOL, (k, name, e)
        assert relerr(g.u.T, pres.u) < 1e-8 or np.max(np.abs(pres.u)) < 1e-14
        for s in ("Q", "R1", "R2", "R3"):
            assert abs(getattr(g, s) - getattr(pres, s)) <= FP64_TOL * max(1.0, abs(getattr(pres, s))), s
        assert abs(st.energy - out["energy"]) <= FP64_TOL * abs(out["energy"])
        assert abs(st.area_ratio - out["area_ratio"]) <= 1e-8 * max(out["area_ratio"], 1e-12) + 1e-15
        assert list(st.krylov_iters) == [s.niter for s in out["krylov"]]


def test_two_vesicles_steps_match_oracle(jv, so):
    wl = jv.workload_c3(n=64, steps=3)
    conf = wl.conf.with_(epsilon=0.12, S1=jv.stabilisers(0.12, 0.1, 1000, 1000)[0],
                         S3=jv.stabilisers(0.12, 0.1, 1000, 1000)[2], dt=1e-4)
    phi, psi = jv.initial_fields(conf, wl.vesicles, wl.c_out)
    for k, pres, old, out, g, st, rc in paired_runs(jv, so, conf, phi, psi, 3):
        for name in ("phi", "psi", "mu", "nu"):
            e = relerr(getattr(g, name).T, getattr(pres, name))
            assert e < FP64_TOL, (k, name, e)
        assert relerr(g.u.T, pres.u) < 1e-9
        assert relerr(g.p.T, pres.p) < 1e-9


def test_3d_steps_match_oracle(jv, so):
    wl = jv.workload_c4(n=16, steps=3)
    conf = wl.conf.with_(epsilon=0.5, S1=jv.stabilisers(0.5, 0.1, 1000, 1000)[0],
                         S3=jv.stabilisers(0.5, 0.1, 1000, 1000)[2])
    phi, psi = jv.initial_fields(conf, wl.vesicles, wl.c_out)
    for k, pres, old, out, g, st, rc in paired_runs(jv, so, conf, phi, psi, 3):
        for name in ("phi", "psi", "mu", "nu"):
            e = relerr(getattr(g, name).T, getattr(pres, name))
            assert e < FP64_TOL, (k, name, e)
        assert relerr(g.u.T, pres.u) < 1e-9
        assert abs(st.energy - out["energy"]) <= FP64_TOL * abs(out["energy"])


def test_fp32_mode_close_to_fp64(jv, so):
    wl = jv.workload_c1(n=128, steps=3)
    conf = wl.conf
    phi, psi = wl.initial_fields()
    errs = []
    for k, pres, old, out, g, st, rc in paired_runs(jv, so, conf, phi, psi, 3, precision=1):
        errs.append(relerr(g.phi.T, pres.phi))
    assert max(errs) < 1e-4, errs
