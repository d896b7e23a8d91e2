"""Mixed-precision mode (SAV_PREC_MIXED): FP64 state and stages, with the 2D psi solve's
inner iterations in FP32 inside an FP64 defect correction that stops on the reference's
preconditioned residual test (Solvers.jl:73-75) taken on the true residual.

Checked against the FP64 oracle (the true residual of each solve, the fields after a few
steps) and against the FP64 device form over whole benchmark runs, where the accuracy
contract is <= 1e-5 relative max-norm error in phi, psi and u and a non-increasing
modified energy.  The FP32-fields form is held to the same contract and is expected to
miss it (its measured errors are in the xfail reason)."""
import numpy as np
import pytest

from tests.helpers import relerr, device_state

pytestmark = pytest.mark.gpu

CONTRACT = 1e-5


def test_mixed_psi_solve_meets_the_reference_tolerance(jv, so):
    wl = jv.workload_c1(n=64, steps=1)
    conf = wl.conf
    phi, psi = wl.initial_fields()
    M = so.get_M_psi(phi, conf)
    rng = np.random.default_rng(11)
    orc = so.Oracle(conf)
    rhs = orc.ops.rfft(rng.standard_normal(phi.shape))
    ref = orc.solve_psi(rhs, M, float(M.max()), 1.5, conf.dt)          # rtol 1e-10, the reference's
    exact = so.Oracle(conf.with_(tol=1e-15), itmax=200).solve_psi(rhs, M, float(M.max()), 1.5, conf.dt)
    ops = jv.Operators(conf, precision=jv.SAV_PREC_MIXED)
    got, info = ops.psi_solve(phi.T, 1.5, conf.dt, rhs.T)
    ops.close()
    assert info["solved"]
    # the reported norm is the FP64 true preconditioned residual, under the reference's threshold
    Pb = orc.krylov_log[-1].residuals[0]          # ||P b||, the oracle's initial residual (x0 = 0)
    assert info["rnorm"] <= 1e-14 + conf.tol * Pb
    # the reference's solve and the mixed one are both within the tolerance of the exact
    # solution: neither is nearer than the other by more than the conditioning allows
    e_ref, e_mix = relerr(ref, exact), relerr(got.T, exact)
    print(f"error vs converged solve: reference tolerance {e_ref:.2e}, mixed {e_mix:.2e}")
    assert e_mix < 1e-9 and e_ref < 1e-9


def test_mixed_steps_match_oracle(jv, so):
    wl = jv.workload_c1(n=64, steps=4)
    conf = wl.conf
    phi, psi = wl.initial_fields()
    ops = jv.Operators(conf, precision=jv.SAV_PREC_MIXED)
    ops.set_state(jv.SAV_PRESENT, device_state(jv, ops, conf, phi, psi))
    ops.init_sav()
    orc = so.Oracle(conf)
    pres, old = orc.init_sav(so.state_from_physical(phi, psi, orc.ops, 2))
    try:
        for n in range(1, 5):
            out = orc.step(pres, old, conf.dt, n)
            rc, st = ops.step(conf.dt, n)
            assert rc == 0 and all(st.krylov_solved)
            g = ops.get_state(jv.SAV_PRESENT)
            for name, tol in (("phi", 1e-9), ("psi", 1e-9), ("mu", 1e-7), ("u", 1e-6)):
                e = relerr(getattr(g, name).T, getattr(pres, name))
                assert e < tol, (n, name, e)
            assert abs(st.energy - out["energy"]) <= 1e-9 * max(1.0, abs(out["energy"]))
    finally:
        ops.close()


def _run(jv, wl, precision, steps):
    conf = wl.conf
    phi, psi = wl.initial_fields()
    ops = jv.Operators(conf, precision=precision)
    try:
        ops.set_state(jv.SAV_PRESENT, device_state(jv, ops, conf, phi, psi))
        ops.init_sav()
        rc, E, A, kit = ops.run(np.arange(steps + 1) * conf.dt, 1)
        return rc, E, ops.get_state(jv.SAV_PRESENT)
    finally:
        ops.close()


CASES = [("single_vesicle_128", 200), ("four_vesicles_512", 300), ("two_vesicles_1024", 100)]


def _workload(jv, name):
    return {"single_vesicle_128": jv.workload_c1, "four_vesicles_512": jv.workload_c2,
            "two_vesicles_1024": jv.workload_c3}[name]()


@pytest.mark.parametrize("name,steps", CASES)
def test_mixed_meets_accuracy_contract(jv, name, steps):
    wl = _workload(jv, name)
    rc64, E64, s64 = _run(jv, wl, jv.SAV_PREC_FP64, steps)
    rcm, Em, sm = _run(jv, wl, jv.SAV_PREC_MIXED, steps)
    errs = {k: relerr(getattr(sm, k), getattr(s64, k)) for k in ("phi", "psi", "u")}
    print(f"{name} {steps} steps: mixed vs FP64 " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    assert rc64 == 0 and rcm == 0
    for k, v in errs.items():
        assert v <= CONTRACT, (k, v)
    assert np.all(np.diff(Em) <= 0), "mixed energy increases"
    assert relerr(Em, E64) < 1e-6


# measured (bench.py --configs, FP32 fields against FP64): phi 1.0e-4 after 1000 steps of
# four_vesicles_512, u 8.1e-4 after 200 steps of single_vesicle_128, energy not monotone
@pytest.mark.xfail(reason="FP32 fields and transforms: per-step FP32 rounding of state-bound data grows "
                          "past 1e-5 (u 8e-4 at 128^2 / 200 steps, phi 1e-4 at 512^2 / 1000 steps)",
                   strict=False)
def test_fp32_fields_accuracy_contract(jv):
    wl = jv.workload_c1()
    rc64, E64, s64 = _run(jv, wl, jv.SAV_PREC_FP64, 200)
    rc32, E32, s32 = _run(jv, wl, jv.SAV_PREC_FP32, 200)
    for k in ("phi", "psi", "u"):
        assert relerr(getattr(s32, k), getattr(s64, k)) <= CONTRACT, k
    assert np.all(np.diff(E32) <= 0)
