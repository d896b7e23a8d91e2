"""Device SAV path against the FP64 oracle at the benchmark workloads' full sizes: the
first two steps (BDF1, then BDF2) of two vesicles on 1024^2 and of two ellipsoids on 128^3,
from the same seeded initial fields.  The oracle takes some tens of seconds per case on the
host cores."""
import os

import numpy as np
import pytest

from tests.helpers import relerr, device_state

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10   # relative max-norm, FP64 mode
U_TOL = 1e-8       # velocity and pressure, relative max-norm


def _compare_two_steps(jv, so, wl):
    conf = wl.conf
    phi, psi = wl.initial_fields()
    ops = jv.Operators(conf, precision=jv.SAV_PREC_FP64)
    try:
        ops.set_state(jv.SAV_PRESENT, device_state(jv, ops, conf, phi, psi))
        ops.init_sav()
        orc = so.Oracle(conf, workers=min(os.cpu_count() or 1, 16))
        pres, old = orc.init_sav(so.state_from_physical(phi, psi, orc.ops, conf.ndim))
        for n in (1, 2):
            out = orc.step(pres, old, conf.dt, n)
            rc, st = ops.step(conf.dt, n)
            print(f"{wl.name}: step {n} done (rc {rc})", flush=True)
            assert rc == 0
            g = ops.get_state(jv.SAV_PRESENT)
            errs = {name: relerr(getattr(g, name).T, getattr(pres, name))
                    for name in ("phi", "psi", "mu", "nu", "phi_hat", "psi_hat", "u", "p")}
            # the modified energy is the O(1-10) remainder of terms of size lambda (C1 + C2 + C3)
            # (1e5-1e6 here) that cancel: it is compared at the scale of those terms
            scale = abs(conf.lamda) * (abs(conf.C1) + abs(conf.C2) + abs(conf.C3))
            errs["energy"] = abs(st.energy - out["energy"]) / max(scale, abs(out["energy"]))
            print(f"{wl.name}: step {n} errors " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
            for name in ("phi", "psi", "mu", "nu", "phi_hat", "psi_hat", "energy"):
                assert errs[name] < FP64_TOL, (n, name, errs[name])
            # the velocity and pressure are driven by grad mu and grad nu: the FFT round-off
            # of the two implementations, multiplied by k up to N/2, sets their floor
            assert errs["u"] < U_TOL and errs["p"] < U_TOL, (n, errs["u"], errs["p"])
            for s in ("Q", "R1", "R2", "R3"):
                assert abs(getattr(g, s) - getattr(pres, s)) <= FP64_TOL * max(1.0, abs(getattr(pres, s))), s
            if conf.ndim == 2:
                assert list(st.krylov_iters) == [k.niter for k in out["krylov"]]
    finally:
        ops.close()


def test_two_vesicles_1024_match_oracle(jv, so):
    _compare_two_steps(jv, so, jv.workload_c3())


def test_two_ellipsoids_128cubed_match_oracle(jv, so):
    _compare_two_steps(jv, so, jv.workload_c4())


def test_four_vesicles_512_match_oracle(jv, so):
    _compare_two_steps(jv, so, jv.workload_c2())


def test_ensemble_members_match_oracle(jv, so):
    # one member of each grid size of the 64-instance workload (384^2 is a radix-3 length),
    # stepped inside an Ensemble from host threads, against the oracle of each member
    wls = jv.workload_c5(steps=2)
    picks = []
    for n in (128, 256, 384, 512):
        size = [w for w in wls if w.conf.Nx == n]
        # the member with the most vesicles (and, among those, the stiffest bending)
        picks.append(max(size, key=lambda w: (w.conf.N, w.conf.gamma_bend)))
    assert any(w.conf.S2 > 0 for w in wls)
    ens = jv.Ensemble([w.conf for w in picks], precision=jv.SAV_PREC_FP64)
    try:
        for m, wl in zip(ens.members, picks):
            phi, psi = wl.initial_fields()
            m.set_state(jv.SAV_PRESENT, device_state(jv, m, wl.conf, phi, psi))
            m.init_sav()
        t = np.arange(3) * picks[0].conf.dt
        rc, E, A, kit, st = ens.run(t, 1, threads=4)
        assert rc == 0 and not st.any()
        got = [m.get_state(jv.SAV_PRESENT) for m in ens.members]
    finally:
        ens.close()
    for i, wl in enumerate(picks):
        conf = wl.conf
        phi, psi = wl.initial_fields()
        orc = so.Oracle(conf, workers=min(os.cpu_count() or 1, 16))
        pres, old = orc.init_sav(so.state_from_physical(phi, psi, orc.ops, 2))
        for n in (1, 2):
            out = orc.step(pres, old, conf.dt, n)
        errs = {name: relerr(getattr(got[i], name).T, getattr(pres, name))
                for name in ("phi", "psi", "mu", "nu", "u", "p")}
        print(f"member {conf.Nx}^2 N={conf.N} S2={conf.S2:.3g}: " +
              " ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
        for name in ("phi", "psi", "mu", "nu"):
            assert errs[name] < FP64_TOL, (conf.Nx, name, errs[name])
        assert errs["u"] < U_TOL and errs["p"] < U_TOL
        assert E[i][-1] == pytest.approx(out["energy"], rel=0, abs=FP64_TOL * abs(conf.lamda) * (conf.C1 + conf.C2 + conf.C3))
