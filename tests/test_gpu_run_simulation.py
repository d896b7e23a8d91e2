"""The drop-in driver run_simulation (simulation.jl:2-155) against the oracle stepped over
the same schedule: the reference's two-phase time nodes 0:1e-7:1e-5 then dt up to T
(simulation.jl:15-19, 110 steps here), BDF1 at the first step and fixed BDF2 coefficients
across the dt switch (simulation.jl:73-76), the energy and area-ratio histories of
simulation.jl:146-148, the final state, and the snapshot written at each save interval
(simulation.jl:113-138)."""
import numpy as np
import pytest

from tests.helpers import relerr

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10


def test_run_simulation_matches_oracle_over_the_reference_schedule(jv, so, tmp_path):
    conf = jv.workload_c1(n=64, steps=1).conf
    T = 1e-3
    Dt = jv.time_nodes(conf.dt, T)
    assert len(Dt) - 1 == 100 + 9
    present, E, A, Dt_out = jv.run_simulation(conf.dt, T, 1, conf=conf, save_path=str(tmp_path),
                                              save_interval=5e-4)
    assert np.array_equal(Dt_out, Dt)
    # the phi image of every 2D snapshot (simulation.jl:130-135)
    assert len(list(tmp_path.glob("phi_t*.png"))) == len(jv.save_steps(Dt, 5e-4))
    phi, psi = jv.initial_fields_2d(conf, 1)
    orc = so.Oracle(conf)
    pres, old = orc.init_sav(so.state_from_physical(phi.T, psi.T, orc.ops, 2))
    E_ref, A_ref, snaps = [], [], {}
    saves = set(jv.save_steps(Dt, 5e-4))
    for n in range(1, len(Dt)):
        out = orc.step(pres, old, float(Dt[n] - Dt[n - 1]), n)
        E_ref.append(out["energy"])
        A_ref.append(out["area_ratio"])
        if n in saves:
            snaps[jv.snapshot_name(conf, float(Dt[n - 1]))] = pres.phi.copy()
    assert len(E) == len(E_ref) == len(Dt) - 1
    scale = abs(conf.lamda) * (conf.C1 + conf.C2 + conf.C3)
    assert np.max(np.abs(np.array(E) - np.array(E_ref))) <= FP64_TOL * scale
    assert np.max(np.abs(np.array(A) - np.array(A_ref))) <= 1e-8 * max(np.max(np.abs(A_ref)), 1e-12) + 1e-15
    assert np.all(np.diff(E) <= 0)
    for name in ("phi", "psi", "mu", "nu"):
        assert relerr(getattr(present, name).T, getattr(pres, name)) < FP64_TOL, name
    assert relerr(present.u.T, pres.u) < 1e-8
    # the snapshots of the save schedule hold the state after the step they are named for
    files = sorted(p.name for p in (tmp_path / "MAT" / "case3").glob("*.mat"))
    assert files == sorted(snaps)
    for fname, ref_phi in snaps.items():
        got = jv.read_snapshot(tmp_path / "MAT" / "case3" / fname)
        assert relerr(np.asarray(got["phi"]).T, ref_phi) < FP64_TOL, fname
