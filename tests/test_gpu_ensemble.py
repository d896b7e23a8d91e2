"""Ensemble of independent instances on the GPU: running members concurrently from
several host threads gives bit-identical results to running each instance alone
(same kernels, deterministic reductions, one stream per member)."""
import numpy as np
import pytest

from helpers import device_state

pytestmark = pytest.mark.gpu


def _members(jv):
    wls = jv.workload_c5(steps=3, sizes=(64, 96), per_size=2)
    return [w.conf for w in wls], wls


@pytest.mark.parametrize("precision", [0, 1])
def test_ensemble_matches_single_runs(jv, precision):
    confs, wls = _members(jv)
    nsteps = 3
    t = np.arange(nsteps + 1) * confs[0].dt
    ens = jv.Ensemble(confs, precision=precision)
    for m, wl in zip(ens.members, wls):
        phi, psi = wl.initial_fields()
        m.set_state(jv.SAV_PRESENT, device_state(jv, m, wl.conf, phi, psi))
        m.init_sav()
    rc, E, A, kit, st = ens.run(t, 1, threads=4)
    assert rc >= 0 and np.all(st >= 0)
    got = [m.get_state(jv.SAV_PRESENT) for m in ens.members]
    ens.close()
    for i, wl in enumerate(wls):
        ops = jv.Operators(wl.conf, precision=precision)
        phi, psi = wl.initial_fields()
        ops.set_state(jv.SAV_PRESENT, device_state(jv, ops, wl.conf, phi, psi))
        ops.init_sav()
        rc1, E1, A1, kit1 = ops.run(t, 1)
        ref = ops.get_state(jv.SAV_PRESENT)
        ops.close()
        assert kit[i] == kit1
        assert np.array_equal(E[i], E1)
        for name in ("phi", "psi", "mu", "nu", "u", "p"):
            assert np.array_equal(getattr(got[i], name), getattr(ref, name)), (i, name)
        assert got[i].Q == ref.Q


@pytest.mark.parametrize("precision", [0, 1])
def test_batched_size_classes_match_single_runs(jv, precision):
    """Size classes (sav_ensemble_set_batched): members with the same grid and vesicle count
    run each step's psi solves as one batched solve over all their systems.  Every member's
    states, energies and iteration counts stay bit-identical to its own run, for grouped
    members and for those left alone; after decoupling (set_batched(False)) the members keep
    solving on their own from the group's buffers, still bit-identical."""
    wls = jv.workload_c5(steps=5, sizes=(64, 96), per_size=6)
    confs = [w.conf for w in wls]
    t = np.arange(6) * confs[0].dt
    ens = jv.Ensemble(confs, precision=precision)
    ng = ens.set_batched(True)
    groups, _ = ens.batched_info()
    print("groups", ng, groups.tolist())
    assert ng >= 2 and np.sum(groups >= 0) >= 4
    for m, wl in zip(ens.members, wls):
        phi, psi = wl.initial_fields()
        m.set_state(jv.SAV_PRESENT, device_state(jv, m, wl.conf, phi, psi))
        m.init_sav()
    rc, E, A, kit, st = ens.run(t[:4], 1)
    assert rc >= 0 and np.all(st >= 0)
    groups, solves = ens.batched_info()
    assert np.all(solves[groups >= 0] == 3)
    mid = [m.get_state(jv.SAV_PRESENT) for m in ens.members]
    # between batched runs a member is stepped through its own handle (solves alone)
    probe = ens.members[int(np.flatnonzero(groups >= 0)[0])]
    saved = (probe.get_state(jv.SAV_PRESENT), probe.get_state(jv.SAV_OLD))
    rcp, _ = probe.step(confs[0].dt, 4)
    assert rcp >= 0
    probe.set_state(jv.SAV_PRESENT, saved[0])
    probe.set_state(jv.SAV_OLD, saved[1])
    ens.set_batched(False)
    rc2, E2, A2, kit2, st2 = ens.run(t[3:], 4, threads=4)
    assert rc2 >= 0
    got = [m.get_state(jv.SAV_PRESENT) for m in ens.members]
    ens.close()
    for i, wl in enumerate(wls):
        ops = jv.Operators(wl.conf, precision=precision)
        phi, psi = wl.initial_fields()
        ops.set_state(jv.SAV_PRESENT, device_state(jv, ops, wl.conf, phi, psi))
        ops.init_sav()
        _, E1, _, kit1 = ops.run(t[:4], 1)
        ref_mid = ops.get_state(jv.SAV_PRESENT)
        _, E1b, _, kit1b = ops.run(t[3:], 4)
        ref = ops.get_state(jv.SAV_PRESENT)
        ops.close()
        assert kit[i] == kit1 and kit2[i] == kit1b, i
        assert np.array_equal(E[i], E1) and np.array_equal(E2[i], E1b), i
        for name in ("phi", "psi", "mu", "nu", "u", "p"):
            assert np.array_equal(getattr(mid[i], name), getattr(ref_mid, name)), (i, name)
            assert np.array_equal(getattr(got[i], name), getattr(ref, name)), (i, name)
