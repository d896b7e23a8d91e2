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
