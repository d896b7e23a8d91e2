"""The physical fields of a state read back after steps (eager and graph-replayed) are the
device back-transforms of its spectra, in the present and in the old state, as
simulation.jl:99-104 defines them (the library transforms mu, nu and p when a state is read,
phi, psi and u with the step that reads them), and reading a state twice, or stepping after a
read, changes nothing."""
import numpy as np
import pytest

from helpers import device_state

pytestmark = pytest.mark.gpu


def _check_state(ops, st):
    for phys, spec in (("phi", "phi_hat"), ("psi", "psi_hat"), ("mu", "mu_hat"), ("nu", "nu_hat"),
                       ("u", "u_hat"), ("p", "p_hat")):
        back = ops.irfft(getattr(st, spec))
        got = getattr(st, phys).reshape(back.shape, order="F")
        assert np.array_equal(got, back), phys


@pytest.mark.parametrize("case,graph", [("2d", False), ("2d", True), ("3d", False)])
def test_physical_fields_are_back_transforms(jv, case, graph):
    wl = jv.workload_c3(n=64, steps=4) if case == "2d" else jv.workload_c4(n=16, steps=4)
    conf = wl.conf
    phi, psi = wl.initial_fields()
    runs = {}
    for read_between in (False, True):
        ops = jv.Operators(conf)
        ops.set_state(jv.SAV_PRESENT, device_state(jv, ops, conf, phi, psi))
        ops.init_sav()
        if graph:
            ops.set_graph(1)
        for n in range(1, 5):
            rc, _ = ops.step(conf.dt, n)
            assert rc >= 0
            if read_between:
                pres = ops.get_state(jv.SAV_PRESENT)
                old = ops.get_state(jv.SAV_OLD)
                _check_state(ops, pres)
                if n >= 2:  # after one step the old state is the initial one, as it was set
                    _check_state(ops, old)
        pres = ops.get_state(jv.SAV_PRESENT)
        again = ops.get_state(jv.SAV_PRESENT)
        old = ops.get_state(jv.SAV_OLD)
        _check_state(ops, pres)
        _check_state(ops, old)
        for k in jv.FieldState.ARRAYS:
            assert np.array_equal(getattr(pres, k), getattr(again, k)), k
        runs[read_between] = pres
        ops.close()
    for k in jv.FieldState.ARRAYS:
        assert np.array_equal(getattr(runs[False], k), getattr(runs[True], k)), k
