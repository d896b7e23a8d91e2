"""A step whose W + C turns negative (the reference's error("FATAL: W+C<0"), Utils.jl:180)
is reported and not committed: present and old stay exactly as they were, in both
precision modes."""
import numpy as np
import pytest

from tests.helpers import device_state

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [0, 1])
def test_fatal_step_is_not_committed(jv, so, precision):
    wl = jv.workload_c1(n=32, steps=2)
    conf = wl.conf
    phi, psi = wl.initial_fields()
    orc = so.Oracle(conf)
    w0 = so.get_W2(phi, orc.ops, conf)
    # a 2 % rescaling of phi that lowers W2, and C2 half-way between the two values of -W2:
    # the SAV initialisation on phi passes, a step from the rescaled field fails
    fac = min((0.98, 1.02), key=lambda f: so.get_W2(f * phi, orc.ops, conf))
    w1 = so.get_W2(fac * phi, orc.ops, conf)
    assert w1 < w0
    conf2 = conf.with_(C2=-(w0 + w1) / 2)
    ops = jv.Operators(conf2, precision=precision)
    try:
        ops.set_state(jv.SAV_PRESENT, device_state(jv, ops, conf2, phi, psi))
        ops.init_sav()
        bad = device_state(jv, ops, conf2, fac * phi, psi)
        ops.set_state(jv.SAV_PRESENT, bad)
        ops.set_state(jv.SAV_OLD, bad)
        before = [ops.get_state(jv.SAV_PRESENT), ops.get_state(jv.SAV_OLD)]
        with pytest.raises(jv.WCNegativeError):
            ops.step(conf2.dt, 1)
        after = [ops.get_state(jv.SAV_PRESENT), ops.get_state(jv.SAV_OLD)]
        for b, a in zip(before, after):
            for k in jv.FieldState.ARRAYS:
                assert np.array_equal(getattr(a, k), getattr(b, k)), k
            for s in ("Q", "R1", "R2", "R3"):
                assert getattr(a, s) == getattr(b, s), s
        # the handle stays usable: a valid state steps normally
        ops.set_state(jv.SAV_PRESENT, device_state(jv, ops, conf2, phi, psi))
        ops.init_sav()
        rc, st = ops.step(conf2.dt, 1)
        assert rc == 0
    finally:
        ops.close()
