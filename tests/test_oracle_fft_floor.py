"""The round-off floor between two correct FFT implementations, on the CPU.

The oracle is run twice from the same inputs, once with scipy.fft's r2c/c2r/c2c and once
with oracle.AltOps (the same transforms by another route).  Every field agrees to round-off,
but u and p, which the projection builds from grad mu and grad nu, differ by far more than
phi and psi: mu carries S1 k^4 times the round-off of phi_hat at high k.  This is the
floor the device (a third FFT implementation) is compared against at full size
(tests/test_gpu_full_size.py).  Measured with this comparison: after two steps of two
vesicles on 1024^2, u 1.0e-9 and p 1.7e-10 (the device against scipy.fft: u 1.1e-9,
p 1.2e-10); on one vesicle at 128^2, u 7.9e-12 (the device: 9.9e-12).  Set
SAV_FLOOR_1024=1 to include the 1024^2 case (a few minutes on 8 cores)."""
import os

import numpy as np
import pytest

from tests.helpers import relerr

CASES = [("c1", 128)] + ([("c3", 1024)] if os.environ.get("SAV_FLOOR_1024") else [])


@pytest.mark.parametrize("name,n", CASES)
def test_two_fft_routes_agree_to_round_off(jv, so, name, n):
    wl = {"c1": jv.workload_c1, "c3": jv.workload_c3}[name](n=n, steps=2)
    conf = wl.conf
    phi, psi = wl.initial_fields()
    states = []
    for alt in (False, True):
        orc = so.Oracle(conf, workers=min(os.cpu_count() or 1, 16))
        if alt:
            orc.ops.__class__ = so.AltOps
        pres, old = orc.init_sav(so.state_from_physical(phi, psi, orc.ops, 2))
        for k in (1, 2):
            orc.step(pres, old, conf.dt, k)
        states.append(pres)
    errs = {k: relerr(getattr(states[1], k), getattr(states[0], k)) for k in ("phi", "psi", "mu", "nu", "u", "p")}
    print(name, n, errs)
    assert errs["phi"] < 1e-13 and errs["psi"] < 1e-13
    assert errs["mu"] < 1e-10 and errs["nu"] < 1e-10
    # the velocity and pressure floor: well above FP64 epsilon, below the 1e-8 of the device test
    assert 1e-13 < errs["u"] < 1e-8 and errs["p"] < 1e-8
