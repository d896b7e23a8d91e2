"""run_simulation with snapshots on the GPU: files on the reference's schedule, and a
resume from a checkpoint that reproduces the uninterrupted run exactly."""
import numpy as np
import pytest

from helpers import device_state

pytestmark = pytest.mark.gpu


def test_snapshots_and_resume(jv, tmp_path):
    wl = jv.workload_c1(n=32, steps=12)
    conf = wl.conf
    phi, psi = wl.initial_fields()
    t = np.arange(13) * conf.dt
    ops = jv.Operators(conf)
    st0 = device_state(jv, ops, conf, phi, psi)
    ops.close()
    full, E, A, Dt = jv.run_simulation(conf.dt, t[-1], 1, conf=conf, state=st0.copy(), t_nodes=t)
    part, E2, A2, _ = jv.run_simulation(conf.dt, t[-1], 1, conf=conf, state=st0.copy(), t_nodes=t,
                                        save_path=str(tmp_path), save_interval=4 * conf.dt, checkpoint=True)
    assert np.array_equal(part.phi, full.phi) and E2 == E
    folder = tmp_path / "MAT" / "case3"
    names = sorted(p.name for p in folder.glob("*.mat"))
    expect = sorted(jv.snapshot_name(conf, float(t[n - 1])) for n in jv.save_steps(t, 4 * conf.dt))
    assert names == expect and len(names) == 3
    # resume after the snapshot of step 5 (written under t[4])
    ck = folder / (jv.snapshot_name(conf, float(t[4]))[:-4] + ".ckpt.npz")
    res, E3, A3 = jv.resume_simulation(ck, conf, t[5:])
    assert np.array_equal(res.phi, full.phi) and np.array_equal(res.psi, full.psi)
    assert np.allclose(E3, E[5:], rtol=0, atol=0)
