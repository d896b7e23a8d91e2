"""Snapshot schedule and files, on the CPU: the save-index rule of simulation.jl:113-114
checked the way test/test_save_interval.jl checks it, and the MAT keys of
test/verify_save_state.jl."""
import numpy as np


def test_save_schedule_hits_every_interval(jv):
    # test_save_interval.jl: Dt = [0, 1e-7, 1e-5, 2e-5, ...] up to 0.015, jittered by 1e-18
    Dt = [0.0, 1e-7]
    t = 1e-5
    while t <= 0.015 + 1e-7:
        Dt.append(t)
        t += 1e-5
    rng = np.random.default_rng(1)
    Dt = np.array(Dt) + (rng.random(len(Dt)) - 0.5) * 1e-18
    # the reference test walks n over all nodes (t = Dt[n]); save_steps' n-th step starts at Dt[n-1]
    steps = jv.save_steps(np.append(Dt, Dt[-1] + 1e-5), 0.001)
    saved = [Dt[n - 1] for n in steps]
    for e in np.arange(0.0, 0.015 + 1e-9, 0.001):
        assert any(abs(s - e) < 1e-10 for s in saved), e
    assert len(saved) == 16


def test_save_schedule_first_step_and_names(jv):
    Dt = jv.time_nodes(1e-4, 3e-3)
    steps = jv.save_steps(Dt, 0.001)
    assert steps[0] == 1                      # curr index 0 > -1 at t = 0
    conf = jv.workload_c1(n=32, steps=1).conf
    assert jv.snapshot_name(conf, 1e-3) == "state_2d_t1.00e-03.mat"


def test_snapshot_keys_and_values(jv, tmp_path):
    conf = jv.workload_c1(n=16, steps=1).conf
    st = jv.FieldState.zeros(conf)
    rng = np.random.default_rng(2)
    for k in ("phi", "psi", "mu", "nu", "p", "u"):
        getattr(st, k)[...] = rng.random(getattr(st, k).shape)
    st.R1, st.R2, st.R3, st.Q = 1.1, 2.2, 3.3, 4.4
    path = tmp_path / "state_2d_final.mat"
    jv.write_snapshot(path, st)
    d = jv.read_snapshot(path)
    assert set(d) == set(jv.SNAPSHOT_KEYS)
    assert d["R1"] == 1.1 and d["Q"] == 4.4
    assert d["phi"].shape == (16, 16, 1)
    assert np.array_equal(d["u"], st.u)


def test_checkpoint_roundtrip(jv, tmp_path):
    conf = jv.workload_c1(n=16, steps=1).conf
    a, b = jv.FieldState.zeros(conf), jv.FieldState.zeros(conf)
    rng = np.random.default_rng(3)
    for st in (a, b):
        for k in jv.FieldState.ARRAYS:
            arr = getattr(st, k)
            arr[...] = rng.random(arr.shape) + (1j * rng.random(arr.shape) if np.iscomplexobj(arr) else 0)
        st.Q, st.R1, st.R2, st.R3, st.dt = rng.random(5)
        st.A0 = rng.random(1)
    jv.write_checkpoint(tmp_path / "c.npz", a, b, 7, 0.25)
    p, o, n, t = jv.read_checkpoint(tmp_path / "c.npz")
    assert n == 7 and t == 0.25
    for x, y in ((a, p), (b, o)):
        for k in jv.FieldState.ARRAYS:
            assert np.array_equal(getattr(x, k), getattr(y, k))
        assert (x.Q, x.R1, x.R2, x.R3, x.dt) == (y.Q, y.R1, y.R2, y.R3, y.dt)


def test_initial_fields_3d(jv):
    base = jv.workload_c4(n=32, steps=1).conf
    conf = base.with_(N=1, goal="s", psi_in_v=[1.0], psi_out_v=[1.0])
    phi, psi = jv.initial_fields_3d(conf, 1)
    assert phi.shape == (32, 32, 32, 1)
    assert phi[16, 16, 16, 0] > 0.99 and phi[0, 0, 0, 0] < -0.99
    assert np.allclose(psi, -0.1 * phi + 0.7)
    # the two-star shape is sized for the unit cube (Init_3d.jl centres at x = 0.73, 1.27)
    conf2 = jv.workload_c4(n=64, steps=1).conf.with_(N=2, goal="g", Lx=2.0, Ly=1.0, Lz=1.0, epsilon=0.01)
    phi2, psi2 = jv.initial_fields_3d(conf2, 7)
    i73 = int(round(0.73 / (conf2.Lx / 64)))
    assert phi2[i73, 32, 32, 0] > 0.99 and phi2[i73, 32, 32, 1] < -0.99
    assert np.allclose(psi2, -0.35 * phi2 + 0.45)
    import pytest
    with pytest.raises(ValueError):
        jv.initial_fields_3d(conf2.with_(goal="s"), 7)
