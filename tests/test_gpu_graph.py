"""Graph replay of whole steps (sav_set_graph): the replayed steps give bit-identical
states, energies and psi iteration counts to eager steps, also across the reference's
time-step switch (dt 1e-7 -> 1e-4, simulation.jl:15-19), where the psi iteration count
jumps past the graph's budget and the step is repeated without the graph."""
import numpy as np
import pytest

from helpers import device_state

pytestmark = pytest.mark.gpu


def _schedule(dt):
    # the reference's two-phase schedule, shortened: 5 steps of 1e-7, then 6 steps of dt
    t1 = np.arange(6) * 1e-7
    return np.concatenate([t1, t1[-1] + np.arange(1, 7) * dt])


def _fresh(jv, conf, phi, psi, precision, graph):
    ops = jv.Operators(conf, precision=precision)
    ops.set_state(jv.SAV_PRESENT, device_state(jv, ops, conf, phi, psi))
    ops.init_sav()
    ops.set_graph(graph)
    return ops


def _same_state(a, b):
    for name in ("phi", "psi", "mu", "nu", "u", "p", "phi_hat", "psi_hat", "u_hat", "p_hat"):
        assert np.array_equal(getattr(a, name), getattr(b, name)), name
    for s in ("Q", "R1", "R2", "R3"):
        assert getattr(a, s) == getattr(b, s), s


@pytest.mark.parametrize("precision", [0, 1])
def test_graph_steps_equal_eager_steps_2d(jv, precision):
    wl = jv.workload_c1(n=64, steps=1)
    conf = wl.conf
    phi, psi = wl.initial_fields()
    t = _schedule(conf.dt)
    eager = _fresh(jv, conf, phi, psi, precision, False)
    graph = _fresh(jv, conf, phi, psi, precision, True)
    for k in range(1, len(t)):
        dt = t[k] - t[k - 1]
        rc0, s0 = eager.step(dt, k)
        rc1, s1 = graph.step(dt, k)
        assert rc0 == rc1
        assert list(s0.krylov_iters) == list(s1.krylov_iters), k
        assert list(s0.krylov_solved) == list(s1.krylov_solved), k
        assert s0.energy == s1.energy and s0.area_ratio == s1.area_ratio, k
        assert (s0.Q, s0.R1, s0.R2, s0.R3) == (s1.Q, s1.R1, s1.R2, s1.R3), k
    _same_state(eager.get_state(jv.SAV_PRESENT), graph.get_state(jv.SAV_PRESENT))
    _same_state(eager.get_state(jv.SAV_OLD), graph.get_state(jv.SAV_OLD))
    info = graph.graph_info()
    print(info)
    # step 1 runs eagerly (it chooses the matvec's column tile); the dt switch needs more
    # psi iterations than the 1e-7 steps did, which the graph's budget does not cover
    assert info["replays"] >= len(t) - 2
    assert info["repeats"] >= 1
    assert 1 <= info["graphs"] <= 12
    assert eager.graph_info()["replays"] == 0
    eager.close()
    graph.close()


def test_graph_run_equals_eager_run_2d_two_vesicles(jv):
    wl = jv.workload_c3(n=64, steps=3)
    conf = wl.conf.with_(epsilon=0.12, S1=jv.stabilisers(0.12, 0.1, 1000, 1000)[0],
                         S3=jv.stabilisers(0.12, 0.1, 1000, 1000)[2], dt=1e-4)
    phi, psi = jv.initial_fields(conf, wl.vesicles, wl.c_out)
    t = np.arange(9) * conf.dt
    out = []
    for graph in (False, True):
        ops = _fresh(jv, conf, phi, psi, 1, graph)
        rc, E, A, kit = ops.run(t, 1)
        out.append((rc, E, A, kit, ops.get_state(jv.SAV_PRESENT), ops.graph_info()))
        ops.close()
    (rc0, E0, A0, k0, s0, _), (rc1, E1, A1, k1, s1, info) = out
    assert rc0 == rc1 and k0 == k1
    assert np.array_equal(E0, E1) and np.array_equal(A0, A1)
    _same_state(s0, s1)
    assert info["replays"] >= 7


def test_graph_steps_equal_eager_steps_3d(jv):
    wl = jv.workload_c4(n=16, steps=3)
    conf = wl.conf.with_(epsilon=0.5, S1=jv.stabilisers(0.5, 0.1, 1000, 1000)[0],
                         S3=jv.stabilisers(0.5, 0.1, 1000, 1000)[2])
    phi, psi = jv.initial_fields(conf, wl.vesicles, wl.c_out)
    t = _schedule(conf.dt)
    eager = _fresh(jv, conf, phi, psi, 0, False)
    graph = _fresh(jv, conf, phi, psi, 0, True)
    rc0, E0, A0, _ = eager.run(t, 1)
    rc1, E1, A1, _ = graph.run(t, 1)
    assert rc0 == rc1
    assert np.array_equal(E0, E1) and np.array_equal(A0, A1)
    _same_state(eager.get_state(jv.SAV_PRESENT), graph.get_state(jv.SAV_PRESENT))
    info = graph.graph_info()
    assert info["replays"] == len(t) - 1 and info["repeats"] == 0
    eager.close()
    graph.close()


def test_graph_ensemble_equals_eager_ensemble(jv):
    wls = jv.workload_c5(steps=3, sizes=(64, 96), per_size=2)
    confs = [w.conf for w in wls]
    t = np.arange(6) * confs[0].dt
    res = []
    for graph in (False, True):
        ens = jv.Ensemble(confs, precision=1)
        for m, wl in zip(ens.members, wls):
            phi, psi = wl.initial_fields()
            m.set_state(jv.SAV_PRESENT, device_state(jv, m, wl.conf, phi, psi))
            m.init_sav()
        ens.set_graph(graph)
        rc, E, A, kit, st = ens.run(t, 1, threads=4)
        states = [m.get_state(jv.SAV_PRESENT) for m in ens.members]
        replays = sum(m.graph_info()["replays"] for m in ens.members)
        ens.close()
        res.append((rc, E, kit, states, replays))
    (rc0, E0, k0, s0, r0), (rc1, E1, k1, s1, r1) = res
    assert rc0 == rc1 and np.array_equal(k0, k1) and np.array_equal(E0, E1)
    for a, b in zip(s0, s1):
        _same_state(a, b)
    assert r0 == 0 and r1 >= len(confs) * 3


@pytest.mark.parametrize("case", ["2d", "3d"])
def test_measured_replay_mode(jv, case):
    """set_graph(GRAPH_AUTO): the handle times eager against replayed steps and keeps one;
    the run is bit-identical to an eager run whichever it keeps, and the choice is made
    within the run (graph_info choice 0 or 1)."""
    wl = jv.workload_c1(n=128, steps=1) if case == "2d" else jv.workload_c4(n=32, steps=1)
    conf = wl.conf
    phi, psi = wl.initial_fields()
    t = np.arange(31) * conf.dt
    runs, infos = {}, {}
    for mode in (0, jv.GRAPH_AUTO):
        ops = _fresh(jv, conf, phi, psi, 0, mode)
        rc, E, A, kit = ops.run(t, 1)
        assert rc >= 0
        runs[mode] = (ops.get_state(jv.SAV_PRESENT), E, kit)
        infos[mode] = ops.graph_info()
        ops.close()
    print(case, infos, flush=True)
    assert infos[0]["choice"] == 0 and infos[0]["replays"] == 0
    assert infos[jv.GRAPH_AUTO]["choice"] in (0, 1)
    assert infos[jv.GRAPH_AUTO]["replays"] >= 5
    _same_state(runs[0][0], runs[jv.GRAPH_AUTO][0])
    assert np.array_equal(runs[0][1], runs[jv.GRAPH_AUTO][1])
    assert runs[0][2] == runs[jv.GRAPH_AUTO][2]
