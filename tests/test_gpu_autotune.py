"""The psi matvec's column-tile autotuner: it picks one of the widths at the first solve,
reports the timings of the widths with distinct kernels, and every width gives
bit-identical solutions (the Krylov dot products are summed per column, whatever the tile)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


# at 256 the 8- and 16-column kernels differ (4 resolves to 8); at 1024 in FP64 the 8- and
# 4-column kernels differ (16 resolves to 8: 160 KB of LDS holds 8 columns of 1024)
@pytest.mark.parametrize("n,precision,timed", [(256, 0, (8, 16)), (256, 1, (8, 16)), (1024, 0, (8, 4))])
def test_column_tile_choice_does_not_change_results(jv, so, n, precision, timed):
    wl = jv.workload_c1(n=n, steps=1)
    conf = wl.conf
    phi, psi = wl.initial_fields()
    orc = so.Oracle(conf)
    rhs = orc.ops.rfft(np.random.default_rng(4).standard_normal(phi.shape))
    ops = jv.Operators(conf, precision=precision)
    assert ops.autotune_info()["columns"] == 0
    auto, info = ops.psi_solve(phi.T, 1.5, conf.dt, rhs.T)
    t = ops.autotune_info()
    print(n, precision, t)
    assert t["columns"] in timed
    for cw in (4, 8, 16):
        assert (t[f"ms_{cw}"] > 0) == (cw in timed), cw
    best = min(timed, key=lambda cw: t[f"ms_{cw}"])
    assert t["columns"] == best
    outs = {}
    for cw in (4, 8, 16):
        ops.set_column_tile(cw)
        outs[cw], inf = ops.psi_solve(phi.T, 1.5, conf.dt, rhs.T)
        assert inf["iters"] == info["iters"]
    ops.close()
    assert np.array_equal(outs[8], outs[16])
    assert np.array_equal(outs[8], outs[4])
    assert np.array_equal(auto, outs[t["columns"]])


@pytest.mark.parametrize("case", ["2d", "3d"])
def test_transform_tile_widths_give_identical_steps(jv, case):
    """The column transforms outside the psi matvec run with 4, 8 or 16 columns per
    workgroup (sav_set_transform_tile): two steps give bit-identical states whatever the
    width, and the autotuner (width 0) picks one of the timed widths."""
    from helpers import device_state
    wl = jv.workload_c1(n=96, steps=2) if case == "2d" else jv.workload_c4(n=16, steps=2)
    conf = wl.conf
    phi, psi = wl.initial_fields()
    states = {}
    for w in (0, 4, 8, 16):
        ops = jv.Operators(conf)
        ops.set_state(jv.SAV_PRESENT, device_state(jv, ops, conf, phi, psi))
        ops.init_sav()
        if w:
            ops.set_transform_tile(w)
        for n in (1, 2):
            rc, st = ops.step(conf.dt, n)
            assert rc == 0
        infos = (ops.transform_tile_info(), ops.transform_tile_fwd_info())
        states[w] = ops.get_state(jv.SAV_PRESENT)
        ops.close()
        for info in infos:  # inverse and forward widths
            if w == 0:
                print(case, info)
                assert info["columns"] in (4, 8, 16)
                timed = [info[f"ms_{c}"] for c in (4, 8, 16)]
                assert max(timed) > 0 and info[f"ms_{info['columns']}"] > 0
                assert info[f"ms_{info['columns']}"] == min(t for t in timed if t > 0)
            else:
                assert info["columns"] == w
    for w in (4, 8, 16):
        for name in ("phi", "psi", "mu", "nu", "u", "p", "phi_hat", "psi_hat"):
            assert np.array_equal(getattr(states[w], name), getattr(states[0], name)), (w, name)


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_matvec_order_gives_identical_steps(jv, precision):
    """The psi matvec's kernels over the three systems of a step's batched solve run either
    all systems per launch (order 1) or system by system over one shared intermediate pair
    (order 2): three steps of two vesicles give bit-identical states and iteration counts in
    both orders, and the autotuner (order 0) times both and keeps the faster."""
    from helpers import device_state
    wl = jv.workload_c3(n=128, steps=3)
    conf = wl.conf
    phi, psi = wl.initial_fields()
    states, iters = {}, {}
    for order in (0, 1, 2):
        ops = jv.Operators(conf, precision=precision)
        ops.set_state(jv.SAV_PRESENT, device_state(jv, ops, conf, phi, psi))
        ops.init_sav()
        if order:
            ops.set_matvec_order(order)
        its = []
        for n in (1, 2, 3):
            rc, st = ops.step(conf.dt, n)
            assert rc == 0
            its.append(tuple(st.krylov_iters))
        info = ops.matvec_order_info()
        states[order] = ops.get_state(jv.SAV_PRESENT)
        iters[order] = its
        ops.close()
        if order == 0:
            print(precision, info)
            assert info["order"] in (1, 2)
            assert info["ms_batched"] > 0 and info["ms_sequenced"] > 0
            best = 1 if info["ms_batched"] <= info["ms_sequenced"] else 2
            assert info["order"] == best
        else:
            assert info["order"] == order
    assert iters[1] == iters[2] == iters[0]
    for o in (0, 2):
        for name in ("phi", "psi", "mu", "nu", "u", "p", "phi_hat", "psi_hat"):
            assert np.array_equal(getattr(states[o], name), getattr(states[1], name)), (o, name)
