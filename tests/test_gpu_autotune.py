"""The psi matvec's column-tile autotuner: it picks one of the two widths at the first
solve, reports both timings, and both widths give bit-identical solutions (the Krylov dot
products are summed per column, whatever the tile)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [0, 1])
def test_column_tile_choice_does_not_change_results(jv, so, precision):
    wl = jv.workload_c1(n=256, steps=1)   # at 256 the two widths give different kernels
    conf = wl.conf
    phi, psi = wl.initial_fields()
    orc = so.Oracle(conf)
    rhs = orc.ops.rfft(np.random.default_rng(4).standard_normal(phi.shape))
    ops = jv.Operators(conf, precision=precision)
    assert ops.autotune_info()["columns"] == 0
    auto, info = ops.psi_solve(phi.T, 1.5, conf.dt, rhs.T)
    t = ops.autotune_info()
    print(t)
    assert t["columns"] in (8, 16) and t["ms_8"] > 0 and t["ms_16"] > 0
    outs = {}
    for cw in (8, 16):
        ops.set_column_tile(cw)
        outs[cw], inf = ops.psi_solve(phi.T, 1.5, conf.dt, rhs.T)
        assert inf["iters"] == info["iters"]
    ops.close()
    assert np.array_equal(outs[8], outs[16])
    assert np.array_equal(auto, outs[t["columns"]])
