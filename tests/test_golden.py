"""The oracle against the committed golden fixtures (tests/golden/, written by
tests/golden/make_golden.py): a change to the oracle that moves any of these
outputs fails here, on the CPU."""
import numpy as np
import pytest

from helpers import relerr

GOLD = __import__("pathlib").Path(__file__).resolve().parent / "golden"


def test_transforms_fixture(so):
    g = np.load(GOLD / "transforms_8x6.npz")

    class F:
        N, Nx, Ny, Lx, Ly, ndim = 1, 8, 6, 2 * np.pi, 2 * np.pi, 2
    o = so.Ops(F)
    assert np.max(np.abs(o.rfft(g["x"]) - g["x_hat"])) < 1e-13
    assert np.max(np.abs(o.irfft(g["X"]) - g["X_c2r"])) < 1e-14


def test_bicgstab_fixture(jv, so):
    g = np.load(GOLD / "bicgstab_8x8.npz")
    conf = jv.workload_c1(n=8, steps=1).conf.with_(epsilon=float(g["epsilon"]))
    orc = so.Oracle(conf)
    M = so.get_M_psi(g["phi"], conf)
    sol = orc.solve_psi(g["rhs"], M, float(M.max()), float(g["a"]), float(g["dt"]))
    st = orc.krylov_log[-1]
    assert st.niter == int(g["niter"])
    assert relerr(sol, g["sol"]) < 1e-12
    assert np.allclose(st.residuals, g["residuals"], rtol=1e-9, atol=1e-18)


@pytest.mark.parametrize("upto", [2])
def test_trajectory_fixture(jv, so, upto):
    g = np.load(GOLD / "single_vesicle_64.npz")
    wl = jv.workload_c1(n=64, steps=upto)
    conf = wl.conf
    phi, psi = wl.initial_fields()
    assert np.array_equal(phi, g["phi0"]) and np.array_equal(psi, g["psi0"])
    orc = so.Oracle(conf)
    pres, old = orc.init_sav(so.state_from_physical(phi, psi, orc.ops, conf.ndim))
    for n in range(1, upto + 1):
        res = orc.step(pres, old, conf.dt, n)
        for k in ("phi", "psi", "mu", "nu"):
            assert relerr(getattr(pres, k), g[f"{k}_{n}"]) < 1e-12, (n, k)
        assert [s.niter for s in res["krylov"]] == list(g[f"iters_{n}"])
        sc = np.array([pres.Q, pres.R1, pres.R2, pres.R3, res["energy"], res["area_ratio"]])
        assert np.allclose(sc, g[f"scalars_{n}"], rtol=1e-11, atol=1e-14)
