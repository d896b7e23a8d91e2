"""The device path (FP64) against the committed golden fixtures: transforms, a psi
solve with its iteration count, and 200 SAV steps of the single-vesicle workload on
64^2 (fields within 1e-10 after 2 steps and 1e-8 after 200, identical psi iterations)."""
import numpy as np
import pytest

from helpers import relerr, device_state

pytestmark = pytest.mark.gpu
GOLD = __import__("pathlib").Path(__file__).resolve().parent / "golden"


def test_transforms_match_fixture(jv):
    g = np.load(GOLD / "transforms_8x6.npz")
    conf = jv.workload_c1(n=8, steps=1).conf.with_(Ny=6, Nx=8)
    ops = jv.Operators(conf)
    got = ops.rfft(g["x"].T)
    back = ops.irfft(np.asfortranarray(g["X"].T))
    ops.close()
    assert np.max(np.abs(got.T - g["x_hat"])) < 1e-13
    assert np.max(np.abs(back.T - g["X_c2r"])) < 1e-14


def test_psi_solve_matches_fixture(jv):
    g = np.load(GOLD / "bicgstab_8x8.npz")
    conf = jv.workload_c1(n=8, steps=1).conf.with_(epsilon=float(g["epsilon"]))
    ops = jv.Operators(conf)
    got, info = ops.psi_solve(g["phi"].T, float(g["a"]), float(g["dt"]), g["rhs"].T)
    ops.close()
    assert info["iters"] == int(g["niter"])
    assert relerr(got.T, g["sol"]) < 1e-11


def test_200_steps_match_fixture(jv):
    g = np.load(GOLD / "single_vesicle_64.npz")
    wl = jv.workload_c1(n=64, steps=200)
    conf = wl.conf
    phi, psi = g["phi0"], g["psi0"]
    ops = jv.Operators(conf)
    ops.set_state(jv.SAV_PRESENT, device_state(jv, ops, conf, phi, psi))
    ops.init_sav()
    for n in range(1, 201):
        rc, st = ops.step(conf.dt, n)
        assert rc == 0
        if n in (1, 2, 200):
            s = ops.get_state(jv.SAV_PRESENT)
            tol = 1e-10 if n <= 2 else 1e-8
            for k in ("phi", "psi", "mu", "nu"):
                e = relerr(getattr(s, k).T, g[f"{k}_{n}"])
                assert e < tol, (n, k, e)
            assert list(st.krylov_iters) == list(g[f"iters_{n}"]), n
            ref = g[f"scalars_{n}"]
            assert abs(st.Q - ref[0]) <= tol * abs(ref[0])
            assert abs(st.energy - ref[4]) <= tol * max(abs(ref[4]), 1.0)
    ops.close()
