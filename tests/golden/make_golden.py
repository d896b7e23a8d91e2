"""Writes the golden fixtures of tests/golden/ from the FP64 oracle (oracle/sav_oracle.py).

The reference's own tests hold no trajectories and Julia cannot run here, so these
vectors come from the build's oracle, itself pinned by the reference's KATs and the
closed-form checks of tests/test_oracle_kat.py.  They freeze the oracle's outputs so
a later change to the oracle or to the device path is caught against fixed data:

* single_vesicle_64.npz: the single-vesicle workload (ellipse 1.2 x 0.8, eps = 0.08,
  dt = 1e-4) on 64^2: initial fields, then phi, psi, mu, nu, u, p and the scalars
  (Q, R1, R2, R3, energy, area ratio, psi iterations) after steps 1, 2 and 200.
* transforms_8x6.npz: a random real 8 x 6 field with its rfft, and a random
  non-Hermitian half spectrum with its c2r back-transform (brute-force definition).
* bicgstab_8x8.npz: a psi solve on 8 x 8 with the per-iteration residual norms,
  alpha and omega of the restated Krylov.jl recurrence.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))


def _trajectory(jv, so):
    wl = jv.workload_c1(n=64, steps=200)
    conf = wl.conf
    phi, psi = wl.initial_fields()
    orc = so.Oracle(conf)
    pres, old = orc.init_sav(so.state_from_physical(phi, psi, orc.ops, conf.ndim))
    out = {"phi0": phi, "psi0": psi}
    for n in range(1, 201):
        res = orc.step(pres, old, conf.dt, n)
        if n in (1, 2, 200):
            for k in ("phi", "psi", "mu", "nu", "u", "p"):
                out[f"{k}_{n}"] = getattr(pres, k).copy()
            out[f"scalars_{n}"] = np.array([pres.Q, pres.R1, pres.R2, pres.R3, res["energy"], res["area_ratio"]])
            out[f"iters_{n}"] = np.array([s.niter for s in res["krylov"]], dtype=np.int64)
    np.savez_compressed(HERE / "single_vesicle_64.npz", **out)


def _transforms(so):
    rng = np.random.default_rng(20240601)
    cfg = type("C", (), dict(ndim=2, N=1, Nx=8, Ny=6, Nz=1, Lx=2 * np.pi, Ly=2 * np.pi, Lz=2 * np.pi))()
    ops = so.Ops(cfg)
    x = rng.standard_normal((1, 6, 8))
    X = rng.standard_normal((1, 6, 5)) + 1j * rng.standard_normal((1, 6, 5))
    # c2r by its definition: c2c inverse over y, then a 1-D c2r over x reading only
    # the real parts of the DC and Nyquist bins (FFTW's c2r on non-Hermitian input)
    Nx, Ny = 8, 6
    out = np.zeros((1, Ny, Nx))
    for yy in range(Ny):
        for xx in range(Nx):
            acc = 0.0
            for kx in range(Nx // 2 + 1):
                col = np.sum(X[0, :, kx] * np.exp(2j * np.pi * np.arange(Ny) * yy / Ny))
                w = 1.0 if kx in (0, Nx // 2) else 2.0
                val = col * np.exp(2j * np.pi * kx * xx / Nx)
                acc += w * (val.real if kx not in (0, Nx // 2) else (col.real * np.cos(2 * np.pi * kx * xx / Nx)
                                                                      - col.imag * np.sin(2 * np.pi * kx * xx / Nx)))
            out[0, yy, xx] = acc / (Nx * Ny)
    np.savez_compressed(HERE / "transforms_8x6.npz", x=x, x_hat=ops.rfft(x), X=X, X_c2r=out)


def _bicgstab(jv, so):
    conf = jv.workload_c1(n=8, steps=1).conf.with_(epsilon=0.5)
    rng = np.random.default_rng(7)
    orc = so.Oracle(conf)
    phi = np.tanh(rng.standard_normal((1, 8, 8)))
    M = so.get_M_psi(phi, conf)
    rhs = orc.ops.rfft(rng.standard_normal((1, 8, 8)))
    sol = orc.solve_psi(rhs, M, float(M.max()), 1.5, conf.dt)
    st = orc.krylov_log[-1]
    np.savez_compressed(HERE / "bicgstab_8x8.npz", phi=phi, rhs=rhs, a=np.array(1.5), dt=np.array(conf.dt),
                        epsilon=np.array(conf.epsilon), sol=sol, residuals=np.array(st.residuals),
                        alphas=np.array(st.alphas), omegas=np.array(st.omegas), niter=np.array(st.niter))


def main():
    import __graft_entry__ as ge
    jv = ge.load_package()
    from oracle import sav_oracle as so
    _transforms(so)
    _bicgstab(jv, so)
    _trajectory(jv, so)


if __name__ == "__main__":
    main()
