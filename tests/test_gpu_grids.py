"""FP64 parity on grids beyond the square powers of two: rectangular and mixed-radix
(factors 2 and 3) sizes, three vesicles, a non-cubic 3D grid, and FP32 in 3D.  Each
case runs the device and the oracle side by side for a few steps."""
import math

import numpy as np
import pytest

from tests.helpers import relerr, paired_runs

pytestmark = pytest.mark.gpu
FP64_TOL = 1e-10


def _conf(jv, Nx, Ny, N, eps, Nz=None, dt=1e-4):
    S1, S2, S3, S4 = jv.stabilisers(eps, 0.1, 1000, 1000)
    L = 2 * math.pi
    kw = dict(N=N, epsilon=eps, M_phi=1.0, M0_psi=0.5, eta=1.0, gamma_surf=1.0, gamma_area=1000,
              gamma_bend=0.1, gamma_in=1000, beta_in=0, psi_in_v=[1.1 + 0.1 * n for n in range(N)],
              gamma_out=1000, beta_out=0, psi_out_v=[1.0] * N, lamda=1, S1=S1, S2=S2, S3=S3, S4=S4,
              C1=1.0, C2=4e6, C3=8e6, dt=dt, T=dt * 3, Nx=Nx, Ny=Ny, Lx=L, Ly=L, tol=1e-10)
    if Nz is not None:
        kw.update(Nz=Nz, Lz=L)
    return jv.Config.make(**kw)


def _vesicles(jv, N, ndim):
    L = 2 * math.pi
    out = []
    for n in range(N):
        c = [L * (n + 0.5) / N, L / 2] + ([L / 2] if ndim == 3 else [])
        ax = [min(1.1, 0.45 * L / N), 0.8] + ([0.9] if ndim == 3 else [])
        out.append(jv.Vesicle(tuple(c), tuple(ax), 1.1 + 0.1 * n))
    return out


def _check(jv, so, conf, ves, nsteps=3, precision=0, tol=FP64_TOL):
    phi, psi = jv.initial_fields(conf, ves, 1.0)
    worst = 0.0
    for k, pres, old, out, g, st, rc in paired_runs(jv, so, conf, phi, psi, nsteps, precision=precision):
        assert rc >= 0
        for name in ("phi", "psi", "mu", "nu"):
            e = relerr(getattr(g, name).T, getattr(pres, name))
            worst = max(worst, e) if name in ("phi", "psi") else worst
            if precision == 0:
                assert e < tol, (k, name, e)
        if precision == 0 and conf.ndim == 2:
            assert list(st.krylov_iters) == [s.niter for s in out["krylov"]]
    return worst


@pytest.mark.parametrize("Nx,Ny", [(48, 32), (96, 64), (32, 96), (24, 12)])
def test_rectangular_and_mixed_radix_grids(jv, so, Nx, Ny):
    conf = _conf(jv, Nx, Ny, 1, eps=0.3)
    _check(jv, so, conf, _vesicles(jv, 1, 2))


# radix-5 lengths (5 2^a: rows of 40 -> 20-point transforms, columns of 80, 160, 320, 1280)
# and 2048-point rows and columns; a few steps each against the oracle
@pytest.mark.parametrize("Nx,Ny", [(80, 40), (40, 160), (320, 320), (1280, 40), (40, 1280), (2048, 64), (64, 2048)])
def test_radix5_and_2048_grids(jv, so, Nx, Ny):
    conf = _conf(jv, Nx, Ny, 1, eps=0.3)
    _check(jv, so, conf, _vesicles(jv, 1, 2), nsteps=2)


@pytest.mark.parametrize("Nx,Ny", [(80, 40), (2048, 64)])
def test_radix5_and_2048_grids_mixed(jv, so, Nx, Ny):
    """The mixed form's FP32 inner psi iterations on the new lengths: two steps within the
    solve tolerance of the oracle's FP64 steps."""
    conf = _conf(jv, Nx, Ny, 1, eps=0.3)
    phi, psi = jv.initial_fields(conf, _vesicles(jv, 1, 2), 1.0)
    for k, pres, old, out, g, st, rc in paired_runs(jv, so, conf, phi, psi, 2, precision=2):
        assert rc >= 0
        for name in ("phi", "psi"):
            assert relerr(getattr(g, name).T, getattr(pres, name)) < 1e-8, (k, name)


def test_three_vesicles(jv, so):
    conf = _conf(jv, 96, 48, 3, eps=0.2)
    _check(jv, so, conf, _vesicles(jv, 3, 2))


def test_3d_non_cubic_grid(jv, so):
    conf = _conf(jv, 24, 16, 1, eps=0.6, Nz=12)
    _check(jv, so, conf, _vesicles(jv, 1, 3))


def test_3d_fp32_close_to_oracle(jv, so):
    conf = _conf(jv, 32, 32, 2, eps=0.5, Nz=32)
    worst = _check(jv, so, conf, _vesicles(jv, 2, 3), precision=1)
    assert worst < 1e-5, worst
