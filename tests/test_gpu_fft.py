"""Device transforms against numpy's FFT and the closed-form c2r definition."""
import numpy as np
import pytest

from tests.helpers import relerr

pytestmark = pytest.mark.gpu


def _conf(jv, Nx, Ny, N=1, Nz=None):
    wl = jv.workload_c1(n=8, steps=1)
    kw = dict(Nx=Nx, Ny=Ny, Lx=2 * np.pi, Ly=2 * np.pi, N=N, psi_in_v=[1.2] * N, psi_out_v=[1.0] * N)
    c = wl.conf
    if Nz is not None:
        return jv.Config.make(N=N, epsilon=c.epsilon, M_phi=1, M0_psi=0.5, eta=1, gamma_surf=1, gamma_area=1,
                              gamma_bend=0.1, gamma_in=1, beta_in=0, psi_in_v=[1.2] * N, gamma_out=1, beta_out=0,
                              psi_out_v=[1.0] * N, lamda=1, S1=1, S2=0, S3=1, S4=1, C1=1, C2=1, C3=1, dt=1e-3, T=1,
                              Nx=Nx, Ny=Ny, Lx=2 * np.pi, Ly=2 * np.pi, tol=1e-10, Nz=Nz, Lz=2 * np.pi)
    return jv.Config.make(N=N, epsilon=c.epsilon, M_phi=1, M0_psi=0.5, eta=1, gamma_surf=1, gamma_area=1,
                          gamma_bend=0.1, gamma_in=1, beta_in=0, psi_in_v=[1.2] * N, gamma_out=1, beta_out=0,
                          psi_out_v=[1.0] * N, lamda=1, S1=1, S2=0, S3=1, S4=1, C1=1, C2=1, C3=1, dt=1e-3, T=1,
                          Nx=Nx, Ny=Ny, Lx=2 * np.pi, Ly=2 * np.pi, tol=1e-10)


SIZES_2D = [(8, 6), (16, 12), (128, 64), (64, 128), (96, 48), (256, 256), (384, 384), (512, 512), (1024, 1024),
            # radix 5 (lengths 5 2^a) and 2048
            (10, 20), (40, 80), (160, 320), (640, 40), (1280, 160), (2048, 64), (64, 2048)]


@pytest.mark.parametrize("Nx,Ny", SIZES_2D)
@pytest.mark.parametrize("prec", [0, 1])
def test_rfft_irfft_vs_numpy(jv, Nx, Ny, prec):
    conf = _conf(jv, Nx, Ny, N=2)
    ops = jv.Operators(conf, precision=prec)
    rng = np.random.default_rng(Nx * 7 + Ny)
    u = rng.standard_normal((2, Ny, Nx))
    ref = np.fft.rfftn(u, axes=(-2, -1))
    got = ops.rfft(u.T).T
    tol = 1e-13 if prec == 0 else 2e-6
    assert relerr(got, ref) < tol
    # non-Hermitian half spectra: irfft follows the c2r definition (numpy's irfftn)
    X = rng.standard_normal(ref.shape) + 1j * rng.standard_normal(ref.shape)
    back = ops.irfft(X.T).T
    assert relerr(back, np.fft.irfftn(X, s=(Ny, Nx), axes=(-2, -1))) < (1e-13 if prec == 0 else 2e-6)
    ops.close()


@pytest.mark.parametrize("Nx,Ny", [(8, 6), (128, 64), (1024, 1024), (80, 40), (2048, 32), (40, 2048), (1280, 20)])
def test_fft_full_vs_numpy(jv, Nx, Ny):
    conf = _conf(jv, Nx, Ny, N=2)
    ops = jv.Operators(conf)
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, Ny, Nx)) + 1j * rng.standard_normal((2, Ny, Nx))
    assert relerr(ops.fft_full(x.T).T, np.fft.fftn(x, axes=(-2, -1))) < 1e-13
    assert relerr(ops.fft_full(x.T, inverse=True).T, np.fft.ifftn(x, axes=(-2, -1))) < 1e-13
    ops.close()


# Square (Nx == Ny) planes of up to 128 run the fused (x, y) plane kernels; the others
# (8, and Nx != Ny) the separate row and column passes.
SIZES_3D = [(8, 8, 8), (16, 16, 16), (32, 32, 32), (64, 64, 16), (96, 96, 24), (128, 128, 128),
            (24, 16, 12), (32, 64, 16), (20, 40, 10), (80, 80, 20)]


@pytest.mark.parametrize("Nx,Ny,Nz", SIZES_3D)
@pytest.mark.parametrize("prec", [0, 1])
def test_rfft_3d_vs_numpy(jv, Nx, Ny, Nz, prec):
    conf = _conf(jv, Nx, Ny, N=2, Nz=Nz)
    ops = jv.Operators(conf, precision=prec)
    rng = np.random.default_rng(Nx + 3 * Ny + 7 * Nz)
    u = rng.standard_normal((2, Nz, Ny, Nx))
    ref = np.fft.rfftn(u, axes=(-3, -2, -1))
    tol = 1e-13 if prec == 0 else 2e-6
    assert relerr(ops.rfft(u.T).T, ref) < tol
    # non-Hermitian half spectra: irfft follows the c2r definition (numpy's irfftn)
    X = rng.standard_normal(ref.shape) + 1j * rng.standard_normal(ref.shape)
    assert relerr(ops.irfft(X.T).T, np.fft.irfftn(X, s=(Nz, Ny, Nx), axes=(-3, -2, -1))) < tol
    ops.close()
