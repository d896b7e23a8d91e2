"""The two (x, y) plane kernel families of 3D transforms over 128 x 128 planes (register
columns with a half plane in LDS, and the whole plane in LDS) against numpy's FFT, against
each other over SAV steps, and the plane-kernel autotuner."""
import numpy as np
import pytest

from tests.helpers import relerr
from tests.test_gpu_fft import _conf

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("Nz", [16, 128])
@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("kind", ["register", "lds"])
def test_plane_kernels_vs_numpy(jv, Nz, prec, kind):
    conf = _conf(jv, 128, 128, N=2, Nz=Nz)
    ops = jv.Operators(conf, precision=prec)
    ops.set_plane_kernel(jv.SAV_PLANE_REGISTER if kind == "register" else jv.SAV_PLANE_LDS)
    rng = np.random.default_rng(11 + Nz + prec)
    u = rng.standard_normal((2, Nz, 128, 128))
    ref = np.fft.rfftn(u, axes=(-3, -2, -1))
    tol = 1e-13 if prec == 0 else 2e-6
    print(f"plane kernels {kind} prec {prec} 128x128x{Nz}", flush=True)
    assert relerr(ops.rfft(u.T).T, ref) < tol
    # the DC and Nyquist x columns (packed together by the register kernels) carry energy
    X = rng.standard_normal(ref.shape) + 1j * rng.standard_normal(ref.shape)
    X[..., 0] *= 10.0
    X[..., -1] *= 10.0
    assert relerr(ops.irfft(X.T).T, np.fft.irfftn(X, s=(Nz, 128, 128), axes=(-3, -2, -1))) < tol
    # a real field with a structured spectrum: one mode in each of the DC and Nyquist columns
    z, y, x = np.meshgrid(np.arange(Nz), np.arange(128), np.arange(128), indexing="ij")
    f = np.cos(2 * np.pi * (3 * y) / 128) + np.cos(np.pi * x) * np.sin(2 * np.pi * (5 * y + 2 * z) / 128)
    f = np.stack([f, 2 * f])
    F = ops.rfft(f.T).T
    assert relerr(F, np.fft.rfftn(f, axes=(-3, -2, -1))) < tol
    assert relerr(ops.irfft(F.T).T, f) < tol
    ops.close()


def test_register_plane_kernel_needs_128_planes(jv):
    conf = _conf(jv, 64, 64, N=1, Nz=16)
    ops = jv.Operators(conf)
    with pytest.raises(jv.SavError):
        ops.set_plane_kernel(jv.SAV_PLANE_REGISTER)
    ops.set_plane_kernel(jv.SAV_PLANE_LDS)
    ops.close()


def _run3(jv, conf, phi, psi, prec, kind):
    ops = jv.Operators(conf, precision=prec)
    st = jv.FieldState.zeros(conf)
    st.phi[...] = phi.T
    st.psi[...] = psi.T
    st.phi_hat[...] = ops.rfft_accurate(st.phi)
    st.psi_hat[...] = ops.rfft_accurate(st.psi)
    st.Q = st.R1 = st.R2 = st.R3 = 1.0
    ops.set_plane_kernel(kind)
    ops.set_state(jv.SAV_PRESENT, st)
    ops.init_sav()
    for n in range(1, 4):
        rc, _ = ops.step(conf.dt, n)
        assert rc == 0
    out, info = ops.get_state(jv.SAV_PRESENT), ops.plane_kernel_info()
    ops.close()
    return out, info


@pytest.mark.parametrize("prec", [0, 1])
def test_plane_kernels_agree_over_steps_and_autotune(jv, prec):
    """FP64: both kernel families give the same steps to round-off.  FP32: each is as close
    to the FP64 steps as the other (the FP32 chemical potentials carry the transforms'
    round-off times S1 k^4, so two FP32 transforms differ there by far more than 1e-5)."""
    wl = jv.workload_c4(n=128, steps=3)
    conf = wl.conf
    phi, psi = wl.initial_fields()
    states, infos = {}, {}
    for kind in (jv.SAV_PLANE_REGISTER, jv.SAV_PLANE_LDS, jv.SAV_PLANE_AUTO):
        states[kind], infos[kind] = _run3(jv, conf, phi, psi, prec, kind)
    print("plane kernel info", infos, flush=True)
    a, b = states[jv.SAV_PLANE_REGISTER], states[jv.SAV_PLANE_LDS]
    if prec == 0:
        for name in ("phi", "psi", "mu", "nu"):
            assert relerr(getattr(a, name), getattr(b, name)) < 1e-12, name
        assert relerr(a.u, b.u) < 1e-10
    else:
        ref, _ = _run3(jv, conf, phi, psi, 0, jv.SAV_PLANE_LDS)
        for name in ("phi", "psi", "mu", "nu", "u"):
            ea = relerr(getattr(a, name), getattr(ref, name))
            eb = relerr(getattr(b, name), getattr(ref, name))
            print(f"fp32 vs fp64 {name}: register {ea:.2e} lds {eb:.2e}", flush=True)
            assert ea < 2.0 * eb + 1e-6, name
        for name in ("phi", "psi"):
            assert relerr(getattr(a, name), getattr(ref, name)) < 1e-5, name
    auto = infos[jv.SAV_PLANE_AUTO]
    assert auto["kind"] in (jv.SAV_PLANE_REGISTER, jv.SAV_PLANE_LDS)
    assert auto["ms_register"] > 0 and auto["ms_lds"] > 0
    best = jv.SAV_PLANE_LDS if auto["ms_lds"] < 0.97 * auto["ms_register"] else jv.SAV_PLANE_REGISTER
    assert auto["kind"] == best
    assert infos[jv.SAV_PLANE_REGISTER]["kind"] == jv.SAV_PLANE_REGISTER
