"""Plots of scripts/visualize.jl (host side, matplotlib): files are written, the isosurface
of plot_iso is geometrically right (a sphere's level set has area 4 pi R^2 and lies at
radius R), and plot_vector refuses mismatched series."""
import math

import numpy as np
import pytest


def _conf2d(jv, n=64):
    return jv.workload_c1(n=n, steps=1).conf


def test_isosurface_of_a_sphere(jv):
    n, L, R = 48, 2.0, 0.6
    h = L / n
    x = np.arange(n) * h - L / 2 + h / 2
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    f = R - np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    tri = jv.plots.isosurface(f, 0.0, (h, h, h))
    area = 0.5 * np.linalg.norm(np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]), axis=1).sum()
    assert area == pytest.approx(4 * math.pi * R ** 2, rel=2e-2)
    centre = (n - 1) * h / 2
    r = np.linalg.norm(tri.reshape(-1, 3) - centre, axis=1)
    assert np.max(np.abs(r - R)) < 0.5 * h
    # a level outside the range gives no triangles
    assert len(jv.plots.isosurface(f, 10.0)) == 0


def test_field_images_and_histories(jv, tmp_path):
    conf = _conf2d(jv)
    wl = jv.workload_c1(n=64, steps=1)
    phi, psi = wl.initial_fields()
    st = jv.FieldState.zeros(conf)
    st.phi[...] = phi.T
    st.psi[...] = psi.T
    st.u[..., 0] = np.sin(np.arange(conf.Nx))[:, None]
    plots = jv.plots
    plots.plot_field(st.phi[..., 0], conf, title="phi", filename=str(tmp_path / "f.png"))
    plots.plot_all_fields(st, conf, 1e-3, filename=str(tmp_path / "all.png"))
    plots.plot_velocity(st.u[..., 0], st.u[..., 1], conf, filename=str(tmp_path / "v.png"))
    t = np.linspace(0, 1e-3, 50)
    plots.plot_vector([np.exp(-t), np.exp(-2 * t)], t, labels=["a", "b"], phase_switch=1e-5,
                      filename=str(tmp_path / "e.png"))
    with pytest.raises(ValueError):
        plots.plot_vector(np.ones(3), t)
    files = plots.visualize(st, conf, list(np.exp(-t[1:])), list(t[1:] * 0), t, 1e-3, str(tmp_path))
    for fn in ["f.png", "all.png", "v.png", "e.png", "phi_final.png", "energy_history.png", "area_history.png"]:
        p = tmp_path / fn
        assert p.exists() and p.stat().st_size > 1000, fn
        assert p.read_bytes()[:8] == b"\x89PNG\r\n\x1a\n"
    assert len(files) == 3
    # phi summed over vesicles + N - 1 (simulation.jl:131): -1 outside every vesicle
    two = np.stack([st.phi[..., 0], -np.ones_like(st.phi[..., 0])], axis=-1)
    np.testing.assert_allclose(plots.phi_sum(two, 2), st.phi[..., 0], rtol=0, atol=1e-15)


def test_plot_iso_writes_a_surface(jv, tmp_path):
    conf = jv.workload_c4(n=16, steps=1).conf
    phi, _ = jv.initial_fields(conf, jv.workload_c4(n=16, steps=1).vesicles, 1.0)
    jv.plots.plot_iso(phi[0].T, conf, filename=str(tmp_path / "iso.png"), title="phi = 0")
    assert (tmp_path / "iso.png").stat().st_size > 1000
    with pytest.warns(UserWarning):
        jv.plots.plot_iso(phi[0].T, conf, isovalue=5.0)
