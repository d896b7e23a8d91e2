"""Plots of scripts/visualize.jl and the image outputs of run_simulation (simulation.jl:130-135)
and of Main.jl's _visualize (scripts/Main.jl:119-170), drawn with matplotlib (Agg, files only).

Host-side only: the fields come from FieldState (Julia shapes, x first), as the MAT
snapshots hold them.  The 3D isosurface of plot_iso is extracted by this module's own
marching-tetrahedra pass (the reference uses Meshing.jl's MarchingTetrahedra)."""
from __future__ import annotations

from typing import Optional, Sequence, Union

import numpy as np


def _plt():
    import matplotlib
    matplotlib.use("Agg", force=True)
    import matplotlib.pyplot as plt
    return plt


def _cmap(name):
    """Colormap by the reference's name; `berlin` and `coolwarm` exist in matplotlib, `bwr` too."""
    import matplotlib
    try:
        return matplotlib.colormaps[name]
    except KeyError:
        return matplotlib.colormaps["RdBu_r"]


def plot_field(field: np.ndarray, conf, title: Optional[str] = None, colormap: str = "viridis",
               filename: Optional[str] = None):
    """plot_field (visualize.jl:2-27): heatmap of an (Nx, Ny) field over [0, Lx] x [0, Ly],
    colour range (-0.5, 0.5), interpolated, no axes decorations."""
    plt = _plt()
    fig, ax = plt.subplots(figsize=(6, 5))
    ax.imshow(np.asarray(field).T, origin="lower", extent=(0.0, conf.Lx, 0.0, conf.Ly), cmap=_cmap(colormap),
              vmin=-0.5, vmax=0.5, interpolation="bilinear", aspect="equal")
    if title is not None:
        ax.set_title(title)
    ax.set_axis_off()
    if filename is not None:
        fig.savefig(filename, dpi=100)
    return fig


def _rowcol(i):
    """fld2rowcol (visualize.jl:63): panel i (1-based) of a 2 x 3 grid."""
    return (i - 1) // 3, (i - 1) % 3


def plot_all_fields(state, conf, t: float, filename: Optional[str] = None, vesicle: int = 0):
    """plot_all_fields (visualize.jl:30-61): phi, psi, mu, nu, p and u_x with colour bars.
    The reference reads `state.ux`, which FieldState does not have; u[:, :, 0] is that
    component.  Multi-vesicle fields show vesicle `vesicle`."""
    plt = _plt()
    specs = [(state.phi[..., vesicle], "berlin", "φ (Phase 1)", (-1.0, 1.0)),
             (state.psi[..., vesicle], "berlin", "ψ (Phase 2)", (-1.0, 1.0)),
             (state.mu[..., vesicle], "viridis", "μ (Chem. Pot)", None),
             (state.nu[..., vesicle], "viridis", "ν (Chem. Pot)", None),
             (state.p, "bwr", "p (Pressure)", None),
             (state.u[..., 0], "coolwarm", "u_x (Velocity)", None)]
    fig, axes = plt.subplots(2, 3, figsize=(10, 7))
    fig.suptitle(f"t = {round(t, 4)}")
    for i, (fld, cmap, ttl, clim) in enumerate(specs, start=1):
        r, c = _rowcol(i)
        ax = axes[r, c]
        f = np.asarray(fld)
        lo, hi = clim if clim is not None else (float(f.min()), float(f.max()))
        im = ax.imshow(f.T, origin="lower", extent=(0.0, conf.Lx, 0.0, conf.Ly), cmap=_cmap(cmap), vmin=lo, vmax=hi,
                       aspect="equal")
        ax.set_title(ttl)
        fig.colorbar(im, ax=ax, fraction=0.046)
    if filename is not None:
        fig.savefig(filename, dpi=100)
    return fig


def plot_velocity(ux: np.ndarray, uy: np.ndarray, conf, stride: int = 8, filename: Optional[str] = None):
    """plot_velocity (visualize.jl:66-88): speed in grey, arrows on every stride-th point."""
    plt = _plt()
    ux, uy = np.asarray(ux), np.asarray(uy)
    speed = np.sqrt(ux ** 2 + uy ** 2)
    fig, ax = plt.subplots(figsize=(8, 8))
    ax.imshow(speed.T, origin="lower", extent=(0.0, conf.Lx, 0.0, conf.Ly), cmap="gray", alpha=0.5, aspect="equal")
    x = np.linspace(0.0, conf.Lx, conf.Nx)[::stride]
    y = np.linspace(0.0, conf.Ly, conf.Ny)[::stride]
    X, Y = np.meshgrid(x, y, indexing="ij")
    ax.quiver(X, Y, ux[::stride, ::stride], uy[::stride, ::stride], angles="xy")
    ax.set_title("Velocity Field")
    if filename is not None:
        fig.savefig(filename, dpi=100)
    return fig


def plot_vector(ys: Union[Sequence[float], Sequence[Sequence[float]]], ts: Sequence[float], ylabel: str = "Value",
                title: Optional[str] = None, filename: Optional[str] = None, phase_switch: Optional[float] = None,
                labels: Optional[Sequence[str]] = None):
    """plot_vector (visualize.jl:91-151): one or several time series against ts, a legend
    when labels are given, a dashed line at phase_switch when it lies inside the range."""
    plt = _plt()
    arr = ys
    curves = [np.asarray(arr, dtype=float)] if np.ndim(arr) == 1 else [np.asarray(c, dtype=float) for c in arr]
    t = np.asarray(ts, dtype=float)
    fig, ax = plt.subplots(figsize=(7.5, 4.8))
    for i, y in enumerate(curves):
        if len(y) != len(t):
            raise ValueError(f"curve {i} has {len(y)} values for {len(t)} times")
        lbl = labels[i] if labels is not None and i < len(labels) else None
        ax.plot(t, y, linewidth=2, label=lbl)
    ax.set_xlabel("t")
    ax.set_ylabel(ylabel)
    if title is not None:
        ax.set_title(title)
    if labels is not None:
        ax.legend(loc="upper right")
    if phase_switch is not None and t[0] < phase_switch < t[-1]:
        ax.axvline(phase_switch, color="gray", linestyle="--", linewidth=1)
    if filename is not None:
        fig.savefig(filename, dpi=100)
    return fig


# Kuhn decomposition of the unit cube into six tetrahedra along the 0-7 diagonal; corner
# k of a cube is (k & 1, (k >> 1) & 1, (k >> 2) & 1).
_TETS = np.array([[0, 1, 3, 7], [0, 3, 2, 7], [0, 2, 6, 7], [0, 6, 4, 7], [0, 4, 5, 7], [0, 5, 1, 7]])
_CORNERS = np.array([[(k >> 0) & 1, (k >> 1) & 1, (k >> 2) & 1] for k in range(8)], dtype=float)


def isosurface(field: np.ndarray, isovalue: float = 0.0, spacing=(1.0, 1.0, 1.0)):
    """Triangles (T, 3, 3) of the level set field = isovalue by marching tetrahedra on the
    grid of `field` (node (i, j, k) at (i dx, j dy, k dz)); every tetrahedron the level set
    crosses contributes one triangle (one vertex on the other side) or two (two and two),
    with vertices interpolated linearly along its edges."""
    f = np.asarray(field, dtype=float) - isovalue
    nx, ny, nz = f.shape
    I, J, K = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1), np.arange(nz - 1), indexing="ij")
    base = np.stack([I.ravel(), J.ravel(), K.ravel()], axis=1)                 # (C, 3)
    corner_idx = base[:, None, :] + _CORNERS[None, :, :].astype(int)           # (C, 8, 3)
    vals = f[corner_idx[..., 0], corner_idx[..., 1], corner_idx[..., 2]]        # (C, 8)
    pos = corner_idx * np.asarray(spacing, dtype=float)                        # (C, 8, 3)
    tv = vals[:, _TETS].reshape(-1, 4)                                         # (C*6, 4)
    tp = pos[:, _TETS].reshape(-1, 4, 3)
    inside = tv < 0
    n_in = inside.sum(axis=1)
    tris = []

    def edge(pa, va, pb, vb):
        w = va / (va - vb)
        return pa + w[:, None] * (pb - pa)

    for n_one, flip in ((1, False), (3, True)):
        sel = n_in == n_one
        if not np.any(sel):
            continue
        ins = inside[sel] if not flip else ~inside[sel]
        p, v = tp[sel], tv[sel]
        lone = np.argmax(ins, axis=1)                       # the vertex alone on its side
        others = np.array([[b for b in range(4) if b != a] for a in range(4)])[lone]
        r = np.arange(len(lone))
        pa, va = p[r, lone], v[r, lone]
        pts = [edge(pa, va, p[r, others[:, q]], v[r, others[:, q]]) for q in range(3)]
        tris.append(np.stack(pts, axis=1))
    sel = n_in == 2
    if np.any(sel):
        ins, p, v = inside[sel], tp[sel], tv[sel]
        order = np.argsort(~ins, axis=1, kind="stable")     # the two inside vertices first
        r = np.arange(len(order))
        a, b, c, d = (order[:, q] for q in range(4))
        e_ac = edge(p[r, a], v[r, a], p[r, c], v[r, c])
        e_ad = edge(p[r, a], v[r, a], p[r, d], v[r, d])
        e_bc = edge(p[r, b], v[r, b], p[r, c], v[r, c])
        e_bd = edge(p[r, b], v[r, b], p[r, d], v[r, d])
        tris.append(np.stack([e_ac, e_ad, e_bd], axis=1))
        tris.append(np.stack([e_ac, e_bd, e_bc], axis=1))
    return np.concatenate(tris, axis=0) if tris else np.zeros((0, 3, 3))


def plot_iso(field: np.ndarray, conf, isovalue: float = 0.0, color="steelblue", alpha: float = 0.85,
             title: Optional[str] = None, filename: Optional[str] = None):
    """plot_iso (visualize.jl:155-218): the closed level surface of a 3D field in physical
    coordinates, shaded, with a faint wireframe when translucent.  Warns and returns an
    empty figure when the isovalue lies outside the field's range."""
    import warnings
    plt = _plt()
    from mpl_toolkits.mplot3d.art3d import Poly3DCollection
    f = np.asarray(field, dtype=float)
    tri = isosurface(f, isovalue, (conf.Lx / conf.Nx, conf.Ly / conf.Ny, conf.Lz / conf.Nz))
    fig = plt.figure(figsize=(7, 6.5))
    if len(tri) == 0:
        warnings.warn(f"plot_iso: empty isosurface, isovalue={isovalue} outside {float(f.min())}..{float(f.max())}")
        return fig
    ax = fig.add_subplot(projection="3d")
    coll = Poly3DCollection(tri, facecolor=color, alpha=alpha,
                            edgecolor=(0, 0, 0, 0.06) if alpha < 0.95 else None, linewidth=0.3)
    ax.add_collection3d(coll)
    ax.set_xlim(0, conf.Lx)
    ax.set_ylim(0, conf.Ly)
    ax.set_zlim(0, conf.Lz)
    ax.set_xlabel("x")
    ax.set_ylabel("y")
    ax.set_zlabel("z")
    ax.view_init(elev=36, azim=229.5)
    if title is not None:
        ax.set_title(title)
    if filename is not None:
        fig.savefig(filename, dpi=100)
    return fig


def phi_sum(phi: np.ndarray, N: int) -> np.ndarray:
    """sum over the vesicles + N - 1 (simulation.jl:131, Main.jl:123): one field whose
    interfaces are those of every vesicle."""
    return np.sum(phi, axis=-1) + N - 1


def visualize(state, conf, energy_history, area_history, Dt, T: float, save_path: str):
    """_visualize (scripts/Main.jl:119-170) without its MAT write (run_simulation writes
    the snapshots): phi_final.png (2D), energy_history.png and area_history.png."""
    import os
    files = []
    if conf.ndim == 2:
        fn = os.path.join(save_path, "phi_final.png")
        plot_field(phi_sum(state.phi, conf.N), conf, title=f"phi  (t = {T:.2e})", filename=fn)
        files.append(fn)
    if len(energy_history) > 1:
        fn = os.path.join(save_path, "energy_history.png")
        plot_vector(energy_history, Dt[1: len(energy_history) + 1], ylabel="Modified energy",
                    title="Modified energy vs. time", filename=fn)
        files.append(fn)
    if len(area_history) > 1:
        fn = os.path.join(save_path, "area_history.png")
        plot_vector(area_history, Dt[1: len(area_history) + 1], ylabel="Surface Ratio",
                    title="Surface Ratio vs. time", filename=fn)
        files.append(fn)
    _plt().close("all")
    return files
