"""Snapshot schedule and files, on the CPU: the save-index rule of simulation.jl:113-114
checked the way test/test_save_interval.jl checks it, the MAT keys of
test/verify_save_state.jl, and the 7.3 (HDF5) MAT format that MAT.jl's matwrite writes."""
import json
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest


def test_save_schedule_hits_every_interval(jv):
    # test_save_interval.jl: Dt = [0, 1e-7, 1e-5, 2e-5, ...] up to 0.015, jittered by 1e-18
    Dt = [0.0, 1e-7]
    t = 1e-5
    while t <= 0.015 + 1e-7:
        Dt.append(t)
        t += 1e-5
    rng = np.random.default_rng(1)
    Dt = np.array(Dt) + (rng.random(len(Dt)) - 0.5) * 1e-18
    # the reference test walks n over all nodes (t = Dt[n]); save_steps' n-th step starts at Dt[n-1]
    steps = jv.save_steps(np.append(Dt, Dt[-1] + 1e-5), 0.001)
    saved = [Dt[n - 1] for n in steps]
    for e in np.arange(0.0, 0.015 + 1e-9, 0.001):
        assert any(abs(s - e) < 1e-10 for s in saved), e
    assert len(saved) == 16


def test_save_schedule_first_step_and_names(jv):
    Dt = jv.time_nodes(1e-4, 3e-3)
    steps = jv.save_steps(Dt, 0.001)
    assert steps[0] == 1                      # curr index 0 > -1 at t = 0
    conf = jv.workload_c1(n=32, steps=1).conf
    assert jv.snapshot_name(conf, 1e-3) == "state_2d_t1.00e-03.mat"


def _dummy_state(jv, conf, seed=2):
    st = jv.FieldState.zeros(conf)
    rng = np.random.default_rng(seed)
    for k in ("phi", "psi", "mu", "nu", "p", "u"):
        getattr(st, k)[...] = rng.random(getattr(st, k).shape)
    st.R1, st.R2, st.R3, st.Q = 1.1, 2.2, 3.3, 4.4
    return st


@pytest.mark.parametrize("version", ["7.3", "5"])
def test_snapshot_keys_and_values(jv, tmp_path, version):
    # verify_save_state.jl: the ten keys, R1 and Q, and phi of size (Nx, Ny, N)
    conf = jv.workload_c1(n=16, steps=1).conf
    st = _dummy_state(jv, conf)
    path = tmp_path / "state_2d_final.mat"
    jv.write_snapshot(path, st, version=version)
    d = jv.read_snapshot(path)
    assert set(d) == set(jv.SNAPSHOT_KEYS)
    assert d["R1"] == 1.1 and d["Q"] == 4.4
    assert d["phi"].shape == (16, 16, 1)
    for k in ("phi", "psi", "mu", "nu", "p", "u"):
        assert np.array_equal(d[k], getattr(st, k)), k


def test_mat73_header(jv, tmp_path):
    """The 128-byte MATLAB header of a 7.3 file: scipy's format probe reads version (2, 0)
    from it (and loadmat then asks for an HDF5 reader), and the HDF5 signature sits at the
    end of the 512-byte user block."""
    import scipy.io
    import scipy.io.matlab as sm
    conf = jv.workload_c1(n=16, steps=1).conf
    path = tmp_path / "s.mat"
    jv.write_snapshot(path, _dummy_state(jv, conf))
    raw = path.read_bytes()
    assert raw.startswith(b"MATLAB 7.3 MAT-file") and raw[512:520] == b"\x89HDF\r\n\x1a\n"
    assert sm.matfile_version(str(path)) == (2, 0)
    with pytest.raises(NotImplementedError):
        scipy.io.loadmat(str(path))


_H5PY = Path("/opt/conda/bin/python3.9")

_H5_READ = r"""
import json, sys
import h5py, numpy as np
f = h5py.File(sys.argv[1], "r")
ref = np.load(sys.argv[2])
out = {"userblock": f.userblock_size, "keys": sorted(f.keys()), "vars": {}}
for k in f:
    d = f[k]
    a = d[()]
    r = ref[k]
    r = r if r.ndim >= 2 else r.reshape(-1, 1)
    out["vars"][k] = {"shape": list(d.shape), "dtype": str(d.dtype),
                      "cls": d.attrs["MATLAB_class"].decode(), "equal": bool(np.array_equal(a.T, r))}
print(json.dumps(out))
"""


def _h5py_available():
    if not _H5PY.exists():
        return False
    env = {k: v for k, v in os.environ.items() if k != "PYTHONPATH"}
    env["PYTHONHOME"] = "/opt/conda"
    r = subprocess.run([str(_H5PY), "-c", "import h5py"], env=env, capture_output=True)
    return r.returncode == 0


_H5_WRITE = r"""
import sys
import h5py, numpy as np
ref = np.load(sys.argv[2])
with h5py.File(sys.argv[1], "w", userblock_size=512, libver="earliest") as f:
    for k in ref.files:
        a = ref[k]
        a = a if a.ndim >= 2 else a.reshape(-1, 1)
        d = f.create_dataset(k, data=np.ascontiguousarray(a.T))
        d.attrs["MATLAB_class"] = np.bytes_("double")
"""


@pytest.mark.skipif(not _h5py_available(), reason="no HDF5 library (h5py) in this image to write the file")
def test_mat73_reader_reads_hdf5_library_files(jv, tmp_path):
    """A 7.3 file written by the HDF5 library (h5py, 512-byte user block, MATLAB_class
    attributes added after each dataset, as MAT.jl lays out its files), with the MATLAB
    header put into the user block: read_snapshot returns every variable in its shape."""
    from importlib import import_module
    matfile = import_module(jv.__name__ + ".matfile")
    conf = jv.workload_c1(n=12, steps=1).conf.with_(Ny=10)
    st = _dummy_state(jv, conf, seed=7)
    ref = tmp_path / "ref.npz"
    np.savez(ref, **{k: np.asarray(getattr(st, k)) for k in jv.SNAPSHOT_KEYS})
    path = tmp_path / "h.mat"
    env = {k: v for k, v in os.environ.items() if k != "PYTHONPATH"}
    env["PYTHONHOME"] = "/opt/conda"
    r = subprocess.run([str(_H5PY), "-c", _H5_WRITE, str(path), str(ref)], env=env, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    with open(path, "r+b") as f:
        f.write(matfile.matlab_header())
    d = jv.read_snapshot(path)
    assert d["R1"] == 1.1 and d["Q"] == 4.4
    for k in ("phi", "psi", "mu", "nu", "p", "u"):
        assert np.array_equal(d[k], getattr(st, k)), k


@pytest.mark.skipif(not _h5py_available(), reason="no independent HDF5 reader (h5py) in this image")
@pytest.mark.parametrize("case", ["2d", "3d"])
def test_mat73_read_by_hdf5_library(jv, tmp_path, case):
    """The HDF5 library itself (h5py 3 over HDF5 1.10, a separate interpreter of the image)
    reads every variable of a snapshot back: a 512-byte user block, the ten keys, float64
    datasets of the reversed (row-major) shape with the column-major data, scalars as 1 x 1,
    and MATLAB_class "double" on each, as MAT.jl and MATLAB write them."""
    conf = jv.workload_c1(n=12, steps=1).conf.with_(Ny=10) if case == "2d" else \
        jv.workload_c4(n=8, steps=1).conf
    st = _dummy_state(jv, conf, seed=5)
    path = tmp_path / "s.mat"
    jv.write_snapshot(path, st)
    ref = tmp_path / "ref.npz"
    np.savez(ref, **{k: np.asarray(getattr(st, k)) for k in jv.SNAPSHOT_KEYS})
    env = {k: v for k, v in os.environ.items() if k != "PYTHONPATH"}
    env["PYTHONHOME"] = "/opt/conda"
    r = subprocess.run([str(_H5PY), "-c", _H5_READ, str(path), str(ref)], env=env, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)
    assert out["userblock"] == 512
    assert out["keys"] == sorted(jv.SNAPSHOT_KEYS)
    for k, v in out["vars"].items():
        shape = np.asarray(getattr(st, k)).shape
        assert v["shape"] == (list(shape[::-1]) if len(shape) >= 2 else [1, 1]), k
        assert v["dtype"] == "float64" and v["cls"] == "double" and v["equal"], (k, v)


def test_checkpoint_roundtrip(jv, tmp_path):
    conf = jv.workload_c1(n=16, steps=1).conf
    a, b = jv.FieldState.zeros(conf), jv.FieldState.zeros(conf)
    rng = np.random.default_rng(3)
    for st in (a, b):
        for k in jv.FieldState.ARRAYS:
            arr = getattr(st, k)
            arr[...] = rng.random(arr.shape) + (1j * rng.random(arr.shape) if np.iscomplexobj(arr) else 0)
        st.Q, st.R1, st.R2, st.R3, st.dt = rng.random(5)
        st.A0 = rng.random(1)
    jv.write_checkpoint(tmp_path / "c.npz", a, b, 7, 0.25)
    p, o, n, t = jv.read_checkpoint(tmp_path / "c.npz")
    assert n == 7 and t == 0.25
    for x, y in ((a, p), (b, o)):
        for k in jv.FieldState.ARRAYS:
            assert np.array_equal(getattr(x, k), getattr(y, k))
        assert (x.Q, x.R1, x.R2, x.R3, x.dt) == (y.Q, y.R1, y.R2, y.R3, y.dt)


def test_initial_fields_3d(jv):
    base = jv.workload_c4(n=32, steps=1).conf
    conf = base.with_(N=1, goal="s", psi_in_v=[1.0], psi_out_v=[1.0])
    phi, psi = jv.initial_fields_3d(conf, 1)
    assert phi.shape == (32, 32, 32, 1)
    assert phi[16, 16, 16, 0] > 0.99 and phi[0, 0, 0, 0] < -0.99
    assert np.allclose(psi, -0.1 * phi + 0.7)
    # the two-star shape is sized for the unit cube (Init_3d.jl centres at x = 0.73, 1.27)
    conf2 = jv.workload_c4(n=64, steps=1).conf.with_(N=2, goal="g", Lx=2.0, Ly=1.0, Lz=1.0, epsilon=0.01)
    phi2, psi2 = jv.initial_fields_3d(conf2, 7)
    i73 = int(round(0.73 / (conf2.Lx / 64)))
    assert phi2[i73, 32, 32, 0] > 0.99 and phi2[i73, 32, 32, 1] < -0.99
    assert np.allclose(psi2, -0.35 * phi2 + 0.45)
    with pytest.raises(ValueError):
        jv.initial_fields_3d(conf2.with_(goal="s"), 7)
