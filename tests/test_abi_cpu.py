"""Host-side checks that need no GPU: the C ABI library loads and exports every
symbol include/sav_abi.h declares, the config mirror and the workload generators."""
import math
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def _header_functions():
    txt = (ROOT / "include" / "sav_abi.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sav_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(jv):
    import ctypes
    lib = ctypes.CDLL(str(jv.LIB_PATH))          # load only: no device call is made
    names = _header_functions()
    assert len(names) >= 16
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(jv.EXPORTS) == names


def test_net_parameter_count_matches_header_and_oracle(jv):
    from oracle import conv_oracle as co
    txt = (ROOT / "include" / "sav_abi.h").read_text()
    n = int(re.search(r"#define SAV_NET_NPARAM (\d+)", txt).group(1))
    count = sum(W.size + b.size for W, b in co.init_params(0))
    assert n == jv.SAV_NET_NPARAM == count


def test_abi_version_without_device(jv):
    assert jv.load_library().sav_abi_version() == 1


def test_missing_library_fails_loudly(jv, tmp_path):
    with pytest.raises(FileNotFoundError):
        jv.load_library(tmp_path / "libsav_hip.so")


def test_config_int_fields_reject_fractions(jv):
    base = jv.set_para_base(1e-6, 1e-4)
    kw = {k: getattr(base, k) for k in ("N", "epsilon", "M_phi", "M0_psi", "eta", "gamma_surf", "gamma_area",
                                        "gamma_bend", "gamma_in", "beta_in", "psi_in_v", "gamma_out", "beta_out",
                                        "psi_out_v", "lamda", "S1", "S2", "S3", "S4", "C1", "C2", "C3", "dt", "T",
                                        "Nx", "Ny", "Lx", "Ly", "tol")}
    kw["gamma_area"] = 1.5
    with pytest.raises(jv.InexactError):
        jv.Config.make(**kw)


def test_set_para_base_matches_reference_defaults(jv):
    c = jv.set_para_base(1e-6, 1e-3)      # Init.jl:8-69
    assert (c.N, c.Nx, c.Ny, c.Lx, c.Ly) == (1, 256, 256, 1.0, 1.0)
    assert c.epsilon == 0.01 and c.gamma_in == 100000 and c.gamma_area == 50000
    assert c.S1 == pytest.approx(3 * math.sqrt(2) / 8 * 0.01 * 0.1)
    assert c.S3 == pytest.approx(0.1 / 0.01 ** 3) and c.S4 == 1e5
    assert c.Nt == 1000 and c.dx == 1 / 256 and c.spe_len == 129 * 256


def test_time_nodes_two_phase_schedule(jv):
    t = jv.time_nodes(1e-5, 1e-4)           # simulation.jl:15-19
    assert len(t) == 101 + 9
    assert t[0] == 0 and t[100] == pytest.approx(1e-5) and t[-1] == pytest.approx(1e-4)
    assert np.allclose(np.diff(t[:101]), 1e-7)


def test_bdf_coeff(jv):
    assert (jv.bdf_coeff(1).a, jv.bdf_coeff(1).b, jv.bdf_coeff(1).c) == (1.0, -1.0, 0.0)
    assert (jv.bdf_coeff(7).a, jv.bdf_coeff(7).b, jv.bdf_coeff(7).c) == (1.5, -2.0, 0.5)


def test_workloads_are_seeded_and_non_overlapping(jv):
    a, b = jv.workload_c2(n=64), jv.workload_c2(n=64)
    assert [v.centre for v in a.vesicles] == [v.centre for v in b.vesicles]
    L = 2 * math.pi
    for i, v in enumerate(a.vesicles):
        for w in a.vesicles[i + 1:]:
            d = [((v.centre[k] - w.centre[k] + L / 2) % L) - L / 2 for k in range(2)]
            assert math.hypot(*d) > max(v.semi_axes) + max(w.semi_axes)
    ens = jv.workload_c5(steps=1)
    assert len(ens) == 64 and sorted({w.conf.Nx for w in ens}) == [128, 256, 384, 512]
    assert all(1 <= w.conf.N <= 3 for w in ens)


def test_initial_fields_sign_and_range(jv):
    wl = jv.workload_c1(n=64)
    phi, psi = wl.initial_fields()
    assert phi.shape == (1, 64, 64)
    assert phi[0, 32, 32] > 0.99 and phi[0, 0, 0] < -0.99          # +1 inside (Init.jl:113)
    assert psi.max() <= 1.2 + 1e-12 and psi.min() >= 1.0 - 1e-12


def test_splitmix64_known_values(jv):
    # SplitMix64 reference outputs for seed 0 (Steele et al.; the same constants as java.util.SplittableRandom)
    r = jv.SplitMix64(0)
    assert r.next_u64() == 0xE220A8397B1DCDAF
    assert r.next_u64() == 0x6E789E6AA1B965F4


def test_julia_shim_sizes_A0_by_the_vesicle_count():
    # the library copies N values of A0 in and out (sav_set_state / sav_get_state), while the
    # reference's FieldState holds A0 = [0.0] (Types.jl:228): the binding must size its
    # buffers by N, never by length(s.A0)
    src = (ROOT / "julia-vesicles-growth-and-shrinkage_amd" / "julia" / "SavHip.jl").read_text()
    body = lambda name: src.split(f"function {name}(")[1].split("\nend\n")[0]
    for fn in ("set_state!", "get_state!"):
        b = body(fn)
        assert "nvesicles(s)" in b, fn
        assert "length(s.A0))" not in b.replace("min(N, length(s.A0))", ""), fn
    assert "nvesicles(s) = size(s.phi, ndims(s.phi))" in src
    doc = (ROOT / "INTEGRATION.md").read_text()
    assert "sc, s.A0, s.dt)" not in doc
