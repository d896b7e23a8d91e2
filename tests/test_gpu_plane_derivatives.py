"""3D back-transforms of derivatives in the register-column plane kernels at 128^3: i kx X and
i ky X share the plain field's z pass and the plane pass applies the multiplier (on the packed
DC / Nyquist column through the Hermitian parts, with the Nyquist row y = 64 set apart).  A
rough (white-noise) field gives the DC and Nyquist planes and rows O(1) weight, so the
functionals, whose f_surf reads grad phi (Utils_3d.jl:36-42), and H1..H3, G are compared with
the oracle on it, in both plane-kernel families."""
import numpy as np
import pytest

from helpers import relerr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["register", "lds"])
def test_rough_field_functionals_at_128cubed(jv, so, kind):
    wl = jv.workload_c4(n=128, steps=1)
    conf = wl.conf
    rng = np.random.default_rng(11)
    shape = (conf.N, conf.Nz, conf.Ny, conf.Nx)
    phi = rng.uniform(-1.0, 1.0, shape)
    psi = rng.uniform(0.2, 0.8, shape)
    ops = jv.Operators(conf)
    ops.set_plane_kernel(jv.SAV_PLANE_REGISTER if kind == "register" else jv.SAV_PLANE_LDS)
    orc = so.Oracle(conf)
    A0 = so.calculate_area(phi, orc.ops, conf) * 0.99
    st = jv.FieldState.zeros(conf)
    st.A0 = A0
    ops.set_state(jv.SAV_PRESENT, st)
    H1, H2, H3, G, W, area = ops.eval_functionals(phi.T, psi.T)
    ops.close()
    ref_area = so.calculate_area(phi, orc.ops, conf)
    print(kind, "area", relerr(area, ref_area), "W2", abs(W[1] - so.get_W2(phi, orc.ops, conf)) / abs(W[1]),
          flush=True)
    assert relerr(area, ref_area) < 1e-12
    assert abs(W[0] - so.get_W1(phi, orc.ops, conf, A0)) < 1e-10 * abs(W[0])
    assert abs(W[1] - so.get_W2(phi, orc.ops, conf)) < 1e-12 * abs(W[1])
    assert relerr(H1.T, so.get_H1(phi, orc.ops, conf, A0)) < 1e-11
    assert relerr(H2.T, so.get_H2(phi, orc.ops, conf)) < 1e-11
    assert relerr(H3.T, so.get_H3(phi, psi, conf, orc.ops)) < 1e-12
    assert relerr(G.T, so.get_MG(phi, psi, conf, orc.ops)) < 1e-12
