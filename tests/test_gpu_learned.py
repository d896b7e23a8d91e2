"""Learned psi preconditioner on the GPU: the BF16-MFMA network against the FP64
conv oracle (forward and parameter gradient), and the flexible BiCGSTAB that
uses it against the oracle's psi solve (same solution within the solve's
tolerance, same diagonally preconditioned stopping test).

BF16 weights and activations: the network checks use relative L2 tolerances of
3e-2 (forward) and 6e-2 (gradients); the solves use the FP64 tolerance of the
reference's solve (rtol 1e-10), so the solutions agree to ~1e-9."""
import numpy as np
import pytest

from helpers import relerr

pytestmark = pytest.mark.gpu


def _flat(params):
    return np.concatenate([np.concatenate([W.reshape(-1), b]) for W, b in params]).astype(np.float32)


def _bf16(a):
    """Round to BF16 (nearest even) and back, as the device does on input."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u << 16).astype(np.uint32).view(np.float32)


def _l2rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


@pytest.fixture(scope="module")
def ops64(jv):
    wl = jv.workload_c1(n=64, steps=1)
    ops = jv.Operators(wl.conf)
    yield ops
    ops.close()


def _params(co, seed):
    p = co.init_params(seed)
    # non-zero biases and a non-trivial output layer
    rng = np.random.default_rng(seed + 100)
    return [(W, 0.1 * rng.standard_normal(b.shape)) for W, b in p]


def test_load_save_roundtrip(jv, ops64):
    from oracle import conv_oracle as co
    p = _flat(_params(co, 1))
    ops64.precond_load(p)
    assert np.array_equal(ops64.precond_save(), p)


@pytest.mark.parametrize("shape", [(2, 32, 64), (1, 48, 96)])
def test_net_forward_matches_oracle(jv, ops64, shape):
    from oracle import conv_oracle as co
    params = _params(co, 2)
    ops64.precond_load(_flat(params))
    nimg, H, W = shape
    rng = np.random.default_rng(3)
    x = rng.standard_normal((nimg, H, W, 2)).astype(np.float32)
    y = ops64.net_forward(x)
    qparams = [(_bf16(W_).astype(np.float64), b) for W_, b in params]
    ref, _ = co.forward(qparams, _bf16(x).astype(np.float64).transpose(0, 3, 1, 2))
    e = _l2rel(y, ref.transpose(0, 2, 3, 1))
    print("forward rel L2", e)
    assert e < 3e-2


def test_net_gradient_matches_oracle(jv, ops64):
    from oracle import conv_oracle as co
    params = _params(co, 4)
    ops64.precond_load(_flat(params))
    rng = np.random.default_rng(5)
    nimg, H, W, border = 2, 32, 64, 4
    x = rng.standard_normal((nimg, H, W, 2)).astype(np.float32)
    t = rng.standard_normal((nimg, H, W, 9)).astype(np.float32)
    loss, g = ops64.net_grad(x, t, border)
    qparams = [(_bf16(W_).astype(np.float64), b) for W_, b in params]
    y, cache = co.forward(qparams, _bf16(x).astype(np.float64).transpose(0, 3, 1, 2))
    mask = np.zeros((nimg, 9, H, W))
    mask[:, :, border:H - border, border:W - border] = 1.0
    nint = mask.sum()
    d = (y - t.transpose(0, 3, 1, 2)) * mask
    ref_loss = float(np.sum(d * d) / nint)
    grads, _ = co.backward(qparams, cache, 2.0 * d / nint)
    assert abs(loss - ref_loss) <= 2e-2 * ref_loss
    off = 0
    for l, (gW, gb) in enumerate(grads):
        nw = gW.size
        eW = _l2rel(g[off:off + nw], gW.reshape(-1))
        eb = _l2rel(g[off + nw:off + nw + gb.size], gb)
        print("layer", l, "grad rel L2", eW, eb)
        assert eW < 6e-2 and eb < 6e-2, (l, eW, eb)
        off += nw + gb.size


def _psi_problem(jv, so, n=64, seed=5):
    wl = jv.workload_c1(n=n, steps=1)
    conf = wl.conf
    phi, psi = wl.initial_fields()
    orc = so.Oracle(conf)
    M = so.get_M_psi(phi, conf)
    rng = np.random.default_rng(seed)
    rhs = orc.ops.rfft(rng.standard_normal(phi.shape))
    ref = orc.solve_psi(rhs, M, float(M.max()), 1.5, conf.dt)
    return conf, phi, rhs, ref, orc.krylov_log[-1]


def test_untrained_learned_solve_matches_oracle(jv, so):
    """Zero output layer: the learned preconditioner equals the diagonal one and the
    flexible solve reaches the oracle's solution in about as many iterations."""
    conf, phi, rhs, ref, ks = _psi_problem(jv, so)
    ops = jv.Operators(conf, precond=jv.SAV_PRECOND_LEARNED)
    got, info = ops.psi_solve(phi.T, 1.5, conf.dt, rhs.T)
    ops.close()
    assert info["solved"]
    assert abs(info["iters"] - ks.niter) <= 1, (info["iters"], ks.niter)
    assert relerr(got.T, ref) < 1e-8


def test_trained_learned_solve_matches_oracle(jv, so):
    """Collect windows from diagonal solves, train, then solve with the network."""
    conf, phi, rhs, ref, ks = _psi_problem(jv, so)
    ops = jv.Operators(conf)
    ops.precond_collect(True, patches_per_snapshot=16, seed=11)
    rng = np.random.default_rng(9)
    for k in range(8):
        r = ops.rfft(rng.standard_normal(phi.T.shape))
        ops.psi_solve(phi.T, 1.5, conf.dt, r)
    ops.precond_collect(False)
    n = ops.precond_samples()
    assert n > 0
    l0, l1 = ops.precond_train(iters=600, batch=16, lr=2e-3, seed=3)
    print("windows", n, "loss", l0, "->", l1)
    assert np.isfinite(l1) and l1 < l0
    ops.set_precond(jv.SAV_PRECOND_LEARNED)
    got, info = ops.psi_solve(phi.T, 1.5, conf.dt, rhs.T)
    ops.set_precond(jv.SAV_PRECOND_STENCIL)
    got_s, info_s = ops.psi_solve(phi.T, 1.5, conf.dt, rhs.T)
    ops.set_precond(jv.SAV_PRECOND_DIAG)
    got_d, info_d = ops.psi_solve(phi.T, 1.5, conf.dt, rhs.T)
    ops.close()
    print("iterations learned", info["iters"], "stencil", info_s["iters"], "diag", info_d["iters"], "oracle", ks.niter)
    assert relerr(got_s.T, ref) < 1e-8
    assert info["solved"] and info_d["solved"]
    assert relerr(got.T, ref) < 1e-8
    assert relerr(got_d.T, ref) < 1e-11


def test_learned_steps_track_diagonal_steps(jv):
    """Whole SAV steps with the learned preconditioner stay within the solve tolerance
    of the diagonal run (FP64)."""
    wl = jv.workload_c1(n=64, steps=3)
    conf = wl.conf
    phi, psi = wl.initial_fields()
    from helpers import device_state
    runs = {}
    for pc in (jv.SAV_PRECOND_DIAG, jv.SAV_PRECOND_LEARNED):
        ops = jv.Operators(conf, precond=pc)
        ops.set_state(jv.SAV_PRESENT, device_state(jv, ops, conf, phi, psi))
        ops.init_sav()
        its = []
        for n in range(1, 4):
            rc, st = ops.step(conf.dt, n)
            assert rc == 0
            its.append(list(st.krylov_iters))
        runs[pc] = (ops.get_state(jv.SAV_PRESENT), its)
        ops.close()
    a, b = runs[jv.SAV_PRECOND_DIAG][0], runs[jv.SAV_PRECOND_LEARNED][0]
    print("iters diag", runs[jv.SAV_PRECOND_DIAG][1], "learned", runs[jv.SAV_PRECOND_LEARNED][1])
    for name in ("phi", "psi"):
        assert relerr(getattr(b, name), getattr(a, name)) < 1e-8, name
    # nu = S4 psi + R3 G: two terms of size S4 * psi ~ 1e3 cancel to nu ~ 1
    assert relerr(b.nu, a.nu) < 1e-8 * 1e3 * np.max(np.abs(a.psi)) / np.max(np.abs(a.nu))


def test_preconditioners_at_1024_measured_bound(jv):
    """The two-vesicle 1024^2 workload (FP64): windows logged from 4 diagonal steps (about
    20k), then 4 steps from one saved state with the diagonal preconditioner, the analytic
    first-order stencil, the network distilled from that stencil on the logged windows, and
    the network trained further on the solver's own error corrections.

    What is asserted is what was measured: the stencil and the distilled network lower the
    psi iteration count below the diagonal preconditioner's with no fallback solves, but
    every application costs two full-spectrum transforms (a third of a matvec's), so no
    physical-space preconditioner makes a step faster than the diagonal one does."""
    import time
    wl = jv.workload_c3(steps=9)
    conf = wl.conf
    phi, psi = wl.initial_fields()
    from helpers import device_state
    ops = jv.Operators(conf)
    try:
        ops.set_state(jv.SAV_PRESENT, device_state(jv, ops, conf, phi, psi))
        ops.init_sav()
        ops.precond_collect(True, patches_per_snapshot=256, seed=17)
        t = np.arange(10) * conf.dt
        ops.run(t[:5], 1)
        ops.precond_collect(False)
        nwin = ops.precond_samples()
        present, old = ops.get_state(jv.SAV_PRESENT), ops.get_state(jv.SAV_OLD)
        res, losses = {}, {}

        def leg(name, pc):
            ops.set_state(jv.SAV_PRESENT, present)
            ops.set_state(jv.SAV_OLD, old)
            ops.set_precond(pc)
            ops.run(t[4:6], 5)                    # one step to warm up (stencil / network refresh)
            fb0 = ops.precond_fallbacks()
            t0 = time.perf_counter()
            rc, E, A, kit = ops.run(t[5:10], 6)
            el = time.perf_counter() - t0
            res[name] = dict(iters=kit / 4, ms=1e3 * el / 4, fallbacks=ops.precond_fallbacks() - fb0, rc=rc)

        leg("diag", jv.SAV_PRECOND_DIAG)
        leg("stencil", jv.SAV_PRECOND_STENCIL)
        ops.precond_train_target(1)
        losses["distilled"] = ops.precond_train(iters=1000, batch=32, lr=2e-3, seed=5)
        leg("distilled", jv.SAV_PRECOND_LEARNED)
        ops.precond_train_target(0)
        losses["trained"] = ops.precond_train(iters=300, batch=32, lr=2e-4, seed=6)
        leg("trained", jv.SAV_PRECOND_LEARNED)
        print(f"windows {nwin} losses {losses}; per step: " +
              "; ".join(f"{k} {v['iters']:.2f} iters {v['ms']:.2f} ms fallbacks {v['fallbacks']}" for k, v in res.items()),
              flush=True)
    finally:
        ops.close()
    assert nwin > 15000
    l0, l1 = losses["distilled"]
    assert l1 < 0.1 * l0
    for name in ("stencil", "distilled"):
        assert res[name]["iters"] < res["diag"]["iters"], name
        assert res[name]["fallbacks"] == 0, name
    for name in ("stencil", "distilled", "trained"):
        assert res[name]["ms"] > res["diag"]["ms"], name
