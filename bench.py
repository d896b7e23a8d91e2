#!/usr/bin/env python3
"""SAV steps/s of the seven-step pseudo-spectral timestep on one MI355X.

Default workload: two seeded vesicles on 1024^2 (eps = 0.015, dt = 2e-5, BiCGSTAB
rtol 1e-10, itmax 20) in FP64, the reference's arithmetic.  The same line reports the
mixed form (FP32 inner psi iterations inside an FP64 defect correction) and the FP32-fields
form with their field errors against FP64, the FP64 and FP32 forms at 128^3 (two
ellipsoids), the psi-solve iteration counts, the roofline of the dominant FP64 kernel,
graph replay, the learned preconditioner, the 64-instance ensemble and the CPU oracle
timed on the host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
TRAFFIC_PROFILE = ROOT / "profiles" / "traffic_psi_matvec_fp64_two_vesicles_1024.md"


def _traffic_ratio(kernel):
    """HBM bytes / algorithmic bytes of one FP64 psi-matvec kernel at 1024^2, measured with
    rocprofv3 FETCH_SIZE/WRITE_SIZE counters scaled by a plain copy (the committed profile
    holds one `| traffic_over_algorithmic <kernel> | ratio |` row per kernel)."""
    try:
        for line in TRAFFIC_PROFILE.read_text().splitlines():
            if line.startswith(f"| traffic_over_algorithmic {kernel} |"):
                return float(line.split("|")[2])
    except (OSError, ValueError):
        pass
    return None


def _sync():
    import torch
    torch.cuda.synchronize()


def copy_ceiling_gbs(nbytes=1 << 30, reps=5):
    """Rate of a plain device copy of `nbytes` (read + write, buffers four times the 256 MB
    infinity cache), best of `reps`: what HBM sustains for a streaming kernel, beside the
    8 TB/s peak the roofline fraction is taken against."""
    import torch
    src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    dst.copy_(src)
    best = float("inf")
    for _ in range(reps):
        e0.record()
        dst.copy_(src)
        e1.record()
        e1.synchronize()
        best = min(best, e0.elapsed_time(e1))
    del src, dst
    torch.cuda.empty_cache()
    return 2.0 * nbytes / (best * 1e-3) / 1e9


def _setup(jv, wl, precision):
    conf = wl.conf
    phi, psi = wl.initial_fields()
    ops = jv.Operators(conf, precision=precision)
    st = jv.FieldState.zeros(conf)
    st.phi[...] = phi.T
    st.psi[...] = psi.T
    st.phi_hat[...] = ops.rfft_accurate(st.phi)
    st.psi_hat[...] = ops.rfft_accurate(st.psi)
    st.Q = st.R1 = st.R2 = st.R3 = 1.0
    ops.set_state(jv.SAV_PRESENT, st)
    ops.init_sav()
    return ops


COLUMN_TILE = 0  # psi-matvec columns per workgroup: 0 = autotuned per handle, else fixed (8 or 16)


def run_case(jv, wl, precision, steps, warmup, timed_kernel=0, graph=False):
    """Warmup steps untimed, then `steps` timed steps (inputs resident in HBM); with
    `graph`, each step is replayed as one captured HIP graph."""
    conf = wl.conf
    ops = _setup(jv, wl, precision)
    if COLUMN_TILE and conf.ndim == 2:
        ops.set_column_tile(COLUMN_TILE)
    ops.set_graph(graph)
    t = np.arange(warmup + steps + 1) * conf.dt
    if warmup:
        ops.run(t[: warmup + 1], 1)
    if timed_kernel:
        ops.kernel_timing(timed_kernel)
    _sync()
    t0 = time.perf_counter()
    rc, E, A, kit = ops.run(t[warmup:], warmup + 1)
    _sync()
    t1 = time.perf_counter()
    res = dict(rc=rc, seconds=t1 - t0, steps_per_s=steps / (t1 - t0), ms_per_step=1e3 * (t1 - t0) / steps,
               krylov_iters=kit, energy_first=float(E[0]), energy_last=float(E[-1]),
               energy_monotone=bool(np.all(np.diff(E) <= 0)))
    if timed_kernel:
        n, ms, b = ops.kernel_timing_get()
        res["kernel"] = dict(name=ops.KERNELS[timed_kernel], launches=n, avg_ms=ms / max(n, 1), bytes=b)
    res["autotune"] = ops.autotune_info() if conf.ndim == 2 else None
    res["matvec_order"] = ops.matvec_order_info() if conf.ndim == 2 else None
    res["transform_tiles"] = dict(inverse=ops.transform_tile_info(), forward=ops.transform_tile_fwd_info())
    res["plane_kernel"] = ops.plane_kernel_info() if conf.ndim == 3 else None
    if graph:
        res["graph"] = ops.graph_info()
    st = ops.get_state(jv.SAV_PRESENT)
    ops.close()
    return res, st


def run_case_learned(jv, wl, precision, steps, warmup, collect_steps=4, train_iters=1000, tune_iters=300):
    """Learned psi preconditioner: windows logged from `collect_steps` diagonal steps, then
    Adam training on the GPU in two stages: distillation of the analytic first-order
    stencil (targets: that stencil's output on the logged windows), then training on the
    solver's own error corrections.  From one saved state, `warmup` + `steps` steps are
    timed with the diagonal preconditioner, the analytic stencil, the distilled network and
    the further-trained network."""
    conf = wl.conf
    ops = _setup(jv, wl, precision)
    # about 20k windows: 4 steps x ~19 diagonal iterations x 256 windows per iterate
    ops.precond_collect(True, patches_per_snapshot=256, seed=17)
    t = np.arange(collect_steps + warmup + steps + 1) * conf.dt
    rc, E0, A0, kit_diag = ops.run(t[: collect_steps + 1], 1)
    ops.precond_collect(False)
    nwin = ops.precond_samples()
    present, old = ops.get_state(jv.SAV_PRESENT), ops.get_state(jv.SAV_OLD)

    def leg(pc):
        ops.set_state(jv.SAV_PRESENT, present)
        ops.set_state(jv.SAV_OLD, old)
        ops.set_precond(pc)
        if warmup:
            ops.run(t[collect_steps: collect_steps + warmup + 1], collect_steps + 1)
        _sync()
        fb0 = ops.precond_fallbacks()
        t0 = time.perf_counter()
        rc, E, A, kit = ops.run(t[collect_steps + warmup:], collect_steps + warmup + 1)
        _sync()
        el = time.perf_counter() - t0
        return dict(rc=rc, steps_per_s=steps / el, ms_per_step=1e3 * el / steps, psi_iters_per_step=kit / steps,
                    diag_fallback_solves=ops.precond_fallbacks() - fb0,
                    energy_monotone=bool(np.all(np.diff(E) <= 0)))

    res = dict(windows=nwin, diag_psi_iters_per_step_during_collection=kit_diag / collect_steps)
    res["diag"] = leg(jv.SAV_PRECOND_DIAG)
    res["stencil"] = leg(jv.SAV_PRECOND_STENCIL)
    ops.precond_train_target(1)
    _sync()
    t0 = time.perf_counter()
    l0, l1 = ops.precond_train(iters=train_iters, batch=32, lr=2e-3, seed=5)
    _sync()
    res["distilled"] = leg(jv.SAV_PRECOND_LEARNED)
    res["distilled"].update(train_iters=train_iters, train_seconds=round(time.perf_counter() - t0, 3),
                            loss_first=l0, loss_last=l1)
    ops.precond_train_target(0)
    t0 = time.perf_counter()
    l0, l1 = ops.precond_train(iters=tune_iters, batch=32, lr=2e-4, seed=6)
    _sync()
    res["trained"] = leg(jv.SAV_PRECOND_LEARNED)
    res["trained"].update(train_iters=tune_iters, train_seconds=round(time.perf_counter() - t0, 3),
                          loss_first=l0, loss_last=l1)
    ops.close()
    return res


def run_ensemble(jv, steps, precision, threads_list=(1, 4, 8, 16), graph=False, batched=True):
    """The 64-instance ensemble (16 each of 128^2, 256^2, 384^2, 512^2, 1-3 vesicles):
    aggregate steps/s for each host-thread count (1 = the sequential per-instance loop),
    after one warm-up step of every member."""
    wls = jv.workload_c5(steps=steps)
    ens = jv.Ensemble([w.conf for w in wls], precision=precision)
    for m, wl in zip(ens.members, wls):
        phi, psi = wl.initial_fields()
        st = jv.FieldState.zeros(wl.conf)
        st.phi[...] = phi.T
        st.psi[...] = psi.T
        st.phi_hat[...] = m.rfft_accurate(st.phi)
        st.psi_hat[...] = m.rfft_accurate(st.psi)
        st.Q = st.R1 = st.R2 = st.R3 = 1.0
        m.set_state(jv.SAV_PRESENT, st)
        m.init_sav()
    ens.set_graph(graph)
    dt = wls[0].conf.dt
    n0 = 1
    ens.run(np.arange(2) * dt, n0, threads=16)
    n0 += 1
    res = {}
    for th in threads_list:
        t = np.arange(n0 - 1, n0 + steps) * dt
        _sync()
        t0 = time.perf_counter()
        rc, E, A, kit, st = ens.run(t, n0, threads=th)
        _sync()
        el = time.perf_counter() - t0
        n0 += steps
        # algorithmic bytes of the run: every member's steps by the byte model, with its own
        # psi iteration count
        sb = 4 if precision == jv.SAV_PREC_FP32 else 8
        B = sum(algorithmic_bytes(w.conf, sb, float(k) / steps) * steps for w, k in zip(wls, kit))
        Bb = sum(algorithmic_bytes_built(w.conf, sb, float(k) / steps) * steps for w, k in zip(wls, kit))
        res[th] = dict(seconds=el, steps_per_s=len(wls) * steps / el, rc=int(rc), psi_iters=int(kit.sum()),
                       algorithmic_tb_per_s=B / el / 1e12, built_kernels_tb_per_s=Bb / el / 1e12)
    # size classes: each step's psi solves batched over the members of equal grid and N
    # (one host thread per member); one warm-up step tunes the groups' matvecs
    if not batched:
        ens.close()
        return dict(instances=len(wls), steps=steps, by_threads=res)
    groups = ens.set_batched(True)
    ens.run(np.arange(n0 - 1, n0 + 1) * dt, n0)
    n0 += 1
    t = np.arange(n0 - 1, n0 + steps) * dt
    _sync()
    t0 = time.perf_counter()
    rc, E, A, kit, st = ens.run(t, n0)
    _sync()
    el = time.perf_counter() - t0
    gof, _ = ens.batched_info()
    sb = 4 if precision == jv.SAV_PREC_FP32 else 8
    B = sum(algorithmic_bytes(w.conf, sb, float(k) / steps) * steps for w, k in zip(wls, kit))
    Bb = sum(algorithmic_bytes_built(w.conf, sb, float(k) / steps) * steps for w, k in zip(wls, kit))
    batched = dict(seconds=el, steps_per_s=len(wls) * steps / el, rc=int(rc), psi_iters=int(kit.sum()),
                   algorithmic_tb_per_s=B / el / 1e12, built_kernels_tb_per_s=Bb / el / 1e12, groups=int(groups),
                   grouped_members=int((gof >= 0).sum()))
    ens.close()
    best = max(res, key=lambda k: res[k]["steps_per_s"])
    sizes = sorted({w.conf.Nx for w in wls})
    sb = 4 if precision == jv.SAV_PREC_FP32 else 8
    per = {n: sum(algorithmic_bytes(w.conf, sb, 20.0) for w in wls if w.conf.Nx == n) for n in sizes}
    tot = sum(per.values())
    return dict(instances=len(wls), steps=steps, by_threads=res, best_threads=best,
                speedup_vs_sequential=res[best]["steps_per_s"] / res[1]["steps_per_s"], batched=batched,
                byte_share_by_grid={f"{n}^2": round(v / tot, 4) for n, v in per.items()})


def cpu_baseline(jv, wl, nsteps=3):
    """The FP64 CPU oracle on the host cores: the first `nsteps` steps (BDF1, then BDF2)
    of the same workload after the SAV initialisation, a bounded sample of the timed work."""
    from oracle import sav_oracle as so
    workers = min(os.cpu_count() or 1, 16)
    conf = wl.conf
    phi, psi = wl.initial_fields()
    orc = so.Oracle(conf, workers=workers)
    pres, old = orc.init_sav(so.state_from_physical(phi, psi, orc.ops, conf.ndim))
    t0 = time.perf_counter()
    for n in range(1, nsteps + 1):
        orc.step(pres, old, conf.dt, n)
    dt = time.perf_counter() - t0
    return dict(value=nsteps / dt, unit="SAV steps/s", cores=workers, kind="port",
                sample=f"{nsteps} steps (BDF1, BDF2) of {wl.name} ({conf.Nx}^{conf.ndim}, N={conf.N}) with the "
                       f"monitor, numpy/scipy FP64 oracle, scipy.fft workers={workers}, pointwise single-threaded; "
                       f"{dt:.1f} s")


def algorithmic_bytes(conf, s, k):
    """Algorithmic HBM bytes of one SAV step: every logical transform reads its input and
    writes its output once, fused pointwise work counts its unique operands, a state
    rotation is free.  u = one real N-field, v = one scalar field (s bytes per value);
    k = psi BiCGSTAB iterations of the step (88 u each: two matvecs and the vector
    updates over full complex spectra).  2D: (282 + 88 k) u + 94 v; 3D: 280 u + 144 v."""
    pts = conf.Nx * conf.Ny * (conf.Nz if conf.ndim == 3 else 1)
    u, v = s * pts * conf.N, s * pts
    if conf.ndim == 3:
        return 280 * u + 144 * v
    return (282 + 88 * k) * u + 94 * v


def algorithmic_bytes_built(conf, s, k):
    """HBM bytes of one SAV step as this library's kernels move them (u = one real N-field,
    v = one scalar field, a half spectrum counted as u or v; a transform of one field is two
    passes that each read and write it, 4u, plus u when it extrapolates 2 X - Y on load).
    Outside the psi iterations, kernel by kernel:
    * 2D: step 1 148u + 8v (extrapolations 7u, M_max u, Hermitian part of phi* 3u, grad phi* and
      Lap phi* 12u, grad psi* 10u, functionals 6u, rfft omega 4u, G/H2/H3 6u + rfft 12u, spectral
      H1/H2 5u, advection 6u + 2v + rfft 8u, right-hand sides 5u, M grad G 24u, full right-hand
      sides 9u, Krylov set-up 24u, solution read-out 6u; velocity extrapolation 6v), step 2 12u,
      step 4 26u + 47v, step 5 11u + 6v, steps 6-7 14u + 8v, back-transform 8u + 8v, monitor
      6u + 5v: 225u + 82v;
    * 3D: step 1 98u + 12v, step 2 12u, step 4 32u + 79v, step 5 11u + 9v, steps 6-7 14u + 11v,
      back-transform 8u + 12v, monitor 6u + 7v: 181u + 130v (the back-transforms of x and y
      derivatives share the field's z pass);
    and per psi system-iteration FP64 72u (k_mv_cols_inv with the s / p update on load 10u + 12u,
    two k_mv_rows 18u, k_mv_cols_fwd 10u + 8u, k_kry_upd2 14u), FP32 76u (separate updates:
    column inverse 6u + 6u, rows 18u, column forward 10u + 8u, s update 6u, k_kry_upd2 14u,
    p update 8u), in place of the survey's 88u."""
    pts = conf.Nx * conf.Ny * (conf.Nz if conf.ndim == 3 else 1)
    u, v = s * pts * conf.N, s * pts
    if conf.ndim == 3:
        return 181 * u + 130 * v
    return (225 + (72 if s == 8 else 76) * k) * u + 82 * v


def cpu_report(jv, which="all", steps=20, warm=2):
    """The FP64 CPU oracle on the host: `steps` steps after `warm` steps of the 1024^2 ("2d")
    and 128^3 ("3d") workloads, with scipy.fft on every host core, on 8 and on 1 (the monitor
    included, as the reference loop runs it every step)."""
    from oracle import sav_oracle as so
    info = _host_info()
    rows = {"host": info}
    allc = min(os.cpu_count() or 1, 16)
    wls = {"2d": [jv.workload_c3()], "3d": [jv.workload_c4()]}.get(which, [jv.workload_c3(), jv.workload_c4()])
    for wl in wls:
        conf = wl.conf
        phi, psi = wl.initial_fields()
        row = {}
        for workers in sorted({allc, 8, 1}, reverse=True):
            orc = so.Oracle(conf, workers=workers)
            pres, old = orc.init_sav(so.state_from_physical(phi, psi, orc.ops, conf.ndim))
            for n in range(1, warm + 1):
                orc.step(pres, old, conf.dt, n)
            t0 = time.perf_counter()
            for n in range(warm + 1, warm + steps + 1):
                orc.step(pres, old, conf.dt, n)
                print(f"cpu {wl.name} workers {workers}: step {n} at {time.perf_counter() - t0:.1f} s",
                      file=sys.stderr, flush=True)
            el = time.perf_counter() - t0
            row[f"fft_workers_{workers}"] = dict(steps_per_s=round(steps / el, 4), seconds=round(el, 1))
            print(f"cpu {wl.name} workers {workers}: {steps / el:.4f} steps/s", file=sys.stderr, flush=True)
        rows[wl.name] = row
    return rows


def run_planes(jv, fields=8, repeats=5):
    """Both (x, y) plane-kernel families of 3D transforms run alone: `repeats` rfft and irfft
    calls of `fields` random 128^3 fields per precision and family (host copies included;
    a kernel trace gives each kernel's duration), with the round-trip error."""
    out = {}
    rng = np.random.default_rng(1)
    u = rng.standard_normal((fields, 128, 128, 128))
    for prec, pname in ((jv.SAV_PREC_FP64, "fp64"), (jv.SAV_PREC_FP32, "fp32")):
        conf = jv.workload_c4(n=128, steps=1).conf.with_(N=fields, psi_in_v=[1.2] * fields,
                                                        psi_out_v=[1.0] * fields)
        ops = jv.Operators(conf, precision=prec)
        for kind, kname in ((jv.SAV_PLANE_REGISTER, "register"), (jv.SAV_PLANE_LDS, "lds")):
            ops.set_plane_kernel(kind)
            for _ in range(repeats):
                back = ops.irfft(ops.rfft(u.T))
            out[f"{pname}_{kname}_roundtrip_max_err"] = float(np.max(np.abs(back - u.T)))
        ops.close()
    return out


def _run_full(jv, wl, precision):
    """The workload's whole step schedule (uniform dt, BDF1 then BDF2) after the SAV
    initialisation; the steps are timed together, with step replay chosen by measurement as
    run_simulation does (graph_choice: 1 replay kept, 0 eager, -1 still measuring)."""
    conf = wl.conf
    ops = _setup(jv, wl, precision)
    ops.set_graph(jv.GRAPH_AUTO)
    t = np.arange(wl.steps + 1) * conf.dt
    _sync()
    t0 = time.perf_counter()
    rc, E, A, kit = ops.run(t, 1)
    _sync()
    el = time.perf_counter() - t0
    st = ops.get_state(jv.SAV_PRESENT)
    choice = ops.graph_info()["choice"]
    ops.close()
    return dict(rc=int(rc), steps_per_s=wl.steps / el, seconds=el, krylov_iters=int(kit), E=np.asarray(E),
                graph_choice=int(choice)), st


def _rel(a, b):
    den = float(np.max(np.abs(b)))
    return float(np.max(np.abs(a - b))) / den if den > 0 else float(np.max(np.abs(a - b)))


def run_configs(jv, cpu=True):
    """One row per benchmark configuration at its full step count: SAV steps/s in FP64,
    in the mixed form (2D: FP32 inner psi iterations in an FP64 defect correction) and in
    the FP32-fields form, psi iterations per step, the max-norm relative error of each form
    against FP64 for phi, psi, u and the modified energy (over every step), whether each
    energy history is non-increasing, algorithmic bytes per step and the rate they imply.
    The single-vesicle 128^2 case also runs the FP64 CPU oracle over all its steps and
    reports the GPU's FP64 error against it."""
    rows = {}
    for wl in (jv.workload_c1(), jv.workload_c2(), jv.workload_c3(), jv.workload_c4()):
        conf = wl.conf
        runs = {"fp64": _run_full(jv, wl, jv.SAV_PREC_FP64)}
        if conf.ndim == 2:
            runs["mixed"] = _run_full(jv, wl, jv.SAV_PREC_MIXED)
        runs["fp32"] = _run_full(jv, wl, jv.SAV_PREC_FP32)
        r64, s64 = runs["fp64"]
        row = dict(grid=f"{conf.Nx}^{conf.ndim}" if conf.ndim == 3 else f"{conf.Nx}x{conf.Ny}", N=conf.N,
                   steps=wl.steps, epsilon=conf.epsilon, dt=conf.dt, S2=conf.S2,
                   status={k: r["rc"] for k, (r, _) in runs.items()})
        for name, (r, st) in runs.items():
            sb = 4 if name == "fp32" else 8
            k = r["krylov_iters"] / wl.steps
            B = algorithmic_bytes(conf, sb, k)
            row[name] = dict(steps_per_s=round(r["steps_per_s"], 2), graph_choice=r["graph_choice"],
                             psi_iters_per_step=round(k, 3),
                             algorithmic_gb_per_step=round(B / 1e9, 4),
                             algorithmic_tb_per_s=round(B * r["steps_per_s"] / 1e12, 3),
                             energy_monotone=bool(np.all(np.diff(r["E"]) <= 0)))
            if name != "mixed":
                Bb = algorithmic_bytes_built(conf, sb, k)
                row[name]["built_kernels_tb_per_s"] = round(Bb * r["steps_per_s"] / 1e12, 3)
            if name != "fp64":
                row[name]["max_rel_err_vs_fp64"] = {"phi": _rel(st.phi, s64.phi), "psi": _rel(st.psi, s64.psi),
                                                    "u": _rel(st.u, s64.u), "energy": _rel(r["E"], r64["E"])}
        if cpu and wl.name == "single_vesicle_128":
            from oracle import sav_oracle as so
            phi, psi = wl.initial_fields()
            orc = so.Oracle(conf, workers=min(os.cpu_count() or 1, 16))
            pres, old = orc.init_sav(so.state_from_physical(phi, psi, orc.ops, conf.ndim))
            t0 = time.perf_counter()
            for n in range(1, wl.steps + 1):
                orc.step(pres, old, conf.dt, n)
            el = time.perf_counter() - t0
            row["cpu_oracle"] = dict(steps_per_s=round(wl.steps / el, 3), cores=min(os.cpu_count() or 1, 16),
                                     gpu_fp64_vs_oracle_max_rel_err={
                                         "phi": _rel(s64.phi.T, pres.phi), "psi": _rel(s64.psi.T, pres.psi)})
        rows[wl.name] = row
        print(f"config {wl.name}: " + " ".join(f"{k} {row[k]['steps_per_s']}" for k in runs) + " steps/s",
              file=sys.stderr, flush=True)
    ens = run_ensemble(jv, 500, jv.SAV_PREC_FP64, threads_list=(1, 16))
    rows["ensemble_64"] = dict(steps=500, precision="fp64",
                               sequential_steps_per_s=round(ens["by_threads"][1]["steps_per_s"], 2),
                               threads16_steps_per_s=round(ens["by_threads"][16]["steps_per_s"], 2),
                               speedup_vs_sequential=round(ens["speedup_vs_sequential"], 3),
                               batched_size_classes_steps_per_s=round(ens["batched"]["steps_per_s"], 2),
                               batched_groups=ens["batched"]["groups"],
                               status=[v["rc"] for v in ens["by_threads"].values()] + [ens["batched"]["rc"]])
    return rows


FIELD_KEYS = ("phi", "psi", "u", "mu", "nu", "p")


def _errors(s, ref, E=None, Eref=None):
    """Max-norm relative error of each field of `s` against `ref` (and of the energy history)."""
    out = {k: _rel(getattr(s, k), getattr(ref, k)) for k in FIELD_KEYS}
    if E is not None:
        out["energy"] = _rel(np.asarray(E), np.asarray(Eref))
    return out


def _host_info():
    info = {"nproc": os.cpu_count()}
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                info["cpu_model"] = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return info


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle leg")
    ap.add_argument("--only", choices=["all", "2d", "3d", "ensemble", "fp64-2d", "mixed-2d", "fp64-3d", "planes",
                                       "roofline-2d"],
                    default="all", help="fp64-2d / mixed-2d / fp64-3d: one run of that form, for profiler runs; "
                                        "roofline-2d: the headline's live-timed run of the roofline kernel alone")
    ap.add_argument("--ensemble-steps", type=int, default=10)
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-preconditioner leg")
    ap.add_argument("--column-tile", type=int, default=0, choices=[0, 4, 8, 16],
                    help="fix the psi matvec's columns per workgroup (0: autotuned; fixing it makes "
                         "separate profiler runs launch the same kernels)")
    ap.add_argument("--cpu-steps", type=int, default=5, help="steps of the CPU oracle sample")
    ap.add_argument("--configs", action="store_true",
                    help="one row per benchmark configuration at its full step count (a separate report, "
                         "not the headline line)")
    ap.add_argument("--cpu-report", nargs="?", const="all", choices=["all", "2d", "3d"],
                    help="time the FP64 CPU oracle over 20 steps of the 1024^2 (2d) and 128^3 (3d) "
                         "workloads with all host cores, 8 and 1 (a separate report; minutes)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a plain device copy of 256 MiB (x4), the reference kernel for "
                         "scaling rocprofv3 FETCH_SIZE/WRITE_SIZE counters to bytes")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(sharding over several GPUs is designed, not built)"}))
        sys.exit(2)
    global COLUMN_TILE
    COLUMN_TILE = a.column_tile
    import torch  # noqa: F401  (device-wide synchronisation around the timed region)
    import __graft_entry__ as ge
    jv = ge.load_package()
    if a.copy_calibration:
        src = torch.empty(64 * 2 ** 20, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(4):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    if a.configs:
        print(json.dumps(run_configs(jv, cpu=not a.no_cpu)), flush=True)
        return
    if a.cpu_report:
        print(json.dumps(cpu_report(jv, a.cpu_report)), flush=True)
        return
    if a.only == "planes":
        print(json.dumps(run_planes(jv)), flush=True)
        return
    if a.only == "ensemble":
        print(json.dumps(run_ensemble(jv, a.ensemble_steps, jv.SAV_PREC_FP64)), flush=True)
        return
    wl3 = jv.workload_c3(n=1024, steps=a.steps)
    wl4 = jv.workload_c4(n=128, steps=a.steps)
    if a.only == "roofline-2d":
        # the same run as the headline's roofline leg: k_mv_cols_inv timed live with HIP events
        r, _ = run_case(jv, wl3, jv.SAV_PREC_FP64, a.steps, a.warmup, timed_kernel=1)
        k = r["kernel"]
        print(json.dumps({"only": a.only, "kernel": k["name"], "avg_launch_ms": k["avg_ms"],
                          "launches": k["launches"], "bytes_per_launch": k["bytes"],
                          "achieved_gb_s": k["bytes"] / (k["avg_ms"] * 1e-3) / 1e9,
                          "steps_per_s": r["steps_per_s"]}), flush=True)
        return
    if a.only in ("fp64-2d", "mixed-2d", "fp64-3d"):
        wl, prec = {"fp64-2d": (wl3, jv.SAV_PREC_FP64), "mixed-2d": (wl3, jv.SAV_PREC_MIXED),
                    "fp64-3d": (wl4, jv.SAV_PREC_FP64)}[a.only]
        r, _ = run_case(jv, wl, prec, a.steps, a.warmup)
        print(json.dumps({"only": a.only, "steps_per_s": r["steps_per_s"], "ms_per_step": r["ms_per_step"],
                          "krylov_iters": r["krylov_iters"]}), flush=True)
        return
    out = {}
    if a.only in ("all", "2d"):
        r64, s64 = run_case(jv, wl3, jv.SAV_PREC_FP64, a.steps, a.warmup)
        # each psi-matvec kernel timed live in a run of its own (the events delay the next dispatch)
        timed = {}
        for kid in (1, 2, 3):
            rk, _ = run_case(jv, wl3, jv.SAV_PREC_FP64, a.steps, a.warmup, timed_kernel=kid)
            timed[kid] = rk["kernel"]
        rmx, smx = run_case(jv, wl3, jv.SAV_PREC_MIXED, a.steps, a.warmup)
        r32, s32 = run_case(jv, wl3, jv.SAV_PREC_FP32, a.steps, a.warmup)
        rg, sg = run_case(jv, wl3, jv.SAV_PREC_FP64, a.steps, a.warmup, graph=True)
        out["2d"] = dict(fp64=r64, timed=timed, mixed=rmx, fp32=r32, graph=rg,
                         mixed_err=_errors(smx, s64), fp32_err=_errors(s32, s64),
                         graph_equal=all(np.array_equal(getattr(sg, k), getattr(s64, k)) for k in FIELD_KEYS),
                         copy_gb_s=copy_ceiling_gbs())
        if not a.no_learned:
            out["2d"]["learned"] = run_case_learned(jv, wl3, jv.SAV_PREC_FP64, a.steps, a.warmup)
    if a.only in ("all", "3d"):
        q64, t64 = run_case(jv, wl4, jv.SAV_PREC_FP64, a.steps, a.warmup)
        q32, t32 = run_case(jv, wl4, jv.SAV_PREC_FP32, a.steps, a.warmup)
        out["3d"] = dict(fp64=q64, fp32=q32, fp32_err=_errors(t32, t64))

    if "2d" in out:
        head, wl_head, grid = out["2d"]["fp64"], wl3, "1024x1024"
    else:
        head, wl_head, grid = out["3d"]["fp64"], wl4, "128^3"
    conf = wl_head.conf
    line = {
        "metric": "SAV steps/s at 1024^2 (two vesicles, FP64, diagonal psi preconditioner)" if "2d" in out
                  else "SAV steps/s at 128^3 (two ellipsoids, FP64)",
        "value": round(head["steps_per_s"], 3),
        "unit": "SAV steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded initial vesicle fields)",
        "config": {"workload": wl_head.name, "grid": grid, "N": conf.N, "epsilon": conf.epsilon, "dt": conf.dt,
                   "S2": round(conf.S2, 6), "krylov_rtol": conf.tol, "itmax": 20, "precond": "diag",
                   "parallelism": "single GPU"},
    }
    if "2d" in out:
        o = out["2d"]
        # the dominant psi-matvec kernel of the FP64 step: k_mv_cols_inv, the largest share of
        # kernel time in the kernel trace (its three variants 19.3 %; k_mv_cols_fwd 19.1 %,
        # k_mv_rows 15.6 %); all three are listed below
        kid = 1
        k = o["timed"][kid]
        achieved = k["bytes"] / (k["avg_ms"] * 1e-3) / 1e9
        ratio = _traffic_ratio(k["name"])
        line["roofline"] = {"bound": "hbm", "kernel": k["name"] + "<double, 1024>", "achieved": round(achieved, 1),
                            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                            "traffic": round(ratio * k["bytes"]) if ratio else None,
                            "traffic_over_algorithmic": ratio,
                            "avg_launch_ms": round(k["avg_ms"], 5), "launches": k["launches"],
                            "bytes_per_launch": round(k["bytes"])}
        if "copy_gb_s" in o:
            line["roofline"]["copy_gb_s_1GiB"] = round(o["copy_gb_s"], 1)
            line["roofline"]["frac_of_copy"] = round(achieved / o["copy_gb_s"], 4)
        line["psi_matvec_kernels_fp64_1024"] = {
            v["name"]: {"avg_launch_ms": round(v["avg_ms"], 5), "bytes_per_launch": round(v["bytes"]),
                        "achieved_gb_s": round(v["bytes"] / (v["avg_ms"] * 1e-3) / 1e9, 1)}
            for v in o["timed"].values()}
        line["psi_iters_per_step_1024"] = {"fp64": o["fp64"]["krylov_iters"] / a.steps,
                                           "mixed_inner_fp32": o["mixed"]["krylov_iters"] / a.steps}
        line["energy_monotone_1024"] = {"fp64": o["fp64"]["energy_monotone"], "mixed": o["mixed"]["energy_monotone"],
                                        "fp32": o["fp32"]["energy_monotone"]}
        line["mixed_1024"] = {"steps_per_s": round(o["mixed"]["steps_per_s"], 3),
                              "speedup_vs_fp64": round(o["mixed"]["steps_per_s"] / o["fp64"]["steps_per_s"], 3),
                              "dtype": "f64 state and stages, f32 inner psi iterations (f64 defect correction)",
                              "max_rel_err_vs_fp64": o["mixed_err"]}
        line["fp32_fields_1024"] = {"steps_per_s": round(o["fp32"]["steps_per_s"], 3),
                                    "dtype": "f32 fields and transforms, f64 reductions",
                                    "max_rel_err_vs_fp64": o["fp32_err"]}
        line["fp64_graph_replay_1024"] = {"steps_per_s": round(o["graph"]["steps_per_s"], 3),
                                          "bit_identical_to_eager": o["graph_equal"], **o["graph"]["graph"]}
        tune = dict(o["fp64"]["autotune"])
        if 0.0 in (tune["ms_4"], tune["ms_8"], tune["ms_16"]):
            tune["note"] = ("a width whose kernels resolve to those of a width already timed is not timed "
                            "(ms 0): at 1024 in FP64, 160 KB of LDS holds at most 8 columns, so 16 is 8")
        line["psi_matvec_column_tile_1024"] = tune
        tt = o["fp64"]["transform_tiles"]
        line["transform_tiles_fp64_1024"] = {d: {"columns": v["columns"], **{f"ms_{c}": round(v[f"ms_{c}"], 4) for c in (4, 8, 16)}} for d, v in tt.items()}
        mo = o["fp64"]["matvec_order"]
        line["psi_matvec_order_1024"] = {
            "order": {1: "all systems per launch", 2: "system by system (shared intermediate pair)"}.get(mo["order"], mo["order"]),
            "ms_batched": mo["ms_batched"], "ms_sequenced": mo["ms_sequenced"]}
        if "learned" in o:
            L = o["learned"]

            def _leg(d):
                keys = ("steps_per_s", "ms_per_step", "psi_iters_per_step", "diag_fallback_solves",
                        "train_iters", "train_seconds", "loss_first", "loss_last")
                return {k: (round(d[k], 4) if isinstance(d[k], float) else d[k]) for k in keys if k in d}
            line["learned_precond_1024"] = {
                "dtype": "f64 solve, bf16 network", "windows": L["windows"],
                "from_one_saved_state": {name: _leg(L[name]) for name in ("diag", "stencil", "distilled", "trained")},
                "note": ("the network is first trained to reproduce the analytic first-order stencil "
                         "(distilled), then on the solver's own error corrections (trained); every "
                         "application of a physical-space preconditioner costs two full-spectrum "
                         "transforms, more than the iterations it saves, so the diagonal "
                         "preconditioner stays the default")}
    if "3d" in out:
        o = out["3d"]
        line["fp64_steps_per_s_128cubed"] = round(o["fp64"]["steps_per_s"], 3)
        pk = o["fp64"]["plane_kernel"]
        line["plane_kernel_128cubed"] = {
            "fp64": {"kind": {1: "register columns", 2: "LDS plane"}.get(pk["kind"], pk["kind"]),
                     "transform_pair_ms_register": round(pk["ms_register"], 4),
                     "transform_pair_ms_lds": round(pk["ms_lds"], 4)}}
        line["mixed_128cubed"] = {"steps_per_s": round(o["fp64"]["steps_per_s"], 3),
                                  "note": "the 3D psi solve is a spectral divide: the mixed form has no FP32 "
                                          "stage in 3D and runs the FP64 step"}
        line["fp32_fields_128cubed"] = {"steps_per_s": round(o["fp32"]["steps_per_s"], 3),
                                        "max_rel_err_vs_fp64": o["fp32_err"]}
        line["energy_monotone_128cubed"] = {"fp64": o["fp64"]["energy_monotone"],
                                            "fp32": o["fp32"]["energy_monotone"]}
    if a.only == "all":
        ens = run_ensemble(jv, a.ensemble_steps, jv.SAV_PREC_FP64)
        ensg = run_ensemble(jv, a.ensemble_steps, jv.SAV_PREC_FP64, threads_list=(1, 16), graph=True, batched=False)
        line["ensemble_64_fp64"] = {"instances": ens["instances"], "steps": ens["steps"],
                                    "sequential_steps_per_s": round(ens["by_threads"][1]["steps_per_s"], 2),
                                    "best_steps_per_s": round(ens["by_threads"][ens["best_threads"]]["steps_per_s"], 2),
                                    "best_threads": ens["best_threads"],
                                    "by_threads": {k: round(v["steps_per_s"], 2) for k, v in ens["by_threads"].items()},
                                    "speedup_vs_sequential": round(ens["speedup_vs_sequential"], 3),
                                    "algorithmic_tb_per_s_by_threads": {
                                        k: round(v["algorithmic_tb_per_s"], 3) for k, v in ens["by_threads"].items()},
                                    "built_kernels_tb_per_s_by_threads": {
                                        k: round(v["built_kernels_tb_per_s"], 3) for k, v in ens["by_threads"].items()},
                                    "byte_share_by_grid": ens["byte_share_by_grid"],
                                    "graph_replay_by_threads": {k: round(v["steps_per_s"], 2)
                                                                for k, v in ensg["by_threads"].items()}}
        bt = ens["batched"]
        line["ensemble_64_fp64"]["batched_size_classes"] = {
            "steps_per_s": round(bt["steps_per_s"], 2), "groups": bt["groups"],
            "grouped_members": bt["grouped_members"], "algorithmic_tb_per_s": round(bt["algorithmic_tb_per_s"], 3),
            "built_kernels_tb_per_s": round(bt["built_kernels_tb_per_s"], 3),
            "speedup_vs_best_threads": round(bt["steps_per_s"] / ens["by_threads"][ens["best_threads"]]["steps_per_s"], 3)}
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(jv, wl_head if "2d" in out else jv.workload_c4(n=128, steps=1),
                                            nsteps=a.cpu_steps if "2d" in out else 2)
        line["cpu_baseline"]["host"] = _host_info()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
