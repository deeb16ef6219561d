#!/usr/bin/env python3
"""Benchmark of the cahn-hilliard time-stepping path on one MI355X.

Headline (`value`): Landau 2D N=1024, explicit Euler, FP64 fields.  A bench
"step" is one print interval of the reference's run loop (src/main.cpp:104-129):
`--inner` (default 100) Euler time steps followed by the FP64 mass /
free-energy reductions, as Manager::run does every print_every steps.  The
field stays resident in HBM for the whole run.

value = cells * inner * steps / wall seconds (Gcell-updates/s), timed with a
device synchronisation on both sides.  The JSON line also carries
  roofline      the dominant kernel's algorithmic HBM rate (16 B per
                cell-step: read psi + write psi, FP64) over its average launch
                duration, timed with HIP events on the library's own stream;
  cpu_baseline  the oracle's single-threaded FP64 Euler (the reference's loop
                order, built -O3 -ffast-math) on a bounded sample of the same
                workload on this host;
  extra         the other configurations of the path: FP32 Euler, the
                pseudospectral integrator (FP64 / FP32) at N=1024, the batch of
                256 mixed-size FP32 replicas, the forward + adjoint / forward
                cost at N=256, and SimpleWertheim 3D N=128 euler_mobility.
`--workload NAME` runs one of them alone (for profiling).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0


def load_package():
    import importlib.util
    if "cahn_hilliard_amd" in sys.modules:
        return sys.modules["cahn_hilliard_amd"]
    pkg = os.path.join(REPO, "cahn-hilliard_amd")
    spec = importlib.util.spec_from_file_location("cahn_hilliard_amd", os.path.join(pkg, "__init__.py"),
                                                  submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["cahn_hilliard_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def sync(ctx):
    import torch
    ctx.synchronize()
    torch.cuda.synchronize()


def timed_run(ctx, sysm, steps, warmup, inner, observe=True):
    """`steps` print intervals of `inner` time steps each.  The FP64 mass and
    free energy of every interval are queued behind its steps and read one
    interval later, as a driver printing them would (same values, same order),
    so the device does not wait for the host between intervals; every pair is
    read inside the timed region."""
    def run(n):
        for _ in range(n):
            if observe:
                sysm.evolve_observe(inner)  # the steps and the queued reductions, one graph
                if sysm.observables_pending > 1:
                    sysm.fetch_observables()
            else:
                sysm.evolve(inner)
        while observe and sysm.observables_pending:
            sysm.fetch_observables()
    run(warmup)
    sync(ctx)
    t0 = time.perf_counter()
    run(steps)
    sync(ctx)
    return time.perf_counter() - t0


def kernel_time_per_launch(ctx, sysm, nsteps, steps_per_launch):
    """HIP events recorded on the stream the kernels run on."""
    import torch
    ext = torch.cuda.ExternalStream(ctx.stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    sysm.evolve(nsteps)  # warm graphs
    ctx.synchronize()
    e0.record(ext)
    sysm.evolve(nsteps)
    e1.record(ext)
    e1.synchronize()
    return e0.elapsed_time(e1) * 1e-3 / (nsteps / steps_per_launch)


def cpu_baseline(N, budget_s=12.0):
    """Oracle FP64 Euler (liboracle_fast.so, 1 thread) on the N x N field."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    p = O.Params(dims=2, N=N, dt=0.01, dx=1.0, k=1.0)
    m = O.Model(epsilon=0.9)
    rho = O.init_density(2, N, 0.0, 0.01, 0, 123456)
    t0 = time.perf_counter()
    rho = O.run(O.EULER, p, m, rho, 1, fast=True)
    one = time.perf_counter() - t0
    steps = max(1, int(budget_s / max(one, 1e-6)))
    t0 = time.perf_counter()
    O.run(O.EULER, p, m, rho, steps, fast=True)
    dt = time.perf_counter() - t0
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"value": round(N * N * steps / dt / 1e9, 5), "unit": "Gcell-updates/s", "cores": 1, "kind": "port",
            "sample": f"{steps} FP64 Euler steps of the Landau 2D N={N} field (oracle, -O3 -ffast-math, "
                      f"1 thread), {dt:.1f} s",
            "host": {"cpus_visible": os.cpu_count(), "cpus_usable": len(os.sched_getaffinity(0)), "cpu_model": model},
            "calibration": "the oracle's Euler matches the reference's own EulerCPU bit for bit and runs at "
                           "0.84-1.10 x its speed on one core (profiles/cpu_baseline_calibration.md)"}


# ------------------------------------------------------------------ workloads
def w_euler(ch, ctx, args, precision="fp64"):
    from cahn_hilliard_amd.init import initial_density
    N = args.N
    rho0 = initial_density(2, N, 0.0, 0.01, 0, 123456)
    desc = ch.make_desc(dims=2, N=N, integrator="euler", precision=precision, dt=0.01, dx=1.0, k=1.0, epsilon=0.9)
    s = ch.System(ctx, desc)
    s.upload(rho0)
    steps = args.steps if precision == "fp64" else max(10, args.steps // 4)
    wall = timed_run(ctx, s, steps, args.warmup, args.inner)
    spl = s.steps_per_launch
    shape = s.region_shape
    t_launch = kernel_time_per_launch(ctx, s, max(1024, 10 * args.inner), spl)
    mass, fe = s.observables()
    s.close()
    return dict(wall=wall, steps=steps, gcell=N * N * args.inner * steps / wall / 1e9, spl=spl, t_launch=t_launch,
                mass=mass, fe=fe, shape=shape)


def w_pseudospectral(ch, ctx, args, precision="fp64"):
    from cahn_hilliard_amd.init import initial_density
    N = args.N
    rho0 = initial_density(2, N, 0.0, 0.01, 0, 123456)
    d = ch.make_desc(dims=2, N=N, integrator="pseudospectral", precision=precision, dt=0.1, dx=1.0, k=1.0,
                     epsilon=0.9)
    s = ch.System(ctx, d)
    s.upload(rho0)
    nst, inner = max(10, args.steps // 4), 20
    w = timed_run(ctx, s, nst, 2, inner)
    s.close()
    return {"Gcell-updates/s": round(N * N * inner * nst / w / 1e9, 3),
            "us_per_time_step": round(w / (nst * inner) * 1e6, 3)}


def c3_replicas(ch, precision="fp32"):
    """256 Landau 2D replicas, 64 each at N = 64, 128, 256, 512: per replica r,
    drand48 seeded 123456 + r draws eps ~ U[0.5, 1], then c ~ U[-0.3, 0.3], then
    the cells psi0 = c + U[-0.01, 0.01] (x fastest)."""
    from cahn_hilliard_amd.init import replica_batch
    Ns, eps, fields = replica_batch()
    descs = [ch.make_desc(dims=2, N=N, integrator="euler", precision=precision, dt=0.01, dx=1.0, k=1.0, epsilon=e)
             for N, e in zip(Ns, eps)]
    return descs, fields


_C3_FP64 = {}


def c3_fp64(ch, ctx, marks):
    """The C3 replicas advanced by the FP64 batch, downloaded after each of `marks` steps."""
    if marks not in _C3_FP64:
        descs, fields = c3_replicas(ch, "fp64")
        b = ch.Batch(ctx, descs)
        for i, f in enumerate(fields):
            b.upload(i, f)
        out, done = {}, 0
        for t in marks:
            b.evolve(t - done)
            done = t
            out[t] = [b.download(i) for i in range(len(descs))]
        b.close()
        _C3_FP64[marks] = out
    return _C3_FP64[marks]


def w_batch(ch, ctx, args, compensation=None):
    """C3 as written: the 256 FP32 replicas advanced 5000 steps from their
    initial fields (timed: the steps, device-synchronised on both sides; the
    downloads after 1000 and 5000 steps are not), and the worst replica's
    relative L-inf error against the FP64 batch after 1000 and 5000 steps.
    `compensation`: None = the library's FP32 default (per-launch), "none",
    "low8" (ch_batch_create_ex)."""
    descs, fields = c3_replicas(ch)
    b = ch.Batch(ctx, descs, compensation=compensation)
    for i, f in enumerate(fields):
        b.upload(i, f)
    cells = sum(d.N * d.N for d in descs)
    shape, spl = b.autotune()  # untimed (it also runs by itself on the first evolve)
    b.evolve(64)  # graph capture, then start again from the initial fields
    for i, f in enumerate(fields):
        b.upload(i, f)
    marks = (1000, 5000)
    outs, done, wall = {}, 0, 0.0
    for t in marks:
        sync(ctx)
        t0 = time.perf_counter()
        b.evolve(t - done)
        sync(ctx)
        wall += time.perf_counter() - t0
        done = t
        outs[t] = [b.download(i) for i in range(len(descs))]
    mode = b.compensation
    b.close()
    ref = c3_fp64(ch, ctx, marks)
    errs = {t: max(float(np.abs(a - c).max() / np.abs(c).max()) for a, c in zip(outs[t], ref[t])) for t in marks}
    return {"Gcell-updates/s": round(cells * marks[-1] / wall / 1e9, 3), "replicas": len(descs), "cells": cells,
            "time_steps": marks[-1], "compensation": mode, "autotuned": {"shape": shape, "steps_per_launch": spl},
            "us_per_time_step": round(wall / marks[-1] * 1e6, 3),
            "fp32_rel_linf_vs_fp64": float(f"{errs[marks[-1]]:.3e}"), "error_after_steps": marks[-1],
            "fp32_rel_linf_vs_fp64_by_steps": {str(t): float(f"{e:.3e}") for t, e in errs.items()}}


def w_copy(ch, ctx, args):
    """Calibration for the HBM byte counters: 20 device copies of 256 MiB (one
    read and one write of 256 MiB per dispatch, well above the 256 MB MALL)."""
    import torch
    n = 256 << 20
    a = torch.ones(n // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(20):
        b.copy_(a)
    torch.cuda.synchronize()
    return {"bytes_per_copy": n, "copies": 20}


def fp32_error(ch, ctx, integrator, dt, N, steps):
    """max |psi_fp32 - psi_fp64| / max |psi_fp64| after `steps` from the same start."""
    from cahn_hilliard_amd.init import initial_density
    rho0 = initial_density(2, N, 0.0, 0.01, 0, 123456)
    res = []
    for prec in ("fp32", "fp64"):
        s = ch.System(ctx, ch.make_desc(dims=2, N=N, integrator=integrator, precision=prec, dt=dt, dx=1.0, k=1.0,
                                        epsilon=0.9))
        s.upload(rho0)
        s.evolve(steps)
        res.append(s.download())
        s.close()
    return float(f"{np.abs(res[0] - res[1]).max() / np.abs(res[1]).max():.3e}")


def w_adjoint(ch, ctx, args):
    """C4 as written: Landau 2D N=256 Euler, dt = 0.01, T = 5000 steps from
    psi0 = U[-0.01, 0.01) (drand48 seed 123456); loss = sum over the radial
    shells of (S(q) - S_target(q))^2 on the final field (utils/sq_from_traj.py:72-103,
    ch_sys_sq_loss), S_target from the same start at eps = 0.8, k = 1.0; the
    gradients wrt eps, k and psi0 at eps = 0.9, k = 1.2, checkpoints every 100
    steps, windows recomputed.  Timed (device-synchronised, best of 3): the
    forward alone (T steps) against record + loss and seed + reverse sweep.  In
    FP64 and FP32; the FP32 gradients are compared with the FP64 ones."""
    from cahn_hilliard_amd.init import initial_density
    N, T, ckpt = 256, 5000, 100
    rho0 = initial_density(2, N, 0.0, 0.01, 0, 123456)
    st = ch.System(ctx, ch.make_desc(dims=2, N=N, integrator="euler", dt=0.01, dx=1.0, k=1.0, epsilon=0.8))
    st.upload(rho0)
    st.evolve(T)
    target = st.structure_factor()
    st.close()
    res, grads = {}, {}
    for prec in ("fp64", "fp32"):
        s = ch.System(ctx, ch.make_desc(dims=2, N=N, integrator="euler", precision=prec, dt=0.01, dx=1.0, k=1.2,
                                        epsilon=0.9))
        s.upload(rho0)
        s.evolve(T)  # graphs
        t_fwd = float("inf")
        for _ in range(3):
            s.upload(rho0)
            sync(ctx)
            t0 = time.perf_counter()
            s.evolve(T)
            sync(ctx)
            t_fwd = min(t_fwd, time.perf_counter() - t0)
        # an untimed first call allocates the checkpoints, tapes and the
        # recompute stream (set-up, like an FFT plan)
        s.upload(rho0)
        s.record(T, ckpt)
        s.adjoint(s.sq_loss(target)[1])
        t_both = float("inf")
        for _ in range(3):
            s.upload(rho0)
            sync(ctx)
            t0 = time.perf_counter()
            s.record(T, ckpt)
            loss, seed = s.sq_loss(target)
            d0, g = s.adjoint(seed)
            t_both = min(t_both, time.perf_counter() - t0)
        s.close()
        grads[prec] = (d0, g)
        res[prec] = {"fwd_plus_adjoint_over_fwd": round(t_both / t_fwd, 3), "fwd_ms": round(t_fwd * 1e3, 3),
                     "fwd_plus_adjoint_ms": round(t_both * 1e3, 3), "loss": float(f"{loss:.9e}"),
                     "dL_depsilon": float(f"{g[ch.GRAD_EPSILON]:.9e}"), "dL_dk": float(f"{g[ch.GRAD_K]:.9e}")}
    (d64, g64), (d32, g32) = grads["fp64"], grads["fp32"]
    res["fp32"]["rel_err_vs_fp64"] = {
        "dL_dpsi0": float(f"{np.abs(d32 - d64).max() / np.abs(d64).max():.3e}"),
        "dL_depsilon": float(f"{abs(g32[ch.GRAD_EPSILON] / g64[ch.GRAD_EPSILON] - 1):.3e}"),
        "dL_dk": float(f"{abs(g32[ch.GRAD_K] / g64[ch.GRAD_K] - 1):.3e}")}
    return {"N": N, "time_steps": T, "checkpoint_every": ckpt, "fwd_plus_adjoint_over_fwd": res["fp64"][
        "fwd_plus_adjoint_over_fwd"], "fp64": res["fp64"], "fp32": res["fp32"],
            "timing": "best of 3; forward + adjoint = record + S(q) loss and its gradient + reverse sweep, "
                      "seed and dL/dpsi0 through host memory"}


def w_wertheim(ch, ctx, args):
    """C5 shape: SimpleWertheim 3D N=128 euler_mobility, valence 4, B2 = 1,
    delta = 1e4, k = 1, dt = 1e-4, M = rho / 0.05, FP64; the 8 replicas run one
    after the other (seeds 123456..123463).  Then, per replica, dF(psi_T)/d delta:
    record, the free-energy gradient as the seed, the reverse sweep, plus the
    explicit dF/d delta.  The record keeps the full tape (every state with its
    mu and mobility, 400 x 3 fields = 20 GB of HBM), so the reverse sweep
    recomputes nothing; the same timing with checkpoints every 50 steps and
    window recompute is reported beside it.  The configuration's 2000 steps
    drive rho below 0 near step 450 (f' = ln rho then fails: the oracle's FP64
    restatement of EulerMobilityCPU gives the same values and the same NaN),
    so the workload runs 400 steps."""
    from cahn_hilliard_amd.init import initial_density
    N, nst = 128, 400
    d = ch.make_desc(dims=3, N=N, integrator="euler_mobility", free_energy="simple_wertheim", B2=1.0, delta=1e4,
                     valence=4, mobility="linear", M=1.0, rho_ref=0.05, dt=1e-4, dx=1.0, k=1.0)
    tot = 0.0
    tot_grad = {0: 0.0, 50: 0.0}  # checkpoint spacing (0: full tape) -> time
    dFdd = []
    s = ch.System(ctx, d)
    # untimed set-up: graphs, checkpoints, tapes and the recompute stream
    s.upload(initial_density(3, N, 0.05, 0.05, 0, 123456))
    s.evolve(nst)
    for ck in tot_grad:
        s.upload(initial_density(3, N, 0.05, 0.05, 0, 123456))
        s.record(nst, ck)
        s.adjoint(s.free_energy_grad()[1])
    for r in range(8):
        rho0 = initial_density(3, N, 0.05, 0.05, 0, 123456 + r)
        s.upload(rho0)
        sync(ctx)
        t0 = time.perf_counter()
        s.evolve(nst)
        sync(ctx)
        tot += time.perf_counter() - t0
        for ck in tot_grad:
            s.upload(rho0)
            sync(ctx)
            t0 = time.perf_counter()
            s.record(nst, ck)
            _, g, dp = s.free_energy_grad()
            _, grads = s.adjoint(g)
            tot_grad[ck] += time.perf_counter() - t0
            if ck == 0:
                dFdd.append(float(dp[ch.GRAD_DELTA] + grads[ch.GRAD_DELTA]))
    s.close()
    return {"Gcell-updates/s": round(8 * N ** 3 * nst / tot / 1e9, 3), "replicas": 8, "time_steps": nst,
            "dF_ddelta": [float(f"{v:.6e}") for v in dFdd],
            "fwd_plus_adjoint_over_fwd": round(tot_grad[0] / tot, 3), "tape": "full (no recompute)",
            "fwd_plus_adjoint_over_fwd_ckpt50": round(tot_grad[50] / tot, 3)}


def w_noise(ch, ctx, args):
    """euler_mobility with conserved noise (mobility.with_noise, Philox normals),
    FP64, one launch per time step: Landau 2D N=1024 (constant M = 1, dt = 0.01,
    noise_rescale_factor = 0.01) and the C5 shape SimpleWertheim 3D N=128
    (M = rho / 0.05, dt = 1e-4, noise_rescale_factor = 1e-3).  Algorithmic HBM
    bytes: one read and one write of the field per time step (16 B / cell)."""
    from cahn_hilliard_amd.init import initial_density
    out = {}
    shapes = {
        "landau2d_N1024": (dict(dims=2, N=1024, integrator="euler_mobility", free_energy="landau", epsilon=0.9,
                                mobility="constant", M=1.0, dt=0.01, dx=1.0, k=1.0, with_noise=True,
                                noise_rescale_factor=0.01, seed=123456), initial_density(2, 1024, 0.0, 0.01, 0, 123456)),
        "wertheim3d_N128": (dict(dims=3, N=128, integrator="euler_mobility", free_energy="simple_wertheim", B2=1.0,
                                 delta=1e4, valence=4, mobility="linear", M=1.0, rho_ref=0.05, dt=1e-4, dx=1.0, k=1.0,
                                 with_noise=True, noise_rescale_factor=1e-3, seed=123456),
                            initial_density(3, 128, 0.05, 0.05, 0, 123456)),
    }
    for name, (kw, rho0) in shapes.items():
        # the 3D shape stays below the step (~450) where its rho turns negative
        warm, nst = (128, 640 if args.steps >= 100 else 128) if kw["dims"] == 2 else (64, 256)
        s = ch.System(ctx, ch.make_desc(**kw))
        s.upload(rho0)
        s.evolve(warm)  # graph capture + warm-up
        sync(ctx)
        t0 = time.perf_counter()
        s.evolve(nst)
        sync(ctx)
        dt = time.perf_counter() - t0
        mass0 = float(np.sum(rho0))
        rho = s.download()
        s.close()
        cells = kw["N"] ** kw["dims"]
        out[name] = {"Gcell-updates/s": round(cells * nst / dt / 1e9, 3), "us_per_time_step": round(dt / nst * 1e6, 3),
                     "effective_GB/s": round(16.0 * cells * nst / dt / 1e9, 1), "time_steps": nst,
                     "mass_rel_drift": float(f"{abs(float(np.sum(rho)) - mass0) / max(abs(mass0), 1.0):.3e}"),
                     "finite": bool(np.isfinite(rho).all())}
    return out


def w_psm(ch, ctx, args):
    """pseudospectral_mobility, Landau 2D N=1024 FP64 (eps = 0.9, k = 1, dx = 1,
    psi0 = 0.5 + U[-0.15, 0.15], regularised mobility M psi / (psi + 0.2)):
    the splitting scheme (dt = 0.02, S = 0.5, dealiased; one captured chain per
    64 steps) and the GMRES scheme (dt = 0.05, gmres_tol = 1e-10)."""
    N = 1024
    rng = np.random.default_rng(123456)
    rho0 = 0.5 + 0.15 * (2 * rng.random(N * N) - 1)
    base = dict(dims=2, N=N, integrator="pseudospectral_mobility", free_energy="landau", epsilon=0.9, k=1.0, dx=1.0,
                mobility="regularised", M=1.0, rho_min=0.2, S=0.5, use_dealias=True)
    out = {}
    s = ch.System(ctx, ch.make_desc(dt=0.02, **base))
    s.upload(rho0)
    s.evolve(64)
    sync(ctx)
    nst = 256 if args.steps >= 100 else 64
    t0 = time.perf_counter()
    s.evolve(nst)
    sync(ctx)
    dt = time.perf_counter() - t0
    rho = s.download()
    s.close()
    out["splitting"] = {"us_per_time_step": round(dt / nst * 1e6, 2), "Gcell-updates/s": round(N * N * nst / dt / 1e9, 3),
                        "time_steps": nst, "finite": bool(np.isfinite(rho).all()),
                        "mass_rel_drift": float(f"{abs(rho.sum() - rho0.sum()) / rho0.sum():.3e}")}
    s = ch.System(ctx, ch.make_desc(dt=0.05, use_gmres=True, gmres_tol=1e-10, **base))
    s.upload(rho0)
    s.evolve(1)
    sync(ctx)
    nst = 8 if args.steps >= 100 else 2
    its = []
    t0 = time.perf_counter()
    for _ in range(nst):
        s.evolve(1)
        its.append(s.gmres_iterations)
    sync(ctx)
    dt = time.perf_counter() - t0
    rho = s.download()
    s.close()
    out["gmres"] = {"ms_per_time_step": round(dt / nst * 1e3, 3), "iterations_per_step": its,
                    "ms_per_iteration": round(dt / max(1, sum(its)) * 1e3, 3), "finite": bool(np.isfinite(rho).all())}
    return out


def w_species(ch, ctx, args):
    """Multi-species Wertheim systems, 2D N=512 FP64, rho_s = 0.03 (1 + U[-0.1, 0.1]):
    SalehWertheim (3 species, valence 4/4/2, B2 = 0.7, delta_AA = 1e3,
    delta_BB = 2e3) with euler, dt = 1e-4, and GenericWertheim with the same
    topology (X by mass-action sweeps, warm-started per cell) with
    euler_mobility and free-energy mobility.  Rate = species-cell updates / s."""
    N = 512
    rng = np.random.default_rng(7)
    out = {}
    shapes = {
        "saleh_euler": (ch.saleh_model(1.0, valence=(4, 4, 2), B2=0.7, delta_AA=1e3, delta_BB=2e3), "euler",
                        "constant", "saleh"),
        "generic_em_free_energy": (ch.generic_model(1.0, [[0] * 4, [1] * 4, [0, 1]], {(0, 0): 1e3, (1, 1): 2e3},
                                                    {(a, b): 0.7 for a in range(3) for b in range(3)}),
                                   "euler_mobility", "free_energy", "generic_wertheim"),
    }
    nst = 256 if args.steps >= 100 else 64
    for name, (m, integ, mob, fe) in shapes.items():
        d = ch.make_desc(dims=2, N=N, integrator=integ, free_energy=fe, mobility=mob, M=1.0, dt=1e-4, dx=1.0)
        s = ch.System(ctx, d, species_model=m)
        rho0 = 0.03 * (1 + 0.1 * (2 * rng.random(3 * N * N) - 1))
        s.upload(rho0)
        s.evolve(64)
        sync(ctx)
        t0 = time.perf_counter()
        s.evolve(nst)
        sync(ctx)
        dt = time.perf_counter() - t0
        rho = s.download()
        s.close()
        out[name] = {"us_per_time_step": round(dt / nst * 1e6, 2),
                     "Gspecies-cell-updates/s": round(3 * N * N * nst / dt / 1e9, 3), "time_steps": nst,
                     "finite": bool(np.isfinite(rho).all())}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--inner", type=int, default=100)
    ap.add_argument("--N", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--workload", default=None,
                    choices=["euler", "euler_fp32", "pseudospectral", "pseudospectral_fp32", "batch", "batch_plain", "batch_low8",
                             "adjoint", "wertheim", "noise", "psm", "species", "copy"])
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: the library has no multi-GPU path (asked for {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    import torch  # device plumbing only (events on the library stream)
    ch = load_package()
    ctx = ch.Context(0)
    table = {"euler_fp32": lambda: w_euler(ch, ctx, args, "fp32"),
             "pseudospectral": lambda: w_pseudospectral(ch, ctx, args, "fp64"),
             "pseudospectral_fp32": lambda: w_pseudospectral(ch, ctx, args, "fp32"),
             "batch": lambda: w_batch(ch, ctx, args), "adjoint": lambda: w_adjoint(ch, ctx, args),
             "batch_plain": lambda: w_batch(ch, ctx, args, "none"),
             "batch_low8": lambda: w_batch(ch, ctx, args, "low8"),
             "wertheim": lambda: w_wertheim(ch, ctx, args), "noise": lambda: w_noise(ch, ctx, args),
             "psm": lambda: w_psm(ch, ctx, args), "species": lambda: w_species(ch, ctx, args),
             "copy": lambda: w_copy(ch, ctx, args)}
    if args.workload and args.workload != "euler":
        r = table[args.workload]()
        print(json.dumps({"workload": args.workload, "result": {k: v for k, v in r.items() if k != "wall"}}), flush=True)
        return

    r = w_euler(ch, ctx, args)
    N = args.N
    achieved = 16.0 * N * N * r["spl"] / r["t_launch"] / 1e9
    # HBM bytes per launch from the committed FETCH_SIZE / WRITE_SIZE runs of this
    # kernel configuration (profiles/hbm_traffic.json), null for any other
    cfg = r["shape"]
    traffic = None
    try:
        with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "hbm_traffic.json")) as f:
            rec = json.load(f).get(f"euler_fp64_N{N}_K{r['spl']}_cfg{cfg}")
        if rec:
            traffic = rec["hbm_bytes_per_launch"]
    except (OSError, ValueError):
        pass
    out = {
        "metric": "Gcell-updates/s & %HBM roofline, Landau 2D N=1024; fwd+adjoint/fwd cost",
        "value": round(r["gcell"], 3), "unit": "Gcell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(r["wall"] / args.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"Landau 2D N={N} periodic, explicit Euler, eps=0.9 k=1 dt=0.01 dx=1, "
                               f"psi0 = U[-0.01,0.01] (drand48 seed 123456); one step = {args.inner} time steps "
                               f"+ FP64 mass/free-energy reductions (queued, read one interval later)",
                   "N": N, "dims": 2, "integrator": "euler", "time_steps_per_step": args.inner},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_stencil2d_wave (fused Euler steps)", "time_steps_per_launch": r["spl"],
                     "avg_launch_us": round(r["t_launch"] * 1e6, 3)},
        "parity": {"mass": r["mass"], "free_energy": r["fe"]},
    }
    if not args.no_extra:
        extra = {}
        e32 = w_euler(ch, ctx, args, "fp32")
        extra["euler_fp32"] = {"Gcell-updates/s": round(e32["gcell"], 3),
                               "us_per_time_step": round(e32["wall"] / (e32["steps"] * args.inner) * 1e6, 3),
                               "fp32_rel_linf_vs_fp64": fp32_error(ch, ctx, "euler", 0.01, N, 1000),
                               "error_after_steps": 1000}
        # a failing extra workload is reported, never allowed to drop the headline line
        for name in ("pseudospectral", "pseudospectral_fp32", "batch", "batch_plain", "batch_low8", "adjoint", "wertheim",
                     "noise", "psm", "species"):
            try:
                extra[name] = table[name]()
            except Exception as e:  # noqa: BLE001
                extra[name] = {"error": f"{type(e).__name__}: {e}"}
        try:
            extra["pseudospectral_fp32"]["fp32_rel_linf_vs_fp64"] = fp32_error(ch, ctx, "pseudospectral", 0.1, N, 1000)
            extra["pseudospectral_fp32"]["error_after_steps"] = 1000
            for name in ("batch", "batch_plain", "batch_low8"):
                extra[name]["vs_fp64_single"] = round(extra[name]["Gcell-updates/s"] / r["gcell"], 3)
        except Exception as e:  # noqa: BLE001
            extra["error"] = f"{type(e).__name__}: {e}"
        out["extra"] = extra
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(N)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
