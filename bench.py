#!/usr/bin/env python3
"""Benchmark: Landau 2D N=1024, explicit Euler, FP64 fields, on one MI355X.

A bench "step" is one print interval of the reference's run loop
(src/main.cpp:104-129): `--inner` (default 100) Euler time steps followed by
the FP64 mass / free-energy reductions, as Manager::run does every
print_every steps.  The field stays resident in HBM for the whole run.

value = cells * inner * steps / wall seconds (Gcell-updates/s), timed with a
device synchronisation on both sides.  The JSON line also carries
  roofline      the dominant kernel's algorithmic HBM rate (16 B per
                cell-step: read psi + write psi, FP64) over its average launch
                duration, timed with HIP events on the library's own stream;
  cpu_baseline  the oracle's single-threaded FP64 Euler (the reference's
                loop order, built -O3 -ffast-math) on a bounded sample of the
                same workload on this host's cores;
  extra         the other hot-path configurations at N=1024 (FP32 Euler,
                pseudospectral FP64/FP32).
"""
import argparse
import json
import os
import sys
import time

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


def timed_run(ch, ctx, sysm, steps, warmup, inner):
    import torch
    for _ in range(warmup):
        sysm.evolve(inner)
        sysm.observables()
    ctx.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sysm.evolve(inner)
        sysm.observables()
    ctx.synchronize()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def kernel_time_per_launch(ctx, sysm, nsteps, steps_per_launch):
    """HIP events recorded on the stream the kernels run on."""
    import torch
    ext = torch.cuda.ExternalStream(ctx.stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ctx.synchronize()
    e0.record(ext)
    sysm.evolve(nsteps)
    e1.record(ext)
    e1.synchronize()
    launches = nsteps / steps_per_launch
    return e0.elapsed_time(e1) * 1e-3 / launches


def cpu_baseline(N, budget_s=12.0):
    """Oracle FP64 Euler (liboracle_fast.so, 1 thread) on the N x N field."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    import numpy as np
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
    return {"value": N * N * steps / dt / 1e9, "unit": "Gcell-updates/s", "cores": 1, "kind": "port",
            "sample": f"{steps} FP64 Euler steps of the Landau 2D N={N} field (oracle, -O3 -ffast-math, "
                      f"1 thread), {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--inner", type=int, default=100)
    ap.add_argument("--N", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: this build has no multi-GPU path (asked for {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    import torch  # device plumbing only (events on the library stream)
    ch = load_package()
    from cahn_hilliard_amd.init import initial_density

    N = args.N
    cells = N * N
    ctx = ch.Context(0)
    rho0 = initial_density(2, N, 0.0, 0.01, 0, 123456)
    desc = ch.make_desc(dims=2, N=N, integrator="euler", precision="fp64", dt=0.01, dx=1.0, k=1.0, epsilon=0.9)
    s = ch.System(ctx, desc)
    s.upload(rho0)
    wall = timed_run(ch, ctx, s, args.steps, args.warmup, args.inner)
    value = cells * args.inner * args.steps / wall / 1e9
    spl = s.steps_per_launch
    t_launch = kernel_time_per_launch(ctx, s, max(1000, 10 * args.inner), spl)
    bytes_per_launch = 16.0 * cells * spl
    achieved = bytes_per_launch / t_launch / 1e9
    mass, fe = s.observables()
    s.close()

    out = {
        "metric": "Gcell-updates/s & %HBM roofline, Landau 2D N=1024; fwd+adjoint/fwd cost",
        "value": round(value, 3), "unit": "Gcell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"Landau 2D N={N} periodic, explicit Euler, eps=0.9 k=1 dt=0.01 dx=1, "
                               f"psi0 = U[-0.01,0.01] (drand48 seed 123456); one step = {args.inner} time steps "
                               f"+ FP64 mass/free-energy reductions",
                   "N": N, "dims": 2, "integrator": "euler", "time_steps_per_step": args.inner},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_stencil (fused Euler step)", "time_steps_per_launch": spl,
                     "avg_launch_us": round(t_launch * 1e6, 3)},
        "parity": {"mass": mass, "free_energy": fe},
    }
    if not args.no_extra:
        extra = {}
        for name, integ, prec, dt, inner in [("euler_fp32", "euler", "fp32", 0.01, args.inner),
                                            ("pseudospectral_fp64", "pseudospectral", "fp64", 0.1, 20),
                                            ("pseudospectral_fp32", "pseudospectral", "fp32", 0.1, 20)]:
            d = ch.make_desc(dims=2, N=N, integrator=integ, precision=prec, dt=dt, dx=1.0, k=1.0, epsilon=0.9)
            sx = ch.System(ctx, d)
            sx.upload(rho0)
            nst = max(10, args.steps // 4)
            w = timed_run(ch, ctx, sx, nst, 2, inner)
            extra[name] = {"Gcell-updates/s": round(cells * inner * nst / w / 1e9, 3),
                           "us_per_time_step": round(w / (nst * inner) * 1e6, 3)}
            sx.close()
        out["extra"] = extra
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(N)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
