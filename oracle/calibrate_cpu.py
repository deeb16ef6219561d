#!/usr/bin/env python3
"""Calibrates the CPU baseline that bench.py times (the oracle's FP64 Euler,
liboracle_fast.so: -O3 -ffast-math, one thread) against the reference's own
EulerCPU (oracle/_ref/ref_2D, compiled from the unmodified sources with the
reference's release flags), on the Landau 2D workload of the benchmark
(eps 0.9, k 1, dt 0.01, dx 1, psi0 = U[-0.01, 0.01) from drand48 seed 123456).

TEST INFRASTRUCTURE ONLY (development container: it needs oracle/_ref).
For each N it runs both for the same number of steps, reports the largest
difference between their fields and each one's cell updates per second (the
reference's process start-up and TOML parsing are measured with a 0-step run
and subtracted).

Usage: python oracle/calibrate_cpu.py [N ...]
"""
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import oracle as O  # noqa: E402

TOML = """integrator = "euler"
free_energy = "landau"
N = {N}
dt = 0.01
dx = 1.0
k = 1.0
initial_density = 0.0
[mobility]
type = "constant"
M = 1.0
[landau]
epsilon = 0.9
"""


def ref_run(N, rho0, steps, tmp):
    cfg, inp, out = (os.path.join(tmp, n) for n in ("c.toml", "in.bin", "out.bin"))
    with open(cfg, "w") as f:
        f.write(TOML.format(N=N))
    rho0.astype("<f8").tofile(inp)
    t0 = time.perf_counter()
    subprocess.run([os.path.join(HERE, "_ref", "ref_2D"), cfg, inp, str(steps), out], check=True)
    return time.perf_counter() - t0, np.fromfile(out, dtype="<f8")


def main():
    Ns = [int(a) for a in sys.argv[1:]] or [256, 1024]
    for N in Ns:
        steps = max(20, int(4e8 / (N * N)))
        rho0 = O.init_density(2, N, 0.0, 0.01, 0, 123456)
        with tempfile.TemporaryDirectory() as tmp:
            t_start, _ = ref_run(N, rho0, 0, tmp)
            t_ref, r_ref = ref_run(N, rho0, steps, tmp)
        p, m = O.Params(dims=2, N=N, dt=0.01, dx=1.0, k=1.0), O.Model(epsilon=0.9)
        O.run(O.EULER, p, m, rho0, 1, fast=True)
        t0 = time.perf_counter()
        r_or = O.run(O.EULER, p, m, rho0, steps, fast=True)
        t_or = time.perf_counter() - t0
        ref_rate = N * N * steps / max(t_ref - t_start, 1e-9) / 1e6
        or_rate = N * N * steps / t_or / 1e6
        print(f"N={N}, {steps} steps: max |oracle - reference| = {np.abs(r_or - r_ref).max():.3e}; "
              f"reference {ref_rate:.1f} Mcell/s, oracle {or_rate:.1f} Mcell/s (ratio {or_rate / ref_rate:.3f})",
              flush=True)


if __name__ == "__main__":
    main()
