#!/usr/bin/env python3
"""Write FP64 golden vectors from the reference's own CPU integrators.

Runs in the development container only (it needs oracle/_ref/ref_{1,2,3}D,
which `make -C oracle ref` compiles from the unmodified reference sources).
For every case in CASES it writes the initial field, builds the reference TOML
configuration, runs the reference's EulerCPU / EulerMobilityCPU (through
oracle/ref_harness.cpp) for 1 step and for `steps` steps, and stores
    tests/golden/<name>.npz   rho0, rho1, rhoT   (float64)
plus tests/golden/cases.json with the parameters.  The initial fields come
from glibc's srand48/drand48 through the reference's recipe
(src/CahnHilliard.cpp:133-151), computed here independently of the oracle.

Usage: python tests/golden/make_golden.py            (the small cases)
       python tests/golden/make_golden.py --large    (the full-size cases, ~2 min)
"""
import ctypes
import json
import math
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF_BIN = os.path.join(REPO, "oracle", "_ref")

CASES = [
    # name, integrator, dims, N, steps, init(density, A, seed), model, mobility, dt, k
    dict(name="euler_landau_1d", integrator="euler", dims=1, N=64, steps=1000, dt=0.01, k=1.0,
         init=(0.0, 0.3, 123456), model=dict(free_energy="landau", epsilon=0.9), mobility=dict(type="constant", M=1.0)),
    dict(name="euler_landau_2d", integrator="euler", dims=2, N=32, steps=1000, dt=0.01, k=1.0,
         init=(0.0, 0.3, 123456), model=dict(free_energy="landau", epsilon=0.9), mobility=dict(type="constant", M=1.0)),
    dict(name="euler_landau_2d_dx", integrator="euler", dims=2, N=16, steps=500, dt=0.02, k=0.7,
         init=(0.1, 0.4, 42), model=dict(free_energy="landau", epsilon=0.6), mobility=dict(type="constant", M=0.8),
         dx=1.3),
    dict(name="euler_landau_3d", integrator="euler", dims=3, N=16, steps=300, dt=0.001, k=1.0,
         init=(0.0, 0.3, 7), model=dict(free_energy="landau", epsilon=0.9), mobility=dict(type="constant", M=1.0)),
    dict(name="eulermob_landau_const_2d", integrator="euler_mobility", dims=2, N=32, steps=500, dt=0.01, k=1.0,
         init=(0.0, 0.3, 99), model=dict(free_energy="landau", epsilon=0.9), mobility=dict(type="constant", M=1.0)),
    dict(name="eulermob_landau_reg_2d", integrator="euler_mobility", dims=2, N=32, steps=500, dt=0.01, k=1.0,
         init=(0.5, 0.05, 5), model=dict(free_energy="landau", epsilon=0.9),
         mobility=dict(type="regularised", M=1.0, rho_min=0.1)),
    dict(name="eulermob_wertheim_fe_3d", integrator="euler_mobility", dims=3, N=16, steps=200, dt=1e-3, k=1.0,
         init=(0.05, 0.3, 123456), model=dict(free_energy="simple_wertheim", valence=4, B2=1.0, delta=10.0),
         mobility=dict(type="free_energy", M=100.0)),
    dict(name="eulermob_wertheim_const_2d", integrator="euler_mobility", dims=2, N=32, steps=300, dt=1e-4, k=1.0,
         init=(0.05, 0.05, 11), model=dict(free_energy="simple_wertheim", valence=4, B2=1.0, delta=1e4),
         mobility=dict(type="constant", M=1.0)),
    dict(name="euler_wertheim_reg_1d", integrator="euler", dims=1, N=64, steps=300, dt=1e-4, k=1.0,
         init=(0.05, 0.5, 3), model=dict(free_energy="simple_wertheim", valence=4, B2=1.0, delta=1e4,
                                         regularisation_delta=0.04),
         mobility=dict(type="constant", M=1.0)),
    dict(name="eulermob_wertheim_fe_1d", integrator="euler_mobility", dims=1, N=64, steps=300, dt=1e-3, k=1.0,
         init=(0.05, 0.05, 8), model=dict(free_energy="simple_wertheim", valence=2, B2=-3.0, delta=500.0),
         mobility=dict(type="free_energy", M=200.0)),
]


# Full-size configurations: the Landau 2D N=1024 Euler workload after 1000 steps
# (stored as every 8th cell in x and y plus sums over the whole field, to keep
# the fixture small) and the N=128 workload after 1000 and 10,000 steps (whole
# field).  Both start from psi0 = U[-0.01, 0.01) (initial_density 0,
# initial_A 0.01, seed 123456) with eps = 0.9, k = 1, dt = 0.01, dx = 1, M = 1.
LARGE_CASES = [
    dict(name="large_landau_2d_N1024", integrator="euler", dims=2, N=1024, marks=[1000], stride=8, dt=0.01,
         k=1.0, init=(0.0, 0.01, 123456), model=dict(free_energy="landau", epsilon=0.9),
         mobility=dict(type="constant", M=1.0)),
    dict(name="large_landau_2d_N128", integrator="euler", dims=2, N=128, marks=[1000, 10000], stride=1, dt=0.01,
         k=1.0, init=(0.0, 0.01, 123456), model=dict(free_energy="landau", epsilon=0.9),
         mobility=dict(type="constant", M=1.0)),
    # SimpleWertheim 3D N=128 euler_mobility with M = rho / 0.05 (the harness's
    # linear mobility through the reference's IMobility), rho0 = 0.05 (1 +
    # U[-0.05, 0.05)), 400 steps (rho turns negative near step 450 on these
    # inputs); every 4th cell in x, y and z
    dict(name="large_wertheim_3d_N128", integrator="euler_mobility", dims=3, N=128, marks=[400], stride=4, dt=1e-4,
         k=1.0, init=(0.05, 0.05, 123456), model=dict(free_energy="simple_wertheim", valence=4, B2=1.0, delta=1e4),
         mobility=dict(type="linear", M=1.0, rho_ref=0.05)),
]


def glibc_init(dims, N, density, A, seed):
    """src/CahnHilliard.cpp:133-151 with N_peaks = 0, using glibc's drand48."""
    libc = ctypes.CDLL("libc.so.6")
    libc.drand48.restype = ctypes.c_double
    libc.srand48.argtypes = [ctypes.c_long]
    libc.srand48(seed)
    rho = np.empty(N ** dims)
    for b in range(N ** dims):
        r = libc.drand48() - 0.5
        rho[b] = density * (1.0 + 2.0 * A * r) if density != 0.0 else 2.0 * A * r
    return rho


def toml_text(c):
    m = c["model"]
    lines = [f'integrator = "{c["integrator"]}"', f'free_energy = "{m["free_energy"]}"',
             f'N = {c["N"]}', f'dt = {c["dt"]!r}', f'dx = {c.get("dx", 1.0)!r}', f'k = {c["k"]!r}',
             f'steps = {c["steps"]}', 'initial_density = 0.0']
    lines.append("[mobility]")
    for key, v in c["mobility"].items():
        lines.append(f'{key} = "{v}"' if isinstance(v, str) else f"{key} = {v!r}")
    if m["free_energy"] == "landau":
        lines += ["[landau]", f'epsilon = {m["epsilon"]!r}']
    else:
        lines += ["[wertheim]", f'valence = {m["valence"]}', f'B2 = {m["B2"]!r}', f'delta = {m["delta"]!r}',
                  f'regularisation_delta = {m.get("regularisation_delta", 0.0)!r}']
    return "\n".join(lines) + "\n"


def run_ref(c, rho0, steps, tmp):
    cfg = os.path.join(tmp, "c.toml")
    with open(cfg, "w") as f:
        f.write(toml_text(c))
    inp, out = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
    rho0.astype("<f8").tofile(inp)
    exe = os.path.join(REF_BIN, f"ref_{c['dims']}D")
    subprocess.run([exe, cfg, inp, str(steps), out], check=True, capture_output=True)
    return np.fromfile(out, dtype="<f8")


def make_large(tmp):
    """LARGE_CASES -> tests/golden/<name>.npz: for each mark t the field after t
    steps at the cells whose coordinates are all multiples of `stride`
    (rho<t>_sub, x fastest), and over the whole field sum(rho<t>) and
    sum(rho<t>^2); plus
    tests/golden/large_cases.json."""
    meta = []
    for c in LARGE_CASES:
        N, st = c["N"], c["stride"]
        rho0 = glibc_init(c["dims"], N, *c["init"])
        out, prev, done = {}, rho0, 0
        for t in c["marks"]:
            cfg = dict(c, steps=t - done)
            prev = run_ref(cfg, prev, t - done, tmp)
            done = t
            assert np.all(np.isfinite(prev)), c["name"]
            f = prev.reshape((N,) * c["dims"])
            out[f"rho{t}_sub"] = f[(slice(None, None, st),) * c["dims"]].copy()
            out[f"sum{t}"] = np.array(prev.sum())
            out[f"sumsq{t}"] = np.array((prev * prev).sum())
            print(f"{c['name']} t={t}: max|rho|={np.abs(prev).max():.4g} sum={prev.sum():.6e}")
        np.savez_compressed(os.path.join(HERE, c["name"] + ".npz"), **out)
        meta.append({k: v for k, v in c.items()})
    with open(os.path.join(HERE, "large_cases.json"), "w") as f:
        json.dump(meta, f, indent=1)


def main():
    if not os.path.exists(os.path.join(REF_BIN, "ref_2D")):
        sys.exit("oracle/_ref is not built: run `make -C oracle ref`")
    if "--large" in sys.argv:
        with tempfile.TemporaryDirectory() as tmp:
            make_large(tmp)
        return
    meta = []
    with tempfile.TemporaryDirectory() as tmp:
        for c in CASES:
            rho0 = glibc_init(c["dims"], c["N"], *c["init"])
            rho1 = run_ref(c, rho0, 1, tmp)
            rhoT = run_ref(c, rho0, c["steps"], tmp)
            assert np.all(np.isfinite(rhoT)), c["name"]
            np.savez_compressed(os.path.join(HERE, c["name"] + ".npz"), rho0=rho0, rho1=rho1, rhoT=rhoT)
            meta.append(dict(c, dx=c.get("dx", 1.0)))
            print(f"{c['name']}: |rho0|max={np.abs(rho0).max():.3g} |rhoT-rho0|max={np.abs(rhoT - rho0).max():.3g}")
    with open(os.path.join(HERE, "cases.json"), "w") as f:
        json.dump(meta, f, indent=1)


if __name__ == "__main__":
    main()
