"""Initial fields with the reference's recipe (src/CahnHilliard.cpp:133-151).

srand48(seed) (src/main.cpp:36), then one drand48() draw per bin, x fastest:
    rho = density (1 + 2 A cos(k0 bin) r)   or   2 A cos(k0 bin) r   (density == 0)
    r = drand48() - 0.5 (initial_N_peaks == 0) else 1 + 0.02 (drand48() - 0.5)
drand48 is the POSIX 48-bit LCG x' = (0x5DEECE66D x + 0xB) mod 2^48.  The stream
is generated with numpy in blocks of B values using the B-step jump
x_{n+B} = a^B x_n + c (a^B - 1)/(a - 1) mod 2^48; uint64 products wrap mod 2^64,
which 2^48 divides, so masking gives the exact residue.
"""
import math

import numpy as np

_A, _C, _MASK = 0x5DEECE66D, 0xB, (1 << 48) - 1


def drand48_stream(seed, n, block=4096):
    x = (((seed & 0xFFFFFFFF) << 16) | 0x330E) & _MASK
    first = np.empty(min(block, n), dtype=np.uint64)
    for i in range(first.size):
        x = (_A * x + _C) & _MASK
        first[i] = x
    out = np.empty(n, dtype=np.uint64)
    out[:first.size] = first
    if n > block:
        aB, cB = 1, 0
        for _ in range(block):  # compose x -> a x + c, block times
            aB, cB = (_A * aB) & _MASK, (_A * cB + _C) & _MASK
        aB, cB, mask = np.uint64(aB), np.uint64(cB), np.uint64(_MASK)
        cur = first
        pos = block
        while pos < n:
            cur = (cur * aB + cB) & mask
            m = min(block, n - pos)
            out[pos:pos + m] = cur[:m]
            pos += m
    return out.astype(np.float64) / float(1 << 48)


def initial_density(dims, N, density=0.0, A=1e-2, N_peaks=0, seed=123456):
    bins = N ** dims
    d = drand48_stream(seed, bins)
    r = (d - 0.5) if N_peaks == 0 else 1.0 + 0.02 * (d - 0.5)
    k0 = 2 * math.pi * N_peaks / N
    modulation = A * np.cos(k0 * np.arange(bins, dtype=np.float64))
    if density != 0.0:
        return density * (1.0 + 2.0 * modulation * r)
    return 2.0 * modulation * r


def replica_batch(sizes=(64, 128, 256, 512), per_size=64, seed=123456):
    """The mixed-size Landau 2D replica workload: `per_size` replicas at each N
    of `sizes`; replica r draws from drand48 seeded `seed` + r: eps ~ U[0.5, 1],
    then c ~ U[-0.3, 0.3], then the cells psi0 = c + U[-0.01, 0.01] (x fastest).
    Returns (Ns, eps, fields)."""
    Ns, eps, fields = [], [], []
    for i, N in enumerate([n for n in sizes for _ in range(per_size)]):
        r = drand48_stream(seed + i, 2 + N * N)
        Ns.append(N)
        eps.append(0.5 + 0.5 * float(r[0]))
        fields.append(-0.3 + 0.6 * r[1] + 0.02 * (r[2:] - 0.5))
    return Ns, eps, fields


def write_replicas(directory, Ns, eps, fields, k=1.0, dt=0.01, steps=5000, print_every=100):
    """Writes replica r of a batch as `replica_<r>.dat`, in the native text
    format that `load_from` reads (src/CahnHilliard.cpp:72-132: a '#' header
    line, then one line of N values per row y, x fastest; 17 significant digits
    so the state is exact), and `replica_<r>.toml`, an input of the reference's
    driver for the same system (Landau 2D euler, epsilon = eps[r], k, dt, M = 1,
    dx = 1, load_from = the .dat file).  Returns the TOML paths."""
    import os
    os.makedirs(directory, exist_ok=True)
    tomls = []
    for r, (N, e, f) in enumerate(zip(Ns, eps, fields)):
        f = np.asarray(f, dtype=np.float64).reshape(N, N)
        dat = os.path.join(directory, f"replica_{r}.dat")
        with open(dat, "w") as out:
            out.write(f"# step = 0, t = 0, size = {N}x{N}x1 | density\n")
            for row in f:
                out.write(" ".join(repr(float(v)) for v in row) + " \n")
        toml = os.path.join(directory, f"replica_{r}.toml")
        with open(toml, "w") as out:
            out.write(f"steps = {steps}\nprint_every = {print_every}\nN = {N}\ndt = {dt}\ndx = 1.0\nk = {k}\nM = 1.0\n"
                      f'free_energy = "landau"\nintegrator = "euler"\nload_from = "{os.path.abspath(dat)}"\n'
                      f"\n[landau]\nepsilon = {e!r}\n")
        tomls.append(toml)
    return tomls

