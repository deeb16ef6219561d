"""GPU euler_mobility with conserved noise (mobility.with_noise) against the oracle.

Both draw eta from Philox4x32-10 keyed by the seed with the counter (cell,
step), so the HIP path is compared with the oracle value by value:
  FP64: <= 1e-10 relative L-inf after 100 steps; total mass unchanged to
        1e-12 (relative to sum |rho_0|, at least 1).
  FP32: <= 2e-4 relative L-inf against the FP64 oracle after 100 steps (FP32
        normals from the same 32-bit words).
The run is split as 70 + 30 steps: a captured chain of 64 launches, its
remainder, then launches without a graph, so the device step index is
checked across graph replays and direct launches.  At N = 1024 the increment
from a uniform field is checked against its analytic variance
4 d M dt r^2 / dx^(d+2) and neighbour correlation -1/(2d).
"""
import numpy as np
import pytest

import oracle as O
from helpers import rel_linf

pytestmark = pytest.mark.gpu

LANDAU = dict(free_energy="landau", epsilon=0.9)
WERTHEIM = dict(free_energy="simple_wertheim", B2=1.0, delta=1e4, valence=4)

CASES = [
    ("landau2d_N64_constant", dict(dims=2, N=64, dt=0.01, dx=1.0, k=1.0, mobility="constant", M=1.0, **LANDAU),
     0.02, (0.0, 0.1)),
    ("landau2d_N32_constant_small_tiles", dict(dims=2, N=32, dt=0.01, dx=1.25, k=1.0, mobility="constant", M=0.7,
                                               **LANDAU), 0.01, (0.0, 0.1)),
    ("landau1d_N512_regularised", dict(dims=1, N=512, dt=0.01, dx=1.0, k=1.0, mobility="regularised", M=1.0,
                                       rho_min=0.1, **LANDAU), 0.02, (0.5, 0.1)),
    ("wertheim2d_N128_free_energy", dict(dims=2, N=128, dt=1e-4, dx=1.0, k=1.0, mobility="free_energy", M=1.0,
                                         **WERTHEIM), 0.01, (0.05, 0.05)),
    ("wertheim3d_N16_linear", dict(dims=3, N=16, dt=1e-4, dx=1.0, k=1.0, mobility="linear", M=1.0, rho_ref=0.05,
                                   **WERTHEIM), 0.01, (0.05, 0.05)),
    ("wertheim3d_N4_linear", dict(dims=3, N=4, dt=1e-4, dx=1.0, k=1.0, mobility="linear", M=1.0, rho_ref=0.05,
                                  **WERTHEIM), 0.01, (0.05, 0.05)),
]


@pytest.fixture(scope="module")
def ctx(ch):
    c = ch.Context(0)
    yield c
    c.close()


def oracle_of(kw):
    mob = {"constant": O.MOB_CONSTANT, "regularised": O.MOB_REGULARISED, "linear": O.MOB_LINEAR,
           "free_energy": O.MOB_FREE_ENERGY}[kw["mobility"]]
    p = O.Params(dims=kw["dims"], N=kw["N"], dt=kw["dt"], dx=kw["dx"], k=kw["k"], mobility=mob, M=kw["M"],
                 rho_min=kw.get("rho_min", 0.0), rho_ref=kw.get("rho_ref", 1.0))
    if kw["free_energy"] == "landau":
        m = O.Model(O.LANDAU, epsilon=kw["epsilon"])
    else:
        m = O.Model(O.SIMPLE_WERTHEIM, B2=kw["B2"], delta=kw["delta"], valence=kw["valence"])
    return p, m


def initial(kw, center, spread, seed=7):
    rng = np.random.default_rng(seed)
    n = kw["N"] ** kw["dims"]
    if center == 0.0:
        return spread * (2 * rng.random(n) - 1)
    return center * (1.0 + spread * (2 * rng.random(n) - 1))


def run_gpu(ch, ctx, kw, rho0, splits, rescale, seed, precision="fp64"):
    d = ch.make_desc(integrator="euler_mobility", precision=precision, with_noise=True,
                     noise_rescale_factor=rescale, seed=seed, **kw)
    s = ch.System(ctx, d)
    assert s.steps_per_launch == 1
    s.upload(rho0)
    for n in splits:
        s.evolve(n)
    out = s.download()
    s.close()
    return out


@pytest.mark.parametrize("name,kw,rescale,init", CASES, ids=[c[0] for c in CASES])
def test_noise_matches_oracle_fp64(ch, ctx, name, kw, rescale, init):
    print(f"case {name}", flush=True)
    rho0 = initial(kw, *init)
    p, m = oracle_of(kw)
    seed = 0x1234_5678_9ABC_DEF0
    want = O.run_noise(p, m, rho0, 100, noise_rescale=rescale, seed=seed)
    got = run_gpu(ch, ctx, kw, rho0, (70, 30), rescale, seed)
    err = rel_linf(got, want)
    print(f"{name}: fp64 rel L-inf vs oracle = {err:.3e}")
    assert err < 1e-10
    assert abs(got.sum() - rho0.sum()) < 1e-12 * max(1.0, np.abs(rho0).sum())
    # the noise moved the field away from the noiseless path
    plain = O.run(O.EULER_MOBILITY, p, m, rho0, 100)
    assert np.abs(want - plain).max() > 100 * np.abs(got - want).max()


@pytest.mark.parametrize("name,kw,rescale,init", CASES[:4], ids=[c[0] for c in CASES[:4]])
def test_noise_fp32_against_fp64_oracle(ch, ctx, name, kw, rescale, init):
    rho0 = initial(kw, *init)
    p, m = oracle_of(kw)
    want = O.run_noise(p, m, rho0, 100, noise_rescale=rescale, seed=99)
    got = run_gpu(ch, ctx, kw, rho0, (100,), rescale, 99, "fp32")
    err = rel_linf(got, want)
    print(f"{name}: fp32 rel L-inf vs fp64 oracle = {err:.3e}")
    assert err < 2e-4


def test_noise_seed_reproducible_and_split_invariant(ch, ctx):
    kw = CASES[0][1]
    rho0 = initial(kw, 0.0, 0.1)
    a = run_gpu(ch, ctx, kw, rho0, (200,), 0.02, 42)
    b = run_gpu(ch, ctx, kw, rho0, (1, 63, 64, 5, 67), 0.02, 42)
    c = run_gpu(ch, ctx, kw, rho0, (200,), 0.02, 43)
    assert np.array_equal(a, b)
    assert np.abs(a - c).max() > 1e-3


@pytest.mark.parametrize("dims,N", [(2, 1024), (3, 128)])
def test_noise_variance_at_scale(ch, ctx, dims, N):
    M, dt, dx, r = 1.3, 0.01, 0.7, 1.7
    kw = dict(dims=dims, N=N, dt=dt, dx=dx, k=1.0, mobility="constant", M=M, **LANDAU)
    rho0 = np.full(N ** dims, 0.3)
    d = run_gpu(ch, ctx, kw, rho0, (1,), r, 2024) - rho0
    expected = 4.0 * dims * M * dt * r * r / dx ** (dims + 2)
    print(f"dims {dims} N {N}: var / expected = {d.var() / expected:.5f}")
    assert abs(d.var() / expected - 1.0) < 0.01
    g = d.reshape((N,) * dims)
    for ax in range(dims):
        corr = np.mean(g * np.roll(g, 1, axis=ax)) / np.mean(g * g)
        assert abs(corr + 1.0 / (2 * dims)) < 0.01
    assert abs(d.sum()) < 1e-12 * d.size * np.sqrt(expected)


def test_noise_rejects_reverse_mode(ch, ctx):
    kw = CASES[0][1]
    d = ch.make_desc(integrator="euler_mobility", with_noise=True, noise_rescale_factor=0.02, seed=1, **kw)
    s = ch.System(ctx, d)
    s.upload(np.zeros(kw["N"] ** 2))
    with pytest.raises(ch.CHError):
        s.record(10, 0)
    s.close()
