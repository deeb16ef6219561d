"""GPU pseudospectral_mobility (PseudospectralMobilityCPU.cpp) against the oracle.

Tolerances (stated per result):
  splitting scheme, FP64: <= 1e-10 relative L-inf after 100 steps (a captured
      chain of 64 steps, then direct steps); mass drift <= 1e-12;
  GMRES scheme, FP64: both sides solve to gmres_tol = 1e-13, <= 1e-10 relative
      L-inf after 5 steps; iteration counts within 2 of the oracle's (inner
      products are summed in another order);
  FP32 fields against the FP64 oracle: <= 2e-4 (splitting, 100 steps), <= 1e-4
      (GMRES solved to 1e-6, 5 steps).
"""
import numpy as np
import pytest

import oracle as O
from helpers import rel_linf

pytestmark = pytest.mark.gpu

LANDAU = dict(free_energy="landau", epsilon=0.9)
WERTHEIM = dict(free_energy="simple_wertheim", B2=1.0, delta=1e4, valence=4)

CASES = [
    ("landau1d_N256_regularised_dealias", dict(dims=1, N=256, dt=0.02, dx=1.0, k=1.0, mobility="regularised", M=1.0,
                                               rho_min=0.2, S=0.5, use_dealias=True, **LANDAU), (0.5, 0.3)),
    ("landau2d_N64_regularised", dict(dims=2, N=64, dt=0.02, dx=1.0, k=1.0, mobility="regularised", M=1.0,
                                      rho_min=0.2, S=0.5, use_dealias=False, **LANDAU), (0.5, 0.3)),
    ("landau2d_N128_regularised_dx", dict(dims=2, N=128, dt=0.01, dx=0.8, k=1.2, mobility="regularised", M=0.7,
                                          rho_min=0.3, S=0.2, use_dealias=True, **LANDAU), (0.5, 0.3)),
    ("landau2d_N64_constant", dict(dims=2, N=64, dt=0.1, dx=1.0, k=1.0, mobility="constant", M=1.0, S=0.0,
                                   use_dealias=True, **LANDAU), (0.0, 0.02)),
    ("wertheim2d_N64_linear", dict(dims=2, N=64, dt=3e-4, dx=1.0, k=1.0, mobility="linear", M=1.0, rho_ref=0.05,
                                   S=0.0, use_dealias=True, **WERTHEIM), (0.05, 0.05)),
    ("wertheim3d_N16_free_energy", dict(dims=3, N=16, dt=1e-3, dx=1.0, k=1.0, mobility="free_energy", M=1.0,
                                        S=0.0, use_dealias=False, **WERTHEIM), (0.05, 0.05)),
    ("landau3d_N32_linear", dict(dims=3, N=32, dt=0.01, dx=1.0, k=1.0, mobility="linear", M=1.0, rho_ref=1.0,
                                 S=0.3, use_dealias=True, **LANDAU), (0.5, 0.3)),
]
GMRES_CASES = [
    ("landau1d_N64_regularised", dict(dims=1, N=64, dt=0.2, dx=0.9, k=1.0, mobility="regularised", M=1.0, rho_min=0.2,
                                      S=0.3, use_dealias=False, **LANDAU), (0.5, 0.3)),
    ("landau2d_N32_regularised", dict(dims=2, N=32, dt=0.05, dx=1.0, k=1.0, mobility="regularised", M=1.0,
                                      rho_min=0.2, S=0.5, use_dealias=False, **LANDAU), (0.5, 0.3)),
    ("landau2d_N32_regularised_dealias", dict(dims=2, N=32, dt=0.05, dx=1.0, k=1.0, mobility="regularised", M=1.0,
                                              rho_min=0.2, S=0.5, use_dealias=True, **LANDAU), (0.5, 0.3)),
    ("wertheim2d_N32_linear", dict(dims=2, N=32, dt=3e-4, dx=1.0, k=1.0, mobility="linear", M=1.0, rho_ref=0.05,
                                   S=0.0, use_dealias=True, **WERTHEIM), (0.05, 0.05)),
    ("landau3d_N8_regularised", dict(dims=3, N=8, dt=0.05, dx=1.0, k=1.0, mobility="regularised", M=1.0,
                                     rho_min=0.2, S=0.3, use_dealias=False, **LANDAU), (0.5, 0.3)),
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
                 rho_min=kw.get("rho_min", 0.0), rho_ref=kw.get("rho_ref", 1.0), S=kw["S"],
                 use_dealias=int(kw["use_dealias"]))
    if kw["free_energy"] == "landau":
        m = O.Model(O.LANDAU, epsilon=kw["epsilon"])
    else:
        m = O.Model(O.SIMPLE_WERTHEIM, B2=kw["B2"], delta=kw["delta"], valence=kw["valence"])
    return p, m


def initial(kw, center, spread, seed=3):
    rng = np.random.default_rng(seed)
    n = kw["N"] ** kw["dims"]
    if center == 0.0:
        return spread * (2 * rng.random(n) - 1)
    return center * (1.0 + spread * (2 * rng.random(n) - 1))


def run_gpu(ch, ctx, kw, rho0, splits, precision="fp64", iters=None, **extra):
    d = ch.make_desc(integrator="pseudospectral_mobility", precision=precision, **kw, **extra)
    s = ch.System(ctx, d)
    assert s.steps_per_launch == 1
    s.upload(rho0)
    for n in splits:
        if iters is None:
            s.evolve(n)
        else:
            for _ in range(n):
                s.evolve(1)
                iters.append(s.gmres_iterations)
    out = s.download()
    s.close()
    return out


@pytest.mark.parametrize("name,kw,init", CASES, ids=[c[0] for c in CASES])
def test_splitting_matches_oracle_fp64(ch, ctx, name, kw, init):
    print(f"case {name}", flush=True)
    rho0 = initial(kw, *init)
    p, m = oracle_of(kw)
    want = O.run_psm(p, m, rho0, 100)
    got = run_gpu(ch, ctx, kw, rho0, (70, 30))
    err = rel_linf(got, want)
    print(f"{name}: splitting fp64 rel L-inf vs oracle = {err:.3e}")
    assert err < 1e-10
    assert abs(got.sum() - rho0.sum()) < 1e-12 * max(1.0, np.abs(rho0).sum())


@pytest.mark.parametrize("name,kw,init", GMRES_CASES, ids=[c[0] for c in GMRES_CASES])
def test_gmres_matches_oracle_fp64(ch, ctx, name, kw, init):
    print(f"case {name}", flush=True)
    rho0 = initial(kw, *init)
    p, m = oracle_of(kw)
    want_it, got_it = [], []
    want = O.run_psm(p, m, rho0, 5, use_gmres=True, restart=30, max_iter=400, tol=1e-13, iters=want_it)
    got = run_gpu(ch, ctx, kw, rho0, (5,), iters=got_it, use_gmres=True, gmres_restart=30, gmres_max_iter=400,
                  gmres_tol=1e-13)
    err = rel_linf(got, want)
    print(f"{name}: gmres fp64 rel L-inf vs oracle = {err:.3e}; iterations gpu {got_it} oracle {want_it}")
    assert err < 1e-10
    assert all(abs(a - b) <= 2 for a, b in zip(got_it, want_it))
    assert abs(got.sum() - rho0.sum()) < 1e-12 * max(1.0, np.abs(rho0).sum())


@pytest.mark.parametrize("name,kw,init", CASES[:3], ids=[c[0] for c in CASES[:3]])
def test_splitting_fp32_against_fp64_oracle(ch, ctx, name, kw, init):
    rho0 = initial(kw, *init)
    p, m = oracle_of(kw)
    want = O.run_psm(p, m, rho0, 100)
    got = run_gpu(ch, ctx, kw, rho0, (100,), "fp32")
    err = rel_linf(got, want)
    print(f"{name}: splitting fp32 rel L-inf vs fp64 oracle = {err:.3e}")
    assert err < 2e-4


def test_gmres_fp32_against_fp64_oracle(ch, ctx):
    name, kw, init = GMRES_CASES[1]
    rho0 = initial(kw, *init)
    p, m = oracle_of(kw)
    want = O.run_psm(p, m, rho0, 5, use_gmres=True, tol=1e-13, max_iter=400)
    got = run_gpu(ch, ctx, kw, rho0, (5,), "fp32", use_gmres=True, gmres_tol=1e-6)
    err = rel_linf(got, want)
    print(f"{name}: gmres fp32 rel L-inf vs fp64 oracle = {err:.3e}")
    assert err < 1e-4


def test_constant_mobility_splitting_equals_pseudospectral(ch, ctx):
    """M - M0 = 0: the splitting scheme is the pseudospectral integrator (dealiasing off)."""
    kw = dict(CASES[3][1], use_dealias=False)
    rho0 = initial(kw, 0.0, 0.02)
    a = run_gpu(ch, ctx, kw, rho0, (100,))
    d = ch.make_desc(integrator="pseudospectral", **kw)
    s = ch.System(ctx, d)
    s.upload(rho0)
    s.evolve(100)
    b = s.download()
    s.close()
    assert rel_linf(a, b) < 1e-12


def test_psm_rejects_reverse_mode(ch, ctx):
    kw = CASES[1][1]
    s = ch.System(ctx, ch.make_desc(integrator="pseudospectral_mobility", **kw))
    s.upload(initial(kw, 0.5, 0.3))
    with pytest.raises(ch.CHError):
        s.record(4, 0)
    s.close()


LARGE = [
    ("landau2d_N1024_regularised_dealias", dict(dims=2, N=1024, dt=0.02, dx=1.0, k=1.0, mobility="regularised", M=1.0,
                                                rho_min=0.2, S=0.5, use_dealias=True, **LANDAU), (0.5, 0.3)),
    ("wertheim2d_N2048_linear", dict(dims=2, N=2048, dt=3e-4, dx=1.0, k=1.0, mobility="linear", M=1.0, rho_ref=0.05,
                                     S=0.0, use_dealias=True, **WERTHEIM), (0.05, 0.05)),
]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("name,kw,init", LARGE, ids=[c[0] for c in LARGE])
def test_fused_divergence_equals_elementwise_kernels(ch, ctx, monkeypatch, name, kw, init, precision):
    """2D N = 1024 and 2048 fold the divergence term's element-wise kernels into
    the transform passes; the result is the same bits as the element-wise path
    (CH_PSM_FUSE=0), for the splitting and the GMRES schemes, FP64 and FP32."""
    print(f"case {name} {precision}", flush=True)
    rho0 = initial(kw, *init)
    outs = {}
    tol = 1e-10 if precision == "fp64" else 1e-6
    for fuse in ("1", "0"):
        monkeypatch.setenv("CH_PSM_FUSE", fuse)
        outs["split", fuse] = run_gpu(ch, ctx, kw, rho0, (5, 2), precision)
        it = []
        outs["gmres", fuse] = run_gpu(ch, ctx, kw, rho0, (2,), precision, iters=it, use_gmres=True, gmres_tol=tol)
        outs["iters", fuse] = it
    assert np.array_equal(outs["split", "1"], outs["split", "0"])
    assert np.array_equal(outs["gmres", "1"], outs["gmres", "0"])
    assert outs["iters", "1"] == outs["iters", "0"]


def test_fused_splitting_matches_oracle_at_1024(ch, ctx):
    """The fused path against the oracle (splitting, FP64, 4 steps): <= 1e-10."""
    name, kw, init = LARGE[0]
    rho0 = initial(kw, *init)
    p, m = oracle_of(kw)
    want = O.run_psm(p, m, rho0, 4)
    got = run_gpu(ch, ctx, kw, rho0, (4,))
    err = rel_linf(got, want)
    print(f"{name}: fused splitting fp64 rel L-inf vs oracle = {err:.3e}")
    assert err < 1e-10
    assert abs(got.sum() - rho0.sum()) < 1e-12 * np.abs(rho0).sum()
