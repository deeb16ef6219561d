"""GPU parity of the HIP path (through the C ABI) against the oracle and the
reference-held / reference-built fixtures.

Tolerances (stated per result):
  FP64 fields: <= 1e-10 relative L-inf against the reference-built golden
               vectors and the oracle (up to 1000 steps); mass drift <= 1e-12.
  FP32 fields: relative L-inf against the FP64 result <= 2e-4 (1000 Euler
               steps) / 1e-4 (pseudospectral), FP64 reductions.
"""
import os

import numpy as np
import pytest

import oracle as O
from helpers import CASES, GOLD, desc_for, golden, oracle_setup, rel_linf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(ch):
    c = ch.Context(0)
    yield c
    c.close()


def run_gpu(ch, ctx, desc, rho0, steps):
    s = ch.System(ctx, desc)
    s.upload(rho0)
    s.evolve(steps)
    out = s.download()
    s.close()
    return out


@pytest.mark.parametrize("c", CASES, ids=[c["name"] for c in CASES])
def test_reference_golden_vectors_fp64(ch, ctx, c):
    """Reference EulerCPU / EulerMobilityCPU outputs (oracle/_ref) at 1 and T steps."""
    g = golden(c["name"])
    d = desc_for(ch, c)
    one = run_gpu(ch, ctx, d, g["rho0"], 1)
    assert rel_linf(one, g["rho1"]) < 1e-13
    out = run_gpu(ch, ctx, d, g["rho0"], c["steps"])
    assert rel_linf(out, g["rhoT"]) < 1e-10
    assert abs(out.sum() - g["rho0"].sum()) < 1e-12 * max(1.0, np.abs(g["rho0"]).sum())


@pytest.mark.parametrize("c", CASES, ids=[c["name"] for c in CASES])
def test_fp32_against_fp64(ch, ctx, c):
    g = golden(c["name"])
    out = run_gpu(ch, ctx, desc_for(ch, c, "fp32"), g["rho0"], c["steps"])
    err = rel_linf(out, g["rhoT"])
    print(f"{c['name']} fp32 rel L-inf vs fp64 = {err:.3e}")
    assert err < 2e-4


@pytest.mark.parametrize("integrator,name", [("euler", "euler"), ("pseudospectral", "pseudospectral")])
@pytest.mark.parametrize("dims", [1, 2, 3])
def test_reference_regression_fixtures(ch, ctx, integrator, name, dims):
    """The reference's own regression test (tests/test_landau.py:55-65): 100 steps of
    tests/data/simple_landau.toml, compared at the fixtures' 6 printed digits."""
    fx = np.load(os.path.join(GOLD, "reference_landau_fixtures.npz"))
    ref = fx[f"{name}_{dims}D"]
    rho0 = O.init_density(dims, 64, 0.01, 0.01, 0, 123456)
    d = ch.make_desc(dims=dims, N=64, integrator=integrator, dt=0.001, dx=1.0, k=1.0, epsilon=0.9)
    out = run_gpu(ch, ctx, d, rho0, 100)
    mism = sum(a != b for a, b in zip(("%g" % v for v in out), ("%g" % v for v in ref)))
    assert mism == 0
    assert abs(rho0.sum() - out.sum()) < 1e-4


@pytest.mark.parametrize("dims,N,steps", [(1, 256, 300), (2, 128, 300), (3, 32, 100)])
def test_pseudospectral_against_oracle(ch, ctx, dims, N, steps):
    rho0 = O.init_density(dims, N, 0.0, 0.3, 0, 123456)
    kw = dict(dims=dims, N=N, dt=0.1, dx=1.0, k=1.0)
    p = O.Params(**kw)
    ref = O.run(O.PSEUDOSPECTRAL, p, O.Model(epsilon=0.9), rho0, steps)
    out = run_gpu(ch, ctx, ch.make_desc(integrator="pseudospectral", epsilon=0.9, **kw), rho0, steps)
    assert rel_linf(out, ref) < 1e-10
    assert abs(out.sum() - rho0.sum()) < 1e-12 * N ** dims
    out32 = run_gpu(ch, ctx, ch.make_desc(integrator="pseudospectral", epsilon=0.9, precision="fp32", **kw), rho0, steps)
    err = rel_linf(out32, ref)
    print(f"pseudospectral {dims}D fp32 rel L-inf vs fp64 = {err:.3e}")
    assert err < 1e-4


@pytest.mark.parametrize("dims,N,steps,integrator", [(2, 2048, 19, "euler"), (2, 4096, 5, "euler"),
                                                     (2, 2048, 11, "euler_mobility"), (3, 256, 2, "euler")])
def test_stencil_large_grids_against_oracle(ch, ctx, dims, N, steps, integrator):
    """The default region shapes at the largest grids (2D N = 2048, 4096: 128 x 80
    regions; 3D N = 256: 16 x 8 x 8 tiles), with step counts that run launches
    of every depth (19 = 8 + 8 + 2 + 1): FP64 within 1e-12 relative L-inf of
    the oracle, mass within 1e-12."""
    print(f"{integrator} {dims}D N={N}, {steps} steps", flush=True)
    rho0 = O.init_density(dims, N, 0.0, 0.3, 0, 7)
    kw = dict(dims=dims, N=N, dt=0.01, dx=1.0, k=1.0, M=1.0)
    ref = O.run(O.EULER if integrator == "euler" else O.EULER_MOBILITY, O.Params(**kw), O.Model(epsilon=0.9),
                rho0, steps)
    s = ch.System(ctx, ch.make_desc(integrator=integrator, epsilon=0.9, **kw))
    if dims == 2:
        assert s.region_shape == 7
    s.upload(rho0)
    s.evolve(steps)
    out = s.download()
    s.close()
    err = rel_linf(out, ref)
    print(f"rel L-inf vs oracle = {err:.3e}")
    assert err < 1e-12
    assert abs(out.sum() - rho0.sum()) < 1e-12 * N ** dims


@pytest.mark.parametrize("N,steps", [(512, 20), (1024, 20), (2048, 6)])
def test_pseudospectral_large_grids_against_oracle(ch, ctx, N, steps):
    """2D grids whose axis passes run the wave kernels: one wave per line
    (N = 512), two waves per line (N = 1024, 2048) and four waves per line
    (the FP32 fused column pass at N = 1024), with the one-wave row passes.  FP64 within 1e-10 relative L-inf of the oracle, mass within 1e-12;
    FP32 within 1e-4."""
    print(f"pseudospectral 2D N={N}, {steps} steps", flush=True)
    rho0 = O.init_density(2, N, 0.0, 0.3, 0, 5)
    kw = dict(dims=2, N=N, dt=0.1, dx=1.0, k=1.0)
    ref = O.run(O.PSEUDOSPECTRAL, O.Params(**kw), O.Model(epsilon=0.9), rho0, steps)
    out = run_gpu(ch, ctx, ch.make_desc(integrator="pseudospectral", epsilon=0.9, **kw), rho0, steps)
    err = rel_linf(out, ref)
    print(f"N={N}: fp64 rel L-inf vs oracle = {err:.3e}")
    assert err < 1e-10
    assert abs(out.sum() - rho0.sum()) < 1e-12 * N * N
    out32 = run_gpu(ch, ctx, ch.make_desc(integrator="pseudospectral", epsilon=0.9, precision="fp32", **kw), rho0,
                    steps)
    assert rel_linf(out32, ref) < 1e-4


def test_pseudospectral_options_against_oracle(ch, ctx):
    """S splitting, dealiasing, dx != 1 and M != 1 (PseudospectralCPU.cpp:23-25,119-133)."""
    for dims in (1, 2, 3):
        N = 16 if dims == 3 else 64
        kw = dict(dims=dims, N=N, dt=0.05, dx=0.7, k=0.8, M=1.3, S=0.4, use_dealias=1)
        rho0 = O.init_density(dims, N, 0.1, 0.5, 0, 17)
        ref = O.run(O.PSEUDOSPECTRAL, O.Params(**kw), O.Model(epsilon=0.6), rho0, 50)
        out = run_gpu(ch, ctx, ch.make_desc(integrator="pseudospectral", epsilon=0.6, **kw), rho0, 50)
        assert rel_linf(out, ref) < 1e-11, dims


def test_euler_c1_config_against_oracle(ch, ctx):
    """Landau 2D N=128, dt=0.01, eps=0.9, k=1, 1000 steps (the C1 workload, FP64)."""
    rho0 = O.init_density(2, 128, 0.0, 0.01, 0, 123456)
    kw = dict(dims=2, N=128, dt=0.01, dx=1.0, k=1.0)
    ref = O.run(O.EULER, O.Params(**kw), O.Model(epsilon=0.9), rho0, 1000)
    out = run_gpu(ch, ctx, ch.make_desc(integrator="euler", epsilon=0.9, **kw), rho0, 1000)
    assert rel_linf(out, ref) < 1e-10
    assert abs(out.sum() - rho0.sum()) < 1e-12


def test_euler_c1_full_length_against_oracle(ch, ctx):
    """The C1 workload at its full length: Landau 2D N=128, dt=0.01, 10,000 steps,
    FP64 against the oracle (the field has left the linear regime and reached
    |psi| ~ 0.9): within 1e-12 relative L-inf (measured 8e-15), mass within
    1e-12, free energy within 1e-12 relative."""
    rho0 = O.init_density(2, 128, 0.0, 0.01, 0, 123456)
    kw = dict(dims=2, N=128, dt=0.01, dx=1.0, k=1.0)
    p = O.Params(**kw)
    ref = O.run(O.EULER, p, O.Model(epsilon=0.9), rho0, 10000)
    s = ch.System(ctx, ch.make_desc(integrator="euler", epsilon=0.9, **kw))
    s.upload(rho0)
    for _ in range(100):
        s.evolve(100)
    out = s.download()
    m, f = s.observables()
    s.close()
    err = rel_linf(out, ref)
    print(f"C1 10,000 steps: fp64 rel L-inf vs oracle = {err:.3e}")
    assert err < 1e-12
    assert abs(out.sum() - rho0.sum()) < 1e-12
    assert f == pytest.approx(O.average_free_energy(p, O.Model(epsilon=0.9), out), rel=1e-12)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_pseudospectral_c2_full_length_invariants(ch, ctx, precision):
    """The C2 workload at its full length (Landau 2D N=1024, dt=0.1, 20,000
    steps; too long for the oracle): the field stays finite and bounded, the
    mass is conserved (FP64 to 1e-12 of the sum of |psi0|, FP32 fields to
    1e-5), and the free energy has decreased from its initial value."""
    rho0 = O.init_density(2, 1024, 0.0, 0.01, 0, 123456)
    s = ch.System(ctx, ch.make_desc(dims=2, N=1024, integrator="pseudospectral", dt=0.1, epsilon=0.9,
                                    precision=precision))
    s.upload(rho0)
    _, f0 = s.observables()
    s.evolve(20000)
    out = s.download()
    m, f = s.observables()
    s.close()
    scale = np.abs(rho0).sum()
    drift = abs(out.sum() - rho0.sum()) / scale
    print(f"C2 {precision} 20,000 steps: mass drift {drift:.3e} of sum|psi0|, F {f0:.6e} -> {f:.6e}, "
          f"max |psi| {np.abs(out).max():.4f}")
    assert np.isfinite(out).all() and np.abs(out).max() < 1.5
    assert drift < (1e-12 if precision == "fp64" else 1e-5)
    assert f < f0


@pytest.mark.parametrize("integrator", ["euler", "euler_mobility", "pseudospectral"])
@pytest.mark.parametrize("dims", [1, 2, 3])
def test_observables_against_oracle(ch, ctx, integrator, dims):
    N = {1: 128, 2: 64, 3: 16}[dims]
    for model in ("landau", "simple_wertheim"):
        # the nonlinear term is explicit: Wertheim's f'' ~ 1/rho needs a small dt
        dt = 0.1 if (integrator == "pseudospectral" and model == "landau") else 0.001
        if model == "landau":
            om, dens, A = O.Model(O.LANDAU, epsilon=0.9), 0.0, 0.3
            mk = dict(epsilon=0.9)
        else:
            om, dens, A = O.Model(O.SIMPLE_WERTHEIM, B2=1.0, delta=1e4, valence=4), 0.05, 0.05
            mk = dict(free_energy="simple_wertheim", B2=1.0, delta=1e4, valence=4)
        kw = dict(dims=dims, N=N, dt=dt, dx=1.2, k=1.1)
        rho0 = O.init_density(dims, N, dens, A, 0, 3)
        s = ch.System(ctx, ch.make_desc(integrator=integrator, **kw, **mk))
        s.upload(rho0)
        s.evolve(10)
        rho = s.download()
        m, f = s.observables()
        p = O.Params(**kw)
        assert m == pytest.approx(O.average_mass(p, rho), rel=1e-12, abs=1e-15)
        assert f == pytest.approx(O.average_free_energy(p, om, rho), rel=1e-12)
        s.close()


def test_queued_observables_equal_synchronous_ones(ch, ctx):
    """ch_sys_observables_queue / _fetch: the values of the synchronous call at
    each queue point, first in first out, behind later evolve calls; at most
    CH_OBS_QUEUE pending; fetching with none pending is an error."""
    rho0 = O.init_density(2, 256, 0.0, 0.3, 0, 4)
    d = ch.make_desc(dims=2, N=256, dt=0.01)
    a = ch.System(ctx, d)
    a.upload(rho0)
    want = []
    for _ in range(5):
        a.evolve(37)
        want.append(a.observables())
    a.close()
    b = ch.System(ctx, d)
    b.upload(rho0)
    got = []
    for i in range(5):
        b.evolve(37)
        b.queue_observables()
        if i >= 1:
            got.append(b.fetch_observables())  # one interval late
    assert b.observables_pending == 1
    got.append(b.fetch_observables())
    assert got == want
    with pytest.raises(ch.CHError):
        b.fetch_observables()
    for _ in range(ch._lib.OBS_QUEUE):
        b.queue_observables()
    with pytest.raises(ch.CHError):
        b.queue_observables()
    vals = [b.fetch_observables() for _ in range(ch._lib.OBS_QUEUE)]
    assert all(v == want[-1] for v in vals) and b.observables_pending == 0
    b.close()


@pytest.mark.parametrize("kw", [dict(dims=2, N=1024), dict(dims=2, N=128, integrator="euler_mobility",
                                                             mobility="regularised", rho_min=0.5),
                                dict(dims=3, N=32, dt=0.001), dict(dims=2, N=64, integrator="pseudospectral", dt=0.05)],
                         ids=["2d_1024", "2d_em", "3d", "spectral"])
def test_evolve_observe_equals_evolve_then_queue(ch, ctx, kw):
    """ch_sys_evolve_observe: the fields and the queued pairs of evolve + queue,
    for a recurring count (direct on first sight, captured on the second,
    replayed after), mixed with other counts, the queue running past its 8
    slots, and the synchronous observables in between."""
    kw = dict(dict(dt=0.01), **kw)
    N, dims = kw["N"], kw["dims"]
    rho0 = O.init_density(dims, N, 0.5 if kw.get("mobility") else 0.0, 0.3, 0, 8)
    counts = [100, 100, 100, 37, 100, 100, 36, 100, 100, 100, 100, 100]
    a = ch.System(ctx, ch.make_desc(**kw))
    a.upload(rho0)
    want, sync_vals = [], []
    for i, n in enumerate(counts):
        a.evolve(n)
        a.queue_observables()
        want.append(a.fetch_observables())
        if i == 5:
            sync_vals.append(a.observables())
    fa = a.download()
    a.close()
    b = ch.System(ctx, ch.make_desc(**kw))
    b.upload(rho0)
    got = []
    for i, n in enumerate(counts):
        b.evolve_observe(n)
        if b.observables_pending > 2:
            got.append(b.fetch_observables())
        if i == 5:
            sync_vals.append(b.observables())
    while b.observables_pending:
        got.append(b.fetch_observables())
    fb = b.download()
    b.close()
    assert got == want
    assert sync_vals[0] == sync_vals[1]
    assert np.array_equal(fa, fb)


def test_observables_nan_status(ch, ctx):
    s = ch.System(ctx, ch.make_desc(dims=2, N=16))
    rho = np.zeros(256)
    rho[37] = np.nan
    s.upload(rho)
    with pytest.raises(ch.CHError) as e:
        s.observables()
    assert e.value.status == 3  # CH_ENAN (src/CahnHilliard.cpp:290-293 exits instead)
    s.close()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-15), ("fp32", 1e-6)])
def test_pointwise_model_and_mobility(ch, ctx, precision, tol):
    import torch
    dt = torch.float64 if precision == "fp64" else torch.float32
    rng = np.random.default_rng(0)
    rho = rng.uniform(0.001, 0.9, 4096)
    r = torch.tensor(rho, dtype=dt, device="cuda")
    out = torch.empty_like(r)
    models = [(O.Model(O.LANDAU, epsilon=0.9), dict(epsilon=0.9)),
              (O.Model(O.SIMPLE_WERTHEIM, B2=1.0, delta=1e4, valence=4, regularisation_delta=0.01),
               dict(free_energy="simple_wertheim", B2=1.0, delta=1e4, valence=4, regularisation_delta=0.01))]
    for om, mk in models:
        d = ch.make_desc(dims=1, N=4096, precision=precision, **mk)
        torch.cuda.synchronize()
        ch.der_bulk_free_energy(ctx, d, r.data_ptr(), out.data_ptr(), r.numel())
        ctx.synchronize()
        ref = O.der_bulk(om, r.double().cpu().numpy())
        assert np.abs(out.double().cpu().numpy() - ref).max() <= tol * max(1.0, np.abs(ref).max()) * 4
    om = O.Model(O.SIMPLE_WERTHEIM, B2=1.0, delta=1e4, valence=4)
    for mob, kw in (("constant", {}), ("free_energy", {}), ("regularised", dict(rho_min=0.1)),
                    ("linear", dict(rho_ref=0.05))):
        d = ch.make_desc(dims=1, N=4096, precision=precision, free_energy="simple_wertheim", B2=1.0, delta=1e4,
                         valence=4, mobility=mob, M=2.0, **kw)
        ch.set_mobility(ctx, d, r.data_ptr(), out.data_ptr(), r.numel())
        ctx.synchronize()
        p = O.Params(dims=1, N=4096, mobility=O.MOB_CONSTANT if mob == "constant" else
                     {"free_energy": 1, "regularised": 2, "linear": 3}[mob], M=2.0, **kw)
        ref = O.mobility(p, om, r.double().cpu().numpy())
        assert np.abs(out.double().cpu().numpy() - ref).max() <= tol * max(1.0, np.abs(ref).max()) * 4, mob


def test_wertheim_fprime_wide_range(ch, ctx):
    """SimpleWertheim f' = 2 B2 rho + ln(rho X^valence) over rho from 1e-9 to 1e3
    (log-uniform) against the oracle's libm-based f', element by element: the
    device logarithm (series of atanh((m - 1)/(m + 1)) with the exponent split
    off) on arguments rho X^4 spanning ~14 decades, within 2e-15 of the terms'
    magnitude.  delta = 1e10 keeps 2 valence delta rho >= 80, where X(rho) =
    (sqrt(1 + 2 valence delta rho) - 1) / (valence delta rho) has no
    cancellation (below ~1e-2 the subtraction loses digits in any
    implementation, and a one-ulp square-root difference shows in X)."""
    import torch
    rng = np.random.default_rng(11)
    rho = 10.0 ** rng.uniform(-9.0, 3.0, 1 << 16)
    r = torch.tensor(rho, dtype=torch.float64, device="cuda")
    out = torch.empty_like(r)
    om = O.Model(O.SIMPLE_WERTHEIM, B2=1.0, delta=1e10, valence=4)
    d = ch.make_desc(dims=1, N=1 << 16, free_energy="simple_wertheim", B2=1.0, delta=1e10, valence=4)
    torch.cuda.synchronize()
    ch.der_bulk_free_energy(ctx, d, r.data_ptr(), out.data_ptr(), r.numel())
    ctx.synchronize()
    got = out.cpu().numpy()
    ref = O.der_bulk(om, rho)
    scale = np.abs(2.0 * rho) + np.abs(ref - 2.0 * rho) + 1.0
    err = np.abs(got - ref) / scale
    print(f"max scaled error {err.max():.2e} at rho = {rho[err.argmax()]:.3e}")
    assert np.all(np.isfinite(got)) and err.max() < 2e-15


def test_linear_mobility_against_oracle(ch, ctx):
    """mobility.type = "linear" (M rho / rho_ref, new) with SimpleWertheim in 3D (C5 shape, small N)."""
    kw = dict(dims=3, N=16, dt=1e-4, dx=1.0, k=1.0, M=1.0, rho_ref=0.05)
    rho0 = O.init_density(3, 16, 0.05, 0.05, 0, 123456)
    om = O.Model(O.SIMPLE_WERTHEIM, B2=1.0, delta=1e4, valence=4)
    ref = O.run(O.EULER_MOBILITY, O.Params(mobility=O.MOB_LINEAR, **kw), om, rho0, 200)
    d = ch.make_desc(integrator="euler_mobility", free_energy="simple_wertheim", B2=1.0, delta=1e4, valence=4,
                     mobility="linear", **kw)
    out = run_gpu(ch, ctx, d, rho0, 200)
    assert rel_linf(out - rho0, ref - rho0) < 1e-9
    assert rel_linf(out, ref) < 1e-12


def test_graph_replay_equals_eager(ch, ctx, monkeypatch):
    """The captured-graph step loop is bit-identical to step-by-step launches."""
    rho0 = O.init_density(2, 64, 0.0, 0.3, 0, 9)
    d = ch.make_desc(dims=2, N=64, dt=0.01)
    s = ch.System(ctx, d)
    s.upload(rho0)
    s.evolve(64 * 3 + 5)
    a = s.download()
    s.upload(rho0)
    for _ in range(64 * 3 + 5):
        s.evolve(1)
    b = s.download()
    s.close()
    assert np.array_equal(a, b)


@pytest.mark.parametrize("N", [64, 1024])
def test_remainder_graph_equals_direct_launches(ch, ctx, N):
    """Recurring counts of at most 256 steps, and the remainders nsteps % 64 of
    longer runs, run as one cached graph per starting buffer (captured on the
    second occurrence, kept until another count recurs): the same bits as the
    direct launches, whatever the sequence of counts and parities."""
    rho0 = O.init_density(2, N, 0.0, 0.3, 0, 9)
    d = ch.make_desc(dims=2, N=N, dt=0.01)
    counts = [100, 100, 100, 20, 100, 20, 20, 100, 9, 9, 9, 36, 36, 36, 100, 300, 300, 300, 256, 256, 7]
    s = ch.System(ctx, d)
    s.upload(rho0)
    for n in counts:
        s.evolve(n)
    a = s.download()
    s.upload(rho0)
    for _ in range(sum(counts)):
        s.evolve(1)
    b = s.download()
    s.close()
    assert np.array_equal(a, b)


@pytest.mark.parametrize("N", [64, 1024])
def test_pseudospectral_cached_graphs_equal_direct_steps(ch, ctx, N):
    """Pseudospectral chains: recurring counts of at most 256 steps replay one
    cached graph, longer runs the 64-step graph and the remainder; the same bits
    as one step per call."""
    rho0 = O.init_density(2, N, 0.0, 0.3, 0, 13)
    d = ch.make_desc(dims=2, N=N, integrator="pseudospectral", dt=0.1, epsilon=0.9)
    counts = [20, 20, 20, 5, 20, 5, 5, 300, 300, 7]
    s = ch.System(ctx, d)
    s.upload(rho0)
    for n in counts:
        s.evolve(n)
    a = s.download()
    s.upload(rho0)
    for _ in range(sum(counts)):
        s.evolve(1)
    b = s.download()
    s.close()
    assert np.array_equal(a, b)


@pytest.mark.parametrize("integrator,mob", [("euler", "constant"), ("euler_mobility", "regularised")])
def test_temporal_blocking_variants_bitwise(ch, ctx, monkeypatch, integrator, mob):
    """Every steps-per-launch depth and rows-per-wave split of the 2D wave kernel
    gives the same bits as one step per launch, and matches the oracle."""
    N, steps = 128, 37
    rho0 = O.init_density(2, N, 0.5, 0.1, 0, 21)
    kw = dict(dims=2, N=N, dt=0.01, dx=1.0, k=1.0, M=1.0, rho_min=0.1)
    outs = {}
    for cfg in ("0", "1", "2", "3", "4", "5", "6", "7", "8"):
        for tb in ("1", "2", "4", "8"):
            if cfg == "1" and tb == "8":
                continue  # a 128 x 32 region has no interior left after 8 steps
            monkeypatch.setenv("CH_TB_STEPS", tb)
            monkeypatch.setenv("CH_WAVE_CFG", cfg)
            s = ch.System(ctx, ch.make_desc(integrator=integrator, mobility=mob, epsilon=0.9, **kw))
            assert s.steps_per_launch == int(tb)
            s.upload(rho0)
            s.evolve(steps)
            outs[(cfg, tb)] = s.download()
            s.close()
    base = outs[("0", "1")]
    for key, v in outs.items():
        assert np.array_equal(v, base), key
    p = O.Params(mobility=O.MOB_CONSTANT if mob == "constant" else O.MOB_REGULARISED, **kw)
    ref = O.run(O.EULER if integrator == "euler" else O.EULER_MOBILITY, p, O.Model(epsilon=0.9), rho0, steps)
    assert rel_linf(base, ref) < 1e-12


@pytest.mark.parametrize("case", ["wertheim_em_linear", "wertheim_euler", "landau_em_regularised", "landau_em_constant"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_split3d_equals_fused_tile(ch, ctx, monkeypatch, case, precision):
    """The 3D two-launch step (mu and the cells' mobilities of every cell, then
    the update: the default for SimpleWertheim) against the fused LDS tile
    kernel, bit for bit, at N = 32 and N = 128 over 24 steps; both evaluate each
    cell with the same arithmetic."""
    from cahn_hilliard_amd.init import initial_density
    kw = {"wertheim_em_linear": dict(integrator="euler_mobility", free_energy="simple_wertheim", B2=1.0, delta=1e4,
                                     valence=4, mobility="linear", M=1.0, rho_ref=0.05, dt=1e-4, k=1.0),
          "wertheim_euler": dict(integrator="euler", free_energy="simple_wertheim", B2=1.0, delta=1e4, valence=4,
                                 dt=1e-4, k=1.0),
          "landau_em_regularised": dict(integrator="euler_mobility", mobility="regularised", rho_min=0.5, dt=1e-3,
                                        k=1.0, epsilon=0.9),
          "landau_em_constant": dict(integrator="euler_mobility", mobility="constant", M=1.0, dt=1e-3, k=1.0,
                                     epsilon=0.9)}[case]
    for N in (32, 128):
        rho0 = initial_density(3, N, 0.05 if "wertheim" in case else 0.3, 0.05, 0, 4242 + N)
        outs = {}
        for t3 in ("1", "3"):
            monkeypatch.setenv("CH_TILE3D", t3)
            s = ch.System(ctx, ch.make_desc(dims=3, N=N, precision=precision, **kw))
            s.upload(rho0)
            s.evolve(24)
            outs[t3] = s.download()
            s.close()
        print(f"{case} {precision} N={N}: max |rho - rho0| {np.abs(outs['3'] - rho0.ravel()).max():.3e}")
        assert np.all(np.isfinite(outs["3"])) and not np.array_equal(outs["3"], rho0.ravel())
        assert np.array_equal(outs["1"], outs["3"]), (case, precision, N)


@pytest.mark.parametrize("name", ["large_landau_2d_N1024", "large_landau_2d_N128", "large_wertheim_3d_N128"])
def test_full_size_against_reference_build(ch, ctx, name):
    """Full-size workloads against the reference's own integrators (oracle/_ref,
    tests/golden/large_*.npz), each through the library's default kernels:
      the headline Landau 2D N=1024 Euler (FP64, default region shape), 1000 steps;
      Landau 2D N=128 Euler, 1000 and 10,000 steps;
      SimpleWertheim 3D N=128 euler_mobility with M = rho / 0.05, 400 steps.
    The stored cells (every 8th / every / every 4th along each axis) within
    1e-10 relative L-inf, the whole-field sum within 1e-12 per cell and the sum
    of squares within 1e-10 relative."""
    from cahn_hilliard_amd.init import initial_density
    from helpers import LARGE
    c = LARGE[name]
    g = golden(name)
    N, st, dims = c["N"], c["stride"], c["dims"]
    s = ch.System(ctx, desc_for(ch, c))
    if name == "large_landau_2d_N1024":
        assert s.region_shape == 8 and s.steps_per_launch == 8
    density, A, seed = c["init"]
    s.upload(initial_density(dims, N, density, A, 0, seed))
    done = 0
    for t in c["marks"]:
        s.evolve(t - done)
        done = t
        f = s.download()
        sub = f.reshape((N,) * dims)[(slice(None, None, st),) * dims]
        err = rel_linf(sub, g[f"rho{t}_sub"])
        print(f"{name} after {t} steps: rel L-inf vs reference {err:.2e}, "
              f"sum diff {abs(f.sum() - float(g[f'sum{t}'])):.2e}")
        assert err < 1e-10
        assert abs(f.sum() - float(g[f"sum{t}"])) < 1e-12 * f.size
        assert abs((f * f).sum() / float(g[f"sumsq{t}"]) - 1) < 1e-10
    s.close()


def test_full_size_properties(ch, ctx):
    """Landau 2D N=1024 FP64 (the headline workload): exact psi -> -psi antisymmetry,
    exact translation equivariance and mass conservation after 2000 steps."""
    N = 1024
    rho0 = O.init_density(2, N, 0.0, 0.01, 0, 123456)
    d = ch.make_desc(dims=2, N=N, dt=0.01, epsilon=0.9, k=1.0)
    a = run_gpu(ch, ctx, d, rho0, 2000)
    b = run_gpu(ch, ctx, d, -rho0, 2000)
    assert np.array_equal(a, -b)
    sh = lambda f: np.roll(f.reshape(N, N), (37, 501), axis=(0, 1)).ravel()
    c = run_gpu(ch, ctx, d, sh(rho0), 2000)
    assert np.array_equal(c, sh(a))
    assert abs(a.sum() - rho0.sum()) < 1e-12 * N
    assert np.abs(a).max() > 0.0


def test_invalid_descriptors(ch, ctx):
    bad = [dict(dims=2, N=48), dict(dims=4, N=16), dict(dims=2, N=2),
           dict(dims=2, N=16, integrator="euler", mobility="regularised"),
           dict(dims=2, N=16, integrator="euler_mobility", mobility="free_energy"),
           dict(dims=2, N=16, dt=-1.0)]
    for kw in bad:
        with pytest.raises(ch.CHError):
            ch.System(ctx, ch.make_desc(**kw))
