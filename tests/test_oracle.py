"""Pins the oracle (oracle/ch_oracle.c) before anything is compared with it.

Each check is chosen so that a plausible mistake (dropped term, wrong sign,
wrong index, factor 2 in k) fails it:
  * reference regression fixtures (tests/data/reference/*.dat, 6 digits): exact
    '%g' token match for Euler and pseudospectral in 1D/2D/3D;
  * golden vectors from the reference's own EulerCPU / EulerMobilityCPU
    (oracle/_ref, tests/golden/make_golden.py): 1e-12 relative;
  * drand48 against glibc's;
  * closed forms: Laplacian eigenmodes, spectral per-mode update, observables
    on uniform and single-cosine fields, Wertheim mass-action law;
  * library/brute force: FFT vs O(n^2) DFT, S(q) vs numpy.fft;
  * finite differences for f', f'' and the adjoint.
"""
import ctypes
import json
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = json.load(open(os.path.join(GOLD, "cases.json")))
MOB = {"constant": O.MOB_CONSTANT, "free_energy": O.MOB_FREE_ENERGY, "regularised": O.MOB_REGULARISED,
       "linear": O.MOB_LINEAR}


def case_setup(c):
    m = c["model"]
    if m["free_energy"] == "landau":
        mod = O.Model(O.LANDAU, epsilon=m["epsilon"])
    else:
        mod = O.Model(O.SIMPLE_WERTHEIM, B2=m["B2"], delta=m["delta"],
                      regularisation_delta=m.get("regularisation_delta", 0.0), valence=m["valence"])
    mb = c["mobility"]
    p = O.Params(dims=c["dims"], N=c["N"], dt=c["dt"], dx=c["dx"], k=c["k"], mobility=MOB[mb["type"]],
                 M=mb["M"], rho_min=mb.get("rho_min", 0.0), rho_ref=mb.get("rho_ref", 1.0))
    integ = O.EULER if c["integrator"] == "euler" else O.EULER_MOBILITY
    return p, mod, integ


def test_drand48_matches_glibc():
    libc = ctypes.CDLL("libc.so.6")
    libc.drand48.restype = ctypes.c_double
    libc.srand48.argtypes = [ctypes.c_long]
    for seed in (123456, 0, 1, 2 ** 31 - 1):
        libc.srand48(seed)
        ref = np.array([libc.drand48() for _ in range(500)])
        assert np.array_equal(O.drand48_stream(seed, 500), ref)


@pytest.mark.parametrize("integ,name", [(O.EULER, "euler"), (O.PSEUDOSPECTRAL, "pseudospectral")])
@pytest.mark.parametrize("dims", [1, 2, 3])
def test_reference_regression_fixtures(integ, name, dims):
    """tests/test_landau.py:55-65 of the reference, on the oracle: 100 steps of
    tests/data/simple_landau.toml, compared token by token."""
    fx = np.load(os.path.join(GOLD, "reference_landau_fixtures.npz"))
    ref = fx[f"{name}_{dims}D"]
    p = O.Params(dims=dims, N=64, dt=0.001, dx=1.0, k=1.0)
    m = O.Model(epsilon=0.9)
    rho0 = O.init_density(dims, 64, 0.01, 0.01, 0, 123456)
    rho = O.run(integ, p, m, rho0, 100)
    assert ["%g" % v for v in rho] == ["%g" % v for v in ref]
    # reference tests/test_landau.py:41-53
    assert abs(rho0.sum() - rho.sum()) < 1e-4


@pytest.mark.parametrize("c", CASES, ids=[c["name"] for c in CASES])
def test_golden_vectors_from_reference_build(c):
    g = np.load(os.path.join(GOLD, c["name"] + ".npz"))
    p, mod, integ = case_setup(c)
    r0 = O.init_density(c["dims"], c["N"], c["init"][0], c["init"][1], 0, c["init"][2])
    assert np.array_equal(r0, g["rho0"])
    r1 = O.run(integ, p, mod, g["rho0"], 1)
    rT = O.run(integ, p, mod, g["rho0"], c["steps"])
    scale = np.abs(g["rhoT"]).max()
    assert np.abs(r1 - g["rho1"]).max() <= 1e-14 * scale
    assert np.abs(rT - g["rhoT"]).max() <= 1e-12 * scale


def lap_eig(N, dx, m):
    return -(4.0 / dx ** 2) * sum(np.sin(np.pi * mi / N) ** 2 for mi in m)


def test_c1_full_length_against_reference_build():
    """The N=128 workload (Landau 2D, eps 0.9, k 1, dt 0.01, psi0 = U[-0.01, 0.01)
    from drand48 seed 123456): the oracle against the reference's own EulerCPU
    (tests/golden/large_landau_2d_N128.npz) after 1000 and 10,000 steps."""
    from helpers import LARGE, golden
    c = LARGE["large_landau_2d_N128"]
    g = golden(c["name"])
    p = O.Params(dims=2, N=128, dt=0.01, dx=1.0, k=1.0)
    m = O.Model(epsilon=0.9)
    rho = O.init_density(2, 128, 0.0, 0.01, 0, 123456)
    done = 0
    for t in c["marks"]:
        rho = O.run(O.EULER, p, m, rho, t - done)
        done = t
        ref = g[f"rho{t}_sub"].ravel()
        err = float(np.abs(rho - ref).max() / np.abs(ref).max())
        print(f"C1 oracle vs reference after {t} steps: rel L-inf {err:.2e}")
        assert err < 1e-10
        assert abs(rho.sum() - float(g[f"sum{t}"])) < 1e-12 * rho.size


@pytest.mark.parametrize("dims,m", [(1, (3,)), (2, (2, 5)), (3, (1, 2, 3))])
def test_euler_linear_eigenmode(dims, m):
    """psi = a cos(2 pi m.x / N), a = 1e-7: one Euler step multiplies it by
    1 + dt M lam (-eps - 2 k lam) (the 2k convention of EulerCPU.cpp:24)."""
    N, dx, dt, eps, k, M = 16, 1.1, 0.003, 0.7, 1.3, 0.9
    grids = np.meshgrid(*[np.arange(N)] * dims, indexing="ij")
    phase = sum(2 * np.pi * mi * g / N for mi, g in zip(m, grids))
    psi = 1e-7 * np.cos(phase).transpose().ravel()  # idx = x + N y + N^2 z
    lam = lap_eig(N, dx, m)
    p = O.Params(dims=dims, N=N, dt=dt, dx=dx, k=k, M=M)
    out = O.run(O.EULER, p, O.Model(epsilon=eps), psi, 1)
    expect = psi * (1 + dt * M * lam * (-eps - 2 * k * lam))
    assert np.abs(out - expect).max() < 1e-11 * np.abs(psi).max()
    outm = O.run(O.EULER_MOBILITY, p, O.Model(epsilon=eps), psi, 1)
    assert np.abs(outm - expect).max() < 1e-11 * np.abs(psi).max()


@pytest.mark.parametrize("dims,m", [(1, (5,)), (2, (3, 1)), (3, (1, 0, 2))])
def test_pseudospectral_mode_closed_form(dims, m):
    """S = 0, small amplitude: rho_hat <- rho_hat (1 + dt M q^2 eps)/(1 + 2 dt M k q^4)
    (PseudospectralCPU.cpp:128-130 with f' = -eps rho)."""
    N, dx, dt, eps, k, M = 16, 0.9, 0.5, 0.8, 1.4, 1.0
    grids = np.meshgrid(*[np.arange(N)] * dims, indexing="ij")
    phase = sum(2 * np.pi * mi * g / N for mi, g in zip(m, grids))
    psi = 1e-7 * np.cos(phase).transpose().ravel()
    q2 = (2 * np.pi / (N * dx)) ** 2 * sum(mi ** 2 for mi in m)
    p = O.Params(dims=dims, N=N, dt=dt, dx=dx, k=k, M=M)
    out = O.run(O.PSEUDOSPECTRAL, p, O.Model(epsilon=eps), psi, 3)
    g = (1 + dt * M * q2 * eps) / (1 + 2 * dt * M * k * q2 ** 2)
    assert np.abs(out - psi * g ** 3).max() < 1e-12 * np.abs(psi).max()


def test_brute_force_dense_operator_2d():
    """Tiny 2D grid: the step equals psi + dt M L (f'(psi) - 2k L psi) with L the
    dense periodic 5-point matrix built from the adjacency, independently of the
    oracle's index arithmetic."""
    N, dx, dt, eps, k, M = 4, 1.0, 0.01, 0.9, 1.0, 1.0
    n = N * N
    L = np.zeros((n, n))
    for y in range(N):
        for x in range(N):
            i = x + N * y
            L[i, i] = -4
            for xx, yy in ((x + 1, y), (x - 1, y), (x, y + 1), (x, y - 1)):
                L[i, (xx % N) + N * (yy % N)] += 1
    L /= dx * dx
    psi = np.random.default_rng(0).uniform(-0.5, 0.5, n)
    mu = -eps * psi + psi ** 3 - 2 * k * L @ psi
    expect = psi + dt * M * L @ mu
    out = O.run(O.EULER, O.Params(dims=2, N=N, dt=dt, dx=dx, k=k, M=M), O.Model(epsilon=eps), psi, 1)
    assert np.abs(out - expect).max() < 1e-15


@pytest.mark.parametrize("integ", [O.EULER, O.EULER_MOBILITY, O.PSEUDOSPECTRAL])
def test_mass_conservation(integ):
    p = O.Params(dims=2, N=32, dt=0.01 if integ != O.PSEUDOSPECTRAL else 0.1)
    psi = O.init_density(2, 32, 0.0, 0.3, 0, 5)
    out = O.run(integ, p, O.Model(), psi, 500)
    assert abs(out.sum() - psi.sum()) < 1e-12


def test_fft_against_brute_dft():
    rng = np.random.default_rng(1)
    for n in (2, 4, 8, 16, 64):
        x = rng.normal(size=n) + 1j * rng.normal(size=n)
        for inv in (False, True):
            assert np.abs(O.fft_c2c(x, inv) - O.dft_brute(x, inv)).max() < 1e-12 * n
    x = rng.normal(size=128) + 1j * rng.normal(size=128)
    assert np.abs(O.fft_c2c(x) - np.fft.fft(x)).max() < 1e-12


def test_observables_closed_forms():
    m = O.Model(epsilon=0.9)
    for dims in (1, 2, 3):
        N, dx, k = 8, 1.5, 1.3
        p = O.Params(dims=dims, N=N, dx=dx, k=k)
        psi = np.full(N ** dims, 0.37)
        vb = dx ** 3  # V_bin = dx^3 in every dimension (CahnHilliard.cpp:163)
        assert O.average_mass(p, psi) == pytest.approx(0.37 * vb, rel=1e-15)
        f = -0.5 * 0.9 * 0.37 ** 2 + 0.25 * 0.37 ** 4
        assert O.average_free_energy(p, m, psi) == pytest.approx(f * vb, rel=1e-14)
    # single cosine along x: forward-difference gradient sum is A^2 (1 - cos(2 pi m/N)) / dx^2 per cell
    N, dx, k, A, mm = 16, 1.0, 1.0, 1e-3, 3
    x = np.arange(N)
    psi = np.tile(A * np.cos(2 * np.pi * mm * x / N), N)
    p = O.Params(dims=2, N=N, dx=dx, k=k)
    bulk = np.mean(-0.5 * 0.9 * psi ** 2 + 0.25 * psi ** 4)
    grad = k * A ** 2 * (1 - np.cos(2 * np.pi * mm / N)) / dx ** 2
    assert O.average_free_energy(p, m, psi) == pytest.approx(bulk + grad, rel=1e-12)


def test_model_derivatives_by_finite_differences():
    models = [O.Model(O.LANDAU, epsilon=0.9),
              O.Model(O.SIMPLE_WERTHEIM, B2=1.0, delta=1e4, valence=4),
              O.Model(O.SIMPLE_WERTHEIM, B2=-2.0, delta=30.0, valence=2, regularisation_delta=1e-3)]
    for m in models:
        for r in (0.02, 0.05, 0.3, 0.7):
            h = 1e-6 * r
            fd = (O.bulk(m, r + h) - O.bulk(m, r - h)) / (2 * h)
            assert fd == pytest.approx(O.der_bulk(m, np.array([r]))[0], rel=1e-7, abs=1e-9)
            fd2 = (O.der_bulk(m, np.array([r + h]))[0] - O.der_bulk(m, np.array([r - h]))[0]) / (2 * h)
            assert fd2 == pytest.approx(O.second_der_bulk(m, r), rel=1e-6)


def test_wertheim_mass_action_law():
    m = O.Model(O.SIMPLE_WERTHEIM, B2=1.0, delta=1e4, valence=4)
    for r in (1e-6, 0.01, 0.05, 1.0):
        X = O.wertheim_X(m, r)
        assert m.valence * m.delta * r * X * X + X - 1 == pytest.approx(0, abs=1e-12)
    assert O.wertheim_X(m, 0.05) == pytest.approx(0.0221, rel=1e-2)


def numpy_radial_sq(field, dx=1.0):
    """Independent numpy statement of the circularly averaged S(q) (library FFT)."""
    L = field.shape[0]
    phi = field - field.mean()
    Sk = np.abs(np.fft.fft2(phi)) ** 2 / (L * L)
    k = 2 * np.pi * np.fft.fftfreq(L, d=dx)
    KX, KY = np.meshgrid(k, k, indexing="ij")
    shell = np.rint(np.sqrt(KX ** 2 + KY ** 2) / (2 * np.pi / (L * dx))).astype(np.int64).ravel()
    sums = np.bincount(shell, weights=Sk.ravel())
    cnt = np.bincount(shell)
    return np.where(cnt > 0, sums / np.maximum(cnt, 1), 0.0)


def test_radial_sq_and_loss_gradient():
    rng = np.random.default_rng(3)
    f = rng.normal(size=(32, 32))
    sq = O.radial_sq(f, 1.0)
    assert len(sq) == 24  # rint(16 sqrt 2) + 1 = 23 + 1
    assert np.abs(sq - numpy_radial_sq(f)).max() < 1e-12 * sq.max()
    assert len(O.radial_sq(np.zeros((256, 256)))) == 182
    T = numpy_radial_sq(rng.normal(size=(32, 32)))
    loss, g = O.sq_loss(f, T)
    d = rng.normal(size=f.shape)
    h = 1e-6
    fd = (O.sq_loss(f + h * d, T, grad=False) - O.sq_loss(f - h * d, T, grad=False)) / (2 * h)
    assert np.sum(g * d) == pytest.approx(fd, rel=1e-6)


def test_euler_adjoint_against_finite_differences():
    N, steps = 16, 40
    p = O.Params(dims=2, N=N, dt=0.01, dx=1.0, k=1.2, M=0.9)
    m = O.Model(epsilon=0.9)
    psi0 = O.init_density(2, N, 0.0, 0.3, 0, 123456)
    w = np.random.default_rng(4).normal(size=N * N)
    lam0, grad = O.euler_adjoint(p, m, psi0, steps, w)

    def loss(pp, mm, x):
        return float(np.dot(w, O.run(O.EULER, pp, mm, x, steps)))

    d = np.random.default_rng(5).normal(size=N * N)
    h = 1e-6
    fd = (loss(p, m, psi0 + h * d) - loss(p, m, psi0 - h * d)) / (2 * h)
    assert np.dot(lam0, d) == pytest.approx(fd, rel=1e-6)
    for i, (attr, obj) in enumerate([("epsilon", "m"), ("k", "p"), ("M", "p")]):
        def shifted(s):
            pp = O.Params(**{**p.__dict__})
            mm = O.Model(**{**m.__dict__})
            target = mm if obj == "m" else pp
            setattr(target, attr, getattr(target, attr) + s)
            return loss(pp, mm, psi0)
        fd = (shifted(h) - shifted(-h)) / (2 * h)
        assert grad[i] == pytest.approx(fd, rel=1e-6)


def test_euler_adjoint_dot_product():
    """<J v, u> = <v, J^T u> for one step (J by central differences of the step)."""
    N = 8
    p = O.Params(dims=2, N=N, dt=0.01, k=1.0)
    m = O.Model(epsilon=0.9)
    rng = np.random.default_rng(6)
    psi = rng.uniform(-0.5, 0.5, N * N)
    v, u = rng.normal(size=N * N), rng.normal(size=N * N)
    h = 1e-6
    Jv = (O.run(O.EULER, p, m, psi + h * v, 1) - O.run(O.EULER, p, m, psi - h * v, 1)) / (2 * h)
    JTu, _ = O.euler_adjoint_step(p, m, psi, u)
    assert np.dot(Jv, u) == pytest.approx(np.dot(v, JTu), rel=1e-8)


def test_parameter_derivatives_by_finite_differences():
    """d f'/d delta, d f/d delta, dM/drho and dM/d delta against central differences."""
    models = [O.Model(O.SIMPLE_WERTHEIM, B2=1.0, delta=1e4, valence=4),
              O.Model(O.SIMPLE_WERTHEIM, B2=-2.0, delta=30.0, valence=2, regularisation_delta=1e-3)]
    for m in models:
        for r in (0.02, 0.05, 0.3):
            hd = 1e-6 * m.delta
            up, dn = O.Model(**{**m.__dict__, "delta": m.delta + hd}), O.Model(**{**m.__dict__, "delta": m.delta - hd})
            fd = (O.der_bulk(up, np.array([r]))[0] - O.der_bulk(dn, np.array([r]))[0]) / (2 * hd)
            assert O.der_bulk_ddelta(m, r) == pytest.approx(fd, rel=1e-6)
            fd = (O.bulk(up, r) - O.bulk(dn, r)) / (2 * hd)
            assert O.bulk_ddelta(m, r) == pytest.approx(fd, rel=1e-6)
            for mob in (O.MOB_CONSTANT, O.MOB_FREE_ENERGY, O.MOB_REGULARISED, O.MOB_LINEAR):
                p = O.Params(mobility=mob, M=1.3, rho_min=0.1, rho_ref=0.05)
                h = 1e-6 * r
                fd = (O.mobility(p, m, [r + h])[0] - O.mobility(p, m, [r - h])[0]) / (2 * h)
                assert O.mobility_drho(p, m, r) == pytest.approx(fd, rel=1e-6, abs=1e-12)
                fd = (O.mobility(p, up, [r])[0] - O.mobility(p, dn, [r])[0]) / (2 * hd)
                assert O.mobility_ddelta(p, m, r) == pytest.approx(fd, rel=1e-6, abs=1e-14)
    lm = O.Model(O.LANDAU, epsilon=0.9)
    assert O.der_bulk_ddelta(lm, 0.3) == 0.0 and O.bulk_ddelta(lm, 0.3) == 0.0


ADJOINT_CASES = [
    # integrator, dims, N, steps, model kwargs, params kwargs, (init density, amplitude)
    (O.EULER_MOBILITY, 3, 4, 15, dict(model=O.SIMPLE_WERTHEIM, B2=1.0, delta=1e4, valence=4),
     dict(dt=1e-4, k=1.0, mobility=O.MOB_LINEAR, M=1.0, rho_ref=0.05), (0.05, 0.05)),
    (O.EULER_MOBILITY, 2, 8, 20, dict(model=O.SIMPLE_WERTHEIM, B2=-2.0, delta=30.0, valence=2),
     dict(dt=2e-4, k=0.7, mobility=O.MOB_FREE_ENERGY, M=1.1), (0.3, 0.2)),
    (O.EULER_MOBILITY, 2, 8, 25, dict(model=O.LANDAU, epsilon=0.9),
     dict(dt=0.01, k=1.2, mobility=O.MOB_REGULARISED, M=0.9, rho_min=0.5), (0.5, 0.2)),
    (O.EULER_MOBILITY, 1, 16, 30, dict(model=O.LANDAU, epsilon=0.8),
     dict(dt=0.01, k=1.0, mobility=O.MOB_CONSTANT, M=1.3, dx=0.7), (0.0, 0.3)),
    (O.EULER, 2, 8, 20, dict(model=O.SIMPLE_WERTHEIM, B2=1.0, delta=1e3, valence=4),
     dict(dt=1e-4, k=1.0, M=1.0), (0.2, 0.1)),
]


@pytest.mark.parametrize("case", range(len(ADJOINT_CASES)))
def test_adjoint_against_finite_differences(case):
    """dL/dpsi0 and dL/d(eps, k, M, delta, B2) of L = <w, psi_T> against central
    differences of the forward map, for both stencil integrators and every
    mobility rule."""
    integ, dims, N, steps, mk, pk, (rho0, A) = ADJOINT_CASES[case]
    p = O.Params(dims=dims, N=N, **pk)
    m = O.Model(**mk)
    psi0 = O.init_density(dims, N, rho0, A, 0, 123456 + case)
    rng = np.random.default_rng(case)
    w, d = rng.normal(size=N ** dims), rng.normal(size=N ** dims)
    lam0, grad = O.adjoint(integ, p, m, psi0, steps, w)

    def loss(pp, mm, x):
        return float(np.dot(w, O.run(integ, pp, mm, x, steps)))

    h = 1e-6 * max(1.0, np.abs(psi0).max())
    fd = (loss(p, m, psi0 + h * d) - loss(p, m, psi0 - h * d)) / (2 * h)
    assert np.dot(lam0, d) == pytest.approx(fd, rel=1e-6)
    names = {O.G_EPSILON: ("epsilon", "m"), O.G_K: ("k", "p"), O.G_M: ("M", "p"), O.G_DELTA: ("delta", "m"),
             O.G_B2: ("B2", "m")}
    for g, (attr, obj) in names.items():
        base = getattr(m if obj == "m" else p, attr)
        hp = 1e-6 * max(abs(base), 1.0)

        def shifted(s):
            pp = O.Params(**{**p.__dict__})
            mm = O.Model(**{**m.__dict__})
            target = mm if obj == "m" else pp
            setattr(target, attr, getattr(target, attr) + s)
            return loss(pp, mm, psi0)

        fd = (shifted(hp) - shifted(-hp)) / (2 * hp)
        has = not ((g == O.G_EPSILON and m.model != O.LANDAU) or (g in (O.G_DELTA, O.G_B2) and m.model == O.LANDAU))
        if has:
            assert grad[g] == pytest.approx(fd, rel=1e-5, abs=1e-9 * np.abs(w).sum()), attr
        else:
            assert grad[g] == 0.0


def test_general_adjoint_reduces_to_euler_adjoint():
    N, steps = 16, 30
    p = O.Params(dims=2, N=N, dt=0.01, dx=1.0, k=1.2, M=0.9)
    m = O.Model(epsilon=0.9)
    psi0 = O.init_density(2, N, 0.0, 0.3, 0, 99)
    w = np.random.default_rng(8).normal(size=N * N)
    lam_a, g_a = O.euler_adjoint(p, m, psi0, steps, w)
    lam_b, g_b = O.adjoint(O.EULER, p, m, psi0, steps, w)
    assert np.abs(lam_a - lam_b).max() <= 1e-13 * np.abs(lam_a).max()
    np.testing.assert_allclose(g_b[:3], g_a, rtol=1e-12)
    # euler_mobility with constant mobility is the same operator up to round-off
    lam_c, g_c = O.adjoint(O.EULER_MOBILITY, p, m, psi0, steps, w)
    assert np.abs(lam_c - lam_b).max() <= 1e-11 * np.abs(lam_b).max()
    np.testing.assert_allclose(g_c, g_b, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("mob", [O.MOB_FREE_ENERGY, O.MOB_LINEAR, O.MOB_REGULARISED])
def test_euler_mobility_adjoint_dot_product(mob):
    """<J v, u> = <v, J^T u> for one euler_mobility step (J by central differences)."""
    N = 8
    p = O.Params(dims=2, N=N, dt=1e-3, k=1.0, mobility=mob, M=1.2, rho_min=0.2, rho_ref=0.3)
    m = O.Model(O.SIMPLE_WERTHEIM, B2=1.0, delta=50.0, valence=3)
    rng = np.random.default_rng(mob)
    psi = rng.uniform(0.1, 0.6, N * N)
    v, u = rng.normal(size=N * N), rng.normal(size=N * N)
    h = 1e-7
    Jv = (O.run(O.EULER_MOBILITY, p, m, psi + h * v, 1) - O.run(O.EULER_MOBILITY, p, m, psi - h * v, 1)) / (2 * h)
    JTu, _ = O.adjoint_step(O.EULER_MOBILITY, p, m, psi, u)
    assert np.dot(Jv, u) == pytest.approx(np.dot(v, JTu), rel=1e-8)


@pytest.mark.parametrize("model", [O.Model(O.LANDAU, epsilon=0.7),
                                   O.Model(O.SIMPLE_WERTHEIM, B2=1.0, delta=1e4, valence=4)])
def test_free_energy_gradient_by_finite_differences(model):
    N, dims = 8, 3
    p = O.Params(dims=dims, N=N, dx=0.8, k=1.3)
    rho = O.init_density(dims, N, 0.05, 0.05, 0, 7) if model.model != O.LANDAU else \
        O.init_density(dims, N, 0.0, 0.3, 0, 7)
    F, g, dp = O.free_energy_grad(p, model, rho)
    assert F == O.average_free_energy(p, model, rho)
    d = np.random.default_rng(2).normal(size=rho.size)
    h = 1e-6 * np.abs(rho).max()
    fd = (O.average_free_energy(p, model, rho + h * d) - O.average_free_energy(p, model, rho - h * d)) / (2 * h)
    assert np.dot(g, d) == pytest.approx(fd, rel=1e-6)
    for gi, (attr, obj) in {O.G_EPSILON: ("epsilon", "m"), O.G_K: ("k", "p"), O.G_DELTA: ("delta", "m"),
                            O.G_B2: ("B2", "m")}.items():
        base = getattr(model if obj == "m" else p, attr)
        # F is linear in epsilon, k and B2: a large step has no truncation error
        hp = 1e-6 * max(abs(base), 1.0) if attr == "delta" else 0.1
        vals = []
        for s in (hp, -hp):
            pp, mm = O.Params(**{**p.__dict__}), O.Model(**{**model.__dict__})
            t = mm if obj == "m" else pp
            setattr(t, attr, getattr(t, attr) + s)
            vals.append(O.average_free_energy(pp, mm, rho))
        fd = (vals[0] - vals[1]) / (2 * hp)
        assert dp[gi] == pytest.approx(fd, rel=1e-6, abs=1e-12), attr
    assert dp[O.G_M] == 0.0


PS_ADJOINT_CASES = [
    # dims, N, steps, model kwargs, params kwargs, (density, A)
    (2, 16, 12, dict(model=O.LANDAU, epsilon=0.9), dict(dt=0.1, k=1.2, M=0.8), (0.0, 0.3)),
    (2, 16, 10, dict(model=O.LANDAU, epsilon=0.7), dict(dt=0.05, k=0.9, M=1.1, S=0.3, use_dealias=1), (0.0, 0.4)),
    (1, 32, 15, dict(model=O.LANDAU, epsilon=0.8), dict(dt=0.1, k=1.0, M=1.0, dx=0.8), (0.1, 0.3)),
    (3, 8, 8, dict(model=O.LANDAU, epsilon=0.9), dict(dt=0.1, k=1.0, M=1.0, use_dealias=1), (0.0, 0.3)),
    (2, 16, 10, dict(model=O.SIMPLE_WERTHEIM, B2=1.0, delta=50.0, valence=4), dict(dt=1e-3, k=1.0, M=1.0),
     (0.3, 0.2)),
]


@pytest.mark.parametrize("case", range(len(PS_ADJOINT_CASES)))
def test_pseudospectral_adjoint_against_finite_differences(case):
    """dL/drho0 and dL/d(eps, k, M, delta, B2) of L = <w, rho_T> for the
    pseudospectral integrator against central differences of the forward map."""
    dims, N, steps, mk, pk, (rho0, A) = PS_ADJOINT_CASES[case]
    p = O.Params(dims=dims, N=N, **pk)
    m = O.Model(**mk)
    psi0 = O.init_density(dims, N, rho0, A, 0, 777 + case)
    rng = np.random.default_rng(50 + case)
    w, d = rng.normal(size=N ** dims), rng.normal(size=N ** dims)
    lam0, grad = O.adjoint(O.PSEUDOSPECTRAL, p, m, psi0, steps, w)

    def loss(pp, mm, x):
        return float(np.dot(w, O.run(O.PSEUDOSPECTRAL, pp, mm, x, steps)))

    h = 1e-6 * max(1.0, np.abs(psi0).max())
    fd = (loss(p, m, psi0 + h * d) - loss(p, m, psi0 - h * d)) / (2 * h)
    assert np.dot(lam0, d) == pytest.approx(fd, rel=1e-6)
    names = {O.G_EPSILON: ("epsilon", "m"), O.G_K: ("k", "p"), O.G_M: ("M", "p"), O.G_DELTA: ("delta", "m"),
             O.G_B2: ("B2", "m")}
    for g, (attr, obj) in names.items():
        base = getattr(m if obj == "m" else p, attr)
        hp = 1e-6 * max(abs(base), 1.0)

        def shifted(s):
            pp = O.Params(**{**p.__dict__})
            mm = O.Model(**{**m.__dict__})
            target = mm if obj == "m" else pp
            setattr(target, attr, getattr(target, attr) + s)
            return loss(pp, mm, psi0)

        fd = (shifted(hp) - shifted(-hp)) / (2 * hp)
        has = not ((g == O.G_EPSILON and m.model != O.LANDAU) or (g in (O.G_DELTA, O.G_B2) and m.model == O.LANDAU))
        if has:
            assert grad[g] == pytest.approx(fd, rel=1e-5, abs=1e-9 * np.abs(w).sum()), attr
        else:
            assert grad[g] == 0.0


def test_pseudospectral_adjoint_dot_product():
    """<J v, u> = <v, J^T u> for one pseudospectral step, with dealiasing and S."""
    N = 16
    p = O.Params(dims=2, N=N, dt=0.1, k=1.0, M=1.0, S=0.2, use_dealias=1)
    m = O.Model(epsilon=0.9)
    rng = np.random.default_rng(9)
    psi = rng.uniform(-0.5, 0.5, N * N)
    v, u = rng.normal(size=N * N), rng.normal(size=N * N)
    h = 1e-6
    Jv = (O.run(O.PSEUDOSPECTRAL, p, m, psi + h * v, 1) - O.run(O.PSEUDOSPECTRAL, p, m, psi - h * v, 1)) / (2 * h)
    JTu, _ = O.adjoint_step(O.PSEUDOSPECTRAL, p, m, psi, u)
    assert np.dot(Jv, u) == pytest.approx(np.dot(v, JTu), rel=1e-8)
