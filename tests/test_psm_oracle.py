"""pseudospectral_mobility oracle (src/integrators/PseudospectralMobilityCPU.cpp).

The reference integrator needs FFTW3, which is absent here, so its build is
unavailable and it ships no fixtures for this integrator.  The oracle's
restatement is pinned by:
  * reduction: with constant mobility the splitting scheme (_evolve_simple)
    is the pseudospectral integrator (M - M0 = 0), checked against the
    oracle's PseudospectralCPU restatement, itself pinned by the reference's
    fixtures;
  * the spectral operator div(M grad g) against closed-form derivatives of
    band-limited fields (sign of i k, normalisation, axis order);
  * one step of each scheme against an independent numpy restatement (numpy's
    rfftn / irfftn, whose c2r keeps only the real part of the DC and Nyquist
    entries of each last-axis line, as FFTW's does), the GMRES scheme against
    a dense direct solve of A x = b;
  * total mass unchanged (the divergence has no k = 0 component);
  * consistency, fixed by the mathematics rather than by a restatement: on a
    band-limited field with varying mobility, (rho_1 - rho_0) / dt of each
    scheme converges at first order to div(M grad mu_0) (mu_0 in closed form,
    the operator pinned by closed-form derivatives), which a wrong sign of the
    (M - M0) term cannot reach, and the two schemes' steps differ by O(dt^2).
"""
import numpy as np
import pytest

import oracle as O


def kvecs(dims, N, dx):
    """k components per axis (slowest first) on the rfftn grid, as the reference signs them."""
    full = np.fft.fftfreq(N) * N
    half = np.fft.rfftfreq(N) * N
    axes = [full] * (dims - 1) + [half]
    grids = np.meshgrid(*axes, indexing="ij")
    return [2.0 * np.pi * g / (N * dx) for g in grids]


def irfftn(x, shape):
    return np.fft.irfftn(x, s=shape, axes=tuple(range(len(shape))))


def np_div_w_grad(src_hat, w, kc, deal, dims, N):
    shape = (N,) * dims
    div = np.zeros_like(src_hat)
    for d in range(dims):
        gh = 1j * kc[d] * src_hat * deal
        g = irfftn(gh, shape)
        fh = np.fft.rfftn(w.reshape(shape) * g) * deal
        div = div + 1j * kc[d] * fh
    return div


def np_setup(dims, N, dx, dealias):
    kc = kvecs(dims, N, dx)
    k2 = sum(k * k for k in kc)
    k_cut = N * np.pi / (N * dx) * 2.0 / 3.0
    deal = (np.sqrt(k2) <= k_cut).astype(float) if dealias else np.ones_like(k2)
    return kc, k2, deal


def landau_fp(rho, eps):
    return -eps * rho + rho ** 3


def test_simple_scheme_with_constant_mobility_is_pseudospectral():
    for dims, N in ((1, 64), (2, 32), (3, 8)):
        p = O.Params(dims=dims, N=N, dt=0.1, dx=0.9, k=1.0, mobility=O.MOB_CONSTANT, M=1.3, S=0.4)
        m = O.Model(O.LANDAU, epsilon=0.9)
        rho0 = 0.02 * (2 * np.random.default_rng(dims).random(N ** dims) - 1)
        a = O.run_psm(p, m, rho0, 30)
        b = O.run(O.PSEUDOSPECTRAL, p, m, rho0, 30)
        assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max()


@pytest.mark.parametrize("dims,N,dx", [(1, 64, 0.5), (2, 32, 0.7), (3, 16, 1.1)])
def test_div_mob_grad_closed_form(dims, N, dx):
    L = N * dx
    x = np.arange(N) * dx
    grids = np.meshgrid(*([x] * dims), indexing="ij")  # slowest axis first; the last is x
    kx, ky = 2 * np.pi / L, 2 * np.pi / L
    X = grids[-1]
    Y = grids[0] if dims >= 2 else np.zeros_like(X)
    # M = 1 + 0.3 cos(kx x) + 0.2 sin(2 ky y);  g = cos(2 kx x) + sin(3 ky y)
    M = 1 + 0.3 * np.cos(kx * X) + (0.2 * np.sin(2 * ky * Y) if dims >= 2 else 0)
    g = np.cos(2 * kx * X) + (np.sin(3 * ky * Y) if dims >= 2 else 0)
    Mx, My = -0.3 * kx * np.sin(kx * X), (0.4 * ky * np.cos(2 * ky * Y) if dims >= 2 else 0)
    gx, gy = -2 * kx * np.sin(2 * kx * X), (3 * ky * np.cos(3 * ky * Y) if dims >= 2 else 0)
    gxx, gyy = -4 * kx * kx * np.cos(2 * kx * X), (-9 * ky * ky * np.sin(3 * ky * Y) if dims >= 2 else 0)
    want = Mx * gx + M * gxx + (My * gy + M * gyy if dims >= 2 else 0)
    p = O.Params(dims=dims, N=N, dx=dx)
    got = O.psm_div_mob_grad(p, M.ravel(), g.ravel())
    assert np.abs(got - np.asarray(want).ravel()).max() < 1e-11 * np.abs(want).max()


@pytest.mark.parametrize("dealias", [False, True])
@pytest.mark.parametrize("dims,N", [(1, 32), (2, 16), (3, 8)])
def test_simple_step_against_numpy(dims, N, dealias):
    dx, dt, kap, S, eps, Mc, rho_ref = 0.8, 0.05, 1.2, 0.5, 0.9, 1.1, 0.5
    p = O.Params(dims=dims, N=N, dt=dt, dx=dx, k=kap, mobility=O.MOB_LINEAR, M=Mc, rho_ref=rho_ref, S=S,
                 use_dealias=int(dealias))
    m = O.Model(O.LANDAU, epsilon=eps)
    rho0 = 0.5 + 0.2 * (2 * np.random.default_rng(4).random(N ** dims) - 1)
    got = O.run_psm(p, m, rho0, 1)
    shape = (N,) * dims
    kc, k2, deal = np_setup(dims, N, dx, dealias)
    mob = Mc * rho0 / rho_ref
    M0 = max(0.0, mob.max())
    rh = np.fft.rfftn(rho0.reshape(shape))
    fh = np.fft.rfftn(landau_fp(rho0, eps).reshape(shape))
    muh = (fh + 2 * kap * k2 * rh) * deal
    div = np_div_w_grad(muh, mob - M0, kc, deal, dims, N)
    den = 1 + dt * M0 * (S * k2 + 2 * kap * k2 * k2)
    new = (rh + dt * div - dt * M0 * k2 * (fh - S * rh)) / den
    want = irfftn(new, shape).ravel()
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("dealias", [False, True])
@pytest.mark.parametrize("dims,N", [(1, 32), (2, 8), (3, 4)])
def test_gmres_step_against_dense_solve(dims, N, dealias):
    dx, dt, kap, S, eps, Mc, rho_min = 0.9, 0.2, 1.0, 0.3, 0.9, 1.0, 0.2
    p = O.Params(dims=dims, N=N, dt=dt, dx=dx, k=kap, mobility=O.MOB_REGULARISED, M=Mc, rho_min=rho_min, S=S,
                 use_dealias=int(dealias))
    m = O.Model(O.LANDAU, epsilon=eps)
    n = N ** dims
    rho0 = 0.5 + 0.3 * (2 * np.random.default_rng(5).random(n) - 1)
    iters = []
    got = O.run_psm(p, m, rho0, 1, use_gmres=True, restart=30, max_iter=400, tol=1e-14, iters=iters)
    shape = (N,) * dims
    kc, k2, deal = np_setup(dims, N, dx, dealias)
    mob = Mc * rho0 / (rho0 + rho_min)

    def apply_A(x):
        q = (S + 2 * kap * k2) * np.fft.rfftn(x.reshape(shape)) * deal
        return x - dt * irfftn(np_div_w_grad(q, mob, kc, deal, dims, N), shape).ravel()

    A = np.stack([apply_A(e) for e in np.eye(n)], axis=1)
    gh = np.fft.rfftn((landau_fp(rho0, eps) - S * rho0).reshape(shape))
    b = rho0 + dt * irfftn(np_div_w_grad(gh, mob, kc, deal, dims, N), shape).ravel()
    want = np.linalg.solve(A, b)
    assert 0 < iters[0] <= 400
    assert np.abs(got - want).max() < 1e-11 * np.abs(want).max()


@pytest.mark.parametrize("use_gmres", [False, True])
def test_mass_is_conserved(use_gmres):
    p = O.Params(dims=2, N=32, dt=0.05, dx=1.0, k=1.0, mobility=O.MOB_REGULARISED, M=1.0, rho_min=0.1, S=0.2,
                 use_dealias=1)
    m = O.Model(O.LANDAU, epsilon=0.9)
    rho0 = 0.4 + 0.1 * (2 * np.random.default_rng(8).random(32 * 32) - 1)
    out = O.run_psm(p, m, rho0, 20, use_gmres=use_gmres)
    assert abs(out.sum() - rho0.sum()) < 1e-12 * rho0.sum()
    assert np.abs(out - rho0).max() > 1e-4


def smooth_state(N, dx, eps, kap):
    """A band-limited field rho0 (a few low modes around 0.5) with its chemical
    potential mu0 = f'(rho0) - 2 kap lap(rho0) in closed form (f' = -eps rho + rho^3
    has three times the modes, still resolved at N = 32)."""
    L = N * dx
    k = 2 * np.pi / L
    x = np.arange(N) * dx
    Y, X = np.meshgrid(x, x, indexing="ij")  # [y, x]: x fastest, as idx = x + N y
    rho = 0.5 + 0.1 * np.cos(k * X) + 0.05 * np.sin(2 * k * Y) + 0.03 * np.cos(k * X + k * Y)
    lap = -k * k * 0.1 * np.cos(k * X) - 4 * k * k * 0.05 * np.sin(2 * k * Y) - 2 * k * k * 0.03 * np.cos(k * X + k * Y)
    mu = -eps * rho + rho ** 3 - 2 * kap * lap
    return rho.ravel(), mu.ravel()


@pytest.mark.parametrize("use_gmres", [False, True])
def test_step_converges_to_div_mobility_grad_mu(use_gmres):
    """Consistency of each scheme with d rho / dt = div(M(rho) grad mu): as dt -> 0,
    (rho_1 - rho_0) / dt tends to div(M grad mu_0) at first order (the error halves
    with dt), mu_0 in closed form and div(M grad .) the operator pinned above by
    closed-form derivatives.  With the sign of the (M - M0) term flipped the limit
    would be div((2 M0 - M) grad mu_0), which this check tells apart."""
    N, dx, eps, kap, S, Mc, rho_min = 32, 1.0, 0.9, 1.0, 0.5, 1.0, 0.2
    rho0, mu0 = smooth_state(N, dx, eps, kap)
    m = O.Model(O.LANDAU, epsilon=eps)
    base = O.Params(dims=2, N=N, dx=dx, k=kap, mobility=O.MOB_REGULARISED, M=Mc, rho_min=rho_min, S=S)
    mob = Mc * rho0 / (rho0 + rho_min)
    want = O.psm_div_mob_grad(base, mob, mu0)
    flipped = O.psm_div_mob_grad(base, 2 * mob.max() - mob, mu0)
    errs = []
    for dt in (4e-3, 2e-3, 1e-3):
        p = O.Params(dims=2, N=N, dt=dt, dx=dx, k=kap, mobility=O.MOB_REGULARISED, M=Mc, rho_min=rho_min, S=S)
        rho1 = O.run_psm(p, m, rho0, 1, use_gmres=use_gmres, restart=40, max_iter=400, tol=1e-14)
        errs.append(np.abs((rho1 - rho0) / dt - want).max())
    print(f"use_gmres={use_gmres}: errors {errs}, |limit| {np.abs(want).max():.3e}, "
          f"|flipped - limit| {np.abs(flipped - want).max():.3e}")
    assert 0.4 < errs[1] / errs[0] < 0.6 and 0.4 < errs[2] / errs[1] < 0.6
    assert errs[2] < 0.02 * np.abs(want).max()
    assert errs[2] < 0.01 * np.abs(flipped - want).max()


def test_splitting_and_gmres_agree_to_second_order():
    """Both schemes are first-order consistent with the same equation, so one
    step of each from the same state differs by O(dt^2): the difference drops
    by four when dt halves."""
    N, dx, eps, kap, S, Mc, rho_min = 32, 1.0, 0.9, 1.0, 0.5, 1.0, 0.2
    rho0, _ = smooth_state(N, dx, eps, kap)
    m = O.Model(O.LANDAU, epsilon=eps)
    diffs = []
    for dt in (4e-3, 2e-3, 1e-3):
        p = O.Params(dims=2, N=N, dt=dt, dx=dx, k=kap, mobility=O.MOB_REGULARISED, M=Mc, rho_min=rho_min, S=S)
        a = O.run_psm(p, m, rho0, 1)
        b = O.run_psm(p, m, rho0, 1, use_gmres=True, restart=40, max_iter=400, tol=1e-14)
        diffs.append(np.abs(a - b).max())
    print("splitting - gmres:", diffs)
    assert 0.2 < diffs[1] / diffs[0] < 0.3 and 0.2 < diffs[2] / diffs[1] < 0.3
