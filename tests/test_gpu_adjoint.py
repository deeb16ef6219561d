"""Reverse mode of the euler and euler_mobility integrators on the GPU
(ch_sys_record / ch_sys_adjoint / ch_sys_free_energy_grad).

Tolerances: FP64 dL/dpsi0 within 1e-10 relative L-inf and dL/d(eps, k, M,
delta, B2) within 1e-9 relative of the oracle's FP64 adjoint
(oracle/ch_oracle.c: och_adjoint); central finite differences of the GPU
forward map (h = 1e-6) within 1e-6 (1e-5 for chained losses); every checkpoint
spacing gives the same dL/dpsi0 bits; FP32 fields within 2e-3 of FP64.
"""
import numpy as np
import pytest

import oracle as O
from helpers import rel_linf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(ch):
    c = ch.Context(0)
    yield c
    c.close()


def gpu_adjoint(ch, ctx, kw, psi0, steps, w, ckpt=100, precision="fp64", eps=0.9):
    s = ch.System(ctx, ch.make_desc(integrator="euler", epsilon=eps, precision=precision, **kw))
    s.upload(psi0)
    s.record(steps, ckpt)
    psiT = s.download()
    lam, g = s.adjoint(w)
    s.close()
    return psiT, lam, g


@pytest.mark.parametrize("dims,N,steps,ckpt", [(2, 64, 60, 16), (2, 32, 40, 0), (1, 128, 80, 25), (3, 16, 30, 7),
                                              (2, 128, 45, 10), (2, 256, 30, 7), (2, 512, 21, 0)])
def test_adjoint_against_oracle(ch, ctx, dims, N, steps, ckpt):
    kw = dict(dims=dims, N=N, dt=0.01 if dims < 3 else 0.002, dx=1.0, k=1.2, M=0.9)
    psi0 = O.init_density(dims, N, 0.0, 0.3, 0, 123456)
    w = np.random.default_rng(dims).normal(size=N ** dims)
    psiT, lam, g = gpu_adjoint(ch, ctx, kw, psi0, steps, w, ckpt)
    ref_lam, ref_g = O.adjoint(O.EULER, O.Params(**kw), O.Model(epsilon=0.9), psi0, steps, w)
    ref_T = O.run(O.EULER, O.Params(**kw), O.Model(epsilon=0.9), psi0, steps)
    assert rel_linf(psiT, ref_T) < 1e-12
    assert rel_linf(lam, ref_lam) < 1e-10
    np.testing.assert_allclose(g, ref_g, rtol=1e-10)
    assert g[O.G_DELTA] == 0.0 and g[O.G_B2] == 0.0


def test_checkpoint_spacing_invariance(ch, ctx):
    kw = dict(dims=2, N=64, dt=0.01, dx=1.0, k=1.0, M=1.0)
    psi0 = O.init_density(2, 64, 0.0, 0.3, 0, 5)
    w = np.random.default_rng(0).normal(size=64 * 64)
    base = None
    for ckpt in (0, 1, 3, 16, 37, 100):
        _, lam, g = gpu_adjoint(ch, ctx, kw, psi0, 100, w, ckpt)
        if base is None:
            base = (lam, g)
        else:
            assert np.array_equal(lam, base[0]), ckpt
            np.testing.assert_allclose(g, base[1], rtol=1e-12)


def test_adjoint_finite_differences(ch, ctx):
    """Central differences (h = 1e-6) of L = <w, psi_T> through the GPU forward."""
    N, steps = 64, 200
    kw = dict(dims=2, N=N, dt=0.01, dx=1.0)
    psi0 = O.init_density(2, N, 0.0, 0.3, 0, 123456)
    w = np.random.default_rng(1).normal(size=N * N)
    d = np.random.default_rng(2).normal(size=N * N)
    params = dict(eps=0.9, k=1.2, M=1.0)

    def loss(x, eps, k, M):
        s = ch.System(ctx, ch.make_desc(integrator="euler", epsilon=eps, k=k, M=M, **kw))
        s.upload(x)
        s.evolve(steps)
        v = float(np.dot(w, s.download()))
        s.close()
        return v

    s = ch.System(ctx, ch.make_desc(integrator="euler", epsilon=0.9, k=1.2, M=1.0, **kw))
    s.upload(psi0)
    s.record(steps, 50)
    lam, g = s.adjoint(w)
    s.close()
    h = 1e-6
    fd = (loss(psi0 + h * d, **params) - loss(psi0 - h * d, **params)) / (2 * h)
    assert np.dot(lam, d) == pytest.approx(fd, rel=1e-6)
    for i, name in enumerate(("eps", "k", "M")):
        up, dn = dict(params), dict(params)
        up[name] += h
        dn[name] -= h
        fd = (loss(psi0, **up) - loss(psi0, **dn)) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=1e-6), name


# N = 64 is one whole-grid region; N = 256 runs the tiled reverse kernel (64
# blocks, 8 backward steps per launch) with its ring of tape buffers
@pytest.mark.parametrize("N", [64, 256])
def test_fp32_adjoint_against_fp64(ch, ctx, N):
    kw = dict(dims=2, N=N, dt=0.01, dx=1.0, k=1.0, M=1.0)
    psi0 = O.init_density(2, N, 0.0, 0.3, 0, 9)
    w = np.random.default_rng(3).normal(size=N * N)
    _, lam64, g64 = gpu_adjoint(ch, ctx, kw, psi0, 100, w, 25)
    _, lam32, g32 = gpu_adjoint(ch, ctx, kw, psi0, 100, w, 25, precision="fp32")
    err = rel_linf(lam32, lam64)
    nz = g64 != 0
    print(f"fp32 adjoint rel L-inf vs fp64 = {err:.3e}, grads {np.abs(g32[nz] / g64[nz] - 1).max():.3e}")
    assert err < 2e-3
    np.testing.assert_allclose(g32, g64, rtol=2e-3)


def test_wertheim_adjoint_against_oracle(ch, ctx):
    """SimpleWertheim f'' in the adjoint; the delta and B2 gradients (no epsilon)."""
    kw = dict(dims=2, N=32, dt=1e-4, dx=1.0, k=1.0, M=1.0)
    mk = dict(free_energy="simple_wertheim", B2=1.0, delta=1e4, valence=4)
    psi0 = O.init_density(2, 32, 0.05, 0.05, 0, 7)
    w = np.random.default_rng(4).normal(size=32 * 32)
    s = ch.System(ctx, ch.make_desc(integrator="euler", **mk, **kw))
    s.upload(psi0)
    s.record(50, 10)
    lam, g = s.adjoint(w)
    s.close()
    om = O.Model(O.SIMPLE_WERTHEIM, B2=1.0, delta=1e4, valence=4)
    ref_lam, ref_g = O.adjoint(O.EULER, O.Params(**kw), om, psi0, 50, w)
    assert rel_linf(lam, ref_lam) < 1e-10
    assert g[O.G_EPSILON] == 0.0
    np.testing.assert_allclose(g, ref_g, rtol=1e-9)


def test_autograd_gradcheck(ch):
    import torch
    from cahn_hilliard_amd.autograd import euler_evolve
    N = 16
    psi0 = torch.tensor(O.init_density(2, N, 0.0, 0.3, 0, 3), dtype=torch.float64, device="cuda", requires_grad=True)
    eps = torch.tensor(0.9, dtype=torch.float64, device="cuda", requires_grad=True)
    k = torch.tensor(1.1, dtype=torch.float64, device="cuda", requires_grad=True)
    M = torch.tensor(0.8, dtype=torch.float64, device="cuda", requires_grad=True)
    fn = lambda a, b, c, d: euler_evolve(a, b, c, d, 20, 5, dims=2, N=N, dt=0.01)
    assert torch.autograd.gradcheck(fn, (psi0, eps, k, M), eps=1e-6, atol=1e-8, rtol=1e-6)
    # composition with a torch loss
    out = fn(psi0, eps, k, M)
    (out ** 2).sum().backward()
    assert psi0.grad is not None and torch.isfinite(psi0.grad).all()


@pytest.mark.parametrize("integrator,precision,tol", [("euler", "fp64", 1e-10), ("pseudospectral", "fp64", 1e-10),
                                                      ("euler", "fp32", 2e-4)])
def test_structure_factor_and_loss_against_oracle(ch, ctx, integrator, precision, tol):
    """Device S(q) and its loss gradient (utils/sq_from_traj.py:72-103) against the
    oracle's restatement, which is itself pinned to numpy.fft."""
    N = 64
    psi = O.init_density(2, N, 0.0, 0.3, 0, 11)
    s = ch.System(ctx, ch.make_desc(dims=2, N=N, integrator=integrator, precision=precision, dt=0.01))
    s.upload(psi)
    sq = s.structure_factor()
    ref = O.radial_sq(psi.reshape(N, N))
    assert len(sq) == len(ref) == 46
    assert np.abs(sq - ref).max() <= tol * ref.max()
    target = O.radial_sq(O.init_density(2, N, 0.0, 0.3, 0, 12).reshape(N, N))
    loss, g = s.sq_loss(target)
    ref_loss, ref_g = O.sq_loss(psi.reshape(N, N), target)
    s.close()
    assert loss == pytest.approx(ref_loss, rel=10 * tol)
    assert np.abs(g - ref_g.ravel()).max() <= 10 * tol * np.abs(ref_g).max()


def test_c4_loss_gradient_chain(ch, ctx):
    """C4 shape at small size: S(q) loss of the final field, its gradient seeded into
    the reverse sweep, gradients wrt eps / k checked by central differences."""
    N, T = 32, 100
    kw = dict(dims=2, N=N, dt=0.01, dx=1.0, M=1.0)
    psi0 = O.init_density(2, N, 0.0, 0.3, 0, 123456)
    tgt_sys = ch.System(ctx, ch.make_desc(integrator="euler", epsilon=0.8, k=1.0, **kw))
    tgt_sys.upload(psi0)
    tgt_sys.evolve(T)
    target = tgt_sys.structure_factor()
    tgt_sys.close()

    def loss_at(eps, k):
        s = ch.System(ctx, ch.make_desc(integrator="euler", epsilon=eps, k=k, **kw))
        s.upload(psi0)
        s.evolve(T)
        v = s.sq_loss(target, grad=False)
        s.close()
        return v

    s = ch.System(ctx, ch.make_desc(integrator="euler", epsilon=0.9, k=1.2, **kw))
    s.upload(psi0)
    s.record(T, 25)
    loss, g = s.sq_loss(target)
    lam, grads = s.adjoint(g)
    s.close()
    h = 1e-6
    fd_eps = (loss_at(0.9 + h, 1.2) - loss_at(0.9 - h, 1.2)) / (2 * h)
    fd_k = (loss_at(0.9, 1.2 + h) - loss_at(0.9, 1.2 - h)) / (2 * h)
    assert grads[0] == pytest.approx(fd_eps, rel=1e-5)
    assert grads[1] == pytest.approx(fd_k, rel=1e-5)


def test_c4_full_size_finite_differences(ch, ctx):
    """C4 at its real size: Landau 2D N=256, dt=0.01, 5000 steps, checkpoints
    every 100; S(q) loss against the target of a run at eps=0.8, k=1.0; the
    gradients at eps=0.9, k=1.2 wrt eps, k and the initial field (along one
    random direction) against central differences with h=1e-6: within 1e-6
    relative."""
    N, T = 256, 5000
    kw = dict(dims=2, N=N, dt=0.01, dx=1.0, M=1.0)
    psi0 = O.init_density(2, N, 0.0, 0.01, 0, 123456)
    tgt_sys = ch.System(ctx, ch.make_desc(integrator="euler", epsilon=0.8, k=1.0, **kw))
    tgt_sys.upload(psi0)
    tgt_sys.evolve(T)
    target = tgt_sys.structure_factor()
    tgt_sys.close()

    def loss_at(eps, k, start):
        s = ch.System(ctx, ch.make_desc(integrator="euler", epsilon=eps, k=k, **kw))
        s.upload(start)
        s.evolve(T)
        v = s.sq_loss(target, grad=False)
        s.close()
        return v

    s = ch.System(ctx, ch.make_desc(integrator="euler", epsilon=0.9, k=1.2, **kw))
    s.upload(psi0)
    s.record(T, 100)
    loss, g = s.sq_loss(target)
    lam0, grads = s.adjoint(g)
    s.close()
    h = 1e-6
    fd_eps = (loss_at(0.9 + h, 1.2, psi0) - loss_at(0.9 - h, 1.2, psi0)) / (2 * h)
    fd_k = (loss_at(0.9, 1.2 + h, psi0) - loss_at(0.9, 1.2 - h, psi0)) / (2 * h)
    v = np.random.default_rng(7).standard_normal(N * N)
    v /= np.linalg.norm(v)
    fd_psi = (loss_at(0.9, 1.2, psi0 + h * v) - loss_at(0.9, 1.2, psi0 - h * v)) / (2 * h)
    ad_psi = float(np.dot(lam0, v))
    print(f"C4 full size: loss {loss:.6e}; d/deps {grads[0]:.9e} vs fd {fd_eps:.9e}; d/dk {grads[1]:.9e} vs fd "
          f"{fd_k:.9e}; d/dpsi0.v {ad_psi:.9e} vs fd {fd_psi:.9e}")
    assert grads[0] == pytest.approx(fd_eps, rel=1e-6)
    assert grads[1] == pytest.approx(fd_k, rel=1e-6)
    assert ad_psi == pytest.approx(fd_psi, rel=1e-6)


EM_CASES = [
    # dims, N, steps, ckpt, model kwargs (desc), mobility kwargs (desc), init (density, A), dt
    (3, 16, 30, 7, dict(free_energy="simple_wertheim", B2=1.0, delta=1e4, valence=4),
     dict(mobility="linear", rho_ref=0.05), (0.05, 0.05), 1e-4),
    (2, 64, 40, 16, dict(free_energy="simple_wertheim", B2=-2.0, delta=30.0, valence=2),
     dict(mobility="free_energy"), (0.3, 0.2), 2e-4),
    (2, 32, 50, 0, dict(epsilon=0.9), dict(mobility="regularised", rho_min=0.5), (0.5, 0.2), 0.01),
    (1, 128, 60, 25, dict(epsilon=0.8), dict(mobility="constant"), (0.0, 0.3), 0.01),
    (2, 128, 24, 5, dict(epsilon=0.9), dict(mobility="linear", rho_ref=0.5), (0.5, 0.1), 0.01),
]


def _oracle_model(mk):
    if mk.get("free_energy") == "simple_wertheim":
        return O.Model(O.SIMPLE_WERTHEIM, B2=mk["B2"], delta=mk["delta"], valence=mk["valence"])
    return O.Model(O.LANDAU, epsilon=mk["epsilon"])


def _oracle_params(dims, N, dt, k, M, mob):
    table = {"constant": O.MOB_CONSTANT, "free_energy": O.MOB_FREE_ENERGY, "regularised": O.MOB_REGULARISED,
             "linear": O.MOB_LINEAR}
    return O.Params(dims=dims, N=N, dt=dt, dx=1.0, k=k, M=M, mobility=table[mob["mobility"]],
                    rho_min=mob.get("rho_min", 0.0), rho_ref=mob.get("rho_ref", 1.0))


@pytest.mark.parametrize("case", range(len(EM_CASES)))
def test_euler_mobility_adjoint_against_oracle(ch, ctx, case):
    dims, N, steps, ckpt, mk, mob, (rho0, A), dt = EM_CASES[case]
    psi0 = O.init_density(dims, N, rho0, A, 0, 123456 + case)
    w = np.random.default_rng(10 + case).normal(size=N ** dims)
    desc = ch.make_desc(dims=dims, N=N, integrator="euler_mobility", dt=dt, k=1.1, M=0.9, **mk, **mob)
    s = ch.System(ctx, desc)
    s.upload(psi0)
    s.record(steps, ckpt)
    psiT = s.download()
    lam, g = s.adjoint(w)
    s.close()
    p, m = _oracle_params(dims, N, dt, 1.1, 0.9, mob), _oracle_model(mk)
    ref_T = O.run(O.EULER_MOBILITY, p, m, psi0, steps)
    ref_lam, ref_g = O.adjoint(O.EULER_MOBILITY, p, m, psi0, steps, w)
    print(f"case {case}: psiT {rel_linf(psiT, ref_T):.2e} lam {rel_linf(lam, ref_lam):.2e} grads {g} ref {ref_g}")
    assert rel_linf(psiT, ref_T) < 1e-12
    assert rel_linf(lam, ref_lam) < 1e-10
    np.testing.assert_allclose(g, ref_g, rtol=1e-9, atol=1e-12 * np.abs(ref_g).max())


@pytest.mark.parametrize("case", ["wertheim_linear", "landau_regularised", "landau_constant"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_euler_mobility_3d_split_adjoint_equals_tile(ch, ctx, monkeypatch, case, precision):
    """The 3D euler_mobility reverse step in launches of one cell per thread (nu
    and eta, then the update, with mu and the mobilities of each state kept on
    the tape; the default for SimpleWertheim), with a checkpointed record (the
    two-launch window recompute) and with the whole tape, against the LDS tile
    adjoint: dL/dpsi0 bit for bit, the parameter gradients (FP64 sums in
    another order) within 1e-12."""
    N, steps, ckpt = 32, 24, 5
    mk, mob, init, dt = {
        "wertheim_linear": (dict(free_energy="simple_wertheim", B2=1.0, delta=1e4, valence=4),
                            dict(mobility="linear", rho_ref=0.05), (0.05, 0.05), 1e-4),
        "landau_regularised": (dict(epsilon=0.9), dict(mobility="regularised", rho_min=0.5), (0.5, 0.2), 1e-3),
        "landau_constant": (dict(epsilon=0.9), dict(mobility="constant"), (0.0, 0.3), 1e-3)}[case]
    psi0 = O.init_density(3, N, init[0], init[1], 0, 777)
    w = np.random.default_rng(5).normal(size=N ** 3)
    out = {}
    for t3, ck in (("1", ckpt), ("3", ckpt), ("3", 0)):  # 0: the whole tape recorded, no recompute
        monkeypatch.setenv("CH_TILE3D", t3)
        s = ch.System(ctx, ch.make_desc(dims=3, N=N, integrator="euler_mobility", dt=dt, k=1.1, M=0.9,
                                        precision=precision, **mk, **mob))
        s.upload(psi0)
        s.record(steps, ck)
        psiT = s.download()
        lam, g = s.adjoint(w)
        s.close()
        out[t3, ck] = (psiT, lam, np.asarray(g))
    ref = out["1", ckpt]
    for key in (("3", ckpt), ("3", 0)):
        print(f"{case} {precision} {key}: grads {out[key][2]} vs tile {ref[2]}")
        assert np.all(np.isfinite(out[key][1]))
        assert np.array_equal(ref[0], out[key][0]), key
        assert np.array_equal(ref[1], out[key][1]), key
        np.testing.assert_allclose(out[key][2], ref[2], rtol=1e-12, atol=1e-14 * np.abs(ref[2]).max())


def test_euler_mobility_checkpoint_invariance(ch, ctx):
    dims, N, steps = 2, 64, 40
    psi0 = O.init_density(dims, N, 0.3, 0.2, 0, 3)
    w = np.random.default_rng(0).normal(size=N * N)
    base = None
    for ckpt in (0, 1, 7, 40):
        s = ch.System(ctx, ch.make_desc(dims=dims, N=N, integrator="euler_mobility", dt=2e-4, free_energy="simple_wertheim",
                                        B2=-2.0, delta=30.0, valence=2, mobility="free_energy"))
        s.upload(psi0)
        s.record(steps, ckpt)
        lam, g = s.adjoint(w)
        s.close()
        if base is None:
            base = (lam, g)
        else:
            assert np.array_equal(lam, base[0]), ckpt
            np.testing.assert_allclose(g, base[1], rtol=1e-12)


@pytest.mark.parametrize("dims,N,model,precision", [(3, 16, "simple_wertheim", "fp64"), (2, 64, "landau", "fp64"),
                                                    (2, 64, "simple_wertheim", "fp32"), (1, 256, "landau", "fp32")])
def test_free_energy_grad_against_oracle(ch, ctx, dims, N, model, precision):
    mk = dict(free_energy="simple_wertheim", B2=1.0, delta=1e4, valence=4) if model == "simple_wertheim" else \
        dict(epsilon=0.7)
    rho = O.init_density(dims, N, 0.05, 0.05, 0, 5) if model == "simple_wertheim" else \
        O.init_density(dims, N, 0.0, 0.3, 0, 5)
    s = ch.System(ctx, ch.make_desc(dims=dims, N=N, integrator="euler", dx=0.8, k=1.3, precision=precision, **mk))
    s.upload(rho)
    F, g, dp = s.free_energy_grad()
    Fobs = s.average_free_energy()
    s.close()
    p = O.Params(dims=dims, N=N, dx=0.8, k=1.3)
    if precision == "fp32":
        rho = rho.astype(np.float32).astype(np.float64)  # the field the device holds
    rF, rg, rdp = O.free_energy_grad(p, _oracle_model(mk), rho)
    tol = 1e-12 if precision == "fp64" else 1e-6
    assert F == pytest.approx(rF, rel=tol, abs=1e-14)
    assert F == pytest.approx(Fobs, rel=1e-12, abs=1e-14)
    assert np.abs(g - rg).max() <= tol * np.abs(rg).max()
    np.testing.assert_allclose(dp, rdp, rtol=tol, atol=1e-15)


def test_c5_free_energy_delta_gradient_chain(ch, ctx):
    """C5 shape at small size: SimpleWertheim 3D euler_mobility with linear
    mobility; d F(psi_T) / d delta = explicit part + adjoint part, against
    central differences of the GPU forward map."""
    N, T = 16, 40
    kw = dict(dims=3, N=N, integrator="euler_mobility", dt=1e-4, k=1.0, M=1.0, free_energy="simple_wertheim", B2=1.0,
              valence=4, mobility="linear", rho_ref=0.05)
    psi0 = O.init_density(3, N, 0.05, 0.05, 0, 123456)
    delta = 1e4

    def F_at(dl):
        s = ch.System(ctx, ch.make_desc(delta=dl, **kw))
        s.upload(psi0)
        s.evolve(T)
        v = s.average_free_energy()
        s.close()
        return v

    s = ch.System(ctx, ch.make_desc(delta=delta, **kw))
    s.upload(psi0)
    s.record(T, 10)
    F, g, dp = s.free_energy_grad()
    _, grads = s.adjoint(g)
    s.close()
    total = dp[ch.GRAD_DELTA] + grads[ch.GRAD_DELTA]
    h = 1e-4 * delta
    fd = (F_at(delta + h) - F_at(delta - h)) / (2 * h)
    print(f"dF/ddelta = {total:.10e} (explicit {dp[ch.GRAD_DELTA]:.3e}), FD {fd:.10e}")
    assert total == pytest.approx(fd, rel=1e-5)


def test_c5_full_size_delta_gradient(ch, ctx):
    """C5 at its size: SimpleWertheim 3D N=128 euler_mobility, valence 4, B2 = 1,
    delta = 1e4, k = 1, dt = 1e-4, M = rho / 0.05, rho0 = 0.05 (1 + U[-0.05, 0.05))
    from drand48 seed 123456, 400 steps (checkpoints every 50): d F(rho_T) / d delta
    (explicit + adjoint) against central differences of the GPU forward map with
    h = 1e-6 delta."""
    N, T = 128, 400
    kw = dict(dims=3, N=N, integrator="euler_mobility", dt=1e-4, k=1.0, M=1.0, free_energy="simple_wertheim", B2=1.0,
              valence=4, mobility="linear", rho_ref=0.05)
    from cahn_hilliard_amd.init import initial_density
    psi0 = initial_density(3, N, 0.05, 0.05, 0, 123456)
    delta = 1e4

    def F_at(dl):
        s = ch.System(ctx, ch.make_desc(delta=dl, **kw))
        s.upload(psi0)
        s.evolve(T)
        v = s.average_free_energy()
        s.close()
        return v

    s = ch.System(ctx, ch.make_desc(delta=delta, **kw))
    s.upload(psi0)
    s.record(T, 50)
    F, g, dp = s.free_energy_grad()
    _, grads = s.adjoint(g)
    s.close()
    total = dp[ch.GRAD_DELTA] + grads[ch.GRAD_DELTA]
    h = 1e-6 * delta
    fd = (F_at(delta + h) - F_at(delta - h)) / (2 * h)
    print(f"C5 N=128: dF/ddelta = {total:.10e} (explicit {dp[ch.GRAD_DELTA]:.3e}), FD {fd:.10e}")
    assert total == pytest.approx(fd, rel=1e-5)


PS_CASES = [
    # dims, N, steps, ckpt, desc kwargs, (density, A)
    (2, 64, 20, 7, dict(epsilon=0.9, dt=0.1, k=1.2, M=0.8), (0.0, 0.3)),
    (2, 128, 12, 0, dict(epsilon=0.7, dt=0.05, k=0.9, M=1.1, S=0.3, use_dealias=True), (0.0, 0.4)),
    (1, 256, 30, 0, dict(epsilon=0.8, dt=0.1, dx=0.8), (0.1, 0.3)),
    (3, 16, 10, 4, dict(epsilon=0.9, dt=0.1, use_dealias=True), (0.0, 0.3)),
    (2, 64, 15, 5, dict(free_energy="simple_wertheim", B2=1.0, delta=50.0, valence=4, dt=1e-3), (0.3, 0.2)),
]


@pytest.mark.parametrize("case", range(len(PS_CASES)))
def test_pseudospectral_adjoint_against_oracle(ch, ctx, case):
    dims, N, steps, ckpt, kw, (rho0, A) = PS_CASES[case]
    psi0 = O.init_density(dims, N, rho0, A, 0, 99 + case)
    w = np.random.default_rng(70 + case).normal(size=N ** dims)
    s = ch.System(ctx, ch.make_desc(dims=dims, N=N, integrator="pseudospectral", **kw))
    s.upload(psi0)
    s.record(steps, ckpt)
    psiT = s.download()
    lam, g = s.adjoint(w)
    s.close()
    okw = {k: v for k, v in kw.items() if k in ("dt", "dx", "k", "M", "S")}
    p = O.Params(dims=dims, N=N, use_dealias=int(kw.get("use_dealias", False)), **okw)
    m = O.Model(O.SIMPLE_WERTHEIM, B2=kw["B2"], delta=kw["delta"], valence=kw["valence"]) \
        if kw.get("free_energy") == "simple_wertheim" else O.Model(O.LANDAU, epsilon=kw["epsilon"])
    ref_T = O.run(O.PSEUDOSPECTRAL, p, m, psi0, steps)
    ref_lam, ref_g = O.adjoint(O.PSEUDOSPECTRAL, p, m, psi0, steps, w)
    print(f"case {case}: psiT {rel_linf(psiT, ref_T):.2e} lam {rel_linf(lam, ref_lam):.2e} g {g} ref {ref_g}")
    assert rel_linf(psiT, ref_T) < 1e-10
    assert rel_linf(lam, ref_lam) < 1e-10
    np.testing.assert_allclose(g, ref_g, rtol=1e-9, atol=1e-12 * np.abs(ref_g).max())


def test_pseudospectral_checkpoint_invariance(ch, ctx):
    N, steps = 64, 24
    psi0 = O.init_density(2, N, 0.0, 0.3, 0, 17)
    w = np.random.default_rng(5).normal(size=N * N)
    base = None
    for ckpt in (0, 1, 5, 24):
        s = ch.System(ctx, ch.make_desc(dims=2, N=N, integrator="pseudospectral", dt=0.1, use_dealias=True))
        s.upload(psi0)
        s.record(steps, ckpt)
        lam, g = s.adjoint(w)
        s.close()
        if base is None:
            base = (lam, g)
        else:
            assert np.array_equal(lam, base[0]), ckpt
            np.testing.assert_allclose(g, base[1], rtol=1e-12)


def test_pseudospectral_adjoint_finite_differences(ch, ctx):
    """Central differences (h = 1e-6) of L = <w, psi_T> through the GPU forward."""
    N, steps = 64, 40
    kw = dict(dims=2, N=N, integrator="pseudospectral", dt=0.1, dx=1.0)
    psi0 = O.init_density(2, N, 0.0, 0.3, 0, 123456)
    w = np.random.default_rng(11).normal(size=N * N)
    d = np.random.default_rng(12).normal(size=N * N)
    params = dict(epsilon=0.9, k=1.2, M=1.0)

    def loss(x, **pp):
        s = ch.System(ctx, ch.make_desc(**pp, **kw))
        s.upload(x)
        s.evolve(steps)
        v = float(np.dot(w, s.download()))
        s.close()
        return v

    s = ch.System(ctx, ch.make_desc(**params, **kw))
    s.upload(psi0)
    s.record(steps, 10)
    lam, g = s.adjoint(w)
    s.close()
    h = 1e-6
    fd = (loss(psi0 + h * d, **params) - loss(psi0 - h * d, **params)) / (2 * h)
    assert np.dot(lam, d) == pytest.approx(fd, rel=1e-6)
    for i, name in ((ch.GRAD_EPSILON, "epsilon"), (ch.GRAD_K, "k"), (ch.GRAD_M, "M")):
        up, dn = dict(params), dict(params)
        up[name] += h
        dn[name] -= h
        fd = (loss(psi0, **up) - loss(psi0, **dn)) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=1e-6), name


def test_pseudospectral_fp32_adjoint_against_fp64(ch, ctx):
    N, steps = 64, 20
    psi0 = O.init_density(2, N, 0.0, 0.3, 0, 21)
    w = np.random.default_rng(13).normal(size=N * N)
    out = {}
    for prec in ("fp64", "fp32"):
        s = ch.System(ctx, ch.make_desc(dims=2, N=N, integrator="pseudospectral", precision=prec, dt=0.1))
        s.upload(psi0)
        s.record(steps, 5)
        out[prec] = s.adjoint(w)
        s.close()
    err = rel_linf(out["fp32"][0], out["fp64"][0])
    print(f"pseudospectral fp32 adjoint rel L-inf vs fp64 = {err:.3e}")
    assert err < 1e-3
    np.testing.assert_allclose(out["fp32"][1][:3], out["fp64"][1][:3], rtol=1e-3)


def test_pseudospectral_autograd_gradcheck(ch):
    import torch
    from cahn_hilliard_amd.autograd import euler_evolve
    N = 16
    psi0 = torch.tensor(O.init_density(2, N, 0.0, 0.3, 0, 3), dtype=torch.float64, device="cuda", requires_grad=True)
    eps = torch.tensor(0.9, dtype=torch.float64, device="cuda", requires_grad=True)
    k = torch.tensor(1.1, dtype=torch.float64, device="cuda", requires_grad=True)
    M = torch.tensor(0.8, dtype=torch.float64, device="cuda", requires_grad=True)
    fn = lambda a, b, c, d: euler_evolve(a, b, c, d, 10, 4, dims=2, N=N, dt=0.1, integrator="pseudospectral")
    assert torch.autograd.gradcheck(fn, (psi0, eps, k, M), eps=1e-6, atol=1e-8, rtol=1e-6)
