"""Multi-species Wertheim oracle (src/models/SalehWertheim.cpp, RicciWertheim.cpp,
GenericWertheim.cpp) and the species loops of the integrators.

The reference's multi-species paths have no fixtures, and the reference build
here (oracle/_ref) compiles the single-species integrators only.  Pinned by:
  * cross-model identities: GenericWertheim with Saleh's patch topology (A
    patches on species 0 and on the linker, B patches on species 1 and on the
    linker) reproduces SalehWertheim's closed-form X, f' and f; with Ricci's
    topology (3 A + 1 B on the nanostar, 2 C on the antibody, A-A and B-C
    bonds) RicciWertheim's; with one species of `valence` identical patches,
    SimpleWertheim's.  Agreement to the mass-action tolerance (1e-8);
  * f'_s = d f / d rho_s by central differences for every model.  Saleh's f'
    carries 3 B3 rho^2 while its f has B3 rho^3 / 2 (SalehWertheim.cpp:74, 91):
    the difference is checked to be exactly that;
  * the multi-species integrators against the single-species ones (pinned by
    the reference build): a one-species Generic system equals the
    SimpleWertheim system through euler, euler_mobility and pseudospectral;
    two species without cross terms evolve as two one-species systems, each
    with its own k;
  * total mass of every species unchanged.
"""
import numpy as np
import pytest

import oracle as O

ALL_B2 = lambda S, v: {(a, b): v for a in range(S) for b in range(S)}  # noqa: E731


def saleh_pair(B3=0.0):
    sal = O.saleh_model(1.0, valence=(4, 4, 2), B2=0.7, B3=B3, delta_AA=1e3, delta_BB=2e3)
    gen = O.generic_model(1.0, [[0] * 4, [1] * 4, [0, 1]], {(0, 0): 1e3, (1, 1): 2e3}, ALL_B2(3, 0.7))
    return sal, gen


def ricci_pair():
    ric = O.ricci_model(1.0, B2=0.7, delta_00=1e3, delta_12=5e2)
    gen = O.generic_model(1.0, [[0, 0, 0, 1], [2, 2]], {(0, 0): 1e3, (1, 2): 5e2}, ALL_B2(2, 0.7))
    return ric, gen


@pytest.mark.parametrize("pair", [saleh_pair, ricci_pair])
def test_generic_reproduces_closed_form_models(pair):
    closed, gen = pair()
    rng = np.random.default_rng(1)
    for _ in range(20):
        r = rng.uniform(0.002, 0.08, closed.n_species)
        a, b = O.ms_der(closed, r)[0], O.ms_der(gen, r)[0]
        assert np.abs(a - b).max() < 1e-6 * np.abs(a).max()
        assert abs(O.ms_bulk(closed, r) - O.ms_bulk(gen, r)) < 1e-7 * abs(O.ms_bulk(closed, r))


def test_generic_one_species_is_simple_wertheim():
    gen = O.generic_model(1.0, [[0, 0, 0, 0]], {(0, 0): 1e4}, {(0, 0): 1.3})
    sw = O.Model(O.SIMPLE_WERTHEIM, B2=1.3, delta=1e4, valence=4)
    for r in (1e-4, 0.01, 0.05, 0.3):
        want = O.lib().och_der_bulk_free_energy(__import__("ctypes").byref(sw.c()), r)
        # the reference stops when an iteration moves X by < 1e-8; near X ~ 0.01 the
        # map contracts by ~0.99 per sweep, leaving ~1e-7 relative in f'
        assert abs(O.ms_der(gen, [r])[0][0] - want) < 1e-6 * abs(want)


def fd_grad(model, r, h=1e-6):
    g = np.zeros(len(r))
    for s in range(len(r)):
        e = np.zeros(len(r))
        e[s] = h * r[s]
        g[s] = (O.ms_bulk(model, r + e) - O.ms_bulk(model, r - e)) / (2 * e[s])
    return g


@pytest.mark.parametrize("model", [saleh_pair()[0], ricci_pair()[0], saleh_pair()[1], ricci_pair()[1],
                                   O.generic_model(1.0, [[0, 1, 1], [1, 2], [0, 2, 2, 3]],
                                                   {(0, 1): 800.0, (1, 2): 300.0, (2, 3): 1500.0, (0, 0): 50.0},
                                                   {(0, 0): 0.5, (0, 1): -0.2, (1, 1): 0.8, (1, 2): 0.1,
                                                    (2, 2): 0.3, (0, 2): 0.0})],
                         ids=["saleh", "ricci", "generic_saleh", "generic_ricci", "generic_3species"])
def test_fprime_is_the_density_derivative_of_f(model):
    # Generic: f' uses X iterated until a sweep moves it by < 1e-8 (a few 1e-6
    # in f' where bonding is strong); the closed forms are exact
    tol = 2e-5 if model.model == O.GENERIC else 2e-6
    rng = np.random.default_rng(2)
    for _ in range(10):
        r = rng.uniform(0.005, 0.08, model.n_species)
        assert np.abs(O.ms_der(model, r)[0] - fd_grad(model, r)).max() < tol


def test_saleh_B3_term_as_in_the_reference():
    sal3 = O.saleh_model(1.0, valence=(4, 4, 2), B2=0.7, B3=2.5, delta_AA=1e3, delta_BB=2e3)
    r = np.array([0.02, 0.03, 0.01])
    diff = O.ms_der(sal3, r)[0] - fd_grad(sal3, r)
    assert np.allclose(diff, 1.5 * 2.5 * r.sum() ** 2, rtol=1e-5)


def one_species_params(integrator, mob):
    if integrator == O.PSEUDOSPECTRAL:
        return O.Params(dims=2, N=32, dt=1e-3, dx=1.0, k=1.0, mobility=O.MOB_CONSTANT, M=1.0, S=0.5, use_dealias=1)
    return O.Params(dims=2, N=32, dt=1e-4, dx=1.0, k=1.0, mobility=mob, M=1.0, rho_ref=0.05)


@pytest.mark.parametrize("integrator,mob", [(O.EULER, O.MOB_CONSTANT), (O.EULER_MOBILITY, O.MOB_LINEAR),
                                            (O.EULER_MOBILITY, O.MOB_FREE_ENERGY), (O.PSEUDOSPECTRAL, O.MOB_CONSTANT)])
def test_one_species_generic_equals_simple_wertheim(integrator, mob):
    p = one_species_params(integrator, mob)
    gen = O.generic_model(p.k, [[0] * 4], {(0, 0): 1e4}, {(0, 0): 1.0})
    sw = O.Model(O.SIMPLE_WERTHEIM, B2=1.0, delta=1e4, valence=4)
    rho0 = 0.05 * (1 + 0.05 * (2 * np.random.default_rng(3).random(p.bins) - 1))
    a = O.run_ms(integrator, p, gen, rho0, 50)
    b = O.run(integrator, p, sw, rho0, 50)
    # the Generic f' carries the mass-action stopping error (~1e-7 relative at delta = 1e4)
    assert np.abs(a - b).max() < 2e-6 * np.abs(b).max()


@pytest.mark.parametrize("integrator,mob", [(O.EULER, O.MOB_CONSTANT), (O.EULER_MOBILITY, O.MOB_REGULARISED),
                                            (O.PSEUDOSPECTRAL, O.MOB_CONSTANT)])
def test_decoupled_species_evolve_independently(integrator, mob):
    p = one_species_params(integrator, mob)
    p.rho_min = 0.01
    ks = (1.0, 2.5)
    pair = O.generic_model(ks, [[0, 0, 0], [1, 1]], {(0, 0): 2e3, (1, 1): 5e3}, {(0, 0): 0.9, (1, 1): 1.4})
    rng = np.random.default_rng(4)
    rho0 = 0.04 * (1 + 0.05 * (2 * rng.random(2 * p.bins) - 1))
    both = O.run_ms(integrator, p, pair, rho0, 40)
    for s, (patches, d, b2) in enumerate((([0, 0, 0], 2e3, 0.9), ([0, 0], 5e3, 1.4))):
        one = O.generic_model(ks[s], [patches], {(0, 0): d}, {(0, 0): b2})
        alone = O.run_ms(integrator, p, one, rho0[s * p.bins:(s + 1) * p.bins], 40)
        part = both[s * p.bins:(s + 1) * p.bins]
        # the pair's X sweeps stop on the slower patch: differences at the
        # mass-action tolerance, not at round-off
        assert np.abs(part - alone).max() < 1e-6 * np.abs(alone).max()


@pytest.mark.parametrize("integrator", [O.EULER, O.EULER_MOBILITY, O.PSEUDOSPECTRAL])
def test_species_mass_conserved(integrator):
    sal, _ = saleh_pair()
    p = one_species_params(integrator, O.MOB_REGULARISED)
    p.rho_min = 0.01
    rho0 = 0.03 * (1 + 0.1 * (2 * np.random.default_rng(5).random(3 * p.bins) - 1))
    out = O.run_ms(integrator, p, sal, rho0, 30)
    for s in range(3):
        sl = slice(s * p.bins, (s + 1) * p.bins)
        assert abs(out[sl].sum() - rho0[sl].sum()) < 1e-12 * rho0[sl].sum()
        assert np.abs(out[sl] - rho0[sl]).max() > 1e-8


def test_observables_of_one_species_generic_match_simple_wertheim():
    p = one_species_params(O.EULER, O.MOB_CONSTANT)
    gen = O.generic_model(1.0, [[0] * 4], {(0, 0): 1e4}, {(0, 0): 1.0})
    sw = O.Model(O.SIMPLE_WERTHEIM, B2=1.0, delta=1e4, valence=4)
    rho0 = 0.05 * (1 + 0.05 * (2 * np.random.default_rng(6).random(p.bins) - 1))
    rho, (mass, fe) = O.run_ms(O.EULER, p, gen, rho0, 5, observe=True)
    want_fe = O.average_free_energy(p, sw, rho)
    assert abs(mass - O.average_mass(p, rho)) < 1e-14
    assert abs(fe - want_fe) < 1e-7 * abs(want_fe)
