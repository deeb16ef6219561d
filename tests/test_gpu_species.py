"""GPU multi-species Wertheim systems (ch_sys_create_species) against the oracle.

Tolerances (stated per result):
  FP64, Saleh and Ricci (closed-form X): <= 1e-10 relative L-inf after 50 steps;
  FP64, Generic: <= 1e-8 (the mass-action sweep stops when no X moves by 1e-8;
      inputs that differ in the last bit can stop one sweep apart, which moves
      f' by up to ~1e-8 relative);
  every species' mass unchanged to 1e-12; observables: mass 1e-13, free
  energy 1e-10 (closed forms) / 1e-8 (Generic) relative;
  FP32 fields against the FP64 oracle: <= 2e-4 after 50 steps.
"""
import numpy as np
import pytest

import oracle as O
from helpers import rel_linf

pytestmark = pytest.mark.gpu


def saleh(cls):
    return O.saleh_model((1.0, 1.5, 2.0), valence=(4, 4, 2), B2=0.7, B3=0.3, delta_AA=1e3, delta_BB=2e3,
                         model_cls=cls)


def ricci(cls):
    return O.ricci_model((1.0, 2.0), B2=0.7, delta_00=1e3, delta_12=5e2, model_cls=cls)


def generic3(cls):
    return O.generic_model((1.0, 1.3, 0.8), [[0, 1, 1], [1, 2], [0, 2, 2, 3]],
                           {(0, 1): 800.0, (1, 2): 300.0, (2, 3): 1500.0, (0, 0): 50.0},
                           {(0, 0): 0.5, (0, 1): -0.2, (1, 1): 0.8, (1, 2): 0.1, (2, 2): 0.3, (0, 2): 0.0},
                           model_cls=cls)


NAMES = {O.SALEH: "saleh", O.RICCI: "ricci", O.GENERIC: "generic_wertheim"}
# (name, model builder, integrator, dims, N, dt, mobility kw, S, dealias)
CASES = [
    ("saleh2d_N64_euler", saleh, "euler", 2, 64, 1e-4, dict(mobility="constant"), 0.0, False),
    ("saleh2d_N64_em_regularised", saleh, "euler_mobility", 2, 64, 1e-4, dict(mobility="regularised", rho_min=0.01),
     0.0, False),
    ("saleh1d_N256_ps", saleh, "pseudospectral", 1, 256, 1e-3, dict(mobility="constant"), 0.5, True),
    ("ricci2d_N64_euler", ricci, "euler", 2, 64, 1e-4, dict(mobility="constant"), 0.0, False),
    ("ricci3d_N16_em_linear", ricci, "euler_mobility", 3, 16, 1e-4, dict(mobility="linear", rho_ref=0.03), 0.0,
     False),
    ("ricci2d_N64_ps", ricci, "pseudospectral", 2, 64, 1e-3, dict(mobility="constant"), 0.3, False),
    ("generic2d_N64_em_free_energy", generic3, "euler_mobility", 2, 64, 1e-4, dict(mobility="free_energy"), 0.0,
     False),
    ("generic2d_N32_euler", generic3, "euler", 2, 32, 1e-4, dict(mobility="constant"), 0.0, False),
    ("generic3d_N8_ps", generic3, "pseudospectral", 3, 8, 1e-3, dict(mobility="constant"), 0.5, True),
]
INTEGRATOR = {"euler": O.EULER, "euler_mobility": O.EULER_MOBILITY, "pseudospectral": O.PSEUDOSPECTRAL}
MOB = {"constant": O.MOB_CONSTANT, "regularised": O.MOB_REGULARISED, "linear": O.MOB_LINEAR,
       "free_energy": O.MOB_FREE_ENERGY}


@pytest.fixture(scope="module")
def ctx(ch):
    c = ch.Context(0)
    yield c
    c.close()


def setup(ch, case):
    name, build, integ, dims, N, dt, mob, S, deal = case
    om = build(O.SpeciesModel)
    gm = build(ch.ChSpeciesModel)
    p = O.Params(dims=dims, N=N, dt=dt, dx=1.0, k=1.0, mobility=MOB[mob["mobility"]], M=1.0,
                 rho_min=mob.get("rho_min", 0.0), rho_ref=mob.get("rho_ref", 1.0), S=S, use_dealias=int(deal))
    desc = ch.make_desc(dims=dims, N=N, integrator=integ, free_energy=NAMES[om.model], dt=dt, dx=1.0, M=1.0, S=S,
                        use_dealias=deal, **mob)
    rng = np.random.default_rng(11)
    n = om.n_species * N ** dims
    rho0 = 0.03 * (1 + 0.1 * (2 * rng.random(n) - 1))
    return om, gm, p, desc, rho0


def tol_of(om):
    return 1e-8 if om.model == O.GENERIC else 1e-10


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_species_matches_oracle_fp64(ch, ctx, case):
    print(f"case {case[0]}", flush=True)
    om, gm, p, desc, rho0 = setup(ch, case)
    integ = INTEGRATOR[case[2]]
    want, (mass_w, fe_w) = O.run_ms(integ, p, om, rho0, 50, observe=True)
    s = ch.System(ctx, desc, species_model=gm)
    assert s.species == om.n_species
    s.upload(rho0)
    s.evolve(30)
    s.evolve(20)
    got = s.download()
    mass, fe = s.observables()
    s.close()
    err = rel_linf(got, want)
    print(f"{case[0]}: fp64 rel L-inf vs oracle = {err:.3e}; fe {fe:.12e} vs {fe_w:.12e}")
    assert err < tol_of(om)
    bins = p.bins
    for sp in range(om.n_species):
        sl = slice(sp * bins, (sp + 1) * bins)
        assert abs(got[sl].sum() - rho0[sl].sum()) < 1e-12 * rho0[sl].sum()
    assert abs(mass - mass_w) < 1e-13 * abs(mass_w)
    assert abs(fe - fe_w) < tol_of(om) * abs(fe_w)


def test_species_graph_replay_equals_direct_steps(ch, ctx):
    om, gm, p, desc, rho0 = setup(ch, CASES[1])
    outs = []
    for splits in ((130,), (1, 63, 2, 64)):
        s = ch.System(ctx, desc, species_model=gm)
        s.upload(rho0)
        for n in splits:
            s.evolve(n)
        outs.append(s.download())
        s.close()
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("case", [CASES[0], CASES[6]], ids=[CASES[0][0], CASES[6][0]])
def test_species_fp32_against_fp64_oracle(ch, ctx, case):
    om, gm, p, desc, rho0 = setup(ch, case)
    desc.precision = ch.FP32
    want = O.run_ms(INTEGRATOR[case[2]], p, om, rho0, 50)
    s = ch.System(ctx, desc, species_model=gm)
    s.upload(rho0)
    s.evolve(50)
    got = s.download()
    s.close()
    err = rel_linf(got, want)
    print(f"{case[0]}: fp32 rel L-inf vs fp64 oracle = {err:.3e}")
    assert err < 2e-4


def test_species_rejects_single_species_features(ch, ctx):
    om, gm, p, desc, rho0 = setup(ch, CASES[0])
    s = ch.System(ctx, desc, species_model=gm)
    s.upload(rho0)
    with pytest.raises(ch.CHError):
        s.record(4, 0)
    with pytest.raises(ch.CHError):
        s.structure_factor()
    s.close()
    bad = ch.make_desc(dims=2, N=32, integrator="euler_mobility", free_energy="saleh", mobility="free_energy")
    with pytest.raises(ch.CHError):
        ch.System(ctx, bad, species_model=saleh(ch.ChSpeciesModel))
