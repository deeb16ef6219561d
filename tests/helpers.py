"""Shared helpers of the test-suite: golden cases -> oracle params / ch_desc."""
import json
import os

import numpy as np

import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = json.load(open(os.path.join(GOLD, "cases.json")))
LARGE = {c["name"]: c for c in json.load(open(os.path.join(GOLD, "large_cases.json")))}
MOB = {"constant": O.MOB_CONSTANT, "free_energy": O.MOB_FREE_ENERGY, "regularised": O.MOB_REGULARISED,
       "linear": O.MOB_LINEAR}


def golden(name):
    return np.load(os.path.join(GOLD, name + ".npz"))


def oracle_setup(c):
    m = c["model"]
    if m["free_energy"] == "landau":
        mod = O.Model(O.LANDAU, epsilon=m["epsilon"])
    else:
        mod = O.Model(O.SIMPLE_WERTHEIM, B2=m["B2"], delta=m["delta"],
                      regularisation_delta=m.get("regularisation_delta", 0.0), valence=m["valence"])
    mb = c["mobility"]
    p = O.Params(dims=c["dims"], N=c["N"], dt=c["dt"], dx=c.get("dx", 1.0), k=c["k"], mobility=MOB[mb["type"]],
                 M=mb["M"], rho_min=mb.get("rho_min", 0.0), rho_ref=mb.get("rho_ref", 1.0), S=c.get("S", 0.0),
                 use_dealias=int(c.get("use_dealias", 0)))
    integ = {"euler": O.EULER, "euler_mobility": O.EULER_MOBILITY, "pseudospectral": O.PSEUDOSPECTRAL}[c["integrator"]]
    return p, mod, integ


def desc_for(ch, c, precision="fp64"):
    m = c["model"]
    mb = c["mobility"]
    kw = dict(dims=c["dims"], N=c["N"], integrator=c["integrator"], free_energy=m["free_energy"],
              precision=precision, dt=c["dt"], dx=c.get("dx", 1.0), k=c["k"], mobility=mb["type"], M=mb["M"],
              rho_min=mb.get("rho_min", 0.0), rho_ref=mb.get("rho_ref", 1.0), S=c.get("S", 0.0),
              use_dealias=c.get("use_dealias", 0))
    if m["free_energy"] == "landau":
        kw["epsilon"] = m["epsilon"]
    else:
        kw.update(B2=m["B2"], delta=m["delta"], valence=m["valence"],
                  regularisation_delta=m.get("regularisation_delta", 0.0))
    return ch.make_desc(**kw)


def rel_linf(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
