"""ctypes bindings for the oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker.  Imported by tests/, by
__graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
product package.  Each wrapper names the reference code it restates (the C
source cites file:line per function).
"""
import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

LANDAU, SIMPLE_WERTHEIM = 0, 1
MOB_CONSTANT, MOB_FREE_ENERGY, MOB_REGULARISED, MOB_LINEAR = 0, 1, 2, 3
EULER, EULER_MOBILITY, PSEUDOSPECTRAL = 0, 1, 2


class _Model(ctypes.Structure):
    _fields_ = [("model", ctypes.c_int), ("epsilon", ctypes.c_double), ("B2", ctypes.c_double),
                ("delta", ctypes.c_double), ("regularisation_delta", ctypes.c_double),
                ("valence", ctypes.c_int)]


class _Params(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int), ("N", ctypes.c_int), ("dt", ctypes.c_double),
                ("dx", ctypes.c_double), ("k", ctypes.c_double), ("mobility", ctypes.c_int),
                ("M", ctypes.c_double), ("rho_min", ctypes.c_double), ("rho_ref", ctypes.c_double),
                ("S", ctypes.c_double), ("use_dealias", ctypes.c_int)]


@dataclass
class Model:
    model: int = LANDAU
    epsilon: float = 0.9
    B2: float = 1.0
    delta: float = 1e4
    regularisation_delta: float = 0.0
    valence: int = 4

    def c(self):
        return _Model(self.model, self.epsilon, self.B2, self.delta, self.regularisation_delta, self.valence)


@dataclass
class Params:
    dims: int = 2
    N: int = 64
    dt: float = 0.01
    dx: float = 1.0
    k: float = 1.0
    mobility: int = MOB_CONSTANT
    M: float = 1.0
    rho_min: float = 0.0
    rho_ref: float = 1.0
    S: float = 0.0
    use_dealias: int = 0

    def c(self):
        return _Params(self.dims, self.N, self.dt, self.dx, self.k, self.mobility, self.M,
                       self.rho_min, self.rho_ref, self.S, self.use_dealias)

    @property
    def bins(self):
        return self.N ** self.dims


_lib = None


def lib(fast=False):
    """Load (fast=False) liboracle.so or (fast=True) liboracle_fast.so."""
    global _lib
    name = "liboracle_fast.so" if fast else "liboracle.so"
    path = os.path.join(_HERE, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: run `make -C oracle` (or __graft_entry__.build())")
    L = ctypes.CDLL(path)
    D, P, I64, I = ctypes.c_double, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
    MP, PP = ctypes.POINTER(_Model), ctypes.POINTER(_Params)
    sigs = {
        "och_drand48": (D, [P]), "och_srand48": (None, [P, I64]),
        "och_init_density": (None, [I, I, D, D, I, I64, P]),
        "och_bulk_free_energy": (D, [MP, D]), "och_der_bulk_free_energy": (D, [MP, D]),
        "och_second_der_bulk_free_energy": (D, [MP, D]), "och_wertheim_X": (D, [MP, D]),
        "och_pressure": (D, [MP, D]),
        "och_der_bulk_free_energy_field": (None, [MP, P, P, I64]),
        "och_update_mobility": (None, [PP, MP, P, P, I64]),
        "och_euler_step": (None, [PP, MP, P]), "och_euler_mobility_step": (None, [PP, MP, P]),
        "och_ps_create": (P, [PP, MP, P]), "och_ps_step": (None, [P, P]), "och_ps_destroy": (None, [P]),
        "och_ps_hat_size": (I64, [P]), "och_ps_get_hat": (None, [P, P]),
        "och_run": (I, [I, PP, MP, P, I64]),
        "och_average_mass": (D, [PP, P]), "och_average_free_energy": (D, [PP, MP, P]),
        "och_fft_c2c": (None, [I, I, P, I]), "och_dft_brute": (None, [I, I, P, P]),
        "och_sq_nshells": (I, [I]), "och_radial_sq": (None, [I, D, P, P]),
        "och_sq_loss": (D, [I, D, P, P, P]),
        "och_euler_adjoint_step": (None, [PP, MP, P, P, P]),
        "och_euler_adjoint": (None, [PP, MP, P, I64, P, P, P]),
        "och_der_bulk_free_energy_ddelta": (D, [MP, D]), "och_bulk_free_energy_ddelta": (D, [MP, D]),
        "och_mobility_drho": (D, [PP, MP, D]), "och_mobility_ddelta": (D, [PP, MP, D]),
        "och_adjoint_step": (None, [I, PP, MP, P, P, P]),
        "och_adjoint": (None, [I, PP, MP, P, I64, P, P, P]),
        "och_free_energy_grad": (D, [PP, MP, P, P, P]),
        "och_ps_adjoint_step": (None, [PP, MP, P, P, P]),
        "och_ps_adjoint": (None, [PP, MP, P, I64, P, P, P]),
        "och_psm_create": (ctypes.c_void_p, [PP, MP, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, D]),
        "och_psm_step": (ctypes.c_int, [ctypes.c_void_p, P]),
        "och_psm_last_M0": (D, [ctypes.c_void_p]),
        "och_psm_destroy": (None, [ctypes.c_void_p]),
        "och_psm_div_mob_grad": (None, [PP, P, P, P]),
        "och_ms_update_X": (None, [P, P, P]),
        "och_ms_der": (None, [P, P, P, P]),
        "och_ms_bulk": (D, [P, D, P]),
        "och_ms_create": (ctypes.c_void_p, [I, PP, P, P]),
        "och_ms_step": (None, [ctypes.c_void_p, P]),
        "och_ms_observables": (None, [ctypes.c_void_p, D, P, P, P]),
        "och_ms_X": (ctypes.c_void_p, [ctypes.c_void_p]),
        "och_ms_destroy": (None, [ctypes.c_void_p]),
        "och_philox4x32_10": (None, [P, P, P]),
        "och_cell_normals": (None, [ctypes.c_uint64, I64, I64, P]),
        "och_euler_mobility_noise_step": (None, [PP, MP, D, ctypes.c_uint64, I64, P]),
    }
    for n, (r, a) in sigs.items():
        f = getattr(L, n)
        f.restype, f.argtypes = r, a
    if not fast:
        _lib = L
    return L


def _L():
    return _lib if _lib is not None else lib()


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Rng(ctypes.Structure):
    _fields_ = [("x", ctypes.c_uint64)]


def drand48_stream(seed, n):
    r = Rng()
    _L().och_srand48(ctypes.byref(r), seed)
    return np.array([_L().och_drand48(ctypes.byref(r)) for _ in range(n)])


def init_density(dims, N, density, A=0.01, N_peaks=0, seed=123456):
    """src/CahnHilliard.cpp:133-151."""
    rho = np.empty(N ** dims)
    _L().och_init_density(dims, N, density, A, N_peaks, seed, _ptr(rho))
    return rho


def der_bulk(model, rho):
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    out = np.empty_like(rho)
    _L().och_der_bulk_free_energy_field(ctypes.byref(model.c()), _ptr(rho), _ptr(out), rho.size)
    return out


def bulk(model, r):
    return _L().och_bulk_free_energy(ctypes.byref(model.c()), float(r))


def second_der_bulk(model, r):
    return _L().och_second_der_bulk_free_energy(ctypes.byref(model.c()), float(r))


def wertheim_X(model, r):
    return _L().och_wertheim_X(ctypes.byref(model.c()), float(r))


def mobility(params, model, rho):
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    out = np.empty_like(rho)
    _L().och_update_mobility(ctypes.byref(params.c()), ctypes.byref(model.c()), _ptr(rho), _ptr(out), rho.size)
    return out


def run(integrator, params, model, rho, steps, fast=False):
    """`steps` calls of mobility update + integrator evolve; returns a new array."""
    rho = np.array(rho, dtype=np.float64, copy=True)
    L = lib(fast=True) if fast else _L()
    rc = L.och_run(integrator, ctypes.byref(params.c()), ctypes.byref(model.c()), _ptr(rho), steps)
    if rc != 0:
        raise ValueError("bad integrator")
    return rho


class Pseudospectral:
    """src/integrators/PseudospectralCPU.cpp:13-136 (state carries rho_hat)."""

    def __init__(self, params, model, rho):
        self._p, self._m = params.c(), model.c()
        self.rho = np.array(rho, dtype=np.float64, copy=True)
        self._s = _L().och_ps_create(ctypes.byref(self._p), ctypes.byref(self._m), _ptr(self.rho))

    def step(self, n=1):
        for _ in range(n):
            _L().och_ps_step(self._s, _ptr(self.rho))
        return self.rho

    def rho_hat(self):
        h = np.empty(2 * _L().och_ps_hat_size(self._s))
        _L().och_ps_get_hat(self._s, _ptr(h))
        return h[0::2] + 1j * h[1::2]

    def __del__(self):
        if getattr(self, "_s", None):
            _L().och_ps_destroy(self._s)
            self._s = None


def average_mass(params, rho):
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    return _L().och_average_mass(ctypes.byref(params.c()), _ptr(rho))


def average_free_energy(params, model, rho):
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    return _L().och_average_free_energy(ctypes.byref(params.c()), ctypes.byref(model.c()), _ptr(rho))


def fft_c2c(x, inverse=False):
    a = np.empty(2 * x.size)
    a[0::2], a[1::2] = x.real, x.imag
    _L().och_fft_c2c(x.size, int(inverse), _ptr(a), 1)
    return a[0::2] + 1j * a[1::2]


def dft_brute(x, inverse=False):
    a = np.empty(2 * x.size)
    a[0::2], a[1::2] = x.real, x.imag
    out = np.empty_like(a)
    _L().och_dft_brute(x.size, int(inverse), _ptr(a), _ptr(out))
    return out[0::2] + 1j * out[1::2]


def radial_sq(field2d, dx=1.0):
    L = field2d.shape[0]
    f = np.ascontiguousarray(field2d, dtype=np.float64)
    out = np.empty(_L().och_sq_nshells(L))
    _L().och_radial_sq(L, dx, _ptr(f), _ptr(out))
    return out


def sq_loss(field2d, target, dx=1.0, grad=True):
    L = field2d.shape[0]
    f = np.ascontiguousarray(field2d, dtype=np.float64)
    t = np.ascontiguousarray(target, dtype=np.float64)
    g = np.empty_like(f) if grad else None
    loss = _L().och_sq_loss(L, dx, _ptr(f), _ptr(t), _ptr(g) if grad else None)
    return (loss, g) if grad else loss


def euler_adjoint(params, model, psi0, steps, dL_dpsiT):
    psi0 = np.ascontiguousarray(psi0, dtype=np.float64)
    w = np.ascontiguousarray(dL_dpsiT, dtype=np.float64)
    out = np.empty_like(psi0)
    grad = np.zeros(3)
    _L().och_euler_adjoint(ctypes.byref(params.c()), ctypes.byref(model.c()), _ptr(psi0), steps,
                           _ptr(w), _ptr(out), _ptr(grad))
    return out, grad


def euler_adjoint_step(params, model, psi, lam):
    psi = np.ascontiguousarray(psi, dtype=np.float64)
    lam = np.array(lam, dtype=np.float64, copy=True)
    grad = np.zeros(3)
    _L().och_euler_adjoint_step(ctypes.byref(params.c()), ctypes.byref(model.c()), _ptr(psi), _ptr(lam), _ptr(grad))
    return lam, grad


# grad[] order of adjoint() / free_energy_grad(): oracle/ch_oracle.h OCH_G_*
G_EPSILON, G_K, G_M, G_DELTA, G_B2, NGRADS = 0, 1, 2, 3, 4, 5


def der_bulk_ddelta(model, r):
    return _L().och_der_bulk_free_energy_ddelta(ctypes.byref(model.c()), float(r))


def bulk_ddelta(model, r):
    return _L().och_bulk_free_energy_ddelta(ctypes.byref(model.c()), float(r))


def mobility_drho(params, model, r):
    return _L().och_mobility_drho(ctypes.byref(params.c()), ctypes.byref(model.c()), float(r))


def mobility_ddelta(params, model, r):
    return _L().och_mobility_ddelta(ctypes.byref(params.c()), ctypes.byref(model.c()), float(r))


def adjoint(integrator, params, model, psi0, steps, dL_dpsiT):
    """Reverse sweep of `steps` EULER / EULER_MOBILITY / PSEUDOSPECTRAL steps for
    the loss <dL_dpsiT, psi_T>: returns (dL/dpsi0, grad[NGRADS])."""
    psi0 = np.ascontiguousarray(psi0, dtype=np.float64)
    w = np.ascontiguousarray(dL_dpsiT, dtype=np.float64)
    out = np.empty_like(psi0)
    grad = np.zeros(NGRADS)
    args = (ctypes.byref(params.c()), ctypes.byref(model.c()), _ptr(psi0), steps, _ptr(w), _ptr(out), _ptr(grad))
    if integrator == PSEUDOSPECTRAL:
        _L().och_ps_adjoint(*args)
    else:
        _L().och_adjoint(integrator, *args)
    return out, grad


def adjoint_step(integrator, params, model, psi, lam):
    psi = np.ascontiguousarray(psi, dtype=np.float64)
    lam = np.array(lam, dtype=np.float64, copy=True)
    grad = np.zeros(NGRADS)
    args = (ctypes.byref(params.c()), ctypes.byref(model.c()), _ptr(psi), _ptr(lam), _ptr(grad))
    if integrator == PSEUDOSPECTRAL:
        _L().och_ps_adjoint_step(*args)
    else:
        _L().och_adjoint_step(integrator, *args)
    return lam, grad


def free_energy_grad(params, model, rho):
    """(F, dF/drho, explicit dF/dparams[NGRADS]) of average_free_energy."""
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    g = np.empty_like(rho)
    dp = np.zeros(NGRADS)
    F = _L().och_free_energy_grad(ctypes.byref(params.c()), ctypes.byref(model.c()), _ptr(rho), _ptr(g), _ptr(dp))
    return F, g, dp


def philox4x32_10(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    _L().och_philox4x32_10(_ptr(c), _ptr(k), _ptr(out))
    return out


def cell_normals(seed, cell, step):
    out = np.zeros(4)
    _L().och_cell_normals(seed, cell, step, _ptr(out))
    return out


def run_noise(params, model, rho, steps, noise_rescale=1.0, seed=123456, step0=0):
    """`steps` euler_mobility steps with conserved noise (step indices step0, ...)."""
    rho = np.array(rho, dtype=np.float64, copy=True)
    for t in range(steps):
        _L().och_euler_mobility_noise_step(ctypes.byref(params.c()), ctypes.byref(model.c()), noise_rescale, seed,
                                           step0 + t, _ptr(rho))
    return rho


def run_psm(params, model, rho, steps, use_gmres=False, restart=30, max_iter=200, tol=1e-10, iters=None):
    """`steps` pseudospectral_mobility steps; GMRES iteration counts appended to `iters`."""
    rho = np.array(rho, dtype=np.float64, copy=True)
    L = _L()
    h = L.och_psm_create(ctypes.byref(params.c()), ctypes.byref(model.c()), _ptr(rho), int(use_gmres), restart,
                         max_iter, tol)
    try:
        for _ in range(steps):
            it = L.och_psm_step(h, _ptr(rho))
            if iters is not None:
                iters.append(it)
    finally:
        L.och_psm_destroy(h)
    return rho


def psm_div_mob_grad(params, mob, g):
    mob = np.ascontiguousarray(mob, dtype=np.float64)
    g = np.ascontiguousarray(g, dtype=np.float64)
    out = np.zeros(params.bins)
    _L().och_psm_div_mob_grad(ctypes.byref(params.c()), _ptr(mob), _ptr(g), _ptr(out))
    return out


# ---- multi-species models (layout of ch_species_model)
SALEH, RICCI, GENERIC = 2, 3, 4
MAX_SPECIES, MAX_PATCHES = 8, 16


class SpeciesModel(ctypes.Structure):
    _fields_ = [("model", ctypes.c_int32), ("n_species", ctypes.c_int32), ("k", ctypes.c_double * MAX_SPECIES),
                ("saleh_valence", ctypes.c_int32 * 3), ("reserved0", ctypes.c_int32),
                ("saleh_B2", ctypes.c_double), ("saleh_B3", ctypes.c_double), ("saleh_delta_AA", ctypes.c_double),
                ("saleh_delta_BB", ctypes.c_double), ("ricci_B2", ctypes.c_double), ("ricci_delta_00", ctypes.c_double),
                ("ricci_delta_12", ctypes.c_double), ("n_patches", ctypes.c_int32), ("reserved1", ctypes.c_int32),
                ("patch_count", (ctypes.c_int32 * MAX_PATCHES) * MAX_SPECIES),
                ("delta", (ctypes.c_double * MAX_PATCHES) * MAX_PATCHES),
                ("B2", (ctypes.c_double * MAX_SPECIES) * MAX_SPECIES)]


def saleh_model(k, valence=(4, 4, 2), B2=1.0, B3=0.0, delta_AA=1e3, delta_BB=2e3, model_cls=None):
    m = (model_cls or SpeciesModel)()
    m.model, m.n_species = SALEH, 3
    for i, v in enumerate(np.broadcast_to(k, 3)):
        m.k[i] = v
    for i, v in enumerate(valence):
        m.saleh_valence[i] = v
    m.saleh_B2, m.saleh_B3, m.saleh_delta_AA, m.saleh_delta_BB = B2, B3, delta_AA, delta_BB
    return m


def ricci_model(k, B2=1.0, delta_00=1e3, delta_12=5e2, model_cls=None):
    m = (model_cls or SpeciesModel)()
    m.model, m.n_species = RICCI, 2
    for i, v in enumerate(np.broadcast_to(k, 2)):
        m.k[i] = v
    m.ricci_B2, m.ricci_delta_00, m.ricci_delta_12 = B2, delta_00, delta_12
    return m


def generic_model(k, species_patches, deltas, B2s, model_cls=None):
    """species_patches: list of patch-type lists; deltas: {(p, q): delta}; B2s: {(s, t): B2}."""
    m = (model_cls or SpeciesModel)()
    m.model, m.n_species = GENERIC, len(species_patches)
    for i, v in enumerate(np.broadcast_to(k, len(species_patches))):
        m.k[i] = v
    m.n_patches = max(max(p) for p in species_patches) + 1
    for s_, pl in enumerate(species_patches):
        for q in pl:
            m.patch_count[s_][q] += 1
    for (p_, q), d in deltas.items():
        m.delta[p_][q] = m.delta[q][p_] = d
    for (a, b), v in B2s.items():
        m.B2[a][b] = m.B2[b][a] = v
    return m


def ms_der(model, rhos, X=None):
    rhos = np.ascontiguousarray(rhos, dtype=np.float64)
    X = np.ones(MAX_PATCHES) if X is None else np.ascontiguousarray(X, dtype=np.float64)
    out = np.zeros(model.n_species)
    _L().och_ms_der(ctypes.byref(model), _ptr(rhos), _ptr(X), _ptr(out))
    return out, X


def ms_bulk(model, rhos, dcf=1.0):
    rhos = np.ascontiguousarray(rhos, dtype=np.float64)
    return _L().och_ms_bulk(ctypes.byref(model), dcf, _ptr(rhos))


def run_ms(integrator, params, model, rho, steps, dcf=1.0, observe=False):
    """`steps` steps of a multi-species system; rho is species-major (S * bins)."""
    rho = np.array(rho, dtype=np.float64, copy=True)
    L = _L()
    h = L.och_ms_create(integrator, ctypes.byref(params.c()), ctypes.byref(model), _ptr(rho))
    try:
        for _ in range(steps):
            L.och_ms_step(h, _ptr(rho))
        obs = None
        if observe:
            mass, fe = ctypes.c_double(), ctypes.c_double()
            L.och_ms_observables(h, dcf, _ptr(rho), ctypes.byref(mass), ctypes.byref(fe))
            obs = (mass.value, fe.value)
    finally:
        L.och_ms_destroy(h)
    return (rho, obs) if observe else rho
