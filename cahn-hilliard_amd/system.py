"""Python handles over the C ABI: Context (device + stream) and System.

`System` plays the role of the reference's CahnHilliard<dims> object driving an
Integrator (src/CahnHilliard.h:24-62): evolve() == CahnHilliard::evolve called
n times, sync()/download() == Integrator::sync, average_mass() /
average_free_energy() == the FP64 observables (src/CahnHilliard.cpp:285-319).
`desc_from_config` reads the reference's TOML keys (README.md:24-53 and the
_config_* calls in Integrator.cpp:11-28, Landau.cpp:20-33,
SimpleWertheim.cpp:16-33, IMobility.cpp:18-57, PseudospectralCPU.cpp:23-25).
"""
import ctypes
import math
import time
import weakref

import numpy as np

from . import _lib as L


def _check(ctx_handle, status):
    if status != L.CH_OK:
        msg = L.lib().ch_ctx_last_error(ctx_handle) if ctx_handle else b""
        raise L.CHError(status, (msg or b"").decode() or L.lib().ch_status_string(status).decode())


class Context:
    def __init__(self, device=0):
        h = ctypes.c_void_p()
        st = L.lib().ch_ctx_create(device, ctypes.byref(h))
        if st != L.CH_OK:
            raise L.CHError(st, f"ch_ctx_create(device={device}) failed (is a GPU visible?)")
        self.handle = h
        # systems and batches created on this context: closed before it
        self._children = weakref.WeakSet()

    def synchronize(self):
        _check(self.handle, L.lib().ch_ctx_synchronize(self.handle))

    @property
    def stream(self):
        return L.lib().ch_ctx_stream(self.handle)

    def close(self):
        if getattr(self, "handle", None):
            for child in list(getattr(self, "_children", ())):
                child.close()
            L.lib().ch_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_desc(dims, N, integrator="euler", free_energy="landau", precision="fp64", dt=0.01, dx=1.0, k=1.0,
              epsilon=0.9, B2=1.0, delta=1e4, valence=4, regularisation_delta=0.0, mobility="constant", M=1.0,
              rho_min=0.0, rho_ref=1.0, S=0.0, use_dealias=False, with_noise=False, noise_rescale_factor=1.0,
              seed=0, use_gmres=False, gmres_restart=30, gmres_max_iter=200, gmres_tol=1e-10,
              density_conversion=1.0):
    d = L.ChDesc()
    d.dims, d.N = dims, N
    d.precision = {"fp64": L.FP64, "fp32": L.FP32}[precision] if isinstance(precision, str) else precision
    d.integrator = L.INTEGRATORS[integrator] if isinstance(integrator, str) else integrator
    d.model = ({**L.MODELS, **L.MODELS_SPECIES}[free_energy] if isinstance(free_energy, str) else free_energy)
    d.mobility = L.MOBILITIES[mobility] if isinstance(mobility, str) else mobility
    d.valence, d.use_dealias = valence, int(bool(use_dealias))
    d.epsilon, d.B2, d.delta, d.regularisation_delta = epsilon, B2, delta, regularisation_delta
    d.M, d.rho_min, d.rho_ref = M, rho_min, rho_ref
    d.dt, d.dx, d.k, d.S = dt, dx, k, S
    d.with_noise, d.noise_rescale_factor, d.seed = int(bool(with_noise)), noise_rescale_factor, seed
    d.use_gmres, d.gmres_restart, d.gmres_max_iter, d.gmres_tol = int(bool(use_gmres)), gmres_restart, gmres_max_iter, \
        gmres_tol
    d.density_conversion = density_conversion
    return d


def _getter(cfg):
    def get(path, default=None, required=False):
        node = cfg
        for part in path.split("."):
            if not isinstance(node, dict) or part not in node:
                if required:
                    raise KeyError(f"mandatory option '{path}' is missing")
                return default
            node = node[part]
        return node
    return get


def delta_value(delta):
    """A bonding volume given as a number or as the {T, deltaH, deltaS, salt,
    sticky_size} / {value} table of src/utils/Delta.cpp:15-42."""
    if isinstance(delta, dict):
        value = delta.get("value", 0.0)
        if value == 0.0:
            T, dH, dS = delta["T"], delta["deltaH"], delta["deltaS"]
            salt, L_DNA = delta.get("salt", 1.0), delta.get("sticky_size", 6)
            dS_salt = 0.368 * (L_DNA - 1.0) * math.log(salt)
            dG = dH - T * (dS + dS_salt)
            value = 1.6606 * math.exp(-dG / (1.9872036 * T))
        return value
    return delta


def _fill_k(m, k):
    for i, v in enumerate(np.broadcast_to(np.asarray(k, dtype=float), (m.n_species,))):
        m.k[i] = v


def saleh_model(k, valence=(4, 4, 2), B2=1.0, B3=0.0, delta_AA=1e3, delta_BB=1e3):
    """ChSpeciesModel of SalehWertheim (internal units; k per species or one k)."""
    m = L.ChSpeciesModel()
    m.model, m.n_species = L.SALEH_WERTHEIM, 3
    _fill_k(m, k)
    for i, v in enumerate(valence):
        m.saleh_valence[i] = int(v)
    m.saleh_B2, m.saleh_B3, m.saleh_delta_AA, m.saleh_delta_BB = B2, B3, delta_AA, delta_BB
    return m


def ricci_model(k, B2=1.0, delta_00=1e3, delta_12=1e3):
    """ChSpeciesModel of RicciWertheim (internal units)."""
    m = L.ChSpeciesModel()
    m.model, m.n_species = L.RICCI_WERTHEIM, 2
    _fill_k(m, k)
    m.ricci_B2, m.ricci_delta_00, m.ricci_delta_12 = B2, delta_00, delta_12
    return m


def generic_model(k, species_patches, deltas, B2s):
    """ChSpeciesModel of GenericWertheim: species_patches lists the patch types
    of each species, deltas maps (p, q) and B2s maps (s, t) to values (both symmetric)."""
    m = L.ChSpeciesModel()
    m.model, m.n_species = L.GENERIC_WERTHEIM, len(species_patches)
    _fill_k(m, k)
    m.n_patches = max(max(pl) for pl in species_patches) + 1
    for s_, pl in enumerate(species_patches):
        for q in pl:
            m.patch_count[s_][q] += 1
    for (p_, q), d in deltas.items():
        m.delta[p_][q] = m.delta[q][p_] = d
    for (a, b_), v in B2s.items():
        m.B2[a][b_] = m.B2[b_][a] = v
    return m


def species_model_from_config(cfg):
    """ChSpeciesModel of a multi-species free energy (saleh, ricci,
    generic_wertheim), or None for the single-species models.  Keys and unit
    scaling as SalehWertheim.cpp:18-37, RicciWertheim.cpp:19-36,
    GenericWertheim.cpp:22-121 and Integrator.cpp:24-27 (k per species ~ m^5)."""
    get = _getter(cfg)
    u2i = 1.0 / get("distance_scaling_factor", 1.0)
    model = get("free_energy", required=True)
    if model not in L.MODELS_SPECIES:
        return None
    m = L.ChSpeciesModel()
    m.model = L.MODELS_SPECIES[model]
    if model == "saleh":
        m.n_species = 3
        val = get("saleh.valence", required=True)
        val = [val] * 3 if not isinstance(val, list) else val
        for i in range(3):
            m.saleh_valence[i] = int(val[i])
        m.saleh_B2 = get("saleh.B2", required=True) * u2i ** 3
        m.saleh_B3 = get("saleh.B3", 0.0) * u2i ** 6
        m.saleh_delta_AA = delta_value(get("saleh.delta_AA", required=True)) * u2i ** 3
        m.saleh_delta_BB = delta_value(get("saleh.delta_BB", required=True)) * u2i ** 3
    elif model == "ricci":
        m.n_species = 2
        m.ricci_B2 = get("ricci.B2", required=True) * u2i ** 3
        m.ricci_delta_00 = delta_value(get("ricci.delta_00", required=True)) * u2i ** 3
        m.ricci_delta_12 = delta_value(get("ricci.delta_12", required=True)) * u2i ** 3
    else:
        species = get("generic_wertheim.species", required=True)
        m.n_species = len(species)
        if m.n_species > L.MAX_SPECIES:
            raise ValueError(f"at most {L.MAX_SPECIES} species are supported")
        top = 0
        for s_, sp in enumerate(species):
            for q in sp["patches"]:
                m.patch_count[s_][int(q)] += 1
                top = max(top, int(q))
        m.n_patches = top + 1

        def pair(text):
            a, b = text.split("-")
            return int(a), int(b)
        for d in get("generic_wertheim.deltas", required=True):
            a, b = pair(d["interaction"])
            m.delta[a][b] = m.delta[b][a] = delta_value(d) * u2i ** 3
        B2s = get("generic_wertheim.B2s", required=True)
        if len(B2s) != m.n_species * (m.n_species + 1) // 2 and not get("generic_wertheim.allow_unspecified_B2s", False):
            raise ValueError("The number of B2 coefficients differs from N (N + 1) / 2; set "
                             "generic_wertheim.allow_unspecified_B2s = true to run regardless")
        for b2 in B2s:
            a, b = pair(b2["interaction"])
            m.B2[a][b] = m.B2[b][a] = b2["value"] * u2i ** 3
    k = get("k", required=True)
    ks = k if isinstance(k, list) else [k] * m.n_species
    for i in range(m.n_species):
        m.k[i] = ks[i] * u2i ** 5
    return m


def desc_from_config(cfg, dims, precision="fp64"):
    """ch_desc from a reference TOML table (already parsed to nested dicts)."""
    get = _getter(cfg)

    u2i = 1.0 / get("distance_scaling_factor", 1.0)
    model = get("free_energy", required=True)
    kw = {}
    if model == "landau":
        eps = get("landau.epsilon")
        if eps is None:
            T = get("landau.T")
            if T is None:
                raise KeyError("The Landau free energy requires either 'epsilon' or 'T' to be set")
            eps = 1 - T
        if u2i != 1.0:
            raise ValueError("The Landau free energy does not support distance_scaling_factor values different from 1.0")
        kw["epsilon"] = eps
    elif model == "simple_wertheim":
        delta = delta_value(get("wertheim.delta", required=True))
        kw.update(valence=int(get("wertheim.valence", required=True)), B2=get("wertheim.B2", required=True) * u2i ** 3,
                  delta=delta * u2i ** 3, regularisation_delta=get("wertheim.regularisation_delta", 0.0) / u2i ** 3)
    elif model not in L.MODELS_SPECIES:
        raise ValueError(f"Unsupported free energy model '{model}'")
    k = get("k", required=True)
    if isinstance(k, list):
        k = k[0]
    mob = get("mobility.type", "constant")
    return make_desc(dims=dims, N=int(get("N", required=True)), integrator=get("integrator", "euler"),
                     free_energy=model, precision=precision, dt=get("dt", required=True),
                     dx=get("dx", 1.0) * u2i, k=k * u2i ** 5, mobility=mob, M=get("mobility.M", 1.0) / u2i,
                     rho_min=get("mobility.rho_min", 0.0) / u2i, rho_ref=get("mobility.rho_ref", 1.0),
                     # pseudospectral rescales S (PseudospectralCPU.cpp:22-23); pseudospectral_mobility
                     # does not (PseudospectralMobilityCPU.cpp:22)
                     S=get("pseudospectral.S", 0.0) * (u2i ** 3 if get("integrator", "euler") == "pseudospectral"
                                                       else 1.0),
                     use_dealias=get("pseudospectral.use_dealias", False),
                     with_noise=get("mobility.with_noise", False),
                     noise_rescale_factor=get("mobility.noise_rescale_factor", 1.0),
                     seed=int(get("seed", int(time.time()))) & 0xFFFFFFFFFFFFFFFF,
                     use_gmres=get("pseudospectral.use_gmres", False),
                     gmres_restart=int(get("pseudospectral.gmres_restart", 30)),
                     gmres_max_iter=int(get("pseudospectral.gmres_max_iter", 200)),
                     gmres_tol=get("pseudospectral.gmres_tol", 1e-10), density_conversion=u2i ** 3, **kw)


class System:
    """One Cahn-Hilliard system resident on the GPU."""

    def __init__(self, ctx, desc, species_model=None):
        """species_model: a ChSpeciesModel for the multi-species Wertheim models;
        fields are then species-major arrays of n_species * N^dims values."""
        self.ctx, self.desc = ctx, desc
        h = ctypes.c_void_p()
        if species_model is None:
            st = L.lib().ch_sys_create(ctx.handle, ctypes.byref(desc), ctypes.byref(h))
        else:
            st = L.lib().ch_sys_create_species(ctx.handle, ctypes.byref(desc), ctypes.byref(species_model),
                                               ctypes.byref(h))
        _check(ctx.handle, st)
        self.handle = h
        ctx._children.add(self)
        self.species = int(L.lib().ch_sys_species(h))
        self.bins = desc.N ** desc.dims * self.species

    def _chk(self, st):
        _check(self.ctx.handle, st)

    def upload(self, rho):
        a = np.ascontiguousarray(rho, dtype=np.float64).ravel()
        if a.size != self.bins:
            raise ValueError(f"expected {self.bins} values, got {a.size}")
        self._chk(L.lib().ch_sys_upload(self.handle, a.ctypes.data_as(ctypes.c_void_p)))

    def download(self):
        out = np.empty(self.bins, dtype=np.float64)
        self._chk(L.lib().ch_sys_download(self.handle, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    sync = download

    def download_mobility(self):
        out = np.empty(self.bins, dtype=np.float64)
        self._chk(L.lib().ch_sys_download_mobility(self.handle, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def upload_device(self, ptr):
        self._chk(L.lib().ch_sys_upload_device(self.handle, ctypes.c_void_p(ptr)))

    def download_device(self, ptr):
        self._chk(L.lib().ch_sys_download_device(self.handle, ctypes.c_void_p(ptr)))

    def evolve(self, nsteps=1):
        self._chk(L.lib().ch_sys_evolve(self.handle, int(nsteps)))

    @property
    def steps_per_launch(self):
        return int(L.lib().ch_sys_steps_per_launch(self.handle))

    @property
    def region_shape(self):
        """Region shape index of the 2D wave kernel, -1 where it is not used."""
        return int(L.lib().ch_sys_region_shape(self.handle))

    @property
    def gmres_iterations(self):
        """GMRES iterations of the last pseudospectral_mobility step (use_gmres)."""
        return int(L.lib().ch_sys_gmres_iterations(self.handle))

    def observables(self):
        m, f = ctypes.c_double(), ctypes.c_double()
        self._chk(L.lib().ch_sys_observables(self.handle, ctypes.byref(m), ctypes.byref(f)))
        return m.value, f.value

    def queue_observables(self):
        """Enqueue the observables of the current field; read them later with
        fetch_observables (first in, first out)."""
        self._chk(L.lib().ch_sys_observables_queue(self.handle))

    def evolve_observe(self, nsteps):
        """evolve(nsteps) then queue_observables(), as one captured graph when
        the count recurs (ch_sys_evolve_observe)."""
        self._chk(L.lib().ch_sys_evolve_observe(self.handle, int(nsteps)))

    def fetch_observables(self):
        m, f = ctypes.c_double(), ctypes.c_double()
        self._chk(L.lib().ch_sys_observables_fetch(self.handle, ctypes.byref(m), ctypes.byref(f)))
        return m.value, f.value

    @property
    def observables_pending(self):
        return int(L.lib().ch_sys_observables_pending(self.handle))

    def average_mass(self):
        return self.observables()[0]

    def average_free_energy(self):
        return self.observables()[1]

    def structure_factor(self):
        """Circularly averaged S(q) of the current 2D field (utils/sq_from_traj.py:72-103)."""
        ns = L.lib().ch_sq_nshells(self.desc.N)
        out = np.empty(ns)
        self._chk(L.lib().ch_sys_structure_factor(self.handle, out.ctypes.data_as(ctypes.c_void_p), ns))
        return out

    def sq_loss(self, target, grad=True):
        """sum_s (S_s - target_s)^2 and (optionally) its gradient with respect to the field."""
        t = np.ascontiguousarray(target, dtype=np.float64)
        loss = ctypes.c_double()
        g = np.empty(self.bins) if grad else None
        self._chk(L.lib().ch_sys_sq_loss(self.handle, t.ctypes.data_as(ctypes.c_void_p), t.size, ctypes.byref(loss),
                                         g.ctypes.data_as(ctypes.c_void_p) if grad else None))
        return (loss.value, g) if grad else loss.value

    def record(self, nsteps, ckpt_every=100):
        """Advance nsteps keeping a checkpoint every ckpt_every steps (0: every state)."""
        self._chk(L.lib().ch_sys_record(self.handle, int(nsteps), int(ckpt_every)))

    def adjoint(self, dL_dpsiT):
        """Reverse sweep of the last record: returns (dL/dpsi0, grads) with
        grads[GRAD_EPSILON, GRAD_K, GRAD_M, GRAD_DELTA, GRAD_B2]."""
        g = np.ascontiguousarray(dL_dpsiT, dtype=np.float64).ravel()
        out = np.empty(self.bins, dtype=np.float64)
        grads = np.zeros(L.NGRADS)
        self._chk(L.lib().ch_sys_adjoint(self.handle, g.ctypes.data_as(ctypes.c_void_p),
                                         out.ctypes.data_as(ctypes.c_void_p), grads.ctypes.data_as(ctypes.c_void_p)))
        return out, grads

    def adjoint_device(self, dL_dpsiT_ptr, dL_dpsi0_ptr):
        grads = np.zeros(L.NGRADS)
        self._chk(L.lib().ch_sys_adjoint_device(self.handle, ctypes.c_void_p(dL_dpsiT_ptr), ctypes.c_void_p(dL_dpsi0_ptr),
                                                grads.ctypes.data_as(ctypes.c_void_p)))
        return grads

    def free_energy_grad(self):
        """(F, dF/dpsi, explicit dF/dparams) of the current field: F as
        average_free_energy (src/CahnHilliard.cpp:300-319)."""
        F = ctypes.c_double()
        g = np.empty(self.bins)
        dp = np.zeros(L.NGRADS)
        self._chk(L.lib().ch_sys_free_energy_grad(self.handle, ctypes.byref(F), g.ctypes.data_as(ctypes.c_void_p),
                                                  dp.ctypes.data_as(ctypes.c_void_p)))
        return F.value, g, dp

    def close(self):
        if getattr(self, "handle", None):
            L.lib().ch_sys_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


COMPENSATION = {"none": 0, "launch": 1, "stored": 2, "low8": 3}


class Batch:
    """Independent 2D Euler replicas (mixed N and parameters) advanced together.

    compensation: None for the library default (FP32: "launch", FP64: "none"),
    or "none" / "launch" / "stored" / "low8" (ch_batch_create_ex)."""

    def __init__(self, ctx, descs, compensation=None):
        self.ctx = ctx
        arr = (L.ChDesc * len(descs))(*descs)
        h = ctypes.c_void_p()
        if compensation is None:
            st = L.lib().ch_batch_create(ctx.handle, len(descs), arr, ctypes.byref(h))
        else:
            st = L.lib().ch_batch_create_ex(ctx.handle, len(descs), arr, COMPENSATION[compensation], ctypes.byref(h))
        _check(ctx.handle, st)
        self.handle, self.n = h, len(descs)
        ctx._children.add(self)
        self.bins = [d.N ** d.dims for d in descs]

    def _chk(self, st):
        _check(self.ctx.handle, st)

    def upload(self, r, rho):
        a = np.ascontiguousarray(rho, dtype=np.float64).ravel()
        if a.size != self.bins[r]:
            raise ValueError(f"replica {r}: expected {self.bins[r]} values, got {a.size}")
        self._chk(L.lib().ch_batch_upload(self.handle, r, a.ctypes.data_as(ctypes.c_void_p)))

    def download(self, r):
        out = np.empty(self.bins[r], dtype=np.float64)
        self._chk(L.lib().ch_batch_download(self.handle, r, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def evolve(self, nsteps=1):
        self._chk(L.lib().ch_batch_evolve(self.handle, int(nsteps)))

    def observables(self):
        m, f = np.empty(self.n), np.empty(self.n)
        self._chk(L.lib().ch_batch_observables(self.handle, m.ctypes.data_as(ctypes.c_void_p),
                                               f.ctypes.data_as(ctypes.c_void_p)))
        return m, f

    @property
    def steps_per_launch(self):
        return int(L.lib().ch_batch_steps_per_launch(self.handle))

    @property
    def compensation(self):
        c = int(L.lib().ch_batch_compensation(self.handle))
        return {v: k for k, v in COMPENSATION.items()}[c]

    def autotune(self):
        """Time the kernel's region splits and launch depths on the current
        fields; returns (shape, steps_per_launch) of the fastest."""
        sh, k = ctypes.c_int32(), ctypes.c_int32()
        self._chk(L.lib().ch_batch_autotune(self.handle, ctypes.byref(sh), ctypes.byref(k)))
        return sh.value, k.value

    def close(self):
        if getattr(self, "handle", None):
            L.lib().ch_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def observables_device_noalloc(ctx, desc, rho_ptr):
    """(mass, free energy) of a device field (int: address of N^dims elements of
    desc's precision) without any allocation (ch_observables_device_noalloc)."""
    m, f = ctypes.c_double(), ctypes.c_double()
    _check(ctx.handle, L.lib().ch_observables_device_noalloc(ctx.handle, ctypes.byref(desc), ctypes.c_void_p(rho_ptr),
                                                             ctypes.byref(m), ctypes.byref(f)))
    return m.value, f.value


def der_bulk_free_energy(ctx, desc, rho_ptr, out_ptr, n):
    """FreeEnergyModel::der_bulk_free_energy on device buffers (ints: device addresses)."""
    _check(ctx.handle, L.lib().ch_der_bulk_free_energy(ctx.handle, ctypes.byref(desc), ctypes.c_void_p(rho_ptr),
                                                       ctypes.c_void_p(out_ptr), int(n)))


def set_mobility(ctx, desc, rho_ptr, mob_ptr, n):
    """IMobility::update_mobility's pointwise rule on device buffers."""
    _check(ctx.handle, L.lib().ch_set_mobility(ctx.handle, ctypes.byref(desc), ctypes.c_void_p(rho_ptr),
                                               ctypes.c_void_p(mob_ptr), int(n)))
