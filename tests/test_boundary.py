"""CPU-side checks of the drop-in boundary: the C-ABI library loads and exports
every symbol include/ch_hip.h declares; descriptors built from reference TOML
keys carry the reference's unit conventions.  No GPU calls."""
import ctypes
import os

import pytest

from conftest import REPO


def test_library_exports_every_header_symbol(ch):
    L = ch._lib.lib()
    names = ch.header_functions()
    assert "ch_sys_evolve" in names and len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert missing == []


def test_every_header_in_include_is_exported(ch):
    L = ch._lib.lib()
    inc = os.path.join(REPO, "include")
    for h in os.listdir(inc):
        if h.endswith(".h"):
            for n in ch.header_functions(os.path.join(inc, h)):
                assert hasattr(L, n), (h, n)


def test_abi_version_and_status_strings_without_gpu(ch):
    L = ch._lib.lib()
    assert L.ch_abi_version() == 9
    assert L.ch_status_string(3) == b"NaN encountered"


def test_ctx_create_rejects_null_out(ch):
    L = ch._lib.lib()
    assert L.ch_ctx_create(0, None) == ch._lib.CH_EINVAL


def test_desc_from_reference_config(ch):
    cfg = {"N": 64, "dt": 0.001, "dx": 1.0, "k": 1.0, "free_energy": "landau", "integrator": "pseudospectral",
           "landau": {"T": 0.1}, "pseudospectral": {"S": 0.5}}
    d = ch.desc_from_config(cfg, dims=2)
    assert (d.dims, d.N, d.integrator, d.model) == (2, 64, ch.PSEUDOSPECTRAL, ch.LANDAU)
    assert d.epsilon == pytest.approx(0.9) and d.S == 0.5 and d.M == 1.0
    w = {"N": 32, "dt": 1e-4, "k": [2.0], "free_energy": "simple_wertheim", "distance_scaling_factor": 2.0,
         "mobility": {"type": "free_energy", "M": 3.0},
         "wertheim": {"valence": 4, "B2": 8.0, "delta": {"T": 300, "deltaH": -42200, "deltaS": -119.1, "salt": 0.5,
                                                        "sticky_size": 6}}}
    d = ch.desc_from_config(w, dims=3)
    # src/integrators/Integrator.cpp:24-27 (k ~ m^5), SimpleWertheim.cpp:29-31 (B2 ~ m^3),
    # IMobility FreeEnergyMobility.h:21 (M ~ m^-1)
    assert d.k == pytest.approx(2.0 / 32) and d.B2 == pytest.approx(1.0) and d.M == pytest.approx(6.0)
    assert d.mobility == 1 and d.valence == 4 and d.delta > 0
    with pytest.raises(KeyError):
        ch.desc_from_config({"N": 8, "dt": 0.1, "free_energy": "landau", "landau": {}}, dims=2)


def test_noise_keys_from_reference_config(ch):
    # src/integrators/EulerMobilityCPU.cpp:10-18: mobility.with_noise,
    # mobility.noise_rescale_factor (default 1.0), seed
    cfg = {"N": 64, "dt": 0.01, "k": 1.0, "free_energy": "landau", "landau": {"epsilon": 0.9},
           "integrator": "euler_mobility", "seed": 1234,
           "mobility": {"type": "constant", "with_noise": True, "noise_rescale_factor": 0.25}}
    d = ch.desc_from_config(cfg, dims=2)
    assert (d.with_noise, d.noise_rescale_factor, d.seed) == (1, 0.25, 1234)
    del cfg["mobility"]["noise_rescale_factor"]
    assert ch.desc_from_config(cfg, dims=2).noise_rescale_factor == 1.0
    assert ch.desc_from_config({**cfg, "mobility": {"type": "constant"}}, dims=2).with_noise == 0


@pytest.mark.parametrize("cname,pyname", [("ch_desc", "ChDesc"), ("ch_species_model", "ChSpeciesModel")])
def test_desc_layout_matches_header(ch, tmp_path, cname, pyname):
    """sizeof / offsetof of every field of the ABI structs, from the C header, equal the ctypes mirror."""
    cls = getattr(ch._lib, pyname)
    fields = [f for f, _ in cls._fields_]
    src = tmp_path / "layout.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "ch_hip.h"\nint main(void) {\n'
                   f'    printf("%zu\\n", sizeof({cname}));\n' +
                   "".join(f'    printf("%zu\\n", offsetof({cname}, {f}));\n' for f in fields) + "    return 0;\n}\n")
    exe = tmp_path / "layout"
    import subprocess
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(REPO, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    want = [ctypes.sizeof(cls)] + [getattr(cls, f).offset for f in fields]
    assert got == want


def test_pseudospectral_mobility_keys_from_reference_config(ch):
    # PseudospectralMobilityCPU.cpp:22-23 (S read without the m^3 rescaling), :203-208 (GMRES keys)
    cfg = {"N": 64, "dt": 0.01, "k": 1.0, "free_energy": "landau", "landau": {"epsilon": 0.9},
           "integrator": "pseudospectral_mobility", "distance_scaling_factor": 1.0,
           "mobility": {"type": "regularised", "rho_min": 0.2},
           "pseudospectral": {"S": 0.5, "use_dealias": True, "use_gmres": True, "gmres_tol": 1e-8}}
    d = ch.desc_from_config(cfg, dims=2)
    assert d.integrator == ch.PSEUDOSPECTRAL_MOBILITY and d.S == 0.5 and d.use_dealias == 1
    assert (d.use_gmres, d.gmres_restart, d.gmres_max_iter, d.gmres_tol) == (1, 30, 200, 1e-8)


def test_species_model_from_reference_config(ch):
    import oracle as O
    cfg = {"N": 32, "dt": 1e-4, "k": [1.0, 2.0, 3.0], "free_energy": "generic_wertheim", "integrator": "euler",
           "generic_wertheim": {"species": [{"patches": [0, 0, 1]}, {"patches": [1, 2]}, {"patches": [2]}],
                                "deltas": [{"interaction": "0-1", "value": 800.0}, {"interaction": "2-2", "value": 50.0}],
                                "B2s": [{"interaction": f"{a}-{b}", "value": 0.1 * (a + b)}
                                        for a in range(3) for b in range(a, 3)]}}
    m = ch.species_model_from_config(cfg)
    assert (m.model, m.n_species, m.n_patches) == (ch.GENERIC_WERTHEIM, 3, 3)
    assert list(m.k)[:3] == [1.0, 2.0, 3.0] and m.patch_count[0][0] == 2 and m.patch_count[0][1] == 1
    assert m.delta[1][0] == 800.0 and m.B2[2][1] == pytest.approx(0.3)
    assert ch.desc_from_config(cfg, dims=2).model == ch.GENERIC_WERTHEIM
    # the product's struct and the oracle's have one layout
    assert ctypes.sizeof(ch.ChSpeciesModel) == ctypes.sizeof(O.SpeciesModel)
    sal = ch.species_model_from_config({"N": 8, "dt": 1e-4, "k": 1.0, "free_energy": "saleh",
                                         "saleh": {"valence": [4, 4, 2], "B2": 1.0, "delta_AA": 1e3,
                                                   "delta_BB": {"value": 2e3}}})
    assert list(sal.saleh_valence) == [4, 4, 2] and sal.saleh_delta_BB == 2e3 and list(sal.k)[:3] == [1.0] * 3


def test_header_constants_match_the_bindings(ch):
    import re
    L = ch._lib
    text = open(L.HEADER).read()
    consts = dict(re.findall(r"#define\s+(CH_[A-Z_]+)\s+(\d+)", text))
    assert int(consts["CH_OBS_QUEUE"]) == L.OBS_QUEUE
    assert int(consts["CH_MAX_SPECIES"]) == L.MAX_SPECIES and int(consts["CH_MAX_PATCHES"]) == L.MAX_PATCHES
