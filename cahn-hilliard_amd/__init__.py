"""cahn-hilliard_amd: MI355X-native time-stepping for the cahn-hilliard phase-field code.

The compute path is libch_hip.so (hand-written gfx950 HIP kernels behind the C
ABI in include/ch_hip.h).  This Python package is a thin ctypes layer used by
the tests, the benchmark and the autograd wrapper; the drop-in C++ host code
lives in host/.
"""
from . import _lib
from ._lib import (CHError, EULER, EULER_MOBILITY, FP32, FP64, GRAD_B2, GRAD_DELTA, GRAD_EPSILON, GRAD_K, GRAD_M,
                   LANDAU, NGRADS, PSEUDOSPECTRAL, PSEUDOSPECTRAL_MOBILITY, SIMPLE_WERTHEIM, header_functions)
from ._lib import ChSpeciesModel, GENERIC_WERTHEIM, RICCI_WERTHEIM, SALEH_WERTHEIM
from .system import (Batch, Context, System, der_bulk_free_energy, desc_from_config, make_desc, set_mobility,
                     observables_device_noalloc, species_model_from_config, saleh_model, ricci_model, generic_model)

__all__ = ["CHError", "ChSpeciesModel", "species_model_from_config", "saleh_model", "ricci_model", "generic_model", "SALEH_WERTHEIM", "RICCI_WERTHEIM",
           "GENERIC_WERTHEIM", "Batch", "Context", "System", "make_desc", "desc_from_config", "der_bulk_free_energy",
           "set_mobility", "observables_device_noalloc", "header_functions", "EULER", "EULER_MOBILITY", "PSEUDOSPECTRAL", "PSEUDOSPECTRAL_MOBILITY", "LANDAU",
           "SIMPLE_WERTHEIM", "FP32", "FP64", "GRAD_EPSILON", "GRAD_K", "GRAD_M", "GRAD_DELTA", "GRAD_B2", "NGRADS"]
