"""MI355X walker step of miniAFQMC: C-ABI library libafqmc_hip.so plus host mirrors."""
from ._native import AfqmcError, MODES, lib  # noqa: F401  (raises if the library is not built)
from .afqmc import REFERENCE_TIMERS, AfqmcGroup, AfqmcSys, reference_timers  # noqa: F401
from .system import (PRESETS, System, Csr, fields, fields_block, make_system, multi_system, multi_systems,  # noqa: F401
                     preset_system)
