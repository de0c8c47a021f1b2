"""MI355X-native particle-life step engine (HIP kernels for gfx950 behind a C ABI).

Public Python surface: :class:`ParticleSimulation` (mirror of the reference's class of the
same name), :class:`Ensemble`, :class:`SimulationParams`, and the seeded workload
generators in :mod:`cellflow_amd.configs`.
"""

from ._abi import CellFlowError, LIB_PATH  # noqa: F401
from .presets import preset_params, read_preset, write_preset  # noqa: F401
from .simulation import (PARTICLE_DTYPE, Ensemble, ParticleSimulation,  # noqa: F401
                         SimulationParams, cutoff_threshold, device_count,
                         force_table_transform, grid_dims, particles_from_state,
                         tuning_cache_lookup, tuning_cache_store)
