"""ctypes declarations of ``include/cellflow.h`` (the engine's C ABI).

Loads ``cellflow_amd/libcellflow.so`` built in-tree by ``cellflow_amd/csrc/Makefile``.
There is no fallback: if the library is missing or a call fails, an exception is raised.
"""

from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcellflow.so")

CF_OK, CF_ERR_ARG, CF_ERR_DEVICE, CF_ERR_CAPACITY, CF_ERR_NONFINITE, CF_ERR_HANDLE, \
    CF_ERR_STATE = range(7)
CF_MAX_TYPES = 10
CF_TUNE_FORCE = 1
CF_MAX_CONNECTIONS = 32

ERROR_NAMES = {1: "CF_ERR_ARG", 2: "CF_ERR_DEVICE", 3: "CF_ERR_CAPACITY",
               4: "CF_ERR_NONFINITE", 5: "CF_ERR_HANDLE", 6: "CF_ERR_STATE"}


class CellFlowError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{ERROR_NAMES.get(code, code)}: {message}")
        self.code = code


class cf_particle(C.Structure):
    """struct Particle (cuda-native/include/SimulationParams.h:6-12), 44 bytes."""
    _fields_ = [("pos", C.c_float * 3), ("vel", C.c_float * 3), ("acc", C.c_float * 3),
                ("ptype", C.c_uint32), ("pad", C.c_float)]


class cf_color(C.Structure):
    _fields_ = [("r", C.c_float), ("g", C.c_float), ("b", C.c_float)]


class cf_params(C.Structure):
    _fields_ = [("radius", C.c_float), ("delta_t", C.c_float), ("friction", C.c_float),
                ("repulsion", C.c_float), ("attraction", C.c_float), ("k", C.c_float),
                ("balance", C.c_float), ("canvas_width", C.c_float),
                ("canvas_height", C.c_float), ("canvas_depth", C.c_float),
                ("num_particle_types", C.c_int), ("ratio_with_lfo", C.c_float),
                ("force_multiplier", C.c_float), ("max_expected_neighbors", C.c_int)]


class cf_edge_opts(C.Structure):
    _fields_ = [("distance", C.c_float), ("max_connections", C.c_int)]


class cf_tuning(C.Structure):
    _fields_ = [("cell_divisor", C.c_int), ("force_variant", C.c_int),
                ("lanes_per_particle", C.c_int), ("block_size", C.c_int),
                ("steps_per_graph", C.c_int),
                ("sort_variant", C.c_int)]


class cf_preset(C.Structure):
    _fields_ = [("present", C.c_uint32), ("particle_count", C.c_int),
                ("num_particle_types", C.c_int)] + \
        [(k, C.c_float) for k in ("radius", "delta_t", "friction", "repulsion", "attraction", "k",
                                  "balance", "force_multiplier", "force_range", "force_bias",
                                  "force_offset", "ratio", "lfo_a", "lfo_s", "canvas_width",
                                  "canvas_height", "canvas_depth", "spawn_region_size")] + \
        [("num_colors", C.c_int), ("colors", cf_color * CF_MAX_TYPES), ("num_radio", C.c_int),
         ("radio_by_type", C.c_float * CF_MAX_TYPES), ("num_raw", C.c_int),
         ("raw_force_table", C.c_float * (CF_MAX_TYPES * CF_MAX_TYPES))]


# presence bits of cf_preset, in declaration order of include/cellflow.h
PRESET_KEYS = ["PARTICLE_COUNT", "numParticleTypes", "radius", "delta_t", "friction", "repulsion",
               "attraction", "k", "balance", "forceMultiplier", "forceRange", "forceBias", "ratio",
               "lfoA", "lfoS", "forceOffset", "canvasWidth", "canvasHeight", "canvasDepth",
               "spawnRegionSize", "particleColors", "radioByType", "rawForceTable"]

assert C.sizeof(cf_particle) == 44

P = C.POINTER
_i, _f, _u64, _ll, _vp = C.c_int, C.c_float, C.c_uint64, C.c_longlong, C.c_void_p

# name -> (restype, argtypes)
SIGNATURES = {
    "cf_create": (_i, [_i, _i, P(_vp)]),
    "cf_destroy": (_i, [_vp]),
    "cf_set_particle_count": (_i, [_vp, _i]),
    "cf_set_num_particle_types": (_i, [_vp, _i]),
    "cf_get_particle_count": (_i, [_vp, P(_i)]),
    "cf_get_num_particle_types": (_i, [_vp, P(_i)]),
    "cf_set_particles": (_i, [_vp, P(cf_particle), _i]),
    "cf_get_particles": (_i, [_vp, P(cf_particle), _i]),
    "cf_set_state_soa": (_i, [_vp, P(_f), P(_f), P(C.c_uint32), _i]),
    "cf_get_state_soa": (_i, [_vp, P(_f), P(_f), P(C.c_uint32), _i]),
    "cf_set_neighbor_counts": (_i, [_vp, P(C.c_int32), _i]),
    "cf_get_neighbor_counts": (_i, [_vp, P(C.c_int32), _i]),
    "cf_init_particles": (_i, [_vp, _u64, P(_f), _f]),
    "cf_move_universe": (_i, [_vp, _f, _f, _f, P(_f)]),
    "cf_set_force_table": (_i, [_vp, P(_f), _i]),
    "cf_get_force_table": (_i, [_vp, P(_f), _i]),
    "cf_set_raw_force_table": (_i, [_vp, P(_f), _i]),
    "cf_get_raw_force_table": (_i, [_vp, P(_f), _i]),
    "cf_update_force_table": (_i, [_vp, _f, _f, _f]),
    "cf_set_radio_by_type": (_i, [_vp, P(_f), _i]),
    "cf_set_radio_by_type_value": (_i, [_vp, _i, _f]),
    "cf_get_radio_by_type": (_i, [_vp, P(_f), _i]),
    "cf_rotate_radio_by_type": (_i, [_vp]),
    "cf_simulate": (_i, [_vp, P(cf_params), _i, P(cf_edge_opts)]),
    "cf_build_edges": (_i, [_vp, _f, _i]),
    "cf_get_edges": (_i, [_vp, P(C.c_int32), _ll, P(_ll)]),
    "cf_get_step_edges": (_i, [_vp, P(C.c_int32), _ll, P(_ll)]),
    "cf_get_line_vertices": (_i, [_vp, P(cf_color), _i, P(_f), _ll, P(_ll)]),
    "cf_build_triangles": (_i, [_vp, _f, _i]),
    "cf_get_triangles": (_i, [_vp, P(C.c_int32), _ll, P(_ll)]),
    "cf_get_triangle_vertices": (_i, [_vp, P(cf_color), _i, P(_f), _ll, P(_ll)]),
    "cf_write_line_vertices_device": (_i, [_vp, P(cf_color), _i, _vp, _ll, P(_ll)]),
    "cf_write_triangle_vertices_device": (_i, [_vp, P(cf_color), _i, _vp, _ll, P(_ll)]),
    "cf_get_grid": (_i, [_vp, P(_i), P(_i)]),
    "cf_get_sort_order": (_i, [_vp, P(C.c_uint32), P(C.c_uint32), _i]),
    "cf_profile_steps": (_i, [_vp, P(cf_params), _i, P(cf_edge_opts), P(C.c_double), _i, P(_i),
                              C.c_char_p, _i]),
    "cf_pair_histogram": (_i, [_vp, _i, _f, P(C.c_double)]),
    "cf_pair_histogram_noalloc": (_i, [_vp, _i, _f, P(C.c_double)]),
    "cf_rollout_grad": (_i, [_vp, P(cf_params), _i, _i, _i, _f, P(C.c_double), P(C.c_double),
                             P(C.c_double), P(C.c_double), P(C.c_double)]),
    "cf_set_tuning": (_i, [_vp, P(cf_tuning)]),
    "cf_get_tuning": (_i, [_vp, P(cf_tuning)]),
    "cf_autotune": (_i, [_vp, P(cf_params), P(cf_edge_opts), _i, C.c_char_p, _i, P(cf_tuning),
                         P(C.c_double), C.c_char_p, _i]),
    "cf_tuning_key": (_i, [_vp, P(cf_params), P(cf_edge_opts), C.c_char_p, _i]),
    "cf_tuning_cache_lookup": (_i, [C.c_char_p, C.c_char_p, P(cf_tuning), P(C.c_double), P(_i)]),
    "cf_tuning_cache_store": (_i, [C.c_char_p, C.c_char_p, P(cf_tuning), C.c_double]),
    "cf_preset_init": (None, [P(cf_preset)]),
    "cf_preset_read": (_i, [C.c_char_p, P(cf_preset)]),
    "cf_preset_write": (_i, [C.c_char_p, P(cf_preset), C.c_char_p]),
    "cf_preset_params": (_i, [P(cf_preset), P(cf_params)]),
    "cf_grid_dims": (_i, [P(cf_params), P(_f), _i, _i, P(_i)]),
    "cf_cutoff_threshold": (_f, [_f]),
    "cf_force_table_transform": (None, [P(_f), _i, _f, _f, _f, P(_f)]),
    "cf_last_error_string": (C.c_char_p, []),
    "cf_version": (_i, []),
    "cf_device_count": (_i, [P(_i)]),
    "cf_ensemble_create": (_i, [_i, P(_i), P(_i), P(_vp)]),
    "cf_ensemble_destroy": (_i, [_vp]),
    "cf_ensemble_set_sim": (_i, [_vp, _i, P(cf_params), P(_f), P(_f), P(cf_particle), _i]),
    "cf_ensemble_simulate": (_i, [_vp, _i]),
    "cf_ensemble_get_particles": (_i, [_vp, _i, P(cf_particle), _i]),
    "cf_ensemble_get_neighbor_counts": (_i, [_vp, _i, P(C.c_int32), _i]),
    "cf_ensemble_set_tuning": (_i, [_vp, P(cf_tuning)]),
    "cf_ensemble_get_tuning": (_i, [_vp, P(cf_tuning)]),
    "cf_ensemble_autotune": (_i, [_vp, _i, C.c_char_p, _i, P(cf_tuning), P(C.c_double),
                                  C.c_char_p, _i]),
}

_lib = None


def lib():
    """The loaded engine library.  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CellFlowError(CF_ERR_DEVICE,
                                f"native library {LIB_PATH} is missing; build it with "
                                "`python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != CF_OK:
        msg = lib().cf_last_error_string()
        raise CellFlowError(rc, msg.decode() if msg else "")
