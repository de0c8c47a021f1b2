"""Preset files of the reference UI over the C ABI (cf_preset_read / cf_preset_write).

A preset is a dict keyed like the file (CellFlowWidget::loadPreset / savePreset,
CellFlowWidget.cpp:1070-1269), holding only the step-path fields the file had:
scalars as float/int, "particleColors" as an (n, 3) float32 array, "radioByType" and
"rawForceTable" as float32 arrays.  Parsing and writing are native (cellflow_amd/csrc/
cf_preset.cpp); files are written in the layout Qt gives the reference's own presets.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _abi
from ._abi import check, lib

_INT_KEYS = {"PARTICLE_COUNT": "particle_count", "numParticleTypes": "num_particle_types"}
_FLOAT_KEYS = {"radius": "radius", "delta_t": "delta_t", "friction": "friction",
               "repulsion": "repulsion", "attraction": "attraction", "k": "k",
               "balance": "balance", "forceMultiplier": "force_multiplier",
               "forceRange": "force_range", "forceBias": "force_bias", "ratio": "ratio",
               "lfoA": "lfo_a", "lfoS": "lfo_s", "forceOffset": "force_offset",
               "canvasWidth": "canvas_width", "canvasHeight": "canvas_height",
               "canvasDepth": "canvas_depth", "spawnRegionSize": "spawn_region_size"}


def _bit(key: str) -> int:
    return 1 << _abi.PRESET_KEYS.index(key)


def _from_c(p: _abi.cf_preset) -> dict:
    out = {}
    for key, field in list(_INT_KEYS.items()) + list(_FLOAT_KEYS.items()):
        if p.present & _bit(key):
            out[key] = getattr(p, field)
    if p.present & _bit("particleColors"):
        out["particleColors"] = np.array([[c.r, c.g, c.b] for c in p.colors[: p.num_colors]],
                                         np.float32).reshape(-1, 3)
    if p.present & _bit("radioByType"):
        out["radioByType"] = np.array(p.radio_by_type[: p.num_radio], np.float32)
    if p.present & _bit("rawForceTable"):
        out["rawForceTable"] = np.array(p.raw_force_table[: p.num_raw], np.float32)
    return out


def _to_c(d: dict) -> _abi.cf_preset:
    p = _abi.cf_preset()
    lib().cf_preset_init(C.byref(p))
    for key, val in d.items():
        if key in _INT_KEYS:
            setattr(p, _INT_KEYS[key], int(val))
        elif key in _FLOAT_KEYS:
            setattr(p, _FLOAT_KEYS[key], float(val))
        elif key == "particleColors":
            cols = np.asarray(val, np.float32).reshape(-1, 3)
            if len(cols) > _abi.CF_MAX_TYPES:
                raise ValueError("at most 10 colours")
            p.num_colors = len(cols)
            for i, (r, g, b) in enumerate(cols):
                p.colors[i] = _abi.cf_color(r, g, b)
        elif key == "radioByType":
            v = np.asarray(val, np.float32).ravel()
            if len(v) > _abi.CF_MAX_TYPES:
                raise ValueError("at most 10 radius modifiers")
            p.num_radio = len(v)
            p.radio_by_type[: len(v)] = v.tolist()
        elif key == "rawForceTable":
            v = np.asarray(val, np.float32).ravel()
            if len(v) > _abi.CF_MAX_TYPES ** 2:
                raise ValueError("at most 100 raw table values")
            p.num_raw = len(v)
            p.raw_force_table[: len(v)] = v.tolist()
        else:
            raise KeyError(f"{key} is not a step-path preset field")
        p.present |= _bit(key)
    return p


def read_preset(path) -> dict:
    p = _abi.cf_preset()
    check(lib().cf_preset_read(os.fsencode(path), C.byref(p)))
    return _from_c(p)


def write_preset(path, preset: dict, template=None) -> None:
    """Write the preset's fields; with `template`, every other member of that file is kept."""
    p = _to_c(preset)
    check(lib().cf_preset_write(os.fsencode(path), C.byref(p),
                                None if template is None else os.fsencode(template)))


def preset_params(preset: dict):
    """SimulationParams of a preset at LFO phase 0 (fields it lacks keep their defaults)."""
    from .simulation import SimulationParams
    p = _to_c(preset)
    cp = _abi.cf_params()
    check(lib().cf_preset_params(C.byref(p), C.byref(cp)))
    return SimulationParams(radius=cp.radius, delta_t=cp.delta_t, friction=cp.friction,
                            repulsion=cp.repulsion, attraction=cp.attraction, k=cp.k,
                            balance=cp.balance, canvasWidth=cp.canvas_width,
                            canvasHeight=cp.canvas_height, canvasDepth=cp.canvas_depth,
                            numParticleTypes=cp.num_particle_types,
                            ratioWithLFO=cp.ratio_with_lfo,
                            forceMultiplier=cp.force_multiplier,
                            maxExpectedNeighbors=cp.max_expected_neighbors)
