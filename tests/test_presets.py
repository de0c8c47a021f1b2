"""Preset files (CellFlowWidget::loadPreset / savePreset, CellFlowWidget.cpp:1070-1269)
through the native reader/writer, without a GPU.

The reference's four preset files were written by Qt's QJsonDocument::toJson; reading one
and writing it back with itself as the template must reproduce it byte for byte, which pins
the writer's layout and number formatting to Qt's.  Python's json module (same layout with
indent=4, sort_keys=True) cross-checks files written from random presets."""

import json
import os

import numpy as np
import pytest
from hypothesis import given, settings
from hypothesis import strategies as hs

from cellflow_amd import CellFlowError, _abi, preset_params, read_preset, write_preset

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "presets")
NAMES = ["settings", "eater", "littlecells", "pulser"]


def golden(name):
    return os.path.join(GOLDEN, name + ".json")


@pytest.mark.parametrize("name", NAMES)
def test_read_matches_json(name):
    doc = json.load(open(golden(name)))
    pr = read_preset(golden(name))
    engine_keys = set(_abi.PRESET_KEYS) & set(doc)
    assert set(pr) == engine_keys
    for k in engine_keys - {"particleColors", "radioByType", "rawForceTable"}:
        assert pr[k] == np.float32(doc[k]) if isinstance(pr[k], float) else pr[k] == doc[k], k
    np.testing.assert_array_equal(pr["rawForceTable"], np.float32(doc["rawForceTable"]))
    np.testing.assert_array_equal(pr["radioByType"], np.float32(doc["radioByType"]))
    want = np.float32([[c["r"], c["g"], c["b"]] for c in doc["particleColors"]])
    np.testing.assert_array_equal(pr["particleColors"], want)
    T = doc["numParticleTypes"]
    assert len(pr["rawForceTable"]) == T * T


@pytest.mark.parametrize("name", NAMES)
def test_roundtrip_with_template_is_byte_identical(tmp_path, name):
    out = tmp_path / "out.json"
    write_preset(out, read_preset(golden(name)), template=golden(name))
    assert out.read_bytes() == open(golden(name), "rb").read()


def test_write_without_template_has_only_engine_fields(tmp_path):
    pr = read_preset(golden("eater"))
    out = tmp_path / "p.json"
    write_preset(out, pr)
    text = out.read_text()
    doc = json.loads(text)
    assert set(doc) == set(pr)
    assert text == json.dumps(doc, indent=4, sort_keys=True) + "\n"
    back = read_preset(out)
    for k in pr:
        np.testing.assert_array_equal(back[k], pr[k])


def test_params_of_a_preset():
    pr = read_preset(golden("eater"))
    p = preset_params(pr)
    assert p.radius == np.float32(397) and p.numParticleTypes == 6
    assert p.delta_t == pr["delta_t"] and p.ratioWithLFO == pr["ratio"]
    assert p.maxExpectedNeighbors == 400
    # fields a preset lacks take SimulationParams' defaults (SimulationParams.h:15-29)
    q = preset_params({"radius": 5.0})
    assert q.radius == 5.0 and q.friction == np.float32(0.51) and q.canvasWidth == 8000.0


@pytest.mark.parametrize("text,what", [
    ("not json", "object"), ("[1, 2]", "object"), ('{"radius": "big"}', "number"),
    ('{"PARTICLE_COUNT": 10.5}', "integer"), ('{"rawForceTable": [1, "x"]}', "non-number"),
    ('{"particleColors": [1]}', "colour"), ('{"radioByType": {}}', "array"),
])
def test_malformed_presets_are_rejected(tmp_path, text, what):
    p = tmp_path / "bad.json"
    p.write_text(text)
    with pytest.raises(CellFlowError) as ei:
        read_preset(p)
    assert ei.value.code == _abi.CF_ERR_ARG and what in str(ei.value)
    with pytest.raises(CellFlowError):
        write_preset(tmp_path / "o.json", {"radius": 1.0}, template=p) if what == "object" \
            else read_preset(p)
    with pytest.raises(CellFlowError):
        read_preset(tmp_path / "missing.json")


_f32 = hs.floats(-1e6, 1e6, allow_nan=False, width=32)


@settings(max_examples=60, deadline=None)
@given(count=hs.integers(1, 2_000_000), T=hs.integers(1, 10), vals=hs.lists(_f32, min_size=18,
                                                                            max_size=18),
       data=hs.data())
def test_property_roundtrip_and_python_layout(tmp_path_factory, count, T, vals, data):
    keys = [k for k in _abi.PRESET_KEYS[2:20]]
    pr = {"PARTICLE_COUNT": count, "numParticleTypes": T}
    pr.update({k: float(v) for k, v in zip(keys, vals)})
    pr["rawForceTable"] = np.float32(data.draw(hs.lists(_f32, min_size=T * T, max_size=T * T)))
    pr["radioByType"] = np.float32(data.draw(hs.lists(_f32, min_size=T, max_size=T)))
    pr["particleColors"] = np.float32(data.draw(hs.lists(_f32, min_size=3 * T,
                                                         max_size=3 * T))).reshape(T, 3)
    out = tmp_path_factory.mktemp("p") / "p.json"
    write_preset(out, pr)
    back = read_preset(out)
    assert set(back) == set(pr)
    for k in pr:
        np.testing.assert_array_equal(back[k], pr[k], err_msg=k)
    text = out.read_text()
    doc = json.loads(text)
    assert text == json.dumps(doc, indent=4, sort_keys=True) + "\n"
    assert doc["rawForceTable"] == [float(x) for x in pr["rawForceTable"]]


@settings(max_examples=40, deadline=None)
@given(vals=hs.lists(hs.floats(allow_nan=False, allow_infinity=False, width=32), min_size=100,
                     max_size=100))
def test_property_number_text_is_python_repr(tmp_path_factory, vals):
    """Every float32 value (subnormal to huge) is written as the shortest decimal that reads
    back to it, in Python repr's layout."""
    out = tmp_path_factory.mktemp("n") / "p.json"
    raw = np.float32(vals)
    write_preset(out, {"numParticleTypes": 10, "rawForceTable": raw})
    text = out.read_text()
    assert text == json.dumps(json.loads(text), indent=4, sort_keys=True) + "\n"
    np.testing.assert_array_equal(read_preset(out)["rawForceTable"], raw)
    doc = json.loads(text)
    assert [float(x) for x in doc["rawForceTable"]] == [float(x) for x in raw]


COMMA_LOCALES = ["de_DE.UTF-8", "de_DE.utf8", "de_DE", "fr_FR.UTF-8", "fr_FR.utf8", "fr_FR",
                 "ru_RU.UTF-8", "nl_NL.UTF-8", "pt_BR.UTF-8", "it_IT.UTF-8"]


@pytest.fixture
def comma_locale():
    """LC_NUMERIC switched to an installed comma-decimal locale (as a Qt application sets it
    from the environment) for the duration of the test; skipped when none is installed."""
    import locale
    old = locale.setlocale(locale.LC_NUMERIC)
    for name in COMMA_LOCALES:
        try:
            locale.setlocale(locale.LC_NUMERIC, name)
        except locale.Error:
            continue
        if locale.localeconv()["decimal_point"] == ",":
            break
        locale.setlocale(locale.LC_NUMERIC, old)
    else:
        pytest.skip("no comma-decimal locale is installed")
    yield
    locale.setlocale(locale.LC_NUMERIC, old)


@pytest.mark.parametrize("name", NAMES)
def test_roundtrip_under_a_comma_decimal_locale(tmp_path, comma_locale, name):
    """The reader and writer convert numbers without the C locale: under a comma-decimal
    LC_NUMERIC a preset still reads its fractions and writes back byte for byte."""
    pr = read_preset(golden(name))
    doc = json.load(open(golden(name)))
    assert pr["radius"] == np.float32(doc["radius"])
    out = tmp_path / "out.json"
    write_preset(out, pr, template=golden(name))
    assert out.read_bytes() == open(golden(name), "rb").read()


def test_tuning_cache_under_a_comma_decimal_locale(tmp_path, comma_locale):
    from cellflow_amd import tuning_cache_lookup, tuning_cache_store
    p = tmp_path / "tune.json"
    t = dict(cell_divisor=1, force_variant=3, lanes_per_particle=8, block_size=256,
             steps_per_graph=32, sort_variant=1)
    tuning_cache_store(p, "k", t, 46.25)
    json.load(open(p))  # valid JSON: '.' as the decimal point
    assert tuning_cache_lookup(p, "k") == (t, 46.25)


@pytest.mark.parametrize("text,value", [("1e-400", 0.0), ("-1e-400", -0.0), ("0.25", 0.25),
                                        ("-2.5E+3", -2500.0), ("1e999", None)])
def test_number_range(tmp_path, text, value):
    """Numbers below the double range read as a zero of their sign; numbers above it are
    rejected rather than read as infinity."""
    p = tmp_path / "p.json"
    p.write_text('{"radius": %s}\n' % text)
    if value is None:
        with pytest.raises(CellFlowError):
            read_preset(p)
    else:
        r = read_preset(p)["radius"]
        assert r == value and np.signbit(r) == np.signbit(value)


@pytest.mark.parametrize("name", NAMES)
def test_sanitized_roundtrip(tmp_path, name):
    """The reader and writer under AddressSanitizer and UndefinedBehaviorSanitizer (a g++
    build of the host-only sources): a golden preset reads and writes back byte for byte."""
    import host_sanitizer
    out = tmp_path / "out.json"
    text = host_sanitizer.run("preset", golden(name), out)
    assert "read rc=0" in text and "write rc=0" in text
    assert out.read_bytes() == open(golden(name), "rb").read()


@pytest.mark.parametrize("text", [
    "", "not json", "[1, 2]", '{"radius": "big"}', '{"PARTICLE_COUNT": 10.5}',
    '{"rawForceTable": [1, "x"]}', '{"particleColors": [1]}', '{"radioByType": {}}',
    '{"radius": 1e999}', '{"radius": ' + "[" * 200 + "]" * 200 + "}", '{"a": "\\ud800"}',
    '{"rawForceTable": [' + ", ".join(["0.5"] * 150) + "]}", '{"numParticleTypes": 99}',
    '{"particleColors": [' + ", ".join(['{"r": 1, "g": 0, "b": 0}'] * 12) + "]}",
    '{"radius": 0.25', '﻿{"radius": 0.25}', '{"radius": 0.25}\x00',
])
def test_sanitized_malformed_presets(tmp_path, text):
    """Malformed and oversized preset files under the sanitizers: each is read (and written
    over with itself as template) without a memory or UB error, whatever status it gets."""
    import host_sanitizer
    p = tmp_path / "bad.json"
    p.write_bytes(text.encode("utf-8", "surrogatepass"))
    out = host_sanitizer.run("preset", p, tmp_path / "o.json")
    assert "read rc=" in out
