"""The autotuner's JSON cache file through the C ABI (host-only calls, no GPU).

The C++ reader is checked against Python's json module: files written by json.dumps in
several styles must read back to the same entries, and malformed files must be rejected
without being overwritten."""

import json
import os

import pytest
from hypothesis import given, settings
from hypothesis import strategies as hs

from cellflow_amd import CellFlowError, _abi, tuning_cache_lookup, tuning_cache_store

T0 = dict(cell_divisor=1, force_variant=0, lanes_per_particle=16, block_size=256,
          steps_per_graph=32, sort_variant=1)
T1 = dict(cell_divisor=2, force_variant=1, lanes_per_particle=4, block_size=64,
          steps_per_graph=1, sort_variant=0)


def test_missing_file_is_empty_cache(tmp_path):
    assert tuning_cache_lookup(tmp_path / "none.json", "k") is None


def test_store_lookup_roundtrip_and_merge(tmp_path):
    p = tmp_path / "tune.json"
    tuning_cache_store(p, "gfx950/sims=0/n=16/rho=1/T=8/e=16", T0, 74.25)
    tuning_cache_store(p, "gfx950/sims=8/n=21/rho=1/T=8/e=0", T1, 1234.5)
    tuning_cache_store(p, "gfx950/sims=0/n=16/rho=1/T=8/e=16", T1, 70.0)  # replaces
    assert tuning_cache_lookup(p, "gfx950/sims=0/n=16/rho=1/T=8/e=16") == (T1, 70.0)
    assert tuning_cache_lookup(p, "gfx950/sims=8/n=21/rho=1/T=8/e=0") == (T1, 1234.5)
    assert tuning_cache_lookup(p, "other") is None
    doc = json.load(open(p))
    assert doc["format"] == "cellflow-tuning" and doc["version"] == 1
    assert list(doc["entries"]) == sorted(doc["entries"])
    assert not [f for f in os.listdir(tmp_path) if ".tmp." in f]


def test_reads_files_written_by_python_json(tmp_path):
    p = tmp_path / "tune.json"
    doc = {"version": 7, "note": ["x", {"y": None, "z": [1.5e3, -2, True, False]}],
           "format": "cellflow-tuning",
           "entries": {"aé中\U0001f600/\"q\"\\": dict(T1, us_per_step=3.25, extra={"k": []}),
                       # no timing: reads as 0; no sort_variant (older files): reads as 1
                       "b": {k: v for k, v in T0.items() if k != "sort_variant"},
                       "bad-values": dict(T0, block_size=96),
                       "bad-sort": dict(T0, sort_variant=2),
                       "not-an-object": 5,
                       "incomplete": {"cell_divisor": 1}}}
    for style in (dict(indent=None), dict(indent=4), dict(ensure_ascii=False),
                  dict(separators=(",", ":"))):
        p.write_text(json.dumps(doc, **style), encoding="utf-8")
        assert tuning_cache_lookup(p, "aé中\U0001f600/\"q\"\\") == (T1, 3.25)
        assert tuning_cache_lookup(p, "b") == (T0, 0.0)
        for k in ("bad-values", "bad-sort", "not-an-object", "incomplete"):
            assert tuning_cache_lookup(p, k) is None
    # a store keeps the readable entries and drops the unusable ones
    tuning_cache_store(p, "c", T0, 1.0)
    ents = json.load(open(p, encoding="utf-8"))["entries"]
    assert set(ents) == {"aé中\U0001f600/\"q\"\\", "b", "c"}


@pytest.mark.parametrize("text", [
    "", "[]", "{", '{"format": "cellflow-tuning",}', '{"format": "other", "entries": {}}',
    '{"entries": {}}', '{"format": "cellflow-tuning", "entries": []}',
    '{"format": "cellflow-tuning", "entries": {"a": {"cell_divisor": 01}}}',
    '{"format": "cellflow-tuning"} trailing', '{"format": "cellflow-tuning\n"}',
    '{"format": "cellflow-tuning", "x": "\\ud800"}',
    '{"format": "cellflow-tuning", "x": ' + "[" * 40 + "]" * 40 + "}",
])
def test_malformed_files_are_rejected_and_kept(tmp_path, text):
    p = tmp_path / "tune.json"
    p.write_text(text)
    with pytest.raises(CellFlowError) as ei:
        tuning_cache_lookup(p, "k")
    assert ei.value.code == _abi.CF_ERR_ARG
    with pytest.raises(CellFlowError):
        tuning_cache_store(p, "k", T0, 1.0)
    assert p.read_text() == text


def test_store_rejects_invalid_tuning(tmp_path):
    with pytest.raises(CellFlowError):
        tuning_cache_store(tmp_path / "t.json", "k", dict(T0, lanes_per_particle=12), 1.0)
    with pytest.raises(CellFlowError):
        tuning_cache_store(tmp_path / "t.json", "k", T0, float("nan"))
    assert not (tmp_path / "t.json").exists()


_tunings = hs.fixed_dictionaries(dict(
    cell_divisor=hs.sampled_from([1, 2]), force_variant=hs.sampled_from([0, 1, 2]),
    lanes_per_particle=hs.sampled_from([4, 8, 16, 32]), block_size=hs.sampled_from([64, 128, 256]),
    steps_per_graph=hs.integers(1, 1 << 20), sort_variant=hs.sampled_from([0, 1])))


@settings(max_examples=60, deadline=None)
@given(ents=hs.dictionaries(hs.text(hs.characters(blacklist_categories=("Cs",),
                                                  blacklist_characters="\x00"),
                                    min_size=1, max_size=12), hs.tuples(
    _tunings, hs.floats(0, 1e7, allow_nan=False).map(lambda x: round(x, 3))), max_size=6),
    indent=hs.sampled_from([None, 0, 2]), ascii_=hs.booleans())
def test_property_python_json_roundtrip(tmp_path_factory, ents, indent, ascii_):
    p = tmp_path_factory.mktemp("c") / "tune.json"
    doc = {"format": "cellflow-tuning", "version": 1,
           "entries": {k: dict(t, us_per_step=u) for k, (t, u) in ents.items()}}
    p.write_text(json.dumps(doc, indent=indent, ensure_ascii=ascii_), encoding="utf-8")
    for k, (t, u) in ents.items():
        assert tuning_cache_lookup(p, k) == (t, u)
    # and the C++ writer's output reads back through Python's json
    tuning_cache_store(p, "\u0001ctl", T0, 2.5)
    back = json.load(open(p, encoding="utf-8"))["entries"]
    assert back["\u0001ctl"] == dict(T0, us_per_step=2.5)
    for k, (t, u) in ents.items():
        if k != "\u0001ctl":
            assert back[k] == dict(t, us_per_step=u)


@pytest.mark.parametrize("text", [
    None, "", "[]", "{", '{"format": "cellflow-tuning",}', '{"format": "other", "entries": {}}',
    '{"format": "cellflow-tuning", "entries": {"a": {"cell_divisor": 01}}}',
    '{"format": "cellflow-tuning", "x": "\\ud800"}',
    '{"format": "cellflow-tuning", "x": ' + "[" * 40 + "]" * 40 + "}",
    json.dumps({"format": "cellflow-tuning", "version": 1,
                "entries": {"k": dict(T0, us_per_step=1.5), "bad": dict(T0, block_size=96),
                            "é\U0001f600": 3}}),
])
def test_sanitized_cache_reader_and_writer(tmp_path, text):
    """The cache reader and writer under AddressSanitizer and UndefinedBehaviorSanitizer (a
    g++ build of the host-only sources), on missing, malformed and valid files."""
    import host_sanitizer
    p = tmp_path / "tune.json"
    if text is not None:
        p.write_text(text, encoding="utf-8")
    look = host_sanitizer.run("cache-lookup", p, "k")
    store = host_sanitizer.run("cache-store", p, "k")
    assert "rc=" in look and "rc=" in store
    if text is None or (text.startswith("{\"format\": \"cellflow-tuning\", \"version\"")):
        assert "rc=0" in store
        assert "rc=0 found=1" in host_sanitizer.run("cache-lookup", p, "k")
