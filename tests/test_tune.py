"""The tile-shape autotuner (neslgpu_amd/tune.py) on the host: its search
space resolves against the HIP sources, and `pick` keeps a macro's current
value unless a candidate whose results matched the oracle's is more than 2%
faster, and never takes a candidate whose results differed."""
import json
import shutil

from neslgpu_amd import tune as T


def test_space_resolves_against_the_sources():
    for key, (values, prim) in T.SPACE.items():
        cur = T.norm(T.key_value(key))
        assert cur in [T.norm(v) for v in values], (key, cur)
        assert len(set(T.norm(v) for v in values)) == len(values), key
        assert prim


def _pick(tmp_path, monkeypatch, key, timings):
    """pick() over a results file in which `key` was searched: `timings`
    maps each value to (ms, correct); the current value is the default."""
    tuned = tmp_path / "tuned.hpp"
    shutil.copy(T.TUNED, tuned)
    monkeypatch.setattr(T, "TUNED", str(tuned))
    monkeypatch.setattr(T, "RECORD", str(tmp_path / "record.md"))
    _, prim = T.SPACE[key]
    cur = T.norm(T.key_value(key))
    results = {}
    for v, (ms, ok) in timings.items():
        name = f"default:{prim}" if T.norm(v) == cur else T.variant_name(key, v)
        results[name] = {"ms": ms, "correct": ok, "primitive": prim}
    path = tmp_path / "results.json"
    path.write_text(json.dumps({"results": results, "reps": 7,
                                "variants": {"default": {"defaults": {key: cur}}}}))
    return T.pick(str(path)), tuned.read_text()


def test_pick_keeps_the_current_value_against_wrong_or_marginal_candidates(tmp_path, monkeypatch):
    key = "CVLHIP_LOOKBACK_K"
    values, _ = T.SPACE[key]
    cur = T.norm(T.key_value(key))
    a, b = [v for v in values if T.norm(v) != cur][:2]
    chosen, text = _pick(tmp_path, monkeypatch, key, {cur: (1.00, True), a: (0.50, False), b: (0.99, True)})
    assert chosen[key] == cur
    assert f"#define {key} {cur}\n" in text


def test_pick_takes_a_correct_candidate_more_than_two_percent_faster(tmp_path, monkeypatch):
    key = "CVLHIP_LOOKBACK_K"
    values, _ = T.SPACE[key]
    cur = T.norm(T.key_value(key))
    a, b = [v for v in values if T.norm(v) != cur][:2]
    chosen, text = _pick(tmp_path, monkeypatch, key, {cur: (1.00, True), a: (0.90, True), b: (0.95, True)})
    assert chosen[key] == a
    assert f"#define {key} {a}\n" in text


def test_pick_writes_every_macro_of_a_joint_key(tmp_path, monkeypatch):
    key = "CVLHIP_RANK_THREADS,CVLHIP_RANK_ITEMS"
    values, _ = T.SPACE[key]
    cur = T.norm(T.key_value(key))
    other = next(T.norm(v) for v in values if T.norm(v) != cur)
    chosen, text = _pick(tmp_path, monkeypatch, key, {cur: (1.00, True), other: (0.80, True)})
    assert chosen[key] == other
    for m, v in zip(T.macros(key), other):
        assert f"#define {m} {v}\n" in text
