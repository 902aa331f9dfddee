"""The autotuner (voronoids_amd/autotune.py): its build constants are the kernels' own macros with
the kernels' defaults, its runtime knobs are vo_tri_config keys, the stored winner reaches the
compiler (build.py) and DelaunayTree, and a trial times a valid triangulation on the GPU."""
import json
import os
import re

import pytest

from voronoids_amd import autotune as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KERNELS = os.path.join(ROOT, "voronoids_amd", "csrc", "kernels_group.hpp")
HEADER = os.path.join(ROOT, "include", "voronoids.h")


def test_build_constants_are_the_kernels_macros_with_their_defaults():
    src = open(KERNELS).read()
    for dim, defaults in A.BUILD_DEFAULTS.items():
        for name, value in defaults.items():
            m = re.search(rf"#ifndef {name}\n#define {name} (\d+)", src)
            assert m, name
            assert int(m.group(1)) == value, name
    for dim, cands in A.BUILD_CANDIDATES.items():
        assert cands[0] == {}
        for c in cands:
            assert set(c) <= set(A.BUILD_DEFAULTS[dim]), c
            for k, v in c.items():
                if "_KC" in k or "_FC" in k:
                    assert v % 16 == 0, (k, v)    # int4 list scans and aligned LDS rows
    # the two dimensions' constants are disjoint, so one build carries both winners
    assert not set(A.BUILD_DEFAULTS[2]) & set(A.BUILD_DEFAULTS[3])


def test_runtime_knobs_are_config_keys():
    hdr = open(HEADER).read()
    for dim in (2, 3):
        for s in A.runtime_candidates(dim):
            for k in s:
                assert f'"{k}"' in hdr, k


def test_stored_winner_reaches_the_compiler(tmp_path, monkeypatch):
    from voronoids_amd import build as B
    t = tmp_path / "tuning.json"
    t.write_text(json.dumps({"3": {"build": {"VO_CAVITY_WAVES": 4}, "setting": {"lanes_target": 1 << 20}},
                             "2": {"build": {"VO_COMMIT_WAVES2": 3}, "setting": {}}}))
    monkeypatch.setattr(A, "TUNING", str(t))
    assert A.tuned_defines() == ("VO_CAVITY_WAVES=4", "VO_COMMIT_WAVES2=3")
    cmds = []
    monkeypatch.setattr(B, "_run", lambda cmd, verbose: cmds.append(cmd))
    monkeypatch.setattr(B.os, "replace", lambda a, b: None)
    monkeypatch.setattr(B.os, "remove", lambda p: None)
    B.build(out=str(tmp_path / "lib.so"), defines=("VO_COMMIT_WAVES2=5",), static=False)
    compile_cmds = [c for c in cmds if "-c" in c]
    assert compile_cmds
    for c in compile_cmds:
        assert "-DVO_CAVITY_WAVES=4" in c and "-DVO_COMMIT_WAVES2=5" in c and "-DVO_COMMIT_WAVES2=3" not in c
    cmds.clear()
    B.build(out=str(tmp_path / "lib.so"), static=False, tuned=False)
    assert not any(a.startswith("-DVO_") for c in cmds for a in c)
    t.write_text(json.dumps({"3": {"build": {"VO_NOT_A_KNOB": 1}}}))
    with pytest.raises(ValueError):
        A.tuned_defines()


def test_shipped_table_is_consistent():
    if not os.path.exists(A.TUNING):
        pytest.skip("no tuning.json shipped")
    table = json.load(open(A.TUNING))
    hdr = open(HEADER).read()
    for d, rec in table.items():
        assert set(rec["build"]) <= set(A.BUILD_DEFAULTS[int(d)])
        for k in rec["setting"]:
            assert f'"{k}"' in hdr
        assert rec["ms"] <= rec["default_ms"]
        chosen = A.pick(rec["candidates"], rec["margin"])
        assert (chosen["build"], chosen["setting"], chosen["ms"]) == (rec["build"], rec["setting"], rec["ms"])


def test_pick_keeps_the_default_within_the_margin():
    c = [{"build": {}, "setting": {}, "ms": 100.0, "valid": True},
         {"build": {"VO_COMMIT_WAVES": 3}, "setting": {}, "ms": 99.5, "valid": True},
         {"build": {}, "setting": {"chunk_rounds": 2}, "ms": 90.0, "valid": False}]
    assert A.pick(c, 0.01) is c[0]
    c[1]["ms"] = 98.0
    assert A.pick(c, 0.01) is c[1]
    c[1]["rejected"] = "slower on clustered points"
    assert A.pick(c, 0.01) is c[0]


@pytest.mark.gpu
@pytest.mark.parametrize("dim", [3, 2])
def test_trial_times_a_valid_triangulation(dim):
    recs = A.trial(dim, 20000, 1, [{}, A.runtime_candidates(dim)[1]])
    assert len(recs) == 2
    for r in recs:
        assert r["error"] is None and r["ms"] > 0 and r["violations"] == 0
    assert recs[0]["simplices"] == recs[1]["simplices"]
