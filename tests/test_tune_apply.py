"""The autotuner's table (host logic, no GPU): apply(side) exports the tuned
launch settings of a side, leaves variables the caller has set alone, and
does nothing for an untuned side or a missing table; the committed table
holds a choice for every tuned side."""
import json
import os

from mantle_convection_lab_amd import tune


def test_apply_respects_caller_and_missing_entries(tmp_path, monkeypatch):
    path = tmp_path / "t.json"
    path.write_text(json.dumps({"best": {"256": {"MCL_MIXED_BLOCKS": 256, "MCL_MG_FUSE": 2, "MCL_MG_SWEEPS": 2}}}))
    for k in tune.KNOBS:
        monkeypatch.delenv(k, raising=False)
    monkeypatch.setenv("MCL_MG_SWEEPS", "1")
    done = tune.apply(256, str(path))
    try:
        assert done == {"MCL_MIXED_BLOCKS": 256, "MCL_MG_FUSE": 2}
        assert os.environ["MCL_MIXED_BLOCKS"] == "256" and os.environ["MCL_MG_SWEEPS"] == "1"
        assert tune.apply(64, str(path)) == {}
        assert tune.apply(256, str(tmp_path / "missing.json")) == {}
    finally:
        for k in done:
            os.environ.pop(k, None)


def test_committed_table():
    with open(tune.DEFAULT_PATH) as f:
        tab = json.load(f)
    assert set(tab["knobs"]) == set(tune.KNOBS)
    for side, best in tab["best"].items():
        rows = tab["table"][side]
        assert len(rows) == len(tune.CANDIDATES)
        fastest = min(rows, key=lambda r: r["ms"])
        assert best == {k: fastest[k] for k in tune.KNOBS}
