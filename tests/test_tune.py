"""The autotuner's table (ptilopsis_amd/tuning.json, when present) holds only
values the library accepts, for every knob and bucket."""
import json
import os

from ptilopsis_amd import tune


def test_tuning_table_values():
    assert set(tune.KNOBS) == {"block.preprocess", "block.isn_cnt", "block.isn_gen", "block.optimize",
                               "block.emit", "block.fix_jumps", "walk_ring", "small_programs"}
    if not os.path.exists(tune.TUNING_PATH):
        return
    t = json.load(open(tune.TUNING_PATH))
    assert t["buckets"] == list(tune.BUCKETS)
    for knob, vals in t["values"].items():
        assert knob in tune.KNOBS
        assert len(vals) == len(tune.BUCKETS)
        assert all(v in tune.KNOBS[knob][0] for v in vals), (knob, vals)
