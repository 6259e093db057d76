"""The autotuner's committed table: well formed, every knob known to the
library's vxhip_tuning, and looked up by the scene's workload name."""
import json

from velox_amd import tuning
from velox_amd.device import Tuning

KNOWN = {name for name, _ in Tuning._fields_}


def test_table_is_well_formed():
    with open(tuning.TABLE) as f:
        table = json.load(f)
    assert table["workloads"], "empty tuning table"
    for name, entry in table["workloads"].items():
        assert set(entry["knobs"]) <= KNOWN, (name, set(entry["knobs"]) - KNOWN)
        assert entry["knobs"]["batchThreads"] in (256, 512)
        assert entry["ms_per_step"] > 0 and entry["default_ms_per_step"] > 0
        assert entry["trials"] and all(set(t["knobs"]) <= KNOWN for t in entry["trials"])
        # The chosen setting is the fastest the tuner saw (or within its 1% rule of it).
        best = min(t["ms_per_step"] for t in entry["trials"])
        chosen = [t["ms_per_step"] for t in entry["trials"] if t["knobs"] == entry["knobs"]]
        assert chosen and min(chosen) <= best * 1.02


def test_knobs_follow_the_scene_workload(tmp_path):
    path = tmp_path / "t.json"
    path.write_text(json.dumps({"workloads": {"debris": {"knobs": {"batchThreads": 512}},
                                              "default": {"knobs": {"islandLimit": 2048}}}}))

    class S:
        name = "debris"

    class T:
        name = "terrain"
    assert tuning.knobs_for(S(), str(path)) == {"batchThreads": 512}
    assert tuning.knobs_for(T(), str(path)) == {"islandLimit": 2048}
    assert tuning.knobs_for(S(), str(tmp_path / "missing.json")) == {}


def test_tuning_struct_round_trip():
    t = Tuning.from_dict({"batchThreads": 512, "graphSubsteps": True, "sideStream": False})
    assert (t.batchThreads, t.graphSubsteps, t.sideStream, t.persistentBatches) == (512, 1, 2, 0)
    d = t.as_dict()
    assert d["graphSubsteps"] is True and d["sideStream"] is False
