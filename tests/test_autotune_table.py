"""The tuning table on the host: keys, defaults, and the first RK stage's own tiling."""
import json

from trapbec_amd import autotune


def test_lookup_defaults_and_recorded_entries(tmp_path):
    table = str(tmp_path / "t.json")
    # nothing recorded: the default tile with an automatic segment, contiguous order, and
    # no tiling of its own for the first stage
    assert autotune.lookup((64, 64, 64), 32, table=table) == autotune.DEFAULT[32] + (1,)
    assert autotune.lookup_first_stage((64, 64, 64), 32, table=table) == (-1, 0)
    key = autotune._key((64, 48, 32), 32, 2, False, False)
    rot = autotune._key((64, 48, 32), 32, 2, False, True)
    with open(table, "w") as f:
        json.dump({key: {"tile": 2, "seg": 128, "remap": 2,
                         "first_stage": {"tile": 1, "seg": 256, "us_per_stage": 1.0}},
                   rot: {"tile": 4, "seg": 64, "remap": 0}}, f)
    assert autotune.lookup((64, 48, 32), 32, 2, table=table) == (2, 128, 2)
    assert autotune.lookup_first_stage((64, 48, 32), 32, 2, table=table) == (1, 256)
    # the rotating entry has its own tiling and none for the first stage
    assert autotune.lookup((64, 48, 32), 32, 2, rot=True, table=table) == (4, 64, 0)
    assert autotune.lookup_first_stage((64, 48, 32), 32, 2, rot=True, table=table) == (-1, 0)
    # FP64 defaults to the reference summation in its key
    assert autotune._key((8, 8, 8), 64) == "8x8x8:f64:s0"
    assert autotune._key((8, 8, 8), 32, w_stored=True) == "8x8x8:f32:s2:ws"


def test_recorded_headline_tuning_is_well_formed():
    """The committed table's 512^3 FP32 headline entry names a valid tile, segment and
    order, and its first-stage record is the fastest of its own sweep."""
    t = json.load(open(autotune.TABLE))
    e = t[autotune._key((512, 512, 512), 32, 2)]
    assert 0 <= e["tile"] < 5 and e["seg"] in autotune.SEGS and e["remap"] in (0, 1, 2)
    fs = e.get("first_stage")
    if fs:
        assert [fs["tile"], fs["seg"]] == fs["all"][0][:2]
        assert fs["us_per_stage"] <= fs["us_per_stage_same_tiling"]
