"""GPU test of the autotuner (trident_amd/autotune.py): on a small R-MAT
product every candidate tuning gives output identical to the first run, the
chosen tuning is one the library accepts, and the JSON round-trips through
device.load_tuning."""
import json

import pytest

from trident_amd import autotune
from trident_amd.device import TUNING_FIELDS, load_tuning

pytestmark = pytest.mark.gpu


def test_autotune_small_product(tmp_path):
    doc = autotune.tune("rmat:12", reps=1, log=lambda *a: None)
    assert doc["candidates"], "no candidate was measured"
    assert all(c["identical_output"] for c in doc["candidates"])
    assert set(doc["tuning"]) == set(TUNING_FIELDS)
    assert doc["best_ms"] <= doc["default_ms"]
    p = tmp_path / "t.json"
    p.write_text(json.dumps(doc))
    assert load_tuning(str(p)) == doc["tuning"]
