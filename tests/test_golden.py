"""The oracle's outputs for fixed scenes are frozen in tests/golden."""
import json
import os

import pytest

from golden.make_golden import digest, scenes

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "oracle_hashes.json")))


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_oracle_matches_golden(name):
    scene, size, bg = scenes()[name]
    assert digest(scene, size, bg) == GOLDEN[name]
