"""The committed trajectory digests are the oracle's: the first steps of the 64^3 dam break and
the first step of the 128^3 one, recomputed here by the oracle, give the fixture's digests."""
import json

import pytest

from tests.golden.make_trajectories import CONFIGS, fixture_path, oracle_steps


@pytest.mark.parametrize("name,upto", [("dam64", 3), ("dam128", 1)])
def test_trajectory_fixture_is_the_oracles(orc, name, upto):
    fx = json.load(open(fixture_path(name)))
    res, record = CONFIGS[name]
    assert fx["scene"]["resolution"] == res and sorted(int(s) for s in fx["steps"]) == record
    for s, k, d in oracle_steps(name, upto):
        assert dict(d, substeps=k) == fx["steps"][str(s)], f"{name} step {s}"
