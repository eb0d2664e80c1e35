"""The committed fixture of the 128x64x4 configuration's 200-step oracle run
(tests/golden/c1_128x64x4_200steps.json, written by tests/golden/make_c1_golden.py) is
the oracle's: its first steps are run again here and compared exactly.  The GPU test
compares all 200 steps and the final fields with it."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_c1_golden as mk  # noqa: E402
import oracle_py as op  # noqa: E402


def test_c1_fixture_is_the_oracles_run():
    with open(mk.OUT) as f:
        gold = json.load(f)
    assert len(gold["steps"]) == gold["case"]["steps"] == 200
    assert set(gold["final_sha256"]) == {"u0", "v0", "w0", "p", "x"}
    sim = mk.c1_sim(op.Oracle())
    for s in range(4):
        assert mk.report(sim.step()) == gold["steps"][s], s
