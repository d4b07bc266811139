"""The committed replay fixture (tests/golden/python_replay.json) is what
its generator writes from the oracle today: the requests, the reference
tests' assertions and the oracle's values (CPU only)."""

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def test_replay_fixture_matches_generator(oracle):
    sys.path.insert(0, os.path.join(HERE, "golden"))
    try:
        import make_python_replay as gen
    finally:
        sys.path.pop(0)
    with open(os.path.join(HERE, "golden", "python_replay.json")) as f:
        committed = json.load(f)["cases"]
    fresh = json.loads(json.dumps(gen.build_cases()))
    assert [c["name"] for c in committed] == [c["name"] for c in fresh]
    for a, b in zip(committed, fresh):
        assert len(a["steps"]) == len(b["steps"])
        for sa, sb in zip(a["steps"], b["steps"]):
            assert sa["request"] == sb["request"]
            ea, eb = sa.get("expect", {}), sb.get("expect", {})
            assert set(ea) == set(eb)
            for key in ea:
                if isinstance(ea[key], dict) and "values" in ea[key]:
                    np.testing.assert_allclose(ea[key]["values"], eb[key]["values"], rtol=1e-12, atol=1e-14)
                elif isinstance(ea[key], dict) and "entries" in ea[key]:
                    assert [e[:2] for e in ea[key]["entries"]] == [e[:2] for e in eb[key]["entries"]]
                    np.testing.assert_allclose([e[2] for e in ea[key]["entries"]], [e[2] for e in eb[key]["entries"]],
                                               atol=1e-14)
                else:
                    assert ea[key] == eb[key]


def test_replay_dimer_anchor(oracle):
    # the reference test's anchor: the spin-1/2 dimer's singlet at -3/4 and
    # triplet at 1/4 (bindings/python/tests/test_qmbed.py:70)
    with open(os.path.join(HERE, "golden", "python_replay.json")) as f:
        cases = {c["name"]: c for c in json.load(f)["cases"]}
    ev = cases["native_eigsh_dimer"]["steps"][0]["expect"]["eigenvalues_oracle"]["values"]
    np.testing.assert_allclose(ev, [-0.75, 0.25], atol=1e-13)
