"""The reference's Python tests on the spin eigsh path, replayed through this
build's C ABI: tests/golden/python_replay.json holds the requests that the
reference's Python package sends for bindings/python/tests/test_qmbed.py:
57-150 (native eigsh, the QuSpin static adapter, a hamiltonian model's
lifecycle, a QuSpin site map against the optimized momentum sector) and what
those tests assert, plus the oracle's values for the same requests
(tests/golden/make_python_replay.py writes it). A later step may name an
earlier step's handle as "$name"."""

import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "python_replay.json")


def load_cases():
    with open(FIXTURE) as f:
        return json.load(f)["cases"]


def substitute(value, saved):
    if isinstance(value, str) and value.startswith("$"):
        return saved[value[1:]]
    if isinstance(value, dict):
        return {k: substitute(v, saved) for k, v in value.items()}
    if isinstance(value, list):
        return [substitute(v, saved) for v in value]
    return value


def check(result, expect):
    for key, want in expect.items():
        if key == "eigenvalues":
            got = result["eigenvalues"][:len(want["values"])]
            np.testing.assert_allclose(got, want["values"], rtol=want.get("rtol", 0), atol=want.get("atol", 0))
        elif key == "eigenvalues_oracle":
            np.testing.assert_allclose(result["eigenvalues"], want["values"], rtol=want["rtol"], atol=1e-12)
        elif key == "count":
            assert len(result["eigenvalues"]) == want
        elif key == "dense":
            n = result["shape"][0]
            got = np.zeros((n, n), dtype=complex)
            for e in result["entries"]:
                got[e["row"], e["column"]] = complex(*e["value"])
            ref = np.zeros((n, n), dtype=complex)
            for i, j, v in want["entries"]:
                ref[i, j] = complex(*v)
            np.testing.assert_allclose(got, ref, rtol=0, atol=want["atol"])
        else:
            assert result[key] == want, (key, result.get(key), want)


@pytest.mark.parametrize("case", load_cases(), ids=lambda c: c["name"])
def test_replay(qa, case):
    print(f"[replay] {case['name']} ({case['source']})", flush=True)
    saved = {}
    for step in case["steps"]:
        request = substitute(step["request"], saved)
        call = qa.run if step["call"] == "run" else qa.command
        if "expect_error" in step:
            with pytest.raises(qa.QmbedError, match=step["expect_error"]):
                call(request)
            continue
        result = call(request)
        if "save" in step:
            saved[step["save"]] = result["handle"]
        check(result, step.get("expect", {}))
