"""Full-size parity against oracle fixtures (tests/golden/*.json, written by
tests/golden/make_golden.py): the benchmark sectors are too large to run the
oracle here, so the fixtures hold size-independent fingerprints — dimension,
exact stored-entry count, hashes of the representative list and orbit sizes,
and bilinear forms w·(Hx) for seeded x, w — plus oracle eigenvalues.
Both H·v modes (orbit walk + sorted-table search, image table) are checked."""

import hashlib
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ["c1", "c2", "c2k5", "c4k0", "c3", "c5", "c5a", "c5b"]


def load(name):
    with open(os.path.join(HERE, f"{name}.json")) as f:
        return json.load(f)


def basis_of(d):
    b = {"kind": "spin", "sites": d["sites"], "pauli": d["normalization"] == "pauli_cartesian"}
    for key in ("up", "momentum", "parity"):
        if d[key] is not None:
            b[key] = d[key]
    return b


def terms_of(d):
    out = []
    for sym, couplings in d["terms"]:
        out.append((sym, [(complex(*c) if isinstance(c, list) else c, s) for c, s in couplings]))
    return out


@pytest.mark.parametrize("mode", ["walk", "image", "stream"])
@pytest.mark.parametrize("name", CASES)
def test_fingerprint(qa, monkeypatch, name, mode):
    d = load(name)
    monkeypatch.setenv("QMBED_HV_MODE", mode)
    print(f"[golden] {name} mode={mode}", flush=True)
    m = qa.Model(basis_of(d), terms_of(d))
    assert m.dim == d["dimension"]
    if mode != "walk":
        assert m.hv_config()["mode"] == mode
    s, o = m.states()
    assert hashlib.sha256(s.astype("<u4").tobytes()).hexdigest() == d["states_sha256"]
    assert hashlib.sha256(o.astype("u1").tobytes()).hexdigest() == d["orbits_sha256"]
    assert m.nnz() == d["nnz"]
    rng = np.random.default_rng(d["x_seed"])
    x = rng.standard_normal(m.dim)
    if d["complex_vectors"]:
        x = x + 1j * rng.standard_normal(m.dim)
    y = m.apply(x)
    assert abs(np.linalg.norm(y) - d["hx_norm"]) <= 1e-12 * d["hx_norm"]
    for seed, (fr, fi) in zip(d["w_seeds"], d["forms"]):
        rw = np.random.default_rng(seed)
        w = rw.standard_normal(m.dim)
        if d["complex_vectors"]:
            w = w + 1j * rw.standard_normal(m.dim)
        f = complex(np.dot(w, y))
        scale = np.linalg.norm(w) * d["hx_norm"]
        assert abs(f - complex(fr, fi)) <= 1e-12 * scale, (f, fr, fi)
    if not d["complex_vectors"]:
        y32 = m.apply(x.astype(np.float32))
        assert np.linalg.norm(y32 - y) <= 2e-6 * np.linalg.norm(y)
    m.close()


@pytest.mark.parametrize("name", ["c1", "c2", "c2k5", "c4k0", "c5", "c5a", "c5b"])
def test_eigenvalues(qa, name):
    d = load(name)
    m = qa.Model(basis_of(d), terms_of(d))
    k = len(d["eigenvalues"])
    for prec in ("fp64", "fp32"):
        r = m.eigsh(k, matvec_precision=prec)
        np.testing.assert_allclose(r["eigenvalues"], d["eigenvalues"], rtol=1e-10, atol=0)
        for lam, res in zip(r["eigenvalues"], r["residuals"]):
            assert res <= 1e-10 * max(abs(lam), 1.0)


def test_majumdar_ghosh_ground_state(qa):
    # J2 = J1/2 (the J1-J2 benchmark sector): exact dimer ground state, E0 = -3L/8
    d = load("c3")
    m = qa.Model(basis_of(d), terms_of(d))
    for prec in ("fp64", "fp32"):
        r = m.eigsh(1, matvec_precision=prec)
        assert abs(r["eigenvalues"][0] + 3 * 28 / 8) <= 1e-10 * 10.5
        assert r["residuals"][0] <= 1e-10 * 10.5


def test_batch_sectors_against_oracle_eigenvalues(qa):
    # two of the 72 XXZ L=24 sectors of the batch, pinned by the oracle's own
    # eigsh (tests/golden/c2k5.json: up=11 k=5, complex; c4k0.json: up=12
    # k=0): the six lowest eigenvalues from the full 72-sector lockstep batch
    L = 24
    keys = [(up, k) for up in (11, 12, 13) for k in range(L)]
    spec = terms_of(load("c4k0"))
    models = [qa.Model({"kind": "spin", "sites": L, "up": up, "momentum": k}, spec) for up, k in keys]
    out = qa.eigsh_batch(models, 6)
    for name, key in (("c2k5", (11, 5)), ("c4k0", (12, 0))):
        d = load(name)
        assert terms_of(d) == spec
        r = out["results"][keys.index(key)]
        np.testing.assert_allclose(r["eigenvalues"], d["eigenvalues"], rtol=1e-10, atol=0, err_msg=name)
        for lam, res in zip(r["eigenvalues"], r["residuals"]):
            assert res <= 1e-10 * max(abs(lam), 1.0)
    for m in models:
        m.close()
