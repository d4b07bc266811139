"""Stored formats on the GPU (OperatorBuilder::build with a stored format,
QMBED src/operator/mod.rs:2091-2216) and the model commands that expose them
through the C ABI: materialize_model (the stored entries in row-major order,
bindings/capi/src/lib.rs:3913-3940), apply_model (:4351-4360),
describe_model (:5990-5997) and eigh_model (:5290-5299), checked against the
oracle's assembled triplets, COO product and dense spectra; eigsh with the
csr/csc formats (sparse products over the stored matrix) against the oracle
and the full-size fixture."""

import json
import os

import numpy as np
import pytest

from tests import cases

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def basis_json(L, up=None, k=None, p=None, pauli=False, **extra):
    b = {"kind": "spin", "sites": L, "pauli": pauli}
    if up is not None:
        b["up"] = up
    if k is not None:
        b["momentum"] = k
    if p is not None:
        b["parity"] = p
    b.update(extra)
    return b


def create(qa, basis, spec, checks=None):
    req = {"operation": "create_model", "basis": basis, "terms": qa.terms_from_spec(spec)}
    if checks is not None:
        req["checks"] = checks
    return qa.command(req)["handle"]


def entries_of(res):
    rows = np.array([e["row"] for e in res["entries"]], dtype=np.int64)
    cols = np.array([e["column"] for e in res["entries"]], dtype=np.int64)
    vals = np.array([complex(*e["value"]) for e in res["entries"]])
    return rows, cols, vals


def assert_same_entries(got, want, tol=1e-13):
    """Row-major order, equal values at common positions, and any entry the
    other side lacks at rounding level: entries that vanish exactly cancel
    to a few ulps of noise, which either side may store (|v| > f64::EPSILON)
    or drop."""
    r, c, v = got
    assert np.all(np.diff(r * (int(c.max(initial=0)) + 1) + c) > 0), "entries are not in row-major order"
    g = {(int(a), int(b)): x for a, b, x in zip(r, c, v)}
    w = {(int(a), int(b)): x for a, b, x in zip(*want)}
    for key in set(g) | set(w):
        assert abs(g.get(key, 0.0) - w.get(key, 0.0)) <= tol, key
    for key in set(g) ^ set(w):
        assert abs(g.get(key, w.get(key))) <= tol, key
    # all but rounding-level entries are shared
    big = {k for k, x in w.items() if abs(x) > 1e-12}
    assert big <= set(g)


def cvec(x):
    return [[float(v.real), float(v.imag)] for v in x]


def from_json(v):
    return np.array([complex(a, b) for a, b in v])


STORED = [
    # (name, L, up, k, p, hamiltonian, extra basis keys)
    ("heis_k0", 12, 6, 0, None, cases.heisenberg, {}),
    ("dm_k3", 12, 6, 3, None, cases.dm_chain, {}),
    ("j1j2_k0p-1", 12, 6, 0, -1, cases.j1j2, {}),
    ("xy_full", 8, None, None, None, cases.xy_field, {}),
    ("xy_k2", 9, None, 2, None, cases.xy_field, {}),
    ("heis_rev_k4", 12, 6, 4, None, cases.heisenberg, {"reverse": True}),
    ("ising_pauli", 10, None, 0, 1, lambda L: cases.ising(L, 0.7, 0.2), {"pauli": True}),
]


@pytest.mark.parametrize("fmt", ["csc", "csr", "dense", "dia"])
@pytest.mark.parametrize("name,L,up,k,p,ham,extra", STORED, ids=[c[0] for c in STORED])
def test_materialize_matches_oracle_triplets(qa, oracle, name, L, up, k, p, ham, extra, fmt):
    print(f"[stored] {name} {fmt}", flush=True)
    spec = ham(L)
    pauli = extra.get("pauli", False)
    om = oracle.Model(L, spec, up=up, momentum=k, parity=p, norm="pauli_cartesian" if pauli else "angular",
                      reverse=extra.get("reverse", False))
    h = create(qa, basis_json(L, up, k, p, **extra), spec)
    res = qa.command({"operation": "materialize_model", "handle": h, "format": fmt})
    assert res["kind"] == "operator" and res["format"] == fmt
    assert res["shape"] == [om.dim, om.dim]
    assert_same_entries(entries_of(res), om.triplets())
    # describe_model: the sector's rows
    d = qa.command({"operation": "describe_model", "handle": h})
    assert d["kind"] == "basis" and d["dimension"] == om.dim
    np.testing.assert_array_equal(np.array([int(s) for s in d["states"]], dtype=np.uint64), om.states()[0])
    qa.command({"operation": "release_model", "handle": h})


@pytest.mark.parametrize("name,L,up,k,p,ham,extra", STORED, ids=[c[0] for c in STORED])
def test_apply_model_actions(qa, oracle, name, L, up, k, p, ham, extra):
    spec = ham(L)
    pauli = extra.get("pauli", False)
    om = oracle.Model(L, spec, up=up, momentum=k, parity=p, norm="pauli_cartesian" if pauli else "angular",
                      reverse=extra.get("reverse", False))
    H = om.dense()
    h = create(qa, basis_json(L, up, k, p, **extra), spec)
    xs = [cases.rand_vec(om.dim, seed=s) for s in (1, 2)]
    want = {"normal": [H @ x for x in xs], "adjoint": [H.conj().T @ x for x in xs],
            "transpose": [H.T @ x for x in xs], "conjugate": [H.conj() @ x for x in xs]}
    for fmt in ("matrix_free", "csr"):
        # the model's action follows the format of its last solve
        qa.command({"operation": "eigsh_model", "handle": h, "format": fmt,
                    "solver": {"eigenpairs": 1, "target": {"kind": "smallest_algebraic"}}})
        for action, ys in want.items():
            res = qa.command({"operation": "apply_model", "handle": h, "action": action,
                              "vectors": [cvec(x) for x in xs]})
            assert res["kind"] == "vectors" and res["dimension"] == om.dim
            for got, y in zip(res["vectors"], ys):
                assert np.linalg.norm(from_json(got) - y) <= 1e-13 * max(np.linalg.norm(y), 1.0), (fmt, action)
    qa.command({"operation": "release_model", "handle": h})


def test_non_hermitian_stored_matrix(qa, oracle):
    # hermiticity checks off and a one-sided hopping: the stored rows come
    # from the columns ordered by row (the transposing path)
    L = 10
    spec = [("+-", [(0.7, [i, (i + 1) % L]) for i in range(L)]), ("z", [(0.3, [i]) for i in range(L)])]
    om = oracle.Model(L, spec, up=5, momentum=2, check_hermiticity=False)
    checks = {"hermiticity": False, "particle_conservation": True, "symmetry_compatibility": True}
    h = create(qa, basis_json(L, 5, 2), spec, checks=checks)
    for fmt in ("csr", "csc"):
        assert_same_entries(entries_of(qa.command({"operation": "materialize_model", "handle": h, "format": fmt})),
                            om.triplets())
    x = cases.rand_vec(om.dim, seed=5)
    res = qa.command({"operation": "apply_model", "handle": h, "vectors": [cvec(x)]})
    np.testing.assert_allclose(from_json(res["vectors"][0]), om.dense() @ x, atol=1e-12)
    res = qa.command({"operation": "apply_model", "handle": h, "action": "conjugate", "vectors": [cvec(x)]})
    np.testing.assert_allclose(from_json(res["vectors"][0]), om.dense().conj() @ x, atol=1e-12)
    qa.command({"operation": "release_model", "handle": h})


@pytest.mark.parametrize("vectors", [False, True])
@pytest.mark.parametrize("name,L,up,k,p,ham,extra", STORED, ids=[c[0] for c in STORED])
def test_eigh_model(qa, oracle, name, L, up, k, p, ham, extra, vectors):
    spec = ham(L)
    pauli = extra.get("pauli", False)
    om = oracle.Model(L, spec, up=up, momentum=k, parity=p, norm="pauli_cartesian" if pauli else "angular",
                      reverse=extra.get("reverse", False))
    H = om.dense()
    w = np.linalg.eigvalsh(H)
    h = create(qa, basis_json(L, up, k, p, **extra), spec)
    res = qa.command({"operation": "eigh_model", "handle": h, "eigenvectors": vectors})
    assert res["kind"] == "eigensystem" and res["dimension"] == om.dim
    assert res["iterations"] == 1 and res["converged"]
    np.testing.assert_allclose(res["eigenvalues"], w, rtol=0, atol=1e-11 * max(1.0, np.abs(w).max()))
    assert max(res["residuals"]) <= 1e-10 * max(1.0, np.abs(w).max())
    if vectors:
        assert len(res["eigenvectors"]) == om.dim
        for lam, vec in zip(res["eigenvalues"], res["eigenvectors"]):
            vv = from_json(vec)
            assert np.linalg.norm(H @ vv - lam * vv) <= 1e-10 * max(abs(lam), 1.0)
    else:
        assert "eigenvectors" not in res
    qa.command({"operation": "release_model", "handle": h})


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("fmt", ["csr", "csc", "matrix_free"])
def test_stored_format_eigsh_full_size(qa, fmt, precision):
    # the XXZ L=24 up=12 k=0 sector through the one-shot request with each
    # format: the oracle's six lowest eigenvalues (fixture) to 1e-10
    with open(os.path.join(GOLDEN, "c2.json")) as f:
        d = json.load(f)
    print(f"[stored eigsh] c2 {fmt} {precision}", flush=True)
    terms = [(sym, [(c, s) for c, s in cpl]) for sym, cpl in d["terms"]]
    res = qa.run({"basis": basis_json(24, 12, 0), "terms": qa.terms_from_spec(terms), "format": fmt,
                  "solver": {"eigenpairs": 6, "target": {"kind": "smallest_algebraic"},
                             "matvec_precision": precision}})
    np.testing.assert_allclose(res["eigenvalues"], d["eigenvalues"], rtol=1e-10)


def test_materialize_family_parameters(qa, oracle):
    # a family's stored matrix follows the evaluation's parameters
    L = 10
    h = qa.command({"operation": "create_model", "basis": basis_json(L, None, 0, 1, pauli=True),
                    "terms": qa.terms_from_spec([("zz", [(-1.0, [i, (i + 1) % L]) for i in range(L)])]),
                    "components": [{"name": "hx", "terms": qa.terms_from_spec([("x", [(-1.0, [i]) for i in range(L)])])},
                                   {"name": "hz", "terms": qa.terms_from_spec([("z", [(-1.0, [i]) for i in range(L)])]),
                                    "default": [0.0, 0.0]}]})["handle"]
    for hx, hz in ((0.5, 0.0), (1.3, 0.4)):
        om = oracle.Model(L, cases.ising(L, hx, hz), momentum=0, parity=1, norm="pauli_cartesian")
        res = qa.command({"operation": "materialize_model", "handle": h, "format": "csr",
                          "parameters": {"hx": [hx, 0.0], "hz": [hz, 0.0]}})
        assert_same_entries(entries_of(res), om.triplets())
    qa.command({"operation": "release_model", "handle": h})


@pytest.mark.parametrize("fmt", ["csr", "csc", "dense"])
def test_typed_model_stored_format(qa, oracle, fmt):
    # the typed path's format key: the products of the stored matrix equal
    # the matrix-free ones (FP64 to 1e-13, FP32 to 2e-6) and the oracle's
    L, up, k = 14, 7, 2
    spec = cases.dm_chain(L)
    print(f"[typed stored] {fmt}", flush=True)
    om = oracle.Model(L, spec, up=up, momentum=k)
    free = qa.Model(basis_json(L, up, k), spec)
    stored = qa.Model(basis_json(L, up, k), spec, format=fmt)
    x = cases.rand_vec(om.dim, seed=5)
    y_ref = om.apply(x)
    y = stored.apply(x.astype(np.complex128))
    assert np.linalg.norm(y - y_ref) <= 1e-13 * np.linalg.norm(y_ref)
    assert np.linalg.norm(y - free.apply(x.astype(np.complex128))) <= 1e-13 * np.linalg.norm(y_ref)
    y32 = stored.apply(x.astype(np.complex64))
    assert np.linalg.norm(y32 - y_ref) <= 2e-6 * np.linalg.norm(y_ref)
