"""The Shift target (shift-invert eigsh, QMBED src/solve/mod.rs:950-999,
1285-1456, 2198-2240): the k eigenpairs nearest σ, ordered by |λ − σ|. The
oracle does not restate the shift-invert iteration; what the mathematics
fixes pins it instead: the dense spectrum of the oracle's sector matrix gives
the k eigenvalues nearest σ (≤ 1e-10), and every returned vector is an
eigenvector of that matrix. Both the dense route (N ≤ 256) and the Krylov
route (MINRES inner solves) run, for real and complex sectors, matrix-free
and stored."""

import numpy as np
import pytest

from tests import cases

pytestmark = pytest.mark.gpu


def basis_json(L, up=None, k=None, p=None):
    b = {"kind": "spin", "sites": L}
    if up is not None:
        b["up"] = up
    if k is not None:
        b["momentum"] = k
    if p is not None:
        b["parity"] = p
    return b


def nearest(w, sigma, k):
    idx = sorted(range(len(w)), key=lambda i: (abs(w[i] - sigma), i))[:k]
    return np.array([w[i] for i in idx])


CASES = [
    # (name, L, up, k, p, hamiltonian, which gap the shift sits in, pairs)
    ("heis_dense", 10, 5, 0, None, cases.heisenberg, 3, 3),
    ("heis_k0", 16, 8, 0, None, cases.heisenberg, 5, 3),
    ("xxz_k0p1", 18, 9, 0, 1, lambda L: cases.xxz(L, 1.5), 2, 2),
    ("dm_k3", 16, 8, 3, None, cases.dm_chain, 4, 3),
    ("j1j2_mid", 16, 8, 0, None, cases.j1j2, "mid", 2),
]


@pytest.mark.parametrize("fmt", ["matrix_free", "csr"])
@pytest.mark.parametrize("name,L,up,k,p,ham,gap,pairs", CASES, ids=[c[0] for c in CASES])
def test_shift_target_nearest_eigenpairs(qa, oracle, name, L, up, k, p, ham, gap, pairs, fmt):
    spec = ham(L)
    om = oracle.Model(L, spec, up=up, momentum=k, parity=p)
    H = om.dense()
    w = np.linalg.eigvalsh(H)
    # a shift between two distinct eigenvalues (the chains' spectra hold
    # degenerate pairs; a shift onto an eigenvalue leaves (A − σI) singular)
    i = len(w) // 2 if gap == "mid" else gap
    while w[i + 1] - w[i] < 1e-3:
        i += 1
    sigma = 0.5 * (w[i] + w[i + 1]) + 0.1 * (w[i + 1] - w[i])
    print(f"[shift] {name} {fmt} N={om.dim} sigma={sigma}", flush=True)
    # a shift inside the bulk leaves (A − σI) strongly indefinite: its inner
    # solves need more products than the default budget of 1000
    solver = {"eigenpairs": pairs, "target": {"kind": "shift", "value": sigma}, "eigenvectors": True}
    if gap == "mid":
        solver["max_iterations"] = 4000
    res = qa.run({"basis": basis_json(L, up, k, p), "terms": qa.terms_from_spec(spec), "format": fmt,
                  "solver": solver})
    assert res["converged"]
    np.testing.assert_allclose(res["eigenvalues"], nearest(w, sigma, pairs), rtol=1e-10, atol=1e-12)
    d = [abs(x - sigma) for x in res["eigenvalues"]]
    assert d == sorted(d)
    for lam, vec in zip(res["eigenvalues"], res["eigenvectors"]):
        v = np.array([complex(a, b) for a, b in vec])
        assert abs(np.linalg.norm(v) - 1) < 1e-12
        assert np.linalg.norm(H @ v - lam * v) <= 1e-9 * max(abs(lam), 1.0)


def test_shift_target_errors(qa):
    spec = qa.terms_from_spec(cases.heisenberg(16))
    b = basis_json(16, 8, 0)
    with pytest.raises(qa.QmbedError, match="shift must be finite|invalid"):
        qa.run({"basis": b, "terms": spec, "solver": {"eigenpairs": 1, "target": {"kind": "shift", "value": float("nan")}}})
    with pytest.raises(qa.QmbedError, match="thick restart"):
        qa.run({"basis": b, "terms": spec,
                "solver": {"eigenpairs": 1, "target": {"kind": "shift", "value": -5.0}, "krylov_dimension": 40}})
