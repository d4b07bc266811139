"""Thick-restarted Lanczos, initial subspaces and the eigsh workspace on the
GPU (QMBED src/solve/mod.rs:1644-2196, :2385-2444, :708-727), against the
oracle's restatement and dense spectra; the reference's control-flow anchors
and the error cases of its tests, through the C ABI."""

import numpy as np
import pytest

from tests import cases
from tests.test_oracle import TR_CASES

pytestmark = pytest.mark.gpu


def _model(qa, L, spec, up=None, momentum=None, parity=None):
    return qa.Model(qa.spin_basis(L, up=up, momentum=momentum, parity=parity), spec)


def _vecs(r):
    return np.array([[complex(a, b) for a, b in v] for v in r["eigenvectors"]])


@pytest.mark.parametrize("case", range(len(TR_CASES)))
def test_thick_restart_matches_oracle(qa, oracle, case):
    L, spec, up, k_mom, k, kd, target = TR_CASES[case]
    print(f"[thick] case {case}: L={L} up={up} k={k_mom} eigenpairs={k} window={kd} {target}", flush=True)
    m = _model(qa, L, spec, up=up, momentum=k_mom)
    om = oracle.Model(L, spec, up=up, momentum=k_mom)
    assert m.dim == om.dim
    r = m.eigsh(k, target=target, krylov_dimension=kd, tolerance=1e-10, eigenvectors=True)
    ro = om.eigsh(k, target=target, krylov_dimension=kd, tolerance=1e-10)
    np.testing.assert_allclose(r["eigenvalues"], ro["eigenvalues"], rtol=1e-10, atol=1e-12)
    assert r["converged"] and r["iterations"] > kd  # restarted, within the budget
    assert r["iterations"] <= 1000
    H = om.dense()
    for v, lam, res in zip(_vecs(r), r["eigenvalues"], r["residuals"]):
        assert abs(np.linalg.norm(v) - 1) < 1e-12
        true_res = np.linalg.norm(H @ v - lam * v)
        assert true_res <= 1e-10 * max(abs(lam), 1) * 1.01 + 1e-14
        assert abs(true_res - res) <= 1e-9 * max(abs(lam), 1)
    m.close()


def test_thick_restart_diagonal_workspace(qa, oracle):
    # tests/solver_extensions.rs:100-123 on the 512-value diagonal i/511:
    # a first solve restarts; a workspace solve from its converged subspace
    # gives the same values at a fraction of the products. From the same
    # workspace vectors the GPU solve and the oracle's take the same
    # iterations to the same residuals (which vectors a first solve ends on
    # depends on rounding; from the oracle's, one window suffices, as the
    # reference's test has it)
    L = 9
    spec = cases.diagonal_chain(L, 1.0 / 511)
    m = _model(qa, L, spec)
    om = oracle.Model(L, spec)
    ws = qa.Workspace()
    first = m.eigsh_workspace(ws, 3, krylov_dimension=16, tolerance=1e-10, seed=19, eigenvectors=True)
    assert first["iterations"] > 16 and len(ws) == 3 and ws.dimension == 512
    np.testing.assert_allclose(first["eigenvalues"], [0, 1 / 511, 2 / 511], atol=1e-10)
    o1 = om.eigsh_subspace(3, [], krylov_dimension=16, tolerance=1e-10, seed=19, workspace=True, vectors=True)
    np.testing.assert_allclose(first["eigenvalues"], o1["eigenvalues"], atol=1e-12)
    assert o1["iterations"] > 16
    second = m.eigsh_workspace(ws, 3, krylov_dimension=16, tolerance=1e-10, seed=19)
    assert 10 * second["iterations"] < first["iterations"]
    np.testing.assert_allclose(second["eigenvalues"], first["eigenvalues"], atol=1e-14)
    assert max(second["residuals"]) <= 1e-10
    o2 = om.eigsh_subspace(3, _vecs(first), krylov_dimension=16, tolerance=1e-10, seed=19, workspace=True)
    assert second["iterations"] == o2["iterations"]
    np.testing.assert_allclose(second["residuals"], o2["residuals"], rtol=1e-6, atol=1e-15)
    ws.set(512, o1["eigenvectors"])
    g3 = m.eigsh_workspace(ws, 3, krylov_dimension=16, tolerance=1e-10, seed=19)
    o3 = om.eigsh_subspace(3, o1["eigenvectors"], krylov_dimension=16, tolerance=1e-10, seed=19, workspace=True)
    assert g3["iterations"] == o3["iterations"] <= 16
    np.testing.assert_allclose(g3["residuals"], o3["residuals"], rtol=1e-6, atol=1e-15)
    np.testing.assert_allclose(g3["eigenvalues"], o3["eigenvalues"], atol=1e-14)
    # a window that cannot converge answers NonConvergence
    # (tests/solver_extensions.rs:73-87)
    r = qa.raw_call("qmbed_run_json", {
        "basis": qa.spin_basis(L), "terms": qa.terms_from_spec(spec),
        "solver": {"eigenpairs": 1, "target": {"kind": "smallest_algebraic"}, "krylov_dimension": 2,
                   "tolerance": 1e-12, "max_iterations": 2, "seed": 7}})
    assert r["status"] == "error" and r["error"].startswith("solver did not converge after 2 iterations")
    ws.close()
    m.close()


def test_large_workspace_window_keeps_the_tridiagonal_path(qa, oracle):
    # tests/solver_extensions.rs:126-154: four isolated low values, window 100
    L = 9
    spec = cases.isolated_low_chain(L)
    m = _model(qa, L, spec)
    ws = qa.Workspace()
    kw = dict(krylov_dimension=100, tolerance=1e-10, max_iterations=100, seed=23)
    first = m.eigsh_workspace(ws, 3, **kw)
    second = m.eigsh_workspace(ws, 3, **kw)
    assert len(ws) == 3
    vals = np.sort(np.diag(oracle.Model(L, spec).dense()).real)
    np.testing.assert_allclose(second["eigenvalues"], vals[:3], atol=1e-10)
    np.testing.assert_allclose(first["eigenvalues"], second["eigenvalues"], atol=1e-12)
    assert max(second["residuals"]) <= 1e-10
    m.close()


def test_initial_subspace_route(qa, oracle):
    L = 16
    m = _model(qa, L, cases.heisenberg(L), up=8, momentum=0)
    om = oracle.Model(L, cases.heisenberg(L), up=8, momentum=0)
    w, V = np.linalg.eigh(om.dense())
    sub = [[[z.real, z.imag] for z in V[:, j]] for j in range(3)]
    r = m.eigsh(2, tolerance=1e-10, initial_subspace=sub)
    np.testing.assert_allclose(r["eigenvalues"], w[:2], rtol=1e-12)
    assert r["iterations"] <= 2 * (4 * 2 + 24)
    ro = om.eigsh_subspace(2, V[:, :3].T, tolerance=1e-10)
    np.testing.assert_allclose(r["eigenvalues"], ro["eigenvalues"], rtol=1e-12)
    # a random 4-vector subspace converges to the same pairs
    rng = np.random.default_rng(3)
    R = rng.standard_normal((4, m.dim))
    r4 = m.eigsh(2, tolerance=1e-10, initial_subspace=[[[x, 0.0] for x in v] for v in R])
    np.testing.assert_allclose(r4["eigenvalues"], w[:2], rtol=1e-10)
    with pytest.raises(qa.QmbedError, match="dimension mismatch: eigsh initial vector does not match"):
        m.eigsh(2, initial_subspace=[[[1.0, 0.0]] * 3])
    m.close()


def test_full_window_iteration_count(qa):
    # tests/visible_contract.rs:1041-1059: diagonal 0..255, window 160,
    # without early stopping the unrestarted window spends 160 products
    L = 8
    m = _model(qa, L, cases.diagonal_chain(L))
    r = m.eigsh(3, krylov_dimension=160, tolerance=1e-8, max_iterations=192, seed=17, early_stop=False)
    np.testing.assert_allclose(r["eigenvalues"], [0, 1, 2], atol=1e-7)
    assert max(r["residuals"]) < 1e-7
    assert r["iterations"] == 160
    m.close()


def test_reference_error_cases_through_the_c_abi(qa):
    # IncompatibleSymmetry: momentum 1 with parity (tests/full_api_foundation.rs:191-197)
    r = qa.raw_call("qmbed_command_json", {"operation": "describe_basis",
                                           "basis": qa.spin_basis(4, up=2, momentum=1, parity=1)})
    assert r["status"] == "error" and r["error"].startswith("incompatible symmetry request")
    # initial vectors (tests/full_api_foundation.rs:308-333): wrong length,
    # zero, non-finite
    L = 9
    req = {"basis": qa.spin_basis(L, up=4), "terms": qa.terms_from_spec(cases.heisenberg(L, nn=1)),
           "solver": {"eigenpairs": 1, "target": {"kind": "smallest_algebraic"}}}
    n = qa.run({**req, "solver": dict(req["solver"])})["dimension"]
    bad = dict(req, solver=dict(req["solver"], initial_vector=[[1.0, 0.0]] * 2))
    assert qa.raw_call("qmbed_run_json", bad)["error"] == \
        "dimension mismatch: eigsh initial vector does not match the operator"
    bad = dict(req, solver=dict(req["solver"], initial_vector=[[0.0, 0.0]] * n))
    assert qa.raw_call("qmbed_run_json", bad)["error"] == \
        "invalid options: eigsh initial vector must have nonzero finite norm"
    ws = qa.Workspace()
    with pytest.raises(qa.QmbedError, match="invalid options: eigsh initial vector must contain only finite values"):
        ws.set(n, [np.full(n, np.nan)])
    with pytest.raises(qa.QmbedError, match="invalid options: eigsh initial vector must have nonzero finite norm"):
        ws.set(n, [np.zeros(n)])
    ws.close()
