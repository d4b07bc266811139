"""Operator families H(c) = H_static + Σ_p c_p H_p (the reference's
create_model components, src/operator/mod.rs:2710-2830, src/interop.rs:941-1326):
re-evaluating the coefficients must give the operator built directly with the
scaled terms, on the same sector tables."""

import numpy as np
import pytest

from tests import cases

pytestmark = pytest.mark.gpu

L = 12


def ising_family():
    bonds = cases.bonds(L)
    static = [("zz", [(-1.0, b) for b in bonds])]
    comps = [("hx", [("x", [(-1.0, [i]) for i in range(L)])], None),
             ("hz", [("z", [(-1.0, [i]) for i in range(L)])], 0.0)]
    return static, comps


def basis(k=0, p=1, pauli=True):
    b = {"kind": "spin", "sites": L, "pauli": pauli, "momentum": k}
    if p is not None:
        b["parity"] = p
    return b


@pytest.mark.parametrize("hx,hz", [(0.5, 0.1), (1.3, 0.0), (2.0, 0.5)])
def test_family_matches_direct_build(qa, oracle, hx, hz):
    static, comps = ising_family()
    fam = qa.Model(basis(), static, components=comps)
    fam.set_parameters(hx=hx, hz=hz)
    direct = qa.Model(basis(), cases.ising(L, hx, hz))
    om = oracle.Model(L, cases.ising(L, hx, hz), momentum=0, parity=1, norm="pauli_cartesian")
    assert fam.dim == direct.dim == om.dim
    x = cases.rand_vec(fam.dim, 3, complex_=False)
    y = fam.apply(x)
    np.testing.assert_allclose(y, direct.apply(x), rtol=0, atol=1e-13 * np.linalg.norm(y))
    np.testing.assert_allclose(y, om.apply(x).real, rtol=0, atol=1e-13 * np.linalg.norm(y))
    e0 = np.linalg.eigvalsh(om.dense())[0]
    r = fam.eigsh(1, parameters={"hx": hx, "hz": hz})
    assert abs(r["eigenvalues"][0] - e0) <= 1e-10 * abs(e0)


def test_family_rebinding_in_sequence(qa, oracle):
    # one family evaluated at several points in turn: every evaluation is
    # independent of the previous one
    static, comps = ising_family()
    fam = qa.Model(basis(), static, components=comps)
    x = cases.rand_vec(fam.dim, 5, complex_=False)
    for hx, hz in [(0.2, 0.5), (1.0, 0.0), (0.2, 0.5), (1.7, 0.25)]:
        om = oracle.Model(L, cases.ising(L, hx, hz), momentum=0, parity=1, norm="pauli_cartesian")
        r = fam.eigsh(1, parameters={"hx": hx, "hz": hz})
        e0 = np.linalg.eigvalsh(om.dense())[0]
        assert abs(r["eigenvalues"][0] - e0) <= 1e-10 * abs(e0)
        fam.set_parameters(hx=hx, hz=hz)
        y = fam.apply(x)
        np.testing.assert_allclose(y, om.apply(x).real, rtol=0, atol=1e-13 * np.linalg.norm(y))


def test_complex_phase_split_over_components(qa, oracle):
    # forward and backward hopping in separate (non-Hermitian) components:
    # the sum is Hermitian only when c_bw = conj(c_fw), which the exact
    # sector-matrix check has to establish
    bonds = cases.bonds(L)
    static = [("zz", [(0.7, b) for b in bonds])]
    comps = [("fw", [("+-", [(0.5, b) for b in bonds])], None),
             ("bw", [("-+", [(0.5, b) for b in bonds])], None)]
    b = {"kind": "spin", "sites": L, "up": 6, "momentum": 2}
    fam = qa.Model(b, static, components=comps)
    phi = 0.37
    c = np.exp(1j * phi)
    r = fam.eigsh(2, parameters={"fw": c, "bw": np.conj(c)})
    spec = [("zz", [(0.7, bb) for bb in bonds]), ("+-", [(0.5 * c, bb) for bb in bonds]),
            ("-+", [(0.5 * np.conj(c), bb) for bb in bonds])]
    om = oracle.Model(L, spec, up=6, momentum=2)
    ex = np.linalg.eigvalsh(om.dense())[:2]
    np.testing.assert_allclose(r["eigenvalues"], ex, rtol=1e-10, atol=0)
    fam.set_parameters(fw=c, bw=np.conj(c))
    x = cases.rand_vec(fam.dim, 7)
    y = fam.apply(x)
    np.testing.assert_allclose(y, om.apply(x), rtol=0, atol=1e-13 * np.linalg.norm(y))
    with pytest.raises(qa.QmbedError, match="a Hermitian operator is required"):
        fam.eigsh(1, parameters={"fw": c, "bw": c})


@pytest.mark.parametrize("mode", ["walk", "image", "stream"])
def test_complex_coefficients_on_diagonal_parts(qa, oracle, monkeypatch, mode):
    # two diagonal components with conjugate complex coefficients (their sum
    # is a real field) and a real third: the per-row diagonal, combined from
    # the parts' stored diagonals with conjugated coefficients, must equal
    # that of the direct build in every H·v mode
    monkeypatch.setenv("QMBED_HV_MODE", mode)
    bonds = cases.bonds(L)
    static = [("zz", [(0.7, b) for b in bonds]), ("+-", [(0.5, b) for b in bonds]), ("-+", [(0.5, b) for b in bonds])]
    field = [("z", [(1.0, [i]) for i in range(L)])]
    comps = [("a", field, None), ("b", field, None), ("c", [("zz", [(1.0, b) for b in bonds])], None)]
    basis_ = {"kind": "spin", "sites": L, "momentum": 3}
    fam = qa.Model(basis_, static, components=comps)
    ca, cb, cc = 0.3 + 0.4j, 0.3 - 0.4j, 0.25
    spec = static + [("z", [(0.6, [i]) for i in range(L)]), ("zz", [(0.25, b) for b in bonds])]
    om = oracle.Model(L, spec, momentum=3)
    fam.set_parameters(a=ca, b=cb, c=cc)
    x = cases.rand_vec(fam.dim, 11)
    y = fam.apply(x)
    np.testing.assert_allclose(y, om.apply(x), rtol=0, atol=1e-13 * np.linalg.norm(y))
    ex = np.linalg.eigvalsh(om.dense())[:2]
    r = fam.eigsh(2, parameters={"a": ca, "b": cb, "c": cc})
    np.testing.assert_allclose(r["eigenvalues"], ex, rtol=1e-10, atol=0)


def test_family_errors(qa):
    static, comps = ising_family()
    fam = qa.Model(basis(), static, components=comps)
    with pytest.raises(qa.QmbedError, match='unknown operator parameter "gamma"'):
        fam.eigsh(1, parameters={"hx": 1.0, "gamma": 1.0})
    with pytest.raises(qa.QmbedError, match='missing required operator parameter "hx"'):
        fam.eigsh(1, parameters={"hz": 0.2})
    with pytest.raises(qa.QmbedError, match="component names must be nonempty and unique"):
        qa.Model(basis(), static, components=[comps[0], ("hx", comps[1][1], None)])
    with pytest.raises(qa.QmbedError, match="component names must be nonempty and unique"):
        qa.Model(basis(), static, components=[("", comps[0][1], None)])
    plain = qa.Model(basis(), cases.ising(L, 1.0, 0.0))
    with pytest.raises(qa.QmbedError, match='unknown operator parameter "hx"'):
        plain.eigsh(1, parameters={"hx": 1.0})


def test_family_through_json_commands(qa, oracle):
    static, comps = ising_family()
    req = {"operation": "create_model", "basis": basis(), "terms": qa.terms_from_spec(static),
           "components": [{"name": n, "terms": qa.terms_from_spec(t), **({"default": [d, 0.0]} if d is not None else {})}
                          for n, t, d in comps]}
    h = qa.command(req)["handle"]
    try:
        r = qa.command({"operation": "eigsh_model", "handle": h,
                        "solver": {"eigenpairs": 1, "target": {"kind": "smallest_algebraic"}},
                        "parameters": {"hx": [0.9, 0.0], "hz": [0.3, 0.0]}})
        om = oracle.Model(L, cases.ising(L, 0.9, 0.3), momentum=0, parity=1, norm="pauli_cartesian")
        e0 = np.linalg.eigvalsh(om.dense())[0]
        assert abs(r["eigenvalues"][0] - e0) <= 1e-10 * abs(e0)
        # the default of hz applies when it is absent
        r = qa.command({"operation": "eigsh_model", "handle": h,
                        "solver": {"eigenpairs": 1, "target": {"kind": "smallest_algebraic"}},
                        "parameters": {"hx": [0.9, 0.0]}})
        om = oracle.Model(L, cases.ising(L, 0.9, 0.0), momentum=0, parity=1, norm="pauli_cartesian")
        e0 = np.linalg.eigvalsh(om.dense())[0]
        assert abs(r["eigenvalues"][0] - e0) <= 1e-10 * abs(e0)
        resp = qa.raw_call("qmbed_command_json", {
            "operation": "eigsh_model", "handle": h,
            "solver": {"eigenpairs": 1, "target": {"kind": "smallest_algebraic"}}})
        assert resp == {"status": "error", "error": 'invalid options: missing required operator parameter "hx"'}
    finally:
        qa.command({"operation": "release_model", "handle": h})
