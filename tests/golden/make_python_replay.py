"""Write tests/golden/python_replay.json: the C-ABI requests that the
reference's Python tests send on the spin eigsh path, as request/expectation
sequences that tests/test_gpu_python_replay.py replays through this build's
library.

The requests are the JSON the reference's Python package builds for those
calls (qmbed.eigsh and qmbed.compat.quspin.eigsh: bindings/python/src/qmbed/
model.py:66-300, 311-345, compat/quspin.py:29-122; the QuSpin-compatible
bases and hamiltonian: quspin/basis/__init__.py:38-265, 1457-1573,
2356-2405, quspin/operators/__init__.py:86-87, 993-1125, 1304-1368, and the
model commands of qmbed/_ffi.py:470-525), written here from the schema; no
reference file is imported or copied. Expectations are the reference tests'
own assertions (bindings/python/tests/test_qmbed.py:57-150) plus values of
this repository's CPU oracle for the same requests (eigenvalues, basis
states, dense matrices), which pin what the assertions leave open.

Run from the repository root after building the oracle:
    python tests/golden/make_python_replay.py
"""

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle_ffi  # noqa: E402

LOCAL = {"I": "identity", "n": "number", "z": "z", "+": "raising", "-": "lowering", "x": "x", "y": "y"}


def term(symbols, rows):
    """One operator string with QuSpin coupling rows [coefficient, site, ...]
    (OpProduct without species splits, Coupling.request)."""
    couplings = []
    for coefficient, *sites in rows:
        c = complex(coefficient)
        couplings.append({"coefficient": [c.real, c.imag], "sites": [int(s) for s in sites]})
    return {"product": {"local": [LOCAL[s] for s in symbols]}, "couplings": couplings}


def native_spin_basis(sites):
    """qmbed.SpinBasis(sites).request()"""
    return {"kind": "spin", "sites": sites, "spin_twice": 1, "up": None, "momentum": None, "parity": None,
            "pauli": False}


def eigsh_options(k, target="smallest_algebraic"):
    """qmbed.EigshOptions(k, target=...).request()"""
    return {"eigenpairs": k, "target": {"kind": target}, "krylov_dimension": None, "tolerance": 1.0e-10,
            "max_iterations": 1000, "seed": 0, "eigenvectors": False}


def rust_site_map(python_map, sites):
    """A QuSpin site map (site i -> python_map[i], negative entries inverting
    the digit) in the Rust site order: site i of QuSpin is Rust site
    sites - 1 - i."""
    dest = [0] * sites
    flip = [False] * sites
    for src, d in enumerate(python_map):
        inverted = d < 0
        pd = -d - 1 if inverted else d
        rs = sites - src - 1
        dest[rs] = sites - pd - 1
        flip[rs] = inverted
    return dest, flip


def symmetry_request(python_map, sector, sites):
    dest, flip = rust_site_map(python_map, sites)
    q = {"destinations": dest, "sector": int(sector)}
    if any(flip):
        q["local_permutations"] = [[1, 0] if f else [0, 1] for f in flip]
    return q


def quspin_spin_basis_1d(L, Nup=None, pauli=True, kblock=None):
    """spin_basis_1d(L, Nup, pauli, kblock) request (no reflection blocks, so
    no matrix-valued sector)."""
    syms = []
    if kblock is not None:
        syms.append(symmetry_request([(i + 1) % L for i in range(L)], kblock, L))
    return {"kind": "spin", "sites": L, "spin_twice": 1, "up": Nup, "up_sectors": None, "momentum": None,
            "parity": None, "normalization": {0: "angular_momentum", 1: "pauli", -1: "pauli_cartesian"}[int(pauli)],
            "symmetries": syms, "matrix_symmetry": None, "reverse": True}


def quspin_spin_basis_general(N, Nup=None, pauli=True, **blocks):
    syms = [symmetry_request(m, q, N) for m, q in blocks.values()]
    return {"kind": "spin", "sites": N, "spin_twice": 1, "up": Nup, "up_sectors": None, "momentum": None,
            "parity": None, "normalization": {0: "angular_momentum", 1: "pauli", -1: "pauli_cartesian"}[int(pauli)],
            "symmetries": syms, "reverse": True}


def hamiltonian_model(basis, static, checks=(True, True, True)):
    """hamiltonian(static, [], basis=...) create_model request"""
    L = basis["sites"]
    return {"operation": "create_model", "basis": basis, "terms": [term(s, rows) for s, rows in static],
            "components": [], "site_permutation": list(range(L - 1, -1, -1)),
            "checks": {"hermiticity": checks[0], "particle_conservation": checks[1],
                       "symmetry_compatibility": checks[2]}}


def oracle_model(basis, terms, permutation=None):
    """The oracle's model for a basis request and JSON terms."""
    norm = basis.get("normalization") or ("pauli_cartesian" if basis.get("pauli") else "angular_momentum")
    norm = {"angular_momentum": "angular", "pauli": "pauli", "pauli_cartesian": "pauli_cartesian"}[norm]
    spec = []
    names = {v: k for k, v in LOCAL.items()}
    for t in terms:
        sym = "".join(names[x] for x in t["product"]["local"])
        cpl = []
        for c in t["couplings"]:
            sites = [permutation[s] for s in c["sites"]] if permutation else c["sites"]
            cpl.append((complex(*c["coefficient"]), sites))
        spec.append((sym, cpl))
    return oracle_ffi.Model(basis["sites"], spec, up=basis.get("up"), momentum=basis.get("momentum"),
                            parity=basis.get("parity"), norm=norm, symmetries=basis.get("symmetries") or None,
                            reverse=bool(basis.get("reverse")), check_hermiticity=False, check_particles=False)


def dense_entries(a):
    return [[int(i), int(j), [float(a[i, j].real), float(a[i, j].imag)]]
            for i in range(a.shape[0]) for j in range(a.shape[1]) if abs(a[i, j]) > 0]


def build_cases():
    cases = []

    # test_qmbed.py:57-72, the native eigsh of the spin-1/2 dimer
    terms = [term("zz", [[1.0, 0, 1]]), term("+-", [[0.5, 0, 1]]), term("-+", [[0.5, 0, 1]])]
    req = {"basis": native_spin_basis(2), "terms": terms, "format": "csc", "solver": eigsh_options(2)}
    w = np.linalg.eigvalsh(oracle_model(req["basis"], terms).dense())
    cases.append({"name": "native_eigsh_dimer", "source": "bindings/python/tests/test_qmbed.py:57-72",
                  "steps": [{"call": "run", "request": req,
                             "expect": {"dimension": 4, "converged": True,
                                        "eigenvalues": {"values": [-0.75], "atol": 1e-10},
                                        "eigenvalues_oracle": {"values": list(w[:2]), "rtol": 1e-10}}}]})

    # test_qmbed.py:74-85, quspin.eigsh(SpinBasis(2), static, k=2, which="SA")
    static = [("zz", [[1.0, 0, 1]]), ("+-", [[0.5, 0, 1]]), ("-+", [[0.5, 0, 1]])]
    req = {"basis": native_spin_basis(2), "terms": [term(s, r) for s, r in static], "format": "csc",
           "solver": eigsh_options(2)}
    cases.append({"name": "quspin_static_adapter", "source": "bindings/python/tests/test_qmbed.py:74-85",
                  "steps": [{"call": "run", "request": req,
                             "expect": {"eigenvalues": {"values": [-0.75], "atol": 1e-10},
                                        "eigenvalues_oracle": {"values": list(w[:2]), "rtol": 1e-10}}}]})

    # test_qmbed.py:87-111, one hamiltonian model: dense matrix, eigvalsh,
    # description, release
    basis = quspin_spin_basis_1d(2)
    create = hamiltonian_model(basis, [("zz", [[1.0, 0, 1]])])
    om = oracle_model(basis, create["terms"], create["site_permutation"])
    dense = om.dense()
    cases.append({"name": "quspin_hamiltonian_lifecycle", "source": "bindings/python/tests/test_qmbed.py:87-111",
                  "steps": [
                      {"call": "command", "request": create, "save": "h", "expect": {"dimension": 4, "kind": "model"}},
                      {"call": "command",
                       "request": {"operation": "materialize_model", "handle": "$h", "parameters": {}, "format": "csc"},
                       "expect": {"shape": [4, 4], "dense": {"entries": dense_entries(dense), "atol": 1e-13}}},
                      {"call": "command",
                       "request": {"operation": "eigh_model", "handle": "$h", "parameters": {}, "eigenvectors": False},
                       "expect": {"eigenvalues": {"values": list(np.linalg.eigvalsh(dense)), "atol": 1e-12},
                                  "count": 4}},
                      {"call": "command", "request": {"operation": "describe_model", "handle": "$h"},
                       "expect": {"dimension": 4}},
                      {"call": "command", "request": {"operation": "release_model", "handle": "$h"},
                       "expect": {"kind": "released"}},
                      {"call": "command", "request": {"operation": "describe_model", "handle": "$h"},
                       "expect_error": "is not registered"},
                  ]})

    # test_qmbed.py:113-150, a QuSpin site map reproduces the optimized
    # momentum sector: equal states and equal matrices
    L = 6
    general = quspin_spin_basis_general(L, Nup=3, pauli=False, translation=([(i + 1) % L for i in range(L)], 1))
    optimized = quspin_spin_basis_1d(L, Nup=3, pauli=False, kblock=1)
    static = [["+-", [[1.0j, s, (s + 1) % L] for s in range(L)]], ["-+", [[-1.0j, s, (s + 1) % L] for s in range(L)]]]
    checks = (False, False, False)
    og = oracle_model(general, [term(s, r) for s, r in static], list(range(L - 1, -1, -1)))
    states = [str(int(x)) for x in og.states()[0]]
    steps = [
        {"call": "command",
         "request": {"operation": "create_basis_plan", "basis": general, "site_permutation": list(range(L - 1, -1, -1)),
                     "checks": {"hermiticity": False, "particle_conservation": False, "symmetry_compatibility": False}},
         "save": "plan", "expect": {"kind": "basis_plan"}},
        {"call": "command", "request": {"operation": "materialize_basis_plan", "plan_handle": "$plan"},
         "save": "basis_model", "expect": {"kind": "model", "dimension": len(states)}},
        {"call": "command", "request": {"operation": "describe_basis", "basis": general},
         "expect": {"states": states}},
        {"call": "command", "request": {"operation": "describe_basis", "basis": optimized},
         "expect": {"states": states}},
    ]
    for name, b in (("g", general), ("o", optimized)):
        create = hamiltonian_model(b, static, checks)
        steps.append({"call": "command", "request": create, "save": name, "expect": {"dimension": len(states)}})
        steps.append({"call": "command",
                      "request": {"operation": "materialize_model", "handle": "$" + name, "parameters": {},
                                  "format": "csc"},
                      "expect": {"dense": {"entries": dense_entries(og.dense()), "atol": 1e-12}}})
    steps.append({"call": "command", "request": {"operation": "release_basis_plan", "plan_handle": "$plan"},
                  "expect": {"kind": "released"}})
    cases.append({"name": "python_site_maps_momentum_sector",
                  "source": "bindings/python/tests/test_qmbed.py:113-150", "steps": steps})

    return cases


def main():
    out = os.path.join(HERE, "python_replay.json")
    with open(out, "w") as f:
        json.dump({"generated_by": "tests/golden/make_python_replay.py", "cases": build_cases()}, f, indent=1)
    print(out)


if __name__ == "__main__":
    main()
