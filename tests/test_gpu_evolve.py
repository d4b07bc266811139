"""Batched Krylov evolution exp(-iHt)ψ through the device-resident entry
point (qmbed_hip_evolve_device): against the oracle's evolve (≤ 1e-10
relative), against the JSON evolve_model path, and at the full XXZ L=24
up=12 k=0 size (64 seeded states) through size-independent invariants."""

import numpy as np
import pytest

from tests import cases

pytestmark = pytest.mark.gpu

TIMES = [0.1, 1.0, 10.0]


@pytest.mark.parametrize("k", [0, 1, 6])
def test_device_evolve_matches_oracle(qa, oracle, k):
    L = 12
    spec = cases.xxz(L, 1.0)
    m = qa.Model({"kind": "spin", "sites": L, "up": 6, "momentum": k}, spec)
    om = oracle.Model(L, spec, up=6, momentum=k)
    states = np.array([cases.seeded_state(m.dim, c) for c in range(5)])
    out, diag = m.evolve(states, TIMES, krylov_dimension=30)
    assert diag["kind"] == "evolution" and diag["accepted_substeps"] > 0
    for c in range(len(states)):
        ref, st = om.evolve(states[c], TIMES, krylov_dimension=30)
        for t in range(len(TIMES)):
            assert np.linalg.norm(out[t, c] - ref[t]) <= 1e-10 * np.linalg.norm(ref[t]), (c, t)


def test_device_evolve_equals_json_path(qa):
    L = 12
    spec = cases.xxz(L, 1.0)
    basis = {"kind": "spin", "sites": L, "up": 6, "momentum": 2}
    m = qa.Model(basis, spec)
    states = np.array([cases.seeded_state(m.dim, c) for c in range(3)])
    out, diag = m.evolve(states, TIMES, krylov_dimension=30)
    h = qa.command({"operation": "create_model", "basis": basis, "terms": qa.terms_from_spec(spec)})["handle"]
    res = qa.command({"operation": "evolve_model", "handle": h,
                      "vectors": [[[z.real, z.imag] for z in v] for v in states],
                      "evolution": {"times": TIMES, "krylov_dimension": 30}})
    for t in range(len(TIMES)):
        for c in range(len(states)):
            got = np.array(res["states"][t][c])
            assert np.array_equal(got[:, 0] + 1j * got[:, 1], out[t, c])
    qa.command({"operation": "release_model", "handle": h})


def test_device_evolve_zero_time_and_zero_state(qa):
    L = 10
    m = qa.Model({"kind": "spin", "sites": L, "up": 5, "momentum": 0}, cases.xxz(L, 1.0))
    psi = cases.seeded_state(m.dim, 7)
    states = np.array([psi, np.zeros(m.dim, dtype=complex)])
    out, _ = m.evolve(states, [0.0, 0.5])
    assert np.array_equal(out[0, 0], psi)
    assert np.all(out[:, 1] == 0)
    with pytest.raises(qa.QmbedError, match="times must be a nonempty finite nondecreasing grid"):
        m.evolve(states, [1.0, 0.5])


def test_expm_64_states_L24(qa, oracle):
    # the batched-expm workload: XXZ Δ=1, L=24, up=12, k=0 (N = 112,720),
    # 64 seeded states, t ∈ {0.1, 1, 10}, Krylov 30, tolerance 1e-10
    L = 24
    spec = cases.xxz(L, 1.0)
    m = qa.Model({"kind": "spin", "sites": L, "up": 12, "momentum": 0}, spec)
    assert m.dim == 112720
    states = np.array([cases.seeded_state(m.dim, c) for c in range(64)])
    out, diag = m.evolve(states, TIMES, krylov_dimension=30)
    print(f"[expm64] projections={diag['lanczos_projections']} matvecs={diag['matrix_vector_products']} "
          f"substeps={diag['accepted_substeps']}", flush=True)
    # unitarity: norms and the Gram matrix of the 64 states are preserved
    g0 = states.conj() @ states.T
    for t in range(len(TIMES)):
        norms = np.linalg.norm(out[t], axis=1)
        assert np.max(np.abs(norms - 1.0)) <= 1e-12
        gt = out[t].conj() @ out[t].T
        assert np.max(np.abs(gt - g0)) <= 1e-9
    # energy conservation of a few columns
    for c in (0, 17, 63):
        e0 = np.vdot(states[c], m.apply(states[c])).real
        et = np.vdot(out[-1, c], m.apply(out[-1, c])).real
        assert abs(et - e0) <= 1e-9 * max(abs(e0), 1.0)
    # two columns against the oracle at full size
    om = oracle.Model(L, spec, up=12, momentum=0)
    for c in (0, 63):
        ref, _ = om.evolve(states[c], TIMES, krylov_dimension=30)
        for t in range(len(TIMES)):
            assert np.linalg.norm(out[t, c] - ref[t]) <= 1e-10 * np.linalg.norm(ref[t]), (c, t)
