"""Sector batch: many momentum/magnetization sectors solved in one lockstep
Lanczos batch (qmbed_hip_eigsh_batch_json), against the oracle's dense
spectra, the single-sector GPU eigsh, and the full (unreduced) sector."""

import numpy as np
import pytest

from tests import cases

pytestmark = pytest.mark.gpu


def _sector_models(qa, L, ups, delta=1.0):
    spec = cases.xxz(L, delta)
    keys = [(up, k) for up in ups for k in range(L)]
    models = [qa.Model({"kind": "spin", "sites": L, "up": up, "momentum": k}, spec) for up, k in keys]
    return spec, keys, models


def test_batch_ground_states_match_dense_oracle(qa, oracle):
    # L=16, up ∈ {7, 8, 9} × k ∈ 0..15: 48 sectors of dimension 715-810
    # (complex Hermitian except k = 0, 8), every one on the Krylov route
    L = 16
    spec, keys, models = _sector_models(qa, L, (7, 8, 9))
    out = qa.eigsh_batch(models, 1)
    batch = out["batch"]
    print(f"[batch] L={L} sectors={batch['sectors']} steps={batch['steps']} ritz={batch['ritz_rounds']} "
          f"fallback={batch['fallback']}", flush=True)
    assert batch["sectors"] == len(keys) and batch["batched"] == len(keys) and batch["fallback"] == 0
    for (up, k), m, r in zip(keys, models, out["results"]):
        om = oracle.Model(L, spec, up=up, momentum=k)
        assert om.dim == m.dim == r["dimension"]
        e0 = np.linalg.eigvalsh(om.dense())[0]
        assert abs(r["eigenvalues"][0] - e0) <= 1e-10 * abs(e0), (up, k)
        assert r["residuals"][0] <= 1e-10 * max(abs(e0), 1.0)
        assert r["converged"] and r["iterations"] > 1
        assert r["reorthogonalization_passes"] >= r["iterations"]


def test_batch_six_pairs_and_vectors(qa, oracle):
    L = 16
    spec, keys, models = _sector_models(qa, L, (8,))
    out = qa.eigsh_batch(models, 6, eigenvectors=True)
    for (up, k), m, r in zip(keys, models, out["results"]):
        om = oracle.Model(L, spec, up=up, momentum=k)
        ref = np.linalg.eigvalsh(om.dense())[:6]
        np.testing.assert_allclose(r["eigenvalues"], ref, rtol=1e-10, atol=1e-10)
        for lam, vec in zip(r["eigenvalues"], r["eigenvectors"]):
            v = np.array(vec)
            v = v[:, 0] + 1j * v[:, 1]
            hv = om.apply(v)
            assert abs(np.linalg.norm(v) - 1.0) <= 1e-12
            assert np.linalg.norm(hv - lam * v) <= 1e-10 * max(abs(lam), 1.0)


def test_batch_matches_single_sector_solves(qa):
    L = 18
    spec, keys, models = _sector_models(qa, L, (9,))
    out = qa.eigsh_batch(models, 2)
    for (up, k), m, r in zip(keys, models, out["results"]):
        single = m.eigsh(2)
        np.testing.assert_allclose(r["eigenvalues"], single["eigenvalues"], rtol=1e-10, atol=0)
    # lockstep batches are deterministic: a second run is bit-identical
    again = qa.eigsh_batch(models, 2)
    for r, s in zip(out["results"], again["results"]):
        assert r["eigenvalues"] == s["eigenvalues"] and r["iterations"] == s["iterations"]


def test_batch_momentum_minimum_is_full_sector_ground_state(qa, oracle):
    # the union of the momentum sectors is the whole up sector: the lowest
    # sector ground state is the ground state of the unreduced up = 8 sector
    L = 16
    spec, keys, models = _sector_models(qa, L, (8,))
    out = qa.eigsh_batch(models, 1)
    e_min = min(r["eigenvalues"][0] for r in out["results"])
    full = qa.Model({"kind": "spin", "sites": L, "up": 8}, spec)
    e_full = full.eigsh(1)["eigenvalues"][0]
    om = oracle.Model(L, spec, up=8)
    e_ref = np.linalg.eigvalsh(om.dense())[0]
    assert abs(e_min - e_ref) <= 1e-10 * abs(e_ref)
    assert abs(e_full - e_ref) <= 1e-10 * abs(e_ref)


def test_batch_forced_second_passes(qa, oracle):
    # η = 1 takes the DGKS second pass on (nearly) every step: the listed
    # second-pass launches run for the flagged sectors only
    L = 16
    spec, keys, models = _sector_models(qa, L, (8,))
    out = qa.eigsh_batch(models, 1, reorthogonalization_eta=1.0)
    seconds = sum(r["conditional_second_passes"] for r in out["results"])
    assert seconds > 0
    for (up, k), r in zip(keys, out["results"]):
        e0 = np.linalg.eigvalsh(oracle.Model(L, spec, up=up, momentum=k).dense())[0]
        assert abs(r["eigenvalues"][0] - e0) <= 1e-10 * abs(e0)


def test_batch_mixed_routes(qa, oracle):
    # small sectors take the dense route and are solved one by one in the
    # same call; results stay in input order
    spec16 = cases.xxz(16, 1.0)
    spec10 = cases.xxz(10, 1.0)
    ms = [qa.Model({"kind": "spin", "sites": 16, "up": 8, "momentum": 3}, spec16),
          qa.Model({"kind": "spin", "sites": 10, "up": 5, "momentum": 1}, spec10),
          qa.Model({"kind": "spin", "sites": 16, "up": 7, "momentum": 0}, spec16)]
    out = qa.eigsh_batch(ms, 1)
    assert out["batch"]["batched"] == 2
    refs = [oracle.Model(16, spec16, up=8, momentum=3), oracle.Model(10, spec10, up=5, momentum=1),
            oracle.Model(16, spec16, up=7, momentum=0)]
    for om, r in zip(refs, out["results"]):
        e0 = np.linalg.eigvalsh(om.dense())[0]
        assert abs(r["eigenvalues"][0] - e0) <= 1e-10 * abs(e0)


def test_batch_chunks_and_targets(qa, oracle):
    # consecutive batches of whole sectors give the same results as one batch
    L = 16
    spec, keys, models = _sector_models(qa, L, (8,))
    one = qa.eigsh_batch(models, 2)
    many = qa.eigsh_batch(models, 2, batch_rows=3 * 2048)
    assert one["batch"]["batches"] == 1 and many["batch"]["batches"] == 6
    for a, b in zip(one["results"], many["results"]):
        assert a["eigenvalues"] == b["eigenvalues"]
    # other extreme targets
    for target in ("largest_algebraic", "both_ends", "largest_magnitude"):
        out = qa.eigsh_batch(models[:4], 2, target=target)
        for (up, k), r in zip(keys[:4], out["results"]):
            ev = np.linalg.eigvalsh(oracle.Model(L, spec, up=up, momentum=k).dense())
            if target == "largest_algebraic":
                ref = sorted(ev[-2:], reverse=True)
            elif target == "both_ends":
                ref = [ev[0], ev[-1]]
            else:
                ref = sorted(ev, key=lambda x: -abs(x))[:2]
            assert np.allclose(sorted(r["eigenvalues"]), sorted(ref), rtol=1e-10, atol=1e-10), (target, k)


def test_batch_errors(qa):
    spec = cases.xxz(16, 1.0)
    a = qa.Model({"kind": "spin", "sites": 16, "up": 8, "momentum": 0}, spec)
    b = qa.Model({"kind": "spin", "sites": 16, "up": 8, "momentum": 0, "parity": 1}, spec)
    with pytest.raises(qa.QmbedError, match="same symmetry kind"):
        qa.eigsh_batch([a, b], 1)
    with pytest.raises(qa.QmbedError, match="appears twice"):
        qa.eigsh_batch([a, a], 1)
    with pytest.raises(qa.QmbedError, match="FP64"):
        qa.eigsh_batch([a], 1, matvec_precision="fp32")
    # a small explicit window thick-restarts, in the batch (its single-sector
    # finish) as in the single solver
    rb = qa.eigsh_batch([a], 1, krylov_dimension=4)["results"][0]
    rs = a.eigsh(1, krylov_dimension=4)
    assert rb["eigenvalues"] == rs["eigenvalues"] and rb["iterations"] == rs["iterations"] > 4
    # a budget too small to converge: the same error as the single solver
    with pytest.raises(qa.QmbedError, match="did not converge") as e1:
        qa.eigsh_batch([a], 1, krylov_dimension=4, max_iterations=8)
    with pytest.raises(qa.QmbedError, match="did not converge") as e2:
        a.eigsh(1, krylov_dimension=4, max_iterations=8)
    assert str(e1.value) == str(e2.value)


def test_batch_json_command(qa, oracle):
    spec = cases.xxz(16, 1.0)
    hs = [qa.command({"operation": "create_model", "basis": {"kind": "spin", "sites": 16, "up": 8, "momentum": k},
                      "terms": qa.terms_from_spec(spec)})["handle"] for k in (0, 5)]
    res = qa.command({"operation": "eigsh_models", "handles": hs,
                      "solver": {"eigenpairs": 1, "target": {"kind": "smallest_algebraic"}}})
    assert res["kind"] == "eigensystems" and len(res["results"]) == 2
    for k, r in zip((0, 5), res["results"]):
        e0 = np.linalg.eigvalsh(oracle.Model(16, spec, up=8, momentum=k).dense())[0]
        assert abs(r["eigenvalues"][0] - e0) <= 1e-10 * abs(e0)
    for h in hs:
        qa.command({"operation": "release_model", "handle": h})


def test_batch_72_sectors_L24(qa):
    # the full-size batch: XXZ Δ=1, L=24, up ∈ {11, 12, 13} × k ∈ 0..23
    # (72 sectors, dimensions 104,006-112,720, Σ = 7,696,444)
    L = 24
    spec, keys, models = _sector_models(qa, L, (11, 12, 13))
    dims = [m.dim for m in models]
    assert (min(dims), max(dims), sum(dims)) == (104006, 112720, 7696444)
    out = qa.eigsh_batch(models, 1)
    b = out["batch"]
    print(f"[batch72] steps={b['steps']} ritz={b['ritz_rounds']} fallback={b['fallback']} "
          f"seconds={b['seconds']:.3f}", flush=True)
    assert b["batched"] == 72 and b["fallback"] == 0
    res = out["results"]
    for r in res:
        assert r["residuals"][0] <= 1e-10 * max(abs(r["eigenvalues"][0]), 1.0)
    # per-sector single solves agree (a real and two complex sectors)
    for idx in (keys.index((12, 0)), keys.index((12, 5)), keys.index((11, 12))):
        single = models[idx].eigsh(1)["eigenvalues"][0]
        assert abs(res[idx]["eigenvalues"][0] - single) <= 1e-10 * abs(single)
    # up ↔ down symmetry of the XXZ chain: sectors up = 11 and 13 are equal
    for k in range(L):
        e11 = res[keys.index((11, k))]["eigenvalues"][0]
        e13 = res[keys.index((13, k))]["eigenvalues"][0]
        assert abs(e11 - e13) <= 1e-10 * abs(e11)
    # momentum k and L - k are complex conjugate sectors with equal spectra
    for up in (11, 12, 13):
        for k in range(1, L // 2):
            a = res[keys.index((up, k))]["eigenvalues"][0]
            c = res[keys.index((up, L - k))]["eigenvalues"][0]
            assert abs(a - c) <= 1e-10 * abs(a)
