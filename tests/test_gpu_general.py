"""General symmetry sectors (the QuSpin-compatible route: lattice maps with
local digit permutations and sector characters, maximum representatives and
descending rows under ``reverse``) on the GPU, against the oracle: the
representatives and orbit sizes bit for bit, the H·v in FP64 (≤ 1e-13) and
FP32, eigenvalues (≤ 1e-10) and the row order of eigenvectors, all through
the C ABI (GeneralBasis + SymmetryReducer, QMBED src/basis/mod.rs:3330-3650,
4096-4507; bindings/capi/src/lib.rs:2849-2920, 3103-3129, 3412-3422)."""

import numpy as np
import pytest

from tests import cases

pytestmark = pytest.mark.gpu


def translation(L, step=1):
    return [(i + step) % L for i in range(L)]


def reflection(L):
    return [L - 1 - i for i in range(L)]


def inversion(L):
    return list(range(L))


def general(dest, sector, flip=False):
    q = {"destinations": list(dest), "sector": sector}
    if flip:
        q["local_permutations"] = [[1, 0]] * len(dest)
    return q


def basis_json(L, up=None, symmetries=(), reverse=False, up_sectors=None, momentum=None, parity=None):
    b = {"kind": "spin", "sites": L, "symmetries": list(symmetries), "reverse": reverse}
    if up is not None:
        b["up"] = up
    if up_sectors is not None:
        b["up_sectors"] = list(up_sectors)
    if momentum is not None:
        b["momentum"] = momentum
    if parity is not None:
        b["parity"] = parity
    return b


# (name, L, up, up_sectors, generators); each runs with reverse False and True
CONFIGS = [
    ("t6k1", 6, 3, None, [general(translation(6), 1)]),
    ("t8k3", 8, 4, None, [general(translation(8), 3)]),
    ("t9k2", 9, 4, None, [general(translation(9), 2)]),
    ("t10k5_p1", 10, 5, None, [general(translation(10), 5), general(reflection(10), 1)]),
    ("t10k0_p0_z1", 10, 5, None,
     [general(translation(10), 0), general(reflection(10), 0), general(inversion(10), 1, flip=True)]),
    ("t12s2k1_z0", 12, 6, None, [general(translation(12, 2), 1), general(inversion(12), 0, flip=True)]),
    ("full10_t0_p1_z0", 10, None, None,
     [general(translation(10), 0), general(reflection(10), 1), general(inversion(10), 0, flip=True)]),
    # spin inversion out of the up=3 parent: some orbits' extreme states
    # lie in the up=5 sector
    ("t8k2_z1_up3", 8, 3, None, [general(translation(8), 2), general(inversion(8), 1, flip=True)]),
    ("z_sectors35", 8, None, [3, 5], [general(inversion(8), 0, flip=True)]),
    ("t11k4_z1", 11, None, None, [general(translation(11), 4), general(inversion(11), 1, flip=True)]),
    ("full9_t2", 9, None, None, [general(translation(9), 2)]),
    # translation and reflection do not commute at k = 1: every orbit
    # conflicts, the sector is empty
    ("t8k1_p0_empty", 8, 4, None, [general(translation(8), 1), general(reflection(8), 0)]),
]

HAMS = {
    "heisenberg": cases.heisenberg,
    "dm": cases.dm_chain,
    "j1j2": cases.j1j2,
    "xyfield": cases.xy_field,
}


def _oracle_model(oracle, L, spec, up, up_sectors, gens, reverse):
    return oracle.Model(L, spec, up=up, up_sectors=up_sectors, symmetries=gens, reverse=reverse)


def _allowed(ham, up, ups, gens):
    """Hamiltonians that commute with the sector's maps (the DM and
    y-field terms are odd under reflection and spin inversion) and keep its
    magnetization sector."""
    translation_only = all(g["destinations"] == translation(len(g["destinations"]), 1) or
                           g["destinations"] == translation(len(g["destinations"]), 2) for g in gens)
    translation_only = translation_only and not any("local_permutations" in g for g in gens)
    if ham in ("dm", "xyfield") and not translation_only:
        return False
    if ham == "xyfield" and (up is not None or ups is not None):
        return False
    return True


@pytest.mark.parametrize("reverse", [False, True])
@pytest.mark.parametrize("name,L,up,ups,gens", CONFIGS, ids=[c[0] for c in CONFIGS])
def test_general_basis_bit_exact(qa, oracle, name, L, up, ups, gens, reverse):
    print(f"[general basis] {name} reverse={reverse}", flush=True)
    om = _oracle_model(oracle, L, [], up, ups, gens, reverse)
    states, orbits = om.states()
    res = qa.command({"operation": "describe_basis", "basis": basis_json(L, up, gens, reverse, ups)})
    assert res["dimension"] == om.dim
    got = np.array([int(s) for s in res["states"]], dtype=np.uint64)
    np.testing.assert_array_equal(got, states)
    if om.dim:
        m = qa.Model(basis_json(L, up, gens, reverse, ups), cases.heisenberg(L))
        s, o = m.states()
        np.testing.assert_array_equal(s.astype(np.uint64), states)
        np.testing.assert_array_equal(o.astype(np.int32), orbits)


@pytest.mark.parametrize("reverse", [False, True])
@pytest.mark.parametrize("ham", sorted(HAMS))
@pytest.mark.parametrize("name,L,up,ups,gens", CONFIGS[:-1], ids=[c[0] for c in CONFIGS[:-1]])
def test_general_hv(qa, oracle, name, L, up, ups, gens, ham, reverse):
    if not _allowed(ham, up, ups, gens):
        pytest.skip("the term set does not commute with the sector's maps or conserve its magnetization")
    print(f"[general hv] {name} {ham} reverse={reverse}", flush=True)
    spec = HAMS[ham](L)
    om = _oracle_model(oracle, L, spec, up, ups, gens, reverse)
    m = qa.Model(basis_json(L, up, gens, reverse, ups), spec)
    assert m.dim == om.dim
    assert m.hv_config()["mode"] == "stream"
    x = cases.rand_vec(m.dim, seed=L + 7)
    y = m.apply(x)
    yr = om.apply(x)
    assert np.linalg.norm(y - yr) <= 1e-13 * max(np.linalg.norm(yr), 1.0)
    if m.is_real:
        xr = cases.rand_vec(m.dim, seed=3, complex_=False)
        np.testing.assert_allclose(m.apply(xr), om.apply(xr.astype(complex)).real, atol=1e-12)
    # FP32 product
    y32 = m.apply(x.astype(np.complex64))
    assert np.linalg.norm(y32 - yr) <= 2e-6 * max(np.linalg.norm(yr), 1.0)


@pytest.mark.parametrize("reverse", [False, True])
@pytest.mark.parametrize("name,L,up,ups,gens", CONFIGS[:-1], ids=[c[0] for c in CONFIGS[:-1]])
def test_general_eigsh(qa, oracle, name, L, up, ups, gens, reverse):
    print(f"[general eigsh] {name} reverse={reverse}", flush=True)
    spec = cases.j1j2(L)
    om = _oracle_model(oracle, L, spec, up, ups, gens, reverse)
    H = om.dense()
    w = np.linalg.eigvalsh(H)
    k = min(3, om.dim - 1)  # the reference requires eigenpairs < dimension
    if k < 1:
        pytest.skip("one-dimensional sector")
    res = qa.run({"basis": basis_json(L, up, gens, reverse, ups), "terms": qa.terms_from_spec(spec),
                  "solver": {"eigenpairs": k, "target": {"kind": "smallest_algebraic"}, "eigenvectors": True}})
    assert res["dimension"] == om.dim
    np.testing.assert_allclose(res["eigenvalues"], w[:k], rtol=1e-10, atol=1e-12)
    # the eigenvectors are in the sector's row order: eigenvectors of the
    # oracle's matrix (rows descending under reverse)
    for lam, vec in zip(res["eigenvalues"], res["eigenvectors"]):
        v = np.array([complex(a, b) for a, b in vec])
        assert np.linalg.norm(H @ v - lam * v) <= 1e-8 * max(abs(lam), 1.0)


def test_general_translation_equals_builtin_momentum(qa, oracle):
    # the reference's runtime-symmetry check (tests/runtime_symmetry.rs:9-39,
    # bindings/capi/src/lib.rs:8074-8110): the translation map at sector k
    # is the built-in momentum-k sector, rows and matrix alike
    res = qa.command({"operation": "describe_basis",
                      "basis": basis_json(6, 3, [general(translation(6), 1)])})
    assert [int(s) for s in res["states"]] == [7, 11, 13]
    L = 12
    spec = cases.dm_chain(L)
    # (under reverse the general sector takes orbit maxima and the built-in
    # one keeps minima, so only the ascending convention coincides)
    for k in (0, 1, 5, 6):
        for reverse in (False,):
            g = qa.Model(basis_json(L, 6, [general(translation(L), k)], reverse), spec)
            b = qa.Model(basis_json(L, 6, momentum=k, reverse=reverse), spec)
            sg, og = g.states()
            sb, ob = b.states()
            np.testing.assert_array_equal(sg, sb)
            np.testing.assert_array_equal(og, ob)
            x = cases.rand_vec(g.dim, seed=k)
            np.testing.assert_allclose(g.apply(x), b.apply(x), atol=1e-13)


@pytest.mark.parametrize("k,p", [(0, 1), (0, -1), (3, None), (6, -1)])
def test_reversed_builtin_sector(qa, oracle, k, p):
    # the reverse convention on a built-in sector: minimum representatives,
    # descending rows; H·v and eigenvectors in that order
    L = 12
    spec = cases.heisenberg(L)
    om = oracle.Model(L, spec, up=6, momentum=k, parity=p, reverse=True)
    asc = oracle.Model(L, spec, up=6, momentum=k, parity=p)
    np.testing.assert_array_equal(om.states()[0], asc.states()[0][::-1])
    m = qa.Model(basis_json(L, 6, momentum=k, parity=p, reverse=True), spec)
    s, o = m.states()
    np.testing.assert_array_equal(s.astype(np.uint64), om.states()[0])
    np.testing.assert_array_equal(o.astype(np.int32), om.states()[1])
    x = cases.rand_vec(m.dim, seed=11)
    yr = om.apply(x)
    assert np.linalg.norm(m.apply(x) - yr) <= 1e-13 * np.linalg.norm(yr)
    r = m.eigsh(2, eigenvectors=True)
    w = np.linalg.eigvalsh(om.dense())
    np.testing.assert_allclose(r["eigenvalues"], w[:2], rtol=1e-10)
    H = om.dense()
    for lam, vec in zip(r["eigenvalues"], r["eigenvectors"]):
        v = np.array([complex(a, b) for a, b in vec])
        assert np.linalg.norm(H @ v - lam * v) <= 1e-8 * max(abs(lam), 1.0)


def test_general_sector_at_size(qa, oracle):
    # L=20 half filling with translation (k=0), reflection (even) and spin
    # inversion (odd), QuSpin rows: the lowest pair against the oracle's
    # Lanczos
    L = 20
    gens = [general(translation(L), 0), general(reflection(L), 0), general(inversion(L), 1, flip=True)]
    spec = cases.xxz(L, 1.5)
    om = oracle.Model(L, spec, up=L // 2, symmetries=gens, reverse=True)
    ref = om.eigsh(2)
    res = qa.run({"basis": basis_json(L, L // 2, gens, True), "terms": qa.terms_from_spec(spec),
                  "solver": {"eigenpairs": 2, "target": {"kind": "smallest_algebraic"}}})
    assert res["dimension"] == om.dim
    np.testing.assert_allclose(res["eigenvalues"], ref["eigenvalues"], rtol=1e-10)


def test_general_sector_errors(qa):
    def err(basis):
        with pytest.raises(qa.QmbedError) as e:
            qa.command({"operation": "describe_basis", "basis": basis})
        return str(e.value)

    assert "symmetry map has 2 sites, expected 6" in err(basis_json(6, 3, [general([1, 0], 0)]))
    assert err(basis_json(6, 3, [general([1, 1, 2, 3, 4, 5], 0)])).startswith(
        "incompatible symmetry request: symmetry site destinations must be a bijection")
    bad_local = general(translation(6), 0)
    bad_local["local_permutations"] = [[0, 0]] * 6
    assert err(basis_json(6, 3, [bad_local])).startswith(
        "incompatible symmetry request: every local-state map must be a permutation")
    assert "built-in and general spin symmetries cannot be mixed" in err(
        basis_json(6, 3, [general(translation(6), 1)], momentum=1))
    b = basis_json(6, 3)
    b["matrix_symmetry"] = {"dimension": 2, "selected_row": 0, "generators": []}
    assert "matrix-valued symmetry sectors materialize as projected operator models" in err(b)


def ladder(L, jr=1.0, jl=0.8):
    """Two-leg ladder on L = 2·rungs sites: rungs (2i, 2i+1), legs (2i, 2i+2)
    and (2i+1, 2i+3) periodic; S·S couplings jr on rungs, jl on legs."""
    bonds = [(jr, [2 * i, 2 * i + 1]) for i in range(L // 2)]
    bonds += [(jl, [2 * i, (2 * i + 2) % L]) for i in range(L // 2)]
    bonds += [(jl, [2 * i + 1, (2 * i + 3) % L]) for i in range(L // 2)]
    return [("zz", [(c, s) for c, s in bonds]), ("+-", [(0.5 * c, s) for c, s in bonds]),
            ("-+", [(0.5 * c, s) for c, s in bonds])]


def leg_swap(L):
    return [i + 1 if i % 2 == 0 else i - 1 for i in range(L)]


# groups with an element that is neither a rotation nor a reflection of the
# ring (the leg swap of a ladder): their elements move the sites one by one
LADDER = [
    ("ladder12_t2k1_s0", 12, 6, [general(translation(12, 2), 1), general(leg_swap(12), 0)]),
    ("ladder12_t2k0_s1_z0", 12, 6,
     [general(translation(12, 2), 0), general(leg_swap(12), 1), general(inversion(12), 0, flip=True)]),
    ("ladder10_s1", 10, 5, [general(leg_swap(10), 1)]),
]


@pytest.mark.parametrize("reverse", [False, True])
@pytest.mark.parametrize("name,L,up,gens", LADDER, ids=[c[0] for c in LADDER])
def test_general_non_ring_group(qa, oracle, name, L, up, gens, reverse):
    print(f"[general ladder] {name} reverse={reverse}", flush=True)
    spec = ladder(L)
    om = _oracle_model(oracle, L, spec, up, None, gens, reverse)
    states, orbits = om.states()
    m = qa.Model(basis_json(L, up, gens, reverse), spec)
    assert m.dim == om.dim
    s, o = m.states()
    np.testing.assert_array_equal(s.astype(np.uint64), states)
    np.testing.assert_array_equal(o.astype(np.int32), orbits)
    x = cases.rand_vec(m.dim, seed=L + 11)
    yr = om.apply(x)
    assert np.linalg.norm(m.apply(x) - yr) <= 1e-13 * max(np.linalg.norm(yr), 1.0)
    w = np.linalg.eigvalsh(om.dense())
    k = min(2, om.dim - 1)
    res = qa.run({"basis": basis_json(L, up, gens, reverse), "terms": qa.terms_from_spec(spec),
                  "solver": {"eigenpairs": k, "target": {"kind": "smallest_algebraic"}}})
    np.testing.assert_allclose(res["eigenvalues"], w[:k], rtol=1e-10, atol=1e-12)
