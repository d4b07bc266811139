"""Pins the CPU oracle (CPU only).

Every check is independent of the oracle's own formulas: the reference's
inline anchors, closed-form sector counts (necklace / bracelet counting),
spectra of Hamiltonians built here from Kronecker products, explicit symmetry
projectors, scipy's dense exponential, and a brute-force search of the
Taylor degree/scaling table.
"""

import math

import numpy as np
import pytest
import scipy.linalg

from tests import cases

# ----------------------------------------------------------- helpers -------

SX = np.array([[0, 1], [1, 0]], dtype=complex) / 2
SY = np.array([[0, 1j], [-1j, 0]], dtype=complex) / 2  # (down, up) ordering: Sy|down> = -i/2 |up>
SZ = np.array([[-1, 0], [0, 1]], dtype=complex) / 2  # digit 0 = down (m=-1/2)
SP = np.array([[0, 0], [1, 0]], dtype=complex)       # raises digit 0 -> 1
SM = SP.T.copy()
ID = np.eye(2, dtype=complex)


def local_matrix(sym, norm="angular"):
    m = {"x": SX, "y": SY, "z": SZ, "+": SP, "-": SM, "I": ID}[sym]
    if norm == "pauli_cartesian" and sym in "xyz":
        m = 2 * m
    if norm == "pauli" and sym != "I":
        m = 2 * m
    return m


def full_hamiltonian(L, spec, norm="angular"):
    """Dense H on all 2^L states, basis index = integer state (bit i = site i)."""
    dim = 2 ** L
    H = np.zeros((dim, dim), dtype=complex)
    for sym, couplings in spec:
        for coeff, sites in couplings:
            op = np.eye(dim, dtype=complex)
            # product O_0 O_1 ... O_{n-1}; O_{n-1} acts first
            for s, site in zip(sym, sites):
                mats = [ID] * L
                mats[site] = local_matrix(s, norm)
                full = np.array([[1.0 + 0j]])
                for q in range(L - 1, -1, -1):  # site L-1 is the most significant bit
                    full = np.kron(full, mats[q])
                op = op @ full
            H += coeff * op
    return H


def rotate(s, k, L):
    k %= L
    mask = (1 << L) - 1
    return ((s << k) & mask) | (s >> (L - k)) if k else s


def reflect(s, L):
    return int(format(s, f"0{L}b")[::-1], 2)


def mobius(n):
    res, p, m = 1, 2, n
    while p * p <= m:
        if m % p == 0:
            m //= p
            if m % p == 0:
                return 0
            res = -res
        p += 1
    return -res if m > 1 else res


def primitive_necklaces(d, w):
    """Aperiodic binary necklaces of length d and weight w."""
    tot = 0
    for e in range(1, d + 1):
        if d % e == 0 and w % e == 0:
            tot += mobius(e) * math.comb(d // e, w // e)
    return tot // d


def momentum_sector_dim(L, w, k):
    """Orbits of period d with (w·d) ≡ 0 (mod L) and k·d ≡ 0 (mod L)."""
    tot = 0
    for d in range(1, L + 1):
        if L % d or (w * d) % L or (k * d) % L:
            continue
        tot += primitive_necklaces(d, w * d // L)
    return tot


def bracelets(n):
    neck = sum(_phi(d) * 2 ** (n // d) for d in range(1, n + 1) if n % d == 0) // n
    if n % 2 == 0:
        return (neck + 3 * 2 ** (n // 2 - 1)) // 2
    return (neck + 2 ** ((n + 1) // 2)) // 2


def _phi(n):
    return sum(1 for i in range(1, n + 1) if math.gcd(i, n) == 1)


# ------------------------------------------------- reference anchors -------

def test_reference_sector_anchors(oracle):
    # tests/visible_contract.rs:114-121, :158-163
    assert oracle.Model(4, [], up=2).dim == 6
    assert oracle.Model(4, [], up=2, momentum=0).dim == 2
    assert oracle.Model(4, [], up=2, momentum=1).dim == 1
    # tests/full_api_foundation.rs:160-167
    assert oracle.Model(4, [], up=2, parity=1).dim == 4
    assert oracle.Model(4, [], up=2, parity=-1).dim == 2
    # tests/full_api_foundation.rs:191-197
    with pytest.raises(oracle.OracleError, match="incompatible symmetry request"):
        oracle.Model(4, [], up=2, momentum=1, parity=1)
    # tests/visible_contract.rs:192-194
    assert oracle.Model(18, [], up=9, momentum=0).dim == 2704


def test_reference_reduction_metadata(oracle):
    # bindings/capi/src/lib.rs:8700-8735: L=4 up=2 k=1
    m = oracle.Model(4, [], up=2, momentum=1)
    rep, phase, orbit = m.image(3)
    assert (rep, orbit) == (3, 4) and abs(phase - 1) < 1e-15
    rep, phase, orbit = m.image(6)
    assert rep == 3 and orbit == 4
    assert abs(phase - (-1j)) < 1e-15  # exp(-i pi/2): state 6 = T^1 3
    assert abs(0.5 * phase - (-0.5j)) < 1e-15  # amplitude 1/sqrt(4) * phase
    assert m.image(5) is None and m.image(0) is None


def test_reference_energy_anchors(oracle):
    # dimer -0.75 (tests/visible_contract.rs:991-1007), L=4 k=0 Heisenberg -2 (:171-179)
    dimer = [("zz", [(1.0, [0, 1])]), ("+-", [(0.5, [0, 1])]), ("-+", [(0.5, [0, 1])])]
    ev = np.linalg.eigvalsh(oracle.Model(2, dimer).dense())
    np.testing.assert_allclose(ev, [-0.75, 0.25, 0.25, 0.25], atol=1e-14)
    ev = np.linalg.eigvalsh(oracle.Model(4, cases.heisenberg(4), up=2, momentum=0).dense())
    assert abs(ev[0] + 2) < 1e-13


# -------------------------------------------------- closed-form counts -----

@pytest.mark.parametrize("L,w,k", [(16, 8, 0), (12, 6, 0), (12, 6, 6), (12, 5, 3), (10, 5, 1), (14, 7, 7), (9, 3, 0),
                                   (18, 9, 0), (15, 6, 5)])
def test_momentum_dims_closed_form(oracle, L, w, k):
    assert oracle.Model(L, [], up=w, momentum=k).dim == momentum_sector_dim(L, w, k)


def test_config_dims_closed_form():
    # the benchmark sectors; these are too large for the oracle's hash map
    assert momentum_sector_dim(16, 8, 0) == 810
    assert momentum_sector_dim(24, 12, 0) == 112720
    assert momentum_sector_dim(28, 14, 0) == 1432860
    dims = [momentum_sector_dim(24, w, k) for w in (11, 12, 13) for k in range(24)]
    assert (min(dims), max(dims), sum(dims)) == (104006, 112720, 7696444)
    assert bracelets(26) == 1296858


@pytest.mark.parametrize("L", [8, 10, 12, 13])
def test_parity_k0_is_bracelets(oracle, L):
    # k=0, p=+1 on the unconstrained space: every dihedral orbit survives
    assert oracle.Model(L, [], momentum=0, parity=1).dim == bracelets(L)


@pytest.mark.parametrize("L,up,k,p", cases.SECTORS)
def test_basis_brute_force(oracle, L, up, k, p):
    """Representatives = orbit minima whose projected coefficient survives."""
    m = oracle.Model(L, [], up=up, momentum=k, parity=p)
    states, orbits = m.states()
    expect, exp_orb = [], []
    for s in range(2 ** L):
        if up is not None and bin(s).count("1") != up:
            continue
        group = []
        shifts = range(L) if k is not None else [0]
        for t in shifts:
            r = rotate(s, t, L)
            group.append((r, t, 1))
            if p is not None:
                group.append((reflect(r, L), t, p))
        if min(g[0] for g in group) != s:
            continue
        coeff = {}
        for r, t, sign in group:
            ch = np.exp(-2j * np.pi * (k or 0) * t / L)
            coeff[r] = coeff.get(r, 0) + sign * ch
        alive = {r: c for r, c in coeff.items() if abs(c) > 1e-9}
        if alive:
            expect.append(s)
            exp_orb.append(len(alive))
    np.testing.assert_array_equal(states, np.array(expect, dtype=np.uint64))
    np.testing.assert_array_equal(orbits, np.array(exp_orb))


# --------------------------------------------------- matrix elements -------

@pytest.mark.parametrize("name", ["heisenberg", "j1j2", "dm", "xyfield"])
def test_unreduced_matrix_is_kron_hamiltonian(oracle, name):
    L = 6
    spec = {"heisenberg": cases.heisenberg, "j1j2": cases.j1j2, "dm": cases.dm_chain, "xyfield": cases.xy_field}[name](L)
    m = oracle.Model(L, spec)
    np.testing.assert_allclose(m.dense(), full_hamiltonian(L, spec), atol=1e-14)
    if name != "xyfield":
        sub = oracle.Model(L, spec, up=3)
        idx = [s for s in range(2 ** L) if bin(s).count("1") == 3]
        np.testing.assert_allclose(sub.dense(), full_hamiltonian(L, spec)[np.ix_(idx, idx)], atol=1e-14)


def test_pauli_normalizations(oracle):
    L = 4
    spec = [("zz", [(0.7, [0, 1]), (0.3, [2, 3])]), ("x", [(0.4, [1])]), ("+-", [(0.2, [1, 2])]), ("-+", [(0.2, [1, 2])])]
    for norm in ("angular", "pauli", "pauli_cartesian"):
        m = oracle.Model(L, spec, norm=norm)
        np.testing.assert_allclose(m.dense(), full_hamiltonian(L, spec, norm), atol=1e-14)


@pytest.mark.parametrize("name,L,up", [("heisenberg", 8, 4), ("j1j2", 8, 4), ("dm", 8, 3), ("xyfield", 7, None)])
def test_momentum_sectors_union_is_full_spectrum(oracle, name, L, up):
    spec = {"heisenberg": cases.heisenberg, "j1j2": cases.j1j2, "dm": cases.dm_chain, "xyfield": cases.xy_field}[name](L)
    full = oracle.Model(L, spec, up=up)
    ref = np.linalg.eigvalsh(full.dense())
    parts = []
    for k in range(L):
        mk = oracle.Model(L, spec, up=up, momentum=k)
        if mk.dim:
            d = mk.dense()
            assert np.allclose(d, d.conj().T, atol=1e-13)
            parts.append(np.linalg.eigvalsh(d))
    np.testing.assert_allclose(np.sort(np.concatenate(parts)), ref, atol=1e-12)


@pytest.mark.parametrize("L,up,k", [(8, 4, 0), (8, 4, 4), (10, 5, 0), (9, None, 0)])
def test_parity_sectors_split_momentum_sector(oracle, L, up, k):
    spec = cases.j1j2(L)
    whole = np.linalg.eigvalsh(oracle.Model(L, spec, up=up, momentum=k).dense())
    parts = [np.linalg.eigvalsh(oracle.Model(L, spec, up=up, momentum=k, parity=p).dense()) for p in (1, -1)]
    np.testing.assert_allclose(np.sort(np.concatenate(parts)), whole, atol=1e-12)


def test_reduced_matrix_equals_projector_matrix(oracle):
    """H_k[i,j] = <ψ_i|H|ψ_j> with ψ_r built from explicit orbit vectors."""
    L, up, k = 8, 4, 3
    spec = cases.dm_chain(L)
    m = oracle.Model(L, spec, up=up, momentum=k)
    H = full_hamiltonian(L, spec)
    states, orbits = m.states()
    vecs = []
    for r in states:
        v = np.zeros(2 ** L, dtype=complex)
        for s in range(L):
            v[rotate(int(r), s, L)] += np.exp(-2j * np.pi * k * s / L)
        v /= np.linalg.norm(v)
        vecs.append(v)
    B = np.array(vecs).T  # columns = symmetric states
    # the reference's element H[s,r] = Σ_t a_t sqrt(N_r/N_s) conj(phase_t) with phase(T^s r) = χ(s) is
    # <s~|H|r~> for the symmetric states |r~> = N^-1/2 Σ_s χ(s) |T^s r>
    Hk = B.conj().T @ H @ B
    np.testing.assert_allclose(m.dense(), Hk, atol=1e-13)


# ----------------------------------------------------- solvers -------------

def test_deterministic_start_closed_form(oracle):
    # src/solve/mod.rs:729-742, evaluated with Python integers
    for seed in (0, 1, 12345, 2 ** 63 + 7):
        state = seed | 1
        vals = []
        for _ in range(50):
            state = (state * 6364136223846793005 + 1442695040888963407) % 2 ** 64
            vals.append((state >> 11) / 2 ** 53 - 0.5)
        v = np.array(vals)
        v /= np.linalg.norm(v)
        np.testing.assert_allclose(oracle.deterministic_start(50, seed), v, rtol=0, atol=1e-16)


def test_hermitian_eig_against_lapack(oracle):
    rng = np.random.default_rng(0)
    a = rng.standard_normal((40, 40)) + 1j * rng.standard_normal((40, 40))
    a = a + a.conj().T
    np.testing.assert_allclose(oracle.hermitian_eigvals(a), np.linalg.eigvalsh(a), atol=1e-11)


@pytest.mark.parametrize("L,up,k,name,nev", [(16, 8, 0, "heisenberg", 4), (12, 6, 3, "dm", 3), (12, 6, 6, "j1j2", 2)])
def test_oracle_eigsh_against_dense(oracle, L, up, k, name, nev):
    spec = {"heisenberg": cases.heisenberg, "j1j2": cases.j1j2, "dm": cases.dm_chain}[name](L)
    m = oracle.Model(L, spec, up=up, momentum=k)
    r = m.eigsh(nev)
    dense = np.linalg.eigvalsh(m.dense())
    np.testing.assert_allclose(r["eigenvalues"], dense[:nev], rtol=1e-10, atol=1e-10)
    assert all(res <= 1e-10 * max(abs(v), 1) for v, res in zip(r["eigenvalues"], r["residuals"]))
    # the reference's non-restarted window always runs in full (src/solve/mod.rs:1344-1377)
    if m.dim <= 256:  # dense route (src/solve/mod.rs:364-372)
        assert r["iterations"] == 1
    else:
        assert r["iterations"] == min(max(8 * nev + 64, 256), m.dim)


def test_oracle_evolve_against_expm(oracle):
    L = 10
    spec = cases.xxz(L, 1.0)
    m = oracle.Model(L, spec, up=5, momentum=0)
    H = m.dense()
    rng = np.random.default_rng(1)
    psi = rng.standard_normal(m.dim) + 1j * rng.standard_normal(m.dim)
    psi /= np.linalg.norm(psi)
    times = [0.1, 1.0, 10.0]
    out, stats = m.evolve(psi, times, krylov_dimension=30)
    for t, got in zip(times, out):
        exact = scipy.linalg.expm(-1j * t * H) @ psi
        assert np.linalg.norm(got - exact) <= 1e-9
    assert stats["matrix_vector_products"] >= stats["lanczos_projections"]


def test_oracle_taylor_against_expm(oracle):
    L = 8
    spec = cases.dm_chain(L)
    m = oracle.Model(L, spec, up=4, momentum=2)
    H = m.dense()
    rng = np.random.default_rng(2)
    v = rng.standard_normal(m.dim) + 1j * rng.standard_normal(m.dim)
    for coef in (-0.5j, 0.3, -2.0 + 1j):
        out, info = m.expm(coef, [v])
        exact = scipy.linalg.expm(coef * H) @ v
        assert np.linalg.norm(out[0] - exact) <= 1e-12 * np.linalg.norm(exact)
        # degree/scaling = argmin m*s over the theta table (brute force here)
        theta = [(1, 2.29e-16), (2, 2.58e-8), (3, 1.39e-5), (4, 3.40e-4), (5, 2.40e-3), (6, 9.07e-3), (7, 2.38e-2),
                 (8, 5.00e-2), (9, 8.96e-2), (10, 1.44e-1), (11, 2.14e-1), (12, 3.00e-1), (13, 4.00e-1),
                 (14, 5.14e-1), (15, 6.41e-1), (16, 7.81e-1), (17, 9.31e-1), (18, 1.09), (19, 1.26), (20, 1.44),
                 (21, 1.62), (22, 1.82), (23, 2.01), (24, 2.22), (25, 2.43), (26, 2.64), (27, 2.86), (28, 3.08),
                 (29, 3.31), (30, 3.54), (35, 4.7), (40, 6.0), (45, 7.2), (50, 8.5), (55, 9.9)]
        mu = np.trace(H) / m.dim
        norm1 = np.abs(H - mu * np.eye(m.dim)).sum(axis=0).max()
        assert abs(info["one_norm"] - norm1) <= 1e-12 * norm1
        x = abs(coef) * norm1
        best = min(theta, key=lambda t: t[0] * max(math.ceil(x / t[1]), 1))
        assert (info["degree"], info["scaling"]) == (best[0], max(math.ceil(x / best[1]), 1))


def test_oracle_nilpotent_taylor(oracle):
    # a single raising operator is nilpotent: exp(A)·|down> = |down> + |up>
    # (nilpotent Taylor anchor, tests/solver_extensions.rs:364-444)
    m = oracle.Model(1, [("+", [(1.0, [0])])], check_hermiticity=False, check_particles=False)
    out, info = m.expm(1.0, [np.array([1.0, 0.0])])
    np.testing.assert_allclose(out[0], [1.0, 1.0], atol=1e-15)


# ------------------------------------------------- thick restart / subspace --
# thick_restarted_eigsh (src/solve/mod.rs:1644-2196), eigsh_with_initial_subspace
# (:2385-2410), eigsh_with_workspace (:2419-2444), balanced_subspace_start
# (:708-727): pinned by dense LAPACK spectra, the reference's own control-flow
# anchors (tests/solver_extensions.rs:100-154, tests/visible_contract.rs:1041-1059)
# and the closed form of the balanced start.

TR_CASES = [
    # (L, spec, up, momentum, k, krylov, target)
    (16, cases.heisenberg(16), 8, 0, 1, 12, "smallest_algebraic"),
    (16, cases.heisenberg(16), 8, 0, 4, 24, "smallest_algebraic"),
    (16, cases.xxz(16, 1.5), 8, 1, 3, 20, "smallest_algebraic"),  # complex sector
    (16, cases.j1j2(16), 8, 0, 2, 40, "largest_algebraic"),
    (14, cases.heisenberg(14), 7, 2, 2, 16, "largest_magnitude"),
    (11, cases.dm_chain(11), None, 3, 2, 30, "both_ends"),
]


@pytest.mark.parametrize("case", range(len(TR_CASES)))
def test_oracle_thick_restart_matches_dense(oracle, case):
    L, spec, up, k_mom, k, kd, target = TR_CASES[case]
    m = oracle.Model(L, spec, up=up, momentum=k_mom)
    dense = np.linalg.eigvalsh(m.dense())
    r = m.eigsh(k, target=target, krylov_dimension=kd, tolerance=1e-10, vectors=True)
    if target == "smallest_algebraic":
        ref = dense[:k]
    elif target == "largest_algebraic":
        ref = dense[::-1][:k]
    elif target == "largest_magnitude":
        ref = dense[np.argsort(-np.abs(dense), kind="stable")][:k]
    else:
        lo, hi = k // 2, k - k // 2
        ref = np.sort(np.concatenate([dense[:lo], dense[len(dense) - hi:]]))
    np.testing.assert_allclose(r["eigenvalues"], ref, rtol=1e-10, atol=1e-12)
    H = m.dense()
    for v, lam, res in zip(r["eigenvectors"], r["eigenvalues"], r["residuals"]):
        assert abs(np.linalg.norm(v) - 1) < 1e-12
        true_res = np.linalg.norm(H @ v - lam * v)
        assert true_res <= 1e-10 * max(abs(lam), 1) * 1.01 + 1e-14
        assert abs(true_res - res) <= 1e-9 * max(abs(lam), 1)
    assert r["iterations"] > kd  # the window restarted
    assert r["iterations"] <= 1000


def test_oracle_thick_restart_diagonal_kats(oracle):
    # tests/solver_extensions.rs:100-123 on a 512-dimensional diagonal with
    # values i/511 (the reference's has 300 values i/299): a first solve
    # restarts, a workspace solve from its converged subspace stays within one
    # window, with the same eigenvalues
    m = oracle.Model(9, cases.diagonal_chain(9, 1.0 / 511))
    np.testing.assert_allclose(np.sort(np.diag(m.dense()).real), np.arange(512) / 511, atol=1e-15)
    first = m.eigsh_subspace(3, [], krylov_dimension=16, tolerance=1e-10, seed=19, workspace=True, vectors=True)
    assert first["iterations"] > 16
    np.testing.assert_allclose(first["eigenvalues"], [0, 1 / 511, 2 / 511], atol=1e-10)
    second = m.eigsh_subspace(3, first["eigenvectors"], krylov_dimension=16, tolerance=1e-10, seed=19,
                              workspace=True)
    assert second["iterations"] <= 16
    np.testing.assert_allclose(second["eigenvalues"], first["eigenvalues"], atol=1e-14)
    assert np.all(second["residuals"] <= 1e-10)
    # a small window that cannot converge answers NonConvergence
    # (tests/solver_extensions.rs:73-87: krylov 2, max_iterations 2)
    with pytest.raises(oracle.OracleError, match="did not converge"):
        m.eigsh(1, krylov_dimension=2, tolerance=1e-12, max_iterations=2, seed=7)


def test_oracle_full_window_iteration_count(oracle):
    # tests/visible_contract.rs:1041-1059: diagonal 0..255, three lowest,
    # window 160 (>= 96: one unrestarted window), 160 iterations
    m = oracle.Model(8, cases.diagonal_chain(8))
    r = m.eigsh(3, krylov_dimension=160, tolerance=1e-8, max_iterations=192, seed=17)
    np.testing.assert_allclose(r["eigenvalues"], [0, 1, 2], atol=1e-7)
    assert np.all(r["residuals"] < 1e-7)
    assert r["iterations"] == 160


def test_oracle_large_workspace_window_keeps_tridiagonal_path(oracle):
    # tests/solver_extensions.rs:126-154 shape: three isolated low values, a
    # window of 100: both solves take the unrestarted route (the second from
    # the balanced start of the first's three eigenvectors)
    L = 9
    m = oracle.Model(L, cases.isolated_low_chain(L))
    vals = np.diag(m.dense()).real
    np.testing.assert_allclose(np.sort(vals)[:5], [0, 1, 2, 3, 220], atol=1e-12)  # four low values, the rest far above
    first = m.eigsh_subspace(3, [], krylov_dimension=100, tolerance=1e-10, max_iterations=100, seed=23,
                             workspace=True, vectors=True)
    second = m.eigsh_subspace(3, first["eigenvectors"], krylov_dimension=100, tolerance=1e-10, max_iterations=100,
                              seed=23, workspace=True)
    np.testing.assert_allclose(second["eigenvalues"], np.sort(vals)[:3], atol=1e-10)
    np.testing.assert_allclose(first["eigenvalues"], second["eigenvalues"], atol=1e-12)


def test_oracle_balanced_start_closed_form(oracle):
    rng = np.random.default_rng(5)
    for J in (1, 2, 3, 5):
        sub = rng.standard_normal((J, 40)) + 1j * rng.standard_normal((J, 40))
        ref = sum(np.exp(2j * np.pi * j / J) / np.sqrt(J) * sub[j] for j in range(J))
        ref /= np.linalg.norm(ref)
        np.testing.assert_allclose(oracle.balanced_start(sub), ref, atol=1e-14)
    # the phases cancel: the sum of a unit vector and itself times e^{iπ}
    # is below f64::EPSILON, and the first vector is used
    v = rng.standard_normal(40)
    v /= np.linalg.norm(v)
    out = oracle.balanced_start(np.array([v, v]))
    np.testing.assert_allclose(out, v / np.linalg.norm(v), atol=1e-14)


def test_oracle_initial_subspace_route(oracle):
    # several start vectors take the thick restart and keep them: starting
    # from exact eigenvectors, the first window converges
    L = 16
    m = oracle.Model(L, cases.heisenberg(L), up=8, momentum=0)
    w, V = np.linalg.eigh(m.dense())
    r = m.eigsh_subspace(2, V[:, :3].T, tolerance=1e-10)
    np.testing.assert_allclose(r["eigenvalues"], w[:2], rtol=1e-12)
    assert r["iterations"] <= 2 * (4 * 2 + 24)
    # one vector is eigsh_with_initial (the expanding window)
    r1 = m.eigsh_subspace(2, V[:, :1].T, tolerance=1e-10)
    np.testing.assert_allclose(r1["eigenvalues"], w[:2], rtol=1e-10)


# ------------------------------------------- general symmetry sectors ------
# GeneralBasis + SymmetryReducer (QMBED src/basis/mod.rs:3330-3650,
# 4096-4507) as the C ABI builds it (bindings/capi/src/lib.rs:2849-2920,
# 3103-3129, 3412-3422).

def translation(L, step=1):
    return [(i + step) % L for i in range(L)]


def reflection(L):
    return [L - 1 - i for i in range(L)]


def general(dest, sector, flip=False):
    q = {"destinations": list(dest), "sector": sector}
    if flip:
        q["local_permutations"] = [[1, 0]] * len(dest)
    return q


def perm_matrix(L, dest, flip=False):
    """U|s> = |g s> on all 2^L states (bit i = site i)."""
    dim = 2 ** L
    U = np.zeros((dim, dim))
    for s in range(dim):
        t = 0
        for i in range(L):
            b = (s >> i) & 1
            if flip:
                b ^= 1
            t |= b << dest[i]
        U[t, s] = 1.0
    return U


def test_general_translation_matches_builtin_sector(oracle):
    # the reference's own check (bindings/capi/src/lib.rs:8073-8100): the
    # runtime translation map [1, 2, 3, 4, 5, 0] at sector 1 describes the
    # built-in momentum-1 basis of L=6 up=3
    spec = cases.heisenberg(6)
    g = oracle.Model(6, spec, up=3, symmetries=[general(translation(6), 1)])
    b = oracle.Model(6, spec, up=3, momentum=1)
    assert list(g.states()[0]) == [7, 11, 13] == list(b.states()[0])
    # and for every momentum of L=8, with equal reduced matrices
    spec = cases.dm_chain(8)
    for k in range(8):
        g = oracle.Model(8, spec, up=4, symmetries=[general(translation(8), k)])
        b = oracle.Model(8, spec, up=4, momentum=k)
        np.testing.assert_array_equal(g.states()[0], b.states()[0])
        np.testing.assert_array_equal(g.states()[1], b.states()[1])
        np.testing.assert_allclose(g.dense(), b.dense(), atol=1e-13)


@pytest.mark.parametrize("reverse", [False, True])
@pytest.mark.parametrize("L,up,zflip", [(8, 4, True), (7, 3, False), (8, None, True)])
def test_general_sectors_union_is_full_spectrum(oracle, L, up, zflip, reverse):
    # translation, reflection and (at half filling / the full space) spin
    # inversion: every sector combination together spans the parent space
    # (translation and reflection commute only at k = 0 and k = L/2, where a
    # reflection character splits the momentum sector; the other momenta
    # stay whole)
    spec = cases.j1j2(L)
    full = np.linalg.eigvalsh(oracle.Model(L, spec, up=up).dense())
    parts = []
    for k in range(L):
        for p in ((0, 1) if (2 * k) % L == 0 else (None,)):
            for z in ((0, 1) if zflip else (None,)):
                gens = [general(translation(L), k)]
                if p is not None:
                    gens.append(general(reflection(L), p))
                if z is not None:
                    gens.append(general(list(range(L)), z, flip=True))
                m = oracle.Model(L, spec, up=up, symmetries=gens, reverse=reverse)
                if m.dim:
                    d = m.dense()
                    assert np.allclose(d, d.conj().T, atol=1e-13)
                    parts.append(np.linalg.eigvalsh(d))
    np.testing.assert_allclose(np.sort(np.concatenate(parts)), full, atol=1e-11)


@pytest.mark.parametrize("reverse", [False, True])
@pytest.mark.parametrize("config", ["t3_dm", "t4_p_z_j1j2"])
def test_general_symmetric_states_carry_the_sector_characters(oracle, reverse, config):
    # |r~> = N^-1/2 Σ_t phase(t)|t> from the oracle's images: orthonormal,
    # U_g |r~> = exp(2πi q/p) |r~> for every generator g of period p and
    # sector q, and <r~_i|H|r~_j> is the reduced matrix
    L, up = 8, 4
    if config == "t3_dm":  # complex characters and couplings
        spec = cases.dm_chain(L)
        gens = [(translation(L), 3, False, 8)]
    else:
        spec = cases.j1j2(L)
        gens = [(translation(L), 4, False, 8), (reflection(L), 1, False, 2), (list(range(L)), 1, True, 2)]
    m = oracle.Model(L, spec, up=up, symmetries=[general(d, q, f) for d, q, f, _ in gens], reverse=reverse)
    states, orbits = m.states()
    assert m.dim > 0
    dim = 2 ** L
    B = np.zeros((dim, m.dim), dtype=complex)
    for t in range(dim):
        if bin(t).count("1") != up:
            continue
        img = m.image(t)
        if img is None:
            continue
        rep, phase, orb = img
        i = list(states).index(rep)
        assert orb == orbits[i]
        B[t, i] = phase / np.sqrt(orb)
    np.testing.assert_allclose(B.conj().T @ B, np.eye(m.dim), atol=1e-13)
    for d, q, f, p in gens:
        U = perm_matrix(L, d, f)
        np.testing.assert_allclose(U @ B, np.exp(2j * np.pi * q / p) * B, atol=1e-13)
    H = full_hamiltonian(L, spec)
    np.testing.assert_allclose(m.dense(), B.conj().T @ H @ B, atol=1e-13)


@pytest.mark.parametrize("reverse", [False, True])
def test_general_representatives_and_row_order(oracle, reverse):
    L, up = 10, 5
    gens = [general(translation(L), 0), general(reflection(L), 0), general(list(range(L)), 0, flip=True)]
    m = oracle.Model(L, cases.heisenberg(L), up=up, symmetries=gens, reverse=reverse)
    states = [int(s) for s in m.states()[0]]
    mask = (1 << L) - 1
    for r in states:
        orbit = {rotate(r, k, L) for k in range(L)}
        orbit |= {reflect(t, L) for t in orbit}
        orbit |= {t ^ mask for t in orbit}
        assert r == (max(orbit) if reverse else min(orbit))
    assert states == sorted(states, reverse=reverse)
    # without symmetries the reverse convention only reverses the rows
    plain = oracle.Model(L, cases.heisenberg(L), up=up, reverse=True)
    asc = oracle.Model(L, cases.heisenberg(L), up=up)
    np.testing.assert_array_equal(plain.states()[0], asc.states()[0][::-1])


def test_general_sector_errors(oracle):
    spec = cases.heisenberg(6)
    with pytest.raises(oracle.OracleError, match="expected 6"):
        oracle.Model(6, spec, up=3, symmetries=[general([1, 0], 0)])
    with pytest.raises(oracle.OracleError, match="bijection"):
        oracle.Model(6, spec, up=3, symmetries=[general([1, 1, 2, 3, 4, 5], 0)])
    with pytest.raises(oracle.OracleError, match="cannot be mixed"):
        oracle.Model(6, spec, up=3, momentum=1, symmetries=[general(translation(6), 1)])
