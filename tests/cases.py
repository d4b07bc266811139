"""Shared Hamiltonians and sectors for the parity tests (spec lists of
(symbols, [(coefficient, sites), ...]) understood by both the oracle and the
library's JSON terms)."""

import numpy as np


def bonds(L, nn=1):
    return [[i, (i + nn) % L] for i in range(L)]


def heisenberg(L, jz=1.0, jxy=0.5, nn=1):
    b = bonds(L, nn)
    return [("zz", [(jz, x) for x in b]), ("+-", [(jxy, x) for x in b]), ("-+", [(jxy, x) for x in b])]


def xxz(L, delta):
    return heisenberg(L, jz=delta, jxy=0.5)


def j1j2(L, j1=1.0, j2=0.5):
    return heisenberg(L, j1, 0.5 * j1, 1) + heisenberg(L, j2, 0.5 * j2, 2)


def ising(L, hx, hz):
    out = [("zz", [(-1.0, x) for x in bonds(L)])]
    if hx:
        out.append(("x", [(-hx, [i]) for i in range(L)]))
    if hz:
        out.append(("z", [(-hz, [i]) for i in range(L)]))
    return out


def dm_chain(L, d=0.3, jz=0.7):
    """XXZ plus a Dzyaloshinskii-Moriya term: complex Hermitian couplings."""
    b = bonds(L)
    return [("zz", [(jz, x) for x in b]),
            ("+-", [(complex(0.5, d), x) for x in b]),
            ("-+", [(complex(0.5, -d), x) for x in b])]


def xy_field(L, hy=0.4):
    """Cartesian xx + yy couplings and a y field (y symbols, complex amplitudes)."""
    b = bonds(L)
    return [("xx", [(1.0, x) for x in b]), ("yy", [(0.8, x) for x in b]), ("y", [(hy, [i]) for i in range(L)]),
            ("z", [(0.3, [i]) for i in range(L)])]


def rand_vec(n, seed, complex_=True):
    rng = np.random.default_rng(seed)
    if complex_:
        return rng.standard_normal(n) + 1j * rng.standard_normal(n)
    return rng.standard_normal(n)


# (L, up, momentum, parity) sectors covering every symmetry kind
SECTORS = [
    (6, 3, None, None), (6, None, None, None), (8, 4, 0, None), (8, 4, 1, None), (8, 4, 4, None),
    (8, 3, 2, None), (8, 4, None, 1), (8, 4, None, -1), (8, 4, 0, 1), (8, 4, 0, -1), (8, 4, 4, 1),
    (8, 4, 4, -1), (10, 5, 3, None), (10, None, 0, 1), (12, 6, 0, None), (12, 6, 5, None), (12, 6, 6, -1),
    (12, 5, 0, 1), (14, 7, 0, None), (14, None, 7, None), (16, 8, 0, None), (9, 4, 2, None), (9, None, 0, -1),
]


def seeded_state(n, seed):
    """Normalized complex Gaussian state (g_re + i g_im)/‖g‖, g ~ N(0, 1) from
    numpy's default generator seeded with `seed` (the 64 evolution states of
    the batched-expm workload use seed = column index)."""
    rng = np.random.default_rng(seed)
    g = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    return g / np.linalg.norm(g)


def diagonal_chain(L, scale=1.0, offset=None):
    """H = Σ_i scale·2^i z_i + offset (angular-momentum z = ±1/2, no
    symmetry): on the 2^L computational states, state s has energy
    scale·(s − (2^L − 1)/2) + offset, so the sector matrix is the diagonal
    operator of the reference's solver tests (value index·scale with
    offset = scale·(2^L − 1)/2, the default)."""
    if offset is None:
        offset = scale * (2 ** L - 1) / 2
    return [("z", [(scale * 2.0 ** i, [i]) for i in range(L)]), ("I", [(offset, [0])])]


def isolated_low_chain(L):
    """Diagonal H = z_0 + 2 z_1 + Σ_{i≥2} (200 + 10 i) z_i + offset: the
    values 0, 1, 2, 3 and every other value at least 100 above them, the
    shape of the reference's isolated-low-values solver test."""
    w = [1.0, 2.0] + [200.0 + 10.0 * i for i in range(2, L)]
    return [("z", [(c, [i]) for i, c in enumerate(w)]), ("I", [(sum(w) / 2, [0])])]
