"""Seeded random sectors and couplings: chains of 6 to 22 sites with a random
up count (or none), momentum and parity, and a random XXZ chain with
next-nearest-neighbour hops, a longitudinal field and (for complex
momenta) a Dzyaloshinskii-Moriya term. For each H·v mode (orbit walk with
the state directory, image table, compiled stream) the product matches the
oracle's COO product (FP64 ≤ 1e-13 relative) and the representatives are
bit-exact."""

import numpy as np
import pytest

from tests import cases

pytestmark = pytest.mark.gpu


def random_case(seed):
    rng = np.random.default_rng(1000 + seed)
    L = int(rng.integers(6, 23))
    up = None if rng.random() < 0.25 and L <= 16 else int(rng.integers(1, L))
    k = None if rng.random() < 0.2 else int(rng.integers(0, L))
    p = None
    if rng.random() < 0.35:  # parity sectors: no momentum, k = 0 or k = L/2
        p = int(rng.choice([-1, 1]))
        k = [None, 0, L // 2 if L % 2 == 0 else 0][int(rng.integers(0, 3))]
    complex_ok = k is not None and k != 0 and 2 * k != L and p is None
    b1 = cases.bonds(L, 1)
    b2 = cases.bonds(L, 2)
    jz, jxy, j2, h = rng.uniform(-1.5, 1.5, 4)
    spec = [("zz", [(float(jz), x) for x in b1]), ("+-", [(0.5 * float(jxy), x) for x in b1]),
            ("-+", [(0.5 * float(jxy), x) for x in b1]), ("+-", [(0.5 * float(j2), x) for x in b2]),
            ("-+", [(0.5 * float(j2), x) for x in b2])]
    if up is not None:
        spec.append(("z", [(float(h), [i]) for i in range(L)]))
    if complex_ok and rng.random() < 0.5:
        d = float(rng.uniform(0.1, 0.6))
        spec.append(("+-", [(complex(0.0, d), x) for x in b1]))
        spec.append(("-+", [(complex(0.0, -d), x) for x in b1]))
    return L, up, k, p, spec


CASES = list(range(40))


@pytest.mark.parametrize("mode", ["walk", "image", "stream"])
@pytest.mark.parametrize("seed", CASES)
def test_random_sector_hv(qa, oracle, monkeypatch, seed, mode):
    L, up, k, p, spec = random_case(seed)
    monkeypatch.setenv("QMBED_HV_MODE", mode)
    print(f"[random sector] seed={seed} L={L} up={up} k={k} p={p} {mode}", flush=True)
    om = oracle.Model(L, spec, up=up, momentum=k, parity=p)
    if om.dim == 0:
        pytest.skip("empty sector")
    b = {"kind": "spin", "sites": L}
    if up is not None:
        b["up"] = up
    if k is not None:
        b["momentum"] = k
    if p is not None:
        b["parity"] = p
    m = qa.Model(b, spec)
    assert m.dim == om.dim
    states, _ = om.states()
    s, _ = m.states()
    np.testing.assert_array_equal(s.astype(np.uint64), states)
    x = cases.rand_vec(om.dim, seed=seed)
    y_ref = om.apply(x)
    y = m.apply(x.astype(np.complex128))
    assert np.linalg.norm(y - y_ref) <= 1e-13 * max(np.linalg.norm(y_ref), 1e-300)
