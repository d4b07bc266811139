"""Chains of 29 to 31 sites on 32-bit states: the longest narrow sectors,
where the state directory's cells cover 2^18 to 2^20 values each and its
sub-bucket shifts reach their widest. Few up spins keep the sectors small
enough for the oracle. For each H·v mode (orbit walk, image table, compiled
stream) the product matches the oracle's (FP64 ≤ 1e-13 relative), and the
representatives and orbit sizes are bit-exact."""

import numpy as np
import pytest

from tests import cases

pytestmark = pytest.mark.gpu

# (L, up, momentum, parity)
SECTORS = [
    (29, 3, 0, None), (29, 4, 5, None), (30, 4, 0, 1), (30, 3, 15, -1),
    (31, 5, 0, None), (31, 4, 7, None), (31, 3, None, 1), (30, 2, None, None),
]

HAMS = {"heisenberg": cases.heisenberg, "j1j2": cases.j1j2, "dm": cases.dm_chain}


def basis_json(L, up, k, p):
    b = {"kind": "spin", "sites": L, "up": up}
    if k is not None:
        b["momentum"] = k
    if p is not None:
        b["parity"] = p
    return b


@pytest.mark.parametrize("mode", ["walk", "image", "stream"])
@pytest.mark.parametrize("ham", sorted(HAMS))
@pytest.mark.parametrize("L,up,k,p", SECTORS, ids=[f"L{s[0]}_up{s[1]}_k{s[2]}_p{s[3]}" for s in SECTORS])
def test_long_narrow_chain_hv(qa, oracle, monkeypatch, L, up, k, p, ham, mode):
    if ham == "dm" and p is not None:
        pytest.skip("the DM term is odd under reflection")
    monkeypatch.setenv("QMBED_HV_MODE", mode)
    print(f"[long chain] L={L} up={up} k={k} p={p} {ham} {mode}", flush=True)
    spec = HAMS[ham](L)
    om = oracle.Model(L, spec, up=up, momentum=k, parity=p)
    if om.dim == 0:
        pytest.skip("empty sector")
    m = qa.Model(basis_json(L, up, k, p), spec)
    assert m.dim == om.dim
    assert m.hv_config()["mode"] == mode
    states, orbits = om.states()
    s, o = m.states()
    np.testing.assert_array_equal(s.astype(np.uint64), states)
    np.testing.assert_array_equal(o.astype(np.int32), orbits)
    x = cases.rand_vec(om.dim, seed=L + up)
    y_ref = om.apply(x)
    y = m.apply(x.astype(np.complex128))
    assert np.linalg.norm(y - y_ref) <= 1e-13 * max(np.linalg.norm(y_ref), 1e-300)
