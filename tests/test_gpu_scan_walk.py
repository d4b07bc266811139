"""The batched scan H·v (y_b = H(c_b) x_b over up to 16 points of a family,
the kernel of lockstep scan solves) in its two forms: reading the compiled
index stream, and on the fly (QMBED_SCAN_HV=walk: per target the orbit walk
and the state-directory lookup, shared by the points). Both add the same
terms in the same order with the same factors, so their outputs must agree
bit for bit; each point's column is also checked against the oracle's
sector matrix at the same couplings (≤ 1e-13 relative, the FP32 form on the
inputs rounded to FP32), for every symmetry kind of the walk (none,
translation, reflection, both) and fixed and free magnetization."""

import os

import numpy as np
import pytest

from tests import cases

pytestmark = pytest.mark.gpu


def ising_family(L):
    static = [("zz", [(-1.0, b) for b in cases.bonds(L)])]
    comps = [("hx", [("x", [(-1.0, [i]) for i in range(L)])], None),
             ("hz", [("z", [(-1.0, [i]) for i in range(L)])], None)]

    def spec(c):
        return cases.ising(L, c[0], c[1])

    return static, comps, spec


def j1j2_family(L):
    # H = Σ_nn (S·S) + J2 Σ_nnn (S·S) + Δ Σ_nn SzSz
    static = cases.heisenberg(L, 1.0, 0.5, 1)
    comps = [("j2", cases.heisenberg(L, 1.0, 0.5, 2), None),
             ("delta", [("zz", [(1.0, b) for b in cases.bonds(L)])], None)]

    def spec(c):
        out = cases.heisenberg(L, 1.0, 0.5, 1) + cases.heisenberg(L, c[0], 0.5 * c[0], 2)
        out.append(("zz", [(c[1], b) for b in cases.bonds(L)]))
        return out

    return static, comps, spec


# (name, family, L, up, momentum, parity, pauli)
CASES = [
    ("ising_k0p1", ising_family, 16, None, 0, 1, True),
    ("ising_k0", ising_family, 14, None, 0, None, True),
    ("ising_p1", ising_family, 12, None, None, 1, True),
    ("ising_plain", ising_family, 10, None, None, None, True),
    ("j1j2_k0", j1j2_family, 18, 9, 0, None, False),
    ("j1j2_k0p-1", j1j2_family, 16, 8, 0, -1, False),
    ("j1j2_up7_p1", j1j2_family, 14, 7, None, 1, False),
]


def basis_json(L, up, k, p, pauli):
    b = {"kind": "spin", "sites": L, "pauli": pauli}
    if up is not None:
        b["up"] = up
    if k is not None:
        b["momentum"] = k
    if p is not None:
        b["parity"] = p
    return b


def model_in_mode(qa, basis, static, comps, walk):
    """A family model whose first batched scan launch (which fixes the
    form) happens with QMBED_SCAN_HV set as asked."""
    m = qa.Model(basis, static, components=comps)
    old = os.environ.pop("QMBED_SCAN_HV", None)
    if walk:
        os.environ["QMBED_SCAN_HV"] = "walk"
    try:
        m.apply_scan_batch(np.zeros((1, 2)), np.zeros((1, m.dim)))
    finally:
        os.environ.pop("QMBED_SCAN_HV", None)
        if old is not None:
            os.environ["QMBED_SCAN_HV"] = old
    return m


@pytest.mark.parametrize("fp32", [False, True], ids=["fp64", "fp32"])
@pytest.mark.parametrize("name,family,L,up,k,p,pauli", CASES, ids=[c[0] for c in CASES])
def test_scan_walk_equals_stream_and_oracle(qa, oracle, name, family, L, up, k, p, pauli, fp32):
    static, comps, spec = family(L)
    basis = basis_json(L, up, k, p, pauli)
    print(f"[scan walk] {name} fp32={fp32}", flush=True)
    ms = model_in_mode(qa, basis, static, comps, walk=False)
    mw = model_in_mode(qa, basis, static, comps, walk=True)
    qa.launch_timer(True)
    rng = np.random.default_rng(L)
    for width in (16, 8, 1):
        pts = np.column_stack([rng.uniform(0.2, 2.0, width), rng.uniform(-0.5, 0.5, width)])
        x = rng.standard_normal((width, ms.dim))
        ys = ms.apply_scan_batch(pts, x, fp32=fp32)
        yw = mw.apply_scan_batch(pts, x, fp32=fp32)
        assert np.array_equal(ys, yw), f"width {width}: max diff {np.max(np.abs(ys - yw))}"
        xin = x.astype(np.float32).astype(np.float64) if fp32 else x
        for b in (0, width - 1):
            om = oracle.Model(L, spec(pts[b]), up=up, momentum=k, parity=p,
                              norm="pauli_cartesian" if pauli else "angular")
            yr = om.apply(xin[b]).real
            assert np.linalg.norm(yw[b] - yr) <= 1e-13 * np.linalg.norm(yr)
    report = qa.launch_timer_report()
    qa.launch_timer(False)
    assert any(k.startswith("hv_scan_walk_kernel") for k in report), report
