"""Writes tests/golden/*.json with the CPU oracle (run in the build container;
the GPU tests only read the JSON).

Full-size sectors are too large to store vectors for, so each fixture keeps
size-independent checksums: the dimension, the exact stored-entry count, a
hash of the sorted representative list and of the orbit sizes, and the
bilinear forms w_k · (H x) for seeded vectors x, w_k (plus ‖H x‖). A wrong
index, phase or amplitude anywhere in H changes these at O(1/√N) relative,
far above the 1e-12 comparison tolerance.

    python tests/golden/make_golden.py [c1 c2 c3 c5 ...]
"""

import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle_ffi as O  # noqa: E402
from tests import cases  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def seeded(n, seed, complex_=False):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(n)
    if complex_:
        v = v + 1j * rng.standard_normal(n)
    return v


def fingerprint(model, nforms=4, complex_=False):
    states, orbits = model.states()
    x = seeded(model.dim, 1234, complex_).astype(np.complex128)
    y = model.apply(x)
    forms = []
    for k in range(nforms):
        w = seeded(model.dim, 100 + k, complex_)
        f = complex(np.dot(w, y))
        forms.append([f.real, f.imag])
    return {
        "dimension": int(model.dim),
        "nnz": int(model.nnz),
        "states_sha256": hashlib.sha256(states.astype("<u4").tobytes()).hexdigest(),
        "orbits_sha256": hashlib.sha256(orbits.astype("u1").tobytes()).hexdigest(),
        "x_seed": 1234,
        "w_seeds": [100 + k for k in range(nforms)],
        "complex_vectors": complex_,
        "forms": forms,
        "hx_norm": float(np.linalg.norm(y)),
    }


CASES = {
    "c1": dict(L=16, spec=lambda: cases.heisenberg(16), up=8, k=0, p=None, norm="angular", eig=4),
    "c2": dict(L=24, spec=lambda: cases.xxz(24, 1.5), up=12, k=0, p=None, norm="angular", eig=6),
    # two of the 72 XXZ L=24 sectors of the batch: up=11 k=5 (complex) and up=12 k=0
    "c2k5": dict(L=24, spec=lambda: cases.xxz(24, 1.0), up=11, k=5, p=None, norm="angular", eig=6),
    "c4k0": dict(L=24, spec=lambda: cases.xxz(24, 1.0), up=12, k=0, p=None, norm="angular", eig=6),
    "c3": dict(L=28, spec=lambda: cases.j1j2(28), up=14, k=0, p=None, norm="angular", eig=0),
    # three points of the L=26 Ising scan family: an interior point, the
    # headline's H·v point and one next to the critical line (hx = 1, hz -> 0)
    "c5": dict(L=26, spec=lambda: cases.ising(26, 0.7, 0.2), up=None, k=0, p=1, norm="pauli_cartesian", eig=1),
    "c5a": dict(L=26, spec=lambda: cases.ising(26, 1.0, 0.25), up=None, k=0, p=1, norm="pauli_cartesian", eig=1),
    "c5b": dict(L=26, spec=lambda: cases.ising(26, 1.0, 0.02), up=None, k=0, p=1, norm="pauli_cartesian", eig=1),
}


def make(name):
    c = CASES[name]
    t0 = time.time()
    m = O.Model(c["L"], c["spec"](), up=c["up"], momentum=c["k"], parity=c["p"], norm=c["norm"])
    out = {"case": name, "sites": c["L"], "up": c["up"], "momentum": c["k"], "parity": c["p"],
           "normalization": c["norm"], "terms": c["spec"](), "built_s": time.time() - t0}
    out.update(fingerprint(m, complex_=not m.is_real))
    if c["eig"]:
        t1 = time.time()
        r = m.eigsh(c["eig"])
        out["eigenvalues"] = [float(v) for v in r["eigenvalues"]]
        out["residuals"] = [float(v) for v in r["residuals"]]
        out["eigsh_iterations"] = r["iterations"]
        out["eigsh_s"] = time.time() - t1
    with open(os.path.join(HERE, f"{name}.json"), "w") as f:
        json.dump(out, f, indent=1, default=lambda z: [z.real, z.imag] if isinstance(z, complex) else z)
    print(name, "done", time.time() - t0, "s", flush=True)


if __name__ == "__main__":
    for n in sys.argv[1:] or list(CASES):
        make(n)
