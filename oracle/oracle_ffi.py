"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

The oracle (``oracle/qmbed_oracle.cpp``) is a single-threaded FP64
restatement of the reference algorithms.  Only the tests, the smoke check and
the CPU-baseline leg of ``bench.py`` load it, and only as the checker; the
product library never does.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

TARGETS = {
    "smallest_algebraic": 0,
    "largest_algebraic": 1,
    "smallest_magnitude": 2,
    "largest_magnitude": 3,
    "both_ends": 4,
}
NORMS = {"angular": 0, "pauli": 1, "pauli_cartesian": 2}


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        c_int, c_long, c_double, c_void_p = ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_void_p
        P = ctypes.POINTER
        L.oracle_last_error.restype = ctypes.c_char_p
        L.oracle_model_new.restype = c_void_p
        L.oracle_model_new.argtypes = [c_int, c_int, P(c_int), c_int, c_int, c_int, c_int, c_int,
                                       ctypes.c_char_p, c_int, c_int]
        L.oracle_model_new_general.restype = c_void_p
        L.oracle_model_new_general.argtypes = [c_int, c_int, P(c_int), c_int, c_int, c_int, c_int, c_int,
                                               ctypes.c_char_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                               c_int]
        L.oracle_model_free.argtypes = [c_void_p]
        L.oracle_dim.restype = c_long
        L.oracle_dim.argtypes = [c_void_p]
        L.oracle_nnz.restype = c_long
        L.oracle_nnz.argtypes = [c_void_p]
        L.oracle_is_real.argtypes = [c_void_p]
        L.oracle_states.argtypes = [c_void_p, c_void_p, c_void_p]
        L.oracle_image.argtypes = [c_void_p, ctypes.c_ulonglong, P(ctypes.c_ulonglong), P(c_double),
                                   P(c_double), P(c_int)]
        L.oracle_triplets.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]
        L.oracle_apply.argtypes = [c_void_p, c_void_p, c_void_p]
        L.oracle_apply_real.argtypes = [c_void_p, c_void_p, c_void_p]
        L.oracle_deterministic_start.argtypes = [c_long, ctypes.c_ulonglong, c_void_p]
        L.oracle_eigsh.argtypes = [c_void_p, c_int, c_int, c_long, c_double, c_long, ctypes.c_ulonglong,
                                   c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]
        L.oracle_evolve.argtypes = [c_void_p, c_void_p, c_void_p, c_int, c_long, c_double, c_long,
                                    c_void_p, c_void_p]
        L.oracle_expm.argtypes = [c_void_p, c_double, c_double, c_int, c_double, c_long, c_int, c_void_p,
                                  c_void_p, c_void_p]
        L.oracle_hermitian_eigvals.argtypes = [c_int, c_void_p, c_void_p]
        L.oracle_eigsh_subspace.argtypes = [c_void_p, c_int, c_int, c_long, c_double, c_long, ctypes.c_ulonglong,
                                            c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]
        L.oracle_balanced_start.argtypes = [c_int, c_long, c_void_p, c_void_p]
        _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def terms_text(terms):
    """terms: list of (symbols, [(coeff, [sites...]), ...])."""
    lines = []
    for symbols, couplings in terms:
        for coeff, sites in couplings:
            c = complex(coeff)
            lines.append(f"{symbols} {c.real!r} {c.imag!r} " + " ".join(str(s) for s in sites))
    return "\n".join(lines)


class Model:
    """Oracle sector + assembled operator (row-sorted triplets)."""

    def __init__(self, sites, terms=(), up=None, up_sectors=None, momentum=None, parity=None,
                 norm="angular", check_hermiticity=True, check_particles=True, symmetries=None, reverse=False):
        """symmetries: the general sector's generators, as the C ABI's
        requests ({"destinations": [...], "local_permutations": [[a, b], ...]
        or None, "sector": q}); reverse: maximum representatives and
        descending rows."""
        L = lib()
        secs = list(up_sectors or [])
        arr = (ctypes.c_int * max(len(secs), 1))(*secs) if secs else None
        args = (sites, -1 if up is None else up, arr, len(secs),
                0 if momentum is None else 1, 0 if momentum is None else momentum,
                0 if parity is None else parity, NORMS[norm],
                terms_text(terms).encode(), int(check_hermiticity), int(check_particles))
        if symmetries or reverse:
            syms = list(symmetries or [])
            dest = np.zeros((max(len(syms), 1), sites), dtype=np.int32)
            local = np.zeros((max(len(syms), 1), sites, 2), dtype=np.int32)
            local[:, :, 1] = 1
            sectors = np.zeros(max(len(syms), 1), dtype=np.int32)
            for g, q in enumerate(syms):
                d = list(q["destinations"])
                if len(d) != sites:
                    raise OracleError(f"invalid options: symmetry map has {len(d)} sites, expected {sites}")
                dest[g] = d
                if q.get("local_permutations") is not None:
                    local[g] = np.asarray(q["local_permutations"], dtype=np.int32)
                sectors[g] = int(q["sector"])
            h = L.oracle_model_new_general(*args, len(syms), _ptr(dest), _ptr(local), _ptr(sectors), int(reverse))
        else:
            h = L.oracle_model_new(*args)
        if not h:
            raise OracleError(L.oracle_last_error().decode())
        self._h = h
        self.dim = L.oracle_dim(h)
        self.nnz = L.oracle_nnz(h)
        self.is_real = bool(L.oracle_is_real(h))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.oracle_model_free(h)
            self._h = None

    def states(self):
        s = np.zeros(self.dim, dtype=np.uint64)
        o = np.zeros(self.dim, dtype=np.int32)
        lib().oracle_states(self._h, _ptr(s), _ptr(o))
        return s, o

    def image(self, state):
        rep = ctypes.c_ulonglong()
        re, im = ctypes.c_double(), ctypes.c_double()
        orb = ctypes.c_int()
        ok = lib().oracle_image(self._h, state, ctypes.byref(rep), ctypes.byref(re), ctypes.byref(im),
                                ctypes.byref(orb))
        if not ok:
            return None
        return int(rep.value), complex(re.value, im.value), int(orb.value)

    def triplets(self):
        n = self.nnz
        rows = np.zeros(n, dtype=np.int64)
        cols = np.zeros(n, dtype=np.int64)
        re = np.zeros(n)
        im = np.zeros(n)
        lib().oracle_triplets(self._h, _ptr(rows), _ptr(cols), _ptr(re), _ptr(im))
        return rows, cols, re + 1j * im

    def dense(self):
        r, c, v = self.triplets()
        a = np.zeros((self.dim, self.dim), dtype=complex)
        a[r, c] = v
        return a

    def apply(self, x):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.zeros(self.dim, dtype=np.complex128)
        lib().oracle_apply(self._h, _ptr(x), _ptr(y))
        return y

    def eigsh(self, k, target="smallest_algebraic", krylov_dimension=None, tolerance=1e-10,
              max_iterations=1000, seed=0, initial=None, vectors=False):
        vals = np.zeros(k)
        res = np.zeros(k)
        vecs = np.zeros((k, self.dim), dtype=np.complex128) if vectors else None
        info = np.zeros(3, dtype=np.int64)
        init = None if initial is None else np.ascontiguousarray(initial, dtype=np.complex128)
        rc = lib().oracle_eigsh(self._h, k, TARGETS[target], krylov_dimension or 0, tolerance, max_iterations,
                                seed, _ptr(init), _ptr(vals), _ptr(res), _ptr(vecs), _ptr(info))
        if rc != 0:
            raise OracleError(lib().oracle_last_error().decode())
        return {"eigenvalues": vals, "residuals": res, "eigenvectors": vecs, "iterations": int(info[0]),
                "reorthogonalization_passes": int(info[1]), "conditional_second_passes": int(info[2])}

    def eigsh_subspace(self, k, subspace, target="smallest_algebraic", krylov_dimension=None, tolerance=1e-10,
                       max_iterations=1000, seed=0, workspace=False, vectors=False):
        """eigsh_with_initial_subspace (workspace=False) or eigsh_with_workspace
        on a workspace holding `subspace` (workspace=True)."""
        sub = np.ascontiguousarray(np.atleast_2d(subspace), dtype=np.complex128) if len(subspace) else \
            np.zeros((0, self.dim), dtype=np.complex128)
        vals = np.zeros(k)
        res = np.zeros(k)
        vecs = np.zeros((k, self.dim), dtype=np.complex128) if vectors else None
        info = np.zeros(3, dtype=np.int64)
        rc = lib().oracle_eigsh_subspace(self._h, k, TARGETS[target], krylov_dimension or 0, tolerance,
                                         max_iterations, seed, sub.shape[0], _ptr(sub), 1 if workspace else 0,
                                         _ptr(vals), _ptr(res), _ptr(vecs), _ptr(info))
        if rc != 0:
            raise OracleError(lib().oracle_last_error().decode())
        return {"eigenvalues": vals, "residuals": res, "eigenvectors": vecs, "iterations": int(info[0]),
                "reorthogonalization_passes": int(info[1]), "conditional_second_passes": int(info[2])}

    def evolve(self, psi, times, krylov_dimension=30, tolerance=1e-10, max_substeps=10000):
        psi = np.ascontiguousarray(psi, dtype=np.complex128)
        times = np.ascontiguousarray(times, dtype=np.float64)
        out = np.zeros((len(times), self.dim), dtype=np.complex128)
        stats = np.zeros(5)
        rc = lib().oracle_evolve(self._h, _ptr(psi), _ptr(times), len(times), krylov_dimension, tolerance,
                                 max_substeps, _ptr(out), _ptr(stats))
        if rc != 0:
            raise OracleError(lib().oracle_last_error().decode())
        return out, {"lanczos_projections": int(stats[0]), "matrix_vector_products": int(stats[1]),
                     "accepted_substeps": int(stats[2]), "rejected_trial_intervals": int(stats[3]),
                     "maximum_estimated_error": float(stats[4])}

    def expm(self, coefficient, vectors, max_degree=55, tolerance=2.0 ** -53, max_substeps=10000):
        v = np.array(vectors, dtype=np.complex128, order="C", ndmin=2).copy()
        info = np.zeros(2, dtype=np.int64)
        dinfo = np.zeros(3)
        c = complex(coefficient)
        rc = lib().oracle_expm(self._h, c.real, c.imag, max_degree, tolerance, max_substeps, v.shape[0], _ptr(v),
                               _ptr(info), _ptr(dinfo))
        if rc != 0:
            raise OracleError(lib().oracle_last_error().decode())
        return v, {"degree": int(info[0]), "scaling": int(info[1]), "shift": complex(dinfo[0], dinfo[1]),
                   "one_norm": float(dinfo[2])}


def deterministic_start(n, seed=0):
    out = np.zeros(n)
    lib().oracle_deterministic_start(n, seed, _ptr(out))
    return out


def balanced_start(subspace):
    sub = np.ascontiguousarray(np.atleast_2d(subspace), dtype=np.complex128)
    out = np.zeros(sub.shape[1], dtype=np.complex128)
    rc = lib().oracle_balanced_start(sub.shape[0], sub.shape[1], _ptr(sub), _ptr(out))
    if rc != 0:
        raise OracleError(lib().oracle_last_error().decode())
    return out


def hermitian_eigvals(a):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    n = a.shape[0]
    v = np.zeros(n)
    lib().oracle_hermitian_eigvals(n, _ptr(a), _ptr(v))
    return v


def heisenberg_terms(L, jz=1.0, jxy=0.5, nn=1, periodic=True):
    bonds = [(i, (i + nn) % L) for i in range(L if periodic else L - nn)]
    return [("zz", [(jz, list(b)) for b in bonds]),
            ("+-", [(jxy, list(b)) for b in bonds]),
            ("-+", [(jxy, list(b)) for b in bonds])]
