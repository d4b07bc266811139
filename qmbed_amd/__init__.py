"""qmbed_amd — MI355X-native spin-sector H·v, eigsh and Krylov exponential engine.

Host-side mirror of the reference's Python FFI (QMBED
bindings/python/src/qmbed/_ffi.py:154-215): ``run`` sends one solve request
through ``qmbed_run_json`` and ``command`` one tagged operation through
``qmbed_command_json``; both return the decoded ``result`` or raise
:class:`QmbedError` with the library's error text.  :class:`Model` wraps the
typed device fast path (``qmbed_hip_*``).

The native library ``libqmbed_hip.so`` is built in-tree by
``qmbed_amd/build.py`` (``__graft_entry__.build()``).  There is no CPU
fallback: importing the bindings without the library raises.
"""

from __future__ import annotations

import ctypes
import json
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.environ.get("QMBED_HIP_LIBRARY", os.path.join(_HERE, "libqmbed_hip.so"))

F64, F32, C128, C64 = 0, 1, 2, 3
_DTYPES = {np.dtype(np.float64): F64, np.dtype(np.float32): F32,
           np.dtype(np.complex128): C128, np.dtype(np.complex64): C64}

_lib = None


class QmbedError(RuntimeError):
    """Error text returned by the native library (reference QmbedError display)."""


def library():
    """Load ``libqmbed_hip.so``; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBRARY_PATH):
        raise RuntimeError(
            f"qmbed_amd native library not found at {LIBRARY_PATH}; build it with "
            "`python qmbed_amd/build.py` (or __graft_entry__.build())")
    L = ctypes.CDLL(LIBRARY_PATH)
    c_char_p, c_void_p, c_int, c_long = ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_long
    for name in ("qmbed_run_json", "qmbed_command_json"):
        f = getattr(L, name)
        f.restype = c_void_p
        f.argtypes = [c_char_p]
    L.qmbed_string_free.argtypes = [c_void_p]
    L.qmbed_hip_basis_archive_save.restype = c_void_p
    L.qmbed_hip_basis_archive_save.argtypes = [c_char_p, c_void_p, ctypes.c_size_t]
    L.qmbed_hip_basis_archive_load.restype = c_void_p
    L.qmbed_hip_basis_archive_load.argtypes = [c_char_p, c_void_p, ctypes.c_size_t, c_int]
    L.qmbed_string_free.restype = None
    L.qmbed_hip_model_create.restype = c_void_p
    L.qmbed_hip_model_create.argtypes = [c_char_p, ctypes.POINTER(c_void_p)]
    L.qmbed_hip_model_free.argtypes = [c_void_p]
    L.qmbed_hip_model_dimension.restype = c_long
    L.qmbed_hip_model_dimension.argtypes = [c_void_p]
    L.qmbed_hip_model_is_real.argtypes = [c_void_p]
    L.qmbed_hip_model_states.argtypes = [c_void_p, c_void_p, c_void_p]
    L.qmbed_hip_model_states64.argtypes = [c_void_p, c_void_p, c_void_p]
    L.qmbed_hip_model_nnz.restype = c_long
    L.qmbed_hip_model_nnz.argtypes = [c_void_p]
    L.qmbed_hip_model_hv_config.argtypes = [c_void_p, ctypes.POINTER(c_int), ctypes.POINTER(c_int)]
    L.qmbed_hip_model_stream.restype = c_void_p
    L.qmbed_hip_model_stream.argtypes = [c_void_p]
    L.qmbed_hip_apply_device.argtypes = [c_void_p, c_int, c_void_p, c_void_p, c_void_p]
    L.qmbed_hip_apply_host.argtypes = [c_void_p, c_int, c_void_p, c_void_p]
    L.qmbed_hip_time_apply.restype = ctypes.c_double
    L.qmbed_hip_time_apply.argtypes = [c_void_p, c_int, c_void_p, c_void_p, c_int]
    L.qmbed_hip_eigsh_json.restype = c_void_p
    L.qmbed_hip_eigsh_json.argtypes = [c_void_p, c_char_p]
    L.qmbed_hip_eigsh_batch_json.restype = c_void_p
    L.qmbed_hip_evolve_device.restype = c_void_p
    L.qmbed_hip_evolve_device.argtypes = [c_void_p, c_void_p, c_int, c_void_p, c_char_p]
    L.qmbed_hip_eigsh_batch_json.argtypes = [ctypes.POINTER(c_void_p), c_int, c_char_p]
    L.qmbed_hip_model_set_parameters.restype = c_void_p
    L.qmbed_hip_model_set_parameters.argtypes = [c_void_p, c_char_p]
    L.qmbed_hip_scan_create.restype = c_void_p
    L.qmbed_hip_scan_create.argtypes = [c_void_p, c_char_p, ctypes.POINTER(c_void_p)]
    L.qmbed_hip_scan_free.argtypes = [c_void_p]
    for f in (L.qmbed_hip_scan_train, L.qmbed_hip_scan_solve):
        f.restype = c_void_p
    L.qmbed_hip_scan_train.argtypes = [c_void_p, c_void_p, c_int, c_char_p]
    L.qmbed_hip_scan_solve.argtypes = [c_void_p, c_void_p, c_int, c_char_p, c_int]
    L.qmbed_hip_scan_predict.argtypes = [c_void_p, c_void_p, c_int, c_void_p]
    L.qmbed_hip_mlp_parameter_count.argtypes = [c_int, c_int, c_int, c_int]
    L.qmbed_hip_mlp_forward.argtypes = [c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_int, c_void_p]
    L.qmbed_hip_mlp_train.argtypes = [c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_int, c_int,
                                      ctypes.c_float, c_void_p, c_void_p]
    L.qmbed_hip_malloc.restype = c_void_p
    L.qmbed_hip_malloc.argtypes = [ctypes.c_size_t]
    L.qmbed_hip_free.argtypes = [c_void_p]
    L.qmbed_hip_memcpy.argtypes = [c_void_p, c_void_p, ctypes.c_size_t, c_int]
    L.qmbed_hip_gemm_atb.restype = c_int
    L.qmbed_hip_gemm_atb.argtypes = [c_void_p, ctypes.c_size_t, c_int, c_void_p, c_int, c_void_p]
    L.qmbed_hip_device_count.restype = c_int
    L.qmbed_hip_workspace_create.restype = c_void_p
    L.qmbed_hip_workspace_create.argtypes = []
    L.qmbed_hip_workspace_free.argtypes = [c_void_p]
    L.qmbed_hip_workspace_dimension.restype = c_long
    L.qmbed_hip_workspace_dimension.argtypes = [c_void_p]
    L.qmbed_hip_workspace_vectors.restype = c_int
    L.qmbed_hip_workspace_vectors.argtypes = [c_void_p]
    L.qmbed_hip_workspace_set.restype = c_void_p
    L.qmbed_hip_workspace_set.argtypes = [c_void_p, c_long, c_int, c_void_p]
    L.qmbed_hip_eigsh_workspace_json.restype = c_void_p
    L.qmbed_hip_eigsh_workspace_json.argtypes = [c_void_p, c_void_p, c_char_p]
    L.qmbed_hip_time_scan_batch.restype = ctypes.c_double
    L.qmbed_hip_time_scan_batch.argtypes = [c_void_p, c_int, c_int, c_void_p, c_int]
    L.qmbed_hip_apply_scan_batch.restype = c_int
    L.qmbed_hip_apply_scan_batch.argtypes = [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p]
    L.qmbed_hip_launch_timer.restype = c_int
    L.qmbed_hip_launch_timer.argtypes = [c_int]
    L.qmbed_hip_launch_timer_report.restype = c_void_p
    L.qmbed_hip_launch_timer_report.argtypes = []
    _lib = L
    return L


def _take(ptr) -> str:
    L = library()
    try:
        return ctypes.string_at(ptr).decode("utf-8")
    finally:
        L.qmbed_string_free(ptr)


def _call(fn_name: str, request) -> dict:
    payload = request if isinstance(request, (bytes, str)) else json.dumps(request)
    if isinstance(payload, str):
        payload = payload.encode("utf-8")
    response = json.loads(_take(getattr(library(), fn_name)(payload)))
    if response.get("status") != "ok":
        raise QmbedError(response.get("error", "unknown error"))
    return response["result"]


def run(request) -> dict:
    """One solve request (``qmbed_run_json``)."""
    return _call("qmbed_run_json", request)


def command(request) -> dict:
    """One tagged command (``qmbed_command_json``)."""
    return _call("qmbed_command_json", request)


def raw_call(fn_name: str, request) -> dict:
    """Decoded response envelope without raising on errors (for tests)."""
    payload = request if isinstance(request, (bytes, str)) or request is None else json.dumps(request)
    if isinstance(payload, str):
        payload = payload.encode("utf-8")
    return json.loads(_take(getattr(library(), fn_name)(payload)))


# --------------------------------------------------------------- requests --

def spin_basis(sites, up=None, momentum=None, parity=None, pauli=False, up_sectors=None, **extra):
    b = {"kind": "spin", "sites": sites, "pauli": pauli}
    if up is not None:
        b["up"] = up
    if up_sectors is not None:
        b["up_sectors"] = list(up_sectors)
    if momentum is not None:
        b["momentum"] = momentum
    if parity is not None:
        b["parity"] = parity
    b.update(extra)
    return b


_NAMES = {"I": "identity", "n": "number", "z": "z", "+": "raising", "-": "lowering", "x": "x", "y": "y"}


def term(symbols, couplings):
    """symbols like "zz" / "+-"; couplings: [(coefficient, [sites...]), ...]."""
    return {"product": {"local": [_NAMES[s] for s in symbols]},
            "couplings": [{"coefficient": [complex(c).real, complex(c).imag], "sites": list(map(int, s))}
                          for c, s in couplings]}


def terms_from_spec(spec):
    """[(symbols, [(coeff, sites)...]), ...] -> JSON term list."""
    return [term(sym, cpl) for sym, cpl in spec]


def chain_terms(L, jz=1.0, jxy=0.5, nn=1):
    bonds = [(i, (i + nn) % L) for i in range(L)]
    return [("zz", [(jz, b) for b in bonds]), ("+-", [(jxy, b) for b in bonds]), ("-+", [(jxy, b) for b in bonds])]


def xxz_terms(L, delta):
    return chain_terms(L, jz=delta, jxy=0.5)


def j1j2_terms(L, j1=1.0, j2=0.5):
    return chain_terms(L, jz=j1, jxy=0.5 * j1, nn=1) + chain_terms(L, jz=j2, jxy=0.5 * j2, nn=2)


def ising_terms(L, hx, hz, j=1.0):
    bonds = [(i, (i + 1) % L) for i in range(L)]
    out = [("zz", [(-j, b) for b in bonds])]
    if hx:
        out.append(("x", [(-hx, [i]) for i in range(L)]))
    if hz:
        out.append(("z", [(-hz, [i]) for i in range(L)]))
    return out


def _terms_json(terms):
    return terms if terms and isinstance(terms[0], dict) else terms_from_spec(terms)


class Model:
    """Device-resident sector + matrix-free operator (typed fast path)."""

    def __init__(self, basis: dict, terms, checks=None, components=None, format=None):
        """components: [(name, terms, default or None), ...] — a family
        H = H_static + Σ_p c_p H_p (the reference's create_model components).
        format: a stored format ("csr", "csc", "dense", "dia") whose matrix
        the products then run on, or None (matrix-free)."""
        L = library()
        req = {"basis": basis, "terms": _terms_json(terms)}
        if checks is not None:
            req["checks"] = checks
        if format is not None:
            req["format"] = format
        if components:
            req["components"] = []
            for name, cterms, default in components:
                c = {"name": name, "terms": _terms_json(cterms)}
                if default is not None:
                    c["default"] = [complex(default).real, complex(default).imag]
                req["components"].append(c)
        err = ctypes.c_void_p()
        h = L.qmbed_hip_model_create(json.dumps(req).encode(), ctypes.byref(err))
        if not h:
            msg = _take(err.value) if err.value else "model creation failed"
            raise QmbedError(msg)
        self._h = h
        self.basis = basis
        self.dim = int(L.qmbed_hip_model_dimension(h))
        self.is_real = bool(L.qmbed_hip_model_is_real(h))

    def close(self):
        if getattr(self, "_h", None):
            library().qmbed_hip_model_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def states(self):
        """(representatives, orbit sizes) in row order: uint32 words, or
        uint64 words for the wide sectors of 32 or more sites."""
        s = np.zeros(self.dim, dtype=np.uint32)
        o = np.zeros(self.dim, dtype=np.uint8)
        rc = library().qmbed_hip_model_states(self._h, s.ctypes.data, o.ctypes.data)
        if rc == -2:
            s = np.zeros(self.dim, dtype=np.uint64)
            rc = library().qmbed_hip_model_states64(self._h, s.ctypes.data, o.ctypes.data)
        if rc != 0:
            raise QmbedError("state copy failed")
        return s, o

    def save_basis_archive(self, path, metadata=None):
        """The sector's representatives, in row order, as a basis manifest of
        width `sites` bits (src/archive.rs:282-295 BasisArchive::from_basis,
        :333 save_basis_zip)."""
        states, _ = self.states()
        return save_basis_archive(path, int(self.basis["sites"]), states, metadata)

    def hv_config(self):
        lpr, img = ctypes.c_int(), ctypes.c_int()
        library().qmbed_hip_model_hv_config(self._h, ctypes.byref(lpr), ctypes.byref(img))
        return {"lanes_per_row": lpr.value, "mode": ("walk", "image", "stream")[img.value]}

    def nnz(self):
        v = int(library().qmbed_hip_model_nnz(self._h))
        if v < 0:
            raise QmbedError("entry count failed")
        return v

    def time_apply(self, dtype, iters=20, seed=0):
        """Average ms per H·v on device-resident random vectors (HIP events)."""
        dt = np.dtype(dtype)
        rng = np.random.default_rng(seed)
        x = rng.standard_normal(self.dim * (2 if dt.kind == "c" else 1)).astype(
            np.float32 if dt in (np.dtype(np.float32), np.dtype(np.complex64)) else np.float64)
        bx, by = DeviceBuffer(x.nbytes), DeviceBuffer(x.nbytes)
        bx.upload(x)
        L = library()
        ms = L.qmbed_hip_time_apply(self._h, _DTYPES[dt], bx.ptr, by.ptr, 2)  # warm
        ms = L.qmbed_hip_time_apply(self._h, _DTYPES[dt], bx.ptr, by.ptr, iters)
        bx.free()
        by.free()
        if ms < 0:
            raise QmbedError("timed apply failed")
        return ms

    def apply_scan_batch(self, points, x, fp32=True):
        """y[b] = H(points[b]) x[b] through the batched scan H·v (the kernel
        of lockstep scan solves) for len(points) lanes (1, 2, 4, 8 or 16);
        x real, shape (lanes, dim), rounded to FP32 when fp32."""
        pts = np.ascontiguousarray(points, dtype=np.float64)
        xs = np.ascontiguousarray(x, dtype=np.float64)
        if xs.shape != (pts.shape[0], self.dim):
            raise QmbedError("x must hold one vector of the dimension per point")
        y = np.zeros_like(xs)
        rc = library().qmbed_hip_apply_scan_batch(self._h, pts.shape[0], 1 if fp32 else 0, pts.ctypes.data,
                                                  xs.ctypes.data, y.ctypes.data)
        if rc == -2:
            raise QmbedError("batched scan launches of this width are not supported by the model")
        if rc < 0:
            raise QmbedError("batched scan launch failed")
        return y

    def time_scan_batch(self, points, fp32=True, iters=20):
        """Average ms per launch of the batched scan H·v over len(points)
        lanes (1, 2, 4, 8 or 16), lane b evaluated at points[b] (one real
        coefficient per component), FP32 or FP64 interleaved inputs."""
        pts = np.ascontiguousarray(points, dtype=np.float64)
        ms = library().qmbed_hip_time_scan_batch(self._h, pts.shape[0], 1 if fp32 else 0, pts.ctypes.data, iters)
        if ms == -2:
            raise QmbedError("batched scan launches of this width are not supported by the model")
        if ms < 0:
            raise QmbedError("timed batched scan launch failed")
        return ms

    def apply_mixed(self, x):
        """FP32 input, FP64 accumulation and output (compiled-stream mode):
        float32 → float64, or complex64 → complex128."""
        x = np.ascontiguousarray(x)
        if x.dtype not in (np.dtype(np.float32), np.dtype(np.complex64)):
            x = x.astype(np.complex64 if np.iscomplexobj(x) else np.float32)
        code = 5 if x.dtype == np.dtype(np.complex64) else 4
        y = np.empty(x.shape, dtype=np.complex128 if code == 5 else np.float64)
        rc = library().qmbed_hip_apply_host(self._h, code, x.ctypes.data, y.ctypes.data)
        if rc == -3:
            raise QmbedError("the mixed-precision action needs the compiled-stream mode")
        if rc != 0:
            raise QmbedError(f"apply failed (code {rc})")
        return y

    def time_apply_mixed(self, complex_=False, iters=20, seed=0):
        """Average ms per mixed-precision H·v (FP32 in, FP64 out) on device data."""
        rng = np.random.default_rng(seed)
        n_in = self.dim * (2 if complex_ else 1)
        x = rng.standard_normal(n_in).astype(np.float32)
        bx, by = DeviceBuffer(x.nbytes), DeviceBuffer(x.nbytes * 2)
        bx.upload(x)
        L = library()
        code = 5 if complex_ else 4
        ms = L.qmbed_hip_time_apply(self._h, code, bx.ptr, by.ptr, 2)  # warm
        ms = L.qmbed_hip_time_apply(self._h, code, bx.ptr, by.ptr, iters)
        bx.free()
        by.free()
        if ms < 0:
            raise QmbedError("timed apply failed")
        return ms

    def apply(self, x):
        x = np.ascontiguousarray(x)
        if x.dtype not in _DTYPES:
            x = x.astype(np.complex128 if np.iscomplexobj(x) else np.float64)
        y = np.empty_like(x)
        rc = library().qmbed_hip_apply_host(self._h, _DTYPES[x.dtype], x.ctypes.data, y.ctypes.data)
        if rc != 0:
            raise QmbedError(f"apply failed (code {rc})")
        return y

    def evolve(self, states, times, **options):
        """Krylov evolution exp(-iHt)ψ of the columns of `states` (ncols × dim,
        complex) through the device-resident entry point; returns
        (outputs [time][column][dim], diagnostics)."""
        x = np.ascontiguousarray(np.atleast_2d(states), dtype=np.complex128)
        nc = x.shape[0]
        if x.shape[1] != self.dim:
            raise QmbedError("states do not match the model dimension")
        bx = DeviceBuffer(max(x.nbytes, 16))
        bo = DeviceBuffer(max(x.nbytes * len(times), 16))
        try:
            bx.upload(x)
            req = {"times": list(map(float, times)), **options}
            res = _envelope(library().qmbed_hip_evolve_device(self._h, bx.ptr, nc, bo.ptr,
                                                              json.dumps(req).encode()))
            out = np.empty((len(times), nc, self.dim), dtype=np.complex128)
            bo.download(out)
        finally:
            bx.free()
            bo.free()
        return out, res

    def evolve_device(self, states_ptr, ncols, out_ptr, times, **options):
        """Device-resident evolution: `states_ptr` holds ncols complex128 columns
        ([column][row]), `out_ptr` receives [time][column][row]; returns the
        diagnostics."""
        req = {"times": list(map(float, times)), **options}
        return _envelope(library().qmbed_hip_evolve_device(self._h, states_ptr, int(ncols), out_ptr,
                                                           json.dumps(req).encode()))

    def set_parameters(self, **values):
        """Evaluate a family at the named coefficients (complex allowed)."""
        req = {k: [complex(v).real, complex(v).imag] for k, v in values.items()}
        response = json.loads(_take(library().qmbed_hip_model_set_parameters(self._h, json.dumps(req).encode())))
        if response.get("status") != "ok":
            raise QmbedError(response.get("error", "unknown error"))
        self.is_real = bool(response["result"]["real"])
        return [complex(a, b) for a, b in response["result"]["values"]]

    def eigsh_workspace(self, workspace, k, target="smallest_algebraic", parameters=None, **solver):
        """eigsh_with_workspace: started from `workspace` (a Workspace), which
        then holds this solve's eigenvectors."""
        req = {"eigenpairs": k, "target": {"kind": target}, **solver}
        if parameters is not None:
            req["parameters"] = {n: [complex(v).real, complex(v).imag] for n, v in parameters.items()}
        return _envelope(library().qmbed_hip_eigsh_workspace_json(self._h, workspace.handle,
                                                                  json.dumps(req).encode()))

    def eigsh(self, k, target="smallest_algebraic", parameters=None, **solver):
        req = {"eigenpairs": k, "target": {"kind": target}, **solver}
        if parameters is not None:
            req["parameters"] = {n: [complex(v).real, complex(v).imag] for n, v in parameters.items()}
        return _call_model(self._h, req)


def launch_timer(on):
    """Switch the library's launch timer on (clearing its records) or off."""
    if library().qmbed_hip_launch_timer(1 if on else 0) != 0:
        raise QmbedError("launch timer switch failed")


def launch_timer_report():
    """{kernel instance: {"launches", "ms_total", "ms_mean"}} of the launches
    recorded since the timer was switched on."""
    return _envelope(library().qmbed_hip_launch_timer_report())


class Workspace:
    """EigshWorkspace (QMBED src/solve/mod.rs:477-526): the eigenvectors of
    the previous workspace solve, reused by the next (Model.eigsh_workspace)."""

    def __init__(self):
        self._h = library().qmbed_hip_workspace_create()

    @property
    def handle(self):
        return self._h

    @property
    def dimension(self):
        return int(library().qmbed_hip_workspace_dimension(self._h))

    def __len__(self):
        return int(library().qmbed_hip_workspace_vectors(self._h))

    def set(self, dimension, vectors):
        """set_initial_subspace: complex vectors (count × dimension)."""
        v = np.ascontiguousarray(np.array(vectors, dtype=np.complex128).reshape(-1, dimension))
        return _envelope(library().qmbed_hip_workspace_set(self._h, int(dimension), v.shape[0], v.ctypes.data))

    def close(self):
        if getattr(self, "_h", None):
            library().qmbed_hip_workspace_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def eigsh_batch(models, k, target="smallest_algebraic", **solver):
    """Solve several models (e.g. the symmetry sectors of one Hamiltonian) in
    one lockstep batch; returns {"results": [one eigsh result per model],
    "batch": {...}}."""
    req = {"eigenpairs": k, "target": {"kind": target}, **solver}
    handles = (ctypes.c_void_p * len(models))(*[m.handle for m in models])
    return _envelope(library().qmbed_hip_eigsh_batch_json(handles, len(models), json.dumps(req).encode()))


def _call_model(h, req):
    L = library()
    response = json.loads(_take(L.qmbed_hip_eigsh_json(h, json.dumps(req).encode())))
    if response.get("status") != "ok":
        raise QmbedError(response.get("error", "unknown error"))
    return response["result"]


def save_basis_archive(path, width_bits, states, metadata=None):
    """Write a basis manifest (src/archive.rs:333-362): states as an unsigned
    integer array (64-bit words) or a list of Python ints / decimal strings of
    any width, in row order."""
    L = library()
    req = {"path": os.fspath(path), "width_bits": int(width_bits), "metadata": dict(metadata or {})}
    if isinstance(states, np.ndarray) and states.dtype.kind == "u":
        words = np.ascontiguousarray(states, dtype=np.uint64)
        return _envelope(L.qmbed_hip_basis_archive_save(json.dumps(req).encode(), words.ctypes.data, words.size))
    req["states"] = [str(int(x)) if not isinstance(x, str) else x for x in states]
    return _envelope(L.qmbed_hip_basis_archive_save(json.dumps(req).encode(), None, 0))


def load_basis_archive(path):
    """(width_bits, states, metadata) of a basis manifest (src/archive.rs:
    364-416): states as a uint64 array when every one fits 64 bits, else a
    list of Python ints."""
    L = library()
    p = os.fspath(path).encode()
    head = _envelope(L.qmbed_hip_basis_archive_load(p, None, 0, 0))
    if head["fits_u64"]:
        states = np.zeros(head["count"], dtype=np.uint64)
        res = _envelope(L.qmbed_hip_basis_archive_load(p, states.ctypes.data, states.size, 0))
        if not res["written"]:
            raise QmbedError("basis archive changed while it was read")
    else:
        res = _envelope(L.qmbed_hip_basis_archive_load(p, None, 0, 1))
        states = [int(x) for x in res["states"]]
    return res["width_bits"], states, res["metadata"]


def _envelope(ptr):
    response = json.loads(_take(ptr))
    if response.get("status") != "ok":
        raise QmbedError(response.get("error", "unknown error"))
    return response["result"]


class Scan:
    """Learned Lanczos warm start over a family's parameter space (typed path):
    train on points solved with the plain start, then solve points from the
    network's prediction (warm=True) or the plain start (warm=False)."""

    def __init__(self, model: "Model", **options):
        err = ctypes.c_void_p()
        h = library().qmbed_hip_scan_create(model.handle, json.dumps(options).encode(), ctypes.byref(err))
        if not h:
            raise QmbedError(_take(err.value) if err.value else "scan creation failed")
        self._h = h
        self._model = model  # the scan re-evaluates this model

    def close(self):
        if getattr(self, "_h", None):
            library().qmbed_hip_scan_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def train(self, points, **solver):
        pts = np.ascontiguousarray(points, dtype=np.float64)
        req = {"eigenpairs": 1, "target": {"kind": "smallest_algebraic"}, **solver}
        return _envelope(library().qmbed_hip_scan_train(self._h, pts.ctypes.data, pts.shape[0],
                                                        json.dumps(req).encode()))

    def solve(self, points, warm=True, **solver):
        """warm: False (plain start), True (network start) or "rayleigh_ritz"
        (the lowest eigenvector of the family projected on the snapshots)."""
        pts = np.ascontiguousarray(points, dtype=np.float64)
        req = {"eigenpairs": 1, "target": {"kind": "smallest_algebraic"}, **solver}
        mode = 2 if warm == "rayleigh_ritz" else (1 if warm else 0)
        return _envelope(library().qmbed_hip_scan_solve(self._h, pts.ctypes.data, pts.shape[0],
                                                        json.dumps(req).encode(), mode))

    def predict(self, points):
        pts = np.ascontiguousarray(points, dtype=np.float64)
        out = np.zeros((pts.shape[0], 64))
        k = library().qmbed_hip_scan_predict(self._h, pts.ctypes.data, pts.shape[0], out.ctypes.data)
        if k < 0:
            raise QmbedError("prediction failed")
        return out.ravel()[: pts.shape[0] * k].reshape(pts.shape[0], k)


def mlp_forward(shape, params, x):
    """Network outputs on the GPU (FP32)."""
    i, h1, h2, o = shape
    p = np.ascontiguousarray(params, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.zeros((x.shape[0], o), dtype=np.float32)
    if library().qmbed_hip_mlp_forward(i, h1, h2, o, p.ctypes.data, x.ctypes.data, x.shape[0], y.ctypes.data) != 0:
        raise QmbedError("mlp forward failed")
    return y


def mlp_gradient(shape, params, x, t):
    """(gradient, loss) of one evaluation by the GPU training kernel."""
    i, h1, h2, o = shape
    p = np.ascontiguousarray(params, dtype=np.float32).copy()
    x = np.ascontiguousarray(x, dtype=np.float32)
    t = np.ascontiguousarray(t, dtype=np.float32)
    g = np.zeros_like(p)
    loss = np.zeros(1, dtype=np.float32)
    if library().qmbed_hip_mlp_train(i, h1, h2, o, p.ctypes.data, x.ctypes.data, t.ctypes.data, x.shape[0], 0,
                                     0.0, g.ctypes.data, loss.ctypes.data) != 0:
        raise QmbedError("mlp gradient failed")
    return g, float(loss[0])


def mlp_train(shape, params, x, t, epochs, lr):
    """(trained parameters, loss of every epoch) from the GPU training kernel."""
    i, h1, h2, o = shape
    p = np.ascontiguousarray(params, dtype=np.float32).copy()
    x = np.ascontiguousarray(x, dtype=np.float32)
    t = np.ascontiguousarray(t, dtype=np.float32)
    hist = np.zeros(epochs, dtype=np.float32)
    if library().qmbed_hip_mlp_train(i, h1, h2, o, p.ctypes.data, x.ctypes.data, t.ctypes.data, x.shape[0], epochs,
                                     float(lr), None, hist.ctypes.data) != 0:
        raise QmbedError("mlp training failed")
    return p, hist


class DeviceBuffer:
    """Raw device allocation (for timing device-resident work)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = library().qmbed_hip_malloc(self.nbytes)
        if not self.ptr:
            raise MemoryError(f"hipMalloc of {nbytes} bytes failed")

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        if library().qmbed_hip_memcpy(self.ptr, arr.ctypes.data, arr.nbytes, 0) != 0:
            raise QmbedError("upload failed")

    def download(self, arr):
        assert arr.nbytes <= self.nbytes
        if library().qmbed_hip_memcpy(arr.ctypes.data, self.ptr, arr.nbytes, 1) != 0:
            raise QmbedError("download failed")
        return arr

    def free(self):
        if self.ptr:
            library().qmbed_hip_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
