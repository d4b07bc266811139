"""CPU-only checks of the native library: it loads, exports every symbol the
public header declares, and answers protocol errors (before any GPU work)
with the reference's envelopes and texts."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "qmbed_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set()
    for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(qmbed_\w+)\s*\(", text, flags=re.M):
        names.add(m.group(1))
    return sorted(names)


def test_header_symbols_exported(qa):
    lib = ctypes.CDLL(qa.LIBRARY_PATH)
    names = declared_symbols()
    assert "qmbed_run_json" in names and "qmbed_string_free" in names and len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n


def test_reference_exports_present(qa):
    # the five symbols the reference's Python loader resolves (bindings/python/src/qmbed/_ffi.py:154-185)
    lib = ctypes.CDLL(qa.LIBRARY_PATH)
    for n in ("qmbed_run_json", "qmbed_command_json", "qmbed_string_free",
              "qmbed_evolve_model_with_drive_json", "qmbed_register_user_basis_32_json",
              "qmbed_register_user_basis_64_json"):
        assert hasattr(lib, n)


def test_null_request(qa):
    # bindings/capi/src/lib.rs:6093-6095
    assert qa.raw_call("qmbed_run_json", None) == {"status": "error", "error": "request pointer is null"}
    assert qa.raw_call("qmbed_command_json", None) == {"status": "error", "error": "request pointer is null"}


def test_protocol_errors(qa):
    r = qa.raw_call("qmbed_run_json", "{bad")
    assert r["status"] == "error" and r["error"].startswith("invalid options: invalid binding request")
    r = qa.raw_call("qmbed_command_json", {"operation": "release_model", "handle": "0"})
    assert r == {"status": "error", "error": "invalid options: model handle must be positive"}
    r = qa.raw_call("qmbed_command_json", {"operation": "release_model", "handle": "abc"})
    assert r["error"] == 'invalid options: model handle "abc" is not a positive integer'
    r = qa.raw_call("qmbed_command_json", {"operation": "release_model", "handle": "17"})
    assert r["error"] == "invalid options: model handle 17 is not registered"
    r = qa.raw_call("qmbed_command_json", {"operation": "apply_expm_action", "expm_handle": "7", "vectors": []})
    assert r["error"] == 'invalid options: exponential-action handle "7" does not use the expm: prefix'
    r = qa.raw_call("qmbed_command_json", {"operation": "release_expm_action", "expm_handle": "expm:0"})
    assert r["error"] == "invalid options: exponential-action handle must be positive"
    r = qa.raw_call("qmbed_command_json", {"operation": "release_expm_action", "expm_handle": "expm:9"})
    assert r["error"] == "invalid options: exponential-action handle expm:9 is not registered"


def test_out_of_path_requests_error(qa):
    r = qa.raw_call("qmbed_command_json", {"operation": "describe_basis",
                                           "basis": {"kind": "boson", "sites": 3, "states_per_site": 2}})
    assert r["status"] == "error" and "unsupported backend" in r["error"]
    r = qa.raw_call("qmbed_command_json", {"operation": "describe_basis",
                                           "basis": {"kind": "spin", "sites": 4, "reverse": True}})
    assert r["status"] == "error"
    r = qa.raw_call("qmbed_command_json", {"operation": "describe_basis",
                                           "basis": {"kind": "spin", "sites": 4, "up": 2, "up_sectors": [1, 2]}})
    assert r["error"] == "invalid options: spin basis accepts either up or up_sectors, not both"


def test_bad_terms_rejected_before_device_work(qa):
    bad_arity = {"basis": {"kind": "spin", "sites": 4},
                 "terms": [{"product": {"local": ["z", "z"]}, "couplings": [{"coefficient": [1, 0], "sites": [0]}]}],
                 "solver": {"eigenpairs": 1, "target": {"kind": "smallest_algebraic"}}}
    r = qa.raw_call("qmbed_run_json", bad_arity)
    assert r["status"] == "error" and r["error"].startswith("invalid coupling: operator \"zz\" has arity 2")
    bad_name = dict(bad_arity)
    bad_name["terms"] = [{"product": {"local": ["spin"]}, "couplings": []}]
    r = qa.raw_call("qmbed_run_json", bad_name)
    assert r["error"] == "invalid or unregistered operator: spin"


def test_no_cpu_fallback_without_library(tmp_path, monkeypatch):
    import importlib

    import qmbed_amd

    monkeypatch.setenv("QMBED_HIP_LIBRARY", str(tmp_path / "missing.so"))
    mod = importlib.reload(qmbed_amd)
    try:
        with pytest.raises(RuntimeError, match="native library not found"):
            mod.library()
    finally:
        monkeypatch.delenv("QMBED_HIP_LIBRARY")
        importlib.reload(qmbed_amd)


def _raw(fn, *args):
    import json

    import qmbed_amd

    lib = qmbed_amd.library()
    f = getattr(lib, fn)
    f.restype = ctypes.c_void_p
    return json.loads(qmbed_amd._take(f(*args)))


def test_user_basis_entry_points_take_the_reference_arguments(qa):
    # six parameters: request, operator, next_state, pre_check, map_callbacks, map_count
    # (bindings/capi/src/lib.rs:6194-6251); the argument checks come first (:6207-6216)
    op32 = ctypes.CFUNCTYPE(ctypes.c_int32, ctypes.c_void_p, ctypes.c_int8, ctypes.c_int32, ctypes.c_int32,
                            ctypes.c_void_p)(lambda *a: 0)
    op64 = ctypes.CFUNCTYPE(ctypes.c_int32, ctypes.c_void_p, ctypes.c_int8, ctypes.c_int32, ctypes.c_int32,
                            ctypes.c_void_p)(lambda *a: 0)
    req = b'{"sites": 4, "states_per_site": 2, "allowed_ops": "xyz"}'
    for fn, op in (("qmbed_register_user_basis_32_json", op32), ("qmbed_register_user_basis_64_json", op64)):
        lib = qa.library()
        getattr(lib, fn).argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_size_t]
        assert _raw(fn, None, op, None, None, None, 0) == {"status": "error", "error": "request pointer is null"}
        assert _raw(fn, req, None, None, None, None, 0) == {"status": "error",
                                                            "error": "user operator callback is null"}
        assert _raw(fn, req, op, None, None, None, 2) == {"status": "error",
                                                          "error": "user symmetry callback array is null"}
        r = _raw(fn, req, op, None, None, None, 0)
        assert r["status"] == "error" and r["error"].startswith("unsupported backend")


def test_drive_entry_point_checks_its_callback(qa):
    # QmbedDriveCallback(context, time, QmbedComplex64* coefficients, count) (lib.rs:464-469, :6153-6179)
    lib = qa.library()
    lib.qmbed_evolve_model_with_drive_json.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p]
    cb = ctypes.CFUNCTYPE(ctypes.c_int32, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p,
                          ctypes.c_size_t)(lambda *a: 0)
    req = b'{"handle": "1"}'
    assert _raw("qmbed_evolve_model_with_drive_json", None, cb, None)["error"] == "request pointer is null"
    assert _raw("qmbed_evolve_model_with_drive_json", req, None, None)["error"] == "drive callback is null"
    r = _raw("qmbed_evolve_model_with_drive_json", req, cb, None)
    assert r["status"] == "error" and r["error"].startswith("unsupported backend")


def test_non_utf8_requests(qa):
    # CStr::to_str failures, displayed as Rust's Utf8Error (lib.rs:6096-6103)
    lib = qa.library()
    for fn in ("qmbed_run_json", "qmbed_command_json"):
        getattr(lib, fn).argtypes = [ctypes.c_char_p]
        r = _raw(fn, b'{"a\xff": 1}')
        assert r == {"status": "error", "error": "request is not UTF-8: invalid utf-8 sequence of 1 bytes from index 3"}
        r = _raw(fn, b'{"\xe2\x82')
        assert r["error"] == "request is not UTF-8: incomplete utf-8 byte sequence from index 2"
        r = _raw(fn, b'{"\xe2\x28\xa1"}')  # bad second byte of a 3-byte sequence
        assert r["error"] == "request is not UTF-8: invalid utf-8 sequence of 1 bytes from index 2"
        r = _raw(fn, b'{"\xe2\x82\x28"}')  # bad third byte
        assert r["error"] == "request is not UTF-8: invalid utf-8 sequence of 2 bytes from index 2"
        r = _raw(fn, b'{"\xed\xa0\x80"}')  # an encoded surrogate
        assert r["error"] == "request is not UTF-8: invalid utf-8 sequence of 1 bytes from index 2"
        r = _raw(fn, '{"é€😀": 1}'.encode())  # valid multi-byte text parses on to the schema checks
        assert not r["error"].startswith("request is not UTF-8")


def test_deep_nesting_and_bad_numbers_are_errors(qa):
    deep = "[" * 5000 + "]" * 5000
    r = qa.raw_call("qmbed_run_json", deep)
    assert r["status"] == "error" and "recursion limit exceeded" in r["error"]
    ok_depth = '{"basis": ' + "[" * 100 + "]" * 100 + "}"
    r = qa.raw_call("qmbed_run_json", ok_depth)
    assert "recursion limit" not in r["error"]
    r = qa.raw_call("qmbed_command_json", {"operation": "describe_basis", "basis": {"kind": "spin", "sites": 1e300}})
    assert r["error"] == "invalid options: sites: integer out of range"
    r = qa.raw_call("qmbed_command_json", {"operation": "describe_basis",
                                           "basis": {"kind": "spin", "sites": 4, "up": 4294967298}})
    assert r["error"] == "invalid options: up: integer out of range"
    r = qa.raw_call("qmbed_run_json", '{"basis": "\\ud800"}')
    assert r["status"] == "error" and "surrogate" in r["error"]
    r = qa.raw_call("qmbed_run_json", '{"basis": "\\ud83d\\ude00"}')  # a surrogate pair is one code point
    assert "surrogate" not in r["error"]


def test_general_symmetry_requests_checked_before_device_work(qa):
    # the lattice-map checks of runtime_symmetry_sector and
    # LatticeSymmetryMap::new (bindings/capi/src/lib.rs:3103-3127,
    # src/basis/mod.rs:2978-3060) and the request-level checks of
    # build_spin_basis (:2849-2880), with the reference's texts
    def err(**extra):
        basis = {"kind": "spin", "sites": 6, "up": 3, **extra}
        r = qa.raw_call("qmbed_command_json", {"operation": "describe_basis", "basis": basis})
        assert r["status"] == "error"
        return r["error"]

    t6 = [1, 2, 3, 4, 5, 0]
    assert err(symmetries=[{"destinations": [1, 0], "sector": 0}]) == \
        "invalid options: symmetry map has 2 sites, expected 6"
    assert err(symmetries=[{"destinations": [1, 1, 2, 3, 4, 5], "sector": 0}]) == \
        "incompatible symmetry request: symmetry site destinations must be a bijection"
    assert err(symmetries=[{"destinations": [0, 1, 2, 3, 4, 9], "sector": 0}]) == \
        "incompatible symmetry request: symmetry site destinations must be a bijection"
    assert err(symmetries=[{"destinations": t6, "local_permutations": [[0, 0]] * 6, "sector": 0}]) == \
        "incompatible symmetry request: every local-state map must be a permutation"
    assert err(symmetries=[{"destinations": t6, "local_permutations": [[1, 0]] * 5, "sector": 0}]) == \
        "incompatible symmetry request: every local-state map must be a permutation"
    assert err(symmetries=[{"destinations": t6, "sector": 1}], momentum=1) == \
        "invalid options: built-in and general spin symmetries cannot be mixed"
    assert err(matrix_symmetry={"dimension": 2, "selected_row": 0, "generators": []}) == \
        "invalid options: matrix-valued symmetry sectors materialize as projected operator models"


def test_site_permutation_and_basis_plan_requests_checked(qa):
    # validate_site_permutation (src/interop.rs:1328-1345) and the basis-plan
    # handles (bindings/capi/src/lib.rs:1706-1718, 1799-1823)
    def err(req):
        r = qa.raw_call("qmbed_command_json", req)
        assert r["status"] == "error"
        return r["error"]

    base = {"operation": "create_model", "basis": {"kind": "spin", "sites": 3},
            "terms": [{"product": {"local": ["z", "z"]}, "couplings": [{"coefficient": [1.0, 0.0], "sites": [0, 1]}]}]}
    assert err({**base, "site_permutation": [0, 3, 1]}) == "site 3 is outside a lattice with 3 sites"
    assert err({**base, "site_permutation": [0, 1, 1]}) == "invalid options: site permutation must be bijective"
    assert err({"operation": "materialize_basis_plan", "plan_handle": "0"}) == \
        "invalid options: basis-plan handle must be positive"
    assert err({"operation": "materialize_basis_plan", "plan_handle": "x"}) == \
        'invalid options: basis-plan handle "x" is not a positive integer'
    assert err({"operation": "release_basis_plan", "plan_handle": "41"}) == \
        "invalid options: basis-plan handle 41 is not registered"
