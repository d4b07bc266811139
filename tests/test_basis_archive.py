"""Basis manifests through the C ABI (host code, no GPU): the reference's own
round trip (tests/archive_extensions.rs:98-123: 256-bit states 0, 3 and 2^200
with two metadata items), the container read back by Python's zipfile and
numpy (so other tools can read what is written), archives written by
zipfile (stored and deflated, numpy's NPY header) read here, and the
reference's error texts (src/archive.rs:222-280, :365-416;
src/basis/mod.rs:4838-4852, :5139-5169)."""

import io
import zipfile

import numpy as np
import pytest

import qmbed_amd as qa

BIG = "1606938044258990275541962092341162602522202993782792835301376"


def test_wide_states_round_trip(tmp_path):
    assert int(BIG) == 2 ** 200
    path = tmp_path / "basis.npz"
    res = qa.save_basis_archive(path, 256, ["0", "3", BIG], {"sites": "200", "kind": "spin"})
    assert res == {"path": str(path), "width_bits": 256, "count": 3}
    width, states, meta = qa.load_basis_archive(path)
    assert width == 256 and states == [0, 3, 2 ** 200]
    assert meta == {"kind": "spin", "sites": "200"}
    # the container, read by other tools
    with zipfile.ZipFile(path) as z:
        assert sorted(z.namelist()) == ["basis_archive_version.npy", "basis_manifest.tsv", "metadata.tsv"]
        assert all(i.compress_type == zipfile.ZIP_DEFLATED for i in z.infolist())
        assert z.testzip() is None
        version = np.load(io.BytesIO(z.read("basis_archive_version.npy")), allow_pickle=False)
        assert version.dtype == np.uint8 and version.tolist() == [1]
        assert z.read("basis_manifest.tsv").decode() == f"width_bits\t256\nstate\t0\nstate\t3\nstate\t{BIG}\n"
        # metadata rows in key order
        assert z.read("metadata.tsv").decode() == "kind\tspin\nsites\t200\n"


def test_u64_words_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    states = np.unique(rng.integers(0, 2 ** 63, 5000, dtype=np.uint64))
    states = np.concatenate([states, np.array([2 ** 64 - 1], dtype=np.uint64)])
    rng.shuffle(states)  # row order is kept, not sorted
    path = tmp_path / "words.zip"
    qa.save_basis_archive(path, 64, states)
    width, back, meta = qa.load_basis_archive(path)
    assert width == 64 and meta == {}
    assert back.dtype == np.uint64
    np.testing.assert_array_equal(back, states)
    # the same states given as integers write the same manifest
    path2 = tmp_path / "ints.zip"
    qa.save_basis_archive(path2, 64, [int(x) for x in states])
    with zipfile.ZipFile(path) as a, zipfile.ZipFile(path2) as b:
        assert a.read("basis_manifest.tsv") == b.read("basis_manifest.tsv")


def _foreign(path, compression, manifest, metadata=None, version=1):
    buf = io.BytesIO()
    np.save(buf, np.array([version], dtype=np.uint8))
    with zipfile.ZipFile(path, "w", compression=compression) as z:
        z.writestr("basis_archive_version.npy", buf.getvalue())
        z.writestr("basis_manifest.tsv", manifest)
        if metadata is not None:
            z.writestr("metadata.tsv", metadata)


@pytest.mark.parametrize("compression", [zipfile.ZIP_STORED, zipfile.ZIP_DEFLATED])
def test_reads_archives_written_by_zipfile(tmp_path, compression):
    path = tmp_path / "foreign.zip"
    # leading zeros, '+' and '_' separators parse as BigUint::parse_bytes does;
    # CRLF line ends and a missing metadata entry are accepted
    _foreign(path, compression, "width_bits\t12\r\nstate\t0007\nstate\t+4_0\nstate\t4095\n")
    width, states, meta = qa.load_basis_archive(path)
    assert width == 12 and states.tolist() == [7, 40, 4095] and meta == {}
    _foreign(path, compression, "width_bits\t8\nstate\t1\n", "a.b-c_1\tx y\n")
    assert qa.load_basis_archive(path)[2] == {"a.b-c_1": "x y"}


def test_zip64_records_from_zipfile(tmp_path):
    # numpy's savez writes its members with force_zip64 (zip64 extra fields)
    path = tmp_path / "z64.zip"
    buf = io.BytesIO()
    np.save(buf, np.array([1], dtype=np.uint8))
    with zipfile.ZipFile(path, "w", compression=zipfile.ZIP_DEFLATED) as z:
        for name, data in (("basis_archive_version.npy", buf.getvalue()),
                           ("basis_manifest.tsv", b"width_bits\t5\nstate\t9\nstate\t2\n")):
            with z.open(name, "w", force_zip64=True) as f:
                f.write(data)
    width, states, _ = qa.load_basis_archive(path)
    assert width == 5 and states.tolist() == [9, 2]


@pytest.mark.parametrize("manifest,message", [
    ("", "archive error: basis width is missing"),
    ("width\t4\n", "archive error: basis width is missing"),
    ("width_bits\t4\nstat\t1\n", "archive error: invalid basis state row"),
    ("width_bits\t4\nstate\t16\n", "unsupported backend: integer needs 5 bits but the requested state width is 4"),
    ("width_bits\t4\nstate\t1\nstate\t01\n", "archive error: basis archives require unique physical states"),
    ("width_bits\t4\nstate\t-1\n", 'invalid options: basis state "-1" is not a nonnegative decimal integer'),
    ("width_bits\t4\nstate\t_1\n", 'invalid options: basis state "_1" is not a nonnegative decimal integer'),
    ("width_bits\t0\nstate\t0\n", "invalid Hilbert-space sector: a basis state needs at least one bit"),
    ("width_bits\t0\n", "archive error: basis states must share one positive logical width"),
    ("width_bits\t16385\nstate\t1\n", "unsupported backend: basis state requires more than 16384 bits"),
])
def test_manifest_errors(tmp_path, manifest, message):
    path = tmp_path / "bad.zip"
    _foreign(path, zipfile.ZIP_DEFLATED, manifest)
    with pytest.raises(qa.QmbedError) as e:
        qa.load_basis_archive(path)
    assert str(e.value) == message


def test_container_errors(tmp_path):
    path = tmp_path / "v2.zip"
    _foreign(path, zipfile.ZIP_DEFLATED, "width_bits\t4\n", version=2)
    with pytest.raises(qa.QmbedError, match="^archive error: unsupported basis archive version 2$"):
        qa.load_basis_archive(path)
    with zipfile.ZipFile(path, "w") as z:
        z.writestr("basis_manifest.tsv", "width_bits\t4\n")
    with pytest.raises(qa.QmbedError, match="^archive error: specified file not found in archive$"):
        qa.load_basis_archive(path)
    with zipfile.ZipFile(path, "w") as z:
        z.writestr("basis_archive_version.npy", b"NUMPY")
    with pytest.raises(qa.QmbedError, match="^archive error: invalid NPY magic$"):
        qa.load_basis_archive(path)
    path.write_bytes(b"not a zip archive at all" * 4)
    with pytest.raises(qa.QmbedError, match="^archive error: invalid Zip archive"):
        qa.load_basis_archive(path)
    with pytest.raises(qa.QmbedError, match=r"^archive error: No such file or directory \(os error 2\)$"):
        qa.load_basis_archive(tmp_path / "missing.zip")


def test_save_errors(tmp_path):
    path = tmp_path / "x.zip"
    with pytest.raises(qa.QmbedError, match="^archive error: basis archives require unique"):
        qa.save_basis_archive(path, 8, np.array([3, 5, 3], dtype=np.uint64))
    with pytest.raises(qa.QmbedError, match="^unsupported backend: integer needs 9 bits"):
        qa.save_basis_archive(path, 8, np.array([256], dtype=np.uint64))
    with pytest.raises(qa.QmbedError, match="^archive error: archive metadata keys may contain only"):
        qa.save_basis_archive(path, 8, [1], {"bad key": "x"})
    with pytest.raises(qa.QmbedError, match="^archive error: archive metadata values may not contain tabs"):
        qa.save_basis_archive(path, 8, [1], {"key": "a\tb"})
    with pytest.raises(qa.QmbedError, match="^archive error: basis states must share one positive logical width"):
        qa.save_basis_archive(path, 0, [])
    assert not path.exists()
