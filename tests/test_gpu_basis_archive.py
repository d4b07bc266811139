"""Basis manifests of sectors built on the GPU: the representatives of a
sector, in row order, written as a basis archive (src/archive.rs:282-295
BasisArchive::from_basis, :333 save_basis_zip) equal, byte for byte, the
archive written from the oracle's states, and read back unchanged; 32-bit
and 64-bit state words."""

import zipfile

import numpy as np
import pytest

from tests import cases

pytestmark = pytest.mark.gpu

# (sites, up, momentum, parity)
SECTORS = [(20, 10, 0, 1), (18, None, 3, None), (40, 3, 0, None), (36, 2, 5, None)]


@pytest.mark.parametrize("L,up,k,p", SECTORS)
def test_sector_manifest_matches_oracle(qa, oracle, tmp_path, L, up, k, p):
    print(f"[basis archive] L={L} up={up} k={k} p={p}", flush=True)
    b = {"kind": "spin", "sites": L}
    for key, v in (("up", up), ("momentum", k), ("parity", p)):
        if v is not None:
            b[key] = v
    spec = cases.heisenberg(L)
    m = qa.Model(b, spec)
    om = oracle.Model(L, spec, up=up, momentum=k, parity=p)
    ref_states, _ = om.states()
    assert m.dim == len(ref_states)
    meta = {"kind": "spin", "sites": str(L)}
    gpu_path, ref_path = tmp_path / "gpu.zip", tmp_path / "oracle.zip"
    res = m.save_basis_archive(gpu_path, meta)
    assert res["count"] == m.dim and res["width_bits"] == L
    qa.save_basis_archive(ref_path, L, np.asarray(ref_states, dtype=np.uint64), meta)
    with zipfile.ZipFile(gpu_path) as g, zipfile.ZipFile(ref_path) as r:
        for name in ("basis_archive_version.npy", "basis_manifest.tsv", "metadata.tsv"):
            assert g.read(name) == r.read(name), name
    width, states, back = qa.load_basis_archive(gpu_path)
    assert width == L and back == meta
    np.testing.assert_array_equal(states, np.asarray(ref_states, dtype=np.uint64))
