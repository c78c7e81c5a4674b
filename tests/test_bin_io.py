"""Map.bin / nMap.bin / maskMap.bin layout and the header hash."""
import hashlib

import numpy as np
import pytest

import midas_amd as ma
from midas_amd import synth
from midas_amd.bin_io import MapHeader, write_map
from helpers import small_cases, to_oracle


def test_sha256_known_answers():
    # FIPS 180-4 examples: the header hash is plain SHA-256 of the string.
    assert hashlib.sha256(b"abc").hexdigest() == (
        "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad")
    assert hashlib.sha256(b"abcdbcdecdefdefgefghfghighijhijkijkljklmklmnlmnomnopnopq").hexdigest() == (
        "248d6a61d204a8cd7ee6eeab7e5ca6f8fe0a7ba00cfc2c12f3d8e36e0e0d4b47"[:0] +
        "248d6a61d20638b8e5c026930c3e6039a33ce45964ff2167f6ecedd419db06c1")


@pytest.mark.parametrize("name", ["flat", "tilt", "subpix", "qmode"])
def test_canonical_string_matches_oracle(name, oracle_lib):
    P = small_cases()[name]
    assert ma.canonical_string(P) == oracle_lib.header_string(to_oracle(P))


def test_header_layout_and_round_trip():
    P = small_cases()["qmode"]
    h = ma.header_for(P)
    b = h.to_bytes()
    assert len(b) == 64 and b[:4] == b"MAP0"
    assert int.from_bytes(b[4:8], "little") == 3
    assert b[40] == 1 and h.wavelength == P.Wavelength
    assert MapHeader.from_bytes(b) == h
    assert MapHeader.from_bytes(b"\0" * 64) is None


def test_write_then_load(tmp_path, oracle_lib):
    P = small_cases()["flat"]
    rng = np.random.default_rng(2)
    mask = (rng.random((P.NrPixelsZ, P.NrPixelsY)) < 0.02).astype(float)
    m = oracle_lib.build_map(to_oracle(P), mask=mask)
    write_map(tmp_path, pxList=m.entries.view(ma.PXLIST_DTYPE), counts=m.counts,
              offsets=m.offsets, mask_flags=m.flags, header=ma.header_for(P))
    pm = ma.load_map(tmp_path / "Map.bin", tmp_path / "nMap.bin", tmp_path / "maskMap.bin")
    assert pm.pxList.tobytes() == m.entries.tobytes()
    np.testing.assert_array_equal(pm.counts, m.counts)
    np.testing.assert_array_equal(pm.mask_flags, m.flags)
    assert pm.map_header.param_hash == ma.header_for(P).param_hash
    assert (tmp_path / "Map.bin").stat().st_size == 64 + 24 * m.entries.size
    assert (tmp_path / "nMap.bin").stat().st_size == 64 + 8 * m.counts.size


def test_legacy_and_bad_files(tmp_path):
    rec = np.zeros(3, dtype=ma.PXLIST_DTYPE)
    write_map(tmp_path, pxList=rec, counts=np.array([1, 2]), offsets=np.array([0, 1]))
    pm = ma.load_map(tmp_path / "Map.bin", tmp_path / "nMap.bin")
    assert pm.map_header is None and pm.n_entries == 3
    (tmp_path / "Map.bin").write_bytes(b"\0" * 40)  # 16-byte legacy records
    with pytest.raises(ValueError, match="multiple of 24"):
        ma.load_map(tmp_path / "Map.bin", tmp_path / "nMap.bin")
