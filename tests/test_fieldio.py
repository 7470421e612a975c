"""Field files of the reference through the library's host code (no device): raw float64 files
as getdata2 reads them (main_3d.cu:153-181) and count-prefixed files as saveVectorToBin writes
them (main_3d.cu:1315-1327)."""
import struct

import numpy as np
import pytest

from dmtz_amd import fieldio, fields


def test_raw_file_is_read_as_numpy_wrote_it(tmp_path):
    f = fields.gaussian_mixture((5, 6, 7), 3, seed=1).astype(np.float64)
    p = tmp_path / "f.bin"
    f.tofile(p)  # raw little-endian doubles, x fastest
    assert np.array_equal(fieldio.read_raw(p, f.shape), f)
    with open(p, "ab") as fh:  # a trailing partial value is ignored
        fh.write(b"\x01\x02\x03")
    assert np.array_equal(fieldio.read_raw(p, f.shape), f)


def test_bin_file_layout_and_round_trip(tmp_path):
    g = np.linspace(-1.0, 2.0, 3 * 4 * 5).reshape(3, 4, 5)
    p = tmp_path / "g.bin"
    fieldio.write_bin(p, g)
    raw = p.read_bytes()
    assert len(raw) == 8 + 8 * g.size
    assert struct.unpack("<Q", raw[:8])[0] == g.size
    assert np.array_equal(np.frombuffer(raw[8:], "<f8"), g.ravel())
    assert np.array_equal(fieldio.read_bin(p, g.shape), g)
    fieldio.write_bin(p, np.zeros(0))
    assert fieldio.read_bin(p).size == 0


def test_malformed_and_missing_files_raise(tmp_path):
    p = tmp_path / "bad.bin"
    p.write_bytes(struct.pack("<Q", 10) + b"\x00" * 16)  # count says 10 values, file holds 2
    with pytest.raises(OSError):
        fieldio.read_bin(p)
    with pytest.raises(OSError):
        fieldio.read_raw(tmp_path / "missing.bin")
