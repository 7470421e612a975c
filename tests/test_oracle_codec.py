"""The oracle's restatement of the reference's edit-stream encoders (main_3d.cu:1399-1757), pinned
by encodings worked by hand and by library routines that fix the same bit orders.

  encodeTo3BitBitmap :1446-1476   three bits per value, least significant bit first, a partial
                                  last byte kept, values outside 0..6 rejected;
  runLengthEncodeBitmap :1399-1426  runs of the bits of each byte, most significant bit first;
  cost :1494-1510                 index = counter of a changed vertex, one value -(g0 - (f - bound))
                                  per index 6;
  original_cost :1649-1671        index differences and g1 - g0 of the changed vertices.
"""
import numpy as np
import pytest

from oracle import oracle as O


def test_encode3_known_answers():
    # 2 3 4 5 6 0 2 3: bits (LSB first) 010 110 001 101 011 000 010 110 -> bytes 26, 107, 104
    assert O.encode3([2, 3, 4, 5, 6, 0, 2, 3]) == bytes([26, 107, 104])
    # 2 0 6 0: 010 000 011 000 -> byte 0 = 2 + 128, byte 1 = 1 (the third bit of the 6 spills over)
    assert O.encode3([2, 0, 6, 0]) == bytes([0x82, 0x01])
    assert O.encode3([5]) == bytes([5])
    assert O.encode3([]) == b""
    assert O.encode3([1, 7]) is None and O.encode3([-1]) is None


def test_encode3_is_lsb_first_packing():
    """numpy's little-endian bit packing of the values' bits, least significant first, gives the
    same bytes for any length (partial last byte included)."""
    rng = np.random.default_rng(3)
    for n in [1, 2, 3, 7, 8, 9, 100, 1001]:
        v = rng.integers(0, 7, n)
        bits = ((v[:, None] >> np.arange(3)) & 1).astype(np.uint8).ravel()
        assert O.encode3(v) == np.packbits(bits, bitorder="little").tobytes(), n


def test_rle_known_answers():
    assert O.rle_bitmap(bytes([0xF0, 0x0F])).tolist() == [[1, 4], [0, 8], [1, 4]]
    assert O.rle_bitmap(bytes([0x00])).tolist() == [[0, 8]]
    assert O.rle_bitmap(bytes([0xFF, 0xFF])).tolist() == [[1, 16]]
    assert O.rle_bitmap(bytes([0xAA])).tolist() == [[1, 1], [0, 1]] * 4
    # 0x01 0x80: 0000000 1 1 0000000 -> (0, 7) (1, 2) (0, 7)
    assert O.rle_bitmap(bytes([0x01, 0x80])).tolist() == [[0, 7], [1, 2], [0, 7]]
    assert O.rle_bitmap(b"").shape == (0, 2)


def test_rle_expands_back_to_the_bitmap():
    """Expanding the runs and packing the bits most significant first (numpy's default order)
    restores the bitmap; runs alternate and cover 8 bits per byte."""
    rng = np.random.default_rng(4)
    for n, p in [(1, 0.5), (17, 0.5), (300, 0.05), (300, 0.95), (4096, 0.3)]:
        bm = np.packbits(rng.random(8 * n) < p).tobytes()
        r = O.rle_bitmap(bm)
        assert r[:, 1].sum() == 8 * n and (r[:, 1] > 0).all()
        assert (r[1:, 0] != r[:-1, 0]).all()
        bits = np.repeat(r[:, 0], r[:, 1]).astype(np.uint8)
        assert np.packbits(bits).tobytes() == bm


def test_cost_streams_known_answer():
    f = [1.0, 2.0, 3.0, 4.0, 5.0]
    g0 = [0.5, 2.0, 2.5, 4.0, 4.5]
    g1 = [0.75, 2.0, 2.0, 4.0, 4.25]
    s = O.cost_streams(f, g0, g1, [2, 5, 6, 0, 3], 1.0)
    assert s["index"].tolist() == [2, 0, 6, 0, 3]  # vertex 1 is unchanged: its counter is not written
    assert s["edited"] == 3
    assert s["edits"].tolist() == [-(2.5 - (3.0 - 1.0))]
    assert s["bitmap"] == O.encode3([2, 0, 6, 0, 3])


def test_original_cost_streams_known_answer():
    s = O.original_cost_streams([0, 0, 0, 0, 0, 0], [0, 1, 0, 0, 2, -1])
    assert s["diffs"].tolist() == [1, 3, 1] and s["edits"].tolist() == [1.0, 2.0, -1.0] and s["edited"] == 3
    s = O.original_cost_streams([1, 2], [1, 2])
    assert s["edited"] == 0 and s["diffs"].size == 0


@pytest.mark.parametrize("seed", [0, 1])
def test_original_cost_diffs_sum_to_indices(seed):
    rng = np.random.default_rng(seed)
    g0 = rng.random(5000)
    g1 = g0.copy()
    ch = rng.random(5000) < 0.3
    g1[ch] -= 0.01
    s = O.original_cost_streams(g0, g1)
    assert np.array_equal(np.cumsum(s["diffs"]), np.flatnonzero(ch))
    assert np.array_equal(s["edits"], (g1 - g0)[ch])
