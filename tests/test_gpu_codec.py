"""The edit stream (include/dmtz.h dmtz_edit_stream_* / dmtz_edit_apply_*) on the device.

Known answers: the 3-bit bitmap of encodeTo3BitBitmap (main_3d.cu:1446-1476) worked by hand for a
short index sequence, and the index differences of original_cost (main_3d.cu:1649-1757). Round
trips: the streams of real corrections decode back to the corrected field, and edited / N is the
fraction of changed values (edit_ratio, main_3d.cu:1553)."""
import numpy as np
import pytest

import dmtz_amd
from dmtz_amd import codec, fields
from dmtz_amd.devmem import DeviceArray

pytestmark = pytest.mark.gpu


def _kat_fields():
    """Eight vertices with g0 = 10, eps = 1: counters 2, 3, 4, 5 reach 9.5, 9.25, 9.125, 9.0625
    (steps -eps/2, -eps/4, ... of main_3d.cu:241), counter 6 holds the snap 9, counter 1 leaves the
    value unchanged."""
    counter = np.array([2, 3, 4, 5, 6, 1, 2, 3], np.uint8)
    reach = {1: 10.0, 2: 9.5, 3: 9.25, 4: 9.125, 5: 9.0625, 6: 9.0}
    g0 = np.full(8, 10.0, np.float32)
    g1 = np.array([reach[int(c)] for c in counter], np.float32)
    return g0, g1, counter


def test_bitmap_known_answer(dev):
    g0, g1, counter = _kat_fields()
    s = codec.edit_stream(DeviceArray.from_numpy(g0), DeviceArray.from_numpy(g1), DeviceArray.from_numpy(counter),
                          eps=1.0, edit_type=1, ctx=dev)
    # indices 2 3 4 5 6 0 2 3, three bits each, least significant first (encodeTo3BitBitmap):
    # byte 0 = 2 | 3 << 3 | (4 << 6 & 255) = 26; byte 1 = 4 >> 2 | 5 << 1 | 6 << 4 | 0 << 7 = 107;
    # byte 2 = 0 >> 1 | 2 << 2 | 3 << 5 = 104
    assert list(s.index) == [26, 107, 104]
    assert s.edits.tolist() == [-1.0] and s.edited == 7
    d = DeviceArray.from_numpy(g0)
    codec.edit_apply(d, s, eps=1.0, ctx=dev)
    assert d.numpy().tobytes() == g1.tobytes()


def test_index_differences_known_answer(dev):
    g0, g1, counter = _kat_fields()
    s = codec.edit_stream(DeviceArray.from_numpy(g0), DeviceArray.from_numpy(g1), None, edit_type=0, ctx=dev)
    # changed vertices 0 1 2 3 4 6 7: the first index, then the gaps (original_cost :1661-1670)
    assert np.frombuffer(s.index, np.int32).tolist() == [0, 1, 1, 1, 1, 2, 1]
    assert s.edits.tolist() == [-0.5, -0.75, -0.875, -0.9375, -1.0, -0.5, -0.75]
    d = DeviceArray.from_numpy(g0)
    codec.edit_apply(d, s, ctx=dev)
    assert d.numpy().tobytes() == g1.tobytes()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("fl", [dict(edit_type=1, preserve_vpath=1, preserve_geometry=1, finish=1),
                                dict(edit_type=0, preserve_vpath=1, preserve_geometry=1),
                                dict(edit_type=1, filtration=1)],
                         ids=["type1_finish", "type0", "type1_filtration"])
def test_stream_round_trip_of_corrections(dev, dtype, fl):
    for seed, shape, rel in [(0, (24, 20, 16), 1e-2), (1, (32, 32, 32), 3e-3)]:
        f = fields.gaussian_mixture(shape, 6, seed=seed, noise=0.02).astype(dtype)
        eps, rng = fields.eps_from_rel(f, rel)
        g0 = fields.quantize(f, eps)
        r = dmtz_amd.correct(f, g0, eps, rng, dmtz_amd.CorrectionFlags(**fl), ctx=dev, return_counter=True)
        assert r.status == 0
        print("stream", shape, dtype.__name__, fl, flush=True)
        s = codec.edit_stream(DeviceArray.from_numpy(g0), DeviceArray.from_numpy(r.g),
                              DeviceArray.from_numpy(r.counter), eps=eps, edit_type=fl["edit_type"], ctx=dev)
        assert s.edited == int(np.count_nonzero(r.g != g0))
        d = DeviceArray.from_numpy(g0)
        codec.edit_apply(d, s, eps=eps, ctx=dev)
        back = d.numpy()
        if dtype == np.float32:  # g1 - g0 of two float32 values is exact in double
            assert back.tobytes() == r.g.tobytes()
        else:  # g0 + (g1 - g0) in double can round: at most one unit in the last place
            assert np.all(np.abs(back - r.g) <= np.spacing(np.abs(r.g)))
        c = codec.costs(s, f, g0, r.g)
        assert c["fixed_psnr"] > 0 and c["index_zstd"] > 0


@pytest.mark.parametrize("geometry", [0, 1], ids=["critical_cells", "geometry"])
def test_streams_match_the_oracle_restatement(dev, oracle, geometry):
    """float64 corrections with the counter-driven edit (edit_type 1, the one the reference
    reports through cost, main_3d.cu:7175-7178) and without filtration or finisher, so every
    changed vertex's counter replays to its value: the edit_type 1 stream is the reference's cost
    stream byte for byte (bitmap and index-6 values, oracle.cost_streams), the edit_type 0 stream
    is original_cost's (oracle.original_cost_streams), and the run-length code of the bitmap is
    runLengthEncodeBitmap's (oracle.rle_bitmap)."""
    fl = dmtz_amd.CorrectionFlags(preserve_vpath=geometry, preserve_geometry=geometry, edit_type=1)
    for seed, shape, rel in [(0, (24, 20, 16), 1e-2), (2, (28, 28, 28), 1e-3), (3, (40, 17, 23), 3e-3)]:
        f = fields.gaussian_mixture(shape, 6, seed=seed, noise=0.02).astype(np.float64)
        eps, rng = fields.eps_from_rel(f, rel)
        g0 = fields.quantize(f, eps)
        r = dmtz_amd.correct(f, g0, eps, rng, fl, ctx=dev, return_counter=True)
        print("oracle streams", shape, rel, "status", r.status, flush=True)
        want = oracle.cost_streams(f, g0, r.g, r.counter, eps)
        s1 = codec.edit_stream(DeviceArray.from_numpy(g0), DeviceArray.from_numpy(r.g),
                               DeviceArray.from_numpy(r.counter), eps=eps, edit_type=1, ctx=dev)
        assert s1.edited == want["edited"] > 0
        assert s1.index == want["bitmap"]
        assert s1.edits.tobytes() == want["edits"].tobytes()
        w0 = oracle.original_cost_streams(g0, r.g)
        s0 = codec.edit_stream(DeviceArray.from_numpy(g0), DeviceArray.from_numpy(r.g), None, edit_type=0, ctx=dev)
        assert s0.index == w0["diffs"].tobytes() and s0.edits.tobytes() == w0["edits"].tobytes()
        assert np.array_equal(codec.bitmap_rle(s1.index, ctx=dev), oracle.rle_bitmap(s1.index))


def test_rle_matches_the_oracle(dev, oracle):
    rng = np.random.default_rng(5)
    cases = [bytes([0xF0, 0x0F]), bytes([0x00]), bytes([0xFF] * 9), bytes([0xAA]), bytes([0x01, 0x80])]
    for n, p in [(1, 0.5), (1000, 0.02), (65537, 0.5), (300000, 0.97), (1 << 20, 0.0)]:
        cases.append(np.packbits(rng.random(8 * n) < p).tobytes())
    for bm in cases:
        print("rle", len(bm), "bytes", flush=True)
        assert np.array_equal(codec.bitmap_rle(bm, ctx=dev), oracle.rle_bitmap(bm)), len(bm)
    assert codec.bitmap_rle(b"", ctx=dev).shape == (0, 2)
    assert codec.bitmap_rle(bytes([0xF0, 0x0F]), ctx=dev).tolist() == [[1, 4], [0, 8], [1, 4]]


@pytest.mark.timeout(600)
def test_streams_round_trip_at_512(dev):
    """The 512^3 fBm field at eps = 1e-2 range with the counter-driven edit, geometry and finisher:
    both edit streams decode the uncorrected field back to the corrected one bit for bit, the
    edited count is the number of changed values, and the run-length code of the bitmap covers
    its bits with alternating runs."""
    f = fields.fbm((512, 512, 512), seed=0)
    eps, rng = fields.eps_from_rel(f, 1e-2)
    g0 = fields.quantize(f, eps)
    fl = dmtz_amd.CorrectionFlags(preserve_vpath=1, preserve_geometry=1, finish=1, edit_type=1)
    print("512^3 correction for the streams", flush=True)
    fd, gd = DeviceArray.from_numpy(f), DeviceArray.from_numpy(g0)
    cd = DeviceArray(f.shape, np.uint8)
    r = dmtz_amd.correct_device(fd, gd, eps, rng, fl, ctx=dev, counter_out=cd)
    assert r.status == 0 and r.counts == (0,) * 8
    g1 = gd.numpy()
    changed = int(np.count_nonzero(g1 != g0))
    g0d = DeviceArray.from_numpy(g0)
    for edit_type in (1, 0):
        print("stream type", edit_type, flush=True)
        s = codec.edit_stream(g0d, gd, cd if edit_type == 1 else None, eps=eps, edit_type=edit_type, ctx=dev)
        assert s.edited == changed
        back = DeviceArray.from_numpy(g0)
        codec.edit_apply(back, s, eps=eps, ctx=dev)
        assert back.numpy().tobytes() == g1.tobytes()
        if edit_type == 1:
            runs = codec.bitmap_rle(s.index, ctx=dev)
            assert runs[:, 1].sum() == 8 * len(s.index) and (runs[1:, 0] != runs[:-1, 0]).all()
