"""Edit streams of a correction and the reference's cost figures (include/dmtz.h
dmtz_edit_stream_* / dmtz_edit_apply_*).

The reference reports, after a correction (cost main_3d.cu:1494-1646 for edit_type 1,
original_cost main_3d.cu:1649-1757 for edit_type 0): edit_ratio = edited / N (:1553), the index
and value streams it writes and compresses with the zstd command-line tool, the overall
compression ratio OCR = original bytes / (compressed data + compressed index + compressed values),
CR, the bit rates OBR = 64 / OCR and BR, and the PSNR before and after (:1597-1642).

The streams are built and decoded by the library on the device. Their compressed sizes come from
libzstd at level 3 (the zstd tool's default level) through pyarrow; the tool's own output frames
carry a 4-byte checksum more.
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _native as N
from .api import _device_arg, default_context
from .devmem import DeviceArray


@dataclass
class EditStream:
    edit_type: int
    n: int
    edited: int
    index: bytes  # 3-bit bitmap (edit_type 1) or int32 index differences (edit_type 0)
    edits: np.ndarray  # float64 values

    @property
    def edit_ratio(self):
        """The reference's edit_ratio: changed vertices over all vertices (main_3d.cu:1553)."""
        return self.edited / self.n


def edit_stream(g0, g1, counter=None, eps=0.0, edit_type=1, ctx=None):
    """The edit stream from the decompressed field g0 to the corrected field g1 (device buffers of
    one float type; counter: the uint8 counter output of the correction, needed for edit_type 1)."""
    ctx = ctx or default_context()
    p0, shape, tag = _device_arg(g0)
    p1, shape1, tag1 = _device_arg(g1)
    if shape1 != shape or tag1 != tag or tag not in ("f32", "f64"):
        raise ValueError("edit_stream: g0 and g1 must be float buffers of one shape and type")
    pc = _device_arg(counter)[0] if counter is not None else None
    n = int(np.prod(shape))
    fn = getattr(ctx._lib, "dmtz_edit_stream_" + tag)
    out = N.EditStream()
    rc = fn(ctx.handle, ctypes.c_void_p(p0), ctypes.c_void_p(p1), ctypes.c_void_p(pc) if pc else None, n,
            float(eps), int(edit_type), None, 0, None, 0, ctypes.byref(out))
    if rc not in (0, -3):
        ctx.check(rc, "dmtz_edit_stream")
    idx = DeviceArray((max(out.index_bytes, 1),), np.uint8)
    val = DeviceArray((max(out.n_edits, 1),), np.float64)
    rc = fn(ctx.handle, ctypes.c_void_p(p0), ctypes.c_void_p(p1), ctypes.c_void_p(pc) if pc else None, n,
            float(eps), int(edit_type), ctypes.c_void_p(idx.ptr), idx.size, ctypes.c_void_p(val.ptr), val.size,
            ctypes.byref(out))
    ctx.check(rc, "dmtz_edit_stream")
    index = idx.numpy()[: out.index_bytes].tobytes()
    edits = val.numpy()[: out.n_edits].copy()
    return EditStream(int(edit_type), n, int(out.edited), index, edits)


def edit_apply(g, stream, eps=0.0, ctx=None):
    """Decode an edit stream onto the device buffer g (the decompressed field, corrected in place)."""
    ctx = ctx or default_context()
    p, shape, tag = _device_arg(g)
    if tag not in ("f32", "f64") or int(np.prod(shape)) != stream.n:
        raise ValueError("edit_apply: g must be a float buffer of the stream's length")
    idx = DeviceArray.from_numpy(np.frombuffer(stream.index, np.uint8)) if stream.index else None
    val = DeviceArray.from_numpy(stream.edits) if len(stream.edits) else None
    fn = getattr(ctx._lib, "dmtz_edit_apply_" + tag)
    rc = fn(ctx.handle, ctypes.c_void_p(p), stream.n, float(eps), stream.edit_type,
            ctypes.c_void_p(idx.ptr) if idx else None, len(stream.index), ctypes.c_void_p(val.ptr) if val else None,
            len(stream.edits))
    ctx.check(rc, "dmtz_edit_apply")


def bitmap_rle(bitmap, ctx=None):
    """runLengthEncodeBitmap (main_3d.cu:1399-1426) on the device: (bit, run length) pairs of the
    bitmap's bits, most significant bit of each byte first, as an int32 array of shape (runs, 2).
    bitmap: bytes (uploaded) or a uint8 device buffer."""
    ctx = ctx or default_context()
    if isinstance(bitmap, (bytes, bytearray, memoryview)):
        raw = bytes(bitmap)
        if not raw:
            return np.zeros((0, 2), np.int32)
        buf = DeviceArray.from_numpy(np.frombuffer(raw, np.uint8))
    else:
        buf = bitmap
    p, shape, tag = _device_arg(buf)
    if tag != "u8":
        raise ValueError("bitmap_rle: the bitmap must be uint8")
    nbytes = int(np.prod(shape))
    runs = ctypes.c_int64(0)
    rc = ctx._lib.dmtz_bitmap_rle(ctx.handle, ctypes.c_void_p(p), nbytes, None, 0, ctypes.byref(runs))
    if rc not in (0, -3):
        ctx.check(rc, "dmtz_bitmap_rle")
    if runs.value == 0:
        return np.zeros((0, 2), np.int32)
    out = DeviceArray((runs.value, 2), np.int32)
    ctx.check(ctx._lib.dmtz_bitmap_rle(ctx.handle, ctypes.c_void_p(p), nbytes, ctypes.c_void_p(out.ptr), runs.value,
                                       ctypes.byref(runs)), "dmtz_bitmap_rle")
    return out.numpy()


def zstd_size(data, level=3):
    """Bytes of data compressed by libzstd at the given level (pyarrow)."""
    import pyarrow as pa

    if not len(data):
        return 0
    return len(pa.Codec("zstd", compression_level=level).compress(bytes(data), asbytes=True))


def psnr(original, x, value_range):
    """calculatePSNR main_3d.cu:1345-1352: -20 log10(sqrt(mse) / range); inf when equal."""
    d = np.asarray(original, np.float64).ravel() - np.asarray(x, np.float64).ravel()
    mse = float(np.mean(d * d))
    return float("inf") if mse == 0 else -20.0 * np.log10(np.sqrt(mse) / value_range)


def costs(stream, f, g0, g1, compressed_data_bytes=None, value_range=None):
    """The reference's report of a correction (main_3d.cu:1597-1642): edit_ratio, raw and zstd
    sizes of the two streams, psnr and fixed_psnr, and, given the compressed size of g0's own
    stream, OCR, CR, OBR and BR against the float64 original (8 bytes per value, as the reference's
    files hold)."""
    value_range = value_range if value_range is not None else float(np.max(f) - np.min(f))
    out = dict(edit_ratio=stream.edit_ratio, edited=stream.edited, index_bytes=len(stream.index),
               edits_bytes=8 * len(stream.edits), index_zstd=zstd_size(stream.index),
               edits_zstd=zstd_size(stream.edits.tobytes()), psnr=psnr(f, g0, value_range),
               fixed_psnr=psnr(f, g1, value_range))
    if compressed_data_bytes:
        original = 8 * stream.n
        ocr = original / (compressed_data_bytes + out["index_zstd"] + out["edits_zstd"])
        out.update(OCR=ocr, CR=original / compressed_data_bytes, OBR=64.0 / ocr,
                   BR=64.0 / (out["index_bytes"] / compressed_data_bytes))
    return out
