"""Field files of the reference, through the library (include/dmtz.h dmtz_read_raw_f64,
dmtz_write_bin_f64, dmtz_read_bin_f64).

* raw: float64 values without header, x fastest (getdata2 main_3d.cu:153-181: the input field and
  the decompressed field are read this way, :6300, :6364);
* bin: an 8-byte count followed by the float64 values (saveVectorToBin main_3d.cu:1315-1327: the
  corrected field is written this way, :7166).
"""
import ctypes

import numpy as np

from . import _native as N

_ERR = {-1: "invalid argument", -3: "buffer too small", -7: "cannot read or write the file, or it is malformed"}


def _check(rc, what, path):
    if rc != 0:
        raise OSError("%s(%s): %s (%d)" % (what, path, _ERR.get(rc, "error"), rc))


def _read(fn, path, shape=None):
    L = N.lib()
    n = ctypes.c_int64(0)
    p = str(path).encode()
    _check(getattr(L, fn)(p, None, 0, ctypes.byref(n)), fn, path)
    out = np.empty(n.value, np.float64)
    _check(getattr(L, fn)(p, out.ctypes.data_as(ctypes.c_void_p), out.size, ctypes.byref(n)), fn, path)
    if shape is not None:
        out = out.reshape(shape)
    return out


def read_raw(path, shape=None):
    """A raw float64 field file (getdata2); shape = (nz, ny, nx) reshapes it."""
    return _read("dmtz_read_raw_f64", path, shape)


def read_bin(path, shape=None):
    """A field written by saveVectorToBin / write_bin."""
    return _read("dmtz_read_bin_f64", path, shape)


def write_bin(path, field):
    """Write a field as saveVectorToBin does: 8-byte count, then float64 values."""
    a = np.ascontiguousarray(field, dtype=np.float64).ravel()
    _check(N.lib().dmtz_write_bin_f64(str(path).encode(), a.ctypes.data_as(ctypes.c_void_p), a.size),
           "dmtz_write_bin_f64", path)
