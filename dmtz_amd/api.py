"""Host-side interface of the correction path over the C ABI of include/dmtz.h.

The reference exposes this path only as a CLI (main_3d.cu:6255-6271):
``main_3d "name,W,H,D" rel_eb compressor preserve_saddles preserve_vpath preserve_geometry
filtration preserve_connectors preserve_type edit_type``. ``correct`` keeps those flags with the
same names and meanings, takes f and g as arrays instead of files, and returns the corrected
field with the mismatch counters of main_3d.cu:68-75 instead of printing them.
"""
import ctypes
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _native as N


class DmtzError(RuntimeError):
    pass


# Non-negative status codes of a correction (include/dmtz.h), and the batch status of a field
# that never started because another field failed.
CONVERGED, NOT_CONVERGED, STALLED = 0, 1, 2
NOT_RUN = -6


@dataclass
class CorrectionFlags:
    preserve_saddles: int = 1
    preserve_vpath: int = 0
    preserve_geometry: int = 0
    filtration: int = 0
    preserve_connectors: int = 0
    preserve_type: int = 0
    edit_type: int = 0
    max_sweeps: int = 100000
    # 1: on a stall (a fixed point with mismatches left, where the reference loops forever),
    # snap growing boxes around the stalled edit sets to f - eps, and restore f everywhere once
    # the boxes would cover the grid (no counterpart in the reference); the run then ends with
    # zero mismatches.
    finish: int = 0

    def c(self):
        return N.Flags(self.preserve_saddles, self.preserve_vpath, self.preserve_geometry,
                       self.filtration, self.preserve_connectors, self.preserve_type,
                       self.edit_type, self.max_sweeps, 0, self.finish)


@dataclass
class CorrectionResult:
    g: object
    counts: tuple
    sweeps: int
    status: int
    counter: Optional[np.ndarray] = None
    trajectory: Optional[np.ndarray] = None
    phases: dict = field(default_factory=dict)
    info: dict = field(default_factory=dict)

    @property
    def converged(self):
        return self.status == 0 and all(c == 0 for c in self.counts)

    @property
    def stalled(self):
        return self.status == STALLED


class Context:
    """A dmtz_ctx: one HIP stream and the scratch memory of the correction loop."""

    def __init__(self, device=0):
        self._lib = N.lib()
        h = ctypes.c_void_p()
        rc = self._lib.dmtz_ctx_create(ctypes.byref(h), int(device))
        if rc != 0:
            raise DmtzError("dmtz_ctx_create failed (%d)" % rc)
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.dmtz_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    def error(self):
        return self._lib.dmtz_last_error(self._h).decode()

    def check(self, rc, what):
        if rc < 0:
            raise DmtzError("%s failed (%d): %s" % (what, rc, self.error()))
        return rc

    def trajectory(self):
        n = self._lib.dmtz_trajectory(self._h, None, 0)
        out = np.zeros(max(n, 1) * 8, np.int64)
        self._lib.dmtz_trajectory(self._h, out.ctypes.data_as(ctypes.c_void_p), n)
        return out[: n * 8].reshape(n, 8)

    def kernel_stats(self, reset=False):
        """Dense detection kernel: dict(launches, ms, bytes, voxels) measured with HIP events."""
        out = np.zeros(4, np.float64)
        self._lib.dmtz_kernel_stats(self._h, out.ctypes.data_as(ctypes.c_void_p), 1 if reset else 0)
        return dict(zip(("launches", "ms", "bytes", "voxels"), out.tolist()))

    def run_info(self):
        """Last correction: dict(stalls, finished, radius, stars) (finisher outcome and the
        number of lower stars re-evaluated incrementally)."""
        out = np.zeros(4, np.int64)
        self._lib.dmtz_run_info(self._h, out.ctypes.data_as(ctypes.c_void_p))
        return dict(zip(("stalls", "finished", "radius", "stars"), (int(x) for x in out)))

    def phases(self):
        out = np.zeros(6, np.float64)
        self._lib.dmtz_phase_times(self._h, out.ctypes.data_as(ctypes.c_void_p))
        names = ("gradient_pairing", "get_f_cp", "get_f_path", "fix_f_cp", "fix_f_vpath", "total")
        return dict(zip(names, out.tolist()))


_DEFAULT_CTX = None


def default_context():
    global _DEFAULT_CTX
    if _DEFAULT_CTX is None:
        _DEFAULT_CTX = Context(0)
    return _DEFAULT_CTX


def _tag(dtype):
    if dtype == np.float32:
        return "f32"
    if dtype == np.float64:
        return "f64"
    raise TypeError("fields must be float32 or float64")


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _same_shape(f, g, what):
    if tuple(g.shape) != tuple(f.shape):
        raise ValueError("%s: g has shape %s, f has %s" % (what, tuple(g.shape), tuple(f.shape)))
    if len(f.shape) != 3:
        raise ValueError("%s: fields must be 3-D [z, y, x]" % what)


def _single_runtime():
    """Torch tensors are handed to the library only when torch and the library run on one HIP
    runtime. torch bundles its own libamdhip64; when torch is imported first, the library binds
    to that copy (same soname), otherwise the process holds two runtimes whose streams do not
    order against each other."""
    paths = set()
    with open("/proc/self/maps") as fh:
        for line in fh:
            parts = line.split()
            if len(parts) >= 6 and "libamdhip64.so" in parts[-1]:
                paths.add(parts[-1])
    if len(paths) > 1:
        raise DmtzError("two HIP runtimes are loaded (%s): import torch before dmtz_amd, or pass "
                        "dmtz_amd.devmem.DeviceArray buffers instead of torch tensors" % ", ".join(sorted(paths)))


def _device_arg(x):
    """(pointer, shape, dtype tag) of a device buffer: a devmem.DeviceArray or a contiguous cuda
    torch tensor (whose pending work is finished first)."""
    if hasattr(x, "ptr") and hasattr(x, "dtype") and not hasattr(x, "data_ptr"):
        tag = {np.dtype(np.float32): "f32", np.dtype(np.float64): "f64", np.dtype(np.uint8): "u8"}
        return x.ptr, tuple(x.shape), tag.get(np.dtype(x.dtype), "other")
    import torch

    assert x.is_cuda and x.is_contiguous()
    _single_runtime()
    torch.cuda.synchronize(x.device)
    tag = {torch.float32: "f32", torch.float64: "f64", torch.uint8: "u8"}.get(x.dtype, "other")
    return x.data_ptr(), tuple(x.shape), tag


def correct(f, g, eps, rng, flags=None, ctx=None, return_counter=False):
    """Correct g toward f within eps until its discrete gradient matches f's.

    f, g: numpy arrays indexed [z, y, x] (x fastest), same float dtype. Returns a
    CorrectionResult with the corrected copy of g."""
    _same_shape(f, g, "correct")
    ctx = ctx or default_context()
    flags = flags or CorrectionFlags()
    f = np.ascontiguousarray(f)
    gout = np.array(g, dtype=f.dtype, copy=True, order="C")
    nz, ny, nx = f.shape
    counts = N.Counts()
    sweeps = ctypes.c_int32(0)
    counter = np.zeros(f.shape, np.uint8) if return_counter else None
    fn = getattr(ctx._lib, "dmtz_correct_" + _tag(f.dtype))
    fl = flags.c()
    rc = fn(ctx.handle, _p(f), _p(gout), nx, ny, nz, float(eps), float(rng), ctypes.byref(fl),
            ctypes.byref(counts), _p(counter) if counter is not None else None, ctypes.byref(sweeps))
    ctx.check(rc, "dmtz_correct")
    return CorrectionResult(gout, counts.as_tuple(), int(sweeps.value), rc, counter,
                            ctx.trajectory(), ctx.phases(), ctx.run_info())


def correct_device(f, g, eps, rng, flags=None, ctx=None, counter_out=None):
    """Correct device-resident fields in place (g is modified). f, g: devmem.DeviceArray buffers
    or contiguous cuda torch tensors [z, y, x] of float32 or float64 on the context's device.
    counter_out: optional uint8 device buffer of the same shape."""
    ctx = ctx or default_context()
    flags = flags or CorrectionFlags()
    pf, shape, tag = _device_arg(f)
    pg, gshape, gtag = _device_arg(g)
    if gshape != shape or gtag != tag or tag not in ("f32", "f64") or len(shape) != 3:
        raise ValueError("correct_device: f and g must be 3-D float32 or float64 buffers of one shape")
    pc = None
    if counter_out is not None:
        pc, cshape, _ = _device_arg(counter_out)
        if int(np.prod(cshape)) != int(np.prod(shape)):
            raise ValueError("correct_device: counter_out must hold one byte per vertex")
    nz, ny, nx = shape
    counts = N.Counts()
    sweeps = ctypes.c_int32(0)
    fn = getattr(ctx._lib, "dmtz_correct_dev_" + tag)
    fl = flags.c()
    rc = fn(ctx.handle, ctypes.c_void_p(pf), ctypes.c_void_p(pg), nx, ny, nz,
            float(eps), float(rng), ctypes.byref(fl), ctypes.byref(counts),
            ctypes.c_void_p(pc) if pc is not None else None, ctypes.byref(sweeps))
    ctx.check(rc, "dmtz_correct_dev")
    return CorrectionResult(g, counts.as_tuple(), int(sweeps.value), rc, None, ctx.trajectory(),
                            ctx.phases(), ctx.run_info())


def gradient(field_, ctx=None):
    """The reference's four pairing arrays (pv, pe, pf, pt) of one field, from the device."""
    ctx = ctx or default_context()
    field_ = np.ascontiguousarray(field_)
    nz, ny, nx = field_.shape
    Nv = nx * ny * nz
    E = ((nx - 1) * ny * nz + nx * (ny - 1) * nz + nx * ny * (nz - 1) + (nx - 1) * (ny - 1) * nz
         + (nx - 1) * ny * (nz - 1) + nx * (ny - 1) * (nz - 1) + (nx - 1) * (ny - 1) * (nz - 1))
    F = 2 * ((nx - 1) * (ny - 1) * nz + nx * (ny - 1) * (nz - 1) + (nx - 1) * ny * (nz - 1)) \
        + 6 * (nx - 1) * (ny - 1) * (nz - 1)
    T = 6 * (nx - 1) * (ny - 1) * (nz - 1)
    outs = [np.empty(n, np.int64) for n in (Nv, E, F, T)]
    fn = getattr(ctx._lib, "dmtz_gradient_" + _tag(field_.dtype))
    ctx.check(fn(ctx.handle, _p(field_), nx, ny, nz, *[_p(o) for o in outs]), "dmtz_gradient")
    return tuple(outs)


def targets(f, g, eps, rng, flags=None, ctx=None):
    """One detection + repair phase: (edit-set mask [z,y,x] bool, counts4)."""
    _same_shape(f, g, "targets")
    ctx = ctx or default_context()
    flags = flags or CorrectionFlags()
    f = np.ascontiguousarray(f)
    g = np.ascontiguousarray(g, dtype=f.dtype)
    nz, ny, nx = f.shape
    tg = np.zeros(f.shape, np.uint8)
    c4 = np.zeros(4, np.int64)
    fn = getattr(ctx._lib, "dmtz_targets_" + _tag(f.dtype))
    fl = flags.c()
    ctx.check(fn(ctx.handle, _p(f), _p(g), nx, ny, nz, float(eps), float(rng), ctypes.byref(fl),
                 _p(tg), _p(c4)), "dmtz_targets")
    return tg.astype(bool), tuple(int(x) for x in c4)


def eps_from_rel(f, rel):
    """(eps, range) as main_3d.cu:6301-6309, computed by the library."""
    f = np.ascontiguousarray(f)
    e = ctypes.c_double()
    r = ctypes.c_double()
    fn = getattr(N.lib(), "dmtz_eps_from_rel_" + _tag(f.dtype))
    if fn(_p(f), f.size, float(rel), ctypes.byref(e), ctypes.byref(r)) != 0:
        raise DmtzError("dmtz_eps_from_rel failed")
    return e.value, r.value


def paths(f, g, flags=None, ctx=None, with_targets=False):
    """Separatrix mismatch lists on (f, g): dict(ds=(edge ids), ds_rec, as_=(tet ids), as_rec,
    targets=edit-set mask or None)."""
    _same_shape(f, g, "paths")
    ctx = ctx or default_context()
    flags = flags or CorrectionFlags(preserve_vpath=1, preserve_geometry=1)
    f = np.ascontiguousarray(f)
    g = np.ascontiguousarray(g, dtype=f.dtype)
    nz, ny, nx = f.shape
    cap = 4 * f.size + 16
    bufs = [np.zeros(cap, np.int64) for _ in range(4)]
    c2 = np.zeros(2, np.int64)
    tg = np.zeros(f.shape, np.uint8) if with_targets else None
    fn = getattr(ctx._lib, "dmtz_paths_" + _tag(f.dtype))
    fl = flags.c()
    ctx.check(fn(ctx.handle, _p(f), _p(g), nx, ny, nz, ctypes.byref(fl), _p(c2), *[_p(b) for b in bufs],
                 cap, _p(tg) if tg is not None else None), "dmtz_paths")
    nds, nas = int(c2[0]), int(c2[1])
    return dict(ds=bufs[0][:nds], ds_rec=bufs[1][:nds], as_=bufs[2][:nas], as_rec=bufs[3][:nas],
                targets=tg.astype(bool) if tg is not None else None)


def walls(f, g, eps=0.0, rng=0.0, flags=None, ctx=None, with_targets=False):
    """2-saddle walls on (f, g) (get_false_ds_wall main_3d.cu:5064-5133, made deterministic):
    dict(face=(f-critical face ids), edge=(first edge whose f and g partners differ),
    targets=edit set of fix_ds_wall main_3d.cu:5297-5413 or None). eps and range enter the
    repair's threshold only."""
    _same_shape(f, g, "walls")
    ctx = ctx or default_context()
    f = np.ascontiguousarray(f)
    g = np.ascontiguousarray(g, dtype=f.dtype)
    nz, ny, nx = f.shape
    cap = 2 * f.size + 16
    face = np.zeros(cap, np.int64)
    edge = np.zeros(cap, np.int64)
    n = np.zeros(1, np.int64)
    tg = np.zeros(f.shape, np.uint8) if with_targets else None
    fl = (flags or CorrectionFlags()).c()
    fn = getattr(ctx._lib, "dmtz_walls_" + _tag(f.dtype))
    ctx.check(fn(ctx.handle, _p(f), _p(g), nx, ny, nz, float(eps), float(rng), ctypes.byref(fl), _p(n), _p(face),
                 _p(edge), cap, _p(tg) if tg is not None else None), "dmtz_walls")
    k = int(n[0])
    return dict(face=face[:k], edge=edge[:k], targets=tg.astype(bool) if tg is not None else None)


def classify(field_, ctx=None, masks=False):
    """The classifier alone: critical cells of one field under the reference's gradient
    (main_3d.cu:1930-2316). Returns (minima, 1-saddles, 2-saddles, maxima) counts and, with
    masks=True, per-vertex masks of the owned critical local cells (lo: 51 bits, hi: 24 bits)."""
    ctx = ctx or default_context()
    field_ = np.ascontiguousarray(field_)
    nz, ny, nx = field_.shape
    counts = np.zeros(4, np.int64)
    lo = np.zeros(field_.shape, np.uint64) if masks else None
    hi = np.zeros(field_.shape, np.uint32) if masks else None
    fn = getattr(ctx._lib, "dmtz_classify_" + _tag(field_.dtype))
    rc = fn(ctx.handle, _p(field_), nx, ny, nz, _p(counts), _p(lo) if masks else None,
            _p(hi) if masks else None)
    ctx.check(rc, "dmtz_classify")
    out = tuple(int(x) for x in counts)
    return (out, lo, hi) if masks else out


def correct_batch(items, flags=None, ctx=None):
    """Correct several fields with the same flags in one call. items: list of (f, g, eps, rng),
    either numpy arrays (g is copied, not modified) or device buffers on the context's device
    (devmem.DeviceArray or contiguous cuda tensors; g is corrected in place, no copies). Returns a
    list of CorrectionResult (without trajectories)."""
    if not items:
        return []
    if isinstance(items[0][0], np.ndarray):
        dt0 = np.asarray(items[0][0]).dtype
        for f, g, _, _ in items:
            if np.asarray(f).dtype != dt0 or np.asarray(g).dtype != dt0:
                raise TypeError("correct_batch: every field must have the dtype of the first (%s)" % dt0)
            _same_shape(f, g, "correct_batch")
    ctx = ctx or default_context()
    flags = flags or CorrectionFlags()
    if not isinstance(items[0][0], np.ndarray):
        return _correct_batch_device(items, flags, ctx)
    dt = np.ascontiguousarray(items[0][0]).dtype
    tag = _tag(dt)
    FieldT = N.FieldF32 if tag == "f32" else N.FieldF64
    keep, arr = [], (FieldT * len(items))()
    for i, (f, g, eps, rng) in enumerate(items):
        f = np.ascontiguousarray(f, dtype=dt)
        g = np.array(g, dtype=dt, copy=True, order="C")
        keep.append((f, g))
        nz, ny, nx = f.shape
        arr[i] = FieldT(_p(f).value, _p(g).value, nx, ny, nz, float(eps), float(rng), 0, 0)
    n = len(items)
    counts = (N.Counts * n)()
    status = np.zeros(n, np.int32)
    sweeps = np.zeros(n, np.int32)
    fl = flags.c()
    rc = getattr(ctx._lib, "dmtz_correct_batch_" + tag)(ctx.handle, arr, n, ctypes.byref(fl),
                                                         ctypes.cast(counts, ctypes.c_void_p),
                                                         _p(status), _p(sweeps))
    return _batch_results(ctx, rc, [keep[i][1] for i in range(n)], counts, status, sweeps)


def _correct_batch_device(items, flags, ctx):
    """Batch over device buffers (devmem.DeviceArray or cuda torch tensors), corrected in place."""
    tag = _device_arg(items[0][0])[2]
    if tag not in ("f32", "f64"):
        raise TypeError("correct_batch: fields must be float32 or float64")
    FieldT = N.FieldF32 if tag == "f32" else N.FieldF64
    arr = (FieldT * len(items))()
    for i, (f, g, eps, rng) in enumerate(items):
        pf, shape, tf = _device_arg(f)
        pg, gshape, tg = _device_arg(g)
        if tf != tag or tg != tag:
            raise TypeError("correct_batch: every field must have the dtype of the first (%s)" % tag)
        if gshape != shape or len(shape) != 3:
            raise ValueError("correct_batch: f and g of field %d must be 3-D of one shape" % i)
        nz, ny, nx = shape
        arr[i] = FieldT(pf, pg, nx, ny, nz, float(eps), float(rng), 1, 0)
    n = len(items)
    counts = (N.Counts * n)()
    status = np.zeros(n, np.int32)
    sweeps = np.zeros(n, np.int32)
    fl = flags.c()
    rc = getattr(ctx._lib, "dmtz_correct_batch_" + tag)(ctx.handle, arr, n, ctypes.byref(fl),
                                                         ctypes.cast(counts, ctypes.c_void_p),
                                                         _p(status), _p(sweeps))
    return _batch_results(ctx, rc, [items[i][1] for i in range(n)], counts, status, sweeps)


def _batch_results(ctx, rc, gs, counts, status, sweeps):
    """Per-field results of a batch call. On an error the DmtzError carries them as .results:
    fields that never started have status NOT_RUN (include/dmtz.h DMTZ_E_NOT_RUN)."""
    out = [CorrectionResult(gs[i], counts[i].as_tuple(), int(sweeps[i]), int(status[i])) for i in range(len(gs))]
    if rc < 0:
        err = DmtzError("dmtz_correct_batch failed (%d): %s" % (rc, ctx.error()))
        err.results = out
        raise err
    return out
