"""The learned warm start (include/dmtz_learned.h): host side.

A four-layer 3x3x3 convolutional network, trained on (g, g') pairs that the plain correction
produces, predicts which vertices the plain loop will edit and by how much. Its prediction is
applied as a first edit of g (never past the snap of f - eps), and the plain loop then runs to
zero mismatches as usual, so it decides the result. Every kernel (features, forward, backward,
Adam, warm start) is the library's own HIP code; torch only holds device memory.
"""
import ctypes

import numpy as np

from . import _native as N
from .api import CorrectionFlags, CorrectionResult, _device_arg, correct_device, default_context

PARAMS = 4 * 6928
CHANNELS = 16
LAYER = 27 * CHANNELS * CHANNELS + CHANNELS  # W[27][16][16] (tap, input, output), then b[16]
EDIT_LOGIT_BIAS = 3 * LAYER + 27 * CHANNELS * CHANNELS  # last layer, output 0 (edit logit); output 1: size


def uniform_params(size=1.0):
    """Parameters of the trivial warm start that the network is measured against: all weights
    zero, so every vertex gets a positive edit logit and the edit size `size` (in units of eps):
    g is lowered by size * eps everywhere, never past the snap of f - eps."""
    p = np.zeros(PARAMS, np.float32)
    p[EDIT_LOGIT_BIAS] = 10.0
    p[EDIT_LOGIT_BIAS + 1] = size
    return p


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class Net:
    """A network on the context's device. Parameters: 4 x (W[27][16][16], b[16]) fp32."""

    def __init__(self, seed=0, ctx=None):
        self.ctx = ctx or default_context()
        self._lib = self.ctx._lib
        h = ctypes.c_void_p()
        self.ctx.check(self._lib.dmtz_net_create(self.ctx.handle, seed, ctypes.byref(h)), "dmtz_net_create")
        self.handle = h

    def close(self):
        if self.handle:
            self._lib.dmtz_net_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.zeros(PARAMS, np.float32)
        self.ctx.check(self._lib.dmtz_net_params(self.ctx.handle, self.handle, out.ctypes.data_as(ctypes.c_void_p)),
                       "dmtz_net_params")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == PARAMS
        self.ctx.check(self._lib.dmtz_net_set_params(self.ctx.handle, self.handle,
                                                     p.ctypes.data_as(ctypes.c_void_p)), "dmtz_net_set_params")

    def features(self, g, eps, rng):
        """Input channels of a float32 cuda field g [z, y, x]: int16 cuda tensor [z, y, x, 16]
        holding the bf16 bit patterns."""
        import torch

        assert g.is_cuda and g.dtype == torch.float32 and g.is_contiguous()
        nz, ny, nx = g.shape
        feat = torch.empty((nz, ny, nx, CHANNELS), dtype=torch.int16, device=g.device)
        torch.cuda.synchronize(g.device)
        self.ctx.check(self._lib.dmtz_net_features_f32(self.ctx.handle, _ptr(g), nx, ny, nz, float(eps), float(rng),
                                                       _ptr(feat)), "dmtz_net_features")
        return feat

    def forward(self, feat):
        """(edit logit, edit size / eps) per vertex: float32 cuda tensor [z, y, x, 2]."""
        import torch

        nz, ny, nx, _ = feat.shape
        out = torch.empty((nz, ny, nx, 2), dtype=torch.float32, device=feat.device)
        torch.cuda.synchronize(feat.device)
        self.ctx.check(self._lib.dmtz_net_forward(self.ctx.handle, self.handle, _ptr(feat), nx, ny, nz, _ptr(out)),
                       "dmtz_net_forward")
        return out

    def gradient(self, feat, labels, pos_weight=1.0, lam=1.0):
        """Parameter gradient (numpy float32) and mean loss terms (bce, size) without an update."""
        import torch

        nz, ny, nx, _ = feat.shape
        g = np.zeros(PARAMS, np.float32)
        loss = np.zeros(2, np.float64)
        torch.cuda.synchronize(feat.device)
        self.ctx.check(self._lib.dmtz_net_gradient(self.ctx.handle, self.handle, _ptr(feat), _ptr(labels), nx, ny, nz,
                                                   float(pos_weight), float(lam), g.ctypes.data_as(ctypes.c_void_p),
                                                   loss.ctypes.data_as(ctypes.c_void_p)), "dmtz_net_gradient")
        return g, loss

    def train_step(self, feat, labels, lr, pos_weight=1.0, lam=1.0):
        import torch

        nz, ny, nx, _ = feat.shape
        loss = np.zeros(2, np.float64)
        torch.cuda.synchronize(feat.device)
        self.ctx.check(self._lib.dmtz_net_train_step(self.ctx.handle, self.handle, _ptr(feat), _ptr(labels), nx, ny,
                                                     nz, float(pos_weight), float(lam), float(lr),
                                                     loss.ctypes.data_as(ctypes.c_void_p)), "dmtz_net_train_step")
        return loss


def labels(g, g_corrected, eps):
    """Training targets of one plain correction: (y, t) per vertex as a float32 cuda tensor
    [z, y, x, 2]; y = 1 where the loop changed g, t = (g - g') / eps."""
    import torch

    y = (g_corrected != g).to(torch.float32)
    t = ((g.double() - g_corrected.double()) / eps).to(torch.float32)
    return torch.stack([y, t], dim=-1).contiguous()


def training_pair(net, f, g, eps, rng, flags=None, ctx=None):
    """Features and labels of one field from a plain correction of (f, g) (cuda tensors)."""
    flags = flags or CorrectionFlags(preserve_vpath=1, preserve_geometry=1, finish=1)
    gc = g.clone()
    r = correct_device(f, gc, eps, rng, flags, ctx=ctx or net.ctx)
    assert r.converged, r.counts
    return net.features(g, eps, rng), labels(g, gc, eps)


def correct_learned(net, f, g, eps, rng, flags=None, counter_out=None):
    """Warm-started correction of float32 device fields f, g (devmem.DeviceArray or cuda torch
    tensors; g corrected in place). Returns (CorrectionResult, number of warm-start edits)."""
    ctx = net.ctx
    flags = flags or CorrectionFlags()
    pf, shape, tf = _device_arg(f)
    pg, gshape, tg = _device_arg(g)
    if tf != "f32" or tg != "f32" or gshape != shape or len(shape) != 3:
        raise ValueError("correct_learned: f and g must be 3-D float32 device buffers of one shape")
    pc = _device_arg(counter_out)[0] if counter_out is not None else None
    nz, ny, nx = shape
    counts = N.Counts()
    sweeps = ctypes.c_int32(0)
    warm = ctypes.c_int64(0)
    fl = flags.c()
    rc = net._lib.dmtz_correct_learned_f32(ctx.handle, net.handle, ctypes.c_void_p(pf), ctypes.c_void_p(pg), nx, ny,
                                           nz, float(eps), float(rng), ctypes.byref(fl), ctypes.byref(counts),
                                           ctypes.c_void_p(pc) if pc is not None else None,
                                           ctypes.byref(sweeps), ctypes.byref(warm))
    ctx.check(rc, "dmtz_correct_learned")
    res = CorrectionResult(g, counts.as_tuple(), int(sweeps.value), rc, None, ctx.trajectory(), ctx.phases(),
                           ctx.run_info())
    return res, int(warm.value)


def correct_batch_learned(net, items, flags=None):
    """The batched form with the warm start: items is a list of (f, g, eps, rng) float32 device
    buffers (g corrected in place). Returns (list of CorrectionResult, warm edits per field)."""
    from .api import _batch_results

    ctx = net.ctx
    flags = flags or CorrectionFlags()
    n = len(items)
    arr = (N.FieldF32 * n)()
    for i, (f, g, eps, rng) in enumerate(items):
        pf, shape, tf = _device_arg(f)
        pg, gshape, tg = _device_arg(g)
        if tf != "f32" or tg != "f32" or gshape != shape or len(shape) != 3:
            raise ValueError("correct_batch_learned: field %d must be 3-D float32 device buffers of one shape" % i)
        nz, ny, nx = shape
        arr[i] = N.FieldF32(pf, pg, nx, ny, nz, float(eps), float(rng), 1, 0)
    counts = (N.Counts * n)()
    status = np.zeros(n, np.int32)
    sweeps = np.zeros(n, np.int32)
    warm = np.zeros(n, np.int64)
    fl = flags.c()
    rc = net._lib.dmtz_correct_batch_learned_f32(ctx.handle, net.handle, arr, n, ctypes.byref(fl),
                                                 ctypes.cast(counts, ctypes.c_void_p),
                                                 status.ctypes.data_as(ctypes.c_void_p),
                                                 sweeps.ctypes.data_as(ctypes.c_void_p),
                                                 warm.ctypes.data_as(ctypes.c_void_p))
    return _batch_results(ctx, rc, [it[1] for it in items], counts, status, sweeps), warm
