"""MI355X-native DMTz topology-correction path.

Given an original 3-D field f, a decompressed field g and an error bound eps, edit g within
|g' - f| <= eps until the discrete gradient of g on the Freudenthal lattice has the critical
cells (and optionally the separatrices) of f's. The work runs in HIP kernels for gfx950 behind
the C ABI of include/dmtz.h; this package is the Python host side.
"""
from .api import (CONVERGED, NOT_CONVERGED, NOT_RUN, STALLED, Context, CorrectionFlags, CorrectionResult, DmtzError,
                  classify, correct, correct_batch, correct_device, default_context, eps_from_rel, gradient,
                  paths, targets, walls)
from . import fields

__all__ = ["CONVERGED", "NOT_CONVERGED", "NOT_RUN", "STALLED", "Context", "CorrectionFlags", "CorrectionResult",
           "DmtzError", "classify", "correct", "correct_batch", "correct_device", "default_context",
           "eps_from_rel", "gradient", "paths", "targets", "walls", "fields"]
