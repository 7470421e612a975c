"""ctypes binding of libdmtz.so (include/dmtz.h). Fails loudly when the library is missing."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdmtz.so")

I64 = ctypes.c_int64
I32 = ctypes.c_int32
P = ctypes.c_void_p
D = ctypes.c_double


class Flags(ctypes.Structure):
    """dmtz_flags: the reference's argv flags (main_3d.cu:6264-6271) plus a sweep cap, the mode
    and the finisher switch."""
    _fields_ = [
        ("preserve_saddles", I32),
        ("preserve_vpath", I32),
        ("preserve_geometry", I32),
        ("filtration", I32),
        ("preserve_connectors", I32),
        ("preserve_type", I32),
        ("edit_type", I32),
        ("max_sweeps", I32),
        ("mode", I32),
        ("finish", I32),
    ]


class Counts(ctypes.Structure):
    """dmtz_counts: count_f_* of main_3d.cu:68-75."""
    _fields_ = [(n, I64) for n in ("min", "saddle1", "saddle2", "max", "ds_vpath", "as_vpath",
                                   "ds_wall", "as_wall")]

    def as_tuple(self):
        return tuple(int(getattr(self, n)) for n, _ in self._fields_)


class EditStream(ctypes.Structure):
    """dmtz_edit_stream: sizes of an edit stream (cost / original_cost, main_3d.cu:1494-1757)."""
    _fields_ = [("edited", I64), ("index_bytes", I64), ("n_edits", I64)]


def _field_struct(name, ctype):
    return type(name, (ctypes.Structure,), {"_fields_": [
        ("f", P), ("g", P), ("nx", I64), ("ny", I64), ("nz", I64), ("eps", D), ("range", D),
        ("on_device", I32), ("reserved", I32)]})


# dmtz_field_f32 / dmtz_field_f64: one field of a batch (host or device pointers).
FieldF32 = _field_struct("FieldF32", ctypes.c_float)
FieldF64 = _field_struct("FieldF64", ctypes.c_double)

_LIB = None


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "libdmtz.so is not built (%s); run `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `python -m dmtz_amd.build`" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    L.dmtz_ctx_create.argtypes = [ctypes.POINTER(P), ctypes.c_int]
    L.dmtz_ctx_create.restype = ctypes.c_int
    L.dmtz_ctx_destroy.argtypes = [P]
    L.dmtz_ctx_destroy.restype = None
    L.dmtz_last_error.argtypes = [P]
    L.dmtz_last_error.restype = ctypes.c_char_p
    for t in ("f32", "f64"):
        getattr(L, "dmtz_eps_from_rel_" + t).argtypes = [P, I64, D, ctypes.POINTER(D), ctypes.POINTER(D)]
        for name in ("dmtz_correct_", "dmtz_correct_dev_"):
            fn = getattr(L, name + t)
            fn.argtypes = [P, P, P, I64, I64, I64, D, D, ctypes.POINTER(Flags), ctypes.POINTER(Counts),
                           P, ctypes.POINTER(I32)]
            fn.restype = ctypes.c_int
        getattr(L, "dmtz_gradient_" + t).argtypes = [P, P, I64, I64, I64, P, P, P, P]
        getattr(L, "dmtz_paths_" + t).argtypes = [P, P, P, I64, I64, I64, ctypes.POINTER(Flags), P, P, P,
                                                  P, P, I64, P]
        getattr(L, "dmtz_targets_" + t).argtypes = [P, P, P, I64, I64, I64, D, D, ctypes.POINTER(Flags),
                                                    P, P]
        getattr(L, "dmtz_classify_" + t).argtypes = [P, P, I64, I64, I64, P, P, P]
        getattr(L, "dmtz_walls_" + t).argtypes = [P, P, P, I64, I64, I64, D, D, ctypes.POINTER(Flags), P, P, P,
                                                  I64, P]
        getattr(L, "dmtz_correct_batch_" + t).argtypes = [P, P, I32, ctypes.POINTER(Flags), P, P, P]
    L.dmtz_classify_noalloc_f32.argtypes = [P, P, I64, I64, I64, P, P, P]
    for t in ("f32", "f64"):
        getattr(L, "dmtz_edit_stream_" + t).argtypes = [P, P, P, P, I64, D, I32, P, I64, P, I64,
                                                        ctypes.POINTER(EditStream)]
        getattr(L, "dmtz_edit_apply_" + t).argtypes = [P, P, I64, D, I32, P, I64, P, I64]
    L.dmtz_bitmap_rle.argtypes = [P, P, I64, P, I64, ctypes.POINTER(I64)]
    L.dmtz_read_raw_f64.argtypes = [ctypes.c_char_p, P, I64, ctypes.POINTER(I64)]
    L.dmtz_write_bin_f64.argtypes = [ctypes.c_char_p, P, I64]
    L.dmtz_read_bin_f64.argtypes = [ctypes.c_char_p, P, I64, ctypes.POINTER(I64)]
    L.dmtz_trajectory.argtypes = [P, P, I64]
    L.dmtz_trajectory.restype = I64
    L.dmtz_phase_times.argtypes = [P, P]
    L.dmtz_kernel_stats.argtypes = [P, P, ctypes.c_int]
    L.dmtz_run_info.argtypes = [P, P]
    # learned warm start (include/dmtz_learned.h)
    U64 = ctypes.c_uint64
    F = ctypes.c_float
    L.dmtz_net_create.argtypes = [P, U64, ctypes.POINTER(P)]
    L.dmtz_net_destroy.argtypes = [P]
    L.dmtz_net_destroy.restype = None
    L.dmtz_net_params.argtypes = [P, P, P]
    L.dmtz_net_set_params.argtypes = [P, P, P]
    L.dmtz_net_features_f32.argtypes = [P, P, I64, I64, I64, D, D, P]
    L.dmtz_net_forward.argtypes = [P, P, P, I64, I64, I64, P]
    L.dmtz_net_gradient.argtypes = [P, P, P, P, I64, I64, I64, F, F, P, P]
    L.dmtz_net_train_step.argtypes = [P, P, P, P, I64, I64, I64, F, F, F, P]
    L.dmtz_correct_learned_f32.argtypes = [P, P, P, P, I64, I64, I64, D, D, ctypes.POINTER(Flags),
                                           ctypes.POINTER(Counts), P, ctypes.POINTER(I32), P]
    L.dmtz_correct_batch_learned_f32.argtypes = [P, P, P, I32, ctypes.POINTER(Flags), P, P, P, P]
    _LIB = L
    return L


# Every symbol declared in include/dmtz.h and include/dmtz_learned.h.
EXPORTS = (
    "dmtz_ctx_create", "dmtz_ctx_destroy", "dmtz_last_error",
    "dmtz_eps_from_rel_f32", "dmtz_eps_from_rel_f64",
    "dmtz_correct_f32", "dmtz_correct_f64", "dmtz_correct_dev_f32", "dmtz_correct_dev_f64",
    "dmtz_trajectory", "dmtz_phase_times", "dmtz_kernel_stats", "dmtz_run_info",
    "dmtz_gradient_f32", "dmtz_gradient_f64",
    "dmtz_targets_f32", "dmtz_targets_f64",
    "dmtz_paths_f32", "dmtz_paths_f64",
    "dmtz_classify_f32", "dmtz_classify_f64", "dmtz_classify_noalloc_f32",
    "dmtz_correct_batch_f32", "dmtz_correct_batch_f64",
    "dmtz_edit_stream_f32", "dmtz_edit_stream_f64", "dmtz_edit_apply_f32", "dmtz_edit_apply_f64", "dmtz_bitmap_rle",
    "dmtz_read_raw_f64", "dmtz_write_bin_f64", "dmtz_read_bin_f64",
    "dmtz_walls_f32", "dmtz_walls_f64",
    "dmtz_net_create", "dmtz_net_destroy", "dmtz_net_params", "dmtz_net_set_params",
    "dmtz_net_features_f32", "dmtz_net_forward", "dmtz_net_gradient", "dmtz_net_train_step",
    "dmtz_correct_learned_f32", "dmtz_correct_batch_learned_f32",
)
