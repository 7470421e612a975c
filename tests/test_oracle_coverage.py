"""Which branches of the oracle's restated repair rules and separatrix checks the parity corpus
reaches (the corpus on which the GPU parity tests compare the HIP path with the oracle).

The oracle counts every branch of fix_minimum / fix_saddle / fix_2saddle / fix_maximum
(main_3d.cu:2554-4303), the path checks (main_3d.cu:4437-5133) and the path and wall repairs
(main_3d.cu:5135-5413). Every branch must be reached, except four that no valid input reaches:

  S1_FP_FACE_A        (:2684-2699) the edge is critical in g, yet g would have a vertex below both
                      of its ends in a common triangle; pass E0 pairs such an edge with a
                      triangle (:2057-2071), so no g-critical edge meets the condition.
  S1_FN_FACE_FVERTEX  (:2919-2937) tests paired_edges[id] != -1 inside the branch taken because
                      paired_edges[id] == -1 (:2830): dead code.
  S2_FP_TET_W         (:3324-3337) the else of a test that compares the g-paired face's vertices
                      with the tet's fourth vertex, which is one of them: the self-comparison is
                      false, so the test always holds.
  ASW_NOT_PAIRED_BACK (:4486-4489) a face whose tet partner does not pair back: impossible in a
                      matching (paired_tetrahedras[t] = f implies paired_faces[f] = t).
The test asserts these four stay at zero, so a hit would expose the reasoning as wrong.
"""
import ctypes

import numpy as np

import corpus

UNREACHABLE = {"S1_FP_FACE_A", "S1_FN_FACE_FVERTEX", "S2_FP_TET_W", "ASW_NOT_PAIRED_BACK"}


def _path_targets(oracle, f, g, eps, rng, fl):
    L = oracle.lib()
    P, I = ctypes.c_void_p, ctypes.c_int64
    L.orc_path_targets_f32.argtypes = [P, P, I, I, I, ctypes.c_double, ctypes.c_double, P, P, P]
    D, H, W = f.shape
    tg = np.zeros(W * H * D, np.uint8)
    c2 = np.zeros(2, np.int64)
    L.orc_path_targets_f32(f.ctypes.data_as(P), g.ctypes.data_as(P), W, H, D, eps, rng, ctypes.byref(fl),
                           tg.ctypes.data_as(P), c2.ctypes.data_as(P))


def test_parity_corpus_reaches_every_branch(oracle):
    oracle.coverage(True)
    try:
        for _, f, g, eps, rng in corpus.gm_inputs() + corpus.fbm_inputs():
            for _, fl in corpus.FLAG_CASES + corpus.WALL_CASES:
                oracle.correct(f, g, eps, rng, oracle.flags(**fl))
        for _, f, g, eps, rng in corpus.edit_set_inputs():
            for edit_type in (0, 1):
                oracle.targets(f, g, eps, rng, oracle.flags(edit_type=edit_type))
        for _, f, g, eps, rng in corpus.path_inputs(oracle):
            for geometry in (1, 0):
                _path_targets(oracle, f, g, eps, rng, oracle.flags(preserve_vpath=1, preserve_geometry=geometry))
        for _, f, g, eps, rng in corpus.wall_inputs(oracle):
            for edit_type in (0, 1):
                oracle.wall_targets(f, g, eps, rng, oracle.flags(preserve_vpath=1, preserve_geometry=1,
                                                                 preserve_connectors=1, edit_type=edit_type))
    finally:
        cov = oracle.coverage(False)
    short = {k.split()[0]: v for k, v in cov.items()}
    assert len(short) >= 70
    missed = sorted(k for k, v in cov.items() if v == 0 and k.split()[0] not in UNREACHABLE)
    assert not missed, "branches the parity corpus does not reach:\n" + "\n".join(missed)
    hit_dead = sorted(k for k in UNREACHABLE if short.get(k, 0) > 0)
    assert not hit_dead, hit_dead
