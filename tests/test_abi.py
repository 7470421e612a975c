"""The C-ABI library builds/loads and exports every entry point include/*.h declares
(no device calls: runs without a GPU)."""
import os
import re

from dmtz_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in ("dmtz.h", "dmtz_learned.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        names |= set(re.findall(r"\b(dmtz_\w+)\s*\(", src))
    return sorted(names)


def test_header_symbols_exported():
    L = _native.lib()
    names = declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(_native.EXPORTS)


def test_eps_from_rel_through_the_abi():
    """dmtz_eps_from_rel_* (main_3d.cu:6301-6309: range = max f - min f, eps = rel * range) is
    host code of the library: checked against numpy on both field types."""
    import numpy as np

    import dmtz_amd
    from dmtz_amd import fields

    for dt in (np.float32, np.float64):
        f = fields.gaussian_mixture((9, 8, 7), 4, seed=3).astype(dt)
        f.flat[17] = -3.25
        f.flat[40] = 2.5
        eps, rng = dmtz_amd.eps_from_rel(f, 1e-3)
        assert rng == 5.75
        assert eps == 5.75 * 1e-3
        assert dmtz_amd.eps_from_rel(f, 0.5) == (2.875, 5.75)


def test_host_side_argument_checks():
    """Mismatched shapes and mixed batch dtypes are refused before any device call."""
    import numpy as np
    import pytest

    import dmtz_amd

    f = np.zeros((4, 5, 6), np.float32)
    g = np.zeros((4, 5, 5), np.float32)
    for call in (lambda: dmtz_amd.correct(f, g, 0.1, 1.0), lambda: dmtz_amd.targets(f, g, 0.1, 1.0),
                 lambda: dmtz_amd.paths(f, g), lambda: dmtz_amd.walls(f, g)):
        with pytest.raises(ValueError):
            call()
    with pytest.raises(TypeError):
        dmtz_amd.correct_batch([(f, f, 0.1, 1.0), (f.astype(np.float64), f.astype(np.float64), 0.1, 1.0)])
