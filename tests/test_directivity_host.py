"""The kernels' T4010A1 directivity formulas, evaluated on the host through
AFH_DirectivityT4010A1, against the oracle's FP64 evaluation of
core/wave_sources/t4010a1.rs:107-120 (theta = acos(c)).

FAST mode (cell table of quadratics in c): <= 1.5e-7 absolute (2.5 f32 ulp) away from the
segment boundaries, and every jump of the reference preserved.
ACCURATE mode (acos + reference binning + cubic in f32): <= 5e-7 absolute
(f32 rounding of theta amplified by the cubic's slope)."""
import ctypes
import math

import numpy as np
import pytest


def evaluate(afc, c, mode):
    c = np.ascontiguousarray(c, np.float32)
    out = np.empty_like(c)
    rc = afc.dll().AFH_DirectivityT4010A1(c.ctypes.data_as(ctypes.c_void_p), c.size, mode,
                                          out.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0
    return out


def reference(oracle, c):
    return oracle.np_directivity(np.arccos(np.clip(np.asarray(c, np.float64), -1, 1)))


BOUNDS = [math.cos(k * 10 * math.pi / 180) for k in range(2, 10)]


def away_from_bounds(c, margin=2e-6):
    m = np.ones(c.shape, bool)
    for b in BOUNDS + [0.0]:
        m &= np.abs(c.astype(np.float64) - b) > margin
    return m


@pytest.mark.parametrize("mode,tol", [(0, 1.5e-7), (1, 5e-7)])
def test_dense_sweep(afc, oracle, mode, tol):
    c = np.linspace(-1.0, 1.0, 400001).astype(np.float32)
    got = evaluate(afc, c, mode)
    want = reference(oracle, c)
    m = away_from_bounds(c)
    assert np.max(np.abs(got[m] - want[m])) <= tol


@pytest.mark.parametrize("mode", [0, 1])
def test_jumps_preserved(afc, oracle, mode):
    for b in BOUNDS[1:]:  # 30 .. 90 degrees (20 degrees is continuous)
        lo, hi = np.float32(b - 1e-5), np.float32(b + 1e-5)
        g = evaluate(afc, np.array([lo, hi], np.float32), mode)
        w = reference(oracle, np.array([lo, hi]))
        assert np.allclose(g, w, atol=2e-5), (b, g, w)


def test_constant_regions(afc):
    for mode in (0, 1):
        assert np.all(evaluate(afc, np.array([0.95, 0.99, 1.0], np.float32), mode) == 1.0)
        assert np.all(evaluate(afc, np.array([-0.01, -0.5, -1.0], np.float32), mode) == 0.0)
