"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the FP64 oracle; the reference holds no fixtures for this path).

CPU checks: the oracle still reproduces every fixture, and each fixture
satisfies a closed form that does not use the oracle's formulas:
  * field_t4010a1: with FocalPoint phases all terms arrive in phase at the
    focus, so |p(f)| = sum_j amp D(theta_j) e^{-alpha r_j} / r_j;
  * field_sphere: the independent numpy FP64 twin gives the same field;
  * gspat_m1: a single focus is reproduced exactly (|p(f)| = p0, gs_pat.rs);
  * gspat_m5: focal amplitudes within 15 % of the targets (GS-PAT's closing
    normalisation gamma / |gamma|^2 p0^2);
  * lm_m1: a single focus reaches the target amplitude after 200 iterations;
  * lm_m4_20it: the 20-iteration iterate lowers F(x) = |B e^{ix}|^2.
"""
import math
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def focal_f64(oracle, src, k, a, foci, amp=None, phase=None):
    s = src.copy()
    if amp is not None:
        s[:, 6] = amp
    if phase is not None:
        s[:, 7] = phase
    return oracle.np_field_f64(1, s, k, a, foci)


def test_oracle_reproduces_field_fixtures(oracle):
    for name, st in (("field_t4010a1.npz", 1), ("field_sphere.npz", 0)):
        g = load(name)
        f64 = oracle.field(st, g["src"], g["k"], g["a"], g["points"], "f64")
        f32 = oracle.field(st, g["src"], g["k"], g["a"], g["points"], "f32")
        assert np.allclose(f64, g["field_f64"], rtol=1e-13, atol=0), name
        assert np.array_equal(f32, g["field_f32"]), name


def test_field_t4010a1_focus_identity(oracle):
    g = load("field_t4010a1.npz")
    src, k, a = g["src"].astype(np.float64), g["k"].astype(np.float64), g["a"].astype(np.float64)
    f = g["points"][-1].astype(np.float64)
    d = f[None, :] - src[:, 0:3]
    r = np.linalg.norm(d, axis=1)
    theta = np.arccos(np.clip(np.sum(d * src[:, 3:6], 1) / r, -1, 1))
    want = np.sum(src[:, 6] * oracle.np_directivity(theta) * np.exp(-a * r) / r)
    # the FocalPoint phases are f32 (phase error ~1e-7 rad per term)
    assert abs(abs(g["field_f64"][-1]) - want) <= 1e-6 * want


def test_field_sphere_matches_numpy_twin(oracle):
    g = load("field_sphere.npz")
    twin = oracle.np_field_f64(0, g["src"], g["k"], g["a"], g["points"])
    assert np.allclose(twin, g["field_f64"], rtol=1e-10, atol=1e-12)


def test_gspat_fixtures(oracle):
    for name in ("gspat_m1.npz", "gspat_m5.npz"):
        g = load(name)
        amp, phase = oracle.gspat(1, g["src"], g["k"], g["a"], g["foci"], g["amps"], "f64")
        assert np.allclose(amp, g["out_amp"], rtol=1e-12, atol=0), name
        assert np.allclose(np.angle(np.exp(1j * (phase - g["out_phase"]))), 0, atol=1e-10), name
    g = load("gspat_m1.npz")
    p = np.abs(focal_f64(oracle, g["src"], g["k"], g["a"], g["foci"], g["out_amp"], g["out_phase"]))
    assert p[0] == pytest.approx(float(g["amps"][0]), rel=1e-9)
    g = load("gspat_m5.npz")
    p = np.abs(focal_f64(oracle, g["src"], g["k"], g["a"], g["foci"], g["out_amp"], g["out_phase"]))
    assert np.max(np.abs(p / g["amps"] - 1)) <= 0.15


def test_lm_fixtures(oracle):
    g = load("lm_m1.npz")
    phase, x, it = oracle.lm(1, g["src"], g["k"], g["a"], g["foci"], g["amps"], "f64", k_max=int(g["k_max"]))
    assert it == int(g["iterations"])
    assert np.allclose(np.angle(np.exp(1j * (phase - g["out_phase"]))), 0, atol=1e-9)
    p = np.abs(focal_f64(oracle, g["src"], g["k"], g["a"], g["foci"], phase=g["out_phase"]))
    assert p[0] == pytest.approx(float(g["amps"][0]), rel=1e-5)
    g = load("lm_m4_20it.npz")
    phase, x, it = oracle.lm(1, g["src"], g["k"], g["a"], g["foci"], g["amps"], "f64", k_max=int(g["k_max"]))
    assert it == int(g["iterations"])
    assert np.allclose(np.angle(np.exp(1j * (phase - g["out_phase"]))), 0, atol=1e-9)
    f0, _ = oracle.lm_eval(1, g["src"], g["k"], g["a"], g["foci"], g["amps"], np.zeros_like(g["out_x"]))
    f1, _ = oracle.lm_eval(1, g["src"], g["k"], g["a"], g["foci"], g["amps"], g["out_x"])
    assert f1 < 0.5 * f0
