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
  * lm_m4_20it: the 20-iteration iterate lowers F(x) = |B e^{ix}|^2;
  * greedy_m3_div16: every phase is a candidate, the objective is below that
    of zero phases;
  * gauss_newton_m3, gradient_descent_m2: the iterate lowers F(x) below a
    fraction of F(0);
  * apo_m3, long_m4, horn_m4: outputs camax-normalised (max_j |Re q_j| +
    |Im q_j| = 1); Long and Horn reproduce the target ratios.
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


def camax_norm(amp, phase):
    """max_j amp_j (|cos phi_j| + |sin phi_j|): 1 for outputs scaled by
    nalgebra's camax (max_j |Re q_j| + |Im q_j|) with unit source amplitudes."""
    return float(np.max(amp * (np.abs(np.cos(phase)) + np.abs(np.sin(phase)))))


def test_greedy_fixture(oracle):
    """Every phase is one of the phase_div candidates 2 pi i / div, and the
    objective sum_i |a_i - |p_i|| is below that of all phases zero."""
    g = load("greedy_m3_div16.npz")
    div = int(g["phase_div"])
    ph = oracle.greedy(1, g["src"], g["k"], g["a"], g["foci"], g["amps"], div, "f64")
    assert np.allclose(np.angle(np.exp(1j * (ph - g["out_phase"]))), 0, atol=1e-12)
    idx = np.mod(g["out_phase"], 2 * np.pi) / (2 * np.pi / div)
    assert np.max(np.abs(idx - np.round(idx))) <= 1e-9

    def objective(phase):
        return float(np.sum(np.abs(g["amps"] - np.abs(focal_f64(oracle, g["src"], g["k"], g["a"], g["foci"],
                                                                phase=phase)))))

    assert objective(g["out_phase"]) < 0.5 * objective(np.zeros_like(g["out_phase"]))


@pytest.mark.parametrize("name,fn,bound", [("gauss_newton_m3.npz", "gauss_newton", 0.01),
                                           ("gradient_descent_m2.npz", "gradient_descent", 0.05)])
def test_nls_fixtures(oracle, name, fn, bound):
    """The fixture's iterate lowers F(x) = |B e^{ix}|^2 below `bound` F(0)."""
    g = load(name)
    phase, x, it = getattr(oracle, fn)(1, g["src"], g["k"], g["a"], g["foci"], g["amps"], "f64",
                                       k_max=int(g["k_max"]))
    assert it == int(g["iterations"])
    assert np.allclose(np.angle(np.exp(1j * (phase - g["out_phase"]))), 0, atol=1e-9)
    f0, _ = oracle.lm_eval(1, g["src"], g["k"], g["a"], g["foci"], g["amps"], np.zeros_like(g["out_x"]))
    f1, _ = oracle.lm_eval(1, g["src"], g["k"], g["a"], g["foci"], g["amps"], g["out_x"])
    assert f1 < bound * f0


@pytest.mark.parametrize("name,spread", [("apo_m3.npz", None), ("long_m4.npz", 0.15), ("horn_m4.npz", 0.05)])
def test_amp_phase_fixtures(oracle, name, spread):
    """The oracle reproduces the fixture; the outputs are camax-normalised;
    Long and Horn reproduce the target amplitude ratios within `spread`."""
    g = load(name)
    if name.startswith("apo"):
        amp, ph, it = oracle.apo(1, g["src"], g["k"], g["a"], g["foci"], g["amps"], float(g["lam"]), "f64")
        assert it == int(g["iterations"])
    elif name.startswith("long"):
        amp, ph = oracle.long_(1, g["src"], g["k"], g["a"], g["foci"], g["amps"])
    else:
        amp, ph = oracle.horn(1, g["src"], g["k"], g["a"], g["foci"], g["amps"], seed=int(g["seed"]),
                              lam=np.float32(0.8), alpha=np.float32(1e-5))
    assert np.allclose(amp, g["out_amp"], rtol=1e-10, atol=1e-14)
    assert np.allclose(np.angle(np.exp(1j * (ph - g["out_phase"]))), 0, atol=1e-9)
    assert camax_norm(g["out_amp"], g["out_phase"]) == pytest.approx(1.0, abs=1e-12)
    if spread is not None:
        p = np.abs(focal_f64(oracle, g["src"], g["k"], g["a"], g["foci"], g["out_amp"], g["out_phase"]))
        r = p / g["amps"]
        assert r.min() >= (1 - spread) * r.max()
