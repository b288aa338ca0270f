"""Pins the oracle's field path (oracle/afc_oracle.c) with checks that do not
restate its formulas: datasheet values, physical laws, invariants, brute-force
enumeration and an independent numpy implementation.

Reference: core/wave_sources/t4010a1.rs:18-120, core/sound_speed.rs:22-24,
system/uniform_medium.rs:111-128, observe_area/grid/{builder,grid_area}.rs,
accurate/system_acc.rs:26-54.
"""
import math

import numpy as np
import pytest

DEG = math.pi / 180.0


def rel_l2(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# ---------------------------------------------------------------- directivity
def test_directivity_segment_starts_are_datasheet_db_steps(oracle):
    # the T4010A1 table starts each 10-degree segment at 0, -1, -3, -6, -9,
    # -12, -14 dB relative to on-axis (amplitude 10^(-dB/20)), from 20 deg on
    db = {3: 0.0, 4: 1.0, 5: 3.0, 6: 6.0, 7: 9.0, 8: 12.0, 9: 14.0}
    for i, d in db.items():
        theta = (i - 1) * 10 * DEG + 1e-6  # just inside segment i
        want = 10 ** (-d / 20.0)
        assert oracle.directivity(theta, "f64") == pytest.approx(want, abs=2e-6), i
        assert oracle.directivity(theta, "f32") == pytest.approx(want, abs=3e-6), i


def test_directivity_constant_regions_and_nan(oracle):
    for prec in ("f32", "f64"):
        for deg in (0.0, 5.0, 10.0, 15.0, 19.99):
            assert oracle.directivity(deg * DEG, prec) == 1.0
        for deg in (90.01, 120.0, 180.0):
            assert oracle.directivity(deg * DEG, prec) == 0.0
        # Rust's saturating `as usize` maps NaN to index 0 -> 1.0
        assert oracle.directivity(float("nan"), prec) == 1.0


def test_directivity_segment_ends(oracle):
    # values at the end of each segment (theta = i * 10 deg, still segment i);
    # the jumps at 30, 40, 60 and 90 degrees are properties of the table
    ends = {3: 0.87524, 4: 0.70497, 5: 0.50119, 6: 0.30791, 7: 0.25119, 8: 0.19953, 9: 0.16370}
    for i, v in ends.items():
        assert oracle.directivity(i * 10 * DEG - 1e-7, "f64") == pytest.approx(v, abs=2e-5), i
    assert oracle.directivity(30 * DEG + 1e-6, "f64") - oracle.directivity(30 * DEG - 1e-6, "f64") > 0.015
    assert oracle.directivity(90 * DEG, "f64") > 0.16
    assert oracle.directivity(90 * DEG + 1e-6, "f64") == 0.0


def test_numpy_twin_directivity_agrees(oracle):
    th = np.linspace(0, math.pi, 20001)
    c = np.array([oracle.directivity(float(t), "f64") for t in th])
    assert np.max(np.abs(c - oracle.np_directivity(th))) < 1e-12


# ---------------------------------------------------------------- medium
def test_sound_speed(oracle):
    assert oracle.sound_speed(273.15, "f64") == pytest.approx(331300.0, rel=1e-15)
    # c scales with sqrt(T)
    assert oracle.sound_speed(4 * 273.15, "f64") == pytest.approx(2 * 331300.0, rel=1e-15)
    assert oracle.sound_speed(oracle.TEMP_340, "f32") == pytest.approx(340000.0, rel=1e-6)


def test_attenuation_f32_matches_f64_and_scales(oracle):
    a32 = oracle.attenuation_coef(40e3, 30, 1, 1, 300.0, "f32")
    a64 = oracle.attenuation_coef(40e3, 30, 1, 1, 300.0, "f64")
    assert a32 == pytest.approx(a64, rel=1e-5)
    # absorption grows with frequency, close to f^2 well below the relaxation frequencies
    lo = [oracle.attenuation_coef(f, 30, 1, 1, 300.0, "f64") for f in (0.25, 0.5)]
    assert lo[1] / lo[0] == pytest.approx(4.0, rel=1e-3)
    assert oracle.attenuation_coef(80e3, 30, 1, 1, 300.0, "f64") > a64 > 0


# ISO 9613-2 Table 2: atmospheric attenuation coefficient (dB/km) at 70 %
# relative humidity, octave midband frequencies 63 Hz .. 8 kHz (the exact
# frequencies 1000 * 10^(3b/10) of ISO 9613-1). core/attenuation.rs:31-48 is
# the ISO 9613-1 (Bass et al.) formula, so it must reproduce the table.
ISO9613_70RH = {
    293.15: [0.1, 0.3, 1.1, 2.8, 5.0, 9.0, 22.9, 76.6],
    283.15: [0.1, 0.4, 1.0, 1.9, 3.7, 9.7, 32.8, 117.0],
    303.15: [0.1, 0.3, 1.0, 3.1, 7.4, 12.7, 23.1, 59.3],
}


@pytest.mark.parametrize("temp", sorted(ISO9613_70RH))
def test_attenuation_matches_iso9613_table(oracle, temp):
    freqs = [1000.0 * 10 ** (3 * b / 10) for b in range(-4, 4)]
    for f, want in zip(freqs, ISO9613_70RH[temp]):
        db_per_km = oracle.attenuation_coef(f, 70, 1, 1, temp, "f64") * 1e6 * 20 / math.log(10)
        # table values are rounded to 0.1 dB/km
        assert abs(db_per_km - want) <= 0.05 + 0.005 * want, (temp, f, db_per_km, want)


# ---------------------------------------------------------------- one source
def _sphere(pos, amp=1.0, phase=0.0, freq=40e3):
    return np.array([[*pos, amp, phase, freq]], np.float32)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_spherical_wave_inverse_distance_and_phase_sign(oracle, prec):
    k, alpha = np.float32(0.7391983), np.float32(1.0e-4)
    src = _sphere((1.0, -2.0, 0.5), amp=2.0, phase=0.3)
    dirs = np.array([[0.6, 0.0, 0.8]], np.float64)
    r = np.array([10.0, 50.0, 137.5, 400.0])
    pts = (np.array([1.0, -2.0, 0.5]) + r[:, None] * dirs).astype(np.float32)
    p = oracle.field(0, src, [k], [alpha], pts, prec)
    rr = np.linalg.norm(pts.astype(np.float64) - np.array([1.0, -2.0, 0.5]), axis=1)
    # |p| r e^{alpha r} = amp
    tol = 1e-5 if prec == "f32" else 1e-12
    assert np.allclose(np.abs(p) * rr * np.exp(float(alpha) * rr), 2.0, rtol=tol)
    # the phase advances by +k dr along a ray (sign of from_polar(r, phase + k r))
    dphi = np.angle(p[1:] * np.conj(p[:-1]))
    want = np.angle(np.exp(1j * float(k) * np.diff(rr)))
    assert np.allclose(dphi, want, atol=5e-4 if prec == "f32" else 1e-9)


def test_t4010a1_on_axis_equals_sphere(oracle):
    k, a = [np.float32(0.74)], [np.float32(9e-5)]
    t = np.array([[0, 0, 0, 0, 0, 1, 1.0, 0.2, 40e3]], np.float32)
    s = _sphere((0, 0, 0), 1.0, 0.2)
    pts = np.array([[0, 0, 100.0], [3.0, 4.0, 80.0]], np.float32)  # theta <= 20 deg
    assert np.array_equal(oracle.field(1, t, k, a, pts, "f64"), oracle.field(0, s, k, a, pts, "f64"))
    behind = np.array([[0, 0, -100.0]], np.float32)
    assert oracle.field(1, t, k, a, behind, "f64")[0] == 0


# ---------------------------------------------------------------- many sources
def test_linearity_in_sources(oracle):
    src, k, a, area = oracle.focus_plane_249()
    pts, _ = oracle.grid_points(**area)
    pts = pts[::37]
    full = oracle.field(1, src, k, a, pts, "f64")
    h = src.shape[0] // 2
    parts = oracle.field(1, src[:h], k[:h], a[:h], pts, "f64") + oracle.field(1, src[h:], k[h:], a[h:], pts, "f64")
    assert rel_l2(parts, full) < 1e-13


def test_translation_invariance(oracle):
    src, k, a, area = oracle.focus_plane_249()
    pts, _ = oracle.grid_points(**area)
    pts = pts[::53].astype(np.float64)
    # positions on a 1/64 mm lattice so that the shifted inputs are exact in f32
    src = src.copy()
    src[:, :3] = np.round(src[:, :3] * 64) / 64
    shift = np.array([256.0, -128.0, 64.0])
    s2 = src.copy()
    s2[:, :3] = (src[:, :3] + shift).astype(np.float32)
    f1 = oracle.field(1, src, k, a, pts.astype(np.float32), "f64")
    f2 = oracle.field(1, s2, k, a, (pts + shift).astype(np.float32), "f64")
    assert rel_l2(f2, f1) < 1e-12


def test_focus_identity(oracle):
    """With FocalPoint phases every term arrives in phase at the focus, so |p(f)|
    equals the sum of the term amplitudes."""
    src, k, a, _ = oracle.focus_plane_249()
    f = np.array([[90.0, 70.0, 150.0]], np.float32)
    p = oracle.field(1, src, k, a, f, "f64")[0]
    pos = src[:, :3].astype(np.float64)
    d = np.linalg.norm(f.astype(np.float64) - pos, axis=1)
    th = np.arccos((f[0, 2] - pos[:, 2]) / d)
    amp_sum = np.sum(oracle.np_directivity(th) * np.exp(-d * a.astype(np.float64)) / d)
    assert abs(p) == pytest.approx(amp_sum, rel=1e-8)
    # and it is the maximum of the focal plane
    pts, _ = oracle.grid_points(order=(0, 1), ranges=((70, 110), (50, 90), (150, 150)), res=1.0)
    plane = np.abs(oracle.field(1, src, k, a, pts, "f64"))
    assert np.argmax(plane) == np.where((pts[:, 0] == 90) & (pts[:, 1] == 70))[0][0]


def test_c_f64_matches_numpy_twin(oracle):
    rng = np.random.default_rng(7)
    pos = rng.uniform(-50, 50, (40, 3)).astype(np.float32)
    pos[:, 2] = 0
    dirs = rng.normal(size=(40, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    src = oracle.t4010a1_sources(pos, amp=rng.uniform(0.5, 1.0, 40), phase=rng.uniform(0, 6.28, 40),
                                 direction=dirs.astype(np.float32))
    k = rng.uniform(0.5, 1.0, 40).astype(np.float32)
    a = rng.uniform(0, 2e-4, 40).astype(np.float32)
    pts = rng.uniform(-80, 80, (300, 3)).astype(np.float32)
    for st, s in ((1, src), (0, src[:, [0, 1, 2, 6, 7, 8]])):
        c = oracle.field(st, s, k, a, pts, "f64")
        n = oracle.np_field_f64(st, s, k, a, pts)
        assert rel_l2(c, n) < 1e-13


def test_fp32_path_error_vs_fp64_focus_plane(oracle):
    src, k, a, area = oracle.focus_plane_249()
    pts, _ = oracle.grid_points(**area)
    e = rel_l2(oracle.field(1, src, k, a, pts, "f32"), oracle.field(1, src, k, a, pts, "f64"))
    assert 1e-7 < e < 1e-5  # FP32 rounding is visible but small


def test_accurate_sum_is_order_independent(oracle):
    """The ascending-order sum (system_acc.rs:46-51) does not depend on the order
    in which sources were added; the naive sum does."""
    src, k, a, area = oracle.focus_plane_249()
    pts, _ = oracle.grid_points(**area)
    pts = pts[::7]
    perm = np.random.default_rng(1).permutation(src.shape[0])
    acc = oracle.field(1, src, k, a, pts, "f32", accurate=True)
    acc_p = oracle.field(1, src[perm], k[perm], a[perm], pts, "f32", accurate=True)
    assert np.array_equal(acc, acc_p)
    naive = oracle.field(1, src, k, a, pts, "f32")
    naive_p = oracle.field(1, src[perm], k[perm], a[perm], pts, "f32")
    assert not np.array_equal(naive, naive_p)
    assert rel_l2(acc, naive) < 1e-5


def test_reductions(oracle):
    cp = np.array([3 + 4j, -1e-3 + 2e-3j, 0], np.complex64)
    assert np.allclose(oracle.reduce_f32(0, cp), [5.0, np.hypot(1e-3, 2e-3), 0.0], rtol=2e-7, atol=0)
    assert np.array_equal(oracle.reduce_f32(2, cp), cp.view(np.float32))
    assert np.allclose(oracle.reduce_f32(1, cp), [25.0, 5e-6, 0.0], rtol=1e-6)


# ---------------------------------------------------------------- grids
def _brute(order, ranges, res):
    res = np.float32(res)
    cnt = []
    for lo, hi in ranges:
        q = (np.float32(hi) - np.float32(lo)) / res
        cnt.append(int(q) + 1 if q > 0 else 1)
    full = list(order) + [a for a in range(3) if a not in order]
    pts = []
    for n2 in range(cnt[full[2]]):
        for n1 in range(cnt[full[1]]):
            for n0 in range(cnt[full[0]]):
                p = [np.float32(r[0]) for r in ranges]
                for ax, n in ((full[0], n0), (full[1], n1), (full[2], n2)):
                    p[ax] = np.float32(np.float32(ranges[ax][0]) + np.float32(np.float32(n) * res))
                pts.append(p)
    return np.array(pts, np.float32), tuple(cnt)


@pytest.mark.parametrize("order", [(0, 1, 2), (2, 0, 1), (1, 2, 0), (0, 2, 1), (1, 0, 2), (2, 1, 0)])
def test_grid3d_order_and_count(oracle, order):
    ranges = ((-3.0, 2.5), (10.0, 12.0), (0.25, 1.9))
    pts, b = oracle.grid_points(order, ranges, 0.5)
    want, cnt = _brute(order, ranges, 0.5)
    assert b == cnt
    assert np.array_equal(pts, want)


@pytest.mark.parametrize("order", [(0, 1), (1, 0), (1, 2), (2, 1), (2, 0), (0, 2)])
def test_grid2d_order_and_count(oracle, order):
    ranges = [(-1.0, 1.0), (5.0, 6.5), (7.0, 7.0)]
    at = [a for a in range(3) if a not in order][0]
    ranges[at] = (3.25, 3.25)
    pts, b = oracle.grid_points(order, ranges, 0.5)
    want, cnt = _brute(order, ranges, 0.5)
    assert b == cnt and np.array_equal(pts, want)


def test_grid_counts_follow_builder(oracle):
    # x_range(40, 139) at 1 mm -> 100 points; a +-50 range -> 101 (builder.rs:198-200)
    _, b = oracle.grid_points((0, 1), ((40, 139), (20, 119), (150, 150)), 1.0)
    assert b == (100, 100, 1)
    _, b = oracle.grid_points((0,), ((-50, 50), (0, 0), (0, 0)), 1.0)
    assert b == (101, 1, 1)
    # reversed range: saturating cast gives one point
    _, b = oracle.grid_points((0,), ((5, -5), (0, 0), (0, 0)), 1.0)
    assert b == (1, 1, 1)
