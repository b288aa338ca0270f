"""C-ABI checks that need no GPU: the library loads, exports every symbol the
header declares, and its host-side logic (handles, per-source constants, grid
counts, packing, error paths) matches the reference's semantics
(capi/src/*.rs, system/uniform_medium.rs:102-107, observe_area/grid/builder.rs).
No kernel is launched here."""
import ctypes
import math
import os
import re
from ctypes import byref, c_uint, c_void_p

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "afc_hip.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b((?:AFC|AFO|AFH|AFS)_\w+)\s*\(", txt)
    return sorted(set(names))


def test_header_symbols_exported(afc):
    lib = afc.dll()
    names = declared_functions()
    assert len(names) >= 43
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_reference_capi_symbols_present(afc):
    # the 38 exports of the reference cdylib (capi/src/*.rs)
    ref = """AFC_CreateCalculator AFC_FreeCalculator AFC_Calculate AFC_CreatePressureField AFC_FreePressureField
    AFC_CreatePowerField AFC_FreePowerField AFC_CreateComplexPressureField AFC_FreeComplexPressureField
    AFC_CreateScatterArea AFC_ScatterAddObservePoint AFC_FreeScatterArea AFC_CreateGridArea1D AFC_CreateGridArea2D
    AFC_CreateGridArea3D AFC_FreeGridArea AFC_GetGridSize AFC_CreateUniformSystem AFC_FreeUniformSystem
    AFC_UniformSystemSoundSpeed AFC_UniformSystemInfo AFC_UniformSystemSourceInfo AFC_AddSphereWaveSource
    AFC_AddT4010A1 AFC_GetWaveSources AFO_FocalPoint AFO_BesselBeam AFO_IFFT AFO_GSPAT AFO_GS AFO_Naive AFO_Horn
    AFO_Long AFO_APO AFO_GaussNewton AFO_LM AFO_GradientDescent AFO_Greedy""".split()
    assert len(ref) == 38
    lib = afc.dll()
    assert all(hasattr(lib, n) for n in ref)


def _system(afc, temperature, positions, source=None):
    sys_ = afc.UniformSystem(temperature)
    for p in positions:
        sys_.add_wave_source(source(p) if source else afc.T4010A1(p, (0, 0, 1), 1.0, 0.0, 40e3))
    return sys_


def test_source_constants_bit_exact(afc, oracle):
    for t in (300.0, oracle.TEMP_340, 287.0):
        s = _system(afc, t, [(0, 0, 0), (10.16, 0, 0)])
        k = np.zeros(2, np.float32)
        a = np.zeros(2, np.float32)
        assert afc.dll().AFH_GetSourceConstants(s.handle, k.ctypes.data_as(c_void_p), a.ctypes.data_as(c_void_p)) == 2
        ko, ao = oracle.source_constants(np.array([40e3, 40e3], np.float32), t)
        assert np.array_equal(k, ko) and np.array_equal(a, ao)
        assert s.sound_speed() == np.float32(oracle.sound_speed(t))


def test_wave_sources_pointer_is_mutable_view(afc):
    s = _system(afc, 300.0, [(0, 0, 0), (1, 2, 3)])
    srcs = s.get_wave_sources()
    assert len(srcs) == 2 and tuple(srcs[1].pos) == (1.0, 2.0, 3.0)
    srcs[1].phase = 1.25
    assert s.source_array()[1, 7] == np.float32(1.25)


def test_info_strings(afc):
    s = _system(afc, 300.0, [(0, 0, 0)])
    assert "Sound Speed" in s.info()
    assert "Wavenumber" in s.info_of_source(0)


@pytest.mark.parametrize("axes", [(1, 2, 3), (3, 1, 2), (2, 3, 1), (1, 3, 2), (2, 1, 3), (3, 2, 1)])
def test_grid_sizes_by_axis(afc, oracle, axes):
    b = afc.GridAreaBuilder()
    rng = {1: (-3.0, 2.5), 2: (10.0, 12.0), 3: (0.25, 1.9)}
    for ax in axes:
        getattr(b, {1: "x_range", 2: "y_range", 3: "z_range"}[ax])(rng[ax])
    area = b.resolution(0.5).generate()
    _, want = oracle.grid_points([a - 1 for a in axes], (rng[1], rng[2], rng[3]), 0.5)
    assert area.get_size() == want


def test_grid_2d_with_at_axis(afc):
    area = afc.GridAreaBuilder().x_range((40, 139)).y_range((20, 119)).z_at(150).resolution(1.0).generate()
    assert area.get_size() == (100, 100, 1)
    area = afc.GridAreaBuilder().z_range((-50, 50)).x_at(1).y_at(2).resolution(1.0).generate()
    assert area.get_size() == (1, 1, 101)


def test_packing_layout(afc, oracle):
    src, k, a, _ = oracle.focus_plane_249()
    from afc import device
    p = device.pack_sources(1, src, k, a)
    n = src.shape[0]
    assert p.shape == (device.packed_count(n), 12) and p.shape[0] % 32 == 0
    assert np.array_equal(p[:n, 0:6], src[:, 0:6]) and np.array_equal(p[:n, 6], src[:, 6])
    # phase in revolutions modulo 1, k/(2 pi) as a float pair, alpha * log2(e)
    ph = src[:, 7].astype(np.float64) / (2 * math.pi)
    assert np.allclose(p[:n, 7], ph - np.floor(ph), atol=6e-8)
    kr = (p[:n, 8].astype(np.float64) + p[:n, 9].astype(np.float64)) * 2 * math.pi
    assert np.allclose(kr, k.astype(np.float64), rtol=1e-15)
    assert np.allclose(p[:n, 10], a.astype(np.float64) * math.log2(math.e), rtol=1e-7)
    assert np.array_equal(p[:n, 11], a)
    # the f32 wavenumber is recoverable exactly from the hi/lo pair
    assert np.array_equal(kr.astype(np.float32), k)
    assert np.all(p[n:, 6] == 0)  # padding has zero amplitude


def test_every_calculator_type_constructs(afc):
    """CalculatorType 0 (Cpu), 1 (Accurate), 2 (Gpu) of type_inference_aux.rs:55-59
    and this backend's 3 all create a handle; an unknown tag is refused."""
    for cls in (afc.CpuCalculator, afc.AccurateCalculator, afc.GpuCalculator, afc.HipCalculator):
        assert cls().handle
    h = c_void_p()
    afc.dll().AFC_CreateCalculator(byref(h), 4)
    assert not h and b"unknown" in afc.dll().AFC_GetLastError()


# The public names of the reference's binding: python/afc/__init__.py:20-40
# (package exports, plot_helper aside) and the methods of python/afc/afc.py
# (UniformSystem :30-85, Calculator :88-150, GridArea :233-301,
# GridAreaBuilder :305-416, Optimizer :419-567).
REF_EXPORTS = ["UniformSystem", "CpuCalculator", "AccurateCalculator", "GpuCalculator", "PressureFieldBuffer",
               "PowerFieldBuffer", "ComplexPressureFieldBuffer", "ScatterArea", "GridAreaBuilder", "SphereWaveSource",
               "T4010A1", "Vector3", "Optimizer"]
REF_METHODS = {
    "UniformSystem": ["set_source_type", "sound_speed", "info", "info_of_source", "add_wave_source",
                      "get_wave_sources"],
    "CpuCalculator": ["calculate", "get_type"],
    "GridAreaBuilder": ["x_at", "y_at", "z_at", "x_range", "y_range", "z_range", "resolution", "generate"],
    "ScatterArea": ["add_observe_point", "get_type"],
    "Optimizer": ["focus", "bessel", "ifft", "gs_pat", "gs", "naive", "horn", "long2014", "apo", "lm", "gauss_newton",
                  "gradient_descent", "greedy"],
}


def test_mirror_has_every_reference_name(afc):
    for name in REF_EXPORTS:
        assert hasattr(afc, name), name
    for cls, methods in REF_METHODS.items():
        for m in methods:
            assert callable(getattr(getattr(afc, cls), m, None)), f"{cls}.{m}"
    for cls in ("GridArea1D", "GridArea2D", "GridArea3D"):
        assert callable(getattr(afc, cls).get_size)
    # struct properties the reference's scripts use (python/main.py:58-63, foci.py:105)
    t = afc.T4010A1((1, 2, 3), (0, 0, 1), 1.0, 0.5, 40e3)
    assert tuple(t.pos) == (1, 2, 3) and tuple(t.dir) == (0, 0, 1)


def test_ifft_reports_not_provided(afc):
    s = _system(afc, 300.0, [(0, 0, 0)])
    before = s.source_array().copy()
    v = afc.Vector3((0, 0, 0))
    afc.dll().AFO_IFFT(s.handle, b"image.png", v, v, v, ctypes.c_float(1.0), ctypes.c_float(100.0), 1)
    assert b"not provided" in afc.dll().AFC_GetLastError()
    with pytest.raises(afc.AFCError, match="not provided"):
        afc.Optimizer.ifft(s, "star.bmp", v, v, v, 10.18, 150.0)
    assert np.array_equal(s.source_array(), before)


def test_bad_tags_do_not_abort(afc):
    h = c_void_p()
    afc.dll().AFC_CreateUniformSystem(byref(h), 300.0, 7)
    assert not h and afc.dll().AFC_GetLastError()
    g = c_void_p()
    afc.dll().AFC_CreateGridArea3D(byref(g), 0, 1, 0, 1, 0, 1, 1.0, 1, 1, 3)
    assert not g and b"permutation" in afc.dll().AFC_GetLastError()


def test_focal_point_matches_oracle(afc, oracle):
    pos = oracle.array_249()
    s = _system(afc, oracle.TEMP_340, [tuple(p) for p in pos])
    afc.Optimizer.focus(s, (90.0, 70.0, 150.0))
    got = s.source_array()[:, 7]
    want = oracle.focal_point_phases(1, s.source_array().copy(), s.sound_speed(), (90.0, 70.0, 150.0))
    assert np.array_equal(got, want)


@pytest.mark.parametrize("direction,theta", [((0.0, 0.0, 1.0), 0.3), ((0.2, -0.1, 0.97), 0.25), ((0.6, 0.0, 0.8), 0.5)])
def test_bessel_beam_matches_oracle(afc, oracle, direction, theta):
    """AFO_BesselBeam (bessel.rs:31-55) sets the same f32 phases as the oracle's
    restatement, bit for bit (both in the reference's operation order)."""
    pos = oracle.array_249()
    s = _system(afc, oracle.TEMP_340, [tuple(p) for p in pos])
    point = (90.0, 70.0, 0.0)
    afc.Optimizer.bessel(s, point, direction, theta)
    got = s.source_array()[:, 7]
    want = oracle.bessel_phases(1, s.source_array().copy(), s.sound_speed(), point, direction, theta)
    assert np.array_equal(got, want)
