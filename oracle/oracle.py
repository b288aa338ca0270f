"""ctypes loader for the C oracle (oracle/afc_oracle.c) plus a small numpy FP64 twin.

TEST INFRASTRUCTURE ONLY. Imported by tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py``, always as the checker. The product library
(``acoustic-field-calculator_amd``) never imports this module.

Every function restates the reference's algorithm; the C file's header lists the
reference file:line each follows. Source arrays use the reference's repr(C)
layouts: SphereWaveSource = 6 f32 {pos[3], amp, phase, freq}
(core/wave_sources/sphere_sources.rs:19-27) and T4010A1 = 9 f32
{pos[3], dir[3], amp, phase, freq} (core/wave_sources/t4010a1.rs:73-82).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_float, c_int, c_uint32, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liborc_afc.so")

SPHERE = 0
T4010A1 = 1

_lib = None


def build() -> str:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        fp = c_void_p
        for sfx, R in (("f32", c_float), ("f64", c_double)):
            getattr(L, f"orc_directivity_t4010a1_{sfx}").restype = R
            getattr(L, f"orc_directivity_t4010a1_{sfx}").argtypes = [R]
            getattr(L, f"orc_sound_speed_{sfx}").restype = R
            getattr(L, f"orc_sound_speed_{sfx}").argtypes = [R]
            getattr(L, f"orc_attenuation_coef_{sfx}").restype = R
            getattr(L, f"orc_attenuation_coef_{sfx}").argtypes = [R, R, R, R, R]
            for nm in ("orc_field", "orc_field_accurate"):
                f = getattr(L, f"{nm}_{sfx}")
                f.restype = None
                f.argtypes = [c_int, fp, fp, fp, c_uint64, fp, c_uint64, fp, c_int]
            f = getattr(L, f"orc_propagation_matrix_{sfx}")
            f.restype = None
            f.argtypes = [c_int, fp, fp, fp, c_uint64, fp, c_uint64, fp]
            for nm in ("orc_gs", "orc_gspat"):
                f = getattr(L, f"{nm}_{sfx}")
                f.restype = None
                f.argtypes = [c_int, fp, fp, fp, c_uint64, fp, fp, c_uint64, c_uint32, fp, fp]
            f = getattr(L, f"orc_naive_{sfx}")
            f.restype = None
            f.argtypes = [c_int, fp, fp, fp, c_uint64, fp, fp, c_uint64, fp, fp]
            for nm in ("orc_gauss_newton", "orc_gradient_descent"):
                f = getattr(L, f"{nm}_{sfx}")
                f.restype = c_uint32
                f.argtypes = [c_int, fp, fp, fp, c_uint64, fp, fp, c_uint64, c_uint32, fp, fp]
            f = getattr(L, f"orc_apo_{sfx}")
            f.restype = c_uint32
            f.argtypes = [c_int, fp, fp, fp, c_uint64, fp, fp, c_uint64, R, c_uint32, fp, fp]
            f = getattr(L, f"orc_long_{sfx}")
            f.restype = None
            f.argtypes = [c_int, fp, fp, fp, c_uint64, fp, fp, c_uint64, fp, fp]
            f = getattr(L, f"orc_horn_{sfx}")
            f.restype = None
            rt = c_float if sfx == "f32" else c_double
            f.argtypes = [c_int, fp, fp, fp, c_uint64, fp, fp, c_uint64, c_uint64, c_uint64, rt, rt, fp, fp]
            f = getattr(L, f"orc_herm_eig_{sfx}")
            f.restype = None
            f.argtypes = [fp, c_uint64, fp, fp]
            f = getattr(L, f"orc_greedy_{sfx}")
            f.restype = None
            f.argtypes = [c_int, fp, fp, fp, c_uint64, fp, fp, c_uint64, c_uint64, fp]
            f = getattr(L, f"orc_lm_{sfx}")
            f.restype = c_uint32
            f.argtypes = [c_int, fp, fp, fp, c_uint64, fp, fp, c_uint64, c_uint32, fp, fp]
            f = getattr(L, f"orc_lm_eval_{sfx}")
            f.restype = R
            f.argtypes = [c_int, fp, fp, fp, c_uint64, fp, fp, c_uint64, fp, fp]
            f = getattr(L, f"orc_focal_point_phases_{sfx}")
            f.restype = None
            f.argtypes = [c_int, fp, c_uint64, R, R, R, R, fp]
            f = getattr(L, f"orc_bessel_phases_{sfx}")
            f.restype = None
            f.argtypes = [c_int, fp, c_uint64, R, R, R, R, R, R, R, R, fp]
        L.orc_grid_points.restype = c_uint64
        L.orc_grid_points.argtypes = [c_int, fp, fp, c_float, fp, fp]
        L.orc_reduce_f32.restype = None
        L.orc_reduce_f32.argtypes = [c_int, fp, c_uint64, fp]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(c_void_p)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _dt(prec: str):
    return (np.float32, "f32") if prec == "f32" else (np.float64, "f64")


# --------------------------------------------------------------------------- scalars
def directivity(theta: float, prec: str = "f32") -> float:
    return getattr(lib(), f"orc_directivity_t4010a1_{_dt(prec)[1]}")(theta)


def sound_speed(t: float, prec: str = "f32") -> float:
    return getattr(lib(), f"orc_sound_speed_{_dt(prec)[1]}")(t)


def attenuation_coef(freq, hr, ps, ps0, t, prec: str = "f32") -> float:
    return getattr(lib(), f"orc_attenuation_coef_{_dt(prec)[1]}")(freq, hr, ps, ps0, t)


def source_constants(freqs: np.ndarray, temperature: float):
    """Per-source (wavenumber, attenuation) as UniformSystem::add_wave_source does
    (system/uniform_medium.rs:102-107): k = 2*PI*f/c(T), alpha at 30 %RH, 1 atm."""
    c = np.float32(sound_speed(temperature))
    two_pi = np.float32(2.0) * np.float32(np.pi)
    k = (two_pi * _f32(freqs)) / c
    a = np.array([attenuation_coef(float(f), 30.0, 1.0, 1.0, temperature) for f in _f32(freqs)], np.float32)
    return k.astype(np.float32), a


# --------------------------------------------------------------------------- field
def field(source_type: int, sources, wavenums, attens, points, prec: str = "f32", threads: int = 0,
          accurate: bool = False) -> np.ndarray:
    """Complex pressure at each point (complex64 for f32, complex128 for f64)."""
    dt, sfx = _dt(prec)
    src, k, a, pts = _f32(sources), _f32(wavenums), _f32(attens), _f32(points).reshape(-1, 3)
    n = k.shape[0]
    assert src.size == n * (9 if source_type == T4010A1 else 6)
    out = np.empty(2 * pts.shape[0], dt)
    threads = threads or os.cpu_count() or 1
    fn = getattr(lib(), f"orc_field{'_accurate' if accurate else ''}_{sfx}")
    fn(source_type, _p(src), _p(k), _p(a), n, _p(pts), pts.shape[0], _p(out), threads)
    return out.view(np.complex64 if prec == "f32" else np.complex128)


def reduce_f32(field_type: int, cp: np.ndarray) -> np.ndarray:
    """FieldBuffer reduction of an FP32 complex pressure (0 |p|, 1 |p|^2, 2 complex)."""
    c = np.ascontiguousarray(cp, np.complex64)
    n = c.shape[0]
    out = np.empty(2 * n if field_type == 2 else n, np.float32)
    lib().orc_reduce_f32(field_type, _p(c.view(np.float32)), n, _p(out))
    return out


def propagation_matrix(source_type, sources, wavenums, attens, foci, prec="f32") -> np.ndarray:
    """G (M x N) as a numpy array, G[i, j] = term(source j -> focus i)."""
    dt, sfx = _dt(prec)
    src, k, a, f = _f32(sources), _f32(wavenums), _f32(attens), _f32(foci).reshape(-1, 3)
    n, m = k.shape[0], f.shape[0]
    out = np.empty(2 * m * n, dt)
    getattr(lib(), f"orc_propagation_matrix_{sfx}")(source_type, _p(src), _p(k), _p(a), n, _p(f), m, _p(out))
    cdt = np.complex64 if prec == "f32" else np.complex128
    return out.view(cdt).reshape(n, m).T.copy()


def _solver(name, source_type, sources, wavenums, attens, foci, amps, prec, repeat=None):
    dt, sfx = _dt(prec)
    src, k, a, f, am = _f32(sources), _f32(wavenums), _f32(attens), _f32(foci).reshape(-1, 3), _f32(amps)
    n, m = k.shape[0], f.shape[0]
    oa, op = np.empty(n, dt), np.empty(n, dt)
    fn = getattr(lib(), f"{name}_{sfx}")
    if repeat is None:
        fn(source_type, _p(src), _p(k), _p(a), n, _p(f), _p(am), m, _p(oa), _p(op))
    else:
        fn(source_type, _p(src), _p(k), _p(a), n, _p(f), _p(am), m, repeat, _p(oa), _p(op))
    return oa, op


def gs(source_type, sources, wavenums, attens, foci, amps, prec="f32", repeat=100):
    """GS (gs.rs:36-77). Returns (amp, phase) per source."""
    return _solver("orc_gs", source_type, sources, wavenums, attens, foci, amps, prec, repeat)


def gspat(source_type, sources, wavenums, attens, foci, amps, prec="f32", repeat=100):
    """GS-PAT (gs_pat.rs:38-81). Returns (amp, phase) per source."""
    return _solver("orc_gspat", source_type, sources, wavenums, attens, foci, amps, prec, repeat)


def naive(source_type, sources, wavenums, attens, foci, amps, prec="f32"):
    """Naive linear synthesis (naive.rs:32-51). Returns (amp, phase) per source."""
    return _solver("orc_naive", source_type, sources, wavenums, attens, foci, amps, prec)


def greedy(source_type, sources, wavenums, attens, foci, amps, phase_div, prec="f32"):
    """Greedy phase search (combination/greedy.rs:23-83). Returns the phase per source."""
    dt, sfx = _dt(prec)
    src, k, a, f, am = _f32(sources), _f32(wavenums), _f32(attens), _f32(foci).reshape(-1, 3), _f32(amps)
    n, m = k.shape[0], f.shape[0]
    ph = np.empty(n, dt)
    getattr(lib(), f"orc_greedy_{sfx}")(source_type, _p(src), _p(k), _p(a), n, _p(f), _p(am), m, int(phase_div),
                                        _p(ph))
    return ph


def lm(source_type, sources, wavenums, attens, foci, amps, prec="f64", k_max=200):
    """Levenberg-Marquardt (lm.rs:40-105). Returns (phases[N], x[N+M], iterations)."""
    dt, sfx = _dt(prec)
    src, k, a, f, am = _f32(sources), _f32(wavenums), _f32(attens), _f32(foci).reshape(-1, 3), _f32(amps)
    n, m = k.shape[0], f.shape[0]
    ph, x = np.empty(n, dt), np.empty(n + m, dt)
    it = getattr(lib(), f"orc_lm_{sfx}")(source_type, _p(src), _p(k), _p(a), n, _p(f), _p(am), m, k_max, _p(ph),
                                         _p(x))
    return ph, x, int(it)


def _nls(name, source_type, sources, wavenums, attens, foci, amps, prec, k_max):
    dt, sfx = _dt(prec)
    src, k, a, f, am = _f32(sources), _f32(wavenums), _f32(attens), _f32(foci).reshape(-1, 3), _f32(amps)
    n, m = k.shape[0], f.shape[0]
    ph, x = np.empty(n, dt), np.empty(n + m, dt)
    it = getattr(lib(), f"{name}_{sfx}")(source_type, _p(src), _p(k), _p(a), n, _p(f), _p(am), m, k_max, _p(ph), _p(x))
    return ph, x, int(it)


def gauss_newton(source_type, sources, wavenums, attens, foci, amps, prec="f64", k_max=200):
    """Gauss-Newton (gauss_newton.rs:35-89). Returns (phases[N], x[N+M], iterations)."""
    return _nls("orc_gauss_newton", source_type, sources, wavenums, attens, foci, amps, prec, k_max)


def gradient_descent(source_type, sources, wavenums, attens, foci, amps, prec="f64", k_max=10000):
    """Gradient descent (gradient_descent.rs:34-61). Returns (phases[N], x[N+M], steps)."""
    return _nls("orc_gradient_descent", source_type, sources, wavenums, attens, foci, amps, prec, k_max)


def long_(source_type, sources, wavenums, attens, foci, amps, prec="f64"):
    """Long et al. (matrix/long.rs:46-106). Returns (amp[N], phase[N])."""
    return _solver("orc_long", source_type, sources, wavenums, attens, foci, amps, prec)


def horn(source_type, sources, wavenums, attens, foci, amps, seed=0, repeat=100, lam=0.8, alpha=1e-5, prec="f64"):
    """Horn / Inoue et al. (matrix/horn.rs:58-129) with the seeded row choice. Returns (amp[N], phase[N])."""
    dt, sfx = _dt(prec)
    src, k, a, f, am = _f32(sources), _f32(wavenums), _f32(attens), _f32(foci).reshape(-1, 3), _f32(amps)
    n, m = k.shape[0], f.shape[0]
    amp, ph = np.empty(n, dt), np.empty(n, dt)
    getattr(lib(), f"orc_horn_{sfx}")(source_type, _p(src), _p(k), _p(a), n, _p(f), _p(am), m, int(seed),
                                      int(repeat), float(lam), float(alpha), _p(amp), _p(ph))
    return amp, ph


def herm_eig(A, prec="f64"):
    """The oracle's Hermitian Jacobi eigen-decomposition: (eigenvalues, eigenvectors as columns)."""
    dt, sfx = _dt(prec)
    cdt = np.complex128 if dt == np.float64 else np.complex64
    n = A.shape[0]
    a = np.asfortranarray(np.asarray(A, cdt)).copy(order="F")
    ev = np.empty(n, dt)
    V = np.empty((n, n), cdt, order="F")
    getattr(lib(), f"orc_herm_eig_{sfx}")(a.ctypes.data_as(c_void_p), n, _p(ev), V.ctypes.data_as(c_void_p))
    return ev, V


def apo(source_type, sources, wavenums, attens, foci, amps, lam, prec="f64", k_max=200):
    """APO (apo.rs:100-158). Returns (amp[N], phase[N], iterations)."""
    dt, sfx = _dt(prec)
    src, k, a, f, am = _f32(sources), _f32(wavenums), _f32(attens), _f32(foci).reshape(-1, 3), _f32(amps)
    n, m = k.shape[0], f.shape[0]
    amp, ph = np.empty(n, dt), np.empty(n, dt)
    it = getattr(lib(), f"orc_apo_{sfx}")(source_type, _p(src), _p(k), _p(a), n, _p(f), _p(am), m, lam, k_max,
                                          _p(amp), _p(ph))
    return amp, ph, int(it)


def lm_eval(source_type, sources, wavenums, attens, foci, amps, x):
    """(F(x), g(x)) of the LM objective in FP64 (nls/macros.rs:17-58)."""
    src, k, a, f, am = _f32(sources), _f32(wavenums), _f32(attens), _f32(foci).reshape(-1, 3), _f32(amps)
    n, m = k.shape[0], f.shape[0]
    xx = np.ascontiguousarray(x, np.float64)
    g = np.empty(n + m, np.float64)
    F = lib().orc_lm_eval_f64(source_type, _p(src), _p(k), _p(a), n, _p(f), _p(am), m, _p(xx), _p(g))
    return F, g


def bessel_phases(source_type, sources, sound_speed_mm_s, point, direction, theta, prec="f32") -> np.ndarray:
    """BesselBeam phases (bessel.rs:31-55) per source."""
    dt, sfx = _dt(prec)
    src = _f32(sources)
    n = src.shape[0]
    out = np.empty(n, dt)
    getattr(lib(), f"orc_bessel_phases_{sfx}")(source_type, _p(src), n, sound_speed_mm_s, *[float(v) for v in point],
                                               *[float(v) for v in direction], float(theta), _p(out))
    return out


def focal_point_phases(source_type, sources, sound_speed_mm_s, focus, prec="f32") -> np.ndarray:
    dt, sfx = _dt(prec)
    src = _f32(sources)
    n = src.size // (9 if source_type == T4010A1 else 6)
    out = np.empty(n, dt)
    getattr(lib(), f"orc_focal_point_phases_{sfx}")(source_type, _p(src), n, sound_speed_mm_s, float(focus[0]),
                                                     float(focus[1]), float(focus[2]), _p(out))
    return out


def grid_points(order, ranges, res):
    """Observation points of a GridArea. order: ranged axes (0=x,1=y,2=z) in the
    order given; ranges: ((xmin,xmax),(ymin,ymax),(zmin,zmax)) with min == max for
    `*_at` axes. Returns (points[P,3] f32, bounds (nx, ny, nz))."""
    o = np.zeros(3, np.int32)
    o[: len(order)] = order
    r = _f32(np.array(ranges, np.float32).reshape(6))
    b = np.zeros(3, np.uint32)
    n = lib().orc_grid_points(len(order), _p(o), _p(r), float(np.float32(res)), None, _p(b))
    pts = np.empty((n, 3), np.float32)
    lib().orc_grid_points(len(order), _p(o), _p(r), float(np.float32(res)), _p(pts), _p(b))
    return pts, tuple(int(v) for v in b)


# --------------------------------------------------------------------------- numpy FP64 twin
_DIR = np.array([
    [0.0, 1.0, 1.0, 1.0, 0.891250938, 0.707945784, 0.501187234, 0.354813389, 0.251188643, 0.199526231],
    [0.0, 0.0, 0.0, -0.2633589359691162, -0.8910683499820542, -1.1924098538381465, -1.0439934287916925,
     -0.7015151838490918, -0.4471050803642494, -0.179254124486545],
    [0.0, 0.0, 0.0, -2.586746661956269, -1.0097633564074366, -0.7167962799004945, 0.027352097655185065,
     0.39509604454751995, 1.0625664749725618, 0.008242027161100222],
    [0.0, 0.0, 0.0, 0.0009174516976956062, 0.0001704413361662757, 4.362034715924911, -2.238476800444034,
     1.274774620435337, -1.1276955684895724, -0.9016569137759131],
], np.float32).astype(np.float64)


def np_directivity(theta: np.ndarray) -> np.ndarray:
    """Vectorised t4010a1.rs:107-120 in float64 (NaN -> index 0 -> 1)."""
    th = np.asarray(theta, np.float64)
    q = np.ceil(th / (np.pi / 18.0))
    i = np.where(np.isnan(q) | (q < 0), 0, q)
    i = np.minimum(i, 10).astype(np.int64)
    ic = np.clip(i, 0, 9)
    x = th - (ic - 1.0) * (np.pi / 18.0)
    a, b, c, d = (_DIR[r][ic] for r in range(4))
    val = a + (b + (c + d * x) * x) * x
    return np.where(i == 0, 1.0, np.where(i > 9, 0.0, val))


def np_field_f64(source_type, sources, wavenums, attens, points) -> np.ndarray:
    """Independent numpy float64 field (vectorised over sources), for cross-checks."""
    st = 9 if source_type == T4010A1 else 6
    s = _f32(sources).reshape(-1, st).astype(np.float64)
    pos = s[:, 0:3]
    if source_type == T4010A1:
        dirv, amp, phase = s[:, 3:6], s[:, 6], s[:, 7]
    else:
        dirv, amp, phase = np.zeros_like(pos), s[:, 3], s[:, 4]
    k = _f32(wavenums).astype(np.float64)
    a = _f32(attens).astype(np.float64)
    pts = _f32(points).reshape(-1, 3).astype(np.float64)
    out = np.empty(pts.shape[0], np.complex128)
    for p0 in range(0, pts.shape[0], 256):
        t = pts[p0:p0 + 256]
        diff = t[:, None, :] - pos[None, :, :]
        dist = np.sqrt((diff * diff).sum(-1))
        with np.errstate(invalid="ignore", divide="ignore"):
            cos_t = (diff * dirv[None]).sum(-1) / dist
            theta = np.arccos(cos_t)
        d = np_directivity(theta) if source_type == T4010A1 else 1.0
        r = amp[None] * d * np.exp(-dist * a[None]) / dist
        phi = phase[None] + k[None] * dist
        out[p0:p0 + 256] = (r * np.exp(1j * phi)).sum(-1)
    return out


# --------------------------------------------------------------------------- workloads
def t4010a1_sources(pos: np.ndarray, amp=1.0, phase=None, freq=40e3, direction=(0.0, 0.0, 1.0)) -> np.ndarray:
    n = pos.shape[0]
    s = np.zeros((n, 9), np.float32)
    s[:, 0:3] = pos
    s[:, 3:6] = direction
    s[:, 6] = amp
    s[:, 7] = 0.0 if phase is None else phase
    s[:, 8] = freq
    return s


def array_249(offset=(0.0, 0.0)) -> np.ndarray:
    """18 x 14 sites at 10.16 mm pitch in z = 0, sites (1,1), (14,1), (16,1) empty."""
    pitch = np.float32(10.16)
    pos = []
    for y in range(14):
        for x in range(18):
            if y == 1 and x in (1, 14, 16):
                continue
            pos.append((np.float32(offset[0]) + pitch * np.float32(x), np.float32(offset[1]) + pitch * np.float32(y),
                        0.0))
    return np.array(pos, np.float32)


def tiled_array(tiles_x: int, tiles_y: int) -> np.ndarray:
    """tiles_x x tiles_y copies of the 249-site array at offsets (182.88 i, 142.24 j)."""
    return np.concatenate([array_249((np.float32(182.88) * i, np.float32(142.24) * j))
                           for j in range(tiles_y) for i in range(tiles_x)])


# c = 340 m/s exactly: T = 273.15 * (340 / 331.3)^2 K
TEMP_340 = float(np.float32(273.15 * (340.0 / 331.3) ** 2))
WAVENUM_40K = float(np.float32(2.0 * np.pi * 40e3 / 340e3))


def focus_plane_249():
    """249-transducer array, focus at (90,70,150) mm, 100x100 XY plane at z=150."""
    pos = array_249()
    alpha = np.float32(attenuation_coef(40e3, 30.0, 1.0, 1.0, TEMP_340))
    src = t4010a1_sources(pos)
    ph = focal_point_phases(T4010A1, src, 340e3, (90.0, 70.0, 150.0))
    src[:, 7] = ph
    n = pos.shape[0]
    k = np.full(n, WAVENUM_40K, np.float32)
    a = np.full(n, alpha, np.float32)
    area = dict(order=(0, 1), ranges=((40.0, 139.0), (20.0, 119.0), (150.0, 150.0)), res=1.0)
    return src, k, a, area


def volume_3984(seed: int = 2):
    """16 tiled arrays (3984 transducers), random phases; 256x256x128 volume at 1 mm."""
    rng = np.random.Generator(np.random.PCG64(seed))
    pos = tiled_array(4, 4)
    n = pos.shape[0]
    src = t4010a1_sources(pos, phase=rng.uniform(0, 2 * np.pi, n).astype(np.float32))
    alpha = np.float32(attenuation_coef(40e3, 30.0, 1.0, 1.0, TEMP_340))
    k = np.full(n, WAVENUM_40K, np.float32)
    a = np.full(n, alpha, np.float32)
    area = dict(order=(0, 1, 2), ranges=((233.0, 488.0), (152.0, 407.0), (50.0, 177.0)), res=1.0)
    return src, k, a, area


def gspat_batch(batch: int, n_tiles=(4, 4), m_range=(2, 64), seed: int = 3):
    """Random GS-PAT instances: foci U([0,170]^2 x [100,250]) mm, targets U[0.5,1]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ms = rng.integers(m_range[0], m_range[1] + 1, batch)
    offs = np.zeros(batch + 1, np.uint32)
    offs[1:] = np.cumsum(ms)
    tot = int(offs[-1])
    foci = np.empty((tot, 3), np.float32)
    foci[:, 0:2] = rng.uniform(0, 170, (tot, 2))
    foci[:, 2] = rng.uniform(100, 250, tot)
    amps = rng.uniform(0.5, 1.0, tot).astype(np.float32)
    return foci, amps, offs
