"""Device-resident entry points of libafc_hip.so (include/afc_hip.h, part 2).

torch is used only for device memory and streams; every computation runs in
the library's HIP kernels through the C ABI.
"""
from __future__ import annotations

import ctypes
from ctypes import byref, c_void_p

import numpy as np

from . import AFCError, dll
from .nativemethods import AFH_PointSet

SOURCE_STRIDE = 12
FIELD_PRESSURE, FIELD_POWER, FIELD_COMPLEX = 0, 1, 2
PREC_FAST, PREC_ACCURATE = 0, 1


def _err(what: str, rc: int):
    if rc != 0:
        raise AFCError(f"{what} failed ({rc}): {dll().AFC_GetLastError().decode()}")


def packed_count(n: int) -> int:
    return int(dll().AFH_PackedCount(n))


def pack_sources(source_type: int, sources: np.ndarray, wavenums: np.ndarray, attens: np.ndarray) -> np.ndarray:
    """Reference source structs (n, 6|9) + k, alpha -> packed (packed_count(n), 12) float32."""
    src = np.ascontiguousarray(sources, np.float32)
    k = np.ascontiguousarray(wavenums, np.float32)
    a = np.ascontiguousarray(attens, np.float32)
    n = k.shape[0]
    out = np.empty((packed_count(n), SOURCE_STRIDE), np.float32)
    rc = dll().AFH_PackSources(src.ctypes.data_as(c_void_p), k.ctypes.data_as(c_void_p), a.ctypes.data_as(c_void_p),
                               n, source_type, out.ctypes.data_as(c_void_p))
    _err("AFH_PackSources", rc)
    return out


def grid_pointset(order, counts, origin, resolution) -> AFH_PointSet:
    """order: axes fastest-first (0=x,1=y,2=z); counts along that order; origin by axis."""
    ps = AFH_PointSet()
    ps.kind = 0
    for i in range(3):
        ps.order[i] = int(order[i])
        ps.count[i] = int(counts[i])
        ps.origin[i] = float(origin[i])
    ps.resolution = float(resolution)
    ps.points = None
    ps.n_points = int(counts[0]) * int(counts[1]) * int(counts[2])
    return ps


def scatter_pointset(points_dev_ptr: int, n: int) -> AFH_PointSet:
    ps = AFH_PointSet()
    ps.kind = 1
    ps.points = points_dev_ptr
    ps.n_points = n
    return ps


def field_eval(packed_dev_ptr: int, n_packed: int, source_type: int, pointset: AFH_PointSet, field_type: int,
               precision: int, out_dev_ptr: int, stream_ptr: int = 0) -> None:
    rc = dll().AFH_FieldEval(c_void_p(packed_dev_ptr), n_packed, source_type, byref(pointset), field_type, precision,
                             c_void_p(out_dev_ptr), c_void_p(stream_ptr))
    _err("AFH_FieldEval", rc)


def grid_from_area(order, ranges, res):
    """(order of ranged axes, ranges by axis, res) as the reference builder takes them ->
    a device grid point set with the builder's counts (builder.rs:198-200)."""
    cnt = []
    for a in range(3):
        lo, hi = np.float32(ranges[a][0]), np.float32(ranges[a][1])
        q = np.float32(hi - lo) / np.float32(res)
        cnt.append(int(q) + 1 if q > 0 else 1)
    full = list(order) + [a for a in range(3) if a not in order]
    return grid_pointset(full, [cnt[a] for a in full], [ranges[a][0] for a in range(3)], res), tuple(cnt)


ALGO_GSPAT, ALGO_GS, ALGO_NAIVE = 0, 1, 2
HOLO_FP32, HOLO_MIXED, HOLO_MIXED_CLOSING, HOLO_MIXED_SPLIT = 0, 1, 2, 3


class Context:
    """A HIP stream plus grow-only workspace owned by the library (AFH_CreateContext)."""

    def __init__(self):
        self.handle = c_void_p()
        _err("AFH_CreateContext", dll().AFH_CreateContext(byref(self.handle)))

    def close(self):
        if self.handle:
            dll().AFH_DestroyContext(self.handle)
            self.handle = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def holo_solve(ctx: Context, packed_dev_ptr: int, n_packed: int, n_sources: int, source_type: int, foci_dev_ptr: int,
               amps_dev_ptr: int, offsets: np.ndarray, algo: int, precision: int, iterations: int,
               out_amp_dev_ptr: int, out_phase_dev_ptr: int, stream_ptr: int = 0) -> None:
    """Batched GS-PAT / GS / Naive on device arrays; offsets (host, uint32, batch + 1)."""
    offs = np.ascontiguousarray(offsets, np.uint32)
    rc = dll().AFH_HoloSolve(ctx.handle, c_void_p(packed_dev_ptr), n_packed, n_sources, source_type,
                             c_void_p(foci_dev_ptr), c_void_p(amps_dev_ptr), offs.ctypes.data_as(c_void_p),
                             offs.shape[0] - 1, algo, precision, iterations, c_void_p(out_amp_dev_ptr),
                             c_void_p(out_phase_dev_ptr), c_void_p(stream_ptr))
    _err("AFH_HoloSolve", rc)


NLS_LM, NLS_GAUSS_NEWTON, NLS_GRADIENT_DESCENT = 0, 1, 2


def nl_solve(ctx: Context, packed_dev_ptr: int, n_packed: int, n_sources: int, source_type: int, foci_dev_ptr: int,
             amps_dev_ptr: int, offsets: np.ndarray, algo: int, k_max: int, out_phase_dev_ptr: int,
             out_iters_dev_ptr: int = 0, stream_ptr: int = 0) -> None:
    """Batched FP64 LM / Gauss-Newton / gradient descent (phases only) on device arrays."""
    offs = np.ascontiguousarray(offsets, np.uint32)
    rc = dll().AFH_NLSolve(ctx.handle, c_void_p(packed_dev_ptr), n_packed, n_sources, source_type,
                           c_void_p(foci_dev_ptr), c_void_p(amps_dev_ptr), offs.ctypes.data_as(c_void_p),
                           offs.shape[0] - 1, algo, k_max, c_void_p(out_phase_dev_ptr),
                           c_void_p(out_iters_dev_ptr or None), c_void_p(stream_ptr))
    _err("AFH_NLSolve", rc)


def lm_solve(ctx: Context, packed_dev_ptr: int, n_packed: int, n_sources: int, source_type: int, foci_dev_ptr: int,
             amps_dev_ptr: int, offsets: np.ndarray, k_max: int, out_phase_dev_ptr: int, out_iters_dev_ptr: int = 0,
             stream_ptr: int = 0) -> None:
    """Batched FP64 Levenberg-Marquardt (phases only) on device arrays."""
    offs = np.ascontiguousarray(offsets, np.uint32)
    rc = dll().AFH_LMSolve(ctx.handle, c_void_p(packed_dev_ptr), n_packed, n_sources, source_type,
                           c_void_p(foci_dev_ptr), c_void_p(amps_dev_ptr), offs.ctypes.data_as(c_void_p),
                           offs.shape[0] - 1, k_max, c_void_p(out_phase_dev_ptr), c_void_p(out_iters_dev_ptr or None),
                           c_void_p(stream_ptr))
    _err("AFH_LMSolve", rc)


def greedy_solve(ctx: Context, packed_dev_ptr: int, n_packed: int, n_sources: int, source_type: int, foci_dev_ptr: int,
                 amps_dev_ptr: int, offsets: np.ndarray, phase_div: int, out_phase_dev_ptr: int,
                 stream_ptr: int = 0) -> None:
    """Batched greedy phase search (phases only) on device arrays."""
    offs = np.ascontiguousarray(offsets, np.uint32)
    rc = dll().AFH_GreedySolve(ctx.handle, c_void_p(packed_dev_ptr), n_packed, n_sources, source_type,
                               c_void_p(foci_dev_ptr), c_void_p(amps_dev_ptr), offs.ctypes.data_as(c_void_p),
                               offs.shape[0] - 1, phase_div, c_void_p(out_phase_dev_ptr), c_void_p(stream_ptr))
    _err("AFH_GreedySolve", rc)


def apo_solve(ctx: Context, packed_dev_ptr: int, n_packed: int, n_sources: int, source_type: int, foci_dev_ptr: int,
              amps_dev_ptr: int, offsets: np.ndarray, lam: float, k_max: int, out_amp_dev_ptr: int,
              out_phase_dev_ptr: int, out_iters_dev_ptr: int = 0, stream_ptr: int = 0) -> None:
    """Batched FP64 APO (amplitudes and phases) on device arrays."""
    offs = np.ascontiguousarray(offsets, np.uint32)
    rc = dll().AFH_APOSolve(ctx.handle, c_void_p(packed_dev_ptr), n_packed, n_sources, source_type,
                            c_void_p(foci_dev_ptr), c_void_p(amps_dev_ptr), offs.ctypes.data_as(c_void_p),
                            offs.shape[0] - 1, lam, k_max, c_void_p(out_amp_dev_ptr), c_void_p(out_phase_dev_ptr),
                            c_void_p(out_iters_dev_ptr or None), c_void_p(stream_ptr))
    _err("AFH_APOSolve", rc)


def long_solve(ctx: Context, packed_dev_ptr: int, n_packed: int, n_sources: int, source_type: int, foci_dev_ptr: int,
               amps_dev_ptr: int, offsets: np.ndarray, out_amp_dev_ptr: int, out_phase_dev_ptr: int,
               stream_ptr: int = 0) -> None:
    """Batched FP64 Long (amplitudes and phases) on device arrays."""
    offs = np.ascontiguousarray(offsets, np.uint32)
    rc = dll().AFH_LongSolve(ctx.handle, c_void_p(packed_dev_ptr), n_packed, n_sources, source_type,
                             c_void_p(foci_dev_ptr), c_void_p(amps_dev_ptr), offs.ctypes.data_as(c_void_p),
                             offs.shape[0] - 1, c_void_p(out_amp_dev_ptr), c_void_p(out_phase_dev_ptr),
                             c_void_p(stream_ptr))
    _err("AFH_LongSolve", rc)


def horn_solve(ctx: Context, packed_dev_ptr: int, n_packed: int, n_sources: int, source_type: int, foci_dev_ptr: int,
               amps_dev_ptr: int, offsets: np.ndarray, out_amp_dev_ptr: int, out_phase_dev_ptr: int, seed: int = 0,
               repeat: int = 100, lam: float = 0.8, tikhonov: float = 1e-5, stream_ptr: int = 0) -> None:
    """Batched FP64 Horn (amplitudes and phases) on device arrays; instance b uses seed + b."""
    offs = np.ascontiguousarray(offsets, np.uint32)
    rc = dll().AFH_HornSolve(ctx.handle, c_void_p(packed_dev_ptr), n_packed, n_sources, source_type,
                             c_void_p(foci_dev_ptr), c_void_p(amps_dev_ptr), offs.ctypes.data_as(c_void_p),
                             offs.shape[0] - 1, seed, repeat, lam, tikhonov, c_void_p(out_amp_dev_ptr),
                             c_void_p(out_phase_dev_ptr), c_void_p(stream_ptr))
    _err("AFH_HornSolve", rc)
