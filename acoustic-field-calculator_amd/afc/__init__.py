"""Python binding of libafc_hip.so with the reference binding's interface.

Mirrors python/afc/afc.py:30-567 (class names, methods, argument meaning): a
script written for the reference's `afc` package runs against this backend by
importing this package instead. Differences:
  * Every calculator runs on the GPU: ``CpuCalculator`` (tag 0),
    ``GpuCalculator`` (2) and this backend's own ``HipCalculator`` (3) take
    the FAST kernels, ``AccurateCalculator`` (1) the ACCURATE ones. There is
    no CPU path.
  * ``Optimizer.ifft`` (an image read and an FFT, outside the hot path)
    raises ``AFCError`` ("not provided") and leaves the system as it was.
  * ``afc.plot_helper`` (matplotlib plotting) is not part of this package.
  * Failures raise ``AFCError`` with the library's message instead of
    aborting the process.
  * There is no binary download: the library is built in-tree
    (``acoustic-field-calculator_amd/build.py``). A missing library raises.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, byref, c_float, c_uint, c_ulong, c_void_p
from enum import IntEnum

import numpy as np

from . import nativemethods
from .nativemethods import SphereWaveSource, T4010A1, Vector3  # noqa: F401

__all__ = [
    "Axis", "UniformSystem", "SphereWaveSource", "T4010A1", "Vector3", "CpuCalculator", "AccurateCalculator",
    "GpuCalculator", "HipCalculator", "PressureFieldBuffer", "PowerFieldBuffer", "ComplexPressureFieldBuffer",
    "ScatterArea", "GridAreaBuilder", "GridArea1D", "GridArea2D", "GridArea3D", "Optimizer", "AFCError", "dll",
]


class AFCError(RuntimeError):
    pass


def dll():
    return nativemethods.load()


def _check(what: str):
    msg = dll().AFC_GetLastError()
    if msg:
        raise AFCError(f"{what}: {msg.decode()}")


class Axis(IntEnum):
    X = 1
    Y = 2
    Z = 3


class UniformSystem:
    """python/afc/afc.py:30-93."""

    def __init__(self, temperature):
        self.handle = c_void_p()
        self.temperature = temperature
        self._source_type = -1

    def __del__(self):
        if self.handle and nativemethods.AFC_DLL is not None:
            nativemethods.AFC_DLL.AFC_FreeUniformSystem(self.handle, self._source_type)
            self.handle = c_void_p()

    def set_source_type(self, v):
        self._source_type = v

    def _check_and_init(self, source):
        if self._source_type == -1:
            self._source_type = source.get_type()
            dll().AFC_CreateUniformSystem(byref(self.handle), self.temperature, self._source_type)
            _check("AFC_CreateUniformSystem")
        elif self._source_type != source.get_type():
            raise TypeError("Source type error! Mixing sound source types is not allowed.")

    def sound_speed(self):
        if self.handle:
            return dll().AFC_UniformSystemSoundSpeed(self.handle, self._source_type)
        raise TypeError("Please add sound source before.")

    def info(self) -> str:
        if self.handle:
            return dll().AFC_UniformSystemInfo(self.handle, self._source_type).decode("utf-8")
        return ""

    def info_of_source(self, idx: int) -> str:
        if self.handle:
            return dll().AFC_UniformSystemSourceInfo(self.handle, idx, self._source_type).decode("utf-8")
        return ""

    def add_wave_source(self, source):
        self._check_and_init(source)
        if isinstance(source, SphereWaveSource):
            dll().AFC_AddSphereWaveSource(self.handle, source)
        elif isinstance(source, T4010A1):
            dll().AFC_AddT4010A1(self.handle, source)
        else:
            raise TypeError("Unknown source type")
        _check("add_wave_source")

    def get_wave_sources(self):
        """Mutable view of the system's own source array (edits are seen by the next calculate)."""
        ptr = c_void_p()
        size = dll().AFC_GetWaveSources(self.handle, byref(ptr), self._source_type)
        if self._source_type == 0:
            return ctypes.cast(ptr, POINTER(SphereWaveSource * size)).contents
        if self._source_type == 1:
            return ctypes.cast(ptr, POINTER(T4010A1 * size)).contents
        raise TypeError("Unknown source type")

    def source_array(self) -> np.ndarray:
        """numpy view (n, 6 or 9) float32 of the same source array."""
        ptr = c_void_p()
        size = dll().AFC_GetWaveSources(self.handle, byref(ptr), self._source_type)
        st = 9 if self._source_type == 1 else 6
        if size == 0:
            return np.zeros((0, st), np.float32)
        buf = (c_float * (size * st)).from_address(ptr.value)
        return np.ctypeslib.as_array(buf).reshape(size, st)


class Calculator:
    """python/afc/afc.py:96-150."""

    def __init__(self):
        self.handle = c_void_p()
        dll().AFC_CreateCalculator(byref(self.handle), self.get_type())
        _check("AFC_CreateCalculator")

    def __del__(self):
        if self.handle and nativemethods.AFC_DLL is not None:
            nativemethods.AFC_DLL.AFC_FreeCalculator(self.handle, self.get_type())
            self.handle = c_void_p()

    def calculate(self, system, observe_area, field):
        ptr = c_void_p()
        ft = field.get_type()
        size = dll().AFC_Calculate(self.handle, system.handle, observe_area.handle, field.handle, byref(ptr),
                                   self.get_type(), system._source_type, observe_area.get_type(), ft)
        _check("AFC_Calculate")
        fptr = ctypes.cast(ptr, POINTER(c_float))
        if size == 0:
            return np.zeros(0, np.complex64 if ft == 2 else np.float32)
        if ft == 2:
            return np.ctypeslib.as_array(fptr, shape=(size, 2)).copy().view(np.complex64).reshape(size)
        return np.ctypeslib.as_array(fptr, shape=(size,)).copy()

    @staticmethod
    def get_type():
        raise NotImplementedError()


class CpuCalculator(Calculator):
    """The reference's CPU calculator tag (0); computed by the GPU FAST kernels here."""

    @staticmethod
    def get_type():
        return 0


class AccurateCalculator(Calculator):
    @staticmethod
    def get_type():
        return 1


class GpuCalculator(Calculator):
    @staticmethod
    def get_type():
        return 2


class HipCalculator(Calculator):
    @staticmethod
    def get_type():
        return 3


class FieldBuffer:
    def __init__(self):
        self.handle = c_void_p()


class PressureFieldBuffer(FieldBuffer):
    def __init__(self):
        super().__init__()
        dll().AFC_CreatePressureField(byref(self.handle))

    def __del__(self):
        if self.handle and nativemethods.AFC_DLL is not None:
            nativemethods.AFC_DLL.AFC_FreePressureField(self.handle)
            self.handle = c_void_p()

    @staticmethod
    def get_type():
        return 0


class PowerFieldBuffer(FieldBuffer):
    def __init__(self):
        super().__init__()
        dll().AFC_CreatePowerField(byref(self.handle))

    def __del__(self):
        if self.handle and nativemethods.AFC_DLL is not None:
            nativemethods.AFC_DLL.AFC_FreePowerField(self.handle)
            self.handle = c_void_p()

    @staticmethod
    def get_type():
        return 1


class ComplexPressureFieldBuffer(FieldBuffer):
    def __init__(self):
        super().__init__()
        dll().AFC_CreateComplexPressureField(byref(self.handle))

    def __del__(self):
        if self.handle and nativemethods.AFC_DLL is not None:
            nativemethods.AFC_DLL.AFC_FreeComplexPressureField(self.handle)
            self.handle = c_void_p()

    @staticmethod
    def get_type():
        return 2


class ScatterArea:
    def __init__(self):
        self.handle = c_void_p()
        dll().AFC_CreateScatterArea(byref(self.handle))

    def __del__(self):
        if self.handle and nativemethods.AFC_DLL is not None:
            nativemethods.AFC_DLL.AFC_FreeScatterArea(self.handle)
            self.handle = c_void_p()

    def add_observe_point(self, p):
        if not isinstance(p, Vector3):
            p = Vector3(p)
        dll().AFC_ScatterAddObservePoint(self.handle, p)

    @staticmethod
    def get_type():
        return 3


class GridArea:
    def __init__(self):
        self.handle = c_void_p()

    def __del__(self):
        if self.handle and nativemethods.AFC_DLL is not None:
            nativemethods.AFC_DLL.AFC_FreeGridArea(self.handle, self.get_dimension())
            self.handle = c_void_p()

    def get_size(self):
        f, s, t = c_uint(), c_uint(), c_uint()
        dll().AFC_GetGridSize(self.handle, self.get_dimension(), byref(f), byref(s), byref(t))
        return (f.value, s.value, t.value)


class GridArea1D(GridArea):
    @staticmethod
    def get_type():
        return 0

    @staticmethod
    def get_dimension():
        return 1


class GridArea2D(GridArea):
    @staticmethod
    def get_type():
        return 1

    @staticmethod
    def get_dimension():
        return 2


class GridArea3D(GridArea):
    @staticmethod
    def get_type():
        return 2

    @staticmethod
    def get_dimension():
        return 3


class GridAreaBuilder:
    """python/afc/afc.py:315-424 (first-specified axis varies fastest)."""

    def __init__(self):
        self._dim_cnt = 0
        self._ranges = []
        self._resolution = None
        self._axes = []

    def _add_axis(self, axis):
        if axis in self._axes:
            raise RuntimeError(f"The axis ({axis}) is already specified.")
        self._axes.append(axis)

    def x_at(self, pos: float):
        self._ranges.append((pos, pos))
        self._add_axis(Axis.X)
        return self

    def y_at(self, pos: float):
        self._ranges.append((pos, pos))
        self._add_axis(Axis.Y)
        return self

    def z_at(self, pos: float):
        self._ranges.append((pos, pos))
        self._add_axis(Axis.Z)
        return self

    def x_range(self, obs_range):
        self._ranges.append(tuple(obs_range))
        self._add_axis(Axis.X)
        self._dim_cnt += 1
        return self

    def y_range(self, obs_range):
        self._ranges.append(tuple(obs_range))
        self._add_axis(Axis.Y)
        self._dim_cnt += 1
        return self

    def z_range(self, obs_range):
        self._ranges.append(tuple(obs_range))
        self._add_axis(Axis.Z)
        self._dim_cnt += 1
        return self

    def resolution(self, r: float):
        if self._resolution is not None:
            raise RuntimeError("Resolution is already specified.")
        self._resolution = r
        return self

    def generate(self):
        if len(self._ranges) != 3:
            raise RuntimeError("Some ranges is not specified.")
        if self._resolution is None:
            raise RuntimeError("Resolution is not specified.")
        fr, sr, tr = self._ranges
        ax = [int(a) for a in self._axes]
        if self._dim_cnt in (0, 1):
            area = GridArea1D()
            dll().AFC_CreateGridArea1D(byref(area.handle), fr[0], fr[1], sr[0], tr[0], self._resolution, *ax)
        elif self._dim_cnt == 2:
            area = GridArea2D()
            dll().AFC_CreateGridArea2D(byref(area.handle), fr[0], fr[1], sr[0], sr[1], tr[0], self._resolution, *ax)
        else:
            area = GridArea3D()
            dll().AFC_CreateGridArea3D(byref(area.handle), fr[0], fr[1], sr[0], sr[1], tr[0], tr[1],
                                       self._resolution, *ax)
        _check("GridAreaBuilder.generate")
        return area


def _pack_foci(foci, amps):
    size = len(foci)
    a = np.ascontiguousarray(np.asarray(amps, np.float32))
    f = np.ascontiguousarray(np.asarray(foci, np.float32).reshape(size, 3))
    return size, f, a


class Optimizer:
    """python/afc/afc.py:427-567."""

    @staticmethod
    def focus(system: UniformSystem, p):
        if not isinstance(p, Vector3):
            p = Vector3(p)
        dll().AFO_FocalPoint(system.handle, p, system._source_type)
        _check("AFO_FocalPoint")

    @staticmethod
    def bessel(system: UniformSystem, p, direction, theta_z: float):
        """Bessel beam phases (python/afc/afc.py:427, acoustic-field-optimizer/src/bessel.rs:31-55)."""
        p = p if isinstance(p, Vector3) else Vector3(p)
        d = direction if isinstance(direction, Vector3) else Vector3(direction)
        dll().AFO_BesselBeam(system.handle, p, d, c_float(theta_z), system._source_type)
        _check("AFO_BesselBeam")

    bessel_beam = bessel

    @staticmethod
    def ifft(system: UniformSystem, path, bottom_left, top_left, bottom_right, spacing, z):
        """python/afc/afc.py:435; not provided by this backend: raises AFCError."""
        args = [v if isinstance(v, Vector3) else Vector3(v) for v in (bottom_left, top_left, bottom_right)]
        dll().AFO_IFFT(system.handle, str(path).encode("utf-8"), *args, c_float(spacing), c_float(z),
                       system._source_type)
        _check("AFO_IFFT")

    @staticmethod
    def _run(name, system, foci, amps):
        size, f, a = _pack_foci(foci, amps)
        getattr(dll(), name)(system.handle, f.ctypes.data_as(c_void_p), a.ctypes.data_as(POINTER(c_float)),
                             c_ulong(size), system._source_type)
        _check(name)

    @staticmethod
    def gs_pat(system: UniformSystem, foci, amps):
        Optimizer._run("AFO_GSPAT", system, foci, amps)

    @staticmethod
    def gs(system: UniformSystem, foci, amps):
        Optimizer._run("AFO_GS", system, foci, amps)

    @staticmethod
    def naive(system: UniformSystem, foci, amps):
        Optimizer._run("AFO_Naive", system, foci, amps)

    @staticmethod
    def lm(system: UniformSystem, foci, amps):
        Optimizer._run("AFO_LM", system, foci, amps)

    @staticmethod
    def gauss_newton(system: UniformSystem, foci, amps):
        Optimizer._run("AFO_GaussNewton", system, foci, amps)

    @staticmethod
    def gradient_descent(system: UniformSystem, foci, amps):
        Optimizer._run("AFO_GradientDescent", system, foci, amps)

    @staticmethod
    def horn(system: UniformSystem, foci, amps):
        Optimizer._run("AFO_Horn", system, foci, amps)

    @staticmethod
    def long2014(system: UniformSystem, foci, amps):
        """python/afc/afc.py:492."""
        Optimizer._run("AFO_Long", system, foci, amps)

    long = long2014

    @staticmethod
    def apo(system: UniformSystem, foci, amps, lambda_reg: float):
        size, f, a = _pack_foci(foci, amps)
        dll().AFO_APO(system.handle, f.ctypes.data_as(c_void_p), a.ctypes.data_as(POINTER(c_float)), c_ulong(size),
                      c_float(lambda_reg), system._source_type)
        _check("AFO_APO")

    @staticmethod
    def greedy(system: UniformSystem, foci, amps, phase_div: int = 16):
        size, f, a = _pack_foci(foci, amps)
        dll().AFO_Greedy(system.handle, f.ctypes.data_as(c_void_p), a.ctypes.data_as(POINTER(c_float)),
                         c_ulong(size), c_ulong(phase_div), system._source_type)
        _check("AFO_Greedy")
