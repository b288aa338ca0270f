"""ctypes declarations of libafc_hip.so.

The first group repeats the argtypes of the reference binding
(python/afc/nativemethods.py:113-234) so that code written against the
reference's `afc` package drives this library unchanged; the second group
declares the entry points this backend adds (include/afc_hip.h, part 2).
"""
import ctypes
import os
from ctypes import POINTER, Structure, c_char_p, c_float, c_int, c_uint, c_uint32, c_uint64, c_ulong, c_void_p

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "libafc_hip.so")

AFC_DLL = None


class Vector3(Structure):
    """nalgebra Vector3<f32> by value (python/afc/nativemethods.py:18-25)."""
    _fields_ = [("x", c_float), ("y", c_float), ("z", c_float)]

    def __init__(self, position=(0.0, 0.0, 0.0)):
        super().__init__()
        self.x, self.y, self.z = (float(v) for v in position)


class SphereWaveSource(Structure):
    """core/wave_sources/sphere_sources.rs:19-27."""
    _fields_ = [("x", c_float), ("y", c_float), ("z", c_float), ("amp", c_float), ("phase", c_float),
                ("freq", c_float)]

    def __init__(self, position=(0.0, 0.0, 0.0), amp=1.0, phase=0.0, freq=40e3):
        super().__init__()
        self.pos = position
        self.amp, self.phase, self.freq = amp, phase, freq

    @property
    def pos(self):
        return (self.x, self.y, self.z)

    @pos.setter
    def pos(self, position):
        self.x, self.y, self.z = (float(v) for v in position)

    @staticmethod
    def get_type():
        return 0


class T4010A1(Structure):
    """core/wave_sources/t4010a1.rs:73-82."""
    _fields_ = [("x", c_float), ("y", c_float), ("z", c_float), ("nx", c_float), ("ny", c_float), ("nz", c_float),
                ("amp", c_float), ("phase", c_float), ("freq", c_float)]

    def __init__(self, position=(0.0, 0.0, 0.0), direction=(0.0, 0.0, 1.0), amp=1.0, phase=0.0, freq=40e3):
        super().__init__()
        self.pos = position
        self.direction = direction
        self.amp, self.phase, self.freq = amp, phase, freq

    @property
    def pos(self):
        return (self.x, self.y, self.z)

    @pos.setter
    def pos(self, position):
        self.x, self.y, self.z = (float(v) for v in position)

    @property
    def direction(self):
        return (self.nx, self.ny, self.nz)

    @direction.setter
    def direction(self, d):
        self.nx, self.ny, self.nz = (float(v) for v in d)

    # the reference binding's getter is named `dir` (python/afc/nativemethods.py:86-88)
    dir = property(lambda self: (self.nx, self.ny, self.nz))

    @staticmethod
    def get_type():
        return 1


class AFS_Config(Structure):
    _fields_ = [("hbm_budget_bytes", c_uint64), ("pinned_budget_bytes", c_uint64), ("n_streams", c_uint32),
                ("max_gspat_batch", c_uint32), ("gspat_precision", ctypes.c_int32), ("gspat_iterations", c_uint32),
                ("gspat_cus", c_uint32), ("gspat_graphs", c_uint32)]


class AFH_PointSet(Structure):
    _fields_ = [("kind", ctypes.c_int32), ("order", ctypes.c_int32 * 3), ("count", c_uint32 * 3),
                ("origin", c_float * 3), ("resolution", c_float), ("points", c_void_p), ("n_points", c_uint64)]


def load(path: str = LIB_PATH):
    """Loads the HIP library. There is no fallback: a missing library is an error."""
    global AFC_DLL
    if AFC_DLL is not None:
        return AFC_DLL
    try:
        # One HIP runtime per process: when torch (which bundles its own
        # libamdhip64.so.7) is present, load it first so that this library's
        # libamdhip64.so.7 dependency resolves to the copy already mapped.
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(path):
        raise ImportError(f"libafc_hip.so not found at {path}: build it with "
                          f"`python acoustic-field-calculator_amd/build.py` (or __graft_entry__.build())")
    dll = ctypes.CDLL(path)
    _declare(dll)
    AFC_DLL = dll
    return dll


def _sig(dll, name, argtypes, restype=None):
    f = getattr(dll, name)
    f.argtypes = argtypes
    f.restype = restype


def _declare(d):
    # ---- reference surface (python/afc/nativemethods.py:113-234)
    _sig(d, "AFC_CreateCalculator", [POINTER(c_void_p), c_int])
    _sig(d, "AFC_FreeCalculator", [c_void_p, c_int])
    _sig(d, "AFC_Calculate", [c_void_p, c_void_p, c_void_p, c_void_p, POINTER(c_void_p), c_int, c_int, c_int, c_int],
         c_ulong)
    for nm in ("Pressure", "Power", "ComplexPressure"):
        _sig(d, f"AFC_Create{nm}Field", [POINTER(c_void_p)])
        _sig(d, f"AFC_Free{nm}Field", [c_void_p])
    _sig(d, "AFC_CreateScatterArea", [POINTER(c_void_p)])
    _sig(d, "AFC_ScatterAddObservePoint", [c_void_p, Vector3])
    _sig(d, "AFC_FreeScatterArea", [c_void_p])
    _sig(d, "AFC_CreateGridArea1D", [POINTER(c_void_p), c_float, c_float, c_float, c_float, c_float, c_int, c_int,
                                     c_int])
    _sig(d, "AFC_CreateGridArea2D", [POINTER(c_void_p), c_float, c_float, c_float, c_float, c_float, c_float, c_int,
                                     c_int, c_int])
    _sig(d, "AFC_CreateGridArea3D", [POINTER(c_void_p), c_float, c_float, c_float, c_float, c_float, c_float, c_float,
                                     c_int, c_int, c_int])
    _sig(d, "AFC_FreeGridArea", [c_void_p, c_int])
    _sig(d, "AFC_GetGridSize", [c_void_p, c_int, POINTER(c_uint), POINTER(c_uint), POINTER(c_uint)])
    _sig(d, "AFO_FocalPoint", [c_void_p, Vector3, c_int])
    _sig(d, "AFO_BesselBeam", [c_void_p, Vector3, Vector3, c_float, c_int])
    _sig(d, "AFO_IFFT", [c_void_p, c_char_p, Vector3, Vector3, Vector3, c_float, c_float, c_int])
    for nm in ("GSPAT", "GS", "Naive", "Horn", "Long", "GaussNewton", "LM", "GradientDescent"):
        _sig(d, f"AFO_{nm}", [c_void_p, c_void_p, POINTER(c_float), c_ulong, c_int])
    _sig(d, "AFO_APO", [c_void_p, c_void_p, POINTER(c_float), c_ulong, c_float, c_int])
    _sig(d, "AFO_Greedy", [c_void_p, c_void_p, POINTER(c_float), c_ulong, c_ulong, c_int])
    _sig(d, "AFC_CreateUniformSystem", [c_void_p, c_float, c_int])
    _sig(d, "AFC_FreeUniformSystem", [c_void_p, c_int])
    _sig(d, "AFC_AddSphereWaveSource", [c_void_p, SphereWaveSource])
    _sig(d, "AFC_AddT4010A1", [c_void_p, T4010A1])
    _sig(d, "AFC_GetWaveSources", [c_void_p, POINTER(c_void_p), c_int], c_ulong)
    _sig(d, "AFC_UniformSystemSoundSpeed", [c_void_p, c_int], c_float)
    _sig(d, "AFC_UniformSystemInfo", [c_void_p, c_int], c_char_p)
    _sig(d, "AFC_UniformSystemSourceInfo", [c_void_p, c_ulong, c_int], c_char_p)
    # ---- additions of this backend (include/afc_hip.h, part 2)
    _sig(d, "AFC_GetLastError", [], c_char_p)
    _sig(d, "AFH_GetSourceConstants", [c_void_p, c_void_p, c_void_p], c_uint64)
    _sig(d, "AFH_PackedCount", [c_uint64], c_uint64)
    _sig(d, "AFH_CreateContext", [POINTER(c_void_p)], ctypes.c_int32)
    _sig(d, "AFH_DestroyContext", [c_void_p])
    _sig(d, "AFH_HoloSolve", [c_void_p, c_void_p, c_uint64, c_uint64, ctypes.c_int32, c_void_p, c_void_p, c_void_p,
                              c_uint32, ctypes.c_int32, ctypes.c_int32, c_uint32, c_void_p, c_void_p, c_void_p],
         ctypes.c_int32)
    _sig(d, "AFH_HoloSolveNoAlloc", [c_void_p, c_uint64, c_uint64, ctypes.c_int32, c_void_p, c_void_p, c_void_p,
                                     c_uint32, ctypes.c_int32, ctypes.c_int32, c_uint32, c_void_p, c_void_p, c_void_p],
         ctypes.c_int32)
    _sig(d, "AFH_LMSolve", [c_void_p, c_void_p, c_uint64, c_uint64, ctypes.c_int32, c_void_p, c_void_p, c_void_p,
                            c_uint32, c_uint32, c_void_p, c_void_p, c_void_p], ctypes.c_int32)
    _sig(d, "AFH_NLSolve", [c_void_p, c_void_p, c_uint64, c_uint64, ctypes.c_int32, c_void_p, c_void_p, c_void_p,
                            c_uint32, ctypes.c_int32, c_uint32, c_void_p, c_void_p, c_void_p], ctypes.c_int32)
    _sig(d, "AFH_APOSolve", [c_void_p, c_void_p, c_uint64, c_uint64, ctypes.c_int32, c_void_p, c_void_p, c_void_p,
                             c_uint32, c_float, c_uint32, c_void_p, c_void_p, c_void_p, c_void_p], ctypes.c_int32)
    _sig(d, "AFH_LongSolve", [c_void_p, c_void_p, c_uint64, c_uint64, ctypes.c_int32, c_void_p, c_void_p, c_void_p,
                              c_uint32, c_void_p, c_void_p, c_void_p], ctypes.c_int32)
    _sig(d, "AFH_HornSolve", [c_void_p, c_void_p, c_uint64, c_uint64, ctypes.c_int32, c_void_p, c_void_p, c_void_p,
                              c_uint32, c_uint64, c_uint32, c_float, c_float, c_void_p, c_void_p, c_void_p],
         ctypes.c_int32)
    _sig(d, "AFH_GreedySolve", [c_void_p, c_void_p, c_uint64, c_uint64, ctypes.c_int32, c_void_p, c_void_p, c_void_p,
                                c_uint32, c_uint32, c_void_p, c_void_p], ctypes.c_int32)
    _sig(d, "AFS_Create", [POINTER(AFS_Config), POINTER(c_void_p)], ctypes.c_int32)
    _sig(d, "AFS_Destroy", [c_void_p])
    _sig(d, "AFS_AddArray", [c_void_p, c_void_p, c_void_p, c_void_p, c_uint64, ctypes.c_int32, POINTER(c_uint32)],
         ctypes.c_int32)
    _sig(d, "AFS_SubmitField", [c_void_p, c_uint32, POINTER(AFH_PointSet), ctypes.c_int32, c_void_p,
                                POINTER(c_uint64)], ctypes.c_int32)
    _sig(d, "AFS_SubmitGSPAT", [c_void_p, c_uint32, c_void_p, c_void_p, c_uint32, c_void_p, c_void_p,
                                POINTER(c_uint64)], ctypes.c_int32)
    _sig(d, "AFS_SubmitGSPATBatch", [c_void_p, c_uint32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_void_p], ctypes.c_int32)
    _sig(d, "AFS_Wait", [c_void_p, c_uint64, POINTER(ctypes.c_double)], ctypes.c_int32)
    _sig(d, "AFS_Stats", [c_void_p, ctypes.c_int32, POINTER(ctypes.c_double), POINTER(ctypes.c_double),
                          POINTER(c_uint64), ctypes.c_int32], ctypes.c_int32)
    _sig(d, "AFS_GraphCounters", [c_void_p, POINTER(c_uint64), POINTER(c_uint64)], ctypes.c_int32)
    _sig(d, "AFH_DirectivityT4010A1", [c_void_p, c_uint64, ctypes.c_int32, c_void_p], ctypes.c_int32)
    _sig(d, "AFH_PackSources", [c_void_p, c_void_p, c_void_p, c_uint64, ctypes.c_int32, c_void_p], ctypes.c_int32)
    _sig(d, "AFH_FieldEval", [c_void_p, c_uint64, ctypes.c_int32, POINTER(AFH_PointSet), ctypes.c_int32,
                              ctypes.c_int32, c_void_p, c_void_p], ctypes.c_int32)
