"""Synthetic inputs of the benchmark configurations, built through this
library (no data files, nothing here evaluates a field):

- the reference's 249-transducer array (18 x 14 sites at 10.16 mm pitch with
  three empty sites, as examples/main.rs lays it out) and its tilings;
- T4010A1 source rows {x, y, z, nx, ny, nz, amp, phase, freq};
- per-source wavenumber and attenuation as a UniformSystem computes them
  (system/uniform_medium.rs:102-107, through AFH_GetSourceConstants);
- FocalPoint phases (focal_point.rs:28-39, through AFO_FocalPoint);
- observation grids and seeded batches of focus sets.
"""

from ctypes import c_void_p

import numpy as np

from . import Optimizer, UniformSystem, dll
from .nativemethods import T4010A1

# c = 340 m/s exactly: T = 273.15 (340 / 331.3)^2 K (sound_speed.rs:22-24)
TEMP_340 = float(np.float32(273.15 * (340.0 / 331.3) ** 2))
PITCH = np.float32(10.16)
EMPTY_SITES = ((1, 1), (14, 1), (16, 1))  # (x, y) sites without a transducer
TILE_OFFSET = (np.float32(182.88), np.float32(142.24))


def array_249(offset=(0.0, 0.0)) -> np.ndarray:
    """Positions (249, 3) in z = 0, x fastest."""
    xs = np.arange(18, dtype=np.float32) * PITCH + np.float32(offset[0])
    ys = np.arange(14, dtype=np.float32) * PITCH + np.float32(offset[1])
    keep = [(x, y) for y in range(14) for x in range(18) if (x, y) not in EMPTY_SITES]
    return np.array([(xs[x], ys[y], 0.0) for x, y in keep], np.float32)


def tiled_array(tiles_x: int, tiles_y: int) -> np.ndarray:
    """tiles_x x tiles_y copies of the 249-site array, row of tiles by row."""
    return np.concatenate([array_249((TILE_OFFSET[0] * i, TILE_OFFSET[1] * j))
                           for j in range(tiles_y) for i in range(tiles_x)])


def t4010a1_rows(pos: np.ndarray, amp=1.0, phase=None, freq=40e3) -> np.ndarray:
    """Source rows (n, 9) facing +z."""
    rows = np.zeros((pos.shape[0], 9), np.float32)
    rows[:, 0:3] = pos
    rows[:, 5] = 1.0
    rows[:, 6] = amp
    if phase is not None:
        rows[:, 7] = phase
    rows[:, 8] = freq
    return rows


def system_of(rows: np.ndarray, temperature: float = TEMP_340) -> UniformSystem:
    """A T4010A1 UniformSystem holding the given rows."""
    s = UniformSystem(temperature)
    for r in rows:
        s.add_wave_source(T4010A1(tuple(r[0:3]), tuple(r[3:6]), float(r[6]), float(r[7]), float(r[8])))
    return s


def source_constants(system: UniformSystem):
    """(wavenumbers [1/mm], attenuations [Np/mm]) of the system's sources."""
    n = system.source_array().shape[0]
    k = np.empty(n, np.float32)
    a = np.empty(n, np.float32)
    dll().AFH_GetSourceConstants(system.handle, k.ctypes.data_as(c_void_p), a.ctypes.data_as(c_void_p))
    return k, a


def focus_plane_249():
    """249 transducers focused at (90, 70, 150) mm by FocalPoint; the 100 x 100
    XY plane at z = 150 mm, 1 mm apart. Returns (rows, k, a, area)."""
    s = system_of(t4010a1_rows(array_249()))
    Optimizer.focus(s, (90.0, 70.0, 150.0))
    k, a = source_constants(s)
    area = dict(order=(0, 1), ranges=((40.0, 139.0), (20.0, 119.0), (150.0, 150.0)), res=1.0)
    return s.source_array().copy(), k, a, area


def volume_3984(seed: int = 2):
    """4 x 4 tiled arrays (3984 transducers) with seeded uniform phases; the
    256 x 256 x 128 volume at 1 mm above them. Returns (rows, k, a, area)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    pos = tiled_array(4, 4)
    rows = t4010a1_rows(pos, phase=rng.uniform(0, 2 * np.pi, pos.shape[0]).astype(np.float32))
    k, a = source_constants(system_of(rows))
    area = dict(order=(0, 1, 2), ranges=((233.0, 488.0), (152.0, 407.0), (50.0, 177.0)), res=1.0)
    return rows, k, a, area


def array_rows(tiles=(1, 1), phase_seed=None):
    """Rows, k and a of a tiled array (optionally with seeded uniform phases)."""
    pos = array_249() if tuple(tiles) == (1, 1) else tiled_array(*tiles)
    phase = None
    if phase_seed is not None:
        phase = np.random.default_rng(phase_seed).uniform(0, 2 * np.pi, pos.shape[0]).astype(np.float32)
    rows = t4010a1_rows(pos, phase=phase)
    k, a = source_constants(system_of(rows))
    return rows, k, a


def focus_batch(batch: int, m_range=(2, 64), seed: int = 3):
    """batch focus sets: M ~ U{m_range}, foci U([0, 170]^2 x [100, 250]) mm,
    target amplitudes U[0.5, 1]. Returns (foci (sum M, 3), amps, offsets)."""
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
