"""GPU parity of the field path against the oracle (oracle/afc_oracle.c).

Tolerances (relative L2 over all points unless stated):
  * GPU vs the reference's FP32 CPU path (the oracle's *_f32): <= 1e-5;
  * GPU FAST mode vs FP64: <= 5e-6; ACCURATE mode vs FP64: <= 2e-6;
  * reductions |p|, |p|^2 and grid-vs-scatter consistency: bit-exact where the
    inputs are bit-identical.
Reference: calculator/cpu_calculator.rs:43-65, system/uniform_medium.rs:111-128,
field_buffer/*.rs, observe_area/grid/*.rs, capi/src/calculator.rs:53-117.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def build_system(afc, src, temperature, source_type=1):
    s = afc.UniformSystem(temperature)
    for row in src:
        if source_type == 1:
            s.add_wave_source(afc.T4010A1(row[0:3], row[3:6], row[6], row[7], row[8]))
        else:
            s.add_wave_source(afc.SphereWaveSource(row[0:3], row[3], row[4], row[5]))
    return s


def consts(afc, s, n):
    import ctypes
    k = np.zeros(n, np.float32)
    a = np.zeros(n, np.float32)
    afc.dll().AFH_GetSourceConstants(s.handle, k.ctypes.data_as(ctypes.c_void_p), a.ctypes.data_as(ctypes.c_void_p))
    return k, a


@pytest.fixture(scope="module")
def plane249(afc, oracle, torch_cuda):
    src, k, a, area = oracle.focus_plane_249()
    s = build_system(afc, src, oracle.TEMP_340)
    kk, aa = consts(afc, s, src.shape[0])
    pts, bounds = oracle.grid_points(**area)
    ref32 = oracle.field(1, src, kk, aa, pts, "f32")
    ref64 = oracle.field(1, src, kk, aa, pts, "f64")
    return dict(src=src, sys=s, k=kk, a=aa, pts=pts, bounds=bounds, ref32=ref32, ref64=ref64)


def plane249_area(afc):
    return afc.GridAreaBuilder().x_range((40, 139)).y_range((20, 119)).z_at(150).resolution(1.0).generate()


def test_plane249_pressure_capi(afc, oracle, plane249):
    log("case: focus_plane_249 PressureField via AFC_Calculate (GpuCalculator)")
    calc = afc.GpuCalculator()
    area = plane249_area(afc)
    assert area.get_size() == plane249["bounds"]
    got = calc.calculate(plane249["sys"], area, afc.PressureFieldBuffer())
    assert got.shape == (10000,)
    want32 = oracle.reduce_f32(0, plane249["ref32"])
    e32 = float(np.linalg.norm(got - want32) / np.linalg.norm(want32))
    e64 = float(np.linalg.norm(got - np.abs(plane249["ref64"])) / np.linalg.norm(np.abs(plane249["ref64"])))
    log(f"  |p| rel-L2 vs FP32 CPU path {e32:.3e}, vs FP64 {e64:.3e}")
    assert e32 <= 1e-5
    assert e64 <= 5e-6
    # the focus is where FocalPoint put it
    assert np.argmax(got) == np.where((plane249["pts"][:, 0] == 90) & (plane249["pts"][:, 1] == 70))[0][0]


def test_plane249_complex_and_power(afc, oracle, plane249):
    log("case: focus_plane_249 Complex/Power via AFC_Calculate")
    calc = afc.GpuCalculator()
    area = plane249_area(afc)
    cp = calc.calculate(plane249["sys"], area, afc.ComplexPressureFieldBuffer())
    e32 = rel_l2(cp, plane249["ref32"])
    e64 = rel_l2(cp, plane249["ref64"])
    log(f"  complex rel-L2 vs FP32 CPU {e32:.3e}, vs FP64 {e64:.3e}; CPU FP32 vs FP64 "
        f"{rel_l2(plane249['ref32'], plane249['ref64']):.3e}")
    assert e32 <= 1e-5 and e64 <= 5e-6
    pw = calc.calculate(plane249["sys"], area, afc.PowerFieldBuffer())
    pr = calc.calculate(plane249["sys"], area, afc.PressureFieldBuffer())
    # the reductions are applied to the same complex sum: |p|^2 and sqrt(|p|^2)
    c32 = cp.astype(np.complex64)
    ns = (c32.real * c32.real + c32.imag * c32.imag).astype(np.float32)
    assert np.array_equal(pw, ns)
    assert np.array_equal(pr, np.sqrt(ns))


def test_plane249_accurate_mode(afc, oracle, plane249):
    log("case: focus_plane_249 AccurateCalculator (polynomial sincos)")
    calc = afc.AccurateCalculator()
    cp = calc.calculate(plane249["sys"], plane249_area(afc), afc.ComplexPressureFieldBuffer())
    e64 = rel_l2(cp, plane249["ref64"])
    log(f"  accurate rel-L2 vs FP64 {e64:.3e}")
    assert e64 <= 2e-6
    assert rel_l2(cp, plane249["ref32"]) <= 1e-5


def test_hip_tag_equals_gpu_tag(afc, plane249):
    a = afc.GpuCalculator().calculate(plane249["sys"], plane249_area(afc), afc.ComplexPressureFieldBuffer())
    b = afc.HipCalculator().calculate(plane249["sys"], plane249_area(afc), afc.ComplexPressureFieldBuffer())
    assert np.array_equal(a, b)


@pytest.mark.parametrize("axes", [(1, 2, 3), (3, 1, 2), (2, 3, 1), (1, 3, 2), (2, 1, 3), (3, 2, 1)])
def test_grid3d_orders(afc, oracle, plane249, axes):
    log(f"case: grid3d axes {axes}")
    rng = {1: (60.0, 75.0), 2: (55.0, 66.0), 3: (140.0, 149.0)}
    b = afc.GridAreaBuilder()
    for ax in axes:
        getattr(b, {1: "x_range", 2: "y_range", 3: "z_range"}[ax])(rng[ax])
    area = b.resolution(1.5).generate()
    got = afc.GpuCalculator().calculate(plane249["sys"], area, afc.ComplexPressureFieldBuffer())
    pts, _ = oracle.grid_points([a - 1 for a in axes], (rng[1], rng[2], rng[3]), 1.5)
    want = oracle.field(1, plane249["src"], plane249["k"], plane249["a"], pts, "f32")
    assert got.shape == want.shape
    assert rel_l2(got, want) <= 1e-5


@pytest.mark.parametrize("builder", ["x", "y", "z", "xz", "zy", "yx"])
def test_grid_1d_2d(afc, oracle, plane249, builder):
    log(f"case: grid {builder}")
    rng = {"x": (60.0, 120.0), "y": (40.0, 100.0), "z": (100.0, 180.0)}
    at = {"x": 90.0, "y": 70.0, "z": 150.0}
    b = afc.GridAreaBuilder()
    order = [("xyz".index(c)) for c in builder]
    for c in builder:
        getattr(b, f"{c}_range")(rng[c])
    for c in "xyz":
        if c not in builder:
            getattr(b, f"{c}_at")(at[c])
    area = b.resolution(0.75).generate()
    got = afc.GpuCalculator().calculate(plane249["sys"], area, afc.ComplexPressureFieldBuffer())
    ranges = [rng[c] if c in builder else (at[c], at[c]) for c in "xyz"]
    pts, _ = oracle.grid_points(order, ranges, 0.75)
    want = oracle.field(1, plane249["src"], plane249["k"], plane249["a"], pts, "f32")
    assert rel_l2(got, want) <= 1e-5


def test_scatter_equals_grid_bitwise(afc, oracle, plane249):
    """Grid points are generated on the device with the reference's rounding, so
    the same points given as a scatter area produce identical sums."""
    log("case: scatter vs grid")
    area = plane249_area(afc)
    grid = afc.GpuCalculator().calculate(plane249["sys"], area, afc.ComplexPressureFieldBuffer())
    sc = afc.ScatterArea()
    for p in plane249["pts"]:
        sc.add_observe_point(tuple(p))
    scat = afc.GpuCalculator().calculate(plane249["sys"], sc, afc.ComplexPressureFieldBuffer())
    assert np.array_equal(grid, scat)


def test_sphere_sources_random(afc, oracle, torch_cuda):
    log("case: sphere sources, random geometry, scatter points")
    rng = np.random.default_rng(11)
    n = 301
    src = np.zeros((n, 6), np.float32)
    src[:, 0:2] = rng.uniform(0, 150, (n, 2))
    src[:, 3] = rng.uniform(0.2, 1.0, n)
    src[:, 4] = rng.uniform(-10, 10, n)  # phases outside [0, 2 pi) too
    src[:, 5] = 40e3
    s = build_system(afc, src, 300.0, source_type=0)
    k, a = consts(afc, s, n)
    pts = rng.uniform((0, 0, 20), (150, 150, 300), (3000, 3)).astype(np.float32)
    sc = afc.ScatterArea()
    for p in pts:
        sc.add_observe_point(tuple(p))
    got = afc.GpuCalculator().calculate(s, sc, afc.ComplexPressureFieldBuffer())
    want = oracle.field(0, src, k, a, pts, "f32")
    w64 = oracle.field(0, src, k, a, pts, "f64")
    log(f"  rel-L2 vs FP32 {rel_l2(got, want):.3e} vs FP64 {rel_l2(got, w64):.3e}")
    assert rel_l2(got, want) <= 1e-5 and rel_l2(got, w64) <= 5e-6


def test_t4010a1_random_directions_all_angles(afc, oracle, torch_cuda):
    """Observation points on all sides of tilted transducers: exercises every
    directivity segment, including theta > 90 degrees (D = 0)."""
    log("case: T4010A1 random directions")
    rng = np.random.default_rng(5)
    n = 97
    pos = rng.uniform(-40, 40, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    src = oracle.t4010a1_sources(pos, amp=rng.uniform(0.5, 1, n), phase=rng.uniform(0, 6.3, n),
                                 direction=d.astype(np.float32))
    s = build_system(afc, src, 300.0)
    k, a = consts(afc, s, n)
    pts = rng.uniform(-120, 120, (4000, 3)).astype(np.float32)
    sc = afc.ScatterArea()
    for p in pts:
        sc.add_observe_point(tuple(p))
    got = afc.GpuCalculator().calculate(s, sc, afc.ComplexPressureFieldBuffer())
    want = oracle.field(1, src, k, a, pts, "f32")
    w64 = oracle.field(1, src, k, a, pts, "f64")
    log(f"  rel-L2 vs FP32 {rel_l2(got, want):.3e} vs FP64 {rel_l2(got, w64):.3e}")
    assert rel_l2(got, want) <= 1e-5 and rel_l2(got, w64) <= 5e-6
    # pointwise: the largest error relative to the local field magnitude scale
    scale = np.abs(w64).max()
    assert np.max(np.abs(got - w64)) <= 1e-5 * scale


def test_in_place_source_edit_is_seen(afc, oracle, plane249):
    """AFC_GetWaveSources hands out the system's own array (python/main.py:58-63)."""
    log("case: in-place phase edit")
    s = build_system(afc, plane249["src"], oracle.TEMP_340)
    calc = afc.GpuCalculator()
    area = plane249_area(afc)
    before = calc.calculate(s, area, afc.ComplexPressureFieldBuffer())
    srcs = s.get_wave_sources()
    for i in range(len(srcs)):
        srcs[i].phase = 0.0
    after = calc.calculate(s, area, afc.ComplexPressureFieldBuffer())
    src0 = plane249["src"].copy()
    src0[:, 7] = 0
    want = oracle.field(1, src0, plane249["k"], plane249["a"], plane249["pts"], "f32")
    assert rel_l2(after, want) <= 1e-5 and rel_l2(before, after) > 0.1


def test_empty_inputs(afc, oracle, plane249):
    log("case: empty source list / empty area")
    calc = afc.GpuCalculator()
    empty = afc.UniformSystem(300.0)
    empty.add_wave_source(afc.T4010A1((0, 0, 0), (0, 0, 1), 1.0, 0.0, 40e3))
    # no points
    assert calc.calculate(plane249["sys"], afc.ScatterArea(), afc.PressureFieldBuffer()).shape == (0,)
    # amplitude zero everywhere -> exactly zero field
    srcs = empty.get_wave_sources()
    srcs[0].amp = 0.0
    out = calc.calculate(empty, plane249_area(afc), afc.PressureFieldBuffer())
    assert np.all(out == 0)


def test_point_on_a_source_is_not_finite(afc, oracle, plane249):
    """At r = 0 the reference divides by zero (uniform_medium.rs:117): non-finite."""
    sc = afc.ScatterArea()
    sc.add_observe_point(tuple(plane249["src"][5, 0:3]))
    sc.add_observe_point((90.0, 70.0, 150.0))
    out = afc.GpuCalculator().calculate(plane249["sys"], sc, afc.PressureFieldBuffer())
    assert not np.isfinite(out[0]) and np.isfinite(out[1])


def test_split_and_point_parallel_kernels_agree(afc, oracle, torch_cuda):
    """Point sets below 200 000 points run the source-split kernel, larger ones the
    point-parallel kernel: the same points through both agree to FP32 rounding
    and both match the FP64 oracle. The points reach ~750 mm from sources,
    where the FP32 rounding of p - s and |p - s|^2 (shared with the reference's
    FP32 path) is ~3e-5 rad of phase: FP64 tolerance 5e-5 here."""
    log("case: 640x640 grid (point-parallel) vs scatter subsets (source-split)")
    import torch
    from afc import device as D
    src, k, a, _ = oracle.volume_3984(seed=7)
    src = src[:996]
    k, a = k[:996], a[:996]
    packed = D.pack_sources(1, src, k, a)
    dev = torch.device("cuda:0")
    src_d = torch.from_numpy(packed).to(dev)
    ranges = ((10.0, 329.75), (0.0, 319.75), (120.0, 120.0))
    ps, cnt = D.grid_from_area((0, 1), ranges, 0.5)
    n_big = cnt[0] * cnt[1] * cnt[2]
    assert n_big >= 200000
    out_big = torch.empty((n_big, 2), dtype=torch.float32, device=dev)
    D.field_eval(src_d.data_ptr(), packed.shape[0], 1, ps, 2, 0, out_big.data_ptr())
    pts_all, _ = oracle.grid_points((0, 1), ranges, 0.5)
    sel = np.random.default_rng(0).choice(n_big, 5000, replace=False)
    pts = torch.from_numpy(pts_all[sel]).to(dev)
    out_small = torch.empty((sel.size, 2), dtype=torch.float32, device=dev)
    D.field_eval(src_d.data_ptr(), packed.shape[0], 1, D.scatter_pointset(pts.data_ptr(), sel.size), 2, 0,
                 out_small.data_ptr())
    torch.cuda.synchronize()
    big = out_big.cpu().numpy().view(np.complex64).reshape(-1)[sel]
    small = out_small.cpu().numpy().view(np.complex64).reshape(-1)
    want = oracle.field(1, src, k, a, pts_all[sel], "f64")
    log(f"  split vs point-parallel {rel_l2(small, big):.2e}; vs FP64 {rel_l2(small, want):.2e} / "
        f"{rel_l2(big, want):.2e}")
    assert rel_l2(small, big) <= 2e-6
    assert rel_l2(small, want) <= 5e-5 and rel_l2(big, want) <= 5e-5


@pytest.mark.parametrize("order,ranges,res", [
    ((0, 1, 2), ((40.0, 189.0), (20.0, 520.5), (150.0, 150.5)), 0.5),      # second axis y, 1002 points
    ((1, 2, 0), ((60.0, 61.0), (30.0, 159.0), (100.0, 358.0)), 0.5),      # second axis z, 517
    ((2, 0, 1), ((0.0, 375.0), (70.0, 70.5), (120.0, 189.0)), 0.25),      # second axis x, 1501
    ((0, 1), ((10.0, 329.5), (0.0, 250.5), (130.0, 130.0)), 0.5),         # 2-D, 502
])
def test_large_grid_runs_vs_oracle(afc, oracle, torch_cuda, order, ranges, res):
    """Grids of 2e5+ points (the point-parallel kernel) in every axis order and
    with counts that are not multiples of 4 or 256: every point is written and
    sampled points match the FP32 CPU path to 1e-5."""
    import torch
    from afc import device as D
    src, k, a, _ = oracle.volume_3984(seed=11)
    src, k, a = src[:249], k[:249], a[:249]
    packed = D.pack_sources(1, src, k, a)
    dev = torch.device("cuda:0")
    src_d = torch.from_numpy(packed).to(dev)
    ps, cnt = D.grid_from_area(order, ranges, res)
    n = cnt[0] * cnt[1] * cnt[2]
    log(f"case: grid order {order} counts {cnt} ({n} points)")
    assert n >= 200000
    out = torch.full((n, 2), float("nan"), dtype=torch.float32, device=dev)
    D.field_eval(src_d.data_ptr(), packed.shape[0], 1, ps, 2, 0, out.data_ptr())
    torch.cuda.synchronize()
    got_all = out.cpu().numpy().view(np.complex64).reshape(-1)
    assert np.all(np.isfinite(got_all))
    pts_all, _ = oracle.grid_points(order, ranges, res)
    sel = np.random.default_rng(1).choice(n, 4000, replace=False)
    want = oracle.field(1, src, k, a, pts_all[sel], "f32")
    err = rel_l2(got_all[sel], want)
    log(f"  rel-L2 vs FP32 CPU path at 4000 sampled points {err:.2e}")
    assert err <= 1e-5


@pytest.mark.parametrize("n", [1, 63, 255, 256, 257, 199999, 200000, 200001])
def test_ragged_point_counts(afc, oracle, torch_cuda, n):
    """Point counts around the workgroup size (256 points) and the source-split
    threshold (200 000): every point is written, none beyond, and sampled
    points match the FP32 CPU path to 1e-5."""
    log(f"case: ragged scatter set of {n} points")
    import torch
    from afc import device as D
    src, k, a, _ = oracle.focus_plane_249()
    packed = D.pack_sources(1, src, k, a)
    dev = torch.device("cuda:0")
    src_d = torch.from_numpy(packed).to(dev)
    rng = np.random.default_rng(n)
    pts = np.stack([rng.uniform(0, 180, n), rng.uniform(0, 140, n), rng.uniform(30, 300, n)], 1).astype(np.float32)
    pts_d = torch.from_numpy(pts).to(dev)
    out = torch.full((n + 64, 2), 7.0, dtype=torch.float32, device=dev)  # 64 guard points after the end
    D.field_eval(src_d.data_ptr(), packed.shape[0], 1, D.scatter_pointset(pts_d.data_ptr(), n), D.FIELD_COMPLEX,
                 D.PREC_FAST, out.data_ptr())
    torch.cuda.synchronize()
    host = out.cpu().numpy()
    assert np.all(host[n:] == 7.0)
    got = host[:n].view(np.complex64).reshape(-1)
    sel = np.unique(np.concatenate([[0, n - 1], rng.choice(n, min(n, 300), replace=False)]))
    want = oracle.field(1, src, k, a, pts[sel], "f32")
    assert np.all(np.isfinite(got))
    assert rel_l2(got[sel], want) <= 1e-5


def test_volume_3984_full_size(afc, oracle, torch_cuda):
    """The whole 3984-source x 256x256x128 volume (8 388 608 points, complex
    output) through AFH_FieldEval: every point finite, and 4096 sampled points
    (plus the eight corners) within 1e-5 relative L2 of the reference's FP32
    CPU path and 5e-6 of FP64 (cpu_calculator.rs:43-65, uniform_medium.rs:122-128)."""
    import torch
    from afc import device as D
    src, k, a, area = oracle.volume_3984()
    packed = D.pack_sources(1, src, k, a)
    dev = torch.device("cuda:0")
    src_d = torch.from_numpy(packed).to(dev)
    ps, cnt = D.grid_from_area(area["order"], area["ranges"], area["res"])
    n = cnt[0] * cnt[1] * cnt[2]
    log(f"case: volume_3984 full size, {src.shape[0]} sources x counts {cnt} ({n} points)")
    assert n == 256 * 256 * 128
    out = torch.full((n, 2), float("nan"), dtype=torch.float32, device=dev)
    D.field_eval(src_d.data_ptr(), packed.shape[0], 1, ps, D.FIELD_COMPLEX, D.PREC_FAST, out.data_ptr())
    torch.cuda.synchronize()
    assert bool(torch.isfinite(out).all())
    got_all = out.cpu().numpy().view(np.complex64).reshape(-1)
    del out
    pts_all, bounds = oracle.grid_points(area["order"], area["ranges"], area["res"])
    assert pts_all.shape[0] == n
    corners = [0, cnt[0] - 1, cnt[0] * (cnt[1] - 1), cnt[0] * cnt[1] - 1]
    corners += [c + cnt[0] * cnt[1] * (cnt[2] - 1) for c in corners]
    sel = np.unique(np.concatenate([corners, np.random.default_rng(2).choice(n, 4096, replace=False)]))
    want32 = oracle.field(1, src, k, a, pts_all[sel], "f32")
    want64 = oracle.field(1, src, k, a, pts_all[sel], "f64")
    e32, e64 = rel_l2(got_all[sel], want32), rel_l2(got_all[sel], want64)
    log(f"  {sel.size} sampled points: rel-L2 vs FP32 CPU path {e32:.2e}, vs FP64 {e64:.2e}")
    assert e32 <= 1e-5 and e64 <= 5e-6
