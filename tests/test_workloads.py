"""The benchmark's inputs (afc.workloads, built through the library's own
UniformSystem, AFH_GetSourceConstants and AFO_FocalPoint) against the oracle's
builders of the same configurations. Host-side only: no kernel is launched."""
import numpy as np


def test_arrays_match_the_oracle(afc, oracle):
    from afc import workloads as W

    assert np.array_equal(W.array_249(), oracle.array_249())
    for tiles in ((2, 2), (4, 4), (5, 4)):
        assert np.array_equal(W.tiled_array(*tiles), oracle.tiled_array(*tiles))
    pos = oracle.tiled_array(2, 2)
    ph = np.linspace(0, 6, pos.shape[0], dtype=np.float32)
    assert np.array_equal(W.t4010a1_rows(pos, phase=ph), oracle.t4010a1_sources(pos, phase=ph))


def test_constants_are_the_reference_formulas(afc, oracle):
    from afc import workloads as W

    src, k, a = W.array_rows((1, 1))
    ko, ao = oracle.source_constants(src[:, 8].copy(), W.TEMP_340)
    assert np.array_equal(k, ko) and np.array_equal(a, ao)
    assert W.TEMP_340 == oracle.TEMP_340


def test_volume_and_focus_batch_match_the_oracle(afc, oracle):
    from afc import workloads as W

    src, k, a, area = W.volume_3984(seed=2)
    so, _, ao, areao = oracle.volume_3984(seed=2)
    assert np.array_equal(src, so) and np.array_equal(a, ao) and area == areao
    for args in ((1024, (2, 64), 3), (64, (16, 16), 4), (256, (2, 16), 8)):
        got = W.focus_batch(args[0], m_range=args[1], seed=args[2])
        want = oracle.gspat_batch(args[0], m_range=args[1], seed=args[2])
        assert all(np.array_equal(g, w) for g, w in zip(got, want))


def test_focus_plane_phases(afc, oracle):
    """FocalPoint through AFO_FocalPoint against the oracle's focal_point.rs
    restatement: every phase within 1e-5 turn (mod 2 pi)."""
    from afc import workloads as W

    src, k, a, area = W.focus_plane_249()
    so, _, _, areao = oracle.focus_plane_249()
    assert np.array_equal(src[:, :7], so[:, :7]) and area == areao
    d = np.angle(np.exp(1j * (src[:, 7].astype(np.float64) - so[:, 7])))
    assert np.max(np.abs(d)) <= 2 * np.pi * 1e-5
