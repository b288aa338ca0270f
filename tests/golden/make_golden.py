"""Writes the golden fixtures of this directory from the FP64 oracle.

The reference ships no fixtures or golden vectors for the field or optimizer
path (its only tests are in fmath), so these vectors are self-derived: FP64
restatements of system/uniform_medium.rs:111-128, gs_pat.rs:38-81,
nls/{lm,gauss_newton,gradient_descent,apo}.rs, combination/greedy.rs and
matrix/{long,horn}.rs on small seeded inputs. Each fixture is also checked by a
property in tests/test_golden.py that does not reuse the oracle's formulas
(focus identity, single-focus closed forms, candidate phases, objective
decrease, camax normalisation), so a wrong fixture cannot pass silently.

Run from the repository root:  python tests/golden/make_golden.py
Output: tests/golden/*.npz (numpy savez, no pickled objects).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402


def field_t4010a1():
    """Config-1 array (249 T4010A1, 40 kHz, c = 340 m/s) with FocalPoint phases
    for (90, 70, 150) mm; 64 seeded scatter points plus the focus (last)."""
    src, k, a, _ = O.focus_plane_249()
    rng = np.random.default_rng(11)
    pts = np.stack([rng.uniform(0, 180, 64), rng.uniform(0, 140, 64), rng.uniform(20, 300, 64)], 1)
    pts = np.vstack([pts, [[90.0, 70.0, 150.0]]]).astype(np.float32)
    return dict(src=src, k=k, a=a, points=pts, field_f64=O.field(1, src, k, a, pts, "f64"),
                field_f32=O.field(1, src, k, a, pts, "f32"))


def field_sphere():
    """48 sphere sources (random positions, amplitudes, phases), 40 points."""
    rng = np.random.default_rng(12)
    n = 48
    src = np.zeros((n, 6), np.float32)
    src[:, 0:2] = rng.uniform(0, 100, (n, 2))
    src[:, 3] = rng.uniform(0.2, 1.5, n)
    src[:, 4] = rng.uniform(0, 2 * np.pi, n)
    src[:, 5] = 40e3
    k = np.full(n, O.WAVENUM_40K, np.float32)
    a = np.full(n, np.float32(O.attenuation_coef(40e3, 30, 1, 1, O.TEMP_340)), np.float32)
    pts = np.stack([rng.uniform(-50, 150, 40), rng.uniform(-50, 150, 40), rng.uniform(10, 250, 40)],
                   1).astype(np.float32)
    return dict(src=src, k=k, a=a, points=pts, field_f64=O.field(0, src, k, a, pts, "f64"),
                field_f32=O.field(0, src, k, a, pts, "f32"))


def optim_problem(m, seed):
    pos = O.array_249()
    src = O.t4010a1_sources(pos)
    n = pos.shape[0]
    k = np.full(n, O.WAVENUM_40K, np.float32)
    a = np.full(n, np.float32(O.attenuation_coef(40e3, 30, 1, 1, O.TEMP_340)), np.float32)
    rng = np.random.default_rng(seed)
    foci = np.stack([rng.uniform(40, 130, m), rng.uniform(30, 110, m), rng.uniform(120, 200, m)],
                    1).astype(np.float32)
    amps = rng.uniform(0.5, 1.0, m).astype(np.float32)
    return src, k, a, foci, amps


def gspat(m, seed):
    src, k, a, foci, amps = optim_problem(m, seed)
    amp, phase = O.gspat(1, src, k, a, foci, amps, "f64")
    return dict(src=src, k=k, a=a, foci=foci, amps=amps, out_amp=amp, out_phase=phase)


def lm(m, seed, k_max):
    src, k, a, foci, amps = optim_problem(m, seed)
    phase, x, iters = O.lm(1, src, k, a, foci, amps, "f64", k_max=k_max)
    return dict(src=src, k=k, a=a, foci=foci, amps=amps, k_max=np.int32(k_max), out_phase=phase, out_x=x,
                iterations=np.int32(iters))


def greedy(m, seed, div):
    src, k, a, foci, amps = optim_problem(m, seed)
    phase = O.greedy(1, src, k, a, foci, amps, div, "f64")
    return dict(src=src, k=k, a=a, foci=foci, amps=amps, phase_div=np.int32(div), out_phase=phase)


def nls(fn, m, seed, k_max):
    src, k, a, foci, amps = optim_problem(m, seed)
    phase, x, iters = fn(1, src, k, a, foci, amps, "f64", k_max=k_max)
    return dict(src=src, k=k, a=a, foci=foci, amps=amps, k_max=np.int32(k_max), out_phase=phase, out_x=x,
                iterations=np.int32(iters))


def apo(m, seed, lam):
    src, k, a, foci, amps = optim_problem(m, seed)
    amps = (amps * 0.2).astype(np.float32)
    lam = np.float32(lam)  # the C ABI takes lambda as f32
    amp, phase, iters = O.apo(1, src, k, a, foci, amps, float(lam), "f64")
    return dict(src=src, k=k, a=a, foci=foci, amps=amps, lam=lam, out_amp=amp, out_phase=phase,
                iterations=np.int32(iters))


def long_(m, seed):
    src, k, a, foci, amps = optim_problem(m, seed)
    amp, phase = O.long_(1, src, k, a, foci, amps)
    return dict(src=src, k=k, a=a, foci=foci, amps=amps, out_amp=amp, out_phase=phase)


def horn(m, seed, rseed):
    src, k, a, foci, amps = optim_problem(m, seed)
    amp, phase = O.horn(1, src, k, a, foci, amps, seed=rseed, lam=np.float32(0.8), alpha=np.float32(1e-5))
    return dict(src=src, k=k, a=a, foci=foci, amps=amps, seed=np.uint64(rseed), out_amp=amp, out_phase=phase)


def main():
    fixtures = {
        "field_t4010a1.npz": field_t4010a1(),
        "field_sphere.npz": field_sphere(),
        "gspat_m1.npz": gspat(1, 21),
        "gspat_m5.npz": gspat(5, 22),
        "lm_m1.npz": lm(1, 23, 200),
        "lm_m4_20it.npz": lm(4, 24, 20),
        "greedy_m3_div16.npz": greedy(3, 25, 16),
        "gauss_newton_m3.npz": nls(O.gauss_newton, 3, 26, 200),
        "gradient_descent_m2.npz": nls(O.gradient_descent, 2, 27, 10000),
        "apo_m3.npz": apo(3, 28, 0.05),
        "long_m4.npz": long_(4, 29),
        "horn_m4.npz": horn(4, 30, 3),
    }
    for name, arrays in fixtures.items():
        np.savez(os.path.join(HERE, name), **arrays)
        print(name, {k: np.asarray(v).shape for k, v in arrays.items()})


if __name__ == "__main__":
    main()
