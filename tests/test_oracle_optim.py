"""Pins the oracle's optimizers (oracle/afc_oracle_body.h) with closed forms
and identities rather than restated formulas.

Reference: acoustic-field-optimizer/src/multiple_foci/linear_synthesis/
{gs,gs_pat,naive}.rs, nls/{lm,macros}.rs, multiple_foci/macros.rs."""
import numpy as np
import pytest


def small_problem(oracle, n_tiles=1, m=1, seed=0):
    rng = np.random.default_rng(seed)
    pos = oracle.array_249()
    src = oracle.t4010a1_sources(pos)
    n = pos.shape[0]
    k = np.full(n, oracle.WAVENUM_40K, np.float32)
    a = np.full(n, np.float32(oracle.attenuation_coef(40e3, 30, 1, 1, oracle.TEMP_340)), np.float32)
    foci = np.stack([rng.uniform(40, 130, m), rng.uniform(30, 110, m), rng.uniform(120, 200, m)], 1).astype(np.float32)
    amps = rng.uniform(0.5, 1.0, m).astype(np.float32)
    return src, k, a, foci, amps


def with_solution(src, amp, phase):
    s = src.copy()
    s[:, 6] = amp
    s[:, 7] = phase
    return s


def focal(oracle, src, k, a, foci):
    return oracle.field(1, src, k, a, foci, "f64")


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_gspat_single_focus_closed_form(oracle, prec):
    """M = 1: R = N, the final p is p0 / N and q_j = conj(G_j) / |G_j|^2 p0 / N,
    so amp_j = p0 / (N |G_j|), phase_j = -arg G_j and the focal field is p0."""
    src, k, a, foci, amps = small_problem(oracle, m=1)
    amp, ph = oracle.gspat(1, src, k, a, foci, amps, prec)
    G = oracle.propagation_matrix(1, with_solution(src, 1.0, 0.0), k, a, foci, "f64")[0]
    n = src.shape[0]
    tol = 2e-5 if prec == "f32" else 1e-12
    assert np.allclose(amp, amps[0] / (n * np.abs(G)), rtol=tol)
    assert np.allclose(np.angle(np.exp(1j * (ph + np.angle(G)))), 0, atol=tol * 10)
    p = focal(oracle, with_solution(src, amp, ph), k, a, foci)[0]
    assert abs(p) == pytest.approx(amps[0], rel=max(tol * 10, 1e-6))  # solution stored as f32


def test_gs_keeps_source_amplitudes(oracle):
    src, k, a, foci, amps = small_problem(oracle, m=5, seed=1)
    amp, _ = oracle.gs(1, src, k, a, foci, amps, "f64")
    assert np.allclose(amp, 1.0, rtol=1e-12)


def gs_problem(oracle, m, seed):
    """small_problem with non-uniform source amplitudes, so that q0 (the current
    amplitudes, gs.rs:48-54) is not the constant 1."""
    src, k, a, foci, amps = small_problem(oracle, m=m, seed=seed)
    src[:, 6] = np.random.default_rng(seed + 100).uniform(0.5, 1.0, src.shape[0]).astype(np.float32)
    # G as gs.rs builds it: the current amplitudes, phases zeroed (macros.rs:17-53)
    G = oracle.propagation_matrix(1, with_solution(src, src[:, 6], 0.0), k, a, foci, "f64")
    return src, k, a, foci, amps, G


def gs_single_focus_pin(amp, ph, q0, g, tol):
    """The M = 1 closed form of gs.rs:56-72: p is a scalar, xi_j = conj(G_j) p, so
    |q_j| = q0_j and arg q_j = -arg G_j + c for one constant c; every term then
    arrives at the focus in phase."""
    c = np.angle(np.exp(1j * (ph + np.angle(g))))
    spread = np.max(np.abs(np.angle(np.exp(1j * (c - c[0])))))
    return bool(np.allclose(amp, q0, rtol=tol) and spread <= 10 * tol)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_gs_single_focus_closed_form(oracle, prec):
    src, k, a, foci, amps, G = gs_problem(oracle, 1, 5)
    amp, ph = oracle.gs(1, src, k, a, foci, amps, prec)
    q0 = src[:, 6].astype(np.float64)
    # f32: G's phase k r (r ~ 200 mm) carries ~1e-5 rad of rounding
    tol = 1e-5 if prec == "f32" else 1e-12
    assert gs_single_focus_pin(amp, ph, q0, G[0], tol)
    # |p(f)| = sum_j q0_j |G_j| with |G_j| the unit-amplitude term's modulus
    g_unit = oracle.propagation_matrix(1, with_solution(src, 1.0, 0.0), k, a, foci, "f64")[0]
    p = focal(oracle, with_solution(src, amp, ph), k, a, foci)[0]
    assert abs(p) == pytest.approx(float(np.sum(q0 * np.abs(g_unit))), rel=max(tol, 1e-6))
    # the pin rejects what two plausible mistakes would produce: xi without the
    # conjugate (arg q_j = +arg G_j + c) and q normalised without q0 (|q_j| = 1)
    assert not gs_single_focus_pin(amp, np.angle(G[0]) + 0.3, q0, G[0], 1e-3)
    assert not gs_single_focus_pin(np.ones_like(amp), ph, q0, G[0], 1e-3)


def gs_ascent_objective(G, p0, q):
    """max over |p_i| = p0_i of Re <p, G q>, i.e. sum_i p0_i |(G q)_i|."""
    return np.sum(p0 * np.abs(G @ q))


@pytest.mark.parametrize("m,seed", [(3, 6), (12, 7), (30, 8)])
def test_gs_is_block_coordinate_ascent(oracle, m, seed):
    """GS maximises f(p, q) = Re <p, G q> over |p_i| = p0_i and |q_j| = q0_j by
    blocks: p = gamma / |gamma| p0 is the best p for the current q, and
    q = xi / |xi| q0 with xi = G^H p the best q for that p (gs.rs:56-69). Hence
    max_p f(p, q_r) never decreases with the iteration count r. A conjugate
    dropped from xi, or a transposed G, breaks the ascent."""
    src, k, a, foci, amps, G = gs_problem(oracle, m, seed)
    p0 = amps.astype(np.float64)
    f = []
    for r in range(0, 21):
        amp, ph = oracle.gs(1, src, k, a, foci, amps, "f64", repeat=r)
        f.append(gs_ascent_objective(G, p0, amp * np.exp(1j * ph)))
    assert np.all(np.diff(f) >= -1e-12 * f[-1]), np.diff(f)
    assert f[-1] > 2 * f[0]
    # a planted mistake in a CPU restatement fails the same check
    q0 = src[:, 6].astype(np.float64)
    q, bad = q0.astype(np.complex128), []
    for _ in range(20):
        g = G @ q
        p = g / np.abs(g) * p0
        xi = G.T @ p  # conjugate dropped
        q = xi / np.abs(xi) * q0
        bad.append(gs_ascent_objective(G, p0, q))
    assert np.min(np.diff(bad)) < 0


def test_naive_single_focus(oracle):
    src, k, a, foci, amps = small_problem(oracle, m=1, seed=2)
    amp, ph = oracle.naive(1, src, k, a, foci, amps, "f64")
    G = oracle.propagation_matrix(1, with_solution(src, 1.0, 0.0), k, a, foci, "f64")[0]
    q = np.conj(G) * amps[0]
    assert np.allclose(amp, np.abs(q), rtol=1e-12) and np.allclose(ph, np.angle(q), atol=1e-12)


def test_gspat_f32_close_to_f64(oracle):
    src, k, a, foci, amps = small_problem(oracle, m=8, seed=3)
    a32, p32 = oracle.gspat(1, src, k, a, foci, amps, "f32")
    a64, p64 = oracle.gspat(1, src, k, a, foci, amps, "f64")
    f32 = np.abs(focal(oracle, with_solution(src, a32, p32), k, a, foci))
    f64 = np.abs(focal(oracle, with_solution(src, a64, p64), k, a, foci))
    assert np.max(np.abs(f32 - f64) / f64) < 1e-4


def test_gspat_reaches_target_ratios(oracle):
    """GS-PAT's closing normalisation (gamma / |gamma|^2 p0^2) puts the focal
    amplitudes close to the targets themselves (GS keeps |q| = q0 and cannot)."""
    src, k, a, foci, amps = small_problem(oracle, m=4, seed=4)
    am, ph = oracle.gspat(1, src, k, a, foci, amps, "f64")
    f = np.abs(focal(oracle, with_solution(src, am, ph), k, a, foci))
    ratio = f / amps
    assert np.all(np.abs(ratio - 1) < 0.1), ratio


def test_lm_gradient_is_half_dF(oracle):
    """g of calc_JtJ_Jtf is dF/dx / 2 with F = Re(t^H BhB t) (pins the sign of T = e^{-ix})."""
    src, k, a, foci, amps = small_problem(oracle, m=3, seed=5)
    src = src[:40]
    k, a = k[:40], a[:40]
    rng = np.random.default_rng(9)
    x = rng.uniform(-np.pi, np.pi, 43)
    F0, g = oracle.lm_eval(1, src, k, a, foci, amps, x)
    h = 1e-6
    for idx in (0, 7, 39, 40, 42):
        e = np.zeros_like(x)
        e[idx] = h
        Fp, _ = oracle.lm_eval(1, src, k, a, foci, amps, x + e)
        Fm, _ = oracle.lm_eval(1, src, k, a, foci, amps, x - e)
        assert (Fp - Fm) / (2 * h) == pytest.approx(2 * g[idx], rel=1e-5, abs=1e-9)


def test_lm_single_focus_reaches_target(oracle):
    """M = 1: F = |sum_j G_j e^{i x_j} - a e^{i x_N}|^2 has minimum 0 whenever
    sum_j |G_j| >= a, so LM ends with |p(f)| = a."""
    src, k, a, foci, amps = small_problem(oracle, m=1, seed=6)
    amps = np.ones(1, np.float32)
    ph, x, it = oracle.lm(1, src, k, a, foci, amps, "f64", k_max=200)
    F, _ = oracle.lm_eval(1, src, k, a, foci, amps, x)
    f = np.abs(focal(oracle, with_solution(src, 1.0, ph), k, a, foci))
    assert it < 200 and F < 1e-9 and np.allclose(f, 1.0, rtol=1e-5), (f, it, F)


def test_lm_decreases_objective(oracle):
    src, k, a, foci, amps = small_problem(oracle, m=2, seed=6)
    amps = np.ones(2, np.float32)
    F0, _ = oracle.lm_eval(1, src, k, a, foci, amps, np.zeros(src.shape[0] + 2))
    _, x, _ = oracle.lm(1, src, k, a, foci, amps, "f64", k_max=50)
    F, _ = oracle.lm_eval(1, src, k, a, foci, amps, x)
    assert F < 0.2 * F0


def test_greedy_single_unreachable_focus_aligns_all_sources(oracle):
    """M = 1 with a target no phase choice reaches: each step maximises |cache|,
    so every term lands within pi / div of the running sum's direction and the
    focal amplitude is sum_j |G_j| cos(pi / div) at least (combination/greedy.rs)."""
    src, k, a, _, _ = small_problem(oracle, m=1)
    foci = np.array([[90, 70, 150]], np.float32)
    G = oracle.propagation_matrix(1, src, k, a, foci, "f64")
    total = np.abs(G).sum()
    for div in (16, 4096):
        ph = oracle.greedy(1, src, k, a, foci, np.array([1e3], np.float32), div, "f64")
        p = np.abs(focal(oracle, with_solution(src, src[:, 6], ph), k, a, foci))[0]
        assert total * np.cos(np.pi / div) - 1e-9 <= p <= total * (1 + 1e-12)


def test_greedy_phases_are_candidates(oracle):
    """Every phase is arg(e^{i 2 pi idx / div}) for some idx, and f32 agrees
    with f64 on which candidate each source takes, relative to the first
    source's choice (a tie on the empty cache, so a global rotation)."""
    src, k, a, foci, amps = small_problem(oracle, m=3, seed=5)
    div = 8
    ph32 = oracle.greedy(1, src, k, a, foci, amps, div, "f32")
    ph64 = oracle.greedy(1, src, k, a, foci, amps, div, "f64")
    idx32 = np.round(np.mod(ph32, 2 * np.pi) / (2 * np.pi / div)).astype(int) % div
    idx64 = np.round(np.mod(ph64, 2 * np.pi) / (2 * np.pi / div)).astype(int) % div
    assert np.allclose(np.angle(np.exp(1j * (ph64 - 2 * np.pi * idx64 / div))), 0, atol=1e-12)
    assert np.mean((idx32 - idx32[0]) % div == (idx64 - idx64[0]) % div) > 0.95


def test_greedy_lowers_objective(oracle):
    """The greedy objective sum_j |a_j - |p_j|| at the end is below that of
    the all-zero phases."""
    src, k, a, foci, amps = small_problem(oracle, m=4, seed=6)
    ph = oracle.greedy(1, src, k, a, foci, amps, 16, "f64")
    p1 = np.abs(focal(oracle, with_solution(src, src[:, 6], ph), k, a, foci))
    p0 = np.abs(focal(oracle, with_solution(src, src[:, 6], np.zeros(len(src))), k, a, foci))
    assert np.sum(np.abs(amps - p1)) < 0.5 * np.sum(np.abs(amps - p0))


def _bhb_numpy(oracle, src, k, a, foci, amps):
    """B^H B with B = [G | -diag(amps)] (nls/macros.rs:17-31) from the pinned G."""
    G = oracle.propagation_matrix(1, src, k, a, foci, "f64")  # M x N
    B = np.concatenate([G, -np.diag(amps.astype(np.float64))], 1)
    return B.conj().T @ B


def test_gauss_newton_step_is_truncated_pseudo_inverse(oracle):
    """One Gauss-Newton step from x = 0 equals -pinv(A, 1e-3) g with A =
    Re(B^H B) and the pseudo-inverse from LAPACK's eigh (independent of the
    oracle's Jacobi solver), g from the pinned gradient."""
    src, k, a, foci, amps = small_problem(oracle, m=3, seed=9)
    _, x1, it = oracle.gauss_newton(1, src, k, a, foci, amps, "f64", k_max=1)
    assert it == 1
    A = _bhb_numpy(oracle, src, k, a, foci, amps).real
    _, g = oracle.lm_eval(1, src, k, a, foci, amps, np.zeros(len(src) + 3))
    lam, V = np.linalg.eigh(A)
    keep = np.abs(lam) > 1e-3
    h = -(V[:, keep] @ ((V[:, keep].T @ g) / lam[keep]))
    assert np.allclose(x1, h, rtol=1e-8, atol=1e-10 * np.max(np.abs(h)))


def test_gradient_descent_first_step_is_scaled_gradient(oracle):
    """x_1 = -0.1 g(0) = -0.05 dF/dx(0), dF/dx by central differences of F."""
    src, k, a, foci, amps = small_problem(oracle, m=2, seed=10)
    n = len(src) + 2
    _, x1, it = oracle.gradient_descent(1, src, k, a, foci, amps, "f64", k_max=1)
    assert it == 1
    x0 = np.zeros(n)
    dF = np.empty(n)
    for i in range(0, n, 17):
        e = np.zeros(n)
        e[i] = 1e-6
        dF[i] = (oracle.lm_eval(1, src, k, a, foci, amps, x0 + e)[0] - oracle.lm_eval(1, src, k, a, foci, amps,
                                                                                      x0 - e)[0]) / 2e-6
    idx = np.arange(0, n, 17)
    assert np.allclose(x1[idx], -0.05 * dF[idx], rtol=1e-5, atol=1e-9)


def test_gradient_descent_and_gauss_newton_lower_F(oracle):
    src, k, a, foci, amps = small_problem(oracle, m=3, seed=11)
    n = len(src) + 3
    F0 = oracle.lm_eval(1, src, k, a, foci, amps, np.zeros(n))[0]
    prev = F0
    for kk in (10, 100, 1000):
        _, x, _ = oracle.gradient_descent(1, src, k, a, foci, amps, "f64", k_max=kk)
        F = oracle.lm_eval(1, src, k, a, foci, amps, x)[0]
        assert F < prev
        prev = F
    # Gauss-Newton stops early once the pinv cut-off (1e-3) removes the
    # remaining descent directions (F / F0 = 0.026 here)
    _, x, _ = oracle.gauss_newton(1, src, k, a, foci, amps, "f64")
    assert oracle.lm_eval(1, src, k, a, foci, amps, x)[0] < 0.1 * F0


def test_apo_start_is_regularised_least_squares(oracle):
    """With no iteration APO returns q0 = (G^H G + lambda I)^{-1} G^H p
    (apo.rs:116-119), from numpy's LAPACK solve here, scaled by camax =
    max_j |Re q_j| + |Im q_j| (nalgebra's camax)."""
    src, k, a, foci, amps = small_problem(oracle, m=3, seed=12)
    lam = 0.05
    amp, ph, it = oracle.apo(1, src, k, a, foci, amps, lam, "f64", k_max=0)
    assert it == 0
    G = oracle.propagation_matrix(1, src, k, a, foci, "f64")
    q0 = np.linalg.solve(G.conj().T @ G + lam * np.eye(G.shape[1]), G.conj().T @ amps.astype(np.float64))
    cm = np.max(np.abs(q0.real) + np.abs(q0.imag))
    assert np.allclose(amp, src[:, 6] * np.abs(q0) / cm, rtol=1e-7, atol=1e-12)
    assert np.allclose(np.angle(np.exp(1j * (ph - np.angle(q0)))), 0, atol=1e-7)


def test_apo_single_focus_stays_matched(oracle):
    """M = 1: q0, the gradient and every update stay proportional to conj(g),
    so the result focuses: phase_j = -arg G_j and amp_j proportional to |G_j|."""
    src, k, a, _, _ = small_problem(oracle, m=1)
    foci = np.array([[90, 70, 150]], np.float32)
    amp, ph, it = oracle.apo(1, src, k, a, foci, np.array([0.3], np.float32), 0.01, "f64")
    G = oracle.propagation_matrix(1, src, k, a, foci, "f64")[0]
    assert np.allclose(np.angle(np.exp(1j * (ph + np.angle(G)))), 0, atol=1e-6)
    r = amp / np.abs(G)
    assert np.max(r) / np.min(r) - 1 < 1e-6


def test_bessel_phases_axial_closed_form(oracle):
    """dir = +z: the axis v = (0, 0, 0) leaves r unturned, so dist = sin(theta)
    rho - cos(theta) (z_j - z_p) with rho the distance from the beam axis, and
    phi_j = -2 pi ((dist mod lambda) / lambda) (C fmod: the sign of dist)."""
    pos = oracle.array_249()
    src = oracle.t4010a1_sources(pos)
    c = 340e3
    point, theta = np.array([90.0, 70.0, -20.0]), 0.3
    ph = oracle.bessel_phases(1, src, c, point, (0.0, 0.0, 1.0), theta, "f64")
    d = src[:, 0:3].astype(np.float64) - point
    dist = np.sin(theta) * np.hypot(d[:, 0], d[:, 1]) - np.cos(theta) * d[:, 2]
    lam = c / 40e3
    want = -2 * np.pi * (np.fmod(dist, lam) / lam)
    assert np.allclose(ph, want, rtol=0, atol=1e-12)
    # the f32 restatement agrees to f32 rounding
    ph32 = oracle.bessel_phases(1, src, c, point, (0.0, 0.0, 1.0), theta, "f32")
    assert np.max(np.abs(ph32 - want)) < 2e-5
