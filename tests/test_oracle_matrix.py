"""Pins the oracle's matrix-method optimizers (Long, Horn) and the Hermitian
eigen-decomposition they share with LAPACK (numpy) routines, closed forms
and invariants.

Reference: acoustic-field-optimizer/src/multiple_foci/matrix/long.rs:46-106,
matrix/horn.rs:58-129. The eigen-decomposition convention is the oracle's
own: eigenvalues ascending, each eigenvector's largest-modulus entry real
positive (nalgebra's SymmetricEigen, which the reference calls, fixes no
documented order or phase)."""
import numpy as np
import pytest

from test_oracle_optim import small_problem


def normalise_columns(V):
    """Scale each column so that its largest-modulus entry (the first on
    ties) is real positive."""
    V = V.copy()
    for k in range(V.shape[1]):
        j = int(np.argmax(np.abs(V[:, k]) ** 2))
        V[:, k] *= np.conj(V[j, k]) / abs(V[j, k])
    return V


def random_hermitian(rng, n):
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return A + A.conj().T


@pytest.mark.parametrize("n", [1, 2, 3, 7, 16, 32])
def test_herm_eig_matches_lapack(oracle, n):
    rng = np.random.default_rng(n)
    A = random_hermitian(rng, n)
    ev, V = oracle.herm_eig(A)
    w, U = np.linalg.eigh(A)
    scale = np.max(np.abs(w))
    assert np.all(np.diff(ev) >= 0)
    assert np.allclose(ev, w, atol=1e-12 * scale)
    assert np.allclose(V.conj().T @ V, np.eye(n), atol=1e-12)
    assert np.allclose(A @ V, V * ev, atol=1e-11 * scale)
    # the phase convention: largest entry real positive; the same vectors as
    # LAPACK's (random spectra are simple) under that convention
    for k in range(n):
        j = int(np.argmax(np.abs(V[:, k])))
        assert V[j, k].imag == 0 and V[j, k].real > 0
    assert np.allclose(V, normalise_columns(U), atol=1e-10)


def test_herm_eig_diagonal_and_ties(oracle):
    """A diagonal input is already converged: the eigenvalues are sorted
    (stably for ties) and the vectors are unit vectors."""
    A = np.diag([3.0, -1.0, 3.0, 0.5]).astype(complex)
    ev, V = oracle.herm_eig(A)
    assert np.array_equal(ev, [-1.0, 0.5, 3.0, 3.0])
    assert np.array_equal(np.abs(V), np.eye(4)[:, [1, 3, 0, 2]])


def _long_numpy(oracle, src, k, a, foci, amps):
    """long.rs:46-106 through LAPACK: eigh of the Hermitian matrix read from
    R's lower triangle (UPLO='L', as SymmetricEigen reads it) and a dense
    solve of the normal equations g^H g q = g^H f."""
    G = oracle.propagation_matrix(1, src, k, a, foci, "f64")
    m, n = G.shape
    am = amps.astype(np.float64)
    den = G.sum(axis=1)
    X = (am[:, None] * G.conj() / den[:, None]).T
    R = G @ X
    w, U = np.linalg.eigh(R, UPLO="L")
    U = normalise_columns(U)
    e_arg = np.angle(U[int(np.argmax(w)), :])
    sigma = np.sqrt((np.abs(G) * am[:, None]).sum(axis=0) / m)
    g = np.vstack([G, np.diag(sigma)])
    f = np.concatenate([am * np.exp(1j * e_arg), np.zeros(n)])
    q = np.linalg.solve(g.conj().T @ g, g.conj().T @ f)
    cm = np.max(np.abs(q.real) + np.abs(q.imag))
    return src[:, 6] * np.abs(q) / cm, np.angle(q)


@pytest.mark.parametrize("m,seed", [(2, 20), (5, 21), (12, 22)])
def test_long_matches_lapack_pipeline(oracle, m, seed):
    src, k, a, foci, amps = small_problem(oracle, m=m, seed=seed)
    amp, ph = oracle.long_(1, src, k, a, foci, amps)
    ra, rp = _long_numpy(oracle, src, k, a, foci, amps)
    assert np.allclose(amp, ra, rtol=1e-8, atol=1e-12)
    assert np.allclose(np.angle(np.exp(1j * (ph - rp))), 0, atol=1e-8)


def test_long_single_focus_closed_form(oracle):
    """M = 1: the eigenvector is [1], so f = [a]; sigma_j^2 = |G_j| a and
    q = (G^H G + diag sigma^2)^{-1} G^H a = conj(G_j) / sigma_j^2 * a / (1 + sum_l |G_l|^2 / sigma_l^2):
    |q_j| = 1 / (1 + sum_l |G_l| / a) for every j, phase_j = -arg G_j."""
    src, k, a, foci, amps = small_problem(oracle, m=1, seed=23)
    amp, ph = oracle.long_(1, src, k, a, foci, amps)
    G = oracle.propagation_matrix(1, src, k, a, foci, "f64")[0]
    q = np.conj(G) / np.abs(G)  # common magnitude drops out of amp / camax
    cm = np.max(np.abs(q.real) + np.abs(q.imag))
    assert np.allclose(amp, np.abs(q) / cm, rtol=1e-10)
    assert np.allclose(np.angle(np.exp(1j * (ph + np.angle(G)))), 0, atol=1e-10)


def test_horn_single_focus_closed_form(oracle):
    """M = 1: no update changes X = [1] (the 0 x 0 product gives gamma = 0),
    u = [1] and q = G^H (|G|^2 + a^2)^{-1} p: phase_j = -arg G_j, amp_j
    proportional to |G_j|."""
    src, k, a, foci, amps = small_problem(oracle, m=1, seed=24)
    amp, ph = oracle.horn(1, src, k, a, foci, amps, seed=5)
    G = oracle.propagation_matrix(1, src, k, a, foci, "f64")[0]
    q = np.conj(G)
    cm = np.max(np.abs(q.real) + np.abs(q.imag))
    assert np.allclose(amp, np.abs(q) / cm, rtol=1e-10)
    assert np.allclose(np.angle(np.exp(1j * (ph + np.angle(G)))), 0, atol=1e-10)


@pytest.mark.parametrize("alpha", [1e-5, 10.0])
def test_horn_pinv_is_regularised_svd_inverse(oracle, alpha):
    """repeat = 0 leaves X = I, whose eigenvalues tie at 1: the first vector
    e_0 is taken, so q = a_0 G_pinv[:, 0] with G_pinv = V diag(s / (s^2 +
    alpha^2)) U^H from LAPACK's SVD (horn.rs:58-68)."""
    src, k, a, foci, amps = small_problem(oracle, m=6, seed=25)
    amp, ph = oracle.horn(1, src, k, a, foci, amps, repeat=0, alpha=alpha)
    G = oracle.propagation_matrix(1, src, k, a, foci, "f64")
    U, s, Vh = np.linalg.svd(G, full_matrices=False)
    Gp = Vh.conj().T @ np.diag(s / (s * s + alpha * alpha)) @ U.conj().T
    q = Gp[:, 0] * amps[0]
    cm = np.max(np.abs(q.real) + np.abs(q.imag))
    assert np.allclose(amp, np.abs(q) / cm, rtol=1e-9, atol=1e-12)
    assert np.allclose(np.angle(np.exp(1j * (ph - np.angle(q)))), 0, atol=1e-9)


def splitmix_rows(seed, m, count):
    """The oracle's row sequence: splitmix64 from the seed, the top 24 bits
    as an f32 in [0, 1), scaled by m in f32 and truncated (clamped to m - 1)."""
    mask = (1 << 64) - 1
    s = seed
    out = []
    for _ in range(count):
        s = (s + 0x9E3779B97F4A7C15) & mask
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        z ^= z >> 31
        r = np.float32(z >> 40) * np.float32(1.0 / 16777216.0)
        out.append(min(int(np.float32(m) * r), m - 1))
    return out


def test_horn_sdp_matches_numpy_restatement(oracle):
    """The SDP row updates (horn.rs:86-119) with numpy's row / column
    deletion, MM from the SVD pseudo-inverse (alpha = 1, where I - G G_pinv
    is well conditioned in f64), the top eigenvector from LAPACK."""
    m, seed, repeat, lam, alpha = 5, 7, 20, 0.8, 1.0
    src, k, a, foci, amps = small_problem(oracle, m=m, seed=26)
    amp, ph = oracle.horn(1, src, k, a, foci, amps, seed=seed, repeat=repeat, lam=lam, alpha=alpha)
    G = oracle.propagation_matrix(1, src, k, a, foci, "f64")
    U, s, Vh = np.linalg.svd(G, full_matrices=False)
    Gp = Vh.conj().T @ np.diag(s / (s * s + alpha * alpha)) @ U.conj().T
    P = np.diag(amps.astype(np.float64))
    MM = P @ (np.eye(m) - G @ Gp) @ P
    X = np.eye(m, dtype=complex)
    for ii in splitmix_rows(seed, m, m * repeat):
        Xc = np.delete(np.delete(X, ii, 0), ii, 1)
        MMc = np.delete(MM[:, ii], ii)
        Xb = Xc @ MMc
        gamma = np.vdot(Xb, MMc)
        rest = [r for r in range(m) if r != ii]
        if gamma.real > 0:
            Xb = -np.sqrt(lam / gamma.real) * Xb
            X[rest, ii] = Xb
            X[ii, rest] = Xb.conj()
        else:
            X[rest, ii] = 0
            X[ii, rest] = 0
    assert np.allclose(np.diag(X), 1) and np.allclose(X, X.conj().T)
    w, V = np.linalg.eigh(X)
    u = normalise_columns(V)[:, int(np.argmax(w))]
    q = Gp @ P @ u
    cm = np.max(np.abs(q.real) + np.abs(q.imag))
    assert np.allclose(amp, np.abs(q) / cm, rtol=1e-8, atol=1e-12)
    assert np.allclose(np.angle(np.exp(1j * (ph - np.angle(q)))), 0, atol=1e-8)


def test_horn_rows_cover_all_indices():
    rows = splitmix_rows(3, 7, 7000)
    counts = np.bincount(rows, minlength=7)
    assert counts.min() > 850 and counts.max() < 1150
