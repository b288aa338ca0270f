"""GPU parity of Levenberg-Marquardt (nls/lm.rs:40-105) against the oracle's
FP64 restatement (dense B^H B, Householder QR, the reference's quirks).

The reference iteration is chaotic: in the FP64 oracle itself, moving one
attenuation by one f32 ulp changes the phases by ~1e-7 after 20 iterations,
~1e-5 after 40, ~1e-3 after 80 and O(1) after 200. So:
  * trajectory parity is asserted over 20 iterations: iteration counts equal,
    phases within 1e-5 rad, focal amplitudes within 1e-6 relative;
  * a converged single focus (M = 1) must reach the target amplitude
    (|p| = a, the closed form) within 1e-5 after the full 200 iterations;
  * longer runs are compared as outcomes: the GPU's final objective is within
    a factor 2 of the oracle's run and the mean focal amplitude within 10 %.
The GPU solves the damped step in its low-rank form (FP64 Cholesky of
mu I + W W^T) and builds G in FP64."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


def setup(oracle, tiles=(1, 1)):
    pos = oracle.array_249() if tiles == (1, 1) else oracle.tiled_array(*tiles)
    src = oracle.t4010a1_sources(pos)
    n = pos.shape[0]
    k = np.full(n, oracle.WAVENUM_40K, np.float32)
    a = np.full(n, np.float32(oracle.attenuation_coef(40e3, 30, 1, 1, oracle.TEMP_340)), np.float32)
    return src, k, a


def focal(oracle, src, k, a, foci, phase):
    s = src.copy()
    s[:, 7] = phase
    return np.abs(oracle.field(1, s, k, a, foci, "f64"))


def wrap(d):
    return np.angle(np.exp(1j * np.asarray(d, np.float64)))


def run_lm(torch, oracle, src, k, a, foci, amps, offs, k_max):
    from afc import device as D
    dev = torch.device("cuda:0")
    packed = D.pack_sources(1, src, k, a)
    src_d = torch.from_numpy(packed).to(dev)
    foci_d = torch.from_numpy(np.ascontiguousarray(foci, np.float32)).to(dev)
    amps_d = torch.from_numpy(np.ascontiguousarray(amps, np.float32)).to(dev)
    B = offs.shape[0] - 1
    n = src.shape[0]
    out = torch.empty((B, n), dtype=torch.float32, device=dev)
    iters = torch.zeros(B, dtype=torch.int32, device=dev)
    ctx = D.Context()
    D.lm_solve(ctx, src_d.data_ptr(), packed.shape[0], n, 1, foci_d.data_ptr(), amps_d.data_ptr(), offs, k_max,
               out.data_ptr(), iters.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    ctx.close()
    return out.cpu().numpy(), iters.cpu().numpy()


def rand_foci(rng, m):
    return np.stack([rng.uniform(40, 130, m), rng.uniform(30, 110, m), rng.uniform(120, 200, m)],
                    1).astype(np.float32)


def test_afo_lm_single_focus_converges(afc, oracle, torch_cuda):
    log("case: AFO_LM, 249 array, M=1 (converges)")
    src, k, a = setup(oracle)
    foci = rand_foci(np.random.default_rng(6), 1)
    s = afc.UniformSystem(oracle.TEMP_340)
    for row in src:
        s.add_wave_source(afc.T4010A1(row[0:3], row[3:6], row[6], row[7], row[8]))
    afc.Optimizer.lm(s, foci, np.ones(1, np.float32))
    got = s.source_array()[:, 7].copy()
    f = focal(oracle, src, k, a, foci, got)
    log(f"  focal amplitude {f}")
    assert abs(f[0] - 1.0) <= 1e-5


@pytest.mark.parametrize("m,seed", [(2, 6), (5, 1), (16, 3)])
def test_lm_trajectory_20_iterations(afc, oracle, torch_cuda, m, seed):
    log(f"case: LM trajectory, 249 array, M={m}, k_max 20")
    src, k, a = setup(oracle)
    foci = rand_foci(np.random.default_rng(seed), m)
    amps = np.ones(m, np.float32)
    phases, iters = run_lm(torch_cuda, oracle, src, k, a, foci, amps, np.array([0, m], np.uint32), 20)
    ph, x, it = oracle.lm(1, src, k, a, foci, amps, "f64", k_max=20)
    err = np.max(np.abs(wrap(phases[0] - ph)))
    fg, fr = focal(oracle, src, k, a, foci, phases[0]), focal(oracle, src, k, a, foci, ph)
    log(f"  iterations gpu {iters[0]} oracle {it}; max phase diff {err:.2e}; "
        f"focal rel diff {np.max(np.abs(fg - fr) / fr):.2e}")
    assert int(iters[0]) == it
    assert err <= 1e-5
    assert np.allclose(fg, fr, rtol=1e-6)


def test_lm_996_array_trajectory(afc, oracle, torch_cuda):
    log("case: LM, 996 array (n = 1012), 16 foci, k_max 10")
    src, k, a = setup(oracle, (2, 2))
    foci, amps, offs = oracle.gspat_batch(1, m_range=(16, 16), seed=5)
    amps = np.ones_like(amps)
    phases, iters = run_lm(torch_cuda, oracle, src, k, a, foci, amps, offs, 10)
    ph, x, it = oracle.lm(1, src, k, a, foci, amps, "f64", k_max=10)
    assert int(iters[0]) == it
    assert np.max(np.abs(wrap(phases[0] - ph))) <= 1e-5


def test_lm_batched_50_iterations_outcome(afc, oracle, torch_cuda):
    """Config-4 shape on the 249 array: 16 foci per instance, 50 iterations,
    compared as outcomes (the iteration is chaotic beyond ~40 steps)."""
    log("case: batched LM, 249 array, 6 instances x 16 foci, k_max 50")
    src, k, a = setup(oracle)
    B = 6
    foci, amps, offs = oracle.gspat_batch(B, m_range=(16, 16), seed=11)
    amps = np.ones_like(amps)
    phases, iters = run_lm(torch_cuda, oracle, src, k, a, foci, amps, offs, 50)
    mg, mr = [], []
    for b in range(B):
        f = foci[offs[b]:offs[b + 1]]
        am = amps[offs[b]:offs[b + 1]]
        ph, x, it = oracle.lm(1, src, k, a, f, am, "f64", k_max=50)
        fg, fr = focal(oracle, src, k, a, f, phases[b]), focal(oracle, src, k, a, f, ph)
        Fg, Fr = np.sum((fg - am) ** 2), np.sum((fr - am) ** 2)
        log(f"  instance {b}: |p| mean gpu {fg.mean():.4f} oracle {fr.mean():.4f}; "
            f"sum (|p|-a)^2 gpu {Fg:.3e} oracle {Fr:.3e}")
        assert Fg <= 2 * Fr + 1e-6
        mg.append(fg.mean())
        mr.append(fr.mean())
    assert abs(np.mean(mg) - np.mean(mr)) <= 0.1 * np.mean(mr)


def test_lm_instance_isolation(afc, oracle, torch_cuda):
    src, k, a = setup(oracle)
    foci, amps, offs = oracle.gspat_batch(4, m_range=(3, 16), seed=2)
    amps = np.ones_like(amps)
    full, it_full = run_lm(torch_cuda, oracle, src, k, a, foci, amps, offs, 30)
    for b in (0, 3):
        lo, hi = offs[b], offs[b + 1]
        one, it_one = run_lm(torch_cuda, oracle, src, k, a, foci[lo:hi], amps[lo:hi], np.array([0, hi - lo], np.uint32),
                             30)
        assert np.array_equal(one[0], full[b]) and it_one[0] == it_full[b]


def ulp_sensitivity(oracle, src, k, a, foci, amps, k_max, prec="f64"):
    """How far the oracle's own focal amplitudes move when one source's
    attenuation changes by one f32 ulp (sources 7 and 100, up and down): the
    reference iteration's sensitivity to its inputs at this length (max
    relative change over the foci and the four perturbations)."""
    ph, _, _ = oracle.lm(1, src, k, a, foci, amps, prec, k_max=k_max)
    f1 = focal(oracle, src, k, a, foci, ph)
    worst = 0.0
    for j in (7, 100):
        for to in (1.0, -1.0):
            a2 = a.copy()
            a2[j] = np.nextafter(a2[j], np.float32(to))
            ph2, _, _ = oracle.lm(1, src, k, a2, foci, amps, prec, k_max=k_max)
            f2 = focal(oracle, src, k, a, foci, ph2)
            worst = max(worst, float(np.max(np.abs(f1 - f2) / f1)))
    return worst


@pytest.mark.parametrize("tiles,batch", [((1, 1), 4), ((2, 2), 1)])
def test_lm_config4_50_iterations(afc, oracle, torch_cuda, tiles, batch):
    """Config 4 at its stated length: 16 foci per instance (foci as the GS-PAT
    batch), target amplitude 1, k_max 50, on the 249 and 996 arrays. Per
    instance: the same iteration count as the FP64 oracle, and focal amplitudes
    within 1e-3 relative of the oracle's (lm.rs:40-105) -- or, on an instance
    where the reference iteration is chaotic by then, within the distance the
    oracle itself moves when one input changes by one f32 ulp. Both sides are
    FP64 from the same f32 inputs; they differ by FP64 rounding (the GPU's
    low-rank Cholesky step against the oracle's dense QR), which stays below
    1e-7 where the iteration is not chaotic."""
    src, k, a = setup(oracle, tiles)
    foci, amps, offs = oracle.gspat_batch(batch, m_range=(16, 16), seed=11)
    amps = np.ones_like(amps)
    log(f"case: LM config 4, N={src.shape[0]}, {batch} instance(s) x 16 foci, k_max 50")
    phases, iters = run_lm(torch_cuda, oracle, src, k, a, foci, amps, offs, 50)
    errs = []
    for b in range(batch):
        f = foci[offs[b]:offs[b + 1]]
        am = amps[offs[b]:offs[b + 1]]
        ph, x, it = oracle.lm(1, src, k, a, f, am, "f64", k_max=50)
        fg, fr = focal(oracle, src, k, a, f, phases[b]), focal(oracle, src, k, a, f, ph)
        e = float(np.max(np.abs(fg - fr) / fr))
        errs.append(e)
        gate = 1e-3
        if e > gate:
            gate = max(gate, ulp_sensitivity(oracle, src, k, a, f, am, 50))
        log(f"  instance {b}: iterations gpu {iters[b]} oracle {it}; focal-amplitude max rel err vs FP64 {e:.2e}; "
            f"max phase diff {np.max(np.abs(wrap(phases[b] - ph))):.2e}; gate {gate:.2e}")
        assert int(iters[b]) == it
        assert e <= gate
    # most instances stay on the oracle's trajectory
    assert np.median(errs) <= 1e-4
