"""GPU parity of Gauss-Newton (nls/gauss_newton.rs:35-89) and gradient
descent (nls/gradient_descent.rs:34-61) against the oracle's FP64
restatements (dense B^H B, Jacobi pseudo-inverse with the 1e-3 cut-off).

The GPU solves both in the low-rank form (W = [Re C; Im C], K = W W^T): the
Gauss-Newton step as -W^T pinv(K, 1e-3) v, the gradient as sum_i Im(conj(C_il) s_i).
Tolerances: iteration counts equal (gradient descent within 2 steps of its
stopping test), phases within 1e-6 rad, and the objective F within 1e-6
relative of the oracle's run; through the C ABI the source amplitudes stay
unchanged.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


def setup(oracle, n_src=None):
    pos = oracle.array_249()
    if n_src:
        pos = pos[:n_src]
    src = oracle.t4010a1_sources(pos)
    n = pos.shape[0]
    k = np.full(n, oracle.WAVENUM_40K, np.float32)
    a = np.full(n, np.float32(oracle.attenuation_coef(40e3, 30, 1, 1, oracle.TEMP_340)), np.float32)
    return src, k, a


def problem(rng, m):
    foci = np.stack([rng.uniform(40, 130, m), rng.uniform(30, 110, m), rng.uniform(120, 200, m)], 1).astype(np.float32)
    amps = rng.uniform(0.5, 1.0, m).astype(np.float32)
    return foci, amps


def run_nls(torch, src, k, a, foci, amps, offs, algo, k_max):
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
    D.nl_solve(ctx, src_d.data_ptr(), packed.shape[0], n, 1, foci_d.data_ptr(), amps_d.data_ptr(), offs, algo, k_max,
               out.data_ptr(), iters.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    ctx.close()
    return out.cpu().numpy(), iters.cpu().numpy()


def wrapdiff(a, b):
    return np.abs(np.angle(np.exp(1j * (np.asarray(a, np.float64) - np.asarray(b, np.float64)))))


@pytest.mark.parametrize("m,seed", [(1, 1), (3, 2), (6, 3)])
def test_gauss_newton_vs_oracle(afc, oracle, torch_cuda, m, seed):
    log(f"case: Gauss-Newton M={m} seed={seed}")
    from afc import device as D
    import torch
    rng = np.random.default_rng(seed)
    src, k, a = setup(oracle)
    foci, amps = problem(rng, m)
    ph, it = run_nls(torch, src, k, a, foci, amps, np.array([0, m], np.uint32), D.NLS_GAUSS_NEWTON, 200)
    ref, x_ref, it_ref = oracle.gauss_newton(1, src, k, a, foci, amps, "f64")
    d = wrapdiff(ph[0], ref).max()
    log(f"  iterations gpu {it[0]} oracle {it_ref}; max phase diff {d:.2e}")
    assert it[0] == it_ref
    assert d <= 1e-6


@pytest.mark.parametrize("m,seed", [(1, 4), (4, 5)])
def test_gradient_descent_vs_oracle(afc, oracle, torch_cuda, m, seed):
    log(f"case: gradient descent M={m} seed={seed}")
    from afc import device as D
    import torch
    rng = np.random.default_rng(seed)
    src, k, a = setup(oracle, 100)
    foci, amps = problem(rng, m)
    foci[:, 0] *= 0.5
    ph, it = run_nls(torch, src, k, a, foci, amps, np.array([0, m], np.uint32), D.NLS_GRADIENT_DESCENT, 10000)
    ref, x_ref, it_ref = oracle.gradient_descent(1, src, k, a, foci, amps, "f64")
    d = wrapdiff(ph[0], ref).max()
    log(f"  steps gpu {it[0]} oracle {it_ref}; max phase diff {d:.2e}")
    assert abs(int(it[0]) - it_ref) <= 2
    assert d <= 1e-6


@pytest.mark.parametrize("tiles,n_src,m,seed", [((2, 1), 300, 3, 11), ((1, 1), 200, 20, 12), ((2, 2), 600, 3, 13)])
def test_gradient_descent_sizes_vs_oracle(afc, oracle, torch_cuda, tiles, n_src, m, seed):
    """Gradient descent at the sizes that select each kernel: G in registers
    for M <= 16 with 256 < N <= 512 (two columns per thread) and for
    16 < M <= 32 with N <= 256, G in global memory for N > 512. 2000 steps."""
    log(f"case: gradient descent tiles={tiles} N={n_src} M={m} seed={seed}")
    from afc import device as D
    import torch
    rng = np.random.default_rng(seed)
    pos = oracle.tiled_array(*tiles)[:n_src]
    src = oracle.t4010a1_sources(pos)
    k = np.full(n_src, oracle.WAVENUM_40K, np.float32)
    a = np.full(n_src, np.float32(oracle.attenuation_coef(40e3, 30, 1, 1, oracle.TEMP_340)), np.float32)
    foci, amps = problem(rng, m)
    ph, it = run_nls(torch, src, k, a, foci, amps, np.array([0, m], np.uint32), D.NLS_GRADIENT_DESCENT, 2000)
    ref, x_ref, it_ref = oracle.gradient_descent(1, src, k, a, foci, amps, "f64", k_max=2000)
    d = wrapdiff(ph[0], ref).max()
    log(f"  steps gpu {it[0]} oracle {it_ref}; max phase diff {d:.2e}")
    assert abs(int(it[0]) - it_ref) <= 2
    assert d <= 1e-6


def test_nls_capi_sets_phases_only(afc, oracle, torch_cuda):
    log("case: AFO_GaussNewton / AFO_GradientDescent through the C ABI")
    rng = np.random.default_rng(6)
    src, k, a = setup(oracle)
    foci, amps = problem(rng, 2)
    for name, ref_fn in (("gauss_newton", oracle.gauss_newton), ("gradient_descent", oracle.gradient_descent)):
        s = afc.UniformSystem(oracle.TEMP_340)
        for row in src:
            s.add_wave_source(afc.T4010A1(row[0:3], row[3:6], row[6], row[7], row[8]))
        getattr(afc.Optimizer, name)(s, foci, amps)
        got = s.source_array().copy()
        assert np.array_equal(got[:, 6], src[:, 6])
        kk = np.zeros(len(src), np.float32)
        aa = np.zeros(len(src), np.float32)
        import ctypes
        afc.dll().AFH_GetSourceConstants(s.handle, kk.ctypes.data_as(ctypes.c_void_p), aa.ctypes.data_as(ctypes.c_void_p))
        ref, _, _ = ref_fn(1, src, kk, aa, foci, amps, "f64")
        d = wrapdiff(got[:, 7], ref).max()
        log(f"  {name}: max phase diff vs FP64 oracle {d:.2e}")
        assert d <= 1e-5


def test_nls_batch_equals_single(afc, oracle, torch_cuda):
    log("case: batched Gauss-Newton and gradient descent equal one-at-a-time")
    from afc import device as D
    import torch
    rng = np.random.default_rng(7)
    src, k, a = setup(oracle, 120)
    ms = [2, 5, 1]
    fs, am = zip(*[problem(rng, m) for m in ms])
    foci, amps = np.concatenate(fs), np.concatenate(am)
    offs = np.concatenate([[0], np.cumsum(ms)]).astype(np.uint32)
    for algo, kmax in ((D.NLS_GAUSS_NEWTON, 200), (D.NLS_GRADIENT_DESCENT, 3000)):
        ph, it = run_nls(torch, src, k, a, foci, amps, offs, algo, kmax)
        for b in range(len(ms)):
            one, it1 = run_nls(torch, src, k, a, foci[offs[b]:offs[b + 1]], amps[offs[b]:offs[b + 1]],
                               np.array([0, ms[b]], np.uint32), algo, kmax)
            assert np.array_equal(one[0], ph[b]) and it1[0] == it[b], (algo, b)


def run_apo(torch, src, k, a, foci, amps, offs, lam, k_max=200):
    from afc import device as D
    dev = torch.device("cuda:0")
    packed = D.pack_sources(1, src, k, a)
    src_d = torch.from_numpy(packed).to(dev)
    foci_d = torch.from_numpy(np.ascontiguousarray(foci, np.float32)).to(dev)
    amps_d = torch.from_numpy(np.ascontiguousarray(amps, np.float32)).to(dev)
    B = offs.shape[0] - 1
    n = src.shape[0]
    oa = torch.empty((B, n), dtype=torch.float32, device=dev)
    op = torch.empty((B, n), dtype=torch.float32, device=dev)
    iters = torch.zeros(B, dtype=torch.int32, device=dev)
    ctx = D.Context()
    D.apo_solve(ctx, src_d.data_ptr(), packed.shape[0], n, 1, foci_d.data_ptr(), amps_d.data_ptr(), offs, lam, k_max,
                oa.data_ptr(), op.data_ptr(), iters.data_ptr())
    torch.cuda.synchronize()
    ctx.close()
    return oa.cpu().numpy(), op.cpu().numpy(), iters.cpu().numpy()


@pytest.mark.parametrize("m,lam,seed", [(1, 0.01, 11), (3, 0.05, 12), (8, 1.0, 13)])
def test_apo_vs_oracle(afc, oracle, torch_cuda, m, lam, seed):
    """APO (nls/apo.rs:100-158) against the FP64 oracle: iteration counts
    equal, amplitudes within 1e-5 relative, phases within 1e-5 rad where the
    amplitude is above 1e-3 of the largest (arg of a vanishing q_j is noise)."""
    log(f"case: APO M={m} lambda={lam}")
    import torch
    rng = np.random.default_rng(seed)
    src, k, a = setup(oracle, 160)
    foci, amps = problem(rng, m)
    foci[:, 0] *= 0.6
    amps *= 0.2
    oa, op, it = run_apo(torch, src, k, a, foci, amps, np.array([0, m], np.uint32), lam)
    ra, rp, rit = oracle.apo(1, src, k, a, foci, amps, lam, "f64")
    big = ra > 1e-3 * ra.max()
    ea = float(np.max(np.abs(oa[0] - ra) / ra.max()))
    ep = float(wrapdiff(op[0][big], rp[big]).max())
    log(f"  iterations gpu {it[0]} oracle {rit}; amp err {ea:.2e}, phase err {ep:.2e}")
    assert it[0] == rit
    assert ea <= 1e-5 and ep <= 1e-5


def test_apo_capi_and_batch(afc, oracle, torch_cuda):
    log("case: AFO_APO through the C ABI; batched APO equals one at a time")
    import torch
    rng = np.random.default_rng(21)
    src, k, a = setup(oracle, 120)
    ms = [2, 1, 4]
    fs, am = zip(*[problem(rng, m) for m in ms])
    foci, amps = np.concatenate(fs), np.concatenate(am) * 0.2
    offs = np.concatenate([[0], np.cumsum(ms)]).astype(np.uint32)
    oa, op, it = run_apo(torch, src, k, a, foci, amps, offs, 0.1)
    for b in range(len(ms)):
        a1, p1, i1 = run_apo(torch, src, k, a, foci[offs[b]:offs[b + 1]], amps[offs[b]:offs[b + 1]],
                             np.array([0, ms[b]], np.uint32), 0.1)
        assert np.array_equal(a1[0], oa[b]) and np.array_equal(p1[0], op[b]) and i1[0] == it[b]
    s = afc.UniformSystem(oracle.TEMP_340)
    for row in src:
        s.add_wave_source(afc.T4010A1(row[0:3], row[3:6], row[6], row[7], row[8]))
    f0, a0 = foci[:2], amps[:2]
    afc.Optimizer.apo(s, f0, a0, 0.1)
    got = s.source_array().copy()
    import ctypes
    kk = np.zeros(len(src), np.float32)
    aa = np.zeros(len(src), np.float32)
    afc.dll().AFH_GetSourceConstants(s.handle, kk.ctypes.data_as(ctypes.c_void_p), aa.ctypes.data_as(ctypes.c_void_p))
    ra, rp, _ = oracle.apo(1, src, kk, aa, f0, a0, 0.1, "f64")
    big = ra > 1e-3 * ra.max()
    assert np.max(np.abs(got[:, 6] - ra) / ra.max()) <= 1e-5
    assert wrapdiff(got[big, 7], rp[big]).max() <= 1e-5
