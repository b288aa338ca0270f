"""GPU parity of the hologram solvers against the oracle.

Tolerances:
  * single-instance AFO_GSPAT / AFO_GS / AFO_Naive (FP32 on the GPU) vs the
    reference's FP32 restatement: focal amplitudes within 1e-3 relative of the
    FP64 oracle's, source phases within 1e-3 rad where amplitudes are not tiny;
  * batched GS-PAT in every precision mode: focal amplitudes within 1e-3
    relative of the FP64 oracle, instance by instance (M up to 128, foci over
    the whole 3984 array), and on all 1024 instances of the config-3 batch;
  * an instance gives bit-identical results alone and inside a batch.
Reference: acoustic-field-optimizer/src/multiple_foci/linear_synthesis/
{gs,gs_pat,naive}.rs, capi/src/optimizer.rs:80-127.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


def build_system(afc, src, temperature):
    s = afc.UniformSystem(temperature)
    for row in src:
        s.add_wave_source(afc.T4010A1(row[0:3], row[3:6], row[6], row[7], row[8]))
    return s


def setup_array(oracle, tiles=(1, 1)):
    pos = oracle.tiled_array(*tiles) if tiles != (1, 1) else oracle.array_249()
    src = oracle.t4010a1_sources(pos)
    n = pos.shape[0]
    k = np.full(n, oracle.WAVENUM_40K, np.float32)
    a = np.full(n, np.float32(oracle.attenuation_coef(40e3, 30, 1, 1, oracle.TEMP_340)), np.float32)
    return src, k, a


def with_solution(src, amp, phase):
    s = src.copy()
    s[:, 6] = amp
    s[:, 7] = phase
    return s


def focal_amps(oracle, src, k, a, foci, amp, phase):
    return np.abs(oracle.field(1, with_solution(src, amp, phase), k, a, foci, "f64"))


def random_foci(rng, m, lo=(40, 30, 120), hi=(130, 110, 200)):
    return np.stack([rng.uniform(lo[0], hi[0], m), rng.uniform(lo[1], hi[1], m), rng.uniform(lo[2], hi[2], m)],
                    1).astype(np.float32)


@pytest.mark.parametrize("method,m", [("gs_pat", 2), ("gs_pat", 9), ("gs", 3), ("naive", 4)])
def test_single_instance_capi(afc, oracle, torch_cuda, method, m):
    log(f"case: AFO_{method} single instance, M={m}")
    rng = np.random.default_rng(m)
    src, k, a = setup_array(oracle)
    foci = random_foci(rng, m)
    amps = rng.uniform(0.5, 1.0, m).astype(np.float32)
    s = build_system(afc, src, oracle.TEMP_340)
    getattr(afc.Optimizer, method)(s, foci, amps)
    got = s.source_array().copy()
    ref = {"gs_pat": oracle.gspat, "gs": oracle.gs, "naive": oracle.naive}[method]
    a64, p64 = ref(1, src, k, a, foci, amps, "f64")
    f_gpu = focal_amps(oracle, src, k, a, foci, got[:, 6], got[:, 7])
    f_ref = focal_amps(oracle, src, k, a, foci, a64, p64)
    err = np.max(np.abs(f_gpu - f_ref) / f_ref)
    log(f"  focal amplitude max rel err vs FP64 oracle: {err:.3e}")
    assert err <= 1e-3
    # the complex drive q_j, relative to the largest one
    q_gpu = got[:, 6].astype(np.float64) * np.exp(1j * got[:, 7].astype(np.float64))
    q_ref = a64 * np.exp(1j * p64)
    assert np.max(np.abs(q_gpu - q_ref)) <= 1e-3 * np.max(np.abs(q_ref))


def run_batch(afc, torch, oracle, src, k, a, foci, amps, offs, algo, precision, iters=100):
    from afc import device as D
    dev = torch.device("cuda:0")
    packed = D.pack_sources(1, src, k, a)
    src_d = torch.from_numpy(packed).to(dev)
    foci_d = torch.from_numpy(np.ascontiguousarray(foci, np.float32)).to(dev)
    amps_d = torch.from_numpy(np.ascontiguousarray(amps, np.float32)).to(dev)
    B = offs.shape[0] - 1
    n = src.shape[0]
    out_a = torch.empty((B, n), dtype=torch.float32, device=dev)
    out_p = torch.empty((B, n), dtype=torch.float32, device=dev)
    ctx = D.Context()
    D.holo_solve(ctx, src_d.data_ptr(), packed.shape[0], n, 1, foci_d.data_ptr(), amps_d.data_ptr(), offs, algo,
                 precision, iters, out_a.data_ptr(), out_p.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    ctx.close()
    return out_a.cpu().numpy(), out_p.cpu().numpy()


@pytest.fixture(scope="module")
def batch16(oracle):
    """16 instances on the 3984-transducer array, M from 2 to 128."""
    src, k, a = setup_array(oracle, (4, 4))
    rng = np.random.default_rng(42)
    ms = np.array([2, 5, 16, 17, 31, 32, 33, 47, 64, 3, 128, 100, 65, 9, 24, 40], np.int64)
    offs = np.zeros(ms.size + 1, np.uint32)
    offs[1:] = np.cumsum(ms)
    foci = random_foci(rng, int(offs[-1]), lo=(0, 0, 100), hi=(730, 570, 250))
    amps = rng.uniform(0.5, 1.0, int(offs[-1])).astype(np.float32)
    ref = []
    for b in range(ms.size):
        f = foci[offs[b]:offs[b + 1]]
        am = amps[offs[b]:offs[b + 1]]
        a64, p64 = oracle.gspat(1, src, k, a, f, am, "f64")
        ref.append(focal_amps(oracle, src, k, a, f, a64, p64))
    return dict(src=src, k=k, a=a, foci=foci, amps=amps, offs=offs, ref=ref)


@pytest.mark.parametrize("precision,tol", [(0, 1e-3), (3, 1e-3), (2, 1e-3), (1, 1e-3)])
def test_gspat_batched_vs_fp64(afc, oracle, torch_cuda, batch16, precision, tol):
    log(f"case: batched GS-PAT precision={precision}, 16 instances, N=3984")
    d = batch16
    amp, ph = run_batch(afc, torch_cuda, oracle, d["src"], d["k"], d["a"], d["foci"], d["amps"], d["offs"], 0,
                        precision)
    errs = []
    for b in range(d["offs"].shape[0] - 1):
        f = d["foci"][d["offs"][b]:d["offs"][b + 1]]
        got = focal_amps(oracle, d["src"], d["k"], d["a"], f, amp[b], ph[b])
        errs.append(float(np.max(np.abs(got - d["ref"][b]) / d["ref"][b])))
    ms = np.diff(d["offs"])
    log(f"  per-instance max rel focal-amplitude error: max {max(errs):.3e} (M={ms[int(np.argmax(errs))]}), "
        f"median {np.median(errs):.3e}")
    assert max(errs) <= tol


@pytest.fixture(scope="module")
def config3(oracle):
    """The 1024-instance GS-PAT batch on the 3984 array (M ~ U[2, 64], foci in
    [0, 170]^2 x [100, 250] mm, targets U[0.5, 1]) with the focal amplitudes of
    the oracle's FP64 run of every instance (one instance per host thread)."""
    from concurrent.futures import ThreadPoolExecutor
    src, k, a = setup_array(oracle, (4, 4))
    foci, amps, offs = oracle.gspat_batch(1024, m_range=(2, 64), seed=3)

    def ref_one(b):
        f, am = foci[offs[b]:offs[b + 1]], amps[offs[b]:offs[b + 1]]
        a64, p64 = oracle.gspat(1, src, k, a, f, am, "f64")
        return focal_amps(oracle, src, k, a, f, a64, p64)

    with ThreadPoolExecutor(max_workers=16) as ex:
        ref = list(ex.map(ref_one, range(1024)))
    return dict(src=src, k=k, a=a, foci=foci, amps=amps, offs=offs, ref=ref)


@pytest.mark.parametrize("precision", [0, 1, 2, 3])
def test_gspat_config3_every_instance(afc, oracle, torch_cuda, config3, precision):
    """gs_pat.rs:38-81 on all 1024 instances of the config-3 batch: every
    instance's focal amplitudes within 1e-3 relative of the FP64 oracle's."""
    d = config3
    log(f"case: GS-PAT config-3 batch, 1024 instances, precision={precision}")
    amp, ph = run_batch(afc, torch_cuda, oracle, d["src"], d["k"], d["a"], d["foci"], d["amps"], d["offs"], 0,
                        precision)
    errs = np.empty(1024)
    for b in range(1024):
        f = d["foci"][d["offs"][b]:d["offs"][b + 1]]
        got = focal_amps(oracle, d["src"], d["k"], d["a"], f, amp[b], ph[b])
        errs[b] = np.max(np.abs(got - d["ref"][b]) / d["ref"][b])
    ms = np.diff(d["offs"].astype(np.int64))
    w = int(np.argmax(errs))
    log(f"  max {errs[w]:.3e} (instance {w}, M={ms[w]}), p99 {np.percentile(errs, 99):.3e}, "
        f"median {np.median(errs):.3e}")
    assert errs[w] <= 1e-3


def test_gspat_instance_isolation(afc, oracle, torch_cuda, batch16):
    """An instance's result does not depend on its batch-mates."""
    d = batch16
    for precision in (0, 1, 2, 3):
        full_a, full_p = run_batch(afc, torch_cuda, oracle, d["src"], d["k"], d["a"], d["foci"], d["amps"], d["offs"],
                                   0, precision)
        for b in (0, 8, 10):
            lo, hi = d["offs"][b], d["offs"][b + 1]
            one = np.array([0, hi - lo], np.uint32)
            a1, p1 = run_batch(afc, torch_cuda, oracle, d["src"], d["k"], d["a"], d["foci"][lo:hi], d["amps"][lo:hi],
                               one, 0, precision)
            assert np.array_equal(a1[0], full_a[b]) and np.array_equal(p1[0], full_p[b])


def test_gs_and_naive_batched(afc, oracle, torch_cuda):
    log("case: batched GS / Naive on the 249 array")
    src, k, a = setup_array(oracle)
    rng = np.random.default_rng(3)
    # M = 40 pads to 64 foci; M = 80 pads to 128, which GS runs on its
    # two-pass kernel
    ms = np.array([2, 7, 20, 40, 80], np.int64)
    offs = np.zeros(len(ms) + 1, np.uint32)
    offs[1:] = np.cumsum(ms)
    foci = random_foci(rng, int(offs[-1]))
    amps = rng.uniform(0.5, 1.0, int(offs[-1])).astype(np.float32)
    for algo, ref in ((1, oracle.gs), (2, oracle.naive)):
        amp, ph = run_batch(afc, torch_cuda, oracle, src, k, a, foci, amps, offs, algo, 0)
        for b in range(len(ms)):
            f, am = foci[offs[b]:offs[b + 1]], amps[offs[b]:offs[b + 1]]
            a64, p64 = ref(1, src, k, a, f, am, "f64")
            got = focal_amps(oracle, src, k, a, f, amp[b], ph[b])
            want = focal_amps(oracle, src, k, a, f, a64, p64)
            assert np.max(np.abs(got - want) / want) <= 1e-3, (algo, b)


def test_gs_large_array(afc, oracle, torch_cuda):
    """GS on 4980 sources (5x4 tiles): more than 4096 padded sources, past
    the size up to which the two-pass kernel keeps q in LDS."""
    log("case: batched GS on the 4980-source array")
    src, k, a = setup_array(oracle, (5, 4))
    rng = np.random.default_rng(5)
    ms = np.array([3, 12], np.int64)
    offs = np.zeros(3, np.uint32)
    offs[1:] = np.cumsum(ms)
    foci = random_foci(rng, int(offs[-1]))
    amps = rng.uniform(0.5, 1.0, int(offs[-1])).astype(np.float32)
    amp, ph = run_batch(afc, torch_cuda, oracle, src, k, a, foci, amps, offs, 1, 0)
    for b in range(2):
        f, am = foci[offs[b]:offs[b + 1]], amps[offs[b]:offs[b + 1]]
        a64, p64 = oracle.gs(1, src, k, a, f, am, "f64")
        got = focal_amps(oracle, src, k, a, f, amp[b], ph[b])
        want = focal_amps(oracle, src, k, a, f, a64, p64)
        err = float(np.max(np.abs(got - want) / want))
        log(f"  M={ms[b]}: max rel focal-amplitude error {err:.2e}")
        assert err <= 1e-3


def greedy_objective(oracle, src, k, a, foci, amps, phase):
    p = focal_amps(oracle, src, k, a, foci, src[:, 6], phase)
    return float(np.sum(np.abs(amps - p))), p


@pytest.mark.parametrize("m,div", [(1, 64), (3, 16), (8, 32)])
def test_greedy_capi_vs_oracle(afc, oracle, torch_cuda, m, div):
    """AFO_Greedy (combination/greedy.rs:23-83) on the GPU against the FP64
    oracle. The choice at each source is discrete and the GPU's terms differ
    from the CPU's by ~1e-7, so a near-tie may resolve differently. The first
    source is always a tie (an empty cache makes every candidate equal), so
    its rounding-noise choice sets a global rotation by a whole number of
    candidate steps that the rest of the run follows; indices are compared
    relative to it. Checked: every phase is one of the div candidates;
    >= 90 % of the sources take the oracle's candidate (relative to the
    first); the objective sum_j |a_j - |p_j|| is
    within 2 % of the oracle's; and for an unreachable single target
    |p| >= sum_j |G_j| cos(pi / div)."""
    log(f"case: AFO_Greedy M={m} div={div}")
    rng = np.random.default_rng(40 + m)
    src, k, a = setup_array(oracle)
    foci = random_foci(rng, m)
    # m = 1: a target of 3, above sum_j |G_j| = 1.43, that no phase choice
    # reaches (a far larger one leaves the f32 objective unable to tell the
    # candidates apart, in the reference's FP32 arithmetic as well)
    amps = (np.full(m, 3.0, np.float32) if m == 1 else rng.uniform(0.5, 1.0, m).astype(np.float32))
    s = build_system(afc, src, oracle.TEMP_340)
    afc.Optimizer.greedy(s, foci, amps, div)
    got = s.source_array().copy()
    assert np.array_equal(got[:, 6], src[:, 6])  # amplitudes untouched
    ph = got[:, 7].astype(np.float64)
    ref = oracle.greedy(1, src, k, a, foci, amps, div, "f64")
    idx = np.round(np.mod(ph, 2 * np.pi) / (2 * np.pi / div)).astype(int) % div
    idx_ref = np.round(np.mod(ref, 2 * np.pi) / (2 * np.pi / div)).astype(int) % div
    assert np.allclose(np.angle(np.exp(1j * (ph - 2 * np.pi * idx / div))), 0, atol=1e-6)
    same = float(np.mean((idx - idx[0]) % div == (idx_ref - idx_ref[0]) % div))
    o_gpu, p = greedy_objective(oracle, src, k, a, foci, amps, ph)
    o_ref, _ = greedy_objective(oracle, src, k, a, foci, amps, ref)
    log(f"  same candidate for {same:.3f} of sources; objective {o_gpu:.6g} vs oracle {o_ref:.6g}")
    assert same >= 0.9
    assert abs(o_gpu - o_ref) <= 0.02 * o_ref
    if m == 1:
        G = oracle.propagation_matrix(1, src, k, a, foci, "f64")
        assert p[0] >= np.abs(G).sum() * np.cos(np.pi / div) * (1 - 1e-6)


def test_greedy_batched_equals_single(afc, oracle, torch_cuda):
    log("case: AFH_GreedySolve batch of 5 vs one at a time")
    import torch
    from afc import device as D
    rng = np.random.default_rng(77)
    src, k, a = setup_array(oracle)
    n = src.shape[0]
    ms = [1, 4, 7, 2, 12]
    foci = np.concatenate([random_foci(rng, m) for m in ms]).astype(np.float32)
    amps = rng.uniform(0.5, 1.0, foci.shape[0]).astype(np.float32)
    offs = np.concatenate([[0], np.cumsum(ms)]).astype(np.uint32)
    dev = torch.device("cuda:0")
    packed = D.pack_sources(1, src, k, a)
    src_d = torch.from_numpy(packed).to(dev)
    foci_d = torch.from_numpy(foci).to(dev)
    amps_d = torch.from_numpy(amps).to(dev)
    ctx = D.Context()
    out = torch.empty((len(ms), n), dtype=torch.float32, device=dev)
    D.greedy_solve(ctx, src_d.data_ptr(), packed.shape[0], n, 1, foci_d.data_ptr(), amps_d.data_ptr(), offs, 16,
                   out.data_ptr())
    torch.cuda.synchronize()
    batch = out.cpu().numpy()
    for b in range(len(ms)):
        one = torch.empty((1, n), dtype=torch.float32, device=dev)
        sub = np.array([0, ms[b]], np.uint32)
        D.greedy_solve(ctx, src_d.data_ptr(), packed.shape[0], n, 1, foci_d[offs[b]:].data_ptr(),
                       amps_d[offs[b]:].data_ptr(), sub, 16, one.data_ptr())
        torch.cuda.synchronize()
        assert np.array_equal(one.cpu().numpy()[0], batch[b]), b
    ctx.close()
