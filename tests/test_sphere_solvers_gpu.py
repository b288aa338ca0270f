"""GPU parity of the hologram solvers with SphereWaveSource arrays (source
type 0, sphere_sources.rs:19-61: no directivity) against the oracle's FP64
restatements, at the tolerances of the T4010A1 tests:

  * GS-PAT batch (gs_pat.rs:38-81), FP32 and MIXED_SPLIT: focal amplitudes
    within 1e-3 relative of the FP64 oracle's solution;
  * Greedy (greedy.rs:23-83): objective sum_j |a_j - |p_j|| within 2 %;
  * LM over 20 iterations (lm.rs:40-105): iteration count equal, phases
    within 1e-5 rad;
  * Gauss-Newton (gauss_newton.rs:35-89): iteration count equal, phases within
    1e-5 rad;
  * APO (apo.rs:100-158), Long (long.rs:46-106), Horn (horn.rs:58-129):
    amplitudes within 1e-5 of the largest, phases within 1e-5 rad where the
    amplitude is above 1e-3 of the largest.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


def sphere_setup(oracle, tiles=(1, 1)):
    pos = oracle.array_249() if tiles == (1, 1) else oracle.tiled_array(*tiles)
    n = pos.shape[0]
    src = np.zeros((n, 6), np.float32)
    src[:, 0:3] = pos
    src[:, 3] = 1.0
    src[:, 5] = 40e3
    k = np.full(n, oracle.WAVENUM_40K, np.float32)
    a = np.full(n, np.float32(oracle.attenuation_coef(40e3, 30, 1, 1, oracle.TEMP_340)), np.float32)
    return src, k, a


def problem(rng, m, scale=1.0):
    foci = np.stack([rng.uniform(40, 130, m), rng.uniform(30, 110, m), rng.uniform(120, 200, m)], 1).astype(np.float32)
    amps = (rng.uniform(0.5, 1.0, m) * scale).astype(np.float32)
    return foci, amps


def focal(oracle, src, k, a, foci, amp, phase):
    s = src.copy()
    s[:, 3] = amp
    s[:, 4] = phase
    return np.abs(oracle.field(0, s, k, a, foci, "f64"))


def wrapdiff(x, y):
    return np.abs(np.angle(np.exp(1j * (np.asarray(x, np.float64) - np.asarray(y, np.float64)))))


class Dev:
    """Device copies of one problem (packed sphere sources, foci, targets)."""

    def __init__(self, torch, src, k, a, foci, amps):
        from afc import device as D
        self.D, self.torch = D, torch
        dev = torch.device("cuda:0")
        packed = D.pack_sources(0, src, k, a)
        self.n_packed, self.n = packed.shape[0], src.shape[0]
        self.src = torch.from_numpy(packed).to(dev)
        self.foci = torch.from_numpy(np.ascontiguousarray(foci, np.float32)).to(dev)
        self.amps = torch.from_numpy(np.ascontiguousarray(amps, np.float32)).to(dev)
        self.dev = dev
        self.ctx = D.Context()

    def out(self, batch, dtype=None):
        return self.torch.empty((batch, self.n), dtype=dtype or self.torch.float32, device=self.dev)

    def args(self):
        return (self.ctx, self.src.data_ptr(), self.n_packed, self.n, 0, self.foci.data_ptr(), self.amps.data_ptr())

    def done(self, *ts):
        self.torch.cuda.synchronize()
        self.ctx.close()
        return [t.cpu().numpy() for t in ts]


@pytest.mark.parametrize("precision", [0, 3])
def test_sphere_gspat_batch(afc, oracle, torch_cuda, precision):
    log(f"case: GS-PAT batch, sphere sources, precision {precision}")
    import torch
    rng = np.random.default_rng(60 + precision)
    src, k, a = sphere_setup(oracle, (2, 2))
    ms = [2, 9, 33, 17]
    fs, am = zip(*[problem(rng, m) for m in ms])
    foci, amps = np.concatenate(fs), np.concatenate(am)
    foci[:, 0] *= 2.0
    offs = np.concatenate([[0], np.cumsum(ms)]).astype(np.uint32)
    d = Dev(torch, src, k, a, foci, amps)
    oa, op = d.out(len(ms)), d.out(len(ms))
    d.D.holo_solve(*d.args(), offs, d.D.ALGO_GSPAT, precision, 100, oa.data_ptr(), op.data_ptr())
    oa, op = d.done(oa, op)
    for b in range(len(ms)):
        sl = slice(offs[b], offs[b + 1])
        a64, p64 = oracle.gspat(0, src, k, a, foci[sl], amps[sl], "f64")
        f_gpu = focal(oracle, src, k, a, foci[sl], oa[b], op[b])
        f_ref = focal(oracle, src, k, a, foci[sl], a64, p64)
        err = float(np.max(np.abs(f_gpu - f_ref) / f_ref))
        log(f"  M={ms[b]}: focal amplitude rel err {err:.2e}")
        assert err <= 1e-3


def test_sphere_greedy(afc, oracle, torch_cuda):
    log("case: Greedy, sphere sources, M=4, div=16")
    import torch
    rng = np.random.default_rng(63)
    src, k, a = sphere_setup(oracle)
    foci, amps = problem(rng, 4)
    d = Dev(torch, src, k, a, foci, amps)
    op = d.out(1)
    d.D.greedy_solve(*d.args(), np.array([0, 4], np.uint32), 16, op.data_ptr())
    (ph,) = d.done(op)
    ref = oracle.greedy(0, src, k, a, foci, amps, 16, "f64")

    def objective(phase):
        return float(np.sum(np.abs(amps - focal(oracle, src, k, a, foci, src[:, 3], phase))))

    o_gpu, o_ref = objective(ph[0]), objective(ref)
    log(f"  objective {o_gpu:.6g} vs oracle {o_ref:.6g}")
    assert abs(o_gpu - o_ref) <= 0.02 * o_ref


@pytest.mark.parametrize("algo,k_max", [(0, 20), (1, 200)])
def test_sphere_lm_gauss_newton(afc, oracle, torch_cuda, algo, k_max):
    name = "LM" if algo == 0 else "Gauss-Newton"
    log(f"case: {name}, sphere sources, M=3, k_max={k_max}")
    import torch
    rng = np.random.default_rng(64 + algo)
    src, k, a = sphere_setup(oracle)
    foci, amps = problem(rng, 3)
    d = Dev(torch, src, k, a, foci, amps)
    op = d.out(1)
    it = torch.zeros(1, dtype=torch.int32, device=d.dev)
    d.D.nl_solve(*d.args(), np.array([0, 3], np.uint32), algo, k_max, op.data_ptr(), it.data_ptr())
    ph, its = d.done(op, it)
    ref_fn = oracle.lm if algo == 0 else oracle.gauss_newton
    ref, _, it_ref = ref_fn(0, src, k, a, foci, amps, "f64", k_max=k_max)
    diff = float(wrapdiff(ph[0], ref).max())
    log(f"  iterations gpu {its[0]} oracle {it_ref}; max phase diff {diff:.2e}")
    assert int(its[0]) == it_ref
    assert diff <= 1e-5


def compare(oa, op, ra, rp):
    big = ra > 1e-3 * ra.max()
    return float(np.max(np.abs(oa - ra) / ra.max())), float(wrapdiff(op[big], rp[big]).max())


@pytest.mark.parametrize("method", ["apo", "long", "horn"])
def test_sphere_amp_phase_solvers(afc, oracle, torch_cuda, method):
    log(f"case: {method}, sphere sources, M=5")
    import torch
    rng = np.random.default_rng(66)
    src, k, a = sphere_setup(oracle)
    foci, amps = problem(rng, 5, 0.2 if method == "apo" else 1.0)
    d = Dev(torch, src, k, a, foci, amps)
    oa, op = d.out(1), d.out(1)
    offs = np.array([0, 5], np.uint32)
    if method == "apo":
        d.D.apo_solve(*d.args(), offs, 0.1, 200, oa.data_ptr(), op.data_ptr())
        ra, rp, _ = oracle.apo(0, src, k, a, foci, amps, np.float32(0.1), "f64")
    elif method == "long":
        d.D.long_solve(*d.args(), offs, oa.data_ptr(), op.data_ptr())
        ra, rp = oracle.long_(0, src, k, a, foci, amps)
    else:
        d.D.horn_solve(*d.args(), offs, oa.data_ptr(), op.data_ptr(), seed=11)
        ra, rp = oracle.horn(0, src, k, a, foci, amps, seed=11, lam=np.float32(0.8), alpha=np.float32(1e-5))
    oa, op = d.done(oa, op)
    ea, ep = compare(oa[0], op[0], ra, rp)
    log(f"  amp err {ea:.2e}, phase err {ep:.2e}")
    assert ea <= 1e-5 and ep <= 1e-5
