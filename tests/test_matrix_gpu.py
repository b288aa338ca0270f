"""GPU parity of Long (matrix/long.rs:46-106) and Horn (matrix/horn.rs:58-129)
against the oracle's FP64 restatements.

The GPU solves Long's N x N normal equations through the M x M identity
q = D^{-1} G^H (I + G D^{-1} G^H)^{-1} f, and Horn's pseudo-inverse as
G^H (G G^H + a^2 I)^{-1}; the oracle uses the dense forms. Both share the
eigenvector convention (ascending eigenvalues, largest entry of each vector
real positive) and Horn's seeded row sequence (instance b: seed + b).
Tolerances: amplitudes within 1e-5 of the largest, phases within 1e-5 rad
where the amplitude is above 1e-3 of the largest (the argument of a
vanishing q_j is noise); a batch equals its instances solved one at a time
bit for bit.
"""
import ctypes

import numpy as np
import pytest

from test_nls_gpu import log, problem, setup, wrapdiff

pytestmark = pytest.mark.gpu


def run_matrix(torch, method, src, k, a, foci, amps, offs, seed=0, repeat=100, lam=0.8, tikhonov=1e-5):
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
    ctx = D.Context()
    if method == "long":
        D.long_solve(ctx, src_d.data_ptr(), packed.shape[0], n, 1, foci_d.data_ptr(), amps_d.data_ptr(), offs,
                     oa.data_ptr(), op.data_ptr())
    else:
        D.horn_solve(ctx, src_d.data_ptr(), packed.shape[0], n, 1, foci_d.data_ptr(), amps_d.data_ptr(), offs,
                     oa.data_ptr(), op.data_ptr(), seed=seed, repeat=repeat, lam=lam, tikhonov=tikhonov)
    torch.cuda.synchronize()
    ctx.close()
    return oa.cpu().numpy(), op.cpu().numpy()


def compare(oa, op, ra, rp):
    big = ra > 1e-3 * ra.max()
    ea = float(np.max(np.abs(oa - ra) / ra.max()))
    ep = float(wrapdiff(op[big], rp[big]).max())
    return ea, ep


@pytest.mark.parametrize("m,seed", [(1, 31), (4, 32), (12, 33), (32, 34)])
def test_long_vs_oracle(afc, oracle, torch_cuda, m, seed):
    log(f"case: Long M={m}")
    import torch
    rng = np.random.default_rng(seed)
    src, k, a = setup(oracle)
    foci, amps = problem(rng, m)
    oa, op = run_matrix(torch, "long", src, k, a, foci, amps, np.array([0, m], np.uint32))
    ra, rp = oracle.long_(1, src, k, a, foci, amps)
    ea, ep = compare(oa[0], op[0], ra, rp)
    log(f"  amp err {ea:.2e}, phase err {ep:.2e}")
    assert ea <= 1e-5 and ep <= 1e-5


@pytest.mark.parametrize("m,seed,rseed", [(1, 41, 0), (4, 42, 7), (12, 43, 8), (32, 44, 9)])
def test_horn_vs_oracle(afc, oracle, torch_cuda, m, seed, rseed):
    log(f"case: Horn M={m} seed={rseed}")
    import torch
    rng = np.random.default_rng(seed)
    src, k, a = setup(oracle)
    foci, amps = problem(rng, m)
    oa, op = run_matrix(torch, "horn", src, k, a, foci, amps, np.array([0, m], np.uint32), seed=rseed)
    ra, rp = oracle.horn(1, src, k, a, foci, amps, seed=rseed, lam=np.float32(0.8), alpha=np.float32(1e-5))
    ea, ep = compare(oa[0], op[0], ra, rp)
    log(f"  amp err {ea:.2e}, phase err {ep:.2e}")
    assert ea <= 1e-5 and ep <= 1e-5


def test_horn_parameters(afc, oracle, torch_cuda):
    """Non-default repeat, lambda and Tikhonov parameter (the reference's
    setters, horn.rs:46-56)."""
    log("case: Horn repeat=7 lambda=0.5 tikhonov=0.5")
    import torch
    rng = np.random.default_rng(45)
    src, k, a = setup(oracle, 160)
    foci, amps = problem(rng, 6)
    oa, op = run_matrix(torch, "horn", src, k, a, foci, amps, np.array([0, 6], np.uint32), seed=3, repeat=7, lam=0.5,
                        tikhonov=0.5)
    ra, rp = oracle.horn(1, src, k, a, foci, amps, seed=3, repeat=7, lam=0.5, alpha=0.5)
    ea, ep = compare(oa[0], op[0], ra, rp)
    log(f"  amp err {ea:.2e}, phase err {ep:.2e}")
    assert ea <= 1e-5 and ep <= 1e-5


@pytest.mark.parametrize("method", ["long", "horn"])
def test_matrix_batch_equals_single(afc, oracle, torch_cuda, method):
    log(f"case: batched {method} equals one instance at a time")
    import torch
    rng = np.random.default_rng(46)
    src, k, a = setup(oracle, 120)
    ms = [3, 1, 7, 32]
    fs, am = zip(*[problem(rng, m) for m in ms])
    foci, amps = np.concatenate(fs), np.concatenate(am)
    offs = np.concatenate([[0], np.cumsum(ms)]).astype(np.uint32)
    oa, op = run_matrix(torch, method, src, k, a, foci, amps, offs, seed=100)
    for b in range(len(ms)):
        sl = slice(offs[b], offs[b + 1])
        a1, p1 = run_matrix(torch, method, src, k, a, foci[sl], amps[sl], np.array([0, ms[b]], np.uint32),
                            seed=100 + b)
        assert np.array_equal(a1[0], oa[b]) and np.array_equal(p1[0], op[b])


@pytest.mark.parametrize("method", ["long", "horn"])
def test_matrix_capi(afc, oracle, torch_cuda, method):
    """AFO_Long / AFO_Horn write amplitudes and phases into the system's
    sources (long.rs:97-104, horn.rs:121-128); Horn draws from seed 0."""
    log(f"case: AFO_{method} through the C ABI")
    rng = np.random.default_rng(47)
    src, _, _ = setup(oracle)
    foci, amps = problem(rng, 5)
    s = afc.UniformSystem(oracle.TEMP_340)
    for row in src:
        s.add_wave_source(afc.T4010A1(row[0:3], row[3:6], row[6], row[7], row[8]))
    getattr(afc.Optimizer, method)(s, foci, amps)
    got = s.source_array().copy()
    kk = np.zeros(len(src), np.float32)
    aa = np.zeros(len(src), np.float32)
    afc.dll().AFH_GetSourceConstants(s.handle, kk.ctypes.data_as(ctypes.c_void_p), aa.ctypes.data_as(ctypes.c_void_p))
    if method == "long":
        ra, rp = oracle.long_(1, src, kk, aa, foci, amps)
    else:
        ra, rp = oracle.horn(1, src, kk, aa, foci, amps, seed=0)
    ea, ep = compare(got[:, 6], got[:, 7], ra, rp)
    log(f"  amp err {ea:.2e}, phase err {ep:.2e}")
    assert ea <= 1e-5 and ep <= 1e-5


@pytest.mark.parametrize("method", ["long", "horn"])
def test_matrix_rejects_bad_sizes(afc, oracle, torch_cuda, method):
    """More than 32 foci, or an instance with none, is refused before any launch."""
    import torch
    rng = np.random.default_rng(48)
    src, k, a = setup(oracle, 64)
    foci, amps = problem(rng, 33)
    for offs in (np.array([0, 33], np.uint32), np.array([0, 2, 2], np.uint32)):
        with pytest.raises(afc.AFCError):
            run_matrix(torch, method, src, k, a, foci, amps, offs)
