"""AFH_HoloSolveNoAlloc (the allocation-free variant of AFH_HoloSolve, with its
workspace in the library's device scratch array) against AFH_HoloSolve itself:
the same kernels, so the same bits, on two calls in a row (the second with a
smaller batch on a smaller array, over the first call's leftovers), for
GS-PAT in every precision mode, GS, Naive; a batch beyond the scratch array is
refused with AFH_ERR_BUDGET before anything is launched."""
import ctypes
from ctypes import c_void_p

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


def problem(oracle, torch, tiles, batch, m_range, seed):
    from afc import device as D
    pos = oracle.array_249() if tiles == (1, 1) else oracle.tiled_array(*tiles)
    src = oracle.t4010a1_sources(pos)
    n = pos.shape[0]
    k = np.full(n, oracle.WAVENUM_40K, np.float32)
    a = np.full(n, np.float32(oracle.attenuation_coef(40e3, 30, 1, 1, oracle.TEMP_340)), np.float32)
    foci, amps, offs = oracle.gspat_batch(batch, n_tiles=tiles, m_range=m_range, seed=seed)
    dev = torch.device("cuda:0")
    packed = D.pack_sources(1, src, k, a)
    return dict(n=n, np=packed.shape[0], src=torch.from_numpy(packed).to(dev), foci=torch.from_numpy(foci).to(dev),
                amps=torch.from_numpy(amps).to(dev), offs=np.ascontiguousarray(offs, np.uint32), batch=batch)


def solve_noalloc(afc, torch, p, algo, prec, out_a, out_p):
    rc = afc.dll().AFH_HoloSolveNoAlloc(c_void_p(p["src"].data_ptr()), p["np"], p["n"], 1,
                                        c_void_p(p["foci"].data_ptr()), c_void_p(p["amps"].data_ptr()),
                                        p["offs"].ctypes.data_as(c_void_p), p["batch"], algo, prec, 100,
                                        c_void_p(out_a.data_ptr()), c_void_p(out_p.data_ptr()),
                                        c_void_p(torch.cuda.current_stream().cuda_stream))
    return rc


def solve_ctx(torch, p, algo, prec, out_a, out_p):
    from afc import device as D
    ctx = D.Context()
    D.holo_solve(ctx, p["src"].data_ptr(), p["np"], p["n"], 1, p["foci"].data_ptr(), p["amps"].data_ptr(), p["offs"],
                 algo, prec, 100, out_a.data_ptr(), out_p.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    ctx.close()


CASES = [(0, 0), (0, 1), (0, 2), (0, 3), (1, 0), (2, 0)]  # (algo, precision): GS-PAT x 4, GS, Naive


@pytest.mark.parametrize("algo,prec", CASES)
def test_no_alloc_matches_context_entry(afc, oracle, torch_cuda, algo, prec):
    torch = torch_cuda
    dev = torch.device("cuda:0")
    first = problem(oracle, torch, (4, 4), 24, (2, 64), 21)
    second = problem(oracle, torch, (2, 2), 5, (2, 128), 22)
    for name, p in (("first", first), ("second (smaller)", second)):
        log(f"case: AFH_HoloSolveNoAlloc algo={algo} precision={prec}, {name} call: "
            f"{p['batch']} instances on {p['n']} sources")
        want_a = torch.empty((p["batch"], p["n"]), dtype=torch.float32, device=dev)
        want_p = torch.empty_like(want_a)
        solve_ctx(torch, p, algo, prec, want_a, want_p)
        got_a = torch.full_like(want_a, float("nan"))
        got_p = torch.full_like(want_a, float("nan"))
        rc = solve_noalloc(afc, torch, p, algo, prec, got_a, got_p)
        assert rc == 0, afc.dll().AFC_GetLastError()
        torch.cuda.synchronize()
        assert torch.equal(got_a, want_a) and torch.equal(got_p, want_p)


def test_no_alloc_refuses_a_batch_beyond_the_array(afc, oracle, torch_cuda):
    torch = torch_cuda
    dev = torch.device("cuda:0")
    p = problem(oracle, torch, (4, 4), 1024, (2, 64), 3)
    out_a = torch.zeros((p["batch"], p["n"]), dtype=torch.float32, device=dev)
    out_p = torch.zeros_like(out_a)
    rc = solve_noalloc(afc, torch, p, 0, 1, out_a, out_p)
    assert rc == 5 and b"scratch" in afc.dll().AFC_GetLastError()
    torch.cuda.synchronize()
    assert bool((out_a == 0).all())
