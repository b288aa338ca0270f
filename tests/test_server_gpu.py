"""Request server: continuous batching, HBM-budget admission and isolation.

Each request's result must be bit-identical when it runs alone and when it
runs in a mix of concurrent field and GS-PAT requests on several arrays (the
kernels keep a fixed per-request reduction order). A server with a small HBM
budget must still complete a stream larger than the budget (admission waits
for memory) and must refuse a request that can never fit."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


def arrays(oracle, tiles_list=((1, 1), (2, 2))):
    out = []
    for tiles in tiles_list:
        pos = oracle.array_249() if tiles == (1, 1) else oracle.tiled_array(*tiles)
        rng = np.random.default_rng(len(pos))
        src = oracle.t4010a1_sources(pos, phase=rng.uniform(0, 6.28, len(pos)).astype(np.float32))
        n = len(pos)
        k = np.full(n, oracle.WAVENUM_40K, np.float32)
        a = np.full(n, np.float32(oracle.attenuation_coef(40e3, 30, 1, 1, oracle.TEMP_340)), np.float32)
        out.append((src, k, a))
    return out


def make_requests(rng, n_arrays, count):
    reqs = []
    for i in range(count):
        aid = int(rng.integers(0, n_arrays))
        if rng.random() < 0.6:
            nx, ny, nz = (int(v) for v in rng.integers(8, 64, 3))
            origin = (float(rng.uniform(0, 100)), float(rng.uniform(0, 80)), float(rng.uniform(60, 200)))
            reqs.append(("field", aid, (nx, ny, nz), origin, int(rng.integers(0, 3))))
        else:
            m = int(rng.integers(2, 129))
            foci = np.stack([rng.uniform(0, 170, m), rng.uniform(0, 130, m), rng.uniform(100, 250, m)],
                            1).astype(np.float32)
            amps = rng.uniform(0.5, 1.0, m).astype(np.float32)
            reqs.append(("gspat", aid, foci, amps))
    return reqs


def submit(srv, ids, arr_n, req):
    from afc import device as D
    if req[0] == "field":
        _, aid, cnt, origin, ft = req
        ps = D.grid_pointset((0, 1, 2), cnt, origin, 1.0)
        out = np.empty(cnt[0] * cnt[1] * cnt[2] * (2 if ft == 2 else 1), np.float32)
        return srv.submit_field(ids[aid], ps, ft, out), (out,)
    _, aid, foci, amps = req
    oa, op = np.empty(arr_n[aid], np.float32), np.empty(arr_n[aid], np.float32)
    return srv.submit_gspat(ids[aid], foci, amps, oa, op), (oa, op)


def test_mix_equals_alone(afc, oracle, torch_cuda):
    from afc.server import Server
    arrs = arrays(oracle)
    rng = np.random.default_rng(0)
    reqs = make_requests(rng, len(arrs), 48)
    # GS-PAT batches on a reserved CU partition, field requests on the rest
    srv = Server(hbm_budget=256 << 20, pinned_budget=128 << 20, n_streams=4, max_gspat_batch=64, gspat_cus=32)
    ids = [srv.add_array(1, *a) for a in arrs]
    arr_n = [a[0].shape[0] for a in arrs]
    # alone: one request in the server at a time
    alone = []
    for r in reqs:
        t, bufs = submit(srv, ids, arr_n, r)
        srv.wait(t)
        alone.append(bufs)
    # mixed: three submitting threads, everything in flight together
    mixed = [None] * len(reqs)
    tickets = [None] * len(reqs)

    def worker(idx):
        for i in idx:
            tickets[i], mixed[i] = submit(srv, ids, arr_n, reqs[i])

    th = [threading.Thread(target=worker, args=(list(range(j, len(reqs), 3)),)) for j in range(3)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for i in range(len(reqs)):
        srv.wait(tickets[i])
    for i, r in enumerate(reqs):
        for x, y in zip(alone[i], mixed[i]):
            assert np.array_equal(x, y), (i, r[0])
    p50, p99, n = srv.stats()
    gp50, gp99, gn = srv.stats(kind=1)
    assert gn > 0 and gn < n
    log(f"  {n} requests completed; p50 {p50:.3f} ms p99 {p99:.3f} ms")
    # field results match the FP64 oracle (random phases: an incoherent field,
    # where the FP32 CPU path itself is ~1e-5 away from FP64)
    k = next(i for i, r in enumerate(reqs) if r[0] == "field" and r[4] == 2)
    _, aid, cnt, origin, _ = reqs[k]
    src, kk, aa = arrs[aid]
    pts, _ = oracle.grid_points((0, 1, 2), [(origin[a], origin[a] + cnt[a] - 1) for a in range(3)], 1.0)
    want = oracle.field(1, src, kk, aa, pts, "f64")
    got = alone[k][0].view(np.complex64).astype(np.complex128)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) <= 5e-6
    srv.close()


def test_small_budget_queues_and_rejects(afc, oracle, torch_cuda):
    from afc import AFCError
    from afc import device as D
    from afc.server import Server
    arrs = arrays(oracle)
    srv = Server(hbm_budget=4 << 20, pinned_budget=4 << 20, n_streams=2, max_gspat_batch=16)
    aid = srv.add_array(1, *arrs[0])
    # 16 requests of 1.6 MB each: 25.6 MB through a 4 MB budget
    outs, ts = [], []
    for i in range(16):
        ps = D.grid_pointset((0, 1, 2), (100, 100, 20), (30.0, 20.0, 100.0 + i), 1.0)
        out = np.empty(100 * 100 * 20 * 2, np.float32)
        ts.append(srv.submit_field(aid, ps, 2, out))
        outs.append(out)
    for t in ts:
        srv.wait(t)
    assert all(np.all(np.isfinite(o)) for o in outs)
    # 8 MB of output can never fit a 4 MB arena
    ps = D.grid_pointset((0, 1, 2), (100, 100, 100), (30.0, 20.0, 100.0), 1.0)
    with pytest.raises(AFCError, match="budget"):
        srv.submit_field(aid, ps, 2, np.empty(100 * 100 * 100 * 2, np.float32))
    srv.close()


def test_mix_equals_alone_3984_and_large_grids(afc, oracle, torch_cuda):
    """The isolation check with the three arrays of the served stream (249, 996
    and 3984 transducers) and field requests of 10^6+ points: every result of
    a mixed, concurrently submitted stream equals the same request run alone,
    bit for bit; one 10^6-point result on the 3984 array matches the FP64
    oracle at sampled points."""
    from afc.server import Server
    arrs = arrays(oracle, ((1, 1), (2, 2), (4, 4)))
    rng = np.random.default_rng(5)
    reqs = make_requests(rng, len(arrs), 24)
    big = [("field", 2, (128, 128, 64), (300.0, 220.0, 90.0), 2), ("field", 2, (256, 128, 32), (200.0, 150.0, 60.0), 0),
           ("field", 1, (100, 100, 101), (60.0, 40.0, 80.0), 1)]
    reqs = reqs[:8] + big[:1] + reqs[8:16] + big[1:2] + reqs[16:] + big[2:]
    reqs += [("gspat", 2, np.stack([rng.uniform(0, 730, m), rng.uniform(0, 570, m), rng.uniform(100, 250, m)],
                                   1).astype(np.float32), rng.uniform(0.5, 1.0, m).astype(np.float32))
             for m in (2, 64, 128)]
    log(f"case: server isolation, arrays 249/996/3984, {len(reqs)} requests incl. 3 grids of 10^6+ points")
    srv = Server(hbm_budget=512 << 20, pinned_budget=256 << 20, n_streams=4, max_gspat_batch=64, gspat_cus=32)
    ids = [srv.add_array(1, *a) for a in arrs]
    arr_n = [a[0].shape[0] for a in arrs]
    alone = []
    for r in reqs:
        t, bufs = submit(srv, ids, arr_n, r)
        srv.wait(t)
        alone.append(bufs)
    mixed = [None] * len(reqs)
    tickets = [None] * len(reqs)

    def worker(idx):
        for i in idx:
            tickets[i], mixed[i] = submit(srv, ids, arr_n, reqs[i])

    th = [threading.Thread(target=worker, args=(list(range(j, len(reqs), 3)),)) for j in range(3)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for i in range(len(reqs)):
        srv.wait(tickets[i])
    for i, r in enumerate(reqs):
        for x, y in zip(alone[i], mixed[i]):
            assert np.array_equal(x, y), (i, r[0])
    i = reqs.index(big[0])
    _, aid, cnt, origin, _ = reqs[i]
    src, kk, aa = arrs[aid]
    pts, _ = oracle.grid_points((0, 1, 2), [(origin[a], origin[a] + cnt[a] - 1) for a in range(3)], 1.0)
    sel = np.random.default_rng(9).choice(pts.shape[0], 2048, replace=False)
    want = oracle.field(1, src, kk, aa, pts[sel], "f64")
    got = alone[i][0].view(np.complex64).astype(np.complex128)[sel]
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    log(f"  10^6-point field on the 3984 array: rel-L2 vs FP64 at 2048 points {err:.2e}")
    assert err <= 5e-6
    srv.close()


@pytest.mark.parametrize("precision", [1, 0, 2])
def test_graph_launch_equals_direct(afc, oracle, torch_cuda, precision):
    """GS-PAT batches whose bucket chains run as cached hipGraphs (grids and
    arguments set per batch) give the same bits as kernel-by-kernel launches:
    every request of three rounds (alone, a concurrent mix that forms batches
    of other shapes, alone again) equals the direct-launch server's result
    alone, and later batches reuse the graphs that earlier ones built."""
    from afc.server import Server
    arrs = arrays(oracle, ((1, 1), (2, 2), (4, 4)))
    rng = np.random.default_rng(11 + precision)
    reqs = []
    for i in range(30):
        aid = i % 3
        m = int(rng.choice([2, 9, 16, 17, 31, 33, 64, 65, 100, 128]))
        span = (170, 130) if aid == 0 else ((360, 280) if aid == 1 else (730, 570))
        foci = np.stack([rng.uniform(0, span[0], m), rng.uniform(0, span[1], m), rng.uniform(100, 250, m)],
                        1).astype(np.float32)
        reqs.append(("gspat", aid, foci, rng.uniform(0.5, 1.0, m).astype(np.float32)))
    log(f"case: graph vs direct GS-PAT launches, precision {precision}, {len(reqs)} requests on 249/996/3984")
    cfg = dict(hbm_budget=256 << 20, pinned_budget=64 << 20, n_streams=4, max_gspat_batch=64, gspat_cus=32,
               gspat_precision=precision)
    direct = Server(**cfg)
    ids = [direct.add_array(1, *a) for a in arrs]
    arr_n = [a[0].shape[0] for a in arrs]
    want = []
    for r in reqs:
        t, bufs = submit(direct, ids, arr_n, r)
        direct.wait(t)
        want.append(bufs)
    assert direct.graph_counters() == (0, 0)
    direct.close()
    srv = Server(gspat_graphs=True, **cfg)
    ids = [srv.add_array(1, *a) for a in arrs]
    for rnd in range(3):
        got = [None] * len(reqs)
        if rnd == 1:
            tickets = [None] * len(reqs)

            def worker(idx):
                for i in idx:
                    tickets[i], got[i] = submit(srv, ids, arr_n, reqs[i])

            th = [threading.Thread(target=worker, args=(list(range(j, len(reqs), 3)),)) for j in range(3)]
            for t in th:
                t.start()
            for t in th:
                t.join()
            for t in tickets:
                srv.wait(t)
        else:
            for i, r in enumerate(reqs):
                t, got[i] = submit(srv, ids, arr_n, r)
                srv.wait(t)
        for i in range(len(reqs)):
            for x, y in zip(want[i], got[i]):
                assert np.array_equal(x, y), (rnd, i, reqs[i][2].shape[0])
    launches, builds = srv.graph_counters()
    log(f"  graph launches {launches}, builds {builds}")
    assert builds >= 1 and launches > 2 * builds
    srv.close()


def test_pinned_and_pageable_results_agree(afc, oracle, torch_cuda):
    """Field results bound for pinned (page-locked) caller buffers are DMA'd
    there directly; results bound for pageable buffers are staged and copied,
    in parts for large ones. Both give the same bits, for grid and scatter
    requests small and large (up to 8.4 MB of complex output)."""
    from afc import device as D
    from afc.server import Server
    torch = torch_cuda
    arrs = arrays(oracle, ((1, 1), (4, 4)))
    srv = Server(hbm_budget=512 << 20, pinned_budget=256 << 20, n_streams=4, max_gspat_batch=64)
    ids = [srv.add_array(1, *a) for a in arrs]
    rng = np.random.default_rng(3)
    cases = [((0, 1, 2), (16, 16, 1), 2), ((0, 1, 2), (512, 512, 4), 2), ((2, 0, 1), (100, 101, 99), 0),
             ((0, 1, 2), (1024, 1024, 1), 1)]
    log(f"case: pinned vs pageable field results, {len(cases)} grids + 1 scatter set on 249/3984")
    for aid in range(2):
        subs = []
        for order, cnt, ft in cases:
            ps = D.grid_pointset(order, cnt, (float(rng.uniform(0, 100)), float(rng.uniform(0, 80)), 60.0), 1.0)
            n = cnt[0] * cnt[1] * cnt[2] * (2 if ft == 2 else 1)
            pin = torch.full((n,), -1.0, dtype=torch.float32, pin_memory=True)
            page = np.full(n, -1.0, np.float32)
            subs.append((srv.submit_field(ids[aid], ps, ft, pin.numpy()), srv.submit_field(ids[aid], ps, ft, page),
                         pin, page))
        pts = np.stack([rng.uniform(0, 170, 300000), rng.uniform(0, 130, 300000), rng.uniform(80, 200, 300000)],
                       1).astype(np.float32)
        sp = D.scatter_pointset(pts.ctypes.data, pts.shape[0])  # host points: the server copies them
        pin = torch.full((600000,), -1.0, dtype=torch.float32, pin_memory=True)
        page = np.full(600000, -1.0, np.float32)
        subs.append((srv.submit_field(ids[aid], sp, 2, pin.numpy()), srv.submit_field(ids[aid], sp, 2, page),
                     pin, page))
        for tp, tg, pin, page in subs:
            srv.wait(tp)
            srv.wait(tg)
            got = pin.numpy()
            assert np.all(np.isfinite(got)) and np.array_equal(got, page)
    srv.close()


def test_batch_submit_equals_single(afc, oracle, torch_cuda):
    """AFS_SubmitGSPATBatch: requests queued at once give the same bits as the
    same requests submitted and waited one by one, and a batch holding an
    invalid request is rejected before any of it is queued."""
    from afc.server import Server
    arrs = arrays(oracle, ((1, 1), (2, 2), (4, 4)))
    rng = np.random.default_rng(23)
    ids_l, foci_l, amps_l = [], [], []
    for i in range(24):
        aid = i % 3
        m = int(rng.choice([2, 16, 17, 33, 64, 65, 128]))
        span = (170, 130) if aid == 0 else ((360, 280) if aid == 1 else (730, 570))
        foci_l.append(np.stack([rng.uniform(0, span[0], m), rng.uniform(0, span[1], m), rng.uniform(100, 250, m)],
                               1).astype(np.float32))
        amps_l.append(rng.uniform(0.5, 1.0, m).astype(np.float32))
        ids_l.append(aid)
    log(f"case: batched vs single GS-PAT submission, {len(ids_l)} requests on 249/996/3984")
    srv = Server(hbm_budget=256 << 20, pinned_budget=64 << 20, n_streams=2, max_gspat_batch=64, gspat_cus=0,
                 gspat_precision=1)
    ids = [srv.add_array(1, *a) for a in arrs]
    arr_n = [a[0].shape[0] for a in arrs]
    want = []
    for aid, f, a in zip(ids_l, foci_l, amps_l):
        oa, op = np.empty(arr_n[aid], np.float32), np.empty(arr_n[aid], np.float32)
        srv.wait(srv.submit_gspat(ids[aid], f, a, oa, op))
        want.append((oa, op))
    got_a = [np.empty(arr_n[aid], np.float32) for aid in ids_l]
    got_p = [np.empty(arr_n[aid], np.float32) for aid in ids_l]
    tickets = srv.submit_gspat_batch([ids[aid] for aid in ids_l], foci_l, amps_l, got_a, got_p)
    assert len(tickets) == len(ids_l) and tickets == sorted(tickets)
    for t in tickets:
        srv.wait(t)
    for i in range(len(ids_l)):
        assert np.array_equal(want[i][0], got_a[i]) and np.array_equal(want[i][1], got_p[i]), i
    # one request with 129 foci: nothing of the batch is queued
    bad_f = foci_l[:2] + [np.zeros((129, 3), np.float32)]
    bad_a = amps_l[:2] + [np.ones(129, np.float32)]
    from afc import AFCError
    with pytest.raises(AFCError):
        srv.submit_gspat_batch([ids[0], ids[1], ids[2]], bad_f, bad_a, got_a[:3], got_p[:3])
    assert srv.stats(kind=1)[2] >= len(ids_l)
    done_before = srv.stats(kind=1)[2]
    t = srv.submit_gspat(ids[0], foci_l[0], amps_l[0], got_a[0], got_p[0])
    srv.wait(t)
    assert srv.stats(kind=1)[2] == done_before + 1
    srv.close()
