"""Field-evaluation benchmark on one MI355X.

Workload: 16 tiled 249-transducer arrays (3984 T4010A1 sources, 40 kHz,
random phases, amplitude 1) observed on a 256 x 256 x 128 grid at 1 mm
(8 388 608 points) from z = 50 mm; complex pressure output, FP32 arithmetic.
One step = one evaluation of the whole volume (3.342e10 source-point terms)
with sources and output resident in HBM and the grid generated on device.

A second section times batched GS-PAT: 1024 instances on the same 3984-source
array, 2..64 foci each (uniform), foci in [0,170]^2 x [100,250] mm, targets in
[0.5, 1], 100 iterations, per precision mode; it reports solves/s, the batch
latency and the focal-amplitude error of sampled instances against the FP64
oracle, beside the unbatched plain FP32 path (AFO_GSPAT, one instance per call).

Prints one JSON line (metric, value in Gsrc*pt/s, roofline of the field
kernel, CPU baseline of the reference's FP32 path restated in C, GS-PAT).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "acoustic-field-calculator_amd"))

FLOP_PER_TERM = 38  # T4010A1 term, counted from uniform_medium.rs:111-119 and t4010a1.rs:107-120
# HBM bytes of one field launch at 3984 sources x 256x256x128 points, measured
# with rocprofv3 (FETCH_SIZE scaled by 2.0 and WRITE_SIZE by 1.0, the factors
# a 256 MiB device copy reads)
FIELD_TRAFFIC_BYTES = 68.8e6
PEAK_FP32 = 157.3  # TFLOP/s, MI355X vector FP32
# Static VALU mix of one term of the timed field kernel (field_zrun_kernel,
# T4010A1, complex output), per source-point term of its inner loop: wave64
# VALU instructions, of which transcendental (rsq, exp, sin, cos) and FMAs with
# three VGPR sources (v_fma_f32, which issue at half rate on gfx950). Read from
# the kernel's gfx950 assembly (145 VALU, 16 transcendental, 36 three-source
# FMAs per 2 sources x 2 points) and checked against rocprofv3 SQ_INSTS_VALU /
# SQ_INSTS_VALU_TRANS_F32 per dispatch (profiles/field_kernel_3984src_256x256x128.md).
FIELD_VALU_PER_TERM = 36.25
FIELD_TRANS_PER_TERM = 4.0
FIELD_FMA3_PER_TERM = 9.0
PEAK_ENGINE_GHZ = 2.4  # MI355X peak engine clock
ISSUE_LIB = os.path.join(ROOT, "acoustic-field-calculator_amd", "libafc_issue.so")
ISSUE_KINDS = {"fma": 0, "mul": 1, "sin": 2, "rsq": 3, "exp": 4, "mix_9fma_1sin": 5, "fmac_vop2": 6,
               "fma_const": 7, "mul_vop3": 8}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def issue_rate(kinds=("fma", "mul", "sin"), waves_per_simd=8, reps=4000):
    """SIMD cycles per wave64 VALU instruction of each class, measured live by
    the calibration kernels of libafc_issue.so (measure/issue_rate.hip): a long
    stream of one instruction class on 16 independent chains per lane,
    waves_per_simd waves on every SIMD; cycles = kernel time (HIP events) x
    shader clock (s_memtime / s_memrealtime of the median wave) x SIMDs /
    wave instructions."""
    import ctypes

    if not os.path.exists(ISSUE_LIB):
        raise ImportError(f"{ISSUE_LIB} not found: build it with `python acoustic-field-calculator_amd/build.py`")
    lib = ctypes.CDLL(ISSUE_LIB)
    fn = lib.afi_issue_rate
    fn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    fn.restype = ctypes.c_int
    out = (ctypes.c_double * 5)()
    res = {}
    for name in kinds:
        rc = fn(ISSUE_KINDS[name], waves_per_simd, reps, out)
        if rc != 0:
            raise RuntimeError(f"issue-rate calibration {name} failed ({rc})")
        res[name] = {"cycles_per_instr": round(out[3], 3), "cycles_per_instr_wave_clock": round(out[4], 3),
                     "clock_ghz": round(out[1], 3), "kernel_ms": round(out[0], 3), "wave_instructions": out[2]}
    return res


def field_issue_model(cal, terms_per_s):
    """Issue-slot ceiling of the field kernel: its static VALU mix per term
    priced with the measured per-class costs (three-source FMA, two-operand
    VALU, transcendental), at the peak engine clock, and the kernel's fraction
    of it."""
    c3, c2, ct = (cal[k]["cycles_per_instr"] for k in ("fma", "mul", "sin"))
    n2 = FIELD_VALU_PER_TERM - FIELD_TRANS_PER_TERM - FIELD_FMA3_PER_TERM
    cyc = FIELD_FMA3_PER_TERM * c3 + n2 * c2 + FIELD_TRANS_PER_TERM * ct
    simds = 1024  # 256 CUs x 4
    ceiling = simds * PEAK_ENGINE_GHZ * 1e9 / cyc * 64
    return {"valu_per_term": FIELD_VALU_PER_TERM, "fma3_per_term": FIELD_FMA3_PER_TERM,
            "trans_per_term": FIELD_TRANS_PER_TERM, "cycles_fma3": c3, "cycles_2op": c2, "cycles_trans": ct,
            "cycles_per_wave_term": round(cyc, 2), "clock_ghz": PEAK_ENGINE_GHZ,
            "ceiling_gsrc_pt_per_s": round(ceiling / 1e9, 1), "frac_of_issue_ceiling": round(terms_per_s / ceiling, 4),
            "source": "per-class cycles measured live (bench.py issue_rate, 8 waves per SIMD); VALU mix per term "
                      "from the kernel's assembly, checked with rocprofv3 counters"}


def cpu_threads():
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        n = min(n, int(env))
    return max(1, n)


def cpu_leg_oracle():
    """The cpu_baseline leg's oracle: the CPU restatement of the reference under
    oracle/ (test infrastructure). This file uses it for two things only, both
    outside every timed GPU region: the CPU baselines it times, and the FP64
    references it checks the GPU results against on samples of the same
    workloads. Every input comes from afc.workloads."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    return oracle


def cpu_baseline(src, k, a, order, counts, origin, res):
    """The reference CPU path (FP32, restated in C, OpenMP over points) on whole
    z-slices of the same volume, until ~10 s of wall time or 8 slices."""
    O = cpu_leg_oracle()
    threads = cpu_threads()
    nx, ny = counts[0], counts[1]
    done_pts, t_total, slices = 0, 0.0, 0
    while slices < 8 and t_total < 10.0:
        z = origin[2] + slices * 16 * res
        pts, _ = O.grid_points((0, 1), ((origin[0], origin[0] + (nx - 1) * res),
                                        (origin[1], origin[1] + (ny - 1) * res), (z, z)), res)
        t0 = time.perf_counter()
        O.field(1, src, k, a, pts, "f32", threads=threads)
        t_total += time.perf_counter() - t0
        done_pts += pts.shape[0]
        slices += 1
    terms = done_pts * src.shape[0]
    return {"value": round(terms / t_total / 1e9, 4), "unit": "Gsrc*pt/s", "cores": threads, "kind": "port",
            "sample": f"{slices} z-slices of 256x256 points x 3984 sources ({terms:.3e} terms), "
                      f"{t_total:.1f} s wall on {threads} threads"}


def cpu_baseline_instances(run_one, count, what):
    """A solver restated in C (FP32, single-threaded per call, as the
    reference's optimizers are) over `count` instances, one instance per host
    thread at a time."""
    from concurrent.futures import ThreadPoolExecutor
    threads = cpu_threads()
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(run_one, range(count)))
    dt = time.perf_counter() - t0
    return {"value": round(count / dt, 1), "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"{what}: {count} instances, {dt:.2f} s wall on {threads} threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", choices=["fast", "accurate"], default="fast")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-gspat", action="store_true")
    ap.add_argument("--gspat-batch", type=int, default=1024)
    ap.add_argument("--no-lm", action="store_true")
    ap.add_argument("--no-greedy", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-qps", default="500,1000,2000,3000")
    ap.add_argument("--serve-seconds", type=float, default=4.0, help="Poisson arrivals per rate, mixed stream")
    ap.add_argument("--serve-gspat-seconds", type=float, default=1.5, help="Poisson arrivals per rate, GS-PAT ramp")
    ap.add_argument("--serve-gspat-qps", default="10000,40000,80000,100000,120000,140000,160000,200000,240000,"
                                                 "320000,400000,500000,640000,800000,1000000")
    ap.add_argument("--serve-gspat-3984-qps",
                    default="10000,20000,40000,60000,80000,90000,100000,110000,120000,140000,160000,200000")
    ap.add_argument("--serve-gspat-cus", type=int, default=32, help="CUs reserved for GS-PAT batches while serving")
    ap.add_argument("--serve-pageable", action="store_true",
                    help="mixed stream: field results into pageable buffers (default: pinned, DMA'd directly)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 256 MiB device-to-device copy (reference for HBM byte counters)")
    ap.add_argument("--calibrate-issue", action="store_true",
                    help="only measure SIMD cycles per wave64 VALU instruction of each class (fma, mul, sin, rsq, "
                         "exp, 9 fma : 1 sin) at 4 and 8 waves per SIMD")
    args = ap.parse_args()
    if args.calibrate_issue:
        import torch  # noqa: F401  (one HIP runtime per process: torch's, loaded first)

        out = {"calibrate_issue": {f"waves_per_simd_{w}": issue_rate(tuple(ISSUE_KINDS), w) for w in (4, 8)}}
        print(json.dumps(out), flush=True)
        return
    if args.calibrate_copy:
        import torch
        x = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": x.numel() * 4}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              f"the field path shards by points with no communication", file=sys.stderr)
        sys.exit(2)

    import torch
    import afc
    from afc import device as D

    from afc import workloads as W

    # the oracle comes in through the cpu_baseline leg only, as the checker and
    # the timed CPU baseline, never for inputs or in a timed GPU region
    O = cpu_leg_oracle()

    src, k, a, area = W.volume_3984(seed=2)
    n = src.shape[0]
    ps, cnt = D.grid_from_area(area["order"], area["ranges"], area["res"])
    n_pts = cnt[0] * cnt[1] * cnt[2]
    packed = D.pack_sources(1, src, k, a)
    dev = torch.device("cuda:0")
    src_dev = torch.from_numpy(packed).to(dev)
    out = torch.empty((n_pts, 2), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream(dev)
    prec = D.PREC_ACCURATE if args.precision == "accurate" else D.PREC_FAST

    def step():
        D.field_eval(src_dev.data_ptr(), packed.shape[0], 1, ps, D.FIELD_COMPLEX, prec, out.data_ptr(),
                     stream.cuda_stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        step()
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = float(np.mean([s.elapsed_time(e) for s, e in zip(starts, ends)]))
    terms = n * n_pts
    value = terms * args.steps / wall / 1e9
    log(f"field kernel: {kern_ms:.3f} ms/launch, {terms / kern_ms / 1e6:.1f} Gsrc*pt/s by events, "
        f"{value:.1f} Gsrc*pt/s by wall clock")

    # PCIe-inclusive rate (what AFC_Calculate delivers to host memory): the
    # same launch followed by the copy of the field into pinned host memory
    pinned = torch.empty((n_pts, 2), dtype=torch.float32, pin_memory=True)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(3):
        step()
        pinned.copy_(out, non_blocking=True)
    e1.record(stream)
    torch.cuda.synchronize()
    with_copy_ms = e0.elapsed_time(e1) / 3
    log(f"field + device-to-host copy: {with_copy_ms:.3f} ms")

    # accuracy of the timed output on a sample of points vs FP64
    host = pinned.numpy().view(np.complex64).reshape(-1)
    rng = np.random.default_rng(0)
    sample = np.sort(rng.choice(n_pts, 512, replace=False))
    nx, ny = cnt[0], cnt[1]
    ix, iy, iz = sample % nx, (sample // nx) % ny, sample // (nx * ny)
    o = [area["ranges"][i][0] for i in range(3)]
    spts = np.stack([np.float32(o[0]) + ix.astype(np.float32), np.float32(o[1]) + iy.astype(np.float32),
                     np.float32(o[2]) + iz.astype(np.float32)], 1).astype(np.float32)
    ref64 = O.field(1, src, k, a, spts, "f64")
    err = float(np.linalg.norm(host[sample] - ref64) / np.linalg.norm(ref64))
    log(f"rel-L2 vs FP64 on 512 sampled points: {err:.3e}")

    achieved = terms * FLOP_PER_TERM / (kern_ms * 1e-3) / 1e12
    issue = field_issue_model(issue_rate(), terms / (kern_ms * 1e-3))
    log(f"issue model: {issue['cycles_fma3']} / {issue['cycles_2op']} / {issue['cycles_trans']} cycles per "
        f"3-source FMA / 2-operand / transcendental, {issue['cycles_per_wave_term']} per wave-term at "
        f"{issue['clock_ghz']} GHz -> ceiling {issue['ceiling_gsrc_pt_per_s']} Gsrc*pt/s, "
        f"kernel at {issue['frac_of_issue_ceiling']:.3f} of it")
    result = {
        "metric": "field Gsrc*pt/s (%VALU peak); GS-PAT solves/s + p50/p99 latency",
        "value": round(value, 3),
        "unit": "Gsrc*pt/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded PCG64 phases, tiled T4010A1 arrays)",
        "config": {"workload": "field 3984 T4010A1 x 256x256x128 grid, complex output", "sources": n,
                   "points": n_pts, "precision": args.precision},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP32, 4), "traffic": FIELD_TRAFFIC_BYTES,
                     "traffic_source": "rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0 per dispatch "
                                       "(profiles/field_kernel_3984src_256x256x128.md)",
                     "algorithmic_bytes": n_pts * 8 + packed.nbytes,
                     "flop_per_term": FLOP_PER_TERM, "kernel_ms": round(kern_ms, 4),
                     "issue_model": issue},
        "accuracy": {"rel_l2_vs_fp64": err, "sample": "512 random points of the timed output"},
        "pcie_inclusive": {"ms": round(with_copy_ms, 4), "gsrc_pt_per_s": round(terms / with_copy_ms / 1e6, 1),
                           "note": "launch + copy of the 8 B/point field into pinned host memory"},
    }
    if not args.no_gspat:
        result["gspat"] = bench_gspat(torch, afc, D, O, src, k, a, args.gspat_batch, not args.no_cpu_baseline)
    result["capi_focus_plane_249"] = bench_capi_focus_plane_249(afc, O, not args.no_cpu_baseline)
    if not args.no_lm:
        result["lm"] = bench_lm(torch, D, O, not args.no_cpu_baseline)
    if not args.no_greedy:
        result["greedy"] = bench_greedy(torch, D, O)
    if not args.no_serve:
        result["serve"] = bench_serve(D, [float(q) for q in args.serve_qps.split(",")], args.serve_seconds,
                                      args.serve_gspat_cus, pinned=not args.serve_pageable)
        result["serve_gspat"] = bench_serve_gspat(afc, D, [float(q) for q in args.serve_gspat_qps.split(",")],
                                                  args.serve_gspat_seconds)
        # the largest array alone: the most work per request
        result["serve_gspat_3984"] = bench_serve_gspat(
            afc, D, [float(q) for q in args.serve_gspat_3984_qps.split(",")], args.serve_gspat_seconds,
            precisions=(("mixed", 1, False), ("fp32", 0, False)), arrays=(2,))
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        result["cpu_baseline"] = cpu_baseline(src, k, a, area["order"], cnt, o, area["res"])
    print(json.dumps(result), flush=True)


def bench_serve(D, qps_list, seconds, gspat_cus=32, pinned=True):
    """Served mixed stream: Poisson arrivals at a fixed QPS, 70 % field
    evaluations (grids of 1e4..1e6 points, log-uniform, 1 mm, complex output)
    and 30 % GS-PAT solves (2..128 foci), each on one of the 249-, 996- or
    3984-transducer arrays chosen uniformly; random seeded phases and targets."""
    from afc import workloads as W

    import queue
    import threading
    from afc.server import Server

    arrs = []
    for tiles in ((1, 1), (2, 2), (4, 4)):
        n = 249 * tiles[0] * tiles[1]
        src, k, a = W.array_rows(tiles, phase_seed=n)
        ext = src[:, 0:3].max(0)
        arrs.append((src, k, a, ext))
    out = {"workload": "Poisson arrivals; 70% field (1e4..1e6 pts log-uniform, complex) / 30% GS-PAT (M~U[2,128], "
                       "MIXED_SPLIT); arrays 249/996/3984 uniform", "gspat_cus": gspat_cus,
           "field_results": "pinned client buffers" if pinned else "pageable client buffers", "runs": []}
    # result buffers from a ring, touched once beforehand, as a serving client
    # reuses them (a fresh np.empty per request would add its first-touch page
    # faults to the server's copy of every result); 256 slots cover the
    # requests in flight (tens at the highest rate)
    ring = 256
    if pinned:
        # page-locked buffers: the server DMAs field results into them directly
        import torch
        fbufs = torch.ones((ring, 2 * 1010 * 1000), dtype=torch.float32, pin_memory=True).numpy()
    else:
        fbufs = np.ones((ring, 2 * 1010 * 1000), np.float32)  # grids drawn below have <= 1.0033e6 points
    gbufs = [np.ones((ring, 2, a[0].shape[0]), np.float32) for a in arrs]
    for qps in qps_list:
        srv = Server(hbm_budget=4 << 30, pinned_budget=1 << 30, n_streams=4, max_gspat_batch=256,
                     gspat_precision=D.HOLO_MIXED_SPLIT, gspat_cus=gspat_cus)
        ids = [srv.add_array(1, src, k, a) for src, k, a, _ in arrs]
        # warm-up: every kernel variant loaded once, outside the measurement
        for aid, (src, k, a, ext) in enumerate(arrs):
            for cnt in ((16, 16, 1), (1024, 1024, 2)):
                ps = D.grid_pointset((0, 1, 2), cnt, (0.0, 0.0, 100.0), 1.0)
                srv.wait(srv.submit_field(ids[aid], ps, 2, np.empty(cnt[0] * cnt[1] * cnt[2] * 2, np.float32)))
            for m in (4, 40, 100):
                f = np.tile(np.array([[10.0, 10.0, 150.0]], np.float32), (m, 1)) + np.arange(m, dtype=np.float32)[:, None]
                srv.wait(srv.submit_gspat(ids[aid], f, np.ones(m, np.float32), np.empty(len(src), np.float32),
                                          np.empty(len(src), np.float32)))
        srv.stats(reset=True)
        rng = np.random.default_rng(int(qps))
        nreq = max(1, int(qps * seconds))
        arrivals = np.cumsum(rng.exponential(1.0 / qps, nreq))
        specs = []
        for _ in range(nreq):
            aid = int(rng.integers(0, 3))
            ext = arrs[aid][3]
            if rng.random() < 0.7:
                P = float(np.exp(rng.uniform(np.log(1e4), np.log(1e6))))
                nz = int(rng.integers(1, 17))
                nxy = max(1, int(round((P / nz) ** 0.5)))
                origin = (float(rng.uniform(0, ext[0])), float(rng.uniform(0, ext[1])), float(rng.uniform(50, 200)))
                specs.append(("f", aid, (nxy, nxy, nz), origin))
            else:
                m = int(rng.integers(2, 129))
                foci = np.stack([rng.uniform(0, ext[0], m), rng.uniform(0, ext[1], m), rng.uniform(100, 250, m)],
                                1).astype(np.float32)
                specs.append(("g", aid, foci, rng.uniform(0.5, 1.0, m).astype(np.float32)))
        q = queue.Queue()
        lat = []

        def waiter():
            while True:
                item = q.get()
                if item is None:
                    return
                lat.append(srv.wait(item))

        th = threading.Thread(target=waiter)
        th.start()
        terms = 0.0
        solves = 0
        t0 = time.perf_counter()
        for i, sp in enumerate(specs):
            delay = t0 + arrivals[i] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            if sp[0] == "f":
                _, aid, cnt, origin = sp
                ps = D.grid_pointset((0, 1, 2), cnt, origin, 1.0)
                buf = fbufs[i % ring, :cnt[0] * cnt[1] * cnt[2] * 2]
                q.put(srv.submit_field(ids[aid], ps, 2, buf))
                terms += cnt[0] * cnt[1] * cnt[2] * arrs[aid][0].shape[0]
            else:
                _, aid, foci, amps = sp
                gb = gbufs[aid][i % ring]
                q.put(srv.submit_gspat(ids[aid], foci, amps, gb[0], gb[1]))
                solves += 1
        q.put(None)
        th.join()
        elapsed = time.perf_counter() - t0
        p50, p99, done = srv.stats()
        fp50, fp99, _ = srv.stats(kind=0)
        gp50, gp99, _ = srv.stats(kind=1)
        srv.close()
        run = {"qps_offered": qps, "requests": nreq, "achieved_req_per_s": round(nreq / elapsed, 1),
               "p50_ms": round(p50, 3), "p99_ms": round(p99, 3), "field_p50_ms": round(fp50, 3),
               "field_p99_ms": round(fp99, 3), "gspat_p50_ms": round(gp50, 3), "gspat_p99_ms": round(gp99, 3),
               "field_gsrc_pt_per_s": round(terms / elapsed / 1e9, 2), "gspat_solves_per_s": round(solves / elapsed, 1)}
        out["runs"].append(run)
        log(f"serve qps {qps}: {run}")
    ok = [r for r in out["runs"] if r["p99_ms"] <= 5.0]
    out["max_qps_p99_le_5ms"] = max((r["qps_offered"] for r in ok), default=None)
    okg = [r for r in out["runs"] if r["gspat_p99_ms"] <= 5.0]
    out["max_qps_gspat_p99_le_5ms"] = max((r["qps_offered"] for r in okg), default=None)
    return out


def bench_serve_gspat(afc, D, qps_list, seconds,
                      precisions=(("mixed", 1, False), ("mixed_graphs", 1, True), ("fp32", 0, False)),
                      plain_calls=200, arrays=(0, 1, 2)):
    """Served GS-PAT alone: Poisson arrivals of GS-PAT requests (M ~ U[2, 128],
    foci over each array, targets U[0.5, 1]) on the 249-, 996- and
    3984-transducer arrays (those listed in `arrays`) chosen uniformly, at a
    ramp of offered rates until p99 passes 20 ms or the submitter falls behind
    (the highest rate with p99 <= 5 ms is reported; the requests that have
    arrived are submitted together with AFS_SubmitGSPATBatch, pass by pass);
    the same ramp with each
    bucket's kernel chain launched as a hipGraph, and with the server in FP32
    mode; and the plain path, AFO_GSPAT one request per call, on requests drawn
    from the same distribution."""
    from ctypes import byref, c_double, c_uint64, c_void_p
    from afc import workloads as W
    from afc.server import Server

    rng = np.random.default_rng(11)
    arrs = []
    for tiles in ((1, 1), (2, 2), (4, 4)):
        src, k, a = W.array_rows(tiles, phase_seed=249 * tiles[0] * tiles[1])
        arrs.append((src, k, a, src[:, 0:3].max(0)))
    # a pool of request specs, reused cyclically: array, foci, targets
    pool = []
    for _ in range(2048):
        aid = int(arrays[int(rng.integers(0, len(arrays)))])
        m = int(rng.integers(2, 129))
        ext = arrs[aid][3]
        foci = np.ascontiguousarray(np.stack([rng.uniform(0, ext[0], m), rng.uniform(0, ext[1], m),
                                              rng.uniform(100, 250, m)], 1).astype(np.float32))
        amps = np.ascontiguousarray(rng.uniform(0.5, 1.0, m).astype(np.float32))
        pool.append((aid, m, foci, amps))
    ring = 4096
    outs = [np.empty((ring, 2, a[0].shape[0]), np.float32) for a in arrs]
    dll = afc.dll()
    # the pool and the output ring as raw addresses
    pool_aid = np.array([p_[0] for p_ in pool], np.int64)
    pool_m = np.array([p_[1] for p_ in pool], np.uint32)
    pool_foci = np.array([p_[2].ctypes.data for p_ in pool], np.uint64)
    pool_amps = np.array([p_[3].ctypes.data for p_ in pool], np.uint64)
    out_ptr = np.array([[(o[i, 0].ctypes.data, o[i, 1].ctypes.data) for i in range(ring)] for o in outs], np.uint64)
    names = "/".join(("249", "996", "3984")[i] for i in arrays)
    out = {"workload": f"GS-PAT requests only: Poisson arrivals, M ~ U[2,128], arrays {names} uniform, "
                       "100 iterations; p99 over the requests of each rate", "runs": {}}
    for name, prec, graphs in precisions:
        runs = []
        for qps in qps_list:
            srv = Server(hbm_budget=2 << 30, pinned_budget=1 << 30, n_streams=2, max_gspat_batch=256,
                         gspat_precision=prec, gspat_cus=0, gspat_graphs=graphs)
            ids = [srv.add_array(1, src, k, a) for src, k, a, _ in arrs]
            for aid, (src, k, a, ext) in enumerate(arrs):
                for m in (4, 20, 40, 100):
                    f = pool[0][2][:1].repeat(m, 0) + np.arange(m, dtype=np.float32)[:, None]
                    srv.wait(srv.submit_gspat(ids[aid], f, np.ones(m, np.float32), np.empty(len(src), np.float32),
                                              np.empty(len(src), np.float32)))
            srv.stats(reset=True)
            nreq = max(1, int(qps * seconds))
            arrivals = np.cumsum(rng.exponential(1.0 / qps, nreq))
            # per request: array, foci, targets, output buffers as raw
            # addresses (AFS_SubmitGSPATBatch takes arrays of them)
            sel = np.arange(nreq) % len(pool)
            which = pool_aid[sel]
            a_id = np.array(ids, np.uint32)[which]
            a_m = pool_m[sel]
            a_foci = pool_foci[sel]
            a_amps = pool_amps[sel]
            slot = np.arange(nreq) % ring
            a_oa = np.ascontiguousarray(out_ptr[which, slot, 0])
            a_op = np.ascontiguousarray(out_ptr[which, slot, 1])
            tickets = np.zeros(nreq, np.uint64)
            base = [x.ctypes.data for x in (a_id, a_foci, a_amps, a_m, a_oa, a_op, tickets)]
            h = srv.handle
            submit = dll.AFS_SubmitGSPATBatch
            i = 0
            t0 = time.perf_counter()
            # every pass submits the requests that have arrived by now, in
            # one call: the Poisson schedule is kept to the pass time (a few
            # microseconds), and one Python call per request (which topped
            # out near 470K submissions/s) does not bound the offered rate
            while i < nreq:
                j = int(np.searchsorted(arrivals, time.perf_counter() - t0, side="right"))
                if j > i:
                    if submit(h, j - i, base[0] + 4 * i, base[1] + 8 * i, base[2] + 8 * i, base[3] + 4 * i,
                              base[4] + 8 * i, base[5] + 8 * i, base[6] + 8 * i) != 0:
                        raise RuntimeError("AFS_SubmitGSPATBatch failed")
                    i = j
            t_sub = time.perf_counter() - t0
            lat = c_double()
            for t in tickets.tolist():
                dll.AFS_Wait(h, t, byref(lat))
            elapsed = time.perf_counter() - t0
            p50, p99, done = srv.stats(kind=1)
            g_launch, g_build = srv.graph_counters()
            srv.close()
            run = {"qps_offered": qps, "requests": nreq, "submitted_per_s": round(nreq / t_sub, 1),
                   "served_solves_per_s": round(nreq / elapsed, 1), "p50_ms": round(p50, 3), "p99_ms": round(p99, 3)}
            if graphs:
                run["graph_launches"], run["graph_builds"] = g_launch, g_build
            runs.append(run)
            log(f"serve gspat {name} qps {qps}: {run}")
            # the ramp goes on past a rate whose p99 is above 5 ms (the tail is
            # not monotonic in the rate at this window length) and stops once
            # the server is clearly saturated or the submitter falls behind
            if p99 > 20.0 or nreq / t_sub < 0.9 * qps:
                break
        ok = [r for r in runs if r["p99_ms"] <= 5.0]
        out["runs"][name] = runs
        out[f"{name}_max_solves_per_s_p99_le_5ms"] = max((r["served_solves_per_s"] for r in ok), default=None)
    # the plain path: AFO_GSPAT, one request per call, the same request distribution
    systems = []
    for src, k, a, _ in arrs:
        s = afc.UniformSystem(W.TEMP_340)
        for row in src:
            s.add_wave_source(afc.T4010A1(row[0:3], row[3:6], row[6], row[7], row[8]))
        systems.append(s)
    lat = []
    for i in range(plain_calls):
        aid, m, foci, amps = pool[i]
        sa = systems[aid].source_array()
        sa[:, 6] = 1.0
        t0 = time.perf_counter()
        afc.Optimizer.gs_pat(systems[aid], foci, amps)
        lat.append(time.perf_counter() - t0)
    out["plain_afo_gspat"] = {"calls": plain_calls, "solves_per_s": round(plain_calls / sum(lat), 1),
                              "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
                              "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3)}
    best = out.get("mixed_max_solves_per_s_p99_le_5ms")
    if best:
        out["served_mixed_vs_plain_afo_gspat"] = round(best / out["plain_afo_gspat"]["solves_per_s"], 2)
        if out.get("fp32_max_solves_per_s_p99_le_5ms"):
            out["served_mixed_vs_served_fp32"] = round(best / out["fp32_max_solves_per_s_p99_le_5ms"], 2)
    log(f"serve gspat: {json.dumps({k: v for k, v in out.items() if k != 'runs'})}")
    return out


def bench_greedy(torch, D, O):
    """Batched greedy phase search (combination/greedy.rs:23-83): 256
    instances on the 996-transducer array, M ~ U[2,16] foci drawn as in the
    GS-PAT workload, 16 phase candidates (the reference's default)."""
    from afc import workloads as W

    dev = torch.device("cuda:0")
    src, k, a = W.array_rows((2, 2))
    n = src.shape[0]
    foci, amps, offs = W.focus_batch(256, m_range=(2, 16), seed=8)
    packed = D.pack_sources(1, src, k, a)
    src_d = torch.from_numpy(packed).to(dev)
    foci_d = torch.from_numpy(foci).to(dev)
    amps_d = torch.from_numpy(amps).to(dev)
    out = torch.empty((256, n), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream(dev)
    ctx = D.Context()

    def run():
        D.greedy_solve(ctx, src_d.data_ptr(), packed.shape[0], n, 1, foci_d.data_ptr(), amps_d.data_ptr(), offs, 16,
                       out.data_ptr(), stream.cuda_stream)
    run()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(3)]
    for e0, e1 in ev:
        e0.record(stream)
        run()
        e1.record(stream)
    torch.cuda.synchronize()
    ms = float(np.median([e0.elapsed_time(e1) for e0, e1 in ev]))
    # objective of instance 0 against the FP64 oracle's run
    ph = out.cpu().numpy()[0].astype(np.float64)
    f = foci[offs[0]:offs[1]]
    am = amps[offs[0]:offs[1]]
    ref = O.greedy(1, src, k, a, f, am, 16, "f64")

    def objective(phase):
        s1 = src.copy()
        s1[:, 7] = phase
        return float(np.sum(np.abs(am - np.abs(O.field(1, s1, k, a, f, "f64")))))
    ctx.close()
    log(f"greedy: {ms:.3f} ms per batch of 256, {256 / (ms * 1e-3):.0f} solves/s")
    return {"workload": "Greedy B=256, N=996, M~U[2,16], 16 phase candidates", "solves_per_s": round(256 / (ms * 1e-3), 1),
            "batch_ms": round(ms, 3), "objective_instance0": objective(ph), "objective_instance0_fp64_oracle": objective(ref)}


def bench_lm(torch, D, O, cpu=True):
    """Batched FP64 LM (lm.rs:40-105): 64 instances on the 249- and on the
    996-transducer array, 16 foci each (as the GS-PAT workload), targets 1,
    50 iterations."""
    from afc import workloads as W

    out = {}
    dev = torch.device("cuda:0")
    for name, tiles in (("n249", (1, 1)), ("n996", (2, 2))):
        src, k, a = W.array_rows(tiles)
        n = src.shape[0]
        foci, amps, offs = W.focus_batch(64, m_range=(16, 16), seed=4)
        amps = np.ones_like(amps)
        packed = D.pack_sources(1, src, k, a)
        src_d = torch.from_numpy(packed).to(dev)
        foci_d = torch.from_numpy(foci).to(dev)
        amps_d = torch.from_numpy(amps).to(dev)
        ph_d = torch.empty((64, n), dtype=torch.float32, device=dev)
        it_d = torch.zeros(64, dtype=torch.int32, device=dev)
        stream = torch.cuda.current_stream(dev)
        ctx = D.Context()

        def run():
            D.lm_solve(ctx, src_d.data_ptr(), packed.shape[0], n, 1, foci_d.data_ptr(), amps_d.data_ptr(), offs, 50,
                       ph_d.data_ptr(), it_d.data_ptr(), stream.cuda_stream)
        run()
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(3)]
        for e0, e1 in ev:
            e0.record(stream)
            run()
            e1.record(stream)
        torch.cuda.synchronize()
        ms = float(np.median([e0.elapsed_time(e1) for e0, e1 in ev]))
        entry = {"solves_per_s": round(64 / (ms * 1e-3), 1), "batch_ms": round(ms, 3), "iterations": 50,
                 "n_params": n + 16}
        if name == "n249":
            # the reference iteration is chaotic beyond ~40 steps (a one-ulp input
            # change reaches O(1) phase differences by 200): agreement with the
            # FP64 oracle is checked on 20-iteration trajectories
            sub = np.array([0, 16, 32], np.uint32)
            D.lm_solve(ctx, src_d.data_ptr(), packed.shape[0], n, 1, foci_d.data_ptr(), amps_d.data_ptr(), sub, 20,
                       ph_d.data_ptr(), it_d.data_ptr(), stream.cuda_stream)
            torch.cuda.synchronize()
            ph = ph_d.cpu().numpy()
            errs, dph = [], []
            for b in range(2):
                f = foci[offs[b]:offs[b + 1]]
                ref, _, _ = O.lm(1, src, k, a, f, amps[offs[b]:offs[b + 1]], "f64", k_max=20)
                dph.append(float(np.max(np.abs(np.angle(np.exp(1j * (ph[b].astype(np.float64) - ref)))))))
                s1, s2 = src.copy(), src.copy()
                s1[:, 7], s2[:, 7] = ph[b], ref
                g1 = np.abs(O.field(1, s1, k, a, f, "f64"))
                g2 = np.abs(O.field(1, s2, k, a, f, "f64"))
                errs.append(float(np.max(np.abs(g1 - g2) / g2)))
            # phases are compared after the FP32 rounding the output format applies
            entry["focal_amp_rel_err_vs_fp64_20it_max"] = max(errs)
            entry["phase_max_abs_diff_vs_fp64_20it_rad"] = max(dph)
        if name == "n249" and cpu:
            entry["cpu_baseline"] = cpu_baseline_instances(
                lambda b: O.lm(1, src, k, a, foci[offs[b]:offs[b + 1]], amps[offs[b]:offs[b + 1]], "f32", k_max=50),
                64, "the 64 instances, 50 iterations, dense f32 QR as lm.rs:72")
            log(f"lm n249 cpu baseline: {entry['cpu_baseline']['value']} solves/s")
        if name == "n249":
            # the other two solvers of nls/ on the same machinery and workload
            for algo, key, kmax in ((D.NLS_GAUSS_NEWTON, "gauss_newton_n249", 200),
                                    (D.NLS_GRADIENT_DESCENT, "gradient_descent_n249", 10000)):
                def run_nls():
                    D.nl_solve(ctx, src_d.data_ptr(), packed.shape[0], n, 1, foci_d.data_ptr(), amps_d.data_ptr(),
                               offs, algo, kmax, ph_d.data_ptr(), it_d.data_ptr(), stream.cuda_stream)
                run_nls()
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                run_nls()
                e1.record(stream)
                torch.cuda.synchronize()
                nms = e0.elapsed_time(e1)
                its = it_d.cpu().numpy()
                out[key] = {"solves_per_s": round(64 / (nms * 1e-3), 1), "batch_ms": round(nms, 3),
                            "k_max": kmax, "iterations_mean": float(its.mean())}
                log(f"{key}: {nms:.3f} ms per batch of 64")
            # APO (apo.rs:100-158), lambda 0.1, targets 0.2 (within reach of the 249 array)
            oa_d = torch.empty((64, n), dtype=torch.float32, device=dev)
            amps_apo = torch.full_like(amps_d, 0.2)

            def run_apo():
                D.apo_solve(ctx, src_d.data_ptr(), packed.shape[0], n, 1, foci_d.data_ptr(), amps_apo.data_ptr(), offs,
                            0.1, 200, oa_d.data_ptr(), ph_d.data_ptr(), it_d.data_ptr(), stream.cuda_stream)
            run_apo()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            run_apo()
            e1.record(stream)
            torch.cuda.synchronize()
            ams = e0.elapsed_time(e1)
            out["apo_n249"] = {"solves_per_s": round(64 / (ams * 1e-3), 1), "batch_ms": round(ams, 3), "k_max": 200,
                               "iterations_mean": float(it_d.cpu().numpy().mean()), "lambda": 0.1}
            log(f"apo_n249: {ams:.3f} ms per batch of 64")
            # Long (long.rs:46-106) and Horn (horn.rs:58-129, REPEAT 100,
            # lambda 0.8, Tikhonov 1e-5) on the same 64 instances
            for key, fn in (("long_n249", lambda: D.long_solve(
                    ctx, src_d.data_ptr(), packed.shape[0], n, 1, foci_d.data_ptr(), amps_d.data_ptr(), offs,
                    oa_d.data_ptr(), ph_d.data_ptr(), stream.cuda_stream)),
                            ("horn_n249", lambda: D.horn_solve(
                    ctx, src_d.data_ptr(), packed.shape[0], n, 1, foci_d.data_ptr(), amps_d.data_ptr(), offs,
                    oa_d.data_ptr(), ph_d.data_ptr(), stream_ptr=stream.cuda_stream))):
                fn()
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                fn()
                e1.record(stream)
                torch.cuda.synchronize()
                xms = e0.elapsed_time(e1)
                out[key] = {"solves_per_s": round(64 / (xms * 1e-3), 1), "batch_ms": round(xms, 3)}
                log(f"{key}: {xms:.3f} ms per batch of 64")
        ctx.close()
        out[name] = entry
        log(f"lm {name}: {ms:.3f} ms per batch of 64, {64 / (ms * 1e-3):.0f} solves/s")
    return out


def bench_capi_focus_plane_249(afc, O, cpu=True):
    """The 249-source focus plane (100x100 points at z = 150 mm) through the reference's own entry point, AFC_Calculate (one
    synchronous call: the sources packed and uploaded, the field computed on the
    GPU, |p| copied back), beside the CPU restatement of the reference's
    calculator on the same volume."""
    from afc import workloads as W

    src, k, a, area = W.focus_plane_249()
    system = afc.UniformSystem(W.TEMP_340)
    for row in src:
        system.add_wave_source(afc.T4010A1(row[0:3], row[3:6], row[6], row[7], row[8]))
    grid = afc.GridAreaBuilder().x_range((40, 139)).y_range((20, 119)).z_at(150).resolution(1.0).generate()
    field = afc.PressureFieldBuffer()
    calc = afc.GpuCalculator()
    for _ in range(5):
        got = calc.calculate(system, grid, field)
    lat = []
    for _ in range(100):
        t0 = time.perf_counter()
        got = calc.calculate(system, grid, field)
        lat.append(time.perf_counter() - t0)
    pts, _ = O.grid_points(area["order"], area["ranges"], area["res"])
    terms = pts.shape[0] * src.shape[0]
    ref = np.abs(O.field(1, src, k, a, pts, "f64"))
    p50 = float(np.percentile(lat, 50))
    out = {"workload": "249 T4010A1 x 100x100 points at z = 150 mm, |p| (PressureField), AFC_Calculate",
           "p50_ms": round(p50 * 1e3, 4), "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 4),
           "gsrc_pt_per_s": round(terms / p50 / 1e9, 3),
           "rel_l2_vs_fp64": float(np.linalg.norm(got - ref) / np.linalg.norm(ref))}
    if cpu:
        threads = cpu_threads()
        runs = []
        for th in (threads, 1):
            t0 = time.perf_counter()
            O.field(1, src, k, a, pts, "f32", threads=th)
            runs.append(time.perf_counter() - t0)
        out["cpu_baseline"] = {"ms": round(runs[0] * 1e3, 3), "ms_single_thread": round(runs[1] * 1e3, 3),
                               "cores": threads, "kind": "port",
                               "sample": "the whole volume (2.49e6 terms), FP32 restatement of the reference's "
                                         "CpuCalculator"}
    log(f"AFC_Calculate, 249 sources x 100x100 points: p50 {out['p50_ms']:.3f} ms; cpu {out.get('cpu_baseline', {}).get('ms')} ms")
    return out


def bench_gspat(torch, afc, D, O, src, k, a, batch, cpu=True):
    """Batched GS-PAT (gs_pat.rs:38-81) on device-resident inputs."""
    from afc import workloads as W

    src = src.copy()
    src[:, 7] = 0.0  # GS-PAT zeroes the phases (gs_pat.rs:39-41)
    n = src.shape[0]
    foci, amps, offs = W.focus_batch(batch, m_range=(2, 64), seed=3)
    dev = torch.device("cuda:0")
    packed = D.pack_sources(1, src, k, a)
    src_d = torch.from_numpy(packed).to(dev)
    foci_d = torch.from_numpy(foci).to(dev)
    amps_d = torch.from_numpy(amps).to(dev)
    out_a = torch.empty((batch, n), dtype=torch.float32, device=dev)
    out_p = torch.empty((batch, n), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream(dev)
    ctx = D.Context()
    rng = np.random.default_rng(1)
    sample = rng.choice(batch, 6, replace=False)
    ms_all = np.diff(offs.astype(np.int64))
    # focal amplitudes of the oracle's FP64 GS-PAT on every instance of the
    # batch (one instance per host thread), the yardstick of every mode
    from concurrent.futures import ThreadPoolExecutor

    def ref_one(b):
        f, am = foci[offs[b]:offs[b + 1]], amps[offs[b]:offs[b + 1]]
        return O.gspat(1, src, k, a, f, am, "f64")

    t_ref = time.perf_counter()
    with ThreadPoolExecutor(max_workers=cpu_threads()) as ex:
        sols = list(ex.map(ref_one, range(batch)))
    ref = {}
    for b in range(batch):
        s2 = src.copy()
        s2[:, 6], s2[:, 7] = sols[b]
        ref[b] = np.abs(O.field(1, s2, k, a, foci[offs[b]:offs[b + 1]], "f64"))
    log(f"gspat FP64 references of all {batch} instances: {time.perf_counter() - t_ref:.1f} s")
    modes = {"fp32": D.HOLO_FP32, "mixed": D.HOLO_MIXED, "mixed_split": D.HOLO_MIXED_SPLIT,
             "mixed_closing": D.HOLO_MIXED_CLOSING}
    out = {"workload": f"GS-PAT B={batch}, N={n}, M~U[2,64] (sum M={int(offs[-1])}), 100 iterations"}
    for name, prec in modes.items():
        def run():
            D.holo_solve(ctx, src_d.data_ptr(), packed.shape[0], n, 1, foci_d.data_ptr(), amps_d.data_ptr(), offs,
                         D.ALGO_GSPAT, prec, 100, out_a.data_ptr(), out_p.data_ptr(), stream.cuda_stream)
        run()
        torch.cuda.synchronize()
        reps = 5
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
        for e0, e1 in ev:
            e0.record(stream)
            run()
            e1.record(stream)
        torch.cuda.synchronize()
        ms = float(np.median([e0.elapsed_time(e1) for e0, e1 in ev]))
        am = out_a.cpu().numpy()
        ph = out_p.cpu().numpy()
        errs = []
        for b in range(batch):
            f = foci[offs[b]:offs[b + 1]]
            s2 = src.copy()
            s2[:, 6], s2[:, 7] = am[b], ph[b]
            got = np.abs(O.field(1, s2, k, a, f, "f64"))
            errs.append(float(np.max(np.abs(got - ref[b]) / ref[b])))
        errs = np.array(errs)
        worst = int(np.argmax(errs))
        out[name] = {"solves_per_s": round(batch / (ms * 1e-3), 1), "batch_ms": round(ms, 4),
                     "focal_amp_rel_err_max": float(errs[worst]), "focal_amp_rel_err_max_m": int(ms_all[worst]),
                     "focal_amp_rel_err_p99": float(np.percentile(errs, 99)),
                     "focal_amp_rel_err_median": float(np.median(errs)), "instances_checked": batch,
                     "instances_over_1e-3": int(np.sum(errs > 1e-3))}
        log(f"gspat {name}: {ms:.3f} ms per batch of {batch}, {batch / (ms * 1e-3):.0f} solves/s, "
            f"focal-amplitude rel err vs FP64 over all {batch} instances: max {errs[worst]:.2e} "
            f"(M={ms_all[worst]}), median {np.median(errs):.2e}, {int(np.sum(errs > 1e-3))} above 1e-3")
    # GS (gs.rs:36-77, 100 iterations) and Naive (naive.rs:32-51) on the same batch, FP32, against the
    # oracle's FP64 runs of the sampled instances
    for name, algo, ref_fn in (("gs", D.ALGO_GS, lambda f, am: O.gs(1, src, k, a, f, am, "f64")),
                               ("naive", D.ALGO_NAIVE, lambda f, am: O.naive(1, src, k, a, f, am, "f64"))):
        def run():
            D.holo_solve(ctx, src_d.data_ptr(), packed.shape[0], n, 1, foci_d.data_ptr(), amps_d.data_ptr(), offs,
                         algo, D.HOLO_FP32, 100, out_a.data_ptr(), out_p.data_ptr(), stream.cuda_stream)
        run()
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(3)]
        for e0, e1 in ev:
            e0.record(stream)
            run()
            e1.record(stream)
        torch.cuda.synchronize()
        ms = float(np.median([e0.elapsed_time(e1) for e0, e1 in ev]))
        am, ph = out_a.cpu().numpy(), out_p.cpu().numpy()
        errs = []
        for b in sample:
            f, amp_t = foci[offs[b]:offs[b + 1]], amps[offs[b]:offs[b + 1]]
            a64, p64 = ref_fn(f, amp_t)
            s_ref, s_got = src.copy(), src.copy()
            s_ref[:, 6], s_ref[:, 7] = a64, p64
            s_got[:, 6], s_got[:, 7] = am[b], ph[b]
            want = np.abs(O.field(1, s_ref, k, a, f, "f64"))
            got = np.abs(O.field(1, s_got, k, a, f, "f64"))
            errs.append(float(np.max(np.abs(got - want) / want)))
        out[name] = {"solves_per_s": round(batch / (ms * 1e-3), 1), "batch_ms": round(ms, 4),
                     "focal_amp_rel_err_max_vs_fp64_run": max(errs), "sample_instances": int(len(sample))}
        log(f"{name}: {ms:.3f} ms per batch of {batch}, {batch / (ms * 1e-3):.0f} solves/s, "
            f"max focal-amplitude rel err vs the FP64 run {max(errs):.2e}")
    ctx.close()
    # the plain FP32 path: AFO_GSPAT, one instance per call through the C ABI
    system = afc.UniformSystem(W.TEMP_340)
    for row in src:
        system.add_wave_source(afc.T4010A1(row[0:3], row[3:6], row[6], row[7], row[8]))
    import time as _t
    reps = 200
    lat = []
    for b in range(reps):
        f, am = foci[offs[b]:offs[b + 1]], amps[offs[b]:offs[b + 1]]
        srcs = system.source_array()
        srcs[:, 6] = 1.0
        t0 = _t.perf_counter()
        afc.Optimizer.gs_pat(system, f, am)
        lat.append(_t.perf_counter() - t0)
    out["plain_fp32_unbatched"] = {"solves_per_s": round(reps / sum(lat), 1),
                                   "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
                                   "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3)}
    out["speedup_mixed_split_vs_plain"] = round(out["mixed_split"]["solves_per_s"] /
                                                out["plain_fp32_unbatched"]["solves_per_s"], 1)
    log(f"gspat plain FP32 unbatched: {reps / sum(lat):.0f} solves/s")
    if cpu:
        count = min(256, batch)
        out["cpu_baseline"] = cpu_baseline_instances(
            lambda b: O.gspat(1, src, k, a, foci[offs[b]:offs[b + 1]], amps[offs[b]:offs[b + 1]], "f32"), count,
            f"the first {count} instances of the batch")
        log(f"gspat cpu baseline: {out['cpu_baseline']['value']} solves/s")
    return out


if __name__ == "__main__":
    main()
