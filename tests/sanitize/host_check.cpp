// Host-side memory and undefined-behaviour check of libafc_hip's C ABI.
//
// Built by tests/sanitize/build.py together with the library's sources, whose
// host code is compiled with AddressSanitizer and UndefinedBehaviorSanitizer
// (device code is not instrumented). A sanitizer report ends the process
// with a nonzero status; the driver's own checks print "FAIL" and exit 1.
//
//   host_check host   the calls that need no GPU: handles, grids, info
//                     strings, source packing, FocalPoint / BesselBeam,
//                     directivity, and the argument checks of every entry
//   host_check all    the above, then the GPU paths: AFC_Calculate in every
//                     tag combination, the AFO_* optimizers, batched solves
//                     on a context and without allocation, and the request
//                     server under concurrent submitters and waiters (with
//                     and without graph launches, with a CU partition, and a
//                     budget small enough that admission waits)
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <random>
#include <thread>
#include <vector>

#include "afc_hip.h"

static int g_fail = 0;

#define CHECK(cond)                                                              \
    do {                                                                         \
        if (!(cond)) {                                                           \
            fprintf(stderr, "FAIL %s:%d: %s (%s)\n", __FILE__, __LINE__, #cond, \
                    AFC_GetLastError());                                         \
            g_fail = 1;                                                          \
        }                                                                        \
    } while (0)

// The 18 x 14 array at 10.16 mm pitch with three sites left empty (249
// transducers), tiled tx x ty times.
static std::vector<AFC_T4010A1> tiled_array(int tx, int ty, float phase_seed) {
    std::vector<AFC_T4010A1> v;
    std::mt19937 rng((unsigned)(phase_seed * 1000));
    std::uniform_real_distribution<float> ph(0.f, 6.2831853f);
    for (int a = 0; a < ty; ++a)
        for (int b = 0; b < tx; ++b)
            for (int y = 0; y < 14; ++y)
                for (int x = 0; x < 18; ++x) {
                    if (y == 1 && (x == 1 || x == 14 || x == 16)) continue;
                    AFC_T4010A1 s;
                    s.x = 10.16f * (float)(18 * b + x);
                    s.y = 10.16f * (float)(14 * a + y);
                    s.z = 0.f;
                    s.nx = 0.f;
                    s.ny = 0.f;
                    s.nz = 1.f;
                    s.amp = 1.f;
                    s.phase = ph(rng);
                    s.freq = 40e3f;
                    v.push_back(s);
                }
    return v;
}

static void* make_system(const std::vector<AFC_T4010A1>& src) {
    void* sys = nullptr;
    AFC_CreateUniformSystem(&sys, 293.15f, AFC_SOURCE_T4010A1);
    for (const auto& s : src) AFC_AddT4010A1(sys, s);
    return sys;
}

static bool all_finite(const float* p, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!isfinite(p[i])) return false;
    return true;
}

// ------------------------------------------------------------------ host part
static void host_part() {
    // systems of both source types, info strings, in-place edits
    void* sph = nullptr;
    AFC_CreateUniformSystem(&sph, 300.f, AFC_SOURCE_SPHERE);
    CHECK(sph != nullptr);
    for (int i = 0; i < 40; ++i) {
        AFC_SphereWaveSource s = {(float)i, 2.f * i, 0.f, 1.f, 0.1f * i, 40e3f};
        AFC_AddSphereWaveSource(sph, s);
    }
    CHECK(AFC_UniformSystemSoundSpeed(sph, AFC_SOURCE_SPHERE) > 3e5f);
    for (int r = 0; r < 3; ++r) CHECK(strlen(AFC_UniformSystemInfo(sph, AFC_SOURCE_SPHERE)) > 0);
    CHECK(strlen(AFC_UniformSystemSourceInfo(sph, 39, AFC_SOURCE_SPHERE)) > 0);
    (void)AFC_UniformSystemSourceInfo(sph, 40, AFC_SOURCE_SPHERE);  // out of range: an error, not a read
    void* view = nullptr;
    CHECK(AFC_GetWaveSources(sph, &view, AFC_SOURCE_SPHERE) == 40);
    static_cast<AFC_SphereWaveSource*>(view)[39].phase = 1.f;
    std::vector<float> k(40), a(40);
    CHECK(AFH_GetSourceConstants(sph, k.data(), a.data()) == 40);
    CHECK(AFH_GetSourceConstants(sph, nullptr, nullptr) == 40);
    AFC_FreeUniformSystem(sph, AFC_SOURCE_SPHERE);

    const std::vector<AFC_T4010A1> arr = tiled_array(1, 1, 1.f);
    void* sys = make_system(arr);
    AFC_Vector3 fp = {90.f, 70.f, 150.f}, dir = {0.f, 0.f, 1.f};
    AFO_FocalPoint(sys, fp, AFC_SOURCE_T4010A1);
    AFO_BesselBeam(sys, fp, dir, 0.1f, AFC_SOURCE_T4010A1);
    AFO_IFFT(sys, "missing.png", fp, fp, fp, 1.f, 100.f, AFC_SOURCE_T4010A1);
    CHECK(strlen(AFC_GetLastError()) > 0);
    void* tv = nullptr;
    const uint64_t n = AFC_GetWaveSources(sys, &tv, AFC_SOURCE_T4010A1);
    CHECK(n == arr.size());

    // packing: exactly AFH_PackedCount(n) entries written
    std::vector<float> kk(n), aa(n);
    AFH_GetSourceConstants(sys, kk.data(), aa.data());
    const uint64_t np = AFH_PackedCount(n);
    CHECK(np % AFH_SOURCE_PAD == 0 && np >= n);
    std::vector<float> packed(np * AFH_SOURCE_STRIDE);
    CHECK(AFH_PackSources(tv, kk.data(), aa.data(), n, AFC_SOURCE_T4010A1, packed.data()) == AFH_OK);
    CHECK(AFH_PackSources(tv, kk.data(), aa.data(), n, 7, packed.data()) == AFH_ERR_ARG);
    CHECK(AFH_PackSources(nullptr, kk.data(), aa.data(), n, AFC_SOURCE_T4010A1, packed.data()) == AFH_ERR_ARG);

    // grids of every dimension and axis order, scatter areas, fields
    const int32_t orders[6][3] = {{1, 2, 3}, {1, 3, 2}, {2, 1, 3}, {2, 3, 1}, {3, 1, 2}, {3, 2, 1}};
    for (const auto& o : orders) {
        void* g1 = nullptr;
        void* g2 = nullptr;
        void* g3 = nullptr;
        AFC_CreateGridArea1D(&g1, 0.f, 10.f, 1.f, 2.f, 0.5f, o[0], o[1], o[2]);
        AFC_CreateGridArea2D(&g2, 0.f, 10.f, -5.f, 5.f, 3.f, 1.f, o[0], o[1], o[2]);
        AFC_CreateGridArea3D(&g3, 0.f, 3.f, 0.f, 4.f, 0.f, 5.f, 1.f, o[0], o[1], o[2]);
        uint32_t c[3];
        AFC_GetGridSize(g1, AFC_AREA_GRID1D, &c[0], &c[1], &c[2]);
        CHECK(c[0] * c[1] * c[2] == 21);
        AFC_GetGridSize(g2, AFC_AREA_GRID2D, &c[0], &c[1], &c[2]);
        CHECK(c[0] * c[1] * c[2] == 121);
        AFC_GetGridSize(g3, AFC_AREA_GRID3D, &c[0], &c[1], &c[2]);
        CHECK(c[0] * c[1] * c[2] == 120);
        AFC_FreeGridArea(g1, AFC_AREA_GRID1D);
        AFC_FreeGridArea(g2, AFC_AREA_GRID2D);
        AFC_FreeGridArea(g3, AFC_AREA_GRID3D);
    }
    void* sc = nullptr;
    AFC_CreateScatterArea(&sc);
    for (int i = 0; i < 1000; ++i) AFC_ScatterAddObservePoint(sc, AFC_Vector3{(float)i, 1.f, 100.f});
    AFC_FreeScatterArea(sc);
    void* f = nullptr;
    AFC_CreatePressureField(&f);
    AFC_FreePressureField(f);
    AFC_CreatePowerField(&f);
    AFC_FreePowerField(f);
    AFC_CreateComplexPressureField(&f);
    AFC_FreeComplexPressureField(f);
    for (int t = 0; t <= 3; ++t) {
        void* c = nullptr;
        AFC_CreateCalculator(&c, t);
        CHECK(c != nullptr);
        AFC_FreeCalculator(c, t);
    }
    void* bad = reinterpret_cast<void*>(1);
    AFC_CreateCalculator(&bad, 9);
    CHECK(bad == nullptr);

    // directivity in both modes over the whole cos range and beyond
    std::vector<float> cs(4001), d(4001);
    for (int i = 0; i <= 4000; ++i) cs[i] = -1.5f + 3.0f * (float)i / 4000.f;
    CHECK(AFH_DirectivityT4010A1(cs.data(), cs.size(), AFH_PREC_FAST, d.data()) == AFH_OK);
    CHECK(AFH_DirectivityT4010A1(cs.data(), cs.size(), AFH_PREC_ACCURATE, d.data()) == AFH_OK);
    CHECK(AFH_DirectivityT4010A1(cs.data(), cs.size(), 5, d.data()) == AFH_ERR_ARG);

    // argument checks that return before any device work
    CHECK(AFH_HoloSolve(nullptr, nullptr, 0, 0, 0, nullptr, nullptr, nullptr, 1, 0, 0, 100, nullptr, nullptr,
                        nullptr) != AFH_OK);
    CHECK(AFS_Create(nullptr, nullptr) == AFH_ERR_ARG);
    AFS_Config bad_cfg = {1 << 20, 1 << 20, 0, 1, 0, 100, 0, 0};
    void* srv = nullptr;
    CHECK(AFS_Create(&bad_cfg, &srv) == AFH_ERR_ARG);
    bad_cfg.n_streams = 2;
    bad_cfg.gspat_graphs = 2;
    CHECK(AFS_Create(&bad_cfg, &srv) == AFH_ERR_ARG);
    uint64_t l = 0, b = 0;
    CHECK(AFS_GraphCounters(nullptr, &l, &b) == AFH_ERR_ARG);
    AFC_FreeUniformSystem(sys, AFC_SOURCE_T4010A1);
}

// ------------------------------------------------------------------- GPU part
static void capi_part() {
    const std::vector<AFC_T4010A1> arr = tiled_array(1, 1, 2.f);
    void* sys = make_system(arr);
    void* grid = nullptr;
    AFC_CreateGridArea2D(&grid, 40.f, 139.f, 20.f, 119.f, 150.f, 1.f, AFC_AXIS_X, AFC_AXIS_Y, AFC_AXIS_Z);
    void* sc = nullptr;
    AFC_CreateScatterArea(&sc);
    for (int i = 0; i < 777; ++i) AFC_ScatterAddObservePoint(sc, AFC_Vector3{0.1f * i, 50.f, 120.f});
    for (int ct = 0; ct <= 3; ++ct) {
        void* calc = nullptr;
        AFC_CreateCalculator(&calc, ct);
        for (int ft = 0; ft <= 2; ++ft) {
            void* field = nullptr;
            if (ft == 0) AFC_CreatePressureField(&field);
            if (ft == 1) AFC_CreatePowerField(&field);
            if (ft == 2) AFC_CreateComplexPressureField(&field);
            for (int rep = 0; rep < 2; ++rep) {
                const void* out = nullptr;
                uint64_t n = AFC_Calculate(calc, sys, grid, field, &out, ct, AFC_SOURCE_T4010A1, AFC_AREA_GRID2D, ft);
                CHECK(n == 10000 && out && all_finite(static_cast<const float*>(out), n * (ft == 2 ? 2 : 1)));
                n = AFC_Calculate(calc, sys, sc, field, &out, ct, AFC_SOURCE_T4010A1, AFC_AREA_SCATTER, ft);
                CHECK(n == 777 && out && all_finite(static_cast<const float*>(out), n * (ft == 2 ? 2 : 1)));
            }
            if (ft == 0) AFC_FreePressureField(field);
            if (ft == 1) AFC_FreePowerField(field);
            if (ft == 2) AFC_FreeComplexPressureField(field);
        }
        AFC_FreeCalculator(calc, ct);
    }
    AFC_FreeScatterArea(sc);
    AFC_FreeGridArea(grid, AFC_AREA_GRID2D);

    // every optimizer of the capi on small focus sets
    std::vector<float> foci, amps;
    for (int i = 0; i < 6; ++i) {
        foci.insert(foci.end(), {40.f + 15.f * i, 60.f + 5.f * i, 150.f});
        amps.push_back(1.f);
    }
    void* view = nullptr;
    AFO_GSPAT(sys, foci.data(), amps.data(), 6, AFC_SOURCE_T4010A1);
    AFO_GS(sys, foci.data(), amps.data(), 6, AFC_SOURCE_T4010A1);
    AFO_Naive(sys, foci.data(), amps.data(), 6, AFC_SOURCE_T4010A1);
    AFO_LM(sys, foci.data(), amps.data(), 4, AFC_SOURCE_T4010A1);
    AFO_Greedy(sys, foci.data(), amps.data(), 6, 16, AFC_SOURCE_T4010A1);
    AFO_GaussNewton(sys, foci.data(), amps.data(), 3, AFC_SOURCE_T4010A1);
    AFO_GradientDescent(sys, foci.data(), amps.data(), 2, AFC_SOURCE_T4010A1);
    AFO_APO(sys, foci.data(), amps.data(), 3, 0.1f, AFC_SOURCE_T4010A1);
    AFO_Horn(sys, foci.data(), amps.data(), 4, AFC_SOURCE_T4010A1);
    AFO_Long(sys, foci.data(), amps.data(), 4, AFC_SOURCE_T4010A1);
    const uint64_t n = AFC_GetWaveSources(sys, &view, AFC_SOURCE_T4010A1);
    CHECK(all_finite(static_cast<const float*>(view), n * 9));
    // too many foci: an error, nothing written
    std::vector<float> many(3 * 40, 100.f), many_a(40, 1.f);
    AFO_LM(sys, many.data(), many_a.data(), 40, AFC_SOURCE_T4010A1);
    CHECK(strlen(AFC_GetLastError()) > 0);
    AFC_FreeUniformSystem(sys, AFC_SOURCE_T4010A1);
}

static void batched_part() {
    const std::vector<AFC_T4010A1> arr = tiled_array(2, 2, 3.f);
    void* sys = make_system(arr);
    void* tv = nullptr;
    const uint64_t n = AFC_GetWaveSources(sys, &tv, AFC_SOURCE_T4010A1);
    std::vector<float> k(n), a(n);
    AFH_GetSourceConstants(sys, k.data(), a.data());
    const uint64_t np = AFH_PackedCount(n);
    std::vector<float> packed(np * AFH_SOURCE_STRIDE);
    AFH_PackSources(tv, k.data(), a.data(), n, AFC_SOURCE_T4010A1, packed.data());
    // 12 instances, M over every bucket
    const uint32_t ms[12] = {2, 16, 17, 32, 33, 64, 65, 128, 5, 40, 90, 1};
    std::vector<uint32_t> off(13, 0);
    std::vector<float> foci, amps;
    std::mt19937 rng(5);
    std::uniform_real_distribution<float> u(0.f, 1.f);
    for (int b = 0; b < 12; ++b) {
        off[b + 1] = off[b] + ms[b];
        for (uint32_t i = 0; i < ms[b]; ++i) {
            foci.insert(foci.end(), {360.f * u(rng), 280.f * u(rng), 100.f + 150.f * u(rng)});
            amps.push_back(0.5f + 0.5f * u(rng));
        }
    }
    float *d_src = nullptr, *d_foci = nullptr, *d_amps = nullptr, *d_oa = nullptr, *d_op = nullptr, *d_ref = nullptr;
    CHECK(hipMalloc(&d_src, packed.size() * 4) == hipSuccess);
    CHECK(hipMalloc(&d_foci, foci.size() * 4) == hipSuccess);
    CHECK(hipMalloc(&d_amps, amps.size() * 4) == hipSuccess);
    CHECK(hipMalloc(&d_oa, 12 * n * 4) == hipSuccess);
    CHECK(hipMalloc(&d_op, 12 * n * 4) == hipSuccess);
    CHECK(hipMalloc(&d_ref, 12 * n * 4) == hipSuccess);
    hipMemcpy(d_src, packed.data(), packed.size() * 4, hipMemcpyHostToDevice);
    hipMemcpy(d_foci, foci.data(), foci.size() * 4, hipMemcpyHostToDevice);
    hipMemcpy(d_amps, amps.data(), amps.size() * 4, hipMemcpyHostToDevice);
    void* ctx = nullptr;
    CHECK(AFH_CreateContext(&ctx) == AFH_OK);
    std::vector<float> h1(12 * n), h2(12 * n);
    for (int algo = 0; algo <= 2; ++algo)
        for (int prec = 0; prec <= 3; ++prec) {
            if (algo != AFH_ALGO_GSPAT && prec != 0) continue;
            CHECK(AFH_HoloSolve(ctx, d_src, np, n, AFC_SOURCE_T4010A1, d_foci, d_amps, off.data(), 12, algo, prec, 100,
                                d_ref, d_op, nullptr) == AFH_OK);
            CHECK(AFH_HoloSolveNoAlloc(d_src, np, n, AFC_SOURCE_T4010A1, d_foci, d_amps, off.data(), 12, algo, prec,
                                       100, d_oa, d_op, nullptr) == AFH_OK);
            hipDeviceSynchronize();
            hipMemcpy(h1.data(), d_ref, h1.size() * 4, hipMemcpyDeviceToHost);
            hipMemcpy(h2.data(), d_oa, h2.size() * 4, hipMemcpyDeviceToHost);
            CHECK(memcmp(h1.data(), h2.data(), h1.size() * 4) == 0 && all_finite(h1.data(), h1.size()));
        }
    // bad instance lists are refused before any launch
    std::vector<uint32_t> bad_off = off;
    bad_off[3] = bad_off[2];
    CHECK(AFH_HoloSolve(ctx, d_src, np, n, AFC_SOURCE_T4010A1, d_foci, d_amps, bad_off.data(), 12, 0, 0, 100, d_oa,
                        d_op, nullptr) != AFH_OK);
    // LM-family and greedy batches on 16-focus instances
    std::vector<uint32_t> off16(5);
    for (int b = 0; b <= 4; ++b) off16[b] = 16 * b;
    for (int algo = 0; algo <= 2; ++algo)
        CHECK(AFH_NLSolve(ctx, d_src, np, n, AFC_SOURCE_T4010A1, d_foci, d_amps, off16.data(), 4, algo,
                          algo == 2 ? 200 : 20, d_op, nullptr, nullptr) == AFH_OK);
    CHECK(AFH_GreedySolve(ctx, d_src, np, n, AFC_SOURCE_T4010A1, d_foci, d_amps, off16.data(), 4, 16, d_op,
                          nullptr) == AFH_OK);
    CHECK(AFH_APOSolve(ctx, d_src, np, n, AFC_SOURCE_T4010A1, d_foci, d_amps, off16.data(), 2, 0.1f, 20, d_oa, d_op,
                       nullptr, nullptr) == AFH_OK);
    CHECK(AFH_LongSolve(ctx, d_src, np, n, AFC_SOURCE_T4010A1, d_foci, d_amps, off16.data(), 4, d_oa, d_op,
                        nullptr) == AFH_OK);
    CHECK(AFH_HornSolve(ctx, d_src, np, n, AFC_SOURCE_T4010A1, d_foci, d_amps, off16.data(), 4, 7, 10, 0.8f, 1e-5f,
                        d_oa, d_op, nullptr) == AFH_OK);
    hipDeviceSynchronize();
    AFH_DestroyContext(ctx);
    hipFree(d_src);
    hipFree(d_foci);
    hipFree(d_amps);
    hipFree(d_oa);
    hipFree(d_op);
    hipFree(d_ref);
    AFC_FreeUniformSystem(sys, AFC_SOURCE_T4010A1);
}

// One server configuration under 4 submitting threads (each waits for its own
// tickets from a second thread) with field grids, scatter sets and GS-PAT
// requests on three arrays.
static void server_run(uint64_t hbm, uint32_t gspat_cus, uint32_t graphs, int per_thread) {
    AFS_Config cfg = {hbm, 64ull << 20, 4, 32, AFH_HOLO_MIXED, 100, gspat_cus, graphs};
    void* srv = nullptr;
    CHECK(AFS_Create(&cfg, &srv) == AFH_OK);
    if (!srv) return;
    uint32_t ids[3];
    uint32_t ns[3];
    for (int t = 0; t < 3; ++t) {
        const int tiles = t == 0 ? 1 : (t == 1 ? 2 : 4);
        const std::vector<AFC_T4010A1> arr = tiled_array(tiles, tiles, 4.f + t);
        void* sys = make_system(arr);
        void* tv = nullptr;
        const uint64_t n = AFC_GetWaveSources(sys, &tv, AFC_SOURCE_T4010A1);
        std::vector<float> k(n), a(n);
        AFH_GetSourceConstants(sys, k.data(), a.data());
        CHECK(AFS_AddArray(srv, tv, k.data(), a.data(), n, AFC_SOURCE_T4010A1, &ids[t]) == AFH_OK);
        ns[t] = (uint32_t)n;
        AFC_FreeUniformSystem(sys, AFC_SOURCE_T4010A1);
    }
    std::atomic<int> bad{0};
    auto submitter = [&](int seed) {
        std::mt19937 rng(seed);
        std::uniform_real_distribution<float> u(0.f, 1.f);
        struct Job {
            uint64_t t;
            std::vector<float> out, out2, pts;
            size_t len;
        };
        std::vector<Job> jobs(per_thread);
        std::atomic<int> submitted{0};
        std::thread waiter([&] {
            for (int i = 0; i < per_thread; ++i) {
                while (submitted.load() <= i) std::this_thread::yield();
                double lat = 0;
                if (AFS_Wait(srv, jobs[i].t, &lat) != AFH_OK || !all_finite(jobs[i].out.data(), jobs[i].len)) ++bad;
            }
        });
        for (int i = 0; i < per_thread; ++i) {
            Job& j = jobs[i];
            const int aid = (int)(u(rng) * 3) % 3;
            const float kind = u(rng);
            int32_t rc;
            if (kind < 0.4f) {
                AFH_PointSet ps;
                memset(&ps, 0, sizeof(ps));
                ps.kind = 0;
                ps.order[0] = 0;
                ps.order[1] = 1;
                ps.order[2] = 2;
                for (int q = 0; q < 3; ++q) ps.count[q] = 8 + (uint32_t)(40 * u(rng));
                ps.origin[0] = 100.f * u(rng);
                ps.origin[1] = 100.f * u(rng);
                ps.origin[2] = 60.f + 100.f * u(rng);
                ps.resolution = 1.f;
                const int ft = (int)(u(rng) * 3) % 3;
                j.len = (size_t)ps.count[0] * ps.count[1] * ps.count[2] * (ft == 2 ? 2 : 1);
                j.out.resize(j.len);
                rc = AFS_SubmitField(srv, ids[aid], &ps, ft, j.out.data(), &j.t);
            } else if (kind < 0.6f) {
                const uint32_t m = 1 + (uint32_t)(3000 * u(rng));
                j.pts.resize(3 * (size_t)m);
                for (uint32_t q = 0; q < m; ++q) {
                    j.pts[3 * q] = 170.f * u(rng);
                    j.pts[3 * q + 1] = 130.f * u(rng);
                    j.pts[3 * q + 2] = 80.f + 100.f * u(rng);
                }
                AFH_PointSet ps;
                memset(&ps, 0, sizeof(ps));
                ps.kind = 1;
                ps.points = j.pts.data();
                ps.n_points = m;
                j.len = 2 * (size_t)m;
                j.out.resize(j.len);
                rc = AFS_SubmitField(srv, ids[aid], &ps, AFC_FIELD_COMPLEX, j.out.data(), &j.t);
            } else {
                const uint32_t m = 1 + (uint32_t)(128 * u(rng)) % 128;
                std::vector<float> f(3 * (size_t)m), am(m);
                for (uint32_t q = 0; q < m; ++q) {
                    f[3 * q] = 170.f * u(rng);
                    f[3 * q + 1] = 130.f * u(rng);
                    f[3 * q + 2] = 100.f + 150.f * u(rng);
                    am[q] = 0.5f + 0.5f * u(rng);
                }
                j.len = ns[aid];
                j.out.resize(j.len);
                j.out2.resize(j.len);
                // the server copies foci and targets at submission
                rc = AFS_SubmitGSPAT(srv, ids[aid], f.data(), am.data(), m, j.out.data(), j.out2.data(), &j.t);
            }
            if (rc != AFH_OK) ++bad;
            submitted.store(i + 1);
        }
        waiter.join();
    };
    std::vector<std::thread> th;
    for (int s = 0; s < 4; ++s) th.emplace_back(submitter, 100 + s);
    for (auto& t : th) t.join();
    CHECK(bad.load() == 0);
    double p50 = 0, p99 = 0;
    uint64_t done = 0;
    CHECK(AFS_Stats(srv, -1, &p50, &p99, &done, 1) == AFH_OK);
    CHECK(done == 4ull * per_thread);
    uint64_t gl = 0, gb = 0;
    CHECK(AFS_GraphCounters(srv, &gl, &gb) == AFH_OK);
    CHECK(graphs ? gl > 0 : gl == 0);
    // refusals: a request larger than the budget, an unknown ticket and array
    AFH_PointSet big;
    memset(&big, 0, sizeof(big));
    big.order[0] = 0;
    big.order[1] = 1;
    big.order[2] = 2;
    big.count[0] = big.count[1] = big.count[2] = 1024;
    big.resolution = 1.f;
    uint64_t t = 0;
    std::vector<float> tiny(1);
    CHECK(AFS_SubmitField(srv, ids[0], &big, AFC_FIELD_COMPLEX, tiny.data(), &t) == AFH_ERR_BUDGET);
    CHECK(AFS_Wait(srv, 123456789, nullptr) == AFH_ERR_ARG);
    CHECK(AFS_SubmitGSPAT(srv, 99, tiny.data(), tiny.data(), 1, tiny.data(), tiny.data(), &t) == AFH_ERR_ARG);
    printf("  server hbm %llu MiB, gspat_cus %u, graphs %u: %llu requests, p50 %.3f ms, p99 %.3f ms\n",
           (unsigned long long)(hbm >> 20), gspat_cus, graphs, (unsigned long long)done, p50, p99);
    fflush(stdout);
    AFS_Destroy(srv);
}

int main(int argc, char** argv) {
    const bool gpu = argc > 1 && strcmp(argv[1], "all") == 0;
    if (argc < 2 || (!gpu && strcmp(argv[1], "host") != 0)) {
        fprintf(stderr, "usage: host_check host|all\n");
        return 2;
    }
    printf("host part\n");
    fflush(stdout);
    host_part();
    if (gpu) {
        printf("capi part\n");
        fflush(stdout);
        capi_part();
        printf("batched part\n");
        fflush(stdout);
        batched_part();
        printf("server part\n");
        fflush(stdout);
        server_run(512ull << 20, 0, 0, 60);
        server_run(512ull << 20, 32, 1, 60);
        server_run(24ull << 20, 32, 0, 40);  // admission waits for memory
    }
    printf(g_fail ? "host_check: FAILED\n" : "host_check: ok\n");
    return g_fail;
}
