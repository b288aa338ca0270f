// Hologram solvers (GS, GS-PAT, Naive, Greedy, LM, Gauss-Newton, gradient
// descent, APO, Long, Horn) on the GPU.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <mutex>
#include <type_traits>
#include <vector>

#include "afh_host.hpp"

namespace afh {

enum HoloAlgo { kAlgoGSPAT = 0, kAlgoGS = 1, kAlgoNaive = 2, kAlgoLM = 3, kAlgoGreedy = 4, kAlgoGN = 5, kAlgoGD = 6 };

// Nonlinear least-squares solvers of lm.hip (= AFH_NLS_* of the C ABI).
enum NlsAlgo { kNlsLM = 0, kNlsGN = 1, kNlsGD = 2 };

// Matrix-method solvers of matrix.hip.
enum MatMethod { kMatLong = 0, kMatHorn = 1 };

// Greedy limits: foci per instance and phase candidates.
constexpr uint32_t kGreedyMaxM = 128;
constexpr uint32_t kGreedyMaxDiv = 4096;

// One problem in the reference's host layout: sources as repr(C) structs,
// foci as 3 floats each.
struct HoloProblem {
    int source_type;
    const float* sources;
    const float* wavenums;
    const float* attens;
    uint64_t n;
    const float* foci;
    const float* amps;
    uint64_t m;
    uint32_t phase_div = 0;  // Greedy only
};

// A batch of instances on the device (all pointers device memory except
// offsets_host): instance b owns foci [offsets[b], offsets[b+1]).
struct HoloBatch {
    const float* sources;
    uint32_t n, n_packed;
    int source_type;
    const float* foci;
    const float* amps;
    const uint32_t* offsets_host;
    uint32_t batch;
    int algo, precision;
    uint32_t iterations;
    uint32_t phase_div = 0;  // Greedy only
    float* out_amp;
    float* out_phase;
};

// One kernel launch kept as a record: the kernel, its launch shape and a copy
// of its arguments (at most 4, stored in `blob` at offsets `off`).
struct KLaunch {
    const void* fn = nullptr;
    dim3 grid, block;
    int nargs = 0;
    uint16_t off[4] = {0, 0, 0, 0};
    alignas(16) unsigned char blob[256];
};

// Launches that run one after another on one stream (at most 4).
struct KChain {
    KLaunch k[4];
    int n = 0;
    template <class... A>
    void add(const void* fn, dim3 grid, dim3 block, const A&... args) {
        static_assert(sizeof...(A) >= 1 && sizeof...(A) <= 4, "one to four kernel arguments");
        KLaunch& l = k[n++];
        l.fn = fn;
        l.grid = grid;
        l.block = block;
        l.nargs = 0;
        size_t pos = 0;
        auto put = [&](const auto& v) {
            using T = std::decay_t<decltype(v)>;
            pos = (pos + alignof(T) - 1) & ~(alignof(T) - 1);
            memcpy(l.blob + pos, &v, sizeof(T));
            l.off[l.nargs++] = (uint16_t)pos;
            pos += sizeof(T);
        };
        (put(args), ...);
    }
};

// A KChain instantiated as a hipGraph (a linear chain of kernel nodes). Later
// chains of the same kernels and block shapes reuse it with their own grids
// and arguments set on its nodes, once its previous launch has completed.
struct ChainGraph {
    const void* fn[4] = {nullptr, nullptr, nullptr, nullptr};
    dim3 block[4];
    int n = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraphNode_t node[4] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t done = nullptr;  // recorded after each launch
    uint64_t last_use = 0;
};

// Stream + grow-only device workspace; calls on one context are serialized.
// A small GS batch runs its padded-focus buckets side by side on the context's
// auxiliary streams (created on first use, with cu_mask when one is set),
// forked from and joined back into the caller's stream by events.
struct Context {
    std::mutex mu;
    hipStream_t stream = nullptr;
    DevBuf work, io;
    HostBuf staging, io_host;
    hipEvent_t upload_done = nullptr;
    std::vector<uint32_t> cu_mask;  // CU mask of the auxiliary streams (empty: all CUs)
    hipStream_t aux[4] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t aux_done[4] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t fork = nullptr;
    // GS-PAT bucket chains issued as hipGraphs (off: launch by launch)
    bool graphs = false;
    std::vector<ChainGraph> chain_graphs;
    uint64_t graph_clock = 0, graph_launches = 0, graph_builds = 0;
    bool ensure_stream() {
        if (stream) return true;
        return hip_ok(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
    }
    bool ensure_aux();
    bool reserve_staging(size_t bytes);
    void drop_graph(ChainGraph& g) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
        if (g.done) (void)hipEventDestroy(g.done);
        g = ChainGraph();
    }
    ~Context() {
        for (auto& g : chain_graphs) {
            if (g.done) (void)hipEventSynchronize(g.done);
            drop_graph(g);
        }
        for (int i = 0; i < 4; ++i) {
            if (aux[i]) (void)hipStreamDestroy(aux[i]);
            if (aux_done[i]) (void)hipEventDestroy(aux_done[i]);
        }
        if (fork) (void)hipEventDestroy(fork);
        if (upload_done) (void)hipEventDestroy(upload_done);
        if (stream) (void)hipStreamDestroy(stream);
        work.release();
        io.release();
        staging.release();
        io_host.release();
    }
};

int32_t holo_solve(Context* ctx, const HoloBatch& hb, hipStream_t st);
// Workspace bytes a GS-PAT holo_solve of `batch` instances on an array of
// n_packed sources needs at most (every instance with 65..128 foci).
size_t gspat_workspace_bound(uint32_t n_packed, uint32_t batch);
Context& default_context();

// Solves one instance on the device and returns per-source amplitude and
// phase (host arrays of n floats). false on error (message set).
bool holo_solve_single(const HoloProblem& p, int algo, float* amp_out, float* phase_out);
int32_t lm_solve_batch(Context* ctx, const float* sources, uint32_t n, uint32_t n_packed, int source_type,
                       const float* foci, const float* amps, const uint32_t* offsets_host, uint32_t batch,
                       uint32_t k_max, float* out_phase, uint32_t* out_iters, hipStream_t st, int algo = kNlsLM);
// LM (lm.rs:40-105), Gauss-Newton (gauss_newton.rs:35-89) or gradient descent
// (gradient_descent.rs:34-61) with the reference's K_MAX: phases of the n sources.
bool lm_solve_single(const HoloProblem& p, float* phase_out, int algo = kNlsLM);
// APO (apo.rs:100-158), K_MAX 200: amplitudes and phases of the n sources.
int32_t apo_solve_batch(Context* ctx, const float* sources, uint32_t n, int source_type, const float* foci,
                        const float* amps, const uint32_t* offsets_host, uint32_t batch, double lambda,
                        uint32_t k_max, float* out_amp, float* out_phase, uint32_t* out_iters, hipStream_t st);
bool apo_solve_single(const HoloProblem& p, float lambda, float* amp_out, float* phase_out);
// Long (long.rs:46-106) or Horn (horn.rs:58-129; instance b draws its rows
// from seed + b): amplitudes and phases of the n sources; M <= 32 per instance.
int32_t matrix_solve_batch(Context* ctx, const float* sources, uint32_t n, int source_type, const float* foci,
                           const float* amps, const uint32_t* offsets_host, uint32_t batch, int method,
                           uint64_t seed, uint32_t repeat, double lambda, double alpha, float* out_amp,
                           float* out_phase, hipStream_t st);
// One instance with the reference's constants (REPEAT 100, lambda 0.8,
// Tikhonov 1e-5) and seed 0.
bool matrix_solve_single(const HoloProblem& p, int method, float* amp_out, float* phase_out);

}  // namespace afh
