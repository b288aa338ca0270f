// Hologram solvers (GS, GS-PAT, Naive, LM) on the GPU.
#pragma once

#include <stdint.h>

namespace afh {

enum HoloAlgo { kAlgoGSPAT = 0, kAlgoGS = 1, kAlgoNaive = 2, kAlgoLM = 3 };

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
};

// Solves one instance on the device and returns per-source amplitude and
// phase (host arrays of n floats). false on error (message set).
bool holo_solve_single(const HoloProblem& p, int algo, float* amp_out, float* phase_out);

}  // namespace afh
