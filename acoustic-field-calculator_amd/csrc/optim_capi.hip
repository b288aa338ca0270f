// Optimizer exports of the C ABI (capi/src/optimizer.rs).
//
// GS, GS-PAT, Naive, Greedy, LM, Gauss-Newton, gradient descent, APO, Long
// and Horn run on the GPU (holo.hip / lm.hip / matrix.hip); FocalPoint and
// BesselBeam set phases on the host, as the reference does. IFFT (an image
// file read and an FFT) is outside this backend: it records "unsupported" and
// leaves the system untouched.
#include <math.h>
#include <string.h>

#include <string>
#include <vector>

#include "afh_host.hpp"
#include "afh_internal.hpp"
#include "afh_holo.hpp"

#pragma clang fp contract(off)

using namespace afh;

namespace {

System* as_system(void* h, int32_t source_type, const char* who) {
    System* s = static_cast<System*>(h);
    if (!s || s->source_type != source_type) {
        set_error(std::string(who) + ": null system or source type mismatch");
        return nullptr;
    }
    return s;
}

void unsupported(const char* who) {
    set_error(std::string(who) + " is not provided by libafc_hip (outside the GPU hot path)");
}

// Shared front end of AFO_GS / AFO_GSPAT / AFO_Naive / AFO_LM: one instance
// through the batched device solver, then amp / phase written back into the
// system's sources the way the reference does (gs.rs:71-76, gs_pat.rs:75-80,
// naive.rs:45-50, lm.rs:102-104).
void run_single(void* handle, const void* foci, const float* amps, uint64_t size, int32_t source_type, int algo,
                const char* who, uint32_t phase_div = 0) {
    clear_error();
    System* s = as_system(handle, source_type, who);
    if (!s) return;
    if (size == 0 || !foci || !amps) {
        set_error(std::string(who) + ": no foci");
        return;
    }
    const uint64_t n = s->count();
    if (n == 0) return;
    std::vector<float> amp_out(n), phase_out(n);
    HoloProblem p;
    p.source_type = source_type;
    p.sources = s->sources.data();
    p.wavenums = s->wavenums.data();
    p.attens = s->attens.data();
    p.n = n;
    p.foci = static_cast<const float*>(foci);
    p.amps = amps;
    p.m = size;
    p.phase_div = phase_div;
    if (!holo_solve_single(p, algo, amp_out.data(), phase_out.data())) return;
    const size_t st = s->stride();
    const size_t ia = source_type == kT4010A1 ? 6 : 3;
    for (uint64_t j = 0; j < n; ++j) {
        float* src = s->sources.data() + st * j;
        // LM, Gauss-Newton, gradient descent and Greedy set phases only
        if (algo != kAlgoLM && algo != kAlgoGN && algo != kAlgoGD && algo != kAlgoGreedy) src[ia] = amp_out[j];
        src[ia + 1] = phase_out[j];
    }
}

// Front end of the solvers that set amplitudes and phases and are not part of
// holo_solve_single (APO, Long, Horn): solve, then write both back.
template <class Solve>
void run_amp_phase(void* handle, const void* foci, const float* amps, uint64_t size, int32_t source_type,
                   const char* who, Solve solve) {
    clear_error();
    System* s = as_system(handle, source_type, who);
    if (!s) return;
    if (size == 0 || !foci || !amps) {
        set_error(std::string(who) + ": no foci");
        return;
    }
    const uint64_t n = s->count();
    if (n == 0) return;
    std::vector<float> amp_out(n), phase_out(n);
    HoloProblem p;
    p.source_type = source_type;
    p.sources = s->sources.data();
    p.wavenums = s->wavenums.data();
    p.attens = s->attens.data();
    p.n = n;
    p.foci = static_cast<const float*>(foci);
    p.amps = amps;
    p.m = size;
    if (!solve(p, amp_out.data(), phase_out.data())) return;
    const size_t st = s->stride();
    const size_t ia = source_type == kT4010A1 ? 6 : 3;
    for (uint64_t j = 0; j < n; ++j) {
        float* src = s->sources.data() + st * j;
        src[ia] = amp_out[j];
        src[ia + 1] = phase_out[j];
    }
}

}  // namespace

extern "C" {

// acoustic-field-optimizer/src/focal_point.rs:28-39 (host-side input set-up,
// as in the reference: phi_j = -2 PI ((|s_j - f| mod lambda) / lambda))
void AFO_FocalPoint(void* handle, AFC_Vector3 point, int32_t source_type) {
    clear_error();
    System* s = as_system(handle, source_type, "AFO_FocalPoint");
    if (!s) return;
    const size_t st = s->stride();
    for (uint64_t j = 0; j < s->count(); ++j) {
        float* src = s->sources.data() + st * j;
        const float dx = src[0] - point.x, dy = src[1] - point.y, dz = src[2] - point.z;
        const float d = sqrtf((dx * dx + dy * dy) + dz * dz);
        const float wl = s->sound_speed / src[st - 1];
        const float ph = fmodf(d, wl) / wl;
        src[st - 2] = -2.0f * kPiF * ph;
    }
}

void AFO_GSPAT(void* handle, const void* foci, const float* amps, uint64_t size, int32_t source_type) {
    run_single(handle, foci, amps, size, source_type, kAlgoGSPAT, "AFO_GSPAT");
}
void AFO_GS(void* handle, const void* foci, const float* amps, uint64_t size, int32_t source_type) {
    run_single(handle, foci, amps, size, source_type, kAlgoGS, "AFO_GS");
}
void AFO_Naive(void* handle, const void* foci, const float* amps, uint64_t size, int32_t source_type) {
    run_single(handle, foci, amps, size, source_type, kAlgoNaive, "AFO_Naive");
}
void AFO_LM(void* handle, const void* foci, const float* amps, uint64_t size, int32_t source_type) {
    run_single(handle, foci, amps, size, source_type, kAlgoLM, "AFO_LM");
}

// acoustic-field-optimizer/src/bessel.rs:31-55 (capi/src/optimizer.rs:47-56),
// host-side like FocalPoint and in the reference's f32 operation order:
// v = (dir_y, -dir_x, 0) (used unnormalised), theta_w = asin |v|, r = s - point
// turned by r cos theta_w + (r x v) sin theta_w + v (v.r (1 - cos theta_w));
// phi_j = -2 PI ((dist mod lambda) / lambda), dist = sin theta |r_xy| - cos theta r_z
void AFO_BesselBeam(void* handle, AFC_Vector3 point, AFC_Vector3 dir, float theta, int32_t source_type) {
    clear_error();
    System* s = as_system(handle, source_type, "AFO_BesselBeam");
    if (!s) return;
    const float vx = dir.y, vy = -dir.x, vz = 0.0f;
    const float tw = asinf(sqrtf((vx * vx + vy * vy) + vz * vz));
    const float ct = cosf(tw), stw = sinf(tw);
    const float sth = sinf(theta), cth = cosf(theta);
    const size_t st = s->stride();
    for (uint64_t j = 0; j < s->count(); ++j) {
        float* src = s->sources.data() + st * j;
        const float rx = src[0] - point.x, ry = src[1] - point.y, rz = src[2] - point.z;
        const float xx = ry * vz - rz * vy, xy = rz * vx - rx * vz, xz = rx * vy - ry * vx;
        const float vr = ((vx * rx + vy * ry) + vz * rz) * (1.0f - ct);
        const float qx = (rx * ct + xx * stw) + vx * vr;
        const float qy = (ry * ct + xy * stw) + vy * vr;
        const float qz = (rz * ct + xz * stw) + vz * vr;
        const float dist = sth * sqrtf(qx * qx + qy * qy) - cth * qz;
        const float wl = s->sound_speed / src[st - 1];
        src[st - 2] = -2.0f * kPiF * (fmodf(dist, wl) / wl);
    }
}
void AFO_IFFT(void*, const char*, AFC_Vector3, AFC_Vector3, AFC_Vector3, float, float, int32_t) {
    unsupported("AFO_IFFT");
}
// matrix/horn.rs:70-129 (capi/src/optimizer.rs:130-141): the reference's
// constants; the random rows come from seed 0 (the reference's thread_rng is
// unseeded)
void AFO_Horn(void* handle, const void* foci, const float* amps, uint64_t size, int32_t source_type) {
    run_amp_phase(handle, foci, amps, size, source_type, "AFO_Horn",
                  [](const HoloProblem& p, float* am, float* ph) { return matrix_solve_single(p, kMatHorn, am, ph); });
}
// matrix/long.rs:46-106 (capi/src/optimizer.rs:144-155)
void AFO_Long(void* handle, const void* foci, const float* amps, uint64_t size, int32_t source_type) {
    run_amp_phase(handle, foci, amps, size, source_type, "AFO_Long",
                  [](const HoloProblem& p, float* am, float* ph) { return matrix_solve_single(p, kMatLong, am, ph); });
}
// nls/apo.rs:100-158 (capi/src/optimizer.rs:158-174): sets amplitudes and phases
void AFO_APO(void* handle, const void* foci, const float* amps, uint64_t size, float lambda, int32_t source_type) {
    run_amp_phase(handle, foci, amps, size, source_type, "AFO_APO", [lambda](const HoloProblem& p, float* am, float* ph) {
        return apo_solve_single(p, lambda, am, ph);
    });
}
// nls/gauss_newton.rs:35-89 (capi/src/optimizer.rs:177-192)
void AFO_GaussNewton(void* handle, const void* foci, const float* amps, uint64_t size, int32_t source_type) {
    run_single(handle, foci, amps, size, source_type, kAlgoGN, "AFO_GaussNewton");
}
// nls/gradient_descent.rs:34-61 (capi/src/optimizer.rs:195-210)
void AFO_GradientDescent(void* handle, const void* foci, const float* amps, uint64_t size, int32_t source_type) {
    run_single(handle, foci, amps, size, source_type, kAlgoGD, "AFO_GradientDescent");
}
// combination/greedy.rs:23-83 (capi/src/optimizer.rs:227-241)
void AFO_Greedy(void* handle, const void* foci, const float* amps, uint64_t size, uint64_t phase_div,
                int32_t source_type) {
    if (phase_div == 0 || phase_div > kGreedyMaxDiv) {
        clear_error();
        set_error("AFO_Greedy: phase_div must be in 1..4096");
        return;
    }
    run_single(handle, foci, amps, size, source_type, kAlgoGreedy, "AFO_Greedy", (uint32_t)phase_div);
}

}  // extern "C"
