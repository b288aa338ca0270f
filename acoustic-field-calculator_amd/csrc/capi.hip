// C ABI of libafc_hip.so: the reference's capi surface (capi/src/*.rs) backed
// by the HIP kernels of this library.
//
// Handles mirror the reference's boxed objects: every create writes a heap
// pointer to *out, every free deletes it, every other call borrows it
// (capi/src/*.rs: Box::from_raw ... forget). Nothing here evaluates a field on
// the CPU: AFC_Calculate always runs the HIP kernel, and a calculator type the
// library does not provide is an error.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "afh_internal.hpp"
#include "afh_host.hpp"

// Host-side per-source constants must round exactly like the reference's f32
// code: no FMA contraction in this file.
#pragma clang fp contract(off)

namespace afh {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

// ------------------------------------------------------------------ physics
// core/sound_speed.rs:22-24 (mm/s)
float sound_speed_f32(float t) { return 331.3f * sqrtf(t / 273.15f) * 1000.0f; }

// core/attenuation.rs:31-48 (Bass et al. 1995), Np/mm, in f32 like the reference
float attenuation_coef_f32(float freq, float hr, float ps, float ps0, float t) {
    const float T0 = 293.15f;
    const float T01 = 273.16f;
    const float ten = 10.0f;
    const float psat = ps0 * powf(ten, -6.8346f * powf(T01 / t, 1.261f) + 4.6151f);
    const float h = ps0 * (hr / ps) * (psat / ps0);
    const float f_ro = (24.f + 4.04e4f * h * (0.02f + h) / (0.391f + h)) / ps0;
    const float f_rn = (1.f / ps0) * (9.f + 280.f * h * expf(-4.17f * (powf(T0 / t, 1.f / 3.f) - 1.f))) *
                       powf(T0 / t, 1.f / 2.f);
    const float f = freq / ps;
    const float alpha =
        (f * f) / ps0 * ps *
        (1.84f * powf(t / T0, 1.f / 2.f) * 1e-11f +
         powf(t / T0, -5.f / 2.f) * (0.01278f * expf(-2239.1f / t) / (f_ro + f * f / f_ro) +
                                     0.1068f * expf(-3352.f / t) / (f_rn + f * f / f_rn)));
    return alpha * 1e-3f;
}

// ------------------------------------------------------------------ packing
uint64_t packed_count(uint64_t n) { return (n + AFH_SOURCE_PAD - 1) / AFH_SOURCE_PAD * AFH_SOURCE_PAD; }

void pack_sources(const float* src, const float* wavenums, const float* attens, uint64_t n, int source_type,
                  float* out) {
    const size_t st = source_type == kT4010A1 ? 9 : 6;
    const double two_pi = 6.283185307179586476925;
    const double log2e = 1.442695040888963407360;
    const uint64_t np = packed_count(n);
    for (uint64_t j = 0; j < np; ++j) {
        float* o = out + AFH_SOURCE_STRIDE * j;
        if (j >= n) {
            // zero-amplitude padding far from any observation point: every
            // intermediate stays finite and the term is exactly 0
            const float pad[AFH_SOURCE_STRIDE] = {-3.0e18f, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
            memcpy(o, pad, sizeof(pad));
            continue;
        }
        const float* s = src + st * j;
        const float amp = source_type == kT4010A1 ? s[6] : s[3];
        const float phase = source_type == kT4010A1 ? s[7] : s[4];
        double ph = (double)phase / two_pi;
        ph -= floor(ph);
        const double kr = (double)wavenums[j] / two_pi;
        const float kh = (float)kr;
        o[0] = s[0];
        o[1] = s[1];
        o[2] = s[2];
        o[3] = source_type == kT4010A1 ? s[3] : 0.0f;
        o[4] = source_type == kT4010A1 ? s[4] : 0.0f;
        o[5] = source_type == kT4010A1 ? s[5] : 0.0f;
        o[6] = amp;
        o[7] = (float)ph;
        o[8] = kh;
        o[9] = (float)(kr - (double)kh);
        o[10] = (float)((double)attens[j] * log2e);
        o[11] = attens[j];  // the f32 attenuation itself, for FP64 consumers
    }
}

// ------------------------------------------------------------------ device buffers
Device& device() {
    static Device d;
    return d;
}

bool Device::ensure() {
    std::lock_guard<std::mutex> g(mu);
    if (ready) return true;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        (void)hipGetLastError();
        set_error("no HIP device visible");
        return false;
    }
    if (hipGetDevice(&id) != hipSuccess) return false;
    ready = true;
    return true;
}

bool DevBuf::reserve(size_t bytes) {
    if (bytes <= cap) return true;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
    size_t want = std::max<size_t>(bytes, 256);
    const hipError_t e = hipMalloc(&ptr, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        set_error("hipMalloc failed for " + std::to_string(want) + " bytes: " + hipGetErrorString(e));
        ptr = nullptr;
        return false;
    }
    cap = want;
    return true;
}
void DevBuf::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
}

bool HostBuf::reserve(size_t bytes) {
    if (bytes <= cap) return true;
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    cap = 0;
    size_t want = std::max<size_t>(bytes, 256);
    const hipError_t e = hipHostMalloc(&ptr, want, hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        set_error("hipHostMalloc failed for " + std::to_string(want) + " bytes: " + hipGetErrorString(e));
        ptr = nullptr;
        return false;
    }
    cap = want;
    return true;
}
void HostBuf::release() {
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    cap = 0;
}

bool hip_ok(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return false;
}

// ------------------------------------------------------------------ grid helper
// builder.rs:198-200: floor((max - min) / res) + 1 with Rust's saturating cast
static uint32_t axis_count(float lo, float hi, float res) {
    const float q = (hi - lo) / res;
    if (!(q > 0.0f)) return 1;
    if (q >= 4294967295.0f) return 0xFFFFFFFFu;
    return (uint32_t)q + 1;
}

static GridArea* make_grid(int ndim, const int axes_tag[3], const float lo[3], const float hi[3], float res) {
    int axes[3];
    bool seen[3] = {false, false, false};
    for (int k = 0; k < 3; ++k) {
        const int t = axes_tag[k];
        if (t < 1 || t > 3 || seen[t - 1]) {
            set_error("grid axes must be a permutation of X=1, Y=2, Z=3");
            return nullptr;
        }
        seen[t - 1] = true;
        axes[k] = t - 1;
    }
    GridArea* g = new GridArea();
    g->ndim = ndim;
    float rlo[3], rhi[3];
    for (int k = 0; k < 3; ++k) {
        // the first ndim axes are ranges, the rest are `*_at` (min == max)
        rlo[axes[k]] = lo[k];
        rhi[axes[k]] = (k < ndim) ? hi[k] : lo[k];
    }
    for (int a = 0; a < 3; ++a) {
        g->bounds[a] = axis_count(rlo[a], rhi[a], res);
        g->origin[a] = rlo[a];
    }
    for (int k = 0; k < 3; ++k) g->order[k] = axes[k];
    g->res = res;
    g->n = (uint64_t)g->bounds[0] * g->bounds[1] * g->bounds[2];
    return g;
}

// Builds kernel arguments for a point set; scatter points must already be on
// the device at `pts_dev`.
FieldArgs field_args_for(const GridArea* g, const float* pts_dev, uint64_t n) {
    FieldArgs a;
    memset(&a, 0, sizeof(a));
    a.n_points = (uint32_t)n;
    if (g) {
        a.ax0 = g->order[0];
        a.ax1 = g->order[1];
        a.ax2 = g->order[2];
        a.count0 = g->bounds[a.ax0];
        a.count1 = g->bounds[a.ax1];
        for (int k = 0; k < 3; ++k) a.origin[k] = g->origin[k];
        a.res = g->res;
    } else {
        a.points = pts_dev;
    }
    return a;
}

}  // namespace afh

using namespace afh;

// ====================================================================== C ABI
extern "C" {

const char* AFC_GetLastError(void) { return g_last_error.c_str(); }

// --------------------------------------------------------------- calculator
void AFC_CreateCalculator(void** out, int32_t calc_type) {
    clear_error();
    if (!out) return;
    // CpuCalculator (0) is accepted and runs the same FAST kernels as Gpu (2)
    // and Hip (3): this library has no CPU path, and the reference's CPU and
    // Vulkan calculators compute the same field (cpu_calculator.rs:43-65)
    if (calc_type < AFC_CALC_CPU || calc_type > AFC_CALC_HIP) {
        set_error("unknown calculator type");
        *out = nullptr;
        return;
    }
    Calculator* c = new Calculator();
    c->calc_type = calc_type;
    *out = c;
}

void AFC_FreeCalculator(void* handle, int32_t) {
    Calculator* c = static_cast<Calculator*>(handle);
    if (!c) return;
    if (c->stream) (void)hipStreamDestroy(c->stream);
    c->src_dev.release();
    c->pts_dev.release();
    c->out_dev.release();
    c->staging.release();
    delete c;
}

uint64_t AFC_Calculate(void* calc_handle, void* system_handle, void* obs_handle, void* field_handle, const void** out,
                       int32_t calc_type, int32_t source_type, int32_t obs_type, int32_t field_type) {
    clear_error();
    if (out) *out = nullptr;
    Calculator* calc = static_cast<Calculator*>(calc_handle);
    System* sys = static_cast<System*>(system_handle);
    FieldBuffer* fb = static_cast<FieldBuffer*>(field_handle);
    if (!calc || !sys || !obs_handle || !fb || !out) {
        set_error("AFC_Calculate: null handle");
        return 0;
    }
    if (calc_type < AFC_CALC_CPU || calc_type > AFC_CALC_HIP || calc_type != calc->calc_type) {
        set_error("AFC_Calculate: unknown calculator type or not matching the handle");
        return 0;
    }
    if (source_type != sys->source_type || (source_type != kSphere && source_type != kT4010A1)) {
        set_error("AFC_Calculate: source type does not match the system");
        return 0;
    }
    if (field_type < AFC_FIELD_PRESSURE || field_type > AFC_FIELD_COMPLEX || field_type != fb->field_type) {
        set_error("AFC_Calculate: field type does not match the field buffer");
        return 0;
    }
    if (obs_type < AFC_AREA_GRID1D || obs_type > AFC_AREA_SCATTER) {
        set_error("AFC_Calculate: unknown observe area type");
        return 0;
    }
    const GridArea* grid = nullptr;
    const ScatterArea* scat = nullptr;
    uint64_t n;
    if (obs_type == AFC_AREA_SCATTER) {
        scat = static_cast<const ScatterArea*>(obs_handle);
        n = scat->pts.size() / 3;
    } else {
        grid = static_cast<const GridArea*>(obs_handle);
        if (grid->ndim != obs_type + 1) {
            set_error("AFC_Calculate: grid dimension does not match the observe area type");
            return 0;
        }
        n = grid->n;
    }
    if (n > kMaxPointsPerCall) {
        set_error("AFC_Calculate: more than 2^31 observation points in one call");
        return 0;
    }
    const size_t per_pt = field_type == AFC_FIELD_COMPLEX ? 2 : 1;
    if (!fb->host.reserve(std::max<uint64_t>(n, 1) * per_pt * sizeof(float))) return 0;
    fb->len = n;
    if (n == 0) {
        *out = fb->host.ptr;
        return 0;
    }
    if (!device().ensure()) return 0;
    if (!calc->stream && !hip_ok(hipStreamCreateWithFlags(&calc->stream, hipStreamNonBlocking), "hipStreamCreate"))
        return 0;
    const uint64_t ns = sys->count();
    const uint64_t np = packed_count(ns);
    const size_t src_bytes = np * AFH_SOURCE_STRIDE * sizeof(float);
    const size_t pts_bytes = scat ? n * 3 * sizeof(float) : 0;
    if (!calc->staging.reserve(src_bytes + pts_bytes + 64)) return 0;
    if (!calc->src_dev.reserve(src_bytes + 64)) return 0;
    if (scat && !calc->pts_dev.reserve(pts_bytes)) return 0;
    if (!calc->out_dev.reserve(n * per_pt * sizeof(float))) return 0;
    // sources are re-read on every call: callers edit them in place through
    // AFC_GetWaveSources (python/main.py:58-63)
    float* stage = static_cast<float*>(calc->staging.ptr);
    pack_sources(sys->sources.data(), sys->wavenums.data(), sys->attens.data(), ns, sys->source_type, stage);
    if (!hip_ok(hipMemcpyAsync(calc->src_dev.ptr, stage, src_bytes, hipMemcpyHostToDevice, calc->stream), "H2D"))
        return 0;
    if (scat) {
        float* sp = stage + np * AFH_SOURCE_STRIDE;
        memcpy(sp, scat->pts.data(), pts_bytes);
        if (!hip_ok(hipMemcpyAsync(calc->pts_dev.ptr, sp, pts_bytes, hipMemcpyHostToDevice, calc->stream), "H2D"))
            return 0;
    }
    FieldArgs a = field_args_for(grid, static_cast<const float*>(calc->pts_dev.ptr), n);
    a.sources = static_cast<const float*>(calc->src_dev.ptr);
    a.n_sources = (uint32_t)np;
    a.out = static_cast<float*>(calc->out_dev.ptr);
    const int prec = calc_type == AFC_CALC_ACCURATE ? kAccurate : kFast;
    if (!hip_ok(launch_field(a, sys->source_type, scat ? kScatter : kGrid, field_type, prec, calc->stream),
                "field kernel launch"))
        return 0;
    if (!hip_ok(hipMemcpyAsync(fb->host.ptr, calc->out_dev.ptr, n * per_pt * sizeof(float), hipMemcpyDeviceToHost,
                               calc->stream),
                "D2H"))
        return 0;
    if (!hip_ok(hipStreamSynchronize(calc->stream), "hipStreamSynchronize")) return 0;
    *out = fb->host.ptr;
    return n;
}

// --------------------------------------------------------------- field buffers
static void create_field(void** out, int type) {
    if (!out) return;
    FieldBuffer* f = new FieldBuffer();
    f->field_type = type;
    *out = f;
}
static void free_field(void* h) {
    FieldBuffer* f = static_cast<FieldBuffer*>(h);
    if (!f) return;
    f->host.release();
    delete f;
}
void AFC_CreatePressureField(void** out) { create_field(out, AFC_FIELD_PRESSURE); }
void AFC_FreePressureField(void* handle) { free_field(handle); }
void AFC_CreatePowerField(void** out) { create_field(out, AFC_FIELD_POWER); }
void AFC_FreePowerField(void* handle) { free_field(handle); }
void AFC_CreateComplexPressureField(void** out) { create_field(out, AFC_FIELD_COMPLEX); }
void AFC_FreeComplexPressureField(void* handle) { free_field(handle); }

// --------------------------------------------------------------- observe areas
void AFC_CreateScatterArea(void** out) {
    if (out) *out = new ScatterArea();
}
void AFC_ScatterAddObservePoint(void* handle, AFC_Vector3 p) {
    ScatterArea* s = static_cast<ScatterArea*>(handle);
    if (!s) return;
    s->pts.push_back(p.x);
    s->pts.push_back(p.y);
    s->pts.push_back(p.z);
}
void AFC_FreeScatterArea(void* handle) { delete static_cast<ScatterArea*>(handle); }

void AFC_CreateGridArea1D(void** out, float first_start, float first_end, float second, float third, float resolution,
                          int32_t first_axis, int32_t second_axis, int32_t third_axis) {
    clear_error();
    if (!out) return;
    const int ax[3] = {first_axis, second_axis, third_axis};
    const float lo[3] = {first_start, second, third};
    const float hi[3] = {first_end, second, third};
    *out = make_grid(1, ax, lo, hi, resolution);
}
void AFC_CreateGridArea2D(void** out, float first_start, float first_end, float second_start, float second_end,
                          float third, float resolution, int32_t first_axis, int32_t second_axis, int32_t third_axis) {
    clear_error();
    if (!out) return;
    const int ax[3] = {first_axis, second_axis, third_axis};
    const float lo[3] = {first_start, second_start, third};
    const float hi[3] = {first_end, second_end, third};
    *out = make_grid(2, ax, lo, hi, resolution);
}
void AFC_CreateGridArea3D(void** out, float first_start, float first_end, float second_start, float second_end,
                          float third_start, float third_end, float resolution, int32_t first_axis,
                          int32_t second_axis, int32_t third_axis) {
    clear_error();
    if (!out) return;
    const int ax[3] = {first_axis, second_axis, third_axis};
    const float lo[3] = {first_start, second_start, third_start};
    const float hi[3] = {first_end, second_end, third_end};
    *out = make_grid(3, ax, lo, hi, resolution);
}
void AFC_FreeGridArea(void* handle, int32_t) { delete static_cast<GridArea*>(handle); }
void AFC_GetGridSize(void* handle, int32_t, uint32_t* first, uint32_t* second, uint32_t* third) {
    const GridArea* g = static_cast<const GridArea*>(handle);
    if (!g) return;
    if (first) *first = g->bounds[0];
    if (second) *second = g->bounds[1];
    if (third) *third = g->bounds[2];
}

// --------------------------------------------------------------- system
void AFC_CreateUniformSystem(void** out, float temperature, int32_t source_type) {
    clear_error();
    if (!out) return;
    if (source_type != kSphere && source_type != kT4010A1) {
        set_error("unknown source type");
        *out = nullptr;
        return;
    }
    System* s = new System();
    s->source_type = source_type;
    s->temperature = temperature;
    s->sound_speed = sound_speed_f32(temperature);
    *out = s;
}
void AFC_FreeUniformSystem(void* handle, int32_t) { delete static_cast<System*>(handle); }
float AFC_UniformSystemSoundSpeed(void* handle, int32_t) {
    const System* s = static_cast<const System*>(handle);
    return s ? s->sound_speed : 0.0f;
}
// system/uniform_medium.rs:73-80
char* AFC_UniformSystemInfo(void* handle, int32_t) {
    System* s = static_cast<System*>(handle);
    if (!s) return nullptr;
    char buf[256];
    snprintf(buf, sizeof(buf), "Uniform Medium:\n Temperature: %g K = %g \xe2\x84\x83\n Sound Speed: %g mm/s",
             (double)s->temperature, (double)(s->temperature - 273.15f), (double)sound_speed_f32(s->temperature));
    s->info = buf;
    return &s->info[0];
}
// system/uniform_medium.rs:82-90
char* AFC_UniformSystemSourceInfo(void* handle, uint64_t idx, int32_t) {
    System* s = static_cast<System*>(handle);
    if (!s) return nullptr;
    if (idx >= s->count()) {
        set_error("source index out of range");
        s->info.clear();
        return &s->info[0];
    }
    const float two_pi = 2.0f * kPiF;
    char buf[256];
    snprintf(buf, sizeof(buf), "%llu-th wave source:\n Wavelength: %g mm\n Wavenumber: %g mm^-1\n Attenuation: %g Np/mm",
             (unsigned long long)idx, (double)(two_pi / s->wavenums[idx]), (double)s->wavenums[idx],
             (double)s->attens[idx]);
    s->info = buf;
    return &s->info[0];
}

// system/uniform_medium.rs:102-107: k = 2 PI f / c(T), alpha at 30 %RH, 1 atm
static void add_source(System* s, const float* fields, size_t nfields, float freq) {
    const float c = sound_speed_f32(s->temperature);
    const float k = 2.0f * kPiF * freq / c;
    const float alpha = attenuation_coef_f32(freq, 30.f, 1.f, 1.f, s->temperature);
    s->sources.insert(s->sources.end(), fields, fields + nfields);
    s->wavenums.push_back(k);
    s->attens.push_back(alpha);
}
void AFC_AddSphereWaveSource(void* handle, AFC_SphereWaveSource source) {
    System* s = static_cast<System*>(handle);
    if (!s || s->source_type != kSphere) {
        set_error("AFC_AddSphereWaveSource: system is not a SphereWaveSource system");
        return;
    }
    add_source(s, &source.x, 6, source.freq);
}
void AFC_AddT4010A1(void* handle, AFC_T4010A1 source) {
    System* s = static_cast<System*>(handle);
    if (!s || s->source_type != kT4010A1) {
        set_error("AFC_AddT4010A1: system is not a T4010A1 system");
        return;
    }
    add_source(s, &source.x, 9, source.freq);
}
uint64_t AFC_GetWaveSources(void* handle, void** out, int32_t) {
    System* s = static_cast<System*>(handle);
    if (!s || !out) return 0;
    *out = s->sources.data();
    return s->count();
}

uint64_t AFH_GetSourceConstants(void* system_handle, float* wavenums, float* attens) {
    const System* s = static_cast<const System*>(system_handle);
    if (!s) return 0;
    if (wavenums) memcpy(wavenums, s->wavenums.data(), s->count() * sizeof(float));
    if (attens) memcpy(attens, s->attens.data(), s->count() * sizeof(float));
    return s->count();
}

// --------------------------------------------------------------- packing (device layout)
uint64_t AFH_PackedCount(uint64_t n) { return packed_count(n); }
int32_t AFH_PackSources(const void* sources, const float* wavenums, const float* attens, uint64_t n,
                        int32_t source_type, float* out) {
    clear_error();
    if ((source_type != kSphere && source_type != kT4010A1) || (!sources && n) || !out) {
        set_error("AFH_PackSources: bad argument");
        return AFH_ERR_ARG;
    }
    pack_sources(static_cast<const float*>(sources), wavenums, attens, n, source_type, out);
    return AFH_OK;
}

int32_t AFH_FieldEval(const float* packed_sources_dev, uint64_t n_packed, int32_t source_type, const AFH_PointSet* pts,
                      int32_t field_type, int32_t precision, float* out_dev, void* stream) {
    clear_error();
    if (!pts || !out_dev || (n_packed && !packed_sources_dev) || n_packed % AFH_SOURCE_PAD != 0 ||
        (source_type != kSphere && source_type != kT4010A1) || field_type < 0 || field_type > 2 ||
        (precision != AFH_PREC_FAST && precision != AFH_PREC_ACCURATE) || n_packed > 0xFFFFFFFFull) {
        set_error("AFH_FieldEval: bad argument");
        return AFH_ERR_ARG;
    }
    uint64_t n;
    GridArea g;
    const GridArea* gp = nullptr;
    if (pts->kind == 0) {
        for (int k = 0; k < 3; ++k) {
            if (pts->order[k] < 0 || pts->order[k] > 2) {
                set_error("AFH_FieldEval: bad grid order");
                return AFH_ERR_ARG;
            }
            g.order[k] = pts->order[k];
            g.bounds[pts->order[k]] = pts->count[k];
            g.origin[k] = pts->origin[k];
        }
        if (g.order[0] == g.order[1] || g.order[1] == g.order[2] || g.order[0] == g.order[2]) {
            set_error("AFH_FieldEval: grid order must be a permutation");
            return AFH_ERR_ARG;
        }
        g.res = pts->resolution;
        n = (uint64_t)pts->count[0] * pts->count[1] * pts->count[2];
        gp = &g;
    } else if (pts->kind == 1) {
        n = pts->n_points;
        if (n && !pts->points) {
            set_error("AFH_FieldEval: null points");
            return AFH_ERR_ARG;
        }
    } else {
        set_error("AFH_FieldEval: bad point-set kind");
        return AFH_ERR_ARG;
    }
    if (n > kMaxPointsPerCall) {
        set_error("AFH_FieldEval: more than 2^31 points in one call");
        return AFH_ERR_ARG;
    }
    FieldArgs a = field_args_for(gp, pts->points, n);
    a.sources = packed_sources_dev;
    a.n_sources = (uint32_t)n_packed;
    a.out = out_dev;
    if (!hip_ok(launch_field(a, source_type, gp ? kGrid : kScatter, field_type,
                             precision == AFH_PREC_ACCURATE ? kAccurate : kFast, (hipStream_t)stream),
                "field kernel launch"))
        return AFH_ERR_HIP;
    return AFH_OK;
}

}  // extern "C"
