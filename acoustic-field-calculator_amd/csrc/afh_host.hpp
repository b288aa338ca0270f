// Host-side handle types behind the C ABI. Each mirrors a boxed object of the
// reference's capi (capi/src/*.rs).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <string>
#include <vector>

#include "afh_internal.hpp"

namespace afh {

// Grow-only device allocation.
struct DevBuf {
    void* ptr = nullptr;
    size_t cap = 0;
    bool reserve(size_t bytes);
    void release();
};

// Grow-only pinned host allocation.
struct HostBuf {
    void* ptr = nullptr;
    size_t cap = 0;
    bool reserve(size_t bytes);
    void release();
};

struct Device {
    std::mutex mu;
    bool ready = false;
    int id = 0;
    bool ensure();
};
Device& device();
bool hip_ok(hipError_t e, const char* what);

// UniformSystem<S> (system/uniform_medium.rs:21-27): the source structs in the
// reference's own layout (6 or 9 floats each) plus per-source k and alpha.
struct System {
    int source_type = 0;
    float temperature = 0.f;
    float sound_speed = 0.f;
    std::vector<float> sources;
    std::vector<float> wavenums;
    std::vector<float> attens;
    std::string info;
    uint64_t count() const { return wavenums.size(); }
    size_t stride() const { return source_type == kT4010A1 ? 9 : 6; }
};

// GridArea<N1|N2|N3> (observe_area/grid/grid_area.rs:236-241): points are not
// materialised; the kernel generates them from these parameters.
struct GridArea {
    int ndim = 0;
    int order[3] = {0, 1, 2};  // axes, fastest first (ranged axes in given order, then `*_at` axes)
    uint32_t bounds[3] = {1, 1, 1};  // counts by axis (x, y, z)
    float origin[3] = {0, 0, 0};
    float res = 1.f;
    uint64_t n = 0;
};

// ScatterArea (observe_area/scatter/scatter_area.rs:17-48)
struct ScatterArea {
    std::vector<float> pts;
};

// PressureField / PowerField / ComplexPressureField (field_buffer/*.rs)
struct FieldBuffer {
    int field_type = 0;
    HostBuf host;
    uint64_t len = 0;
};

// GpuCalculator slot: one stream plus grow-only device buffers.
struct Calculator {
    int calc_type = 0;
    hipStream_t stream = nullptr;
    DevBuf src_dev, pts_dev, out_dev;
    HostBuf staging;
};

float sound_speed_f32(float t);
float attenuation_coef_f32(float freq, float hr, float ps, float ps0, float t);
uint64_t packed_count(uint64_t n);
void pack_sources(const float* src, const float* wavenums, const float* attens, uint64_t n, int source_type,
                  float* out);
FieldArgs field_args_for(const GridArea* g, const float* pts_dev, uint64_t n);

}  // namespace afh
