// Internal declarations shared by the HIP translation units of libafc_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/afc_hip.h"
#include "afh_term.hpp"

namespace afh {

// Point indices are 32-bit in the kernels (block base + lane stays below 2^32).
constexpr uint64_t kMaxPointsPerCall = 1ull << 31;

enum PointKind { kGrid = 0, kScatter = 1 };
enum OutKind { kAbs = 0, kAbs2 = 1, kComplex = 2 };  // = AFC_FIELD_* tags

// Kernel arguments of the field evaluation (passed by value).
struct FieldArgs {
    const float* sources;  // packed, AFH_SOURCE_STRIDE floats each
    uint32_t n_sources;    // packed count (multiple of AFH_SOURCE_PAD)
    uint32_t n_points;
    // grid: fastest-varying axis first
    int ax0, ax1, ax2;
    uint32_t count0, count1;
    float origin[3];
    float res;
    const float* points;  // scatter: 3 floats per point
    float* out;
    const DirCell* cells;  // FAST-mode directivity table (set by launch_field)
};

// Host builder of the FAST-mode directivity cells and the per-device copy.
void build_dir_cells(DirCell* cells);
const DirCell* dir_cells_device();

hipError_t launch_field(const FieldArgs& a, int src, int pts, int out, int prec, hipStream_t stream);

// Thread-local last-error string of the C ABI.
void set_error(const std::string& msg);
void clear_error();

}  // namespace afh
