// FP64 device pieces shared by the per-instance dense solvers (lm.hip,
// matrix.hip): the propagation term in FP64 and fixed-order workgroup
// reductions over kLB threads.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "afh_term.hpp"

namespace afh {

constexpr int kLB = 256;

static __constant__ DirEntry c_lm_dir[11] = {AFH_DIR_ENTRIES(AFH_DIR_ENTRY)};

// Reference directivity in FP64: theta binned by ceil(theta / (pi/18)), the
// f32 coefficients of t4010a1.rs:18-69 promoted.
__device__ inline double dir_f64(double theta) {
    const double ad = 3.14159265358979323846 / 18.0;
    const double q = ceil(theta / ad);
    const int i = q > 0 ? (int)fmin(q, 10.0) : 0;
    if (i == 0) return 1.0;
    if (i > 9) return 0.0;
    const DirEntry e = c_lm_dir[i];
    const double x = theta - (i - 1.0) * ad;
    return (double)e.a + ((double)e.b + ((double)e.c + (double)e.d * x) * x) * x;
}

// G_ij in FP64 (uniform_medium.rs:111-120 with phase 0), from the packed source:
// the f32 wavenumber is recovered exactly by rounding (k/2pi hi + lo) 2 pi to
// f32, and slot 11 carries the f32 attenuation, so G uses the same inputs as
// the reference.
__device__ inline double2 term_f64(const float* src, int source_type, double tx, double ty, double tz) {
    const double dx = tx - (double)src[0], dy = ty - (double)src[1], dz = tz - (double)src[2];
    const double dist = sqrt((dx * dx + dy * dy) + dz * dz);
    double d = 1.0;
    if (source_type == kT4010A1) {
        const double c = (((double)src[3] * dx + (double)src[4] * dy) + (double)src[5] * dz) / dist;
        d = dir_f64(acos(c));
    }
    const double k = (double)(float)(((double)src[8] + (double)src[9]) * 6.283185307179586476925);
    const double alpha = (double)src[11];
    const double r = (double)src[6] * d * exp(-dist * alpha) / dist;
    double sn, cs;
    sincos(k * dist, &sn, &cs);
    return make_double2(r * cs, r * sn);
}

// Workgroup barrier that orders LDS accesses only (__syncthreads() also
// waits for outstanding global stores).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Block reductions in a fixed order: lanes by xor shuffles, then the 4 waves.
__device__ inline double block_sum(double v, double* red) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    return ((red[0] + red[1]) + red[2]) + red[3];
}

__device__ inline double block_max(double v, double* red) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    return fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

__device__ inline double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }

__device__ inline double2 block_sum2(double2 v, double* red) {
    const double re = block_sum(v.x, red);
    const double im = block_sum(v.y, red);
    return make_double2(re, im);
}

}  // namespace afh
