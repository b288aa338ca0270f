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

// v of lane (lane ^ O). O = 1, 2 and 8 are single DPP moves within a 16-lane
// row (quad permutes, row rotate by 8); O = 4 takes lane + 4 (row shift left)
// or lane - 4 (row shift right) by the lane's bit 2; larger O use the LDS
// crossbar (__shfl_xor). Every lane of the wave must be active.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, 0xf, 0xf, false);
}

template <int O>
__device__ __forceinline__ double xor_lane_f64(double v) {
    if constexpr (O >= 16) {
        return __shfl_xor(v, O);
    } else {
        const uint64_t u = __builtin_bit_cast(uint64_t, v);
        uint32_t lo = (uint32_t)u, hi = (uint32_t)(u >> 32);
        if constexpr (O == 1) {
            lo = dpp_u32<0xb1>(lo);  // quad_perm [1, 0, 3, 2]
            hi = dpp_u32<0xb1>(hi);
        } else if constexpr (O == 2) {
            lo = dpp_u32<0x4e>(lo);  // quad_perm [2, 3, 0, 1]
            hi = dpp_u32<0x4e>(hi);
        } else if constexpr (O == 8) {
            lo = dpp_u32<0x128>(lo);  // row_ror:8
            hi = dpp_u32<0x128>(hi);
        } else {
            static_assert(O == 4, "xor offset");
            const bool up = (threadIdx.x & 4) != 0;
            const uint32_t l_dn = dpp_u32<0x104>(lo), h_dn = dpp_u32<0x104>(hi);  // row_shl:4, lane + 4
            const uint32_t l_up = dpp_u32<0x114>(lo), h_up = dpp_u32<0x114>(hi);  // row_shr:4, lane - 4
            lo = up ? l_up : l_dn;
            hi = up ? h_up : h_dn;
        }
        return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
    }
}

// One step of a transposed butterfly over N values per lane: the lane with
// bit O set keeps the upper half of its list and sends the lower half to
// lane ^ O, which keeps the lower half; each lane adds what it receives to
// what it keeps (N / 2 values remain).
template <int O, int N>
__device__ __forceinline__ void butterfly_step(double* v, int lane) {
    const bool up = (lane & O) != 0;
#pragma unroll
    for (int j = 0; j < N / 2; ++j) {
        const double send = up ? v[j] : v[j + N / 2];
        const double keep = up ? v[j + N / 2] : v[j];
        v[j] = keep + xor_lane_f64<O>(send);
    }
}

// Sum (maximum) over the 64 lanes by xor steps at offsets 32, 16, .. 1 (the
// last four by DPP); every lane gets the same bits. All lanes must be active.
__device__ __forceinline__ double wave_sum_f64(double v) {
    v += xor_lane_f64<32>(v);
    v += xor_lane_f64<16>(v);
    v += xor_lane_f64<8>(v);
    v += xor_lane_f64<4>(v);
    v += xor_lane_f64<2>(v);
    v += xor_lane_f64<1>(v);
    return v;
}

__device__ __forceinline__ double wave_max_f64(double v) {
    v = fmax(v, xor_lane_f64<32>(v));
    v = fmax(v, xor_lane_f64<16>(v));
    v = fmax(v, xor_lane_f64<8>(v));
    v = fmax(v, xor_lane_f64<4>(v));
    v = fmax(v, xor_lane_f64<2>(v));
    v = fmax(v, xor_lane_f64<1>(v));
    return v;
}

// Block reductions in a fixed order: lanes by xor steps, then the 4 waves.
__device__ inline double block_sum(double v, double* red) {
    v = wave_sum_f64(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    return ((red[0] + red[1]) + red[2]) + red[3];
}

__device__ inline double block_max(double v, double* red) {
    v = wave_max_f64(v);
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
