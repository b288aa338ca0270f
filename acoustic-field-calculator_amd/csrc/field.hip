// Field kernel: complex pressure of all sources at every observation point.
//
// Replaces CpuCalculator::calculate (calculator/cpu_calculator.rs:43-65) over
// UniformSystem::propagate_all (system/uniform_medium.rs:122-128) and the
// Vulkan shaders assets/shaders/cpressure_base.comp:43-71 (+ cpressure.comp,
// pressure.comp, power.comp output stages).
//
// Layout and schedule (gfx950):
//   * workgroup = 256 lanes (4 wave64); each lane owns PPL observation points
//     strided by 256 so loads of scatter points and all output stores coalesce;
//   * grid points are generated in registers from (origin, resolution,
//     counts), with the reference's own f32 rounding (origin + n * res, no
//     contraction), so no point array is read from HBM;
//   * packed sources (12 floats each) are staged through LDS in tiles of
//     kTile; every lane of the workgroup reads the same LDS address
//     (broadcast), so the per-term operands cost one ds_read_b128 x3 per
//     source per wave;
//   * sums: plain FP32 over chunks of kChunk sources, then TwoSum-compensated
//     across chunks, so the rounding error does not grow with N.
#include "afh_internal.hpp"

namespace afh {

constexpr int kBlock = 256;
constexpr int kTile = 128;
constexpr int kChunk = 32;
static_assert(kTile % kChunk == 0, "tile must hold whole chunks");
static_assert(AFH_SOURCE_PAD % kChunk == 0, "packed sources must be whole chunks");

#define AFH_DIR_ENTRY(i, A, B, C, D) \
    { A, B, C, D, (float(i) - 1.0f) * kAngleDiv }
// t4010a1.rs:18-69, entries 0..9, and entry 10 for theta beyond 90 degrees.
__constant__ DirEntry c_dir_tab[11] = {
    AFH_DIR_ENTRY(0, 1.0f, 0.f, 0.f, 0.f),
    AFH_DIR_ENTRY(1, 1.0f, 0.f, 0.f, 0.f),
    AFH_DIR_ENTRY(2, 1.0f, 0.f, 0.f, 0.f),
    AFH_DIR_ENTRY(3, 1.0f, -0.2633589359691162f, -2.586746661956269f, 0.0009174516976956062f),
    AFH_DIR_ENTRY(4, 0.891250938f, -0.8910683499820542f, -1.0097633564074366f, 0.0001704413361662757f),
    AFH_DIR_ENTRY(5, 0.707945784f, -1.1924098538381465f, -0.7167962799004945f, 4.362034715924911f),
    AFH_DIR_ENTRY(6, 0.501187234f, -1.0439934287916925f, 0.027352097655185065f, -2.238476800444034f),
    AFH_DIR_ENTRY(7, 0.354813389f, -0.7015151838490918f, 0.39509604454751995f, 1.274774620435337f),
    AFH_DIR_ENTRY(8, 0.251188643f, -0.4471050803642494f, 1.0625664749725618f, -1.1276955684895724f),
    AFH_DIR_ENTRY(9, 0.199526231f, -0.179254124486545f, 0.008242027161100222f, -0.9016569137759131f),
    AFH_DIR_ENTRY(10, 0.0f, 0.f, 0.f, 0.f),
};

// Coordinates of point idx (< n_points).
template <int PTS>
__device__ __forceinline__ void point_at(const FieldArgs& a, uint32_t idx, float& x, float& y, float& z) {
    if (PTS == kScatter) {
        x = a.points[3 * (uint64_t)idx];
        y = a.points[3 * (uint64_t)idx + 1];
        z = a.points[3 * (uint64_t)idx + 2];
    } else {
#pragma clang fp contract(off)
        const uint32_t n0 = idx % a.count0;
        const uint32_t rest = idx / a.count0;
        const uint32_t n1 = rest % a.count1;
        const uint32_t n2 = rest / a.count1;
        float p[3] = {a.origin[0], a.origin[1], a.origin[2]};
        // grid_area.rs: origin[axis] + (n as Float * resolution), two roundings
        // (contraction is off in this block, so no fma)
        p[a.ax2] = a.origin[a.ax2] + (float)n2 * a.res;
        p[a.ax1] = a.origin[a.ax1] + (float)n1 * a.res;
        p[a.ax0] = a.origin[a.ax0] + (float)n0 * a.res;
        x = p[0];
        y = p[1];
        z = p[2];
    }
}

template <int SRC, int PTS, int OUT, int PREC, int PPL>
__global__ __launch_bounds__(kBlock) void field_kernel(FieldArgs a) {
    __shared__ float4 s_src[kTile * 3];
    __shared__ DirEntry s_dir[11];
    const int tid = threadIdx.x;
    if (SRC == kT4010A1 && tid < 11) s_dir[tid] = c_dir_tab[tid];

    const uint32_t base = blockIdx.x * (uint32_t)(kBlock * PPL) + tid;
    float px[PPL], py[PPL], pz[PPL];
    float sr[PPL], si[PPL], cr[PPL], ci[PPL];
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        uint32_t idx = base + k * kBlock;
        idx = idx < a.n_points ? idx : a.n_points - 1;  // tail lanes recompute the last point
        point_at<PTS>(a, idx, px[k], py[k], pz[k]);
        sr[k] = si[k] = cr[k] = ci[k] = 0.0f;
    }

    const float4* __restrict__ gsrc = reinterpret_cast<const float4*>(a.sources);
    for (uint32_t t0 = 0; t0 < a.n_sources; t0 += kTile) {
        const uint32_t cnt = min((uint32_t)kTile, a.n_sources - t0);
        __syncthreads();
        for (uint32_t v = tid; v < cnt * 3; v += kBlock) s_src[v] = gsrc[(uint64_t)t0 * 3 + v];
        __syncthreads();
        for (uint32_t j0 = 0; j0 < cnt; j0 += kChunk) {
            float tr[PPL], ti[PPL];
#pragma unroll
            for (int k = 0; k < PPL; ++k) tr[k] = ti[k] = 0.0f;
#pragma unroll 2
            for (int j = 0; j < kChunk; ++j) {
                const float4 q0 = s_src[(j0 + j) * 3];
                const float4 q1 = s_src[(j0 + j) * 3 + 1];
                const float4 q2 = s_src[(j0 + j) * 3 + 2];
                const Src s = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
#pragma unroll
                for (int k = 0; k < PPL; ++k) add_term<SRC, PREC>(s, px[k], py[k], pz[k], s_dir, tr[k], ti[k]);
            }
#pragma unroll
            for (int k = 0; k < PPL; ++k) {
                two_sum_acc(sr[k], cr[k], tr[k]);
                two_sum_acc(si[k], ci[k], ti[k]);
            }
        }
    }

#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        const uint32_t idx = base + k * kBlock;
        if (idx >= a.n_points) continue;
        const float re = sr[k] + cr[k];
        const float im = si[k] + ci[k];
        if (OUT == kComplex) {
            reinterpret_cast<float2*>(a.out)[idx] = make_float2(re, im);
        } else {
#pragma clang fp contract(off)
            // pressure_field.rs:35-37 norm_sqr().sqrt(); power_field.rs:35-37 norm_sqr()
            const float ns = re * re + im * im;
            a.out[idx] = (OUT == kAbs) ? sqrtf(ns) : ns;  // correctly rounded (HIP default)
        }
    }
}

template <int SRC, int PTS, int OUT, int PREC>
static hipError_t launch4(const FieldArgs& a, hipStream_t stream) {
    constexpr int PPL = 4;
    const uint32_t per_block = kBlock * PPL;
    const uint32_t blocks = (uint32_t)(((uint64_t)a.n_points + per_block - 1) / per_block);
    hipLaunchKernelGGL((field_kernel<SRC, PTS, OUT, PREC, PPL>), dim3(blocks), dim3(kBlock), 0, stream, a);
    return hipGetLastError();
}

template <int SRC, int PTS, int OUT>
static hipError_t launch_prec(const FieldArgs& a, int prec, hipStream_t s) {
    return prec == kAccurate ? launch4<SRC, PTS, OUT, kAccurate>(a, s) : launch4<SRC, PTS, OUT, kFast>(a, s);
}
template <int SRC, int PTS>
static hipError_t launch_out(const FieldArgs& a, int out, int prec, hipStream_t s) {
    switch (out) {
        case kAbs: return launch_prec<SRC, PTS, kAbs>(a, prec, s);
        case kAbs2: return launch_prec<SRC, PTS, kAbs2>(a, prec, s);
        default: return launch_prec<SRC, PTS, kComplex>(a, prec, s);
    }
}
template <int SRC>
static hipError_t launch_pts(const FieldArgs& a, int pts, int out, int prec, hipStream_t s) {
    return pts == kScatter ? launch_out<SRC, kScatter>(a, out, prec, s) : launch_out<SRC, kGrid>(a, out, prec, s);
}

hipError_t launch_field(const FieldArgs& a, int src, int pts, int out, int prec, hipStream_t stream) {
    if (a.n_points == 0) return hipSuccess;
    return src == kT4010A1 ? launch_pts<kT4010A1>(a, pts, out, prec, stream)
                           : launch_pts<kSphere>(a, pts, out, prec, stream);
}

}  // namespace afh
