// VALU issue-rate calibration for the field kernel's roofline (bench.py
// --calibrate-issue). Not part of the library's interface: a separate
// shared object with one C entry point.
//
// Each kind runs a long unrolled stream of one VALU instruction class on 16
// independent FP32 chains per lane (so that issue, not latency, bounds it)
// on W waves per SIMD (256 CUs x 4 SIMDs x W waves, 256-lane workgroups),
// the operands in VGPRs as in the field kernel:
//   0  v_fma_f32 / v_fmac_f32   a = a * x + y
//   1  v_mul_f32                a = a * x
//   2  v_sin_f32                a = sin(2 pi a)
//   3  v_rsq_f32                a = 1 / sqrt(a)   (a stays in (0.5, 2))
//   4  v_exp_f32                a = 2^a           (a in (-1, 0) first, then finite)
//   5  9 v_fma_f32 : 1 v_sin_f32
//   6  v_fmac_f32 (VOP2)        a = x * y + a     (the accumulator is the destination)
//   7  v_fma_f32, one constant  a = a * x + 0.5
//   8  v_mul_f32 (VOP3)         a = |a| * x       (a source modifier forces VOP3)
// Every wave reads the shader clock (s_memtime) and the 100 MHz constant
// clock (s_memrealtime) around its stream; the host reports the kernel time
// (HIP events), the shader clock frequency and the SIMD cycles per wave64
// instruction, (kernel time x clock x SIMDs) / (wave instructions).
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

constexpr int kChains = 16;
constexpr int kInner = 8;  // instructions per chain per repetition

template <int KIND>
__device__ __forceinline__ float step(float a, float x, float y) {
    if constexpr (KIND == 0) return fmaf(a, x, y);
    if constexpr (KIND == 1) return a * x;
    if constexpr (KIND == 2) return __builtin_amdgcn_sinf(a);
    if constexpr (KIND == 3) return __builtin_amdgcn_rsqf(a);
    if constexpr (KIND == 4) return __builtin_amdgcn_exp2f(a);
    if constexpr (KIND == 6) return fmaf(x, y, a);
    if constexpr (KIND == 7) return fmaf(a, x, 0.5f);
    return __builtin_fabsf(a) * x;
}

template <int KIND>
__global__ __launch_bounds__(256) void issue_kernel(uint32_t reps, float seed, float* sink,
                                                    unsigned long long* clocks) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const float lane = (float)(threadIdx.x & 63) * 1e-3f;
    float a[kChains];
#pragma unroll
    for (int i = 0; i < kChains; ++i)
        a[i] = KIND == 3 ? 1.0f + lane + 0.01f * i : (KIND == 4 ? -0.5f - lane : seed + lane + 0.01f * i);
    const float x = 0.999f + lane * 1e-3f, y = 1e-3f * seed;
    const unsigned long long c0 = clock64(), w0 = wall_clock64();
    for (uint32_t r = 0; r < reps; ++r) {
#pragma unroll
        for (int u = 0; u < kInner; ++u) {
            if constexpr (KIND == 5) {
                if (u == kInner - 1) {
#pragma unroll
                    for (int i = 0; i < kChains; ++i) a[i] = __builtin_amdgcn_sinf(a[i]);
                } else {
#pragma unroll
                    for (int i = 0; i < kChains; ++i) a[i] = fmaf(a[i], x, y);
                }
            } else {
#pragma unroll
                for (int i = 0; i < kChains; ++i) a[i] = step<KIND>(a[i], x, y);
            }
        }
    }
    const unsigned long long c1 = clock64(), w1 = wall_clock64();
    float s = 0.0f;
#pragma unroll
    for (int i = 0; i < kChains; ++i) s += a[i];
    sink[gid] = s;
    if ((threadIdx.x & 63) == 0) {
        const uint32_t wave = gid >> 6;
        clocks[2 * wave] = c1 - c0;
        clocks[2 * wave + 1] = w1 - w0;
    }
}

template <int KIND>
hipError_t launch(uint32_t blocks, uint32_t reps, float* sink, unsigned long long* clocks, hipStream_t st) {
    hipLaunchKernelGGL(issue_kernel<KIND>, dim3(blocks), dim3(256), 0, st, reps, 0.25f, sink, clocks);
    return hipGetLastError();
}

}  // namespace

// out[0] kernel ms (HIP events), out[1] shader clock GHz (median wave),
// out[2] wave64 VALU instructions issued in the timed streams, out[3] SIMD
// cycles per wave64 instruction from the kernel time, out[4] the same from
// the median wave's own clock (its stream shared the SIMD with W - 1 others).
extern "C" int afi_issue_rate(int kind, int waves_per_simd, int reps, double* out) {
    if (kind < 0 || kind > 8 || waves_per_simd < 1 || waves_per_simd > 8 || reps < 1 || !out) return -1;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return -2;
    const uint32_t cus = (uint32_t)prop.multiProcessorCount;
    const uint32_t blocks = cus * (uint32_t)waves_per_simd;  // 4 waves per block, one per SIMD
    const uint32_t waves = blocks * 4;
    float* sink = nullptr;
    unsigned long long* clocks = nullptr;
    if (hipMalloc(&sink, (size_t)blocks * 256 * sizeof(float)) != hipSuccess) return -3;
    if (hipMalloc(&clocks, (size_t)waves * 2 * sizeof(unsigned long long)) != hipSuccess) {
        (void)hipFree(sink);
        return -3;
    }
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    auto run = [&](int r) -> hipError_t {
        switch (kind) {
            case 0: return launch<0>(blocks, (uint32_t)r, sink, clocks, nullptr);
            case 1: return launch<1>(blocks, (uint32_t)r, sink, clocks, nullptr);
            case 2: return launch<2>(blocks, (uint32_t)r, sink, clocks, nullptr);
            case 3: return launch<3>(blocks, (uint32_t)r, sink, clocks, nullptr);
            case 4: return launch<4>(blocks, (uint32_t)r, sink, clocks, nullptr);
            case 5: return launch<5>(blocks, (uint32_t)r, sink, clocks, nullptr);
            case 6: return launch<6>(blocks, (uint32_t)r, sink, clocks, nullptr);
            case 7: return launch<7>(blocks, (uint32_t)r, sink, clocks, nullptr);
            default: return launch<8>(blocks, (uint32_t)r, sink, clocks, nullptr);
        }
    };
    int rc = 0;
    float ms = 0.0f;
    if (run(1) != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = -4;  // warm-up
    if (!rc && (hipEventRecord(e0, nullptr) != hipSuccess || run(reps) != hipSuccess ||
                hipEventRecord(e1, nullptr) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        rc = -5;
    if (!rc) {
        unsigned long long* h = new unsigned long long[(size_t)waves * 2];
        if (hipMemcpy(h, clocks, (size_t)waves * 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost) != hipSuccess) {
            rc = -6;
        } else {
            // median wave by shader cycles
            unsigned long long* cyc = new unsigned long long[waves];
            for (uint32_t w = 0; w < waves; ++w) cyc[w] = h[2 * w];
            uint32_t med = 0;
            {
                // partial selection of the median index
                unsigned long long* tmp = new unsigned long long[waves];
                for (uint32_t w = 0; w < waves; ++w) tmp[w] = cyc[w];
                for (uint32_t i = 0; i <= waves / 2; ++i) {
                    uint32_t m = i;
                    for (uint32_t j = i + 1; j < waves; ++j)
                        if (tmp[j] < tmp[m]) m = j;
                    const unsigned long long t = tmp[i];
                    tmp[i] = tmp[m];
                    tmp[m] = t;
                }
                const unsigned long long mv = tmp[waves / 2];
                for (uint32_t w = 0; w < waves; ++w)
                    if (cyc[w] == mv) {
                        med = w;
                        break;
                    }
                delete[] tmp;
            }
            const double wave_cyc = (double)h[2 * med], wave_ns = (double)h[2 * med + 1] * 10.0;
            const double ghz = wave_cyc / wave_ns;
            const double per_wave = (double)reps * kInner * kChains;
            const double instr = per_wave * waves;
            out[0] = ms;
            out[1] = ghz;
            out[2] = instr;
            out[3] = (double)ms * 1e6 * ghz * (double)(cus * 4) / instr;
            out[4] = wave_cyc / (per_wave * waves_per_simd);
            delete[] cyc;
        }
        delete[] h;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(sink);
    (void)hipFree(clocks);
    return rc;
}
