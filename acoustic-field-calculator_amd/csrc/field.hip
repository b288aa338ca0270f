// Field kernel: complex pressure of all sources at every observation point.
//
// Replaces CpuCalculator::calculate (calculator/cpu_calculator.rs:43-65) over
// UniformSystem::propagate_all (system/uniform_medium.rs:122-128) and the
// Vulkan shaders assets/shaders/cpressure_base.comp:43-71 (+ cpressure.comp,
// pressure.comp, power.comp output stages).
//
// Layout and schedule (gfx950):
//   * workgroup = 256 lanes (4 wave64); each lane owns PPL observation points
//     (PPL = 1) strided by 256 so loads of scatter points and all output
//     stores coalesce; on grids whose slowest axis is z a lane owns a run of
//     2 points along z instead (field_zrun_kernel), which share the x / y
//     parts of every term;
//   * grid points are generated in registers from (origin, resolution,
//     counts), with the reference's own f32 rounding (origin + n * res, no
//     contraction), so no point array is read from HBM;
//   * packed sources (12 floats each) are staged through LDS in tiles of
//     kTile; every lane of the workgroup reads the same LDS address
//     (broadcast), so the per-term operands cost one ds_read_b128 x3 per
//     source per wave;
//   * sums: plain FP32 over chunks of kChunk sources, then TwoSum-compensated
//     across chunks, so the rounding error does not grow with N.
#include <math.h>
#include <stdlib.h>

#include <mutex>

#include "afh_internal.hpp"

namespace afh {

constexpr int kBlock = 256;
constexpr int kTile = 64;
constexpr int kChunk = 32;
static_assert(kTile % kChunk == 0, "tile must hold whole chunks");
static_assert(AFH_SOURCE_PAD % kChunk == 0, "packed sources must be whole chunks");

// t4010a1.rs:18-69 as device constants (ACCURATE mode) and a host copy.
__constant__ DirEntry c_dir_tab[11] = {AFH_DIR_ENTRIES(AFH_DIR_ENTRY)};
static const DirEntry h_dir_tab[11] = {AFH_DIR_ENTRIES(AFH_DIR_ENTRY)};

// Reference directivity (t4010a1.rs:107-120) in double on the f32 constants.
static double dir_ref(double theta, int seg) {
    static const float A[10] = {0.f, 1.0f, 1.0f, 1.0f, 0.891250938f, 0.707945784f, 0.501187234f, 0.354813389f,
                                0.251188643f, 0.199526231f};
    static const float B[10] = {0.f, 0.f, 0.f, -0.2633589359691162f, -0.8910683499820542f, -1.1924098538381465f,
                                -1.0439934287916925f, -0.7015151838490918f, -0.4471050803642494f,
                                -0.179254124486545f};
    static const float C[10] = {0.f, 0.f, 0.f, -2.586746661956269f, -1.0097633564074366f, -0.7167962799004945f,
                                0.027352097655185065f, 0.39509604454751995f, 1.0625664749725618f,
                                0.008242027161100222f};
    static const float D[10] = {0.f, 0.f, 0.f, 0.0009174516976956062f, 0.0001704413361662757f,
                                4.362034715924911f, -2.238476800444034f, 1.274774620435337f,
                                -1.1276955684895724f, -0.9016569137759131f};
    if (seg <= 0) return 1.0;
    if (seg > 9) return 0.0;
    const double x = theta - (seg - 1) * (double)kAngleDiv;
    return A[seg] + (B[seg] + (C[seg] + D[seg] * x) * x) * x;
}

// Quadratic in t through Chebyshev nodes of [t0, t1] of D on segment `seg`,
// with c = (cell_lo + t) / 512.
static void fit_piece(double cell_lo, double t0, double t1, int seg, float* p) {
    double tn[3], v[3];
    for (int k = 0; k < 3; ++k) {
        tn[k] = 0.5 * (t0 + t1) + 0.5 * (t1 - t0) * cos((2 * k + 1) * M_PI / 6.0);
        const double c = (cell_lo + tn[k]) / 512.0;
        v[k] = dir_ref(acos(c < -1 ? -1 : (c > 1 ? 1 : c)), seg);
    }
    // Newton divided differences -> monomial coefficients
    const double d01 = (v[1] - v[0]) / (tn[1] - tn[0]);
    const double d12 = (v[2] - v[1]) / (tn[2] - tn[1]);
    const double d012 = (d12 - d01) / (tn[2] - tn[0]);
    // v0 + d01 (t - t0) + d012 (t - t0)(t - t1)
    p[2] = (float)d012;
    p[1] = (float)(d01 - d012 * (tn[0] + tn[1]));
    p[0] = (float)(v[0] - d01 * tn[0] + d012 * tn[0] * tn[1]);
}

// Segment of the reference for c (theta = acos c), i = ceil(theta / ANGLE_DIV).
static int segment_of(double c) {
    const double th = acos(c < -1 ? -1 : (c > 1 ? 1 : c));
    const double q = ceil(th / (double)kAngleDiv);
    return q <= 0 ? 0 : (q > 10 ? 10 : (int)q);
}

void build_dir_cells(DirCell* cells) {
    const float two24 = 16777216.0f;
    for (int j = 0; j < kDirCells; ++j) {
        DirCell& e = cells[j];
        e.p0 = e.p1 = e.p2 = 0.f;
        e.q0 = e.q1 = e.q2 = 0.f;
        e.k = -2.0f * two24;  // step = 0 everywhere in the cell
        e.s = two24;
        if (j == 0) continue;  // D = 0
        if (j == kDirCells - 1) {
            e.p0 = 1.f;
            continue;
        }
        const double lo = (double)(j - 2);  // in units of 1/512
        // boundaries cos(k * ANGLE_DIV), k = 2..9, inside this cell
        double split = -1;
        for (int k = 2; k <= 9; ++k) {
            const double cb = cos(k * (double)kAngleDiv) * 512.0 - lo;
            if (cb > 0.0 && cb < 1.0) split = cb;
        }
        if (split < 0) {
            fit_piece(lo, 0.0, 1.0, segment_of((lo + 0.5) / 512.0), &e.p0);
        } else {
            // ts on the 2^-24 grid, so that step(t) = clamp(2^24 (t - ts) + 1)
            // is exactly 0 or 1 for t on that grid
            const double ts = nearbyint(split * 16777216.0) / 16777216.0;
            float b[3];
            fit_piece(lo, 0.0, ts, segment_of((lo + 0.5 * ts) / 512.0), &e.p0);
            fit_piece(lo, ts, 1.0, segment_of((lo + 0.5 * (1.0 + ts)) / 512.0), b);
            e.q0 = (float)((double)b[0] - (double)e.p0);
            e.q1 = (float)((double)b[1] - (double)e.p1);
            e.q2 = (float)((double)b[2] - (double)e.p2);
            e.k = (float)(1.0 - ts * 16777216.0);  // integer below 2^24: exact
        }
    }
}

// Per-device copy of the cell table (device memory, uploaded once; the
// kernels stage it into LDS).
const DirCell* dir_cells_device() {
    static std::mutex mu;
    static DirCell* dev_tab[64] = {};
    static DirCell host[kDirCells];
    static bool built = false;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    std::lock_guard<std::mutex> g(mu);
    if (dev_tab[dev]) return dev_tab[dev];
    if (!built) {
        build_dir_cells(host);
        built = true;
    }
    DirCell* p = nullptr;
    if (hipMalloc(&p, sizeof(host)) != hipSuccess) return nullptr;
    if (hipMemcpy(p, host, sizeof(host), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(p);
        return nullptr;
    }
    dev_tab[dev] = p;
    return p;
}

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

// Copies the FAST-mode cell table into LDS as two float4 arrays (the first
// and second half of each DirCell).
__device__ __forceinline__ void stage_cells(const DirCell* __restrict__ cells, float4* cp, float4* cq) {
    const float4* g = reinterpret_cast<const float4*>(cells);
    for (int j = threadIdx.x; j < kDirCells; j += kBlock) {
        cp[j] = g[2 * j];
        cq[j] = g[2 * j + 1];
    }
}

// Copies cnt packed sources (3 float4 each) into an LDS tile: unchanged for
// the ACCURATE mode, as SrcF (the FAST mode's derived operands) otherwise.
template <int PREC>
__device__ __forceinline__ void stage_tile(const float4* __restrict__ g, uint32_t cnt, float4* tile) {
    if (PREC == kFast) {
        for (uint32_t j = threadIdx.x; j < cnt; j += kBlock) {
            const float4 q0 = g[3 * j], q1 = g[3 * j + 1], q2 = g[3 * j + 2];
            const SrcF f = stage_fast({q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w});
            tile[3 * j] = make_float4(f.x, f.y, f.z, f.la);
            tile[3 * j + 1] = make_float4(f.nx, f.ny, f.nz, f.ph);
            tile[3 * j + 2] = make_float4(f.kh, f.kl, f.nal, f.kh2);
        }
    } else {
        for (uint32_t v = threadIdx.x; v < cnt * 3; v += kBlock) tile[v] = g[v];
    }
}

template <int SRC, int PTS, int OUT, int PREC, int PPL>
__global__ __launch_bounds__(kBlock) void field_kernel(FieldArgs a) {
    constexpr int kCells = (SRC == kT4010A1 && PREC == kFast) ? kDirCells : 1;
    __shared__ float4 s_src[kTile * 3];
    __shared__ DirEntry s_dir[11];
    __shared__ float4 s_cp[kCells], s_cq[kCells];
    const DirCell* s_cells = nullptr;  // the ACCURATE mode reads s_dir instead
    const int tid = threadIdx.x;
    if (SRC == kT4010A1 && PREC == kAccurate && tid < 11) s_dir[tid] = c_dir_tab[tid];
    if (SRC == kT4010A1 && PREC == kFast) stage_cells(a.cells, s_cp, s_cq);

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
        stage_tile<PREC>(gsrc + (uint64_t)t0 * 3, cnt, s_src);
        __syncthreads();
        for (uint32_t j0 = 0; j0 < cnt; j0 += kChunk) {
            float tr[PPL], ti[PPL];
#pragma unroll
            for (int k = 0; k < PPL; ++k) tr[k] = ti[k] = 0.0f;
#pragma unroll 4
            for (int j = 0; j < kChunk; ++j) {
                const float4 q0 = s_src[(j0 + j) * 3];
                const float4 q1 = s_src[(j0 + j) * 3 + 1];
                const float4 q2 = s_src[(j0 + j) * 3 + 2];
                if (PREC == kFast) {
                    const SrcF s = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
                    add_terms_staged<SRC, PPL>(s, px, py, pz, s_cp, s_cq, tr, ti);
                } else {
                    const Src s = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
#pragma unroll
                    for (int k = 0; k < PPL; ++k)
                        add_term<SRC, PREC>(s, px[k], py[k], pz[k], s_dir, s_cells, tr[k], ti[k]);
                }
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

// Grid variant for grids whose slowest axis is z (FAST mode): lane
// (n0, n1, q) owns the ZR points n2 = ZR q + k of one column, which share x
// and y, so add_terms_zrun forms dx, dy, dx^2 + dy^2 and the x / y part of
// n . (p - s) once per source for all ZR of them (and each source's LDS
// operands are read once per ZR terms). Consecutive lanes take
// consecutive n0 (coalesced stores, one store per point run row). Same
// per-term arithmetic, source order and compensated sums as field_kernel:
// every point gets the bits it gets there.
template <int SRC, int OUT, int ZR>
__global__ __launch_bounds__(kBlock) void field_zrun_kernel(FieldArgs a) {
    constexpr int kCells = SRC == kT4010A1 ? kDirCells : 1;
    __shared__ float4 s_src[kTile * 3];
    __shared__ float4 s_cp[kCells], s_cq[kCells];
    const int tid = threadIdx.x;
    if (SRC == kT4010A1) stage_cells(a.cells, s_cp, s_cq);

    const uint32_t cols = a.count0 * a.count1;
    const uint32_t count2 = a.n_points / cols;
    const uint32_t runs = (count2 + ZR - 1) / ZR;
    const uint64_t total = (uint64_t)cols * runs;
    const uint64_t g = (uint64_t)blockIdx.x * kBlock + tid;
    const uint64_t gc = g < total ? g : total - 1;  // tail lanes recompute the last run
    const uint32_t n0 = (uint32_t)(gc % a.count0);
    const uint64_t rest = gc / a.count0;
    const uint32_t n1 = (uint32_t)(rest % a.count1);
    const uint32_t q = (uint32_t)(rest / a.count1);
    float px, py, pz[ZR];
    {
#pragma clang fp contract(off)
        // grid_area.rs: origin[axis] + (n as Float * resolution), two roundings
        float p[3] = {a.origin[0], a.origin[1], a.origin[2]};
        p[a.ax1] = a.origin[a.ax1] + (float)n1 * a.res;
        p[a.ax0] = a.origin[a.ax0] + (float)n0 * a.res;
        px = p[0];
        py = p[1];
#pragma unroll
        for (int k = 0; k < ZR; ++k) {
            const uint32_t n2 = min(q * ZR + k, count2 - 1);
            pz[k] = a.origin[2] + (float)n2 * a.res;
        }
    }
    float sr[ZR], si[ZR], cr[ZR], ci[ZR];
#pragma unroll
    for (int k = 0; k < ZR; ++k) sr[k] = si[k] = cr[k] = ci[k] = 0.0f;

    const float4* __restrict__ gsrc = reinterpret_cast<const float4*>(a.sources);
    for (uint32_t t0 = 0; t0 < a.n_sources; t0 += kTile) {
        const uint32_t cnt = min((uint32_t)kTile, a.n_sources - t0);
        __syncthreads();
        stage_tile<kFast>(gsrc + (uint64_t)t0 * 3, cnt, s_src);
        __syncthreads();
        for (uint32_t j0 = 0; j0 < cnt; j0 += kChunk) {
            float tr[ZR], ti[ZR];
#pragma unroll
            for (int k = 0; k < ZR; ++k) tr[k] = ti[k] = 0.0f;
#pragma unroll 2
            for (int j = 0; j < kChunk; ++j) {
                const float4 q0 = s_src[(j0 + j) * 3];
                const float4 q1 = s_src[(j0 + j) * 3 + 1];
                const float4 q2 = s_src[(j0 + j) * 3 + 2];
                const SrcF sf = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
                add_terms_zrun<SRC, ZR>(sf, px, py, pz, s_cp, s_cq, tr, ti);
            }
#pragma unroll
            for (int k = 0; k < ZR; ++k) {
                two_sum_acc(sr[k], cr[k], tr[k]);
                two_sum_acc(si[k], ci[k], ti[k]);
            }
        }
    }

#pragma unroll
    for (int k = 0; k < ZR; ++k) {
        const uint32_t n2 = q * ZR + k;
        if (g >= total || n2 >= count2) continue;
        const uint32_t idx = n0 + a.count0 * (n1 + a.count1 * n2);
        const float re = sr[k] + cr[k];
        const float im = si[k] + ci[k];
        if (OUT == kComplex) {
            reinterpret_cast<float2*>(a.out)[idx] = make_float2(re, im);
        } else {
#pragma clang fp contract(off)
            // pressure_field.rs:35-37 norm_sqr().sqrt(); power_field.rs:35-37 norm_sqr()
            const float ns = re * re + im * im;
            a.out[idx] = (OUT == kAbs) ? sqrtf(ns) : ns;
        }
    }
}

// Source-split variant for small point sets: each wave owns PPW points and its
// 64 lanes split the sources (lane l takes sources l, l + 64, ... of each LDS
// tile); per-lane sums are TwoSum-compensated per tile and the 64 partial
// sums are combined by xor shuffles in a fixed order. Small grids then spread
// over many waves instead of leaving most CUs idle, and the result of a point
// set depends only on the point set itself.
constexpr int kTileS = 256;  // sources per LDS tile (4 per lane)
constexpr int kSplitPoints = 4;  // points per wave

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

template <int SRC, int PTS, int OUT, int PREC, int PPW>
__global__ __launch_bounds__(kBlock) void field_split_kernel(FieldArgs a) {
    constexpr int kCells = (SRC == kT4010A1 && PREC == kFast) ? kDirCells : 1;
    __shared__ float4 s_src[kTileS * 3];
    __shared__ DirEntry s_dir[11];
    __shared__ float4 s_cp[kCells], s_cq[kCells];
    const DirCell* s_cells = nullptr;  // the ACCURATE mode reads s_dir instead
    const int tid = threadIdx.x, lane = tid & 63;
    if (SRC == kT4010A1 && PREC == kAccurate && tid < 11) s_dir[tid] = c_dir_tab[tid];
    if (SRC == kT4010A1 && PREC == kFast) stage_cells(a.cells, s_cp, s_cq);
    const uint32_t wave = blockIdx.x * (kBlock / 64) + (tid >> 6);
    const uint32_t base = wave * PPW;
    float px[PPW], py[PPW], pz[PPW];
    float sr[PPW], si[PPW], cr[PPW], ci[PPW];
#pragma unroll
    for (int k = 0; k < PPW; ++k) {
        uint32_t idx = base + k;
        idx = idx < a.n_points ? idx : a.n_points - 1;
        point_at<PTS>(a, idx, px[k], py[k], pz[k]);
        sr[k] = si[k] = cr[k] = ci[k] = 0.0f;
    }
    const float4* __restrict__ gsrc = reinterpret_cast<const float4*>(a.sources);
    for (uint32_t t0 = 0; t0 < a.n_sources; t0 += kTileS) {
        const uint32_t cnt = min((uint32_t)kTileS, a.n_sources - t0);
        __syncthreads();
        stage_tile<PREC>(gsrc + (uint64_t)t0 * 3, cnt, s_src);
        __syncthreads();
        float tr[PPW], ti[PPW];
#pragma unroll
        for (int k = 0; k < PPW; ++k) tr[k] = ti[k] = 0.0f;
        // cnt is a multiple of 32 (packed padding): lanes past it idle
        for (uint32_t j = lane; j < cnt; j += 64) {
            const float4 q0 = s_src[j * 3], q1 = s_src[j * 3 + 1], q2 = s_src[j * 3 + 2];
            if (PREC == kFast) {
                const SrcF src = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
#pragma unroll
                for (int k = 0; k < PPW; k += 2)
                    add_terms_staged<SRC, 2>(src, px + k, py + k, pz + k, s_cp, s_cq, tr + k, ti + k);
            } else {
                const Src src = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
#pragma unroll
                for (int k = 0; k < PPW; ++k)
                    add_term<SRC, PREC>(src, px[k], py[k], pz[k], s_dir, s_cells, tr[k], ti[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < PPW; ++k) {
            two_sum_acc(sr[k], cr[k], tr[k]);
            two_sum_acc(si[k], ci[k], ti[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < PPW; ++k) {
        const float re = wave_sum_f(sr[k] + cr[k]);
        const float im = wave_sum_f(si[k] + ci[k]);
        const uint32_t idx = base + k;
        if (lane == k && idx < a.n_points) {
            if (OUT == kComplex) {
                reinterpret_cast<float2*>(a.out)[idx] = make_float2(re, im);
            } else {
#pragma clang fp contract(off)
                const float ns = re * re + im * im;
                a.out[idx] = (OUT == kAbs) ? sqrtf(ns) : ns;
            }
        }
    }
}

// Point sets below this size use the source-split kernel (AFH_FIELD_SPLIT_BELOW
// overrides it, for measurements).
static uint32_t split_threshold() {
    static uint32_t thr = [] {
        const char* e = getenv("AFH_FIELD_SPLIT_BELOW");
        // below ~2e5 points the point-parallel grid leaves CUs idle (measured
        // on MI355X with 3984 sources: 6.5x slower at 1e4 points, even at
        // 2.3e5, 1.05x faster at 3e5)
        return e ? (uint32_t)strtoul(e, nullptr, 10) : (uint32_t)200000;
    }();
    return thr;
}

template <int SRC, int PTS, int OUT, int PREC, int PPL>
static hipError_t launch_ppl(const FieldArgs& a, hipStream_t stream) {
    const uint32_t per_block = kBlock * PPL;
    const uint32_t blocks = (uint32_t)(((uint64_t)a.n_points + per_block - 1) / per_block);
    hipLaunchKernelGGL((field_kernel<SRC, PTS, OUT, PREC, PPL>), dim3(blocks), dim3(kBlock), 0, stream, a);
    return hipGetLastError();
}

template <int SRC, int PTS, int OUT, int PREC>
static hipError_t launch4(const FieldArgs& a, hipStream_t stream) {
    if (a.n_points < split_threshold()) {
        constexpr int PPW = kSplitPoints;
        const uint32_t per_block = (kBlock / 64) * PPW;
        const uint32_t blocks = (uint32_t)(((uint64_t)a.n_points + per_block - 1) / per_block);
        hipLaunchKernelGGL((field_split_kernel<SRC, PTS, OUT, PREC, PPW>), dim3(blocks), dim3(kBlock), 0, stream, a);
        return hipGetLastError();
    }
    if (PTS == kGrid && PREC == kFast && a.ax2 == 2) {
        // z runs of 2 points per lane (AFH_FIELD_ZRUN=1 turns them off, for
        // measurements). Measured on MI355X, 3984 sources x 256 x 256 x 128:
        // 2 points (63 VGPRs, 8 waves per SIMD) 6.5% faster than one point per
        // lane; 4 points (94-128 VGPRs, 4-5 waves) gained less or nothing
        static const bool zrun = [] {
            const char* e = getenv("AFH_FIELD_ZRUN");
            return !(e && atoi(e) == 1);
        }();
        if (zrun) {
            constexpr int ZR = 2;
            const uint32_t cols = a.count0 * a.count1;
            const uint64_t total = (uint64_t)cols * ((a.n_points / cols + ZR - 1) / ZR);
            const uint32_t blocks = (uint32_t)((total + kBlock - 1) / kBlock);
            hipLaunchKernelGGL((field_zrun_kernel<SRC, OUT, ZR>), dim3(blocks), dim3(kBlock), 0, stream, a);
            return hipGetLastError();
        }
    }
    // one point per lane: few VGPRs, 8 waves per SIMD
    return launch_ppl<SRC, PTS, OUT, PREC, 1>(a, stream);
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

hipError_t launch_field(const FieldArgs& a_in, int src, int pts, int out, int prec, hipStream_t stream) {
    if (a_in.n_points == 0) return hipSuccess;
    FieldArgs a = a_in;
    if (src == kT4010A1 && prec == kFast) {
        a.cells = dir_cells_device();
        if (!a.cells) return hipErrorOutOfMemory;
    }
    return src == kT4010A1 ? launch_pts<kT4010A1>(a, pts, out, prec, stream)
                           : launch_pts<kSphere>(a, pts, out, prec, stream);
}

}  // namespace afh

// Host evaluation of the device directivity formulas, for tests: mode
// AFH_PREC_FAST reads the cell table, AFH_PREC_ACCURATE evaluates the
// reference binning and cubic (both exactly as the kernels do).
extern "C" int32_t AFH_DirectivityT4010A1(const float* cos_theta, uint64_t n, int32_t mode, float* out) {
    using namespace afh;
    static DirCell cells[kDirCells];
    static std::once_flag once;
    std::call_once(once, [] { build_dir_cells(cells); });
    if (((!cos_theta || !out) && n) || (mode != AFH_PREC_FAST && mode != AFH_PREC_ACCURATE)) return AFH_ERR_ARG;
    for (uint64_t i = 0; i < n; ++i)
        out[i] = mode == AFH_PREC_ACCURATE ? directivity_band(cos_theta[i], h_dir_tab)
                                           : directivity_table(cos_theta[i], cells);
    return AFH_OK;
}
