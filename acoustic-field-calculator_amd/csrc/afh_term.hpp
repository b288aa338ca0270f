// Device-side source-to-point term for gfx950.
//
// Reference: system/uniform_medium.rs:111-120 (same as
// acoustic-field-optimizer/src/multiple_foci/macros.rs:17-30):
//     diff = t - s; r = |diff|; theta = acos(dir.diff / r); D = directivity(theta)
//     term = amp * D * exp(-r * alpha) / r * exp(i (phase + k r))
//
// Formulation here (FP32 throughout):
//   * r comes from the hardware rsqrt plus one Newton residual, kept as a
//     (r, r_lo) pair so the phase sees a distance accurate to well below an ulp;
//   * the phase is carried in revolutions: k r / (2 pi) with k / (2 pi) split
//     into hi + lo floats and the product's rounding error recovered with an
//     fma, the integer part dropped exactly with v_fract; phase / (2 pi) is
//     precomputed per source modulo 1;
//   * exp(-alpha r) = exp2(-(alpha log2 e) r) on v_exp_f32;
//   * sin / cos: v_sin_f32 / v_cos_f32 (take revolutions) in the FAST mode, a
//     quadrant-reduced Taylor polynomial in the ACCURATE mode;
//   * T4010A1 directivity (t4010a1.rs:107-120): the ACCURATE mode takes
//     theta = acos(c) and the reference's piecewise cubic with its
//     ceil(theta / (pi/18)) binning (D = 1 for c >= 0.94, i.e. theta <= 20 deg
//     where segments 0-2 are identically 1; D = 0 for c <= -0.001); the FAST
//     mode reads D(c) from a cell table of quadratic pieces (DirCell).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace afh {

enum SourceModel { kSphere = 0, kT4010A1 = 1 };
enum Precision { kFast = 0, kAccurate = 1 };

// Packed per-source parameters (AFH_SOURCE_STRIDE = 12 floats).
struct Src {
    float x, y, z, nx, ny, nz, amp, ph, kh, kl, al, pad;
};

// f32 constants of t4010a1.rs: ANGLE_DIV = PI_f32 / 18 in f32 arithmetic.
constexpr float kPiF = 3.14159265358979323846f;
constexpr float kAngleDiv = kPiF / 18.0f;
constexpr float kPio2Hi = 1.57079637050628662109375f;  // nearest f32 to pi/2
constexpr float kPio2Lo = -4.37113900018624283e-8f;    // pi/2 - kPio2Hi
constexpr float kCosBandHi = 0.94f;                    // > cos(20 deg) = 0.93969
constexpr float kCosBandLo = -0.001f;                  // < cos(90 deg)

// Directivity table entry i (i = ceil(theta / ANGLE_DIV) clamped to 0..10):
// {a, b, c, d} and the segment start (i - 1) * ANGLE_DIV computed in f32.
struct DirEntry {
    float a, b, c, d, start;
};

// Directivity as a function of c = cos(theta), for the FAST mode. Cells of
// width exactly 1/512 in c, indexed by u = 512 c (exact): cell j
// (1..kDirCells-2) covers u in [j - 2, j - 1); cell 0 holds u < -1
// (theta > 90 deg, D = 0) and the last cell u >= 482 (theta < 20 deg, D = 1).
// Inside a cell, with t = u - floor(u) in [0, 1), D = P(t) + step(t) Q(t)
// with quadratics P, Q and step(t) = clamp(t s + k, 0, 1): a cell that
// contains a segment boundary of the reference's piecewise cubic (theta = 20,
// 30, ..., 90 deg) at t = ts has s = 2^24 and k = 1 - 2^24 ts (ts rounded to a
// multiple of 2^-24), so step is 0 below ts and 1 from ts on and the jumps of
// the table are kept; every other cell has Q = 0 and k = -2^25. The form
// needs no compare or select (no VCC traffic in the kernels' inner loops).
// Max deviation from the reference formula evaluated in FP64: 1.2e-7.
struct alignas(16) DirCell {
    float p0, p1, p2, k;  // P(t) = p0 + p1 t + p2 t^2, step offset
    float q0, q1, q2, s;  // Q(t) = q0 + q1 t + q2 t^2 (piece after ts minus P), step slope
};
constexpr int kDirCells = 485;
constexpr float kDirScale = 512.0f;

// acos for c in [-0.01, 1]: asin polynomial of the Cephes Math Library
// (S. L. Moshier, asinf.c, minimax on [0, 0.5]) with the half-angle identity
// acos(c) = 2 asin(sqrt((1 - c) / 2)) for c > 0.5 and pi/2 - asin(c) below.
__host__ __device__ inline float acos_band(float c) {
    const bool big = c > 0.5f;
    const float z = big ? fmaf(-0.5f, c, 0.5f) : c * c;
#if defined(__HIP_DEVICE_COMPILE__)
    const float s = big ? __builtin_amdgcn_sqrtf(z) : c;
#else
    const float s = big ? sqrtf(z) : c;
#endif
    float p = fmaf(4.2163199048e-2f, z, 2.4181311049e-2f);
    p = fmaf(p, z, 4.5470025998e-2f);
    p = fmaf(p, z, 7.4953002686e-2f);
    p = fmaf(p, z, 1.6666752422e-1f);
    const float a = fmaf(p * z, s, s);  // asin(s)
    return big ? 2.0f * a : (kPio2Hi - a) + kPio2Lo;
}

// sin and cos of 2 pi u: u reduced to the nearest quarter turn, then Taylor
// series of sin (to x^9) and cos (to x^10) on |x| <= pi/4, whose truncation
// errors are below 2e-9.
__host__ __device__ inline void sincos_rev_poly(float u, float* s, float* c) {
    const float q = rintf(4.0f * u);
    const float w = fmaf(-0.25f, q, u);  // exact
    const float x = w * 6.28318530717958647692f;
    const float x2 = x * x;
    float sp = fmaf(x2, 2.7557319223985893e-6f, -1.9841269841269841e-4f);
    sp = fmaf(sp, x2, 8.3333333333333333e-3f);
    sp = fmaf(sp, x2, -1.6666666666666667e-1f);
    sp = fmaf(sp * x2, x, x);
    float cp = fmaf(x2, -2.7557319223985888e-7f, 2.4801587301587302e-5f);
    cp = fmaf(cp, x2, -1.3888888888888889e-3f);
    cp = fmaf(cp, x2, 4.1666666666666667e-2f);
    cp = fmaf(cp, x2, -0.5f);
    cp = fmaf(cp, x2, 1.0f);
    const int qi = ((int)q) & 3;
    const float s0 = (qi & 1) ? cp : sp;
    const float c0 = (qi & 1) ? sp : cp;
    *s = (qi & 2) ? -s0 : s0;
    *c = ((qi + 1) & 2) ? -c0 : c0;
}

// Reference coefficient table (t4010a1.rs:18-69) with the segment start
// (i - 1) * ANGLE_DIV in f32; entry 10 stands for theta beyond 90 degrees.
#define AFH_DIR_ENTRIES(E)                                                                                   \
    E(0, 1.0f, 0.f, 0.f, 0.f)                                                                                \
    E(1, 1.0f, 0.f, 0.f, 0.f)                                                                                \
    E(2, 1.0f, 0.f, 0.f, 0.f)                                                                                \
    E(3, 1.0f, -0.2633589359691162f, -2.586746661956269f, 0.0009174516976956062f)                           \
    E(4, 0.891250938f, -0.8910683499820542f, -1.0097633564074366f, 0.0001704413361662757f)                  \
    E(5, 0.707945784f, -1.1924098538381465f, -0.7167962799004945f, 4.362034715924911f)                      \
    E(6, 0.501187234f, -1.0439934287916925f, 0.027352097655185065f, -2.238476800444034f)                    \
    E(7, 0.354813389f, -0.7015151838490918f, 0.39509604454751995f, 1.274774620435337f)                      \
    E(8, 0.251188643f, -0.4471050803642494f, 1.0625664749725618f, -1.1276955684895724f)                     \
    E(9, 0.199526231f, -0.179254124486545f, 0.008242027161100222f, -0.9016569137759131f)                    \
    E(10, 0.0f, 0.f, 0.f, 0.f)
#define AFH_DIR_ENTRY(i, A, B, C, D) {A, B, C, D, (float(i) - 1.0f) * kAngleDiv},

// Exact reference binning of theta (t4010a1.rs:108): i = ceil(theta / ANGLE_DIV)
// with Rust's saturating cast, NaN -> 0. Division is IEEE (correctly rounded).
__host__ __device__ inline float directivity_exact_t(float c, const DirEntry* tab) {
    const float theta = acos_band(c);
    const float q = ceilf(theta / kAngleDiv);
    int i = (q > 0.0f) ? (int)fminf(q, 10.0f) : 0;
    const DirEntry e = tab[i];
    const float x = theta - e.start;
    return e.a + (e.b + (e.c + e.d * x) * x) * x;
}

__host__ __device__ inline float directivity_band(float c, const DirEntry* tab) {
    float d = 1.0f;  // also the NaN case (c undefined at r = 0)
    if (c < kCosBandHi) {
        d = 0.0f;
        if (c > kCosBandLo) d = directivity_exact_t(c, tab);
    }
    return d;
}

// D from u = 512 c (see DirCell). fract(u) is v_fract_f32 on the device
// (u - floor(u), kept below 1) and the same formula on the host.
__host__ __device__ inline float directivity_cells(float u, const DirCell* cells) {
#if defined(__HIP_DEVICE_COMPILE__)
    const float t = __builtin_amdgcn_fractf(u);
#else
    const float t = fminf(u - floorf(u), 0x1.fffffep-1f);
#endif
    int j = (int)floorf(u) + 2;
    j = j < 0 ? 0 : (j > kDirCells - 1 ? kDirCells - 1 : j);
    const DirCell& e = cells[j];
    const float p = fmaf(fmaf(e.p2, t, e.p1), t, e.p0);
    const float q = fmaf(fmaf(e.q2, t, e.q1), t, e.q0);
    const float step = fminf(fmaxf(fmaf(t, e.s, e.k), 0.0f), 1.0f);
    return fmaf(step, q, p);
}

__host__ __device__ inline float directivity_table(float c, const DirCell* cells) {
    return directivity_cells(c * kDirScale, cells);  // exact scaling
}

#if defined(__HIPCC__)

// One source's term at one point, (re, im) = A (cos phi, sin phi). The FAST
// mode takes the directivity from the cell table, the ACCURATE mode evaluates
// the reference's acos / ceil binning / cubic.
template <int SRC, int PREC>
__device__ __forceinline__ void term_value(const Src& s, float px, float py, float pz, const DirEntry* tab,
                                           const DirCell* cells, float& a, float& cs, float& sn) {
    const float dx = px - s.x;
    const float dy = py - s.y;
    const float dz = pz - s.z;
    const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    const float rinv = __builtin_amdgcn_rsqf(d2);
    const float r = d2 * rinv;
    const float r_lo = fmaf(-r, r, d2) * (0.5f * rinv);
    float amp = s.amp;
    if (SRC == kT4010A1) {
        const float c = fmaf(s.nx, dx, fmaf(s.ny, dy, s.nz * dz)) * rinv;
        amp *= (PREC == kFast) ? directivity_table(c, cells) : directivity_band(c, tab);
    }
    const float att = __builtin_amdgcn_exp2f(-s.al * r);
    a = amp * att * rinv;
    const float t_hi = r * s.kh;
    const float t_lo = fmaf(r, s.kh, -t_hi) + fmaf(r_lo, s.kh, r * s.kl);
    const float u = __builtin_amdgcn_fractf(__builtin_amdgcn_fractf(t_hi) + (t_lo + s.ph));
    if (PREC == kFast) {
        sn = __builtin_amdgcn_sinf(u);
        cs = __builtin_amdgcn_cosf(u);
    } else {
        sincos_rev_poly(u, &sn, &cs);
    }
}

// FAST-mode source operands as the field kernels stage them in LDS (derived
// once per source from Src):
//   la  = log2|amp| (-inf for amp = 0), the amplitude enters the exponent;
//   n   = 512 * direction, so n.diff / r is the cell coordinate u = 512 c;
//   ph  = phase / 2 pi in [0, 1), plus 1/2 (mod 1) when amp < 0;
//   nal = -alpha log2 e, kh2 = kh / 2.
struct SrcF {
    float x, y, z, la;
    float nx, ny, nz, ph;
    float kh, kl, nal, kh2;
};

__device__ __forceinline__ SrcF stage_fast(const Src& s) {
    SrcF f;
    f.x = s.x;
    f.y = s.y;
    f.z = s.z;
    f.la = log2f(fabsf(s.amp));
    f.nx = s.nx * kDirScale;
    f.ny = s.ny * kDirScale;
    f.nz = s.nz * kDirScale;
    f.ph = s.amp < 0.0f ? __builtin_amdgcn_fractf(s.ph + 0.5f) : s.ph;
    f.kh = s.kh;
    f.kl = s.kl;
    f.nal = -s.al;
    f.kh2 = 0.5f * s.kh;
    return f;
}

// floor(u) as an integer in one instruction (v_cvt_flr_i32_f32; NaN gives 0).
__device__ __forceinline__ int floor_to_int(float u) {
    int j;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(j) : "v"(u));
    return j;
}

// The terms of one staged source at NP points (coordinates px[k], py[k],
// pz[k]), added to (re[k], im[k]): (re, im) += a (cos phi, sin phi) with
//   a   = D(u) / r * 2^(la - alpha log2e r),  u = n . (p - s) / r (scaled);
//   phi = 2 pi (fract(r kh) + [err(r kh) + r kl + ph + r_lo kh]), with
//         r_lo kh = (d2 - r^2) (kh / 2) / r; sin / cos take revolutions
//         (argument in [0, 2)),
// every VALU operand a VGPR or an inline constant. d2 and n . (p - s) sum the
// x and y parts first and add the z part last, so that add_terms_zrun (which
// shares the x and y parts between points of one column) gives the same
// bits. Arranged in stages over the points: distances and cell lookups for
// every point first, so the LDS reads of the cell table are in flight while
// the exponentials, phases and sin / cos are computed. The table is read as two float4 arrays (cp = p0 p1 p2 k,
// cq = q0 q1 q2 s of each DirCell), 16 bytes per read.
// With a non-null weight, the amplitude of point k is also multiplied by
// weight[k] (0 or 1 masks padding rows without a select); with a non-null
// amp_out, amp_out[k] receives that amplitude (the term's modulus, >= 0).
template <int SRC, int NP>
__device__ __forceinline__ void add_terms_staged(const SrcF& s, const float* px, const float* py, const float* pz,
                                                 const float4* cp, const float4* cq, float* re, float* im,
                                                 const float* weight = nullptr, float* amp_out = nullptr) {
    float d2[NP], rinv[NP], t[NP];
    float4 P[NP], Q[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const float dx = px[k] - s.x;
        const float dy = py[k] - s.y;
        const float dz = pz[k] - s.z;
        d2[k] = fmaf(dz, dz, fmaf(dx, dx, dy * dy));
        rinv[k] = __builtin_amdgcn_rsqf(d2[k]);
        if (SRC == kT4010A1) {
            const float u = fmaf(s.nz, dz, fmaf(s.nx, dx, s.ny * dy)) * rinv[k];
            t[k] = __builtin_amdgcn_fractf(u);
            int j = floor_to_int(u) + 2;
            j = j < 0 ? 0 : (j > kDirCells - 1 ? kDirCells - 1 : j);
            P[k] = cp[j];
            Q[k] = cq[j];
        }
    }
    float w[NP], sn[NP], cs[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const float r = d2[k] * rinv[k];
        w[k] = rinv[k] * __builtin_amdgcn_exp2f(fmaf(s.nal, r, s.la));
        const float e = fmaf(-r, r, d2[k]);
        const float t_hi = r * s.kh;
        const float err = fmaf(r, s.kh, -t_hi);
        float lo = fmaf(r, s.kl, s.ph);
        lo = fmaf(e, rinv[k] * s.kh2, lo);
        const float u = __builtin_amdgcn_fractf(t_hi) + (lo + err);
        sn[k] = __builtin_amdgcn_sinf(u);
        cs[k] = __builtin_amdgcn_cosf(u);
    }
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        float a = w[k];
        if (SRC == kT4010A1) {
            const float p = fmaf(fmaf(P[k].z, t[k], P[k].y), t[k], P[k].x);
            const float q = fmaf(fmaf(Q[k].z, t[k], Q[k].y), t[k], Q[k].x);
            const float step = fminf(fmaxf(fmaf(t[k], Q[k].w, P[k].w), 0.0f), 1.0f);
            a *= fmaf(step, q, p);
        }
        if (weight) a *= weight[k];
        if (amp_out) amp_out[k] = a;
        re[k] = fmaf(a, cs[k], re[k]);
        im[k] = fmaf(a, sn[k], im[k]);
    }
}

// add_terms_staged for NP points that share x and y (a run along z, heights
// pz[k]): dx, dy, dx^2 + dy^2 and the x / y part of n . (p - s) once per
// source instead of once per point, the rest per point exactly as in
// add_terms_staged (same operations, same order, same bits).
template <int SRC, int NP>
__device__ __forceinline__ void add_terms_zrun(const SrcF& s, float px, float py, const float* pz, const float4* cp,
                                               const float4* cq, float* re, float* im) {
    const float dx = px - s.x;
    const float dy = py - s.y;
    const float sxy = fmaf(dx, dx, dy * dy);
    const float nxy = fmaf(s.nx, dx, s.ny * dy);
    float d2[NP], rinv[NP], t[NP];
    float4 P[NP], Q[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const float dz = pz[k] - s.z;
        d2[k] = fmaf(dz, dz, sxy);
        rinv[k] = __builtin_amdgcn_rsqf(d2[k]);
        if (SRC == kT4010A1) {
            const float u = fmaf(s.nz, dz, nxy) * rinv[k];
            t[k] = __builtin_amdgcn_fractf(u);
            int j = floor_to_int(u) + 2;
            j = j < 0 ? 0 : (j > kDirCells - 1 ? kDirCells - 1 : j);
            P[k] = cp[j];
            Q[k] = cq[j];
        }
    }
    float w[NP], sn[NP], cs[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const float r = d2[k] * rinv[k];
        w[k] = rinv[k] * __builtin_amdgcn_exp2f(fmaf(s.nal, r, s.la));
        const float e = fmaf(-r, r, d2[k]);
        const float t_hi = r * s.kh;
        const float err = fmaf(r, s.kh, -t_hi);
        float lo = fmaf(r, s.kl, s.ph);
        lo = fmaf(e, rinv[k] * s.kh2, lo);
        const float u = __builtin_amdgcn_fractf(t_hi) + (lo + err);
        sn[k] = __builtin_amdgcn_sinf(u);
        cs[k] = __builtin_amdgcn_cosf(u);
    }
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        float a = w[k];
        if (SRC == kT4010A1) {
            const float p = fmaf(fmaf(P[k].z, t[k], P[k].y), t[k], P[k].x);
            const float q = fmaf(fmaf(Q[k].z, t[k], Q[k].y), t[k], Q[k].x);
            const float step = fminf(fmaxf(fmaf(t[k], Q[k].w, P[k].w), 0.0f), 1.0f);
            a *= fmaf(step, q, p);
        }
        re[k] = fmaf(a, cs[k], re[k]);
        im[k] = fmaf(a, sn[k], im[k]);
    }
}

// A position split into integer millimetres and the remainder: x = xh + xl
// with xh = rint(x) and |xl| <= 1/2, both exact in FP32.
struct Split3 {
    float xh, yh, zh, xl, yl, zl;
};

__device__ __forceinline__ Split3 split3(float x, float y, float z) {
    Split3 p;
    p.xh = rintf(x);
    p.yh = rintf(y);
    p.zh = rintf(z);
    p.xl = x - p.xh;
    p.yl = y - p.yh;
    p.zl = z - p.zh;
    return p;
}

// add_terms_staged with the distance of the exact coordinate difference. The
// FP32 difference p - s of the plain term rounds to an ulp of |p - s| (3e-5 mm
// at 500 mm, 2e-5 rad of phase at 40 kHz), an error the reference's FP32 path
// shares; here (p - s)^2 is formed as the integer part dh . dh (exact for
// distances below 4 m) plus the small remainder 2 dh . dl + dl . dl, and the
// phase takes r from their sum with the residual e = (dh . dh - r^2) + rem, so
// that k r is accurate to ~1e-7 rad. The directivity and the amplitude use
// the rounded difference, as before. `ss` is the source's split position,
// pt[k] the points'.
template <int SRC, int NP>
__device__ __forceinline__ void add_terms_exact(const SrcF& s, const Split3& ss, const Split3* pt, const float4* cp,
                                                const float4* cq, float* re, float* im, const float* weight,
                                                float* amp_out) {
    float d2[NP], rem[NP], d2h[NP], rinv[NP], t[NP];
    float4 P[NP], Q[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const float dxh = pt[k].xh - ss.xh, dyh = pt[k].yh - ss.yh, dzh = pt[k].zh - ss.zh;
        const float dxl = pt[k].xl - ss.xl, dyl = pt[k].yl - ss.yl, dzl = pt[k].zl - ss.zl;
        d2h[k] = fmaf(dxh, dxh, fmaf(dyh, dyh, dzh * dzh));
        rem[k] = fmaf(dxl, fmaf(2.0f, dxh, dxl), fmaf(dyl, fmaf(2.0f, dyh, dyl), dzl * fmaf(2.0f, dzh, dzl)));
        d2[k] = d2h[k] + rem[k];
        rinv[k] = __builtin_amdgcn_rsqf(d2[k]);
        if (SRC == kT4010A1) {
            const float dx = dxh + dxl, dy = dyh + dyl, dz = dzh + dzl;
            const float u = fmaf(s.nx, dx, fmaf(s.ny, dy, s.nz * dz)) * rinv[k];
            t[k] = __builtin_amdgcn_fractf(u);
            int j = floor_to_int(u) + 2;
            j = j < 0 ? 0 : (j > kDirCells - 1 ? kDirCells - 1 : j);
            P[k] = cp[j];
            Q[k] = cq[j];
        }
    }
    float w[NP], sn[NP], cs[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const float r = d2[k] * rinv[k];
        w[k] = rinv[k] * __builtin_amdgcn_exp2f(fmaf(s.nal, r, s.la));
        const float e = fmaf(-r, r, d2h[k]) + rem[k];
        const float t_hi = r * s.kh;
        const float err = fmaf(r, s.kh, -t_hi);
        float lo = fmaf(r, s.kl, s.ph);
        lo = fmaf(e, rinv[k] * s.kh2, lo);
        const float u = __builtin_amdgcn_fractf(t_hi) + (lo + err);
        sn[k] = __builtin_amdgcn_sinf(u);
        cs[k] = __builtin_amdgcn_cosf(u);
    }
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        float a = w[k];
        if (SRC == kT4010A1) {
            const float p = fmaf(fmaf(P[k].z, t[k], P[k].y), t[k], P[k].x);
            const float q = fmaf(fmaf(Q[k].z, t[k], Q[k].y), t[k], Q[k].x);
            const float step = fminf(fmaxf(fmaf(t[k], Q[k].w, P[k].w), 0.0f), 1.0f);
            a *= fmaf(step, q, p);
        }
        if (weight) a *= weight[k];
        if (amp_out) amp_out[k] = a;
        re[k] = fmaf(a, cs[k], re[k]);
        im[k] = fmaf(a, sn[k], im[k]);
    }
}

// Adds one source's term at one point to (re, im).
template <int SRC, int PREC>
__device__ __forceinline__ void add_term(const Src& s, float px, float py, float pz, const DirEntry* tab,
                                         const DirCell* cells, float& re, float& im) {
    float a, cs, sn;
    term_value<SRC, PREC>(s, px, py, pz, tab, cells, a, cs, sn);
    re = fmaf(a, cs, re);
    im = fmaf(a, sn, im);
}

// Error-free addition of x into (s, c): Knuth's TwoSum, the rounding error of
// s + x accumulated into c.
__device__ __forceinline__ void two_sum_acc(float& s, float& c, float x) {
    const float t = s + x;
    const float bp = t - s;
    const float err = (s - (t - bp)) + (x - bp);
    s = t;
    c += err;
}

#endif  // __HIPCC__

}  // namespace afh
