/*
 * afc_oracle -- CPU restatement of shinolab/acoustic-field-calculator's hot
 * paths, used as the oracle.
 *
 * TEST INFRASTRUCTURE ONLY: tests/, the smoke check and bench.py's CPU
 * baseline leg load this library as the checker. The product library never
 * links, loads or calls it.
 *
 * Two instantiations of afc_oracle_body.h:
 *   *_f32  the reference's FP32 CPU path in its own operation order (built
 *          with -ffp-contract=off, no fast-math, glibc libm like Rust's std),
 *   *_f64  the same formulas in double precision on the same f32 inputs.
 *
 * Functions and the reference file:line each follows:
 *   orc_directivity_t4010a1  core/wave_sources/t4010a1.rs:107-120
 *   orc_sound_speed          core/sound_speed.rs:22-24
 *   orc_attenuation_coef     core/attenuation.rs:31-48
 *   orc_field                system/uniform_medium.rs:111-128 + calculator/cpu_calculator.rs:43-65
 *   orc_field_accurate       accurate/system_acc.rs:26-54
 *   orc_propagation_matrix   acoustic-field-optimizer/src/multiple_foci/macros.rs:17-53
 *   orc_gs / orc_gspat       multiple_foci/linear_synthesis/gs.rs:36-77, gs_pat.rs:38-81
 *   orc_naive                multiple_foci/linear_synthesis/naive.rs:32-51
 *   orc_greedy               multiple_foci/combination/greedy.rs:23-83
 *   orc_lm                   multiple_foci/nls/lm.rs:40-105, nls/macros.rs:17-58
 *   orc_gauss_newton         multiple_foci/nls/gauss_newton.rs:35-89
 *   orc_gradient_descent     multiple_foci/nls/gradient_descent.rs:34-61
 *   orc_apo                  multiple_foci/nls/apo.rs:36-158
 *   orc_long                 multiple_foci/matrix/long.rs:46-106
 *   orc_horn                 multiple_foci/matrix/horn.rs:58-129 (seeded row choice)
 *   orc_herm_eig             the Hermitian eigen-decomposition both use (Jacobi)
 *   orc_bessel_phases        acoustic-field-optimizer/src/bessel.rs:31-55
 *   orc_focal_point_phases   acoustic-field-optimizer/src/focal_point.rs:28-39
 *   orc_grid_points          observe_area/grid/builder.rs:144-209, grid_area.rs:268-438
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* t4010a1.rs:18-69, as f32 literals */
static const float ORC_DIR_A[10] = {0.f,         1.0f,        1.0f,        1.0f,        0.891250938f,
                                    0.707945784f, 0.501187234f, 0.354813389f, 0.251188643f, 0.199526231f};
static const float ORC_DIR_B[10] = {0.f,
                                    0.f,
                                    0.f,
                                    -0.2633589359691162f,
                                    -0.8910683499820542f,
                                    -1.1924098538381465f,
                                    -1.0439934287916925f,
                                    -0.7015151838490918f,
                                    -0.4471050803642494f,
                                    -0.179254124486545f};
static const float ORC_DIR_C[10] = {0.f,
                                    0.f,
                                    0.f,
                                    -2.586746661956269f,
                                    -1.0097633564074366f,
                                    -0.7167962799004945f,
                                    0.027352097655185065f,
                                    0.39509604454751995f,
                                    1.0625664749725618f,
                                    0.008242027161100222f};
static const float ORC_DIR_D[10] = {0.f,
                                    0.f,
                                    0.f,
                                    0.0009174516976956062f,
                                    0.0001704413361662757f,
                                    4.362034715924911f,
                                    -2.238476800444034f,
                                    1.274774620435337f,
                                    -1.1276955684895724f,
                                    -0.9016569137759131f};

/* Rust: const ANGLE_DIV: f32 = PI / 18.0 (f32 arithmetic) */
#define ORC_PI_f32 3.14159265358979323846f
#define ORC_PI_f64 3.14159265358979323846
#define ORC_ANGLE_DIV_f32 (ORC_PI_f32 / 18.0f)
#define ORC_ANGLE_DIV_f64 (ORC_PI_f64 / 18.0)

/* floats per source in the reference's repr(C) structs */
static inline size_t orc_stride(int source_type) { return source_type == 1 ? 9 : 6; }

/* ---- FP32 instantiation ---- */
#define R float
#define SFX(name) name##_f32
#define RACOS acosf
#define RASIN asinf
#define RCOS cosf
#define RSIN sinf
#define REXP expf
#define RSQRT sqrtf
#define RSQRT_F sqrtf
#define RCEIL ceilf
#define RHYPOT hypotf
#define RATAN2 atan2f
#define RFMOD fmodf
#define RPOW powf
#define RABS fabsf
#include "afc_oracle_body.h"
#undef R
#undef SFX
#undef RACOS
#undef RASIN
#undef RCOS
#undef RSIN
#undef REXP
#undef RSQRT
#undef RSQRT_F
#undef RCEIL
#undef RHYPOT
#undef RATAN2
#undef RFMOD
#undef RPOW
#undef RABS

/* ---- FP64 instantiation ---- */
#define R double
#define SFX(name) name##_f64
#define RACOS acos
#define RASIN asin
#define RCOS cos
#define RSIN sin
#define REXP exp
#define RSQRT sqrt
#define RSQRT_F sqrt
#define RCEIL ceil
#define RHYPOT hypot
#define RATAN2 atan2
#define RFMOD fmod
#define RPOW pow
#define RABS fabs
#include "afc_oracle_body.h"

/* ---------------------------------------------------------------------------
 * Grid observation points (FP32 only, as the reference builds them).
 *
 * ranges[axis] = (min, max) for axes 0..2 (x, y, z); an `*_at` axis has
 * min == max. order[0..ndim) lists the ranged axes in the order they were
 * given (the first given varies fastest, grid_area.rs:197-213).
 * Count per axis: floor((max - min) / res) + 1 in f32 with Rust's saturating
 * `as usize` (builder.rs:198-200). Coordinates: origin + (n as f32) * res.
 * Returns the number of points; writes them to out (3 floats each) when out
 * is non-null and the per-axis counts (x, y, z) to bounds.
 * ------------------------------------------------------------------------- */
static uint64_t orc_axis_count(float lo, float hi, float res) {
    float q = (hi - lo) / res;
    uint64_t n;
    if (!(q > 0.0f)) {
        n = 0;
    } else if (q >= 18446744073709551615.0f) {
        n = UINT64_MAX - 1;
    } else {
        n = (uint64_t)q;
    }
    return n + 1;
}

uint64_t orc_grid_points(int ndim, const int* order, const float* ranges, float res, float* out, uint32_t* bounds) {
    uint64_t b[3], total, n0, n1, n2, idx = 0;
    float origin[3];
    int a, ax0, ax1, ax2;
    for (a = 0; a < 3; ++a) {
        b[a] = orc_axis_count(ranges[2 * a], ranges[2 * a + 1], res);
        origin[a] = ranges[2 * a];
        if (bounds) bounds[a] = (uint32_t)b[a];
    }
    total = b[0] * b[1] * b[2];
    if (!out) return total;
    if (ndim == 1) {
        ax0 = order[0];
        for (n0 = 0; n0 < total; ++n0) {
            float p[3] = {origin[0], origin[1], origin[2]};
            p[ax0] = origin[ax0] + ((float)n0 * res);
            memcpy(out + 3 * n0, p, sizeof(p));
        }
    } else if (ndim == 2) {
        ax0 = order[0];
        ax1 = order[1];
        for (n1 = 0; n1 < b[ax1]; ++n1)
            for (n0 = 0; n0 < b[ax0]; ++n0) {
                float p[3] = {origin[0], origin[1], origin[2]};
                p[ax1] = origin[ax1] + ((float)n1 * res);
                p[ax0] = origin[ax0] + ((float)n0 * res);
                memcpy(out + 3 * idx, p, sizeof(p));
                ++idx;
            }
    } else {
        ax0 = order[0];
        ax1 = order[1];
        ax2 = order[2];
        for (n2 = 0; n2 < b[ax2]; ++n2)
            for (n1 = 0; n1 < b[ax1]; ++n1)
                for (n0 = 0; n0 < b[ax0]; ++n0) {
                    float p[3];
                    p[ax2] = origin[ax2] + ((float)n2 * res);
                    p[ax1] = origin[ax1] + ((float)n1 * res);
                    p[ax0] = origin[ax0] + ((float)n0 * res);
                    memcpy(out + 3 * idx, p, sizeof(p));
                    ++idx;
                }
    }
    return total;
}

/* |p| and |p|^2 reductions of the FP32 complex pressure:
 * field_buffer/pressure_field.rs:35-37 (norm_sqr().sqrt()),
 * field_buffer/power_field.rs:35-37 (norm_sqr()). */
void orc_reduce_f32(int field_type, const float* cplx, uint64_t n, float* out) {
    uint64_t i;
    for (i = 0; i < n; ++i) {
        float re = cplx[2 * i], im = cplx[2 * i + 1];
        float ns = re * re + im * im;
        if (field_type == 0)
            out[i] = sqrtf(ns);
        else if (field_type == 1)
            out[i] = ns;
        else {
            out[2 * i] = re;
            out[2 * i + 1] = im;
        }
    }
}
