/*
 * Oracle body, instantiated twice by afc_oracle.c: once with R = float (the
 * reference's FP32 CPU path, same operation order, no FMA contraction) and
 * once with R = double (the FP64 twin used to measure error).
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product library links or calls
 * this code; it is the checker for tests/, the smoke check and the CPU
 * baseline leg of bench.py.
 *
 * Macros expected from the includer:
 *   R         real type
 *   SFX(name) name mangling (name##_f32 / name##_f64)
 *   RACOS RCOS RSIN REXP RSQRT RCEIL RHYPOT RATAN2 RFMOD RPOW RABS  libm calls
 */

typedef struct {
    R re, im;
} SFX(cplx);

static inline SFX(cplx) SFX(c_make)(R re, R im) {
    SFX(cplx) z;
    z.re = re;
    z.im = im;
    return z;
}
/* num-complex Mul: (a.re*b.re - a.im*b.im, a.re*b.im + a.im*b.re) */
static inline SFX(cplx) SFX(c_mul)(SFX(cplx) a, SFX(cplx) b) {
    return SFX(c_make)(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline SFX(cplx) SFX(c_add)(SFX(cplx) a, SFX(cplx) b) {
    return SFX(c_make)(a.re + b.re, a.im + b.im);
}
static inline SFX(cplx) SFX(c_conj)(SFX(cplx) a) { return SFX(c_make)(a.re, -a.im); }
static inline SFX(cplx) SFX(c_divr)(SFX(cplx) a, R r) { return SFX(c_make)(a.re / r, a.im / r); }
/* nalgebra ComplexField::abs for Complex = num-complex norm() = hypot */
static inline R SFX(c_abs)(SFX(cplx) a) { return RHYPOT(a.re, a.im); }
static inline R SFX(c_arg)(SFX(cplx) a) { return RATAN2(a.im, a.re); }

/* ---------------------------------------------------------------------------
 * T4010A1 directivity: core/wave_sources/t4010a1.rs:107-120, tables :18-69.
 * The tables are the reference's f32 literals; the f64 twin promotes them.
 * ------------------------------------------------------------------------- */
R SFX(orc_directivity_t4010a1)(R theta) {
    const R angle_div = SFX(ORC_ANGLE_DIV);
    R q = RCEIL(theta / angle_div);
    /* Rust `as usize` saturates: NaN and negatives become 0. */
    size_t i;
    if (!(q > 0)) {
        i = 0;
    } else if (q > (R)1e9) {
        i = 1000000000u;
    } else {
        i = (size_t)q;
    }
    if (i == 0) return (R)1.0;
    if (i > 9) return (R)0.0;
    {
        R a = (R)ORC_DIR_A[i];
        R b = (R)ORC_DIR_B[i];
        R c = (R)ORC_DIR_C[i];
        R d = (R)ORC_DIR_D[i];
        R x = theta - ((R)i - (R)1.0) * angle_div;
        return a + (b + ((c + d * x) * x)) * x;
    }
}

/* core/sound_speed.rs:22-24, in mm/s */
R SFX(orc_sound_speed)(R t) { return (R)331.3 * RSQRT(t / (R)273.15) * (R)1000.0; }

/* core/attenuation.rs:31-48 (Bass et al. 1995), Np/mm */
R SFX(orc_attenuation_coef)(R freq, R hr, R ps, R ps0, R t) {
    const R T0 = (R)293.15;
    const R T01 = (R)273.16;
    R ten = (R)10.0;
    R psat = ps0 * RPOW(ten, (R)-6.8346 * RPOW(T01 / t, (R)1.261) + (R)4.6151);
    R h = ps0 * (hr / ps) * (psat / ps0);
    R f_ro = ((R)24. + (R)4.04e4 * h * ((R)0.02 + h) / ((R)0.391 + h)) / ps0;
    R f_rn = ((R)1. / ps0) * ((R)9. + (R)280. * h * REXP((R)-4.17 * (RPOW(T0 / t, (R)1. / (R)3.) - (R)1.))) *
             RPOW(T0 / t, (R)1. / (R)2.);
    R f = freq / ps;
    R alpha = (f * f) / ps0 * ps *
              ((R)1.84 * RPOW(t / T0, (R)1. / (R)2.) * (R)1e-11 +
               RPOW(t / T0, (R)-5. / (R)2.) * ((R)0.01278 * REXP((R)-2239.1 / t) / (f_ro + f * f / f_ro) +
                                                (R)0.1068 * REXP((R)-3352. / t) / (f_rn + f * f / f_rn)));
    return alpha * (R)1e-3;
}

/* ---------------------------------------------------------------------------
 * One source-to-point term: system/uniform_medium.rs:111-120
 * (same as acoustic-field-optimizer/src/multiple_foci/macros.rs:17-30).
 * src points at the reference struct: SphereWaveSource = {pos[3], amp, phase,
 * freq} (sphere_sources.rs:19-27); T4010A1 = {pos[3], dir[3], amp, phase, freq}
 * (t4010a1.rs:73-82).
 * ------------------------------------------------------------------------- */
static inline SFX(cplx) SFX(term)(int source_type, const float* src, R wavenum, R atten, R tx, R ty, R tz,
                                  R phase_override, int use_override) {
    R px = (R)src[0], py = (R)src[1], pz = (R)src[2];
    R dirx = 0, diry = 0, dirz = 0, amp, phase;
    R dx, dy, dz, dist, cos_t, theta, d, r, phi;
    if (source_type == 1) {
        dirx = (R)src[3];
        diry = (R)src[4];
        dirz = (R)src[5];
        amp = (R)src[6];
        phase = (R)src[7];
    } else {
        amp = (R)src[3];
        phase = (R)src[4];
    }
    if (use_override) phase = phase_override;
    /* fmath::sub is plain subtraction in the capi build (fmath gated off by `accurate`) */
    dx = tx - px;
    dy = ty - py;
    dz = tz - pz;
    /* nalgebra norm of a 3-vector: sqrt((x*x + y*y) + z*z) */
    dist = RSQRT_F((dx * dx + dy * dy) + dz * dz);
    cos_t = ((dirx * dx + diry * dy) + dirz * dz) / dist;
    theta = RACOS(cos_t);
    d = (source_type == 1) ? SFX(orc_directivity_t4010a1)(theta) : (R)1.0;
    r = amp * d * REXP(-dist * atten) / dist;
    phi = phase + wavenum * dist;
    return SFX(c_make)(r * RCOS(phi), r * RSIN(phi));
}

/* ---------------------------------------------------------------------------
 * Field: CpuCalculator::calculate (calculator/cpu_calculator.rs:43-65) over
 * UniformSystem::propagate_all (system/uniform_medium.rs:122-128): per point a
 * serial complex sum in source order. out holds interleaved (re, im).
 * ------------------------------------------------------------------------- */
void SFX(orc_field)(int source_type, const float* sources, const float* wavenums, const float* attens,
                    uint64_t n_sources, const float* points, uint64_t n_points, R* out, int threads) {
    const size_t st = orc_stride(source_type);
    int64_t p;
    if (threads <= 0) threads = 1;
#pragma omp parallel for schedule(static) num_threads(threads)
    for (p = 0; p < (int64_t)n_points; ++p) {
        SFX(cplx) c = SFX(c_make)(0, 0);
        R tx = (R)points[3 * p], ty = (R)points[3 * p + 1], tz = (R)points[3 * p + 2];
        uint64_t i;
        for (i = 0; i < n_sources; ++i) {
            SFX(cplx) t = SFX(term)(source_type, sources + st * i, (R)wavenums[i], (R)attens[i], tx, ty, tz, 0, 0);
            c.re += t.re;
            c.im += t.im;
        }
        out[2 * p] = c.re;
        out[2 * p + 1] = c.im;
    }
}

static int SFX(cmp_asc)(const void* a, const void* b) {
    R x = *(const R*)a, y = *(const R*)b;
    return (x > y) - (x < y);
}

/* AccurateCalculator: accurate/system_acc.rs:26-54 -- each term's re and im
 * are summed in ascending signed order (min-heap drain), starting from 0. */
void SFX(orc_field_accurate)(int source_type, const float* sources, const float* wavenums, const float* attens,
                             uint64_t n_sources, const float* points, uint64_t n_points, R* out, int threads) {
    const size_t st = orc_stride(source_type);
    int64_t p;
    if (threads <= 0) threads = 1;
#pragma omp parallel num_threads(threads)
    {
        R* re = (R*)malloc(sizeof(R) * (n_sources + 1));
        R* im = (R*)malloc(sizeof(R) * (n_sources + 1));
#pragma omp for schedule(static)
        for (p = 0; p < (int64_t)n_points; ++p) {
            R tx = (R)points[3 * p], ty = (R)points[3 * p + 1], tz = (R)points[3 * p + 2];
            R sr = 0, si = 0;
            uint64_t i;
            for (i = 0; i < n_sources; ++i) {
                SFX(cplx) t = SFX(term)(source_type, sources + st * i, (R)wavenums[i], (R)attens[i], tx, ty, tz, 0, 0);
                re[i] = t.re;
                im[i] = t.im;
            }
            qsort(re, n_sources, sizeof(R), SFX(cmp_asc));
            qsort(im, n_sources, sizeof(R), SFX(cmp_asc));
            for (i = 0; i < n_sources; ++i) sr += re[i];
            for (i = 0; i < n_sources; ++i) si += im[i];
            out[2 * p] = sr;
            out[2 * p + 1] = si;
        }
        free(re);
        free(im);
    }
}

/* ---------------------------------------------------------------------------
 * Propagation matrix G (M x N, column-major, element (i, j) at i + j*M):
 * acoustic-field-optimizer/src/multiple_foci/macros.rs:32-53. The caller's
 * phases are used as stored (the optimizers zero them first).
 * ------------------------------------------------------------------------- */
void SFX(orc_propagation_matrix)(int source_type, const float* sources, const float* wavenums, const float* attens,
                                 uint64_t n, const float* foci, uint64_t m, R* out_interleaved) {
    const size_t st = orc_stride(source_type);
    uint64_t i, j;
    for (j = 0; j < n; ++j) {
        for (i = 0; i < m; ++i) {
            SFX(cplx) t = SFX(term)(source_type, sources + st * j, (R)wavenums[j], (R)attens[j], (R)foci[3 * i],
                                    (R)foci[3 * i + 1], (R)foci[3 * i + 2], 0, 0);
            out_interleaved[2 * (i + j * m)] = t.re;
            out_interleaved[2 * (i + j * m) + 1] = t.im;
        }
    }
}

/* G with every phase forced to 0, as GS / GS-PAT / LM do before building it
 * (gs.rs:37-39, gs_pat.rs:39-41, lm.rs:41-43). */
static SFX(cplx) * SFX(make_G)(int source_type, const float* sources, const float* wavenums, const float* attens,
                               uint64_t n, const float* foci, uint64_t m) {
    const size_t st = orc_stride(source_type);
    SFX(cplx)* G = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * m * n);
    uint64_t i, j;
    for (j = 0; j < n; ++j)
        for (i = 0; i < m; ++i)
            G[i + j * m] = SFX(term)(source_type, sources + st * j, (R)wavenums[j], (R)attens[j], (R)foci[3 * i],
                                     (R)foci[3 * i + 1], (R)foci[3 * i + 2], 0, 1);
    return G;
}

static R SFX(src_amp)(int source_type, const float* sources, uint64_t j) {
    return (R)sources[orc_stride(source_type) * j + (source_type == 1 ? 6 : 3)];
}

/* y = A x for column-major A (rows x cols): nalgebra gemv accumulates column by
 * column, y += x_j * A[:, j]. */
static void SFX(gemv)(const SFX(cplx) * A, uint64_t rows, uint64_t cols, const SFX(cplx) * x, SFX(cplx) * y) {
    uint64_t i, j;
    for (i = 0; i < rows; ++i) y[i] = SFX(c_make)(0, 0);
    for (j = 0; j < cols; ++j)
        for (i = 0; i < rows; ++i) y[i] = SFX(c_add)(SFX(c_mul)(x[j], A[i + j * rows]), y[i]);
}

/* ---------------------------------------------------------------------------
 * GS: acoustic-field-optimizer/src/multiple_foci/linear_synthesis/gs.rs:36-77.
 * Outputs per-source amp = |q| and phase = arg q.
 * ------------------------------------------------------------------------- */
void SFX(orc_gs)(int source_type, const float* sources, const float* wavenums, const float* attens, uint64_t n,
                 const float* foci, const float* amps, uint64_t m, uint32_t repeat, R* out_amp, R* out_phase) {
    SFX(cplx)* G = SFX(make_G)(source_type, sources, wavenums, attens, n, foci, m);
    SFX(cplx)* Gh = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * m * n); /* N x M column-major */
    SFX(cplx)* q0 = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * n);
    SFX(cplx)* q = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * n);
    SFX(cplx)* xi = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * n);
    SFX(cplx)* p0 = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * m);
    SFX(cplx)* p = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * m);
    SFX(cplx)* gamma = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * m);
    uint64_t i, j;
    uint32_t it;
    for (j = 0; j < n; ++j)
        for (i = 0; i < m; ++i) Gh[j + i * n] = SFX(c_conj)(G[i + j * m]);
    for (i = 0; i < m; ++i) p0[i] = SFX(c_make)((R)amps[i], 0);
    for (j = 0; j < n; ++j) q0[j] = SFX(c_make)(SFX(src_amp)(source_type, sources, j), 0);
    for (j = 0; j < n; ++j) q[j] = q0[j];
    for (it = 0; it < repeat; ++it) {
        SFX(gemv)(G, m, n, q, gamma);
        for (i = 0; i < m; ++i) p[i] = SFX(c_mul)(SFX(c_divr)(gamma[i], SFX(c_abs)(gamma[i])), p0[i]);
        SFX(gemv)(Gh, n, m, p, xi);
        for (j = 0; j < n; ++j) q[j] = SFX(c_mul)(SFX(c_divr)(xi[j], SFX(c_abs)(xi[j])), q0[j]);
    }
    for (j = 0; j < n; ++j) {
        out_amp[j] = SFX(c_abs)(q[j]);
        out_phase[j] = SFX(c_arg)(q[j]);
    }
    free(G);
    free(Gh);
    free(q0);
    free(q);
    free(xi);
    free(p0);
    free(p);
    free(gamma);
}

/* ---------------------------------------------------------------------------
 * GS-PAT: acoustic-field-optimizer/src/multiple_foci/linear_synthesis/gs_pat.rs:38-81.
 * denomi_j = sum_i |G_ij| (row_sum: one entry per transducer column),
 * B = (conj(G_ij) / denomi_j^2)^T (N x M), R = G B (M x M).
 * ------------------------------------------------------------------------- */
void SFX(orc_gspat)(int source_type, const float* sources, const float* wavenums, const float* attens, uint64_t n,
                    const float* foci, const float* amps, uint64_t m, uint32_t repeat, R* out_amp, R* out_phase) {
    SFX(cplx)* G = SFX(make_G)(source_type, sources, wavenums, attens, n, foci, m);
    SFX(cplx)* B = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * m * n); /* N x M column-major */
    SFX(cplx)* Rm = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * m * m);
    SFX(cplx)* p0 = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * m);
    SFX(cplx)* p = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * m);
    SFX(cplx)* gamma = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * m);
    SFX(cplx)* q = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * n);
    R* denomi = (R*)malloc(sizeof(R) * n);
    uint64_t i, j, k;
    uint32_t it;
    for (j = 0; j < n; ++j) {
        R s = 0;
        for (i = 0; i < m; ++i) s += SFX(c_abs)(G[i + j * m]);
        denomi[j] = s;
    }
    for (j = 0; j < n; ++j)
        for (i = 0; i < m; ++i) B[j + i * n] = SFX(c_divr)(SFX(c_conj)(G[i + j * m]), denomi[j] * denomi[j]);
    /* R = G * B, column by column (gemm as repeated gemv) */
    for (k = 0; k < m; ++k) SFX(gemv)(G, m, n, B + k * n, Rm + k * m);
    for (i = 0; i < m; ++i) p0[i] = SFX(c_make)((R)amps[i], 0);
    for (i = 0; i < m; ++i) p[i] = p0[i];
    SFX(gemv)(Rm, m, m, p, gamma);
    for (it = 0; it < repeat; ++it) {
        for (i = 0; i < m; ++i) p[i] = SFX(c_mul)(SFX(c_divr)(gamma[i], SFX(c_abs)(gamma[i])), p0[i]);
        SFX(gemv)(Rm, m, m, p, gamma);
    }
    for (i = 0; i < m; ++i) {
        R a = SFX(c_abs)(gamma[i]);
        p[i] = SFX(c_mul)(SFX(c_mul)(SFX(c_divr)(gamma[i], a * a), p0[i]), p0[i]);
    }
    SFX(gemv)(B, n, m, p, q);
    for (j = 0; j < n; ++j) {
        out_amp[j] = SFX(c_abs)(q[j]);
        out_phase[j] = SFX(c_arg)(q[j]);
    }
    free(G);
    free(B);
    free(Rm);
    free(p0);
    free(p);
    free(gamma);
    free(q);
    free(denomi);
}

/* Naive linear synthesis: linear_synthesis/naive.rs:32-51, q = G^H p. */
void SFX(orc_naive)(int source_type, const float* sources, const float* wavenums, const float* attens, uint64_t n,
                    const float* foci, const float* amps, uint64_t m, R* out_amp, R* out_phase) {
    SFX(cplx)* G = SFX(make_G)(source_type, sources, wavenums, attens, n, foci, m);
    uint64_t i, j;
    for (j = 0; j < n; ++j) {
        SFX(cplx) q = SFX(c_make)(0, 0);
        for (i = 0; i < m; ++i) q = SFX(c_add)(SFX(c_mul)(SFX(c_make)((R)amps[i], 0), SFX(c_conj)(G[i + j * m])), q);
        out_amp[j] = SFX(c_abs)(q);
        out_phase[j] = SFX(c_arg)(q);
    }
    free(G);
}

/* Greedy: combination/greedy.rs:23-83. Candidates e^{i 2 pi idx / div}
 * (Complex::exp of (0, 2 pi idx / div), evaluated left to right); for each
 * source in index order the first candidate minimising
 * sum_j |amps_j - |G_ji c + cache_j|| is kept (strict <, so ties keep the
 * earliest), cache += G_:i c, and the source's phase becomes arg(c). The
 * amplitudes are not changed. */
void SFX(orc_greedy)(int source_type, const float* sources, const float* wavenums, const float* attens, uint64_t n,
                     const float* foci, const float* amps, uint64_t m, uint64_t phase_div, R* out_phase) {
    SFX(cplx)* G = SFX(make_G)(source_type, sources, wavenums, attens, n, foci, m);
    SFX(cplx)* cand = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * phase_div);
    SFX(cplx)* cache = (SFX(cplx)*)calloc(m, sizeof(SFX(cplx)));
    const R pi = SFX(ORC_PI);
    uint64_t i, j, idx;
    for (idx = 0; idx < phase_div; ++idx) {
        const R th = (R)2.0 * pi * (R)idx / (R)phase_div;
        cand[idx] = SFX(c_make)(REXP((R)0) * RCOS(th), REXP((R)0) * RSIN(th));
    }
    for (i = 0; i < n; ++i) {
        uint64_t min_idx = 0;
        R min_v = (R)INFINITY;
        for (idx = 0; idx < phase_div; ++idx) {
            R v = 0;
            for (j = 0; j < m; ++j) {
                const SFX(cplx) tmp = SFX(c_mul)(G[j + i * m], cand[idx]);
                v += RABS((R)amps[j] - SFX(c_abs)(SFX(c_add)(tmp, cache[j])));
            }
            if (v < min_v) {
                min_v = v;
                min_idx = idx;
            }
        }
        for (j = 0; j < m; ++j) cache[j] = SFX(c_add)(cache[j], SFX(c_mul)(G[j + i * m], cand[min_idx]));
        out_phase[i] = SFX(c_arg)(cand[min_idx]);
    }
    free(G);
    free(cand);
    free(cache);
}

/* ---------------------------------------------------------------------------
 * Levenberg-Marquardt: acoustic-field-optimizer/src/multiple_foci/nls/lm.rs:40-105
 * with make_BhB / make_T / calc_JtJ_Jtf / calc_Fx from nls/macros.rs:17-58.
 * The damped system is solved by Householder QR (nalgebra qr().solve), which
 * reports failure when a diagonal entry of R is zero.
 * ------------------------------------------------------------------------- */
static int SFX(qr_solve)(R* A, uint64_t n, R* b) {
    /* A column-major n x n, overwritten; b overwritten by the solution */
    uint64_t k, i, j;
    R* diag = (R*)malloc(sizeof(R) * n);
    for (k = 0; k < n; ++k) {
        R norm = 0, alpha, s;
        for (i = k; i < n; ++i) norm = RHYPOT(norm, A[i + k * n]);
        if (norm == 0) {
            diag[k] = 0;
            continue;
        }
        alpha = (A[k + k * n] > 0) ? -norm : norm;
        /* v = x - alpha e1 stored in A[k:, k] */
        A[k + k * n] -= alpha;
        s = 0;
        for (i = k; i < n; ++i) s += A[i + k * n] * A[i + k * n];
        if (s != 0) {
            for (j = k + 1; j < n; ++j) {
                R dot = 0;
                for (i = k; i < n; ++i) dot += A[i + k * n] * A[i + j * n];
                dot = (R)2 * dot / s;
                for (i = k; i < n; ++i) A[i + j * n] -= dot * A[i + k * n];
            }
            {
                R dot = 0;
                for (i = k; i < n; ++i) dot += A[i + k * n] * b[i];
                dot = (R)2 * dot / s;
                for (i = k; i < n; ++i) b[i] -= dot * A[i + k * n];
            }
        }
        diag[k] = alpha;
    }
    for (k = 0; k < n; ++k)
        if (diag[k] == 0) {
            free(diag);
            return 0;
        }
    /* back substitution with R (diag on diag[], strict upper in A) */
    for (i = n; i-- > 0;) {
        R s = b[i];
        for (j = i + 1; j < n; ++j) s -= A[i + j * n] * b[j];
        b[i] = s / diag[i];
    }
    free(diag);
    return 1;
}

static void SFX(lm_A_g)(const SFX(cplx) * BhB, const SFX(cplx) * T, uint64_t np, R* A, R* g) {
    uint64_t k, l;
    for (k = 0; k < np; ++k) g[k] = 0;
    /* column_sum: g = sum over columns l of Im(BhB o T T^H)[:, l], column by column */
    for (l = 0; l < np; ++l) {
        SFX(cplx) tl = SFX(c_conj)(T[l]);
        for (k = 0; k < np; ++k) {
            SFX(cplx) tt = SFX(c_mul)(tl, T[k]);
            SFX(cplx) e = SFX(c_mul)(BhB[k + l * np], tt);
            A[k + l * np] = e.re;
            g[k] += e.im;
        }
    }
}

static R SFX(lm_Fx)(const SFX(cplx) * BhB, const R* x, uint64_t np, SFX(cplx) * t, SFX(cplx) * row) {
    uint64_t k, l;
    SFX(cplx) acc = SFX(c_make)(0, 0);
    for (k = 0; k < np; ++k) t[k] = SFX(c_make)(RCOS(x[k]), RSIN(x[k]));
    /* row = t^H BhB (1 x np), then row * t */
    for (l = 0; l < np; ++l) {
        SFX(cplx) s = SFX(c_make)(0, 0);
        for (k = 0; k < np; ++k) s = SFX(c_add)(SFX(c_mul)(SFX(c_conj)(t[k]), BhB[k + l * np]), s);
        row[l] = s;
    }
    for (l = 0; l < np; ++l) acc = SFX(c_add)(SFX(c_mul)(t[l], row[l]), acc);
    return acc.re;
}

/* Returns the number of iterations taken; out_phase gets x[0..n). */
uint32_t SFX(orc_lm)(int source_type, const float* sources, const float* wavenums, const float* attens, uint64_t n,
                     const float* foci, const float* amps, uint64_t m, uint32_t k_max, R* out_phase, R* out_x) {
    const R EPS_1 = (R)1e-8, EPS_2 = (R)1e-8, TAU = (R)1e-3;
    const uint64_t np = n + m;
    SFX(cplx)* G = SFX(make_G)(source_type, sources, wavenums, attens, n, foci, m);
    SFX(cplx)* Bm = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * m * np); /* M x np */
    SFX(cplx)* BhB = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * np * np);
    SFX(cplx)* T = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * np);
    SFX(cplx)* tv = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * np);
    SFX(cplx)* row = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * np);
    R* A = (R*)malloc(sizeof(R) * np * np);
    R* M = (R*)malloc(sizeof(R) * np * np);
    R* g = (R*)malloc(sizeof(R) * np);
    R* h = (R*)malloc(sizeof(R) * np);
    R* x = (R*)calloc(np, sizeof(R));
    R* xn = (R*)malloc(sizeof(R) * np);
    R nu = 2, mu, Fx, amax, gmax;
    uint64_t i, j, k;
    uint32_t it = 0;
    int found;
    memset(Bm, 0, sizeof(SFX(cplx)) * m * np);
    for (j = 0; j < n; ++j)
        for (i = 0; i < m; ++i) Bm[i + j * m] = G[i + j * m];
    for (i = 0; i < m; ++i) Bm[i + (n + i) * m] = SFX(c_make)(-(R)amps[i], 0);
    /* BhB = B^H B: column l = sum_i B^H[:, i] * B[i, l] */
    for (j = 0; j < np; ++j) {
        for (k = 0; k < np; ++k) BhB[k + j * np] = SFX(c_make)(0, 0);
        for (i = 0; i < m; ++i) {
            SFX(cplx) bil = Bm[i + j * m];
            for (k = 0; k < np; ++k)
                BhB[k + j * np] = SFX(c_add)(SFX(c_mul)(bil, SFX(c_conj)(Bm[i + k * m])), BhB[k + j * np]);
        }
    }
    for (k = 0; k < np; ++k) T[k] = SFX(c_make)(RCOS(x[k]), -RSIN(x[k]));
    SFX(lm_A_g)(BhB, T, np, A, g);
    amax = A[0];
    for (k = 1; k < np; ++k)
        if (A[k + k * np] > amax) amax = A[k + k * np];
    gmax = g[0];
    for (k = 1; k < np; ++k)
        if (g[k] > gmax) gmax = g[k];
    mu = TAU * amax;
    found = gmax <= EPS_1;
    Fx = SFX(lm_Fx)(BhB, x, np, tv, row);
    for (it = 0; it < k_max; ++it) {
        R hn = 0, xnorm = 0;
        if (found) break;
        memcpy(M, A, sizeof(R) * np * np);
        for (k = 0; k < np; ++k) M[k + k * np] += mu;
        for (k = 0; k < np; ++k) h[k] = g[k];
        if (!SFX(qr_solve)(M, np, h)) break;
        for (k = 0; k < np; ++k) h[k] = -h[k];
        for (k = 0; k < np; ++k) hn += h[k] * h[k];
        for (k = 0; k < np; ++k) xnorm += x[k] * x[k];
        hn = RSQRT(hn);
        xnorm = RSQRT(xnorm);
        if (hn <= EPS_2 * (xnorm + EPS_2)) {
            found = 1;
        } else {
            R Fx_new, L, rho;
            for (k = 0; k < np; ++k) xn[k] = x[k] + h[k];
            Fx_new = SFX(lm_Fx)(BhB, xn, np, tv, row);
            L = 0;
            for (k = 0; k < np; ++k) L += h[k] * (mu * h[k] - g[k]);
            L = (R)0.5 * L;
            rho = (Fx - Fx_new) / L;
            Fx = Fx_new; /* assigned even when the step is rejected (lm.rs:85) */
            if (rho > 0) {
                R f;
                memcpy(x, xn, sizeof(R) * np);
                for (k = 0; k < np; ++k) T[k] = SFX(c_make)(RCOS(x[k]), -RSIN(x[k]));
                SFX(lm_A_g)(BhB, T, np, A, g);
                gmax = g[0];
                for (k = 1; k < np; ++k)
                    if (g[k] > gmax) gmax = g[k];
                found = gmax <= EPS_1;
                f = (R)1 - RPOW((R)2 * rho - (R)1, (R)3);
                mu *= ((R)1 / (R)3 > f) ? (R)1 / (R)3 : f;
                nu = 2;
            } else {
                mu *= nu;
                nu *= 2;
            }
        }
    }
    for (j = 0; j < n; ++j) out_phase[j] = x[j];
    if (out_x)
        for (k = 0; k < np; ++k) out_x[k] = x[k];
    free(G);
    free(Bm);
    free(BhB);
    free(T);
    free(tv);
    free(row);
    free(A);
    free(M);
    free(g);
    free(h);
    free(x);
    free(xn);
    return it;
}

/* BhB = B^H B with B = [G | -diag(amps)] (nls/macros.rs:17-31), column-major
 * np x np, caller frees. */
static SFX(cplx) * SFX(nls_BhB)(int source_type, const float* sources, const float* wavenums, const float* attens,
                                 uint64_t n, const float* foci, const float* amps, uint64_t m) {
    const uint64_t np = n + m;
    SFX(cplx)* G = SFX(make_G)(source_type, sources, wavenums, attens, n, foci, m);
    SFX(cplx)* Bm = (SFX(cplx)*)calloc(m * np, sizeof(SFX(cplx)));
    SFX(cplx)* BhB = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * np * np);
    uint64_t i, j, k;
    for (j = 0; j < n; ++j)
        for (i = 0; i < m; ++i) Bm[i + j * m] = G[i + j * m];
    for (i = 0; i < m; ++i) Bm[i + (n + i) * m] = SFX(c_make)(-(R)amps[i], 0);
    for (j = 0; j < np; ++j) {
        for (k = 0; k < np; ++k) BhB[k + j * np] = SFX(c_make)(0, 0);
        for (i = 0; i < m; ++i) {
            SFX(cplx) bil = Bm[i + j * m];
            for (k = 0; k < np; ++k)
                BhB[k + j * np] = SFX(c_add)(SFX(c_mul)(bil, SFX(c_conj)(Bm[i + k * m])), BhB[k + j * np]);
        }
    }
    free(G);
    free(Bm);
    return BhB;
}

/* Eigen-decomposition of a symmetric matrix (column-major n x n, destroyed)
 * by cyclic Jacobi rotations: evals[k], eigenvectors in the columns of V. */
static void SFX(sym_eig)(R* A, uint64_t n, R* evals, R* V) {
    uint64_t p, q, k, sweep;
    for (p = 0; p < n; ++p)
        for (q = 0; q < n; ++q) V[p + q * n] = (p == q) ? (R)1 : (R)0;
    for (sweep = 0; sweep < 100; ++sweep) {
        R off = 0, diag = 0;
        for (q = 0; q < n; ++q)
            for (p = 0; p < n; ++p) {
                if (p != q) off += A[p + q * n] * A[p + q * n];
                else diag += A[p + p * n] * A[p + p * n];
            }
        if (off <= diag * (R)1e-30 || off == 0) break;
        for (p = 0; p + 1 < n; ++p)
            for (q = p + 1; q < n; ++q) {
                const R apq = A[p + q * n];
                R theta, t, c, sn;
                if (apq == 0) continue;
                theta = (A[q + q * n] - A[p + p * n]) / ((R)2 * apq);
                t = (theta >= 0 ? (R)1 : (R)-1) / (RABS(theta) + RSQRT(theta * theta + (R)1));
                c = (R)1 / RSQRT(t * t + (R)1);
                sn = t * c;
                for (k = 0; k < n; ++k) {
                    const R akp = A[k + p * n], akq = A[k + q * n];
                    A[k + p * n] = c * akp - sn * akq;
                    A[k + q * n] = sn * akp + c * akq;
                }
                for (k = 0; k < n; ++k) {
                    const R apk = A[p + k * n], aqk = A[q + k * n];
                    A[p + k * n] = c * apk - sn * aqk;
                    A[q + k * n] = sn * apk + c * aqk;
                }
                for (k = 0; k < n; ++k) {
                    const R vkp = V[k + p * n], vkq = V[k + q * n];
                    V[k + p * n] = c * vkp - sn * vkq;
                    V[k + q * n] = sn * vkp + c * vkq;
                }
            }
    }
    for (k = 0; k < n; ++k) evals[k] = A[k + k * n];
}

/* Gauss-Newton: nls/gauss_newton.rs:35-89. h = -pinv(A, 1e-3) g, the
 * pseudo-inverse keeping singular values above 1e-3 (A is symmetric positive
 * semi-definite, so its singular values are |eigenvalues|). Returns the
 * iterations taken. */
uint32_t SFX(orc_gauss_newton)(int source_type, const float* sources, const float* wavenums, const float* attens,
                               uint64_t n, const float* foci, const float* amps, uint64_t m, uint32_t k_max,
                               R* out_phase, R* out_x) {
    const R EPS_1 = (R)1e-8, EPS_2 = (R)1e-8, PINV_EPS = (R)1e-3;
    const uint64_t np = n + m;
    SFX(cplx)* BhB = SFX(nls_BhB)(source_type, sources, wavenums, attens, n, foci, amps, m);
    SFX(cplx)* T = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * np);
    R* A = (R*)malloc(sizeof(R) * np * np);
    R* V = (R*)malloc(sizeof(R) * np * np);
    R* ev = (R*)malloc(sizeof(R) * np);
    R* w = (R*)malloc(sizeof(R) * np);
    R* g = (R*)malloc(sizeof(R) * np);
    R* h = (R*)malloc(sizeof(R) * np);
    R* x = (R*)calloc(np, sizeof(R));
    uint64_t k, l;
    uint32_t it;
    int found;
    R gmax;
    for (k = 0; k < np; ++k) T[k] = SFX(c_make)(RCOS(x[k]), -RSIN(x[k]));
    SFX(lm_A_g)(BhB, T, np, A, g);
    gmax = g[0];
    for (k = 1; k < np; ++k)
        if (g[k] > gmax) gmax = g[k];
    found = gmax <= EPS_1;
    for (it = 0; it < k_max; ++it) {
        R hn = 0, xn = 0;
        if (found) break;
        SFX(sym_eig)(A, np, ev, V);
        for (k = 0; k < np; ++k) {
            R d = 0;
            for (l = 0; l < np; ++l) d += V[l + k * np] * g[l];
            w[k] = RABS(ev[k]) > PINV_EPS ? d / ev[k] : (R)0;
        }
        for (l = 0; l < np; ++l) {
            R acc = 0;
            for (k = 0; k < np; ++k) acc += V[l + k * np] * w[k];
            h[l] = -acc;
        }
        for (k = 0; k < np; ++k) hn += h[k] * h[k];
        for (k = 0; k < np; ++k) xn += x[k] * x[k];
        if (RSQRT(hn) <= EPS_2 * (RSQRT(xn) + EPS_2)) {
            found = 1;
        } else {
            for (k = 0; k < np; ++k) x[k] += h[k];
            for (k = 0; k < np; ++k) T[k] = SFX(c_make)(RCOS(x[k]), -RSIN(x[k]));
            SFX(lm_A_g)(BhB, T, np, A, g);
            gmax = g[0];
            for (k = 1; k < np; ++k)
                if (g[k] > gmax) gmax = g[k];
            found = gmax <= EPS_1;
        }
    }
    for (k = 0; k < n; ++k) out_phase[k] = x[k];
    if (out_x)
        for (k = 0; k < np; ++k) out_x[k] = x[k];
    free(BhB);
    free(T);
    free(A);
    free(V);
    free(ev);
    free(w);
    free(g);
    free(h);
    free(x);
    return it;
}

/* Gradient descent: nls/gradient_descent.rs:34-61. x -= 0.1 g until
 * max(g) <= 1e-6 (signed max), at most k_max steps. Returns the steps taken. */
uint32_t SFX(orc_gradient_descent)(int source_type, const float* sources, const float* wavenums, const float* attens,
                                   uint64_t n, const float* foci, const float* amps, uint64_t m, uint32_t k_max,
                                   R* out_phase, R* out_x) {
    const R EPS = (R)1e-6;
    const uint64_t np = n + m;
    SFX(cplx)* BhB = SFX(nls_BhB)(source_type, sources, wavenums, attens, n, foci, amps, m);
    SFX(cplx)* T = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * np);
    R* A = (R*)malloc(sizeof(R) * np * np);
    R* g = (R*)malloc(sizeof(R) * np);
    R* x = (R*)calloc(np, sizeof(R));
    uint64_t k;
    uint32_t it;
    for (it = 0; it < k_max; ++it) {
        R gmax;
        for (k = 0; k < np; ++k) T[k] = SFX(c_make)(RCOS(x[k]), -RSIN(x[k]));
        SFX(lm_A_g)(BhB, T, np, A, g);
        gmax = g[0];
        for (k = 1; k < np; ++k)
            if (g[k] > gmax) gmax = g[k];
        if (gmax <= EPS) break;
        for (k = 0; k < np; ++k) x[k] = x[k] - (R)0.1 * g[k];
    }
    for (k = 0; k < n; ++k) out_phase[k] = x[k];
    if (out_x)
        for (k = 0; k < np; ++k) out_x[k] = x[k];
    free(BhB);
    free(T);
    free(A);
    free(g);
    free(x);
    return it;
}

/* Complex dense solve A x = b (column-major n x n, destroyed; b overwritten)
 * by Gaussian elimination with partial pivoting on |.|. */
static int SFX(c_solve)(SFX(cplx) * A, uint64_t n, SFX(cplx) * b) {
    uint64_t k, i, j;
    for (k = 0; k < n; ++k) {
        uint64_t piv = k;
        R best = SFX(c_abs)(A[k + k * n]);
        SFX(cplx) inv, t;
        for (i = k + 1; i < n; ++i)
            if (SFX(c_abs)(A[i + k * n]) > best) {
                best = SFX(c_abs)(A[i + k * n]);
                piv = i;
            }
        if (best == 0) return 0;
        if (piv != k) {
            for (j = 0; j < n; ++j) {
                t = A[k + j * n];
                A[k + j * n] = A[piv + j * n];
                A[piv + j * n] = t;
            }
            t = b[k];
            b[k] = b[piv];
            b[piv] = t;
        }
        {
            const R d = A[k + k * n].re * A[k + k * n].re + A[k + k * n].im * A[k + k * n].im;
            inv = SFX(c_make)(A[k + k * n].re / d, -A[k + k * n].im / d);
        }
        for (i = k + 1; i < n; ++i) {
            const SFX(cplx) f = SFX(c_mul)(A[i + k * n], inv);
            if (f.re == 0 && f.im == 0) continue;
            for (j = k; j < n; ++j) {
                const SFX(cplx) pr = SFX(c_mul)(f, A[k + j * n]);
                A[i + j * n] = SFX(c_make)(A[i + j * n].re - pr.re, A[i + j * n].im - pr.im);
            }
            {
                const SFX(cplx) pr = SFX(c_mul)(f, b[k]);
                b[i] = SFX(c_make)(b[i].re - pr.re, b[i].im - pr.im);
            }
        }
    }
    for (k = n; k-- > 0;) {
        SFX(cplx) acc = b[k];
        R d;
        for (j = k + 1; j < n; ++j) {
            const SFX(cplx) pr = SFX(c_mul)(A[k + j * n], b[j]);
            acc = SFX(c_make)(acc.re - pr.re, acc.im - pr.im);
        }
        d = A[k + k * n].re * A[k + k * n].re + A[k + k * n].im * A[k + k * n].im;
        b[k] = SFX(c_make)((acc.re * A[k + k * n].re + acc.im * A[k + k * n].im) / d,
                           (acc.im * A[k + k * n].re - acc.re * A[k + k * n].im) / d);
    }
    return 1;
}

/* J(q) of apo.rs:45-53: sum_i (|g_i q|^2 - p_i^2)^2 + lambda |q . q| (the
 * non-conjugated dot of nalgebra's dot()); g_i = row i of G (M x N). */
static R SFX(apo_J)(const SFX(cplx) * G, uint64_t m, uint64_t n, const R* p2, const SFX(cplx) * q, R lambda) {
    uint64_t i, j;
    R acc = 0;
    SFX(cplx) qq = SFX(c_make)(0, 0);
    for (i = 0; i < m; ++i) {
        SFX(cplx) gq = SFX(c_make)(0, 0);
        R sdiff;
        for (j = 0; j < n; ++j) gq = SFX(c_add)(gq, SFX(c_mul)(G[i + j * m], q[j]));
        sdiff = gq.re * gq.re + gq.im * gq.im - p2[i];
        acc += sdiff * sdiff;
    }
    for (j = 0; j < n; ++j) qq = SFX(c_add)(qq, SFX(c_mul)(q[j], q[j]));
    return acc + SFX(c_abs)(qq) * lambda;
}

/* grad of apo.rs:56-70: sum_i (R_i q) (p_i^2 - |q^H R_i q|) + lambda q, with
 * R_i = G^H D_i G, so R_i q = conj(g_i)^T (g_i q) and q^H R_i q = |g_i q|^2. */
static void SFX(apo_grad)(const SFX(cplx) * G, uint64_t m, uint64_t n, const R* p2, const SFX(cplx) * q, R lambda,
                          SFX(cplx) * out) {
    uint64_t i, j;
    for (j = 0; j < n; ++j) out[j] = SFX(c_make)(lambda * q[j].re, lambda * q[j].im);
    for (i = 0; i < m; ++i) {
        SFX(cplx) gq = SFX(c_make)(0, 0);
        R w;
        for (j = 0; j < n; ++j) gq = SFX(c_add)(gq, SFX(c_mul)(G[i + j * m], q[j]));
        w = p2[i] - (gq.re * gq.re + gq.im * gq.im);
        for (j = 0; j < n; ++j) {
            const SFX(cplx) r = SFX(c_mul)(SFX(c_conj)(G[i + j * m]), gq);
            out[j] = SFX(c_make)(out[j].re + r.re * w, out[j].im + r.im * w);
        }
    }
}

/* APO: nls/apo.rs:100-158 (Hasegawa et al. 2020). q0 = (G^H G + lambda I)^{-1}
 * G^H p; quasi-Newton with H = I initially, d = -H grad, alpha the first
 * minimiser of J(q + a d) over a = 0, 0.01, ..., 0.99, stop when |alpha d| <
 * 1e-8; H += y y^T / (y . s) - (H s)(H s)^T / (s^T H s) with y = alpha d,
 * s = grad_new - grad (transposes and dots without conjugation, as the
 * reference writes them). Output amp_j <- amp_j |q_j| / camax(q), phase
 * arg q_j, camax = max_j (|Re q_j| + |Im q_j|) (nalgebra's camax). Returns
 * the iterations taken. */
uint32_t SFX(orc_apo)(int source_type, const float* sources, const float* wavenums, const float* attens, uint64_t n,
                      const float* foci, const float* amps, uint64_t m, R lambda, uint32_t k_max, R* out_amp,
                      R* out_phase) {
    const R EPS = (R)1e-8;
    const uint64_t LS = 100;
    const size_t st = orc_stride(source_type);
    SFX(cplx)* G = SFX(make_G)(source_type, sources, wavenums, attens, n, foci, m);
    SFX(cplx)* A = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * n * n);
    SFX(cplx)* H = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * n * n);
    SFX(cplx)* q = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * n);
    SFX(cplx)* qn = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * n);
    SFX(cplx)* d = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * n);
    SFX(cplx)* gr = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * n);
    SFX(cplx)* grn = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * n);
    SFX(cplx)* sv = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * n);
    SFX(cplx)* u = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * n);
    R* p2 = (R*)malloc(sizeof(R) * m);
    uint64_t i, j, k;
    uint32_t it;
    R cmax = 0;
    for (i = 0; i < m; ++i) p2[i] = (R)amps[i] * (R)amps[i];
    /* (G^H G + lambda I) q0 = G^H p */
    for (j = 0; j < n; ++j) {
        SFX(cplx) b = SFX(c_make)(0, 0);
        for (k = 0; k < n; ++k) {
            SFX(cplx) acc = SFX(c_make)(0, 0);
            for (i = 0; i < m; ++i) acc = SFX(c_add)(acc, SFX(c_mul)(SFX(c_conj)(G[i + k * m]), G[i + j * m]));
            if (k == j) acc.re += lambda;
            A[k + j * n] = acc;
        }
        for (i = 0; i < m; ++i) b = SFX(c_add)(b, SFX(c_mul)(SFX(c_conj)(G[i + j * m]), SFX(c_make)((R)amps[i], 0)));
        q[j] = b;
    }
    SFX(c_solve)(A, n, q);
    for (j = 0; j < n; ++j)
        for (k = 0; k < n; ++k) H[j + k * n] = SFX(c_make)(j == k ? (R)1 : (R)0, 0);
    SFX(apo_grad)(G, m, n, p2, q, lambda, gr);
    for (it = 0; it < k_max; ++it) {
        R alpha = 0, best = (R)INFINITY, dn = 0;
        uint64_t l;
        SFX(cplx) ys = SFX(c_make)(0, 0), su = SFX(c_make)(0, 0);
        for (j = 0; j < n; ++j) {
            SFX(cplx) acc = SFX(c_make)(0, 0);
            for (k = 0; k < n; ++k) acc = SFX(c_add)(acc, SFX(c_mul)(H[j + k * n], gr[k]));
            d[j] = SFX(c_make)(-acc.re, -acc.im);
        }
        for (l = 0; l < LS; ++l) {
            const R a = (R)l / (R)LS;
            R v;
            for (j = 0; j < n; ++j) qn[j] = SFX(c_make)(q[j].re + d[j].re * a, q[j].im + d[j].im * a);
            v = SFX(apo_J)(G, m, n, p2, qn, lambda);
            if (v < best) {
                alpha = a;
                best = v;
            }
        }
        for (j = 0; j < n; ++j) d[j] = SFX(c_make)(d[j].re * alpha, d[j].im * alpha);
        for (j = 0; j < n; ++j) dn += d[j].re * d[j].re + d[j].im * d[j].im;
        if (RSQRT(dn) < EPS) break;
        for (j = 0; j < n; ++j) qn[j] = SFX(c_add)(q[j], d[j]);
        SFX(apo_grad)(G, m, n, p2, qn, lambda, grn);
        for (j = 0; j < n; ++j) sv[j] = SFX(c_make)(grn[j].re - gr[j].re, grn[j].im - gr[j].im);
        /* u = H s, y . s, s^T H s = s . u */
        for (j = 0; j < n; ++j) {
            SFX(cplx) acc = SFX(c_make)(0, 0);
            for (k = 0; k < n; ++k) acc = SFX(c_add)(acc, SFX(c_mul)(H[j + k * n], sv[k]));
            u[j] = acc;
        }
        for (j = 0; j < n; ++j) ys = SFX(c_add)(ys, SFX(c_mul)(d[j], sv[j]));
        for (j = 0; j < n; ++j) su = SFX(c_add)(su, SFX(c_mul)(sv[j], u[j]));
        {
            const R dy = ys.re * ys.re + ys.im * ys.im, du = su.re * su.re + su.im * su.im;
            const SFX(cplx) iy = SFX(c_make)(ys.re / dy, -ys.im / dy), iu = SFX(c_make)(su.re / du, -su.im / du);
            for (k = 0; k < n; ++k)
                for (j = 0; j < n; ++j) {
                    const SFX(cplx) a1 = SFX(c_mul)(SFX(c_mul)(d[j], d[k]), iy);
                    const SFX(cplx) a2 = SFX(c_mul)(SFX(c_mul)(u[j], u[k]), iu);
                    H[j + k * n] = SFX(c_make)(H[j + k * n].re + a1.re - a2.re, H[j + k * n].im + a1.im - a2.im);
                }
        }
        memcpy(q, qn, sizeof(SFX(cplx)) * n);
        memcpy(gr, grn, sizeof(SFX(cplx)) * n);
    }
    for (j = 0; j < n; ++j) {
        const R c1 = RABS(q[j].re) + RABS(q[j].im);
        if (c1 > cmax) cmax = c1;
    }
    for (j = 0; j < n; ++j) {
        const R amp = (source_type == 1) ? (R)sources[st * j + 6] : (R)sources[st * j + 3];
        out_amp[j] = amp * SFX(c_abs)(q[j]) / cmax;
        out_phase[j] = SFX(c_arg)(q[j]);
    }
    free(G);
    free(A);
    free(H);
    free(q);
    free(qn);
    free(d);
    free(gr);
    free(grn);
    free(sv);
    free(u);
    free(p2);
    return it;
}

/* Eigen-decomposition of a Hermitian matrix (column-major n x n, the full
 * matrix given, destroyed) by cyclic complex Jacobi rotations: for each pair
 * (p, q) the off-diagonal entry a_pq = |a_pq| e^{i phi} is made real by
 * scaling column q by e^{-i phi} (row q by e^{i phi}), then a real rotation
 * zeroes it. evals[k] real, eigenvectors in the columns of V. Sweeps stop when
 * the off-diagonal mass is below 1e-30 of the diagonal's. Output convention
 * (nalgebra's SymmetricEigen, which the reference uses, fixes neither order
 * nor phases in a documented way): eigenvalues ascending, and each
 * eigenvector's largest-modulus entry real positive. */
void SFX(orc_herm_eig)(SFX(cplx) * A, uint64_t n, R* evals, SFX(cplx) * V) {
    uint64_t p, q, k, sweep;
    for (p = 0; p < n; ++p)
        for (q = 0; q < n; ++q) V[p + q * n] = SFX(c_make)(p == q ? (R)1 : (R)0, 0);
    for (sweep = 0; sweep < 100; ++sweep) {
        R off = 0, dg = 0;
        for (q = 0; q < n; ++q)
            for (p = 0; p < n; ++p) {
                const R m2 = A[p + q * n].re * A[p + q * n].re + A[p + q * n].im * A[p + q * n].im;
                if (p != q) off += m2;
                else dg += m2;
            }
        if (off <= dg * (R)1e-30 || off == 0) break;
        for (p = 0; p + 1 < n; ++p)
            for (q = p + 1; q < n; ++q) {
                const SFX(cplx) apq = A[p + q * n];
                const R b = SFX(c_abs)(apq);
                R theta, t, c, sn;
                SFX(cplx) ph, phc;
                if (b == 0) continue;
                ph = SFX(c_make)(apq.re / b, apq.im / b); /* e^{i phi} */
                phc = SFX(c_conj)(ph);
                for (k = 0; k < n; ++k) A[k + q * n] = SFX(c_mul)(A[k + q * n], phc);
                for (k = 0; k < n; ++k) A[q + k * n] = SFX(c_mul)(A[q + k * n], ph);
                for (k = 0; k < n; ++k) V[k + q * n] = SFX(c_mul)(V[k + q * n], phc);
                theta = (A[q + q * n].re - A[p + p * n].re) / ((R)2 * b);
                t = (theta >= 0 ? (R)1 : (R)-1) / (RABS(theta) + RSQRT(theta * theta + (R)1));
                c = (R)1 / RSQRT(t * t + (R)1);
                sn = t * c;
                for (k = 0; k < n; ++k) {
                    const SFX(cplx) akp = A[k + p * n], akq = A[k + q * n];
                    A[k + p * n] = SFX(c_make)(c * akp.re - sn * akq.re, c * akp.im - sn * akq.im);
                    A[k + q * n] = SFX(c_make)(sn * akp.re + c * akq.re, sn * akp.im + c * akq.im);
                }
                for (k = 0; k < n; ++k) {
                    const SFX(cplx) apk = A[p + k * n], aqk = A[q + k * n];
                    A[p + k * n] = SFX(c_make)(c * apk.re - sn * aqk.re, c * apk.im - sn * aqk.im);
                    A[q + k * n] = SFX(c_make)(sn * apk.re + c * aqk.re, sn * apk.im + c * aqk.im);
                }
                for (k = 0; k < n; ++k) {
                    const SFX(cplx) vkp = V[k + p * n], vkq = V[k + q * n];
                    V[k + p * n] = SFX(c_make)(c * vkp.re - sn * vkq.re, c * vkp.im - sn * vkq.im);
                    V[k + q * n] = SFX(c_make)(sn * vkp.re + c * vkq.re, sn * vkp.im + c * vkq.im);
                }
            }
    }
    /* ascending eigenvalues (stable for ties), each eigenvector scaled so
     * that its largest-modulus entry (the first on ties) is real positive;
     * A is reused for the permuted vectors */
    R* dg = (R*)malloc(sizeof(R) * (n ? n : 1));
    for (k = 0; k < n; ++k) dg[k] = A[k + k * n].re;
    for (k = 0; k < n; ++k) {
        const R ek = dg[k];
        uint64_t rank = 0, j, jm = 0;
        R best = -1;
        SFX(cplx) ph;
        for (j = 0; j < n; ++j) {
            const R ej = dg[j];
            if (ej < ek || (ej == ek && j < k)) ++rank;
        }
        for (j = 0; j < n; ++j) {
            const R a2 = V[j + k * n].re * V[j + k * n].re + V[j + k * n].im * V[j + k * n].im;
            if (a2 > best) {
                best = a2;
                jm = j;
            }
        }
        ph = SFX(c_conj)(SFX(c_divr)(V[jm + k * n], RSQRT(best)));
        evals[rank] = ek;
        for (j = 0; j < n; ++j) A[j + rank * n] = SFX(c_mul)(V[j + k * n], ph);
        A[jm + rank * n].im = 0;
    }
    memcpy(V, A, sizeof(SFX(cplx)) * n * n);
    free(dg);
}

/* Index of the largest entry (the first one on ties), as nalgebra's imax. */
static uint64_t SFX(imax)(const R* v, uint64_t n) {
    uint64_t k, best = 0;
    for (k = 1; k < n; ++k)
        if (v[k] > v[best]) best = k;
    return best;
}

/* amp_j |q_j| / camax(q), arg q_j into the outputs (camax: nalgebra's
 * max_j |Re q_j| + |Im q_j|). */
static void SFX(write_amp_phase)(int source_type, const float* sources, uint64_t n, const SFX(cplx) * q, R* out_amp,
                                 R* out_phase) {
    const size_t st = orc_stride(source_type);
    uint64_t j;
    R cmax = 0;
    for (j = 0; j < n; ++j) {
        const R c1 = RABS(q[j].re) + RABS(q[j].im);
        if (c1 > cmax) cmax = c1;
    }
    for (j = 0; j < n; ++j) {
        const R amp = (source_type == 1) ? (R)sources[st * j + 6] : (R)sources[st * j + 3];
        out_amp[j] = amp * SFX(c_abs)(q[j]) / cmax;
        out_phase[j] = SFX(c_arg)(q[j]);
    }
}

/* Long: matrix/long.rs:46-106 (Long et al. 2014). denomi = G.column_sum()
 * (row sums, length M); X = (amps_i conj(G_ij) / denomi_i)^T; R = G X; the
 * Hermitian matrix read from R's lower triangle (SymmetricEigen reads only
 * that triangle; the diagonal's real part) is decomposed; e_arg_k =
 * arg(U[imax][k]) -- row imax of the eigenvector matrix, as written;
 * sigma_j = sqrt(sum_i |G_ij| amps_i / M); q solves (g^H g) q = g^H f with
 * g = [G; diag(sigma)], f = [amps e^{i e_arg}; 0] (dense complex solve). */
void SFX(orc_long)(int source_type, const float* sources, const float* wavenums, const float* attens, uint64_t n,
                   const float* foci, const float* amps, uint64_t m, R* out_amp, R* out_phase) {
    SFX(cplx)* G = SFX(make_G)(source_type, sources, wavenums, attens, n, foci, m);
    SFX(cplx)* den = (SFX(cplx)*)calloc(m, sizeof(SFX(cplx)));
    SFX(cplx)* Rm = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * m * m);
    SFX(cplx)* U = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * m * m);
    R* ev = (R*)malloc(sizeof(R) * m);
    R* sig2 = (R*)malloc(sizeof(R) * n);
    SFX(cplx)* A = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * n * n);
    SFX(cplx)* q = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * n);
    SFX(cplx)* f = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * m);
    uint64_t i, j, k, im;
    for (j = 0; j < n; ++j)
        for (i = 0; i < m; ++i) den[i] = SFX(c_add)(den[i], G[i + j * m]);
    for (k = 0; k < m; ++k)
        for (i = 0; i < m; ++i) {
            /* R_ik = sum_j G_ij X_jk, X_jk = amps_k conj(G_kj) / denomi_k */
            SFX(cplx) acc = SFX(c_make)(0, 0);
            const R dd = den[k].re * den[k].re + den[k].im * den[k].im;
            const SFX(cplx) inv = SFX(c_make)(den[k].re / dd, -den[k].im / dd);
            for (j = 0; j < n; ++j) {
                const SFX(cplx) x = SFX(c_mul)(SFX(c_mul)(SFX(c_make)((R)amps[k], 0), SFX(c_conj)(G[k + j * m])), inv);
                acc = SFX(c_add)(acc, SFX(c_mul)(G[i + j * m], x));
            }
            Rm[i + k * m] = acc;
        }
    /* Hermitian matrix from the lower triangle */
    for (k = 0; k < m; ++k)
        for (i = 0; i < m; ++i) {
            if (i > k) U[i + k * m] = Rm[i + k * m];
            else if (i < k) U[i + k * m] = SFX(c_conj)(Rm[k + i * m]);
            else U[i + k * m] = SFX(c_make)(Rm[i + i * m].re, 0);
        }
    memcpy(Rm, U, sizeof(SFX(cplx)) * m * m);
    SFX(orc_herm_eig)(Rm, m, ev, U);
    im = SFX(imax)(ev, m);
    for (k = 0; k < m; ++k) {
        const R e = SFX(c_arg)(U[im + k * m]);
        f[k] = SFX(c_make)((R)amps[k] * RCOS(e), (R)amps[k] * RSIN(e));
    }
    for (j = 0; j < n; ++j) {
        R acc = 0;
        for (i = 0; i < m; ++i) acc += SFX(c_abs)(G[i + j * m]) * (R)amps[i];
        sig2[j] = acc / (R)m; /* sigma_j^2 */
    }
    for (j = 0; j < n; ++j) {
        SFX(cplx) b = SFX(c_make)(0, 0);
        for (k = 0; k < n; ++k) {
            SFX(cplx) acc = SFX(c_make)(0, 0);
            for (i = 0; i < m; ++i) acc = SFX(c_add)(acc, SFX(c_mul)(SFX(c_conj)(G[i + k * m]), G[i + j * m]));
            if (k == j) acc.re += sig2[j];
            A[k + j * n] = acc;
        }
        for (i = 0; i < m; ++i) b = SFX(c_add)(b, SFX(c_mul)(SFX(c_conj)(G[i + j * m]), f[i]));
        q[j] = b;
    }
    SFX(c_solve)(A, n, q);
    SFX(write_amp_phase)(source_type, sources, n, q, out_amp, out_phase);
    free(G);
    free(den);
    free(Rm);
    free(U);
    free(ev);
    free(sig2);
    free(A);
    free(q);
    free(f);
}

/* The random row choice of Horn: splitmix64 over (seed, instance), a value
 * in [0, 1) with 24 bits, ii = (uint)(M * r) in f32 as horn.rs:93 writes it.
 * (The reference draws from thread_rng, unseeded.) */
#ifndef ORC_SPLITMIX_DEFINED
#define ORC_SPLITMIX_DEFINED
static uint64_t orc_splitmix(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
#endif

/* Horn: matrix/horn.rs:58-129 (Inoue et al. 2015); the reference's
 * constants are REPEAT 100, lambda 0.8, Tikhonov alpha 1e-5. With the
 * thin SVD G = U S V^H, G_pinv = V diag(s / (s^2 + a^2)) U^H equals
 * G^H K^{-1} with K = G G^H + a^2 I (push-through), and so
 * I - G G_pinv = a^2 K^{-1} exactly; both are evaluated in that form (in
 * f32 the reference's I - G G_pinv cancels to rounding noise at a = 1e-5).
 * MM = P a^2 K^{-1} P; m * repeat random single-row updates of X (the row
 * drawn by orc_splitmix from the seed, as rand's [0, 1) f32 scaled by m;
 * the reference draws from thread_rng, unseeded); q = G_pinv P u with u the
 * eigenvector of X's largest eigenvalue. */
void SFX(orc_horn)(int source_type, const float* sources, const float* wavenums, const float* attens, uint64_t n,
                   const float* foci, const float* amps, uint64_t m, uint64_t seed, uint64_t repeat, R lambda,
                   R alpha, R* out_amp, R* out_phase) {
    SFX(cplx)* G = SFX(make_G)(source_type, sources, wavenums, attens, n, foci, m);
    SFX(cplx)* K = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * m * m);
    SFX(cplx)* Ki = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * m * m);
    SFX(cplx)* col = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * m);
    SFX(cplx)* MM = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * m * m);
    SFX(cplx)* X = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * m * m);
    SFX(cplx)* U = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * m * m);
    SFX(cplx)* xb = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * m);
    SFX(cplx)* z = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * m);
    SFX(cplx)* q = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * n);
    R* ev = (R*)malloc(sizeof(R) * m);
    uint64_t i, j, k, r, it, im;
    uint64_t rs = seed;
    /* K^{-1} column by column (K = G G^H + a^2 I is destroyed by each solve) */
    for (k = 0; k < m; ++k) {
        for (j = 0; j < m; ++j)
            for (i = 0; i < m; ++i) {
                SFX(cplx) acc = SFX(c_make)(0, 0);
                uint64_t l;
                for (l = 0; l < n; ++l) acc = SFX(c_add)(acc, SFX(c_mul)(G[i + l * m], SFX(c_conj)(G[j + l * m])));
                if (i == j) acc.re += alpha * alpha;
                K[i + j * m] = acc;
            }
        for (i = 0; i < m; ++i) col[i] = SFX(c_make)(i == k ? (R)1 : (R)0, 0);
        SFX(c_solve)(K, m, col);
        for (i = 0; i < m; ++i) Ki[i + k * m] = col[i];
    }
    /* MM = P (I - G G_pinv) P = a^2 P K^{-1} P */
    for (k = 0; k < m; ++k)
        for (i = 0; i < m; ++i) {
            const R sc = alpha * alpha * (R)amps[i] * (R)amps[k];
            MM[i + k * m] = SFX(c_make)(Ki[i + k * m].re * sc, Ki[i + k * m].im * sc);
        }
    for (k = 0; k < m; ++k)
        for (i = 0; i < m; ++i) X[i + k * m] = SFX(c_make)(i == k ? (R)1 : (R)0, 0);
    for (it = 0; it < m * repeat; ++it) {
        const float rnd = (float)(orc_splitmix(&rs) >> 40) * (1.0f / 16777216.0f);
        uint64_t ii = (uint64_t)((float)m * rnd);
        SFX(cplx) gamma = SFX(c_make)(0, 0);
        if (ii > m - 1) ii = m - 1;
        /* Xb = Xc MMc over the rows / columns other than ii */
        for (r = 0; r < m; ++r) {
            SFX(cplx) acc = SFX(c_make)(0, 0);
            if (r == ii) continue;
            for (k = 0; k < m; ++k)
                if (k != ii) acc = SFX(c_add)(acc, SFX(c_mul)(X[r + k * m], MM[k + ii * m]));
            xb[r] = acc;
        }
        for (r = 0; r < m; ++r)
            if (r != ii) gamma = SFX(c_add)(gamma, SFX(c_mul)(SFX(c_conj)(xb[r]), MM[r + ii * m]));
        if (gamma.re > 0) {
            const R sc = -RSQRT(lambda / gamma.re);
            for (r = 0; r < m; ++r) {
                if (r == ii) continue;
                X[r + ii * m] = SFX(c_make)(xb[r].re * sc, xb[r].im * sc);
                X[ii + r * m] = SFX(c_conj)(X[r + ii * m]);
            }
        } else {
            for (r = 0; r < m; ++r) {
                if (r == ii) continue;
                X[r + ii * m] = SFX(c_make)(0, 0);
                X[ii + r * m] = SFX(c_make)(0, 0);
            }
        }
    }
    memcpy(MM, X, sizeof(SFX(cplx)) * m * m);
    SFX(orc_herm_eig)(MM, m, ev, U);
    im = SFX(imax)(ev, m);
    /* z = K^{-1} P u, q = G^H z */
    for (i = 0; i < m; ++i) {
        SFX(cplx) acc = SFX(c_make)(0, 0);
        for (k = 0; k < m; ++k)
            acc = SFX(c_add)(acc, SFX(c_mul)(Ki[i + k * m], SFX(c_mul)(SFX(c_make)((R)amps[k], 0), U[k + im * m])));
        z[i] = acc;
    }
    for (j = 0; j < n; ++j) {
        SFX(cplx) acc = SFX(c_make)(0, 0);
        for (i = 0; i < m; ++i) acc = SFX(c_add)(acc, SFX(c_mul)(SFX(c_conj)(G[i + j * m]), z[i]));
        q[j] = acc;
    }
    SFX(write_amp_phase)(source_type, sources, n, q, out_amp, out_phase);
    free(G);
    free(K);
    free(Ki);
    free(col);
    free(MM);
    free(X);
    free(U);
    free(xb);
    free(z);
    free(q);
    free(ev);
}

/* F(x) and g(x) of the LM objective for given x (nls/macros.rs:17-58), for
 * gradient checks: returns F, writes g (n + m). */
R SFX(orc_lm_eval)(int source_type, const float* sources, const float* wavenums, const float* attens, uint64_t n,
                   const float* foci, const float* amps, uint64_t m, const R* x, R* g_out) {
    const uint64_t np = n + m;
    SFX(cplx)* G = SFX(make_G)(source_type, sources, wavenums, attens, n, foci, m);
    SFX(cplx)* Bm = (SFX(cplx)*)calloc(m * np, sizeof(SFX(cplx)));
    SFX(cplx)* BhB = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * np * np);
    SFX(cplx)* T = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * np);
    SFX(cplx)* tv = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * np);
    SFX(cplx)* row = (SFX(cplx)*)malloc(sizeof(SFX(cplx)) * np);
    R* A = (R*)malloc(sizeof(R) * np * np);
    R F;
    uint64_t i, j, k;
    for (j = 0; j < n; ++j)
        for (i = 0; i < m; ++i) Bm[i + j * m] = G[i + j * m];
    for (i = 0; i < m; ++i) Bm[i + (n + i) * m] = SFX(c_make)(-(R)amps[i], 0);
    for (j = 0; j < np; ++j) {
        for (k = 0; k < np; ++k) BhB[k + j * np] = SFX(c_make)(0, 0);
        for (i = 0; i < m; ++i)
            for (k = 0; k < np; ++k)
                BhB[k + j * np] = SFX(c_add)(SFX(c_mul)(Bm[i + j * m], SFX(c_conj)(Bm[i + k * m])), BhB[k + j * np]);
    }
    for (k = 0; k < np; ++k) T[k] = SFX(c_make)(RCOS(x[k]), -RSIN(x[k]));
    SFX(lm_A_g)(BhB, T, np, A, g_out);
    F = SFX(lm_Fx)(BhB, x, np, tv, row);
    free(G);
    free(Bm);
    free(BhB);
    free(T);
    free(tv);
    free(row);
    free(A);
    return F;
}

/* FocalPoint phases: acoustic-field-optimizer/src/focal_point.rs:28-39 */
/* BesselBeam phases: acoustic-field-optimizer/src/bessel.rs:31-55, in its
 * operation order. v = (dir_y, -dir_x, 0), theta_w = asin |v|; r = s - point;
 * r' = r cos theta_w + (r x v) sin theta_w + v (v.r (1 - cos theta_w)) (the
 * axis v is used unnormalised, as written); dist = sin theta |r'_xy| -
 * cos theta r'_z; phase = -2 pi ((dist mod lambda) / lambda). */
void SFX(orc_bessel_phases)(int source_type, const float* sources, uint64_t n, R sound_speed, R px, R py, R pz,
                            R dx, R dy, R dz, R theta, R* out_phase) {
    const size_t st = orc_stride(source_type);
    const R pi = SFX(ORC_PI);
    const R vx = dy, vy = -dx, vz = 0;
    const R tw = RASIN(RSQRT_F((vx * vx + vy * vy) + vz * vz));
    const R ct = RCOS(tw), stw = RSIN(tw);
    uint64_t j;
    (void)dz;
    for (j = 0; j < n; ++j) {
        const float* s = sources + st * j;
        const R rx = (R)s[0] - px, ry = (R)s[1] - py, rz = (R)s[2] - pz;
        const R xx = ry * vz - rz * vy, xy = rz * vx - rx * vz, xz = rx * vy - ry * vx;
        const R vr = ((vx * rx + vy * ry) + vz * rz) * ((R)1 - ct);
        const R qx = (rx * ct + xx * stw) + vx * vr;
        const R qy = (ry * ct + xy * stw) + vy * vr;
        const R qz = (rz * ct + xz * stw) + vz * vr;
        const R dist = RSIN(theta) * RSQRT_F(qx * qx + qy * qy) - RCOS(theta) * qz;
        const R wl = sound_speed / (R)s[st - 1];
        out_phase[j] = (R)-2.0 * pi * (RFMOD(dist, wl) / wl);
    }
}

void SFX(orc_focal_point_phases)(int source_type, const float* sources, uint64_t n, R sound_speed, R fx, R fy, R fz,
                                 R* out_phase) {
    const size_t st = orc_stride(source_type);
    const R pi = SFX(ORC_PI);
    uint64_t j;
    for (j = 0; j < n; ++j) {
        const float* s = sources + st * j;
        R dx = (R)s[0] - fx, dy = (R)s[1] - fy, dz = (R)s[2] - fz;
        R d = RSQRT_F((dx * dx + dy * dy) + dz * dz);
        R freq = (R)s[st - 1];
        R wl = sound_speed / freq;
        R ph = RFMOD(d, wl) / wl;
        out_phase[j] = (R)-2.0 * pi * ph;
    }
}
