// Matrix-method hologram solvers on the GPU, batched, FP64: Long
// (acoustic-field-optimizer/src/multiple_foci/matrix/long.rs:46-106) and
// Horn (matrix/horn.rs:58-129).
//
// Both reduce to M x M complex work once G (M x N) is known, and run one
// workgroup per instance with the M x M matrices in LDS:
//
// Long: denomi_i = sum_j G_ij; R = (G G^H) diag(amps / denomi); the Hermitian
//   matrix read from R's lower triangle (diagonal real part) is decomposed
//   and e_arg_k = arg U[imax][k] (a row of the eigenvector matrix, as
//   long.rs:76-80 reads it); f = amps e^{i e_arg}; with D = diag(sigma^2),
//   sigma_j^2 = sum_i |G_ij| amps_i / M, the reference's N x N normal
//   equations (G^H G + D) q = G^H f are solved through the identity
//   q = D^{-1} G^H (I + G D^{-1} G^H)^{-1} f (an M x M system).
// Horn: K = G G^H + a^2 I, G_pinv = G^H K^{-1} (= V diag(s / (s^2 + a^2)) U^H
//   of horn.rs:58-68) and I - G G_pinv = a^2 K^{-1}, so MM = a^2 P K^{-1} P
//   (in f32 the reference's I - G G_pinv cancels to rounding noise at
//   a = 1e-5); m * repeat single-row updates of X with the row drawn from a
//   seeded splitmix64 stream (the reference uses rand's thread_rng,
//   unseeded); q = G^H K^{-1} P u, u the eigenvector of X's largest
//   eigenvalue.
//
// Eigen-decompositions: one wave, cyclic complex Jacobi (each pivot made real
// by a diagonal phase, then a real rotation). Output convention: eigenvalues
// ascending, each eigenvector's largest-modulus entry real positive
// (nalgebra's SymmetricEigen, used by the reference, fixes neither). Linear
// systems (Hermitian positive definite): one-wave Gauss-Jordan, a lane per
// column. Outputs are scaled by camax = max_j |Re q_j| + |Im q_j| as the
// reference does (long.rs:97-104, horn.rs:121-128).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <string>

#include "afh_dense.hpp"
#include "afh_holo.hpp"
#include "afh_internal.hpp"

namespace afh {

constexpr int kMatMaxM = 32;
constexpr int kMS = kMatMaxM + 1;      // row stride (double2) of M x M LDS matrices
constexpr int kAS = 2 * kMatMaxM + 1;  // row stride of the augmented [K | B]

struct MatArgs {
    const float* sources;  // packed (AFH_SOURCE_STRIDE floats each)
    int source_type;
    uint32_t n;
    const float* foci;
    const float* amps;
    const uint32_t* offsets;
    uint64_t seed;  // Horn: instance b draws from seed + b
    uint32_t repeat;
    double lambda, alpha;
    char* scratch;  // per instance, scratch_stride bytes
    uint64_t scratch_stride;
    float* out_amp;    // batch x N
    float* out_phase;  // batch x N
};

// Scratch of one instance: G (M x N, double2) | q (N, double2) | sigma^2 (N).
__host__ __device__ inline uint64_t mat_scratch_bytes(uint64_t N, uint64_t M) {
    auto al = [](uint64_t b) { return (b + 255) & ~(uint64_t)255; };
    return al(M * N * 16) + al(N * 16) + al(N * 8);
}

struct MatScratch {
    double2 *G, *q;
    double* s2;
};

__device__ inline MatScratch mat_scratch(char* base, uint64_t N, uint64_t M) {
    auto al = [](uint64_t b) { return (b + 255) & ~(uint64_t)255; };
    MatScratch s;
    s.G = reinterpret_cast<double2*>(base);
    s.q = reinterpret_cast<double2*>(base + al(M * N * 16));
    s.s2 = reinterpret_cast<double*>(base + al(M * N * 16) + al(N * 16));
    return s;
}

__global__ __launch_bounds__(kLB) void mat_gen_g(MatArgs a) {
    const uint32_t b = blockIdx.y;
    const uint32_t j = blockIdx.x * kLB + threadIdx.x;
    const uint32_t base = a.offsets[b], M = a.offsets[b + 1] - base;
    if (j >= a.n) return;
    MatScratch s = mat_scratch(a.scratch + b * a.scratch_stride, a.n, M);
    const float* src = a.sources + AFH_SOURCE_STRIDE * (uint64_t)j;
    for (uint32_t i = 0; i < M; ++i)
        s.G[(uint64_t)i * a.n + j] = term_f64(src, a.source_type, a.foci[3 * (uint64_t)(base + i)],
                                              a.foci[3 * (uint64_t)(base + i) + 1], a.foci[3 * (uint64_t)(base + i) + 2]);
}

__device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }

__device__ inline void wave_fence() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

// Eigen-decomposition of the Hermitian M x M matrix A (LDS, row stride kMS,
// full matrix) by wave 0 (threadIdx.x < 64; lane k owns row / column k).
// On return A holds the eigenvectors as columns and ev[0..M) the eigenvalues,
// in the convention above; V and raw are work arrays (V M x M, stride kMS).
// Site gives each calling kernel its own instance, so that the LDS arrays stay
// compile-time addresses inside the function.
template <int Site>
__device__ __attribute__((noinline)) void herm_eig_wave(double2* A, double2* V, double* ev, double* raw, uint32_t M) {
    const int k = threadIdx.x;
    const bool live = k < (int)M;
    if (live)
        for (uint32_t j = 0; j < M; ++j) V[k * kMS + j] = make_double2(j == (uint32_t)k ? 1.0 : 0.0, 0.0);
    wave_fence();
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, dg = 0.0;
        if (live)
            for (uint32_t j = 0; j < M; ++j) {
                const double2 v = A[k * kMS + j];
                const double m2 = v.x * v.x + v.y * v.y;
                if (j == (uint32_t)k)
                    dg += m2;
                else
                    off += m2;
            }
        off = wave_sum_f64(off);
        dg = wave_sum_f64(dg);
        if (off <= dg * 1e-30 || off == 0.0) break;
        for (uint32_t p = 0; p + 1 < M; ++p)
            for (uint32_t q = p + 1; q < M; ++q) {
                const double2 apq = A[p * kMS + q];
                const double bm = hypot(apq.x, apq.y);
                if (bm == 0.0) continue;
                const double2 ph = make_double2(apq.x / bm, apq.y / bm), phc = cconj(ph);
                wave_fence();
                if (live) A[k * kMS + q] = cmul(A[k * kMS + q], phc);  // column q
                wave_fence();
                if (live) {
                    A[q * kMS + k] = cmul(A[q * kMS + k], ph);  // row q
                    V[k * kMS + q] = cmul(V[k * kMS + q], phc);
                }
                wave_fence();
                const double theta = (A[q * kMS + q].x - A[p * kMS + p].x) / (2.0 * bm);
                const double tt = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double c = 1.0 / sqrt(tt * tt + 1.0), sn = tt * c;
                wave_fence();
                if (live) {
                    const double2 akp = A[k * kMS + p], akq = A[k * kMS + q];
                    A[k * kMS + p] = make_double2(c * akp.x - sn * akq.x, c * akp.y - sn * akq.y);
                    A[k * kMS + q] = make_double2(sn * akp.x + c * akq.x, sn * akp.y + c * akq.y);
                }
                wave_fence();
                if (live) {
                    const double2 apk = A[p * kMS + k], aqk = A[q * kMS + k];
                    A[p * kMS + k] = make_double2(c * apk.x - sn * aqk.x, c * apk.y - sn * aqk.y);
                    A[q * kMS + k] = make_double2(sn * apk.x + c * aqk.x, sn * apk.y + c * aqk.y);
                    const double2 vkp = V[k * kMS + p], vkq = V[k * kMS + q];
                    V[k * kMS + p] = make_double2(c * vkp.x - sn * vkq.x, c * vkp.y - sn * vkq.y);
                    V[k * kMS + q] = make_double2(sn * vkp.x + c * vkq.x, sn * vkp.y + c * vkq.y);
                }
                wave_fence();
            }
    }
    if (live) raw[k] = A[k * kMS + k].x;
    wave_fence();
    if (live) {
        const double ek = raw[k];
        uint32_t rank = 0, jm = 0;
        double best = -1.0;
        for (uint32_t j = 0; j < M; ++j) {
            const double ej = raw[j];
            if (ej < ek || (ej == ek && j < (uint32_t)k)) ++rank;
            const double2 v = V[j * kMS + k];
            const double a2 = v.x * v.x + v.y * v.y;
            if (a2 > best) {
                best = a2;
                jm = j;
            }
        }
        const double nb = sqrt(best);
        const double2 vm = V[jm * kMS + k];
        const double2 ph = make_double2(vm.x / nb, -vm.y / nb);
        ev[rank] = ek;
        for (uint32_t j = 0; j < M; ++j) A[j * kMS + rank] = j == jm ? make_double2(nb, 0.0) : cmul(V[j * kMS + k], ph);
    }
    wave_fence();
}

// Solves the Hermitian positive definite system in the augmented matrix
// [K | B] (LDS, M rows, ncols <= 64 columns, row stride kAS) by Gauss-Jordan
// elimination without pivoting, by wave 0 (lane c owns column c): on return
// columns M..ncols-1 hold K^{-1} B. colbuf holds >= M entries.
template <int Site>
__device__ __attribute__((noinline)) void gj_solve_wave(double2* A, uint32_t M, uint32_t ncols, double2* colbuf) {
    const int c = threadIdx.x;
    for (uint32_t p = 0; p < M; ++p) {
        wave_fence();
        if (c < (int)M) colbuf[c] = A[c * kAS + p];
        wave_fence();
        const double2 piv = colbuf[p];
        const double n2 = piv.x * piv.x + piv.y * piv.y;
        const double2 inv = make_double2(piv.x / n2, -piv.y / n2);
        if (c < (int)ncols) {
            const double2 rp = cmul(A[p * kAS + c], inv);
            A[p * kAS + c] = rp;
            for (uint32_t r = 0; r < M; ++r) {
                if (r == p) continue;
                const double2 f = colbuf[r], v = A[r * kAS + c];
                const double2 fr = cmul(f, rp);
                A[r * kAS + c] = make_double2(v.x - fr.x, v.y - fr.y);
            }
        }
    }
    wave_fence();
}

// sum_l G_il conj(G_kl) w_l (w = null: weight 1) for the lower triangle k <= i,
// into A (row stride ld) with the conjugate above the diagonal, plus diag_add
// on the diagonal (whose imaginary part is zero by symmetry).
__device__ inline void gram_herm(const double2* G, const double* w, uint32_t N, uint32_t M, double diag_add,
                                 double2* A, int ld) {
    for (uint32_t e = threadIdx.x; e < M * M; e += kLB) {
        const uint32_t i = e / M, k = e % M;
        if (k > i) continue;
        double re = 0.0, im = 0.0;
#pragma unroll 8
        for (uint32_t l = 0; l < N; ++l) {
            const double2 gi = G[(uint64_t)i * N + l], gk = G[(uint64_t)k * N + l];
            double pr = gi.x * gk.x + gi.y * gk.y, pi = gi.y * gk.x - gi.x * gk.y;
            if (w) {
                pr /= w[l];
                pi /= w[l];
            }
            re += pr;
            im += pi;
        }
        if (i == k) {
            A[i * ld + i] = make_double2(re + diag_add, 0.0);
        } else {
            A[i * ld + k] = make_double2(re, im);
            A[k * ld + i] = make_double2(re, -im);
        }
    }
}

// q_j (scratch) -> amp_j = source amp |q_j| / camax, phase_j = arg q_j.
__device__ inline void write_camax(const MatArgs& a, const MatScratch& s, uint32_t b, uint32_t N, double* red) {
    double cm = 0.0;
    for (uint32_t j = threadIdx.x; j < N; j += kLB) cm = fmax(cm, fabs(s.q[j].x) + fabs(s.q[j].y));
    cm = block_max(cm, red);
    for (uint32_t j = threadIdx.x; j < N; j += kLB) {
        const double2 qv = s.q[j];
        const double amp = (double)a.sources[AFH_SOURCE_STRIDE * (uint64_t)j + 6];
        a.out_amp[(uint64_t)b * N + j] = (float)(amp * sqrt(qv.x * qv.x + qv.y * qv.y) / cm);
        a.out_phase[(uint64_t)b * N + j] = (float)atan2(qv.y, qv.x);
    }
}

// Index of the largest eigenvalue (the first on ties), as nalgebra's imax.
__device__ inline uint32_t imax_of(const double* ev, uint32_t M) {
    uint32_t im = 0;
    for (uint32_t k = 1; k < M; ++k)
        if (ev[k] > ev[im]) im = k;
    return im;
}

__global__ __launch_bounds__(kLB) void long_solve(MatArgs a) {
    __shared__ double2 s_A[kMatMaxM * kAS];
    __shared__ double2 s_V[kMatMaxM * kMS];
    __shared__ double2 s_den[kMatMaxM], s_col[64];
    __shared__ double s_ev[kMatMaxM], s_raw[kMatMaxM], s_red[4];
    const int t = threadIdx.x;
    const uint32_t b = blockIdx.x;
    const uint32_t base = a.offsets[b], M = a.offsets[b + 1] - base;
    const uint32_t N = a.n;
    const float* amps = a.amps + base;
    MatScratch s = mat_scratch(a.scratch + b * a.scratch_stride, N, M);
    // sigma_j^2 = sum_i |G_ij| amps_i / M
    for (uint32_t j = t; j < N; j += kLB) {
        double acc = 0.0;
#pragma unroll 4
        for (uint32_t i = 0; i < M; ++i) {
            const double2 g = s.G[(uint64_t)i * N + j];
            acc += sqrt(g.x * g.x + g.y * g.y) * (double)amps[i];
        }
        s.s2[j] = acc / (double)M;
    }
    // denomi_i = sum_j G_ij (G.column_sum(): a wave per row)
    {
        const int w = t >> 6, l = t & 63;
        for (uint32_t i = w; i < M; i += kLB / 64) {
            double re = 0.0, im = 0.0;
            for (uint32_t j = l; j < N; j += 64) {
                const double2 g = s.G[(uint64_t)i * N + j];
                re += g.x;
                im += g.y;
            }
            re = wave_sum_f64(re);
            im = wave_sum_f64(im);
            if (l == 0) s_den[i] = make_double2(re, im);
        }
    }
    // R = (G G^H) diag(amps / denomi), read through its lower triangle
    gram_herm(s.G, nullptr, N, M, 0.0, s_A, kMS);
    __syncthreads();
    for (uint32_t e = t; e < M * M; e += kLB) {
        const uint32_t i = e / M, k = e % M;
        if (k > i) continue;
        const double2 dk = s_den[k];
        const double dd = dk.x * dk.x + dk.y * dk.y;
        const double2 sk = make_double2((double)amps[k] * dk.x / dd, -(double)amps[k] * dk.y / dd);
        const double2 r = cmul(s_A[i * kMS + k], sk);
        if (i == k) {
            s_A[i * kMS + i] = make_double2(r.x, 0.0);
        } else {
            s_A[i * kMS + k] = r;
            s_A[k * kMS + i] = cconj(r);
        }
    }
    __syncthreads();
    if (t < 64) herm_eig_wave<0>(s_A, s_V, s_ev, s_raw, M);
    __syncthreads();
    // f_k = amps_k e^{i arg U[imax][k]}
    if (t < (int)M) {
        const double2 u = s_A[imax_of(s_ev, M) * kMS + t];
        double sn, cs;
        sincos(atan2(u.y, u.x), &sn, &cs);
        s_col[t] = make_double2((double)amps[t] * cs, (double)amps[t] * sn);
    }
    __syncthreads();
    // (I + G D^{-1} G^H) y = f
    gram_herm(s.G, s.s2, N, M, 1.0, s_A, kAS);
    for (uint32_t i = t; i < M; i += kLB) s_A[i * kAS + M] = s_col[i];
    __syncthreads();
    if (t < 64) gj_solve_wave<0>(s_A, M, M + 1, s_col);
    __syncthreads();
    // q = D^{-1} G^H y
    for (uint32_t j = t; j < N; j += kLB) {
        double re = 0.0, im = 0.0;
#pragma unroll 4
        for (uint32_t i = 0; i < M; ++i) {
            const double2 g = s.G[(uint64_t)i * N + j], y = s_A[i * kAS + M];
            re += g.x * y.x + g.y * y.y;
            im += g.x * y.y - g.y * y.x;
        }
        s.q[j] = make_double2(re / s.s2[j], im / s.s2[j]);
    }
    __syncthreads();
    write_camax(a, s, b, N, s_red);
}

__device__ inline uint64_t splitmix64(uint64_t& st) {
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(kLB) void horn_solve(MatArgs a) {
    __shared__ double2 s_A[kMatMaxM * kAS];
    __shared__ double2 s_MM[kMatMaxM * kMS];
    __shared__ double2 s_X[kMatMaxM * kMS];
    __shared__ double2 s_V[kMatMaxM * kMS];
    __shared__ double2 s_col[64], s_z[kMatMaxM];
    __shared__ double s_ev[kMatMaxM], s_raw[kMatMaxM], s_red[4];
    const int t = threadIdx.x;
    const uint32_t b = blockIdx.x;
    const uint32_t base = a.offsets[b], M = a.offsets[b + 1] - base;
    const uint32_t N = a.n;
    const float* amps = a.amps + base;
    const double al2 = a.alpha * a.alpha;
    MatScratch s = mat_scratch(a.scratch + b * a.scratch_stride, N, M);
    // [K | I], K = G G^H + a^2 I
    gram_herm(s.G, nullptr, N, M, al2, s_A, kAS);
    for (uint32_t e = t; e < M * M; e += kLB) {
        const uint32_t i = e / M, k = e % M;
        s_A[i * kAS + M + k] = make_double2(i == k ? 1.0 : 0.0, 0.0);
    }
    __syncthreads();
    if (t < 64) gj_solve_wave<1>(s_A, M, 2 * M, s_col);
    __syncthreads();
    // MM = a^2 P K^{-1} P, X = I
    for (uint32_t e = t; e < M * M; e += kLB) {
        const uint32_t i = e / M, k = e % M;
        const double sc = al2 * (double)amps[i] * (double)amps[k];
        const double2 v = s_A[i * kAS + M + k];
        s_MM[i * kMS + k] = make_double2(v.x * sc, v.y * sc);
        s_X[i * kMS + k] = make_double2(i == k ? 1.0 : 0.0, 0.0);
    }
    __syncthreads();
    if (t < 64) {
        // m * repeat single-row updates (horn.rs:88-119); lane r owns row r
        const int r = t;
        uint64_t rs = a.seed + b;
        const uint64_t count = (uint64_t)M * a.repeat;
        for (uint64_t it = 0; it < count; ++it) {
            const float rnd = (float)(splitmix64(rs) >> 40) * (1.0f / 16777216.0f);
            uint32_t ii = (uint32_t)((float)M * rnd);
            if (ii > M - 1) ii = M - 1;
            const bool act = r < (int)M && r != (int)ii;
            double2 xb = make_double2(0.0, 0.0);
            double gre = 0.0;
            if (act) {
                for (uint32_t k = 0; k < M; ++k) {
                    if (k == ii) continue;
                    const double2 p = cmul(s_X[r * kMS + k], s_MM[k * kMS + ii]);
                    xb.x += p.x;
                    xb.y += p.y;
                }
                const double2 mr = s_MM[r * kMS + ii];
                gre = xb.x * mr.x + xb.y * mr.y;  // Re(conj(xb_r) MM[r][ii])
            }
            gre = wave_sum_f64(gre);
            if (act) {
                double2 v = make_double2(0.0, 0.0);
                if (gre > 0.0) {
                    const double sc = -sqrt(a.lambda / gre);
                    v = make_double2(xb.x * sc, xb.y * sc);
                }
                s_X[r * kMS + ii] = v;
                s_X[ii * kMS + r] = cconj(v);
            }
            wave_fence();
        }
        herm_eig_wave<1>(s_X, s_V, s_ev, s_raw, M);
        // z = K^{-1} P u
        const uint32_t im = imax_of(s_ev, M);
        if (r < (int)M) {
            double2 z = make_double2(0.0, 0.0);
            for (uint32_t k = 0; k < M; ++k) {
                const double2 u = s_X[k * kMS + im];
                const double2 w = make_double2((double)amps[k] * u.x, (double)amps[k] * u.y);
                const double2 p = cmul(s_A[r * kAS + M + k], w);
                z.x += p.x;
                z.y += p.y;
            }
            s_z[r] = z;
        }
    }
    __syncthreads();
    // q = G^H z
    for (uint32_t j = t; j < N; j += kLB) {
        double re = 0.0, im = 0.0;
#pragma unroll 4
        for (uint32_t i = 0; i < M; ++i) {
            const double2 g = s.G[(uint64_t)i * N + j], z = s_z[i];
            re += g.x * z.x + g.y * z.y;
            im += g.x * z.y - g.y * z.x;
        }
        s.q[j] = make_double2(re, im);
    }
    __syncthreads();
    write_camax(a, s, b, N, s_red);
}

int32_t matrix_solve_batch(Context* ctx, const float* sources, uint32_t n, int source_type, const float* foci,
                           const float* amps, const uint32_t* offsets_host, uint32_t batch, int method,
                           uint64_t seed, uint32_t repeat, double lambda, double alpha, float* out_amp,
                           float* out_phase, hipStream_t st) {
    const char* who = method == kMatHorn ? "Horn" : "Long";
    if (batch == 0) return AFH_OK;
    uint32_t maxm = 0;
    for (uint32_t b = 0; b < batch; ++b) {
        const uint32_t m = offsets_host[b + 1] - offsets_host[b];
        if (m == 0 || offsets_host[b + 1] < offsets_host[b]) {
            set_error(std::string(who) + ": an instance has no foci");
            return AFH_ERR_ARG;
        }
        maxm = std::max(maxm, m);
    }
    if (maxm > (uint32_t)kMatMaxM) {
        set_error(std::string(who) + ": more than 32 foci in one instance is not supported");
        return AFH_ERR_UNSUPPORTED;
    }
    const uint64_t stride = mat_scratch_bytes(n, maxm);
    const size_t off_bytes = (batch + 1) * sizeof(uint32_t);
    const size_t o_scr = (off_bytes + 255) & ~(size_t)255;
    if (o_scr + stride * batch > ctx->work.cap && !hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize"))
        return AFH_ERR_HIP;
    if (!ctx->work.reserve(o_scr + stride * batch)) return AFH_ERR_HIP;
    if (!ctx->reserve_staging(off_bytes)) return AFH_ERR_HIP;
    memcpy(ctx->staging.ptr, offsets_host, off_bytes);
    char* w = static_cast<char*>(ctx->work.ptr);
    if (!hip_ok(hipMemcpyAsync(w, ctx->staging.ptr, off_bytes, hipMemcpyHostToDevice, st), "H2D offsets"))
        return AFH_ERR_HIP;
    if (!ctx->upload_done && !hip_ok(hipEventCreateWithFlags(&ctx->upload_done, hipEventDisableTiming), "event"))
        return AFH_ERR_HIP;
    if (!hip_ok(hipEventRecord(ctx->upload_done, st), "hipEventRecord")) return AFH_ERR_HIP;
    MatArgs a;
    memset(&a, 0, sizeof(a));
    a.sources = sources;
    a.source_type = source_type;
    a.n = n;
    a.foci = foci;
    a.amps = amps;
    a.offsets = reinterpret_cast<const uint32_t*>(w);
    a.seed = seed;
    a.repeat = repeat;
    a.lambda = lambda;
    a.alpha = alpha;
    a.scratch = w + o_scr;
    a.scratch_stride = stride;
    a.out_amp = out_amp;
    a.out_phase = out_phase;
    hipLaunchKernelGGL(mat_gen_g, dim3((n + kLB - 1) / kLB, batch), dim3(kLB), 0, st, a);
    if (method == kMatHorn)
        hipLaunchKernelGGL(horn_solve, dim3(batch), dim3(kLB), 0, st, a);
    else
        hipLaunchKernelGGL(long_solve, dim3(batch), dim3(kLB), 0, st, a);
    if (!hip_ok(hipGetLastError(), "matrix solver launch")) return AFH_ERR_HIP;
    return AFH_OK;
}

bool matrix_solve_single(const HoloProblem& p, int method, float* amp_out, float* phase_out) {
    if (!device().ensure()) return false;
    Context& ctx = default_context();
    std::lock_guard<std::mutex> g(ctx.mu);
    if (!ctx.ensure_stream()) return false;
    const uint64_t n = p.n, np = packed_count(n), m = p.m;
    const size_t src_b = np * AFH_SOURCE_STRIDE * sizeof(float), foci_b = m * 3 * sizeof(float);
    const size_t amps_b = m * sizeof(float), out_b = n * sizeof(float);
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t o_foci = align(src_b), o_amps = o_foci + align(foci_b), o_amp = o_amps + align(amps_b);
    const size_t o_ph = o_amp + align(out_b), total = o_ph + align(out_b);
    if (!ctx.io.reserve(total) || !ctx.io_host.reserve(total)) return false;
    char* h = static_cast<char*>(ctx.io_host.ptr);
    char* d = static_cast<char*>(ctx.io.ptr);
    pack_sources(p.sources, p.wavenums, p.attens, n, p.source_type, reinterpret_cast<float*>(h));
    memcpy(h + o_foci, p.foci, foci_b);
    memcpy(h + o_amps, p.amps, amps_b);
    if (!hip_ok(hipMemcpyAsync(d, h, o_amp, hipMemcpyHostToDevice, ctx.stream), "H2D")) return false;
    const uint32_t offs[2] = {0, (uint32_t)m};
    // REPEAT_SDP 100, LAMBDA_SDP 0.8, TIKHONOV_DEFAULT 1e-5 (horn.rs:19-21); seed 0
    if (matrix_solve_batch(&ctx, reinterpret_cast<const float*>(d), (uint32_t)n, p.source_type,
                           reinterpret_cast<const float*>(d + o_foci), reinterpret_cast<const float*>(d + o_amps),
                           offs, 1, method, 0, 100, 0.8, 1e-5, reinterpret_cast<float*>(d + o_amp),
                           reinterpret_cast<float*>(d + o_ph), ctx.stream) != AFH_OK)
        return false;
    if (!hip_ok(hipMemcpyAsync(h + o_amp, d + o_amp, o_ph + out_b - o_amp, hipMemcpyDeviceToHost, ctx.stream), "D2H"))
        return false;
    if (!hip_ok(hipStreamSynchronize(ctx.stream), "hipStreamSynchronize")) return false;
    memcpy(amp_out, h + o_amp, out_b);
    memcpy(phase_out, h + o_ph, out_b);
    return true;
}

}  // namespace afh

using namespace afh;

static int32_t matrix_entry(const char* who, void* ctx, const float* packed_sources_dev, uint64_t n_packed,
                            uint64_t n_sources, int32_t source_type, const float* foci_dev, const float* amps_dev,
                            const uint32_t* offsets_host, uint32_t batch, int method, uint64_t seed, uint32_t repeat,
                            float lambda, float alpha, float* out_amp_dev, float* out_phase_dev, void* stream) {
    clear_error();
    Context* c = static_cast<Context*>(ctx);
    if (!c || !packed_sources_dev || !foci_dev || !amps_dev || !offsets_host || !out_amp_dev || !out_phase_dev ||
        n_packed % AFH_SOURCE_PAD != 0 || n_packed < n_sources || n_sources > 0xFFFFFFFFull ||
        (source_type != kSphere && source_type != kT4010A1) || offsets_host[0] != 0) {
        set_error(std::string(who) + ": bad argument");
        return AFH_ERR_ARG;
    }
    std::lock_guard<std::mutex> g(c->mu);
    return matrix_solve_batch(c, packed_sources_dev, (uint32_t)n_sources, source_type, foci_dev, amps_dev,
                              offsets_host, batch, method, seed, repeat, (double)lambda, (double)alpha, out_amp_dev,
                              out_phase_dev, (hipStream_t)stream);
}

extern "C" int32_t AFH_LongSolve(void* ctx, const float* packed_sources_dev, uint64_t n_packed, uint64_t n_sources,
                                 int32_t source_type, const float* foci_dev, const float* amps_dev,
                                 const uint32_t* offsets_host, uint32_t batch, float* out_amp_dev,
                                 float* out_phase_dev, void* stream) {
    return matrix_entry("AFH_LongSolve", ctx, packed_sources_dev, n_packed, n_sources, source_type, foci_dev,
                        amps_dev, offsets_host, batch, kMatLong, 0, 0, 0.0f, 0.0f, out_amp_dev, out_phase_dev,
                        stream);
}

extern "C" int32_t AFH_HornSolve(void* ctx, const float* packed_sources_dev, uint64_t n_packed, uint64_t n_sources,
                                 int32_t source_type, const float* foci_dev, const float* amps_dev,
                                 const uint32_t* offsets_host, uint32_t batch, uint64_t seed, uint32_t repeat,
                                 float lambda, float tikhonov, float* out_amp_dev, float* out_phase_dev,
                                 void* stream) {
    return matrix_entry("AFH_HornSolve", ctx, packed_sources_dev, n_packed, n_sources, source_type, foci_dev,
                        amps_dev, offsets_host, batch, kMatHorn, seed, repeat, lambda, tikhonov, out_amp_dev,
                        out_phase_dev, stream);
}
