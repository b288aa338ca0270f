// Levenberg-Marquardt hologram solver on the GPU, batched, FP64.
//
// Reference: acoustic-field-optimizer/src/multiple_foci/nls/lm.rs:40-105 with
// make_BhB / make_T / calc_JtJ_Jtf / calc_Fx (nls/macros.rs:17-58):
//   B = [G | -diag(amps)] (M x n, n = N + M), BhB = B^H B,
//   T = e^{-ix}, A = Re(BhB o T T^H), g_k = sum_l Im(BhB o T T^H)_kl,
//   F(x) = Re(t^H BhB t), t = e^{ix}; mu0 = 1e-3 max diag A;
//   h = -(A + mu I)^{-1} g; stop if |h| <= 1e-8 (|x| + 1e-8) or max_k g_k <= 1e-8;
//   rho = (F - F_new) / (h.(mu h - g) / 2); F = F_new (also on rejection);
//   accept if rho > 0: mu *= max(1/3, 1 - (2 rho - 1)^3), nu = 2; else mu *= nu, nu *= 2.
//
// Same quantities, low-rank form. With C = B diag(e^{ix}) (M x n) and
// W = [Re C; Im C] (2M x n):
//   A = Re(C^H C) = W^T W,     s = C 1 = B e^{ix} (the focal field),
//   g = Im(C^H s) = W^T v with v = [Im s; -Re s],     F = |s|^2,
// so the damped step is h = -(mu I + W^T W)^{-1} W^T v = -W^T (mu I + W W^T)^{-1} v
// (push-through identity): an FP64 Cholesky of the 2M x 2M matrix
// K = mu I + W W^T replaces the n x n solve, and an iteration costs O(M^2 n).
// Everything is FP64, G included: the reference's acceptance rule makes the
// iterate path sensitive to perturbations of a few FP32 ulp.
//
// The same kernel runs the other two optimizers of nls/ on this machinery:
//   Gauss-Newton (gauss_newton.rs:35-89): h = -pinv(A, 1e-3) g; with A = W^T W
//     the pseudo-inverse keeps the eigenvalues of K = W W^T above 1e-3 (A and K
//     share their nonzero eigenvalues), so h = -W^T U diag(1/lambda) U^T v over
//     those, U and lambda from a parallel (round-robin) Jacobi
//     eigen-decomposition of K;
//   gradient descent (gradient_descent.rs:34-61): x -= 0.1 g until max g <= 1e-6,
//     g_l = sum_i Im(conj(C_il) s_i) without forming W.
//
// APO (apo.rs:100-158) runs in its own kernel on the same pieces: q0 from the
// push-through form G^H (G G^H + lambda I)^{-1} p (the M x M complex system
// real-ified and solved by the one-wave Cholesky), then the reference's
// quasi-Newton iteration with a dense N x N H.
//
// One workgroup per instance; K, v, y in LDS; G, W, x, g, h, e^{ix} in a
// per-instance global scratch. Every reduction has a fixed order.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "afh_dense.hpp"
#include "afh_holo.hpp"
#include "afh_internal.hpp"

namespace afh {

constexpr int kLmMaxM = 32;
constexpr int kLmMaxD = 2 * kLmMaxM;
constexpr int kApoLineSearch = 100;  // LINE_SEARCH_MAX (apo.rs:19)

struct LmArgs {
    const float* sources;  // packed (AFH_SOURCE_STRIDE floats each)
    int source_type;
    uint32_t n;  // sources N
    const float* foci;
    const float* amps;
    const uint32_t* offsets;
    uint32_t k_max;
    int algo;       // kNlsLM, kNlsGN, kNlsGD
    double lambda;  // APO's regularisation
    float* out_amp; // APO: batch x N
    char* scratch;  // per instance, scratch_stride bytes
    uint64_t scratch_stride;
    float* out_phase;     // batch x N
    uint32_t* out_iters;  // batch (may be null)
};

// Scratch layout of one instance (n = N + M): G (M x N, double2) | W (2M x n) |
// x | xn | g | h (n each) | E | En (n, double2).
struct LmScratch {
    double2* G;
    double* W;
    double *x, *xn, *g, *h;
    double2 *E, *En;
};

__host__ __device__ inline uint64_t lm_scratch_bytes(uint64_t N, uint64_t M) {
    const uint64_t n = N + M;
    auto al = [](uint64_t b) { return (b + 255) & ~(uint64_t)255; };
    return al(M * N * 16) + al(2 * M * n * 8) + 4 * al(n * 8) + 2 * al(n * 16);
}

__device__ inline LmScratch lm_scratch(char* base, uint64_t N, uint64_t M) {
    const uint64_t n = N + M;
    auto al = [](uint64_t b) { return (b + 255) & ~(uint64_t)255; };
    LmScratch s;
    char* p = base;
    s.G = reinterpret_cast<double2*>(p);
    p += al(M * N * 16);
    s.W = reinterpret_cast<double*>(p);
    p += al(2 * M * n * 8);
    s.x = reinterpret_cast<double*>(p);
    p += al(n * 8);
    s.xn = reinterpret_cast<double*>(p);
    p += al(n * 8);
    s.g = reinterpret_cast<double*>(p);
    p += al(n * 8);
    s.h = reinterpret_cast<double*>(p);
    p += al(n * 8);
    s.E = reinterpret_cast<double2*>(p);
    p += al(n * 16);
    s.En = reinterpret_cast<double2*>(p);
    return s;
}

__global__ __launch_bounds__(kLB) void lm_gen_g(LmArgs a) {
    const uint32_t b = blockIdx.y;
    const uint32_t j = blockIdx.x * kLB + threadIdx.x;
    const uint32_t base = a.offsets[b], M = a.offsets[b + 1] - base;
    if (j >= a.n) return;
    LmScratch s = lm_scratch(a.scratch + b * a.scratch_stride, a.n, M);
    const float* src = a.sources + AFH_SOURCE_STRIDE * (uint64_t)j;
    for (uint32_t i = 0; i < M; ++i)
        s.G[(uint64_t)i * a.n + j] = term_f64(src, a.source_type, a.foci[3 * (uint64_t)(base + i)],
                                              a.foci[3 * (uint64_t)(base + i) + 1], a.foci[3 * (uint64_t)(base + i) + 2]);
}

// s_i = sum_l B_il E_l (E = e^{ix}); a wave per row, lanes over columns.
__device__ inline void field_at_foci(const LmScratch& s, const double2* E, uint32_t N, uint32_t M, const float* amps,
                                     double2* s_out) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    for (uint32_t i = w; i < M; i += 4) {
        double re = 0.0, im = 0.0;
        for (uint32_t j = l; j < N; j += 64) {
            const double2 gv = s.G[(uint64_t)i * N + j], e = E[j];
            re += gv.x * e.x - gv.y * e.y;
            im += gv.x * e.y + gv.y * e.x;
        }
        re = wave_sum_f64(re);
        im = wave_sum_f64(im);
        if (l == 0) {
            // the -diag(amps) block of B: column N + i
            const double2 e = E[N + i];
            const double am = (double)amps[i];
            s_out[i] = make_double2(re - am * e.x, im - am * e.y);
        }
    }
}

typedef double doublex4 __attribute__((ext_vector_type(4)));

// Gram matrix W W^T of the D x n matrix W (row-major, D <= 64), lower
// triangle (p >= q) into sK (row stride kLmMaxD) and the diagonal into sgd,
// with FP64 MFMA (v_mfma_f64_16x16x4f64, A 16 x 4, B 4 x 16): the
// lower-triangle 16 x 16 tiles (I, J), I >= J, of the D x D
// result are dealt to the 4 waves in turn; a tile sums W_I W_J^T over the
// n / 4 column blocks, lane l supplying W[16 I + l % 16][4 kb + l / 16] as A
// and W[16 J + l % 16][4 kb + l / 16] as B and receiving the tile's rows
// l / 16 + 4 e, e = 0..3, column l % 16 (the FP64 MFMA interleaves its
// output rows, unlike the FP32 16 x 16 MFMAs' 4 (l / 16) + e). Even and odd column blocks go to two
// accumulators (summed at the end), so consecutive MFMAs do not wait on each
// other. Rows >= D and columns >= n enter as zeros.
__device__ void gram_mfma(const double* __restrict__ W, uint32_t n, uint32_t D, double* sK, double* sgd) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int T = (int)(D + 15) / 16, ntiles = T * (T + 1) / 2;
    const uint32_t kq = l >> 4;
    for (int tile = w; tile < ntiles; tile += kLB / 64) {
        int I = 0, J = tile;
        while (J > I) {
            J -= I + 1;
            ++I;
        }
        const uint32_t ra = 16 * I + (l & 15), rb = 16 * J + (l & 15);
        const bool va = ra < D, vb = rb < D;
        const double* __restrict__ wa = W + (uint64_t)(va ? ra : 0) * n;
        const double* __restrict__ wb = W + (uint64_t)(vb ? rb : 0) * n;
        doublex4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = acc0;
        uint32_t k0 = 0;
#pragma unroll 2
        for (; k0 + 8 <= n; k0 += 8) {
            const uint32_t k = k0 + kq;
            const double a0 = va ? wa[k] : 0.0, b0 = vb ? wb[k] : 0.0;
            const double a1 = va ? wa[k + 4] : 0.0, b1 = vb ? wb[k + 4] : 0.0;
            acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc1, 0, 0, 0);
        }
        for (; k0 < n; k0 += 4) {
            const uint32_t k = k0 + kq;
            const double a0 = (va && k < n) ? wa[k] : 0.0, b0 = (vb && k < n) ? wb[k] : 0.0;
            acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc0, 0, 0, 0);
        }
        const doublex4 acc = acc0 + acc1;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const uint32_t p = 16 * I + kq + 4 * e, q = 16 * J + (l & 15);
            if (p < D && q <= p) {
                sK[p * kLmMaxD + q] = acc[e];
                if (p == q) sgd[p] = acc[e];
            }
        }
    }
}

// Lane `lane`'s value of v (a wave-uniform result, read with v_readlane).
__device__ __forceinline__ double readlane_f64(double v, int lane) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// One wave (lanes 0..63) factors K = L L^T (D x D, lower triangle of sK with
// row stride kLmMaxD) and solves L z = v, L^T y = z. Lane i keeps row i of
// the matrix in registers (DP >= D entries, rows i >= D are identity rows, so
// the padded system is block-diagonal); the loops are unrolled, so the values
// one lane needs from another (a pivot, a column entry, a solved unknown) are
// read with v_readlane, without LDS round trips or fences. Right-looking
// elimination, forward substitution in increasing j, backward in decreasing
// j (the backward step reads row j of L, written back to sK), every update an
// fma, each division by L_ii a multiplication by its reciprocal. Writes L back
// to sK (the lower triangle only: the upper triangle holds the Gram copy) and
// y to sy[0..D). Returns false when a pivot is not positive (or NaN). Site
// gives each calling kernel its own instance, so that the LDS arrays it passes
// stay compile-time addresses inside the function (with two callers they
// become runtime pointers and every unrolled access pays for address
// arithmetic).
template <int DP, int Site>
__device__ __attribute__((noinline)) bool chol_solve_wave(double* sK, const double* sv, double* sy, double* col, uint32_t D) {
    (void)col;
    const int i = threadIdx.x;  // < 64, the whole wave active
    const bool real = i < (int)D;
    double r[DP];
#pragma unroll
    for (int j = 0; j < DP; ++j) r[j] = real ? (j <= i ? sK[i * kLmMaxD + j] : 0.0) : (j == i ? 1.0 : 0.0);
    bool ok = true;
#pragma unroll
    for (int k = 0; k < DP; ++k) {
        const double dkk = readlane_f64(r[k], k);
        ok = ok && dkk > 0.0;
        const double lkk = sqrt(dkk);
        if (i > k) r[k] = r[k] / lkk;
        if (i == k) r[k] = lkk;
#pragma unroll
        for (int j = k + 1; j < DP; ++j) {
            const double ljk = readlane_f64(r[k], j);
            if (i >= j) r[j] = fma(-r[k], ljk, r[j]);
        }
    }
    // L back to LDS for the backward substitution (row j read across lanes)
    if (real)
#pragma unroll
        for (int j = 0; j < DP; ++j)
            if (j <= i) sK[i * kLmMaxD + j] = r[j];
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // 1 / L_ii, once per lane
    double diag = 1.0;
#pragma unroll
    for (int j = 0; j < DP; ++j)
        if (i == j) diag = r[j];
    const double dinv = 1.0 / diag;
    // L z = v
    double acc = real ? sv[i] : 0.0, z = 0.0;
#pragma unroll
    for (int j = 0; j < DP; ++j) {
        const double zj = readlane_f64(acc * dinv, j);
        if (i == j) z = zj;
        if (i > j) acc = fma(-r[j], zj, acc);
    }
    // L^T y = z
    acc = z;
    double y = 0.0;
#pragma unroll
    for (int j = DP - 1; j >= 0; --j) {
        const double yj = readlane_f64(acc * dinv, j);
        if (i == j) y = yj;
        // L[j][i] for i < j: row j of L, read by lane i
        if (i < j && j < (int)D) acc = fma(-sK[j * kLmMaxD + i], yj, acc);
    }
    if (real) sy[i] = y;
    return ok;
}

// The same factorization and solves for 32 < D <= 64 with the rows kept in
// LDS (work, row stride kLmMaxD + 1 doubles, so that a column is spread over
// the banks) instead of registers; identical arithmetic in the same order.
template <int Site>
__device__ __attribute__((noinline)) bool chol_solve_wave_lds(double* sK, const double* sv, double* sy, double* col, double* work,
                                    uint32_t D) {
    constexpr int DP = kLmMaxD, WS = kLmMaxD + 1;
    const int i = threadIdx.x;  // < 64
    const bool real = i < (int)D;
    for (int j = 0; j < DP; ++j) work[i * WS + j] = real ? (j <= i ? sK[i * kLmMaxD + j] : 0.0) : (j == i ? 1.0 : 0.0);
    bool ok = true;
    for (int k = 0; k < DP; ++k) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const double dkk = work[k * WS + k];
        ok = ok && dkk > 0.0;
        const double lkk = sqrt(dkk);
        double rk = work[i * WS + k];
        if (i > k) rk = rk / lkk;
        if (i == k) rk = lkk;
        work[i * WS + k] = rk;
        col[i] = rk;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        for (int j = k + 1; j < DP; ++j) {
            const double ljk = col[j];
            if (i >= j) work[i * WS + j] = fma(-rk, ljk, work[i * WS + j]);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (real)
        for (int j = 0; j <= i; ++j) sK[i * kLmMaxD + j] = work[i * WS + j];
    const double dinv = 1.0 / work[i * WS + i];  // 1 / L_ii (1 on the identity rows)
    double acc = real ? sv[i] : 0.0, z = 0.0;
    for (int j = 0; j < DP; ++j) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        if (i == j) col[j] = acc * dinv;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const double zj = col[j];
        if (i == j) z = zj;
        if (i > j) acc = fma(-work[i * WS + j], zj, acc);
    }
    acc = z;
    double y = 0.0;
    for (int j = DP - 1; j >= 0; --j) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        if (i == j) col[j] = acc * dinv;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const double yj = col[j];
        if (i == j) y = yj;
        if (i < j && j < (int)D) acc = fma(-work[j * WS + i], yj, acc);
    }
    if (real) sy[i] = y;
    return ok;
}

// y = pinv(K, eps) v for the Gram matrix K (lower triangle of sK with row
// stride kLmMaxD, diagonal in sgd), D <= DP, by the whole workgroup (kLB
// threads): parallel cyclic Jacobi on a copy of K padded to DP x DP with
// zeros (A, row stride kLmMaxD + 1), the eigenvectors accumulated in V. The
// padding rows and columns stay exactly zero (every rotation that pairs a
// padding index has a_pq = 0 and is the identity), so they add eigenvalues 0,
// which the threshold drops. A sweep is DP - 1 rounds of the round-robin
// (circle) schedule: in round t index DP - 1 meets t and (t + r) mod (DP - 1)
// meets (t - r) mod (DP - 1), r = 1 .. DP/2 - 1, so each round rotates DP/2
// disjoint pairs at once. Threads r < DP/2 compute the round's rotations from
// its A (into cs and pq); then each 2 x 2 block of A at pairs (r, s) becomes
// R_r^T A_rs R_s (the column rotation of pair s first, then the row rotation
// of pair r), independently of every other block, and V's rows take the
// column rotations: thread T holds pair s = T mod DP/2 throughout, the blocks
// and V rows of its group. Sweeps run until the off-diagonal mass is below
// 1e-30 of the diagonal's; eigenvalues |lambda| <= eps are dropped (as the
// reference's SVD pseudo-inverse drops singular values <= eps).
template <int DP>
__device__ __attribute__((noinline)) void jacobi_pinv_block(const double* sK, const double* sgd, const double* sv,
                                                            double* sy, double* A, double* V, double* col,
                                                            double2* cs, int2* pq, double* red, uint32_t D,
                                                            double eps) {
    constexpr int WS = kLmMaxD + 1, H = DP / 2, NM = DP - 1;
    constexpr int RSTEP = kLB / H;              // block rows / V rows a thread steps by
    constexpr int NB = H * H / kLB;             // blocks per thread; 2 NB V row pairs
    static_assert(kLB == 256 && (DP == 32 || DP == 64), "4 waves; 1 or 4 blocks per thread");
    const int T = threadIdx.x;
    const int s = T % H, r0 = T / H;
    for (int e = T; e < DP * DP; e += kLB) {
        const int k = e / DP, j = e % DP;
        const bool in = k < (int)D && j < (int)D;
        A[k * WS + j] = !in ? 0.0 : (j == k ? sgd[k] : (j < k ? sK[k * kLmMaxD + j] : sK[j * kLmMaxD + k]));
        V[k * WS + j] = (j == k) ? 1.0 : 0.0;
    }
    lds_barrier();
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, dg = 0.0;
        for (int e = T; e < DP * DP; e += kLB) {
            const int k = e / DP, j = e % DP;
            const double v = A[k * WS + j];
            if (j == k)
                dg += v * v;
            else
                off += v * v;
        }
        off = block_sum(off, red);
        dg = block_sum(dg, red);
        if (off <= dg * 1e-30 || off == 0.0) break;
        for (int t = 0; t < NM; ++t) {
            if (T < H) {
                const int p = T == 0 ? NM : (t + T) % NM;
                const int q = T == 0 ? t : (t - T + NM) % NM;
                const double apq = A[p * WS + q];
                double c = 1.0, sn = 0.0;
                if (apq != 0.0) {
                    // t = sgn(theta) / (|theta| + sqrt(theta^2 + 1)) with theta = delta / tau,
                    // delta = a_qq - a_pp, tau = 2 a_pq, as sgn(theta) |tau| / (|delta| +
                    // hypot(delta, tau)); sgn(+-0) = +1. The angle is taken in FP32 from
                    // delta and tau scaled by a power of two (exact), which leaves a_pq at
                    // ~1e-7 of its value instead of 0 and does not change the sweep count;
                    // c and s are then renormalised in FP64 (c^2 + s^2 - 1 = e, times
                    // 1 - e/2 + 3e^2/8), so every rotation, and V, stays orthogonal to FP64
                    // rounding
                    const double delta = A[q * WS + q] - A[p * WS + p], tau = 2.0 * apq;
                    const int ex = __builtin_amdgcn_frexp_exp(fmax(fabs(delta), fabs(tau)));
                    const float d32 = (float)__builtin_amdgcn_ldexp(delta, -ex);
                    const float t32 = (float)__builtin_amdgcn_ldexp(tau, -ex);
                    const float sg = (d32 == 0.f || (d32 > 0.f) == (t32 > 0.f)) ? 1.f : -1.f;
                    const float tt = sg * fabsf(t32) *
                                     __builtin_amdgcn_rcpf(fabsf(d32) + __builtin_amdgcn_sqrtf(fmaf(d32, d32, t32 * t32)));
                    const float c32 = __builtin_amdgcn_rsqf(fmaf(tt, tt, 1.f));
                    const double c0 = c32, s0 = (double)(tt * c32);
                    const double e2 = fma(c0, c0, fma(s0, s0, -1.0));
                    const double f = fma(e2, fma(0.375, e2, -0.5), 1.0);
                    c = c0 * f;
                    sn = s0 * f;
                }
                cs[T] = make_double2(c, sn);
                pq[T] = make_int2(p, q);
            }
            lds_barrier();
            // every load of the thread's group is issued before its first
            // store (the elements are disjoint, which the compiler cannot
            // see), so the group pays one LDS round trip
            const int2 es = pq[s];
            const double2 rs = cs[s];
            int2 er[NB];
            double2 rr[NB];
            double va[NB], vb[NB], vc[NB], vd[NB], vp[2 * NB], vq[2 * NB];
#pragma unroll
            for (int u = 0; u < NB; ++u) {
                const int r = r0 + RSTEP * u;
                er[u] = pq[r];
                rr[u] = cs[r];
            }
#pragma unroll
            for (int u = 0; u < NB; ++u) {
                va[u] = A[er[u].x * WS + es.x];
                vb[u] = A[er[u].x * WS + es.y];
                vc[u] = A[er[u].y * WS + es.x];
                vd[u] = A[er[u].y * WS + es.y];
            }
#pragma unroll
            for (int u = 0; u < 2 * NB; ++u) {
                const int row = r0 + RSTEP * u;
                vp[u] = V[row * WS + es.x];
                vq[u] = V[row * WS + es.y];
            }
#pragma unroll
            for (int u = 0; u < NB; ++u) {
                const double a1 = rs.x * va[u] - rs.y * vb[u], b1 = rs.y * va[u] + rs.x * vb[u];
                const double c1 = rs.x * vc[u] - rs.y * vd[u], d1 = rs.y * vc[u] + rs.x * vd[u];
                A[er[u].x * WS + es.x] = rr[u].x * a1 - rr[u].y * c1;
                A[er[u].y * WS + es.x] = rr[u].y * a1 + rr[u].x * c1;
                A[er[u].x * WS + es.y] = rr[u].x * b1 - rr[u].y * d1;
                A[er[u].y * WS + es.y] = rr[u].y * b1 + rr[u].x * d1;
            }
#pragma unroll
            for (int u = 0; u < 2 * NB; ++u) {
                const int row = r0 + RSTEP * u;
                V[row * WS + es.x] = rs.x * vp[u] - rs.y * vq[u];
                V[row * WS + es.y] = rs.y * vp[u] + rs.x * vq[u];
            }
            lds_barrier();
        }
    }
    // w_k = (V_:k . v) / lambda_k over the kept eigenvalues, y = V w
    if (T < DP) {
        double d = 0.0;
#pragma unroll 8
        for (int j = 0; j < DP; ++j) d += V[j * WS + T] * (j < (int)D ? sv[j] : 0.0);
        const double lam = A[T * WS + T];
        col[T] = fabs(lam) > eps ? d / lam : 0.0;
    }
    lds_barrier();
    if (T < (int)D) {
        double y = 0.0;
#pragma unroll 8
        for (int j = 0; j < DP; ++j) y += V[T * WS + j] * col[j];
        sy[T] = y;
    }
}

__global__ __launch_bounds__(kLB) void lm_solve(LmArgs a) {
    __shared__ double s_K[kLmMaxD * kLmMaxD];
    __shared__ double s_v[kLmMaxD];
    __shared__ double s_gd[kLmMaxD];  // Gram diagonal (the factor overwrites s_K's)
    __shared__ double s_y[kLmMaxD];
    __shared__ double2 s_s[kLmMaxM];
    __shared__ double s_red[4];
    __shared__ double s_col[kLmMaxD];
    __shared__ double s_work[kLmMaxD * (kLmMaxD + 1)];
    __shared__ double s_vec[kLmMaxD * (kLmMaxD + 1)];
    __shared__ double2 s_rot[kLmMaxD / 2];
    __shared__ int2 s_pair[kLmMaxD / 2];
    __shared__ int s_ok;
    const int t = threadIdx.x;
    const uint32_t b = blockIdx.x;
    const uint32_t base = a.offsets[b], M = a.offsets[b + 1] - base;
    const uint32_t N = a.n, n = N + M, D = 2 * M;
    const float* amps = a.amps + base;
    LmScratch s = lm_scratch(a.scratch + b * a.scratch_stride, N, M);
    const double EPS_1 = 1e-8, EPS_2 = 1e-8, TAU = 1e-3;

    // W = [Re C; Im C] with C = B diag(E), v = [Im s; -Re s], g = W^T v,
    // the Gram matrix W W^T into s_K (lower triangle and diagonal), F = |s|^2.
    // Returns (F, max_l g_l, max_l diag(A)_l).
    auto linearize = [&](double& F, double& gmax, double& amax) {
        // G and W are separate regions of the scratch: restrict lets the next
        // G loads issue ahead of this row's W stores
        const double2* __restrict__ Gp = s.G;
        double* __restrict__ Wp = s.W;
        for (uint32_t l = t; l < n; l += kLB) {
            const double2 e = s.E[l];
            double colsq = 0.0;
#pragma unroll 4
            for (uint32_t i = 0; i < M; ++i) {
                double2 bv;
                if (l < N) {
                    bv = Gp[(uint64_t)i * N + l];
                } else {
                    bv = make_double2(l - N == i ? -(double)amps[i] : 0.0, 0.0);
                }
                const double cr = bv.x * e.x - bv.y * e.y, ci = bv.x * e.y + bv.y * e.x;
                Wp[(uint64_t)i * n + l] = cr;
                Wp[(uint64_t)(M + i) * n + l] = ci;
                colsq += cr * cr + ci * ci;
            }
            s.h[l] = colsq;  // diag(A)_l, reused as scratch here
        }
        __syncthreads();
        field_at_foci(s, s.E, N, M, amps, s_s);
        __syncthreads();
        if (t < (int)M) {
            s_v[t] = s_s[t].y;
            s_v[M + t] = -s_s[t].x;
        }
        double fl = 0.0;
        for (uint32_t i = 0; i < M; ++i) fl += s_s[i].x * s_s[i].x + s_s[i].y * s_s[i].y;
        F = fl;
        __syncthreads();
        double gm = -INFINITY, am = -INFINITY;
        for (uint32_t l = t; l < n; l += kLB) {
            double gl = 0.0;
#pragma unroll 8
            for (uint32_t q = 0; q < D; ++q) gl += s.W[(uint64_t)q * n + l] * s_v[q];
            s.g[l] = gl;
            gm = fmax(gm, gl);
            am = fmax(am, s.h[l]);
        }
        gmax = block_max(gm, s_red);
        amax = block_max(am, s_red);
        // Gram matrix W W^T (p >= q) into s_K and its diagonal into s_gd
        gram_mfma(s.W, n, D, s_K, s_gd);
        __syncthreads();
    };

    for (uint32_t l = t; l < n; l += kLB) {
        s.x[l] = 0.0;
        s.E[l] = make_double2(1.0, 0.0);
    }
    __syncthreads();
    if (a.algo == kNlsGD) {
        // gradient_descent.rs:49-56
        uint32_t it = 0;
        for (; it < a.k_max; ++it) {
            field_at_foci(s, s.E, N, M, amps, s_s);
            __syncthreads();
            double gm = -INFINITY;
            for (uint32_t l = t; l < n; l += kLB) {
                const double2 e = s.E[l];
                double gl = 0.0;
#pragma unroll 4
                for (uint32_t i = 0; i < M; ++i) {
                    double2 bv;
                    if (l < N)
                        bv = s.G[(uint64_t)i * N + l];
                    else
                        bv = make_double2(l - N == i ? -(double)amps[i] : 0.0, 0.0);
                    const double cr = bv.x * e.x - bv.y * e.y, ci = bv.x * e.y + bv.y * e.x;
                    gl += cr * s_s[i].y - ci * s_s[i].x;  // Im(conj(C_il) s_i)
                }
                s.g[l] = gl;
                gm = fmax(gm, gl);
            }
            const double gmax = block_max(gm, s_red);
            if (gmax <= 1e-6) break;
            for (uint32_t l = t; l < n; l += kLB) {
                const double xl = s.x[l] - 0.1 * s.g[l];
                s.x[l] = xl;
                double sn, cs;
                sincos(xl, &sn, &cs);
                s.E[l] = make_double2(cs, sn);
            }
            __syncthreads();
        }
        for (uint32_t j = t; j < N; j += kLB) a.out_phase[(uint64_t)b * N + j] = (float)s.x[j];
        if (t == 0 && a.out_iters) a.out_iters[b] = it;
        return;
    }
    double F, gmax, amax;
    linearize(F, gmax, amax);
    // Gauss-Newton, gauss_newton.rs:49-84 (LM runs in lm_wide)
    {
        bool found = gmax <= EPS_1;
        uint32_t it = 0;
        for (; it < a.k_max; ++it) {
            if (found) break;
            if (D <= 32)
                jacobi_pinv_block<32>(s_K, s_gd, s_v, s_y, s_work, s_vec, s_col, s_rot, s_pair, s_red, D, 1e-3);
            else
                jacobi_pinv_block<64>(s_K, s_gd, s_v, s_y, s_work, s_vec, s_col, s_rot, s_pair, s_red, D, 1e-3);
            __syncthreads();
            double hh = 0.0, xx = 0.0;
            for (uint32_t l = t; l < n; l += kLB) {
                double acc = 0.0;
#pragma unroll 8
                for (uint32_t q = 0; q < D; ++q) acc += s.W[(uint64_t)q * n + l] * s_y[q];
                const double hl = -acc;
                const double xl = s.x[l];
                hh += hl * hl;
                xx += xl * xl;
                s.xn[l] = xl + hl;
            }
            const double hnorm = sqrt(block_sum(hh, s_red));
            const double xnorm = sqrt(block_sum(xx, s_red));
            if (hnorm <= EPS_2 * (xnorm + EPS_2)) {
                found = true;
                continue;
            }
            for (uint32_t l = t; l < n; l += kLB) {
                const double xl = s.xn[l];
                s.x[l] = xl;
                double sn, cs;
                sincos(xl, &sn, &cs);
                s.E[l] = make_double2(cs, sn);
            }
            __syncthreads();
            linearize(F, gmax, amax);
            found = gmax <= EPS_1;
        }
        for (uint32_t j = t; j < N; j += kLB) a.out_phase[(uint64_t)b * N + j] = (float)s.x[j];
        if (t == 0 && a.out_iters) a.out_iters[b] = it;
        return;
    }
}

// ---------------------------------------------------------------- LM, wide
// The LM iteration (lm.rs:40-105) on a workgroup of kLW threads (8 waves)
// per instance. The chain of one iteration is latency-bound (a few hundred
// FP64 flops per thread between barriers), so every step is spread over all
// waves: the columns l of x, g, h and e^{ix} over the threads (one or two
// per thread for n <= 2 kLW, their loads in flight together); the focus rows
// of the focal field over the waves; the Gram tiles over the waves. W = [Re C;
// Im C] (C = B diag(e^{ix})) is never stored: g, h and the Gram matrix form
// it from G and e^{ix}, so that an instance's working set is G and a few
// n-vectors (it stays in L2 for the 996-source array, where G and W did not).
// Besides:
//   - g = W^T v is formed in the same pass over G as diag(A) (for mu0, first
//     linearization only: |e^{ix}| = 1, so it never changes);
//   - v comes from the focal field of the accepted point, which the
//     acceptance test has just computed;
//   - x / x_new and e^{ix} / e^{ix_new} swap roles on acceptance;
//   - the Gram matrix goes to the lower triangle of s_K and its diagonal to
//     s_gd, where it survives rejected steps; the L D L^T solve reads
//     mu I + Gram from there and writes L to its own array.
// Every reduction has a fixed order; the iteration count, stopping rules and
// mu/nu updates are the reference's.
constexpr int kLW = 512;
constexpr int kLWaves = kLW / 64;
constexpr int kLmPart = kLmMaxD * (kLmMaxD + 1);  // doubles of the second Gram staging buffer

// Fixed-order workgroup sums (maxima) of NV values: lanes by xor steps, then
// the wave results by xor steps within each 16-lane row (lane q of a row
// reads wave q's value, lanes past the last wave the identity). Every thread
// gets the same bits.
template <int NV, bool MAX>
__device__ __forceinline__ void lw_reduce(double* v, double* red) {
    static_assert(kLWaves <= 16, "one wave result per lane of a DPP row");
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = MAX ? wave_max_f64(v[k]) : wave_sum_f64(v[k]);
    const int w = threadIdx.x >> 6, q = threadIdx.x & 15;
    lds_barrier();
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) red[k * kLWaves + w] = v[k];
    lds_barrier();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double u = q < kLWaves ? red[k * kLWaves + q] : (MAX ? -INFINITY : 0.0);
        u = MAX ? fmax(u, xor_lane_f64<8>(u)) : u + xor_lane_f64<8>(u);
        u = MAX ? fmax(u, xor_lane_f64<4>(u)) : u + xor_lane_f64<4>(u);
        u = MAX ? fmax(u, xor_lane_f64<2>(u)) : u + xor_lane_f64<2>(u);
        u = MAX ? fmax(u, xor_lane_f64<1>(u)) : u + xor_lane_f64<1>(u);
        v[k] = u;
    }
}

// s_i = sum_l B_il E_l: the rows i = w, w + kLWaves, ... of wave w together
// (one load of E_j for all of them), lanes over the columns, each row summed
// in column order and reduced over the lanes.
__device__ inline void lw_field(const double2* __restrict__ G, const double2* __restrict__ E, uint32_t N, uint32_t M,
                                const float* amps, double2* s_out) {
    constexpr int R = (kLmMaxM + kLWaves - 1) / kLWaves;
    const uint32_t w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const uint32_t nr = M > w ? (M - w + kLWaves - 1) / kLWaves : 0;  // rows of this wave
    double re[R], im[R];
#pragma unroll
    for (int r = 0; r < R; ++r) re[r] = im[r] = 0.0;
#pragma unroll 2
    for (uint32_t j = l; j < N; j += 64) {
        const double2 e = E[j];
#pragma unroll
        for (int r = 0; r < R; ++r)
            if ((uint32_t)r < nr) {
                const double2 gv = G[(uint64_t)(w + r * kLWaves) * N + j];
                re[r] += gv.x * e.x - gv.y * e.y;
                im[r] += gv.x * e.y + gv.y * e.x;
            }
    }
#pragma unroll
    for (int r = 0; r < R; ++r)
        if ((uint32_t)r < nr) {
            const uint32_t i = w + r * kLWaves;
            const double sr = wave_sum_f64(re[r]), si = wave_sum_f64(im[r]);
            if (l == 0) {
                const double2 e = E[N + i];  // the -diag(amps) block of B: column N + i
                const double am = (double)amps[i];
                s_out[i] = make_double2(sr - am * e.x, si - am * e.y);
            }
        }
}

// Gram matrix W W^T for W = [Re C; Im C], C = B diag(E) (D = 2M rows, n =
// N + M columns), with FP64 MFMA (v_mfma_f64_16x16x4f64 as in gram_mfma),
// without W in memory. The N columns of G go through LDS in chunks of cw
// (128 for M <= 16, 64 above, 4 elements of C per thread): all threads form the chunk's W (thread per element of C, G read
// coalesced along the columns) into one of two buffers while the waves run
// the MFMAs of the previous chunk on the other, one barrier per chunk; the
// next chunk's G and E are loaded before those MFMAs. Wave w accumulates
// unit u = w (and u = w + kLWaves while u < ntiles): lower-triangle tile
// u % ntiles over the 4-column blocks kb = u / ntiles (mod S) of every
// chunk, S = the number of units of that tile. The units' tiles are summed
// in unit order, then the last M columns (-diag(amps), one nonzero C_i,N+i
// each) are added to the entries they touch: (i, i), (M + i, M + i) and
// (M + i, i). Writes the off-diagonal entries to the lower triangle
// (sK[p * kLmMaxD + q], q < p) and the diagonal to sgd. buf0 and buf1 hold
// kLmMaxD (kLmMaxD + 1) doubles each (D rows of cw + 1).
constexpr int kGramPer = 4;  // elements of C per thread and chunk

__device__ __attribute__((noinline)) void lw_gram(const double2* __restrict__ G, const double2* __restrict__ E, uint32_t N, uint32_t M,
                        const float* amps, double* sK, double* sgd, double* buf0, double* buf1) {
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    const uint32_t D = 2 * M;
    const int T = (int)(D + 15) / 16, ntiles = T * (T + 1) / 2;
    const int units = ntiles > kLWaves ? ntiles : kLWaves;
    const uint32_t cw = M <= 16 ? 128 : 64;  // M x cw <= kGramPer x kLW
    const uint32_t ws = cw + 1;              // LDS row stride (rows spread over the banks)
    const uint32_t nchunks = (N + cw - 1) / cw;
    const uint32_t kq = l >> 4;
    // chunk staging: element e = t + kLW j of the M x kGramCW block of C
    double2 pg[kGramPer], pe[kGramPer];
    auto fetch = [&](uint32_t c) {
#pragma unroll
        for (int j = 0; j < kGramPer; ++j) {
            const uint32_t e = t + kLW * j, i = e / cw, k = c * cw + e % cw;
            const bool in = i < M && k < N;
            pg[j] = in ? G[(uint64_t)i * N + k] : make_double2(0.0, 0.0);
            pe[j] = in ? E[k] : make_double2(0.0, 0.0);
        }
    };
    auto store = [&](double* buf) {
#pragma unroll
        for (int j = 0; j < kGramPer; ++j) {
            const uint32_t e = t + kLW * j, i = e / cw, k = e % cw;
            if (i < M) {
                buf[i * ws + k] = pg[j].x * pe[j].x - pg[j].y * pe[j].y;
                buf[(M + i) * ws + k] = pg[j].x * pe[j].y + pg[j].y * pe[j].x;
            }
        }
    };
    // the (at most two) units of this wave
    int ut[2], uk[2], us[2], uI[2], uJ[2];
    bool uv[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int u = w + q * kLWaves;
        uv[q] = u < units;
        ut[q] = u % ntiles;
        uk[q] = u / ntiles;
        us[q] = (units - 1 - ut[q]) / ntiles + 1;
        int I = 0, J = ut[q];
        while (J > I) {
            J -= I + 1;
            ++I;
        }
        uI[q] = I;
        uJ[q] = J;
    }
    doublex4 acc[2] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
    fetch(0);
    store(buf0);
    lds_barrier();
    for (uint32_t c = 0; c < nchunks; ++c) {
        double* cur = (c & 1) ? buf1 : buf0;
        double* nxt = (c & 1) ? buf0 : buf1;
        if (c + 1 < nchunks) fetch(c + 1);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            if (!uv[q]) continue;
            const uint32_t ra = 16 * uI[q] + (l & 15), rb = 16 * uJ[q] + (l & 15);
            const bool va = ra < D, vb = rb < D;
            const double* pa = cur + (va ? ra : 0) * ws + kq;
            const double* pb = cur + (vb ? rb : 0) * ws + kq;
            for (int kb = uk[q]; kb < (int)cw / 4; kb += us[q]) {
                const double a0 = va ? pa[4 * kb] : 0.0, b0 = vb ? pb[4 * kb] : 0.0;
                acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[q], 0, 0, 0);
            }
        }
        if (c + 1 < nchunks) store(nxt);
        lds_barrier();
    }
    // partial tiles into buf0 (free after the last barrier), unit u at u * 256
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        if (!uv[q]) continue;
        const int u = w + q * kLWaves;
#pragma unroll
        for (int e = 0; e < 4; ++e) buf0[u * 256 + (kq + 4 * e) * 16 + (l & 15)] = acc[q][e];
    }
    lds_barrier();
    for (int e = t; e < ntiles * 256; e += kLW) {
        const int tile = e >> 8, pl = (e >> 4) & 15, ql = e & 15;
        int I = 0, J = tile;
        while (J > I) {
            J -= I + 1;
            ++I;
        }
        const uint32_t p = 16 * I + pl, q = 16 * J + ql;
        if (p >= D || q > p) continue;
        double v = buf0[tile * 256 + pl * 16 + ql];
        for (int u = tile + ntiles; u < units; u += ntiles) v += buf0[u * 256 + pl * 16 + ql];
        if (p == q || p == q + M) {
            const uint32_t i0 = q < M ? q : q - M;
            const double am = (double)amps[i0];
            const double2 ee = E[N + i0];
            const double cr = -am * ee.x, ci = -am * ee.y;
            v += p != q ? ci * cr : (p < M ? cr * cr : ci * ci);
        }
        if (p == q)
            sgd[p] = v;
        else
            sK[p * kLmMaxD + q] = v;
    }
}

// 1 / x to FP64 rounding: v_rcp_f64 and two Newton steps.
__device__ __forceinline__ double rcp_f64(double x) {
    double y = __builtin_amdgcn_rcp(x);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const double e = fma(-x, y, 1.0);
        y = fma(y, e, y);
    }
    return y;
}

// One wave solves (mu I + Gram) y = v (D <= DP <= 32; the Gram matrix in the
// lower triangle of sK and its diagonal in sgd) by K = L diag(d) L^T with
// unit lower L: lane i keeps row i in registers (rows i >= D are identity
// rows, so the padded system is block-diagonal), the pivot and the column
// entries of other lanes by v_readlane. Column k scales by 1/d_k (v_rcp_f64
// and two Newton steps); the updates read the unscaled column, so they do
// not wait for the reciprocal, and are not masked by lane: a lane's entries
// right of its diagonal receive values that nothing reads. Branch-free:
// every lane loads its whole row of sK (128-bit reads, fenced in groups of
// four by an empty asm so that the compiler neither moves them into a
// branch nor waits for each one) and selects, and writes all of its row of
// L to sL (row stride kLmMaxD), from where the backward substitution reads
// its column. Then
// L z = v, L^T y = z / d. Returns false when a pivot is not positive (or
// NaN).
template <int DP>
__device__ __attribute__((noinline)) bool chol_lm_wave(const double* sK, const double* sgd, double mu,
                                                       const double* sv, double* sy, double* sL, uint32_t D) {
    // the lane index through an empty asm keeps the per-lane masks (i < j,
    // i == k, ...) from being hoisted into long-lived SGPR pairs
    int i = threadIdx.x;  // < 64, the whole wave active
    __asm__ volatile("" : "+v"(i));
    const bool real = i < (int)D;
    const double di = sgd[i] + mu;
    typedef double d2v __attribute__((ext_vector_type(2)));
    const d2v* row = reinterpret_cast<const d2v*>(sK + i * kLmMaxD);
    d2v kv[DP / 2];
#pragma unroll
    for (int h = 0; h < DP / 2; ++h) kv[h] = row[h];
#pragma unroll
    for (int g = 0; g < DP / 8; ++g)
        __asm__ volatile("" : "+v"(kv[4 * g]), "+v"(kv[4 * g + 1]), "+v"(kv[4 * g + 2]), "+v"(kv[4 * g + 3]));
    double r[DP];
#pragma unroll
    for (int j = 0; j < DP; ++j) {
        const double kij = kv[j / 2][j % 2];
        r[j] = real ? (j < i ? kij : (j == i ? di : 0.0)) : (j == i ? 1.0 : 0.0);
    }
    bool ok = true;
    double dinv = 1.0;
#pragma unroll
    for (int k = 0; k < DP; ++k) {
        const double dkk = readlane_f64(r[k], k);
        ok = ok && dkk > 0.0;
        const double inv = rcp_f64(dkk);
        dinv = i == k ? inv : dinv;
        const double lik = r[k] * inv;
#pragma unroll
        for (int j = k + 1; j < DP; ++j) r[j] = fma(-lik, readlane_f64(r[k], j), r[j]);
        r[k] = i == k ? dkk : lik;
    }
#pragma unroll
    for (int j = 0; j < DP; ++j) sL[i * kLmMaxD + j] = r[j];
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // L z = v (a lane's acc is final when its own step reads it)
    double acc = real ? sv[i] : 0.0, z = 0.0;
#pragma unroll
    for (int j = 0; j < DP; ++j) {
        const double zj = readlane_f64(acc, j);
        z = i == j ? zj : z;
        acc = fma(-r[j], zj, acc);
    }
    // L^T y = z / d with column i of L (L_ji, j > i) from sL
    double lc[DP];
#pragma unroll
    for (int j = 0; j < DP; ++j) lc[j] = sL[j * kLmMaxD + i];
    acc = z * dinv;
    double y = 0.0;
#pragma unroll
    for (int j = DP - 1; j >= 0; --j) {
        const double yj = readlane_f64(acc, j);
        y = i == j ? yj : y;
        acc = fma(-lc[j], yj, acc);
    }
    if (real) sy[i] = y;
    return ok;
}

// The same L diag(d) L^T solve for 32 < D <= 64 with the rows in LDS (work,
// row stride kLmMaxD + 1, so that a column spreads over the banks) and the
// column broadcast through col.
__device__ __attribute__((noinline)) bool chol_lm_wave_lds(const double* sK, const double* sgd, double mu,
                                                           const double* sv, double* sy, double* col, double* work,
                                                           uint32_t D) {
    constexpr int DP = kLmMaxD, WS = kLmMaxD + 1;
    int i = threadIdx.x;  // < 64
    __asm__ volatile("" : "+v"(i));  // as in chol_lm_wave
    const bool real = i < (int)D;
    for (int j = 0; j < DP; ++j)
        work[i * WS + j] = real ? (j < i ? sK[i * kLmMaxD + j] : (j == i ? sgd[i] + mu : 0.0)) : (j == i ? 1.0 : 0.0);
    bool ok = true;
    double dinv = 1.0;
    for (int k = 0; k < DP; ++k) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const double dkk = work[k * WS + k];
        ok = ok && dkk > 0.0;
        const double inv = rcp_f64(dkk);
        dinv = i == k ? inv : dinv;
        const double rk = work[i * WS + k];
        col[i] = rk;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const double lik = rk * inv;
        for (int j = k + 1; j < DP; ++j) work[i * WS + j] = fma(-lik, col[j], work[i * WS + j]);
        work[i * WS + k] = i == k ? dkk : lik;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    double acc = real ? sv[i] : 0.0, z = 0.0;
    for (int j = 0; j < DP; ++j) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        if (i == j) col[j] = acc;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const double zj = col[j];
        z = i == j ? zj : z;
        acc = fma(-work[i * WS + j], zj, acc);
    }
    acc = z * dinv;
    double y = 0.0;
    for (int j = DP - 1; j >= 0; --j) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        if (i == j) col[j] = acc;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const double yj = col[j];
        y = i == j ? yj : y;
        acc = fma(-work[j * WS + i], yj, acc);
    }
    if (real) sy[i] = y;
    return ok;
}

// A workgroup-uniform pointer, read from the first lane (so it lives in
// SGPRs) and passed through an empty asm: addresses derived from it are
// recomputed where they are used instead of being hoisted out of the
// iteration loop, where dozens of them would stay live across every step.
template <typename T>
__device__ __forceinline__ T* fresh(T* p) {
    const uint64_t u = reinterpret_cast<uint64_t>(p);
    uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)u);
    uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(u >> 32));
    __asm__ volatile("" : "+s"(lo), "+s"(hi));
    return reinterpret_cast<T*>(((uint64_t)hi << 32) | lo);
}

__global__ __launch_bounds__(kLW) void lm_wide(LmArgs a) {
    __shared__ double s_K[kLmMaxD * kLmMaxD];
    __shared__ double s_gd[kLmMaxD];
    __shared__ double s_v[kLmMaxD];
    __shared__ double s_y[kLmMaxD];
    __shared__ double2 s_s[kLmMaxM];
    __shared__ double s_red[3 * kLWaves];
    __shared__ double s_col[kLmMaxD];
    __shared__ double s_work[kLmMaxD * (kLmMaxD + 1)];
    __shared__ double s_part[kLmPart];
    __shared__ int s_ok;
    const int t = threadIdx.x;
    const uint32_t b = blockIdx.x;
    const uint32_t base = a.offsets[b], M = a.offsets[b + 1] - base;
    const uint32_t N = a.n, n = N + M, D = 2 * M;
    const float* amps = a.amps + base;
    const LmScratch s = lm_scratch(a.scratch + b * a.scratch_stride, N, M);
    double* x = s.x;
    double* xn = s.xn;
    double2* E = s.E;
    double2* En = s.En;
    const double EPS_1 = 1e-8, EPS_2 = 1e-8, TAU = 1e-3;

    // From the focal field s of the current point (s_s): v = [Im s; -Re s],
    // g = W^T v with W = [Re C; Im C], C = B diag(E), and the Gram matrix
    // W W^T (lower triangle of s_K, diagonal in s_gd). Returns max_l g_l and,
    // when asked, max_l diag(A)_l = max_l sum_i |B_il|^2.
    auto linearize = [&](bool want_amax, double& gmax, double& amax) {
        if (t < (int)M) {
            s_v[t] = s_s[t].y;
            s_v[M + t] = -s_s[t].x;
        }
        lds_barrier();
        const double2* __restrict__ Gp = fresh(s.G);
        double* __restrict__ gp = fresh(s.g);
        double mx[2] = {-INFINITY, -INFINITY};
        for (uint32_t l = t; l < n; l += kLW) {
            const double2 e = E[l];
            double gr = 0.0, gi = 0.0, colsq = 0.0;
#pragma unroll 16
            for (uint32_t i = 0; i < M; ++i) {
                double2 bv;
                if (l < N)
                    bv = Gp[(uint64_t)i * N + l];
                else
                    bv = make_double2(l - N == i ? -(double)amps[i] : 0.0, 0.0);
                const double cr = bv.x * e.x - bv.y * e.y, ci = bv.x * e.y + bv.y * e.x;
                gr = fma(cr, s_v[i], gr);
                gi = fma(ci, s_v[M + i], gi);
                if (want_amax) colsq += cr * cr + ci * ci;
            }
            const double gl = gr + gi;
            gp[l] = gl;
            mx[0] = fmax(mx[0], gl);
            mx[1] = fmax(mx[1], colsq);
        }
        lw_reduce<2, true>(mx, s_red);
        gmax = mx[0];
        amax = mx[1];
        lw_gram(Gp, E, N, M, amps, s_K, s_gd, s_work, s_part);
        lds_barrier();
    };

    for (uint32_t l = t; l < n; l += kLW) {
        x[l] = 0.0;
        E[l] = make_double2(1.0, 0.0);
    }
    __syncthreads();
    lw_field(s.G, E, N, M, amps, s_s);
    lds_barrier();
    double F = 0.0;
    for (uint32_t i = 0; i < M; ++i) F += s_s[i].x * s_s[i].x + s_s[i].y * s_s[i].y;
    double gmax, amax;
    linearize(true, gmax, amax);
    double mu = TAU * amax, nu = 2.0;
    bool found = gmax <= EPS_1;
    uint32_t it = 0;
    for (; it < a.k_max; ++it) {
        if (found) break;
        x = fresh(x);
        xn = fresh(xn);
        E = fresh(E);
        En = fresh(En);
        const double* __restrict__ gm = fresh(s.g);
        // (mu I + W W^T) y = v by wave 0
        if (t < 64) {
            bool okw;
            if (D <= 16)
                okw = chol_lm_wave<16>(s_K, s_gd, mu, s_v, s_y, s_work, D);
            else if (D <= 32)
                okw = chol_lm_wave<32>(s_K, s_gd, mu, s_v, s_y, s_work, D);
            else
                okw = chol_lm_wave_lds(s_K, s_gd, mu, s_v, s_y, s_col, s_work, D);
            if (t == 0) s_ok = okw ? 1 : 0;
        }
        lds_barrier();
        if (!s_ok) break;  // the reference stops when its solve fails (lm.rs:72-77)
        // h = -W^T y; |h|^2, |x|^2, h.(mu h - g); x_new, e^{i x_new}
        double sums[3] = {0.0, 0.0, 0.0};
        auto step = [&](uint32_t l, double acc) {
            const double hl = -acc;
            const double xl = x[l];
            sums[0] += hl * hl;
            sums[1] += xl * xl;
            sums[2] += hl * (mu * hl - gm[l]);
            const double xv = xl + hl;
            xn[l] = xv;
            double sn, cs;
            sincos(xv, &sn, &cs);
            En[l] = make_double2(cs, sn);
        };
        // h_l = -(W^T y)_l = -sum_i (Re C_il y_i + Im C_il y_{M+i}), C_il = B_il E_l from G
        const double2* __restrict__ Gh = fresh(s.G);
        auto hcol = [&](uint32_t l, double2 e, uint32_t i, double& hr, double& hi) {
            const double2 bv = l < N ? Gh[(uint64_t)i * N + l] : make_double2(l - N == i ? -(double)amps[i] : 0.0, 0.0);
            hr = fma(bv.x * e.x - bv.y * e.y, s_y[i], hr);
            hi = fma(bv.x * e.y + bv.y * e.x, s_y[M + i], hi);
        };
        if (n <= (uint32_t)kLW) {
            // one column per thread: all of its loads in flight at once
            if ((uint32_t)t < n) {
                const double2 e0 = E[t];
                double hr0 = 0.0, hi0 = 0.0;
#pragma unroll 8
                for (uint32_t i = 0; i < M; ++i) hcol(t, e0, i, hr0, hi0);
                step(t, hr0 + hi0);
            }
        } else {
            // columns l0 and l0 + kLW together, so that both columns' loads are in flight
            for (uint32_t l0 = t; l0 < n; l0 += 2 * kLW) {
                const uint32_t l1 = l0 + kLW;
                const bool two = l1 < n;
                const uint32_t l1c = two ? l1 : l0;
                const double2 e0 = E[l0], e1 = E[l1c];
                double hr0 = 0.0, hi0 = 0.0, hr1 = 0.0, hi1 = 0.0;
#pragma unroll 4
                for (uint32_t i = 0; i < M; ++i) {
                    hcol(l0, e0, i, hr0, hi0);
                    hcol(l1c, e1, i, hr1, hi1);
                }
                step(l0, hr0 + hi0);
                if (two) step(l1, hr1 + hi1);
            }
        }
        lw_reduce<3, false>(sums, s_red);
        const double hnorm = sqrt(sums[0]), xnorm = sqrt(sums[1]), L0 = 0.5 * sums[2];
        if (hnorm <= EPS_2 * (xnorm + EPS_2)) {
            found = true;
            continue;
        }
        __syncthreads();  // e^{i x_new} before every wave reads it
        lw_field(fresh(s.G), En, N, M, amps, s_s);
        lds_barrier();
        double Fn = 0.0;
        for (uint32_t i = 0; i < M; ++i) Fn += s_s[i].x * s_s[i].x + s_s[i].y * s_s[i].y;
        const double rho = (F - Fn) / L0;
        F = Fn;  // assigned on rejection too (lm.rs:85)
        if (rho > 0.0) {
            double* tx = x;
            x = xn;
            xn = tx;
            double2* te = E;
            E = En;
            En = te;
            double am2;
            linearize(false, gmax, am2);
            found = gmax <= EPS_1;
            const double c = 2.0 * rho - 1.0;
            mu *= fmax(1.0 / 3.0, 1.0 - c * c * c);
            nu = 2.0;
        } else {
            mu *= nu;
            nu *= 2.0;
        }
    }
    for (uint32_t j = t; j < N; j += kLW) a.out_phase[(uint64_t)b * N + j] = (float)x[j];
    if (t == 0 && a.out_iters) a.out_iters[b] = it;
}

// ---------------------------------------------------------------- APO
// Scratch of one APO instance: G (M x N) first, as lm_gen_g writes it, then
// H (N x N, column-major) and the N-vectors q, q_new, d, grad, grad_new, s, u.
__host__ __device__ inline uint64_t apo_scratch_bytes(uint64_t N, uint64_t M) {
    auto al = [](uint64_t b) { return (b + 255) & ~(uint64_t)255; };
    return al(M * N * 16) + al(N * N * 16) + 7 * al(N * 16);
}

struct ApoScratch {
    double2 *G, *H, *q, *qn, *d, *gr, *grn, *sv, *u;
};

__device__ inline ApoScratch apo_scratch(char* base, uint64_t N, uint64_t M) {
    auto al = [](uint64_t b) { return (b + 255) & ~(uint64_t)255; };
    ApoScratch s;
    char* p = base;
    s.G = reinterpret_cast<double2*>(p);
    p += al(M * N * 16);
    s.H = reinterpret_cast<double2*>(p);
    p += al(N * N * 16);
    double2** v[7] = {&s.q, &s.qn, &s.d, &s.gr, &s.grn, &s.sv, &s.u};
    for (int k = 0; k < 7; ++k) {
        *v[k] = reinterpret_cast<double2*>(p);
        p += al(N * 16);
    }
    return s;
}

// gq_i = g_i q for every focus (a wave per row, lanes over sources).
__device__ inline void rows_times(const double2* G, const double2* q, uint32_t N, uint32_t M, double2* out) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    for (uint32_t i = w; i < M; i += kLB / 64) {
        double re = 0.0, im = 0.0;
        for (uint32_t j = l; j < N; j += 64) {
            const double2 g = G[(uint64_t)i * N + j], v = q[j];
            re += g.x * v.x - g.y * v.y;
            im += g.x * v.y + g.y * v.x;
        }
        re = wave_sum_f64(re);
        im = wave_sum_f64(im);
        if (l == 0) out[i] = make_double2(re, im);
    }
}

// Gradient descent (gradient_descent.rs:49-56) with G held in registers:
// thread T owns the columns l = T + kLB c (c < CPT, l < N) of G for MP >= M
// foci (zero rows beyond M), its phases x_l and E_l = e^{i x_l}; thread i < M
// also owns the amplitude parameter x_{N+i} (column N + i of B = [G |
// -diag(amps)]). A step: the thread's partial field sums over its columns, 2MP
// values, reduced over the wave by a transposed butterfly (at offset o = 1, 2,
// .. 32 the lane with bit o set keeps the upper half of its list and sends the
// lower half, so 2MP values cost 2MP - 1 lane moves, most of them DPP, and
// lane L ends with the value whose index is L's bits reversed), then over the
// 4 waves in a fixed
// order; g_l = Im(conj(E_l) sum_i conj(G_il) s_i); the block maximum of g
// (the reference's stop test); x -= 0.1 g. No global memory is touched inside
// the loop.
template <int MP, int CPT>
__global__ __launch_bounds__(kLB) void gd_solve(LmArgs a) {
    constexpr int K = 2 * MP;  // re s_i at i, im s_i at MP + i
    static_assert(K == 32 || K == 64, "16 or 32 foci");
    __shared__ double s_part[4][64];
    __shared__ double s_f[K];
    __shared__ double2 s_ea[MP];
    __shared__ double s_red[4];
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    const uint32_t b = blockIdx.x;
    const uint32_t base = a.offsets[b], M = a.offsets[b + 1] - base;
    const uint32_t N = a.n;
    LmScratch sc = lm_scratch(a.scratch + b * a.scratch_stride, N, M);
    double gr[CPT][MP], gi[CPT][MP], x[CPT];
    double2 e[CPT];
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
        const uint32_t l = t + kLB * c;
#pragma unroll
        for (int i = 0; i < MP; ++i) {
            const double2 v = (l < N && i < (int)M) ? sc.G[(uint64_t)i * N + l] : make_double2(0.0, 0.0);
            gr[c][i] = v.x;
            gi[c][i] = v.y;
        }
        x[c] = 0.0;
        e[c] = make_double2(1.0, 0.0);
    }
    const bool own_amp = t < (int)M;
    const double amp = own_amp ? (double)a.amps[base + t] : 0.0;
    const bool amp_row = t < K && t % MP < (int)M;  // thread t < K sums value t
    const double amp_f = amp_row ? (double)a.amps[base + t % MP] : 0.0;
    double xa = 0.0;
    if (t < MP) s_ea[t] = make_double2(1.0, 0.0);
    uint32_t it = 0;
    for (; it < a.k_max; ++it) {
        // partial sums over this thread's columns: s_i = sum_l G_il E_l
        double v[K];
#pragma unroll
        for (int i = 0; i < MP; ++i) {
            double re = 0.0, im = 0.0;
#pragma unroll
            for (int c = 0; c < CPT; ++c) {
                re += gr[c][i] * e[c].x - gi[c][i] * e[c].y;
                im += gr[c][i] * e[c].y + gi[c][i] * e[c].x;
            }
            v[i] = re;
            v[MP + i] = im;
        }
        // transposed butterfly over the wave, low lane bits first: the steps
        // with many values move them by DPP, only the last one or two values
        // cross 16-lane rows through the LDS crossbar
        butterfly_step<1, K>(v, lane);
        butterfly_step<2, K / 2>(v, lane);
        butterfly_step<4, K / 4>(v, lane);
        butterfly_step<8, K / 8>(v, lane);
        butterfly_step<16, K / 16>(v, lane);
        if constexpr (K == 64) butterfly_step<32, 2>(v, lane);
        double f = v[0];
        if constexpr (K == 32) f += xor_lane_f64<32>(f);
        // lane bit b chose the half of weight K / 2^(b + 1): the value index is
        // the lane's bits reversed (6 bits for K = 64, bits 0..4 for K = 32)
        const int idx = K == 64 ? (((lane & 1) << 5) | ((lane & 2) << 3) | ((lane & 4) << 1) | ((lane & 8) >> 1) |
                                   ((lane & 16) >> 3) | ((lane & 32) >> 5))
                                : (((lane & 1) << 4) | ((lane & 2) << 2) | (lane & 4) | ((lane & 8) >> 2) |
                                   ((lane & 16) >> 4));
        if (K == 64 || (lane & 32) == 0) s_part[w][idx] = f;
        lds_barrier();
        if (t < K) {
            double sum = ((s_part[0][t] + s_part[1][t]) + s_part[2][t]) + s_part[3][t];
            if (amp_row) {
                // the -diag(amps) block of B: column N + t mod MP
                const double2 ea = s_ea[t % MP];
                sum -= amp_f * (t < MP ? ea.x : ea.y);
            }
            s_f[t] = sum;
        }
        lds_barrier();
        // g_l = Im(conj(C_il) s_i) summed over i, C_il = G_il E_l
        double gm = -INFINITY, g[CPT];
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
            double tr = 0.0, ti = 0.0;
#pragma unroll
            for (int i = 0; i < MP; ++i) {
                const double fr = s_f[i], fi = s_f[MP + i];
                tr += gr[c][i] * fr + gi[c][i] * fi;
                ti += gr[c][i] * fi - gi[c][i] * fr;
            }
            g[c] = e[c].x * ti - e[c].y * tr;
            if (t + kLB * c < N) gm = fmax(gm, g[c]);
        }
        double ga = 0.0;
        if (own_amp) {
            const double2 ea = s_ea[t];
            ga = -amp * (ea.x * s_f[MP + t] - ea.y * s_f[t]);
            gm = fmax(gm, ga);
        }
        gm = fmax(gm, xor_lane_f64<1>(gm));
        gm = fmax(gm, xor_lane_f64<2>(gm));
        gm = fmax(gm, xor_lane_f64<4>(gm));
        gm = fmax(gm, xor_lane_f64<8>(gm));
        gm = fmax(gm, xor_lane_f64<16>(gm));
        gm = fmax(gm, xor_lane_f64<32>(gm));
        if (lane == 0) s_red[w] = gm;
        lds_barrier();
        const double gmax = fmax(fmax(s_red[0], s_red[1]), fmax(s_red[2], s_red[3]));
        if (gmax <= 1e-6) break;
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
            x[c] -= 0.1 * g[c];
            double sn, cs;
            sincos(x[c], &sn, &cs);
            e[c] = make_double2(cs, sn);
        }
        if (own_amp) {
            xa -= 0.1 * ga;
            double sn, cs;
            sincos(xa, &sn, &cs);
            s_ea[t] = make_double2(cs, sn);
        }
    }
#pragma unroll
    for (int c = 0; c < CPT; ++c)
        if (t + kLB * c < N) a.out_phase[(uint64_t)b * N + t + kLB * c] = (float)x[c];
    if (t == 0 && a.out_iters) a.out_iters[b] = it;
}

__global__ __launch_bounds__(kLB) void apo_solve(LmArgs a) {
    __shared__ double s_K[kLmMaxD * kLmMaxD];
    __shared__ double s_v[kLmMaxD], s_y[kLmMaxD], s_col[kLmMaxD];
    __shared__ double s_work[kLmMaxD * (kLmMaxD + 1)];
    __shared__ double2 s_gq[kLmMaxM], s_gd[kLmMaxM];
    __shared__ double s_p2[kLmMaxM];
    __shared__ double s_J[kApoLineSearch];
    __shared__ double s_red[4];
    __shared__ double s_alpha;
    const int t = threadIdx.x;
    const uint32_t b = blockIdx.x;
    const uint32_t base = a.offsets[b], M = a.offsets[b + 1] - base;
    const uint32_t N = a.n, D = 2 * M;
    const double lam = a.lambda;
    const float* amps = a.amps + base;
    ApoScratch s = apo_scratch(a.scratch + b * a.scratch_stride, N, M);
    for (uint32_t i = t; i < M; i += kLB) s_p2[i] = (double)amps[i] * (double)amps[i];

    // q0 = G^H (G G^H + lambda I)^{-1} p: the Hermitian M x M system as the
    // real symmetric 2M x 2M system [[Re K, -Im K], [Im K, Re K]] (lower
    // triangle into s_K)
    for (uint32_t e = t; e < M * M; e += kLB) {
        const uint32_t i = e / M, k = e % M;
        if (k > i) continue;
        double re = 0.0, im = 0.0;
#pragma unroll 8
        for (uint32_t l = 0; l < N; ++l) {
            const double2 gi = s.G[(uint64_t)i * N + l], gk = s.G[(uint64_t)k * N + l];
            re += gi.x * gk.x + gi.y * gk.y;  // G_il conj(G_kl)
            im += gi.y * gk.x - gi.x * gk.y;
        }
        if (i == k) re += lam;
        s_K[i * kLmMaxD + k] = re;
        s_K[(i + M) * kLmMaxD + (k + M)] = re;
        s_K[(i + M) * kLmMaxD + k] = im;
        if (i != k) s_K[(k + M) * kLmMaxD + i] = -im;
    }
    for (uint32_t i = t; i < D; i += kLB) s_v[i] = i < M ? (double)amps[i] : 0.0;
    __syncthreads();
    if (t < 64) {
        if (D <= 16)
            (void)chol_solve_wave<16, 1>(s_K, s_v, s_y, s_col, D);
        else if (D <= 32)
            (void)chol_solve_wave<32, 1>(s_K, s_v, s_y, s_col, D);
        else
            (void)chol_solve_wave_lds<1>(s_K, s_v, s_y, s_col, s_work, D);
    }
    __syncthreads();
    for (uint32_t j = t; j < N; j += kLB) {
        double re = 0.0, im = 0.0;
        for (uint32_t i = 0; i < M; ++i) {
            const double2 g = s.G[(uint64_t)i * N + j];
            const double zr = s_y[i], zi = s_y[M + i];
            re += g.x * zr + g.y * zi;  // conj(G_ij) z_i
            im += g.x * zi - g.y * zr;
        }
        s.q[j] = make_double2(re, im);
    }
#pragma unroll 4
    for (uint64_t e = t; e < (uint64_t)N * N; e += kLB)
        s.H[e] = make_double2(e % ((uint64_t)N + 1) == 0 ? 1.0 : 0.0, 0.0);
    __syncthreads();

    // grad J (apo.rs:56-70) at v into out
    auto grad = [&](const double2* v, double2* out) {
        rows_times(s.G, v, N, M, s_gq);
        __syncthreads();
        for (uint32_t j = t; j < N; j += kLB) {
            double re = 0.0, im = 0.0;
#pragma unroll 4
            for (uint32_t i = 0; i < M; ++i) {
                const double2 g = s.G[(uint64_t)i * N + j], gq = s_gq[i];
                const double w = s_p2[i] - (gq.x * gq.x + gq.y * gq.y);
                re += (g.x * gq.x + g.y * gq.y) * w;  // conj(G_ij) gq_i w_i
                im += (g.x * gq.y - g.y * gq.x) * w;
            }
            out[j] = make_double2(re + lam * v[j].x, im + lam * v[j].y);
        }
        __syncthreads();
    };
    grad(s.q, s.gr);
    uint32_t it = 0;
    for (; it < a.k_max; ++it) {
        // d = -H grad
        for (uint32_t j = t; j < N; j += kLB) {
            double re = 0.0, im = 0.0;
#pragma unroll 8
            for (uint32_t k = 0; k < N; ++k) {
                const double2 h = s.H[(uint64_t)k * N + j], g = s.gr[k];
                re += h.x * g.x - h.y * g.y;
                im += h.x * g.y + h.y * g.x;
            }
            s.d[j] = make_double2(-re, -im);
        }
        __syncthreads();
        // line search (apo.rs:76-97): J(q + a d), a = l / 100, first minimum
        rows_times(s.G, s.q, N, M, s_gq);
        rows_times(s.G, s.d, N, M, s_gd);
        double2 qq = make_double2(0.0, 0.0), qd = qq, dd = qq;
        for (uint32_t j = t; j < N; j += kLB) {
            const double2 qv = s.q[j], dv = s.d[j];
            const double2 a1 = cmul(qv, qv), a2 = cmul(qv, dv), a3 = cmul(dv, dv);
            qq.x += a1.x;
            qq.y += a1.y;
            qd.x += a2.x;
            qd.y += a2.y;
            dd.x += a3.x;
            dd.y += a3.y;
        }
        qq = block_sum2(qq, s_red);
        qd = block_sum2(qd, s_red);
        dd = block_sum2(dd, s_red);
        for (uint32_t l = t; l < kApoLineSearch; l += kLB) {
            const double al = (double)l / (double)kApoLineSearch;
            double v = 0.0;
#pragma unroll 4
            for (uint32_t i = 0; i < M; ++i) {
                const double re = s_gq[i].x + al * s_gd[i].x, im = s_gq[i].y + al * s_gd[i].y;
                const double sdiff = re * re + im * im - s_p2[i];
                v += sdiff * sdiff;
            }
            // (q + a d) . (q + a d) without conjugation
            const double dr = qq.x + 2.0 * al * qd.x + al * al * dd.x, di = qq.y + 2.0 * al * qd.y + al * al * dd.y;
            s_J[l] = v + sqrt(dr * dr + di * di) * lam;
        }
        __syncthreads();
        if (t < 64) {
            double bv = INFINITY;
            int bi = kApoLineSearch;
            for (int l = t; l < kApoLineSearch; l += 64)
                if (s_J[l] < bv) {
                    bv = s_J[l];
                    bi = l;
                }
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
                const double ov = __shfl_xor(bv, off);
                const int oi = __shfl_xor(bi, off);
                if (ov < bv || (ov == bv && oi < bi)) {
                    bv = ov;
                    bi = oi;
                }
            }
            if (t == 0) s_alpha = bi < kApoLineSearch ? (double)bi / (double)kApoLineSearch : 0.0;
        }
        __syncthreads();
        const double alpha = s_alpha;
        double dn = 0.0;
        for (uint32_t j = t; j < N; j += kLB) {
            const double2 dv = make_double2(s.d[j].x * alpha, s.d[j].y * alpha);
            s.d[j] = dv;
            dn += dv.x * dv.x + dv.y * dv.y;
            s.qn[j] = make_double2(s.q[j].x + dv.x, s.q[j].y + dv.y);
        }
        dn = block_sum(dn, s_red);
        if (sqrt(dn) < 1e-8) break;
        grad(s.qn, s.grn);
        for (uint32_t j = t; j < N; j += kLB) s.sv[j] = make_double2(s.grn[j].x - s.gr[j].x, s.grn[j].y - s.gr[j].y);
        __syncthreads();
        // u = H s, y . s, s . u (no conjugation)
        double2 ys = make_double2(0.0, 0.0), su = ys;
        for (uint32_t j = t; j < N; j += kLB) {
            double re = 0.0, im = 0.0;
#pragma unroll 8
            for (uint32_t k = 0; k < N; ++k) {
                const double2 h = s.H[(uint64_t)k * N + j], v = s.sv[k];
                re += h.x * v.x - h.y * v.y;
                im += h.x * v.y + h.y * v.x;
            }
            const double2 uj = make_double2(re, im);
            s.u[j] = uj;
            const double2 a1 = cmul(s.d[j], s.sv[j]), a2 = cmul(s.sv[j], uj);
            ys.x += a1.x;
            ys.y += a1.y;
            su.x += a2.x;
            su.y += a2.y;
        }
        ys = block_sum2(ys, s_red);
        su = block_sum2(su, s_red);
        const double ny = ys.x * ys.x + ys.y * ys.y, nu = su.x * su.x + su.y * su.y;
        const double2 iy = make_double2(ys.x / ny, -ys.y / ny), iu = make_double2(su.x / nu, -su.y / nu);
        // H += y y^T / (y . s) - u u^T / (s . u)
#pragma unroll 4
        for (uint64_t e = t; e < (uint64_t)N * N; e += kLB) {
            const uint32_t j = (uint32_t)(e % N), k = (uint32_t)(e / N);
            const double2 a1 = cmul(cmul(s.d[j], s.d[k]), iy), a2 = cmul(cmul(s.u[j], s.u[k]), iu);
            const double2 h = s.H[e];
            s.H[e] = make_double2(h.x + a1.x - a2.x, h.y + a1.y - a2.y);
        }
        for (uint32_t j = t; j < N; j += kLB) {
            s.q[j] = s.qn[j];
            s.gr[j] = s.grn[j];
        }
        __syncthreads();
    }
    // camax (nalgebra): max_j |Re q_j| + |Im q_j|
    double cm = 0.0;
    for (uint32_t j = t; j < N; j += kLB) cm = fmax(cm, fabs(s.q[j].x) + fabs(s.q[j].y));
    cm = block_max(cm, s_red);
    for (uint32_t j = t; j < N; j += kLB) {
        const double2 qv = s.q[j];
        const double amp = (double)a.sources[AFH_SOURCE_STRIDE * (uint64_t)j + 6];
        a.out_amp[(uint64_t)b * N + j] = (float)(amp * sqrt(qv.x * qv.x + qv.y * qv.y) / cm);
        a.out_phase[(uint64_t)b * N + j] = (float)atan2(qv.y, qv.x);
    }
    if (t == 0 && a.out_iters) a.out_iters[b] = it;
}

int32_t apo_solve_batch(Context* ctx, const float* sources, uint32_t n, int source_type, const float* foci,
                        const float* amps, const uint32_t* offsets_host, uint32_t batch, double lambda,
                        uint32_t k_max, float* out_amp, float* out_phase, uint32_t* out_iters, hipStream_t st) {
    if (batch == 0) return AFH_OK;
    uint32_t maxm = 0;
    for (uint32_t b = 0; b < batch; ++b) {
        const uint32_t m = offsets_host[b + 1] - offsets_host[b];
        if (m == 0) {
            set_error("APO: an instance has no foci");
            return AFH_ERR_ARG;
        }
        maxm = std::max(maxm, m);
    }
    if (maxm > (uint32_t)kLmMaxM) {
        set_error("APO: more than 32 foci in one instance is not supported");
        return AFH_ERR_UNSUPPORTED;
    }
    const uint64_t stride = std::max(apo_scratch_bytes(n, maxm), lm_scratch_bytes(n, maxm));
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
    LmArgs a;
    memset(&a, 0, sizeof(a));
    a.sources = sources;
    a.source_type = source_type;
    a.n = n;
    a.foci = foci;
    a.amps = amps;
    a.offsets = reinterpret_cast<const uint32_t*>(w);
    a.k_max = k_max;
    a.lambda = lambda;
    a.scratch = w + o_scr;
    a.scratch_stride = stride;
    a.out_amp = out_amp;
    a.out_phase = out_phase;
    a.out_iters = out_iters;
    hipLaunchKernelGGL(lm_gen_g, dim3((n + kLB - 1) / kLB, batch), dim3(kLB), 0, st, a);
    hipLaunchKernelGGL(apo_solve, dim3(batch), dim3(kLB), 0, st, a);
    if (!hip_ok(hipGetLastError(), "APO kernel launch")) return AFH_ERR_HIP;
    return AFH_OK;
}

int32_t lm_solve_batch(Context* ctx, const float* sources, uint32_t n, uint32_t n_packed, int source_type,
                       const float* foci, const float* amps, const uint32_t* offsets_host, uint32_t batch,
                       uint32_t k_max, float* out_phase, uint32_t* out_iters, hipStream_t st, int algo) {
    (void)n_packed;
    if (batch == 0) return AFH_OK;
    uint32_t maxm = 0;
    for (uint32_t b = 0; b < batch; ++b) {
        const uint32_t m = offsets_host[b + 1] - offsets_host[b];
        if (m == 0) {
            set_error("LM: an instance has no foci");
            return AFH_ERR_ARG;
        }
        maxm = std::max(maxm, m);
    }
    if (maxm > (uint32_t)kLmMaxM) {
        set_error("LM: more than 32 foci in one instance is not supported");
        return AFH_ERR_UNSUPPORTED;
    }
    const uint64_t stride = lm_scratch_bytes(n, maxm);
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
    LmArgs a;
    memset(&a, 0, sizeof(a));
    a.sources = sources;
    a.source_type = source_type;
    a.n = n;
    a.foci = foci;
    a.amps = amps;
    a.offsets = reinterpret_cast<const uint32_t*>(w);
    a.k_max = k_max;
    a.algo = algo;
    a.scratch = w + o_scr;
    a.scratch_stride = stride;
    a.out_phase = out_phase;
    a.out_iters = out_iters;
    hipLaunchKernelGGL(lm_gen_g, dim3((n + kLB - 1) / kLB, batch), dim3(kLB), 0, st, a);
    if (algo == kNlsGD && maxm <= 16 && n <= kLB)
        hipLaunchKernelGGL((gd_solve<16, 1>), dim3(batch), dim3(kLB), 0, st, a);
    else if (algo == kNlsGD && maxm <= 16 && n <= 2 * kLB)
        hipLaunchKernelGGL((gd_solve<16, 2>), dim3(batch), dim3(kLB), 0, st, a);
    else if (algo == kNlsGD && n <= kLB)
        hipLaunchKernelGGL((gd_solve<32, 1>), dim3(batch), dim3(kLB), 0, st, a);
    else if (algo == kNlsLM)
        hipLaunchKernelGGL(lm_wide, dim3(batch), dim3(kLW), 0, st, a);
    else
        hipLaunchKernelGGL(lm_solve, dim3(batch), dim3(kLB), 0, st, a);
    if (!hip_ok(hipGetLastError(), "LM kernel launch")) return AFH_ERR_HIP;
    return AFH_OK;
}

bool lm_solve_single(const HoloProblem& p, float* phase_out, int algo) {
    if (!device().ensure()) return false;
    Context& ctx = default_context();
    std::lock_guard<std::mutex> g(ctx.mu);
    if (!ctx.ensure_stream()) return false;
    const uint64_t n = p.n, np = packed_count(n), m = p.m;
    const size_t src_b = np * AFH_SOURCE_STRIDE * sizeof(float), foci_b = m * 3 * sizeof(float);
    const size_t amps_b = m * sizeof(float), out_b = n * sizeof(float);
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t o_foci = align(src_b), o_amps = o_foci + align(foci_b), o_ph = o_amps + align(amps_b);
    const size_t total = o_ph + align(out_b);
    if (!ctx.io.reserve(total) || !ctx.io_host.reserve(total)) return false;
    char* h = static_cast<char*>(ctx.io_host.ptr);
    char* d = static_cast<char*>(ctx.io.ptr);
    pack_sources(p.sources, p.wavenums, p.attens, n, p.source_type, reinterpret_cast<float*>(h));
    memcpy(h + o_foci, p.foci, foci_b);
    memcpy(h + o_amps, p.amps, amps_b);
    if (!hip_ok(hipMemcpyAsync(d, h, o_ph, hipMemcpyHostToDevice, ctx.stream), "H2D")) return false;
    const uint32_t offs[2] = {0, (uint32_t)m};
    // K_MAX = 200 (lm.rs:20, gauss_newton.rs:19), 10 000 (gradient_descent.rs:18)
    const uint32_t k_max = algo == kNlsGD ? 10000 : 200;
    if (lm_solve_batch(&ctx, reinterpret_cast<const float*>(d), (uint32_t)n, (uint32_t)np, p.source_type,
                       reinterpret_cast<const float*>(d + o_foci), reinterpret_cast<const float*>(d + o_amps), offs,
                       1, k_max, reinterpret_cast<float*>(d + o_ph), nullptr, ctx.stream, algo) != AFH_OK)
        return false;
    if (!hip_ok(hipMemcpyAsync(h + o_ph, d + o_ph, out_b, hipMemcpyDeviceToHost, ctx.stream), "D2H")) return false;
    if (!hip_ok(hipStreamSynchronize(ctx.stream), "hipStreamSynchronize")) return false;
    memcpy(phase_out, h + o_ph, out_b);
    return true;
}

bool apo_solve_single(const HoloProblem& p, float lambda, float* amp_out, float* phase_out) {
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
    // K_MAX = 200 (apo.rs:18)
    if (apo_solve_batch(&ctx, reinterpret_cast<const float*>(d), (uint32_t)n, p.source_type,
                        reinterpret_cast<const float*>(d + o_foci), reinterpret_cast<const float*>(d + o_amps), offs,
                        1, (double)lambda, 200, reinterpret_cast<float*>(d + o_amp), reinterpret_cast<float*>(d + o_ph),
                        nullptr, ctx.stream) != AFH_OK)
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

static int32_t nls_entry(const char* who, void* ctx, const float* packed_sources_dev, uint64_t n_packed,
                         uint64_t n_sources, int32_t source_type, const float* foci_dev, const float* amps_dev,
                         const uint32_t* offsets_host, uint32_t batch, int32_t algo, uint32_t k_max,
                         float* out_phase_dev, uint32_t* out_iters_dev, void* stream) {
    clear_error();
    Context* c = static_cast<Context*>(ctx);
    if (!c || !packed_sources_dev || !foci_dev || !amps_dev || !offsets_host || !out_phase_dev ||
        n_packed % AFH_SOURCE_PAD != 0 || n_packed < n_sources || n_sources > 0xFFFFFFFFull ||
        (source_type != kSphere && source_type != kT4010A1) || offsets_host[0] != 0 || algo < kNlsLM ||
        algo > kNlsGD) {
        set_error(std::string(who) + ": bad argument");
        return AFH_ERR_ARG;
    }
    std::lock_guard<std::mutex> g(c->mu);
    return lm_solve_batch(c, packed_sources_dev, (uint32_t)n_sources, (uint32_t)n_packed, source_type, foci_dev,
                          amps_dev, offsets_host, batch, k_max, out_phase_dev, out_iters_dev, (hipStream_t)stream,
                          algo);
}

extern "C" int32_t AFH_LMSolve(void* ctx, const float* packed_sources_dev, uint64_t n_packed, uint64_t n_sources,
                               int32_t source_type, const float* foci_dev, const float* amps_dev,
                               const uint32_t* offsets_host, uint32_t batch, uint32_t k_max, float* out_phase_dev,
                               uint32_t* out_iters_dev, void* stream) {
    return nls_entry("AFH_LMSolve", ctx, packed_sources_dev, n_packed, n_sources, source_type, foci_dev, amps_dev,
                     offsets_host, batch, kNlsLM, k_max, out_phase_dev, out_iters_dev, stream);
}

extern "C" int32_t AFH_NLSolve(void* ctx, const float* packed_sources_dev, uint64_t n_packed, uint64_t n_sources,
                               int32_t source_type, const float* foci_dev, const float* amps_dev,
                               const uint32_t* offsets_host, uint32_t batch, int32_t algo, uint32_t k_max,
                               float* out_phase_dev, uint32_t* out_iters_dev, void* stream) {
    return nls_entry("AFH_NLSolve", ctx, packed_sources_dev, n_packed, n_sources, source_type, foci_dev, amps_dev,
                     offsets_host, batch, algo, k_max, out_phase_dev, out_iters_dev, stream);
}

extern "C" int32_t AFH_APOSolve(void* ctx, const float* packed_sources_dev, uint64_t n_packed, uint64_t n_sources,
                                int32_t source_type, const float* foci_dev, const float* amps_dev,
                                const uint32_t* offsets_host, uint32_t batch, float lambda, uint32_t k_max,
                                float* out_amp_dev, float* out_phase_dev, uint32_t* out_iters_dev, void* stream) {
    clear_error();
    Context* c = static_cast<Context*>(ctx);
    if (!c || !packed_sources_dev || !foci_dev || !amps_dev || !offsets_host || !out_amp_dev || !out_phase_dev ||
        n_packed % AFH_SOURCE_PAD != 0 || n_packed < n_sources || n_sources > 0xFFFFFFFFull ||
        (source_type != kSphere && source_type != kT4010A1) || offsets_host[0] != 0) {
        set_error("AFH_APOSolve: bad argument");
        return AFH_ERR_ARG;
    }
    std::lock_guard<std::mutex> g(c->mu);
    return apo_solve_batch(c, packed_sources_dev, (uint32_t)n_sources, source_type, foci_dev, amps_dev, offsets_host,
                           batch, (double)lambda, k_max, out_amp_dev, out_phase_dev, out_iters_dev,
                           (hipStream_t)stream);
}
