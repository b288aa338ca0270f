// Hologram solvers on the GPU: GS-PAT, GS and naive linear synthesis, batched
// over instances that share one transducer array.
//
// Reference algorithms (acoustic-field-optimizer/src/multiple_foci/):
//   macros.rs:32-53            G_ij = term(source j -> focus i), phases forced to 0
//   linear_synthesis/gs_pat.rs:38-81
//       d_j = sum_i |G_ij|; B = (conj(G_ij) / d_j^2)^T; R = G B;
//       gamma = R p0; 100 x {p = gamma / |gamma| p0; gamma = R p};
//       p = gamma / |gamma|^2 p0^2; q = B p; amp = |q|, phase = arg q
//   linear_synthesis/gs.rs:36-77
//       100 x {gamma = G q; p = gamma / |gamma| p0; xi = G^H p; q = xi / |xi| q0}
//   linear_synthesis/naive.rs:32-51   q = G^H p0
//   combination/greedy.rs:23-83       per source in order, the phase candidate
//       minimising sum_j |a_j - |G_ji c + cache_j|| (first minimum), cache += G_:i c
//
// GS-PAT formulation here: H_ij = G_ij / d_j (|H_ij| <= 1, the per-column
// scaling), so R = H H^H. Three kernels per bucket of instances with the same
// padded focus count MP (16, 32, 64, 128):
//   A  generates G in registers from the exact coordinate difference of
//      source and focus (add_terms_exact), reduces d_j and accumulates R:
//        - mixed modes: H = s G / d_j split into FP16 hi + lo, R += hi hi^H +
//          hi lo^H + lo hi^H on v_mfma_f32_16x16x32_f16 with FP32
//          accumulation, fused with the generation (gspat_wave for MP <= 32:
//          each wave takes every fourth chunk of the range for all MP foci;
//          gspat_pair for MP = 64: two waves share a chunk, 32 foci each;
//          gspat_fused for MP = 128: one wave per 16 foci);
//        - FP32: gspat_gen stores FP32 G and d_j, gspat_r32 forms R on
//          v_mfma_f32_16x16x4f32 (MP >= 32); gspat_r_fp32 on the VALU (MP = 16);
//      and keeps kernel C's copy: FP16 hi of H (MIXED) or FP32 G (the others),
//      rows < M only;
//   B  runs the iterations with R held in registers, one workgroup per
//      instance, reductions by fixed-order lane shuffles (gspat_iter);
//   C  computes q = B p in FP32 from that copy and d_j (reading it back costs
//      less than generating G again); in MIXED_CLOSING a first pass redoes the
//      last gamma = R p exactly as G (B p) in FP32 (gspat_q).
// The buckets' kernels A run in turn; B and C run once for every bucket up to
// MP = 64 (gspat_iter_set, gspat_q_set), the MP = 128 bucket on its own.
// A and C run a.split workgroups per instance over ranges of source chunks
// (a function of the array and the bucket); A's partial R and C's partial
// gamma are summed in a fixed order.
// Every reduction is per instance and in a fixed order, so an instance's result
// does not depend on the other instances of the batch.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <type_traits>
#include <vector>

#include "afh_holo.hpp"
#include "afh_host.hpp"
#include "afh_internal.hpp"

namespace afh {

constexpr int kHB = 256;  // threads per workgroup
constexpr int kKC = 32;   // G columns per chunk
constexpr int kMaxMP = 128;  // largest padded focus count (bucket)

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef uint32_t uintx4 __attribute__((ext_vector_type(4)));

struct HoloArgs {
    const float* sources;  // packed, AFH_SOURCE_STRIDE floats each
    uint32_t n;            // sources (outputs per instance)
    uint32_t n_packed;     // multiple of kKC
    const float* foci;     // 3 floats per focus, all instances
    const float* amps;     // target amplitude per focus
    const uint32_t* offsets;  // batch + 1
    const uint32_t* inst;     // instance ids of this bucket
    const float4* staged;     // SrcF of every packed source (3 float4 each), phases zeroed
    const DirCell* cells;
    float2* R;   // bucket: split partial MP * MP matrices per instance (row-major)
    uint32_t split;  // GS-PAT kernel A workgroups per instance (source chunk ranges)
    float2* p;   // bucket: MP per instance
    float2* G;   // GS, GS-PAT: MP x n_packed per instance (row-major), bucket-relative
    float* d;    // GS-PAT: d_j = sum_i |G_ij|, n_packed per instance, bucket-relative
    float2* q;   // GS: n_packed per instance
    float2* gam; // GS-PAT MIXED_CLOSING: partial gamma (MP) per kernel C workgroup
    float* out_amp;    // batch x n
    float* out_phase;  // batch x n
};

// Source j in the FAST-mode operand form (SrcF), from the staged copy.
__device__ __forceinline__ SrcF load_staged(const float4* staged, uint32_t j) {
    const float4 q0 = staged[3 * (uint64_t)j], q1 = staged[3 * (uint64_t)j + 1], q2 = staged[3 * (uint64_t)j + 2];
    return {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
}

// Workgroup barrier that orders LDS accesses only. __syncthreads() also waits
// for every outstanding global store (its release fence covers global memory),
// which stalls kernel A on its G tile stores once per chunk; the barriers of
// the GS-PAT kernels only guard LDS data (results in global memory are read by
// later kernels).
__device__ __forceinline__ void lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Stages every packed source once per solve: the optimizers zero every phase
// before building G (gs_pat.rs:39-41, gs.rs:37-39, naive.rs:34-36), so the
// staged phase is 0 (1/2 turn for a negative amplitude, see stage_fast).
__global__ __launch_bounds__(kHB) void stage_holo_sources(const float* __restrict__ sources, uint32_t np,
                                                          float4* __restrict__ staged) {
    const uint32_t j = blockIdx.x * kHB + threadIdx.x;
    if (j >= np) return;
    const float4* p = reinterpret_cast<const float4*>(sources) + 3 * (uint64_t)j;
    const float4 q0 = p[0], q1 = p[1], q2 = p[2];
    Src s = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
    s.ph = 0.0f;
    const SrcF f = stage_fast(s);
    staged[3 * (uint64_t)j] = make_float4(f.x, f.y, f.z, f.la);
    staged[3 * (uint64_t)j + 1] = make_float4(f.nx, f.ny, f.nz, f.ph);
    staged[3 * (uint64_t)j + 2] = make_float4(f.kh, f.kl, f.nal, f.kh2);
}

// The FAST-mode cell table in LDS as two float4 arrays (see DirCell).
struct CellsLds {
    float4 p[kDirCells], q[kDirCells];
};

__device__ __forceinline__ void stage_cells_lds(const DirCell* __restrict__ cells, CellsLds* l) {
    const float4* g = reinterpret_cast<const float4*>(cells);
    for (int j = threadIdx.x; j < kDirCells; j += kHB) {
        l->p[j] = g[2 * j];
        l->q[j] = g[2 * j + 1];
    }
}

// Per-instance foci in LDS, positions split as by split3; rows >= M are
// padding.
template <int MP>
struct Foci {
    float xh[MP], yh[MP], zh[MP], xl[MP], yl[MP], zl[MP], a[MP], v[MP];  // v: 1 for a focus, 0 for padding
};

template <int SRC, int MP>
__device__ __forceinline__ void load_instance(const HoloArgs& a, uint32_t b, Foci<MP>& f, CellsLds* cells,
                                              uint32_t& M, uint32_t& base) {
    base = a.offsets[b];
    M = a.offsets[b + 1] - base;
    for (int i = threadIdx.x; i < MP; i += kHB) {
        const bool v = i < (int)M;
        const Split3 q = split3(v ? a.foci[3 * (uint64_t)(base + i)] : 0.f,
                                v ? a.foci[3 * (uint64_t)(base + i) + 1] : 0.f,
                                v ? a.foci[3 * (uint64_t)(base + i) + 2] : 1.f);
        f.xh[i] = q.xh;
        f.yh[i] = q.yh;
        f.zh[i] = q.zh;
        f.xl[i] = q.xl;
        f.yl[i] = q.yl;
        f.zl[i] = q.zl;
        f.a[i] = v ? a.amps[base + i] : 0.f;
        f.v[i] = v ? 1.f : 0.f;
    }
    if (SRC == kT4010A1) stage_cells_lds(a.cells, cells);
}

// G for the thread's column (source s, staged) at rows (t >> 5) + 8 r,
// r < MP / 8 (zero for padding rows and padding columns), RT rows at a time
// through the staged term; steps whose rows are all padding (>= M) are set to
// zero without evaluating the terms. Kernel A takes RT = 2: with 4 the MP = 64
// kernels reach their 168-register bound and ran slower.
template <int SRC, int MP, int RT>
__device__ __forceinline__ void g_tile(const SrcF& s, const Split3& ss, const Foci<MP>& f, const CellsLds* cells,
                                       uint32_t M, float2 (&g)[MP / 8]) {
    static_assert(RT <= MP / 8 && (MP / 8) % RT == 0, "rows per step");
    const int t = threadIdx.x;
    const int grp = t >> 5;
#pragma unroll
    for (int r = 0; r < MP / 8; r += RT) {
        const int i0 = grp + 8 * r;
        if (i0 >= (int)M) {
            // every row of the step is padding (the rest are above i0): zero
            // without computing the terms
#pragma unroll
            for (int u = 0; u < RT; ++u) g[r + u] = make_float2(0.f, 0.f);
            continue;
        }
        Split3 pt[RT];
        float w[RT], re[RT], im[RT];
#pragma unroll
        for (int u = 0; u < RT; ++u) {
            const int i = i0 + 8 * u;
            pt[u] = {f.xh[i], f.yh[i], f.zh[i], f.xl[i], f.yl[i], f.zl[i]};
            w[u] = f.v[i];
            re[u] = 0.f;
            im[u] = 0.f;
        }
        add_terms_exact<SRC, RT>(s, ss, pt, cells->p, cells->q, re, im, w, nullptr);
#pragma unroll
        for (int u = 0; u < RT; ++u) g[r + u] = make_float2(re[u], im[u]);
    }
}

// Kernel A's staged sources, kSB chunks at a time in LDS (SoA: v[q][j] is
// float4 q of source j of the block), loaded one block ahead. vmcnt counts
// loads and stores together and the compiler waits for zero, so a load issued
// between G tile stores would make every chunk wait for the previous chunk's
// stores; with the block loads issued kSB chunks before their data is written
// to LDS, that wait falls once per block.
constexpr int kSB = 8;
static_assert(kSB * kKC * 3 == 3 * kHB, "one block is three float4 per thread");

struct SrcBlock {
    float4 v[3][kSB * kKC];
};

// The block of kSB chunks starting at chunk c (staged sources 32 c ...).
__device__ __forceinline__ void src_block_fetch(const HoloArgs& a, uint32_t c, float4 (&pre)[3]) {
    const uint64_t base = (uint64_t)c * kKC * 3, lim = (uint64_t)a.n_packed * 3;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const uint64_t e = base + threadIdx.x + q * kHB;
        pre[q] = e < lim ? a.staged[e] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

__device__ __forceinline__ void src_block_put(SrcBlock& sb, const float4 (&pre)[3]) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const int e = threadIdx.x + q * kHB;
        sb.v[e % 3][e / 3] = pre[q];
    }
}

__device__ __forceinline__ SrcF src_from_block(const SrcBlock& sb, int cb) {
    const int j = cb * kKC + (threadIdx.x & 31);
    const float4 q0 = sb.v[0][j], q1 = sb.v[1][j], q2 = sb.v[2][j];
    return {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
}

// 1 / d of the thread's column j0 + (t & 31), d = sum_i |G_ij| summed in a
// fixed order (each row group's partial, then the 8 groups in order); 0 for
// columns beyond n. Threads t < kKC write d to d_out[j0 + t]. Every thread sums
// the 8 partials of its own column, so one barrier suffices; s_part is written
// again only in the next chunk, after the chunk's tile barrier, which every
// thread reaches after its reads here.
template <int MP>
__device__ __forceinline__ float column_inv(const float2 (&g)[MP / 8], uint32_t j0, uint32_t n,
                                            float (*s_part)[kKC], float* d_out) {
    const int t = threadIdx.x;
    float part = 0.f;
#pragma unroll
    for (int r = 0; r < MP / 8; ++r) part += sqrtf(g[r].x * g[r].x + g[r].y * g[r].y);
    s_part[t >> 5][t & 31] = part;
    lds_sync();
    float d = 0.f;
#pragma unroll
    for (int q = 0; q < 8; ++q) d += s_part[q][t & 31];
    if (t < kKC) d_out[j0 + t] = d;
    // a real column with d = 0 gives NaN, as conj(G) / 0 does in the reference
    return (j0 + (t & 31) < n) ? 1.0f / d : 0.f;
}

// Kernel A's copy of the G tile for kernel C: rows < M only (rows >= M are
// zero and never read back), coalesced over the 32 columns of the chunk.
template <int MP>
__device__ __forceinline__ void store_g_tile(const HoloArgs& a, uint32_t k, uint32_t j0, uint32_t M,
                                             const float2 (&g)[MP / 8]) {
    const int t = threadIdx.x;
    float2* G = a.G + (uint64_t)k * MP * a.n_packed + j0 + (t & 31);
#pragma unroll
    for (int r = 0; r < MP / 8; ++r) {
        const int i = (t >> 5) + 8 * r;
        if (i < (int)M) {
            // streamed: read once, by kernel C
            __builtin_nontemporal_store(g[r].x, &G[(uint64_t)i * a.n_packed].x);
            __builtin_nontemporal_store(g[r].y, &G[(uint64_t)i * a.n_packed].y);
        }
    }
}

// Kernel C's stored inputs of one chunk as loaded: the G tile (FP32 G, or
// H16 words: the FP16 copy of H = s G_ij / d_j that the fused kernel A stores
// in MIXED mode) and, for threads t < kKC, d of column j0 + t.
template <int MP, bool H16>
struct StoredChunk {
    std::conditional_t<H16, uint32_t, float2> g[MP / 8];
    float d;
};

// cache-policy operand of a buffer load: bit 1 is the non-temporal hint (nt)
constexpr int kNonTemporal = 2;

template <int MP, bool H16>
__device__ __forceinline__ void load_chunk(const HoloArgs& a, __amdgpu_buffer_rsrc_t gr, uint32_t k, uint32_t j0,
                                           uint32_t M, StoredChunk<MP, H16>& c) {
    const int t = threadIdx.x;
    const uint32_t col = j0 + (t & 31);
    if constexpr (H16) {
        // every load unconditional: the descriptor ends at row M (rows >= M
        // read as 0 without a memory access), and d is read by every thread
        // (only t < kKC use it); with a select per load and the ring rotated
        // by copies, the compiler waited for each chunk's loads at once
#pragma unroll
        for (int r = 0; r < MP / 8; ++r) {
            const uint32_t e = (uint32_t)((t >> 5) + 8 * r) * a.n_packed + col;
            c.g[r] = __builtin_amdgcn_raw_buffer_load_b32(gr, (int)(e * 4), 0, kNonTemporal);
        }
        c.d = a.d[(uint64_t)k * a.n_packed + col];
    } else {
        // FP32 G, one chunk ahead: rows >= M are skipped (measured faster here
        // than the bounded buffer loads, which still issue those rows)
        const float2* G = a.G + (uint64_t)k * MP * a.n_packed + col;
#pragma unroll
        for (int r = 0; r < MP / 8; ++r) {
            const int i = (t >> 5) + 8 * r;
            const float2* g = G + (uint64_t)i * a.n_packed;
            c.g[r] = i < (int)M ? make_float2(__builtin_nontemporal_load(&g->x), __builtin_nontemporal_load(&g->y))
                                : make_float2(0.f, 0.f);
        }
        c.d = t < kKC ? a.d[(uint64_t)k * a.n_packed + col] : 0.f;
    }
}

// The chunk's G tile in FP32 (H16 words widened).
template <int MP, bool H16>
__device__ __forceinline__ void chunk_g(const StoredChunk<MP, H16>& c, float2 (&g)[MP / 8]) {
#pragma unroll
    for (int r = 0; r < MP / 8; ++r) {
        if constexpr (H16)
            g[r] = make_float2((float)__builtin_bit_cast(_Float16, (uint16_t)(c.g[r] & 0xffffu)),
                               (float)__builtin_bit_cast(_Float16, (uint16_t)(c.g[r] >> 16)));
        else
            g[r] = c.g[r];
    }
}

// ---------------------------------------------------------------- kernel A (FP32)
template <int SRC, int MP>
__global__ __launch_bounds__(kHB) void gspat_r_fp32(HoloArgs a) {
    __shared__ std::conditional_t<SRC == kT4010A1, CellsLds, float> s_cells_mem;
    CellsLds* s_cells = reinterpret_cast<CellsLds*>(&s_cells_mem);
    __shared__ Foci<MP> f;
    __shared__ float s_part[8][kKC];
    __shared__ float2 s_h[MP][kKC + 1];
    __shared__ SrcBlock s_src;
    const int t = threadIdx.x;
    const uint32_t k = blockIdx.x / a.split, part = blockIdx.x % a.split, b = a.inst[k];
    const uint32_t n_chunks = a.n_packed / kKC;
    const uint32_t c_begin = part * n_chunks / a.split, c_end = (part + 1) * n_chunks / a.split;
    uint32_t M, base;
    load_instance<SRC, MP>(a, b, f, s_cells, M, base);
    __syncthreads();
    constexpr int BS = MP / 16;
    const int rb = (t >> 4) * BS, cb = (t & 15) * BS;
    float2 acc[BS][BS];
#pragma unroll
    for (int x = 0; x < BS; ++x)
#pragma unroll
        for (int y = 0; y < BS; ++y) acc[x][y] = make_float2(0.f, 0.f);

    float* d_out = a.d + (uint64_t)k * a.n_packed;
    float4 pre[3];
    src_block_fetch(a, c_begin, pre);
    for (uint32_t chunk = c_begin; chunk < c_end; ++chunk) {
        const uint32_t j0 = chunk * kKC, in_blk = (chunk - c_begin) % kSB;
        if (in_blk == 0) {
            // every wave is past its last read of the previous block (the
            // previous chunk's tile barrier)
            src_block_put(s_src, pre);
            lds_sync();
            if (chunk + kSB < c_end) src_block_fetch(a, chunk + kSB, pre);
        }
        const SrcF sc = src_from_block(s_src, in_blk);
        float2 g[MP / 8];
        g_tile<SRC, MP, 2>(sc, split3(sc.x, sc.y, sc.z), f, s_cells, M, g);
        store_g_tile<MP>(a, k, j0, M, g);
        const float inv = column_inv<MP>(g, j0, a.n, s_part, d_out);
#pragma unroll
        for (int r = 0; r < MP / 8; ++r) s_h[(t >> 5) + 8 * r][t & 31] = make_float2(g[r].x * inv, g[r].y * inv);
        // the tile barrier: s_h is written again only after the next chunk's
        // column barrier, which every thread reaches after the products below
        lds_sync();
        for (int jj = 0; jj < kKC; ++jj) {
            float2 hr[BS], hc[BS];
#pragma unroll
            for (int x = 0; x < BS; ++x) hr[x] = s_h[rb + x][jj];
#pragma unroll
            for (int y = 0; y < BS; ++y) hc[y] = s_h[cb + y][jj];
#pragma unroll
            for (int x = 0; x < BS; ++x)
#pragma unroll
                for (int y = 0; y < BS; ++y) {
                    // R_ik += H_ij conj(H_kj)
                    acc[x][y].x = fmaf(hr[x].x, hc[y].x, fmaf(hr[x].y, hc[y].y, acc[x][y].x));
                    acc[x][y].y = fmaf(hr[x].y, hc[y].x, fmaf(-hr[x].x, hc[y].y, acc[x][y].y));
                }
        }
    }
    float2* R = a.R + (uint64_t)blockIdx.x * MP * MP;  // partial `part` of instance k
#pragma unroll
    for (int x = 0; x < BS; ++x)
#pragma unroll
        for (int y = 0; y < BS; ++y) R[(rb + x) * MP + cb + y] = acc[x][y];
}

// ---------------------------------------------------------------- FP32 mode: kernels A1, A2
// A1: G of every instance of the bucket, one thread per source column j and
// the instance's foci in turn (four at a time through the exact-difference
// term), stored for rows < M as FP32 G_ij (8 bytes, the layout kernel C
// reads), non-temporal, and d_j = sum_i |G_ij| over i in order (FP32, from
// each term's modulus).
template <int SRC, int MP>
__global__ __launch_bounds__(kHB) void gspat_gen(HoloArgs a) {
    __shared__ std::conditional_t<SRC == kT4010A1, CellsLds, float> s_cells_mem;
    CellsLds* s_cells = reinterpret_cast<CellsLds*>(&s_cells_mem);
    __shared__ float4 s_fh[MP], s_fl[MP];  // the foci split as by split3
    const uint32_t k = blockIdx.y, b = a.inst[k];
    const uint32_t base = a.offsets[b], M = a.offsets[b + 1] - base;
    for (int i = threadIdx.x; i < MP; i += kHB) {
        const uint64_t q = 3 * (uint64_t)(base + i);
        const Split3 p = i < (int)M ? split3(a.foci[q], a.foci[q + 1], a.foci[q + 2]) : split3(0.f, 0.f, 1.f);
        s_fh[i] = make_float4(p.xh, p.yh, p.zh, 0.f);
        s_fl[i] = make_float4(p.xl, p.yl, p.zl, 0.f);
    }
    if (SRC == kT4010A1) stage_cells_lds(a.cells, s_cells);
    __syncthreads();
    const uint32_t j = blockIdx.x * kHB + threadIdx.x;
    if (j >= a.n_packed) return;
    const SrcF src = load_staged(a.staged, j);
    const Split3 ss = split3(src.x, src.y, src.z);
    float2* G32 = a.G + (uint64_t)k * MP * a.n_packed + j;
    float d = 0.f;
    for (uint32_t i = 0; i < M; i += 4) {
        // i + 3 < MP (i is a multiple of 4 below M <= MP, MP a multiple of
        // 16); rows from M on are padding foci, evaluated and not stored
        Split3 pt[4];
        float re[4], im[4], am[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const float4 h = s_fh[i + u], lo = s_fl[i + u];
            pt[u] = {h.x, h.y, h.z, lo.x, lo.y, lo.z};
            re[u] = im[u] = 0.f;
        }
        add_terms_exact<SRC, 4>(src, ss, pt, s_cells->p, s_cells->q, re, im, nullptr, am);
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (i + u < M) {
                d += fabsf(am[u]);
                float2* g = G32 + (uint64_t)(i + u) * a.n_packed;
                __builtin_nontemporal_store(re[u], &g->x);
                __builtin_nontemporal_store(im[u], &g->y);
            }
    }
    a.d[(uint64_t)k * a.n_packed + j] = d;
}

// A2 for FP32 mode (MP >= 32): R = H H^H over the split's source chunk range
// from A1's FP32 G, H_ij = G_ij / d_j in FP32 (an LDS tile, double-buffered by
// chunk parity as in A2), on FP32 MFMA (v_mfma_f32_16x16x4f32: lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16] and receives D[4 (l / 16) + e][l %
// 16]). R is Hermitian, so only the tiles (I, J) with I <= J are formed, dealt
// to the 4 waves in turn, and each with three real products (P1 = Hr Hr^T, P2 =
// Hi Hi^T, P3 = (Hr + Hi)(Hr - Hi)^T; Re R = P1 + P2, Im R = P3 - P1 + P2, the
// sums in FP32); the tile (J, I) is written as the conjugate transpose. The sum
// over columns is taken in a permuted order so that each lane reads 4 adjacent
// columns at once: MFMA step (m, q) gives lane group c = l / 16 the column
// 16 m + 4 c + q, for both operands (rows 36 floats apart: the 128-bit reads of
// 8 or 16 lanes fall in distinct banks of a 32- or 64-bank LDS).
template <int MP>
__global__ __launch_bounds__(kHB) void gspat_r32(HoloArgs a) {
    static_assert(MP >= 32, "one 16x16 tile per wave at least");
    constexpr int T = MP / 16;
    constexpr int NT = T * (T + 1) / 2;  // upper-triangle tiles
    constexpr int TPW = (NT + 3) / 4;
    constexpr int HS = kKC + 4;
    __shared__ __attribute__((aligned(16))) float s_hr[2][MP * HS];
    __shared__ __attribute__((aligned(16))) float s_hi[2][MP * HS];
    const int t = threadIdx.x;
    const int w = t >> 6, l = t & 63;
    const uint32_t k = blockIdx.x / a.split, part = blockIdx.x % a.split, b = a.inst[k];
    const uint32_t n_chunks = a.n_packed / kKC;
    const uint32_t c_begin = part * n_chunks / a.split, c_end = (part + 1) * n_chunks / a.split;
    const uint32_t M = a.offsets[b + 1] - a.offsets[b];
    const float2* G = a.G + (uint64_t)k * MP * a.n_packed + (t & 31);
    const float* dk = a.d + (uint64_t)k * a.n_packed + (t & 31);
    // the wave's tiles: tau = w + 4 s in row-major order of the upper triangle;
    // a wave with fewer than TPW tiles forms tile 0 again in the last slot, so
    // every wave runs the same straight-line MFMA sequence, and does not store it
    int tI[TPW], tJ[TPW];
#pragma unroll
    for (int s = 0; s < TPW; ++s) {
        int r = w + 4 * s < NT ? w + 4 * s : 0, I = 0;
        while (r >= T - I) {
            r -= T - I;
            ++I;
        }
        tI[s] = I;
        tJ[s] = I + r;
    }
    floatx4 p1[TPW], p2[TPW], p3[TPW];
#pragma unroll
    for (int s = 0; s < TPW; ++s) {
        p1[s] = floatx4{0.f, 0.f, 0.f, 0.f};
        p2[s] = p1[s];
        p3[s] = p1[s];
    }
    // the thread's G (rows (t >> 5) + 8 r, column j0 + t % 32) and d_j, one chunk ahead
    float2 ng[MP / 8];
    float nd = 0.f;
    auto fetch = [&](uint32_t j0) {
#pragma unroll
        for (int r = 0; r < MP / 8; ++r) {
            const int i = (t >> 5) + 8 * r;
            const float2* g = G + (uint64_t)i * a.n_packed + j0;
            ng[r] = i < (int)M ? make_float2(__builtin_nontemporal_load(&g->x), __builtin_nontemporal_load(&g->y))
                               : make_float2(0.f, 0.f);
        }
        nd = dk[j0];
    };
    if (c_begin < c_end) fetch(c_begin * kKC);
    for (uint32_t chunk = c_begin; chunk < c_end; ++chunk) {
        const uint32_t j0 = chunk * kKC;
        const int pb = (chunk - c_begin) & 1;
        float2 g[MP / 8];
#pragma unroll
        for (int r = 0; r < MP / 8; ++r) g[r] = ng[r];
        const float dj = nd;
        if (chunk + 1 < c_end) fetch(j0 + kKC);
        // a real column with d = 0 gives NaN, as conj(G) / 0 does in the reference
        const float inv = (j0 + (t & 31) < a.n) ? 1.0f / dj : 0.f;
#pragma unroll
        for (int r = 0; r < MP / 8; ++r) {
            const int i = (t >> 5) + 8 * r;
            s_hr[pb][i * HS + (t & 31)] = g[r].x * inv;
            s_hi[pb][i * HS + (t & 31)] = g[r].y * inv;
        }
        lds_sync();
#pragma unroll
        for (int m = 0; m < kKC / 16; ++m) {
            const int cb = 16 * m + 4 * (l >> 4);
#pragma unroll
            for (int s = 0; s < TPW; ++s) {
                const int arow = tI[s] * 16 + (l & 15);
                const int brow = tJ[s] * 16 + (l & 15);
                const float4 ar = *reinterpret_cast<const float4*>(&s_hr[pb][arow * HS + cb]);
                const float4 ai = *reinterpret_cast<const float4*>(&s_hi[pb][arow * HS + cb]);
                const float4 br = *reinterpret_cast<const float4*>(&s_hr[pb][brow * HS + cb]);
                const float4 bi = *reinterpret_cast<const float4*>(&s_hi[pb][brow * HS + cb]);
                const float xr[4] = {ar.x, ar.y, ar.z, ar.w}, xi[4] = {ai.x, ai.y, ai.z, ai.w};
                const float yr[4] = {br.x, br.y, br.z, br.w}, yi[4] = {bi.x, bi.y, bi.z, bi.w};
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    p1[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(xr[q], yr[q], p1[s], 0, 0, 0);
                    p2[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(xi[q], yi[q], p2[s], 0, 0, 0);
                    p3[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(xr[q] + xi[q], yr[q] - yi[q], p3[s], 0, 0, 0);
                }
            }
        }
    }
    float2* R = a.R + (uint64_t)blockIdx.x * MP * MP;  // partial `part` of instance k
#pragma unroll
    for (int s = 0; s < TPW; ++s) {
        if (w + 4 * s >= NT) continue;  // wave-uniform
        const int col = tJ[s] * 16 + (l & 15);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int row = tI[s] * 16 + 4 * (l >> 4) + e;
            const float re = p1[s][e] + p2[s][e], im = p3[s][e] - p1[s][e] + p2[s][e];
            R[row * MP + col] = make_float2(re, im);
            if (tI[s] != tJ[s]) R[col * MP + row] = make_float2(re, -im);
        }
    }
}

// ---------------------------------------------------------------- kernel B
// v summed over the TPR consecutive lanes of a row group (TPR = 1, 2, 4, 8 or
// 16) with DPP lane moves instead of LDS-crossbar shuffles: lane i ^ 8 (row
// rotate by 8), 7 - i within 8 lanes (row half mirror), i ^ 2 and i ^ 1 (quad
// permutes). Each step pairs lanes symmetrically, so every lane of the group
// ends with the same bits.
#define AFH_DPP_ADD(v, ctrl) \
    ((v) + __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, (v)), (ctrl), 0xf, 0xf, false)))
template <int TPR>
__device__ __forceinline__ float row_group_sum(float v) {
    static_assert(TPR == 1 || TPR == 2 || TPR == 4 || TPR == 8 || TPR == 16, "row group within a DPP row");
    if constexpr (TPR >= 16) v = AFH_DPP_ADD(v, 0x128);  // row_ror:8
    if constexpr (TPR >= 8) v = AFH_DPP_ADD(v, 0x141);   // row_half_mirror
    if constexpr (TPR >= 4) v = AFH_DPP_ADD(v, 0x4e);    // quad_perm [2, 3, 0, 1]
    if constexpr (TPR >= 2) v = AFH_DPP_ADD(v, 0xb1);    // quad_perm [1, 0, 3, 2]
    return v;
}
#undef AFH_DPP_ADD

// gamma = R p0; iters x {p = gamma / |gamma| p0; gamma = R p}. Writes the
// last p (closing) or p = gamma / |gamma|^2 p0^2 to a.p.
// Workgroup blk of the bucket (one instance) with the p vectors in s_p.
template <int MP>
__device__ __forceinline__ void gspat_iter_run(const HoloArgs& a, uint32_t blk, uint32_t iters, int closing,
                                               float2 (*s_p)[kMaxMP]) {
    constexpr int TPR = kHB / MP;  // lanes per row (consecutive, same wave)
    constexpr int CPT = MP / TPR;  // columns per lane
    static_assert(TPR >= 1 && TPR <= 64, "row group within a wave");
    const int t = threadIdx.x;
    const int row = t / TPR, h = t % TPR;
    const uint32_t k = blk, b = a.inst[k];
    const uint32_t base = a.offsets[b], M = a.offsets[b + 1] - base;
    // R = sum of kernel A's partials over source chunk ranges, in order
    const float2* R = a.R + (uint64_t)k * a.split * MP * MP + (uint64_t)row * MP + h * CPT;
    float2 r[CPT];
#pragma unroll
    for (int c = 0; c < CPT; ++c) r[c] = R[c];
    for (uint32_t sp = 1; sp < a.split; ++sp)
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
            const float2 v = R[(uint64_t)sp * MP * MP + c];
            r[c].x += v.x;
            r[c].y += v.y;
        }
    const bool valid = row < (int)M;
    const float p0 = valid ? a.amps[base + row] : 0.f;
    if (h == 0) s_p[0][row] = make_float2(p0, 0.f);
    __syncthreads();
    int cur = 0;
    auto matvec = [&](int buf) {
        // two accumulators (even and odd columns) halve the FMA chain
        float2 a0 = make_float2(0.f, 0.f), a1 = make_float2(0.f, 0.f);
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
            const float2 pv = s_p[buf][h * CPT + c];
            float2& acc = (c & 1) ? a1 : a0;
            acc.x = fmaf(r[c].x, pv.x, fmaf(-r[c].y, pv.y, acc.x));
            acc.y = fmaf(r[c].x, pv.y, fmaf(r[c].y, pv.x, acc.y));
        }
        return make_float2(row_group_sum<TPR>(a0.x + a1.x), row_group_sum<TPR>(a0.y + a1.y));
    };
    float2 gam = matvec(cur);
    for (uint32_t it = 0; it < iters; ++it) {
        // every lane of the row group holds the same gamma; one writes p
        const float inv = p0 * __builtin_amdgcn_rsqf(fmaf(gam.x, gam.x, gam.y * gam.y));
        if (h == 0) s_p[cur ^ 1][row] = valid ? make_float2(gam.x * inv, gam.y * inv) : make_float2(0.f, 0.f);
        lds_sync();
        cur ^= 1;
        gam = matvec(cur);
    }
    if (h == 0) {
        float2 out;
        if (closing) {
            out = s_p[cur][row];
        } else {
            const float inv = p0 * p0 / (gam.x * gam.x + gam.y * gam.y);
            out = valid ? make_float2(gam.x * inv, gam.y * inv) : make_float2(0.f, 0.f);
        }
        a.p[(uint64_t)k * MP + row] = out;
    }
}

template <int MP>
__global__ __launch_bounds__(kHB) void gspat_iter(HoloArgs a, uint32_t iters, int closing) {
    __shared__ float2 s_p[2][kMaxMP];
    gspat_iter_run<MP>(a, blockIdx.x, iters, closing, s_p);
}

// ---------------------------------------------------------------- kernel A, fused (MFMA modes)
// One workgroup of T = MP / 16 waves per instance and source-chunk range. Wave
// w owns the foci 16 w .. 16 w + 15 (row tile w): lane l evaluates the terms
// of focus 16 w + l % 16 at the 8 sources 8 (l / 16) .. + 7 of each chunk of
// 32, which is where v_mfma_f32_16x16x32_f16 expects a row of H in its A and B
// operands. Per chunk:
//   1. the 8 terms (sources staged in LDS, one chunk ahead); a tile whose
//      rows are all padding (>= M) skips them;
//   2. d_j = sum_i |G_ij|: the 16 rows of the wave by DPP (every lane of the
//      row group ends with the same bits), then the T waves' partials through
//      LDS in wave order;
//   3. H = s G / d_j (s = 2^floor(log2 M) / 2, a power of two that keeps the
//      low halves in the FP16 normal range) split into FP16 hi + lo: the
//      wave's own operands stay in registers, and are written to LDS for the
//      other waves; G (FP32) or H hi (FP16) is stored for kernel C;
//   4. R += hi hi^H + hi lo^H + lo hi^H on MFMA with FP32 accumulation, over
//      the wave's tiles (w, (w + d) mod T): d = 0 .. (T - 1) / 2, plus d = T / 2
//      for w < T / 2, so each unordered pair of row tiles is formed once, with
//      the wave's own registers as the A operand. A wave with one tile fewer
//      forms a duplicate (never written) so every wave runs the same
//      straight-line MFMA sequence.
// R (the partial of this source range, times 1 / s^2) is written in full: each
// tile and, off the diagonal, its conjugate transpose.
template <int T>
struct FusedTiles {
    static constexpr int kMax = 1 + (T - 1) / 2 + (T % 2 == 0 ? 1 : 0);
};

typedef float floatx2 __attribute__((ext_vector_type(2)));
typedef _Float16 halfx2 __attribute__((ext_vector_type(2)));

// x[v] f[v] (v < 8) split into FP16 hi + lo halves, two halves a word in v
// order (the half8 operand layout of mfma16): hi = fp16(x f), lo = fp16(x f
// - hi), x f - hi formed exactly by one fused multiply-add and rounded once.
// hi + lo carries x f to ~22 bits. Two values per v_cvt_pk_f16_f32, lo by
// v_fma_mix{lo,hi}_f16 straight into the packed word.
__device__ __forceinline__ void split_h8(const float (&x)[8], const float (&f)[8], uintx4& hi, uintx4& lo) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float p0 = x[2 * i] * f[2 * i], p1 = x[2 * i + 1] * f[2 * i + 1];
        const halfx2 h = __builtin_convertvector((floatx2){p0, p1}, halfx2);
        const halfx2 r = {(_Float16)fmaf(x[2 * i], f[2 * i], -(float)h.x),
                          (_Float16)fmaf(x[2 * i + 1], f[2 * i + 1], -(float)h.y)};
        hi[i] = __builtin_bit_cast(uint32_t, h);
        lo[i] = __builtin_bit_cast(uint32_t, r);
    }
}

// Kernel C's FP16 copy of 8 elements, (re, im) halves a word, from the packed
// real and imaginary hi words: element 2 i is the low halves of word i,
// element 2 i + 1 the high halves. Words 4 h .. 4 h + 3 of the copy.
__device__ __forceinline__ uintx4 interleave_h8(const uintx4& re, const uintx4& im, int h) {
    const uint32_t r0 = re[2 * h], r1 = re[2 * h + 1], i0 = im[2 * h], i1 = im[2 * h + 1];
    return uintx4{(r0 & 0xffffu) | (i0 << 16), (r0 >> 16) | (i0 & 0xffff0000u), (r1 & 0xffffu) | (i1 << 16),
                  (r1 >> 16) | (i1 & 0xffff0000u)};
}

// The sign bits of the 8 FP16 halves of an operand flipped (exact negation).
__device__ __forceinline__ uintx4 neg_h8(const uintx4& x) { return x ^ 0x80008000u; }

// D += A B on v_mfma_f32_16x16x32_f16 (K = 32, 8 halves per lane): lane l
// supplies A[l % 16][8 (l / 16) + v] and B[8 (l / 16) + v][l % 16] and
// receives D[4 (l / 16) + e][l % 16].
__device__ __forceinline__ floatx4 mfma16(const uintx4& a, const uintx4& b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(half8, a), __builtin_bit_cast(half8, b), c, 0,
                                                   0, 0);
}

template <int SRC, int MP, bool G32>
__global__ __launch_bounds__(4 * MP, 4) void gspat_fused(HoloArgs a) {
    constexpr int T = MP / 16;
    constexpr int NTW = FusedTiles<T>::kMax;
    constexpr int RS = kKC + 8;  // LDS row stride in halves (80 bytes)
    __shared__ std::conditional_t<SRC == kT4010A1, CellsLds, float> s_cells_mem;
    CellsLds* s_cells = reinterpret_cast<CellsLds*>(&s_cells_mem);
    __shared__ float4 s_src[2][3][kKC];                            // staged sources of a chunk, by float4 index
    __shared__ __attribute__((aligned(16))) float s_part[T][kKC];  // per-wave column sums of |G|
    __shared__ __attribute__((aligned(16))) _Float16 s_b[4][MP * RS];  // hi re, hi im, lo re, lo im
    const int t = threadIdx.x;
    // the wave index as a scalar: the wave's tile list and liveness in SGPRs
    const int w = __builtin_amdgcn_readfirstlane(t >> 6), l = t & 63, r16 = l & 15, g = l >> 4;
    const uint32_t k = blockIdx.x / a.split, part = blockIdx.x % a.split, b = a.inst[k];
    const uint32_t n_chunks = a.n_packed / kKC;
    const uint32_t c_begin = part * n_chunks / a.split, c_end = (part + 1) * n_chunks / a.split;
    const uint32_t base = a.offsets[b], M = a.offsets[b + 1] - base;
    if (SRC == kT4010A1) {
        const float4* gc = reinterpret_cast<const float4*>(a.cells);
        for (int j = t; j < kDirCells; j += 4 * MP) {
            s_cells->p[j] = gc[2 * j];
            s_cells->q[j] = gc[2 * j + 1];
        }
    }
    // the lane's focus (rows >= M are padding: weight 0)
    const int row = 16 * w + r16;
    float px = 0.f, py = 0.f, pz = 1.f, wt = 0.f;
    if (row < (int)M) {
        const uint64_t q = 3 * (uint64_t)(base + row);
        px = a.foci[q];
        py = a.foci[q + 1];
        pz = a.foci[q + 2];
        wt = 1.f;
    }
    const Split3 pf = split3(px, py, pz);
    const bool live = 16 * w < (int)M;  // wave-uniform
    // uniform: kept in an SGPR
    const float scale = __builtin_bit_cast(
        float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, fmaxf(1.0f, exp2f(floorf(log2f((float)M)) - 1.0f)))));
    // the chunk's staged sources: thread t < 3 kKC loads float4 t of the chunk
    constexpr int kStageT = 3 * kKC;
    float4 pre = make_float4(0.f, 0.f, 0.f, 0.f);
    auto fetch_src = [&](uint32_t chunk) {
        if (t < kStageT) pre = a.staged[(uint64_t)chunk * kKC * 3 + t];
    };
    auto put_src = [&](int buf) {
        if (t < kStageT) s_src[buf][t % 3][t / 3] = pre;
    };
    if (c_begin < c_end) {
        fetch_src(c_begin);
        put_src(0);
        if (c_begin + 1 < c_end) fetch_src(c_begin + 1);
    }
    __syncthreads();
    // tiles of this wave: J = (w + d) mod T
    int tJ[NTW];
#pragma unroll
    for (int s = 0; s < NTW; ++s) {
        int d = s;
        if (s == NTW - 1 && T % 2 == 0 && w >= T / 2) d = 0;  // the duplicate slot
        tJ[s] = (w + d) % T;
    }
    const bool has_last = !(T % 2 == 0 && w >= T / 2);
    floatx4 accRe[NTW], accIm[NTW];
#pragma unroll
    for (int s = 0; s < NTW; ++s) {
        accRe[s] = floatx4{0.f, 0.f, 0.f, 0.f};
        accIm[s] = accRe[s];
    }
    float* d_out = a.d + (uint64_t)k * a.n_packed;
    for (uint32_t chunk = c_begin; chunk < c_end; ++chunk) {
        const uint32_t j0 = chunk * kKC;
        const int buf = (chunk - c_begin) & 1;
        // 1. terms of the lane's focus at sources j0 + 8 g + v
        float re[8], im[8], am[8];
#pragma unroll
        for (int v = 0; v < 8; ++v) re[v] = im[v] = am[v] = 0.f;
        if (live) {
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                const int c = 8 * g + v;
                const float4 q0 = s_src[buf][0][c], q1 = s_src[buf][1][c], q2 = s_src[buf][2][c];
                const SrcF sv = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
                add_terms_exact<SRC, 1>(sv, split3(q0.x, q0.y, q0.z), &pf, s_cells->p, s_cells->q, &re[v], &im[v], &wt,
                                        &am[v]);
                // one term at a time where the 4-wave bound would otherwise
                // spill (the FP16-copy variant; the FP32-G one fits as is)
                if constexpr (!G32) __builtin_amdgcn_sched_barrier(0);
            }
        }
        // 2. column sums over the wave's 16 rows, then over the waves
        float cs[8];
#pragma unroll
        for (int v = 0; v < 8; ++v) cs[v] = row_group_sum<16>(am[v]);
        if (r16 == 0) {
            *reinterpret_cast<float4*>(&s_part[w][8 * g]) = make_float4(cs[0], cs[1], cs[2], cs[3]);
            *reinterpret_cast<float4*>(&s_part[w][8 * g + 4]) = make_float4(cs[4], cs[5], cs[6], cs[7]);
        }
        lds_sync();
        // every wave is past its reads of s_src[buf ^ 1] (chunk - 1) and of s_b
        if (chunk + 1 < c_end) {
            put_src(buf ^ 1);
            if (chunk + 2 < c_end) fetch_src(chunk + 2);
        }
        float d[8];
#pragma unroll
        for (int v = 0; v < 8; ++v) d[v] = 0.f;
#pragma unroll
        for (int ww = 0; ww < T; ++ww) {
            const float4 p0 = *reinterpret_cast<const float4*>(&s_part[ww][8 * g]);
            const float4 p1 = *reinterpret_cast<const float4*>(&s_part[ww][8 * g + 4]);
            d[0] += p0.x;
            d[1] += p0.y;
            d[2] += p0.z;
            d[3] += p0.w;
            d[4] += p1.x;
            d[5] += p1.y;
            d[6] += p1.z;
            d[7] += p1.w;
        }
        if (w == 0 && r16 == 0) {
            float4* dp = reinterpret_cast<float4*>(d_out + j0 + 8 * g);
            dp[0] = make_float4(d[0], d[1], d[2], d[3]);
            dp[1] = make_float4(d[4], d[5], d[6], d[7]);
        }
        // 3. H = s G / d_j split into FP16 hi + lo (a real column with d = 0
        // gives NaN, as conj(G) / 0 does in the reference)
        float inv[8];
#pragma unroll
        for (int v = 0; v < 8; ++v) {
            // 1 / d from v_rcp_f32 and one Newton step (within an ulp)
            float r = __builtin_amdgcn_rcpf(d[v]);
            r = fmaf(r, fmaf(-d[v], r, 1.0f), r);
            inv[v] = (j0 + 8 * g + v < a.n) ? scale * r : 0.f;
        }
        uintx4 ahr, ahi, alr, ali;
        split_h8(re, inv, ahr, alr);
        split_h8(im, inv, ahi, ali);
        {
            const int o = row * RS + 8 * g;
            *reinterpret_cast<uintx4*>(&s_b[0][o]) = ahr;
            *reinterpret_cast<uintx4*>(&s_b[1][o]) = ahi;
            *reinterpret_cast<uintx4*>(&s_b[2][o]) = alr;
            *reinterpret_cast<uintx4*>(&s_b[3][o]) = ali;
        }
        // kernel C's copy of the tile, rows < M: FP32 G, or the hi half of H
        if (row < (int)M) {
            if constexpr (G32) {
                floatx4* gp = reinterpret_cast<floatx4*>(a.G + ((uint64_t)k * MP + row) * a.n_packed + j0 + 8 * g);
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    gp[v] = floatx4{re[2 * v], im[2 * v], re[2 * v + 1], im[2 * v + 1]};
            } else {
                uintx4* hp = reinterpret_cast<uintx4*>(reinterpret_cast<uint32_t*>(a.G) +
                                                       ((uint64_t)k * MP + row) * a.n_packed + j0 + 8 * g);
                hp[0] = interleave_h8(ahr, ahi, 0);
                hp[1] = interleave_h8(ahr, ahi, 1);
            }
        }
        lds_sync();
        // 4. R tiles: A = this wave's rows (registers), B = rows of tile J
        const uintx4 nhr = neg_h8(ahr), nlr = neg_h8(alr);
#pragma unroll
        for (int s = 0; s < NTW; ++s) {
            uintx4 bhr = ahr, bhi = ahi, blr = alr, bli = ali;
            if (tJ[s] != w) {
                const int o = (16 * tJ[s] + r16) * RS + 8 * g;
                bhr = *reinterpret_cast<const uintx4*>(&s_b[0][o]);
                bhi = *reinterpret_cast<const uintx4*>(&s_b[1][o]);
                blr = *reinterpret_cast<const uintx4*>(&s_b[2][o]);
                bli = *reinterpret_cast<const uintx4*>(&s_b[3][o]);
            }
            // the small cross products first
            accRe[s] = mfma16(ahr, blr, accRe[s]);
            accRe[s] = mfma16(alr, bhr, accRe[s]);
            accRe[s] = mfma16(ahi, bli, accRe[s]);
            accRe[s] = mfma16(ali, bhi, accRe[s]);
            accIm[s] = mfma16(ahi, blr, accIm[s]);
            accIm[s] = mfma16(ali, bhr, accIm[s]);
            accIm[s] = mfma16(nhr, bli, accIm[s]);
            accIm[s] = mfma16(nlr, bhi, accIm[s]);
            accRe[s] = mfma16(ahr, bhr, accRe[s]);
            accRe[s] = mfma16(ahi, bhi, accRe[s]);
            accIm[s] = mfma16(ahi, bhr, accIm[s]);
            accIm[s] = mfma16(nhr, bhi, accIm[s]);
        }
    }
    // lane l holds rows 4 (l / 16) + e, column l % 16 of each tile
    const float unscale = 1.0f / (scale * scale);  // exact: scale is a power of two
    float2* R = a.R + (uint64_t)blockIdx.x * MP * MP;
#pragma unroll
    for (int s = 0; s < NTW; ++s) {
        if (s == NTW - 1 && !has_last) break;
        const int J = tJ[s];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int ri = 16 * w + 4 * g + e, cj = 16 * J + r16;
            const float2 v = make_float2(accRe[s][e] * unscale, accIm[s][e] * unscale);
            R[ri * MP + cj] = v;
            if (J != w) R[cj * MP + ri] = make_float2(v.x, -v.y);
        }
    }
}

// ---------------------------------------------------------------- kernel A, one wave per instance row set
// The terms of a lane's first NL rows (points pf[0 .. NL - 1]; rows NL .. NQ - 1
// are zero) at the staged sources j0 .. j0 + 7 of sb: re / im[q][v], and cs[v]
// the sum of the amplitudes over q in order. NL is a template constant so the
// 8 NL terms form one straight-line block in which the LDS reads and the long
// per-term chains overlap: the NL rows of a source in one staged evaluation,
// sources in pairs; with SERIAL (kernels that also keep FP32 G, where the
// longer live ranges spill) each term on its own and one source at a time.
template <int SRC, int NQ, int NL, bool SERIAL>
__device__ __forceinline__ void lane_terms(const SrcBlock& sb, int j0, const Split3 (&pf)[NQ], const float (&wt)[NQ],
                                           const CellsLds* cells, float (&re)[NQ][8], float (&im)[NQ][8],
                                           float (&cs)[8]) {
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        float rr[NL > 0 ? NL : 1], ii[NL > 0 ? NL : 1], am[NL > 0 ? NL : 1];
        if constexpr (NL > 0) {
            const int j = j0 + v;
            const float4 q0 = sb.v[0][j], q1 = sb.v[1][j], q2 = sb.v[2][j];
            const SrcF sv = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
#pragma unroll
            for (int q = 0; q < NL; ++q) rr[q] = ii[q] = 0.f;
            if constexpr (SERIAL) {
                const Split3 ss = split3(q0.x, q0.y, q0.z);
#pragma unroll
                for (int q = 0; q < NL; ++q)
                    add_terms_exact<SRC, 1>(sv, ss, &pf[q], cells->p, cells->q, &rr[q], &ii[q], &wt[q], &am[q]);
            } else {
                add_terms_exact<SRC, NL>(sv, split3(q0.x, q0.y, q0.z), pf, cells->p, cells->q, rr, ii, wt, am);
            }
        }
        cs[v] = 0.f;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            if (q < NL) {
                re[q][v] = rr[q];
                im[q][v] = ii[q];
                cs[v] += am[q];
            } else {
                re[q][v] = im[q][v] = 0.f;
            }
        }
        if (NL > 1 && (v + 1) % (SERIAL ? 1 : 2) == 0) __builtin_amdgcn_sched_barrier(0);
    }
}

// lane_terms with the live rows tested term by term (each term its own
// block): the shortest live ranges, for the wave-pair kernel that keeps FP32 G.
template <int SRC, int NQ>
__device__ __forceinline__ void lane_terms_each(int nl, const SrcBlock& sb, int j0, const Split3 (&pf)[NQ],
                                                const float (&wt)[NQ], const CellsLds* cells, float (&re)[NQ][8],
                                                float (&im)[NQ][8], float (&cs)[8]) {
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        cs[v] = 0.f;
        const int j = j0 + v;
        const float4 q0 = sb.v[0][j], q1 = sb.v[1][j], q2 = sb.v[2][j];
        const SrcF sv = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
        const Split3 ss = split3(q0.x, q0.y, q0.z);
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            re[q][v] = im[q][v] = 0.f;
            if (q < nl) {
                float am;
                add_terms_exact<SRC, 1>(sv, ss, &pf[q], cells->p, cells->q, &re[q][v], &im[q][v], &wt[q], &am);
                cs[v] += am;
            }
        }
    }
}

// lane_terms for the nl (wave-uniform) live rows of the lane's NQ <= 2.
template <int SRC, int NQ, bool SERIAL>
__device__ __forceinline__ void lane_terms_live(int nl, const SrcBlock& sb, int j0, const Split3 (&pf)[NQ],
                                                const float (&wt)[NQ], const CellsLds* cells, float (&re)[NQ][8],
                                                float (&im)[NQ][8], float (&cs)[8]) {
    static_assert(NQ == 1 || NQ == 2, "one or two rows per lane");
    if constexpr (NQ == 2) {
        if (nl >= 2) {
            lane_terms<SRC, 2, 2, SERIAL>(sb, j0, pf, wt, cells, re, im, cs);
            return;
        }
    }
    if (nl >= 1)
        lane_terms<SRC, NQ, 1, SERIAL>(sb, j0, pf, wt, cells, re, im, cs);
    else
        lane_terms<SRC, NQ, 0, SERIAL>(sb, j0, pf, wt, cells, re, im, cs);
}


// For MP <= 32 (T = MP / 16 <= 2 row tiles): a workgroup of 4 waves takes one
// instance and one source-chunk range; wave w handles the chunks c_begin + w,
// c_begin + w + 4, ... of the range, each for all MP foci, so the column sums
// d_j and every MFMA operand stay inside the wave. A chunk of 32 sources is one
// K = 32 step (SPL = 8 sources per lane). Lane l evaluates focus 16 t + l % 16
// (t < T) at the sources SPL (l / 16) .. + SPL - 1 of the chunk: each staged
// source is read once and used for the T foci. Per step:
//   1. the SPL T terms (tiles whose rows are all padding skip them);
//   2. d_j: the lane's T rows in order, then the 16 rows of the lane group by
//      DPP (every lane of the group ends with the same bits);
//   3. kernel C's copy (FP32 G, or the hi half of H) for rows < M;
//   4. H = s G / d_j (s = 2^floor(log2 M) / 2) split into FP16 hi + lo;
//   5. R(I, J) += hi_I hi_J^H + hi_I lo_J^H + lo_I hi_J^H on MFMA (FP32
//      accumulation) for the upper-triangle tiles I <= J with live rows.
// Sources are staged in LDS kSB chunks at a time, loaded a block ahead and
// written after the block (one barrier per block). At the end the four waves'
// partial R are summed in LDS in wave order, and the workgroup writes R (times
// 1 / s^2) with the lower tiles as conjugate transposes. Occupancy: 3 waves
// per SIMD (at most 168 registers, the accumulators in VGPRs, no spills).
// 64 foci run on wave pairs (gspat_pair).
template <int SRC, int MP, bool G32>
__global__ __launch_bounds__(kHB, 3) void gspat_wave(HoloArgs a) {
    constexpr int T = MP / 16;
    constexpr int NU = T * (T + 1) / 2;  // upper-triangle tiles
    constexpr int SPL = 8;               // sources per lane per step
    constexpr int KS = 4 * SPL;          // columns per step
    static_assert(T >= 1 && T <= 2, "one wave holds at most 2 row tiles");
    __shared__ std::conditional_t<SRC == kT4010A1, CellsLds, float> s_cells_mem;
    CellsLds* s_cells = reinterpret_cast<CellsLds*>(&s_cells_mem);
    constexpr size_t kSrcBytes = 2 * sizeof(SrcBlock), kRBytes = (size_t)MP * MP * sizeof(float2);
    __shared__ __attribute__((aligned(16))) unsigned char s_buf[kSrcBytes > kRBytes ? kSrcBytes : kRBytes];
    SrcBlock* s_src = reinterpret_cast<SrcBlock*>(s_buf);
    float2* s_R = reinterpret_cast<float2*>(s_buf);
    const int t = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6), l = t & 63, r16 = l & 15, g = l >> 4;
    const uint32_t k = blockIdx.x / a.split, part = blockIdx.x % a.split, b = a.inst[k];
    const uint32_t n_chunks = a.n_packed / kKC;
    const uint32_t c_begin = part * n_chunks / a.split, c_end = (part + 1) * n_chunks / a.split;
    const uint32_t base = a.offsets[b], M = a.offsets[b + 1] - base;
    if (SRC == kT4010A1) stage_cells_lds(a.cells, s_cells);
    Split3 pf[T];
    float wt[T];
    bool live[T];
#pragma unroll
    for (int q = 0; q < T; ++q) {
        const int row = 16 * q + r16;
        live[q] = 16 * q < (int)M;  // wave-uniform
        float px = 0.f, py = 0.f, pz = 1.f;
        wt[q] = 0.f;
        if (row < (int)M) {
            const uint64_t o = 3 * (uint64_t)(base + row);
            px = a.foci[o];
            py = a.foci[o + 1];
            pz = a.foci[o + 2];
            wt[q] = 1.f;
        }
        pf[q] = split3(px, py, pz);
    }
    int n_live = 0;  // live rows tiles are a prefix
#pragma unroll
    for (int q = 0; q < T; ++q) n_live += live[q] ? 1 : 0;
    const float scale = fmaxf(1.0f, exp2f(floorf(log2f((float)M)) - 1.0f));
    floatx4 accRe[NU], accIm[NU];
#pragma unroll
    for (int s = 0; s < NU; ++s) {
        accRe[s] = floatx4{0.f, 0.f, 0.f, 0.f};
        accIm[s] = accRe[s];
    }
    float* d_out = a.d + (uint64_t)k * a.n_packed;
    float4 pre[3];
    if (c_begin < c_end) {
        src_block_fetch(a, c_begin, pre);
        src_block_put(s_src[0], pre);
        if (c_begin + kSB < c_end) src_block_fetch(a, c_begin + kSB, pre);
    }
    __syncthreads();
    for (uint32_t c0 = c_begin; c0 < c_end; c0 += kSB) {
        const int bb = ((c0 - c_begin) / kSB) & 1;
        for (uint32_t cb = w; cb < kSB && c0 + cb < c_end; cb += 4) {
#pragma unroll
            for (int h = 0; h < kKC / KS; ++h) {
                const uint32_t jb = (c0 + cb) * kKC + h * KS + SPL * g;  // the lane's first column
                // 1. terms; cs: the lane's rows in order
                float re[T][SPL], im[T][SPL], cs[SPL];
                lane_terms_live<SRC, T, G32>(n_live, s_src[bb], cb * kKC + h * KS + SPL * g, pf, wt, s_cells, re, im,
                                             cs);
                // 2. d_j over the 16 lanes of the group
                float d[SPL];
#pragma unroll
                for (int v = 0; v < SPL; ++v) d[v] = row_group_sum<16>(cs[v]);
                if (r16 == 0) {
                    floatx4* dp = reinterpret_cast<floatx4*>(d_out + jb);
#pragma unroll
                    for (int v = 0; v < SPL / 4; ++v) dp[v] = floatx4{d[4 * v], d[4 * v + 1], d[4 * v + 2], d[4 * v + 3]};
                }
                // 3. kernel C's copy, rows < M (FP32 G here; H16 after the split)
                if constexpr (G32) {
#pragma unroll
                    for (int q = 0; q < T; ++q) {
                        const int row = 16 * q + r16;
                        if (row < (int)M) {
                            floatx4* gp = reinterpret_cast<floatx4*>(a.G + ((uint64_t)k * MP + row) * a.n_packed + jb);
#pragma unroll
                            for (int v = 0; v < SPL / 2; ++v)
                                gp[v] = floatx4{re[q][2 * v], im[q][2 * v], re[q][2 * v + 1], im[q][2 * v + 1]};
                        }
                    }
                }
                // 4. H = s G / d_j in FP16 hi + lo (a real column with d = 0
                // gives NaN, as conj(G) / 0 does in the reference)
                float inv[SPL];
#pragma unroll
                for (int v = 0; v < SPL; ++v) {
                    float r = __builtin_amdgcn_rcpf(d[v]);
                    r = fmaf(r, fmaf(-d[v], r, 1.0f), r);
                    inv[v] = (jb + v < a.n) ? scale * r : 0.f;
                }
                uintx4 hr[T], hi[T], lr[T], li[T];
#pragma unroll
                for (int q = 0; q < T; ++q) {
                    split_h8(re[q], inv, hr[q], lr[q]);
                    split_h8(im[q], inv, hi[q], li[q]);
                    if constexpr (!G32) {
                        const int row = 16 * q + r16;
                        if (row < (int)M) {
                            uintx4* hp = reinterpret_cast<uintx4*>(reinterpret_cast<uint32_t*>(a.G) +
                                                                   ((uint64_t)k * MP + row) * a.n_packed + jb);
#pragma unroll
                            for (int v = 0; v < SPL / 4; ++v) hp[v] = interleave_h8(hr[q], hi[q], v);
                        }
                    }
                }
                // 5. the upper-triangle tiles with live rows
                int s = 0;
#pragma unroll
                for (int I = 0; I < T; ++I) {
                    const uintx4 nhr = neg_h8(hr[I]), nlr = neg_h8(lr[I]);
#pragma unroll
                    for (int J = I; J < T; ++J, ++s) {
                        if (!(live[I] && live[J])) continue;
                        accRe[s] = mfma16(hr[I], lr[J], accRe[s]);
                        accRe[s] = mfma16(lr[I], hr[J], accRe[s]);
                        accRe[s] = mfma16(hi[I], li[J], accRe[s]);
                        accRe[s] = mfma16(li[I], hi[J], accRe[s]);
                        accIm[s] = mfma16(hi[I], lr[J], accIm[s]);
                        accIm[s] = mfma16(li[I], hr[J], accIm[s]);
                        accIm[s] = mfma16(nhr, li[J], accIm[s]);
                        accIm[s] = mfma16(nlr, hi[J], accIm[s]);
                        accRe[s] = mfma16(hr[I], hr[J], accRe[s]);
                        accRe[s] = mfma16(hi[I], hi[J], accRe[s]);
                        accIm[s] = mfma16(hi[I], hr[J], accIm[s]);
                        accIm[s] = mfma16(nhr, hi[J], accIm[s]);
                    }
                }
            }
        }
        // every wave is past its reads of this block: stage the next one
        if (c0 + kSB < c_end) {
            src_block_put(s_src[bb ^ 1], pre);
            if (c0 + 2 * kSB < c_end) src_block_fetch(a, c0 + 2 * kSB, pre);
        }
        lds_sync();
    }
    // the four waves' partials summed in LDS in wave order; lane l holds rows
    // 4 (l / 16) + e, column l % 16 of each tile
    for (int ww = 0; ww < 4; ++ww) {
        if (w == ww) {
            int s = 0;
#pragma unroll
            for (int I = 0; I < T; ++I)
#pragma unroll
                for (int J = I; J < T; ++J, ++s)
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int ri = 16 * I + 4 * g + e, cj = 16 * J + r16;
                        float2 v = make_float2(accRe[s][e], accIm[s][e]);
                        if (ww > 0) {
                            const float2 o = s_R[ri * MP + cj];
                            v = make_float2(o.x + v.x, o.y + v.y);
                        }
                        s_R[ri * MP + cj] = v;
                    }
        }
        lds_sync();
    }
    const float unscale = 1.0f / (scale * scale);  // exact: scale is a power of two
    float2* R = a.R + (uint64_t)blockIdx.x * MP * MP;
    for (int e = t; e < MP * MP; e += kHB) {
        const int ri = e / MP, cj = e % MP;
        // the upper triangle of tiles holds the sums; lower tiles by symmetry
        const bool up = ri / 16 <= cj / 16;
        const float2 v = up ? s_R[e] : s_R[cj * MP + ri];
        R[e] = make_float2(v.x * unscale, up ? v.y * unscale : -v.y * unscale);
    }
}

// ---------------------------------------------------------------- kernel A, wave pairs (64 foci)
// Kernel A for MP = 64 on pairs of waves: the two waves of a pair take the same
// chunk of 32 sources, each for its own 32 foci (row tiles 2 hw and 2 hw + 1,
// hw the wave's half of the pair), as gspat_wave lays out one wave for 32 foci
// (K = 32 MFMA steps, 8 sources per lane). Per chunk:
//   1. the 16 terms of each lane (2 row tiles x 8 sources);
//   2. the pair's column sums d_j from the two waves' partials over their 32
//      rows, exchanged through LDS (both waves add them in the same order);
//   3. H = s G / d_j split into FP16 hi + lo, kernel C's copy stored;
//   4. the operands the partner needs exchanged through LDS (hw = 0 sends tile
//      1, hw = 1 tiles 2 and 3), then 5 of the 10 upper-triangle tiles per
//      wave: hw = 0 (0,0) (0,1) (1,1) (0,2) (0,3), hw = 1 (2,2) (2,3) (3,3)
//      (1,2) (1,3).
// Half the accumulators and operands of the one-wave layout per wave: 3 waves
// per SIMD (53 KB of LDS per workgroup of 2 pairs, the source block single-
// buffered). The two pairs take alternate chunks of the workgroup's range and
// run in lockstep (two barriers per chunk); their partial R are summed in
// LDS in pair order at the end. Kernel C's copy is written with plain stores:
// the FP32 copy with non-temporal ones (16-byte pieces at a 64-byte stride per
// row and instruction) took 40% longer.
template <int SRC, bool G32>
__global__ __launch_bounds__(kHB, 3) void gspat_pair(HoloArgs a) {
    constexpr int MP = 64, SPL = 8;
    static_assert(4 * SPL == kKC, "one K = 32 step per chunk");
    __shared__ std::conditional_t<SRC == kT4010A1, CellsLds, float> s_cells_mem;
    CellsLds* s_cells = reinterpret_cast<CellsLds*>(&s_cells_mem);
    struct Xch {
        SrcBlock src;
        uintx4 op[2][3][4][64];  // [pair][slot: tile 1, 2, 3][hr hi lr li][lane]
        float dpart[2][2][kKC];  // [pair][hw][column]
    };
    constexpr size_t kXBytes = sizeof(Xch), kRBytes = (size_t)MP * MP * sizeof(float2);
    __shared__ __attribute__((aligned(16))) unsigned char s_buf[kXBytes > kRBytes ? kXBytes : kRBytes];
    Xch* X = reinterpret_cast<Xch*>(s_buf);
    float2* s_R = reinterpret_cast<float2*>(s_buf);
    const int t = threadIdx.x;
    const int w = t >> 6, l = t & 63, r16 = l & 15, g = l >> 4;
    const int pr = w >> 1, hw = w & 1;
    const uint32_t k = blockIdx.x / a.split, part = blockIdx.x % a.split, b = a.inst[k];
    const uint32_t n_chunks = a.n_packed / kKC;
    const uint32_t c_begin = part * n_chunks / a.split, c_end = (part + 1) * n_chunks / a.split;
    const uint32_t base = a.offsets[b], M = a.offsets[b + 1] - base;
    if (SRC == kT4010A1) stage_cells_lds(a.cells, s_cells);
    // own row tiles 2 hw + q, q < 2
    Split3 pf[2];
    float wt[2];
    bool live[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int tl = 2 * hw + q, row = 16 * tl + r16;
        live[q] = 16 * tl < (int)M;  // wave-uniform
        float px = 0.f, py = 0.f, pz = 1.f;
        wt[q] = 0.f;
        if (row < (int)M) {
            const uint64_t o = 3 * (uint64_t)(base + row);
            px = a.foci[o];
            py = a.foci[o + 1];
            pz = a.foci[o + 2];
            wt[q] = 1.f;
        }
        pf[q] = split3(px, py, pz);
    }
    const int n_live = (live[0] ? 1 : 0) + (live[1] ? 1 : 0);  // live tiles are a prefix
    const bool live1 = 16 < (int)M, live2 = 32 < (int)M, live3 = 48 < (int)M;  // tile 0 is always live
    const float scale = fmaxf(1.0f, exp2f(floorf(log2f((float)M)) - 1.0f));
    // accumulators: slots 0..2 the wave's own tiles (q, q') = (0,0) (0,1) (1,1),
    // slots 3, 4 the cross tiles
    floatx4 accRe[5], accIm[5];
#pragma unroll
    for (int s = 0; s < 5; ++s) {
        accRe[s] = floatx4{0.f, 0.f, 0.f, 0.f};
        accIm[s] = accRe[s];
    }
    float* d_out = a.d + (uint64_t)k * a.n_packed;
    // kernel C's FP16 copy of this instance (MP n_packed words < 2^32 B, checked on the host)
    const __amdgpu_buffer_rsrc_t gw = __builtin_amdgcn_make_buffer_rsrc(
        reinterpret_cast<uint32_t*>(a.G) + (uint64_t)k * MP * a.n_packed, (short)0, (int)(MP * a.n_packed * 4u),
        0x00020000);
    float4 pre[3];
    if (c_begin < c_end) {
        src_block_fetch(a, c_begin, pre);
        src_block_put(X->src, pre);
    }
    __syncthreads();
    for (uint32_t c0 = c_begin; c0 < c_end; c0 += kSB) {
        if (c0 + kSB < c_end) src_block_fetch(a, c0 + kSB, pre);
        for (int rd = 0; rd < kSB / 2; ++rd) {
            const int cb = 2 * rd + pr;  // the pair's chunk within the block
            const bool act = c0 + cb < c_end;  // pair-uniform
            const uint32_t jb = (c0 + cb) * kKC + SPL * g;  // the lane's first column
            uintx4 hr[2], hi[2], lr[2], li[2];
            if (act) {
                // 1. terms; cs: the lane's rows in order
                float re[2][SPL], im[2][SPL], cs[SPL];
                if constexpr (G32)
                    lane_terms_each<SRC, 2>(n_live, X->src, cb * kKC + SPL * g, pf, wt, s_cells, re, im, cs);
                else
                    lane_terms_live<SRC, 2, false>(n_live, X->src, cb * kKC + SPL * g, pf, wt, s_cells, re, im, cs);
                // 2. the wave's partial column sums over its 32 rows
#pragma unroll
                for (int v = 0; v < SPL; ++v) cs[v] = row_group_sum<16>(cs[v]);
                if (r16 == 0) {
                    floatx4* dp = reinterpret_cast<floatx4*>(&X->dpart[pr][hw][SPL * g]);
                    dp[0] = floatx4{cs[0], cs[1], cs[2], cs[3]};
                    dp[1] = floatx4{cs[4], cs[5], cs[6], cs[7]};
                }
                lds_sync();
                float d[SPL];
#pragma unroll
                for (int v = 0; v < SPL; ++v) d[v] = X->dpart[pr][0][SPL * g + v] + X->dpart[pr][1][SPL * g + v];
                if (hw == 0 && r16 == 0) {
                    floatx4* dg = reinterpret_cast<floatx4*>(d_out + jb);
                    dg[0] = floatx4{d[0], d[1], d[2], d[3]};
                    dg[1] = floatx4{d[4], d[5], d[6], d[7]};
                }
                // 3. H = s G / d_j in FP16 hi + lo; kernel C's copy
                float inv[SPL];
#pragma unroll
                for (int v = 0; v < SPL; ++v) {
                    float r = __builtin_amdgcn_rcpf(d[v]);
                    r = fmaf(r, fmaf(-d[v], r, 1.0f), r);
                    inv[v] = (jb + v < a.n) ? scale * r : 0.f;
                }
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    split_h8(re[q], inv, hr[q], lr[q]);
                    split_h8(im[q], inv, hi[q], li[q]);
                    const int row = 16 * (2 * hw + q) + r16;
                    if (row < (int)M) {
                        if constexpr (G32) {
                            floatx4* gp = reinterpret_cast<floatx4*>(a.G + ((uint64_t)k * MP + row) * a.n_packed + jb);
#pragma unroll
                            for (int v = 0; v < SPL / 2; ++v)
                                gp[v] = floatx4{re[q][2 * v], im[q][2 * v], re[q][2 * v + 1], im[q][2 * v + 1]};
                        } else {
                            // through the instance's buffer resource (a 32-bit offset per
                            // lane instead of a 64-bit address: one VGPR, not two)
                            const uint32_t off = ((uint32_t)row * a.n_packed + jb) * 4u;
#pragma unroll
                            for (int v = 0; v < SPL / 4; ++v)
                                __builtin_amdgcn_raw_buffer_store_b128(interleave_h8(hr[q], hi[q], v), gw,
                                                                       (int)(off + 16u * v), 0, 0);
                        }
                    }
                }
                // 4. the operands the partner needs: hw = 0 sends tile 1 (slot
                // 0), hw = 1 tiles 2 and 3 (slots 1, 2)
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    if (hw == 0 && q == 0) continue;
                    const int slot = hw == 0 ? 0 : 1 + q;
                    X->op[pr][slot][0][l] = hr[q];
                    X->op[pr][slot][1][l] = hi[q];
                    X->op[pr][slot][2][l] = lr[q];
                    X->op[pr][slot][3][l] = li[q];
                }
            } else {
                lds_sync();  // the d exchange of the other pair
            }
            lds_sync();
            if (act) {
                // R(I, J) += H_I H_J^H in split FP16: the 12 products of
                // gspat_wave, with hr / hi / lr / li the real and imaginary
                // hi and lo halves. Im R = Hi_I Hr_J^T - Hr_I Hi_J^T: the
                // three negative products take (m1a, m1b), (m2a, m2b),
                // (m3a, m3b), the minus on A's real halves (wave 0, whose A
                // tiles are its own) or on B's imaginary halves (wave 1, whose
                // B tiles are its own), so each wave negates only its own two
                // tiles, once per chunk
                auto tile = [&](int s, const uintx4& ahr, const uintx4& ahi, const uintx4& alr, const uintx4& ali,
                                const uintx4& bhr, const uintx4& bhi, const uintx4& blr, const uintx4& bli,
                                const uintx4& m1a, const uintx4& m1b, const uintx4& m2a, const uintx4& m2b,
                                const uintx4& m3a, const uintx4& m3b) {
                    accRe[s] = mfma16(ahr, blr, accRe[s]);
                    accRe[s] = mfma16(alr, bhr, accRe[s]);
                    accRe[s] = mfma16(ahi, bli, accRe[s]);
                    accRe[s] = mfma16(ali, bhi, accRe[s]);
                    accIm[s] = mfma16(ahi, blr, accIm[s]);
                    accIm[s] = mfma16(ali, bhr, accIm[s]);
                    accIm[s] = mfma16(m1a, m1b, accIm[s]);  // -hr_I li_J
                    accIm[s] = mfma16(m2a, m2b, accIm[s]);  // -lr_I hi_J
                    accRe[s] = mfma16(ahr, bhr, accRe[s]);
                    accRe[s] = mfma16(ahi, bhi, accRe[s]);
                    accIm[s] = mfma16(ahi, bhr, accIm[s]);
                    accIm[s] = mfma16(m3a, m3b, accIm[s]);  // -hr_I hi_J
                };
                auto xop = [&](int slot, int c) { return X->op[pr][slot][c][l]; };
                if (hw == 0) {
                    // own tiles 0, 1; partner tiles 2, 3; the minus on A
                    uintx4 nhr[2], nlr[2];
#pragma unroll
                    for (int q = 0; q < 2; ++q) {
                        nhr[q] = neg_h8(hr[q]);
                        nlr[q] = neg_h8(lr[q]);
                    }
                    tile(0, hr[0], hi[0], lr[0], li[0], hr[0], hi[0], lr[0], li[0], nhr[0], li[0], nlr[0], hi[0],
                         nhr[0], hi[0]);
                    if (live1) {
                        tile(1, hr[0], hi[0], lr[0], li[0], hr[1], hi[1], lr[1], li[1], nhr[0], li[1], nlr[0], hi[1],
                             nhr[0], hi[1]);
                        tile(2, hr[1], hi[1], lr[1], li[1], hr[1], hi[1], lr[1], li[1], nhr[1], li[1], nlr[1], hi[1],
                             nhr[1], hi[1]);
                    }
                    if (live2) {
                        const uintx4 bhr = xop(1, 0), bhi = xop(1, 1), blr = xop(1, 2), bli = xop(1, 3);
                        tile(3, hr[0], hi[0], lr[0], li[0], bhr, bhi, blr, bli, nhr[0], bli, nlr[0], bhi, nhr[0], bhi);
                    }
                    if (live3) {
                        const uintx4 bhr = xop(2, 0), bhi = xop(2, 1), blr = xop(2, 2), bli = xop(2, 3);
                        tile(4, hr[0], hi[0], lr[0], li[0], bhr, bhi, blr, bli, nhr[0], bli, nlr[0], bhi, nhr[0], bhi);
                    }
                } else {
                    // own tiles 2, 3; partner tile 1; the minus on B
                    uintx4 nhi[2], nli[2];
#pragma unroll
                    for (int q = 0; q < 2; ++q) {
                        nhi[q] = neg_h8(hi[q]);
                        nli[q] = neg_h8(li[q]);
                    }
                    if (live2)
                        tile(0, hr[0], hi[0], lr[0], li[0], hr[0], hi[0], lr[0], li[0], hr[0], nli[0], lr[0], nhi[0],
                             hr[0], nhi[0]);
                    if (live3) {
                        tile(1, hr[0], hi[0], lr[0], li[0], hr[1], hi[1], lr[1], li[1], hr[0], nli[1], lr[0], nhi[1],
                             hr[0], nhi[1]);
                        tile(2, hr[1], hi[1], lr[1], li[1], hr[1], hi[1], lr[1], li[1], hr[1], nli[1], lr[1], nhi[1],
                             hr[1], nhi[1]);
                    }
                    const uintx4 ahr = xop(0, 0), ahi = xop(0, 1), alr = xop(0, 2), ali = xop(0, 3);
                    if (live2)
                        tile(3, ahr, ahi, alr, ali, hr[0], hi[0], lr[0], li[0], ahr, nli[0], alr, nhi[0], ahr, nhi[0]);
                    if (live3)
                        tile(4, ahr, ahi, alr, ali, hr[1], hi[1], lr[1], li[1], ahr, nli[1], alr, nhi[1], ahr, nhi[1]);
                }
            }
        }
        // every wave is past its reads of this block: stage the next one
        lds_sync();
        if (c0 + kSB < c_end) src_block_put(X->src, pre);
        lds_sync();
    }
    // the two pairs' partials summed in LDS in pair order; lane l holds rows
    // 4 (l / 16) + e, column l % 16 of each tile
    const int tI[5] = {hw ? 2 : 0, hw ? 2 : 0, hw ? 3 : 1, hw ? 1 : 0, hw ? 1 : 0};
    const int tJ[5] = {hw ? 2 : 0, hw ? 3 : 1, hw ? 3 : 1, hw ? 2 : 2, hw ? 3 : 3};
    for (int pp = 0; pp < 2; ++pp) {
        if (pr == pp) {
#pragma unroll
            for (int s = 0; s < 5; ++s)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int ri = 16 * tI[s] + 4 * g + e, cj = 16 * tJ[s] + r16;
                    float2 v = make_float2(accRe[s][e], accIm[s][e]);
                    if (pp > 0) {
                        const float2 o = s_R[ri * MP + cj];
                        v = make_float2(o.x + v.x, o.y + v.y);
                    }
                    s_R[ri * MP + cj] = v;
                }
        }
        lds_sync();
    }
    const float unscale = 1.0f / (scale * scale);  // exact: scale is a power of two
    float2* R = a.R + (uint64_t)blockIdx.x * MP * MP;
    for (int e = t; e < MP * MP; e += kHB) {
        const int ri = e / MP, cj = e % MP;
        const bool up = ri / 16 <= cj / 16;
        const float2 v = up ? s_R[e] : s_R[cj * MP + ri];
        R[e] = make_float2(v.x * unscale, up ? v.y * unscale : -v.y * unscale);
    }
}

// ---------------------------------------------------------------- kernel C
// q = B p over a range of source chunks: a.split workgroups per instance,
// each with the chunk range kernel A's workgroup of the same index had.
// pass kQGamma (MIXED_CLOSING only): q from the iterated p, then the range's
//   partial gamma_i = sum_j G_ij q_j into a.gam;
// pass kQFinal: p = a.p, or in MIXED_CLOSING p = gamma / |gamma|^2 p0^2 with
//   gamma the sum of the partials in order; q_j -> amplitude and phase.
enum { kQGamma = 0, kQFinal = 1 };

// Kernel C's LDS. s_qp and s_q are double-buffered by chunk parity: a buffer
// is written again two chunks later, after a barrier that every reader of its
// previous contents has reached.
struct QLds {
    float2 qp[2][8][kKC];
    float2 q[2][kKC];
    float2 qall[8][kKC];  // kQFinal: q of up to 8 chunks, written out together
    float2 p[kMaxMP];
};

// Workgroup blk of the bucket (instance blk / split, chunk range blk % split).
template <int MP, bool H16>
__device__ __forceinline__ void gspat_q_run(const HoloArgs& a, uint32_t blk, int closing, int pass, QLds& L) {
    auto& s_qp = L.qp;
    auto& s_q = L.q;
    auto& s_qall = L.qall;
    auto& s_p = L.p;
    const int t = threadIdx.x;
    const int grp = t >> 5, jj = t & 31;
    const uint32_t k = blk / a.split, part = blk % a.split, b = a.inst[k];
    const uint32_t n_chunks = a.n_packed / kKC;
    const uint32_t c_begin = part * n_chunks / a.split, c_end = (part + 1) * n_chunks / a.split;
    const uint32_t base = a.offsets[b], M = a.offsets[b + 1] - base;
    for (int i = t; i < MP; i += kHB) {
        if (pass == kQFinal && closing) {
            float2 gam = make_float2(0.f, 0.f);
            for (uint32_t sp = 0; sp < a.split; ++sp) {
                const float2 v = a.gam[((uint64_t)k * a.split + sp) * MP + i];
                gam.x += v.x;
                gam.y += v.y;
            }
            const float p0 = i < (int)M ? a.amps[base + i] : 0.f;
            const float inv = p0 * p0 / (gam.x * gam.x + gam.y * gam.y);
            s_p[i] = i < (int)M ? make_float2(gam.x * inv, gam.y * inv) : make_float2(0.f, 0.f);
        } else {
            s_p[i] = a.p[(uint64_t)k * MP + i];
        }
    }
    lds_sync();
    // the fused kernel A's scale of H (a power of two)
    const float hscale = fmaxf(1.0f, exp2f(floorf(log2f((float)M)) - 1.0f));

    float2 gam[MP / 8];
#pragma unroll
    for (int r = 0; r < MP / 8; ++r) gam[r] = make_float2(0.f, 0.f);
    // chunks loaded ahead of their use: H16 holds 9 words a chunk and reads
    // 4 bytes per element, so it keeps three chunks in flight to cover the
    // memory latency; FP32 G (17 words, 8 bytes) keeps one
    constexpr int PD = H16 ? 3 : 1;
    constexpr uint32_t kElem = H16 ? 4 : 8;
    // H16: rows < M of the instance's G words (MP n_packed kElem < 2^32 is checked on the host)
    const __amdgpu_buffer_rsrc_t gr = __builtin_amdgcn_make_buffer_rsrc(
        reinterpret_cast<char*>(a.G) + (uint64_t)k * MP * a.n_packed * kElem, (short)0,
        (int)(M * a.n_packed * kElem), 0x00020000);
    // one chunk of columns from its G tile (FP32, H16 words widened) and d_j
    auto run_chunk = [&](const float2 (&g)[MP / 8], float dj, uint32_t chunk) {
        const uint32_t j0 = chunk * kKC;
        const int pb = (chunk - c_begin) & 1, slot = (chunk - c_begin) % 8;
        // q_j = w_j sum_i conj(G_ij) p_i: this row group's partial, then the 8
        // groups in a fixed order; w_j = 1 / d_j^2 (FP32 G) or 1 / (s d_j) (H16,
        // the copy holds H = s G / d_j)
        float2 pq = make_float2(0.f, 0.f);
#pragma unroll
        for (int r = 0; r < MP / 8; ++r) {
            const float2 pv = s_p[grp + 8 * r];
            pq.x = fmaf(g[r].x, pv.x, fmaf(g[r].y, pv.y, pq.x));
            pq.y = fmaf(g[r].x, pv.y, fmaf(-g[r].y, pv.x, pq.y));
        }
        s_qp[pb][grp][jj] = pq;
        lds_sync();
        if (t < kKC) {
            float2 q = make_float2(0.f, 0.f);
#pragma unroll
            for (int q8 = 0; q8 < 8; ++q8) {
                q.x += s_qp[pb][q8][t].x;
                q.y += s_qp[pb][q8][t].y;
            }
            const float w = (j0 + t < a.n) ? (H16 ? 1.0f / (hscale * dj) : 1.0f / (dj * dj)) : 0.f;
            q = make_float2(q.x * w, q.y * w);
            if (pass == kQGamma)
                s_q[pb][t] = q;
            else
                s_qall[slot][t] = q;
        }
        if (pass == kQFinal && (slot == 7 || chunk + 1 == c_end)) {
            // amplitude and phase of the group's columns, one per thread (slot
            // sl holds chunk - (slot - sl)); a slot is written again in the
            // next group, after the next chunk's barrier, which every thread
            // reaches after its read here
            lds_sync();
            const int sl = t >> 5;
            const uint32_t j = j0 - (uint32_t)(slot - sl) * kKC + jj;
            if (sl <= slot && j < a.n) {
                const float2 q = s_qall[sl][jj];
                a.out_amp[(uint64_t)b * a.n + j] = hypotf(q.x, q.y);
                a.out_phase[(uint64_t)b * a.n + j] = atan2f(q.y, q.x);
            }
        }
        if (pass == kQGamma) {
            lds_sync();
            const float2 qj = s_q[pb][jj];
#pragma unroll
            for (int r = 0; r < MP / 8; ++r) {
                gam[r].x = fmaf(g[r].x, qj.x, fmaf(-g[r].y, qj.y, gam[r].x));
                gam[r].y = fmaf(g[r].x, qj.y, fmaf(g[r].y, qj.x, gam[r].y));
            }
        }
    };
    if constexpr (H16) {
        // slot u of the ring always holds chunks c_begin + u (mod PD): the
        // loop is unrolled over the slots, so no loaded register is copied to
        // another before its use; past the range a slot re-reads the last chunk
        auto clamp_chunk = [&](uint32_t c) { return c < c_end ? c : c_end - 1; };
        StoredChunk<MP, H16> cq[PD];
        if (c_begin < c_end) {
#pragma unroll
            for (int u = 0; u < PD; ++u) load_chunk<MP, H16>(a, gr, k, clamp_chunk(c_begin + u) * kKC, M, cq[u]);
        }
        for (uint32_t c0 = c_begin; c0 < c_end; c0 += PD) {
#pragma unroll
            for (int u = 0; u < PD; ++u) {
                const uint32_t chunk = c0 + u;
                float2 g[MP / 8];
                chunk_g<MP, H16>(cq[u], g);
                const float dj = cq[u].d;
                load_chunk<MP, H16>(a, gr, k, clamp_chunk(chunk + PD) * kKC, M, cq[u]);
                if (chunk < c_end) run_chunk(g, dj, chunk);  // wave-uniform
            }
        }
    } else {
        StoredChunk<MP, H16> cq;
        if (c_begin < c_end) load_chunk<MP, H16>(a, gr, k, c_begin * kKC, M, cq);
        for (uint32_t chunk = c_begin; chunk < c_end; ++chunk) {
            float2 g[MP / 8];
            chunk_g<MP, H16>(cq, g);
            const float dj = cq.d;
            if (chunk + 1 < c_end) load_chunk<MP, H16>(a, gr, k, (chunk + 1) * kKC, M, cq);
            run_chunk(g, dj, chunk);
        }
    }
    if (pass == kQGamma) {
        // reduce over the 32 columns held by the half-wave of this row group
#pragma unroll
        for (int r = 0; r < MP / 8; ++r) {
#pragma unroll
            for (int off = 16; off >= 1; off >>= 1) {
                gam[r].x += __shfl_xor(gam[r].x, off);
                gam[r].y += __shfl_xor(gam[r].y, off);
            }
        }
        if (jj == 0)
#pragma unroll
            for (int r = 0; r < MP / 8; ++r) a.gam[(uint64_t)blk * MP + grp + 8 * r] = gam[r];
    }
}

template <int MP, bool H16>
__global__ __launch_bounds__(kHB) void gspat_q(HoloArgs a, int closing, int pass) {
    __shared__ QLds L;
    gspat_q_run<MP, H16>(a, blockIdx.x, closing, pass, L);
}

// Kernels B and C of every bucket of a batch in one launch each: bucket u
// takes the workgroups first[u] .. first[u + 1] - 1 (instances for B,
// instances times split for C), buckets in the order listed.
struct BucketSet {
    uint32_t n;
    uint32_t mp[4];
    uint32_t first[5];
    const uint32_t* inst[4];
    float2* R[4];
    float2* p[4];
    float2* G[4];
    float* d[4];
    float2* gam[4];
};

// The bucket of workgroup blockIdx.x: its arguments in a, its workgroup index
// within the bucket in blk; returns the bucket's padded focus count.
__device__ __forceinline__ uint32_t set_bucket(const BucketSet& s, HoloArgs& a, uint32_t& blk) {
    uint32_t u = 0;
    blk = blockIdx.x;
    while (u + 1 < s.n && blk >= s.first[u + 1]) ++u;
    blk -= s.first[u];
    a.inst = s.inst[u];
    a.R = s.R[u];
    a.p = s.p[u];
    a.G = s.G[u];
    a.d = s.d[u];
    a.gam = s.gam[u];
    return s.mp[u];
}

// Buckets of MP <= 64 only (one thread count, the register use of MP = 64).
__global__ __launch_bounds__(kHB) void gspat_iter_set(HoloArgs a, BucketSet s, uint32_t iters, int closing) {
    __shared__ float2 s_p[2][kMaxMP];
    uint32_t blk;
    const uint32_t mp = set_bucket(s, a, blk);
    if (mp == 16)
        gspat_iter_run<16>(a, blk, iters, closing, s_p);
    else if (mp == 32)
        gspat_iter_run<32>(a, blk, iters, closing, s_p);
    else
        gspat_iter_run<64>(a, blk, iters, closing, s_p);
}

template <bool H16>
__global__ __launch_bounds__(kHB) void gspat_q_set(HoloArgs a, BucketSet s, int closing, int pass) {
    __shared__ QLds L;
    uint32_t blk;
    const uint32_t mp = set_bucket(s, a, blk);
    if (mp == 16)
        gspat_q_run<16, H16>(a, blk, closing, pass, L);
    else if (mp == 32)
        gspat_q_run<32, H16>(a, blk, closing, pass, L);
    else
        gspat_q_run<64, H16>(a, blk, closing, pass, L);
}

// ---------------------------------------------------------------- GS
// G (FP32, M x n_packed row-major) of every instance of the bucket.
template <int SRC>
__global__ __launch_bounds__(kHB) void gen_g(HoloArgs a, uint32_t g_stride_inst) {
    __shared__ std::conditional_t<SRC == kT4010A1, CellsLds, float> s_cells_mem;
    CellsLds* s_cells = reinterpret_cast<CellsLds*>(&s_cells_mem);
    if (SRC == kT4010A1) stage_cells_lds(a.cells, s_cells);
    __syncthreads();
    const uint32_t k = blockIdx.y, b = a.inst[k];
    const uint32_t j = blockIdx.x * kHB + threadIdx.x;
    if (j >= a.n_packed) return;
    const uint32_t base = a.offsets[b], M = a.offsets[b + 1] - base;
    const SrcF s = load_staged(a.staged, j);
    float2* G = a.G + (uint64_t)k * g_stride_inst;
    for (uint32_t i = 0; i < M; ++i) {
        const float* fp = a.foci + 3 * (uint64_t)(base + i);
        float re = 0.f, im = 0.f;
        add_terms_staged<SRC, 1>(s, fp, fp + 1, fp + 2, s_cells->p, s_cells->q, &re, &im);
        G[(uint64_t)i * a.n_packed + j] = make_float2(re, im);
    }
}

__device__ __forceinline__ float2 wave_sum(float2 v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        v.x += __shfl_xor(v.x, off);
        v.y += __shfl_xor(v.y, off);
    }
    return v;
}

// QLDS: q (n_packed <= kGsLdsQ) lives in LDS, so that the G q pass reads only G
// from memory; otherwise in the workspace.
constexpr uint32_t kGsLdsQ = 4096;

template <int MP, bool QLDS>
__global__ __launch_bounds__(kHB) void gs_iter(HoloArgs a, uint32_t iters, uint32_t g_stride_inst) {
    __shared__ float2 s_gam[MP];
    __shared__ float2 s_p[MP];
    __shared__ float2 s_qbuf[QLDS ? kGsLdsQ : 1];
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    const uint32_t k = blockIdx.x, b = a.inst[k];
    const uint32_t base = a.offsets[b], M = a.offsets[b + 1] - base;
    const uint32_t n = a.n, np = a.n_packed;
    const float2* G = a.G + (uint64_t)k * g_stride_inst;
    float2* q = QLDS ? s_qbuf : a.q + (uint64_t)k * np;
    // q0 = current source amplitudes (gs.rs:48-54)
    for (uint32_t j = t; j < np; j += kHB) q[j] = make_float2(j < n ? a.sources[AFH_SOURCE_STRIDE * (uint64_t)j + 6] : 0.f, 0.f);
    __syncthreads();
    for (uint32_t it = 0; it < iters; ++it) {
        // gamma = G q, a wave per row, lanes over columns
        for (uint32_t i = w; i < M; i += 4) {
            float2 acc = make_float2(0.f, 0.f);
            // unrolled so that several loads of G and q are in flight per lane
#pragma unroll 8
            for (uint32_t j = l; j < np; j += 64) {
                const float2 gv = G[(uint64_t)i * np + j], qv = q[j];
                acc.x = fmaf(gv.x, qv.x, fmaf(-gv.y, qv.y, acc.x));
                acc.y = fmaf(gv.x, qv.y, fmaf(gv.y, qv.x, acc.y));
            }
            acc = wave_sum(acc);
            if (l == 0) s_gam[i] = acc;
        }
        __syncthreads();
        if (t < (int)M) {
            const float2 g = s_gam[t];
            const float inv = a.amps[base + t] / sqrtf(g.x * g.x + g.y * g.y);
            s_p[t] = make_float2(g.x * inv, g.y * inv);
        }
        __syncthreads();
        // xi = G^H p, q = xi / |xi| q0
        for (uint32_t j = t; j < n; j += kHB) {
            float2 xi = make_float2(0.f, 0.f);
#pragma unroll 8
            for (uint32_t i = 0; i < M; ++i) {
                const float2 gv = G[(uint64_t)i * np + j], pv = s_p[i];
                xi.x = fmaf(gv.x, pv.x, fmaf(gv.y, pv.y, xi.x));
                xi.y = fmaf(gv.x, pv.y, fmaf(-gv.y, pv.x, xi.y));
            }
            const float q0 = a.sources[AFH_SOURCE_STRIDE * (uint64_t)j + 6];
            const float inv = q0 / sqrtf(xi.x * xi.x + xi.y * xi.y);
            q[j] = make_float2(xi.x * inv, xi.y * inv);
        }
        __threadfence_block();
        __syncthreads();
    }
    for (uint32_t j = t; j < n; j += kHB) {
        const float2 qv = q[j];
        a.out_amp[(uint64_t)b * n + j] = hypotf(qv.x, qv.y);
        a.out_phase[(uint64_t)b * n + j] = atan2f(qv.y, qv.x);
    }
}

// GS with one pass over G per iteration (gs.rs:56-72, the same steps in the
// same order): q_{t+1,j} = xi_{t,j} / |xi_{t,j}| q0_j depends on column j
// alone, so the pass that forms xi_t = G^H p_t column by column also forms
// gamma_{t+1} = G q_{t+1} from the same loaded values of G. Pass 0 forms
// gamma_0 = G q0; pass t + 1 forms xi_t, q_{t+1} and, before the last, gamma_{t+1};
// the last writes |q| and arg q. W waves per instance (more for small MP, so
// that a bucket of a few hundred instances still fills the CUs); wave w takes
// the 16-column groups w, w + W, ...; lane (rg, c) = (l / 16, l % 16) holds
// rows rg + 4 r of column 16 g + c, the next group's loaded during this one.
// G is read through a buffer descriptor of M rows, so rows >= M read as 0
// without a memory access. Every load is issued unconditionally and its value
// kept as loaded (columns >= n read source n - 1, past the last group the
// current one is read again; the 0 / 1 mask multiplies q0 where it is used):
// a select right after a load let the compiler move the load under a branch
// and wait for it there. xi_j sums the 4 row groups by lane exchange; gamma's
// partials are summed over a wave's 16 columns by lane exchange and over the W
// waves in LDS, in fixed order.
template <int MP, int W>
__global__ __launch_bounds__(64 * W) void gs_fused(HoloArgs a, uint32_t iters, uint32_t g_stride_inst) {
    constexpr int RPT = MP / 4;
    __shared__ float2 s_gam[W][MP];
    __shared__ float2 s_p[MP];
    const int t = threadIdx.x, w = t >> 6, l = t & 63, rg = l >> 4, c = l & 15;
    const uint32_t k = blockIdx.x, b = a.inst[k];
    const uint32_t base = a.offsets[b], M = a.offsets[b + 1] - base;  // 1 <= M <= MP
    const uint32_t n = a.n, np = a.n_packed;                          // np: a multiple of kKC
    // rows < M of the instance's G: M np 8 bytes (MP np 8 < 2^32 is checked at launch)
    const __amdgpu_buffer_rsrc_t gr = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float2*>(a.G + (uint64_t)k * g_stride_inst), (short)0, (int)(M * np * 8), 0x00020000);
    const float* src_amp = a.sources + 6;
    float2 p[RPT], g[RPT], ng[RPT];
    float nq0 = 0.f;
#pragma unroll
    for (int r = 0; r < RPT; ++r) p[r] = make_float2(0.f, 0.f);
    auto fetch = [&](uint32_t j) {
#pragma unroll
        for (int r = 0; r < RPT; ++r)
            ng[r] = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(
                                                   gr, (int)(((rg + 4 * r) * np + j) * 8), 0, 0));
        nq0 = src_amp[AFH_SOURCE_STRIDE * (uint64_t)(j < n ? j : n - 1)];
    };
    for (uint32_t pass = 0; pass <= iters; ++pass) {
        const bool first = pass == 0, last = pass == iters;
        float2 gam[RPT];
#pragma unroll
        for (int r = 0; r < RPT; ++r) gam[r] = make_float2(0.f, 0.f);
        fetch((16 * w < np ? 16 * w : 0) + c);  // a wave past np has no group
        for (uint32_t g0 = 16 * w; g0 < np; g0 += 16 * W) {
            const uint32_t j = g0 + c;
#pragma unroll
            for (int r = 0; r < RPT; ++r) g[r] = ng[r];
            const float q0 = nq0 * (j < n ? 1.f : 0.f);
            fetch((g0 + 16 * W < np ? g0 + 16 * W : g0) + c);
            float2 q = make_float2(q0, 0.f);
            if (!first) {
                float2 xi = make_float2(0.f, 0.f);
#pragma unroll
                for (int r = 0; r < RPT; ++r) {
                    xi.x = fmaf(g[r].x, p[r].x, fmaf(g[r].y, p[r].y, xi.x));
                    xi.y = fmaf(g[r].x, p[r].y, fmaf(-g[r].y, p[r].x, xi.y));
                }
                xi.x += __shfl_xor(xi.x, 16);
                xi.y += __shfl_xor(xi.y, 16);
                xi.x += __shfl_xor(xi.x, 32);
                xi.y += __shfl_xor(xi.y, 32);
                const float inv = q0 / sqrtf(xi.x * xi.x + xi.y * xi.y);
                q = j < n ? make_float2(xi.x * inv, xi.y * inv) : make_float2(0.f, 0.f);
                if (last) {
                    if (rg == 0 && j < n) {
                        a.out_amp[(uint64_t)b * n + j] = hypotf(q.x, q.y);
                        a.out_phase[(uint64_t)b * n + j] = atan2f(q.y, q.x);
                    }
                    continue;
                }
            }
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                gam[r].x = fmaf(g[r].x, q.x, fmaf(-g[r].y, q.y, gam[r].x));
                gam[r].y = fmaf(g[r].x, q.y, fmaf(g[r].y, q.x, gam[r].y));
            }
        }
        if (last) break;
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) {
                gam[r].x += __shfl_xor(gam[r].x, off);
                gam[r].y += __shfl_xor(gam[r].y, off);
            }
        }
        if (c == 0) {
#pragma unroll
            for (int r = 0; r < RPT; ++r) s_gam[w][rg + 4 * r] = gam[r];
        }
        __syncthreads();
        if (t < (int)M) {
            float2 gm = s_gam[0][t];
#pragma unroll
            for (int v = 1; v < W; ++v) {
                gm.x += s_gam[v][t].x;
                gm.y += s_gam[v][t].y;
            }
            const float inv = a.amps[base + t] / sqrtf(gm.x * gm.x + gm.y * gm.y);
            s_p[t] = make_float2(gm.x * inv, gm.y * inv);
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const uint32_t i = rg + 4 * r;
            p[r] = i < M ? s_p[i] : make_float2(0.f, 0.f);
        }
    }
}

// ---------------------------------------------------------------- naive
template <int SRC>
__global__ __launch_bounds__(kHB) void naive_q(HoloArgs a) {
    __shared__ std::conditional_t<SRC == kT4010A1, CellsLds, float> s_cells_mem;
    CellsLds* s_cells = reinterpret_cast<CellsLds*>(&s_cells_mem);
    if (SRC == kT4010A1) stage_cells_lds(a.cells, s_cells);
    __syncthreads();
    const uint32_t b = blockIdx.y;
    const uint32_t j = blockIdx.x * kHB + threadIdx.x;
    if (j >= a.n) return;
    const uint32_t base = a.offsets[b], M = a.offsets[b + 1] - base;
    const SrcF s = load_staged(a.staged, j);
    float2 q = make_float2(0.f, 0.f);
    for (uint32_t i = 0; i < M; ++i) {
        const float* fp = a.foci + 3 * (uint64_t)(base + i);
        float re = 0.f, im = 0.f;
        add_terms_staged<SRC, 1>(s, fp, fp + 1, fp + 2, s_cells->p, s_cells->q, &re, &im);
        const float p = a.amps[base + i];
        // q += conj(G_ij) p_i
        q.x = fmaf(re, p, q.x);
        q.y = fmaf(-im, p, q.y);
    }
    a.out_amp[(uint64_t)b * a.n + j] = hypotf(q.x, q.y);
    a.out_phase[(uint64_t)b * a.n + j] = atan2f(q.y, q.x);
}

// ---------------------------------------------------------------- greedy
// One workgroup per instance, sources in index order (greedy.rs:60-82): the M
// terms of source i in parallel, one lane per candidate summing
// |a_j - |G_ji c + cache_j|| over j in increasing order, the first minimum
// over the candidates (strict <: ties keep the lower index, as the
// reference's scan does) by a fixed-order wave reduction, then the cache
// update and the phase arg(c) from the host's table. cand / cand_phase: div
// candidates e^{i 2 pi idx / div} and their arguments.
template <int SRC>
__global__ __launch_bounds__(kHB) void greedy_kernel(HoloArgs a, const float2* __restrict__ cand,
                                                     const float* __restrict__ cand_phase, uint32_t div) {
    __shared__ std::conditional_t<SRC == kT4010A1, CellsLds, float> s_cells_mem;
    CellsLds* s_cells = reinterpret_cast<CellsLds*>(&s_cells_mem);
    __shared__ float s_fx[kGreedyMaxM], s_fy[kGreedyMaxM], s_fz[kGreedyMaxM], s_a[kGreedyMaxM];
    __shared__ float2 s_g[kGreedyMaxM], s_cache[kGreedyMaxM];
    __shared__ float2 s_cand[kGreedyMaxDiv];
    __shared__ float s_v[kGreedyMaxDiv];
    __shared__ int s_best;
    const int t = threadIdx.x;
    const uint32_t b = blockIdx.x;
    const uint32_t base = a.offsets[b], M = a.offsets[b + 1] - base;
    if (SRC == kT4010A1) stage_cells_lds(a.cells, s_cells);
    for (uint32_t j = t; j < M; j += kHB) {
        s_fx[j] = a.foci[3 * (uint64_t)(base + j)];
        s_fy[j] = a.foci[3 * (uint64_t)(base + j) + 1];
        s_fz[j] = a.foci[3 * (uint64_t)(base + j) + 2];
        s_a[j] = a.amps[base + j];
        s_cache[j] = make_float2(0.f, 0.f);
    }
    for (uint32_t c = t; c < div; c += kHB) s_cand[c] = cand[c];
    __syncthreads();
    for (uint32_t i = 0; i < a.n; ++i) {
        const SrcF s = load_staged(a.staged, i);
        for (uint32_t j = t; j < M; j += kHB) {
            float re = 0.f, im = 0.f;
            add_terms_staged<SRC, 1>(s, &s_fx[j], &s_fy[j], &s_fz[j], s_cells->p, s_cells->q, &re, &im);
            s_g[j] = make_float2(re, im);
        }
        __syncthreads();
        for (uint32_t c = t; c < div; c += kHB) {
            const float2 cc = s_cand[c];
            float v = 0.f;
            for (uint32_t j = 0; j < M; ++j) {
                const float2 g = s_g[j], ca = s_cache[j];
                const float re = g.x * cc.x - g.y * cc.y + ca.x;
                const float im = g.x * cc.y + g.y * cc.x + ca.y;
                v += fabsf(s_a[j] - hypotf(re, im));
            }
            s_v[c] = v;
        }
        __syncthreads();
        if (t < 64) {
            // first minimum: scan this lane's candidates in increasing order,
            // then combine lanes preferring the smaller value, then the lower index
            float bv = INFINITY;
            int bi = (int)div;
            for (uint32_t c = t; c < div; c += 64) {
                const float v = s_v[c];
                if (v < bv) {
                    bv = v;
                    bi = (int)c;
                }
            }
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
                const float ov = __shfl_xor(bv, off);
                const int oi = __shfl_xor(bi, off);
                if (ov < bv || (ov == bv && oi < bi)) {
                    bv = ov;
                    bi = oi;
                }
            }
            if (t == 0) s_best = bi < (int)div ? bi : 0;  // every v NaN or infinite: index 0
        }
        __syncthreads();
        const int best = s_best;
        const float2 cb = s_cand[best];
        for (uint32_t j = t; j < M; j += kHB) {
            const float2 g = s_g[j];
            s_cache[j].x += g.x * cb.x - g.y * cb.y;
            s_cache[j].y += g.x * cb.y + g.y * cb.x;
        }
        if (t == 0) a.out_phase[(uint64_t)b * a.n + i] = cand_phase[best];
        __syncthreads();
    }
}

// ---------------------------------------------------------------- host side
// Kernel A workgroups per GS-PAT instance: each takes a contiguous range of
// source chunks and writes a partial R that kernel B sums. A function of the
// array size only (the largest power of two <= chunks / 16, at most 8), so an
// instance is computed the same way in any batch. One workgroup per instance
// leaves most of the chip idle or waiting on a tail round (R's chunk loop is
// serial, with four barriers per chunk).
// Kernel A / C workgroups per instance (source-chunk ranges) of a bucket: a
// function of the array and the bucket only, so an instance's result does
// not depend on its batch. MP = 128 takes twice the ranges (at most 8): its
// workgroups of 8 waves are the longest, and a small batch of them (as a
// server forms) is otherwise latency-bound on their chunk loop.
static uint32_t gspat_split(uint32_t n_packed, int mp) {
    const uint32_t chunks = n_packed / kKC;
    uint32_t s = 1;
    while (s < 8 && 2 * s * 16 <= chunks) s *= 2;
    return mp == 128 ? std::min<uint32_t>(8, 2 * s) : s;
}

size_t gspat_workspace_bound(uint32_t n_packed, uint32_t batch) {
    // holo_solve's layout with every instance in the MP = 128 bucket
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t B = batch, np = n_packed, mp = 128, sp = gspat_split(n_packed, 128);
    return al((B + 1) * sizeof(uint32_t)) + al(B * sizeof(uint32_t)) + al(np * AFH_SOURCE_STRIDE * sizeof(float)) +
           al(B * mp * mp * sp * sizeof(float2)) + al(B * mp * sizeof(float2)) + al(B * mp * np * sizeof(float2)) +
           al(B * mp * sp * sizeof(float2)) + al(B * np * sizeof(float));
}

static int padded_m(uint32_t m) {
    if (m <= 16) return 16;
    if (m <= 32) return 32;
    if (m <= 64) return 64;
    if (m <= 128) return 128;
    return 0;
}

// A kernel's host-side handle, as hipLaunchKernel and graph nodes take it.
template <class F>
static const void* kfn(F* f) {
    return reinterpret_cast<const void*>(f);
}

// Launch by launch on `st`.
static hipError_t issue_direct(const KChain& ch, hipStream_t st) {
    for (int i = 0; i < ch.n; ++i) {
        const KLaunch& l = ch.k[i];
        void* args[4];
        for (int j = 0; j < l.nargs; ++j) args[j] = const_cast<unsigned char*>(l.blob) + l.off[j];
        const hipError_t e = hipLaunchKernel(l.fn, l.grid, l.block, args, 0, st);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Cached chain graphs per context; the least recently used idle one makes
// room for a new shape.
constexpr size_t kChainGraphs = 32;

static bool same_shape(const ChainGraph& g, const KChain& ch) {
    if (g.n != ch.n) return false;
    for (int i = 0; i < ch.n; ++i)
        if (g.fn[i] != ch.k[i].fn || g.block[i].x != ch.k[i].block.x || g.block[i].y != ch.k[i].block.y ||
            g.block[i].z != ch.k[i].block.z)
            return false;
    return true;
}

// Node parameters of launch l; args receives the argument pointers (the
// runtime copies the values when the node is added or set).
static hipKernelNodeParams node_params(const KLaunch& l, void* (&args)[4]) {
    for (int j = 0; j < l.nargs; ++j) args[j] = const_cast<unsigned char*>(l.blob) + l.off[j];
    hipKernelNodeParams p;
    memset(&p, 0, sizeof(p));
    p.func = const_cast<void*>(l.fn);
    p.gridDim = l.grid;
    p.blockDim = l.block;
    p.sharedMemBytes = 0;
    p.kernelParams = args;
    p.extra = nullptr;
    return p;
}

static hipError_t build_chain_graph(const KChain& ch, ChainGraph& g) {
    g.n = ch.n;
    hipError_t e = hipGraphCreate(&g.graph, 0);
    for (int i = 0; e == hipSuccess && i < ch.n; ++i) {
        g.fn[i] = ch.k[i].fn;
        g.block[i] = ch.k[i].block;
        void* args[4];
        const hipKernelNodeParams p = node_params(ch.k[i], args);
        e = hipGraphAddKernelNode(&g.node[i], g.graph, i ? &g.node[i - 1] : nullptr, i ? 1 : 0, &p);
    }
    if (e == hipSuccess) e = hipGraphInstantiateWithFlags(&g.exec, g.graph, 0);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&g.done, hipEventDisableTiming);
    return e;
}

// The chain as one graph launch on `st`: a cached graph of the same shape whose
// last launch has completed gets this chain's grids and arguments, otherwise a
// new graph is built (and cached). Without an idle graph to replace in a full
// cache, the chain is launched directly.
static hipError_t issue_graph(Context* ctx, const KChain& ch, hipStream_t st) {
    ChainGraph* use = nullptr;
    for (auto& g : ctx->chain_graphs) {
        if (!same_shape(g, ch)) continue;
        const hipError_t q = hipEventQuery(g.done);
        if (q == hipErrorNotReady) continue;
        if (q != hipSuccess) return q;
        bool ok = true;
        for (int i = 0; ok && i < ch.n; ++i) {
            void* args[4];
            const hipKernelNodeParams p = node_params(ch.k[i], args);
            ok = hipGraphExecKernelNodeSetParams(g.exec, g.node[i], &p) == hipSuccess;
        }
        if (!ok) {
            // a graph whose nodes refuse the new parameters is rebuilt
            (void)hipGetLastError();
            ctx->drop_graph(g);
            continue;
        }
        use = &g;
        break;
    }
    if (!use) {
        // a free slot: a dropped entry, room in the cache, or the least recently
        // used idle graph
        ChainGraph* slot = nullptr;
        for (auto& g : ctx->chain_graphs)
            if (!g.exec) {
                slot = &g;
                break;
            }
        if (!slot && ctx->chain_graphs.size() < kChainGraphs) {
            ctx->chain_graphs.emplace_back();
            slot = &ctx->chain_graphs.back();
        }
        if (!slot) {
            for (auto& g : ctx->chain_graphs)
                if (hipEventQuery(g.done) == hipSuccess && (!slot || g.last_use < slot->last_use)) slot = &g;
            (void)hipGetLastError();
            if (!slot) return issue_direct(ch, st);
            ctx->drop_graph(*slot);
        }
        const hipError_t e = build_chain_graph(ch, *slot);
        if (e != hipSuccess) {
            ctx->drop_graph(*slot);
            return e;
        }
        ++ctx->graph_builds;
        use = slot;
    }
    hipError_t e = hipGraphLaunch(use->exec, st);
    if (e == hipSuccess) e = hipEventRecord(use->done, st);
    use->last_use = ++ctx->graph_clock;
    ++ctx->graph_launches;
    return e;
}

static hipError_t issue_chain(Context* ctx, const KChain& ch, hipStream_t st) {
    return ctx && ctx->graphs ? issue_graph(ctx, ch, st) : issue_direct(ch, st);
}

// The kernel chains of one GS-PAT bucket, kept as records so that each can be
// issued on a stream launch by launch or as one hipGraph (KChain,
// issue_chain): ch the kernel A launches (one or two), tail the iterations
// and the final pass (two or three launches).
template <int MP>
static void record_gspat_bucket(KChain& ch, KChain& tail, const HoloArgs& a, int src, int precision, uint32_t nb,
                                uint32_t iters) {
    const dim3 grid(nb), grid_a(nb * a.split), blk(kHB);
    const dim3 grid_g(a.n_packed / kHB + (a.n_packed % kHB ? 1 : 0), nb);
    const bool t4 = src == kT4010A1;
    if (precision == AFH_HOLO_FP32) {
        if constexpr (MP >= 32) {
            // A1 (FP32 G, one thread per source column) then A2 on FP32 MFMA
            ch.add(t4 ? kfn(gspat_gen<kT4010A1, MP>) : kfn(gspat_gen<kSphere, MP>), grid_g, blk, a);
            ch.add(kfn(gspat_r32<MP>), grid_a, blk, a);
        } else {
            ch.add(t4 ? kfn(gspat_r_fp32<kT4010A1, MP>) : kfn(gspat_r_fp32<kSphere, MP>), grid_a, blk, a);
        }
    } else {
        // kernel A: G, d_j and the split-FP16 MFMA R in one pass; MIXED keeps
        // the FP16 H for kernel C, the other modes FP32 G
        const bool g32 = precision != AFH_HOLO_MIXED;
        if constexpr (MP == 64) {
            const void* fn = t4 ? (g32 ? kfn(gspat_pair<kT4010A1, true>) : kfn(gspat_pair<kT4010A1, false>))
                                : (g32 ? kfn(gspat_pair<kSphere, true>) : kfn(gspat_pair<kSphere, false>));
            ch.add(fn, grid_a, blk, a);
        } else if constexpr (MP <= 32) {
            const void* fn = t4 ? (g32 ? kfn(gspat_wave<kT4010A1, MP, true>) : kfn(gspat_wave<kT4010A1, MP, false>))
                                : (g32 ? kfn(gspat_wave<kSphere, MP, true>) : kfn(gspat_wave<kSphere, MP, false>));
            ch.add(fn, grid_a, blk, a);
        } else {
            const void* fn = t4 ? (g32 ? kfn(gspat_fused<kT4010A1, MP, true>) : kfn(gspat_fused<kT4010A1, MP, false>))
                                : (g32 ? kfn(gspat_fused<kSphere, MP, true>) : kfn(gspat_fused<kSphere, MP, false>));
            ch.add(fn, grid_a, dim3(4 * MP), a);
        }
    }
    const int closing = precision == AFH_HOLO_MIXED_CLOSING ? 1 : 0;
    tail.add(kfn(gspat_iter<MP>), grid, blk, a, iters, closing);
    if (closing) tail.add(kfn(gspat_q<MP, false>), grid_a, blk, a, closing, (int)kQGamma);
    if (precision == AFH_HOLO_MIXED)
        tail.add(kfn(gspat_q<MP, true>), grid_a, blk, a, closing, (int)kQFinal);
    else
        tail.add(kfn(gspat_q<MP, false>), grid_a, blk, a, closing, (int)kQFinal);
}

template <int MP>
static hipError_t launch_gs_bucket(HoloArgs a, int src, uint32_t nb, uint32_t iters, uint32_t g_stride,
                                   hipStream_t st) {
    const dim3 ggrid((a.n_packed + kHB - 1) / kHB, nb);
    if (src == kT4010A1)
        hipLaunchKernelGGL((gen_g<kT4010A1>), ggrid, dim3(kHB), 0, st, a, g_stride);
    else
        hipLaunchKernelGGL((gen_g<kSphere>), ggrid, dim3(kHB), 0, st, a, g_stride);
    if constexpr (MP <= 64) {
        // gs_fused addresses an instance's G with 32-bit buffer offsets
        if ((uint64_t)MP * a.n_packed * sizeof(float2) <= 0xFFFFFFFFull) {
            // waves per instance: 16 at MP = 16 (1.52 against 2.56 ms with 4 for a
            // bucket of 253 instances); 4 at MP = 32 (8 measured slower)
            constexpr int W = MP == 16 ? 16 : 4;
            hipLaunchKernelGGL((gs_fused<MP, W>), dim3(nb), dim3(64 * W), 0, st, a, iters, g_stride);
            return hipGetLastError();
        }
    }
    if (a.n_packed <= kGsLdsQ)
        hipLaunchKernelGGL((gs_iter<MP, true>), dim3(nb), dim3(kHB), 0, st, a, iters, g_stride);
    else
        hipLaunchKernelGGL((gs_iter<MP, false>), dim3(nb), dim3(kHB), 0, st, a, iters, g_stride);
    return hipGetLastError();
}

bool Context::reserve_staging(size_t bytes) {
    if (upload_done) {
        // the previous call's uploads must have left the staging buffer
        if (!hip_ok(hipEventSynchronize(upload_done), "hipEventSynchronize")) return false;
    }
    return staging.reserve(bytes);
}

// Workspace plan of one holo_solve call: instances per padded-focus bucket and
// the byte offsets of offsets | instance lists | staged sources | R | p | G |
// q | d | Greedy candidates.
struct HoloPlan {
    uint32_t nb[4] = {0, 0, 0, 0};  // instances of the MP = 16, 32, 64, 128 buckets
    uint32_t div = 0;               // Greedy phase candidates
    size_t o_off = 0, o_list = 0, o_stg = 0, o_r = 0, o_p = 0, o_g = 0, o_q = 0, o_d = 0, o_cand = 0, total = 0;
};

constexpr int kBucketMP[4] = {16, 32, 64, 128};

static int bucket_of(const HoloBatch& hb, uint32_t b) {
    const uint32_t m = hb.offsets_host[b + 1] - hb.offsets_host[b];
    const int mp = padded_m(m);
    for (int q = 0; q < 4; ++q)
        if (kBucketMP[q] == mp) return q;
    return -1;
}

static int32_t plan_holo(const HoloBatch& hb, HoloPlan& pl) {
    const uint32_t B = hb.batch, np = hb.n_packed;
    for (uint32_t b = 0; b < B; ++b) {
        const uint32_t m = hb.offsets_host[b + 1] - hb.offsets_host[b];
        if (m == 0 || hb.offsets_host[b + 1] < hb.offsets_host[b]) {
            set_error("holo: an instance has no foci");
            return AFH_ERR_ARG;
        }
        const int q = bucket_of(hb, b);
        if (q < 0) {
            set_error("holo: more than 128 foci in one instance is not supported");
            return AFH_ERR_UNSUPPORTED;
        }
        if (hb.algo != kAlgoGreedy) ++pl.nb[q];  // Greedy runs without buckets
    }
    size_t r_elems = 0, p_elems = 0, g_elems = 0, q_elems = 0, d_elems = 0;
    for (int q = 0; q < 4; ++q) {
        const size_t nb = pl.nb[q], mp = kBucketMP[q];
        if (hb.algo == kAlgoGSPAT) {
            r_elems += nb * mp * mp * gspat_split(np, (int)mp);
            p_elems += nb * mp;
            g_elems += nb * mp * np;
            d_elems += nb * np;
            q_elems += nb * mp * gspat_split(np, (int)mp);  // partial gammas
        } else if (hb.algo == kAlgoGS) {
            g_elems += nb * mp * np;
            q_elems += nb * np;
        }
    }
    pl.div = hb.algo == kAlgoGreedy ? hb.phase_div : 0;
    if (hb.algo == kAlgoGreedy && (pl.div == 0 || pl.div > kGreedyMaxDiv)) {
        set_error("greedy: phase_div must be in 1..4096");
        return AFH_ERR_ARG;
    }
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    pl.o_off = 0;
    pl.o_list = align((B + 1) * sizeof(uint32_t));
    pl.o_stg = pl.o_list + align(B * sizeof(uint32_t));
    pl.o_r = pl.o_stg + align((size_t)np * AFH_SOURCE_STRIDE * sizeof(float));
    pl.o_p = pl.o_r + align(r_elems * sizeof(float2));
    pl.o_g = pl.o_p + align(p_elems * sizeof(float2));
    pl.o_q = pl.o_g + align(g_elems * sizeof(float2));
    pl.o_d = pl.o_q + align(q_elems * sizeof(float2));
    pl.o_cand = pl.o_d + align(d_elems * sizeof(float));
    pl.total = pl.o_cand + align((size_t)pl.div * (sizeof(float2) + sizeof(float)));
    return AFH_OK;
}

// Calls put(b) for every instance in bucket order (the instance lists).
template <class F>
static void for_each_listed(const HoloBatch& hb, F&& put) {
    if (hb.algo == kAlgoGreedy) return;
    for (int q = 0; q < 4; ++q)
        for (uint32_t b = 0; b < hb.batch; ++b)
            if (bucket_of(hb, b) == q) put(b);
}

// greedy.rs:23-29: Complex::new(0, 2 PI i / div).exp() in f32, and its arg().
static void greedy_candidate(uint32_t q, uint32_t div, float2& c, float& ph) {
    const float th = 2.0f * 3.14159265358979323846f * (float)q / (float)div;
    c = make_float2(expf(0.0f) * cosf(th), expf(0.0f) * sinf(th));
    ph = atan2f(c.y, c.x);
}

bool Context::ensure_aux() {
    if (fork) return true;
    for (int i = 0; i < 4; ++i) {
        bool made = false;
        if (!cu_mask.empty()) {
            made = hipExtStreamCreateWithCUMask(&aux[i], (uint32_t)cu_mask.size(), cu_mask.data()) == hipSuccess;
            if (!made) (void)hipGetLastError();
        }
        if (!made && !hip_ok(hipStreamCreateWithFlags(&aux[i], hipStreamNonBlocking), "hipStreamCreate")) return false;
        if (!hip_ok(hipEventCreateWithFlags(&aux_done[i], hipEventDisableTiming), "hipEventCreate")) return false;
    }
    return hip_ok(hipEventCreateWithFlags(&fork, hipEventDisableTiming), "hipEventCreate");
}

// GS batches up to this many instances run their buckets side by side on the
// context's auxiliary streams (they are independent). GS-PAT batches of any
// size run on the caller's stream: the buckets' kernel A in turn, largest MP
// first, then kernels B and C once for all buckets up to MP = 64
// (gspat_iter_set, gspat_q_set), one launch each instead of one per bucket,
// the small buckets' latency-bound iterations under the large bucket's. The
// MP = 128 bucket keeps its per-bucket chain: its kernels B and C hold 166-168
// VGPRs, which in a merged launch would set every bucket's occupancy. (Served
// GS-PAT batches forked over the auxiliary streams had a p50 of 1.8-2.5 ms
// from 40K to 400K solves/s against 0.57-1.24 ms on one stream: a server's
// lanes, each with its own streams, already run batches side by side, and the
// forked chains shared the process's 4 hardware queues.)
constexpr uint32_t kForkMaxBatch = 256;

// Launches the solver on a workspace whose offsets, instance lists and Greedy
// candidates are already on the device at wbase (laid out by `pl`). With a
// context (ctx not null), a small GS-PAT or GS batch runs its buckets on the
// context's auxiliary streams.
static int32_t run_holo(const HoloBatch& hb, const HoloPlan& pl, char* wbase, const DirCell* cells, hipStream_t st,
                        Context* ctx = nullptr) {
    const uint32_t B = hb.batch, np = hb.n_packed;
    HoloArgs a;
    memset(&a, 0, sizeof(a));
    a.sources = hb.sources;
    a.n = hb.n;
    a.n_packed = np;
    a.foci = hb.foci;
    a.amps = hb.amps;
    a.offsets = reinterpret_cast<const uint32_t*>(wbase + pl.o_off);
    a.staged = reinterpret_cast<const float4*>(wbase + pl.o_stg);
    a.cells = cells;
    a.out_amp = hb.out_amp;
    a.out_phase = hb.out_phase;
    a.split = gspat_split(np, 64);  // per GS-PAT bucket below
    if (!a.cells) {
        set_error("holo: directivity table allocation failed");
        return AFH_ERR_HIP;
    }
    hipLaunchKernelGGL(stage_holo_sources, dim3((np + kHB - 1) / kHB), dim3(kHB), 0, st, hb.sources, np,
                       reinterpret_cast<float4*>(wbase + pl.o_stg));
    if (!hip_ok(hipGetLastError(), "holo staging launch")) return AFH_ERR_HIP;
    int buckets = 0;
    for (int q = 0; q < 4; ++q) buckets += pl.nb[q] > 0;
    const bool fork = ctx && buckets > 1 && B <= kForkMaxBatch && hb.algo == kAlgoGS;
    int small_buckets = 0;  // buckets that kernels B and C of a merged batch take together
    for (int q = 0; q < 4; ++q) small_buckets += pl.nb[q] > 0 && kBucketMP[q] <= 64;
    const bool merged = small_buckets > 1 && hb.algo == kAlgoGSPAT;
    if (fork && (!ctx->ensure_aux() || !hip_ok(hipEventRecord(ctx->fork, st), "hipEventRecord"))) return AFH_ERR_HIP;
    BucketSet set_b, set_c;  // merged: the buckets' workgroup ranges for kernels B and C
    memset(&set_b, 0, sizeof(set_b));
    memset(&set_c, 0, sizeof(set_c));
    const hipStream_t caller = st;
    // each bucket's workspace ranges, in bucket order
    size_t list_at[4], r_at[4], p_at[4], g_at[4], q_at[4], d_at[4];
    {
        size_t list_pos = 0, r_pos = 0, p_pos = 0, g_pos = 0, q_pos = 0, d_pos = 0;
        for (int q = 0; q < 4; ++q) {
            const size_t nb = pl.nb[q], mp = kBucketMP[q];
            list_at[q] = list_pos;
            r_at[q] = r_pos;
            p_at[q] = p_pos;
            g_at[q] = g_pos;
            q_at[q] = q_pos;
            d_at[q] = d_pos;
            list_pos += nb;
            if (hb.algo == kAlgoGSPAT) {
                r_pos += nb * mp * mp * gspat_split(np, (int)mp);
                p_pos += nb * mp;
                g_pos += nb * mp * np;
                d_pos += nb * np;
                q_pos += nb * mp * gspat_split(np, (int)mp);
            } else if (hb.algo == kAlgoGS) {
                g_pos += nb * mp * np;
                q_pos += nb * np;
            }
        }
    }
    for (int qi = 0; qi < 4; ++qi) {
        const int q = merged ? 3 - qi : qi;  // merged: largest buckets first
        const uint32_t nb = pl.nb[q];
        const int mp = kBucketMP[q];
        if (nb == 0) continue;
        if (fork) {
            st = ctx->aux[q];
            if (!hip_ok(hipStreamWaitEvent(st, ctx->fork, 0), "hipStreamWaitEvent")) return AFH_ERR_HIP;
        }
        a.inst = reinterpret_cast<const uint32_t*>(wbase + pl.o_list) + list_at[q];
        a.R = reinterpret_cast<float2*>(wbase + pl.o_r) + r_at[q];
        a.p = reinterpret_cast<float2*>(wbase + pl.o_p) + p_at[q];
        a.G = reinterpret_cast<float2*>(wbase + pl.o_g) + g_at[q];
        a.q = reinterpret_cast<float2*>(wbase + pl.o_q) + q_at[q];
        a.d = reinterpret_cast<float*>(wbase + pl.o_d) + d_at[q];
        a.gam = reinterpret_cast<float2*>(wbase + pl.o_q) + q_at[q];
        hipError_t e = hipSuccess;
        if (hb.algo == kAlgoGSPAT) {
            a.split = gspat_split(np, mp);
            // kernel C reads an instance's G through 32-bit buffer offsets
            if ((uint64_t)mp * np * sizeof(float2) > 0xFFFFFFFFull) {
                set_error("holo: GS-PAT supports at most 2^32 / (8 * 128) padded sources");
                return AFH_ERR_UNSUPPORTED;
            }
            KChain ch, tail;
            switch (mp) {
                case 16: record_gspat_bucket<16>(ch, tail, a, hb.source_type, hb.precision, nb, hb.iterations); break;
                case 32: record_gspat_bucket<32>(ch, tail, a, hb.source_type, hb.precision, nb, hb.iterations); break;
                case 64: record_gspat_bucket<64>(ch, tail, a, hb.source_type, hb.precision, nb, hb.iterations); break;
                default:
                    record_gspat_bucket<128>(ch, tail, a, hb.source_type, hb.precision, nb, hb.iterations);
                    break;
            }
            e = issue_chain(ctx, ch, st);
            if (merged && mp <= 64) {
                for (BucketSet* bs : {&set_b, &set_c}) {
                    const uint32_t u = bs->n++;
                    bs->mp[u] = (uint32_t)mp;
                    bs->first[u + 1] = bs->first[u] + (bs == &set_b ? nb : nb * a.split);
                    bs->inst[u] = a.inst;
                    bs->R[u] = a.R;
                    bs->p[u] = a.p;
                    bs->G[u] = a.G;
                    bs->d[u] = a.d;
                    bs->gam[u] = a.gam;
                }
            } else if (e == hipSuccess) {
                e = issue_chain(ctx, tail, st);
            }
        } else if (hb.algo == kAlgoGS) {
            const uint32_t gstride = (uint32_t)(mp * np);
            switch (mp) {
                case 16: e = launch_gs_bucket<16>(a, hb.source_type, nb, hb.iterations, gstride, st); break;
                case 32: e = launch_gs_bucket<32>(a, hb.source_type, nb, hb.iterations, gstride, st); break;
                case 64: e = launch_gs_bucket<64>(a, hb.source_type, nb, hb.iterations, gstride, st); break;
                default: e = launch_gs_bucket<128>(a, hb.source_type, nb, hb.iterations, gstride, st); break;
            }
        }
        if (!hip_ok(e, "holo kernel launch")) return AFH_ERR_HIP;
        if (fork && (!hip_ok(hipEventRecord(ctx->aux_done[q], st), "hipEventRecord") ||
                     !hip_ok(hipStreamWaitEvent(caller, ctx->aux_done[q], 0), "hipStreamWaitEvent")))
            return AFH_ERR_HIP;
    }
    if (merged) {
        // kernels B and C of every bucket, as record_gspat_bucket's tail
        const int closing = hb.precision == AFH_HOLO_MIXED_CLOSING ? 1 : 0;
        a.split = gspat_split(np, 64);  // every bucket of the set
        const dim3 gb(set_b.first[set_b.n]), gc(set_c.first[set_c.n]), blk(kHB);
        hipLaunchKernelGGL(gspat_iter_set, gb, blk, 0, st, a, set_b, hb.iterations, closing);
        if (closing) hipLaunchKernelGGL(gspat_q_set<false>, gc, blk, 0, st, a, set_c, closing, (int)kQGamma);
        if (hb.precision == AFH_HOLO_MIXED)
            hipLaunchKernelGGL(gspat_q_set<true>, gc, blk, 0, st, a, set_c, closing, (int)kQFinal);
        else
            hipLaunchKernelGGL(gspat_q_set<false>, gc, blk, 0, st, a, set_c, closing, (int)kQFinal);
        if (!hip_ok(hipGetLastError(), "holo kernel launch")) return AFH_ERR_HIP;
    }
    st = caller;
    if (hb.algo == kAlgoGreedy) {
        const float2* cand = reinterpret_cast<const float2*>(wbase + pl.o_cand);
        const float* cph = reinterpret_cast<const float*>(wbase + pl.o_cand + (size_t)pl.div * sizeof(float2));
        for (uint32_t b = 0; b < B; ++b)
            if (hb.offsets_host[b + 1] - hb.offsets_host[b] > kGreedyMaxM) {
                set_error("greedy: more than 128 foci in one instance is not supported");
                return AFH_ERR_UNSUPPORTED;
            }
        if (hb.source_type == kT4010A1)
            hipLaunchKernelGGL((greedy_kernel<kT4010A1>), dim3(B), dim3(kHB), 0, st, a, cand, cph, pl.div);
        else
            hipLaunchKernelGGL((greedy_kernel<kSphere>), dim3(B), dim3(kHB), 0, st, a, cand, cph, pl.div);
        if (!hip_ok(hipGetLastError(), "greedy kernel launch")) return AFH_ERR_HIP;
    }
    if (hb.algo == kAlgoNaive) {
        const dim3 grid((hb.n + kHB - 1) / kHB, B);
        if (hb.source_type == kT4010A1)
            hipLaunchKernelGGL((naive_q<kT4010A1>), grid, dim3(kHB), 0, st, a);
        else
            hipLaunchKernelGGL((naive_q<kSphere>), grid, dim3(kHB), 0, st, a);
        if (!hip_ok(hipGetLastError(), "naive kernel launch")) return AFH_ERR_HIP;
    }
    return AFH_OK;
}

int32_t holo_solve(Context* ctx, const HoloBatch& hb, hipStream_t st) {
    const uint32_t B = hb.batch;
    if (B == 0) return AFH_OK;
    HoloPlan pl;
    const int32_t rc = plan_holo(hb, pl);
    if (rc != AFH_OK) return rc;
    const size_t off_bytes = (B + 1) * sizeof(uint32_t), list_bytes = B * sizeof(uint32_t);
    const size_t cand_bytes = (size_t)pl.div * (sizeof(float2) + sizeof(float));
    // the candidates' float2 table starts 16-byte aligned in the staging buffer
    const size_t cand_off = (off_bytes + list_bytes + 15) & ~(size_t)15;
    // growing the workspace frees the old block: work queued on this stream
    // may still use it
    if (pl.total > ctx->work.cap && !hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize")) return AFH_ERR_HIP;
    if (!ctx->work.reserve(pl.total)) return AFH_ERR_HIP;
    if (!ctx->reserve_staging(cand_off + cand_bytes)) return AFH_ERR_HIP;
    uint32_t* stage = static_cast<uint32_t*>(ctx->staging.ptr);
    memcpy(stage, hb.offsets_host, off_bytes);
    {
        uint32_t* lp = stage + (B + 1);
        for_each_listed(hb, [&](uint32_t b) { *lp++ = b; });
    }
    char* wbase = static_cast<char*>(ctx->work.ptr);
    if (!hip_ok(hipMemcpyAsync(wbase + pl.o_off, stage, off_bytes, hipMemcpyHostToDevice, st), "H2D offsets") ||
        !hip_ok(hipMemcpyAsync(wbase + pl.o_list, stage + (B + 1), list_bytes, hipMemcpyHostToDevice, st),
                "H2D lists"))
        return AFH_ERR_HIP;
    if (pl.div) {
        char* cs = reinterpret_cast<char*>(stage) + cand_off;
        float2* c2 = reinterpret_cast<float2*>(cs);
        float* cph = reinterpret_cast<float*>(cs + (size_t)pl.div * sizeof(float2));
        for (uint32_t q = 0; q < pl.div; ++q) greedy_candidate(q, pl.div, c2[q], cph[q]);
        if (!hip_ok(hipMemcpyAsync(wbase + pl.o_cand, cs, cand_bytes, hipMemcpyHostToDevice, st), "H2D candidates"))
            return AFH_ERR_HIP;
    }
    if (!ctx->upload_done && !hip_ok(hipEventCreateWithFlags(&ctx->upload_done, hipEventDisableTiming), "event"))
        return AFH_ERR_HIP;
    if (!hip_ok(hipEventRecord(ctx->upload_done, st), "hipEventRecord")) return AFH_ERR_HIP;
    return run_holo(hb, pl, wbase, dir_cells_device(), st, ctx);
}

// ---------------------------------------------------------------- allocation-free entry point
// AFH_HoloSolveNoAlloc runs holo_solve's kernels on a file-scope device array
// instead of a context's workspace, for callers that cannot allocate (code
// under graph capture, serving loops that forbid hipMalloc). Each device
// holds its own copy of `scratch` until the process ends: the first
// kHoloScratchTable bytes keep the directivity table (written once per
// device), the rest is the workspace of one call. Host data (offsets,
// instance lists, Greedy candidates, the table) reaches it as kernel
// arguments, so no staging buffer is needed either. Calls that overlap on one
// device share the array and race.
constexpr size_t kHoloScratchBytes = 128ull << 20;
constexpr size_t kHoloScratchTable = 16384;
static_assert(sizeof(DirCell) * kDirCells <= kHoloScratchTable, "directivity table fits its slot");

static __device__ __align__(256) unsigned char scratch[kHoloScratchBytes];

constexpr uint32_t kUploadWords = 960;  // 3.8 KB of kernel arguments per launch
struct UploadChunk {
    uint32_t* dst;
    uint32_t n;
    uint32_t w[kUploadWords];
};

static_assert(sizeof(UploadChunk) <= 4096, "kernel arguments stay within 4 KiB");

__global__ __launch_bounds__(kHB) void upload_words(UploadChunk c) {
    for (uint32_t i = threadIdx.x; i < c.n; i += kHB) c.dst[i] = c.w[i];
}

// Writes a sequence of 32-bit words to consecutive device addresses from dst
// on, kUploadWords per launch.
struct WordUploader {
    hipStream_t st;
    uint32_t* dst;
    UploadChunk c;
    bool ok = true;
    WordUploader(hipStream_t s, void* d) : st(s), dst(static_cast<uint32_t*>(d)) {
        c.dst = dst;
        c.n = 0;
    }
    void flush() {
        if (c.n == 0 || !ok) return;
        hipLaunchKernelGGL(upload_words, dim3(1), dim3(kHB), 0, st, c);
        ok = hip_ok(hipGetLastError(), "upload launch");
        dst += c.n;
        c.dst = dst;
        c.n = 0;
    }
    void push(uint32_t v) {
        c.w[c.n++] = v;
        if (c.n == kUploadWords) flush();
    }
    void push(float v) { push(__builtin_bit_cast(uint32_t, v)); }
};

// The device's scratch address, from hipGetSymbolAddress once per device,
// with the directivity table written into its first bytes.
static char* holo_scratch(hipStream_t st) {
    static std::mutex mu;
    static char* ptr[64] = {};
    int dev = 0;
    if (!hip_ok(hipGetDevice(&dev), "hipGetDevice")) return nullptr;
    if (dev < 0 || dev >= 64) {
        set_error("holo: device index out of range");
        return nullptr;
    }
    std::lock_guard<std::mutex> g(mu);
    if (ptr[dev]) return ptr[dev];
    char* p = nullptr;
    if (!hip_ok(hipGetSymbolAddress(reinterpret_cast<void**>(&p), HIP_SYMBOL(scratch)), "hipGetSymbolAddress"))
        return nullptr;
    static DirCell host[kDirCells];
    build_dir_cells(host);
    WordUploader up(st, p);
    const uint32_t* words = reinterpret_cast<const uint32_t*>(host);
    for (size_t i = 0; i < kDirCells * sizeof(DirCell) / sizeof(uint32_t); ++i) up.push(words[i]);
    up.flush();
    // the table must be in place before a call on another stream reads it
    if (!up.ok || !hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize")) return nullptr;
    ptr[dev] = p;
    return p;
}

int32_t holo_solve_noalloc(const HoloBatch& hb, hipStream_t st) {
    const uint32_t B = hb.batch;
    if (B == 0) return AFH_OK;
    HoloPlan pl;
    const int32_t rc = plan_holo(hb, pl);
    if (rc != AFH_OK) return rc;
    if (pl.total > kHoloScratchBytes - kHoloScratchTable) {
        set_error("AFH_HoloSolveNoAlloc: the batch needs " + std::to_string(pl.total) + " bytes of workspace, more than the " +
                  std::to_string(kHoloScratchBytes - kHoloScratchTable) + " bytes of the device scratch array");
        return AFH_ERR_BUDGET;
    }
    char* base = holo_scratch(st);
    if (!base) return AFH_ERR_HIP;
    char* wbase = base + kHoloScratchTable;
    {
        WordUploader up(st, wbase + pl.o_off);
        for (uint32_t b = 0; b <= B; ++b) up.push(hb.offsets_host[b]);
        up.flush();
        if (!up.ok) return AFH_ERR_HIP;
    }
    if (hb.algo != kAlgoGreedy) {
        WordUploader up(st, wbase + pl.o_list);
        for_each_listed(hb, [&](uint32_t b) { up.push(b); });
        up.flush();
        if (!up.ok) return AFH_ERR_HIP;
    }
    if (pl.div) {
        WordUploader up(st, wbase + pl.o_cand);
        for (uint32_t q = 0; q < pl.div; ++q) {
            float2 c;
            float ph;
            greedy_candidate(q, pl.div, c, ph);
            up.push(c.x);
            up.push(c.y);
        }
        for (uint32_t q = 0; q < pl.div; ++q) {
            float2 c;
            float ph;
            greedy_candidate(q, pl.div, c, ph);
            up.push(ph);
        }
        up.flush();
        if (!up.ok) return AFH_ERR_HIP;
    }
    return run_holo(hb, pl, wbase, reinterpret_cast<const DirCell*>(base), st);
}

// ---------------------------------------------------------------- single instance (AFO_*)
Context& default_context() {
    static Context c;
    return c;
}

bool holo_solve_single(const HoloProblem& p, int algo, float* amp_out, float* phase_out) {
    if (algo == kAlgoLM) return lm_solve_single(p, phase_out, kNlsLM);
    if (algo == kAlgoGN) return lm_solve_single(p, phase_out, kNlsGN);
    if (algo == kAlgoGD) return lm_solve_single(p, phase_out, kNlsGD);
    if (!device().ensure()) return false;
    Context& ctx = default_context();
    std::lock_guard<std::mutex> g(ctx.mu);
    if (!ctx.ensure_stream()) return false;
    const uint64_t n = p.n, np = packed_count(n), m = p.m;
    const size_t src_b = np * AFH_SOURCE_STRIDE * sizeof(float), foci_b = m * 3 * sizeof(float);
    const size_t amps_b = m * sizeof(float), out_b = n * sizeof(float);
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t o_src = 0, o_foci = align(src_b), o_amps = o_foci + align(foci_b), o_amp = o_amps + align(amps_b);
    const size_t o_ph = o_amp + align(out_b), total = o_ph + align(out_b);
    if (!ctx.io.reserve(total) || !ctx.io_host.reserve(total)) return false;
    char* h = static_cast<char*>(ctx.io_host.ptr);
    char* d = static_cast<char*>(ctx.io.ptr);
    pack_sources(p.sources, p.wavenums, p.attens, n, p.source_type, reinterpret_cast<float*>(h + o_src));
    memcpy(h + o_foci, p.foci, foci_b);
    memcpy(h + o_amps, p.amps, amps_b);
    if (!hip_ok(hipMemcpyAsync(d, h, o_amp, hipMemcpyHostToDevice, ctx.stream), "H2D")) return false;
    const uint32_t offs[2] = {0, (uint32_t)m};
    HoloBatch hb;
    hb.sources = reinterpret_cast<const float*>(d + o_src);
    hb.n = (uint32_t)n;
    hb.n_packed = (uint32_t)np;
    hb.source_type = p.source_type;
    hb.foci = reinterpret_cast<const float*>(d + o_foci);
    hb.amps = reinterpret_cast<const float*>(d + o_amps);
    hb.offsets_host = offs;
    hb.batch = 1;
    hb.algo = algo;
    hb.precision = AFH_HOLO_FP32;
    hb.iterations = 100;  // REPEAT in gs.rs:17 and gs_pat.rs:17
    hb.phase_div = p.phase_div;
    hb.out_amp = reinterpret_cast<float*>(d + o_amp);
    hb.out_phase = reinterpret_cast<float*>(d + o_ph);
    if (holo_solve(&ctx, hb, ctx.stream) != AFH_OK) return false;
    if (!hip_ok(hipMemcpyAsync(h + o_amp, d + o_amp, o_ph + out_b - o_amp, hipMemcpyDeviceToHost, ctx.stream), "D2H"))
        return false;
    if (!hip_ok(hipStreamSynchronize(ctx.stream), "hipStreamSynchronize")) return false;
    memcpy(amp_out, h + o_amp, out_b);
    memcpy(phase_out, h + o_ph, out_b);
    return true;
}

}  // namespace afh

using namespace afh;

extern "C" {

int32_t AFH_CreateContext(void** out) {
    clear_error();
    if (!out) return AFH_ERR_ARG;
    *out = nullptr;
    if (!device().ensure()) return AFH_ERR_NO_DEVICE;
    Context* c = new Context();
    if (!c->ensure_stream()) {
        delete c;
        return AFH_ERR_HIP;
    }
    *out = c;
    return AFH_OK;
}

void AFH_DestroyContext(void* ctx) {
    Context* c = static_cast<Context*>(ctx);
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    delete c;
}

int32_t AFH_HoloSolve(void* ctx, const float* packed_sources_dev, uint64_t n_packed, uint64_t n_sources,
                      int32_t source_type, const float* foci_dev, const float* amps_dev, const uint32_t* offsets_host,
                      uint32_t batch, int32_t algo, int32_t precision, uint32_t iterations, float* out_amp_dev,
                      float* out_phase_dev, void* stream) {
    clear_error();
    Context* c = static_cast<Context*>(ctx);
    if (!c || !packed_sources_dev || !foci_dev || !amps_dev || !offsets_host || !out_amp_dev || !out_phase_dev ||
        n_packed % AFH_SOURCE_PAD != 0 || n_packed < n_sources || n_packed > 0xFFFFFFFFull ||
        (source_type != kSphere && source_type != kT4010A1) || algo < kAlgoGSPAT || algo > kAlgoNaive ||
        precision < AFH_HOLO_FP32 || precision > AFH_HOLO_MIXED_SPLIT ||
        (algo != kAlgoGSPAT && precision != AFH_HOLO_FP32) || offsets_host[0] != 0) {
        set_error("AFH_HoloSolve: bad argument");
        return AFH_ERR_ARG;
    }
    std::lock_guard<std::mutex> g(c->mu);
    HoloBatch hb;
    hb.sources = packed_sources_dev;
    hb.n = (uint32_t)n_sources;
    hb.n_packed = (uint32_t)n_packed;
    hb.source_type = source_type;
    hb.foci = foci_dev;
    hb.amps = amps_dev;
    hb.offsets_host = offsets_host;
    hb.batch = batch;
    hb.algo = algo;
    hb.precision = precision;
    hb.iterations = iterations;
    hb.out_amp = out_amp_dev;
    hb.out_phase = out_phase_dev;
    return holo_solve(c, hb, (hipStream_t)stream);
}

int32_t AFH_HoloSolveNoAlloc(const float* packed_sources_dev, uint64_t n_packed, uint64_t n_sources,
                             int32_t source_type, const float* foci_dev, const float* amps_dev,
                             const uint32_t* offsets_host, uint32_t batch, int32_t algo, int32_t precision,
                             uint32_t iterations, float* out_amp_dev, float* out_phase_dev, void* stream) {
    clear_error();
    if (!packed_sources_dev || !foci_dev || !amps_dev || !offsets_host || !out_amp_dev || !out_phase_dev ||
        n_packed % AFH_SOURCE_PAD != 0 || n_packed < n_sources || n_packed > 0xFFFFFFFFull ||
        (source_type != kSphere && source_type != kT4010A1) || algo < kAlgoGSPAT || algo > kAlgoNaive ||
        precision < AFH_HOLO_FP32 || precision > AFH_HOLO_MIXED_SPLIT ||
        (algo != kAlgoGSPAT && precision != AFH_HOLO_FP32) || offsets_host[0] != 0) {
        set_error("AFH_HoloSolveNoAlloc: bad argument");
        return AFH_ERR_ARG;
    }
    HoloBatch hb;
    hb.sources = packed_sources_dev;
    hb.n = (uint32_t)n_sources;
    hb.n_packed = (uint32_t)n_packed;
    hb.source_type = source_type;
    hb.foci = foci_dev;
    hb.amps = amps_dev;
    hb.offsets_host = offsets_host;
    hb.batch = batch;
    hb.algo = algo;
    hb.precision = precision;
    hb.iterations = iterations;
    hb.out_amp = out_amp_dev;
    hb.out_phase = out_phase_dev;
    return holo_solve_noalloc(hb, (hipStream_t)stream);
}

int32_t AFH_GreedySolve(void* ctx_handle, const float* packed, uint64_t n_packed, uint64_t n, int32_t source_type,
                        const float* foci, const float* amps, const uint32_t* offsets, uint32_t batch,
                        uint32_t phase_div, float* out_phase, void* stream) {
    clear_error();
    Context* ctx = static_cast<Context*>(ctx_handle);
    if (!ctx || !packed || !foci || !amps || !offsets || !out_phase || n == 0 || n_packed % AFH_SOURCE_PAD != 0 ||
        n_packed < n || n_packed > 0xFFFFFFFFull || (source_type != kSphere && source_type != kT4010A1) ||
        offsets[0] != 0) {
        set_error("AFH_GreedySolve: bad arguments");
        return AFH_ERR_ARG;
    }
    for (uint32_t b = 0; b < batch; ++b)
        if (offsets[b + 1] <= offsets[b]) {
            set_error("AFH_GreedySolve: an instance has no foci");
            return AFH_ERR_ARG;
        }
    HoloBatch hb;
    hb.sources = packed;
    hb.n = (uint32_t)n;
    hb.n_packed = (uint32_t)n_packed;
    hb.source_type = source_type;
    hb.foci = foci;
    hb.amps = amps;
    hb.offsets_host = offsets;
    hb.batch = batch;
    hb.algo = kAlgoGreedy;
    hb.precision = AFH_HOLO_FP32;
    hb.iterations = 0;
    hb.phase_div = phase_div;
    hb.out_amp = nullptr;
    hb.out_phase = out_phase;
    std::lock_guard<std::mutex> g(ctx->mu);
    return holo_solve(ctx, hb, static_cast<hipStream_t>(stream));
}

}  // extern "C"
