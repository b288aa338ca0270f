/*
 * afc_hip.h -- C ABI of libafc_hip.so, the MI355X (gfx950) backend for the
 * acoustic field sum and the GS / GS-PAT / LM hologram solvers.
 *
 * Part 1 is the drop-in surface: the same symbols, argument order, value
 * structs, type tags and ownership rules as the reference's cdylib `afccapi`
 * (the .rs files of capi/src; ctypes argtypes in
 * python/afc/nativemethods.py:113-234).
 * Each declaration cites the reference export it replaces.
 *
 * Part 2 adds entry points the reference has no counterpart for: device-
 * resident evaluation, batched solvers and the request server. They take
 * plain pointers and sizes and return an int32 status (AFH_OK on success).
 *
 * Differences from the reference (intentional):
 *   - Every CalculatorType runs on the GPU: 0 (Cpu), 2 (Gpu) and this
 *     backend's own 3 (AFC_CALC_HIP) take the FAST kernels, 1 (Accurate) the
 *     ACCURATE ones. There is no CPU path in this library.
 *   - Errors never abort the process: calls that fail record a message
 *     retrievable with AFC_GetLastError() (thread-local).
 *   - AFO_IFFT (an image file read and an FFT) is exported so that bindings
 *     load, but it only records "not provided" and leaves the system as is.
 *     Every other AFO_* symbol is implemented.
 *   - Size limits: LM, Gauss-Newton, gradient descent, APO, Long and Horn
 *     take at most 32 foci; GS, GS-PAT, Naive and Greedy at most 128; one
 *     AFC_Calculate at most 2^31 observation points.
 */
#ifndef AFC_HIP_H
#define AFC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- value types */
/* nalgebra Vector3<f32> as passed by value (nativemethods.py:18-25) */
typedef struct {
    float x, y, z;
} AFC_Vector3;

/* core/wave_sources/sphere_sources.rs:19-27 (repr(C), 24 bytes) */
typedef struct {
    float x, y, z;
    float amp, phase, freq;
} AFC_SphereWaveSource;

/* core/wave_sources/t4010a1.rs:73-82 (repr(C), 36 bytes) */
typedef struct {
    float x, y, z;
    float nx, ny, nz;
    float amp, phase, freq;
} AFC_T4010A1;

/* type tags: capi/src/type_inference_aux.rs:53-103, grid/dimension.rs:23-28 */
enum { AFC_CALC_CPU = 0, AFC_CALC_ACCURATE = 1, AFC_CALC_GPU = 2, AFC_CALC_HIP = 3 };
enum { AFC_SOURCE_SPHERE = 0, AFC_SOURCE_T4010A1 = 1 };
enum { AFC_AREA_GRID1D = 0, AFC_AREA_GRID2D = 1, AFC_AREA_GRID3D = 2, AFC_AREA_SCATTER = 3 };
enum { AFC_FIELD_PRESSURE = 0, AFC_FIELD_POWER = 1, AFC_FIELD_COMPLEX = 2 };
enum { AFC_AXIS_X = 1, AFC_AXIS_Y = 2, AFC_AXIS_Z = 3 };

/* ======================================================= part 1: capi surface */

/* capi/src/calculator.rs:21 */
void AFC_CreateCalculator(void** out, int32_t calc_type);
/* capi/src/calculator.rs:39 */
void AFC_FreeCalculator(void* handle, int32_t calc_type);
/* capi/src/calculator.rs:53-63. Returns the number of observation points and
 * points *out at the field handle's own buffer (1 float per point, or an
 * interleaved (re, im) pair for AFC_FIELD_COMPLEX), valid until the next
 * calculate on that field or its free. 0 with *out = NULL on error. */
uint64_t AFC_Calculate(void* calc_handle, void* system_handle, void* obs_handle, void* field_handle, const void** out,
                       int32_t calc_type, int32_t source_type, int32_t obs_type, int32_t field_type);

/* capi/src/field_buffer.rs:37-49 */
void AFC_CreatePressureField(void** out);
void AFC_FreePressureField(void* handle);
void AFC_CreatePowerField(void** out);
void AFC_FreePowerField(void* handle);
void AFC_CreateComplexPressureField(void** out);
void AFC_FreeComplexPressureField(void* handle);

/* capi/src/observe_area.rs:21 */
void AFC_CreateScatterArea(void** out);
/* capi/src/observe_area.rs:31 */
void AFC_ScatterAddObservePoint(void* handle, AFC_Vector3 p);
/* capi/src/observe_area.rs:38 */
void AFC_FreeScatterArea(void* handle);
/* capi/src/observe_area.rs:43-53 */
void AFC_CreateGridArea1D(void** out, float first_start, float first_end, float second, float third, float resolution,
                          int32_t first_axis, int32_t second_axis, int32_t third_axis);
/* capi/src/observe_area.rs:88-99 */
void AFC_CreateGridArea2D(void** out, float first_start, float first_end, float second_start, float second_end,
                          float third, float resolution, int32_t first_axis, int32_t second_axis, int32_t third_axis);
/* capi/src/observe_area.rs:134-146 */
void AFC_CreateGridArea3D(void** out, float first_start, float first_end, float second_start, float second_end,
                          float third_start, float third_end, float resolution, int32_t first_axis,
                          int32_t second_axis, int32_t third_axis);
/* capi/src/observe_area.rs:181 */
void AFC_FreeGridArea(void* handle, int32_t dimension);
/* capi/src/observe_area.rs:197-203 -- counts by axis (x, y, z), not by order */
void AFC_GetGridSize(void* handle, int32_t dimension, uint32_t* first, uint32_t* second, uint32_t* third);

/* capi/src/system.rs:22-26 */
void AFC_CreateUniformSystem(void** out, float temperature, int32_t source_type);
/* capi/src/system.rs:44 */
void AFC_FreeUniformSystem(void* handle, int32_t source_type);
/* capi/src/system.rs:58 -- mm/s */
float AFC_UniformSystemSoundSpeed(void* handle, int32_t source_type);
/* capi/src/system.rs:75 -- string owned by the handle, valid until the next info call on it */
char* AFC_UniformSystemInfo(void* handle, int32_t source_type);
/* capi/src/system.rs:95 */
char* AFC_UniformSystemSourceInfo(void* handle, uint64_t source_idx, int32_t source_type);
/* capi/src/system.rs:117 */
void AFC_AddSphereWaveSource(void* handle, AFC_SphereWaveSource source);
/* capi/src/system.rs:125 */
void AFC_AddT4010A1(void* handle, AFC_T4010A1 source);
/* capi/src/system.rs:132-136 -- mutable pointer into the system's own source
 * array; callers may edit it in place and the next calculate sees the edits */
uint64_t AFC_GetWaveSources(void* handle, void** out, int32_t source_type);

/* capi/src/optimizer.rs:43 */
void AFO_FocalPoint(void* handle, AFC_Vector3 point, int32_t source_type);
/* capi/src/optimizer.rs:80-95 -- foci packed as 3 floats per focus */
void AFO_GSPAT(void* handle, const void* foci, const float* amps, uint64_t size, int32_t source_type);
/* capi/src/optimizer.rs:98-109 */
void AFO_GS(void* handle, const void* foci, const float* amps, uint64_t size, int32_t source_type);
/* capi/src/optimizer.rs:112-127 */
void AFO_Naive(void* handle, const void* foci, const float* amps, uint64_t size, int32_t source_type);
/* capi/src/optimizer.rs:213-224 */
void AFO_LM(void* handle, const void* foci, const float* amps, uint64_t size, int32_t source_type);
/* capi/src/optimizer.rs:227-241, combination/greedy.rs:23-83: sets the
 * phases (amplitudes unchanged); phase_div in 1..4096, at most 128 foci */
void AFO_Greedy(void* handle, const void* foci, const float* amps, uint64_t size, uint64_t phase_div,
                int32_t source_type);
/* capi/src/optimizer.rs:177-192, nls/gauss_newton.rs:35-89 (phases only,
 * FP64, at most 32 foci) */
void AFO_GaussNewton(void* handle, const void* foci, const float* amps, uint64_t size, int32_t source_type);
/* capi/src/optimizer.rs:195-210, nls/gradient_descent.rs:34-61 (phases only,
 * FP64, at most 32 foci) */
void AFO_GradientDescent(void* handle, const void* foci, const float* amps, uint64_t size, int32_t source_type);
/* capi/src/optimizer.rs:158-174, nls/apo.rs:100-158: sets amplitudes
 * (amp |q_j| / camax q) and phases (FP64, at most 32 foci) */
void AFO_APO(void* handle, const void* foci, const float* amps, uint64_t size, float lambda, int32_t source_type);
/* capi/src/optimizer.rs:130-141, matrix/horn.rs:70-129: sets amplitudes
 * and phases (FP64, at most 32 foci); REPEAT 100, lambda 0.8, Tikhonov 1e-5
 * as in the reference, rows drawn from a splitmix64 stream with seed 0 (the
 * reference's thread_rng is unseeded) */
void AFO_Horn(void* handle, const void* foci, const float* amps, uint64_t size, int32_t source_type);
/* capi/src/optimizer.rs:144-155, matrix/long.rs:46-106: sets amplitudes and
 * phases (FP64, at most 32 foci). The eigenvector matrix whose row imax the
 * reference reads is taken with ascending eigenvalues and each vector's
 * largest-modulus entry real positive (nalgebra fixes neither). */
void AFO_Long(void* handle, const void* foci, const float* amps, uint64_t size, int32_t source_type);
/* capi/src/optimizer.rs:47-56, acoustic-field-optimizer/src/bessel.rs:31-55:
 * sets the phases on the host, in the reference's f32 operation order */
void AFO_BesselBeam(void* handle, AFC_Vector3 point, AFC_Vector3 dir, float theta, int32_t source_type);
/* Exported for binding compatibility only; records "unsupported".
 * capi/src/optimizer.rs:59-77 */
void AFO_IFFT(void* handle, const char* path, AFC_Vector3 bottom_left, AFC_Vector3 top_left,
              AFC_Vector3 bottom_right, float spacing, float z, int32_t source_type);

/* ============================================== part 2: no reference counterpart */

enum {
    AFH_OK = 0,
    AFH_ERR_ARG = 1,      /* bad tag, size or pointer */
    AFH_ERR_HIP = 2,      /* HIP runtime error */
    AFH_ERR_NO_DEVICE = 3,/* no GPU visible */
    AFH_ERR_UNSUPPORTED = 4,
    AFH_ERR_BUDGET = 5    /* request does not fit the HBM budget */
};

/* Thread-local message of the last failing call ("" if none). */
const char* AFC_GetLastError(void);

/* Copies the per-source wavenumber [1/mm] and attenuation [Np/mm] the system
 * computed when each source was added (system/uniform_medium.rs:102-107).
 * Returns the number of sources; either output may be NULL. */
uint64_t AFH_GetSourceConstants(void* system_handle, float* wavenums, float* attens);

/* Floats per packed source in device memory (see AFH_PackSources). */
#define AFH_SOURCE_STRIDE 12
/* Packed source arrays are padded to a multiple of this many entries. */
#define AFH_SOURCE_PAD 32

/* Packs n reference structs (AFC_SphereWaveSource or AFC_T4010A1 by
 * source_type) with their wavenumbers [1/mm] and attenuations [Np/mm] into
 * the device layout: per source {x, y, z, nx, ny, nz, amp, phase/(2pi) mod 1,
 * k/(2pi) hi, k/(2pi) lo, alpha*log2(e), alpha}. Writes AFH_PackedCount(n)
 * entries (the tail is zero-amplitude padding) to out (host memory). */
uint64_t AFH_PackedCount(uint64_t n);
int32_t AFH_PackSources(const void* sources, const float* wavenums, const float* attens, uint64_t n,
                        int32_t source_type, float* out);

/* Observation points of a field evaluation. kind 0: a grid generated on the
 * device from (origin, resolution, counts in axis order); kind 1: scatter
 * points read from `points` (device pointer, 3 floats per point). */
typedef struct {
    int32_t kind;
    int32_t order[3];   /* axes 0=x 1=y 2=z, fastest-varying first */
    uint32_t count[3];  /* points along order[0], order[1], order[2] */
    float origin[3];    /* by axis (x, y, z) */
    float resolution;
    const float* points;
    uint64_t n_points;
} AFH_PointSet;

enum { AFH_PREC_FAST = 0, AFH_PREC_ACCURATE = 1 };

/* Field of packed device sources at a point set, written to device `out`
 * (field_type as AFC_FIELD_*: 1 float per point, or 2 for complex). All
 * pointers are device memory; the call is asynchronous on `stream`
 * (a hipStream_t, NULL = default stream). */
int32_t AFH_FieldEval(const float* packed_sources_dev, uint64_t n_packed, int32_t source_type, const AFH_PointSet* pts,
                      int32_t field_type, int32_t precision, float* out_dev, void* stream);

/* ---- hologram solvers, batched over instances sharing one transducer array */

/* A context owns a HIP stream and a grow-only device workspace on the current
 * device; calls on one context are serialized (use one context per thread or
 * stream for concurrency). */
int32_t AFH_CreateContext(void** out);
void AFH_DestroyContext(void* ctx);

enum { AFH_ALGO_GSPAT = 0, AFH_ALGO_GS = 1, AFH_ALGO_NAIVE = 2 };
/* GS-PAT precision. FP32 forms R = G diag(1/d^2) G^H in FP32 (FP32 MFMA, or
 * the VALU for at most 16 foci) and q = B p from FP32 G. The mixed modes form
 * R in one pass with G on FP16 MFMA with FP32 accumulation, each entry of
 * H = s G / d (per-column scaling, |H| <= s) carried as an FP16 hi + lo pair
 * (hi hi^H + hi lo^H + lo hi^H, near-FP32 R); MIXED then takes q = B p from
 * the FP16 hi copy of H, MIXED_SPLIT from FP32 G, and MIXED_CLOSING redoes the
 * last R p product as G (B p) in FP32 before q = B p. Every GS-PAT mode takes
 * the source-focus distance from the exact coordinate difference. GS and
 * Naive are FP32 only. */
enum { AFH_HOLO_FP32 = 0, AFH_HOLO_MIXED = 1, AFH_HOLO_MIXED_CLOSING = 2, AFH_HOLO_MIXED_SPLIT = 3 };

/* Solves `batch` instances. packed_sources_dev: AFH_PackSources layout
 * (phases ignored: the optimizers zero them, gs_pat.rs:39-41); instance b
 * has foci foci_dev[3*offsets_host[b] ...] (3 floats each) with target
 * amplitudes amps_dev[offsets_host[b] ...], offsets_host[0] = 0, at most 128
 * foci per instance. Writes per-source amplitude |q| and phase arg q of
 * instance b to out_amp_dev / out_phase_dev [b * n_sources ...]
 * (gs_pat.rs:75-80, gs.rs:71-76, naive.rs:45-50). Asynchronous on `stream`
 * (a hipStream_t; NULL is the default stream). iterations: 100 in the
 * reference (gs.rs:17, gs_pat.rs:17). */
int32_t AFH_HoloSolve(void* ctx, const float* packed_sources_dev, uint64_t n_packed, uint64_t n_sources,
                      int32_t source_type, const float* foci_dev, const float* amps_dev, const uint32_t* offsets_host,
                      uint32_t batch, int32_t algo, int32_t precision, uint32_t iterations, float* out_amp_dev,
                      float* out_phase_dev, void* stream);

/* No counterpart in the reference. AFH_HoloSolve without a context, for
 * callers that cannot allocate (code under HIP graph capture, serving loops
 * that forbid hipMalloc): the same kernels and results, with the workspace in
 * a 128 MiB device array of the library (each GPU holds its own copy until
 * the process ends) and the host-side tables passed as kernel arguments, so
 * the call allocates neither device nor pinned memory. A batch whose
 * workspace exceeds the array (128 MiB less 16 KiB) returns AFH_ERR_BUDGET.
 * Calls that overlap on one GPU share the array and race: serialize them.
 * The first call on a device writes the directivity table into the array and
 * synchronizes `stream` once. Arguments as AFH_HoloSolve. */
int32_t AFH_HoloSolveNoAlloc(const float* packed_sources_dev, uint64_t n_packed, uint64_t n_sources,
                             int32_t source_type, const float* foci_dev, const float* amps_dev,
                             const uint32_t* offsets_host, uint32_t batch, int32_t algo, int32_t precision,
                             uint32_t iterations, float* out_amp_dev, float* out_phase_dev, void* stream);

/* Greedy phase search (combination/greedy.rs:23-83) for `batch` instances:
 * candidates e^{i 2 pi idx / phase_div}, sources in index order; writes the
 * phase of every source of instance b to out_phase_dev[b * n_sources ...].
 * phase_div in 1..4096, at most 128 foci per instance. Same argument
 * conventions as AFH_HoloSolve. */
int32_t AFH_GreedySolve(void* ctx, const float* packed_sources_dev, uint64_t n_packed, uint64_t n_sources,
                        int32_t source_type, const float* foci_dev, const float* amps_dev,
                        const uint32_t* offsets_host, uint32_t batch, uint32_t phase_div, float* out_phase_dev,
                        void* stream);

/* Levenberg-Marquardt (nls/lm.rs:40-105) for `batch` instances, FP64: writes
 * the source phases x[0..N) of instance b to out_phase_dev[b * n_sources ...]
 * (amplitudes are left to the caller, as the reference leaves them) and, if
 * out_iters_dev is not NULL, the iteration count. At most 32 foci per
 * instance; k_max = 200 in the reference (lm.rs:20). Same argument
 * conventions as AFH_HoloSolve. */
int32_t AFH_LMSolve(void* ctx, const float* packed_sources_dev, uint64_t n_packed, uint64_t n_sources,
                    int32_t source_type, const float* foci_dev, const float* amps_dev, const uint32_t* offsets_host,
                    uint32_t batch, uint32_t k_max, float* out_phase_dev, uint32_t* out_iters_dev, void* stream);

/* The nonlinear least-squares solvers of nls/ on the LM machinery, batched:
 * algo AFH_NLS_LM (as AFH_LMSolve), AFH_NLS_GAUSS_NEWTON (gauss_newton.rs:35-89,
 * K_MAX 200) or AFH_NLS_GRADIENT_DESCENT (gradient_descent.rs:34-61, K_MAX
 * 10 000). Same arguments and outputs as AFH_LMSolve. */
enum { AFH_NLS_LM = 0, AFH_NLS_GAUSS_NEWTON = 1, AFH_NLS_GRADIENT_DESCENT = 2 };
int32_t AFH_NLSolve(void* ctx, const float* packed_sources_dev, uint64_t n_packed, uint64_t n_sources,
                    int32_t source_type, const float* foci_dev, const float* amps_dev, const uint32_t* offsets_host,
                    uint32_t batch, int32_t algo, uint32_t k_max, float* out_phase_dev, uint32_t* out_iters_dev,
                    void* stream);

/* APO (nls/apo.rs:100-158) for `batch` instances, FP64, K_MAX 200 in the
 * reference: writes the amplitude amp_j |q_j| / camax(q) and phase arg q_j
 * of every source of instance b to out_amp_dev / out_phase_dev
 * [b * n_sources ...], and the iteration count if out_iters_dev is not NULL.
 * At most 32 foci per instance; the N x N quasi-Newton matrix takes 16 N^2
 * bytes of workspace per instance. Same argument conventions as AFH_LMSolve. */
int32_t AFH_APOSolve(void* ctx, const float* packed_sources_dev, uint64_t n_packed, uint64_t n_sources,
                     int32_t source_type, const float* foci_dev, const float* amps_dev, const uint32_t* offsets_host,
                     uint32_t batch, float lambda, uint32_t k_max, float* out_amp_dev, float* out_phase_dev,
                     uint32_t* out_iters_dev, void* stream);

/* Long (matrix/long.rs:46-106) for `batch` instances, FP64: amplitudes and
 * phases of every source of instance b into out_amp_dev / out_phase_dev
 * [b * n_sources ...]. At most 32 foci per instance. Same argument
 * conventions as AFH_LMSolve. */
int32_t AFH_LongSolve(void* ctx, const float* packed_sources_dev, uint64_t n_packed, uint64_t n_sources,
                      int32_t source_type, const float* foci_dev, const float* amps_dev, const uint32_t* offsets_host,
                      uint32_t batch, float* out_amp_dev, float* out_phase_dev, void* stream);

/* Horn (matrix/horn.rs:58-129) for `batch` instances, FP64, with the
 * reference's parameters as arguments (REPEAT_SDP 100, LAMBDA_SDP 0.8,
 * TIKHONOV_DEFAULT 1e-5 there): instance b draws its rows from the
 * splitmix64 stream seeded with seed + b. Outputs as AFH_LongSolve. */
int32_t AFH_HornSolve(void* ctx, const float* packed_sources_dev, uint64_t n_packed, uint64_t n_sources,
                      int32_t source_type, const float* foci_dev, const float* amps_dev, const uint32_t* offsets_host,
                      uint32_t batch, uint64_t seed, uint32_t repeat, float lambda, float tikhonov,
                      float* out_amp_dev, float* out_phase_dev, void* stream);

/* ---- request server: field evaluations and GS-PAT solves, continuously
 * batched on one GPU (see the design notes in server.hip) */
typedef struct {
    uint64_t hbm_budget_bytes;     /* device arena for request inputs / outputs */
    uint64_t pinned_budget_bytes;  /* pinned host arena for result staging */
    uint32_t n_streams;            /* 1..16 */
    uint32_t max_gspat_batch;      /* instances per batched GS-PAT launch */
    int32_t gspat_precision;       /* AFH_HOLO_* */
    uint32_t gspat_iterations;     /* 100 in the reference */
    uint32_t gspat_cus;            /* CUs reserved for GS-PAT batches (0: shared) */
    uint32_t gspat_graphs;         /* 1: each GS-PAT bucket's kernel chain is one
                                      hipGraph launch (graphs cached per shape,
                                      grids and arguments set per batch); 0:
                                      launched kernel by kernel */
} AFS_Config;

int32_t AFS_Create(const AFS_Config* cfg, void** out);
void AFS_Destroy(void* srv);
/* Registers a transducer array (reference structs + k + alpha, as
 * AFH_PackSources takes them); it stays resident on the device. */
int32_t AFS_AddArray(void* srv, const void* sources, const float* wavenums, const float* attens, uint64_t n,
                     int32_t source_type, uint32_t* array_id);
/* Field of a registered array at a point set (grid, or scatter with HOST
 * points), FAST precision, written to out_host when the request completes
 * (1 or 2 floats per point as AFC_FIELD_*). When out_host lies in pinned
 * (page-locked) host memory the result is DMA'd there directly; otherwise it
 * is staged in the pinned arena and copied (in parts, by several threads,
 * for results above 1 MiB). */
int32_t AFS_SubmitField(void* srv, uint32_t array_id, const AFH_PointSet* pts, int32_t field_type, float* out_host,
                        uint64_t* ticket);
/* GS-PAT on a registered array: m (1..128) foci (host, 3 floats each) and
 * target amplitudes; per-source amplitude / phase written to out_amp /
 * out_phase (n floats each) when the request completes. */
int32_t AFS_SubmitGSPAT(void* srv, uint32_t array_id, const float* foci, const float* amps, uint32_t m,
                        float* out_amp, float* out_phase, uint64_t* ticket);
/* count GS-PAT requests at once (request i as AFS_SubmitGSPAT with
 * array_ids[i], foci[i], amps[i], m[i], out_amp[i], out_phase[i]; its ticket
 * in tickets[i]), queued in order under one lock with one wake-up of the
 * worker. All or nothing: an invalid request rejects the batch before any is
 * queued. No counterpart in the reference (its optimizers are called one
 * problem at a time); for clients that hold several requests at once. */
int32_t AFS_SubmitGSPATBatch(void* srv, uint32_t count, const uint32_t* array_ids, const float* const* foci,
                             const float* const* amps, const uint32_t* m, float* const* out_amp,
                             float* const* out_phase, uint64_t* tickets);
/* Blocks until the request completes; returns its status and latency
 * (submit to result in the caller's buffer). Each ticket is waited once. */
int32_t AFS_Wait(void* srv, uint64_t ticket, double* latency_ms);
/* Latency percentiles over the requests completed since the last reset, of
 * one kind (0 field, 1 GS-PAT) or of both (-1). */
int32_t AFS_Stats(void* srv, int32_t kind, double* p50_ms, double* p99_ms, uint64_t* completed, int32_t reset);
/* GS-PAT graph launches so far and how many of them had to build (instantiate)
 * a graph rather than reuse a cached one (both 0 with gspat_graphs = 0). */
int32_t AFS_GraphCounters(void* srv, uint64_t* launches, uint64_t* builds);

/* T4010A1 directivity as a function of cos(theta), evaluated on the host with
 * exactly the formulas the kernels use in each precision mode (cell table
 * for AFH_PREC_FAST, reference binning for AFH_PREC_ACCURATE). For tests and
 * inspection; reference formula: core/wave_sources/t4010a1.rs:107-120. */
int32_t AFH_DirectivityT4010A1(const float* cos_theta, uint64_t n, int32_t mode, float* out);

#ifdef __cplusplus
}
#endif
#endif /* AFC_HIP_H */
