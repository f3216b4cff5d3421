/*
 * pfsoft_amd.h -- C ABI of the MI355X implementation of the discrete SO(3)
 * Fourier transform (DSOFT) and its inverse (iDSOFT).
 *
 * Plain C: pointers, sizes and int status codes.  No torch or HIP types cross
 * this boundary (streams are passed as void*, a hipStream_t or NULL).
 *
 * Layouts (N = 2B, complex128 = two interleaved doubles):
 *   grid    N^3 complex, element (row r, col c, slice k) at N*N*k + N*c + r --
 *           the reference's grid3D<complex<double>> memory
 *           (include/PFSOFTlib_headers/grid3D_cx.hpp:293-303).
 *   coeffs  B(4B^2-1)/3 complex, packed by degree: block l starts at
 *           (4l^3 - l)/3; element (l, M, M') at that offset + iM + (2l+1)*iM',
 *           iX = X >= 0 ? X : 2l+1+X -- the reference's per-degree column-major
 *           matrices laid end to end (src/dsoft_fourier_coefficients.cpp:48-112).
 *
 * Every function returns PFS_OK (0) or a PFS_ERR_* code; pfs_last_error()
 * gives the message of the last failure on the calling thread.  Calls on one
 * context are serialised by the context; separate contexts may run
 * concurrently.
 */
#ifndef PFSOFT_AMD_H
#define PFSOFT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pfs_context pfs_context;

enum pfs_status {
    PFS_OK = 0,
    PFS_ERR_ARG = 1,        /* null pointer, bad size or bad enum value        */
    PFS_ERR_SHAPE = 2,      /* grid not cubic / not even / wrong bandwidth      */
    PFS_ERR_DEVICE = 3,     /* HIP runtime error                                */
    PFS_ERR_MEMORY = 4,     /* device allocation failed                         */
    PFS_ERR_UNSUPPORTED = 5,/* bandwidth or size outside the supported range    */
    PFS_ERR_BUDGET = 6,     /* request does not fit the context's HBM budget    */
    PFS_ERR_TIMEOUT = 7     /* a wait ended before the request finished; the
                               request is still queued or running          */
};

enum pfs_precision {
    PFS_F64 = 0,            /* FP64 throughout (the plain path)                 */
    PFS_F32 = 1             /* FP32 FFT and DWT, FP64 recurrence tables         */
};

/* Message for the last failing call on this thread ("" if none). */
const char* pfs_last_error(void);

/* Library version string. */
const char* pfs_version(void);

/* Context on one HIP device.  hbm_budget_bytes = 0 means no limit.
 * Replaces the reference's implicit process-wide state (FFTW plans, OpenMP
 * pool); there is no counterpart function in the reference. */
int pfs_ctx_create(int device, size_t hbm_budget_bytes, pfs_context** out);
int pfs_ctx_destroy(pfs_context* ctx);

/* Build and keep resident the per-bandwidth tables (quadrature weights,
 * cos/sin of the Chebyshev nodes, Wigner base norms, FFT twiddles).  Called
 * implicitly by the transforms. */
int pfs_plan_warm(pfs_context* ctx, int bandwidth);

/* Number of coefficients B(4B^2-1)/3 and the packed index of (l, M, M'). */
long pfs_coeff_count(int bandwidth);
long pfs_coeff_index(int l, int M, int Mp);

/* Forward transform, host buffers.  Replaces
 * pfsoft::FourierTransforms::DSOFT(grid3D<complex<double>> sample,
 *   DSOFTFourierCoefficients& fc, int threads)
 *   (include/PFSOFTlib_headers/fn_fourier_transforms.hpp:75, src/fn_dsoft.cpp:76).
 * grid: N^3 complex128 (read only -- the reference takes it by value);
 * coeffs: pfs_coeff_count(B) complex128, written. */
int pfs_dsoft_f64(pfs_context* ctx, int bandwidth, const double* grid, double* coeffs);

/* Inverse transform, host buffers.  Replaces
 * pfsoft::FourierTransforms::IDSOFT(const DSOFTFourierCoefficients& fc,
 *   grid3D<complex<double>>& synthesis, int threads)
 *   (fn_fourier_transforms.hpp:78, src/fn_idsoft.cpp:72).
 * grid is read and written: as in the reference, the entries with row == B or
 * col == B are never produced by the Wigner stage, so the inverse FFT reads
 * what the caller's grid held there (zero for a fresh grid). */
int pfs_idsoft_f64(pfs_context* ctx, int bandwidth, const double* coeffs, double* grid);

/* Same transforms on device buffers (HBM-resident), enqueued on stream
 * (a hipStream_t; NULL = the HIP null stream); the calls return once
 * enqueued.  The pipeline work buffers belong to the context: calls that
 * overlap in time on different streams of one context race on them, so use
 * one stream per context (or one context per stream). */
int pfs_dsoft_f64_dev(pfs_context* ctx, int bandwidth, const void* grid_dev, void* coeffs_dev, void* stream);
int pfs_idsoft_f64_dev(pfs_context* ctx, int bandwidth, const void* coeffs_dev, void* grid_dev, void* stream);

/* Forward transform on device buffers that allocates nothing.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and results as
 * pfs_dsoft_f64_dev (bit for bit), but its two work buffers come from a
 * file-scope device array of 64 MiB per GPU instead of the context, so:
 *   - bandwidth <= 64 (2B <= 128); larger ones return PFS_ERR_UNSUPPORTED;
 *   - the bandwidth's plan must already be on the context
 *     (pfs_plan_warm(ctx, bandwidth)), else PFS_ERR_ARG;
 *   - every call on one GPU shares that GPU's array, whatever its context:
 *     calls that overlap in time on one GPU race;
 *   - the calls launch kernels only (no allocation, no timing events, no
 *     memset), so they can be captured into a HIP graph. */
int pfs_dsoft_f64_dev_noalloc(pfs_context* ctx, int bandwidth, const void* grid_dev, void* coeffs_dev, void* stream);

/* Batched mixed-precision transforms of nsig signals of one bandwidth
 * (no counterpart in the reference, which transforms one complex128 grid per
 * call).  grids: nsig consecutive N^3 complex64 grids (reference layout);
 * coeffs: nsig consecutive packed complex64 coefficient sets.  Slice FFTs and
 * the Wigner contraction run in FP32 (MFMA); Wigner tables, quadrature weights
 * and twiddles are FP64-computed.  The inverse writes every grid entry; those
 * with row or column index B (never produced by the Wigner stage) are 0.
 * Supported B: 8, 16, 32, 64, 128. */
int pfs_dsoft_batch_f32_dev(pfs_context* ctx, int bandwidth, int nsig, const void* grids_dev, void* coeffs_dev,
                            void* stream);
int pfs_idsoft_batch_f32_dev(pfs_context* ctx, int bandwidth, int nsig, const void* coeffs_dev, void* grids_dev,
                             void* stream);

/* Batched mixed-precision transforms of nsig signals of mixed bandwidth in
 * one call (no counterpart in the reference, which transforms one grid of one
 * bandwidth per call).  bandwidths[i] (host array) is signal i's bandwidth
 * (8, 16, 32, 64 or 128); its input and output are at the device addresses
 * in[i] and out[i] (host arrays of device pointers): complex64 grids
 * (reference layout) and packed complex64 coefficients as in
 * pfs_dsoft_batch_f32_dev.  The slice FFTs and coefficient transposes run per
 * bandwidth group; the Wigner contraction of all groups is one persistent
 * launch over a work queue of (bandwidth, pair, signal group) items, heaviest
 * first.  The launch sequence of each batch composition (direction and the
 * signal count of each bandwidth) is captured once into a HIP graph and
 * replayed after that (context tune key "batch_graphs", 1 default, 0 =
 * launch directly).  Results equal those of the single-bandwidth entry
 * points bit for bit.  Enqueued on stream; the address tables are uploaded
 * on that stream before the batch. */
int pfs_dsoft_batch_multi_f32_dev(pfs_context* ctx, int nsig, const int* bandwidths, const void* const* grids_dev,
                                  void* const* coeffs_dev, void* stream);
int pfs_idsoft_batch_multi_f32_dev(pfs_context* ctx, int nsig, const int* bandwidths, const void* const* coeffs_dev,
                                   void* const* grids_dev, void* stream);

/* Autotuner of the batched path (no counterpart in the reference): times every
 * work shape of the two Wigner contraction kernels (waves per workgroup x
 * signals per wave) for batches of 1, 2, 4, ... max_signals signals at this
 * bandwidth and keeps the fastest per direction and size bucket in the
 * context; later batches of a covered size use it, others a fixed rule.
 * Results do not depend on the shape.  pfs_batch_shape reports the shape a
 * batch of nsig signals would use. */
int pfs_batch_autotune(pfs_context* ctx, int bandwidth, int max_signals);
int pfs_batch_shape(pfs_context* ctx, int bandwidth, int forward, int nsig, int* waves_per_block,
                    int* signals_per_wave);

/* Layer-wise unnormalised 2D DFT of lays square slices, in place, host
 * memory.  Replaces uzl_fftw_layer_wise_DFT2_grid3D / _IDFT2_grid3D
 * (include/PFSOFTlib_headers/fftw_wrapper.hpp:28-33, src/fftw_wrapper.cpp:47-172):
 * same arguments, same layout (n0 = cols slow, n1 = rows fast), forward sign
 * -1 / backward +1, no normalisation.  threads is accepted and ignored.
 * Only cols == rows is supported; failures are reported through
 * pfs_last_error() since the reference signature returns void.  Uses a
 * process-wide context on device 0. */
void pfs_layer_wise_DFT2_grid3D(int cols, int rows, int lays, double* arr, int threads);
void pfs_layer_wise_IDFT2_grid3D(int cols, int rows, int lays, double* arr, int threads);

/* The same two functions under the reference's own symbol names
 * (include/PFSOFTlib_headers/fftw_wrapper.hpp:31-32, declared extern "C"
 * inside namespace pfsoft, so the linker names carry no namespace): a
 * reference build that links -lpfsoft_amd instead of src/fftw_wrapper.cpp +
 * libfftw3 resolves grid3D::layer_wise_DFT2/IDFT2 (grid3D_cx.hpp:339,400)
 * to these. */
void uzl_fftw_layer_wise_DFT2_grid3D(int cols, int rows, int lays, double* arr, int threads);
void uzl_fftw_layer_wise_IDFT2_grid3D(int cols, int rows, int lays, double* arr, int threads);

/* grid3D<complex<double>>::layer_wise_DFT2 / layer_wise_IDFT2(scale, threads)
 * (grid3D_cx.hpp:81-82, :309-366, :370-427) on host memory: the seam's
 * transform (sign -1 forward, +1 backward), then every element multiplied by
 * the complex scale (scale_re, scale_im), as the reference's serial loop
 * does after the FFTW call; the scale is applied on the device.  Returns a
 * status (the reference returns void).  Uses the process-wide context of the
 * seam. */
int pfs_layer_wise_dft2(int cols, int rows, int lays, double* arr, int sign, double scale_re, double scale_im,
                        int threads);

/* Device form of the seam: in place on arr_dev, result multiplied by scale
 * (the reference's grid3D::layer_wise_DFT2/IDFT2 scale argument,
 * grid3D_cx.hpp:309-427). sign -1 forward, +1 backward. */
int pfs_layer_wise_dft2_dev(pfs_context* ctx, int n, int lays, void* arr_dev, int sign, double scale,
                            void* stream);

/* Public Wigner helpers on the device (include/PFSOFTlib_headers/fn_dwt.hpp:64,
 * :107, :239).  out is host memory.
 *   pfs_quadrature_weights: 2B values w_B(j).
 *   pfs_wigner_d_matrix:   (B - max(|M|,|M'|)) x 2B, column-major like the
 *                          reference matrix, rows l = max(|M|,|M'|) .. B-1,
 *                          columns beta_k; weighted != 0 multiplies column k
 *                          by w_B(k).  Values carry the reference's base sign
 *                          -1 (fn_dwt.hpp:124-125).  Requires M >= M' >= 0,
 *                          the only orders the transforms use. */
int pfs_quadrature_weights(pfs_context* ctx, int bandwidth, double* out);
int pfs_wigner_d_matrix(pfs_context* ctx, int bandwidth, int M, int Mp, int weighted, double* out);

/* Pipeline stages on device buffers (for stage timing and stage parity).
 * pencils: N^3 complex in the layout F(p, q, k) at (p*N + q)*N + k, the 2D
 * spectrum of every beta-slice with the slice index fastest. */
int pfs_stage_fft_fwd_dev(pfs_context* ctx, int bandwidth, const void* grid_dev, void* pencils_dev, void* stream);
int pfs_stage_dwt_fwd_dev(pfs_context* ctx, int bandwidth, const void* pencils_dev, void* coeffs_dev, void* stream);
int pfs_stage_dwt_inv_dev(pfs_context* ctx, int bandwidth, const void* coeffs_dev, void* pencils_dev, void* stream);
int pfs_stage_fft_inv_dev(pfs_context* ctx, int bandwidth, const void* pencils_dev, void* grid_dev, void* stream);

/* Tuning knobs (no counterpart in the reference).  Keys:
 *   "batch_interleave"  1 (default): batched contraction on signal-interleaved
 *                       coefficients with transposes; 0: on the caller's layout.
 *   "dwt_variant"       work shape of the FP64 DWT kernels for 2B <= 256 (2B = 512
 *                       always runs the node-blocked forward)
 *                       (0 default, 1..6 alternatives; results agree to
 *                       rounding).  Forward at 2B = 128, 256: 0 the
 *                       node-blocked kernel (one 64-node block at a time,
 *                       all degree chunks in accumulators), 4 / 5 the same
 *                       held to three / two waves per SIMD, 6 and 1..3 the
 *                       chain-parallel kernel (all nodes at once) in its
 *                       shapes.
 *   "dwt_inv_kernel"    inverse FP64 DWT kernel: 15 (default) the staged
 *                       vector-ALU kernel (coefficients staged in LDS a chunk
 *                       of 16 degrees at a time), 0 the MFMA kernel; 1..3 /
 *                       5..7 select the unstaged vector-ALU kernel's work
 *                       shapes (2 chains x 2 waves over the nodes, 4 x 2 over
 *                       the right-hand sides, 4 x 1; 5..7 with the row
 *                       broadcasts folded into the FMAs), 13 / 15 the staged
 *                       kernel in shapes 1 / 3 (bit-identical to 5 / 7);
 *                       results agree to rounding.
 *   "dwt_pair_order"    pair launch order of the FP64 DWT kernels, bit 0 the
 *                       forward's, bit 1 the inverse's: 0 = every XCD takes a
 *                       contiguous M' range of each M (default, both), 1 = runs
 *                       of 8 consecutive M' per XCD; results equal.
 *   "server_graphs"     see the request server below.
 *   "fft_chunk_slices"  beta-slices per FFT chunk (0 = automatic: one chunk);
 *                       rounded to a power of two >= 16 dividing 2B.  Smaller
 *                       chunks keep the row-pass output in the Infinity
 *                       Cache at the price of more, smaller launches.
 *   "batch_chunk_signals" signals per sub-batch of the batched path (work
 *                       buffers hold this many complex64 grids; default 64).
 *   "batch_fft_wide"    complex64 FFT passes with 16-byte accesses (two values
 *                       per thread): 2 (default) the forward passes of 2B <= 128,
 *                       1 every pass, 0 none; results equal bit for bit.
 *   "batch_fft_onepass" 1: the batched complex64 FFT at 2B = 64 and 128 as one
 *                       launch per direction (whole slices in LDS); 0 (default)
 *                       two passes; results equal bit for bit.
 *   "batch_graphs"      mixed-bandwidth batches replay a captured HIP graph
 *                       per composition (1, default) or launch directly (0).
 *   "batch_precision"   products of the batched Wigner contraction: 0 FP32
 *                       MFMA (default); 3 split-bf16 x3 (h h' + h m' + m h'
 *                       on bf16 MFMA, FP32 accumulation); 6 split-bf16 x6
 *                       (adds h l' + m m' + l h').  The bf16 forms keep the
 *                       bf16 terms of the Wigner tables resident (6 bytes per
 *                       table value and table) and need the context's bf16
 *                       MFMA self-test to have passed (else
 *                       PFS_ERR_UNSUPPORTED). */
int pfs_tune_set(pfs_context* ctx, const char* key, long value);
int pfs_tune_get(pfs_context* ctx, const char* key, long* value);

/* Per-kernel timing (no counterpart in the reference, whose only timer is a
 * wall clock around whole calls, src/stopwatch.cpp).  While enabled, every
 * kernel launch of the context is bracketed by HIP events on its launch
 * stream.  pfs_timing_collect waits for them and writes one CSV line per
 * kernel, "name,launches,total_ms", into buf, then forgets them. */
int pfs_timing_enable(pfs_context* ctx, int on);
int pfs_timing_collect(pfs_context* ctx, char* buf, size_t len);
/* Restricts timing to the kernels named in a comma-separated list (the names
 * pfs_timing_collect writes); NULL or "" times every kernel again.  Each timed
 * launch costs about 3.5 us of device time (21 us over the six launches of a
 * B = 128 round trip of 580 us), so a measurement times only what it needs.
 * No counterpart in the reference. */
int pfs_timing_select(pfs_context* ctx, const char* names);

/* Device memory helpers for callers without their own allocator. */
int pfs_device_alloc(pfs_context* ctx, size_t bytes, void** out);
int pfs_device_free(pfs_context* ctx, void* ptr);
int pfs_memcpy_h2d(pfs_context* ctx, void* dst_dev, const void* src_host, size_t bytes);
int pfs_memcpy_d2h(pfs_context* ctx, void* dst_host, const void* src_dev, size_t bytes);
/* dst <- src on stream, bytes a multiple of 16, in 16-byte accesses (the
 * access width of the complex128 FFT passes) or, with flags bit 1 (2), 8-byte
 * accesses (the complex64 passes); flags bit 0 (1): non-temporal loads and
 * stores, the cache hints of those passes.  A copy of known size for
 * calibrating hardware byte counters against the kernels' own access pattern
 * (no counterpart in the reference). */
int pfs_copy_dev(pfs_context* ctx, void* dst_dev, const void* src_dev, size_t bytes, int flags, void* stream);
/* Waits for stream; NULL waits for the context's own stream and the null stream. */
int pfs_synchronize(pfs_context* ctx, void* stream);

/* ---- request server ----------------------------------------------------
 * No counterpart in the reference, whose transforms are blocking calls
 * (include/PFSOFTlib_headers/fn_fourier_transforms.hpp:75,78) that are not
 * safe to call concurrently (src/fftw_wrapper.cpp:58-59,83).  A server on a
 * context accepts forward (complex64 grid -> packed complex64 coefficients)
 * and inverse (coefficients -> grid, Nyquist entries 0) requests of mixed
 * bandwidth on caller-owned device buffers, which must stay valid until the
 * request completes, and runs them with continuous batching on the batched
 * mixed-precision path: up to `slots` batches in flight on separate streams;
 * a free slot takes the oldest waiting request plus the waiting requests of
 * the same direction and bandwidth, up to its capacity.
 * hbm_budget_bytes (0 = only max_batch limits) bounds the server's own device
 * memory: resident FP32 tables of the served bandwidths plus the slots' work
 * buffers (two complex64 grids per signal of a batch).  A result does not
 * depend on which requests share its batch.  Each batch is ordered after the
 * work enqueued on the HIP null stream before it starts; work on other streams
 * that writes a request's input must be complete when it is submitted.  The
 * launch sequence of each (direction, bandwidth, batch size) is captured once
 * per slot into a HIP graph and replayed (context tune key "server_graphs",
 * read at creation: 1 default, 0 = launch kernels directly).  Submit and collect/wait may be called from any thread; a ticket
 * is consumed by pfs_server_wait or by pfs_server_collect, not both. */
typedef struct pfs_server pfs_server;

enum pfs_request_kind {
    PFS_REQ_FORWARD = 0, /* DSOFT:  in = grid, out = coefficients  */
    PFS_REQ_INVERSE = 1  /* iDSOFT: in = coefficients, out = grid  */
};

typedef struct pfs_server_stats_t {
    uint64_t completed;            /* requests finished (ok or failed)      */
    uint64_t batches;              /* batches launched                      */
    uint64_t signals;              /* signals over all batches              */
    uint64_t pending;              /* requests waiting for a slot           */
    int max_inflight_batches;      /* most batches on the device at once    */
    size_t device_bytes;           /* tables + work buffers held            */
    size_t slot_bytes;             /* work-buffer bytes per slot            */
    uint64_t graphs;               /* batch graphs captured (replayed after) */
    uint64_t tracked;              /* tickets the server still holds: not yet
                                      consumed by pfs_server_wait or
                                      pfs_server_collect                    */
} pfs_server_stats_t;

int pfs_server_create(pfs_context* ctx, const int* bandwidths, int nbw, int slots, int max_batch,
                      size_t hbm_budget_bytes, pfs_server** out);
/* Finishes every submitted request, then frees the server. */
int pfs_server_destroy(pfs_server* srv);
int pfs_server_submit(pfs_server* srv, int kind, int bandwidth, const void* in_dev, void* out_dev, uint64_t* ticket);
/* Blocks until the request is done (timeout_s < 0: no limit); latency_s =
 * completion - submission, host clock.  Returns the request's status, and
 * then forgets the ticket.  PFS_ERR_TIMEOUT: the request is still queued or
 * running -- its buffers are still in use and the ticket stays valid (wait
 * again, or take it from pfs_server_collect). */
int pfs_server_wait(pfs_server* srv, uint64_t ticket, double timeout_s, double* latency_s);
/* Up to cap requests finished since the last call, oldest first. */
int pfs_server_collect(pfs_server* srv, uint64_t* tickets, double* latency_s, int* status, int cap, int* n_out);
/* Blocks until nothing is waiting or running (PFS_ERR_TIMEOUT if the limit
 * passes first). */
int pfs_server_drain(pfs_server* srv, double timeout_s);
int pfs_server_stats(pfs_server* srv, pfs_server_stats_t* out);
/* Signals per batch the server admits at this bandwidth. */
int pfs_server_capacity(pfs_server* srv, int bandwidth, int* signals_per_batch);

#ifdef __cplusplus
}
#endif

#endif /* PFSOFT_AMD_H */
