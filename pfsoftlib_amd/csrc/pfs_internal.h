// Internal declarations shared by the HIP translation units of libpfsoft_amd.
//
// Layouts in HBM (N = 2B, complex values interleaved re/im):
//   grid    x[k][c][r]   element (row r, col c, beta-slice k) at N*N*k + N*c + r
//                        (the reference's grid3D index, grid3D_cx.hpp:293-303)
//   work W  W[k][p][c]   slice-transposed after the first FFT pass
//   pencils G[p][q][k]   F(p, q, k) at (p*N + q)*N + k: the 2D spectrum of every
//                        beta-slice with the beta index fastest, so the DWT reads
//                        each (p, q) beta-vector as one contiguous 16*N byte run
//   coeffs  packed       degree-l block at offset(l) = (4l^3 - l)/3, element
//                        (l, M, M') at offset(l) + iM + (2l+1)*iM', iX = X>=0 ? X : 2l+1+X
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <stdint.h>
#include <stddef.h>
#include <tuple>
#include <utility>

namespace pfs {

// ---- kernel launches -----------------------------------------------------
// Start/stop events the next kernel launch of this host thread records in its
// own dispatch packet (hipExtLaunchKernel), set by the context's timing wrapper
// around exactly one launch.  No barrier packets go into the stream, so timing
// a sequence of kernels does not slow it down.
struct LaunchTiming {
    hipEvent_t start, stop;
};
LaunchTiming& pending_timing();

template <typename... Formals, size_t... I>
inline hipError_t launch_args(void (*kernel)(Formals...), dim3 grid, dim3 block, size_t smem, hipStream_t st,
                              std::tuple<Formals...>& f, std::index_sequence<I...>)
{
    void* args[] = {(void*)&std::get<I>(f)..., nullptr};
    LaunchTiming& t = pending_timing();
    const hipEvent_t a = t.start, b = t.stop;
    t.start = t.stop = nullptr;
    if (a || b) return hipExtLaunchKernel((const void*)kernel, grid, block, args, smem, st, a, b, 0);
    return hipLaunchKernel((const void*)kernel, grid, block, args, smem, st);
}

// Launches kernel<<<grid, block, smem, st>>>(actuals...), the actuals converted
// to the kernel's parameter types.
template <typename... Formals, typename... Actuals>
inline hipError_t launch(void (*kernel)(Formals...), dim3 grid, dim3 block, size_t smem, hipStream_t st,
                         Actuals&&... actuals)
{
    static_assert(sizeof...(Formals) == sizeof...(Actuals), "kernel argument count");
    std::tuple<Formals...> f(std::forward<Actuals>(actuals)...);
    return launch_args(kernel, grid, block, smem, st, f, std::index_sequence_for<Formals...>{});
}

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// Lane rr of each 16-lane row of v, in every lane of that row: one DPP move
// (row_newbcast:rr) into a VGPR.  Two v_readlane into SGPRs cost twice the
// VALU issue, and a VALU op may read only one SGPR pair, so the SGPR form
// also needed a copy back to a VGPR per step (measured: 116.8 -> 112.6 us
// forward, 125.7 -> 121.8 us inverse at B = 128).
template <int RR> __device__ __forceinline__ double row_bcast_c(double v)
{
    return __builtin_amdgcn_mov_dpp(v, 0x150 + RR, 0xf, 0xf, false);
}
// rr is a constant once the caller's loop is unrolled; the switch folds away.
__device__ __forceinline__ double row_bcast(double v, int rr)
{
    switch (rr) {
    case 0: return row_bcast_c<0>(v);
    case 1: return row_bcast_c<1>(v);
    case 2: return row_bcast_c<2>(v);
    case 3: return row_bcast_c<3>(v);
    case 4: return row_bcast_c<4>(v);
    case 5: return row_bcast_c<5>(v);
    case 6: return row_bcast_c<6>(v);
    case 7: return row_bcast_c<7>(v);
    case 8: return row_bcast_c<8>(v);
    case 9: return row_bcast_c<9>(v);
    case 10: return row_bcast_c<10>(v);
    case 11: return row_bcast_c<11>(v);
    case 12: return row_bcast_c<12>(v);
    case 13: return row_bcast_c<13>(v);
    case 14: return row_bcast_c<14>(v);
    default: return row_bcast_c<15>(v);
    }
}

__host__ __device__ inline long coeff_offset(int l) { return ((long)4 * l * l * l - l) / 3; }

__host__ __device__ inline long coeff_index(int l, int M, int Mp)
{
    long n = 2 * l + 1;
    long iM = M >= 0 ? M : n + M;
    long iMp = Mp >= 0 ? Mp : n + Mp;
    return coeff_offset(l) + iM + n * iMp;
}

// 32-bit form for device index math: exact for l < 512 (4 l^3 < 2^31).
__device__ __forceinline__ int coeff_index32(int l, int M, int Mp)
{
    int n = 2 * l + 1;
    int iM = M >= 0 ? M : n + M;
    int iMp = Mp >= 0 ? Mp : n + Mp;
    return (4 * l * l * l - l) / 3 + iM + n * iMp;
}

// x^e by binary powering (e >= 0).  The plan's power tables are made with
// it, so a table entry equals the value a kernel would compute in place.
__host__ __device__ __forceinline__ double ipow(double x, int e)
{
    double r = 1.0;
    while (e) {
        if (e & 1) r *= x;
        x *= x;
        e >>= 1;
    }
    return r;
}

// Recurrence coefficients for the step from degree l = M + r to l + 1
// (fn_dwt.hpp:167-192), FP64: (f1, f1*f2, c1), zero from the last row on so
// the padded rows stay 0.
__host__ __device__ __forceinline__ void recurrence_step(int M, int Mp, int r, int R, double& f1, double& f1f2, double& c1)
{
    f1 = 0, f1f2 = 0, c1 = 0;
    if (r < R - 1) {
        double idx = M + r;
        double nrm = sqrt((2.0 * idx + 3.0) / (2.0 * idx + 1.0));
        double nom = (idx + 1.0) * (2.0 * idx + 1.0);
        double den = 1.0 / sqrt(((idx + 1) * (idx + 1) - (double)M * M) * ((idx + 1) * (idx + 1) - (double)Mp * Mp));
        f1 = nrm * nom * den;
        double f2 = 0;
        if (M + r != 0) {
            double t1 = sqrt((2.0 * idx + 3.0) / (2.0 * idx - 1.0)) * (idx + 1.0) / idx;
            double t2 = sqrt((idx * idx - (double)M * M) * (idx * idx - (double)Mp * Mp));
            c1 = -t1 * t2 * den;
            f2 = -(double)(M * Mp) / (idx * (idx + 1.0));
        }
        f1f2 = f1 * f2;
    }
}

// Device-resident tables for one bandwidth (all FP64 unless noted).
struct PlanTables {
    int B, N;
    int npairs;
    const double* cosb;    // cos(beta_k)
    const double* sinh;    // sin(beta_k / 2)
    const double* cosh;    // cos(beta_k / 2)
    const double* wts;     // quadrature weights w_B(k)
    const double* pws;     // sin(beta_k / 2)^e, [e * N + k], e < B (ipow)
    const double* pwc;     // cos(beta_k / 2)^e, [e * N + k], e < 2B - 1 (ipow)
    const double* normfac; // sqrt((2M+1)/2) * sqrt(C(2M, M-M')), [M*B + M'], M >= M' >= 0
    // The rescaled recurrence of pair (M, M'), rows r < RP = 16 ceil((B-M)/16)
    // at rsm[M] + M' RP + r: e_{r+1} = (a_r cos(beta) + b_r) e_r - e_{r-1}
    // with rab[.] = (a_r, b_r), and d^{M+r} = rp[.] e_r (zero rows from B - M
    // on).  Made for 2B = 512 plans only (null otherwise).
    const long* rsm;
    const double2* rab;
    const double* rp;
    const int2* pairs;     // canonical (M, M'), M >= M' >= 0, heaviest first
    const int2* xpairs;    // the same pairs in the FP64 DWT launch order (workgroup b runs xpairs[b])
    const double2* tw;     // exp(-2 pi i j / N), j < N
    const float2* tw32;    // same, rounded to FP32
};

// Output lane layout of v_mfma_f32_16x16x4f32 found by the context self-test:
// 0: lane L holds D[4v + L/16][L%16]; 1: D[4(L/16) + v][L%16].
extern int g_mfma_f32_layout;

// Largest dynamic LDS allocation a workgroup may request on the current device.
size_t lds_limit();

// Where the signals of a batch live: base + s * stride (complex elements), or,
// when ptrs is set, at the device addresses ptrs[s] (a device array).
struct SigSet {
    float2* base;
    long stride;
    float2* const* ptrs;
    SigSet at(int s0) const
    {
        return ptrs ? SigSet{nullptr, stride, ptrs + s0} : SigSet{base + (long)s0 * stride, stride, nullptr};
    }
};

// ---- FFT stage -----------------------------------------------------------
// A batch of length-N line transforms over a 3D array.  Line (u, v) element n
// sits at u*su + v*sv + n*sn (complex elements) on each side; u runs over
// blockIdx.y, v over the tiles of blockIdx.x, signals over blockIdx.z.
struct LineMap {
    long in_su, in_sv, in_sn;
    long out_su, out_sv, out_sn;
    long in_sig, out_sig;  // per-signal strides
    int nu, nv;            // extents of u and v
    // when set, signal s of the input (output) starts at in_ptrs[s] (out_ptrs[s])
    // instead of in + s * in_sig; device arrays of device addresses
    const void* const* in_ptrs;
    void* const* out_ptrs;
    long in_off, out_off;  // complex elements added to every signal's start
};

enum LineKind { LINES_ROW_IN_TILE_OUT = 0, LINES_TILE_IN_ROW_OUT = 1 };

// sign: -1 forward (FFTW_FORWARD), +1 backward.  scale multiplies the result.
hipError_t launch_fft_lines_f64(const double2* in, double2* out, const LineMap& m, int N, int sign,
                                double scale, const double2* tw, int kind, int nsig, hipStream_t st);
// paired: 16-byte accesses (two complex64 values per thread on each side)
// where the map's row- and tile-side strides are 1, at N = 64, 128, 256;
// 1 = every such pass, 2 = the forward-kind passes of N <= 128 only, 0 =
// none.  The results equal the 8-byte kernel's bit for bit.
hipError_t launch_fft_lines_f32(const float2* in, float2* out, const LineMap& m, int N, int sign,
                                float scale, const float2* tw, int kind, int nsig, hipStream_t st,
                                int paired = 0);
// One-pass slice FFT (complex64, N = 64 or 128): both line passes of every
// slice's 2D FFT in one launch.  sign -1: grids (input side of m) -> pencils
// G[p][q][k] (output side); sign +1: pencils -> grids, times scale.
bool fft_slice_supported(int N, size_t elem_bytes);
hipError_t launch_fft_slice_f32(const float2* in, float2* out, const LineMap& m, int N, int sign, float scale,
                                const float2* tw, int nsig, hipStream_t st);

// ---- DWT stage -----------------------------------------------------------
// Forward: G pencils (FP64) -> packed coefficients, scaled by pi/(2B^2).
// variant: work shape of the 2B <= 256 kernels (chains per lane x waves per
// pair); 0 is the default, see dwt.hip.
hipError_t launch_dwt_fwd_f64(const PlanTables& t, const double2* G, double2* coeffs, double norm,
                              hipStream_t st, int variant = 0);
// Inverse: packed coefficients -> G pencils.  Entries with p == B or q == B
// are untouched, or, for 2B <= 256 with nq_grid or nq_compact set, copied as
// launch_nyquist_gather_f64 would by the same launch.
hipError_t launch_dwt_inv_f64(const PlanTables& t, const double2* coeffs, double2* G, double norm,
                              hipStream_t st, int variant = 0, const double2* nq_grid = nullptr,
                              const double2* nq_compact = nullptr);


// ---- batched mixed-precision path (batch.hip) ------------------------------
bool batch_supported(int N);
// FP32 tables of the batched path for one bandwidth (layout in batch.hip).
struct BatchTables {
    const long* roff;    // per pair: first table row (multiple of 16)
    const float* tf;     // Wigner-d tiles times the quadrature weights, row-major within a tile
    const float* ti;     // the Wigner-d tiles (unweighted), column-major within a tile
    // bf16 terms (hi, mid, lo) of tf and ti for the split precisions, same
    // element order; null until a split precision is first used
    const unsigned short* tf16[3];
    const unsigned short* ti16[3];
};
// Whether the context self-test found the bf16 MFMA operand packing of the
// split precisions to give the FP32 MFMA's products (set by bf16_selftest).
extern bool g_mfma_bf16_ok;
int bf16_selftest(hipStream_t st);
// Bf16 terms h, m, l of n FP32 table values.
hipError_t launch_split_tables_bf16(const float* T, long n, unsigned short* h, unsigned short* m, unsigned short* l,
                                    hipStream_t st);
// Number of table rows (16-padded per pair) for bandwidth B.
long batch_table_rows(int B);
hipError_t launch_wigner_tiles_f32(const PlanTables& t, const BatchTables& bt, hipStream_t st);
// Work shape of a batched contraction launch: waves per workgroup (1, 2, 4) and
// signals per wave (1, 2, 4; at most 2 at 2B = 256 and for the inverse at
// 2B = 128).  {0, 0} = the default for the batch size.
struct BatchKernelShape {
    int wpb, spw;
};
BatchKernelShape batch_default_shape(int N, bool fwd, int nsig);
bool batch_shape_valid(int N, bool fwd, BatchKernelShape sh);
// Coefficient (s, idx) of the batch sits at s * cdist + idx * cstride (complex
// elements): (count, 1) is the per-signal layout, (1, ld) the interleaved one.
// prec: 0 FP32 MFMA; 3 / 6 split-bf16 products (three or six bf16 MFMAs per
// FP32 step group, FP32 accumulation; needs the bf16 tables and a passing
// bf16 self-test).
hipError_t launch_dwt_fwd_batch_f32(const PlanTables& t, const BatchTables& bt, const float2* G, float2* coeffs,
                                    int nsig, long cdist, long cstride, float norm, BatchKernelShape sh,
                                    hipStream_t st, int prec = 0);
hipError_t launch_dwt_inv_batch_f32(const PlanTables& t, const BatchTables& bt, const float2* coeffs, float2* G,
                                    int nsig, long cdist, long cstride, float norm, BatchKernelShape sh,
                                    hipStream_t st, int prec = 0);
// One bandwidth group of a mixed-bandwidth batched contraction
// (launch_dwt_batch_multi): the group's plan and FP32 tile table (Tf forward,
// Ti inverse), its input and output (forward: pencils -> coefficients,
// inverse: coefficients -> pencils), and the coefficient layout
// s * cdist + idx * cstride as in launch_dwt_fwd_batch_f32.  Lives in device
// memory: the kernel reads the group of each work item from there.
struct BatchMultiGroup {
    PlanTables t;
    const long* roff;
    const float* tab;
    const void* in;
    float2* out;
    int nsig;
    long cdist, cstride;
    float norm;
};
// Work shape of the mixed-bandwidth contraction: 4 waves x 2 signals per item.
constexpr int MULTI_WPB = 4, MULTI_SPW = 2;
// One persistent launch over work items int4{group, pair index, signal group,
// 0} (group < 0: padding), grid workgroups walking items b, b + grid, ...;
// nnmax: largest 2B among the groups.
hipError_t launch_dwt_batch_multi(bool fwd, const BatchMultiGroup* groups, const int4* items, int nitems, int nnmax,
                                  int grid, hipStream_t st);
// Between the per-signal layout (per_signal + s * count, or sptr[s] when sptr
// is set) and the interleaved layout [idx][ld], in the direction given.
hipError_t launch_coeff_transpose(float2* per_signal, float2* const* sptr, float2* interleaved, long count, int n,
                                  int ld, bool to_interleaved, hipStream_t st);
hipError_t launch_nyquist_zero_f32(float2* G, int B, int nsig, hipStream_t st);

// Copy the grid entries the inverse DWT never writes (row == B or col == B)
// into the pencil buffer, so the inverse FFT reads the caller's values there.
hipError_t launch_nyquist_gather_f64(const double2* grid, double2* G, int B, hipStream_t st,
                                     const double2* compact = nullptr);

// ---- batched pipeline entry points used by the request server ------------
}  // namespace pfs

struct pfs_context;

namespace pfs {
// Plan and FP32 tables for bandwidth B (built once; takes the context lock).
int batch_warm(pfs_context* c, int B);
// Bytes of each of the two work buffers for cap signals at bandwidth B.
size_t batch_work_bytes(int B, int cap);
// One batched forward (grids -> coefficients) or inverse transform of n <= cap
// complex64 signals on stream st with caller-owned work buffers W and G of
// batch_work_bytes(B, cap) bytes each.  Takes the context lock while enqueueing.
int batch_run(pfs_context* c, bool forward, int B, int n, int cap, SigSet in, SigSet out, float2* W, float2* G,
              hipStream_t st);
// Set this thread's pfs_last_error() text; returns code.
int set_error(int code, const char* msg);
// This thread's pfs_last_error() text.
const char* error_text();
// Device ordinal of a context.
int context_device(const pfs_context* c);
// Forward FP64 transform with caller-provided work buffers W and G (N^3
// complex128 each) for a bandwidth whose plan the context already holds;
// launches kernels only: no allocation, no timing events.  Takes the lock.
int dsoft_prepared(pfs_context* c, int B, const double2* grid, double2* coeffs, double2* W, double2* G,
                   hipStream_t st);
// Looks up the device address of the allocation-free entry point's scratch
// array on `device` (which must be current), once per device (noalloc.hip).
void noalloc_setup(int device);
// Whether the batched-shape autotuner has covered batches of up to `upto`
// signals at bandwidth B on this context.
bool batch_is_tuned(pfs_context* c, int B, int upto);
}  // namespace pfs

