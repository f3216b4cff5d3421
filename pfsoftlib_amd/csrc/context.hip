// Context, per-bandwidth plans and the C ABI (include/pfsoft_amd.h).
//
// A context owns one HIP stream on one device, the resident per-bandwidth
// tables and the work buffers of the single-signal pipeline:
//   forward  grid --FFT rows--> W --FFT cols--> pencils G --DWT--> coeffs
//   inverse  coeffs --iDWT--> G --IFFT cols--> W --IFFT rows, x 1/(4B^2)--> grid
// Plan tables are computed on the device in FP64: the Chebyshev nodes, the
// quadrature weights w_B (fn_dwt.hpp:62-83, compensated summation), the base
// norms sqrt((2M+1)/2) sqrt(C(2M, M-M')) (fn_dwt.hpp:117-121) and the FFT
// twiddles.
#include "pfs_internal.h"
#include "../../include/pfsoft_amd.h"

#include <algorithm>
#include <map>
#include <set>
#include <mutex>
#include <string>
#include <vector>
#include <thread>
#include <cstdio>
#include <cstdarg>
#include <cstring>
#include <cmath>

using namespace pfs;

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(e_ == hipErrorOutOfMemory ? PFS_ERR_MEMORY : PFS_ERR_DEVICE,       \
                        "%s failed: %s", #expr, hipGetErrorString(e_));                    \
    } while (0)

// Largest bandwidth the transforms accept: every kernel configuration up to
// 2B = 512 runs in the GPU tests; 2B = 1024 has never run, so it is refused.
constexpr int MAX_BANDWIDTH = 256;

__global__ void k_plan_nodes(int B, double* cosb, double* sinh, double* cosh, double* wts, double2* tw, float2* tw32)
{
    const int N = 2 * B;
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    double x = (2.0 * k + 1.0) / (4.0 * B);  // beta_k / pi
    cosb[k] = cospi(x);
    sinh[k] = sinpi(0.5 * x);
    cosh[k] = cospi(0.5 * x);
    // w_B(i) = 2/B sin(pi(2i+1)/(4B)) sum_{m<B} sin((2i+1)(2m+1) pi/(4B)) / (2m+1)
    int i = k < B ? k : N - 1 - k;
    double s = 0.0, comp = 0.0;
    for (int m = 0; m < B; ++m) {
        double a = (double)((2L * i + 1) * (2L * m + 1) % (8L * B)) / (4.0 * B);
        double term = sinpi(a) / (2.0 * m + 1.0) - comp;
        double tsum = s + term;
        comp = (tsum - s) - term;
        s = tsum;
    }
    wts[k] = 2.0 / B * sinpi((2.0 * i + 1.0) / (4.0 * B)) * s;
    double2 t;
    t.x = cospi(2.0 * k / N);
    t.y = -sinpi(2.0 * k / N);
    tw[k] = t;
    float2 tf;
    tf.x = (float)t.x;
    tf.y = (float)t.y;
    tw32[k] = tf;
}

// pws[e][k] = sin(beta_k / 2)^e (e < B), pwc[e][k] = cos(beta_k / 2)^e
// (e < 2B - 1): the base-row powers of every canonical pair, made once per
// bandwidth instead of in every transform.
__global__ void k_plan_powers(int B, const double* sinh, const double* cosh, double* pws, double* pwc)
{
    const int N = 2 * B, e = blockIdx.y;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    if (e < B) pws[(long)e * N + k] = ipow(sinh[k], e);
    if (e < 2 * B - 1) pwc[(long)e * N + k] = ipow(cosh[k], e);
}

// The rescaled recurrence of each pair (M, M'), one thread per pair walking
// its rows in order.  With d_r = P_r e_r, P_0 = P_1 = 1 and P_{r+1} =
// -c1_r P_{r-1} (c1_r < 0 for 1 <= r < R - 1; c1_0 = 0 and the row before 0
// is zero), the step d_{r+1} = (f1 cb + f1f2) d_r + c1 d_{r-1} becomes
// e_{r+1} = (a cb + b) e_r - e_{r-1} with (a, b) = (f1, f1f2) P_r / P_{r+1}:
// one fused multiply-add per node and degree fewer.  P stays within
// [0.11, 1.13] for every pair up to B = 256.  Rows from B - M - 1 on have
// a = b = 0; P is zero from B - M on (the pair's padding rows).
__global__ void k_plan_rescale(int B, const long* rsm, double2* rab, double* rp)
{
    const int M = blockIdx.x;
    for (int Mp = threadIdx.x; Mp <= M; Mp += blockDim.x) {
        const int R = B - M, RP = ((R + 15) >> 4) << 4;
        const long o = rsm[M] + (long)Mp * RP;
        double pm = 1.0, pc = 1.0;  // P_{r-1}, P_r
        for (int r = 0; r < RP; ++r) {
            double f1, f1f2, c1;
            recurrence_step(M, Mp, r, R, f1, f1f2, c1);
            if (r < R - 1) {
                const double pn = r == 0 ? 1.0 : -c1 * pm;
                const double rho = pc / pn;
                rab[o + r] = make_double2(f1 * rho, f1f2 * rho);
                rp[o + r] = pc;
                pm = pc;
                pc = pn;
            } else {
                rab[o + r] = make_double2(0.0, 0.0);
                rp[o + r] = r < R ? pc : 0.0;
            }
        }
    }
}

__global__ void k_plan_normfac(int B, double* nf)
{
    int M = blockIdx.x, Mp = threadIdx.x;
    for (; Mp <= M; Mp += blockDim.x) {
        double f = sqrt((2.0 * M + 1.0) / 2.0);
        for (int i = 0; i < M - Mp; ++i) f *= sqrt((2.0 * M - i) / (i + 1.0));
        nf[M * B + Mp] = f;
    }
}

// Wigner table for one (M, M'), column-major (B - M) x 2B, with the
// reference's base sign -1 (fn_dwt.hpp:124-125, :152-199).
__global__ void k_wigner_table(int B, int M, int Mp, const double* cosb, const double* sinh, const double* cosh,
                               const double* wts, const double* nf, int weighted, double* out)
{
    const int N = 2 * B, R = B - M;
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    double b = -nf[M * B + Mp];
    for (int e = 0; e < M - Mp; ++e) b *= sinh[k];
    for (int e = 0; e < M + Mp; ++e) b *= cosh[k];
    if (weighted) b *= wts[k];
    double cb = cosb[k], dp = 0.0, d = b;
    for (int r = 0; r < R; ++r) {
        out[(long)k * R + r] = d;
        if (r == R - 1) break;
        double idx = M + r;
        double nrm = sqrt((2.0 * idx + 3.0) / (2.0 * idx + 1.0));
        double nom = (idx + 1.0) * (2.0 * idx + 1.0);
        double den = 1.0 / sqrt(((idx + 1) * (idx + 1) - (double)M * M) * ((idx + 1) * (idx + 1) - (double)Mp * Mp));
        double f1 = nrm * nom * den, f2 = 0.0, c1 = 0.0;
        if (M + r != 0) {
            double t1 = sqrt((2.0 * idx + 3.0) / (2.0 * idx - 1.0)) * (idx + 1.0) / idx;
            double t2 = sqrt((idx * idx - (double)M * M) * (idx * idx - (double)Mp * Mp));
            c1 = -t1 * t2 * den;
            f2 = -(double)(M * Mp) / (idx * (idx + 1.0));
        }
        double dn = c1 * dp + d * f1 * (f2 + cb);
        dp = d;
        d = dn;
    }
}

// Self-test of the FP64 MFMA lane layout the DWT kernels rely on:
// A (16x4): lane i + 16kk holds A[i][kk]; B (4x16): lane j + 16kk holds
// B[kk][j]; D (16x16): lane L holds D[4v + L/16][L%16].  With a single
// non-zero K index ksel, A[i][ksel] = 2^i and B[ksel][j] = 3^j, every output is
// 2^i 3^j, which names its row and column uniquely.
__global__ void k_mfma_probe(double* out)
{
    const int lane = threadIdx.x;
    for (int t = 0; t < 2; ++t) {
        const int ksel = t == 0 ? 0 : 3;
        double a = (lane >> 4) == ksel ? ldexp(1.0, lane & 15) : 0.0;
        double b = 1.0;
        for (int e = 0; e < (lane & 15); ++e) b *= 3.0;
        if ((lane >> 4) != ksel) b = 0.0;
        v4d acc = {0.0, 0.0, 0.0, 0.0};
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
        for (int v = 0; v < 4; ++v) out[t * 256 + lane * 4 + v] = acc[v];
    }
}

// FP32 16x16x4 probe: A[i][ksel] = i + 1, B[ksel][j] = 2^(6j); the product
// (i+1) 2^(6j) is exact in FP32 and names row and column.
__global__ void k_mfma_probe_f32(float* out)
{
    const int lane = threadIdx.x;
    for (int t = 0; t < 2; ++t) {
        const int ksel = t == 0 ? 0 : 3;
        float a = (lane >> 4) == ksel ? (float)((lane & 15) + 1) : 0.0f;
        float b = (lane >> 4) == ksel ? ldexpf(1.0f, 6 * (lane & 15)) : 0.0f;
        v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
        for (int v = 0; v < 4; ++v) out[t * 256 + lane * 4 + v] = acc[v];
    }
}

int mfma_selftest_f32(hipStream_t st)
{
    float* d = nullptr;
    HIP_TRY(hipMalloc(&d, 512 * sizeof(float)));
    hipError_t e = pfs::launch(k_mfma_probe_f32, dim3(1), dim3(64), 0, st, d);
    float h[512];
    if (e == hipSuccess) e = hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(PFS_ERR_DEVICE, "MFMA f32 self-test: %s", hipGetErrorString(e));
    bool ok[2] = {true, true};  // layout 0: row 4v + L/16; layout 1: row 4(L/16) + v
    for (int t = 0; t < 2; ++t)
        for (int lane = 0; lane < 64; ++lane)
            for (int v = 0; v < 4; ++v) {
                float x = h[t * 256 + lane * 4 + v];
                int ex = 0;
                frexpf(x, &ex);
                int jj = (ex - 1) / 6;
                float im1 = ldexpf(x, -6 * jj);
                int i = (int)im1 - 1;
                bool exact = x > 0 && im1 == (float)(i + 1) && i >= 0 && i < 16;
                int col = lane & 15;
                if (!(exact && jj == col && i == 4 * v + (lane >> 4))) ok[0] = false;
                if (!(exact && jj == col && i == 4 * (lane >> 4) + v)) ok[1] = false;
            }
    if (ok[0]) g_mfma_f32_layout = 0;
    else if (ok[1]) g_mfma_f32_layout = 1;
    else return fail(PFS_ERR_UNSUPPORTED, "MFMA f32 16x16x4 output layout matches neither known form");
    return PFS_OK;
}

int mfma_selftest(hipStream_t st)
{
    double* d = nullptr;
    HIP_TRY(hipMalloc(&d, 512 * sizeof(double)));
    hipError_t e = pfs::launch(k_mfma_probe, dim3(1), dim3(64), 0, st, d);
    double h[512];
    if (e == hipSuccess) e = hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(PFS_ERR_DEVICE, "MFMA self-test: %s", hipGetErrorString(e));
    for (int t = 0; t < 2; ++t)
        for (int lane = 0; lane < 64; ++lane)
            for (int v = 0; v < 4; ++v) {
                double x = h[t * 256 + lane * 4 + v];
                int i = 0, j = 0;
                while (x > 0 && fmod(x, 2.0) == 0.0) { x /= 2.0; ++i; }
                while (x > 1 && fmod(x, 3.0) == 0.0) { x /= 3.0; ++j; }
                int ei = 4 * v + (lane >> 4), ej = lane & 15;
                if (x != 1.0 || i != ei || j != ej)
                    return fail(PFS_ERR_UNSUPPORTED,
                                "MFMA f64 16x16x4 layout differs from the assumed one: ksel=%d lane=%d v=%d gives "
                                "row %d col %d (rest %g), expected row %d col %d",
                                t ? 3 : 0, lane, v, i, j, x, ei, ej);
            }
    return PFS_OK;
}

struct Plan {
    int B = 0;
    PlanTables t{};
    void* mem = nullptr;  // one allocation for all tables
    size_t bytes = 0;
    // FP32 Wigner tables of the batched path (built on first use)
    void* d32 = nullptr;  // row offsets, FP32 weights and both tile tables
    size_t d32_bytes = 0;
    // bf16 terms of both tile tables for the split precisions (built on first use)
    void* d16 = nullptr;
    size_t d16_bytes = 0;
    BatchTables bt{};
    // batched contraction shapes picked by the autotuner per direction and
    // batch-size bucket (ceil(log2 n), capped at 7); {0, 0} = default rule
    BatchKernelShape tuned[2][8] = {};
    // launch orders of the pairs for the FP64 DWT kernels: 0 = per-XCD M'
    // ranges (xcd_pair_order), 1 = runs of 8 consecutive M' per XCD within
    // windows of 64 pairs
    const int2* xorder[2] = {nullptr, nullptr};
};

}  // namespace

struct pfs_context {
    int device = 0;
    size_t budget = 0;
    size_t used = 0;
    hipStream_t stream = nullptr;
    std::map<int, Plan> plans;
    std::mutex mu;
    // single-signal work buffers (grown on demand)
    void* work[2] = {nullptr, nullptr};
    size_t work_bytes = 0;
    // device staging for the host-buffer entry points
    void* io_grid = nullptr;
    size_t io_grid_bytes = 0;
    void* io_coeff = nullptr;
    size_t io_coeff_bytes = 0;
    // pinned staging for the host-buffer entry points: two chunks, each with
    // the event of the last copy that used it
    void* stage[2] = {nullptr, nullptr};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    // allocations handed out by pfs_device_alloc
    std::map<void*, size_t> user_allocs;
    // per-kernel timing: a start/stop event pair recorded around every launch
    // on the launch stream while enabled
    bool timing = false;
    std::set<std::string> timing_only;  // kernels to time (empty: all)
    std::vector<hipEvent_t> ev_pool;
    struct Rec { const char* name; hipEvent_t a, b; };
    std::vector<Rec> recs;
    // tuning knobs (pfs_tune_set); 0 = automatic
    long fft_chunk_slices = 0;
    long batch_chunk_signals = 64;
    long batch_interleave = 1;  // contraction kernels on signal-interleaved coefficients
    // batched contraction precision: 0 FP32 MFMA (default), 3 / 6 split-bf16
    // products on bf16 MFMA with FP32 accumulation
    long batch_precision = 0;
    long dwt_variant = 0;       // work shape of the FP64 DWT kernels (2B <= 256)
    // inverse DWT kernel: 0 MFMA; vector ALU with 2 chains x 2 waves over the
    // nodes, 4 chains x 2 waves over the right-hand sides, or 4 x 1 per pair
    // (2B <= 256; 4 x 2 at 2B = 512): 1..3 with separate broadcast moves,
    // 5..7 with the broadcasts folded into the FMAs; 13 and 15 the staged
    // kernel (k_dwt_inv_f64_staged) in shapes 1 and 3.
    // Measured: 7 takes B = 128 from 120 to 109-113 us, B = 256 from 1.49 to
    // 1.30 ms; 15 takes B = 128 from 110.2-111.5 to about 97 us (its FP64 FMA
    // count is 7's, its other VALU and SALU work per degree much smaller).
    long dwt_inv_kernel = 15;
    // pair launch order of the FP64 DWT kernels (Plan::xorder): bit 0 the
    // forward's, bit 1 the inverse's.  Measured at B = 128: the per-XCD M'
    // ranges halve the inverse's HBM fetch (183 -> 96 MB) and cut the forward's
    // writes (71 -> 54 MB) at the same kernel times (within 1 us), so 0 (both).
    long dwt_pair_order = 0;
    long server_graphs = 1;     // request servers replay captured HIP graphs
    long server_autotune = 1;   // request servers autotune the batched shapes at creation
    long batch_force_shape = 0; // 0: tuned / default shapes; else 10 * waves + signals per wave
    long batch_fft_signals = 0; // batched FFT: signals per chunk (0 = the whole sub-batch)
    long batch_fft_slices = 0;  // batched FFT: beta-slices per chunk (0 = all); a multiple of 32
    long batch_fft_onepass = 0; // batched FFT at 2B = 64, 128: one launch per direction, slices in LDS
    long batch_fft_wide = 2;    // batched FFT passes with 16-byte accesses: 0 none, 1 all, 2 forward passes of 2B <= 128
    // batched-path work buffers (complex64, batch_chunk_signals grids each)
    void* bwork[2] = {nullptr, nullptr};
    size_t bwork_bytes = 0;
    // mixed-bandwidth batches: group descriptors, work items and per-signal
    // address tables in device memory, filled from a pinned host copy; the
    // event of the last upload; the captured pipelines by batch composition
    void* mb_dev = nullptr;
    void* mb_host = nullptr;
    size_t mb_bytes = 0;
    hipEvent_t mb_ev = nullptr;
    hipStream_t mb_cap = nullptr;
    std::map<std::string, hipGraphExec_t> mb_graphs;
    long batch_graphs = 1;  // mixed-bandwidth batches replay captured HIP graphs
};

LaunchTiming& pfs::pending_timing()
{
    static thread_local LaunchTiming t = {nullptr, nullptr};
    return t;
}

namespace {

hipEvent_t take_event(pfs_context* c)
{
    if (!c->ev_pool.empty()) {
        hipEvent_t e = c->ev_pool.back();
        c->ev_pool.pop_back();
        return e;
    }
    // timing only: no system-scope fence (cache writeback + invalidate) at each
    // record, which would otherwise slow the launch that follows
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
    return e;
}

// Wraps one kernel launch: when timing is enabled the launch records a start
// and a stop event in its own dispatch packet.
template <typename F>
hipError_t timed(pfs_context* c, const char* name, hipStream_t st, F&& launch)
{
    if (!c->timing || (!c->timing_only.empty() && !c->timing_only.count(name))) return launch();
    hipEvent_t a = take_event(c), b = take_event(c);
    LaunchTiming& t = pending_timing();
    t.start = a;
    t.stop = b;
    hipError_t e = launch();
    const bool taken = t.start == nullptr && t.stop == nullptr;
    t.start = t.stop = nullptr;
    if (a && b && taken && e == hipSuccess) {
        c->recs.push_back({name, a, b});
    } else {
        if (a) c->ev_pool.push_back(a);
        if (b) c->ev_pool.push_back(b);
    }
    return e;
}

int ctx_alloc(pfs_context* c, void** p, size_t bytes)
{
    if (c->budget && c->used + bytes > c->budget)
        return fail(PFS_ERR_BUDGET, "allocation of %zu bytes exceeds the context budget (%zu of %zu used)", bytes,
                    c->used, c->budget);
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(PFS_ERR_MEMORY, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    c->used += bytes;
    return PFS_OK;
}

void ctx_free(pfs_context* c, void*& p, size_t bytes)
{
    if (p) {
        (void)hipFree(p);
        c->used -= bytes;
        p = nullptr;
    }
}

int check_bandwidth(int B)
{
    if (B < 1 || B > MAX_BANDWIDTH)
        return fail(PFS_ERR_UNSUPPORTED, "bandwidth %d outside the supported range 1..%d", B, MAX_BANDWIDTH);
    int N = 2 * B;
    if ((N & (N - 1)) != 0 && N > 512)
        return fail(PFS_ERR_UNSUPPORTED, "bandwidth %d: grid side %d is neither a power of two nor <= 512", B, N);
    return PFS_OK;
}

// Launch order of the canonical pairs for the FP64 DWT kernels (one pair per
// workgroup; workgroup b runs on XCD b % 8, the round-robin dispatch of the
// device).  Pairs stay heaviest first (M ascending).  The pairs M' = 0..M of
// each M are split into 8 contiguous M' ranges and range x goes to XCD x, so
// one XCD works at a time on neighbouring M' of one M and on the same M'
// range of the next few M.  In the packed column-major coefficient blocks,
// (l, +-M', .) are neighbours for consecutive M' and (l, +-M, .) for
// consecutive M: one L2 then serves the coefficient lines of all eight cases
// of neighbouring pairs.  Which XCD gets a range's extra pair rotates with M,
// so the eight queues stay within one pair of each other.  (Alternating each
// queue's heaviest and lightest pairs instead, so that pairs that mostly
// compute run beside pairs that mostly store, measured slower at B = 128:
// inverse 103 -> 128 us, forward 104 -> 137 us.)
std::vector<int2> xcd_pair_order(int B)
{
    constexpr int X = 8;
    std::vector<std::vector<int2>> q(X);
    int rot = 0;
    for (int M = 0; M < B; ++M) {
        const int n = M + 1, base = n / X, rem = n % X;
        int Mp = 0;
        for (int x = 0; x < X; ++x) {
            const int sz = base + (((x - rot + X) % X) < rem ? 1 : 0);
            for (int i = 0; i < sz; ++i, ++Mp) q[x].push_back(make_int2(M, Mp));
        }
        rot = (rot + rem) % X;
    }
    std::vector<int2> order;
    order.reserve((size_t)B * (B + 1) / 2);
    size_t maxlen = 0;
    for (auto& v : q) maxlen = std::max(maxlen, v.size());
    for (size_t j = 0; j < maxlen; ++j)
        for (int x = 0; x < X; ++x)
            if (j < q[x].size()) order.push_back(q[x][j]);
    return order;
}

int get_plan(pfs_context* c, int B, Plan** out)
{
    int rc = check_bandwidth(B);
    if (rc) return rc;
    auto it = c->plans.find(B);
    if (it != c->plans.end()) {
        *out = &it->second;
        return PFS_OK;
    }
    const int N = 2 * B;
    std::vector<int2> pairs;
    for (int M = 0; M < B; ++M)
        for (int Mp = 0; Mp <= M; ++Mp) pairs.push_back(make_int2(M, Mp));
    // heaviest first: rows B - M
    std::stable_sort(pairs.begin(), pairs.end(), [](int2 a, int2 b) { return a.x < b.x; });

    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    size_t off_cosb = 0;
    size_t off_sinh = off_cosb + align(N * sizeof(double));
    size_t off_cosh = off_sinh + align(N * sizeof(double));
    size_t off_wts = off_cosh + align(N * sizeof(double));
    size_t off_nf = off_wts + align(N * sizeof(double));
    size_t off_pairs = off_nf + align((size_t)B * B * sizeof(double));
    const std::vector<int2> xpairs = xcd_pair_order(B);
    std::vector<int2> wpairs(pairs.size());
    for (size_t b = 0; b < pairs.size(); ++b) {
        const size_t g = b >> 6, r = b & 63;
        wpairs[b] = (g + 1) * 64 > pairs.size() ? pairs[b] : pairs[g * 64 + (r & 7) * 8 + (r >> 3)];
    }
    size_t off_xpairs = off_pairs + align(pairs.size() * sizeof(int2));
    size_t off_wpairs = off_xpairs + align(xpairs.size() * sizeof(int2));
    size_t off_tw = off_wpairs + align(wpairs.size() * sizeof(int2));
    size_t off_tw32 = off_tw + align(N * sizeof(double2));
    size_t off_pws = off_tw32 + align(N * sizeof(float2));
    size_t off_pwc = off_pws + align((size_t)B * N * sizeof(double));
    // rescaled-recurrence rows (16 ceil((B - M) / 16) per pair, pairs in
    // (M, M') order), read by the inverse DWT at 2B = 512 only (67 MB at
    // B = 256)
    const bool rescaled = N > 256;
    std::vector<long> rsm(B, 0);
    size_t nrows = 0;
    for (int M = 0; M < B && rescaled; ++M) {
        rsm[M] = (long)nrows;
        nrows += (size_t)(M + 1) * (((B - M + 15) >> 4) << 4);
    }
    size_t off_rsm = off_pwc + align((size_t)(2 * B - 1) * N * sizeof(double));
    size_t off_rab = off_rsm + align(rsm.size() * sizeof(long));
    size_t off_rp = off_rab + align(nrows * sizeof(double2));
    size_t total = off_rp + align(nrows * sizeof(double));

    Plan p;
    p.B = B;
    rc = ctx_alloc(c, &p.mem, total);
    if (rc) return rc;
    p.bytes = total;
    char* base = (char*)p.mem;
    p.t.B = B;
    p.t.N = N;
    p.t.npairs = (int)pairs.size();
    p.t.cosb = (double*)(base + off_cosb);
    p.t.sinh = (double*)(base + off_sinh);
    p.t.cosh = (double*)(base + off_cosh);
    p.t.wts = (double*)(base + off_wts);
    p.t.normfac = (double*)(base + off_nf);
    p.t.pairs = (int2*)(base + off_pairs);
    p.t.xpairs = (int2*)(base + off_xpairs);
    p.xorder[0] = p.t.xpairs;
    p.xorder[1] = (int2*)(base + off_wpairs);
    p.t.tw = (double2*)(base + off_tw);
    p.t.tw32 = (float2*)(base + off_tw32);
    p.t.pws = (double*)(base + off_pws);
    p.t.pwc = (double*)(base + off_pwc);
    p.t.rsm = rescaled ? (const long*)(base + off_rsm) : nullptr;
    p.t.rab = rescaled ? (const double2*)(base + off_rab) : nullptr;
    p.t.rp = rescaled ? (const double*)(base + off_rp) : nullptr;
    hipError_t e = hipMemcpyAsync((void*)p.t.pairs, pairs.data(), pairs.size() * sizeof(int2), hipMemcpyHostToDevice,
                                  c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync((void*)p.t.xpairs, xpairs.data(), xpairs.size() * sizeof(int2), hipMemcpyHostToDevice,
                           c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync((void*)p.xorder[1], wpairs.data(), wpairs.size() * sizeof(int2), hipMemcpyHostToDevice,
                           c->stream);
    if (e == hipSuccess)
        e = pfs::launch(k_plan_nodes, dim3((N + 127) / 128), dim3(128), 0, c->stream, B, (double*)p.t.cosb,
                        (double*)p.t.sinh, (double*)p.t.cosh, (double*)p.t.wts, (double2*)p.t.tw, (float2*)p.t.tw32);
    if (e == hipSuccess) e = pfs::launch(k_plan_normfac, dim3(B), dim3(64), 0, c->stream, B, (double*)p.t.normfac);
    if (e == hipSuccess)
        e = pfs::launch(k_plan_powers, dim3((N + 127) / 128, 2 * B - 1), dim3(128), 0, c->stream, B, p.t.sinh, p.t.cosh,
                        (double*)p.t.pws, (double*)p.t.pwc);
    if (e == hipSuccess && rescaled)
        e = hipMemcpyAsync((void*)p.t.rsm, rsm.data(), rsm.size() * sizeof(long), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && rescaled)
        e = pfs::launch(k_plan_rescale, dim3(B), dim3(64), 0, c->stream, B, p.t.rsm, (double2*)p.t.rab,
                        (double*)p.t.rp);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        ctx_free(c, p.mem, p.bytes);
        return fail(PFS_ERR_DEVICE, "plan tables for B=%d: %s", B, hipGetErrorString(e));
    }
    auto ins = c->plans.emplace(B, p);
    *out = &ins.first->second;
    return PFS_OK;
}

int ensure_work(pfs_context* c, size_t bytes)
{
    if (c->work_bytes >= bytes) return PFS_OK;
    // the old buffers may still be in use on any stream a caller passed
    (void)hipDeviceSynchronize();
    ctx_free(c, c->work[0], c->work_bytes);
    ctx_free(c, c->work[1], c->work_bytes);
    c->work_bytes = 0;
    int rc = ctx_alloc(c, &c->work[0], bytes);
    if (rc) return rc;
    rc = ctx_alloc(c, &c->work[1], bytes);
    if (rc) {
        ctx_free(c, c->work[0], bytes);
        return rc;
    }
    c->work_bytes = bytes;
    return PFS_OK;
}

int ensure_io(pfs_context* c, size_t gbytes, size_t cbytes)
{
    if (c->io_grid_bytes < gbytes) {
        ctx_free(c, c->io_grid, c->io_grid_bytes);
        c->io_grid_bytes = 0;
        int rc = ctx_alloc(c, &c->io_grid, gbytes);
        if (rc) return rc;
        c->io_grid_bytes = gbytes;
    }
    if (c->io_coeff_bytes < cbytes) {
        ctx_free(c, c->io_coeff, c->io_coeff_bytes);
        c->io_coeff_bytes = 0;
        int rc = ctx_alloc(c, &c->io_coeff, cbytes);
        if (rc) return rc;
        c->io_coeff_bytes = cbytes;
    }
    return PFS_OK;
}

// ---- host <-> device copies through pinned staging ------------------------
// A pageable copy goes through the runtime's own single-threaded staging.  The
// host-buffer entry points instead copy chunks of the caller's buffer into two
// pinned buffers with several threads and overlap that with the DMA of the
// other buffer.
constexpr size_t STAGE_BYTES = 32u << 20;

int ensure_staging(pfs_context* c)
{
    for (int i = 0; i < 2; ++i) {
        if (!c->stage[i] && hipHostMalloc(&c->stage[i], STAGE_BYTES, hipHostMallocDefault) != hipSuccess) {
            c->stage[i] = nullptr;
            return fail(PFS_ERR_MEMORY, "hipHostMalloc(%zu) for pinned staging failed", STAGE_BYTES);
        }
        if (!c->stage_ev[i] && hipEventCreateWithFlags(&c->stage_ev[i], hipEventDisableTiming) != hipSuccess) {
            c->stage_ev[i] = nullptr;
            return fail(PFS_ERR_DEVICE, "hipEventCreate for pinned staging failed");
        }
    }
    return PFS_OK;
}

// memcpy split over a few threads (one host thread streams ~10 GB/s)
void par_memcpy(void* dst, const void* src, size_t n)
{
    const size_t min_part = 4u << 20;
    unsigned nt = std::min<unsigned>(8, std::max(1u, std::thread::hardware_concurrency()));
    nt = (unsigned)std::min<size_t>(nt, std::max<size_t>(1, n / min_part));
    if (nt <= 1) {
        std::memcpy(dst, src, n);
        return;
    }
    const size_t part = (n / nt + 63) & ~(size_t)63;
    std::vector<std::thread> th;
    size_t done = part;  // [0, done) is this thread's, the rest the workers' or, failing them, this thread's
    try {
        for (unsigned t = 1; t < nt; ++t) {
            const size_t off = t * part;
            if (off >= n) break;
            th.emplace_back([=] { std::memcpy((char*)dst + off, (const char*)src + off, std::min(part, n - off)); });
            done = std::min(n, off + part);
        }
    } catch (...) {
        // no more threads: copy what is left here (no exception crosses the C ABI)
    }
    std::memcpy(dst, src, std::min(part, n));
    if (done < n) std::memcpy((char*)dst + done, (const char*)src + done, n - done);
    for (auto& x : th) x.join();
}

int copy_h2d(pfs_context* c, void* dev, const void* host, size_t n, hipStream_t st)
{
    int rc = ensure_staging(c);
    if (rc) return rc;
    // the host writes the staging buffers: no earlier copy may still read them
    for (int b = 0; b < 2; ++b) HIP_TRY(hipEventSynchronize(c->stage_ev[b]));
    for (size_t off = 0, i = 0; off < n; off += STAGE_BYTES, ++i) {
        const int b = (int)(i & 1);
        const size_t len = std::min(STAGE_BYTES, n - off);
        if (i >= 2) HIP_TRY(hipEventSynchronize(c->stage_ev[b]));  // its last DMA is done
        par_memcpy(c->stage[b], (const char*)host + off, len);
        HIP_TRY(hipMemcpyAsync((char*)dev + off, c->stage[b], len, hipMemcpyHostToDevice, st));
        HIP_TRY(hipEventRecord(c->stage_ev[b], st));
    }
    return PFS_OK;
}

// Returns once the data is in host memory.
int copy_d2h(pfs_context* c, void* host, const void* dev, size_t n, hipStream_t st)
{
    int rc = ensure_staging(c);
    if (rc) return rc;
    const size_t nchunk = (n + STAGE_BYTES - 1) / STAGE_BYTES;
    auto issue = [&](size_t i) -> hipError_t {
        const int b = (int)(i & 1);
        const size_t off = i * STAGE_BYTES, len = std::min(STAGE_BYTES, n - off);
        hipError_t e = hipMemcpyAsync(c->stage[b], (const char*)dev + off, len, hipMemcpyDeviceToHost, st);
        return e == hipSuccess ? hipEventRecord(c->stage_ev[b], st) : e;
    };
    for (size_t i = 0; i < std::min<size_t>(2, nchunk); ++i) HIP_TRY(issue(i));
    for (size_t i = 0; i < nchunk; ++i) {
        const int b = (int)(i & 1);
        const size_t off = i * STAGE_BYTES, len = std::min(STAGE_BYTES, n - off);
        HIP_TRY(hipEventSynchronize(c->stage_ev[b]));
        par_memcpy((char*)host + off, c->stage[b], len);
        if (i + 2 < nchunk) HIP_TRY(issue(i + 2));
    }
    return PFS_OK;
}

inline size_t grid_bytes(int B) { size_t N = 2 * (size_t)B; return N * N * N * sizeof(double2); }
inline size_t coeff_bytes(int B) { return (size_t)coeff_offset(B) * sizeof(double2); }

// Forward row pass: x[k][c][r] -> W[k][p][c]  (lines u = k, v = c).
LineMap map_rows_fwd(int N)
{
    long n = N, n2 = n * n;
    LineMap m{n2, n, 1, n2, 1, n, 0, 0, N, N};
    return m;
}
// Forward column pass: W[k][p][c] -> G[p][q][k]  (lines u = p, v = k).
LineMap map_cols_fwd(int N)
{
    long n = N, n2 = n * n;
    LineMap m{n, n2, 1, n2, 1, n, 0, 0, N, N};
    return m;
}
// Inverse column pass: G[p][q][k] -> W[k][p][c]  (lines u = p, v = k).
LineMap map_cols_inv(int N)
{
    long n = N, n2 = n * n;
    LineMap m{n2, 1, n, n, n2, 1, 0, 0, N, N};
    return m;
}
// Inverse row pass: W[k][p][c] -> x[k][c][r]  (lines u = k, v = c).
LineMap map_rows_inv(int N)
{
    long n = N, n2 = n * n;
    LineMap m{n2, 1, n, n2, n, 1, 0, 0, N, N};
    return m;
}
// In-place 2D DFT, second pass: W[k][p][c] -> x[k][q][p]  (lines u = k, v = p).
LineMap map_cols_seam(int N)
{
    long n = N, n2 = n * n;
    LineMap m{n2, n, 1, n2, 1, n, 0, 0, N, N};
    return m;
}

// Device entry points take a caller stream; NULL is the HIP null stream.
hipStream_t pick(pfs_context* c, void* s)
{
    (void)c;
    return (hipStream_t)s;
}

// Slices per FFT chunk.  The row pass of a chunk writes W, the column pass
// reads it right back, so a chunk of a few tens of MB keeps W in the
// Infinity Cache (MALL).  Measured on MI355X at 2B = 256 the smaller launches
// cost more than the cache saves, so the automatic choice is one chunk.
int fft_chunk(const pfs_context* c, int N)
{
    long want = c->fft_chunk_slices;
    if (want <= 0) return N;
    int k = 16;  // the column pass moves 16 slices per workgroup
    while (k * 2 <= want && k * 2 <= N) k *= 2;
    if (k > N || N % k) k = N;
    return k;
}

int fft_fwd(pfs_context* c, const Plan& p, const double2* grid, double2* W, double2* G, hipStream_t st)
{
    const int N = p.t.N;
    const int K = fft_chunk(c, N);
    const long n2 = (long)N * N;
    for (int k0 = 0; k0 < N; k0 += K) {
        LineMap a = map_rows_fwd(N);
        a.nu = K;
        LineMap b = map_cols_fwd(N);
        b.nv = K;
        const double2* in = grid + k0 * n2;
        HIP_TRY(timed(c, "fft_fwd_rows", st, [&] { return launch_fft_lines_f64(in, W, a, N, -1, 1.0, p.t.tw, LINES_ROW_IN_TILE_OUT, 1, st); }));
        HIP_TRY(timed(c, "fft_fwd_cols", st, [&] { return launch_fft_lines_f64(W, G + k0, b, N, -1, 1.0, p.t.tw, LINES_ROW_IN_TILE_OUT, 1, st); }));
    }
    return PFS_OK;
}

int fft_inv(pfs_context* c, const Plan& p, const double2* G, double2* W, double2* grid, hipStream_t st)
{
    const int N = p.t.N;
    const int B = p.B;
    const int K = fft_chunk(c, N);
    const long n2 = (long)N * N;
    for (int k0 = 0; k0 < N; k0 += K) {
        LineMap b = map_cols_inv(N);
        b.nv = K;
        LineMap a = map_rows_inv(N);
        a.nu = K;
        double2* out = grid + k0 * n2;
        HIP_TRY(timed(c, "fft_inv_cols", st, [&] { return launch_fft_lines_f64(G + k0, W, b, N, +1, 1.0, p.t.tw, LINES_TILE_IN_ROW_OUT, 1, st); }));
        HIP_TRY(timed(c, "fft_inv_rows", st, [&] {
            return launch_fft_lines_f64(W, out, a, N, +1, 1. / (4. * B * B), p.t.tw, LINES_TILE_IN_ROW_OUT, 1, st);
        }));
    }
    return PFS_OK;
}

// The plan's tables with the FP64 DWT launch order the context selects for
// one direction.
PlanTables dwt_tables(const pfs_context* c, const Plan& p, bool fwd)
{
    PlanTables t = p.t;
    t.xpairs = p.xorder[(c->dwt_pair_order >> (fwd ? 0 : 1)) & 1];
    return t;
}

// Inverse DWT launch variant: the work shape in the low bits, the kernel kind above.
int inv_variant(const pfs_context* c) { return (int)c->dwt_variant | (int)(c->dwt_inv_kernel << 8); }

// pi / (B * 2B) and (B * 2B) / pi in double (src/fn_dsoft.cpp:119, src/fn_idsoft.cpp:109)
inline double fwd_norm(int B) { return 3.14159265358979323846 / (B * (2 * B)); }
inline double inv_norm(int B) { return (B * (2 * B)) / 3.14159265358979323846; }

int dsoft_dev(pfs_context* c, int B, const double2* grid, double2* coeffs, hipStream_t st)
{
    Plan* p;
    int rc = get_plan(c, B, &p);
    if (rc) return rc;
    rc = ensure_work(c, grid_bytes(B));
    if (rc) return rc;
    double2* W = (double2*)c->work[0];
    double2* G = (double2*)c->work[1];
    rc = fft_fwd(c, *p, grid, W, G, st);
    if (rc) return rc;
    HIP_TRY(timed(c, "dwt_fwd", st, [&] { return launch_dwt_fwd_f64(dwt_tables(c, *p, true), G, coeffs, fwd_norm(B), st, (int)c->dwt_variant); }));
    return PFS_OK;
}

// stale: the packed Nyquist entries of the output grid (see
// k_nyquist_gather_f64), or null to read them from grid itself.
int idsoft_dev(pfs_context* c, int B, const double2* coeffs, double2* grid, hipStream_t st,
               const double2* stale = nullptr)
{
    Plan* p;
    int rc = get_plan(c, B, &p);
    if (rc) return rc;
    rc = ensure_work(c, grid_bytes(B));
    if (rc) return rc;
    double2* W = (double2*)c->work[0];
    double2* G = (double2*)c->work[1];
    // up to 2B = 256 the inverse DWT launch copies the stale Nyquist entries itself
    const bool fused = p->t.N <= 256;
    if (!fused) HIP_TRY(timed(c, "nyquist_gather", st, [&] { return launch_nyquist_gather_f64(grid, G, B, st, stale); }));
    HIP_TRY(timed(c, "dwt_inv", st, [&] {
        return launch_dwt_inv_f64(dwt_tables(c, *p, false), coeffs, G, inv_norm(B), st, inv_variant(c), fused && !stale ? grid : nullptr,
                                  fused ? stale : nullptr);
    }));
    return fft_inv(c, *p, G, W, grid, st);
}

// x *= (sr + i si) over n complex values (the reference's serial scale loop
// after a layer-wise DFT2 with a complex scale, grid3D_cx.hpp:342-356).
__global__ void k_scale_cx(double2* x, long n, double sr, double si)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double2 v = x[i];
        x[i] = make_double2(v.x * sr - v.y * si, v.x * si + v.y * sr);
    }
}

// scale_im != 0: the transform runs unscaled, then one pass multiplies by the
// complex scale; a real scale is folded into the transform's second pass.
int dft2_dev(pfs_context* c, int n, int lays, void* arr_dev, int sign, double scale, hipStream_t st,
             double scale_im = 0.0)
{
    if (!c || !arr_dev || n < 2 || (n & 1) || lays < 1 || (sign != 1 && sign != -1))
        return fail(PFS_ERR_ARG, "pfs_layer_wise_dft2_dev: bad arguments (n must be even and >= 2)");
    int B = n / 2;
    Plan* p;
    int rc = get_plan(c, B, &p);
    if (rc) return rc;
    size_t bytes = (size_t)n * n * lays * sizeof(double2);
    rc = ensure_work(c, bytes);
    if (rc) return rc;
    double2* a = (double2*)arr_dev;
    double2* W = (double2*)c->work[0];
    const int K = std::min(fft_chunk(c, n), lays);
    const long n2 = (long)n * n;
    for (int k0 = 0; k0 < lays; k0 += K) {
        const int kc = std::min(K, lays - k0);
        LineMap m1 = map_rows_fwd(n);
        m1.nu = kc;
        LineMap m2 = map_cols_seam(n);
        m2.nu = kc;
        HIP_TRY(launch_fft_lines_f64(a + k0 * n2, W, m1, n, sign, 1.0, p->t.tw, LINES_ROW_IN_TILE_OUT, 1, st));
        HIP_TRY(launch_fft_lines_f64(W, a + k0 * n2, m2, n, sign, scale_im != 0.0 ? 1.0 : scale, p->t.tw,
                                     LINES_ROW_IN_TILE_OUT, 1, st));
    }
    if (scale_im != 0.0) {
        const long tot = n2 * lays;
        const int blocks = (int)std::min<long>((tot + 255) / 256, 8192);
        HIP_TRY(pfs::launch(k_scale_cx, dim3(blocks), dim3(256), 0, st, a, tot, scale, scale_im));
    }
    return PFS_OK;
}

// ---- batched mixed-precision pipeline ---------------------------------------

int ensure_table32(pfs_context* c, Plan& p, hipStream_t st)
{
    if (p.d32) return PFS_OK;
    const int B = p.B, N = p.t.N;
    std::vector<int2> pairs;
    for (int M = 0; M < B; ++M)
        for (int Mp = 0; Mp <= M; ++Mp) pairs.push_back(make_int2(M, Mp));
    std::stable_sort(pairs.begin(), pairs.end(), [](int2 a, int2 b) { return a.x < b.x; });
    std::vector<long> roff(pairs.size());
    long rows = 0;
    for (size_t i = 0; i < pairs.size(); ++i) {
        roff[i] = rows;
        rows += ((B - pairs[i].x + 15) / 16) * 16;
    }
    if (rows != batch_table_rows(B)) return fail(PFS_ERR_ARG, "batched table: row count mismatch");
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t o_tf = al(pairs.size() * sizeof(long));
    const size_t tab = (size_t)rows * N * sizeof(float);
    const size_t o_ti = o_tf + al(tab);
    const size_t bytes = o_ti + al(tab);
    int rc = ctx_alloc(c, &p.d32, bytes);
    if (rc) return rc;
    p.d32_bytes = bytes;
    char* base = (char*)p.d32;
    p.bt.roff = (const long*)base;
    p.bt.tf = (const float*)(base + o_tf);
    p.bt.ti = (const float*)(base + o_ti);
    HIP_TRY(hipMemcpyAsync((void*)p.bt.roff, roff.data(), pairs.size() * sizeof(long), hipMemcpyHostToDevice, st));
    HIP_TRY(launch_wigner_tiles_f32(p.t, p.bt, st));
    HIP_TRY(hipStreamSynchronize(st));
    return PFS_OK;
}

// The bf16 terms (hi, mid, lo) of both FP32 tile tables, for the split-bf16
// contraction: 3 x 2 bytes per table value and table.
int ensure_table_bf16(pfs_context* c, Plan& p, hipStream_t st)
{
    if (p.d16) return PFS_OK;
    if (!g_mfma_bf16_ok)
        return fail(PFS_ERR_UNSUPPORTED, "split-bf16 contraction: the bf16 MFMA self-test did not pass on this device");
    const long n = batch_table_rows(p.B) * p.t.N;
    const size_t part = ((size_t)n * sizeof(unsigned short) + 255) & ~(size_t)255;
    int rc = ctx_alloc(c, &p.d16, 6 * part);
    if (rc) return rc;
    p.d16_bytes = 6 * part;
    unsigned short* q[6];
    for (int i = 0; i < 6; ++i) q[i] = (unsigned short*)((char*)p.d16 + i * part);
    HIP_TRY(launch_split_tables_bf16(p.bt.tf, n, q[0], q[1], q[2], st));
    HIP_TRY(launch_split_tables_bf16(p.bt.ti, n, q[3], q[4], q[5], st));
    HIP_TRY(hipStreamSynchronize(st));
    for (int i = 0; i < 3; ++i) {
        p.bt.tf16[i] = q[i];
        p.bt.ti16[i] = q[3 + i];
    }
    return PFS_OK;
}

// Captured mixed-bandwidth pipelines name the work buffers and device tables
// directly: they are dropped whenever either moves.
void drop_multi_graphs(pfs_context* c)
{
    for (auto& kv : c->mb_graphs) (void)hipGraphExecDestroy(kv.second);
    c->mb_graphs.clear();
}

int ensure_bwork(pfs_context* c, size_t bytes)
{
    if (c->bwork_bytes >= bytes) return PFS_OK;
    (void)hipDeviceSynchronize();
    drop_multi_graphs(c);
    ctx_free(c, c->bwork[0], c->bwork_bytes);
    ctx_free(c, c->bwork[1], c->bwork_bytes);
    c->bwork_bytes = 0;
    int rc = ctx_alloc(c, &c->bwork[0], bytes);
    if (rc) return rc;
    rc = ctx_alloc(c, &c->bwork[1], bytes);
    if (rc) {
        ctx_free(c, c->bwork[0], bytes);
        return rc;
    }
    c->bwork_bytes = bytes;
    return PFS_OK;
}

int batch_prepare(pfs_context* c, int B, int nsig, Plan** pp, int* sb, hipStream_t st)
{
    if (nsig < 1) return fail(PFS_ERR_ARG, "batched transform: nsig must be >= 1");
    int rc = get_plan(c, B, pp);
    if (rc) return rc;
    if (!batch_supported((*pp)->t.N))
        return fail(PFS_ERR_UNSUPPORTED, "batched mixed-precision path supports B = 8, 16, 32, 64, 128 (got %d)", B);
    rc = ensure_table32(c, **pp, st);
    if (rc) return rc;
    if (c->batch_precision) {
        rc = ensure_table_bf16(c, **pp, st);
        if (rc) return rc;
    }
    *sb = (int)std::min<long>(c->batch_chunk_signals, nsig);
    const size_t N = 2 * (size_t)B;
    return ensure_bwork(c, (size_t)*sb * N * N * N * sizeof(float2));
}

int size_bucket(int n)
{
    int b = 0;
    while ((1 << b) < n && b < 7) ++b;
    return b;
}

// Contraction shape for a batch of n signals: the autotuned one for n's size
// bucket when the autotuner has covered it, else the default rule.
BatchKernelShape shape_for(const pfs_context* c, const Plan& p, bool fwd, int n)
{
    if (c->batch_force_shape) {
        const BatchKernelShape sh{(int)c->batch_force_shape / 10, (int)c->batch_force_shape % 10};
        if (batch_shape_valid(p.t.N, fwd, sh)) return sh;  // else the rule below
    }
    return p.tuned[fwd ? 1 : 0][size_bucket(n)];  // {0, 0} until the autotuner covered n's bucket
}

// Row length of the interleaved coefficient layout for a sub-batch of n
// signals in work buffers sized for cap signals: padded to 8 signals (32-byte
// runs) when the buffer has room.
int interleave_ld(int n, int cap, long count, long N3)
{
    const int ld = (n + 7) & ~7;
    return (long)ld * count <= (long)cap * N3 ? ld : n;
}

// Times every valid contraction shape on work buffers of n signals and keeps
// the fastest per direction and size bucket (n = 1, 2, 4, ... up to max_n).
int batch_autotune(pfs_context* c, int B, int max_n)
{
    Plan* p;
    int sb;
    int rc = batch_prepare(c, B, max_n, &p, &sb, c->stream);
    if (rc) return rc;
    const int N = p->t.N;
    const long N3 = (long)N * N * N, count = coeff_offset(B);
    float2* W = (float2*)c->bwork[0];
    float2* G = (float2*)c->bwork[1];
    const float fnorm = (float)fwd_norm(B), inorm = (float)inv_norm(B);
    hipEvent_t e0 = take_event(c), e1 = take_event(c);
    if (!e0 || !e1) return fail(PFS_ERR_DEVICE, "batch_autotune: event creation failed");
    // defined inputs: zero pencils and coefficients (timing does not depend on values)
    HIP_TRY(hipMemsetAsync(W, 0, (size_t)sb * N3 * sizeof(float2), c->stream));
    HIP_TRY(hipMemsetAsync(G, 0, (size_t)sb * N3 * sizeof(float2), c->stream));
    for (int n = 1; n <= sb; n *= 2) {
        const int ld = interleave_ld(n, sb, count, N3);
        for (int dir = 0; dir < 2; ++dir) {
            float best = 1e30f;
            BatchKernelShape pick{0, 0};
            for (int w : {1, 2, 4})
                for (int sp : {1, 2, 4}) {
                    const BatchKernelShape sh{w, sp};
                    if (!batch_shape_valid(N, dir == 1, sh)) continue;
                    float t = 1e30f;
                    for (int rep = 0; rep < 3; ++rep) {
                        HIP_TRY(hipEventRecord(e0, c->stream));
                        if (dir == 1)
                            HIP_TRY(launch_dwt_fwd_batch_f32(p->t, p->bt, G, W, n, 1, ld, fnorm, sh, c->stream));
                        else
                            HIP_TRY(launch_dwt_inv_batch_f32(p->t, p->bt, W, G, n, 1, ld, inorm, sh, c->stream));
                        HIP_TRY(hipEventRecord(e1, c->stream));
                        HIP_TRY(hipEventSynchronize(e1));
                        float ms = 0;
                        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
                        t = std::min(t, ms);
                    }
                    if (t < best) {
                        best = t;
                        pick = sh;
                    }
                }
            p->tuned[dir][size_bucket(n)] = pick;
        }
    }
    c->ev_pool.push_back(e0);
    c->ev_pool.push_back(e1);
    return PFS_OK;
}

int batch_fft_signals(const pfs_context* c, int n)
{
    return c->batch_fft_signals > 0 ? (int)std::min<long>(c->batch_fft_signals, n) : n;
}

int batch_fft_slices(const pfs_context* c, int N)
{
    const long k = c->batch_fft_slices;
    // the complex64 register passes move 32 slices per workgroup
    return (k >= 32 && k < N && N % k == 0) ? (int)k : N;
}

// The one-pass slice FFT serves the batched complex64 passes at 2B = 64 and
// 128 when the context turns it on.  It is off by default: measured on MI355X
// (512 signals at 2B = 128) it takes 8.2 ms forward and 13.9 ms inverse against
// 6.8 and 5.9 ms for the two passes.  Its pencil-side accesses are single
// 8-byte values (one slice per workgroup, beta fastest in the pencils), one
// L2 request per lane; with those accesses made contiguous (a timing probe
// writing and reading the wrong layout) the same kernel takes 4.8 / 4.1 ms.
bool use_slice_fft(const pfs_context* c, int N) { return c->batch_fft_onepass && fft_slice_supported(N, sizeof(float2)); }

// Forward batch of n <= cap signals: FFT rows (grids -> W), FFT columns
// (W -> G), DWT (G -> coefficients).  W and G hold cap grids each.
int dsoft_batch_run(pfs_context* c, Plan* p, int n, int cap, SigSet grids, SigSet coeffs, float2* W, float2* G,
                    hipStream_t st)
{
    const int B = p->B, N = p->t.N;
    const long N3 = (long)N * N * N, count = coeff_offset(B);
    const float norm = (float)fwd_norm(B);
    // slice FFTs in chunks of S signals x K slices: the row pass of a chunk
    // writes a compact W block that the column pass reads back at once, so a
    // chunk small enough stays in the Infinity Cache instead of going through HBM
    const int S = use_slice_fft(c, N) ? n : batch_fft_signals(c, n), K = batch_fft_slices(c, N);
    const long n2 = (long)N * N;
    if (use_slice_fft(c, N)) {
        LineMap m{};
        m.in_sig = grids.stride;
        m.in_ptrs = (const void* const*)grids.ptrs;
        m.out_sig = N3;
        HIP_TRY(timed(c, "batch_fft_fwd_slice", st, [&] { return launch_fft_slice_f32(grids.base, G, m, N, -1, 1.0f, p->t.tw32, n, st); }));
    }
    for (int g0 = 0; g0 < n && !use_slice_fft(c, N); g0 += S) {
        const int ns = std::min(S, n - g0);
        const SigSet gs = grids.at(g0);
        for (int k0 = 0; k0 < N; k0 += K) {
            LineMap a = map_rows_fwd(N), b = map_cols_fwd(N);
            a.nu = K;
            a.in_sig = gs.stride;
            a.in_ptrs = (const void* const*)gs.ptrs;
            a.in_off = k0 * n2;
            a.out_sig = b.in_sig = K * n2;
            b.nv = K;
            b.out_sig = N3;
            float2* Gc = G + g0 * N3 + k0;
            HIP_TRY(timed(c, "batch_fft_fwd_rows", st, [&] { return launch_fft_lines_f32(gs.base, W, a, N, -1, 1.0f, p->t.tw32, LINES_ROW_IN_TILE_OUT, ns, st, (int)c->batch_fft_wide); }));
            HIP_TRY(timed(c, "batch_fft_fwd_cols", st, [&] { return launch_fft_lines_f32(W, Gc, b, N, -1, 1.0f, p->t.tw32, LINES_ROW_IN_TILE_OUT, ns, st, (int)c->batch_fft_wide); }));
        }
    }
    if (c->batch_interleave || coeffs.ptrs) {
        // W is free once the column pass has run: the coefficients go there interleaved
        const int ld = interleave_ld(n, cap, count, N3);
        HIP_TRY(timed(c, "batch_dwt_fwd", st, [&] { return launch_dwt_fwd_batch_f32(p->t, p->bt, G, W, n, 1, ld, norm, shape_for(c, *p, true, n), st, (int)c->batch_precision); }));
        HIP_TRY(timed(c, "batch_coeff_transpose", st, [&] { return launch_coeff_transpose(coeffs.base, coeffs.ptrs, W, count, n, ld, false, st); }));
    } else {
        HIP_TRY(timed(c, "batch_dwt_fwd", st, [&] { return launch_dwt_fwd_batch_f32(p->t, p->bt, G, coeffs.base, n, coeffs.stride, 1, norm, shape_for(c, *p, true, n), st, (int)c->batch_precision); }));
    }
    return PFS_OK;
}

// Inverse batch of n <= cap signals: DWT (coefficients -> G, Nyquist pencils
// zeroed), FFT columns (G -> W), FFT rows (W -> grids) x 1/(4B^2).
int idsoft_batch_run(pfs_context* c, Plan* p, int n, int cap, SigSet coeffs, SigSet grids, float2* W, float2* G,
                     hipStream_t st)
{
    const int B = p->B, N = p->t.N;
    const long N3 = (long)N * N * N, count = coeff_offset(B);
    const float norm = (float)inv_norm(B);
    const float scale = (float)(1. / (4. * B * B));
    HIP_TRY(timed(c, "batch_nyquist_zero", st, [&] { return launch_nyquist_zero_f32(G, B, n, st); }));
    if (c->batch_interleave || coeffs.ptrs) {
        const int ld = interleave_ld(n, cap, count, N3);
        HIP_TRY(timed(c, "batch_coeff_transpose", st, [&] { return launch_coeff_transpose(coeffs.base, coeffs.ptrs, W, count, n, ld, true, st); }));
        HIP_TRY(timed(c, "batch_dwt_inv", st, [&] { return launch_dwt_inv_batch_f32(p->t, p->bt, W, G, n, 1, ld, norm, shape_for(c, *p, false, n), st, (int)c->batch_precision); }));
    } else {
        HIP_TRY(timed(c, "batch_dwt_inv", st, [&] { return launch_dwt_inv_batch_f32(p->t, p->bt, coeffs.base, G, n, coeffs.stride, 1, norm, shape_for(c, *p, false, n), st, (int)c->batch_precision); }));
    }
    const int S = use_slice_fft(c, N) ? n : batch_fft_signals(c, n), K = batch_fft_slices(c, N);
    const long n2 = (long)N * N;
    if (use_slice_fft(c, N)) {
        LineMap m{};
        m.in_sig = N3;
        m.out_sig = grids.stride;
        m.out_ptrs = (void* const*)grids.ptrs;
        HIP_TRY(timed(c, "batch_fft_inv_slice", st, [&] { return launch_fft_slice_f32(G, grids.base, m, N, +1, scale, p->t.tw32, n, st); }));
    }
    for (int g0 = 0; g0 < n && !use_slice_fft(c, N); g0 += S) {
        const int ns = std::min(S, n - g0);
        const SigSet gs = grids.at(g0);
        for (int k0 = 0; k0 < N; k0 += K) {
            LineMap b = map_cols_inv(N), a = map_rows_inv(N);
            b.nv = K;
            b.in_sig = N3;
            b.out_sig = a.in_sig = K * n2;
            a.nu = K;
            a.out_sig = gs.stride;
            a.out_ptrs = (void* const*)gs.ptrs;
            a.out_off = k0 * n2;
            const float2* Gc = G + g0 * N3 + k0;
            HIP_TRY(timed(c, "batch_fft_inv_cols", st, [&] { return launch_fft_lines_f32(Gc, W, b, N, +1, 1.0f, p->t.tw32, LINES_TILE_IN_ROW_OUT, ns, st, (int)c->batch_fft_wide); }));
            HIP_TRY(timed(c, "batch_fft_inv_rows", st, [&] { return launch_fft_lines_f32(W, gs.base, a, N, +1, scale, p->t.tw32, LINES_TILE_IN_ROW_OUT, ns, st, (int)c->batch_fft_wide); }));
        }
    }
    return PFS_OK;
}

int dsoft_batch_f32(pfs_context* c, int B, int nsig, const float2* grids, float2* coeffs, hipStream_t st)
{
    Plan* p;
    int sb;
    int rc = batch_prepare(c, B, nsig, &p, &sb, st);
    if (rc) return rc;
    const long N3 = (long)p->t.N * p->t.N * p->t.N, count = coeff_offset(B);
    const SigSet in{const_cast<float2*>(grids), N3, nullptr}, out{coeffs, count, nullptr};
    for (int s0 = 0; s0 < nsig; s0 += sb) {
        rc = dsoft_batch_run(c, p, std::min(sb, nsig - s0), sb, in.at(s0), out.at(s0), (float2*)c->bwork[0],
                             (float2*)c->bwork[1], st);
        if (rc) return rc;
    }
    return PFS_OK;
}

int idsoft_batch_f32(pfs_context* c, int B, int nsig, const float2* coeffs, float2* grids, hipStream_t st)
{
    Plan* p;
    int sb;
    int rc = batch_prepare(c, B, nsig, &p, &sb, st);
    if (rc) return rc;
    const long N3 = (long)p->t.N * p->t.N * p->t.N, count = coeff_offset(B);
    const SigSet in{const_cast<float2*>(coeffs), count, nullptr}, out{grids, N3, nullptr};
    for (int s0 = 0; s0 < nsig; s0 += sb) {
        rc = idsoft_batch_run(c, p, std::min(sb, nsig - s0), sb, in.at(s0), out.at(s0), (float2*)c->bwork[0],
                              (float2*)c->bwork[1], st);
        if (rc) return rc;
    }
    return PFS_OK;
}

// ---- batches of mixed bandwidth ----------------------------------------------

// One bandwidth group of a mixed batch as the host enqueues it.
struct MultiRun {
    Plan* p;
    int n;             // signals
    int ld;            // interleaved coefficient row length
    float2* W;         // the group's part of the work buffers
    float2* G;
    float2* const* in_ptrs;   // device tables of the group's signal addresses
    float2* const* out_ptrs;
};

int ensure_multi_tables(pfs_context* c, size_t bytes)
{
    if (c->mb_bytes >= bytes) return PFS_OK;
    if (c->mb_ev) HIP_TRY(hipEventSynchronize(c->mb_ev));
    (void)hipDeviceSynchronize();
    drop_multi_graphs(c);
    ctx_free(c, c->mb_dev, c->mb_bytes);
    if (c->mb_host) (void)hipHostFree(c->mb_host);
    c->mb_host = nullptr;
    c->mb_bytes = 0;
    const size_t want = std::max<size_t>(bytes, (size_t)1 << 20);
    int rc = ctx_alloc(c, &c->mb_dev, want);
    if (rc) return rc;
    if (hipHostMalloc(&c->mb_host, want, hipHostMallocDefault) != hipSuccess) {
        ctx_free(c, c->mb_dev, want);
        return fail(PFS_ERR_MEMORY, "mixed batch: pinned table allocation failed");
    }
    c->mb_bytes = want;
    if (!c->mb_ev) HIP_TRY(hipEventCreateWithFlags(&c->mb_ev, hipEventDisableTiming));
    return PFS_OK;
}

// The kernels of one mixed batch (no allocation, no host synchronisation, so
// the sequence can be captured into a graph).
int multi_enqueue(pfs_context* c, bool fwd, const std::vector<MultiRun>& runs, const BatchMultiGroup* dgroups,
                  const int4* ditems, int nitems, int nnmax, int grid, hipStream_t st)
{
    auto fft = [&](const MultiRun& r, bool forward) -> int {
        const int N = r.p->t.N, B = r.p->B;
        const long N3 = (long)N * N * N;
        if (use_slice_fft(c, N)) {
            LineMap m{};
            if (forward) {
                m.in_ptrs = (const void* const*)r.in_ptrs;
                m.out_sig = N3;
                HIP_TRY(timed(c, "batch_fft_fwd_slice", st, [&] { return launch_fft_slice_f32(nullptr, r.G, m, N, -1, 1.0f, r.p->t.tw32, r.n, st); }));
            } else {
                m.in_sig = N3;
                m.out_ptrs = (void* const*)r.out_ptrs;
                const float scale = (float)(1. / (4. * B * B));
                HIP_TRY(timed(c, "batch_fft_inv_slice", st, [&] { return launch_fft_slice_f32(r.G, nullptr, m, N, +1, scale, r.p->t.tw32, r.n, st); }));
            }
        } else if (forward) {
            LineMap a = map_rows_fwd(N), b = map_cols_fwd(N);
            a.in_ptrs = (const void* const*)r.in_ptrs;
            a.out_sig = b.in_sig = N3;
            b.out_sig = N3;
            HIP_TRY(timed(c, "batch_fft_fwd_rows", st, [&] { return launch_fft_lines_f32(nullptr, r.W, a, N, -1, 1.0f, r.p->t.tw32, LINES_ROW_IN_TILE_OUT, r.n, st, (int)c->batch_fft_wide); }));
            HIP_TRY(timed(c, "batch_fft_fwd_cols", st, [&] { return launch_fft_lines_f32(r.W, r.G, b, N, -1, 1.0f, r.p->t.tw32, LINES_ROW_IN_TILE_OUT, r.n, st, (int)c->batch_fft_wide); }));
        } else {
            LineMap b = map_cols_inv(N), a = map_rows_inv(N);
            b.in_sig = N3;
            b.out_sig = a.in_sig = N3;
            a.out_ptrs = (void* const*)r.out_ptrs;
            const float scale = (float)(1. / (4. * B * B));
            HIP_TRY(timed(c, "batch_fft_inv_cols", st, [&] { return launch_fft_lines_f32(r.G, r.W, b, N, +1, 1.0f, r.p->t.tw32, LINES_TILE_IN_ROW_OUT, r.n, st, (int)c->batch_fft_wide); }));
            HIP_TRY(timed(c, "batch_fft_inv_rows", st, [&] { return launch_fft_lines_f32(r.W, nullptr, a, N, +1, scale, r.p->t.tw32, LINES_TILE_IN_ROW_OUT, r.n, st, (int)c->batch_fft_wide); }));
        }
        return PFS_OK;
    };
    if (fwd) {
        for (const MultiRun& r : runs) {
            int rc = fft(r, true);
            if (rc) return rc;
        }
        HIP_TRY(timed(c, "batch_dwt_fwd_multi", st, [&] { return launch_dwt_batch_multi(true, dgroups, ditems, nitems, nnmax, grid, st); }));
        for (const MultiRun& r : runs)
            HIP_TRY(timed(c, "batch_coeff_transpose", st, [&] { return launch_coeff_transpose(nullptr, r.out_ptrs, r.W, coeff_offset(r.p->B), r.n, r.ld, false, st); }));
    } else {
        for (const MultiRun& r : runs) {
            HIP_TRY(timed(c, "batch_coeff_transpose", st, [&] { return launch_coeff_transpose(nullptr, (float2* const*)r.in_ptrs, r.W, coeff_offset(r.p->B), r.n, r.ld, true, st); }));
            HIP_TRY(timed(c, "batch_nyquist_zero", st, [&] { return launch_nyquist_zero_f32(r.G, r.p->B, r.n, st); }));
        }
        HIP_TRY(timed(c, "batch_dwt_inv_multi", st, [&] { return launch_dwt_batch_multi(false, dgroups, ditems, nitems, nnmax, grid, st); }));
        for (const MultiRun& r : runs) {
            int rc = fft(r, false);
            if (rc) return rc;
        }
    }
    return PFS_OK;
}

// Forward (grids -> coefficients) or inverse transform of nsig complex64
// signals of mixed bandwidth, signal i at bws[i] with its input and output at
// the device addresses in[i] and out[i].  Called with the context lock held.
int batch_multi(pfs_context* c, bool fwd, int nsig, const int* bws, const void* const* in, void* const* out,
                hipStream_t st)
{
    const char* who = fwd ? "pfs_dsoft_batch_multi_f32_dev" : "pfs_idsoft_batch_multi_f32_dev";
    if (nsig < 1 || !bws || !in || !out) return fail(PFS_ERR_ARG, "%s: bad arguments", who);
    // groups by bandwidth, widest first (their pairs are the heaviest work items)
    std::map<int, std::vector<int>, std::greater<int>> by_b;
    for (int i = 0; i < nsig; ++i) {
        if (!in[i] || !out[i]) return fail(PFS_ERR_ARG, "%s: null buffer address of signal %d", who, i);
        by_b[bws[i]].push_back(i);
    }
    std::vector<MultiRun> runs;
    size_t work = 0;
    int nnmax = 0;
    for (auto& kv : by_b) {
        Plan* p;
        int rc = get_plan(c, kv.first, &p);
        if (rc) return rc;
        if (!batch_supported(p->t.N))
            return fail(PFS_ERR_UNSUPPORTED, "%s: the batched path supports B = 8, 16, 32, 64, 128 (got %d)", who,
                        kv.first);
        rc = ensure_table32(c, *p, c->stream);
        if (rc) return rc;
        const size_t N3 = (size_t)p->t.N * p->t.N * p->t.N;
        runs.push_back(MultiRun{p, (int)kv.second.size(), 0, nullptr, nullptr, nullptr, nullptr});
        work += kv.second.size() * N3;
        nnmax = std::max(nnmax, p->t.N);
    }
    int rc = ensure_bwork(c, work * sizeof(float2));
    if (rc) return rc;
    // table layout: groups | in addresses | out addresses | items
    const int ng = (int)runs.size();
    std::vector<int4> items;
    {
        struct PairCost { long cost; int g, pi; };
        std::vector<PairCost> pc;
        for (int g = 0; g < ng; ++g) {
            const int B = runs[g].p->B, N = runs[g].p->t.N;
            int pi = 0;
            for (int M = 0; M < B; ++M)  // the plan's pair order: M ascending, then M'
                for (int Mp = 0; Mp <= M; ++Mp, ++pi) pc.push_back(PairCost{(long)(B - M) * N, g, pi});
        }
        std::stable_sort(pc.begin(), pc.end(), [](const PairCost& a, const PairCost& b) { return a.cost > b.cost; });
        std::vector<std::vector<int4>> q(8);
        for (size_t k = 0; k < pc.size(); ++k) {
            const int per = MULTI_WPB * MULTI_SPW;
            const int nsg = (runs[pc[k].g].n + per - 1) / per;
            for (int sg = 0; sg < nsg; ++sg) q[k % 8].push_back(make_int4(pc[k].g, pc[k].pi, sg, 0));
        }
        size_t len = 0;
        for (auto& v : q) len = std::max(len, v.size());
        items.assign(8 * len, make_int4(-1, 0, 0, 0));
        for (int x = 0; x < 8; ++x)
            for (size_t j = 0; j < q[x].size(); ++j) items[8 * j + x] = q[x][j];
    }
    const int nitems = (int)items.size();
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t o_in = al(ng * sizeof(BatchMultiGroup)), o_out = o_in + al(nsig * sizeof(void*)),
                 o_it = o_out + al(nsig * sizeof(void*)), bytes = o_it + al(items.size() * sizeof(int4));
    rc = ensure_multi_tables(c, bytes);
    if (rc) return rc;
    HIP_TRY(hipEventSynchronize(c->mb_ev));  // the previous upload from the pinned copy is done
    char* h = (char*)c->mb_host;
    char* d = (char*)c->mb_dev;
    BatchMultiGroup* hg = (BatchMultiGroup*)h;
    void** hin = (void**)(h + o_in);
    void** hout = (void**)(h + o_out);
    float2* W = (float2*)c->bwork[0];
    float2* G = (float2*)c->bwork[1];
    size_t off = 0, pos = 0;
    int g = 0;
    for (auto& kv : by_b) {
        MultiRun& r = runs[g];
        const Plan& p = *r.p;
        const long N3 = (long)p.t.N * p.t.N * p.t.N, count = coeff_offset(p.B);
        r.ld = interleave_ld(r.n, r.n, count, N3);
        r.W = W + off;
        r.G = G + off;
        r.in_ptrs = (float2* const*)(d + o_in) + pos;
        r.out_ptrs = (float2* const*)(d + o_out) + pos;
        for (int i : kv.second) {
            hin[pos] = const_cast<void*>(in[i]);
            hout[pos] = out[i];
            ++pos;
        }
        BatchMultiGroup& bg = hg[g];
        bg.t = p.t;
        bg.roff = p.bt.roff;
        bg.tab = fwd ? p.bt.tf : p.bt.ti;
        bg.in = fwd ? (const void*)r.G : (const void*)r.W;
        bg.out = fwd ? r.W : r.G;
        bg.nsig = r.n;
        bg.cdist = 1;
        bg.cstride = r.ld;
        bg.norm = (float)(fwd ? fwd_norm(p.B) : inv_norm(p.B));
        off += (size_t)r.n * N3;
        ++g;
    }
    memcpy(h + o_it, items.data(), items.size() * sizeof(int4));
    HIP_TRY(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, st));
    HIP_TRY(hipEventRecord(c->mb_ev, st));
    const BatchMultiGroup* dgroups = (const BatchMultiGroup*)d;
    const int4* ditems = (const int4*)(d + o_it);
    const int grid = std::min(nitems, 1024);  // persistent: a multiple of 8 (XCD residues kept)

    if (!c->batch_graphs) return multi_enqueue(c, fwd, runs, dgroups, ditems, nitems, nnmax, grid, st);
    // the kernels' arguments depend only on the composition (the addresses are
    // in the device tables), so one captured graph per composition is replayed
    std::string key = fwd ? "f" : "i";
    for (const MultiRun& r : runs) key += "|" + std::to_string(r.p->B) + "x" + std::to_string(r.n);
    auto it = c->mb_graphs.find(key);
    if (it == c->mb_graphs.end()) {
        if (!c->mb_cap) HIP_TRY(hipStreamCreateWithFlags(&c->mb_cap, hipStreamNonBlocking));
        const bool timing = c->timing;
        c->timing = false;  // no timing events inside a graph
        HIP_TRY(hipStreamBeginCapture(c->mb_cap, hipStreamCaptureModeThreadLocal));
        rc = multi_enqueue(c, fwd, runs, dgroups, ditems, nitems, nnmax, grid, c->mb_cap);
        hipGraph_t graph = nullptr;
        const hipError_t ec = hipStreamEndCapture(c->mb_cap, &graph);
        c->timing = timing;
        if (rc || ec != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc ? rc : fail(PFS_ERR_DEVICE, "%s: graph capture failed: %s", who, hipGetErrorString(ec));
        }
        hipGraphExec_t exec = nullptr;
        const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) return fail(PFS_ERR_DEVICE, "%s: graph instantiation failed: %s", who, hipGetErrorString(ei));
        it = c->mb_graphs.emplace(key, exec).first;
    }
    HIP_TRY(hipGraphLaunch(it->second, st));
    return PFS_OK;
}

std::mutex g_seam_mu;
pfs_context* g_seam_ctx = nullptr;

int seam_dft2(int cols, int rows, int lays, double* arr, int sign, double scale_re = 1.0, double scale_im = 0.0)
{
    g_last_error.clear();  // the seam returns void: callers read pfs_last_error()
    if (!arr || cols < 1 || rows < 1 || lays < 0) return fail(PFS_ERR_ARG, "layer_wise DFT2: bad arguments");
    if (cols != rows) return fail(PFS_ERR_UNSUPPORTED, "layer_wise DFT2: only square slices (cols == rows) are supported");
    if (lays == 0) return PFS_OK;
    std::lock_guard<std::mutex> g(g_seam_mu);
    if (!g_seam_ctx) {
        int rc = pfs_ctx_create(0, 0, &g_seam_ctx);
        if (rc) return rc;
    }
    pfs_context* c = g_seam_ctx;
    std::lock_guard<std::mutex> g2(c->mu);
    size_t bytes = (size_t)cols * rows * lays * sizeof(double2);
    int rc = ensure_io(c, bytes, 0);
    if (rc) return rc;
    rc = ensure_work(c, bytes);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(c->device));
    rc = copy_h2d(c, c->io_grid, arr, bytes, c->stream);
    if (rc) return rc;
    rc = dft2_dev(c, cols, lays, c->io_grid, sign, scale_re, c->stream, scale_im);
    if (rc) return rc;
    return copy_d2h(c, arr, c->io_grid, bytes, c->stream);
}

}  // namespace

// ===========================================================================
// dst <- src in 16-byte elements, ordinary or non-temporal loads and stores:
// the access widths and cache hints of the FFT passes.
template <bool NT>
__global__ void k_copy16(const double2* __restrict__ src, double2* __restrict__ dst, long n)
{
    typedef double v2d __attribute__((ext_vector_type(2)));
    const v2d* s = reinterpret_cast<const v2d*>(src);
    v2d* d = reinterpret_cast<v2d*>(dst);
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        if (NT) __builtin_nontemporal_store(__builtin_nontemporal_load(s + i), d + i);
        else d[i] = s[i];
    }
}

// The same with 8-byte accesses (one complex64 value per lane, the width of
// the batched FP32 FFT passes).
template <bool NT>
__global__ void k_copy8(const float2* __restrict__ src, float2* __restrict__ dst, long n)
{
    typedef float v2f __attribute__((ext_vector_type(2)));
    const v2f* s = reinterpret_cast<const v2f*>(src);
    v2f* d = reinterpret_cast<v2f*>(dst);
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        if (NT) __builtin_nontemporal_store(__builtin_nontemporal_load(s + i), d + i);
        else d[i] = s[i];
    }
}

extern "C" {

const char* pfs_last_error(void) { return g_last_error.c_str(); }

const char* pfs_version(void) { return "pfsoft_amd 0.1.0 (gfx950)"; }

int pfs_ctx_create(int device, size_t hbm_budget_bytes, pfs_context** out)
{
    if (!out) return fail(PFS_ERR_ARG, "pfs_ctx_create: out is NULL");
    *out = nullptr;
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(PFS_ERR_ARG, "pfs_ctx_create: device %d of %d", device, n);
    HIP_TRY(hipSetDevice(device));
    pfs_context* c = new pfs_context;
    c->device = device;
    c->budget = hbm_budget_bytes;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(PFS_ERR_DEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    int rc = mfma_selftest(c->stream);
    if (!rc) rc = mfma_selftest_f32(c->stream);
    // the split-bf16 contraction is available only if its operand packing
    // reproduces the FP32 MFMA's products here (checked once per process)
    if (!rc) {
        static std::once_flag bf16_once;
        std::call_once(bf16_once, [&] { (void)pfs::bf16_selftest(c->stream); });
    }
    // the device address of this device's allocation-free scratch array (once
    // per device; a failure is reported by the entry point that needs it)
    if (!rc) pfs::noalloc_setup(device);
    if (rc) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return rc;
    }
    *out = c;
    return PFS_OK;
}

int pfs_ctx_destroy(pfs_context* c)
{
    if (!c) return PFS_OK;
    {
        std::lock_guard<std::mutex> g(c->mu);
        (void)hipSetDevice(c->device);
        (void)hipStreamSynchronize(c->stream);
        for (auto& kv : c->plans) {
            ctx_free(c, kv.second.mem, kv.second.bytes);
            ctx_free(c, kv.second.d32, kv.second.d32_bytes);
            ctx_free(c, kv.second.d16, kv.second.d16_bytes);
        }
        ctx_free(c, c->bwork[0], c->bwork_bytes);
        ctx_free(c, c->bwork[1], c->bwork_bytes);
        for (auto& kv : c->mb_graphs) (void)hipGraphExecDestroy(kv.second);
        c->mb_graphs.clear();
        ctx_free(c, c->mb_dev, c->mb_bytes);
        if (c->mb_host) (void)hipHostFree(c->mb_host);
        if (c->mb_ev) (void)hipEventDestroy(c->mb_ev);
        if (c->mb_cap) (void)hipStreamDestroy(c->mb_cap);
        c->plans.clear();
        ctx_free(c, c->work[0], c->work_bytes);
        ctx_free(c, c->work[1], c->work_bytes);
        ctx_free(c, c->io_grid, c->io_grid_bytes);
        ctx_free(c, c->io_coeff, c->io_coeff_bytes);
        for (int i = 0; i < 2; ++i) {
            if (c->stage[i]) (void)hipHostFree(c->stage[i]);
            if (c->stage_ev[i]) (void)hipEventDestroy(c->stage_ev[i]);
            c->stage[i] = nullptr;
            c->stage_ev[i] = nullptr;
        }
        for (auto& kv : c->user_allocs) {
            void* ptr = kv.first;
            ctx_free(c, ptr, kv.second);
        }
        c->user_allocs.clear();
        for (auto& r : c->recs) {
            c->ev_pool.push_back(r.a);
            c->ev_pool.push_back(r.b);
        }
        c->recs.clear();
        for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
        c->ev_pool.clear();
        (void)hipStreamDestroy(c->stream);
    }
    delete c;
    return PFS_OK;
}

int pfs_plan_warm(pfs_context* c, int B)
{
    if (!c) return fail(PFS_ERR_ARG, "null context");
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    Plan* p;
    return get_plan(c, B, &p);
}

long pfs_coeff_count(int B) { return B < 0 ? 0 : coeff_offset(B); }

long pfs_coeff_index(int l, int M, int Mp)
{
    if (l < 0 || M > l || M < -l || Mp > l || Mp < -l) return -1;
    return coeff_index(l, M, Mp);
}

int pfs_dsoft_f64(pfs_context* c, int B, const double* grid, double* coeffs)
{
    if (!c || !grid || !coeffs) return fail(PFS_ERR_ARG, "pfs_dsoft_f64: null argument");
    int rc = check_bandwidth(B);
    if (rc) return rc;
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    rc = ensure_io(c, grid_bytes(B), coeff_bytes(B));
    if (rc) return rc;
    rc = copy_h2d(c, c->io_grid, grid, grid_bytes(B), c->stream);
    if (rc) return rc;
    rc = dsoft_dev(c, B, (const double2*)c->io_grid, (double2*)c->io_coeff, c->stream);
    if (rc) return rc;
    return copy_d2h(c, coeffs, c->io_coeff, coeff_bytes(B), c->stream);
}

int pfs_idsoft_f64(pfs_context* c, int B, const double* coeffs, double* grid)
{
    if (!c || !grid || !coeffs) return fail(PFS_ERR_ARG, "pfs_idsoft_f64: null argument");
    int rc = check_bandwidth(B);
    if (rc) return rc;
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    rc = ensure_io(c, grid_bytes(B), coeff_bytes(B));
    if (rc) return rc;
    rc = copy_h2d(c, c->io_coeff, coeffs, coeff_bytes(B), c->stream);
    if (rc) return rc;
    // Of the caller's grid only the stale Nyquist entries are read (the
    // reference transforms in place: src/fn_idsoft.cpp:100-283 never writes
    // them, grid3D_cx.hpp:393-427 then reads them).  They go up packed, into
    // the front of the output buffer, which the last FFT pass overwrites after
    // they have been read: 2 MB instead of the 268 MB grid at B = 128.
    const long N = 2L * B;
    std::vector<double2> stale((size_t)(2 * N - 1) * N);
    const double2* hg = reinterpret_cast<const double2*>(grid);
    for (long k = 0; k < N; ++k) {
        const double2* slice = hg + N * N * k;
        for (long e = 0; e < 2 * N - 1; ++e) {
            long p, q;
            if (e < N) { p = B; q = e; }
            else { q = B; p = e - N; if (p >= B) ++p; }
            stale[(size_t)(e * N + k)] = slice[N * q + p];
        }
    }
    HIP_TRY(hipMemcpyAsync(c->io_grid, stale.data(), stale.size() * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    rc = idsoft_dev(c, B, (const double2*)c->io_coeff, (double2*)c->io_grid, c->stream, (const double2*)c->io_grid);
    if (rc) return rc;
    return copy_d2h(c, grid, c->io_grid, grid_bytes(B), c->stream);
}

int pfs_dsoft_f64_dev(pfs_context* c, int B, const void* grid_dev, void* coeffs_dev, void* stream)
{
    if (!c || !grid_dev || !coeffs_dev) return fail(PFS_ERR_ARG, "pfs_dsoft_f64_dev: null argument");
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    return dsoft_dev(c, B, (const double2*)grid_dev, (double2*)coeffs_dev, pick(c, stream));
}

int pfs_idsoft_f64_dev(pfs_context* c, int B, const void* coeffs_dev, void* grid_dev, void* stream)
{
    if (!c || !grid_dev || !coeffs_dev) return fail(PFS_ERR_ARG, "pfs_idsoft_f64_dev: null argument");
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    return idsoft_dev(c, B, (const double2*)coeffs_dev, (double2*)grid_dev, pick(c, stream));
}

void pfs_layer_wise_DFT2_grid3D(int cols, int rows, int lays, double* arr, int threads)
{
    (void)threads;
    int rc = seam_dft2(cols, rows, lays, arr, -1);
    if (rc) fprintf(stderr, "** [pfsoft_amd error] layer_wise_DFT2_grid3D: %s **\n", pfs_last_error());
}

void pfs_layer_wise_IDFT2_grid3D(int cols, int rows, int lays, double* arr, int threads)
{
    (void)threads;
    int rc = seam_dft2(cols, rows, lays, arr, +1);
    if (rc) fprintf(stderr, "** [pfsoft_amd error] layer_wise_IDFT2_grid3D: %s **\n", pfs_last_error());
}

void uzl_fftw_layer_wise_DFT2_grid3D(int cols, int rows, int lays, double* arr, int threads)
{
    pfs_layer_wise_DFT2_grid3D(cols, rows, lays, arr, threads);
}

void uzl_fftw_layer_wise_IDFT2_grid3D(int cols, int rows, int lays, double* arr, int threads)
{
    pfs_layer_wise_IDFT2_grid3D(cols, rows, lays, arr, threads);
}

int pfs_layer_wise_dft2(int cols, int rows, int lays, double* arr, int sign, double scale_re, double scale_im,
                        int threads)
{
    (void)threads;
    if (sign != 1 && sign != -1) return fail(PFS_ERR_ARG, "pfs_layer_wise_dft2: sign must be -1 or +1");
    return seam_dft2(cols, rows, lays, arr, sign, scale_re, scale_im);
}

int pfs_layer_wise_dft2_dev(pfs_context* c, int n, int lays, void* arr_dev, int sign, double scale, void* stream)
{
    if (!c) return fail(PFS_ERR_ARG, "null context");
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    return dft2_dev(c, n, lays, arr_dev, sign, scale, pick(c, stream));
}

int pfs_quadrature_weights(pfs_context* c, int B, double* out)
{
    if (!c || !out) return fail(PFS_ERR_ARG, "pfs_quadrature_weights: null argument");
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    Plan* p;
    int rc = get_plan(c, B, &p);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(out, p->t.wts, 2 * (size_t)B * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return PFS_OK;
}

int pfs_wigner_d_matrix(pfs_context* c, int B, int M, int Mp, int weighted, double* out)
{
    if (!c || !out) return fail(PFS_ERR_ARG, "pfs_wigner_d_matrix: null argument");
    if (!(M >= Mp && Mp >= 0 && M < B)) return fail(PFS_ERR_ARG, "pfs_wigner_d_matrix: need B > M >= M' >= 0");
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    Plan* p;
    int rc = get_plan(c, B, &p);
    if (rc) return rc;
    size_t bytes = (size_t)(B - M) * 2 * B * sizeof(double);
    rc = ensure_io(c, bytes, 0);
    if (rc) return rc;
    HIP_TRY(pfs::launch(k_wigner_table, dim3((2 * B + 63) / 64), dim3(64), 0, c->stream, B, M, Mp, p->t.cosb,
                        p->t.sinh, p->t.cosh, p->t.wts, p->t.normfac, weighted, (double*)c->io_grid));
    HIP_TRY(hipMemcpyAsync(out, c->io_grid, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return PFS_OK;
}

int pfs_stage_fft_fwd_dev(pfs_context* c, int B, const void* grid_dev, void* pencils_dev, void* stream)
{
    if (!c || !grid_dev || !pencils_dev) return fail(PFS_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    Plan* p;
    int rc = get_plan(c, B, &p);
    if (rc) return rc;
    rc = ensure_work(c, grid_bytes(B));
    if (rc) return rc;
    return fft_fwd(c, *p, (const double2*)grid_dev, (double2*)c->work[0], (double2*)pencils_dev, pick(c, stream));
}

int pfs_stage_dwt_fwd_dev(pfs_context* c, int B, const void* pencils_dev, void* coeffs_dev, void* stream)
{
    if (!c || !pencils_dev || !coeffs_dev) return fail(PFS_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    Plan* p;
    int rc = get_plan(c, B, &p);
    if (rc) return rc;
    HIP_TRY(launch_dwt_fwd_f64(dwt_tables(c, *p, true), (const double2*)pencils_dev, (double2*)coeffs_dev, fwd_norm(B), pick(c, stream),
                               (int)c->dwt_variant));
    return PFS_OK;
}

int pfs_stage_dwt_inv_dev(pfs_context* c, int B, const void* coeffs_dev, void* pencils_dev, void* stream)
{
    if (!c || !pencils_dev || !coeffs_dev) return fail(PFS_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    Plan* p;
    int rc = get_plan(c, B, &p);
    if (rc) return rc;
    HIP_TRY(launch_dwt_inv_f64(dwt_tables(c, *p, false), (const double2*)coeffs_dev, (double2*)pencils_dev, inv_norm(B), pick(c, stream),
                               inv_variant(c)));
    return PFS_OK;
}

int pfs_stage_fft_inv_dev(pfs_context* c, int B, const void* pencils_dev, void* grid_dev, void* stream)
{
    if (!c || !pencils_dev || !grid_dev) return fail(PFS_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    Plan* p;
    int rc = get_plan(c, B, &p);
    if (rc) return rc;
    rc = ensure_work(c, grid_bytes(B));
    if (rc) return rc;
    return fft_inv(c, *p, (const double2*)pencils_dev, (double2*)c->work[0], (double2*)grid_dev, pick(c, stream));
}

int pfs_dsoft_batch_f32_dev(pfs_context* c, int B, int nsig, const void* grids_dev, void* coeffs_dev, void* stream)
{
    if (!c || !grids_dev || !coeffs_dev) return fail(PFS_ERR_ARG, "pfs_dsoft_batch_f32_dev: null argument");
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    return dsoft_batch_f32(c, B, nsig, (const float2*)grids_dev, (float2*)coeffs_dev, pick(c, stream));
}

int pfs_idsoft_batch_f32_dev(pfs_context* c, int B, int nsig, const void* coeffs_dev, void* grids_dev, void* stream)
{
    if (!c || !grids_dev || !coeffs_dev) return fail(PFS_ERR_ARG, "pfs_idsoft_batch_f32_dev: null argument");
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    return idsoft_batch_f32(c, B, nsig, (const float2*)coeffs_dev, (float2*)grids_dev, pick(c, stream));
}

int pfs_dsoft_batch_multi_f32_dev(pfs_context* c, int nsig, const int* bandwidths, const void* const* grids_dev,
                                  void* const* coeffs_dev, void* stream)
{
    if (!c) return fail(PFS_ERR_ARG, "null context");
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    return batch_multi(c, true, nsig, bandwidths, grids_dev, coeffs_dev, pick(c, stream));
}

int pfs_idsoft_batch_multi_f32_dev(pfs_context* c, int nsig, const int* bandwidths, const void* const* coeffs_dev,
                                   void* const* grids_dev, void* stream)
{
    if (!c) return fail(PFS_ERR_ARG, "null context");
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    return batch_multi(c, false, nsig, bandwidths, coeffs_dev, grids_dev, pick(c, stream));
}

int pfs_tune_set(pfs_context* c, const char* key, long value)
{
    if (!c || !key) return fail(PFS_ERR_ARG, "pfs_tune_set: null argument");
    std::lock_guard<std::mutex> g(c->mu);
    if (!strcmp(key, "fft_chunk_slices")) {
        if (value < 0) return fail(PFS_ERR_ARG, "fft_chunk_slices must be >= 0");
        c->fft_chunk_slices = value;
        return PFS_OK;
    }
    if (!strcmp(key, "batch_chunk_signals")) {
        if (value < 1) return fail(PFS_ERR_ARG, "batch_chunk_signals must be >= 1");
        c->batch_chunk_signals = value;
        return PFS_OK;
    }
    if (!strcmp(key, "batch_interleave")) {
        if (value != 0 && value != 1) return fail(PFS_ERR_ARG, "batch_interleave must be 0 or 1");
        c->batch_interleave = value;
        return PFS_OK;
    }
    if (!strcmp(key, "batch_graphs")) {
        if (value != 0 && value != 1) return fail(PFS_ERR_ARG, "batch_graphs must be 0 or 1");
        c->batch_graphs = value;
        return PFS_OK;
    }
    if (!strcmp(key, "batch_precision")) {
        if (value != 0 && value != 3 && value != 6) return fail(PFS_ERR_ARG, "batch_precision must be 0, 3 or 6");
        if (value && !g_mfma_bf16_ok)
            return fail(PFS_ERR_UNSUPPORTED, "batch_precision %ld: the bf16 MFMA self-test did not pass", value);
        c->batch_precision = value;
        return PFS_OK;
    }
    if (!strcmp(key, "dwt_inv_kernel")) {
        if (value < 0 || value > 15 || value == 4 || (value >= 8 && value != 13 && value != 15))
            return fail(PFS_ERR_ARG, "dwt_inv_kernel must be 0 (MFMA), 1..3 or 5..7 (vector ALU), 13 or 15 (staged)");
        c->dwt_inv_kernel = value;
        return PFS_OK;
    }
    if (!strcmp(key, "dwt_pair_order")) {
        if (value < 0 || value > 3) return fail(PFS_ERR_ARG, "dwt_pair_order must be in 0..3 (bit 0 forward, bit 1 inverse)");
        c->dwt_pair_order = value;
        return PFS_OK;
    }
    if (!strcmp(key, "dwt_variant")) {
        if (value < 0 || value > 6) return fail(PFS_ERR_ARG, "dwt_variant must be in 0..6");
        c->dwt_variant = value;
        return PFS_OK;
    }
    if (!strcmp(key, "server_graphs")) {
        if (value != 0 && value != 1) return fail(PFS_ERR_ARG, "server_graphs must be 0 or 1");
        c->server_graphs = value;
        return PFS_OK;
    }
    if (!strcmp(key, "server_autotune")) {
        if (value != 0 && value != 1) return fail(PFS_ERR_ARG, "server_autotune must be 0 or 1");
        c->server_autotune = value;
        return PFS_OK;
    }
    if (!strcmp(key, "batch_fft_signals")) {
        if (value < 0) return fail(PFS_ERR_ARG, "batch_fft_signals must be >= 0");
        c->batch_fft_signals = value;
        return PFS_OK;
    }
    if (!strcmp(key, "batch_fft_slices")) {
        if (value < 0 || (value && (value % 32))) return fail(PFS_ERR_ARG, "batch_fft_slices must be 0 or a multiple of 32");
        c->batch_fft_slices = value;
        return PFS_OK;
    }
    if (!strcmp(key, "batch_fft_wide")) {
        if (value < 0 || value > 2) return fail(PFS_ERR_ARG, "batch_fft_wide must be 0, 1 or 2");
        if (value != c->batch_fft_wide) {
            (void)hipDeviceSynchronize();
            drop_multi_graphs(c);
        }
        c->batch_fft_wide = value;
        return PFS_OK;
    }
    if (!strcmp(key, "batch_fft_onepass")) {
        if (value != 0 && value != 1) return fail(PFS_ERR_ARG, "batch_fft_onepass must be 0 or 1");
        if (value != c->batch_fft_onepass) {
            // captured mixed batches hold the other kernels
            (void)hipDeviceSynchronize();
            drop_multi_graphs(c);
        }
        c->batch_fft_onepass = value;
        return PFS_OK;
    }
    if (!strcmp(key, "batch_force_shape")) {
        const BatchKernelShape sh{(int)value / 10, (int)value % 10};
        if (value != 0 && !(batch_shape_valid(256, true, sh) || batch_shape_valid(64, true, sh)))
            return fail(PFS_ERR_ARG, "batch_force_shape must be 0 or 10 * waves + signals per wave (1/2/4 each)");
        c->batch_force_shape = value;
        return PFS_OK;
    }
    return fail(PFS_ERR_ARG, "pfs_tune_set: unknown key '%s'", key);
}

int pfs_tune_get(pfs_context* c, const char* key, long* value)
{
    if (!c || !key || !value) return fail(PFS_ERR_ARG, "pfs_tune_get: null argument");
    std::lock_guard<std::mutex> g(c->mu);
    if (!strcmp(key, "fft_chunk_slices")) {
        *value = c->fft_chunk_slices;
        return PFS_OK;
    }
    if (!strcmp(key, "batch_chunk_signals")) {
        *value = c->batch_chunk_signals;
        return PFS_OK;
    }
    if (!strcmp(key, "batch_interleave")) {
        *value = c->batch_interleave;
        return PFS_OK;
    }
    if (!strcmp(key, "batch_precision")) {
        *value = c->batch_precision;
        return PFS_OK;
    }
    if (!strcmp(key, "batch_graphs")) {
        *value = c->batch_graphs;
        return PFS_OK;
    }
    if (!strcmp(key, "dwt_inv_kernel")) {
        *value = c->dwt_inv_kernel;
        return PFS_OK;
    }
    if (!strcmp(key, "dwt_pair_order")) {
        *value = c->dwt_pair_order;
        return PFS_OK;
    }
    if (!strcmp(key, "dwt_variant")) {
        *value = c->dwt_variant;
        return PFS_OK;
    }
    if (!strcmp(key, "server_graphs")) {
        *value = c->server_graphs;
        return PFS_OK;
    }
    if (!strcmp(key, "server_autotune")) {
        *value = c->server_autotune;
        return PFS_OK;
    }
    if (!strcmp(key, "batch_fft_signals")) {
        *value = c->batch_fft_signals;
        return PFS_OK;
    }
    if (!strcmp(key, "batch_fft_slices")) {
        *value = c->batch_fft_slices;
        return PFS_OK;
    }
    if (!strcmp(key, "batch_fft_onepass")) {
        *value = c->batch_fft_onepass;
        return PFS_OK;
    }
    if (!strcmp(key, "batch_fft_wide")) {
        *value = c->batch_fft_wide;
        return PFS_OK;
    }
    if (!strcmp(key, "batch_force_shape")) {
        *value = c->batch_force_shape;
        return PFS_OK;
    }
    return fail(PFS_ERR_ARG, "pfs_tune_get: unknown key '%s'", key);
}

int pfs_timing_enable(pfs_context* c, int on)
{
    if (!c) return fail(PFS_ERR_ARG, "null context");
    std::lock_guard<std::mutex> g(c->mu);
    c->timing = on != 0;
    return PFS_OK;
}

int pfs_timing_select(pfs_context* c, const char* names)
{
    if (!c) return fail(PFS_ERR_ARG, "null context");
    std::lock_guard<std::mutex> g(c->mu);
    c->timing_only.clear();
    if (!names) return PFS_OK;
    std::string cur;
    for (const char* q = names;; ++q) {
        if (*q == ',' || *q == 0) {
            if (!cur.empty()) c->timing_only.insert(cur);
            cur.clear();
            if (*q == 0) break;
        } else if (*q != ' ') {
            cur += *q;
        }
    }
    return PFS_OK;
}

int pfs_timing_collect(pfs_context* c, char* buf, size_t len)
{
    if (!c || !buf || len == 0) return fail(PFS_ERR_ARG, "pfs_timing_collect: bad arguments");
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    std::map<std::string, std::pair<long, double>> acc;
    for (auto& r : c->recs) {
        HIP_TRY(hipEventSynchronize(r.b));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, r.a, r.b));
        auto& e = acc[r.name];
        e.first += 1;
        e.second += ms;
        c->ev_pool.push_back(r.a);
        c->ev_pool.push_back(r.b);
    }
    c->recs.clear();
    std::string out;
    for (auto& kv : acc) {
        char line[160];
        snprintf(line, sizeof(line), "%s,%ld,%.6f\n", kv.first.c_str(), kv.second.first, kv.second.second);
        out += line;
    }
    if (out.size() + 1 > len) return fail(PFS_ERR_ARG, "pfs_timing_collect: buffer too small");
    memcpy(buf, out.c_str(), out.size() + 1);
    return PFS_OK;
}

int pfs_device_alloc(pfs_context* c, size_t bytes, void** out)
{
    if (!c || !out) return fail(PFS_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    int rc = ctx_alloc(c, out, bytes);
    if (rc == PFS_OK) c->user_allocs[*out] = bytes;
    return rc;
}

int pfs_device_free(pfs_context* c, void* ptr)
{
    if (!c) return fail(PFS_ERR_ARG, "null argument");
    if (!ptr) return PFS_OK;
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    auto it = c->user_allocs.find(ptr);
    if (it == c->user_allocs.end()) return fail(PFS_ERR_ARG, "pfs_device_free: pointer not allocated by this context");
    size_t bytes = it->second;
    c->user_allocs.erase(it);
    ctx_free(c, ptr, bytes);
    return PFS_OK;
}

int pfs_memcpy_h2d(pfs_context* c, void* dst, const void* src, size_t bytes)
{
    if (!c || !dst || !src) return fail(PFS_ERR_ARG, "null argument");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return PFS_OK;
}

int pfs_memcpy_d2h(pfs_context* c, void* dst, const void* src, size_t bytes)
{
    if (!c || !dst || !src) return fail(PFS_ERR_ARG, "null argument");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return PFS_OK;
}

int pfs_copy_dev(pfs_context* c, void* dst_dev, const void* src_dev, size_t bytes, int flags, void* stream)
{
    if (!c || !dst_dev || !src_dev || (bytes & 15) || ((uintptr_t)dst_dev & 15) || ((uintptr_t)src_dev & 15))
        return fail(PFS_ERR_ARG, "pfs_copy_dev: null argument, or a size or address not a multiple of 16 bytes");
    if (flags & ~3) return fail(PFS_ERR_ARG, "pfs_copy_dev: flags other than 1 (non-temporal) and 2 (8-byte accesses)");
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    const bool nontemporal = flags & 1;
    if (flags & 2) {
        const long n8 = (long)(bytes / 8);
        const int blocks8 = (int)std::min<long>((n8 + 255) / 256, 8192);
        if (nontemporal)
            HIP_TRY(timed(c, "copy8_nt", pick(c, stream), [&] { return pfs::launch(k_copy8<true>, dim3(blocks8), dim3(256), 0, pick(c, stream), (const float2*)src_dev, (float2*)dst_dev, n8); }));
        else
            HIP_TRY(timed(c, "copy8", pick(c, stream), [&] { return pfs::launch(k_copy8<false>, dim3(blocks8), dim3(256), 0, pick(c, stream), (const float2*)src_dev, (float2*)dst_dev, n8); }));
        return PFS_OK;
    }
    const long n = (long)(bytes / 16);
    const int blocks = (int)std::min<long>((n + 255) / 256, 8192);
    if (n == 0) return PFS_OK;
    if (nontemporal)
        HIP_TRY(timed(c, "copy16_nt", pick(c, stream), [&] { return pfs::launch(k_copy16<true>, dim3(blocks), dim3(256), 0, pick(c, stream), (const double2*)src_dev, (double2*)dst_dev, n); }));
    else
        HIP_TRY(timed(c, "copy16", pick(c, stream), [&] { return pfs::launch(k_copy16<false>, dim3(blocks), dim3(256), 0, pick(c, stream), (const double2*)src_dev, (double2*)dst_dev, n); }));
    return PFS_OK;
}

int pfs_synchronize(pfs_context* c, void* stream)
{
    if (!c) return fail(PFS_ERR_ARG, "null argument");
    HIP_TRY(hipSetDevice(c->device));
    if (stream) {
        HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    } else {
        HIP_TRY(hipStreamSynchronize(c->stream));
        HIP_TRY(hipStreamSynchronize(nullptr));
    }
    return PFS_OK;
}

}  // extern "C"

// ---- batched pipeline for the request server -------------------------------

namespace pfs {

int batch_warm(pfs_context* c, int B)
{
    std::lock_guard<std::mutex> g(c->mu);
    Plan* p;
    int rc = get_plan(c, B, &p);
    if (rc) return rc;
    if (!batch_supported(p->t.N))
        return fail(PFS_ERR_UNSUPPORTED, "batched mixed-precision path supports B = 8, 16, 32, 64, 128 (got %d)", B);
    rc = ensure_table32(c, *p, c->stream);
    if (!rc && c->batch_precision) rc = ensure_table_bf16(c, *p, c->stream);
    return rc;
}

size_t batch_work_bytes(int B, int cap)
{
    const size_t N = 2 * (size_t)B;
    return (size_t)cap * N * N * N * sizeof(float2);
}

int batch_run(pfs_context* c, bool forward, int B, int n, int cap, SigSet in, SigSet out, float2* W, float2* G,
              hipStream_t st)
{
    std::lock_guard<std::mutex> g(c->mu);
    if (n < 1 || n > cap) return fail(PFS_ERR_ARG, "batch_run: n = %d outside [1, %d]", n, cap);
    auto it = c->plans.find(B);
    if (it == c->plans.end() || !it->second.d32) return fail(PFS_ERR_ARG, "batch_run: bandwidth %d not warmed", B);
    Plan* p = &it->second;
    // no timing events here: the server may be capturing this sequence into a graph
    const bool timing = c->timing;
    c->timing = false;
    const int rc = forward ? dsoft_batch_run(c, p, n, cap, in, out, W, G, st)
                           : idsoft_batch_run(c, p, n, cap, in, out, W, G, st);
    c->timing = timing;
    return rc;
}

}  // namespace pfs

namespace pfs {

int set_error(int code, const char* msg)
{
    return fail(code, "%s", msg);
}

const char* error_text() { return g_last_error.c_str(); }

int dsoft_prepared(pfs_context* c, int B, const double2* grid, double2* coeffs, double2* W, double2* G,
                   hipStream_t st)
{
    std::lock_guard<std::mutex> g(c->mu);
    int rc = check_bandwidth(B);
    if (rc) return rc;
    auto it = c->plans.find(B);  // a lookup only: building a plan allocates
    if (it == c->plans.end())
        return fail(PFS_ERR_ARG, "bandwidth %d: plan not prepared on this context (call pfs_plan_warm first)", B);
    HIP_TRY(hipSetDevice(c->device));
    // the same launches as dsoft_dev, without timing events (creating one allocates)
    const bool timing = c->timing;
    c->timing = false;
    rc = fft_fwd(c, it->second, grid, W, G, st);
    if (!rc) {
        const hipError_t e = launch_dwt_fwd_f64(dwt_tables(c, it->second, true), G, coeffs, fwd_norm(B), st,
                                                (int)c->dwt_variant);
        if (e != hipSuccess) rc = fail(PFS_ERR_DEVICE, "forward DWT launch failed: %s", hipGetErrorString(e));
    }
    c->timing = timing;
    return rc;
}

int context_device(const pfs_context* c) { return c->device; }

bool batch_is_tuned(pfs_context* c, int B, int upto)
{
    std::lock_guard<std::mutex> g(c->mu);
    auto it = c->plans.find(B);
    if (it == c->plans.end()) return false;
    int b = 0;
    while ((1 << (b + 1)) <= upto && b < 7) ++b;  // largest power-of-two bucket <= upto
    return it->second.tuned[0][b].wpb && it->second.tuned[1][b].wpb;
}

}  // namespace pfs

extern "C" {

int pfs_batch_autotune(pfs_context* c, int B, int max_signals)
{
    if (!c || max_signals < 1) return fail(PFS_ERR_ARG, "pfs_batch_autotune: bad argument");
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    const long keep = c->batch_chunk_signals;
    const size_t had = c->bwork_bytes;
    c->batch_chunk_signals = std::max<long>(keep, max_signals);
    const long force = c->batch_force_shape;
    c->batch_force_shape = 0;
    const int rc = batch_autotune(c, B, max_signals);
    c->batch_force_shape = force;
    c->batch_chunk_signals = keep;
    if (c->bwork_bytes > had) {
        // give back what tuning grew; the next batched call sizes them again
        (void)hipStreamSynchronize(c->stream);
        ctx_free(c, c->bwork[0], c->bwork_bytes);
        ctx_free(c, c->bwork[1], c->bwork_bytes);
        c->bwork[0] = c->bwork[1] = nullptr;
        c->bwork_bytes = 0;
    }
    return rc;
}

int pfs_batch_shape(pfs_context* c, int B, int forward, int nsig, int* waves_per_block, int* signals_per_wave)
{
    if (!c || !waves_per_block || !signals_per_wave || nsig < 1) return fail(PFS_ERR_ARG, "pfs_batch_shape: bad argument");
    std::lock_guard<std::mutex> g(c->mu);
    auto it = c->plans.find(B);
    BatchKernelShape sh{0, 0};
    if (it != c->plans.end()) sh = shape_for(c, it->second, forward != 0, nsig);
    if (!sh.wpb) sh = batch_default_shape(2 * B, forward != 0, nsig);
    *waves_per_block = sh.wpb;
    *signals_per_wave = sh.spw;
    return PFS_OK;
}

}  // extern "C"

