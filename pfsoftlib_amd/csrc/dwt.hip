// Discrete Wigner transform stage, FP64, Wigner-d generated on the fly.
//
// One workgroup per canonical pair (M, M'), M >= M' >= 0, with one wave per
// 64 beta nodes.  Every lane runs the three-term recurrence of the
// L2-normalised Wigner d-function for its own node k in FP64 registers
// (fn_dwt.hpp:152-199: base row, then rows l = M+1 .. B-1), 16 degrees at a
// time, into a per-wave LDS tile.  The tile feeds v_mfma_f64_16x16x4_f64
// against the 16 real right-hand sides of the pair: the 8 symmetric cases
// (l, +-M, +-M') / (l, +-M', +-M) of the reference (src/fn_dsoft.cpp:112-283,
// src/fn_idsoft.cpp:100-283) share one table D+ and differ only in which
// (p, q) beta-vector they read, whether that vector is read reversed (the
// d(pi - beta) cases) and a sign (+1, (-1)^(M-M'), (-1)^(l-M) or both):
//
//   case  coefficient   pencil (p, q)   sign           beta reversed
//   0     ( M ,  M')    ( M',  M )      +1             no
//   1     ( M',  M )    ( M ,  M')      e              no
//   2     (-M , -M')    (-M', -M )      e              no
//   3     (-M', -M )    (-M , -M')      +1             no
//   4     ( M', -M )    (-M ,  M')      s_l            yes
//   5     ( M , -M')    (-M',  M )      s_l            yes
//   6     (-M',  M )    ( M , -M')      e s_l          yes
//   7     (-M ,  M')    ( M', -M )      e s_l          yes
//   e = (-1)^(M-M'), s_l = (-1)^(l-M); M' = 0 uses cases 0-3, M' = M uses
//   0, 2, 4, 6 and (0, 0) case 0 only.
//
// The reference builds the table with base sign -1 and undoes it per case
// (fn_dwt.hpp:124-125); here D+ carries the positive base and the table
// above already holds the net sign.  The forward weights w_k multiply the
// input vectors instead of the table (w is symmetric, so reversed cases see
// the same weight).
#include "pfs_internal.h"

namespace pfs {

constexpr int TSTR = 68;          // LDS tile row stride (doubles): 64 nodes + 4 pad
constexpr int OSTR = 18;          // output staging row stride (doubles)
constexpr int WAVE_LDS = 1152;    // doubles per wave: max(16*TSTR, 64*OSTR)

struct CaseInfo {
    long vec;     // pencil offset (complex elements) of the (p, q) beta-vector
    int Mc, Mpc;  // coefficient orders
    int sgn;      // +1 / -1 before the row sign
    int flip;     // reversed beta and row sign s_l
    int active;
};

__device__ __forceinline__ CaseInfo case_info(int c, int M, int Mp, int N)
{
    CaseInfo ci;
    const int eps = ((M - Mp) & 1) ? -1 : 1;
    int row, col;
    switch (c) {
    case 0: ci.Mc = M;   ci.Mpc = Mp;  row = Mp;  col = M;   ci.sgn = 1;   ci.flip = 0; break;
    case 1: ci.Mc = Mp;  ci.Mpc = M;   row = M;   col = Mp;  ci.sgn = eps; ci.flip = 0; break;
    case 2: ci.Mc = -M;  ci.Mpc = -Mp; row = -Mp; col = -M;  ci.sgn = eps; ci.flip = 0; break;
    case 3: ci.Mc = -Mp; ci.Mpc = -M;  row = -M;  col = -Mp; ci.sgn = 1;   ci.flip = 0; break;
    case 4: ci.Mc = Mp;  ci.Mpc = -M;  row = -M;  col = Mp;  ci.sgn = 1;   ci.flip = 1; break;
    case 5: ci.Mc = M;   ci.Mpc = -Mp; row = -Mp; col = M;   ci.sgn = 1;   ci.flip = 1; break;
    case 6: ci.Mc = -Mp; ci.Mpc = M;   row = M;   col = -Mp; ci.sgn = eps; ci.flip = 1; break;
    default: ci.Mc = -M; ci.Mpc = Mp;  row = Mp;  col = -M;  ci.sgn = eps; ci.flip = 1; break;
    }
    if (M == 0) ci.active = (c == 0);
    else if (Mp == 0) ci.active = (c < 4);
    else if (Mp == M) ci.active = !(c & 1);
    else ci.active = 1;
    int p = row < 0 ? row + N : row;
    int q = col < 0 ? col + N : col;
    ci.vec = ((long)p * N + q) * N;
    return ci;
}

__device__ __forceinline__ double ipow(double x, int e)
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
// (fn_dwt.hpp:167-192), FP64.  Stored as (f1, f1*f2, c1); zero past the
// last row so the padded rows stay 0.
__device__ __forceinline__ void recurrence_coeffs(int M, int Mp, int r, int R, double* cf, int stride)
{
    double f1 = 0, f1f2 = 0, c1 = 0;
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
    cf[r] = f1;
    cf[stride + r] = f1f2;
    cf[2 * stride + r] = c1;
}

// MFMA f64 16x16x4 output layout (gfx950, checked by the context's
// self-test): lane L holds D[4v + L/16][L%16] in accumulator register v < 4.
__device__ __forceinline__ int mfma_out_row(int lane, int v) { return 4 * v + (lane >> 4); }
__device__ __forceinline__ int mfma_out_col(int lane) { return lane & 15; }

__global__ __launch_bounds__(1024) void k_dwt_fwd_f64(PlanTables t, const double2* __restrict__ G,
                                                      double2* __restrict__ coeffs, double norm)
{
    extern __shared__ __align__(16) double sm[];
    const int B = t.B, N = t.N;
    const int2 pr = t.pairs[blockIdx.x];
    const int M = pr.x, Mp = pr.y;
    const int R = B - M;
    const int nch = (R + 15) >> 4;
    const int RP = nch * 16;
    const int nthr = blockDim.x, waves = nthr >> 6;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;

    double* tile = sm + w * WAVE_LDS;
    double* red = sm + waves * WAVE_LDS;       // waves * 256
    double* cf = red + waves * 256;            // 3 * RP

    for (int r = tid; r < RP; r += nthr) recurrence_coeffs(M, Mp, r, R, cf, RP);

    // this lane's beta node for the recurrence
    const int k = 64 * w + lane;
    const bool kv = k < N;
    const double cb = kv ? t.cosb[k] : 0.0;
    double d_cur = kv ? t.normfac[M * B + Mp] * ipow(t.sinh[k], M - Mp) * ipow(t.cosh[k], M + Mp) : 0.0;
    double d_prev = 0.0;

    // B operand: lane (j + 16 kk) holds S[64w + 4m + kk][j], j = 2 case + re/im
    const int j = lane & 15, kk = lane >> 4;
    const CaseInfo ci = case_info(j >> 1, M, Mp, N);
    const double* vsrc = reinterpret_cast<const double*>(G + ci.vec) + (j & 1);
    double sreg[16];
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        int ks = 64 * w + 4 * m + kk;
        double v = 0.0;
        if (ci.active && ks < N) {
            int kr = ci.flip ? N - 1 - ks : ks;
            v = vsrc[2 * kr] * t.wts[ks];
        }
        sreg[m] = v;
    }
    __syncthreads();

    for (int ch = 0; ch < nch; ++ch) {
#pragma unroll
        for (int rr = 0; rr < 16; ++rr) {
            int r = ch * 16 + rr;
            tile[rr * TSTR + lane] = d_cur;
            double u = fma(cf[r], cb, cf[RP + r]);
            double dn = fma(d_cur, u, cf[2 * RP + r] * d_prev);
            d_prev = d_cur;
            d_cur = dn;
        }
        __syncthreads();
        v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            double a = tile[(lane & 15) * TSTR + 4 * m + (lane >> 4)];
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, sreg[m], acc, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) red[w * 256 + mfma_out_row(lane, v) * 16 + mfma_out_col(lane)] = acc[v];
        __syncthreads();
        for (int e = tid; e < 256; e += nthr) {
            double s = 0.0;
            for (int ww = 0; ww < waves; ++ww) s += red[ww * 256 + e];
            int i = e >> 4, jj = e & 15;
            int r = ch * 16 + i;
            CaseInfo co = case_info(jj >> 1, M, Mp, N);
            if (r < R && co.active) {
                int l = M + r;
                double sg = (co.flip && (r & 1)) ? -co.sgn : co.sgn;
                reinterpret_cast<double*>(coeffs + coeff_index(l, co.Mc, co.Mpc))[jj & 1] = norm * (sg * s);
            }
        }
        // the next chunk writes tile (read above before the barrier) and red
        // (read by this loop, which every thread finishes before the next
        // chunk's first barrier)
    }
}

__global__ __launch_bounds__(1024) void k_dwt_inv_f64(PlanTables t, const double2* __restrict__ coeffs,
                                                      double2* __restrict__ G, double norm)
{
    extern __shared__ __align__(16) double sm[];
    const int B = t.B, N = t.N;
    const int2 pr = t.pairs[blockIdx.x];
    const int M = pr.x, Mp = pr.y;
    const int R = B - M;
    const int nch = (R + 15) >> 4;
    const int RP = nch * 16;
    const int nthr = blockDim.x, waves = nthr >> 6;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;

    double* tile = sm + w * WAVE_LDS;
    double* cv = sm + waves * WAVE_LDS;   // RP * 16
    double* cf = cv + RP * 16;            // 3 * RP

    for (int r = tid; r < RP; r += nthr) recurrence_coeffs(M, Mp, r, R, cf, RP);
    // coefficient right-hand sides, with norm and the per-case signs applied
    for (int e = tid; e < RP * 16; e += nthr) {
        int r = e >> 4, jj = e & 15;
        CaseInfo co = case_info(jj >> 1, M, Mp, N);
        double v = 0.0;
        if (r < R && co.active) {
            double sg = (co.flip && (r & 1)) ? -co.sgn : co.sgn;
            v = sg * (norm * reinterpret_cast<const double*>(coeffs + coeff_index(M + r, co.Mc, co.Mpc))[jj & 1]);
        }
        cv[e] = v;
    }

    const int k = 64 * w + lane;
    const bool kv = k < N;
    const double cb = kv ? t.cosb[k] : 0.0;
    double d_cur = kv ? t.normfac[M * B + Mp] * ipow(t.sinh[k], M - Mp) * ipow(t.cosh[k], M + Mp) : 0.0;
    double d_prev = 0.0;

    v4d acc[4];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) acc[kb] = (v4d){0.0, 0.0, 0.0, 0.0};
    __syncthreads();

    for (int ch = 0; ch < nch; ++ch) {
#pragma unroll
        for (int rr = 0; rr < 16; ++rr) {
            int r = ch * 16 + rr;
            tile[rr * TSTR + lane] = d_cur;
            double u = fma(cf[r], cb, cf[RP + r]);
            double dn = fma(d_cur, u, cf[2 * RP + r] * d_prev);
            d_prev = d_cur;
            d_cur = dn;
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            double b = cv[(ch * 16 + 4 * s + (lane >> 4)) * 16 + (lane & 15)];
#pragma unroll
            for (int kb = 0; kb < 4; ++kb) {
                double a = tile[(4 * s + (lane >> 4)) * TSTR + 16 * kb + (lane & 15)];
                acc[kb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[kb], 0, 0, 0);
            }
        }
        __syncthreads();
    }

    // stage this wave's 64 x 16 block and write each case's beta-vector run
#pragma unroll
    for (int kb = 0; kb < 4; ++kb)
#pragma unroll
        for (int v = 0; v < 4; ++v)
            tile[(16 * kb + mfma_out_row(lane, v)) * OSTR + mfma_out_col(lane)] = acc[kb][v];
    __syncthreads();
    if (kv) {
        for (int c = 0; c < 8; ++c) {
            CaseInfo co = case_info(c, M, Mp, N);
            if (!co.active) continue;
            double2 val;
            val.x = tile[lane * OSTR + 2 * c];
            val.y = tile[lane * OSTR + 2 * c + 1];
            int kr = co.flip ? N - 1 - k : k;
            G[co.vec + kr] = val;
        }
    }
}

static int dwt_waves(int N) { return (N + 63) / 64; }

hipError_t launch_dwt_fwd_f64(const PlanTables& t, const double2* G, double2* coeffs, double norm, hipStream_t st)
{
    int waves = dwt_waves(t.N);
    if (waves > 16) return hipErrorInvalidValue;
    int RP = ((t.B + 15) / 16) * 16;
    size_t smem = sizeof(double) * ((size_t)waves * WAVE_LDS + waves * 256 + 3 * RP);
    if (smem > lds_limit()) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(k_dwt_fwd_f64, dim3(t.npairs), dim3(64 * waves), smem, st, t, G, coeffs, norm);
    return hipGetLastError();
}

hipError_t launch_dwt_inv_f64(const PlanTables& t, const double2* coeffs, double2* G, double norm, hipStream_t st)
{
    int waves = dwt_waves(t.N);
    if (waves > 16) return hipErrorInvalidValue;
    int RP = ((t.B + 15) / 16) * 16;
    size_t smem = sizeof(double) * ((size_t)waves * WAVE_LDS + (size_t)RP * 16 + 3 * RP);
    if (smem > lds_limit()) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(k_dwt_inv_f64, dim3(t.npairs), dim3(64 * waves), smem, st, t, coeffs, G, norm);
    return hipGetLastError();
}

// G[p][q][k] <- grid(row p, col q, slice k) for p == B or q == B.
__global__ void k_nyquist_gather_f64(const double2* __restrict__ grid, double2* __restrict__ G, int B)
{
    const int N = 2 * B;
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long total = (long)(2 * N - 1) * N;
    if (idx >= total) return;
    int e = (int)(idx / N), k = (int)(idx % N);
    int p, q;
    if (e < N) { p = B; q = e; }
    else { q = B; p = e - N; if (p >= B) ++p; }
    G[((long)p * N + q) * N + k] = grid[(long)N * N * k + (long)N * q + p];
}

hipError_t launch_nyquist_gather_f64(const double2* grid, double2* G, int B, hipStream_t st)
{
    const int N = 2 * B;
    long total = (long)(2 * N - 1) * N;
    hipLaunchKernelGGL(k_nyquist_gather_f64, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, grid, G, B);
    return hipGetLastError();
}

}  // namespace pfs
