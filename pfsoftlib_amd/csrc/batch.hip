// Batched mixed-precision transforms: many complex64 signals of one bandwidth.
//
// Precision: the Wigner-d tables come from the FP64 three-term recurrence
// (as in dwt.hip) and are stored once per bandwidth in FP32; the quadrature
// weights and FFT twiddles are FP64-computed and rounded to FP32; the slice
// FFTs and the DWT contraction run in FP32 (v_mfma_f32_16x16x4_f32, FP32
// accumulation).
//
// DWT mapping: one workgroup per (canonical pair, group of signals).  For
// every 16-degree chunk the pair's FP32 table rows are staged once in LDS
// and shared by the workgroup's waves; each wave contracts them against its
// own signals, whose 8 case beta-vectors (forward) it keeps staged in LDS.
// The 8 symmetric cases, reversed-beta reads and signs follow the table in
// dwt.hip (reference: src/fn_dsoft.cpp:112-283, src/fn_idsoft.cpp:100-283).
#include "pfs_internal.h"

namespace pfs {

int g_mfma_f32_layout = 0;

namespace {

struct BCase {
    long vec;
    int Mc, Mpc, sgn, flip, active;
};

__device__ __forceinline__ BCase bcase(int c, int M, int Mp, int N)
{
    BCase ci;
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
    const int p = row < 0 ? row + N : row;
    const int q = col < 0 ? col + N : col;
    ci.vec = ((long)p * N + q) * N;
    return ci;
}

template <int LAYOUT>
__device__ __forceinline__ int out_row(int lane, int v)
{
    return LAYOUT == 0 ? 4 * v + (lane >> 4) : 4 * (lane >> 4) + v;
}

// FP32 table: for pair index pi, rows r < RP(pi) (16-padded, zero beyond R)
// at D[(roff[pi] + r) * N + k], unweighted, positive base.  FP64 recurrence.
__global__ void k_wigner_table_f32(PlanTables t, const long* __restrict__ roff, float* __restrict__ D)
{
    const int N = t.N, B = t.B;
    const int2 pr = t.pairs[blockIdx.x];
    const int M = pr.x, Mp = pr.y, R = B - M, RP = ((R + 15) >> 4) << 4;
    float* out = D + roff[blockIdx.x] * N;
    const double nf = t.normfac[M * B + Mp];
    for (int k = threadIdx.x; k < N; k += blockDim.x) {
        double d = nf, dp = 0.0;
        const double sh = t.sinh[k], ch = t.cosh[k], cb = t.cosb[k];
        for (int e = 0; e < M - Mp; ++e) d *= sh;
        for (int e = 0; e < M + Mp; ++e) d *= ch;
        for (int r = 0; r < RP; ++r) {
            out[(long)r * N + k] = r < R ? (float)d : 0.0f;
            if (r >= R - 1) continue;
            const double idx = M + r;
            const double nrm = sqrt((2.0 * idx + 3.0) / (2.0 * idx + 1.0));
            const double nom = (idx + 1.0) * (2.0 * idx + 1.0);
            const double den =
                1.0 / sqrt(((idx + 1) * (idx + 1) - (double)M * M) * ((idx + 1) * (idx + 1) - (double)Mp * Mp));
            const double f1 = nrm * nom * den;
            double f2 = 0.0, c1 = 0.0;
            if (M + r != 0) {
                const double t1 = sqrt((2.0 * idx + 3.0) / (2.0 * idx - 1.0)) * (idx + 1.0) / idx;
                const double t2 = sqrt((idx * idx - (double)M * M) * (idx * idx - (double)Mp * Mp));
                c1 = -t1 * t2 * den;
                f2 = -(double)(M * Mp) / (idx * (idx + 1.0));
            }
            const double dn = c1 * dp + d * f1 * (f2 + cb);
            dp = d;
            d = dn;
        }
    }
}

constexpr int DPAD = 4;  // LDS row padding of the staged table chunk (floats)

// Stage degrees 16 ch .. 16 ch + 15 of the pair's table into LDS.
template <int NN, int NT>
__device__ __forceinline__ void stage_table(float* Dc, const float* __restrict__ Dpair, int ch)
{
    const float* src = Dpair + (long)ch * 16 * NN;
    for (int e = threadIdx.x; e < 16 * NN; e += NT) {
        const int r = e / NN, k = e - r * NN;
        Dc[r * (NN + DPAD) + k] = src[e];
    }
}

template <int NN, int WPB, int SPW, int LAYOUT>
__global__ __launch_bounds__(64 * WPB) void k_dwt_fwd_batch_f32(PlanTables t, const long* __restrict__ roff,
                                                                const float* __restrict__ D,
                                                                const float2* __restrict__ G, float2* __restrict__ coeffs,
                                                                int nsig, long count, float norm)
{
    constexpr int NT = 64 * WPB;
    extern __shared__ __align__(16) float smf[];
    float* Dc = smf;                                                  // 16 x (NN + DPAD)
    float2* St = reinterpret_cast<float2*>(smf + 16 * (NN + DPAD));   // [WPB][SPW][8][NN]
    const int B = t.B;
    const long N3 = (long)NN * NN * NN;
    const int2 pr = t.pairs[blockIdx.y];
    const int M = pr.x, Mp = pr.y, R = B - M, nch = (R + 15) >> 4;
    const float* Dpair = D + roff[blockIdx.y] * NN;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int s0 = (blockIdx.x * WPB + w) * SPW;
    float2* Sw = St + (long)w * SPW * 8 * NN;

    // stage this wave's signals: weighted, reversed for the d(pi - beta) cases
    for (int sp = 0; sp < SPW; ++sp) {
        const int s = s0 + sp;
        for (int c = 0; c < 8; ++c) {
            const BCase ci = bcase(c, M, Mp, NN);
            for (int k = lane; k < NN; k += 64) {
                float2 v = make_float2(0.0f, 0.0f);
                if (s < nsig && ci.active) {
                    v = G[s * N3 + ci.vec + (ci.flip ? NN - 1 - k : k)];
                    const float wk = (float)t.wts[k];
                    v.x *= wk;
                    v.y *= wk;
                }
                Sw[(sp * 8 + c) * NN + k] = v;
            }
        }
    }
    // per-lane output case (column j = 2 case + re/im)
    const int j = lane & 15, kk = lane >> 4;
    const BCase co = bcase(j >> 1, M, Mp, NN);

    for (int ch = 0; ch < nch; ++ch) {
        __syncthreads();
        stage_table<NN, NT>(Dc, Dpair, ch);
        __syncthreads();
        for (int sp = 0; sp < SPW; ++sp) {
            const int s = s0 + sp;
            if (s >= nsig) break;  // wave-uniform
            const float2* Ss = Sw + (sp * 8 + (j >> 1)) * NN;
            v4f acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 8
            for (int m = 0; m < NN / 4; ++m) {
                const float a = Dc[(lane & 15) * (NN + DPAD) + 4 * m + kk];
                const float2 bv = Ss[4 * m + kk];
                const float b = (j & 1) ? bv.y : bv.x;
                if (m & 1) acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc1, 0, 0, 0);
                else acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc0, 0, 0, 0);
            }
            const v4f acc = acc0 + acc1;
            if (co.active) {
                float* cs = reinterpret_cast<float*>(coeffs + s * count) + (j & 1);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = ch * 16 + out_row<LAYOUT>(lane, v);
                    if (r < R) {
                        const float sg = (float)((co.flip && (r & 1)) ? -co.sgn : co.sgn);
                        cs[2 * coeff_index32(M + r, co.Mc, co.Mpc)] = norm * (sg * acc[v]);
                    }
                }
            }
        }
    }
}

constexpr int OSTRF = 17;  // output staging row stride (floats), NN rows of 16

template <int NN, int WPB, int SPW, int LAYOUT>
__global__ __launch_bounds__(64 * WPB) void k_dwt_inv_batch_f32(PlanTables t, const long* __restrict__ roff,
                                                                const float* __restrict__ D,
                                                                const float2* __restrict__ coeffs, float2* __restrict__ G,
                                                                int nsig, long count, float norm)
{
    constexpr int NT = 64 * WPB;
    constexpr int NKB = NN / 16;
    extern __shared__ __align__(16) float smf[];
    float* Dc = smf;                                      // 16 x (NN + DPAD)
    float* stage = smf + 16 * (NN + DPAD);                // [WPB][NN][OSTRF]
    const int B = t.B;
    const long N3 = (long)NN * NN * NN;
    const int2 pr = t.pairs[blockIdx.y];
    const int M = pr.x, Mp = pr.y, R = B - M, nch = (R + 15) >> 4;
    const float* Dpair = D + roff[blockIdx.y] * NN;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int s0 = (blockIdx.x * WPB + w) * SPW;
    const int j = lane & 15, bsel = lane >> 4;
    const BCase ci = bcase(j >> 1, M, Mp, NN);

    v4f acc[SPW][NKB];
#pragma unroll
    for (int sp = 0; sp < SPW; ++sp)
#pragma unroll
        for (int kb = 0; kb < NKB; ++kb) acc[sp][kb] = (v4f){0.f, 0.f, 0.f, 0.f};

    for (int ch = 0; ch < nch; ++ch) {
        __syncthreads();
        stage_table<NN, NT>(Dc, Dpair, ch);
        __syncthreads();
#pragma unroll
        for (int sp = 0; sp < SPW; ++sp) {
            const int s = s0 + sp;
            if (s >= nsig) break;  // wave-uniform
            float bv[4];
            const float* cs = reinterpret_cast<const float*>(coeffs + s * count) + (j & 1);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = ch * 16 + 4 * q + bsel;
                float v = 0.f;
                if (ci.active && r < R) {
                    const float sg = (float)((ci.flip && (r & 1)) ? -ci.sgn : ci.sgn);
                    v = sg * (norm * cs[2 * coeff_index32(M + r, ci.Mc, ci.Mpc)]);
                }
                bv[q] = v;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int kb = 0; kb < NKB; ++kb) {
                    const float a = Dc[(4 * q + bsel) * (NN + DPAD) + 16 * kb + j];
                    acc[sp][kb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bv[q], acc[sp][kb], 0, 0, 0);
                }
        }
    }

    // write each signal's 8 beta-vectors, staged so every case run is contiguous
    float* sw = stage + (long)w * NN * OSTRF;
#pragma unroll
    for (int sp = 0; sp < SPW; ++sp) {
        const int s = s0 + sp;
        if (s >= nsig) break;
#pragma unroll
        for (int kb = 0; kb < NKB; ++kb)
#pragma unroll
            for (int v = 0; v < 4; ++v) sw[(16 * kb + out_row<LAYOUT>(lane, v)) * OSTRF + j] = acc[sp][kb][v];
        __builtin_amdgcn_s_waitcnt(0xc07f);
        __builtin_amdgcn_wave_barrier();
        for (int c = 0; c < 8; ++c) {
            const BCase co = bcase(c, M, Mp, NN);
            if (!co.active) continue;
            for (int k = lane; k < NN; k += 64)
                G[s * N3 + co.vec + (co.flip ? NN - 1 - k : k)] = make_float2(sw[k * OSTRF + 2 * c], sw[k * OSTRF + 2 * c + 1]);
        }
        __builtin_amdgcn_s_waitcnt(0xc07f);
        __builtin_amdgcn_wave_barrier();
    }
}

// Zero the pencil entries with p == B or q == B for nsig signals.
__global__ void k_nyquist_zero_f32(float2* __restrict__ G, int B, int nsig)
{
    const int N = 2 * B;
    const long per = (long)(2 * N - 1) * N;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= per * nsig) return;
    const long s = idx / per, rem = idx - s * per;
    const int e = (int)(rem / N), k = (int)(rem % N);
    int p, q;
    if (e < N) { p = B; q = e; }
    else { q = B; p = e - N; if (p >= B) ++p; }
    G[s * (long)N * N * N + ((long)p * N + q) * N + k] = make_float2(0.f, 0.f);
}

template <int NN> struct BatchShape;
template <> struct BatchShape<16> { static constexpr int WPB = 4, SPW = 4; };
template <> struct BatchShape<32> { static constexpr int WPB = 4, SPW = 4; };
template <> struct BatchShape<64> { static constexpr int WPB = 4, SPW = 2; };
template <> struct BatchShape<128> { static constexpr int WPB = 4, SPW = 2; };
template <> struct BatchShape<256> { static constexpr int WPB = 2, SPW = 1; };

template <int NN>
hipError_t fwd_nn(const PlanTables& t, const long* roff, const float* D, const float2* G, float2* coeffs, int nsig,
                  long count, float norm, hipStream_t st)
{
    constexpr int WPB = BatchShape<NN>::WPB, SPW = BatchShape<NN>::SPW;
    const size_t smem = sizeof(float) * 16 * (NN + DPAD) + sizeof(float2) * (size_t)WPB * SPW * 8 * NN;
    if (smem > lds_limit()) return hipErrorInvalidConfiguration;
    dim3 grid((nsig + WPB * SPW - 1) / (WPB * SPW), t.npairs), block(64 * WPB);
    if (g_mfma_f32_layout == 0)
        hipLaunchKernelGGL((k_dwt_fwd_batch_f32<NN, WPB, SPW, 0>), grid, block, smem, st, t, roff, D, G, coeffs, nsig, count, norm);
    else
        hipLaunchKernelGGL((k_dwt_fwd_batch_f32<NN, WPB, SPW, 1>), grid, block, smem, st, t, roff, D, G, coeffs, nsig, count, norm);
    return hipGetLastError();
}

template <int NN>
hipError_t inv_nn(const PlanTables& t, const long* roff, const float* D, const float2* coeffs, float2* G, int nsig,
                  long count, float norm, hipStream_t st)
{
    constexpr int WPB = BatchShape<NN>::WPB, SPW = BatchShape<NN>::SPW;
    const size_t smem = sizeof(float) * (16 * (NN + DPAD) + (size_t)WPB * NN * OSTRF);
    if (smem > lds_limit()) return hipErrorInvalidConfiguration;
    dim3 grid((nsig + WPB * SPW - 1) / (WPB * SPW), t.npairs), block(64 * WPB);
    if (g_mfma_f32_layout == 0)
        hipLaunchKernelGGL((k_dwt_inv_batch_f32<NN, WPB, SPW, 0>), grid, block, smem, st, t, roff, D, coeffs, G, nsig, count, norm);
    else
        hipLaunchKernelGGL((k_dwt_inv_batch_f32<NN, WPB, SPW, 1>), grid, block, smem, st, t, roff, D, coeffs, G, nsig, count, norm);
    return hipGetLastError();
}

}  // namespace

bool batch_supported(int N) { return N == 16 || N == 32 || N == 64 || N == 128 || N == 256; }

hipError_t launch_wigner_table_f32(const PlanTables& t, const long* roff, float* D, hipStream_t st)
{
    hipLaunchKernelGGL(k_wigner_table_f32, dim3(t.npairs), dim3(t.N < 256 ? t.N : 256), 0, st, t, roff, D);
    return hipGetLastError();
}

hipError_t launch_dwt_fwd_batch_f32(const PlanTables& t, const long* roff, const float* D, const float2* G,
                                    float2* coeffs, int nsig, float norm, hipStream_t st)
{
    const long count = coeff_offset(t.B);
    switch (t.N) {
    case 16: return fwd_nn<16>(t, roff, D, G, coeffs, nsig, count, norm, st);
    case 32: return fwd_nn<32>(t, roff, D, G, coeffs, nsig, count, norm, st);
    case 64: return fwd_nn<64>(t, roff, D, G, coeffs, nsig, count, norm, st);
    case 128: return fwd_nn<128>(t, roff, D, G, coeffs, nsig, count, norm, st);
    case 256: return fwd_nn<256>(t, roff, D, G, coeffs, nsig, count, norm, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_dwt_inv_batch_f32(const PlanTables& t, const long* roff, const float* D, const float2* coeffs,
                                    float2* G, int nsig, float norm, hipStream_t st)
{
    const long count = coeff_offset(t.B);
    switch (t.N) {
    case 16: return inv_nn<16>(t, roff, D, coeffs, G, nsig, count, norm, st);
    case 32: return inv_nn<32>(t, roff, D, coeffs, G, nsig, count, norm, st);
    case 64: return inv_nn<64>(t, roff, D, coeffs, G, nsig, count, norm, st);
    case 128: return inv_nn<128>(t, roff, D, coeffs, G, nsig, count, norm, st);
    case 256: return inv_nn<256>(t, roff, D, coeffs, G, nsig, count, norm, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_nyquist_zero_f32(float2* G, int B, int nsig, hipStream_t st)
{
    const int N = 2 * B;
    const long total = (long)(2 * N - 1) * N * nsig;
    hipLaunchKernelGGL(k_nyquist_zero_f32, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, G, B, nsig);
    return hipGetLastError();
}

}  // namespace pfs
