// Batched mixed-precision transforms: many complex64 signals of one bandwidth.
//
// Precision: the Wigner-d tables come from the FP64 three-term recurrence
// (as in dwt.hip) and are stored once per bandwidth in FP32, the forward's
// multiplied by the FP64 quadrature weights before rounding; the FFT
// twiddles are FP64-computed and rounded to FP32; the slice
// FFTs and the DWT contraction run in FP32 (v_mfma_f32_16x16x4_f32, FP32
// accumulation).
//
// DWT mapping: one workgroup per (canonical pair, group of signals); each
// wave owns SPW signals and computes, per signal, the pair's whole
// (degrees x 16) forward product or (betas x 16) inverse product in MFMA
// accumulators -- 16 columns = 8 symmetric cases x (re, im).  Table operands
// come from 16 x 16 tiles laid out so that every lane loads a float4 and a
// wave loads 1 KB contiguous; the workgroup's waves and the pair's other
// signal groups (same XCD) reuse them from cache.  Forward signal operands
// are read straight from the pencils, no LDS.  The 8 symmetric cases,
// reversed-beta reads and signs follow the table in dwt.hip (reference:
// src/fn_dsoft.cpp:112-283, src/fn_idsoft.cpp:100-283).
#include "pfs_internal.h"

#include <algorithm>

namespace pfs {

typedef float v2f __attribute__((ext_vector_type(2)));
typedef short v4s __attribute__((ext_vector_type(4)));

int g_mfma_f32_layout = 0;
bool g_mfma_bf16_ok = false;

// ---- split-BF16 products ------------------------------------------------------
// A value x is split into bf16 terms h + m (+ l): h = bf16_rne(x), m =
// bf16_rne(x - h), l = bf16_rne(x - h - m); each remainder is exact in FP32,
// and each term keeps 8 more significant bits.  PREC 3 sums the products
// h h' + h m' + m h' (the terms of order 2^-16 dropped), PREC 6 adds
// h l' + m m' + l h' (order 2^-24 dropped), on v_mfma_f32_16x16x16_bf16 with
// FP32 accumulation.  PREC 0 is the plain FP32 MFMA path.

// bf16 bit pattern of round-to-nearest-even(x), in the high 16 bits (finite x).
__device__ __forceinline__ unsigned bf16_rne_bits(float x)
{
    const unsigned u = __float_as_uint(x);
    return (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;
}

// Bf16 terms of four values, packed as the K-ordered operand of one
// 16x16x16 bf16 MFMA: element u of part p is term p of x[u].
template <int PARTS>
__device__ __forceinline__ void split4(const float (&x)[4], v4s (&out)[PARTS])
{
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        float r = x[u];
#pragma unroll
        for (int pp = 0; pp < PARTS; ++pp) {
            const unsigned hb = bf16_rne_bits(r);
            out[pp][u] = (short)(hb >> 16);
            r -= __uint_as_float(hb);
        }
    }
}

// acc += A B over the split terms (A: ap parts, B: bp parts), smallest first.
template <int PREC>
__device__ __forceinline__ v4f mfma_split(const v4s* ap, const v4s* bp, v4f acc)
{
    if constexpr (PREC == 6) {
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ap[0], bp[2], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ap[1], bp[1], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ap[2], bp[0], acc, 0, 0, 0);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ap[0], bp[1], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ap[1], bp[0], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ap[0], bp[0], acc, 0, 0, 0);
    return acc;
}

constexpr int split_parts(int prec) { return prec == 6 ? 3 : 2; }

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

// Output row of accumulator register v in lane `lane` for the FP32 MFMA
// output layout the context's self-test found (0 or 1).
__device__ __forceinline__ int out_row(int lane, int v, int layout)
{
    return layout == 0 ? 4 * v + (lane >> 4) : 4 * (lane >> 4) + v;
}

// FP32 tables for pair index pi, rows r < RP(pi) (16-padded, zero beyond R),
// positive base, from the FP64 recurrence: Tf holds d * w_k (the forward's
// weighted table, the product rounded once), Ti holds d.  Both tables are cut
// into 16 x 16 tiles; tile (q, kb) of pair pi (rows 16q.., betas 16kb..) sits
// at tile index (roff[pi]/16 + q) * N/16 + kb.  Tf stores a tile row-major
// (r, k) -> 16 r + k, Ti column-major (r, k) -> 16 k + r, so that a lane of
// either contraction reads its four operands as one float4.
__global__ void k_wigner_tiles_f32(PlanTables t, const long* __restrict__ roff, float* __restrict__ Tf,
                                   float* __restrict__ Ti)
{
    const int N = t.N, B = t.B, NKB = N >> 4;
    const int2 pr = t.pairs[blockIdx.x];
    const int M = pr.x, Mp = pr.y, R = B - M, RP = ((R + 15) >> 4) << 4;
    const long tile0 = (roff[blockIdx.x] >> 4) * NKB;
    const double nf = t.normfac[M * B + Mp];
    for (int k = threadIdx.x; k < N; k += blockDim.x) {
        double d = nf, dp = 0.0;
        const double sh = t.sinh[k], ch = t.cosh[k], cb = t.cosb[k];
        for (int e = 0; e < M - Mp; ++e) d *= sh;
        for (int e = 0; e < M + Mp; ++e) d *= ch;
        const double wk = t.wts[k];
        for (int r = 0; r < RP; ++r) {
            const long tb = (tile0 + (long)(r >> 4) * NKB + (k >> 4)) * 256;
            Tf[tb + 16 * (r & 15) + (k & 15)] = r < R ? (float)(d * wk) : 0.0f;
            Ti[tb + 16 * (k & 15) + (r & 15)] = r < R ? (float)d : 0.0f;
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

// Block -> (pair, signal group).  Blocks are dealt to the 8 XCDs round-robin,
// so the signal groups of one pair are given consecutive slots of one XCD and
// share its L2 copy of the pair's table tiles.
struct BlockWork {
    int pi, sg;
};

__device__ __forceinline__ BlockWork block_work(int nsg)
{
    const int b = blockIdx.x, x = b & 7, tb = b >> 3;
    BlockWork bw;
    bw.sg = tb % nsg;
    bw.pi = 8 * (tb / nsg) + x;
    return bw;
}

// Forward contraction of one canonical pair for SPW signals per wave:
// C[r][j] = sum_k (D[r][k] w_k) * G_s[pencil(j/2)][k or N-1-k].(re|im by j&1),
// as v_mfma_f32_16x16x4f32 over NC 16-degree chunks (A = weighted table tile
// rows, read from the row-major tiles) and 16-beta blocks (B = the signal's 8 case
// vectors, read straight from the pencils).  Within a 16-beta block, MFMA
// step t contracts betas 16 kb + 4 kk + t over the lane groups kk.  The loads
// of block kb + 1 are issued before the MFMAs of block kb (two register
// stages), so memory latency overlaps the contraction.
template <int NN, int SPW, int NC, int PREC = 0>
struct FwdStage {
    float4 a[PREC ? 1 : NC];                      // FP32 table operands (weighted)
    v4s as[PREC ? NC : 1][split_parts(PREC)];     // split table operands
    float4 g[SPW];  // two consecutive complex values of this lane's case pencil
};

// Exchange with the partner lane (lane ^ 1) through DPP quad_perm [1,0,3,2].
__device__ __forceinline__ float swap_pair(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
}

// Lane pair (2c, 2c+1) of lane group kk covers the four betas 16 kb + 4 kk + t
// (t = 0..3) of case c: for the direct cases lane 2c+p loads the complex pair
// at 16 kb + 4 kk + 2p, for the reversed cases the pair at N - 4 - 16 kb - 4 kk
// + 2(1-p) -- one 16-byte load per lane, 1 KB contiguous per wave (8 pencils
// x 128 bytes).  The partners then swap halves so that the even lane holds the
// real parts and the odd lane the imaginary parts of the four betas, in t order.
// Table tile parts of the split precisions: same element order as the FP32
// tiles, one array of bf16 bit patterns per term.
struct SplitTab {
    const unsigned short* p[3];
};

template <int NN, int SPW, int NC, int PREC>
__device__ __forceinline__ void fwd_load(FwdStage<NN, SPW, NC, PREC>& S, const float* const (&gp)[SPW], const float* tf,
                                         const SplitTab& ts, long toff, int e0, int de, int kb)
{
    constexpr int NKB = NN / 16;
    if constexpr (PREC == 0) {
#pragma unroll
        for (int ch = 0; ch < NC; ++ch) S.a[ch] = *reinterpret_cast<const float4*>(tf + (ch * NKB + kb) * 256);
    } else {
#pragma unroll
        for (int ch = 0; ch < NC; ++ch)
#pragma unroll
            for (int pp = 0; pp < split_parts(PREC); ++pp)
                S.as[ch][pp] = *reinterpret_cast<const v4s*>(ts.p[pp] + toff + (ch * NKB + kb) * 256);
    }
    const int e = e0 + kb * de;
    // each pencil value is read once: a non-temporal load does not allocate, so
    // the streamed pencils (8.6 GB for 64 grids at 2B = 256) do not evict the
    // table tiles the other signal groups of the pair read from L2
#pragma unroll
    for (int sp = 0; sp < SPW; ++sp) {
        const v4f g = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(gp[sp] + e));
        S.g[sp] = make_float4(g.x, g.y, g.z, g.w);
    }
}

template <int NN, int SPW, int NC, int PREC>
__device__ __forceinline__ void fwd_mfma(v4f (&acc)[NC][SPW], const FwdStage<NN, SPW, NC, PREC>& S, bool odd, bool flip)
{
    float b[SPW][4];
#pragma unroll
    for (int sp = 0; sp < SPW; ++sp) {
        const float4 o = S.g[sp];
        float4 x;
        x.x = swap_pair(o.x);
        x.y = swap_pair(o.y);
        x.z = swap_pair(o.z);
        x.w = swap_pair(o.w);
        // direct:   even (o.x, o.z, x.x, x.z), odd (x.y, x.w, o.y, o.w)
        // reversed: even (o.z, o.x, x.z, x.x), odd (x.w, x.y, o.w, o.y)
        const float d0 = odd ? x.y : o.x, d1 = odd ? x.w : o.z, d2 = odd ? o.y : x.x, d3 = odd ? o.w : x.z;
        b[sp][0] = flip ? d1 : d0;
        b[sp][1] = flip ? d0 : d1;
        b[sp][2] = flip ? d3 : d2;
        b[sp][3] = flip ? d2 : d3;
    }
    if constexpr (PREC == 0) {
#pragma unroll
        for (int ch = 0; ch < NC; ++ch)
#pragma unroll
            for (int sp = 0; sp < SPW; ++sp) {
                acc[ch][sp] = __builtin_amdgcn_mfma_f32_16x16x4f32(S.a[ch].x, b[sp][0], acc[ch][sp], 0, 0, 0);
                acc[ch][sp] = __builtin_amdgcn_mfma_f32_16x16x4f32(S.a[ch].y, b[sp][1], acc[ch][sp], 0, 0, 0);
                acc[ch][sp] = __builtin_amdgcn_mfma_f32_16x16x4f32(S.a[ch].z, b[sp][2], acc[ch][sp], 0, 0, 0);
                acc[ch][sp] = __builtin_amdgcn_mfma_f32_16x16x4f32(S.a[ch].w, b[sp][3], acc[ch][sp], 0, 0, 0);
            }
    } else {
        // the four betas of an FP32 step group are the K = 16 of one bf16 MFMA
        v4s bs[SPW][split_parts(PREC)];
#pragma unroll
        for (int sp = 0; sp < SPW; ++sp) split4<split_parts(PREC)>(b[sp], bs[sp]);
#pragma unroll
        for (int ch = 0; ch < NC; ++ch)
#pragma unroll
            for (int sp = 0; sp < SPW; ++sp) acc[ch][sp] = mfma_split<PREC>(S.as[ch], bs[sp], acc[ch][sp]);
    }
}

template <int NN, int SPW, int NC, int PREC>
__device__ __forceinline__ void fwd_pair(const float* const (&gp)[SPW], const float* tf, const SplitTab& ts, long toff,
                                         int e0, int de, float2* __restrict__ coeffs, const BCase& cb,
                                         int M, int R, int s0, int nsig, long cdist, long cstride, float norm, int lane,
                                         int layout)
{
    constexpr int NKB = NN / 16;
    v4f acc[NC][SPW];
#pragma unroll
    for (int ch = 0; ch < NC; ++ch)
#pragma unroll
        for (int sp = 0; sp < SPW; ++sp) acc[ch][sp] = (v4f){0.f, 0.f, 0.f, 0.f};

    const bool odd = lane & 1, flip = cb.flip;
    // ST register stages: block kb + ST - 1 loads while block kb contracts.  A
    // pair with one or two chunks does only 8 SPW or 16 SPW MFMAs per block,
    // too little to cover a load's latency with one block in flight.
    constexpr int ST0 = NC <= 2 ? 4 : 2;
    constexpr int ST = ST0 < NKB ? ST0 : NKB;
    FwdStage<NN, SPW, NC, PREC> S[ST];
#pragma unroll
    for (int st = 0; st + 1 < ST; ++st) fwd_load(S[st], gp, tf, ts, toff, e0, de, st);
    if constexpr (ST == 1) {
        fwd_load(S[0], gp, tf, ts, toff, e0, de, 0);
        fwd_mfma(acc, S[0], odd, flip);
    } else {
        // The scheduling barriers keep each stage's loads ahead of the other
        // stages' MFMAs (left alone, the scheduler sinks them next to their use).
        for (int kb = 0; kb < NKB; kb += ST) {
#pragma unroll
            for (int st = 0; st < ST; ++st) {
                // unconditional (past the end it re-reads the last block): a
                // branch here would make the compiler drain every outstanding
                // load before the next MFMAs
                const int nk = kb + st + ST - 1;
                fwd_load(S[(st + ST - 1) % ST], gp, tf, ts, toff, e0, de, nk < NKB ? nk : NKB - 1);
                __builtin_amdgcn_sched_barrier(0);
                fwd_mfma(acc, S[st], odd, flip);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    }

    if (!cb.active) return;  // lanes 2c, 2c+1 (case c, re / im) leave together
    const int part = lane & 1;
    constexpr int STEP = SPW % 4 == 0 ? 4 : 2;
    if (SPW % 2 == 0 && cdist == 1 && cstride % STEP == 0) {
        // interleaved layout, row length a multiple of STEP: signals s0 + sp,
        // s0 + sp + 1 of a coefficient are adjacent complex values, written by
        // the re lane as one 16-byte store after taking the im values from its
        // partner (a slot past nsig lies in the row's padding, since s0 + sp
        // < nsig <= cstride and both are multiples of STEP, and is never read
        // back).  With four signals per wave the re lane writes the first two,
        // the im lane the other two, so every lane stores.
#pragma unroll
        for (int sp = 0; sp < SPW; sp += STEP) {
            if (s0 + sp >= nsig) break;
            float* cp = reinterpret_cast<float*>(coeffs + (s0 + sp + (STEP == 4 ? 2 * part : 0)));
#pragma unroll
            for (int ch = 0; ch < NC; ++ch)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = ch * 16 + out_row(lane, v, layout);  // the same in both lanes of a pair
                    if (r < R) {
                        const float sg = (float)((cb.flip && (r & 1)) ? -cb.sgn : cb.sgn);
                        float a[STEP], b[STEP];
#pragma unroll
                        for (int q = 0; q < STEP; ++q) {
                            a[q] = norm * (sg * acc[ch][sp + q][v]);
                            b[q] = swap_pair(a[q]);
                        }
                        float4* dst = reinterpret_cast<float4*>(cp + 2 * cstride * coeff_index32(M + r, cb.Mc, cb.Mpc));
                        if (STEP == 4)
                            *dst = part ? make_float4(b[2], a[2], b[3], a[3]) : make_float4(a[0], b[0], a[1], b[1]);
                        else if (!part)
                            *dst = make_float4(a[0], b[0], a[1], b[1]);
                    }
                }
        }
        return;
    }
#pragma unroll
    for (int sp = 0; sp < SPW; ++sp) {
        if (s0 + sp >= nsig) break;
        float* cs = reinterpret_cast<float*>(coeffs + (long)(s0 + sp) * cdist) + part;
#pragma unroll
        for (int ch = 0; ch < NC; ++ch)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = ch * 16 + out_row(lane, v, layout);
                if (r < R) {
                    const float sg = (float)((cb.flip && (r & 1)) ? -cb.sgn : cb.sgn);
                    cs[2 * cstride * coeff_index32(M + r, cb.Mc, cb.Mpc)] = norm * (sg * acc[ch][sp][v]);
                }
            }
    }
}

// Run fwd_pair with the pair's chunk count as a template argument.
template <int NN, int SPW, int NC, int PREC>
__device__ __forceinline__ void fwd_dispatch(int nch, const float* const (&gp)[SPW], const float* tf, const SplitTab& ts,
                                             long toff, int e0, int de, float2* __restrict__ coeffs,
                                             const BCase& cb, int M, int R, int s0, int nsig, long cdist, long cstride,
                                             float norm, int lane, int layout)
{
    if constexpr (NC <= (NN / 2 + 15) / 16) {
        if (nch == NC)
            fwd_pair<NN, SPW, NC, PREC>(gp, tf, ts, toff, e0, de, coeffs, cb, M, R, s0, nsig, cdist, cstride, norm,
                                        lane, layout);
        else
            fwd_dispatch<NN, SPW, NC + 1, PREC>(nch, gp, tf, ts, toff, e0, de, coeffs, cb, M, R, s0, nsig, cdist,
                                                cstride, norm, lane, layout);
    }
}

// One work item of the forward contraction: pair bw.pi for signal group
// bw.sg (WPB waves x SPW signals).
template <int NN, int WPB, int SPW, int PREC>
__device__ __forceinline__ void fwd_item(const PlanTables& t, const long* __restrict__ roff,
                                         const float* __restrict__ Tf, const SplitTab& ts, const float2* __restrict__ G, float2* __restrict__ coeffs,
                                         int nsig, long cdist, long cstride, float norm, int layout, BlockWork bw)
{
    constexpr int NKB = NN / 16;
    if (bw.pi >= t.npairs) return;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int s0 = (bw.sg * WPB + w) * SPW;
    if (s0 >= nsig) return;
    const int2 pr = t.pairs[bw.pi];
    const int M = pr.x, Mp = pr.y, R = t.B - M, nch = (R + 15) >> 4;
    const int i = lane & 15, kk = lane >> 4;
    const BCase cb = bcase(i >> 1, M, Mp, NN);
    const long N3 = (long)NN * NN * NN;

    const float* gp[SPW];
#pragma unroll
    for (int sp = 0; sp < SPW; ++sp)
        gp[sp] = reinterpret_cast<const float*>(G + (long)min(s0 + sp, nsig - 1) * N3 + cb.vec);
    const long toff = (roff[bw.pi] >> 4) * NKB * 256 + 16 * i + 4 * kk;
    const float* tf = Tf + toff;
    // float offset of this lane's complex pair in beta block 0, and its step per block
    const int p = i & 1;
    const int e0 = 2 * (cb.flip ? NN - 4 - 4 * kk + 2 * (1 - p) : 4 * kk + 2 * p);
    const int de = cb.flip ? -32 : 32;

    fwd_dispatch<NN, SPW, 1, PREC>(nch, gp, tf, ts, toff, e0, de, coeffs, cb, M, R, s0, nsig, cdist, cstride, norm,
                                   lane, layout);
}

template <int NN, int WPB, int SPW, int PREC>
__global__ __launch_bounds__(64 * WPB) void k_dwt_fwd_batch_f32(PlanTables t, const long* __restrict__ roff,
                                                                const float* __restrict__ Tf, SplitTab ts,
                                                                const float2* __restrict__ G, float2* __restrict__ coeffs,
                                                                int nsig, int nsg, long cdist, long cstride, float norm,
                                                                int layout)
{
    fwd_item<NN, WPB, SPW, PREC>(t, roff, Tf, ts, G, coeffs, nsig, cdist, cstride, norm, layout,
                                 block_work(nsg));
}

constexpr int OSTRF = 17;  // output staging row stride (floats), NN rows of 16

// Coefficient operands of one degree chunk as loaded (v) and the per-row
// factor norm * sign (0 past R and for inactive cases) applied at use, so the
// loads can stay in flight across the previous chunk's MFMAs.
template <int SPW>
struct InvStage {
    float v[SPW][4];
    float f[4];
};

// PAIRED (interleaved layout with an even row length, SPW even): signals
// s0 + sp and s0 + sp + 1 are adjacent complex values, so one 16-byte load
// from cp[sp / 2] gives both and the lane keeps its part of each -- half the
// gather instructions, each covering the same cache lines.
template <int SPW, bool PAIRED>
__device__ __forceinline__ void inv_gather(InvStage<SPW>& S, const float* const (&cs)[SPW], const float* const (&cp)[SPW],
                                           const BCase& cb, int M, int R, long cstride, float norm, int kk, int q,
                                           int part)
{
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int r = 16 * q + 4 * kk + u;
        const bool live = cb.active && r < R;
        const long ci = live ? 2 * cstride * coeff_index32(M + r, cb.Mc, cb.Mpc) : 0;
        S.f[u] = live ? norm * (float)((cb.flip && (r & 1)) ? -cb.sgn : cb.sgn) : 0.0f;
        if constexpr (PAIRED) {
#pragma unroll
            for (int sp = 0; sp < SPW; sp += 2) {
                const float4 v = *reinterpret_cast<const float4*>(cp[sp] + ci);
                S.v[sp][u] = part ? v.y : v.x;
                S.v[sp + 1][u] = part ? v.w : v.z;
            }
        } else {
#pragma unroll
            for (int sp = 0; sp < SPW; ++sp) S.v[sp][u] = cs[sp][ci];
        }
    }
}

// Table operands of the inverse: FP32 tiles, or their split parts.
template <int NKB, int PREC>
struct InvTab {
    float4 a[PREC ? 1 : NKB];
    v4s as[PREC ? NKB : 1][split_parts(PREC)];
};

template <int NKB, int PREC>
__device__ __forceinline__ void inv_tab_load(InvTab<NKB, PREC>& T, int kb, const float* tnext, const SplitTab& ts,
                                             long tnoff)
{
    if constexpr (PREC == 0) {
        T.a[kb] = *reinterpret_cast<const float4*>(tnext + kb * 256);
    } else {
#pragma unroll
        for (int pp = 0; pp < split_parts(PREC); ++pp)
            T.as[kb][pp] = *reinterpret_cast<const v4s*>(ts.p[pp] + tnoff + kb * 256);
    }
}

template <int NKB, int SPW, int PREC>
__device__ __forceinline__ void inv_contract(v4f (&acc)[NKB][SPW], const InvStage<SPW>& S, InvTab<NKB, PREC>& T,
                                             const float* tnext, const SplitTab& ts, long tnoff)
{
    float b[SPW][4];
#pragma unroll
    for (int sp = 0; sp < SPW; ++sp)
#pragma unroll
        for (int u = 0; u < 4; ++u) b[sp][u] = S.v[sp][u] * S.f[u];
    if constexpr (PREC != 0) {
        // the four degrees of an FP32 step group are the K = 16 of one bf16 MFMA
        v4s bs[SPW][split_parts(PREC)];
#pragma unroll
        for (int sp = 0; sp < SPW; ++sp) split4<split_parts(PREC)>(b[sp], bs[sp]);
        __builtin_amdgcn_sched_barrier(0);
        constexpr int GS = NKB < 4 ? NKB : 4;
#pragma unroll
        for (int g = 0; g < NKB / GS; ++g) {
#pragma unroll
            for (int k4 = 0; k4 < GS; ++k4)
#pragma unroll
                for (int sp = 0; sp < SPW; ++sp)
                    acc[GS * g + k4][sp] = mfma_split<PREC>(T.as[GS * g + k4], bs[sp], acc[GS * g + k4][sp]);
#pragma unroll
            for (int k4 = 0; k4 < GS; ++k4) inv_tab_load<NKB, PREC>(T, GS * g + k4, tnext, ts, tnoff);
            __builtin_amdgcn_sched_group_barrier(0x008, (PREC == 6 ? 6 : 3) * GS * SPW, 0);  // the group's MFMAs
            __builtin_amdgcn_sched_group_barrier(0x020, split_parts(PREC) * GS, 0);       // then its tile loads
        }
    } else {
    float4 (&a)[NKB] = T.a;
    __builtin_amdgcn_sched_barrier(0);
    // Tiles in groups of four: a group's GS x SPW independent accumulator
    // chains hide the MFMA latency, and its registers take the next chunk's
    // tiles as soon as its MFMAs have read them, a chunk's worth of MFMAs
    // before they are used (measured: 3.92 -> 3.61 ms for 64 signals at
    // B = 128 against loading each chunk's tiles right before its MFMAs).
    constexpr int GS = NKB < 4 ? NKB : 4;
#pragma unroll
    for (int g = 0; g < NKB / GS; ++g) {
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int k4 = 0; k4 < GS; ++k4)
#pragma unroll
                for (int sp = 0; sp < SPW; ++sp) {
                    const int kb = GS * g + k4;
                    const float av = u == 0 ? a[kb].x : u == 1 ? a[kb].y : u == 2 ? a[kb].z : a[kb].w;
                    acc[kb][sp] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, b[sp][u], acc[kb][sp], 0, 0, 0);
                }
#pragma unroll
        for (int k4 = 0; k4 < GS; ++k4)
            a[GS * g + k4] = *reinterpret_cast<const float4*>(tnext + (GS * g + k4) * 256);
        __builtin_amdgcn_sched_group_barrier(0x008, 4 * GS * SPW, 0);  // the group's MFMAs
        __builtin_amdgcn_sched_group_barrier(0x020, GS, 0);            // then its next-chunk tile loads
    }
    }
}

// inv_contract (FP32 tiles) with the chunk's A operands read from the
// workgroup's LDS copy of its tiles (lt: this lane's float4 of tile 0; tile
// kb at lt + 256 kb), one group of four tiles ahead of its MFMAs.
template <int NKB, int SPW>
__device__ __forceinline__ void inv_contract_lds(v4f (&acc)[NKB][SPW], const InvStage<SPW>& S, const float* lt)
{
    float b[SPW][4];
#pragma unroll
    for (int sp = 0; sp < SPW; ++sp)
#pragma unroll
        for (int u = 0; u < 4; ++u) b[sp][u] = S.v[sp][u] * S.f[u];
    constexpr int GS = NKB < 4 ? NKB : 4;
    v4f a[GS], an[GS];
#pragma unroll
    for (int k4 = 0; k4 < GS; ++k4) a[k4] = *reinterpret_cast<const v4f*>(lt + k4 * 256);
#pragma unroll
    for (int g = 0; g < NKB / GS; ++g) {
        if (g + 1 < NKB / GS) {
#pragma unroll
            for (int k4 = 0; k4 < GS; ++k4) an[k4] = *reinterpret_cast<const v4f*>(lt + (GS * (g + 1) + k4) * 256);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int k4 = 0; k4 < GS; ++k4)
#pragma unroll
                for (int sp = 0; sp < SPW; ++sp) {
                    const int kb = GS * g + k4;
                    acc[kb][sp] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[k4][u], b[sp][u], acc[kb][sp], 0, 0, 0);
                }
#pragma unroll
        for (int k4 = 0; k4 < GS; ++k4) a[k4] = an[k4];
    }
}

// Inverse contraction of one canonical pair for SPW signals per wave:
// P[k][j] = sum_r D[r][k] * c_s(M + r, case j/2).(re|im by j&1) * sign * norm,
// A = column-major table tiles (k rows, r columns), B = the coefficients,
// gathered per lane.  The 8 beta-vectors of each signal are staged per wave
// in LDS and written out as contiguous pencils.
// One work item of the inverse contraction; smf: WPB x NN x OSTRF floats of LDS.
// Workgroup-staged coefficients (GATHER 2): the WPB x SPW = SG signals of the
// workgroup are adjacent in each interleaved coefficient row, so one chunk's
// operands (16 degrees x 8 cases x SG signals) are SG / 2 16-byte loads per
// (degree, case), spread over the workgroup's threads, and reach the lanes
// through LDS as [degree][case][re / im][signal]: one coefficient index per
// thread and chunk, and each cache line read once per workgroup instead of
// once per wave.  Two chunk buffers (CF floats each) alias the output
// staging, which is not in use until the last chunk.
template <int NN, int WPB, int SPW, int PREC, int GATHER>
__device__ __forceinline__ void inv_item_body(const PlanTables& t, const long* __restrict__ roff,
                                              const float* __restrict__ Ti, const SplitTab& ts,
                                              const float2* __restrict__ coeffs, float2* __restrict__ G, int nsig,
                                              long cdist, long cstride, float norm, int layout, BlockWork bw, float* smf,
                                              int wreg)
{
    constexpr bool PAIRED = GATHER == 1;
    constexpr int NKB = NN / 16;
    constexpr int SG = WPB * SPW;
    if (bw.pi >= t.npairs) return;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int s0 = (bw.sg * WPB + w) * SPW;
    // with staged coefficients every wave takes part in the chunk barriers
    if (GATHER != 2 && s0 >= nsig) return;
    const int2 pr = t.pairs[bw.pi];
    const int M = pr.x, Mp = pr.y, R = t.B - M, nch = (R + 15) >> 4;
    const int i = lane & 15, kk = lane >> 4, part = i & 1;
    const BCase cb = bcase(i >> 1, M, Mp, NN);
    const long N3 = (long)NN * NN * NN;

    const float* cs[SPW];
    const float* cp[SPW];  // PAIRED: the even signal of each pair (a valid one past nsig); row length even
#pragma unroll
    for (int sp = 0; sp < SPW; ++sp) {
        cs[sp] = reinterpret_cast<const float*>(coeffs + (long)min(s0 + sp, nsig - 1) * cdist) + part;
        cp[sp] = reinterpret_cast<const float*>(coeffs + ((s0 + (sp & ~1)) < nsig ? s0 + (sp & ~1) : 0));
    }
    const long toff = (roff[bw.pi] >> 4) * NKB * 256 + 16 * i + 4 * kk;
    const float* ti = Ti + toff;

    v4f acc[NKB][SPW];
#pragma unroll
    for (int kb = 0; kb < NKB; ++kb)
#pragma unroll
        for (int sp = 0; sp < SPW; ++sp) acc[kb][sp] = (v4f){0.f, 0.f, 0.f, 0.f};

    // Coefficient operands of degree chunk q are gathered one chunk ahead of
    // the chunk's MFMAs.  The loads are unconditional (coefficient 0 stands in
    // past R and for inactive cases, scaled by 0): a branch around them would
    // make the compiler drain all outstanding loads.
    InvStage<SPW> bc, bn;
    InvTab<NKB, PREC> tab;
    // staged coefficients with FP32 tiles: the chunk's tiles are staged per
    // workgroup too when both buffer pairs fit in the output staging (NKB /
    // WPB 16-byte loads per thread instead of NKB per lane, under the same
    // per-chunk barrier).  Measured (interleaved A/B, two runs each, same
    // results): 64 signals at 2B = 256 inverse 3.41-3.43 -> 3.10-3.11 ms,
    // 512 signals at 2B = 128 2.83-2.84 -> 2.58 ms.
    constexpr bool LDS_TILES = GATHER == 2 && PREC == 0 && NKB % WPB == 0 &&
                               16 * 8 * 2 * SG * 2 + 2 * NKB * 256 <= WPB * NN * OSTRF;
    if constexpr (!LDS_TILES) {
#pragma unroll
        for (int kb = 0; kb < NKB; ++kb) inv_tab_load<NKB, PREC>(tab, kb, ti, ts, toff);
    }
    if constexpr (GATHER == 2) {
        // loader role: items e = tid + 64 WPB j (j < SPW) of the chunk's
        // 16 x 8 x SG / 2 16-byte loads: (degree rl, case cl, signal pair h)
        constexpr int CF = 16 * 8 * 2 * SG;
        const int tid = threadIdx.x, sbase = bw.sg * SG;
        // per item: local degree, signal pair, the case's (M, M') in the
        // index walk, and the case's degree limit (0 for an inactive case)
        int lrl[SPW], lh[SPW], lmc[SPW], lmpc[SPW], lrr[SPW];
#pragma unroll
        for (int j = 0; j < SPW; ++j) {
            const int e = tid + 64 * WPB * j;
            lh[j] = e % (SG / 2);
            const int rc = e / (SG / 2);
            lrl[j] = rc >> 3;
            const BCase lc = bcase(rc & 7, M, Mp, NN);
            lmc[j] = lc.Mc;
            lmpc[j] = lc.Mpc;
            lrr[j] = lc.active ? R : 0;
        }
        const float* cbase = reinterpret_cast<const float*>(coeffs) + 2 * sbase;
        float4 ld[SPW];
        auto load = [&](int q) __attribute__((always_inline)) {
#pragma unroll
            for (int j = 0; j < SPW; ++j) {
                const int r = 16 * q + lrl[j];
                const long ci = r < lrr[j] ? 2 * cstride * coeff_index32(M + r, lmc[j], lmpc[j]) : 0;
                ld[j] = *reinterpret_cast<const float4*>(cbase + ci + 4 * lh[j]);
            }
        };
        auto put = [&](int buf) __attribute__((always_inline)) {
            float* st = smf + buf * CF;
#pragma unroll
            for (int j = 0; j < SPW; ++j) {
                const int e = tid + 64 * WPB * j;
                const int rc = e / (SG / 2);  // rl * 8 + case
                *reinterpret_cast<float2*>(st + (2 * rc) * SG + 2 * lh[j]) = make_float2(ld[j].x, ld[j].z);
                *reinterpret_cast<float2*>(st + (2 * rc + 1) * SG + 2 * lh[j]) = make_float2(ld[j].y, ld[j].w);
            }
        };
        // this lane's operands of chunk q: degrees 16 q + 4 kk + u, its case
        // and part, signals w SPW .. w SPW + SPW - 1
        auto take = [&](InvStage<SPW>& S, int buf, int q) __attribute__((always_inline)) {
            const float* st = smf + buf * CF;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int r = 16 * q + 4 * kk + u;
                const bool live = cb.active && r < R;
                S.f[u] = live ? norm * (float)((cb.flip && (r & 1)) ? -cb.sgn : cb.sgn) : 0.0f;
                const float* src = st + (2 * ((4 * kk + u) * 8 + (i >> 1)) + part) * SG + w * SPW;
#pragma unroll
                for (int sp = 0; sp < SPW; ++sp) S.v[sp][u] = src[sp];
            }
        };
        // LDS_TILES: chunk q's NKB tiles (NKB x 256 floats, contiguous per
        // pair) as NKB / WPB float4 per thread, into tile buffer q & 1
        constexpr int TF = NKB * 256, TJ = LDS_TILES ? NKB / WPB : 1;
        const float* tsrc = Ti + (roff[bw.pi] >> 4) * NKB * 256;
        float* tb = smf + 2 * CF;
        v4f tl[TJ];
        auto tload = [&](int q) __attribute__((always_inline)) {
#pragma unroll
            for (int j = 0; j < TJ; ++j)
                tl[j] = *reinterpret_cast<const v4f*>(tsrc + (long)q * TF + 4 * (tid + 64 * WPB * j));
        };
        auto tput = [&](int buf) __attribute__((always_inline)) {
#pragma unroll
            for (int j = 0; j < TJ; ++j) *reinterpret_cast<v4f*>(tb + buf * TF + 4 * (tid + 64 * WPB * j)) = tl[j];
        };
        load(0);
        if constexpr (LDS_TILES) tload(0);
        // a workgroup's waves may still be reading a previous item's output
        // staging, which the chunk buffers alias
        __syncthreads();
        put(0);
        if constexpr (LDS_TILES) tput(0);
        for (int q = 0; q < nch; ++q) {
            if (q + 1 < nch) {
                load(q + 1);
                if constexpr (LDS_TILES) tload(q + 1);
            }
            __syncthreads();  // chunk q staged by every thread; chunk q - 1's reads done
            take(bc, q & 1, q);
            if constexpr (LDS_TILES) {
                inv_contract_lds<NKB, SPW>(acc, bc, tb + (q & 1) * TF + 16 * i + 4 * kk);
            } else {
                const int qn = q + 1 < nch ? q + 1 : q;
                inv_contract<NKB, SPW, PREC>(acc, bc, tab, ti + qn * NKB * 256, ts, toff + (long)qn * NKB * 256);
            }
            if (q + 1 < nch) {
                put((q + 1) & 1);
                if constexpr (LDS_TILES) tput((q + 1) & 1);
            }
        }
        __syncthreads();  // the last chunk's reads done before the output staging reuses the region
        if (s0 >= nsig) return;
    } else {
        inv_gather<SPW, PAIRED>(bc, cs, cp, cb, M, R, cstride, norm, kk, 0, part);
        for (int q = 0; q < nch; ++q) {
            const int qn = q + 1 < nch ? q + 1 : q;
            inv_gather<SPW, PAIRED>(bn, cs, cp, cb, M, R, cstride, norm, kk, qn, part);
            __builtin_amdgcn_sched_barrier(0);
            inv_contract<NKB, SPW, PREC>(acc, bc, tab, ti + qn * NKB * 256, ts, toff + (long)qn * NKB * 256);
            __builtin_amdgcn_sched_barrier(0);
            bc = bn;
        }
    }

    // write each signal's 8 beta-vectors, staged so every case run is contiguous
    // wave w's region: wreg floats (>= NN OSTRF; the largest NN of a
    // persistent launch, so items of different sizes never share a region)
    float* sw = smf + (long)w * wreg;
#pragma unroll
    for (int sp = 0; sp < SPW; ++sp) {
        const int s = s0 + sp;
        if (s >= nsig) break;
#pragma unroll
        for (int kb = 0; kb < NKB; ++kb)
#pragma unroll
            for (int v = 0; v < 4; ++v) sw[(16 * kb + out_row(lane, v, layout)) * OSTRF + i] = acc[kb][sp][v];
        __builtin_amdgcn_s_waitcnt(0xc07f);
        __builtin_amdgcn_wave_barrier();
        for (int c = 0; c < 8; ++c) {
            const BCase co = bcase(c, M, Mp, NN);
            if (!co.active) continue;
            float2* dst = G + (long)s * N3 + co.vec;
            // non-temporal: a batch's pencils are far beyond the Infinity Cache
            for (int k = lane; k < NN; k += 64) {
                v2f val;
                val.x = sw[k * OSTRF + 2 * c];
                val.y = sw[k * OSTRF + 2 * c + 1];
                __builtin_nontemporal_store(val, reinterpret_cast<v2f*>(dst + (co.flip ? NN - 1 - k : k)));
            }
        }
        __builtin_amdgcn_s_waitcnt(0xc07f);
        __builtin_amdgcn_wave_barrier();
    }
}

// The paired gathers need the interleaved layout (cdist 1) with an even row
// length, so 16-byte loads stay aligned and inside the row.  Measured per 64
// signals (interleaved A/B, two runs each, same results): 2B = 128 inverse
// 3.30 -> 3.01-3.03 ms (512 x B = 64: 25.0 k -> 25.4 k pairs/s); 2B = 256
// 3.61 -> 3.66-3.67 ms, so 2B = 256 keeps the 8-byte gathers.
template <int NN, int WPB, int SPW, int PREC, bool LDSC = false>
__device__ __forceinline__ void inv_item(const PlanTables& t, const long* __restrict__ roff,
                                         const float* __restrict__ Ti, const SplitTab& ts,
                                         const float2* __restrict__ coeffs, float2* __restrict__ G, int nsig,
                                         long cdist, long cstride, float norm, int layout, BlockWork bw, float* smf,
                                         int wreg = NN * OSTRF)
{
    // workgroup-staged coefficients (LDSC: the single-bandwidth kernel): the
    // two chunk buffers fit in the output staging (2 x 256 SG <= WPB NN OSTRF
    // floats for NN >= 128), every signal group inside the padded row.
    // Measured (interleaved A/B, two runs each, same results): 64 signals at
    // 2B = 256 3.52-3.53 -> 3.34-3.36 ms, 512 signals at 2B = 128 2.91 ->
    // 2.74 ms, against the gathers below
    if constexpr (LDSC && NN >= 128 && (WPB * SPW) % 2 == 0) {
        if (cdist == 1 && cstride % (WPB * SPW) == 0) {
            inv_item_body<NN, WPB, SPW, PREC, 2>(t, roff, Ti, ts, coeffs, G, nsig, cdist, cstride, norm, layout, bw,
                                                 smf, wreg);
            return;
        }
    }
    if constexpr (SPW % 2 == 0 && NN <= 128) {
        if (cdist == 1 && (cstride & 1) == 0) {
            inv_item_body<NN, WPB, SPW, PREC, 1>(t, roff, Ti, ts, coeffs, G, nsig, cdist, cstride, norm, layout, bw,
                                                 smf, wreg);
            return;
        }
    }
    inv_item_body<NN, WPB, SPW, PREC, 0>(t, roff, Ti, ts, coeffs, G, nsig, cdist, cstride, norm, layout, bw, smf,
                                         wreg);
}

template <int NN, int WPB, int SPW, int PREC>
__global__ __launch_bounds__(64 * WPB) void k_dwt_inv_batch_f32(PlanTables t, const long* __restrict__ roff,
                                                                const float* __restrict__ Ti, SplitTab ts,
                                                                const float2* __restrict__ coeffs, float2* __restrict__ G,
                                                                int nsig, int nsg, long cdist, long cstride, float norm,
                                                                int layout)
{
    extern __shared__ __align__(16) float smf[];
    inv_item<NN, WPB, SPW, PREC, true>(t, roff, Ti, ts, coeffs, G, nsig, cdist, cstride, norm, layout, block_work(nsg),
                                       smf);
}

// ---- contraction over a batch of mixed bandwidths -----------------------------
// One persistent launch per direction for signals of several bandwidths: every
// workgroup walks a host-built queue of work items (bandwidth group, pair,
// signal group), items b, b + gridDim.x, ...  The host orders the queue
// heaviest first across all groups and puts the signal groups of one pair in
// positions of one residue mod 8, so (with a grid that is a multiple of 8 and
// round-robin dispatch) they share one XCD's L2 copy of the pair's tiles.
// Each item runs the single-bandwidth contraction body of its group's 2B.
template <bool FWD, int NN>
__device__ __forceinline__ void multi_item(const BatchMultiGroup& g, BlockWork bw, float* smf, int layout, int wreg)
{
    if constexpr (FWD)
        fwd_item<NN, MULTI_WPB, MULTI_SPW, 0>(g.t, g.roff, g.tab, SplitTab{}, (const float2*)g.in, g.out,
                                              g.nsig, g.cdist, g.cstride, g.norm, layout, bw);
    else
        // per-lane gathers: the staged coefficients' per-chunk barriers
        // measured slower in this persistent kernel (fixed 64-request batch
        // of the serving mix 1.545 -> 1.636 ms), whose waves otherwise drift
        // apart across items
        inv_item<NN, MULTI_WPB, MULTI_SPW, 0, false>(g.t, g.roff, g.tab, SplitTab{}, (const float2*)g.in, g.out, g.nsig,
                                              g.cdist, g.cstride, g.norm, layout, bw, smf, wreg);
}

template <bool FWD, int NNMAX>
__global__ __launch_bounds__(64 * MULTI_WPB) void k_dwt_batch_multi(const BatchMultiGroup* __restrict__ groups,
                                                                    const int4* __restrict__ items, int nitems,
                                                                    int layout)
{
    extern __shared__ __align__(16) float smf[];
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const int4 w = items[it];
        if (w.x < 0) continue;  // padding
        const BatchMultiGroup& g = groups[w.x];
        const BlockWork bw{w.y, w.z};
        switch (g.t.N) {
        case 16: multi_item<FWD, 16>(g, bw, smf, layout, NNMAX * OSTRF); break;
        case 32: if constexpr (NNMAX >= 32) multi_item<FWD, 32>(g, bw, smf, layout, NNMAX * OSTRF); break;
        case 64: if constexpr (NNMAX >= 64) multi_item<FWD, 64>(g, bw, smf, layout, NNMAX * OSTRF); break;
        case 128: if constexpr (NNMAX >= 128) multi_item<FWD, 128>(g, bw, smf, layout, NNMAX * OSTRF); break;
        default: if constexpr (NNMAX >= 256) multi_item<FWD, 256>(g, bw, smf, layout, NNMAX * OSTRF); break;
        }
    }
}

// Coefficient sets between the per-signal layout [s][count] (s < n) and the
// signal-interleaved layout [idx][ld] the contraction kernels use, so that
// their per-coefficient accesses cover the signals of a workgroup in one
// contiguous run.  64 coefficients x 32 signals per workgroup, through LDS.
// The per-signal side is either contiguous (ps + s * count) or, with sptr
// set, signal s at sptr[s].
template <bool TO_INTERLEAVED>
__global__ __launch_bounds__(256) void k_coeff_transpose(float2* __restrict__ ps, float2* const* __restrict__ sptr,
                                                         float2* __restrict__ il, long count, int n, int ld)
{
    __shared__ float2 tile[32][65];
    const long i0 = (long)blockIdx.x * 64;
    const int s0 = blockIdx.y * 32, t = threadIdx.x;
    auto row = [&](int s) { return sptr ? sptr[s] : ps + (long)s * count; };
    if (TO_INTERLEAVED) {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int sl = (t >> 6) + 4 * r, il_ = t & 63;
            if (s0 + sl < n && i0 + il_ < count) tile[sl][il_] = row(s0 + sl)[i0 + il_];
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int ix = (t >> 5) + 8 * r, sl = t & 31;
            if (s0 + sl < n && i0 + ix < count) il[(i0 + ix) * ld + s0 + sl] = tile[sl][ix];
        }
    } else {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int ix = (t >> 5) + 8 * r, sl = t & 31;
            if (s0 + sl < n && i0 + ix < count) tile[sl][ix] = il[(i0 + ix) * ld + s0 + sl];
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int sl = (t >> 6) + 4 * r, il_ = t & 63;
            if (s0 + sl < n && i0 + il_ < count) row(s0 + sl)[i0 + il_] = tile[sl][il_];
        }
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

inline dim3 batch_grid(int npairs, int nsg) { return dim3((unsigned)(8 * ((npairs + 7) / 8) * nsg)); }

// Signals per wave a kernel instance may hold: at 2B = 256 four signals of
// accumulators (8 chunks x 4 x 4 forward, 16 blocks x 4 x 4 inverse) would
// leave one wave per SIMD.
constexpr int max_spw(int NN, bool fwd) { return NN >= 256 ? 2 : (fwd || NN <= 64 ? 4 : 2); }

inline SplitTab split_tab(const BatchTables& bt, bool fwd)
{
    SplitTab ts;
    for (int pp = 0; pp < 3; ++pp) ts.p[pp] = fwd ? bt.tf16[pp] : bt.ti16[pp];
    return ts;
}

template <int NN, int WPB, int SPW, int PREC>
hipError_t fwd_launch(const PlanTables& t, const BatchTables& bt, const float2* G, float2* coeffs, int nsig, long cdist,
                      long cstride, float norm, hipStream_t st)
{
    if constexpr (SPW > max_spw(NN, true)) {
        return hipErrorInvalidValue;
    } else {
        const int nsg = (nsig + WPB * SPW - 1) / (WPB * SPW);
        return pfs::launch((k_dwt_fwd_batch_f32<NN, WPB, SPW, PREC>), batch_grid(t.npairs, nsg), dim3(64 * WPB), 0, st,
                           t, bt.roff, bt.tf, split_tab(bt, true), G, coeffs, nsig, nsg, cdist, cstride, norm,
                           g_mfma_f32_layout);
    }
}

template <int NN, int WPB, int SPW, int PREC>
hipError_t inv_launch(const PlanTables& t, const BatchTables& bt, const float2* coeffs, float2* G, int nsig, long cdist,
                      long cstride, float norm, hipStream_t st)
{
    if constexpr (SPW > max_spw(NN, false)) {
        return hipErrorInvalidValue;
    } else {
        const size_t smem = sizeof(float) * (size_t)WPB * NN * OSTRF;
        if (smem > lds_limit()) return hipErrorInvalidConfiguration;
        const int nsg = (nsig + WPB * SPW - 1) / (WPB * SPW);
        return pfs::launch((k_dwt_inv_batch_f32<NN, WPB, SPW, PREC>), batch_grid(t.npairs, nsg), dim3(64 * WPB), smem,
                           st, t, bt.roff, bt.ti, split_tab(bt, false), coeffs, G, nsig, nsg, cdist, cstride, norm,
                           g_mfma_f32_layout);
    }
}

// Split precisions run at most two signals per wave (their operands take more
// registers), so their instances are the shapes of one and two.
template <int NN, int PREC>
hipError_t fwd_nn(const PlanTables& t, const BatchTables& bt, const float2* G, float2* coeffs, int nsig, long cdist,
                  long cstride, float norm, BatchKernelShape sh, hipStream_t st)
{
#define PFS_B(W, S) \
    if (sh.wpb == W && sh.spw == S) return fwd_launch<NN, W, S, PREC>(t, bt, G, coeffs, nsig, cdist, cstride, norm, st);
    PFS_B(1, 1) PFS_B(1, 2) PFS_B(2, 1) PFS_B(2, 2) PFS_B(4, 1) PFS_B(4, 2)
    if constexpr (PREC == 0) { PFS_B(1, 4) PFS_B(2, 4) PFS_B(4, 4) }
#undef PFS_B
    return hipErrorInvalidValue;
}

template <int NN, int PREC>
hipError_t inv_nn(const PlanTables& t, const BatchTables& bt, const float2* coeffs, float2* G, int nsig, long cdist,
                  long cstride, float norm, BatchKernelShape sh, hipStream_t st)
{
#define PFS_B(W, S) \
    if (sh.wpb == W && sh.spw == S) return inv_launch<NN, W, S, PREC>(t, bt, coeffs, G, nsig, cdist, cstride, norm, st);
    PFS_B(1, 1) PFS_B(1, 2) PFS_B(2, 1) PFS_B(2, 2) PFS_B(4, 1) PFS_B(4, 2)
    if constexpr (PREC == 0) { PFS_B(1, 4) PFS_B(2, 4) PFS_B(4, 4) }
#undef PFS_B
    return hipErrorInvalidValue;
}

template <int NN>
hipError_t fwd_prec(const PlanTables& t, const BatchTables& bt, const float2* G, float2* coeffs, int nsig, long cdist,
                    long cstride, float norm, BatchKernelShape sh, int prec, hipStream_t st)
{
    switch (prec) {
    case 0: return fwd_nn<NN, 0>(t, bt, G, coeffs, nsig, cdist, cstride, norm, sh, st);
    case 3: return fwd_nn<NN, 3>(t, bt, G, coeffs, nsig, cdist, cstride, norm, sh, st);
    case 6: return fwd_nn<NN, 6>(t, bt, G, coeffs, nsig, cdist, cstride, norm, sh, st);
    default: return hipErrorInvalidValue;
    }
}

template <int NN>
hipError_t inv_prec(const PlanTables& t, const BatchTables& bt, const float2* coeffs, float2* G, int nsig, long cdist,
                    long cstride, float norm, BatchKernelShape sh, int prec, hipStream_t st)
{
    switch (prec) {
    case 0: return inv_nn<NN, 0>(t, bt, coeffs, G, nsig, cdist, cstride, norm, sh, st);
    case 3: return inv_nn<NN, 3>(t, bt, coeffs, G, nsig, cdist, cstride, norm, sh, st);
    case 6: return inv_nn<NN, 6>(t, bt, coeffs, G, nsig, cdist, cstride, norm, sh, st);
    default: return hipErrorInvalidValue;
    }
}

// Bf16 terms of every FP32 table value (the tiles keep their element order).
__global__ void k_split_tables_bf16(const float* __restrict__ T, long n, unsigned short* __restrict__ h,
                                    unsigned short* __restrict__ m, unsigned short* __restrict__ l)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        float r = T[i];
        unsigned b = bf16_rne_bits(r);
        h[i] = (unsigned short)(b >> 16);
        r -= __uint_as_float(b);
        b = bf16_rne_bits(r);
        m[i] = (unsigned short)(b >> 16);
        r -= __uint_as_float(b);
        l[i] = (unsigned short)(bf16_rne_bits(r) >> 16);
    }
}

// Self-test probe of the split path's operand packing: a 16 x 16 x 16 product
// of small integers (exact in bf16 and FP32) as four FP32 16x16x4 MFMAs and
// as one bf16 16x16x16 MFMA, with the K-ordered packing split4 produces.
__global__ void k_mfma_probe_bf16(float* out)
{
    const int lane = threadIdx.x, i = lane & 15, kk = lane >> 4;
    float a[4], b[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int k = 4 * kk + u;
        a[u] = (float)(((i * 7 + k * 3) % 9) - 4);   // A[i][k]
        b[u] = (float)(((k * 5 + i * 11) % 7) - 3);  // B[k][n = i]
    }
    v4f f = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int u = 0; u < 4; ++u) f = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u], b[u], f, 0, 0, 0);
    v4s as[1], bs[1];
    split4<1>(a, as);
    split4<1>(b, bs);
    v4f h = {0.f, 0.f, 0.f, 0.f};
    h = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(as[0], bs[0], h, 0, 0, 0);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        out[lane * 4 + v] = f[v];
        out[256 + lane * 4 + v] = h[v];
    }
}

}  // namespace

long batch_table_rows(int B)
{
    long rows = 0;
    for (int M = 0; M < B; ++M) rows += (long)(M + 1) * (((B - M + 15) / 16) * 16);
    return rows;
}

bool batch_supported(int N) { return N == 16 || N == 32 || N == 64 || N == 128 || N == 256; }

hipError_t launch_wigner_tiles_f32(const PlanTables& t, const BatchTables& bt, hipStream_t st)
{
    return pfs::launch(k_wigner_tiles_f32, dim3(t.npairs), dim3(t.N < 256 ? t.N : 256), 0, st, t, bt.roff,
                       (float*)bt.tf, (float*)bt.ti);
}

BatchKernelShape batch_default_shape(int N, bool fwd, int n)
{
    // the full shape (4 waves of the most signals a wave may hold) once the batch
    // fills it; fewer signals per wave and fewer waves for small batches, so no
    // wave idles and no wave repeats a clamped signal
    const int smax = N >= 256 ? 2 : (fwd || N <= 64 ? 4 : 2);
    const int spw = n >= smax ? smax : (n >= 2 ? 2 : 1);
    const int per = (n + spw - 1) / spw;
    return {per >= 4 ? 4 : (per >= 2 ? 2 : 1), spw};
}

bool batch_shape_valid(int N, bool fwd, BatchKernelShape sh)
{
    const int smax = N >= 256 ? 2 : (fwd || N <= 64 ? 4 : 2);
    return (sh.wpb == 1 || sh.wpb == 2 || sh.wpb == 4) && (sh.spw == 1 || sh.spw == 2 || sh.spw == 4) &&
           sh.spw <= smax;
}

// The default shape for a batch, with at most two signals per wave in the
// split precisions.
static BatchKernelShape prec_shape(int N, bool fwd, int nsig, BatchKernelShape sh, int prec)
{
    if (sh.wpb == 0) sh = batch_default_shape(N, fwd, nsig);
    if (prec != 0 && sh.spw > 2) sh.spw = 2;
    return sh;
}

hipError_t launch_dwt_fwd_batch_f32(const PlanTables& t, const BatchTables& bt, const float2* G, float2* coeffs,
                                    int nsig, long cdist, long cstride, float norm, BatchKernelShape sh, hipStream_t st,
                                    int prec)
{
    if (prec != 0 && (!g_mfma_bf16_ok || !bt.tf16[0])) return hipErrorInvalidValue;
    sh = prec_shape(t.N, true, nsig, sh, prec);
    switch (t.N) {
    case 16: return fwd_prec<16>(t, bt, G, coeffs, nsig, cdist, cstride, norm, sh, prec, st);
    case 32: return fwd_prec<32>(t, bt, G, coeffs, nsig, cdist, cstride, norm, sh, prec, st);
    case 64: return fwd_prec<64>(t, bt, G, coeffs, nsig, cdist, cstride, norm, sh, prec, st);
    case 128: return fwd_prec<128>(t, bt, G, coeffs, nsig, cdist, cstride, norm, sh, prec, st);
    case 256: return fwd_prec<256>(t, bt, G, coeffs, nsig, cdist, cstride, norm, sh, prec, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_dwt_inv_batch_f32(const PlanTables& t, const BatchTables& bt, const float2* coeffs, float2* G,
                                    int nsig, long cdist, long cstride, float norm, BatchKernelShape sh, hipStream_t st,
                                    int prec)
{
    if (prec != 0 && (!g_mfma_bf16_ok || !bt.ti16[0])) return hipErrorInvalidValue;
    sh = prec_shape(t.N, false, nsig, sh, prec);
    switch (t.N) {
    case 16: return inv_prec<16>(t, bt, coeffs, G, nsig, cdist, cstride, norm, sh, prec, st);
    case 32: return inv_prec<32>(t, bt, coeffs, G, nsig, cdist, cstride, norm, sh, prec, st);
    case 64: return inv_prec<64>(t, bt, coeffs, G, nsig, cdist, cstride, norm, sh, prec, st);
    case 128: return inv_prec<128>(t, bt, coeffs, G, nsig, cdist, cstride, norm, sh, prec, st);
    case 256: return inv_prec<256>(t, bt, coeffs, G, nsig, cdist, cstride, norm, sh, prec, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_dwt_batch_multi(bool fwd, const BatchMultiGroup* groups, const int4* items, int nitems, int nnmax,
                                  int grid, hipStream_t st)
{
    if (nitems < 1 || grid < 1) return hipErrorInvalidValue;
    const dim3 g((unsigned)grid), b(64 * MULTI_WPB);
    // the inverse's per-wave output regions are sized for the kernel's NNM
    // (every item of the launch uses the same regions)
#define PFS_M(NNM)                                                                                           \
    if (nnmax <= NNM) {                                                                                      \
        const size_t smem = fwd ? 0 : sizeof(float) * (size_t)MULTI_WPB * NNM * OSTRF;                       \
        if (smem > lds_limit()) return hipErrorInvalidConfiguration;                                         \
        return fwd ? pfs::launch((k_dwt_batch_multi<true, NNM>), g, b, smem, st, groups, items, nitems,     \
                                 g_mfma_f32_layout)                                                         \
                   : pfs::launch((k_dwt_batch_multi<false, NNM>), g, b, smem, st, groups, items, nitems,    \
                                 g_mfma_f32_layout);                                                        \
    }
    PFS_M(32) PFS_M(64) PFS_M(128) PFS_M(256)
#undef PFS_M
    return hipErrorInvalidValue;
}

hipError_t launch_split_tables_bf16(const float* T, long n, unsigned short* h, unsigned short* m, unsigned short* l,
                                    hipStream_t st)
{
    const unsigned blocks = (unsigned)std::min<long>((n + 255) / 256, 16384);
    return pfs::launch(k_split_tables_bf16, dim3(blocks), dim3(256), 0, st, T, n, h, m, l);
}

int bf16_selftest(hipStream_t st)
{
    float* d = nullptr;
    if (hipMalloc(&d, 512 * sizeof(float)) != hipSuccess) return -1;
    hipError_t e = pfs::launch(k_mfma_probe_bf16, dim3(1), dim3(64), 0, st, d);
    float h[512];
    if (e == hipSuccess) e = hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(d);
    if (e != hipSuccess) return -1;
    bool same = true;
    for (int i = 0; i < 256; ++i) same = same && h[i] == h[256 + i];
    g_mfma_bf16_ok = same;
    return same ? 0 : 1;
}

hipError_t launch_coeff_transpose(float2* per_signal, float2* const* sptr, float2* interleaved, long count, int n,
                                  int ld, bool to_interleaved, hipStream_t st)
{
    if (n < 1 || ld < n) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((count + 63) / 64), (unsigned)((n + 31) / 32));
    if (to_interleaved)
        return pfs::launch(k_coeff_transpose<true>, grid, dim3(256), 0, st, per_signal, sptr, interleaved, count, n, ld);
    return pfs::launch(k_coeff_transpose<false>, grid, dim3(256), 0, st, per_signal, sptr, interleaved, count, n, ld);
}

hipError_t launch_nyquist_zero_f32(float2* G, int B, int nsig, hipStream_t st)
{
    const int N = 2 * B;
    const long total = (long)(2 * N - 1) * N * nsig;
    return pfs::launch(k_nyquist_zero_f32, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, G, B, nsig);
}

}  // namespace pfs
