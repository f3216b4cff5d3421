// Discrete Wigner transform stage, FP64, Wigner-d generated on the fly.
//
// One wave per canonical pair (M, M'), M >= M' >= 0 (the default kernels: the
// node-blocked forward k_dwt_fwd_f64_nb and the staged vector-ALU inverse
// k_dwt_inv_f64_staged, at every 2B up to 512; the chain-parallel kernels up
// to 2B = 256).  Lane i of the wave runs the three-term recurrence of the
// L2-normalised Wigner d-function (fn_dwt.hpp:152-199: base row, then rows
// l = M+1 .. B-1) in FP64 registers for the beta nodes k = i, i+64, ... --
// one recurrence chain per 64 nodes.  Sixteen degrees of all chains at a time
// go through a 16 x (64 x chains) LDS tile into the
// v_mfma_f64_16x16x4_f64 operand layout; the MFMAs contract them against the
// 16 real right-hand sides of the pair.  The 8 symmetric cases
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
// above already holds the net sign.  The forward transform uses the weighted
// table: like weighted_wigner_d_matrix (fn_dwt.hpp:162) the quadrature weight
// multiplies the base row and propagates through the linear recurrence.
#include "pfs_internal.h"

#include <type_traits>

namespace pfs {

constexpr int OSTR = 18;            // output staging row stride (doubles), 64 rows

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

// The same, stored as three rows of `stride` values.
__device__ __forceinline__ void recurrence_coeffs(int M, int Mp, int r, int R, double* cf, int stride)
{
    double f1, f1f2, c1;
    recurrence_step(M, Mp, r, R, f1, f1f2, c1);
    cf[r] = f1;
    cf[stride + r] = f1f2;
    cf[2 * stride + r] = c1;
}

// MFMA f64 16x16x4 output layout (gfx950, checked by the context's
// self-test): lane L holds D[4v + L/16][L%16] in accumulator register v < 4.
__device__ __forceinline__ int mfma_out_row(int lane, int v) { return 4 * v + (lane >> 4); }
__device__ __forceinline__ int mfma_out_col(int lane) { return lane & 15; }

// Per-lane recurrence chains: chain c covers node k = kbase + 64 c.
template <int NCH>
struct Chains {
    double cb[NCH], cur[NCH], prev[NCH];
};

// Loads the node tables of every chain for the pair (M, M'): cos(beta), the
// base-row powers sin(beta/2)^(M-M') and cos(beta/2)^(M+M') from the plan's
// power tables, and the weight.  Separate from init_chains so the loads can
// all be in flight before any of them is used.
template <int NCH>
struct ChainTables {
    double cb[NCH], ps[NCH], pc[NCH], w[NCH];
};

template <int NCH>
__device__ __forceinline__ void load_chain_tables(ChainTables<NCH>& ct, const PlanTables& t, int kbase, bool weighted,
                                                  int M, int Mp)
{
    const double* ps = t.pws + (M - Mp) * t.N;
    const double* pc = t.pwc + (M + Mp) * t.N;
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        const int k = min(kbase + 64 * c, t.N - 1);
        ct.cb[c] = t.cosb[k];
        ct.ps[c] = ps[k];
        ct.pc[c] = pc[k];
        ct.w[c] = weighted ? t.wts[k] : 1.0;
    }
}

// Base row d^M_{M M'}(beta) = nf sin(beta/2)^(M-M') cos(beta/2)^(M+M')
// (times the weight for the forward), fn_dwt.hpp:152-160.
template <int NCH>
__device__ __forceinline__ void init_chains(Chains<NCH>& ch, const ChainTables<NCH>& ct, double nf, int kbase, int N)
{
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        const bool kv = kbase + 64 * c < N;
        ch.cb[c] = kv ? ct.cb[c] : 0.0;
        ch.cur[c] = kv ? nf * ct.ps[c] * ct.pc[c] * ct.w[c] : 0.0;
        ch.prev[c] = 0.0;
    }
}

// Workgroup b handles the pair t.xpairs[b]: the host orders the pairs for the
// 8 XCDs (round-robin workgroup dispatch) so that neighbouring pairs, whose
// coefficients share cache lines, run on one XCD at about the same time.

// Orders one wave's LDS accesses around a point: its earlier LDS operations
// complete before any later one (no other wave is waited for).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
    __builtin_amdgcn_wave_barrier();
}

// The forward reads each pencil once: a non-temporal load does not allocate in
// the caches, so it neither evicts the pencils the FFT's last pass left in the
// Infinity Cache nor the coefficient lines neighbouring pairs share in L2.
// (Pencil stores of the inverse stay ordinary: the next FFT pass reads what
// they leave in the Infinity Cache.)
typedef double d2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 load_nt(const double2* p)
{
    const d2v v = __builtin_nontemporal_load(reinterpret_cast<const d2v*>(p));
    double2 r;
    r.x = v.x;
    r.y = v.y;
    return r;
}

// acc += (lane q of this lane's 16-lane row of c) * d, as one v_fmac_f64_dpp
// (row_newbcast: DPP on 64-bit ALU ops; the compiler does not fold a separate
// broadcast move into the FMA by itself).  The statements are volatile, so
// they keep their program order: the first one after c is written (FIRST)
// starts with the two wait states a DPP read of a VALU-written VGPR needs,
// the same s_nop 1 the compiler places before its own DPP moves.
#define PFS_FMAC_BCAST(Q)                                                                                   \
    template <bool FIRST> __device__ __forceinline__ void fmac_bcast_##Q(double& acc, double c, double d)   \
    {                                                                                                       \
        if (FIRST)                                                                                          \
            asm volatile("s_nop 1\n\tv_fmac_f64_dpp %0, %1, %2 row_newbcast:" #Q " row_mask:0xf bank_mask:0xf" \
                         : "+v"(acc) : "v"(c), "v"(d));                                                     \
        else                                                                                                \
            asm volatile("v_fmac_f64_dpp %0, %1, %2 row_newbcast:" #Q " row_mask:0xf bank_mask:0xf"           \
                         : "+v"(acc) : "v"(c), "v"(d));                                                     \
    }
PFS_FMAC_BCAST(0) PFS_FMAC_BCAST(1) PFS_FMAC_BCAST(2) PFS_FMAC_BCAST(3)
PFS_FMAC_BCAST(4) PFS_FMAC_BCAST(5) PFS_FMAC_BCAST(6) PFS_FMAC_BCAST(7)
PFS_FMAC_BCAST(8) PFS_FMAC_BCAST(9) PFS_FMAC_BCAST(10) PFS_FMAC_BCAST(11)
PFS_FMAC_BCAST(12) PFS_FMAC_BCAST(13) PFS_FMAC_BCAST(14) PFS_FMAC_BCAST(15)
#undef PFS_FMAC_BCAST

// q is a constant once the caller's loop is unrolled; the switch folds away.
template <bool FIRST> __device__ __forceinline__ void fmac_bcast(int q, double& acc, double c, double d)
{
    switch (q) {
    case 0: fmac_bcast_0<FIRST>(acc, c, d); break;
    case 1: fmac_bcast_1<FIRST>(acc, c, d); break;
    case 2: fmac_bcast_2<FIRST>(acc, c, d); break;
    case 3: fmac_bcast_3<FIRST>(acc, c, d); break;
    case 4: fmac_bcast_4<FIRST>(acc, c, d); break;
    case 5: fmac_bcast_5<FIRST>(acc, c, d); break;
    case 6: fmac_bcast_6<FIRST>(acc, c, d); break;
    case 7: fmac_bcast_7<FIRST>(acc, c, d); break;
    case 8: fmac_bcast_8<FIRST>(acc, c, d); break;
    case 9: fmac_bcast_9<FIRST>(acc, c, d); break;
    case 10: fmac_bcast_10<FIRST>(acc, c, d); break;
    case 11: fmac_bcast_11<FIRST>(acc, c, d); break;
    case 12: fmac_bcast_12<FIRST>(acc, c, d); break;
    case 13: fmac_bcast_13<FIRST>(acc, c, d); break;
    case 14: fmac_bcast_14<FIRST>(acc, c, d); break;
    default: fmac_bcast_15<FIRST>(acc, c, d); break;
    }
}

// Writes degrees r0 .. r0+15 of chains C0 .. C0+NC-1 into the tile (row rr,
// column 64 (c - C0) + lane) and advances them.  Lane L fetches the step
// coefficients of degree r0 + L % 16 in one LDS read each, so every 16-lane
// row holds all 16; each step then broadcasts its own within the rows, and no
// LDS round trip sits on the recurrence chain.
template <int NCH, int C0, int NC>
__device__ __forceinline__ void run_chains(Chains<NCH>& ch, double* __restrict__ tile, int tstr, int lane,
                                           const double* __restrict__ cf, int RP, int r0)
{
    const int rl = r0 + (lane & 15);
    const double vf1 = cf[rl], vf1f2 = cf[RP + rl], vc1 = cf[2 * RP + rl];
#pragma unroll
    for (int rr = 0; rr < 16; ++rr) {
        const double f1 = row_bcast(vf1, rr), f1f2 = row_bcast(vf1f2, rr), c1 = row_bcast(vc1, rr);
#pragma unroll
        for (int c = C0; c < C0 + NC; ++c) {
            tile[rr * tstr + 64 * (c - C0) + lane] = ch.cur[c];
            double u = fma(f1, ch.cb[c], f1f2);
            double dn = fma(ch.cur[c], u, c1 * ch.prev[c]);
            ch.prev[c] = ch.cur[c];
            ch.cur[c] = dn;
        }
    }
}

// run_chains with two of its three row broadcasts folded into DPP FMAs:
// u = f1 f2 + f1 cb (the f1 f2 broadcast moved, f1 folded) and dn = cur u +
// c1 prev (c1 folded), four VALU instructions per chain and degree where
// run_chains takes six.  dn rounds cur u before adding c1 prev (run_chains
// rounds c1 prev instead): the same accuracy, a different last bit.
template <int NCH, int C0, int NC>
__device__ __forceinline__ void run_chains_fold(Chains<NCH>& ch, double* __restrict__ tile, int tstr, int lane,
                                                const double* __restrict__ cf, int RP, int r0)
{
    const int rl = r0 + (lane & 15);
    const double vf1 = cf[rl], vf1f2 = cf[RP + rl], vc1 = cf[2 * RP + rl];
#pragma unroll
    for (int rr = 0; rr < 16; ++rr) {
        const double f1f2 = row_bcast(vf1f2, rr);
#pragma unroll
        for (int c = C0; c < C0 + NC; ++c) {
            tile[rr * tstr + 64 * (c - C0) + lane] = ch.cur[c];
            double u = f1f2;
            // the DPP sources (vf1, vc1) come from LDS, not from a VALU write;
            // the chunk's first fold waits the two states a DPP read needs anyway
            if (rr == 0 && c == C0) fmac_bcast<true>(rr, u, vf1, ch.cb[c]);
            else fmac_bcast<false>(rr, u, vf1, ch.cb[c]);
            double dn = ch.cur[c] * u;
            fmac_bcast<false>(rr, dn, vc1, ch.prev[c]);
            ch.prev[c] = ch.cur[c];
            ch.cur[c] = dn;
        }
    }
}

// Step coefficients of degrees r = first, first + stride, ... < RP,
// interleaved per degree (cf4[4 r] = f1, f1 f2, c1, 0).
__device__ __forceinline__ void recurrence_coeffs4(int M, int Mp, int R, int RP, double* cf4, int first, int stride)
{
    for (int r = first; r < RP; r += stride) {
        recurrence_step(M, Mp, r, R, cf4[4 * r], cf4[4 * r + 1], cf4[4 * r + 2]);
        cf4[4 * r + 3] = 0.0;
    }
}

// The rescaled recurrence rows of pair (M, M') interleaved per degree:
// cf4[4 r] = (a_r, b_r, P_r, 0), r = first, first + stride, ... < RP.
__device__ __forceinline__ void rescaled_coeffs4(const PlanTables& t, int M, int Mp, int RP, double* cf4, int first,
                                                 int stride)
{
    const long o = t.rsm[M] + (long)Mp * RP;
    for (int r = first; r < RP; r += stride) {
        const double2 ab = t.rab[o + r];
        *reinterpret_cast<double2*>(cf4 + 4 * r) = ab;
        *reinterpret_cast<double2*>(cf4 + 4 * r + 2) = make_double2(t.rp[o + r], 0.0);
    }
}

// LDS tile row stride for a tile of nc chains: 64 nc nodes + 1 (odd, so the
// 16 rows an MFMA operand read touches fall in distinct banks).
__host__ __device__ constexpr int tile_stride(int nc) { return 64 * nc + 1; }

// Row stride of the forward kernels' tile.  gfx950 serves a 64-bit LDS read
// 16 lanes at a time over 32 banks (measured, SQ_LDS_BANK_CONFLICT: a 2-word
// row shift read conflict-free, a 4-word shift 2-way).  The forward's
// A-operand read has lanes (row r, node kk) at words 2 r TS + 2 kk, a quarter
// being one kk and r = 0..15: with TS = 64 nc + 1 the 16 rows sit 2 words
// apart (mod 32) and cover all 32 banks once.
__host__ __device__ constexpr int fwd_tile_stride(int nc) { return 64 * nc + 1; }

// Pencil staging of the forward preamble: 8 cases x 64 nodes x (re, im) for one
// chain, case rows 130 doubles apart (260 words = 4 mod 32): the
// operand-layout read of a 16-lane quarter (case j / 2, re/im j % 2) then
// covers 8 case slots x 4 words = all 32 banks once.  (At 136 doubles, 16 mod
// 32, four cases shared each bank group: 6.3 M conflict cycles per launch.)
constexpr int FSTG_CASE = 130;
constexpr int FSTG = 8 * FSTG_CASE;

// Doubles of LDS per wave of the forward kernel: its 16-row tile, which during
// the preamble holds the pencil staging.
__host__ __device__ constexpr int fwd_wave_region(int nh)
{
    return 16 * fwd_tile_stride(nh) > FSTG ? 16 * fwd_tile_stride(nh) : FSTG;
}

// Doubles of LDS per wave of the inverse kernel: its 16-row tile, which after
// the contraction holds the 64 x OSTR output staging of one chain.
__host__ __device__ constexpr int inv_wave_region(int nc)
{
    return 16 * tile_stride(nc) > 64 * OSTR ? 16 * tile_stride(nc) : 64 * OSTR;
}

// Per-lane view of one case of the pair (hoisted out of the degree loop).
struct LaneCase {
    int Mc, Mpc, active;
    double sgn;      // sign for even rows
    double sgn_odd;  // sign for odd rows (differs for the reversed cases)
};

// Packed indices of the coefficients (l, Mc, Mpc) for l = l0, l0 + 4, l0 + 8,
// ...: for fixed orders the index is a cubic polynomial in l, so it advances
// by forward differences, three integer additions a step in place of the
// multiplications of coeff_index32 (exact while l < 512).
struct IdxWalk {
    int f, d1, d2, d3;
    __device__ __forceinline__ IdxWalk(int l0, int Mc, int Mpc)
    {
        const int f0 = coeff_index32(l0, Mc, Mpc), f1 = coeff_index32(l0 + 4, Mc, Mpc);
        const int f2 = coeff_index32(l0 + 8, Mc, Mpc), f3 = coeff_index32(l0 + 12, Mc, Mpc);
        f = f0;
        d1 = f1 - f0;
        d2 = f2 - 2 * f1 + f0;
        d3 = f3 - 3 * f2 + 3 * f1 - f0;
    }
    __device__ __forceinline__ void step()
    {
        f += d1;
        d1 += d2;
        d2 += d3;
    }
};

__device__ __forceinline__ LaneCase lane_case(int c, int M, int Mp, int N, long* vec)
{
    CaseInfo ci = case_info(c, M, Mp, N);
    LaneCase lc;
    lc.Mc = ci.Mc;
    lc.Mpc = ci.Mpc;
    lc.active = ci.active;
    lc.sgn = ci.sgn;
    lc.sgn_odd = ci.flip ? -ci.sgn : ci.sgn;
    if (vec) *vec = ci.vec;
    return lc;
}

// Forward DWT.  NCH chains of 64 nodes per lane, processed NH at a time
// through a 16 x 64 NH tile; WPP waves per pair split the nodes and sum
// their partial blocks through LDS.  SMALL: 2B < 64, only the first 2B
// nodes (rounded up to 4) carry data.
template <int NCH, int NH, int WPP, bool SMALL>
__global__ __launch_bounds__(64 * WPP) __attribute__((amdgpu_waves_per_eu(2))) void k_dwt_fwd_f64(PlanTables t, const double2* __restrict__ G,
                                                          double2* __restrict__ coeffs, double norm)
{
    extern __shared__ __align__(16) double sm[];
    constexpr int TS = fwd_tile_stride(NH);
    constexpr int WREG = fwd_wave_region(NH);
    const int B = t.B, N = t.N;
    const int2 pr = t.xpairs[blockIdx.x];
    const int M = pr.x, Mp = pr.y;
    const int R = B - M;
    const int nch = (R + 15) >> 4;
    const int RP = nch * 16;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int nmf = SMALL ? ((N + 3) >> 2) : 16;  // MFMAs per chain and chunk

    double* tile = sm + w * WREG;
    double* red = sm + WPP * WREG;   // WPP * 256 (only when WPP > 1)
    double* cf = red + (WPP > 1 ? WPP * 256 : 0);

    const int kw = w * 64 * NCH;

    // issue every global load of the preamble before any of them is used:
    // the node tables first (init_chains waits for them), then the pencils
    const int j = lane & 15, kk = lane >> 4;
    const LaneCase lc = lane_case(j >> 1, M, Mp, N, nullptr);
    ChainTables<NCH> ct;
    load_chain_tables<NCH>(ct, t, kw + lane, true, M, Mp);
    const double nf = t.normfac[M * B + Mp];
    // B operands: lane (j + 16 kk) holds S[kw + 4m + kk][j], j = 2 case + re/im
    double sreg[NCH * 16];
    if (SMALL) {
        long vec;
        (void)lane_case(j >> 1, M, Mp, N, &vec);
        const bool flip = lc.sgn != lc.sgn_odd;
        const double* vsrc = reinterpret_cast<const double*>(G + vec) + (j & 1);
#pragma unroll
        for (int m = 0; m < NCH * 16; ++m) {
            const int ks = kw + 4 * m + kk;
            const int kr = min(flip ? N - 1 - ks : ks, N - 1);
            sreg[m] = vsrc[2 * max(kr, 0)];
        }
    }
    // Non-SMALL: each case's beta-vector is read with whole-wave 16-byte loads
    // (lane = node, 1 KB contiguous per load, reversed for the d(pi - beta)
    // cases) and moved into the operand layout through the wave's tile.
    double2 raw[SMALL ? 1 : NCH][8];
    int cact = 0;
    if (!SMALL) {
#pragma unroll
        for (int cs = 0; cs < 8; ++cs) {
            const CaseInfo ci = case_info(cs, M, Mp, N);
            cact |= ci.active << cs;
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                const int n = min(kw + 64 * c + lane, N - 1);
                raw[c][cs] = load_nt(G + ci.vec + (ci.flip ? N - 1 - n : n));
            }
        }
    }

    for (int r = tid; r < RP; r += 64 * WPP) recurrence_coeffs(M, Mp, r, R, cf, RP);
    Chains<NCH> ch;
    init_chains<NCH>(ch, ct, nf, kw + lane, N);
    if (SMALL) {
#pragma unroll
        for (int m = 0; m < NCH * 16; ++m)
            if (!lc.active || kw + 4 * m + kk >= N) sreg[m] = 0.0;
    } else {
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            const bool live = kw + 64 * c + lane < N;
            if (c) wave_sync();  // the previous chain's operand reads are done
#pragma unroll
            for (int cs = 0; cs < 8; ++cs) {
                const bool on = live && ((cact >> cs) & 1);
                double2 v;
                v.x = on ? raw[c][cs].x : 0.0;
                v.y = on ? raw[c][cs].y : 0.0;
                *reinterpret_cast<double2*>(tile + cs * FSTG_CASE + 2 * lane) = v;
            }
            wave_sync();
#pragma unroll
            for (int m = 0; m < 16; ++m) sreg[16 * c + m] = tile[(j >> 1) * FSTG_CASE + 2 * (4 * m + kk) + (j & 1)];
        }
        wave_sync();  // operand reads done before the tile is rewritten
    }
    // output rows of this lane: 4v + lane/16 within each chunk
    const int orow = lane >> 4;

    for (int chk = 0; chk < nch; ++chk) {
        v4d acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int h = 0; h < NCH / NH; ++h) {
            __syncthreads();  // coefficient table ready / previous tile reads done
            if (h == 0) run_chains<NCH, 0, NH>(ch, tile, TS, lane, cf, RP, chk * 16);
            else if (h == 1) run_chains<NCH, (NH < NCH ? NH : 0), NH>(ch, tile, TS, lane, cf, RP, chk * 16);
            else if (h == 2) run_chains<NCH, (2 * NH < NCH ? 2 * NH : 0), NH>(ch, tile, TS, lane, cf, RP, chk * 16);
            else run_chains<NCH, (3 * NH < NCH ? 3 * NH : 0), NH>(ch, tile, TS, lane, cf, RP, chk * 16);
            __syncthreads();
#pragma unroll
            for (int c = 0; c < NH; ++c)
#pragma unroll
                for (int m = 0; m < 16; ++m) {
                    if (!SMALL || m < nmf) {
                        double a = tile[(lane & 15) * TS + 64 * c + 4 * m + kk];
                        double b = sreg[16 * (h * NH + c) + m];
                        if (m & 1) acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc1, 0, 0, 0);
                        else acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc0, 0, 0, 0);
                    }
                }
        }
        v4d acc = acc0 + acc1;
        if (WPP > 1) {
#pragma unroll
            for (int v = 0; v < 4; ++v) red[w * 256 + (4 * v + orow) * 16 + j] = acc[v];
            __syncthreads();
            if (w == 0) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    double s = 0.0;
                    for (int ww = 0; ww < WPP; ++ww) s += red[ww * 256 + (4 * v + orow) * 16 + j];
                    acc[v] = s;
                }
            }
        }
        if (w == 0 && lc.active) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = chk * 16 + 4 * v + orow;
                if (r < R) {
                    const double sg = (r & 1) ? lc.sgn_odd : lc.sgn;
                    reinterpret_cast<double*>(coeffs + coeff_index32(M + r, lc.Mc, lc.Mpc))[j & 1] = norm * (sg * acc[v]);
                }
            }
        }
    }
}

// Forward DWT, node blocks outermost (2B = 64 NB): one wave per pair walks
// the pair's NB blocks of 64 beta nodes, and for each block runs the
// recurrence over all of the pair's NCHK 16-degree chunks, accumulating every
// chunk's 16 x 16 output block in its own MFMA accumulator.  Only one block's
// operands (16 doubles per lane) are live, so the wave holds about half the
// registers of the chain-parallel kernel, and the next block's pencils and
// node tables are loaded while the current block's chunks compute.  The LDS
// region per wave is the 16 x 65 recurrence tile, which also stages each
// block's pencils into the MFMA operand layout, plus the step coefficients.
template <int NB, int NCHK>
__device__ __forceinline__ void fwd_nb_pair(const PlanTables& t, const double2* __restrict__ G,
                                            double2* __restrict__ coeffs, double norm, int M, int Mp, double* sm)
{
    constexpr int TS = fwd_tile_stride(1);
    constexpr int WREG = 16 * TS > FSTG ? 16 * TS : FSTG;
    constexpr int RP = NCHK * 16;
    double* tile = sm;
    double* cf = sm + WREG;
    const int B = t.B, N = t.N, R = B - M;
    const int lane = threadIdx.x & 63, j = lane & 15, kk = lane >> 4;
    const LaneCase lc = lane_case(j >> 1, M, Mp, N, nullptr);
    long cvec[8];
    int cflip = 0;
#pragma unroll
    for (int cs = 0; cs < 8; ++cs) {
        const CaseInfo ci = case_info(cs, M, Mp, N);
        cvec[cs] = ci.vec;
        cflip |= ci.flip << cs;
    }
    // block b of case cs: node n = 64 b + lane, reversed for the d(pi - beta) cases
    auto load_block = [&](int b, double2 (&raw)[8]) {
        const int n = 64 * b + lane;
#pragma unroll
        for (int cs = 0; cs < 8; ++cs) raw[cs] = load_nt(G + cvec[cs] + (((cflip >> cs) & 1) ? N - 1 - n : n));
    };
    double2 raw[8];
    load_block(0, raw);
    const double nf = t.normfac[M * B + Mp];
    for (int r = lane; r < RP; r += 64) recurrence_coeffs(M, Mp, r, R, cf, RP);

    v4d acc[NCHK];
#pragma unroll
    for (int c = 0; c < NCHK; ++c) acc[c] = (v4d){0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
    for (int b = 0; b < NB; ++b) {
        // this block's node tables (small, L2-resident), loaded here rather
        // than a block ahead so that they hold no registers across the chunks
        ChainTables<1> ct;
        load_chain_tables<1>(ct, t, 64 * b + lane, true, M, Mp);
        wave_sync();  // step coefficients written / the previous block's tile reads done
        // an inactive case reads a valid pencil (another case's): its output
        // columns are computed like the others and never stored
#pragma unroll
        for (int cs = 0; cs < 8; ++cs) *reinterpret_cast<double2*>(tile + cs * FSTG_CASE + 2 * lane) = raw[cs];
        wave_sync();
        // B operands: lane (j + 16 kk) holds S[64 b + 4 m + kk][j], j = 2 case + re/im
        double sreg[16];
#pragma unroll
        for (int m = 0; m < 16; ++m) sreg[m] = tile[(j >> 1) * FSTG_CASE + 2 * (4 * m + kk) + (j & 1)];
        Chains<1> ch;
        init_chains<1>(ch, ct, nf, 64 * b + lane, N);
        // the next block's pencils stay in flight across this block's chunks
        if (b + 1 < NB) load_block(b + 1, raw);
        wave_sync();  // operand reads done before the recurrence writes the tile
#pragma unroll
        for (int chk = 0; chk < NCHK; ++chk) {
            run_chains_fold<1, 0, 1>(ch, tile, TS, lane, cf, RP, chk * 16);
            wave_sync();
            v4d a1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int m = 0; m < 16; ++m) {
                const double a = tile[(lane & 15) * TS + 4 * m + kk];
                if (m & 1) a1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, sreg[m], a1, 0, 0, 0);
                else acc[chk] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, sreg[m], acc[chk], 0, 0, 0);
            }
            acc[chk] += a1;
            wave_sync();  // tile reads done before the next chunk's recurrence writes
        }
    }
    if (!lc.active) return;
    // this lane's rows r = 16 chk + 4 v + orow: degrees M + orow + 4 (4 chk + v),
    // all of the parity of orow (norm * sign is exact, sign = +-1)
    int orow = lane >> 4;
    asm volatile("" : "+v"(orow));  // the walk's set-up stays here, out of the loop's registers
    const double ns = norm * ((orow & 1) ? lc.sgn_odd : lc.sgn);
    double* cdst = reinterpret_cast<double*>(coeffs) + (j & 1);
    IdxWalk iw(M + orow, lc.Mc, lc.Mpc);
#pragma unroll
    for (int chk = 0; chk < NCHK; ++chk)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            if (chk * 16 + 4 * v + orow < R) cdst[2 * iw.f] = ns * acc[chk][v];
            iw.step();
        }
}

template <int NB, int NCHK>
__device__ __forceinline__ void fwd_nb_dispatch(int nch, const PlanTables& t, const double2* __restrict__ G,
                                                double2* __restrict__ coeffs, double norm, int M, int Mp, double* sm)
{
    if constexpr (NCHK <= 2 * NB) {  // B / 16 = 2 NB chunks at most
        if (nch == NCHK) fwd_nb_pair<NB, NCHK>(t, G, coeffs, norm, M, Mp, sm);
        else fwd_nb_dispatch<NB, NCHK + 1>(nch, t, G, coeffs, norm, M, Mp, sm);
    }
}

template <int NB, int WPE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WPE))) void k_dwt_fwd_f64_nb(PlanTables t, const double2* __restrict__ G,
                                                       double2* __restrict__ coeffs, double norm)
{
    extern __shared__ __align__(16) double sm[];
    const int2 pr = t.xpairs[blockIdx.x];
    const int nch = (t.B - pr.x + 15) >> 4;
    fwd_nb_dispatch<NB, 1>(nch, t, G, coeffs, norm, pr.x, pr.y, sm);
}

// LDS bytes of the node-blocked forward: tile / staging region + step coefficients.
static size_t fwd_nb_smem(int B)
{
    const int ts = fwd_tile_stride(1);
    const int wreg = 16 * ts > FSTG ? 16 * ts : FSTG;
    return (size_t)(wreg + 3 * (((B + 15) / 16) * 16)) * sizeof(double);
}

// G[p][q][k] <- grid(row p, col q, slice k) for p == B or q == B.
// The (2N - 1) N stale Nyquist entries -- spectral row p = B and column q = B
// of every slice -- either read from the caller's grid or, when compact is
// set, from a packed copy [e][k] (e < N: (p, q) = (B, e); e >= N: q = B,
// p = e - N, skipping p = B) that a host caller uploads instead of the grid.
__device__ __forceinline__ void nyquist_gather_one(const double2* __restrict__ grid, double2* __restrict__ G, int B,
                                                   const double2* __restrict__ compact, long idx)
{
    const int N = 2 * B;
    const long total = (long)(2 * N - 1) * N;
    if (idx >= total) return;
    int e = (int)(idx / N), k = (int)(idx % N);
    int p, q;
    if (e < N) { p = B; q = e; }
    else { q = B; p = e - N; if (p >= B) ++p; }
    G[((long)p * N + q) * N + k] = compact ? compact[idx] : grid[(long)N * N * k + (long)N * q + p];
}

// Inverse DWT: every wave owns its nodes outright (no reduction).  Workgroups
// past the last pair copy the stale Nyquist entries (nq_grid or nq_compact
// set): no pair writes those entries, and in the kernel's tail the copy costs
// less than a kernel of its own.
template <int NCH, int WPP, bool SMALL>
__global__ __launch_bounds__(64 * WPP) __attribute__((amdgpu_waves_per_eu(2))) void k_dwt_inv_f64(PlanTables t, const double2* __restrict__ coeffs,
                                                          double2* __restrict__ G, double norm,
                                                          const double2* __restrict__ nq_grid,
                                                          const double2* __restrict__ nq_compact)
{
    if ((int)blockIdx.x >= t.npairs) {
        nyquist_gather_one(nq_grid, G, t.B, nq_compact, (long)(blockIdx.x - t.npairs) * (64 * WPP) + threadIdx.x);
        return;
    }
    extern __shared__ __align__(16) double sm[];
    constexpr int TS = tile_stride(NCH);
    constexpr int WREG = inv_wave_region(NCH);  // a wave's tile, reused as its output staging
    const int B = t.B, N = t.N;
    const int2 pr = t.xpairs[blockIdx.x];
    const int M = pr.x, Mp = pr.y;
    const int R = B - M;
    const int nch = (R + 15) >> 4;
    const int RP = nch * 16;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int nkb = SMALL ? ((N + 15) >> 4) : 4;  // 16-node blocks per chain

    double* tile = sm + w * WREG;
    double* cf = sm + WPP * WREG;

    const int kw = w * 64 * NCH;
    ChainTables<NCH> ct;
    load_chain_tables<NCH>(ct, t, kw + lane, false, M, Mp);
    const double nf = t.normfac[M * B + Mp];

    // B operand rows: lane (j + 16 b) holds n' * sign * coeff(l = M + r0 + 4s + b, case j/2).(re|im)
    const int j = lane & 15, bsel = lane >> 4;
    const LaneCase lc = lane_case(j >> 1, M, Mp, N, nullptr);
    const double* csrc = reinterpret_cast<const double*>(coeffs) + (j & 1);
    // Gathers of chunk chk: the raw values and their factor n' * sign (0 past R
    // and for inactive cases; coefficient 0 stands in for the load).  The
    // loads are unconditional and scaled only at use: a branch around them, or
    // an immediate use, makes the compiler wait for each load right away.
    auto load_b = [&](int chk, double* raw, double* fac) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int r = chk * 16 + 4 * s + bsel;
            const bool live = lc.active && r < R;
            const int ci = live ? 2 * coeff_index32(M + r, lc.Mc, lc.Mpc) : 0;
            fac[s] = live ? norm * ((r & 1) ? lc.sgn_odd : lc.sgn) : 0.0;
            raw[s] = csrc[ci];
        }
    };
    double rcur[4], fcur[4], rnext[4], fnext[4];
    load_b(0, rcur, fcur);

    for (int r = tid; r < RP; r += 64 * WPP) recurrence_coeffs(M, Mp, r, R, cf, RP);
    Chains<NCH> ch;
    init_chains<NCH>(ch, ct, nf, kw + lane, N);

    v4d acc[NCH * 4];
#pragma unroll
    for (int i = 0; i < NCH * 4; ++i) acc[i] = (v4d){0.0, 0.0, 0.0, 0.0};
    __syncthreads();  // the step-coefficient table is complete; from here on a
                      // wave touches only its own tile, so it syncs with itself

    for (int chk = 0; chk < nch; ++chk) {
        load_b(chk + 1 < nch ? chk + 1 : chk, rnext, fnext);
        wave_sync();
        run_chains<NCH, 0, NCH>(ch, tile, TS, lane, cf, RP, chk * 16);
        wave_sync();
        double bcur[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) bcur[s] = rcur[s] * fcur[s];
        // 4-degree steps of this chunk that hold a degree below R (uniform).
        // The steps past it only add products with a zero coefficient factor;
        // skipping them saves 12 % of the MFMAs at B = 128.
        const int nst = min(4, (R - chk * 16 + 3) >> 2);
        // all of the chunk's A operands first, then the MFMAs: left alone the
        // compiler reads one operand pair per MFMA pair and waits on each read
        double a[4][NCH][4];
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int c = 0; c < NCH; ++c)
#pragma unroll
                for (int kb = 0; kb < 4; ++kb)
                    a[s][c][kb] = (!SMALL || kb < nkb) ? tile[(4 * s + bsel) * TS + 64 * c + 16 * kb + j] : 0.0;
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            if (s > 0 && s >= nst) continue;
#pragma unroll
            for (int c = 0; c < NCH; ++c)
#pragma unroll
                for (int kb = 0; kb < 4; ++kb)
                    if (!SMALL || kb < nkb)
                        acc[4 * c + kb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s][c][kb], bcur[s], acc[4 * c + kb], 0, 0, 0);
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            rcur[s] = rnext[s];
            fcur[s] = fnext[s];
        }
    }

    // stage each chain's 64 x 16 block and write each case's beta-vector run
    long cvec[8];
    int cflip = 0, cact = 0;
#pragma unroll
    for (int cs = 0; cs < 8; ++cs) {
        const CaseInfo co = case_info(cs, M, Mp, N);
        cvec[cs] = co.vec;
        cflip |= co.flip << cs;
        cact |= co.active << cs;
    }
    double* stage = tile;  // 64 rows x OSTR fit in the wave's region
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        wave_sync();
#pragma unroll
        for (int kb = 0; kb < 4; ++kb)
            if (!SMALL || kb < nkb)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    stage[(16 * kb + mfma_out_row(lane, v)) * OSTR + mfma_out_col(lane)] = acc[4 * c + kb][v];
        wave_sync();
        const int k = kw + 64 * c + lane;
        if (k < N) {
#pragma unroll
            for (int cs = 0; cs < 8; ++cs) {
                if (!((cact >> cs) & 1)) continue;
                double2 val;
                val.x = stage[lane * OSTR + 2 * cs];
                val.y = stage[lane * OSTR + 2 * cs + 1];
                G[cvec[cs] + (((cflip >> cs) & 1) ? N - 1 - k : k)] = val;
            }
        }
    }
}

// Inverse DWT on the vector ALU.  In the inverse, lane k owns node k's
// outputs outright: for each degree l, out_j(k) += d_l(k) * c_l,j over the 16
// right-hand sides j (8 cases x re/im), and c_l,j is the same for every lane.
// So the recurrence value never leaves its register: each FMA takes c_l,j
// broadcast from lane j of its 16-lane row (DPP row_newbcast), no LDS tile,
// no operand shuffles, no per-chunk syncs.  The FP64 vector and matrix peaks
// of the device are equal, so the MFMA form has no arithmetic advantage here.
//
// Coefficients: every 16-lane row of ring[rr] holds the 16 values of degree
// r = 16 chk + rr (lane j: case j/2, re/im j&1), so the four rows load the
// same 128 bytes.  Slot rr is refilled with the next chunk's degree right
// after its last use: each load has a whole chunk of FMAs to arrive.
// Signs and the norm are per lane (case) and per degree parity: the d(pi -
// beta) cases 4-7 carry (-1)^(l-M).  Workgroups past the last pair copy the
// stale Nyquist entries, as in k_dwt_inv_f64.
// QS = 1: the WPP waves of a pair split its nodes.  QS = 2 (WPP = 2): both
// waves run all NCH x 64 nodes and each takes 8 of the 16 right-hand sides,
// which halves the accumulators (one more wave per SIMD) at the price of
// running the recurrence twice.
template <int NCH, int WPP, bool DPPFMA, int QS>
__global__ __launch_bounds__(64 * WPP) void k_dwt_inv_f64_valu(PlanTables t, const double2* __restrict__ coeffs,
                                                               double2* __restrict__ G, double norm,
                                                               const double2* __restrict__ nq_grid,
                                                               const double2* __restrict__ nq_compact)
{
    if ((int)blockIdx.x >= t.npairs) {
        nyquist_gather_one(nq_grid, G, t.B, nq_compact, (long)(blockIdx.x - t.npairs) * (64 * WPP) + threadIdx.x);
        return;
    }
    extern __shared__ __align__(16) double sm[];
    const int B = t.B, N = t.N;
    const int2 pr = t.xpairs[blockIdx.x];
    const int M = pr.x, Mp = pr.y;
    const int R = B - M;
    const int nch = (R + 15) >> 4;
    const int RP = nch * 16;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    constexpr int QW = 16 / QS;               // right-hand sides per wave
    const int kw = QS == 1 ? w * 64 * NCH : 0;
    const int q0 = QS == 1 ? 0 : QW * w;      // this wave's first right-hand side
    double* cf = sm;

    ChainTables<NCH> ct;
    load_chain_tables<NCH>(ct, t, kw + lane, false, M, Mp);
    const double nf = t.normfac[M * B + Mp];

    // this lane's right-hand side j = lane % 16: case j/2, part j&1
    const int j = lane & 15;
    const LaneCase lc = lane_case(j >> 1, M, Mp, N, nullptr);
    const double* csrc = reinterpret_cast<const double*>(coeffs) + (j & 1);
    const double fac_even = lc.active ? norm * lc.sgn : 0.0;
    const double fac_odd = lc.active ? norm * lc.sgn_odd : 0.0;
    auto cidx = [&](int r) { return r < R ? 2 * coeff_index32(M + r, lc.Mc, lc.Mpc) : 0; };
    double ring[16];
#pragma unroll
    for (int rr = 0; rr < 16; ++rr) ring[rr] = csrc[cidx(rr)];

    for (int r = tid; r < RP; r += 64 * WPP) recurrence_coeffs(M, Mp, r, R, cf, RP);
    Chains<NCH> ch;
    init_chains<NCH>(ch, ct, nf, kw + lane, N);

    double acc[NCH][QW];
#pragma unroll
    for (int c = 0; c < NCH; ++c)
#pragma unroll
        for (int q = 0; q < QW; ++q) acc[c][q] = 0.0;
    __syncthreads();  // the step-coefficient table is complete

    // acc[c][q - Q0] += c_q * d(c) for this wave's right-hand sides q (Q0 is a
    // constant, so every broadcast lane is an immediate)
    auto contract = [&](auto q0c, double cv) {
        constexpr int Q0 = decltype(q0c)::value;
        if (DPPFMA) {
#pragma unroll
            for (int q = Q0; q < Q0 + QW; ++q)
#pragma unroll
                for (int c = 0; c < NCH; ++c) {
                    if (q == Q0 && c == 0) fmac_bcast<true>(q, acc[c][q - Q0], cv, ch.cur[c]);
                    else fmac_bcast<false>(q, acc[c][q - Q0], cv, ch.cur[c]);
                }
        } else {
#pragma unroll
            for (int q = Q0; q < Q0 + QW; ++q) {
                const double cq = row_bcast(cv, q);
#pragma unroll
                for (int c = 0; c < NCH; ++c) acc[c][q - Q0] = fma(ch.cur[c], cq, acc[c][q - Q0]);
            }
        }
    };

    for (int chk = 0; chk < nch; ++chk) {
        const int rl = chk * 16 + (lane & 15);
        const double vf1 = cf[rl], vf1f2 = cf[RP + rl], vc1 = cf[2 * RP + rl];
        const int rn = (chk + 1) * 16;  // first degree of the next chunk
#pragma unroll
        for (int rr = 0; rr < 16; ++rr) {
            if (chk * 16 + rr < R) {  // uniform
                const double cv = ring[rr] * ((rr & 1) ? fac_odd : fac_even);
                const double f1 = row_bcast(vf1, rr), f1f2 = row_bcast(vf1f2, rr), c1 = row_bcast(vc1, rr);
                if (q0 == 0) contract(std::integral_constant<int, 0>(), cv);
                else contract(std::integral_constant<int, QW>(), cv);  // QS = 2, second wave
#pragma unroll
                for (int c = 0; c < NCH; ++c) {
                    const double u = fma(f1, ch.cb[c], f1f2);
                    const double dn = fma(ch.cur[c], u, c1 * ch.prev[c]);
                    ch.prev[c] = ch.cur[c];
                    ch.cur[c] = dn;
                }
            }
            ring[rr] = csrc[cidx(rn + rr)];
        }
    }

    // each lane writes its nodes' values straight into the 8 pencils
    // (1 KB contiguous per case and chain, reversed for cases 4-7).  Up to
    // 2B = 256 the stores stay ordinary: the next FFT pass finds part of the
    // 268 MB pencil buffer in the Infinity Cache.  At 2B = 512 (2.1 GB) it
    // cannot, and the stores are non-temporal.
    const bool nts = N > 256;
#pragma unroll
    for (int cl = 0; cl < QW / 2; ++cl) {
        const int cs = q0 / 2 + cl;
        const CaseInfo co = case_info(cs, M, Mp, N);
        if (!co.active) continue;
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            const int k = kw + 64 * c + lane;
            if (k < N) {
                d2v val;
                val.x = acc[c][2 * cl];
                val.y = acc[c][2 * cl + 1];
                d2v* dst = reinterpret_cast<d2v*>(G + co.vec + (co.flip ? N - 1 - k : k));
                if (nts) __builtin_nontemporal_store(val, dst);
                else *dst = val;
            }
        }
    }
}

// The inverse of k_dwt_inv_f64_valu with its per-degree overhead moved out of
// the degree loop:
//   - the coefficients of a chunk's 16 degrees (16 x 16 values, norm and sign
//     applied) are loaded one chunk ahead, four per lane, and staged in the
//     wave's LDS slice; each degree reads its 16 values back with one
//     ds_read_b64 (16-lane rows read the same 128 bytes), so the 16 index
//     computations per lane and chunk become 4, and no ring of 16 registers
//     is held;
//   - the step coefficients sit interleaved (f1, f1 f2, c1, 0) per degree and
//     every lane reads them at one address (an LDS broadcast) instead of a
//     DPP row broadcast per value;
//   - full chunks run unrolled with no bounds test; the last R % 16 degrees
//     of a pair run one at a time (no padded degree is computed).
// Results equal k_dwt_inv_f64_valu's: the same operations in the same order.
// With RESC the chains run the plan's rescaled recurrence instead (e_{r+1} =
// (a_r cb + b_r) e_r - e_{r-1}, d_r = P_r e_r, the cf4 rows (a, b, P, 0)):
// two FMAs per chain and degree where the reference's form takes three, P_r
// folded into the staged coefficients, and results equal to rounding.
// Measured (interleaved A/B, same rounding-level results): 2B = 512 inverse
// 1.092 -> 1.063 ms; 2B = 256 92.4 -> 92.9 us (the table costs more than the
// instruction saves there), so only 2B = 512 runs it.
template <int NCH, int WPP, bool RESC>
__global__ __launch_bounds__(64 * WPP) void k_dwt_inv_f64_staged(PlanTables t, const double2* __restrict__ coeffs,
                                                                 double2* __restrict__ G, double norm,
                                                                 const double2* __restrict__ nq_grid,
                                                                 const double2* __restrict__ nq_compact)
{
    if ((int)blockIdx.x >= t.npairs) {
        nyquist_gather_one(nq_grid, G, t.B, nq_compact, (long)(blockIdx.x - t.npairs) * (64 * WPP) + threadIdx.x);
        return;
    }
    extern __shared__ __align__(16) double sm[];
    const int B = t.B, N = t.N;
    const int2 pr = t.xpairs[blockIdx.x];
    const int M = pr.x, Mp = pr.y;
    const int R = B - M;
    const int nch = (R + 15) >> 4;
    const int RP = nch * 16;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int kw = w * 64 * NCH;
    double* cf4 = sm;                            // RP x (f1, f1 f2, c1, 0)
    double* stg = sm + 4 * RP + 256 * w;         // this wave's 16 degrees x 16 values

    ChainTables<NCH> ct;
    load_chain_tables<NCH>(ct, t, kw + lane, false, M, Mp);
    const double nf = t.normfac[M * B + Mp];

    // the right-hand side this lane broadcasts (j = lane % 16: case j/2,
    // part j&1) and the staging role: lane L loads degrees L/16 + 4i of a
    // chunk, whose parity is that of L/16
    const int j = lane & 15, dl = lane >> 4;
    const LaneCase lc = lane_case(j >> 1, M, Mp, N, nullptr);
    const double* csrc = reinterpret_cast<const double*>(coeffs) + (j & 1);
    const double fac = lc.active ? norm * ((dl & 1) ? lc.sgn_odd : lc.sgn) : 0.0;
    // the lane's loads walk degrees M + dl + 4 t in order (t = 4 chunk + i)
    IdxWalk iw(M + dl, lc.Mc, lc.Mpc);
    auto cidx = [&](int r) {
        const int x = r < R ? 2 * iw.f : 0;
        iw.step();
        return x;
    };
    double nxt[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) nxt[i] = csrc[cidx(dl + 4 * i)];

    if constexpr (RESC) rescaled_coeffs4(t, M, Mp, RP, cf4, tid, 64 * WPP);
    else recurrence_coeffs4(M, Mp, R, RP, cf4, tid, 64 * WPP);
    Chains<NCH> ch;
    init_chains<NCH>(ch, ct, nf, kw + lane, N);

    double acc[NCH][16];
#pragma unroll
    for (int c = 0; c < NCH; ++c)
#pragma unroll
        for (int q = 0; q < 16; ++q) acc[c][q] = 0.0;
    __syncthreads();  // the step-coefficient table is complete

    // one degree: acc[c][q] += c_q * d(c) (RESC: (P_r c_q) * e(c)) over the
    // 16 right-hand sides, then the chains step to the next degree with the
    // step coefficients (x, y, z) = (f1, f1 f2, c1) (RESC: (a, b, -))
    auto degree = [&](double cv, double x, double y, double z) {
#pragma unroll
        for (int q = 0; q < 16; ++q)
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                if (q == 0 && c == 0) fmac_bcast<true>(q, acc[c][q], cv, ch.cur[c]);
                else fmac_bcast<false>(q, acc[c][q], cv, ch.cur[c]);
            }
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            const double u = fma(x, ch.cb[c], y);
            const double dn = RESC ? fma(ch.cur[c], u, -ch.prev[c]) : fma(ch.cur[c], u, z * ch.prev[c]);
            ch.prev[c] = ch.cur[c];
            ch.cur[c] = dn;
        }
    };
    // stage a chunk's values (loaded a chunk ago; RESC: each times its
    // degree's P_r); the wave's LDS operations run in order, so the writes
    // land after the previous chunk's reads and before this chunk's
    auto stage = [&](int chk) {
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int i = 0; i < 4; ++i)
            stg[lane + 64 * i] = RESC ? nxt[i] * (fac * cf4[4 * (16 * chk + dl + 4 * i) + 2]) : nxt[i] * fac;
        __builtin_amdgcn_wave_barrier();
    };

    // full chunks of 16 degrees, unrolled, each degree's LDS values read one
    // degree ahead
    const int nfull = R >> 4, rem = R & 15;
    for (int chk = 0; chk < nfull; ++chk) {
        stage(chk);
        const int rn = (chk + 1) * 16;
#pragma unroll
        for (int i = 0; i < 4; ++i) nxt[i] = csrc[cidx(rn + dl + 4 * i)];
        const double* cfc = cf4 + 4 * 16 * chk;
        double cvn = stg[j], c1n = RESC ? 0.0 : cfc[2];
        double2 s01n = *reinterpret_cast<const double2*>(cfc);
#pragma unroll
        for (int rr = 0; rr < 16; ++rr) {
            const double cv = cvn, c1 = c1n;
            const double2 s01 = s01n;
            if (rr < 15) {
                cvn = stg[16 * (rr + 1) + j];
                s01n = *reinterpret_cast<const double2*>(cfc + 4 * (rr + 1));
                if (!RESC) c1n = cfc[4 * (rr + 1) + 2];
            }
            degree(cv, s01.x, s01.y, c1);
        }
    }
    // the last rem < 16 degrees one at a time (no padded degrees run)
    if (rem) {
        stage(nfull);
        const double* cfc = cf4 + 4 * 16 * nfull;
#pragma nounroll
        for (int rr = 0; rr < rem; ++rr) {
            const double2 s01 = *reinterpret_cast<const double2*>(cfc + 4 * rr);
            degree(stg[16 * rr + j], s01.x, s01.y, RESC ? 0.0 : cfc[4 * rr + 2]);
        }
    }

    // each lane writes its nodes' values straight into the 8 pencils, as in
    // k_dwt_inv_f64_valu (non-temporal only at 2B = 512)
    const bool nts = N > 256;
#pragma unroll
    for (int cs = 0; cs < 8; ++cs) {
        const CaseInfo co = case_info(cs, M, Mp, N);
        if (!co.active) continue;
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            const int k = kw + 64 * c + lane;
            if (k < N) {
                d2v val;
                val.x = acc[c][2 * cs];
                val.y = acc[c][2 * cs + 1];
                d2v* dst = reinterpret_cast<d2v*>(G + co.vec + (co.flip ? N - 1 - k : k));
                if (nts) __builtin_nontemporal_store(val, dst);
                else *dst = val;
            }
        }
    }
}

struct DwtShape {
    int nch, nh, wpp;
};

// Chain-parallel kernels: grids up to 256 nodes.  Default forward
// shape: one wave per pair, chains two at a time through a 16 KB tile so two
// waves fit per SIMD; inverse: two chains per lane and one wave per 128 nodes
// (its waves need no reduction).  Other variants (tuning knob "dwt_variant",
// forward and inverse alike) split the nodes over more waves per pair.
static DwtShape dwt_shape(int N, bool fwd, int variant)
{
    if (N <= 64) return {1, 1, 1};
    if (N <= 128) return variant == 1 ? DwtShape{1, 1, 2} : DwtShape{2, 2, 1};
    if (fwd) {
        switch (variant) {
        case 1: return {2, 2, 2};
        case 2: return {2, 1, 2};
        case 3: return {1, 1, 4};
        default: return {4, 2, 1};
        }
    }
    switch (variant) {
    case 1: return {1, 1, 4};
    default: return {2, 2, 2};
    }
}

static size_t dwt_smem(int B, int N, bool fwd, int variant)
{
    DwtShape s = dwt_shape(N, fwd, variant);
    int RP = ((B + 15) / 16) * 16;
    size_t d = (size_t)s.wpp * (fwd ? fwd_wave_region(s.nh) : inv_wave_region(s.nch)) + 3 * RP;
    if (fwd && s.wpp > 1) d += (size_t)s.wpp * 256;
    return d * sizeof(double);
}

hipError_t launch_dwt_fwd_f64(const PlanTables& t, const double2* G, double2* coeffs, double norm, hipStream_t st,
                              int variant)
{
    if (t.N > 512) return hipErrorInvalidValue;
    if (t.N == 512) {
        // every variant: the node-blocked kernel, 8 blocks of 64 nodes and up
        // to 16 chunk accumulators per lane (248 VGPRs, two waves per SIMD).
        // Measured at B = 256: 1.25 ms against 1.75 ms for the chain-parallel
        // four-chains-by-two-waves kernel that served 2B = 512 before.
        return pfs::launch((k_dwt_fwd_f64_nb<8, 2>), dim3(t.npairs), dim3(64), fwd_nb_smem(t.B), st, t, G, coeffs, norm);
    }
    // Default at 2B = 128 and 256: the node-blocked kernel, held to three waves
    // per SIMD at 2B = 256 (168 VGPRs, 52 bytes of spills; measured 101.5 us
    // against 105.0 us for the chain-parallel kernel and 109.5 us at two waves)
    // and two at 2B = 128 (16.8 us against 18.0 and 18.6 us).  Variants 4 and
    // 5 force three or two waves; 6 is the chain-parallel kernel.
    if ((variant == 0 || variant == 4 || variant == 5) && (t.N == 128 || t.N == 256)) {
        const size_t smem = fwd_nb_smem(t.B);
        const dim3 grid(t.npairs), block(64);
        const bool three = variant == 4 || (variant == 0 && t.N == 256);
        if (three)
            return t.N == 128 ? pfs::launch((k_dwt_fwd_f64_nb<2, 3>), grid, block, smem, st, t, G, coeffs, norm)
                              : pfs::launch((k_dwt_fwd_f64_nb<4, 3>), grid, block, smem, st, t, G, coeffs, norm);
        return t.N == 128 ? pfs::launch((k_dwt_fwd_f64_nb<2, 2>), grid, block, smem, st, t, G, coeffs, norm)
                          : pfs::launch((k_dwt_fwd_f64_nb<4, 2>), grid, block, smem, st, t, G, coeffs, norm);
    }
    if (variant >= 4) variant = 0;  // the chain-parallel kernel's default shape
    const DwtShape s = dwt_shape(t.N, true, variant);
    const size_t smem = dwt_smem(t.B, t.N, true, variant);
    if (smem > lds_limit()) return hipErrorInvalidConfiguration;
    dim3 grid(t.npairs), block(64 * s.wpp);
#define PFS_FWD(NC, NH_, W)                                                                                 \
    else if (s.nch == NC && s.nh == NH_ && s.wpp == W)                                                      \
        return pfs::launch((k_dwt_fwd_f64<NC, NH_, W, false>), grid, block, smem, st, t, G, coeffs, norm);
    if (t.N < 64) return pfs::launch((k_dwt_fwd_f64<1, 1, 1, true>), grid, block, smem, st, t, G, coeffs, norm);
    PFS_FWD(1, 1, 1)
    PFS_FWD(2, 2, 1)
    PFS_FWD(4, 2, 1)
    PFS_FWD(1, 1, 2)
    PFS_FWD(2, 2, 2)
    PFS_FWD(2, 1, 2)
    PFS_FWD(1, 1, 4)
#undef PFS_FWD
    return hipErrorInvalidValue;
}

hipError_t launch_dwt_inv_f64(const PlanTables& t, const double2* coeffs, double2* G, double norm, hipStream_t st,
                              int variant, const double2* nq_grid, const double2* nq_compact)
{
    if (t.N > 512) return hipErrorInvalidValue;
    if (t.N > 256) {
        if (nq_grid || nq_compact) return hipErrorInvalidValue;  // gathered by its own kernel there
        // the MFMA inverse serves 2B <= 256; the vector-ALU kernels take 2B = 512
        if (!((variant >> 8) & 3)) variant = (variant & 0xff) | (15 << 8);
    }
    // bits 8-9 of variant: 0 = the MFMA kernel (shape from the low bits),
    // 1..3 = the vector-ALU kernel with 2 chains x 2 waves over the nodes,
    // 4 chains x 2 waves over the right-hand sides, or 4 chains x 1 wave per
    // pair (grids of 2B <= 128 use the shape that covers their nodes); bit 10:
    // its broadcasts folded into the FMAs
    const int kind = (variant >> 8) & 3;
    const bool dppfma = (variant >> 10) & 1;  // the broadcast folded into each FMA
    const bool staged = (variant >> 11) & 1;  // k_dwt_inv_f64_staged (broadcasts folded as well)
    variant &= 0xff;
    if (kind) {
        // 2B = 512: four chains x two waves per pair, whatever the kind
        const int nch = t.N <= 128 ? 1 : t.N > 256 ? 4 : (kind == 1 ? 2 : 4);
        const int wpp = t.N <= 64 ? 1 : t.N <= 128 ? 2 : t.N > 256 ? 2 : (kind == 3 ? 1 : 2);
        const int qs = (t.N > 128 && t.N <= 256 && kind == 2) ? 2 : 1;
        const size_t smem = 3 * (size_t)(((t.B + 15) / 16) * 16) * sizeof(double);
        const long nq_total = (nq_grid || nq_compact) ? (long)(2 * t.N - 1) * t.N : 0;
        const int nq_blocks = (int)((nq_total + 64 * wpp - 1) / (64 * wpp));
        dim3 grid(t.npairs + nq_blocks), block(64 * wpp);
        if (staged) {
            if (qs != 1) return hipErrorInvalidValue;
            const size_t ssmem = (4 * (size_t)(((t.B + 15) / 16) * 16) + 256 * (size_t)wpp) * sizeof(double);
            // 2B = 512 (four chains x two waves): the rescaled recurrence
            if (t.N > 256)
                return nch == 4 && wpp == 2 && t.rab
                           ? pfs::launch((k_dwt_inv_f64_staged<4, 2, true>), grid, block, ssmem, st, t, coeffs, G,
                                         norm, nq_grid, nq_compact)
                           : hipErrorInvalidValue;
#define PFS_INVS(NC, W)                                                                                     \
            if (nch == NC && wpp == W)                                                                      \
                return pfs::launch((k_dwt_inv_f64_staged<NC, W, false>), grid, block, ssmem, st, t, coeffs, G, norm, \
                                   nq_grid, nq_compact);
            PFS_INVS(1, 1)
            PFS_INVS(1, 2)
            PFS_INVS(2, 2)
            PFS_INVS(4, 1)
            PFS_INVS(4, 2)
#undef PFS_INVS
            return hipErrorInvalidValue;
        }
#define PFS_INVV(NC, W, QS)                                                                                 \
        if (nch == NC && wpp == W && qs == QS) {                                                            \
            if (dppfma)                                                                                     \
                return pfs::launch((k_dwt_inv_f64_valu<NC, W, true, QS>), grid, block, smem, st, t, coeffs, G, \
                                   norm, nq_grid, nq_compact);                                              \
            return pfs::launch((k_dwt_inv_f64_valu<NC, W, false, QS>), grid, block, smem, st, t, coeffs, G,    \
                               norm, nq_grid, nq_compact);                                                  \
        }
        PFS_INVV(1, 1, 1)
        PFS_INVV(1, 2, 1)
        PFS_INVV(2, 2, 1)
        PFS_INVV(4, 1, 1)
        PFS_INVV(4, 2, 1)
        PFS_INVV(4, 2, 2)
#undef PFS_INVV
        return hipErrorInvalidValue;
    }
    const DwtShape s = dwt_shape(t.N, false, variant);
    const size_t smem = dwt_smem(t.B, t.N, false, variant);
    if (smem > lds_limit()) return hipErrorInvalidConfiguration;
    const long nq_total = (nq_grid || nq_compact) ? (long)(2 * t.N - 1) * t.N : 0;
    const int nq_blocks = (int)((nq_total + 64 * s.wpp - 1) / (64 * s.wpp));
    dim3 grid(t.npairs + nq_blocks), block(64 * s.wpp);
#define PFS_INV(NC, W)                                                                                      \
    else if (s.nch == NC && s.wpp == W)                                                                     \
        return pfs::launch((k_dwt_inv_f64<NC, W, false>), grid, block, smem, st, t, coeffs, G, norm, nq_grid, nq_compact);
    if (t.N < 64)
        return pfs::launch((k_dwt_inv_f64<1, 1, true>), grid, block, smem, st, t, coeffs, G, norm, nq_grid, nq_compact);
    PFS_INV(1, 1)
    PFS_INV(2, 1)
    PFS_INV(2, 2)
    PFS_INV(1, 2)
    PFS_INV(1, 4)
#undef PFS_INV
    return hipErrorInvalidValue;
}

// The stale Nyquist entries on their own (grids of 2B >= 512, whose inverse
// DWT kernel does not carry them).
__global__ void k_nyquist_gather_f64(const double2* __restrict__ grid, double2* __restrict__ G, int B,
                                     const double2* __restrict__ compact)
{
    nyquist_gather_one(grid, G, B, compact, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

hipError_t launch_nyquist_gather_f64(const double2* grid, double2* G, int B, hipStream_t st, const double2* compact)
{
    const int N = 2 * B;
    long total = (long)(2 * N - 1) * N;
    return pfs::launch(k_nyquist_gather_f64, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, grid, G, B, compact);
}

}  // namespace pfs
