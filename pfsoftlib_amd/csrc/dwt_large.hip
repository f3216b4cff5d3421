// Discrete Wigner transform for grids of 2B >= 512 nodes (B >= 256).
//
// The same algorithm as dwt.hip (on-the-fly FP64 Wigner recurrence, LDS
// tile, FP64 MFMA contraction, 8 symmetric cases per canonical pair; see the
// header there), in the kernel variant that passes the B = 256 parity and
// round-trip tests on MI355X: step coefficients read from LDS per degree,
// four chains per lane and two waves per pair (forward), two chains per lane
// and four waves per pair (inverse).  The variant in dwt.hip with the
// coefficients broadcast through SGPRs is used for 2B <= 256 only: at
// B = 256 it failed with hipErrorLaunchFailure whose cause is not yet found,
// so its large-grid path is switched off.
#include "pfs_internal.h"

namespace pfs {
namespace lg {


constexpr int OSTR = 18;            // output staging row stride (doubles), 64 rows

// Orders one wave's LDS accesses around a point (no other wave is waited for).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
    __builtin_amdgcn_wave_barrier();
}

// Workgroup b handles the pair t.xpairs[b] (an XCD-aware order built on the host).

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

// Per-lane recurrence chains: chain c covers node k = kbase + 64 c.
template <int NCH>
struct Chains {
    double cb[NCH], cur[NCH], prev[NCH];
};

template <int NCH>
__device__ __forceinline__ void init_chains(Chains<NCH>& ch, const PlanTables& t, int M, int Mp, int kbase,
                                            bool weighted)
{
    const double nf = t.normfac[M * t.B + Mp];
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        int k = kbase + 64 * c;
        bool kv = k < t.N;
        ch.cb[c] = kv ? t.cosb[k] : 0.0;
        double b = kv ? nf * ipow(t.sinh[k], M - Mp) * ipow(t.cosh[k], M + Mp) : 0.0;
        if (weighted && kv) b *= t.wts[k];
        ch.cur[c] = b;
        ch.prev[c] = 0.0;
    }
}

// Writes degrees r0 .. r0+15 of chains C0 .. C0+NC-1 into the tile (row rr,
// column 64 (c - C0) + lane) and advances them.  Lane L reads the step
// coefficients of degree r0 + L % 16 once; each step broadcasts its own
// within the 16-lane rows.  Measured at B = 256 against three LDS broadcast
// reads per step: forward 1849 -> 1740 us, inverse 1604 -> 1480 us.
template <int NCH, int C0, int NC>
__device__ __forceinline__ void run_chains(Chains<NCH>& ch, double* tile, int tstr, int lane, const double* cf,
                                           int RP, int r0)
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

// LDS tile row stride for a tile of nc chains: 64 nc nodes + 1 (odd, so the
// 16 rows an MFMA operand read touches fall in distinct banks).
__host__ __device__ constexpr int tile_stride(int nc) { return 64 * nc + 1; }

// Pencil staging of the forward preamble: 8 cases x 64 nodes x (re, im) for
// one chain, case rows padded to 136 doubles (272 words = 16 mod 64) so the
// operand-layout reads spread over all banks.
constexpr int FSTG_CASE = 136;
constexpr int FSTG = 8 * FSTG_CASE;
typedef double d2v __attribute__((ext_vector_type(2)));

// Per-lane view of one case of the pair (hoisted out of the degree loop).
struct LaneCase {
    int Mc, Mpc, active;
    double sgn;      // sign for even rows
    double sgn_odd;  // sign for odd rows (differs for the reversed cases)
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
__global__ __launch_bounds__(64 * WPP) void k_dwt_fwd_f64(PlanTables t, const double2* __restrict__ G,
                                                          double2* __restrict__ coeffs, double norm)
{
    extern __shared__ __align__(16) double sm[];
    constexpr int TS = tile_stride(NH);
    const int B = t.B, N = t.N;
    const int2 pr = t.xpairs[blockIdx.x];
    const int M = pr.x, Mp = pr.y;
    const int R = B - M;
    const int nch = (R + 15) >> 4;
    const int RP = nch * 16;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int nmf = SMALL ? ((N + 3) >> 2) : 16;  // MFMAs per chain and chunk

    double* tile = sm + w * 16 * TS;
    double* red = sm + WPP * 16 * TS;   // WPP * 256 (only when WPP > 1)
    double* cf = red + (WPP > 1 ? WPP * 256 : 0);

    const int kw = w * 64 * NCH;
    const int j = lane & 15, kk = lane >> 4;
    long vec;
    const LaneCase lc = lane_case(j >> 1, M, Mp, N, &vec);
    // B operands: lane (j + 16 kk) holds S[kw + 4m + kk][j], j = 2 case + re/im.
    // With room in the tile, each case's beta-vector is read with whole-wave
    // 16-byte loads (lane = node, 1 KB contiguous per load, reversed for the
    // d(pi - beta) cases), issued before the rest of the preamble, and moved
    // into the operand layout through the tile.
    constexpr bool STAGED = !SMALL && 16 * TS >= FSTG;
    double2 raw[STAGED ? NCH : 1][8];
    int cact = 0;
    if (STAGED) {
#pragma unroll
        for (int cs = 0; cs < 8; ++cs) {
            const CaseInfo ci = case_info(cs, M, Mp, N);
            cact |= ci.active << cs;
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                const int n = min(kw + 64 * c + lane, N - 1);
                const double2* src = G + ci.vec + (ci.flip ? N - 1 - n : n);
                const d2v v = __builtin_nontemporal_load(reinterpret_cast<const d2v*>(src));  // read once
                raw[c][cs].x = v.x;
                raw[c][cs].y = v.y;
            }
        }
    }

    for (int r = tid; r < RP; r += 64 * WPP) recurrence_coeffs(M, Mp, r, R, cf, RP);
    Chains<NCH> ch;
    init_chains<NCH>(ch, t, M, Mp, kw + lane, true);

    double sreg[NCH * 16];
    if (STAGED) {
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
    } else {
        const bool flip = lc.sgn != lc.sgn_odd;
        const double* vsrc = reinterpret_cast<const double*>(G + vec) + (j & 1);
#pragma unroll
        for (int m = 0; m < NCH * 16; ++m) {
            int ks = kw + 4 * m + kk;
            double v = 0.0;
            if (lc.active && ks < N) v = vsrc[2 * (flip ? N - 1 - ks : ks)];
            sreg[m] = v;
        }
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

// Inverse DWT: every wave owns its nodes outright (no reduction).
template <int NCH, int WPP, bool SMALL>
__global__ __launch_bounds__(64 * WPP) void k_dwt_inv_f64(PlanTables t, const double2* __restrict__ coeffs,
                                                          double2* __restrict__ G, double norm)
{
    extern __shared__ __align__(16) double sm[];
    constexpr int TS = tile_stride(NCH);
    const int B = t.B, N = t.N;
    const int2 pr = t.xpairs[blockIdx.x];
    const int M = pr.x, Mp = pr.y;
    const int R = B - M;
    const int nch = (R + 15) >> 4;
    const int RP = nch * 16;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int nkb = SMALL ? ((N + 15) >> 4) : 4;  // 16-node blocks per chain

    double* tile = sm + w * 16 * TS;
    double* cf = sm + WPP * 16 * TS;

    for (int r = tid; r < RP; r += 64 * WPP) recurrence_coeffs(M, Mp, r, R, cf, RP);

    const int kw = w * 64 * NCH;
    Chains<NCH> ch;
    init_chains<NCH>(ch, t, M, Mp, kw + lane, false);

    // B operand rows: lane (j + 16 b) holds n' * sign * coeff(l = M + r0 + 4s + b, case j/2).(re|im)
    const int j = lane & 15, bsel = lane >> 4;
    const LaneCase lc = lane_case(j >> 1, M, Mp, N, nullptr);
    const double* csrc = reinterpret_cast<const double*>(coeffs) + (j & 1);
    // Gathers of chunk chk: raw values and their factor n' * sign (0 past R and
    // for inactive cases, where coefficient 0 stands in); unconditional loads,
    // scaled at use, so they stay in flight across the chunk's work.
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

    v4d acc[NCH * 4];
#pragma unroll
    for (int i = 0; i < NCH * 4; ++i) acc[i] = (v4d){0.0, 0.0, 0.0, 0.0};
    double rcur[4], fcur[4], rnext[4], fnext[4];
    load_b(0, rcur, fcur);
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
        // all of the chunk's A operands first, then the MFMAs
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
    double* stage = tile;  // 64 * OSTR <= 16 * TS for the two-chain shapes used here
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
            for (int cs = 0; cs < 8; ++cs) {
                CaseInfo co = case_info(cs, M, Mp, N);
                if (!co.active) continue;
                double2 val;
                val.x = stage[lane * OSTR + 2 * cs];
                val.y = stage[lane * OSTR + 2 * cs + 1];
                G[co.vec + (co.flip ? N - 1 - k : k)] = val;
            }
        }
    }
}

struct DwtShape {
    int nch, nh, wpp;
};

// Forward: one wave per pair up to 256 nodes, chains two at a time through
// the tile (16 KB) so two waves fit per SIMD; above 256 nodes the nodes are
// split over waves.  Inverse: two chains per wave and one wave per 128 nodes
// (its waves need no reduction; fewer accumulators keep occupancy up).
static DwtShape dwt_shape(int N, bool fwd)
{
    if (N <= 64) return {1, 1, 1};
    if (fwd) {
        if (N <= 128) return {2, 2, 1};
        if (N <= 256) return {4, 2, 1};
        if (N <= 512) return {4, 2, 2};
        return {4, 2, 4};
    }
    if (N <= 128) return {2, 2, 1};
    if (N <= 256) return {2, 2, 2};
    if (N <= 512) return {2, 2, 4};
    return {2, 2, 8};
}

static size_t dwt_smem(int B, int N, bool fwd)
{
    DwtShape s = dwt_shape(N, fwd);
    int RP = ((B + 15) / 16) * 16;
    size_t d = (size_t)s.wpp * 16 * tile_stride(s.nh) + 3 * RP;
    if (fwd && s.wpp > 1) d += (size_t)s.wpp * 256;
    return d * sizeof(double);
}

hipError_t launch_fwd(const PlanTables& t, const double2* G, double2* coeffs, double norm, hipStream_t st)
{
    if (t.N > 1024) return hipErrorInvalidValue;
    DwtShape s = dwt_shape(t.N, true);
    size_t smem = dwt_smem(t.B, t.N, true);
    if (smem > lds_limit()) return hipErrorInvalidConfiguration;
    dim3 grid(t.npairs), block(64 * s.wpp);
    if (t.N < 64) return pfs::launch((k_dwt_fwd_f64<1, 1, 1, true>), grid, block, smem, st, t, G, coeffs, norm);
    else if (s.nch == 1) return pfs::launch((k_dwt_fwd_f64<1, 1, 1, false>), grid, block, smem, st, t, G, coeffs, norm);
    else if (s.nch == 2) return pfs::launch((k_dwt_fwd_f64<2, 2, 1, false>), grid, block, smem, st, t, G, coeffs, norm);
    else if (s.wpp == 1) return pfs::launch((k_dwt_fwd_f64<4, 2, 1, false>), grid, block, smem, st, t, G, coeffs, norm);
    else if (s.wpp == 2) return pfs::launch((k_dwt_fwd_f64<4, 2, 2, false>), grid, block, smem, st, t, G, coeffs, norm);
    else return pfs::launch((k_dwt_fwd_f64<4, 2, 4, false>), grid, block, smem, st, t, G, coeffs, norm);
}

hipError_t launch_inv(const PlanTables& t, const double2* coeffs, double2* G, double norm, hipStream_t st)
{
    if (t.N > 1024) return hipErrorInvalidValue;
    DwtShape s = dwt_shape(t.N, false);
    size_t smem = dwt_smem(t.B, t.N, false);
    if (smem > lds_limit()) return hipErrorInvalidConfiguration;
    dim3 grid(t.npairs), block(64 * s.wpp);
    if (t.N < 64) return pfs::launch((k_dwt_inv_f64<1, 1, true>), grid, block, smem, st, t, coeffs, G, norm);
    else if (s.nch == 1) return pfs::launch((k_dwt_inv_f64<1, 1, false>), grid, block, smem, st, t, coeffs, G, norm);
    else if (s.wpp == 1) return pfs::launch((k_dwt_inv_f64<2, 1, false>), grid, block, smem, st, t, coeffs, G, norm);
    else if (s.wpp == 2) return pfs::launch((k_dwt_inv_f64<2, 2, false>), grid, block, smem, st, t, coeffs, G, norm);
    else if (s.wpp == 4) return pfs::launch((k_dwt_inv_f64<2, 4, false>), grid, block, smem, st, t, coeffs, G, norm);
    else return pfs::launch((k_dwt_inv_f64<2, 8, false>), grid, block, smem, st, t, coeffs, G, norm);
}

}  // namespace lg

hipError_t launch_dwt_fwd_f64_large(const PlanTables& t, const double2* G, double2* coeffs, double norm,
                                    hipStream_t st)
{
    return lg::launch_fwd(t, G, coeffs, norm, st);
}

hipError_t launch_dwt_inv_f64_large(const PlanTables& t, const double2* coeffs, double2* G, double norm,
                                    hipStream_t st)
{
    return lg::launch_inv(t, coeffs, G, norm, st);
}

}  // namespace pfs
