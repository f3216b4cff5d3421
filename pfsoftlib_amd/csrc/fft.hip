// Slice FFT stage: batched length-N line DFTs for the layer-wise 2D FFT.
//
// A 2D FFT of every beta-slice is done as two passes of line transforms, each
// one HBM read + one HBM write of the grid.  A workgroup owns LINES lines:
// it loads them into LDS (coalesced: either whole rows, or LINES adjacent
// elements per n), runs a Stockham autosort FFT (radix 8/4/2 passes, one
// barrier-separated in-place LDS exchange per pass) and stores the lines in
// the other orientation, which is how the slice transposes and the
// beta-fastest pencil layout are produced without extra passes.
// Non-power-of-two N (any even N up to 512) uses a direct O(N^2) DFT per line
// with the same load/store paths.
//
// Contract restated from src/fftw_wrapper.cpp:47-172: unnormalised,
// FFTW_FORWARD = exp(-2 pi i jk/N) (sign -1), FFTW_BACKWARD sign +1.
#include "pfs_internal.h"

namespace pfs {

template <typename T> struct CT;
template <> struct CT<double> { typedef double2 C; };
template <> struct CT<float> { typedef float2 C; };

template <typename C> __device__ __forceinline__ C cadd(C a, C b) { C r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { C r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
// Products feeding sums are written as explicit fused multiply-adds, so every
// kernel that runs a line's DFT performs the same roundings whatever the
// compiler would contract in its context (the kernels that share these
// helpers give bit-identical results).
template <typename C> __device__ __forceinline__ C cmul(C a, C b)
{
    C r;
    r.x = fma(a.x, b.x, -(a.y * b.y));
    r.y = fma(a.x, b.y, a.y * b.x);
    return r;
}
// a * (SIGN * i)
template <int SIGN, typename C> __device__ __forceinline__ C cmul_i(C a)
{
    C r;
    if (SIGN > 0) { r.x = -a.y; r.y = a.x; }
    else          { r.x = a.y;  r.y = -a.x; }
    return r;
}

template <int SIGN, typename C> __device__ __forceinline__ void dft2(C* v)
{
    C t = v[0];
    v[0] = cadd(t, v[1]);
    v[1] = csub(t, v[1]);
}

template <int SIGN, typename C> __device__ __forceinline__ void dft4(C& x0, C& x1, C& x2, C& x3)
{
    C a = cadd(x0, x2), b = csub(x0, x2), c = cadd(x1, x3), d = cmul_i<SIGN>(csub(x1, x3));
    x0 = cadd(a, c);
    x2 = csub(a, c);
    x1 = cadd(b, d);
    x3 = csub(b, d);
}

template <int SIGN, typename C> __device__ __forceinline__ void dft8(C* v)
{
    typedef decltype(v[0].x) R;
    const R h = (R)0.70710678118654752440084436210484904;
    C e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    C o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<SIGN>(e0, e1, e2, e3);
    dft4<SIGN>(o0, o1, o2, o3);
    // o_k *= w^k, w = exp(SIGN 2 pi i / 8): o1 w = (o1 + SIGN i o1) h, o3 w^3 = (-o3 + SIGN i o3) h,
    // each product fused into the sums with e1, e3
    C a1, a3;
    a1.x = o1.x - SIGN * o1.y;
    a1.y = o1.y + SIGN * o1.x;
    C t2 = cmul_i<SIGN>(o2);
    a3.x = -o3.x - SIGN * o3.y;
    a3.y = -o3.y + SIGN * o3.x;
    v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
    v[1].x = fma(a1.x, h, e1.x); v[1].y = fma(a1.y, h, e1.y);
    v[5].x = fma(-a1.x, h, e1.x); v[5].y = fma(-a1.y, h, e1.y);
    v[2] = cadd(e2, t2); v[6] = csub(e2, t2);
    v[3].x = fma(a3.x, h, e3.x); v[3].y = fma(a3.y, h, e3.y);
    v[7].x = fma(-a3.x, h, e3.x); v[7].y = fma(-a3.y, h, e3.y);
}

template <int R, int SIGN, typename C> __device__ __forceinline__ void dftR(C* v)
{
    if (R == 2) dft2<SIGN>(v);
    else if (R == 4) dft4<SIGN>(v[0], v[1], v[2], v[3]);
    else dft8<SIGN>(v);
}

__host__ __device__ constexpr int n_passes(int logn) { return logn / 3 + (logn % 3 ? 1 : 0); }
__host__ __device__ constexpr int radix_at(int logn, int p) { return p < logn / 3 ? 8 : (logn % 3 == 2 ? 4 : 2); }

// One in-place Stockham pass over LINES lines of length N held in LDS.
template <int N, int LINES, int R, int SIGN, typename C>
__device__ __forceinline__ void stockham_pass(C* lds, int ls, int ns, const C* __restrict__ tw)
{
    constexpr int NB = N / R;                 // butterflies per line
    constexpr int TOT = NB * LINES;
    constexpr int PER = (TOT + 255) / 256;
    C v[PER][R];
    int tid = threadIdx.x;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        int b = tid + 256 * q;
        if (TOT % 256 == 0 || b < TOT) {
            int line = b / NB, j = b % NB;
            const C* src = lds + line * ls;
            int k = j % ns;
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = src[j + r * NB];
            if (ns > 1) {
                int step = N / (ns * R);
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    C w = tw[r * k * step];
                    if (SIGN > 0) w.y = -w.y;
                    v[q][r] = cmul(v[q][r], w);
                }
            }
            dftR<R, SIGN>(v[q]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        int b = tid + 256 * q;
        if (TOT % 256 == 0 || b < TOT) {
            int line = b / NB, j = b % NB;
            C* dst = lds + line * ls;
            int k = j % ns;
            int base = (j / ns) * ns * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) dst[base + r * ns] = v[q][r];
        }
    }
    __syncthreads();
}

template <int LOGN, int LINES, int SIGN, typename C>
__device__ __forceinline__ void fft_in_lds(C* lds, int ls, const C* __restrict__ tw)
{
    constexpr int N = 1 << LOGN;
    int ns = 1;
#pragma unroll
    for (int p = 0; p < n_passes(LOGN); ++p) {
        constexpr int dummy = 0;
        (void)dummy;
        const int R = radix_at(LOGN, p);
        if (R == 8) stockham_pass<N, LINES, (LOGN >= 3 ? 8 : 2), SIGN>(lds, ls, ns, tw);
        else if (R == 4) stockham_pass<N, LINES, (LOGN >= 2 ? 4 : 2), SIGN>(lds, ls, ns, tw);
        else stockham_pass<N, LINES, 2, SIGN>(lds, ls, ns, tw);
        ns *= R;
    }
}

// Load LINES lines of the tile into LDS.
template <int KIND, typename C>
__device__ __forceinline__ void load_tile(C* lds, int ls, const C* __restrict__ in, long base, long sv, long sn,
                                          int N, int lines)
{
    int tid = threadIdx.x;
    int tot = lines * N;
    if (KIND == LINES_ROW_IN_TILE_OUT) {
        for (int e = tid; e < tot; e += 256) {
            int line = e / N, n = e - line * N;
            lds[line * ls + n] = in[base + line * sv + (long)n * sn];
        }
    } else {
        for (int e = tid; e < tot; e += 256) {
            int n = e / lines, line = e - n * lines;
            lds[line * ls + n] = in[base + line * sv + (long)n * sn];
        }
    }
}

template <int KIND, typename C, typename T>
__device__ __forceinline__ void store_tile(const C* lds, int ls, C* __restrict__ out, long base, long sv, long sn,
                                           int N, int lines, T scale)
{
    int tid = threadIdx.x;
    int tot = lines * N;
    if (KIND == LINES_ROW_IN_TILE_OUT) {
        for (int e = tid; e < tot; e += 256) {
            int n = e / lines, line = e - n * lines;
            C v = lds[line * ls + n];
            v.x *= scale;
            v.y *= scale;
            out[base + line * sv + (long)n * sn] = v;
        }
    } else {
        for (int e = tid; e < tot; e += 256) {
            int line = e / N, n = e - line * N;
            C v = lds[line * ls + n];
            v.x *= scale;
            v.y *= scale;
            out[base + line * sv + (long)n * sn] = v;
        }
    }
}

// First element of signal sig on either side of a line map.
template <typename C>
__device__ __forceinline__ const C* sig_in(const C* in, const LineMap& m, long sig)
{
    return (m.in_ptrs ? reinterpret_cast<const C*>(m.in_ptrs[sig]) : in + sig * m.in_sig) + m.in_off;
}
template <typename C>
__device__ __forceinline__ C* sig_out(C* out, const LineMap& m, long sig)
{
    return (m.out_ptrs ? reinterpret_cast<C*>(m.out_ptrs[sig]) : out + sig * m.out_sig) + m.out_off;
}

// Grid data lives in device global memory; saying so lets the compiler use
// global (not flat) loads and stores, which count only against vmcnt.  The
// accesses go through clang vector types (the HIP double2 / float2 classes
// have no assignment from another address space).
template <typename C> struct GV;
template <> struct GV<double2> { typedef double T __attribute__((ext_vector_type(2))); };
template <> struct GV<float2> { typedef float T __attribute__((ext_vector_type(2))); };
template <typename C> using gvec = __attribute__((address_space(1))) typename GV<C>::T;
template <typename C> __device__ __forceinline__ C gload(const C* base, unsigned i)
{
    const typename GV<C>::T v = ((const gvec<C>*)base)[i];
    C r;
    r.x = v.x;
    r.y = v.y;
    return r;
}
template <typename C, bool NT = false> __device__ __forceinline__ void gstore(C* base, unsigned i, C r)
{
    typename GV<C>::T v;
    v.x = r.x;
    v.y = r.y;
    if constexpr (NT) __builtin_nontemporal_store(v, (gvec<C>*)base + i);
    else ((gvec<C>*)base)[i] = v;
}
// Grid data is read once per pass with non-temporal (streaming) loads: they
// do not allocate in the caches, so the part of the input that the previous
// kernel's ordinary stores left in the 256 MB Infinity Cache stays there until
// it is read.  Measured on MI355X at 2B = 256 (four passes of a round trip):
// 408 us with ordinary loads and stores, 391 us with both non-temporal, 356 us
// with non-temporal loads and ordinary stores.
template <typename C> __device__ __forceinline__ C gload_nt(const C* base, unsigned i)
{
    const typename GV<C>::T v = __builtin_nontemporal_load((const gvec<C>*)base + i);
    C r;
    r.x = v.x;
    r.y = v.y;
    return r;
}

template <typename T, int LOGN, int LINES, int KIND, int SIGN>
__global__ __launch_bounds__(256) void k_fft_lines(const typename CT<T>::C* __restrict__ in,
                                                   typename CT<T>::C* __restrict__ out, LineMap m,
                                                   const typename CT<T>::C* __restrict__ tw, T scale)
{
    typedef typename CT<T>::C C;
    constexpr int N = 1 << LOGN;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* lds = reinterpret_cast<C*>(smem_raw);
    const int ls = N + 1;
    const long u = blockIdx.y, v0 = (long)blockIdx.x * LINES, sig = blockIdx.z;
    const long ib = u * m.in_su + v0 * m.in_sv;
    const long ob = u * m.out_su + v0 * m.out_sv;
    load_tile<KIND>(lds, ls, sig_in(in, m, sig), ib, m.in_sv, m.in_sn, N, LINES);
    __syncthreads();
    fft_in_lds<LOGN, LINES, SIGN>(lds, ls, tw);
    store_tile<KIND>(lds, ls, sig_out(out, m, sig), ob, m.out_sv, m.out_sn, N, LINES, scale);
}

// Direct DFT for even N that is not a power of two.
template <typename T, int KIND, int SIGN>
__global__ __launch_bounds__(256) void k_dft_lines(const typename CT<T>::C* __restrict__ in,
                                                   typename CT<T>::C* __restrict__ out, LineMap m,
                                                   const typename CT<T>::C* __restrict__ tw, T scale, int N,
                                                   int lines)
{
    typedef typename CT<T>::C C;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* a = reinterpret_cast<C*>(smem_raw);
    const int ls = N + 1;
    C* b = a + lines * ls;
    const long u = blockIdx.y, v0 = (long)blockIdx.x * lines, sig = blockIdx.z;
    const long ib = u * m.in_su + v0 * m.in_sv;
    const long ob = u * m.out_su + v0 * m.out_sv;
    out = sig_out(out, m, sig);
    load_tile<KIND>(a, ls, sig_in(in, m, sig), ib, m.in_sv, m.in_sn, N, lines);
    __syncthreads();
    for (int e = threadIdx.x; e < lines * N; e += 256) {
        int line = e / N, k = e - line * N;
        const C* x = a + line * ls;
        T sr = 0, si = 0;
        int idx = 0;
        for (int j = 0; j < N; ++j) {
            C w = tw[idx];
            if (SIGN > 0) w.y = -w.y;
            sr += x[j].x * w.x - x[j].y * w.y;
            si += x[j].x * w.y + x[j].y * w.x;
            idx += k;
            if (idx >= N) idx -= N;
        }
        C r;
        r.x = sr;
        r.y = si;
        b[line * ls + k] = r;
    }
    __syncthreads();
    store_tile<KIND>(b, ls, out, ob, m.out_sv, m.out_sn, N, lines, scale);
}

// ---------------------------------------------------------------------------
// Register four-step line FFT for N = 64, 128, 256 (TPL = N/16 threads per
// line, 16 complex values per thread).  With n = n1 + TPL n2 and
// k = k2 + 16 k1:
//   step A  (thread n1)  Y[n1][k2] = sum_n2 x[n1 + TPL n2] W_16^(n2 k2)   -- read from HBM
//           Y[n1][k2] *= W_N^(n1 k2)
//   LDS transpose
//   step B  (thread t)   X[k2 + 16 k1] = sum_n1 Y[n1][k2] W_TPL^(n1 k1),  k2 = t + TPL q  -- written to HBM
// One LDS exchange per line instead of one per radix pass.

template <int SIGN, typename C> __device__ __forceinline__ C tw16(C a, int e)
{
    // a * exp(SIGN 2 pi i e / 16), e a compile-time constant after unrolling
    typedef decltype(a.x) R;
    const R c1 = (R)0.92387953251128675613, s1 = (R)0.38268343236508977173, h = (R)0.70710678118654752440;
    R c, s;
    switch (e & 15) {
    case 0: return a;
    case 1: c = c1; s = s1; break;
    case 2: c = h; s = h; break;
    case 3: c = s1; s = c1; break;
    case 4: c = 0; s = 1; break;
    case 5: c = -s1; s = c1; break;
    case 6: c = -h; s = h; break;
    case 7: c = -c1; s = s1; break;
    case 8: c = -1; s = 0; break;
    case 9: c = -c1; s = -s1; break;
    case 10: c = -h; s = -h; break;
    case 11: c = -s1; s = -c1; break;
    case 12: c = 0; s = -1; break;
    case 13: c = s1; s = -c1; break;
    case 14: c = h; s = -h; break;
    default: c = c1; s = -s1; break;
    }
    C w;
    w.x = c;
    w.y = SIGN * s;
    return cmul(a, w);
}

// 16-point DFT in registers; on return X[k] sits in v[(k >> 2) + 4 (k & 3)].
template <int SIGN, typename C> __device__ __forceinline__ void dft16_perm(C* v)
{
#pragma unroll
    for (int n1 = 0; n1 < 4; ++n1) dft4<SIGN>(v[n1], v[n1 + 4], v[n1 + 8], v[n1 + 12]);
    // v[n1 + 4 k2] = Y[n1][k2]; twiddle W16^(n1 k2)
#pragma unroll
    for (int n1 = 1; n1 < 4; ++n1)
#pragma unroll
        for (int k2 = 1; k2 < 4; ++k2) v[n1 + 4 * k2] = tw16<SIGN>(v[n1 + 4 * k2], n1 * k2);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) dft4<SIGN>(v[4 * k2], v[4 * k2 + 1], v[4 * k2 + 2], v[4 * k2 + 3]);
}

template <int Rn> __device__ __forceinline__ constexpr int perm_index(int k)
{
    return Rn == 16 ? ((k >> 2) + 4 * (k & 3)) : k;
}

template <int Rn, int SIGN, typename C> __device__ __forceinline__ void dft_small(C* v)
{
    if (Rn == 16) dft16_perm<SIGN>(v);
    else if (Rn == 8) dft8<SIGN>(v);
    else if (Rn == 4) dft4<SIGN>(v[0], v[1], v[2], v[3]);
    else dft2<SIGN>(v);
}

// Lines per workgroup of the register kernel: the tile-side accesses move
// whole runs of one column element over the workgroup's lines, 256 bytes at
// 16 complex128 or 32 complex64 lines.  Measured on MI355X at 2B = 256: four
// complex128 passes 350.1 us at 8 lines for the tile writes (128-byte runs),
// 347.2 us at 16; complex64 passes, 64 grids per launch, 4-9 % faster at 32
// lines than at 16.
template <typename T, int KIND> constexpr int r16_lines()
{
    return sizeof(T) == 4 ? 32 : 16;
}
// Tile stores of complex64 passes are non-temporal.  Those passes stream many
// grids (8.6 GB for 64 grids at 2B = 256), far beyond the Infinity Cache, and
// measured 3-9 % faster so; the complex64 row stores (inverse passes) and all
// complex128 stores stay ordinary (for one complex128 grid the next kernel
// reads part of the output from the Infinity Cache).
// (The 512-point kernel, k_fft_lines_r16x2, stores non-temporally in every
// pass: at 2B = 512 one complex128 grid is 2.1 GB, eight times the cache.)
template <typename T, int KIND> constexpr bool r16_nt_store()
{
    return sizeof(T) == 4 && KIND == LINES_ROW_IN_TILE_OUT;
}
// y * w for a twiddle w = tw[e] of the forward table exp(-2 pi i e / N),
// conjugated for SIGN > 0.  tw[0] = 1 exactly, so e = 0 needs no branch (a
// branch around the load would make every load wait for the previous one).
template <int SIGN, typename C> __device__ __forceinline__ C twiddle(C y, C w)
{
    if (SIGN > 0) w.y = -w.y;
    return cmul(y, w);
}

// The register kernel exchanges the real parts through LDS, then the
// imaginary parts, in the same buffer: half the LDS of a complex exchange, so
// twice the workgroups fit on a CU (its registers allow four waves per SIMD)
// for two more barriers.
template <typename T> constexpr bool r16_split() { return sizeof(T) == 8; }
// LDS bytes of the register kernel: L lines of 16 rows x (TPL + 1) + 1 values.
template <typename T, int KIND> constexpr size_t r16_lds(int logn)
{
    return (size_t)r16_lines<T, KIND>() * (16 * ((1 << logn) / 16 + 1) + 1) * (r16_split<T>() ? 1 : 2) * sizeof(T);
}
// The complex128 register kernel reads its twiddles from an LDS table the
// workgroup fills once, instead of 16 global loads per thread held in
// registers: 158 -> 84 VGPRs, four waves per SIMD (its LDS allows four
// workgroups per CU).  Measured at B = 128 (interleaved A/B, same results):
// passes 86.2-90.3 -> 84.6-87.8 us, round trip 549 -> 542.5 us.  The
// complex64 passes measured the same either way, and the 512-point kernel
// (160 -> 86-90 VGPRs with the table) slightly slower at B = 256 (round trip
// 5.51 -> 5.52-5.54 ms): both keep register twiddles.
template <typename T> constexpr bool r16_lds_twiddles() { return sizeof(T) == 8; }
template <typename T> constexpr int r16_wpe() { return r16_lds_twiddles<T>() ? 4 : 1; }
// LDS bytes with the twiddle table (N complex values after the exchange
// buffer, 16-byte aligned) where the kernel keeps one.
template <typename T, int KIND> constexpr size_t r16_lds_tw(int logn)
{
    return r16_lds_twiddles<T>()
               ? ((r16_lds<T, KIND>(logn) + 15) & ~(size_t)15) + (size_t)(1 << logn) * 2 * sizeof(T)
               : r16_lds<T, KIND>(logn);
}

template <typename T, int LOGN, int KIND, int SIGN>
__global__ __launch_bounds__((r16_lines<T, KIND>() * (1 << LOGN) / 16)) __attribute__((amdgpu_waves_per_eu(r16_wpe<T>()))) void k_fft_lines_r16(const typename CT<T>::C* __restrict__ in,
                                                                          typename CT<T>::C* __restrict__ out,
                                                                          LineMap m,
                                                                          const typename CT<T>::C* __restrict__ tw,
                                                                          T scale)
{
    typedef typename CT<T>::C C;
    constexpr int N = 1 << LOGN;
    constexpr int TPL = N / 16;          // threads per line; also the step-B DFT size
    constexpr int L = r16_lines<T, KIND>();  // lines per workgroup
    constexpr int YS = TPL + 1;          // LDS row stride (values) for one k2
    constexpr int LS = 16 * YS + 1;      // LDS stride between lines
    extern __shared__ __align__(16) unsigned char smem_r16[];

    const int tid = threadIdx.x;
    const long u = blockIdx.y, v0 = (long)blockIdx.x * L, sig = blockIdx.z;
    const C* src = sig_in(in, m, sig) + u * m.in_su + v0 * m.in_sv;
    C* dst = sig_out(out, m, sig) + u * m.out_su + v0 * m.out_sv;

    // step A: rows read with n1 fastest, tiles with the line fastest
    int la, n1;
    if (KIND == LINES_ROW_IN_TILE_OUT) { n1 = tid % TPL; la = tid / TPL; }
    else { la = tid % L; n1 = tid / L; }
    C v[16];
    {
        // offsets within one signal's grid stay below N^3 <= 2^27: 32-bit
        const unsigned o = (unsigned)(la * m.in_sv + n1 * m.in_sn), step = (unsigned)(TPL * m.in_sn);
#pragma unroll
        for (int n2 = 0; n2 < 16; ++n2) v[n2] = gload_nt(src, o + n2 * step);
    }
    // the twiddles W^(n1 k2) of this thread's 16 outputs: from the workgroup's
    // LDS table [k2][n1] (consecutive n1 in consecutive slots: no bank
    // conflicts), or loaded into registers in flight with the line loads
    C* ltw = reinterpret_cast<C*>(smem_r16 + ((r16_lds<T, KIND>(LOGN) + 15) & ~(size_t)15));
    C w[r16_lds_twiddles<T>() ? 1 : 16];
    if constexpr (r16_lds_twiddles<T>()) {
        for (int e = tid; e < N; e += L * TPL) ltw[e] = gload(tw, (e % TPL) * (e / TPL));
        __syncthreads();
    } else {
#pragma unroll
        for (int k2 = 0; k2 < 16; ++k2) w[k2] = gload(tw, n1 * k2);
    }
    auto wt = [&](int k2) { return r16_lds_twiddles<T>() ? ltw[k2 * TPL + n1] : w[k2 % (r16_lds_twiddles<T>() ? 1 : 16)]; };
    dft16_perm<SIGN>(v);

    // step B: 16 / TPL DFTs of size TPL per thread; row stores with k2 fastest,
    // tile stores with the line fastest
    int lb, t;
    if (KIND == LINES_ROW_IN_TILE_OUT) { lb = tid % L; t = tid / L; }
    else { t = tid % TPL; lb = tid / TPL; }
    constexpr int Q = 16 / TPL;
    C x[Q][TPL];
    if (r16_split<T>()) {
        // real parts go to LDS as they are twiddled; imaginary parts wait in registers
        T* ly = reinterpret_cast<T*>(smem_r16) + la * LS;
        const T* lyb = reinterpret_cast<const T*>(smem_r16) + lb * LS;
        T yi[16];
#pragma unroll
        for (int k2 = 0; k2 < 16; ++k2) {
            C y = twiddle<SIGN>(v[perm_index<16>(k2)], wt(k2));
            ly[k2 * YS + n1] = y.x;
            yi[k2] = y.y;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < Q; ++q)
#pragma unroll
            for (int a = 0; a < TPL; ++a) x[q][a].x = lyb[(t + TPL * q) * YS + a];
        __syncthreads();
#pragma unroll
        for (int k2 = 0; k2 < 16; ++k2) ly[k2 * YS + n1] = yi[k2];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < Q; ++q)
#pragma unroll
            for (int a = 0; a < TPL; ++a) x[q][a].y = lyb[(t + TPL * q) * YS + a];
    } else {
        C* ly = reinterpret_cast<C*>(smem_r16) + la * LS;
        const C* lyb = reinterpret_cast<const C*>(smem_r16) + lb * LS;
#pragma unroll
        for (int k2 = 0; k2 < 16; ++k2) ly[k2 * YS + n1] = twiddle<SIGN>(v[perm_index<16>(k2)], wt(k2));
        __syncthreads();
#pragma unroll
        for (int q = 0; q < Q; ++q)
#pragma unroll
            for (int a = 0; a < TPL; ++a) x[q][a] = lyb[(t + TPL * q) * YS + a];
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int k2 = t + TPL * q;
        dft_small<TPL, SIGN>(x[q]);
        const unsigned o = (unsigned)(lb * m.out_sv + k2 * m.out_sn), step = (unsigned)(16 * m.out_sn);
#pragma unroll
        for (int k1 = 0; k1 < TPL; ++k1) {
            C r = x[q][perm_index<TPL>(k1)];
            r.x *= scale;
            r.y *= scale;
            gstore<C, r16_nt_store<T, KIND>()>(dst, o + k1 * step, r);
        }
    }
}

// complex64 register kernel with 16-byte global accesses.  Each thread carries
// two of k_fft_lines_r16's slots: on the row side two adjacent n1 (step A) or
// two adjacent t (step B), on the tile side two adjacent lines, so every load
// and store moves two neighbouring complex64 values at once (half the address
// work per byte of the 8-byte accesses, which keep the texture addresser of
// the complex64 passes busy 92-95 % of the time).  Every line runs the same
// operations in the same order as in k_fft_lines_r16: the results are equal
// bit for bit.  A workgroup whose signal start on either side is not 16-byte
// aligned moves its values 8 bytes at a time.
template <int LOGN, int KIND, int SIGN>
__global__ __launch_bounds__((r16_lines<float, KIND>() * (1 << LOGN) / 32)) void k_fft_lines_r16p(const float2* __restrict__ in,
                                                                                            float2* __restrict__ out,
                                                                                            LineMap m,
                                                                                            const float2* __restrict__ tw,
                                                                                            float scale)
{
    typedef float2 C;
    typedef float f4 __attribute__((ext_vector_type(4)));
    typedef __attribute__((address_space(1))) f4 gf4;
    constexpr int N = 1 << LOGN;
    constexpr int TPL = N / 16;
    constexpr int L = r16_lines<float, KIND>();
    constexpr int YS = TPL + 1;
    constexpr int LS = 16 * YS + 1;
    constexpr int Q = 16 / TPL;
    constexpr bool ROWIN = KIND == LINES_ROW_IN_TILE_OUT;
    extern __shared__ __align__(16) unsigned char smem_r16p[];
    C* lds = reinterpret_cast<C*>(smem_r16p);

    const int tid = threadIdx.x;
    const long u = blockIdx.y, v0 = (long)blockIdx.x * L, sig = blockIdx.z;
    const C* src = sig_in(in, m, sig) + u * m.in_su + v0 * m.in_sv;
    C* dst = sig_out(out, m, sig) + u * m.out_su + v0 * m.out_sv;
    const bool wide = !(((uintptr_t)src | (uintptr_t)dst) & 15);

    // step A: row side pairs (line la, n1 = 2j, 2j + 1), tile side pairs (lines 2j, 2j + 1, n1)
    const int la0 = ROWIN ? tid / (TPL / 2) : 2 * (tid % (L / 2));
    const int n10 = ROWIN ? 2 * (tid % (TPL / 2)) : tid / (L / 2);
    const int la1 = ROWIN ? la0 : la0 + 1, n11 = ROWIN ? n10 + 1 : n10;
    C v[2][16];
    {
        const unsigned o = (unsigned)(la0 * m.in_sv + n10 * m.in_sn), step = (unsigned)(TPL * m.in_sn);
        if (wide) {
#pragma unroll
            for (int n2 = 0; n2 < 16; ++n2) {
                const f4 x = __builtin_nontemporal_load((const gf4*)(src + o + n2 * step));
                v[0][n2].x = x.x;
                v[0][n2].y = x.y;
                v[1][n2].x = x.z;
                v[1][n2].y = x.w;
            }
        } else {
#pragma unroll
            for (int n2 = 0; n2 < 16; ++n2) {
                v[0][n2] = gload_nt(src, o + n2 * step);
                v[1][n2] = gload_nt(src, o + 1 + n2 * step);
            }
        }
    }
    C w0[16], w1[16];
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) {
        w0[k2] = gload(tw, n10 * k2);
        w1[k2] = ROWIN ? gload(tw, n11 * k2) : w0[k2];
    }
    dft16_perm<SIGN>(v[0]);
    dft16_perm<SIGN>(v[1]);
    {
        C* ly0 = lds + la0 * LS;
        C* ly1 = lds + la1 * LS;
#pragma unroll
        for (int k2 = 0; k2 < 16; ++k2) {
            ly0[k2 * YS + n10] = twiddle<SIGN>(v[0][perm_index<16>(k2)], w0[k2]);
            ly1[k2 * YS + n11] = twiddle<SIGN>(v[1][perm_index<16>(k2)], w1[k2]);
        }
    }
    __syncthreads();

    // step B: tile-side stores pair lines (2j, 2j + 1) at one t, row-side stores pair t = 2j, 2j + 1
    const int lb0 = ROWIN ? 2 * (tid % (L / 2)) : tid / (TPL / 2);
    const int t0 = ROWIN ? tid / (L / 2) : 2 * (tid % (TPL / 2));
    const int lb1 = ROWIN ? lb0 + 1 : lb0, t1 = ROWIN ? t0 : t0 + 1;
    C x[2][Q][TPL];
#pragma unroll
    for (int q = 0; q < Q; ++q)
#pragma unroll
        for (int a = 0; a < TPL; ++a) {
            x[0][q][a] = lds[lb0 * LS + (t0 + TPL * q) * YS + a];
            x[1][q][a] = lds[lb1 * LS + (t1 + TPL * q) * YS + a];
        }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        dft_small<TPL, SIGN>(x[0][q]);
        dft_small<TPL, SIGN>(x[1][q]);
        const int k2 = t0 + TPL * q;
        const unsigned o = (unsigned)(lb0 * m.out_sv + k2 * m.out_sn), step = (unsigned)(16 * m.out_sn);
#pragma unroll
        for (int k1 = 0; k1 < TPL; ++k1) {
            C r0 = x[0][q][perm_index<TPL>(k1)], r1 = x[1][q][perm_index<TPL>(k1)];
            r0.x *= scale;
            r0.y *= scale;
            r1.x *= scale;
            r1.y *= scale;
            if (wide) {
                f4 y;
                y.x = r0.x;
                y.y = r0.y;
                y.z = r1.x;
                y.w = r1.y;
                if constexpr (r16_nt_store<float, KIND>()) __builtin_nontemporal_store(y, (gf4*)(dst + o + k1 * step));
                else *(gf4*)(dst + o + k1 * step) = y;
            } else {
                gstore<C, r16_nt_store<float, KIND>()>(dst, o + k1 * step, r0);
                gstore<C, r16_nt_store<float, KIND>()>(dst, o + 1 + k1 * step, r1);
            }
        }
    }
}

// 512-point lines.  Step A is the register kernel's (32 threads per line, a
// 16-point DFT each, twiddle, exchange through LDS).  Step B splits each of
// the 16 DFTs of size 32 over a pair of lanes: lane half h takes the inputs
// n1 = 2i + h, runs a 16-point DFT, the odd half multiplies by W32^k, and one
// lane exchange gives X[k] = E0 + W E1 (h = 0) and X[k + 16] = E0 - W E1
// (h = 1).  Pairs sit 8 lanes apart for tile stores, 16 apart for row stores.
template <typename T, int KIND, int SIGN>
__global__ __launch_bounds__((r16_lines<T, KIND>() * 32)) void k_fft_lines_r16x2(const typename CT<T>::C* __restrict__ in,
                                                                               typename CT<T>::C* __restrict__ out,
                                                                               LineMap m,
                                                                               const typename CT<T>::C* __restrict__ tw,
                                                                               T scale)
{
    typedef typename CT<T>::C C;
    constexpr int TPL = 32;
    constexpr int L = r16_lines<T, KIND>();
    constexpr int YS = TPL + 1;
    constexpr int LS = 16 * YS + 1;
    extern __shared__ __align__(16) unsigned char smem_r16[];

    const int tid = threadIdx.x;
    const long u = blockIdx.y, v0 = (long)blockIdx.x * L, sig = blockIdx.z;
    const C* src = sig_in(in, m, sig) + u * m.in_su + v0 * m.in_sv;
    C* dst = sig_out(out, m, sig) + u * m.out_su + v0 * m.out_sv;

    // step A
    int la, n1;
    if (KIND == LINES_ROW_IN_TILE_OUT) { n1 = tid % TPL; la = tid / TPL; }
    else { la = tid % L; n1 = tid / L; }
    C v[16];
    {
        const unsigned o = (unsigned)(la * m.in_sv + n1 * m.in_sn), step = (unsigned)(TPL * m.in_sn);
#pragma unroll
        for (int n2 = 0; n2 < 16; ++n2) v[n2] = gload_nt(src, o + n2 * step);
    }
    C w[16];
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) w[k2] = gload(tw, n1 * k2);
    dft16_perm<SIGN>(v);

    // step B lanes: (line, k2, half)
    int lb, k2b, h;
    if (KIND == LINES_ROW_IN_TILE_OUT) { lb = tid % L; const int t = tid / L; k2b = t >> 1; h = t & 1; }
    else { const int t = tid % TPL; lb = tid / TPL; k2b = t & 15; h = t >> 4; }
    constexpr int XM = (KIND == LINES_ROW_IN_TILE_OUT) ? L : 16;  // lane distance to the partner half
    C x[16];
    if (r16_split<T>()) {
        T* ly = reinterpret_cast<T*>(smem_r16) + la * LS;
        const T* lyb = reinterpret_cast<const T*>(smem_r16) + lb * LS + k2b * YS + h;
        T yi[16];
#pragma unroll
        for (int k2 = 0; k2 < 16; ++k2) {
            C y = twiddle<SIGN>(v[perm_index<16>(k2)], w[k2]);
            ly[k2 * YS + n1] = y.x;
            yi[k2] = y.y;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 16; ++i) x[i].x = lyb[2 * i];
        __syncthreads();
#pragma unroll
        for (int k2 = 0; k2 < 16; ++k2) ly[k2 * YS + n1] = yi[k2];
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 16; ++i) x[i].y = lyb[2 * i];
    } else {
        C* ly = reinterpret_cast<C*>(smem_r16) + la * LS;
        const C* lyb = reinterpret_cast<const C*>(smem_r16) + lb * LS + k2b * YS + h;
#pragma unroll
        for (int k2 = 0; k2 < 16; ++k2) ly[k2 * YS + n1] = twiddle<SIGN>(v[perm_index<16>(k2)], w[k2]);
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 16; ++i) x[i] = lyb[2 * i];
    }
    dft16_perm<SIGN>(x);
    const unsigned o = (unsigned)(lb * m.out_sv + (k2b + 256 * h) * m.out_sn), step = (unsigned)(16 * m.out_sn);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        C e = x[perm_index<16>(k)];
        const C wk = gload(tw, 16 * k);
        const C t = twiddle<SIGN>(e, wk);
        // lane half 0 keeps E0[k] and receives W E1[k]; half 1 the reverse
        C mine;
        mine.x = h ? t.x : e.x;
        mine.y = h ? t.y : e.y;
        C theirs;
        theirs.x = __shfl_xor(mine.x, XM);
        theirs.y = __shfl_xor(mine.y, XM);
        C r;
        r.x = (h ? theirs.x - mine.x : mine.x + theirs.x) * scale;
        r.y = (h ? theirs.y - mine.y : mine.y + theirs.y) * scale;
        gstore<C, true>(dst, o + k * step, r);  // 2B = 512: the grid is far beyond the Infinity Cache
    }
}

// The paired complex64 kernel needs the two values a thread moves to be
// neighbours in memory on both sides: unit element stride along the row side
// (n) and along the tile side (lines).
static bool pairs_adjacent(const LineMap& m, int kind)
{
    return kind == LINES_ROW_IN_TILE_OUT ? (m.in_sn == 1 && m.out_sv == 1) : (m.in_sv == 1 && m.out_sn == 1);
}

template <typename T, int LOGN, int KIND, int SIGN>
static hipError_t launch_pow2(const typename CT<T>::C* in, typename CT<T>::C* out, const LineMap& m,
                              const typename CT<T>::C* tw, T scale, int nsig, hipStream_t st, int paired)
{
    constexpr int N = 1 << LOGN;
    if (LOGN == 9) {
        constexpr int L = r16_lines<T, KIND>();
        if (m.nv % L != 0) return hipErrorInvalidValue;
        dim3 grid(m.nv / L, m.nu, nsig);
        const size_t smem = r16_lds<T, KIND>(LOGN);
        if (smem > lds_limit()) return hipErrorInvalidConfiguration;
        return pfs::launch((k_fft_lines_r16x2<T, KIND, SIGN>), grid, dim3(L * 32), smem, st, in, out, m, tw, scale);
    }
    if (LOGN >= 6 && LOGN <= 8) {
        constexpr int L = r16_lines<T, KIND>();
        if (m.nv % L != 0) return hipErrorInvalidValue;
        dim3 grid(m.nv / L, m.nu, nsig);
        const size_t smem = r16_lds<T, KIND>(LOGN);
        if (smem > lds_limit()) return hipErrorInvalidConfiguration;
        if constexpr (sizeof(T) == 4) {
            // wide 2 (the default): the forward-kind passes of 2B <= 128 only.  Measured on
            // MI355X (per 64 signals): forward passes at 2B = 128 421/434 -> 395/402 us; every
            // other pass slower with the paired kernel's 124-144 VGPRs (3-4 waves per SIMD
            // against 7): 2B = 128 inverse 359/388 -> 411/501 us, 2B = 256 forward
            // 3.04/2.97 -> 3.22/2.96 ms, inverse 3.38/3.10 -> 3.45/3.17 ms.
            const bool use = paired == 1 || (paired == 2 && KIND == LINES_ROW_IN_TILE_OUT && LOGN <= 7);
            if (use && pairs_adjacent(m, KIND))
                return pfs::launch((k_fft_lines_r16p<(LOGN >= 6 && LOGN <= 8 ? LOGN : 8), KIND, SIGN>), grid,
                                   dim3(L * N / 32), smem, st, in, out, m, tw, scale);
        }
        const size_t smem_tw = r16_lds_tw<T, KIND>(LOGN);
        if (smem_tw > lds_limit()) return hipErrorInvalidConfiguration;
        return pfs::launch((k_fft_lines_r16<T, (LOGN >= 6 && LOGN <= 8 ? LOGN : 8), KIND, SIGN>), grid,
                           dim3(L * N / 16), smem_tw, st, in, out, m, tw, scale);
    }
    constexpr int LINES = N >= 8 ? (N >= 1024 ? 2 : (N >= 512 ? 4 : 8)) : N;
    size_t smem = (size_t)LINES * (N + 1) * sizeof(typename CT<T>::C);
    if (smem > lds_limit()) return hipErrorInvalidConfiguration;
    if (m.nv % LINES != 0) return hipErrorInvalidValue;
    dim3 grid(m.nv / LINES, m.nu, nsig);
    return pfs::launch((k_fft_lines<T, LOGN, LINES, KIND, SIGN>), grid, dim3(256), smem, st, in, out, m, tw, scale);
}

template <typename T, int KIND, int SIGN>
static hipError_t dispatch_pow2(int logn, const typename CT<T>::C* in, typename CT<T>::C* out, const LineMap& m,
                                const typename CT<T>::C* tw, T scale, int nsig, hipStream_t st, int paired)
{
    switch (logn) {
    case 1: return launch_pow2<T, 1, KIND, SIGN>(in, out, m, tw, scale, nsig, st, paired);
    case 2: return launch_pow2<T, 2, KIND, SIGN>(in, out, m, tw, scale, nsig, st, paired);
    case 3: return launch_pow2<T, 3, KIND, SIGN>(in, out, m, tw, scale, nsig, st, paired);
    case 4: return launch_pow2<T, 4, KIND, SIGN>(in, out, m, tw, scale, nsig, st, paired);
    case 5: return launch_pow2<T, 5, KIND, SIGN>(in, out, m, tw, scale, nsig, st, paired);
    case 6: return launch_pow2<T, 6, KIND, SIGN>(in, out, m, tw, scale, nsig, st, paired);
    case 7: return launch_pow2<T, 7, KIND, SIGN>(in, out, m, tw, scale, nsig, st, paired);
    case 8: return launch_pow2<T, 8, KIND, SIGN>(in, out, m, tw, scale, nsig, st, paired);
    case 9: return launch_pow2<T, 9, KIND, SIGN>(in, out, m, tw, scale, nsig, st, paired);
    case 10: return launch_pow2<T, 10, KIND, SIGN>(in, out, m, tw, scale, nsig, st, paired);
    default: return hipErrorInvalidValue;
    }
}

template <typename T>
static hipError_t launch_lines(const typename CT<T>::C* in, typename CT<T>::C* out, const LineMap& m, int N,
                               int sign, T scale, const typename CT<T>::C* tw, int kind, int nsig, hipStream_t st,
                               int paired = 0)
{
    if (N < 2 || (N & 1) || m.nv < 1 || m.nu < 1 || nsig < 1) return hipErrorInvalidValue;
    int logn = 0;
    while ((1 << logn) < N) ++logn;
    if ((1 << logn) == N) {
        if (logn > 10) return hipErrorInvalidValue;
        if (kind == LINES_ROW_IN_TILE_OUT)
            return sign < 0 ? dispatch_pow2<T, LINES_ROW_IN_TILE_OUT, -1>(logn, in, out, m, tw, scale, nsig, st, paired)
                            : dispatch_pow2<T, LINES_ROW_IN_TILE_OUT, 1>(logn, in, out, m, tw, scale, nsig, st, paired);
        return sign < 0 ? dispatch_pow2<T, LINES_TILE_IN_ROW_OUT, -1>(logn, in, out, m, tw, scale, nsig, st, paired)
                        : dispatch_pow2<T, LINES_TILE_IN_ROW_OUT, 1>(logn, in, out, m, tw, scale, nsig, st, paired);
    }
    if (N > 512) return hipErrorInvalidValue;
    int lines = 0;
    for (int cand : {8, 4, 2})
        if (N % cand == 0 && 2 * (size_t)cand * (N + 1) * sizeof(typename CT<T>::C) <= 65536) {
            lines = cand;
            break;
        }
    if (!lines) return hipErrorInvalidValue;
    size_t smem = 2 * (size_t)lines * (N + 1) * sizeof(typename CT<T>::C);
    if (m.nv % lines != 0) return hipErrorInvalidValue;
    dim3 grid(m.nv / lines, m.nu, nsig);
    if (kind == LINES_ROW_IN_TILE_OUT)
        return sign < 0 ? pfs::launch((k_dft_lines<T, LINES_ROW_IN_TILE_OUT, -1>), grid, dim3(256), smem, st, in, out, m,
                                      tw, scale, N, lines)
                        : pfs::launch((k_dft_lines<T, LINES_ROW_IN_TILE_OUT, 1>), grid, dim3(256), smem, st, in, out, m,
                                      tw, scale, N, lines);
    return sign < 0 ? pfs::launch((k_dft_lines<T, LINES_TILE_IN_ROW_OUT, -1>), grid, dim3(256), smem, st, in, out, m, tw,
                                  scale, N, lines)
                    : pfs::launch((k_dft_lines<T, LINES_TILE_IN_ROW_OUT, 1>), grid, dim3(256), smem, st, in, out, m, tw,
                                  scale, N, lines);
}

size_t lds_limit()
{
    static int cached[64];
    static bool init[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 65536;
    if (!init[dev]) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0) v = 65536;
        cached[dev] = v;
        init[dev] = true;
    }
    return (size_t)cached[dev];
}

// ---------------------------------------------------------------------------
// One-pass slice FFT: a workgroup holds one whole beta-slice in LDS and runs
// both line passes of its 2D FFT there, so the grid crosses HBM once per
// direction instead of twice (complex64, N = 64 and 128: a slice is 32 or
// 128 KB).
//
// Forward, slice k of one signal:
//   pass 1 (lines c, over r):  x[k][c][r]  (HBM, whole rows) -> Y[c][p]
//   LDS transpose:             Y[c][p] -> Y'[p][c]
//   pass 2 (lines p, over c):  Y'[p][c] -> G[p][q][k]  (HBM)
// Inverse, slice k:
//   pass 1 (lines p, over q):  G[p][q][k]  (HBM) -> Z[p][c]
//   LDS transpose:             Z[p][c] -> Z'[c][p]
//   pass 2 (lines c, over p):  Z'[c][p] -> x[k][c][r] * scale  (HBM, whole rows)
// Each line runs the register kernel's four-step algorithm with the same
// operations in the same order (16-point DFT, twiddle W_N^(n1 k2), exchange,
// N/16-point DFTs), so the results equal the two-pass kernels' bit for bit.
//
// The pencil side moves one 8-byte value per (p, q) and slice: a 128-byte
// line of G[p][q][k0..k0+15] is shared by 16 slices.  The workgroups are
// therefore dealt so that the 16 slices of one such k-block run together on
// one XCD (workgroup b runs on XCD b mod 8): their stores meet in that XCD's
// L2 and leave it as whole lines, and their loads (inverse) fetch each line
// once.  Both use ordinary (allocating) accesses for that reason.
template <typename T, int LOGN> struct SliceGeo {
    static constexpr int N = 1 << LOGN;
    static constexpr int TPL = N / 16;            // threads per line
    static constexpr int THREADS = N * TPL;
    static constexpr int YS = TPL + 1;            // exchange: stride between k2 rows
    static constexpr int LS = 16 * YS + 1;        // exchange: stride between lines
    static constexpr int RS = N + 8;              // transpose: row stride
    static constexpr size_t LDS = (size_t)N * (LS > RS ? LS : RS) * sizeof(typename CT<T>::C);
};

// Workgroup b -> (signal, slice): XCD x = b mod 8 takes the k-blocks
// g = 8 i + x, the 16 slices of block g at consecutive local positions.
__device__ __forceinline__ bool slice_of_block(int b, long nblocks, int kblocks, long& sig, int& k)
{
    const int x = b & 7, j = b >> 3;
    const long g = (long)(j >> 4) * 8 + x;
    if (g >= nblocks) return false;
    sig = g / kblocks;
    k = (int)(g % kblocks) * 16 + (j & 15);
    return true;
}

template <typename T, int LOGN, int SIGN>
__global__ __launch_bounds__((SliceGeo<T, LOGN>::THREADS)) void k_fft_slice(const typename CT<T>::C* __restrict__ in,
                                                                       typename CT<T>::C* __restrict__ out,
                                                                       LineMap m,
                                                                       const typename CT<T>::C* __restrict__ tw,
                                                                       T scale, long nblocks)
{
    typedef typename CT<T>::C C;
    typedef SliceGeo<T, LOGN> Geo;
    constexpr int N = Geo::N, TPL = Geo::TPL, YS = Geo::YS, LS = Geo::LS, RS = Geo::RS;
    constexpr int Q = 16 / TPL;
    constexpr bool FWD = SIGN < 0;
    extern __shared__ __align__(16) unsigned char smem_slice[];
    C* S = reinterpret_cast<C*>(smem_slice);

    long sig;
    int k;
    if (!slice_of_block(blockIdx.x, nblocks, N / 16, sig, k)) return;
    const int tid = threadIdx.x;
    const int line = tid / TPL, n = tid % TPL;  // line of either pass; n1 in step A, t in step B
    // forward: x rows on the input side, pencils on the output side; inverse the reverse
    const C* src = sig_in(in, m, sig) + (FWD ? (long)k * N * N : (long)k);
    C* dst = sig_out(out, m, sig) + (FWD ? (long)k : (long)k * N * N);

    C w[16];
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) w[k2] = gload(tw, n * k2);
    C v[16];
    if (FWD) {
        // whole rows x[k][line][n1 + TPL n2]
#pragma unroll
        for (int n2 = 0; n2 < 16; ++n2) v[n2] = gload_nt(src, (unsigned)(line * N + n + TPL * n2));
    } else {
        // G[line][n1 + TPL n2][k]: one value per pencil, shared with the k-block's other slices
#pragma unroll
        for (int n2 = 0; n2 < 16; ++n2) v[n2] = gload(src, (unsigned)((line * N + n + TPL * n2) * N));
    }
    C x[Q][TPL];
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) {
            // step A input of the second pass: row `line` of the transposed slice
#pragma unroll
            for (int n2 = 0; n2 < 16; ++n2) v[n2] = S[line * RS + n + TPL * n2];
            __syncthreads();  // every thread has read its row before the exchange overwrites it
        }
        dft16_perm<SIGN>(v);
#pragma unroll
        for (int k2 = 0; k2 < 16; ++k2) S[line * LS + k2 * YS + n] = twiddle<SIGN>(v[perm_index<16>(k2)], w[k2]);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < Q; ++q)
#pragma unroll
            for (int a = 0; a < TPL; ++a) x[q][a] = S[line * LS + (n + TPL * q) * YS + a];
#pragma unroll
        for (int q = 0; q < Q; ++q) dft_small<TPL, SIGN>(x[q]);
        if (pass == 0) {
            __syncthreads();  // exchange reads done before the transpose writes
            // output (line, f = k2 + 16 k1) of pass 1 goes to row f, column `line`
#pragma unroll
            for (int q = 0; q < Q; ++q)
#pragma unroll
                for (int k1 = 0; k1 < TPL; ++k1) S[(n + TPL * q + 16 * k1) * RS + line] = x[q][perm_index<TPL>(k1)];
            __syncthreads();
        }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q)
#pragma unroll
        for (int k1 = 0; k1 < TPL; ++k1) {
            const int f = n + TPL * q + 16 * k1;
            C r = x[q][perm_index<TPL>(k1)];
            r.x *= scale;
            r.y *= scale;
            if (FWD) gstore(dst, (unsigned)((line * N + f) * N), r);  // G[line][f][k]
            else gstore(dst, (unsigned)(line * N + f), r);            // x[k][line][f]
        }
}

bool fft_slice_supported(int N, size_t elem_bytes)
{
    return elem_bytes == sizeof(float2) && (N == 64 || N == 128) &&
           (N == 64 ? SliceGeo<float, 6>::LDS : SliceGeo<float, 7>::LDS) <= lds_limit();
}

template <int LOGN>
static hipError_t launch_slice_f32(const float2* in, float2* out, const LineMap& m, int sign, float scale,
                                   const float2* tw, int nsig, hipStream_t st)
{
    typedef SliceGeo<float, LOGN> Geo;
    const long nblocks = (long)nsig * (Geo::N / 16);
    const long nwg = (nblocks + 7) / 8 * 8 * 16;
    if (nwg > 0x7fffffffL) return hipErrorInvalidValue;
    if (sign < 0)
        return pfs::launch((k_fft_slice<float, LOGN, -1>), dim3((unsigned)nwg), dim3(Geo::THREADS), Geo::LDS, st, in, out,
                           m, tw, scale, nblocks);
    return pfs::launch((k_fft_slice<float, LOGN, 1>), dim3((unsigned)nwg), dim3(Geo::THREADS), Geo::LDS, st, in, out, m,
                       tw, scale, nblocks);
}

hipError_t launch_fft_slice_f32(const float2* in, float2* out, const LineMap& m, int N, int sign, float scale,
                                const float2* tw, int nsig, hipStream_t st)
{
    if (nsig < 1 || !fft_slice_supported(N, sizeof(float2))) return hipErrorInvalidValue;
    return N == 64 ? launch_slice_f32<6>(in, out, m, sign, scale, tw, nsig, st)
                   : launch_slice_f32<7>(in, out, m, sign, scale, tw, nsig, st);
}

hipError_t launch_fft_lines_f64(const double2* in, double2* out, const LineMap& m, int N, int sign, double scale,
                                const double2* tw, int kind, int nsig, hipStream_t st)
{
    return launch_lines<double>(in, out, m, N, sign, scale, tw, kind, nsig, st);
}

hipError_t launch_fft_lines_f32(const float2* in, float2* out, const LineMap& m, int N, int sign, float scale,
                                const float2* tw, int kind, int nsig, hipStream_t st, int paired)
{
    return launch_lines<float>(in, out, m, N, sign, scale, tw, kind, nsig, st, paired);
}

}  // namespace pfs
