/*
 * CPU oracle for the SO(3) Fourier transform path (DSOFT / iDSOFT).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a plain-C restatement of the
 * reference algorithm (PFSOFTlib 1.0.0).  It is the checker the HIP path is
 * compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product library never links
 * it and has no CPU fallback.
 *
 * Precision mix, as the reference compiles on x86-64 with g++:
 *   - Wigner-d tables and the DWT contraction are long double (x87, 80 bit);
 *   - every transcendental (sin, cos, pow, sqrt) is the C *double* function:
 *     the reference calls them unqualified inside namespace pfsoft with only
 *     <cmath> included, which binds the C library double versions, so every
 *     long double argument is rounded to double first.  C does the same
 *     implicit conversion, so the calls below are written exactly as there;
 *   - pi is a double literal widened to long double
 *     (include/PFSOFTlib_headers/forward_declarations.hpp:227-228).
 * The reference's layer-wise 2D FFT goes through FFTW 3.3.4
 * (src/fftw_wrapper.cpp:47-172), which is not vendored and not installed
 * here; oracle_fft2_layers() restates its contract (in place, unnormalised,
 * FFTW_FORWARD = exp(-2 pi i ...), n0 = cols slow, n1 = rows fast) with a
 * long double radix-2 FFT (power-of-two sizes) or a direct DFT (other
 * sizes).  Pinning: the tests compare it with numpy's fft2 and a brute-force
 * DFT.
 *
 * Layouts (shared with the product library):
 *   grid   : interleaved complex double, element (row, col, lay) at
 *            N*N*lay + N*col + row        (grid3D_cx.hpp:293-303)
 *   coeffs : packed, degree l block at offset(l) = (4 l^3 - l) / 3 complex,
 *            element (l, M, M') at offset(l) + iM + (2l+1) * iM' with
 *            iX = X >= 0 ? X : 2l+1+X     (dsoft_fourier_coefficients.cpp:48-112,
 *            matrix_cx column-major storage)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef long double ld;

/* constants<long double>::pi: a double literal widened (forward_declarations.hpp:227-228) */
static const ld PI_LD = (ld)3.14159265358979323846;
/* constants<double>::pi */
static const double PI_D = 3.14159265358979323846;

/* DSOFT_THRESHOLD (cmake_aux/compiler_config.hpp.in:55) */
#define ORACLE_DSOFT_THRESHOLD 20

static inline long coeff_offset(int l) { return ((long)4 * l * l * l - l) / 3; }

static inline long coeff_index(int l, int M, int Mp)
{
    long n = 2 * l + 1;
    long iM = M >= 0 ? M : n + M;
    long iMp = Mp >= 0 ? Mp : n + Mp;
    return coeff_offset(l) + iM + n * iMp;
}

long oracle_coeff_count(int B) { return coeff_offset(B); }

long oracle_coeff_index(int l, int M, int Mp) { return coeff_index(l, M, Mp); }

/* ------------------------------------------------------------------------- */
/* Seeded synthetic coefficients: SplitMix64, u = (x >> 11) * 2^-53,
 * value = (2u - 1) * (l + 1)^(-power).  Fill order l up, M = -l..l,
 * M' = -l..l, re then im -- the order of the reference's rand()
 * (include/PFSOFTlib_headers/random.hpp:105-116), with a reproducible
 * generator in place of its time-seeded mt19937_64.                       */
static inline uint64_t splitmix64_next(uint64_t* s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void oracle_seeded_coeffs(int B, uint64_t seed, double power, double* coeffs)
{
    uint64_t s = seed;
    for (int l = 0; l < B; ++l) {
        double scale = pow((double)(l + 1), -power);
        for (int M = -l; M <= l; ++M)
            for (int Mp = -l; Mp <= l; ++Mp) {
                long e = coeff_index(l, M, Mp);
                double u0 = (double)(splitmix64_next(&s) >> 11) * 0x1.0p-53;
                double u1 = (double)(splitmix64_next(&s) >> 11) * 0x1.0p-53;
                coeffs[2 * e] = (2.0 * u0 - 1.0) * scale;
                coeffs[2 * e + 1] = (2.0 * u1 - 1.0) * scale;
            }
    }
}

/* ------------------------------------------------------------------------- */
/* DWT::quadrature_weights (include/PFSOFTlib_headers/fn_dwt.hpp:62-83).
 * w has 2B entries.                                                         */
void oracle_quadrature_weights(int B, ld* w)
{
    for (int i = 0; i < B; ++i) {
        /* 2.0/bandwidth * sin(...) is a double product (fn_dwt.hpp:72) */
        ld wi = 2.0 / B * sin(PI_LD * (2.0 * i + 1.0) / (4.0 * B));
        ld sum = 0;
        for (int k = 0; k < B; ++k)
            sum += 1.0 / (2.0 * k + 1.0) * sin((2.0 * i + 1.0) * (2.0 * k + 1.0) * PI_LD / (4.0 * B));
        wi *= sum;
        w[i] = wi;
        w[2 * B - 1 - i] = wi;
    }
}

void oracle_quadrature_weights_d(int B, double* w)
{
    ld* t = (ld*)malloc(sizeof(ld) * 2 * B);
    oracle_quadrature_weights(B, t);
    for (int i = 0; i < 2 * B; ++i) w[i] = (double)t[i];
    free(t);
}

/* ------------------------------------------------------------------------- */
/* DWT::weighted_wigner_d_matrix / DWT::wigner_d_matrix
 * (fn_dwt.hpp:105-200 and :237-332).  wig is column-major
 * (B - minJ) x 2B (matrix.hpp:413-416): wig[j * rows + i].
 * weights == NULL gives the unweighted form.                                */
void oracle_wigner_d(int B, int M, int Mp, const ld* weights, ld* wig)
{
    int aM = abs(M), aMp = abs(Mp);
    int minJ = aM > aMp ? aM : aMp;
    int mn = aM < aMp ? aM : aMp;
    int rows = B - minJ, cols = 2 * B;
#define W(i, j) wig[(long)(j) * rows + (i)]

    ld normFactor = sqrt((2.0 * minJ + 1.0) / 2.0);
    for (int i = 0; i < minJ - mn; ++i)
        normFactor *= sqrt((2.0 * minJ - i) / (i + 1.0));

    /* fn_dwt.hpp:124-125: the second assignment resets the sign to -1
     * whenever minJ == |M|, so the base-case sign is -1 in every call the
     * transforms make. */
    ld sinSign = (minJ == aM && M >= 0 && ((minJ - Mp) & 1)) ? 1 : -1;
    sinSign = (minJ != aM && Mp < 0 && ((minJ - Mp) & 1)) ? sinSign : -1;

    ld cosPower, sinPower;
    if (minJ == aM && M >= 0) { cosPower = minJ + Mp; sinPower = minJ - Mp; }
    else if (minJ == aM)      { cosPower = minJ - Mp; sinPower = minJ + Mp; }
    else if (Mp >= 0)         { cosPower = minJ + M;  sinPower = minJ - M;  }
    else                      { cosPower = minJ - M;  sinPower = minJ + M;  }

    ld* cosBeta = (ld*)malloc(sizeof(ld) * cols);
    for (int i = 0; i < cols; ++i) {
        ld sinHalfBeta = sin(0.5 * ((2.0 * i + 1.0) * PI_LD) / (4.0 * B));
        ld cosHalfBeta = cos(0.5 * ((2.0 * i + 1.0) * PI_LD) / (4.0 * B));
        cosBeta[i] = cos(((2.0 * i + 1.0) * PI_LD) / (4.0 * B));
        ld v = normFactor * sinSign * (ld)pow(sinHalfBeta, sinPower) * (ld)pow(cosHalfBeta, cosPower);
        if (weights) v = v * weights[i];
        W(0, i) = v;
    }

    for (int i = 0; i < rows - 1; ++i) {
        ld c1 = 0;
        ld idx = minJ + i;
        ld norm = sqrt((2.0 * idx + 3.0) / (2.0 * idx + 1.0));
        ld nom = (idx + 1.0) * (2. * idx + 1.0);
        /* 1.0 / sqrt(...) is a double quotient (fn_dwt.hpp:178) */
        ld den = 1.0 / sqrt(((idx + 1) * (idx + 1) - M * M) * ((idx + 1) * (idx + 1) - Mp * Mp));
        ld f1 = norm * nom * den;
        ld f2 = 0;
        if (minJ + i != 0) {
            ld t1 = sqrt((2.0 * idx + 3.0) / (2.0 * idx - 1.0)) * (idx + 1.0) / idx;
            ld t2 = sqrt((idx * idx - M * M) * (idx * idx - Mp * Mp));
            c1 = -t1 * t2 * den;
            f2 = -M * Mp / (idx * (idx + 1.));
        }
        for (int j = 0; j < cols; ++j)
            W(i + 1, j) = c1 * (i == 0 ? 0 : W(i - 1, j)) + W(i, j) * f1 * (f2 + cosBeta[j]);
    }
    free(cosBeta);
#undef W
}

/* Double-rounded table for the tests (same layout). */
void oracle_wigner_d_d(int B, int M, int Mp, int weighted, double* out)
{
    int minJ = abs(M) > abs(Mp) ? abs(M) : abs(Mp);
    long n = (long)(B - minJ) * 2 * B;
    ld* t = (ld*)malloc(sizeof(ld) * (n > 0 ? n : 1));
    ld* w = NULL;
    if (weighted) {
        w = (ld*)malloc(sizeof(ld) * 2 * B);
        oracle_quadrature_weights(B, w);
    }
    oracle_wigner_d(B, M, Mp, w, t);
    for (long i = 0; i < n; ++i) out[i] = (double)t[i];
    free(t);
    free(w);
}

/* ------------------------------------------------------------------------- */
/* Flip helpers (include/PFSOFTlib_headers/fn_flip.hpp:44-305) on a
 * column-major rows x cols long double matrix.                              */
static void fliplr(ld* m, int rows, int cols, int negate_parity)
{
    /* negate_parity: -1 plain fliplr (:46), 0 fliplr_ne2nderow (:88,
     * negates rows with even index), 1 fliplr_ne2ndorow (:138, odd rows) */
    for (int j = 0; j < cols / 2; ++j)
        for (int k = 0; k < rows; ++k) {
            ld* a = &m[(long)j * rows + k];
            ld* b = &m[(long)(cols - j - 1) * rows + k];
            ld tmp = *a;
            int neg = negate_parity >= 0 && (k & 1) == negate_parity;
            if (neg) { *a = -*b; *b = -tmp; }
            else     { *a = *b;  *b = tmp;  }
        }
}

static void flipud(ld* m, int rows, int cols, int negate_parity)
{
    /* negate_parity: -1 plain flipud (:186), 0 flipud_ne2ndecol (:227,
     * negates columns with even index), 1 flipud_ne2ndocol (:277) */
    for (int k = 0; k < cols; ++k)
        for (int j = 0; j < rows / 2; ++j) {
            ld* a = &m[(long)k * rows + j];
            ld* b = &m[(long)k * rows + rows - j - 1];
            ld tmp = *a;
            int neg = negate_parity >= 0 && (k & 1) == negate_parity;
            if (neg) { *a = -*b; *b = -tmp; }
            else     { *a = *b;  *b = tmp;  }
        }
}

/* matrix<long double>::transpose (matrix.hpp:443-464): column-major
 * rows x cols -> column-major cols x rows.                                  */
static void transpose(const ld* m, int rows, int cols, ld* out)
{
    for (int i = 0; i < rows; ++i)
        for (int j = 0; j < cols; ++j)
            out[(long)i * cols + j] = m[(long)j * rows + i];
}

/* matrix<long double> * vector<complex<long double>> (matrix.hpp:358-377):
 * y[j] = sum_i A[j,i] x[i], outer loop over columns i, accumulation in long
 * double starting from zero; the double input is widened
 * (fn_fourier_transforms.hpp:48-72) and the result is rounded to double.  */
static void matvec(const ld* A, int rows, int cols, const double* x /* 2*cols */,
                   double* y /* 2*rows */, ld* acc /* 2*rows scratch */)
{
    for (int j = 0; j < 2 * rows; ++j) acc[j] = 0;
    for (int i = 0; i < cols; ++i) {
        ld xr = x[2 * i], xi = x[2 * i + 1];
        const ld* col = A + (long)i * rows;
        for (int j = 0; j < rows; ++j) {
            acc[2 * j] += col[j] * xr;
            acc[2 * j + 1] += col[j] * xi;
        }
    }
    for (int j = 0; j < 2 * rows; ++j) y[j] = (double)acc[j];
}

/* ------------------------------------------------------------------------- */
/* Layer-wise unnormalised 2D DFT, in place (src/fftw_wrapper.cpp:47-172 and
 * grid3D_cx.hpp:309-427).  For layer t the element (i0, i1) sits at
 * arr[t*n0*n1 + i0*n1 + i1] with n0 = cols, n1 = rows.  sign = -1 forward
 * (FFTW_FORWARD), +1 backward.                                              */
static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

static void dft1(ld* re, ld* im, int n, int sign, const ld* cs, const ld* sn, ld* tr, ld* ti)
{
    /* cs[j] = cos(2 pi j / n), sn[j] = sin(2 pi j / n) */
    if (is_pow2(n)) {
        /* iterative radix-2, bit-reversed input order */
        for (int i = 1, j = 0; i < n; ++i) {
            int bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) {
                ld t = re[i]; re[i] = re[j]; re[j] = t;
                t = im[i]; im[i] = im[j]; im[j] = t;
            }
        }
        for (int len = 2; len <= n; len <<= 1) {
            int step = n / len;
            for (int i = 0; i < n; i += len)
                for (int k = 0; k < len / 2; ++k) {
                    ld wr = cs[k * step], wi = sign * sn[k * step];
                    int a = i + k, b = i + k + len / 2;
                    ld xr = re[b] * wr - im[b] * wi;
                    ld xi = re[b] * wi + im[b] * wr;
                    re[b] = re[a] - xr; im[b] = im[a] - xi;
                    re[a] += xr; im[a] += xi;
                }
        }
    } else {
        for (int k = 0; k < n; ++k) {
            ld sr = 0, si = 0;
            for (int j = 0; j < n; ++j) {
                long e = ((long)j * k) % n;
                ld wr = cs[e], wi = sign * sn[e];
                sr += re[j] * wr - im[j] * wi;
                si += re[j] * wi + im[j] * wr;
            }
            tr[k] = sr; ti[k] = si;
        }
        memcpy(re, tr, sizeof(ld) * n);
        memcpy(im, ti, sizeof(ld) * n);
    }
}

static void trig_table(int n, ld* cs, ld* sn)
{
    for (int j = 0; j < n; ++j) {
        ld a = 2.0L * 3.14159265358979323846264338327950288L * (ld)j / (ld)n;
        cs[j] = cosl(a);
        sn[j] = sinl(a);
    }
}

static void fft2_layer(double* arr, int n0, int n1, int sign)
{
    int nm = n0 > n1 ? n0 : n1;
    ld* re = (ld*)malloc(sizeof(ld) * (size_t)n0 * n1);
    ld* im = (ld*)malloc(sizeof(ld) * (size_t)n0 * n1);
    ld* cs0 = (ld*)malloc(sizeof(ld) * n0);
    ld* sn0 = (ld*)malloc(sizeof(ld) * n0);
    ld* cs1 = (ld*)malloc(sizeof(ld) * n1);
    ld* sn1 = (ld*)malloc(sizeof(ld) * n1);
    ld* lr = (ld*)malloc(sizeof(ld) * nm);
    ld* li = (ld*)malloc(sizeof(ld) * nm);
    ld* tr = (ld*)malloc(sizeof(ld) * nm);
    ld* ti = (ld*)malloc(sizeof(ld) * nm);
    trig_table(n0, cs0, sn0);
    trig_table(n1, cs1, sn1);
    for (long e = 0; e < (long)n0 * n1; ++e) { re[e] = arr[2 * e]; im[e] = arr[2 * e + 1]; }
    /* along i1 (contiguous) */
    for (int i0 = 0; i0 < n0; ++i0)
        dft1(re + (long)i0 * n1, im + (long)i0 * n1, n1, sign, cs1, sn1, tr, ti);
    /* along i0 */
    for (int i1 = 0; i1 < n1; ++i1) {
        for (int i0 = 0; i0 < n0; ++i0) { lr[i0] = re[(long)i0 * n1 + i1]; li[i0] = im[(long)i0 * n1 + i1]; }
        dft1(lr, li, n0, sign, cs0, sn0, tr, ti);
        for (int i0 = 0; i0 < n0; ++i0) { re[(long)i0 * n1 + i1] = lr[i0]; im[(long)i0 * n1 + i1] = li[i0]; }
    }
    for (long e = 0; e < (long)n0 * n1; ++e) { arr[2 * e] = (double)re[e]; arr[2 * e + 1] = (double)im[e]; }
    free(re); free(im); free(cs0); free(sn0); free(cs1); free(sn1);
    free(lr); free(li); free(tr); free(ti);
}

void oracle_fft2_layers(int cols, int rows, int lays, double* arr, int sign, int threads)
{
    long slice = 2L * rows * cols;
    (void)threads;
#pragma omp parallel for schedule(dynamic) num_threads(threads)
    for (int t = 0; t < lays; ++t)
        fft2_layer(arr + t * slice, cols, rows, sign);
}

/* ------------------------------------------------------------------------- */
/* Per-thread work buffers for the DWT loops.                               */
typedef struct {
    ld* d;      /* wigner matrix, up to B x 2B */
    ld* dt;     /* transposed copy */
    ld* acc;    /* matvec accumulator */
    double* s;  /* 2B complex */
    double* sh; /* B complex */
} dwt_scratch;

static void scratch_init(dwt_scratch* w, int B)
{
    w->d = (ld*)malloc(sizeof(ld) * (size_t)B * 2 * B);
    w->dt = (ld*)malloc(sizeof(ld) * (size_t)B * 2 * B);
    w->acc = (ld*)malloc(sizeof(ld) * 4 * (size_t)B);
    w->s = (double*)malloc(sizeof(double) * 4 * (size_t)B);
    w->sh = (double*)malloc(sizeof(double) * 4 * (size_t)B);
}

static void scratch_free(dwt_scratch* w)
{
    free(w->d); free(w->dt); free(w->acc); free(w->s); free(w->sh);
}

#define G(arr, N, row, col, lay) ((arr) + 2 * ((long)(N) * (N) * (lay) + (long)(N) * (col) + (row)))

/* gather s[k] = sample(row, col, k) */
static void gather(const double* F, int N, int row, int col, double* s)
{
    for (int k = 0; k < N; ++k) {
        const double* p = G(F, N, row, col, k);
        s[2 * k] = p[0];
        s[2 * k + 1] = p[1];
    }
}

/* fc(bandwidth - (sh.end() - e), M, Mp) = norm * sh[e]: sh has B - m rows,
 * element e belongs to degree l = m + e. */
static void store_fwd(double* fc, int B, int M, int Mp, const double* sh, int n, double norm)
{
    for (int e = 0; e < n; ++e) {
        int l = B - (n - e);
        long i = coeff_index(l, M, Mp);
        fc[2 * i] = norm * sh[2 * e];
        fc[2 * i + 1] = norm * sh[2 * e + 1];
    }
}

static void neg(double* v, int n, int start, int step)
{
    for (int e = start; e < n; e += step) { v[2 * e] = -v[2 * e]; v[2 * e + 1] = -v[2 * e + 1]; }
}

static void scale_ld(ld* m, long n, ld f)
{
    for (long i = 0; i < n; ++i) m[i] *= f;
}

/* Forward transform: FourierTransforms::DSOFT (src/fn_dsoft.cpp:76-285).
 * grid is not modified (the reference takes it by value, :76).             */
int oracle_dsoft(int B, const double* grid, double* fc, int threads)
{
    if (B <= 0) return -1;
    const int N = 2 * B;
    const long gsz = 2L * N * N * N;
    double* F = (double*)malloc(sizeof(double) * gsz);
    if (!F) return -2;
    memcpy(F, grid, sizeof(double) * gsz);

    /* :104 layer-wise FFT2, scale 1 */
    oracle_fft2_layers(N, N, N, F, -1, threads);

    /* :109-110 */
    ld* weights = (ld*)malloc(sizeof(ld) * N);
    oracle_quadrature_weights(B, weights);

    /* :119 */
    const double norm = PI_D / (B * N);

    /* (0,0): :112-129 */
    {
        dwt_scratch w;
        scratch_init(&w, B);
        oracle_wigner_d(B, 0, 0, weights, w.d);
        scale_ld(w.d, (long)B * N, -1);
        gather(F, N, 0, 0, w.s);
        matvec(w.d, B, N, w.s, w.sh, w.acc);
        store_fwd(fc, B, 0, 0, w.sh, B, norm);
        scratch_free(&w);
    }

    const int par = B >= ORACLE_DSOFT_THRESHOLD;
    (void)par;
#pragma omp parallel if (par) num_threads(threads)
    {
        dwt_scratch w;
        scratch_init(&w, B);

        /* :137-206 */
#pragma omp for schedule(dynamic) nowait
        for (int M = 1; M < B; ++M) {
            int R = B - M;
            oracle_wigner_d(B, M, 0, weights, w.d);
            scale_ld(w.d, (long)R * N, -1);

            gather(F, N, 0, M, w.s);                 /* f_{M,0} */
            matvec(w.d, R, N, w.s, w.sh, w.acc);
            store_fwd(fc, B, M, 0, w.sh, R, norm);

            gather(F, N, M, 0, w.s);                 /* f_{0,M} */
            matvec(w.d, R, N, w.s, w.sh, w.acc);
            if (M & 1) neg(w.sh, R, 0, 1);
            store_fwd(fc, B, 0, M, w.sh, R, norm);

            fliplr(w.d, R, N, -1);
            gather(F, N, 0, N - M, w.s);             /* f_{-M,0} */
            matvec(w.d, R, N, w.s, w.sh, w.acc);
            if (M & 1) neg(w.sh, R, 0, 2);
            else       neg(w.sh, R, 1, 2);
            store_fwd(fc, B, -M, 0, w.sh, R, norm);

            gather(F, N, N - M, 0, w.s);             /* f_{0,-M} */
            matvec(w.d, R, N, w.s, w.sh, w.acc);
            neg(w.sh, R, 1, 2);
            store_fwd(fc, B, 0, -M, w.sh, R, norm);

            oracle_wigner_d(B, M, M, weights, w.d);
            scale_ld(w.d, (long)R * N, -1);

            gather(F, N, M, M, w.s);                 /* f_{M,M} */
            matvec(w.d, R, N, w.s, w.sh, w.acc);
            store_fwd(fc, B, M, M, w.sh, R, norm);

            gather(F, N, N - M, N - M, w.s);         /* f_{-M,-M} */
            matvec(w.d, R, N, w.s, w.sh, w.acc);
            store_fwd(fc, B, -M, -M, w.sh, R, norm);

            fliplr(w.d, R, N, 1);

            gather(F, N, N - M, M, w.s);             /* f_{M,-M} */
            matvec(w.d, R, N, w.s, w.sh, w.acc);
            store_fwd(fc, B, M, -M, w.sh, R, norm);

            gather(F, N, M, N - M, w.s);             /* f_{-M,M} */
            matvec(w.d, R, N, w.s, w.sh, w.acc);
            store_fwd(fc, B, -M, M, w.sh, R, norm);
        }

        /* :214-283, the fused (M, M') loop with 1 <= M' < M < B */
#pragma omp for schedule(dynamic) nowait
        for (int MMp = 0; MMp < (B - 2) * (B - 1) / 2; ++MMp) {
            int i = MMp / (B - 1) + 1;
            int j = MMp % (B - 1) + 1;
            int M = j > i ? B - i : i + 1;
            int Mp = j > i ? B - j : j;
            int R = B - (M > Mp ? M : Mp);
            oracle_wigner_d(B, M, Mp, weights, w.d);

            gather(F, N, Mp, M, w.s);                /* f_{M,Mp} */
            matvec(w.d, R, N, w.s, w.sh, w.acc);
            neg(w.sh, R, 0, 1);
            store_fwd(fc, B, M, Mp, w.sh, R, norm);

            gather(F, N, M, Mp, w.s);                /* f_{Mp,M} */
            matvec(w.d, R, N, w.s, w.sh, w.acc);
            if (!((M - Mp) & 1)) neg(w.sh, R, 0, 1);
            store_fwd(fc, B, Mp, M, w.sh, R, norm);

            gather(F, N, N - Mp, N - M, w.s);        /* f_{-M,-Mp} */
            matvec(w.d, R, N, w.s, w.sh, w.acc);
            if (!((M - Mp) & 1)) neg(w.sh, R, 0, 1);
            store_fwd(fc, B, -M, -Mp, w.sh, R, norm);

            gather(F, N, N - M, N - Mp, w.s);        /* f_{-Mp,-M} */
            matvec(w.d, R, N, w.s, w.sh, w.acc);
            neg(w.sh, R, 0, 1);
            store_fwd(fc, B, -Mp, -M, w.sh, R, norm);

            fliplr(w.d, R, N, 0);

            gather(F, N, N - M, Mp, w.s);            /* f_{Mp,-M} */
            matvec(w.d, R, N, w.s, w.sh, w.acc);
            store_fwd(fc, B, Mp, -M, w.sh, R, norm);

            gather(F, N, N - Mp, M, w.s);            /* f_{M,-Mp} */
            matvec(w.d, R, N, w.s, w.sh, w.acc);
            store_fwd(fc, B, M, -Mp, w.sh, R, norm);

            if ((M - Mp) & 1) scale_ld(w.d, (long)R * N, -1);

            gather(F, N, M, N - Mp, w.s);            /* f_{-Mp,M} */
            matvec(w.d, R, N, w.s, w.sh, w.acc);
            store_fwd(fc, B, -Mp, M, w.sh, R, norm);

            gather(F, N, Mp, N - M, w.s);            /* f_{-M,Mp} */
            matvec(w.d, R, N, w.s, w.sh, w.acc);
            store_fwd(fc, B, -M, Mp, w.sh, R, norm);
        }
        scratch_free(&w);
    }
    free(weights);
    free(F);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* sh[e] = norm * fc(bandwidth - (sh.end() - e), M, Mp) */
static void load_inv(const double* fc, int B, int M, int Mp, double* sh, int n, double norm)
{
    for (int e = 0; e < n; ++e) {
        int l = B - (n - e);
        long i = coeff_index(l, M, Mp);
        sh[2 * e] = norm * fc[2 * i];
        sh[2 * e + 1] = norm * fc[2 * i + 1];
    }
}

static void scatter(double* grid, int N, int row, int col, const double* s)
{
    for (int k = 0; k < N; ++k) {
        double* p = G(grid, N, row, col, k);
        p[0] = s[2 * k];
        p[1] = s[2 * k + 1];
    }
}

/* Inverse transform: FourierTransforms::IDSOFT (src/fn_idsoft.cpp:72-290).
 * Writes the grid in place; entries with row == B or col == B keep what the
 * caller's grid held before the inverse FFT reads them, as in the reference. */
int oracle_idsoft(int B, const double* fc, double* grid, int threads)
{
    if (B <= 0) return -1;
    const int N = 2 * B;
    /* :109 */
    const double norm = (B * N) / PI_D;

    /* (0,0): :100-119 */
    {
        dwt_scratch w;
        scratch_init(&w, B);
        oracle_wigner_d(B, 0, 0, NULL, w.d);
        scale_ld(w.d, (long)B * N, -1);
        transpose(w.d, B, N, w.dt);               /* now N x B */
        load_inv(fc, B, 0, 0, w.sh, B, norm);
        matvec(w.dt, N, B, w.sh, w.s, w.acc);
        scatter(grid, N, 0, 0, w.s);
        scratch_free(&w);
    }

    const int par = B >= ORACLE_DSOFT_THRESHOLD;
    (void)par;
#pragma omp parallel if (par) num_threads(threads)
    {
        dwt_scratch w;
        scratch_init(&w, B);

        /* :127-203 */
#pragma omp for schedule(dynamic) nowait
        for (int M = 1; M < B; ++M) {
            int R = B - M;
            oracle_wigner_d(B, M, 0, NULL, w.d);
            scale_ld(w.d, (long)R * N, -1);
            transpose(w.d, R, N, w.dt);           /* N x R */

            load_inv(fc, B, M, 0, w.sh, R, norm);  /* f_{M,0} */
            matvec(w.dt, N, R, w.sh, w.s, w.acc);
            scatter(grid, N, 0, M, w.s);

            load_inv(fc, B, 0, M, w.sh, R, norm);  /* f_{0,M} */
            if (M & 1) neg(w.sh, R, 0, 1);
            matvec(w.dt, N, R, w.sh, w.s, w.acc);
            scatter(grid, N, M, 0, w.s);

            flipud(w.dt, N, R, -1);

            load_inv(fc, B, -M, 0, w.sh, R, norm); /* f_{-M,0} */
            if (M & 1) neg(w.sh, R, 0, 2);
            else       neg(w.sh, R, 1, 2);
            matvec(w.dt, N, R, w.sh, w.s, w.acc);
            scatter(grid, N, 0, N - M, w.s);

            load_inv(fc, B, 0, -M, w.sh, R, norm); /* f_{0,-M} */
            neg(w.sh, R, 1, 2);
            matvec(w.dt, N, R, w.sh, w.s, w.acc);
            scatter(grid, N, N - M, 0, w.s);

            oracle_wigner_d(B, M, M, NULL, w.d);
            scale_ld(w.d, (long)R * N, -1);
            transpose(w.d, R, N, w.dt);

            load_inv(fc, B, M, M, w.sh, R, norm);  /* f_{M,M} */
            matvec(w.dt, N, R, w.sh, w.s, w.acc);
            scatter(grid, N, M, M, w.s);

            load_inv(fc, B, -M, -M, w.sh, R, norm); /* f_{-M,-M} */
            matvec(w.dt, N, R, w.sh, w.s, w.acc);
            scatter(grid, N, N - M, N - M, w.s);

            flipud(w.dt, N, R, 1);

            load_inv(fc, B, M, -M, w.sh, R, norm); /* f_{M,-M} */
            matvec(w.dt, N, R, w.sh, w.s, w.acc);
            scatter(grid, N, N - M, M, w.s);

            load_inv(fc, B, -M, M, w.sh, R, norm); /* f_{-M,M} */
            matvec(w.dt, N, R, w.sh, w.s, w.acc);
            scatter(grid, N, M, N - M, w.s);
        }

        /* :211-283 */
#pragma omp for schedule(dynamic) nowait
        for (int MMp = 0; MMp < (B - 2) * (B - 1) / 2; ++MMp) {
            int i = MMp / (B - 1) + 1;
            int j = MMp % (B - 1) + 1;
            int M = j > i ? B - i : i + 1;
            int Mp = j > i ? B - j : j;
            int R = B - (M > Mp ? M : Mp);
            oracle_wigner_d(B, M, Mp, NULL, w.d);
            transpose(w.d, R, N, w.dt);

            load_inv(fc, B, M, Mp, w.sh, R, norm);   /* f_{M,Mp} */
            neg(w.sh, R, 0, 1);
            matvec(w.dt, N, R, w.sh, w.s, w.acc);
            scatter(grid, N, Mp, M, w.s);

            load_inv(fc, B, Mp, M, w.sh, R, norm);   /* f_{Mp,M} */
            if (!((M - Mp) & 1)) neg(w.sh, R, 0, 1);
            matvec(w.dt, N, R, w.sh, w.s, w.acc);
            scatter(grid, N, M, Mp, w.s);

            load_inv(fc, B, -M, -Mp, w.sh, R, norm); /* f_{-M,-Mp} */
            if (!((M - Mp) & 1)) neg(w.sh, R, 0, 1);
            matvec(w.dt, N, R, w.sh, w.s, w.acc);
            scatter(grid, N, N - Mp, N - M, w.s);

            load_inv(fc, B, -Mp, -M, w.sh, R, norm); /* f_{-Mp,-M} */
            neg(w.sh, R, 0, 1);
            matvec(w.dt, N, R, w.sh, w.s, w.acc);
            scatter(grid, N, N - M, N - Mp, w.s);

            flipud(w.dt, N, R, 0);

            load_inv(fc, B, Mp, -M, w.sh, R, norm);  /* f_{Mp,-M} */
            matvec(w.dt, N, R, w.sh, w.s, w.acc);
            scatter(grid, N, N - M, Mp, w.s);

            load_inv(fc, B, M, -Mp, w.sh, R, norm);  /* f_{M,-Mp} */
            matvec(w.dt, N, R, w.sh, w.s, w.acc);
            scatter(grid, N, N - Mp, M, w.s);

            if ((M - Mp) & 1) scale_ld(w.dt, (long)R * N, -1);

            load_inv(fc, B, -Mp, M, w.sh, R, norm);  /* f_{-Mp,M} */
            matvec(w.dt, N, R, w.sh, w.s, w.acc);
            scatter(grid, N, M, N - Mp, w.s);

            load_inv(fc, B, -M, Mp, w.sh, R, norm);  /* f_{-M,Mp} */
            matvec(w.dt, N, R, w.sh, w.s, w.acc);
            scatter(grid, N, Mp, N - M, w.s);
        }
        scratch_free(&w);
    }

    /* :289 layer-wise IFFT2 then scale by 1/(4B^2) (grid3D_cx.hpp:400-411) */
    oracle_fft2_layers(N, N, N, grid, +1, threads);
    const double sc = 1. / (4. * B * B);
    const long n = (long)N * N * N;
#pragma omp parallel for num_threads(threads)
    for (long e = 0; e < n; ++e) {
        grid[2 * e] *= sc;
        grid[2 * e + 1] *= sc;
    }
    return 0;
}

int oracle_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
