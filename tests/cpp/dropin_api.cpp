// The drop-in C++ surface beyond the round trip, written the way reference
// user code uses it:
//   - DSOFT / IDSOFT have the reference's exact function types
//     (include/PFSOFTlib_headers/fn_fourier_transforms.hpp:75,78), checked at
//     compile time;
//   - the grid and coefficient constructors (grid3D_cx.hpp:61-70,
//     dsoft_fourier_coefficients.hpp:60-61) and the single-value complex
//     constructor (complex.hpp:133-138);
//   - grid3D::layer_wise_DFT2 / layer_wise_IDFT2 with a complex scale
//     (grid3D_cx.hpp:81-82, 309-427), against a direct O(n^4) DFT per slice;
//   - both stream operators (grid3D_cx.hpp:430, dsoft_fourier_coefficients.hpp:71);
//   - the seam under the reference's own symbol name
//     (fftw_wrapper.hpp:31-32).
// Prints one line per check and exits non-zero on the first failure.
#include "pfsoft_amd.hpp"

#include <cmath>
#include <cstdio>
#include <sstream>
#include <type_traits>
#include <vector>

using namespace pfsoft;

static_assert(std::is_same<decltype(&FourierTransforms::DSOFT),
                           void (*)(grid3D<complex<double> >, DSOFTFourierCoefficients&, int)>::value,
              "DSOFT has the reference's type");
static_assert(std::is_same<decltype(&FourierTransforms::IDSOFT),
                           void (*)(const DSOFTFourierCoefficients&, grid3D<complex<double> >&, int)>::value,
              "IDSOFT has the reference's type");
static_assert(sizeof(complex<double>) == 2 * sizeof(double), "complex<double> is two contiguous doubles");

static int fails = 0;
static void check(bool ok, const char* what)
{
    printf("%-60s %s\n", what, ok ? "ok" : "FAILED");
    if (!ok) ++fails;
}

// Direct DFT of one n x n slice s[col][row] (row fastest), sign -1 or +1.
static std::vector<complex<double> > dft2(const complex<double>* s, int n, int sign)
{
    std::vector<complex<double> > out((size_t)n * n);
    const double two_pi = 6.283185307179586476925286766559;
    for (int q = 0; q < n; ++q)
        for (int p = 0; p < n; ++p) {
            long double re = 0, im = 0;
            for (int c = 0; c < n; ++c)
                for (int r = 0; r < n; ++r) {
                    const double a = sign * two_pi * (double)((q * c + p * r) % n) / n;
                    const double wr = std::cos(a), wi = std::sin(a);
                    const complex<double>& x = s[(size_t)n * c + r];
                    re += (long double)x.re * wr - (long double)x.im * wi;
                    im += (long double)x.re * wi + (long double)x.im * wr;
                }
            out[(size_t)n * q + p] = complex<double>((double)re, (double)im);
        }
    return out;
}

int main()
{
    // constructors
    {
        complex<double> both(1.5);
        check(both.re == 1.5 && both.im == 1.5, "complex(x) sets re = im = x");
        grid3D<complex<double> > e;
        check(e.rows == 0 && e.mem == nullptr, "grid3D() is empty");
        grid3D<complex<double> > a(4, complex<double>(1, -2));
        check(a(3, 2, 1).re == 1 && a(3, 2, 1).im == -2, "grid3D(rcl, complex initial)");
        // (the index rows*cols*lay + cols*col + row stays inside a non-square
        // layer only for cols*col + row < rows*cols, grid3D_cx.hpp:295)
        grid3D<complex<double> > b(2, 3, 4, 7.0);
        check(b.rows == 2 && b.cols == 3 && b.lays == 4 && b(1, 1, 3).re == 7 && b(1, 1, 3).im == 0 &&
                  b.mem[23].re == 7,
              "grid3D(rows, cols, lays, real initial)");
        grid3D<complex<double> > c(2, 3, 4);
        check(c(1, 1, 3).re == 0 && c(1, 1, 3).im == 0, "grid3D(rows, cols, lays) is zero");
        grid3D<complex<double> > d(std::move(a));
        check(d(3, 2, 1).im == -2 && a.mem == nullptr, "grid3D move keeps the values");
        DSOFTFourierCoefficients none;
        check(none.bandwidth == 0, "DSOFTFourierCoefficients() has bandwidth 0");
    }

    // layer-wise DFT2 / IDFT2 with a complex scale against the direct DFT
    {
        const int n = 8, lays = 3;
        grid3D<complex<double> > g(n);
        unsigned s = 12345;
        for (int k = 0; k < lays; ++k)
            for (int col = 0; col < n; ++col)
                for (int row = 0; row < n; ++row) {
                    s = s * 1103515245u + 12345u;
                    const double re = (double)(s >> 8) / (1 << 24) - 0.5;
                    s = s * 1103515245u + 12345u;
                    const double im = (double)(s >> 8) / (1 << 24) - 0.5;
                    g(row, col, k) = complex<double>(re, im);
                }
        grid3D<complex<double> > h(g);
        const complex<double> scale(0.75, -0.5);
        h.layer_wise_DFT2(scale);
        double err = 0, ref = 0;
        for (int k = 0; k < n; ++k) {
            std::vector<complex<double> > f = dft2(&g(0, 0, k), n, -1);
            for (int i = 0; i < n * n; ++i) {
                const complex<double> want = f[i] * scale;
                const complex<double> got = h.mem[(size_t)n * n * k + i];
                err = std::fmax(err, std::fabs(got.re - want.re) + std::fabs(got.im - want.im));
                ref = std::fmax(ref, std::fabs(want.re) + std::fabs(want.im));
            }
        }
        printf("  layer_wise_DFT2(scale) max error %.3e (max value %.3e)\n", err, ref);
        check(err <= 1e-13 * ref, "layer_wise_DFT2 with a complex scale");
        // back: IDFT2 with 1 / (n^2 * scale) restores the grid
        const double m2 = scale.re * scale.re + scale.im * scale.im;
        h.layer_wise_IDFT2(complex<double>(scale.re / m2 / (n * n), -scale.im / m2 / (n * n)));
        double back = 0;
        for (size_t i = 0; i < (size_t)n * n * n; ++i)
            back = std::fmax(back, std::fabs(h.mem[i].re - g.mem[i].re) + std::fabs(h.mem[i].im - g.mem[i].im));
        printf("  IDFT2 round trip max error %.3e\n", back);
        check(back <= 1e-14, "layer_wise_IDFT2 undoes layer_wise_DFT2");

        // the seam under the reference's symbol name, in place, unscaled
        grid3D<complex<double> > u(g);
        uzl_fftw_layer_wise_DFT2_grid3D(n, n, n, reinterpret_cast<double*>(const_cast<complex<double>*>(u.mem)), 1);
        double serr = 0;
        std::vector<complex<double> > f0 = dft2(&g(0, 0, 2), n, -1);
        for (int i = 0; i < n * n; ++i)
            serr = std::fmax(serr, std::fabs(u.mem[(size_t)n * n * 2 + i].re - f0[i].re) +
                                       std::fabs(u.mem[(size_t)n * n * 2 + i].im - f0[i].im));
        check(serr <= 1e-13, "uzl_fftw_layer_wise_DFT2_grid3D (reference symbol)");
    }

    // stream operators
    {
        grid3D<complex<double> > g(2, complex<double>(1.25, -3));
        std::ostringstream o;
        o << g;
        const std::string t = o.str();
        check(t.find("layer[1]") != std::string::npos && t.find("1.2500 - 3.0000i") != std::string::npos,
              "operator<< of a grid");
        DSOFTFourierCoefficients fc(2);
        fc(1, -1, 1) = complex<double>(2, 0.5);
        std::ostringstream p;
        p << fc;
        const std::string u = p.str();
        check(u.find("~> [l = 1]") != std::string::npos && u.find("2.0000 + 0.5000i") != std::string::npos,
              "operator<< of the coefficients");
    }
    printf("%s\n", fails ? "FAILED" : "OK");
    return fails ? 1 : 0;
}
