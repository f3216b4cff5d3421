// pfsoft_amd.hpp -- drop-in C++ interface for the SO(3) Fourier transform
// path of PFSOFTlib, computed on an MI355X through the C ABI of pfsoft_amd.h.
//
// Code written against the reference's
//     pfsoft::FourierTransforms::DSOFT(grid3D<complex<double>>, DSOFTFourierCoefficients&, int)
//     pfsoft::FourierTransforms::IDSOFT(const DSOFTFourierCoefficients&, grid3D<complex<double>>&, int)
// (include/PFSOFTlib_headers/fn_fourier_transforms.hpp:75,78) compiles against
// this header unchanged for those calls, and the two functions have the
// reference's exact types.  The types keep the reference's members,
// constructors and index contracts (grid3D_cx.hpp:47-83,293-303;
// dsoft_fourier_coefficients.hpp:50-72; complex.hpp:60-97), including the
// layer-wise DFT2/IDFT2 members of the grid (grid3D_cx.hpp:81-82) and the
// stream operators (grid3D_cx.hpp:430; dsoft_fourier_coefficients.hpp:71).
// The checks keep the reference's warn-and-return / print-and-exit behaviour
// (cmake_aux/compiler_config.hpp.in:93-120).  The threads argument is
// accepted and ignored.  Link with -lpfsoft_amd.
#ifndef PFSOFT_AMD_HPP
#define PFSOFT_AMD_HPP

#include "pfsoft_amd.h"

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <mutex>
#include <ostream>
#include <sstream>
#include <string>

#ifndef PFSOFT_MAX_THREADS
#define PFSOFT_MAX_THREADS 1
#endif

#define PFSOFT_AMD_WARN_RET(cond, msg)                                                                     \
    do {                                                                                                   \
        if (cond) {                                                                                        \
            fprintf(stderr, "** [PFSOFTLib warning] %s:%d:%s(): %s **\n", __FILE__, __LINE__, __func__, msg); \
            return;                                                                                        \
        }                                                                                                  \
    } while (0)

namespace pfsoft {

// Two contiguous T, real part first: the layout the reference relies on when
// it reinterprets a complex<double> array as double* (grid3D_cx.hpp:321).
template <typename T>
struct complex {
    typedef T pod_type;
    T re, im;
    complex() : re(0), im(0) {}
    complex(T r, T i) : re(r), im(i) {}
    // As in the reference (complex.hpp:133-138), one value sets BOTH parts.
    complex(const T& real_imag) : re(real_imag), im(real_imag) {}

    T abs() const { return std::sqrt(re * re + im * im); }
    void polar(const T& rho, const T& theta = 0)
    {
        re = rho * std::cos(theta);
        im = rho * std::sin(theta);
    }

    complex& operator+=(const complex& o) { re += o.re; im += o.im; return *this; }
    complex& operator-=(const complex& o) { re -= o.re; im -= o.im; return *this; }
    complex& operator*=(const complex& o)
    {
        const T r = re * o.re - im * o.im;
        im = re * o.im + im * o.re;
        re = r;
        return *this;
    }
    complex& operator*=(T s) { re *= s; im *= s; return *this; }
    bool operator==(const complex& o) const { return re == o.re && im == o.im; }
    bool operator!=(const complex& o) const { return !(*this == o); }
};
template <typename T> inline complex<T> operator+(complex<T> a, const complex<T>& b) { return a += b; }
template <typename T> inline complex<T> operator-(complex<T> a, const complex<T>& b) { return a -= b; }
template <typename T> inline complex<T> operator*(complex<T> a, const complex<T>& b) { return a *= b; }
template <typename T> inline complex<T> operator*(complex<T> a, T s) { return a *= s; }
template <typename T> inline complex<T> operator*(T s, complex<T> a) { return a *= s; }

namespace detail {
// "re + imi" / "re - |im|i" in fixed or scientific notation, 4 digits, right
// aligned: the element text of the reference's grid and matrix printers.
template <typename T>
inline std::string element_text(const complex<T>& v, bool sci)
{
    std::ostringstream s;
    if (sci) s << std::scientific;
    else s << std::fixed;
    s << std::setprecision(4) << v.re << (v.im < 0 ? " - " : " + ") << (v.im == 0 ? T(0) : std::fabs(v.im)) << "i";
    return s.str();
}
// Column width (and notation) from the largest magnitude printed.
template <typename T>
inline int element_width(const complex<T>* v, size_t n, bool* sci)
{
    double m = 0;
    for (size_t i = 0; i < n; ++i) m = std::max(m, std::max(std::fabs((double)v[i].re), std::fabs((double)v[i].im)));
    *sci = m >= 1000;
    return m >= 1000 ? 28 : m >= 100 ? 24 : m >= 10 ? 22 : 20;
}
}  // namespace detail

// Only the complex<double> grid is on the path.
template <typename T>
struct grid3D;

template <>
struct grid3D<complex<double> > {
    typedef double pod_type;
    const size_t rows, cols, lays;
    const complex<double>* mem;  // element (row, col, lay) at rows*cols*lay + cols*col + row

    // grid3D_cx.hpp:61-70.  New grids are zero (complex.hpp:117-120).
    grid3D() : rows(0), cols(0), lays(0), mem(nullptr) {}
    explicit grid3D(const size_t& rcl) : rows(rcl), cols(rcl), lays(rcl), mem(alloc(rcl * rcl * rcl)) {}
    grid3D(const size_t& rcl, const complex<double>& initial) : grid3D(rcl) { fill(initial); }
    grid3D(const size_t& rcl, const double& initial) : grid3D(rcl) { fill(complex<double>(initial, 0)); }
    grid3D(const size_t& r, const size_t& c, const size_t& l) : rows(r), cols(c), lays(l), mem(alloc(r * c * l)) {}
    grid3D(const size_t& r, const size_t& c, const size_t& l, const complex<double>& initial) : grid3D(r, c, l)
    {
        fill(initial);
    }
    grid3D(const size_t& r, const size_t& c, const size_t& l, const double& initial) : grid3D(r, c, l)
    {
        fill(complex<double>(initial, 0));
    }
    grid3D(const grid3D& o) : rows(o.rows), cols(o.cols), lays(o.lays), mem(alloc(o.rows * o.cols * o.lays))
    {
        if (size()) std::memcpy(const_cast<complex<double>*>(mem), o.mem, size() * sizeof(complex<double>));
    }
    // Takes the storage; the source is left empty (its destructor frees nothing).
    grid3D(grid3D&& o) : rows(o.rows), cols(o.cols), lays(o.lays), mem(o.mem) { o.mem = nullptr; }
    // The reference's assignments write its const dimensions and cannot be
    // instantiated (grid3D_cx.hpp:236-289); here they are not declared usable.
    grid3D& operator=(const grid3D&) = delete;
    ~grid3D() { delete[] mem; }

    complex<double>& operator()(const size_t& row, const size_t& col, const size_t& lay)
    {
        return const_cast<complex<double>*>(mem)[rows * cols * lay + cols * col + row];
    }
    const complex<double>& operator()(const size_t& row, const size_t& col, const size_t& lay) const
    {
        return mem[rows * cols * lay + cols * col + row];
    }

    // grid3D_cx.hpp:309-366 / 370-427: unnormalised 2D DFT (forward sign -1,
    // backward +1) of each of the lays slices of cols x rows values, in place,
    // then every value times scale.  On the device; square slices only.
    void layer_wise_DFT2(const complex<double>& scale = complex<double>(1, 0), int threads = 1)
    {
        transform(-1, scale, threads);
    }
    void layer_wise_IDFT2(const complex<double>& scale = complex<double>(1, 0), int threads = 1)
    {
        transform(+1, scale, threads);
    }

private:
    static complex<double>* alloc(size_t n) { return n ? new complex<double>[n] : nullptr; }  // zero-initialised
    size_t size() const { return rows * cols * lays; }
    void fill(const complex<double>& v) { std::fill(const_cast<complex<double>*>(mem), const_cast<complex<double>*>(mem) + size(), v); }
    void transform(int sign, const complex<double>& scale, int threads)
    {
        if (!size()) return;
        if (pfs_layer_wise_dft2((int)cols, (int)rows, (int)lays, reinterpret_cast<double*>(const_cast<complex<double>*>(mem)),
                                sign, scale.re, scale.im, threads) != PFS_OK) {
            fprintf(stderr, "** [PFSOFTLib error]   layer_wise_%sDFT2: %s **\n", sign < 0 ? "" : "I", pfs_last_error());
            exit(EXIT_FAILURE);
        }
    }
};

// Prints every layer as "layer[z]" and its rows x cols values, row by row
// (the reference's grid printer, grid3D_cx.hpp:430-529).
inline std::ostream& operator<<(std::ostream& o, const grid3D<complex<double> >& g)
{
    bool sci = false;
    const int w = detail::element_width(g.mem, g.rows * g.cols * g.lays, &sci);
    o << std::endl;
    for (size_t z = 0; z < g.lays; ++z) {
        o << "layer[" << z << "]" << std::endl;
        for (size_t x = 0; x < g.rows; ++x) {
            for (size_t y = 0; y < g.cols; ++y) o << std::setfill(' ') << std::right << std::setw(w) << detail::element_text(g(x, y, z), sci);
            o << std::endl;
        }
        o << std::endl;
    }
    return o;
}

// Coefficients f^l_{M,M'}, packed per degree (pfs_coeff_index): the
// reference's per-degree column-major (2l+1)^2 matrices laid end to end.
struct DSOFTFourierCoefficients {
    const int bandwidth;

    // dsoft_fourier_coefficients.cpp:31-35: an empty container.
    DSOFTFourierCoefficients() : bandwidth(0), mem(nullptr) {}
    explicit DSOFTFourierCoefficients(int bandlimit)
        : bandwidth(bandlimit), mem(bandlimit > 0 ? new complex<double>[pfs_coeff_count(bandlimit)] : nullptr)
    {
    }
    // Deep copy (the reference's container has none and would double-free a
    // copy, dsoft_fourier_coefficients.cpp:51,65).
    DSOFTFourierCoefficients(const DSOFTFourierCoefficients& o)
        : bandwidth(o.bandwidth), mem(o.bandwidth > 0 ? new complex<double>[pfs_coeff_count(o.bandwidth)] : nullptr)
    {
        if (bandwidth > 0) std::memcpy(mem, o.mem, pfs_coeff_count(bandwidth) * sizeof(complex<double>));
    }
    DSOFTFourierCoefficients& operator=(const DSOFTFourierCoefficients&) = delete;
    ~DSOFTFourierCoefficients() { delete[] mem; }

    complex<double>& operator()(const int& l, const int& M, const int& Mp) { return mem[index(l, M, Mp)]; }
    const complex<double>& operator()(const int& l, const int& M, const int& Mp) const
    {
        return mem[index(l, M, Mp)];
    }

    complex<double>* data() { return mem; }
    const complex<double>* data() const { return mem; }

    friend std::ostream& operator<<(std::ostream& o, const DSOFTFourierCoefficients& fc);

private:
    complex<double>* mem;

    long index(int l, int M, int Mp) const
    {
        // src/dsoft_fourier_coefficients.cpp:82: print and exit on a bad order
        if (M > l || Mp > l || M < -l || Mp < -l || l < 0 || l >= bandwidth) {
            fprintf(stderr, "** [PFSOFTLib error]   %s:%d:%s(): %s **\n", __FILE__, __LINE__, "operator()",
                    "illegal parameter for DSOFTFourierCoefficients. Condition |M|,|Mp| <= l violated.");
            exit(EXIT_FAILURE);
        }
        return pfs_coeff_index(l, M, Mp);
    }
};

// Every degree block as its (2l+1) x (2l+1) matrix, rows M = 0, 1, ..., -1 and
// columns M' in the same order, under the reference's block heading
// (dsoft_fourier_coefficients.cpp:124-137).
inline std::ostream& operator<<(std::ostream& o, const DSOFTFourierCoefficients& fc)
{
    o << std::endl;
    for (int l = 0; l < fc.bandwidth; ++l) {
        o << "DSOFTFourierCoefficients[M_{0,1,2,...,-2,-1} x M'_{0,1,2,...,-2,-1}] ~> [l = " << l << "]" << std::endl;
        const int n = 2 * l + 1;
        const complex<double>* blk = fc.mem + pfs_coeff_index(l, 0, 0);
        bool sci = false;
        const int w = detail::element_width(blk, (size_t)n * n, &sci);
        o << std::endl;
        for (int i = 0; i < n; ++i) {
            for (int j = 0; j < n; ++j) o << std::setfill(' ') << std::right << std::setw(w) << detail::element_text(blk[i + (long)n * j], sci);
            o << std::endl;
        }
        o << std::endl;
    }
    return o;
}

namespace detail {
inline pfs_context* default_context()
{
    static std::once_flag once;
    static pfs_context* ctx = nullptr;
    std::call_once(once, [] {
        if (pfs_ctx_create(0, 0, &ctx) != PFS_OK) {
            fprintf(stderr, "** [PFSOFTLib error]   pfsoft_amd: %s **\n", pfs_last_error());
            exit(EXIT_FAILURE);
        }
    });
    return ctx;
}

inline void check(int rc, const char* what)
{
    if (rc != PFS_OK) {
        fprintf(stderr, "** [PFSOFTLib error]   %s: %s **\n", what, pfs_last_error());
        exit(EXIT_FAILURE);
    }
}
}  // namespace detail

namespace FourierTransforms {

// src/fn_dsoft.cpp:76, the reference's exact type: the grid is taken by value,
// so the caller's grid is never modified (a temporary is moved in).
inline void DSOFT(grid3D<complex<double> > sample, DSOFTFourierCoefficients& fc, int threads = PFSOFT_MAX_THREADS)
{
    (void)threads;
    PFSOFT_AMD_WARN_RET(sample.rows != sample.cols || sample.rows != sample.lays,
                        "all DSOFT sample grid dimensions should be equal.");
    PFSOFT_AMD_WARN_RET(sample.rows & 1, "DSOFT sample grid dimensions are not even.");
    const int bandwidth = static_cast<int>(sample.cols / 2);
    PFSOFT_AMD_WARN_RET(bandwidth != fc.bandwidth,
                        "DSOFT Fourier coefficients container bandwidth does not match to sample grid bandwidth.");
    detail::check(pfs_dsoft_f64(detail::default_context(), bandwidth, reinterpret_cast<const double*>(sample.mem),
                                reinterpret_cast<double*>(fc.data())),
                  "DSOFT");
}

// src/fn_idsoft.cpp:72.  Written in place; entries with row or column index
// B keep the caller's values for the inverse FFT, as in the reference.
inline void IDSOFT(const DSOFTFourierCoefficients& fc, grid3D<complex<double> >& synthesis,
                   int threads = PFSOFT_MAX_THREADS)
{
    (void)threads;
    PFSOFT_AMD_WARN_RET(synthesis.rows != synthesis.cols || synthesis.rows != synthesis.lays,
                        "all IDSOFT synthesis grid dimensions should be equal.");
    PFSOFT_AMD_WARN_RET(synthesis.rows & 1, "IDSOFT synthesis grid dimensions are not even.");
    const int bandwidth = static_cast<int>(synthesis.cols / 2);
    PFSOFT_AMD_WARN_RET(bandwidth != fc.bandwidth,
                        "IDSOFT Fourier coefficients container bandwidth does not match to synthesis grid bandwidth.");
    detail::check(pfs_idsoft_f64(detail::default_context(), bandwidth, reinterpret_cast<const double*>(fc.data()),
                                 reinterpret_cast<double*>(const_cast<complex<double>*>(synthesis.mem))),
                  "IDSOFT");
}

}  // namespace FourierTransforms
}  // namespace pfsoft

#undef PFSOFT_AMD_WARN_RET

#endif  // PFSOFT_AMD_HPP
