// pfsoft_amd.hpp -- drop-in C++ interface for the SO(3) Fourier transform
// path of PFSOFTlib, computed on an MI355X through the C ABI of pfsoft_amd.h.
//
// Code written against the reference's
//     pfsoft::FourierTransforms::DSOFT(grid3D<complex<double>>, DSOFTFourierCoefficients&, int)
//     pfsoft::FourierTransforms::IDSOFT(const DSOFTFourierCoefficients&, grid3D<complex<double>>&, int)
// (include/PFSOFTlib_headers/fn_fourier_transforms.hpp:75,78) compiles against
// this header unchanged for those calls: the types keep the reference's
// members and index contracts (grid3D_cx.hpp:47-83,293-303;
// dsoft_fourier_coefficients.hpp:50-72; complex.hpp:60-97), and the checks
// keep its warn-and-return / print-and-exit behaviour
// (cmake_aux/compiler_config.hpp.in:93-120).  The threads argument is
// accepted and ignored.  Link with -lpfsoft_amd.
#ifndef PFSOFT_AMD_HPP
#define PFSOFT_AMD_HPP

#include "pfsoft_amd.h"

#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

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

template <typename T>
struct complex {
    T re, im;
    complex() : re(0), im(0) {}
    complex(T r, T i) : re(r), im(i) {}
};

// Only the complex<double> grid is on the path.
template <typename T>
struct grid3D;

template <>
struct grid3D<complex<double> > {
    typedef double pod_type;
    const size_t rows, cols, lays;
    const complex<double>* mem;  // element (row, col, lay) at rows*cols*lay + cols*col + row

    explicit grid3D(const size_t& rcl) : rows(rcl), cols(rcl), lays(rcl), mem(alloc(rcl * rcl * rcl)) {}
    grid3D(const size_t& r, const size_t& c, const size_t& l) : rows(r), cols(c), lays(l), mem(alloc(r * c * l)) {}
    grid3D(const grid3D& o) : rows(o.rows), cols(o.cols), lays(o.lays), mem(alloc(o.rows * o.cols * o.lays))
    {
        std::memcpy(const_cast<complex<double>*>(mem), o.mem, rows * cols * lays * sizeof(complex<double>));
    }
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

private:
    static complex<double>* alloc(size_t n) { return new complex<double>[n]; }  // zero-initialised
};

// Coefficients f^l_{M,M'}, packed per degree (pfsoft_amd.h layout): the
// reference's per-degree column-major (2l+1)^2 matrices laid end to end.
struct DSOFTFourierCoefficients {
    const int bandwidth;

    explicit DSOFTFourierCoefficients(int bandlimit)
        : bandwidth(bandlimit), mem(new complex<double>[pfs_coeff_count(bandlimit)])
    {
    }
    DSOFTFourierCoefficients(const DSOFTFourierCoefficients& o)
        : bandwidth(o.bandwidth), mem(new complex<double>[pfs_coeff_count(o.bandwidth)])
    {
        std::memcpy(mem, o.mem, pfs_coeff_count(bandwidth) * sizeof(complex<double>));
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

// src/fn_dsoft.cpp:76.  The reference takes the grid by value so the
// caller's grid is never modified; a const reference keeps that and saves
// the copy, and callers compile unchanged.
inline void DSOFT(const grid3D<complex<double> >& sample, DSOFTFourierCoefficients& fc,
                  int threads = PFSOFT_MAX_THREADS)
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
