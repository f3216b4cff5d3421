// Forward transform that allocates nothing: pfs_dsoft_f64_dev_noalloc.
//
// The plain forward transform (grid -> FFT rows -> W -> FFT columns -> pencils
// G -> DWT -> coefficients) needs two work buffers of N^3 complex128 each.
// pfs_dsoft_f64_dev takes them from its context, which allocates them on first
// use.  For callers that cannot allocate device memory (code under HIP graph
// capture, serving loops, harnesses that forbid hipMalloc) this entry point
// takes them from one file-scope device array, `scratch`, sized for 2B <= 128:
// W then G, 2 * 128^3 complex128 = 64 MiB.  Every GPU the library runs on
// holds its own copy of the array from the library's first use on that device
// until the process ends.
//
// In host code `scratch` names a host-side stand-in, not device memory: its
// device address is looked up once per device (hipGetSymbolAddress, at context
// creation, with that device current) and only that address is handed to
// kernels.  Every call on one GPU shares that GPU's array, whichever context
// makes it: calls that overlap in time on one GPU race on it.  Both buffers
// are written in full before they are read (the row pass writes all of W, the
// column pass all of G), so nothing needs clearing between calls.
#include "pfs_internal.h"
#include "pfsoft_amd.h"

#include <cstdio>
#include <mutex>
#include <string>

namespace {

constexpr int NOALLOC_MAX_B = 64;
constexpr size_t NOALLOC_N3 = (size_t)(2 * NOALLOC_MAX_B) * (2 * NOALLOC_MAX_B) * (2 * NOALLOC_MAX_B);
static __device__ __align__(256) double2 scratch[2 * NOALLOC_N3];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void* g_scratch[MAX_DEVICES];             // device address of `scratch`, per device
hipError_t g_setup_error[MAX_DEVICES];   // the lookup's failure, if it failed

}  // namespace

namespace pfs {

void noalloc_setup(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return;
    std::lock_guard<std::mutex> g(g_mu);
    if (g_scratch[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        g_setup_error[device] = e != hipSuccess ? e : hipErrorInvalidSymbol;
        return;
    }
    g_scratch[device] = p;
    g_setup_error[device] = hipSuccess;
}

}  // namespace pfs

extern "C" int pfs_dsoft_f64_dev_noalloc(pfs_context* ctx, int bandwidth, const void* grid_dev, void* coeffs_dev,
                                         void* stream)
{
    if (!ctx || !grid_dev || !coeffs_dev) return pfs::set_error(PFS_ERR_ARG, "pfs_dsoft_f64_dev_noalloc: null argument");
    if (bandwidth < 1 || bandwidth > NOALLOC_MAX_B) {
        char msg[160];
        snprintf(msg, sizeof(msg),
                 "pfs_dsoft_f64_dev_noalloc: bandwidth %d outside 1..%d (its scratch holds 2B <= %d, 64 MiB)",
                 bandwidth, NOALLOC_MAX_B, 2 * NOALLOC_MAX_B);
        return pfs::set_error(PFS_ERR_UNSUPPORTED, msg);
    }
    const int dev = pfs::context_device(ctx);
    void* base = nullptr;
    hipError_t why = hipSuccess;
    {
        std::lock_guard<std::mutex> g(g_mu);
        if (dev >= 0 && dev < MAX_DEVICES) {
            base = g_scratch[dev];
            why = g_setup_error[dev];
        }
    }
    if (!base) {
        const std::string msg = std::string("pfs_dsoft_f64_dev_noalloc: no scratch address on this device (") +
                                (why != hipSuccess ? hipGetErrorString(why) : "context not set up") + ")";
        return pfs::set_error(PFS_ERR_DEVICE, msg.c_str());
    }
    const size_t n3 = (size_t)(2 * bandwidth) * (2 * bandwidth) * (2 * bandwidth);  // <= NOALLOC_N3
    double2* W = static_cast<double2*>(base);
    double2* G = W + n3;
    return pfs::dsoft_prepared(ctx, bandwidth, static_cast<const double2*>(grid_dev), static_cast<double2*>(coeffs_dev),
                               W, G, static_cast<hipStream_t>(stream));
}
