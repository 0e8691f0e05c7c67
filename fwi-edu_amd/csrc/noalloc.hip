// An entry point that allocates no device memory: compute_model_perturbation
// (optimization.c:178-298) on arrays the caller already holds on the device,
// for callers that cannot allocate (code that forbids hipMalloc, serving loops
// that keep allocation out of their steady state).
//
// The kernels' scratch (min/max block partials, the 1001 histogram counters,
// the FP64 block partials of the slowness sum and the results) lives in the
// file-scope device array `scratch` below: SCRATCH_BYTES per GPU, held for the
// life of the process.  The array's name is a host-side stand-in, not a device
// address; its address on a device comes from hipGetSymbolAddress, taken once
// per device with that device current, and only that address is handed to
// kernels.  Every call on one GPU shares that GPU's array: calls that overlap
// on one GPU race.  The counters are cleared by a fill kernel in every call;
// the partials are overwritten whole.
#include <hip/hip_runtime.h>
#include <limits.h>

#include <mutex>

#include "engine.hpp"

using namespace fwiedu;

namespace {

constexpr size_t SCRATCH_BYTES = 32768;
constexpr int MAX_DEVICES = 64;

__device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

std::mutex g_mu;
void *g_ptr[MAX_DEVICES] = {};

constexpr size_t round256(size_t n) { return (n + 255) / 256 * 256; }

// Device address of this device's scratch, looked up on the first call.
unsigned char *scratch_on(int dev)
{
    if (dev < 0 || dev >= MAX_DEVICES)
        fail("<error>: fwiedu_model_perturbation_noalloc: device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            fail("<error>: fwiedu_model_perturbation_noalloc: hipGetSymbolAddress failed: %s",
                 hipGetErrorString(e));
        g_ptr[dev] = p;
    }
    return static_cast<unsigned char *>(g_ptr[dev]);
}

}  // namespace

extern "C" float fwiedu_model_perturbation_noalloc(float fraction, const float *slowness,
                                                   const float *gradient, float *out, int len)
{
    if (len <= 0) fail("<error>: fwiedu_model_perturbation_noalloc: len %d must be positive", len);
    if (!slowness || !gradient || !out)
        fail("<error>: fwiedu_model_perturbation_noalloc: NULL array");
    if (gradient == out)
        fail("<error>: fwiedu_model_perturbation_noalloc: gradient and out may not alias");
    ensure_ready();
    Ctx &c = ctx();
    int cur = -1;
    HIPCHECK(hipGetDevice(&cur));
    if (cur != c.device)
        fail("<error>: fwiedu_model_perturbation_noalloc: device %d is current, the library "
             "runs on device %d",
             cur, c.device);
    // layout of the scratch: min/max partials and result, histogram, FP64
    // partials and sums
    const size_t mm_bytes = round256((2 * 512 + 2) * sizeof(float));
    const size_t hist_bytes = round256(1001 * sizeof(int));
    const size_t red_bytes = round256((reduce_scratch_doubles() + 2) * sizeof(double));
    if (mm_bytes + hist_bytes + red_bytes > SCRATCH_BYTES)
        fail("<error>: fwiedu_model_perturbation_noalloc: scratch of %zu bytes too small (%zu "
             "needed)",
             SCRATCH_BYTES, mm_bytes + hist_bytes + red_bytes);
    unsigned char *base = scratch_on(c.device);
    float *mm = reinterpret_cast<float *>(base);
    int *hist = reinterpret_cast<int *>(base + mm_bytes);
    double *red = reinterpret_cast<double *>(base + mm_bytes + hist_bytes);
    const float s = model_perturbation_with(fraction, slowness, gradient, out, len, mm, hist, red);
    sync();
    return s;
}
