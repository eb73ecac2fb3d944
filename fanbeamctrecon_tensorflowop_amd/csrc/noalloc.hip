// noalloc.hip -- fbct_fbp_noalloc: filtered backprojection with its filtered sinograms
// in a device array the library holds, for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// fbct_fbp takes a caller workspace of one filtered sinogram per slice.  Here the
// filtered sinograms go to `scratch`, a file-scope device array of kScratchBytes: each
// GPU the library runs on holds its own copy until the process ends.  Slices are taken
// in chunks whose filtered sinograms fit it; each chunk is the same two launches as
// fbct_fbp (MFMA ramp filter into the scratch, 1/L^2-weighted backprojection out of it),
// stream-ordered, so a chunk's filter runs after the previous chunk's backprojection.
// The filter writes every element of a chunk's rows before the backprojection reads
// them, so nothing in the array has to start at zero.  Calls that overlap on one GPU
// (on different streams or host threads) share the array and race.

#include <hip/hip_runtime.h>

#include <mutex>

#include "fbct_internal.hpp"

namespace {

constexpr size_t kScratchBytes = FBCT_FBP_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

// Filtered sinograms (FP32) of the chunk in flight.  Host code never uses this name as
// an address: it is a host-side stand-in; the device address comes from
// hipGetSymbolAddress, per device.
__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per device with that
// device current (the first call on a device); later calls only read the table.
fbct_status scratch_ptr(void **out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fbct::fail(FBCT_ERR_DEVICE, "hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return fbct::fail(FBCT_ERR_UNSUPPORTED, "device index %d", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fbct::fail(FBCT_ERR_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return FBCT_OK;
}

}  // namespace

extern "C" {

fbct_status fbct_fbp_noalloc(fbct_plan *plan, const float *sino, float *image, int n_slices, void *stream)
{
    const char *what = "fbct_fbp_noalloc";
    if (!plan) return fbct::fail(FBCT_ERR_INVALID_ARGUMENT, "%s: null plan", what);
    if (n_slices < 0) return fbct::fail(FBCT_ERR_INVALID_ARGUMENT, "%s: negative slice count %d", what, n_slices);
    if (n_slices == 0) return FBCT_OK;
    if (!sino || !image) return fbct::fail(FBCT_ERR_INVALID_ARGUMENT, "%s: null data pointer", what);
    if (n_slices > 65535) return fbct::fail(FBCT_ERR_UNSUPPORTED, "%s: at most 65535 slices per call", what);
    const size_t sino_elems = (size_t)plan->g.nViews * plan->g.nCols;
    const size_t img_elems = (size_t)plan->g.nX * plan->g.nY;
    const size_t per_slice = sino_elems * sizeof(float);
    if (per_slice > kScratchBytes)
        return fbct::fail(FBCT_ERR_WORKSPACE,
                          "%s: a slice's filtered sinogram takes %zu bytes, more than the %zu-byte scratch array "
                          "(use fbct_fbp with a workspace)",
                          what, per_slice, kScratchBytes);
    fbct_status s = fbct::check_disjoint(image, (size_t)n_slices * img_elems * sizeof(float), sino,
                                         (size_t)n_slices * per_slice, what, "image", "sinogram");
    if (s != FBCT_OK) return s;
    const fbct::DeviceTables *t = nullptr;
    if ((s = fbct::resident_tables(plan, &t)) != FBCT_OK) return s;
    void *q = nullptr;
    if ((s = scratch_ptr(&q)) != FBCT_OK) return s;
    hipStream_t st = static_cast<hipStream_t>(stream);
    // chunks of whole 16-slice groups where they fit (the backprojection's largest group)
    int chunk = (int)(kScratchBytes / per_slice);
    if (chunk >= 16) chunk -= chunk % 16;
    for (int s0 = 0; s0 < n_slices; s0 += chunk) {
        const int c = n_slices - s0 < chunk ? n_slices - s0 : chunk;
        float *qf = static_cast<float *>(q);
        if ((s = fbct::launch_ramp_filter_mfma(*plan, *t, sino + (size_t)s0 * sino_elems, qf,
                                               (int64_t)c * plan->g.nViews, st)) != FBCT_OK)
            return s;
        if ((s = fbct::launch_bp(*plan, *t, qf, image + (size_t)s0 * img_elems, c, true, st)) != FBCT_OK) return s;
    }
    return FBCT_OK;
}

}  // extern "C"
