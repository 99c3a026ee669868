// noalloc.hip -- s1g_rcmc_matched_filter_noalloc: the RCMC + matched-filter
// stage for callers that cannot allocate device memory.  Host code only
// (the kernels are the azimuth stage's own).
//
// The stage's scratch is its per-column geometry table (AzColumn[n],
// AzUniform, float amplitude[n]).  s1g_rcmc_matched_filter keeps it in
// context memory it allocates; here it lives in `scratch`, a file-scope
// device array: each GPU holds its own copy until the process ends, and each
// call rewrites the part it uses in full (nothing has to start at zero).
// Calls that overlap on one GPU share that GPU's copy and race.
#include <stdio.h>

#include <mutex>

#include "s1g_internal.h"

using namespace s1g;

static __device__ __align__(256) unsigned char scratch[az_table_bytes(S1G_NOALLOC_MAX_SAMPLES)];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
// Device address of `scratch` per device, from hipGetSymbolAddress with that
// device current.  In host code the array's name is a host-side stand-in,
// not a device address: only these pointers are passed to kernels.
void* g_ptr[kMaxDevices] = {};

hipError_t scratch_of(int device, void** out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess) return e;
        g_ptr[device] = p;
    }
    *out = g_ptr[device];
    return hipSuccess;
}

}  // namespace

extern "C" int s1g_rcmc_matched_filter_noalloc(s1g_ctx* c, const s1g_azimuth_desc* d, int prec, const void* in,
                                               int64_t ld_in, void* out, int64_t ld_out) {
    if (!c || !d) return S1G_EARG;
    if (prec != S1G_F32 && prec != S1G_F64) return ctx_error(c, S1G_EARG, "precision must be S1G_F32 or S1G_F64");
    if (!in || !out || in == out) return ctx_error(c, S1G_EARG, "rcmc/matched filter is out-of-place");
    if (ld_in < d->n_samples || ld_out < d->n_samples) return ctx_error(c, S1G_EARG, "leading dimension too small");
    if (d->n_lines <= 1 || d->n_samples <= 0) return ctx_error(c, S1G_EARG, "azimuth stage needs >= 2 lines");
    if (!d->slant_ranges || !d->v_rel || !d->f_dc || !d->t_corr)
        return ctx_error(c, S1G_EARG, "null geometry array");
    if (d->n_samples > S1G_NOALLOC_MAX_SAMPLES) {
        char buf[128];
        snprintf(buf, sizeof buf, "n_samples %d above the no-alloc maximum of %d", d->n_samples,
                 S1G_NOALLOC_MAX_SAMPLES);
        return ctx_error(c, S1G_EARG, buf);
    }
    void* table = nullptr;
    hipError_t e = scratch_of(ctx_device(c), &table);
    if (e != hipSuccess) {
        char buf[160];
        snprintf(buf, sizeof buf, "no-alloc column table address (hipGetSymbolAddress): %s", hipGetErrorString(e));
        return ctx_error(c, S1G_EHIP, buf);
    }
    AzColumn* cols = (AzColumn*)table;
    hipStream_t s = ctx_stream(c);
    int rc = launch_az_columns(s, *d, cols);
    if (rc) return ctx_error(c, S1G_EHIP, hipGetErrorString((hipError_t)rc));
    // the tuner's usual picks for the IW bursts: the FP32 linear-geometry
    // kernel (runs of 8), the FP64 256 x 4 tile
    const int v = ctx_rcmc_variant(c, prec, d->n_lines, d->n_samples, d->apply_rcmc, prec == S1G_F32 ? 6 : 1);
    rc = launch_rcmc_mf(s, prec, *d, cols, in, ld_in, out, ld_out, v);
    if (rc) return ctx_error(c, S1G_EHIP, hipGetErrorString((hipError_t)rc));
    return S1G_OK;
}
