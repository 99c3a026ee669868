// fft_global.hip -- out-of-place Stockham FFT passes on global memory.
//
// Used where a transform does not fit one workgroup's LDS and is off the
// streaming path: the FP64 FFT of the overlap-save range filter (once per
// chirp geometry) and the full-length ECCM detection transforms.  Pass with
// radix R after Ns = product of earlier radices (Stockham autosort):
//   out[(t/Ns) Ns R + j + r Ns] = DFT_R( in[t + r n/R] * w_{Ns R}^{j r} )_r,
//   j = t mod Ns, t < n/R.
#include "s1g_internal.h"

namespace s1g {

template <typename T, int R, bool INV>
__global__ __launch_bounds__(256) void stockham_pass_kernel(const cpx<T>* __restrict__ in, cpx<T>* __restrict__ out,
                                                            int n, int Ns, int batch, int64_t dist,
                                                            const cpx<T>* __restrict__ tw) {
    const int nR = n / R;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)nR * batch) return;
    const int b = (int)(idx / nR);
    const int t = (int)(idx - (int64_t)b * nR);
    const int j = t % Ns;
    const int tstep = n / (Ns * R);
    const cpx<T>* src = in + b * dist;
    cpx<T>* dst = out + b * dist;
    cpx<T> a[R];
#pragma unroll
    for (int r = 0; r < R; ++r) a[r] = src[t + r * nR];
    if (j) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
            cpx<T> w = tw[j * r * tstep];
            a[r] = INV ? cmulc(a[r], w) : cmul(a[r], w);
        }
    }
    Dft<T, R, INV>::run(a);
    const int base = (t / Ns) * Ns * R + j;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[base + r * Ns] = a[r];
}

template <typename T, bool INV>
static void launch_pass(hipStream_t s, int R, const cpx<T>* in, cpx<T>* out, int n, int Ns, int batch, int64_t dist,
                        const cpx<T>* tw) {
    const int64_t total = (int64_t)(n / R) * batch;
    dim3 block(256), grid((unsigned)((total + 255) / 256));
    switch (R) {
#define S1G_CASE(r)                                                                                     \
    case r:                                                                                             \
        hipLaunchKernelGGL((stockham_pass_kernel<T, r, INV>), grid, block, 0, s, in, out, n, Ns, batch, \
                           dist, tw);                                                                   \
        break;
        S1G_RADIX_LIST(S1G_CASE)
#undef S1G_CASE
        default:
            break;
    }
}

template <typename T>
static int run_fft_global(hipStream_t s, const DevFft& f, const void* in_, void* out_, void* tmp_, int batch,
                          int64_t dist, int inverse) {
    const cpx<T>* in = (const cpx<T>*)in_;
    cpx<T>* bufs[2];
    // choose the ping-pong order so the last pass lands in `out`
    const int np = f.plan.npass;
    bufs[(np - 1) % 2] = (cpx<T>*)out_;
    bufs[np % 2] = (cpx<T>*)tmp_;
    const cpx<T>* src = in;
    int Ns = 1;
    for (int p = 0; p < np; ++p) {
        cpx<T>* dst = bufs[p % 2];
        const int R = f.plan.radix[p];
        if (inverse)
            launch_pass<T, true>(s, R, src, dst, f.plan.n, Ns, batch, dist, tw_of<T>(f));
        else
            launch_pass<T, false>(s, R, src, dst, f.plan.n, Ns, batch, dist, tw_of<T>(f));
        src = dst;
        Ns *= R;
    }
    if (np == 0) (void)hipMemcpyAsync(out_, in_, sizeof(cpx<T>) * (size_t)dist * batch, hipMemcpyDeviceToDevice, s);
    return (int)hipGetLastError();
}

int launch_fft_global(hipStream_t s, int prec, const DevFft& f, const void* in, void* out, void* tmp, int batch,
                      int64_t dist, int inverse) {
    if (prec == S1G_F32) return run_fft_global<float>(s, f, in, out, tmp, batch, dist, inverse);
    return run_fft_global<double>(s, f, in, out, tmp, batch, dist, inverse);
}

}  // namespace s1g
