// eccm.hip -- RFI detection and notch (ECCM), out of place.
//
// Reference: eccm (signal_processing.cpp:913-1004), called before range
// compression with fft_size 128, stride 64 (s1_decoder.cpp:470-474):
//   * a line is processed iff max_k |FFT_N(line)[k]| >= detection threshold;
//   * num_ffts = floor((N / fft_size) * (fft_size / stride)) windows, window i
//     starts at i*stride and ends at start + fft_size, or at N - 1 when that
//     overruns N; each is zero-padded to fft_size and transformed;
//   * per bin, |gradient| along frequency (central difference inside, one
//     sided at both ends); bins with |gradient| >= mitigation threshold are
//     zeroed; inverse transform (x 1/fft_size);
//   * windows are written back in order, so a sample keeps the value of the
//     last window covering it; samples no window covers keep their input.
//
// Detection: the N-point spectrum maximum is found as the maximum over r
// interleaved sub-spectra X[r k + q] = FFT_{N/r}(y_q)[k] with
// y_q[n] = sum_m x[n + m N/r] w_r^{q m} w_N^{q n}, each of which fits LDS.
// Notch: one workgroup per (line, run of windows); unflagged lines are copied.
#include "s1g_internal.h"

namespace s1g {

template <typename T, int NT>
__global__ __launch_bounds__(NT) void eccm_detect_kernel(const cpx<T>* __restrict__ in, int64_t ld, int N, int r,
                                                         FftPlan sub, const cpx<T>* __restrict__ twN,
                                                         const cpx<T>* __restrict__ twSub,
                                                         const int32_t* __restrict__ permSub,
                                                         T* __restrict__ partial) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<T>* buf = (cpx<T>*)smem_raw;
    __shared__ T red[NT / 64];
    const int line = blockIdx.x / r, q = blockIdx.x % r;
    const int tid = threadIdx.x;
    const int n1 = N / r;
    const cpx<T>* row = in + (int64_t)line * ld;
    for (int n = tid; n < n1; n += NT) {
        cpx<T> y = mk((T)0, (T)0);
        for (int m = 0; m < r; ++m) {
            const cpx<T> x = row[n + m * n1];
            const int k = (int)(((int64_t)q * m % r) * n1);  // w_r^{qm} = w_N^{qm N/r}
            y = y + cmul(x, twN[k]);
        }
        if (q) y = cmul(y, twN[q * n]);
        buf[permSub[n]] = y;
    }
    __syncthreads();
    fft_dit_lds<T, false>(buf, sub, 1, 0, 1, twSub, tid, NT);
    T mx = 0;
    for (int k = tid; k < n1; k += NT) {
        const cpx<T> v = buf[k];
        mx = fmax(mx, v.x * v.x + v.y * v.y);
    }
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
    if ((tid & 63) == 0) red[tid >> 6] = mx;
    __syncthreads();
    if (tid == 0) {
        for (int w = 1; w < NT / 64; ++w) mx = fmax(mx, red[w]);
        partial[blockIdx.x] = sqrt(mx);
    }
}

// owner window of sample n (-1: none), windows start at i*stride
__device__ __forceinline__ int eccm_owner(int n, int N, int fsz, int stride, int nf) {
    int i = min(n / stride, nf - 1);
    for (int k = 0; k < 3 && i >= 0; ++k, --i) {
        const int s = i * stride;
        int e = s + fsz;
        if (e > N) e = N - 1;
        if (n >= s && n < e) return i;
        if (n >= s + fsz) break;
    }
    return -1;
}

template <typename T, int NT>
__global__ __launch_bounds__(NT) void eccm_notch_kernel(const cpx<T>* __restrict__ in, int64_t ld_in,
                                                        cpx<T>* __restrict__ out, int64_t ld_out, int N, int fsz,
                                                        int stride, int nf, int wpc, int nchunk, double det,
                                                        double mit, const T* __restrict__ partial, int r,
                                                        uint8_t* __restrict__ flags, FftPlan pw,
                                                        const cpx<T>* __restrict__ tw,
                                                        const int32_t* __restrict__ perm) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int line = blockIdx.x / nchunk, chunk = blockIdx.x % nchunk;
    const int tid = threadIdx.x;
    const cpx<T>* row = in + (int64_t)line * ld_in;
    cpx<T>* orow = out + (int64_t)line * ld_out;
    double mx = 0;
    for (int q = 0; q < r; ++q) mx = fmax(mx, (double)partial[line * r + q]);
    const bool flagged = !(mx < det);
    if (flags && chunk == 0 && tid == 0) flags[line] = flagged ? 1 : 0;
    const int s0 = chunk * wpc * stride;
    const int s1 = min(s0 + wpc * stride, N);
    if (!flagged) {
        for (int n = s0 + tid; n < s1; n += NT) orow[n] = row[n];
        return;
    }
    // windows whose samples can own [s0, s1): the run [w0, w1)
    const int w0 = max(chunk * wpc - ((fsz + stride - 1) / stride - 1), 0);
    const int w1 = min((chunk + 1) * wpc, nf);
    const int nw = w1 - w0;
    cpx<T>* buf = (cpx<T>*)smem_raw;                       // nw * fsz
    uint8_t* keep = (uint8_t*)(buf + (size_t)nw * fsz);    // nw * fsz
    for (int idx = tid; idx < nw * fsz; idx += NT) {
        const int w = idx / fsz, j = idx - w * fsz;
        const int s = (w0 + w) * stride;
        int e = s + fsz;
        if (e > N) e = N - 1;
        cpx<T> v = mk((T)0, (T)0);
        if (s + j < e) v = row[s + j];
        buf[w * fsz + perm[j]] = v;
    }
    __syncthreads();
    fft_dit_lds<T, false, true>(buf, pw, nw, fsz, 1, tw, tid, NT);
    for (int idx = tid; idx < nw * fsz; idx += NT) {
        const int w = idx / fsz, j = idx - w * fsz;
        const cpx<T>* X = buf + w * fsz;
        cpx<T> g;
        if (j != 0 && j != fsz - 1) {
            const cpx<T> d = X[j + 1] - X[j - 1];
            g = mk(d.x / (T)2, d.y / (T)2);
        } else if (j == 0) {
            g = X[1] - X[0];
        } else {
            g = X[j] - X[j - 1];
        }
        keep[idx] = (T)sqrt((double)g.x * g.x + (double)g.y * g.y) < (T)mit ? 1 : 0;
    }
    __syncthreads();
    for (int idx = tid; idx < nw * fsz; idx += NT)
        if (!keep[idx]) buf[idx] = mk((T)0, (T)0);
    __syncthreads();
    fft_dif_lds<T, true, true>(buf, pw, nw, fsz, 1, tw, tid, NT);
    const T inv = (T)(1.0 / (double)fsz);
    for (int n = s0 + tid; n < s1; n += NT) {
        const int o = eccm_owner(n, N, fsz, stride, nf);
        if (o >= w0 && o < w1) {
            const int j = n - o * stride;
            orow[n] = scl(buf[(o - w0) * fsz + perm[j]], inv);
        } else {
            orow[n] = row[n];
        }
    }
}

template <typename T>
static int run_eccm(hipStream_t s, const s1g_eccm_desc& d, const DevFft& fN, const DevFft& fsub, int r,
                    const DevFft& fw, const void* in, int64_t ld_in, void* out, int64_t ld_out, uint8_t* flags,
                    void* partial) {
    constexpr int NT = 512;
    const int L = d.n_lines, N = d.n_samples;
    const size_t lds_det = sizeof(cpx<T>) * (size_t)(N / r);
    hipLaunchKernelGGL((eccm_detect_kernel<T, NT>), dim3((unsigned)(L * r)), dim3(NT), lds_det, s,
                       (const cpx<T>*)in, ld_in, N, r, fsub.plan, tw_of<T>(fN), tw_of<T>(fsub), fsub.perm,
                       (T*)partial);
    const int nf = (int)floor(((double)N / (double)d.fft_size) * ((double)d.fft_size / (double)d.stride));
    const int wpc = 16;
    const int nchunk = (N + wpc * d.stride - 1) / (wpc * d.stride);
    const int nw_max = wpc + (d.fft_size + d.stride - 1) / d.stride - 1;
    const size_t lds = (sizeof(cpx<T>) + 1) * (size_t)nw_max * d.fft_size;
    hipLaunchKernelGGL((eccm_notch_kernel<T, NT>), dim3((unsigned)(L * nchunk)), dim3(NT), lds, s,
                       (const cpx<T>*)in, ld_in, (cpx<T>*)out, ld_out, N, d.fft_size, d.stride, nf > 0 ? nf : 0, wpc,
                       nchunk, d.detection_threshold, d.mitigation_threshold, (const T*)partial, r, flags, fw.plan,
                       tw_of<T>(fw), fw.perm);
    return (int)hipGetLastError();
}

int launch_eccm(hipStream_t s, int prec, const s1g_eccm_desc& d, const DevFft& fN, const DevFft& fsub, int r,
                 const DevFft& fw, const void* in, int64_t ld_in, void* out, int64_t ld_out, uint8_t* flags,
                 void* partial) {
    if (prec == S1G_F32) return run_eccm<float>(s, d, fN, fsub, r, fw, in, ld_in, out, ld_out, flags, partial);
    return run_eccm<double>(s, d, fN, fsub, r, fw, in, ld_in, out, ld_out, flags, partial);
}

}  // namespace s1g
