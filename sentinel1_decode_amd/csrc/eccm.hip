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
// y_q[n] = sum_m x[n + m N/r] w_r^{q m} w_N^{q n}.  One workgroup per line
// forms and transforms the r sub-spectra in turn (the line comes from HBM
// once; the later sub-spectra re-read it from cache), each in LDS with the
// sub-length's twiddles in LDS too; the maximum does not depend on the
// order of the bins, so natural-order input and DIF passes need no
// digit-reversal.
// Notch: one workgroup per (line, run of windows); unflagged lines are copied.
// In a burst chain (in and out the same rows) the detection kernel also
// copies each flagged line to a staging buffer, and the notch reads the
// flagged lines from there and leaves the unflagged ones alone: the 80-90%
// of lines without RFI are then neither read nor written again.
#include "s1g_internal.h"
#include "s1g_lds_fft.h"

namespace s1g {

template <typename T, int NT, int MAXR>
__global__ __launch_bounds__(NT) void eccm_detect_kernel(const cpx<T>* __restrict__ in, int64_t ld, int N, int r,
                                                         FftPlan sub, const cpx<T>* __restrict__ twN,
                                                         const cpx<T>* __restrict__ twSub, T* __restrict__ peak,
                                                         double det, cpx<T>* __restrict__ stage, int64_t ld_stage) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int n1 = N / r;
    cpx<T>* buf = (cpx<T>*)smem_raw;   // n1
    cpx<T>* lo = buf + n1;             // kTwLo
    cpx<T>* hi = lo + kTwLo;           // tw_hi_count(n1)
    cpx<T>* wr = hi + tw_hi_count(n1); // r: w_r^k = w_N^{k N / r}
    __shared__ T red[NT / 64];
    const int line = blockIdx.x;
    const int tid = threadIdx.x;
    const cpx<T>* row = in + (int64_t)line * ld;
    load_tw2(twSub, n1, lo, hi, tid, NT);
    for (int k = tid; k < r; k += NT) wr[k] = twN[(int64_t)k * n1];
    __syncthreads();
    T mx = 0;
    for (int q = 0; q < r; ++q) {
        // four elements per thread at a time, their loads issued together
        for (int n0 = 0; n0 < n1; n0 += 4 * NT) {
            cpx<T> y[4], t[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int n = n0 + tid + NT * u;
                y[u] = mk((T)0, (T)0);
                t[u] = mk((T)1, (T)0);
                if (n < n1) {
                    y[u] = row[n];
                    if (q) t[u] = twN[q * n];
                }
            }
            int k = 0;  // (q m) mod r
            for (int m = 1; m < r; ++m) {
                k += q;
                if (k >= r) k -= r;
                cpx<T> x[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int n = n0 + tid + NT * u;
                    x[u] = n < n1 ? row[n + m * n1] : mk((T)0, (T)0);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) y[u] = y[u] + cmul(x[u], wr[k]);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int n = n0 + tid + NT * u;
                if (n < n1) buf[n] = q ? cmul(y[u], t[u]) : y[u];
            }
        }
        __syncthreads();
        fft_dif_lds_tw2<T, false, MAXR>(buf, sub, lo, hi, tid, NT);
        for (int k = tid; k < n1; k += NT) {
            const cpx<T> v = buf[k];
            mx = fmax(mx, v.x * v.x + v.y * v.y);
        }
        __syncthreads();  // every lane has read the sub-spectrum
    }
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
    if ((tid & 63) == 0) red[tid >> 6] = mx;
    __syncthreads();
    if (tid == 0) {
        for (int w = 1; w < NT / 64; ++w) mx = fmax(mx, red[w]);
        peak[line] = sqrt(mx);
        red[0] = mx;
    }
    if (stage) {
        // a flagged line (the notch kernel's test on the same peak value)
        // to the staging buffer, four loads per thread issued together
        __syncthreads();
        if (!((double)sqrt(red[0]) < det)) {
            cpx<T>* srow = stage + (int64_t)line * ld_stage;
            for (int n0 = 0; n0 < N; n0 += 4 * NT) {
                cpx<T> v[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int n = n0 + tid + NT * u;
                    if (n < N) v[u] = row[n];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int n = n0 + tid + NT * u;
                    if (n < N) srow[n] = v[u];
                }
            }
        }
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
                                                        double mit, const T* __restrict__ peak,
                                                        uint8_t* __restrict__ flags, FftPlan pw,
                                                        const cpx<T>* __restrict__ tw,
                                                        const int32_t* __restrict__ perm, int copy_unflagged) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int line = blockIdx.x / nchunk, chunk = blockIdx.x % nchunk;
    const int tid = threadIdx.x;
    const cpx<T>* row = in + (int64_t)line * ld_in;
    cpx<T>* orow = out + (int64_t)line * ld_out;
    const bool flagged = !((double)peak[line] < det);
    if (flags && chunk == 0 && tid == 0) flags[line] = flagged ? 1 : 0;
    const int s0 = chunk * wpc * stride;
    const int s1 = min(s0 + wpc * stride, N);
    if (!flagged) {
        if (copy_unflagged)
            for (int n = s0 + tid; n < s1; n += NT) orow[n] = row[n];
        return;
    }
    // windows whose samples can own [s0, s1): the run [w0, w1)
    const int w0 = max(chunk * wpc - ((fsz + stride - 1) / stride - 1), 0);
    const int w1 = min((chunk + 1) * wpc, nf);
    const int nw = w1 - w0;
    // window w at buf + w * vs: one pad element per window, so that lanes on
    // neighbouring windows (the passes' fastest index) hit different banks
    const int vs = fsz + 1;
    cpx<T>* buf = (cpx<T>*)smem_raw;                       // nw * vs
    uint8_t* keep = (uint8_t*)(buf + (size_t)nw * vs);     // nw * fsz
    // the window samples, loaded four per thread at a time before their LDS
    // stores (one memory wait per group)
    for (int i0 = 0; i0 < nw * fsz; i0 += 4 * NT) {
        cpx<T> v[4];
        int pos[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int idx = i0 + tid + NT * u;
            const int w = idx / fsz, j = idx - w * fsz;
            const int s = (w0 + w) * stride;
            int e = s + fsz;
            if (e > N) e = N - 1;
            v[u] = mk((T)0, (T)0);
            pos[u] = -1;
            if (idx < nw * fsz) {
                pos[u] = w * vs + perm[j];
                if (s + j < e) v[u] = row[s + j];
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (pos[u] >= 0) buf[pos[u]] = v[u];
    }
    __syncthreads();
    fft_dit_lds<T, false, true>(buf, pw, nw, vs, 1, tw, tid, NT);
    for (int idx = tid; idx < nw * fsz; idx += NT) {
        const int w = idx / fsz, j = idx - w * fsz;
        const cpx<T>* X = buf + w * vs;
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
        if (!keep[idx]) buf[idx + idx / fsz] = mk((T)0, (T)0);  // window idx / fsz, bin idx mod fsz
    __syncthreads();
    fft_dif_lds<T, true, true>(buf, pw, nw, vs, 1, tw, tid, NT);
    const T inv = (T)(1.0 / (double)fsz);
    for (int n = s0 + tid; n < s1; n += NT) {
        const int o = eccm_owner(n, N, fsz, stride, nf);
        if (o >= w0 && o < w1) {
            const int j = n - o * stride;
            orow[n] = scl(buf[(o - w0) * vs + perm[j]], inv);
        } else {
            orow[n] = row[n];
        }
    }
}

// LDS of the detection kernel for sub-length n1 = N / r.
static size_t detect_lds_bytes(size_t es, int n1, int r) { return es * (size_t)(n1 + kTwLo + tw_hi_count(n1) + r); }

template <typename T>
static int run_eccm(hipStream_t s, const s1g_eccm_desc& d, const DevFft& fN, const DevFft& fsub, int r,
                    const DevFft& fw, const void* in, int64_t ld_in, void* out, int64_t ld_out, uint8_t* flags,
                    void* peak, void* stage, int64_t ld_stage) {
    constexpr int NT = 512;
    const int L = d.n_lines, N = d.n_samples;
    const size_t lds_det = detect_lds_bytes(sizeof(cpx<T>), N / r, r);
#define S1G_DETECT(CLS)                                                                                        \
    hipLaunchKernelGGL((eccm_detect_kernel<T, NT, kRadixClassMax[CLS]>), dim3((unsigned)L), dim3(NT), lds_det, s, \
                       (const cpx<T>*)in, ld_in, N, r, fsub.plan, tw_of<T>(fN), tw_of<T>(fsub), (T*)peak,      \
                       d.detection_threshold, (cpx<T>*)stage, ld_stage)
    switch (radix_class(fsub.plan)) {
        case 0: S1G_DETECT(0); break;
        case 1: S1G_DETECT(1); break;
        case 2: S1G_DETECT(2); break;
        default: S1G_DETECT(3); break;
    }
#undef S1G_DETECT
    const int nf = (int)floor(((double)N / (double)d.fft_size) * ((double)d.fft_size / (double)d.stride));
    const int wpc = 16;
    const int nchunk = (N + wpc * d.stride - 1) / (wpc * d.stride);
    const int nw_max = wpc + (d.fft_size + d.stride - 1) / d.stride - 1;
    const size_t lds = sizeof(cpx<T>) * (size_t)nw_max * (d.fft_size + 1) + (size_t)nw_max * d.fft_size;
    // staged: the notch reads the flagged lines from the staging buffer
    const void* nin = stage ? stage : in;
    const int64_t nld = stage ? ld_stage : ld_in;
    hipLaunchKernelGGL((eccm_notch_kernel<T, NT>), dim3((unsigned)(L * nchunk)), dim3(NT), lds, s,
                       (const cpx<T>*)nin, nld, (cpx<T>*)out, ld_out, N, d.fft_size, d.stride, nf > 0 ? nf : 0, wpc,
                       nchunk, d.detection_threshold, d.mitigation_threshold, (const T*)peak, flags, fw.plan,
                       tw_of<T>(fw), fw.perm, stage ? 0 : 1);
    return (int)hipGetLastError();
}

int eccm_detect_split(int prec, int N, bool (*factor)(int, FftPlan&)) {
    // the smallest r with N / r factoring over the pass radices and the
    // workgroup's LDS within 52 KB (three workgroups per CU)
    static const int cand[] = {1, 2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 32, 40, 48, 64};
    const size_t es = prec == S1G_F32 ? sizeof(cpx<float>) : sizeof(cpx<double>);
    for (int k : cand) {
        if (N % k) continue;
        FftPlan p;
        if (detect_lds_bytes(es, N / k, k) <= 52 * 1024 && factor(N / k, p)) return k;
    }
    return 0;
}

int launch_eccm(hipStream_t s, int prec, const s1g_eccm_desc& d, const DevFft& fN, const DevFft& fsub, int r,
                 const DevFft& fw, const void* in, int64_t ld_in, void* out, int64_t ld_out, uint8_t* flags,
                 void* partial, void* stage, int64_t ld_stage) {
    if (prec == S1G_F32)
        return run_eccm<float>(s, d, fN, fsub, r, fw, in, ld_in, out, ld_out, flags, partial, stage, ld_stage);
    return run_eccm<double>(s, d, fN, fsub, r, fw, in, ld_in, out, ld_out, flags, partial, stage, ld_stage);
}

}  // namespace s1g
