// autofocus.hip -- image entropy of the azimuth-compressed burst and its
// hand-written adjoint with respect to a per-Doppler-bin phase vector.
// (No counterpart in the reference; it builds on the azimuth IFFT of
// s1_decoder.cpp:682.)
//
// Forward, with rd the range-Doppler data after RCMC + matched filter
// (rows = Doppler bins) and phi[L]:
//   img = IFFT_az(rd * e^{i phi}) / L,  I = |img|^2,  S = sum I,  T = sum I log I,
//   E = -sum (I/S) log(I/S) = log S - T / S.
// Adjoint (dE/dI = -log(I)/S + T/S^2 =: w; the IFFT/L is adjointed by FFT/L):
//   g = FFT_az(w * img) / L,  dE/dphi_i = 2 sum_j Im(g_ij conj(rd_ij e^{i phi_i})).
// The spectrum rd is the checkpoint: the image is never stored between the
// two passes; the gradient pass recomputes it from rd (one read of rd per
// pass, nothing written but per-workgroup partial sums).  Sums in FP64.
#include "s1g_internal.h"

namespace s1g {

__device__ __forceinline__ int af_panel(int bid, int npanels) {
    const int per = (npanels + 7) / 8;
    return (bid & 7) * per + (bid >> 3);
}

template <typename T>
__global__ void af_phase_table_kernel(const double* __restrict__ phi, cpx<T>* __restrict__ tab, int L) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L) return;
    double s, c;
    sincos(phi[i], &s, &c);
    tab[i] = mk((T)c, (T)s);
}

// img = IFFT(rd * e^{i phi}) / L per column panel; per-panel partial sums of
// I and I log I into part[2 * panel].
template <typename T, int NT>
__global__ __launch_bounds__(NT) void af_forward_kernel(const cpx<T>* __restrict__ rd, int64_t ld, cpx<T>* __restrict__ img,
                                                         int L, int N, int W, int npanels, FftPlan plan,
                                                         const cpx<T>* __restrict__ tw, const int32_t* __restrict__ perm,
                                                         const cpx<T>* __restrict__ ph, double* __restrict__ part) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<T>* buf = (cpx<T>*)smem_raw;
    __shared__ double red[2][NT / 64];
    const int panel = af_panel(blockIdx.x, npanels);
    if (panel >= npanels) return;
    const int j0 = panel * W, w = min(W, N - j0), tid = threadIdx.x, total = L * W;
    for (int idx = tid; idx < total; idx += NT) {
        const int e = idx / W, v = idx - e * W;
        cpx<T> x = mk((T)0, (T)0);
        if (v < w) x = cmul(rd[(int64_t)e * ld + j0 + v], ph[e]);
        buf[idx] = x;
    }
    __syncthreads();
    fft_dif_lds<T, true>(buf, plan, W, 1, W, tw, tid, NT);
    const T sc = (T)(1.0 / (double)L);
    double s = 0, t = 0;
    for (int idx = tid; idx < total; idx += NT) {
        const int k = idx / W, v = idx - k * W;
        if (v >= w) continue;
        const cpx<T> y = scl(buf[perm[k] * W + v], sc);
        img[(int64_t)k * N + j0 + v] = y;
        const T I = y.x * y.x + y.y * y.y;
        s += (double)I;
        if (I > (T)0) t += (double)(I * log(I));
    }
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_xor(s, o);
        t += __shfl_xor(t, o);
    }
    if ((tid & 63) == 0) {
        red[0][tid >> 6] = s;
        red[1][tid >> 6] = t;
    }
    __syncthreads();
    if (tid == 0) {
        for (int q = 1; q < NT / 64; ++q) {
            s += red[0][q];
            t += red[1][q];
        }
        part[2 * panel] = s;
        part[2 * panel + 1] = t;
    }
}

// scal[0] = S, scal[1] = T, scal[2] = E
__global__ void af_reduce_scalars_kernel(const double* __restrict__ part, int npanels, double* __restrict__ scal) {
    __shared__ double rs[256], rt[256];
    double s = 0, t = 0;
    for (int p = threadIdx.x; p < npanels; p += blockDim.x) {
        s += part[2 * p];
        t += part[2 * p + 1];
    }
    rs[threadIdx.x] = s;
    rt[threadIdx.x] = t;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            rs[threadIdx.x] += rs[threadIdx.x + o];
            rt[threadIdx.x] += rt[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double S = rs[0], Tt = rt[0];
        scal[0] = S;
        scal[1] = Tt;
        scal[2] = log(S) - Tt / S;
    }
}

// g = FFT(w * img) / L per panel; gpart[panel * L + i] = 2 sum_{j in panel} Im(g_ij conj(rd_ij e^{i phi_i}))
template <typename T, int NT>
__global__ __launch_bounds__(NT) void af_adjoint_kernel(const cpx<T>* __restrict__ rd, int64_t ld,
                                                         const cpx<T>* __restrict__ img, int L, int N, int W,
                                                         int npanels, FftPlan plan, const cpx<T>* __restrict__ tw,
                                                         const int32_t* __restrict__ perm,
                                                         const cpx<T>* __restrict__ ph, const double* __restrict__ scal,
                                                         double* __restrict__ gpart) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<T>* buf = (cpx<T>*)smem_raw;
    const int panel = af_panel(blockIdx.x, npanels);
    if (panel >= npanels) return;
    const int j0 = panel * W, w = min(W, N - j0), tid = threadIdx.x, total = L * W;
    const double S = scal[0], Tt = scal[1];
    const T a = (T)(-1.0 / S), b = (T)(Tt / (S * S));
    for (int idx = tid; idx < total; idx += NT) {
        const int e = idx / W, v = idx - e * W;
        cpx<T> y = mk((T)0, (T)0);
        if (v < w) {
            y = img[(int64_t)e * N + j0 + v];
            const T I = y.x * y.x + y.y * y.y;
            const T wt = I > (T)0 ? a * log(I) + b : (T)0;
            y = scl(y, wt);
        }
        buf[perm[e] * W + v] = y;
    }
    __syncthreads();
    fft_dit_lds<T, false>(buf, plan, W, 1, W, tw, tid, NT);
    const T sc = (T)(1.0 / (double)L);
    // one row per thread (rows strided by NT): sum the panel's columns
    for (int i = tid; i < L; i += NT) {
        double acc = 0;
        const cpx<T> p = ph[i];
        for (int v = 0; v < w; ++v) {
            const cpx<T> g = scl(buf[i * W + v], sc);
            const cpx<T> z = cmul(rd[(int64_t)i * ld + j0 + v], p);
            acc += 2.0 * ((double)g.y * z.x - (double)g.x * z.y);
        }
        gpart[(int64_t)panel * L + i] = acc;
    }
}

template <typename T>
static int run_af(hipStream_t s, int L, int N, const void* rd, int64_t ld, const double* phase, const DevFft& fl,
                  void* img, void* img_tmp, double* work, int want_grad, double* grad, double* scal) {
    constexpr int prec = sizeof(T) == 4 ? S1G_F32 : S1G_F64;
    cpx<T>* ph = (cpx<T>*)work;
    double* part = work + (size_t)L * 2;
    hipLaunchKernelGGL((af_phase_table_kernel<T>), dim3((L + 255) / 256), dim3(256), 0, s, phase, ph, L);
    if (af_fast(L, N, ld, prec)) {
        // work: phase table | 2 G entropy partials | G L gradient partials
        const int G = af_reg_workgroups(N, prec);
        double* gpart = part + 2 * (size_t)G;
        int rc = launch_af_reg(s, prec, 0, L, N, rd, ld, fl, ph, scal, img, part, gpart);
        hipLaunchKernelGGL(af_reduce_scalars_kernel, dim3(1), dim3(256), 0, s, part, G, scal);
        if (!rc && want_grad) {
            rc = launch_af_reg(s, prec, 2, L, N, rd, ld, fl, ph, scal, nullptr, part, gpart);
            if (!rc) rc = launch_af_sum_partials(s, gpart, G, L, grad);
        }
        return rc ? rc : (int)hipGetLastError();
    }
    // other lengths: one workgroup per panel through the plan's LDS passes,
    // the image kept in img_tmp between the two kernels
    constexpr int NT = 512;
    int W = sizeof(T) == 4 ? 8 : 4;
    while (W > 1 && (size_t)L * W * sizeof(cpx<T>) > 150 * 1024) W /= 2;
    if ((size_t)L * W * sizeof(cpx<T>) > 150 * 1024) return -3;
    const int npanels = (N + W - 1) / W;
    const int per = (npanels + 7) / 8;
    // work: phase table | 2 npanels partials | npanels L gradient partials
    double* gpart = part + 2 * (size_t)npanels;
    const size_t lds = (size_t)L * W * sizeof(cpx<T>);
    hipLaunchKernelGGL((af_forward_kernel<T, NT>), dim3(8 * per), dim3(NT), lds, s, (const cpx<T>*)rd, ld,
                       (cpx<T>*)img_tmp, L, N, W, npanels, fl.plan, tw_of<T>(fl), fl.perm, ph, part);
    hipLaunchKernelGGL(af_reduce_scalars_kernel, dim3(1), dim3(256), 0, s, part, npanels, scal);
    if (want_grad) {
        hipLaunchKernelGGL((af_adjoint_kernel<T, NT>), dim3(8 * per), dim3(NT), lds, s, (const cpx<T>*)rd, ld,
                           (const cpx<T>*)img_tmp, L, N, W, npanels, fl.plan, tw_of<T>(fl), fl.perm, ph, scal,
                           gpart);
        const int rc = launch_af_sum_partials(s, gpart, npanels, L, grad);
        if (rc) return rc;
    }
    return (int)hipGetLastError();
}

size_t af_work_doubles(int L, int N, int64_t ld, int prec) {
    size_t nparts;
    if (af_fast(L, N, ld, prec)) {
        nparts = (size_t)af_reg_workgroups(N, prec);
    } else {
        int W = prec == S1G_F32 ? 8 : 4;
        const size_t esz = prec == S1G_F32 ? sizeof(cpx<float>) : sizeof(cpx<double>);
        while (W > 1 && (size_t)L * W * esz > 150 * 1024) W /= 2;
        nparts = (size_t)(N + W - 1) / W;
    }
    return (size_t)L * 2 + 2 * nparts + nparts * (size_t)L + 8;
}

int launch_af(hipStream_t s, int prec, int L, int N, const void* rd, int64_t ld, const double* phase, const DevFft& fl,
              void* img, void* img_tmp, double* work, int want_grad, double* grad, double* scal) {
    if (prec == S1G_F32) return run_af<float>(s, L, N, rd, ld, phase, fl, img, img_tmp, work, want_grad, grad, scal);
    return run_af<double>(s, L, N, rd, ld, phase, fl, img, img_tmp, work, want_grad, grad, scal);
}

}  // namespace s1g
