// range.hip -- range compression against the replica chirp, overlap-save.
//
// Reference: S1_Decoder::_range_compress (s1_decoder.cpp:476-497) pads each
// line of N samples to P = N + 2 nc, multiplies its FFT by
// get_reference_function(get_replica_chirp()) (image_formation.cpp:4-25,
// packet.cpp:345-373), inverts, fftshifts and crops nc samples each side.
// Worked through, that is the linear correlation
//     out[j] = nc^2 P sum_{u<nc} x[j + u - ceil(nc/2)] e^{-i theta_u},
//     theta_u = 2 PI (TXPSF t_u + TXPRR/2 t_u^2),  t_u = u txpl / (nc - 1),
// (x = 0 outside [0, N)).  Here it is evaluated by overlap-save with an
// FFT length M (a power of two that fits LDS): block k computes the
// V = M - nc + 1 outputs j0 = k V .. j0 + V - 1 from x[j0 - c .. j0 - c + M).
// Per block: load (digit-reversed into LDS) -> DIT FFT -> multiply by the
// filter Hm -> DIF inverse FFT -> store the valid part.  Hm = conj(FFT_M(g))/M
// with g[u] = nc^2 P e^{+i theta_u} (u < nc) is built in FP64 once per chirp
// geometry and rounded to FP32 for the FP32 chain.
#include "s1g_internal.h"

namespace s1g {

// g[u] = scale * e^{i theta_u} for u < nc, 0 up to M; FP64 phase with the
// reference's expression order (packet.cpp:359-370) and PI = 3.14159265358979.
__global__ void chirp_filter_seed_kernel(cpx<double>* __restrict__ g, int m, int nc, double phi1, double phi2,
                                         double txpl, double scale) {
#pragma clang fp contract(off)
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= m) return;
    if (u >= nc) {
        g[u] = mk(0.0, 0.0);
        return;
    }
    const double PI = 3.14159265358979;
    const double step = (txpl - 0.0) / (nc - 1);  // linspace(0, txpl, nc)
    const double t = 0.0 + u * step;
    const double ph = (2.0 * PI) * ((phi1 * t) + phi2 * (t * t));
    double sn, cs;
    sincos(ph, &sn, &cs);
    g[u] = mk(scale * cs, scale * sn);
}

template <typename T>
__global__ void filter_finalize_kernel(const cpx<double>* __restrict__ G, cpx<T>* __restrict__ h, int m) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m) return;
    const double inv = 1.0 / (double)m;
    const cpx<double> v = G[f];
    h[f] = mk((T)(v.x * inv), (T)(-v.y * inv));
}

// scratch: 3 * M complex doubles.
int launch_range_filter(hipStream_t s, const DevFft& fm, double txpsf, double txprr, double txpl, double fs_mhz,
                        int n_samples, int nc, void* scratch, void* h_out, int prec) {
    const int m = fm.plan.n;
    const double p = (double)n_samples + 2.0 * nc;
    const double scale = (double)nc * (double)nc * p;
    dim3 b(256), gr((m + 255) / 256);
    cpx<double>* g1 = (cpx<double>*)scratch;
    cpx<double>* g2 = g1 + m;
    cpx<double>* g3 = g2 + m;
    hipLaunchKernelGGL(chirp_filter_seed_kernel, gr, b, 0, s, g1, m, nc, txpsf, txprr * 0.5, txpl, scale);
    launch_fft_global(s, S1G_F64, fm, g1, g2, g3, 1, m, 0);  // G = FFT_M(g) in g2
    if (prec == S1G_F32)
        hipLaunchKernelGGL(filter_finalize_kernel<float>, gr, b, 0, s, (const cpx<double>*)g2, (cpx<float>*)h_out,
                           m);
    else
        hipLaunchKernelGGL(filter_finalize_kernel<double>, gr, b, 0, s, (const cpx<double>*)g2,
                           (cpx<double>*)h_out, m);
    return (int)hipGetLastError();
}

// One workgroup per (line, block).  Blocks of a line are consecutive in the
// launch order so their overlapping input halos meet in L2.
template <typename T, int NT>
__global__ __launch_bounds__(NT) void range_os_kernel(const cpx<T>* __restrict__ in, int64_t ld_in,
                                                      cpx<T>* __restrict__ out, int64_t ld_out, int n_lines, int N,
                                                      int cshift, int V, int nblk, FftPlan plan,
                                                      const cpx<T>* __restrict__ tw,
                                                      const int32_t* __restrict__ perm,
                                                      const cpx<T>* __restrict__ h) {
    extern __shared__ unsigned char smem_raw[];
    cpx<T>* buf = (cpx<T>*)smem_raw;
    const int tid = threadIdx.x;
    const int M = plan.n;
    const int line = blockIdx.x / nblk;
    const int blk = blockIdx.x - line * nblk;
    if (line >= n_lines) return;
    const int j0 = blk * V;
    const cpx<T>* row = in + (int64_t)line * ld_in;
    for (int q = tid; q < M; q += NT) {
        const int src = j0 - cshift + q;
        cpx<T> v = mk((T)0, (T)0);
        if (src >= 0 && src < N) v = row[src];
        buf[perm[q]] = v;
    }
    __syncthreads();
    fft_dit_lds<T, false, true>(buf, plan, 1, 0, 1, tw, tid, NT);
    for (int f = tid; f < M; f += NT) buf[f] = cmul(buf[f], h[f]);
    __syncthreads();
    fft_dif_lds<T, true, true>(buf, plan, 1, 0, 1, tw, tid, NT);
    cpx<T>* orow = out + (int64_t)line * ld_out;
    const int cnt = min(V, N - j0);
    for (int k = tid; k < cnt; k += NT) orow[j0 + k] = buf[perm[k]];
}

int launch_range_os(hipStream_t s, int prec, const DevFft& fm, const void* in, int64_t ld_in, void* out,
                    int64_t ld_out, int n_lines, int n_samples, int nc, const void* h) {
    const int M = fm.plan.n;
    const int V = M - nc + 1;
    if (V <= 0) return -1;
    const int nblk = (n_samples + V - 1) / V;
    const int cshift = (nc + 1) / 2;
    dim3 grid((unsigned)(nblk * n_lines));
    constexpr int NT = 512;
    if (prec == S1G_F32) {
        size_t lds = sizeof(cpx<float>) * (size_t)M;
        hipLaunchKernelGGL((range_os_kernel<float, NT>), grid, dim3(NT), lds, s, (const cpx<float>*)in, ld_in,
                           (cpx<float>*)out, ld_out, n_lines, n_samples, cshift, V, nblk, fm.plan, fm.tw32, fm.perm,
                           (const cpx<float>*)h);
    } else {
        size_t lds = sizeof(cpx<double>) * (size_t)M;
        hipLaunchKernelGGL((range_os_kernel<double, NT>), grid, dim3(NT), lds, s, (const cpx<double>*)in, ld_in,
                           (cpx<double>*)out, ld_out, n_lines, n_samples, cshift, V, nblk, fm.plan, fm.tw64,
                           fm.perm, (const cpx<double>*)h);
    }
    return (int)hipGetLastError();
}

}  // namespace s1g
