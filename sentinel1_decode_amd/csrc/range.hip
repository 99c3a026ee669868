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

#include <algorithm>

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

// Register-pass kernels: block length M = R0 R1 R2 handled by NT threads.
struct OsShape {
    int m, r0, r1, r2, nt;
};

// The register-pass configuration for (precision, M), or m = 0 when M is
// handled by the generic LDS kernel.
static OsShape os_shape(int prec, int m) {
    if (prec == S1G_F32 && m == 16384) return {16384, 32, 32, 16, 512};
    if (prec == S1G_F32 && m == 8192) return {8192, 32, 16, 16, 256};
    if (prec == S1G_F64 && m == 8192) return {8192, 32, 16, 16, 256};
    return {0, 0, 0, 0, 0};
}

// h[dst(f)] = conj(G[f]) / m.  For the register-pass kernels the filter is
// stored in the order their last forward pass leaves the spectrum: bin
// f = k0 + R0 k1 + R0 R1 k2 sits at k2 (m / R2) + k0 R1 + k1.
template <typename T>
__global__ void filter_finalize_kernel(const cpx<double>* __restrict__ G, cpx<T>* __restrict__ h, int m, int r0,
                                       int r1, int r2) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m) return;
    const double inv = 1.0 / (double)m;
    const cpx<double> v = G[f];
    int dst = f;
    if (r0) {
        const int k0 = f % r0, k1 = (f / r0) % r1, k2 = f / (r0 * r1);
        dst = k2 * (m / r2) + k0 * r1 + k1;
    }
    h[dst] = mk((T)(v.x * inv), (T)(-v.y * inv));
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
    const OsShape sh = os_shape(prec, m);
    if (prec == S1G_F32)
        hipLaunchKernelGGL(filter_finalize_kernel<float>, gr, b, 0, s, (const cpx<double>*)g2, (cpx<float>*)h_out,
                           m, sh.r0, sh.r1, sh.r2);
    else
        hipLaunchKernelGGL(filter_finalize_kernel<double>, gr, b, 0, s, (const cpx<double>*)g2,
                           (cpx<double>*)h_out, m, sh.r0, sh.r1, sh.r2);
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
    extern __shared__ __align__(16) unsigned char smem_raw[];
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


// LDS index with one pad element per 16, so that the last pass's
// contiguous 16/32-element reads by neighbouring lanes hit distinct banks.
__device__ __forceinline__ int os_pad(int a) { return a + (a >> 4); }

// Overlap-save block with three register-resident radix passes.
// Forward (in-place DIF on the index n = n' + S0 n0, S0 = M / R0, S1 = S0 / R1):
//   pass 0: DFT_R0 over n0 for each n' (loaded straight from global memory),
//           times w_M^{n' k0};
//   pass 1: per k0, DFT_R1 over the stride-S1 elements, times w_S0^{r k1};
//   pass 2: per (k0, k1), DFT_R2 over R2 contiguous elements; bin
//           f = k0 + R0 k1 + R0 R1 k2 is in registers, multiplied by the
//           filter (stored in that order) and transformed back by the
//           conjugate pass at once.
// The inverse passes mirror the forward ones in reverse order, so the last
// one holds outputs n' + S0 n0 and stores them straight to global memory.
// Twiddles: w_M^e = lo[e mod 128] * hi[e / 128] and w_S0^{r k1} from a
// [k1][r] table, all in LDS.
//
// One launch may serve several bursts (RangeBatch: per-instance lines,
// filter and buffers; workgroups numbered consecutively across instances).
template <typename T, int M, int R0, int R1, int R2, int NT>
__global__ __launch_bounds__(NT, (sizeof(T) == 4 ? 2 : 1)) void range_os3_kernel(RangeBatch bt, const cpx<T>* __restrict__ tw) {
    constexpr int S0 = M / R0, S1 = S0 / R1;
    static_assert(S1 == R2 && R0 * R1 * R2 == M, "three passes cover M");
    constexpr int D0 = S0, D1 = M / R1, D2 = M / R2;
    static_assert(D0 == NT && D1 % NT == 0 && D2 % NT == 0, "whole butterflies per thread");
    constexpr int P1 = D1 / NT, P2 = D2 / NT;
    constexpr int E = M / NT;
    static_assert(E == R0 && E == P1 * R1 && E == P2 * R2, "E elements per thread in every pass");
    constexpr int NHI = M / 128;
    constexpr int LDATA = M + M / 16;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<T>* buf = (cpx<T>*)smem_raw;
    cpx<T>* tlo = buf + LDATA;
    cpx<T>* thi = tlo + 128;
    cpx<T>* t1 = thi + NHI;  // [k1][r], R1 * S1 entries
    const int tid = threadIdx.x;
    // the workgroup's instance (uniform: a scan of the kernel arguments)
    int k = 0;
    while (k + 1 < bt.n && (int)blockIdx.x >= bt.inst[k + 1].wg0) ++k;
    const RangeInst& I = bt.inst[k];
    const int N = I.N, V = I.V, nblk = I.nblk, cshift = I.cshift;
    const int local = (int)blockIdx.x - I.wg0;
    const int line = local / nblk;
    const int blk = local - line * nblk;
    if (line >= I.n_lines) return;
    const int j0 = blk * V;
    const cpx<T>* row = (const cpx<T>*)I.in + (int64_t)line * I.ld_in;
    const cpx<T>* __restrict__ hs = (const cpx<T>*)I.h;

    cpx<T> a[E];
    // pass 0 input: n = tid + S0 n0, in [j0 - cshift, j0 - cshift + M)
    {
        const int src0 = j0 - cshift + tid;
#pragma unroll
        for (int n0 = 0; n0 < R0; ++n0) {
            const int src = src0 + S0 * n0;
            a[n0] = (src >= 0 && src < N) ? row[src] : mk((T)0, (T)0);
        }
    }
    // twiddle tables: every load issued (beside the data loads) before the
    // first LDS store, so the workgroup waits for memory once
    {
        static_assert(128 <= NT && NHI <= NT && R1 * S1 <= NT, "one table entry per thread");
        const cpx<T> vlo = tid < 128 ? tw[tid] : mk((T)0, (T)0);
        const cpx<T> vhi = tid < NHI ? tw[128 * tid] : mk((T)0, (T)0);
        const int k1 = tid / S1, r = tid - k1 * S1;
        const cpx<T> v1 = tid < R1 * S1 ? tw[((r * k1) % S0) * R0] : mk((T)0, (T)0);
        if (tid < 128) tlo[tid] = vlo;
        if (tid < NHI) thi[tid] = vhi;
        if (tid < R1 * S1) t1[tid] = v1;
    }
    __syncthreads();

    // ---- forward pass 0
    Dft<T, R0, false>::run(a);
#pragma unroll
    for (int k0 = 1; k0 < R0; ++k0) {
        const int e = (tid * k0) & (M - 1);
        a[k0] = cmul(a[k0], cmul(tlo[e & 127], thi[e >> 7]));
    }
#pragma unroll
    for (int k0 = 0; k0 < R0; ++k0) buf[os_pad(tid + S0 * k0)] = a[k0];
    __syncthreads();

    // ---- forward pass 1
#pragma unroll
    for (int p = 0; p < P1; ++p) {
        const int d = tid + NT * p;
        const int k0 = d / S1, r = d - k0 * S1;
        const int base = k0 * S0 + r;
        cpx<T>* v = a + p * R1;
#pragma unroll
        for (int n1 = 0; n1 < R1; ++n1) v[n1] = buf[os_pad(base + S1 * n1)];
        Dft<T, R1, false>::run(v);
#pragma unroll
        for (int k1 = 1; k1 < R1; ++k1) v[k1] = cmul(v[k1], t1[k1 * S1 + r]);
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) buf[os_pad(base + S1 * k1)] = v[k1];
    }
    __syncthreads();

    // ---- pass 2, filter, inverse pass 2
#pragma unroll
    for (int p = 0; p < P2; ++p) {
        const int d = tid + NT * p;
        const int base = d * R2;
        cpx<T>* v = a + p * R2;
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) v[n2] = buf[os_pad(base + n2)];
        Dft<T, R2, false>::run(v);
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) v[k2] = cmul(v[k2], hs[k2 * D2 + d]);
        Dft<T, R2, true>::run(v);
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) buf[os_pad(base + n2)] = v[n2];
    }
    __syncthreads();

    // ---- inverse pass 1
#pragma unroll
    for (int p = 0; p < P1; ++p) {
        const int d = tid + NT * p;
        const int k0 = d / S1, r = d - k0 * S1;
        const int base = k0 * S0 + r;
        cpx<T>* v = a + p * R1;
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) v[k1] = buf[os_pad(base + S1 * k1)];
#pragma unroll
        for (int k1 = 1; k1 < R1; ++k1) v[k1] = cmulc(v[k1], t1[k1 * S1 + r]);
        Dft<T, R1, true>::run(v);
#pragma unroll
        for (int n1 = 0; n1 < R1; ++n1) buf[os_pad(base + S1 * n1)] = v[n1];
    }
    __syncthreads();

    // ---- inverse pass 0 and store
#pragma unroll
    for (int k0 = 0; k0 < R0; ++k0) a[k0] = buf[os_pad(tid + S0 * k0)];
#pragma unroll
    for (int k0 = 1; k0 < R0; ++k0) {
        const int e = (tid * k0) & (M - 1);
        a[k0] = cmulc(a[k0], cmul(tlo[e & 127], thi[e >> 7]));
    }
    Dft<T, R0, true>::run(a);
    cpx<T>* orow = (cpx<T>*)I.out + (int64_t)line * I.ld_out;
    const int cnt = min(V, N - j0);
#pragma unroll
    for (int n0 = 0; n0 < R0; ++n0) {
        const int k = tid + S0 * n0;
        if (k < cnt) orow[j0 + k] = a[n0];
    }
}

template <typename T, int M, int R0, int R1, int R2, int NT>
static void launch_os3(hipStream_t s, const RangeBatch& bt, int nwg, const DevFft& fm) {
    const size_t lds = sizeof(cpx<T>) * (size_t)(M + M / 16 + 128 + M / 128 + R1 * (M / R0 / R1));
    hipLaunchKernelGGL((range_os3_kernel<T, M, R0, R1, R2, NT>), dim3((unsigned)nwg), dim3(NT), lds, s, bt,
                       tw_of<T>(fm));
}

// One instance of a batch (workgroup offset set by the launcher).
RangeInst range_inst(const void* in, int64_t ld_in, void* out, int64_t ld_out, int n_lines, int n_samples,
                            int M, int nc, const void* h) {
    RangeInst r;
    r.in = in;
    r.out = out;
    r.h = h;
    r.ld_in = ld_in;
    r.ld_out = ld_out;
    r.n_lines = n_lines;
    r.N = n_samples;
    r.V = M - nc + 1;
    r.nblk = r.V > 0 ? (n_samples + r.V - 1) / r.V : 0;
    r.cshift = (nc + 1) / 2;
    r.wg0 = 0;
    return r;
}

bool range_batchable(int prec, int M) { return os_shape(prec, M).m != 0; }

// Several instances whose block length M has a register-pass kernel, in
// launches of at most kMaxBatch instances.
int launch_range_os_batch(hipStream_t s, int prec, const DevFft& fm, const RangeInst* inst, int n) {
    const int M = fm.plan.n;
    if (!range_batchable(prec, M)) return -1;
    for (int b0 = 0; b0 < n; b0 += kMaxBatch) {
        RangeBatch bt;
        bt.n = std::min(kMaxBatch, n - b0);
        int nwg = 0;
        for (int k = 0; k < bt.n; ++k) {
            bt.inst[k] = inst[b0 + k];
            if (bt.inst[k].V <= 0) return -1;
            bt.inst[k].wg0 = nwg;
            nwg += bt.inst[k].nblk * bt.inst[k].n_lines;
        }
        if (nwg == 0) continue;
        if (prec == S1G_F32 && M == 16384) launch_os3<float, 16384, 32, 32, 16, 512>(s, bt, nwg, fm);
        else if (prec == S1G_F32) launch_os3<float, 8192, 32, 16, 16, 256>(s, bt, nwg, fm);
        else launch_os3<double, 8192, 32, 16, 16, 256>(s, bt, nwg, fm);
    }
    return (int)hipGetLastError();
}

int launch_range_os(hipStream_t s, int prec, const DevFft& fm, const void* in, int64_t ld_in, void* out,
                    int64_t ld_out, int n_lines, int n_samples, int nc, const void* h) {
    const int M = fm.plan.n;
    const int V = M - nc + 1;
    if (V <= 0) return -1;
    const int nblk = (n_samples + V - 1) / V;
    const int cshift = (nc + 1) / 2;
    dim3 grid((unsigned)(nblk * n_lines));
    if (range_batchable(prec, M)) {
        const RangeInst r = range_inst(in, ld_in, out, ld_out, n_lines, n_samples, M, nc, h);
        return launch_range_os_batch(s, prec, fm, &r, 1);
    }
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
