// azimuth.hip -- the range-Doppler azimuth chain.
//
//   azimuth FFT   column panels of W range samples x all L lines in LDS;
//                 natural-order load, DIF passes, digit-reversed gather on
//                 store with the fftshift (forward, s1_decoder.cpp:499-505)
//                 or the 1/L scale (inverse, s1_decoder.cpp:682,
//                 signal_processing.cpp:581-605) folded in.
//   RCMC + MF     one Doppler row tile per workgroup: FP64 migration offset
//                 and phase per sample, 17-tap windowed-sinc interpolation
//                 from an LDS tile, amplitude weight and matched filter
//                 (formulas of s1_decoder.cpp:599-676, image_formation.cpp:
//                 422-476, signal_processing.cpp:228-257).
#include "s1g_internal.h"

#include <initializer_list>

namespace s1g {

// Blocks are dealt round-robin to the 8 XCDs; give each XCD a contiguous
// run of panels so neighbouring panels share that XCD's L2.
__device__ __forceinline__ int xcd_panel(int bid, int npanels) {
    const int per = (npanels + 7) / 8;
    return (bid & 7) * per + (bid >> 3);
}

// In-place DIF passes with a compile-time radix sequence (DIT order
// R0, R1, ...; DIF runs them last to first), so every pass is straight-line
// code with its own register allocation.
template <typename T, bool INV, int... Rs>
struct DifSeq;

template <typename T, bool INV>
struct DifSeq<T, INV> {
    static __device__ __forceinline__ void run(cpx<T>*, int, int, int, int, int, const cpx<T>*, int, int) {}
};

template <typename T, bool INV, int R, int... Rest>
struct DifSeq<T, INV, R, Rest...> {
    // Ns = product of the radices before R (those DIF applies after R)
    static __device__ __forceinline__ void run(cpx<T>* buf, int n, int Ns, int nvec, int vstride, int estride,
                                               const cpx<T>* tw, int tid, int nthr) {
        DifSeq<T, INV, Rest...>::run(buf, n, Ns * R, nvec, vstride, estride, tw, tid, nthr);
        dif_pass<T, R, INV>(buf, n, Ns, nvec, vstride, estride, tw, tid, nthr);
        __syncthreads();
    }
};

// One workgroup per panel of W adjacent range columns (W * sizeof(cpx<T>)
// = 64 contiguous bytes per line), for a length whose radices are known at
// compile time.  Every thread issues all its panel loads before the first
// use (at most MAXE per thread); the twiddles and the digit-reversal map
// are staged in LDS; natural-order input, DIF passes, and the store gathers
// the digit-reversed output with the fftshift (forward) or the 1/L scale
// (inverse) folded in.
template <typename T, bool INV, int NT, int W, int MAXE, int... Rs>
__global__ __launch_bounds__(NT) void az_fft_fixed_kernel(cpx<T>* __restrict__ data, int64_t ld, int N,
                                                          int npanels, const cpx<T>* __restrict__ tw_g,
                                                          const int32_t* __restrict__ perm_g, T scale) {
    constexpr int L = (Rs * ...);
    constexpr int total = L * W;
    static_assert(total <= NT * MAXE, "panel fits the per-thread registers");
    extern __shared__ unsigned char smem_raw[];
    cpx<T>* buf = (cpx<T>*)smem_raw;        // L * W
    cpx<T>* tw = buf + (size_t)L * W;       // L
    int32_t* perm = (int32_t*)(tw + L);     // L
    const int panel = xcd_panel(blockIdx.x, npanels);
    if (panel >= npanels) return;
    const int j0 = panel * W;
    const int w = min(W, N - j0);
    const int tid = threadIdx.x;
    cpx<T> r[MAXE];
#pragma unroll
    for (int i = 0; i < MAXE; ++i) {
        const int idx = tid + NT * i;
        const int e = idx / W, v = idx % W;
        r[i] = mk((T)0, (T)0);
        if (idx < total && v < w) r[i] = data[(int64_t)e * ld + j0 + v];
    }
    for (int k = tid; k < L; k += NT) {
        tw[k] = tw_g[k];
        perm[k] = perm_g[k];
    }
#pragma unroll
    for (int i = 0; i < MAXE; ++i) {
        const int idx = tid + NT * i;
        if (idx < total) buf[idx] = r[i];
    }
    __syncthreads();
    DifSeq<T, INV, Rs...>::run(buf, L, 1, W, 1, W, tw, tid, NT);
    // forward: out[i] = X[(i + L - L/2) mod L] (fftshift); inverse: x[k] / L
    constexpr int sh = INV ? 0 : L - L / 2;
#pragma unroll
    for (int i = 0; i < MAXE; ++i) {
        const int idx = tid + NT * i;
        const int k = idx / W, v = idx % W;
        if (idx < total && v < w) {
            int src = k + sh;
            if (src >= L) src -= L;
            const cpx<T> y = buf[perm[src] * W + v];
            data[(int64_t)k * ld + j0 + v] = INV ? scl(y, scale) : y;
        }
    }
}

template <typename T, bool INV, int W, int... Rs>
static void launch_az_fixed(hipStream_t s, const DevFft& fl, cpx<T>* data, int64_t ld, int N, T scale) {
    constexpr int NT = 512;
    constexpr int L = (Rs * ...);
    constexpr int MAXE = (L * W + NT - 1) / NT;
    const int npanels = (N + W - 1) / W;
    const int per = (npanels + 7) / 8;
    const size_t lds = (size_t)L * W * sizeof(cpx<T>) + (size_t)L * (sizeof(cpx<T>) + sizeof(int32_t));
    hipLaunchKernelGGL((az_fft_fixed_kernel<T, INV, NT, W, MAXE, Rs...>), dim3((unsigned)(8 * per)), dim3(NT), lds,
                       s, data, ld, N, npanels, tw_of<T>(fl), fl.perm, scale);
}

// Plans with a compiled radix sequence (the IW1/IW2/IW3 burst lengths).
template <typename T, bool INV, int W>
static bool az_fixed(hipStream_t s, const DevFft& fl, cpx<T>* data, int64_t ld, int N, T scale) {
    const FftPlan& p = fl.plan;
    auto is = [&](std::initializer_list<int> rs) {
        if ((int)rs.size() != p.npass) return false;
        int k = 0;
        for (int r : rs)
            if (p.radix[k++] != r) return false;
        return true;
    };
    if (is({37, 5, 8})) launch_az_fixed<T, INV, W, 37, 5, 8>(s, fl, data, ld, N, scale);
    else if (is({71, 5, 4})) launch_az_fixed<T, INV, W, 71, 5, 4>(s, fl, data, ld, N, scale);
    else if (is({19, 5, 8, 2})) launch_az_fixed<T, INV, W, 19, 5, 8, 2>(s, fl, data, ld, N, scale);
    else return false;
    return true;
}

template <typename T, bool INV, int NT>
__global__ __launch_bounds__(NT) void az_fft_generic_kernel(cpx<T>* __restrict__ data, int64_t ld, int L, int N, int W,
                                                    int npanels, FftPlan plan, const cpx<T>* __restrict__ tw,
                                                    const int32_t* __restrict__ perm, T scale) {
    extern __shared__ unsigned char smem_raw[];
    cpx<T>* buf = (cpx<T>*)smem_raw;
    const int panel = xcd_panel(blockIdx.x, npanels);
    if (panel >= npanels) return;
    const int j0 = panel * W;
    const int w = min(W, N - j0);
    const int tid = threadIdx.x;
    const int total = L * W;
    if (!INV) {
        for (int idx = tid; idx < total; idx += NT) {
            const int e = idx / W, v = idx - e * W;
            cpx<T> x = mk((T)0, (T)0);
            if (v < w) x = data[(int64_t)e * ld + j0 + v];
            buf[perm[e] * W + v] = x;
        }
        __syncthreads();
        fft_dit_lds<T, false>(buf, plan, W, 1, W, tw, tid, NT);
        const int sh = L - L / 2;  // fftshift: out[i] = X[(i + L - L/2) mod L]
        for (int idx = tid; idx < total; idx += NT) {
            const int i = idx / W, v = idx - i * W;
            int src = i + sh;
            if (src >= L) src -= L;
            if (v < w) data[(int64_t)i * ld + j0 + v] = buf[src * W + v];
        }
    } else {
        for (int idx = tid; idx < total; idx += NT) {
            const int e = idx / W, v = idx - e * W;
            cpx<T> x = mk((T)0, (T)0);
            if (v < w) x = data[(int64_t)e * ld + j0 + v];
            buf[idx] = x;
        }
        __syncthreads();
        fft_dif_lds<T, true>(buf, plan, W, 1, W, tw, tid, NT);
        for (int idx = tid; idx < total; idx += NT) {
            const int k = idx / W, v = idx - k * W;
            if (v < w) data[(int64_t)k * ld + j0 + v] = scl(buf[perm[k] * W + v], scale);
        }
    }
}

static int launch_az_generic(hipStream_t s, int prec, const DevFft& fl, void* data, int64_t ld, int L, int N,
                             int inverse) {
    constexpr int NT = 512;
    // panel width: 64-byte row segments (8 complex64 / 4 complex128) within the LDS budget
    int W = prec == S1G_F32 ? 8 : 4;
    const size_t esz = prec == S1G_F32 ? sizeof(cpx<float>) : sizeof(cpx<double>);
    while (W > 1 && (size_t)L * W * esz > 150 * 1024) W /= 2;
    if ((size_t)L * W * esz > 150 * 1024) return -3;
    const int npanels = (N + W - 1) / W;
    const int per = (npanels + 7) / 8;
    dim3 grid((unsigned)(8 * per));
    const size_t lds = (size_t)L * W * esz;
    if (prec == S1G_F32) {
        const float sc = (float)(1.0 / (double)L);
        if (inverse)
            hipLaunchKernelGGL((az_fft_generic_kernel<float, true, NT>), grid, dim3(NT), lds, s, (cpx<float>*)data, ld, L, N,
                               W, npanels, fl.plan, fl.tw32, fl.perm, sc);
        else
            hipLaunchKernelGGL((az_fft_generic_kernel<float, false, NT>), grid, dim3(NT), lds, s, (cpx<float>*)data, ld, L,
                               N, W, npanels, fl.plan, fl.tw32, fl.perm, 1.0f);
    } else {
        const double sc = 1.0 / (L);
        if (inverse)
            hipLaunchKernelGGL((az_fft_generic_kernel<double, true, NT>), grid, dim3(NT), lds, s, (cpx<double>*)data, ld, L,
                               N, W, npanels, fl.plan, fl.tw64, fl.perm, sc);
        else
            hipLaunchKernelGGL((az_fft_generic_kernel<double, false, NT>), grid, dim3(NT), lds, s, (cpx<double>*)data, ld, L,
                               N, W, npanels, fl.plan, fl.tw64, fl.perm, 1.0);
    }
    return (int)hipGetLastError();
}

int launch_azimuth_fft(hipStream_t s, int prec, const DevFft& fl, void* data, int64_t ld, int L, int N,
                       int inverse) {
    // panel width: 64-byte line segments (8 complex64 / 4 complex128)
    bool done;
    if (prec == S1G_F32) {
        const float sc = (float)(1.0 / (double)L);
        done = inverse ? az_fixed<float, true, 8>(s, fl, (cpx<float>*)data, ld, N, sc)
                       : az_fixed<float, false, 8>(s, fl, (cpx<float>*)data, ld, N, 1.0f);
    } else {
        const double sc = 1.0 / (double)L;
        done = inverse ? az_fixed<double, true, 4>(s, fl, (cpx<double>*)data, ld, N, sc)
                       : az_fixed<double, false, 4>(s, fl, (cpx<double>*)data, ld, N, 1.0);
    }
    if (done) return (int)hipGetLastError();
    return launch_az_generic(s, prec, fl, data, ld, L, N, inverse);
}

// ----------------------------------------------------------- column table
__global__ void az_columns_kernel(s1g_azimuth_desc d, AzColumn* __restrict__ cols) {
#pragma clang fp contract(off)
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= d.n_samples) return;
    const double PI = 3.14159265358979;
    const double R = d.slant_ranges[j];
    const double v = d.v_rel[j];
    AzColumn c;
    c.a = (d.wavelength * d.wavelength) / (4 * v * v);
    c.b = ((4.0 * PI) * R) / d.wavelength;
    c.c = (2.0 * PI) * d.t_corr[j];
    c.rdr = R / d.range_spacing;
    c.fdc = d.f_dc[j];
    c.amp = 1.0 / sqrt(pow(d.ref_range / R, 3.0));
    cols[j] = c;
}

int launch_az_columns(hipStream_t s, const s1g_azimuth_desc& d, AzColumn* cols) {
    dim3 b(256), g((d.n_samples + 255) / 256);
    hipLaunchKernelGGL(az_columns_kernel, g, b, 0, s, d, cols);
    return (int)hipGetLastError();
}

// ------------------------------------------------------------ RCMC + MF
constexpr int kTaps = 8;  // half width of the windowed sinc (17 taps)

// Interpolation weights at fractional offset fr in [0, 1): tap k (-8..8)
// gets sinc(fr - k) * 0.5 (1 + cos(PI (fr - k) / 8)); with
// sin(PI (fr - k)) = (-1)^k sin(PI fr) and the angle-addition form of the
// window only two sin/cos pairs are evaluated per sample.
// cos and sin of PI k / 8 for k = -8..8
constexpr double kCosK8[17] = {
    turn_cos(-8, 16), turn_cos(-7, 16), turn_cos(-6, 16), turn_cos(-5, 16), turn_cos(-4, 16), turn_cos(-3, 16),
    turn_cos(-2, 16), turn_cos(-1, 16), turn_cos(0, 16),  turn_cos(1, 16),  turn_cos(2, 16),  turn_cos(3, 16),
    turn_cos(4, 16),  turn_cos(5, 16),  turn_cos(6, 16),  turn_cos(7, 16),  turn_cos(8, 16)};
constexpr double kSinK8[17] = {
    turn_sin(-8, 16), turn_sin(-7, 16), turn_sin(-6, 16), turn_sin(-5, 16), turn_sin(-4, 16), turn_sin(-3, 16),
    turn_sin(-2, 16), turn_sin(-1, 16), turn_sin(0, 16),  turn_sin(1, 16),  turn_sin(2, 16),  turn_sin(3, 16),
    turn_sin(4, 16),  turn_sin(5, 16),  turn_sin(6, 16),  turn_sin(7, 16),  turn_sin(8, 16)};

template <typename T>
struct SincWeights;

template <>
struct SincWeights<float> {
    // sin(pi (fr - k)) = (-1)^k sinpi(fr); sinpi/cospi reduce their argument
    // exactly, so the weights keep full relative accuracy for fr near 0 or 1.
    static __device__ __forceinline__ void eval(float fr, float* w) {
        const float pi = 3.14159265358979f;
        float S, C, s8, c8;
        sincospif(fr, &S, &C);
        sincospif(fr * 0.125f, &s8, &c8);
#pragma unroll
        for (int k = -kTaps; k <= kTaps; ++k) {
            const float x = fr - (float)k;
            const float win = 0.5f * (1.0f + (c8 * (float)kCosK8[k + kTaps] + s8 * (float)kSinK8[k + kTaps]));
            const float sgn = (k & 1) ? -1.0f : 1.0f;
            const float sinc = (x == 0.0f) ? 1.0f : (sgn * S) / (pi * x);
            w[k + kTaps] = sinc * win;
        }
    }
};

template <>
struct SincWeights<double> {
    // The reference evaluates sin(PI (fr - k)) per tap; the two nearest taps
    // take exactly its arguments: PI fr (k = 0) and -PI (1 - fr) (k = 1).
    // Even taps reuse sin(PI fr), odd taps -sin(PI (1 - fr)).
    static __device__ __forceinline__ void eval(double fr, double* w) {
#pragma clang fp contract(off)
        const double PI = 3.14159265358979;
        const double Se = sin(PI * fr);
        const double So = sin(PI * (1.0 - fr));
        double s8, c8;
        sincos(PI * fr / kTaps, &s8, &c8);
#pragma unroll
        for (int k = -kTaps; k <= kTaps; ++k) {
            const double x = fr - (double)k;
            const double win = 0.5 * (1 + (c8 * kCosK8[k + kTaps] + s8 * kSinK8[k + kTaps]));
            const double num = (k & 1) ? -So : Se;
            const double sinc = (x == 0.0) ? 1.0 : num / (PI * x);
            w[k + kTaps] = sinc * win;
        }
    }
};

// Per-sample geometry.  FP64 (reference expression order, used by the FP64
// path): t = j + (R/D - R)/dR, phases ph1 = 4 PI R D / lambda and
// ph2 = 2 PI (f + fdc) tc, amplitude 1/sqrt((Rref/R)^3).
struct SampleGeo {
    double t;
    double c, s;  // e^{i (ph1 + ph2)}
    double amp;
};

__device__ __forceinline__ SampleGeo geo_f64(const s1g_azimuth_desc& d, double fa, int j) {
#pragma clang fp contract(off)
    const double PI = 3.14159265358979;
    const double R = d.slant_ranges[j], v = d.v_rel[j], fdc = d.f_dc[j], tc = d.t_corr[j];
    const double lam = d.wavelength;
    const double f = fa + fdc;
    const double D = sqrt(1 - ((lam * lam) * (f * f)) / (4 * v * v));
    SampleGeo g;
    const double shift = (R / D) - R;
    g.t = (double)j + shift / d.range_spacing;
    g.amp = sqrt(pow(d.ref_range / R, 3.0));
    const double ph1 = (((4.0 * PI) * R) * D) / lam;
    const double ph2 = ((2.0 * PI) * (fa + fdc)) * tc;
    double s1, c1, s2, c2;
    sincos(ph1, &s1, &c1);
    sincos(ph2, &s2, &c2);
    g.c = c1 * c2 - s1 * s2;
    g.s = c1 * s2 + s1 * c2;
    return g;
}

// FP32 chain: same quantities from the per-column FP64 table, the phase
// reduced mod 2 pi in FP64 before the FP32 sincos.
struct SampleGeo32 {
    double t;
    float c, s, amp;
};

__device__ __forceinline__ SampleGeo32 geo_f32(const AzColumn& col, double fa, int j) {
    const double f = fa + col.fdc;
    const double D = sqrt(1.0 - col.a * (f * f));
    SampleGeo32 g;
    g.t = (double)j + col.rdr * (1.0 / D - 1.0);
    const double ph = reduce_2pi(col.b * D + col.c * f);
    float sn, cs;
    sincosf((float)ph, &sn, &cs);
    g.c = cs;
    g.s = sn;
    g.amp = (float)col.amp;
    return g;
}

template <typename T, int NT, int EPT>
__global__ __launch_bounds__(NT) void rcmc_mf_kernel(s1g_azimuth_desc d, const AzColumn* __restrict__ cols,
                                                     const cpx<T>* __restrict__ in, int64_t ld_in,
                                                     cpx<T>* __restrict__ out, int64_t ld_out, int cap) {
    extern __shared__ unsigned char smem_raw[];
    cpx<T>* tile = (cpx<T>*)smem_raw;
    __shared__ int s_lo, s_hi;
    const int N = d.n_samples, L = d.n_lines;
    const int i = blockIdx.y;
    const int tid = threadIdx.x;
    const int j0 = blockIdx.x * NT * EPT;
    // Doppler frequency of row i: linspace(-prf/2, prf/2, L)[i]
    double fa;
    {
#pragma clang fp contract(off)
        const double start = -d.prf / 2, end = d.prf / 2;
        const double step = (end - start) / (L - 1);
        fa = start + i * step;
    }
    const cpx<T>* row = in + (int64_t)i * ld_in;
    double tt[EPT];
    T cw[EPT], sw[EPT], aw[EPT];
    int lo = 1 << 30, hi = -(1 << 30);
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int j = j0 + e * NT + tid;
        tt[e] = 0;
        cw[e] = 0;
        sw[e] = 0;
        aw[e] = 0;
        if (j < N) {
            if (sizeof(T) == 8) {
                SampleGeo g = geo_f64(d, fa, j);
                tt[e] = d.apply_rcmc ? g.t : (double)j;
                cw[e] = (T)g.c;
                sw[e] = (T)g.s;
                aw[e] = (T)g.amp;
            } else {
                SampleGeo32 g = geo_f32(cols[j], fa, j);
                tt[e] = d.apply_rcmc ? g.t : (double)j;
                cw[e] = (T)g.c;
                sw[e] = (T)g.s;
                aw[e] = (T)g.amp;
            }
            const int n0 = (int)floor(tt[e]);
            lo = min(lo, n0 - kTaps);
            hi = max(hi, n0 + kTaps);
        }
    }
    if (tid == 0) {
        s_lo = 1 << 30;
        s_hi = -(1 << 30);
    }
    __syncthreads();
    atomicMin(&s_lo, lo);
    atomicMax(&s_hi, hi);
    __syncthreads();
    const int tlo = max(s_lo, 0), thi = min(s_hi, N - 1);
    const bool use_tile = (thi - tlo + 1) <= cap;
    if (use_tile)
        for (int k = tlo + tid; k <= thi; k += NT) tile[k - tlo] = row[k];
    __syncthreads();
    cpx<T>* orow = out + (int64_t)i * ld_out;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int j = j0 + e * NT + tid;
        if (j >= N) continue;
        cpx<T> acc = mk((T)0, (T)0);
        if (d.apply_rcmc) {
            const double t = tt[e];
            const int n0 = (int)floor(t);
            T w[2 * kTaps + 1];
            SincWeights<T>::eval((T)(t - (double)n0), w);
#pragma unroll
            for (int k = -kTaps; k <= kTaps; ++k) {
                const int idx = n0 + k;
                if (idx < 0 || idx >= N) continue;
                const cpx<T> x = use_tile ? tile[idx - tlo] : row[idx];
                acc.x += x.x * w[k + kTaps];
                acc.y += x.y * w[k + kTaps];
            }
        } else {
            acc = row[j];
        }
        if (sizeof(T) == 8) {
            acc = mk(acc.x / aw[e], acc.y / aw[e]);
        } else {
            acc = scl(acc, aw[e]);
        }
        orow[j] = cmul(acc, mk(cw[e], sw[e]));
    }
}

int launch_rcmc_mf(hipStream_t s, int prec, const s1g_azimuth_desc& d, const AzColumn* cols, const void* in,
                   int64_t ld_in, void* out, int64_t ld_out) {
    constexpr int NT = 256, EPT = 8;
    const int per = NT * EPT;
    const int cap = per + 1024;
    dim3 grid((unsigned)((d.n_samples + per - 1) / per), (unsigned)d.n_lines);
    if (prec == S1G_F32)
        hipLaunchKernelGGL((rcmc_mf_kernel<float, NT, EPT>), grid, dim3(NT), sizeof(cpx<float>) * cap, s, d, cols,
                           (const cpx<float>*)in, ld_in, (cpx<float>*)out, ld_out, cap);
    else
        hipLaunchKernelGGL((rcmc_mf_kernel<double, NT, EPT>), grid, dim3(NT), sizeof(cpx<double>) * cap, s, d,
                           cols, (const cpx<double>*)in, ld_in, (cpx<double>*)out, ld_out, cap);
    return (int)hipGetLastError();
}

}  // namespace s1g
