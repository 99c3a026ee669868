// rcmc.hip -- RCMC and the azimuth matched filter of the range-Doppler
// chain: one Doppler row tile per workgroup, FP64 migration offset and phase
// per sample, 17-tap windowed-sinc interpolation from an LDS tile, amplitude
// weight and matched filter (formulas of s1_decoder.cpp:599-676,
// image_formation.cpp:422-476, signal_processing.cpp:228-257), and the
// per-column table those kernels read.
#include "s1g_internal.h"

#include <algorithm>

namespace s1g {

// ----------------------------------------------------------- column table
__global__ void az_columns_kernel(s1g_azimuth_desc d, AzColumn* __restrict__ cols, float* __restrict__ amp32) {
#pragma clang fp contract(off)
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= d.n_samples) return;
    const double PI = 3.14159265358979;
    const double inv2pi = 0.15915494309189535;  // 1 / (2 pi), true pi: phases in turns
    const double R = d.slant_ranges[j];
    const double v = d.v_rel[j];
    AzColumn c;
    c.a = (d.wavelength * d.wavelength) / (4 * v * v);
    c.b = ((4.0 * PI) * R) / d.wavelength;
    c.c = (2.0 * PI) * d.t_corr[j];
    c.rdr = R / d.range_spacing;
    c.fdc = d.f_dc[j];
    c.sq = sqrt(pow(d.ref_range / R, 3.0));
    c.amp = 1.0 / c.sq;
    c.bt = c.b * inv2pi;
    c.ct = c.c * inv2pi;
    c.btf = c.bt - rint(c.bt);
    cols[j] = c;
    amp32[j] = (float)c.amp;
}

// Tolerances of the linear-geometry test: the migration R_j / dR to 1e-4
// samples (a shift error below 1e-4 u / 2), the phase coefficient bt to
// 5e-8 turns and t_corr to 5e-8 turns over the Doppler band; the FP32
// chain's phases are good to about 1e-7 rad besides.
constexpr double kLinTolRdr = 1e-4;
constexpr double kLinTolTurns = 5e-8;

// One workgroup: whether the columns share v_rel and f_dc and have R_j / dR,
// bt and t_corr linear in j (fit through the first and last column).
__global__ __launch_bounds__(1024) void az_uniform_kernel(const AzColumn* __restrict__ cols, int N, double prf,
                                                          AzUniform* __restrict__ out) {
    __shared__ int s_bad;
    const int tid = threadIdx.x;
    const AzColumn c0 = cols[0], c1 = cols[N - 1];
    const double den = N > 1 ? (double)(N - 1) : 1.0;
    const double drdr = (c1.rdr - c0.rdr) / den, dbt = (c1.bt - c0.bt) / den, dct = (c1.ct - c0.ct) / den;
    const double tol_ct = kLinTolTurns / (0.5 * fabs(prf) + fabs(c0.fdc) + 1.0);
    if (tid == 0) s_bad = 0;
    __syncthreads();
    int bad = 0;
    for (int j = tid; j < N; j += blockDim.x) {
        const AzColumn c = cols[j];
        bad |= c.a != c0.a || c.fdc != c0.fdc || !(fabs(c.rdr - (c0.rdr + j * drdr)) <= kLinTolRdr) ||
               !(fabs(c.bt - (c0.bt + j * dbt)) <= kLinTolTurns) || !(fabs(c.ct - (c0.ct + j * dct)) <= tol_ct);
    }
    if (bad) atomicOr(&s_bad, 1);
    __syncthreads();
    if (tid == 0) {
        AzUniform u;
        u.ok = s_bad ? 0 : 1;
        u.a = c0.a;
        u.fdc = c0.fdc;
        u.rdr0 = c0.rdr;
        u.drdr = drdr;
        u.bt0 = c0.bt;
        u.dbt = dbt;
        u.ct0 = c0.ct;
        u.dct = dct;
        *out = u;
    }
}

int launch_az_columns(hipStream_t s, const s1g_azimuth_desc& d, AzColumn* cols) {
    const int n = d.n_samples;
    dim3 b(256), g((n + 255) / 256);
    hipLaunchKernelGGL(az_columns_kernel, g, b, 0, s, d, cols, az_amp_of(cols, n));
    hipLaunchKernelGGL(az_uniform_kernel, dim3(1), dim3(1024), 0, s, (const AzColumn*)cols, n, d.prf,
                       az_uniform_of(cols, n));
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

// Per-sample quantities of the range-Doppler stage: the interpolation
// position n0 + fr (t = j + (R/D - R)/dR, split so that fr keeps full
// precision), the matched-filter phasor and the amplitude factor.
template <typename T>
struct SampleGeo {
    int n0;
    T fr;
    T c, s;
    T amp;  // FP32: multiply; FP64: divide (the reference divides)
};

// FP32 chain.  FP64: f = f_a + f_dc, u = lambda^2 f^2 / (4 v^2),
// D = sqrt(1 - u); the migration (R/D - R)/dR = rdr u / ((1 + D) D) avoids
// the cancellation in 1/D - 1; the phase 4 PI R D / lambda + 2 PI f t_c is
// formed in turns and reduced to [-1/2, 1/2] before the FP32 hardware
// sin/cos (which take turns).
__device__ __forceinline__ SampleGeo<float> geo32(const AzColumn& col, double fa, int j, bool rcmc) {
    SampleGeo<float> g;
    const double f = fa + col.fdc;
    const double u = col.a * (f * f);
    const double D = sqrt(1.0 - u);
    const double den = (1.0 + D) * D;
    double r = __builtin_amdgcn_rcp(den);
    r = r * (2.0 - den * r);
    const double shift = col.rdr * (u * r);
    const double fl = floor(shift);
    g.n0 = rcmc ? j + (int)fl : j;
    g.fr = rcmc ? (float)(shift - fl) : 0.0f;
    const double pt = col.bt * D + col.ct * f;
    const float pr = (float)(pt - rint(pt));
    g.s = __builtin_amdgcn_sinf(pr);
    g.c = __builtin_amdgcn_cosf(pr);
    g.amp = (float)col.amp;
    return g;
}

// FP64 path: the reference's expressions and order (s1_decoder.cpp:599-676,
// image_formation.cpp:439-441,469); the amplitude sqrt((Rref/R)^3) is the
// per-column value of the same expression, and the phase 4 PI R D / lambda
// (about 2e8 rad) is reduced mod 2 pi (three-part 2 pi) before sincos.
__device__ __forceinline__ SampleGeo<double> geo64(const AzColumn& col, const s1g_azimuth_desc& d, double fa, int j,
                                                   bool rcmc) {
#pragma clang fp contract(off)
    const double PI = 3.14159265358979;
    const double R = d.slant_ranges[j], v = d.v_rel[j], fdc = col.fdc, tc = d.t_corr[j];
    const double lam = d.wavelength;
    const double f = fa + fdc;
    const double D = sqrt(1 - ((lam * lam) * (f * f)) / (4 * v * v));
    SampleGeo<double> g;
    const double shift = (R / D) - R;
    const double t = (double)j + shift / d.range_spacing;
    const double fl = floor(t);
    g.n0 = rcmc ? (int)fl : j;
    g.fr = rcmc ? t - fl : 0.0;
    g.amp = col.sq;
    const double ph1 = reduce_2pi((((4.0 * PI) * R) * D) / lam);
    const double ph2 = ((2.0 * PI) * (fa + fdc)) * tc;
    double s1, c1, s2, c2;
    sincos(ph1, &s1, &c1);
    sincos(ph2, &s2, &c2);
    g.c = c1 * c2 - s1 * s2;
    g.s = c1 * s2 + s1 * c2;
    return g;
}

// w[k + 8] = sinc(fr - k) * 0.5 (1 + cos(PI (fr - k) / 8))
//          = (-1)^k (sin(PI fr) / 2 PI) (1 + cos(PI fr/8) cos(PI k/8) + sin(PI fr/8) sin(PI k/8)) / (fr - k).
// sinpi keeps full relative accuracy of sin(PI fr) for fr near 0 or 1.
__device__ __forceinline__ void sinc_weights(float fr, float* w) {
    const float S = sinpif(fr);
    const float q = fr * 0.0625f;  // PI fr / 8 in turns
    const float s8 = __builtin_amdgcn_sinf(q), c8 = __builtin_amdgcn_cosf(q);
    const float A = S * 0.15915494309189535f;
#pragma unroll
    for (int k = -kTaps; k <= kTaps; ++k) {
        const float num = 1.0f + (c8 * (float)kCosK8[k + kTaps] + s8 * (float)kSinK8[k + kTaps]);
        const float v = A * num * __builtin_amdgcn_rcpf(fr - (float)k);
        w[k + kTaps] = (k & 1) ? -v : v;
    }
    // the two exact sample positions: 0 * inf above.  fr = 1 is reached when
    // an FP64 offset within 2^-25 below an integer rounds up to 1.0f; its
    // weights select sample n0 + 1, the reference's result at t = n0 + 1.
    if (fr == 0.0f) w[kTaps] = 1.0f;
    if (fr == 1.0f) w[kTaps + 1] = 1.0f;
}

// FP64: the reference's per-tap expression sin(PI x)/(PI x) * window with
// x = fr - k; the two nearest taps take exactly its arguments, PI fr
// (k = 0) and -PI (1 - fr) (k = 1); even taps reuse sin(PI fr), odd taps
// -sin(PI (1 - fr)).
__device__ __forceinline__ void sinc_weights(double fr, double* w) {
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

// One workgroup per (Doppler row i, run of NT * EPT range samples).  The
// row segment the run's 17-tap windows reach is staged in LDS with zeros
// outside [0, N), so the taps need no bounds tests; a run whose migration
// span exceeds the tile falls back to bounds-checked global reads.
template <typename T, int NT, int EPT>
__global__ __launch_bounds__(NT) void rcmc_mf_kernel(s1g_azimuth_desc d, const AzColumn* __restrict__ cols,
                                                     const cpx<T>* __restrict__ in, int64_t ld_in,
                                                     cpx<T>* __restrict__ out, int64_t ld_out, int cap) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<T>* tile = (cpx<T>*)smem_raw;
    __shared__ int s_lo, s_hi;
    const int N = d.n_samples, L = d.n_lines;
    const int i = blockIdx.y;
    const int tid = threadIdx.x;
    const int j0 = blockIdx.x * NT * EPT;
    const bool rcmc = d.apply_rcmc != 0;
    // Doppler frequency of row i: linspace(-prf/2, prf/2, L)[i]
    double fa;
    {
#pragma clang fp contract(off)
        const double start = -d.prf / 2, end = d.prf / 2;
        const double step = (end - start) / (L - 1);
        fa = start + i * step;
    }
    const cpx<T>* row = in + (int64_t)i * ld_in;
    SampleGeo<T> g[EPT];
    int lo = 1 << 30, hi = -(1 << 30);
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int j = j0 + e * NT + tid;
        g[e].n0 = j;
        g[e].fr = 0;
        g[e].c = g[e].s = g[e].amp = 0;
        if (j < N) {
            if constexpr (sizeof(T) == 4)
                g[e] = geo32(cols[j], fa, j, rcmc);
            else
                g[e] = geo64(cols[j], d, fa, j, rcmc);
            lo = min(lo, g[e].n0 - kTaps);
            hi = max(hi, g[e].n0 + kTaps);
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        lo = min(lo, __shfl_xor(lo, o));
        hi = max(hi, __shfl_xor(hi, o));
    }
    if (tid == 0) {
        s_lo = 1 << 30;
        s_hi = -(1 << 30);
    }
    __syncthreads();
    if ((tid & 63) == 0) {
        atomicMin(&s_lo, lo);
        atomicMax(&s_hi, hi);
    }
    __syncthreads();
    const int tlo = s_lo, thi = s_hi;
    const bool use_tile = (thi - tlo + 1) <= cap;
    if (use_tile)
        for (int k = tlo + tid; k <= thi; k += NT) tile[k - tlo] = (k >= 0 && k < N) ? row[k] : mk((T)0, (T)0);
    __syncthreads();
    cpx<T>* orow = out + (int64_t)i * ld_out;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int j = j0 + e * NT + tid;
        if (j >= N) continue;
        cpx<T> acc = mk((T)0, (T)0);
        if (rcmc) {
            T w[2 * kTaps + 1];
            sinc_weights(g[e].fr, w);
            const int n0 = g[e].n0;
            if (use_tile) {
                const cpx<T>* tp = tile + (n0 - kTaps - tlo);
#pragma unroll
                for (int k = 0; k <= 2 * kTaps; ++k) {
                    const cpx<T> x = tp[k];
                    acc.x += x.x * w[k];
                    acc.y += x.y * w[k];
                }
            } else {
#pragma unroll
                for (int k = 0; k <= 2 * kTaps; ++k) {
                    const int idx = n0 - kTaps + k;
                    if (idx < 0 || idx >= N) continue;
                    const cpx<T> x = row[idx];
                    acc.x += x.x * w[k];
                    acc.y += x.y * w[k];
                }
            }
        } else {
            acc = row[j];
        }
        if constexpr (sizeof(T) == 8)
            acc = mk(acc.x / g[e].amp, acc.y / g[e].amp);
        else
            acc = scl(acc, g[e].amp);
        orow[j] = cmul(acc, mk(g[e].c, g[e].s));
    }
}

// FP32 chain, runs of consecutive samples.  For u = lambda^2 f^2 / (4 v^2)
// up to kSeriesU the geometry takes the series of 1/D - 1 and 1 - D
// (D = sqrt(1 - u)) to u^5, whose truncation is below 5e-16 relative; the
// phase bt D + ct f is formed as (bt - rint(bt)) - bt (1 - D) + ct f.
constexpr double kSeriesU = 1e-3;

__device__ __forceinline__ void geo32_series(const AzColumn& col, double fa, int j, bool rcmc, int& n0, double& fr,
                                             float& c, float& s, float& amp) {
    const double f = fa + col.fdc;
    const double u = col.a * (f * f);
    double idm1, omd;  // 1/D - 1 and 1 - D
    if (u <= kSeriesU) {
        idm1 = u * (0.5 + u * (0.375 + u * (0.3125 + u * (0.2734375 + u * 0.24609375))));
        omd = u * (0.5 + u * (0.125 + u * (0.0625 + u * (0.0390625 + u * 0.02734375))));
    } else {
        const double D = sqrt(1.0 - u);
        omd = u / (1.0 + D);
        idm1 = omd / D;
    }
    const double shift = col.rdr * idm1;
    const double fl = floor(shift);
    n0 = rcmc ? j + (int)fl : j;
    fr = rcmc ? shift - fl : 0.0;
    const double pt = (col.btf - col.bt * omd) + col.ct * f;
    const float pr = (float)(pt - rint(pt));
    s = __builtin_amdgcn_sinf(pr);
    c = __builtin_amdgcn_cosf(pr);
    amp = (float)col.amp;
}

// Largest spread of the fractional offset over one run for which the
// weights are interpolated: the windowed sinc has |w''| < 20, so the linear
// interpolation error stays below 20 * (2.5e-4)^2 / 8 = 1.6e-7.
constexpr double kInterpSpan = 2.5e-4;

// One workgroup per (Doppler row i, NT runs of EPT consecutive samples).
// Along a row the migration changes by at most u/2 per sample, so within a
// run the tap set usually moves with the sample (n0 = j + const) and the
// fractional offset by less than kInterpSpan: the run then evaluates the
// windowed-sinc weights exactly at its first and last samples, interpolates
// them linearly in the offset for the others, and reads its 2*8 + EPT taps
// from the LDS tile once.  Other runs take the per-sample path.  The
// geometry (column table reads) and the matched-filter epilogue (stores)
// run with consecutive lanes on consecutive samples; LDS hands n0 and the
// offset to the runs and the runs' sums back (EPT odd keeps the lanes' runs
// on distinct LDS banks).
template <int NT, int EPT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void rcmc32_run_kernel(s1g_azimuth_desc d, const AzColumn* __restrict__ cols,
                                                        const cpx<float>* __restrict__ in, int64_t ld_in,
                                                        cpx<float>* __restrict__ out, int64_t ld_out, int cap) {
    constexpr int NW = 2 * kTaps + EPT;
    constexpr int PER = NT * EPT;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float* s_fr = (float*)smem_raw;                    // PER
    int* s_n0 = (int*)(s_fr + PER);                    // PER
    cpx<float>* tile = (cpx<float>*)(s_n0 + PER);      // cap (>= PER): taps, then the runs' sums
    __shared__ int s_lo[NT / 64], s_hi[NT / 64];
    const int N = d.n_samples, L = d.n_lines;
    const int i = blockIdx.y;
    const int tid = threadIdx.x;
    const int j0 = blockIdx.x * PER;
    const bool rcmc = d.apply_rcmc != 0;
    double fa;
    {
#pragma clang fp contract(off)
        const double start = -d.prf / 2, end = d.prf / 2;
        const double step = (end - start) / (L - 1);
        fa = start + i * step;
    }
    const cpx<float>* row = in + (int64_t)i * ld_in;
    // ---- geometry, consecutive lanes on consecutive samples
    float cs[EPT], sn[EPT], amp[EPT];
    int lo = 1 << 30, hi = -(1 << 30);
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int q = e * NT + tid;
        const int j = j0 + q;
        int n0 = j;
        double fr = 0.0;
        cs[e] = sn[e] = amp[e] = 0.0f;
        if (j < N) {
            geo32_series(cols[j], fa, j, rcmc, n0, fr, cs[e], sn[e], amp[e]);
            lo = min(lo, n0 - kTaps);
            hi = max(hi, n0 + kTaps);
        }
        s_fr[q] = (float)fr;
        s_n0[q] = n0;
    }
    for (int o = 32; o > 0; o >>= 1) {
        lo = min(lo, __shfl_xor(lo, o));
        hi = max(hi, __shfl_xor(hi, o));
    }
    if ((tid & 63) == 0) {
        s_lo[tid >> 6] = lo;
        s_hi[tid >> 6] = hi;
    }
    __syncthreads();
    int tlo = s_lo[0], thi = s_hi[0];
#pragma unroll
    for (int w = 1; w < NT / 64; ++w) {
        tlo = min(tlo, s_lo[w]);
        thi = max(thi, s_hi[w]);
    }
    const bool use_tile = (thi - tlo + 1) <= cap;
    if (use_tile)
        for (int k = tlo + tid; k <= thi; k += NT) tile[k - tlo] = (k >= 0 && k < N) ? row[k] : mk(0.0f, 0.0f);
    __syncthreads();
    // ---- the run of samples j0 + tid * EPT + e
    const int r0 = tid * EPT;
    const int js = j0 + r0;
    cpx<float> acc[EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) acc[e] = mk(0.0f, 0.0f);
    if (js < N) {
        int n0[EPT];
        float fr[EPT];
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
            n0[e] = s_n0[r0 + e];
            fr[e] = s_fr[r0 + e];
        }
        bool fast = use_tile && rcmc && js + EPT <= N && fabsf(fr[EPT - 1] - fr[0]) <= (float)kInterpSpan;
#pragma unroll
        for (int e = 1; e < EPT; ++e) fast = fast && n0[e] == n0[0] + e;
        if (fast) {
            float w0[2 * kTaps + 1], dw[2 * kTaps + 1];
            sinc_weights(fr[0], w0);
            sinc_weights(fr[EPT - 1], dw);
#pragma unroll
            for (int k = 0; k <= 2 * kTaps; ++k) dw[k] -= w0[k];
            const float dfr = fr[EPT - 1] - fr[0];
            const float inv = dfr != 0.0f ? 1.0f / dfr : 0.0f;
            const cpx<float>* tp = tile + (n0[0] - kTaps - tlo);
            cpx<float> x[NW];
#pragma unroll
            for (int m = 0; m < NW; ++m) x[m] = tp[m];
#pragma unroll
            for (int e = 0; e < EPT; ++e) {
                const float t = (fr[e] - fr[0]) * inv;
#pragma unroll
                for (int k = 0; k <= 2 * kTaps; ++k) {
                    const float w = fmaf(t, dw[k], w0[k]);
                    acc[e].x += x[e + k].x * w;
                    acc[e].y += x[e + k].y * w;
                }
            }
        } else {
#pragma unroll
            for (int e = 0; e < EPT; ++e) {
                const int j = js + e;
                if (j >= N) continue;
                if (!rcmc) {
                    acc[e] = row[j];
                    continue;
                }
                float w[2 * kTaps + 1];
                sinc_weights(fr[e], w);
                if (use_tile) {
                    const cpx<float>* tp = tile + (n0[e] - kTaps - tlo);
#pragma unroll
                    for (int k = 0; k <= 2 * kTaps; ++k) {
                        acc[e].x += tp[k].x * w[k];
                        acc[e].y += tp[k].y * w[k];
                    }
                } else {
#pragma unroll
                    for (int k = 0; k <= 2 * kTaps; ++k) {
                        const int idx = n0[e] - kTaps + k;
                        if (idx < 0 || idx >= N) continue;
                        const cpx<float> xv = row[idx];
                        acc[e].x += xv.x * w[k];
                        acc[e].y += xv.y * w[k];
                    }
                }
            }
        }
    }
    __syncthreads();  // every run has read its taps: the tile now takes the sums
#pragma unroll
    for (int e = 0; e < EPT; ++e) tile[r0 + e] = acc[e];
    __syncthreads();
    // ---- amplitude and matched filter, consecutive lanes on consecutive samples
    cpx<float>* orow = out + (int64_t)i * ld_out;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int q = e * NT + tid;
        const int j = j0 + q;
        if (j < N) orow[j] = cmul(scl(tile[q], amp[e]), mk(cs[e], sn[e]));
    }
}

template <int NT, int EPT>
static void launch_rcmc32_run(hipStream_t s, const s1g_azimuth_desc& d, const AzColumn* cols, const void* in,
                              int64_t ld_in, void* out, int64_t ld_out) {
    const int per = NT * EPT;
    const int cap = per + 2 * kTaps + 256;  // migration spans beyond 256 samples read global memory
    const size_t lds = (sizeof(float) + sizeof(int)) * (size_t)per + sizeof(cpx<float>) * (size_t)cap;
    dim3 grid((unsigned)((d.n_samples + per - 1) / per), (unsigned)d.n_lines);
    hipLaunchKernelGGL((rcmc32_run_kernel<NT, EPT>), grid, dim3(NT), lds, s, d, cols, (const cpx<float>*)in, ld_in,
                       (cpx<float>*)out, ld_out, cap);
}

// 1/D - 1 and 1 - D (D = sqrt(1 - u)): the series to u^5 up to kSeriesU,
// else the square root without cancellation.
__device__ __forceinline__ void migration_terms(double u, double& idm1, double& omd) {
    if (u <= kSeriesU) {
        idm1 = u * (0.5 + u * (0.375 + u * (0.3125 + u * (0.2734375 + u * 0.24609375))));
        omd = u * (0.5 + u * (0.125 + u * (0.0625 + u * (0.0390625 + u * 0.02734375))));
    } else {
        const double D = sqrt(1.0 - u);
        omd = u / (1.0 + D);
        idm1 = omd / D;
    }
}

// Tile slack for the spread of the integer migration over one workgroup
// (at most a few samples at Sentinel-1 geometries).
constexpr int kLinSlack = 64;

// LDS position of tile element a: one pad element per 16.  A thread's run
// starts EPT = 8 elements after its neighbour's, a 64-byte lane stride that
// would put the lanes of one access on a few of the 32 banks (the LDS was
// stalled by bank conflicts for a third of the kernel); the pad rotates
// each 128-byte row of the tile by one element.
__device__ __forceinline__ int lin_pad(int a) { return a + (a >> 4); }

// FP32 chain, linear column geometry (AzUniform::ok).  With v_rel and f_dc
// the same in every column, D, 1/D - 1 and 1 - D depend on the Doppler row
// only; R_j / dR, bt and t_corr are linear in j, so the migration is
// shift_j = (rdr0 + j drdr)(1/D - 1) and the phase in turns
// pt_j = P0 + j P1 with P0 = bt0 D + ct0 f, P1 = dbt D + dct f (FP64, the
// fraction taken before the FP32 sin/cos).  A thread takes EPT consecutive
// samples: the migration moves by (1/D - 1) drdr (below 6e-6 samples here)
// per sample, so when the run keeps one integer shift and its offset spread
// is within kInterpSpan the windowed-sinc weights are evaluated at its first
// and last sample and interpolated linearly for the others (the offset is
// linear in the sample), and the run reads its 2*8 + EPT taps from the LDS
// row tile once.  Other runs, and tiles that would exceed kLinSlack, take
// exact per-sample weights.  Without linear geometry every sample takes the
// column table's geometry and exact weights.
//
// One launch may serve several bursts (RcmcBatch: per-instance buffers,
// column table and shape; workgroups numbered consecutively across
// instances, nbx per Doppler row).
template <int NT, int EPT>
__global__ __launch_bounds__(NT) void rcmc32_lin_kernel(RcmcBatch bt) {
    constexpr int PER = NT * EPT;
    constexpr int CAP = PER + 2 * kTaps + kLinSlack;
    constexpr int NW = 2 * kTaps + EPT;
    __shared__ cpx<float> tile[CAP + CAP / 16 + 1];  // indexed through lin_pad
    // the workgroup's instance (uniform: a scan of the kernel arguments)
    int kI = 0;
    while (kI + 1 < bt.n && (int)blockIdx.x >= bt.inst[kI + 1].wg0) ++kI;
    const RcmcInst& I = bt.inst[kI];
    const AzColumn* __restrict__ cols = I.cols;
    const AzUniform* __restrict__ uni = I.uni;
    const float* __restrict__ amp32 = I.amp32;
    const cpx<float>* __restrict__ in = (const cpx<float>*)I.in;
    cpx<float>* __restrict__ out = (cpx<float>*)I.out;
    const int64_t ld_in = I.ld_in, ld_out = I.ld_out;
    const int N = I.N, L = I.n_lines;
    const int local = (int)blockIdx.x - I.wg0;
    const int i = local / I.nbx;
    const int tid = threadIdx.x;
    const int j0 = (local - i * I.nbx) * PER;
    const bool rcmc = I.apply_rcmc != 0;
    double fa;
    {
#pragma clang fp contract(off)
        const double start = -I.prf / 2, end = I.prf / 2;
        const double step = (end - start) / (L - 1);
        fa = start + i * step;
    }
    const cpx<float>* row = in + (int64_t)i * ld_in;
    cpx<float>* orow = out + (int64_t)i * ld_out;
    const int js = j0 + tid * EPT;
    cpx<float> acc[EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) acc[e] = mk(0.0f, 0.0f);
    // per-sample taps and exact weights (bounds-checked global reads)
    auto sample = [&](int e, int n0, float fr) {
        float w[2 * kTaps + 1];
        sinc_weights(fr, w);
#pragma unroll
        for (int k = 0; k <= 2 * kTaps; ++k) {
            const int idx = n0 - kTaps + k;
            if (idx < 0 || idx >= N) continue;
            const cpx<float> x = row[idx];
            acc[e].x += x.x * w[k];
            acc[e].y += x.y * w[k];
        }
    };
    if (uni->ok) {
        const AzUniform g = *uni;
        const double f = fa + g.fdc;
        double idm1, omd;
        migration_terms(g.a * (f * f), idm1, omd);
        if (!rcmc) idm1 = 0.0;
        const double D = 1.0 - omd;
        const double P0 = g.bt0 * D + g.ct0 * f;
        const double P1 = g.dbt * D + g.dct * f;
        // the workgroup's tile: every tap its runs can reach
        const int jl = min(j0 + PER, N) - 1;
        const double shA = (g.rdr0 + j0 * g.drdr) * idm1, shB = (g.rdr0 + jl * g.drdr) * idm1;
        const int tlo = j0 + (int)floor(fmin(shA, shB)) - kTaps;
        const int thi = jl + (int)floor(fmax(shA, shB)) + kTaps;
        const bool use_tile = rcmc && thi - tlo + 1 <= CAP;
        if (use_tile) {
            // every load of the tile issued before the first LDS store (raw
            // buffer loads: offsets outside the row read zero, the padding)
            constexpr int NLD = (CAP + NT - 1) / NT;
            const __amdgpu_buffer_rsrc_t rs = buf_rsrc(row, (uint32_t)N * (uint32_t)sizeof(cpx<float>));
            cpx<float> v[NLD];
#pragma unroll
            for (int u = 0; u < NLD; ++u) {
                const int k = tlo + tid + NT * u;
                v[u] = buf_load<float>(rs, (k >= 0 && k <= thi) ? (uint32_t)k * (uint32_t)sizeof(cpx<float>) : kBufSkip);
            }
#pragma unroll
            for (int u = 0; u < NLD; ++u)
                if (tid + NT * u <= thi - tlo) tile[lin_pad(tid + NT * u)] = v[u];
        }
        __syncthreads();
        if (js < N) {
            const double s0 = (g.rdr0 + js * g.drdr) * idm1;
            const double s1 = (g.rdr0 + (js + EPT - 1) * g.drdr) * idm1;
            const double f0 = floor(s0), f1 = floor(s1);
            if (!rcmc) {
#pragma unroll
                for (int e = 0; e < EPT; ++e)
                    if (js + e < N) acc[e] = row[js + e];
            } else if (use_tile && f0 == f1 && js + EPT <= N && fabs(s1 - s0) <= kInterpSpan) {
                float w0[2 * kTaps + 1], dw[2 * kTaps + 1];
                sinc_weights((float)(s0 - f0), w0);
                sinc_weights((float)(s1 - f0), dw);
#pragma unroll
                for (int k = 0; k <= 2 * kTaps; ++k) dw[k] -= w0[k];
                const int tb = js + (int)f0 - kTaps - tlo;
                cpx<float> x[NW];
#pragma unroll
                for (int m = 0; m < NW; ++m) x[m] = tile[lin_pad(tb + m)];
#pragma unroll
                for (int e = 0; e < EPT; ++e) {
                    const float t = (float)e / (float)(EPT - 1);
#pragma unroll
                    for (int k = 0; k <= 2 * kTaps; ++k) {
                        const float w = fmaf(t, dw[k], w0[k]);
                        acc[e].x += x[e + k].x * w;
                        acc[e].y += x[e + k].y * w;
                    }
                }
            } else {
#pragma unroll
                for (int e = 0; e < EPT; ++e) {
                    const int j = js + e;
                    if (j >= N) continue;
                    const double sh = (g.rdr0 + j * g.drdr) * idm1;
                    const double fl = floor(sh);
                    sample(e, j + (int)fl, (float)(sh - fl));
                }
            }
#pragma unroll
            for (int e = 0; e < EPT; ++e) {
                const int j = js + e;
                if (j >= N) continue;
                const double pt = fma((double)j, P1, P0);
                const float pr = (float)(pt - rint(pt));
                const cpx<float> ph = mk(__builtin_amdgcn_cosf(pr), __builtin_amdgcn_sinf(pr));
                acc[e] = cmul(scl(acc[e], amp32[j]), ph);
            }
        }
        // hand the runs' samples back through LDS so that consecutive lanes
        // store consecutive samples (whole lines, no partial-line writes)
        __syncthreads();
#pragma unroll
        for (int e = 0; e < EPT; ++e) tile[lin_pad(tid * EPT + e)] = acc[e];
        __syncthreads();
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
            const int q = e * NT + tid;
            if (j0 + q < N) orow[j0 + q] = tile[lin_pad(q)];
        }
    } else if (js < N) {
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
            const int j = js + e;
            if (j >= N) continue;
            int n0;
            double fr;
            float c, sn, a;
            geo32_series(cols[j], fa, j, rcmc, n0, fr, c, sn, a);
            if (rcmc) sample(e, n0, (float)fr);
            else acc[e] = row[j];
            orow[j] = cmul(scl(acc[e], a), mk(c, sn));
        }
    }
}

template <int NT, int EPT>
static void launch_rcmc32_lin(hipStream_t s, const RcmcInst* inst, int n) {
    constexpr int PER = NT * EPT;
    for (int b0 = 0; b0 < n; b0 += kMaxBatch) {
        RcmcBatch bt;
        bt.n = std::min(kMaxBatch, n - b0);
        int nwg = 0;
        for (int k = 0; k < bt.n; ++k) {
            bt.inst[k] = inst[b0 + k];
            bt.inst[k].nbx = (bt.inst[k].N + PER - 1) / PER;
            bt.inst[k].wg0 = nwg;
            nwg += bt.inst[k].nbx * bt.inst[k].n_lines;
        }
        if (nwg > 0) hipLaunchKernelGGL((rcmc32_lin_kernel<NT, EPT>), dim3((unsigned)nwg), dim3(NT), 0, s, bt);
    }
}

static RcmcInst rcmc_inst(const s1g_azimuth_desc& d, const AzColumn* cols, const void* in, int64_t ld_in, void* out,
                          int64_t ld_out) {
    RcmcInst r;
    AzColumn* c = const_cast<AzColumn*>(cols);
    r.in = in;
    r.out = out;
    r.ld_in = ld_in;
    r.ld_out = ld_out;
    r.cols = cols;
    r.uni = az_uniform_of(c, d.n_samples);
    r.amp32 = az_amp_of(c, d.n_samples);
    r.prf = d.prf;
    r.n_lines = d.n_lines;
    r.N = d.n_samples;
    r.apply_rcmc = d.apply_rcmc;
    r.nbx = 0;
    r.wg0 = 0;
    return r;
}

template <int NT, int EPT>
static void launch_rcmc32_lin(hipStream_t s, const s1g_azimuth_desc& d, const AzColumn* cols, const void* in,
                              int64_t ld_in, void* out, int64_t ld_out) {
    const RcmcInst r = rcmc_inst(d, cols, in, ld_in, out, ld_out);
    launch_rcmc32_lin<NT, EPT>(s, &r, 1);
}

bool rcmc_batchable(int prec, int variant) { return prec == S1G_F32 && (variant == 5 || variant == 6); }

int launch_rcmc_batch(hipStream_t s, const RcmcInst* inst, int n, int variant) {
    if (variant == 5) launch_rcmc32_lin<256, 16>(s, inst, n);
    else if (variant == 6) launch_rcmc32_lin<256, 8>(s, inst, n);
    else return -1;
    return (int)hipGetLastError();
}

int rcmc_variants(int prec) { return prec == S1G_F32 ? 7 : 3; }

template <typename T, int NT, int EPT>
static void launch_rcmc_t(hipStream_t s, const s1g_azimuth_desc& d, const AzColumn* cols, const void* in,
                          int64_t ld_in, void* out, int64_t ld_out) {
    const int per = NT * EPT;
    const int cap = per + 1024;
    dim3 grid((unsigned)((d.n_samples + per - 1) / per), (unsigned)d.n_lines);
    hipLaunchKernelGGL((rcmc_mf_kernel<T, NT, EPT>), grid, dim3(NT), sizeof(cpx<T>) * cap, s, d, cols,
                       (const cpx<T>*)in, ld_in, (cpx<T>*)out, ld_out, cap);
}

// Variants (threads x samples per thread): 0 = 256 x 8, 1 = 256 x 4, 2 = 512 x 4;
// FP32 also runs of 9 consecutive samples per thread (rcmc32_run_kernel):
// 3 = 256 threads, 4 = 128 threads; and the linear-geometry kernel
// (rcmc32_lin_kernel), 256 threads with runs of 5 = 16, 6 = 8 samples.
int launch_rcmc_mf(hipStream_t s, int prec, const s1g_azimuth_desc& d, const AzColumn* cols, const void* in,
                   int64_t ld_in, void* out, int64_t ld_out, int variant) {
    if (prec == S1G_F32) {
        if (variant == 5) launch_rcmc32_lin<256, 16>(s, d, cols, in, ld_in, out, ld_out);
        else if (variant == 6) launch_rcmc32_lin<256, 8>(s, d, cols, in, ld_in, out, ld_out);
        else if (variant == 3) launch_rcmc32_run<256, 9>(s, d, cols, in, ld_in, out, ld_out);
        else if (variant == 4) launch_rcmc32_run<128, 9>(s, d, cols, in, ld_in, out, ld_out);
        else if (variant == 1) launch_rcmc_t<float, 256, 4>(s, d, cols, in, ld_in, out, ld_out);
        else if (variant == 2) launch_rcmc_t<float, 512, 4>(s, d, cols, in, ld_in, out, ld_out);
        else launch_rcmc_t<float, 256, 8>(s, d, cols, in, ld_in, out, ld_out);
    } else {
        if (variant == 1) launch_rcmc_t<double, 256, 4>(s, d, cols, in, ld_in, out, ld_out);
        else if (variant == 2) launch_rcmc_t<double, 512, 4>(s, d, cols, in, ld_in, out, ld_out);
        else launch_rcmc_t<double, 256, 8>(s, d, cols, in, ld_in, out, ld_out);
    }
    return (int)hipGetLastError();
}

}  // namespace s1g
