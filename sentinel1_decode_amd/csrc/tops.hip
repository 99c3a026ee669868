// tops.hip -- TOPS (IW) focusing: the Doppler-centroid estimate, the
// azimuth-frequency and azimuth-time unfolding (UFR) around the range-Doppler
// line filter, as the TOPS branch of S1_Decoder::_azimuth_compress
// (s1_decoder.cpp:511-701) composes them.
//
// Layout: the unfolding works per range sample along azimuth, so each UFR
// transposes the burst once to [samples][lines] (azimuth contiguous), runs
// its tile / ramp / FFT / gate / FFT / ramp steps as row operations there,
// and transposes back.  Ramps that depend on one index only (the frequency
// ramps, the time axes) and the resampling positions are tables built on the
// host in FP64 with the reference's own loops, so the device applies exactly
// the reference's weights; per-element phases are FP64 on the device.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "s1g_internal.h"

using namespace s1g;

namespace {

constexpr double kPi = 3.14159265358979;  // the reference's PI (structs.h:23)
constexpr double kC = 299792458.0;
constexpr double kF0 = 5.405000454334350e9;
constexpr double kWavelength = kC / kF0;

// Slots of the context's aux buffers used here.
enum { kBufA = 0, kBufB = 1, kBufC = 2, kTables = 3, kFmT = 4, kOut = 5, kVrel = 6, kFactors = 7 };

int status(s1g_ctx* c, const char* what) {
    const hipError_t e = hipGetLastError();
    if (e == hipSuccess) return S1G_OK;
    char buf[160];
    snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    return ctx_error(c, S1G_EHIP, buf);
}

// Row DFTs (ctx_dft_rows batches rows so that the work arrays stay bounded).
int dft_rows_batched(s1g_ctx* c, int prec, int rows, int n, void* data, int64_t ld, int inverse, bool normalise) {
    return ctx_dft_rows(c, prec, rows, n, data, ld, inverse, normalise);
}

// linspace (signal_processing.cpp:17-28): start + i step.
void linspace(double start, double end, int size, double* out) {
#pragma clang fp contract(off)
    const double step = (end - start) / (size - 1);
    for (int i = 0; i < size; ++i) out[i] = start + i * step;
}

// Where linear_resample (signal_processing.cpp:32-63) evaluates each output:
// its running abscissa x and input segment, walked as the reference walks
// them; seg = -1 for an output the walk never reaches (left 0), -2 for the
// last output (the last input).
struct LinTab {
    int32_t seg;
    double x;
};
std::vector<LinTab> linear_walk(int n_in, int n_out) {
#pragma clang fp contract(off)
    std::vector<LinTab> t((size_t)n_out, LinTab{-1, 0.0});
    if (n_in == n_out) {
        for (int k = 0; k < n_out; ++k) t[k] = LinTab{-3, 0.0};  // identity
        return t;
    }
    double x = 0.0;
    const double dx_in = 1.0 / (n_in - 1), dx_out = 1.0 / (n_out - 1);
    int index = 0;
    for (int i = 0; i < n_in - 1; i++) {
        const double x1 = (i + 1) * dx_in;
        while (x < x1 && index < n_out) {
            t[index] = LinTab{i, x};
            x += dx_out;
            index += 1;
        }
    }
    t[n_out - 1] = LinTab{-2, 0.0};
    return t;
}

// One output of linear_resample from its table entry and the input row.
__device__ __forceinline__ double linear_at(const LinTab& e, int k, const double* __restrict__ row, int n_in,
                                            double dx_in) {
#pragma clang fp contract(off)
    if (e.seg == -3) return row[k];
    if (e.seg == -2) return row[n_in - 1];
    if (e.seg < 0) return 0.0;
    const double x0 = e.seg * dx_in, x1 = (e.seg + 1) * dx_in;
    const double y0 = row[e.seg], y1 = row[e.seg + 1];
    return y0 + ((e.x - x0) * (y1 - y0) / (x1 - x0));
}

// quadratic_resample (signal_processing.cpp:109-155): the three points and
// Lagrange weights of output j.
struct QuadTab {
    int32_t i0, i1, i2;
    double l0, l1, l2;
};
QuadTab quadratic_at(int j, int n_in, int n_out) {
#pragma clang fp contract(off)
    const double dx_in = 1.0 / (n_in - 1), dx_out = 1.0 / (n_out - 1);
    const double x = j * dx_out;
    const int i = std::min(int(x / dx_in), n_in - 2);
    const int i0 = std::max(i - 1, 0), i1 = i, i2 = std::min(i + 1, n_in - 1);
    const double x0 = i0 * dx_in, x1 = i1 * dx_in, x2 = i2 * dx_in;
    QuadTab q;
    q.i0 = i0;
    q.i1 = i1;
    q.i2 = i2;
    q.l0 = (x - x1) * (x - x2) / ((x0 - x1) * (x0 - x2));
    q.l1 = (x - x0) * (x - x2) / ((x1 - x0) * (x1 - x2));
    q.l2 = (x - x0) * (x - x1) / ((x2 - x0) * (x2 - x1));
    return q;
}

template <typename T>
__device__ __forceinline__ cpx<double> ld64(const cpx<T>& v) {
    cpx<double> r;
    r.x = (double)v.x;
    r.y = (double)v.y;
    return r;
}
template <typename T>
__device__ __forceinline__ cpx<T> st(double re, double im) {
    cpx<T> r;
    r.x = (T)re;
    r.y = (T)im;
    return r;
}

// --------------------------------------------------- Doppler centroid
// Per range sample j: sum over lines of y_i conj(y_{i+1}), y_i = x_ij dr_i,
// serially in line order and without contractions, as get_wrapped_estimates
// (doppler.cpp:4-39) sums.  A thread per sample; consecutive threads read
// consecutive samples of a line.
template <typename T>
__global__ __launch_bounds__(256) void dc_lag_kernel(const cpx<T>* __restrict__ x, int64_t ld, int L, int n,
                                                     const cpx<double>* __restrict__ dr,
                                                     cpx<double>* __restrict__ sums) {
#pragma clang fp contract(off)
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    auto ramp = [&](int i, const cpx<T>& v) {
#pragma clang fp contract(off)
        const cpx<double> a = ld64(v), d = dr[i];
        cpx<double> r;
        r.x = a.x * d.x - a.y * d.y;
        r.y = a.x * d.y + a.y * d.x;
        return r;
    };
    cpx<double> prev = ramp(0, x[j]);
    double sr = 0.0, si = 0.0;
    constexpr int U = 8;
    int i = 1;
    for (; i + U <= L; i += U) {
        cpx<T> v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = x[(int64_t)(i + u) * ld + j];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const cpx<double> cur = ramp(i + u, v[u]);
            // prev * conj(cur)
            sr += prev.x * cur.x - prev.y * (-cur.y);
            si += prev.x * (-cur.y) + prev.y * cur.x;
            prev = cur;
        }
    }
    for (; i < L; ++i) {
        const cpx<double> cur = ramp(i, x[(int64_t)i * ld + j]);
        sr += prev.x * cur.x - prev.y * (-cur.y);
        si += prev.x * (-cur.y) + prev.y * cur.x;
        prev = cur;
    }
    sums[j].x = sr;
    sums[j].y = si;
}

// The reference's 3 x 3 normal-equation fit (polyfit, signal_processing.cpp:
// 159-214): coefficients of 1, x, x^2.
void polyfit2(const std::vector<double>& x, const std::vector<double>& y, double c[3]) {
#pragma clang fp contract(off)
    double Sx[5] = {0, 0, 0, 0, 0}, Sy[3] = {0, 0, 0};
    for (size_t i = 0; i < x.size(); ++i) {
        double xi = 1.0;
        for (int j = 0; j <= 4; ++j) {
            if (j <= 2) Sy[j] += xi * y[i];
            Sx[j] += xi;
            xi *= x[i];
        }
    }
    double A[3][3] = {{Sx[0], Sx[1], Sx[2]}, {Sx[1], Sx[2], Sx[3]}, {Sx[2], Sx[3], Sx[4]}};
    double b[3] = {Sy[0], Sy[1], Sy[2]};
    for (int i = 0; i < 3; ++i) {
        for (int k = i + 1; k < 3; ++k) {
            if (fabs(A[k][i]) > fabs(A[i][i])) {
                for (int j = 0; j < 3; ++j) std::swap(A[i][j], A[k][j]);
                std::swap(b[i], b[k]);
            }
        }
        for (int k = i + 1; k < 3; ++k) {
            const double f = A[k][i] / A[i][i];
            for (int j = i; j < 3; ++j) A[k][j] -= f * A[i][j];
            b[k] -= f * b[i];
        }
    }
    for (int i = 2; i >= 0; --i) {
        c[i] = b[i];
        for (int j = i + 1; j < 3; ++j) c[i] -= A[i][j] * c[j];
        c[i] /= A[i][i];
    }
}

// get_unwrapped_estimates (doppler.cpp:42-95) with its 8x dense grid: the
// peak of the DFT of floor(|exp(2 pi i f / prf)|) sets a phase ramp a t + b
// (radians), and each estimate is re-wrapped about it.
std::vector<double> unwrap_estimates(const std::vector<double>& fine, const std::vector<double>& t, double prf) {
#pragma clang fp contract(off)
    const int m = (int)fine.size() * 8;
    const float dt = (float)(t[1] - t[0]);
    std::vector<double> Fr((size_t)m, 0.0), Fi((size_t)m, 0.0);
    std::vector<double> g(fine.size());
    for (size_t b = 0; b < fine.size(); ++b) {
        const double th = 2.0 * kPi * fine[b] / prf;
        g[b] = floor(hypot(cos(th), sin(th)));
    }
    // the DFT of a 0/1 sequence: its power peaks at bin 0 (sum of the ones)
    int best = 0;
    double best_p = -1.0, best_r = 0.0, best_i = 0.0;
    for (int k = 0; k < m; ++k) {
        double sr = 0.0, si = 0.0;
        for (size_t b = 0; b < fine.size(); ++b) {
            if (g[b] == 0.0) continue;
            const double w = -2.0 * kPi * (double)((int64_t)k * (int64_t)b % m) / m;
            sr += g[b] * cos(w);
            si += g[b] * sin(w);
        }
        const double a = hypot(sr, si), p = a * a;
        if (p > best_p) {
            best_p = p;
            best = k;
            best_r = sr;
            best_i = si;
        }
    }
    // fftfreq (signal_processing.cpp:501-513)
    const double val = 1.0 / (m * (double)dt);
    const int half = (m - 1) / 2 + 1;
    const double v = best < half ? best * val : (best - m) * val;
    const double a = v / dt;
    const double bb = atan2(best_i, best_r) / (2.0 * kPi);
    std::vector<double> out(fine.size());
    for (size_t b = 0; b < fine.size(); ++b) {
        const double th1 = 2.0 * kPi * fine[b] / prf, th2 = -(a * t[b] + bb);
        const double c1 = cos(th1), s1 = sin(th1), c2 = cos(th2), s2 = sin(th2);
        const double r = atan2(c1 * s2 + s1 * c2, c1 * c2 - s1 * s2);
        out[b] = (a * t[b] + bb + r / (2.0 * kPi)) * prf;
    }
    return out;
}

// ------------------------------------------------ azimuth frequency UFR
// Row j of [n][L] holds the azimuth spectrum of sample j; the output row is
// that spectrum fftshifted (std::rotate by L / 2), tiled over `filled`
// positions of `shape` (zero beyond), times the deramp conj(ramp[a]).
template <typename T>
__global__ void freq_tile_kernel(const cpx<T>* __restrict__ spec, int L, int n, int shape, int filled,
                                 const cpx<double>* __restrict__ ramp, cpx<T>* __restrict__ out) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (a >= shape || j >= n) return;
    double re = 0.0, im = 0.0;
    if (a < filled) {
        const int src = ((a % L) + L - L / 2) % L;
        const cpx<double> s = ld64(spec[(int64_t)j * L + src]);
        const double dr = ramp[a].x, di = -ramp[a].y;
        re = dr * s.x - di * s.y;
        im = dr * s.y + di * s.x;
    }
    out[(int64_t)j * shape + a] = st<T>(re, im);
}

// The low-pass gate: out[r][a] = in[r][start + a] for a < count, else 0.
template <typename T>
__global__ void gate_kernel(const cpx<T>* __restrict__ in, cpx<T>* __restrict__ out, int rows, int len, int start,
                            int count) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (a >= len || r >= rows) return;
    cpx<T> v;
    v.x = (T)0;
    v.y = (T)0;
    if (a < count) v = in[(int64_t)r * len + start + a];
    out[(int64_t)r * len + a] = v;
}

// dst[a][j] = ramp[a] src[j][a]: the re-ramp fused into the transpose back
// to [lines][samples], through a 32 x 32 LDS tile.
template <typename T>
__global__ __launch_bounds__(256) void transpose_ramp_kernel(const cpx<T>* __restrict__ src, int rows, int cols,
                                                             const cpx<double>* __restrict__ ramp,
                                                             cpx<T>* __restrict__ dst, int64_t ld_dst) {
    __shared__ cpx<T> tile[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int k = ty; k < 32; k += 8) {
        const int r = r0 + k, cc = c0 + tx;
        if (r < rows && cc < cols) tile[k][tx] = src[(int64_t)r * cols + cc];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int cc = c0 + k, r = r0 + tx;  // output line cc, sample r
        if (r < rows && cc < cols) {
            const cpx<double> s = ld64(tile[tx][k]), p = ramp[cc];
            dst[(int64_t)cc * ld_dst + r] = st<T>(p.x * s.x - p.y * s.y, p.x * s.y + p.y * s.x);
        }
    }
}

// ----------------------------------------------------- TOPS line filter
// After the unnormalised range IFFT: amplitude weight, RCMC phase, azimuth
// filter and FM rate (s1_decoder.cpp:651-675), per (line i, sample j).
template <typename T>
__global__ void tops_filter_kernel(cpx<T>* __restrict__ data, int64_t ld, int rows, int n,
                                   const double* __restrict__ vrel, const double* __restrict__ D,
                                   const double* __restrict__ R, const double* __restrict__ fr,
                                   const double* __restrict__ fdc, const double* __restrict__ tc,
                                   const double* __restrict__ faz, double ref_range, double* __restrict__ fm) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= n || i >= rows) return;
    const int64_t k = (int64_t)i * n + j;
    const double r = R[j], d = D[k], v = vrel[k];
    cpx<double> x = ld64(data[(int64_t)i * ld + j]);
    const double amp = sqrt(cube<T>(ref_range / r));
    x.x /= amp;
    x.y /= amp;
    // apply_range_cell_migration_correction (image_formation.cpp:458-476)
    const double shift = (r / d) - r;
    double s, c;
    phase_sincos<T>((((-4.0 * kPi) * fr[j]) * shift) / kC, &s, &c);
    cpx<double> y;
    y.x = x.x * c - x.y * s;
    y.y = x.x * s + x.y * c;
    // exp(4 i pi R D / lambda) exp(2 i pi (f_a + f_dc) t_corr)
    double s1, c1, s2, c2;
    phase_sincos<T>((((4.0 * kPi) * r) * d) / kWavelength, &s1, &c1);
    phase_sincos<T>(((2.0 * kPi) * (faz[i] + fdc[j])) * tc[j], &s2, &c2);
    const double hr = c1 * c2 - s1 * s2, hi = c1 * s2 + s1 * c2;
    data[(int64_t)i * ld + j] = st<T>(y.x * hr - y.y * hi, y.x * hi + y.y * hr);
    if constexpr (sizeof(T) == 8) fm[k] = -2 * pow(v, 2.0) * pow(d, 3.0) / (kWavelength * r);
    else fm[k] = -2 * (v * v) * (d * d * d) / (kWavelength * r);
}

// ---------------------------------------------------- azimuth time UFR
// Row r of [n][n_az] (sample r along azimuth), tiled to `shape`, times
// exp(i (p1 + p2)) with the FM rate ka linearly resampled from fmT[r],
// p1 = -pi kt t^2, p2 = 2 pi (kt + ka) dct t, kt = -ka rate / (rate - ka),
// dct = -f_dc / ka, t = dct + t1[a] (image_formation.cpp:303-327).
template <typename T>
__global__ void time_deramp_kernel(const cpx<T>* __restrict__ in, const double* __restrict__ fmT, int n_az, int n,
                                   int shape, const LinTab* __restrict__ lin, const double* __restrict__ t1,
                                   const double* __restrict__ fdc, double rate, cpx<T>* __restrict__ out) {
#pragma clang fp contract(off)
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (a >= shape || r >= n) return;
    const double dx_in = 1.0 / (n_az - 1);
    const double ka = linear_at(lin[a], a, fmT + (int64_t)r * n_az, n_az, dx_in);
    const double kt = -ka * rate / (rate - ka);
    const double dct = -fdc[r] / ka;
    const double t = dct + t1[a];
    const double p1 = ((-1.0 * kPi) * kt) * (t * t);
    const double p2 = (((2.0 * kPi) * (kt + ka)) * dct) * t;
    double s, c;
    phase_sincos<T>(p1 + p2, &s, &c);
    const cpx<double> x = ld64(in[(int64_t)r * n_az + (a % n_az)]);
    out[(int64_t)r * shape + a] = st<T>(c * x.x - s * x.y, c * x.y + s * x.x);
}

// Quadratic resampling of row r onto the cropped output lines and the
// re-ramp exp(i (p1 - p2)), p1 = pi kt t^2, t = dct + t2[j]
// (image_formation.cpp:350-380); out: [n][n_out] (row r, line j - crop).
template <typename T>
__global__ void time_resample_kernel(const cpx<T>* __restrict__ in, int shape, const double* __restrict__ fmT,
                                     int n_az, int n, int crop, int n_out, const QuadTab* __restrict__ quad,
                                     const LinTab* __restrict__ lin, const double* __restrict__ t2,
                                     const double* __restrict__ fdc, double rate, cpx<T>* __restrict__ out) {
#pragma clang fp contract(off)
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (k >= n_out || r >= n) return;
    const int j = crop + k;
    const QuadTab q = quad[k];
    const cpx<T>* row = in + (int64_t)r * shape;
    const cpx<double> y0 = ld64(row[q.i0]), y1 = ld64(row[q.i1]), y2 = ld64(row[q.i2]);
    const double vr = q.l0 * y0.x + q.l1 * y1.x + q.l2 * y2.x;
    const double vi = q.l0 * y0.y + q.l1 * y1.y + q.l2 * y2.y;
    const double dx_in = 1.0 / (n_az - 1);
    const double ka = linear_at(lin[k], j, fmT + (int64_t)r * n_az, n_az, dx_in);
    const double kt = -ka * rate / (rate - ka);
    const double dct = -fdc[r] / ka;
    const double t = dct + t2[k];
    const double p1 = ((1.0 * kPi) * kt) * (t * t);
    const double p2 = (((2.0 * kPi) * (kt + ka)) * dct) * t;
    double s, c;
    phase_sincos<T>(p1 - p2, &s, &c);
    out[(int64_t)r * n_out + k] = st<T>(c * vr - s * vi, c * vi + s * vr);
}

template <typename T>
int launch_tops_filter(hipStream_t s, void* data, int64_t ld, int rows, int n, const double* vrel, const double* D,
                       const double* R, const double* fr, const double* fdc, const double* tc, const double* faz,
                       double ref_range, double* fm) {
    dim3 grid((unsigned)((n + 255) / 256), (unsigned)rows);
    hipLaunchKernelGGL((tops_filter_kernel<T>), grid, dim3(256), 0, s, (cpx<T>*)data, ld, rows, n, vrel, D, R, fr,
                       fdc, tc, faz, ref_range, fm);
    return 0;
}

// Table uploads: host vectors into the context's table buffer, ordered on
// the stream and complete before the host vectors go away.
struct TablePack {
    std::vector<char> bytes;
    template <typename V>
    size_t add(const std::vector<V>& v) {
        const size_t off = (bytes.size() + 255) & ~(size_t)255;
        bytes.resize(off + sizeof(V) * v.size());
        if (!v.empty()) memcpy(bytes.data() + off, v.data(), sizeof(V) * v.size());
        return off;
    }
};
int upload(s1g_ctx* c, const TablePack& t, char** dev) {
    void* p = nullptr;
    int rc = ctx_aux(c, kTables, std::max<size_t>(t.bytes.size(), 256), &p);
    if (rc) return rc;
    hipStream_t s = ctx_stream(c);
    if (hipMemcpyAsync(p, t.bytes.data(), t.bytes.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return status(c, "table upload");
    *dev = (char*)p;
    return S1G_OK;
}

int check_common(s1g_ctx* c, int prec, int n_lines, int n_samples) {
    if (prec != S1G_F32 && prec != S1G_F64) return ctx_error(c, S1G_EARG, "precision must be S1G_F32 or S1G_F64");
    if (n_lines < 0 || n_samples < 0) return ctx_error(c, S1G_EARG, "bad shape");
    return S1G_OK;
}

template <typename T>
int freq_ufr_t(s1g_ctx* c, int L, int n, const void* in, int64_t ld_in, double nr, double inv_rate, double prf,
               double bw, void* out, int64_t ld_out) {
    const int prec = sizeof(T) == 4 ? S1G_F32 : S1G_F64;
    hipStream_t s = ctx_stream(c);
    const int shape = (int)floor(nr * L);
    const int mid = shape / 2;
    const int start = (int)(mid - (floor(bw) / 2)), end = (int)(mid + (floor(bw) / 2));
    if (shape < 1 || start < 0 || end > shape || end < start)
        return ctx_error(c, S1G_EARG, "azimuth frequency UFR: the low-pass window leaves the spectrum");
    const int whole = (int)floor(nr);
    const int rest = floor(nr) != nr ? (int)floor((nr - floor(nr)) * L) : 0;
    const int filled = (int)std::min<int64_t>((int64_t)whole * L + rest, shape);
    // re-ramp exp(i pi f^2 / rate), f = linspace(-B, B, shape); the deramp
    // is its conjugate (its phase is the exact negative)
    std::vector<double> f((size_t)shape);
    linspace(-(nr * prf / 2.0), nr * prf / 2.0, shape, f.data());
    std::vector<cpx<double>> ramp((size_t)shape);
    for (int a = 0; a < shape; ++a) {
        const double ph = (kPi * inv_rate) * (f[a] * f[a]);
        ramp[a].x = cos(ph);
        ramp[a].y = sin(ph);
    }
    TablePack tp;
    const size_t o_ramp = tp.add(ramp);
    char* tab;
    int rc = upload(c, tp, &tab);
    if (rc) return rc;
    const cpx<double>* d_ramp = (const cpx<double>*)(tab + o_ramp);
    void *A, *B, *C;
    const size_t es = sizeof(cpx<T>);
    if ((rc = ctx_aux(c, kBufA, es * (size_t)n * L, &A)) || (rc = ctx_aux(c, kBufB, es * (size_t)n * shape, &B)) ||
        (rc = ctx_aux(c, kBufC, es * (size_t)n * shape, &C)))
        return rc;
    // [L][n] -> [n][L], azimuth FFT per sample
    rc = launch_transpose(s, prec, in, ld_in, A, L, L, n);
    if (rc) return status(c, "azimuth frequency UFR");
    if ((rc = dft_rows_batched(c, prec, n, L, A, L, 0, true))) return rc;
    hipLaunchKernelGGL((freq_tile_kernel<T>), dim3((unsigned)((shape + 255) / 256), (unsigned)n), dim3(256), 0, s,
                       (const cpx<T>*)A, L, n, shape, filled, d_ramp, (cpx<T>*)B);
    if ((rc = dft_rows_batched(c, prec, n, shape, B, shape, 0, true))) return rc;
    hipLaunchKernelGGL((gate_kernel<T>), dim3((unsigned)((shape + 255) / 256), (unsigned)n), dim3(256), 0, s,
                       (const cpx<T>*)B, (cpx<T>*)C, n, shape, start, end - start);
    if ((rc = dft_rows_batched(c, prec, n, shape, C, shape, 1, true))) return rc;
    hipLaunchKernelGGL((transpose_ramp_kernel<T>), dim3((unsigned)((shape + 31) / 32), (unsigned)((n + 31) / 32)),
                       dim3(256), 0, s, (const cpx<T>*)C, n, shape, d_ramp, (cpx<T>*)out, ld_out);
    return status(c, "azimuth frequency UFR");
}

// Sample 0 of every line of fm, gathered into a contiguous array.
__global__ void gather_col0_kernel(const double* __restrict__ fm, int64_t ld, int n_az, double* __restrict__ out) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a < n_az) out[a] = fm[(int64_t)a * ld];
}

// The set-up's inputs on the host: sample 0 of fm on every line, f_dc at
// the middle sample, and the FM rate n_ref divides by.
int read_params_inputs(s1g_ctx* c, int n_az, int n, const double* dc, const double* fm, std::vector<double>& col0,
                       double* dc_mid, double* fm_ref) {
    hipStream_t s = ctx_stream(c);
    col0.assign((size_t)n_az, 0.0);
    // the FM rate n_ref divides by (the reference's transposed index where
    // it lies inside the array)
    int64_t ref = (n_az / 2 < n && n / 2 < n_az) ? (int64_t)(n / 2) * n + n_az / 2 : (int64_t)(n_az / 2) * n + n / 2;
    void* d_col = nullptr;
    int rc = ctx_aux(c, 9, sizeof(double) * ((size_t)n_az + 2), &d_col);
    if (rc) return rc;
    double* col = (double*)d_col;
    hipLaunchKernelGGL(gather_col0_kernel, dim3((unsigned)((n_az + 255) / 256)), dim3(256), 0, s, fm, (int64_t)n,
                       n_az, col);
    if ((rc = status(c, "azimuth time UFR set-up"))) return rc;
    if (hipMemcpyAsync(col + n_az, dc + n / 2, sizeof(double), hipMemcpyDeviceToDevice, s) != hipSuccess ||
        hipMemcpyAsync(col + n_az + 1, fm + ref, sizeof(double), hipMemcpyDeviceToDevice, s) != hipSuccess)
        return status(c, "azimuth time UFR set-up");
    std::vector<double> h((size_t)n_az + 2);
    if (hipMemcpyAsync(h.data(), col, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return status(c, "azimuth time UFR set-up");
    std::copy(h.begin(), h.begin() + n_az, col0.begin());
    *dc_mid = h[n_az];
    *fm_ref = h[n_az + 1];
    return S1G_OK;
}

int time_params(const std::vector<double>& col0, int n_az, double dc_mid, double fm_ref, double nr, double prf,
                double T, double bw, int swath, s1g_time_ufr_params* p) {
#pragma clang fp contract(off)
    const double bandwidth = (nr * prf) - bw;
    double sum = 0.0;
    for (int a = 0; a < n_az; ++a) sum += col0[a];
    const double ka_mean = fabs(sum) / n_az;
    double focused_time = T + ((bandwidth - (2 * bw)) / ka_mean);
    const double n_ref = dc_mid / fm_ref;
    focused_time *= swath != 11 ? 0.9 : 1.05;
    const int pos = (int)ceil((n_ref + focused_time / 2) / T);
    const int neg = (int)floor((n_ref - focused_time / 2) / T);
    const int ext = swath == 11 ? (int)floor(bw * 1.15) : (int)floor(bw * 0.26);
    p->bandwidth = bandwidth;
    p->focused_time = focused_time;
    p->n_ref = n_ref;
    p->ka_mean = ka_mean;
    p->num_tiles = pos - neg;
    p->shape = p->num_tiles * n_az;
    const int mid = p->shape / 2;
    p->lpf_start = (int)(mid - bandwidth / 2);
    p->lpf_end = (int)(mid + bandwidth / 2);
    p->output_shape = (int)floor(prf * T);
    p->downsample_shape = p->output_shape + ext;
    const int cmid = p->downsample_shape / 2;
    p->crop_start = cmid - (p->output_shape / 2);
    p->crop_end = cmid + (p->output_shape / 2);
    const bool ok = p->num_tiles >= 1 && (int64_t)p->num_tiles * n_az < (1 << 30) && p->lpf_start >= 0 &&
                    p->lpf_end <= p->shape && p->lpf_end >= p->lpf_start && p->shape >= 3 &&
                    p->downsample_shape >= 3 && p->crop_start >= 0 && p->crop_end <= p->downsample_shape;
    return ok ? S1G_OK : S1G_EARG;
}

template <typename T>
int time_ufr_t(s1g_ctx* c, int n_az, int n, const void* in, int64_t ld_in, const double* dc, const double* fm,
               double rate, const s1g_time_ufr_params& p, void* out, int64_t ld_out) {
    const int prec = sizeof(T) == 4 ? S1G_F32 : S1G_F64;
    hipStream_t s = ctx_stream(c);
    const int shape = p.shape, ds = p.downsample_shape, c0 = p.crop_start, n_out = p.crop_end - p.crop_start;
    // tables: linear-resampling walks (n_az -> shape, n_az -> ds), time
    // axes, quadratic weights of the kept outputs
    std::vector<LinTab> lin1 = linear_walk(n_az, shape), lin2all = linear_walk(n_az, ds);
    std::vector<LinTab> lin2(lin2all.begin() + c0, lin2all.begin() + c0 + n_out);
    std::vector<double> t1((size_t)shape), t2all((size_t)ds);
    linspace(-p.focused_time / 2, p.focused_time / 2, shape, t1.data());
    linspace(p.n_ref - p.focused_time / 2, p.n_ref + p.focused_time / 2, ds, t2all.data());
    std::vector<double> t2(t2all.begin() + c0, t2all.begin() + c0 + n_out);
    std::vector<QuadTab> quad((size_t)n_out);
    for (int k = 0; k < n_out; ++k) {
        if (shape == ds) {
            // quadratic_resample returns its input when the lengths agree
            quad[k] = QuadTab{c0 + k, c0 + k, c0 + k, 1.0, 0.0, 0.0};
        } else {
            quad[k] = quadratic_at(c0 + k, shape, ds);
        }
    }
    TablePack tp;
    const size_t o1 = tp.add(lin1), o2 = tp.add(lin2), o3 = tp.add(t1), o4 = tp.add(t2), o5 = tp.add(quad);
    char* tab;
    int rc = upload(c, tp, &tab);
    if (rc) return rc;
    void *A, *B, *C, *FT, *R;
    const size_t es = sizeof(cpx<T>);
    if ((rc = ctx_aux(c, kBufA, es * (size_t)n * n_az, &A)) || (rc = ctx_aux(c, kBufB, es * (size_t)n * shape, &B)) ||
        (rc = ctx_aux(c, kBufC, es * (size_t)n * shape, &C)) ||
        (rc = ctx_aux(c, kFmT, sizeof(double) * (size_t)n * n_az, &FT)) ||
        (rc = ctx_aux(c, kOut, es * (size_t)n * n_out, &R)))
        return rc;
    // [n_az][n] -> [n][n_az] for the data and the FM rate (8-byte elements
    // move as complex64)
    rc = launch_transpose(s, prec, in, ld_in, A, n_az, n_az, n);
    if (!rc) rc = launch_transpose(s, S1G_F32, fm, n, FT, n_az, n_az, n);
    if (rc) return status(c, "azimuth time UFR");
    hipLaunchKernelGGL((time_deramp_kernel<T>), dim3((unsigned)((shape + 255) / 256), (unsigned)n), dim3(256), 0, s,
                       (const cpx<T>*)A, (const double*)FT, n_az, n, shape, (const LinTab*)(tab + o1),
                       (const double*)(tab + o3), dc, rate, (cpx<T>*)B);
    if ((rc = dft_rows_batched(c, prec, n, shape, B, shape, 0, true))) return rc;
    hipLaunchKernelGGL((gate_kernel<T>), dim3((unsigned)((shape + 255) / 256), (unsigned)n), dim3(256), 0, s,
                       (const cpx<T>*)B, (cpx<T>*)C, n, shape, p.lpf_start, p.lpf_end - p.lpf_start);
    if ((rc = dft_rows_batched(c, prec, n, shape, C, shape, 1, true))) return rc;
    hipLaunchKernelGGL((time_resample_kernel<T>), dim3((unsigned)((n_out + 255) / 256), (unsigned)n), dim3(256), 0,
                       s, (const cpx<T>*)C, shape, (const double*)FT, n_az, n, c0, n_out,
                       (const QuadTab*)(tab + o5), (const LinTab*)(tab + o2), (const double*)(tab + o4), dc, rate,
                       (cpx<T>*)R);
    rc = launch_transpose(s, prec, R, n_out, out, ld_out, n, n_out);
    if (rc) return status(c, "azimuth time UFR");
    // the tables are read by kernels still queued: finish before the next
    // call may rewrite them
    if (hipStreamSynchronize(s) != hipSuccess) return status(c, "azimuth time UFR");
    return status(c, "azimuth time UFR");
}

}  // namespace

extern "C" {

double s1g_azimuth_beam_angle(int address) {
    if (address < 0 || address > 1023) return NAN;
    std::vector<double> a(1024);
    linspace(-0.018, 0.018, 1024, a.data());
    return a[address];
}

double s1g_doppler_centroid_rate(double pri_us, int first_beam_address, int last_beam_address, int n_packets,
                                 double velocity) {
#pragma clang fp contract(off)
    const double pri = pri_us * 1e-6;
    const double steer = (s1g_azimuth_beam_angle(last_beam_address) - s1g_azimuth_beam_angle(first_beam_address)) /
                         (pri * n_packets);
    return (-2 * velocity * steer) / kWavelength;
}

int s1g_doppler_centroid(s1g_ctx* c, int prec, int n_lines, int n_samples, const void* data, int64_t ld, double rate,
                         double prf, double t_near, double t_far, int n_blocks, double* dc) {
#pragma clang fp contract(off)
    if (!c) return S1G_EARG;
    int rc = check_common(c, prec, n_lines, n_samples);
    if (rc) return rc;
    if (n_lines < 2 || n_blocks < 2 || n_samples < n_blocks || ld < n_samples)
        return ctx_error(c, S1G_EARG, "Doppler centroid: need 2 lines and at least n_blocks >= 2 samples");
    if (!data || !dc) return ctx_error(c, S1G_EARG, "null pointer");
    hipStream_t s = ctx_stream(c);
    // get_deramping_signal (doppler.cpp:98-116) over the fixed 1.06 s
    std::vector<double> t((size_t)n_lines);
    linspace(0.0, 1.06, n_lines, t.data());
    std::vector<cpx<double>> dr((size_t)n_lines);
    for (int i = 0; i < n_lines; ++i) {
        const double d = t[i] - (0.5 * 1.06);
        const double ph = (kPi * rate) * (d * d);
        dr[i].x = cos(ph);
        dr[i].y = sin(ph);
    }
    TablePack tp;
    const size_t o_dr = tp.add(dr);
    const size_t o_sum = tp.add(std::vector<cpx<double>>((size_t)n_samples));
    char* tab;
    if ((rc = upload(c, tp, &tab))) return rc;
    cpx<double>* sums = (cpx<double>*)(tab + o_sum);
    const dim3 grid((unsigned)((n_samples + 255) / 256));
    if (prec == S1G_F32)
        hipLaunchKernelGGL(dc_lag_kernel<float>, grid, dim3(256), 0, s, (const cpx<float>*)data, ld, n_lines,
                           n_samples, (const cpx<double>*)(tab + o_dr), sums);
    else
        hipLaunchKernelGGL(dc_lag_kernel<double>, grid, dim3(256), 0, s, (const cpx<double>*)data, ld, n_lines,
                           n_samples, (const cpx<double>*)(tab + o_dr), sums);
    if ((rc = status(c, "Doppler centroid"))) return rc;
    std::vector<cpx<double>> h((size_t)n_samples);
    if (hipMemcpyAsync(h.data(), sums, sizeof(cpx<double>) * n_samples, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return status(c, "Doppler centroid");
    // wrapped estimates, block means, unwrapping, fit (doppler.cpp:28-212)
    std::vector<double> w((size_t)n_samples);
    for (int j = 0; j < n_samples; ++j) w[j] = -prf * atan2(h[j].y, h[j].x) / (2 * kPi);
    const int bs = n_samples / n_blocks;
    std::vector<double> fine((size_t)n_blocks), bt((size_t)n_blocks), tt((size_t)n_samples);
    for (int b = 0; b < n_blocks; ++b) {
        const int st0 = b * bs, en = st0 + bs < n_samples ? st0 + bs : n_samples - 1;
        double sum = 0.0;
        for (int k = st0; k < en; ++k) sum += w[k];
        fine[b] = sum / (double(en) - double(st0));
    }
    linspace(t_near, t_far, n_blocks, bt.data());
    const double b0 = bt[0];
    for (auto& v : bt) v -= b0;
    std::vector<double> un = unwrap_estimates(fine, bt, prf);
    double cf[3];
    polyfit2(bt, un, cf);
    linspace(t_near, t_far, n_samples, tt.data());
    const double t0 = tt[0];
    std::vector<double> out((size_t)n_samples);
    for (int j = 0; j < n_samples; ++j) {
        const double x = tt[j] - t0;
        out[j] = (cf[0] + cf[1] * x + cf[2] * x * x);
    }
    if (hipMemcpyAsync(dc, out.data(), sizeof(double) * n_samples, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return status(c, "Doppler centroid");
    return S1G_OK;
}

int s1g_tops_frequency_lines(int n_lines, double rate, double burst_duration, double prf) {
    const double nr = fabs(rate * burst_duration / prf);
    const double v = floor(nr * n_lines);
    return v >= 0 && v < 2147483647.0 ? (int)v : -1;
}

int s1g_azimuth_frequency_ufr(s1g_ctx* c, int prec, int n_lines, int n_samples, const void* in, int64_t ld_in,
                              double rate, double burst_duration, double prf, double doppler_bandwidth, void* out,
                              int64_t ld_out) {
    if (!c) return S1G_EARG;
    int rc = check_common(c, prec, n_lines, n_samples);
    if (rc) return rc;
    if (!n_lines || !n_samples) return S1G_OK;
    if (!in || !out || ld_in < n_samples || ld_out < n_samples) return ctx_error(c, S1G_EARG, "bad buffer");
    if (!(rate != 0.0) || !(prf > 0.0)) return ctx_error(c, S1G_EARG, "azimuth frequency UFR: rate and prf");
    const double nr = fabs(rate * burst_duration / prf);
    if (s1g_tops_frequency_lines(n_lines, rate, burst_duration, prf) < 1)
        return ctx_error(c, S1G_EARG, "azimuth frequency UFR: no output lines");
    return prec == S1G_F32
               ? freq_ufr_t<float>(c, n_lines, n_samples, in, ld_in, nr, 1 / rate, prf, doppler_bandwidth, out, ld_out)
               : freq_ufr_t<double>(c, n_lines, n_samples, in, ld_in, nr, 1 / rate, prf, doppler_bandwidth, out,
                                    ld_out);
}

int s1g_tops_filter_lines(s1g_ctx* c, int prec, int n_lines, int n_samples, void* data, int64_t ld,
                          const double* positions, const double* velocity_norms, const double* az_freqs,
                          const double* slant_ranges, const double* range_freqs, const double* f_dc,
                          const double* t_corr, double ref_range, double* fm) {
    if (!c) return S1G_EARG;
    int rc = check_common(c, prec, n_lines, n_samples);
    if (rc) return rc;
    if (!n_lines || !n_samples) return S1G_OK;
    if (!data || ld < n_samples || !positions || !velocity_norms || !az_freqs || !slant_ranges || !range_freqs ||
        !f_dc || !t_corr || !fm)
        return ctx_error(c, S1G_EARG, "bad buffer");
    void *V, *D;
    const size_t plane = sizeof(double) * (size_t)n_lines * n_samples;
    if ((rc = ctx_aux(c, kVrel, plane, &V)) || (rc = ctx_aux(c, kFactors, plane, &D))) return rc;
    // get_effective_velocities per line, the range FFT, secondary range
    // compression (factors D), the unnormalised range IFFT
    if ((rc = s1g_effective_velocities(c, positions, velocity_norms, n_lines, slant_ranges, n_samples, (double*)V,
                                       n_samples)))
        return rc;
    if ((rc = dft_rows_batched(c, prec, n_lines, n_samples, data, ld, 0, false))) return rc;
    if ((rc = s1g_secondary_range_compression(c, prec, n_lines, n_samples, data, ld, (const double*)V, slant_ranges,
                                              range_freqs, f_dc, az_freqs, (double*)D)))
        return rc;
    if ((rc = dft_rows_batched(c, prec, n_lines, n_samples, data, ld, 1, false))) return rc;
    hipStream_t s = ctx_stream(c);
    if (prec == S1G_F32)
        launch_tops_filter<float>(s, data, ld, n_lines, n_samples, (const double*)V, (const double*)D, slant_ranges,
                                  range_freqs, f_dc, t_corr, az_freqs, ref_range, fm);
    else
        launch_tops_filter<double>(s, data, ld, n_lines, n_samples, (const double*)V, (const double*)D, slant_ranges,
                                   range_freqs, f_dc, t_corr, az_freqs, ref_range, fm);
    return status(c, "TOPS line filter");
}

int s1g_azimuth_time_ufr_params(s1g_ctx* c, int n_lines, int n_samples, const double* f_dc, const double* fm,
                                double rate, double burst_duration, double prf, double doppler_bandwidth,
                                int swath_number, s1g_time_ufr_params* out) {
    if (!c || !out) return S1G_EARG;
    if (n_lines < 3 || n_samples < 1 || !f_dc || !fm) return ctx_error(c, S1G_EARG, "bad shape or buffer");
    std::vector<double> col0;
    double dc_mid = 0.0, fm_ref = 0.0;
    int rc = read_params_inputs(c, n_lines, n_samples, f_dc, fm, col0, &dc_mid, &fm_ref);
    if (rc) return rc;
    const double nr = fabs(rate * burst_duration / prf);
    rc = time_params(col0, n_lines, dc_mid, fm_ref, nr, prf, burst_duration, doppler_bandwidth, swath_number, out);
    if (rc) {
        char buf[400];
        snprintf(buf, sizeof buf,
                 "azimuth time UFR: the low-pass window or the crop leaves the arrays (tiles %d, shape %d, "
                 "window [%d, %d), crop [%d, %d) of %d; focused time %.9g s, n_ref %.9g s, mean FM rate %.9g, "
                 "f_dc %.9g, FM rate at the reference point %.9g)",
                 out->num_tiles, out->shape, out->lpf_start, out->lpf_end, out->crop_start, out->crop_end,
                 out->downsample_shape, out->focused_time, out->n_ref, out->ka_mean, dc_mid, fm_ref);
        return ctx_error(c, S1G_EARG, buf);
    }
    return S1G_OK;
}

int s1g_azimuth_time_ufr(s1g_ctx* c, int prec, int n_lines, int n_samples, const void* in, int64_t ld_in,
                         const double* f_dc, const double* fm, double rate, double burst_duration, double prf,
                         double doppler_bandwidth, int swath_number, void* out, int64_t ld_out) {
    if (!c) return S1G_EARG;
    int rc = check_common(c, prec, n_lines, n_samples);
    if (rc) return rc;
    if (!in || !out || ld_in < n_samples || ld_out < n_samples) return ctx_error(c, S1G_EARG, "bad buffer");
    s1g_time_ufr_params p;
    if ((rc = s1g_azimuth_time_ufr_params(c, n_lines, n_samples, f_dc, fm, rate, burst_duration, prf,
                                          doppler_bandwidth, swath_number, &p)))
        return rc;
    return prec == S1G_F32 ? time_ufr_t<float>(c, n_lines, n_samples, in, ld_in, f_dc, fm, rate, p, out, ld_out)
                           : time_ufr_t<double>(c, n_lines, n_samples, in, ld_in, f_dc, fm, rate, p, out, ld_out);
}

}  // extern "C"
