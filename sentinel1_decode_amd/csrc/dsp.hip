// dsp.hip -- kernels behind the general signal-processing entry points of
// the C ABI: batched DFTs of any length (Stockham passes, or Bluestein's
// chirp-z form when a prime factor exceeds the pass radices), transposes,
// elementwise products, the reference function, the replica chirp and
// windowed-sinc interpolation.
//
// Reference behaviour: compute_1d_dft / compute_axis_dft
// (signal_processing.cpp:516-681: FFTW_FORWARD unnormalised, inverse x 1/n),
// transpose (:463-498), get_reference_function / pulse_compression
// (image_formation.cpp:4-43), L0Packet::get_replica_chirp
// (packet.cpp:345-373), sinc_interpolate (signal_processing.cpp:228-257).
#include "s1g_internal.h"
#include "s1g_lds_fft.h"

namespace s1g {

// ------------------------------------------------------------------ copies
template <typename T>
__global__ void copy_scale_kernel(const cpx<T>* __restrict__ src, int64_t lds, cpx<T>* __restrict__ dst, int64_t ldd,
                                  int rows, int n, T scale) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)rows * n) return;
    const int r = (int)(i / n), k = (int)(i - (int64_t)r * n);
    dst[r * ldd + k] = scl(src[r * lds + k], scale);
}

int launch_copy_scale(hipStream_t s, int prec, const void* src, int64_t lds, void* dst, int64_t ldd, int rows, int n,
                      double scale) {
    const int64_t total = (int64_t)rows * n;
    if (total == 0) return 0;
    dim3 b(256), g((unsigned)((total + 255) / 256));
    if (prec == S1G_F32)
        hipLaunchKernelGGL(copy_scale_kernel<float>, g, b, 0, s, (const cpx<float>*)src, lds, (cpx<float>*)dst, ldd,
                           rows, n, (float)scale);
    else
        hipLaunchKernelGGL(copy_scale_kernel<double>, g, b, 0, s, (const cpx<double>*)src, lds, (cpx<double>*)dst,
                           ldd, rows, n, scale);
    return (int)hipGetLastError();
}

// ------------------------------------------------------- LDS row transforms
// One workgroup per row, in place: the row is loaded into LDS at its
// digit-reversed positions, the plan's DIT passes run there with the
// two-level twiddle tables in LDS, and the store writes the natural-order
// spectrum with the scale folded in.  One read and one write of the row,
// against one per pass plus a copy for the global Stockham passes.
template <typename T, bool INV, int NT, int MAXR, int MINW>
__global__ __launch_bounds__(NT, MINW) void fft_rows_lds_kernel(cpx<T>* __restrict__ data, int64_t ld, int rows,
                                                                FftPlan plan, const cpx<T>* __restrict__ tw,
                                                                const int32_t* __restrict__ perm, T scale) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int n = plan.n, tid = threadIdx.x;
    cpx<T>* buf = (cpx<T>*)smem_raw;
    cpx<T>* lo = buf + n;
    cpx<T>* hi = lo + kTwLo;
    const int row = blockIdx.x;
    if (row >= rows) return;
    cpx<T>* x = data + (int64_t)row * ld;
    load_tw2(tw, n, lo, hi, tid, NT);
    // loads in batches of kRowBatch per thread, all issued before the
    // batch's LDS stores (one memory wait per batch, not per element)
    constexpr int kRowBatch = 8;
    for (int q0 = 0; q0 < n; q0 += NT * kRowBatch) {
        cpx<T> v[kRowBatch];
        int pos[kRowBatch];
#pragma unroll
        for (int u = 0; u < kRowBatch; ++u) {
            const int q = q0 + tid + NT * u;
            if (q < n) {
                v[u] = x[q];
                pos[u] = perm[q];
            }
        }
#pragma unroll
        for (int u = 0; u < kRowBatch; ++u)
            if (q0 + tid + NT * u < n) buf[pos[u]] = v[u];
    }
    __syncthreads();
    fft_dit_lds_tw2<T, INV, MAXR>(buf, plan, lo, hi, tid, NT);
    for (int q = tid; q < n; q += NT) x[q] = scl(buf[q], scale);
}

static size_t rows_lds_bytes(int prec, int n) {
    return (size_t)(n + kTwLo + tw_hi_count(n)) * (prec == S1G_F32 ? sizeof(cpx<float>) : sizeof(cpx<double>));
}

bool fft_rows_lds_fits(int prec, int n) { return n > 0 && rows_lds_bytes(prec, n) <= kRowLdsBytes; }

// Minimum waves per SIMD asked of the compiler per radix class: the
// register budget of a class's largest butterfly (FP64 needs twice the
// registers of FP32).  Rows above half the LDS leave one workgroup per CU,
// which then runs 512 threads (two waves per SIMD: up to 256 registers).
template <typename T, int CLS, int NT>
constexpr int rows_min_waves() {
    if constexpr (NT == 512) return 2;
    else if constexpr (sizeof(T) == 4) return CLS <= 1 ? 4 : 2;
    else return CLS == 0 ? 4 : CLS == 1 ? 2 : 1;
}

template <typename T, int CLS, int NT>
static void launch_rows_cls(hipStream_t s, const DevFft& f, void* data, int64_t ld, int rows, int inverse,
                            double scale, size_t lds) {
    constexpr int MAXR = kRadixClassMax[CLS], MINW = rows_min_waves<T, CLS, NT>();
    const dim3 g((unsigned)rows);
    if (inverse)
        hipLaunchKernelGGL((fft_rows_lds_kernel<T, true, NT, MAXR, MINW>), g, dim3(NT), lds, s, (cpx<T>*)data, ld,
                           rows, f.plan, tw_of<T>(f), f.perm, (T)scale);
    else
        hipLaunchKernelGGL((fft_rows_lds_kernel<T, false, NT, MAXR, MINW>), g, dim3(NT), lds, s, (cpx<T>*)data, ld,
                           rows, f.plan, tw_of<T>(f), f.perm, (T)scale);
}

template <typename T, int NT>
static void launch_rows_nt(hipStream_t s, const DevFft& f, void* data, int64_t ld, int rows, int inverse,
                           double scale, size_t lds) {
    switch (radix_class(f.plan)) {
        case 0: launch_rows_cls<T, 0, NT>(s, f, data, ld, rows, inverse, scale, lds); break;
        case 1: launch_rows_cls<T, 1, NT>(s, f, data, ld, rows, inverse, scale, lds); break;
        case 2: launch_rows_cls<T, 2, NT>(s, f, data, ld, rows, inverse, scale, lds); break;
        default: launch_rows_cls<T, 3, NT>(s, f, data, ld, rows, inverse, scale, lds); break;
    }
}

template <typename T>
static void launch_rows_t(hipStream_t s, const DevFft& f, void* data, int64_t ld, int rows, int inverse,
                          double scale, size_t lds) {
    if (lds > kRowLdsBytes / 2) launch_rows_nt<T, 512>(s, f, data, ld, rows, inverse, scale, lds);
    else launch_rows_nt<T, 256>(s, f, data, ld, rows, inverse, scale, lds);
}

int launch_fft_rows_lds(hipStream_t s, int prec, const DevFft& f, void* data, int64_t ld, int rows, int inverse,
                        double scale) {
    const int n = f.plan.n;
    if (rows <= 0) return 0;
    if (f.plan.npass < 0 || !fft_rows_lds_fits(prec, n)) return -1;
    const size_t lds = rows_lds_bytes(prec, n);
    if (prec == S1G_F32) launch_rows_t<float>(s, f, data, ld, rows, inverse, scale, lds);
    else launch_rows_t<double>(s, f, data, ld, rows, inverse, scale, lds);
    return (int)hipGetLastError();
}

// --------------------------------------------------------------- transpose
// dst[c][r] = src[r][c], 32 x 32 tiles through LDS (one pad column).
template <typename T>
__global__ __launch_bounds__(256) void transpose_kernel(const cpx<T>* __restrict__ src, int64_t lds,
                                                        cpx<T>* __restrict__ dst, int64_t ldd, int rows, int cols) {
    __shared__ cpx<T> tile[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int k = ty; k < 32; k += 8) {
        const int r = r0 + k, c = c0 + tx;
        if (r < rows && c < cols) tile[k][tx] = src[(int64_t)r * lds + c];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int c = c0 + k, r = r0 + tx;
        if (r < rows && c < cols) dst[(int64_t)c * ldd + r] = tile[tx][k];
    }
}

int launch_transpose(hipStream_t s, int prec, const void* src, int64_t lds, void* dst, int64_t ldd, int rows,
                     int cols) {
    if (rows <= 0 || cols <= 0) return 0;
    dim3 b(256), g((unsigned)((cols + 31) / 32), (unsigned)((rows + 31) / 32));
    if (prec == S1G_F32)
        hipLaunchKernelGGL(transpose_kernel<float>, g, b, 0, s, (const cpx<float>*)src, lds, (cpx<float>*)dst, ldd,
                           rows, cols);
    else
        hipLaunchKernelGGL(transpose_kernel<double>, g, b, 0, s, (const cpx<double>*)src, lds, (cpx<double>*)dst,
                           ldd, rows, cols);
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------- Bluestein
// X_k = w_k sum_j (x_j w_j) conj(w_{k-j}),  w_m = e^{-i pi m^2 / n} (forward;
// conjugated for the inverse).  m^2 is reduced mod 2n in integers, so the
// chirp angle stays below 2 pi and sincospi keeps full precision.
__device__ __forceinline__ void chirp_w(int64_t m, int n, bool inverse, double* c, double* sn) {
    const int64_t r = (m * m) % (2 * (int64_t)n);
    double sv, cv;
    sincospi((double)r / (double)n, &sv, &cv);
    *c = cv;
    *sn = inverse ? sv : -sv;
}

// a[r][j] = x[r][j] w_j (j < n), 0 up to M
template <typename T>
__global__ void bluestein_pre_kernel(const cpx<T>* __restrict__ x, int64_t ldx, cpx<T>* __restrict__ a, int rows,
                                     int n, int M, int inverse) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)rows * M) return;
    const int r = (int)(i / M), j = (int)(i - (int64_t)r * M);
    cpx<T> v = mk((T)0, (T)0);
    if (j < n) {
        double c, sn;
        chirp_w(j, n, inverse != 0, &c, &sn);
        v = cmul(x[r * ldx + j], mk((T)c, (T)sn));
    }
    a[i] = v;
}

// b[m] = conj(w_m) for |m| < n (index m, and M - m for negative m), 0 elsewhere
__global__ void bluestein_b_kernel(cpx<double>* __restrict__ b, int n, int M, int inverse) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    int m = -1;
    if (i < n) m = i;
    else if (i > M - n) m = M - i;
    cpx<double> v = mk(0.0, 0.0);
    if (m >= 0) {
        double c, sn;
        chirp_w(m, n, inverse != 0, &c, &sn);
        v = mk(c, -sn);
    }
    b[i] = v;
}

// Bhat (FP64) / M to the working precision
template <typename T>
__global__ void bluestein_bhat_kernel(const cpx<double>* __restrict__ B, cpx<T>* __restrict__ out, int M) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const double inv = 1.0 / (double)M;
    out[i] = mk((T)(B[i].x * inv), (T)(B[i].y * inv));
}

// a[r][k] *= bhat[k]
template <typename T>
__global__ void mul_rows_kernel(cpx<T>* __restrict__ a, int64_t lda, const cpx<T>* __restrict__ v, int rows, int n,
                                int conj_v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)rows * n) return;
    const int r = (int)(i / n), k = (int)(i - (int64_t)r * n);
    cpx<T>* p = a + r * lda + k;
    *p = conj_v ? cmulc(*p, v[k]) : cmul(*p, v[k]);
}

// x[r][k] = w_k c[r][k] * scale (k < n)
template <typename T>
__global__ void bluestein_post_kernel(const cpx<T>* __restrict__ cbuf, int M, cpx<T>* __restrict__ x, int64_t ldx,
                                      int rows, int n, int inverse, T scale) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)rows * n) return;
    const int r = (int)(i / n), k = (int)(i - (int64_t)r * n);
    double c, sn;
    chirp_w(k, n, inverse != 0, &c, &sn);
    x[r * ldx + k] = scl(cmul(cbuf[(int64_t)r * M + k], mk((T)c, (T)sn)), scale);
}

// Bluestein filter of length M for an n-point transform, FP64 spectrum in
// `spec` (scratch: 3 M complex doubles), result (scaled by 1/M) in bhat.
int launch_bluestein_filter(hipStream_t s, int prec, const DevFft& fM, int n, int inverse, void* scratch, void* bhat) {
    const int M = fM.plan.n;
    cpx<double>* b = (cpx<double>*)scratch;
    cpx<double>* B = b + M;
    cpx<double>* t = B + M;
    dim3 blk(256), g((M + 255) / 256);
    hipLaunchKernelGGL(bluestein_b_kernel, g, blk, 0, s, b, n, M, inverse);
    int rc = launch_fft_global(s, S1G_F64, fM, b, B, t, 1, M, 0);
    if (rc) return rc;
    if (prec == S1G_F32)
        hipLaunchKernelGGL(bluestein_bhat_kernel<float>, g, blk, 0, s, B, (cpx<float>*)bhat, M);
    else
        hipLaunchKernelGGL(bluestein_bhat_kernel<double>, g, blk, 0, s, B, (cpx<double>*)bhat, M);
    return (int)hipGetLastError();
}

// x (rows x n, leading dimension ldx) <- DFT_n along rows, via three
// M-point transforms.  w1, w2: rows * M scratch each.
int launch_bluestein(hipStream_t s, int prec, const DevFft& fM, const void* bhat, void* x, int64_t ldx, int rows,
                     int n, int inverse, double scale, void* w1, void* w2, void* w3) {
    const int M = fM.plan.n;
    const int64_t totM = (int64_t)rows * M, totN = (int64_t)rows * n;
    dim3 blk(256), gM((unsigned)((totM + 255) / 256)), gN((unsigned)((totN + 255) / 256));
    int rc;
    if (prec == S1G_F32) {
        hipLaunchKernelGGL(bluestein_pre_kernel<float>, gM, blk, 0, s, (const cpx<float>*)x, ldx, (cpx<float>*)w1,
                           rows, n, M, inverse);
        rc = launch_fft_global(s, prec, fM, w1, w2, w3, rows, M, 0);
        if (rc) return rc;
        hipLaunchKernelGGL(mul_rows_kernel<float>, gM, blk, 0, s, (cpx<float>*)w2, (int64_t)M,
                           (const cpx<float>*)bhat, rows, M, 0);
        rc = launch_fft_global(s, prec, fM, w2, w1, w3, rows, M, 1);
        if (rc) return rc;
        hipLaunchKernelGGL(bluestein_post_kernel<float>, gN, blk, 0, s, (const cpx<float>*)w1, M, (cpx<float>*)x,
                           ldx, rows, n, inverse, (float)scale);
    } else {
        hipLaunchKernelGGL(bluestein_pre_kernel<double>, gM, blk, 0, s, (const cpx<double>*)x, ldx,
                           (cpx<double>*)w1, rows, n, M, inverse);
        rc = launch_fft_global(s, prec, fM, w1, w2, w3, rows, M, 0);
        if (rc) return rc;
        hipLaunchKernelGGL(mul_rows_kernel<double>, gM, blk, 0, s, (cpx<double>*)w2, (int64_t)M,
                           (const cpx<double>*)bhat, rows, M, 0);
        rc = launch_fft_global(s, prec, fM, w2, w1, w3, rows, M, 1);
        if (rc) return rc;
        hipLaunchKernelGGL(bluestein_post_kernel<double>, gN, blk, 0, s, (const cpx<double>*)w1, M,
                           (cpx<double>*)x, ldx, rows, n, inverse, scale);
    }
    return (int)hipGetLastError();
}

int launch_mul_rows(hipStream_t s, int prec, void* a, int64_t lda, const void* v, int rows, int n, int conj_v) {
    const int64_t total = (int64_t)rows * n;
    if (total == 0) return 0;
    dim3 b(256), g((unsigned)((total + 255) / 256));
    if (prec == S1G_F32)
        hipLaunchKernelGGL(mul_rows_kernel<float>, g, b, 0, s, (cpx<float>*)a, lda, (const cpx<float>*)v, rows, n,
                           conj_v);
    else
        hipLaunchKernelGGL(mul_rows_kernel<double>, g, b, 0, s, (cpx<double>*)a, lda, (const cpx<double>*)v, rows,
                           n, conj_v);
    return (int)hipGetLastError();
}

// ------------------------------------------------------- reference function
// energy = sum |c|^4 / n as the reference accumulates it: norm = |c|^2 per
// sample (magnitude_1d), norm *= norm / n, summed (image_formation.cpp:6-14).
template <typename T>
__global__ __launch_bounds__(256) void chirp_energy_kernel(const cpx<T>* __restrict__ c, int n,
                                                           double* __restrict__ energy) {
    __shared__ double red[256];
    double acc = 0;
    for (int i = threadIdx.x; i < n; i += 256) {
        const double p = (double)c[i].x * (double)c[i].x + (double)c[i].y * (double)c[i].y;
        acc += p * (p / (double)n);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *energy = red[0];
}

// ref = conj(ref) / energy
template <typename T>
__global__ void conj_div_kernel(cpx<T>* __restrict__ ref, int n, const double* __restrict__ energy) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double e = *energy;
    const cpx<T> v = ref[i];
    ref[i] = mk((T)((double)v.x / e), (T)(-(double)v.y / e));
}

int launch_chirp_energy(hipStream_t s, int prec, const void* chirp, int n, double* energy) {
    if (prec == S1G_F32)
        hipLaunchKernelGGL(chirp_energy_kernel<float>, dim3(1), dim3(256), 0, s, (const cpx<float>*)chirp, n, energy);
    else
        hipLaunchKernelGGL(chirp_energy_kernel<double>, dim3(1), dim3(256), 0, s, (const cpx<double>*)chirp, n,
                           energy);
    return (int)hipGetLastError();
}

int launch_conj_div(hipStream_t s, int prec, void* ref, int n, const double* energy) {
    dim3 b(256), g((n + 255) / 256);
    if (prec == S1G_F32)
        hipLaunchKernelGGL(conj_div_kernel<float>, g, b, 0, s, (cpx<float>*)ref, n, energy);
    else
        hipLaunchKernelGGL(conj_div_kernel<double>, g, b, 0, s, (cpx<double>*)ref, n, energy);
    return (int)hipGetLastError();
}

// ------------------------------------------------------------ replica chirp
// chirp[i] = (1 / nc) exp(i 2 PI (txpsf t + txprr/2 t^2)) for i in
// [m, m + nc), m = floor((len - nc) / 2), t = linspace(0, txpl, nc)[i - m];
// zero elsewhere.  FP64 phase, the reference's PI and expression order.
template <typename T>
__global__ void replica_chirp_kernel(cpx<T>* __restrict__ out, int len, int nc, double phi1, double phi2,
                                     double txpl) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    const int m = (len - nc) / 2;
    cpx<T> v = mk((T)0, (T)0);
    if (i >= m && i < m + nc) {
        const double PI = 3.14159265358979;
        const double step = (txpl - 0.0) / (nc - 1);
        const double t = 0.0 + (i - m) * step;
        const double ph = (2.0 * PI) * ((phi1 * t) + phi2 * (t * t));
        double sn, cs;
        sincos(ph, &sn, &cs);
        const double a = 1.0 / nc;
        v = mk((T)(a * cs), (T)(a * sn));
    }
    out[i] = v;
}

int launch_replica_chirp(hipStream_t s, int prec, void* out, int len, int nc, double txpsf, double txprr, double txpl) {
    dim3 b(256), g((len + 255) / 256);
    if (prec == S1G_F32)
        hipLaunchKernelGGL(replica_chirp_kernel<float>, g, b, 0, s, (cpx<float>*)out, len, nc, txpsf, txprr * 0.5,
                           txpl);
    else
        hipLaunchKernelGGL(replica_chirp_kernel<double>, g, b, 0, s, (cpx<double>*)out, len, nc, txpsf, txprr * 0.5,
                           txpl);
    return (int)hipGetLastError();
}

// ------------------------------------------------------- sinc interpolation
// out[q] = sum_{k=-L..L, 0 <= n+k < len} x[n+k] sinc(t-n-k) 0.5 (1 + cos(PI (t-n-k)/L)),
// n = floor(t_q) (signal_processing.cpp:228-257, FP64, the reference's PI).
template <typename T>
__global__ void sinc_interp_kernel(const cpx<T>* __restrict__ x, int len, const double* __restrict__ t, int m,
                                   int L, cpx<T>* __restrict__ out) {
#pragma clang fp contract(off)
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m) return;
    const double PI = 3.14159265358979;
    const double tq = t[q];
    const int n = (int)floor(tq);
    double re = 0, im = 0;
    for (int k = -L; k <= L; ++k) {
        const int idx = n + k;
        if (idx < 0 || idx >= len) continue;
        const double xx = tq - idx;
        const double w = 0.5 * (1 + cos(PI * xx / L));
        const double sc = xx == 0.0 ? 1.0 : sin(PI * xx) / (PI * xx);
        re += ((double)x[idx].x * sc) * w;  // (signal[idx] * sinc(x)) * w
        im += ((double)x[idx].y * sc) * w;
    }
    out[q] = mk((T)re, (T)im);
}

int launch_sinc_interp(hipStream_t s, int prec, const void* x, int len, const double* t, int m, int L, void* out) {
    if (m <= 0) return 0;
    dim3 b(256), g((m + 255) / 256);
    if (prec == S1G_F32)
        hipLaunchKernelGGL(sinc_interp_kernel<float>, g, b, 0, s, (const cpx<float>*)x, len, t, m, L,
                           (cpx<float>*)out);
    else
        hipLaunchKernelGGL(sinc_interp_kernel<double>, g, b, 0, s, (const cpx<double>*)x, len, t, m, L,
                           (cpx<double>*)out);
    return (int)hipGetLastError();
}

}  // namespace s1g
