// output.hip -- the image's way out: display scaling (norm, log, magnitude,
// real / imaginary part) on the GPU, complex128 -> complex64 for the cf32
// dumps, and a BigTIFF writer for the scaled float image.
//
// The scaling is two passes over the image: a max |x| reduction (norm
// modes), then one elementwise pass; both are HBM-bound (read c bytes, write
// 4 bytes per sample).
#include <math.h>
#include <stdio.h>

#include <vector>

#include "s1g_internal.h"

using namespace s1g;

namespace {

template <typename T>
__device__ __forceinline__ double absd(const cpx<T>& x) {
    return hypot((double)x.x, (double)x.y);
}

__global__ void zero_u64_kernel(unsigned long long* p) { *p = 0ull; }

// max |x| over the image: block reduction, then an atomic max on the bits of
// the (non-negative) double, which order like the values.
template <typename T>
__global__ __launch_bounds__(256) void max_abs_kernel(const cpx<T>* __restrict__ x, int64_t ld, int rows, int cols,
                                                      unsigned long long* __restrict__ result) {
    __shared__ double s[256];
    double m = 0.0;
    const int64_t total = (int64_t)rows * cols;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = k / cols, c = k - r * cols;
        m = fmax(m, absd(x[r * ld + c]));
    }
    s[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(result, (unsigned long long)__double_as_longlong(s[0]));
}

// norm_2d (signal_processing.cpp:367-401): val = |x| + 1e-12, then val / max
// or 20 log10(val / max), in double, stored as float; magnitude_2d
// (:440-460): pow(|x|, 2); real / imag parts.
template <typename T>
__global__ void scale_kernel(const cpx<T>* __restrict__ x, int64_t ld, int rows, int cols, int mode,
                             const unsigned long long* __restrict__ maxbits, float* __restrict__ out,
                             int64_t ld_out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (c >= cols || r >= rows) return;
    const cpx<T> v = x[(int64_t)r * ld + c];
    float o;
    if (mode == S1G_SCALE_REAL) {
        o = (float)v.x;
    } else if (mode == S1G_SCALE_IMAG) {
        o = (float)v.y;
    } else if (mode == S1G_SCALE_MAG) {
        const double a = absd(v);
        o = (float)(a * a);
    } else {
        const double mx = __longlong_as_double((long long)*maxbits);
        const double val = absd(v) + 1e-12;
        o = mode == S1G_SCALE_NORM_LOG ? (float)(20.0 * log10(val / mx)) : (float)(val / mx);
    }
    out[(int64_t)r * ld_out + c] = o;
}

__global__ void to_cf32_kernel(const cpx<double>* __restrict__ x, int64_t ld, int rows, int cols,
                               cpx<float>* __restrict__ out, int64_t ld_out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (c >= cols || r >= rows) return;
    const cpx<double> v = x[(int64_t)r * ld + c];
    cpx<float> o;
    o.x = (float)v.x;
    o.y = (float)v.y;
    out[(int64_t)r * ld_out + c] = o;
}

int status(s1g_ctx* c, const char* what) {
    const hipError_t e = hipGetLastError();
    if (e == hipSuccess) return S1G_OK;
    char buf[160];
    snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    return ctx_error(c, S1G_EHIP, buf);
}

template <typename T>
int scale_t(s1g_ctx* c, int mode, int rows, int cols, const void* data, int64_t ld, float* out, int64_t ld_out) {
    hipStream_t s = ctx_stream(c);
    unsigned long long* mx = nullptr;
    if (mode == S1G_SCALE_NORM || mode == S1G_SCALE_NORM_LOG) {
        void* p = nullptr;
        int rc = ctx_scratch(c, sizeof(unsigned long long), &p);
        if (rc) return rc;
        mx = (unsigned long long*)p;
        // a fill kernel (not a memset node) so the call may be captured
        hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(1), 0, s, mx);
        const int64_t total = (int64_t)rows * cols;
        const int blocks = (int)std::min<int64_t>((total + 255) / 256, 4096);
        hipLaunchKernelGGL((max_abs_kernel<T>), dim3(blocks), dim3(256), 0, s, (const cpx<T>*)data, ld, rows, cols,
                           mx);
    }
    dim3 grid((unsigned)((cols + 255) / 256), (unsigned)rows);
    hipLaunchKernelGGL((scale_kernel<T>), grid, dim3(256), 0, s, (const cpx<T>*)data, ld, rows, cols, mode, mx, out,
                       ld_out);
    return status(c, "scale");
}

// Little-endian BigTIFF: 16-byte header, the image data, then one IFD of
// 20-byte entries (tag, type, count, value / offset).
void put16(std::vector<uint8_t>& b, uint16_t v) {
    for (int i = 0; i < 2; ++i) b.push_back((uint8_t)(v >> (8 * i)));
}
void put64(std::vector<uint8_t>& b, uint64_t v) {
    for (int i = 0; i < 8; ++i) b.push_back((uint8_t)(v >> (8 * i)));
}
void entry(std::vector<uint8_t>& b, uint16_t tag, uint16_t type, uint64_t count, uint64_t value) {
    put16(b, tag);
    put16(b, type);
    put64(b, count);
    put64(b, value);  // short / long / long8 values sit left-justified in the 8 bytes
}

}  // namespace

extern "C" {

int s1g_scale(s1g_ctx* c, int prec, int mode, int n_rows, int n_cols, const void* data, int64_t ld, float* out,
              int64_t ld_out) {
    if (!c) return S1G_EARG;
    if (prec != S1G_F32 && prec != S1G_F64) return ctx_error(c, S1G_EARG, "precision must be S1G_F32 or S1G_F64");
    if (mode < S1G_SCALE_NORM_LOG || mode > S1G_SCALE_IMAG) return ctx_error(c, S1G_EARG, "unknown scaling mode");
    if (n_rows < 0 || n_cols < 0 || ld < n_cols || ld_out < n_cols) return ctx_error(c, S1G_EARG, "bad shape");
    if (!n_rows || !n_cols) return S1G_OK;
    if (!data || !out) return ctx_error(c, S1G_EARG, "null pointer");
    return prec == S1G_F32 ? scale_t<float>(c, mode, n_rows, n_cols, data, ld, out, ld_out)
                           : scale_t<double>(c, mode, n_rows, n_cols, data, ld, out, ld_out);
}

int s1g_to_cf32(s1g_ctx* c, int n_rows, int n_cols, const void* data, int64_t ld, void* out, int64_t ld_out) {
    if (!c) return S1G_EARG;
    if (n_rows < 0 || n_cols < 0 || ld < n_cols || ld_out < n_cols) return ctx_error(c, S1G_EARG, "bad shape");
    if (!n_rows || !n_cols) return S1G_OK;
    if (!data || !out) return ctx_error(c, S1G_EARG, "null pointer");
    dim3 grid((unsigned)((n_cols + 255) / 256), (unsigned)n_rows);
    hipLaunchKernelGGL(to_cf32_kernel, grid, dim3(256), 0, ctx_stream(c), (const cpx<double>*)data, ld, n_rows, n_cols,
                       (cpx<float>*)out, ld_out);
    return status(c, "to_cf32");
}

int s1g_write_tiff(const char* path, const float* img, int n_rows, int n_cols, int flip_rows) {
    if (!path || n_rows <= 0 || n_cols <= 0 || !img) return S1G_EARG;
    const uint64_t row_bytes = 4ull * (uint64_t)n_cols, data_bytes = row_bytes * (uint64_t)n_rows;
    const uint64_t data_off = 16, ifd_off = data_off + data_bytes;
    std::vector<uint8_t> head;
    head.push_back('I');
    head.push_back('I');
    put16(head, 43);  // BigTIFF
    put16(head, 8);   // offset size
    put16(head, 0);
    put64(head, ifd_off);
    std::vector<uint8_t> ifd;
    const int n_entries = 12;
    put64(ifd, n_entries);
    entry(ifd, 256, 16, 1, (uint64_t)n_cols);   // ImageWidth (LONG8)
    entry(ifd, 257, 16, 1, (uint64_t)n_rows);   // ImageLength
    entry(ifd, 258, 3, 1, 32);                  // BitsPerSample
    entry(ifd, 259, 3, 1, 1);                   // Compression: none
    entry(ifd, 262, 3, 1, 1);                   // Photometric: min-is-black
    entry(ifd, 273, 16, 1, data_off);           // StripOffsets (one strip)
    entry(ifd, 274, 3, 1, 1);                   // Orientation: top-left
    entry(ifd, 277, 3, 1, 1);                   // SamplesPerPixel
    entry(ifd, 278, 16, 1, (uint64_t)n_rows);   // RowsPerStrip
    entry(ifd, 279, 16, 1, data_bytes);         // StripByteCounts
    entry(ifd, 284, 3, 1, 1);                   // PlanarConfiguration: contiguous
    entry(ifd, 339, 3, 1, 3);                   // SampleFormat: IEEE float
    put64(ifd, 0);                              // no next IFD
    FILE* f = fopen(path, "wb");
    if (!f) return S1G_EARG;
    bool ok = fwrite(head.data(), 1, head.size(), f) == head.size();
    for (int r = 0; r < n_rows && ok; ++r) {
        const int src = flip_rows ? n_rows - 1 - r : r;
        ok = fwrite(img + (size_t)src * n_cols, 4, (size_t)n_cols, f) == (size_t)n_cols;
    }
    ok = ok && fwrite(ifd.data(), 1, ifd.size(), f) == ifd.size();
    ok = (fclose(f) == 0) && ok;
    return ok ? S1G_OK : S1G_EHIP;
}

}  // extern "C"
