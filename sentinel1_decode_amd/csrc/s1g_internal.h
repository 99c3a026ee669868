// s1g_internal.h -- declarations shared between the kernel files and the
// C-ABI implementation (host side).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/s1gpu.h"
#include "s1g_device.h"

namespace s1g {

// Device-resident FFT plan of one length: radix list plus twiddle tables
// (FP64 and FP32, e^{-2 pi i k / n}) and the mixed-radix digit-reversal map
// perm[i] = position of input element i for the in-place DIT passes.
struct DevFft {
    FftPlan plan;
    cpx<double>* tw64 = nullptr;
    cpx<float>* tw32 = nullptr;
    int32_t* perm = nullptr;
};

template <typename T>
__host__ __device__ inline const cpx<T>* tw_of(const DevFft& f);
template <>
__host__ __device__ inline const cpx<double>* tw_of<double>(const DevFft& f) { return f.tw64; }
template <>
__host__ __device__ inline const cpx<float>* tw_of<float>(const DevFft& f) { return f.tw32; }

// Per-column constants of the range-Doppler stage.
struct AzColumn {
    double a;     // lambda^2 / (4 v_rel^2)
    double b;     // 4 pi R_j / lambda
    double c;     // 2 pi t_corr_j
    double rdr;   // R_j / range_spacing
    double fdc;   // Doppler centroid
    double amp;   // 1 / sqrt((R_ref / R_j)^3)
    double sq;    // sqrt((R_ref / R_j)^3), the reference's divisor
    double bt;    // b / (2 pi): phase coefficients in turns
    double ct;    // c / (2 pi)
    double btf;   // bt - rint(bt)
};

// Summary of the column table for the FP32 chain's linear-geometry kernel:
// `ok` when v_rel and f_dc are the same in every column and R_j / dR, the
// phase coefficient bt and t_corr are linear in j (to the tolerances of
// az_uniform_kernel), with the fit's base values and slopes.
struct AzUniform {
    int ok;
    double a, fdc;        // lambda^2 / (4 v^2), Doppler centroid
    double rdr0, drdr;    // R_j / dR = rdr0 + j drdr
    double bt0, dbt;      // bt_j = bt0 + j dbt (turns)
    double ct0, dct;      // t_corr_j = ct0 + j dct
};

// Bytes of the column table of n columns: AzColumn[n], AzUniform, float amplitude[n].
constexpr size_t az_table_bytes(int n) {
    return ((sizeof(AzColumn) * (size_t)n + sizeof(AzUniform) + 15) & ~(size_t)15) + sizeof(float) * (size_t)n;
}
inline AzUniform* az_uniform_of(AzColumn* cols, int n) { return (AzUniform*)(cols + n); }
inline float* az_amp_of(AzColumn* cols, int n) {
    return (float*)((char*)cols + ((sizeof(AzColumn) * (size_t)n + sizeof(AzUniform) + 15) & ~(size_t)15));
}

// Bit widths of the 8 primary (48 bits) and 46 secondary (496 bits) header
// fields of a space packet, in header order (structs.h:42-115).
#define S1G_HEADER_FIELD_BITS                                                                                   \
    3, 1, 1, 7, 4, 2, 14, 16, 32, 16, 32, 32, 8, 1, 3, 4, 32, 8, 16, 32, 32, 1, 2, 5, 8, 8, 8, 8, 1, 15, 1, 15, 24, \
        3, 5, 24, 24, 24, 1, 3, 2, 2, 4, 2, 10, 2, 1, 5, 4, 3, 1, 8, 16, 8

// Context internals (capi.hip) for entry points defined in other files.
hipStream_t ctx_stream(const s1g_ctx* c);
int ctx_device(const s1g_ctx* c);
// Sets the context's error message and returns code.
int ctx_error(s1g_ctx* c, int code, const char* msg);
// The context's general scratch buffer, grown to at least `bytes` (shared
// with the DFT entry points; valid until the next call that uses it).
int ctx_scratch(s1g_ctx* c, size_t bytes, void** p);
// Buffer `slot` (0..15) of the context, grown to at least `bytes`; kept
// until the context is destroyed (valid until the next call that grows it).
int ctx_aux(s1g_ctx* c, int slot, size_t bytes, void** p);
// In-place DFT of `rows` rows of length n (Stockham, else Bluestein); the
// inverse is scaled by 1 / n when `normalise`.
int ctx_dft_rows(s1g_ctx* c, int prec, int rows, int n, void* data, int64_t ld, int inverse, bool normalise);
// RCMC launch variant for a shape without timing any: the forced variant,
// else the tuner's cached choice, else `fallback`; recorded as the stage's
// last variant.
int ctx_rcmc_variant(s1g_ctx* c, int prec, int n_lines, int n_samples, int apply_rcmc, int fallback);

// Decode the packets whose length l satisfies lo_bytes < l <= hi_bytes (the
// others are left to another launch), LDS sized for max_bytes.
int launch_fdbaq_decode(hipStream_t stream, int prec, const uint8_t* ud, const uint64_t* offs, const uint32_t* lens,
                        const int32_t* nqs, int n_packets, int max_bytes, int max_quads, void* out, int64_t ld_out,
                        int32_t* status, uint32_t lo_bytes = 0u, uint32_t hi_bytes = 0xFFFFFFFFu);
// Packets up to this length decode with LDS sized for it when a burst also
// holds longer ones (launch_fdbaq_decode per size class).
constexpr int kDecodeSplitBytes = 20 * 1024;

// Out-of-place Stockham FFT of `batch` vectors of length n (vector b at
// in + b*dist) in global memory, one launch per pass; the result ends in
// `out` (`tmp` is scratch of the same size).
int launch_fft_global(hipStream_t stream, int prec, const DevFft& f, const void* in, void* out, void* tmp, int batch,
                      int64_t dist, int inverse);

// Several bursts in one launch: per-instance parameters passed by value in
// the kernel arguments (a workgroup finds its instance by a scan of them),
// at most kMaxBatch instances per launch.
constexpr int kMaxBatch = 16;
struct RangeInst {
    const void* in;
    void* out;
    const void* h;  // the instance's overlap-save filter
    int64_t ld_in, ld_out;
    int n_lines, N, V, nblk, cshift;
    int wg0;  // first workgroup of the instance (set by the launcher)
};
struct RangeBatch {
    int n;
    RangeInst inst[kMaxBatch];
};
struct AzInst {
    void* data;
    int64_t ld;
    int N;
    int panel0;  // first panel of the instance (set by the launcher)
};
struct AzBatch {
    int n, npanels;
    AzInst inst[kMaxBatch];
};
struct AzColumn;
struct AzUniform;
struct RcmcInst {
    const void* in;
    void* out;
    int64_t ld_in, ld_out;
    const AzColumn* cols;  // the instance's column table (launch_az_columns)
    const AzUniform* uni;  // its linear-geometry summary and FP32 amplitudes
    const float* amp32;
    double prf;
    int n_lines, N, apply_rcmc;
    int nbx;  // workgroups per Doppler row (set by the launcher)
    int wg0;  // first workgroup of the instance (set by the launcher)
};
struct RcmcBatch {
    int n;
    RcmcInst inst[kMaxBatch];
};

// Range compression: build the overlap-save filter, then filter the lines.
int launch_range_filter(hipStream_t stream, const DevFft& fm, double txpsf, double txprr, double txpl, double fs_mhz,
                        int n_samples, int nc, void* scratch, void* h_out, int prec);
int launch_range_os(hipStream_t stream, int prec, const DevFft& fm, const void* in, int64_t ld_in, void* out,
                    int64_t ld_out, int n_lines, int n_samples, int nc, const void* h);
// Batched form: instances sharing the block length fm.plan.n, which must
// have a register-pass kernel (range_batchable).  range_inst fills one
// instance.
bool range_batchable(int prec, int block_len);
RangeInst range_inst(const void* in, int64_t ld_in, void* out, int64_t ld_out, int n_lines, int n_samples,
                     int block_len, int nc, const void* h);
int launch_range_os_batch(hipStream_t stream, int prec, const DevFft& fm, const RangeInst* inst, int n);

// Azimuth stages.
// Launch variants (candidates of the autotuner): variant counts and the
// launch with a given variant.
int azimuth_fft_variants(const DevFft& fl);
int rcmc_variants(int prec);
// shift = 1: fftshift after the forward transform (the azimuth chain);
// shift = 0: plain transform.  The inverse is scaled by 1/n_lines.
int launch_azimuth_fft(hipStream_t stream, int prec, const DevFft& fl, void* data, int64_t ld, int n_lines,
                       int n_samples, int inverse, int variant, int shift = 1);
// Batched form (instances of one azimuth length): the persistent variants
// (azimuth_batchable) walk the panels of every instance in one launch.
bool azimuth_batchable(const DevFft& fl, int variant);
int launch_azimuth_fft_batch(hipStream_t stream, int prec, const DevFft& fl, const AzInst* inst, int n, int inverse,
                             int variant);
int launch_az_columns(hipStream_t stream, const s1g_azimuth_desc& d, AzColumn* cols);
// Batched RCMC + matched filter for the FP32 linear-geometry variants
// (rcmc_batchable); each instance's column table filled beforehand.
bool rcmc_batchable(int prec, int variant);
int launch_rcmc_batch(hipStream_t stream, const RcmcInst* inst, int n, int variant);
int launch_rcmc_mf(hipStream_t stream, int prec, const s1g_azimuth_desc& d, const AzColumn* cols, const void* in,
                   int64_t ld_in, void* out, int64_t ld_out, int variant);

// ECCM.
// Out of place (windows overlap, so workgroups would race in place).
// fN: plan of length N (twiddles only), fsub: plan of N / r, fw: window plan;
// partial: n_lines values of the detection maxima.
// With `stage` (rows of ld_stage elements, L of them): in and out may be the
// same rows; flagged lines are copied to stage by the detection kernel and
// notched from there into out, unflagged lines are not touched.
int launch_eccm(hipStream_t stream, int prec, const s1g_eccm_desc& d, const DevFft& fN, const DevFft& fsub, int r,
                const DevFft& fw, const void* in, int64_t ld_in, void* out, int64_t ld_out, uint8_t* flags,
                void* partial, void* stage = nullptr, int64_t ld_stage = 0);
// Detection split r for lines of N samples (0: none), given the planner.
int eccm_detect_split(int prec, int N, bool (*factor)(int, FftPlan&));

// General signal processing (dsp.hip).
int launch_copy_scale(hipStream_t stream, int prec, const void* src, int64_t lds, void* dst, int64_t ldd, int rows,
                      int n, double scale);
int launch_transpose(hipStream_t stream, int prec, const void* src, int64_t lds, void* dst, int64_t ldd, int rows,
                     int cols);
// Rows that fit one workgroup's LDS (at most kRowLdsBytes per row): in-place
// DFT of `rows` rows of length f.plan.n, one workgroup per row, the result
// multiplied by `scale`.  -1 when the plan or the length does not qualify.
constexpr size_t kRowLdsBytes = 160 * 1024;
bool fft_rows_lds_fits(int prec, int n);
int launch_fft_rows_lds(hipStream_t stream, int prec, const DevFft& f, void* data, int64_t ld, int rows, int inverse,
                        double scale);
// Bluestein: filter spectrum for an n-point transform through M points
// (scratch 3 M complex doubles), then the transform of `rows` rows of x in
// place (w1..w3: rows * M complex each); `scale` multiplies the result.
int launch_bluestein_filter(hipStream_t stream, int prec, const DevFft& fM, int n, int inverse, void* scratch,
                            void* bhat);
int launch_bluestein(hipStream_t stream, int prec, const DevFft& fM, const void* bhat, void* x, int64_t ldx, int rows,
                     int n, int inverse, double scale, void* w1, void* w2, void* w3);
int launch_mul_rows(hipStream_t stream, int prec, void* a, int64_t lda, const void* v, int rows, int n, int conj_v);
int launch_chirp_energy(hipStream_t stream, int prec, const void* chirp, int n, double* energy);
int launch_conj_div(hipStream_t stream, int prec, void* ref, int n, const double* energy);
int launch_replica_chirp(hipStream_t stream, int prec, void* out, int len, int nc, double txpsf, double txprr,
                         double txpl);
int launch_sinc_interp(hipStream_t stream, int prec, const void* x, int len, const double* t, int m, int L, void* out);

// Autofocus.  work: af_work_doubles() doubles; scal: 3 doubles (S, T, E).
// img (may be null): the image, n_lines * n_samples complex (dense).
// Lengths with compiled radices (af_fast) recompute the image from rd in
// each pass; other lengths keep it in img_tmp (n_lines * n_samples complex)
// between the entropy and gradient kernels.
bool af_fast(int n_lines, int n_samples, int64_t ld, int prec);
size_t af_work_doubles(int n_lines, int n_samples, int64_t ld, int prec);
// the compiled-radix passes (autofocus_reg.hip): pass 0 entropy partials
// (and the image when img is set), pass 2 gradient partials; one partial
// per workgroup, af_reg_workgroups() of them
int af_reg_workgroups(int n_samples, int prec);
int launch_af_reg(hipStream_t stream, int prec, int pass, int n_lines, int n_samples, const void* rd, int64_t ld,
                  const DevFft& fl, const void* ph, const double* scal, void* img, double* part, double* gpart);
int launch_af_sum_partials(hipStream_t stream, const double* gpart, int nparts, int n_lines, double* grad);
int launch_af(hipStream_t stream, int prec, int n_lines, int n_samples, const void* rd, int64_t ld,
              const double* phase, const DevFft& fl, void* img, void* img_tmp, double* work, int want_grad,
              double* grad, double* scal);

}  // namespace s1g
