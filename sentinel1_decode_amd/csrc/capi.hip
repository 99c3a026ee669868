// capi.hip -- the C ABI (include/s1gpu.h): contexts, FFT plans, cached range
// filters, workspace, and the chain entry points.  Host code only; kernels
// live in the other .hip files.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <cstdlib>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <tuple>
#include <vector>

#include "s1g_fdbaq.h"
#include "s1g_internal.h"

using namespace s1g;

struct FilterKey {
    double txpsf, txprr, txpl, fs;
    int n, m, prec;
    bool operator<(const FilterKey& o) const {
        return std::tie(txpsf, txprr, txpl, fs, n, m, prec) < std::tie(o.txpsf, o.txprr, o.txpl, o.fs, o.n, o.m, o.prec);
    }
};

// Autotuner cache key: stage, precision and the shape parameters that
// decide the stage's launch (lines, samples, replica length).
struct TuneKey {
    int stage, prec, a, b, c;
    bool operator<(const TuneKey& o) const {
        return std::tie(stage, prec, a, b, c) < std::tie(o.stage, o.prec, o.a, o.b, o.c);
    }
};

struct s1g_ctx {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    std::string err;
    std::map<int, DevFft> ffts;
    std::map<std::pair<int, int>, DevFft> fft_orders;  // (n, order) -> plan with reordered radices
    std::map<FilterKey, void*> filters;
    void* filter_scratch = nullptr;
    size_t filter_scratch_bytes = 0;
    void* ws = nullptr;  // chain workspace
    size_t ws_bytes = 0;
    AzColumn* cols = nullptr;
    size_t cols_n = 0;
    void* eccm_scratch = nullptr;
    size_t eccm_scratch_bytes = 0;
    double* af_scalars = nullptr;  // autofocus scalars and partial sums
    size_t af_n = 0;
    void* af_img = nullptr;
    size_t af_img_bytes = 0;
    // Captured batch graphs bake the workspace pointers (ws, cols,
    // eccm_scratch) into their kernel arguments: each remembers the
    // workspace generation it was captured under, and ensure() moves the
    // generation on whenever it reallocates, so a stale graph is captured
    // again instead of replayed into freed memory.
    struct CachedGraph {
        hipGraphExec_t exec = nullptr;
        hipGraph_t graph = nullptr;
        uint64_t gen = 0;
    };
    std::map<std::string, CachedGraph> graphs;
    uint64_t ws_gen = 0;
    // stage timers
    bool profiling = false;
    bool capturing = false;
    std::vector<hipEvent_t> event_pool;
    // batches: the stream decodes run on, and ordering events
    hipStream_t side = nullptr;
    // the long packets' decode class runs here beside the short ones
    hipStream_t dec_side = nullptr;
    hipEvent_t dec_fork = nullptr, dec_join = nullptr;
    std::vector<hipEvent_t> sync_events;
    struct Mark {
        int stage;
        hipEvent_t a, b;
    };
    std::vector<Mark> marks;
    double stage_ms[S1G_NSTAGES] = {0};
    int64_t stage_n[S1G_NSTAGES] = {0};
    // general DSP entry points
    void* dft_ws = nullptr;
    size_t dft_ws_bytes = 0;
    void* blue_scratch = nullptr;
    size_t blue_scratch_bytes = 0;
    std::map<std::tuple<int, int, int>, void*> blue_filters;  // (n, prec, inverse)
    double* dsp_scal = nullptr;
    // buffers of the entry points defined in other files (ctx_aux)
    void* aux[16] = {};
    size_t aux_bytes[16] = {};
    // autotuner
    bool autotune = true;
    int force[S1G_NSTAGES] = {-1, -1, -1, -1, -1, -1, -1, -1, -1};
    int last_var[S1G_NSTAGES] = {0};
    int last_nvar[S1G_NSTAGES] = {0};
    struct Tuned {
        int variant, n_variants;
        float ms;  // the winner's time when it was tuned (0: loaded without one)
    };
    std::map<TuneKey, Tuned> tuned;
    std::string tune_file;  // S1G_TUNE_CACHE at creation: loaded then, saved at destruction
};

// Brackets one stage's launches with events when profiling is on.
struct StageTimer {
    s1g_ctx* c;
    int stage;
    hipEvent_t a = nullptr, b = nullptr;
    hipEvent_t take() {
        if (!c->event_pool.empty()) {
            hipEvent_t e = c->event_pool.back();
            c->event_pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
    StageTimer(s1g_ctx* ctx, int st) : c(ctx), stage(st) {
        if (!c->profiling || c->capturing) return;
        a = take();
        b = take();
        if (a) (void)hipEventRecord(a, c->stream);
    }
    ~StageTimer() {
        if (!a || !b) return;
        (void)hipEventRecord(b, c->stream);
        c->marks.push_back({stage, a, b});
    }
};

namespace {

int fail(s1g_ctx* c, int code, const char* fmt, const char* detail = "") {
    if (c) {
        char buf[512];
        snprintf(buf, sizeof buf, fmt, detail);
        c->err = buf;
    }
    return code;
}

int hip_fail(s1g_ctx* c, hipError_t e, const char* where) {
    if (c) c->err = std::string(where) + ": " + hipGetErrorString(e);
    return S1G_EHIP;
}

#define HIPCHK(ctx, call)                                       \
    do {                                                        \
        hipError_t _e = (call);                                 \
        if (_e != hipSuccess) return hip_fail(ctx, _e, #call);  \
    } while (0)

// Factor n over the LDS radices; biggest odd factors first (their pass then
// runs with Ns = 1 and needs no twiddles), then 8s, a 4 or 2 for the rest.
bool factorise(int n, FftPlan& p) {
    static const int odd[] = {71, 67, 61, 59, 53, 47, 43, 41, 37, 31, 29, 23, 19, 17, 13, 11, 7, 5, 3};
    p.n = n;
    p.npass = 0;
    int r = n;
    std::vector<int> rad;
    for (int f : odd)
        while (r % f == 0) {
            rad.push_back(f);
            r /= f;
        }
    while (r % 8 == 0) {
        rad.push_back(8);
        r /= 8;
    }
    if (r % 4 == 0) {
        rad.push_back(4);
        r /= 4;
    }
    if (r % 2 == 0) {
        rad.push_back(2);
        r /= 2;
    }
    if (r != 1 || rad.size() > 24) return false;
    for (size_t i = 0; i < rad.size(); ++i) p.radix[i] = rad[i];
    p.npass = (int)rad.size();
    return true;
}

// Digit-reversed position of input i for DIT passes radix[0..S-1].
std::vector<int32_t> digit_reversal(const FftPlan& p) {
    const int n = p.n;
    std::vector<int32_t> perm(n);
    for (int i = 0; i < n; ++i) {
        int64_t pos = i, rem = i, m = n;
        if (p.npass > 0) pos = 0;
        for (int s = p.npass - 1; s >= 0; --s) {
            const int R = p.radix[s];
            m /= R;
            pos += (rem % R) * m;
            rem /= R;
        }
        perm[i] = (int32_t)pos;
    }
    return perm;
}

int get_fft(s1g_ctx* c, int n, const DevFft** out) {
    auto it = c->ffts.find(n);
    if (it != c->ffts.end()) {
        *out = &it->second;
        return 0;
    }
    DevFft f;
    if (!factorise(n, f.plan)) {
        // twiddle table only (the length has a prime factor above 71): usable
        // as a table of e^{-2 pi i k / n}, not as an LDS transform plan
        f.plan.n = n;
        f.plan.npass = -1;
    }
    std::vector<cpx<double>> t64(n);
    std::vector<cpx<float>> t32(n);
    for (int k = 0; k < n; ++k) {
        // exact angle: reduce k/n to the first half turn before evaluating
        long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)n;
        double cr = (double)cosl(a), ci = (double)sinl(a);
        t64[k].x = cr;
        t64[k].y = ci;
        t32[k].x = (float)cr;
        t32[k].y = (float)ci;
    }
    const std::vector<int32_t> perm = digit_reversal(f.plan);
    HIPCHK(c, hipMalloc(&f.tw64, sizeof(cpx<double>) * n));
    HIPCHK(c, hipMalloc(&f.tw32, sizeof(cpx<float>) * n));
    HIPCHK(c, hipMalloc(&f.perm, sizeof(int32_t) * n));
    HIPCHK(c, hipMemcpy(f.tw64, t64.data(), sizeof(cpx<double>) * n, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(f.tw32, t32.data(), sizeof(cpx<float>) * n, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(f.perm, perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    auto ins = c->ffts.emplace(n, f);
    *out = &ins.first->second;
    return 0;
}

// Radix orders the DFT tuner tries for a plan (the planner's own order
// first): the same radices reversed, ascending, and with the powers of two
// first; orders equal to an earlier one are dropped.  Any order is a valid
// plan; it changes the passes' strides and twiddle patterns.
int radix_orders(const FftPlan& p, FftPlan* out, int max_orders) {
    std::vector<std::vector<int>> seen;
    auto add = [&](std::vector<int> r) {
        for (const auto& q : seen)
            if (q == r) return;
        seen.push_back(r);
    };
    std::vector<int> base(p.radix, p.radix + p.npass);
    add(base);
    std::vector<int> rev(base.rbegin(), base.rend());
    add(rev);
    std::vector<int> asc = base;
    std::sort(asc.begin(), asc.end());
    add(asc);
    std::vector<int> pow2, odd;
    for (int r : base) ((r & (r - 1)) == 0 ? pow2 : odd).push_back(r);
    pow2.insert(pow2.end(), odd.begin(), odd.end());
    add(pow2);
    int k = 0;
    for (const auto& r : seen) {
        if (k == max_orders) break;
        out[k].n = p.n;
        out[k].npass = (int)r.size();
        for (size_t i = 0; i < r.size(); ++i) out[k].radix[i] = r[i];
        ++k;
    }
    return k;
}

// The plan of length n with its radices in order `order` of radix_orders
// (order 0: the default plan), sharing the default plan's twiddle tables.
int get_fft_order(s1g_ctx* c, int n, int order, const DevFft** out) {
    const DevFft* f0;
    int rc = get_fft(c, n, &f0);
    if (rc || order == 0 || f0->plan.npass < 0) {
        *out = f0;
        return rc;
    }
    auto it = c->fft_orders.find({n, order});
    if (it != c->fft_orders.end()) {
        *out = &it->second;
        return 0;
    }
    FftPlan ord[4];
    const int no = radix_orders(f0->plan, ord, 4);
    if (order >= no) return fail(c, S1G_EARG, "no such radix order%s");
    DevFft f;
    f.plan = ord[order];
    f.tw64 = f0->tw64;
    f.tw32 = f0->tw32;
    const std::vector<int32_t> perm = digit_reversal(f.plan);
    HIPCHK(c, hipMalloc(&f.perm, sizeof(int32_t) * n));
    HIPCHK(c, hipMemcpy(f.perm, perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    auto ins = c->fft_orders.emplace(std::make_pair(n, order), f);
    *out = &ins.first->second;
    return 0;
}

int ensure(s1g_ctx* c, void** p, size_t* have, size_t need) {
    if (*have >= need) return 0;
    // everything a captured chain uses is sized before the capture starts
    if (c->capturing) return fail(c, S1G_EHIP, "workspace reallocation during graph capture%s");
    ++c->ws_gen;
    if (*p) hipFree(*p);
    *p = nullptr;
    *have = 0;
    hipError_t e = hipMalloc(p, need);
    if (e != hipSuccess) return fail(c, S1G_ENOMEM, "device allocation failed%s");
    *have = need;
    return 0;
}

size_t esize(int prec) { return prec == S1G_F32 ? sizeof(cpx<float>) : sizeof(cpx<double>); }

bool valid_prec(int prec) { return prec == S1G_F32 || prec == S1G_F64; }

// Overlap-save FFT length: the smallest power of two holding a whole line
// plus the chirp, else the largest one up to mmax (at most what fits LDS:
// 16384 complex64, 8192 complex128).
int range_block_len(int prec, int n, int nc, int mmax = 0) {
    if (mmax <= 0) mmax = prec == S1G_F32 ? 16384 : 8192;
    int m = 1;
    while (m < n + nc - 1 && m < mmax) m <<= 1;
    if (m <= nc) return -1;
    return m;
}

// Candidate overlap-save lengths (autotuner variants): the longest block
// (fewest blocks per line) and, for FP32, the 8192 block that leaves room
// for two workgroups per CU.
int range_candidates(int prec, int n, int nc, int* cand) {
    int k = 0;
    const int caps[2] = {prec == S1G_F32 ? 16384 : 8192, 8192};
    for (int cap : caps) {
        const int m = range_block_len(prec, n, nc, cap);
        if (m > 0 && (k == 0 || cand[k - 1] != m)) cand[k++] = m;
    }
    return k;
}

// Variant of a stage for this shape: forced, cached, or timed now.  run(v)
// launches variant v on the context stream and returns 0 or an error; the
// candidates are timed (best of two after a warm-up launch) only outside
// graph capture, and the winner is cached.
template <typename F>
int choose_variant(s1g_ctx* c, const TuneKey& k, int nvar, F&& run) {
    int v = 0;
    if (c->force[k.stage] >= 0) {
        v = std::min(c->force[k.stage], nvar - 1);
    } else {
        auto it = c->tuned.find(k);
        if (it != c->tuned.end() && it->second.n_variants == nvar) {
            v = it->second.variant;
        } else if (c->autotune && !c->capturing && nvar > 1) {
            hipEvent_t a = nullptr, b = nullptr;
            if (hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess) {
                float best = 1e30f;
                for (int cand = 0; cand < nvar; ++cand) {
                    if (run(cand)) continue;
                    float t = 1e30f;
                    for (int rep = 0; rep < 2; ++rep) {
                        (void)hipEventRecord(a, c->stream);
                        if (run(cand)) break;
                        (void)hipEventRecord(b, c->stream);
                        float ms = 0.f;
                        if (hipEventSynchronize(b) == hipSuccess && hipEventElapsedTime(&ms, a, b) == hipSuccess)
                            t = std::min(t, ms);
                    }
                    if (t < best) {
                        best = t;
                        v = cand;
                    }
                }
                c->tuned[k] = s1g_ctx::Tuned{v, nvar, best};
            }
            if (a) hipEventDestroy(a);
            if (b) hipEventDestroy(b);
            c->err.clear();
        }
    }
    c->last_var[k.stage] = v;
    c->last_nvar[k.stage] = nvar;
    return v;
}

int get_plan(s1g_ctx* c, int n, const DevFft** out) {
    int rc = get_fft(c, n, out);
    if (rc) return rc;
    if ((*out)->plan.npass < 0) return fail(c, S1G_EUNSUPPORTED, "FFT length has a prime factor above 71%s");
    return 0;
}

int get_filter(s1g_ctx* c, const s1g_range_desc& d, int prec, int nc, const DevFft* fm, void** h) {
    FilterKey k{d.txpsf_mhz, d.txprr_mhz_us, d.txpl_us, d.fs_mhz, d.n_samples, fm->plan.n, prec};
    auto it = c->filters.find(k);
    if (it != c->filters.end()) {
        *h = it->second;
        return 0;
    }
    const int m = fm->plan.n;
    int rc = ensure(c, &c->filter_scratch, &c->filter_scratch_bytes, sizeof(cpx<double>) * 3 * (size_t)m);
    if (rc) return rc;
    void* buf = nullptr;
    HIPCHK(c, hipMalloc(&buf, esize(prec) * m));
    rc = launch_range_filter(c->stream, *fm, d.txpsf_mhz, d.txprr_mhz_us, d.txpl_us, d.fs_mhz, d.n_samples, nc,
                             c->filter_scratch, buf, prec);
    if (rc) return hip_fail(c, (hipError_t)rc, "range filter");
    c->filters[k] = buf;
    *h = buf;
    return 0;
}

int check_launch(s1g_ctx* c, int rc, const char* what) {
    if (rc == -3) return fail(c, S1G_EUNSUPPORTED, "%s: shape does not fit the kernel", what);
    if (rc == -1) return fail(c, S1G_EUNSUPPORTED, "%s: unsupported size", what);
    if (rc) return hip_fail(c, (hipError_t)rc, what);
    return 0;
}

// The range stage's block length (forced, cached or timed now on these
// buffers) with its plan and filter; *nc_out the replica length.
static int range_setup(s1g_ctx* c, const s1g_range_desc& d, int prec, const void* in, int64_t ld_in, void* out,
                       int64_t ld_out, const DevFft** fm_out, void** h_out, int* nc_out) {
    const int nc = (int)floor(d.fs_mhz * d.txpl_us);
    if (nc < 2) return fail(c, S1G_EARG, "replica chirp shorter than 2 samples%s");
    int cand[2];
    const int nvar = range_candidates(prec, d.n_samples, nc, cand);
    if (nvar == 0) return fail(c, S1G_EUNSUPPORTED, "replica chirp longer than the overlap-save block%s");
    auto prepare = [&](int v, const DevFft** fm, void** h) {
        int rc = get_plan(c, cand[v], fm);
        if (!rc) rc = get_filter(c, d, prec, nc, *fm, h);
        return rc;
    };
    auto run = [&](int v) {
        const DevFft* fm;
        void* h;
        int rc = prepare(v, &fm, &h);
        if (rc) return rc;
        return check_launch(c, launch_range_os(c->stream, prec, *fm, in, ld_in, out, ld_out, d.n_lines, d.n_samples,
                                               nc, h),
                            "range compression");
    };
    const int v = choose_variant(c, TuneKey{S1G_STAGE_RANGE, prec, d.n_lines, d.n_samples, nc}, nvar, run);
    *nc_out = nc;
    StageTimer tm(c, S1G_STAGE_FILTER);
    return prepare(v, fm_out, h_out);
}

int do_range(s1g_ctx* c, const s1g_range_desc& d, int prec, const void* in, int64_t ld_in, void* out,
             int64_t ld_out) {
    if (d.n_lines <= 0 || d.n_samples <= 0) return d.n_lines < 0 || d.n_samples < 0 ? fail(c, S1G_EARG, "bad range shape%s") : 0;
    const DevFft* fm;
    void* h;
    int nc;
    int rc = range_setup(c, d, prec, in, ld_in, out, ld_out, &fm, &h, &nc);
    if (rc) return rc;
    StageTimer tm(c, S1G_STAGE_RANGE);
    return check_launch(c, launch_range_os(c->stream, prec, *fm, in, ld_in, out, ld_out, d.n_lines, d.n_samples, nc, h),
                        "range compression");
}

// The azimuth transform's plan and launch variant for an L x N array
// (forced, cached, or timed now on a scratch array: the transform is in
// place).
static int az_setup(s1g_ctx* c, int prec, int L, int N, int inverse, const DevFft** fl_out, int* v_out) {
    const DevFft* fl;
    int rc = get_plan(c, L, &fl);
    if (rc) return rc;
    const int stage = inverse ? S1G_STAGE_AZ_IFFT : S1G_STAGE_AZ_FFT;
    void* tmp = nullptr;
    auto run = [&](int v) {
        if (!tmp) {
            if (hipMalloc(&tmp, esize(prec) * (size_t)L * N) != hipSuccess) {
                tmp = nullptr;
                return (int)S1G_ENOMEM;
            }
            (void)hipMemsetAsync(tmp, 0, esize(prec) * (size_t)L * N, c->stream);
        }
        return check_launch(c, launch_azimuth_fft(c->stream, prec, *fl, tmp, N, L, N, inverse, v), "azimuth fft");
    };
    const int v = choose_variant(c, TuneKey{stage, prec, L, N, 0}, azimuth_fft_variants(*fl), run);
    if (tmp) {
        (void)hipStreamSynchronize(c->stream);
        hipFree(tmp);
    }
    *fl_out = fl;
    *v_out = v;
    return 0;
}

int do_azimuth_fft(s1g_ctx* c, int prec, int L, int N, void* data, int64_t ld, int inverse) {
    if (L <= 0 || N <= 0) return 0;
    const DevFft* fl;
    int v;
    int rc = az_setup(c, prec, L, N, inverse, &fl, &v);
    if (rc) return rc;
    StageTimer tm(c, inverse ? S1G_STAGE_AZ_IFFT : S1G_STAGE_AZ_FFT);
    return check_launch(c, launch_azimuth_fft(c->stream, prec, *fl, data, ld, L, N, inverse, v), "azimuth fft");
}

// The column table of d in cols, and the RCMC stage's launch variant
// (forced, cached or timed now on these buffers).
static int rcmc_setup(s1g_ctx* c, const s1g_azimuth_desc& d, int prec, AzColumn* cols, const void* in,
                      int64_t ld_in, void* out, int64_t ld_out, int* v_out) {
    if (d.n_lines <= 1 || d.n_samples <= 0) return fail(c, S1G_EARG, "azimuth stage needs >= 2 lines%s");
    if (!d.slant_ranges || !d.v_rel || !d.f_dc || !d.t_corr) return fail(c, S1G_EARG, "null geometry array%s");
    int rc = launch_az_columns(c->stream, d, cols);
    if (rc) return hip_fail(c, (hipError_t)rc, "azimuth columns");
    auto run = [&](int v) {
        return check_launch(c, launch_rcmc_mf(c->stream, prec, d, cols, in, ld_in, out, ld_out, v),
                            "rcmc/matched filter");
    };
    *v_out = choose_variant(c, TuneKey{S1G_STAGE_RCMC_MF, prec, d.n_lines, d.n_samples, d.apply_rcmc},
                            rcmc_variants(prec), run);
    return 0;
}

int do_rcmc(s1g_ctx* c, const s1g_azimuth_desc& d, int prec, const void* in, int64_t ld_in, void* out,
            int64_t ld_out) {
    if (d.n_lines <= 1 || d.n_samples <= 0) return fail(c, S1G_EARG, "azimuth stage needs >= 2 lines%s");
    int rc = ensure(c, (void**)&c->cols, &c->cols_n, az_table_bytes(d.n_samples));
    if (rc) return rc;
    int v;
    rc = rcmc_setup(c, d, prec, c->cols, in, ld_in, out, ld_out, &v);
    if (rc) return rc;
    StageTimer tm(c, S1G_STAGE_RCMC_MF);
    return check_launch(c, launch_rcmc_mf(c->stream, prec, d, c->cols, in, ld_in, out, ld_out, v),
                        "rcmc/matched filter");
}

int do_eccm(s1g_ctx* c, const s1g_eccm_desc& d, int prec, const void* in, int64_t ld_in, void* out, int64_t ld_out,
            uint8_t* flags, void* stage = nullptr, int64_t ld_stage = 0) {
    const int L = d.n_lines, N = d.n_samples;
    if (L <= 0 || N <= 0) return L < 0 || N < 0 ? fail(c, S1G_EARG, "bad ECCM shape%s") : 0;
    const int fsz = d.fft_size;
    if (fsz < 4 || fsz > 1024 || (fsz & (fsz - 1)) || d.stride < 1 || d.stride > fsz)
        return fail(c, S1G_EUNSUPPORTED, "ECCM window must be a power of two in [4, 1024] with 1 <= stride <= size%s");
    if ((fsz + d.stride - 1) / d.stride > 8) return fail(c, S1G_EUNSUPPORTED, "ECCM window overlap too deep%s");
    // detection split r: each sub-spectrum N/r is transformed in LDS (a
    // 128 KB bound, r = 2 for the IW2 FP32 line, ran one workgroup per CU)
    const int r = eccm_detect_split(prec, N, factorise);
    if (!r) return fail(c, S1G_EUNSUPPORTED, "ECCM detection length does not split over the LDS radices%s");
    const DevFft *fN, *fsub, *fw;
    int rc = get_fft(c, N, &fN);
    if (!rc) rc = get_plan(c, N / r, &fsub);
    if (!rc) rc = get_plan(c, fsz, &fw);
    if (rc) return rc;
    rc = ensure(c, &c->eccm_scratch, &c->eccm_scratch_bytes, sizeof(double) * (size_t)L);
    if (rc) return rc;
    StageTimer tm(c, S1G_STAGE_ECCM);
    return check_launch(c, launch_eccm(c->stream, prec, d, *fN, *fsub, r, *fw, in, ld_in, out, ld_out, flags,
                                       c->eccm_scratch, stage, ld_stage),
                        "eccm");
}

}  // namespace

// Context internals for entry points defined in other files.
namespace s1g {
hipStream_t ctx_stream(const s1g_ctx* c) { return c->stream; }
int ctx_device(const s1g_ctx* c) { return c->device; }
int ctx_error(s1g_ctx* c, int code, const char* msg) {
    c->err = msg;
    return code;
}
int ctx_scratch(s1g_ctx* c, size_t bytes, void** p) {
    int rc = ensure(c, &c->dft_ws, &c->dft_ws_bytes, bytes);
    *p = rc ? nullptr : c->dft_ws;
    return rc;
}
int ctx_aux(s1g_ctx* c, int slot, size_t bytes, void** p) {
    *p = nullptr;
    if (slot < 0 || slot >= 16) return fail(c, S1G_EARG, "bad buffer slot%s");
    int rc = ensure(c, &c->aux[slot], &c->aux_bytes[slot], bytes);
    if (!rc) *p = c->aux[slot];
    return rc;
}
int ctx_rcmc_variant(s1g_ctx* c, int prec, int n_lines, int n_samples, int apply_rcmc, int fallback) {
    const int nvar = rcmc_variants(prec);
    int v = fallback;
    if (c->force[S1G_STAGE_RCMC_MF] >= 0) {
        v = c->force[S1G_STAGE_RCMC_MF];
    } else {
        auto it = c->tuned.find(TuneKey{S1G_STAGE_RCMC_MF, prec, n_lines, n_samples, apply_rcmc});
        if (it != c->tuned.end() && it->second.n_variants == nvar) v = it->second.variant;
    }
    v = std::max(0, std::min(v, nvar - 1));
    c->last_var[S1G_STAGE_RCMC_MF] = v;
    c->last_nvar[S1G_STAGE_RCMC_MF] = nvar;
    return v;
}
}  // namespace s1g

// ===================================================================== ABI
extern "C" {

int s1g_create(s1g_ctx** out, int device) {
    if (!out) return S1G_EARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return S1G_EHIP;
    if (device < 0 || device >= n) return S1G_EARG;
    s1g_ctx* c = new (std::nothrow) s1g_ctx();
    if (!c) return S1G_ENOMEM;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return S1G_EHIP;
    }
    c->stream = c->own;
    if (const char* path = getenv("S1G_TUNE_CACHE")) {
        c->tune_file = path;
        (void)s1g_tune_load(c, path);  // a missing or foreign file leaves the cache empty
        c->err.clear();
    }
    *out = c;
    return S1G_OK;
}

int s1g_destroy(s1g_ctx* c) {
    if (!c) return S1G_OK;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    if (!c->tune_file.empty()) (void)s1g_tune_save(c, c->tune_file.c_str());
    for (auto& m : c->marks) {
        hipEventDestroy(m.a);
        hipEventDestroy(m.b);
    }
    for (auto e : c->event_pool) hipEventDestroy(e);
    if (c->side) hipStreamSynchronize(c->side);
    if (c->dec_side) hipStreamSynchronize(c->dec_side);
    if (c->dec_fork) hipEventDestroy(c->dec_fork);
    if (c->dec_join) hipEventDestroy(c->dec_join);
    for (auto e : c->sync_events) hipEventDestroy(e);
    for (auto& kv : c->graphs) {
        hipGraphExecDestroy(kv.second.exec);
        hipGraphDestroy(kv.second.graph);
    }
    for (auto& kv : c->ffts) {
        hipFree(kv.second.tw64);
        hipFree(kv.second.tw32);
        hipFree(kv.second.perm);
    }
    for (auto& kv : c->fft_orders) hipFree(kv.second.perm);  // twiddles are the default plan's
    for (auto& kv : c->filters) hipFree(kv.second);
    hipFree(c->filter_scratch);
    hipFree(c->ws);
    hipFree(c->cols);
    hipFree(c->eccm_scratch);
    hipFree(c->af_scalars);
    hipFree(c->af_img);
    hipFree(c->dft_ws);
    hipFree(c->blue_scratch);
    for (auto& kv : c->blue_filters) hipFree(kv.second);
    hipFree(c->dsp_scal);
    for (int k = 0; k < 16; ++k) hipFree(c->aux[k]);
    if (c->own) hipStreamDestroy(c->own);
    if (c->side) hipStreamDestroy(c->side);
    if (c->dec_side) hipStreamDestroy(c->dec_side);
    delete c;
    return S1G_OK;
}

const char* s1g_last_error(const s1g_ctx* c) { return c ? c->err.c_str() : "null context"; }

int s1g_set_stream(s1g_ctx* c, void* stream) {
    if (!c) return S1G_EARG;
    c->stream = (hipStream_t)stream;  // NULL: the null stream
    return S1G_OK;
}

int s1g_synchronize(s1g_ctx* c) {
    if (!c) return S1G_EARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return S1G_OK;
}

int s1g_profile_enable(s1g_ctx* c, int on) {
    if (!c) return S1G_EARG;
    c->profiling = on != 0;
    return S1G_OK;
}

namespace {
// The tuning cache file: a header naming the device it was measured on,
// then one line per (stage, precision, shape): the chosen variant, the
// number of variants it was chosen among, and its time.
std::string tune_device_tag(int device) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device) != hipSuccess) return "unknown";
    char buf[160];
    snprintf(buf, sizeof buf, "%s:%d", p.gcnArchName, p.multiProcessorCount);
    for (char* q = buf; *q; ++q)
        if (*q == ' ') *q = '_';
    return buf;
}
}  // namespace

int s1g_tune_save(const s1g_ctx* c, const char* path) {
    if (!c || !path) return S1G_EARG;
    FILE* f = fopen(path, "w");
    if (!f) return S1G_EARG;
    fprintf(f, "s1gpu-tune 1 %s %zu\n", tune_device_tag(c->device).c_str(), c->tuned.size());
    for (const auto& kv : c->tuned)
        fprintf(f, "%d %d %d %d %d %d %d %.6f\n", kv.first.stage, kv.first.prec, kv.first.a, kv.first.b, kv.first.c,
                kv.second.variant, kv.second.n_variants, (double)kv.second.ms);
    return fclose(f) == 0 ? S1G_OK : S1G_EARG;
}

int s1g_tune_load(s1g_ctx* c, const char* path) {
    if (!c || !path) return S1G_EARG;
    FILE* f = fopen(path, "r");
    if (!f) return fail(c, S1G_EARG, "cannot open the tuning cache file%s");
    char tag[160] = {0};
    int version = 0;
    size_t n = 0;
    if (fscanf(f, "s1gpu-tune %d %159s %zu", &version, tag, &n) != 3 || version != 1) {
        fclose(f);
        return fail(c, S1G_EARG, "not a tuning cache file%s");
    }
    if (tune_device_tag(c->device) != tag) {
        fclose(f);
        return fail(c, S1G_EUNSUPPORTED, "tuning cache file measured on another device%s");
    }
    std::map<TuneKey, s1g_ctx::Tuned> got;
    for (size_t i = 0; i < n; ++i) {
        TuneKey k;
        s1g_ctx::Tuned t;
        double ms = 0;
        if (fscanf(f, "%d %d %d %d %d %d %d %lf", &k.stage, &k.prec, &k.a, &k.b, &k.c, &t.variant, &t.n_variants,
                   &ms) != 8 ||
            k.stage < 0 || k.stage >= S1G_NSTAGES || t.variant < 0 || t.variant >= t.n_variants) {
            fclose(f);
            return fail(c, S1G_EARG, "malformed tuning cache file%s");
        }
        t.ms = (float)ms;
        got[k] = t;
    }
    fclose(f);
    for (const auto& kv : got) c->tuned[kv.first] = kv.second;
    return S1G_OK;
}

int s1g_set_autotune(s1g_ctx* c, int on) {
    if (!c) return S1G_EARG;
    c->autotune = on != 0;
    return S1G_OK;
}

int s1g_force_variant(s1g_ctx* c, int stage, int variant) {
    if (!c || stage < 0 || stage >= S1G_NSTAGES || variant < -1) return S1G_EARG;
    c->force[stage] = variant;
    return S1G_OK;
}

int s1g_last_variant(const s1g_ctx* c, int stage, int* variant, int* n_variants) {
    if (!c || stage < 0 || stage >= S1G_NSTAGES) return S1G_EARG;
    if (variant) *variant = c->last_var[stage];
    if (n_variants) *n_variants = c->last_nvar[stage];
    return S1G_OK;
}

int s1g_profile_read(s1g_ctx* c, double* ms, int64_t* launches, int reset) {
    if (!c) return S1G_EARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (auto& m : c->marks) {
        float t = 0.f;
        if (hipEventElapsedTime(&t, m.a, m.b) == hipSuccess) {
            c->stage_ms[m.stage] += t;
            c->stage_n[m.stage] += 1;
        }
        c->event_pool.push_back(m.a);
        c->event_pool.push_back(m.b);
    }
    c->marks.clear();
    for (int i = 0; i < S1G_NSTAGES; ++i) {
        if (ms) ms[i] = c->stage_ms[i];
        if (launches) launches[i] = c->stage_n[i];
        if (reset) {
            c->stage_ms[i] = 0;
            c->stage_n[i] = 0;
        }
    }
    return S1G_OK;
}

int s1g_fdbaq_symbol(int brc, uint32_t code10, int* code_len, int* m_code) {
    if (brc < 0 || brc > 4 || code10 > 1023u) return S1G_EARG;
    // the decoder's own table-driven rule
    const int clo = fdbaq_clo(code10);
    const uint32_t x4 = brc == 4 ? 1u : 0u;
    if (code_len) *code_len = fdbaq_sym_len(code10, clo, fdbaq_len_table(brc), x4) - 1;
    if (m_code) *m_code = fdbaq_mcode_tab(code10, clo, kMBase[brc], kMShift[brc], kMMask[brc], x4);
    return S1G_OK;
}

int s1g_replica_chirp_length(double fs_mhz, double txpl_us) { return (int)floor(fs_mhz * txpl_us); }

// ---------------------------------------------------------------- headers
// Field widths of the primary (48 bits) and secondary (496 bits) headers.
static const int kFieldBits[S1G_NFIELDS] = {S1G_HEADER_FIELD_BITS};
static const int* const kPrimW = kFieldBits;
static const int* const kSecW = kFieldBits + 8;
static const double kFRef = 37.53472224;
static const double kRangeDecimation[12] = {112.6041667, 100.0925926, 0.0,         83.41049387,
                                            66.72839509, 56.30208336, 50.04629632, 25.02314816,
                                            64.34523813, 46.91840280, 17.32371796, 54.59595962};

static uint64_t be_field(const uint8_t* p, int64_t bit, int n) {
    uint64_t v = 0;
    for (int i = 0; i < n; ++i) {
        const int64_t b = bit + i;
        v = (v << 1) | ((p[b >> 3] >> (7 - (b & 7))) & 1u);
    }
    return v;
}

static int derive_packet_info(s1g_packet_info* info);

int s1g_parse_packet_header(const uint8_t* pk, int64_t nbytes, s1g_packet_info* info) {
    if (!pk || !info || nbytes < 68) return S1G_EARG;
    memset(info, 0, sizeof *info);
    int64_t bit = 0;
    for (int i = 0; i < 8; ++i) {
        info->primary[i] = (int64_t)be_field(pk, bit, kPrimW[i]);
        bit += kPrimW[i];
    }
    bit = 48;
    for (int i = 0; i < 46; ++i) {
        info->secondary[i] = (int64_t)be_field(pk, bit, kSecW[i]);
        bit += kSecW[i];
    }
    return derive_packet_info(info);
}

int s1g_packet_info_from_fields(const int64_t* primary, const int64_t* secondary, s1g_packet_info* info) {
    if (!primary || !secondary || !info) return S1G_EARG;
    memset(info, 0, sizeof *info);
    memcpy(info->primary, primary, sizeof info->primary);
    memcpy(info->secondary, secondary, sizeof info->secondary);
    return derive_packet_info(info);
}

// Quantities derived from the header fields (packet.cpp:29-100, 265-281).
static int derive_packet_info(s1g_packet_info* info) {
    const int64_t* s = info->secondary;
    info->num_quads = (int32_t)s[44];
    info->baq_mode = (int32_t)s[15];
    info->range_decimation = (int32_t)s[18];
    info->rank = (int32_t)s[26];
    info->swath_number = (int32_t)s[43];
    info->user_data_length = (int32_t)(info->primary[7] + 1 - 62);
    info->baq_block_length = (int32_t)(8 * (s[16] + 1));
    // data format (packet.cpp:265-274)
    const int baq = info->baq_mode, test = (int)s[6];
    if (baq == 0)
        info->data_format = (test % 3 == 0) ? 'A' : 'B';
    else if (baq >= 3 && baq <= 5)
        info->data_format = 'C';
    else if (baq >= 12 && baq <= 14)
        info->data_format = 'D';
    else
        return S1G_EARG;
    // physical values (packet.cpp:29-100, 277-281)
    info->pri_us = (double)s[27] / kFRef;
    info->pulse_length_us = (double)s[24] / kFRef;
    info->swl_us = (double)s[29] / kFRef;
    info->swst_us = (double)s[28] / kFRef;
    info->rx_gain_db = -0.5 * (double)s[19];
    info->txprr_mhz_us = (s[20] == 0 ? -1 : 1) * (double)s[21] * (pow(kFRef, 2) / 2097152);
    const int sgn = s[22] == 0 ? -1 : 1;
    info->txpsf_mhz = ((double)(sgn * (int)s[23]) * (kFRef / 16384)) + (info->txprr_mhz_us / (4 * kFRef));
    info->fs_mhz = (s[18] >= 0 && s[18] < 12) ? kRangeDecimation[s[18]] : 0.0;
    info->replica_chirp_length = (int32_t)floor(info->fs_mhz * info->pulse_length_us);
    return S1G_OK;
}

int s1g_slant_ranges(const s1g_packet_info* info, int num_ranges, double* out) {
    if (!info || !out) return S1G_EARG;
    const double c = 299792458.0;
    const double dts = (320 / (8 * kFRef)) * 1e-6;
    if (num_ranges <= 0) num_ranges = 2 * info->num_quads;
    const double nrs = 2 * info->num_quads;
    const double fs = info->fs_mhz * 1e+6;
    const double start = info->swst_us * 1e-6;
    const double pri = info->pri_us * 1e-6;
    const double rank = info->rank;
    const double near = rank * pri + start + dts;
    const double far = near + (nrs / fs);
    const double a = near * c / 2, b = far * c / 2;
    const double step = (b - a) / (num_ranges - 1);
    for (int i = 0; i < num_ranges; ++i) out[i] = a + i * step;
    return S1G_OK;
}

int s1g_timing_corrections(const s1g_packet_info* info, double* out) {
    if (!info || !out) return S1G_EARG;
    const int n = 2 * info->num_quads;
    const double fs = info->fs_mhz * 1e+6;
    const double swst = info->swst_us * 1e-6;
    const double rank = info->rank;
    const double pri = info->pri_us * 1e-6;
    const double dt = 1 / fs;
    for (int i = 0; i < n; ++i) out[i] = (swst + 1.439 * 1e-6) + ((swst + (rank * pri) - ((double)i * dt)) / 2);
    return S1G_OK;
}

int s1g_decode_fdbaq(s1g_ctx* c, const uint8_t* ud, const uint64_t* offs, const uint32_t* lens, const int32_t* nq,
                     int n_packets, int max_bytes, int max_quads, int prec, void* out, int64_t ld_out,
                     int32_t* status) {
    if (!c) return S1G_EARG;
    if (!valid_prec(prec)) return fail(c, S1G_EARG, "precision must be S1G_F32 or S1G_F64%s");
    if (n_packets < 0 || max_bytes < 0 || max_quads < 0) return fail(c, S1G_EARG, "negative size%s");
    if (n_packets == 0) return S1G_OK;
    if (!ud || !offs || !lens || !nq || !out || !status) return fail(c, S1G_EARG, "null pointer%s");
    if (ld_out < 2 * (int64_t)max_quads) return fail(c, S1G_EARG, "ld_out < 2 * max_num_quads%s");
    StageTimer tm(c, S1G_STAGE_DECODE);
    int rc;
    if (max_bytes <= kDecodeSplitBytes + 4096) {
        rc = launch_fdbaq_decode(c->stream, prec, ud, offs, lens, nq, n_packets, max_bytes, max_quads, out, ld_out,
                                 status);
    } else {
        // A workgroup's LDS is sized for the launch's largest packet, which
        // sets how many packets a CU holds.  With a few long packets (RFI
        // lines: ~44 KB of mostly BRC-4 codes against ~19 KB) the packets up
        // to kDecodeSplitBytes decode with LDS for that size, and the longer
        // ones at once on a second stream (default priority), so both
        // classes are resident together; the context stream waits for both.
        if (!c->dec_side) HIPCHK(c, hipStreamCreateWithFlags(&c->dec_side, hipStreamNonBlocking));
        if (!c->dec_fork) HIPCHK(c, hipEventCreateWithFlags(&c->dec_fork, hipEventDisableTiming));
        if (!c->dec_join) HIPCHK(c, hipEventCreateWithFlags(&c->dec_join, hipEventDisableTiming));
        HIPCHK(c, hipEventRecord(c->dec_fork, c->stream));
        HIPCHK(c, hipStreamWaitEvent(c->dec_side, c->dec_fork, 0));
        rc = launch_fdbaq_decode(c->dec_side, prec, ud, offs, lens, nq, n_packets, max_bytes, max_quads, out, ld_out,
                                 status, (uint32_t)kDecodeSplitBytes, 0xFFFFFFFFu);
        const int rc2 = launch_fdbaq_decode(c->stream, prec, ud, offs, lens, nq, n_packets, kDecodeSplitBytes,
                                            max_quads, out, ld_out, status, 0u, (uint32_t)kDecodeSplitBytes);
        // join even after a failed launch, so the side stream never runs on
        // unordered with later work of the context stream
        HIPCHK(c, hipEventRecord(c->dec_join, c->dec_side));
        HIPCHK(c, hipStreamWaitEvent(c->stream, c->dec_join, 0));
        if (!rc) rc = rc2;
    }
    if (rc == -1) return fail(c, S1G_EUNSUPPORTED, "packet larger than the on-chip decode buffer%s");
    if (rc) return hip_fail(c, (hipError_t)rc, "fdbaq decode");
    return S1G_OK;
}

int s1g_check_status(s1g_ctx* c, const int32_t* status, int n_packets, int* first_failed) {
    if (!c) return S1G_EARG;
    if (first_failed) *first_failed = -1;
    if (n_packets < 0 || (n_packets && !status)) return fail(c, S1G_EARG, "bad status array%s");
    if (n_packets == 0) return S1G_OK;
    std::vector<int32_t> h((size_t)n_packets);
    HIPCHK(c, hipMemcpyAsync(h.data(), status, sizeof(int32_t) * (size_t)n_packets, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    static const char* const what[6] = {"ok", "BRC value above 4", "read past the end of the user data",
                                        "Huffman code not found", "magnitude code above the BRC flag",
                                        "packet above the stated maxima"};
    for (int i = 0; i < n_packets; ++i) {
        if (h[i] == S1G_PKT_OK) continue;
        if (first_failed) *first_failed = i;
        char buf[160];
        snprintf(buf, sizeof buf, "packet %d: %s (status %d)", i, what[std::min(std::max(h[i], 0), 5)], h[i]);
        c->err = buf;
        return S1G_EDECODE;
    }
    return S1G_OK;
}

int s1g_range_compress(s1g_ctx* c, const s1g_range_desc* d, int prec, const void* in, int64_t ld_in, void* out,
                       int64_t ld_out) {
    if (!c || !d) return S1G_EARG;
    if (!valid_prec(prec)) return fail(c, S1G_EARG, "precision must be S1G_F32 or S1G_F64%s");
    if (!in || !out || in == out) return fail(c, S1G_EARG, "range compression is out-of-place%s");
    if (ld_in < d->n_samples || ld_out < d->n_samples) return fail(c, S1G_EARG, "leading dimension too small%s");
    return do_range(c, *d, prec, in, ld_in, out, ld_out);
}

int s1g_azimuth_fft(s1g_ctx* c, int prec, int L, int N, void* data, int64_t ld, int inverse) {
    if (!c) return S1G_EARG;
    if (!valid_prec(prec)) return fail(c, S1G_EARG, "precision must be S1G_F32 or S1G_F64%s");
    if (L < 0 || N < 0 || ld < N || (!data && L * N)) return fail(c, S1G_EARG, "bad azimuth fft arguments%s");
    return do_azimuth_fft(c, prec, L, N, data, ld, inverse);
}

int s1g_rcmc_matched_filter(s1g_ctx* c, const s1g_azimuth_desc* d, int prec, const void* in, int64_t ld_in,
                            void* out, int64_t ld_out) {
    if (!c || !d) return S1G_EARG;
    if (!valid_prec(prec)) return fail(c, S1G_EARG, "precision must be S1G_F32 or S1G_F64%s");
    if (!in || !out || in == out) return fail(c, S1G_EARG, "rcmc/matched filter is out-of-place%s");
    if (ld_in < d->n_samples || ld_out < d->n_samples) return fail(c, S1G_EARG, "leading dimension too small%s");
    return do_rcmc(c, *d, prec, in, ld_in, out, ld_out);
}

int s1g_eccm(s1g_ctx* c, const s1g_eccm_desc* d, int prec, void* data, int64_t ld, uint8_t* flags) {
    if (!c || !d) return S1G_EARG;
    if (!valid_prec(prec)) return fail(c, S1G_EARG, "precision must be S1G_F32 or S1G_F64%s");
    if (!data || ld < d->n_samples) return fail(c, S1G_EARG, "bad ECCM buffer%s");
    if (d->n_lines <= 0 || d->n_samples <= 0) return S1G_OK;
    const size_t row = esize(prec) * (size_t)d->n_samples;
    int rc = ensure(c, &c->ws, &c->ws_bytes, row * d->n_lines);
    if (rc) return rc;
    // in place: flagged lines are staged in the workspace and notched back
    // into data; unflagged lines are left as they are
    return do_eccm(c, *d, prec, data, ld, data, ld, flags, c->ws, d->n_samples);
}

// ------------------------------------------------------------ general DSP
// Rows x n DFT along rows, in place: when n factors over the pass radices,
// one workgroup per row in LDS if the row fits, else global Stockham passes;
// otherwise Bluestein through a power-of-two length.  The global forms go in
// batches of rows that keep the work arrays within kDftWorkBudget bytes.
constexpr size_t kDftWorkBudget = size_t(4) << 30;

static int dft_rows(s1g_ctx* c, int prec, int rows, int n, void* data, int64_t ld, int inverse, bool normalise = true) {
    if (rows <= 0 || n <= 0) return 0;
    const DevFft* f;
    int rc = get_fft(c, n, &f);
    if (rc) return rc;
    const size_t es = esize(prec);
    const double scale = inverse && normalise ? 1.0 / (double)n : 1.0;
    if (f->plan.npass >= 0 && fft_rows_lds_fits(prec, n)) {
        // whole rows in LDS: one read and one write per row; the radix order
        // is the tuner's (its trials transform up to 2,048 rows of the
        // scratch buffer, whatever it holds, so the caller's data is
        // transformed once)
        FftPlan ord[4];
        const int nvar = radix_orders(f->plan, ord, 4);
        const int trial_rows = std::min(rows, 2048);
        auto run = [&](int v) {
            const DevFft* fv;
            int r = get_fft_order(c, n, v, &fv);
            if (!r) r = ensure(c, &c->dft_ws, &c->dft_ws_bytes, es * (size_t)trial_rows * n);
            if (r) return r;
            return check_launch(c, launch_fft_rows_lds(c->stream, prec, *fv, c->dft_ws, n, trial_rows, inverse, scale),
                                "dft");
        };
        const int v = choose_variant(c, TuneKey{S1G_STAGE_DFT, prec, n, inverse, 0}, nvar, run);
        const DevFft* fv;
        rc = get_fft_order(c, n, v, &fv);
        if (rc) return rc;
        rc = launch_fft_rows_lds(c->stream, prec, *fv, data, ld, rows, inverse, scale);
        return check_launch(c, rc, "dft");
    }
    if (f->plan.npass >= 0) {
        const int chunk = (int)std::max<size_t>(1, std::min<size_t>(rows, kDftWorkBudget / (2 * es * (size_t)ld)));
        rc = ensure(c, &c->dft_ws, &c->dft_ws_bytes, 2 * es * (size_t)chunk * ld);
        if (rc) return rc;
        char* w1 = (char*)c->dft_ws;
        char* w2 = w1 + es * (size_t)chunk * ld;
        for (int r0 = 0; r0 < rows && !rc; r0 += chunk) {
            const int nr = std::min(chunk, rows - r0);
            char* x = (char*)data + es * (size_t)ld * r0;
            rc = launch_fft_global(c->stream, prec, *f, x, w1, w2, nr, ld, inverse);
            if (!rc) rc = launch_copy_scale(c->stream, prec, w1, ld, x, ld, nr, n, scale);
        }
        return check_launch(c, rc, "dft");
    }
    int M = 1;
    while (M < 2 * n - 1) M <<= 1;
    const DevFft* fM;
    rc = get_plan(c, M, &fM);
    if (rc) return rc;
    auto key = std::make_tuple(n, prec, inverse ? 1 : 0);
    auto it = c->blue_filters.find(key);
    void* bhat;
    if (it != c->blue_filters.end()) {
        bhat = it->second;
    } else {
        rc = ensure(c, &c->blue_scratch, &c->blue_scratch_bytes, sizeof(cpx<double>) * 3 * (size_t)M);
        if (rc) return rc;
        HIPCHK(c, hipMalloc(&bhat, es * M));
        rc = launch_bluestein_filter(c->stream, prec, *fM, n, inverse, c->blue_scratch, bhat);
        if (rc) {
            hipFree(bhat);
            return check_launch(c, rc, "dft filter");
        }
        c->blue_filters[key] = bhat;
    }
    const int chunk = (int)std::max<size_t>(1, std::min<size_t>(rows, kDftWorkBudget / (3 * es * (size_t)M)));
    rc = ensure(c, &c->dft_ws, &c->dft_ws_bytes, 3 * es * (size_t)chunk * M);
    if (rc) return rc;
    char* w1 = (char*)c->dft_ws;
    char* w2 = w1 + es * (size_t)chunk * M;
    char* w3 = w2 + es * (size_t)chunk * M;
    for (int r0 = 0; r0 < rows && !rc; r0 += chunk) {
        const int nr = std::min(chunk, rows - r0);
        rc = launch_bluestein(c->stream, prec, *fM, bhat, (char*)data + es * (size_t)ld * r0, ld, nr, n, inverse,
                              scale, w1, w2, w3);
    }
    return check_launch(c, rc, "dft");
}

// Along columns: the LDS column kernel when the length has a plan, else a
// transpose to rows and back.
static int dft_cols(s1g_ctx* c, int prec, int rows, int cols, void* data, int64_t ld, int inverse) {
    if (rows <= 0 || cols <= 0) return 0;
    const DevFft* fl;
    int rc = get_fft(c, rows, &fl);
    if (rc) return rc;
    if (fl->plan.npass >= 0) {
        rc = launch_azimuth_fft(c->stream, prec, *fl, data, ld, rows, cols, inverse, 0, 0);
        if (rc != -3) return check_launch(c, rc, "dft");
    }
    void* t = nullptr;
    const size_t es = esize(prec);
    HIPCHK(c, hipMalloc(&t, es * (size_t)rows * cols));
    rc = launch_transpose(c->stream, prec, data, ld, t, rows, rows, cols);
    if (!rc) rc = dft_rows(c, prec, cols, rows, t, rows, inverse);
    if (!rc) rc = launch_transpose(c->stream, prec, t, rows, data, ld, cols, rows);
    (void)hipStreamSynchronize(c->stream);
    hipFree(t);
    return rc > 0 ? check_launch(c, rc, "dft") : rc;
}

extern "C++" {
namespace s1g {
int ctx_dft_rows(s1g_ctx* c, int prec, int rows, int n, void* data, int64_t ld, int inverse, bool normalise) {
    return dft_rows(c, prec, rows, n, data, ld, inverse, normalise);
}
}  // namespace s1g
}

extern "C" int s1g_dft(s1g_ctx* c, int prec, int n_rows, int n_cols, void* data, int64_t ld, int axis, int inverse) {
    if (!c) return S1G_EARG;
    if (!valid_prec(prec)) return fail(c, S1G_EARG, "precision must be S1G_F32 or S1G_F64%s");
    if (n_rows < 0 || n_cols < 0 || (axis != 0 && axis != 1)) return fail(c, S1G_EARG, "bad DFT shape or axis%s");
    if (!data || ld < n_cols) return fail(c, S1G_EARG, "bad DFT buffer%s");
    return axis == 1 ? dft_rows(c, prec, n_rows, n_cols, data, ld, inverse != 0)
                     : dft_cols(c, prec, n_rows, n_cols, data, ld, inverse != 0);
}

struct s1g_plan {
    s1g_ctx* ctx;
    int prec, n, inverse;
};

extern "C" int s1g_plan_create(s1g_ctx* c, int prec, int n, int inverse, s1g_plan** plan) {
    if (!c || !plan) return S1G_EARG;
    *plan = nullptr;
    if (!valid_prec(prec)) return fail(c, S1G_EARG, "precision must be S1G_F32 or S1G_F64%s");
    if (n <= 0) return fail(c, S1G_EARG, "plan length must be positive%s");
    // build the twiddle tables (and the Bluestein filter) now, as FFTW
    // plans do their work at creation
    const DevFft* f;
    int rc = get_fft(c, n, &f);
    if (rc) return rc;
    s1g_plan* p = new (std::nothrow) s1g_plan{c, prec, n, inverse ? 1 : 0};
    if (!p) return fail(c, S1G_ENOMEM, "plan allocation failed%s");
    *plan = p;
    return S1G_OK;
}

extern "C" int s1g_plan_execute(s1g_plan* p, void* data, int64_t ld, int n_rows) {
    if (!p) return S1G_EARG;
    if (n_rows < 0 || !data || ld < p->n) return fail(p->ctx, S1G_EARG, "bad plan execution arguments%s");
    return dft_rows(p->ctx, p->prec, n_rows, p->n, data, ld, p->inverse, false);
}

extern "C" int s1g_plan_destroy(s1g_plan* p) {
    delete p;
    return S1G_OK;
}

extern "C" int s1g_transpose(s1g_ctx* c, int prec, int n_rows, int n_cols, const void* src, int64_t ld_src, void* dst,
                             int64_t ld_dst) {
    if (!c) return S1G_EARG;
    if (!valid_prec(prec)) return fail(c, S1G_EARG, "precision must be S1G_F32 or S1G_F64%s");
    if (n_rows < 0 || n_cols < 0 || !src || !dst || src == dst || ld_src < n_cols || ld_dst < n_rows)
        return fail(c, S1G_EARG, "bad transpose arguments%s");
    return check_launch(c, launch_transpose(c->stream, prec, src, ld_src, dst, ld_dst, n_rows, n_cols), "transpose");
}

extern "C" int s1g_multiply(s1g_ctx* c, int prec, int n_rows, int n, void* data, int64_t ld, const void* vec,
                            int conj_vec) {
    if (!c) return S1G_EARG;
    if (!valid_prec(prec)) return fail(c, S1G_EARG, "precision must be S1G_F32 or S1G_F64%s");
    if (n_rows < 0 || n < 0 || !data || !vec || ld < n) return fail(c, S1G_EARG, "bad multiply arguments%s");
    return check_launch(c, launch_mul_rows(c->stream, prec, data, ld, vec, n_rows, n, conj_vec != 0), "multiply");
}

extern "C" int s1g_reference_function(s1g_ctx* c, int prec, const void* chirp, int n, void* out) {
    if (!c) return S1G_EARG;
    if (!valid_prec(prec)) return fail(c, S1G_EARG, "precision must be S1G_F32 or S1G_F64%s");
    if (n <= 0 || !chirp || !out) return fail(c, S1G_EARG, "bad reference-function arguments%s");
    if (!c->dsp_scal) HIPCHK(c, hipMalloc((void**)&c->dsp_scal, 4 * sizeof(double)));
    int rc = launch_chirp_energy(c->stream, prec, chirp, n, c->dsp_scal);
    if (rc) return check_launch(c, rc, "reference function");
    if (out != chirp) HIPCHK(c, hipMemcpyAsync(out, chirp, esize(prec) * n, hipMemcpyDeviceToDevice, c->stream));
    rc = dft_rows(c, prec, 1, n, out, n, 0);
    if (rc) return rc;
    return check_launch(c, launch_conj_div(c->stream, prec, out, n, c->dsp_scal), "reference function");
}

extern "C" int s1g_replica_chirp(s1g_ctx* c, int prec, int n_samples, double txpsf_mhz, double txprr_mhz_us,
                                 double txpl_us, double fs_mhz, void* out) {
    if (!c) return S1G_EARG;
    if (!valid_prec(prec)) return fail(c, S1G_EARG, "precision must be S1G_F32 or S1G_F64%s");
    const int nc = (int)floor(fs_mhz * txpl_us);
    if (n_samples < 0 || nc < 2 || !out) return fail(c, S1G_EARG, "bad replica-chirp arguments%s");
    return check_launch(c, launch_replica_chirp(c->stream, prec, out, n_samples + 2 * nc, nc, txpsf_mhz, txprr_mhz_us,
                                                txpl_us),
                        "replica chirp");
}

extern "C" int s1g_sinc_interpolate(s1g_ctx* c, int prec, const void* signal, int len, const double* t, int m,
                                    int half_width, void* out) {
    if (!c) return S1G_EARG;
    if (!valid_prec(prec)) return fail(c, S1G_EARG, "precision must be S1G_F32 or S1G_F64%s");
    if (len < 0 || m < 0 || half_width < 1 || (m && (!signal || !t || !out)))
        return fail(c, S1G_EARG, "bad sinc-interpolation arguments%s");
    return check_launch(c, launch_sinc_interp(c->stream, prec, signal, len, t, m, half_width, out),
                        "sinc interpolation");
}

// A burst's chain in two parts: decode (with ECCM, through the workspace)
// into b.out, then range compression onward (workspace and b.out).
static int burst_decode(s1g_ctx* c, const s1g_burst_desc& b, int prec) {
    const int L = b.range.n_lines, N = b.range.n_samples;
    if (L <= 1 || N <= 0) return fail(c, S1G_EARG, "burst needs >= 2 lines%s");
    if (b.azimuth.n_lines != L || b.azimuth.n_samples != N) return fail(c, S1G_EARG, "descriptor shapes disagree%s");
    if (b.ld_out < N || !b.out) return fail(c, S1G_EARG, "bad output buffer%s");
    int rc = ensure(c, &c->ws, &c->ws_bytes, esize(prec) * (size_t)L * N);
    if (rc) return rc;
    if (b.use_eccm) {
        // decode -> out; ECCM on out's rows, its flagged lines staged in the
        // workspace (unflagged lines stay as decoded)
        if (b.eccm.n_lines != L || b.eccm.n_samples != N) return fail(c, S1G_EARG, "ECCM shape disagrees%s");
        rc = s1g_decode_fdbaq(c, b.user_data, b.ud_offsets, b.ud_lengths, b.num_quads, L, b.max_user_data_bytes,
                              b.max_num_quads, prec, b.out, b.ld_out, b.status);
        if (rc) return rc;
        rc = do_eccm(c, b.eccm, prec, b.out, b.ld_out, b.out, b.ld_out, b.eccm_flags, c->ws, N);
    } else {
        rc = s1g_decode_fdbaq(c, b.user_data, b.ud_offsets, b.ud_lengths, b.num_quads, L, b.max_user_data_bytes,
                              b.max_num_quads, prec, b.out, b.ld_out, b.status);
    }
    return rc;
}

static int burst_focus(s1g_ctx* c, const s1g_burst_desc& b, int prec) {
    const int L = b.range.n_lines, N = b.range.n_samples;
    int rc = do_range(c, b.range, prec, b.out, b.ld_out, c->ws, N);
    if (rc) return rc;
    rc = do_azimuth_fft(c, prec, L, N, c->ws, N, 0);
    if (rc) return rc;
    rc = do_rcmc(c, b.azimuth, prec, c->ws, N, b.out, b.ld_out);
    if (rc) return rc;
    return do_azimuth_fft(c, prec, L, N, b.out, b.ld_out, 1);
}

static int burst_chain(s1g_ctx* c, const s1g_burst_desc& b, int prec) {
    int rc = burst_decode(c, b, prec);
    return rc ? rc : burst_focus(c, b, prec);
}

// Bytes [lo, hi) a burst's decode and chain write: the SLC rows and the
// packet status words.
static void burst_extent(const s1g_burst_desc& b, int prec, uintptr_t r[2][2]) {
    const size_t L = (size_t)std::max(b.range.n_lines, 0), N = (size_t)std::max(b.range.n_samples, 0);
    r[0][0] = (uintptr_t)b.out;
    r[0][1] = r[0][0] + (L ? esize(prec) * ((L - 1) * (size_t)b.ld_out + N) : 0);
    r[1][0] = (uintptr_t)b.status;
    r[1][1] = r[1][0] + sizeof(int32_t) * L;
}

// A batch may decode burst i + 1 beside the focusing of burst i when no
// decode goes through the shared workspace (no ECCM) and no burst writes
// memory another one writes.
static bool batch_can_overlap(const s1g_burst_desc* d, int n, int prec) {
    if (n < 2) return false;
    std::vector<std::array<uintptr_t, 2>> ext;
    for (int i = 0; i < n; ++i) {
        if (d[i].use_eccm) return false;
        uintptr_t r[2][2];
        burst_extent(d[i], prec, r);
        ext.push_back({r[0][0], r[0][1]});
        ext.push_back({r[1][0], r[1][1]});
    }
    for (size_t i = 0; i < ext.size(); ++i)
        for (size_t j = i + 1; j < ext.size(); ++j)
            if (ext[i][0] < ext[j][1] && ext[j][0] < ext[i][1]) return false;
    return true;
}

static int sync_event(s1g_ctx* c, size_t i, hipEvent_t* e) {
    while (c->sync_events.size() <= i) {
        hipEvent_t ev = nullptr;
        HIPCHK(c, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        c->sync_events.push_back(ev);
    }
    *e = c->sync_events[i];
    return S1G_OK;
}

// The bursts' chains in order on c->stream.  When batch_can_overlap holds,
// the decodes run on a second stream (default priority), each focusing
// waiting on its own burst's decode only, so the decode of burst i + 1
// (latency-bound, a few waves per SIMD) runs beside the focusing of burst i.
// Both streams order after the work already on c->stream, and c->stream
// orders after the last decode again before returning, so the batch stays
// one ordered unit on c->stream (and one graph when captured).
static int batch_chains(s1g_ctx* c, const s1g_burst_desc* d, int n, int prec) {
    if (!batch_can_overlap(d, n, prec)) {
        for (int i = 0; i < n; ++i) {
            int rc = burst_chain(c, d[i], prec);
            if (rc) return rc;
        }
        return S1G_OK;
    }
    if (!c->side) HIPCHK(c, hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    hipEvent_t fork;
    int rc = sync_event(c, 0, &fork);
    if (rc) return rc;
    hipStream_t main_s = c->stream;
    HIPCHK(c, hipEventRecord(fork, main_s));
    HIPCHK(c, hipStreamWaitEvent(c->side, fork, 0));
    for (int i = 0; i < n && !rc; ++i) {
        hipEvent_t dec;
        rc = sync_event(c, 1 + (size_t)i, &dec);
        if (rc) break;
        c->stream = c->side;
        rc = burst_decode(c, d[i], prec);
        c->stream = main_s;
        if (rc) break;
        HIPCHK(c, hipEventRecord(dec, c->side));
        HIPCHK(c, hipStreamWaitEvent(main_s, dec, 0));
        rc = burst_focus(c, d[i], prec);
    }
    c->stream = main_s;
    if (rc) {
        // the side stream rejoins c->stream on the error path too
        hipEvent_t join;
        if (!sync_event(c, (size_t)n + 1, &join) && hipEventRecord(join, c->side) == hipSuccess)
            (void)hipStreamWaitEvent(main_s, join, 0);
        return rc;
    }
    return S1G_OK;  // c->stream already waits on the last decode
}

// Bytes of burst i's workspace slice and column table in a stage-ordered
// batch (256-byte aligned).
static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
static void staged_layout(const s1g_burst_desc* d, int n, int prec, std::vector<size_t>& ws_off,
                          std::vector<size_t>& col_off, size_t* ws_total, size_t* col_total) {
    ws_off.resize(n);
    col_off.resize(n);
    size_t w = 0, k = 0;
    for (int i = 0; i < n; ++i) {
        ws_off[i] = w;
        col_off[i] = k;
        w += align256(esize(prec) * (size_t)d[i].range.n_lines * (size_t)d[i].range.n_samples);
        k += align256(az_table_bytes(d[i].range.n_samples));
    }
    *ws_total = w;
    *col_total = k;
}

// Stage-ordered batch (S1G_BATCH_STAGED): every burst's decode, then each
// later stage as one launch over all bursts -- the persistent,
// descriptor-driven kernels take a table of instances (range compression
// per block length, the azimuth transforms per azimuth length and variant,
// RCMC + matched filter per variant); a stage whose chosen variant has no
// batched form runs burst by burst in its place.  Each burst keeps its own workspace slice (c->ws holds
// them all) and column table (c->cols).  Callers check batch_can_overlap
// (no ECCM, disjoint outputs) and size c->ws and c->cols by staged_layout.
static int batch_staged(s1g_ctx* c, const s1g_burst_desc* d, int n, int prec) {
    std::vector<size_t> wso, colo;
    size_t ws_total, col_total;
    staged_layout(d, n, prec, wso, colo, &ws_total, &col_total);
    if (c->ws_bytes < ws_total || c->cols_n < col_total) return fail(c, S1G_EARG, "batch workspace not prepared%s");
    auto ws_of = [&](int i) { return (void*)((char*)c->ws + wso[i]); };
    auto cols_of = [&](int i) { return (AzColumn*)((char*)c->cols + colo[i]); };
    for (int i = 0; i < n; ++i) {
        const int L = d[i].range.n_lines, N = d[i].range.n_samples;
        if (L <= 1 || N <= 0) return fail(c, S1G_EARG, "burst needs >= 2 lines%s");
        if (d[i].azimuth.n_lines != L || d[i].azimuth.n_samples != N)
            return fail(c, S1G_EARG, "descriptor shapes disagree%s");
        if (d[i].ld_out < N || !d[i].out) return fail(c, S1G_EARG, "bad output buffer%s");
    }
    int rc = 0;
    // decode, burst by burst (one workgroup per packet already fills the chip)
    for (int i = 0; i < n && !rc; ++i) rc = burst_decode(c, d[i], prec);
    if (rc) return rc;
    // range compression: out -> workspace, grouped by block length
    {
        std::vector<const DevFft*> fm(n);
        std::vector<RangeInst> inst(n);
        for (int i = 0; i < n && !rc; ++i) {
            void* h;
            int nc;
            const int N = d[i].range.n_samples;
            rc = range_setup(c, d[i].range, prec, d[i].out, d[i].ld_out, ws_of(i), N, &fm[i], &h, &nc);
            if (!rc) inst[i] = range_inst(d[i].out, d[i].ld_out, ws_of(i), N, d[i].range.n_lines, N, fm[i]->plan.n, nc, h);
        }
        if (rc) return rc;
        StageTimer tm(c, S1G_STAGE_RANGE);
        std::vector<bool> done(n, false);
        for (int i = 0; i < n && !rc; ++i) {
            if (done[i]) continue;
            std::vector<RangeInst> group;
            for (int k = i; k < n; ++k)
                if (!done[k] && fm[k] == fm[i]) {
                    group.push_back(inst[k]);
                    done[k] = true;
                }
            if (range_batchable(prec, fm[i]->plan.n)) {
                rc = check_launch(c, launch_range_os_batch(c->stream, prec, *fm[i], group.data(), (int)group.size()),
                                  "range compression");
            } else {
                for (const RangeInst& r : group) {
                    if (rc) break;
                    const int nc = fm[i]->plan.n - r.V + 1;
                    rc = check_launch(c, launch_range_os(c->stream, prec, *fm[i], r.in, r.ld_in, r.out, r.ld_out,
                                                         r.n_lines, r.N, nc, r.h),
                                      "range compression");
                }
            }
        }
        if (rc) return rc;
    }
    // the azimuth transforms (in place), grouped by length and variant
    auto azimuth = [&](int inverse, auto data_of, auto ld_of) -> int {
        std::vector<const DevFft*> fl(n);
        std::vector<int> var(n);
        int r = 0;
        for (int i = 0; i < n && !r; ++i)
            r = az_setup(c, prec, d[i].range.n_lines, d[i].range.n_samples, inverse, &fl[i], &var[i]);
        if (r) return r;
        StageTimer tm(c, inverse ? S1G_STAGE_AZ_IFFT : S1G_STAGE_AZ_FFT);
        std::vector<bool> done(n, false);
        for (int i = 0; i < n && !r; ++i) {
            if (done[i]) continue;
            std::vector<AzInst> group;
            for (int k = i; k < n; ++k)
                if (!done[k] && fl[k] == fl[i] && var[k] == var[i]) {
                    AzInst a;
                    a.data = data_of(k);
                    a.ld = ld_of(k);
                    a.N = d[k].range.n_samples;
                    a.panel0 = 0;
                    group.push_back(a);
                    done[k] = true;
                }
            if (azimuth_batchable(*fl[i], var[i])) {
                r = check_launch(c, launch_azimuth_fft_batch(c->stream, prec, *fl[i], group.data(), (int)group.size(),
                                                             inverse, var[i]),
                                 "azimuth fft");
            } else {
                for (const AzInst& a : group) {
                    if (r) break;
                    r = check_launch(c, launch_azimuth_fft(c->stream, prec, *fl[i], a.data, a.ld, fl[i]->plan.n, a.N,
                                                           inverse, var[i]),
                                     "azimuth fft");
                }
            }
        }
        return r;
    };
    rc = azimuth(0, [&](int i) { return ws_of(i); }, [&](int i) { return (int64_t)d[i].range.n_samples; });
    if (rc) return rc;
    // RCMC + matched filter: workspace -> out, grouped by variant
    {
        std::vector<int> var(n);
        for (int i = 0; i < n && !rc; ++i)
            rc = rcmc_setup(c, d[i].azimuth, prec, cols_of(i), ws_of(i), d[i].range.n_samples, d[i].out, d[i].ld_out,
                            &var[i]);
        if (rc) return rc;
        StageTimer tm(c, S1G_STAGE_RCMC_MF);
        std::vector<bool> done(n, false);
        for (int i = 0; i < n && !rc; ++i) {
            if (done[i]) continue;
            std::vector<int> group;
            for (int k = i; k < n; ++k)
                if (!done[k] && var[k] == var[i]) {
                    group.push_back(k);
                    done[k] = true;
                }
            if (rcmc_batchable(prec, var[i])) {
                std::vector<RcmcInst> inst;
                for (int k : group) {
                    const s1g_azimuth_desc& a = d[k].azimuth;
                    RcmcInst r;
                    r.in = ws_of(k);
                    r.out = d[k].out;
                    r.ld_in = a.n_samples;
                    r.ld_out = d[k].ld_out;
                    r.cols = cols_of(k);
                    r.uni = az_uniform_of(cols_of(k), a.n_samples);
                    r.amp32 = az_amp_of(cols_of(k), a.n_samples);
                    r.prf = a.prf;
                    r.n_lines = a.n_lines;
                    r.N = a.n_samples;
                    r.apply_rcmc = a.apply_rcmc;
                    r.nbx = 0;
                    r.wg0 = 0;
                    inst.push_back(r);
                }
                rc = check_launch(c, launch_rcmc_batch(c->stream, inst.data(), (int)inst.size(), var[i]),
                                  "rcmc/matched filter");
            } else {
                for (int k : group) {
                    if (rc) break;
                    rc = check_launch(c, launch_rcmc_mf(c->stream, prec, d[k].azimuth, cols_of(k), ws_of(k),
                                                        d[k].range.n_samples, d[k].out, d[k].ld_out, var[i]),
                                      "rcmc/matched filter");
                }
            }
        }
        if (rc) return rc;
    }
    return azimuth(1, [&](int i) { return d[i].out; }, [&](int i) { return d[i].ld_out; });
}

int s1g_focus_burst(s1g_ctx* c, const s1g_burst_desc* b, int prec) {
    if (!c || !b) return S1G_EARG;
    if (!valid_prec(prec)) return fail(c, S1G_EARG, "precision must be S1G_F32 or S1G_F64%s");
    return burst_chain(c, *b, prec);
}

// Everything a burst chain allocates or builds on the host (FFT plans, range
// filters, workspace, column table, ECCM scratch), done ahead of a capture.
static int prepare_burst(s1g_ctx* c, const s1g_burst_desc& b, int prec, size_t* ws_need, size_t* cols_need) {
    const int L = b.range.n_lines, N = b.range.n_samples;
    if (L <= 1 || N <= 0) return fail(c, S1G_EARG, "burst needs >= 2 lines%s");
    const int nc = (int)floor(b.range.fs_mhz * b.range.txpl_us);
    if (nc < 2) return fail(c, S1G_EARG, "replica chirp shorter than 2 samples%s");
    int cand[2];
    const int nvar = range_candidates(prec, N, nc, cand);
    if (nvar == 0) return fail(c, S1G_EUNSUPPORTED, "replica chirp longer than the overlap-save block%s");
    const DevFft *fm, *fl;
    int rc = get_plan(c, L, &fl);
    for (int v = 0; v < nvar && !rc; ++v) {
        rc = get_plan(c, cand[v], &fm);
        void* h;
        if (!rc) rc = get_filter(c, b.range, prec, nc, fm, &h);
    }
    if (!rc && b.use_eccm) {
        const DevFft* f;
        rc = get_fft(c, N, &f);
        if (!rc) rc = get_plan(c, b.eccm.fft_size, &f);
        const int k = eccm_detect_split(prec, N, factorise);
        if (!rc && k) {
            rc = get_plan(c, N / k, &f);
            if (!rc) rc = ensure(c, &c->eccm_scratch, &c->eccm_scratch_bytes, sizeof(double) * (size_t)L);
        }
    }
    *ws_need = std::max(*ws_need, esize(prec) * (size_t)L * N);
    *cols_need = std::max(*cols_need, az_table_bytes(N));
    return rc;
}

int s1g_focus_batch(s1g_ctx* c, const s1g_burst_desc* descs, int n, int prec, int flags) {
    if (!c || (!descs && n)) return S1G_EARG;
    if (!valid_prec(prec)) return fail(c, S1G_EARG, "precision must be S1G_F32 or S1G_F64%s");
    if (n == 0) return S1G_OK;
    size_t ws_need = 0, cols_need = 0;
    for (int i = 0; i < n; ++i) {
        int rc = prepare_burst(c, descs[i], prec, &ws_need, &cols_need);
        if (rc) return rc;
    }
    const bool staged = (flags & S1G_BATCH_STAGED) && batch_can_overlap(descs, n, prec);
    if (staged) {
        std::vector<size_t> wso, colo;
        size_t wt, ct;
        staged_layout(descs, n, prec, wso, colo, &wt, &ct);
        ws_need = std::max(ws_need, wt);
        cols_need = std::max(cols_need, ct);
    }
    int rc = ensure(c, &c->ws, &c->ws_bytes, ws_need);
    if (!rc) rc = ensure(c, (void**)&c->cols, &c->cols_n, cols_need);
    if (rc) return rc;
    auto run_batch = [&]() { return staged ? batch_staged(c, descs, n, prec) : batch_chains(c, descs, n, prec); };
    if (!(flags & S1G_BATCH_GRAPH)) return run_batch();
    // one hipGraph per distinct descriptor set (and forced variants and mode)
    std::string key((const char*)descs, sizeof(s1g_burst_desc) * (size_t)n);
    key.append((const char*)&prec, sizeof prec);
    key.append((const char*)c->force, sizeof c->force);
    key.append(staged ? "S" : "B");
    auto it = c->graphs.find(key);
    if (it != c->graphs.end() && it->second.gen != c->ws_gen) {
        // captured before a workspace reallocation: its pointers are stale
        HIPCHK(c, hipStreamSynchronize(c->stream));
        hipGraphExecDestroy(it->second.exec);
        hipGraphDestroy(it->second.graph);
        c->graphs.erase(it);
        it = c->graphs.end();
    }
    if (it == c->graphs.end()) {
        // one eager pass first: it picks (and caches) each stage's launch
        // variant and builds everything the captured launches use
        rc = run_batch();
        if (rc) return rc;
        hipStream_t user = c->stream;
        HIPCHK(c, hipStreamSynchronize(user));  // filters built above are complete
        c->stream = c->own;
        c->capturing = true;
        hipError_t e = hipStreamBeginCapture(c->own, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) {
            c->capturing = false;
            c->stream = user;
            return hip_fail(c, e, "hipStreamBeginCapture");
        }
        rc = run_batch();
        hipGraph_t g = nullptr;
        e = hipStreamEndCapture(c->own, &g);
        c->capturing = false;
        c->stream = user;
        if (rc) {
            if (g) hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return hip_fail(c, e, "hipStreamEndCapture");
        hipGraphExec_t ge = nullptr;
        e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            hipGraphDestroy(g);
            return hip_fail(c, e, "hipGraphInstantiate");
        }
        s1g_ctx::CachedGraph cg;
        cg.exec = ge;
        cg.graph = g;
        cg.gen = c->ws_gen;
        it = c->graphs.emplace(key, cg).first;
    }
    HIPCHK(c, hipGraphLaunch(it->second.exec, c->stream));
    return S1G_OK;
}

static int af_run(s1g_ctx* c, int prec, int L, int N, const void* rd, int64_t ld, const double* phase, void* img,
                  double* entropy, double* grad) {
    if (!valid_prec(prec)) return fail(c, S1G_EARG, "precision must be S1G_F32 or S1G_F64%s");
    if (L < 2 || N < 1 || !rd || !phase || ld < N || !entropy) return fail(c, S1G_EARG, "bad autofocus arguments%s");
    const DevFft* fl;
    int rc = get_plan(c, L, &fl);
    if (rc) return rc;
    const bool fast = af_fast(L, N, ld, prec);
    const size_t work = sizeof(double) * (af_work_doubles(L, N, ld, prec) + 4);
    if (!fast) rc = ensure(c, &c->af_img, &c->af_img_bytes, esize(prec) * (size_t)L * N);
    if (!rc) rc = ensure(c, (void**)&c->af_scalars, &c->af_n, work);
    if (rc) return rc;
    double* scal = c->af_scalars;
    double* wk = c->af_scalars + 4;
    {
        StageTimer tm(c, S1G_STAGE_AUTOFOCUS);
        rc = launch_af(c->stream, prec, L, N, rd, ld, phase, *fl, fast ? img : nullptr, c->af_img, wk,
                       grad != nullptr, grad, scal);
    }
    if (rc) return check_launch(c, rc, "autofocus");
    if (img && !fast) {
        const size_t row = esize(prec) * (size_t)N;
        HIPCHK(c, hipMemcpy2DAsync(img, row, c->af_img, row, row, L, hipMemcpyDeviceToDevice, c->stream));
    }
    double h[3];
    HIPCHK(c, hipMemcpyAsync(h, scal, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *entropy = h[2];
    return S1G_OK;
}

int s1g_af_forward(s1g_ctx* c, int prec, int L, int N, const void* rd, int64_t ld, const double* phase, void* img,
                   double* entropy) {
    if (!c) return S1G_EARG;
    return af_run(c, prec, L, N, rd, ld, phase, img, entropy, nullptr);
}

int s1g_af_adjoint(s1g_ctx* c, int prec, int L, int N, const void* rd, int64_t ld, const double* phase,
                   double* entropy, double* grad) {
    if (!c) return S1G_EARG;
    if (!grad) return fail(c, S1G_EARG, "null gradient%s");
    return af_run(c, prec, L, N, rd, ld, phase, nullptr, entropy, grad);
}

}  // extern "C"
