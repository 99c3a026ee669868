// s1g_device.h -- device-side building blocks shared by the gfx950 kernels:
// complex arithmetic, compile-time trigonometry, radix-R DFT butterflies and
// the in-place mixed-radix DIT / DIF passes that run on LDS-resident data.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

namespace s1g {

// ---------------------------------------------------------------- complex
template <typename T>
struct alignas(2 * sizeof(T)) cpx {
    T x, y;
};

template <typename T>
__device__ __forceinline__ cpx<T> mk(T a, T b) {
    cpx<T> r;
    r.x = a;
    r.y = b;
    return r;
}
template <typename T>
__device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
    return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <typename T>
__device__ __forceinline__ cpx<T> cmulc(cpx<T> a, cpx<T> b) {
    return mk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename T>
__device__ __forceinline__ cpx<T> scl(cpx<T> a, T s) { return mk(a.x * s, a.y * s); }

// ------------------------------------------------------- buffer accesses
// Raw buffer loads and stores through a resource of `bytes` bytes at
// `base`: an offset at or beyond the range reads zero and drops the store,
// so a masked element needs no branch and every wave issues the same
// number of memory instructions (the waits on them can then be counted).
// Offsets are 32-bit: ranges below 2 GiB, with kBufSkip as the offset of a
// masked element.
constexpr uint32_t kBufSkip = 0x80000000u;
constexpr int kBufConfig = 0x00020000;  // raw buffer, 32-bit data format (gfx9 family)

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, kBufConfig);
}
typedef int s1g_v2i __attribute__((ext_vector_type(2)));
typedef int s1g_v4i __attribute__((ext_vector_type(4)));
template <typename T>
__device__ __forceinline__ cpx<T> buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off);
template <>
__device__ __forceinline__ cpx<float> buf_load<float>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    const s1g_v2i v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0);
    return mk(__int_as_float(v.x), __int_as_float(v.y));
}
template <>
__device__ __forceinline__ cpx<double> buf_load<double>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    const s1g_v4i v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
    return mk(__hiloint2double(v.y, v.x), __hiloint2double(v.w, v.z));
}
__device__ __forceinline__ void buf_store(cpx<float> a, __amdgpu_buffer_rsrc_t r, uint32_t off) {
    s1g_v2i v;
    v.x = __float_as_int(a.x);
    v.y = __float_as_int(a.y);
    __builtin_amdgcn_raw_buffer_store_b64(v, r, (int)off, 0, 0);
}
__device__ __forceinline__ void buf_store(cpx<double> a, __amdgpu_buffer_rsrc_t r, uint32_t off) {
    s1g_v4i v;
    v.x = __double2loint(a.x);
    v.y = __double2hiint(a.x);
    v.z = __double2loint(a.y);
    v.w = __double2hiint(a.y);
    __builtin_amdgcn_raw_buffer_store_b128(v, r, (int)off, 0, 0);
}

// ------------------------------------------------- compile-time trigonometry
// cos(2 pi num / den) and sin(2 pi num / den) for integers, evaluated at
// compile time: exact octant reduction on the rational angle, then Taylor
// series on |x| <= pi/4 (26 terms, beyond double precision).
constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double taylor_sin(double x) {
    double term = x, sum = x;
    for (int k = 1; k < 26; ++k) {
        term *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double taylor_cos(double x) {
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 26; ++k) {
        term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
        sum += term;
    }
    return sum;
}
// angle = num/den of a full turn, 0 <= num < den
constexpr double turn_cos(long num, long den);
constexpr double turn_sin(long num, long den) {
    num %= den;
    if (num < 0) num += den;
    if (2 * num > den) return -turn_sin(den - num, den);      // sin(2pi - a) = -sin a
    if (4 * num > den) return turn_sin(den - 2 * num, 2 * den);  // sin(pi - a) = sin a
    if (8 * num > den) return turn_cos(den - 4 * num, 4 * den);  // sin(pi/2 - a) = cos a
    return taylor_sin(kTwoPi * (double)num / (double)den);
}
constexpr double turn_cos(long num, long den) {
    num %= den;
    if (num < 0) num += den;
    if (2 * num > den) return turn_cos(den - num, den);
    if (4 * num > den) return -turn_cos(den - 2 * num, 2 * den);
    if (8 * num > den) return turn_sin(den - 4 * num, 4 * den);
    return taylor_cos(kTwoPi * (double)num / (double)den);
}

// ------------------------------------------------------------ butterflies
// In-register DFT of length R.  INV = false: X_k = sum x_m e^{-2 pi i mk/R};
// INV = true uses e^{+...}.  a[] is overwritten with the result.
template <typename T, int R, bool INV>
struct Dft;

template <typename T, bool INV>
struct Dft<T, 2, INV> {
    static __device__ __forceinline__ void run(cpx<T>* a) {
        cpx<T> t = a[0];
        a[0] = t + a[1];
        a[1] = t - a[1];
    }
};

template <typename T, bool INV>
struct Dft<T, 4, INV> {
    static __device__ __forceinline__ void run(cpx<T>* a) {
        cpx<T> s0 = a[0] + a[2], d0 = a[0] - a[2];
        cpx<T> s1 = a[1] + a[3], d1 = a[1] - a[3];
        // forward: X1 = d0 - i d1, X3 = d0 + i d1
        cpx<T> jd1 = INV ? mk(-d1.y, d1.x) : mk(d1.y, -d1.x);
        a[0] = s0 + s1;
        a[2] = s0 - s1;
        a[1] = d0 + jd1;
        a[3] = d0 - jd1;
    }
};

template <typename T, bool INV>
struct Dft<T, 8, INV> {
    static __device__ __forceinline__ void run(cpx<T>* a) {
        const T h = (T)0.70710678118654752440084436210485;
        cpx<T> e[4] = {a[0], a[2], a[4], a[6]};
        cpx<T> o[4] = {a[1], a[3], a[5], a[7]};
        Dft<T, 4, INV>::run(e);
        Dft<T, 4, INV>::run(o);
        // twiddles w8^k for k = 1, 2, 3
        cpx<T> o1 = INV ? mk(h * (o[1].x - o[1].y), h * (o[1].x + o[1].y))
                        : mk(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
        cpx<T> o2 = INV ? mk(-o[2].y, o[2].x) : mk(o[2].y, -o[2].x);
        cpx<T> o3 = INV ? mk(-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y))
                        : mk(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));
        a[0] = e[0] + o[0];
        a[4] = e[0] - o[0];
        a[1] = e[1] + o1;
        a[5] = e[1] - o1;
        a[2] = e[2] + o2;
        a[6] = e[2] - o2;
        a[3] = e[3] + o3;
        a[7] = e[3] - o3;
    }
};

// Odd R: pair outputs k and R-k,
//   X_k     = x0 + sum_m (x_m + x_{R-m}) cos(2 pi mk/R) -/+ i sum_m (x_m - x_{R-m}) sin(2 pi mk/R)
//   X_{R-k} = x0 + ...                                  +/- i ...
// The (k, m) loops are template recursions so every cos/sin is a
// compile-time constant.
template <typename T, int R, int K, int M>
__device__ __forceinline__ void odd_acc(const cpx<T>* s, const cpx<T>* d, T& ar, T& ai, T& br, T& bi) {
    if constexpr (M <= (R - 1) / 2) {
        constexpr T c = (T)turn_cos((long)M * K, R);
        constexpr T sn = (T)turn_sin((long)M * K, R);
        ar += s[M - 1].x * c;
        ai += s[M - 1].y * c;
        br += d[M - 1].x * sn;
        bi += d[M - 1].y * sn;
        odd_acc<T, R, K, M + 1>(s, d, ar, ai, br, bi);
    }
}

template <typename T, int R, bool INV, int K>
__device__ __forceinline__ void odd_out(cpx<T>* a, const cpx<T>* s, const cpx<T>* d, cpx<T> x0) {
    if constexpr (K <= (R - 1) / 2) {
        T ar = x0.x, ai = x0.y, br = 0, bi = 0;
        odd_acc<T, R, K, 1>(s, d, ar, ai, br, bi);
        // forward: X_k = A - i B ; inverse: X_k = A + i B
        if (INV) {
            a[K] = mk(ar - bi, ai + br);
            a[R - K] = mk(ar + bi, ai - br);
        } else {
            a[K] = mk(ar + bi, ai - br);
            a[R - K] = mk(ar - bi, ai + br);
        }
        odd_out<T, R, INV, K + 1>(a, s, d, x0);
    }
}

// Same pairs, handed to sink(k, X_k) as soon as they are formed, so the R
// outputs never need registers at the same time.
template <typename T, int R, bool INV, int K, typename Sink>
__device__ __forceinline__ void odd_out_sink(const cpx<T>* s, const cpx<T>* d, cpx<T> x0, Sink& sink) {
    if constexpr (K <= (R - 1) / 2) {
        T ar = x0.x, ai = x0.y, br = 0, bi = 0;
        odd_acc<T, R, K, 1>(s, d, ar, ai, br, bi);
        if (INV) {
            sink(K, mk(ar - bi, ai + br));
            sink(R - K, mk(ar + bi, ai - br));
        } else {
            sink(K, mk(ar + bi, ai - br));
            sink(R - K, mk(ar - bi, ai + br));
        }
        odd_out_sink<T, R, INV, K + 1>(s, d, x0, sink);
    }
}

// Pairs K0 .. K1 - 1 of odd_out_sink only.
template <typename T, int R, bool INV, int K, int K1, typename Sink>
__device__ __forceinline__ void odd_out_sink_range(const cpx<T>* s, const cpx<T>* d, cpx<T> x0, Sink& sink) {
    if constexpr (K < K1) {
        T ar = x0.x, ai = x0.y, br = 0, bi = 0;
        odd_acc<T, R, K, 1>(s, d, ar, ai, br, bi);
        if (INV) {
            sink(K, mk(ar - bi, ai + br));
            sink(R - K, mk(ar + bi, ai - br));
        } else {
            sink(K, mk(ar + bi, ai - br));
            sink(R - K, mk(ar - bi, ai + br));
        }
        odd_out_sink_range<T, R, INV, K + 1, K1>(s, d, x0, sink);
    }
}

template <typename T, int R, bool INV>
struct Dft {
    static_assert(R % 2 == 1, "generic butterfly handles odd lengths");
    static constexpr int H = (R - 1) / 2;
    // Part of the outputs: X_0 when DC, and the pairs (k, R - k) for
    // K0 <= k < K1 (two lanes can share one butterfly this way).
    template <int K0, int K1, bool DC, typename Sink>
    static __device__ __forceinline__ void run_sink_part(const cpx<T>* a, Sink& sink) {
        cpx<T> s[H], d[H];
        const cpx<T> x0 = a[0];
        cpx<T> sum = a[0];
#pragma unroll
        for (int m = 1; m <= H; ++m) {
            s[m - 1] = a[m] + a[R - m];
            d[m - 1] = a[m] - a[R - m];
            if constexpr (DC) sum = sum + s[m - 1];
        }
        if constexpr (DC) sink(0, sum);
        odd_out_sink_range<T, R, INV, K0, K1>(s, d, x0, sink);
    }
    template <typename Sink>
    static __device__ __forceinline__ void run_sink(const cpx<T>* a, Sink& sink) {
        cpx<T> s[H], d[H];
        const cpx<T> x0 = a[0];
        cpx<T> sum = a[0];
#pragma unroll
        for (int m = 1; m <= H; ++m) {
            s[m - 1] = a[m] + a[R - m];
            d[m - 1] = a[m] - a[R - m];
            sum = sum + s[m - 1];
        }
        sink(0, sum);
        odd_out_sink<T, R, INV, 1>(s, d, x0, sink);
    }
    static __device__ __forceinline__ void run(cpx<T>* a) {
        cpx<T> s[H], d[H];
        const cpx<T> x0 = a[0];
        cpx<T> sum = a[0];
#pragma unroll
        for (int m = 1; m <= H; ++m) {
            s[m - 1] = a[m] + a[R - m];
            d[m - 1] = a[m] - a[R - m];
            sum = sum + s[m - 1];
        }
        a[0] = sum;
        odd_out<T, R, INV, 1>(a, s, d, x0);
    }
};

// Multiply by the compile-time twiddle w_R^E (forward: e^{-2 pi i E/R};
// INV: e^{+2 pi i E/R}); trivial angles cost no multiplies.
template <typename T, int R, int E, bool INV>
__device__ __forceinline__ cpx<T> cwk(cpx<T> x) {
    constexpr int e = ((E % R) + R) % R;
    if constexpr (e == 0) {
        return x;
    } else if constexpr (2 * e == R) {
        return mk(-x.x, -x.y);
    } else if constexpr (4 * e == R) {
        return INV ? mk(-x.y, x.x) : mk(x.y, -x.x);
    } else if constexpr (4 * e == 3 * R) {
        return INV ? mk(x.y, -x.x) : mk(-x.y, x.x);
    } else {
        constexpr T c = (T)turn_cos(e, R);
        constexpr T sn = (T)(INV ? turn_sin(e, R) : -turn_sin(e, R));
        return mk(x.x * c - x.y * sn, x.x * sn + x.y * c);
    }
}

// Cooley-Tukey DFT of length R = A * B in registers: n = b + B a,
// k = k1 + A k2;  X[k1 + A k2] = sum_b w_B^{b k2} w_R^{b k1} DFT_A(x[b + B .])[k1].
// Index packs are compile-time, so every twiddle is a constant.
template <typename T, int A, int B, bool INV>
struct DftCT {
    static constexpr int R = A * B;
    template <int b, int... as>
    static __device__ __forceinline__ void inner(const cpx<T>* x, cpx<T>* y, std::integer_sequence<int, as...>) {
        cpx<T> t[A] = {x[b + B * as]...};
        Dft<T, A, INV>::run(t);
        ((y[b * A + as] = cwk<T, R, b * as, INV>(t[as])), ...);
    }
    template <int... bs>
    static __device__ __forceinline__ void inners(const cpx<T>* x, cpx<T>* y, std::integer_sequence<int, bs...>) {
        (inner<bs>(x, y, std::make_integer_sequence<int, A>{}), ...);
    }
    template <int k1, int... bs>
    static __device__ __forceinline__ void outer(const cpx<T>* y, cpx<T>* x, std::integer_sequence<int, bs...>) {
        cpx<T> t[B] = {y[bs * A + k1]...};
        Dft<T, B, INV>::run(t);
        ((x[k1 + A * bs] = t[bs]), ...);
    }
    template <int... k1s>
    static __device__ __forceinline__ void outers(const cpx<T>* y, cpx<T>* x, std::integer_sequence<int, k1s...>) {
        (outer<k1s>(y, x, std::make_integer_sequence<int, B>{}), ...);
    }
    static __device__ __forceinline__ void run(cpx<T>* x) {
        cpx<T> y[R];
        inners(x, y, std::make_integer_sequence<int, B>{});
        outers(y, x, std::make_integer_sequence<int, A>{});
    }
};

template <typename T, bool INV>
struct Dft<T, 16, INV> : DftCT<T, 4, 4, INV> {};
template <typename T, bool INV>
struct Dft<T, 32, INV> : DftCT<T, 8, 4, INV> {};

// q = a / b, r = a % b for 0 <= a < 2^22, 1 <= b < 2^22, from a float
// reciprocal: the product is within one of the quotient, then corrected.
__device__ __forceinline__ int fast_div(int a, int b, float inv_b, int& r) {
    int q = (int)((float)a * inv_b);
    r = a - q * b;
    if (r < 0) {
        --q;
        r += b;
    } else if (r >= b) {
        ++q;
        r -= b;
    }
    return q;
}

// Radices the LDS passes support; the host planner factors lengths over them.
#define S1G_RADIX_LIST(X) X(2) X(3) X(4) X(5) X(7) X(8) X(11) X(13) X(17) X(19) X(23) X(29) X(31) X(37) X(41) X(43) X(47) X(53) X(59) X(61) X(67) X(71)

// ----------------------------------------------------- LDS pass primitives
// Data layout in LDS: element e of vector v lives at buf[v * vstride + e * estride].
// Twiddle table tw[k] = e^{-2 pi i k / n} (k < n) in global memory.
//
// In-place decimation-in-time pass with radix R after Ns = product of the
// radices already applied: butterfly (g, j) combines elements
// g*Ns*R + j + r*Ns, r < R, multiplying input r by w_{Ns R}^{j r} first.
// The input of the first pass must be in mixed-radix digit-reversed order.
template <typename T, int R, bool INV>
__device__ __forceinline__ void dit_pass(cpx<T>* buf, int n, int Ns, int nvec, int vstride, int estride,
                                         const cpx<T>* __restrict__ tw, int tid, int nthr) {
    const int nbf = n / R;
    const int total = nbf * nvec;
    const int tstep = n / (Ns * R);
    const float inv_nvec = 1.0f / (float)nvec, inv_ns = 1.0f / (float)Ns;
    for (int idx = tid; idx < total; idx += nthr) {
        int v, j;
        const int bf = fast_div(idx, nvec, inv_nvec, v);
        const int g = fast_div(bf, Ns, inv_ns, j);
        const int base = (g * Ns * R + j) * estride + v * vstride;
        cpx<T> a[R];
#pragma unroll
        for (int r = 0; r < R; ++r) a[r] = buf[base + r * Ns * estride];
        if (j) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                cpx<T> w = tw[j * r * tstep];
                a[r] = INV ? cmulc(a[r], w) : cmul(a[r], w);
            }
        }
        if constexpr (R % 2 == 1 && R >= 11) {
            auto sink = [&](int k, cpx<T> v) { buf[base + k * Ns * estride] = v; };
            Dft<T, R, INV>::run_sink(a, sink);
        } else {
            Dft<T, R, INV>::run(a);
#pragma unroll
            for (int r = 0; r < R; ++r) buf[base + r * Ns * estride] = a[r];
        }
    }
}

// In-place decimation-in-frequency pass (transpose of dit_pass): butterfly
// first, then multiply output r by w_{Ns R}^{j r}.  Running the DIT passes'
// geometry in reverse order on natural-order input leaves the transform in
// digit-reversed order.
template <typename T, int R, bool INV>
__device__ __forceinline__ void dif_pass(cpx<T>* buf, int n, int Ns, int nvec, int vstride, int estride,
                                         const cpx<T>* __restrict__ tw, int tid, int nthr) {
    const int nbf = n / R;
    const int total = nbf * nvec;
    const int tstep = n / (Ns * R);
    const float inv_nvec = 1.0f / (float)nvec, inv_ns = 1.0f / (float)Ns;
    for (int idx = tid; idx < total; idx += nthr) {
        int v, j;
        const int bf = fast_div(idx, nvec, inv_nvec, v);
        const int g = fast_div(bf, Ns, inv_ns, j);
        const int base = (g * Ns * R + j) * estride + v * vstride;
        cpx<T> a[R];
#pragma unroll
        for (int r = 0; r < R; ++r) a[r] = buf[base + r * Ns * estride];
        if constexpr (R % 2 == 1 && R >= 11) {
            auto sink = [&](int k, cpx<T> v) {
                if (k && j) {
                    const cpx<T> w = tw[j * k * tstep];
                    v = INV ? cmulc(v, w) : cmul(v, w);
                }
                buf[base + k * Ns * estride] = v;
            };
            Dft<T, R, INV>::run_sink(a, sink);
        } else {
            Dft<T, R, INV>::run(a);
            if (j) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    cpx<T> w = tw[j * r * tstep];
                    a[r] = INV ? cmulc(a[r], w) : cmul(a[r], w);
                }
            }
#pragma unroll
            for (int r = 0; r < R; ++r) buf[base + r * Ns * estride] = a[r];
        }
    }
}

// A factorisation of n into LDS radices (first pass first).
struct FftPlan {
    int n;
    int npass;
    int radix[24];
};

#define S1G_POW2_LIST(X) X(2) X(4) X(8)

// POW2 = true restricts the dispatch to radices 2, 4, 8 (power-of-two plans).
template <typename T, bool INV, bool POW2>
__device__ __forceinline__ void dispatch_dit(int R, cpx<T>* buf, int n, int Ns, int nvec, int vstride,
                                             int estride, const cpx<T>* tw, int tid, int nthr) {
#define S1G_CASE(r)                                                                  \
    case r:                                                                          \
        dit_pass<T, r, INV>(buf, n, Ns, nvec, vstride, estride, tw, tid, nthr);      \
        break;
    if constexpr (POW2) {
        switch (R) { S1G_POW2_LIST(S1G_CASE) default: break; }
    } else {
        switch (R) { S1G_RADIX_LIST(S1G_CASE) default: break; }
    }
#undef S1G_CASE
}

template <typename T, bool INV, bool POW2>
__device__ __forceinline__ void dispatch_dif(int R, cpx<T>* buf, int n, int Ns, int nvec, int vstride,
                                             int estride, const cpx<T>* tw, int tid, int nthr) {
#define S1G_CASE(r)                                                                  \
    case r:                                                                          \
        dif_pass<T, r, INV>(buf, n, Ns, nvec, vstride, estride, tw, tid, nthr);      \
        break;
    if constexpr (POW2) {
        switch (R) { S1G_POW2_LIST(S1G_CASE) default: break; }
    } else {
        switch (R) { S1G_RADIX_LIST(S1G_CASE) default: break; }
    }
#undef S1G_CASE
}

// Full forward/inverse transform of nvec LDS-resident vectors whose input
// already sits in digit-reversed order; output in natural order.
template <typename T, bool INV, bool POW2 = false>
__device__ __forceinline__ void fft_dit_lds(cpx<T>* buf, const FftPlan& p, int nvec, int vstride, int estride,
                                            const cpx<T>* tw, int tid, int nthr) {
    int Ns = 1;
    for (int s = 0; s < p.npass; ++s) {
        const int R = p.radix[s];
        dispatch_dit<T, INV, POW2>(R, buf, p.n, Ns, nvec, vstride, estride, tw, tid, nthr);
        Ns *= R;
        __syncthreads();
    }
}

// Natural-order input; output in digit-reversed order.
template <typename T, bool INV, bool POW2 = false>
__device__ __forceinline__ void fft_dif_lds(cpx<T>* buf, const FftPlan& p, int nvec, int vstride, int estride,
                                            const cpx<T>* tw, int tid, int nthr) {
    int Ns = p.n;
    for (int s = p.npass - 1; s >= 0; --s) {
        const int R = p.radix[s];
        Ns /= R;
        dispatch_dif<T, INV, POW2>(R, buf, p.n, Ns, nvec, vstride, estride, tw, tid, nthr);
        __syncthreads();
    }
}

// ------------------------------------------------------- FP64 phase helpers
// e^{i ph} for a phase formed in FP64.  The FP64 chain takes FP64 sincos
// (the reference's arithmetic); the FP32 chain reduces the phase to a
// fraction of a turn in FP64 and takes the FP32 hardware sine and cosine of
// it, as its RCMC does, so phases of 1e8 rad lose nothing before the FP32
// trig.
template <typename T>
__device__ __forceinline__ void phase_sincos(double ph, double* s, double* c) {
    if constexpr (sizeof(T) == 8) {
        sincos(ph, s, c);
    } else {
        const double t = ph * 0.15915494309189535;  // turns
        const float f = (float)(t - rint(t));
        *s = (double)__builtin_amdgcn_sinf(f);
        *c = (double)__builtin_amdgcn_cosf(f);
    }
}

// x^3 as the reference writes it (pow) in FP64, a product in the FP32 chain.
template <typename T>
__device__ __forceinline__ double cube(double x) {
    if constexpr (sizeof(T) == 8) return pow(x, 3.0);
    else return x * x * x;
}
// Reduction of a large FP64 phase to [-pi, pi] (Cody & Waite 1980 additive
// reduction): 2 pi = C1 + C2 + C3 with C1, C2 holding 27 significant bits, so
// k*C1 and k*C2 are exact for |k| < 2^26 (|phase| < 4e8 rad).
__device__ __forceinline__ double reduce_2pi(double ph) {
#pragma clang fp contract(off)
    const double inv2pi = 0.15915494309189535;
    const double C1 = 6.283185303211212;       // 27 significant bits
    const double C2 = 3.968374295837407e-09;   // next 27 bits
    const double C3 = 2.2884754904439327e-17;  // remainder
    double k = rint(ph * inv2pi);
    double r = ph - k * C1;
    r = r - k * C2;
    r = r - k * C3;
    return r;
}

}  // namespace s1g
