// s1g_lds_fft.h -- LDS-resident mixed-radix passes over a runtime plan whose
// twiddles also sit in LDS, as a two-level table: w_n^e = lo[e mod 64] *
// hi[e / 64] with lo[i] = w_n^i (i < 64) and hi[i] = w_n^{64 i}.  One complex
// product per twiddle instead of a global-memory read, whose latency the
// passes cannot hide (a pass waits for its twiddles before its butterflies).
#pragma once

#include "s1g_device.h"

namespace s1g {

constexpr int kTwLoBits = 6;
constexpr int kTwLo = 1 << kTwLoBits;

// Entries of the hi table for length n.
__host__ __device__ constexpr int tw_hi_count(int n) { return (n + kTwLo - 1) / kTwLo; }

// Fill lo / hi from the n-point global table (all threads of the workgroup).
template <typename T>
__device__ __forceinline__ void load_tw2(const cpx<T>* __restrict__ tw, int n, cpx<T>* lo, cpx<T>* hi, int tid,
                                         int nthr) {
    for (int i = tid; i < kTwLo; i += nthr) lo[i] = i < n ? tw[i] : mk((T)1, (T)0);
    for (int i = tid; i < tw_hi_count(n); i += nthr) hi[i] = tw[kTwLo * i];
}

template <typename T>
__device__ __forceinline__ cpx<T> tw2(const cpx<T>* lo, const cpx<T>* hi, int e) {
    return cmul(lo[e & (kTwLo - 1)], hi[e >> kTwLoBits]);
}

// dit_pass / dif_pass of one contiguous vector (s1g_device.h) with the
// two-level LDS twiddles.
template <typename T, int R, bool INV>
__device__ __forceinline__ void dit_pass_tw2(cpx<T>* buf, int n, int Ns, const cpx<T>* lo, const cpx<T>* hi, int tid,
                                             int nthr) {
    const int nbf = n / R;
    const int tstep = n / (Ns * R);
    const float inv_ns = 1.0f / (float)Ns;
    for (int bf = tid; bf < nbf; bf += nthr) {
        int j;
        const int g = fast_div(bf, Ns, inv_ns, j);
        const int base = g * Ns * R + j;
        cpx<T> a[R];
#pragma unroll
        for (int r = 0; r < R; ++r) a[r] = buf[base + r * Ns];
        if (j) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                const cpx<T> w = tw2(lo, hi, j * r * tstep);
                a[r] = INV ? cmulc(a[r], w) : cmul(a[r], w);
            }
        }
        if constexpr (R % 2 == 1 && R >= 11) {
            auto sink = [&](int k, cpx<T> v) { buf[base + k * Ns] = v; };
            Dft<T, R, INV>::run_sink(a, sink);
        } else {
            Dft<T, R, INV>::run(a);
#pragma unroll
            for (int r = 0; r < R; ++r) buf[base + r * Ns] = a[r];
        }
    }
}

template <typename T, int R, bool INV>
__device__ __forceinline__ void dif_pass_tw2(cpx<T>* buf, int n, int Ns, const cpx<T>* lo, const cpx<T>* hi, int tid,
                                             int nthr) {
    const int nbf = n / R;
    const int tstep = n / (Ns * R);
    const float inv_ns = 1.0f / (float)Ns;
    for (int bf = tid; bf < nbf; bf += nthr) {
        int j;
        const int g = fast_div(bf, Ns, inv_ns, j);
        const int base = g * Ns * R + j;
        cpx<T> a[R];
#pragma unroll
        for (int r = 0; r < R; ++r) a[r] = buf[base + r * Ns];
        if constexpr (R % 2 == 1 && R >= 11) {
            auto sink = [&](int k, cpx<T> v) {
                if (k && j) {
                    const cpx<T> w = tw2(lo, hi, j * k * tstep);
                    v = INV ? cmulc(v, w) : cmul(v, w);
                }
                buf[base + k * Ns] = v;
            };
            Dft<T, R, INV>::run_sink(a, sink);
        } else {
            Dft<T, R, INV>::run(a);
            if (j) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    const cpx<T> w = tw2(lo, hi, j * r * tstep);
                    a[r] = INV ? cmulc(a[r], w) : cmul(a[r], w);
                }
            }
#pragma unroll
            for (int r = 0; r < R; ++r) buf[base + r * Ns] = a[r];
        }
    }
}

// The largest radix of a plan; kernels are compiled per class of it
// (radices up to MAXR only), since a pass's registers grow with its radix
// and a kernel's register allocation is that of its largest pass.
inline int plan_max_radix(const FftPlan& p) {
    int m = 1;
    for (int s = 0; s < p.npass; ++s) m = p.radix[s] > m ? p.radix[s] : m;
    return m;
}
// Radix classes: 0 = up to 8, 1 = up to 23, 2 = up to 47, 3 = up to 71.
constexpr int kRadixClassMax[4] = {8, 23, 47, 71};
inline int radix_class(const FftPlan& p) {
    const int m = plan_max_radix(p);
    return m <= 8 ? 0 : m <= 23 ? 1 : m <= 47 ? 2 : 3;
}

// Digit-reversed input, natural-order output (fft_dit_lds with LDS twiddles);
// radices above MAXR are not compiled in (the caller checks the class).
template <typename T, bool INV, int MAXR = 71>
__device__ __forceinline__ void fft_dit_lds_tw2(cpx<T>* buf, const FftPlan& p, const cpx<T>* lo, const cpx<T>* hi,
                                                int tid, int nthr) {
    int Ns = 1;
    for (int s = 0; s < p.npass; ++s) {
        const int R = p.radix[s];
        switch (R) {
#define S1G_CASE(r)                                                                    \
    case r:                                                                            \
        if constexpr (r <= MAXR) dit_pass_tw2<T, r, INV>(buf, p.n, Ns, lo, hi, tid, nthr); \
        break;
            S1G_RADIX_LIST(S1G_CASE)
#undef S1G_CASE
            default:
                break;
        }
        Ns *= R;
        __syncthreads();
    }
}

// Natural-order input, digit-reversed output (fft_dif_lds with LDS twiddles).
template <typename T, bool INV, int MAXR = 71>
__device__ __forceinline__ void fft_dif_lds_tw2(cpx<T>* buf, const FftPlan& p, const cpx<T>* lo, const cpx<T>* hi,
                                                int tid, int nthr) {
    int Ns = p.n;
    for (int s = p.npass - 1; s >= 0; --s) {
        const int R = p.radix[s];
        Ns /= R;
        switch (R) {
#define S1G_CASE(r)                                                                    \
    case r:                                                                            \
        if constexpr (r <= MAXR) dif_pass_tw2<T, r, INV>(buf, p.n, Ns, lo, hi, tid, nthr); \
        break;
            S1G_RADIX_LIST(S1G_CASE)
#undef S1G_CASE
            default:
                break;
        }
        __syncthreads();
    }
}

// Several vectors at once (element e of vector v at buf[v * vstride +
// e * estride], as dit_pass / dif_pass in s1g_device.h) with the two-level
// LDS twiddles.
template <typename T, int R, bool INV>
__device__ __forceinline__ void dit_pass_tw2v(cpx<T>* buf, int n, int Ns, int nvec, int vstride, int estride,
                                              const cpx<T>* lo, const cpx<T>* hi, int tid, int nthr) {
    const int total = (n / R) * nvec;
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
                const cpx<T> w = tw2(lo, hi, j * r * tstep);
                a[r] = INV ? cmulc(a[r], w) : cmul(a[r], w);
            }
        }
        if constexpr (R % 2 == 1 && R >= 11) {
            auto sink = [&](int k, cpx<T> x) { buf[base + k * Ns * estride] = x; };
            Dft<T, R, INV>::run_sink(a, sink);
        } else {
            Dft<T, R, INV>::run(a);
#pragma unroll
            for (int r = 0; r < R; ++r) buf[base + r * Ns * estride] = a[r];
        }
    }
}

template <typename T, int R, bool INV>
__device__ __forceinline__ void dif_pass_tw2v(cpx<T>* buf, int n, int Ns, int nvec, int vstride, int estride,
                                              const cpx<T>* lo, const cpx<T>* hi, int tid, int nthr) {
    const int total = (n / R) * nvec;
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
            auto sink = [&](int k, cpx<T> x) {
                if (k && j) {
                    const cpx<T> w = tw2(lo, hi, j * k * tstep);
                    x = INV ? cmulc(x, w) : cmul(x, w);
                }
                buf[base + k * Ns * estride] = x;
            };
            Dft<T, R, INV>::run_sink(a, sink);
        } else {
            Dft<T, R, INV>::run(a);
            if (j) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    const cpx<T> w = tw2(lo, hi, j * r * tstep);
                    a[r] = INV ? cmulc(a[r], w) : cmul(a[r], w);
                }
            }
#pragma unroll
            for (int r = 0; r < R; ++r) buf[base + r * Ns * estride] = a[r];
        }
    }
}

template <typename T, bool INV, int MAXR = 71>
__device__ __forceinline__ void fft_dit_lds_tw2v(cpx<T>* buf, const FftPlan& p, int nvec, int vstride, int estride,
                                                 const cpx<T>* lo, const cpx<T>* hi, int tid, int nthr) {
    int Ns = 1;
    for (int s = 0; s < p.npass; ++s) {
        const int R = p.radix[s];
        switch (R) {
#define S1G_CASE(r)                                                                                          \
    case r:                                                                                                  \
        if constexpr (r <= MAXR) dit_pass_tw2v<T, r, INV>(buf, p.n, Ns, nvec, vstride, estride, lo, hi, tid, nthr); \
        break;
            S1G_RADIX_LIST(S1G_CASE)
#undef S1G_CASE
            default:
                break;
        }
        Ns *= R;
        __syncthreads();
    }
}

template <typename T, bool INV, int MAXR = 71>
__device__ __forceinline__ void fft_dif_lds_tw2v(cpx<T>* buf, const FftPlan& p, int nvec, int vstride, int estride,
                                                 const cpx<T>* lo, const cpx<T>* hi, int tid, int nthr) {
    int Ns = p.n;
    for (int s = p.npass - 1; s >= 0; --s) {
        const int R = p.radix[s];
        Ns /= R;
        switch (R) {
#define S1G_CASE(r)                                                                                          \
    case r:                                                                                                  \
        if constexpr (r <= MAXR) dif_pass_tw2v<T, r, INV>(buf, p.n, Ns, nvec, vstride, estride, lo, hi, tid, nthr); \
        break;
            S1G_RADIX_LIST(S1G_CASE)
#undef S1G_CASE
            default:
                break;
        }
        __syncthreads();
    }
}

}  // namespace s1g
