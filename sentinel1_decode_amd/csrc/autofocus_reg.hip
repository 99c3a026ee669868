// autofocus_reg.hip -- the autofocus entropy and gradient passes for
// azimuth lengths with compiled radices (see autofocus.hip for the
// quantities).  Register-resident first and last radix passes, one LDS pass
// between them, persistent workgroups over column panels.
#include "s1g_internal.h"

#include <algorithm>
#include <type_traits>

namespace s1g {

// ------------------------------------------------------------------ fast path
// Lengths L = RZ * RM * RA with compiled radices.  Persistent workgroups (one
// per CU, a multiple of 8 so that every XCD holds a contiguous run of
// panels at each step) walk column panels of W range columns.  A panel sits
// in LDS with W pad elements after every 16 rows, so that the contiguous
// butterflies of the last DIF pass (rows 16 apart across the lanes) spread
// over the banks.  The first DIF pass runs in registers on the loaded
// elements (rows j + LZ r of a column), the middle pass in LDS, and the last
// one (RA contiguous rows) in registers again; the gradient pass turns
// straight from there into the forward DIT passes (the last DIF pass leaves
// the order that the first DIT pass takes), whose last pass lands on the
// rows the thread loaded, where rd e^{i phi} is still in its registers.
template <int W>
__device__ __forceinline__ int afp(int row, int v) { return (row + (row >> 4)) * W + v; }

// w_L^e (forward sign) from the split table: lo[e mod 64] * hi[e / 64]
template <typename T>
__device__ __forceinline__ cpx<T> af_tw(const cpx<T>* lo, const cpx<T>* hi, int e) {
    return cmul(lo[e & 63], hi[e >> 6]);
}

// PASS 0: entropy sums; 1: entropy sums and the image (row stride N);
// 2: gradient (reads S and T from scal).  Per-workgroup results: part[2 b],
// part[2 b + 1] (S, T) or gpart[b * L + row].
template <typename T, int PASS, int NT, int W, int RZ, int RM, int RA>
__global__ __launch_bounds__(NT) void af_reg_kernel(const cpx<T>* __restrict__ rd, int64_t ld, int N, int npanels,
                                                    const cpx<T>* __restrict__ tw_g, const cpx<T>* __restrict__ ph,
                                                    const double* __restrict__ scal, cpx<T>* __restrict__ img,
                                                    double* __restrict__ part, double* __restrict__ gpart) {
    constexpr int L = RZ * RM * RA, LZ = L / RZ, LA = L / RA, LM = RM * RA, NHI = (L + 63) / 64;
    constexpr int QF = (LZ * W + NT - 1) / NT;        // first / last-DIT butterflies per thread
    constexpr int QM = ((L / RM) * W + NT - 1) / NT;  // middle-pass butterflies per thread
    constexpr int QL = (LA * W + NT - 1) / NT;        // last-DIF butterflies per thread
    constexpr int LDATA = (L + L / 16) * W;
    static_assert(L >= 64 && NT % W == 0, "panel geometry");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<T>* buf = (cpx<T>*)smem_raw;
    cpx<T>* tlo = buf + LDATA;  // w_L^i, i < 64
    cpx<T>* thi = tlo + 64;     // w_L^{64 i}
    cpx<T>* tmid = thi + NHI;   // w_LM^i, i < LM
    __shared__ double red[2][NT / 64];
    const int tid0 = threadIdx.x;
    const int G = (int)gridDim.x;
    const int first = (int)(blockIdx.x & 7) * (G >> 3) + (int)(blockIdx.x >> 3);
    const uint32_t bytes = (uint32_t)(((int64_t)(L - 1) * ld + N) * (int64_t)sizeof(cpx<T>));
    const auto rsrc = buf_rsrc(rd, bytes);
    for (int k = tid0; k < 64; k += NT) tlo[k] = tw_g[k];
    for (int k = tid0; k < NHI; k += NT) thi[k] = tw_g[64 * k];
    for (int k = tid0; k < LM; k += NT) tmid[k] = tw_g[RZ * k];
    const T sc = (T)(1.0 / (double)L);
    T wa = (T)0, wb = (T)0;
    if constexpr (PASS == 2) {
        const double S = scal[0], Tt = scal[1];
        wa = (T)(-1.0 / S);
        wb = (T)(Tt / (S * S));
    }

    cpx<T> r[QF * RZ];  // loaded elements; after fill, rd e^{i phi} of rows j + LZ rr
    auto load = [&](int panel, int tid) {
        const int j0 = panel * W;
        const int w = panel < npanels ? min(W, N - j0) : 0;
#pragma unroll
        for (int q = 0; q < QF; ++q) {
            const int idx = tid + NT * q;
            const int v = idx % W, j = idx / W;
            const bool ok = idx < LZ * W && v < w;
#pragma unroll
            for (int rr = 0; rr < RZ; ++rr) {
                const uint32_t off =
                    ok ? (uint32_t)(((int64_t)(j + LZ * rr) * ld + j0 + v) * (int64_t)sizeof(cpx<T>)) : kBufSkip;
                r[q * RZ + rr] = buf_load<T>(rsrc, off);
            }
        }
    };
    // phase, then the first DIF pass of the inverse transform (radix RZ,
    // stride LZ, outputs times conj(w_L^{j k})) into LDS
    auto fill = [&](int tid) {
#pragma unroll
        for (int q = 0; q < QF; ++q) {
            const int idx = tid + NT * q;
            if (idx < LZ * W) {
                const int v = idx % W, j = idx / W;
                cpx<T> a[RZ];
#pragma unroll
                for (int rr = 0; rr < RZ; ++rr) {
                    r[q * RZ + rr] = cmul(r[q * RZ + rr], ph[j + LZ * rr]);
                    a[rr] = r[q * RZ + rr];
                }
                Dft<T, RZ, true>::run(a);
                if (j) {
#pragma unroll
                    for (int rr = 1; rr < RZ; ++rr) a[rr] = cmulc(a[rr], af_tw(tlo, thi, j * rr));
                }
#pragma unroll
                for (int rr = 0; rr < RZ; ++rr) buf[afp<W>(j + LZ * rr, v)] = a[rr];
            }
        }
    };
    // middle pass, radix RM over stride RA: DIF (inverse, twiddles after)
    // or DIT (forward, twiddles before)
    auto mid = [&](int tid, auto dit) {
        constexpr bool DIT = decltype(dit)::value;
#pragma unroll
        for (int q = 0; q < QM; ++q) {
            const int idx = tid + NT * q;
            if (idx < (L / RM) * W) {
                const int v = idx % W, bf = idx / W;
                const int g = bf / RA, jj = bf - g * RA;
                const int base = g * LM + jj;
                cpx<T> a[RM];
#pragma unroll
                for (int m = 0; m < RM; ++m) a[m] = buf[afp<W>(base + m * RA, v)];
                if constexpr (DIT) {
                    if (jj) {
#pragma unroll
                        for (int m = 1; m < RM; ++m) a[m] = cmul(a[m], tmid[jj * m]);
                    }
                    Dft<T, RM, false>::run(a);
                } else {
                    Dft<T, RM, true>::run(a);
                    if (jj) {
#pragma unroll
                        for (int m = 1; m < RM; ++m) a[m] = cmulc(a[m], tmid[jj * m]);
                    }
                }
#pragma unroll
                for (int m = 0; m < RM; ++m) buf[afp<W>(base + m * RA, v)] = a[m];
            }
        }
    };
    double s_acc = 0.0, t_acc = 0.0;
    // last DIF pass (radix RA, contiguous, no twiddles): position g RA + k
    // holds image row g / RM + RZ (g mod RM) + RZ RM k
    auto last_entropy = [&](int panel, int tid) {
        const int j0 = panel * W;
        const int w = min(W, N - j0);
#pragma unroll
        for (int q = 0; q < QL; ++q) {
            const int idx = tid + NT * q;
            if (idx < LA * W) {
                const int v = idx % W, g = idx / W;
                cpx<T> a[RA];
#pragma unroll
                for (int k = 0; k < RA; ++k) a[k] = buf[afp<W>(g * RA + k, v)];
                Dft<T, RA, true>::run(a);
                const int row0 = g / RM + RZ * (g % RM);
#pragma unroll
                for (int k = 0; k < RA; ++k) {
                    const cpx<T> y = scl(a[k], sc);
                    const T I = y.x * y.x + y.y * y.y;
                    s_acc += (double)I;
                    if (I > (T)0) t_acc += (double)(I * log(I));
                    if constexpr (PASS == 1) {
                        if (v < w) img[(int64_t)(row0 + RZ * RM * k) * N + j0 + v] = y;
                    }
                }
            }
        }
    };
    // gradient pass: last DIF pass, image weights, first DIT pass (radix RA
    // on the same contiguous rows, no twiddles) back into LDS
    auto last_weight = [&](int tid) {
#pragma unroll
        for (int q = 0; q < QL; ++q) {
            const int idx = tid + NT * q;
            if (idx < LA * W) {
                const int v = idx % W, g = idx / W;
                cpx<T> a[RA];
#pragma unroll
                for (int k = 0; k < RA; ++k) a[k] = buf[afp<W>(g * RA + k, v)];
                Dft<T, RA, true>::run(a);
#pragma unroll
                for (int k = 0; k < RA; ++k) {
                    const cpx<T> y = scl(a[k], sc);
                    const T I = y.x * y.x + y.y * y.y;
                    a[k] = scl(y, I > (T)0 ? (T)(wa * (T)log(I) + wb) : (T)0);
                }
                Dft<T, RA, false>::run(a);
#pragma unroll
                for (int k = 0; k < RA; ++k) buf[afp<W>(g * RA + k, v)] = a[k];
            }
        }
    };
    double acc[PASS == 2 ? QF * RZ : 1];
#pragma unroll
    for (int i = 0; i < (PASS == 2 ? QF * RZ : 1); ++i) acc[i] = 0.0;
    // last DIT pass (radix RZ over stride LZ, inputs times w_L^{j r}): the
    // gradient of rows j + LZ k, summed per lane over the panels
    auto last_grad = [&](int tid) {
#pragma unroll
        for (int q = 0; q < QF; ++q) {
            const int idx = tid + NT * q;
            if (idx < LZ * W) {
                const int v = idx % W, j = idx / W;
                cpx<T> a[RZ];
#pragma unroll
                for (int rr = 0; rr < RZ; ++rr) a[rr] = buf[afp<W>(j + LZ * rr, v)];
                if (j) {
#pragma unroll
                    for (int rr = 1; rr < RZ; ++rr) a[rr] = cmul(a[rr], af_tw(tlo, thi, j * rr));
                }
                Dft<T, RZ, false>::run(a);
#pragma unroll
                for (int k = 0; k < RZ; ++k) {
                    const cpx<T> gk = scl(a[k], sc);
                    const cpx<T> z = r[q * RZ + k];
                    acc[q * RZ + k] += 2.0 * ((double)gk.y * (double)z.x - (double)gk.x * (double)z.y);
                }
            }
        }
    };

    if constexpr (PASS == 2) {
        for (int panel = first; panel < npanels; panel += G) {
            int tid = tid0;
            asm volatile("" : "+v"(tid));
            load(panel, tid);
            fill(tid);
            __syncthreads();
            mid(tid, std::false_type{});
            __syncthreads();
            last_weight(tid);
            __syncthreads();
            mid(tid, std::true_type{});
            __syncthreads();
            last_grad(tid);
            __syncthreads();
        }
        // sum the W lanes of each row (adjacent lanes), then one write per row
#pragma unroll
        for (int q = 0; q < QF; ++q) {
            const int idx = tid0 + NT * q;
            const int v = idx % W, j = idx / W;
#pragma unroll
            for (int k = 0; k < RZ; ++k) {
                double x = acc[q * RZ + k];
#pragma unroll
                for (int o = 1; o < W; o <<= 1) x += __shfl_xor(x, o);
                if (idx < LZ * W && v == 0) gpart[(int64_t)blockIdx.x * L + j + LZ * k] = x;
            }
        }
    } else {
        load(first, tid0);
        fill(tid0);
        __syncthreads();
        for (int panel = first; panel < npanels; panel += G) {
            int tid = tid0;
            asm volatile("" : "+v"(tid));
            load(panel + G, tid);  // in flight during the passes (masked past the last panel)
            mid(tid, std::false_type{});
            __syncthreads();
            last_entropy(panel, tid);
            __syncthreads();
            fill(tid);
            __syncthreads();
        }
        for (int o = 32; o > 0; o >>= 1) {
            s_acc += __shfl_xor(s_acc, o);
            t_acc += __shfl_xor(t_acc, o);
        }
        if ((tid0 & 63) == 0) {
            red[0][tid0 >> 6] = s_acc;
            red[1][tid0 >> 6] = t_acc;
        }
        __syncthreads();
        if (tid0 == 0) {
            for (int q = 1; q < NT / 64; ++q) {
                s_acc += red[0][q];
                t_acc += red[1][q];
            }
            part[2 * blockIdx.x] = s_acc;
            part[2 * blockIdx.x + 1] = t_acc;
        }
    }
}

// grad[i] = sum over nparts partial rows gpart[p * L + i]: 64 rows per
// workgroup, 16 waves each summing every 16th partial.
__global__ __launch_bounds__(1024) void af_sum_partials_kernel(const double* __restrict__ gpart, int nparts, int L,
                                                               double* __restrict__ grad) {
    __shared__ double red[16][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + lane;
    double s0 = 0.0, s1 = 0.0;
    if (i < L) {
        int p = wv;
        for (; p + 16 < nparts; p += 32) {
            s0 += gpart[(int64_t)p * L + i];
            s1 += gpart[(int64_t)(p + 16) * L + i];
        }
        if (p < nparts) s0 += gpart[(int64_t)p * L + i];
    }
    red[wv][lane] = s0 + s1;
    __syncthreads();
    if (wv == 0 && i < L) {
        double s = 0.0;
        for (int q = 0; q < 16; ++q) s += red[q][lane];
        grad[i] = s;
    }
}

// Workgroups of the persistent kernels: one per CU, a multiple of 8.
static int af_workgroups_n(int npanels) {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        ncu = 256;
    const int g = std::max(8, (ncu / 8) * 8);
    return std::min(g, ((npanels + 7) / 8) * 8);
}

// Compiled lengths: L -> (RZ, RM, RA).
static int af_fast_case(int L) {
    switch (L) {
        case 4096: return 1;  // 16 16 16
        case 2048: return 2;  // 16 16 8
        case 1024: return 3;  // 16 8 8
        case 512: return 4;   // 8 8 8
        case 256: return 5;   // 8 8 4
        default: return 0;
    }
}

template <typename T>
static constexpr int af_width() { return sizeof(T) == 4 ? 4 : 2; }  // 32-byte row segments

bool af_fast(int L, int N, int64_t ld, int prec) {
    const size_t esz = prec == S1G_F32 ? sizeof(cpx<float>) : sizeof(cpx<double>);
    return af_fast_case(L) != 0 && ((int64_t)(L - 1) * ld + N) * (int64_t)esz < ((int64_t)1 << 31);
}

template <typename T, int RZ, int RM, int RA>
static void launch_af_reg(hipStream_t s, int pass, int N, const void* rd, int64_t ld, const DevFft& fl,
                          const cpx<T>* ph, const double* scal, void* img, double* part, double* gpart) {
    constexpr int NT = 512, W = af_width<T>();
    constexpr int L = RZ * RM * RA;
    const int npanels = (N + W - 1) / W;
    const int G = af_workgroups_n(npanels);
    const size_t lds = sizeof(cpx<T>) * ((size_t)(L + L / 16) * W + 64 + (L + 63) / 64 + RM * RA);
    const cpx<T>* tw = tw_of<T>(fl);
#define S1G_AF(P)                                                                                             \
    hipLaunchKernelGGL((af_reg_kernel<T, P, NT, W, RZ, RM, RA>), dim3((unsigned)G), dim3(NT), lds, s,         \
                       (const cpx<T>*)rd, ld, N, npanels, tw, ph, scal, (cpx<T>*)img, part, gpart)
    if (pass == 2) S1G_AF(2);
    else if (img) S1G_AF(1);
    else S1G_AF(0);
#undef S1G_AF
}

template <typename T>
static void af_reg_dispatch(hipStream_t s, int pass, int L, int N, const void* rd, int64_t ld, const DevFft& fl,
                            const cpx<T>* ph, const double* scal, void* img, double* part, double* gpart) {
    switch (af_fast_case(L)) {
        case 1: launch_af_reg<T, 16, 16, 16>(s, pass, N, rd, ld, fl, ph, scal, img, part, gpart); break;
        case 2: launch_af_reg<T, 16, 16, 8>(s, pass, N, rd, ld, fl, ph, scal, img, part, gpart); break;
        case 3: launch_af_reg<T, 16, 8, 8>(s, pass, N, rd, ld, fl, ph, scal, img, part, gpart); break;
        case 4: launch_af_reg<T, 8, 8, 8>(s, pass, N, rd, ld, fl, ph, scal, img, part, gpart); break;
        default: launch_af_reg<T, 8, 8, 4>(s, pass, N, rd, ld, fl, ph, scal, img, part, gpart); break;
    }
}


int af_reg_workgroups(int N, int prec) {
    const int W = prec == S1G_F32 ? af_width<float>() : af_width<double>();
    return af_workgroups_n((N + W - 1) / W);
}

int launch_af_reg(hipStream_t s, int prec, int pass, int L, int N, const void* rd, int64_t ld, const DevFft& fl,
                  const void* ph, const double* scal, void* img, double* part, double* gpart) {
    if (prec == S1G_F32)
        af_reg_dispatch<float>(s, pass, L, N, rd, ld, fl, (const cpx<float>*)ph, scal, img, part, gpart);
    else
        af_reg_dispatch<double>(s, pass, L, N, rd, ld, fl, (const cpx<double>*)ph, scal, img, part, gpart);
    return (int)hipGetLastError();
}

int launch_af_sum_partials(hipStream_t s, const double* gpart, int nparts, int L, double* grad) {
    hipLaunchKernelGGL(af_sum_partials_kernel, dim3((L + 63) / 64), dim3(1024), 0, s, gpart, nparts, L, grad);
    return (int)hipGetLastError();
}

}  // namespace s1g
