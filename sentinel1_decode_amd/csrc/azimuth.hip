// azimuth.hip -- the azimuth FFTs of the range-Doppler chain: column panels
// of W range samples x all L lines in LDS; natural-order load, DIF passes,
// digit-reversed gather on store with the fftshift (forward,
// s1_decoder.cpp:499-505) or the 1/L scale (inverse, s1_decoder.cpp:682,
// signal_processing.cpp:581-605) folded in.
#include "s1g_internal.h"
#include "s1g_lds_fft.h"

#include <algorithm>
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
                                                          const int32_t* __restrict__ perm_g, T scale, int sh) {
    constexpr int L = (Rs * ...);
    constexpr int total = L * W;
    static_assert(total <= NT * MAXE, "panel fits the per-thread registers");
    extern __shared__ __align__(16) unsigned char smem_raw[];
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
    // forward with sh = L - L/2: out[i] = X[(i + sh) mod L] (fftshift); inverse: x[k] / L
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
static void launch_az_fixed(hipStream_t s, const DevFft& fl, cpx<T>* data, int64_t ld, int N, T scale, int sh) {
    constexpr int NT = 512;
    constexpr int L = (Rs * ...);
    constexpr int MAXE = (L * W + NT - 1) / NT;
    const int npanels = (N + W - 1) / W;
    const int per = (npanels + 7) / 8;
    const size_t lds = (size_t)L * W * sizeof(cpx<T>) + (size_t)L * (sizeof(cpx<T>) + sizeof(int32_t));
    hipLaunchKernelGGL((az_fft_fixed_kernel<T, INV, NT, W, MAXE, Rs...>), dim3((unsigned)(8 * per)), dim3(NT), lds,
                       s, data, ld, N, npanels, tw_of<T>(fl), fl.perm, scale, sh);
}

// A global panel index of a batch: its instance's buffer resource, leading
// dimension, first column and width (w = 0 past the last panel, so every
// access of it is masked).  Uniform: a scan of the kernel arguments.
struct AzPanel {
    __amdgpu_buffer_rsrc_t rsrc;
    int64_t ld;
    int j0, w;
};
template <typename T, int W>
__device__ __forceinline__ AzPanel az_panel(const AzBatch& bt, int L, int panel) {
    int k = 0;
    while (k + 1 < bt.n && panel >= bt.inst[k + 1].panel0) ++k;
    const AzInst& I = bt.inst[k];
    AzPanel p;
    p.ld = I.ld;
    p.j0 = (panel - I.panel0) * W;
    p.w = panel < bt.npanels ? min(W, I.N - p.j0) : 0;
    p.rsrc = buf_rsrc(I.data, (uint32_t)(((int64_t)(L - 1) * I.ld + I.N) * (int64_t)sizeof(cpx<T>)));
    return p;
}

// Persistent form of az_fft_fixed_kernel: a panel fills most of a CU's LDS,
// so workgroups cannot overlap one another's HBM traffic; instead one
// workgroup per CU walks its panels and issues the next panel's loads into
// registers before the LDS passes of the current one.  Panels are dealt so
// that at every step each XCD holds a contiguous run of them (neighbours
// share that XCD's L2 lines).  Loads and stores are raw buffer accesses
// whose masked elements carry an out-of-range offset: with no branch around
// them every thread issues exactly MAXE loads then MAXE stores per panel, so
// the wait for a panel's prefetched registers counts past the previous
// panel's stores, which keep draining while the next panel is transformed.
// The array must span less than 2 GiB (checked by the launcher).
// One launch may serve several arrays of the same length (AzBatch: per
// instance data, leading dimension, columns and first global panel); the
// workgroups deal the panels of all of them.
template <typename T, bool INV, int NT, int W, int MAXE, int... Rs>
__global__ __launch_bounds__(NT) void az_fft_pipe_kernel(AzBatch bt, const cpx<T>* __restrict__ tw_g,
                                                         const int32_t* __restrict__ perm_g, T scale, int sh) {
    constexpr int L = (Rs * ...);
    constexpr int total = L * W;
    static_assert(total <= NT * MAXE, "panel fits the per-thread registers");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<T>* buf = (cpx<T>*)smem_raw;        // L * W
    cpx<T>* tw = buf + (size_t)L * W;       // L
    int32_t* perm = (int32_t*)(tw + L);     // L
    const int tid0 = threadIdx.x;
    const int G = (int)gridDim.x;           // a multiple of 8
    const int first = (int)(blockIdx.x & 7) * (G >> 3) + (int)(blockIdx.x >> 3);
    const int npanels = bt.npanels;
    for (int k = tid0; k < L; k += NT) {
        tw[k] = tw_g[k];
        perm[k] = perm_g[k];
    }
    cpx<T> r[MAXE];
    // element offsets of a panel (kBufSkip past the array, the panel or the
    // last panel's columns)
    auto load = [&](int panel, int tid) {
        const AzPanel P = az_panel<T, W>(bt, L, panel);
        const int j0 = P.j0, w = P.w;
        const int64_t ld = P.ld;
        const auto rsrc = P.rsrc;
#pragma unroll
        for (int i = 0; i < MAXE; ++i) {
            const int idx = tid + NT * i;
            const int e = idx / W, v = idx % W;
            const bool ok = idx < total && v < w;
            const uint32_t off = ok ? (uint32_t)(((int64_t)e * ld + j0 + v) * (int64_t)sizeof(cpx<T>)) : kBufSkip;
            r[i] = buf_load<T>(rsrc, off);
        }
    };
    auto fill = [&](int tid) {
#pragma unroll
        for (int i = 0; i < MAXE; ++i) {
            const int idx = tid + NT * i;
            if (idx < total) buf[idx] = r[i];
        }
    };
    load(first, tid0);
    fill(tid0);
    __syncthreads();
    // Each step: prefetch the next panel, transform and store this one, then
    // move the prefetched registers into LDS.  The move waits for the
    // prefetch only: the stores issued after it are left to drain.
    for (int panel = first; panel < npanels; panel += G) {
        // The index arithmetic of the loads, passes and stores depends only
        // on the thread: hidden from the optimiser per panel so that it is
        // recomputed rather than hoisted out of the loop and kept live.
        int tid = tid0;
        asm volatile("" : "+v"(tid));
        load(panel + G, tid);  // in flight during the passes (all masked past the last panel)
        DifSeq<T, INV, Rs...>::run(buf, L, 1, W, 1, W, tw, tid, NT);
        const AzPanel P = az_panel<T, W>(bt, L, panel);
        const int j0 = P.j0, w = P.w;
        const int64_t ld = P.ld;
        const auto rsrc = P.rsrc;
#pragma unroll
        for (int i = 0; i < MAXE; ++i) {
            const int idx = tid + NT * i;
            const int k = idx / W, v = idx % W;
            const bool ok = idx < total && v < w;
            int src = k + sh;
            if (src >= L) src -= L;
            const cpx<T> y = buf[ok ? perm[src] * W + v : 0];
            const uint32_t off = ok ? (uint32_t)(((int64_t)k * ld + j0 + v) * (int64_t)sizeof(cpx<T>)) : kBufSkip;
            buf_store(INV ? scl(y, scale) : y, rsrc, off);
        }
        __syncthreads();  // every lane has read its outputs from LDS
        fill(tid);
        __syncthreads();
    }
}

// Persistent kernel with the first and last DIF passes in registers.  For a
// radix sequence (DIT order) RA, MIDS..., RZ the DIF passes run RZ (stride
// L / RZ), the MIDS, and RA (contiguous, no twiddles).  Here each thread
// loads whole first-pass butterflies (elements j + (L / RZ) r of a column)
// straight from global memory, transforms and twiddles them in registers
// and writes them to LDS; the middle passes run in LDS; and the last pass
// reads its RA contiguous elements from LDS and stores each output straight
// to its final row: buffer position p holds frequency ipos[p] (the inverse
// of the digit-reversal map), which lands on row (ipos[p] - sh) mod L (the
// fftshift, forward) or is scaled by 1/L (inverse).  That is two LDS passes
// and two barriers per panel fewer than az_fft_pipe_kernel.  Prefetch,
// XCD-aware dealing and the counted waits (buffer accesses, masked by
// out-of-range offsets) are as there; the array must span less than 2 GiB.
// One launch may serve several arrays of the same length, as the pipe kernel.
template <typename T, bool INV, int NT, int W, int RA, int RZ, int... MIDS>
__global__ __launch_bounds__(NT) void az_fft_reg_kernel(AzBatch bt, const cpx<T>* __restrict__ tw_g,
                                                        const int32_t* __restrict__ perm_g, T scale, int sh) {
    constexpr int L = RA * (MIDS * ... * 1) * RZ;
    constexpr int LZ = L / RZ;                       // first-pass butterflies per column (their stride)
    constexpr int QF = (LZ * W + NT - 1) / NT;       // first-pass butterflies per thread
    constexpr int LA = L / RA;                       // last-pass butterflies per column
    constexpr int QL = (LA * W + NT - 1) / NT;
    // Long odd last-pass butterflies are split over two lanes when whole
    // waves can take one half each (SPLIT): the pass's time is that of the
    // SIMD with the most butterfly waves, and LA W = 320 butterflies (five
    // waves) on eight waves would leave two full butterflies on one SIMD.
    constexpr bool SPLIT = RA >= 31 && RA % 2 == 1 && (LA * W) % 64 == 0;
    constexpr int QS = (2 * LA * W + NT - 1) / NT;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<T>* buf = (cpx<T>*)smem_raw;        // L * W
    cpx<T>* tw = buf + (size_t)L * W;       // L
    int32_t* ipos = (int32_t*)(tw + L);     // L
    const int tid0 = threadIdx.x;
    const int G = (int)gridDim.x;           // a multiple of 8
    const int first = (int)(blockIdx.x & 7) * (G >> 3) + (int)(blockIdx.x >> 3);
    const int npanels = bt.npanels;
    for (int k = tid0; k < L; k += NT) {
        tw[k] = tw_g[k];
        ipos[perm_g[k]] = k;
    }
    cpx<T> r[QF * RZ];
    auto load = [&](int panel, int tid) {
        const AzPanel P = az_panel<T, W>(bt, L, panel);
        const int j0 = P.j0, w = P.w;
        const int64_t ld = P.ld;
        const auto rsrc = P.rsrc;
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
    // first DIF pass (radix RZ, Ns = L / RZ, twiddle w_L^{j r}) into LDS
    auto fill = [&](int tid) {
#pragma unroll
        for (int q = 0; q < QF; ++q) {
            const int idx = tid + NT * q;
            if (idx < LZ * W) {
                const int v = idx % W, j = idx / W;
                cpx<T> a[RZ];
#pragma unroll
                for (int rr = 0; rr < RZ; ++rr) a[rr] = r[q * RZ + rr];
                Dft<T, RZ, INV>::run(a);
                if (j) {
#pragma unroll
                    for (int rr = 1; rr < RZ; ++rr) {
                        const cpx<T> wv = tw[j * rr];
                        a[rr] = INV ? cmulc(a[rr], wv) : cmul(a[rr], wv);
                    }
                }
#pragma unroll
                for (int rr = 0; rr < RZ; ++rr) buf[(j + LZ * rr) * W + v] = a[rr];
            }
        }
    };
    // last DIF pass (radix RA, Ns = 1) and the store
    auto last = [&](int panel, int tid) {
        const AzPanel P = az_panel<T, W>(bt, L, panel);
        const int j0 = P.j0, w = P.w;
        const int64_t ld = P.ld;
        const auto rsrc = P.rsrc;
        if constexpr (SPLIT) {
            // item i < LA W: the DC term and pairs 1 .. HS of butterfly i;
            // item LA W + i: pairs HS + 1 .. (RA - 1) / 2 of butterfly i.
            // LA W is a multiple of 64, so each wave takes one kind, and
            // waves past the last item skip the pass.
            constexpr int HS = (RA - 1) / 4;
#pragma unroll
            for (int q = 0; q < QS; ++q) {
                const int item = tid + NT * q;
                if (item < 2 * LA * W) {
                    const bool second = item >= LA * W;
                    const int idx = second ? item - LA * W : item;
                    const int v = idx % W;
                    const bool ok = v < w;
                    const int g = idx / W;
                    cpx<T> a[RA];
#pragma unroll
                    for (int rr = 0; rr < RA; ++rr) a[rr] = buf[(g * RA + rr) * W + v];
                    auto sink = [&](int k, cpx<T> y) {
                        int row = ipos[g * RA + k] - sh;
                        if (row < 0) row += L;
                        const uint32_t off =
                            ok ? (uint32_t)(((int64_t)row * ld + j0 + v) * (int64_t)sizeof(cpx<T>)) : kBufSkip;
                        buf_store(INV ? scl(y, scale) : y, rsrc, off);
                    };
                    if (second) Dft<T, RA, INV>::template run_sink_part<HS + 1, (RA - 1) / 2 + 1, false>(a, sink);
                    else Dft<T, RA, INV>::template run_sink_part<1, HS + 1, true>(a, sink);
                }
            }
            return;
        }
#pragma unroll
        for (int q = 0; q < QL; ++q) {
            const int idx = tid + NT * q;
            const int v = idx % W;
            const bool ok = idx < LA * W && v < w;
            const int g = ok ? idx / W : 0;
            cpx<T> a[RA];
#pragma unroll
            for (int rr = 0; rr < RA; ++rr) a[rr] = buf[(g * RA + rr) * W + v];
            auto sink = [&](int k, cpx<T> y) {
                int row = ipos[g * RA + k] - sh;
                if (row < 0) row += L;
                const uint32_t off =
                    ok ? (uint32_t)(((int64_t)row * ld + j0 + v) * (int64_t)sizeof(cpx<T>)) : kBufSkip;
                buf_store(INV ? scl(y, scale) : y, rsrc, off);
            };
            Dft<T, RA, INV>::run_sink(a, sink);
        }
    };
    load(first, tid0);
    fill(tid0);
    __syncthreads();
    for (int panel = first; panel < npanels; panel += G) {
        int tid = tid0;
        asm volatile("" : "+v"(tid));
        load(panel + G, tid);  // in flight during the passes (all masked past the last panel)
        DifSeq<T, INV, MIDS...>::run(buf, L, RA, W, 1, W, tw, tid, NT);
        last(panel, tid);
        __syncthreads();  // every lane has read this panel from LDS
        fill(tid);
        __syncthreads();
    }
}

// Workgroups of the persistent kernel: one per CU (the LDS allows no more),
// a multiple of 8 so every XCD gets the same count.
static int pipe_workgroups(int npanels) {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        ncu = 256;
    const int g = std::max(8, (ncu / 8) * 8);
    return std::min(g, ((npanels + 7) / 8) * 8);
}

// Every array of a batch spans less than 2 GiB (32-bit buffer offsets).
template <typename T>
static bool az_spans_ok(const AzInst* inst, int n, int L) {
    for (int k = 0; k < n; ++k)
        if (((int64_t)(L - 1) * inst[k].ld + inst[k].N) * (int64_t)sizeof(cpx<T>) >= ((int64_t)1 << 31)) return false;
    return true;
}

// The instances of a persistent launch (at most kMaxBatch), panels numbered
// across them; returns the panel count.
template <int W>
static int az_batch_of(const AzInst* inst, int n, AzBatch& bt) {
    bt.n = n;
    int np = 0;
    for (int k = 0; k < n; ++k) {
        bt.inst[k] = inst[k];
        bt.inst[k].panel0 = np;
        np += (inst[k].N + W - 1) / W;
    }
    bt.npanels = np;
    return np;
}

template <typename T, bool INV, int W, int... Rs>
static void launch_az_pipe(hipStream_t s, const DevFft& fl, const AzInst* inst, int n, T scale, int sh) {
    constexpr int NT = 512;
    constexpr int L = (Rs * ...);
    if (!az_spans_ok<T>(inst, n, L)) {
        // 32-bit buffer offsets: larger arrays take the one-panel kernel
        for (int k = 0; k < n; ++k)
            launch_az_fixed<T, INV, W, Rs...>(s, fl, (cpx<T>*)inst[k].data, inst[k].ld, inst[k].N, scale, sh);
        return;
    }
    constexpr int MAXE = (L * W + NT - 1) / NT;
    const size_t lds = (size_t)L * W * sizeof(cpx<T>) + (size_t)L * (sizeof(cpx<T>) + sizeof(int32_t));
    for (int b0 = 0; b0 < n; b0 += kMaxBatch) {
        AzBatch bt;
        const int np = az_batch_of<W>(inst + b0, std::min(kMaxBatch, n - b0), bt);
        if (np == 0) continue;
        hipLaunchKernelGGL((az_fft_pipe_kernel<T, INV, NT, W, MAXE, Rs...>), dim3((unsigned)pipe_workgroups(np)),
                           dim3(NT), lds, s, bt, tw_of<T>(fl), fl.perm, scale, sh);
    }
}

// Whether a plan's radix sequence is one compiled into az_fft_fixed_kernel.
static int fixed_sequence(const FftPlan& p) {
    auto is = [&](std::initializer_list<int> rs) {
        if ((int)rs.size() != p.npass) return false;
        int k = 0;
        for (int r : rs)
            if (p.radix[k++] != r) return false;
        return true;
    };
    if (is({37, 5, 8})) return 1;
    if (is({71, 5, 4})) return 2;
    if (is({19, 5, 8, 2})) return 3;
    return 0;
}

template <typename T, bool INV, int W, int RA, int RZ, int... MIDS>
static void launch_az_reg(hipStream_t s, const DevFft& fl, const AzInst* inst, int n, T scale, int sh) {
    constexpr int NT = 512;
    constexpr int L = RA * (MIDS * ... * 1) * RZ;
    if (!az_spans_ok<T>(inst, n, L)) {
        // 32-bit buffer offsets: larger arrays take the one-panel kernel
        for (int k = 0; k < n; ++k)
            launch_az_fixed<T, INV, W, RA, MIDS..., RZ>(s, fl, (cpx<T>*)inst[k].data, inst[k].ld, inst[k].N, scale,
                                                        sh);
        return;
    }
    const size_t lds = (size_t)L * W * sizeof(cpx<T>) + (size_t)L * (sizeof(cpx<T>) + sizeof(int32_t));
    for (int b0 = 0; b0 < n; b0 += kMaxBatch) {
        AzBatch bt;
        const int np = az_batch_of<W>(inst + b0, std::min(kMaxBatch, n - b0), bt);
        if (np == 0) continue;
        hipLaunchKernelGGL((az_fft_reg_kernel<T, INV, NT, W, RA, RZ, MIDS...>), dim3((unsigned)pipe_workgroups(np)),
                           dim3(NT), lds, s, bt, tw_of<T>(fl), fl.perm, scale, sh);
    }
}

template <typename T, bool INV, int W>
static bool az_reg(hipStream_t s, const DevFft& fl, const AzInst* inst, int n, T scale, int sh) {
    switch (fixed_sequence(fl.plan)) {
    case 1: launch_az_reg<T, INV, W, 37, 8, 5>(s, fl, inst, n, scale, sh); return true;
    case 2: launch_az_reg<T, INV, W, 71, 4, 5>(s, fl, inst, n, scale, sh); return true;
    case 3: launch_az_reg<T, INV, W, 19, 2, 5, 8>(s, fl, inst, n, scale, sh); return true;
    default: return false;
    }
}

template <typename T, bool INV, int W>
static bool az_pipe(hipStream_t s, const DevFft& fl, const AzInst* inst, int n, T scale, int sh) {
    switch (fixed_sequence(fl.plan)) {
    case 1: launch_az_pipe<T, INV, W, 37, 5, 8>(s, fl, inst, n, scale, sh); return true;
    case 2: launch_az_pipe<T, INV, W, 71, 5, 4>(s, fl, inst, n, scale, sh); return true;
    case 3: launch_az_pipe<T, INV, W, 19, 5, 8, 2>(s, fl, inst, n, scale, sh); return true;
    default: return false;
    }
}

// Plans with a compiled radix sequence (the IW1/IW2/IW3 burst lengths).
template <typename T, bool INV, int W>
static bool az_fixed(hipStream_t s, const DevFft& fl, cpx<T>* data, int64_t ld, int N, T scale, int sh) {
    switch (fixed_sequence(fl.plan)) {
    case 1: launch_az_fixed<T, INV, W, 37, 5, 8>(s, fl, data, ld, N, scale, sh); return true;
    case 2: launch_az_fixed<T, INV, W, 71, 5, 4>(s, fl, data, ld, N, scale, sh); return true;
    case 3: launch_az_fixed<T, INV, W, 19, 5, 8, 2>(s, fl, data, ld, N, scale, sh); return true;
    default: return false;
    }
}

template <typename T, bool INV, int NT, int MAXR>
__global__ __launch_bounds__(NT) void az_fft_generic_kernel(cpx<T>* __restrict__ data, int64_t ld, int L, int N, int W,
                                                    int npanels, FftPlan plan, const cpx<T>* __restrict__ tw,
                                                    const int32_t* __restrict__ perm, T scale, int sh) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<T>* buf = (cpx<T>*)smem_raw;
    cpx<T>* lo = buf + (size_t)L * W;  // two-level twiddle tables
    cpx<T>* hi = lo + kTwLo;
    const int panel = xcd_panel(blockIdx.x, npanels);
    if (panel >= npanels) return;
    const int j0 = panel * W;
    const int w = min(W, N - j0);
    const int tid = threadIdx.x;
    const int total = L * W;
    load_tw2(tw, L, lo, hi, tid, NT);
    // loads in batches of four per thread before their LDS stores
    constexpr int B = 4;
    for (int i0 = 0; i0 < total; i0 += NT * B) {
        cpx<T> x[B];
        int pos[B];
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const int idx = i0 + tid + NT * u;
            const int e = idx / W, v = idx - e * W;
            x[u] = mk((T)0, (T)0);
            pos[u] = -1;
            if (idx < total) {
                if (v < w) x[u] = data[(int64_t)e * ld + j0 + v];
                pos[u] = (INV ? e : perm[e]) * W + v;
            }
        }
#pragma unroll
        for (int u = 0; u < B; ++u)
            if (pos[u] >= 0) buf[pos[u]] = x[u];
    }
    __syncthreads();
    if (!INV) {
        fft_dit_lds_tw2v<T, false, MAXR>(buf, plan, W, 1, W, lo, hi, tid, NT);
        // sh = L - L/2 for the fftshift: out[i] = X[(i + L - L/2) mod L]
        for (int idx = tid; idx < total; idx += NT) {
            const int i = idx / W, v = idx - i * W;
            int src = i + sh;
            if (src >= L) src -= L;
            if (v < w) data[(int64_t)i * ld + j0 + v] = buf[src * W + v];
        }
    } else {
        fft_dif_lds_tw2v<T, true, MAXR>(buf, plan, W, 1, W, lo, hi, tid, NT);
        for (int idx = tid; idx < total; idx += NT) {
            const int k = idx / W, v = idx - k * W;
            if (v < w) data[(int64_t)k * ld + j0 + v] = scl(buf[perm[k] * W + v], scale);
        }
    }
}

template <typename T, bool INV>
static void launch_az_generic_t(hipStream_t s, const DevFft& fl, T* data, int64_t ld, int L, int N, int W,
                                int npanels, dim3 grid, size_t lds, T scale, int sh) {
    constexpr int NT = 512;
#define S1G_GEN(CLS)                                                                                            \
    hipLaunchKernelGGL((az_fft_generic_kernel<T, INV, NT, kRadixClassMax[CLS]>), grid, dim3(NT), lds, s,          \
                       (cpx<T>*)data, ld, L, N, W, npanels, fl.plan, tw_of<T>(fl), fl.perm, scale, sh)
    switch (radix_class(fl.plan)) {
        case 0: S1G_GEN(0); break;
        case 1: S1G_GEN(1); break;
        case 2: S1G_GEN(2); break;
        default: S1G_GEN(3); break;
    }
#undef S1G_GEN
}

static int launch_az_generic(hipStream_t s, int prec, const DevFft& fl, void* data, int64_t ld, int L, int N,
                             int inverse, int sh) {
    // panel width: 64-byte row segments (8 complex64 / 4 complex128) within the LDS budget
    int W = prec == S1G_F32 ? 8 : 4;
    const size_t esz = prec == S1G_F32 ? sizeof(cpx<float>) : sizeof(cpx<double>);
    const size_t tabs = esz * (size_t)(kTwLo + tw_hi_count(L));
    while (W > 1 && (size_t)L * W * esz + tabs > 150 * 1024) W /= 2;
    if ((size_t)L * W * esz + tabs > 150 * 1024) return -3;
    const int npanels = (N + W - 1) / W;
    const int per = (npanels + 7) / 8;
    dim3 grid((unsigned)(8 * per));
    const size_t lds = (size_t)L * W * esz + tabs;
    if (prec == S1G_F32) {
        const float sc = (float)(1.0 / (double)L);
        if (inverse) launch_az_generic_t<float, true>(s, fl, (float*)data, ld, L, N, W, npanels, grid, lds, sc, sh);
        else launch_az_generic_t<float, false>(s, fl, (float*)data, ld, L, N, W, npanels, grid, lds, 1.0f, sh);
    } else {
        const double sc = 1.0 / (L);
        if (inverse) launch_az_generic_t<double, true>(s, fl, (double*)data, ld, L, N, W, npanels, grid, lds, sc, sh);
        else launch_az_generic_t<double, false>(s, fl, (double*)data, ld, L, N, W, npanels, grid, lds, 1.0, sh);
    }
    return (int)hipGetLastError();
}

int azimuth_fft_variants(const DevFft& fl) { return fixed_sequence(fl.plan) ? 5 : 1; }

// Variants: with a compiled radix sequence 0 = 64-byte panels, 1 = 32-byte
// panels (twice the workgroups per CU), 2 = the generic kernel, 3 = 64-byte
// panels, persistent with the next panel's loads overlapping the passes,
// 4 = as 3 with the first and last passes in registers (az_fft_reg_kernel);
// otherwise only the generic kernel.
// The persistent variants (3, 4) for a list of arrays of length fl.plan.n.
template <typename T>
static void az_persistent(hipStream_t s, const DevFft& fl, const AzInst* inst, int n, int inverse, int variant,
                          int sh) {
    constexpr int W = sizeof(T) == 4 ? 8 : 4;
    const T sc = (T)(1.0 / (double)fl.plan.n);
    if (variant == 4)
        inverse ? az_reg<T, true, W>(s, fl, inst, n, sc, sh) : az_reg<T, false, W>(s, fl, inst, n, (T)1, sh);
    else
        inverse ? az_pipe<T, true, W>(s, fl, inst, n, sc, sh) : az_pipe<T, false, W>(s, fl, inst, n, (T)1, sh);
}

bool azimuth_batchable(const DevFft& fl, int variant) {
    return fixed_sequence(fl.plan) != 0 && (variant == 3 || variant == 4);
}

int launch_azimuth_fft_batch(hipStream_t s, int prec, const DevFft& fl, const AzInst* inst, int n, int inverse,
                             int variant) {
    if (!azimuth_batchable(fl, variant)) return -1;
    const int L = fl.plan.n;
    const int sh = inverse ? 0 : L - L / 2;  // the chain's fftshift after the forward transform
    if (prec == S1G_F32) az_persistent<float>(s, fl, inst, n, inverse, variant, sh);
    else az_persistent<double>(s, fl, inst, n, inverse, variant, sh);
    return (int)hipGetLastError();
}

int launch_azimuth_fft(hipStream_t s, int prec, const DevFft& fl, void* data, int64_t ld, int L, int N,
                       int inverse, int variant, int shift) {
    const int sh = (!inverse && shift) ? L - L / 2 : 0;
    if (fixed_sequence(fl.plan) && variant != 2) {
        AzInst one;
        one.data = data;
        one.ld = ld;
        one.N = N;
        one.panel0 = 0;
        if (variant == 3 || variant == 4) {
            if (prec == S1G_F32) az_persistent<float>(s, fl, &one, 1, inverse, variant, sh);
            else az_persistent<double>(s, fl, &one, 1, inverse, variant, sh);
        } else if (prec == S1G_F32) {
            const float sc = (float)(1.0 / (double)L);
            cpx<float>* d = (cpx<float>*)data;
            if (variant == 1)
                inverse ? az_fixed<float, true, 4>(s, fl, d, ld, N, sc, sh)
                        : az_fixed<float, false, 4>(s, fl, d, ld, N, 1.0f, sh);
            else
                inverse ? az_fixed<float, true, 8>(s, fl, d, ld, N, sc, sh)
                        : az_fixed<float, false, 8>(s, fl, d, ld, N, 1.0f, sh);
        } else {
            const double sc = 1.0 / (double)L;
            cpx<double>* d = (cpx<double>*)data;
            if (variant == 1)
                inverse ? az_fixed<double, true, 2>(s, fl, d, ld, N, sc, sh)
                        : az_fixed<double, false, 2>(s, fl, d, ld, N, 1.0, sh);
            else
                inverse ? az_fixed<double, true, 4>(s, fl, d, ld, N, sc, sh)
                        : az_fixed<double, false, 4>(s, fl, d, ld, N, 1.0, sh);
        }
        return (int)hipGetLastError();
    }
    return launch_az_generic(s, prec, fl, data, ld, L, N, inverse, sh);
}
}  // namespace s1g
