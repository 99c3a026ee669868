// fdbaq_decode.hip -- Level-0 FDBAQ (data format D) decode, one wavefront per
// space packet.
//
// Reference behaviour (AndrewPlayer3/sentinel1_decode): L0Packet::_decode
// (packet.cpp:404-453), _set_quad_type_d (:779-818), _get_h_code_type_d
// (:755-776), _get_s_values_type_d (:678-693), _get_signal_type_d (:696-752),
// huffman_decode / read_n_bits (decoding_utils.cpp:44-134).
//
// Work split inside the wavefront:
//   stage   the packet's user data is copied into LDS as big-endian 32-bit
//           words (all 64 lanes, coalesced);
//   phase 1 the block structure is walked channel by channel (IE, IO, QE, QO):
//           BRC/THIDX headers are read and the end of each block's symbols
//           is found by a speculative, self-synchronising walk (walk_block):
//           each lane decodes the symbols starting in its own slice of the
//           block, neighbours then agree on true slice entries, and a prefix
//           sum of symbol counts locates the block's last symbol; this yields
//           the start bit of every (channel, block) and all error checks;
//   phase 2 lanes take (channel, block) pairs and decode their symbols
//           serially from the recorded starts, reconstruct the samples in
//           FP64 (NRL*SF rounded once) and store them interleaved:
//           out[2k] = IE_k + i QE_k, out[2k+1] = IO_k + i QO_k.
#include "s1g_device.h"
#include "s1g_fdbaq.h"
#include "s1g_internal.h"

namespace s1g {

__constant__ double c_simple[5][10] = {
    {3.0000, 3.0000, 3.1600, 3.5300},
    {4.0000, 4.0000, 4.0800, 4.3700},
    {6.0000, 6.0000, 6.0000, 6.1500, 6.5000, 6.8800},
    {9.0000, 9.0000, 9.0000, 9.0000, 9.3600, 9.5000, 10.1000},
    {15.0000, 15.0000, 15.0000, 15.0000, 15.0000, 15.0000, 15.2200, 15.5000, 16.0600}};
__constant__ double c_nrl[5][16] = {
    {0.3637, 1.0915, 1.8208, 2.6406},
    {0.3042, 0.9127, 1.5216, 2.1313, 2.8426},
    {0.2305, 0.6916, 1.1528, 1.6140, 2.0754, 2.5369, 3.1191},
    {0.1702, 0.5107, 0.8511, 1.1916, 1.5321, 1.8726, 2.2131, 2.5536, 2.8942, 3.3744},
    {0.1130, 0.3389, 0.5649, 0.7908, 1.0167, 1.2428, 1.4687, 1.6947, 1.9206, 2.1466, 2.3725, 2.5985, 2.8244,
     3.0504, 3.2764, 3.6623}};
__constant__ double c_sf[256] = {
    0.00,   0.63,   1.25,   1.88,   2.51,   3.13,   3.76,   4.39,   5.01,   5.64,   6.27,   6.89,   7.52,
    8.15,   8.77,   9.40,   10.03,  10.65,  11.28,  11.91,  12.53,  13.16,  13.79,  14.41,  15.04,  15.67,
    16.29,  16.92,  17.55,  18.17,  18.80,  19.43,  20.06,  20.68,  21.31,  21.93,  22.56,  23.19,  23.81,
    24.44,  25.06,  25.69,  26.32,  26.94,  27.57,  28.20,  28.83,  29.45,  30.08,  30.71,  31.33,  31.96,
    32.59,  33.21,  33.84,  34.47,  35.09,  35.72,  36.35,  36.97,  37.60,  38.23,  38.85,  39.48,  40.11,
    40.73,  41.36,  41.99,  42.61,  43.24,  43.87,  44.49,  45.12,  45.75,  46.37,  47.00,  47.63,  48.25,
    48.88,  49.51,  50.13,  50.76,  51.39,  52.01,  52.64,  53.27,  53.89,  54.52,  55.15,  55.77,  56.40,
    57.03,  57.65,  58.28,  58.91,  59.54,  60.16,  60.79,  61.41,  62.04,  62.98,  64.24,  65.49,  66.74,
    68.00,  69.25,  70.50,  71.76,  73.01,  74.26,  75.52,  76.77,  78.02,  79.28,  80.53,  81.78,  83.04,
    84.29,  85.54,  86.80,  88.05,  89.30,  90.56,  91.81,  93.07,  94.32,  95.57,  96.83,  98.08,  99.33,
    100.58, 101.84, 103.09, 104.34, 105.60, 106.85, 108.10, 109.36, 110.61, 111.86, 113.11, 114.37, 115.62,
    116.87, 118.13, 119.38, 120.63, 121.89, 123.14, 124.39, 125.65, 126.90, 128.15, 129.41, 130.66, 131.91,
    133.17, 134.42, 135.67, 136.93, 138.18, 139.43, 140.69, 141.94, 143.19, 144.45, 145.70, 146.95, 148.21,
    149.46, 150.71, 151.97, 153.22, 154.47, 155.73, 156.98, 158.23, 159.49, 160.74, 161.99, 163.25, 164.50,
    165.75, 167.01, 168.26, 169.51, 170.77, 172.02, 173.27, 174.53, 175.78, 177.03, 178.29, 179.54, 180.79,
    182.05, 183.30, 184.55, 185.81, 187.06, 188.31, 189.57, 190.82, 192.07, 193.33, 194.58, 195.83, 197.09,
    198.34, 199.59, 200.85, 202.10, 203.35, 204.61, 205.86, 207.11, 208.37, 209.62, 210.87, 212.13, 213.38,
    214.63, 215.89, 217.14, 218.39, 219.65, 220.90, 222.15, 223.41, 224.66, 225.91, 227.17, 228.42, 229.67,
    230.93, 232.18, 233.43, 234.69, 235.94, 237.19, 238.45, 239.70, 240.95, 242.21, 243.46, 244.71, 245.97,
    247.22, 248.47, 249.73, 250.98, 252.23, 253.49, 254.74, 255.99, 255.99};

// Zero words staged after a packet: a walk region reaches at most 64 * 32
// bits beyond the last checked position, plus the 3-word window.
constexpr int kPadWords = 72;

// `n` (1..32) bits starting at bit `pos` of the big-endian word array.
__device__ __forceinline__ uint32_t peek_bits(const uint32_t* w, int64_t pos, int n) {
    const int64_t wi = pos >> 5;
    const int sh = (int)(pos & 31);
    uint64_t v = ((uint64_t)w[wi] << 32) | (uint64_t)w[wi + 1];
    return (uint32_t)((v << sh) >> (64 - n));
}

// FP64 reconstruction (packet.cpp:678-693); the sign is applied by negation
// so a zero magnitude with sign 1 gives -0.0 like pow(-1, 1) * 0.
__device__ __forceinline__ double recon(int brc, int thidx, int sign, int m) {
    double v;
    if (thidx <= kThidxLimit[brc])
        v = m < kMFlag[brc] ? (double)m : c_simple[brc][thidx];
    else
        v = __dmul_rn(c_nrl[brc][m], c_sf[thidx]);
    return sign ? -v : v;
}

// Bits [o, o + 11) of the 96-bit big-endian window (a, b, c), o + 11 <= 96.
__device__ __forceinline__ uint32_t win11(uint32_t a, uint32_t b, uint32_t c, int o) {
    const uint64_t hi = ((uint64_t)a << 32) | b;
    const uint64_t lo = ((uint64_t)b << 32) | c;
    const uint64_t v = o < 32 ? hi : lo;
    const int sh = o & 31;
    return (uint32_t)((v << sh) >> 53);
}

// Value of the lane to the left (lane 0 gets 0): DPP wave_shr:1.
__device__ __forceinline__ int from_left(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x138, 0xF, 0xF, false); }

// Inclusive prefix sum across the 64 lanes: row_shr 1, 2, 4, 8 inside each
// row of 16 lanes, then the totals of the lower rows (read from lanes 15,
// 31, 47) are added.
__device__ __forceinline__ int wave_inclusive_sum(int x, int lane) {
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xF, 0xF, true);
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xF, 0xF, true);
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xF, 0xF, true);
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xF, 0xF, true);
    const int r0 = __builtin_amdgcn_readlane(x, 15);
    const int r1 = __builtin_amdgcn_readlane(x, 31);
    const int r2 = __builtin_amdgcn_readlane(x, 47);
    const int row = lane >> 4;
    return x + (row > 0 ? r0 : 0) + (row > 1 ? r1 : 0) + (row > 2 ? r2 : 0);
}

__device__ __forceinline__ int sym_len(uint32_t a, uint32_t b, uint32_t c, int o, int brc) {
    return 1 + fdbaq_code_len(brc, win11(a, b, c, o) & 0x3FFu);
}

// Position after `nsym` symbols of bit-rate code `brc` starting at bit
// `start` (uniform across the wave).  Lane l owns the slice
// [start + l*delta, start + (l+1)*delta) of a region of 64*delta bits:
//   1. it decodes speculatively from its slice start, recording in `mask`
//      the slice offsets where its symbols start, until it leaves the slice;
//   2. its true entry is its left neighbour's exit; if that offset is one it
//      already visited, its walk has synchronised and its exit stands,
//      otherwise it walks again from the entry (a slice that a symbol from
//      the left covers entirely passes the entry on as its exit); repeat
//      until no exit changes;
//   3. a prefix sum of per-lane symbol counts finds the lane holding the
//      nsym-th symbol, which walks to its end.
// Regions repeat when one does not hold all nsym symbols.  delta <= 32.
// Reads stay below word `wlim` (the staged words); once `start` passes
// `nbits` the block cannot end inside the packet and the walk stops there.
__device__ int64_t walk_block(const uint32_t* __restrict__ words, int64_t wlim, int64_t nbits, int64_t start, int brc,
                              int nsym, int delta, int lane) {
    for (;;) {
        if (start > nbits) return start;
        const int64_t s = start + (int64_t)lane * delta;
        const int64_t w0 = s >> 5;
        const uint32_t a = w0 < wlim ? words[w0] : 0u;
        const uint32_t b = w0 + 1 < wlim ? words[w0 + 1] : 0u;
        const uint32_t c = w0 + 2 < wlim ? words[w0 + 2] : 0u;
        const int base = (int)(s & 31);
        uint32_t mask = 0;
        int o = 0;
        while (o < delta) {
            mask |= 1u << o;
            o += sym_len(a, b, c, base + o, brc);
        }
        int wex = o;  // exit of the recorded walk
        int ex = o;   // exit handed to the right neighbour
        int entry = 0;
        // Exits settle left to right, at least one more lane per round: 64
        // rounds suffice, the bound only guards against a malformed table.
        for (int round = 0; round <= 64; ++round) {
            const int up = from_left(ex);
            entry = lane == 0 ? 0 : up - delta;
            bool changed = false;
            int nex;
            if (entry >= delta) {
                nex = entry;  // a symbol from the left covers this whole slice
            } else {
                if (!((mask >> entry) & 1u)) {
                    mask = 0;
                    o = entry;
                    while (o < delta) {
                        mask |= 1u << o;
                        o += sym_len(a, b, c, base + o, brc);
                    }
                    wex = o;
                    changed = true;
                }
                nex = wex;
            }
            changed |= nex != ex;
            ex = nex;
            if (__ballot(changed) == 0) break;
        }
        const int cnt = entry >= delta ? 0 : __popc(mask >> entry);
        const int P = wave_inclusive_sum(cnt, lane);
        const int total = __builtin_amdgcn_readlane(P, 63);
        if (total >= nsym) {
            const unsigned long long bal = __ballot(P >= nsym);
            const int ls = __ffsll((long long)bal) - 1;
            int endrel = 0;
            if (lane == ls) {
                const int k = nsym - (P - cnt);
                o = entry;
                for (int i = 0; i < k; ++i) o += sym_len(a, b, c, base + o, brc);
                endrel = o;
            }
            endrel = __builtin_amdgcn_readlane(endrel, ls);
            return start + (int64_t)ls * delta + endrel;
        }
        nsym -= total;
        start += (int64_t)63 * delta + __builtin_amdgcn_readlane(ex, 63);
    }
}

template <typename T>
__global__ __launch_bounds__(64) void fdbaq_decode_kernel(const uint8_t* __restrict__ ud,
                                                          const uint64_t* __restrict__ offs,
                                                          const uint32_t* __restrict__ lens,
                                                          const int32_t* __restrict__ nqs, int n_packets,
                                                          int max_words, int max_blocks, T* __restrict__ out,
                                                          int64_t ld_out, int32_t* __restrict__ status) {
    extern __shared__ uint32_t smem[];
    const int p = blockIdx.x;
    const int lane = threadIdx.x;
    if (p >= n_packets) return;

    uint32_t* words = smem;                                      // max_words + kPadWords
    int32_t* bstart = (int32_t*)(smem + max_words + kPadWords);  // 4 * max_blocks
    uint8_t* bbrc = (uint8_t*)(bstart + 4 * max_blocks);         // max_blocks
    uint8_t* bthr = bbrc + max_blocks;                           // max_blocks

    const int nq = nqs[p];
    const int64_t nbytes = (int64_t)lens[p];
    const int nb = nq > 0 ? (nq + 127) / 128 : 0;
    T* orow = out + (int64_t)2 * ld_out * p;
    int st = S1G_PKT_OK;
    if (nq < 0 || nb > max_blocks || (nbytes + 3) / 4 > max_words) st = S1G_PKT_ERR_ARG;

    if (st == S1G_PKT_OK) {
        // ---- stage the user data as big-endian words, zero past the end
        const uint8_t* a = ud + offs[p];
        const uintptr_t au = (uintptr_t)a;
        const uint32_t* al = (const uint32_t*)(au & ~(uintptr_t)3);
        const int off = (int)(au & 3);
        const int64_t last_al = ((int64_t)off + nbytes - 1) >> 2;  // last aligned word touching the packet
        const int nwords = (int)((nbytes + 3) >> 2);
        for (int i = lane; i < nwords + kPadWords; i += 64) {
            uint32_t be = 0;
            if (i < nwords) {
                uint32_t lo = (i <= last_al) ? al[i] : 0u;
                uint32_t hi = (off && i + 1 <= last_al) ? al[i + 1] : 0u;
                uint64_t v = ((uint64_t)hi << 32) | lo;
                uint32_t le = (uint32_t)(v >> (8 * off));
                be = __builtin_bswap32(le);
                const int64_t valid = nbytes - 4 * (int64_t)i;  // bytes of this word inside the packet
                if (valid < 4) be &= ~(0xFFFFFFFFu >> (8 * valid));
            }
            words[i] = be;
        }
    }
    __syncthreads();

    // ---- phase 1: block structure (uniform control flow)
    const int64_t nbits = 8 * nbytes;
    const int64_t nwords_staged = (nbytes + 3) / 4 + kPadWords;
    int64_t pos = 0;
    int prev_len = 0;
    for (int ch = 0; ch < 4 && st == S1G_PKT_OK; ++ch) {
        for (int b = 0; b < nb && st == S1G_PKT_OK; ++b) {
            int brc;
            if (ch == 0) {
                if (pos + 3 > nbits) {
                    st = S1G_PKT_ERR_OVERRUN;
                    break;
                }
                const uint32_t v = peek_bits(words, pos, 3);
                if (v > 4) {
                    st = S1G_PKT_ERR_BRC;
                    break;
                }
                if (lane == 0) bbrc[b] = (uint8_t)v;
                brc = (int)v;
                pos += 3;
            } else {
                brc = bbrc[b];
            }
            if (ch == 2) {
                if (pos + 8 > nbits) {
                    st = S1G_PKT_ERR_OVERRUN;
                    break;
                }
                const uint32_t v = peek_bits(words, pos, 8);
                if (lane == 0) bthr[b] = (uint8_t)v;
                pos += 8;
            }
            if (lane == 0) bstart[ch * max_blocks + b] = (int32_t)pos;
            const int nsym = (b == nb - 1) ? nq - 128 * (nb - 1) : 128;
            // slice width from the previous block of this channel (or a BRC guess)
            const int est = b ? prev_len : nsym * (brc + 3);
            const int delta = min(max((est + 63) / 64 + 1, 2), 32);
            const int64_t end = walk_block(words, nwords_staged, nbits, pos, brc, nsym, delta, lane);
            prev_len = (int)(end - pos);
            pos = end;
            if (pos > nbits) st = S1G_PKT_ERR_OVERRUN;
        }
        pos = (pos + 15) & ~(int64_t)15;
        __syncthreads();
    }
    __syncthreads();

    const int N = 2 * (nq > 0 ? nq : 0);
    if (st != S1G_PKT_OK) {
        for (int i = lane; i < 2 * N && i < 2 * ld_out; i += 64) orow[i] = (T)0;
        if (lane == 0) status[p] = st;
        return;
    }

    // ---- phase 2: lanes decode (channel, block) pairs
    for (int item = lane; item < 4 * nb; item += 64) {
        const int ch = item & 3, b = item >> 2;
        const int brc = bbrc[b], th = bthr[b];
        int64_t bp = bstart[ch * max_blocks + b];
        const int q0 = 128 * b;
        const int cnt = (b == nb - 1) ? nq - q0 : 128;
        // out element 2q + (ch & 1), component ch >> 1 (IE: re even, IO: re odd, QE: im even, QO: im odd)
        T* dst = orow + 2 * (2 * (int64_t)q0 + (ch & 1)) + (ch >> 1);
        for (int s = 0; s < cnt; ++s) {
            const uint32_t bits = peek_bits(words, bp, 11);
            const uint32_t code = bits & 0x3FFu;
            const int sign = (int)(bits >> 10);
            const int clen = fdbaq_code_len(brc, code);
            const int m = fdbaq_mcode(brc, code);
            bp += 1 + clen;
            dst[4 * (int64_t)s] = (T)recon(brc, th, sign, m);
        }
    }
    if (lane == 0) status[p] = S1G_PKT_OK;
}

int launch_fdbaq_decode(hipStream_t stream, int prec, const uint8_t* ud, const uint64_t* offs, const uint32_t* lens,
                        const int32_t* nqs, int n_packets, int max_bytes, int max_quads, void* out, int64_t ld_out,
                        int32_t* status) {
    if (n_packets <= 0) return 0;
    const int max_words = (max_bytes + 3) / 4;
    const int max_blocks = (max_quads + 127) / 128 + 1;
    size_t lds = sizeof(uint32_t) * (size_t)(max_words + kPadWords) + sizeof(int32_t) * 4 * (size_t)max_blocks +
                 2 * (size_t)max_blocks;
    lds = (lds + 15) & ~(size_t)15;
    if (lds > 160 * 1024) return -1;
    dim3 grid(n_packets), block(64);
    if (prec == S1G_F32)
        hipLaunchKernelGGL(fdbaq_decode_kernel<float>, grid, block, lds, stream, ud, offs, lens, nqs, n_packets,
                           max_words, max_blocks, (float*)out, ld_out, status);
    else
        hipLaunchKernelGGL(fdbaq_decode_kernel<double>, grid, block, lds, stream, ud, offs, lens, nqs, n_packets,
                           max_words, max_blocks, (double*)out, ld_out, status);
    return (int)hipGetLastError();
}

}  // namespace s1g
