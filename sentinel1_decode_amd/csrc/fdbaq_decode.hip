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
//           words (all 64 lanes, coalesced), up to 20 KB per packet so that
//           occupancy does not depend on the launch's largest packet; words
//           beyond that are converted from global memory when read;
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

#include <algorithm>

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
// bits beyond the last checked position, plus the 4-word window.
constexpr int kPadWords = 72;

// Bit positions are 32-bit: user data up to 2^27 bytes (a Sentinel-1 space
// packet carries at most 65534).
constexpr int kMaxPacketBytes = 1 << 27;

// Bits a lane's speculative walk starts before its slice, so that it has
// usually fallen into step with the true symbol boundaries on arrival.
// kOverlap + 32 (the widest slice) + 11 (the longest symbol) <= 64 keeps
// every walk inside one 64-bit window.
constexpr int kOverlap = 12;
constexpr int kOverlapB4 = 24;

// Packets with more BAQ blocks than this are rejected (S1G_PKT_ERR_ARG):
// the block BRCs are kept in 4 registers per lane.
constexpr int kMaxBlocks = 256;

// Words of LDS a wavefront stages: packets up to this size (20 KB) are
// staged whole, so at least 7 wavefronts fit a CU whatever the largest
// packet of the launch; words beyond it come from global memory on demand.
constexpr int kLdsWords = 5120;

// A packet's user data as big-endian 32-bit words: word i holds bytes
// 4i..4i+3, zero past the packet (and word `wlim` onwards).  Words below
// `cap` are read from the LDS copy, the rest converted from global memory.
template <bool ALL>
struct PacketWords {
    const uint32_t* lds;
    int cap;
    const uint32_t* al;  // 4-byte aligned, at or before the packet's first byte
    int off;             // packet start - al, bytes
    int last_al;         // last aligned word touching the packet
    int nbytes;
    int wlim;

    // Word i from the aligned words lo = al[i], hi = al[i + 1] (zero when
    // past last_al).
    __device__ __forceinline__ uint32_t word_of(uint32_t lo, uint32_t hi, int i) const {
        const int valid = nbytes - 4 * i;  // bytes of this word inside the packet
        if (valid <= 0) return 0u;
        const uint64_t v = ((uint64_t)hi << 32) | lo;
        uint32_t be = __builtin_bswap32((uint32_t)(v >> (8 * off)));
        if (valid < 4) be &= ~(0xFFFFFFFFu >> (8 * valid));
        return be;
    }
    __device__ __forceinline__ uint32_t global_word(int i) const {
        if (4 * i >= nbytes) return 0u;
        const uint32_t lo = (i <= last_al) ? al[i] : 0u;
        const uint32_t hi = (off && i + 1 <= last_al) ? al[i + 1] : 0u;
        return word_of(lo, hi, i);
    }
    // With the whole packet staged (ALL) the reads need no bounds test: the
    // walk reads at most 67 words past the data (lane 63 of a block starting
    // at the last bit, slices of at most 32 bits, a 5-word window) and the
    // kPadWords staged zero words cover that.
    __device__ __forceinline__ uint32_t operator()(int i) const {
        if constexpr (ALL) return lds[i];
        if (i >= wlim) return 0u;
        return i < cap ? lds[i] : global_word(i);
    }
};

// `n` (1..32) bits starting at bit `pos`.
template <bool ALL>
__device__ __forceinline__ uint32_t peek_bits(const PacketWords<ALL>& w, int pos, int n) {
    const int wi = pos >> 5;
    const int sh = (int)(pos & 31);
    uint64_t v = ((uint64_t)w(wi) << 32) | (uint64_t)w(wi + 1);
    return (uint32_t)((v << sh) >> (64 - n));
}

// 160-bit big-endian window of five words.
struct Win5 {
    uint32_t w0, w1, w2, w3, w4;
};

// One of three words by i in {0, 1, 2}, with masks (kept as register
// arithmetic; a select chain may be lowered to a private-memory table).
__device__ __forceinline__ uint32_t sel3(int i, uint32_t a, uint32_t b, uint32_t c) {
    const uint32_t m1 = 0u - (uint32_t)(i >= 1), m2 = 0u - (uint32_t)(i >= 2);
    return (a & ~m1) | (b & m1 & ~m2) | (c & m2);
}

// Bits [o, o + 64) (0 <= o <= 95) of the window.
__device__ __forceinline__ uint64_t win64(const Win5& w, int o) {
    const int i = o >> 5, sh = o & 31;
    const uint32_t x = sel3(i, w.w0, w.w1, w.w2);
    const uint32_t y = sel3(i, w.w1, w.w2, w.w3);
    const uint32_t z = sel3(i, w.w2, w.w3, w.w4);
    const uint32_t hi = (uint32_t)(((((uint64_t)x << 32) | y) << sh) >> 32);
    const uint32_t lo = (uint32_t)(((((uint64_t)y << 32) | z) << sh) >> 32);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t len_table(int brc) { return fdbaq_len_table(brc); }

// Length of the symbol whose code bits head `x`, for code BRC: the walks
// keep their 64-bit window complemented and already past the sign bit, so
// the count of leading ones of the code is a leading-zero count of x (32
// when x is 0).  BRC 0-3: the code length is med3(clo + 1, lo, hi) with
// (lo, hi) = (1, 3), (1, 4), (1, 6), (2, 8), so the symbol length is
// 2 + med3(clo, lo - 1, hi - 1).  BRC 4 reads its nibble table by
// min(clo, 10) and adds the second code bit for its '0x' codes.
template <int BRC>
__device__ __forceinline__ int sym_len_n(uint32_t x) {
    const uint32_t clo = x ? (uint32_t)__builtin_clz(x) : 32u;
    if constexpr (BRC < 3) {
        constexpr uint32_t hi = BRC == 0 ? 2 : BRC == 1 ? 3 : 5;
        return 2 + (int)min(clo, hi);
    } else if constexpr (BRC == 3) {
        return 2 + (int)min(max(clo, 1u), 7u);
    } else {
        const int c = (int)min(clo, 10u);
        const int len = (int)((kLen4 >> (4 * c)) & 15u);
        const uint32_t extra = c == 0 ? ((~x >> 30) & 1u) : 0u;
        return 1 + len + (int)extra;
    }
}

// The walk window: the 64 bits from a symbol's sign bit, complemented and
// shifted past that sign bit.
__device__ __forceinline__ uint64_t walk_window(uint64_t bits) { return ~(bits << 1); }


// Value of the lane to the left (lane 0 gets 0): DPP wave_shr:1.
__device__ __forceinline__ int from_left(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x138, 0xF, 0xF, false); }

// Inclusive prefix sum across the 64 lanes: row_shr 1, 2, 4, 8 inside each
// row of 16 lanes, then row_bcast:15 adds lane 15's sum into the next row
// (rows 1 and 3) and row_bcast:31 adds lane 31's into rows 2 and 3; rows
// outside a DPP's row mask add the `old` operand, 0.
__device__ __forceinline__ int wave_inclusive_sum(int x, int lane) {
    (void)lane;
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xF, 0xF, true);
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xF, 0xF, true);
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xF, 0xF, true);
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xF, 0xF, true);
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xA, 0xF, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xC, 0xF, false);
    return x;
}

// Position of the j-th (1-based, 1 <= j <= popc(w)) set bit of w: a binary
// search on popcounts, with no branch.
__device__ __forceinline__ int select_bit(uint32_t w, int j) {
    int pos = 0;
#pragma unroll
    for (int half = 16; half >= 1; half >>= 1) {
        const int c = __popc(w & ((1u << half) - 1u));
        const bool up = j > c;
        j -= up ? c : 0;
        w = up ? w >> half : w;
        pos += up ? half : 0;
    }
    return pos;
}

// Position after `nsym` symbols of one BRC starting at bit `start`, found by
// all 64 lanes.  Lane l owns the slice [start + l*delta, start + (l+1)*delta)
// of a region of 64*delta bits:
//   1. it decodes speculatively from kOverlap bits before its slice (or from
//      `start`, a true boundary, if that is nearer), recording in `mask` the
//      slice offsets where its symbols start, until it leaves the slice;
//   2. its true entry is its left neighbour's exit; if that offset is one it
//      recorded, its exit stands, otherwise it walks again from the entry
//      (a slice that a symbol from the left covers entirely passes the entry
//      on as its exit); repeat until no exit changes;
//   3. a prefix sum of per-lane symbol counts finds the lane holding the
//      nsym-th symbol; the block ends where its next recorded symbol starts,
//      or at its exit.
// Regions repeat when one does not hold all nsym symbols.  delta <= 32.
// With want_hdr (IE, QE) *hdr receives the 8 bits that follow the end (the
// next block's BRC or THIDX field), else 0.  Reads stay below word `wlim`; once `start` passes `nbits`
// the block cannot end inside the packet and the walk stops there.
template <int BRC, bool ALL>
__device__ int walk_block(const PacketWords<ALL>& words, int nbits, int start, int nsym, int delta, int lane,
                          bool want_hdr, uint32_t* hdr) {
    constexpr bool B4 = BRC == 4;
    for (;;) {
        if (start > nbits) {
            *hdr = 0;
            return start;
        }
        const int s = start + lane * delta;
        // BRC 4 codes resynchronise more slowly: a longer lead-in, within the
        // 64-bit window (lead-in + delta + 11 <= 64)
        const int lead = B4 ? min(kOverlapB4, 53 - delta) : kOverlap;
        const int org = max(s - lead, start);
        const int w0 = org >> 5;
        Win5 w;
        w.w0 = words(w0);
        w.w1 = words(w0 + 1);
        w.w2 = words(w0 + 2);
        w.w3 = words(w0 + 3);
        w.w4 = words(w0 + 4);
        // window offset of slice offset o is o + zero (zero <= 31 + kOverlap)
        const int zero = (int)(org & 31) + (int)(s - org);
        uint64_t win = walk_window(win64(w, (int)(org & 31)));
        uint32_t mask = 0;
        int o = (int)(org - s);
        // lead-in: the overlap before the slice (nothing recorded)
        while (o < 0) {
            const int l = sym_len_n<BRC>((uint32_t)(win >> 32));
            o += l;
            win <<= l;
        }
        while (o < delta) {
            mask |= 1u << o;
            const int l = sym_len_n<BRC>((uint32_t)(win >> 32));
            o += l;
            win <<= l;
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
                    uint64_t wr = walk_window(win64(w, zero + entry));
                    while (o < delta) {
                        mask |= 1u << o;
                        const int l = sym_len_n<BRC>((uint32_t)(wr >> 32));
                        o += l;
                        wr <<= l;
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
        const uint32_t live = entry >= delta ? 0u : (mask >> entry);
        const int cnt = __popc(live);
        const int P = wave_inclusive_sum(cnt, lane);
        const int total = __builtin_amdgcn_readlane(P, 63);
        if (total >= nsym) {
            const unsigned long long bal = __ballot(P >= nsym);
            const int ls = __ffsll((long long)bal) - 1;
            // every lane forms its candidate without branching; lane ls's is
            // the block's: k block symbols start in its slice (1..cnt), so the
            // block ends at its (k+1)-th start, or at its exit when k == cnt
            const int k = nsym - (P - cnt);
            const int endrel = (k >= 1 && k < cnt) ? entry + select_bit(live, k + 1) : ex;
            if (want_hdr) {  // uniform: IE and QE blocks are followed by a header field
                const uint32_t h = (uint32_t)(win64(w, min(zero + endrel, 95)) >> 56);
                *hdr = (uint32_t)__builtin_amdgcn_readlane((int)h, ls);
            } else {
                *hdr = 0;
            }
            return start + ls * delta + __builtin_amdgcn_readlane(endrel, ls);
        }
        nsym -= total;
        start += 63 * delta + __builtin_amdgcn_readlane(ex, 63);
    }
}

// BRC of block b (uniform) from the lane-distributed registers.
__device__ __forceinline__ int brc_of(int b, uint32_t vb) {
    return (int)((uint32_t)__builtin_amdgcn_readlane((int)vb, b & 63) >> (3 * (b >> 6))) & 7;
}

// Decode the symbols of one (channel, block) item from a register bit buffer
// and reconstruct them (packet.cpp:678-693): for THIDX up to the BRC's
// limit, m below the flag gives m and m at the flag the simple level;
// otherwise NRL[brc][m] * SF[THIDX], one FP64 product.  Output element
// 2q + (ch & 1), component ch >> 1 (IE: re even, IO: re odd, QE: im even,
// QO: im odd).
template <typename T, bool ALL>
__device__ __forceinline__ void decode_item(const PacketWords<ALL>& words, const int32_t* bstart,
                                            const uint8_t* bbrc, const uint8_t* bthr, const double* nrl,
                                            int max_blocks, int nq, int nb, int ch, int b, T* orow) {
    const int brc = bbrc[b], th = bthr[b];
    const bool simple = th <= kThidxLimit[brc];
    const int flag = kMFlag[brc];
    const double sv = simple ? c_simple[brc][th] : 0.0;
    const double sf = c_sf[th];
    const double* nrow = nrl + 16 * brc;
    const uint64_t tab = len_table(brc);
    const uint32_t x4 = brc == 4 ? 1u : 0u;
    const uint64_t mbase = kMBase[brc], mshift = kMShift[brc];
    const uint32_t mmask = kMMask[brc];
    const int bp = bstart[ch * max_blocks + b];
    int wi = (bp >> 5) + 2;
    uint64_t cur = (((uint64_t)words(wi - 2) << 32) | words(wi - 1)) << (bp & 31);
    int valid = 64 - (int)(bp & 31);
    uint32_t nxt = words(wi);
    const int q0 = 128 * b;
    const int cnt = (b == nb - 1) ? nq - q0 : 128;
    T* dst = orow + 2 * (2 * (int64_t)q0 + (ch & 1)) + (ch >> 1);
    // Software pipeline: the reconstruction level of symbol s is read from
    // LDS while symbol s + 1 is decoded, and symbol s is stored after that,
    // so the LDS latency overlaps the next symbol's bit work.
    double lvl = 0.0;
    int m_prev = 0;
    uint32_t neg_prev = 0;
    for (int s = 0; s < cnt; ++s) {
        if (valid < 32) {
            cur |= (uint64_t)nxt << (32 - valid);
            valid += 32;
            ++wi;
            nxt = words(wi);
        }
        const uint32_t v = (uint32_t)(cur >> 53);
        const uint32_t code10 = v & 0x3FFu;
        const int clo = fdbaq_clo(code10);
        const int len = fdbaq_sym_len(code10, clo, tab, x4);
        const int m = fdbaq_mcode_tab(code10, clo, mbase, mshift, mmask, x4);
        const double lvl_next = nrow[m];
        cur <<= len;
        valid -= len;
        if (s > 0) {
            const double mg = simple ? (m_prev < flag ? (double)m_prev : sv) : __dmul_rn(lvl, sf);
            dst[4 * (int64_t)(s - 1)] = (T)(neg_prev ? -mg : mg);
        }
        lvl = lvl_next;
        m_prev = m;
        neg_prev = v >> 10;
    }
    if (cnt > 0) {
        const double mg = simple ? (m_prev < flag ? (double)m_prev : sv) : __dmul_rn(lvl, sf);
        dst[4 * (int64_t)(cnt - 1)] = (T)(neg_prev ? -mg : mg);
    }
}

// Two wavefronts per packet.  Wave 0 walks the block structure of IE, IO
// and QE; QE carries the THIDX every channel's reconstruction needs, so at
// that barrier wave 1 starts decoding the three channels' values while wave
// 0 walks QO; after the second barrier both waves decode QO and the rest of
// the three channels' items.  The walk is the packet's critical path, so the
// walking wave raises its issue priority (s_setprio) over the value decodes
// sharing its SIMD.  Only workgroup barriers hand work over, and every wave
// passes each of them once.
template <typename T, bool ALL>
__global__ __launch_bounds__(128) void fdbaq_decode_kernel(const uint8_t* __restrict__ ud,
                                                           const uint64_t* __restrict__ offs,
                                                           const uint32_t* __restrict__ lens,
                                                           const int32_t* __restrict__ nqs, int n_packets,
                                                           int max_words, int lds_words, int max_blocks,
                                                           int max_quads, T* __restrict__ out, int64_t ld_out,
                                                           int32_t* __restrict__ status) {
    extern __shared__ __align__(16) uint32_t smem[];  // base 16-aligned after the static words
    __shared__ int s_st, s_st_qo;  // status after IE..QE; status of the QO walk
    const int p = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63;
    // wave index as a scalar: the block walk under `wave == 0` is then
    // uniform control flow rather than exec-masked
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    if (p >= n_packets) return;

    uint32_t* staged = smem;                                     // lds_words
    int32_t* bstart = (int32_t*)(smem + lds_words);              // 4 * max_blocks
    uint8_t* bthr = (uint8_t*)(bstart + 4 * max_blocks);         // max_blocks
    uint8_t* bbrc = bthr + max_blocks;                           // max_blocks
    // 5 * 16 levels, 8-byte aligned; offset in words from the LDS base (a
    // cast through an integer would turn the table reads into FLAT loads,
    // which wait on every outstanding global store)
    const int nrl_words = (lds_words + 4 * max_blocks + (2 * max_blocks + 3) / 4 + 1) & ~1;
    double* nrl = (double*)(smem + nrl_words);
    for (int i = tid; i < 80; i += 128) nrl[i] = c_nrl[i >> 4][i & 15];

    const int nq = nqs[p];
    const uint32_t nbytes_u = lens[p];
    const int nbytes = (int)min(nbytes_u, (uint32_t)kMaxPacketBytes);
    const int nb = nq > 0 ? (nq + 127) / 128 : 0;
    T* orow = out + (int64_t)2 * ld_out * p;
    int st = S1G_PKT_OK;
    // a packet above the stated maxima would write past its row (ld_out is
    // only checked against 2 * max_quads) or overflow the LDS tables
    if (nq < 0 || nq > max_quads || nb > max_blocks || nb > kMaxBlocks ||
        nbytes_u > (uint32_t)kMaxPacketBytes || (nbytes + 3) / 4 > max_words)
        st = S1G_PKT_ERR_ARG;

    PacketWords<ALL> words;
    {
        const uint8_t* a = ud + offs[p];
        const int off = (int)((uintptr_t)a & 3);
        words.lds = staged;
        words.cap = lds_words;
        // aligned by pointer arithmetic, so the reads stay global loads
        words.al = (const uint32_t*)(a - off);
        words.off = off;
        words.last_al = (words.off + nbytes - 1) >> 2;
        words.nbytes = nbytes;
        words.wlim = (nbytes + 3) / 4 + kPadWords;
    }
    if (st == S1G_PKT_OK) {
        // ---- stage the leading words (all of them for packets up to the cap)
        // Raw buffer loads over the aligned words (zero past last_al): no
        // branch around a load, so each thread issues all kStageBatch pairs of
        // a batch before the first wait.
        const int n_stage = min(lds_words, words.wlim);
        const __amdgpu_buffer_rsrc_t prs = buf_rsrc(words.al, 4u * (uint32_t)(words.last_al + 1));
        constexpr int kStageBatch = 8;
        for (int b = 0; b < n_stage; b += 128 * kStageBatch) {
            uint32_t lo[kStageBatch], hi[kStageBatch];
#pragma unroll
            for (int u = 0; u < kStageBatch; ++u) {
                const int i = b + tid + 128 * u;
                const uint32_t o = i < n_stage ? 4u * (uint32_t)i : kBufSkip;
                lo[u] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(prs, (int)o, 0, 0);
                hi[u] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(prs, (int)(o + 4u), 0, 0);
            }
#pragma unroll
            for (int u = 0; u < kStageBatch; ++u) {
                const int i = b + tid + 128 * u;
                if (i < n_stage) staged[i] = words.word_of(lo[u], hi[u], i);
            }
        }
    }
    if (tid == 0) {
        s_st = st;
        s_st_qo = S1G_PKT_OK;
    }
    __syncthreads();

    // ---- phase 1 (wave 0): block structure, uniform control flow
    const int nbits = 8 * nbytes;
    int pos = 0;
    int prev_len = 0;
    uint32_t vb = 0;  // bits 3j..3j+2: BRC of block lane + 64 j
    auto walk_channel = [&](int ch) __attribute__((always_inline)) {
        // header field of the channel's first block from the packet; later
        // ones come out of the walk of the block before
        const bool hdr_field = ch == 0 || ch == 2;
        uint32_t hdr = hdr_field ? peek_bits(words, pos, 8) : 0u;
        for (int b = 0; b < nb && st == S1G_PKT_OK; ++b) {
            int brc;
            if (ch == 0) {
                if (pos + 3 > nbits) {
                    st = S1G_PKT_ERR_OVERRUN;
                    break;
                }
                const uint32_t v = hdr >> 5;
                if (v > 4) {
                    st = S1G_PKT_ERR_BRC;
                    break;
                }
                brc = (int)v;
                if (lane == (b & 63)) vb |= v << (3 * (b >> 6));
                if (lane == 0) bbrc[b] = (uint8_t)v;
                pos += 3;
            } else {
                brc = brc_of(b, vb);
            }
            if (ch == 2) {
                if (pos + 8 > nbits) {
                    st = S1G_PKT_ERR_OVERRUN;
                    break;
                }
                if (lane == 0) bthr[b] = (uint8_t)hdr;
                pos += 8;
            }
            if (lane == 0) bstart[ch * max_blocks + b] = (int32_t)pos;
            const int nsym = (b == nb - 1) ? nq - 128 * (nb - 1) : 128;
            // slice width from the previous block of this channel (or a BRC guess)
            const int est = b ? prev_len : nsym * (brc + 3);
            const int delta = min(max((est + 63) / 64 + 1, 2), 32);
            int end;
            switch (brc) {
                case 0: end = walk_block<0, ALL>(words, nbits, pos, nsym, delta, lane, hdr_field, &hdr); break;
                case 1: end = walk_block<1, ALL>(words, nbits, pos, nsym, delta, lane, hdr_field, &hdr); break;
                case 2: end = walk_block<2, ALL>(words, nbits, pos, nsym, delta, lane, hdr_field, &hdr); break;
                case 3: end = walk_block<3, ALL>(words, nbits, pos, nsym, delta, lane, hdr_field, &hdr); break;
                default: end = walk_block<4, ALL>(words, nbits, pos, nsym, delta, lane, hdr_field, &hdr); break;
            }
            prev_len = (int)(end - pos);
            pos = end;
            if (pos > nbits) st = S1G_PKT_ERR_OVERRUN;
        }
        pos = (pos + 15) & ~15;
    };
    if (wave == 0 && st == S1G_PKT_OK) {
        __builtin_amdgcn_s_setprio(2);
        for (int ch = 0; ch < 3 && st == S1G_PKT_OK; ++ch) walk_channel(ch);
        __builtin_amdgcn_s_setprio(0);
        if (lane == 0 && st != S1G_PKT_OK) s_st = st;
    }
    __syncthreads();
    st = s_st;
    // IE/IO/QE values: wave 1 decodes items [0, n2) beside the QO walk, the
    // rest fill the lanes the QO blocks leave free in the last round (128
    // lanes for nb QO blocks), so that round stays one round
    const int n2 = min(3 * nb, max(4 * nb - 128, 0));
    if (st == S1G_PKT_OK) {
        if (wave == 0) {
            // the QO walk is the packet's critical path; the value decodes
            // beside it (this packet's and other packets') have slack, so
            // the walking wave takes issue priority on its SIMD meanwhile
            __builtin_amdgcn_s_setprio(2);
            walk_channel(3);
            __builtin_amdgcn_s_setprio(0);
            if (lane == 0 && st != S1G_PKT_OK) s_st_qo = st;
        } else {
            // ---- phase 2 (wave 1): IE, IO, QE items [0, n2) while wave 0
            // walks QO
            for (int item = lane; item < n2; item += 64)
                decode_item(words, bstart, bbrc, bthr, nrl, max_blocks, nq, nb, item % 3, item / 3, orow);
        }
    }
    __syncthreads();
    st = s_st != S1G_PKT_OK ? s_st : s_st_qo;
    const int N = 2 * (nq > 0 ? nq : 0);
    if (st != S1G_PKT_OK) {
        for (int i = tid; i < 2 * N && i < 2 * ld_out; i += 128) orow[i] = (T)0;
        if (tid == 0) status[p] = st;
        return;
    }
    // ---- phase 3 (both waves): QO, then the IE/IO/QE items from n2 on
    const int n3 = nb + (3 * nb - n2);
    for (int t = tid; t < n3; t += 128) {
        const int item = n2 + (t - nb);
        const int ch = t < nb ? 3 : item % 3;
        const int b = t < nb ? t : item / 3;
        decode_item(words, bstart, bbrc, bthr, nrl, max_blocks, nq, nb, ch, b, orow);
    }
    if (tid == 0) status[p] = S1G_PKT_OK;
}

int launch_fdbaq_decode(hipStream_t stream, int prec, const uint8_t* ud, const uint64_t* offs, const uint32_t* lens,
                        const int32_t* nqs, int n_packets, int max_bytes, int max_quads, void* out, int64_t ld_out,
                        int32_t* status) {
    if (n_packets <= 0) return 0;
    const int max_words = (max_bytes + 3) / 4;
    const int lds_words = std::min(max_words + kPadWords, kLdsWords);
    const int max_blocks = (max_quads + 127) / 128 + 1;
    // staged words, block starts, THIDX and BRC bytes, then the 80 levels at
    // the same word offset the kernel computes
    const int nrl_words = (lds_words + 4 * max_blocks + (2 * max_blocks + 3) / 4 + 1) & ~1;
    const size_t lds = sizeof(uint32_t) * (size_t)nrl_words + sizeof(double) * 80;
    if (lds > 160 * 1024) return -1;
    dim3 grid(n_packets), block(128);
    const bool all = max_words + kPadWords <= kLdsWords;
#define S1G_LAUNCH(T, A)                                                                                       \
    hipLaunchKernelGGL((fdbaq_decode_kernel<T, A>), grid, block, lds, stream, ud, offs, lens, nqs, n_packets, \
                       max_words, lds_words, max_blocks, max_quads, (T*)out, ld_out, status)
    if (prec == S1G_F32) {
        if (all) S1G_LAUNCH(float, true);
        else S1G_LAUNCH(float, false);
    } else {
        if (all) S1G_LAUNCH(double, true);
        else S1G_LAUNCH(double, false);
    }
#undef S1G_LAUNCH
    return (int)hipGetLastError();
}

}  // namespace s1g
