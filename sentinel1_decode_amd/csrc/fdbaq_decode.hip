// fdbaq_decode.hip -- Level-0 FDBAQ (data format D) decode, one workgroup per
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
//           out[2k] = IE_k + i QE_k, out[2k+1] = IO_k + i QO_k.  Lanes
//           4j..4j+3 take the four channels of one block, so each store
//           instruction writes whole samples, and a block's values are
//           decoded together once its QO start is known, following the QO
//           walk.
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
constexpr int kOverlapB4 = 40;  // swept on an RFI burst (mostly BRC 4): 24, 32, 40, 48 bits gave 1.82, 1.70, 1.68, 1.75 ms

// Packets with more BAQ blocks than this are rejected (S1G_PKT_ERR_ARG):
// the block BRCs are kept in 4 registers per lane.
constexpr int kMaxBlocks = 256;

// Words of LDS a workgroup stages: packets up to this size (64 KB, more than
// a space packet's user data can hold) are staged whole.  The launch sizes
// the LDS for its largest packet, so a burst of ~19 KB packets keeps 7
// workgroups per CU; one with RFI (larger codes, ~25-30 KB packets) keeps
// fewer but still reads every word from LDS -- words from global memory
// cost the walk a memory latency per window (2.0 ms against 0.6 ms per
// burst with a 20 KB cap).  Larger packets take their words beyond the cap
// from global memory on demand.
constexpr int kLdsWords = 16384 + 80;

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

// One lane's part of a region walk: the region is 64 slices of `delta`
// bits from `rstart`; lane l owns [rstart + l*delta, rstart + (l+1)*delta).
struct Region {
    Win5 w;
    int zero;       // window offset of slice offset 0
    uint32_t mask;  // slice offsets where the settled chain's symbols start
    int wex;        // exit of the lane's recorded walk
    int ex;         // exit handed to the right neighbour
    int entry;      // entry of the settled chain (>= delta: none)
};

// 1. Speculative decode: each lane decodes from kOverlap bits before its
//    slice (or from rstart if that is nearer), recording in `mask` the slice
//    offsets where its symbols start, until it leaves the slice.
template <int BRC, bool ALL>
__device__ __forceinline__ void region_walk(const PacketWords<ALL>& words, int rstart, int delta, int lane,
                                            Region& r) {
    constexpr bool B4 = BRC == 4;
    const int s = rstart + lane * delta;
    // BRC 4 codes resynchronise more slowly: a longer lead-in, within the
    // 64-bit window (lead-in + delta + 11 <= 64)
    const int lead = B4 ? min(kOverlapB4, 53 - delta) : kOverlap;
    const int org = max(s - lead, rstart);
    const int w0 = org >> 5;
    r.w.w0 = words(w0);
    r.w.w1 = words(w0 + 1);
    r.w.w2 = words(w0 + 2);
    r.w.w3 = words(w0 + 3);
    r.w.w4 = words(w0 + 4);
    // window offset of slice offset o is o + zero (zero <= 31 + kOverlap)
    r.zero = (int)(org & 31) + (int)(s - org);
    uint64_t win = walk_window(win64(r.w, (int)(org & 31)));
    uint32_t mask = 0;
    int o = (int)(org - s);
    // lead-in: the overlap before the slice (nothing recorded), then the
    // slice; two symbols per trip, the second under a select rather than a
    // branch, which halves the loop's exec-mask bookkeeping
    while (o < 0) {
        const int l = sym_len_n<BRC>((uint32_t)(win >> 32));
        o += l;
        win <<= l;
        const bool more = o < 0;
        const int l2 = sym_len_n<BRC>((uint32_t)(win >> 32));
        o = more ? o + l2 : o;
        win = more ? win << l2 : win;
    }
    while (o < delta) {
        mask |= 1u << o;
        const int l = sym_len_n<BRC>((uint32_t)(win >> 32));
        o += l;
        win <<= l;
        const bool more = o < delta;
        mask |= more ? 1u << (o & 31) : 0u;
        const int l2 = sym_len_n<BRC>((uint32_t)(win >> 32));
        o = more ? o + l2 : o;
        win = more ? win << l2 : win;
    }
    r.mask = mask;
    r.wex = o;
    r.ex = o;
    r.entry = 0;
}

// 2. Settle the chain that enters anchor lane `k` at slice offset `e` (a true
//    symbol boundary): a lane's entry is its left neighbour's exit; if that
//    offset is one it recorded, its exit stands, otherwise it walks again from
//    the entry (a slice that a symbol from the left covers entirely passes the
//    entry on as its exit); repeat until no exit changes.  Lanes below k take
//    no part.  Exits settle left to right, at least one more lane per round:
//    64 rounds suffice, the bound only guards against a malformed table.
template <int BRC>
__device__ __forceinline__ void region_settle(Region& r, int k, int e, int delta, int lane) {
    for (int round = 0; round <= 64; ++round) {
        const int up = from_left(r.ex);
        const int entry = lane < k ? 2 * delta : (lane == k ? e : up - delta);
        bool changed = false;
        int nex;
        if (entry >= delta) {
            nex = entry;  // a symbol from the left covers this whole slice
        } else {
            if (!((r.mask >> entry) & 1u)) {
                uint32_t mask = 0;
                int o = entry;
                uint64_t wr = walk_window(win64(r.w, r.zero + entry));
                while (o < delta) {
                    mask |= 1u << o;
                    const int l = sym_len_n<BRC>((uint32_t)(wr >> 32));
                    o += l;
                    wr <<= l;
                    const bool more = o < delta;
                    mask |= more ? 1u << (o & 31) : 0u;
                    const int l2 = sym_len_n<BRC>((uint32_t)(wr >> 32));
                    o = more ? o + l2 : o;
                    wr = more ? wr << l2 : wr;
                }
                r.mask = mask;
                r.wex = o;
                changed = true;
            }
            nex = r.wex;
        }
        changed |= nex != r.ex;
        r.ex = nex;
        r.entry = entry;
        if (__ballot(changed) == 0) break;
    }
}

// 3. Count: a prefix sum of per-lane symbol counts of the settled chain finds
//    the lane holding the nsym-th symbol; the block ends where its next
//    recorded symbol starts, or at its exit.  Returns false when the region
//    holds fewer than nsym symbols (*counted of them, the chain leaving the
//    region at *rend).  With want_hdr (IE, QE) *hdr receives the 8 bits that
//    follow the end (the next block's BRC or THIDX field).
__device__ __forceinline__ bool region_count(const Region& r, int rstart, int delta, int nsym, int lane,
                                             bool want_hdr, int* end, uint32_t* hdr, int* counted, int* rend) {
    const uint32_t live = r.entry >= delta ? 0u : (r.mask >> r.entry);
    const int cnt = __popc(live);
    const int P = wave_inclusive_sum(cnt, lane);
    const int total = __builtin_amdgcn_readlane(P, 63);
    if (total >= nsym) {
        const unsigned long long bal = __ballot(P >= nsym);
        const int ls = __ffsll((long long)bal) - 1;
        // every lane forms its candidate without branching; lane ls's is the
        // block's: k block symbols start in its slice (1..cnt), so the block
        // ends at its (k+1)-th start, or at its exit when k == cnt
        const int kk = nsym - (P - cnt);
        const int endrel = (kk >= 1 && kk < cnt) ? r.entry + select_bit(live, kk + 1) : r.ex;
        if (want_hdr) {  // uniform: IE and QE blocks are followed by a header field
            const uint32_t h = (uint32_t)(win64(r.w, min(r.zero + endrel, 95)) >> 56);
            *hdr = (uint32_t)__builtin_amdgcn_readlane((int)h, ls);
        } else {
            *hdr = 0;
        }
        *end = rstart + ls * delta + __builtin_amdgcn_readlane(endrel, ls);
        return true;
    }
    *counted = total;
    *rend = rstart + 63 * delta + __builtin_amdgcn_readlane(r.ex, 63);
    return false;
}

// Position after `nsym` symbols of one BRC starting at bit `start`, found by
// all 64 lanes: region walks from `start` (lane 0 enters at offset 0), repeated
// when one region does not hold all nsym symbols.  delta <= 32.  Reads stay
// below word `wlim`; once `start` passes `nbits` the block cannot end inside
// the packet and the walk stops there.
template <int BRC, bool ALL>
__device__ int walk_block(const PacketWords<ALL>& words, int nbits, int start, int nsym, int delta, int lane,
                          bool want_hdr, uint32_t* hdr) {
    for (;;) {
        if (start > nbits) {
            *hdr = 0;
            return start;
        }
        Region r;
        region_walk<BRC, ALL>(words, start, delta, lane, r);
        region_settle<BRC>(r, 0, 0, delta, lane);
        int end, counted, rend;
        if (region_count(r, start, delta, nsym, lane, want_hdr, &end, hdr, &counted, &rend)) return end;
        nsym -= counted;
        start = rend;
    }
}

// Inclusive prefix over the lanes of the settled chain's symbol counts.
__device__ __forceinline__ int region_prefix(const Region& r, int delta, int lane) {
    const uint32_t live = r.entry >= delta ? 0u : (r.mask >> r.entry);
    return wave_inclusive_sum(__popc(live), lane);
}

// A block walked from a region speculated before its start was known: the
// region from `pred` was walked and settled from lane 0 (prefix counts
// `pre`).  Given the true start `start`:
//   - if the settled chain passes through it (lane k holds it as a recorded
//     symbol start at or after its entry), the chain from `start` is that
//     chain from there on, so its counts are the settled ones shifted by a
//     constant: the block's end follows from one ballot, without settling
//     or summing again;
//   - otherwise the chain is settled again from lane k and counted;
//   - a start outside the region falls back to a plain walk, and a block
//     that outruns the region goes on from the region's end.
template <int BRC, bool ALL>
__device__ int resolve_block(const PacketWords<ALL>& words, int nbits, Region& r, int pre, int pred, int start,
                             int nsym, int delta, uint32_t dmagic, int lane, bool want_hdr, uint32_t* hdr) {
    const int off = start - pred;
    if (start > nbits || off < 0 || off >= 64 * delta)
        return walk_block<BRC, ALL>(words, nbits, start, nsym, delta, lane, want_hdr, hdr);
    // off / delta as a multiply-high by dmagic = floor((2^32 - 1) / delta) + 1,
    // exact for off < 64 delta and delta <= 32 (the division is on the
    // block's critical path; dmagic was formed while the region was walked)
    const int k = (int)__umulhi((uint32_t)off, dmagic), e = off - k * delta;
    const uint32_t mk = (uint32_t)__builtin_amdgcn_readlane((int)r.mask, k);
    const int ek = __builtin_amdgcn_readlane(r.entry, k);
    if (ek <= e && ((mk >> e) & 1u)) {
        const int ck = __popc(mk >> e);
        const int c0 = __builtin_amdgcn_readlane(pre, k) - ck;  // counts from `start`: pre - c0 (lanes >= k)
        const int total = __builtin_amdgcn_readlane(pre, 63) - c0;
        if (total >= nsym) {
            const int P = pre - c0;
            const unsigned long long bal = __ballot(lane >= k && P >= nsym);
            const int ls = __ffsll((long long)bal) - 1;
            const int ent = lane == k ? e : r.entry;
            const uint32_t live = ent >= delta ? 0u : (r.mask >> ent);
            const int cnt = __popc(live);
            const int kk = nsym - (P - cnt);
            const int endrel = (kk >= 1 && kk < cnt) ? ent + select_bit(live, kk + 1) : r.ex;
            if (want_hdr) {
                const uint32_t h = (uint32_t)(win64(r.w, min(r.zero + endrel, 95)) >> 56);
                *hdr = (uint32_t)__builtin_amdgcn_readlane((int)h, ls);
            } else {
                *hdr = 0;
            }
            return pred + ls * delta + __builtin_amdgcn_readlane(endrel, ls);
        }
        return walk_block<BRC, ALL>(words, nbits, pred + 63 * delta + __builtin_amdgcn_readlane(r.ex, 63),
                                    nsym - total, delta, lane, want_hdr, hdr);
    }
    region_settle<BRC>(r, k, e, delta, lane);
    int end, counted, rend;
    if (region_count(r, pred, delta, nsym, lane, want_hdr, &end, hdr, &counted, &rend)) return end;
    return walk_block<BRC, ALL>(words, nbits, rend, nsym - counted, delta, lane, want_hdr, hdr);
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

// Bits the speculative region of a block starts before the predicted start
// (a known block start plus IE's lengths of the blocks in between): channels
// of one block share its BRC and statistics, so the lengths differ by ~16
// bits per block; the margin grows with the number of blocks predicted over.
constexpr int kSpecMargin = 40;
constexpr int kSpecMarginStep = 24;
// The long class (kWalkersLong walkers) predicts six blocks ahead of a block
// it knows, over BRC-4 blocks of ~700 bits: on the 9-burst RFI swath margin
// steps of 12, 24, 40, 56, 72 and 96 bits gave 21.6, 21.3, 20.9, 20.9, 20.9
// and 20.8 ms.
constexpr int kSpecMarginStepLong = 56;
// Waves that walk the IO, QE and QO blocks, block b on wave b mod W; one
// more wave decodes values.  kWalkers for a launch over all of a burst's
// packets (or its short ones): four waves per packet keep 7 packets per CU.
// A launch over only the long packets of a burst (a few hundred RFI lines,
// about one packet per CU) is bound by each packet's walk latency, and more
// walking waves shorten the pipelined channels' walks: on a 9-burst IW2
// swath with ~300 RFI lines per burst, 3, 4, 5, 7, 9, 11 and 15 walkers gave
// 22.7, 21.9, 21.8, 21.6, 21.5, 21.6 and 25.5 ms.  With nearly every packet
// long (LDS-bound at 3 packets per CU) 7 walkers cost 3.5% against 3.
constexpr int kWalkers = 3;
constexpr int kWalkersLong = 7;
// Progress value of a finished (or abandoned) channel.
constexpr int kDone = 1 << 20;

__device__ __forceinline__ int lds_load_acq(int* p) {
    return __builtin_amdgcn_readfirstlane(__hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
}
// Progress only grows (a posted kDone is never taken back).
__device__ __forceinline__ void lds_store_rel(int* p, int v) {
    __hip_atomic_fetch_max(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Waits (wave-uniform) until *flag >= need or an error is posted; false on
// error.  Every wave that others wait on posts its progress, and kDone
// with the error when it stops, so the wait ends; the bound only guards
// against a broken invariant (it posts S1G_PKT_ERR_ARG).
// SLEEP: s_sleep units (64 cycles) between polls -- long for the value
// wave's wait on the walks, so that its polling takes few issue slots from
// the walking waves sharing the CU.
template <int SLEEP>
__device__ __forceinline__ bool wait_for(int* flag, int need, int* err, int lane) {
    for (int spin = 0; spin < (1 << 24); ++spin) {
        if (lds_load_acq(err) != S1G_PKT_OK) return false;
        if (lds_load_acq(flag) >= need) return true;
        __builtin_amdgcn_s_sleep(SLEEP);
    }
    if (lane == 0) __hip_atomic_store(err, S1G_PKT_ERR_ARG, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    return false;
}

// Waits (wave-uniform, no sleep: it is on the walk's critical path) until
// the in-channel progress word reaches `need`, with one LDS load per poll.
// Progress within a channel never reaches kDone before block need - 1 is
// published, so kDone means an error was posted: false.  The bound only
// guards against a broken invariant (it posts S1G_PKT_ERR_ARG).
__device__ __forceinline__ bool wait_start(int* flag, int need, int* err, int lane) {
    for (int spin = 0; spin < (1 << 26); ++spin) {
        const int v = __builtin_amdgcn_readfirstlane(
            __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
        if (v >= need) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");  // the block start written before v
            return v < kDone;
        }
    }
    if (lane == 0) __hip_atomic_store(err, S1G_PKT_ERR_ARG, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    return false;
}

// W + 1 wavefronts per packet.
//   wave 0 walks IE block by block (each block's BRC is in the stream);
//   waves 0 .. W - 1 then walk IO, QE and QO as a pipeline: block b
//     is walked by wave b mod W.  Its BRC is known from IE, so before
//     its start is known the wave speculates: it walks and settles a region
//     from the latest block start it knows (the block after the one it last
//     resolved) plus IE's lengths of the blocks in between, less a margin
//     that grows with their number.  When the wave before it publishes the
//     start (an LDS word with release / acquire at workgroup scope), the
//     chain from there is the settled chain whenever the start is one of its
//     recorded symbol starts -- one ballot -- so each block costs about
//     1/W of a walk on the critical path;
//   wave W decodes values, in groups of 16 blocks (lane 4j + ch takes
//     channel ch of the group's block j, so each store instruction writes
//     whole samples), as soon as the group's QO starts are published; the
//     walking waves join it when the walks end.
// Errors are found in block order (a block is resolved only after the one
// before it), posted once, and end every wait.  The walks are the packet's
// critical path, so the walking waves raise their issue priority
// (s_setprio) over the value decodes sharing their SIMDs.
template <typename T, bool ALL, int W>
__global__ __launch_bounds__(64 * (W + 1)) void fdbaq_decode_kernel(const uint8_t* __restrict__ ud,
                                                           const uint64_t* __restrict__ offs,
                                                           const uint32_t* __restrict__ lens,
                                                           const int32_t* __restrict__ nqs, int n_packets,
                                                           int max_words, int lds_words, int max_blocks,
                                                           int max_quads, T* __restrict__ out, int64_t ld_out,
                                                           int32_t* __restrict__ status, uint32_t lo_bytes,
                                                           uint32_t hi_bytes) {
    constexpr int kThreads = 64 * (W + 1);
    extern __shared__ __align__(16) uint32_t smem[];  // base 16-aligned after the static words
    __shared__ int s_err;          // first error status
    __shared__ int s_pub[4];       // blocks of each channel whose start is published; kDone when finished
    __shared__ int s_chan_end[4];  // end of each channel's last block, before the 16-bit alignment
    __shared__ int s_next_group;   // next value-decode group
    const int p = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63;
    // wave index as a scalar: the walks under `wave == ...` are then uniform
    // control flow rather than exec-masked
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    if (p >= n_packets) return;
    // this launch's size class: lo_bytes < length <= hi_bytes (another
    // launch of the burst takes the other packets)
    {
        const uint32_t len = lens[p];
        if (len <= lo_bytes || len > hi_bytes) return;
    }

    uint32_t* staged = smem;                                     // lds_words
    int32_t* bstart = (int32_t*)(smem + lds_words);              // 4 * max_blocks
    uint8_t* bthr = (uint8_t*)(bstart + 4 * max_blocks);         // max_blocks
    uint8_t* bbrc = bthr + max_blocks;                           // max_blocks
    // 5 * 16 levels, 8-byte aligned; offset in words from the LDS base (a
    // cast through an integer would turn the table reads into FLAT loads,
    // which wait on every outstanding global store)
    const int nrl_words = (lds_words + 4 * max_blocks + (2 * max_blocks + 3) / 4 + 1) & ~1;
    double* nrl = (double*)(smem + nrl_words);
    for (int i = tid; i < 80; i += kThreads) nrl[i] = c_nrl[i >> 4][i & 15];

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
        for (int b = 0; b < n_stage; b += kThreads * kStageBatch) {
            uint32_t lo[kStageBatch], hi[kStageBatch];
#pragma unroll
            for (int u = 0; u < kStageBatch; ++u) {
                const int i = b + tid + kThreads * u;
                const uint32_t o = i < n_stage ? 4u * (uint32_t)i : kBufSkip;
                lo[u] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(prs, (int)o, 0, 0);
                hi[u] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(prs, (int)(o + 4u), 0, 0);
            }
#pragma unroll
            for (int u = 0; u < kStageBatch; ++u) {
                const int i = b + tid + kThreads * u;
                if (i < n_stage) staged[i] = words.word_of(lo[u], hi[u], i);
            }
        }
    }
    if (tid == 0) {
        s_err = st;
        for (int c = 0; c < 4; ++c) {
            s_pub[c] = 0;
            s_chan_end[c] = 0;
        }
        s_next_group = 0;
    }
    __syncthreads();

    const int nbits = 8 * nbytes;
    // posts error `e` (first one wins) and ends every wait
    auto fail = [&](int e) __attribute__((always_inline)) {
        if (lane == 0) {
            int expect = S1G_PKT_OK;
            __hip_atomic_compare_exchange_strong(&s_err, &expect, e, __ATOMIC_RELEASE, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_WORKGROUP);
            for (int c = 0; c < 4; ++c) lds_store_rel(&s_pub[c], kDone);
        }
    };
    // the block walk of one BRC (uniform switch)
    auto walk_any = [&](int brc, int start, int nsym, int delta, bool want_hdr, uint32_t* hdr)
                        __attribute__((always_inline)) {
        switch (brc) {
            case 0: return walk_block<0, ALL>(words, nbits, start, nsym, delta, lane, want_hdr, hdr);
            case 1: return walk_block<1, ALL>(words, nbits, start, nsym, delta, lane, want_hdr, hdr);
            case 2: return walk_block<2, ALL>(words, nbits, start, nsym, delta, lane, want_hdr, hdr);
            case 3: return walk_block<3, ALL>(words, nbits, start, nsym, delta, lane, want_hdr, hdr);
            default: return walk_block<4, ALL>(words, nbits, start, nsym, delta, lane, want_hdr, hdr);
        }
    };

    if (s_err == S1G_PKT_OK && wave == 0) {
        // ---- IE, block by block: the BRC of each block heads it
        __builtin_amdgcn_s_setprio(2);
        int pos = 0, prev_len = 0;
        uint32_t hdr = peek_bits(words, pos, 8);
        bool ok = true;
        for (int b = 0; b < nb; ++b) {
            if (pos + 3 > nbits) {
                fail(S1G_PKT_ERR_OVERRUN);
                ok = false;
                break;
            }
            const uint32_t v = hdr >> 5;
            if (v > 4) {
                fail(S1G_PKT_ERR_BRC);
                ok = false;
                break;
            }
            if (lane == 0) bbrc[b] = (uint8_t)v;
            pos += 3;
            if (lane == 0) bstart[b] = (int32_t)pos;
            const int nsym = (b == nb - 1) ? nq - 128 * (nb - 1) : 128;
            // slice width from the previous block (or a BRC guess)
            const int est = b ? prev_len : nsym * ((int)v + 3);
            const int delta = min(max((est + 63) / 64 + 1, 2), 32);
            const int end = walk_any((int)v, pos, nsym, delta, true, &hdr);
            prev_len = end - pos;
            pos = end;
            if (pos > nbits) {
                fail(S1G_PKT_ERR_OVERRUN);
                ok = false;
                break;
            }
        }
        if (ok && lane == 0) {
            s_chan_end[0] = pos;
            lds_store_rel(&s_pub[0], kDone);
        }
        __builtin_amdgcn_s_setprio(0);
    }

    if (wave < W) {
        // ---- IO, QE, QO: block b on wave b mod W
        __builtin_amdgcn_s_setprio(2);
        for (int ch = 1; ch < 4; ++ch) {
            if (!wait_for<4>(&s_pub[ch - 1], kDone, &s_err, lane)) break;
            if (lds_load_acq(&s_err) != S1G_PKT_OK) break;
            // IE's length of block b, the estimate for every channel's block b
            auto ie_len = [&](int b) __attribute__((always_inline)) {
                const int e = b + 1 < nb ? bstart[b + 1] - 3 : s_chan_end[0];
                return __builtin_amdgcn_readfirstlane(e - bstart[b]);
            };
            int s0 = (__builtin_amdgcn_readfirstlane(s_chan_end[ch - 1]) + 15) & ~15;
            const int hbits = ch == 2 ? 8 : 0;
            if (nb == 0) {  // no blocks: the channel is empty
                if (wave == 0 && lane == 0) {
                    s_chan_end[ch] = s0;
                    lds_store_rel(&s_pub[ch], kDone);
                }
                continue;
            }
            if (ch == 2) {
                if (s0 + 8 > nbits) {
                    if (wave == 0) fail(S1G_PKT_ERR_OVERRUN);
                    break;
                }
                if (wave == 0 && lane == 0) bthr[0] = (uint8_t)peek_bits(words, s0, 8);
                s0 += 8;
            }
            if (wave == 0 && lane == 0) {
                bstart[ch * max_blocks] = s0;
                lds_store_rel(&s_pub[ch], 1);
            }
            // the latest block start this wave knows: block kb starts at base
            int base = s0, kb = 0;
            bool stop = false;
            for (int b = wave; b < nb && !stop; b += W) {
                const int brc = __builtin_amdgcn_readfirstlane((int)bbrc[b]);
                const int nsym = (b == nb - 1) ? nq - 128 * (nb - 1) : 128;
                uint32_t hdr = 0;
                int end;
                if (b == 0) {
                    const int delta = min(max((ie_len(0) + 63) / 64 + 1, 2), 32);
                    end = walk_any(brc, s0, nsym, delta, ch == 2, &hdr);
                } else {
                    // predicted from block kb's start over IE's lengths of
                    // blocks kb .. b - 1; clamped to the data: a region from
                    // at most nbits reads no further than the staged zero
                    // words reach
                    int est = base;
                    for (int q = kb; q < b; ++q) est += ie_len(q) + hbits;
                    const int margin =
                        kSpecMargin + (W == kWalkersLong ? kSpecMarginStepLong : kSpecMarginStep) * (b - kb - 1);
                    const int pred = min(max(est - margin, base), nbits);
                    const int delta = min(max((ie_len(b) + 2 * margin + 16 + 63) / 64 + 1, 2), 32);
                    const uint32_t dmagic = 0xFFFFFFFFu / (uint32_t)delta + 1u;
                    Region r;
                    int pre;
                    switch (brc) {
                        case 0: region_walk<0, ALL>(words, pred, delta, lane, r); region_settle<0>(r, 0, 0, delta, lane); break;
                        case 1: region_walk<1, ALL>(words, pred, delta, lane, r); region_settle<1>(r, 0, 0, delta, lane); break;
                        case 2: region_walk<2, ALL>(words, pred, delta, lane, r); region_settle<2>(r, 0, 0, delta, lane); break;
                        case 3: region_walk<3, ALL>(words, pred, delta, lane, r); region_settle<3>(r, 0, 0, delta, lane); break;
                        default: region_walk<4, ALL>(words, pred, delta, lane, r); region_settle<4>(r, 0, 0, delta, lane); break;
                    }
                    pre = region_prefix(r, delta, lane);
                    if (!wait_start(&s_pub[ch], b + 1, &s_err, lane)) break;
                    const int start = __builtin_amdgcn_readfirstlane(bstart[ch * max_blocks + b]);
                    switch (brc) {
                        case 0: end = resolve_block<0, ALL>(words, nbits, r, pre, pred, start, nsym, delta, dmagic, lane, ch == 2, &hdr); break;
                        case 1: end = resolve_block<1, ALL>(words, nbits, r, pre, pred, start, nsym, delta, dmagic, lane, ch == 2, &hdr); break;
                        case 2: end = resolve_block<2, ALL>(words, nbits, r, pre, pred, start, nsym, delta, dmagic, lane, ch == 2, &hdr); break;
                        case 3: end = resolve_block<3, ALL>(words, nbits, r, pre, pred, start, nsym, delta, dmagic, lane, ch == 2, &hdr); break;
                        default: end = resolve_block<4, ALL>(words, nbits, r, pre, pred, start, nsym, delta, dmagic, lane, ch == 2, &hdr); break;
                    }
                }
                if (end > nbits) {
                    fail(S1G_PKT_ERR_OVERRUN);
                    stop = true;
                    break;
                }
                int next = end;
                if (b + 1 < nb) {
                    if (ch == 2) {
                        if (next + 8 > nbits) {
                            fail(S1G_PKT_ERR_OVERRUN);
                            stop = true;
                            break;
                        }
                        if (lane == 0) bthr[b + 1] = (uint8_t)hdr;
                        next += 8;
                    }
                    if (lane == 0) {
                        bstart[ch * max_blocks + b + 1] = next;
                        lds_store_rel(&s_pub[ch], b + 2);
                    }
                } else if (lane == 0) {
                    s_chan_end[ch] = end;
                    lds_store_rel(&s_pub[ch], kDone);
                }
                base = next;
                kb = b + 1;
            }
            if (stop) break;
        }
        __builtin_amdgcn_s_setprio(0);
    }

    // ---- values, by groups of 16 blocks once their QO starts are published
    {
        const int n_groups = (nb + 15) / 16;
        const bool walked_qe = wait_for<16>(&s_pub[2], kDone, &s_err, lane);
        for (;;) {
            if (!walked_qe) break;
            int g = 0;
            if (lane == 0)
                g = __hip_atomic_fetch_add(&s_next_group, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            g = __builtin_amdgcn_readfirstlane(g);
            if (g >= n_groups) break;
            if (!wait_for<8>(&s_pub[3], min(16 * g + 16, nb), &s_err, lane)) break;
            const int b = 16 * g + (lane >> 2);
            if (b < nb) decode_item(words, bstart, bbrc, bthr, nrl, max_blocks, nq, nb, lane & 3, b, orow);
        }
    }
    __syncthreads();
    st = s_err;
    const int N = 2 * (nq > 0 ? nq : 0);
    if (st != S1G_PKT_OK) {
        for (int i = tid; i < 2 * N && i < 2 * ld_out; i += kThreads) orow[i] = (T)0;
        if (tid == 0) status[p] = st;
        return;
    }
    if (tid == 0) status[p] = S1G_PKT_OK;
}

// One launch over the packets whose length l satisfies lo < l <= hi, its
// LDS sized for max_bytes.
static int launch_decode_class(hipStream_t stream, int prec, const uint8_t* ud, const uint64_t* offs,
                               const uint32_t* lens, const int32_t* nqs, int n_packets, int max_bytes,
                               int max_quads, void* out, int64_t ld_out, int32_t* status, uint32_t lo,
                               uint32_t hi) {
    const int max_words = (max_bytes + 3) / 4;
    const int lds_words = std::min(max_words + kPadWords, kLdsWords);
    const int max_blocks = (max_quads + 127) / 128 + 1;
    // staged words, block starts, THIDX and BRC bytes, then the 80 levels at
    // the same word offset the kernel computes
    const int nrl_words = (lds_words + 4 * max_blocks + (2 * max_blocks + 3) / 4 + 1) & ~1;
    const size_t lds = sizeof(uint32_t) * (size_t)nrl_words + sizeof(double) * 80;
    if (lds > 160 * 1024) return -1;
    const bool all = max_words + kPadWords <= kLdsWords;
    // the long class of a split burst (lo > 0) takes more walking waves
    const int walkers = lo > 0 ? kWalkersLong : kWalkers;
#define S1G_LAUNCH(T, A, W)                                                                                    \
    hipLaunchKernelGGL((fdbaq_decode_kernel<T, A, W>), dim3(n_packets), dim3(64 * (W + 1)), lds, stream, ud, offs, \
                       lens, nqs, n_packets, max_words, lds_words, max_blocks, max_quads, (T*)out, ld_out, status,   \
                       lo, hi)
#define S1G_LAUNCH_W(T, A)                      \
    do {                                        \
        if (walkers == kWalkersLong)            \
            S1G_LAUNCH(T, A, kWalkersLong);     \
        else                                    \
            S1G_LAUNCH(T, A, kWalkers);         \
    } while (0)
    if (prec == S1G_F32) {
        if (all) S1G_LAUNCH_W(float, true);
        else S1G_LAUNCH_W(float, false);
    } else {
        if (all) S1G_LAUNCH_W(double, true);
        else S1G_LAUNCH_W(double, false);
    }
#undef S1G_LAUNCH_W
#undef S1G_LAUNCH
    return (int)hipGetLastError();
}

int launch_fdbaq_decode(hipStream_t stream, int prec, const uint8_t* ud, const uint64_t* offs, const uint32_t* lens,
                        const int32_t* nqs, int n_packets, int max_bytes, int max_quads, void* out, int64_t ld_out,
                        int32_t* status, uint32_t lo_bytes, uint32_t hi_bytes) {
    if (n_packets <= 0) return 0;
    return launch_decode_class(stream, prec, ud, offs, lens, nqs, n_packets, max_bytes, max_quads, out, ld_out,
                               status, lo_bytes, hi_bytes);
}

}  // namespace s1g
