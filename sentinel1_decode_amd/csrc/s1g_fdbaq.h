// s1g_fdbaq.h -- FDBAQ (BAQ modes 12-14, data format D) symbol tables and
// symbol decoding shared by the device decoder and the host-side checks.
//
// Symbol = sign bit followed by the Huffman code of the magnitude code m for
// the block's bit-rate code (BRC 0-4).  Code words (S1 SPPDU Figs. 4-7..4-11,
// structs.h:170-227 of the reference):
//   BRC0: 0 10 110 111                                  -> m 0..3
//   BRC1: 0 10 110 1110 1111                            -> m 0..4
//   BRC2: 0 10 110 1110 11110 111110 111111             -> m 0..6
//   BRC3: 00 01 10 110 1110 11110 111110 1111110 11111110 11111111 -> m 0..9
//   BRC4: 00 010 011 100 101 1100 1101 1110 11110 111110 11111100 11111101
//         111111100 111111101 111111110 111111111       -> m 0..15
// Each code is decided by its count of leading ones (clo) plus at most two
// further bits, which the tables below encode:
//   BRC0-2: m = min(clo, flag);  BRC3: '0x' -> x, '10' -> 2, clo >= 2 -> min(clo + 1, 9);
//   BRC4: '00' -> 0, '01x' -> 1 + x, '10x' -> 3 + x, '110x' -> 5 + x, clo 3..5 -> 7..9,
//         '1111110x' -> 10 + x, '1111111xy' -> 12 + 2x + y.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define S1G_HD __host__ __device__ __forceinline__
#else
#define S1G_HD inline
#endif

namespace s1g {

// Code length by clo (0..10), 4 bits per entry, entry k at bits [4k, 4k+4).
// BRC3 clo=0 and BRC4 clo=0 are the 2-bit '0x' cases (BRC4 '01x' adds one).
#define S1G_NIB11(a, b, c, d, e, f, g, h, i, j, k)                                                   \
    ((uint64_t)(a) | ((uint64_t)(b) << 4) | ((uint64_t)(c) << 8) | ((uint64_t)(d) << 12) |          \
     ((uint64_t)(e) << 16) | ((uint64_t)(f) << 20) | ((uint64_t)(g) << 24) | ((uint64_t)(h) << 28) | \
     ((uint64_t)(i) << 32) | ((uint64_t)(j) << 36) | ((uint64_t)(k) << 40))

constexpr uint64_t kLen0 = S1G_NIB11(1, 2, 3, 3, 3, 3, 3, 3, 3, 3, 3);
constexpr uint64_t kLen1 = S1G_NIB11(1, 2, 3, 4, 4, 4, 4, 4, 4, 4, 4);
constexpr uint64_t kLen2 = S1G_NIB11(1, 2, 3, 4, 5, 6, 6, 6, 6, 6, 6);
constexpr uint64_t kLen3 = S1G_NIB11(2, 2, 3, 4, 5, 6, 7, 8, 8, 8, 8);
constexpr uint64_t kLen4 = S1G_NIB11(2, 3, 4, 4, 5, 6, 8, 9, 9, 9, 9);

// count of leading ones of a 10-bit code (0..10)
S1G_HD int fdbaq_clo(uint32_t code10) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __clz(~(code10 << 22));
#else
    return __builtin_clz(~(code10 << 22));
#endif
}

// Branch-free rule used by the decoder.  Code length: 1 (sign) + table
// entry by clo, plus the second code bit for BRC 4's '0x' codes.
S1G_HD uint64_t fdbaq_len_table(int brc) {
    return brc == 0 ? kLen0 : brc == 1 ? kLen1 : brc == 2 ? kLen2 : brc == 3 ? kLen3 : kLen4;
}
// v: sign bit followed by the 10 code bits; x4 = 1 for BRC 4.
S1G_HD int fdbaq_sym_len(uint32_t code10, int clo, uint64_t len_tab, uint32_t x4) {
    const int len = (int)((len_tab >> (4 * clo)) & 15u);
    const uint32_t extra = clo == 0 ? ((code10 >> 8) & x4) : 0u;
    return 1 + len + (int)extra;
}
// Magnitude code: m = base[clo] + ((code10 >> shift[clo]) & mask[clo]) with
// 4-bit base and shift entries and 2-bit masks by clo; BRC 4 with clo 0
// ('00' -> 0, '01x' -> 1 + x) is the one exception.
constexpr uint64_t kMBase[5] = {0x33333333210ull, 0x44444443210ull, 0x66666543210ull, 0x99987654320ull,
                                0xcccca987530ull};
constexpr uint64_t kMShift[5] = {0ull, 0ull, 0ull, 0x8ull, 0x11112000670ull};
constexpr uint32_t kMMask[5] = {0u, 0u, 0u, 0x1u, 0x3fd014u};
S1G_HD int fdbaq_mcode_tab(uint32_t code10, int clo, uint64_t base, uint64_t shift, uint32_t mask, uint32_t x4) {
    const int b = (int)((base >> (4 * clo)) & 15u);
    const int sh = (int)((shift >> (4 * clo)) & 15u);
    const int mk = (int)((mask >> (2 * clo)) & 3u);
    int m = b + (int)((code10 >> sh) & (uint32_t)mk);
    const int v2 = (int)((code10 >> 7) & 3u);
    const int m4 = v2 >= 2 ? v2 - 1 : 0;
    return (x4 && clo == 0) ? m4 : m;
}

// ----------------------------------------------------- reconstruction law
// packet.cpp:678-693 of the reference: for THIDX <= kThidxLimit[brc] the
// sample is +-m (m below the flag) or +-SIMPLE[brc][THIDX] (m at the flag);
// otherwise +-NRL[brc][m] * SF[THIDX], one FP64 product.
constexpr int kThidxLimit[5] = {3, 3, 5, 6, 8};
constexpr int kMFlag[5] = {3, 4, 6, 9, 15};

}  // namespace s1g
