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
// further bits, which the tables below encode.
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

// Length in bits of the Huffman code at the top of `code10` (the 10 bits
// after the sign bit).
S1G_HD int fdbaq_code_len(int brc, uint32_t code10) {
    const int clo = fdbaq_clo(code10);
    const uint64_t t = brc == 0 ? kLen0 : brc == 1 ? kLen1 : brc == 2 ? kLen2 : brc == 3 ? kLen3 : kLen4;
    int len = (int)((t >> (4 * clo)) & 15u);
    if (brc == 4 && clo == 0) len += (int)((code10 >> 8) & 1u);
    return len;
}

// Magnitude code m of the Huffman code at the top of `code10`.
S1G_HD int fdbaq_mcode(int brc, uint32_t code10) {
    const int clo = fdbaq_clo(code10);
    const int b1 = (int)((code10 >> 8) & 1u), b2 = (int)((code10 >> 7) & 1u);
    const int b3 = (int)((code10 >> 6) & 1u), b7 = (int)((code10 >> 2) & 1u), b8 = (int)((code10 >> 1) & 1u);
    switch (brc) {
        case 0: return clo < 3 ? clo : 3;
        case 1: return clo < 4 ? clo : 4;
        case 2: return clo < 6 ? clo : 6;
        case 3: return clo == 0 ? b1 : clo == 1 ? 2 : (clo + 1 < 9 ? clo + 1 : 9);
        default:
            switch (clo) {
                case 0: return b1 ? 1 + b2 : 0;
                case 1: return 3 + b2;
                case 2: return 5 + b3;
                case 3: return 7;
                case 4: return 8;
                case 5: return 9;
                case 6: return 10 + b7;
                default: return 12 + 2 * b7 + b8;
            }
    }
}

// ----------------------------------------------------- reconstruction law
// packet.cpp:678-693 of the reference: for THIDX <= kThidxLimit[brc] the
// sample is +-m (m below the flag) or +-SIMPLE[brc][THIDX] (m at the flag);
// otherwise +-NRL[brc][m] * SF[THIDX], one FP64 product.
constexpr int kThidxLimit[5] = {3, 3, 5, 6, 8};
constexpr int kMFlag[5] = {3, 4, 6, 9, 15};

}  // namespace s1g
