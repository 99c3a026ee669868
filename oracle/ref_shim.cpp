// C-ABI shim over the reference's own bit reader and Huffman decoder.
// Built only where the reference sources are present (never shipped as
// product): g++ compiles decoding_utils.cpp from its place in the reference
// tree together with this file into oracle/_ref/libref_decoding.so.
// TEST INFRASTRUCTURE ONLY.
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "decoding_utils.h"

extern "C" {

// decoding_utils.cpp:95-134.  Returns 0 on success, 2 where it throws.
int ref_read_n_bits(const uint8_t* data, int64_t nbytes, int start_bit, int n, uint64_t* out) {
    UINT8_VEC_1D v(data, data + nbytes);
    try {
        *out = read_n_bits(v, start_bit, n);
    } catch (const std::out_of_range&) {
        return 2;
    } catch (const std::invalid_argument&) {
        return 5;
    }
    return 0;
}

// decoding_utils.cpp:44-77.  Returns 0 on success, 2 on a read past the end,
// 3 on a code longer than 10 bits.
int ref_huffman_decode(const uint8_t* data, int64_t nbytes, int brc, int* bit_index, int* m_code) {
    UINT8_VEC_1D v(data, data + nbytes);
    try {
        int b = brc;
        *m_code = huffman_decode(v, b, *bit_index);
    } catch (const std::out_of_range& e) {
        std::string what = e.what();
        return what.find("Max bit length") != std::string::npos ? 3 : 2;
    }
    return 0;
}

}

extern "C" {

// The reference's header field widths (structs.h:42-51, :68-115).
int ref_header_widths(int* primary, int* secondary) {
    for (size_t i = 0; i < PRIMARY_HEADER.size(); ++i) primary[i] = PRIMARY_HEADER[i];
    for (size_t i = 0; i < SECONDARY_HEADER.size(); ++i) secondary[i] = SECONDARY_HEADER[i];
    return (int)(PRIMARY_HEADER.size() * 1000 + SECONDARY_HEADER.size());
}

// parse_packet_header (packet.cpp:829-846): consecutive read_n_bits over the
// reference's width tables; bytes = 6 primary + 62 secondary header bytes.
int ref_parse_headers(const uint8_t* bytes, int64_t* primary, int64_t* secondary) {
    UINT8_VEC_1D p(bytes, bytes + 6), s(bytes + 6, bytes + 68);
    try {
        int bit = 0;
        for (size_t i = 0; i < PRIMARY_HEADER.size(); ++i) {
            primary[i] = (int)read_n_bits(p, bit, PRIMARY_HEADER[i]);
            bit += PRIMARY_HEADER[i];
        }
        bit = 0;
        for (size_t i = 0; i < SECONDARY_HEADER.size(); ++i) {
            secondary[i] = (int)read_n_bits(s, bit, SECONDARY_HEADER[i]);
            bit += SECONDARY_HEADER[i];
        }
    } catch (const std::exception&) {
        return 2;
    }
    return 0;
}

// The FDBAQ decode tables and constants as the reference holds them
// (structs.h:21-38,170-354,561-574), flattened for comparison:
//   thidx[5], mflag[5], simple[5*10], nrl[5*16], sf[256], rgdec[12],
//   consts = {PI, SPEED_OF_LIGHT, CENTER_FREQ, WAVELENGTH, F_REF,
//             DELTA_T_SUPPRESSED, LATCH_TIME_OFFSET, WORD_SIZE},
//   huff[5*16*2] = (code value, m) pairs, huff_n[5], huff_start[5].
void ref_tables(int* thidx, int* mflag, double* simple, double* nrl, double* sf, double* rgdec, double* consts,
                int* huff, int* huff_n, int* huff_start) {
    for (int b = 0; b < 5; ++b) {
        thidx[b] = BRC_TO_THIDX[b];
        mflag[b] = BRC_TO_M_CODE[b];
        huff_start[b] = BRC_TO_HUFFMAN_START_BIT_LEN[b];
        for (int k = 0; k < 10; ++k) simple[b * 10 + k] = k < (int)SIMPLE_RECONSTRUCTION[1][b].size() ? SIMPLE_RECONSTRUCTION[1][b][k] : 0.0;
        for (int k = 0; k < 16; ++k) nrl[b * 16 + k] = NORMALIZED_RECONSTRUCTION[1][b][k];
        int n = 0;
        for (const auto& kv : HUFFMAN_CODINGS[b]) {
            huff[(b * 16 + n) * 2] = kv.first;
            huff[(b * 16 + n) * 2 + 1] = kv.second;
            ++n;
        }
        huff_n[b] = n;
    }
    for (int t = 0; t < 256; ++t) sf[t] = THIDX_TO_SF[t];
    for (int r = 0; r < 12; ++r) rgdec[r] = RANGE_DECIMATION[r];
    consts[0] = PI;
    consts[1] = SPEED_OF_LIGHT;
    consts[2] = CENTER_FREQ;
    consts[3] = WAVELENGTH;
    consts[4] = F_REF;
    consts[5] = DELTA_T_SUPPRESSED;
    consts[6] = LATCH_TIME_OFFSET;
    consts[7] = WORD_SIZE;
}

}
