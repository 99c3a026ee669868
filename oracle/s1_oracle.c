/*
 * s1_oracle.c -- FP64 CPU restatement of the Sentinel-1 L0 -> SLC hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker and the timed CPU baseline).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library; the product never does.
 *
 * Each function follows the reference function cited beside it (file:line in
 * AndrewPlayer3/sentinel1_decode src/).  Arithmetic is kept in the
 * reference's operation order where rounding matters (the large phases), and
 * the file is compiled with -ffp-contract=off so no FMA is formed.
 * pow(x, 2.0) of the reference is written x*x (the correctly rounded square).
 */
#include "s1_oracle.h"

/* Planted mistakes, for the tests that show the geometry and azimuth checks
 * catch them (they build variants with -DORC_MUTATION=k into a scratch
 * directory).  The oracle itself is built with ORC_MUTATION 0:
 *   1 sign of 4 PI R D / lambda flipped     2 factor 2 inside D
 *   3 amplitude weight dropped              4 sign of the t_corr phase flipped
 *   5 RCMC shift sign flipped               6 f_dc left out of D
 *   7 az_freqs spacing PRF/L (not PRF/(L-1)) 8 timing-correction j term sign flipped
 *   9 slant-range linspace without its endpoint */
#ifndef ORC_MUTATION
#define ORC_MUTATION 0
#endif

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Constants (structs.h:21-38)                                              */
/* ------------------------------------------------------------------------ */
static const double K_PI = 3.14159265358979; /* truncated as in structs.h:23 */
static const double K_C = 299792458.0;
static const double K_F0 = 5.405000454334350e+9;
static const double K_FREF = 37.53472224;
#define K_WAVELENGTH (K_C / K_F0)
#define K_DT_SUPPRESSED ((320 / (8 * K_FREF)) * 1e-6)
#define K_LATCH (1.439 * 1e-6)

/* Range decimation code -> sample rate in MHz (structs.h:561-574). */
static const double K_RGDEC[12] = {112.6041667, 100.0925926, 0.0,         83.41049387,
                                   66.72839509, 56.30208336, 50.04629632, 25.02314816,
                                   64.34523813, 46.91840280, 17.32371796, 54.59595962};

/* ------------------------------------------------------------------------ */
/* FDBAQ decode tables (structs.h:170-354): FDBAQ rows only                  */
/* ------------------------------------------------------------------------ */
static const int K_HUFF_START[5] = {1, 1, 1, 2, 2};
/* code-word values (as read MSB-first) and their m-codes, per BRC */
static const int K_HUFF_NCODES[5] = {4, 5, 7, 10, 16};
static const uint16_t K_HUFF_VAL[5][16] = {
    {0, 2, 6, 7},
    {0, 2, 6, 14, 15},
    {0, 2, 6, 14, 30, 62, 63},
    {0, 1, 2, 6, 14, 30, 62, 126, 254, 255},
    {0, 2, 3, 4, 5, 12, 13, 14, 30, 62, 252, 253, 508, 509, 510, 511}};
static const uint8_t K_HUFF_M[5][16] = {
    {0, 1, 2, 3},
    {0, 1, 2, 3, 4},
    {0, 1, 2, 3, 4, 5, 6},
    {0, 1, 2, 3, 4, 5, 6, 7, 8, 9},
    {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15}};
static const int K_BRC_THIDX[5] = {3, 3, 5, 6, 8};
static const int K_BRC_MFLAG[5] = {3, 4, 6, 9, 15};
static const double K_SIMPLE[5][10] = {
    {3.0000, 3.0000, 3.1600, 3.5300},
    {4.0000, 4.0000, 4.0800, 4.3700},
    {6.0000, 6.0000, 6.0000, 6.1500, 6.5000, 6.8800},
    {9.0000, 9.0000, 9.0000, 9.0000, 9.3600, 9.5000, 10.1000},
    {15.0000, 15.0000, 15.0000, 15.0000, 15.0000, 15.0000, 15.2200, 15.5000, 16.0600}};
static const double K_NRL[5][16] = {
    {0.3637, 1.0915, 1.8208, 2.6406},
    {0.3042, 0.9127, 1.5216, 2.1313, 2.8426},
    {0.2305, 0.6916, 1.1528, 1.6140, 2.0754, 2.5369, 3.1191},
    {0.1702, 0.5107, 0.8511, 1.1916, 1.5321, 1.8726, 2.2131, 2.5536, 2.8942, 3.3744},
    {0.1130, 0.3389, 0.5649, 0.7908, 1.0167, 1.2428, 1.4687, 1.6947, 1.9206, 2.1466,
     2.3725, 2.5985, 2.8244, 3.0504, 3.2764, 3.6623}};
static const double K_SF[256] = {
    0.00,   0.63,   1.25,   1.88,   2.51,   3.13,   3.76,   4.39,   5.01,   5.64,   6.27,
    6.89,   7.52,   8.15,   8.77,   9.40,   10.03,  10.65,  11.28,  11.91,  12.53,  13.16,
    13.79,  14.41,  15.04,  15.67,  16.29,  16.92,  17.55,  18.17,  18.80,  19.43,  20.06,
    20.68,  21.31,  21.93,  22.56,  23.19,  23.81,  24.44,  25.06,  25.69,  26.32,  26.94,
    27.57,  28.20,  28.83,  29.45,  30.08,  30.71,  31.33,  31.96,  32.59,  33.21,  33.84,
    34.47,  35.09,  35.72,  36.35,  36.97,  37.60,  38.23,  38.85,  39.48,  40.11,  40.73,
    41.36,  41.99,  42.61,  43.24,  43.87,  44.49,  45.12,  45.75,  46.37,  47.00,  47.63,
    48.25,  48.88,  49.51,  50.13,  50.76,  51.39,  52.01,  52.64,  53.27,  53.89,  54.52,
    55.15,  55.77,  56.40,  57.03,  57.65,  58.28,  58.91,  59.54,  60.16,  60.79,  61.41,
    62.04,  62.98,  64.24,  65.49,  66.74,  68.00,  69.25,  70.50,  71.76,  73.01,  74.26,
    75.52,  76.77,  78.02,  79.28,  80.53,  81.78,  83.04,  84.29,  85.54,  86.80,  88.05,
    89.30,  90.56,  91.81,  93.07,  94.32,  95.57,  96.83,  98.08,  99.33,  100.58, 101.84,
    103.09, 104.34, 105.60, 106.85, 108.10, 109.36, 110.61, 111.86, 113.11, 114.37, 115.62,
    116.87, 118.13, 119.38, 120.63, 121.89, 123.14, 124.39, 125.65, 126.90, 128.15, 129.41,
    130.66, 131.91, 133.17, 134.42, 135.67, 136.93, 138.18, 139.43, 140.69, 141.94, 143.19,
    144.45, 145.70, 146.95, 148.21, 149.46, 150.71, 151.97, 153.22, 154.47, 155.73, 156.98,
    158.23, 159.49, 160.74, 161.99, 163.25, 164.50, 165.75, 167.01, 168.26, 169.51, 170.77,
    172.02, 173.27, 174.53, 175.78, 177.03, 178.29, 179.54, 180.79, 182.05, 183.30, 184.55,
    185.81, 187.06, 188.31, 189.57, 190.82, 192.07, 193.33, 194.58, 195.83, 197.09, 198.34,
    199.59, 200.85, 202.10, 203.35, 204.61, 205.86, 207.11, 208.37, 209.62, 210.87, 212.13,
    213.38, 214.63, 215.89, 217.14, 218.39, 219.65, 220.90, 222.15, 223.41, 224.66, 225.91,
    227.17, 228.42, 229.67, 230.93, 232.18, 233.43, 234.69, 235.94, 237.19, 238.45, 239.70,
    240.95, 242.21, 243.46, 244.71, 245.97, 247.22, 248.47, 249.73, 250.98, 252.23, 253.49,
    254.74, 255.99, 255.99};

/* ------------------------------------------------------------------------ */
/* Bit reader: decoding_utils.cpp:95-134 (big-endian, MSB first, throws past */
/* the end).  Returns ORC_ERR_OVERRUN where the reference throws.            */
/* ------------------------------------------------------------------------ */
int orc_read_n_bits(const uint8_t* data, int64_t nbytes, int64_t start_bit, int n,
                    uint64_t* out) {
    if (n < 1 || n > 64) return ORC_ERR_ARG;
    int64_t byte_index = start_bit / 8;
    int bit_offset = (int)(start_bit % 8);
    if (byte_index >= nbytes) return ORC_ERR_OVERRUN;
    uint64_t result = 0;
    int bits_read = 0;
    while (bits_read < n) {
        if (byte_index >= nbytes) return ORC_ERR_OVERRUN;
        int left = 8 - bit_offset;
        int take = (n - bits_read) < left ? (n - bits_read) : left;
        uint8_t mask = (uint8_t)((1u << take) - 1u);
        uint8_t part = (uint8_t)((data[byte_index] >> (left - take)) & mask);
        result = (result << take) | part;
        bits_read += take;
        bit_offset = 0;
        byte_index++;
    }
    *out = result;
    return ORC_OK;
}

static int huff_lookup(int brc, uint64_t bits, int* m) {
    for (int i = 0; i < K_HUFF_NCODES[brc]; i++)
        if (K_HUFF_VAL[brc][i] == bits) {
            *m = K_HUFF_M[brc][i];
            return 1;
        }
    return 0;
}

/* Prefix walk of decoding_utils.cpp:44-77: start at the BRC's shortest length,
 * grow one bit until the value read is a code word of the BRC (at most 10
 * bits); BRC 4 first tests for the 2-bit code '00'. */
int orc_huffman_decode(const uint8_t* data, int64_t nbytes, int brc, int64_t* bit_index,
                       int* m_code) {
    if (brc < 0 || brc > 4) return ORC_ERR_BRC;
    int bit_len = K_HUFF_START[brc];
    uint64_t bits = 0;
    int st;
    if (brc == 4) {
        st = orc_read_n_bits(data, nbytes, *bit_index, bit_len, &bits);
        if (st) return st;
        if (bits == 0) {
            *bit_index += 2;
            *m_code = 0;
            return ORC_OK;
        }
        bit_len += 1;
    }
    int m = -1;
    for (int i = 0; i < 10; i++) {
        st = orc_read_n_bits(data, nbytes, *bit_index, bit_len, &bits);
        if (st) return st;
        if (!huff_lookup(brc, bits, &m))
            bit_len += 1;
        else
            break;
        if (bit_len > 10) return ORC_ERR_HUFFMAN;
    }
    *bit_index += bit_len;
    *m_code = m;
    return ORC_OK;
}

/* packet.cpp:678-693: simple reconstruction for small THIDX, otherwise
 * normalised reconstruction level times sigma factor, evaluated as
 * pow(-1, sign) * NRL * SF from left to right (one rounding). */
double orc_s_value_d(int brc, int thidx, int sign, int m_code, int* status) {
    double s = sign ? -1.0 : 1.0;
    if (thidx <= K_BRC_THIDX[brc]) {
        int flag = K_BRC_MFLAG[brc];
        if (m_code < flag) return s * (double)m_code;
        if (m_code == flag) return s * K_SIMPLE[brc][thidx];
        *status = ORC_ERR_MCODE;
        return 0.0;
    }
    return s * K_NRL[brc][m_code] * K_SF[thidx];
}

/* Type-D decode of one packet: packet.cpp:404-453 (channel order IE, IO, QE,
 * QO), _set_quad_type_d :779-818 (3-bit BRC per block in IE, 8-bit THIDX per
 * block in QE, 16-bit word alignment after each channel), _get_h_code_type_d
 * :755-776 (sign bit then Huffman code), _get_signal_type_d :696-752
 * (out[2k] = IE_k + i QE_k, out[2k+1] = IO_k + i QO_k).
 * `out` receives 2*num_quads complex values (4*num_quads doubles). */
int orc_decode_packet_d(const uint8_t* ud, int64_t nbytes, int num_quads, double* out,
                        uint8_t* brc_out, uint8_t* thidx_out) {
    if (num_quads < 0) return ORC_ERR_ARG;
    int nb = (int)ceil((2.0 * (double)num_quads) / 256.0);
    uint8_t* brc = (uint8_t*)calloc((size_t)nb + 1, 1);
    uint8_t* thr = (uint8_t*)calloc((size_t)nb + 1, 1);
    uint8_t* sg = (uint8_t*)malloc((size_t)num_quads * 4 + 4);
    uint8_t* mc = (uint8_t*)malloc((size_t)num_quads * 4 + 4);
    int64_t bit = 0;
    int st = ORC_OK;
    for (int ch = 0; ch < 4 && st == ORC_OK; ch++) {
        for (int b = 0; b < nb && st == ORC_OK; b++) {
            uint64_t v;
            if (ch == 0) {
                st = orc_read_n_bits(ud, nbytes, bit, 3, &v);
                if (st) break;
                if (v > 4) {
                    st = ORC_ERR_BRC;
                    break;
                }
                brc[b] = (uint8_t)v;
                bit += 3;
            } else if (ch == 2) {
                st = orc_read_n_bits(ud, nbytes, bit, 8, &v);
                if (st) break;
                thr[b] = (uint8_t)v;
                bit += 8;
            }
            int ncode = (b == nb - 1) ? num_quads - 128 * (nb - 1) : 128;
            for (int i = 0; i < ncode; i++) {
                uint64_t s;
                st = orc_read_n_bits(ud, nbytes, bit, 1, &s);
                if (st) break;
                bit += 1;
                int m;
                st = orc_huffman_decode(ud, nbytes, brc[b], &bit, &m);
                if (st) break;
                int64_t q = (int64_t)(128 * b + i);
                sg[q * 4 + ch] = (uint8_t)s;
                mc[q * 4 + ch] = (uint8_t)m;
            }
        }
        if (bit % 16) bit += 16 - (bit % 16);
    }
    if (st == ORC_OK) {
        for (int b = 0; b < nb && st == ORC_OK; b++) {
            int ncode = (b == nb - 1) ? num_quads - 128 * (nb - 1) : 128;
            for (int i = 0; i < ncode; i++) {
                int64_t q = (int64_t)(128 * b + i);
                double v[4];
                for (int ch = 0; ch < 4; ch++)
                    v[ch] = orc_s_value_d(brc[b], thr[b], sg[q * 4 + ch], mc[q * 4 + ch], &st);
                out[4 * q + 0] = v[0]; /* re out[2q]   = IE */
                out[4 * q + 1] = v[2]; /* im out[2q]   = QE */
                out[4 * q + 2] = v[1]; /* re out[2q+1] = IO */
                out[4 * q + 3] = v[3]; /* im out[2q+1] = QO */
            }
        }
    }
    if (brc_out)
        for (int b = 0; b < nb; b++) brc_out[b] = brc[b];
    if (thidx_out)
        for (int b = 0; b < nb; b++) thidx_out[b] = thr[b];
    free(brc);
    free(thr);
    free(sg);
    free(mc);
    return st;
}

/* Batch decode (s1_decoder.cpp:462-468 decodes packets in an OpenMP loop).
 * A failing packet gets its status and zeros instead of terminating. */
void orc_decode_packets(const uint8_t* base, const uint64_t* offsets, const uint32_t* lengths,
                        const int32_t* num_quads, int n_packets, double* out, int64_t ld,
                        int32_t* status) {
#pragma omp parallel for schedule(dynamic, 4)
    for (int p = 0; p < n_packets; p++) {
        double* row = out + 2 * ld * (int64_t)p;
        int st = orc_decode_packet_d(base + offsets[p], lengths[p], num_quads[p], row, NULL, NULL);
        if (st != ORC_OK) memset(row, 0, sizeof(double) * 4 * (size_t)(num_quads[p] > 0 ? num_quads[p] : 0));
        status[p] = st;
    }
}

/* ------------------------------------------------------------------------ */
/* Headers: field widths of structs.h:42-51 and :68-115, parsed MSB-first    */
/* like parse_packet_header (packet.cpp:829-846).                            */
/* ------------------------------------------------------------------------ */
static const int K_PRIM_W[8] = {3, 1, 1, 7, 4, 2, 14, 16};
static const int K_SEC_W[46] = {32, 16, 32, 32, 8, 1, 3, 4, 32, 8, 16, 32, 32, 1, 2, 5,
                                8,  8,  8,  8,  1, 15, 1, 15, 24, 3, 5, 24, 24, 24, 1, 3,
                                2,  2,  4,  2,  10, 2, 1,  5, 4,  3, 1, 8,  16, 8};
/* secondary field indices used below */
enum { S_BAQ_MODE = 15, S_BAQ_BLOCK = 16, S_RGDEC = 18, S_RXGAIN = 19, S_TXPRR_SIGN = 20,
       S_TXPRR_MAG = 21, S_TXPSF_SIGN = 22, S_TXPSF_MAG = 23, S_PULSE_LEN = 24, S_RANK = 26,
       S_PRI = 27, S_SWST = 28, S_SWL = 29, S_NQ = 44 };

int orc_parse_headers(const uint8_t* pkt, int64_t nbytes, int64_t* prim, int64_t* sec) {
    int64_t bit = 0;
    uint64_t v;
    for (int i = 0; i < 8; i++) {
        if (orc_read_n_bits(pkt, nbytes < 6 ? nbytes : 6, bit, K_PRIM_W[i], &v)) return ORC_ERR_OVERRUN;
        prim[i] = (int64_t)v;
        bit += K_PRIM_W[i];
    }
    bit = 0;
    if (nbytes < 68) return ORC_ERR_OVERRUN;
    for (int i = 0; i < 46; i++) {
        if (orc_read_n_bits(pkt + 6, 62, bit, K_SEC_W[i], &v)) return ORC_ERR_OVERRUN;
        sec[i] = (int64_t)v;
        bit += K_SEC_W[i];
    }
    return ORC_OK;
}

/* packet.cpp:29-100, 277-281, 376-380 */
void orc_packet_params(const int64_t* sec, double* out) {
    double txprr = (sec[S_TXPRR_SIGN] == 0 ? -1 : 1) * (double)sec[S_TXPRR_MAG] *
                   ((K_FREF * K_FREF) / 2097152);
    int sgn = sec[S_TXPSF_SIGN] == 0 ? -1 : 1;
    double txpsf = ((double)(sgn * (int)sec[S_TXPSF_MAG]) * (K_FREF / 16384)) + (txprr / (4 * K_FREF));
    double txpl = (double)sec[S_PULSE_LEN] / K_FREF;
    double fs = K_RGDEC[sec[S_RGDEC]];
    out[0] = (double)sec[S_PRI] / K_FREF;
    out[1] = txpl;
    out[2] = (double)sec[S_SWL] / K_FREF;
    out[3] = (double)sec[S_SWST] / K_FREF;
    out[4] = -0.5 * (double)sec[S_RXGAIN];
    out[5] = txprr;
    out[6] = txpsf;
    out[7] = fs;
    out[8] = (double)orc_replica_chirp_length(fs, txpl);
    out[9] = 8.0 * (double)(sec[S_BAQ_BLOCK] + 1);
}

/* signal_processing.cpp:18-29 */
int orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

void orc_linspace(double start, double end, int size, double* out) {
    double step = (end - start) / (size - 1);
    for (int i = 0; i < size; ++i) out[i] = start + i * step;
}

/* packet.cpp:284-301 */
void orc_slant_ranges(const int64_t* sec, int num_ranges, double* out) {
    int nq = (int)sec[S_NQ];
    if (num_ranges <= 0) num_ranges = 2 * nq;
    double n_rng = 2 * nq;
    double fs = K_RGDEC[sec[S_RGDEC]] * 1e+6;
    double start = ((double)sec[S_SWST] / K_FREF) * 1e-6;
    double pri = ((double)sec[S_PRI] / K_FREF) * 1e-6;
    double rank = (double)sec[S_RANK];
    double near = rank * pri + start + K_DT_SUPPRESSED;
    double far = near + (n_rng / fs);
#if ORC_MUTATION == 9
    for (int i = 0; i < num_ranges; i++) out[i] = near * K_C / 2 + i * ((far - near) * K_C / 2) / num_ranges;
    return;
#endif
    orc_linspace(near * K_C / 2, far * K_C / 2, num_ranges, out);
}

/* packet.cpp:322-342 */
void orc_timing_corrections(const int64_t* sec, double* out) {
    int n = 2 * (int)sec[S_NQ];
    double fs = K_RGDEC[sec[S_RGDEC]] * 1e+6;
    double swst = ((double)sec[S_SWST] / K_FREF) * 1e-6;
    double rank = (double)sec[S_RANK];
    double pri = ((double)sec[S_PRI] / K_FREF) * 1e-6;
    double dt = 1 / fs;
#if ORC_MUTATION == 8
    for (int i = 0; i < n; i++) out[i] = (swst + K_LATCH) + ((swst + (rank * pri) + ((double)i * dt)) / 2);
    return;
#endif
    for (int i = 0; i < n; i++) out[i] = (swst + K_LATCH) + ((swst + (rank * pri) - ((double)i * dt)) / 2);
}

/* ------------------------------------------------------------------------ */
/* FFT: FFTW_FORWARD = e^{-2 pi i jk/n}, FFTW_BACKWARD = e^{+...}, neither   */
/* normalised (signal_processing.cpp:516-578 apply 1/n themselves).  Own    */
/* mixed-radix Stockham with an exact-angle twiddle table; Bluestein's      */
/* chirp-z (Bluestein 1970) for lengths with a prime factor above 100.      */
/* ------------------------------------------------------------------------ */
typedef struct {
    int n, nrad, blue, m;
    int rad[40];
    double* tw;    /* n complex: e^{-2 pi i k / n} */
    double* bchirp; /* bluestein: n complex w_k = e^{-i pi k^2/n} */
    double* bfilt;  /* bluestein: FFT_m of conj chirp, m complex */
} fft_plan;

#define PLAN_CACHE 64
static fft_plan* g_plans[PLAN_CACHE];
static int g_nplans = 0;

static void exact_twiddles(double* tw, int n) {
    for (int k = 0; k < n; k++) {
        long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
        tw[2 * k] = (double)cosl(a);
        tw[2 * k + 1] = (double)sinl(a);
    }
}

static fft_plan* plan_create(int n);

static fft_plan* get_plan(int n) {
    fft_plan* p = NULL;
#pragma omp critical(orc_fft_plans)
    {
        for (int i = 0; i < g_nplans; i++)
            if (g_plans[i]->n == n) p = g_plans[i];
        if (!p) {
            p = plan_create(n);
            if (g_nplans < PLAN_CACHE) g_plans[g_nplans++] = p;
        }
    }
    return p;
}

static void stockham_run(fft_plan* p, double* x, double* y, int inverse);

static fft_plan* plan_create(int n) {
    fft_plan* p = (fft_plan*)calloc(1, sizeof(fft_plan));
    p->n = n;
    int r = n, big = 0;
    int tmp[40], nt = 0;
    while (r > 1 && nt < 40) {
        if (r % 4 == 0) { tmp[nt++] = 4; r /= 4; continue; }
        if (r % 2 == 0) { tmp[nt++] = 2; r /= 2; continue; }
        int f = 3;
        while (r % f) f += 2;
        if (f > 100) big = 1;
        tmp[nt++] = f;
        r /= f;
    }
    if (big) {
        p->blue = 1;
        int m = 1;
        while (m < 2 * n - 1) m <<= 1;
        p->m = m;
        p->bchirp = (double*)malloc(sizeof(double) * 2 * (size_t)n);
        for (int k = 0; k < n; k++) {
            long long q = ((long long)k * k) % (2LL * n);
            long double a = -3.14159265358979323846264338327950288L * (long double)q / (long double)n;
            p->bchirp[2 * k] = (double)cosl(a);
            p->bchirp[2 * k + 1] = (double)sinl(a);
        }
        double* f = (double*)calloc(2 * (size_t)m, sizeof(double));
        for (int k = 0; k < n; k++) {
            f[2 * k] = p->bchirp[2 * k];
            f[2 * k + 1] = -p->bchirp[2 * k + 1];
            if (k) {
                f[2 * (m - k)] = p->bchirp[2 * k];
                f[2 * (m - k) + 1] = -p->bchirp[2 * k + 1];
            }
        }
        fft_plan* pm = plan_create(m);
        double* w = (double*)malloc(sizeof(double) * 2 * (size_t)m);
        stockham_run(pm, f, w, 0);
        p->bfilt = f;
        free(w);
        free(pm->tw);
        free(pm);
        return p;
    }
    p->nrad = nt;
    memcpy(p->rad, tmp, sizeof(int) * (size_t)nt);
    p->tw = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    exact_twiddles(p->tw, n);
    return p;
}

/* one Stockham pass: out[(t/Ns)*Ns*R + j + r*Ns] = DFT_R(in[t + r*n/R] * w^{r j}) */
static void stockham_run(fft_plan* p, double* x, double* y, int inverse) {
    int n = p->n;
    double* src = x;
    double* dst = y;
    double sgn = inverse ? -1.0 : 1.0; /* conjugates the forward table */
    int Ns = 1;
    double a[2 * 128], b[2 * 128];
    double* ga = NULL;
    double* gb = NULL;
    for (int s = 0; s < p->nrad; s++) {
        int R = p->rad[s];
        double* va = a;
        double* vb = b;
        if (R > 128) {
            ga = (double*)realloc(ga, sizeof(double) * 2 * (size_t)R);
            gb = (double*)realloc(gb, sizeof(double) * 2 * (size_t)R);
            va = ga;
            vb = gb;
        }
        int nR = n / R;
        int tstep = n / (Ns * R);
        int rstep = n / R; /* table step of W_R */
        for (int t = 0; t < nR; t++) {
            int j = t % Ns;
            for (int r = 0; r < R; r++) {
                double xr = src[2 * (t + r * nR)], xi = src[2 * (t + r * nR) + 1];
                if (r && j) {
                    int ti = (r * j * tstep) % n;
                    double wr = p->tw[2 * ti], wi = sgn * p->tw[2 * ti + 1];
                    double tr = xr * wr - xi * wi;
                    xi = xr * wi + xi * wr;
                    xr = tr;
                }
                va[2 * r] = xr;
                va[2 * r + 1] = xi;
            }
            if (R == 2) {
                vb[0] = va[0] + va[2];
                vb[1] = va[1] + va[3];
                vb[2] = va[0] - va[2];
                vb[3] = va[1] - va[3];
            } else if (R == 4) {
                double s0r = va[0] + va[4], s0i = va[1] + va[5];
                double d0r = va[0] - va[4], d0i = va[1] - va[5];
                double s1r = va[2] + va[6], s1i = va[3] + va[7];
                double d1r = va[2] - va[6], d1i = va[3] - va[7];
                /* forward: X1 = d0 - i d1, X3 = d0 + i d1 */
                vb[0] = s0r + s1r;
                vb[1] = s0i + s1i;
                vb[4] = s0r - s1r;
                vb[5] = s0i - s1i;
                vb[2] = d0r + sgn * d1i;
                vb[3] = d0i - sgn * d1r;
                vb[6] = d0r - sgn * d1i;
                vb[7] = d0i + sgn * d1r;
            } else {
                for (int q = 0; q < R; q++) {
                    double sr = 0.0, si = 0.0;
                    for (int r = 0; r < R; r++) {
                        int ti = (int)(((long long)r * q % R) * rstep);
                        double wr = p->tw[2 * ti], wi = sgn * p->tw[2 * ti + 1];
                        sr += va[2 * r] * wr - va[2 * r + 1] * wi;
                        si += va[2 * r] * wi + va[2 * r + 1] * wr;
                    }
                    vb[2 * q] = sr;
                    vb[2 * q + 1] = si;
                }
            }
            int base = (t / Ns) * Ns * R + j;
            for (int r = 0; r < R; r++) {
                dst[2 * (base + r * Ns)] = vb[2 * r];
                dst[2 * (base + r * Ns) + 1] = vb[2 * r + 1];
            }
        }
        double* tmp = src;
        src = dst;
        dst = tmp;
        Ns *= R;
    }
    if (src != x) memcpy(x, src, sizeof(double) * 2 * (size_t)n);
    free(ga);
    free(gb);
}

static void bluestein_run(fft_plan* p, double* x, int inverse) {
    int n = p->n, m = p->m;
    fft_plan* pm = get_plan(m);
    double* u = (double*)calloc(2 * (size_t)m, sizeof(double));
    double* w = (double*)malloc(sizeof(double) * 2 * (size_t)m);
    double cs = inverse ? -1.0 : 1.0;
    for (int k = 0; k < n; k++) {
        double cr = p->bchirp[2 * k], ci = cs * p->bchirp[2 * k + 1];
        double xr = x[2 * k], xi = x[2 * k + 1];
        u[2 * k] = xr * cr - xi * ci;
        u[2 * k + 1] = xr * ci + xi * cr;
    }
    stockham_run(pm, u, w, 0);
    for (int k = 0; k < m; k++) {
        /* filter for the inverse is the conjugate chirp: FFT(conj h)[k] = conj(FFT(h)[-k]) */
        double fr, fi;
        if (!inverse) {
            fr = p->bfilt[2 * k];
            fi = p->bfilt[2 * k + 1];
        } else {
            int kk = (m - k) % m;
            fr = p->bfilt[2 * kk];
            fi = -p->bfilt[2 * kk + 1];
        }
        double ur = u[2 * k], ui = u[2 * k + 1];
        u[2 * k] = ur * fr - ui * fi;
        u[2 * k + 1] = ur * fi + ui * fr;
    }
    stockham_run(pm, u, w, 1);
    double sc = 1.0 / (double)m;
    for (int k = 0; k < n; k++) {
        double cr = p->bchirp[2 * k], ci = cs * p->bchirp[2 * k + 1];
        double ur = u[2 * k] * sc, ui = u[2 * k + 1] * sc;
        x[2 * k] = ur * cr - ui * ci;
        x[2 * k + 1] = ur * ci + ui * cr;
    }
    free(u);
    free(w);
}

void orc_fft(double* x, int n, int inverse) {
    if (n <= 1) return;
    fft_plan* p = get_plan(n);
    if (p->blue) {
        bluestein_run(p, x, inverse);
        return;
    }
    double* w = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    stockham_run(p, x, w, inverse);
    free(w);
}

void orc_fft_many(double* x, int n, int count, int64_t dist, int inverse) {
    if (n > 1) get_plan(n);
#pragma omp parallel for schedule(dynamic, 8)
    for (int i = 0; i < count; i++) orc_fft(x + 2 * dist * (int64_t)i, n, inverse);
}

/* ------------------------------------------------------------------------ */
/* Range compression                                                         */
/* ------------------------------------------------------------------------ */
/* packet.cpp:376-380 */
int orc_replica_chirp_length(double fs_mhz, double txpl_us) { return (int)floor(fs_mhz * txpl_us); }

/* packet.cpp:345-373: chirp of length N + 2 nc, non-zero on [m, m + nc) with
 * m = floor((P - nc)/2), value (1/nc) exp(i 2 PI (phi1 t + phi2 t^2)),
 * t = linspace(0, txpl, nc). */
void orc_replica_chirp(double txpsf, double txprr, double txpl, double fs_mhz, int num_range,
                       double* chirp) {
    double phi1 = txpsf, phi2 = txprr * 0.5;
    int nc = orc_replica_chirp_length(fs_mhz, txpl);
    int p = num_range + 2 * nc;
    int mi = (int)floor((double)((p - nc) / 2));
    double* t = (double*)malloc(sizeof(double) * (size_t)(nc > 0 ? nc : 1));
    orc_linspace(0.0, txpl, nc, t);
    memset(chirp, 0, sizeof(double) * 2 * (size_t)p);
    double amp = (double)(1.0 / nc);
    for (int i = mi; i < mi + nc; i++) {
        double tt = t[i - mi];
        /* exp(I * 2.0 * PI * x): the imaginary argument is ((2 PI) x) */
        double ph = (2.0 * K_PI) * ((phi1 * tt) + phi2 * (tt * tt));
        chirp[2 * i] = amp * cos(ph);
        chirp[2 * i + 1] = amp * sin(ph);
    }
    free(t);
}

/* image_formation.cpp:4-25: conj(FFT(chirp)) / energy, energy = sum(|c|^4)/P */
void orc_reference_function(const double* chirp, int p, double* ref) {
    double energy = 0.0;
    for (int k = 0; k < p; k++) {
        double a = sqrt(chirp[2 * k] * chirp[2 * k] + chirp[2 * k + 1] * chirp[2 * k + 1]);
        double mag2 = a * a; /* pow(abs(n), 2) */
        energy += mag2 * (mag2 / (double)p);
    }
    memcpy(ref, chirp, sizeof(double) * 2 * (size_t)p);
    orc_fft(ref, p, 0);
    for (int k = 0; k < p; k++) {
        ref[2 * k] = ref[2 * k] / energy;
        ref[2 * k + 1] = -ref[2 * k + 1] / energy;
    }
}

/* s1_decoder.cpp:476-497 with pad/crop :420-445: pad N -> P = N + 2 nc with
 * the data at offset nc, FFT_P, multiply by the reference function, IFFT_P
 * (x 1/P), fftshift (rotate right by P/2), crop nc from both ends. */
void orc_range_compress(double* data, int n_lines, int n, int64_t ld, double txpsf, double txprr,
                        double txpl, double fs_mhz) {
    int nc = orc_replica_chirp_length(fs_mhz, txpl);
    int p = n + 2 * nc;
    double* chirp = (double*)malloc(sizeof(double) * 2 * (size_t)p);
    double* ref = (double*)malloc(sizeof(double) * 2 * (size_t)p);
    orc_replica_chirp(txpsf, txprr, txpl, fs_mhz, n, chirp);
    orc_reference_function(chirp, p, ref);
    get_plan(p);
    double norm = 1.0 / (double)p;
#pragma omp parallel
    {
        double* buf = (double*)malloc(sizeof(double) * 2 * (size_t)p);
#pragma omp for schedule(dynamic, 4)
        for (int l = 0; l < n_lines; l++) {
            double* row = data + 2 * ld * (int64_t)l;
            memset(buf, 0, sizeof(double) * 2 * (size_t)p);
            memcpy(buf + 2 * nc, row, sizeof(double) * 2 * (size_t)n);
            orc_fft(buf, p, 0);
            for (int k = 0; k < p; k++) {
                double xr = buf[2 * k], xi = buf[2 * k + 1];
                double hr = ref[2 * k], hi = ref[2 * k + 1];
                buf[2 * k] = hr * xr - hi * xi;
                buf[2 * k + 1] = hr * xi + hi * xr;
            }
            orc_fft(buf, p, 1);
            /* out[j] = shifted[j + nc] = buf[(j + nc - p/2) mod p] * (1/p) */
            int half = p / 2;
            for (int j = 0; j < n; j++) {
                int src = j + nc - half;
                src = ((src % p) + p) % p;
                row[2 * j] = buf[2 * src] * norm;
                row[2 * j + 1] = buf[2 * src + 1] * norm;
            }
        }
        free(buf);
    }
    free(chirp);
    free(ref);
}

/* ------------------------------------------------------------------------ */
/* Azimuth chain                                                             */
/* ------------------------------------------------------------------------ */
/* s1_decoder.cpp:499-505: per range column FFT_L, then fftshift (rotate
 * right by floor(L/2)). */
void orc_azimuth_fft(double* data, int L, int n, int64_t ld) {
    get_plan(L);
#pragma omp parallel
    {
        double* col = (double*)malloc(sizeof(double) * 2 * (size_t)L);
#pragma omp for schedule(static)
        for (int j = 0; j < n; j++) {
            for (int i = 0; i < L; i++) {
                col[2 * i] = data[2 * (ld * (int64_t)i + j)];
                col[2 * i + 1] = data[2 * (ld * (int64_t)i + j) + 1];
            }
            orc_fft(col, L, 0);
            int sh = L / 2;
            for (int i = 0; i < L; i++) {
                int src = (i + L - sh) % L;
                data[2 * (ld * (int64_t)i + j)] = col[2 * src];
                data[2 * (ld * (int64_t)i + j) + 1] = col[2 * src + 1];
            }
        }
        free(col);
    }
}

/* s1_decoder.cpp:682: compute_axis_dft_in_place(data, 0, 0, true): IFFT_L per
 * column times 1/L, no ifftshift. */
void orc_azimuth_ifft(double* data, int L, int n, int64_t ld) {
    get_plan(L);
    double norm = 1.0 / (L);
#pragma omp parallel
    {
        double* col = (double*)malloc(sizeof(double) * 2 * (size_t)L);
#pragma omp for schedule(static)
        for (int j = 0; j < n; j++) {
            for (int i = 0; i < L; i++) {
                col[2 * i] = data[2 * (ld * (int64_t)i + j)];
                col[2 * i + 1] = data[2 * (ld * (int64_t)i + j) + 1];
            }
            orc_fft(col, L, 1);
            for (int i = 0; i < L; i++) {
                data[2 * (ld * (int64_t)i + j)] = col[2 * i] * norm;
                data[2 * (ld * (int64_t)i + j) + 1] = col[2 * i + 1] * norm;
            }
        }
        free(col);
    }
}

/* signal_processing.cpp:228-257 */
static double sinc_ref(double x) { return x == 0.0 ? 1.0 : sin(K_PI * x) / (K_PI * x); }
static double hann_ref(double x, int L) { return 0.5 * (1 + cos(K_PI * x / L)); }

void orc_sinc_interpolate(const double* row, int n, double t, int L, double* out2) {
    int c = (int)floor(t);
    double rr = 0.0, ri = 0.0;
    for (int k = -L; k <= L; ++k) {
        int idx = c + k;
        if (idx < 0 || idx >= n) continue;
        double x = t - idx;
        double w = hann_ref(x, L);
        double s = sinc_ref(x);
        /* signal[idx] * sinc(x) * w, complex times real, left to right */
        rr += (row[2 * idx] * s) * w;
        ri += (row[2 * idx + 1] * s) * w;
    }
    out2[0] = rr;
    out2[1] = ri;
}

/* Range-Doppler azimuth compression, one Doppler row i at a time, following
 * the formulas of s1_decoder.cpp:599-676 and image_formation.cpp:422-476:
 *   f_i  = linspace(-PRF/2, PRF/2, L)[i]                          (:604)
 *   D    = sqrt(1 - (lambda^2 (f_i + fdc_j)^2) / (4 v_j v_j))     (i_f.cpp:439-441)
 *   RCMC : x(j) <- sinc_interpolate(row, j + (R_j/D - R_j)/dR, 8) (i_f.cpp:469, sp.cpp:238-257)
 *   x   /= sqrt((R_ref / R_j)^3)                                  (:643-646)
 *   x   *= exp(i 4 PI R_j D / lambda) exp(i 2 PI (f_i + fdc_j) tc_j)  (:658-676)
 * The secondary range compression phase is not applied (bounded by 0.04 rad
 * at these geometries). */
void orc_rcmc_mf(double* data, int L, int n, int64_t ld, double prf, double dr, double ref_range,
                 const double* R, const double* v, const double* fdc, const double* tc,
                 int apply_rcmc) {
    double* fa = (double*)malloc(sizeof(double) * (size_t)L);
    orc_linspace(-prf / 2, prf / 2, L, fa);
#if ORC_MUTATION == 7
    for (int i = 0; i < L; i++) fa[i] = -prf / 2 + i * (prf / L);
#endif
    const double lam = K_WAVELENGTH;
#pragma omp parallel
    {
        double* row = (double*)malloc(sizeof(double) * 2 * (size_t)n);
#pragma omp for schedule(dynamic, 4)
        for (int i = 0; i < L; i++) {
            double* out = data + 2 * ld * (int64_t)i;
            memcpy(row, out, sizeof(double) * 2 * (size_t)n);
            for (int j = 0; j < n; j++) {
                double f = fa[i] + fdc[j];
#if ORC_MUTATION == 6
                f = fa[i];
#endif
#if ORC_MUTATION == 2
                double D = sqrt(1 - ((lam * lam) * (f * f)) / (2 * v[j] * v[j]));
#else
                double D = sqrt(1 - ((lam * lam) * (f * f)) / (4 * v[j] * v[j]));
#endif
                double xr, xi;
                if (apply_rcmc) {
                    double shift = (R[j] / D) - R[j];
#if ORC_MUTATION == 5
                    shift = -shift;
#endif
                    double t = (double)j + shift / dr;
                    double o[2];
                    orc_sinc_interpolate(row, n, t, 8, o);
                    xr = o[0];
                    xi = o[1];
                } else {
                    xr = row[2 * j];
                    xi = row[2 * j + 1];
                }
                double q = ref_range / R[j];
                double amp = sqrt(pow(q, 3.0));
#if ORC_MUTATION == 3
                amp = 1.0;
#endif
                xr = xr / amp;
                xi = xi / amp;
                double ph1 = (((4.0 * K_PI) * R[j]) * D) / lam;
                double ph2 = ((2.0 * K_PI) * (fa[i] + fdc[j])) * tc[j];
#if ORC_MUTATION == 1
                ph1 = -ph1;
#endif
#if ORC_MUTATION == 4
                ph2 = -ph2;
#endif
                double c1 = cos(ph1), s1 = sin(ph1), c2 = cos(ph2), s2 = sin(ph2);
                double hr = c1 * c2 - s1 * s2, hi = c1 * s2 + s1 * c2;
                out[2 * j] = xr * hr - xi * hi;
                out[2 * j + 1] = xr * hi + xi * hr;
            }
        }
        free(row);
    }
    free(fa);
}

/* ------------------------------------------------------------------------ */
/* ECCM: signal_processing.cpp:913-1004                                      */
/* ------------------------------------------------------------------------ */
void orc_eccm(double* data, int n_lines, int n, int64_t ld, int fs, int stride, double det,
              double mit, uint8_t* flagged) {
    int num_ffts = (int)floor(((double)n / (double)fs) * ((double)fs / (double)stride));
    get_plan(n);
    get_plan(fs);
#pragma omp parallel
    {
        double* line = (double*)malloc(sizeof(double) * 2 * (size_t)n);
        double* spec = (double*)malloc(sizeof(double) * 2 * (size_t)fs * (size_t)(num_ffts > 0 ? num_ffts : 1));
        double* grads = (double*)malloc(sizeof(double) * (size_t)fs * (size_t)(num_ffts > 0 ? num_ffts : 1));
#pragma omp for schedule(dynamic, 2)
        for (int l = 0; l < n_lines; l++) {
            double* sig = data + 2 * ld * (int64_t)l;
            memcpy(line, sig, sizeof(double) * 2 * (size_t)n);
            orc_fft(line, n, 0);
            double mx = 0.0;
            for (int k = 0; k < n; k++) {
                double a = hypot(line[2 * k], line[2 * k + 1]);
                mx = a > mx ? a : mx;
            }
            if (flagged) flagged[l] = (uint8_t)!(mx < det);
            if (mx < det) continue;
            for (int i = 0; i < num_ffts; i++) {
                int start = i * stride, end = start + fs;
                if (end > n) end = n - 1;
                double* w = spec + 2 * (size_t)fs * i;
                memset(w, 0, sizeof(double) * 2 * (size_t)fs);
                if (end > start) memcpy(w, sig + 2 * start, sizeof(double) * 2 * (size_t)(end - start));
                orc_fft(w, fs, 0);
            }
            for (int i = 0; i < num_ffts; i++) {
                double* w = spec + 2 * (size_t)fs * i;
                for (int j = 0; j < fs; j++) {
                    double gr, gi;
                    if (j != 0 && j != fs - 1) {
                        gr = (w[2 * (j + 1)] - w[2 * (j - 1)]) / 2.0;
                        gi = (w[2 * (j + 1) + 1] - w[2 * (j - 1) + 1]) / 2.0;
                    } else if (j == 0) {
                        gr = w[2] - w[0];
                        gi = w[3] - w[1];
                    } else {
                        gr = w[2 * j] - w[2 * (j - 1)];
                        gi = w[2 * j + 1] - w[2 * (j - 1) + 1];
                    }
                    grads[(size_t)i * fs + j] = hypot(gr, gi);
                }
            }
            for (int i = 0; i < num_ffts; i++) {
                double* w = spec + 2 * (size_t)fs * i;
                for (int j = 0; j < fs; j++) {
                    double k = (fabs(grads[(size_t)i * fs + j]) < mit) ? 1.0 : 0.0;
                    w[2 * j] *= k;
                    w[2 * j + 1] *= k;
                }
                orc_fft(w, fs, 1);
                double nrm = 1.0 / (fs);
                for (int j = 0; j < fs; j++) {
                    w[2 * j] *= nrm;
                    w[2 * j + 1] *= nrm;
                }
            }
            for (int i = 0; i < num_ffts; i++) {
                int start = i * stride, end = start + fs;
                if (end > n) end = n - 1;
                double* w = spec + 2 * (size_t)fs * i;
                for (int j = start; j < end; j++)
                    if (j < n) {
                        sig[2 * j] = w[2 * (j - start)];
                        sig[2 * j + 1] = w[2 * (j - start) + 1];
                    }
            }
        }
        free(line);
        free(spec);
        free(grads);
    }
}

/* ------------------------------------------------------------------------ */
/* Autofocus (no reference counterpart).  rd = range-Doppler data after     */
/* RCMC and matched filter (rows = Doppler bins).  With a per-row phase phi: */
/*   img = IFFT_az(rd * e^{i phi}) / L,  I = |img|^2,  S = sum I,            */
/*   E = -sum (I/S) log(I/S)                                                 */
/* and the adjoint                                                          */
/*   w = -log(I)/S + sum(I log I)/S^2,  g = FFT_az(w * img) / L,             */
/*   dE/dphi_i = 2 sum_j Im(g_ij conj(rd_ij e^{i phi_i})).                   */
/* ------------------------------------------------------------------------ */
double orc_af_entropy_grad(const double* rd, int L, int n, int64_t ld, const double* phase,
                           double* grad) {
    double* img = (double*)malloc(sizeof(double) * 2 * (size_t)L * (size_t)n);
    for (int i = 0; i < L; i++) {
        double c = cos(phase[i]), s = sin(phase[i]);
        for (int j = 0; j < n; j++) {
            double xr = rd[2 * (ld * (int64_t)i + j)], xi = rd[2 * (ld * (int64_t)i + j) + 1];
            img[2 * ((int64_t)i * n + j)] = xr * c - xi * s;
            img[2 * ((int64_t)i * n + j) + 1] = xr * s + xi * c;
        }
    }
    orc_azimuth_ifft(img, L, n, n);
    double S = 0.0, SIl = 0.0;
    for (int64_t k = 0; k < (int64_t)L * n; k++) {
        double I = img[2 * k] * img[2 * k] + img[2 * k + 1] * img[2 * k + 1];
        S += I;
        if (I > 0) SIl += I * log(I);
    }
    double E = -SIl / S + log(S);
    if (grad) {
        for (int64_t k = 0; k < (int64_t)L * n; k++) {
            double I = img[2 * k] * img[2 * k] + img[2 * k + 1] * img[2 * k + 1];
            double w = I > 0 ? (-log(I) / S + SIl / (S * S)) : 0.0;
            img[2 * k] *= w;
            img[2 * k + 1] *= w;
        }
        /* adjoint of (IFFT / L) is FFT / L */
        get_plan(L);
        double* col = (double*)malloc(sizeof(double) * 2 * (size_t)L);
        for (int i = 0; i < L; i++) grad[i] = 0.0;
        for (int j = 0; j < n; j++) {
            for (int i = 0; i < L; i++) {
                col[2 * i] = img[2 * ((int64_t)i * n + j)];
                col[2 * i + 1] = img[2 * ((int64_t)i * n + j) + 1];
            }
            orc_fft(col, L, 0);
            for (int i = 0; i < L; i++) {
                double gr = col[2 * i] / L, gi = col[2 * i + 1] / L;
                double c = cos(phase[i]), s = sin(phase[i]);
                double xr = rd[2 * (ld * (int64_t)i + j)], xi = rd[2 * (ld * (int64_t)i + j) + 1];
                double zr = xr * c - xi * s, zi = xr * s + xi * c;
                /* Im(g * conj(z)) */
                grad[i] += 2.0 * (gi * zr - gr * zi);
            }
        }
        free(col);
    }
    free(img);
    return E;
}

/* The oracle's decode tables and constants, in the layout of the reference
 * shim's ref_tables() so tests can compare them entry by entry. */
void orc_tables(int* thidx, int* mflag, double* simple, double* nrl, double* sf, double* rgdec, double* consts,
                int* huff, int* huff_n, int* huff_start) {
    for (int b = 0; b < 5; ++b) {
        thidx[b] = K_BRC_THIDX[b];
        mflag[b] = K_BRC_MFLAG[b];
        huff_start[b] = K_HUFF_START[b];
        for (int k = 0; k < 10; ++k) simple[b * 10 + k] = K_SIMPLE[b][k];
        for (int k = 0; k < 16; ++k) nrl[b * 16 + k] = K_NRL[b][k];
        for (int k = 0; k < K_HUFF_NCODES[b]; ++k) {
            huff[(b * 16 + k) * 2] = K_HUFF_VAL[b][k];
            huff[(b * 16 + k) * 2 + 1] = K_HUFF_M[b][k];
        }
        huff_n[b] = K_HUFF_NCODES[b];
    }
    for (int t = 0; t < 256; ++t) sf[t] = K_SF[t];
    for (int r = 0; r < 12; ++r) rgdec[r] = K_RGDEC[r];
    consts[0] = K_PI;
    consts[1] = K_C;
    consts[2] = K_F0;
    consts[3] = K_WAVELENGTH;
    consts[4] = K_FREF;
    consts[5] = K_DT_SUPPRESSED;
    consts[6] = K_LATCH;
    consts[7] = 16;
}

/* ------------------------------------------------------------------------ */
/* Burst geometry: image_formation.cpp:391-476                               */
/* ------------------------------------------------------------------------ */
static const double K_SEMI_MAJOR = 6378137.0;    /* structs.h:26 */
static const double K_SEMI_MINOR = 6356752.3142; /* structs.h:27 */

/* get_effective_velocities (image_formation.cpp:391-419); lat = z / x (Q12). */
void orc_effective_velocities(const double* pos, double v, const double* R, int n, double* out) {
    double lat = pos[2] / pos[0];
    double t1 = K_SEMI_MAJOR * K_SEMI_MAJOR * cos(lat), t2 = K_SEMI_MINOR * K_SEMI_MINOR * sin(lat);
    double num = t1 * t1 + t2 * t2;
    double t3 = K_SEMI_MAJOR * cos(lat), t4 = K_SEMI_MINOR * sin(lat);
    double den = t3 * t3 + t4 * t4;
    double re = sqrt(num / den);
    double alt = sqrt(pos[0] * pos[0] + pos[1] * pos[1] + pos[2] * pos[2]);
    for (int j = 0; j < n; j++) {
        double r = R[j];
        double numerator = re * re + alt * alt - r * r;
        double denominator = 2 * re * alt;
        double beta = numerator / denominator;
        double vg = re * v * beta / alt;
        out[j] = sqrt(v * vg);
    }
}

/* apply_secondary_range_compression (image_formation.cpp:422-455) on one
 * line; returns the factors D in `factors`. */
void orc_secondary_range_compression(double* line, int n, const double* vrel, const double* R,
                                     const double* fr, const double* fdc, double az, double* factors) {
    const double lam = K_WAVELENGTH;
    for (int j = 0; j < n; j++) {
        double v = vrel[j];
        double f = az + fdc[j];
        double D = sqrt(1 - ((lam * lam) * (f * f)) / (4 * v * v));
        factors[j] = D;
        double k = 2.0 * (v * v) * pow(K_F0, 3.0) * (D * D);
        k /= (K_C * R[j] * (f * f));
        double ph = -1.0 * K_PI * (fr[j] * fr[j]) / k;
        double c = cos(ph), s = sin(ph);
        double xr = line[2 * j], xi = line[2 * j + 1];
        line[2 * j] = xr * c - xi * s;
        line[2 * j + 1] = xr * s + xi * c;
    }
}

/* apply_range_cell_migration_correction (image_formation.cpp:458-476). */
void orc_rcmc_phase(double* line, int n, const double* R, const double* fr, const double* factors) {
    for (int j = 0; j < n; j++) {
        double shift = (R[j] / factors[j]) - R[j];
        double ph = -4.0 * K_PI * fr[j] * shift / K_C;
        double c = cos(ph), s = sin(ph);
        double xr = line[2 * j], xi = line[2 * j + 1];
        line[2 * j] = xr * c - xi * s;
        line[2 * j + 1] = xr * s + xi * c;
    }
}

/* ------------------------------------------------------------------------ */
/* TOPS: doppler.cpp:4-232, image_formation.cpp:87-388, s1_decoder.cpp:511-701 */
/* ------------------------------------------------------------------------ */
static void cmul(double ar, double ai, double br, double bi, double* r, double* i) {
    *r = ar * br - ai * bi;
    *i = ar * bi + ai * br;
}

/* get_wrapped_estimates over dce_preconditioning (doppler.cpp:4-138): the
 * burst deramped with exp(i PI rate (t - 0.53)^2), t = linspace(0, 1.06, L),
 * then per range column -prf arg(sum_i x_i conj(x_{i+1})) / (2 PI). */
void orc_dc_wrapped(const double* rc, int L, int n, int64_t ld, double rate, double prf, double* wrapped) {
    double* t = (double*)malloc(sizeof(double) * (size_t)L);
    double* dr = (double*)malloc(sizeof(double) * 2 * (size_t)L);
    orc_linspace(0.0, 1.06, L, t);
    for (int i = 0; i < L; i++) {
        double d = t[i] - (0.5 * 1.06);
        double ph = K_PI * rate * (d * d);
        dr[2 * i] = cos(ph);
        dr[2 * i + 1] = sin(ph);
    }
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; j++) {
        double sr = 0.0, si = 0.0;
        for (int i = 0; i < L - 1; i++) {
            double ar, ai, br, bi, pr, pi_;
            cmul(rc[2 * (ld * i + j)], rc[2 * (ld * i + j) + 1], dr[2 * i], dr[2 * i + 1], &ar, &ai);
            cmul(rc[2 * (ld * (i + 1) + j)], rc[2 * (ld * (i + 1) + j) + 1], dr[2 * (i + 1)], dr[2 * (i + 1) + 1],
                 &br, &bi);
            cmul(ar, ai, br, -bi, &pr, &pi_);
            sr += pr;
            si += pi_;
        }
        wrapped[j] = -prf * atan2(si, sr) / (2 * K_PI);
    }
    free(t);
    free(dr);
}

/* The rest of get_doppler_centroid (doppler.cpp:141-218): block means of
 * the wrapped estimates, get_unwrapped_estimates (:43-90, its
 * floor(|exp(i theta)|) included), a quadratic polyfit over the blocks'
 * range times and its value at every range sample. */
static void orc_polyfit2(const double* x, const double* y, int N, double* c);
int orc_dc_unwrap(const double* wrapped, int n, int n_blocks, const double* blk_times, const double* times,
                  double prf, double* dc) {
    if (n_blocks < 2 || n < n_blocks) return -1;
    int bs = n / n_blocks;
    double* fine = (double*)malloc(sizeof(double) * (size_t)n_blocks);
    double* bt = (double*)malloc(sizeof(double) * (size_t)n_blocks);
    for (int b = 0; b < n_blocks; b++) {
        int s = b * bs, e = s + bs < n ? s + bs : n - 1;
        double sum = 0.0;
        for (int k = s; k < e; k++) sum += wrapped[k];
        fine[b] = sum / ((double)e - (double)s);
        bt[b] = blk_times[b] - blk_times[0];
    }
    int m = n_blocks * 8;
    float dt = (float)(bt[1] - bt[0]);
    double* F = (double*)calloc(2 * (size_t)m, sizeof(double));
    for (int b = 0; b < n_blocks; b++) {
        double th = 2.0 * K_PI * fine[b] / prf;
        F[2 * b] = floor(hypot(cos(th), sin(th)));
    }
    orc_fft(F, m, 0);
    int best = 0;
    double bv = -1.0;
    for (int k = 0; k < m; k++) {
        double a = hypot(F[2 * k], F[2 * k + 1]);
        double v = a * a;
        if (v > bv) {
            bv = v;
            best = k;
        }
    }
    /* fftfreq(m, dt) */
    double val = 1.0 / (m * (double)dt);
    int half = (m - 1) / 2 + 1;
    double v = best < half ? best * val : (best - m) * val;
    double a = v / (double)dt;
    double bb = atan2(F[2 * best + 1], F[2 * best]) / (2.0 * K_PI);
    double* un = (double*)malloc(sizeof(double) * (size_t)n_blocks);
    for (int b = 0; b < n_blocks; b++) {
        double th1 = 2.0 * K_PI * fine[b] / prf, th2 = -(a * bt[b] + bb);
        double r = atan2(sin(th1) * cos(th2) + cos(th1) * sin(th2), cos(th1) * cos(th2) - sin(th1) * sin(th2));
        un[b] = (a * bt[b] + bb + r / (2.0 * K_PI)) * prf;
    }
    double c[3];
    orc_polyfit2(bt, un, n_blocks, c);
    double t0 = times[0];
    for (int j = 0; j < n; j++) {
        double t = times[j] - t0;
        dc[j] = (c[0] + c[1] * t + c[2] * t * t);
    }
    free(fine);
    free(bt);
    free(F);
    free(un);
    return 0;
}

/* polyfit (signal_processing.cpp:159-213): normal equations of degree 2,
 * Gaussian elimination with partial pivoting. */
static void orc_polyfit2(const double* x, const double* y, int N, double* c) {
    double Sx[5] = {0}, Sy[3] = {0};
    for (int i = 0; i < N; ++i) {
        double xi = 1.0;
        for (int j = 0; j <= 4; ++j) {
            if (j <= 2) Sy[j] += xi * y[i];
            Sx[j] += xi;
            xi *= x[i];
        }
    }
    double A[3][3] = {{Sx[0], Sx[1], Sx[2]}, {Sx[1], Sx[2], Sx[3]}, {Sx[2], Sx[3], Sx[4]}};
    double b[3] = {Sy[0], Sy[1], Sy[2]};
    for (int i = 0; i < 3; ++i) {
        for (int k = i + 1; k < 3; ++k) {
            if (fabs(A[k][i]) > fabs(A[i][i])) {
                for (int j = 0; j < 3; ++j) {
                    double tmp = A[i][j];
                    A[i][j] = A[k][j];
                    A[k][j] = tmp;
                }
                double tmp = b[i];
                b[i] = b[k];
                b[k] = tmp;
            }
        }
        for (int k = i + 1; k < 3; ++k) {
            double f = A[k][i] / A[i][i];
            for (int j = i; j < 3; ++j) A[k][j] -= f * A[i][j];
            b[k] -= f * b[i];
        }
    }
    for (int i = 2; i >= 0; --i) {
        c[i] = b[i];
        for (int j = i + 1; j < 3; ++j) c[i] -= A[i][j] * c[j];
        c[i] /= A[i][i];
    }
}

/* azimuth_frequency_ufr (image_formation.cpp:122-235): per range column,
 * the fftshifted azimuth spectrum tiled nr = |rate T / prf| times
 * (get_tiled_signal, :87-119), deramped by exp(-i PI f^2 / rate) over
 * f = linspace(-nr prf / 2, nr prf / 2, shape), low-passed to the
 * floor(bw) bins around the centre (moved to the start), re-ramped.
 * out: [shape][n], shape = floor(nr L).  nr and 1 / rate are separate
 * arguments so that the chain can be driven without the ramp. */
int orc_tops_freq_shape(int L, double nr) { return (int)floor(nr * L); }
void orc_az_freq_ufr(const double* in, int L, int n, int64_t ld, double nr, double inv_rate, double prf, double bw,
                     double* out) {
    double B = nr * prf / 2.0;
    int shape = (int)floor(nr * L);
    int mid = shape / 2;
    int start = (int)(mid - (floor(bw) / 2)), end = (int)(mid + (floor(bw) / 2));
    double* f = (double*)malloc(sizeof(double) * (size_t)shape);
    orc_linspace(-B, B, shape, f);
    int whole = (int)floor(nr);
    int rest = floor(nr) != nr ? (int)floor((nr - floor(nr)) * L) : 0;
    int filled = whole * L + rest < shape ? whole * L + rest : shape;
    double sc = 1.0 / (double)shape;
#pragma omp parallel
    {
        double* col = (double*)malloc(sizeof(double) * 2 * (size_t)L);
        double* sh = (double*)malloc(sizeof(double) * 2 * (size_t)L);
        double* u = (double*)malloc(sizeof(double) * 2 * (size_t)shape);
#pragma omp for schedule(dynamic, 8)
        for (int j = 0; j < n; j++) {
            for (int i = 0; i < L; i++) {
                col[2 * i] = in[2 * (ld * i + j)];
                col[2 * i + 1] = in[2 * (ld * i + j) + 1];
            }
            orc_fft(col, L, 0);
            /* fftshift = std::rotate(begin, end - L / 2, end) */
            for (int i = 0; i < L; i++) {
                int src = (i + L - L / 2) % L;
                sh[2 * i] = col[2 * src];
                sh[2 * i + 1] = col[2 * src + 1];
            }
            for (int a = 0; a < shape; a++) {
                double xr = 0.0, xi = 0.0;
                if (a < filled) {
                    xr = sh[2 * (a % L)];
                    xi = sh[2 * (a % L) + 1];
                }
                double ph = (-K_PI * inv_rate) * (f[a] * f[a]);
                double c = cos(ph), s = sin(ph);
                u[2 * a] = c * xr - s * xi;
                u[2 * a + 1] = c * xi + s * xr;
            }
            orc_fft(u, shape, 0);
            for (int a = 0; a < shape; a++) {
                if (a < end - start) {
                    u[2 * a] = u[2 * (start + a)];
                    u[2 * a + 1] = u[2 * (start + a) + 1];
                } else {
                    u[2 * a] = 0.0;
                    u[2 * a + 1] = 0.0;
                }
            }
            orc_fft(u, shape, 1);
            for (int a = 0; a < shape; a++) {
                double ur = u[2 * a] * sc, ui = u[2 * a + 1] * sc;
                double ph = (K_PI * inv_rate) * (f[a] * f[a]);
                double c = cos(ph), s = sin(ph);
                out[2 * ((int64_t)a * n + j)] = c * ur - s * ui;
                out[2 * ((int64_t)a * n + j) + 1] = c * ui + s * ur;
            }
        }
        free(col);
        free(sh);
        free(u);
    }
    free(f);
}

/* The TOPS branch of the Doppler-row loop (s1_decoder.cpp:616-677): per
 * row, unnormalised range FFT, secondary range compression, unnormalised
 * range IFFT, amplitude weight, RCMC phase ramp, azimuth filter; the
 * azimuth FM rate -2 v^2 D^3 / (lambda R) per sample into fm. */
void orc_tops_rows(double* data, int rows, int n, int64_t ld, const double* vrel, const double* R,
                   const double* fr, const double* fdc, const double* faz, const double* tc, double ref_range,
                   double* fm) {
    const double lam = K_WAVELENGTH;
#pragma omp parallel
    {
        double* fac = (double*)malloc(sizeof(double) * (size_t)n);
#pragma omp for schedule(dynamic, 4)
        for (int i = 0; i < rows; i++) {
            double* x = data + 2 * ld * (int64_t)i;
            orc_fft(x, n, 0);
            orc_secondary_range_compression(x, n, vrel + (int64_t)i * n, R, fr, fdc, faz[i], fac);
            orc_fft(x, n, 1);
            for (int j = 0; j < n; j++) {
                double amp = sqrt(pow(ref_range / R[j], 3.0));
                x[2 * j] /= amp;
                x[2 * j + 1] /= amp;
            }
            orc_rcmc_phase(x, n, R, fr, fac);
            for (int j = 0; j < n; j++) {
                double v = vrel[(int64_t)i * n + j], D = fac[j];
                double ph2 = ((2.0 * K_PI) * (faz[i] + fdc[j])) * tc[j];
                double ph1 = (((4.0 * K_PI) * R[j]) * D) / lam;
                double c1 = cos(ph1), s1 = sin(ph1), c2 = cos(ph2), s2 = sin(ph2);
                double hr = c1 * c2 - s1 * s2, hi = c1 * s2 + s1 * c2;
                double xr = x[2 * j], xi = x[2 * j + 1];
                x[2 * j] = xr * hr - xi * hi;
                x[2 * j + 1] = xr * hi + xi * hr;
                fm[(int64_t)i * n + j] = -2 * pow(v, 2.0) * pow(D, 3.0) / (lam * R[j]);
            }
        }
        free(fac);
    }
}

/* linear_resample (signal_processing.cpp:32-63) of a real sequence, as
 * written: a running abscissa x += dx_out walked over the input segments,
 * the last output the last input. */
static void orc_linear_resample(const double* in, int n_in, int n_out, double* out) {
    if (n_in == n_out) {
        memcpy(out, in, sizeof(double) * (size_t)n_in);
        return;
    }
    double x = 0.0, dx_in = 1.0 / (n_in - 1), dx_out = 1.0 / (n_out - 1);
    int index = 0;
    for (int i = 0; i < n_in - 1; i++) {
        double x0 = i * dx_in, x1 = (i + 1) * dx_in, y0 = in[i], y1 = in[i + 1];
        while (x < x1 && index < n_out) {
            out[index] = y0 + ((x - x0) * (y1 - y0) / (x1 - x0));
            x += dx_out;
            index += 1;
        }
    }
    out[n_out - 1] = in[n_in - 1];
}
void orc_linear_resample_many(const double* in, int n_in, int n_out, int count, double* out) {
    for (int k = 0; k < count; k++) orc_linear_resample(in + (int64_t)k * n_in, n_in, n_out, out + (int64_t)k * n_out);
}

/* quadratic_resample (signal_processing.cpp:109-155): three-point Lagrange
 * through i0 = max(i - 1, 0), i, min(i + 1, n_in - 1) with i = min(int(x /
 * dx_in), n_in - 2); outputs whose points coincide (x < dx_in) are 0 / 0. */
void orc_quadratic_resample(const double* in, int n_in, int n_out, double* out) {
    if (n_in == n_out) {
        memcpy(out, in, sizeof(double) * 2 * (size_t)n_in);
        return;
    }
    double dx_in = 1.0 / (n_in - 1), dx_out = 1.0 / (n_out - 1);
    for (int j = 0; j < n_out; ++j) {
        double x = j * dx_out;
        int i = (int)(x / dx_in);
        if (i > n_in - 2) i = n_in - 2;
        int i0 = i - 1 > 0 ? i - 1 : 0, i1 = i, i2 = i + 1 < n_in - 1 ? i + 1 : n_in - 1;
        double x0 = i0 * dx_in, x1 = i1 * dx_in, x2 = i2 * dx_in;
        double L0 = (x - x1) * (x - x2) / ((x0 - x1) * (x0 - x2));
        double L1 = (x - x0) * (x - x2) / ((x1 - x0) * (x1 - x2));
        double L2 = (x - x0) * (x - x1) / ((x2 - x0) * (x2 - x1));
        out[2 * j] = L0 * in[2 * i0] + L1 * in[2 * i1] + L2 * in[2 * i2];
        out[2 * j + 1] = L0 * in[2 * i0 + 1] + L1 * in[2 * i1 + 1] + L2 * in[2 * i2 + 1];
    }
}

/* The scalar set-up of azimuth_time_ufr (image_formation.cpp:238-292) for a
 * burst of n_az (frequency-UFR) lines and n range samples.  fm_col0 is the
 * azimuth FM rate of range sample 0 on every line, fm_ref the rate at the
 * reference point of n_ref = f_dc[n / 2] / fm_ref.
 *   d: [bandwidth, focused_time, n_ref, ka_mean]
 *   k: [num_tiles, shape, lpf_start, lpf_end, downsample_shape,
 *       output_shape, crop_start, crop_end]
 * Returns 0, or -1 when the low-pass window or the crop leaves the arrays. */
int orc_tops_time_params(const double* fm_col0, int n_az, int n, double dc_mid, double fm_ref, double nr, double prf,
                         double T, double bw, int swath, double* d, int* k) {
    double bandwidth = (nr * prf) - bw;
    double sum = 0.0;
    for (int a = 0; a < n_az; a++) sum += fm_col0[a];
    double ka_mean = fabs(sum) / n_az;
    double focused_time = T + ((bandwidth - (2 * bw)) / ka_mean);
    double n_ref = dc_mid / fm_ref;
    if (swath != 11)
        focused_time *= 0.9;
    else
        focused_time *= 1.05;
    int pos = (int)ceil((n_ref + focused_time / 2) / T);
    int neg = (int)floor((n_ref - focused_time / 2) / T);
    int tiles = pos - neg;
    int ext = swath == 11 ? (int)floor(bw * 1.15) : (int)floor(bw * 0.26);
    int out_shape = (int)floor(prf * T);
    int ds = out_shape + ext;
    int shape = tiles * n_az;
    int mid = shape / 2;
    int start = (int)(mid - bandwidth / 2), end = (int)(mid + bandwidth / 2);
    int cmid = ds / 2;
    d[0] = bandwidth;
    d[1] = focused_time;
    d[2] = n_ref;
    d[3] = ka_mean;
    k[0] = tiles;
    k[1] = shape;
    k[2] = start;
    k[3] = end;
    k[4] = ds;
    k[5] = out_shape;
    k[6] = cmid - (out_shape / 2);
    k[7] = cmid + (out_shape / 2);
    if (tiles < 1 || start < 0 || end > shape || end < start || ds < 3 || shape < 3 || k[6] < 0 || k[7] > ds)
        return -1;
    return 0;
}

/* azimuth_time_ufr (image_formation.cpp:238-388) after the set-up: per
 * range column r, the n_az lines tiled k[0] times, deramped with the FM rate
 * resampled to the tiled length, low-passed, resampled (quadratic) to the
 * downsample length, re-ramped and cropped to rows [k[6], k[7]).
 * in: [n_az][n] (ld), dc[n], fm: [n_az][n] (row-major, FM rate of line a at
 * range r), out: [k[7] - k[6]][n]. */
void orc_az_time_ufr(const double* in, int n_az, int n, int64_t ld, const double* dc, const double* fm, double rate,
                     const double* d, const int* k, double* out) {
    int tiles = k[0], shape = k[1], start = k[2], end = k[3], ds = k[4], c0 = k[6], c1 = k[7];
    double ft = d[1], n_ref = d[2];
    double* t1 = (double*)malloc(sizeof(double) * (size_t)shape);
    double* t2 = (double*)malloc(sizeof(double) * (size_t)ds);
    orc_linspace(-ft / 2, ft / 2, shape, t1);
    orc_linspace(n_ref - ft / 2, n_ref + ft / 2, ds, t2);
    double sc = 1.0 / (double)shape;
#pragma omp parallel
    {
        double* col = (double*)malloc(sizeof(double) * (size_t)n_az);
        double* ka = (double*)malloc(sizeof(double) * (size_t)(shape > ds ? shape : ds));
        double* u = (double*)malloc(sizeof(double) * 2 * (size_t)shape);
        double* v = (double*)malloc(sizeof(double) * 2 * (size_t)ds);
#pragma omp for schedule(dynamic, 4)
        for (int r = 0; r < n; r++) {
            for (int a = 0; a < n_az; a++) col[a] = fm[(int64_t)a * n + r];
            orc_linear_resample(col, n_az, shape, ka);
            double fdc = dc[r];
            (void)tiles;
            for (int a = 0; a < shape; a++) {
                int src = a % n_az;
                double xr = in[2 * (ld * src + r)], xi = in[2 * (ld * src + r) + 1];
                double kt = -ka[a] * rate / (rate - ka[a]);
                double dct = -fdc / ka[a];
                double t = dct + t1[a];
                double p1 = ((-1.0 * K_PI) * kt) * pow(t, 2.0);
                double p2 = (((2.0 * K_PI) * (kt + ka[a])) * dct) * t;
                double ph = p1 + p2;
                double c = cos(ph), s = sin(ph);
                u[2 * a] = c * xr - s * xi;
                u[2 * a + 1] = c * xi + s * xr;
            }
            orc_fft(u, shape, 0);
            for (int a = 0; a < shape; a++) {
                if (a < end - start) {
                    u[2 * a] = u[2 * (start + a)];
                    u[2 * a + 1] = u[2 * (start + a) + 1];
                } else {
                    u[2 * a] = 0.0;
                    u[2 * a + 1] = 0.0;
                }
            }
            orc_fft(u, shape, 1);
            for (int a = 0; a < shape; a++) {
                u[2 * a] *= sc;
                u[2 * a + 1] *= sc;
            }
            orc_quadratic_resample(u, shape, ds, v);
            orc_linear_resample(col, n_az, ds, ka);
            for (int j = c0; j < c1; j++) {
                double kt = -ka[j] * rate / (rate - ka[j]);
                double dct = -fdc / ka[j];
                double t = dct + t2[j];
                double p1 = ((1.0 * K_PI) * kt) * pow(t, 2.0);
                double p2 = (((2.0 * K_PI) * (kt + ka[j])) * dct) * t;
                double ph = p1 - p2;
                double c = cos(ph), s = sin(ph);
                double xr = v[2 * j], xi = v[2 * j + 1];
                out[2 * ((int64_t)(j - c0) * n + r)] = c * xr - s * xi;
                out[2 * ((int64_t)(j - c0) * n + r) + 1] = c * xi + s * xr;
            }
        }
        free(col);
        free(ka);
        free(u);
        free(v);
    }
    free(t1);
    free(t2);
}
