/*
 * s1synth.c -- synthetic Sentinel-1 IW/SM raw scenes and their Level-0
 * FDBAQ (BAQ mode 12, data format D) space packets.
 *
 * The scene simulator places point targets analytically in the raw
 * (range-time x azimuth-time) domain: each target contributes an LFM pulse
 * delayed by 2 R(eta) / c with the two-way phase -4 pi R(eta) / lambda,
 * R(eta) = sqrt(R0^2 + V^2 eta^2), over the lines whose instantaneous Doppler
 * lies inside the processed band around the Doppler centroid.  Clutter and
 * receiver noise are white circular complex Gaussian samples in the raw
 * domain.  All randomness comes from the Philox4x32-10 counter-based
 * generator (Salmon et al., SC'11), so every sample is a pure function of
 * (seed, line, sample) and the result does not depend on the thread count.
 *
 * The encoder writes the packet layout the decoder reads (channel order IE,
 * IO, QE, QO; a 3-bit BRC ahead of each IE block, an 8-bit THIDX ahead of
 * each QE block; sign bit then Huffman code per sample; each channel padded
 * to a 16-bit word) with the FDBAQ reconstruction laws of the S1 SPPDU
 * document (the same tables the decoder uses).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- tables */
static const int T_THIDX_LIMIT[5] = {3, 3, 5, 6, 8};
static const int T_MFLAG[5] = {3, 4, 6, 9, 15};
/* Huffman code words (MSB first) and lengths, indexed [brc][m] */
static const uint16_t T_CODE[5][16] = {
    {0x0, 0x2, 0x6, 0x7},
    {0x0, 0x2, 0x6, 0xE, 0xF},
    {0x0, 0x2, 0x6, 0xE, 0x1E, 0x3E, 0x3F},
    {0x0, 0x1, 0x2, 0x6, 0xE, 0x1E, 0x3E, 0x7E, 0xFE, 0xFF},
    {0x0, 0x2, 0x3, 0x4, 0x5, 0xC, 0xD, 0xE, 0x1E, 0x3E, 0xFC, 0xFD, 0x1FC, 0x1FD, 0x1FE, 0x1FF}};
static const uint8_t T_LEN[5][16] = {
    {1, 2, 3, 3},
    {1, 2, 3, 4, 4},
    {1, 2, 3, 4, 5, 6, 6},
    {2, 2, 2, 3, 4, 5, 6, 7, 8, 8},
    {2, 3, 3, 3, 3, 4, 4, 4, 5, 6, 8, 8, 9, 9, 9, 9}};
static const double T_SIMPLE[5][10] = {
    {3.00, 3.00, 3.16, 3.53},
    {4.00, 4.00, 4.08, 4.37},
    {6.00, 6.00, 6.00, 6.15, 6.50, 6.88},
    {9.00, 9.00, 9.00, 9.00, 9.36, 9.50, 10.10},
    {15.00, 15.00, 15.00, 15.00, 15.00, 15.00, 15.22, 15.50, 16.06}};
static const double T_NRL[5][16] = {
    {0.3637, 1.0915, 1.8208, 2.6406},
    {0.3042, 0.9127, 1.5216, 2.1313, 2.8426},
    {0.2305, 0.6916, 1.1528, 1.6140, 2.0754, 2.5369, 3.1191},
    {0.1702, 0.5107, 0.8511, 1.1916, 1.5321, 1.8726, 2.2131, 2.5536, 2.8942, 3.3744},
    {0.1130, 0.3389, 0.5649, 0.7908, 1.0167, 1.2428, 1.4687, 1.6947, 1.9206, 2.1466, 2.3725,
     2.5985, 2.8244, 3.0504, 3.2764, 3.6623}};
static double T_SF[256];
static int T_SF_READY = 0;

static void sf_init(void) {
    /* Sigma factor per THIDX, Table 5.2-3 of the SPPDU document. */
    static const double vals[256] = {
        0.00,   0.63,   1.25,   1.88,   2.51,   3.13,   3.76,   4.39,   5.01,   5.64,   6.27,   6.89,
        7.52,   8.15,   8.77,   9.40,   10.03,  10.65,  11.28,  11.91,  12.53,  13.16,  13.79,  14.41,
        15.04,  15.67,  16.29,  16.92,  17.55,  18.17,  18.80,  19.43,  20.06,  20.68,  21.31,  21.93,
        22.56,  23.19,  23.81,  24.44,  25.06,  25.69,  26.32,  26.94,  27.57,  28.20,  28.83,  29.45,
        30.08,  30.71,  31.33,  31.96,  32.59,  33.21,  33.84,  34.47,  35.09,  35.72,  36.35,  36.97,
        37.60,  38.23,  38.85,  39.48,  40.11,  40.73,  41.36,  41.99,  42.61,  43.24,  43.87,  44.49,
        45.12,  45.75,  46.37,  47.00,  47.63,  48.25,  48.88,  49.51,  50.13,  50.76,  51.39,  52.01,
        52.64,  53.27,  53.89,  54.52,  55.15,  55.77,  56.40,  57.03,  57.65,  58.28,  58.91,  59.54,
        60.16,  60.79,  61.41,  62.04,  62.98,  64.24,  65.49,  66.74,  68.00,  69.25,  70.50,  71.76,
        73.01,  74.26,  75.52,  76.77,  78.02,  79.28,  80.53,  81.78,  83.04,  84.29,  85.54,  86.80,
        88.05,  89.30,  90.56,  91.81,  93.07,  94.32,  95.57,  96.83,  98.08,  99.33,  100.58, 101.84,
        103.09, 104.34, 105.60, 106.85, 108.10, 109.36, 110.61, 111.86, 113.11, 114.37, 115.62, 116.87,
        118.13, 119.38, 120.63, 121.89, 123.14, 124.39, 125.65, 126.90, 128.15, 129.41, 130.66, 131.91,
        133.17, 134.42, 135.67, 136.93, 138.18, 139.43, 140.69, 141.94, 143.19, 144.45, 145.70, 146.95,
        148.21, 149.46, 150.71, 151.97, 153.22, 154.47, 155.73, 156.98, 158.23, 159.49, 160.74, 161.99,
        163.25, 164.50, 165.75, 167.01, 168.26, 169.51, 170.77, 172.02, 173.27, 174.53, 175.78, 177.03,
        178.29, 179.54, 180.79, 182.05, 183.30, 184.55, 185.81, 187.06, 188.31, 189.57, 190.82, 192.07,
        193.33, 194.58, 195.83, 197.09, 198.34, 199.59, 200.85, 202.10, 203.35, 204.61, 205.86, 207.11,
        208.37, 209.62, 210.87, 212.13, 213.38, 214.63, 215.89, 217.14, 218.39, 219.65, 220.90, 222.15,
        223.41, 224.66, 225.91, 227.17, 228.42, 229.67, 230.93, 232.18, 233.43, 234.69, 235.94, 237.19,
        238.45, 239.70, 240.95, 242.21, 243.46, 244.71, 245.97, 247.22, 248.47, 249.73, 250.98, 252.23,
        253.49, 254.74, 255.99, 255.99};
    if (!T_SF_READY) {
        memcpy(T_SF, vals, sizeof(vals));
        T_SF_READY = 1;
    }
}

/* ------------------------------------------------------------ Philox4x32 */
static inline void philox_round(uint32_t* c, const uint32_t* k) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t h0 = (uint32_t)(p0 >> 32), l0 = (uint32_t)p0;
    uint32_t h1 = (uint32_t)(p1 >> 32), l1 = (uint32_t)p1;
    uint32_t n0 = h1 ^ c[1] ^ k[0];
    uint32_t n2 = h0 ^ c[3] ^ k[1];
    c[0] = n0;
    c[1] = l1;
    c[2] = n2;
    c[3] = l0;
}

static void philox4x32_10(const uint32_t ctr[4], uint64_t seed, uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t k[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int r = 0; r < 10; r++) {
        philox_round(c, k);
        k[0] += 0x9E3779B9u;
        k[1] += 0xBB67AE85u;
    }
    memcpy(out, c, sizeof(c));
}

static inline double u01(uint32_t a, uint32_t b) {
    uint64_t x = ((uint64_t)a << 21) ^ (uint64_t)b; /* 53 random bits */
    x &= ((uint64_t)1 << 53) - 1;
    return ((double)x + 0.5) * (1.0 / 9007199254740992.0);
}

/* Two independent N(0,1) draws for stream `stream`, coordinates (a, b). */
static void gauss2(uint64_t seed, uint32_t stream, uint32_t a, uint32_t b, double* g0, double* g1) {
    uint32_t ctr[4] = {a, b, stream, 0x5EED5EEDu};
    uint32_t o[4];
    philox4x32_10(ctr, seed, o);
    double u1 = u01(o[0], o[1]), u2 = u01(o[2], o[3]);
    double r = sqrt(-2.0 * log(u1));
    *g0 = r * cos(2.0 * M_PI * u2);
    *g1 = r * sin(2.0 * M_PI * u2);
}

/* Uniform [0,1) draws for stream `stream`, coordinates (a, b). */
double syn_uniform(uint64_t seed, uint32_t stream, uint32_t a, uint32_t b) {
    uint32_t ctr[4] = {a, b, stream, 0x0001F00Du};
    uint32_t o[4];
    philox4x32_10(ctr, seed, o);
    return u01(o[0], o[1]);
}

void syn_gauss(uint64_t seed, uint32_t stream, uint32_t a, uint32_t b, double* out2) {
    gauss2(seed, stream, a, b, &out2[0], &out2[1]);
}

/* ------------------------------------------------------------ raw scene */
typedef struct {
    double fs_hz;      /* range sample rate                         */
    double k_hz_s;     /* chirp rate                                */
    double tau_s;      /* pulse length                              */
    double f_start_hz; /* chirp start frequency                     */
    double prf_hz;
    double wavelength;
    double v_ms;       /* effective velocity                        */
    double r_near;     /* slant range of sample 0                   */
    double f_dc;       /* Doppler centroid                          */
    double band_frac;  /* processed Doppler band / PRF              */
    double clutter_sigma;
    double noise_sigma;
    double gain;       /* raw amplitude scale applied at the end    */
} syn_scene;

/* targets: ntg rows of (range sample, azimuth line, amplitude, phase) */
void syn_raw_scene(double* out, int L, int N, const syn_scene* sc, const double* tg, int ntg,
                   uint64_t seed) {
    const double c = 299792458.0;
    const double dr = c / (2.0 * sc->fs_hz);
    const double half = 0.5 * sc->tau_s * sc->fs_hz;
    const double bw = sc->band_frac * sc->prf_hz;
#pragma omp parallel for schedule(dynamic, 4)
    for (int l = 0; l < L; l++) {
        double* row = out + 2 * (int64_t)N * l;
        for (int j = 0; j < N; j++) {
            double g0 = 0, g1 = 0, h0 = 0, h1 = 0;
            if (sc->clutter_sigma > 0) gauss2(seed, 1u, (uint32_t)l, (uint32_t)j, &g0, &g1);
            if (sc->noise_sigma > 0) gauss2(seed, 2u, (uint32_t)l, (uint32_t)j, &h0, &h1);
            /* CN(0, s^2): each quadrature N(0, s^2 / 2) */
            row[2 * j] = (sc->clutter_sigma * g0 + sc->noise_sigma * h0) * M_SQRT1_2;
            row[2 * j + 1] = (sc->clutter_sigma * g1 + sc->noise_sigma * h1) * M_SQRT1_2;
        }
        for (int t = 0; t < ntg; t++) {
            double n_t = tg[4 * t], a_t = tg[4 * t + 1], amp = tg[4 * t + 2], psi = tg[4 * t + 3];
            double r0 = sc->r_near + n_t * dr;
            double eta = ((double)l - a_t) / sc->prf_hz;
            double R = sqrt(r0 * r0 + sc->v_ms * sc->v_ms * eta * eta);
            double fd = -2.0 * sc->v_ms * sc->v_ms * eta / (sc->wavelength * R);
            if (fabs(fd - sc->f_dc) > 0.5 * bw) continue;
            double n0 = 2.0 * (R - sc->r_near) / c * sc->fs_hz;
            int first = (int)ceil(n0 - half), last = (int)floor(n0 + half);
            if (first < 0) first = 0;
            if (last > N - 1) last = N - 1;
            if (first > last) continue;
            double ph0 = -4.0 * M_PI * R / sc->wavelength + psi;
            /* pulse phase 2 pi (f0 t + K/2 t^2), t = (n - n0)/fs + tau/2 */
            for (int n = first; n <= last; n += 256) {
                int stop = n + 256 <= last + 1 ? n + 256 : last + 1;
                double t = ((double)n - n0) / sc->fs_hz + 0.5 * sc->tau_s;
                double ph = ph0 + 2.0 * M_PI * (sc->f_start_hz * t + 0.5 * sc->k_hz_s * t * t);
                double zr = amp * cos(ph), zi = amp * sin(ph);
                double d = 1.0 / sc->fs_hz;
                double b = 2.0 * M_PI * (sc->f_start_hz * d + sc->k_hz_s * t * d);
                double g = 2.0 * M_PI * 0.5 * sc->k_hz_s * d * d;
                double rr = cos(b + g), ri = sin(b + g);
                double qr = cos(2.0 * g), qi = sin(2.0 * g);
                for (int m = n; m < stop; m++) {
                    row[2 * m] += zr;
                    row[2 * m + 1] += zi;
                    double nr = zr * rr - zi * ri;
                    zi = zr * ri + zi * rr;
                    zr = nr;
                    double sr = rr * qr - ri * qi;
                    ri = rr * qi + ri * qr;
                    rr = sr;
                }
            }
        }
        for (int j = 0; j < 2 * N; j++) row[j] *= sc->gain;
    }
}

/* ------------------------------------------------------------ bit writer */
typedef struct {
    uint8_t* buf;
    int64_t cap;
    int64_t bit;
    int overflow;
} bitw;

static inline void put_bits(bitw* w, uint32_t v, int n) {
    for (int i = n - 1; i >= 0; i--) {
        int64_t byte = w->bit >> 3;
        if (byte >= w->cap) {
            w->overflow = 1;
            return;
        }
        if ((v >> i) & 1u) w->buf[byte] |= (uint8_t)(0x80u >> (w->bit & 7));
        w->bit++;
    }
}

/* Header fields MSB first, widths of the SPPDU primary (8 fields) and
 * secondary (46 fields) headers. */
static const int W_PRIM[8] = {3, 1, 1, 7, 4, 2, 14, 16};
static const int W_SEC[46] = {32, 16, 32, 32, 8, 1, 3, 4, 32, 8, 16, 32, 32, 1, 2, 5,
                              8,  8,  8,  8,  1, 15, 1, 15, 24, 3, 5, 24, 24, 24, 1, 3,
                              2,  2,  4,  2,  10, 2, 1,  5, 4,  3, 1, 8,  16, 8};

void syn_pack_headers(uint8_t* out68, const int64_t* prim, const int64_t* sec) {
    memset(out68, 0, 68);
    bitw w = {out68, 6, 0, 0};
    for (int i = 0; i < 8; i++) put_bits(&w, (uint32_t)prim[i], W_PRIM[i]);
    bitw s = {out68 + 6, 62, 0, 0};
    for (int i = 0; i < 46; i++) put_bits(&s, (uint32_t)sec[i], W_SEC[i]);
}

/* ------------------------------------------------------------ quantiser */
static inline int quantise(double v, int brc, int thidx, int* sign) {
    double a = fabs(v);
    *sign = v < 0.0;
    int flag = T_MFLAG[brc];
    int best = 0;
    double be = 1e300;
    if (thidx <= T_THIDX_LIMIT[brc]) {
        for (int m = 0; m <= flag; m++) {
            double lv = m < flag ? (double)m : T_SIMPLE[brc][thidx];
            double e = fabs(a - lv);
            if (e < be) { be = e; best = m; }
        }
    } else {
        for (int m = 0; m <= flag; m++) {
            double e = fabs(a - T_NRL[brc][m] * T_SF[thidx]);
            if (e < be) { be = e; best = m; }
        }
    }
    return best;
}

/* Per-block BRC and THIDX.  THIDX is the sigma-factor index closest to the
 * block RMS; BRC grows with the block RMS relative to `sigma_ref`, with a
 * seeded +-1 jitter; a seeded fraction `simple_frac` of blocks is forced to
 * the simple-reconstruction THIDX range.  `scale_out` receives the factor the
 * block's samples are multiplied by before quantisation (1, or the factor
 * that brings a forced-simple block into the integer range). */
void syn_block_codes(const double* iq, int nq, uint64_t seed, int line, double sigma_ref,
                     double simple_frac, uint8_t* brc, uint8_t* thidx, double* scale_out) {
    sf_init();
    int nb = (nq + 127) / 128;
    for (int b = 0; b < nb; b++) {
        int q0 = 128 * b, q1 = q0 + 128 < nq ? q0 + 128 : nq;
        double s2 = 0.0;
        for (int q = q0; q < q1; q++) {
            const double* p = iq + 4 * (int64_t)q;
            s2 += p[0] * p[0] + p[1] * p[1] + p[2] * p[2] + p[3] * p[3];
        }
        double sig = sqrt(s2 / (4.0 * (q1 - q0)));
        double u = syn_uniform(seed, 10u, (uint32_t)line, (uint32_t)b);
        double u2 = syn_uniform(seed, 11u, (uint32_t)line, (uint32_t)b);
        int br = (int)floor(2.0 + 1.5 * log2((sig + 1e-12) / sigma_ref));
        br += (u < 0.2) ? -1 : (u > 0.8 ? 1 : 0);
        if (br < 0) br = 0;
        if (br > 4) br = 4;
        int th = 0;
        double be = 1e300;
        for (int t = 0; t < 256; t++) {
            double e = fabs(T_SF[t] - sig);
            if (e < be) { be = e; th = t; }
        }
        double sc = 1.0;
        if (u2 < simple_frac) {
            th = (int)floor(syn_uniform(seed, 12u, (uint32_t)line, (uint32_t)b) * (T_THIDX_LIMIT[br] + 1));
            if (th > T_THIDX_LIMIT[br]) th = T_THIDX_LIMIT[br];
            sc = sig > 0 ? (0.4 * T_MFLAG[br]) / sig : 1.0;
        }
        brc[b] = (uint8_t)br;
        thidx[b] = (uint8_t)th;
        scale_out[b] = sc;
    }
}

/* Encode one packet's user data.  iq holds 2*nq complex samples (4*nq
 * doubles); channel values are IE_k = Re x[2k], IO_k = Re x[2k+1],
 * QE_k = Im x[2k], QO_k = Im x[2k+1].  Returns the byte length (padded to a
 * multiple of 4), or -1 when `cap` is too small. */
int64_t syn_encode_packet(const double* iq, int nq, const uint8_t* brc, const uint8_t* thidx,
                          const double* scale, uint8_t* out, int64_t cap) {
    sf_init();
    memset(out, 0, (size_t)cap);
    bitw w = {out, cap, 0, 0};
    int nb = (nq + 127) / 128;
    /* component index inside a quad: IE=0 (re even), IO=2 (re odd), QE=1 (im even), QO=3 */
    static const int comp[4] = {0, 2, 1, 3};
    for (int ch = 0; ch < 4; ch++) {
        for (int b = 0; b < nb; b++) {
            if (ch == 0) put_bits(&w, brc[b], 3);
            if (ch == 2) put_bits(&w, thidx[b], 8);
            int q0 = 128 * b, q1 = q0 + 128 < nq ? q0 + 128 : nq;
            for (int q = q0; q < q1; q++) {
                double v = iq[4 * (int64_t)q + comp[ch]] * scale[b];
                int sg;
                int m = quantise(v, brc[b], thidx[b], &sg);
                put_bits(&w, (uint32_t)sg, 1);
                put_bits(&w, T_CODE[brc[b]][m], T_LEN[brc[b]][m]);
            }
        }
        if (w.bit % 16) w.bit += 16 - (w.bit % 16);
    }
    if (w.overflow) return -1;
    int64_t bytes = (w.bit + 7) / 8;
    bytes = (bytes + 3) & ~(int64_t)3;
    if (bytes > cap) return -1;
    return bytes;
}

/* Encode a whole burst: raw [L][N] complex (N = 2 nq) -> concatenated space
 * packets (6-byte primary + 62-byte secondary header + user data).
 * sec_template holds the 46 secondary fields; per line the PRI count, space
 * packet count, sequence count and fine time advance.  Returns the stream
 * length or -1 on overflow.  ud_offsets/ud_lengths receive the position and
 * length of each packet's user data inside the stream. */
int64_t syn_encode_burst(const double* raw, int L, int N, const int64_t* prim_template,
                         const int64_t* sec_template, uint64_t seed, double sigma_ref,
                         double simple_frac, uint8_t* stream, int64_t cap, uint64_t* ud_offsets,
                         uint32_t* ud_lengths) {
    sf_init();
    int nq = N / 2;
    int nb = (nq + 127) / 128;
    int64_t per = (int64_t)nq * 4 * 10 / 8 + 4 * nb + 64; /* worst case user data bytes */
    uint8_t* tmp = (uint8_t*)malloc((size_t)per * (size_t)L);
    int64_t* lens = (int64_t*)malloc(sizeof(int64_t) * (size_t)L);
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(| : bad)
    for (int l = 0; l < L; l++) {
        uint8_t* br = (uint8_t*)malloc((size_t)nb);
        uint8_t* th = (uint8_t*)malloc((size_t)nb);
        double* scl = (double*)malloc(sizeof(double) * (size_t)nb);
        const double* iq = raw + 2 * (int64_t)N * l;
        syn_block_codes(iq, nq, seed, l, sigma_ref, simple_frac, br, th, scl);
        lens[l] = syn_encode_packet(iq, nq, br, th, scl, tmp + per * (int64_t)l, per);
        if (lens[l] < 0) bad = 1;
        free(br);
        free(th);
        free(scl);
    }
    int64_t pos = 0;
    if (!bad) {
        for (int l = 0; l < L; l++) {
            if (pos + 68 + lens[l] > cap) {
                bad = 1;
                break;
            }
            int64_t prim[8], sec[46];
            memcpy(prim, prim_template, sizeof(prim));
            memcpy(sec, sec_template, sizeof(sec));
            prim[6] = (prim_template[6] + l) & 0x3FFF;
            prim[7] = 62 + lens[l] - 1; /* packet data length */
            sec[1] = (sec_template[1] + 7 * l) & 0xFFFF;
            sec[11] = sec_template[11] + l;
            sec[12] = sec_template[12] + l;
            sec[44] = nq;
            syn_pack_headers(stream + pos, prim, sec);
            memcpy(stream + pos + 68, tmp + per * (int64_t)l, (size_t)lens[l]);
            ud_offsets[l] = (uint64_t)(pos + 68);
            ud_lengths[l] = (uint32_t)lens[l];
            pos += 68 + lens[l];
        }
    }
    free(tmp);
    free(lens);
    return bad ? -1 : pos;
}
