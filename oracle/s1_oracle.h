/*
 * s1_oracle.h -- FP64 CPU restatement of the Sentinel-1 Level-0 -> SLC hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker and the timed
 * CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product (sentinel1_decode_amd) never
 * links, loads or calls anything in oracle/.
 *
 * Every function restates the reference algorithm named in its comment
 * (file:line into AndrewPlayer3/sentinel1_decode src/).  Complex arrays are
 * interleaved (re, im) doubles, row-major, one radar line per row, leading
 * dimension `ld` counted in complex elements.
 */
#ifndef S1_ORACLE_H
#define S1_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Per-packet status codes; the same values are returned by the GPU decoder. */
enum {
    ORC_OK = 0,
    ORC_ERR_BRC = 1,      /* BRC field > 4            packet.cpp:795-798          */
    ORC_ERR_OVERRUN = 2,  /* bit read past user data  decoding_utils.cpp:107-119  */
    ORC_ERR_HUFFMAN = 3,  /* code longer than 10 bits decoding_utils.cpp:69-72    */
    ORC_ERR_MCODE = 4,    /* m_code above BRC flag    packet.cpp:690              */
    ORC_ERR_ARG = 5       /* bad argument (num_quads < 0, ...)                    */
};

/* ---- bit reading and Huffman (decoding_utils.cpp:44-134) ---------------- */
int orc_read_n_bits(const uint8_t* data, int64_t nbytes, int64_t start_bit, int n,
                    uint64_t* out);
int orc_huffman_decode(const uint8_t* data, int64_t nbytes, int brc,
                       int64_t* bit_index, int* m_code);

/* ---- FDBAQ (type D) sample reconstruction and packet decode ------------- */
double orc_s_value_d(int brc, int thidx, int sign, int m_code, int* status);
int orc_decode_packet_d(const uint8_t* user_data, int64_t nbytes, int num_quads,
                        double* out, uint8_t* brc_out, uint8_t* thidx_out);
void orc_decode_packets(const uint8_t* base, const uint64_t* offsets,
                        const uint32_t* lengths, const int32_t* num_quads, int n_packets,
                        double* out, int64_t ld, int32_t* status);

/* ---- headers and physical parameters (packet.cpp:15-100,277-380) -------- */
int orc_parse_headers(const uint8_t* packet, int64_t nbytes, int64_t* primary8,
                      int64_t* secondary46);
/* out[0..9] = pri_us, pulse_length_us, swl_us, swst_us, rx_gain_db, txprr, txpsf,
 *             fs_mhz, replica_chirp_length, baq_block_length                     */
void orc_packet_params(const int64_t* secondary46, double* out);
void orc_slant_ranges(const int64_t* secondary46, int num_ranges, double* out);
void orc_timing_corrections(const int64_t* secondary46, double* out);
void orc_linspace(double start, double end, int size, double* out);
/* OpenMP threads for the oracle's parallel loops (timing the CPU baseline
 * at a stated thread count); returns the count in effect. */
int orc_set_threads(int n);

/* ---- FFT with FFTW's conventions (unnormalised both ways) --------------- */
void orc_fft(double* x, int n, int inverse);
void orc_fft_many(double* x, int n, int count, int64_t dist, int inverse);

/* ---- range compression (s1_decoder.cpp:448-497, image_formation.cpp:4-43) */
int orc_replica_chirp_length(double fs_mhz, double txpl_us);
void orc_replica_chirp(double txpsf, double txprr, double txpl, double fs_mhz,
                       int num_range, double* chirp_out);
void orc_reference_function(const double* chirp, int p, double* ref_out);
void orc_range_compress(double* data, int n_lines, int n_samples, int64_t ld,
                        double txpsf, double txprr, double txpl, double fs_mhz);

/* ---- azimuth chain (s1_decoder.cpp:499-505,599-682) ---------------------- */
void orc_azimuth_fft(double* data, int n_lines, int n_samples, int64_t ld);
void orc_azimuth_ifft(double* data, int n_lines, int n_samples, int64_t ld);
void orc_rcmc_mf(double* data, int n_lines, int n_samples, int64_t ld, double prf,
                 double range_spacing, double ref_range, const double* slant_ranges,
                 const double* v_rel, const double* f_dc, const double* t_corr,
                 int apply_rcmc);
void orc_sinc_interpolate(const double* row, int n, double t, int half_width,
                          double* out2);

/* ---- ECCM (signal_processing.cpp:913-1004) ------------------------------ */
void orc_eccm(double* data, int n_lines, int n_samples, int64_t ld, int fft_size,
              int stride, double detection_threshold, double mitigation_threshold,
              uint8_t* flagged_out);

/* ---- the oracle's tables, for comparison with the reference's ----------- */
void orc_tables(int* thidx, int* mflag, double* simple, double* nrl, double* sf, double* rgdec,
                double* consts, int* huff, int* huff_n, int* huff_start);

/* ---- autofocus reference (no reference counterpart) --------------------- */
double orc_af_entropy_grad(const double* rd, int n_lines, int n_samples, int64_t ld,
                           const double* phase, double* grad_out);

/* image_formation.cpp:391-476 */
void orc_effective_velocities(const double* pos, double v, const double* R, int n, double* out);
void orc_secondary_range_compression(double* line, int n, const double* vrel, const double* R,
                                     const double* fr, const double* fdc, double az, double* factors);
void orc_rcmc_phase(double* line, int n, const double* R, const double* fr, const double* factors);

/* TOPS: doppler.cpp:4-232, image_formation.cpp:87-388, s1_decoder.cpp:511-701,
 * signal_processing.cpp:32-214 */
void orc_dc_wrapped(const double* rc, int L, int n, int64_t ld, double rate, double prf, double* wrapped);
int orc_dc_unwrap(const double* wrapped, int n, int n_blocks, const double* blk_times, const double* times,
                  double prf, double* dc);
int orc_tops_freq_shape(int L, double nr);
void orc_az_freq_ufr(const double* in, int L, int n, int64_t ld, double nr, double inv_rate, double prf, double bw,
                     double* out);
void orc_tops_rows(double* data, int rows, int n, int64_t ld, const double* vrel, const double* R,
                   const double* fr, const double* fdc, const double* faz, const double* tc, double ref_range,
                   double* fm);
void orc_linear_resample_many(const double* in, int n_in, int n_out, int count, double* out);
void orc_quadratic_resample(const double* in, int n_in, int n_out, double* out);
int orc_tops_time_params(const double* fm_col0, int n_az, int n, double dc_mid, double fm_ref, double nr, double prf,
                         double T, double bw, int swath, double* d, int* k);
void orc_az_time_ufr(const double* in, int n_az, int n, int64_t ld, const double* dc, const double* fm, double rate,
                     const double* d, const int* k, double* out);

#ifdef __cplusplus
}
#endif


#endif
