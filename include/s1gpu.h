/*
 * s1gpu.h -- C ABI of the MI355X (gfx950) Sentinel-1 Level-0 -> SLC hot path.
 *
 * Plain C: pointers, sizes and POD descriptors only.  Unless stated
 * otherwise every data pointer is a DEVICE pointer (hipMalloc'd or from
 * another HIP allocator) and every call is asynchronous on the context's
 * stream.  Complex arrays are interleaved (re, im), row-major, one radar
 * line (echo / range line) per row; leading dimensions count complex
 * elements.  Precision is chosen per call:
 *   S1G_F64  complex128 data, FP64 arithmetic (the reference's arithmetic);
 *   S1G_F32  complex64 data, FP32 arithmetic and accumulation, with the
 *            replica-chirp phase, the range-Doppler phase and the RCMC
 *            offsets computed in FP64 and reduced mod 2 pi before FP32 use.
 *
 * Return value: 0 on success, a negative S1G_E* code otherwise;
 * s1g_last_error() returns a message.  No C++ exception crosses this ABI.
 *
 * Reference citations are file:line in AndrewPlayer3/sentinel1_decode src/.
 */
#ifndef S1GPU_H
#define S1GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define S1G_ABI_VERSION 1

/* precision selectors */
#define S1G_F64 0
#define S1G_F32 1

/* return codes */
#define S1G_OK 0
#define S1G_EARG (-1)        /* invalid argument                        */
#define S1G_EHIP (-2)        /* HIP runtime error                       */
#define S1G_EUNSUPPORTED (-3) /* length/shape this build cannot handle   */
#define S1G_ENOMEM (-4)      /* device allocation failed                */
#define S1G_EDECODE (-5)     /* >= 1 packet failed (s1g_check_status)   */

/* per-packet decode status (status[] of s1g_decode_fdbaq) */
#define S1G_PKT_OK 0
#define S1G_PKT_ERR_BRC 1     /* BRC field > 4 (packet.cpp:795-798)                */
#define S1G_PKT_ERR_OVERRUN 2 /* read past the user data (decoding_utils.cpp:107-119) */
#define S1G_PKT_ERR_HUFFMAN 3 /* code longer than 10 bits (decoding_utils.cpp:69-72) */
#define S1G_PKT_ERR_MCODE 4   /* m_code above the BRC flag (packet.cpp:690)         */
#define S1G_PKT_ERR_ARG 5     /* num_quads < 0, or packet above the stated maxima    */

typedef struct s1g_ctx s1g_ctx;

/* ------------------------------------------------------------ lifecycle */
/* Create a context on HIP device `device`.  A context owns a stream, the
 * FFT plans (twiddle tables, digit-reversal maps), the cached range
 * reference functions, a growable workspace and captured burst graphs.  A
 * context is not thread-safe; use one per host thread. */
int s1g_create(s1g_ctx** ctx, int device);
int s1g_destroy(s1g_ctx* ctx);
const char* s1g_last_error(const s1g_ctx* ctx);
/* Run subsequent work on `stream` (a hipStream_t; NULL is the null stream).
 * Until this is called a context uses a stream of its own. */
int s1g_set_stream(s1g_ctx* ctx, void* stream);
int s1g_synchronize(s1g_ctx* ctx);

/* ------------------------------------------------------- FDBAQ decode */
/* Decode n_packets FDBAQ (BAQ mode 12-14, data format D) packets.
 * Replaces L0Packet::get_signal (packet.h:205, packet.cpp:393-464) and the
 * per-packet decode loops decode_packets / decode_packets_in_place
 * (packet.h:228,238) and S1_Decoder::_range_compress (s1_decoder.cpp:462-468).
 *   user_data          byte buffer holding the packets' user data fields
 *   ud_offsets[i]      byte offset of packet i's user data in user_data
 *   ud_lengths[i]      its length in bytes (packet_data_length + 1 - 62)
 *   num_quads[i]       the packet's num_quadratures header field
 *   max_user_data_bytes, max_num_quads  host-side bounds over all packets
 *                      (they size on-chip buffers; larger packets fail
 *                      with S1G_PKT_ERR_ARG)
 *   out                complex [n_packets][ld_out] (F32: float2, F64:
 *                      double2), 2*num_quads[i] samples written per row
 *   status[i]          S1G_PKT_* per packet (rows of failed packets are
 *                      zero-filled)
 * Decoding is bit-exact: F64 output equals the reference's doubles, F32
 * output is those doubles rounded once to float. */
int s1g_decode_fdbaq(s1g_ctx* ctx, const uint8_t* user_data, const uint64_t* ud_offsets,
                     const uint32_t* ud_lengths, const int32_t* num_quads, int n_packets,
                     int max_user_data_bytes, int max_num_quads, int prec, void* out,
                     int64_t ld_out, int32_t* status);

/* Waits for the context's stream, then reads status[0..n_packets) (device
 * pointer, as written by s1g_decode_fdbaq or a burst chain).  Returns
 * S1G_EDECODE if any packet failed, with the index of the first one in
 * *first_failed (may be NULL; -1 when none failed) and its status in the
 * message of s1g_last_error.  The reference throws from the decode instead
 * (packet.cpp:795-798, decoding_utils.cpp:69-72,107-119); the decode calls
 * stay asynchronous and report through status[], and this call is the
 * synchronising check a caller makes where the reference would have thrown. */
int s1g_check_status(s1g_ctx* ctx, const int32_t* status, int n_packets, int* first_failed);

/* Host function: length (bits) and magnitude code of the FDBAQ Huffman code
 * word at the top of the 10-bit value `code10` for bit-rate code `brc`
 * (the decoder's symbol rule; huffman_decode, decoding_utils.cpp:44-77). */
int s1g_fdbaq_symbol(int brc, uint32_t code10, int* code_len, int* m_code);

/* -------------------------------------------------- packet headers (host) */
/* Decoded primary (6 B) and secondary (62 B) header of one space packet and
 * the physical quantities the reference derives from them (parse_packet_header,
 * packet.cpp:829-846; accessors packet.cpp:15-100,277-342,376-380). */
typedef struct {
    int64_t primary[8];     /* fields in the order of structs.h:42-51   */
    int64_t secondary[46];  /* fields in the order of structs.h:68-115  */
    int32_t num_quads;
    int32_t baq_mode;
    int32_t range_decimation;
    int32_t rank;
    int32_t swath_number;
    int32_t user_data_length;     /* packet_data_length + 1 - 62                 */
    int32_t replica_chirp_length; /* floor(fs * txpl)                            */
    int32_t baq_block_length;     /* 8 * (field + 1)                             */
    char data_format;             /* 'A'/'B'/'C'/'D' (packet.cpp:265-274)        */
    char reserved[7];
    double pri_us, pulse_length_us, swl_us, swst_us, rx_gain_db;
    double txprr_mhz_us, txpsf_mhz, fs_mhz;
} s1g_packet_info;

/* Host function.  Returns S1G_EARG if nbytes < 68 or the BAQ mode is not a
 * known one. */
int s1g_parse_packet_header(const uint8_t* packet, int64_t nbytes, s1g_packet_info* info);
/* Host functions: slant ranges (linspace over the sampling window,
 * packet.cpp:284-301) and timing corrections (packet.cpp:322-342) of a
 * packet's 2*num_quads range samples (num_ranges <= 0 means 2*num_quads). */
int s1g_slant_ranges(const s1g_packet_info* info, int num_ranges, double* out);
int s1g_timing_corrections(const s1g_packet_info* info, double* out);

/* ------------------------------------------ Level-0 files: framing */
/* A Level-0 file is a sequence of space packets: a 6-byte primary header
 * whose packet_data_length field (bytes 4-5) gives the packet's length,
 * 6 + packet_data_length + 1 bytes, a 62-byte secondary header, and the
 * user data field (get_next_packet / get_packets, packet.cpp:850-905). */
#define S1G_NFIELDS 54 /* 8 primary + 46 secondary header fields */

/* Host function.  Walks the packets of a Level-0 byte stream (host memory)
 * and writes the byte offset of each packet's first header byte to
 * offsets[0 .. min(count, cap)); *n_packets = count.  Stops at the end of
 * the stream; a last packet that runs past it is dropped, as the
 * reference's reader drops it. */
int s1g_index_level0(const uint8_t* stream, int64_t nbytes, uint64_t* offsets, int64_t cap, int64_t* n_packets);

/* Parses the 6 + 62 header bytes of n_packets packets on the GPU (one
 * wavefront per packet, one lane per field; parse_packet_header,
 * packet.cpp:829-846).  stream, offsets: device pointers (offsets[k] =
 * first header byte of packet k).  Outputs (device, any may be NULL):
 *   fields[k * S1G_NFIELDS + f]  header field f (primary fields 0-7, then
 *                                secondary fields in structs.h:68-115 order)
 *   ud_offsets[k], ud_lengths[k], num_quads[k]  the user-data offset
 *                                (offsets[k] + 68), length (packet_data_length
 *                                + 1 - 62) and num_quadratures: the arrays
 *                                s1g_decode_fdbaq takes. */
int s1g_parse_headers(s1g_ctx* ctx, const uint8_t* stream, const uint64_t* offsets, int n_packets,
                      int64_t* fields, uint64_t* ud_offsets, uint32_t* ud_lengths, int32_t* num_quads);

/* Host function.  The bursts of one swath (get_packets_in_bursts,
 * packet.cpp:1026-1076) over the packets' header fields (host, as
 * s1g_parse_headers writes them): among packets whose swath_number equals
 * `swath_number`, a burst ends where azimuth_beam_address decreases or
 * pri_count jumps by more than one; a burst that ends with fewer than 1000
 * packets is dropped, the last one is kept whatever its size; with cal != 0
 * all the swath's packets form one burst.  burst_of_packet[k] = the burst
 * index of packet k, or -1 (other swaths, dropped bursts); *n_bursts = the
 * number of bursts. */
int s1g_segment_bursts(const int64_t* fields, int n_packets, int swath_number, int cal, int32_t* burst_of_packet,
                       int* n_bursts);

/* ------------------------------------ orbit geometry (ancillary data) */
/* One state vector: the time stamp (seconds), ECEF position (m) and
 * velocity (m/s), and the raw attitude words (quaternion q0..q3 and angular
 * rates as transmitted; their conversion to Euler angles is not provided). */
typedef struct {
    double time;
    double position[3];
    double velocity[3];
    double quaternion[4];
    double angular_rate[3];
} s1g_state_vector;

/* Host function.  The state vectors carried by the packets' sub-commutated
 * ancillary words (sc_data_word_index / sc_data_word: build_data_word_dicts
 * and STATE_VECTORS::parse_subcomm_dicts, state_vectors.cpp:19-91): one per
 * complete cycle of 64 words, kept when its time stamp exceeds every earlier
 * one.  fields: host [n_packets][S1G_NFIELDS].  Writes at most cap vectors;
 * *n_out = the number found. */
int s1g_state_vectors(const int64_t* fields, int n_packets, s1g_state_vector* out, int cap, int* n_out);

/* Host function.  STATE_VECTORS::interpolate (state_vectors.cpp:141-200):
 * linear interpolation between the vectors bracketing `time` (bisection as
 * the reference does it), the first / last vector outside their span. */
int s1g_interpolate_state(const s1g_state_vector* vectors, int n, double time, s1g_state_vector* out);

/* Host function.  The platform state at t0 + dt * i for i < count (the
 * per-line states of S1_Decoder::_azimuth_compress, s1_decoder.cpp:
 * 579-591): positions[3 i .. 3 i + 2] and velocity_norms[i] =
 * sqrt(vx^2 + vy^2 + vz^2) of s1g_interpolate_state's result. */
int s1g_line_states(const s1g_state_vector* vectors, int n, double t0, double dt, int count, double* positions,
                    double* velocity_norms);

/* Effective velocities for n_lines platform states and n_samples slant
 * ranges (get_effective_velocities, image_formation.cpp:391-419):
 * out[i * ld_out + j] = v_rel of line i at range j, from positions[3 i ..
 * 3 i + 2] (ECEF, m) and velocity_norms[i] (m/s) over the WGS84 ellipsoid.
 * Device pointers, FP64. */
int s1g_effective_velocities(s1g_ctx* ctx, const double* positions, const double* velocity_norms, int n_lines,
                             const double* slant_ranges, int n_samples, double* out, int64_t ld_out);

/* apply_secondary_range_compression (image_formation.cpp:422-455) on
 * n_rows range lines: for row i and sample j, with f = az_freqs[i] +
 * f_dc[j], D = sqrt(1 - lambda^2 f^2 / (4 v^2)) (v = v_rel[i * n_samples +
 * j]) and K = 2 v^2 f0^3 D^2 / (c R_j f^2),
 *   data[i][j] *= exp(-i PI range_freqs[j]^2 / K),
 *   rcmc_factors[i * n_samples + j] = D (may be NULL).
 * Device pointers; data complex of `prec`, the rest FP64. */
int s1g_secondary_range_compression(s1g_ctx* ctx, int prec, int n_rows, int n_samples, void* data, int64_t ld,
                                    const double* v_rel, const double* slant_ranges, const double* range_freqs,
                                    const double* f_dc, const double* az_freqs, double* rcmc_factors);

/* apply_range_cell_migration_correction (image_formation.cpp:458-476) on
 * n_rows lines: data[i][j] *= exp(-4 i PI range_freqs[j] (R_j / D_ij - R_j) / c),
 * D_ij = rcmc_factors[i * n_samples + j].  Device pointers. */
int s1g_rcmc_phase(s1g_ctx* ctx, int prec, int n_rows, int n_samples, void* data, int64_t ld,
                   const double* slant_ranges, const double* range_freqs, const double* rcmc_factors);

/* ---------------------------------------------------- range compression */
/* Replica-chirp parameters of a burst (from its first packet:
 * L0Packet::get_start_frequency / get_tx_ramp_rate / get_pulse_length /
 * get_range_sample_rate, packet.cpp:36-100,277-281). */
typedef struct {
    int32_t n_lines;      /* echo lines (packets)                         */
    int32_t n_samples;    /* complex samples per line, 2 * num_quads       */
    double txpsf_mhz;     /* TXPSF, MHz                                    */
    double txprr_mhz_us;  /* TXPRR, MHz/us                                 */
    double txpl_us;       /* pulse length, us                              */
    double fs_mhz;        /* range sample rate, MHz                        */
} s1g_range_desc;

/* Range-compress n_lines lines (out-of-place).  Replaces the range part of
 * S1_Decoder::_range_compress (s1_decoder.cpp:476-497): pad to
 * P = N + 2 nc, FFT, multiply by get_reference_function(get_replica_chirp())
 * (image_formation.cpp:4-43, packet.cpp:345-380), IFFT, fftshift, crop.
 * The result equals that linear correlation including its scale
 * nc^2 * P; the FFT length used internally is chosen by the library. */
int s1g_range_compress(s1g_ctx* ctx, const s1g_range_desc* desc, int prec, const void* in,
                       int64_t ld_in, void* out, int64_t ld_out);

/* Replica chirp length nc = floor(fs * txpl) (L0Packet::get_replica_chirp_length,
 * packet.cpp:376-380), host function. */
int s1g_replica_chirp_length(double fs_mhz, double txpl_us);

/* ------------------------------------------------------------- azimuth */
/* Azimuth (column) FFT of an [n_lines][n_samples] burst, in place.
 * inverse = 0: FFT along azimuth then fftshift (s1_decoder.cpp:499-505);
 * inverse = 1: inverse FFT times 1/n_lines, no shift
 * (compute_axis_dft_in_place(data, 0, 0, true), s1_decoder.cpp:682,
 * signal_processing.cpp:581-605). */
int s1g_azimuth_fft(s1g_ctx* ctx, int prec, int n_lines, int n_samples, void* data, int64_t ld,
                    int inverse);

/* Range-Doppler processing of one burst.  Row i of the input is the
 * fftshifted Doppler bin f_i = linspace(-prf/2, prf/2, n_lines)[i]
 * (s1_decoder.cpp:604).  Per sample (i, j):
 *   D  = sqrt(1 - lambda^2 (f_i + f_dc[j])^2 / (4 v_rel[j]^2))
 *        (apply_secondary_range_compression, image_formation.cpp:439-441)
 *   RCMC: x <- sinc_interpolate(row_i, j + (R_j / D - R_j) / range_spacing, 8)
 *        (signal_processing.cpp:238-257 with the migration of
 *         image_formation.cpp:469)  [if apply_rcmc]
 *   x /= sqrt((ref_range / R_j)^3)                       (s1_decoder.cpp:643-646)
 *   x *= exp(i 4 pi R_j D / lambda) exp(i 2 pi (f_i + f_dc[j]) t_corr[j])
 *                                                          (s1_decoder.cpp:658-676)
 * Arrays are device pointers of n_samples doubles. */
typedef struct {
    int32_t n_lines;
    int32_t n_samples;
    int32_t apply_rcmc;
    int32_t reserved;
    double prf;            /* Hz                                              */
    double wavelength;     /* m                                               */
    double range_spacing;  /* slant-range step of the range grid, m           */
    double ref_range;      /* R_ref = slant_ranges[n_samples / 2], m          */
    const double* slant_ranges; /* R_j  (L0Packet::get_slant_ranges)          */
    const double* v_rel;        /* effective velocity per range sample, m/s   */
    const double* f_dc;         /* Doppler centroid per range sample, Hz      */
    const double* t_corr;       /* L0Packet::get_timing_corrections, s        */
} s1g_azimuth_desc;

int s1g_rcmc_matched_filter(s1g_ctx* ctx, const s1g_azimuth_desc* desc, int prec, const void* in,
                            int64_t ld_in, void* out, int64_t ld_out);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same stage and the same kernels as
 * s1g_rcmc_matched_filter, but its per-column geometry table lives in a
 * fixed device array of the library (one per GPU, about 5.5 MB, held until
 * the process ends) instead of context memory, and the call makes no device
 * allocation, frees nothing and times no launch variants (it uses a forced
 * variant, the one the tuner already chose for the shape, or the tuner's
 * usual pick).  n_samples at most S1G_NOALLOC_MAX_SAMPLES (S1G_EARG above
 * it).  Every call on a GPU shares that GPU's table, which each call
 * rewrites in full: calls that overlap on one GPU (other streams or other
 * contexts) race. */
#define S1G_NOALLOC_MAX_SAMPLES 65536
int s1g_rcmc_matched_filter_noalloc(s1g_ctx* ctx, const s1g_azimuth_desc* desc, int prec, const void* in,
                                    int64_t ld_in, void* out, int64_t ld_out);

/* ------------------------------------------------- TOPS (IW) focusing */
/* The TOPS branch of S1_Decoder::_azimuth_compress (s1_decoder.cpp:511-701):
 * Doppler centroid estimation, azimuth-frequency unfolding, the per-line
 * filter, azimuth-time unfolding.  Device arrays are [lines][samples]
 * (range samples contiguous); all stages compute in FP64 (prec selects the
 * element type of the complex data: S1G_F32 or S1G_F64). */

/* L0Packet::get_azimuth_beam_angle (packet.cpp:83-90):
 * linspace(-0.018, 0.018, 1024)[address], radians (host). */
double s1g_azimuth_beam_angle(int address);

/* get_doppler_centroid_rate (doppler.cpp:224-232): -2 v k / lambda with the
 * steering rate k = (angle(last) - angle(first)) / (pri n_packets), Hz/s
 * (host). */
double s1g_doppler_centroid_rate(double pri_us, int first_beam_address, int last_beam_address, int n_packets,
                                 double velocity);

/* get_doppler_centroid (doppler.cpp:146-221) of a range-compressed burst
 * (device, [n_lines][n_samples]): the burst deramped at `rate` over 1.06 s,
 * lag-one phase per range sample, block means over n_blocks range blocks
 * (30 in the reference), get_unwrapped_estimates, a quadratic fit over the
 * slant-range times linspace(t_near, t_far, .) (L0Packet::
 * get_slant_range_times of the first packet) and its value at every
 * sample: dc[n_samples] (device, FP64).  The lag sums are serial per sample
 * and the rest runs on the host in FP64, so the result is that of the
 * reference's arithmetic on the same input. */
int s1g_doppler_centroid(s1g_ctx* ctx, int prec, int n_lines, int n_samples, const void* data, int64_t ld,
                         double rate, double prf, double t_near, double t_far, int n_blocks, double* dc);

/* Lines of azimuth_frequency_ufr's output: floor(|rate T / prf| n_lines). */
int s1g_tops_frequency_lines(int n_lines, double rate, double burst_duration, double prf);

/* azimuth_frequency_ufr (image_formation.cpp:122-235): per range sample,
 * the fftshifted azimuth spectrum tiled |rate T / prf| times, deramped
 * exp(-i pi f^2 / rate), low-passed to floor(doppler_bandwidth) bins moved to
 * the start, re-ramped.  in: [n_lines][n_samples], out:
 * [s1g_tops_frequency_lines(...)][n_samples]; device. */
int s1g_azimuth_frequency_ufr(s1g_ctx* ctx, int prec, int n_lines, int n_samples, const void* in, int64_t ld_in,
                              double rate, double burst_duration, double prf, double doppler_bandwidth, void* out,
                              int64_t ld_out);

/* The TOPS branch of the per-line loop of S1_Decoder::_azimuth_compress
 * (s1_decoder.cpp:626-677) on n_lines lines in place: range FFT, secondary
 * range compression at az_freqs[i] + f_dc[j] with v_rel from
 * get_effective_velocities(positions[i], velocity_norms[i], R), unnormalised
 * range IFFT, division by sqrt((ref_range / R_j)^3), RCMC phase
 * exp(-4 i pi range_freqs[j] (R_j / D - R_j) / c), azimuth filter
 * exp(4 i pi R_j D / lambda) exp(2 i pi (az_freqs[i] + f_dc[j]) t_corr[j]);
 * fm[i][j] = -2 v^2 D^3 / (lambda R_j) (the azimuth FM rate).  Device:
 * positions [n_lines][3], velocity_norms, az_freqs [n_lines], slant_ranges,
 * range_freqs, f_dc, t_corr [n_samples], fm [n_lines][n_samples], FP64. */
int s1g_tops_filter_lines(s1g_ctx* ctx, int prec, int n_lines, int n_samples, void* data, int64_t ld,
                          const double* positions, const double* velocity_norms, const double* az_freqs,
                          const double* slant_ranges, const double* range_freqs, const double* f_dc,
                          const double* t_corr, double ref_range, double* fm);

/* The scalar set-up of azimuth_time_ufr (image_formation.cpp:238-292). */
typedef struct {
    double bandwidth;        /* nr prf - doppler_bandwidth, Hz              */
    double focused_time;     /* s                                           */
    double n_ref;            /* f_dc / k_a at the reference point, s        */
    double ka_mean;          /* |mean FM rate| of range sample 0, Hz/s      */
    int32_t num_tiles;       /* time replicas                               */
    int32_t shape;           /* num_tiles * n_lines                         */
    int32_t lpf_start, lpf_end;
    int32_t downsample_shape;
    int32_t output_shape;    /* floor(prf T)                                */
    int32_t crop_start, crop_end; /* output lines [crop_start, crop_end)    */
} s1g_time_ufr_params;

/* azimuth_time_ufr's set-up from f_dc[n_samples] and fm[n_lines][n_samples]
 * (device).  n_ref divides f_dc[n_samples / 2] by fm at line n_samples / 2,
 * sample n_lines / 2 (the reference's index on its transposed array) where
 * those are in range, else by fm at line n_lines / 2, sample n_samples / 2
 * (the reference reads past its array there).  S1G_EARG when the low-pass
 * window or the crop leaves the arrays. */
int s1g_azimuth_time_ufr_params(s1g_ctx* ctx, int n_lines, int n_samples, const double* f_dc, const double* fm,
                                double rate, double burst_duration, double prf, double doppler_bandwidth,
                                int swath_number, s1g_time_ufr_params* out);

/* azimuth_time_ufr (image_formation.cpp:238-388) on an azimuth-compressed
 * burst in: [n_lines][n_samples]: per range sample, tiled num_tiles times,
 * deramped with the FM rate linearly resampled to the tiled length,
 * low-passed, quadratically resampled to downsample_shape lines, re-ramped
 * and cropped.  out: [crop_end - crop_start][n_samples]; device. */
int s1g_azimuth_time_ufr(s1g_ctx* ctx, int prec, int n_lines, int n_samples, const void* in, int64_t ld_in,
                         const double* f_dc, const double* fm, double rate, double burst_duration, double prf,
                         double doppler_bandwidth, int swath_number, void* out, int64_t ld_out);

/* ---------------------------------------------------------------- ECCM */
/* RFI detection and notch, in place (eccm, signal_processing.cpp:913-1004,
 * as called from s1_decoder.cpp:470-474 with fft_size 128, stride 64).
 * A line is processed when max |FFT_N(line)| >= detection_threshold; then
 * each 128-sample window's spectrum bins whose |central gradient| reaches
 * mitigation_threshold are zeroed and the windows are written back in
 * order.  flags (device, n_lines bytes, may be NULL) receives 1 for
 * processed lines. */
typedef struct {
    int32_t n_lines;
    int32_t n_samples;
    int32_t fft_size;
    int32_t stride;
    double detection_threshold;
    double mitigation_threshold;
} s1g_eccm_desc;

int s1g_eccm(s1g_ctx* ctx, const s1g_eccm_desc* desc, int prec, void* data, int64_t ld,
             uint8_t* flags);

/* --------------------------------------------------- whole burst chain */
/* One burst from packets to SLC: decode -> [ECCM] -> range compression ->
 * azimuth FFT -> RCMC + matched filter -> azimuth IFFT
 * (S1_Decoder::_azimuth_compress, s1_decoder.cpp:511-701, on its
 * non-TOPS formulas).  Intermediate arrays live in the context workspace. */
typedef struct {
    const uint8_t* user_data;
    const uint64_t* ud_offsets;
    const uint32_t* ud_lengths;
    const int32_t* num_quads;
    int32_t* status;            /* n_lines per-packet decode status      */
    int32_t max_user_data_bytes;
    int32_t max_num_quads;
    int32_t use_eccm;
    int32_t reserved;
    s1g_range_desc range;
    s1g_azimuth_desc azimuth;
    s1g_eccm_desc eccm;
    uint8_t* eccm_flags;        /* may be NULL                           */
    void* out;                  /* SLC [n_lines][ld_out]                 */
    int64_t ld_out;
} s1g_burst_desc;

int s1g_focus_burst(s1g_ctx* ctx, const s1g_burst_desc* desc, int prec);

/* Several bursts (uneven sizes allowed) in one submission
 * (_get_azimuth_compressed_swath_iw's burst loop, s1_decoder.cpp:307-349).
 * flags:
 *   S1G_BATCH_GRAPH   capture the whole batch as one hipGraph and replay it
 *                     (the graph is cached per descriptor set);
 *   S1G_BATCH_STAGED  run the batch stage by stage: every decode, then range
 *                     compression, the azimuth FFT, RCMC + matched filter
 *                     and the azimuth IFFT each as one launch over all
 *                     bursts of a shape class (the persistent kernels take a
 *                     table of per-burst descriptors), with a workspace of
 *                     the batch's total size in the context.  Needs no ECCM
 *                     and disjoint outputs (else the default order is used).
 * By default the chains run burst by burst; when no burst uses ECCM and no
 * two bursts' SLC rows or status words overlap, the decode of burst i+1
 * runs on a second stream of the context beside the focusing of burst i
 * (each burst's focusing waits only on its own decode).  Burst by burst is
 * the faster order on MI355X: a burst's intermediate array (284 MB for
 * IW2 FP32) is read back by the next stage while it is still in the
 * 256 MB last-level cache, which a stage over all bursts gives up.  Either
 * way the batch is ordered on the context's stream as one unit.  A burst's
 * SLC must not alias another burst's packet data. */
#define S1G_BATCH_GRAPH 1
#define S1G_BATCH_STAGED 2
int s1g_focus_batch(s1g_ctx* ctx, const s1g_burst_desc* descs, int n_bursts, int prec, int flags);

/* ------------------------------------------- general signal processing */
/* The reference's building blocks, for callers that use them one by one.
 * All in the context stream, device pointers, precision per call. */

/* DFT of every row (axis = 1, length n_cols) or every column (axis = 0,
 * length n_rows) of an [n_rows][ld] array, in place: forward unnormalised
 * e^{-2 pi i jk/n}, inverse e^{+...} times 1/n -- compute_1d_dft,
 * compute_1d_dft_in_place, _compute_axis_dft, compute_axis_dft(_in_place)
 * (signal_processing.cpp:516-681, FFTW conventions).  Any length: Stockham
 * passes when n factors over radices up to 71, else Bluestein's chirp-z
 * transform through a power-of-two length. */
int s1g_dft(s1g_ctx* ctx, int prec, int n_rows, int n_cols, void* data, int64_t ld, int axis, int inverse);

/* An FFT plan: the opaque handle that stands in for the per-row FFTW plans
 * of get_fftw_plans / destroy_fftw_plans (signal_processing.cpp:684-710).
 * Like those plans it is unnormalised in both directions (inverse = 0:
 * e^{-2 pi i jk/n}, 1: e^{+...}, no 1/n).  A plan belongs to its context's
 * stream; s1g_plan_execute transforms n_rows rows of length n of a device
 * array in place. */
typedef struct s1g_plan s1g_plan;
int s1g_plan_create(s1g_ctx* ctx, int prec, int n, int inverse, s1g_plan** plan);
int s1g_plan_execute(s1g_plan* plan, void* data, int64_t ld, int n_rows);
int s1g_plan_destroy(s1g_plan* plan);

/* dst[c][r] = src[r][c] (transpose, signal_processing.cpp:463-498); out of place. */
int s1g_transpose(s1g_ctx* ctx, int prec, int n_rows, int n_cols, const void* src, int64_t ld_src, void* dst,
                  int64_t ld_dst);

/* data[r][k] *= vec[k] (or conj(vec[k])) for k < n: the product of
 * pulse_compression (image_formation.cpp:28-43) applied to every row. */
int s1g_multiply(s1g_ctx* ctx, int prec, int n_rows, int n, void* data, int64_t ld, const void* vec, int conj_vec);

/* out = conj(DFT(chirp)) / E, E = sum |c|^4 / n (get_reference_function,
 * image_formation.cpp:4-25); out may equal chirp. */
int s1g_reference_function(s1g_ctx* ctx, int prec, const void* chirp, int n, void* out);

/* The replica chirp of L0Packet::get_replica_chirp (packet.cpp:345-373):
 * n_samples + 2 nc values, nc = floor(fs txpl), zero except the nc samples
 * (1/nc) exp(i 2 PI (txpsf t + txprr/2 t^2)) centred in the vector. */
int s1g_replica_chirp(s1g_ctx* ctx, int prec, int n_samples, double txpsf_mhz, double txprr_mhz_us, double txpl_us,
                      double fs_mhz, void* out);

/* out[q] = sinc_interpolate(signal, t[q], half_width) for q < m
 * (signal_processing.cpp:228-257): Hann-windowed sinc over the 2 L + 1
 * samples around floor(t).  t: device doubles. */
int s1g_sinc_interpolate(s1g_ctx* ctx, int prec, const void* signal, int len, const double* t, int m, int half_width,
                         void* out);

/* Host function: s1g_packet_info from decoded header field values (for
 * callers holding the reference's header maps, packet.h:142-167). */
int s1g_packet_info_from_fields(const int64_t* primary, const int64_t* secondary, s1g_packet_info* info);

/* ------------------------------------------------------- image output */
/* Scaling of a complex image to float for display (scale, norm_2d,
 * magnitude_2d, signal_processing.cpp:367-460,1007-1071):
 *   S1G_SCALE_NORM_LOG  20 log10((|x| + 1e-12) / max|x|)
 *   S1G_SCALE_NORM      (|x| + 1e-12) / max|x|
 *   S1G_SCALE_MAG       |x|^2
 *   S1G_SCALE_REAL / S1G_SCALE_IMAG  the real / imaginary part
 * |x| is evaluated in FP64 for either precision and the result rounded once
 * to float.  data: device [n_rows][ld] complex of `prec`; out: device float
 * [n_rows][ld_out]. */
#define S1G_SCALE_NORM_LOG 0
#define S1G_SCALE_NORM 1
#define S1G_SCALE_MAG 2
#define S1G_SCALE_REAL 3
#define S1G_SCALE_IMAG 4
int s1g_scale(s1g_ctx* ctx, int prec, int mode, int n_rows, int n_cols, const void* data, int64_t ld, float* out,
              int64_t ld_out);

/* complex128 -> complex64, each part rounded to nearest (the cf32 dumps of
 * image_write_main.cpp:186-324).  Device pointers. */
int s1g_to_cf32(s1g_ctx* ctx, int n_rows, int n_cols, const void* data, int64_t ld, void* out, int64_t ld_out);

/* Host function: writes a float image as an uncompressed BigTIFF (32-bit
 * IEEE float samples, one per pixel; _write_tif, image_write.cpp:3-25), with
 * its rows in reverse order when flip_rows != 0 (the reference writes the
 * last data row first).  host_data: [n_rows][n_cols] floats. */
int s1g_write_tiff(const char* path, const float* host_data, int n_rows, int n_cols, int flip_rows);

/* ----------------------------------------------------------- autofocus */
/* (No reference counterpart.)  rd: range-Doppler data after RCMC and
 * matched filter, [n_lines][ld] (rows = Doppler bins).  phase: device
 * double[n_lines], per-row azimuth phase correction.
 * s1g_af_forward: image = IFFT_az(rd * e^{i phase}) / n_lines written to img
 *   (may be NULL), entropy E = -sum p log p, p = |img|^2 / sum |img|^2,
 *   returned through *entropy (host pointer; the call synchronises).
 * s1g_af_adjoint: additionally grad (device double[n_lines]) = dE/dphase. */
int s1g_af_forward(s1g_ctx* ctx, int prec, int n_lines, int n_samples, const void* rd, int64_t ld,
                   const double* phase, void* img, double* entropy);
int s1g_af_adjoint(s1g_ctx* ctx, int prec, int n_lines, int n_samples, const void* rd, int64_t ld,
                   const double* phase, double* entropy, double* grad);

/* ------------------------------------------------------- stage timers */
/* When enabled, every stage launch is bracketed by HIP events on the
 * context's stream; s1g_profile_read synchronises, returns per-stage total
 * milliseconds and launch counts (arrays of S1G_NSTAGES) and optionally
 * resets them.  Timers are not recorded while a batch is being captured. */
#define S1G_STAGE_DECODE 0
#define S1G_STAGE_ECCM 1
#define S1G_STAGE_RANGE 2
#define S1G_STAGE_AZ_FFT 3
#define S1G_STAGE_RCMC_MF 4
#define S1G_STAGE_AZ_IFFT 5
#define S1G_STAGE_FILTER 6
#define S1G_STAGE_AUTOFOCUS 7
#define S1G_STAGE_DFT 8
#define S1G_NSTAGES 9
int s1g_profile_enable(s1g_ctx* ctx, int on);
int s1g_profile_read(s1g_ctx* ctx, double* ms, int64_t* launches, int reset);

/* ----------------------------------------------------------- autotuner */
/* (No reference counterpart.)  Stages with several launch variants --
 * S1G_STAGE_RANGE (overlap-save block length), S1G_STAGE_AZ_FFT and
 * S1G_STAGE_AZ_IFFT (panel width, compiled radix sequence or generic
 * kernel), S1G_STAGE_RCMC_MF (workgroup tile), S1G_STAGE_DFT (the radix
 * order of s1g_dft's rows that fit one workgroup's LDS; its timer slot stays
 * empty) -- time every candidate on the first eager call for a shape (the
 * in-place transforms on scratch arrays), keep the fastest in the context,
 * and reuse it; all variants give the same result to rounding.  Launches recorded into a graph use the
 * cached choice; s1g_focus_batch runs the chain once eagerly before it
 * captures.  Autotuning is on by default. */
int s1g_set_autotune(s1g_ctx* ctx, int on);
/* Use variant v (>= 0, clamped to the stage's count) for every later call of
 * a stage, or return to the automatic choice with v = -1. */
int s1g_force_variant(s1g_ctx* ctx, int stage, int variant);
/* The variant the stage's last call used and how many it had to choose from. */
int s1g_last_variant(const s1g_ctx* ctx, int stage, int* variant, int* n_variants);
/* The tuning cache as a file (host functions): save writes every choice of
 * the context -- (stage, precision, shape) -> variant, the number of
 * variants it was chosen among and its time -- under a header naming the
 * device; load merges a file saved on the same kind of device (S1G_EUNSUPPORTED
 * otherwise) and its entries are used instead of timing, as long as the
 * stage still offers the same number of variants.  With the environment
 * variable S1G_TUNE_CACHE set, s1g_create loads that file (when present) and
 * s1g_destroy saves the context's cache to it. */
int s1g_tune_save(const s1g_ctx* ctx, const char* path);
int s1g_tune_load(s1g_ctx* ctx, const char* path);

#ifdef __cplusplus
}
#endif
#endif /* S1GPU_H */
