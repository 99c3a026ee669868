/*
 * image_formation.h -- drop-in image-formation functions of the reference's
 * image_formation.h (lines 14-69) used on the L0 -> SLC path, on the GPU
 * (FP64), including the TOPS unfolding (azimuth_frequency_ufr,
 * azimuth_time_ufr).
 */
#pragma once

#include <utility>

#include "misc_types.h"
#include "packet.h"

/* signal (a line's range spectrum) times reference, elementwise over the
 * reference's length, truncated to the signal's length
 * (image_formation.cpp:28-43).  The reference may not be longer than the
 * signal (std::invalid_argument). */
CF_VEC_1D pulse_compression(const CF_VEC_1D& signal, const CF_VEC_1D& reference);

/* conj(DFT(replica_chirp)) / E, E = sum |c|^4 / n (image_formation.cpp:4-25). */
CF_VEC_1D get_reference_function(const CF_VEC_1D& replica_chirp);

/* Stripmap packets split where the line length (2 * num_quads) changes,
 * with the longest line length (image_formation.cpp:46-86). */
std::pair<PACKET_VEC_2D, int> get_azimuth_blocks(PACKET_VEC_1D& packets);

/* Effective velocity per slant range from the platform's ECEF position and
 * speed over the WGS84 ellipsoid (image_formation.cpp:391-419). */
D_VEC_1D get_effective_velocities(const D_VEC_1D& position, const double& velocity, const D_VEC_1D& slant_ranges);

/* Multiplies the range line by the secondary range compression filter
 * exp(-i PI f_r^2 / K_src) and returns the range-cell-migration factors
 * D = sqrt(1 - lambda^2 (f_a + f_dc)^2 / (4 v^2)) (image_formation.cpp:422-455). */
D_VEC_1D apply_secondary_range_compression(CF_VEC_1D& range_line, const D_VEC_1D& effective_velocities,
                                           const D_VEC_1D& slant_ranges, const D_VEC_1D& range_freqs,
                                           const D_VEC_1D& doppler_centroids, const double& az_freq);

/* Multiplies the range line by exp(-4 i PI f_r (R / D - R) / c)
 * (image_formation.cpp:458-476). */
void apply_range_cell_migration_correction(CF_VEC_1D& range_line, const D_VEC_1D& slant_ranges,
                                           const D_VEC_1D& range_freqs, const D_VEC_1D& rcmc_factors);

/* floor(num_replicas) copies of the signal followed by the first
 * floor(frac(num_replicas) * size) samples, in floor(num_replicas * size)
 * samples (get_tiled_signal, image_formation.cpp:87-119). */
CF_VEC_1D get_tiled_signal(CF_VEC_1D& signal, const double& num_replicas);

/* TOPS azimuth-frequency unfolding of a range-compressed burst
 * (azimuth_frequency_ufr, image_formation.cpp:122-235):
 * floor(|dc_rate T / prf| lines) output lines.  The reference transposes
 * and transforms its argument in place on the way; here the argument is left
 * as it was.  dc_estimates and initial_packet are unused, as there. */
CF_VEC_2D azimuth_frequency_ufr(CF_VEC_2D& range_compressed, D_VEC_1D& dc_estimates, L0Packet& initial_packet,
                                const double& dc_rate, const double& burst_duration, const double& prf,
                                const double& doppler_bandwidth);

/* TOPS azimuth-time unfolding of an azimuth-compressed burst
 * (azimuth_time_ufr, image_formation.cpp:238-388) with the azimuth FM rate
 * az_fm_rate[line][sample]: floor(prf T) / 2 * 2 output lines.  The
 * arguments are left as they were (the reference transposes two of them).
 * Where the reference's reference-point index leaves the FM-rate array (for
 * bursts with more than twice as many samples as lines) the FM rate at the
 * middle line and sample is used. */
CF_VEC_2D azimuth_time_ufr(CF_VEC_2D& azimuth_compressed, D_VEC_1D& dc_estimates, D_VEC_2D& az_fm_rate,
                           L0Packet& initial_packet, const double& dc_rate, const double& burst_duration,
                           const double& prf, const double& doppler_bandwidth, const int& swath_number);
