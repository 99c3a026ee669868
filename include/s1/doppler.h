/*
 * doppler.h -- drop-in Doppler-centroid functions of the reference's
 * doppler.h (lines 5-30), on the GPU through the C ABI.
 */
#pragma once

#include "misc_types.h"
#include "packet.h"

/* The Doppler centroid per range sample of a range-compressed burst
 * (get_doppler_centroid, doppler.cpp:146-221): the burst deramped at
 * doppler_centroid_rate over the reference's fixed 1.06 s, the lag-one
 * phase per sample, means over num_rng_blocks range blocks, the
 * reference's unwrapping, and a quadratic fit over the first packet's
 * slant-range times evaluated at every sample.  burst_duration is accepted
 * and unused, as in the reference. */
D_VEC_1D get_doppler_centroid(CF_VEC_2D& range_compressed, const double& doppler_centroid_rate,
                              const double& burst_duration, L0Packet& first_packet, const int& num_rng_blocks = 15);

/* -2 v k / lambda, k the azimuth steering rate from the first and last
 * packets' beam angles over pri * packets (get_doppler_centroid_rate,
 * doppler.cpp:224-232). */
double get_doppler_centroid_rate(PACKET_VEC_1D& burst_packets, const double& velocity);
