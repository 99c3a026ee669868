/*
 * s1_decoder.h -- the burst-level work of S1_Decoder for one burst's
 * packets (the bodies of S1_Decoder::_range_compress and _azimuth_compress,
 * s1_decoder.cpp:448-701, on their non-TOPS formulas), on the GPU.  File
 * reading, swath/burst bookkeeping and orbit processing of S1_Decoder are
 * outside the path: the caller passes the burst's packets and the effective
 * velocity / Doppler centroid per range sample (one value broadcasts).
 */
#pragma once

#include "packet.h"

struct S1_ECCM_THRESHOLDS {
    double detection;
    double mitigation;
};

/* Decode (+ ECCM) + range compression, optionally followed by the azimuth
 * FFT with fftshift (S1_Decoder::_range_compress, s1_decoder.cpp:448-508). */
CF_VEC_2D range_compress_burst(PACKET_VEC_1D& packets, const bool& do_azimuth_fft = false,
                               const S1_ECCM_THRESHOLDS* eccm = nullptr);

/* The whole chain to the focused burst: decode (+ ECCM), range compression,
 * azimuth FFT, RCMC + matched filter, azimuth IFFT
 * (S1_Decoder::_azimuth_compress, s1_decoder.cpp:511-701). */
CF_VEC_2D azimuth_compress_burst(PACKET_VEC_1D& packets, const D_VEC_1D& effective_velocities,
                                 const D_VEC_1D& doppler_centroids, const S1_ECCM_THRESHOLDS* eccm = nullptr);
