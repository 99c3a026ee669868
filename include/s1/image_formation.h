/*
 * image_formation.h -- drop-in get_reference_function and pulse_compression
 * (image_formation.h:14-21 of the reference), on the GPU (FP64).
 */
#pragma once

#include "misc_types.h"

/* signal (a line's range spectrum) times reference, elementwise over the
 * reference's length, truncated to the signal's length
 * (image_formation.cpp:28-43).  The reference may not be longer than the
 * signal (std::invalid_argument). */
CF_VEC_1D pulse_compression(const CF_VEC_1D& signal, const CF_VEC_1D& reference);

/* conj(DFT(replica_chirp)) / E, E = sum |c|^4 / n (image_formation.cpp:4-25). */
CF_VEC_1D get_reference_function(const CF_VEC_1D& replica_chirp);
